/*
 * clg_hip.h -- C ABI of the MI355X-native fermion-matrix library (libclg_hip.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes.  The
 * library never aborts; every entry point returns a clg_status and leaves a
 * message for clg_last_error().
 *
 * Each entry point names the reference (CudaLatticeGauge / CLGLib) interface it
 * replaces.  Host arrays exchanged with the library use the reference's own AoS
 * order in FP64 (interleaved re/im):
 *   site = ((x*Ly + y)*Lz + z)*Lt + t                          CCommonData.h:274-282
 *   Wilson spinor psi[site][spin][colour]        24 doubles     Vectors.h:575-888
 *   staggered vector chi[site][colour]            6 doubles     Vectors.h:287-571
 *   gauge U[site][mu][row][col]                  72 doubles     SU3.h:60 (m_me[3*row+col])
 * Half-lattice (EVEN/ODD) fields exchange V/2 sites in the reference's even/odd
 * index order: the e-th even (odd) site is whichever of the lexicographic sites
 * 2e, 2e+1 has even (odd) parity (CLGLib_Private.h:72-116).  This needs Lt even.
 */
#ifndef CLG_HIP_H
#define CLG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clg_context_s* clg_context;
typedef struct clg_field_s* clg_field;

/* ---- status codes: OK, NOT_CONVERGED (x still written, as CSLASolver::Solve returns
 *      FALSE with the best iterate, CSolverGMRES.cpp:249-259), bad argument, HIP error */
typedef enum {
    CLG_OK = 0,
    CLG_NOT_CONVERGED = 1,
    CLG_BAD_ARG = 2,
    CLG_HIP_ERROR = 3,
    CLG_UNSUPPORTED = 4,
    CLG_OUT_OF_MEMORY = 5
} clg_status;

/* ---- field kinds / precisions / subsets ---- */
typedef enum {
    CLG_WILSON = 0,   /* CFieldFermionWilsonSquareSU3: 4 spin x 3 colour per site */
    CLG_KS = 1,       /* CFieldFermionKSSU3: 3 colour per site */
    CLG_GAUGE = 2,    /* CFieldGaugeSU3: 4 links x 3x3 per site, 18 reals per link */
    CLG_GAUGE12 = 3   /* same field stored as 12 reals per link (rows 0,1; row 2 rebuilt) */
} clg_kind;

typedef enum { CLG_F64 = 0, CLG_F32 = 1, CLG_F16 = 2 } clg_prec;
/* clg_solver_params.inner_prec takes a clg_prec (CLG_F16 there: FP32 vectors with FP16 12-real
 * links) or this value: FP16 storage of the inner Krylov vectors as well as the links, FP32
 * arithmetic (Wilson GMRES, MaxDim <= 16) */
typedef enum { CLG_INNER_F16_VECTORS = 3 } clg_inner_prec;
typedef enum { CLG_FULL = 0, CLG_EVEN = 1, CLG_ODD = 2 } clg_subset;

/* EFieldOperator (CField.h:37-56), same numbering */
typedef enum {
    EFO_F_D = 0,
    EFO_F_Ddagger = 1,
    EFO_F_DD = 2,
    EFO_F_DDdagger = 3,
    EFO_F_InverseD = 4,
    EFO_F_InverseDdagger = 5,
    EFO_F_InverseDD = 6,
    EFO_F_InverseDDdagger = 7,
    EFO_F_RationalD = 8,          /* the C++ facade's ApplyOperator only (takes a rational approximation) */
    EFO_F_D_WithMass = 9,         /* the facade's ApplyOperator only: D with another mass / hopping */
    EFO_F_Ddagger_WithMass = 10,
    EFO_F_DD_WithMass = 11,
    EFO_F_DDdagger_WithMass = 12
} clg_field_operator;

/* EOperatorCoefficientType (CField.h:62-68), same numbering */
typedef enum { EOCT_None = 0, EOCT_Minus = 1, EOCT_Real = 2, EOCT_Complex = 3 } clg_coeff_type;

/* ============================== runtime ============================== */
/* replaces appInitialCLG / appQuitCLG device set-up (CLGLibManager.cpp:1142-1163) */
int clg_init(int device, clg_context* out);
int clg_finalize(clg_context ctx);
int clg_synchronize(clg_context ctx);
const char* clg_status_string(int status);
const char* clg_last_error(void);
/* writes "gfx950:256CU" style text; no reference counterpart (diagnostics) */
int clg_device_info(clg_context ctx, char* buf, int len);

/* ============================== fields ============================== */
/* replaces appCreate(FieldName) + CFieldFermionWilsonSquareSU3 / CFieldFermionKSSU3 /
 * CFieldGaugeSU3 constructors (CFieldFermionWilsonSquareSU3.cu:570-583) */
int clg_field_create(clg_context ctx, int kind, int prec, int subset, const int dims[4],
                     clg_field* out);
int clg_field_destroy(clg_field f);
/* fermion boundary signs (YAML `Period`, CLGLibManager.cpp:552-572); default 1,1,1,-1 */
int clg_field_set_bc(clg_field f, const int bc[4]);
/* host <-> device in the reference AoS order described above */
int clg_field_upload(clg_field f, const double* host);
int clg_field_download(clg_field f, double* host);
/* CField::CopyTo (CField.h:185); converts precision when dst and src differ */
int clg_field_copy(clg_field dst, clg_field src);
/* CField::Zero (CField.h:127) */
int clg_field_zero(clg_field f);
/* a half-lattice view (EVEN=1 / ODD=2) of a FULL fermion field; shares storage */
int clg_field_parity_view(clg_field full, int parity_subset, clg_field* out);
/* number of doubles clg_field_upload/download exchange */
long clg_field_host_doubles(clg_field f);
/* device bytes held by the field */
long clg_field_bytes(clg_field f);

/* ---- configuration files: EFieldFileType in the reference's order (CField.h:26-35) ---- */
typedef enum {
    CLG_FILE_BRIDGEPP_TXT = 0,       /* Bridge++ text: one real a line, x fastest, 18 per link */
    CLG_FILE_BRIDGEPP_BIN = 1,       /* the same values as big-endian doubles */
    CLG_FILE_CLG_BIN = 2,            /* the host AoS order above in Real = double (a double build's
                                      * EFFT_CLGBin) */
    CLG_FILE_CLG_BIN_COMPRESSED = 3, /* gauge only: per link the 9 doubles of A = log U (principal
                                      * logarithm; Re/Im A01, A02, A12, Im A00, A11, A22), read back as
                                      * exp(A) (CFieldGaugeSU3.cu:1225-1265, 1651-1690) */
    CLG_FILE_CLG_BIN_FLOAT = 4,      /* the host AoS order in float */
    CLG_FILE_CLG_BIN_DOUBLE = 5      /* the host AoS order in double */
} clg_file_type;
/* Host only (no device needed).  Reads a full-lattice field of `kind` (CLG_GAUGE[12], CLG_WILSON,
 * CLG_KS) on `dims` into the host AoS order.  Bridge++ files are gauge only
 * (CFieldGaugeSU3::InitialFieldWithFile, CFieldGaugeSU3.cu:1080-1185); fermion fields read the CLG
 * binary format only (CFieldFermionWilsonSquareSU3.cu:602-613), in double (CLG_FILE_CLG_BIN, _DOUBLE)
 * or float (CLG_FILE_CLG_BIN_FLOAT: a single-precision build's files).  A file whose size (or
 * number of text values) does not match the lattice is CLG_BAD_ARG. */
int clg_file_read(const char* path, int file_type, int kind, const int dims[4], double* host);
/* Host only.  Writes the host AoS order in one of the CLG binary formats (CField::SaveToFile,
 * CField.cpp:32-72; gauge fields also compressed, CFieldGaugeSU3::SaveToCompressedFile); md5_out (NULL or 33 chars) receives the upper-case hex MD5 of the bytes
 * written, as SaveToFile returns it (CLGMD5.h:566-590). */
int clg_file_write(const char* path, int file_type, int kind, const int dims[4], const double* host,
                   char md5_out[33]);
/* CLGMD5Hash (CLGMD5.h:566-590): MD5 of n bytes, upper-case hex.  The length word is the bit
 * count modulo 2^32 (CLGMD5.h:376), which equals standard MD5 below 512 MiB. */
int clg_md5(const void* data, unsigned long n, char md5_out[33]);
/* CField::InitialFieldWithFile on a FULL field: read + upload; Bridge++ gauge files are then
 * re-unitarised (SetByArray -> ElementNormalize, CFieldGaugeSU3.cu:1262-1285).  Blocking. */
int clg_field_load_file(clg_field f, const char* path, int file_type);
/* CField::SaveToFile on a FULL field: download + clg_file_write.  Blocking. */
int clg_field_save_file(clg_field f, const char* path, int file_type, char md5_out[33]);

/* ============================== synthetic inputs ============================== */
/* U = reunitarise(exp(i eps H)), H traceless Hermitian with Gaussian entries from a
 * counter-based Philox4x32-10 keyed by (seed, link); replaces
 * CFieldGaugeSU3::InitialField(EFIT_Random) (CFieldGaugeSU3.cu:32-82) */
int clg_gauge_random(clg_field gauge, double eps, unsigned seed);
int clg_gauge_unit(clg_field gauge);
/* Gaussian source, E|psi_i|^2 = 1; replaces InitialField(EFIT_RandomGaussian)
 * (CFieldFermionWilsonSquareSU3.cu:463-500) */
int clg_field_gaussian(clg_field f, unsigned seed);
/* point source (EFS_Point, CFieldFermionWilsonSquareSU3.cu:1110-1160): 1 at (site, spin,
 * colour); spin is ignored for staggered fields */
int clg_field_point_source(clg_field f, const int coord[4], int spin, int colour);

/* ============================== operators ============================== */
/* Full-lattice Wilson D, D^dagger, DD, DD^dagger with an EOperatorCoefficientType factor.
 * Replaces CFieldFermionWilsonSquareSU3::DS/DdaggerS/DDS/DDdaggerS
 * (CFieldFermionWilsonSquareSU3.cu:909-1014 -> _kernelDFermionWilsonSquareSU3 :32-163).
 * out and in must be distinct FULL Wilson fields of one precision. */
int clg_wilson_apply(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa,
                     int op, int coeff_type, double coeff_re, double coeff_im);
/* Half-lattice hopping term: out = b*x + a*H_{p,1-p} in, where out (and x) live on parity
 * p and in on the other parity; H is the Wilson hop sum_mu[(1-g_mu)U psi(x+mu) +
 * (1+g_mu)U^dag psi(x-mu)] (projectors swapped when dagger != 0).  x may be NULL.
 * The even-odd building block the reference only sketches (evenodd.tex; the compiled-out
 * path at CFieldFermionWilsonSquareSU3.cu:1266-1356). */
int clg_wilson_hop(clg_context ctx, clg_field out, clg_field in, clg_field gauge, int dagger,
                   double a_re, double a_im, clg_field x, double b_re, double b_im);
/* Even-site Schur complement Mhat = 1 - kappa^2 H_eo H_oe (op = EFO_F_D), Mhat^dagger
 * (EFO_F_Ddagger), Mhat Mhat (EFO_F_DD), Mhat Mhat^dagger (EFO_F_DDdagger), applied with
 * a scratch odd field owned by the context. */
int clg_wilson_schur(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa,
                     int op);
/* Staggered D, D^dagger, DD, DD^dagger with mass term 2am (YAML `Mass`, CField.cpp:154);
 * replaces CFieldFermionKSSU3::DS... (CFieldFermionKSSU3.cu:1344-1575 -> _kernelDFermionKS
 * :34-117) */
int clg_ks_apply(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am,
                 int op, int coeff_type, double coeff_re, double coeff_im);
/* Staggered half-lattice hop: out = b*x + a*H_{p,1-p} in, where out (and x) live on parity p and in on
 * the other parity; H = sum_mu eta_mu [U_mu(x) psi(x+mu) - U_mu(x-mu)^dag psi(x-mu)] is the hopping
 * part of D = 2am + H (_kernelDFermionKS, CFieldFermionKSSU3.cu:34-117), anti-Hermitian, so
 * D^dagger = 2am - H.  x may be NULL. */
int clg_ks_hop(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double a, clg_field x, double b);
/* The parity-p block of D D^dagger = D^dagger D = (2am)^2 - H^2 for staggered fermions:
 * out = (2am)^2 in - H_{p,1-p} H_{1-p,p} in, out and in on the same parity (EVEN or ODD); Hermitian
 * positive definite.  The staggered form of even-odd preconditioning (the reference applies
 * DD^dagger on the full lattice, CFieldFermionKSSU3.cu:1550-1575). */
int clg_ks_schur(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am);

/* ============================== BLAS ============================== */
/* CField::Axpy (complex a), AxpyPlus/AxpyMinus (a = +-1) (CField.h:127-160) */
int clg_axpy(clg_context ctx, clg_field y, double a_re, double a_im, clg_field x);
/* CField::ScalarMultply (CField.h:160-170) */
int clg_scale(clg_context ctx, clg_field y, double a_re, double a_im);
/* CField::Dot -> cuDoubleComplex, sum_x a(x)^dagger b(x) accumulated in FP64
 * (_kernelDotFermionWilsonSquareSU3, CFieldFermionWilsonSquareSU3.cu:394-411 +
 * CCudaHelper::ThreadBufferSum, CudaHelper.cu:626-710).  Blocks until the value is on
 * the host, as the reference's Dot does.  The field BLAS (axpy, scale, dot, norm2) also take two
 * 18-real FP64/FP32 gauge fields as flat arrays of link elements (CFieldGaugeSU3::Dot / Axpy). */
int clg_dot(clg_context ctx, clg_field a, clg_field b, double out[2]);
int clg_norm2(clg_context ctx, clg_field a, double* out);
/* Krylov BLAS of the solvers, FP64 accumulation, one pass over w per 16 vectors:
 * out[2k..2k+1] = <V_k, w> for k < n <= 1000 -- the Arnoldi inner products the reference takes
 * as n separate CField::Dot calls (CSolverGMRES.cpp:175-189).  Blocks. */
int clg_multidot(clg_context ctx, int n, clg_field* V, clg_field w, double* out);
/* w += sum_k c_k V_k, c_k = coeffs[2k] + i coeffs[2k+1], n <= 992 -- the solution update the
 * reference runs as n CField::Axpy calls (CSolverGMRES.cpp:222-229).  No V_k may be w. */
int clg_multiaxpy(clg_context ctx, clg_field w, int n, clg_field* V, const double* coeffs);
/* clg_multidot that allocates nothing.  No counterpart in the reference: for callers that cannot
 * call hipMalloc (code under HIP graph capture, serving loops, harnesses that forbid allocation).
 * out (DEVICE memory, 2n doubles): out[2k..2k+1] = <V_k, w>, k < n <= 64, written on the context's
 * stream without blocking (the same kernels and sums as clg_multidot).  The partial sums live in
 * a static 2,162,688-byte device array, one per device, whose address clg_init looks up; every call
 * on a device shares it, so calls that overlap on one device (other streams or threads) race. */
int clg_multidot_noalloc(clg_context ctx, int n, clg_field* V, clg_field w, double* out);
/* Gram matrix out[2 (i c + j) .. +1] = <A_i, B_j> (host, interleaved complex) of FP64 fermion fields of one
 * shape, 1 <= r, c <= 32, on the matrix cores (v_mfma_f64_16x16x4f64, FP64 sums in a fixed order).  No
 * single counterpart in the reference: GCRO-DR forms these matrices with one CField::Dot per entry
 * (CSolverGCRODR.cpp:700-760); the solver here uses the same kernels on device data.  Blocks. */
int clg_gram_matrix(clg_context ctx, int r, clg_field* A, int c, clg_field* B, double* out);
/* CFieldMatrixOperation::VectorMultiplyMatrix (CField.h:372-380; _kernelMatrixMultiply,
 * CFieldFermionWilsonSquareSU3.cu:1533-1560): with W = (res_0..res_{nres-1}, left_0..left_{nleft-1}),
 * res_i <- sum_j W_j M[j*nres + i], M a host (nres+nleft) x nres complex matrix (interleaved
 * re/im, row j = input j).  FP64 fields of one shape, nres <= 31, nres + nleft <= 32. */
int clg_vector_multiply_matrix(clg_context ctx, int nres, clg_field* res, int nleft, clg_field* left,
                               const double* matrix);

/* ============================== solvers ============================== */
typedef enum { CLG_SOLVER_GMRES = 0, CLG_SOLVER_BICGSTAB = 1 } clg_solver_kind;

/* stop rules: REFERENCE = the reference solver's own criterion (GMRES: |g_m|^2 < Accuracy *
 * (AbsoluteAccuracy ? 1 : |b|^2) at cycle end, CSolverGMRES.cpp:216-232; BiCGStab:
 * |s|^2/|b|^2 < Accuracy, CSolverBiCGstab.cpp:466-480); TRUE_RESIDUAL = |b - A x|/|b| <
 * accuracy computed in FP64 */
typedef enum { CLG_STOP_REFERENCE = 0, CLG_STOP_TRUE_RESIDUAL = 1 } clg_stop_rule;

typedef struct {
    int solver;            /* clg_solver_kind */
    int max_dim;           /* GMRES MaxDim (Krylov dimension per cycle), <= 64 */
    int restart;           /* GMRES Restart (max cycles) / BiCGStab Restart */
    int max_step;          /* BiCGStab MaxStep (steps per restart) */
    int deviation_step;    /* BiCGStab DiviationStep */
    double accuracy;       /* YAML Accuracy (squared-norm units for STOP_REFERENCE) */
    int absolute_accuracy; /* YAML AbsoluteAccuracy */
    int stop_rule;         /* clg_stop_rule */
    int even_odd;          /* 1: even-odd preconditioning.  Wilson: solve through the even-site Schur
                              complement (2: its normal equations).  Staggered: D D^dagger is solved
                              as its two parity blocks (2am)^2 - H H, each from a zero guess; D and
                              D^dagger solves become x = D^dagger y or D y with D D^dagger y = b */
    int inner_prec;        /* CLG_F64: plain solve; CLG_F32 / CLG_F16 (FP32 vectors, FP16 12-real links;
                              Wilson and staggered) / CLG_INNER_F16_VECTORS (Wilson GMRES): defect
                              correction */
    double inner_accuracy; /* relative residual target of each inner solve (e.g. 1e-5) */
    int max_outer;         /* defect-correction outer iterations (mixed precision) */
    int kind;              /* CLG_WILSON or CLG_KS */
} clg_solver_params;

typedef struct {
    int converged;
    int cycles;            /* GMRES cycles (or BiCGStab iterations) summed over inner solves */
    int outer;             /* defect-correction iterations (1 for a plain solve) */
    int operator_applies;  /* operator applications, counted in the solve precision */
    double true_residual;  /* |b - A x| / |b| in FP64 at exit */
    double last_deviation; /* the reference rule's last deviation */
    double seconds;        /* wall time of the solve (device synchronised) */
} clg_solve_report;

/* CSLASolver::Solve (CSLASolver.h:42-50): x = op^{-1} b, op in {EFO_F_D, EFO_F_Ddagger,
 * EFO_F_DD, EFO_F_DDdagger} of the Wilson (kappa) or staggered (2am) fermion; start may
 * be NULL (then x0 = b as CSolverGMRES.cpp:140-147).  x may alias b. */
int clg_solve(clg_context ctx, const clg_solver_params* params, clg_field x, clg_field b,
              clg_field gauge, double kappa_or_mass, int op, clg_field start,
              clg_solve_report* report);
/* defaults matching the reference YAML defaults (CSolverGMRES.cpp:35-99) */
int clg_solver_params_default(clg_solver_params* p);

/* ---- GCRO-DR with a recycled subspace (CSLASolverGCRODR, CSolverGCRODR.h:24-140, .cpp:67-700) ---- */
typedef struct clg_gcrodr_s* clg_gcrodr;
/* DeflateSpace (EEigenDeflationType, CSolverGCRODR.h:16-20) */
typedef enum { CLG_DEFLATE_REV = 0, CLG_DEFLATE_HEV = 1, CLG_DEFLATE_SVD = 2 } clg_deflation;
/* ESolverPhase (CCommonData.h:168-172); In/End trajectory reuse the recycled subspace */
typedef enum {
    CLG_PHASE_START_TRAJECTORY = 0,
    CLG_PHASE_IN_TRAJECTORY = 1,
    CLG_PHASE_END_TRAJECTORY = 2,
    CLG_PHASE_ONCE = 3
} clg_solver_phase;
typedef struct {
    int m_dim;          /* MDim, 5..31 */
    int k_dim;          /* KDim, 2..MDim-2 (recycled vectors) */
    int restart;        /* Restart: GCRO-DR cycles after the first stage */
    int recalculate_r;  /* RecalculateR: recompute b - A x every n cycles (0: recurrence only) */
    double accuracy;    /* Accuracy: stop when |r| < accuracy * (absolute ? 1 : |b|^2) */
    int absolute_accuracy;
    int deflate;        /* clg_deflation */
    int kind;           /* CLG_WILSON or CLG_KS */
    int even_odd;       /* Wilson D / D^dag only: run on the even-site Schur complement */
} clg_gcrodr_params;
/* defaults of CSolverGCRODR.cpp:67-118 (MDim 10, KDim MDim-2, SVD deflation) */
int clg_gcrodr_params_default(clg_gcrodr_params* p);
int clg_gcrodr_create(clg_context ctx, const clg_gcrodr_params* params, clg_gcrodr* out);
/* destroy a solver before clg_finalize of its context; clg_finalize destroys any solver (and graph)
 * still alive on the context, after which their handles are invalid */
int clg_gcrodr_destroy(clg_gcrodr solver);
/* forget the recycled subspace */
int clg_gcrodr_reset(clg_gcrodr solver);
/* CSLASolverGCRODR::Solve: x = A^-1 b with A = op (EFO_F_D .. EFO_F_DDdagger) of the params.kind
 * fermion on FULL FP64 fields; phase as ESolverPhase; start may be NULL (x0 = b).  Returns
 * CLG_NOT_CONVERGED with the last iterate after Restart cycles. */
int clg_gcrodr_solve(clg_gcrodr solver, clg_field x, clg_field b, clg_field gauge, double kappa_or_mass, int op,
                     int phase, clg_field start, clg_solve_report* report);

/* ---- multi-shift solvers and the rational approximation (RHMC) ---- */
typedef enum {
    CLG_MS_FOM = 0,      /* CMultiShiftFOM */
    CLG_MS_BICGSTAB = 1, /* CMultiShiftBiCGStab */
    CLG_MS_GMRES = 2,    /* CMultiShiftGMRES (CMultiShiftGMRES.cpp:112-330), at most 31 shifts */
    CLG_MS_NESTED = 3    /* CMultiShiftNested (CMultiShiftNested.cpp:69-150): staggered, real shifts */
} clg_multishift_kind;

/* CMultiShiftSolver::Solve (CMultiShiftSolver.h:38-47): xs[k] = (A + c_k)^-1 b for all k, with
 * A = op of the params.kind fermion and c_k = shifts[2k] + i shifts[2k+1].  CLG_MS_FOM follows
 * CMultiShiftFOM.cpp:97-270 (params max_dim <= 60, restart, accuracy, absolute_accuracy), CLG_MS_BICGSTAB
 * follows CMultiShiftBiCGStab.cpp:73-270 (max_step, deviation_step, accuracy, absolute_accuracy).
 * prec CLG_F32 runs the whole solve on FP32 copies (solutions written back as FP64); CLG_F16 the same
 * with FP16 12-real copies of the links (not for CLG_MS_NESTED).  The xs must
 * not alias b.  Returns CLG_NOT_CONVERGED with the iterates written when the budget runs out. */
int clg_multishift_solve(clg_context ctx, const clg_solver_params* params, int kind, int prec, clg_field* xs,
                         const double* shifts, int nshift, clg_field b, clg_field gauge, double kappa_or_mass,
                         int op, clg_solve_report* report);
/* CFieldFermion::RationalApproximation (CField.cpp:114-148): out = c in + sum_k a_k (A + b_k)^-1 in,
 * coeffs = [c, a_1..a_n, b_1..b_n] (CRationalApproximation.h:38-117), shift sums in FP64.
 * out may alias in. */
int clg_rational_apply(clg_context ctx, const clg_solver_params* params, int kind, int prec, clg_field out,
                       clg_field in, clg_field gauge, double kappa_or_mass, int op, const double* coeffs, int degree,
                       clg_solve_report* report);

/* ============================== HMC forces and energies ============================== */
/* Forces live in FP64 CLG_GAUGE-shaped fields (traceless anti-Hermitian per link) and are
 * accumulated (F += ...), as the reference's force kernels add into pForce. */
/* _kernelDWilsonForceSU3 (CFieldFermionWilsonSquareSU3.cu:170-231): inv_dddag = (DD^dag)^-1 phi,
 * inv_d = D^dag inv_dddag */
int clg_wilson_force_kernel(clg_context ctx, clg_field force, clg_field gauge, clg_field inv_dddag, clg_field inv_d,
                            double kappa);
/* _kernelDFermionKSForce (CFieldFermionKSSU3.cu:216-264) over npole (phi_i, D0 phi_i) pairs */
int clg_ks_force_kernel(clg_context ctx, clg_field force, clg_field gauge, const clg_field* phis, const clg_field* dphis,
                        const double* numerators, int npole);
/* CFieldFermionWilsonSquareSU3::CalculateForceS (CFieldFermionWilsonSquareSU3.cu:1016-1100); start
 * (nullable) warm-starts the DD^dag solve as the reference's CFieldCache does; x_out (nullable)
 * receives (DD^dag)^-1 phi */
int clg_wilson_force(clg_context ctx, const clg_solver_params* params, clg_field force, clg_field phi, clg_field gauge,
                     double kappa, clg_field start, clg_field x_out, clg_solve_report* report);
/* CFieldFermionKSSU3::CalculateForceS (CFieldFermionKSSU3.cu:412-497), MD coefficients
 * [c, a_1..a_n, b_1..b_n] */
int clg_ks_force(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field force, clg_field phi,
                 clg_field gauge, double two_am, const double* coeffs, int degree, clg_solve_report* report);
/* CActionFermionWilsonNf2::Energy (CActionFermionWilsonNf2.cpp:53-68): |D^-1 phi|^2 */
int clg_wilson_pf_energy(clg_context ctx, const clg_solver_params* params, clg_field phi, clg_field gauge, double kappa,
                         double* energy, clg_solve_report* report);
/* CActionFermionKS::Energy (CActionFermionKS.cpp:52-71): Re phi^dag r_MD(DD^dag) phi */
int clg_ks_pf_energy(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field phi,
                     clg_field gauge, double two_am, const double* coeffs, int degree, double* energy,
                     clg_solve_report* report);

/* ============================== HMC ============================== */
/* Gauge, momentum and force fields are FP64 18-real CLG_GAUGE fields.  Momenta are traceless
 * anti-Hermitian; the kinetic energy is sum ReTr P^dag P; P += eps F; U <- exp(eps P) U. */
#define CLG_HMC_MAX_ACTIONS 8

/* CFieldGaugeSU3::MakeRandomGenerator (CFieldGaugeSU3.cu:32-82 with EFIT_RandomGenerator,
 * SU3.h:193-226): 8 Gaussian generator weights of variance 1/4 per link, drawn from a Philox
 * stream keyed by (seed, counter, link) */
int clg_gauge_momentum_random(clg_field momentum, unsigned seed, unsigned counter);
/* CFieldGaugeSU3::CalculateKinematicEnergy (CFieldGaugeSU3.cu:492-516): sum ReTr P^dag P, FP64 */
int clg_gauge_kinetic_energy(clg_context ctx, clg_field momentum, double* out);
/* CFieldGaugeSU3::Axpy with a real coefficient (AddForce, CIntegrator.h:101-128): y += a x */
int clg_gauge_axpy(clg_context ctx, clg_field y, double a, clg_field x);
/* CFieldGaugeSU3::ExpMult (CFieldGaugeSU3.cu:452-490, 1457-1477): U <- exp(a P) U per link;
 * exp_precision 0 = SU3::QuickExp (SU3.h:1361-1640), n > 0 = order-n Taylor (SU3.h:1286-1318),
 * as the YAML key ExponentialPrecision selects (CLGLibManager.cpp:282-289) */
int clg_gauge_exp_mult(clg_context ctx, clg_field momentum, clg_field gauge, double a, int exp_precision);
/* CFieldGaugeSU3::ElementNormalize (CFieldGaugeSU3.cu:518-528, SU3::Norm) */
int clg_gauge_normalize(clg_context ctx, clg_field gauge);
/* CGaugeFixingRandom (CGaugeFixingRandom.cu; used by CLGTest's TestGaugeInvarience,
 * TestCommon.cpp:206-239): a random local gauge transformation g(x) = reunitarise(exp(i eps H(x))),
 * one SU(3) per site from its own Philox stream, held in the mu = 0 links of the FP64 gauge field gt */
int clg_gauge_transform_random(clg_field gt, double eps, unsigned seed);
/* GaugeFixing / AlsoFixingFermion of that transformation: a gauge field (any format) takes
 * U_mu(x) -> g(x) U_mu(x) g(x+mu)^dagger, a Wilson or staggered field (any subset, precision)
 * psi(x) -> g(x) psi(x) */
int clg_gauge_transform_apply(clg_field gt, clg_field target);
/* CActionGaugePlaquette::Energy (CActionGaugePlaquette.cpp:14-107, CFieldGaugeSU3.cu:308-371):
 * beta/3 sum_{x, mu<nu} (3 - ReTr U_mu nu(x)), FP64 */
int clg_gauge_plaquette_energy(clg_context ctx, clg_field gauge, double beta, double* out);
/* CActionGaugePlaquette::CalculateForce (CFieldGaugeSU3.cu:194-257): F += -(beta/6) Ta[U staple] */
int clg_gauge_force(clg_context ctx, clg_field force, clg_field gauge, double beta);
/* CFieldFermionWilsonSquareSU3::PrepareForHMCS (CFieldFermionWilsonSquareSU3.cu:867-906):
 * phi = D eta, eta Gaussian with E|eta_i|^2 = 1 */
int clg_wilson_prepare_hmc(clg_context ctx, clg_field phi, clg_field gauge, double kappa, unsigned seed);
/* CFieldFermionKSSU3::PrepareForHMC (CFieldFermionKSSU3.cu:391-407): phi = r_MC(DD^dag) eta */
int clg_ks_prepare_hmc(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field phi,
                       clg_field gauge, double two_am, const double* mc, int degree, unsigned seed,
                       clg_solve_report* report);

/* CAction subclasses of a trajectory */
typedef enum {
    CLG_ACTION_GAUGE_PLAQUETTE = 0,    /* CActionGaugePlaquette (Beta) */
    CLG_ACTION_FERMION_WILSON_NF2 = 1, /* CActionFermionWilsonNf2 (Hopping) */
    CLG_ACTION_FERMION_KS = 2          /* CActionFermionKS (Mass = 2am, MC, MD) */
} clg_action_kind;
typedef struct {
    int kind;                 /* clg_action_kind */
    double beta;              /* plaquette Beta */
    double kappa_or_mass;     /* Wilson Hopping or staggered Mass (= 2am) */
    clg_field phi;            /* pseudofermion: FULL FP64 field of the fermion kind (its bc apply) */
    const double* mc;         /* KS: MC rational [c, a_1..a_n, b_1..b_n] (heat-bath) */
    int mc_degree;
    const double* md;         /* KS: MD rational (force and energy) */
    int md_degree;
    clg_solver_params solver; /* Wilson: the DD^dag / D solves; KS: the multi-shift solve */
    double energy_accuracy;   /* accuracy of the energy solves; 0 = solver.accuracy */
    int ms_kind;              /* KS: clg_multishift_kind */
    int ms_prec;              /* KS: CLG_F64, or CLG_F32 multi-shift (FP64 shift sums) in the force and
                                 heat-bath solves, or CLG_F16: force solves with FP32 vectors and FP16
                                 12-real links, heat bath in CLG_F32; the energy solves are always FP64 */
} clg_action;

/* IntegratorType */
typedef enum {
    CLG_INT_LEAPFROG = 0,              /* CIntegratorLeapFrog (CIntegratorLeapFrog.h) */
    CLG_INT_NESTED_FORCE_GRADIENT = 1  /* CIntegratorNestedForceGradient (CIntegratorNestedForceGradient.cpp:18-157) */
} clg_integrator_kind;
typedef struct {
    int integrator;      /* clg_integrator_kind */
    double step_length;  /* IntegratorStepLength: trajectory length (CIntegrator.cpp:63-66) */
    int steps;           /* IntegratorStep */
    int nested_steps;    /* NestedStep (CIntegrator.cpp:281-297) */
    int inner_leapfrog;  /* InnerLeapfrog: nested leapfrog inside the force-gradient steps */
    int exp_precision;   /* ExponentialPrecision: 0 quick exponential, n > 0 Taylor order n */
    int metropolis;      /* Metropolis: 1 accept on exp(h_diff), 0 accept every trajectory (h_diff is
                            reported either way) */
} clg_hmc_params;
typedef struct {
    double h_before, h_after; /* kinetic + every action, FP64 (CIntegrator::GetEnergy) */
    double h_diff;            /* h_before - h_after (CHMC.cpp:118-124) */
    double kinetic_before, kinetic_after;
    double action_before[CLG_HMC_MAX_ACTIONS], action_after[CLG_HMC_MAX_ACTIONS];
    int accepted;
    int force_evaluations;    /* force passes (gauge or fermion) */
    int solver_iterations;    /* solver cycles summed over every solve of the trajectory */
    int unconverged_solves;
    double seconds;
} clg_hmc_report;

/* defaults: nested force gradient, tau 1, 8 steps, 3 nested, quick exponential, Metropolis on */
int clg_hmc_params_default(clg_hmc_params* p);
/* CIntegrator::Evaluate between two GetEnergy calls (CHMC.cpp:100-108) with the caller's momentum
 * and pseudofermions (no heat-bath); gauge and momentum are updated in place and the gauge field
 * is re-unitarised at the end (FinishEvaluate, CIntegrator.cpp:244-254). */
int clg_hmc_evolve(clg_context ctx, clg_field gauge, clg_field momentum, const clg_hmc_params* params,
                   const clg_action* actions, int nactions, clg_hmc_report* report);
/* CHMC::Update for one trajectory (CHMC.cpp:76-219): pseudofermion heat-bath of every fermion
 * action, momentum noise, H before, integrator, H after, Metropolis accept on exp(h_diff); a
 * rejected trajectory restores the gauge field.  All randomness derives from seed. */
int clg_hmc_trajectory(clg_context ctx, clg_field gauge, const clg_hmc_params* params, clg_action* actions,
                       int nactions, unsigned long long seed, clg_hmc_report* report);

/* ============================== batched operators and solves ============================== */
/* Batched Wilson hop: out_i = a_i H in_i + b x_i for n <= 16 slots in one launch (slots may live on
 * lattices of different extents; x may be NULL).  Same parity/layout rules as clg_wilson_hop for
 * half fields; FULL fields give the hop over both parities.  FP64 spinors with FP64 links, FP32
 * spinors with FP32 18-real or FP16 12-real links, or FP16-storage spinors with FP16 12-real links.
 * tile = right-hand sides per thread sharing one link load (1, 2, 4 (not FP64)); 0 lets the
 * autotuner choose.  The reference applies D to one
 * field per launch (CFieldFermionWilsonSquareSU3.cu:909-1014); this is its batched form. */
int clg_wilson_hop_batch(clg_context ctx, int n, clg_field* out, clg_field* in, clg_field* gauge, int dagger,
                         const double* a, clg_field* x, double b, int tile);
/* Batched solve of n <= 16 Wilson systems D x_i = b_i (op EFO_F_InverseD) or D^dag x_i = b_i
 * (EFO_F_InverseDdagger), each on its own FP64 gauge field (lattices may differ), one kappa.  Even-odd
 * Schur GMRES(MaxDim <= 16) per system in lock step, inner precision params.inner_prec inside FP64
 * defect correction (CLG_F64 = plain), true-residual stop rule per system.  Replaces the reference's
 * per-source loop of CSLASolver::Solve over GetSourcesAtSiteFromPool
 * (CFieldFermionWilsonSquareSU3.cu:1222-1255).  reports (nullable) holds n entries. */
int clg_solve_batch(clg_context ctx, const clg_solver_params* params, int n, clg_field* x, clg_field* b,
                    clg_field* gauge, double kappa, int op, clg_solve_report* reports);

/* ============================== launch autotuning ============================== */
/* The Wilson hop launches (single and batched) time their candidate launch shapes (pin mode,
 * workgroup size, XCD remap; batched: right-hand sides per thread) on the first call of each
 * operator shape and keep the fastest.  The reference fixes one decomposition at start-up
 * (_HC_DecompX.., CCommonData.h:72-91); these calls inspect and steer the choice. */
/* mode 1 tune, 0 use defaults, -1 follow the CLG_TUNE environment variable (default on) */
int clg_autotune_enable(int mode);
/* fix every single-system hop launch: pin_mode 1|2 (register pinning per half / per direction), 3|4 (the
 * same with non-temporal link loads), 5..8 (as 1..4 with the z and t neighbours staged in LDS; needs
 * block = L2*L3/2 with L3 even, and a launch that does not fit is refused), block 64|128|256, xcd 0
 * (hardware workgroup order) | 8 (each XCD sweeps a contiguous slab of sites; full D interleaves the
 * parities block by block) | 1 (XCD slabs, parities one after the other) | 10*TY + m with TY in 2, 4, 8
 * (staged modes only: the (x,y) planes cut into 8 tiles, 8/TY along x and TY along y, XCD k sweeping
 * tile k x-major (m = 0) or y-major (m = 1)), sites_per_thread 1|2|4 (0 = 1: consecutive blocks of
 * sites per workgroup); pin_mode 0 clears */
int clg_autotune_force(int pin_mode, int block, int xcd, int sites_per_thread);
/* Tuned configurations persist across processes when the environment variable CLG_TUNE_CACHE names a
 * file: every tuned key is appended as one line "<device arch> <tuner> <key> : <configuration>", and
 * the tuners take the lines of the current device architecture from the file before timing a key
 * (again after clg_autotune_reset).  Lines that do not parse, or configurations no launch accepts,
 * are skipped. */
/* The staggered D tunes its own launch (block 64|128|256, hardware or XCD-slab order, pinning);
 * the environment variable CLG_TUNE_FORCE_KS="block,xcd,pin" (xcd 0|8, pin 0|2) fixes it. */
/* forget every tuned shape (the next call of each shape tunes again) */
int clg_autotune_reset(void);
/* number of tuned shapes, and a one-line description of entry `index` */
int clg_autotune_count(void);
int clg_autotune_entry(int index, char* buf, int len);

/* ============================== graphs ============================== */
/* Record the launches of non-blocking entry points made on a context between begin and end
 * (operators, hops, batched hops, axpy/scale, copies, zero, gauge updates and forces), then
 * replay them with one launch.  Entry points that read results back to the host or upload host
 * data (dot, norm2, multi-dot, solves, energies, HMC) fail with CLG_BAD_ARG while recording; the
 * recording stays open (clg_dot, clg_norm2, clg_multidot and clg_solve refuse before launching
 * anything, the others may leave launches of their own in it).  A recorded operator keeps the launch shape chosen before recording (tune
 * outside first).  Fields used by a graph must outlive it.  No reference counterpart: the
 * reference launches every kernel from the host (CSolverGMRES.cpp:150-247). */
typedef struct clg_graph_s* clg_graph;
int clg_graph_begin(clg_context ctx);
int clg_graph_end(clg_context ctx, clg_graph* out);
/* enqueue the recorded work on the context stream; returns without waiting (clg_synchronize) */
int clg_graph_launch(clg_graph g);
/* recorded operations (graph nodes) */
int clg_graph_nodes(clg_graph g);
int clg_graph_destroy(clg_graph g);

/* ============================== measurement ============================== */
/* Times `reps` back-to-back launches of one Wilson operator on the context stream with HIP
 * events; writes the average milliseconds per launch.  variant: 0 full-lattice D,
 * 1 half-lattice hop, 2 Schur operator.  No reference counterpart (bench support). */
int clg_time_wilson(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa,
                    int variant, int reps, double* ms_avg);
/* The same for the staggered operator with mass term two_am: variant 0 the full-lattice D
 * (clg_ks_apply, EFO_F_D), 1 the parity Schur block (clg_ks_schur, half fields).  No reference
 * counterpart (bench support). */
int clg_time_ks(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am, int variant,
                int reps, double* ms_avg);

#ifdef __cplusplus
}
#endif
#endif
