/*
 * clg_oracle.h -- CPU oracle for the fermion-matrix hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++ restatement of the
 * reference (CudaLatticeGauge / CLGLib) algorithms, used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py as the
 * checker.  The product (cudalatticegauge_amd/, libclg_hip.so) never links,
 * loads or calls it.
 *
 * Host data formats (the reference's own AoS order, FP64 interleaved re/im):
 *   site   = ((x*Ly + y)*Lz + z)*Lt + t              (CCommonData.h:274-282)
 *   Wilson spinor  psi[site][spin 4][colour 3]  -> 24 doubles per site
 *   KS vector      chi[site][colour 3]          ->  6 doubles per site
 *   SU(3) link     U[site][mu 4][row 3][col 3]  -> 72 doubles per site
 *
 * Every function returns 0 on success and a negative number on bad input.
 * Functions whose results no reference-held artefact pins are pinned by the
 * invariant tests in tests/test_oracle_*.py (plane waves, gamma5-hermiticity,
 * gauge covariance, dense-matrix brute force, finite differences of the actions).
 */
#ifndef CLG_ORACLE_H
#define CLG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- index, boundary, eta, even-odd map ---------------- */
int orc_site_index(const int dims[4], int x, int y, int z, int t);
/* move tables (CIndexSquare.cu:212-255 + CBoundaryConditionTorusSquare.cu:19-64):
 * fwd_site/bwd_site[(site*4+mu)] neighbour site, fwd_sign/bwd_sign = +1/-1 */
int orc_build_moves(const int dims[4], const int bc[4], int* fwd_site, int* bwd_site,
                    int* fwd_sign, int* bwd_sign);
/* eta bits (CIndexSquare.cu:257-267): bit mu set <=> sum_{nu<mu} x_nu odd */
int orc_build_eta(const int dims[4], unsigned char* eta);
/* even/odd index map, reference convention intokernal_even/odd
 * (CLGLib_Private.h:72-116): out[e] = lexicographic site of the e-th site of parity p */
int orc_eo_sites(const int dims[4], int parity, int* out);

/* gamma table, chiral set (GammaMatrix.h:447-475): (G psi)[r] = i^z4[r] psi[idx[r]]
 * which: 0..3 = gamma1..gamma4, 4 = gamma5 */
int orc_gamma(int which, int idx[4], int z4[4]);

/* ---------------- synthetic inputs (counter-based Philox4x32-10) ---------------- */
int orc_philox(const unsigned ctr[4], const unsigned key[2], unsigned out[4]);
/* U = reunitarise(exp(i*eps*H)), H traceless Hermitian Gaussian, keyed by (seed, link) */
int orc_gauge_random(const int dims[4], double eps, unsigned seed, double* U);
int orc_gauge_unit(const int dims[4], double* U);
/* Gaussian complex source, each real part ~ N(0,1/2) (so E|psi_i|^2 = 1) */
int orc_wilson_gaussian(const int dims[4], unsigned seed, double* psi);
int orc_ks_gaussian(const int dims[4], unsigned seed, double* chi);
/* reference SU3::Norm re-unitarisation (SU3.h:1058-1088), in place on n matrices */
int orc_su3_norm(double* m, long n);

/* ---------------- operators (reference restatement, full lattice) ---------------- */
/* EFieldOperator order (CField.h:37-56) */
enum { ORC_D = 0, ORC_DDAG = 1, ORC_DD = 2, ORC_DDDAG = 3 };
/* coeff_type: 0 none, 1 minus, 2 real (coeff_re), 3 complex (coeff_re, coeff_im) */
int orc_wilson_op(const int dims[4], const int bc[4], double kappa, const double* U,
                  const double* in, double* out, int op, int coeff_type,
                  double coeff_re, double coeff_im);
int orc_wilson_op_f32(const int dims[4], const int bc[4], float kappa, const float* U,
                      const float* in, float* out, int op);
int orc_ks_op(const int dims[4], const int bc[4], double two_am, const double* U,
              const double* in, double* out, int op, int coeff_type,
              double coeff_re, double coeff_im);
int orc_ks_op_f32(const int dims[4], const int bc[4], float two_am, const float* U,
                  const float* in, float* out, int op);

/* ---------------- BLAS / dot ---------------- */
/* sum_i conj(a_i) b_i over n complex numbers (long double accumulation) */
int orc_dot(const double* a, const double* b, long n, double out[2]);

/* ---------------- solvers ---------------- */
/* stop_rule 0: reference GMRES rule |g_m|^2 < accuracy * blength, checked at the end of
 *              a cycle, x0 = b or start (CSolverGMRES.cpp:111-260)
 * stop_rule 1: true relative residual |b - A x| / |b| < accuracy, checked each cycle */
typedef struct {
    int kind;          /* 0 Wilson, 1 KS */
    int op;            /* ORC_D .. ORC_DDDAG */
    int dims[4];
    int bc[4];
    double mass_or_kappa;
} orc_operator;

int orc_gmres(const orc_operator* A, const double* U, const double* b, double* x,
              const double* start, int max_dim, int restart, double accuracy,
              double blength, int stop_rule, int* cycles_out, double* last_dev_out);
int orc_bicgstab(const orc_operator* A, const double* U, const double* b, double* x,
                 const double* start, int step_count, int deviation_check, int retry,
                 double accuracy, double small_rho, int absolute, int* iters_out,
                 double* last_dev_out);
/* multi-shift FOM (CMultiShiftFOM.cpp:97-270): x_n = (A + c_n)^-1 b */
int orc_msfom(const orc_operator* A, const double* U, const double* b, double* xs,
              const double* shifts, int nshift, int max_dim, int restart,
              double accuracy, double blength, int* cycles_out, double* max_err_out);

/* GCRO-DR with SVD deflation (CSolverGCRODR.cpp:196-700): x0 = b, urec holds the k recycled vectors
 * (k * 2n doubles) between calls with *have_u set after the first; phase as ESolverPhase (1, 2 reuse);
 * stop |r| < accuracy * (absolute ? 1 : |b|^2).  Returns 0 converged, 1 not converged. */
int orc_gcrodr(const orc_operator* A, const double* U, const double* b, double* x, double* urec, int* have_u,
               int m, int k, int restart, int recalc, double accuracy, int absolute, int phase, int* cycles_out,
               double* dev_out);

/* ---------------- forces ---------------- */
/* Wilson force kernel (CFieldFermionWilsonSquareSU3.cu:170-231): F += ... */
int orc_wilson_force_kernel(const int dims[4], const int bc[4], double kappa, const double* U,
                            const double* inverse_d, const double* inverse_dddag, double* F);
/* KS force kernel (CFieldFermionKSSU3.cu:216-264) over npole (phi_i, D phi_i) pairs */
int orc_ks_force_kernel(const int dims[4], const int bc[4], const double* U,
                        const double* phis, const double* dphis, const double* numer,
                        int npole, double* F);

/* ---------------- gauge action ---------------- */
/* sum over sites and mu<nu of (3 - ReTr plaquette), periodic gauge links; S_g = beta/3 * this */
int orc_plaquette_sum(const int dims[4], const double* U, double* out);
/* plaquette-action force F += -(beta/6) Ta[U C] (CFieldGaugeSU3.cu:194-257) */
int orc_gauge_force(const int dims[4], const double* U, double beta, double* F);

/* ---------------- HMC ---------------- */
/* out = exp(a M) for one traceless anti-Hermitian 3x3 M: prec 0 = SU3::QuickExp
 * (SU3.h:1361-1640), prec n > 0 = order-n Taylor polynomial SU3::ExpReal (SU3.h:1286-1318) */
int orc_su3_exp(const double* M, double a, int prec, double* out);

/* one action of the trajectory (CAction subclasses) */
typedef struct {
    int kind;              /* 0 CActionGaugePlaquette, 1 CActionFermionWilsonNf2, 2 CActionFermionKS */
    double beta;           /* plaquette action */
    double kappa_or_mass;  /* Wilson hopping / KS Mass (= 2am) */
    int bc[4];             /* fermion boundary signs */
    const double* phi;     /* pseudofermion (AoS) */
    const double* md;      /* KS MD rational [c, a_1..a_n, b_1..b_n] */
    int md_degree;
    double accuracy;       /* oracle solve accuracy (relative) */
} orc_action;
enum { ORC_INT_LEAPFROG = 0, ORC_INT_NESTED_FORCE_GRADIENT = 1, ORC_INT_NESTED_FG_INNER_LEAPFROG = 2 };
/* CIntegrator::Evaluate with the given momentum P and pseudofermions (no heat-bath); U and P
 * are updated in place, U re-unitarised at the end; h_before/h_after = kinetic + actions */
int orc_hmc_evolve(const int dims[4], double* U, double* P, const orc_action* acts, int nact, int integrator,
                   double tau, int steps, int nested, int exp_prec, double* h_before, double* h_after);

/* ---------------- support ---------------- */
int orc_num_threads(void);
/* wall-clock seconds of `reps` Wilson D applications, FP64 (prec 0) or FP32 (prec 1) */
double orc_time_wilson_d(const int dims[4], int prec, int reps, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
