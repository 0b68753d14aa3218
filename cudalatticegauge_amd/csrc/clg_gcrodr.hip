// clg_gcrodr.hip -- GCRO-DR: restarted GMRES with deflation by a recycled subspace (Parks, de Sturler,
// Mackey, Johnson, Maiti 2006), as CSLASolverGCRODR runs it (CSolverGCRODR.cpp:196-398).
//
// State: k recycled vectors U (kept between solves, as the reference keeps m_lstU) and, within a
// solve, C = A U with C^H C = I, the Krylov vectors V and the (m+1) x m matrix G = [[D, B], [0, H]].
// Per solve:
//   * with a recycled U and phase In/End trajectory: C = A U, thin QR, U <- U R^-1 (GenerateCUFirstTime);
//     otherwise one GMRES(m) cycle, harmonic Ritz vectors P_k, U = V_m P_k R^-1, C = V_{m+1} Q with
//     H P_k = Q R (FirstTimeGMERESSolve / FindPk1 / GenerateCU);
//   * cycles: x += U C^H r, r -= C C^H r, Arnoldi on (1 - C C^H) A from r/|r|, least squares in
//     G, new P_k from the generalized eigenproblem of the cycle (FindPk2), U and C regenerated.
// The deflation space (DeflateSpace) is REV (Ritz), HEV (harmonic) or SVD (default) as in FindPk1 /
// FindPk2.  Stop rule: |r| < Accuracy * (AbsoluteAccuracy ? 1 : |b|^2), checked after the first
// stage and after every cycle, as the reference checks it.
// The MI355X form: the Arnoldi step is classical Gram-Schmidt with a conditional second pass (one
// multi-dot and one multi-axpy launch over C and V together), the G column stays on the device,
// and the tall-skinny products U <- [U V] P R^-1, C <- [C V] Q run as one in-place kernel each
// (the reference's _kernelMatrixMultiply plus FieldSolveY's k^2 axpys, CFieldFermionWilsonSquareSU3.cu:1525-1606).
// The small dense problems (least squares, QR, eigenproblems; m <= 31) run in FP64 on the host.
#include "clg_internal.h"
#include "clg_smallla_dev.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int blas_dot_dev(clg_context, const clg_field_s*, const clg_field_s*, double*);
int blas_multidot_dev(clg_context, clg_field_s* const*, int, const clg_field_s*, double*, double*, const int*);
int blas_multiaxpy_dev(clg_context, clg_field_s*, clg_field_s* const*, int, const double*, double, double*, const int*,
                       const double*, double, int*);
int fetch_doubles(clg_context, const double*, int, double*);
int workspace(clg_context, size_t, double**);
int blas_vmm_dev(clg_context, clg_field_s* const*, int, clg_field_s* const*, int, const double*);
int linop_apply(clg_context, int, int, const clg_field_s*, double, clg_field_s*, const clg_field_s*);
int wilson_hop_half(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, bool, double, double,
                    const clg_field_s*, double, double);
int wilson_schur_once(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, clg_field_s*);
int wilson_full_hop(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool);
int arn_step(clg_context, clg_field_s* const*, int, clg_field_s*, double*, int, bool, clg_field_s*, double*);
int blas_gram_dev(clg_context, clg_field_s* const*, int, clg_field_s* const*, int, double*, int);
int blas_scale_inv_sqrt(clg_context, clg_field_s*, const clg_field_s*, const double*);
}  // namespace clg

struct clg_gcrodr_s {
    clg_context ctx = nullptr;
    clg_gcrodr_params p{};
    std::vector<clg_field_s*> U;  // recycled vectors (owned)
    int dims[4] = {0, 0, 0, 0};
    int subset = -1, kind = -1;
    bool have_u = false;
};

namespace {

constexpr int kMaxDimDR = 31;  // MDim limit of the reference (_kMaxDimDR, CSolverGCRODR.h)

// ------------------------------------------------------------------ the small problems of a cycle, on the device
// One workgroup solves a cycle's dense problems on the (m+1) x m matrix G the Arnoldi steps left in
// device memory (column-major, ld m+1): the least squares for y, the residual update coefficients, the
// deflation eigenvectors P (FindPk1 / FindPk2, CSolverGCRODR.cpp:541-700, CLinearAlgebraHelper.h:
// 150-237) and the QR of G P with the matrices GenerateCU multiplies the basis by.  Only |r| goes back
// to the host, for the stop test.
struct CycleArgs {
    const double* G;   // (m+1) x m, column-major
    int m, k, mode;    // mode 0: after the first GMRES cycle (H = G, rhs beta e_0); 1: GCRO-DR cycle
    int deflate;       // CLG_DEFLATE_REV / HEV / SVD
    const double* nrm;  // mode 1: |U_i|^2 before the normalisation (the diagonal block of G is 1/|U_i|)
    const double* r2;   // mode 0: beta^2; mode 1: |r|^2 (rhs = sqrt(.) e_k)
    const double* UU;   // k x k, row-major ld 32 (mode 1 SVD and REV)
    const double* WU;   // (m+1) x k (mode 1 HEV)
    const double* UW;   // k x (m+1) (mode 1 REV)
    double* y;          // m coefficients of the solution update
    double* cw;         // m+1 coefficients of the residual update: Wr += sum_i cw_i W_i
    double* MU;         // m x k row-major: the matrix U is recomputed with (P R^-1)
    double* MC;         // (m+1) x k row-major: the matrix C is recomputed with (Q)
    double* status;     // 0 on success, 1 if the deflation eigenproblem failed
};

__device__ __forceinline__ clgd::sld::cz ldz(const double* p, size_t e) { return clgd::sld::cz{p[2 * e], p[2 * e + 1]}; }
__device__ __forceinline__ void stz(double* p, size_t e, clgd::sld::cz v) {
    p[2 * e] = v.x;
    p[2 * e + 1] = v.y;
}

__global__ __launch_bounds__(256) void k_gcro_cycle(CycleArgs a) {
    using namespace clgd::sld;
    __shared__ cz mats[9][kLD * kLD];
    __shared__ cz vy[kLD], vr[kLD];
    __shared__ double red[4];
    cz *Gs = mats[0], *A = mats[1], *B = mats[2], *P = mats[3];
    cz *S0 = mats[4], *S1 = mats[5], *S2 = mats[6], *S3 = mats[7], *S4 = mats[8];
    const int m = a.m, k = a.k, t = threadIdx.x, T = blockDim.x;
    for (int e = t; e < (m + 1) * m; e += T) {
        const int i = e / m, j = e % m;
        Gs[i * kLD + j] = ldz(a.G, (size_t)j * (m + 1) + i);
    }
    __syncthreads();
    if (a.mode == 1) {
        // the recycled block: G(i, j < k) = 0 but G(i, i) = 1 / |U_i| (NormUkAndSetD)
        for (int e = t; e < k * k; e += T) {
            const int i = e / k, j = e % k;
            Gs[i * kLD + j] = i == j ? cz{1.0 / sqrt(a.nrm[i]), 0.0} : zz();
        }
        __syncthreads();
    }
    // least squares min |rhs - G y|, rhs = beta e_0 or |r| e_k
    const int ir = a.mode == 0 ? 0 : k;
    const double beta = sqrt(a.r2[0]);
    copy(S0, Gs, m + 1, m);
    for (int i = t; i <= m; i += T) vr[i] = i == ir ? cz{beta, 0.0} : zz();
    __syncthreads();
    hessenberg_lsq(S0, vr, vy, m, red);
    for (int j = t; j < m; j += T) stz(a.y, j, vy[j]);
    // residual update: mode 0 Wr = -(G y - beta e_0) over W; mode 1 Wr = |r| v_0 - (G y) over W
    for (int i = t; i <= m; i += T) {
        cz s = zz();
        for (int j = 0; j < m; ++j) cfma(s, Gs[i * kLD + j], vy[j]);
        if (a.mode == 0 && i == 0) s = csub(s, cz{beta, 0.0});
        stz(a.cw, i, cz{-s.x, -s.y});
    }
    __syncthreads();
    // deflation vectors P (m x k)
    bool ok = true;
    if (a.mode == 0) {
        if (a.deflate == CLG_DEFLATE_REV) {
            copy(A, Gs, m, m);
            ok = select_eigvecs(A, nullptr, false, m, k, true, P, S0, S1, S2, S3, S4, red);
        } else if (a.deflate == CLG_DEFLATE_HEV) {
            copy(A, Gs, m, m);
            matmul(B, Gs, Gs, m, m + 1, m, true);  // H^H H
            ok = select_eigvecs(A, B, false, m, k, false, P, S0, S1, S2, S3, S4, red);
        } else {
            matmul(A, Gs, Gs, m, m + 1, m, true);
            ok = select_eigvecs(A, nullptr, true, m, k, true, P, S0, S1, S2, S3, S4, red);
        }
    } else if (a.deflate == CLG_DEFLATE_SVD) {
        // G^H G p = lam B p, B = the Gram matrix of [U, V] (identity but for the U U^H block)
        matmul(A, Gs, Gs, m, m + 1, m, true);
        for (int e = t; e < m * m; e += T) {
            const int i = e / m, j = e % m;
            B[i * kLD + j] = (i < k && j < k && i != j) ? ldz(a.UU, (size_t)i * kLD + j) : (i == j ? one() : zz());
        }
        __syncthreads();
        ok = select_eigvecs(A, B, true, m, k, true, P, S0, S1, S2, S3, S4, red);
    } else if (a.deflate == CLG_DEFLATE_HEV) {
        // ALeft ((m+1) x m): columns j < k from W^H U, identity elsewhere; G^H AL p = lam G^H G p
        for (int e = t; e < (m + 1) * m; e += T) {
            const int i = e / m, j = e % m;
            S0[i * kLD + j] = j < k ? ldz(a.WU, (size_t)i * kLD + j) : (i == j ? one() : zz());
        }
        __syncthreads();
        matmul(A, Gs, S0, m, m + 1, m, true);
        matmul(B, Gs, Gs, m, m + 1, m, true);
        ok = select_eigvecs(A, B, false, m, k, false, P, S0, S1, S2, S3, S4, red);
    } else {
        // REV: ALeft (m x (m+1)) rows i < k from U^H W, identity elsewhere; B the Gram matrix of [U, V]
        for (int e = t; e < m * (m + 1); e += T) {
            const int i = e / (m + 1), j = e % (m + 1);
            S0[i * kLD + j] = i < k ? ldz(a.UW, (size_t)i * kLD + j) : (i == j ? one() : zz());
        }
        __syncthreads();
        for (int e = t; e < m * m; e += T) {
            const int i = e / m, j = e % m;
            cz v;
            if (i < k && j < k) v = i == j ? one() : ldz(a.UU, (size_t)i * kLD + j);
            else if (i < k) v = S0[i * kLD + j];
            else if (j < k) v = cconj(S0[j * kLD + i]);
            else v = i == j ? one() : zz();
            B[i * kLD + j] = v;
        }
        __syncthreads();
        matmul(A, S0, Gs, m, m + 1, m, false);
        ok = select_eigvecs(A, B, false, m, k, true, P, S0, S1, S2, S3, S4, red);
    }
    if (!ok) {
        if (t == 0) a.status[0] = 1.0;
        return;
    }
    // GenerateCU: G P = Q R; U <- basis P R^-1, C <- W Q
    matmul(A, Gs, P, m + 1, m, k, false);
    thin_qr(A, B, S0, m + 1, k, red);
    upper_inverse(S0, S1, k);
    matmul(S2, P, S1, m, k, k, false);
    for (int e = t; e < m * k; e += T) stz(a.MU, e, S2[(e / k) * kLD + e % k]);
    for (int e = t; e < (m + 1) * k; e += T) stz(a.MC, e, B[(e / k) * kLD + e % k]);
    if (t == 0) a.status[0] = 0.0;
}

// GenerateCUFirstTime's R: R(j, i) = the two Gram-Schmidt passes' projections, R(i, i) = |C_i'|; the
// matrix U is multiplied by (R^-1, row-major k x k)
__global__ __launch_bounds__(64) void k_gcro_rinv(int k, const double* proj, const double* nrm, double* Rinv) {
    using namespace clgd::sld;
    __shared__ cz R[kLD * kLD], X[kLD * kLD];
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
        const int j = e / k, i = e % k;  // R(j, i)
        cz v = zz();
        if (j < i) v = cadd(ldz(proj, ((size_t)0 * kLD + i) * kLD + j), ldz(proj, ((size_t)1 * kLD + i) * kLD + j));
        else if (j == i) v = cz{sqrt(nrm[i]), 0.0};
        R[j * kLD + i] = v;
    }
    __syncthreads();
    upper_inverse(R, X, k);
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) stz(Rinv, e, X[(e / k) * kLD + e % k]);
}

__global__ void k_fill0(double* p, int n) {
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0.0;
}

struct Gcro {
    clg_gcrodr_s* S;
    clg_context ctx;
    const clg_gcrodr_params* p;
    int m, k;
    const clg_field_s* gauge;
    double param;
    int op;        // operator A (EFO_F_D .. EFO_F_DDdagger), or the Schur D / D^dag with schur
    bool schur;
    clg_field_s* todd = nullptr;
    std::vector<clg_field_s*> C, V;  // C[0..k), V[0..m-k]
    clg_field_s *X = nullptr, *Wr = nullptr, *tmp = nullptr;
    double* dws = nullptr;  // device scratch: G, coefficients, Gram matrices, scalars
    // offsets (doubles) into dws
    size_t offG = 0, offY = 0, offCW = 0, offMU = 0, offMC = 0, offUU = 0, offWU = 0, offUW = 0, offNrm = 0,
           offR2 = 0, offSt = 0, offC1 = 0, offProj = 0, offMX = 0, offS = 0, offA = 0, end = 0;
    int applies = 0, cycles = 0;
    double dev = 0;

    void layout() {
        constexpr size_t M32 = 2 * 32 * 32;  // one 32 x 32 complex matrix
        offG = 0;
        offY = offG + 2 * (size_t)(m + 1) * m;
        offCW = offY + 2 * 32;
        offMU = offCW + 2 * 33;
        offMC = offMU + M32;
        offUU = offMC + M32;
        offWU = offUU + M32;
        offUW = offWU + M32;
        offNrm = offUW + M32;
        offR2 = offNrm + 32;
        offSt = offR2 + 4;
        offC1 = offSt + 4;
        offProj = offC1 + 2 * 32;
        offMX = offProj + 2 * M32;
        offS = offMX + M32;
        offA = offS + 16;
        end = offA + kArnScratch;
    }
    clg_field_s* W(int j) { return j < k ? C[j] : V[j - k]; }           // GetW
    clg_field_s* Vv(int j) { return j < k ? S->U[j] : V[j - k]; }       // GetV

    int apply(clg_field_s* out, const clg_field_s* in) {
        ++applies;
        if (schur) return wilson_schur_once(ctx, out, in, gauge, param, op == EFO_F_Ddagger, todd);
        return linop_apply(ctx, p->kind, op, gauge, param, out, in);
    }
    // Wilson D (D^dagger) and its even-site Schur complement are A = 1 - K with K = kappa H or
    // kappa^2 H_eo H_oe: the Arnoldi step then runs on K, whose projections cancel nothing (A W_j keeps
    // most of its norm along W_j, which makes the reorthogonalisation pass fire at every step), and the
    // column of G follows as e_j - (projections of K W_j): the C rows are exact because the Krylov
    // vectors are orthogonal to C.
    bool shiftable() const { return p->kind == CLG_WILSON && (op == EFO_F_D || op == EFO_F_Ddagger); }
    int apply_k(clg_field_s* out, const clg_field_s* in) {
        ++applies;
        const bool dag = op == EFO_F_Ddagger;
        if (!schur) return wilson_full_hop(ctx, out, in, gauge, param, dag);
        CLG_TRY(wilson_hop_half(ctx, todd, in, gauge, dag, 1.0, 0.0, nullptr, 0.0, 0.0));
        return wilson_hop_half(ctx, out, todd, gauge, dag, param * param, 0.0, nullptr, 0.0, 0.0);
    }
    // one Arnoldi step of the cycle: W_{j+1} from A W_j, orthogonalised against W_0..W_j, and column j
    // of G (the reference's B block above its Hessenberg block, CSolverGCRODR.cpp:277-316)
    int arnoldi(int j) {
        const bool sh = shiftable();
        CLG_TRY(sh ? apply_k(W(j + 1), W(j)) : apply(W(j + 1), W(j)));
        std::vector<clg_field_s*> list(j + 1);
        for (int i = 0; i <= j; ++i) list[i] = W(i);
        return arn_step(ctx, list.data(), j + 1, W(j + 1), dws + offG + 2 * (size_t)j * (m + 1), j, sh, nullptr,
                        dws + offA);
    }
    // |f| on the host (the stop tests)
    int norm(const clg_field_s* f, double& out) {
        CLG_TRY(blas_norm2_dev(ctx, f, dws + offS));
        double v;
        CLG_TRY(fetch_doubles(ctx, dws + offS, 1, &v));
        out = std::sqrt(v);
        return CLG_OK;
    }
    int fill0(double* d, int n) {
        k_fill0<<<std::max(1, std::min(64, (n + 255) / 256)), 256, 0, ctx->stream>>>(d, n);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    // r = b - A x
    int residual(clg_field_s* r, const clg_field_s* x, const clg_field_s* b) {
        CLG_TRY(apply(r, x));
        return blas_axpby(ctx, r, 1.0, 0.0, b, -1.0, 0.0);
    }
    // the cycle's dense problems (see k_gcro_cycle)
    int small(int mode) {
        if (mode == 1) {
            std::vector<clg_field_s*> Us(S->U.begin(), S->U.end());
            if (p->deflate != CLG_DEFLATE_HEV) CLG_TRY(blas_gram_dev(ctx, Us.data(), k, Us.data(), k, dws + offUU, 32));
            std::vector<clg_field_s*> Ws;
            for (int j = 0; j <= m; ++j) Ws.push_back(W(j));
            if (p->deflate == CLG_DEFLATE_HEV) CLG_TRY(blas_gram_dev(ctx, Ws.data(), m + 1, Us.data(), k, dws + offWU, 32));
            if (p->deflate == CLG_DEFLATE_REV) CLG_TRY(blas_gram_dev(ctx, Us.data(), k, Ws.data(), m + 1, dws + offUW, 32));
        }
        CycleArgs a;
        a.G = dws + offG;
        a.m = m;
        a.k = k;
        a.mode = mode;
        a.deflate = p->deflate;
        a.nrm = dws + offNrm;
        a.r2 = dws + offR2;
        a.UU = dws + offUU;
        a.WU = dws + offWU;
        a.UW = dws + offUW;
        a.y = dws + offY;
        a.cw = dws + offCW;
        a.MU = dws + offMU;
        a.MC = dws + offMC;
        a.status = dws + offSt;
        k_gcro_cycle<<<1, 256, 0, ctx->stream>>>(a);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    // |Wr| and the status of the cycle's eigenproblem in one read
    int fetch_dev() {
        CLG_TRY(blas_norm2_dev(ctx, Wr, dws + offSt + 1));
        double v[2];
        CLG_TRY(fetch_doubles(ctx, dws + offSt, 2, v));
        CLG_CHECK_ARG(v[0] == 0.0, "GCRODR: deflation eigenproblem failed");
        dev = std::sqrt(v[1]);
        return CLG_OK;
    }
    // GenerateCU (CSolverGCRODR.cpp:541-560) with the device matrices: U <- basis (P R^-1); C <- W Q
    int generate_cu(bool update_c, bool just_after_gmres) {
        std::vector<clg_field_s*> in, outU(S->U.begin(), S->U.end());
        for (int j = 0; j < m; ++j) in.push_back(just_after_gmres ? W(j) : Vv(j));
        CLG_TRY(blas_vmm_dev(ctx, in.data(), m, outU.data(), k, dws + offMU));
        if (update_c) {
            std::vector<clg_field_s*> inC, outC(C.begin(), C.end());
            for (int j = 0; j <= m; ++j) inC.push_back(W(j));
            CLG_TRY(blas_vmm_dev(ctx, inC.data(), m + 1, outC.data(), k, dws + offMC));
        }
        S->have_u = true;
        return CLG_OK;
    }
    // GenerateCUFirstTime: C = A U orthonormalised (Gram-Schmidt with a second pass, the projections
    // kept on the device), U <- U R^-1, r = b - A x
    int generate_cu_first(clg_field_s* x, clg_field_s* r, const clg_field_s* b) {
        for (int i = 0; i < k; ++i) CLG_TRY(apply(C[i], S->U[i]));
        for (int i = 0; i < k; ++i) {
            for (int pass = 0; pass < 2 && i > 0; ++pass) {
                double* pr = dws + offProj + 2 * ((size_t)pass * 32 + i) * 32;
                std::vector<clg_field_s*> prev(C.begin(), C.begin() + i);
                CLG_TRY(blas_multidot_dev(ctx, prev.data(), i, C[i], pr, nullptr, nullptr));
                CLG_TRY(blas_multiaxpy_dev(ctx, C[i], prev.data(), i, pr, -1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
            }
            CLG_TRY(blas_norm2_dev(ctx, C[i], dws + offNrm + i));
            CLG_TRY(blas_scale_inv_sqrt(ctx, C[i], C[i], dws + offNrm + i));
        }
        k_gcro_rinv<<<1, 64, 0, ctx->stream>>>(k, dws + offProj, dws + offNrm, dws + offMX);
        CLG_HIP(hipGetLastError());
        std::vector<clg_field_s*> Us(S->U.begin(), S->U.end());
        CLG_TRY(blas_vmm_dev(ctx, Us.data(), k, Us.data(), k, dws + offMX));
        CLG_TRY(residual(r, x, b));
        return norm(r, dev);
    }

    // the solve on the system vectors (full lattice or even sites)
    int run(clg_field_s* xout, const clg_field_s* b, const clg_field_s* start, int phase, double blen, bool& conv) {
        const double tol = p->accuracy * blen;
        CLG_TRY(clg_field_copy(X, const_cast<clg_field_s*>(start ? start : b)));
        const bool reuse = S->have_u && (phase == CLG_PHASE_IN_TRAJECTORY || phase == CLG_PHASE_END_TRAJECTORY);
        clg_field_s* r = V[0];  // the reference keeps r in the first Krylov vector
        CLG_TRY(fill0(dws + offProj, 2 * 2 * 32 * 32));
        if (reuse) {
            CLG_TRY(generate_cu_first(X, r, b));
        } else {
            // one GMRES(m) cycle with its basis in W (FirstTimeGMERESSolve)
            CLG_TRY(fill0(dws + offG, 2 * (m + 1) * m));
            clg_field_s* w0 = W(0);
            CLG_TRY(residual(w0, X, b));
            CLG_TRY(blas_norm2_dev(ctx, w0, dws + offR2));
            CLG_TRY(blas_scale_inv_sqrt(ctx, w0, w0, dws + offR2));
            for (int j = 0; j < m; ++j) CLG_TRY(arnoldi(j));
            ++cycles;
            CLG_TRY(small(0));
            std::vector<clg_field_s*> basis;
            for (int j = 0; j < m; ++j) basis.push_back(W(j));
            CLG_TRY(blas_multiaxpy_dev(ctx, X, basis.data(), m, dws + offY, 1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
            if (p->recalculate_r > 1) {
                // r = -sum_j (H y - beta e_0)_j W_j
                std::vector<clg_field_s*> full;
                for (int j = 0; j <= m; ++j) full.push_back(W(j));
                CLG_TRY(clg_field_zero(Wr));
                CLG_TRY(blas_multiaxpy_dev(ctx, Wr, full.data(), m + 1, dws + offCW, 1.0, nullptr, nullptr, nullptr, 0.0,
                                           nullptr));
            } else {
                CLG_TRY(residual(Wr, X, b));
            }
            CLG_TRY(fetch_dev());
            CLG_TRY(generate_cu(dev >= tol, true));
            CLG_TRY(clg_field_copy(r, Wr));
        }
        if (dev < tol) {
            conv = true;
            return clg_field_copy(xout, X);
        }
        for (int it = 0; it < p->restart; ++it) {
            // OrthognalXR: x += U C^H r, r -= C C^H r (coefficients from the unmodified r)
            {
                std::vector<clg_field_s*> Cs(C.begin(), C.end()), Us(S->U.begin(), S->U.end());
                CLG_TRY(blas_multidot_dev(ctx, Cs.data(), k, r, dws + offC1, nullptr, nullptr));
                CLG_TRY(blas_multiaxpy_dev(ctx, X, Us.data(), k, dws + offC1, 1.0, nullptr, nullptr, nullptr, 0.0,
                                           nullptr));
                CLG_TRY(blas_multiaxpy_dev(ctx, r, Cs.data(), k, dws + offC1, -1.0, nullptr, nullptr, nullptr, 0.0,
                                           nullptr));
            }
            // NormUkAndSetD: U_i normalised (the diagonal block of G is set from the norms on the device)
            CLG_TRY(fill0(dws + offG, 2 * (m + 1) * m));
            for (int i = 0; i < k; ++i) {
                CLG_TRY(blas_norm2_dev(ctx, S->U[i], dws + offNrm + i));
                CLG_TRY(blas_scale_inv_sqrt(ctx, S->U[i], S->U[i], dws + offNrm + i));
            }
            // Wr = |r| v_0 (the residual before this cycle's correction), v_0 = r / |r|
            CLG_TRY(blas_norm2_dev(ctx, r, dws + offR2));
            CLG_TRY(clg_field_copy(Wr, r));
            CLG_TRY(blas_scale_inv_sqrt(ctx, r, r, dws + offR2));
            for (int j = k; j < m; ++j) CLG_TRY(arnoldi(j));
            ++cycles;
            CLG_TRY(small(1));
            std::vector<clg_field_s*> basis;
            for (int j = 0; j < m; ++j) basis.push_back(Vv(j));
            CLG_TRY(blas_multiaxpy_dev(ctx, X, basis.data(), m, dws + offY, 1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
            if (p->recalculate_r != 0 && it != 0 && it % p->recalculate_r == 0) {
                CLG_TRY(residual(Wr, X, b));
            } else {
                std::vector<clg_field_s*> full;
                for (int j = 0; j <= m; ++j) full.push_back(W(j));
                CLG_TRY(blas_multiaxpy_dev(ctx, Wr, full.data(), m + 1, dws + offCW, 1.0, nullptr, nullptr, nullptr, 0.0,
                                           nullptr));
            }
            CLG_TRY(fetch_dev());
            if (dev >= tol || phase != CLG_PHASE_END_TRAJECTORY) CLG_TRY(generate_cu(dev >= tol, false));
            if (dev < tol) {
                conv = true;
                return clg_field_copy(xout, X);
            }
            CLG_TRY(clg_field_copy(r, Wr));
        }
        conv = false;
        return clg_field_copy(xout, X);
    }
};

int check_params(const clg_gcrodr_params* p) {
    CLG_CHECK_ARG(p, "null params");
    CLG_CHECK_ARG(p->m_dim >= 5 && p->m_dim <= kMaxDimDR, "MDim must be in [5, %d]", kMaxDimDR);
    CLG_CHECK_ARG(p->k_dim >= 2 && p->k_dim <= p->m_dim - 2, "KDim must be in [2, MDim - 2]");
    CLG_CHECK_ARG(p->restart >= 0 && p->recalculate_r >= 0, "Restart / RecalculateR must be >= 0");
    CLG_CHECK_ARG(p->accuracy > 0, "Accuracy must be positive");
    CLG_CHECK_ARG(p->deflate >= CLG_DEFLATE_REV && p->deflate <= CLG_DEFLATE_SVD, "unknown DeflateSpace");
    CLG_CHECK_ARG(p->kind == CLG_WILSON || p->kind == CLG_KS, "kind must be CLG_WILSON or CLG_KS");
    CLG_CHECK_ARG(p->even_odd == 0 || p->kind == CLG_WILSON, "even_odd applies to Wilson fermions");
    return CLG_OK;
}

}  // namespace

extern "C" {

int clg_gcrodr_params_default(clg_gcrodr_params* p) {
    CLG_CHECK_ARG(p, "null params");
    std::memset(p, 0, sizeof(*p));
    p->m_dim = 10;  // CSolverGCRODR.cpp:69-78 defaults
    p->k_dim = 8;
    p->restart = 100;
    p->recalculate_r = 0;
    p->accuracy = 1e-8;
    p->absolute_accuracy = 0;
    p->deflate = CLG_DEFLATE_SVD;
    p->kind = CLG_WILSON;
    p->even_odd = 0;
    return CLG_OK;
}

static void gcrodr_free(void* obj) {
    clg_gcrodr_s* s = static_cast<clg_gcrodr_s*>(obj);
    for (clg_field_s* f : s->U) clg_field_destroy(f);
    delete s;
}

int clg_gcrodr_create(clg_context ctx, const clg_gcrodr_params* params, clg_gcrodr* out) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    CLG_TRY(check_params(params));
    clg_gcrodr_s* s = new clg_gcrodr_s;
    s->ctx = ctx;
    s->p = *params;
    ctx_adopt(ctx, s, gcrodr_free);
    *out = s;
    return CLG_OK;
}

int clg_gcrodr_destroy(clg_gcrodr s) {
    if (!s) return CLG_OK;
    ctx_release(s->ctx, s);
    gcrodr_free(s);
    return CLG_OK;
}

int clg_gcrodr_reset(clg_gcrodr s) {
    CLG_CHECK_ARG(s, "null solver");
    s->have_u = false;
    return CLG_OK;
}

int clg_gcrodr_solve(clg_gcrodr s, clg_field x, clg_field b, clg_field gauge, double kappa_or_mass, int op, int phase,
                     clg_field start, clg_solve_report* rep) {
    CLG_CHECK_ARG(s && x && b && gauge, "null argument");
    clg_context ctx = s->ctx;
    const clg_gcrodr_params* p = &s->p;
    CLG_CHECK_ARG(op >= EFO_F_D && op <= EFO_F_DDdagger, "op is the operator A (EFO_F_D .. EFO_F_DDdagger)");
    CLG_CHECK_ARG(!p->even_odd || op == EFO_F_D || op == EFO_F_Ddagger, "even-odd GCRODR takes D or D^dagger");
    CLG_CHECK_ARG(phase >= CLG_PHASE_START_TRAJECTORY && phase <= CLG_PHASE_ONCE, "unknown solver phase");
    CLG_CHECK_ARG(x->kind == p->kind && b->kind == p->kind, "field kind does not match params.kind");
    CLG_CHECK_ARG(x->prec == CLG_F64 && b->prec == CLG_F64 && x->subset == CLG_FULL && b->subset == CLG_FULL,
                  "GCRODR takes FULL FP64 fields");
    CLG_TRY(check_same_lattice(x, b));
    CLG_TRY(check_same_lattice(x, gauge));
    CLG_CHECK_ARG(gauge->kind == CLG_GAUGE && gauge->prec == CLG_F64, "GCRODR takes FP64 18-real links");
    if (start) {
        CLG_CHECK_ARG(start->kind == p->kind && start->prec == CLG_F64 && start->subset == CLG_FULL, "bad start field");
        CLG_TRY(check_same_lattice(start, b));
    }
    (void)hipSetDevice(ctx->device);
    const auto t0 = std::chrono::steady_clock::now();
    const int sub = p->even_odd ? CLG_EVEN : CLG_FULL;
    // (re)allocate the recycled space for this lattice
    bool same = s->subset == sub && s->kind == p->kind;
    for (int i = 0; i < 4; ++i) same = same && s->dims[i] == b->dims[i];
    if (!same || (int)s->U.size() != p->k_dim) {
        for (clg_field_s* f : s->U) clg_field_destroy(f);
        s->U.clear();
        for (int i = 0; i < p->k_dim; ++i) {
            clg_field_s* f = nullptr;
            CLG_TRY(clg_field_create(ctx, p->kind, CLG_F64, sub, b->dims, &f));
            s->U.push_back(f);
        }
        for (int i = 0; i < 4; ++i) s->dims[i] = b->dims[i];
        s->subset = sub;
        s->kind = p->kind;
        s->have_u = false;
    }
    for (clg_field_s* f : s->U)
        for (int i = 0; i < 4; ++i) f->bc[i] = b->bc[i];
    Gcro g;
    g.S = s;
    g.ctx = ctx;
    g.p = p;
    g.m = p->m_dim;
    g.k = p->k_dim;
    g.gauge = gauge;
    g.param = kappa_or_mass;
    g.schur = p->even_odd != 0;
    g.op = op;
    g.C.resize(g.k);
    g.V.resize(g.m - g.k + 1);
    for (int i = 0; i < g.k; ++i) CLG_TRY(scratch_field(ctx, b, p->kind, CLG_F64, sub, 30000 + i, &g.C[i]));
    for (int i = 0; i <= g.m - g.k; ++i) CLG_TRY(scratch_field(ctx, b, p->kind, CLG_F64, sub, 30100 + i, &g.V[i]));
    CLG_TRY(scratch_field(ctx, b, p->kind, CLG_F64, sub, 30200, &g.X));
    CLG_TRY(scratch_field(ctx, b, p->kind, CLG_F64, sub, 30201, &g.Wr));
    if (g.schur) CLG_TRY(scratch_field(ctx, b, CLG_WILSON, CLG_F64, CLG_ODD, 30202, &g.todd));
    g.layout();
    CLG_TRY(workspace(ctx, g.end, &g.dws));
    // |b|^2 of the full system for the reference stop rule
    double bn2;
    CLG_TRY(blas_norm2_dev(ctx, b, g.dws + g.offS));
    CLG_TRY(fetch_doubles(ctx, g.dws + g.offS, 1, &bn2));
    const double blen = p->absolute_accuracy ? 1.0 : bn2;
    bool conv = false;
    if (!g.schur) {
        CLG_TRY(g.run(x, b, start, phase, blen, conv));
    } else {
        // bhat_e = b_e + kappa H_eo b_o ; Mhat x_e = bhat_e ; x_o = b_o + kappa H_oe x_e
        const bool dag = op == EFO_F_Ddagger;
        clg_field_s *be, *bo, *xe, *xo, *se = nullptr;
        CLG_TRY(clg_field_parity_view(b, CLG_EVEN, &be));
        CLG_TRY(clg_field_parity_view(b, CLG_ODD, &bo));
        CLG_TRY(clg_field_parity_view(x, CLG_EVEN, &xe));
        CLG_TRY(clg_field_parity_view(x, CLG_ODD, &xo));
        if (start) CLG_TRY(clg_field_parity_view(start, CLG_EVEN, &se));
        struct Guard {
            clg_field_s* v[5];
            ~Guard() {
                for (auto f : v)
                    if (f) clg_field_destroy(f);
            }
        } guard{{be, bo, xe, xo, se}};
        clg_field_s* bhat = nullptr;
        CLG_TRY(scratch_field(ctx, b, CLG_WILSON, CLG_F64, CLG_EVEN, 30203, &bhat));
        CLG_TRY(wilson_hop_half(ctx, bhat, bo, gauge, dag, kappa_or_mass, 0.0, be, 1.0, 0.0));
        // relative accuracy is measured on the even-site right-hand side the Schur solve is given
        // (the reference's even-field solve: CFieldFermionWilsonSquareSU3.cu:1289-1292)
        double blen_e = blen;
        if (!p->absolute_accuracy) {
            CLG_TRY(blas_norm2_dev(ctx, bhat, g.dws + g.offS));
            CLG_TRY(fetch_doubles(ctx, g.dws + g.offS, 1, &blen_e));
        }
        CLG_TRY(g.run(xe, bhat, se, phase, blen_e, conv));
        CLG_TRY(wilson_hop_half(ctx, xo, xe, gauge, dag, kappa_or_mass, 0.0, bo, 1.0, 0.0));
    }
    // true residual of the requested full system
    double r2 = 0;
    {
        clg_field_s* t = nullptr;
        CLG_TRY(scratch_field(ctx, b, p->kind, CLG_F64, CLG_FULL, 30204, &t));
        const int fop = g.schur ? op : op;
        CLG_TRY(linop_apply(ctx, p->kind, fop, gauge, kappa_or_mass, t, x));
        CLG_TRY(blas_axpby(ctx, t, 1.0, 0.0, b, -1.0, 0.0));
        CLG_TRY(blas_norm2_dev(ctx, t, g.dws + g.offS));
        CLG_TRY(fetch_doubles(ctx, g.dws + g.offS, 1, &r2));
    }
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->converged = conv ? 1 : 0;
        rep->cycles = g.cycles;
        rep->outer = 1;
        rep->operator_applies = g.applies;
        rep->true_residual = bn2 > 0 ? std::sqrt(r2 / bn2) : std::sqrt(r2);
        rep->last_deviation = g.dev;
        rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return conv ? CLG_OK : CLG_NOT_CONVERGED;
}

}  // extern "C"
