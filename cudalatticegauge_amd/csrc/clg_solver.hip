// clg_solver.hip -- Krylov solvers on top of the operators: GMRES(m), BiCGStab, even-odd
// Schur wrapping and FP32 / FP16 inner solves inside FP64 defect correction.
//
// Reference semantics kept (CSolverGMRES.cpp:111-358, CSolverBiCGstab.cpp:403-607):
//   * x0 = b unless a start field is given; x may alias b;
//   * GMRES stop rule |g_m|^2 < Accuracy * (AbsoluteAccuracy ? 1 : |b|^2), checked at the end
//     of each cycle, FALSE (CLG_NOT_CONVERGED) with the best iterate after Restart cycles;
//   * BiCGStab: rho~ = conj(r0) (Saad), check |s|^2/|b|^2 every DiviationStep steps, restart
//     from the last x on breakdown.
// What differs, by design: the Arnoldi step orthogonalises with classical Gram-Schmidt
// (one multi-dot pass that also yields |w|^2, one multi-axpy pass that yields |w'|^2), and a
// second pass runs only when the device-side test |w'|^2 < |w|^2/2 asks for it, instead of
// modified Gram-Schmidt's j+1 dependent dot/axpy pairs.  Every Hessenberg entry, Givens
// rotation and step length stays in FP64 device memory; the host reads one scalar per cycle.
#include "clg_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int blas_axpy_dev(clg_context, clg_field_s*, const double*, double, const clg_field_s*);
int blas_scale_inv_sqrt(clg_context, clg_field_s*, const clg_field_s*, const double*);
int arn_step(clg_context, clg_field_s* const*, int, clg_field_s*, double*, int, bool, clg_field_s*, double*);
int blas_dot_dev(clg_context, const clg_field_s*, const clg_field_s*, double*);
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int blas_multidot_dev(clg_context, clg_field_s* const*, int, const clg_field_s*, double*, double*, const int*);
int blas_multiaxpy_dev(clg_context, clg_field_s*, clg_field_s* const*, int, const double*, double, double*, const int*,
                       const double*, double, int*);
int blas_gs_step(clg_context, clg_field_s* const*, int, clg_field_s*, double*, int, bool, clg_field_s*, double*);
int fetch_doubles(clg_context, const double*, int, double*);
int workspace(clg_context, size_t, double**);
int blas_conj(clg_context, clg_field_s*, const clg_field_s*);
int wilson_full_hop(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool);
int wilson_full(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, double, double);
int wilson_hop_half(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, bool, double, double,
                    const clg_field_s*, double, double);
int wilson_schur_once(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, clg_field_s*);
int ks_full(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, double, double);
int ks_schur_once(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, clg_field_s*);
int gauge_convert(clg_context, clg_field_s* dst, const clg_field_s* src);
}  // namespace clg

namespace {

// ------------------------------------------------------------------ small device kernels
// Givens rotations of the (m+1) x m Hessenberg matrix with g = beta e1 and back
// substitution (RotateH / SolveY, CSolverGMRES.cpp:262-358), FP64, one thread.
// out: y[0..m), dev[0] = |g_m|^2, dev[1] = |g_m|
__global__ void k_gmres_lsq(double* H, int ld, int m, const double* beta2, double* g, double* y, double* dev) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    typedef C2<double> c;
    auto Hp = [&](int r, int col) -> c* { return reinterpret_cast<c*>(H + 2 * (col * ld + r)); };
    c* G = reinterpret_cast<c*>(g);
    c* Y = reinterpret_cast<c*>(y);
    G[0] = c{sqrt(beta2[0]), 0.0};
    for (int k = 1; k <= m; ++k) G[k] = c{0.0, 0.0};
    for (int k = 0; k < m; ++k) {
        const c hkk = *Hp(k, k), hk1 = *Hp(k + 1, k);
        const double den = 1.0 / sqrt(hkk.x * hkk.x + hkk.y * hkk.y + hk1.x * hk1.x + hk1.y * hk1.y);
        const c cs = cscale(hkk, den), sn = cscale(hk1, den);
        const c csh = cconj(cs), snh = cconj(sn);
        for (int j = k; j < m; ++j) {
            const c hkj = *Hp(k, j), h1j = *Hp(k + 1, j);
            *Hp(k, j) = cadd(cmul(csh, hkj), cmul(snh, h1j));
            *Hp(k + 1, j) = csub(cmul(cs, h1j), cmul(sn, hkj));
        }
        const c mg = c{-G[k].x, -G[k].y};
        G[k] = cmul(csh, G[k]);
        G[k + 1] = cmul(sn, mg);
    }
    dev[0] = G[m].x * G[m].x + G[m].y * G[m].y;
    dev[1] = sqrt(dev[0]);
    for (int k = m - 1; k >= 0; --k) {
        c acc = G[k];
        for (int j = k + 1; j < m; ++j) acc = csub(acc, cmul(*Hp(k, j), Y[j]));
        const c d = *Hp(k, k);
        const double dd = d.x * d.x + d.y * d.y;
        Y[k] = c{(acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd};
    }
}

// BiCGStab scalar updates (all FP64 on device).  s[] layout:
// 0-1 rho, 2-3 alpha, 4-5 omega, 6-7 beta, 8-9 (rh.v), 10-11 (s.t), 12 (t.t), 13-14 rho_new,
// 16-17 -alpha, 18-19 -omega, 20 |s|^2
__global__ void k_bicg_alpha(double* s) {
    // alpha = rho / (rh . v)
    const double nr = s[0], ni = s[1], dr = s[8], di = s[9];
    const double dd = dr * dr + di * di;
    s[2] = (nr * dr + ni * di) / dd;
    s[3] = (ni * dr - nr * di) / dd;
    s[16] = -s[2];
    s[17] = -s[3];
}
__global__ void k_bicg_omega(double* s) {
    // omega = (s.t) / (t.t)
    s[4] = s[10] / s[12];
    s[5] = s[11] / s[12];
    s[18] = -s[4];
    s[19] = -s[5];
}
__global__ void k_bicg_beta(double* s) {
    // beta = alpha / (omega rho) * rho_new ; rho = rho_new
    const double ar = s[2], ai = s[3], orr = s[4], oi = s[5], rr = s[0], ri = s[1];
    const double mr = orr * rr - oi * ri, mi = orr * ri + oi * rr;
    const double dd = mr * mr + mi * mi;
    const double qr = (ar * mr + ai * mi) / dd, qi = (ai * mr - ar * mi) / dd;
    const double nr = s[13], ni = s[14];
    s[6] = qr * nr - qi * ni;
    s[7] = qr * ni + qi * nr;
    s[0] = nr;
    s[1] = ni;
}
// p = r + beta (p - omega v)
template <typename T>
__global__ void k_bicg_p(long n, const double* __restrict__ s, const T* __restrict__ r, const T* __restrict__ v, T* p) {
    const double br = s[6], bi = s[7], orr = s[4], oi = s[5];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double px = (double)p[i].x, py = (double)p[i].y, vx = (double)v[i].x, vy = (double)v[i].y;
        const double qx = px - (orr * vx - oi * vy), qy = py - (orr * vy + oi * vx);
        p[i].x = r[i].x + (br * qx - bi * qy);
        p[i].y = r[i].y + (br * qy + bi * qx);
    }
}

// ------------------------------------------------------------------ operators
struct LinOp {
    clg_context ctx = nullptr;
    int kind = CLG_WILSON;
    bool schur = false;  // even-site Schur complement (Wilson)
    int op = EFO_F_D;    // D, Ddagger, DD, DDdagger of the (full or Schur) operator
    const clg_field_s* gauge = nullptr;
    double param = 0;    // kappa or 2am
    int applies = 0;

    int apply(clg_field_s* out, const clg_field_s* in) {
        ++applies;
        if (kind == CLG_WILSON) {
            if (!schur) {
                if (op == EFO_F_D || op == EFO_F_Ddagger)
                    return wilson_full(ctx, out, in, gauge, param, op == EFO_F_Ddagger, 1.0, 0.0);
                clg_field_s* tmp = nullptr;
                CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_FULL, 1, &tmp));
                CLG_TRY(wilson_full(ctx, tmp, in, gauge, param, op == EFO_F_DDdagger, 1.0, 0.0));
                return wilson_full(ctx, out, tmp, gauge, param, false, 1.0, 0.0);
            }
            clg_field_s* todd = nullptr;
            CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_ODD, 1, &todd));
            if (op == EFO_F_D || op == EFO_F_Ddagger)
                return wilson_schur_once(ctx, out, in, gauge, param, op == EFO_F_Ddagger, todd);
            clg_field_s* teven = nullptr;
            CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_EVEN, 1, &teven));
            // op 2: Mhat Mhat; op 3: Mhat Mhat^dagger; op 5 (internal): Mhat^dagger Mhat
            if (op == 5) {
                CLG_TRY(wilson_schur_once(ctx, teven, in, gauge, param, false, todd));
                return wilson_schur_once(ctx, out, teven, gauge, param, true, todd);
            }
            CLG_TRY(wilson_schur_once(ctx, teven, in, gauge, param, op == EFO_F_DDdagger, todd));
            return wilson_schur_once(ctx, out, teven, gauge, param, false, todd);
        }
        if (schur) {
            // staggered: the parity block (2am)^2 - H H of D D^dagger, Hermitian (D = D^dagger here)
            clg_field_s* other = nullptr;
            CLG_TRY(scratch_field(ctx, in, CLG_KS, in->prec, in->subset == CLG_EVEN ? CLG_ODD : CLG_EVEN, 1, &other));
            for (int i = 0; i < 4; ++i) other->bc[i] = in->bc[i];
            if (op == EFO_F_D || op == EFO_F_Ddagger) return ks_schur_once(ctx, out, in, gauge, param, other);
            clg_field_s* same = nullptr;
            CLG_TRY(scratch_field(ctx, in, CLG_KS, in->prec, in->subset, 1, &same));
            for (int i = 0; i < 4; ++i) same->bc[i] = in->bc[i];
            CLG_TRY(ks_schur_once(ctx, same, in, gauge, param, other));
            return ks_schur_once(ctx, out, same, gauge, param, other);
        }
        // staggered, full lattice
        if (op == EFO_F_D || op == EFO_F_Ddagger)
            return ks_full(ctx, out, in, gauge, param, op == EFO_F_Ddagger, 1.0, 0.0);
        clg_field_s* tmp = nullptr;
        CLG_TRY(scratch_field(ctx, in, CLG_KS, in->prec, CLG_FULL, 1, &tmp));
        CLG_TRY(ks_full(ctx, tmp, in, gauge, param, op == EFO_F_DDdagger, 1.0, 0.0));
        return ks_full(ctx, out, tmp, gauge, param, false, 1.0, 0.0);
    }

    // Wilson D and the even-site Schur complement are 1 - K with K = kappa H or kappa^2 H_eo H_oe.
    bool shiftable() const { return kind == CLG_WILSON && (op == EFO_F_D || op == EFO_F_Ddagger); }
    // out = K in = in - A in, without forming the difference
    int apply_k(clg_field_s* out, const clg_field_s* in) {
        ++applies;
        const bool dag = op == EFO_F_Ddagger;
        if (!schur) return wilson_full_hop(ctx, out, in, gauge, param, dag);
        clg_field_s* todd = nullptr;
        CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_ODD, 1, &todd));
        CLG_TRY(wilson_hop_half(ctx, todd, in, gauge, dag, 1.0, 0.0, nullptr, 0.0, 0.0));
        return wilson_hop_half(ctx, out, todd, gauge, dag, param * param, 0.0, nullptr, 0.0, 0.0);
    }
};

struct GmresOut {
    int cycles = 0;
    double last_dev = 0;      // reference deviation |g_m|^2
    double rel_res = 0;       // |b - A x| / |b| at the last cycle start (stop rule 1)
    bool converged = false;
};

// Restarted GMRES(m) on A x = b with x holding the initial guess.
// stop_rule 0: |g_m|^2 < acc * blen ; stop_rule 1: |b - A x| < abs_tol (checked at cycle start,
// plus once more after the last cycle); stop_rule 2: |g_m| < abs_tol (Arnoldi estimate, used
// by the inner solves of defect correction, x0 = 0 so r0 = b is not recomputed).
int gmres(clg_context ctx, LinOp& A, clg_field_s* x, const clg_field_s* b, int m, int max_cycles, int stop_rule,
          double acc, double blen, double abs_tol, bool x_is_zero, GmresOut& out) {
    CLG_CHECK_ARG(m >= 1 && m <= 64, "MaxDim must be in [1, 64]");
    std::vector<clg_field_s*> V(m + 1);
    for (int k = 0; k <= m; ++k)
        CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 1000 + k, &V[k]));
    clg_field_s* w = V[m];
    const int ld = m + 1;
    // device workspace: H, c2, g, y, scalars
    const size_t offH = 0, offC = offH + 2 * (size_t)ld * m, offG = offC + 2 * (size_t)m,
                 offY = offG + 2 * (size_t)(m + 1), offS = offY + 2 * (size_t)m, offGS = offS + 16;
    double* wsp = nullptr;
    CLG_TRY(workspace(ctx, offGS + kArnScratch, &wsp));
    double* H = wsp + offH;
    double* g = wsp + offG;
    double* y = wsp + offY;
    double* sc = wsp + offS;  // sc[0] |r|^2, sc[1] |w|^2, sc[2..3] deviation, sc[4] |w|^2 before GS
    const bool shifted = A.shiftable();
    out = GmresOut();
    for (int cyc = 0; cyc < max_cycles; ++cyc) {
        // r = b - A x
        if (x_is_zero && cyc == 0) {
            CLG_TRY(blas_axpby(ctx, V[0], 1.0, 0.0, b, 0.0, 0.0));
        } else {
            CLG_TRY(A.apply(V[0], x));
            CLG_TRY(blas_axpby(ctx, V[0], 1.0, 0.0, b, -1.0, 0.0));
        }
        CLG_TRY(blas_norm2_dev(ctx, V[0], sc));
        if (stop_rule == 1) {
            double r2;
            CLG_TRY(fetch_doubles(ctx, sc, 1, &r2));
            out.rel_res = std::sqrt(r2);
            if (std::sqrt(r2) < abs_tol) { out.converged = true; return CLG_OK; }
            if (!(r2 > 0)) { out.converged = r2 == 0; return CLG_OK; }
        }
        CLG_TRY(blas_scale_inv_sqrt(ctx, V[0], V[0], sc));
        ++out.cycles;
        for (int j = 0; j < m; ++j) {
            if (shifted) CLG_TRY(A.apply_k(w, V[j]));
            else CLG_TRY(A.apply(w, V[j]));
            double* Hj = H + 2 * (size_t)j * ld;
            // classical Gram-Schmidt, repeated only when the first pass cancelled more than half
            // of |w|^2 (|w'|^2 < |w|^2 / 2, the Daniel-Gragg-Kaufman-Stewart test)
            if (m <= 16) {  // fused form: three passes, V[j+1] written normalised
                CLG_TRY(blas_gs_step(ctx, V.data(), j + 1, w, Hj, j, shifted, j < m - 1 ? V[j + 1] : nullptr,
                                     wsp + offGS));
                continue;
            }
            CLG_TRY(arn_step(ctx, V.data(), j + 1, w, Hj, j, shifted, j < m - 1 ? V[j + 1] : nullptr, wsp + offGS));
        }
        k_gmres_lsq<<<1, 64, 0, ctx->stream>>>(H, ld, m, sc, g, y, sc + 2);
        CLG_HIP(hipGetLastError());
        CLG_TRY(blas_multiaxpy_dev(ctx, x, V.data(), m, y, 1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
        x_is_zero = false;
        double dev[2];
        CLG_TRY(fetch_doubles(ctx, sc + 2, 2, dev));
        out.last_dev = dev[0];
        if (stop_rule == 0 && dev[0] < acc * blen) { out.converged = true; return CLG_OK; }
        if (stop_rule == 2 && dev[1] < abs_tol) { out.converged = true; return CLG_OK; }
        if (!std::isfinite(dev[0])) break;
    }
    if (stop_rule == 1) {
        CLG_TRY(A.apply(V[0], x));
        CLG_TRY(blas_axpby(ctx, V[0], 1.0, 0.0, b, -1.0, 0.0));
        CLG_TRY(blas_norm2_dev(ctx, V[0], sc));
        double r2;
        CLG_TRY(fetch_doubles(ctx, sc, 1, &r2));
        out.rel_res = std::sqrt(r2);
        out.converged = std::sqrt(r2) < abs_tol;
    }
    return CLG_OK;
}

// BiCGStab (CSolverBiCGstab.cpp:403-607) with device-resident scalars; the host reads |s|^2
// every deviation_check steps and rho at the restart test.
int bicgstab(clg_context ctx, LinOp& A, clg_field_s* x, const clg_field_s* b, int step_count, int dev_check,
             int retry, double accuracy, double blen, int& iters, double& last_dev, bool& converged) {
    clg_field_s *R, *P, *Vv, *Rh, *S, *T;
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 40, &R));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 41, &P));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 42, &Vv));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 43, &Rh));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 44, &S));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 45, &T));
    double* s = nullptr;
    CLG_TRY(workspace(ctx, 32, &s));
    const long n = (b->subset == CLG_FULL ? 2 : 1) * b->Vh * (b->kind == CLG_WILSON ? 12 : 3);
    const unsigned g = (unsigned)std::min<long>((n + 255) / 256, (long)ctx->num_cu * 16);
    iters = 0;
    converged = false;
    CLG_TRY(blas_axpby(ctx, R, 1.0, 0.0, x, 0.0, 0.0));
    for (int i = 0; i < retry; ++i) {
        // r = b - A r  (r holds x at entry)
        CLG_TRY(A.apply(T, R));
        CLG_TRY(blas_axpby(ctx, R, 1.0, 0.0, b, 0.0, 0.0));
        CLG_TRY(blas_axpby(ctx, R, -1.0, 0.0, T, 1.0, 0.0));
        CLG_TRY(blas_conj(ctx, Rh, R));
        CLG_TRY(blas_axpby(ctx, P, 1.0, 0.0, R, 0.0, 0.0));
        bool restart = false;
        for (int j = 0; j < step_count * dev_check; ++j) {
            ++iters;
            if (j == 0) CLG_TRY(blas_dot_dev(ctx, Rh, R, s + 0));
            double rho[2];
            CLG_TRY(fetch_doubles(ctx, s, 2, rho));
            const double small = (retry == i + 1) ? 1.175494351e-38 : 1e-20;
            if (rho[0] * rho[0] + rho[1] * rho[1] < small) { restart = true; ++i; break; }
            CLG_TRY(A.apply(Vv, P));
            CLG_TRY(blas_dot_dev(ctx, Rh, Vv, s + 8));
            k_bicg_alpha<<<1, 1, 0, ctx->stream>>>(s);
            CLG_TRY(blas_axpby(ctx, S, 1.0, 0.0, R, 0.0, 0.0));
            CLG_TRY(blas_axpy_dev(ctx, S, s + 16, 1.0, Vv));
            if ((j + 1) % dev_check == 0) {
                CLG_TRY(blas_norm2_dev(ctx, S, s + 20));
                double d;
                CLG_TRY(fetch_doubles(ctx, s + 20, 1, &d));
                last_dev = d / blen;
                if (last_dev < accuracy) { converged = true; return CLG_OK; }
            }
            CLG_TRY(A.apply(T, S));
            CLG_TRY(blas_dot_dev(ctx, S, T, s + 10));
            CLG_TRY(blas_norm2_dev(ctx, T, s + 12));
            k_bicg_omega<<<1, 1, 0, ctx->stream>>>(s);
            CLG_TRY(blas_axpby(ctx, R, 1.0, 0.0, S, 0.0, 0.0));
            CLG_TRY(blas_axpy_dev(ctx, R, s + 18, 1.0, T));
            CLG_TRY(blas_dot_dev(ctx, Rh, R, s + 13));
            k_bicg_beta<<<1, 1, 0, ctx->stream>>>(s);
            CLG_TRY(blas_axpy_dev(ctx, x, s + 2, 1.0, P));
            CLG_TRY(blas_axpy_dev(ctx, x, s + 4, 1.0, S));
            if (b->prec == CLG_F64)
                k_bicg_p<double2><<<g, 256, 0, ctx->stream>>>(n, s, (const double2*)R->data, (const double2*)Vv->data,
                                                                (double2*)P->data);
            else
                k_bicg_p<float2><<<g, 256, 0, ctx->stream>>>(n, s, (const float2*)R->data, (const float2*)Vv->data,
                                                               (float2*)P->data);
            CLG_HIP(hipGetLastError());
        }
        (void)restart;
        CLG_TRY(blas_axpby(ctx, R, 1.0, 0.0, x, 0.0, 0.0));
    }
    return CLG_OK;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int validate_params(const clg_solver_params* p) {
    CLG_CHECK_ARG(p, "null params");
    CLG_CHECK_ARG(p->solver == CLG_SOLVER_GMRES || p->solver == CLG_SOLVER_BICGSTAB, "unknown solver %d", p->solver);
    CLG_CHECK_ARG(p->max_dim >= 1 && p->max_dim <= 64, "MaxDim must be in [1,64]");
    CLG_CHECK_ARG(p->restart >= 1, "Restart must be >= 1");
    CLG_CHECK_ARG(p->accuracy > 0, "Accuracy must be > 0");
    CLG_CHECK_ARG(p->inner_prec >= CLG_F64 && p->inner_prec <= CLG_INNER_F16_VECTORS, "bad inner precision");
    if (p->inner_prec == CLG_INNER_F16_VECTORS)
        CLG_CHECK_ARG(p->kind == CLG_WILSON && p->solver == CLG_SOLVER_GMRES && p->max_dim <= 16 && p->even_odd != 2,
                      "FP16-vector inner solves take Wilson GMRES with MaxDim <= 16 (no normal equations)");
    return CLG_OK;
}

// Single system A x = b (A = D or D^dagger of the fermion, full lattice or through the
// even-site Schur complement), with the requested precision strategy.
struct Solve1 {
    clg_context ctx;
    const clg_solver_params* p;
    const clg_field_s* gauge;
    double param;
    int kind;
    bool dag;
    clg_solve_report* rep;
    int sum_cycles = 0;
    int sum_applies = 0;
    int outer = 0;
    double last_dev = 0;

    // D D^dagger y = b for staggered fermions: D D^dagger = (2am)^2 - H^2 does not couple the
    // parities, so each parity is one Schur-block solve (from a zero guess), each to a residual
    // below accuracy |b| / sqrt(2)
    int run_ks_dd(clg_field_s* y, const clg_field_s* b, double bnorm2, bool& conv) {
        conv = true;
        for (int par : {CLG_EVEN, CLG_ODD}) {
            clg_field_s *yp = nullptr, *bp = nullptr;
            CLG_TRY(clg_field_parity_view(y, par, &yp));
            const int st = clg_field_parity_view(const_cast<clg_field_s*>(b), par, &bp);
            if (st != CLG_OK) {
                clg_field_destroy(yp);
                return st;
            }
            bool c = false;
            const int rs = run_system(yp, bp, true, bnorm2 / 2, true, c);
            clg_field_destroy(yp);
            clg_field_destroy(bp);
            CLG_TRY(rs);
            conv = conv && c;
        }
        return CLG_OK;
    }

    // solve full-lattice A x = b with A = D or D^dagger; x holds the initial guess
    int run(clg_field_s* x, const clg_field_s* b, bool x_zero, double bnorm2, bool& conv) {
        if (p->even_odd != 0 && kind == CLG_KS) {
            // D x = b: x = D^dagger y with D D^dagger y = b; D^dagger x = b: x = D y (D and D^dagger
            // commute for staggered fermions)
            clg_field_s* y = nullptr;
            CLG_TRY(scratch_field(ctx, b, CLG_KS, b->prec, CLG_FULL, 53, &y));
            for (int i = 0; i < 4; ++i) y->bc[i] = b->bc[i];
            CLG_TRY(run_ks_dd(y, b, bnorm2, conv));
            return ks_full(ctx, x, y, gauge, param, !dag, 1.0, 0.0);
        }
        const bool eo = p->even_odd != 0 && kind == CLG_WILSON;
        if (!eo) return run_system(x, b, x_zero, bnorm2, false, conv);
        // even-odd: bhat_e = b_e + kappa H_eo b_o ; Mhat x_e = bhat_e ; x_o = b_o + kappa H_oe x_e
        clg_field_s *be, *bo, *xe, *xo;
        CLG_TRY(clg_field_parity_view(const_cast<clg_field_s*>(b), CLG_EVEN, &be));
        CLG_TRY(clg_field_parity_view(const_cast<clg_field_s*>(b), CLG_ODD, &bo));
        CLG_TRY(clg_field_parity_view(x, CLG_EVEN, &xe));
        CLG_TRY(clg_field_parity_view(x, CLG_ODD, &xo));
        struct ViewGuard {
            clg_field_s* v[4];
            ~ViewGuard() { for (auto f : v) clg_field_destroy(f); }
        } guard{{be, bo, xe, xo}};
        clg_field_s* bhat = nullptr;
        CLG_TRY(scratch_field(ctx, b, CLG_WILSON, b->prec, CLG_EVEN, 30, &bhat));
        CLG_TRY(wilson_hop_half(ctx, bhat, bo, gauge, dag, param, 0.0, be, 1.0, 0.0));
        if (!x_zero && p->even_odd == 1) {
            // keep the caller's x_e as the guess
        } else if (!x_zero) {
            // normal-equation variant: same guess
        }
        // the reference's relative stop rule measures the right-hand side the even-site solver is
        // given (fBLength = |bhat_e|^2, CFieldFermionWilsonSquareSU3.cu:1289-1292 with
        // CSolverGMRES.cpp:122 / CSolverBiCGstab.cpp:422-429); the true-residual rule keeps |b|
        double ref_blen2 = -1.0;
        if (p->stop_rule == CLG_STOP_REFERENCE && !p->absolute_accuracy && p->inner_prec == CLG_F64) {
            double* d = ctx->red_result + 9;
            CLG_TRY(blas_norm2_dev(ctx, bhat, d));
            CLG_TRY(fetch_doubles(ctx, d, 1, &ref_blen2));
        }
        bool c = false;
        CLG_TRY(run_system(xe, bhat, x_zero, bnorm2, true, c, ref_blen2));
        CLG_TRY(wilson_hop_half(ctx, xo, xe, gauge, dag, param, 0.0, bo, 1.0, 0.0));
        conv = c;
        return CLG_OK;
    }

    // A x = b on one lattice subset, precision strategy applied here
    int run_system(clg_field_s* x, const clg_field_s* b, bool x_zero, double bnorm2, bool schur, bool& conv,
                   double ref_blen2 = -1.0) {
        LinOp A;
        A.ctx = ctx;
        A.kind = kind;
        A.schur = schur;
        A.gauge = gauge;
        A.param = param;
        A.op = dag ? EFO_F_Ddagger : EFO_F_D;
        const bool normal = schur && p->even_odd == 2 && kind == CLG_WILSON;  // staggered blocks are HPD already
        const double bn = std::sqrt(bnorm2);  // |b| of the full system
        const double tol_abs = p->accuracy * bn;
        if (p->inner_prec == CLG_F64) {
            if (normal) return run_normal(A, x, b, x_zero, bn, conv);
            return run_plain(A, x, b, x_zero, ref_blen2 >= 0 ? ref_blen2 : bnorm2, tol_abs, conv);
        }
        return run_mixed(A, x, b, x_zero, bn, conv);
    }

    int run_plain(LinOp& A, clg_field_s* x, const clg_field_s* b, bool x_zero, double bnorm2, double tol_abs,
                  bool& conv) {
        if (p->solver == CLG_SOLVER_BICGSTAB) {
            int it = 0;
            double dev = 0;
            bool c = false;
            const double blen = p->absolute_accuracy ? 1.0 : bnorm2;
            if (x_zero) CLG_TRY(blas_axpby(ctx, x, 0.0, 0.0, b, 0.0, 0.0));
            CLG_TRY(bicgstab(ctx, A, x, b, p->max_step, p->deviation_step, p->restart, p->accuracy, blen, it, dev, c));
            sum_cycles += it;
            sum_applies += A.applies;
            last_dev = dev;
            conv = c;
            return CLG_OK;
        }
        GmresOut go;
        const double blen = p->absolute_accuracy ? 1.0 : bnorm2;
        const int rule = p->stop_rule == CLG_STOP_REFERENCE ? 0 : 1;
        CLG_TRY(gmres(ctx, A, x, b, p->max_dim, p->restart, rule, p->accuracy, blen, tol_abs, x_zero, go));
        sum_cycles += go.cycles;
        sum_applies += A.applies;
        last_dev = go.last_dev;
        conv = go.converged;
        outer = 1;
        return CLG_OK;
    }

    // Mhat^dagger Mhat x = Mhat^dagger b by GMRES, then report convergence of Mhat x = b
    int run_normal(LinOp& A, clg_field_s* x, const clg_field_s* b, bool x_zero, double bn, bool& conv) {
        clg_field_s* bb = nullptr;
        CLG_TRY(scratch_field(ctx, b, CLG_WILSON, b->prec, CLG_EVEN, 31, &bb));
        A.op = dag ? EFO_F_D : EFO_F_Ddagger;  // adjoint of the system operator
        CLG_TRY(A.apply(bb, b));
        LinOp N = A;
        N.op = dag ? EFO_F_DDdagger : 5;  // dag: Mhat Mhat^dagger ; else Mhat^dagger Mhat
        // tolerance on the normal residual that guarantees |b - Mhat x| < accuracy |b| is not
        // available without the spectrum; iterate cycles and test the true system residual.
        LinOp S = A;
        S.op = dag ? EFO_F_Ddagger : EFO_F_D;
        clg_field_s* r = nullptr;
        CLG_TRY(scratch_field(ctx, b, CLG_WILSON, b->prec, CLG_EVEN, 32, &r));
        double* d = ctx->red_result + 8;
        conv = false;
        for (int round = 0; round < p->restart; ++round) {
            GmresOut go;
            CLG_TRY(gmres(ctx, N, x, bb, p->max_dim, 1, 3, 0, 1, 0, x_zero && round == 0, go));
            sum_cycles += go.cycles;
            CLG_TRY(S.apply(r, x));
            CLG_TRY(blas_axpby(ctx, r, 1.0, 0.0, b, -1.0, 0.0));
            CLG_TRY(blas_norm2_dev(ctx, r, d));
            double r2;
            CLG_TRY(fetch_doubles(ctx, d, 1, &r2));
            last_dev = r2;
            if (std::sqrt(r2) < p->accuracy * bn) { conv = true; break; }
        }
        sum_applies += N.applies * 2 + S.applies + A.applies;
        outer = 1;
        return CLG_OK;
    }

    // FP64 defect correction around FP32 / FP16-storage inner GMRES
    int run_mixed(LinOp& A, clg_field_s* x, const clg_field_s* b, bool x_zero, double bn, bool& conv) {
        const int lp = p->inner_prec;
        // inner gauge: FP32 (for FP32) or FP16 12-real (FP16 links, FP16 links and vectors)
        clg_field_s* g_lo = nullptr;
        const int gprec = lp == CLG_F32 ? CLG_F32 : CLG_F16;
        const int gk = gprec == CLG_F16 ? CLG_GAUGE12 : gauge->kind;
        CLG_TRY(scratch_field(ctx, gauge, gk, gprec, CLG_FULL, 0, &g_lo));
        CLG_TRY(gauge_convert(ctx, g_lo, gauge));
        // Krylov vectors of the inner solve
        const int vprec = lp == CLG_INNER_F16_VECTORS ? CLG_F16 : CLG_F32;
        clg_field_s *r, *r_lo, *e_lo, *e_hi;
        CLG_TRY(scratch_field(ctx, b, b->kind, CLG_F64, b->subset, 33, &r));
        CLG_TRY(scratch_field(ctx, b, b->kind, vprec, b->subset, 34, &r_lo));
        CLG_TRY(scratch_field(ctx, b, b->kind, vprec, b->subset, 35, &e_lo));
        CLG_TRY(scratch_field(ctx, b, b->kind, CLG_F64, b->subset, 36, &e_hi));
        LinOp Alo = A;
        Alo.gauge = g_lo;
        Alo.applies = 0;
        double* d = ctx->red_result + 16;
        conv = false;
        if (x_zero) CLG_TRY(blas_axpby(ctx, x, 0.0, 0.0, b, 0.0, 0.0));
        const int max_outer = p->max_outer > 0 ? p->max_outer : 50;
        const double inner_tol = p->inner_accuracy > 0 ? p->inner_accuracy : 1e-5;
        for (int it = 0; it < max_outer; ++it) {
            // r = b - A x in FP64
            if (it == 0 && x_zero) {
                CLG_TRY(blas_axpby(ctx, r, 1.0, 0.0, b, 0.0, 0.0));
            } else {
                CLG_TRY(A.apply(r, x));
                CLG_TRY(blas_axpby(ctx, r, 1.0, 0.0, b, -1.0, 0.0));
            }
            CLG_TRY(blas_norm2_dev(ctx, r, d));
            double r2;
            CLG_TRY(fetch_doubles(ctx, d, 1, &r2));
            last_dev = r2;
            if (std::sqrt(r2) < p->accuracy * bn) { conv = true; break; }
            if (!std::isfinite(r2)) break;
            ++outer;
            CLG_TRY(clg_field_copy(r_lo, r));
            CLG_TRY(clg_field_zero(e_lo));
            GmresOut go;
            const double rn = std::sqrt(r2);
            CLG_TRY(gmres(ctx, Alo, e_lo, r_lo, p->max_dim, p->restart, 2, 0, 1, inner_tol * rn, true, go));
            sum_cycles += go.cycles;
            CLG_TRY(clg_field_copy(e_hi, e_lo));
            CLG_TRY(blas_axpby(ctx, x, 1.0, 0.0, e_hi, 1.0, 0.0));
        }
        sum_applies += A.applies + Alo.applies;
        return CLG_OK;
    }
};

}  // namespace

namespace clg {
// full-lattice operator application for other solver families (multi-shift, forces)
int linop_apply(clg_context ctx, int kind, int op, const clg_field_s* gauge, double param, clg_field_s* out,
                const clg_field_s* in) {
    LinOp A;
    A.ctx = ctx;
    A.kind = kind;
    A.schur = false;
    A.op = op;
    A.gauge = gauge;
    A.param = param;
    return A.apply(out, in);
}
}  // namespace clg

extern "C" {

int clg_solver_params_default(clg_solver_params* p) {
    CLG_CHECK_ARG(p, "null params");
    std::memset(p, 0, sizeof(*p));
    p->solver = CLG_SOLVER_GMRES;
    p->max_dim = 20;         // CSolverGMRES.cpp:45 default MaxDim
    p->restart = 3;          // CSolverGMRES.cpp:52 default Restart
    p->max_step = 1000;      // CSolverBiCGstab.cpp:40 MaxStep
    p->deviation_step = 10;  // CSolverBiCGstab.cpp:45 DiviationStep
    p->accuracy = 1e-8;      // CSolverGMRES.cpp:57 Accuracy
    p->absolute_accuracy = 0;
    p->stop_rule = CLG_STOP_REFERENCE;
    p->even_odd = 0;
    p->inner_prec = CLG_F64;
    p->inner_accuracy = 1e-5;
    p->max_outer = 50;
    p->kind = CLG_WILSON;
    return CLG_OK;
}

int clg_solve(clg_context ctx, const clg_solver_params* p, clg_field x, clg_field b, clg_field gauge,
              double param, int op, clg_field start, clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && x && b && gauge, "null argument");
    CLG_CHECK_ARG(!stream_capturing(ctx), "a solve blocks; not while recording a graph");
    CLG_TRY(validate_params(p));
    CLG_CHECK_ARG(b->kind == p->kind && x->kind == p->kind, "field kind does not match params.kind");
    CLG_CHECK_ARG(b->subset == CLG_FULL && x->subset == CLG_FULL, "solve takes FULL fields");
    // FP64 fields (inner_prec selects defect correction), or a plain FP32 solve: FP32 fields and an
    // FP32 gauge field with everything in FP32 arithmetic except the FP64 reductions, Hessenberg and
    // Givens step (the reference's single-precision build, CSolverGMRES.cpp with Real = float)
    CLG_CHECK_ARG(x->prec == b->prec && (b->prec == CLG_F64 || b->prec == CLG_F32),
                  "solve takes FP64 fields (inner_prec selects the rest) or FP32 fields");
    CLG_CHECK_ARG(gauge->prec == b->prec, "the gauge field must have the precision of the fermion fields");
    CLG_CHECK_ARG(b->prec == CLG_F64 || p->inner_prec == CLG_F64,
                  "an FP32 solve runs in FP32 throughout (inner_prec must be CLG_F64, no defect correction)");
    CLG_TRY(check_same_lattice(x, b));
    CLG_TRY(check_same_lattice(x, gauge));
    if (start) CLG_TRY(check_same_lattice(x, start));
    int base = op;
    if (op >= EFO_F_InverseD && op <= EFO_F_InverseDDdagger) base = op - EFO_F_InverseD;
    CLG_CHECK_ARG(base >= EFO_F_D && base <= EFO_F_DDdagger, "operator %d cannot be inverted", op);
    if (p->even_odd) {
        for (int i = 0; i < 4; ++i) CLG_CHECK_ARG((b->dims[i] & 1) == 0, "even-odd needs even extents");
    }
    (void)hipSetDevice(ctx->device);
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    const double t0 = now_s();
    for (int i = 0; i < 4; ++i) x->bc[i] = b->bc[i];

    // keep b intact when x aliases it
    const clg_field_s* bb = b;
    if (x == b || x->data == b->data) {
        clg_field_s* bc = nullptr;
        CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, CLG_FULL, 50, &bc));
        CLG_TRY(clg_field_copy(bc, b));
        bb = bc;
    }
    double bn2;
    CLG_TRY(blas_norm2_dev(ctx, bb, ctx->red_result));
    CLG_TRY(fetch_doubles(ctx, ctx->red_result, 1, &bn2));

    Solve1 S{ctx, p, gauge, param, p->kind, false, rep};
    bool conv = true;
    // initial guess: start, or b (CSolverGMRES.cpp:140-147)
    if (start) CLG_TRY(clg_field_copy(x, start));
    else if (x != bb) CLG_TRY(clg_field_copy(x, const_cast<clg_field_s*>(bb)));

    // op decomposition: D, D^dagger: one solve; DD = D D: D^-1 twice; DD^dagger = D D^dagger:
    // (D^dagger)^-1 D^-1
    bool c1 = false, c2 = true;
    if (p->kind == CLG_KS && p->even_odd && base == EFO_F_DDdagger) {
        CLG_TRY(S.run_ks_dd(x, bb, bn2, c1));  // the parity blocks of D D^dagger directly
    } else if (base == EFO_F_D || base == EFO_F_Ddagger) {
        S.dag = base == EFO_F_Ddagger;
        CLG_TRY(S.run(x, bb, false, bn2, c1));
    } else {
        clg_field_s* mid = nullptr;
        CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, CLG_FULL, 51, &mid));
        CLG_TRY(clg_field_copy(mid, const_cast<clg_field_s*>(bb)));
        S.dag = false;
        CLG_TRY(S.run(mid, bb, false, bn2, c1));
        double mn2;
        CLG_TRY(blas_norm2_dev(ctx, mid, ctx->red_result));
        CLG_TRY(fetch_doubles(ctx, ctx->red_result, 1, &mn2));
        S.dag = base == EFO_F_DDdagger;
        if (!start) CLG_TRY(clg_field_copy(x, mid));
        CLG_TRY(S.run(x, mid, false, mn2, c2));
    }
    conv = c1 && c2;

    // true residual of the requested system (FP64; FP32 arithmetic with FP64 sums for an FP32 solve)
    {
        clg_field_s* r = nullptr;
        CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, CLG_FULL, 52, &r));
        LinOp A;
        A.ctx = ctx;
        A.kind = p->kind;
        A.gauge = gauge;
        A.param = param;
        A.op = base;
        CLG_TRY(A.apply(r, x));
        CLG_TRY(blas_axpby(ctx, r, 1.0, 0.0, bb, -1.0, 0.0));
        double r2;
        CLG_TRY(blas_norm2_dev(ctx, r, ctx->red_result));
        CLG_TRY(fetch_doubles(ctx, ctx->red_result, 1, &r2));
        if (rep) rep->true_residual = bn2 > 0 ? std::sqrt(r2 / bn2) : std::sqrt(r2);
    }
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    if (rep) {
        rep->converged = conv ? 1 : 0;
        rep->cycles = S.sum_cycles;
        rep->outer = S.outer > 0 ? S.outer : 1;
        rep->operator_applies = S.sum_applies;
        rep->last_deviation = S.last_dev;
        rep->seconds = now_s() - t0;
    }
    return conv ? CLG_OK : CLG_NOT_CONVERGED;
}

}  // extern "C"
