// clg_multishift.hip -- multi-shift Krylov solvers (x_n = (A + c_n)^-1 b for all shifts at
// once) and the rational-approximation driver of RHMC.
//
// Reference algorithms kept:
//   * CMultiShiftFOM::Solve (CMultiShiftFOM.cpp:97-270): one Arnoldi basis per restart shared
//     by all shifts, per-shift FOM solve (H + c_n) y = beta0_n e1, restart from v_m with
//     beta0_n <- -h_{m,m-1} y_{m-1}; stop when max_n |beta0_n| < Accuracy * (1 or |b|^2).
//   * CMultiShiftBiCGStab::Solve (CMultiShiftBiCGStab.cpp:73-270): shifted BiCGStab with the
//     zeta / beta / alpha / chi / rho recurrences per shift, checked every DiviationStep steps:
//     max(|s|, |s_n|) < Accuracy * (1 or |b|).
//   * CFieldFermion::RationalApproximation (CField.cpp:114-148): out = c in + sum_i a_i x_i with
//     x_i = (A + b_i)^-1 in and coefficients [c, a_1..a_n, b_1..b_n] (CRationalApproximation.h).
// What differs: the per-shift vector updates of a step run as one batched kernel over all
// active shifts (the shared vectors are read once per step, not once per shift), converged
// shifts are left out of the launches, the FOM small systems are solved on device, the shifted
// BiCGStab keeps every scalar on the device (clg_msbicg.hip), and the shift sums accumulate in
// FP64 whatever the storage precision of the solutions.
#include "clg_internal.h"
#include "clg_smallla.h"

#include <algorithm>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int blas_dot_dev(clg_context, const clg_field_s*, const clg_field_s*, double*);
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int blas_conj(clg_context, clg_field_s*, const clg_field_s*);
int blas_multidot_dev(clg_context, clg_field_s* const*, int, const clg_field_s*, double*, double*, const int*);
int blas_multiaxpy_dev(clg_context, clg_field_s*, clg_field_s* const*, int, const double*, double, double*, const int*,
                       const double*, double, int*);
int blas_scale_inv_sqrt(clg_context, clg_field_s*, const clg_field_s*, const double*);
int fetch_doubles(clg_context, const double*, int, double*);
int workspace(clg_context, size_t, double**);
long n_complex(const clg_field_s*);
int linop_apply(clg_context ctx, int kind, int op, const clg_field_s* gauge, double param, clg_field_s* out,
                const clg_field_s* in);
int gauge_convert(clg_context, clg_field_s*, const clg_field_s*);
int arn_step(clg_context, clg_field_s* const*, int, clg_field_s*, double*, int, bool, clg_field_s*, double*);
int msbicgstab_dev(clg_context ctx, int kind, int op, const clg_field_s* gauge, double param,
                   std::vector<clg_field_s*>& X, const std::vector<double>& shifts, const clg_field_s* b, int step_count,
                   int dev_check, double accuracy, double blen, int& iters, double& maxerr, int& applies);
}  // namespace clg

namespace {

const int kMaxShift = 32;
typedef C2<double> cx;

// ---- FP64 arithmetic over 16-byte packets of F64 or F32 storage
template <int P> struct MS;
template <> struct MS<CLG_F64> {
    typedef double2 T;
    static constexpr int NC = 1;
    __device__ static void ld(const T* p, long i, double* v) { const double2 a = p[i]; v[0] = a.x; v[1] = a.y; }
    __device__ static void st(T* p, long i, const double* v) { p[i] = make_double2(v[0], v[1]); }
};
template <> struct MS<CLG_F32> {
    typedef float4 T;
    static constexpr int NC = 2;
    __device__ static void ld(const T* p, long i, double* v) {
        const float4 a = p[i];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
    __device__ static void st(T* p, long i, const double* v) {
        p[i] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    }
};

struct ShiftPtrs {
    void* x[kMaxShift];
    void* s[kMaxShift];
};
struct VPtrs {
    const void* v[64];
};
template <int NC>
__device__ __forceinline__ void cmad(double* acc, const double* c, const double* v) {
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        acc[2 * q] += c[0] * v[2 * q] - c[1] * v[2 * q + 1];
        acc[2 * q + 1] += c[0] * v[2 * q + 1] + c[1] * v[2 * q];
    }
}

// MS-FOM: x_k += sum_j y[k][j] V_j for every active shift k (y[k*2m + 2j])
template <int P>
__global__ void k_msfom_update(long n, int ns, int m, ShiftPtrs sp, VPtrs vp, const double* __restrict__ y) {
    constexpr int NC = MS<P>::NC;
    extern __shared__ double ys[];
    for (int t = threadIdx.x; t < ns * 2 * m; t += blockDim.x) ys[t] = y[t];
    __syncthreads();
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        for (int k = 0; k < ns; ++k) {
            double acc[2 * NC];
            MS<P>::ld((const typename MS<P>::T*)sp.x[k], i, acc);
            for (int j = 0; j < m; ++j) {
                double v[2 * NC];
                MS<P>::ld((const typename MS<P>::T*)vp.v[j], i, v);
                cmad<NC>(acc, ys + (size_t)k * 2 * m + 2 * j, v);
            }
            MS<P>::st((typename MS<P>::T*)sp.x[k], i, acc);
        }
    }
}
// per-shift FOM solve (CMultiShiftFOM RotateHSolveY, CMultiShiftFOM.cpp:272-330): workgroup k (one
// wave) copies H (m x m, column-major, ld = m+1) into LDS, adds c_k on the diagonal, applies the m-1
// Givens rotations to g = beta0_k e1 (each rotation spread over the columns it touches), back-
// substitutes with wave-wide sums, then beta0_k <- -h_{m,m-1} y_{m-1}.
__global__ __launch_bounds__(64) void k_msfom_lsq(int ns, int m, const double* __restrict__ H,
                                                  const double* __restrict__ shifts, double* y, double* beta0,
                                                  double* errs) {
    const int k = blockIdx.x;
    if (k >= ns) return;
    extern __shared__ double lds[];
    cx* Hc = reinterpret_cast<cx*>(lds);  // m x m, column-major
    __shared__ cx G[64], Y[64], rot[2];
    const int t = threadIdx.x, ld = m + 1;
    const cx* Hs = reinterpret_cast<const cx*>(H);
    for (int e = t; e < m * m; e += 64) {
        const int col = e / m, r = e - col * m;
        cx v = Hs[col * ld + r];
        if (r == col) v = cadd(v, cx{shifts[2 * k], shifts[2 * k + 1]});
        Hc[e] = v;
    }
    for (int d = t; d < m; d += 64) G[d] = cx{0.0, 0.0};
    __syncthreads();
    if (t == 0) G[0] = cx{beta0[2 * k], beta0[2 * k + 1]};
    for (int i = 0; i < m - 1; ++i) {
        __syncthreads();
        if (t == 0) {
            const cx hii = Hc[i * m + i], hi1 = Hc[i * m + i + 1];
            const double den = 1.0 / sqrt(hii.x * hii.x + hii.y * hii.y + hi1.x * hi1.x + hi1.y * hi1.y);
            rot[0] = cscale(hii, den);
            rot[1] = cscale(hi1, den);
            const cx mg = cx{-G[i].x, -G[i].y};
            G[i] = cmul(cconj(rot[0]), G[i]);
            G[i + 1] = cmul(rot[1], mg);
        }
        __syncthreads();
        const cx c = rot[0], sn = rot[1], ch = cconj(c), sh = cconj(sn);
        for (int jj = i + t; jj < m; jj += 64) {
            const cx hij = Hc[jj * m + i], h1j = Hc[jj * m + i + 1];
            Hc[jj * m + i] = cadd(cmul(ch, hij), cmul(sh, h1j));
            Hc[jj * m + i + 1] = csub(cmul(c, h1j), cmul(sn, hij));
        }
    }
    __syncthreads();
    for (int i = m - 1; i >= 0; --i) {
        cx part{0.0, 0.0};
        for (int jj = i + 1 + t; jj < m; jj += 64) part = cadd(part, cmul(Hc[jj * m + i], Y[jj]));
        part.x = wave_sum(part.x);
        part.y = wave_sum(part.y);
        if (t == 0) {
            const cx acc = csub(G[i], part);
            const cx d = Hc[i * m + i];
            const double dd = d.x * d.x + d.y * d.y;
            Y[i] = cx{(acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd};
        }
        __syncthreads();
    }
    cx* Yk = reinterpret_cast<cx*>(y + (size_t)k * 2 * m);
    for (int d = t; d < m; d += 64) Yk[d] = Y[d];
    if (t == 0) {
        const double hm = Hs[(m - 1) * ld + m].x;  // h_{m, m-1}, real and positive
        beta0[2 * k] = -hm * Y[m - 1].x;
        beta0[2 * k + 1] = -hm * Y[m - 1].y;
        errs[k] = sqrt(beta0[2 * k] * beta0[2 * k] + beta0[2 * k + 1] * beta0[2 * k + 1]);
    }
}

unsigned ew_grid(clg_context ctx, long n) {
    long g = (n + 255) / 256;
    return (unsigned)std::max<long>(1, std::min<long>(g, (long)ctx->num_cu * 8));
}
// 16-byte packets of a fermion field (F64: one complex; F32: two complex, so n must be even)
long n_packets(const clg_field_s* f) {
    const long n = n_complex(f);
    return f->prec == CLG_F64 ? n : n / 2;
}
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct OpRef {
    int kind, op;
    const clg_field_s* gauge;
    double param;
    int applies = 0;
    int apply(clg_context ctx, clg_field_s* out, const clg_field_s* in) {
        ++applies;
        return linop_apply(ctx, kind, op, gauge, param, out, in);
    }
};

// ------------------------------------------------------------------ MS-FOM
int msfom(clg_context ctx, OpRef& A, std::vector<clg_field_s*>& X, const std::vector<cx>& shifts,
          const clg_field_s* b, int m, int restart, double accuracy, double blen, int& cycles, double& maxerr) {
    const int ns = (int)shifts.size();
    CLG_CHECK_ARG(m >= 2 && m <= 60, "MS-FOM MaxDim must be in [2, 60]");  // the shifted Hessenberg sits in LDS
    // V[0..m-1] basis, V[m] = w, vnext = v_m seeding the next cycle
    std::vector<clg_field_s*> V(m + 1);
    for (int k = 0; k <= m; ++k) CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 2000 + k, &V[k]));
    clg_field_s* vnext = nullptr;
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 59, &vnext));
    clg_field_s* w = V[m];
    const int ld = m + 1;
    const size_t oH = 0, oY = oH + 2 * (size_t)ld * m, oB = oY + 2 * (size_t)m * ns, oE = oB + 2 * (size_t)ns,
                 oSh = oE + ns, oS = oSh + 2 * (size_t)ns, oA = oS + 16, tot = oA + kArnScratch;
    double* ws = nullptr;
    CLG_TRY(workspace(ctx, tot, &ws));
    double* H = ws + oH;
    double* sc = ws + oS;
    std::vector<double> hb(2 * ns), hsh(2 * ns);
    for (int k = 0; k < ns; ++k) { hsh[2 * k] = shifts[k].x; hsh[2 * k + 1] = shifts[k].y; }
    for (auto x : X) CLG_TRY(clg_field_zero(x));
    std::vector<int> active(ns, 1);
    cycles = 0;
    maxerr = 0;
    // v0 = b / |b|, beta0_k = |b|
    CLG_TRY(blas_norm2_dev(ctx, b, sc));
    CLG_TRY(blas_scale_inv_sqrt(ctx, V[0], b, sc));
    double bn2;
    CLG_TRY(fetch_doubles(ctx, sc, 1, &bn2));
    for (int k = 0; k < ns; ++k) { hb[2 * k] = std::sqrt(bn2); hb[2 * k + 1] = 0.0; }
    // the m Arnoldi steps of a cycle: the same launches on the same storage every cycle (v_m is copied
    // into V[0] between cycles), so after the first cycle, which tunes the launches and grows the
    // scratch, they replay as one recorded graph
    auto arnoldi = [&]() -> int {
        for (int j = 0; j < m; ++j) {
            CLG_TRY(A.apply(ctx, w, V[j]));
            CLG_TRY(arn_step(ctx, V.data(), j + 1, w, H + 2 * (size_t)j * ld, j, false, j < m - 1 ? V[j + 1] : vnext,
                             ws + oA));
        }
        return CLG_OK;
    };
    std::vector<unsigned long long> key = {0x4d53464fULL, (unsigned long long)A.kind, (unsigned long long)A.op,
                                           key_bits(A.param), (unsigned long long)(uintptr_t)A.gauge->data,
                                           (unsigned long long)(uintptr_t)vnext->data, (unsigned long long)(uintptr_t)ws,
                                           (unsigned long long)(uintptr_t)ctx->red_partial, (unsigned long long)m,
                                           (unsigned long long)b->kind, (unsigned long long)b->prec,
                                           (unsigned long long)b->subset, (unsigned long long)b->Vh};
    for (int d = 0; d < 4; ++d) key.push_back((unsigned long long)(long long)b->bc[d]);
    for (int k = 0; k <= m; ++k) key.push_back((unsigned long long)(uintptr_t)V[k]->data);
    for (int cyc = 0; cyc < restart; ++cyc) {
        ++cycles;
        if (cyc == 0) {
            CLG_TRY(arnoldi());
        } else {
            const int a0 = A.applies;
            CLG_TRY(cycle_graph_run(ctx, key, arnoldi));
            A.applies = a0 + m;
        }
        std::vector<int> idx;
        for (int k = 0; k < ns; ++k)
            if (active[k]) idx.push_back(k);
        const int na = (int)idx.size();
        if (na == 0) break;
        std::vector<double> bsh(2 * na), bb(2 * na);
        for (int q = 0; q < na; ++q) {
            bsh[2 * q] = hsh[2 * idx[q]]; bsh[2 * q + 1] = hsh[2 * idx[q] + 1];
            bb[2 * q] = hb[2 * idx[q]]; bb[2 * q + 1] = hb[2 * idx[q] + 1];
        }
        CLG_HIP(hipMemcpyAsync(ws + oSh, bsh.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice, ctx->stream));
        CLG_HIP(hipMemcpyAsync(ws + oB, bb.data(), sizeof(double) * 2 * na, hipMemcpyHostToDevice, ctx->stream));
        k_msfom_lsq<<<na, 64, sizeof(cx) * (size_t)m * m, ctx->stream>>>(na, m, H, ws + oSh, ws + oY, ws + oB, ws + oE);
        CLG_HIP(hipGetLastError());
        ShiftPtrs sp;
        VPtrs vp;
        for (int q = 0; q < na; ++q) { sp.x[q] = X[idx[q]]->data; sp.s[q] = nullptr; }
        for (int j = 0; j < m; ++j) vp.v[j] = V[j]->data;
        const long np = n_packets(b);
        const unsigned g = ew_grid(ctx, np);
        const size_t shm = sizeof(double) * 2 * (size_t)m * na;
        if (b->prec == CLG_F64) k_msfom_update<CLG_F64><<<g, 256, shm, ctx->stream>>>(np, na, m, sp, vp, ws + oY);
        else k_msfom_update<CLG_F32><<<g, 256, shm, ctx->stream>>>(np, na, m, sp, vp, ws + oY);
        CLG_HIP(hipGetLastError());
        // beta0 and the errors in one read: [oB, oB + 2 na) and [oE, oE + na) of the region [oB, oE + ns)
        std::vector<double> be(3 * (size_t)ns);
        CLG_TRY(fetch_doubles(ctx, ws + oB, 3 * ns, be.data()));
        const double* bnew = be.data();
        const double* e = be.data() + (oE - oB);
        maxerr = 0;
        for (int q = 0; q < na; ++q) {
            const int k = idx[q];
            hb[2 * k] = bnew[2 * q];
            hb[2 * k + 1] = bnew[2 * q + 1];
            maxerr = std::max(maxerr, e[q]);
            if (e[q] < accuracy * blen) active[k] = 0;
        }
        if (maxerr < accuracy * blen) return CLG_OK;
        CLG_TRY(clg_field_copy(V[0], vnext));  // v_m seeds the next cycle
    }
    return CLG_NOT_CONVERGED;
}

// ------------------------------------------------------------------ MS-GMRES (CMultiShiftGMRES.cpp:112-330)
// Restarted GMRES on the unshifted system; each shifted residual is kept collinear with the
// unshifted one, r_n = beta_n r, by solving the square system [Hbar + c_n I~ | z] (y_n, beta_n') =
// beta_n beta e_1 with z = beta e_1 - Hbar y (Frommer-Glaessner).  Stop at the end of a cycle when
// every shifted residual |beta_n'| |z| < Accuracy * (AbsoluteAccuracy ? 1 : |b|^2).  (The reference
// compares the collinearity factor |beta_n'| alone, which does not bound the residual and grows as
// the unshifted system converges; its own test of this solver is disabled for that reason,
// TestMultiShiftSolver.cpp:318.)
int msgmres(clg_context ctx, OpRef& A, std::vector<clg_field_s*>& X, const std::vector<cx>& shifts,
            const clg_field_s* b, int m, int restart, double accuracy, double blen, int& cycles, double& maxerr) {
    using clg::sla::Mat;
    typedef std::complex<double> zc;
    const int ns = (int)shifts.size();
    CLG_CHECK_ARG(m >= 2 && m <= 63, "MS-GMRES MaxDim must be in [2, 63]");
    std::vector<clg_field_s*> V(m + 1);
    for (int k = 0; k <= m; ++k) CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 2000 + k, &V[k]));
    clg_field_s *xb = nullptr, *w = V[m];
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 2100, &xb));
    const int ld = m + 1;
    const size_t oH = 0, oY = oH + 2 * (size_t)ld * m, oS = oY + 2 * (size_t)m * (ns + 1), oA = oS + 16,
                 tot = oA + kArnScratch;
    double* ws = nullptr;
    CLG_TRY(workspace(ctx, tot, &ws));
    double* H = ws + oH;
    double* sc = ws + oS;
    CLG_TRY(clg_field_zero(xb));
    for (auto x : X) CLG_TRY(clg_field_zero(x));
    std::vector<zc> beta0(ns, zc(1.0, 0.0));
    cycles = 0;
    maxerr = 0;
    for (int cyc = 0; cyc < restart; ++cyc) {
        ++cycles;
        if (cyc == 0) {
            CLG_TRY(blas_axpby(ctx, V[0], 1.0, 0.0, b, 0.0, 0.0));
        } else {
            CLG_TRY(A.apply(ctx, V[0], xb));
            CLG_TRY(blas_axpby(ctx, V[0], 1.0, 0.0, b, -1.0, 0.0));
        }
        CLG_TRY(blas_norm2_dev(ctx, V[0], sc));
        double b2;
        CLG_TRY(fetch_doubles(ctx, sc, 1, &b2));
        const double beta = std::sqrt(b2);
        CLG_TRY(blas_scale_inv_sqrt(ctx, V[0], V[0], sc));
        for (int j = 0; j < m; ++j) {
            CLG_TRY(A.apply(ctx, w, V[j]));
            CLG_TRY(arn_step(ctx, V.data(), j + 1, w, H + 2 * (size_t)j * ld, j, false, j < m - 1 ? V[j + 1] : nullptr,
                             ws + oA));
        }
        std::vector<double> hh(2 * (size_t)ld * m);
        CLG_TRY(fetch_doubles(ctx, H, (int)hh.size(), hh.data()));
        Mat Hb(m + 1, m);
        // columns hold rows 0..j+1 only; the device entries below the sub-diagonal are stale
        for (int j = 0; j < m; ++j)
            for (int i = 0; i <= j + 1; ++i) Hb(i, j) = zc(hh[2 * (j * ld + i)], hh[2 * (j * ld + i) + 1]);
        // unshifted least squares, x_base += V y, z = beta e_1 - Hbar y
        std::vector<zc> rhs(m + 1, zc(0.0, 0.0)), y;
        rhs[0] = beta;
        clg::sla::hessenberg_lsq(Hb, rhs, y);
        std::vector<zc> z(m + 1, zc(0.0, 0.0));
        z[0] = beta;
        for (int i = 0; i <= m; ++i)
            for (int j = 0; j < m; ++j) z[i] -= Hb(i, j) * y[j];
        std::vector<double> up(2 * (size_t)m * (ns + 1));
        for (int j = 0; j < m; ++j) { up[2 * j] = y[j].real(); up[2 * j + 1] = y[j].imag(); }
        // shifted systems
        maxerr = 0;
        for (int n = 0; n < ns; ++n) {
            Mat M(m + 1, m + 1);
            for (int i = 0; i <= m; ++i) {
                for (int j = 0; j < m; ++j) M(i, j) = Hb(i, j) + (i == j ? zc(shifts[n].x, shifts[n].y) : zc(0.0, 0.0));
                M(i, m) = z[i];
            }
            Mat Q, R;
            clg::sla::thin_qr(M, Q, R);
            std::vector<zc> g(m + 1, zc(0.0, 0.0)), yh(m + 1, zc(0.0, 0.0));
            const zc g0 = beta0[n] * beta;
            for (int i = 0; i <= m; ++i) g[i] = std::conj(Q(0, i)) * g0;  // Q^H (g0 e_1)
            for (int i = m; i >= 0; --i) {
                zc acc = g[i];
                for (int l = i + 1; l <= m; ++l) acc -= R(i, l) * yh[l];
                yh[i] = acc / R(i, i);
            }
            beta0[n] = yh[m];
            double zn = 0;
            for (int i = 0; i <= m; ++i) zn += std::norm(z[i]);
            maxerr = std::max(maxerr, std::abs(yh[m]) * std::sqrt(zn));
            for (int j = 0; j < m; ++j) {
                up[2 * (size_t)m * (n + 1) + 2 * j] = yh[j].real();
                up[2 * (size_t)m * (n + 1) + 2 * j + 1] = yh[j].imag();
            }
        }
        CLG_HIP(hipMemcpyAsync(ws + oY, up.data(), sizeof(double) * up.size(), hipMemcpyHostToDevice, ctx->stream));
        ShiftPtrs sp;
        VPtrs vp;
        sp.x[0] = xb->data;
        sp.s[0] = nullptr;
        for (int n = 0; n < ns; ++n) { sp.x[n + 1] = X[n]->data; sp.s[n + 1] = nullptr; }
        for (int j = 0; j < m; ++j) vp.v[j] = V[j]->data;
        const long np = n_packets(b);
        const unsigned gg = ew_grid(ctx, np);
        const size_t shm = sizeof(double) * 2 * (size_t)m * (ns + 1);
        if (b->prec == CLG_F64) k_msfom_update<CLG_F64><<<gg, 256, shm, ctx->stream>>>(np, ns + 1, m, sp, vp, ws + oY);
        else k_msfom_update<CLG_F32><<<gg, 256, shm, ctx->stream>>>(np, ns + 1, m, sp, vp, ws + oY);
        CLG_HIP(hipGetLastError());
        CLG_HIP(hipStreamSynchronize(ctx->stream));  // up is a host temporary
        if (maxerr < accuracy * blen) return CLG_OK;
    }
    return CLG_NOT_CONVERGED;
}

enum { MS_FOM = 0, MS_BICGSTAB = 1, MS_GMRES = 2, MS_NESTED = 3 };

int run_multishift(clg_context ctx, const clg_solver_params* p, int kind_solver, int prec, std::vector<clg_field_s*>& X,
                   const std::vector<cx>& shifts, const clg_field_s* b, const clg_field_s* gauge, double param, int op,
                   clg_solve_report* rep) {
    const double t0 = now_s();
    // working precision: FP64 as given, or FP32 copies of b / solutions with FP32 (prec CLG_F32) or FP16
    // 12-real (prec CLG_F16) copies of the links
    const clg_field_s* bw = b;
    const clg_field_s* gw = gauge;
    std::vector<clg_field_s*> Xw = X;
    if (prec == CLG_F32 || prec == CLG_F16) {
        clg_field_s *b32, *g32;
        CLG_TRY(scratch_field(ctx, b, b->kind, CLG_F32, b->subset, 150, &b32));
        CLG_TRY(clg_field_copy(b32, const_cast<clg_field_s*>(b)));
        if (prec == CLG_F16) CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE12, CLG_F16, CLG_FULL, 0, &g32));
        else CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE, CLG_F32, CLG_FULL, 0, &g32));
        CLG_TRY(gauge_convert(ctx, g32, gauge));
        for (size_t k = 0; k < X.size(); ++k)
            CLG_TRY(scratch_field(ctx, b, b->kind, CLG_F32, b->subset, 4000 + (int)k, &Xw[k]));
        bw = b32;
        gw = g32;
    }
    OpRef A{p->kind, op, gw, param};
    double bn2 = 1.0;
    if (!p->absolute_accuracy) {
        CLG_TRY(blas_norm2_dev(ctx, b, ctx->red_result + 24));
        CLG_TRY(fetch_doubles(ctx, ctx->red_result + 24, 1, &bn2));
    }
    int it = 0;
    double err = 0;
    int st;
    if (kind_solver == MS_GMRES) {
        // CMultiShiftGMRES compares max |beta_n| with Accuracy * (AbsoluteAccuracy ? 1 : |b|^2)
        st = msgmres(ctx, A, Xw, shifts, bw, p->max_dim, p->restart, p->accuracy, p->absolute_accuracy ? 1.0 : bn2, it,
                     err);
    } else if (kind_solver == MS_FOM) {
        // FOM compares |beta0| with Accuracy * (AbsoluteAccuracy ? 1 : |b|^2) (CMultiShiftFOM.cpp:119-124)
        st = msfom(ctx, A, Xw, shifts, bw, p->max_dim, p->restart, p->accuracy, p->absolute_accuracy ? 1.0 : bn2, it,
                   err);
    } else {
        // BiCGStab compares |s| with Accuracy * (AbsoluteAccuracy ? 1 : |b|)
        std::vector<double> sh(2 * shifts.size());
        for (size_t k = 0; k < shifts.size(); ++k) { sh[2 * k] = shifts[k].x; sh[2 * k + 1] = shifts[k].y; }
        st = msbicgstab_dev(ctx, A.kind, A.op, A.gauge, A.param, Xw, sh, bw, p->max_step, p->deviation_step,
                            p->accuracy, p->absolute_accuracy ? 1.0 : std::sqrt(bn2), it, err, A.applies);
    }
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    if (prec != CLG_F64)
        for (size_t k = 0; k < X.size(); ++k) CLG_TRY(clg_field_copy(X[k], Xw[k]));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->converged = st == CLG_OK;
        rep->cycles = it;
        rep->outer = 1;
        rep->operator_applies = A.applies;
        rep->last_deviation = err;
        rep->true_residual = -1.0;
        rep->seconds = now_s() - t0;
    }
    return st;
}

}  // namespace

extern "C" {

int clg_multishift_solve(clg_context ctx, const clg_solver_params* p, int kind_solver, int prec, clg_field* xs,
                         const double* shifts, int nshift, clg_field b, clg_field gauge, double param, int op,
                         clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && p && xs && shifts && b && gauge, "null argument");
    CLG_CHECK_ARG(nshift >= 1 && nshift <= kMaxShift, "nshift must be in [1, %d]", kMaxShift);
    CLG_CHECK_ARG(kind_solver >= MS_FOM && kind_solver <= MS_NESTED, "unknown multi-shift solver %d", kind_solver);
    CLG_CHECK_ARG(kind_solver != MS_GMRES || nshift < kMaxShift, "MS-GMRES takes at most %d shifts", kMaxShift - 1);
    CLG_CHECK_ARG(prec == CLG_F64 || prec == CLG_F32 || prec == CLG_F16,
                  "multi-shift precision must be FP64, FP32, or FP16 links under FP32 vectors");
    CLG_CHECK_ARG(prec != CLG_F16 || kind_solver != MS_NESTED, "MS-Nested runs in FP64 or FP32");
    CLG_CHECK_ARG(b->kind == p->kind && b->subset == CLG_FULL && b->prec == CLG_F64, "b must be a FULL FP64 field of params.kind");
    CLG_CHECK_ARG(gauge->prec == CLG_F64 && gauge->kind == CLG_GAUGE, "gauge must be the FP64 18-real field");
    CLG_CHECK_ARG(op >= EFO_F_D && op <= EFO_F_DDdagger, "operator %d not supported", op);
    std::vector<clg_field_s*> X(nshift);
    std::vector<cx> sh(nshift);
    for (int k = 0; k < nshift; ++k) {
        CLG_CHECK_ARG(xs[k] && xs[k]->kind == b->kind && xs[k]->subset == CLG_FULL && xs[k]->prec == CLG_F64,
                      "solution %d must match b", k);
        CLG_CHECK_ARG(xs[k]->data != b->data, "solutions must not alias b (CMultiShiftSolver.h:35)");
        CLG_TRY(check_same_lattice(xs[k], b));
        X[k] = xs[k];
        sh[k] = cx{shifts[2 * k], shifts[2 * k + 1]};
    }
    CLG_TRY(check_same_lattice(b, gauge));
    (void)hipSetDevice(ctx->device);
    if (kind_solver == MS_NESTED) {
        // CMultiShiftNested::Solve (CMultiShiftNested.cpp:69-150): staggered fermions and real shifts
        // only; one solve of the operator with the shifted mass per shift (D: m + c, DD^dagger:
        // sqrt(m^2 + c)), params acting as the nested solver's
        CLG_CHECK_ARG(p->kind == CLG_KS, "MS-Nested supports staggered fermions only");
        CLG_CHECK_ARG(op == EFO_F_D || op == EFO_F_Ddagger || op == EFO_F_DDdagger, "MS-Nested operator not supported");
        const double t0 = now_s();
        clg_solver_params q = *p;
        if (prec == CLG_F32 && q.inner_prec == CLG_F64) q.inner_prec = CLG_F32;
        int cycles = 0, applies = 0, st_all = CLG_OK;
        double worst = 0;
        for (int k = 0; k < nshift; ++k) {
            CLG_CHECK_ARG(std::fabs(sh[k].y) <= 1e-7, "MS-Nested takes real shifts only");
            const double m2 = op == EFO_F_DDdagger ? param * param + sh[k].x : param + sh[k].x;
            CLG_CHECK_ARG(m2 >= 1e-7, "MS-Nested needs mass + shift > 0");
            const double mass = op == EFO_F_DDdagger ? std::sqrt(m2) : m2;
            clg_solve_report r{};
            const int st = clg_solve(ctx, &q, X[k], b, gauge, mass, op + 4, nullptr, &r);
            if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
            if (st != CLG_OK) st_all = CLG_NOT_CONVERGED;
            cycles += r.cycles;
            applies += r.operator_applies;
            worst = std::max(worst, r.true_residual);
        }
        if (rep) {
            std::memset(rep, 0, sizeof(*rep));
            rep->converged = st_all == CLG_OK;
            rep->cycles = cycles;
            rep->outer = nshift;
            rep->operator_applies = applies;
            rep->true_residual = worst;
            rep->seconds = now_s() - t0;
        }
        return st_all;
    }
    return run_multishift(ctx, p, kind_solver, prec, X, sh, b, gauge, param, op, rep);
}

int clg_rational_apply(clg_context ctx, const clg_solver_params* p, int kind_solver, int prec, clg_field out,
                       clg_field in, clg_field gauge, double param, int op, const double* coeffs, int degree,
                       clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && p && out && in && gauge && coeffs, "null argument");
    CLG_CHECK_ARG(degree >= 1 && degree <= kMaxShift, "degree must be in [1, %d]", kMaxShift);
    CLG_CHECK_ARG(out->kind == in->kind && out->prec == CLG_F64 && out->subset == CLG_FULL, "out must match in");
    (void)hipSetDevice(ctx->device);
    std::vector<clg_field_s*> X(degree);
    std::vector<double> sh(2 * degree);
    for (int k = 0; k < degree; ++k) {
        CLG_TRY(scratch_field(ctx, in, in->kind, CLG_F64, CLG_FULL, 5000 + k, &X[k]));
        sh[2 * k] = coeffs[1 + degree + k];
        sh[2 * k + 1] = 0.0;
    }
    std::vector<clg_field> xs(X.begin(), X.end());
    const int st = clg_multishift_solve(ctx, p, kind_solver, prec, xs.data(), sh.data(), degree, in, gauge, param, op, rep);
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    // out = c in + sum_k a_k x_k, accumulated in FP64 (CField.cpp:140-146)
    double* ws = nullptr;
    CLG_TRY(workspace(ctx, 2 * (size_t)degree, &ws));
    std::vector<double> a(2 * degree);
    for (int k = 0; k < degree; ++k) { a[2 * k] = coeffs[1 + k]; a[2 * k + 1] = 0.0; }
    CLG_HIP(hipMemcpyAsync(ws, a.data(), sizeof(double) * 2 * degree, hipMemcpyHostToDevice, ctx->stream));
    if (out != in) CLG_TRY(clg_field_copy(out, in));
    CLG_TRY(blas_axpby(ctx, out, 0.0, 0.0, out, coeffs[0], 0.0));
    CLG_TRY(blas_multiaxpy_dev(ctx, out, X.data(), degree, ws, 1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    return st;
}

}  // extern "C"
