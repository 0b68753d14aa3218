// clg_msbicg.hip -- the shifted BiCGStab of the multi-shift solvers (CMultiShiftBiCGStab::Solve,
// CMultiShiftBiCGStab.cpp:73-270) with every scalar of the recurrences on the device.
//
// The reference keeps delta, phi, beta, alpha, chi and the per-shift zeta / beta / alpha / chi / rho
// recurrences on the host and reads four inner products back per step.  Here a step is a fixed
// sequence of launches with no host wait (state: kBcgState doubles of the context workspace):
//   W  = R + beta SA                                     k_bcg_w             (:166-167)
//   WA = A W                                             operator            (:168-169)
//   partials of <WA, W> and |WA|^2                       k_bcg_dot<P, true>
//   chi; per shift chi_k, the first update's coefficients, rho_k        fin (:170-190)
//   per active shift x_k += -beta_k s_k + (chi rho zeta)_k W,
//     s_k += -(chi rho zeta / beta)_k W + (chi rho zeta_old / beta)_k R;
//   R = W - chi WA; partials of <W0, R>                  k_bcg_upd1          (:182-194)
//   delta', alpha, delta; per shift alpha_k, the second update's coefficients   fin (:195-206)
//   S = alpha (S - chi SA) + R                           k_bcg_upd2          (:212-214)
//   SA = A S                                             operator            (:216-217)
//   partials of <W0, SA>                                 k_bcg_dot<P, false>
//   phi = <W0, SA> / delta; the next step's beta and per-shift zeta, beta_k      fin (:218, :133-164)
// so DiviationStep steps replay as one recorded graph.  The check after them (|S| and |s_k| of the
// shifts still active, :220-242) is one more reduction and the only read back to the host; a shift
// whose |s_k| fell below the bound leaves the active set on the device.  The per-shift vector updates
// of a step run for all active shifts in one launch (the shared vectors are read once per element),
// each lane loading its shifts' x_k and s_k in groups of 8 before it stores any of them.  The second
// per-shift update, s_k = alpha_k s_k + (zeta rho)_k R (:208-209), is not stored: the next step's first
// update and the check apply it as they read s_k (R is still the vector it names then), so s_k is
// streamed once per step instead of twice.  Inner
// products and all scalars are FP64; vectors are stored in the working precision (FP64 or FP32).
#include "clg_internal.h"
#include "clg_packets.h"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int blas_conj(clg_context, clg_field_s*, const clg_field_s*);
int fetch_doubles(clg_context, const double*, int, double*);
int workspace(clg_context, size_t, double**);
int pk_kind(const clg_field_s* f, long& npk);
int linop_apply(clg_context ctx, int kind, int op, const clg_field_s* gauge, double param, clg_field_s* out,
                const clg_field_s* in);
}  // namespace clg

namespace {

constexpr int kBcgMaxShift = 32;
constexpr int kBcgBlock = 256;
constexpr int kBcgMaxGrid = 1024;
constexpr int kBcgGroup = 8;  // shifts whose x_k, s_k a lane holds at once
constexpr int kBcgNormRows = kBcgMaxShift + 1;

// device state (doubles): globals, then per-shift arrays of kBcgMaxShift entries
enum : int {
    kDelta = 0, kPhi = 2, kBeta = 4, kAlpha = 6, kChi = 8, kMaxErr = 10, kDone = 11,
    oCn = 16,
    oZeta = oCn + 2 * kBcgMaxShift,
    oZold = oZeta + 2 * kBcgMaxShift,
    oBetas = oZold + 2 * kBcgMaxShift,
    oRho = oBetas + 2 * kBcgMaxShift,
    oC8 = oRho + 2 * kBcgMaxShift,
    oC4 = oC8 + 8 * kBcgMaxShift,
    oSl = oC4 + 4 * kBcgMaxShift,
    oAct = oSl + kBcgMaxShift,
    kBcgState = oAct + kBcgMaxShift
};

struct dcx {
    double x, y;
};
__device__ __forceinline__ dcx ld2(const double* p) { return dcx{p[0], p[1]}; }
__device__ __forceinline__ void st2(double* p, dcx a) { p[0] = a.x; p[1] = a.y; }
__device__ __forceinline__ dcx cmul(dcx a, dcx b) { return dcx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ dcx cdiv(dcx a, dcx b) {
    const double d = b.x * b.x + b.y * b.y;
    return dcx{(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ __forceinline__ dcx csub(dcx a, dcx b) { return dcx{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ dcx cadd(dcx a, dcx b) { return dcx{a.x + b.x, a.y + b.y}; }

// acc += c v for NC complex numbers
template <int NC>
__device__ __forceinline__ void cmad(double* acc, dcx c, const double* v) {
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        acc[2 * q] += c.x * v[2 * q] - c.y * v[2 * q + 1];
        acc[2 * q + 1] += c.x * v[2 * q + 1] + c.y * v[2 * q];
    }
}
// acc += conj(a) b
template <int NC>
__device__ __forceinline__ void cdot(double* acc, const double* a, const double* b) {
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        acc[0] += a[2 * q] * b[2 * q] + a[2 * q + 1] * b[2 * q + 1];
        acc[1] += a[2 * q] * b[2 * q + 1] - a[2 * q + 1] * b[2 * q];
    }
}

template <int NR>
__device__ void bcg_partial(double (&v)[NR], double* part) {
    __shared__ double sm[kBcgBlock / 64][NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = v[r];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < NR; r += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < kBcgBlock / 64; ++q) s += sm[q][r];
        part[(size_t)blockIdx.x * NR + r] = s;
    }
}

struct BcgShifts {
    void* x[kBcgMaxShift];
    void* s[kBcgMaxShift];
};

// W = R + beta SA
template <int P>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_w(long n, const double* __restrict__ st,
                                                     const typename Pk<P>::T* __restrict__ R,
                                                     const typename Pk<P>::T* __restrict__ SA,
                                                     typename Pk<P>::T* __restrict__ W) {
    constexpr int NC = Pk<P>::NC;
    const dcx beta = ld2(st + kBeta);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double r[2 * NC], sa[2 * NC];
        Pk<P>::ld(R, i, r);
        Pk<P>::ld(SA, i, sa);
        cmad<NC>(r, beta, sa);
        Pk<P>::st(W, i, r);
    }
}

// partials of <A, B> (2 rows) and, with NORM, |A|^2 (a third)
template <int P, bool NORM>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_dot(long n, const typename Pk<P>::T* __restrict__ A,
                                                       const typename Pk<P>::T* __restrict__ B, double* part) {
    constexpr int NC = Pk<P>::NC;
    constexpr int NR = NORM ? 3 : 2;
    double v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double a[2 * NC], b[2 * NC];
        Pk<P>::ld(A, i, a);
        Pk<P>::ld(B, i, b);
        cdot<NC>(v, a, b);
        if (NORM) {
#pragma unroll
            for (int q = 0; q < 2 * NC; ++q) v[NR - 1] += a[q] * a[q];
        }
    }
    bcg_partial<NR>(v, part);
}

// the first update of every active shift, R = W - chi WA, partials of <W0, R>.  s_k is stored as it was
// before the previous step's second update s_k <- alpha_k s_k + (zeta rho)_k R (coefficients c4, R the
// vector this kernel reads before it overwrites it), which is applied here on the fly.
template <int P>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_upd1(long n, int ns, BcgShifts sp, const double* __restrict__ st,
                                                        const typename Pk<P>::T* __restrict__ W,
                                                        const typename Pk<P>::T* __restrict__ WA,
                                                        const typename Pk<P>::T* __restrict__ W0,
                                                        typename Pk<P>::T* __restrict__ R, double* part) {
    typedef typename Pk<P>::T T;
    constexpr int NC = Pk<P>::NC;
    __shared__ double cs[8 * kBcgMaxShift];
    __shared__ double c2[4 * kBcgMaxShift];
    __shared__ int act[kBcgMaxShift];
    for (int t = threadIdx.x; t < 8 * ns; t += blockDim.x) cs[t] = st[oC8 + t];
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) c2[t] = st[oC4 + t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) act[t] = st[oAct + t] != 0.0;
    __syncthreads();
    const dcx mchi = {-st[kChi], -st[kChi + 1]};
    double v[2] = {0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double w[2 * NC], r[2 * NC];
        Pk<P>::ld(W, i, w);
        Pk<P>::ld(R, i, r);
        for (int k0 = 0; k0 < ns; k0 += kBcgGroup) {
            T xr[kBcgGroup], sr[kBcgGroup];
#pragma unroll
            for (int g = 0; g < kBcgGroup; ++g) {
                const int k = k0 + g;
                if (k < ns && act[k]) {
                    xr[g] = ((const T*)sp.x[k])[i];
                    sr[g] = ((const T*)sp.s[k])[i];
                }
            }
#pragma unroll
            for (int g = 0; g < kBcgGroup; ++g) {
                const int k = k0 + g;
                if (k < ns && act[k]) {
                    const double* c = cs + 8 * k;
                    const double* d = c2 + 4 * k;
                    double x[2 * NC], s0[2 * NC], s[2 * NC], sn[2 * NC];
                    Pk<P>::cvt(xr[g], x);
                    Pk<P>::cvt(sr[g], s0);
#pragma unroll
                    for (int q = 0; q < 2 * NC; ++q) s[q] = 0.0;
                    cmad<NC>(s, dcx{d[0], d[1]}, s0);
                    cmad<NC>(s, dcx{d[2], d[3]}, r);
#pragma unroll
                    for (int q = 0; q < 2 * NC; ++q) sn[q] = s[q];
                    cmad<NC>(x, dcx{c[0], c[1]}, s);
                    cmad<NC>(x, dcx{c[2], c[3]}, w);
                    cmad<NC>(sn, dcx{c[4], c[5]}, w);
                    cmad<NC>(sn, dcx{c[6], c[7]}, r);
                    Pk<P>::st((T*)sp.x[k], i, x);
                    Pk<P>::st((T*)sp.s[k], i, sn);
                }
            }
        }
        double wa[2 * NC], w0[2 * NC];
        Pk<P>::ld(WA, i, wa);
        Pk<P>::ld(W0, i, w0);
        cmad<NC>(w, mchi, wa);
        Pk<P>::st(R, i, w);  // w keeps the stored R
        cdot<NC>(v, w0, w);
    }
    bcg_partial<2>(v, part);
}

// S = alpha (S - chi SA) + R
template <int P>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_upd2(long n, const double* __restrict__ st,
                                                        const typename Pk<P>::T* __restrict__ R,
                                                        const typename Pk<P>::T* __restrict__ SA,
                                                        typename Pk<P>::T* __restrict__ S) {
    constexpr int NC = Pk<P>::NC;
    const dcx mchi = {-st[kChi], -st[kChi + 1]};
    const dcx alpha = ld2(st + kAlpha);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double r[2 * NC], s[2 * NC], sa[2 * NC];
        Pk<P>::ld(R, i, r);
        Pk<P>::ld(S, i, s);
        Pk<P>::ld(SA, i, sa);
        cmad<NC>(s, mchi, sa);
        cmad<NC>(r, alpha, s);
        Pk<P>::st(S, i, r);
    }
}

// the check's partials: row 0 |S|^2, row 1 + k |s_k|^2 of the active shifts (0 for the others), s_k with
// its pending second update applied
template <int P>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_norms(long n, int ns, BcgShifts sp, const double* __restrict__ st,
                                                         const typename Pk<P>::T* __restrict__ S,
                                                         const typename Pk<P>::T* __restrict__ R, double* part) {
    typedef typename Pk<P>::T T;
    constexpr int NC = Pk<P>::NC;
    __shared__ double c2[4 * kBcgMaxShift];
    __shared__ int act[kBcgMaxShift];
    for (int t = threadIdx.x; t < 4 * ns; t += blockDim.x) c2[t] = st[oC4 + t];
    for (int t = threadIdx.x; t < ns; t += blockDim.x) act[t] = st[oAct + t] != 0.0;
    __syncthreads();
    double v[kBcgNormRows];
#pragma unroll
    for (int r = 0; r < kBcgNormRows; ++r) v[r] = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double a[2 * NC], r[2 * NC];
        Pk<P>::ld(S, i, a);
        Pk<P>::ld(R, i, r);
#pragma unroll
        for (int q = 0; q < 2 * NC; ++q) v[0] += a[q] * a[q];
#pragma unroll
        for (int k = 0; k < kBcgMaxShift; ++k) {
            if (k < ns && act[k]) {
                double s0[2 * NC], s[2 * NC];
                Pk<P>::ld((const T*)sp.s[k], i, s0);
#pragma unroll
                for (int q = 0; q < 2 * NC; ++q) s[q] = 0.0;
                cmad<NC>(s, dcx{c2[4 * k], c2[4 * k + 1]}, s0);
                cmad<NC>(s, dcx{c2[4 * k + 2], c2[4 * k + 3]}, r);
#pragma unroll
                for (int q = 0; q < 2 * NC; ++q) v[1 + k] += s[q] * s[q];
            }
        }
    }
    bcg_partial<kBcgNormRows>(v, part);
}

enum { FIN_INIT = 0, FIN_PHI = 1, FIN_CHI = 2, FIN_ALPHA = 3, FIN_CHECK = 4 };

// totals of the G workgroups' partials (row-major part[b * NR + r], summed in a fixed order), then
// the scalar stage.  One workgroup: thread 0 the global scalars, thread k < ns shift k.
template <int NR, int STAGE>
__global__ __launch_bounds__(kBcgBlock) void k_bcg_fin(const double* __restrict__ part, int G, double* st, int ns,
                                                       double bound, int first) {
    __shared__ double sm[kBcgBlock / 64][NR];
    __shared__ double tot[NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, t = threadIdx.x;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (int b = t; b < G; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += part[(size_t)b * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = wave_sum(acc[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = acc[r];
    }
    __syncthreads();
    if (t < NR) {
        double s = 0.0;
        for (int q = 0; q < kBcgBlock / 64; ++q) s += sm[q][t];
        tot[t] = s;
    }
    __syncthreads();
    const bool mine = t < ns;
    const bool active = mine && st[oAct + t] != 0.0;
    if (STAGE == FIN_INIT) {
        // delta = <W0, R>; per shift zeta = zeta_old = beta_k = rho = 1, |s_k| = 1 (:88-97, :124)
        if (t == 0) {
            st[kDelta] = tot[0];
            st[kDelta + 1] = tot[1];
            st[kBeta] = st[kBeta + 1] = st[kAlpha] = st[kAlpha + 1] = 0.0;
        }
        if (mine) {
            st2(st + oZeta + 2 * t, dcx{1.0, 0.0});
            st2(st + oZold + 2 * t, dcx{1.0, 0.0});
            st2(st + oBetas + 2 * t, dcx{1.0, 0.0});
            st2(st + oRho + 2 * t, dcx{1.0, 0.0});
            st[oSl + t] = 1.0;
            st[oAct + t] = (1.0 < bound) ? 0.0 : 1.0;
            double* c = st + oC4 + 4 * t;  // no second update pending before the first step
            c[0] = 1.0;
            c[1] = c[2] = c[3] = 0.0;
        }
    } else if (STAGE == FIN_PHI) {
        // phi = <W0, SA> / delta (:218); the next step's newbeta = -1 / phi and per-shift zeta, beta_k
        // (:133-164), then beta = newbeta
        const dcx phi = cdiv(dcx{tot[0], tot[1]}, ld2(st + kDelta));
        const dcx nb = cdiv(dcx{-1.0, 0.0}, phi);
        if (active) {
            const dcx cn = ld2(st + oCn + 2 * t);
            const dcx one_cnb = csub(dcx{1.0, 0.0}, cmul(cn, nb));
            if (first) {
                const dcx z = cdiv(dcx{1.0, 0.0}, one_cnb);
                st2(st + oZeta + 2 * t, z);
                st2(st + oBetas + 2 * t, cmul(nb, z));
            } else {
                const dcx beta = ld2(st + kBeta), alpha = ld2(st + kAlpha);
                const dcx znm1 = ld2(st + oZold + 2 * t);
                const dcx zold = ld2(st + oZeta + 2 * t);
                const dcx zb = cmul(znm1, beta);
                const dcx z = cdiv(cmul(zold, zb), cadd(cmul(alpha, cmul(nb, csub(znm1, zold))), cmul(zb, one_cnb)));
                st2(st + oZold + 2 * t, zold);
                st2(st + oZeta + 2 * t, z);
                st2(st + oBetas + 2 * t, cdiv(cmul(nb, z), zold));
            }
        }
        __syncthreads();  // every shift has read beta
        if (t == 0) {
            st2(st + kPhi, phi);
            st2(st + kBeta, nb);
        }
    } else if (STAGE == FIN_CHI) {
        // chi = <WA, W> / |WA|^2 (:170); per shift (:178-189)
        const dcx chi = {tot[0] / tot[2], tot[1] / tot[2]};
        if (t == 0) st2(st + kChi, chi);
        if (active) {
            const dcx cn = ld2(st + oCn + 2 * t);
            const dcx dinom = cadd(dcx{1.0, 0.0}, cmul(cn, chi));
            const dcx chis = cdiv(chi, dinom);
            const dcx rho = ld2(st + oRho + 2 * t);
            const dcx betas = ld2(st + oBetas + 2 * t);
            const dcx chirho = cmul(chis, rho);
            const dcx crz = cmul(chirho, ld2(st + oZeta + 2 * t));
            const dcx crzb = cdiv(crz, betas);
            const dcx crzob = cdiv(cmul(chirho, ld2(st + oZold + 2 * t)), betas);
            double* c = st + oC8 + 8 * t;
            c[0] = -betas.x;
            c[1] = -betas.y;
            c[2] = crz.x;
            c[3] = crz.y;
            c[4] = -crzb.x;
            c[5] = -crzb.y;
            c[6] = crzob.x;
            c[7] = crzob.y;
            st2(st + oRho + 2 * t, cdiv(rho, dinom));
        }
    } else if (STAGE == FIN_ALPHA) {
        // delta' = <W0, R>, alpha = beta (-delta') / (chi delta), delta = delta' (:194-196); per shift (:205-209)
        const dcx nd = {tot[0], tot[1]};
        const dcx beta = ld2(st + kBeta), chi = ld2(st + kChi);
        const dcx alpha = cdiv(cmul(beta, dcx{-nd.x, -nd.y}), cmul(chi, ld2(st + kDelta)));
        if (active) {
            const dcx z = ld2(st + oZeta + 2 * t);
            const dcx as = cdiv(cmul(cmul(alpha, z), ld2(st + oBetas + 2 * t)), cmul(ld2(st + oZold + 2 * t), beta));
            const dcx zr = cmul(z, ld2(st + oRho + 2 * t));
            double* c = st + oC4 + 4 * t;
            c[0] = as.x;
            c[1] = as.y;
            c[2] = zr.x;
            c[3] = zr.y;
        }
        __syncthreads();  // every shift has read delta
        if (t == 0) {
            st2(st + kAlpha, alpha);
            st2(st + kDelta, nd);
        }
    } else {
        // the check (:220-242): |S| and the active shifts' |s_k|; a shift below the bound leaves the set
        __shared__ double mx[kBcgMaxShift + 1];
        if (t <= kBcgMaxShift) mx[t] = 0.0;
        __syncthreads();
        if (active) {
            const double sl = sqrt(tot[1 + t]);
            st[oSl + t] = sl;
            mx[1 + t] = sl;
            st[oAct + t] = (sl < bound) ? 0.0 : 1.0;
        }
        __syncthreads();
        if (t == 0) {
            double m = sqrt(tot[0]);
            for (int k = 0; k < ns; ++k) m = fmax(m, mx[1 + k]);
            st[kMaxErr] = m;
            st[kDone] = (m < bound) ? 1.0 : 0.0;
        }
    }
}

template <int P>
struct BcgRun {
    typedef typename Pk<P>::T T;
    clg_context ctx;
    long n;
    unsigned G;
    int ns;
    BcgShifts sp;
    double* st;
    double* part;
    double bound;
    int kind, op;
    const clg_field_s* gauge;
    double param;
    clg_field_s *R, *W, *W0, *S, *SA, *WA;
    int applies = 0;

    static const T* c(const clg_field_s* f) { return (const T*)f->data; }
    static T* m(clg_field_s* f) { return (T*)f->data; }
    int apply(clg_field_s* out, const clg_field_s* in) {
        ++applies;
        return linop_apply(ctx, kind, op, gauge, param, out, in);
    }
    template <int NR, int STAGE>
    void fin(int first = 0) {
        k_bcg_fin<NR, STAGE><<<1, kBcgBlock, 0, ctx->stream>>>(part, (int)G, st, ns, bound, first);
    }
    // one step of the recurrences: launches only
    int step() {
        hipStream_t s = ctx->stream;
        k_bcg_w<P><<<G, kBcgBlock, 0, s>>>(n, st, c(R), c(SA), m(W));
        CLG_TRY(apply(WA, W));
        k_bcg_dot<P, true><<<G, kBcgBlock, 0, s>>>(n, c(WA), c(W), part);
        fin<3, FIN_CHI>();
        k_bcg_upd1<P><<<G, kBcgBlock, 0, s>>>(n, ns, sp, st, c(W), c(WA), c(W0), m(R), part);
        fin<2, FIN_ALPHA>();
        k_bcg_upd2<P><<<G, kBcgBlock, 0, s>>>(n, st, c(R), c(SA), m(S));
        CLG_TRY(apply(SA, S));
        k_bcg_dot<P, false><<<G, kBcgBlock, 0, s>>>(n, c(W0), c(SA), part);
        fin<2, FIN_PHI>(0);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    int check(double& maxerr, bool& done) {
        k_bcg_norms<P><<<G, kBcgBlock, 0, ctx->stream>>>(n, ns, sp, st, c(S), c(R), part);
        fin<kBcgNormRows, FIN_CHECK>();
        CLG_HIP(hipGetLastError());
        double h[2];
        CLG_TRY(fetch_doubles(ctx, st + kMaxErr, 2, h));
        maxerr = h[0];
        done = h[1] != 0.0;
        return CLG_OK;
    }
};

}  // namespace

namespace clg {

int scratch_field(clg_context ctx, const clg_field_s* like, int kind, int prec, int subset, int slot, clg_field_s** out);

// Shifted BiCGStab for x_k = (A + c_k)^-1 b, k < ns <= 32, every vector in b's precision (FP64 or FP32;
// X[k] zeroed here).  step_count checks, each after dev_check steps; converged when max(|S|, |s_k|) <
// accuracy * blen.  iters: steps run, maxerr: the last check's maximum.  A: kind/op/gauge/param as
// linop_apply takes them; applies: operator applications.
int msbicgstab_dev(clg_context ctx, int kind, int op, const clg_field_s* gauge, double param,
                   std::vector<clg_field_s*>& X, const std::vector<double>& shifts, const clg_field_s* b, int step_count,
                   int dev_check, double accuracy, double blen, int& iters, double& maxerr, int& applies) {
    const int ns = (int)X.size();
    CLG_CHECK_ARG(ns >= 1 && ns <= kBcgMaxShift && (int)shifts.size() == 2 * ns, "shifted BiCGStab takes 1..%d shifts",
                  kBcgMaxShift);
    CLG_CHECK_ARG(step_count >= 1 && dev_check >= 1, "MaxStep and DiviationStep must be positive");
    CLG_CHECK_ARG(b->prec == CLG_F64 || b->prec == CLG_F32, "shifted BiCGStab vectors are FP64 or FP32");
    long n;
    const int pk = pk_kind(b, n);
    CLG_CHECK_ARG(pk == PK_D || pk == PK_F4, "FP32 vectors of an even number of complex numbers");
    std::vector<clg_field_s*> Sk(ns);
    for (int k = 0; k < ns; ++k) CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 3000 + k, &Sk[k]));
    clg_field_s *R, *W, *W0, *S, *SA, *WA;
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 120, &R));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 121, &W));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 122, &W0));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 123, &S));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 124, &SA));
    CLG_TRY(scratch_field(ctx, b, b->kind, b->prec, b->subset, 125, &WA));
    double* st = nullptr;
    CLG_TRY(workspace(ctx, kBcgState, &st));
    const unsigned G = (unsigned)std::max<long>(1, std::min<long>((n + kBcgBlock - 1) / kBcgBlock, kBcgMaxGrid));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kBcgMaxGrid * kBcgNormRows));

    auto run = [&](auto tag) -> int {
        constexpr int P = decltype(tag)::value;
        BcgRun<P> r;
        r.ctx = ctx;
        r.n = n;
        r.G = G;
        r.ns = ns;
        for (int k = 0; k < kBcgMaxShift; ++k) {
            r.sp.x[k] = k < ns ? X[k]->data : nullptr;
            r.sp.s[k] = k < ns ? Sk[k]->data : nullptr;
        }
        r.st = st;
        r.part = ctx->red_partial;
        r.bound = accuracy * blen;
        r.kind = kind;
        r.op = op;
        r.gauge = gauge;
        r.param = param;
        r.R = R;
        r.W = W;
        r.W0 = W0;
        r.S = S;
        r.SA = SA;
        r.WA = WA;
        hipStream_t s = ctx->stream;
        // x_k = 0, s_k = R = S = b, W0 = b^dagger (:90, :100, :117-123)
        for (int k = 0; k < ns; ++k) {
            CLG_TRY(clg_field_zero(X[k]));
            CLG_TRY(blas_axpby(ctx, Sk[k], 1.0, 0.0, b, 0.0, 0.0));
        }
        CLG_TRY(blas_axpby(ctx, R, 1.0, 0.0, b, 0.0, 0.0));
        CLG_TRY(blas_axpby(ctx, S, 1.0, 0.0, b, 0.0, 0.0));
        CLG_TRY(blas_conj(ctx, W0, b));
        CLG_HIP(hipMemcpyAsync(st + oCn, shifts.data(), sizeof(double) * 2 * ns, hipMemcpyHostToDevice, s));
        // delta = <W0, R>; SA = A S; phi = <W0, SA> / delta and the first step's zeta, beta_k (:124-126)
        k_bcg_dot<P, false><<<G, kBcgBlock, 0, s>>>(n, r.c(W0), r.c(R), r.part);
        r.template fin<2, FIN_INIT>();
        CLG_TRY(r.apply(SA, S));
        k_bcg_dot<P, false><<<G, kBcgBlock, 0, s>>>(n, r.c(W0), r.c(SA), r.part);
        r.template fin<2, FIN_PHI>(1);
        CLG_HIP(hipGetLastError());
        // the steps between two checks launch the same work every time: after the first group, which
        // runs directly (operator tuning, scratch growth), they replay as one recorded graph
        auto group = [&]() -> int {
            for (int i = 0; i < dev_check; ++i) CLG_TRY(r.step());
            return CLG_OK;
        };
        std::vector<unsigned long long> key = {0x4d534243ULL, (unsigned long long)kind, (unsigned long long)op,
                                               key_bits(param), (unsigned long long)(uintptr_t)gauge->data,
                                               (unsigned long long)(uintptr_t)st,
                                               (unsigned long long)(uintptr_t)r.part, (unsigned long long)n,
                                               (unsigned long long)P, (unsigned long long)ns,
                                               (unsigned long long)dev_check, (unsigned long long)G,
                                               (unsigned long long)b->kind, (unsigned long long)b->subset,
                                               (unsigned long long)b->Vh};
        for (int d = 0; d < 4; ++d) key.push_back((unsigned long long)(long long)b->bc[d]);
        for (int k = 0; k < ns; ++k) {
            key.push_back((unsigned long long)(uintptr_t)X[k]->data);
            key.push_back((unsigned long long)(uintptr_t)Sk[k]->data);
        }
        for (clg_field_s* f : {R, W, W0, S, SA, WA}) key.push_back((unsigned long long)(uintptr_t)f->data);
        iters = 0;
        maxerr = 0.0;
        bool done = false;
        for (int c = 0; c < step_count && !done; ++c) {
            if (c == 0) {
                CLG_TRY(group());
            } else {
                const int a0 = r.applies;
                CLG_TRY(cycle_graph_run(ctx, key, group));
                r.applies = a0 + 2 * dev_check;
            }
            iters += dev_check;
            CLG_TRY(r.check(maxerr, done));
        }
        applies += r.applies;
        return done ? CLG_OK : CLG_NOT_CONVERGED;
    };
    if (pk == PK_D) return run(std::integral_constant<int, PK_D>{});
    return run(std::integral_constant<int, PK_F4>{});
}

}  // namespace clg
