// clg_blas.hip -- level-1 BLAS and FP64 reductions on fermion fields.
//
// F64 and F32 fermion fields are, whatever their site layout, one contiguous array of
// complex numbers, so the element-wise kernels run over that flat array in 16-byte packets
// (one FP64 complex or two FP32 complex per lane and load).  Reductions accumulate in FP64
// per lane, reduce across the 64-lane wavefront with shuffles, across the workgroup through
// LDS, and across workgroups with a last-block pass; the result is written to device memory,
// so the solvers never wait on the host for a dot product (the reference's Dot does ~20
// launches and a blocking copy, CudaHelper.cu:626-710).
#include "clg_internal.h"
#include "clg_packets.h"

#include <algorithm>
#include <cstring>
#include <mutex>

using namespace clg;
using namespace clgd;

namespace {

const int kMaxRedGrid = 2048;
const int kMaxK = 16;  // vectors per multi-dot / multi-axpy launch
// the context's 4096-double device result buffer: results from 0, host-supplied coefficients of
// the BLAS entry points from 2112 (1984 doubles)
const int kHostArgs = 2112;

#define GRID_STRIDE(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)
// four independent packets per lane per trip, so every wave keeps four loads in flight
template <typename F>
__device__ __forceinline__ void stride4(long n, F&& body) {
    const long st = (long)gridDim.x * blockDim.x;
    for (long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * st) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long i = i0 + u * st;
            if (i < n) body(i);
        }
    }
}

// y = a x + b y   (a, b complex, host scalars).  READY = false: b == 0 and y is only written,
// so stale NaN/Inf in y never leaks into the result.
template <int P, bool READY>
__global__ void k_axpby(long n, double ar, double ai, const typename Pk<P>::T* __restrict__ x, double br, double bi,
                        typename Pk<P>::T* y) {
    constexpr int NC = Pk<P>::NC;
    stride4(n, [&](long i) {
        double xv[2 * NC], yv[2 * NC], o[2 * NC];
        Pk<P>::ld(x, i, xv);
        if (READY) Pk<P>::ld(y, i, yv);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            o[2 * c] = ar * xv[2 * c] - ai * xv[2 * c + 1];
            o[2 * c + 1] = ar * xv[2 * c + 1] + ai * xv[2 * c];
            if (READY) {
                o[2 * c] += br * yv[2 * c] - bi * yv[2 * c + 1];
                o[2 * c + 1] += br * yv[2 * c + 1] + bi * yv[2 * c];
            }
        }
        Pk<P>::st(y, i, o);
    });
}
// y += s * c x with c read from device memory (complex), s = +-1
template <int P>
__global__ void k_axpy_dev(long n, const double* __restrict__ cc, double s, const typename Pk<P>::T* __restrict__ x,
                           typename Pk<P>::T* y) {
    constexpr int NC = Pk<P>::NC;
    const double ar = s * cc[0], ai = s * cc[1];
    stride4(n, [&](long i) {
        double xv[2 * NC], yv[2 * NC];
        Pk<P>::ld(x, i, xv);
        Pk<P>::ld(y, i, yv);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            yv[2 * c] += ar * xv[2 * c] - ai * xv[2 * c + 1];
            yv[2 * c + 1] += ar * xv[2 * c + 1] + ai * xv[2 * c];
        }
        Pk<P>::st(y, i, yv);
    });
}
// dst = src / sqrt(nrm2[0])   (nrm2 on device; zero-safe)
template <int P>
__global__ void k_scale_inv_sqrt(long n, const double* __restrict__ nrm2, const typename Pk<P>::T* __restrict__ src,
                                 typename Pk<P>::T* dst) {
    constexpr int NC = Pk<P>::NC;
    const double v = nrm2[0];
    const double f = v > 0 ? 1.0 / sqrt(v) : 0.0;
    stride4(n, [&](long i) {
        double a[2 * NC];
        Pk<P>::ld(src, i, a);
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) a[c] *= f;
        Pk<P>::st(dst, i, a);
    });
}
// b = conj(a)
template <int P>
__global__ void k_conj(long n, const typename Pk<P>::T* __restrict__ a, typename Pk<P>::T* b) {
    constexpr int NC = Pk<P>::NC;
    stride4(n, [&](long i) {
        double v[2 * NC];
        Pk<P>::ld(a, i, v);
#pragma unroll
        for (int c = 0; c < NC; ++c) v[2 * c + 1] = -v[2 * c + 1];
        Pk<P>::st(b, i, v);
    });
}
// out = sum conj(a) b  (2 doubles)
template <int P>
__global__ __launch_bounds__(kRedBlock) void k_dot(long n, const typename Pk<P>::T* __restrict__ a,
                                                    const typename Pk<P>::T* __restrict__ b, double* partial) {
    constexpr int NC = Pk<P>::NC;
    double v[2] = {0, 0};
    stride4(n, [&](long i) {
        double x[2 * NC], y[2 * NC];
        Pk<P>::ld(a, i, x);
        Pk<P>::ld(b, i, y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            v[0] += x[2 * c] * y[2 * c] + x[2 * c + 1] * y[2 * c + 1];
            v[1] += x[2 * c] * y[2 * c + 1] - x[2 * c + 1] * y[2 * c];
        }
    });
    block_partial<2>(v, partial);
}
template <int P>
__global__ __launch_bounds__(kRedBlock) void k_norm2(long n, const typename Pk<P>::T* __restrict__ a, double* partial) {
    constexpr int NC = Pk<P>::NC;
    double v[1] = {0};
    stride4(n, [&](long i) {
        double x[2 * NC];
        Pk<P>::ld(a, i, x);
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) v[0] += x[c] * x[c];
    });
    block_partial<1>(v, partial);
}

template <int P> struct Ptrs {
    const typename Pk<P>::T* p[kMaxK];
};
// partials of sum conj(V_k) w for k < K (2K values) and, WN, of |w|^2 (one more); k_red_finish
// completes them.  cond (nullable): when *cond == 0 the whole grid exits at once (a skipped pass).
template <int P, int K, bool WN>
__global__ __launch_bounds__(kRedBlock) void k_multidot(long n, Ptrs<P> V, const typename Pk<P>::T* __restrict__ w,
                                                         double* partial, const int* cond) {
    constexpr int NC = Pk<P>::NC;
    if (cond && *cond == 0) return;
    double v[2 * K + 1];
#pragma unroll
    for (int k = 0; k < 2 * K + 1; ++k) v[k] = 0;
    GRID_STRIDE(i, n) {
        // issue every packet load of this element first (raw 16-byte packets), then convert: the
        // loads stay in flight together instead of one vector at a time
        typename Pk<P>::T raw[K];
#pragma unroll
        for (int k = 0; k < K; ++k) raw[k] = V.p[k][i];
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double x[2 * NC];
            Pk<P>::cvt(raw[k], x);
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                re += x[2 * c] * y[2 * c] + x[2 * c + 1] * y[2 * c + 1];
                im += x[2 * c] * y[2 * c + 1] - x[2 * c + 1] * y[2 * c];
            }
            v[2 * k] += re;
            v[2 * k + 1] += im;
        }
        if constexpr (WN) {
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) v[2 * K] += y[c] * y[c];
        }
    }
    block_partial<2 * K + 1>(v, partial);
}
// w += s * sum_k c_k V_k  (c on device, 2K doubles); NORM: partials of |w|^2 of the stored result
// (k_red_finish completes them and sets the Gram-Schmidt reorthogonalisation flag).  cond
// (nullable): skip the pass when *cond == 0.
template <int P, int K, bool NORM>
__global__ __launch_bounds__(kRedBlock) void k_multiaxpy(long n, Ptrs<P> V, const double* __restrict__ cc, double s,
                                                          typename Pk<P>::T* w, double* partial, const int* cond) {
    constexpr int NC = Pk<P>::NC;
    if (cond && *cond == 0) return;
    double cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = s * cc[2 * k]; ci[k] = s * cc[2 * k + 1]; }
    double v[1] = {0};
    GRID_STRIDE(i, n) {
        typename Pk<P>::T raw[K];
#pragma unroll
        for (int k = 0; k < K; ++k) raw[k] = V.p[k][i];
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double x[2 * NC];
            Pk<P>::cvt(raw[k], x);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                y[2 * c] += cr[k] * x[2 * c] - ci[k] * x[2 * c + 1];
                y[2 * c + 1] += cr[k] * x[2 * c + 1] + ci[k] * x[2 * c];
            }
        }
        Pk<P>::st(w, i, y);
        if constexpr (NORM) {
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) v[0] += y[c] * y[c];
        }
    }
    if constexpr (NORM) block_partial<1>(v, partial);
}

// Chunked forms for 17..kMaxChunkV vectors (long Arnoldi bases: multi-shift FOM, GMRES(m > 16),
// GCRO-DR): one launch per pass instead of one per 16 vectors.
constexpr int kMaxChunkV = 64;
template <int P> struct PtrsN {
    const typename Pk<P>::T* p[kMaxChunkV];
};
// multi-dot over chunk blockIdx.y (vectors 16c .. 16c+15, the last chunk partial); partials
// [chunk][block][2*kMaxK + 1], the last slot |w|^2
template <int P>
__global__ __launch_bounds__(kRedBlock) void k_multidot_ch(long n, PtrsN<P> V, int nv, const typename Pk<P>::T* __restrict__ w,
                                                            double* partial, const int* cond) {
    constexpr int K = kMaxK, NC = Pk<P>::NC;
    if (cond && *cond == 0) return;
    const int k0 = (int)blockIdx.y * K, kk = min(K, nv - k0);
    double v[2 * K + 1];
#pragma unroll
    for (int k = 0; k < 2 * K + 1; ++k) v[k] = 0;
    GRID_STRIDE(i, n) {
        typename Pk<P>::T raw[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < kk) raw[k] = V.p[k0 + k][i];
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < kk) {
                double x[2 * NC];
                Pk<P>::cvt(raw[k], x);
                double re = 0.0, im = 0.0;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    re += x[2 * c] * y[2 * c] + x[2 * c + 1] * y[2 * c + 1];
                    im += x[2 * c] * y[2 * c + 1] - x[2 * c + 1] * y[2 * c];
                }
                v[2 * k] += re;
                v[2 * k + 1] += im;
            }
        }
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) v[2 * K] += y[c] * y[c];
    }
    block_partial<2 * K + 1>(v, partial + (size_t)blockIdx.y * gridDim.x * (2 * K + 1));
}
// w += s * sum_k c_k V_k over nv <= kMaxChunkV vectors in one pass (w read and written once);
// NORM: partials of |w|^2 of the stored result
template <int P, bool NORM>
__global__ __launch_bounds__(kRedBlock) void k_multiaxpy_ch(long n, PtrsN<P> V, int nv, const double* __restrict__ cc,
                                                             double s, typename Pk<P>::T* w, double* partial,
                                                             const int* cond) {
    constexpr int K = kMaxK, NC = Pk<P>::NC;
    if (cond && *cond == 0) return;
    double v[1] = {0};
    GRID_STRIDE(i, n) {
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
        for (int k0 = 0; k0 < nv; k0 += K) {
            typename Pk<P>::T raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k0 + k < nv) raw[k] = V.p[k0 + k][i];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (k0 + k < nv) {
                    const double cr = s * cc[2 * (k0 + k)], ci = s * cc[2 * (k0 + k) + 1];
                    double x[2 * NC];
                    Pk<P>::cvt(raw[k], x);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        y[2 * c] += cr * x[2 * c] - ci * x[2 * c + 1];
                        y[2 * c + 1] += cr * x[2 * c + 1] + ci * x[2 * c];
                    }
                }
            }
        }
        Pk<P>::st(w, i, y);
        if constexpr (NORM) {
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) v[0] += y[c] * y[c];
        }
    }
    if constexpr (NORM) block_partial<1>(v, partial);
}
// finish of k_multidot_ch: one workgroup per chunk (blockIdx.y); chunk c writes its 2*kk sums to
// out[2*kMaxK*c ..], chunk 0 also |w|^2 to wn (nullable)
__global__ __launch_bounds__(kRedBlock) void k_red_finish_ch(const double* __restrict__ partial, int nblk, int nv,
                                                             double* out, double* wn, const int* cond) {
    constexpr int NR = 2 * kMaxK + 1;
    if (cond && *cond == 0) return;
    const int ch = (int)blockIdx.y, kk = min(kMaxK, nv - ch * kMaxK);
    const double* pp = partial + (size_t)ch * nblk * NR;
    __shared__ double sm[kRedBlock / 64][NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += pp[(size_t)b * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = wave_sum(acc[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x < NR) {
        const int r = threadIdx.x;
        double t = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sm[q][r];
        if (r < 2 * kk) out[2 * kMaxK * ch + r] = t;
        else if (r == NR - 1 && ch == 0 && wn) wn[0] = t;
    }
}

// Second stage of the BLAS reductions (one workgroup): total_r = sum over nblk partials of
// partial[b * NR + r] in a fixed order; totals r < nout go to out[r], the one after them to
// out2[0] (when given).  flag (nullable): *flag = (total_0 < eta * ref[0]).  cond (nullable): skip
// when *cond == 0, as the first stage did.  Every load of a thread is independent (block-major
// sweep with NR accumulators), so they all stay in flight.
template <int NR>
__global__ __launch_bounds__(kRedBlock) void k_red_finish(const double* __restrict__ partial, int nblk, double* out,
                                                          int nout, double* out2, const int* cond, const double* ref,
                                                          double eta, int* flag) {
    if (cond && *cond == 0) return;
    __shared__ double sm[kRedBlock / 64][NR];
    __shared__ double tot[NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += partial[(size_t)b * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = wave_sum(acc[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x < NR) {
        const int r = threadIdx.x;
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w][r];
        tot[r] = t;
        if (r < nout) out[r] = t;
        else if (out2) out2[r - nout] = t;
    }
    if (flag) {
        __syncthreads();
        if (threadIdx.x == 0) *flag = tot[0] < eta * ref[0] ? 1 : 0;
    }
}

// out_i = sum_{j<X} in_j M[j*Y + i] (i < Y <= YM) for FP64 complex vectors of n elements, M in
// LDS; out may be the first Y inputs (each thread reads every input of its element before it
// writes).  One pass: X reads and Y writes per element.
constexpr int kMaxVmmIn = 32, kMaxVmmOut = 31;
struct VmmArgs {
    const double2* in[kMaxVmmIn];
    double2* out[kMaxVmmOut];
    int X, Y;
    long n;
};
template <int YM>
__global__ __launch_bounds__(256) void k_vmm(VmmArgs a, const double* __restrict__ M) {
    __shared__ double2 sm[kMaxVmmIn * YM];
    const int tot = a.X * a.Y;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) sm[i] = make_double2(M[2 * i], M[2 * i + 1]);
    __syncthreads();
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < a.n; e += (long)gridDim.x * blockDim.x) {
        double2 acc[YM];
#pragma unroll
        for (int i = 0; i < YM; ++i) acc[i] = make_double2(0.0, 0.0);
        for (int j = 0; j < a.X; ++j) {
            const double2 v = a.in[j][e];
            const double2* row = sm + j * a.Y;
#pragma unroll
            for (int i = 0; i < YM; ++i) {
                if (i < a.Y) {
                    const double2 m = row[i];
                    acc[i].x = fma(v.x, m.x, fma(-v.y, m.y, acc[i].x));
                    acc[i].y = fma(v.x, m.y, fma(v.y, m.x, acc[i].y));
                }
            }
        }
#pragma unroll
        for (int i = 0; i < YM; ++i)
            if (i < a.Y) a.out[i][e] = acc[i];
    }
}

// ---- fused Gram-Schmidt step of the Arnoldi process (see blas_gs_step)
// w' = w - sum_k h_k V_k (h = the first-pass projections, on device) is stored, and the same pass
// yields the partials of V_k^H w' (2K values) and |w'|^2 (the last one).
// With vnext given, the pass first takes |w'|^2 from the first pass as |w|^2 - |h|^2 (V orthonormal).
// When that keeps at least half of |w|^2 (no second pass will be needed), the step's fast form
// applies: w' / |w'| goes straight to vnext and no partials are taken.  Workgroup 0 records the choice
// in sc ([1] |w'|^2, [2] 1/|w'|, [3] 2 for the fast form, else 0) for the finish and finalize.
template <int P, int K>
__global__ __launch_bounds__(kRedBlock) void k_gs_project(long n, Ptrs<P> V, const double* __restrict__ h,
                                                           typename Pk<P>::T* w, double* partial, double* sc,
                                                           typename Pk<P>::T* vnext) {
    constexpr int NC = Pk<P>::NC;
    double hr[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { hr[k] = h[2 * k]; hi[k] = h[2 * k + 1]; }
    if (vnext) {
        const double w2 = sc[0];
        double b2 = w2;
#pragma unroll
        for (int k = 0; k < K; ++k) b2 -= hr[k] * hr[k] + hi[k] * hi[k];
        const bool fast = b2 >= 0.5 * w2 && b2 > 1e-300;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc[3] = fast ? 2.0 : 0.0;
            if (fast) {
                sc[1] = b2;
                sc[2] = 1.0 / sqrt(b2);
            }
        }
        if (fast) {
            const double inv = 1.0 / sqrt(b2);
            GRID_STRIDE(i, n) {
                typename Pk<P>::T raw[K];
#pragma unroll
                for (int k = 0; k < K; ++k) raw[k] = V.p[k][i];
                double y[2 * NC];
                Pk<P>::ld(w, i, y);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    double x[2 * NC];
                    Pk<P>::cvt(raw[k], x);
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        y[2 * c] -= hr[k] * x[2 * c] - hi[k] * x[2 * c + 1];
                        y[2 * c + 1] -= hr[k] * x[2 * c + 1] + hi[k] * x[2 * c];
                    }
                }
#pragma unroll
                for (int c = 0; c < 2 * NC; ++c) y[c] *= inv;
                Pk<P>::st(vnext, i, y);
            }
            return;
        }
    }
    double v[2 * K + 1];
#pragma unroll
    for (int k = 0; k < 2 * K + 1; ++k) v[k] = 0;
    GRID_STRIDE(i, n) {
        typename Pk<P>::T raw[K];
#pragma unroll
        for (int k = 0; k < K; ++k) raw[k] = V.p[k][i];
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double x[2 * NC];
            Pk<P>::cvt(raw[k], x);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                y[2 * c] -= hr[k] * x[2 * c] - hi[k] * x[2 * c + 1];
                y[2 * c + 1] -= hr[k] * x[2 * c + 1] + hi[k] * x[2 * c];
            }
        }
        Pk<P>::st(w, i, y);  // y now holds the stored (rounded) values
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double x[2 * NC];
            Pk<P>::cvt(raw[k], x);
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                re += x[2 * c] * y[2 * c] + x[2 * c + 1] * y[2 * c + 1];
                im += x[2 * c] * y[2 * c + 1] - x[2 * c + 1] * y[2 * c];
            }
            v[2 * k] += re;
            v[2 * k + 1] += im;
        }
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) v[2 * K] += y[c] * y[c];
    }
    block_partial<2 * K + 1>(v, partial);
}

// Completes k_gs_project (one workgroup) and updates Hessenberg column j (hcol, interleaved complex,
// rows 0..j+1) with the Daniel-Gragg-Kaufman-Stewart test: the second pass counts only when the
// first cancelled more than half of |w|^2 (|w'|^2 < |w|^2 / 2).  Then
//   H(i, j) = h_i + flag c2_i (shifted: delta_ij - that), |w''|^2 = |w'|^2 - flag |c2|^2
// (V orthonormal), H(j+1, j) = +-|w''|.  sc: [0] |w|^2 (in), [1] |w''|^2, [2] 1/|w''|, [3] flag,
// [4..4+2K) c2.
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_gs_finish(const double* __restrict__ partial, int nblk, double* hcol,
                                                          int j, int shifted, double* sc, int fast_ok) {
    constexpr int NR = 2 * K + 1;
    if (fast_ok && sc[3] == 2.0) {
        // the projection pass took the fast form: H(i, j) = h_i, H(j+1, j) = |w'| from sc
        const int i = threadIdx.x;
        if (i < K && shifted) {
            hcol[2 * i] = (i == j ? 1.0 : 0.0) - hcol[2 * i];
            hcol[2 * i + 1] = -hcol[2 * i + 1];
        }
        if (i == 0) {
            const double f = sqrt(sc[1]);
            hcol[2 * (j + 1)] = shifted ? -f : f;
            hcol[2 * (j + 1) + 1] = 0.0;
        }
        return;
    }
    __shared__ double sm[kRedBlock / 64][NR];
    __shared__ double tot[NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += partial[(size_t)b * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = wave_sum(acc[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x < NR) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w][threadIdx.x];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
    const double w2 = tot[2 * K];
    const bool flag = w2 < 0.5 * sc[0];
    const int i = threadIdx.x;
    if (i < K) {
        double re = hcol[2 * i], im = hcol[2 * i + 1];
        if (flag) {
            re += tot[2 * i];
            im += tot[2 * i + 1];
        }
        if (shifted) {
            re = (i == j ? 1.0 : 0.0) - re;
            im = -im;
        }
        hcol[2 * i] = re;
        hcol[2 * i + 1] = im;
        sc[4 + 2 * i] = tot[2 * i];
        sc[4 + 2 * i + 1] = tot[2 * i + 1];
    }
    if (i == 0) {
        double c2 = 0;
        for (int k = 0; k < 2 * K; ++k) c2 += tot[k] * tot[k];
        double f2 = flag ? w2 - c2 : w2;
        if (!(f2 > 1e-300)) f2 = 1e-300;
        const double f = sqrt(f2);
        hcol[2 * (j + 1)] = shifted ? -f : f;
        hcol[2 * (j + 1) + 1] = 0.0;
        sc[1] = f2;
        sc[2] = 1.0 / f;
        sc[3] = flag ? 1.0 : 0.0;
    }
}

// vnext = (w' - flag sum_k c2_k V_k) / |w''| with flag, c2 and 1/|w''| from k_gs_finish
template <int P, int K>
__global__ __launch_bounds__(kRedBlock) void k_gs_finalize(long n, Ptrs<P> V, const typename Pk<P>::T* __restrict__ w,
                                                            typename Pk<P>::T* vnext, const double* __restrict__ sc) {
    constexpr int NC = Pk<P>::NC;
    if (sc[3] == 2.0) return;  // the projection pass wrote vnext (fast form)
    const bool flag = sc[3] != 0.0;
    const double inv = sc[2];
    if (flag) {
        double cr[K], ci[K];
#pragma unroll
        for (int k = 0; k < K; ++k) { cr[k] = sc[4 + 2 * k]; ci[k] = sc[4 + 2 * k + 1]; }
        GRID_STRIDE(i, n) {
            typename Pk<P>::T raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) raw[k] = V.p[k][i];
            double y[2 * NC];
            Pk<P>::ld(w, i, y);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double x[2 * NC];
                Pk<P>::cvt(raw[k], x);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    y[2 * c] -= cr[k] * x[2 * c] - ci[k] * x[2 * c + 1];
                    y[2 * c + 1] -= cr[k] * x[2 * c + 1] + ci[k] * x[2 * c];
                }
            }
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) y[c] *= inv;
            Pk<P>::st(vnext, i, y);
        }
    } else {
        GRID_STRIDE(i, n) {
            double y[2 * NC];
            Pk<P>::ld(w, i, y);
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) y[c] *= inv;
            Pk<P>::st(vnext, i, y);
        }
    }
}

unsigned red_grid(clg_context ctx, long n) {
    long g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    const long cap = std::min<long>(kMaxRedGrid, (long)ctx->num_cu * 4);
    return (unsigned)std::max<long>(1, std::min(g, cap));
}
unsigned ew_grid(clg_context ctx, long n) {
    long g = (n + 255) / 256;
    const long cap = (long)ctx->num_cu * 16;
    return (unsigned)std::max<long>(1, std::min(g, cap));
}

}  // namespace

namespace clg {

// FP16-storage Wilson fields (clg_blas16.hip)
int h16_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int h16_norm2_dev(clg_context, const clg_field_s*, double*);
int h16_scale_inv_sqrt(clg_context, clg_field_s*, const clg_field_s*, const double*);
int h16_multiaxpy(clg_context, clg_field_s*, clg_field_s* const*, int, const double*, double);
int h16_gs_step(clg_context, clg_field_s* const*, int, clg_field_s*, double*, int, bool, clg_field_s*, double*);

// ---- first-stage buffer of clg_multidot_noalloc: a file-scope device array, so that entry point
// allocates nothing.  Each device holds its own copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress returns on that device (kept per device below).
constexpr size_t kNoAllocPartials = (size_t)kMaxRedGrid * (2 * kMaxK + 1) * (kMaxChunkV / kMaxK);  // 2,162,688 B
static __device__ __align__(256) double scratch[kNoAllocPartials];
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kMaxDevices] = {};

// called by clg_init with `device` current: looks the array's address up once per device
int noalloc_scratch_init(int device) {
    CLG_CHECK_ARG(device >= 0 && device < kMaxDevices, "device %d beyond the %d devices the library tracks", device,
                  kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return CLG_OK;
    void* ptr = nullptr;
    CLG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[device] = static_cast<double*>(ptr);
    return CLG_OK;
}
static double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

int ensure_reduction_scratch(clg_context ctx, size_t doubles) {
    if (ctx->red_partial_cap >= doubles) return CLG_OK;
    if (ctx->red_partial) {
        CLG_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->red_partial);
        ctx->red_partial = nullptr;
    }
    CLG_HIP(hipMalloc((void**)&ctx->red_partial, doubles * sizeof(double)));
    ctx->red_partial_cap = doubles;
    return CLG_OK;
}

int reduction_ready(clg_context ctx) { return ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)); }

long n_complex(const clg_field_s* f) {
    if (f->kind == CLG_GAUGE) return 8 * 9 * f->Vh;  // 18-real links: 8 (parity, mu) blocks of 9 complex planes
    const long nblk = f->subset == CLG_FULL ? 2 : 1;
    return nblk * f->Vh * (f->kind == CLG_WILSON ? 12 : 3);
}

// packet kind of a field and its packet count
static int pk_of(const clg_field_s* f, long& npk) {
    const long n = n_complex(f);
    if (f->prec == CLG_F64) { npk = n; return PK_D; }
    if ((n & 1) == 0) { npk = n / 2; return PK_F4; }
    npk = n;
    return PK_F2;
}
// the same for other files' kernels
int pk_kind(const clg_field_s* f, long& npk) { return pk_of(f, npk); }

// fermion fields, and 18-real gauge fields (a flat array of link elements, as CFieldGaugeSU3::Dot /
// Axpy / ScalarMultply see them)
static int check_flat(const clg_field_s* a) {
    CLG_CHECK_ARG(a && (is_fermion(a->kind) || a->kind == CLG_GAUGE), "need a fermion field or an 18-real gauge field");
    CLG_CHECK_ARG(a->prec != CLG_F16, "BLAS on FP16-storage fields is not supported; convert first");
    return CLG_OK;
}
static int check_pair(const clg_field_s* a, const clg_field_s* b) {
    CLG_TRY(check_flat(a));
    CLG_TRY(check_flat(b));
    CLG_CHECK_ARG(a->kind == b->kind && a->subset == b->subset && a->prec == b->prec,
                  "fields differ in kind, subset or precision");
    return check_same_lattice(a, b);
}

#define PK_DISPATCH(pk, MACRO)                  \
    switch (pk) {                               \
        case PK_D: MACRO(PK_D); break;          \
        case PK_F4: MACRO(PK_F4); break;        \
        default: MACRO(PK_F2); break;           \
    }

int blas_axpby(clg_context ctx, clg_field_s* y, double ar, double ai, const clg_field_s* x, double br, double bi) {
    if (y && y->prec == CLG_F16) return h16_axpby(ctx, y, ar, ai, x, br, bi);
    CLG_TRY(check_pair(y, x));
    long n;
    const int pk = pk_of(y, n);
    const unsigned g = ew_grid(ctx, n);
    const bool ready = !(br == 0.0 && bi == 0.0);
#define L(P)                                                                                                       \
    if (ready)                                                                                                     \
        k_axpby<P, true><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const Pk<P>::T*)x->data, br, bi, (Pk<P>::T*)y->data); \
    else                                                                                                           \
        k_axpby<P, false><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const Pk<P>::T*)x->data, br, bi, (Pk<P>::T*)y->data);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_axpy_dev(clg_context ctx, clg_field_s* y, const double* c, double s, const clg_field_s* x) {
    CLG_TRY(check_pair(y, x));
    long n;
    const int pk = pk_of(y, n);
    const unsigned g = ew_grid(ctx, n);
#define L(P) k_axpy_dev<P><<<g, 256, 0, ctx->stream>>>(n, c, s, (const Pk<P>::T*)x->data, (Pk<P>::T*)y->data);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_scale_inv_sqrt(clg_context ctx, clg_field_s* dst, const clg_field_s* src, const double* nrm2) {
    if (dst && dst->prec == CLG_F16) return h16_scale_inv_sqrt(ctx, dst, src, nrm2);
    CLG_TRY(check_pair(dst, src));
    long n;
    const int pk = pk_of(dst, n);
    const unsigned g = ew_grid(ctx, n);
#define L(P) k_scale_inv_sqrt<P><<<g, 256, 0, ctx->stream>>>(n, nrm2, (const Pk<P>::T*)src->data, (Pk<P>::T*)dst->data);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_conj(clg_context ctx, clg_field_s* dst, const clg_field_s* src) {
    CLG_TRY(check_pair(dst, src));
    long n;
    const int pk = pk_of(dst, n);
    const unsigned g = ew_grid(ctx, n);
#define L(P) k_conj<P><<<g, 256, 0, ctx->stream>>>(n, (const Pk<P>::T*)src->data, (Pk<P>::T*)dst->data);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// second stage of a reduction whose first stage wrote nblk * nr partials to ctx->red_partial
// second stage of a reduction whose first stage wrote nblk x nr partials to `partial`
static int red_finish_at(clg_context ctx, const double* partial, unsigned nblk, int nr, double* out, int nout,
                         double* out2, const int* cond, const double* ref, double eta, int* flag) {
#define FIN(N)                                                                                                    \
    case N:                                                                                                       \
        k_red_finish<N><<<1, kRedBlock, 0, ctx->stream>>>(partial, (int)nblk, out, nout, out2, cond, ref, eta, flag); \
        break;
    switch (nr) {
        FIN(1) FIN(2) FIN(3) FIN(5) FIN(7) FIN(9) FIN(11) FIN(13) FIN(15) FIN(17) FIN(19) FIN(21) FIN(23) FIN(25)
        FIN(27) FIN(29) FIN(31) FIN(33)
        default: CLG_CHECK_ARG(false, "no reduction finish for %d values", nr);
    }
#undef FIN
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
static int red_finish(clg_context ctx, unsigned nblk, int nr, double* out, int nout, double* out2, const int* cond,
                      const double* ref, double eta, int* flag) {
    return red_finish_at(ctx, ctx->red_partial, nblk, nr, out, nout, out2, cond, ref, eta, flag);
}

int blas_dot_dev(clg_context ctx, const clg_field_s* a, const clg_field_s* b, double* out) {
    CLG_TRY(check_pair(a, b));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)));
    long n;
    const int pk = pk_of(a, n);
    const unsigned g = red_grid(ctx, n);
#define L(P) \
    k_dot<P><<<g, kRedBlock, 0, ctx->stream>>>(n, (const Pk<P>::T*)a->data, (const Pk<P>::T*)b->data, ctx->red_partial);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return red_finish(ctx, g, 2, out, 2, nullptr, nullptr, nullptr, 0.0, nullptr);
}

int blas_norm2_dev(clg_context ctx, const clg_field_s* a, double* out) {
    if (a && a->prec == CLG_F16) return h16_norm2_dev(ctx, a, out);
    CLG_TRY(check_flat(a));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)));
    long n;
    const int pk = pk_of(a, n);
    const unsigned g = red_grid(ctx, n);
#define L(P) k_norm2<P><<<g, kRedBlock, 0, ctx->stream>>>(n, (const Pk<P>::T*)a->data, ctx->red_partial);
    PK_DISPATCH(pk, L)
#undef L
    CLG_HIP(hipGetLastError());
    return red_finish(ctx, g, 1, out, 1, nullptr, nullptr, nullptr, 0.0, nullptr);
}

#define SWITCH_K(K, CALL)                                         \
    switch (K) {                                                  \
        case 1: CALL(1); break;  case 2: CALL(2); break;          \
        case 3: CALL(3); break;  case 4: CALL(4); break;          \
        case 5: CALL(5); break;  case 6: CALL(6); break;          \
        case 7: CALL(7); break;  case 8: CALL(8); break;          \
        case 9: CALL(9); break;  case 10: CALL(10); break;        \
        case 11: CALL(11); break; case 12: CALL(12); break;       \
        case 13: CALL(13); break; case 14: CALL(14); break;       \
        case 15: CALL(15); break; default: CALL(16); break;       \
    }

// partial: the first-stage buffer, kMaxRedGrid x (2 kMaxK + 1) doubles per 16-vector chunk (nullptr: the
// context's, grown here as needed)
template <int P>
static int multidot_p(clg_context ctx, clg_field_s* const* V, int nv, const clg_field_s* w, double* out, double* wnorm,
                      const int* cond, long n, double* partial) {
    const unsigned g = red_grid(ctx, n);
    if (nv > kMaxK && nv <= kMaxChunkV) {
        const int nch = (nv + kMaxK - 1) / kMaxK;
        if (!partial) {
            CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1) * nch));
            partial = ctx->red_partial;
        }
        PtrsN<P> Pn;
        for (int k = 0; k < nv; ++k) Pn.p[k] = (const typename Pk<P>::T*)V[k]->data;
        k_multidot_ch<P><<<dim3(g, nch), kRedBlock, 0, ctx->stream>>>(n, Pn, nv, (const typename Pk<P>::T*)w->data,
                                                                      partial, cond);
        CLG_HIP(hipGetLastError());
        k_red_finish_ch<<<dim3(1, nch), kRedBlock, 0, ctx->stream>>>(partial, (int)g, nv, out, wnorm, cond);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    if (!partial) partial = ctx->red_partial;
    for (int k0 = 0; k0 < nv; k0 += kMaxK) {
        const int kk = std::min(kMaxK, nv - k0);
        Ptrs<P> Pp;
        for (int k = 0; k < kk; ++k) Pp.p[k] = (const typename Pk<P>::T*)V[k0 + k]->data;
        const typename Pk<P>::T* wp = (const typename Pk<P>::T*)w->data;
        double* wn = k0 == 0 ? wnorm : nullptr;
        double* o = out + 2 * k0;
#define CALL(KK)                                                                            \
    if (wn)                                                                                 \
        k_multidot<P, KK, true><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, wp, partial, cond); \
    else                                                                                    \
        k_multidot<P, KK, false><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, wp, partial, cond);
        SWITCH_K(kk, CALL)
#undef CALL
        CLG_HIP(hipGetLastError());
        // the first stage writes 2K + 1 partials per block (the last is |w|^2, kept only when asked)
        CLG_TRY(red_finish_at(ctx, partial, g, 2 * kk + 1, o, 2 * kk, wn, cond, nullptr, 0.0, nullptr));
    }
    return CLG_OK;
}

// out[2k], out[2k+1] = <V_k, w>, k < nv (chunks of 16 vectors per launch); wnorm (nullable)
// receives |w|^2; cond (nullable, device int) skips the work when zero.
int blas_multidot_dev(clg_context ctx, clg_field_s* const* V, int nv, const clg_field_s* w, double* out, double* wnorm,
                      const int* cond) {
    CLG_CHECK_ARG(nv >= 1, "need at least one vector");
    for (int k = 0; k < nv; ++k) CLG_TRY(check_pair(V[k], w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)));
    long n;
    const int pk = pk_of(w, n);
    if (pk == PK_D) return multidot_p<PK_D>(ctx, V, nv, w, out, wnorm, cond, n, nullptr);
    if (pk == PK_F4) return multidot_p<PK_F4>(ctx, V, nv, w, out, wnorm, cond, n, nullptr);
    return multidot_p<PK_F2>(ctx, V, nv, w, out, wnorm, cond, n, nullptr);
}

template <int P>
static int multiaxpy_p(clg_context ctx, clg_field_s* w, clg_field_s* const* V, int nv, const double* c, double s,
                       double* nrm, const int* cond, const double* ref, double eta, int* flag, long n) {
    if (nv > kMaxK && nv <= kMaxChunkV) {
        PtrsN<P> Pn;
        for (int k = 0; k < nv; ++k) Pn.p[k] = (const typename Pk<P>::T*)V[k]->data;
        typename Pk<P>::T* wp = (typename Pk<P>::T*)w->data;
        if (nrm) {
            const unsigned g = red_grid(ctx, n);
            k_multiaxpy_ch<P, true><<<g, kRedBlock, 0, ctx->stream>>>(n, Pn, nv, c, s, wp, ctx->red_partial, cond);
            CLG_HIP(hipGetLastError());
            return red_finish(ctx, g, 1, nrm, 1, nullptr, cond, ref, eta, flag);
        }
        const unsigned g = ew_grid(ctx, n);
        k_multiaxpy_ch<P, false><<<g, kRedBlock, 0, ctx->stream>>>(n, Pn, nv, c, s, wp, nullptr, cond);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    for (int k0 = 0; k0 < nv; k0 += kMaxK) {
        const int kk = std::min(kMaxK, nv - k0);
        const bool want = k0 + kk >= nv && nrm != nullptr;
        Ptrs<P> Pp;
        for (int k = 0; k < kk; ++k) Pp.p[k] = (const typename Pk<P>::T*)V[k0 + k]->data;
        typename Pk<P>::T* wp = (typename Pk<P>::T*)w->data;
        const double* ck = c + 2 * k0;
        if (want) {
            const unsigned g = red_grid(ctx, n);
#define CALL(KK) k_multiaxpy<P, KK, true><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, ck, s, wp, ctx->red_partial, cond);
            SWITCH_K(kk, CALL)
#undef CALL
            CLG_HIP(hipGetLastError());
            CLG_TRY(red_finish(ctx, g, 1, nrm, 1, nullptr, cond, ref, eta, flag));
        } else {
            const unsigned g = ew_grid(ctx, n);
#define CALL(KK) k_multiaxpy<P, KK, false><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, ck, s, wp, nullptr, cond);
            SWITCH_K(kk, CALL)
#undef CALL
        }
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

// w += s * sum_k c_k V_k ; if nrm != null, |w|^2 of the result -> nrm[0] and, if flag != null,
// *flag = (|w|^2 < eta * ref[0]).  cond (nullable) skips the work when *cond == 0.
int blas_multiaxpy_dev(clg_context ctx, clg_field_s* w, clg_field_s* const* V, int nv, const double* c, double s,
                       double* nrm, const int* cond, const double* ref, double eta, int* flag) {
    CLG_CHECK_ARG(nv >= 1, "need at least one vector");
    if (w && w->prec == CLG_F16) {
        CLG_CHECK_ARG(!nrm && !cond, "FP16 multi-axpy has no norm or condition");
        return h16_multiaxpy(ctx, w, V, nv, c, s);
    }
    for (int k = 0; k < nv; ++k) CLG_TRY(check_pair(V[k], w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)));
    long n;
    const int pk = pk_of(w, n);
    if (pk == PK_D) return multiaxpy_p<PK_D>(ctx, w, V, nv, c, s, nrm, cond, ref, eta, flag, n);
    if (pk == PK_F4) return multiaxpy_p<PK_F4>(ctx, w, V, nv, c, s, nrm, cond, ref, eta, flag, n);
    return multiaxpy_p<PK_F2>(ctx, w, V, nv, c, s, nrm, cond, ref, eta, flag, n);
}

// out_i = sum_{j<nin} in_j M[j*nout + i] over FP64 fields of one shape (M: device, interleaved
// complex, nin x nout row-major); out may be the first nout inputs.
int blas_vmm_dev(clg_context ctx, clg_field_s* const* in, int nin, clg_field_s* const* out, int nout, const double* M) {
    CLG_CHECK_ARG(nin >= 1 && nin <= kMaxVmmIn && nout >= 1 && nout <= kMaxVmmOut && nout <= nin,
                  "need 1 <= nout <= nin, nin <= %d, nout <= %d", kMaxVmmIn, kMaxVmmOut);
    for (int j = 0; j < nin; ++j) {
        CLG_TRY(check_pair(in[j], in[0]));
        CLG_CHECK_ARG(in[j]->prec == CLG_F64, "field combinations take FP64 fields");
    }
    for (int i = 0; i < nout; ++i) CLG_TRY(check_pair(out[i], in[0]));
    VmmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.X = nin;
    a.Y = nout;
    a.n = n_complex(in[0]);
    for (int j = 0; j < nin; ++j) a.in[j] = (const double2*)in[j]->data;
    for (int i = 0; i < nout; ++i) a.out[i] = (double2*)out[i]->data;
    const unsigned grid = (unsigned)std::max<long>(1, std::min<long>((a.n + 255) / 256, (long)ctx->num_cu * 8));
    if (nout <= 8) k_vmm<8><<<grid, 256, 0, ctx->stream>>>(a, M);
    else if (nout <= 16) k_vmm<16><<<grid, 256, 0, ctx->stream>>>(a, M);
    else k_vmm<kMaxVmmOut><<<grid, 256, 0, ctx->stream>>>(a, M);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

template <int P>
static int gs_step_p(clg_context ctx, clg_field_s* const* V, int nv, clg_field_s* w, double* hcol, int j, bool shifted,
                     clg_field_s* vnext, double* sc, long n) {
    const unsigned g = red_grid(ctx, n);
    Ptrs<P> Pp;
    for (int k = 0; k < nv; ++k) Pp.p[k] = (const typename Pk<P>::T*)V[k]->data;
    const typename Pk<P>::T* wc = (const typename Pk<P>::T*)w->data;
    typename Pk<P>::T* wp = (typename Pk<P>::T*)w->data;
    // (1) h = V^H w into the column, |w|^2 into sc[0]
#define CALL(KK) k_multidot<P, KK, true><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, wc, ctx->red_partial, nullptr);
    SWITCH_K(nv, CALL)
#undef CALL
    CLG_HIP(hipGetLastError());
    CLG_TRY(red_finish(ctx, g, 2 * nv + 1, hcol, 2 * nv, sc, nullptr, nullptr, 0.0, nullptr));
    // (2) w' = w - V h with V^H w' and |w'|^2, then the column (or, in the fast form, vnext directly)
    typename Pk<P>::T* vp = vnext ? (typename Pk<P>::T*)vnext->data : nullptr;
#define CALL(KK)                                                                                     \
    k_gs_project<P, KK><<<g, kRedBlock, 0, ctx->stream>>>(n, Pp, hcol, wp, ctx->red_partial, sc, vp);  \
    k_gs_finish<KK><<<1, kRedBlock, 0, ctx->stream>>>(ctx->red_partial, (int)g, hcol, j, shifted ? 1 : 0, sc, vp ? 1 : 0);
    SWITCH_K(nv, CALL)
#undef CALL
    CLG_HIP(hipGetLastError());
    // (3) the next basis vector
    if (vnext) {
        const unsigned ge = ew_grid(ctx, n);
#define CALL(KK) k_gs_finalize<P, KK><<<ge, kRedBlock, 0, ctx->stream>>>(n, Pp, wc, vp, sc);
        SWITCH_K(nv, CALL)
#undef CALL
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

// the reduction finish and the Gram-Schmidt finish for kernels of other files
int red_finish_partials(clg_context ctx, unsigned nblk, int nr, double* out, int nout, double* out2, const int* cond,
                        const double* ref, double eta, int* flag) {
    return red_finish(ctx, nblk, nr, out, nout, out2, cond, ref, eta, flag);
}
int gs_finish_launch(clg_context ctx, unsigned nblk, int nv, double* hcol, int j, bool shifted, double* sc,
                     bool fast_ok) {
#define CALL(KK) k_gs_finish<KK><<<1, kRedBlock, 0, ctx->stream>>>(ctx->red_partial, (int)nblk, hcol, j, shifted ? 1 : 0, sc, fast_ok ? 1 : 0);
    SWITCH_K(nv, CALL)
#undef CALL
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// One Arnoldi orthogonalisation: w against the orthonormal V_0..V_{nv-1} (nv <= 16) by classical
// Gram-Schmidt with the DGKS-conditioned second pass, in three streaming passes (projections;
// subtraction fused with the second-pass projections; normalisation, which applies the second pass
// when it is needed) plus two one-workgroup finishes.  With vnext given and no second pass needed
// (|w|^2 - |h|^2 >= |w|^2 / 2, decided on the device), the subtraction pass writes vnext normalised
// by that norm and the other passes return at once; w then keeps w (not w').  hcol: column j of the Hessenberg matrix
// (device, interleaved complex, rows 0..j+1), written with h + [flag] c2 (shifted: delta_ij minus
// that) and H(j+1, j) = +-|w''|.  vnext (nullable) receives w'' / |w''|; w keeps w'.  sc: device
// scratch of 4 + 32 doubles.
int blas_gs_step(clg_context ctx, clg_field_s* const* V, int nv, clg_field_s* w, double* hcol, int j, bool shifted,
                 clg_field_s* vnext, double* sc) {
    CLG_CHECK_ARG(nv >= 1 && nv <= kMaxK, "fused Gram-Schmidt takes 1..%d vectors", kMaxK);
    if (w && w->prec == CLG_F16) return h16_gs_step(ctx, V, nv, w, hcol, j, shifted, vnext, sc);
    for (int k = 0; k < nv; ++k) CLG_TRY(check_pair(V[k], w));
    if (vnext) CLG_TRY(check_pair(vnext, w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * (2 * kMaxK + 1)));
    long n;
    const int pk = pk_of(w, n);
    if (pk == PK_D) return gs_step_p<PK_D>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
    if (pk == PK_F4) return gs_step_p<PK_F4>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
    return gs_step_p<PK_F2>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
}

// copies `count` doubles from device to the context's pinned buffer and waits
int fetch_doubles(clg_context ctx, const double* dev, int count, double* host) {
    CLG_CHECK_ARG(count > 0 && count <= 4096, "bad fetch size");
    CLG_CHECK_ARG(!stream_capturing(ctx), "a blocking call (a reduction read back to the host) while recording a graph");
    CLG_HIP(hipMemcpyAsync(ctx->host_result, dev, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < count; ++i) host[i] = ctx->host_result[i];
    return CLG_OK;
}

// persistent device workspace of the context (grow-only, zero-filled on growth)
int workspace(clg_context ctx, size_t doubles, double** out) {
    if (ctx->ws_cap < doubles) {
        if (ctx->ws) {
            CLG_HIP(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->ws);
            ctx->ws = nullptr;
            ctx->ws_cap = 0;
        }
        CLG_HIP(hipMalloc((void**)&ctx->ws, doubles * sizeof(double)));
        CLG_HIP(hipMemsetAsync(ctx->ws, 0, doubles * sizeof(double), ctx->stream));
        ctx->ws_cap = doubles;
    }
    *out = ctx->ws;
    return CLG_OK;
}

}  // namespace clg

extern "C" {

int clg_axpy(clg_context ctx, clg_field y, double ar, double ai, clg_field x) {
    CLG_CHECK_ARG(ctx, "null context");
    (void)hipSetDevice(ctx->device);
    return blas_axpby(ctx, y, ar, ai, x, 1.0, 0.0);
}

int clg_scale(clg_context ctx, clg_field y, double ar, double ai) {
    CLG_CHECK_ARG(ctx, "null context");
    (void)hipSetDevice(ctx->device);
    return blas_axpby(ctx, y, 0.0, 0.0, y, ar, ai);
}

int clg_dot(clg_context ctx, clg_field a, clg_field b, double out[2]) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    CLG_CHECK_ARG(!stream_capturing(ctx), "blocking call while recording a graph");
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_dot_dev(ctx, a, b, ctx->red_result));
    return fetch_doubles(ctx, ctx->red_result, 2, out);
}

int clg_norm2(clg_context ctx, clg_field a, double* out) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    CLG_CHECK_ARG(!stream_capturing(ctx), "blocking call while recording a graph");
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_norm2_dev(ctx, a, ctx->red_result));
    return fetch_doubles(ctx, ctx->red_result, 1, out);
}

int clg_multidot(clg_context ctx, int n, clg_field* V, clg_field w, double* out) {
    CLG_CHECK_ARG(ctx && V && w && out, "null argument");
    CLG_CHECK_ARG(!stream_capturing(ctx), "blocking call while recording a graph");
    CLG_CHECK_ARG(n >= 1 && n <= 1000, "n must be in [1, 1000]");  // results below the staging area
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_multidot_dev(ctx, V, n, w, ctx->red_result, nullptr, nullptr));
    return fetch_doubles(ctx, ctx->red_result, 2 * n, out);
}

int clg_multiaxpy(clg_context ctx, clg_field w, int n, clg_field* V, const double* coeffs) {
    CLG_CHECK_ARG(ctx && V && w && coeffs, "null argument");
    CLG_CHECK_ARG(n >= 1 && n <= 992, "n must be in [1, 992]");
    for (int k = 0; k < n; ++k) CLG_CHECK_ARG(V[k] && V[k]->data != w->data, "V[%d] must be a field other than w", k);
    (void)hipSetDevice(ctx->device);
    CLG_CHECK_ARG(!stream_capturing(ctx), "clg_multiaxpy uploads host coefficients; not while recording a graph");
    double* c = ctx->red_result + kHostArgs;
    CLG_HIP(hipMemcpyAsync(c, coeffs, sizeof(double) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
    CLG_TRY(blas_multiaxpy_dev(ctx, w, V, n, c, 1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
    CLG_HIP(hipStreamSynchronize(ctx->stream));  // coeffs is the caller's host memory
    return CLG_OK;
}

int clg_vector_multiply_matrix(clg_context ctx, int nres, clg_field* res, int nleft, clg_field* left,
                               const double* matrix) {
    CLG_CHECK_ARG(ctx && res && matrix && (nleft == 0 || left), "null argument");
    CLG_CHECK_ARG(nres >= 1 && nleft >= 0 && nres + nleft <= kMaxVmmIn && nres <= kMaxVmmOut,
                  "need 1 <= nres <= %d and nres + nleft <= %d", kMaxVmmOut, kMaxVmmIn);
    CLG_CHECK_ARG(!stream_capturing(ctx), "clg_vector_multiply_matrix uploads host data; not while recording a graph");
    (void)hipSetDevice(ctx->device);
    std::vector<clg_field_s*> in;
    for (int j = 0; j < nres; ++j) in.push_back(res[j]);
    for (int j = 0; j < nleft; ++j) in.push_back(left[j]);
    for (size_t j = 0; j < in.size(); ++j) {
        CLG_CHECK_ARG(in[j], "null field %d", (int)j);
        for (size_t k = 0; k < j; ++k) CLG_CHECK_ARG(in[k]->data != in[j]->data, "fields %d and %d alias", (int)k, (int)j);
    }
    const int nin = nres + nleft;
    double* m = ctx->red_result + kHostArgs;
    CLG_HIP(hipMemcpyAsync(m, matrix, sizeof(double) * 2 * nin * nres, hipMemcpyHostToDevice, ctx->stream));
    CLG_TRY(blas_vmm_dev(ctx, in.data(), nin, in.data(), nres, m));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    return CLG_OK;
}

int clg_multidot_noalloc(clg_context ctx, int n, clg_field* V, clg_field w, double* out) {
    CLG_CHECK_ARG(ctx && V && w && out, "null argument");
    CLG_CHECK_ARG(n >= 1 && n <= kMaxChunkV, "n must be in [1, %d]: the static partial-sum buffer holds %d vectors",
                  kMaxChunkV, kMaxChunkV);
    for (int k = 0; k < n; ++k) {
        CLG_CHECK_ARG(V[k], "null field V[%d]", k);
        CLG_TRY(check_pair(V[k], w));
    }
    CLG_CHECK_ARG(w->ctx->device == ctx->device, "fields live on device %d, the context on %d", w->ctx->device,
                  ctx->device);
    (void)hipSetDevice(ctx->device);
    double* part = noalloc_scratch(ctx->device);
    CLG_CHECK_ARG(part, "no static partial-sum buffer on device %d (clg_init was not called there)", ctx->device);
    long np;
    const int pk = pk_of(w, np);
    // at most kMaxRedGrid workgroups x (2 kMaxK + 1) partials per chunk of 16 vectors, 4 chunks: the array's size
    if (pk == PK_D) return multidot_p<PK_D>(ctx, V, n, w, out, nullptr, nullptr, np, part);
    if (pk == PK_F4) return multidot_p<PK_F4>(ctx, V, n, w, out, nullptr, nullptr, np, part);
    return multidot_p<PK_F2>(ctx, V, n, w, out, nullptr, nullptr, np, part);
}

}  // extern "C"
