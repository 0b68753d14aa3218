// clg_arnoldi.hip -- one Arnoldi orthogonalisation step for long Krylov bases (up to 64 vectors):
// w against the orthonormal V_0..V_{nv-1} by classical Gram-Schmidt with the Daniel-Gragg-Kaufman-
// Stewart second pass, the next basis vector written normalised.
//
// The reference orthogonalises with one CField::Dot and one CField::Axpy per basis vector
// (CSolverGMRES.cpp:175-200, CMultiShiftFOM.cpp:150-180, CSolverGCRODR.cpp:620-660): 2 nv launches
// and nv blocking reads per step.  Here a step is five launches with no host wait:
//   A     partials of h = V^H w and |w|^2 (16 vectors per chunk, chunks along blockIdx.y);
//   finA  (one workgroup) h, |w'|^2 = |w|^2 - |h|^2 (V orthonormal), the DGKS flag |w'|^2 < |w|^2 / 2,
//         and, without the flag, the Hessenberg column and 1/|w'|;
//   B     w' = w - V h: without the flag written as vnext = w' / |w'|; with it kept in w, and the same
//         pass (each lane holding its few elements of w' in registers while it reads their V_k again)
//         gives the partials of c2 = V^H w' and |w'|^2;
//   finB  (flag only) c2, the column h + c2, |w''|^2 = |w'|^2 - |c2|^2;
//   C     (flag only) vnext = (w' - V c2) / |w''|.
// (Vectors longer than 256 x 256 x 4 packets take a sixth launch instead of the fused second
// projection: B stores w', then the A kernel projects it.)  So the basis is streamed twice without the
// second pass and three times with it, which a Hermitian operator's Arnoldi process needs at almost
// every step (its w = A v_j keeps most of its norm in v_{j-1} and v_j).  The flag lives on the device,
// so the launches are the same every step (finB and C return at once without it) and a cycle can be
// recorded as a graph.  All sums are FP64 in a fixed order.
#include "clg_internal.h"
#include "clg_packets.h"

#include <algorithm>

using namespace clg;
using namespace clgd;

namespace {

constexpr int kArnBlock = 256;
constexpr int kArnK = 16;               // vectors per chunk
constexpr int kArnMaxV = 64;            // basis vectors per step
constexpr int kArnRows = 2 * kArnK + 1;  // partial rows per chunk: 16 complex dots and |w|^2
constexpr int kArnMaxGrid = 512;        // workgroups (per chunk) of a pass that writes partials
constexpr int kArnFinBlock = 1024;
constexpr int kArnMaxE = 4;             // elements per lane of the fused second-projection pass

template <int P> struct ArnPtrs {
    const typename Pk<P>::T* p[kArnMaxV];
};

// scratch layout (doubles): [0] |w|^2, [1] |w''|^2, [2] 1/|w''|, [3] the flag (as an int), [4, 4+128)
// c2, [132, 132+128) h
constexpr int kScFlag = 3, kScC2 = 4, kScH = 4 + 2 * kArnMaxV;

__device__ __forceinline__ int arn_flag(const double* sc) { return *reinterpret_cast<const int*>(sc + kScFlag); }

// pass A / A': partials of <V_k, w> for the chunk blockIdx.y and, in chunk 0, of |w|^2, written
// transposed: part[(c * kArnRows + r) * G + b] for workgroup b of G.  cond (nullable): run only when
// the flag in sc is set.
template <int P>
__global__ __launch_bounds__(kArnBlock) void k_arn_dots(long n, ArnPtrs<P> V, int nv, const typename Pk<P>::T* __restrict__ w,
                                                        double* part, const double* cond_sc) {
    constexpr int NC = Pk<P>::NC;
    if (cond_sc && !arn_flag(cond_sc)) return;
    const int c = (int)blockIdx.y, k0 = c * kArnK, kk = min(kArnK, nv - k0);
    double v[kArnRows];
#pragma unroll
    for (int r = 0; r < kArnRows; ++r) v[r] = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        typename Pk<P>::T raw[kArnK];
#pragma unroll
        for (int k = 0; k < kArnK; ++k)
            if (k < kk) raw[k] = V.p[k0 + k][i];
        double y[2 * NC];
        Pk<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < kArnK; ++k) {
            if (k < kk) {
                double x[2 * NC];
                Pk<P>::cvt(raw[k], x);
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    v[2 * k] += x[2 * q] * y[2 * q] + x[2 * q + 1] * y[2 * q + 1];
                    v[2 * k + 1] += x[2 * q] * y[2 * q + 1] - x[2 * q + 1] * y[2 * q];
                }
            }
        }
        if (c == 0) {
#pragma unroll
            for (int q = 0; q < 2 * NC; ++q) v[2 * kArnK] += y[q] * y[q];
        }
    }
    __shared__ double sm[kArnBlock / 64][kArnRows];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kArnRows; ++r) v[r] = wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < kArnRows; ++r) sm[wv][r] = v[r];
    }
    __syncthreads();
    const int G = (int)gridDim.x;
    for (int r = threadIdx.x; r < kArnRows; r += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < kArnBlock / 64; ++q) s += sm[q][r];
        part[((size_t)c * kArnRows + r) * G + blockIdx.x] = s;
    }
}

// finA (stage 0) and finB (stage 1, only with the flag): totals of the partials of G workgroups and
// nch chunks, then the step's scalars.  hcol: column j of the Hessenberg matrix (interleaved
// complex; rows 0..nv-1 and j+1), shifted: the column of 1 - K for the operator K the Arnoldi
// process runs on.
__global__ __launch_bounds__(kArnFinBlock) void k_arn_fin(const double* __restrict__ part, int G, int nch, int nv, int j,
                                                          int shifted, double* hcol, double* sc, int stage) {
    if (stage == 1 && !arn_flag(sc)) return;
    __shared__ double tot[4 * kArnRows];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const int rows = nch * kArnRows;
    // wave wv sums rows wv, wv + nw, ... (at most kRowsPerWave), lane l columns l, l + 64, ... (G <=
    // kArnMaxGrid), kCols column blocks per round: every load of a round is issued before the first sum,
    // so the partials, which come from other XCDs' caches through memory, cost one round trip per round
    constexpr int kRowsPerWave = (4 * kArnRows + kArnFinBlock / 64 - 1) / (kArnFinBlock / 64);
    constexpr int kCols = 4;
    double acc[kRowsPerWave];
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) acc[q] = 0.0;
    for (int b0 = 0; b0 < G; b0 += 64 * kCols) {
        double v[kCols][kRowsPerWave];
#pragma unroll
        for (int cb = 0; cb < kCols; ++cb) {
            const int b = b0 + lane + 64 * cb;
#pragma unroll
            for (int q = 0; q < kRowsPerWave; ++q) {
                const int L = wv + q * nw;
                v[cb][q] = (b < G && L < rows) ? part[(size_t)L * G + b] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < kRowsPerWave; ++q) {
#pragma unroll
            for (int cb = 0; cb < kCols; ++cb) acc[q] += v[cb][q];
        }
    }
#pragma unroll
    for (int q = 0; q < kRowsPerWave; ++q) {
        const int L = wv + q * nw;
        const double t = wave_sum(acc[q]);
        if (lane == 0 && L < rows) tot[L] = t;
    }
    __syncthreads();
    // the step's scalars, spread over the threads: thread k < nv handles coefficient k (LDS and global
    // accesses of one thread would otherwise run one after another)
    __shared__ double red[kArnFinBlock / 64];
    const int k = threadIdx.x;
    // dot k sits in chunk k / 16, rows 2 (k % 16) and 2 (k % 16) + 1; |w|^2 in chunk 0, row 32
    double dr = 0.0, di = 0.0;
    if (k < nv) {
        dr = tot[(k / kArnK) * kArnRows + 2 * (k % kArnK)];
        di = tot[(k / kArnK) * kArnRows + 2 * (k % kArnK) + 1];
    }
    const double pw = wave_sum(dr * dr + di * di);
    if (lane == 0) red[wv] = pw;
    __syncthreads();
    double p2 = 0.0;
    for (int q = 0; q < nw; ++q) p2 += red[q];  // the same order in every thread
    const double w2 = tot[2 * kArnK];
    double f2 = w2 - p2;
    double hr = 0.0, hi = 0.0, cr = 0.0, ci = 0.0;
    if (stage == 0) {
        const int flag = !(f2 >= 0.5 * w2);
        if (k < nv) {
            sc[kScH + 2 * k] = dr;
            sc[kScH + 2 * k + 1] = di;
            sc[kScC2 + 2 * k] = 0.0;
            sc[kScC2 + 2 * k + 1] = 0.0;
        }
        if (k == 0) {
            sc[0] = w2;
            *reinterpret_cast<int*>(sc + kScFlag) = flag;
        }
        if (flag) return;  // the column waits for the second projection
        hr = dr;
        hi = di;
    } else {
        if (k < nv) {
            sc[kScC2 + 2 * k] = dr;
            sc[kScC2 + 2 * k + 1] = di;
            hr = sc[kScH + 2 * k];
            hi = sc[kScH + 2 * k + 1];
        }
        cr = dr;
        ci = di;
    }
    if (k < nv) {
        double re = hr + cr, im = hi + ci;
        if (shifted) {
            re = (k == j ? 1.0 : 0.0) - re;
            im = -im;
        }
        hcol[2 * k] = re;
        hcol[2 * k + 1] = im;
    }
    if (k == 0) {
        if (!(f2 > 1e-300)) f2 = 1e-300;  // a breakdown (w in the span) still gives a finite column
        const double f = sqrt(f2);
        hcol[2 * (j + 1)] = shifted ? -f : f;
        hcol[2 * (j + 1) + 1] = 0.0;
        sc[1] = f2;
        sc[2] = 1.0 / f;
    }
}

// the subtraction y = w - sum_k c_k V_k of one element, vectors 16 at a time
template <int P>
__device__ __forceinline__ void arn_subtract(const ArnPtrs<P>& V, int nv, const typename Pk<P>::T* w,
                                             const double* coef, long i, double* y) {
    constexpr int NC = Pk<P>::NC;
    Pk<P>::ld(w, i, y);
    for (int k0 = 0; k0 < nv; k0 += kArnK) {
        typename Pk<P>::T raw[kArnK];
#pragma unroll
        for (int k = 0; k < kArnK; ++k)
            if (k0 + k < nv) raw[k] = V.p[k0 + k][i];
#pragma unroll
        for (int k = 0; k < kArnK; ++k) {
            if (k0 + k < nv) {
                const double cr = coef[2 * (k0 + k)], ci = coef[2 * (k0 + k) + 1];
                double x[2 * NC];
                Pk<P>::cvt(raw[k], x);
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    y[2 * q] -= cr * x[2 * q] - ci * x[2 * q + 1];
                    y[2 * q + 1] -= cr * x[2 * q + 1] + ci * x[2 * q];
                }
            }
        }
    }
}

// y = w - sum_k c_k V_k over the n elements (grid-stride).  mode 0 (pass B), coefficients h: without the
// flag vnext = y / |w'|; with it (unfused form) y = w' is stored in w.  mode 1 (pass C, only with the
// flag), coefficients c2: vnext = y / |w''|.
template <int P>
__global__ __launch_bounds__(kArnBlock) void k_arn_sub(long n, ArnPtrs<P> V, int nv, typename Pk<P>::T* w,
                                                       typename Pk<P>::T* vnext, const double* __restrict__ sc,
                                                       int mode) {
    constexpr int NC = Pk<P>::NC;
    const int flag = arn_flag(sc);
    if (mode == 1 && !flag) return;
    const double* coef = sc + (mode == 0 ? kScH : kScC2);
    const bool to_next = mode == 1 || !flag;
    const double inv = sc[2];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double y[2 * NC];
        arn_subtract<P>(V, nv, w, coef, i, y);
        if (to_next) {
#pragma unroll
            for (int q = 0; q < 2 * NC; ++q) y[q] *= inv;
            Pk<P>::st(vnext, i, y);
        } else {
            Pk<P>::st(w, i, y);
        }
    }
}

// pass B in its fused form: as k_arn_sub mode 0, and with the flag the same pass also yields the
// partials of c2 = V^H w' and |w'|^2 (transposed, as the projection pass writes them): the E <= kArnMaxE
// elements of a lane (element b*256 + t + r*G*256) keep w' in registers while their V_k elements are
// read again, 16 vectors at a time, right after the subtraction.
template <int P>
__global__ __launch_bounds__(kArnBlock) void k_arn_subfuse(long n, ArnPtrs<P> V, int nv, typename Pk<P>::T* w,
                                                           typename Pk<P>::T* vnext, const double* __restrict__ sc,
                                                           double* part, int E) {
    constexpr int NC = Pk<P>::NC;
    const int flag = arn_flag(sc);
    const double* coef = sc + kScH;
    const long stride = (long)gridDim.x * blockDim.x;
    const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!flag) {
        const double inv = sc[2];
        for (int r = 0; r < E; ++r) {
            const long i = i0 + r * stride;
            if (i < n) {
                double y[2 * NC];
                arn_subtract<P>(V, nv, w, coef, i, y);
#pragma unroll
                for (int q = 0; q < 2 * NC; ++q) y[q] *= inv;
                Pk<P>::st(vnext, i, y);
            }
        }
        return;
    }
    double y[kArnMaxE][2 * NC];
#pragma unroll
    for (int r = 0; r < kArnMaxE; ++r) {
#pragma unroll
        for (int q = 0; q < 2 * NC; ++q) y[r][q] = 0.0;
        const long i = i0 + r * stride;
        if (r < E && i < n) {
            arn_subtract<P>(V, nv, w, coef, i, y[r]);
            Pk<P>::st(w, i, y[r]);  // y keeps the stored (rounded) w'
        }
    }
    __shared__ double sm[kArnBlock / 64][kArnRows];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int G = (int)gridDim.x, nch = (nv + kArnK - 1) / kArnK;
    for (int c = 0; c < nch; ++c) {
        const int k0 = c * kArnK, kk = min(kArnK, nv - k0);
        double a[kArnRows];
#pragma unroll
        for (int t = 0; t < kArnRows; ++t) a[t] = 0.0;
#pragma unroll
        for (int r = 0; r < kArnMaxE; ++r) {
            const long i = i0 + r * stride;
            if (r < E && i < n) {
                typename Pk<P>::T raw[kArnK];
#pragma unroll
                for (int k = 0; k < kArnK; ++k)
                    if (k < kk) raw[k] = V.p[k0 + k][i];
#pragma unroll
                for (int k = 0; k < kArnK; ++k) {
                    if (k < kk) {
                        double x[2 * NC];
                        Pk<P>::cvt(raw[k], x);
#pragma unroll
                        for (int q = 0; q < NC; ++q) {
                            a[2 * k] += x[2 * q] * y[r][2 * q] + x[2 * q + 1] * y[r][2 * q + 1];
                            a[2 * k + 1] += x[2 * q] * y[r][2 * q + 1] - x[2 * q + 1] * y[r][2 * q];
                        }
                    }
                }
                if (c == 0) {
#pragma unroll
                    for (int q = 0; q < 2 * NC; ++q) a[2 * kArnK] += y[r][q] * y[r][q];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < kArnRows; ++t) a[t] = wave_sum(a[t]);
        if (c > 0) __syncthreads();  // the previous chunk's sums have been read
        if (lane == 0) {
#pragma unroll
            for (int t = 0; t < kArnRows; ++t) sm[wv][t] = a[t];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kArnRows; t += blockDim.x) {
            double s = 0.0;
            for (int q = 0; q < kArnBlock / 64; ++q) s += sm[q][t];
            part[((size_t)c * kArnRows + t) * G + blockIdx.x] = s;
        }
    }
}

template <int P>
int arn_step_p(clg_context ctx, clg_field_s* const* V, int nv, clg_field_s* w, double* hcol, int j, bool shifted,
               clg_field_s* vnext, double* sc, long n) {
    ArnPtrs<P> A;
    for (int k = 0; k < nv; ++k) A.p[k] = (const typename Pk<P>::T*)V[k]->data;
    typename Pk<P>::T* wp = (typename Pk<P>::T*)w->data;
    typename Pk<P>::T* vp = (typename Pk<P>::T*)(vnext ? vnext->data : w->data);
    const int nch = (nv + kArnK - 1) / kArnK;
    const long lanes = (n + kArnBlock - 1) / kArnBlock;  // workgroups of one element per lane
    // projections: about 256 workgroups in all (each lane sweeps several elements, so the per-workgroup
    // reduction stays a small part of the pass); at most kArnMaxGrid per chunk for the finish
    const int GA = (int)std::max<long>(1, std::min<long>(lanes, std::max(1, 256 / nch)));
    // the streaming passes without partials: enough workgroups to keep the memory system busy
    const unsigned GS = (unsigned)std::max<long>(1, std::min<long>(lanes, (long)ctx->num_cu * 8));
    // the fused second projection: kArnMaxGrid workgroups of E <= kArnMaxE elements per lane
    const int GF = (int)std::max<long>(1, std::min<long>(lanes, kArnMaxGrid));
    const int E = (int)((n + (long)GF * kArnBlock - 1) / ((long)GF * kArnBlock));
    const bool fused = E <= kArnMaxE;
    double* part = ctx->red_partial;
    double* part2 = ctx->red_partial + (size_t)kArnMaxGrid * kArnRows * (kArnMaxV / kArnK);
    hipStream_t st = ctx->stream;
    const int sh = shifted ? 1 : 0;
    k_arn_dots<P><<<dim3(GA, nch), kArnBlock, 0, st>>>(n, A, nv, wp, part, nullptr);
    k_arn_fin<<<1, kArnFinBlock, 0, st>>>(part, GA, nch, nv, j, sh, hcol, sc, 0);
    if (fused) {
        k_arn_subfuse<P><<<GF, kArnBlock, 0, st>>>(n, A, nv, wp, vp, sc, part2, E);
        k_arn_fin<<<1, kArnFinBlock, 0, st>>>(part2, GF, nch, nv, j, sh, hcol, sc, 1);
    } else {
        k_arn_sub<P><<<GS, kArnBlock, 0, st>>>(n, A, nv, wp, vp, sc, 0);
        k_arn_dots<P><<<dim3(GA, nch), kArnBlock, 0, st>>>(n, A, nv, wp, part2, sc);
        k_arn_fin<<<1, kArnFinBlock, 0, st>>>(part2, GA, nch, nv, j, sh, hcol, sc, 1);
    }
    k_arn_sub<P><<<GS, kArnBlock, 0, st>>>(n, A, nv, wp, vp, sc, 1);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // namespace

namespace clg {

int pk_kind(const clg_field_s* f, long& npk);

// One Arnoldi step (see the file comment): hcol gets column j (rows 0..nv-1 and j+1, interleaved
// complex; shifted: the column of 1 - K), vnext (nullptr: w itself) the normalised next vector.  w
// is overwritten (it keeps w' when the second projection ran).  sc: device scratch of at least
// kArnScratch doubles, [1] |w''|^2 and [2] 1/|w''| on return.  Fermion fields of one shape, FP64 or
// FP32, 1 <= nv <= 64.  Capturable: no host synchronisation, no allocation once the context's
// reduction buffer exists.
int arn_step(clg_context ctx, clg_field_s* const* V, int nv, clg_field_s* w, double* hcol, int j, bool shifted,
             clg_field_s* vnext, double* sc) {
    CLG_CHECK_ARG(nv >= 1 && nv <= kArnMaxV, "an Arnoldi step takes 1..%d basis vectors", kArnMaxV);
    CLG_CHECK_ARG(w && w->prec != CLG_F16, "Arnoldi steps on FP16-storage vectors go through the 16-vector form");
    for (int k = 0; k < nv; ++k) {
        CLG_CHECK_ARG(V[k] && V[k]->kind == w->kind && V[k]->prec == w->prec && V[k]->subset == w->subset,
                      "basis vector %d differs from w in kind, precision or subset", k);
        CLG_CHECK_ARG(V[k]->data != w->data, "basis vector %d is w", k);
    }
    if (vnext) CLG_CHECK_ARG(vnext->kind == w->kind && vnext->prec == w->prec && vnext->subset == w->subset,
                             "vnext differs from w in kind, precision or subset");
    CLG_TRY(ensure_reduction_scratch(ctx, 2 * (size_t)kArnMaxGrid * kArnRows * (kArnMaxV / kArnK)));
    long n;
    const int pk = pk_kind(w, n);
    if (pk == PK_D) return arn_step_p<PK_D>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
    if (pk == PK_F4) return arn_step_p<PK_F4>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
    return arn_step_p<PK_F2>(ctx, V, nv, w, hcol, j, shifted, vnext, sc, n);
}

}  // namespace clg
