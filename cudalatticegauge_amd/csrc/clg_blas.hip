// clg_blas.hip -- level-1 BLAS and FP64 reductions on fermion fields.
//
// F64 and F32 fermion fields are, whatever their site layout, one contiguous array of
// complex numbers (double2 / float2), so the element-wise kernels run over that flat array.
// Reductions accumulate in FP64 per lane, reduce across the 64-lane wavefront with
// shuffles, across the workgroup through LDS, and across workgroups with a last-block pass;
// the result is written to device memory, so the solvers never wait on the host for a dot
// product (the reference's Dot does ~20 launches and a blocking copy, CudaHelper.cu:626-710).
#include "clg_internal.h"

using namespace clg;
using namespace clgd;

namespace {

const int kRedBlock = 256;
const int kMaxRedGrid = 1024;
const int kMaxK = 16;  // vectors per multi-dot / multi-axpy launch

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double load_coherent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduce NR per-lane values over the grid; the last workgroup to finish writes out[0..NR).
template <int NR>
__device__ void reduce_finish(double (&v)[NR], double* partial, unsigned* counter, double* out) {
    __shared__ double sm[kRedBlock / 64][NR > 0 ? NR : 1];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = v[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            double s = 0;
            for (int w = 0; w < nw; ++w) s += sm[w][r];
            partial[(size_t)blockIdx.x * NR + r] = s;
        }
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += load_coherent(&partial[(size_t)b * NR + r]);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = wave_sum(acc[r]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int r = 0; r < NR; ++r) {
            double s = 0;
            for (int w = 0; w < nw; ++w) s += sm[w][r];
            out[r] = s;
        }
        *counter = 0u;
    }
}

template <typename T> struct Elem;
template <> struct Elem<double2> {
    __device__ static double2 ld(const double2* p, long i) { return p[i]; }
    __device__ static void st(double2* p, long i, double x, double y) { p[i] = make_double2(x, y); }
};
template <> struct Elem<float2> {
    __device__ static double2 ld(const float2* p, long i) {
        const float2 a = p[i];
        return make_double2((double)a.x, (double)a.y);
    }
    __device__ static void st(float2* p, long i, double x, double y) { p[i] = make_float2((float)x, (float)y); }
};

// y = a x + b y   (a, b complex, host scalars)
// READY = false: b == 0, y is only written (so stale NaN/Inf in y never leaks into the result)
template <typename T, bool READY>
__global__ void k_axpby(long n, double ar, double ai, const T* __restrict__ x, double br, double bi, T* y) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 xv = Elem<T>::ld(x, i);
        const double2 yv = READY ? Elem<T>::ld(y, i) : make_double2(0.0, 0.0);
        const double rx = ar * xv.x - ai * xv.y + br * yv.x - bi * yv.y;
        const double ry = ar * xv.y + ai * xv.x + br * yv.y + bi * yv.x;
        Elem<T>::st(y, i, rx, ry);
    }
}
// y += s * c x with c read from device memory (complex), s = +-1
template <typename T>
__global__ void k_axpy_dev(long n, const double* __restrict__ c, double s, const T* __restrict__ x, T* y) {
    const double ar = s * c[0], ai = s * c[1];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 xv = Elem<T>::ld(x, i), yv = Elem<T>::ld(y, i);
        Elem<T>::st(y, i, yv.x + ar * xv.x - ai * xv.y, yv.y + ar * xv.y + ai * xv.x);
    }
}
// dst = src / sqrt(nrm2[0])   (nrm2 on device; zero-safe)
template <typename T>
__global__ void k_scale_inv_sqrt(long n, const double* __restrict__ nrm2, const T* __restrict__ src, T* dst) {
    const double v = nrm2[0];
    const double f = v > 0 ? 1.0 / sqrt(v) : 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 a = Elem<T>::ld(src, i);
        Elem<T>::st(dst, i, a.x * f, a.y * f);
    }
}
// out = sum conj(a) b  (2 doubles)
template <typename T>
__global__ __launch_bounds__(kRedBlock) void k_dot(long n, const T* __restrict__ a, const T* __restrict__ b,
                                                    double* partial, unsigned* counter, double* out) {
    double v[2] = {0, 0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 x = Elem<T>::ld(a, i), y = Elem<T>::ld(b, i);
        v[0] += x.x * y.x + x.y * y.y;
        v[1] += x.x * y.y - x.y * y.x;
    }
    reduce_finish<2>(v, partial, counter, out);
}
template <typename T>
__global__ __launch_bounds__(kRedBlock) void k_norm2(long n, const T* __restrict__ a, double* partial,
                                                      unsigned* counter, double* out) {
    double v[1] = {0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 x = Elem<T>::ld(a, i);
        v[0] += x.x * x.x + x.y * x.y;
    }
    reduce_finish<1>(v, partial, counter, out);
}

template <typename T> struct Ptrs {
    const T* p[kMaxK];
};
// out[2k..2k+1] = sum conj(V_k) w  for k < K
template <typename T, int K>
__global__ __launch_bounds__(kRedBlock) void k_multidot(long n, Ptrs<T> V, const T* __restrict__ w, double* partial,
                                                         unsigned* counter, double* out) {
    double v[2 * K];
#pragma unroll
    for (int k = 0; k < 2 * K; ++k) v[k] = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 y = Elem<T>::ld(w, i);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double2 x = Elem<T>::ld(V.p[k], i);
            v[2 * k] += x.x * y.x + x.y * y.y;
            v[2 * k + 1] += x.x * y.y - x.y * y.x;
        }
    }
    reduce_finish<2 * K>(v, partial, counter, out);
}
// w += s * sum_k c_k V_k  (c on device, 2K doubles); optionally |w|^2 -> nrm_out
template <typename T, int K, bool NORM>
__global__ __launch_bounds__(kRedBlock) void k_multiaxpy(long n, Ptrs<T> V, const double* __restrict__ c, double s,
                                                          T* w, double* partial, unsigned* counter, double* nrm_out) {
    double cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = s * c[2 * k]; ci[k] = s * c[2 * k + 1]; }
    double v[1] = {0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double2 y = Elem<T>::ld(w, i);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double2 x = Elem<T>::ld(V.p[k], i);
            y.x += cr[k] * x.x - ci[k] * x.y;
            y.y += cr[k] * x.y + ci[k] * x.x;
        }
        Elem<T>::st(w, i, y.x, y.y);
        if constexpr (NORM) {
            const double2 q = Elem<T>::ld(w, i);  // norm of the stored (rounded) value
            v[0] += q.x * q.x + q.y * q.y;
        }
    }
    if constexpr (NORM) reduce_finish<1>(v, partial, counter, nrm_out);
}

unsigned red_grid(clg_context ctx, long n) {
    long g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    const long cap = std::min<long>(kMaxRedGrid, (long)ctx->num_cu * 4);
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}
unsigned ew_grid(clg_context ctx, long n) {
    long g = (n + 255) / 256;
    const long cap = (long)ctx->num_cu * 16;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace

namespace clg {

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

long n_complex(const clg_field_s* f) {
    const long nblk = f->subset == CLG_FULL ? 2 : 1;
    return nblk * f->Vh * (f->kind == CLG_WILSON ? 12 : 3);
}

static int check_flat(const clg_field_s* a) {
    CLG_CHECK_ARG(a && is_fermion(a->kind), "need a fermion field");
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

int blas_axpby(clg_context ctx, clg_field_s* y, double ar, double ai, const clg_field_s* x, double br, double bi) {
    CLG_TRY(check_pair(y, x));
    const long n = n_complex(y);
    const unsigned g = ew_grid(ctx, n);
    const bool ready = !(br == 0.0 && bi == 0.0);
    if (y->prec == CLG_F64) {
        if (ready) k_axpby<double2, true><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const double2*)x->data, br, bi, (double2*)y->data);
        else k_axpby<double2, false><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const double2*)x->data, br, bi, (double2*)y->data);
    } else {
        if (ready) k_axpby<float2, true><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const float2*)x->data, br, bi, (float2*)y->data);
        else k_axpby<float2, false><<<g, 256, 0, ctx->stream>>>(n, ar, ai, (const float2*)x->data, br, bi, (float2*)y->data);
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_axpy_dev(clg_context ctx, clg_field_s* y, const double* c, double s, const clg_field_s* x) {
    CLG_TRY(check_pair(y, x));
    const long n = n_complex(y);
    const unsigned g = ew_grid(ctx, n);
    if (y->prec == CLG_F64)
        k_axpy_dev<double2><<<g, 256, 0, ctx->stream>>>(n, c, s, (const double2*)x->data, (double2*)y->data);
    else
        k_axpy_dev<float2><<<g, 256, 0, ctx->stream>>>(n, c, s, (const float2*)x->data, (float2*)y->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_scale_inv_sqrt(clg_context ctx, clg_field_s* dst, const clg_field_s* src, const double* nrm2) {
    CLG_TRY(check_pair(dst, src));
    const long n = n_complex(dst);
    const unsigned g = ew_grid(ctx, n);
    if (dst->prec == CLG_F64)
        k_scale_inv_sqrt<double2><<<g, 256, 0, ctx->stream>>>(n, nrm2, (const double2*)src->data, (double2*)dst->data);
    else
        k_scale_inv_sqrt<float2><<<g, 256, 0, ctx->stream>>>(n, nrm2, (const float2*)src->data, (float2*)dst->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_dot_dev(clg_context ctx, const clg_field_s* a, const clg_field_s* b, double* out) {
    CLG_TRY(check_pair(a, b));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * 2 * kMaxK));
    const long n = n_complex(a);
    const unsigned g = red_grid(ctx, n);
    if (a->prec == CLG_F64)
        k_dot<double2><<<g, kRedBlock, 0, ctx->stream>>>(n, (const double2*)a->data, (const double2*)b->data,
                                                         ctx->red_partial, ctx->red_counter, out);
    else
        k_dot<float2><<<g, kRedBlock, 0, ctx->stream>>>(n, (const float2*)a->data, (const float2*)b->data,
                                                        ctx->red_partial, ctx->red_counter, out);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int blas_norm2_dev(clg_context ctx, const clg_field_s* a, double* out) {
    CLG_TRY(check_flat(a));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * 2 * kMaxK));
    const long n = n_complex(a);
    const unsigned g = red_grid(ctx, n);
    if (a->prec == CLG_F64)
        k_norm2<double2><<<g, kRedBlock, 0, ctx->stream>>>(n, (const double2*)a->data, ctx->red_partial,
                                                           ctx->red_counter, out);
    else
        k_norm2<float2><<<g, kRedBlock, 0, ctx->stream>>>(n, (const float2*)a->data, ctx->red_partial,
                                                          ctx->red_counter, out);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

template <typename T, int K>
static void multidot_k(clg_context ctx, long n, const Ptrs<T>& P, const T* w, double* out) {
    k_multidot<T, K><<<red_grid(ctx, n), kRedBlock, 0, ctx->stream>>>(n, P, w, ctx->red_partial, ctx->red_counter, out);
}
template <typename T, bool NORM, int K>
static void multiaxpy_k(clg_context ctx, long n, const Ptrs<T>& P, const double* c, double s, T* w, double* nrm) {
    const unsigned g = NORM ? red_grid(ctx, n) : ew_grid(ctx, n);
    k_multiaxpy<T, K, NORM><<<g, kRedBlock, 0, ctx->stream>>>(n, P, c, s, w, ctx->red_partial, ctx->red_counter, nrm);
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

// out[2k], out[2k+1] = <V_k, w>, k < nv (chunks of 16 vectors per launch)
int blas_multidot_dev(clg_context ctx, clg_field_s* const* V, int nv, const clg_field_s* w, double* out) {
    CLG_CHECK_ARG(nv >= 1, "need at least one vector");
    for (int k = 0; k < nv; ++k) CLG_TRY(check_pair(V[k], w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * 2 * kMaxK));
    const long n = n_complex(w);
    for (int k0 = 0; k0 < nv; k0 += kMaxK) {
        const int kk = std::min(kMaxK, nv - k0);
        if (w->prec == CLG_F64) {
            Ptrs<double2> P;
            for (int k = 0; k < kk; ++k) P.p[k] = (const double2*)V[k0 + k]->data;
            const double2* wp = (const double2*)w->data;
#define CALL_D(KK) multidot_k<double2, KK>(ctx, n, P, wp, out + 2 * k0)
            SWITCH_K(kk, CALL_D)
#undef CALL_D
        } else {
            Ptrs<float2> P;
            for (int k = 0; k < kk; ++k) P.p[k] = (const float2*)V[k0 + k]->data;
            const float2* wp = (const float2*)w->data;
#define CALL_F(KK) multidot_k<float2, KK>(ctx, n, P, wp, out + 2 * k0)
            SWITCH_K(kk, CALL_F)
#undef CALL_F
        }
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

// w += s * sum_k c_k V_k ; if nrm != null, |w|^2 of the result -> nrm[0]
int blas_multiaxpy_dev(clg_context ctx, clg_field_s* w, clg_field_s* const* V, int nv, const double* c, double s,
                       double* nrm) {
    CLG_CHECK_ARG(nv >= 1, "need at least one vector");
    for (int k = 0; k < nv; ++k) CLG_TRY(check_pair(V[k], w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxRedGrid * 2 * kMaxK));
    const long n = n_complex(w);
    for (int k0 = 0; k0 < nv; k0 += kMaxK) {
        const int kk = std::min(kMaxK, nv - k0);
        const bool last = k0 + kk >= nv;
        const bool want = last && nrm != nullptr;
        if (w->prec == CLG_F64) {
            Ptrs<double2> P;
            for (int k = 0; k < kk; ++k) P.p[k] = (const double2*)V[k0 + k]->data;
            double2* wp = (double2*)w->data;
            if (want) {
#define CALL_DN(KK) multiaxpy_k<double2, true, KK>(ctx, n, P, c + 2 * k0, s, wp, nrm)
                SWITCH_K(kk, CALL_DN)
#undef CALL_DN
            } else {
#define CALL_D0(KK) multiaxpy_k<double2, false, KK>(ctx, n, P, c + 2 * k0, s, wp, nullptr)
                SWITCH_K(kk, CALL_D0)
#undef CALL_D0
            }
        } else {
            Ptrs<float2> P;
            for (int k = 0; k < kk; ++k) P.p[k] = (const float2*)V[k0 + k]->data;
            float2* wp = (float2*)w->data;
            if (want) {
#define CALL_FN(KK) multiaxpy_k<float2, true, KK>(ctx, n, P, c + 2 * k0, s, wp, nrm)
                SWITCH_K(kk, CALL_FN)
#undef CALL_FN
            } else {
#define CALL_F0(KK) multiaxpy_k<float2, false, KK>(ctx, n, P, c + 2 * k0, s, wp, nullptr)
                SWITCH_K(kk, CALL_F0)
#undef CALL_F0
            }
        }
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

// copies `count` doubles from device to the context's pinned buffer and waits
int fetch_doubles(clg_context ctx, const double* dev, int count, double* host) {
    CLG_CHECK_ARG(count > 0 && count <= 4096, "bad fetch size");
    CLG_HIP(hipMemcpyAsync(ctx->host_result, dev, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < count; ++i) host[i] = ctx->host_result[i];
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
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_dot_dev(ctx, a, b, ctx->red_result));
    return fetch_doubles(ctx, ctx->red_result, 2, out);
}

int clg_norm2(clg_context ctx, clg_field a, double* out) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_norm2_dev(ctx, a, ctx->red_result));
    return fetch_doubles(ctx, ctx->red_result, 1, out);
}

}  // extern "C"
