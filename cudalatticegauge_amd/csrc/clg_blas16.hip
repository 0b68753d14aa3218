// clg_blas16.hip -- Krylov BLAS on FP16-storage Wilson fields, for inner solves that keep their
// Krylov vectors in half precision ("FP16 storage, FP32 arithmetic").
//
// Storage (the hop kernels' WilsonIO<CLG_F16>): per parity block three planes of Vh 16-byte chunks
// (eight halves = four complex numbers each) and one float scale per site, the site's max |component|;
// a value is half * scale.  One thread handles one site: it sees all 24 components, so a store can
// pick the new scale.  Each kernel walks the three chunks of a site in turn, keeping only one chunk
// of each basis vector in registers.
//
// Arithmetic: components are expanded to FP32, per-site sums of products are FP32, and every sum
// over sites is FP64 (partials per workgroup, then the one-workgroup finishes of clg_blas.hip).
// The Gram-Schmidt projections of the second pass are taken against the unrounded w' (the rounding
// to FP16 adds a relative 2^-11 per component that no FP16 vector can resolve anyway).
#include "clg_internal.h"

#include <algorithm>
#include <cstring>

using namespace clg;
using namespace clgd;

namespace clg {
int red_finish_partials(clg_context, unsigned, int, double*, int, double*, const int*, const double*, double, int*);
int gs_finish_launch(clg_context, unsigned, int, double*, int, bool, double*, bool);
int ensure_reduction_scratch(clg_context, size_t);
}  // namespace clg

namespace {

constexpr int kMaxV = 16;     // basis vectors per launch
constexpr int kMaxGrid = 2048;

struct HPtr {
    const uint4* c;  // chunks
    const float* s;  // per-site scales
};
struct HTab {
    HPtr v[kMaxV];
};

#define SITE_LOOP(i, n) for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// y = a y + b x (complex a, b); a == 0 overwrites y without reading it
template <bool READY>
__global__ __launch_bounds__(256) void k_h_axpby(unsigned n, unsigned Vh, double ar, double ai, HPtr x, double br, double bi, uint4* yc,
                          float* ys) {
    SITE_LOOP(i, n) {
        float xv[24], yv[24];
        h16_load_site(x.c, x.s, Vh, i, xv);
        if (READY) h16_load_site(yc, ys, Vh, i, yv);
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const float xr = xv[2 * q], xi = xv[2 * q + 1];
            float re = (float)ar * xr - (float)ai * xi, im = (float)ar * xi + (float)ai * xr;
            if (READY) {
                const float yr = yv[2 * q], yi = yv[2 * q + 1];
                re += (float)br * yr - (float)bi * yi;
                im += (float)br * yi + (float)bi * yr;
            }
            yv[2 * q] = re;
            yv[2 * q + 1] = im;
        }
        h16_store_site(yc, ys, Vh, i, yv);
    }
}

__global__ __launch_bounds__(kRedBlock) void k_h_norm2(unsigned n, unsigned Vh, HPtr x, double* partial) {
    double acc[1] = {0.0};
    SITE_LOOP(i, n) {
        float v[24];
        h16_load_site(x.c, x.s, Vh, i, v);
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < 24; ++q) t += v[q] * v[q];
        acc[0] += t;
    }
    block_partial<1>(acc, partial);
}

// dst = src / sqrt(nrm2[0])
__global__ __launch_bounds__(256) void k_h_scale_inv_sqrt(unsigned n, unsigned Vh, const double* __restrict__ nrm2, HPtr src, uint4* dc,
                                   float* ds) {
    const float f = (float)(1.0 / sqrt(nrm2[0]));
    SITE_LOOP(i, n) {
        float v[24];
        h16_load_site(src.c, src.s, Vh, i, v);
#pragma unroll
        for (int q = 0; q < 24; ++q) v[q] *= f;
        h16_store_site(dc, ds, Vh, i, v);
    }
}

// partials of V_k^H w (2K values) and |w|^2 (one more)
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_h_multidot(unsigned n, unsigned Vh, HTab V, HPtr w, double* partial) {
    double acc[2 * K + 1];
#pragma unroll
    for (int r = 0; r < 2 * K + 1; ++r) acc[r] = 0.0;
    SITE_LOOP(i, n) {
        float sv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) sv[k] = V.v[k].s[i];
        const float sw = w.s[i];
        float re[K], im[K], ww = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k) re[k] = im[k] = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t ci = h16_chunk(i, Vh, ch);
            uint4 raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) raw[k] = V.v[k].c[ci];
            float y[8];
            h16_unpack8(w.c[ci], sw, y);
#pragma unroll
            for (int q = 0; q < 8; ++q) ww += y[q] * y[q];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float x[8];
                h16_unpack8(raw[k], sv[k], x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    re[k] += x[2 * q] * y[2 * q] + x[2 * q + 1] * y[2 * q + 1];
                    im[k] += x[2 * q] * y[2 * q + 1] - x[2 * q + 1] * y[2 * q];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) { acc[2 * k] += re[k]; acc[2 * k + 1] += im[k]; }
        acc[2 * K] += ww;
    }
    block_partial<2 * K + 1>(acc, partial);
}

// w' = w - sum_k h_k V_k, stored to (oc, os) (w itself or the next basis vector), with the
// partials of V_k^H w' and |w'|^2 of the unrounded w'.  fast_ok (oc is the next basis vector): the
// fast form of the FP64 / FP32 step (k_gs_project) -- when |w|^2 - |h|^2 keeps half of |w|^2, w'/|w'|
// is stored and no partials are taken; workgroup 0 records the choice in sc[1..3].
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_h_gs_project(unsigned n, unsigned Vh, HTab V, const double* __restrict__ h,
                                                             const uint4* wc, const float* ws, uint4* oc, float* os,
                                                             double* partial, double* sc, int fast_ok) {
    float hr[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { hr[k] = (float)h[2 * k]; hi[k] = (float)h[2 * k + 1]; }
    if (fast_ok) {
        const double w2 = sc[0];
        double b2 = w2;
#pragma unroll
        for (int k = 0; k < K; ++k) b2 -= h[2 * k] * h[2 * k] + h[2 * k + 1] * h[2 * k + 1];
        const bool fast = b2 >= 0.5 * w2 && b2 > 1e-300;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc[3] = fast ? 2.0 : 0.0;
            if (fast) {
                sc[1] = b2;
                sc[2] = 1.0 / sqrt(b2);
            }
        }
        if (fast) {
            const float inv = (float)(1.0 / sqrt(b2));
            SITE_LOOP(i, n) {
                float sv[K];
#pragma unroll
                for (int k = 0; k < K; ++k) sv[k] = V.v[k].s[i];
                const float sw = ws[i];
                float y[24];
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    const size_t ci = h16_chunk(i, Vh, ch);
                    uint4 raw[K];
#pragma unroll
                    for (int k = 0; k < K; ++k) raw[k] = V.v[k].c[ci];
                    float* yc = y + 8 * ch;
                    h16_unpack8(wc[ci], sw, yc);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        float x[8];
                        h16_unpack8(raw[k], sv[k], x);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            yc[2 * q] -= hr[k] * x[2 * q] - hi[k] * x[2 * q + 1];
                            yc[2 * q + 1] -= hr[k] * x[2 * q + 1] + hi[k] * x[2 * q];
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < 24; ++q) y[q] *= inv;
                h16_store_site(oc, os, Vh, i, y);
            }
            return;
        }
    }
    double acc[2 * K + 1];
#pragma unroll
    for (int r = 0; r < 2 * K + 1; ++r) acc[r] = 0.0;
    SITE_LOOP(i, n) {
        float sv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) sv[k] = V.v[k].s[i];
        const float sw = ws[i];
        float y[24];
        float re[K], im[K], ww = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k) re[k] = im[k] = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t ci = h16_chunk(i, Vh, ch);
            uint4 raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) raw[k] = V.v[k].c[ci];
            float* yc = y + 8 * ch;
            h16_unpack8(wc[ci], sw, yc);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float x[8];
                h16_unpack8(raw[k], sv[k], x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    yc[2 * q] -= hr[k] * x[2 * q] - hi[k] * x[2 * q + 1];
                    yc[2 * q + 1] -= hr[k] * x[2 * q + 1] + hi[k] * x[2 * q];
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float x[8];
                h16_unpack8(raw[k], sv[k], x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    re[k] += x[2 * q] * yc[2 * q] + x[2 * q + 1] * yc[2 * q + 1];
                    im[k] += x[2 * q] * yc[2 * q + 1] - x[2 * q + 1] * yc[2 * q];
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) ww += yc[q] * yc[q];
        }
        h16_store_site(oc, os, Vh, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) { acc[2 * k] += re[k]; acc[2 * k + 1] += im[k]; }
        acc[2 * K] += ww;
    }
    block_partial<2 * K + 1>(acc, partial);
}

// in place on vnext, which holds w': vnext = (w' - flag sum_k c2_k V_k) / |w''| (sc as written by
// the Gram-Schmidt finish).  Without the second pass only the per-site scales change: the stored
// halves of w'/|w'| are those of w'.
template <int K>
__global__ __launch_bounds__(256) void k_h_gs_finalize(unsigned n, unsigned Vh, HTab V, uint4* vc, float* vs,
                                const double* __restrict__ sc) {
    if (sc[3] == 2.0) return;  // the projection pass stored w'/|w'| (fast form)
    const bool flag = sc[3] != 0.0;
    const float inv = (float)sc[2];
    if (!flag) {
        SITE_LOOP(i, n) vs[i] *= inv;
        return;
    }
    float cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = (float)sc[4 + 2 * k]; ci[k] = (float)sc[4 + 2 * k + 1]; }
    SITE_LOOP(i, n) {
        float y[24];
        h16_load_site(vc, vs, Vh, i, y);
        {
            float sv[K];
#pragma unroll
            for (int k = 0; k < K; ++k) sv[k] = V.v[k].s[i];
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const size_t cidx = h16_chunk(i, Vh, ch);
                float* yc = y + 8 * ch;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    float x[8];
                    h16_unpack8(V.v[k].c[cidx], sv[k], x);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        yc[2 * q] -= cr[k] * x[2 * q] - ci[k] * x[2 * q + 1];
                        yc[2 * q + 1] -= cr[k] * x[2 * q + 1] + ci[k] * x[2 * q];
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 24; ++q) y[q] *= inv;
        h16_store_site(vc, vs, Vh, i, y);
    }
}

// y += s * sum_k c_k V_k (c on device)
template <int K>
__global__ __launch_bounds__(256) void k_h_multiaxpy(unsigned n, unsigned Vh, HTab V, const double* __restrict__ c, double s, uint4* yc,
                              float* ys) {
    float cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = (float)(s * c[2 * k]); ci[k] = (float)(s * c[2 * k + 1]); }
    SITE_LOOP(i, n) {
        float y[24];
        h16_load_site(yc, ys, Vh, i, y);
        float sv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) sv[k] = V.v[k].s[i];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t cidx = h16_chunk(i, Vh, ch);
            float* yp = y + 8 * ch;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float x[8];
                h16_unpack8(V.v[k].c[cidx], sv[k], x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    yp[2 * q] += cr[k] * x[2 * q] - ci[k] * x[2 * q + 1];
                    yp[2 * q + 1] += cr[k] * x[2 * q + 1] + ci[k] * x[2 * q];
                }
            }
        }
        h16_store_site(yc, ys, Vh, i, y);
    }
}

#define SWITCH_V(K, CALL)                                   \
    switch (K) {                                            \
        case 1: CALL(1); break;  case 2: CALL(2); break;    \
        case 3: CALL(3); break;  case 4: CALL(4); break;    \
        case 5: CALL(5); break;  case 6: CALL(6); break;    \
        case 7: CALL(7); break;  case 8: CALL(8); break;    \
        case 9: CALL(9); break;  case 10: CALL(10); break;  \
        case 11: CALL(11); break; case 12: CALL(12); break; \
        case 13: CALL(13); break; case 14: CALL(14); break; \
        case 15: CALL(15); break; default: CALL(16); break; \
    }

HPtr hptr(const clg_field_s* f) { return HPtr{(const uint4*)f->data, f->norm}; }
unsigned nsites(const clg_field_s* f) { return (unsigned)((f->subset == CLG_FULL ? 2 : 1) * f->Vh); }
unsigned ew_grid16(clg_context ctx, unsigned n) {
    return std::max(1u, std::min((unsigned)ctx->num_cu * 8, (n + 255) / 256));
}
// reductions: two or more sites per lane at solver sizes (one site per lane spends as long on the
// 2K+1-value workgroup reduction as on the loads; measured 3.5 % of an FP16-vector solve at 32^3x16)
unsigned red_grid16(clg_context ctx, unsigned n) {
    const unsigned cap = std::min((unsigned)kMaxGrid, (unsigned)ctx->num_cu * 2);
    return std::max(1u, std::min(cap, (n + 255) / 256));
}
int check16(const clg_field_s* a, const clg_field_s* b) {
    CLG_CHECK_ARG(a && a->kind == CLG_WILSON && a->prec == CLG_F16 && a->norm,
                  "FP16 BLAS takes FP16-storage Wilson fields");
    if (b) {
        CLG_CHECK_ARG(b->kind == CLG_WILSON && b->prec == CLG_F16 && b->subset == a->subset,
                      "FP16 BLAS fields differ in kind, subset or precision");
        CLG_TRY(check_same_lattice(a, b));
    }
    return CLG_OK;
}
int table(clg_field_s* const* V, int nv, const clg_field_s* w, HTab& t) {
    CLG_CHECK_ARG(nv >= 1 && nv <= kMaxV, "FP16 BLAS takes 1..%d vectors per call", kMaxV);
    std::memset(&t, 0, sizeof(t));
    for (int k = 0; k < nv; ++k) {
        CLG_TRY(check16(V[k], w));
        t.v[k] = hptr(V[k]);
    }
    return CLG_OK;
}

}  // namespace

namespace clg {

int h16_axpby(clg_context ctx, clg_field_s* y, double ar, double ai, const clg_field_s* x, double br, double bi) {
    CLG_TRY(check16(y, x));
    const unsigned n = nsites(y), g = ew_grid16(ctx, n);
    if (br == 0.0 && bi == 0.0)
        k_h_axpby<false><<<g, 256, 0, ctx->stream>>>(n, (unsigned)y->Vh, ar, ai, hptr(x), br, bi, (uint4*)y->data,
                                                     y->norm);
    else
        k_h_axpby<true><<<g, 256, 0, ctx->stream>>>(n, (unsigned)y->Vh, ar, ai, hptr(x), br, bi, (uint4*)y->data,
                                                    y->norm);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int h16_norm2_dev(clg_context ctx, const clg_field_s* a, double* out) {
    CLG_TRY(check16(a, nullptr));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxGrid * (2 * kMaxV + 1)));
    const unsigned n = nsites(a), g = red_grid16(ctx, n);
    k_h_norm2<<<g, kRedBlock, 0, ctx->stream>>>(n, (unsigned)a->Vh, hptr(a), ctx->red_partial);
    CLG_HIP(hipGetLastError());
    return red_finish_partials(ctx, g, 1, out, 1, nullptr, nullptr, nullptr, 0.0, nullptr);
}

int h16_scale_inv_sqrt(clg_context ctx, clg_field_s* dst, const clg_field_s* src, const double* nrm2) {
    CLG_TRY(check16(dst, src));
    const unsigned n = nsites(dst), g = ew_grid16(ctx, n);
    k_h_scale_inv_sqrt<<<g, 256, 0, ctx->stream>>>(n, (unsigned)dst->Vh, nrm2, hptr(src), (uint4*)dst->data,
                                                   dst->norm);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int h16_multiaxpy(clg_context ctx, clg_field_s* w, clg_field_s* const* V, int nv, const double* c, double s) {
    for (int k0 = 0; k0 < nv; k0 += kMaxV) {
        const int kk = std::min(kMaxV, nv - k0);
        HTab t;
        CLG_TRY(table(V + k0, kk, w, t));
        const unsigned n = nsites(w), g = ew_grid16(ctx, n);
        const double* ck = c + 2 * k0;
#define CALL(KK) \
    k_h_multiaxpy<KK><<<g, 256, 0, ctx->stream>>>(n, (unsigned)w->Vh, t, ck, s, (uint4*)w->data, w->norm);
        SWITCH_V(kk, CALL)
#undef CALL
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

// the FP16 form of blas_gs_step (same arguments and results)
int h16_gs_step(clg_context ctx, clg_field_s* const* V, int nv, clg_field_s* w, double* hcol, int j, bool shifted,
                clg_field_s* vnext, double* sc) {
    HTab t;
    CLG_TRY(table(V, nv, w, t));
    if (vnext) CLG_TRY(check16(vnext, w));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kMaxGrid * (2 * kMaxV + 1)));
    const unsigned n = nsites(w), g = red_grid16(ctx, n), Vh = (unsigned)w->Vh;
#define CALL(KK) k_h_multidot<KK><<<g, kRedBlock, 0, ctx->stream>>>(n, Vh, t, hptr(w), ctx->red_partial);
    SWITCH_V(nv, CALL)
#undef CALL
    CLG_HIP(hipGetLastError());
    CLG_TRY(red_finish_partials(ctx, g, 2 * nv + 1, hcol, 2 * nv, sc, nullptr, nullptr, 0.0, nullptr));
    // w' goes straight into the next basis vector when there is one (finalize then works in place)
    clg_field_s* o = vnext ? vnext : w;
#define CALL(KK)                                                                                             \
    k_h_gs_project<KK><<<g, kRedBlock, 0, ctx->stream>>>(n, Vh, t, hcol, (const uint4*)w->data, w->norm, \
                                                         (uint4*)o->data, o->norm, ctx->red_partial, sc, vnext ? 1 : 0);
    SWITCH_V(nv, CALL)
#undef CALL
    CLG_HIP(hipGetLastError());
    CLG_TRY(gs_finish_launch(ctx, g, nv, hcol, j, shifted, sc, vnext != nullptr));
    if (vnext) {
        const unsigned ge = ew_grid16(ctx, n);
#define CALL(KK) k_h_gs_finalize<KK><<<ge, 256, 0, ctx->stream>>>(n, Vh, t, (uint4*)vnext->data, vnext->norm, sc);
        SWITCH_V(nv, CALL)
#undef CALL
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

}  // namespace clg
