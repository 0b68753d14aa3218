// clg_gram.hip -- Gram matrices C = A^H B of two lists of FP64 fermion fields (r, c <= 32 fields of n
// complex elements each) on the matrix cores: the tall-skinny complex GEMM of the deflated solver
// (GCRO-DR's U^H U, U^H W and W^H U, which the reference forms one CField::Dot at a time,
// CSolverGCRODR.cpp:700-760, CFieldFermionWilsonSquareSU3.cu:1533-1606).
//
// The element index is the GEMM's long inner dimension.  One v_mfma_f64_16x16x4f64 multiplies a 16 x 4
// tile of A^T (16 fields, 4 elements) by a 4 x 16 tile of B (4 elements, 16 fields); a complex product
// is four real ones,
//   Re C += Ar^T Br + Ai^T Bi,    Im C += Ar^T Bi - Ai^T Br,
// so a wave keeps (r/16) x (c/16) x 2 accumulator tiles and steps through its share of the elements 4
// at a time.  Lane l holds field (l % 16) of a tile at element (l / 16) of the step, and the result tile
// has row (l / 16) + 4 q, column l % 16 in value q of lane l (the builtin's layout, checked on the device).  Each workgroup writes its partial C
// (summed over its 4 waves in LDS); a one-workgroup finish adds the partials in a fixed order.
#include "clg_internal.h"

#include <algorithm>

using namespace clg;
using namespace clgd;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kGramBlock = 256;
constexpr int kGramMaxGrid = 256;
constexpr int kGramMax = 32;
constexpr int kGramU = 4;  // 4-element steps per block of loads

struct GramPtrs {
    const double2* a[kGramMax];
    const double2* b[kGramMax];
};

// partial C of this workgroup: part[blockIdx.x][i][j] (r x c complex, interleaved)
template <int TR, int TC>
__global__ __launch_bounds__(kGramBlock) void k_gram(long n, GramPtrs P, int r, int c, double* part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int fi = lane & 15, ke = lane >> 4;
    v4d acc_re[TR][TC], acc_im[TR][TC];
#pragma unroll
    for (int ti = 0; ti < TR; ++ti)
#pragma unroll
        for (int tj = 0; tj < TC; ++tj) {
            acc_re[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
            acc_im[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
        }
    // waves of the grid take consecutive blocks of kGramU 4-element steps; a lane issues every load of a
    // block before the block's first product, so each wave keeps (TR + TC) kGramU loads in flight
    const long nw = (long)gridDim.x * (kGramBlock / 64);
    const long w0 = (long)blockIdx.x * (kGramBlock / 64) + wv;
    for (long e0 = 4 * kGramU * w0; e0 < n; e0 += 4 * kGramU * nw) {
        double2 va[kGramU][TR], vb[kGramU][TC];
#pragma unroll
        for (int u = 0; u < kGramU; ++u) {
            const long e = e0 + 4 * u + ke;
            const bool in = e < n;
#pragma unroll
            for (int ti = 0; ti < TR; ++ti) {
                const int f = ti * 16 + fi;
                va[u][ti] = (in && f < r) ? P.a[f][e] : double2{0.0, 0.0};
            }
#pragma unroll
            for (int tj = 0; tj < TC; ++tj) {
                const int f = tj * 16 + fi;
                vb[u][tj] = (in && f < c) ? P.b[f][e] : double2{0.0, 0.0};
            }
        }
#pragma unroll
        for (int u = 0; u < kGramU; ++u)
#pragma unroll
            for (int ti = 0; ti < TR; ++ti)
#pragma unroll
                for (int tj = 0; tj < TC; ++tj) {
                    const double ar = va[u][ti].x, ai = va[u][ti].y, br = vb[u][tj].x, bi = vb[u][tj].y;
                    acc_re[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, acc_re[ti][tj], 0, 0, 0);
                    acc_re[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, acc_re[ti][tj], 0, 0, 0);
                    acc_im[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, acc_im[ti][tj], 0, 0, 0);
                    acc_im[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, br, acc_im[ti][tj], 0, 0, 0);
                }
    }
    // sum the 4 waves' tiles in LDS, then write the workgroup's partial
    __shared__ double sm[kGramBlock / 64][kGramMax * kGramMax * 2];
#pragma unroll
    for (int ti = 0; ti < TR; ++ti)
#pragma unroll
        for (int tj = 0; tj < TC; ++tj)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = ti * 16 + ke + 4 * q, j = tj * 16 + fi;
                sm[wv][2 * (i * kGramMax + j)] = acc_re[ti][tj][q];
                sm[wv][2 * (i * kGramMax + j) + 1] = acc_im[ti][tj][q];
            }
    __syncthreads();
    for (int t = threadIdx.x; t < r * c; t += blockDim.x) {
        const int i = t / c, j = t % c;
        double re = 0.0, im = 0.0;
        for (int q = 0; q < kGramBlock / 64; ++q) {
            re += sm[q][2 * (i * kGramMax + j)];
            im += sm[q][2 * (i * kGramMax + j) + 1];
        }
        part[(size_t)blockIdx.x * 2 * r * c + 2 * t] = re;
        part[(size_t)blockIdx.x * 2 * r * c + 2 * t + 1] = im;
    }
}

// out[i * ldo + j] (complex, interleaved) = sum over G partials, thread per entry
__global__ __launch_bounds__(256) void k_gram_finish(const double* __restrict__ part, int G, int r, int c, double* out,
                                                     int ldo) {
    for (int t = threadIdx.x; t < 2 * r * c; t += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < G; ++b) s += part[(size_t)b * 2 * r * c + t];
        const int e = t / 2, i = e / c, j = e % c;
        out[2 * (i * ldo + j) + (t & 1)] = s;
    }
}

}  // namespace

namespace clg {

// C = A^H B for FP64 fermion fields of one shape: A r fields, B c fields (1 <= r, c <= 32); out: device,
// r x c complex, row i at out + 2 i ldo.  Neither blocks nor allocates once the reduction buffer exists.
int blas_gram_dev(clg_context ctx, clg_field_s* const* A, int r, clg_field_s* const* B, int c, double* out, int ldo) {
    CLG_CHECK_ARG(r >= 1 && r <= kGramMax && c >= 1 && c <= kGramMax && ldo >= c, "Gram matrix of 1..%d by 1..%d fields",
                  kGramMax, kGramMax);
    const clg_field_s* f0 = A[0];
    CLG_CHECK_ARG(f0 && is_fermion(f0->kind) && f0->prec == CLG_F64, "Gram matrices take FP64 fermion fields");
    for (int i = 0; i < r + c; ++i) {
        const clg_field_s* f = i < r ? A[i] : B[i - r];
        CLG_CHECK_ARG(f && f->kind == f0->kind && f->prec == CLG_F64 && f->subset == f0->subset, "field %d differs", i);
        CLG_TRY(check_same_lattice(f, f0));
    }
    const long n = (f0->subset == CLG_FULL ? 2 : 1) * f0->Vh * (f0->kind == CLG_WILSON ? 12 : 3);
    GramPtrs P;
    for (int i = 0; i < r; ++i) P.a[i] = (const double2*)A[i]->data;
    for (int j = 0; j < c; ++j) P.b[j] = (const double2*)B[j]->data;
    const long per_wg = 4L * kGramU * kGramBlock;  // elements one workgroup takes per pass of its loop
    const int G = (int)std::max<long>(1, std::min<long>(kGramMaxGrid, (n + per_wg - 1) / per_wg));
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)kGramMaxGrid * 2 * kGramMax * kGramMax));
    double* part = ctx->red_partial;
    hipStream_t st = ctx->stream;
    if (r <= 16 && c <= 16) k_gram<1, 1><<<G, kGramBlock, 0, st>>>(n, P, r, c, part);
    else if (r <= 16) k_gram<1, 2><<<G, kGramBlock, 0, st>>>(n, P, r, c, part);
    else if (c <= 16) k_gram<2, 1><<<G, kGramBlock, 0, st>>>(n, P, r, c, part);
    else k_gram<2, 2><<<G, kGramBlock, 0, st>>>(n, P, r, c, part);
    k_gram_finish<<<1, 256, 0, st>>>(part, G, r, c, out, ldo);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // namespace clg

namespace clg {
int fetch_doubles(clg_context, const double*, int, double*);
}

extern "C" {

int clg_gram_matrix(clg_context ctx, int r, clg_field* A, int c, clg_field* B, double* out) {
    CLG_CHECK_ARG(ctx && A && B && out, "null argument");
    CLG_CHECK_ARG(r >= 1 && r <= kGramMax && c >= 1 && c <= kGramMax, "r and c must be in [1, %d]", kGramMax);
    CLG_CHECK_ARG(!stream_capturing(ctx), "clg_gram_matrix blocks; not while recording a graph");
    (void)hipSetDevice(ctx->device);
    CLG_TRY(blas_gram_dev(ctx, A, r, B, c, ctx->red_result, c));
    return fetch_doubles(ctx, ctx->red_result, 2 * r * c, out);
}

}  // extern "C"

