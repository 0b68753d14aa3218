// clg_gauge_transform.hip -- random local gauge transformations of gauge and fermion fields.
//
// A transformation is one SU(3) matrix g(x) per site, drawn as reunitarise(exp(i eps H(x))) from its
// own Philox stream and kept in the mu = 0 links of an FP64 gauge-format field.  Applied to links it
// gives U_mu(x) -> g(x) U_mu(x) g(x+mu)^dagger (gauge links are periodic, so no boundary sign), and
// to fermions psi(x) -> g(x) psi(x) on every spin row.  The actions and the fermion operators are
// invariant under it (D[U^g] g psi = g D[U] psi): this is what the reference's random gauge fixing
// (CGaugeFixingRandom::GaugeFixing / AlsoFixingFermion) provides to its gauge-invariance test
// (TestCommon.cpp:206-239).
#include "clg_internal.h"

using namespace clg;
using namespace clgd;

namespace {

constexpr int kGtBlock = 256;

__device__ __forceinline__ bool gt_site(const Geo& g, int& p, int& h) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 2L * g.Vh) return false;
    p = (int)(gid / g.Vh);
    h = (int)(gid - (long)p * g.Vh);
    return true;
}

__global__ void k_gt_make(Geo g, double eps, unsigned seed, char* gt) {
    int p, h;
    if (!gt_site(g, p, h)) return;
    const Site s = site_of(g, p, h);
    cd m[9];
    random_su3((unsigned)s.lex, eps, seed, STREAM_GAUGE_TRANSFORM, m);
    LinkIO<CLG_F64, 18>::store(gt, g.Vh, (unsigned)((p * 4 + 0) * 9 * g.Vh + h), m);
}

// lexicographic neighbour x + mu on the periodic lattice -> (parity, half index)
__device__ __forceinline__ void periodic_fwd(const Geo& g, const Site& s, int mu, int& np, int& nh) {
    const int lex = s.c[mu] == g.L[mu] - 1 ? s.lex - (g.L[mu] - 1) * g.stride[mu] : s.lex + g.stride[mu];
    int c = 0;
    int r = lex;
    for (int d = 3; d >= 0; --d) {
        c += r % g.L[d];
        r /= g.L[d];
    }
    np = c & 1;
    nh = lex >> 1;
}

template <int PREC, int RECON>
__global__ void k_gt_links(Geo g, const char* gt, char* U) {
    int p, h;
    if (!gt_site(g, p, h)) return;
    const Site s = site_of(g, p, h);
    cd gx[9];
    LinkIO<CLG_F64, 18>::load(gt, g.Vh, (unsigned)((p * 4 + 0) * 9 * g.Vh + h), gx);
    for (int mu = 0; mu < 4; ++mu) {
        int np, nh;
        periodic_fwd(g, s, mu, np, nh);
        cd gy[9], u[9], t[9], r[9];
        LinkIO<CLG_F64, 18>::load(gt, g.Vh, (unsigned)((np * 4 + 0) * 9 * g.Vh + nh), gy);
        const unsigned e = (unsigned)((p * 4 + mu) * LinkIO<PREC, RECON>::NPL * g.Vh + h);
        LinkIO<PREC, RECON>::load(U, g.Vh, e, u);
        su3_mm(gx, u, t);
        // r = t gy^dagger
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                cd a = cd{0.0, 0.0};
                for (int k = 0; k < 3; ++k) cfma(a, t[3 * i + k], cconj(gy[3 * j + k]));
                r[3 * i + j] = a;
            }
        LinkIO<PREC, RECON>::store(U, g.Vh, e, r);
    }
}

// psi(x) -> g(x) psi(x), NS spin rows of 3 colours (Wilson 4, staggered 1)
template <int PREC, int NS>
__global__ void k_gt_fermion(Geo g, int pbase, int nblk, const char* gt, char* f, float* fn) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)nblk * g.Vh) return;
    const int blk = (int)(gid / g.Vh);
    const int h = (int)(gid - (long)blk * g.Vh);
    const int p = nblk == 2 ? blk : pbase;
    cd gx[9];
    LinkIO<CLG_F64, 18>::load(gt, g.Vh, (unsigned)((p * 4 + 0) * 9 * g.Vh + h), gx);
    cd v[3 * NS], w[3 * NS];
    const unsigned e = (unsigned)(blk * (NS == 4 ? WilsonIO<PREC>::NPL : KSIO<PREC>::NPL) * g.Vh + h);
    const unsigned ne = (unsigned)(blk * g.Vh + h);
    if constexpr (NS == 4) WilsonIO<PREC>::load(f, fn, g.Vh, e, ne, v);
    else KSIO<PREC>::load(f, fn, g.Vh, e, ne, v);
    for (int sp = 0; sp < NS; ++sp) su3_mv(gx, v + 3 * sp, w + 3 * sp);
    if constexpr (NS == 4) WilsonIO<PREC>::store(f, fn, g.Vh, e, ne, w);
    else KSIO<PREC>::store(f, fn, g.Vh, e, ne, w);
}

}  // namespace

extern "C" {

int clg_gauge_transform_random(clg_field gt, double eps, unsigned seed) {
    CLG_CHECK_ARG(gt && gt->kind == CLG_GAUGE && gt->prec == CLG_F64, "the transformation is held in an FP64 gauge field");
    CLG_CHECK_ARG(eps >= 0, "eps must be >= 0");
    (void)hipSetDevice(gt->ctx->device);
    const Geo g = make_geo(gt);
    k_gt_make<<<grid_for(2L * g.Vh, kGtBlock), kGtBlock, 0, gt->ctx->stream>>>(g, eps, seed, (char*)gt->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_gauge_transform_apply(clg_field gt, clg_field target) {
    CLG_CHECK_ARG(gt && target, "null field");
    CLG_CHECK_ARG(gt->kind == CLG_GAUGE && gt->prec == CLG_F64, "the transformation is held in an FP64 gauge field");
    CLG_CHECK_ARG(gt != target && !fields_overlap(gt, target), "target must not be the transformation");
    CLG_TRY(check_same_lattice(gt, target));
    (void)hipSetDevice(gt->ctx->device);
    hipStream_t st = target->ctx->stream;
    const Geo g = make_geo(target);
    const char* G = (const char*)gt->data;
    if (is_gauge(target->kind)) {
        const unsigned grid = grid_for(2L * g.Vh, kGtBlock);
        const bool r12 = target->kind == CLG_GAUGE12;
        char* U = (char*)target->data;
        if (target->prec == CLG_F64) {
            if (r12) k_gt_links<CLG_F64, 12><<<grid, kGtBlock, 0, st>>>(g, G, U);
            else k_gt_links<CLG_F64, 18><<<grid, kGtBlock, 0, st>>>(g, G, U);
        } else if (target->prec == CLG_F32) {
            if (r12) k_gt_links<CLG_F32, 12><<<grid, kGtBlock, 0, st>>>(g, G, U);
            else k_gt_links<CLG_F32, 18><<<grid, kGtBlock, 0, st>>>(g, G, U);
        } else {
            if (r12) k_gt_links<CLG_F16, 12><<<grid, kGtBlock, 0, st>>>(g, G, U);
            else k_gt_links<CLG_F16, 18><<<grid, kGtBlock, 0, st>>>(g, G, U);
        }
    } else {
        const int nblk = target->subset == CLG_FULL ? 2 : 1;
        const int pbase = target->subset == CLG_ODD ? 1 : 0;
        const unsigned grid = grid_for((long)nblk * g.Vh, kGtBlock);
        char* f = (char*)target->data;
        float* fn = target->norm;
#define GT_F(P)                                                                                   \
    if (target->kind == CLG_WILSON) k_gt_fermion<P, 4><<<grid, kGtBlock, 0, st>>>(g, pbase, nblk, G, f, fn); \
    else k_gt_fermion<P, 1><<<grid, kGtBlock, 0, st>>>(g, pbase, nblk, G, f, fn);
        if (target->prec == CLG_F64) { GT_F(CLG_F64) }
        else if (target->prec == CLG_F32) { GT_F(CLG_F32) }
        else { GT_F(CLG_F16) }
#undef GT_F
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // extern "C"
