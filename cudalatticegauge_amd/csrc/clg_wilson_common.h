// clg_wilson_common.h -- spin projection and register-pinning helpers shared by the Wilson
// hopping kernels (single right-hand side and batched).
//
// Chiral gamma table (GammaMatrix.h:447-475): (g_mu psi)[r] = i^z[r] psi[idx[r]].  For a
// projector (1 + s g_mu), s = +-1:  h_j = psi_j + s i^z_j psi_idx_j (j = 0,1) and the full
// 4-spinor is (h_0, h_1, s i^z_2 h_idx_2, s i^z_3 h_idx_3), because g_mu^2 = 1.
#pragma once

#include "clg_internal.h"

namespace clgd {

template <int MU, int R> struct G;  // idx / z4 of gamma_{MU+1}, row R
#define GDEF(MU_, R_, I_, Z_) \
    template <> struct G<MU_, R_> { static constexpr int idx = I_; static constexpr int z = Z_; };
GDEF(0, 0, 3, 3) GDEF(0, 1, 2, 3) GDEF(0, 2, 1, 1) GDEF(0, 3, 0, 1)
GDEF(1, 0, 3, 2) GDEF(1, 1, 2, 0) GDEF(1, 2, 1, 0) GDEF(1, 3, 0, 2)
GDEF(2, 0, 2, 3) GDEF(2, 1, 3, 1) GDEF(2, 2, 0, 1) GDEF(2, 3, 1, 3)
GDEF(3, 0, 2, 0) GDEF(3, 1, 3, 0) GDEF(3, 2, 0, 0) GDEF(3, 3, 1, 0)
#undef GDEF

// SG = 0 for the projector (1 + g_mu), 2 for (1 - g_mu) (adds i^2 = -1 to the phase)
template <int MU, int SG, typename R>
__device__ __forceinline__ void project(const C2<R>* psi, C2<R>* hh) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        hh[c] = cadd(psi[c], mulz4<G<MU, 0>::z + SG>(psi[G<MU, 0>::idx * 3 + c]));
        hh[3 + c] = cadd(psi[3 + c], mulz4<G<MU, 1>::z + SG>(psi[G<MU, 1>::idx * 3 + c]));
    }
}
// neighbour spinor -> spin-projected half-spinor with the boundary sign (and, for FP16 storage,
// the site scale) applied to the 6 projected components rather than the 12 loaded ones
template <class SIO, int MU, int S, typename R>
__device__ __forceinline__ void load_projected(const char* b, const float* nb, int Vh, unsigned e, unsigned ne, int sign,
                                               C2<R>* psi, C2<R>* hh) {
    if constexpr (SIO::SCALED) {
        const float sc = SIO::load_unscaled(b, nb, Vh, e, ne, psi);
        project<MU, S>(psi, hh);
        const R f = (R)(sign < 0 ? -sc : sc);
#pragma unroll
        for (int k = 0; k < 6; ++k) hh[k] = cmk(f * hh[k].x, f * hh[k].y);
    } else {
        SIO::load(b, nb, Vh, e, ne, psi);
        project<MU, S>(psi, hh);
        if (sign < 0) {
#pragma unroll
            for (int k = 0; k < 6; ++k) hh[k] = cmk(-hh[k].x, -hh[k].y);
        }
    }
}

template <int MU, int SG, typename R>
__device__ __forceinline__ void reconstruct_add(C2<R>* acc, const C2<R>* uh) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[c] = cadd(acc[c], uh[c]);
        acc[3 + c] = cadd(acc[3 + c], uh[3 + c]);
        acc[6 + c] = cadd(acc[6 + c], mulz4<G<MU, 2>::z + SG>(uh[G<MU, 2>::idx * 3 + c]));
        acc[9 + c] = cadd(acc[9 + c], mulz4<G<MU, 3>::z + SG>(uh[G<MU, 3>::idx * 3 + c]));
    }
}

// Register-pressure control.  Without it the compiler hoists the loads and projections of
// all eight neighbours ahead of the link products, holding eight half-spinors live (FP64:
// ~390 VGPR+AGPR, one wave per SIMD).  An empty asm that reads and writes the accumulator
// and clobbers memory orders the work: loads of later hops cannot move above it and
// arithmetic of earlier hops cannot sink below it.  PM = 1 pins after every half-hop,
// PM = 2 after every direction (two hops in flight); the tuner picks per shape.
template <typename R>
__device__ __forceinline__ void pin_acc(C2<R>* acc) {
    asm volatile("" : "+v"(acc[0].x), "+v"(acc[0].y), "+v"(acc[1].x), "+v"(acc[1].y), "+v"(acc[2].x), "+v"(acc[2].y),
                 "+v"(acc[3].x), "+v"(acc[3].y), "+v"(acc[4].x), "+v"(acc[4].y), "+v"(acc[5].x), "+v"(acc[5].y)::"memory");
    asm volatile("" : "+v"(acc[6].x), "+v"(acc[6].y), "+v"(acc[7].x), "+v"(acc[7].y), "+v"(acc[8].x), "+v"(acc[8].y),
                 "+v"(acc[9].x), "+v"(acc[9].y), "+v"(acc[10].x), "+v"(acc[10].y), "+v"(acc[11].x), "+v"(acc[11].y)::"memory");
}


}  // namespace clgd
