// clg_wilson.hip -- Wilson-square SU(3) hopping kernels (D, D^dagger, even-odd hops, Schur).
//
// Operator (reference _kernelDFermionWilsonSquareSU3, CFieldFermionWilsonSquareSU3.cu:32-163):
//   (D psi)(x) = psi(x) - kappa sum_mu [ (1-g_mu) U_mu(x) s+ psi(x+mu) + (1+g_mu) U_mu(x-mu)^dag s- psi(x-mu) ]
//   D^dagger = g5 D g5, i.e. the same sum with the two projectors exchanged.
// The reference applies full 4x4 gamma products to 4-spinors.  Here each hop is spin
// projected: (1 -+ g_mu) has rank 2, so psi(x+-mu) is projected to a 2-spinor, the link is
// applied to 2 colour vectors instead of 4, and the other two spin rows are rebuilt from
// the gamma table's phases: 1320 flops per site instead of ~2700.
//
// Chiral gamma table (GammaMatrix.h:447-475): (g_mu psi)[r] = i^z[r] psi[idx[r]].  For a
// projector (1 + s g_mu), s = +-1:  h_j = psi_j + s i^z_j psi_idx_j (j = 0,1) and the full
// 4-spinor is (h_0, h_1, s i^z_2 h_idx_2, s i^z_3 h_idx_3), because g_mu^2 = 1.
#include "clg_internal.h"
#include "clg_wilson_common.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <iterator>
#include <map>
#include <mutex>
#include <vector>

using namespace clg;
using namespace clgd;

namespace {

struct HopArgs {
    Geo g;
    const char* in;      // input field base (parity-0 block of a FULL field, or the half field)
    const float* inn;
    char* out;
    float* outn;
    const char* x;       // optional b*x term (same layout as out)
    const float* xn;
    const char* gauge;   // 8 (parity, mu) blocks
    int in_full, out_full;  // 1 if the field holds both parity blocks
    double ar, ai, br, bi;
    int pbase, nblk;     // output parities handled
    int xcd, pinter;     // work decomposition (clgd::map_site)
    int spt;             // logical blocks (sites per thread) per workgroup
    int tile_y, tile_ymaj;  // tiled plane schedule (clgd::tile_block) when tile_y > 0
    int nt_out;             // output written with the non-temporal policy
};

// Neighbour spinors come from the input field, or (STAGED, z and t only) from the copy of this
// (x,y) plane's other-parity sites in LDS: `st`/`stn` hold NPL planes of blockDim.x 16-byte chunks
// (and the FP16 site scales); the site with index h sits at local index h - hbase.

template <int MU, bool DAG, int PM, bool STAGED, typename R, class SIO, class LIO>
__device__ __forceinline__ void hop_mu(const HopArgs& a, const Site& s, int p, int h, C2<R>* acc, const char* st,
                                       const float* stn, int hbase) {
    C2<R> psi[12], hh[6], uh[6], u[9];
    const int Vh = a.g.Vh;
    constexpr bool LDS = STAGED && MU >= 2;
    // forward: (1 - g_mu) U_mu(x) psi(x+mu)  [dagger: (1 + g_mu)]
    {
        const Nbr n = fwd_nbr<MU>(a.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        if constexpr (LDS)
            load_projected<SIO, MU, DAG ? 0 : 2>(st, stn, (int)blockDim.x, (unsigned)(n.h - hbase), (unsigned)(n.h - hbase), n.sign, psi, hh);
        else
            load_projected<SIO, MU, DAG ? 0 : 2>(a.in, a.inn, Vh, pe * SIO::NPL * Vh + n.h, pe * Vh + n.h, n.sign, psi, hh);
        LIO::template load<(PM >= 3)>(a.gauge, Vh, (unsigned)(p * 4 + MU) * LIO::NPL * Vh + h, u);
        su3_mv(u, hh, uh);
        su3_mv(u, hh + 3, uh + 3);
        reconstruct_add<MU, DAG ? 0 : 2>(acc, uh);
    }
    if constexpr (PM == 1 || PM == 3) pin_acc(acc);
    // backward: (1 + g_mu) U_mu(x-mu)^dag psi(x-mu)  [dagger: (1 - g_mu)]
    {
        const Nbr n = bwd_nbr<MU>(a.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        if constexpr (LDS)
            load_projected<SIO, MU, DAG ? 2 : 0>(st, stn, (int)blockDim.x, (unsigned)(n.h - hbase), (unsigned)(n.h - hbase), n.sign, psi, hh);
        else
            load_projected<SIO, MU, DAG ? 2 : 0>(a.in, a.inn, Vh, pe * SIO::NPL * Vh + n.h, pe * Vh + n.h, n.sign, psi, hh);
        LIO::template load<(PM >= 3)>(a.gauge, Vh, (unsigned)(n.p * 4 + MU) * LIO::NPL * Vh + n.h, u);
        su3_mdv(u, hh, uh);
        su3_mdv(u, hh + 3, uh + 3);
        reconstruct_add<MU, DAG ? 2 : 0>(acc, uh);
    }
    if constexpr (PM >= 1) pin_acc(acc);
}

// PM 1..4: register pinning and link cache policy (hop_mu); PM 5..8: PM 1..4 with the z and t
// neighbours staged in LDS.  A staged launch needs one workgroup per (x,y) plane of
// one parity (blockDim = L2*L3/2, L3 even, no padded workgroups): a plane's z and t neighbours
// then all lie in the same plane of the other parity, which the workgroup copies into LDS with
// one coalesced read before any hop (stage_bytes of dynamic LDS).
template <typename R, int SP, int GP, int RECON, bool DAG, bool HASX, int PM>
__global__ __launch_bounds__(256) void k_wilson_hop(HopArgs a) {
    typedef WilsonIO<SP> SIO;
    typedef LinkIO<GP, RECON> LIO;
    constexpr bool STAGED = PM >= 5;
    constexpr int PMH = STAGED ? PM - 4 : PM;
    const unsigned lw = a.tile_y > 0 ? tile_block(a.g, a.nblk, a.tile_y, a.tile_ymaj) : logical_wg(a.xcd);
    for (int vs = 0; vs < a.spt; ++vs) {
    int p, h;
    const bool valid = block_site(lw * (unsigned)a.spt + (unsigned)vs, a.g.Vh, a.nblk, a.pbase, a.pinter, p, h);
    const char* st = nullptr;
    const float* stn = nullptr;
    int hbase = 0;
    const unsigned po = a.out_full ? (unsigned)p : 0u;
    const unsigned eo = po * SIO::NPL * a.g.Vh + h, no = po * a.g.Vh + h;
    C2<R> xv[12];
    if constexpr (STAGED) {
        extern __shared__ uint4 stage[];
        const int B = (int)blockDim.x;
        float* stage_n = reinterpret_cast<float*>(stage + SIO::NPL * B);
        // padded blocks (xcd remap, spt) are whole: every lane of a workgroup agrees on valid,
        // and blocks past the end stay past it
        if (!valid) break;
        // a staged workgroup's own-parity sites have usually left L2 by the end of its hops (its
        // other reads all go to the stage or to other planes): read the x term first, in flight
        // together with the stage loads
        if constexpr (HASX) SIO::load(a.x, a.xn, a.g.Vh, eo, no, xv);
        hbase = h - (int)threadIdx.x;
        const unsigned pe = a.in_full ? (unsigned)(1 - p) : 0u;
#pragma unroll
        for (int k = 0; k < SIO::NPL; ++k) stage[k * B + threadIdx.x] = ldp<uint4>(a.in, a.g.Vh, k, pe * SIO::NPL * a.g.Vh + h);
        if constexpr (SIO::SCALED) stage_n[threadIdx.x] = a.inn[pe * a.g.Vh + h];
        __syncthreads();
        st = reinterpret_cast<const char*>(stage);
        stn = stage_n;
    } else {
        if (!valid) continue;
    }
    const Site s = site_of(a.g, p, h);
    C2<R> acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = cmk<R>(0, 0);
    hop_mu<0, DAG, PMH, STAGED, R, SIO, LIO>(a, s, p, h, acc, st, stn, hbase);
    hop_mu<1, DAG, PMH, STAGED, R, SIO, LIO>(a, s, p, h, acc, st, stn, hbase);
    hop_mu<2, DAG, PMH, STAGED, R, SIO, LIO>(a, s, p, h, acc, st, stn, hbase);
    hop_mu<3, DAG, PMH, STAGED, R, SIO, LIO>(a, s, p, h, acc, st, stn, hbase);
    const C2<R> ca = cmk<R>((R)a.ar, (R)a.ai);
    C2<R> o[12];
    if constexpr (HASX) {
        const C2<R> cb = cmk<R>((R)a.br, (R)a.bi);
        if constexpr (!STAGED) SIO::load(a.x, a.xn, a.g.Vh, eo, no, xv);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            o[k] = cmul(cb, xv[k]);
            cfma(o[k], ca, acc[k]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) o[k] = cmul(ca, acc[k]);
    }
    if (a.nt_out) SIO::template store<R, true>(a.out, a.outn, a.g.Vh, eo, no, o);
    else SIO::store(a.out, a.outn, a.g.Vh, eo, no, o);
    if constexpr (STAGED) __syncthreads();  // the next block rewrites the stage
    }
}

struct HopCfg {
    int pm;     // pin mode: 1 after each half of a direction, 2 after each direction; 3 / 4 the same
                // with non-temporal link loads (links read once per launch stream past the caches);
                // 5..8 as 1..4 with the z and t neighbours staged in LDS (block = L2*L3/2)
    int block;  // threads per workgroup
    int xcd;    // 0: hardware order; 8: XCD-contiguous slabs, for full D with the two parities
                // interleaved block by block; 1: XCD-contiguous slabs, parities one after the other;
                // 10*TY + m (TY in 2, 4, 8; m = 0 x-major, 1 y-major): tiled plane schedule, staged
                // launches only (clgd::tile_block), parities interleaved
    int spt = 1;  // sites per thread: consecutive logical blocks per workgroup (1, 2 or 4)
    int nt = 0;   // 1: the output spinor is stored with the non-temporal policy (it is not read again by
                  // the launch, so it need not take L2 space from the neighbour spinors and links)
};
inline bool tiled_xcd(int xcd) { return xcd >= 20 && xcd <= 81 && (xcd / 10 == 2 || xcd / 10 == 4 || xcd / 10 == 8) && xcd % 10 <= 1; }

template <typename R, int SP, int GP, int RECON, int PM>
void launch_pm(const HopArgs& a0, bool dag, bool hasx, hipStream_t st, int block, int xcd, int spt, int nt) {
    HopArgs a = a0;
    a.nt_out = nt;
    unsigned grid;
    if (tiled_xcd(xcd)) {
        a.xcd = 0;
        a.pinter = 1;
        a.spt = 1;
        a.tile_y = xcd / 10;
        a.tile_ymaj = xcd % 10;
        grid = (unsigned)(a.nblk * a.g.L[0] * a.g.L[1]);
    } else {
        a.xcd = xcd ? 8 : 0;
        a.pinter = xcd == 8 ? 1 : 0;
        a.spt = spt;
        a.tile_y = 0;
        a.tile_ymaj = 0;
        grid = sched_grid(a.g.Vh, a.nblk, block, a.xcd, a.pinter, spt);
    }
    // staged launches: NPL 16-byte chunks per site of the plane, plus the FP16 site scales
    const size_t shm = PM >= 5 ? (size_t)block * (16 * WilsonIO<SP>::NPL + (WilsonIO<SP>::SCALED ? 4 : 0)) : 0;
    if (dag) {
        if (hasx) k_wilson_hop<R, SP, GP, RECON, true, true, PM><<<grid, block, shm, st>>>(a);
        else k_wilson_hop<R, SP, GP, RECON, true, false, PM><<<grid, block, shm, st>>>(a);
    } else {
        if (hasx) k_wilson_hop<R, SP, GP, RECON, false, true, PM><<<grid, block, shm, st>>>(a);
        else k_wilson_hop<R, SP, GP, RECON, false, false, PM><<<grid, block, shm, st>>>(a);
    }
}

template <typename R, int SP, int GP, int RECON>
int launch_t(const HopArgs& a, bool dag, bool hasx, hipStream_t st, HopCfg c) {
    if (c.pm == 8) launch_pm<R, SP, GP, RECON, 8>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 7) launch_pm<R, SP, GP, RECON, 7>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 6) launch_pm<R, SP, GP, RECON, 6>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 5) launch_pm<R, SP, GP, RECON, 5>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 4) launch_pm<R, SP, GP, RECON, 4>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 3) launch_pm<R, SP, GP, RECON, 3>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else if (c.pm == 2) launch_pm<R, SP, GP, RECON, 2>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    else launch_pm<R, SP, GP, RECON, 1>(a, dag, hasx, st, c.block, c.xcd, c.spt, c.nt);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// the workgroup size of an LDS-staged launch on this lattice (one (x,y) plane of one parity), or 0
int staged_block(const Geo& g) {
    if (g.L[3] % 2) return 0;
    const int b = g.L[2] * g.L[3] / 2;
    return (b == 64 || b == 128 || b == 256) ? b : 0;
}
int launch_cfg(const HopArgs& a, int prec, int gprec, int recon, bool dag, bool hasx, hipStream_t st, HopCfg c) {
    // a staged launch needs one workgroup per (x,y) plane of one parity; the tuner only proposes
    // fitting ones, so a misfit comes from a forced configuration and is refused, not replaced
    if (c.pm >= 5) {
        const int sb = staged_block(a.g);
        CLG_CHECK_ARG(sb != 0 && c.block == sb,
                      "LDS-staged pin mode %d needs block = L2*L3/2 in {64,128,256} (lattice %dx%dx%dx%d gives %d), "
                      "asked for block %d",
                      c.pm, a.g.L[0], a.g.L[1], a.g.L[2], a.g.L[3], a.g.L[2] * a.g.L[3] / 2, c.block);
    }
    if (tiled_xcd(c.xcd)) {
        const int TY = c.xcd / 10, TX = 8 / TY;
        CLG_CHECK_ARG(c.pm >= 5 && c.spt == 1 && a.g.L[0] % TX == 0 && a.g.L[1] % TY == 0,
                      "tiled schedule %d needs an LDS-staged pin mode, one site per thread, L0 %% %d == 0 and "
                      "L1 %% %d == 0 (lattice %dx%dx%dx%d)",
                      c.xcd, TX, TY, a.g.L[0], a.g.L[1], a.g.L[2], a.g.L[3]);
    }
    if (prec == CLG_F64) {
        return recon == 12 ? launch_t<double, CLG_F64, CLG_F64, 12>(a, dag, hasx, st, c)
                           : launch_t<double, CLG_F64, CLG_F64, 18>(a, dag, hasx, st, c);
    } else if (prec == CLG_F32 && gprec == CLG_F32) {
        return recon == 12 ? launch_t<float, CLG_F32, CLG_F32, 12>(a, dag, hasx, st, c)
                           : launch_t<float, CLG_F32, CLG_F32, 18>(a, dag, hasx, st, c);
    } else if (prec == CLG_F32) {  // FP16 links, FP32 spinors
        return recon == 12 ? launch_t<float, CLG_F32, CLG_F16, 12>(a, dag, hasx, st, c)
                           : launch_t<float, CLG_F32, CLG_F16, 18>(a, dag, hasx, st, c);
    }
    return recon == 12 ? launch_t<float, CLG_F16, CLG_F16, 12>(a, dag, hasx, st, c)
                       : launch_t<float, CLG_F16, CLG_F16, 18>(a, dag, hasx, st, c);
}

// ---- autotuner: the first launch of a (precision, recon, dagger, x, shape) key times each
//      candidate with HIP events on the launch stream and keeps the fastest.  Outputs are
//      rewritten identically by every candidate, so tuning is invisible to the caller unless
//      out aliases x (then the default is used).  CLG_TUNE=0 disables tuning.
struct TuneKey {
    int prec, gprec, recon, dag, hasx, nblk, L0, L1, L2, L3;
    bool operator<(const TuneKey& o) const {
        return std::memcmp(this, &o, sizeof(TuneKey)) < 0;
    }
};
std::map<TuneKey, HopCfg>& tune_table() {
    static std::map<TuneKey, HopCfg> t;
    return t;
}
std::mutex& tune_mutex() {
    static std::mutex m;
    return m;
}
// the tuning-cache file whose entries are in tune_table() (under tune_mutex)
std::string& tune_loaded() {
    static std::string p;
    return p;
}
std::vector<long> key_vec(const TuneKey& k) {
    return {k.prec, k.gprec, k.recon, k.dag, k.hasx, k.nblk, k.L0, k.L1, k.L2, k.L3};
}
bool valid_cfg(int pm, int blk, int xcd, int spt) {
    return pm >= 1 && pm <= 8 && (blk == 64 || blk == 128 || blk == 256) &&
           (xcd == 0 || xcd == 1 || xcd == 8 || tiled_xcd(xcd)) && (spt == 1 || spt == 2 || spt == 4);
}
// process-wide overrides set through clg_autotune_enable / clg_autotune_force
std::atomic<int> g_tune_mode{-1};  // -1: CLG_TUNE environment (default on), 0 off, 1 on
std::atomic<int> g_forced{0};      // packed pin | block << 4 | xcd << 16 | spt << 24 | nt << 28, 0 = none
// CLG_TUNE_FORCE="pin,block,xcd[,spt[,nt]]" (or clg_autotune_force) fixes the launch configuration of
// every hop launch (used to profile one configuration under counters, where event timings are
// not representative)
bool forced_cfg(HopCfg& c) {
    const int f = g_forced.load();
    if (f) {
        c = HopCfg{f & 15, (f >> 4) & 4095, (f >> 16) & 255, (f >> 24) & 7, (f >> 28) & 1};
        return true;
    }
    const char* e = std::getenv("CLG_TUNE_FORCE");
    if (!e) return false;
    int pm = 0, blk = 0, xcd = 0, spt = 1, nt = 0;
    const int got = std::sscanf(e, "%d,%d,%d,%d,%d", &pm, &blk, &xcd, &spt, &nt);
    if (got < 3 || !valid_cfg(pm, blk, xcd, spt) || (nt != 0 && nt != 1)) return false;
    c = HopCfg{pm, blk, xcd, spt, nt};
    return true;
}

int launch_hop(const HopArgs& a, int prec, int gprec, int recon, bool dag, bool hasx, hipStream_t st) {
    const HopCfg dflt{1, 256, 8};
    TuneKey k;
    std::memset(&k, 0, sizeof(k));
    k.prec = prec; k.gprec = gprec; k.recon = recon; k.dag = dag; k.hasx = hasx; k.nblk = a.nblk;
    k.L0 = a.g.L[0]; k.L1 = a.g.L[1]; k.L2 = a.g.L[2]; k.L3 = a.g.L[3];
    HopCfg cfg = dflt;
    if (forced_cfg(cfg)) return launch_cfg(a, prec, gprec, recon, dag, hasx, st, cfg);
    bool need_tune = false;
    const std::string cpath = tune_cache_path();
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        auto it = tune_table().find(k);
        if (it == tune_table().end() && !cpath.empty() && cpath != tune_loaded()) {
            std::vector<std::pair<std::vector<long>, std::vector<long>>> ent;
            tune_cache_read(cpath, "wilson_hop", ent);
            for (const auto& e : ent) {
                const std::vector<long>& v = e.first;
                const std::vector<long>& c = e.second;
                if (v.size() != 10 || (c.size() != 4 && c.size() != 5) ||
                    !valid_cfg((int)c[0], (int)c[1], (int)c[2], (int)c[3]))
                    continue;
                TuneKey t;
                std::memset(&t, 0, sizeof(t));
                t.prec = (int)v[0]; t.gprec = (int)v[1]; t.recon = (int)v[2]; t.dag = (int)v[3]; t.hasx = (int)v[4];
                t.nblk = (int)v[5]; t.L0 = (int)v[6]; t.L1 = (int)v[7]; t.L2 = (int)v[8]; t.L3 = (int)v[9];
                tune_table()[t] = HopCfg{(int)c[0], (int)c[1], (int)c[2], (int)c[3], c.size() == 5 ? (int)(c[4] != 0) : 0};
            }
            tune_loaded() = cpath;
            it = tune_table().find(k);
        }
        if (it != tune_table().end()) cfg = it->second;
        else need_tune = true;
    }
    const bool aliased = hasx && a.x == a.out;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    if (need_tune && tuning_enabled() && !aliased && cs == hipStreamCaptureStatusNone && (long)a.nblk * a.g.Vh >= 4096) {
        const HopCfg cands[] = {{1, 64, 0}, {1, 128, 0}, {1, 256, 0}, {2, 64, 0}, {2, 128, 0}, {2, 256, 0},
                                {1, 64, 8}, {1, 128, 8}, {1, 256, 8}, {2, 64, 8}, {2, 128, 8}, {2, 256, 8},
                                {3, 64, 8}, {3, 128, 8}, {3, 256, 8}, {4, 64, 8}, {4, 128, 8}, {4, 256, 8},
                                {1, 128, 1}, {1, 256, 1}, {2, 128, 1}, {2, 256, 1},
                                {3, 64, 1}, {3, 128, 1}, {3, 256, 1}, {4, 64, 1}, {4, 128, 1}, {4, 256, 1}};
        hipEvent_t e0, e1;
        CLG_HIP(hipEventCreate(&e0));
        CLG_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        std::vector<HopCfg> all(std::begin(cands), std::end(cands));
        // several sites per thread for the XCD-slab launches
        for (int spt : {2, 4})
            for (int pm : {2, 4})
                for (int blk : {128, 256}) all.push_back(HopCfg{pm, blk, 8, spt});
        if (const int sb = staged_block(a.g)) {
            for (int pm = 5; pm <= 8; ++pm) all.push_back(HopCfg{pm, sb, 8});
            for (int pm : {6, 8}) all.push_back(HopCfg{pm, sb, 8, 2});
            // tiled plane schedules where the plane grid divides
            for (int TY : {2, 4, 8})
                for (int m : {0, 1})
                    if (a.g.L[0] % (8 / TY) == 0 && a.g.L[1] % TY == 0)
                        for (int pm = 5; pm <= 8; ++pm) {
                            all.push_back(HopCfg{pm, sb, 10 * TY + m});
                            if (m == 0 && TY >= 4) all.push_back(HopCfg{pm, sb, 10 * TY + m, 1, 1});
                        }
        }
        for (const HopCfg& c : all) {
            if (c.xcd == 1 && a.nblk != 2) continue;  // one parity: the same launch as xcd 8
            CLG_TRY(launch_cfg(a, prec, gprec, recon, dag, hasx, st, c));  // warm
            CLG_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 3; ++r) CLG_TRY(launch_cfg(a, prec, gprec, recon, dag, hasx, st, c));
            CLG_HIP(hipEventRecord(e1, st));
            CLG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CLG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; cfg = c; }
            const char* v2 = std::getenv("CLG_TUNE_VERBOSE");
            if (v2 && v2[0] == '2')  // every candidate
                std::fprintf(stderr, "[clg tune]   candidate pin=%d block=%d xcd=%d spt=%d nt=%d: %.1f us\n", c.pm,
                             c.block, c.xcd, c.spt, c.nt, ms * 1e3 / 3);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const char* verbose = std::getenv("CLG_TUNE_VERBOSE");
        if (verbose && (verbose[0] == '1' || verbose[0] == '2'))
            std::fprintf(stderr, "[clg tune] wilson hop prec=%d gprec=%d recon=%d dag=%d x=%d nblk=%d L=%dx%dx%dx%d -> pin=%d block=%d xcd=%d spt=%d nt=%d (%.1f us)\n",
                         prec, gprec, recon, (int)dag, (int)hasx, a.nblk, a.g.L[0], a.g.L[1], a.g.L[2], a.g.L[3],
                         cfg.pm, cfg.block, cfg.xcd, cfg.spt, cfg.nt, best * 1e3 / 3);
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_table()[k] = cfg;
        tune_cache_append(cpath, "wilson_hop", key_vec(k), {cfg.pm, cfg.block, cfg.xcd, cfg.spt, cfg.nt});
    }
    return launch_cfg(a, prec, gprec, recon, dag, hasx, st, cfg);
}

int check_gauge(const clg_field_s* f, const clg_field_s* gauge) {
    CLG_CHECK_ARG(gauge && is_gauge(gauge->kind), "need a gauge field");
    CLG_TRY(check_same_lattice(f, gauge));
    CLG_CHECK_ARG(gauge->prec == f->prec || (f->prec == CLG_F32 && gauge->prec == CLG_F16),
                  "gauge precision must match the fermion precision (or FP16 links with FP32 spinors)");
    return CLG_OK;
}

HopArgs base_args(const clg_field_s* out, const clg_field_s* gauge) {
    HopArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = make_geo(out);
    a.gauge = (const char*)gauge->data;
    return a;
}

}  // namespace

namespace clg {

bool tuning_enabled() {
    const int m = g_tune_mode.load();
    if (m >= 0) return m != 0;
    const char* e = std::getenv("CLG_TUNE");
    return !(e && e[0] == '0');
}

// out(full) = coeff * (in - kappa H in) with dagger projectors when dag
int wilson_full(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, double kappa,
                bool dag, double cre, double cim) {
    HopArgs a = base_args(out, gauge);
    a.in = (const char*)in->data;
    a.inn = in->norm;
    a.x = (const char*)in->data;
    a.xn = in->norm;
    a.out = (char*)out->data;
    a.outn = out->norm;
    a.in_full = 1;
    a.out_full = 1;
    a.br = cre;
    a.bi = cim;
    a.ar = -kappa * cre;
    a.ai = -kappa * cim;
    a.pbase = 0;
    a.nblk = 2;
    return launch_hop(a, out->prec, gauge->prec, gauge->kind == CLG_GAUGE12 ? 12 : 18, dag, true, ctx->stream);
}

// out(full) = a H in (no diagonal term): the K of D = 1 - kappa H when a = kappa
int wilson_full_hop(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, double a_re,
                    bool dag) {
    HopArgs a = base_args(out, gauge);
    a.in = (const char*)in->data;
    a.inn = in->norm;
    a.out = (char*)out->data;
    a.outn = out->norm;
    a.in_full = 1;
    a.out_full = 1;
    a.ar = a_re;
    a.pbase = 0;
    a.nblk = 2;
    return launch_hop(a, out->prec, gauge->prec, gauge->kind == CLG_GAUGE12 ? 12 : 18, dag, false, ctx->stream);
}

// out(parity p) = b x + a H_{p,1-p} in
int wilson_hop_half(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, bool dag,
                    double ar, double ai, const clg_field_s* x, double br, double bi) {
    const int p = field_parity(out);
    HopArgs a = base_args(out, gauge);
    a.in = (const char*)in->data;
    a.inn = in->norm;
    a.out = (char*)out->data;
    a.outn = out->norm;
    if (x) {
        a.x = (const char*)x->data;
        a.xn = x->norm;
    }
    a.ar = ar;
    a.ai = ai;
    a.br = br;
    a.bi = bi;
    a.pbase = p;
    a.nblk = 1;
    return launch_hop(a, out->prec, gauge->prec, gauge->kind == CLG_GAUGE12 ? 12 : 18, dag, x != nullptr, ctx->stream);
}

int check_hop_fields(const clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge) {
    CLG_CHECK_ARG(out && in, "null field");
    CLG_CHECK_ARG(out->kind == CLG_WILSON && in->kind == CLG_WILSON, "need Wilson fields");
    CLG_CHECK_ARG(out->subset != CLG_FULL && in->subset != CLG_FULL, "hop needs half-lattice fields");
    CLG_CHECK_ARG(out->subset != in->subset, "hop maps one parity to the other");
    CLG_CHECK_ARG(out->prec == in->prec, "precision mismatch");
    CLG_TRY(check_same_lattice(out, in));
    for (int i = 0; i < 4; ++i) CLG_CHECK_ARG((out->dims[i] & 1) == 0, "even-odd operators need even extents");
    CLG_TRY(check_gauge(out, gauge));
    return CLG_OK;
}

// Schur complement on even sites: Mhat = 1 - kappa^2 H_eo H_oe (dag: projectors swapped)
int wilson_schur_once(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge,
                      double kappa, bool dag, clg_field_s* tmp_odd) {
    CLG_TRY(wilson_hop_half(ctx, tmp_odd, in, gauge, dag, 1.0, 0.0, nullptr, 0.0, 0.0));
    return wilson_hop_half(ctx, out, tmp_odd, gauge, dag, -kappa * kappa, 0.0, in, 1.0, 0.0);
}

}  // namespace clg

extern "C" {

int clg_wilson_apply(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa, int op,
                     int coeff_type, double cre, double cim) {
    CLG_CHECK_ARG(ctx && out && in, "null argument");
    CLG_CHECK_ARG(out != in && out->data != in->data, "out and in must be distinct fields");
    CLG_CHECK_ARG(out->kind == CLG_WILSON && in->kind == CLG_WILSON, "need Wilson fields");
    CLG_CHECK_ARG(out->subset == CLG_FULL && in->subset == CLG_FULL, "need FULL fields");
    CLG_CHECK_ARG(out->prec == in->prec, "precision mismatch");
    CLG_TRY(check_same_lattice(out, in));
    CLG_TRY(check_gauge(out, gauge));
    CLG_CHECK_ARG(op >= EFO_F_D && op <= EFO_F_DDdagger, "operator %d not supported here", op);
    double fr = 1.0, fi = 0.0;
    switch (coeff_type) {
        case EOCT_None: break;
        case EOCT_Minus: fr = -1.0; break;
        case EOCT_Real: fr = cre; break;
        case EOCT_Complex: fr = cre; fi = cim; break;
        default: CLG_CHECK_ARG(false, "bad coefficient type %d", coeff_type);
    }
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < 4; ++i) out->bc[i] = in->bc[i];
    if (op == EFO_F_D || op == EFO_F_Ddagger) return wilson_full(ctx, out, in, gauge, kappa, op == EFO_F_Ddagger, fr, fi);
    // DD = D(D psi), DD^dagger = D(D^dagger psi); the coefficient is applied by the second D
    // (CFieldFermionWilsonSquareSU3.cu:962-1014)
    clg_field_s* tmp = nullptr;
    CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_FULL, 0, &tmp));
    CLG_TRY(wilson_full(ctx, tmp, in, gauge, kappa, op == EFO_F_DDdagger, 1.0, 0.0));
    return wilson_full(ctx, out, tmp, gauge, kappa, false, fr, fi);
}

int clg_wilson_hop(clg_context ctx, clg_field out, clg_field in, clg_field gauge, int dagger, double a_re,
                   double a_im, clg_field x, double b_re, double b_im) {
    CLG_CHECK_ARG(ctx, "null context");
    CLG_TRY(check_hop_fields(out, in, gauge));
    if (x) {
        CLG_CHECK_ARG(x->kind == CLG_WILSON && x->subset == out->subset && x->prec == out->prec,
                      "x must match out");
    }
    CLG_CHECK_ARG(out->data != in->data, "out and in must be distinct");
    (void)hipSetDevice(ctx->device);
    return wilson_hop_half(ctx, out, in, gauge, dagger != 0, a_re, a_im, x, b_re, b_im);
}

int clg_wilson_schur(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa, int op) {
    CLG_CHECK_ARG(ctx && out && in, "null argument");
    CLG_CHECK_ARG(out->subset == CLG_EVEN && in->subset == CLG_EVEN, "Schur operator acts on EVEN fields");
    CLG_CHECK_ARG(out->kind == CLG_WILSON && in->kind == CLG_WILSON && out->prec == in->prec, "need Wilson fields");
    CLG_CHECK_ARG(out->data != in->data, "out and in must be distinct");
    CLG_TRY(check_same_lattice(out, in));
    for (int i = 0; i < 4; ++i) CLG_CHECK_ARG((out->dims[i] & 1) == 0, "even-odd operators need even extents");
    CLG_TRY(check_gauge(out, gauge));
    CLG_CHECK_ARG(op >= EFO_F_D && op <= EFO_F_DDdagger, "operator %d not supported", op);
    (void)hipSetDevice(ctx->device);
    clg_field_s* todd = nullptr;
    CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_ODD, 0, &todd));
    for (int i = 0; i < 4; ++i) { out->bc[i] = in->bc[i]; todd->bc[i] = in->bc[i]; }
    if (op == EFO_F_D || op == EFO_F_Ddagger)
        return wilson_schur_once(ctx, out, in, gauge, kappa, op == EFO_F_Ddagger, todd);
    clg_field_s* teven = nullptr;
    CLG_TRY(scratch_field(ctx, in, CLG_WILSON, in->prec, CLG_EVEN, 0, &teven));
    for (int i = 0; i < 4; ++i) teven->bc[i] = in->bc[i];
    CLG_TRY(wilson_schur_once(ctx, teven, in, gauge, kappa, op == EFO_F_DDdagger, todd));
    return wilson_schur_once(ctx, out, teven, gauge, kappa, false, todd);
}

int clg_time_wilson(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double kappa, int variant,
                    int reps, double* ms_avg) {
    CLG_CHECK_ARG(ctx && ms_avg && reps > 0, "bad arguments");
    (void)hipSetDevice(ctx->device);
    auto one = [&]() -> int {
        if (variant == 0) return clg_wilson_apply(ctx, out, in, gauge, kappa, EFO_F_D, EOCT_None, 1.0, 0.0);
        if (variant == 1) return clg_wilson_hop(ctx, out, in, gauge, 0, 1.0, 0.0, nullptr, 0.0, 0.0);
        return clg_wilson_schur(ctx, out, in, gauge, kappa, EFO_F_D);
    };
    CLG_TRY(one());  // validates arguments, warms up
    CLG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps; ++r) CLG_TRY(one());
    CLG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    CLG_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    CLG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_avg = (double)ms / reps;
    return CLG_OK;
}

// ---- launch autotuner control
int clg_autotune_enable(int mode) {
    CLG_CHECK_ARG(mode >= -1 && mode <= 1, "mode must be -1 (environment), 0 or 1");
    g_tune_mode.store(mode);
    return CLG_OK;
}

int clg_autotune_force(int pin_mode, int block, int xcd, int sites_per_thread) {
    if (pin_mode == 0) {
        g_forced.store(0);
        return CLG_OK;
    }
    if (sites_per_thread == 0) sites_per_thread = 1;
    CLG_CHECK_ARG(valid_cfg(pin_mode, block, xcd, sites_per_thread),
                  "pin_mode 1..8, block 64|128|256, xcd 0|1|8 or a tiled plane schedule 20|21|40|41|80|81, "
                  "sites_per_thread 1|2|4");
    g_forced.store(pin_mode | block << 4 | xcd << 16 | sites_per_thread << 24);
    return CLG_OK;
}

int clg_autotune_reset(void) {
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_table().clear();
        tune_loaded().clear();
    }
    batch_tune_reset();
    ks_tune_reset();
    return CLG_OK;
}

int clg_autotune_count(void) {
    int n = 0;
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        n = (int)tune_table().size();
    }
    return n + batch_tune_count() + ks_tune_count();
}

int clg_autotune_entry(int index, char* buf, int len) {
    CLG_CHECK_ARG(buf && len > 0 && index >= 0, "bad arguments");
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        if (index < (int)tune_table().size()) {
            auto it = tune_table().begin();
            std::advance(it, index);
            const TuneKey& k = it->first;
            const HopCfg& c = it->second;
            std::snprintf(buf, (size_t)len,
                          "wilson_hop prec=%d gauge_prec=%d recon=%d dagger=%d x=%d parities=%d dims=%dx%dx%dx%d "
                          "-> pin=%d block=%d xcd=%d spt=%d nt=%d",
                          k.prec, k.gprec, k.recon, k.dag, k.hasx, k.nblk, k.L0, k.L1, k.L2, k.L3, c.pm, c.block, c.xcd,
                          c.spt, c.nt);
            return CLG_OK;
        }
        index -= (int)tune_table().size();
    }
    const int nb = batch_tune_count();
    if (index < nb) return batch_tune_entry(index, buf, len);
    return ks_tune_entry(index - nb, buf, len);
}

}  // extern "C"
