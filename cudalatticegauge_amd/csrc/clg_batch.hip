// clg_batch.hip -- batched Wilson operators and solves: many right-hand sides on one lattice and
// several lattices of different extents advance in lock step, with one launch per operator or
// BLAS step for the whole batch.
//
// The reference solves right-hand sides one at a time (e.g. the 12 point sources of
// CFieldFermionWilsonSquareSU3::GetSourcesAtSiteFromPool, CFieldFermionWilsonSquareSU3.cu:1222-1255,
// each through CSLASolver::Solve).  Here:
//   * the hopping kernel takes a table of segments, each (lattice, gauge field, up to NR
//     right-hand sides); a thread handles one site for all NR right-hand sides of its segment, so
//     every link is loaded once per site and tile (the autotuner picks NR in {1, 2, 4}, the
//     workgroup size and the XCD remap per batch shape);
//   * the Krylov BLAS runs on a 2-D grid (blocks x systems) with per-system reductions;
//   * each system keeps its own Krylov space and GMRES(m) recurrences, so every solution is the
//     one the single-system solver produces (same algorithm, same stopping rule per system).
#include "clg_internal.h"
#include "clg_wilson_common.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int fetch_doubles(clg_context, const double*, int, double*);
int workspace(clg_context, size_t, double**);
int ensure_reduction_scratch(clg_context, size_t);
int blas_axpby(clg_context, clg_field_s*, double, double, const clg_field_s*, double, double);
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int gauge_convert(clg_context ctx, clg_field_s* dst, const clg_field_s* src);
}  // namespace clg

namespace {

constexpr int kMaxSys = 16;  // systems (right-hand sides) per batch
constexpr int kMaxM = 16;    // Krylov dimension of the batched GMRES

// ================================================================== batched hopping kernel
struct BSeg {
    Geo g;
    const char* gauge;
    int slot0, nr;  // right-hand-side slots [slot0, slot0 + nr) of this segment
    int blk0;       // first logical workgroup of the segment
    double ar;      // output = ar * hop + b * x
};
struct BHopArgs {
    BSeg seg[kMaxSys];
    const char* in[kMaxSys];
    char* out[kMaxSys];
    const char* x[kMaxSys];
    const float* inn[kMaxSys];  // per-site scales of FP16-storage fields (null otherwise)
    float* outn[kMaxSys];
    const float* xn[kMaxSys];
    int nseg;
    int in_full, out_full;  // 1: the field holds both parity blocks
    int pbase, nblk;        // output parities
    int xcd;
    double br;
};

// NT: links with the non-temporal cache policy (a half-lattice hop reads every link once)
template <int MU, bool DAG, int NR, bool NT, typename R, class SIO, class LIO>
__device__ __forceinline__ void bhop_mu(const BHopArgs& a, const BSeg& sg, const Site& s, int p, int h,
                                        C2<R> (&acc)[NR][12]) {
    const int Vh = sg.g.Vh;
    C2<R> u[9];
    {
        const Nbr n = fwd_nbr<MU>(sg.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        LIO::template load<NT>(sg.gauge, Vh, (unsigned)(p * 4 + MU) * LIO::NPL * Vh + h, u);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r < sg.nr) {
                C2<R> psi[12], hh[6], uh[6];
                load_projected<SIO, MU, DAG ? 0 : 2>(a.in[sg.slot0 + r], a.inn[sg.slot0 + r], Vh, pe * SIO::NPL * Vh + n.h,
                                              pe * Vh + n.h, n.sign, psi, hh);
                su3_mv(u, hh, uh);
                su3_mv(u, hh + 3, uh + 3);
                reconstruct_add<MU, DAG ? 0 : 2>(acc[r], uh);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) pin_acc(acc[r]);
    {
        const Nbr n = bwd_nbr<MU>(sg.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        LIO::template load<NT>(sg.gauge, Vh, (unsigned)(n.p * 4 + MU) * LIO::NPL * Vh + n.h, u);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r < sg.nr) {
                C2<R> psi[12], hh[6], uh[6];
                load_projected<SIO, MU, DAG ? 2 : 0>(a.in[sg.slot0 + r], a.inn[sg.slot0 + r], Vh, pe * SIO::NPL * Vh + n.h,
                                              pe * Vh + n.h, n.sign, psi, hh);
                su3_mdv(u, hh, uh);
                su3_mdv(u, hh + 3, uh + 3);
                reconstruct_add<MU, DAG ? 2 : 0>(acc[r], uh);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) pin_acc(acc[r]);
}

template <typename R, int SP, int GP, int RECON, bool DAG, bool HASX, int NR, bool NT>
__global__ __launch_bounds__(256) void k_wilson_hop_batch(BHopArgs a) {
    typedef WilsonIO<SP> SIO;
    typedef LinkIO<GP, RECON> LIO;
    unsigned wg = blockIdx.x;
    if (a.xcd > 1) {
        const unsigned per = gridDim.x / (unsigned)a.xcd;
        wg = (wg % (unsigned)a.xcd) * per + wg / (unsigned)a.xcd;
    }
    int j = 0;
    while (j + 1 < a.nseg && (int)wg >= a.seg[j + 1].blk0) ++j;
    const BSeg& sg = a.seg[j];
    const long gid = (long)((int)wg - sg.blk0) * blockDim.x + threadIdx.x;
    if (gid < 0 || gid >= (long)a.nblk * sg.g.Vh) return;
    const int Vh = sg.g.Vh;
    const int p = a.nblk == 2 ? (int)(gid / Vh) : a.pbase;
    const int h = (int)(gid - (a.nblk == 2 ? (long)p * Vh : 0L));
    const Site s = site_of(sg.g, p, h);
    C2<R> acc[NR][12];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int k = 0; k < 12; ++k) acc[r][k] = cmk<R>(0, 0);
    bhop_mu<0, DAG, NR, NT, R, SIO, LIO>(a, sg, s, p, h, acc);
    bhop_mu<1, DAG, NR, NT, R, SIO, LIO>(a, sg, s, p, h, acc);
    bhop_mu<2, DAG, NR, NT, R, SIO, LIO>(a, sg, s, p, h, acc);
    bhop_mu<3, DAG, NR, NT, R, SIO, LIO>(a, sg, s, p, h, acc);
    const unsigned po = a.out_full ? (unsigned)p : 0u;
    const unsigned eo = po * SIO::NPL * Vh + h, no = po * Vh + h;
    const R ca = (R)sg.ar;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r < sg.nr) {
            C2<R> o[12];
            if constexpr (HASX) {
                C2<R> xv[12];
                SIO::load(a.x[sg.slot0 + r], a.xn[sg.slot0 + r], Vh, eo, no, xv);
                const R cb = (R)a.br;
#pragma unroll
                for (int k = 0; k < 12; ++k) o[k] = cmk<R>(cb * xv[k].x + ca * acc[r][k].x, cb * xv[k].y + ca * acc[r][k].y);
            } else {
#pragma unroll
                for (int k = 0; k < 12; ++k) o[k] = cmk<R>(ca * acc[r][k].x, ca * acc[r][k].y);
            }
            SIO::store(a.out[sg.slot0 + r], a.outn[sg.slot0 + r], Vh, eo, no, o);
        }
    }
}

struct BCfg {
    int nr, block, xcd;
    int nt = 0;  // 1: non-temporal link loads
};

template <typename R, int SP, int GP, int RECON, bool DAG, bool HASX, bool NT>
int launch_nr(const BHopArgs& a, unsigned grid, int block, int nr, hipStream_t st) {
    if (nr == 4) {
        if constexpr (SP == CLG_F64) return CLG_UNSUPPORTED;
        else k_wilson_hop_batch<R, SP, GP, RECON, DAG, HASX, 4, NT><<<grid, block, 0, st>>>(a);
    } else if (nr == 2) {
        k_wilson_hop_batch<R, SP, GP, RECON, DAG, HASX, 2, NT><<<grid, block, 0, st>>>(a);
    } else {
        k_wilson_hop_batch<R, SP, GP, RECON, DAG, HASX, 1, NT><<<grid, block, 0, st>>>(a);
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
template <typename R, int SP, int GP, int RECON, bool NT>
int launch_dxn(const BHopArgs& a, bool dag, bool hasx, unsigned grid, int block, int nr, hipStream_t st) {
    if (dag) {
        return hasx ? launch_nr<R, SP, GP, RECON, true, true, NT>(a, grid, block, nr, st)
                    : launch_nr<R, SP, GP, RECON, true, false, NT>(a, grid, block, nr, st);
    }
    return hasx ? launch_nr<R, SP, GP, RECON, false, true, NT>(a, grid, block, nr, st)
                : launch_nr<R, SP, GP, RECON, false, false, NT>(a, grid, block, nr, st);
}
template <typename R, int SP, int GP, int RECON>
int launch_dx(const BHopArgs& a, bool dag, bool hasx, unsigned grid, int block, int nr, int nt, hipStream_t st) {
    return nt ? launch_dxn<R, SP, GP, RECON, true>(a, dag, hasx, grid, block, nr, st)
              : launch_dxn<R, SP, GP, RECON, false>(a, dag, hasx, grid, block, nr, st);
}

// One batched hop launch description: systems in launch order with their fields.
struct HopBatch {
    int n = 0;
    const clg_field_s* gauge[kMaxSys];
    const clg_field_s* in[kMaxSys];
    clg_field_s* out[kMaxSys];
    const clg_field_s* x[kMaxSys];
    double a[kMaxSys];
};

// build the kernel arguments for a tile size; returns the padded grid
int build_args(const HopBatch& B, int nr, int block, int xcd, bool hasx, double br, int pbase, int nblk,
               BHopArgs& a, unsigned& grid) {
    std::memset(&a, 0, sizeof(a));
    // group systems by (gauge field, coefficient), keeping first-appearance order
    std::vector<int> order;
    std::vector<bool> used(B.n, false);
    for (int i = 0; i < B.n; ++i) {
        if (used[i]) continue;
        for (int k = i; k < B.n; ++k) {
            if (!used[k] && B.gauge[k] == B.gauge[i] && B.a[k] == B.a[i]) {
                used[k] = true;
                order.push_back(k);
            }
        }
    }
    int slot = 0, blk = 0, nseg = 0;
    size_t i = 0;
    while (i < order.size()) {
        const int s0 = order[i];
        int cnt = 0;
        while (i + cnt < order.size() && cnt < nr && B.gauge[order[i + cnt]] == B.gauge[s0] &&
               B.a[order[i + cnt]] == B.a[s0])
            ++cnt;
        CLG_CHECK_ARG(nseg < kMaxSys, "too many batch segments");
        BSeg& sg = a.seg[nseg++];
        sg.g = make_geo(B.in[s0]);
        sg.gauge = (const char*)B.gauge[s0]->data;
        sg.slot0 = slot;
        sg.nr = cnt;
        sg.blk0 = blk;
        sg.ar = B.a[s0];
        for (int r = 0; r < cnt; ++r) {
            const int sys = order[i + r];
            a.in[slot + r] = (const char*)B.in[sys]->data;
            a.out[slot + r] = (char*)B.out[sys]->data;
            a.x[slot + r] = hasx ? (const char*)B.x[sys]->data : nullptr;
            a.inn[slot + r] = B.in[sys]->norm;
            a.outn[slot + r] = B.out[sys]->norm;
            a.xn[slot + r] = hasx ? B.x[sys]->norm : nullptr;
        }
        slot += cnt;
        blk += (int)grid_for((long)nblk * sg.g.Vh, block);
        i += cnt;
    }
    a.nseg = nseg;
    a.in_full = B.in[0]->subset == CLG_FULL ? 1 : 0;
    a.out_full = B.out[0]->subset == CLG_FULL ? 1 : 0;
    a.pbase = pbase;
    a.nblk = nblk;
    a.xcd = xcd;
    a.br = br;
    grid = (unsigned)blk;
    if (xcd > 1) grid = (grid + xcd - 1) / xcd * xcd;
    return CLG_OK;
}

int launch_batch_cfg(const HopBatch& B, bool dag, bool hasx, double br, int pbase, int nblk, const BCfg& c,
                     hipStream_t st) {
    BHopArgs a;
    unsigned grid = 0;
    CLG_TRY(build_args(B, c.nr, c.block, c.xcd, hasx, br, pbase, nblk, a, grid));
    const clg_field_s* f = B.in[0];
    const clg_field_s* gf = B.gauge[0];
    const int recon = gf->kind == CLG_GAUGE12 ? 12 : 18;
    if (f->prec == CLG_F64) return launch_dx<double, CLG_F64, CLG_F64, 18>(a, dag, hasx, grid, c.block, c.nr, c.nt, st);
    if (f->prec == CLG_F16) return launch_dx<float, CLG_F16, CLG_F16, 12>(a, dag, hasx, grid, c.block, c.nr, c.nt, st);
    if (gf->prec == CLG_F16) return launch_dx<float, CLG_F32, CLG_F16, 12>(a, dag, hasx, grid, c.block, c.nr, c.nt, st);
    (void)recon;
    return launch_dx<float, CLG_F32, CLG_F32, 18>(a, dag, hasx, grid, c.block, c.nr, c.nt, st);
}

// tuner: key = precision combination, dagger, x-term, parities and the batch shape
std::map<std::vector<long>, BCfg>& btune() {
    static std::map<std::vector<long>, BCfg> t;
    return t;
}
std::mutex& btune_mutex() {
    static std::mutex m;
    return m;
}
std::string& btune_loaded() {  // the tuning-cache file whose entries are in btune()
    static std::string p;
    return p;
}

int check_batch(const HopBatch& B) {
    CLG_CHECK_ARG(B.n >= 1 && B.n <= kMaxSys, "batch needs 1..%d systems", kMaxSys);
    const clg_field_s* f0 = B.in[0];
    for (int i = 0; i < B.n; ++i) {
        CLG_CHECK_ARG(B.in[i] && B.out[i] && B.gauge[i], "null field in batch slot %d", i);
        const clg_field_s *in = B.in[i], *out = B.out[i], *g = B.gauge[i];
        CLG_CHECK_ARG(in->kind == CLG_WILSON && out->kind == CLG_WILSON, "batched hops take Wilson fields");
        CLG_CHECK_ARG(in->prec == f0->prec && out->prec == f0->prec && in->subset == f0->subset &&
                          out->subset == B.out[0]->subset,
                      "all batch slots need one precision and subset layout");
        CLG_CHECK_ARG(is_gauge(g->kind), "slot %d: need a gauge field", i);
        if (in->prec == CLG_F64)
            CLG_CHECK_ARG(g->prec == CLG_F64 && g->kind == CLG_GAUGE, "FP64 batch needs FP64 18-real links");
        else if (in->prec == CLG_F16)
            CLG_CHECK_ARG(g->prec == CLG_F16 && g->kind == CLG_GAUGE12, "FP16 batch needs FP16 12-real links");
        else
            CLG_CHECK_ARG((g->prec == CLG_F32 && g->kind == CLG_GAUGE) || (g->prec == CLG_F16 && g->kind == CLG_GAUGE12),
                          "FP32 batch needs FP32 18-real or FP16 12-real links");
        CLG_CHECK_ARG(g->prec == B.gauge[0]->prec && g->kind == B.gauge[0]->kind, "one link format per batch");
        CLG_TRY(check_same_lattice(in, g));
        CLG_TRY(check_same_lattice(out, g));
        for (int d = 0; d < 4; ++d)
            CLG_CHECK_ARG(in->bc[d] == out->bc[d], "slot %d: in/out boundary conditions differ", i);
        if (B.x[i]) {
            CLG_CHECK_ARG(B.x[i]->prec == in->prec && B.x[i]->subset == out->subset, "slot %d: x layout differs", i);
            CLG_TRY(check_same_lattice(B.x[i], g));
        }
        CLG_CHECK_ARG(in->data != out->data, "slot %d: out must not alias in", i);
    }
    // every workgroup reads in[j] (neighbours) and x[j] while others write out[i]: an output may
    // share storage with nothing but its own slot's x (read and written by the same thread)
    for (int i = 0; i < B.n; ++i)
        for (int j = 0; j < B.n; ++j) {
            CLG_CHECK_ARG(!fields_overlap(B.out[i], B.in[j]), "out of slot %d overlaps in of slot %d", i, j);
            if (j != i) {
                CLG_CHECK_ARG(!fields_overlap(B.out[i], B.out[j]), "outs of slots %d and %d overlap", i, j);
                CLG_CHECK_ARG(!B.x[j] || !fields_overlap(B.out[i], B.x[j]), "out of slot %d overlaps x of slot %d",
                              i, j);
            }
        }
    return CLG_OK;
}

// out_i = a_i * H in_i + br * x_i for every slot, in one launch
int batch_hop(clg_context ctx, const HopBatch& B, bool dag, double br, int tile) {
    CLG_TRY(check_batch(B));
    const clg_field_s* in0 = B.in[0];
    const clg_field_s* out0 = B.out[0];
    const bool hasx = B.x[0] != nullptr;
    for (int i = 0; i < B.n; ++i) CLG_CHECK_ARG((B.x[i] != nullptr) == hasx, "x must be given for all slots or none");
    int pbase = 0, nblk = 2;
    if (out0->subset != CLG_FULL) {
        nblk = 1;
        pbase = out0->subset == CLG_EVEN ? 0 : 1;
        if (in0->subset != CLG_FULL)
            CLG_CHECK_ARG(in0->subset != out0->subset, "a half-lattice hop maps one parity to the other");
    }
    // an in-place x term (x_i sharing out_i's storage) accumulates on every repeated tuning launch
    bool aliased = false;
    for (int i = 0; i < B.n; ++i)
        for (int j = 0; j < B.n; ++j)
            if (hasx && fields_overlap(B.x[i], B.out[j])) aliased = true;
    hipStream_t st = ctx->stream;
    const bool f64 = in0->prec == CLG_F64;
    if (tile > 0) {
        CLG_CHECK_ARG(tile == 1 || tile == 2 || (tile == 4 && !f64), "tile must be 1, 2 or 4 (not FP64)");
        return launch_batch_cfg(B, dag, hasx, br, pbase, nblk, BCfg{tile, 128, 8}, st);
    }
    std::vector<long> key = {in0->prec, B.gauge[0]->prec, B.gauge[0]->kind, dag, hasx, nblk, pbase,
                             in0->subset, out0->subset};
    {
        // the distinct lattice shapes of the batch (not their multiplicities, so the subsets that
        // remain while systems converge reuse the choice)
        std::vector<std::vector<long>> shapes;
        for (int i = 0; i < B.n; ++i)
            shapes.push_back({(long)B.in[i]->dims[0], (long)B.in[i]->dims[1], (long)B.in[i]->dims[2],
                              (long)B.in[i]->dims[3]});
        std::sort(shapes.begin(), shapes.end());
        shapes.erase(std::unique(shapes.begin(), shapes.end()), shapes.end());
        for (auto& s : shapes) key.insert(key.end(), s.begin(), s.end());
    }
    BCfg cfg{2, 128, 8};
    bool need = false;
    const std::string cpath = tune_cache_path();
    {
        std::lock_guard<std::mutex> lk(btune_mutex());
        auto it = btune().find(key);
        if (it == btune().end() && !cpath.empty() && cpath != btune_loaded()) {
            std::vector<std::pair<std::vector<long>, std::vector<long>>> ent;
            tune_cache_read(cpath, "wilson_hop_batch", ent);
            for (const auto& e : ent) {
                const std::vector<long>& c = e.second;
                if (e.first.size() < 13 || (e.first.size() - 9) % 4 || c.size() != 4) continue;
                const bool okc = (c[0] == 1 || c[0] == 2 || (c[0] == 4 && e.first[0] != CLG_F64)) &&
                                 (c[1] == 64 || c[1] == 128 || c[1] == 256) && (c[2] == 0 || c[2] == 8) &&
                                 (c[3] == 0 || c[3] == 1);
                if (okc) btune()[e.first] = BCfg{(int)c[0], (int)c[1], (int)c[2], (int)c[3]};
            }
            btune_loaded() = cpath;
            it = btune().find(key);
        }
        if (it != btune().end()) cfg = it->second;
        else need = true;
    }
    const bool tuning = tuning_enabled();
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    if (need && tuning && !aliased && cs == hipStreamCaptureStatusNone) {
        std::vector<BCfg> cands;
        for (int nr : {1, 2, 4}) {
            if (nr == 4 && f64) continue;
            for (int blk : {64, 128, 256})
                for (int xcd : {0, 8})
                    for (int nt : {0, 1}) cands.push_back(BCfg{nr, blk, xcd, nt});
        }
        hipEvent_t e0, e1;
        CLG_HIP(hipEventCreate(&e0));
        CLG_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        for (const BCfg& c : cands) {
            CLG_TRY(launch_batch_cfg(B, dag, hasx, br, pbase, nblk, c, st));
            CLG_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 3; ++r) CLG_TRY(launch_batch_cfg(B, dag, hasx, br, pbase, nblk, c, st));
            CLG_HIP(hipEventRecord(e1, st));
            CLG_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CLG_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; cfg = c; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const char* verbose = std::getenv("CLG_TUNE_VERBOSE");
        if (verbose && verbose[0] == '1')
            std::fprintf(stderr, "[clg tune] batched wilson hop: %d systems prec=%d dag=%d x=%d nblk=%d -> tile=%d block=%d xcd=%d nt=%d (%.1f us)\n",
                         B.n, in0->prec, (int)dag, (int)hasx, nblk, cfg.nr, cfg.block, cfg.xcd, cfg.nt, best * 1e3 / 3);
        std::lock_guard<std::mutex> lk(btune_mutex());
        btune()[key] = cfg;
        tune_cache_append(cpath, "wilson_hop_batch", key, {cfg.nr, cfg.block, cfg.xcd, cfg.nt});
    }
    return launch_batch_cfg(B, dag, hasx, br, pbase, nblk, cfg, st);
}

// ================================================================== batched BLAS
// Per-system vector tables in device memory: V[s][0..m] Krylov basis (V[s][m] = w), X[s][0] the
// inner solution, X[s][1] the inner right-hand side.  Vectors of one batch share one packet type.
struct BTab {
    long npk[kMaxSys];
    void* V[kMaxSys][kMaxM + 1];
    void* X[kMaxSys][2];
    // FP16-storage vectors: per-site scales and the sites per vector (one parity block)
    float* Vn[kMaxSys][kMaxM + 1];
    float* Xn[kMaxSys][2];
    unsigned vh[kMaxSys];
};
struct Act {
    int n;
    int idx[kMaxSys];
};

template <int P> struct BP;
template <> struct BP<CLG_F64> {
    typedef double2 T;
    static constexpr int NC = 1;
    __device__ static void ld(const T* p, long i, double* v) {
        const double2 a = p[i];
        v[0] = a.x; v[1] = a.y;
    }
    __device__ static void st(T* p, long i, const double* v) { p[i] = make_double2(v[0], v[1]); }
};
template <> struct BP<CLG_F32> {
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

__device__ __forceinline__ void* vec_of(const BTab* T, int s, int sel) {
    return sel >= 0 ? T->V[s][sel] : T->X[s][-sel - 1];
}

// partials of the per-system |v|^2 (k_bred_finish adds them up)
template <int P>
__global__ __launch_bounds__(kRedBlock) void k_bnorm2(const BTab* T, Act act, int sel, double* partial) {
    const int s = act.idx[blockIdx.y];
    typedef typename BP<P>::T PT;
    const PT* v = (const PT*)vec_of(T, s, sel);
    const long n = T->npk[s];
    double acc[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double a[2 * BP<P>::NC];
        BP<P>::ld(v, i, a);
#pragma unroll
        for (int c = 0; c < 2 * BP<P>::NC; ++c) acc[0] += a[c] * a[c];
    }
    block_partial<1>(acc, partial + (size_t)blockIdx.y * gridDim.x);
}

// dst = src / sqrt(n2[s * nstride]); cond (nullable): systems with cond[s * nstride] == 0 skip
template <int P>
__global__ void k_bscale(const BTab* T, Act act, int dsel, int ssel, const double* n2, int nstride,
                         const double* cond = nullptr) {
    const int s = act.idx[blockIdx.y];
    if (cond && cond[(size_t)s * nstride] == 0.0) return;
    typedef typename BP<P>::T PT;
    const PT* src = (const PT*)vec_of(T, s, ssel);
    PT* dst = (PT*)vec_of(T, s, dsel);
    const double f = 1.0 / sqrt(n2[(size_t)s * nstride]);
    const long n = T->npk[s];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double a[2 * BP<P>::NC];
        BP<P>::ld(src, i, a);
#pragma unroll
        for (int c = 0; c < 2 * BP<P>::NC; ++c) a[c] *= f;
        BP<P>::st(dst, i, a);
    }
}

// partials of V_j^dagger w for j < K (w = V[s][wsel]) and, with WN, of |w|^2 (k_bred_finish adds
// them up).  cond (nullable): systems with cond[s * cstride] == 0 skip the pass entirely.
template <int P, int K, bool WN>
__global__ __launch_bounds__(kRedBlock) void k_bmultidot(const BTab* T, Act act, int wsel, double* partial,
                                                         const double* cond, int cstride) {
    const int s = act.idx[blockIdx.y];
    if (cond && cond[(size_t)s * cstride] == 0.0) return;
    typedef typename BP<P>::T PT;
    constexpr int NC = BP<P>::NC;
    constexpr int NR = 2 * K + (WN ? 1 : 0);
    const PT* w = (const PT*)T->V[s][wsel];
    const long n = T->npk[s];
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double wv[2 * NC];
        BP<P>::ld(w, i, wv);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double vv[2 * NC];
            BP<P>::ld((const PT*)T->V[s][j], i, vv);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                acc[2 * j] += vv[2 * c] * wv[2 * c] + vv[2 * c + 1] * wv[2 * c + 1];
                acc[2 * j + 1] += vv[2 * c] * wv[2 * c + 1] - vv[2 * c + 1] * wv[2 * c];
            }
        }
        if constexpr (WN) {
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) acc[2 * K] += wv[c] * wv[c];
        }
    }
    block_partial<NR>(acc, partial + (size_t)blockIdx.y * gridDim.x * NR);
}

// The fast form of a batched Arnoldi step, per system: with h = V^H w (K values at ws[s*stride + offh])
// and |w|^2 = sc[4] from the projections, |w'|^2 is taken as |w|^2 - |h|^2 (V orthonormal).  When that
// keeps at least half of |w|^2 (no second pass needed), w'/|w'| goes straight to V[s][dsel], sc[1] =
// |w'|^2, sc[8] = 0 (no second pass) and sc[9] = 0 (the subtraction and normalisation passes skip the
// system); otherwise sc[9] = 1 and nothing else is written.  sc = ws + s*stride + offs.
template <int P, int K>
__global__ __launch_bounds__(kRedBlock) void k_bgs_fast(const BTab* T, Act act, int wsel, int dsel, double* ws,
                                                        int stride, int offh, int offs) {
    const int s = act.idx[blockIdx.y];
    typedef typename BP<P>::T PT;
    constexpr int NC = BP<P>::NC;
    double* sc = ws + (size_t)s * stride + offs;
    const double* h = ws + (size_t)s * stride + offh;
    const double w2 = sc[4];
    double b2 = w2;
    double hr[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        hr[k] = h[2 * k];
        hi[k] = h[2 * k + 1];
        b2 -= hr[k] * hr[k] + hi[k] * hi[k];
    }
    const bool fast = b2 >= 0.5 * w2 && b2 > 1e-300;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc[9] = fast ? 0.0 : 1.0;
        if (fast) {
            sc[1] = b2;
            sc[8] = 0.0;
        }
    }
    if (!fast) return;
    const double inv = 1.0 / sqrt(b2);
    const PT* w = (const PT*)T->V[s][wsel];
    PT* d = (PT*)T->V[s][dsel];
    const long n = T->npk[s];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double y[2 * NC];
        BP<P>::ld(w, i, y);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double x[2 * NC];
            BP<P>::ld((const PT*)T->V[s][k], i, x);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                y[2 * c] -= hr[k] * x[2 * c] - hi[k] * x[2 * c + 1];
                y[2 * c + 1] -= hr[k] * x[2 * c + 1] + hi[k] * x[2 * c];
            }
        }
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) y[c] *= inv;
        BP<P>::st(d, i, y);
    }
}

// y = target (V[s][tsel] or X), y -= sgn * sum_j c_j V_j with c at c[s*cstride + 2j]; NORM: partials
// of |y|^2 (k_bred_finish forms |y|^2 and the second-pass flag from them).  cond (nullable): systems
// with cond[s * cstride] == 0 skip the pass.
template <int P, int K, bool NORM>
__global__ __launch_bounds__(kRedBlock) void k_bmultiaxpy(const BTab* T, Act act, int tsel, const double* cc,
                                                          int cstride, double sgn, double* partial,
                                                          const double* cond) {
    const int s = act.idx[blockIdx.y];
    if (cond && cond[(size_t)s * cstride] == 0.0) return;
    typedef typename BP<P>::T PT;
    constexpr int NC = BP<P>::NC;
    PT* y = (PT*)vec_of(T, s, tsel);
    const long n = T->npk[s];
    double cr[K], ci[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        cr[j] = sgn * cc[(size_t)s * cstride + 2 * j];
        ci[j] = sgn * cc[(size_t)s * cstride + 2 * j + 1];
    }
    double acc[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double yv[2 * NC];
        BP<P>::ld(y, i, yv);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double vv[2 * NC];
            BP<P>::ld((const PT*)T->V[s][j], i, vv);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                yv[2 * c] -= cr[j] * vv[2 * c] - ci[j] * vv[2 * c + 1];
                yv[2 * c + 1] -= cr[j] * vv[2 * c + 1] + ci[j] * vv[2 * c];
            }
        }
        BP<P>::st(y, i, yv);
        if constexpr (NORM) {
            // norm of the stored (rounded) values
            double q[2 * NC];
            BP<P>::ld(y, i, q);
#pragma unroll
            for (int c = 0; c < 2 * NC; ++c) acc[0] += q[c] * q[c];
        }
    }
    if constexpr (NORM) block_partial<1>(acc, partial + (size_t)blockIdx.y * gridDim.x);
}

// Second stage of the batched reductions: one workgroup per active system (blockIdx.y) sums its
// nblk partials (nr values each) in a fixed order; totals r < nout go to out[s*stride + r], the one
// after them to out2[s*stride]; flag (nullable) gets total_0 < eta * ref[s*stride].  Systems with
// cond[s*stride] == 0 are skipped, as in the first stage.
__global__ __launch_bounds__(kRedBlock) void k_bred_finish(Act act, const double* __restrict__ partial, int nblk,
                                                           int nr, double* out, int nout, double* out2,
                                                           const double* cond, const double* ref, double eta,
                                                           double* flag, int stride) {
    const int s = act.idx[blockIdx.y];
    if (cond && cond[(size_t)s * stride] == 0.0) return;
    const double* pp = partial + (size_t)blockIdx.y * nblk * nr;
    __shared__ double sm[kRedBlock / 64][2 * kMaxM + 1];
    __shared__ double tot[2 * kMaxM + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int r = 0; r < nr; ++r) {
        double a = 0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += pp[(size_t)b * nr + r];
        a = wave_sum(a);
        if (lane == 0) sm[wv][r] = a;
    }
    __syncthreads();
    if ((int)threadIdx.x < nr) {
        const int r = threadIdx.x;
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w][r];
        tot[r] = t;
        if (r < nout) out[(size_t)s * stride + r] = t;
        else if (out2) out2[(size_t)s * stride] = t;
    }
    __syncthreads();
    if (flag && threadIdx.x == 0) flag[(size_t)s * stride] = tot[0] < eta * ref[(size_t)s * stride] ? 1.0 : 0.0;
}

// y = a y + b x over (V or X) selections, real coefficients
template <int P>
__global__ void k_baxpby(const BTab* T, Act act, int ysel, double a, int xsel, double b) {
    const int s = act.idx[blockIdx.y];
    typedef typename BP<P>::T PT;
    PT* y = (PT*)vec_of(T, s, ysel);
    const PT* x = (const PT*)vec_of(T, s, xsel);
    const long n = T->npk[s];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double yv[2 * BP<P>::NC] = {0}, xv[2 * BP<P>::NC];
        if (a != 0.0) BP<P>::ld(y, i, yv);  // a = 0 overwrites y without reading it
        BP<P>::ld(x, i, xv);
#pragma unroll
        for (int c = 0; c < 2 * BP<P>::NC; ++c) yv[c] = a * yv[c] + b * xv[c];
        BP<P>::st(y, i, yv);
    }
}
template <int P>
__global__ void k_bzero(const BTab* T, Act act, int ysel) {
    const int s = act.idx[blockIdx.y];
    typedef typename BP<P>::T PT;
    PT* y = (PT*)vec_of(T, s, ysel);
    const long n = T->npk[s];
    double z[2 * BP<P>::NC] = {0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        BP<P>::st(y, i, z);
}

// ---- per-system small kernels (one workgroup per system): Hessenberg update and least squares
// Per-system workspace: H (2 ld m), c2 (2m), g (2(m+1)), y (2m), sc (16) doubles, then an int area.
struct WsLayout {
    int m, ld;
    size_t offH, offC, offG, offY, offS, offGS, stride;
    WsLayout(int m_) : m(m_), ld(m_ + 1) {
        offH = 0;
        offC = offH + 2 * (size_t)ld * m;
        offG = offC + 2 * (size_t)m;
        offY = offG + 2 * (size_t)(m + 1);
        offS = offY + 2 * (size_t)m;
        offGS = offS + 16;
        stride = offGS + 40;
    }
};
// sc: [0] |r|^2, [1] |w'|^2, [2] dev, [3] |dev|, [4] |w|^2 before GS, [5] |b|^2, [6] |r_outer|^2,
// [7] |b - D x|^2, [8] second Gram-Schmidt pass flag (1.0 / 0.0), [9] 0.0 when the step took the fast
// form (k_bgs_fast), else 1.0

__global__ void k_bhess_update(double* ws, size_t stride, size_t offH, size_t offC, size_t offS, int ld, int j,
                               Act act) {
    const int s = act.idx[blockIdx.y];
    double* base = ws + (size_t)s * stride;
    double* H = base + offH;
    const double* c2 = base + offC;
    const double* sc = base + offS;
    const bool reorth = sc[8] != 0.0;
    const int i = threadIdx.x;
    if (i <= j) {
        double re = H[2 * (j * ld + i)], im = H[2 * (j * ld + i) + 1];
        if (reorth) {
            re += c2[2 * i];
            im += c2[2 * i + 1];
        }
        re = (i == j ? 1.0 : 0.0) - re;  // Arnoldi on K = 1 - A: column j of A's Hessenberg
        im = -im;
        H[2 * (j * ld + i)] = re;
        H[2 * (j * ld + i) + 1] = im;
    }
    if (i == 0) {
        H[2 * (j * ld + j + 1)] = -sqrt(sc[1]);
        H[2 * (j * ld + j + 1) + 1] = 0.0;
    }
}

// Givens QR of the (m+1) x m Hessenberg with g = |r| e1 and back substitution (RotateH / SolveY,
// CSolverGMRES.cpp:262-358) for each system, FP64
__global__ void k_bgmres_lsq(double* ws, size_t stride, WsLayout L, Act act) {
    if (threadIdx.x != 0) return;
    const int s = act.idx[blockIdx.y];
    double* base = ws + (size_t)s * stride;
    typedef C2<double> c;
    const int ld = L.ld, m = L.m;
    double* H = base + L.offH;
    auto Hp = [&](int r, int col) -> c* { return reinterpret_cast<c*>(H + 2 * (col * ld + r)); };
    c* G = reinterpret_cast<c*>(base + L.offG);
    c* Y = reinterpret_cast<c*>(base + L.offY);
    double* sc = base + L.offS;
    G[0] = c{sqrt(sc[0]), 0.0};
    for (int k = 1; k <= m; ++k) G[k] = c{0.0, 0.0};
    for (int k = 0; k < m; ++k) {
        const c hkk = *Hp(k, k), hk1 = *Hp(k + 1, k);
        const double den = 1.0 / sqrt(hkk.x * hkk.x + hkk.y * hkk.y + hk1.x * hk1.x + hk1.y * hk1.y);
        const c cs = cscale(hkk, den), sn = cscale(hk1, den);
        const c csh = cconj(cs), snh = cconj(sn);
        for (int jj = k; jj < m; ++jj) {
            const c hkj = *Hp(k, jj), h1j = *Hp(k + 1, jj);
            *Hp(k, jj) = cadd(cmul(csh, hkj), cmul(snh, h1j));
            *Hp(k + 1, jj) = csub(cmul(cs, h1j), cmul(sn, hkj));
        }
        const c mg = c{-G[k].x, -G[k].y};
        G[k] = cmul(csh, G[k]);
        G[k + 1] = cmul(sn, mg);
    }
    sc[2] = G[m].x * G[m].x + G[m].y * G[m].y;
    sc[3] = sqrt(sc[2]);
    for (int k = m - 1; k >= 0; --k) {
        c acc = G[k];
        for (int jj = k + 1; jj < m; ++jj) acc = csub(acc, cmul(*Hp(k, jj), Y[jj]));
        const c d = *Hp(k, k);
        const double dd = d.x * d.x + d.y * d.y;
        Y[k] = c{(acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd};
    }
}

// ---- batched BLAS on FP16-storage vectors (one parity block each; see clg_blas16.hip for the
//      format and the arithmetic), one workgroup row per active system (blockIdx.y)
__device__ __forceinline__ uint4* hvec(const BTab* T, int s, int sel) {
    return (uint4*)(sel >= 0 ? T->V[s][sel] : T->X[s][-sel - 1]);
}
__device__ __forceinline__ float* hnrm(const BTab* T, int s, int sel) {
    return sel >= 0 ? T->Vn[s][sel] : T->Xn[s][-sel - 1];
}
#define BSITE_LOOP(i, n) for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

__global__ __launch_bounds__(kRedBlock) void k_bh_norm2(const BTab* T, Act act, int sel, double* partial) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const uint4* c = hvec(T, s, sel);
    const float* sn = hnrm(T, s, sel);
    double acc[1] = {0.0};
    BSITE_LOOP(i, n) {
        float v[24];
        h16_load_site(c, sn, n, i, v);
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < 24; ++q) t += v[q] * v[q];
        acc[0] += t;
    }
    block_partial<1>(acc, partial + (size_t)blockIdx.y * gridDim.x);
}
// dst = src / sqrt(n2[s * stride])
__global__ __launch_bounds__(256) void k_bh_scale(const BTab* T, Act act, int dsel, int ssel, const double* n2, int stride) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const float f = (float)(1.0 / sqrt(n2[(size_t)s * stride]));
    const uint4* sc = hvec(T, s, ssel);
    const float* sn = hnrm(T, s, ssel);
    uint4* dc = hvec(T, s, dsel);
    float* dn = hnrm(T, s, dsel);
    BSITE_LOOP(i, n) {
        float v[24];
        h16_load_site(sc, sn, n, i, v);
#pragma unroll
        for (int q = 0; q < 24; ++q) v[q] *= f;
        h16_store_site(dc, dn, n, i, v);
    }
}
// y = a y + b x (real); a == 0 overwrites y without reading it; b == 0 with a == 0 zeroes y
__global__ __launch_bounds__(256) void k_bh_axpby(const BTab* T, Act act, int ysel, float a, int xsel, float b) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    uint4* yc = hvec(T, s, ysel);
    float* yn = hnrm(T, s, ysel);
    const uint4* xc = hvec(T, s, xsel);
    const float* xn = hnrm(T, s, xsel);
    BSITE_LOOP(i, n) {
        float y[24], x[24];
        if (b != 0.f) {
            h16_load_site(xc, xn, n, i, x);
        } else {
#pragma unroll
            for (int q = 0; q < 24; ++q) x[q] = 0.f;
        }
        if (a != 0.f) {
            h16_load_site(yc, yn, n, i, y);
#pragma unroll
            for (int q = 0; q < 24; ++q) y[q] = a * y[q] + b * x[q];
        } else {
#pragma unroll
            for (int q = 0; q < 24; ++q) y[q] = b * x[q];
        }
        h16_store_site(yc, yn, n, i, y);
    }
}
// partials of V_k^H w (2K) and |w|^2 (one more), w = V[s][wsel]
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_bh_multidot(const BTab* T, Act act, int wsel, double* partial) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const uint4* wc = hvec(T, s, wsel);
    const float* wn = hnrm(T, s, wsel);
    double acc[2 * K + 1];
#pragma unroll
    for (int r = 0; r < 2 * K + 1; ++r) acc[r] = 0.0;
    BSITE_LOOP(i, n) {
        float sv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) sv[k] = T->Vn[s][k][i];
        const float sw = wn[i];
        float re[K], im[K], ww = 0.f;
#pragma unroll
        for (int k = 0; k < K; ++k) re[k] = im[k] = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t ci = h16_chunk(i, n, ch);
            uint4 raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) raw[k] = ((const uint4*)T->V[s][k])[ci];
            float y[8];
            h16_unpack8(wc[ci], sw, y);
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
    block_partial<2 * K + 1>(acc, partial + (size_t)blockIdx.y * gridDim.x * (2 * K + 1));
}
// w' = w - sum_k h_k V_k, stored to vector osel (w itself or the next basis vector), with the
// partials of V_k^H w' and |w'|^2; h at ws[s*stride + offh].  fast_ok (osel is the next basis
// vector): per system, when |w|^2 - |h|^2 keeps half of |w|^2 (no second pass needed), w'/|w'| is
// stored and no partials are taken; the system's workgroup 0 records the choice in its Gram-Schmidt
// block (gs[1] |w'|^2, gs[2] 1/|w'|, gs[3] 2 for the fast form, else 0).
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_bh_gs_project(const BTab* T, Act act, int wsel, int osel, double* ws,
                                                              int stride, int offh, double* partial, int offgs,
                                                              int fast_ok) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const double* h = ws + (size_t)s * stride + offh;
    const uint4* wc = hvec(T, s, wsel);
    const float* wn = hnrm(T, s, wsel);
    uint4* oc = hvec(T, s, osel);
    float* on = hnrm(T, s, osel);
    float hr[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { hr[k] = (float)h[2 * k]; hi[k] = (float)h[2 * k + 1]; }
    if (fast_ok) {
        double* gs = ws + (size_t)s * stride + offgs;
        const double w2 = gs[0];
        double b2 = w2;
#pragma unroll
        for (int k = 0; k < K; ++k) b2 -= h[2 * k] * h[2 * k] + h[2 * k + 1] * h[2 * k + 1];
        const bool fast = b2 >= 0.5 * w2 && b2 > 1e-300;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            gs[3] = fast ? 2.0 : 0.0;
            if (fast) {
                gs[1] = b2;
                gs[2] = 1.0 / sqrt(b2);
            }
        }
        if (fast) {
            const float inv = (float)(1.0 / sqrt(b2));
            BSITE_LOOP(i, n) {
                float sv[K];
#pragma unroll
                for (int k = 0; k < K; ++k) sv[k] = T->Vn[s][k][i];
                const float sw = wn[i];
                float y[24];
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    const size_t ci = h16_chunk(i, n, ch);
                    uint4 raw[K];
#pragma unroll
                    for (int k = 0; k < K; ++k) raw[k] = ((const uint4*)T->V[s][k])[ci];
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
                h16_store_site(oc, on, n, i, y);
            }
            return;
        }
    }
    double acc[2 * K + 1];
#pragma unroll
    for (int r = 0; r < 2 * K + 1; ++r) acc[r] = 0.0;
    BSITE_LOOP(i, n) {
        float sv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) sv[k] = T->Vn[s][k][i];
        const float sw = wn[i];
        float y[24];
        float ww = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t ci = h16_chunk(i, n, ch);
            uint4 raw[K];
#pragma unroll
            for (int k = 0; k < K; ++k) raw[k] = ((const uint4*)T->V[s][k])[ci];
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
                float re = 0.f, im = 0.f;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    re += x[2 * q] * yc[2 * q] + x[2 * q + 1] * yc[2 * q + 1];
                    im += x[2 * q] * yc[2 * q + 1] - x[2 * q + 1] * yc[2 * q];
                }
                acc[2 * k] += re;
                acc[2 * k + 1] += im;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) ww += yc[q] * yc[q];
        }
        h16_store_site(oc, on, n, i, y);
        acc[2 * K] += ww;
    }
    block_partial<2 * K + 1>(acc, partial + (size_t)blockIdx.y * gridDim.x * (2 * K + 1));
}
// per system: completes k_bh_gs_project and updates Hessenberg column j of the Arnoldi process on
// K = 1 - A (the batched inner solve): H(i, j) = delta_ij - (h_i + flag c2_i), H(j+1, j) = -|w''|.
// gs (per system, 36 doubles): [0] |w|^2 (in), [1] |w''|^2, [2] 1/|w''|, [3] flag, [4..) c2.
template <int K>
__global__ __launch_bounds__(kRedBlock) void k_bgs_finish(Act act, const double* __restrict__ partial, int nblk,
                                                           double* ws, int stride, int offh, int offgs, int fast_ok) {
    constexpr int NR = 2 * K + 1;
    const int s = act.idx[blockIdx.y];
    const double* pp = partial + (size_t)blockIdx.y * nblk * NR;
    double* hcol = ws + (size_t)s * stride + offh;
    double* gs = ws + (size_t)s * stride + offgs;
    const int j = K - 1;
    if (fast_ok && gs[3] == 2.0) {
        // the projection pass took the fast form: H(i, j) = delta_ij - h_i, H(j+1, j) = -|w'|
        const int i = threadIdx.x;
        if (i < K) {
            hcol[2 * i] = (i == j ? 1.0 : 0.0) - hcol[2 * i];
            hcol[2 * i + 1] = -hcol[2 * i + 1];
        }
        if (i == 0) {
            hcol[2 * (j + 1)] = -sqrt(gs[1]);
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
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w][threadIdx.x];
        tot[threadIdx.x] = t;
    }
    __syncthreads();
    const double w2 = tot[2 * K];
    const bool flag = w2 < 0.5 * gs[0];
    const int i = threadIdx.x;
    if (i < K) {
        double re = hcol[2 * i], im = hcol[2 * i + 1];
        if (flag) {
            re += tot[2 * i];
            im += tot[2 * i + 1];
        }
        hcol[2 * i] = (i == j ? 1.0 : 0.0) - re;
        hcol[2 * i + 1] = -im;
        gs[4 + 2 * i] = tot[2 * i];
        gs[4 + 2 * i + 1] = tot[2 * i + 1];
    }
    if (i == 0) {
        double c2 = 0;
        for (int k = 0; k < 2 * K; ++k) c2 += tot[k] * tot[k];
        double f2 = flag ? w2 - c2 : w2;
        if (!(f2 > 1e-300)) f2 = 1e-300;
        const double f = sqrt(f2);
        hcol[2 * (j + 1)] = -f;
        hcol[2 * (j + 1) + 1] = 0.0;
        gs[1] = f2;
        gs[2] = 1.0 / f;
        gs[3] = flag ? 1.0 : 0.0;
    }
}
// in place on V[s][dsel], which holds w': (w' - flag sum_k c2_k V_k) / |w''|.  Without the second
// pass only the per-site scales change: the stored halves of w'/|w'| are those of w'.
template <int K>
__global__ __launch_bounds__(256) void k_bh_gs_finalize(const BTab* T, Act act, int dsel, const double* ws, int stride,
                                                        int offgs) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const double* gs = ws + (size_t)s * stride + offgs;
    if (gs[3] == 2.0) return;  // the projection pass stored w'/|w'| (fast form)
    const bool flag = gs[3] != 0.0;
    const float inv = (float)gs[2];
    uint4* dc = hvec(T, s, dsel);
    float* dn = hnrm(T, s, dsel);
    if (!flag) {
        BSITE_LOOP(i, n) dn[i] *= inv;
        return;
    }
    float cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = (float)gs[4 + 2 * k]; ci[k] = (float)gs[4 + 2 * k + 1]; }
    BSITE_LOOP(i, n) {
        float y[24];
        h16_load_site(dc, dn, n, i, y);
        if (flag) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const size_t cidx = h16_chunk(i, n, ch);
                float* yc = y + 8 * ch;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    float x[8];
                    h16_unpack8(((const uint4*)T->V[s][k])[cidx], T->Vn[s][k][i], x);
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
        h16_store_site(dc, dn, n, i, y);
    }
}
// y += sum_k c_k V_k, y = V[s][tsel] or X, c at ws[s*stride + offc]
template <int K>
__global__ __launch_bounds__(256) void k_bh_multiaxpy(const BTab* T, Act act, int tsel, const double* ws, int stride, int offc) {
    const int s = act.idx[blockIdx.y];
    const unsigned n = T->vh[s];
    const double* c = ws + (size_t)s * stride + offc;
    float cr[K], ci[K];
#pragma unroll
    for (int k = 0; k < K; ++k) { cr[k] = (float)c[2 * k]; ci[k] = (float)c[2 * k + 1]; }
    uint4* yc = hvec(T, s, tsel);
    float* yn = hnrm(T, s, tsel);
    BSITE_LOOP(i, n) {
        float y[24];
        h16_load_site(yc, yn, n, i, y);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const size_t cidx = h16_chunk(i, n, ch);
            float* yp = y + 8 * ch;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                float x[8];
                h16_unpack8(((const uint4*)T->V[s][k])[cidx], T->Vn[s][k][i], x);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    yp[2 * q] += cr[k] * x[2 * q] - ci[k] * x[2 * q + 1];
                    yp[2 * q + 1] += cr[k] * x[2 * q + 1] + ci[k] * x[2 * q];
                }
            }
        }
        h16_store_site(yc, yn, n, i, y);
    }
}

// out[i * cnt + c] = ws[s_i * stride + off + c] for the systems s_i of act (one read-back per step)
__global__ void k_bgather(const double* ws, int stride, int off, int cnt, Act act, double* out) {
    const int i = blockIdx.x, c = threadIdx.x;
    if (i < act.n && c < cnt) out[(size_t)i * cnt + c] = ws[(size_t)act.idx[i] * stride + off + c];
}

unsigned red_grid_b(clg_context ctx, long n, int nact) {
    long g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    const long cap = std::max<long>(8, (long)ctx->num_cu * 4 / std::max(1, nact));
    return (unsigned)std::max<long>(1, std::min(g, std::min<long>(cap, 2048)));
}
unsigned ew_grid_b(clg_context ctx, long n, int nact) {
    long g = (n + 255) / 256;
    const long cap = std::max<long>(16, (long)ctx->num_cu * 16 / std::max(1, nact));
    return (unsigned)std::max<long>(1, std::min(g, cap));
}

// ================================================================== the batched solver
struct BatchSolver {
    clg_context ctx;
    const clg_solver_params* p;
    int ns;
    double kappa;
    bool dag;
    int vprec;  // inner (Krylov) precision: CLG_F64 or CLG_F32
    int m;
    long max_npk = 0;
    // per system
    clg_field_s *x[kMaxSys], *b[kMaxSys], *gauge[kMaxSys];
    clg_field_s *xe[kMaxSys], *xo[kMaxSys], *be[kMaxSys], *bo[kMaxSys];
    clg_field_s *bhat[kMaxSys], *r64[kMaxSys], *t64[kMaxSys];  // FP64 EVEN, EVEN, ODD
    clg_field_s *V[kMaxSys][kMaxM + 1], *elo[kMaxSys], *rlo[kMaxSys], *tlo[kMaxSys];
    clg_field_s* glo[kMaxSys];  // inner links per system (shared between systems of one gauge field)
    double bnorm[kMaxSys];
    BTab* dtab = nullptr;  // device table (inside the context workspace)
    double* ws = nullptr;  // per-system FP64 workspace
    WsLayout L{1};
    int cycles[kMaxSys], outer[kMaxSys];
    bool conv[kMaxSys];
    std::vector<clg_field_s*> views;

    ~BatchSolver() {
        for (clg_field_s* v : views) clg_field_destroy(v);
    }

    int slot(int s, int k) const { return 20000 + s * 64 + k; }

    int setup(clg_field_s* const* xs, clg_field_s* const* bs, clg_field_s* const* gs) {
        m = p->max_dim;
        L = WsLayout(m);
        vprec = p->inner_prec == CLG_F64 ? CLG_F64 : p->inner_prec == CLG_INNER_F16_VECTORS ? CLG_F16 : CLG_F32;
        // inner links, one conversion per distinct gauge field
        std::vector<std::pair<const clg_field_s*, clg_field_s*>> conv_g;
        for (int s = 0; s < ns; ++s) {
            x[s] = xs[s];
            b[s] = bs[s];
            gauge[s] = gs[s];
            CLG_TRY(clg_field_parity_view(x[s], CLG_EVEN, &xe[s]));
            views.push_back(xe[s]);
            CLG_TRY(clg_field_parity_view(x[s], CLG_ODD, &xo[s]));
            views.push_back(xo[s]);
            CLG_TRY(clg_field_parity_view(b[s], CLG_EVEN, &be[s]));
            views.push_back(be[s]);
            CLG_TRY(clg_field_parity_view(b[s], CLG_ODD, &bo[s]));
            views.push_back(bo[s]);
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, CLG_F64, CLG_EVEN, slot(s, 60), &bhat[s]));
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, CLG_F64, CLG_EVEN, slot(s, 61), &r64[s]));
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, CLG_F64, CLG_ODD, slot(s, 62), &t64[s]));
            for (int k = 0; k <= m; ++k)
                CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, vprec, CLG_EVEN, slot(s, k), &V[s][k]));
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, vprec, CLG_EVEN, slot(s, 50), &elo[s]));
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, vprec, CLG_EVEN, slot(s, 51), &rlo[s]));
            CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, vprec, CLG_ODD, slot(s, 52), &tlo[s]));
            glo[s] = nullptr;
            if (vprec == CLG_F64) {
                glo[s] = gauge[s];
            } else {
                for (auto& pr : conv_g)
                    if (pr.first == gauge[s]) glo[s] = pr.second;
                if (!glo[s]) {
                    const int lp = p->inner_prec == CLG_F32 ? CLG_F32 : CLG_F16;
                    const int gk = lp == CLG_F16 ? CLG_GAUGE12 : CLG_GAUGE;
                    clg_field_s* g = nullptr;
                    CLG_TRY(scratch_field(ctx, gauge[s], gk, lp, CLG_FULL, 20000 + (int)conv_g.size(), &g));
                    CLG_TRY(gauge_convert(ctx, g, gauge[s]));
                    conv_g.push_back({gauge[s], g});
                    glo[s] = g;
                }
            }
        }
        // device tables: BTab then the per-system workspace
        const size_t tab_doubles = (sizeof(BTab) + sizeof(double) - 1) / sizeof(double);
        double* w = nullptr;
        CLG_TRY(workspace(ctx, tab_doubles + L.stride * kMaxSys + 256, &w));  // + read-back staging
        dtab = reinterpret_cast<BTab*>(w);
        ws = w + tab_doubles;
        BTab h;
        std::memset(&h, 0, sizeof(h));
        for (int s = 0; s < ns; ++s) {
            const long nc = 12L * V[s][0]->Vh;
            h.npk[s] = vprec == CLG_F64 ? nc : nc / 2;
            max_npk = std::max(max_npk, h.npk[s]);
            for (int k = 0; k <= m; ++k) {
                h.V[s][k] = V[s][k]->data;
                h.Vn[s][k] = V[s][k]->norm;
            }
            h.X[s][0] = elo[s]->data;
            h.X[s][1] = rlo[s]->data;
            h.Xn[s][0] = elo[s]->norm;
            h.Xn[s][1] = rlo[s]->norm;
            h.vh[s] = (unsigned)V[s][0]->Vh;
        }
        CLG_HIP(hipMemcpyAsync(dtab, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
        CLG_HIP(hipStreamSynchronize(ctx->stream));  // h is a stack object
        CLG_TRY(ensure_reduction_scratch(ctx, (size_t)2048 * kMaxSys * (2 * kMaxM + 3)));
        return CLG_OK;
    }

    // ---- batched operator pieces
    // out_s = a * H in_s (+ br x_s) for the systems in act
    int hop(const std::vector<int>& act, clg_field_s* const* out, clg_field_s* const* in, clg_field_s* const* xx,
            double a, double br, bool lo) {
        HopBatch B;
        B.n = (int)act.size();
        for (int i = 0; i < B.n; ++i) {
            const int s = act[i];
            B.gauge[i] = lo ? glo[s] : gauge[s];
            B.in[i] = in[s];
            B.out[i] = out[s];
            B.x[i] = xx ? xx[s] : nullptr;
            B.a[i] = a;
        }
        return batch_hop(ctx, B, dag, br, 0);
    }
    Act make_act(const std::vector<int>& act) const {
        Act a;
        std::memset(&a, 0, sizeof(a));
        a.n = (int)act.size();
        for (int i = 0; i < a.n; ++i) a.idx[i] = act[i];
        return a;
    }
    double* sc(int s) const { return ws + (size_t)s * L.stride + L.offS; }
    // scalars sc(s)[off .. off+cnt) of the systems in act, one device-to-host copy
    int fetch_sc(const std::vector<int>& act, int off, int cnt, std::vector<double>& out) {
        Act A = make_act(act);
        CLG_CHECK_ARG(cnt >= 1 && cnt <= 16, "bad read-back size");
        double* stage = ws + L.stride * kMaxSys;  // staging after the per-system blocks (256 doubles)
        k_bgather<<<A.n, 64, 0, ctx->stream>>>(ws, (int)L.stride, (int)L.offS + off, cnt, A, stage);
        CLG_HIP(hipGetLastError());
        out.resize((size_t)A.n * cnt);
        return fetch_doubles(ctx, stage, A.n * cnt, out.data());
    }

    template <int P, int K>
    int multidot(const Act& A, double* out_base, size_t ostride_off, double* wn, const double* cond) {
        const unsigned g = red_grid_b(ctx, max_npk, A.n);
        dim3 grid(g, A.n);
        if (wn)
            k_bmultidot<P, K, true><<<grid, kRedBlock, 0, ctx->stream>>>(dtab, A, m, ctx->red_partial, cond,
                                                                         (int)L.stride);
        else
            k_bmultidot<P, K, false><<<grid, kRedBlock, 0, ctx->stream>>>(dtab, A, m, ctx->red_partial, cond,
                                                                          (int)L.stride);
        CLG_HIP(hipGetLastError());
        k_bred_finish<<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)g, 2 * K + (wn ? 1 : 0),
                                                                   out_base + ostride_off, 2 * K, wn, cond, nullptr,
                                                                   0.0, nullptr, (int)L.stride);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    template <int P, int K>
    int multiaxpy(const Act& A, int tsel, const double* c, double sgn, double* nrm, const double* cond,
                  const double* ref, double eta, double* flag) {
        const unsigned g = red_grid_b(ctx, max_npk, A.n);
        dim3 grid(g, A.n);
        if (nrm)
            k_bmultiaxpy<P, K, true><<<grid, kRedBlock, 0, ctx->stream>>>(dtab, A, tsel, c, (int)L.stride, sgn,
                                                                          ctx->red_partial, cond);
        else
            k_bmultiaxpy<P, K, false><<<grid, kRedBlock, 0, ctx->stream>>>(dtab, A, tsel, c, (int)L.stride, sgn,
                                                                           ctx->red_partial, cond);
        CLG_HIP(hipGetLastError());
        if (nrm) {
            k_bred_finish<<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)g, 1, nrm, 1, nullptr,
                                                                       cond, ref, eta, flag, (int)L.stride);
            CLG_HIP(hipGetLastError());
        }
        return CLG_OK;
    }
    template <int P, int K>
    int gs_fast(const Act& A, int dsel, size_t offh) {
        const unsigned g = red_grid_b(ctx, max_npk, A.n);
        k_bgs_fast<P, K><<<dim3(g, A.n), kRedBlock, 0, ctx->stream>>>(dtab, A, m, dsel, ws, (int)L.stride, (int)offh,
                                                                      (int)L.offS);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    template <int P>
    int gs_fast_k(int k, const Act& A, int dsel, size_t offh) {
        switch (k) {
#define GFK(N) case N: return gs_fast<P, N>(A, dsel, offh);
            GFK(1) GFK(2) GFK(3) GFK(4) GFK(5) GFK(6) GFK(7) GFK(8) GFK(9) GFK(10) GFK(11) GFK(12) GFK(13) GFK(14)
            GFK(15) GFK(16)
#undef GFK
        }
        return CLG_BAD_ARG;
    }
    // dispatch on the runtime vector count k (1..kMaxM)
    template <int P>
    int multidot_k(int k, const Act& A, double* base, size_t off, double* wn, const double* cond) {
        switch (k) {
#define MDK(N) case N: return multidot<P, N>(A, base, off, wn, cond);
            MDK(1) MDK(2) MDK(3) MDK(4) MDK(5) MDK(6) MDK(7) MDK(8) MDK(9) MDK(10) MDK(11) MDK(12) MDK(13) MDK(14)
            MDK(15) MDK(16)
#undef MDK
        }
        return CLG_BAD_ARG;
    }
    template <int P>
    int multiaxpy_k(int k, const Act& A, int tsel, const double* c, double sgn, double* nrm, const double* cond,
                    const double* ref, double eta, double* flag) {
        switch (k) {
#define MAK(N) case N: return multiaxpy<P, N>(A, tsel, c, sgn, nrm, cond, ref, eta, flag);
            MAK(1) MAK(2) MAK(3) MAK(4) MAK(5) MAK(6) MAK(7) MAK(8) MAK(9) MAK(10) MAK(11) MAK(12) MAK(13) MAK(14)
            MAK(15) MAK(16)
#undef MAK
        }
        return CLG_BAD_ARG;
    }

    // ---- inner batched GMRES(m) on Mhat e = r (precision P), x0 = 0, stop |g_m| < tol[s] per system
    template <int P>
    int inner(std::vector<int> act, const double* tol) {
        const int maxc = p->restart;
        const double k2 = kappa * kappa;
        std::vector<clg_field_s*> Vj(kMaxSys), W(kMaxSys);
        for (int cyc = 0; cyc < maxc && !act.empty(); ++cyc) {
            Act A = make_act(act);
            const unsigned eg = ew_grid_b(ctx, max_npk, A.n);
            dim3 egrid(eg, A.n);
            const unsigned rg = red_grid_b(ctx, max_npk, A.n);
            dim3 rgrid(rg, A.n);
            // V0 = r - Mhat e (cycle 0: e = 0, V0 = r)
            if (cyc == 0) {
                k_bzero<P><<<egrid, 256, 0, ctx->stream>>>(dtab, A, -1);
                k_baxpby<P><<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 0.0, -2, 1.0);
            } else {
                // V0 = kappa^2 H_eo H_oe e - e  (= -Mhat e), then V0 += r
                for (int s : act) { Vj[s] = V[s][0]; W[s] = elo[s]; }
                CLG_TRY(hop(act, tlo, elo, nullptr, 1.0, 0.0, true));
                CLG_TRY(hop(act, Vj.data(), tlo, elo, k2, -1.0, true));
                k_baxpby<P><<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 1.0, -2, 1.0);
            }
            CLG_HIP(hipGetLastError());
            k_bnorm2<P><<<rgrid, kRedBlock, 0, ctx->stream>>>(dtab, A, 0, ctx->red_partial);
            k_bred_finish<<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)rg, 1, ws + L.offS, 1,
                                                                       nullptr, nullptr, nullptr, 0.0, nullptr,
                                                                       (int)L.stride);
            k_bscale<P><<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 0, ws + L.offS, (int)L.stride);
            CLG_HIP(hipGetLastError());
            for (int s : act) cycles[s]++;
            for (int j = 0; j < m; ++j) {
                // w = K V_j = kappa^2 H_eo H_oe V_j
                for (int s : act) { Vj[s] = V[s][j]; W[s] = V[s][m]; }
                CLG_TRY(hop(act, tlo, Vj.data(), nullptr, 1.0, 0.0, true));
                CLG_TRY(hop(act, W.data(), tlo, nullptr, k2, 0.0, true));
                double* Hj0 = ws + L.offH + 2 * (size_t)j * L.ld;
                double* reorth = ws + L.offS + 8;
                // systems whose step takes the fast form (sc[9] == 0) have V_{j+1} written after the
                // projections and skip the subtraction and normalisation passes; the last step keeps
                // w' in place (no next vector) and runs them for every system
                const double* regular = j < m - 1 ? ws + L.offS + 9 : nullptr;
                CLG_TRY(multidot_k<P>(j + 1, A, ws, L.offH + 2 * (size_t)j * L.ld, ws + L.offS + 4, nullptr));
                if (j < m - 1) CLG_TRY(gs_fast_k<P>(j + 1, A, j + 1, L.offH + 2 * (size_t)j * L.ld));
                CLG_TRY(multiaxpy_k<P>(j + 1, A, m, Hj0, 1.0, ws + L.offS + 1, regular, ws + L.offS + 4, 0.5, reorth));
                CLG_TRY(multidot_k<P>(j + 1, A, ws, L.offC, nullptr, reorth));
                CLG_TRY(multiaxpy_k<P>(j + 1, A, m, ws + L.offC, 1.0, ws + L.offS + 1, reorth, nullptr, 0.0, nullptr));
                k_bhess_update<<<dim3(1, A.n), 64, 0, ctx->stream>>>(ws, L.stride, L.offH, L.offC, L.offS, L.ld, j, A);
                if (j < m - 1)
                    k_bscale<P><<<egrid, 256, 0, ctx->stream>>>(dtab, A, j + 1, m, ws + L.offS + 1, (int)L.stride,
                                                                regular);
                CLG_HIP(hipGetLastError());
            }
            k_bgmres_lsq<<<dim3(1, A.n), 64, 0, ctx->stream>>>(ws, L.stride, L, A);
            CLG_HIP(hipGetLastError());
            // e += V y
            CLG_TRY(multiaxpy_k<P>(m, A, -1, ws + L.offY, -1.0, nullptr, nullptr, nullptr, 0.0, nullptr));
            // per-system deviation |g_m|
            std::vector<double> dev(ns * 2, 0.0);
            CLG_HIP(hipStreamSynchronize(ctx->stream));
            std::vector<int> next;
            std::vector<double> dv;
            CLG_TRY(fetch_sc(act, 2, 2, dv));
            for (size_t q = 0; q < act.size(); ++q) {
                const int s = act[q];
                if (!(dv[2 * q + 1] < tol[s]) && std::isfinite(dv[2 * q])) next.push_back(s);
            }
            act.swap(next);
        }
        return CLG_OK;
    }

    // ---- inner batched GMRES(m) with FP16-storage vectors: the fused Gram-Schmidt step (projections;
    //      subtraction with the second-pass projections; normalisation) per Arnoldi step
    int inner16(std::vector<int> act, const double* tol) {
        const int maxc = p->restart;
        const double k2 = kappa * kappa;
        std::vector<clg_field_s*> Vj(kMaxSys), W(kMaxSys);
        const int stride = (int)L.stride;
        for (int cyc = 0; cyc < maxc && !act.empty(); ++cyc) {
            Act A = make_act(act);
            const unsigned eg = ew_grid_b(ctx, max_npk, A.n);
            dim3 egrid(eg, A.n);
            const unsigned rg = red_grid_b(ctx, max_npk, A.n);
            dim3 rgrid(rg, A.n);
            if (cyc == 0) {
                k_bh_axpby<<<egrid, 256, 0, ctx->stream>>>(dtab, A, -1, 0.f, -1, 0.f);  // e = 0
                k_bh_axpby<<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 0.f, -2, 1.f);   // V0 = r
            } else {
                // V0 = kappa^2 H_eo H_oe e - e (= -Mhat e), then V0 += r
                for (int s : act) Vj[s] = V[s][0];
                CLG_TRY(hop(act, tlo, elo, nullptr, 1.0, 0.0, true));
                CLG_TRY(hop(act, Vj.data(), tlo, elo, k2, -1.0, true));
                k_bh_axpby<<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 1.f, -2, 1.f);
            }
            CLG_HIP(hipGetLastError());
            k_bh_norm2<<<rgrid, kRedBlock, 0, ctx->stream>>>(dtab, A, 0, ctx->red_partial);
            k_bred_finish<<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)rg, 1, ws + L.offS, 1,
                                                                       nullptr, nullptr, nullptr, 0.0, nullptr, stride);
            k_bh_scale<<<egrid, 256, 0, ctx->stream>>>(dtab, A, 0, 0, ws + L.offS, stride);
            CLG_HIP(hipGetLastError());
            for (int s : act) cycles[s]++;
            for (int j = 0; j < m; ++j) {
                // w = K V_j = kappa^2 H_eo H_oe V_j
                for (int s : act) { Vj[s] = V[s][j]; W[s] = V[s][m]; }
                CLG_TRY(hop(act, tlo, Vj.data(), nullptr, 1.0, 0.0, true));
                CLG_TRY(hop(act, W.data(), tlo, nullptr, k2, 0.0, true));
                const int offh = (int)(L.offH + 2 * (size_t)j * L.ld);
                const int nv = j + 1;
#define CALL(KK)                                                                                                 \
    k_bh_multidot<KK><<<rgrid, kRedBlock, 0, ctx->stream>>>(dtab, A, m, ctx->red_partial);                       \
    k_bred_finish<<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)rg, 2 * KK + 1,         \
                                                               ws + offh, 2 * KK, ws + L.offGS, nullptr, nullptr, \
                                                               0.0, nullptr, stride);                            \
    k_bh_gs_project<KK><<<rgrid, kRedBlock, 0, ctx->stream>>>(dtab, A, m, j < m - 1 ? j + 1 : m, ws, stride, offh, \
                                                              ctx->red_partial, (int)L.offGS, j < m - 1 ? 1 : 0); \
    k_bgs_finish<KK><<<dim3(1, A.n), kRedBlock, 0, ctx->stream>>>(A, ctx->red_partial, (int)rg, ws, stride, offh, \
                                                                  (int)L.offGS, j < m - 1 ? 1 : 0);              \
    if (j < m - 1)                                                                                               \
        k_bh_gs_finalize<KK><<<egrid, 256, 0, ctx->stream>>>(dtab, A, j + 1, ws, stride, (int)L.offGS);
                switch (nv) {
                    case 1: CALL(1) break; case 2: CALL(2) break; case 3: CALL(3) break; case 4: CALL(4) break;
                    case 5: CALL(5) break; case 6: CALL(6) break; case 7: CALL(7) break; case 8: CALL(8) break;
                    case 9: CALL(9) break; case 10: CALL(10) break; case 11: CALL(11) break; case 12: CALL(12) break;
                    case 13: CALL(13) break; case 14: CALL(14) break; case 15: CALL(15) break; default: CALL(16) break;
                }
#undef CALL
                CLG_HIP(hipGetLastError());
            }
            k_bgmres_lsq<<<dim3(1, A.n), 64, 0, ctx->stream>>>(ws, L.stride, L, A);
            CLG_HIP(hipGetLastError());
            // e += V y
#define CALL(KK) k_bh_multiaxpy<KK><<<egrid, 256, 0, ctx->stream>>>(dtab, A, -1, ws, stride, (int)L.offY);
            switch (m) {
                case 1: CALL(1) break; case 2: CALL(2) break; case 3: CALL(3) break; case 4: CALL(4) break;
                case 5: CALL(5) break; case 6: CALL(6) break; case 7: CALL(7) break; case 8: CALL(8) break;
                case 9: CALL(9) break; case 10: CALL(10) break; case 11: CALL(11) break; case 12: CALL(12) break;
                case 13: CALL(13) break; case 14: CALL(14) break; case 15: CALL(15) break; default: CALL(16) break;
            }
#undef CALL
            CLG_HIP(hipGetLastError());
            std::vector<int> next;
            std::vector<double> dv;
            CLG_TRY(fetch_sc(act, 2, 2, dv));
            for (size_t q = 0; q < act.size(); ++q) {
                const int s = act[q];
                if (!(dv[2 * q + 1] < tol[s]) && std::isfinite(dv[2 * q])) next.push_back(s);
            }
            act.swap(next);
        }
        return CLG_OK;
    }

    // ---- outer FP64 defect correction on Mhat x_e = bhat for all systems, then x_o
    int run(double* true_res) {
        std::vector<int> all;
        for (int s = 0; s < ns; ++s) {
            all.push_back(s);
            cycles[s] = 0;
            outer[s] = 0;
            conv[s] = false;
        }
        // |b|^2 of the full systems
        for (int s = 0; s < ns; ++s) CLG_TRY(blas_norm2_dev(ctx, b[s], sc(s) + 5));
        {
            std::vector<double> v;
            CLG_TRY(fetch_sc(all, 5, 1, v));
            for (int s = 0; s < ns; ++s) bnorm[s] = std::sqrt(v[s]);
        }
        // bhat = b_e + kappa H_eo b_o ; x_e = 0
        CLG_TRY(hop(all, bhat, bo, be, kappa, 1.0, false));
        for (int s = 0; s < ns; ++s) CLG_TRY(blas_axpby(ctx, xe[s], 0.0, 0.0, bhat[s], 0.0, 0.0));
        const int max_outer = p->max_outer > 0 ? p->max_outer : 50;
        const double inner_tol = p->inner_accuracy > 0 ? p->inner_accuracy : 1e-5;
        std::vector<int> act = all;
        std::vector<double> tol(kMaxSys, 0.0);
        for (int it = 0; it < max_outer && !act.empty(); ++it) {
            // r = bhat - Mhat x_e = bhat - x_e + kappa^2 H_eo H_oe x_e
            if (it == 0) {
                for (int s : act) CLG_TRY(blas_axpby(ctx, r64[s], 1.0, 0.0, bhat[s], 0.0, 0.0));
            } else {
                CLG_TRY(hop(act, t64, xe, nullptr, 1.0, 0.0, false));
                CLG_TRY(hop(act, r64, t64, xe, kappa * kappa, -1.0, false));
                for (int s : act) CLG_TRY(blas_axpby(ctx, r64[s], 1.0, 0.0, bhat[s], 1.0, 0.0));
            }
            for (int s : act) CLG_TRY(blas_norm2_dev(ctx, r64[s], sc(s) + 6));
            std::vector<int> next;
            std::vector<double> rv;
            CLG_TRY(fetch_sc(act, 6, 1, rv));
            for (size_t q = 0; q < act.size(); ++q) {
                const int s = act[q];
                const double r2 = rv[q];
                if (std::sqrt(r2) < p->accuracy * bnorm[s]) {
                    conv[s] = true;
                } else if (std::isfinite(r2)) {
                    next.push_back(s);
                    tol[s] = inner_tol * std::sqrt(r2);
                }
            }
            act.swap(next);
            if (act.empty()) break;
            for (int s : act) {
                outer[s]++;
                CLG_TRY(clg_field_copy(rlo[s], r64[s]));
            }
            if (vprec == CLG_F64) CLG_TRY(inner<CLG_F64>(act, tol.data()));
            else if (vprec == CLG_F16) CLG_TRY(inner16(act, tol.data()));
            else CLG_TRY(inner<CLG_F32>(act, tol.data()));
            // x_e += e
            for (int s : act) {
                CLG_TRY(clg_field_copy(r64[s], elo[s]));
                CLG_TRY(blas_axpby(ctx, xe[s], 1.0, 0.0, r64[s], 1.0, 0.0));
            }
        }
        // x_o = b_o + kappa H_oe x_e
        CLG_TRY(hop(all, xo, xe, bo, kappa, 1.0, false));
        // true residual of the full systems: |b - D x| / |b|, D x = x - kappa H x
        std::vector<clg_field_s*> full(kMaxSys);
        for (int s = 0; s < ns; ++s) CLG_TRY(scratch_field(ctx, b[s], CLG_WILSON, CLG_F64, CLG_FULL, slot(s, 63), &full[s]));
        CLG_TRY(hop(all, full.data(), x, x, kappa, -1.0, false));  // kappa H x - x = -D x
        for (int s = 0; s < ns; ++s) {
            CLG_TRY(blas_axpby(ctx, full[s], 1.0, 0.0, b[s], 1.0, 0.0));
            CLG_TRY(blas_norm2_dev(ctx, full[s], sc(s) + 7));
        }
        std::vector<double> r2;
        CLG_TRY(fetch_sc(all, 7, 1, r2));
        for (int s = 0; s < ns; ++s) true_res[s] = bnorm[s] > 0 ? std::sqrt(r2[s]) / bnorm[s] : std::sqrt(r2[s]);
        return CLG_OK;
    }
};

}  // namespace

namespace clg {

int batch_tune_count() {
    std::lock_guard<std::mutex> lk(btune_mutex());
    return (int)btune().size();
}

// key layout: prec, gauge prec, gauge kind, dagger, x, parities, pbase, in subset, out subset, then
// the distinct lattice shapes (4 extents each)
int batch_tune_entry(int index, char* buf, int len) {
    std::lock_guard<std::mutex> lk(btune_mutex());
    CLG_CHECK_ARG(index >= 0 && index < (int)btune().size(), "no tuned entry %d", index);
    auto it = btune().begin();
    std::advance(it, index);
    const std::vector<long>& k = it->first;
    std::string shapes;
    for (size_t i = 9; i + 3 < k.size(); i += 4) {
        char t[64];
        std::snprintf(t, sizeof(t), "%s%ldx%ldx%ldx%ld", shapes.empty() ? "" : ",", k[i], k[i + 1], k[i + 2], k[i + 3]);
        shapes += t;
    }
    std::snprintf(buf, (size_t)len,
                  "wilson_hop_batch prec=%ld gauge_prec=%ld dagger=%ld x=%ld parities=%ld dims=%s -> tile=%d block=%d "
                  "xcd=%d nt=%d",
                  k[0], k[1], k[3], k[4], k[5], shapes.c_str(), it->second.nr, it->second.block, it->second.xcd,
                  it->second.nt);
    return CLG_OK;
}

void batch_tune_reset() {
    std::lock_guard<std::mutex> lk(btune_mutex());
    btune().clear();
    btune_loaded().clear();
}

}  // namespace clg

// ================================================================== C ABI
extern "C" {

int clg_wilson_hop_batch(clg_context ctx, int n, clg_field* out, clg_field* in, clg_field* gauge, int dagger,
                         const double* a, clg_field* x, double b, int tile) {
    CLG_CHECK_ARG(ctx && out && in && gauge && a, "null argument");
    CLG_CHECK_ARG(n >= 1 && n <= kMaxSys, "batch needs 1..%d systems", kMaxSys);
    CLG_CHECK_ARG(tile >= 0, "tile must be >= 0");
    (void)hipSetDevice(ctx->device);
    HopBatch B;
    B.n = n;
    for (int i = 0; i < n; ++i) {
        B.gauge[i] = gauge[i];
        B.in[i] = in[i];
        B.out[i] = out[i];
        B.x[i] = x ? x[i] : nullptr;
        B.a[i] = a[i];
    }
    return batch_hop(ctx, B, dagger != 0, b, tile);
}

int clg_solve_batch(clg_context ctx, const clg_solver_params* params, int n, clg_field* x, clg_field* b,
                    clg_field* gauge, double kappa, int op, clg_solve_report* reports) {
    CLG_CHECK_ARG(ctx && params && x && b && gauge, "null argument");
    CLG_CHECK_ARG(n >= 1 && n <= kMaxSys, "batch needs 1..%d systems", kMaxSys);
    CLG_CHECK_ARG(op == EFO_F_InverseD || op == EFO_F_InverseDdagger, "batched solves take InverseD or InverseDdagger");
    CLG_CHECK_ARG(params->kind == CLG_WILSON, "batched solves are Wilson solves");
    CLG_CHECK_ARG(params->solver == CLG_SOLVER_GMRES, "batched solves use GMRES");
    CLG_CHECK_ARG(params->stop_rule == CLG_STOP_TRUE_RESIDUAL, "batched solves use the true-residual stop rule");
    CLG_CHECK_ARG(params->max_dim >= 1 && params->max_dim <= kMaxM, "MaxDim must be in [1, %d] for batches", kMaxM);
    CLG_CHECK_ARG(params->restart >= 1 && params->accuracy > 0, "bad Restart / Accuracy");
    for (int i = 0; i < n; ++i) {
        CLG_CHECK_ARG(x[i] && b[i] && gauge[i], "null field in system %d", i);
        CLG_CHECK_ARG(x[i]->kind == CLG_WILSON && b[i]->kind == CLG_WILSON, "system %d: Wilson fields", i);
        CLG_CHECK_ARG(x[i]->prec == CLG_F64 && b[i]->prec == CLG_F64 && x[i]->subset == CLG_FULL &&
                          b[i]->subset == CLG_FULL,
                      "system %d: x and b must be FULL FP64 fields", i);
        CLG_CHECK_ARG(gauge[i]->kind == CLG_GAUGE && gauge[i]->prec == CLG_F64, "system %d: FP64 18-real links", i);
        CLG_TRY(check_same_lattice(x[i], gauge[i]));
        CLG_TRY(check_same_lattice(b[i], gauge[i]));
        CLG_CHECK_ARG(x[i] != b[i], "system %d: x must not alias b", i);
        for (int d = 0; d < 4; ++d) CLG_CHECK_ARG(x[i]->bc[d] == b[i]->bc[d], "system %d: x and b bc differ", i);
        for (int k = 0; k < i; ++k) CLG_CHECK_ARG(!fields_overlap(x[k], x[i]), "systems %d and %d share x", k, i);
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k)
            CLG_CHECK_ARG(!fields_overlap(x[i], b[k]), "x of system %d overlaps b of system %d", i, k);
    (void)hipSetDevice(ctx->device);
    const auto t0 = std::chrono::steady_clock::now();
    BatchSolver S;
    S.ctx = ctx;
    S.p = params;
    S.ns = n;
    S.kappa = kappa;
    S.dag = op == EFO_F_InverseDdagger;
    CLG_TRY(S.setup(x, b, gauge));
    double res[kMaxSys];
    CLG_TRY(S.run(res));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    bool all = true;
    for (int i = 0; i < n; ++i) {
        const bool c = res[i] < params->accuracy;
        all = all && c;
        if (reports) {
            clg_solve_report& r = reports[i];
            std::memset(&r, 0, sizeof(r));
            r.converged = c ? 1 : 0;
            r.cycles = S.cycles[i];
            r.outer = S.outer[i];
            r.true_residual = res[i];
            r.seconds = secs;
        }
    }
    return all ? CLG_OK : CLG_NOT_CONVERGED;
}

}  // extern "C"
