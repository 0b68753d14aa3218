// clg_staggered.hip -- staggered (Kogut-Susskind) SU(3) operator.
//
// Reference _kernelDFermionKS (CFieldFermionKSSU3.cu:34-117):
//   (D psi)(x) = 2am psi(x) + sum_mu eta_mu(x) [ U_mu(x) s+ psi(x+mu) - U_mu(x-mu)^dag s- psi(x-mu) ]
//   D^dagger flips the sign of the hopping sum (D(m=0) is anti-Hermitian).
// eta_mu(x) = (-1)^(x_0 + ... + x_{mu-1}) is computed from the coordinates instead of the
// reference's per-site byte table (CIndexSquare.cu:257-267).
//
// The links (576 bytes per site in FP64) outweigh the colour vectors (48 bytes), and every link
// is read twice: forward by x and backward by x+mu.  The launch therefore keeps both users of a
// link in one L2 (XCD-contiguous slabs with the two parities interleaved block by block) and reads
// links through the cache; a small per-shape tuner picks the workgroup size and the schedule.
#include "clg_internal.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>

using namespace clg;
using namespace clgd;

namespace {

struct KSArgs {
    Geo g;
    const char* in;
    const float* inn;
    char* out;
    float* outn;
    const char* gauge;
    int in_full, out_full;
    double two_am;   // mass term coefficient, applied to in(x)
    double hs;       // +1 for D, -1 for D^dagger
    double cr, ci;   // overall coefficient (hop mode: cr = a)
    int pbase, nblk;
    int xcd, pinter;  // work decomposition (clgd::map_site)
    const char* x;    // hop mode 2: out = a H in + two_am * x
    const float* xn;
    int mode;         // 0: D on full fields; 1: out = a H in; 2: out = a H in + b x (half fields)
};

template <int MU>
__device__ __forceinline__ int eta_sign(const Site& s) {
    int sum = 0;
#pragma unroll
    for (int nu = 0; nu < MU; ++nu) sum += s.c[nu];
    return (sum & 1) ? -1 : 1;
}

template <typename R>
__device__ __forceinline__ void pin3(C2<R>* a) {
    asm volatile("" : "+v"(a[0].x), "+v"(a[0].y), "+v"(a[1].x), "+v"(a[1].y), "+v"(a[2].x), "+v"(a[2].y)::"memory");
}

// PM 2 pins the accumulator after each half of a direction (bounded register use, in-order
// loads); PM 0 leaves the schedule to the compiler
template <int MU, int PM, typename R, class SIO, class LIO>
__device__ __forceinline__ void ks_mu(const KSArgs& a, const Site& s, int p, int h, C2<R>* acc) {
    C2<R> v[3], u[9], t[3];
    const int Vh = a.g.Vh;
    const R e = (R)eta_sign<MU>(s);
    {
        const Nbr n = fwd_nbr<MU>(a.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        SIO::load(a.in, a.inn, Vh, pe * SIO::NPL * Vh + n.h, pe * Vh + n.h, v);
        LIO::load(a.gauge, Vh, (unsigned)(p * 4 + MU) * LIO::NPL * Vh + h, u);
        su3_mv(u, v, t);
        const R f = e * (R)n.sign;
#pragma unroll
        for (int c = 0; c < 3; ++c) { acc[c].x += f * t[c].x; acc[c].y += f * t[c].y; }
    }
    if constexpr (PM == 2) pin3(acc);
    {
        const Nbr n = bwd_nbr<MU>(a.g, s, p);
        const unsigned pe = a.in_full ? (unsigned)n.p : 0u;
        SIO::load(a.in, a.inn, Vh, pe * SIO::NPL * Vh + n.h, pe * Vh + n.h, v);
        LIO::load(a.gauge, Vh, (unsigned)(n.p * 4 + MU) * LIO::NPL * Vh + n.h, u);
        su3_mdv(u, v, t);
        const R f = -e * (R)n.sign;
#pragma unroll
        for (int c = 0; c < 3; ++c) { acc[c].x += f * t[c].x; acc[c].y += f * t[c].y; }
    }
    if constexpr (PM == 2) pin3(acc);
}

// MODE 0: out = c (2am in + hs H in) over both parities; MODE 1/2: the half-lattice hop
// out_p = a H_{p,1-p} in (+ b x_p), where H = sum_mu eta_mu [U s+ psi(x+mu) - U^dag s- psi(x-mu)].
// SP: spinor storage, GP: link storage (FP32 spinors may take FP16 links)
template <typename R, int SP, int GP, int RECON, int PM, int MODE>
__global__ __launch_bounds__(256) void k_ks(KSArgs a) {
    typedef KSIO<SP> SIO;
    typedef LinkIO<GP, RECON> LIO;
    int p, h;
    if (!map_site(a.g.Vh, a.nblk, a.pbase, a.xcd, a.pinter, p, h)) return;
    const Site s = site_of(a.g, p, h);
    C2<R> acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = cmk<R>(0, 0);
    ks_mu<0, PM, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<1, PM, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<2, PM, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<3, PM, R, SIO, LIO>(a, s, p, h, acc);
    const unsigned po = a.out_full ? (unsigned)p : 0u;
    const unsigned eo = po * SIO::NPL * a.g.Vh + h, no = po * a.g.Vh + h;
    C2<R> o[3];
    if constexpr (MODE == 0) {
        const unsigned pi = a.in_full ? (unsigned)p : 0u;
        C2<R> self[3];
        SIO::load(a.in, a.inn, a.g.Vh, pi * SIO::NPL * a.g.Vh + h, pi * a.g.Vh + h, self);
        const C2<R> cc = cmk<R>((R)a.cr, (R)a.ci);
        const R m = (R)a.two_am, hs = (R)a.hs;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const C2<R> t = cmk<R>(m * self[c].x + hs * acc[c].x, m * self[c].y + hs * acc[c].y);
            o[c] = cmul(cc, t);
        }
    } else {
        const R ar = (R)a.cr;
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = cmk<R>(ar * acc[c].x, ar * acc[c].y);
        if constexpr (MODE == 2) {
            C2<R> xv[3];
            SIO::load(a.x, a.xn, a.g.Vh, eo, no, xv);
            const R br = (R)a.two_am;
#pragma unroll
            for (int c = 0; c < 3; ++c) { o[c].x += br * xv[c].x; o[c].y += br * xv[c].y; }
        }
    }
    SIO::store(a.out, a.outn, a.g.Vh, eo, no, o);
}

struct KSCfg {
    int block;  // threads per workgroup
    int xcd;    // 0: hardware order; 8: XCD-contiguous slabs, parities interleaved block by block
    int pm;     // 0 or 2 (ks_mu)
};

template <typename R, int SP, int GP, int RECON, int PM>
void ks_launch_pm(const KSArgs& a0, hipStream_t st, const KSCfg& c) {
    KSArgs a = a0;
    a.xcd = c.xcd;
    a.pinter = c.xcd > 1 ? 1 : 0;
    const unsigned grid = sched_grid(a.g.Vh, a.nblk, c.block, a.xcd, a.pinter);
    if (a.mode == 2) k_ks<R, SP, GP, RECON, PM, 2><<<grid, c.block, 0, st>>>(a);
    else if (a.mode == 1) k_ks<R, SP, GP, RECON, PM, 1><<<grid, c.block, 0, st>>>(a);
    else k_ks<R, SP, GP, RECON, PM, 0><<<grid, c.block, 0, st>>>(a);
}

template <typename R, int SP, int GP, int RECON>
void ks_launch_t(const KSArgs& a, hipStream_t st, const KSCfg& c) {
    if (c.pm == 2) ks_launch_pm<R, SP, GP, RECON, 2>(a, st, c);
    else ks_launch_pm<R, SP, GP, RECON, 0>(a, st, c);
}

// spinor precision prec, link precision gprec: equal, or FP32 spinors with FP16 links
int ks_launch_cfg(const KSArgs& a, int prec, int gprec, bool r12, hipStream_t st, const KSCfg& c) {
    if (prec == CLG_F64) {
        if (r12) ks_launch_t<double, CLG_F64, CLG_F64, 12>(a, st, c);
        else ks_launch_t<double, CLG_F64, CLG_F64, 18>(a, st, c);
    } else if (prec == CLG_F32 && gprec == CLG_F16) {
        if (r12) ks_launch_t<float, CLG_F32, CLG_F16, 12>(a, st, c);
        else ks_launch_t<float, CLG_F32, CLG_F16, 18>(a, st, c);
    } else if (prec == CLG_F32) {
        if (r12) ks_launch_t<float, CLG_F32, CLG_F32, 12>(a, st, c);
        else ks_launch_t<float, CLG_F32, CLG_F32, 18>(a, st, c);
    } else {
        if (r12) ks_launch_t<float, CLG_F16, CLG_F16, 12>(a, st, c);
        else ks_launch_t<float, CLG_F16, CLG_F16, 18>(a, st, c);
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// per-shape tuner: the first launch of a (precision, links, lattice) key times each candidate with
// HIP events on the launch stream and keeps the fastest (every candidate writes the same output)
struct KSKey {
    int prec, gprec, r12, mode, L0, L1, L2, L3;
    bool operator<(const KSKey& o) const { return std::memcmp(this, &o, sizeof(KSKey)) < 0; }
};
std::map<KSKey, KSCfg>& ks_table() {
    static std::map<KSKey, KSCfg> t;
    return t;
}
std::mutex& ks_mutex() {
    static std::mutex m;
    return m;
}
std::string& ks_loaded() {  // the tuning-cache file whose entries are in ks_table()
    static std::string p;
    return p;
}

bool ks_valid(const KSCfg& c) {
    return (c.block == 64 || c.block == 128 || c.block == 256) && (c.xcd == 0 || c.xcd == 8) && (c.pm == 0 || c.pm == 2);
}

int ks_launch(const KSArgs& a, int prec, int gprec, bool r12, hipStream_t st) {
    KSCfg cfg{256, 8, 0};
    // CLG_TUNE_FORCE_KS="block,xcd,pin" fixes the launch configuration (testing and profiling)
    if (const char* e = std::getenv("CLG_TUNE_FORCE_KS")) {
        KSCfg f{0, 0, 0};
        if (std::sscanf(e, "%d,%d,%d", &f.block, &f.xcd, &f.pm) == 3 && ks_valid(f))
            return ks_launch_cfg(a, prec, gprec, r12, st, f);
    }
    KSKey k;
    std::memset(&k, 0, sizeof(k));
    k.prec = prec; k.gprec = gprec; k.r12 = r12; k.mode = a.mode;
    k.L0 = a.g.L[0]; k.L1 = a.g.L[1]; k.L2 = a.g.L[2]; k.L3 = a.g.L[3];
    bool need_tune = false;
    const std::string cpath = tune_cache_path();
    {
        std::lock_guard<std::mutex> lk(ks_mutex());
        auto it = ks_table().find(k);
        if (it == ks_table().end() && !cpath.empty() && cpath != ks_loaded()) {
            std::vector<std::pair<std::vector<long>, std::vector<long>>> ent;
            tune_cache_read(cpath, "staggered_d", ent);
            for (const auto& e : ent) {
                const std::vector<long>& v = e.first;
                const std::vector<long>& c = e.second;
                const KSCfg kc{(int)c[0], c.size() > 1 ? (int)c[1] : -1, c.size() > 2 ? (int)c[2] : -1};
                if (v.size() != 8 || c.size() != 3 || !ks_valid(kc)) continue;
                KSKey t;
                std::memset(&t, 0, sizeof(t));
                t.prec = (int)v[0]; t.gprec = (int)v[1]; t.r12 = (int)v[2]; t.mode = (int)v[3];
                t.L0 = (int)v[4]; t.L1 = (int)v[5]; t.L2 = (int)v[6]; t.L3 = (int)v[7];
                ks_table()[t] = kc;
            }
            ks_loaded() = cpath;
            it = ks_table().find(k);
        }
        if (it != ks_table().end()) cfg = it->second;
        else need_tune = true;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    const bool aliased = a.mode == 2 && a.x == a.out;  // repeated candidate launches would compound
    if (need_tune && tuning_enabled() && !aliased && cs == hipStreamCaptureStatusNone && (long)a.nblk * a.g.Vh >= 4096) {
        const KSCfg cands[] = {{64, 0, 2}, {128, 0, 2}, {256, 0, 2}, {64, 8, 2}, {128, 8, 2}, {256, 8, 2},
                               {64, 8, 0}, {128, 8, 0}, {256, 8, 0}};
        hipEvent_t e0, e1;
        CLG_HIP(hipEventCreate(&e0));
        CLG_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        for (const KSCfg& c : cands) {
            CLG_TRY(ks_launch_cfg(a, prec, gprec, r12, st, c));  // warm
            CLG_HIP(hipEventRecord(e0, st));
            for (int r = 0; r < 3; ++r) CLG_TRY(ks_launch_cfg(a, prec, gprec, r12, st, c));
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
            std::fprintf(stderr, "[clg tune] staggered D prec=%d gprec=%d recon=%d mode=%d L=%dx%dx%dx%d -> block=%d xcd=%d pin=%d (%.1f us)\n",
                         prec, gprec, r12 ? 12 : 18, a.mode, a.g.L[0], a.g.L[1], a.g.L[2], a.g.L[3], cfg.block, cfg.xcd,
                         cfg.pm, best * 1e3 / 3);
        std::lock_guard<std::mutex> lk(ks_mutex());
        ks_table()[k] = cfg;
        tune_cache_append(cpath, "staggered_d", {k.prec, k.gprec, k.r12, k.mode, k.L0, k.L1, k.L2, k.L3},
                          {cfg.block, cfg.xcd, cfg.pm});
    }
    return ks_launch_cfg(a, prec, gprec, r12, st, cfg);
}

}  // namespace

namespace clg {

int ks_tune_count() {
    std::lock_guard<std::mutex> lk(ks_mutex());
    return (int)ks_table().size();
}

int ks_tune_entry(int index, char* buf, int len) {
    std::lock_guard<std::mutex> lk(ks_mutex());
    CLG_CHECK_ARG(index >= 0 && index < (int)ks_table().size(), "no tuned entry %d", index);
    auto it = ks_table().begin();
    std::advance(it, index);
    const KSKey& k = it->first;
    std::snprintf(buf, (size_t)len,
                  "staggered_d prec=%d gauge_prec=%d recon=%d mode=%d dims=%dx%dx%dx%d -> block=%d xcd=%d pin=%d",
                  k.prec, k.gprec, k.r12 ? 12 : 18, k.mode, k.L0, k.L1, k.L2, k.L3, it->second.block, it->second.xcd,
                  it->second.pm);
    return CLG_OK;
}

void ks_tune_reset() {
    std::lock_guard<std::mutex> lk(ks_mutex());
    ks_table().clear();
    ks_loaded().clear();
}

int ks_full(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, double two_am, bool dag,
            double cr, double ci) {
    KSArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = make_geo(out);
    a.in = (const char*)in->data;
    a.inn = in->norm;
    a.out = (char*)out->data;
    a.outn = out->norm;
    a.gauge = (const char*)gauge->data;
    a.in_full = 1;
    a.out_full = 1;
    a.two_am = two_am;
    a.hs = dag ? -1.0 : 1.0;
    a.cr = cr;
    a.ci = ci;
    a.pbase = 0;
    a.nblk = 2;
    return ks_launch(a, out->prec, gauge->prec, gauge->kind == CLG_GAUGE12, ctx->stream);
}

// out_p = a H_{p,1-p} in + b x_p on half fields (x nullable)
int ks_hop_half(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, double av,
                const clg_field_s* x, double bv) {
    KSArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = make_geo(out);
    a.in = (const char*)in->data;
    a.inn = in->norm;
    a.out = (char*)out->data;
    a.outn = out->norm;
    a.gauge = (const char*)gauge->data;
    a.in_full = 0;
    a.out_full = 0;
    a.cr = av;
    a.two_am = bv;
    a.pbase = field_parity(out);
    a.nblk = 1;
    a.mode = x ? 2 : 1;
    if (x) {
        a.x = (const char*)x->data;
        a.xn = x->norm;
    }
    return ks_launch(a, out->prec, gauge->prec, gauge->kind == CLG_GAUGE12, ctx->stream);
}

// the parity-p block of D D^dagger = (2am)^2 - H^2: out = (2am)^2 in - H_{p,1-p} H_{1-p,p} in
int ks_schur_once(clg_context ctx, clg_field_s* out, const clg_field_s* in, const clg_field_s* gauge, double two_am,
                  clg_field_s* tmp_other) {
    CLG_TRY(ks_hop_half(ctx, tmp_other, in, gauge, 1.0, nullptr, 0.0));
    return ks_hop_half(ctx, out, tmp_other, gauge, -1.0, in, two_am * two_am);
}

}  // namespace clg

extern "C" {

// links in the spinor precision, or FP16 links under FP32 spinors
static bool ks_gauge_prec_ok(int prec, int gprec) { return gprec == prec || (prec == CLG_F32 && gprec == CLG_F16); }

int clg_ks_apply(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am, int op, int coeff_type,
                 double cre, double cim) {
    CLG_CHECK_ARG(ctx && out && in && gauge, "null argument");
    CLG_CHECK_ARG(out != in && out->data != in->data, "out and in must be distinct fields");
    CLG_CHECK_ARG(out->kind == CLG_KS && in->kind == CLG_KS, "need staggered fields");
    CLG_CHECK_ARG(out->subset == CLG_FULL && in->subset == CLG_FULL, "need FULL fields");
    CLG_CHECK_ARG(out->prec == in->prec && ks_gauge_prec_ok(in->prec, gauge->prec), "precision mismatch");
    CLG_CHECK_ARG(is_gauge(gauge->kind), "need a gauge field");
    CLG_TRY(check_same_lattice(out, in));
    CLG_TRY(check_same_lattice(out, gauge));
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
    if (op == EFO_F_D || op == EFO_F_Ddagger) return ks_full(ctx, out, in, gauge, two_am, op == EFO_F_Ddagger, fr, fi);
    // DD = D(D psi), DD^dagger = D(D^dagger psi) (CFieldFermionKSSU3.cu:1493-1575)
    clg_field_s* tmp = nullptr;
    CLG_TRY(scratch_field(ctx, in, CLG_KS, in->prec, CLG_FULL, 0, &tmp));
    CLG_TRY(ks_full(ctx, tmp, in, gauge, two_am, op == EFO_F_DDdagger, 1.0, 0.0));
    return ks_full(ctx, out, tmp, gauge, two_am, false, fr, fi);
}

int clg_ks_hop(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double a, clg_field x, double b) {
    CLG_CHECK_ARG(ctx && out && in && gauge, "null argument");
    CLG_CHECK_ARG(out->kind == CLG_KS && in->kind == CLG_KS, "need staggered fields");
    CLG_CHECK_ARG(out->subset != CLG_FULL && in->subset != CLG_FULL && out->subset != in->subset,
                  "the hop maps one parity (half field) to the other");
    CLG_CHECK_ARG(out->prec == in->prec && ks_gauge_prec_ok(in->prec, gauge->prec), "precision mismatch");
    CLG_CHECK_ARG(is_gauge(gauge->kind), "need a gauge field");
    CLG_CHECK_ARG(out->data != in->data, "out and in must be distinct");
    CLG_TRY(check_same_lattice(out, in));
    CLG_TRY(check_same_lattice(out, gauge));
    for (int i = 0; i < 4; ++i) CLG_CHECK_ARG((out->dims[i] & 1) == 0, "even-odd operators need even extents");
    if (x) CLG_CHECK_ARG(x->kind == CLG_KS && x->subset == out->subset && x->prec == out->prec, "x must match out");
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < 4; ++i) out->bc[i] = in->bc[i];
    return ks_hop_half(ctx, out, in, gauge, a, x, b);
}

int clg_ks_schur(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am) {
    CLG_CHECK_ARG(ctx && out && in && gauge, "null argument");
    CLG_CHECK_ARG(out->kind == CLG_KS && in->kind == CLG_KS, "need staggered fields");
    CLG_CHECK_ARG(in->subset != CLG_FULL && out->subset == in->subset, "the Schur block acts on one parity");
    CLG_CHECK_ARG(out->prec == in->prec && ks_gauge_prec_ok(in->prec, gauge->prec), "precision mismatch");
    CLG_CHECK_ARG(is_gauge(gauge->kind), "need a gauge field");
    CLG_CHECK_ARG(out->data != in->data, "out and in must be distinct");
    CLG_TRY(check_same_lattice(out, in));
    CLG_TRY(check_same_lattice(out, gauge));
    for (int i = 0; i < 4; ++i) CLG_CHECK_ARG((out->dims[i] & 1) == 0, "even-odd operators need even extents");
    (void)hipSetDevice(ctx->device);
    clg_field_s* tmp = nullptr;
    CLG_TRY(scratch_field(ctx, in, CLG_KS, in->prec, in->subset == CLG_EVEN ? CLG_ODD : CLG_EVEN, 0, &tmp));
    for (int i = 0; i < 4; ++i) { out->bc[i] = in->bc[i]; tmp->bc[i] = in->bc[i]; }
    return ks_schur_once(ctx, out, in, gauge, two_am, tmp);
}

int clg_time_ks(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am, int variant, int reps,
                double* ms_avg) {
    CLG_CHECK_ARG(ctx && ms_avg && reps > 0 && (variant == 0 || variant == 1), "bad arguments");
    (void)hipSetDevice(ctx->device);
    auto one = [&]() -> int {
        if (variant == 1) return clg_ks_schur(ctx, out, in, gauge, two_am);
        return clg_ks_apply(ctx, out, in, gauge, two_am, EFO_F_D, EOCT_None, 1.0, 0.0);
    };
    CLG_TRY(one());  // validates arguments, tunes, warms up
    CLG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps; ++r) CLG_TRY(one());
    CLG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    CLG_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    CLG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_avg = (double)ms / reps;
    return CLG_OK;
}

}  // extern "C"
