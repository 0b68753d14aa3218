// clg_runtime.hip -- context, fields, host<->device layout conversion, synthetic inputs.
#include "clg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace clg {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
const char* last_error() { return g_err.c_str(); }

Geo make_geo(const clg_field_s* f) {
    Geo g;
    for (int i = 0; i < 4; ++i) {
        g.L[i] = f->dims[i];
        g.sgn[i] = f->bc[i];
    }
    g.stride[3] = 1;
    g.stride[2] = g.L[3];
    g.stride[1] = g.L[2] * g.L[3];
    g.stride[0] = g.L[1] * g.L[2] * g.L[3];
    g.Vh = (int)f->Vh;
    return g;
}

int check_same_lattice(const clg_field_s* a, const clg_field_s* b) {
    CLG_CHECK_ARG(a && b, "null field");
    for (int i = 0; i < 4; ++i)
        CLG_CHECK_ARG(a->dims[i] == b->dims[i], "fields live on different lattices");
    return CLG_OK;
}

}  // namespace clg

using namespace clg;
using namespace clgd;

// ============================================================== device kernels
namespace {

// zero fill with 16-byte stores (a kernel rather than a memset node, so that recorded graphs
// replay it like any other launch)
__global__ void k_zero16(uint4* p, long n16) {
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long)gridDim.x * blockDim.x) p[i] = z;
}
__global__ void k_zero4(unsigned* p, long n4) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

// reunitarise(exp(i eps H)) of the Gaussian draws of this link (clgd::random_su3)
__device__ void random_link(unsigned link, double eps, unsigned seed, cd* out) {
    random_su3(link, eps, seed, STREAM_GAUGE, out);
}

// thread -> (parity, h) over the blocks a field stores
struct Blocks {
    int pbase;  // parity of block 0
    int nblk;   // 1 or 2
};
__device__ __forceinline__ bool thread_site(const Geo& g, Blocks b, int& p, int& h) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)b.nblk * g.Vh) return false;
    const int blk = (int)(gid / g.Vh);
    h = (int)(gid - (long)blk * g.Vh);
    p = b.nblk == 2 ? blk : b.pbase;
    return true;
}

// host AoS (reference order) -> device layout
template <int PREC>
__global__ void k_wilson_from_aos(Geo g, Blocks b, bool full, const double* __restrict__ src, char* blk0,
                                  size_t blk_bytes, float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    const long hi = full ? s.lex : h;
    cd v[12];
    for (int k = 0; k < 12; ++k) v[k] = cd{src[hi * 24 + 2 * k], src[hi * 24 + 2 * k + 1]};
    const int ib = b.nblk == 2 ? p : 0;
    WilsonIO<PREC>::store(blk0, norm0, g.Vh, (unsigned)(ib * WilsonIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
}
template <int PREC>
__global__ void k_wilson_to_aos(Geo g, Blocks b, bool full, double* __restrict__ dst, const char* blk0,
                                size_t blk_bytes, const float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    const long hi = full ? s.lex : h;
    cd v[12];
    const int ib = b.nblk == 2 ? p : 0;
    WilsonIO<PREC>::load(blk0, norm0, g.Vh, (unsigned)(ib * WilsonIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
    for (int k = 0; k < 12; ++k) { dst[hi * 24 + 2 * k] = v[k].x; dst[hi * 24 + 2 * k + 1] = v[k].y; }
}
template <int PREC>
__global__ void k_ks_from_aos(Geo g, Blocks b, bool full, const double* __restrict__ src, char* blk0,
                              size_t blk_bytes, float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    const long hi = full ? s.lex : h;
    cd v[3];
    for (int k = 0; k < 3; ++k) v[k] = cd{src[hi * 6 + 2 * k], src[hi * 6 + 2 * k + 1]};
    const int ib = b.nblk == 2 ? p : 0;
    KSIO<PREC>::store(blk0, norm0, g.Vh, (unsigned)(ib * KSIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
}
template <int PREC>
__global__ void k_ks_to_aos(Geo g, Blocks b, bool full, double* __restrict__ dst, const char* blk0,
                            size_t blk_bytes, const float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    const long hi = full ? s.lex : h;
    cd v[3];
    const int ib = b.nblk == 2 ? p : 0;
    KSIO<PREC>::load(blk0, norm0, g.Vh, (unsigned)(ib * KSIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
    for (int k = 0; k < 3; ++k) { dst[hi * 6 + 2 * k] = v[k].x; dst[hi * 6 + 2 * k + 1] = v[k].y; }
}
template <int PREC, int RECON>
__global__ void k_gauge_from_aos(Geo g, const double* __restrict__ src, char* data, size_t blk_bytes) {
    int p, h;
    if (!thread_site(g, Blocks{0, 2}, p, h)) return;
    const Site s = site_of(g, p, h);
    for (int mu = 0; mu < 4; ++mu) {
        cd m[9];
        const double* q = src + ((long)s.lex * 4 + mu) * 18;
        for (int k = 0; k < 9; ++k) m[k] = cd{q[2 * k], q[2 * k + 1]};
        LinkIO<PREC, RECON>::store(data, g.Vh, (unsigned)((p * 4 + mu) * LinkIO<PREC, RECON>::NPL * g.Vh + h), m);
    }
}
template <int PREC, int RECON>
__global__ void k_gauge_to_aos(Geo g, double* __restrict__ dst, const char* data, size_t blk_bytes) {
    int p, h;
    if (!thread_site(g, Blocks{0, 2}, p, h)) return;
    const Site s = site_of(g, p, h);
    for (int mu = 0; mu < 4; ++mu) {
        cd m[9];
        LinkIO<PREC, RECON>::load(data, g.Vh, (unsigned)((p * 4 + mu) * LinkIO<PREC, RECON>::NPL * g.Vh + h), m);
        double* q = dst + ((long)s.lex * 4 + mu) * 18;
        for (int k = 0; k < 9; ++k) { q[2 * k] = m[k].x; q[2 * k + 1] = m[k].y; }
    }
}
template <int PREC, int RECON>
__global__ void k_gauge_random(Geo g, double eps, unsigned seed, char* data, size_t blk_bytes, int unit) {
    int p, h;
    if (!thread_site(g, Blocks{0, 2}, p, h)) return;
    const Site s = site_of(g, p, h);
    for (int mu = 0; mu < 4; ++mu) {
        cd m[9];
        if (unit) {
            for (int k = 0; k < 9; ++k) m[k] = cd{(k % 4 == 0) ? 1.0 : 0.0, 0.0};
        } else {
            random_link((unsigned)(s.lex * 4 + mu), eps, seed, m);
        }
        LinkIO<PREC, RECON>::store(data, g.Vh, (unsigned)((p * 4 + mu) * LinkIO<PREC, RECON>::NPL * g.Vh + h), m);
    }
}
template <int PREC>
__global__ void k_wilson_gaussian(Geo g, Blocks b, unsigned seed, char* blk0, size_t blk_bytes, float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    cd v[12];
    const double is2 = 0.70710678118654752440084436210485;
    for (int k = 0; k < 12; ++k) {
        double gg[2];
        gauss2((unsigned)s.lex, (unsigned)k, STREAM_WILSON, seed, gg);
        v[k] = cd{gg[0] * is2, gg[1] * is2};
    }
    const int ib = b.nblk == 2 ? p : 0;
    WilsonIO<PREC>::store(blk0, norm0, g.Vh, (unsigned)(ib * WilsonIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
}
template <int PREC>
__global__ void k_ks_gaussian(Geo g, Blocks b, unsigned seed, char* blk0, size_t blk_bytes, float* norm0) {
    int p, h;
    if (!thread_site(g, b, p, h)) return;
    const Site s = site_of(g, p, h);
    cd v[3];
    const double is2 = 0.70710678118654752440084436210485;
    for (int k = 0; k < 3; ++k) {
        double gg[2];
        gauss2((unsigned)s.lex, (unsigned)k, STREAM_KS, seed, gg);
        v[k] = cd{gg[0] * is2, gg[1] * is2};
    }
    const int ib = b.nblk == 2 ? p : 0;
    KSIO<PREC>::store(blk0, norm0, g.Vh, (unsigned)(ib * KSIO<PREC>::NPL * g.Vh + h), (unsigned)(ib * g.Vh + h), v);
}
// precision conversion copy of fermion blocks (same subset layout)
template <int PD, int PS, int NC>
__global__ void k_convert(int n_per_blk, int nblk, char* d0, size_t dbb, float* dn0, const char* s0, size_t sbb,
                          const float* sn0) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)nblk * n_per_blk) return;
    const int ib = (int)(gid / n_per_blk);
    const int h = (int)(gid - (long)ib * n_per_blk);
    if constexpr (NC == 12) {
        cf v[12];
        WilsonIO<PS>::load(s0, sn0, n_per_blk, (unsigned)(ib * WilsonIO<PS>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
        WilsonIO<PD>::store(d0, dn0, n_per_blk, (unsigned)(ib * WilsonIO<PD>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
    } else {
        cf v[3];
        KSIO<PS>::load(s0, sn0, n_per_blk, (unsigned)(ib * KSIO<PS>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
        KSIO<PD>::store(d0, dn0, n_per_blk, (unsigned)(ib * KSIO<PD>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
    }
}
template <int PD, int PS, int NC>
__global__ void k_convert_d(int n_per_blk, int nblk, char* d0, size_t dbb, float* dn0, const char* s0, size_t sbb,
                            const float* sn0) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)nblk * n_per_blk) return;
    const int ib = (int)(gid / n_per_blk);
    const int h = (int)(gid - (long)ib * n_per_blk);
    if constexpr (NC == 12) {
        cd v[12];
        WilsonIO<PS>::load(s0, sn0, n_per_blk, (unsigned)(ib * WilsonIO<PS>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
        WilsonIO<PD>::store(d0, dn0, n_per_blk, (unsigned)(ib * WilsonIO<PD>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
    } else {
        cd v[3];
        KSIO<PS>::load(s0, sn0, n_per_blk, (unsigned)(ib * KSIO<PS>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
        KSIO<PD>::store(d0, dn0, n_per_blk, (unsigned)(ib * KSIO<PD>::NPL * n_per_blk + h), (unsigned)(ib * n_per_blk + h), v);
    }
}
__global__ void k_point(char* blk, int Vh, int h, int comp, int prec, int kind, float* nrm) {
    // writes 1 at one complex component of one site of a zeroed block
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (kind == CLG_WILSON) {
        cd v[12];
        for (int k = 0; k < 12; ++k) v[k] = cd{k == comp ? 1.0 : 0.0, 0.0};
        if (prec == CLG_F64) WilsonIO<CLG_F64>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
        else if (prec == CLG_F32) WilsonIO<CLG_F32>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
        else WilsonIO<CLG_F16>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
    } else {
        cd v[3];
        for (int k = 0; k < 3; ++k) v[k] = cd{k == comp ? 1.0 : 0.0, 0.0};
        if (prec == CLG_F64) KSIO<CLG_F64>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
        else if (prec == CLG_F32) KSIO<CLG_F32>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
        else KSIO<CLG_F16>::store(blk, nrm, Vh, (unsigned)h, (unsigned)h, v);
    }
}

// gauge precision / compression conversion
template <int PD, int RD, int PS, int RS>
__global__ void k_gauge_convert(int Vh, char* dst, const char* src) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 8L * Vh) return;
    const int blk = (int)(gid / Vh);  // (parity, mu) block
    const int h = (int)(gid - (long)blk * Vh);
    cd m[9];
    LinkIO<PS, RS>::load(src, Vh, (unsigned)(blk * LinkIO<PS, RS>::NPL * Vh + h), m);
    LinkIO<PD, RD>::store(dst, Vh, (unsigned)(blk * LinkIO<PD, RD>::NPL * Vh + h), m);
}

Blocks blocks_of(const clg_field_s* f) {
    if (f->subset == CLG_FULL) return Blocks{0, 2};
    return Blocks{field_parity(f), 1};
}

const int kBlock = 256;

}  // namespace

// ============================================================== scratch fields
namespace clg {
int gauge_convert(clg_context ctx, clg_field_s* dst, const clg_field_s* src) {
    CLG_CHECK_ARG(dst && src && is_gauge(dst->kind) && is_gauge(src->kind), "need gauge fields");
    CLG_TRY(check_same_lattice(dst, src));
    const int Vh = (int)src->Vh;
    const unsigned grid = grid_for(8L * Vh, kBlock);
    hipStream_t st = ctx->stream;
    const int rd = dst->kind == CLG_GAUGE12 ? 12 : 18;
    CLG_CHECK_ARG(src->prec == CLG_F64 && src->kind == CLG_GAUGE, "conversion source must be the FP64 18-real field");
    char* d = (char*)dst->data;
    const char* s = (const char*)src->data;
    if (dst->prec == CLG_F64) {
        if (rd == 12) k_gauge_convert<CLG_F64, 12, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
        else k_gauge_convert<CLG_F64, 18, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
    } else if (dst->prec == CLG_F32) {
        if (rd == 12) k_gauge_convert<CLG_F32, 12, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
        else k_gauge_convert<CLG_F32, 18, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
    } else {
        if (rd == 12) k_gauge_convert<CLG_F16, 12, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
        else k_gauge_convert<CLG_F16, 18, CLG_F64, 18><<<grid, kBlock, 0, st>>>(Vh, d, s);
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int scratch_field(clg_context ctx, const clg_field_s* like, int kind, int prec, int subset, int slot,
                  clg_field_s** out) {
    // one context-owned field per (kind, prec, subset, lattice, slot id)
    for (clg_field_s* f : ctx->scratch) {
        if (f->scratch_slot == slot && f->kind == kind && f->prec == prec && f->subset == subset &&
            f->dims[0] == like->dims[0] && f->dims[1] == like->dims[1] && f->dims[2] == like->dims[2] &&
            f->dims[3] == like->dims[3]) {
            for (int i = 0; i < 4; ++i) f->bc[i] = like->bc[i];
            *out = f;
            return CLG_OK;
        }
    }
    clg_field_s* f = nullptr;
    CLG_TRY(clg_field_create(ctx, kind, prec, subset, like->dims, &f));
    f->scratch_slot = slot;
    ctx->scratch.push_back(f);
    for (int i = 0; i < 4; ++i) f->bc[i] = like->bc[i];
    *out = f;
    return CLG_OK;
}

bool stream_capturing(clg_context ctx) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess) return false;
    return cs != hipStreamCaptureStatusNone;
}

static bool cycle_graphs_enabled() {
    const char* e = std::getenv("CLG_GRAPHS");
    return !(e && e[0] == '0');
}

int cycle_graph_run(clg_context ctx, const std::vector<unsigned long long>& key, const std::function<int()>& record) {
    // inside a caller's own recording the work is simply recorded there
    if (!cycle_graphs_enabled() || stream_capturing(ctx)) return record();
    auto it = ctx->cycle_graphs.find(key);
    if (it == ctx->cycle_graphs.end()) {
        constexpr size_t kMaxCycleGraphs = 32;
        if (ctx->cycle_graphs.size() >= kMaxCycleGraphs) {
            CLG_HIP(hipStreamSynchronize(ctx->stream));
            for (auto& kv : ctx->cycle_graphs) (void)hipGraphExecDestroy(kv.second);
            ctx->cycle_graphs.clear();
        }
        CLG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        const int st = record();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st != CLG_OK || e != hipSuccess || !g) {
            (void)hipGetLastError();
            if (g) (void)hipGraphDestroy(g);
            if (st != CLG_OK) return st;
            set_error("recording a solver cycle failed: %s", hipGetErrorString(e));
            return CLG_HIP_ERROR;
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            set_error("hipGraphInstantiate (solver cycle): %s", hipGetErrorString(ei));
            return CLG_HIP_ERROR;
        }
        it = ctx->cycle_graphs.emplace(key, ex).first;
    }
    CLG_HIP(hipGraphLaunch(it->second, ctx->stream));
    return CLG_OK;
}

// ---- persistent tuning cache
std::string tune_cache_path() {
    const char* e = std::getenv("CLG_TUNE_CACHE");
    return e ? std::string(e) : std::string();
}

std::string device_arch() {
    static std::mutex m;
    static std::map<int, std::string> names;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(m);
    auto it = names.find(dev);
    if (it != names.end()) return it->second;
    hipDeviceProp_t prop;
    std::string name = "unknown";
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        name = prop.gcnArchName;
        for (char& c : name)
            if (std::isspace((unsigned char)c)) c = '_';
    }
    names[dev] = name;
    return name;
}

static std::mutex& cache_file_mutex() {
    static std::mutex m;
    return m;
}

int tune_cache_read(const std::string& path, const char* tuner, std::vector<std::pair<std::vector<long>, std::vector<long>>>& out) {
    out.clear();
    std::lock_guard<std::mutex> lk(cache_file_mutex());
    std::FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return 0;
    const std::string arch = device_arch();
    char line[2048];
    while (std::fgets(line, sizeof(line), f)) {
        std::istringstream in(line);
        std::string a, t, tok;
        if (!(in >> a >> t) || a != arch || t != tuner) continue;
        std::vector<long> key, cfg;
        bool colon = false, ok = true;
        while (in >> tok) {
            if (tok == ":") { colon = true; continue; }
            char* end = nullptr;
            const long v = std::strtol(tok.c_str(), &end, 10);
            if (!end || *end) { ok = false; break; }
            (colon ? cfg : key).push_back(v);
        }
        if (ok && colon && !key.empty() && !cfg.empty()) out.emplace_back(key, cfg);
    }
    std::fclose(f);
    return (int)out.size();
}

void tune_cache_append(const std::string& path, const char* tuner, const std::vector<long>& key, const std::vector<long>& cfg) {
    if (path.empty()) return;
    std::ostringstream o;
    o << device_arch() << ' ' << tuner;
    for (long v : key) o << ' ' << v;
    o << " :";
    for (long v : cfg) o << ' ' << v;
    o << '\n';
    std::lock_guard<std::mutex> lk(cache_file_mutex());
    if (std::FILE* f = std::fopen(path.c_str(), "a")) {
        std::fputs(o.str().c_str(), f);
        std::fclose(f);
    }
}
}  // namespace clg

struct clg_graph_s {
    clg_context ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nodes = 0;
};

// ============================================================== C ABI
extern "C" {

const char* clg_status_string(int s) {
    switch (s) {
        case CLG_OK: return "ok";
        case CLG_NOT_CONVERGED: return "solver did not converge (best iterate written)";
        case CLG_BAD_ARG: return "bad argument";
        case CLG_HIP_ERROR: return "HIP runtime error";
        case CLG_UNSUPPORTED: return "unsupported combination";
        case CLG_OUT_OF_MEMORY: return "out of device memory";
    }
    return "unknown status";
}

const char* clg_last_error(void) { return clg::last_error(); }

}  // extern "C"

namespace {
int init_context(clg_context c) {
    hipDeviceProp_t prop;
    CLG_HIP(hipGetDeviceProperties(&prop, c->device));
    c->num_cu = prop.multiProcessorCount;
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    CLG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CLG_HIP(hipMalloc((void**)&c->red_counter, 64 * sizeof(unsigned)));
    CLG_HIP(hipMemsetAsync(c->red_counter, 0, 64 * sizeof(unsigned), c->stream));
    CLG_HIP(hipMalloc((void**)&c->red_result, 4096 * sizeof(double)));
    CLG_HIP(hipHostMalloc((void**)&c->host_result, 4096 * sizeof(double), hipHostMallocDefault));
    CLG_HIP(hipEventCreate(&c->ev0));
    CLG_HIP(hipEventCreate(&c->ev1));
    CLG_TRY(noalloc_scratch_init(c->device));
    CLG_HIP(hipStreamSynchronize(c->stream));
    return CLG_OK;
}
}  // namespace

extern "C" {

int clg_init(int device, clg_context* out) {
    CLG_CHECK_ARG(out, "null output handle");
    int n = 0;
    CLG_HIP(hipGetDeviceCount(&n));
    CLG_CHECK_ARG(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
    CLG_HIP(hipSetDevice(device));
    clg_context c = new clg_context_s();
    c->device = device;
    // a failure part way releases whatever was made so far (clg_finalize skips what is still null)
    const int st = init_context(c);
    if (st != CLG_OK) {
        clg_finalize(c);
        return st;
    }
    *out = c;
    return CLG_OK;
}

int clg_finalize(clg_context c) {
    if (!c) return CLG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // solvers and graphs the caller left: newest first, while the context is still whole
    std::vector<std::pair<void*, void (*)(void*)>> owned;
    owned.swap(c->owned);
    for (auto it = owned.rbegin(); it != owned.rend(); ++it) it->second(it->first);
    for (auto& kv : c->cycle_graphs) (void)hipGraphExecDestroy(kv.second);
    c->cycle_graphs.clear();
    for (clg_field_s* f : c->scratch) clg_field_destroy(f);
    c->scratch.clear();
    if (c->red_partial) (void)hipFree(c->red_partial);
    if (c->red_counter) (void)hipFree(c->red_counter);
    if (c->red_result) (void)hipFree(c->red_result);
    if (c->host_result) (void)hipHostFree(c->host_result);
    if (c->ws) (void)hipFree(c->ws);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CLG_OK;
}

int clg_synchronize(clg_context c) {
    CLG_CHECK_ARG(c, "null context");
    CLG_CHECK_ARG(!stream_capturing(c), "clg_synchronize while the context records a graph");
    CLG_HIP(hipStreamSynchronize(c->stream));
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// ---- graphs: record the launches of non-blocking entry points on the context stream, replay them
int clg_graph_begin(clg_context c) {
    CLG_CHECK_ARG(c, "null context");
    CLG_CHECK_ARG(!stream_capturing(c), "the context already records a graph");
    (void)hipSetDevice(c->device);
    // recorded reductions keep the partial-sum buffer they were recorded with: allocate it first
    CLG_TRY(reduction_ready(c));
    CLG_HIP(hipStreamSynchronize(c->stream));
    CLG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    return CLG_OK;
}

static void graph_free(void* obj);

int clg_graph_end(clg_context c, clg_graph* out) {
    CLG_CHECK_ARG(c && out, "null argument");
    CLG_CHECK_ARG(stream_capturing(c), "the context records no graph");
    (void)hipSetDevice(c->device);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || !g) {
        (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
        set_error("recording failed (%s): a blocking call or an allocation-growing call was made while recording",
                  hipGetErrorString(e));
        return CLG_HIP_ERROR;
    }
    size_t nodes = 0;
    (void)hipGraphGetNodes(g, nullptr, &nodes);
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ei));
        return CLG_HIP_ERROR;
    }
    clg_graph_s* r = new clg_graph_s();
    r->ctx = c;
    r->graph = g;
    r->exec = ex;
    r->nodes = (int)nodes;
    ctx_adopt(c, r, graph_free);
    *out = r;
    return CLG_OK;
}

int clg_graph_launch(clg_graph g) {
    CLG_CHECK_ARG(g && g->exec, "null graph");
    CLG_CHECK_ARG(!stream_capturing(g->ctx), "cannot replay a graph while recording");
    (void)hipSetDevice(g->ctx->device);
    CLG_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
    return CLG_OK;
}

int clg_graph_nodes(clg_graph g) { return g ? g->nodes : -1; }

static void graph_free(void* obj) {
    clg_graph_s* g = static_cast<clg_graph_s*>(obj);
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

int clg_graph_destroy(clg_graph g) {
    if (!g) return CLG_OK;
    ctx_release(g->ctx, g);
    graph_free(g);
    return CLG_OK;
}

int clg_device_info(clg_context c, char* buf, int len) {
    CLG_CHECK_ARG(c && buf && len > 0, "bad arguments");
    std::snprintf(buf, (size_t)len, "%s:%dCU", c->arch, c->num_cu);
    return CLG_OK;
}

int clg_field_create(clg_context ctx, int kind, int prec, int subset, const int dims[4], clg_field* out) {
    CLG_CHECK_ARG(ctx && out && dims, "null argument");
    CLG_CHECK_ARG(kind >= CLG_WILSON && kind <= CLG_GAUGE12, "bad field kind %d", kind);
    CLG_CHECK_ARG(prec >= CLG_F64 && prec <= CLG_F16, "bad precision %d", prec);
    CLG_CHECK_ARG(subset >= CLG_FULL && subset <= CLG_ODD, "bad subset %d", subset);
    CLG_CHECK_ARG(!(is_gauge(kind) && subset != CLG_FULL), "gauge fields are full-lattice");
    long V = 1;
    for (int i = 0; i < 4; ++i) {
        CLG_CHECK_ARG(dims[i] >= 1 && dims[i] <= 127, "lattice extent %d out of range [1,127]", dims[i]);
        V *= dims[i];
    }
    CLG_CHECK_ARG((dims[3] & 1) == 0, "Lt must be even for the even/odd site layout");
    CLG_CHECK_ARG(V * 127L < (1L << 31), "lattice too large for 32-bit site indexing");
    clg_field_s* f = new clg_field_s();
    f->ctx = ctx;
    f->kind = kind;
    f->prec = prec;
    f->subset = subset;
    for (int i = 0; i < 4; ++i) f->dims[i] = dims[i];
    f->V = V;
    f->Vh = V / 2;
    f->nsites = subset == CLG_FULL ? V : V / 2;
    f->block_bytes = site_bytes(kind, prec) * (size_t)f->Vh;
    const int nblk = is_gauge(kind) ? 8 : (subset == CLG_FULL ? 2 : 1);
    f->bytes = f->block_bytes * nblk;
    // kernels address a field by a 32-bit element index (block, plane and site; below 72 * Vh) and a
    // 64-bit byte offset, so its size is bounded by the site-count check above, not by 4 GiB
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(&f->data, f->bytes);
    if (e != hipSuccess) {
        delete f;
        set_error("hipMalloc(%zu) failed: %s", (size_t)(f->bytes), hipGetErrorString(e));
        return CLG_OUT_OF_MEMORY;
    }
    if (prec == CLG_F16 && is_fermion(kind)) {
        e = hipMalloc((void**)&f->norm, sizeof(float) * (size_t)f->Vh * nblk);
        if (e != hipSuccess) {
            (void)hipFree(f->data);
            delete f;
            set_error("hipMalloc(norm) failed: %s", hipGetErrorString(e));
            return CLG_OUT_OF_MEMORY;
        }
        f->bytes += sizeof(float) * (size_t)f->Vh * nblk;
    }
    *out = f;
    return clg_field_zero(f);
}

int clg_field_destroy(clg_field f) {
    if (!f) return CLG_OK;
    if (f->owns) {
        (void)hipSetDevice(f->ctx->device);
        (void)hipStreamSynchronize(f->ctx->stream);
        if (f->data) (void)hipFree(f->data);
        if (f->norm) (void)hipFree(f->norm);
    }
    delete f;
    return CLG_OK;
}

int clg_field_set_bc(clg_field f, const int bc[4]) {
    CLG_CHECK_ARG(f && bc, "null argument");
    for (int i = 0; i < 4; ++i) {
        CLG_CHECK_ARG(bc[i] == 1 || bc[i] == -1, "boundary sign must be +1 or -1");
        f->bc[i] = bc[i];
    }
    return CLG_OK;
}

long clg_field_host_doubles(clg_field f) {
    if (!f) return -1;
    if (is_gauge(f->kind)) return f->V * 72;
    return f->nsites * (f->kind == CLG_WILSON ? 24 : 6);
}

long clg_field_bytes(clg_field f) { return f ? (long)f->bytes : -1; }

int clg_field_parity_view(clg_field full, int parity_subset, clg_field* out) {
    CLG_CHECK_ARG(full && out, "null argument");
    CLG_CHECK_ARG(is_fermion(full->kind) && full->subset == CLG_FULL, "view needs a FULL fermion field");
    CLG_CHECK_ARG(parity_subset == CLG_EVEN || parity_subset == CLG_ODD, "subset must be EVEN or ODD");
    clg_field_s* v = new clg_field_s(*full);
    v->owns = false;
    v->parent = full;
    v->subset = parity_subset;
    v->nsites = full->Vh;
    const int p = parity_subset == CLG_ODD ? 1 : 0;
    v->data = (char*)full->data + (size_t)p * full->block_bytes;
    v->norm = full->norm ? full->norm + (size_t)p * full->Vh : nullptr;
    v->bytes = full->block_bytes;
    *out = v;
    return CLG_OK;
}

int clg_field_zero(clg_field f) {
    CLG_CHECK_ARG(f, "null field");
    (void)hipSetDevice(f->ctx->device);
    const size_t nblk = is_gauge(f->kind) ? 8 : (f->subset == CLG_FULL ? 2 : 1);
    hipStream_t st = f->ctx->stream;
    const size_t bytes = f->block_bytes * nblk;
    const unsigned cap = (unsigned)f->ctx->num_cu * 16;
    if (bytes % 16 == 0 && ((size_t)f->data & 15) == 0) {
        const long n = (long)(bytes / 16);
        k_zero16<<<std::max(1u, std::min(cap, (unsigned)((n + 255) / 256))), 256, 0, st>>>((uint4*)f->data, n);
    } else {
        const long n = (long)(bytes / 4);  // block sizes are multiples of 4 bytes
        k_zero4<<<std::max(1u, std::min(cap, (unsigned)((n + 255) / 256))), 256, 0, st>>>((unsigned*)f->data, n);
    }
    CLG_HIP(hipGetLastError());
    if (f->norm) {
        const long n = (long)(f->Vh * nblk);
        k_zero4<<<std::max(1u, std::min(cap, (unsigned)((n + 255) / 256))), 256, 0, st>>>((unsigned*)f->norm, n);
        CLG_HIP(hipGetLastError());
    }
    return CLG_OK;
}

#define DISPATCH_PREC(prec, KERNEL, ...)                                      \
    do {                                                                      \
        if ((prec) == CLG_F64) KERNEL<CLG_F64><<<grid, kBlock, 0, st>>>(__VA_ARGS__); \
        else if ((prec) == CLG_F32) KERNEL<CLG_F32><<<grid, kBlock, 0, st>>>(__VA_ARGS__); \
        else KERNEL<CLG_F16><<<grid, kBlock, 0, st>>>(__VA_ARGS__);           \
    } while (0)

#define DISPATCH_GAUGE(f, KERNEL, ...)                                                          \
    do {                                                                                        \
        const bool r12 = (f)->kind == CLG_GAUGE12;                                              \
        if ((f)->prec == CLG_F64) {                                                             \
            if (r12) KERNEL<CLG_F64, 12><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                 \
            else KERNEL<CLG_F64, 18><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                     \
        } else if ((f)->prec == CLG_F32) {                                                      \
            if (r12) KERNEL<CLG_F32, 12><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                 \
            else KERNEL<CLG_F32, 18><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                     \
        } else {                                                                                \
            if (r12) KERNEL<CLG_F16, 12><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                 \
            else KERNEL<CLG_F16, 18><<<grid, kBlock, 0, st>>>(__VA_ARGS__);                     \
        }                                                                                       \
    } while (0)

int clg_field_upload(clg_field f, const double* host) {
    CLG_CHECK_ARG(f && host, "null argument");
    (void)hipSetDevice(f->ctx->device);
    hipStream_t st = f->ctx->stream;
    const size_t nd = (size_t)clg_field_host_doubles(f);
    double* stage = nullptr;
    CLG_HIP(hipMalloc((void**)&stage, nd * sizeof(double)));
    hipError_t e = hipMemcpyAsync(stage, host, nd * sizeof(double), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) { (void)hipFree(stage); CLG_HIP(e); }
    const Geo g = make_geo(f);
    if (is_gauge(f->kind)) {
        const unsigned grid = grid_for(f->V, kBlock);
        DISPATCH_GAUGE(f, k_gauge_from_aos, g, stage, (char*)f->data, f->block_bytes);
    } else {
        const Blocks b = blocks_of(f);
        const unsigned grid = grid_for((long)b.nblk * f->Vh, kBlock);
        const bool full = f->subset == CLG_FULL;
        if (f->kind == CLG_WILSON)
            DISPATCH_PREC(f->prec, k_wilson_from_aos, g, b, full, stage, (char*)f->data, f->block_bytes, f->norm);
        else
            DISPATCH_PREC(f->prec, k_ks_from_aos, g, b, full, stage, (char*)f->data, f->block_bytes, f->norm);
    }
    e = hipStreamSynchronize(st);
    (void)hipFree(stage);
    CLG_HIP(e);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_field_download(clg_field f, double* host) {
    CLG_CHECK_ARG(f && host, "null argument");
    (void)hipSetDevice(f->ctx->device);
    hipStream_t st = f->ctx->stream;
    const size_t nd = (size_t)clg_field_host_doubles(f);
    double* stage = nullptr;
    CLG_HIP(hipMalloc((void**)&stage, nd * sizeof(double)));
    const Geo g = make_geo(f);
    if (is_gauge(f->kind)) {
        const unsigned grid = grid_for(f->V, kBlock);
        DISPATCH_GAUGE(f, k_gauge_to_aos, g, stage, (const char*)f->data, f->block_bytes);
    } else {
        const Blocks b = blocks_of(f);
        const unsigned grid = grid_for((long)b.nblk * f->Vh, kBlock);
        const bool full = f->subset == CLG_FULL;
        if (f->kind == CLG_WILSON)
            DISPATCH_PREC(f->prec, k_wilson_to_aos, g, b, full, stage, (const char*)f->data, f->block_bytes, f->norm);
        else
            DISPATCH_PREC(f->prec, k_ks_to_aos, g, b, full, stage, (const char*)f->data, f->block_bytes, f->norm);
    }
    hipError_t e = hipMemcpyAsync(host, stage, nd * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(stage);
    CLG_HIP(e);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_field_copy(clg_field dst, clg_field src) {
    CLG_CHECK_ARG(dst && src, "null argument");
    CLG_TRY(check_same_lattice(dst, src));
    (void)hipSetDevice(dst->ctx->device);
    hipStream_t st = dst->ctx->stream;
    if (dst == src) return CLG_OK;
    if (is_gauge(dst->kind) && is_gauge(src->kind) && (dst->kind != src->kind || dst->prec != src->prec))
        return gauge_convert(dst->ctx, dst, src);
    CLG_CHECK_ARG(dst->kind == src->kind && dst->subset == src->subset, "copy needs the same kind and subset");
    if (dst->prec == src->prec) {
        const size_t nblk = is_gauge(dst->kind) ? 8 : (dst->subset == CLG_FULL ? 2 : 1);
        CLG_HIP(hipMemcpyAsync(dst->data, src->data, dst->block_bytes * nblk, hipMemcpyDeviceToDevice, st));
        if (dst->norm)
            CLG_HIP(hipMemcpyAsync(dst->norm, src->norm, sizeof(float) * dst->Vh * nblk, hipMemcpyDeviceToDevice, st));
        return CLG_OK;
    }
    const int nblk = dst->subset == CLG_FULL ? 2 : 1;
    const unsigned grid = grid_for((long)nblk * dst->Vh, kBlock);
    const int n = (int)dst->Vh;
    char* d0 = (char*)dst->data;
    const char* s0 = (const char*)src->data;
#define CONV(PD, PS)                                                                                          \
    if (dst->prec == PD && src->prec == PS) {                                                                 \
        if (PD == CLG_F64 || PS == CLG_F64) {                                                                 \
            if (dst->kind == CLG_WILSON)                                                                      \
                k_convert_d<PD, PS, 12><<<grid, kBlock, 0, st>>>(n, nblk, d0, dst->block_bytes, dst->norm, s0, \
                                                                 src->block_bytes, src->norm);                \
            else                                                                                              \
                k_convert_d<PD, PS, 3><<<grid, kBlock, 0, st>>>(n, nblk, d0, dst->block_bytes, dst->norm, s0,  \
                                                                src->block_bytes, src->norm);                 \
        } else {                                                                                              \
            if (dst->kind == CLG_WILSON)                                                                      \
                k_convert<PD, PS, 12><<<grid, kBlock, 0, st>>>(n, nblk, d0, dst->block_bytes, dst->norm, s0,   \
                                                               src->block_bytes, src->norm);                  \
            else                                                                                              \
                k_convert<PD, PS, 3><<<grid, kBlock, 0, st>>>(n, nblk, d0, dst->block_bytes, dst->norm, s0,    \
                                                              src->block_bytes, src->norm);                   \
        }                                                                                                     \
    }
    CONV(CLG_F64, CLG_F32)
    CONV(CLG_F64, CLG_F16)
    CONV(CLG_F32, CLG_F64)
    CONV(CLG_F32, CLG_F16)
    CONV(CLG_F16, CLG_F64)
    CONV(CLG_F16, CLG_F32)
#undef CONV
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_gauge_random(clg_field f, double eps, unsigned seed) {
    CLG_CHECK_ARG(f && is_gauge(f->kind), "need a gauge field");
    (void)hipSetDevice(f->ctx->device);
    hipStream_t st = f->ctx->stream;
    const Geo g = make_geo(f);
    const unsigned grid = grid_for(f->V, kBlock);
    DISPATCH_GAUGE(f, k_gauge_random, g, eps, seed, (char*)f->data, f->block_bytes, 0);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_gauge_unit(clg_field f) {
    CLG_CHECK_ARG(f && is_gauge(f->kind), "need a gauge field");
    (void)hipSetDevice(f->ctx->device);
    hipStream_t st = f->ctx->stream;
    const Geo g = make_geo(f);
    const unsigned grid = grid_for(f->V, kBlock);
    DISPATCH_GAUGE(f, k_gauge_random, g, 0.0, 0u, (char*)f->data, f->block_bytes, 1);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_field_gaussian(clg_field f, unsigned seed) {
    CLG_CHECK_ARG(f && is_fermion(f->kind), "need a fermion field");
    (void)hipSetDevice(f->ctx->device);
    hipStream_t st = f->ctx->stream;
    const Geo g = make_geo(f);
    const Blocks b = blocks_of(f);
    const unsigned grid = grid_for((long)b.nblk * f->Vh, kBlock);
    if (f->kind == CLG_WILSON)
        DISPATCH_PREC(f->prec, k_wilson_gaussian, g, b, seed, (char*)f->data, f->block_bytes, f->norm);
    else
        DISPATCH_PREC(f->prec, k_ks_gaussian, g, b, seed, (char*)f->data, f->block_bytes, f->norm);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_field_point_source(clg_field f, const int coord[4], int spin, int colour) {
    CLG_CHECK_ARG(f && coord && is_fermion(f->kind), "need a fermion field");
    for (int i = 0; i < 4; ++i) CLG_CHECK_ARG(coord[i] >= 0 && coord[i] < f->dims[i], "coordinate out of range");
    CLG_CHECK_ARG(colour >= 0 && colour < 3, "colour out of range");
    CLG_CHECK_ARG(f->kind != CLG_WILSON || (spin >= 0 && spin < 4), "spin out of range");
    CLG_TRY(clg_field_zero(f));
    const long lex = ((long)(coord[0] * f->dims[1] + coord[1]) * f->dims[2] + coord[2]) * f->dims[3] + coord[3];
    const int p = (coord[0] + coord[1] + coord[2] + coord[3]) & 1;
    if (f->subset != CLG_FULL && p != field_parity(f)) return CLG_OK;  // site not on this half
    const int comp = f->kind == CLG_WILSON ? spin * 3 + colour : colour;
    hipStream_t st = f->ctx->stream;
    k_point<<<1, 64, 0, st>>>(block_ptr(f, p), (int)f->Vh, (int)(lex >> 1), comp, f->prec, f->kind, norm_ptr(f, p));
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // extern "C"
