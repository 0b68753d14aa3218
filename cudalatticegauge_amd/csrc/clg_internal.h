// clg_internal.h -- shared definitions of libclg_hip.so (not installed, not part of the ABI).
//
// Data layout in HBM (all fields):
//   * Sites are split by parity.  The e-th site of parity p is the lexicographic site
//     2e or 2e+1 with parity p (the reference's even/odd numbering, CLGLib_Private.h:72-116),
//     so a half-lattice index h maps to coordinates with one division chain and Lt even.
//   * A FULL fermion field is [parity-0 block | parity-1 block]; an EVEN/ODD field is one
//     block.  A block is structure-of-arrays: `nplane` planes of Vh 16-byte (or 8/4-byte)
//     chunks, so lane i of a wavefront reads chunk i of a plane: every load is a contiguous
//     64-lane burst.
//   * Gauge fields hold 8 blocks, (parity, mu), each a set of link planes over Vh sites:
//     U_mu(x) for x of that parity.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/clg_hip.h"

namespace clg {

// ------------------------------------------------------------------ errors
void set_error(const char* fmt, ...);
const char* last_error();

#define CLG_CHECK_ARG(cond, ...)              \
    do {                                      \
        if (!(cond)) {                        \
            ::clg::set_error(__VA_ARGS__);    \
            return CLG_BAD_ARG;               \
        }                                     \
    } while (0)

#define CLG_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            ::clg::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e),   \
                             __FILE__, __LINE__);                                      \
            return CLG_HIP_ERROR;                                                      \
        }                                                                              \
    } while (0)

#define CLG_TRY(call)                 \
    do {                              \
        int _s = (call);              \
        if (_s != CLG_OK) return _s;  \
    } while (0)

// ------------------------------------------------------------------ geometry
struct Geo {
    int L[4];       // x, y, z, t extents
    int stride[4];  // lexicographic strides (t fastest)
    int sgn[4];     // fermion boundary sign per direction
    int Vh;         // half volume
};

// ------------------------------------------------------------------ handles
}  // namespace clg

struct clg_field_s;

struct clg_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    char arch[64] = {0};
    // reduction scratch: per-block partials, a done-counter, device results
    double* red_partial = nullptr;
    size_t red_partial_cap = 0;  // doubles
    unsigned* red_counter = nullptr;
    double* red_result = nullptr;  // 4096 doubles of device-resident results
    double* host_result = nullptr;  // pinned
    // solver workspace (Hessenberg, Givens, scalars), grow-only
    double* ws = nullptr;
    size_t ws_cap = 0;
    // scratch fields keyed by shape (owned)
    std::vector<clg_field_s*> scratch;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // objects made on this context (GCRO-DR solvers, graphs) with their destroy functions:
    // clg_finalize destroys those the caller has not
    std::vector<std::pair<void*, void (*)(void*)>> owned;
    // recorded solver cycles (clg::cycle_graph_run): key -> instantiated graph
    std::map<std::vector<unsigned long long>, hipGraphExec_t> cycle_graphs;
};

struct clg_field_s {
    clg_context ctx = nullptr;
    int kind = 0, prec = 0, subset = 0;
    int dims[4] = {0, 0, 0, 0};
    int bc[4] = {1, 1, 1, -1};
    long V = 0;
    long Vh = 0;
    long nsites = 0;       // V (FULL) or Vh
    void* data = nullptr;  // device storage
    float* norm = nullptr; // FP16 fields: per-site scale, [parity block][Vh]
    size_t bytes = 0;
    bool owns = true;
    clg_field_s* parent = nullptr;  // views
    // size of one parity block in bytes (fermions) or one (parity,mu) block (gauge)
    size_t block_bytes = 0;
    int scratch_slot = -1;  // context-owned scratch fields: the slot id they were created for
};

namespace clg {

// ------------------------------------------------------------------ layout sizes
// bytes of one site's worth of a fermion parity block / one link of a gauge block
inline size_t site_bytes(int kind, int prec) {
    switch (kind) {
        case CLG_WILSON: return prec == CLG_F64 ? 192 : prec == CLG_F32 ? 96 : 48;
        case CLG_KS: return prec == CLG_F64 ? 48 : prec == CLG_F32 ? 24 : 12;
        case CLG_GAUGE: return prec == CLG_F64 ? 144 : prec == CLG_F32 ? 72 : 36;
        case CLG_GAUGE12: return prec == CLG_F64 ? 96 : prec == CLG_F32 ? 48 : 24;
    }
    return 0;
}
inline bool is_gauge(int kind) { return kind == CLG_GAUGE || kind == CLG_GAUGE12; }
inline bool is_fermion(int kind) { return kind == CLG_WILSON || kind == CLG_KS; }
inline int parity_of_subset(int subset) { return subset == CLG_ODD ? 1 : 0; }

Geo make_geo(const clg_field_s* f);
int check_same_lattice(const clg_field_s* a, const clg_field_s* b);
// device pointer of parity block p of a fermion field (FULL: p in {0,1}; half: p must match)
inline char* block_ptr(const clg_field_s* f, int p) {
    if (f->subset == CLG_FULL) return (char*)f->data + (size_t)p * f->block_bytes;
    return (char*)f->data;
}
inline float* norm_ptr(const clg_field_s* f, int p) {
    if (!f->norm) return nullptr;
    if (f->subset == CLG_FULL) return f->norm + (size_t)p * f->Vh;
    return f->norm;
}
inline int field_parity(const clg_field_s* f) { return f->subset == CLG_ODD ? 1 : 0; }
// true when the device storage of a and b overlaps (data bytes or FP16 site scales); parity views
// of one FULL field overlap it, and the two parity views of one field do not overlap each other
inline size_t data_bytes(const clg_field_s* f) {
    return f->block_bytes * (size_t)(is_gauge(f->kind) ? 8 : (f->subset == CLG_FULL ? 2 : 1));
}
inline bool fields_overlap(const clg_field_s* a, const clg_field_s* b) {
    if (!a || !b) return false;
    const char *a0 = (const char*)a->data, *b0 = (const char*)b->data;
    if (a0 < b0 + data_bytes(b) && b0 < a0 + data_bytes(a)) return true;
    if (a->norm && b->norm) {
        const size_t na = (size_t)a->nsites, nb = (size_t)b->nsites;
        if (a->norm < b->norm + nb && b->norm < a->norm + na) return true;
    }
    return false;
}
inline void ctx_adopt(clg_context c, void* obj, void (*destroy)(void*)) { c->owned.emplace_back(obj, destroy); }
inline void ctx_release(clg_context c, void* obj) {
    for (auto it = c->owned.begin(); it != c->owned.end(); ++it)
        if (it->first == obj) {
            c->owned.erase(it);
            return;
        }
}

// reductions (clg_blas.hip)
int ensure_reduction_scratch(clg_context ctx, size_t doubles);
// allocates the partial-sum buffer every reduction kernel uses (its final size)
int reduction_ready(clg_context ctx);
// looks up, once per device, the device address of clg_multidot_noalloc's static partial-sum array
int noalloc_scratch_init(int device);
// device scratch (doubles) one Arnoldi orthogonalisation step (arn_step, clg_arnoldi.hip) needs
constexpr int kArnScratch = 264;
// scratch field of the same shape as `like` (kind/prec/subset/dims), index slot
int scratch_field(clg_context ctx, const clg_field_s* like, int kind, int prec, int subset,
                  int slot, clg_field_s** out);
// true while the context stream records a graph (blocking calls are refused then)
bool stream_capturing(clg_context ctx);
// Runs the device work of one solver cycle as a recorded graph: the first call with a key records
// what `record` launches on the context stream (it must neither block nor allocate) and instantiates
// it; every call (the first included) then replays it.  The key names every pointer and scalar the
// recorded launches take.  CLG_GRAPHS=0 runs `record` directly instead.
int cycle_graph_run(clg_context ctx, const std::vector<unsigned long long>& key, const std::function<int()>& record);
// the bits of a double, for graph keys
inline unsigned long long key_bits(double v) {
    unsigned long long u;
    static_assert(sizeof(u) == sizeof(v), "64-bit double");
    __builtin_memcpy(&u, &v, sizeof(u));
    return u;
}

// launch autotuner (clg_wilson.hip, clg_batch.hip): process-wide switch and tuned-shape tables
bool tuning_enabled();
int batch_tune_count();
int batch_tune_entry(int index, char* buf, int len);
void batch_tune_reset();
// Persistent tuning cache (CLG_TUNE_CACHE=<file>): one text line per tuned key,
//   "<device arch> <tuner> <key integers> : <configuration integers>".
// A tuner reads the lines of its name and of the current device's architecture the first time it
// misses a key under a cache path (again after clg_autotune_reset), and appends every key it tunes.
std::string tune_cache_path();
std::string device_arch();
int tune_cache_read(const std::string& path, const char* tuner,
                    std::vector<std::pair<std::vector<long>, std::vector<long>>>& out);
void tune_cache_append(const std::string& path, const char* tuner, const std::vector<long>& key,
                       const std::vector<long>& cfg);
int ks_tune_count();
int ks_tune_entry(int index, char* buf, int len);
void ks_tune_reset();

// launch helpers
inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }
// grid for clgd::map_site: nblk parity blocks of Vh sites, optional parity interleave, spt logical
// blocks per workgroup and XCD padding
inline unsigned sched_grid(long Vh, int nblk, int block, int xcd, int pinter, int spt = 1) {
    long g = (nblk == 2 && pinter) ? 2 * ((Vh + block - 1) / block) : (nblk * Vh + block - 1) / block;
    g = (g + spt - 1) / spt;
    if (xcd > 1) g = (g + xcd - 1) / xcd * xcd;
    return (unsigned)g;
}

}  // namespace clg

// ======================================================================= device side
namespace clgd {

template <typename R> struct C2 { R x, y; };
typedef C2<double> cd;
typedef C2<float> cf;

template <typename R> __device__ __forceinline__ C2<R> cmk(R a, R b) { return C2<R>{a, b}; }
template <typename R> __device__ __forceinline__ C2<R> cadd(C2<R> a, C2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __device__ __forceinline__ C2<R> csub(C2<R> a, C2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename R> __device__ __forceinline__ C2<R> cmulc(C2<R> a, C2<R> b) {
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename R> __device__ __forceinline__ C2<R> cconj(C2<R> a) { return {a.x, -a.y}; }
template <typename R> __device__ __forceinline__ C2<R> cscale(C2<R> a, R s) { return {a.x * s, a.y * s}; }
// acc += a*b
template <typename R> __device__ __forceinline__ void cfma(C2<R>& acc, C2<R> a, C2<R> b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// acc += conj(a)*b
template <typename R> __device__ __forceinline__ void cfmac(C2<R>& acc, C2<R> a, C2<R> b) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
}
// i^k * a for compile-time k
template <int K, typename R> __device__ __forceinline__ C2<R> mulz4(C2<R> a) {
    if constexpr ((K & 3) == 1) return {-a.y, a.x};
    else if constexpr ((K & 3) == 2) return {-a.x, -a.y};
    else if constexpr ((K & 3) == 3) return {a.y, -a.x};
    else return a;
}

// ------------------------------------------------------------------ site geometry
struct Site {
    int c[4];
    int lex;
};
// coordinates of half-index h of parity p (Lt even)
__device__ __forceinline__ Site site_of(const clg::Geo& g, int p, int h) {
    Site s;
    const int lex0 = 2 * h;
    const int t0 = lex0 % g.L[3];
    int r = lex0 / g.L[3];
    s.c[2] = r % g.L[2];
    r /= g.L[2];
    s.c[1] = r % g.L[1];
    s.c[0] = r / g.L[1];
    const int b = (p ^ (s.c[0] + s.c[1] + s.c[2])) & 1;
    s.c[3] = t0 + b;
    s.lex = lex0 + b;
    return s;
}
struct Nbr {
    int p, h, sign;
};
template <int MU>
__device__ __forceinline__ Nbr fwd_nbr(const clg::Geo& g, const Site& s, int p) {
    Nbr n;
    int lex;
    const bool wrap = s.c[MU] == g.L[MU] - 1;
    if (wrap) {
        lex = s.lex - (g.L[MU] - 1) * g.stride[MU];
        n.sign = g.sgn[MU];
        n.p = p ^ 1 ^ (g.L[MU] & 1);
    } else {
        lex = s.lex + g.stride[MU];
        n.sign = 1;
        n.p = p ^ 1;
    }
    n.h = lex >> 1;
    return n;
}
template <int MU>
__device__ __forceinline__ Nbr bwd_nbr(const clg::Geo& g, const Site& s, int p) {
    Nbr n;
    int lex;
    const bool wrap = s.c[MU] == 0;
    if (wrap) {
        lex = s.lex + (g.L[MU] - 1) * g.stride[MU];
        n.sign = g.sgn[MU];
        n.p = p ^ 1 ^ (g.L[MU] & 1);
    } else {
        lex = s.lex - g.stride[MU];
        n.sign = 1;
        n.p = p ^ 1;
    }
    n.h = lex >> 1;
    return n;
}

// ------------------------------------------------------------------ work decomposition
// Workgroups are handed to the 8 XCDs round-robin (wg -> XCD wg % 8), and each XCD has its
// own L2.  With xcd = 8 the hardware id is remapped so that XCD k sweeps the contiguous
// logical range [k*n/8, (k+1)*n/8): neighbouring sites, whose spinors and links are shared,
// then meet in one L2.  With pinter = 1 a two-parity launch alternates parities block by
// block, so a link's forward user (parity p) and backward user (parity 1-p) are co-resident.
// The grid must be padded to a multiple of xcd (sched_grid) for the remap to be a bijection.
// A launch with spt logical blocks per workgroup (sites per thread) gives workgroup w the
// logical blocks w*spt .. w*spt+spt-1 after the remap, so each XCD still sweeps one range.
__device__ __forceinline__ unsigned logical_wg(int xcd) {
    unsigned wg = blockIdx.x;
    if (xcd > 1) {
        const unsigned per = gridDim.x / (unsigned)xcd;
        wg = (wg % (unsigned)xcd) * per + wg / (unsigned)xcd;
    }
    return wg;
}
// site (p, h) of this thread in logical block lb of blockDim.x sites
__device__ __forceinline__ bool block_site(unsigned lb, int Vh, int nblk, int pbase, int pinter, int& p, int& h) {
    if (nblk == 2 && pinter) {
        p = (int)(lb & 1u);
        h = (int)((lb >> 1) * blockDim.x + threadIdx.x);
        return h < Vh;
    }
    const long gid = (long)lb * blockDim.x + threadIdx.x;
    if (gid >= (long)nblk * Vh) return false;
    const int blk = (int)(gid / Vh);
    h = (int)(gid - (long)blk * Vh);
    p = nblk == 2 ? blk : pbase;
    return true;
}
// Tiled plane schedule for launches of one workgroup per (x,y) plane (blockDim = L2*L3/2 sites of one
// parity; nblk = 2 interleaves the parities plane by plane): the L0 x L1 grid of planes is cut into
// 8 tiles, 8/TY along x and TY along y, and XCD k (hardware workgroups k, k+8, ...) sweeps tile k,
// x-major (y fastest) or y-major.  A plane's x+-1 and y+-1 neighbour planes then come up within
// 2*L1/TY (x-major) workgroups of it in the same XCD's L2, instead of a whole x-slab later.  Needs
// L0 % (8/TY) == 0, L1 % TY == 0 and a grid of exactly nblk*L0*L1 workgroups.
__device__ __forceinline__ unsigned tile_block(const clg::Geo& g, int nblk, int TY, int ymajor) {
    const unsigned wg = blockIdx.x;
    const unsigned k = wg & 7u, j = wg >> 3;
    const int TX = 8 / TY;
    const int tx = g.L[0] / TX, ty = g.L[1] / TY;
    const int ix = (int)k / TY, iy = (int)k % TY;
    const unsigned par = nblk == 2 ? (j & 1u) : 0u;
    const int q = (int)(nblk == 2 ? (j >> 1) : j);
    int xl, yl;
    if (ymajor) {
        yl = q / tx;
        xl = q - yl * tx;
    } else {
        xl = q / ty;
        yl = q - xl * ty;
    }
    const unsigned P = (unsigned)((ix * tx + xl) * g.L[1] + iy * ty + yl);
    return nblk == 2 ? 2u * P + par : P;
}
__device__ __forceinline__ bool map_site(int Vh, int nblk, int pbase, int xcd, int pinter, int& p, int& h) {
    return block_site(logical_wg(xcd), Vh, nblk, pbase, pinter, p, h);
}

// ------------------------------------------------------------------ storage traits
// Every field is a uniform base pointer plus a set of planes.  An element index e selects
// one site inside the planes (e = block * NPL * Vh + h for a field of several blocks), and
// plane k of that site sits at base + k*Vh*CH + e*CH.  The element index is 32-bit; the byte
// offset is formed in 64 bits, so a field may exceed 4 GiB.
__device__ __forceinline__ size_t plane_off(int Vh, int k, unsigned e, size_t ch) {
    return ((size_t)k * (size_t)Vh + (size_t)e) * ch;
}
template <typename T>
__device__ __forceinline__ T ldp(const char* base, int Vh, int k, unsigned e) {
    return *reinterpret_cast<const T*>(base + plane_off(Vh, k, e, sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stp(char* base, int Vh, int k, unsigned e, T v) {
    *reinterpret_cast<T*>(base + plane_off(Vh, k, e, sizeof(T))) = v;
}

// Non-temporal form of ldp (the "nt" cache policy: the line is marked for early eviction), for data
// a kernel reads exactly once, so it does not displace data that is read again (neighbour spinors).
template <typename T> struct NtVec;
template <> struct NtVec<double2> { typedef double V __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float2> { typedef float V __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float4> { typedef float V __attribute__((ext_vector_type(4))); };
template <> struct NtVec<uint2> { typedef unsigned V __attribute__((ext_vector_type(2))); };
template <> struct NtVec<uint4> { typedef unsigned V __attribute__((ext_vector_type(4))); };
template <> struct NtVec<unsigned> { typedef unsigned V; };
template <typename T, bool NT>
__device__ __forceinline__ T ldq(const char* base, int Vh, int k, unsigned e) {
    if constexpr (NT) {
        typedef typename NtVec<T>::V V;
        const V* p = reinterpret_cast<const V*>(base + plane_off(Vh, k, e, sizeof(T)));
        const V v = __builtin_nontemporal_load(p);
        T out;
        __builtin_memcpy(&out, &v, sizeof(T));
        return out;
    } else {
        return ldp<T>(base, Vh, k, e);
    }
}

// store form of ldq: with NT the line is written with the non-temporal policy (streamed towards memory
// and marked for early eviction), for outputs no later read of the same launch wants in the cache
template <typename T, bool NT>
__device__ __forceinline__ void stq(char* base, int Vh, int k, unsigned e, T v) {
    if constexpr (NT) {
        typedef typename NtVec<T>::V V;
        V* p = reinterpret_cast<V*>(base + plane_off(Vh, k, e, sizeof(T)));
        V x;
        __builtin_memcpy(&x, &v, sizeof(T));
        __builtin_nontemporal_store(x, p);
    } else {
        stp<T>(base, Vh, k, e, v);
    }
}

// FP16 storage helpers
__device__ __forceinline__ unsigned pack_h2(float a, float b) {
    __half2 h = __floats2half2_rn(a, b);
    return *reinterpret_cast<unsigned*>(&h);
}
__device__ __forceinline__ float2 unpack_h2(unsigned u) {
    __half2 h = *reinterpret_cast<__half2*>(&u);
    return __half22float2(h);
}

// FP16-storage Wilson sites as flat BLAS sees them: site i of a field of Vh-site parity blocks has
// its three 16-byte chunks at chunk index (3 * (i / Vh) + k) * Vh + i % Vh and its scale at s[i].
__device__ __forceinline__ size_t h16_chunk(unsigned i, unsigned Vh, int k) {
    const unsigned b = i / Vh, h = i - b * Vh;
    return (size_t)(3 * b + k) * Vh + h;
}
__device__ __forceinline__ void h16_unpack8(uint4 a, float s, float* v) {
    float2 f;
    f = unpack_h2(a.x); v[0] = f.x * s; v[1] = f.y * s;
    f = unpack_h2(a.y); v[2] = f.x * s; v[3] = f.y * s;
    f = unpack_h2(a.z); v[4] = f.x * s; v[5] = f.y * s;
    f = unpack_h2(a.w); v[6] = f.x * s; v[7] = f.y * s;
}
// all 24 components of site i
__device__ __forceinline__ void h16_load_site(const uint4* c, const float* s, unsigned Vh, unsigned i, float* v) {
    const float sc = s[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) h16_unpack8(c[h16_chunk(i, Vh, k)], sc, v + 8 * k);
}
// stores site i with a new scale; v is left holding the stored (rounded) values
__device__ __forceinline__ void h16_store_site(uint4* c, float* s, unsigned Vh, unsigned i, float* v) {
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < 24; ++q) m = fmaxf(m, fabsf(v[q]));
    const float inv = m > 0.f ? 1.0f / m : 0.f;
    s[i] = m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint4 a;
        a.x = pack_h2(v[8 * k + 0] * inv, v[8 * k + 1] * inv);
        a.y = pack_h2(v[8 * k + 2] * inv, v[8 * k + 3] * inv);
        a.z = pack_h2(v[8 * k + 4] * inv, v[8 * k + 5] * inv);
        a.w = pack_h2(v[8 * k + 6] * inv, v[8 * k + 7] * inv);
        c[h16_chunk(i, Vh, k)] = a;
        h16_unpack8(a, m, v + 8 * k);
    }
}


// Wilson spinor: 12 complex numbers per site.
//   F64: 12 planes of double2; F32: 6 planes of float4 (two complex); F16: 3 planes of 8
//   halves plus a per-site float scale (max |component|), element index ne for the scale.
template <int PREC> struct WilsonIO;

template <> struct WilsonIO<CLG_F64> {
    static constexpr int NPL = 12;
    static constexpr bool SCALED = false;
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float*, int Vh, unsigned e, unsigned, C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const double2 a = ldp<double2>(b, Vh, k, e);
            v[k] = C2<R>{(R)a.x, (R)a.y};
        }
    }
    template <typename R, bool NT = false>
    __device__ __forceinline__ static void store(char* b, float*, int Vh, unsigned e, unsigned, const C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 12; ++k) stq<double2, NT>(b, Vh, k, e, make_double2((double)v[k].x, (double)v[k].y));
    }
};

template <> struct WilsonIO<CLG_F32> {
    static constexpr int NPL = 6;
    static constexpr bool SCALED = false;
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float*, int Vh, unsigned e, unsigned, C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float4 a = ldp<float4>(b, Vh, k, e);
            v[2 * k] = C2<R>{(R)a.x, (R)a.y};
            v[2 * k + 1] = C2<R>{(R)a.z, (R)a.w};
        }
    }
    template <typename R, bool NT = false>
    __device__ __forceinline__ static void store(char* b, float*, int Vh, unsigned e, unsigned, const C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
            stq<float4, NT>(b, Vh, k, e,
                            make_float4((float)v[2 * k].x, (float)v[2 * k].y, (float)v[2 * k + 1].x, (float)v[2 * k + 1].y));
    }
};

template <> struct WilsonIO<CLG_F16> {
    static constexpr int NPL = 3;
    static constexpr bool SCALED = true;
    // the stored halves without the site scale, which is returned (a hop scales the projected
    // half-spinor instead: 12 products instead of 24)
    template <typename R>
    __device__ __forceinline__ static float load_unscaled(const char* b, const float* nb, int Vh, unsigned e, unsigned ne,
                                                          C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint4 a = ldp<uint4>(b, Vh, k, e);
            float2 f;
            f = unpack_h2(a.x); v[4 * k + 0] = C2<R>{(R)f.x, (R)f.y};
            f = unpack_h2(a.y); v[4 * k + 1] = C2<R>{(R)f.x, (R)f.y};
            f = unpack_h2(a.z); v[4 * k + 2] = C2<R>{(R)f.x, (R)f.y};
            f = unpack_h2(a.w); v[4 * k + 3] = C2<R>{(R)f.x, (R)f.y};
        }
        return nb[ne];
    }
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float* nb, int Vh, unsigned e, unsigned ne, C2<R>* v) {
        const float s = nb[ne];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint4 a = ldp<uint4>(b, Vh, k, e);
            float2 f;
            f = unpack_h2(a.x); v[4 * k + 0] = C2<R>{(R)(f.x * s), (R)(f.y * s)};
            f = unpack_h2(a.y); v[4 * k + 1] = C2<R>{(R)(f.x * s), (R)(f.y * s)};
            f = unpack_h2(a.z); v[4 * k + 2] = C2<R>{(R)(f.x * s), (R)(f.y * s)};
            f = unpack_h2(a.w); v[4 * k + 3] = C2<R>{(R)(f.x * s), (R)(f.y * s)};
        }
    }
    template <typename R, bool NT = false>
    __device__ __forceinline__ static void store(char* b, float* nb, int Vh, unsigned e, unsigned ne, const C2<R>* v) {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 12; ++k) m = fmaxf(m, fmaxf(fabsf((float)v[k].x), fabsf((float)v[k].y)));
        const float inv = m > 0.f ? 1.0f / m : 0.f;
        nb[ne] = m;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint4 a;
            a.x = pack_h2((float)v[4 * k + 0].x * inv, (float)v[4 * k + 0].y * inv);
            a.y = pack_h2((float)v[4 * k + 1].x * inv, (float)v[4 * k + 1].y * inv);
            a.z = pack_h2((float)v[4 * k + 2].x * inv, (float)v[4 * k + 2].y * inv);
            a.w = pack_h2((float)v[4 * k + 3].x * inv, (float)v[4 * k + 3].y * inv);
            stq<uint4, NT>(b, Vh, k, e, a);
        }
    }
};

// Staggered vector: 3 complex per site, one complex per plane (16 / 8 / 4 bytes)
template <int PREC> struct KSIO;
template <> struct KSIO<CLG_F64> {
    static constexpr int NPL = 3;
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float*, int Vh, unsigned e, unsigned, C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { const double2 a = ldp<double2>(b, Vh, k, e); v[k] = C2<R>{(R)a.x, (R)a.y}; }
    }
    template <typename R>
    __device__ __forceinline__ static void store(char* b, float*, int Vh, unsigned e, unsigned, const C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 3; ++k) stp<double2>(b, Vh, k, e, make_double2((double)v[k].x, (double)v[k].y));
    }
};
template <> struct KSIO<CLG_F32> {
    static constexpr int NPL = 3;
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float*, int Vh, unsigned e, unsigned, C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { const float2 a = ldp<float2>(b, Vh, k, e); v[k] = C2<R>{(R)a.x, (R)a.y}; }
    }
    template <typename R>
    __device__ __forceinline__ static void store(char* b, float*, int Vh, unsigned e, unsigned, const C2<R>* v) {
#pragma unroll
        for (int k = 0; k < 3; ++k) stp<float2>(b, Vh, k, e, make_float2((float)v[k].x, (float)v[k].y));
    }
};
template <> struct KSIO<CLG_F16> {
    static constexpr int NPL = 3;
    template <typename R>
    __device__ __forceinline__ static void load(const char* b, const float* nb, int Vh, unsigned e, unsigned ne, C2<R>* v) {
        const float s = nb[ne];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float2 f = unpack_h2(ldp<unsigned>(b, Vh, k, e));
            v[k] = C2<R>{(R)(f.x * s), (R)(f.y * s)};
        }
    }
    template <typename R, bool NT = false>
    __device__ __forceinline__ static void store(char* b, float* nb, int Vh, unsigned e, unsigned ne, const C2<R>* v) {
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) m = fmaxf(m, fmaxf(fabsf((float)v[k].x), fabsf((float)v[k].y)));
        const float inv = m > 0.f ? 1.0f / m : 0.f;
        nb[ne] = m;
#pragma unroll
        for (int k = 0; k < 3; ++k) stp<unsigned>(b, Vh, k, e, pack_h2((float)v[k].x * inv, (float)v[k].y * inv));
    }
};

// Gauge links: 8 blocks (parity, mu) of Vh links each; element index e = (p*4+mu)*NPL*Vh + h.
// 18-real: 9 planes of one complex; 12-real: rows 0,1 only, row 2 = conj(r0 x r1) as
// SU3::Norm builds it (SU3.h:1083-1085).
template <typename R>
__device__ __forceinline__ void rebuild_row2(C2<R>* m) {
    m[6] = cconj(csub(cmul(m[1], m[5]), cmul(m[2], m[4])));
    m[7] = cconj(csub(cmul(m[2], m[3]), cmul(m[0], m[5])));
    m[8] = cconj(csub(cmul(m[0], m[4]), cmul(m[1], m[3])));
}

template <int PREC, int RECON> struct LinkIO;
template <> struct LinkIO<CLG_F64, 18> {
    static constexpr int NPL = 9;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) { const double2 a = ldq<double2, NT>(b, Vh, k, e); m[k] = C2<R>{(R)a.x, (R)a.y}; }
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) stp<double2>(b, Vh, k, e, make_double2(m[k].x, m[k].y));
    }
};
template <> struct LinkIO<CLG_F64, 12> {
    static constexpr int NPL = 6;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 6; ++k) { const double2 a = ldq<double2, NT>(b, Vh, k, e); m[k] = C2<R>{(R)a.x, (R)a.y}; }
        rebuild_row2(m);
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 6; ++k) stp<double2>(b, Vh, k, e, make_double2(m[k].x, m[k].y));
    }
};
template <> struct LinkIO<CLG_F32, 18> {
    static constexpr int NPL = 9;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) { const float2 a = ldq<float2, NT>(b, Vh, k, e); m[k] = C2<R>{(R)a.x, (R)a.y}; }
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) stp<float2>(b, Vh, k, e, make_float2((float)m[k].x, (float)m[k].y));
    }
};
template <> struct LinkIO<CLG_F32, 12> {
    static constexpr int NPL = 3;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float4 a = ldq<float4, NT>(b, Vh, k, e);
            m[2 * k] = C2<R>{(R)a.x, (R)a.y};
            m[2 * k + 1] = C2<R>{(R)a.z, (R)a.w};
        }
        rebuild_row2(m);
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            stp<float4>(b, Vh, k, e,
                        make_float4((float)m[2 * k].x, (float)m[2 * k].y, (float)m[2 * k + 1].x, (float)m[2 * k + 1].y));
    }
};
template <> struct LinkIO<CLG_F16, 18> {
    static constexpr int NPL = 9;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) { const float2 f = unpack_h2(ldq<unsigned, NT>(b, Vh, k, e)); m[k] = C2<R>{(R)f.x, (R)f.y}; }
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 9; ++k) stp<unsigned>(b, Vh, k, e, pack_h2((float)m[k].x, (float)m[k].y));
    }
};
template <> struct LinkIO<CLG_F16, 12> {
    static constexpr int NPL = 3;
    template <bool NT = false, typename R>
    __device__ __forceinline__ static void load(const char* b, int Vh, unsigned e, C2<R>* m) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint2 a = ldq<uint2, NT>(b, Vh, k, e);
            const float2 f0 = unpack_h2(a.x), f1 = unpack_h2(a.y);
            m[2 * k] = C2<R>{(R)f0.x, (R)f0.y};
            m[2 * k + 1] = C2<R>{(R)f1.x, (R)f1.y};
        }
        rebuild_row2(m);
    }
    __device__ __forceinline__ static void store(char* b, int Vh, unsigned e, const cd* m) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            stp<uint2>(b, Vh, k, e,
                       make_uint2(pack_h2((float)m[2 * k].x, (float)m[2 * k].y),
                                  pack_h2((float)m[2 * k + 1].x, (float)m[2 * k + 1].y)));
    }
};

// y = U x and y = U^dagger x on colour 3-vectors
template <typename R>
__device__ __forceinline__ void su3_mv(const C2<R>* u, const C2<R>* x, C2<R>* y) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        C2<R> a = cmul(u[3 * i], x[0]);
        cfma(a, u[3 * i + 1], x[1]);
        cfma(a, u[3 * i + 2], x[2]);
        y[i] = a;
    }
}
template <typename R>
__device__ __forceinline__ void su3_mdv(const C2<R>* u, const C2<R>* x, C2<R>* y) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        C2<R> a = cmulc(u[i], x[0]);
        cfmac(a, u[3 + i], x[1]);
        cfmac(a, u[6 + i], x[2]);
        y[i] = a;
    }
}

// ------------------------------------------------------------------ SU(3) helpers
// r = a b
template <typename R>
__device__ __forceinline__ void su3_mm(const C2<R>* a, const C2<R>* b, C2<R>* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            C2<R> acc = cmul(a[3 * i], b[j]);
            cfma(acc, a[3 * i + 1], b[3 + j]);
            cfma(acc, a[3 * i + 2], b[6 + j]);
            r[3 * i + j] = acc;
        }
}
// traceless anti-Hermitian part (SU3::Ta, SU3.h:953-984)
__device__ __forceinline__ void su3_ta(cd* m) {
    const cd n1 = csub(m[1], cconj(m[3]));
    const cd n2 = csub(m[2], cconj(m[6]));
    const cd n5 = csub(m[5], cconj(m[7]));
    m[1] = cscale(n1, 0.5);
    m[3] = cd{-m[1].x, m[1].y};
    m[2] = cscale(n2, 0.5);
    m[6] = cd{-m[2].x, m[2].y};
    m[5] = cscale(n5, 0.5);
    m[7] = cd{-m[5].x, m[5].y};
    const double tr = (m[0].y + m[4].y + m[8].y) / 3.0;
    m[0] = cd{0.0, m[0].y - tr};
    m[4] = cd{0.0, m[4].y - tr};
    m[8] = cd{0.0, m[8].y - tr};
}
// re-unitarisation as SU3::Norm does it (SU3.h:1058-1088): Gram-Schmidt on rows 0 and 1,
// row 2 = conj(r0 x r1)
__device__ __forceinline__ void su3_reunit(cd* m) {
    const double d1 = 1.0 / sqrt(m[0].x * m[0].x + m[0].y * m[0].y + m[1].x * m[1].x + m[1].y * m[1].y +
                                 m[2].x * m[2].x + m[2].y * m[2].y);
    m[0] = cscale(m[0], d1); m[1] = cscale(m[1], d1); m[2] = cscale(m[2], d1);
    cd sp = cadd(cadd(cmul(m[0], cconj(m[3])), cmul(m[1], cconj(m[4]))), cmul(m[2], cconj(m[5])));
    sp = cconj(sp);
    m[3] = csub(m[3], cmul(sp, m[0])); m[4] = csub(m[4], cmul(sp, m[1])); m[5] = csub(m[5], cmul(sp, m[2]));
    const double d2 = 1.0 / sqrt(m[3].x * m[3].x + m[3].y * m[3].y + m[4].x * m[4].x + m[4].y * m[4].y +
                                 m[5].x * m[5].x + m[5].y * m[5].y);
    m[3] = cscale(m[3], d2); m[4] = cscale(m[4], d2); m[5] = cscale(m[5], d2);
    rebuild_row2(m);
}

// ------------------------------------------------------------------ random numbers
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter = (index, draw, stream, c3),
// key = (seed, 0x434C4721); two uniforms of 53 bits -> one Box-Muller pair of N(0,1)
__device__ __forceinline__ void philox(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                       unsigned k1, unsigned out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
__device__ __forceinline__ double u53(unsigned a, unsigned b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ void gauss2(unsigned idx, unsigned draw, unsigned stream, unsigned seed, double g[2],
                                       unsigned c3 = 0u) {
    unsigned w[4];
    philox(idx, draw, stream, c3, seed, 0x434C4721u, w);
    const double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(6.283185307179586476925286766559 * u2, &s, &c);
    g[0] = r * c;
    g[1] = r * s;
}
enum { STREAM_GAUGE = 0, STREAM_WILSON = 1, STREAM_KS = 2, STREAM_MOMENTUM = 3, STREAM_GAUGE_TRANSFORM = 4 };

// reunitarise(exp(i eps H)) for H traceless Hermitian with Gaussian entries drawn from stream `stream`
// at counter `idx`, by Taylor scaling-and-squaring (the same recipe as the oracle's generator)
__device__ inline void random_su3(unsigned idx, double eps, unsigned seed, unsigned stream, cd* out) {
    double g[10];
    for (int d = 0; d < 5; ++d) gauss2(idx, (unsigned)d, stream, seed, g + 2 * d);
    const double tr = (g[0] + g[1] + g[2]) / 3.0;
    const double is2 = 0.70710678118654752440084436210485;
    cd H[9];
    H[0] = cd{g[0] - tr, 0.0};
    H[4] = cd{g[1] - tr, 0.0};
    H[8] = cd{g[2] - tr, 0.0};
    H[1] = cd{g[3] * is2, g[4] * is2};
    H[2] = cd{g[5] * is2, g[6] * is2};
    H[5] = cd{g[7] * is2, g[8] * is2};
    H[3] = cconj(H[1]);
    H[6] = cconj(H[2]);
    H[7] = cconj(H[5]);
    cd A[9];
    double nrm = 0;
    for (int i = 0; i < 9; ++i) {
        A[i] = cd{-eps * H[i].y, eps * H[i].x};
        nrm += A[i].x * A[i].x + A[i].y * A[i].y;
    }
    nrm = sqrt(nrm);
    int s = 0;
    while (nrm > 0.5) { nrm *= 0.5; ++s; }
    const double f = ldexp(1.0, -s);
    for (int i = 0; i < 9; ++i) A[i] = cscale(A[i], f);
    cd T[9], S[9], tmp[9];
    for (int i = 0; i < 9; ++i) T[i] = cd{0.0, 0.0};
    T[0] = T[4] = T[8] = cd{1.0, 0.0};
    for (int i = 0; i < 9; ++i) S[i] = T[i];
    for (int k = 1; k <= 20; ++k) {
        su3_mm(T, A, tmp);
        for (int i = 0; i < 9; ++i) { T[i] = cscale(tmp[i], 1.0 / k); S[i] = cadd(S[i], T[i]); }
    }
    for (int k = 0; k < s; ++k) {
        su3_mm(S, S, tmp);
        for (int i = 0; i < 9; ++i) S[i] = tmp[i];
    }
    su3_reunit(S);
    for (int i = 0; i < 9; ++i) out[i] = S[i];
}


// ------------------------------------------------------------------ grid reductions
constexpr int kRedBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-level stage of a two-kernel grid reduction: sums NR per-lane FP64 values over the
// workgroup (blockDim.x <= kRedBlock, a multiple of 64) and writes them to
// partial[blockIdx.x * NR + r].  No fence or counter: the finishing kernel launched after it on the
// same stream sees every partial.  (A device-scope release fence per workgroup costs a
// cross-XCD L2 write-back on gfx950: tens of microseconds over a 1024-block grid.)
template <int NR>
__device__ void block_partial(double (&v)[NR], double* partial) {
    __shared__ double sm[kRedBlock / 64][NR];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sm[wv][r] = v[r];
    }
    __syncthreads();
    for (int r = threadIdx.x; r < NR; r += blockDim.x) {
        double s = 0;
        for (int w = 0; w < nw; ++w) s += sm[w][r];
        partial[(size_t)blockIdx.x * NR + r] = s;
    }
}

// Reduce NR per-lane FP64 values over the grid (blockDim.x <= kRedBlock, a multiple of 64);
// the last workgroup to finish writes out[0..NR), resets the counter and returns true on its
// thread 0.  partial holds gridDim.x * NR doubles.
template <int NR>
__device__ bool reduce_finish(double (&v)[NR], double* partial, unsigned* counter, double* out) {
    __shared__ double sm[kRedBlock / 64][NR];
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
    if (!is_last) return false;
    // acquire: the fence invalidates this CU's vector L1, so the plain loads below see every
    // block's partials (each published by a release fence before its counter increment)
    __threadfence();
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += partial[(size_t)b * NR + r];
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
        return true;
    }
    return false;
}

}  // namespace clgd

