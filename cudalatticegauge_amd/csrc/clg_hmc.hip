// clg_hmc.hip -- gauge-field kernels and the HMC trajectory (integrators, energies, Metropolis).
//
// Conventions (CIntegrator / CFieldGaugeSU3): momenta P are traceless anti-Hermitian per link,
// generated as sum_a r_a (i lambda_a / 2)-type generators with r_a ~ N(0, 1/4) in the layout of
// deviceSU3::makeSU3RandomGenerator (SU3.h:193-226); the kinetic energy is sum ReTr P^dagger P;
// P <- P + eps F; U <- exp(eps P) U with the quick exponential (ExponentialPrecision 0) or an
// order-n Taylor polynomial (n > 0); the plaquette action is beta/3 sum (3 - ReTr U_p).
// Gauge, momentum and force fields share the FP64 18-real link layout: 8 (parity, mu) blocks of
// 9 planes of Vh double2.  The gauge links are periodic (fermion boundary signs live in the
// fermion fields).
#include "clg_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int fetch_doubles(clg_context, const double*, int, double*);
int ensure_reduction_scratch(clg_context, size_t);
int wilson_full(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, double, double);
}  // namespace clg

namespace {

const int kBlk = 256;
const int kMaxRedGrid = 2048;

// r = a b^dagger, r = a^dagger b
__device__ __forceinline__ void su3_mmd(const cd* a, const cd* b, cd* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cd acc{0.0, 0.0};
            acc = cadd(acc, cmul(a[3 * i], cconj(b[3 * j])));
            acc = cadd(acc, cmul(a[3 * i + 1], cconj(b[3 * j + 1])));
            acc = cadd(acc, cmul(a[3 * i + 2], cconj(b[3 * j + 2])));
            r[3 * i + j] = acc;
        }
}
__device__ __forceinline__ void su3_mdm(const cd* a, const cd* b, cd* r) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            cd acc = cmulc(a[i], b[j]);
            cfmac(acc, a[3 + i], b[3 + j]);
            cfmac(acc, a[6 + i], b[6 + j]);
            r[3 * i + j] = acc;
        }
}

// (parity, half index) of the site with coordinates c (Lt even: h = lex / 2)
struct PH {
    int p, h;
};
__device__ __forceinline__ PH ph_of(const Geo& g, const int* c) {
    const int lex = ((c[0] * g.L[1] + c[1]) * g.L[2] + c[2]) * g.L[3] + c[3];
    return PH{(c[0] + c[1] + c[2] + c[3]) & 1, lex >> 1};
}
__device__ __forceinline__ void load_link(const char* U, const Geo& g, PH s, int mu, cd* m) {
    LinkIO<CLG_F64, 18>::load(U, g.Vh, (unsigned)((s.p * 4 + mu) * 9 * g.Vh + s.h), m);
}

// SU(2) Cayley map (1 + X)(1 - X)^-1 of X = [[i y, z], [-z*, -i y]] embedded at (i, j), applied
// from the left to m: m <- C m
__device__ __forceinline__ void cayley_left(int i, int j, double y, cd z, cd* m) {
    const double l = y * y + z.x * z.x + z.y * z.y, d = 1.0 / (1.0 + l);
    const cd c{d * (1.0 - l), 2.0 * y * d};
    const cd o{2.0 * d * z.x, 2.0 * d * z.y};
    const cd mo{-o.x, o.y};  // -conj(o)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const cd ri = m[3 * i + k], rj = m[3 * j + k];
        m[3 * i + k] = cadd(cmul(c, ri), cmul(o, rj));
        m[3 * j + k] = cadd(cmul(mo, ri), cmul(cconj(c), rj));
    }
}
// e = exp(a M) as SU3::QuickExp (SU3.h:1361-1640) defines it: the symmetric product
// C01 C02 C12 C02 C01 of SU(2) Cayley maps, exact in SU(3) and exp(a M) + O(a^3)
__device__ __forceinline__ void quick_exp(const cd* M, double a, cd* e) {
    const double y1 = a * (M[0].y - M[4].y) * (1.0 / 12.0);
    const double y2 = a * (M[0].y - M[8].y) * (1.0 / 12.0);
    const double y3 = a * (M[4].y - M[8].y) * (1.0 / 6.0);
    const cd z1 = cscale(M[1], 0.25 * a), z2 = cscale(M[2], 0.25 * a), z3 = cscale(M[5], 0.5 * a);
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = cd{0.0, 0.0};
    e[0] = e[4] = e[8] = cd{1.0, 0.0};
    cayley_left(0, 1, y1, z1, e);
    cayley_left(0, 2, y2, z2, e);
    cayley_left(1, 2, y3, z3, e);
    cayley_left(0, 2, y2, z2, e);
    cayley_left(0, 1, y1, z1, e);
}
// e = sum_{k<=n} (a M)^k / k! in Horner form, SU3::ExpReal (SU3.h:1286-1318)
__device__ __forceinline__ void taylor_exp(const cd* M, double a, int n, cd* e) {
#pragma unroll
    for (int k = 0; k < 9; ++k) e[k] = cd{0.0, 0.0};
    e[0] = e[4] = e[8] = cd{1.0, 0.0};
    for (int i = 0; i < n; ++i) {
        const double f = a / (double)(n - i);
        cd t[9], s[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) s[k] = cscale(M[k], f);
        su3_mm(s, e, t);
#pragma unroll
        for (int k = 0; k < 9; ++k) e[k] = t[k];
        e[0].x += 1.0; e[4].x += 1.0; e[8].x += 1.0;
    }
}

// momentum heat-bath, one thread per (parity, mu, h) link
__global__ __launch_bounds__(kBlk) void k_momentum(Geo g, char* P, unsigned seed, unsigned counter) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 8L * g.Vh) return;
    const int blk = (int)(gid / g.Vh), h = (int)(gid - (long)blk * g.Vh);
    const int p = blk >> 2, mu = blk & 3;
    const Site s = site_of(g, p, h);
    const unsigned link = (unsigned)(s.lex * 4 + mu);
    double r[8];
    for (int d = 0; d < 4; ++d) gauss2(link, (unsigned)d, STREAM_MOMENTUM, seed, r + 2 * d, counter);
    for (int d = 0; d < 8; ++d) r[d] *= 0.5;  // N(0, 1/4), _deviceRandomGaussFSqrt2
    const double is3 = 0.57735026918962576450914878050196, is3x2 = 1.1547005383792515290182975610039;
    cd m[9];
    m[0] = cd{0.0, r[7] * is3 + r[2]};
    m[1] = cd{r[1], r[0]};
    m[2] = cd{r[4], r[3]};
    m[3] = cd{-r[1], r[0]};
    m[4] = cd{0.0, r[7] * is3 - r[2]};
    m[5] = cd{r[6], r[5]};
    m[6] = cd{-r[4], r[3]};
    m[7] = cd{-r[6], r[5]};
    m[8] = cd{0.0, -r[7] * is3x2};
    LinkIO<CLG_F64, 18>::store(P, g.Vh, (unsigned)(blk * 9 * g.Vh + h), m);
}

// flat FP64 element-wise kernels over gauge-shaped fields (n double2 values)
__global__ void k_link_axpy(long n, double2* __restrict__ y, const double2* __restrict__ x, double a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 xv = x[i];
        double2 yv = y[i];
        yv.x += a * xv.x;
        yv.y += a * xv.y;
        y[i] = yv;
    }
}
__global__ void k_link_fill0(long n, double2* __restrict__ y) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = make_double2(0.0, 0.0);
}
__global__ __launch_bounds__(kRedBlock) void k_link_norm2(long n, const double2* __restrict__ x, double* partial,
                                                          unsigned* counter, double* out) {
    double v[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 a = x[i];
        v[0] += a.x * a.x + a.y * a.y;
    }
    reduce_finish<1>(v, partial, counter, out);
}

// U <- exp(a M) U per link (CFieldGaugeSU3::ExpMult, CFieldGaugeSU3.cu:452-490)
__global__ __launch_bounds__(kBlk) void k_exp_mult(long nlink, int Vh, const char* M, char* U, double a, int prec) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nlink) return;
    const int blk = (int)(gid / Vh), h = (int)(gid - (long)blk * Vh);
    const unsigned e = (unsigned)(blk * 9 * Vh + h);
    cd m[9], x[9], u[9], r[9];
    LinkIO<CLG_F64, 18>::load(M, Vh, e, m);
    if (prec == 0) quick_exp(m, a, x);
    else taylor_exp(m, a, prec, x);
    LinkIO<CLG_F64, 18>::load(U, Vh, e, u);
    su3_mm(x, u, r);
    LinkIO<CLG_F64, 18>::store(U, Vh, e, r);
}

// SU3::Norm on every link (CFieldGaugeSU3::ElementNormalize, CFieldGaugeSU3.cu:518-528)
__global__ __launch_bounds__(kBlk) void k_normalize(long nlink, int Vh, char* U) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nlink) return;
    const int blk = (int)(gid / Vh), h = (int)(gid - (long)blk * Vh);
    const unsigned e = (unsigned)(blk * 9 * Vh + h);
    cd u[9];
    LinkIO<CLG_F64, 18>::load(U, Vh, e, u);
    su3_reunit(u);
    LinkIO<CLG_F64, 18>::store(U, Vh, e, u);
}

// sum over sites and mu < nu of (3 - ReTr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag)
// (CFieldGaugeSU3.cu:308-371)
__global__ __launch_bounds__(kRedBlock) void k_plaquette(Geo g, const char* U, double* partial, unsigned* counter,
                                                         double* out) {
    double v[1] = {0.0};
    for (long gid = (long)blockIdx.x * blockDim.x + threadIdx.x; gid < 2L * g.Vh;
         gid += (long)gridDim.x * blockDim.x) {
        const int p = (int)(gid / g.Vh), h = (int)(gid - (long)p * g.Vh);
        const Site s = site_of(g, p, h);
        const PH x{p, h};
        for (int mu = 0; mu < 4; ++mu) {
            int cm[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
            cm[mu] = cm[mu] + 1 == g.L[mu] ? 0 : cm[mu] + 1;
            const PH xm = ph_of(g, cm);
            cd umu[9];
            load_link(U, g, x, mu, umu);
            for (int nu = mu + 1; nu < 4; ++nu) {
                int cn[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
                cn[nu] = cn[nu] + 1 == g.L[nu] ? 0 : cn[nu] + 1;
                const PH xn = ph_of(g, cn);
                cd a[9], b[9], t1[9], t2[9];
                load_link(U, g, xm, nu, a);
                su3_mm(umu, a, t1);  // U_mu(x) U_nu(x+mu)
                load_link(U, g, x, nu, a);
                load_link(U, g, xn, mu, b);
                su3_mm(a, b, t2);  // U_nu(x) U_mu(x+nu)
                double tr = 0.0;
                for (int k = 0; k < 9; ++k) tr += t1[k].x * t2[k].x + t1[k].y * t2[k].y;  // ReTr t1 t2^dag
                v[0] += 3.0 - tr;
            }
        }
    }
    reduce_finish<1>(v, partial, counter, out);
}

// F_mu(x) += coef Ta[U_mu(x) C_mu(x)], C = sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
//                                            + U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
// (_kernelStapleAtSiteSU3CacheIndex + CActionGaugePlaquette force, CFieldGaugeSU3.cu:194-257)
__global__ __launch_bounds__(kBlk) void k_gauge_force(Geo g, const char* U, char* F, double coef) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 8L * g.Vh) return;
    const int blk = (int)(gid / g.Vh), h = (int)(gid - (long)blk * g.Vh);
    const int p = blk >> 2, mu = blk & 3;
    const Site s = site_of(g, p, h);
    const PH x{p, h};
    int cpm[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
    cpm[mu] = cpm[mu] + 1 == g.L[mu] ? 0 : cpm[mu] + 1;
    const PH xpm = ph_of(g, cpm);
    cd C[9];
    for (int k = 0; k < 9; ++k) C[k] = cd{0.0, 0.0};
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        int cpn[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
        cpn[nu] = cpn[nu] + 1 == g.L[nu] ? 0 : cpn[nu] + 1;
        int cmn[4] = {s.c[0], s.c[1], s.c[2], s.c[3]};
        cmn[nu] = cmn[nu] == 0 ? g.L[nu] - 1 : cmn[nu] - 1;
        int cpmmn[4] = {cpm[0], cpm[1], cpm[2], cpm[3]};
        cpmmn[nu] = cpmmn[nu] == 0 ? g.L[nu] - 1 : cpmmn[nu] - 1;
        const PH xpn = ph_of(g, cpn), xmn = ph_of(g, cmn), xpmmn = ph_of(g, cpmmn);
        cd a[9], b[9], t[9], r[9];
        // upper staple
        load_link(U, g, xpm, nu, a);
        load_link(U, g, xpn, mu, b);
        su3_mmd(a, b, t);
        load_link(U, g, x, nu, a);
        su3_mmd(t, a, r);
        for (int k = 0; k < 9; ++k) C[k] = cadd(C[k], r[k]);
        // lower staple: (U_mu(x-nu) U_nu(x+mu-nu))^dag U_nu(x-nu)
        load_link(U, g, xmn, mu, a);
        load_link(U, g, xpmmn, nu, b);
        su3_mm(a, b, t);
        load_link(U, g, xmn, nu, a);
        su3_mdm(t, a, r);
        for (int k = 0; k < 9; ++k) C[k] = cadd(C[k], r[k]);
    }
    cd u[9], f[9];
    load_link(U, g, x, mu, u);
    su3_mm(u, C, f);
    su3_ta(f);
    const unsigned e = (unsigned)(blk * 9 * g.Vh + h);
    cd cur[9];
    LinkIO<CLG_F64, 18>::load(F, g.Vh, e, cur);
    for (int k = 0; k < 9; ++k) cur[k] = cd{cur[k].x + coef * f[k].x, cur[k].y + coef * f[k].y};
    LinkIO<CLG_F64, 18>::store(F, g.Vh, e, cur);
}

int check_link_field(const clg_field_s* f, const char* what) {
    CLG_CHECK_ARG(f, "null %s field", what);
    CLG_CHECK_ARG(f->kind == CLG_GAUGE && f->prec == CLG_F64, "%s must be an FP64 18-real gauge-shaped field", what);
    return CLG_OK;
}
long link_values(const clg_field_s* f) { return 8L * 9L * f->Vh; }  // double2 per gauge-shaped field
unsigned ew_grid(clg_context ctx, long n) {
    const long g = (n + kBlk - 1) / kBlk, cap = (long)ctx->num_cu * 16;
    return (unsigned)std::max<long>(1, std::min(g, cap));
}
unsigned red_grid(clg_context ctx, long n) {
    const long g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    const long cap = std::min<long>(kMaxRedGrid, (long)ctx->num_cu * 4);
    return (unsigned)std::max<long>(1, std::min(g, cap));
}

}  // namespace

// ================================================================ host building blocks
namespace clg {

int link_axpy(clg_context ctx, clg_field_s* y, double a, const clg_field_s* x) {
    const long n = link_values(y);
    k_link_axpy<<<ew_grid(ctx, n), kBlk, 0, ctx->stream>>>(n, (double2*)y->data, (const double2*)x->data, a);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
int link_zero(clg_context ctx, clg_field_s* y) {
    const long n = link_values(y);
    k_link_fill0<<<ew_grid(ctx, n), kBlk, 0, ctx->stream>>>(n, (double2*)y->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
int link_norm2_dev(clg_context ctx, const clg_field_s* x, double* out) {
    const long n = link_values(x);
    const unsigned g = red_grid(ctx, n);
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)g));
    k_link_norm2<<<g, kRedBlock, 0, ctx->stream>>>(n, (const double2*)x->data, ctx->red_partial, ctx->red_counter,
                                                    out);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
int exp_mult(clg_context ctx, const clg_field_s* M, clg_field_s* U, double a, int prec) {
    const long n = 8L * U->Vh;
    k_exp_mult<<<grid_for(n, kBlk), kBlk, 0, ctx->stream>>>(n, (int)U->Vh, (const char*)M->data, (char*)U->data, a,
                                                            prec);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
int gauge_force(clg_context ctx, clg_field_s* F, const clg_field_s* U, double beta) {
    const Geo g = make_geo(U);
    k_gauge_force<<<grid_for(8L * g.Vh, kBlk), kBlk, 0, ctx->stream>>>(g, (const char*)U->data, (char*)F->data,
                                                                        -beta / 6.0);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}
int plaquette_dev(clg_context ctx, const clg_field_s* U, double* out) {
    const Geo g = make_geo(U);
    const unsigned grid = red_grid(ctx, 2L * g.Vh * 6);
    CLG_TRY(ensure_reduction_scratch(ctx, (size_t)grid));
    k_plaquette<<<grid, kRedBlock, 0, ctx->stream>>>(g, (const char*)U->data, ctx->red_partial, ctx->red_counter,
                                                     out);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // namespace clg

namespace {

uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int check_actions(const clg_field_s* gauge, const clg_action* acts, int nact) {
    CLG_CHECK_ARG(acts && nact >= 1 && nact <= CLG_HMC_MAX_ACTIONS, "need 1..%d actions", CLG_HMC_MAX_ACTIONS);
    for (int i = 0; i < nact; ++i) {
        const clg_action& a = acts[i];
        if (a.kind == CLG_ACTION_GAUGE_PLAQUETTE) continue;
        CLG_CHECK_ARG(a.kind == CLG_ACTION_FERMION_WILSON_NF2 || a.kind == CLG_ACTION_FERMION_KS,
                      "action %d: unknown kind %d", i, a.kind);
        const clg_field_s* phi = a.phi;
        CLG_CHECK_ARG(phi, "action %d: null pseudofermion field", i);
        CLG_CHECK_ARG(phi->prec == CLG_F64 && phi->subset == CLG_FULL, "action %d: phi must be a FULL FP64 field", i);
        CLG_CHECK_ARG(phi->kind == (a.kind == CLG_ACTION_FERMION_KS ? CLG_KS : CLG_WILSON),
                      "action %d: phi kind does not match the action", i);
        CLG_TRY(check_same_lattice(phi, gauge));
        if (a.kind == CLG_ACTION_FERMION_KS) {
            CLG_CHECK_ARG(a.md && a.md_degree >= 1 && a.md_degree <= 32, "action %d: MD rational needs 1..32 poles", i);
        }
    }
    return CLG_OK;
}

void add_report(clg_hmc_report* rep, const clg_solve_report& r) {
    if (rep) rep->solver_iterations += r.cycles;
}

// One integrator pass over the trajectory.  Mirrors CIntegrator (CIntegrator.cpp:190-357) and
// CIntegratorNestedForceGradient (CIntegratorNestedForceGradient.cpp:26-157).
struct Integrator {
    clg_context ctx;
    clg_field_s *U, *P, *F, *Ubak;
    const clg_hmc_params* hp;
    const clg_action* acts;
    int nact;
    std::vector<clg_field_s*> xcache;  // Wilson: last (DD^dagger)^-1 phi, the next solve's start
    clg_hmc_report* rep;

    int force_into(bool fermion) {
        for (int i = 0; i < nact; ++i) {
            const clg_action& a = acts[i];
            if (!fermion && a.kind == CLG_ACTION_GAUGE_PLAQUETTE) CLG_TRY(gauge_force(ctx, F, U, a.beta));
            if (!fermion) continue;
            clg_solve_report r{};
            int st = CLG_OK;
            if (a.kind == CLG_ACTION_FERMION_WILSON_NF2) {
                st = clg_wilson_force(ctx, &a.solver, F, a.phi, U, a.kappa_or_mass, xcache[i], xcache[i], &r);
            } else if (a.kind == CLG_ACTION_FERMION_KS) {
                st = clg_ks_force(ctx, &a.solver, a.ms_kind, a.ms_prec, F, a.phi, U, a.kappa_or_mass, a.md, a.md_degree,
                                  &r);
            }
            if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
            add_report(rep, r);
            if (rep && !r.converged && a.kind != CLG_ACTION_GAUGE_PLAQUETTE) rep->unconverged_solves++;
        }
        if (rep) rep->force_evaluations++;
        return CLG_OK;
    }
    bool has(bool fermion) const {
        for (int i = 0; i < nact; ++i)
            if ((acts[i].kind != CLG_ACTION_GAUGE_PLAQUETTE) == fermion) return true;
        return false;
    }
    // UpdatePF / UpdatePG: P += step F(fermion or gauge actions)
    int update_p(double step, bool fermion) {
        CLG_TRY(link_zero(ctx, F));
        CLG_TRY(force_into(fermion));
        return link_axpy(ctx, P, step, F);
    }
    // all actions in one force (CIntegrator::UpdateP, CIntegrator.cpp:219-242)
    int update_p_all(double step) {
        CLG_TRY(link_zero(ctx, F));
        CLG_TRY(force_into(false));
        CLG_TRY(force_into(true));
        return link_axpy(ctx, P, step, F);
    }
    // ZeroForce, force, PreserveFields, AddForceToFieldDirectly(fg), Update(step), RecoverFields
    int fg_update(double fg, double step, bool fermion) {
        CLG_TRY(link_zero(ctx, F));
        CLG_TRY(force_into(fermion));
        CLG_TRY(clg_field_copy(Ubak, U));
        CLG_TRY(exp_mult(ctx, F, U, fg, hp->exp_precision));
        CLG_TRY(update_p(step, fermion));
        return clg_field_copy(U, Ubak);
    }
    int update_u(double step) { return exp_mult(ctx, P, U, step, hp->exp_precision); }

    int nested_fg(double h, int n) {
        CLG_TRY(update_p(h / 6.0, false));
        for (int s = 1; s <= n; ++s) {
            CLG_TRY(update_u(0.5 * h));
            CLG_TRY(fg_update(h * h / 24.0, 2.0 * h / 3.0, false));
            CLG_TRY(update_u(0.5 * h));
            CLG_TRY(update_p(s < n ? h / 3.0 : h / 6.0, false));
        }
        return CLG_OK;
    }
    int nested_leapfrog(double h, int n) {
        CLG_TRY(update_p(0.5 * h, false));
        for (int s = 1; s <= n; ++s) {
            CLG_TRY(update_u(h));
            CLG_TRY(update_p(s < n ? h : 0.5 * h, false));
        }
        return CLG_OK;
    }
    int evaluate() {
        const double eps = hp->step_length / hp->steps;
        if (hp->integrator == CLG_INT_LEAPFROG) {
            CLG_TRY(update_p_all(0.5 * eps));
            for (int s = 1; s <= hp->steps; ++s) {
                CLG_TRY(update_u(eps));
                CLG_TRY(update_p_all(s < hp->steps ? eps : 0.5 * eps));
            }
        } else {
            const double hn = 0.5 * eps / hp->nested_steps;
            auto inner = [&]() {
                return hp->inner_leapfrog ? nested_leapfrog(hn, hp->nested_steps) : nested_fg(hn, hp->nested_steps);
            };
            CLG_TRY(update_p(eps / 6.0, true));
            for (int s = 1; s <= hp->steps; ++s) {
                CLG_TRY(inner());
                CLG_TRY(fg_update(eps * eps / 24.0, 2.0 * eps / 3.0, true));
                CLG_TRY(inner());
                CLG_TRY(update_p(s < hp->steps ? eps / 3.0 : eps / 6.0, true));
            }
        }
        // FinishEvaluate: re-unitarise
        const long n = 8L * U->Vh;
        k_normalize<<<grid_for(n, kBlk), kBlk, 0, ctx->stream>>>(n, (int)U->Vh, (char*)U->data);
        CLG_HIP(hipGetLastError());
        return CLG_OK;
    }
    // CIntegrator::GetEnergy (CIntegrator.cpp:256-279): kinetic + every action, FP64
    int energy(double* total, double* kin, double* per_action) {
        double* dv = ctx->red_result + 40;
        CLG_TRY(link_norm2_dev(ctx, P, dv));
        CLG_TRY(fetch_doubles(ctx, dv, 1, kin));
        double sum = *kin;
        for (int i = 0; i < nact; ++i) {
            const clg_action& a = acts[i];
            double e = 0.0;
            clg_solve_report r{};
            int st = CLG_OK;
            clg_solver_params ep = a.solver;
            if (a.energy_accuracy > 0) ep.accuracy = a.energy_accuracy;
            if (a.kind == CLG_ACTION_GAUGE_PLAQUETTE) {
                CLG_TRY(plaquette_dev(ctx, U, dv));
                CLG_TRY(fetch_doubles(ctx, dv, 1, &e));
                e *= a.beta / 3.0;
            } else if (a.kind == CLG_ACTION_FERMION_WILSON_NF2) {
                st = clg_wilson_pf_energy(ctx, &ep, a.phi, U, a.kappa_or_mass, &e, &r);
            } else {
                // energies enter H_before - H_after, so their shift solves stay FP64 whatever ms_prec
                st = clg_ks_pf_energy(ctx, &ep, a.ms_kind, CLG_F64, a.phi, U, a.kappa_or_mass, a.md,
                                      a.md_degree, &e, &r);
            }
            if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
            add_report(rep, r);
            if (per_action) per_action[i] = e;
            sum += e;
        }
        *total = sum;
        return CLG_OK;
    }
};

int make_integrator(clg_context ctx, clg_field_s* gauge, clg_field_s* mom, const clg_hmc_params* hp,
                    const clg_action* acts, int nact, clg_hmc_report* rep, Integrator& it) {
    CLG_CHECK_ARG(ctx && hp, "null argument");
    CLG_TRY(check_link_field(gauge, "gauge"));
    CLG_TRY(check_link_field(mom, "momentum"));
    CLG_TRY(check_same_lattice(gauge, mom));
    CLG_TRY(check_actions(gauge, acts, nact));
    CLG_CHECK_ARG(hp->integrator == CLG_INT_LEAPFROG || hp->integrator == CLG_INT_NESTED_FORCE_GRADIENT,
                  "unknown integrator %d", hp->integrator);
    CLG_CHECK_ARG(hp->steps >= 1 && hp->nested_steps >= 1, "IntegratorStep and NestedStep must be >= 1");
    CLG_CHECK_ARG(hp->exp_precision >= 0 && hp->exp_precision <= 64, "exp precision must be in [0, 64]");
    CLG_CHECK_ARG(hp->step_length > 0, "trajectory length must be positive");
    it.ctx = ctx;
    it.U = gauge;
    it.P = mom;
    it.hp = hp;
    it.acts = acts;
    it.nact = nact;
    it.rep = rep;
    CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE, CLG_F64, CLG_FULL, 11, &it.F));
    CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE, CLG_F64, CLG_FULL, 12, &it.Ubak));
    it.xcache.assign(nact, nullptr);
    for (int i = 0; i < nact; ++i) {
        if (acts[i].kind != CLG_ACTION_FERMION_WILSON_NF2) continue;
        CLG_TRY(scratch_field(ctx, acts[i].phi, CLG_WILSON, CLG_F64, CLG_FULL, 8000 + i, &it.xcache[i]));
        CLG_TRY(clg_field_copy(it.xcache[i], acts[i].phi));  // start as PrepareForHMCS caches phi
    }
    return CLG_OK;
}

int evolve(Integrator& it, clg_hmc_report* rep) {
    double h0 = 0, h1 = 0, k0 = 0, k1 = 0;
    double e0[CLG_HMC_MAX_ACTIONS] = {0}, e1[CLG_HMC_MAX_ACTIONS] = {0};
    CLG_TRY(it.energy(&h0, &k0, e0));
    CLG_TRY(it.evaluate());
    CLG_TRY(it.energy(&h1, &k1, e1));
    if (rep) {
        rep->h_before = h0;
        rep->h_after = h1;
        rep->h_diff = h0 - h1;
        rep->kinetic_before = k0;
        rep->kinetic_after = k1;
        for (int i = 0; i < it.nact; ++i) {
            rep->action_before[i] = e0[i];
            rep->action_after[i] = e1[i];
        }
    }
    return CLG_OK;
}

}  // namespace

// ================================================================ C ABI
extern "C" {

int clg_hmc_params_default(clg_hmc_params* p) {
    CLG_CHECK_ARG(p, "null argument");
    p->integrator = CLG_INT_NESTED_FORCE_GRADIENT;
    p->step_length = 1.0;
    p->steps = 8;
    p->nested_steps = 3;
    p->inner_leapfrog = 0;
    p->exp_precision = 0;
    p->metropolis = 1;
    return CLG_OK;
}

int clg_gauge_momentum_random(clg_field P, unsigned seed, unsigned counter) {
    CLG_TRY(check_link_field(P, "momentum"));
    (void)hipSetDevice(P->ctx->device);
    const Geo g = make_geo(P);
    k_momentum<<<grid_for(8L * g.Vh, kBlk), kBlk, 0, P->ctx->stream>>>(g, (char*)P->data, seed, counter);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_gauge_kinetic_energy(clg_context ctx, clg_field P, double* out) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    CLG_TRY(check_link_field(P, "momentum"));
    (void)hipSetDevice(ctx->device);
    CLG_TRY(link_norm2_dev(ctx, P, ctx->red_result + 40));
    return fetch_doubles(ctx, ctx->red_result + 40, 1, out);
}

int clg_gauge_axpy(clg_context ctx, clg_field y, double a, clg_field x) {
    CLG_CHECK_ARG(ctx, "null argument");
    CLG_TRY(check_link_field(y, "y"));
    CLG_TRY(check_link_field(x, "x"));
    CLG_TRY(check_same_lattice(x, y));
    (void)hipSetDevice(ctx->device);
    return link_axpy(ctx, y, a, x);
}

int clg_gauge_exp_mult(clg_context ctx, clg_field P, clg_field gauge, double a, int exp_precision) {
    CLG_CHECK_ARG(ctx, "null argument");
    CLG_TRY(check_link_field(P, "momentum"));
    CLG_TRY(check_link_field(gauge, "gauge"));
    CLG_TRY(check_same_lattice(P, gauge));
    CLG_CHECK_ARG(exp_precision >= 0 && exp_precision <= 64, "exp precision must be in [0, 64]");
    CLG_CHECK_ARG(P != gauge, "P and gauge must differ");
    (void)hipSetDevice(ctx->device);
    return exp_mult(ctx, P, gauge, a, exp_precision);
}

int clg_gauge_normalize(clg_context ctx, clg_field gauge) {
    CLG_CHECK_ARG(ctx, "null argument");
    CLG_TRY(check_link_field(gauge, "gauge"));
    (void)hipSetDevice(ctx->device);
    const long n = 8L * gauge->Vh;
    k_normalize<<<grid_for(n, kBlk), kBlk, 0, ctx->stream>>>(n, (int)gauge->Vh, (char*)gauge->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_gauge_plaquette_energy(clg_context ctx, clg_field gauge, double beta, double* out) {
    CLG_CHECK_ARG(ctx && out, "null argument");
    CLG_TRY(check_link_field(gauge, "gauge"));
    (void)hipSetDevice(ctx->device);
    CLG_TRY(plaquette_dev(ctx, gauge, ctx->red_result + 40));
    double s = 0;
    CLG_TRY(fetch_doubles(ctx, ctx->red_result + 40, 1, &s));
    *out = beta / 3.0 * s;
    return CLG_OK;
}

int clg_gauge_force(clg_context ctx, clg_field force, clg_field gauge, double beta) {
    CLG_CHECK_ARG(ctx, "null argument");
    CLG_TRY(check_link_field(force, "force"));
    CLG_TRY(check_link_field(gauge, "gauge"));
    CLG_TRY(check_same_lattice(force, gauge));
    CLG_CHECK_ARG(force != gauge, "force and gauge must differ");
    (void)hipSetDevice(ctx->device);
    return gauge_force(ctx, force, gauge, beta);
}

// CFieldFermionWilsonSquareSU3::PrepareForHMCS (CFieldFermionWilsonSquareSU3.cu:867-906): phi = D eta
int clg_wilson_prepare_hmc(clg_context ctx, clg_field phi, clg_field gauge, double kappa, unsigned seed) {
    CLG_CHECK_ARG(ctx && phi && gauge, "null argument");
    CLG_CHECK_ARG(phi->kind == CLG_WILSON && phi->prec == CLG_F64 && phi->subset == CLG_FULL,
                  "phi must be a FULL FP64 Wilson field");
    (void)hipSetDevice(ctx->device);
    clg_field_s* eta;
    CLG_TRY(scratch_field(ctx, phi, CLG_WILSON, CLG_F64, CLG_FULL, 204, &eta));
    CLG_TRY(clg_field_gaussian(eta, seed));
    return wilson_full(ctx, phi, eta, gauge, kappa, false, 1.0, 0.0);
}

// CFieldFermionKSSU3::PrepareForHMC (CFieldFermionKSSU3.cu:391-407): phi = r_MC(DD^dagger) eta
int clg_ks_prepare_hmc(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field phi,
                       clg_field gauge, double two_am, const double* mc, int degree, unsigned seed,
                       clg_solve_report* report) {
    CLG_CHECK_ARG(ctx && params && phi && gauge && mc, "null argument");
    CLG_CHECK_ARG(phi->kind == CLG_KS && phi->prec == CLG_F64 && phi->subset == CLG_FULL,
                  "phi must be a FULL FP64 staggered field");
    (void)hipSetDevice(ctx->device);
    CLG_TRY(clg_field_gaussian(phi, seed));
    return clg_rational_apply(ctx, params, ms_kind, prec, phi, phi, gauge, two_am, EFO_F_DDdagger, mc, degree, report);
}

int clg_hmc_evolve(clg_context ctx, clg_field gauge, clg_field momentum, const clg_hmc_params* hp,
                   const clg_action* actions, int nactions, clg_hmc_report* report) {
    CLG_CHECK_ARG(ctx, "null argument");
    (void)hipSetDevice(ctx->device);
    if (report) std::memset(report, 0, sizeof(*report));
    const auto t0 = std::chrono::steady_clock::now();
    Integrator it;
    CLG_TRY(make_integrator(ctx, gauge, momentum, hp, actions, nactions, report, it));
    CLG_TRY(evolve(it, report));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    if (report) {
        report->accepted = 1;
        report->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return CLG_OK;
}

// CHMC::Update for one trajectory (CHMC.cpp:76-219): heat-bath of every fermion action
// (PrepareForHMC), momentum noise, H before, integrator, H after, Metropolis on
// exp(H_before - H_after), and the gauge field restored when rejected.
int clg_hmc_trajectory(clg_context ctx, clg_field gauge, const clg_hmc_params* hp, clg_action* actions,
                       int nactions, unsigned long long seed, clg_hmc_report* report) {
    CLG_CHECK_ARG(ctx && hp, "null argument");
    CLG_TRY(check_link_field(gauge, "gauge"));
    CLG_TRY(check_actions(gauge, actions, nactions));
    (void)hipSetDevice(ctx->device);
    clg_hmc_report local;
    clg_hmc_report* rep = report ? report : &local;
    std::memset(rep, 0, sizeof(*rep));
    const auto t0 = std::chrono::steady_clock::now();
    clg_field_s *mom, *U0;
    CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE, CLG_F64, CLG_FULL, 13, &mom));
    CLG_TRY(scratch_field(ctx, gauge, CLG_GAUGE, CLG_F64, CLG_FULL, 14, &U0));
    CLG_TRY(clg_field_copy(U0, gauge));
    // from here on, any failure puts the caller's gauge field back before returning its status
    // (the rejected-trajectory rule, applied to errors as well)
    auto restore = [&](int status) -> int {
        (void)clg_field_copy(gauge, U0);
        (void)hipStreamSynchronize(ctx->stream);
        return status;
    };
    const uint64_t s0 = mix64(seed);
    for (int i = 0; i < nactions; ++i) {
        clg_action& a = actions[i];
        const unsigned si = (unsigned)mix64(s0 + 1 + (uint64_t)i);
        if (a.kind == CLG_ACTION_FERMION_WILSON_NF2) {
            const int st = clg_wilson_prepare_hmc(ctx, a.phi, gauge, a.kappa_or_mass, si);
            if (st != CLG_OK) return restore(st);
        } else if (a.kind == CLG_ACTION_FERMION_KS) {
            CLG_CHECK_ARG(a.mc && a.mc_degree >= 1 && a.mc_degree <= 32, "action %d: MC rational needs 1..32 poles", i);
            clg_solve_report r{};
            // the heat bath draws phi = r_MC(D D^dagger) eta; it keeps full-precision links even when the
            // force solves take FP16 ones, so phi follows the action the energies measure
            const int hb_prec = a.ms_prec == CLG_F16 ? CLG_F32 : a.ms_prec;
            const int st = clg_ks_prepare_hmc(ctx, &a.solver, a.ms_kind, hb_prec, a.phi, gauge, a.kappa_or_mass, a.mc,
                                              a.mc_degree, si, &r);
            if (st != CLG_OK && st != CLG_NOT_CONVERGED) return restore(st);
            add_report(rep, r);
        }
    }
    {
        int st = clg_gauge_momentum_random(mom, (unsigned)s0, (unsigned)(s0 >> 32));
        if (st != CLG_OK) return restore(st);
    }
    Integrator it;
    {
        int st = make_integrator(ctx, gauge, mom, hp, actions, nactions, rep, it);
        if (st == CLG_OK) st = evolve(it, rep);
        if (st != CLG_OK) return restore(st);
    }
    const double dh = rep->h_diff;
    int accepted = 1;
    if (hp->metropolis) {
        const uint64_t r = mix64(s0 ^ 0xA5A5A5A5DEADBEEFull);
        const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0);
        accepted = u <= std::exp(dh) ? 1 : 0;
    }
    if (!accepted) CLG_TRY(clg_field_copy(gauge, U0));
    CLG_HIP(hipStreamSynchronize(ctx->stream));
    rep->accepted = accepted;
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CLG_OK;
}

}  // extern "C"
