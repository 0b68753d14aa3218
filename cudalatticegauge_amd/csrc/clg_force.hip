// clg_force.hip -- HMC fermion forces and pseudofermion energies.
//
// Conventions (reference CIntegrator / CFieldGaugeSU3): momenta P are traceless anti-Hermitian,
// the kinetic energy is sum ReTr P^dagger P, P <- P + eps F and U <- exp(eps P) U.  A force F is
// therefore consistent with an action S when dS/deps = -2 Re sum Tr(F^dagger T) along U -> exp(eps T) U,
// which tests/test_oracle_forces.py checks by finite differences on the oracle.
//
// Wilson (CFieldFermionWilsonSquareSU3.cu:170-231, driver :1016-1100):
//   X = (D D^dagger)^-1 phi, Y = D^dagger X,
//   F_mu(x) += s kappa Ta[ U_mu(x) ( sum_s ((1-g_mu) Y(x+mu))_s X(x)_s^dagger + (X(x+mu))_s ((1+g_mu) Y(x))_s^dagger ) ]
// with s the boundary sign of the forward hop.
// Staggered (CFieldFermionKSSU3.cu:216-264, driver :412-497), over the MD rational poles:
//   phi_i = (D D^dagger + b_i)^-1 phi, d_i = D0 phi_i (massless),
//   F_mu(x) -= Ta[ a_i eta_mu(x) s ( (U phi_i(x+mu)) d_i(x)^dagger + phi_i(x) (U d_i(x+mu))^dagger ) ]
#include "clg_internal.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace clg;
using namespace clgd;

namespace clg {
int blas_dot_dev(clg_context, const clg_field_s*, const clg_field_s*, double*);
int blas_norm2_dev(clg_context, const clg_field_s*, double*);
int fetch_doubles(clg_context, const double*, int, double*);
int wilson_full(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, double, double);
int ks_full(clg_context, clg_field_s*, const clg_field_s*, const clg_field_s*, double, bool, double, double);
}  // namespace clg

namespace {

typedef C2<double> cx;

// gamma_mu rows (chiral set, GammaMatrix.h:447-475): (g psi)[r] = i^z[r] psi[idx[r]]
__constant__ int kGIdx[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}};
__constant__ int kGZ4[4][4] = {{3, 3, 1, 1}, {2, 0, 0, 2}, {3, 1, 1, 3}, {0, 0, 0, 0}};
__device__ __forceinline__ cx mulz4r(cx a, int k) {
    switch (k & 3) {
        case 1: return cx{-a.y, a.x};
        case 2: return cx{-a.x, -a.y};
        case 3: return cx{a.y, -a.x};
        default: return a;
    }
}
// out = (1 + sg gamma_mu) psi   (sg = +1 or -1)
__device__ __forceinline__ void proj_full(const cx* psi, int mu, int sg, cx* out) {
    for (int r = 0; r < 4; ++r) {
        const int src = kGIdx[mu][r];
        const int z = kGZ4[mu][r] + (sg < 0 ? 2 : 0);
        for (int c = 0; c < 3; ++c) out[3 * r + c] = cadd(psi[3 * r + c], mulz4r(psi[3 * src + c], z));
    }
}

__device__ __forceinline__ void add_link(char* F, int Vh, unsigned e, const cx* m, double f) {
    cx cur[9];
    LinkIO<CLG_F64, 18>::load(F, Vh, e, cur);
    for (int k = 0; k < 9; ++k) cur[k] = cx{cur[k].x + f * m[k].x, cur[k].y + f * m[k].y};
    LinkIO<CLG_F64, 18>::store(F, Vh, e, cur);
}

// Launch shape of the force kernels: 128-thread workgroups (the register budget that fits their
// per-site state without spilling), XCD-contiguous slabs with the parities interleaved block by
// block, so the x+mu neighbour reads of one parity meet the site reads of the other in one L2.
constexpr int kForceBlock = 128;

// Wilson force kernel; X = (DD^dagger)^-1 phi ("left"), Y = D^dagger X ("right"), FULL FP64 fields
__global__ __launch_bounds__(kForceBlock) void k_wilson_force(Geo g, const char* X, const char* Y, const char* U, char* F,
                                                              double kappa) {
    int p, h;
    if (!map_site(g.Vh, 2, 0, 8, 1, p, h)) return;
    const Site s = site_of(g, p, h);
    const int Vh = g.Vh;
    cx xl[12], xr[12];
    WilsonIO<CLG_F64>::load(X, nullptr, Vh, (unsigned)(p * 12 * Vh + h), 0u, xl);
    WilsonIO<CLG_F64>::load(Y, nullptr, Vh, (unsigned)(p * 12 * Vh + h), 0u, xr);
    for (int mu = 0; mu < 4; ++mu) {
        Nbr n;
        switch (mu) {
            case 0: n = fwd_nbr<0>(g, s, p); break;
            case 1: n = fwd_nbr<1>(g, s, p); break;
            case 2: n = fwd_nbr<2>(g, s, p); break;
            default: n = fwd_nbr<3>(g, s, p); break;
        }
        cx pr[12], pl[12], r1[12], r2[12];
        WilsonIO<CLG_F64>::load(Y, nullptr, Vh, (unsigned)(n.p * 12 * Vh + n.h), 0u, pr);
        WilsonIO<CLG_F64>::load(X, nullptr, Vh, (unsigned)(n.p * 12 * Vh + n.h), 0u, pl);
        proj_full(pr, mu, -1, r1);  // (1 - g_mu) Y(x+mu)
        proj_full(xr, mu, +1, r2);  // (1 + g_mu) Y(x)
        // mid[3a+c] = sum_s r1_s[a] conj(xl_s[c]) + pl_s[a] conj(r2_s[c])   (makeSU3Contract, SU3.h:395-503)
        cx mid[9];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                cx acc{0.0, 0.0};
                for (int sp = 0; sp < 4; ++sp) {
                    cfmac(acc, xl[3 * sp + c], r1[3 * sp + a]);
                    cfmac(acc, r2[3 * sp + c], pl[3 * sp + a]);
                }
                mid[3 * a + c] = acc;
            }
        cx u[9], f[9];
        const unsigned el = (unsigned)((p * 4 + mu) * 9 * Vh + h);
        LinkIO<CLG_F64, 18>::load(U, Vh, el, u);
        su3_mm(u, mid, f);
        su3_ta(f);
        add_link(F, Vh, el, f, n.sign < 0 ? -kappa : kappa);
    }
}

// staggered force kernel over npole (phi_i, d_i) pairs with numerators a_i
struct PolePtrs {
    const char* phi[32];
    const char* d[32];
    double a[32];
};
__global__ __launch_bounds__(kForceBlock) void k_ks_force(Geo g, PolePtrs pp, int npole, const char* U, char* F) {
    int p, h;
    if (!map_site(g.Vh, 2, 0, 8, 1, p, h)) return;
    const Site s = site_of(g, p, h);
    const int Vh = g.Vh;
    for (int mu = 0; mu < 4; ++mu) {
        Nbr n;
        switch (mu) {
            case 0: n = fwd_nbr<0>(g, s, p); break;
            case 1: n = fwd_nbr<1>(g, s, p); break;
            case 2: n = fwd_nbr<2>(g, s, p); break;
            default: n = fwd_nbr<3>(g, s, p); break;
        }
        int sum = 0;
        for (int nu = 0; nu < mu; ++nu) sum += s.c[nu];
        const double eta = (sum & 1) ? -1.0 : 1.0;
        cx u[9];
        const unsigned el = (unsigned)((p * 4 + mu) * 9 * Vh + h);
        LinkIO<CLG_F64, 18>::load(U, Vh, el, u);
        cx tot[9];
        for (int k = 0; k < 9; ++k) tot[k] = cx{0.0, 0.0};
        for (int r = 0; r < npole; ++r) {
            cx phin[3], dn[3], phix[3], dx[3], t1[3], t2[3];
            KSIO<CLG_F64>::load(pp.phi[r], nullptr, Vh, (unsigned)(n.p * 3 * Vh + n.h), 0u, phin);
            KSIO<CLG_F64>::load(pp.d[r], nullptr, Vh, (unsigned)(n.p * 3 * Vh + n.h), 0u, dn);
            KSIO<CLG_F64>::load(pp.phi[r], nullptr, Vh, (unsigned)(p * 3 * Vh + h), 0u, phix);
            KSIO<CLG_F64>::load(pp.d[r], nullptr, Vh, (unsigned)(p * 3 * Vh + h), 0u, dx);
            su3_mv(u, phin, t1);
            su3_mv(u, dn, t2);
            // term[3a+c] = t1[a] conj(d(x)[c]) + phi(x)[a] conj(t2[c])   (makeSU3ContractV, SU3.h:505-523)
            cx term[9];
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) {
                    cx acc{0.0, 0.0};
                    cfmac(acc, dx[c], t1[a]);
                    cfmac(acc, t2[c], phix[a]);
                    term[3 * a + c] = acc;
                }
            const double f = eta * pp.a[r] * (n.sign < 0 ? -1.0 : 1.0);
            for (int k = 0; k < 9; ++k) term[k] = cscale(term[k], f);
            su3_ta(term);
            for (int k = 0; k < 9; ++k) tot[k] = cadd(tot[k], term[k]);
        }
        add_link(F, Vh, el, tot, -1.0);
    }
}

int check_force_args(const clg_field_s* force, const clg_field_s* gauge) {
    CLG_CHECK_ARG(force && gauge, "null field");
    CLG_CHECK_ARG(force->kind == CLG_GAUGE && force->prec == CLG_F64, "force must be an FP64 18-real gauge-shaped field");
    CLG_CHECK_ARG(gauge->kind == CLG_GAUGE && gauge->prec == CLG_F64, "gauge must be the FP64 18-real field");
    return check_same_lattice(force, gauge);
}

}  // namespace

extern "C" {

int clg_wilson_force_kernel(clg_context ctx, clg_field force, clg_field gauge, clg_field inv_dddag, clg_field inv_d,
                            double kappa) {
    CLG_CHECK_ARG(ctx && inv_dddag && inv_d, "null argument");
    CLG_TRY(check_force_args(force, gauge));
    CLG_CHECK_ARG(inv_dddag->kind == CLG_WILSON && inv_d->kind == CLG_WILSON && inv_dddag->prec == CLG_F64 &&
                      inv_d->prec == CLG_F64 && inv_dddag->subset == CLG_FULL && inv_d->subset == CLG_FULL,
                  "need FULL FP64 Wilson fields");
    Geo g = make_geo(inv_dddag);
    (void)hipSetDevice(ctx->device);
    k_wilson_force<<<sched_grid(g.Vh, 2, kForceBlock, 8, 1), kForceBlock, 0, ctx->stream>>>(g, (const char*)inv_dddag->data,
                                                                       (const char*)inv_d->data,
                                                                       (const char*)gauge->data, (char*)force->data, kappa);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

int clg_ks_force_kernel(clg_context ctx, clg_field force, clg_field gauge, const clg_field* phis, const clg_field* dphis,
                        const double* numerators, int npole) {
    CLG_CHECK_ARG(ctx && phis && dphis && numerators, "null argument");
    CLG_CHECK_ARG(npole >= 1 && npole <= 32, "npole must be in [1, 32]");
    CLG_TRY(check_force_args(force, gauge));
    PolePtrs pp;
    std::memset(&pp, 0, sizeof(pp));
    for (int r = 0; r < npole; ++r) {
        CLG_CHECK_ARG(phis[r] && dphis[r] && phis[r]->kind == CLG_KS && dphis[r]->kind == CLG_KS &&
                          phis[r]->prec == CLG_F64 && dphis[r]->prec == CLG_F64 && phis[r]->subset == CLG_FULL &&
                          dphis[r]->subset == CLG_FULL,
                      "pole fields must be FULL FP64 staggered fields");
        pp.phi[r] = (const char*)phis[r]->data;
        pp.d[r] = (const char*)dphis[r]->data;
        pp.a[r] = numerators[r];
    }
    Geo g = make_geo(phis[0]);
    (void)hipSetDevice(ctx->device);
    k_ks_force<<<sched_grid(g.Vh, 2, kForceBlock, 8, 1), kForceBlock, 0, ctx->stream>>>(g, pp, npole, (const char*)gauge->data,
                                                                   (char*)force->data);
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

// CFieldFermionWilsonSquareSU3::CalculateForceS (CFieldFermionWilsonSquareSU3.cu:1016-1100):
// X = (DD^dagger)^-1 phi with the configured solver (start may be the cached previous X),
// Y = D^dagger X, F += kernel.  x_out (nullable) receives X for the caller's cache.
int clg_wilson_force(clg_context ctx, const clg_solver_params* params, clg_field force, clg_field phi, clg_field gauge,
                     double kappa, clg_field start, clg_field x_out, clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && params && phi, "null argument");
    CLG_TRY(check_force_args(force, gauge));
    (void)hipSetDevice(ctx->device);
    clg_field_s *X, *Y;
    CLG_TRY(scratch_field(ctx, phi, CLG_WILSON, CLG_F64, CLG_FULL, 200, &X));
    CLG_TRY(scratch_field(ctx, phi, CLG_WILSON, CLG_F64, CLG_FULL, 201, &Y));
    for (int i = 0; i < 4; ++i) { X->bc[i] = phi->bc[i]; Y->bc[i] = phi->bc[i]; }
    const int st = clg_solve(ctx, params, X, phi, gauge, kappa, EFO_F_InverseDDdagger, start, rep);
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    CLG_TRY(wilson_full(ctx, Y, X, gauge, kappa, true, 1.0, 0.0));
    CLG_TRY(clg_wilson_force_kernel(ctx, force, gauge, X, Y, kappa));
    if (x_out) CLG_TRY(clg_field_copy(x_out, X));
    return st;
}

// CFieldFermionKSSU3::CalculateForceS (CFieldFermionKSSU3.cu:412-497) with the MD rational
// coefficients [c, a_1..a_n, b_1..b_n]
int clg_ks_force(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field force, clg_field phi,
                 clg_field gauge, double two_am, const double* coeffs, int degree, clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && params && phi && coeffs, "null argument");
    CLG_CHECK_ARG(degree >= 1 && degree <= 32, "degree must be in [1, 32]");
    CLG_TRY(check_force_args(force, gauge));
    (void)hipSetDevice(ctx->device);
    std::vector<clg_field_s*> P(degree), Dp(degree);
    std::vector<double> sh(2 * degree), a(degree);
    for (int k = 0; k < degree; ++k) {
        CLG_TRY(scratch_field(ctx, phi, CLG_KS, CLG_F64, CLG_FULL, 6000 + k, &P[k]));
        CLG_TRY(scratch_field(ctx, phi, CLG_KS, CLG_F64, CLG_FULL, 7000 + k, &Dp[k]));
        for (int i = 0; i < 4; ++i) { P[k]->bc[i] = phi->bc[i]; Dp[k]->bc[i] = phi->bc[i]; }
        a[k] = coeffs[1 + k];
        sh[2 * k] = coeffs[1 + degree + k];
        sh[2 * k + 1] = 0.0;
    }
    std::vector<clg_field> xs(P.begin(), P.end());
    const int st = clg_multishift_solve(ctx, params, ms_kind, prec, xs.data(), sh.data(), degree, phi, gauge, two_am,
                                        EFO_F_DDdagger, rep);
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    for (int k = 0; k < degree; ++k) CLG_TRY(ks_full(ctx, Dp[k], P[k], gauge, 0.0, false, 1.0, 0.0));  // D0 phi_i
    std::vector<clg_field> ps(P.begin(), P.end()), ds(Dp.begin(), Dp.end());
    CLG_TRY(clg_ks_force_kernel(ctx, force, gauge, ps.data(), ds.data(), a.data(), degree));
    return st;
}

// CActionFermionWilsonNf2::Energy (CActionFermionWilsonNf2.cpp:53-68): |D^-1 phi|^2 = phi^dagger (DD^dagger)^-1 phi
int clg_wilson_pf_energy(clg_context ctx, const clg_solver_params* params, clg_field phi, clg_field gauge, double kappa,
                         double* energy, clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && params && phi && gauge && energy, "null argument");
    (void)hipSetDevice(ctx->device);
    clg_field_s* X;
    CLG_TRY(scratch_field(ctx, phi, CLG_WILSON, CLG_F64, CLG_FULL, 202, &X));
    for (int i = 0; i < 4; ++i) X->bc[i] = phi->bc[i];
    const int st = clg_solve(ctx, params, X, phi, gauge, kappa, EFO_F_InverseD, nullptr, rep);
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    CLG_TRY(blas_norm2_dev(ctx, X, ctx->red_result + 32));
    CLG_TRY(fetch_doubles(ctx, ctx->red_result + 32, 1, energy));
    return st;
}

// CActionFermionKS::Energy (CActionFermionKS.cpp:52-71): Re phi^dagger r_MD(DD^dagger) phi
int clg_ks_pf_energy(clg_context ctx, const clg_solver_params* params, int ms_kind, int prec, clg_field phi,
                     clg_field gauge, double two_am, const double* coeffs, int degree, double* energy,
                     clg_solve_report* rep) {
    CLG_CHECK_ARG(ctx && params && phi && gauge && coeffs && energy, "null argument");
    (void)hipSetDevice(ctx->device);
    clg_field_s* T;
    CLG_TRY(scratch_field(ctx, phi, CLG_KS, CLG_F64, CLG_FULL, 203, &T));
    for (int i = 0; i < 4; ++i) T->bc[i] = phi->bc[i];
    const int st = clg_rational_apply(ctx, params, ms_kind, prec, T, phi, gauge, two_am, EFO_F_DDdagger, coeffs, degree,
                                      rep);
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) return st;
    CLG_TRY(blas_dot_dev(ctx, T, phi, ctx->red_result + 32));
    double z[2];
    CLG_TRY(fetch_doubles(ctx, ctx->red_result + 32, 2, z));
    *energy = z[0];
    return st;
}

}  // extern "C"
