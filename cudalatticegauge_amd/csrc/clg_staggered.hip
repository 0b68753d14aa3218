// clg_staggered.hip -- staggered (Kogut-Susskind) SU(3) operator.
//
// Reference _kernelDFermionKS (CFieldFermionKSSU3.cu:34-117):
//   (D psi)(x) = 2am psi(x) + sum_mu eta_mu(x) [ U_mu(x) s+ psi(x+mu) - U_mu(x-mu)^dag s- psi(x-mu) ]
//   D^dagger flips the sign of the hopping sum (D(m=0) is anti-Hermitian).
// eta_mu(x) = (-1)^(x_0 + ... + x_{mu-1}) is computed from the coordinates instead of the
// reference's per-site byte table (CIndexSquare.cu:257-267).
#include "clg_internal.h"

#include <cstring>

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
    double cr, ci;   // overall coefficient
    int pbase, nblk;
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

template <int MU, typename R, class SIO, class LIO>
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
    pin3(acc);
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
    pin3(acc);
}

template <typename R, int SP, int RECON>
__global__ __launch_bounds__(256) void k_ks(KSArgs a) {
    typedef KSIO<SP> SIO;
    typedef LinkIO<SP, RECON> LIO;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)a.nblk * a.g.Vh) return;
    const int blk = (int)(gid / a.g.Vh);
    const int h = (int)(gid - (long)blk * a.g.Vh);
    const int p = a.nblk == 2 ? blk : a.pbase;
    const Site s = site_of(a.g, p, h);
    C2<R> acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = cmk<R>(0, 0);
    ks_mu<0, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<1, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<2, R, SIO, LIO>(a, s, p, h, acc);
    ks_mu<3, R, SIO, LIO>(a, s, p, h, acc);
    const unsigned po = a.out_full ? (unsigned)p : 0u;
    const unsigned eo = po * SIO::NPL * a.g.Vh + h, no = po * a.g.Vh + h;
    const unsigned pi = a.in_full ? (unsigned)p : 0u;
    C2<R> self[3], o[3];
    SIO::load(a.in, a.inn, a.g.Vh, pi * SIO::NPL * a.g.Vh + h, pi * a.g.Vh + h, self);
    const C2<R> cc = cmk<R>((R)a.cr, (R)a.ci);
    const R m = (R)a.two_am, hs = (R)a.hs;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const C2<R> t = cmk<R>(m * self[c].x + hs * acc[c].x, m * self[c].y + hs * acc[c].y);
        o[c] = cmul(cc, t);
    }
    SIO::store(a.out, a.outn, a.g.Vh, eo, no, o);
}

}  // namespace

namespace clg {

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
    const unsigned grid = grid_for(2L * a.g.Vh, 256);
    hipStream_t st = ctx->stream;
    const bool r12 = gauge->kind == CLG_GAUGE12;
    if (out->prec == CLG_F64) {
        if (r12) k_ks<double, CLG_F64, 12><<<grid, 256, 0, st>>>(a);
        else k_ks<double, CLG_F64, 18><<<grid, 256, 0, st>>>(a);
    } else if (out->prec == CLG_F32) {
        if (r12) k_ks<float, CLG_F32, 12><<<grid, 256, 0, st>>>(a);
        else k_ks<float, CLG_F32, 18><<<grid, 256, 0, st>>>(a);
    } else {
        if (r12) k_ks<float, CLG_F16, 12><<<grid, 256, 0, st>>>(a);
        else k_ks<float, CLG_F16, 18><<<grid, 256, 0, st>>>(a);
    }
    CLG_HIP(hipGetLastError());
    return CLG_OK;
}

}  // namespace clg

extern "C" {

int clg_ks_apply(clg_context ctx, clg_field out, clg_field in, clg_field gauge, double two_am, int op, int coeff_type,
                 double cre, double cim) {
    CLG_CHECK_ARG(ctx && out && in && gauge, "null argument");
    CLG_CHECK_ARG(out != in && out->data != in->data, "out and in must be distinct fields");
    CLG_CHECK_ARG(out->kind == CLG_KS && in->kind == CLG_KS, "need staggered fields");
    CLG_CHECK_ARG(out->subset == CLG_FULL && in->subset == CLG_FULL, "need FULL fields");
    CLG_CHECK_ARG(out->prec == in->prec && gauge->prec == in->prec, "precision mismatch");
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

}  // extern "C"
