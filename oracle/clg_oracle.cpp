// clg_oracle.cpp -- CPU oracle (test infrastructure only, never linked into the product).
//
// A plain restatement of the reference algorithms on the fermion-matrix hot path,
// written from the reference source text.  Each function names the reference lines
// it follows.  Arithmetic is FP64 unless a *_f32 entry point is used; reductions use
// long double.  The operators deliberately follow the reference's own structure
// (baked neighbour tables with sign tags, full 4-spinor gamma products, no spin
// projection) so that the product's different algorithm is checked against an
// independent formulation.

#include "clg_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <chrono>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename R>
struct cx {
    R re, im;
};
template <typename R> inline cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <typename R> inline cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> inline cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> inline cx<R> operator*(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> inline cx<R> conj(cx<R> a) { return {a.re, -a.im}; }
template <typename R> inline cx<R> scale(cx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> inline R abs2(cx<R> a) { return a.re * a.re + a.im * a.im; }
// i^k * a   (Vectors.h:485-510 MulZ4)
template <typename R> inline cx<R> mulz4(cx<R> a, int k) {
    switch (k & 3) {
        case 1: return {-a.im, a.re};
        case 2: return {-a.re, -a.im};
        case 3: return {a.im, -a.re};
        default: return a;
    }
}

struct Lat {
    int L[4];
    long V;
    long mult[4];
    explicit Lat(const int d[4]) {
        for (int i = 0; i < 4; ++i) L[i] = d[i];
        mult[3] = 1;
        mult[2] = L[3];
        mult[1] = (long)L[2] * L[3];
        mult[0] = (long)L[1] * L[2] * L[3];
        V = (long)L[0] * mult[0];
    }
    void coords(long s, int c[4]) const {
        c[0] = (int)(s / mult[0]);
        c[1] = (int)((s % mult[0]) / mult[1]);
        c[2] = (int)((s % mult[1]) / mult[2]);
        c[3] = (int)(s % mult[2]);
    }
    long index(const int c[4]) const {
        return ((long)(c[0] * (long)L[1] + c[1]) * L[2] + c[2]) * L[3] + c[3];
    }
};

bool dims_ok(const int* d) {
    if (!d) return false;
    for (int i = 0; i < 4; ++i)
        if (d[i] < 1 || d[i] > 127) return false;  // SBYTE coordinates in the reference
    return true;
}

// Neighbour tables: site(x +- mu) with the torus sign, walking the coordinate back into
// the box exactly as _kernalBakeEdgeTorusBoundary does (CBoundaryConditionTorusSquare.cu:19-64).
struct Moves {
    std::vector<int> fwd, bwd, fsg, bsg;
};
Moves build_moves(const Lat& lat, const int bc[4]) {
    Moves m;
    m.fwd.resize(lat.V * 4);
    m.bwd.resize(lat.V * 4);
    m.fsg.resize(lat.V * 4);
    m.bsg.resize(lat.V * 4);
    for (long s = 0; s < lat.V; ++s) {
        int c[4];
        lat.coords(s, c);
        for (int mu = 0; mu < 4; ++mu) {
            for (int dir = 0; dir < 2; ++dir) {
                int w[4] = {c[0], c[1], c[2], c[3]};
                w[mu] += dir == 0 ? 1 : -1;
                int sign = 1;
                for (int d = 0; d < 4; ++d) {
                    while (w[d] < 0) { w[d] += lat.L[d]; sign *= bc[d]; }
                    while (w[d] >= lat.L[d]) { w[d] -= lat.L[d]; sign *= bc[d]; }
                }
                const long n = lat.index(w);
                if (dir == 0) { m.fwd[s * 4 + mu] = (int)n; m.fsg[s * 4 + mu] = sign; }
                else { m.bwd[s * 4 + mu] = (int)n; m.bsg[s * 4 + mu] = sign; }
            }
        }
    }
    return m;
}

// chiral gamma set, GammaMatrix.h:447-475: (G psi)[row] = i^z[row] psi[idx[row]]
const int kGIdx[5][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {2, 3, 0, 1}, {0, 1, 2, 3}};
const int kGZ4[5][4] = {{3, 3, 1, 1}, {2, 0, 0, 2}, {3, 1, 1, 3}, {0, 0, 0, 0}, {0, 0, 2, 2}};

template <typename R> struct Spinor { cx<R> c[4][3]; };
template <typename R> struct Mat { cx<R> m[9]; };
template <typename R> struct Vec3 { cx<R> c[3]; };

template <typename R> inline Spinor<R> gmul(int g, const Spinor<R>& s) {
    Spinor<R> r;
    for (int row = 0; row < 4; ++row)
        for (int a = 0; a < 3; ++a) r.c[row][a] = mulz4(s.c[kGIdx[g][row]][a], kGZ4[g][row]);
    return r;
}
template <typename R> inline void sadd(Spinor<R>& a, const Spinor<R>& b) {
    for (int s = 0; s < 4; ++s) for (int c = 0; c < 3; ++c) a.c[s][c] = a.c[s][c] + b.c[s][c];
}
template <typename R> inline void ssub(Spinor<R>& a, const Spinor<R>& b) {
    for (int s = 0; s < 4; ++s) for (int c = 0; c < 3; ++c) a.c[s][c] = a.c[s][c] - b.c[s][c];
}
// SU3.h:771-788 MulVector / MulWilsonVector
template <typename R> inline Vec3<R> mv(const Mat<R>& u, const Vec3<R>& v) {
    Vec3<R> r;
    for (int i = 0; i < 3; ++i) r.c[i] = u.m[3 * i] * v.c[0] + u.m[3 * i + 1] * v.c[1] + u.m[3 * i + 2] * v.c[2];
    return r;
}
template <typename R> inline Spinor<R> mvw(const Mat<R>& u, const Spinor<R>& v) {
    Spinor<R> r;
    for (int s = 0; s < 4; ++s) {
        Vec3<R> t;
        for (int c = 0; c < 3; ++c) t.c[c] = v.c[s][c];
        t = mv(u, t);
        for (int c = 0; c < 3; ++c) r.c[s][c] = t.c[c];
    }
    return r;
}
template <typename R> inline Mat<R> dag(const Mat<R>& u) {
    Mat<R> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[3 * i + j] = conj(u.m[3 * j + i]);
    return r;
}
template <typename R> inline Mat<R> mmul(const Mat<R>& a, const Mat<R>& b) {
    Mat<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}
template <typename R> inline Mat<R> load_mat(const R* p) {
    Mat<R> m;
    for (int i = 0; i < 9; ++i) m.m[i] = mk(p[2 * i], p[2 * i + 1]);
    return m;
}
template <typename R> inline void store_mat(R* p, const Mat<R>& m) {
    for (int i = 0; i < 9; ++i) { p[2 * i] = m.m[i].re; p[2 * i + 1] = m.m[i].im; }
}
template <typename R> inline Spinor<R> load_sp(const R* p) {
    Spinor<R> s;
    for (int k = 0; k < 12; ++k) s.c[k / 3][k % 3] = mk(p[2 * k], p[2 * k + 1]);
    return s;
}
template <typename R> inline void store_sp(R* p, const Spinor<R>& s) {
    for (int k = 0; k < 12; ++k) { p[2 * k] = s.c[k / 3][k % 3].re; p[2 * k + 1] = s.c[k / 3][k % 3].im; }
}
template <typename R> inline Vec3<R> load_v(const R* p) {
    Vec3<R> v;
    for (int k = 0; k < 3; ++k) v.c[k] = mk(p[2 * k], p[2 * k + 1]);
    return v;
}

// SU3.h:1058-1088 Norm: Gram-Schmidt rows 0,1; row 2 = conj(r0 x r1)
template <typename R> void su3_norm(Mat<R>& u) {
    cx<R>* m = u.m;
    const R d1 = R(1) / std::sqrt(abs2(m[0]) + abs2(m[1]) + abs2(m[2]));
    m[0] = scale(m[0], d1); m[1] = scale(m[1], d1); m[2] = scale(m[2], d1);
    const cx<R> sp1 = conj(m[0] * conj(m[3]) + m[1] * conj(m[4]) + m[2] * conj(m[5]));
    m[3] = m[3] - sp1 * m[0]; m[4] = m[4] - sp1 * m[1]; m[5] = m[5] - sp1 * m[2];
    const R d2 = R(1) / std::sqrt(abs2(m[3]) + abs2(m[4]) + abs2(m[5]));
    m[3] = scale(m[3], d2); m[4] = scale(m[4], d2); m[5] = scale(m[5], d2);
    m[6] = conj(m[1] * m[5] - m[2] * m[4]);
    m[7] = conj(m[2] * m[3] - m[0] * m[5]);
    m[8] = conj(m[0] * m[4] - m[1] * m[3]);
}

// SU3.h:953-984 Ta: traceless anti-Hermitian part (U - U^dagger)/2 - tr/3
template <typename R> void su3_ta(Mat<R>& u) {
    cx<R>* m = u.m;
    const cx<R> n1 = m[1] - conj(m[3]);
    const cx<R> n2 = m[2] - conj(m[6]);
    const cx<R> n5 = m[5] - conj(m[7]);
    m[1] = scale(n1, R(0.5)); m[3] = mk(-m[1].re, m[1].im);
    m[2] = scale(n2, R(0.5)); m[6] = mk(-m[2].re, m[2].im);
    m[5] = scale(n5, R(0.5)); m[7] = mk(-m[5].re, m[5].im);
    const R tr = (m[0].im + m[4].im + m[8].im) / R(3);
    m[0] = mk(R(0), m[0].im - tr);
    m[4] = mk(R(0), m[4].im - tr);
    m[8] = mk(R(0), m[8].im - tr);
}

// ---------------- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----------------
inline void philox(const unsigned ctr[4], const unsigned key[2], unsigned out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
const unsigned kKeyHi = 0x434C4721u;  // second key word, fixed
enum { STREAM_GAUGE = 0, STREAM_WILSON = 1, STREAM_KS = 2 };

inline double u53(unsigned a, unsigned b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}
// two N(0,1) numbers from one Philox block (Box-Muller)
inline void gauss2(unsigned idx, unsigned draw, unsigned stream, unsigned seed, double g[2]) {
    const unsigned ctr[4] = {idx, draw, stream, 0u};
    const unsigned key[2] = {seed, kKeyHi};
    unsigned w[4];
    philox(ctr, key, w);
    const double u1 = u53(w[0], w[1]);
    const double u2 = u53(w[2], w[3]);
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double ph = 6.283185307179586476925286766559 * u2;
    g[0] = r * std::cos(ph);
    g[1] = r * std::sin(ph);
}

// exp(A) for a 3x3 complex matrix by scaling and squaring of the Taylor series.
void expm3(const Mat<double>& A, Mat<double>& E) {
    double nrm = 0;
    for (int i = 0; i < 9; ++i) nrm += abs2(A.m[i]);
    nrm = std::sqrt(nrm);
    int s = 0;
    while (nrm > 0.5) { nrm *= 0.5; ++s; }
    const double f = std::ldexp(1.0, -s);
    Mat<double> B;
    for (int i = 0; i < 9; ++i) B.m[i] = scale(A.m[i], f);
    Mat<double> T, S;
    for (int i = 0; i < 9; ++i) { T.m[i] = mk(0.0, 0.0); }
    T.m[0] = T.m[4] = T.m[8] = mk(1.0, 0.0);
    S = T;
    for (int k = 1; k <= 20; ++k) {
        T = mmul(T, B);
        for (int i = 0; i < 9; ++i) { T.m[i] = scale(T.m[i], 1.0 / k); S.m[i] = S.m[i] + T.m[i]; }
    }
    for (int k = 0; k < s; ++k) S = mmul(S, S);
    E = S;
}

void random_link(unsigned link, double eps, unsigned seed, double* out) {
    double g[10];
    for (int d = 0; d < 5; ++d) gauss2(link, (unsigned)d, STREAM_GAUGE, seed, g + 2 * d);
    const double tr = (g[0] + g[1] + g[2]) / 3.0;
    const double is2 = 0.70710678118654752440084436210485;
    Mat<double> H;
    H.m[0] = mk(g[0] - tr, 0.0);
    H.m[4] = mk(g[1] - tr, 0.0);
    H.m[8] = mk(g[2] - tr, 0.0);
    H.m[1] = mk(g[3] * is2, g[4] * is2);
    H.m[2] = mk(g[5] * is2, g[6] * is2);
    H.m[5] = mk(g[7] * is2, g[8] * is2);
    H.m[3] = conj(H.m[1]);
    H.m[6] = conj(H.m[2]);
    H.m[7] = conj(H.m[5]);
    Mat<double> A, E;
    for (int i = 0; i < 9; ++i) A.m[i] = mk(-eps * H.m[i].im, eps * H.m[i].re);  // i*eps*H
    expm3(A, E);
    su3_norm(E);
    store_mat(out, E);
}

// ---------------- Wilson operator, CFieldFermionWilsonSquareSU3.cu:32-163 ----------------
template <typename R>
void wilson_kernel(const Lat& lat, const Moves& mv_, R kappa, const R* U, const R* in, R* out,
                   bool dagger, int coeff_type, R cre, R cim) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < lat.V; ++s) {
        Spinor<R> self = load_sp(in + s * 24);
        if (dagger) self = gmul(4, self);
        Spinor<R> res;
        for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) res.c[a][c] = mk(R(0), R(0));
        for (int mu = 0; mu < 4; ++mu) {
            const long link = s * 4 + mu;
            const Mat<R> Ux = load_mat(U + link * 18);
            // gauge move: link at x-mu with the dagger tag (CIndexSquare.cu:212-236)
            const Mat<R> Um = dag(load_mat(U + ((long)mv_.bwd[link] * 4 + mu) * 18));
            Spinor<R> pp = load_sp(in + (long)mv_.fwd[link] * 24);
            Spinor<R> pm = load_sp(in + (long)mv_.bwd[link] * 24);
            if (dagger) { pp = gmul(4, pp); pm = gmul(4, pm); }
            const Spinor<R> up = mvw(Ux, pp);
            if (mv_.fsg[link] < 0) { ssub(res, up); sadd(res, gmul(mu, up)); }
            else { sadd(res, up); ssub(res, gmul(mu, up)); }
            const Spinor<R> um = mvw(Um, pm);
            if (mv_.bsg[link] < 0) { ssub(res, um); ssub(res, gmul(mu, um)); }
            else { sadd(res, um); sadd(res, gmul(mu, um)); }
        }
        for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) res.c[a][c] = scale(res.c[a][c], kappa);
        ssub(self, res);
        if (dagger) self = gmul(4, self);
        if (coeff_type == 1 || coeff_type == 2) {
            const R f = coeff_type == 1 ? R(-1) : cre;
            for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) self.c[a][c] = scale(self.c[a][c], f);
        } else if (coeff_type == 3) {
            for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) self.c[a][c] = self.c[a][c] * mk(cre, cim);
        }
        store_sp(out + s * 24, self);
    }
}

// DS/DdaggerS/DDS/DDdaggerS wrappers, CFieldFermionWilsonSquareSU3.cu:909-1014
template <typename R>
void wilson_op(const Lat& lat, const Moves& m, R kappa, const R* U, const R* in, R* out, int op,
               int ct, R cre, R cim) {
    if (op == ORC_D || op == ORC_DDAG) {
        wilson_kernel<R>(lat, m, kappa, U, in, out, op == ORC_DDAG, ct, cre, cim);
        return;
    }
    std::vector<R> tmp(lat.V * 24);
    wilson_kernel<R>(lat, m, kappa, U, in, tmp.data(), op == ORC_DDDAG, 0, R(1), R(0));
    wilson_kernel<R>(lat, m, kappa, U, tmp.data(), out, false, ct, cre, cim);
}

// ---------------- staggered operator, CFieldFermionKSSU3.cu:34-117 ----------------
template <typename R>
void ks_kernel(const Lat& lat, const Moves& m, const std::vector<unsigned char>& eta, R two_am,
               const R* U, const R* in, R* out, bool dagger, int ct, R cre, R cim) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < lat.V; ++s) {
        Vec3<R> res;
        for (int c = 0; c < 3; ++c) res.c[c] = mk(R(0), R(0));
        for (int mu = 0; mu < 4; ++mu) {
            const R e = ((eta[s] >> mu) & 1) ? R(-1) : R(1);
            const long link = s * 4 + mu;
            const Mat<R> Ux = load_mat(U + link * 18);
            const Mat<R> Um = dag(load_mat(U + ((long)m.bwd[link] * 4 + mu) * 18));
            Vec3<R> a = mv(Ux, load_v(in + (long)m.fwd[link] * 6));
            if (m.fsg[link] < 0) for (int c = 0; c < 3; ++c) a.c[c] = scale(a.c[c], R(-1));
            const Vec3<R> b = mv(Um, load_v(in + (long)m.bwd[link] * 6));
            for (int c = 0; c < 3; ++c) a.c[c] = m.bsg[link] < 0 ? a.c[c] + b.c[c] : a.c[c] - b.c[c];
            for (int c = 0; c < 3; ++c) res.c[c] = res.c[c] + scale(a.c[c], e);
        }
        Vec3<R> o = load_v(in + s * 6);
        for (int c = 0; c < 3; ++c) {
            o.c[c] = scale(o.c[c], two_am);
            o.c[c] = dagger ? o.c[c] - res.c[c] : o.c[c] + res.c[c];
            if (ct == 1) o.c[c] = scale(o.c[c], R(-1));
            else if (ct == 2) o.c[c] = scale(o.c[c], cre);
            else if (ct == 3) o.c[c] = o.c[c] * mk(cre, cim);
            out[s * 6 + 2 * c] = o.c[c].re;
            out[s * 6 + 2 * c + 1] = o.c[c].im;
        }
    }
}

std::vector<unsigned char> make_eta(const Lat& lat) {
    std::vector<unsigned char> eta(lat.V);
    for (long s = 0; s < lat.V; ++s) {
        int c[4];
        lat.coords(s, c);
        unsigned char b = 0;
        for (int mu = 0; mu < 5; ++mu) {
            int sum = 0;
            for (int nu = 0; nu < mu && nu < 4; ++nu) sum += c[nu];
            if (sum & 1) b |= (unsigned char)(1u << mu);
        }
        eta[s] = b;
    }
    return eta;
}

// CFieldFermionKSSU3.cu:1344-1575 DS/DdaggerS/DDS/DDdaggerS
template <typename R>
void ks_op(const Lat& lat, const Moves& m, const std::vector<unsigned char>& eta, R two_am,
           const R* U, const R* in, R* out, int op, int ct, R cre, R cim) {
    if (op == ORC_D || op == ORC_DDAG) {
        ks_kernel<R>(lat, m, eta, two_am, U, in, out, op == ORC_DDAG, ct, cre, cim);
        return;
    }
    std::vector<R> tmp(lat.V * 6);
    ks_kernel<R>(lat, m, eta, two_am, U, in, tmp.data(), op == ORC_DDDAG, 0, R(1), R(0));
    ks_kernel<R>(lat, m, eta, two_am, U, tmp.data(), out, false, ct, cre, cim);
}

// ---------------- generic field helpers for the solvers ----------------
typedef std::vector<double> F;

struct Op {
    const orc_operator* A;
    const double* U;
    Lat lat;
    Moves m;
    std::vector<unsigned char> eta;
    long n;  // complex numbers per field
    Op(const orc_operator* a, const double* u) : A(a), U(u), lat(a->dims) {
        m = build_moves(lat, a->bc);
        if (a->kind == 1) eta = make_eta(lat);
        n = lat.V * (a->kind == 0 ? 12 : 3);
    }
    // out = coeff * Op(in)
    void apply(const double* in, double* out, int ct = 0) const {
        if (A->kind == 0)
            wilson_op<double>(lat, m, A->mass_or_kappa, U, in, out, A->op, ct, 1.0, 0.0);
        else
            ks_op<double>(lat, m, eta, A->mass_or_kappa, U, in, out, A->op, ct, 1.0, 0.0);
    }
};

cx<double> dotc(const F& a, const F& b) {
    long double re = 0, im = 0;
    const long n = (long)a.size() / 2;
    for (long i = 0; i < n; ++i) {
        const double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        re += (long double)ar * br + (long double)ai * bi;
        im += (long double)ar * bi - (long double)ai * br;
    }
    return mk((double)re, (double)im);
}
// y += a x
void axpy(F& y, cx<double> a, const F& x) {
    const long n = (long)y.size() / 2;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const double xr = x[2 * i], xi = x[2 * i + 1];
        y[2 * i] += a.re * xr - a.im * xi;
        y[2 * i + 1] += a.re * xi + a.im * xr;
    }
}
void scal(F& y, cx<double> a) {
    const long n = (long)y.size() / 2;
    for (long i = 0; i < n; ++i) {
        const double r = y[2 * i], m = y[2 * i + 1];
        y[2 * i] = a.re * r - a.im * m;
        y[2 * i + 1] = a.re * m + a.im * r;
    }
}
cx<double> cdiv(cx<double> a, cx<double> b) {
    const double d = abs2(b);
    return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

}  // namespace

// =====================================================================================
extern "C" {

int orc_site_index(const int dims[4], int x, int y, int z, int t) {
    if (!dims_ok(dims)) return -1;
    return ((x * dims[1] + y) * dims[2] + z) * dims[3] + t;
}

int orc_build_moves(const int dims[4], const int bc[4], int* fwd_site, int* bwd_site,
                    int* fwd_sign, int* bwd_sign) {
    if (!dims_ok(dims) || !bc) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    std::memcpy(fwd_site, m.fwd.data(), sizeof(int) * m.fwd.size());
    std::memcpy(bwd_site, m.bwd.data(), sizeof(int) * m.bwd.size());
    std::memcpy(fwd_sign, m.fsg.data(), sizeof(int) * m.fsg.size());
    std::memcpy(bwd_sign, m.bsg.data(), sizeof(int) * m.bsg.size());
    return 0;
}

int orc_build_eta(const int dims[4], unsigned char* eta) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    std::vector<unsigned char> e = make_eta(lat);
    std::memcpy(eta, e.data(), e.size());
    return 0;
}

// intokernal_even / intokernal_odd, CLGLib_Private.h:72-116: candidate 2e (even) or
// 2e+1 (odd); move by one site when the candidate has the other parity.
int orc_eo_sites(const int dims[4], int parity, int* out) {
    if (!dims_ok(dims) || (dims[3] & 1)) return -1;
    Lat lat(dims);
    for (long e = 0; e < lat.V / 2; ++e) {
        long s = 2 * e + (parity ? 1 : 0);
        int c[4];
        lat.coords(s, c);
        const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
        if (parity == 0 && p) s = s + 1;
        if (parity == 1 && !p) s = s - 1;
        out[e] = (int)s;
    }
    return 0;
}

int orc_gamma(int which, int idx[4], int z4[4]) {
    if (which < 0 || which > 4) return -1;
    for (int r = 0; r < 4; ++r) { idx[r] = kGIdx[which][r]; z4[r] = kGZ4[which][r]; }
    return 0;
}

int orc_philox(const unsigned ctr[4], const unsigned key[2], unsigned out[4]) {
    philox(ctr, key, out);
    return 0;
}

int orc_gauge_random(const int dims[4], double eps, unsigned seed, double* U) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
#pragma omp parallel for schedule(static)
    for (long l = 0; l < lat.V * 4; ++l) random_link((unsigned)l, eps, seed, U + l * 18);
    return 0;
}

int orc_gauge_unit(const int dims[4], double* U) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    std::memset(U, 0, sizeof(double) * lat.V * 72);
    for (long l = 0; l < lat.V * 4; ++l) { U[l * 18 + 0] = 1; U[l * 18 + 8] = 1; U[l * 18 + 16] = 1; }
    return 0;
}

int orc_wilson_gaussian(const int dims[4], unsigned seed, double* psi) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    const double is2 = 0.70710678118654752440084436210485;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < lat.V; ++s)
        for (int k = 0; k < 12; ++k) {
            double g[2];
            gauss2((unsigned)s, (unsigned)k, STREAM_WILSON, seed, g);
            psi[s * 24 + 2 * k] = g[0] * is2;
            psi[s * 24 + 2 * k + 1] = g[1] * is2;
        }
    return 0;
}

int orc_ks_gaussian(const int dims[4], unsigned seed, double* chi) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    const double is2 = 0.70710678118654752440084436210485;
    for (long s = 0; s < lat.V; ++s)
        for (int k = 0; k < 3; ++k) {
            double g[2];
            gauss2((unsigned)s, (unsigned)k, STREAM_KS, seed, g);
            chi[s * 6 + 2 * k] = g[0] * is2;
            chi[s * 6 + 2 * k + 1] = g[1] * is2;
        }
    return 0;
}

int orc_su3_norm(double* m, long n) {
    for (long i = 0; i < n; ++i) {
        Mat<double> u = load_mat(m + 18 * i);
        su3_norm(u);
        store_mat(m + 18 * i, u);
    }
    return 0;
}

int orc_wilson_op(const int dims[4], const int bc[4], double kappa, const double* U,
                  const double* in, double* out, int op, int ct, double cre, double cim) {
    if (!dims_ok(dims) || op < 0 || op > 3 || in == out) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    wilson_op<double>(lat, m, kappa, U, in, out, op, ct, cre, cim);
    return 0;
}

int orc_wilson_op_f32(const int dims[4], const int bc[4], float kappa, const float* U,
                      const float* in, float* out, int op) {
    if (!dims_ok(dims) || op < 0 || op > 3 || in == out) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    wilson_op<float>(lat, m, kappa, U, in, out, op, 0, 1.0f, 0.0f);
    return 0;
}

int orc_ks_op(const int dims[4], const int bc[4], double two_am, const double* U,
              const double* in, double* out, int op, int ct, double cre, double cim) {
    if (!dims_ok(dims) || op < 0 || op > 3 || in == out) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    std::vector<unsigned char> eta = make_eta(lat);
    ks_op<double>(lat, m, eta, two_am, U, in, out, op, ct, cre, cim);
    return 0;
}

int orc_ks_op_f32(const int dims[4], const int bc[4], float two_am, const float* U,
                  const float* in, float* out, int op) {
    if (!dims_ok(dims) || op < 0 || op > 3 || in == out) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    std::vector<unsigned char> eta = make_eta(lat);
    ks_op<float>(lat, m, eta, two_am, U, in, out, op, 0, 1.0f, 0.0f);
    return 0;
}

int orc_dot(const double* a, const double* b, long n, double out[2]) {
    long double re = 0, im = 0;
    for (long i = 0; i < n; ++i) {
        re += (long double)a[2 * i] * b[2 * i] + (long double)a[2 * i + 1] * b[2 * i + 1];
        im += (long double)a[2 * i] * b[2 * i + 1] - (long double)a[2 * i + 1] * b[2 * i];
    }
    out[0] = (double)re;
    out[1] = (double)im;
    return 0;
}

// ---------------- GMRES, CSolverGMRES.cpp:111-358 ----------------
int orc_gmres(const orc_operator* A, const double* U, const double* b, double* x,
              const double* start, int max_dim, int restart, double accuracy, double blength,
              int stop_rule, int* cycles_out, double* last_dev_out) {
    if (!A || !dims_ok(A->dims) || max_dim < 1 || max_dim > 100 || restart < 1) return -1;
    Op op(A, U);
    const long n2 = op.n * 2;
    F B(b, b + n2), X(start ? start : b, (start ? start : b) + n2), W(n2);
    std::vector<F> Vb(max_dim, F(n2));
    const int m = max_dim;
    std::vector<cx<double>> h((m + 1) * m), g(m + 1), y(m);
    auto H = [&](int r, int c) -> cx<double>& { return h[r * m + c]; };
    const double bnorm = std::sqrt(dotc(B, B).re);
    double last = 0;
    int cyc = 0;
    bool ok = false;
    for (int i = 0; i < restart; ++i) {
        ++cyc;
        // r = b - A x
        op.apply(X.data(), Vb[0].data(), 1);
        axpy(Vb[0], mk(1.0, 0.0), B);
        const double beta = std::sqrt(dotc(Vb[0], Vb[0]).re);
        if (stop_rule == 1 && beta / bnorm < accuracy) { last = beta / bnorm; ok = true; break; }
        scal(Vb[0], mk(1.0 / beta, 0.0));
        for (int j = 0; j < m; ++j) {
            op.apply(Vb[j].data(), W.data(), 0);
            for (int k = 0; k <= j; ++k) {
                const cx<double> d = dotc(Vb[k], W);
                H(k, j) = d;
                axpy(W, mk(-d.re, -d.im), Vb[k]);
            }
            const double wn = std::sqrt(dotc(W, W).re);
            H(j + 1, j) = mk(wn, 0.0);
            if (j < m - 1) {
                scal(W, mk(1.0 / wn, 0.0));
                Vb[j + 1] = W;
            }
        }
        // RotateH (:262-327): Givens rotations on H and g
        g[0] = mk(beta, 0.0);
        for (int k = 1; k <= m; ++k) g[k] = mk(0.0, 0.0);
        for (int k = 0; k < m; ++k) {
            const double den = 1.0 / std::sqrt(abs2(H(k, k)) + abs2(H(k + 1, k)));
            const cx<double> cs = scale(H(k, k), den), sn = scale(H(k + 1, k), den);
            const cx<double> csh = conj(cs), snh = conj(sn);
            for (int j = k; j < m; ++j) {
                const cx<double> hkj = H(k, j);
                H(k, j) = csh * hkj + snh * H(k + 1, j);
                H(k + 1, j) = cs * H(k + 1, j) - sn * hkj;
            }
            const cx<double> mg = mk(-g[k].re, -g[k].im);
            g[k] = csh * g[k];
            g[k + 1] = sn * mg;
        }
        const double dev = abs2(g[m]);
        // SolveY (:329-358): back substitution
        for (int k = m - 1; k >= 0; --k) {
            for (int j = k + 1; j < m; ++j) g[k] = g[k] - H(k, j) * y[j];
            y[k] = cdiv(g[k], H(k, k));
        }
        for (int j = 0; j < m; ++j) axpy(X, y[j], Vb[j]);
        if (stop_rule == 0) {
            last = dev;
            if (dev < accuracy * blength) { ok = true; break; }
        } else {
            F R(n2);
            op.apply(X.data(), R.data(), 1);
            axpy(R, mk(1.0, 0.0), B);
            last = std::sqrt(dotc(R, R).re) / bnorm;
            if (last < accuracy) { ok = true; break; }
        }
    }
    std::memcpy(x, X.data(), sizeof(double) * n2);
    if (cycles_out) *cycles_out = cyc;
    if (last_dev_out) *last_dev_out = last;
    return ok ? 0 : 1;
}

// ---------------- BiCGStab, CSolverBiCGstab.cpp:403-607 ----------------
int orc_bicgstab(const orc_operator* A, const double* U, const double* b, double* x,
                 const double* start, int step_count, int deviation_check, int retry,
                 double accuracy, double small_rho, int absolute, int* iters_out,
                 double* last_dev_out) {
    if (!A || !dims_ok(A->dims) || deviation_check < 1 || retry < 1) return -1;
    Op op(A, U);
    const long n2 = op.n * 2;
    F B(b, b + n2);
    F X(start ? start : b, (start ? start : b) + n2);
    F R = X, P(n2), V(n2), Rh(n2), S(n2), T(n2);
    const double blen = absolute ? 1.0 : std::sqrt(abs2(dotc(B, B)));
    int iters = 0;
    double dev = 0;
    for (int i = 0; i < retry; ++i) {
        F tmp(n2);
        op.apply(R.data(), tmp.data(), 1);
        R = tmp;
        axpy(R, mk(1.0, 0.0), B);
        Rh = R;
        for (long k = 0; k < n2 / 2; ++k) Rh[2 * k + 1] = -Rh[2 * k + 1];  // Dagger(): conjugate
        P = R;
        cx<double> rho = mk(0.0, 0.0), alpha = rho, beta = rho, omega = rho;
        for (int j = 0; j < step_count * deviation_check; ++j) {
            ++iters;
            if (j == 0) rho = dotc(Rh, R);
            if (abs2(rho) < ((retry == i + 1) ? 1.175494351e-38 : small_rho)) { i = i + 1; break; }
            op.apply(P.data(), V.data(), 0);
            alpha = cdiv(rho, dotc(Rh, V));
            S = R;
            axpy(S, mk(-alpha.re, -alpha.im), V);
            if ((j + 1) % deviation_check == 0) {
                dev = dotc(S, S).re / blen;
                if (dev < accuracy) {
                    std::memcpy(x, X.data(), sizeof(double) * n2);
                    if (iters_out) *iters_out = iters;
                    if (last_dev_out) *last_dev_out = dev;
                    return 0;
                }
            }
            op.apply(S.data(), T.data(), 0);
            const cx<double> ts = dotc(S, T);
            const double tt = dotc(T, T).re;
            omega = mk(ts.re / tt, ts.im / tt);
            R = S;
            axpy(R, mk(-omega.re, -omega.im), T);
            beta = cdiv(alpha, omega * rho);
            rho = dotc(Rh, R);
            beta = beta * rho;
            axpy(X, alpha, P);
            axpy(X, omega, S);
            scal(V, omega);
            axpy(P, mk(-1.0, 0.0), V);
            scal(P, beta);
            axpy(P, mk(1.0, 0.0), R);
        }
        R = X;
    }
    std::memcpy(x, X.data(), sizeof(double) * n2);
    if (iters_out) *iters_out = iters;
    if (last_dev_out) *last_dev_out = dev;
    return 1;
}

// ---------------- multi-shift FOM, CMultiShiftFOM.cpp:97-270 ----------------
int orc_msfom(const orc_operator* A, const double* U, const double* b, double* xs,
              const double* shifts, int nshift, int max_dim, int restart, double accuracy,
              double blength, int* cycles_out, double* max_err_out) {
    if (!A || !dims_ok(A->dims) || nshift < 1 || max_dim < 1 || restart < 1) return -1;
    Op op(A, U);
    const long n2 = op.n * 2;
    const int m = max_dim;
    std::vector<F> X(nshift, F(n2, 0.0)), Vb(m, F(n2));
    F Rv(b, b + n2), W(n2);
    std::vector<cx<double>> beta0(nshift, mk(1.0, 0.0));
    std::vector<double> blen0(nshift, 1.0);
    std::vector<cx<double>> h((m + 1) * m), hc(m * m), g(m), y(m);
    auto Hh = [&](int r, int c) -> cx<double>& { return h[r * m + c]; };
    double maxerr = 0;
    int cyc = 0;
    for (int i = 0; i < restart; ++i) {
        ++cyc;
        Vb[0] = Rv;
        if (i == 0) {
            const double fb = std::sqrt(dotc(Vb[0], Vb[0]).re);
            for (int n = 0; n < nshift; ++n) { beta0[n] = mk(fb, 0.0); blen0[n] = fb; }
            scal(Vb[0], mk(1.0 / fb, 0.0));
        }
        for (int j = 0; j < m; ++j) {
            op.apply(Vb[j].data(), W.data(), 0);
            for (int k = 0; k <= j; ++k) {
                const cx<double> d = dotc(Vb[k], W);
                Hh(k, j) = d;
                axpy(W, mk(-d.re, -d.im), Vb[k]);
            }
            const double wn = std::sqrt(dotc(W, W).re);
            Hh(j + 1, j) = mk(wn, 0.0);
            scal(W, mk(1.0 / wn, 0.0));
            if (j < m - 1) Vb[j + 1] = W; else Rv = W;
        }
        maxerr = 0;
        for (int n = 0; n < nshift; ++n) {
            if (blen0[n] < accuracy * blength) continue;
            for (int r = 0; r < m; ++r)
                for (int c = 0; c < m; ++c) hc[r * m + c] = Hh(r, c);
            for (int k = 0; k < m; ++k) hc[k * m + k] = hc[k * m + k] + mk(shifts[2 * n], shifts[2 * n + 1]);
            // RotateHSolveY: square FOM system (H + c_n) y = beta0_n e_1, by Givens + back substitution
            for (int k = 0; k < m; ++k) g[k] = mk(0.0, 0.0);
            g[0] = beta0[n];
            for (int k = 0; k < m - 1; ++k) {
                const double den = 1.0 / std::sqrt(abs2(hc[k * m + k]) + abs2(hc[(k + 1) * m + k]));
                const cx<double> cs = scale(hc[k * m + k], den), sn = scale(hc[(k + 1) * m + k], den);
                const cx<double> csh = conj(cs), snh = conj(sn);
                for (int j = k; j < m; ++j) {
                    const cx<double> hkj = hc[k * m + j];
                    hc[k * m + j] = csh * hkj + snh * hc[(k + 1) * m + j];
                    hc[(k + 1) * m + j] = cs * hc[(k + 1) * m + j] - sn * hkj;
                }
                const cx<double> mg = mk(-g[k].re, -g[k].im);
                g[k] = csh * g[k];
                g[k + 1] = sn * mg;
            }
            for (int k = m - 1; k >= 0; --k) {
                for (int j = k + 1; j < m; ++j) g[k] = g[k] - hc[k * m + j] * y[j];
                y[k] = cdiv(g[k], hc[k * m + k]);
            }
            beta0[n] = scale(y[m - 1], -Hh(m, m - 1).re);
            const double e = std::sqrt(abs2(beta0[n]));
            blen0[n] = e;
            if (e > maxerr) maxerr = e;
            for (int j = 0; j < m; ++j) axpy(X[n], y[j], Vb[j]);
        }
        if (maxerr < accuracy * blength) break;
    }
    for (int n = 0; n < nshift; ++n) std::memcpy(xs + (long)n * n2, X[n].data(), sizeof(double) * n2);
    if (cycles_out) *cycles_out = cyc;
    if (max_err_out) *max_err_out = maxerr;
    return maxerr < accuracy * blength ? 0 : 1;
}

// ---------------- Wilson force kernel, CFieldFermionWilsonSquareSU3.cu:170-231 ----------------
// makeSU3Contract(left, right)[3r+c] = sum_s conj(left_s[c]) right_s[r]   (SU3.h:395-503)
int orc_wilson_force_kernel(const int dims[4], const int bc[4], double kappa, const double* U,
                            const double* inverse_d, const double* inverse_dddag, double* Fo) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    auto contract = [](const Spinor<double>& l, const Spinor<double>& r) {
        Mat<double> o;
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                cx<double> acc = mk(0.0, 0.0);
                for (int s = 0; s < 4; ++s) acc = acc + conj(l.c[s][c]) * r.c[s][a];
                o.m[3 * a + c] = acc;
            }
        return o;
    };
#pragma omp parallel for schedule(static)
    for (long s = 0; s < lat.V; ++s) {
        const Spinor<double> xl = load_sp(inverse_dddag + s * 24);
        const Spinor<double> xr = load_sp(inverse_d + s * 24);
        for (int mu = 0; mu < 4; ++mu) {
            const long link = s * 4 + mu;
            const long n = m.fwd[link];
            const Spinor<double> pr = load_sp(inverse_d + n * 24);
            const Spinor<double> pl = load_sp(inverse_dddag + n * 24);
            Spinor<double> r1 = pr;
            ssub(r1, gmul(mu, pr));
            Mat<double> mid = contract(xl, r1);
            Spinor<double> r2 = xr;
            sadd(r2, gmul(mu, xr));
            const Mat<double> m2 = contract(r2, pl);
            for (int k = 0; k < 9; ++k) mid.m[k] = mid.m[k] + m2.m[k];
            Mat<double> f = mmul(load_mat(U + link * 18), mid);
            su3_ta(f);
            const double k = m.fsg[link] < 0 ? -kappa : kappa;
            for (int q = 0; q < 9; ++q) {
                Fo[link * 18 + 2 * q] += f.m[q].re * k;
                Fo[link * 18 + 2 * q + 1] += f.m[q].im * k;
            }
        }
    }
    return 0;
}

// ---------------- KS force kernel, CFieldFermionKSSU3.cu:216-264 ----------------
// makeSU3ContractV(left, right)[3r+c] = conj(left[c]) right[r]   (SU3.h:505-523)
int orc_ks_force_kernel(const int dims[4], const int bc[4], const double* U, const double* phis,
                        const double* dphis, const double* numer, int npole, double* Fo) {
    if (!dims_ok(dims) || npole < 1) return -1;
    Lat lat(dims);
    Moves m = build_moves(lat, bc);
    std::vector<unsigned char> eta = make_eta(lat);
    const long V = lat.V;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < V; ++s) {
        for (int mu = 0; mu < 4; ++mu) {
            const double e = ((eta[s] >> mu) & 1) ? -1.0 : 1.0;
            const long link = s * 4 + mu;
            const long n = m.fwd[link];
            const Mat<double> Ux = load_mat(U + link * 18);
            for (int r = 0; r < npole; ++r) {
                const double* phi = phis + (long)r * V * 6;
                const double* phd = dphis + (long)r * V * 6;
                Vec3<double> tc = mv(Ux, load_v(phi + n * 6));
                const Vec3<double> pd = load_v(phd + s * 6);
                Mat<double> term;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) term.m[3 * a + c] = conj(pd.c[c]) * tc.c[a];
                tc = mv(Ux, load_v(phd + n * 6));
                const Vec3<double> pi = load_v(phi + s * 6);
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) term.m[3 * a + c] = term.m[3 * a + c] + conj(tc.c[c]) * pi.c[a];
                const double f = e * numer[r] * (m.fsg[link] < 0 ? -1.0 : 1.0);
                for (int q = 0; q < 9; ++q) term.m[q] = scale(term.m[q], f);
                su3_ta(term);
                for (int q = 0; q < 9; ++q) {
                    Fo[link * 18 + 2 * q] -= term.m[q].re;
                    Fo[link * 18 + 2 * q + 1] -= term.m[q].im;
                }
            }
        }
    }
    return 0;
}


// ---------------- gauge action (CActionGaugePlaquette.cpp:14-107, CFieldGaugeSU3.cu:194-371) ----------------
// sum_x sum_{mu<nu} (3 - ReTr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag), periodic links
int orc_plaquette_sum(const int dims[4], const double* U, double* out) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    const int per[4] = {1, 1, 1, 1};
    Moves m = build_moves(lat, per);
    long double acc = 0;
    for (long s = 0; s < lat.V; ++s)
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                const long xm = m.fwd[s * 4 + mu], xn = m.fwd[s * 4 + nu];
                Mat<double> P = mmul(load_mat(U + (s * 4 + mu) * 18), load_mat(U + (xm * 4 + nu) * 18));
                P = mmul(P, dag(load_mat(U + (xn * 4 + mu) * 18)));
                P = mmul(P, dag(load_mat(U + (s * 4 + nu) * 18)));
                acc += 3.0 - (P.m[0].re + P.m[4].re + P.m[8].re);
            }
    *out = (double)acc;
    return 0;
}

// F_mu(x) += -(beta/6) Ta[U_mu(x) C_mu(x)], C = sum over the 6 staples that close the plaquettes
// through the link (_kernelStapleAtSiteSU3CacheIndex, CFieldGaugeSU3.cu:194-257, betaOverN * -1/2)
int orc_gauge_force(const int dims[4], const double* U, double beta, double* Fo) {
    if (!dims_ok(dims)) return -1;
    Lat lat(dims);
    const int per[4] = {1, 1, 1, 1};
    Moves m = build_moves(lat, per);
    for (long s = 0; s < lat.V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            Mat<double> C;
            for (int k = 0; k < 9; ++k) C.m[k] = mk(0.0, 0.0);
            const long xpm = m.fwd[s * 4 + mu];
            for (int nu = 0; nu < 4; ++nu) {
                if (nu == mu) continue;
                const long xpn = m.fwd[s * 4 + nu], xmn = m.bwd[s * 4 + nu];
                const long xpm_mn = m.bwd[xpm * 4 + nu];
                // upper: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
                Mat<double> up = mmul(load_mat(U + (xpm * 4 + nu) * 18), dag(load_mat(U + (xpn * 4 + mu) * 18)));
                up = mmul(up, dag(load_mat(U + (s * 4 + nu) * 18)));
                // lower: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
                Mat<double> lo = mmul(dag(load_mat(U + (xpm_mn * 4 + nu) * 18)), dag(load_mat(U + (xmn * 4 + mu) * 18)));
                lo = mmul(lo, load_mat(U + (xmn * 4 + nu) * 18));
                for (int k = 0; k < 9; ++k) C.m[k] = C.m[k] + up.m[k] + lo.m[k];
            }
            Mat<double> f = mmul(load_mat(U + (s * 4 + mu) * 18), C);
            su3_ta(f);
            for (int q = 0; q < 9; ++q) {
                Fo[(s * 4 + mu) * 18 + 2 * q] += -beta / 6.0 * f.m[q].re;
                Fo[(s * 4 + mu) * 18 + 2 * q + 1] += -beta / 6.0 * f.m[q].im;
            }
        }
    return 0;
}


// ---------------- HMC support: exponentials, kinetic energy, integrators ----------------
namespace {

// SU(2) Cayley map embedded at rows/cols (i, j): for X = [[i y, z], [-z*, -i y]],
// (1 + X)(1 - X)^-1 = ((1 - l) + 2X) / (1 + l), l = y^2 + |z|^2
Mat<double> cayley_embed(int i, int j, double y, cx<double> z) {
    Mat<double> r;
    for (int k = 0; k < 9; ++k) r.m[k] = mk(0.0, 0.0);
    r.m[0] = r.m[4] = r.m[8] = mk(1.0, 0.0);
    const double l = y * y + abs2(z), d = 1.0 / (1.0 + l);
    const cx<double> c = mk(d * (1.0 - l), 2.0 * y * d);
    const cx<double> o = scale(z, 2.0 * d);
    r.m[3 * i + i] = c;
    r.m[3 * j + j] = conj(c);
    r.m[3 * i + j] = o;
    r.m[3 * j + i] = mk(-o.re, o.im);
    return r;
}

// SU3::QuickExp (SU3.h:1361-1640), for traceless anti-Hermitian M: the symmetric product
// C01 C02 C12 C02 C01 of SU(2) Cayley maps; C01 and C02 carry a quarter, C12 half of the
// off-diagonal entries, and the diagonal is split so the product reproduces exp(a M) to O(a^3)
Mat<double> quick_exp(const Mat<double>& M, double a) {
    const double y1 = a * (M.m[0].im - M.m[4].im) / 12.0;
    const double y2 = a * (M.m[0].im - M.m[8].im) / 12.0;
    const double y3 = a * (M.m[4].im - M.m[8].im) / 6.0;
    const Mat<double> c01 = cayley_embed(0, 1, y1, scale(M.m[1], a / 4.0));
    const Mat<double> c02 = cayley_embed(0, 2, y2, scale(M.m[2], a / 4.0));
    const Mat<double> c12 = cayley_embed(1, 2, y3, scale(M.m[5], a / 2.0));
    return mmul(c01, mmul(c02, mmul(c12, mmul(c02, c01))));
}

// SU3::ExpReal (SU3.h:1286-1318): Horner form of the order-n Taylor polynomial of exp(a M)
Mat<double> taylor_exp(const Mat<double>& M, double a, int n) {
    Mat<double> t;
    for (int k = 0; k < 9; ++k) t.m[k] = mk(0.0, 0.0);
    t.m[0] = t.m[4] = t.m[8] = mk(1.0, 0.0);
    for (int i = 0; i < n; ++i) {
        Mat<double> s;
        for (int k = 0; k < 9; ++k) s.m[k] = scale(M.m[k], a / (double)(n - i));
        t = mmul(s, t);
        t.m[0].re += 1.0; t.m[4].re += 1.0; t.m[8].re += 1.0;
    }
    return t;
}

Mat<double> su3_exp(const Mat<double>& M, double a, int prec) {
    return prec == 0 ? quick_exp(M, a) : taylor_exp(M, a, prec);
}

struct Hmc {
    const int* dims;
    long V;
    double* U;
    double* P;
    const orc_action* acts;
    int nact;
    int exp_prec;
    std::vector<double> F, Ubak;

    // U <- exp(a M) U for every link (CFieldGaugeSU3::ExpMult, CFieldGaugeSU3.cu:452-490,1457-1477)
    void exp_mult(const double* M, double a, double* W) {
        for (long l = 0; l < V * 4; ++l) {
            const Mat<double> e = su3_exp(load_mat(M + 18 * l), a, exp_prec);
            store_mat(W + 18 * l, mmul(e, load_mat(W + 18 * l)));
        }
    }
    void fermion_force(const orc_action& A, double* Fo) {
        if (A.kind == 1) {  // Wilson Nf=2, CFieldFermionWilsonSquareSU3.cu:1016-1100
            const long n = V * 24;
            std::vector<double> X(n), Y(n);
            orc_operator op{0, ORC_DDDAG, {dims[0], dims[1], dims[2], dims[3]},
                            {A.bc[0], A.bc[1], A.bc[2], A.bc[3]}, A.kappa_or_mass};
            std::vector<double> zero(n, 0.0);
            int cyc; double dev;
            orc_gmres(&op, U, A.phi, X.data(), zero.data(), 30, 400, A.accuracy, 1.0, 1, &cyc, &dev);
            orc_wilson_op(dims, A.bc, A.kappa_or_mass, U, X.data(), Y.data(), ORC_DDAG, 0, 1.0, 0.0);
            orc_wilson_force_kernel(dims, A.bc, A.kappa_or_mass, U, Y.data(), X.data(), Fo);
        } else if (A.kind == 2) {  // KS rational, CFieldFermionKSSU3.cu:412-497
            const int np = A.md_degree;
            const long n = V * 6;
            std::vector<double> xs((size_t)np * n), dx((size_t)np * n), sh(2 * np);
            for (int k = 0; k < np; ++k) { sh[2 * k] = A.md[1 + np + k]; sh[2 * k + 1] = 0.0; }
            orc_operator op{1, ORC_DDDAG, {dims[0], dims[1], dims[2], dims[3]},
                            {A.bc[0], A.bc[1], A.bc[2], A.bc[3]}, A.kappa_or_mass};
            double b2[2];
            orc_dot(A.phi, A.phi, V * 3, b2);
            int cyc; double err;
            orc_msfom(&op, U, A.phi, xs.data(), sh.data(), np, 40, 400, A.accuracy, std::sqrt(b2[0]), &cyc, &err);
            for (int k = 0; k < np; ++k)
                orc_ks_op(dims, A.bc, 0.0, U, xs.data() + (size_t)k * n, dx.data() + (size_t)k * n, ORC_D, 0, 1.0, 0.0);
            orc_ks_force_kernel(dims, A.bc, U, xs.data(), dx.data(), A.md + 1, np, Fo);
        }
    }
    // UpdatePF / UpdatePG (CIntegrator.cpp:305-357): P += step * F(fermion or gauge actions)
    void update_p(double step, bool fermion) {
        std::fill(F.begin(), F.end(), 0.0);
        force_into(F.data(), fermion);
        for (long i = 0; i < V * 72; ++i) P[i] += step * F[i];
    }
    void force_into(double* Fo, bool fermion) {
        for (int i = 0; i < nact; ++i) {
            if (fermion && acts[i].kind != 0) fermion_force(acts[i], Fo);
            if (!fermion && acts[i].kind == 0) orc_gauge_force(dims, U, acts[i].beta, Fo);
        }
    }
    // force-gradient step: P += step * F(exp(fg F) U), U restored afterwards
    void fg_update(double fg, double step, bool fermion) {
        std::fill(F.begin(), F.end(), 0.0);
        force_into(F.data(), fermion);
        std::copy(U, U + V * 72, Ubak.begin());
        exp_mult(F.data(), fg, U);
        update_p(step, fermion);
        std::copy(Ubak.begin(), Ubak.end(), U);
    }
    void update_u(double step) { exp_mult(P, step, U); }

    // CIntegratorNestedForceGradient::NestedEvaluate (CIntegratorNestedForceGradient.cpp:110-157)
    void nested_fg(double h, int nsteps) {
        update_p(h / 6.0, false);
        for (int s = 1; s <= nsteps; ++s) {
            update_u(0.5 * h);
            fg_update(h * h / 24.0, 2.0 * h / 3.0, false);
            update_u(0.5 * h);
            update_p(s < nsteps ? h / 3.0 : h / 6.0, false);
        }
    }
    // CNestedIntegrator::NestedEvaluateLeapfrog (CIntegrator.cpp:359-380)
    void nested_leapfrog(double h, int nsteps) {
        update_p(0.5 * h, false);
        for (int s = 1; s <= nsteps; ++s) {
            update_u(h);
            update_p(s < nsteps ? h : 0.5 * h, false);
        }
    }
    double energy() {
        double e = 0.0;
        for (long i = 0; i < V * 72; ++i) e += P[i] * P[i];
        for (int i = 0; i < nact; ++i) {
            const orc_action& A = acts[i];
            if (A.kind == 0) {
                double ps;
                orc_plaquette_sum(dims, U, &ps);
                e += A.beta / 3.0 * ps;
            } else if (A.kind == 1) {  // |D^-1 phi|^2, CActionFermionWilsonNf2.cpp:53-68
                const long n = V * 24;
                std::vector<double> X(n), zero(n, 0.0);
                orc_operator op{0, ORC_D, {dims[0], dims[1], dims[2], dims[3]},
                                {A.bc[0], A.bc[1], A.bc[2], A.bc[3]}, A.kappa_or_mass};
                int cyc; double dev;
                orc_gmres(&op, U, A.phi, X.data(), zero.data(), 30, 400, A.accuracy, 1.0, 1, &cyc, &dev);
                double d[2];
                orc_dot(X.data(), X.data(), n / 2, d);
                e += d[0];
            } else {  // Re phi^dag r_MD(DD^dag) phi, CActionFermionKS.cpp:52-71
                const int np = A.md_degree;
                const long n = V * 6;
                std::vector<double> xs((size_t)np * n), sh(2 * np);
                for (int k = 0; k < np; ++k) { sh[2 * k] = A.md[1 + np + k]; sh[2 * k + 1] = 0.0; }
                orc_operator op{1, ORC_DDDAG, {dims[0], dims[1], dims[2], dims[3]},
                                {A.bc[0], A.bc[1], A.bc[2], A.bc[3]}, A.kappa_or_mass};
                double b2[2];
                orc_dot(A.phi, A.phi, V * 3, b2);
                int cyc; double err;
                orc_msfom(&op, U, A.phi, xs.data(), sh.data(), np, 40, 400, A.accuracy, std::sqrt(b2[0]), &cyc, &err);
                double acc = A.md[0] * b2[0];
                for (int k = 0; k < np; ++k) {
                    double d[2];
                    orc_dot(A.phi, xs.data() + (size_t)k * n, V * 3, d);
                    acc += A.md[1 + k] * d[0];
                }
                e += acc;
            }
        }
        return e;
    }
};

}  // namespace

int orc_su3_exp(const double* M, double a, int prec, double* out) {
    if (!M || !out || prec < 0) return -1;
    store_mat(out, su3_exp(load_mat(M), a, prec));
    return 0;
}

int orc_hmc_evolve(const int dims[4], double* U, double* P, const orc_action* acts, int nact, int integrator,
                   double tau, int steps, int nested, int exp_prec, double* h_before, double* h_after) {
    if (!dims_ok(dims) || !U || !P || nact < 1 || steps < 1 || nested < 1 || exp_prec < 0) return -1;
    Hmc h;
    h.dims = dims;
    h.V = Lat(dims).V;
    h.U = U;
    h.P = P;
    h.acts = acts;
    h.nact = nact;
    h.exp_prec = exp_prec;
    h.F.assign(h.V * 72, 0.0);
    h.Ubak.assign(h.V * 72, 0.0);
    if (h_before) *h_before = h.energy();
    const double eps = tau / steps;
    if (integrator == ORC_INT_LEAPFROG) {
        // CIntegratorLeapFrog::Evaluate (CIntegratorLeapFrog.h): all actions in one force
        auto update_all = [&](double st) {
            std::fill(h.F.begin(), h.F.end(), 0.0);
            h.force_into(h.F.data(), false);
            h.force_into(h.F.data(), true);
            for (long i = 0; i < h.V * 72; ++i) P[i] += st * h.F[i];
        };
        update_all(0.5 * eps);
        for (int s = 1; s <= steps; ++s) {
            h.update_u(eps);
            update_all(s < steps ? eps : 0.5 * eps);
        }
    } else if (integrator == ORC_INT_NESTED_FORCE_GRADIENT || integrator == ORC_INT_NESTED_FG_INNER_LEAPFROG) {
        // CIntegratorNestedForceGradient::Evaluate (CIntegratorNestedForceGradient.cpp:26-108);
        // the nested step is half the outer one over `nested` sub-steps (its Initial, :18-23)
        const double hn = 0.5 * eps / nested;
        auto inner = [&]() {
            if (integrator == ORC_INT_NESTED_FORCE_GRADIENT) h.nested_fg(hn, nested);
            else h.nested_leapfrog(hn, nested);
        };
        h.update_p(eps / 6.0, true);
        for (int s = 1; s <= steps; ++s) {
            inner();
            h.fg_update(eps * eps / 24.0, 2.0 * eps / 3.0, true);
            inner();
            h.update_p(s < steps ? eps / 3.0 : eps / 6.0, true);
        }
    } else {
        return -1;
    }
    // FinishEvaluate (CIntegrator.cpp:244-254): re-unitarise every link
    orc_su3_norm(U, h.V * 4);
    if (h_after) *h_after = h.energy();
    return 0;
}


// ---------------- GCRO-DR, SVD deflation (CSolverGCRODR.cpp:196-700, CLinearAlgebraHelper.h) ----------------
namespace {

typedef std::vector<std::vector<cx<double>>> DM;  // dense row-major complex matrix
DM dm(int r, int c) { return DM(r, std::vector<cx<double>>(c, mk(0.0, 0.0))); }
DM dmul(const DM& A, const DM& B) {
    DM C = dm((int)A.size(), (int)B[0].size());
    for (size_t i = 0; i < A.size(); ++i)
        for (size_t k = 0; k < B.size(); ++k)
            for (size_t j = 0; j < B[0].size(); ++j) C[i][j] = C[i][j] + A[i][k] * B[k][j];
    return C;
}
DM dadj(const DM& A) {
    DM B = dm((int)A[0].size(), (int)A.size());
    for (size_t i = 0; i < A.size(); ++i)
        for (size_t j = 0; j < A[0].size(); ++j) B[j][i] = conj(A[i][j]);
    return B;
}
double cabs(cx<double> a) { return std::sqrt(a.re * a.re + a.im * a.im); }

// RotateHenssenbergHost + SolveYHost: Givens least squares of an upper-Hessenberg (m+1) x m system
std::vector<cx<double>> lsq(DM H, std::vector<cx<double>> g) {
    const int m = (int)H[0].size();
    for (int k = 0; k < m; ++k) {
        const cx<double> a = H[k][k], b = H[k + 1][k];
        const double den = std::sqrt(abs2(a) + abs2(b));
        const cx<double> c = scale(a, 1.0 / den), s = scale(b, 1.0 / den);
        for (int j = k; j < m; ++j) {
            const cx<double> x = H[k][j], y = H[k + 1][j];
            H[k][j] = conj(c) * x + conj(s) * y;
            H[k + 1][j] = c * y - s * x;
        }
        const cx<double> x = g[k], y = g[k + 1];
        g[k] = conj(c) * x + conj(s) * y;
        g[k + 1] = c * y - s * x;
    }
    std::vector<cx<double>> y(m);
    for (int k = m - 1; k >= 0; --k) {
        cx<double> acc = g[k];
        for (int j = k + 1; j < m; ++j) acc = acc - H[k][j] * y[j];
        y[k] = cdiv(acc, H[k][k]);
    }
    return y;
}
// Householder-free thin QR by modified Gram-Schmidt (ThinQRFactorizationHost)
void dqr(const DM& T, DM& Q, DM& R) {
    const int n = (int)T.size(), k = (int)T[0].size();
    Q = T;
    R = dm(k, k);
    for (int i = 0; i < k; ++i) {
        double nn = 0;
        for (int r = 0; r < n; ++r) nn += abs2(Q[r][i]);
        nn = std::sqrt(nn);
        R[i][i] = mk(nn, 0.0);
        for (int r = 0; r < n; ++r) Q[r][i] = scale(Q[r][i], 1.0 / nn);
        for (int j = i + 1; j < k; ++j) {
            cx<double> d = mk(0.0, 0.0);
            for (int r = 0; r < n; ++r) d = d + conj(Q[r][i]) * Q[r][j];
            R[i][j] = d;
            for (int r = 0; r < n; ++r) Q[r][j] = Q[r][j] - d * Q[r][i];
        }
    }
}
// Hermitian eigenproblem by Jacobi sweeps; eigenvectors as columns of V, eigenvalues ascending
void djacobi(DM A, std::vector<double>& w, DM& V) {
    const int n = (int)A.size();
    V = dm(n, n);
    for (int i = 0; i < n; ++i) V[i][i] = mk(1.0, 0.0);
    for (int sweep = 0; sweep < 80; ++sweep) {
        double off = 0, tot = 0;
        for (int p = 0; p < n; ++p)
            for (int q = 0; q < n; ++q) { tot += abs2(A[p][q]); if (p != q) off += abs2(A[p][q]); }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double b = cabs(A[p][q]);
                if (b == 0.0) continue;
                // unitary 2x2 rotation that zeroes A[p][q]: phase-align, then a real rotation
                const cx<double> e = scale(A[p][q], 1.0 / b);
                const double tau = (A[q][q].re - A[p][p].re) / (2.0 * b);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                // columns: new_p = c A_p - sn conj(e) A_q ; new_q = sn e... written as J = [[c, sn], [-sn e*, c e*]]
                const cx<double> J00 = mk(c, 0.0), J01 = mk(sn, 0.0), J10 = scale(conj(e), -sn), J11 = scale(conj(e), c);
                for (int r = 0; r < n; ++r) {
                    const cx<double> ap = A[r][p], aq = A[r][q];
                    A[r][p] = ap * J00 + aq * J10;
                    A[r][q] = ap * J01 + aq * J11;
                    const cx<double> vp = V[r][p], vq = V[r][q];
                    V[r][p] = vp * J00 + vq * J10;
                    V[r][q] = vp * J01 + vq * J11;
                }
                for (int r = 0; r < n; ++r) {
                    const cx<double> ap = A[p][r], aq = A[q][r];
                    A[p][r] = conj(J00) * ap + conj(J10) * aq;
                    A[q][r] = conj(J01) * ap + conj(J11) * aq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A[a][a].re < A[b][b].re; });
    DM Vs = dm(n, n);
    w.resize(n);
    for (int j = 0; j < n; ++j) {
        w[j] = A[idx[j]][idx[j]].re;
        for (int r = 0; r < n; ++r) Vs[r][j] = V[r][idx[j]];
    }
    V = Vs;
}
// k eigenvectors of A p = lam B p (A Hermitian, B Hermitian positive definite) with the smallest lam
DM dgen_small(const DM& A, const DM& B, int k) {
    const int n = (int)A.size();
    DM L = dm(n, n);
    for (int j = 0; j < n; ++j) {
        double d = B[j][j].re;
        for (int q = 0; q < j; ++q) d -= abs2(L[j][q]);
        L[j][j] = mk(std::sqrt(d), 0.0);
        for (int i = j + 1; i < n; ++i) {
            cx<double> s = B[i][j];
            for (int q = 0; q < j; ++q) s = s - L[i][q] * conj(L[j][q]);
            L[i][j] = scale(s, 1.0 / L[j][j].re);
        }
    }
    // C = L^-1 A L^-H
    auto lsolve = [&](DM M) {
        for (size_t c = 0; c < M[0].size(); ++c)
            for (int i = 0; i < n; ++i) {
                cx<double> s = M[i][c];
                for (int q = 0; q < i; ++q) s = s - L[i][q] * M[q][c];
                M[i][c] = scale(s, 1.0 / L[i][i].re);
            }
        return M;
    };
    const DM Cm = lsolve(dadj(lsolve(dadj(A))));
    DM Ch = dm(n, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ch[i][j] = scale(Cm[i][j] + conj(Cm[j][i]), 0.5);
    std::vector<double> w;
    DM Y;
    djacobi(Ch, w, Y);
    DM P = dm(n, k);
    for (int c = 0; c < k; ++c)
        for (int i = n - 1; i >= 0; --i) {  // p = L^-H y
            cx<double> s = Y[i][c];
            for (int q = i + 1; q < n; ++q) s = s - conj(L[q][i]) * P[q][c];
            P[i][c] = scale(s, 1.0 / L[i][i].re);
        }
    return P;
}
DM dupper_inv(const DM& R) {
    const int k = (int)R.size();
    DM X = dm(k, k);
    for (int j = 0; j < k; ++j) {
        X[j][j] = cdiv(mk(1.0, 0.0), R[j][j]);
        for (int i = j - 1; i >= 0; --i) {
            cx<double> acc = mk(0.0, 0.0);
            for (int l = i + 1; l <= j; ++l) acc = acc + R[i][l] * X[l][j];
            X[i][j] = cdiv(mk(-acc.re, -acc.im), R[i][i]);
        }
    }
    return X;
}
// out_i = sum_j in_j M[j][i]
std::vector<F> vmul(const std::vector<const F*>& in, const DM& M) {
    std::vector<F> out(M[0].size(), F(in[0]->size(), 0.0));
    for (size_t i = 0; i < M[0].size(); ++i)
        for (size_t j = 0; j < in.size(); ++j) axpy(out[i], M[j][i], *in[j]);
    return out;
}

}  // namespace

int orc_gcrodr(const orc_operator* A, const double* Ug, const double* b, double* x, double* urec, int* have_u,
               int m, int k, int restart, int recalc, double accuracy, int absolute, int phase, int* cycles_out,
               double* dev_out) {
    if (!A || !dims_ok(A->dims) || m < 5 || m > 31 || k < 2 || k > m - 2 || !urec || !have_u) return -1;
    Op op(A, Ug);
    const long n2 = op.n * 2;
    F B(b, b + n2);
    const double bl = absolute ? 1.0 : dotc(B, B).re;
    const double tol = accuracy * bl;
    std::vector<F> U(k, F(n2)), Cv(k, F(n2)), V(m - k + 1, F(n2));
    for (int i = 0; i < k; ++i) std::copy(urec + i * n2, urec + (i + 1) * n2, U[i].begin());
    auto W = [&](int j) -> F& { return j < k ? Cv[j] : V[j - k]; };
    auto Vv = [&](int j) -> F& { return j < k ? U[j] : V[j - k]; };
    auto applyA = [&](const F& in, F& out) { op.apply(in.data(), out.data()); };
    F X = B, R(n2), Wr(n2), T(n2);
    auto resid = [&](F& r) {  // r = b - A X
        applyA(X, r);
        for (long i = 0; i < n2; ++i) r[i] = B[i] - r[i];
    };
    auto nrm = [&](const F& v) { return std::sqrt(dotc(v, v).re); };
    int cycles = 0;
    double dev = 0;
    // GenerateCU: G P = Q R, U = [basis] P R^-1, C = [C V] Q
    auto gen_cu = [&](const DM& G, const DM& P, bool upd, bool after_gmres) {
        DM Q, Rr;
        dqr(dmul(G, P), Q, Rr);
        const DM PR = dmul(P, dupper_inv(Rr));
        std::vector<const F*> in;
        for (int j = 0; j < m; ++j) in.push_back(after_gmres ? &W(j) : &Vv(j));
        std::vector<F> nu = vmul(in, PR);
        if (upd) {
            std::vector<const F*> inc;
            for (int j = 0; j <= m; ++j) inc.push_back(&W(j));
            std::vector<F> nc = vmul(inc, Q);
            for (int i = 0; i < k; ++i) Cv[i] = nc[i];
        }
        for (int i = 0; i < k; ++i) U[i] = nu[i];
        *have_u = 1;
    };
    const bool reuse = *have_u && (phase == 1 || phase == 2);
    F& r = V[0];
    if (reuse) {
        // GenerateCUFirstTime: C = A U, MGS QR, U = U R^-1, r = b - A x
        DM Rr = dm(k, k);
        for (int i = 0; i < k; ++i) applyA(U[i], Cv[i]);
        for (int i = 0; i < k; ++i) {
            const double l = nrm(Cv[i]);
            Rr[i][i] = mk(l, 0.0);
            scal(Cv[i], mk(1.0 / l, 0.0));
            for (int j = i + 1; j < k; ++j) {
                Rr[i][j] = dotc(Cv[i], Cv[j]);
                axpy(Cv[j], mk(-Rr[i][j].re, -Rr[i][j].im), Cv[i]);
            }
        }
        std::vector<const F*> in;
        for (int i = 0; i < k; ++i) in.push_back(&U[i]);
        std::vector<F> nu = vmul(in, dupper_inv(Rr));
        for (int i = 0; i < k; ++i) U[i] = nu[i];
        resid(r);
        dev = nrm(r);
    } else {
        // FirstTimeGMERESSolve: one GMRES(m) cycle (MGS), basis in W
        DM H = dm(m + 1, m);
        F& v0 = W(0);
        resid(v0);
        const double beta = nrm(v0);
        scal(v0, mk(1.0 / beta, 0.0));
        for (int j = 0; j < m; ++j) {
            applyA(W(j), W(j + 1));
            for (int q = 0; q <= j; ++q) {
                const cx<double> d = dotc(W(q), W(j + 1));
                H[q][j] = d;
                axpy(W(j + 1), mk(-d.re, -d.im), W(q));
            }
            const double wn = nrm(W(j + 1));
            H[j + 1][j] = mk(wn, 0.0);
            scal(W(j + 1), mk(1.0 / wn, 0.0));
        }
        ++cycles;
        std::vector<cx<double>> g(m + 1, mk(0.0, 0.0));
        g[0] = mk(beta, 0.0);
        const std::vector<cx<double>> y = lsq(H, g);
        for (int j = 0; j < m; ++j) axpy(X, y[j], W(j));
        if (recalc > 1) {
            std::fill(Wr.begin(), Wr.end(), 0.0);
            for (int i = 0; i <= m; ++i) {
                cx<double> hy = mk(0.0, 0.0);
                for (int j = 0; j < m; ++j) hy = hy + H[i][j] * y[j];
                if (i == 0) hy.re -= beta;
                axpy(Wr, mk(-hy.re, -hy.im), W(i));
            }
        } else {
            resid(Wr);
        }
        dev = nrm(Wr);
        // FindPk1, SVD: smallest right singular vectors of H (eigenvectors of H^H H)
        DM I = dm(m, m);
        for (int i = 0; i < m; ++i) I[i][i] = mk(1.0, 0.0);
        const DM P = dgen_small(dmul(dadj(H), H), I, k);
        gen_cu(H, P, dev >= tol, true);
        r = Wr;
    }
    int status = 1;
    if (dev < tol) status = 0;
    for (int it = 0; it < restart && status != 0; ++it) {
        // OrthognalXR with coefficients from the unmodified r
        T = r;
        for (int i = 0; i < k; ++i) {
            const cx<double> c = dotc(Cv[i], T);
            axpy(X, c, U[i]);
            axpy(r, mk(-c.re, -c.im), Cv[i]);
        }
        DM G = dm(m + 1, m);
        for (int i = 0; i < k; ++i) {
            const double l = 1.0 / nrm(U[i]);
            G[i][i] = mk(l, 0.0);
            scal(U[i], mk(l, 0.0));
        }
        dev = nrm(r);
        scal(r, mk(1.0 / dev, 0.0));
        for (int j = k; j < m; ++j) {
            applyA(W(j), T);
            W(j + 1) = T;
            for (int q = 0; q < k; ++q) {  // CGS against C (coefficients from A v_j)
                const cx<double> d = dotc(Cv[q], T);
                G[q][j] = d;
                axpy(W(j + 1), mk(-d.re, -d.im), Cv[q]);
            }
            for (int q = k; q <= j; ++q) {  // MGS against V
                const cx<double> d = dotc(W(q), W(j + 1));
                G[q][j] = d;
                axpy(W(j + 1), mk(-d.re, -d.im), W(q));
            }
            const double wn = nrm(W(j + 1));
            G[j + 1][j] = mk(wn, 0.0);
            scal(W(j + 1), mk(1.0 / wn, 0.0));
        }
        ++cycles;
        Wr = r;
        scal(Wr, mk(dev, 0.0));
        std::vector<cx<double>> g(m + 1, mk(0.0, 0.0));
        g[k] = mk(dev, 0.0);
        const std::vector<cx<double>> y = lsq(G, g);
        for (int j = 0; j < m; ++j) axpy(X, y[j], Vv(j));
        if (recalc != 0 && it != 0 && it % recalc == 0) {
            resid(Wr);
        } else {
            for (int i = 0; i <= m; ++i) {
                cx<double> gy = mk(0.0, 0.0);
                for (int j = 0; j < m; ++j) gy = gy + G[i][j] * y[j];
                axpy(Wr, mk(-gy.re, -gy.im), W(i));
            }
        }
        dev = nrm(Wr);
        // FindPk2, SVD: G^H G p = lam diag(U^H U block, I) p, smallest lam
        DM Bm = dm(m, m);
        for (int i = 0; i < m; ++i) Bm[i][i] = mk(1.0, 0.0);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j)
                if (i != j) Bm[i][j] = dotc(U[i], U[j]);
        const DM P = dgen_small(dmul(dadj(G), G), Bm, k);
        if (dev >= tol || phase != 2) gen_cu(G, P, dev >= tol, false);
        if (dev < tol) { status = 0; break; }
        r = Wr;
    }
    std::copy(X.begin(), X.end(), x);
    for (int i = 0; i < k; ++i) std::copy(U[i].begin(), U[i].end(), urec + i * n2);
    if (cycles_out) *cycles_out = cycles;
    if (dev_out) *dev_out = dev;
    return status;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double orc_time_wilson_d(const int dims[4], int prec, int reps, int nthreads) {
    if (!dims_ok(dims) || reps < 1) return -1.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    Lat lat(dims);
    const int bc[4] = {1, 1, 1, -1};
    Moves m = build_moves(lat, bc);
    std::vector<double> U(lat.V * 72), in(lat.V * 24), out(lat.V * 24);
    orc_gauge_random(dims, 0.5, 1u, U.data());
    orc_wilson_gaussian(dims, 2u, in.data());
    double sec;
    if (prec == 0) {
        wilson_op<double>(lat, m, 0.1355, U.data(), in.data(), out.data(), ORC_D, 0, 1.0, 0.0);
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r)
            wilson_op<double>(lat, m, 0.1355, U.data(), in.data(), out.data(), ORC_D, 0, 1.0, 0.0);
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
        std::vector<float> Uf(U.begin(), U.end()), inf(in.begin(), in.end()), outf(out.size());
        wilson_op<float>(lat, m, 0.1355f, Uf.data(), inf.data(), outf.data(), ORC_D, 0, 1.0f, 0.0f);
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r)
            wilson_op<float>(lat, m, 0.1355f, Uf.data(), inf.data(), outf.data(), ORC_D, 0, 1.0f, 0.0f);
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return sec;
}

}  // extern "C"
