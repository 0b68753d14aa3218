// Checks of the host small dense linear algebra used by the deflated solver: every routine is
// tested against its defining identity on random complex matrices (residuals printed, the test
// driver compares them with bounds).
#include <cstdio>
#include <random>

#include "clg_smallla.h"

using namespace clg::sla;

static std::mt19937_64 rng(7);
static cx rnd() {
    std::normal_distribution<double> g;
    return cx(g(rng), g(rng));
}
static Mat random(int r, int c) {
    Mat m(r, c);
    for (auto& z : m.a) z = rnd();
    return m;
}

int main() {
    double worst_lsq = 0, worst_qr = 0, worst_h = 0, worst_g = 0, worst_gen = 0, worst_sel = 0;
    for (int trial = 0; trial < 20; ++trial) {
        const int n = 4 + trial % 25;
        // least squares on an upper-Hessenberg (n+1) x n matrix: normal equations H^H (H y - b) = 0
        Mat H = random(n + 1, n);
        for (int i = 0; i < n + 1; ++i)
            for (int j = 0; j < n; ++j)
                if (i > j + 1) H(i, j) = 0.0;
        std::vector<cx> b(n + 1), y;
        for (auto& z : b) z = rnd();
        const double res = hessenberg_lsq(H, b, y);
        Mat Y(n, 1);
        for (int i = 0; i < n; ++i) Y(i, 0) = y[i];
        Mat R = mul(H, Y);
        double rr = 0;
        for (int i = 0; i <= n; ++i) { R(i, 0) -= b[i]; rr += std::norm(R(i, 0)); }
        Mat ne = mul(adj(H), R);
        worst_lsq = std::max(worst_lsq, fro(ne) / fro(H) / std::sqrt(rr + 1e-300));
        worst_lsq = std::max(worst_lsq, std::fabs(std::sqrt(rr) - res) / (std::sqrt(rr) + 1e-300));
        // thin QR
        const int k = 2 + trial % 6;
        Mat T = random(n + 1, k), Q, Rr;
        thin_qr(T, Q, Rr);
        Mat QR = mul(Q, Rr);
        double d = 0;
        for (size_t i = 0; i < QR.a.size(); ++i) d = std::max(d, std::abs(QR.a[i] - T.a[i]));
        Mat QhQ = mul(adj(Q), Q);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) d = std::max(d, std::abs(QhQ(i, j) - (i == j ? 1.0 : 0.0)));
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < i; ++j) d = std::max(d, std::abs(Rr(i, j)));
        worst_qr = std::max(worst_qr, d);
        Mat Ri = upper_inverse(Rr), I = mul(Rr, Ri);
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) worst_qr = std::max(worst_qr, std::abs(I(i, j) - (i == j ? 1.0 : 0.0)));
        // Hermitian eigenproblem
        Mat A = random(n, n);
        Mat Hm = mul(adj(A), A);
        std::vector<double> w;
        Mat V;
        herm_eig(Hm, w, V);
        Mat AV = mul(Hm, V);
        for (int j = 0; j < n; ++j) {
            if (j > 0 && w[j] < w[j - 1]) worst_h = 1.0;
            for (int i = 0; i < n; ++i) worst_h = std::max(worst_h, std::abs(AV(i, j) - w[j] * V(i, j)) / fro(Hm));
        }
        // general eigenproblem
        std::vector<cx> lam;
        Mat X;
        gen_eig(A, lam, X);
        Mat AX = mul(A, X);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) worst_g = std::max(worst_g, std::abs(AX(i, j) - lam[j] * X(i, j)) / fro(A));
        // generalized Hermitian-definite and general-vs-HPD problems through select_eigvecs
        Mat Bm = mul(adj(random(n, n)), random(n, n));
        Mat B = mul(adj(Bm), Bm);
        for (int i = 0; i < n; ++i) B(i, i) += 1.0;
        for (int herm = 0; herm < 2; ++herm) {
            const Mat& AA = herm ? Hm : A;
            Mat P;
            if (!select_eigvecs(AA, &B, herm != 0, k < n ? k : n - 1, trial % 2 == 0, P)) worst_gen = 1.0;
            Mat AP = mul(AA, P), BP = mul(B, P);
            for (int j = 0; j < P.c; ++j) {
                // Rayleigh quotient lam = (p^H A p) / (p^H B p), residual |A p - lam B p|
                cx num(0.0, 0.0), den(0.0, 0.0);
                for (int i = 0; i < n; ++i) { num += std::conj(P(i, j)) * AP(i, j); den += std::conj(P(i, j)) * BP(i, j); }
                const cx l = num / den;
                double r = 0, s = 0;
                for (int i = 0; i < n; ++i) { r += std::norm(AP(i, j) - l * BP(i, j)); s += std::norm(AP(i, j)); }
                worst_gen = std::max(worst_gen, std::sqrt(r / (s + 1e-300)));
            }
        }
        // selection order: the smallest-modulus eigenvalues of a diagonal matrix
        Mat Dg(n, n);
        for (int i = 0; i < n; ++i) Dg(i, i) = cx(n - i, 0.25);
        Mat P;
        select_eigvecs(Dg, nullptr, false, 2, true, P);
        // columns must pick the last two diagonal entries (smallest |.|)
        const double e1 = std::abs(P(n - 1, 0)) + std::abs(P(n - 2, 1));
        worst_sel = std::max(worst_sel, std::fabs(e1 - 2.0));
    }
    std::printf("lsq %.3e qr %.3e herm %.3e gen %.3e geneig %.3e select %.3e\n", worst_lsq, worst_qr, worst_h,
                worst_g, worst_gen, worst_sel);
    return 0;
}
