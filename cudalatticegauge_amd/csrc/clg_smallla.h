// clg_smallla.h -- small dense complex FP64 linear algebra on the host (dimensions <= 32) for the
// deflated solver: Givens least squares on upper-Hessenberg matrices, thin QR, Hermitian and
// general eigenproblems, Hermitian-definite generalized eigenproblems.  Written for clarity
// at these sizes (microseconds per call); the matrices are row-major.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

namespace clg {
namespace sla {

typedef std::complex<double> cx;

struct Mat {
    int r = 0, c = 0;
    std::vector<cx> a;
    Mat() {}
    Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, cx(0.0, 0.0)) {}
    cx& operator()(int i, int j) { return a[(size_t)i * c + j]; }
    const cx& operator()(int i, int j) const { return a[(size_t)i * c + j]; }
    static Mat eye(int n) {
        Mat m(n, n);
        for (int i = 0; i < n; ++i) m(i, i) = 1.0;
        return m;
    }
};

inline Mat mul(const Mat& A, const Mat& B) {
    Mat C(A.r, B.c);
    for (int i = 0; i < A.r; ++i)
        for (int k = 0; k < A.c; ++k) {
            const cx aik = A(i, k);
            if (aik == cx(0.0, 0.0)) continue;
            for (int j = 0; j < B.c; ++j) C(i, j) += aik * B(k, j);
        }
    return C;
}
inline Mat adj(const Mat& A) {
    Mat B(A.c, A.r);
    for (int i = 0; i < A.r; ++i)
        for (int j = 0; j < A.c; ++j) B(j, i) = std::conj(A(i, j));
    return B;
}
inline double fro(const Mat& A) {
    double s = 0;
    for (const cx& z : A.a) s += std::norm(z);
    return std::sqrt(s);
}

// min_y |rhs - H y| for an (m+1) x m upper-Hessenberg H by Givens rotations; returns |residual|
// (the modulus of the last rotated right-hand side entry).  H and rhs are copies.
inline double hessenberg_lsq(Mat H, std::vector<cx> rhs, std::vector<cx>& y) {
    const int m = H.c;
    for (int k = 0; k < m; ++k) {
        const cx a = H(k, k), b = H(k + 1, k);
        const double den = std::sqrt(std::norm(a) + std::norm(b));
        if (den == 0.0) continue;
        const cx c = a / den, s = b / den;  // G = [[conj c, conj s], [-s, c]]
        for (int j = k; j < m; ++j) {
            const cx hk = H(k, j), h1 = H(k + 1, j);
            H(k, j) = std::conj(c) * hk + std::conj(s) * h1;
            H(k + 1, j) = c * h1 - s * hk;
        }
        const cx gk = rhs[k], g1 = rhs[k + 1];
        rhs[k] = std::conj(c) * gk + std::conj(s) * g1;
        rhs[k + 1] = c * g1 - s * gk;
    }
    y.assign(m, cx(0.0, 0.0));
    for (int k = m - 1; k >= 0; --k) {
        cx acc = rhs[k];
        for (int j = k + 1; j < m; ++j) acc -= H(k, j) * y[j];
        y[k] = acc / H(k, k);
    }
    return std::abs(rhs[m]);
}

// T (rows x k) = Q R with orthonormal Q (rows x k) and upper-triangular R (k x k); Gram-Schmidt
// with a second pass
inline void thin_qr(const Mat& T, Mat& Q, Mat& R) {
    const int n = T.r, k = T.c;
    Q = T;
    R = Mat(k, k);
    for (int i = 0; i < k; ++i) {
        for (int pass = 0; pass < 2; ++pass) {
            for (int j = 0; j < i; ++j) {
                cx d(0.0, 0.0);
                for (int r = 0; r < n; ++r) d += std::conj(Q(r, j)) * Q(r, i);
                R(j, i) += d;
                for (int r = 0; r < n; ++r) Q(r, i) -= d * Q(r, j);
            }
        }
        double nn = 0;
        for (int r = 0; r < n; ++r) nn += std::norm(Q(r, i));
        nn = std::sqrt(nn);
        R(i, i) = nn;
        for (int r = 0; r < n; ++r) Q(r, i) /= nn;
    }
}

// inverse of an upper-triangular matrix
inline Mat upper_inverse(const Mat& R) {
    const int k = R.r;
    Mat X(k, k);
    for (int j = 0; j < k; ++j) {
        X(j, j) = 1.0 / R(j, j);
        for (int i = j - 1; i >= 0; --i) {
            cx acc(0.0, 0.0);
            for (int l = i + 1; l <= j; ++l) acc += R(i, l) * X(l, j);
            X(i, j) = -acc / R(i, i);
        }
    }
    return X;
}

// Hermitian eigenproblem by cyclic Jacobi rotations: A = V diag(w) V^H, eigenvalues ascending
inline void herm_eig(Mat A, std::vector<double>& w, Mat& V) {
    const int n = A.r;
    V = Mat::eye(n);
    const double scale = std::max(fro(A), 1e-300);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += std::norm(A(p, q));
        if (std::sqrt(off) < 1e-16 * scale) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double b = std::abs(A(p, q));
                if (b < 1e-300) continue;
                const cx ph = A(p, q) / b;  // A(p,q) = b e^{i phi}
                const double app = A(p, p).real(), aqq = A(q, q).real();
                const double theta = (aqq - app) / (2.0 * b);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                // J = diag(1, conj(ph)) [[c, s], [-s, c]]: (J^H A J)(p, q) = 0
                const cx j00 = c, j01 = s, j10 = -s * std::conj(ph), j11 = c * std::conj(ph);
                for (int r = 0; r < n; ++r) {
                    const cx arp = A(r, p), arq = A(r, q);
                    A(r, p) = arp * j00 + arq * j10;
                    A(r, q) = arp * j01 + arq * j11;
                }
                for (int r = 0; r < n; ++r) {
                    const cx apr = A(p, r), aqr = A(q, r);
                    A(p, r) = std::conj(j00) * apr + std::conj(j10) * aqr;
                    A(q, r) = std::conj(j01) * apr + std::conj(j11) * aqr;
                }
                A(p, q) = A(q, p) = 0.0;
                A(p, p) = A(p, p).real();
                A(q, q) = A(q, q).real();
                for (int r = 0; r < n; ++r) {
                    const cx vrp = V(r, p), vrq = V(r, q);
                    V(r, p) = vrp * j00 + vrq * j10;
                    V(r, q) = vrp * j01 + vrq * j11;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return A(a, a).real() < A(b, b).real(); });
    w.resize(n);
    Mat Vs(n, n);
    for (int j = 0; j < n; ++j) {
        w[j] = A(idx[j], idx[j]).real();
        for (int r = 0; r < n; ++r) Vs(r, j) = V(r, idx[j]);
    }
    V = Vs;
}

// general complex eigenproblem: Householder Hessenberg reduction, single-shift QR iteration with
// Wilkinson shifts to the Schur form T = Z^H A Z, eigenvectors of T by back substitution
inline void gen_eig(Mat A, std::vector<cx>& lam, Mat& X) {
    const int n = A.r;
    Mat Z = Mat::eye(n);
    const double anorm = std::max(fro(A), 1e-300);
    for (int k = 0; k + 2 < n; ++k) {
        double xn = 0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(A(i, k));
        xn = std::sqrt(xn);
        if (xn < 1e-300) continue;
        std::vector<cx> v(n, cx(0.0, 0.0));
        const cx x0 = A(k + 1, k);
        const cx ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cx(1.0, 0.0);
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] += ph * xn;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        if (vn < 1e-300) continue;
        // A <- (I - 2 v v^H / |v|^2) A (I - 2 v v^H / |v|^2); Z <- Z (I - ...)
        for (int j = 0; j < n; ++j) {
            cx d(0.0, 0.0);
            for (int i = k + 1; i < n; ++i) d += std::conj(v[i]) * A(i, j);
            d *= 2.0 / vn;
            for (int i = k + 1; i < n; ++i) A(i, j) -= d * v[i];
        }
        for (int i = 0; i < n; ++i) {
            cx d(0.0, 0.0);
            for (int j = k + 1; j < n; ++j) d += A(i, j) * v[j];
            d *= 2.0 / vn;
            for (int j = k + 1; j < n; ++j) A(i, j) -= d * std::conj(v[j]);
        }
        for (int i = 0; i < n; ++i) {
            cx d(0.0, 0.0);
            for (int j = k + 1; j < n; ++j) d += Z(i, j) * v[j];
            d *= 2.0 / vn;
            for (int j = k + 1; j < n; ++j) Z(i, j) -= d * std::conj(v[j]);
        }
        for (int i = k + 2; i < n; ++i) A(i, k) = 0.0;
    }
    int hi = n - 1, iter = 0, total = 0;
    while (hi > 0 && total < 10000) {
        int l = hi;
        while (l > 0) {
            const double s = std::abs(A(l - 1, l - 1)) + std::abs(A(l, l));
            if (std::abs(A(l, l - 1)) <= 1e-16 * (s > 0 ? s : anorm)) {
                A(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == hi) {
            --hi;
            iter = 0;
            continue;
        }
        ++iter;
        ++total;
        // Wilkinson shift from the trailing 2x2 block
        const cx a = A(hi - 1, hi - 1), b = A(hi - 1, hi), c = A(hi, hi - 1), d = A(hi, hi);
        cx mu;
        if (iter % 11 == 10) {
            mu = d + cx(std::abs(A(hi, hi - 1)), 0.0);  // exceptional shift
        } else {
            const cx tr = a + d, det = a * d - b * c;
            const cx disc = std::sqrt(tr * tr * 0.25 - det);
            const cx e1 = tr * 0.5 + disc, e2 = tr * 0.5 - disc;
            mu = std::abs(e1 - d) < std::abs(e2 - d) ? e1 : e2;
        }
        // QR step on rows/cols l..hi of A - mu I by Givens rotations, applied to the full matrix
        std::vector<cx> cs(hi - l), sn(hi - l);
        for (int i = l; i <= hi; ++i) A(i, i) -= mu;
        for (int k = l; k < hi; ++k) {
            const cx x = A(k, k), y = A(k + 1, k);
            const double den = std::sqrt(std::norm(x) + std::norm(y));
            const cx cc = den > 0 ? x / den : cx(1.0, 0.0), ss = den > 0 ? y / den : cx(0.0, 0.0);
            cs[k - l] = cc;
            sn[k - l] = ss;
            for (int j = k; j < n; ++j) {
                const cx p = A(k, j), q = A(k + 1, j);
                A(k, j) = std::conj(cc) * p + std::conj(ss) * q;
                A(k + 1, j) = cc * q - ss * p;
            }
        }
        for (int k = l; k < hi; ++k) {
            const cx cc = cs[k - l], ss = sn[k - l];
            for (int i = 0; i <= std::min(hi, k + 2); ++i) {
                const cx p = A(i, k), q = A(i, k + 1);
                A(i, k) = p * cc + q * ss;
                A(i, k + 1) = -p * std::conj(ss) + q * std::conj(cc);
            }
            for (int i = 0; i < n; ++i) {
                const cx p = Z(i, k), q = Z(i, k + 1);
                Z(i, k) = p * cc + q * ss;
                Z(i, k + 1) = -p * std::conj(ss) + q * std::conj(cc);
            }
        }
        for (int i = l; i <= hi; ++i) A(i, i) += mu;
    }
    lam.resize(n);
    for (int i = 0; i < n; ++i) lam[i] = A(i, i);
    // eigenvectors of the triangular T, then X = Z Y
    Mat Y(n, n);
    const double small = 1e-15 * anorm;
    for (int i = 0; i < n; ++i) {
        Y(i, i) = 1.0;
        for (int j = i - 1; j >= 0; --j) {
            cx acc(0.0, 0.0);
            for (int k = j + 1; k <= i; ++k) acc += A(j, k) * Y(k, i);
            cx den = A(j, j) - lam[i];
            if (std::abs(den) < small) den = cx(small, 0.0);
            Y(j, i) = -acc / den;
        }
    }
    X = mul(Z, Y);
    for (int j = 0; j < n; ++j) {
        double nn = 0;
        for (int i = 0; i < n; ++i) nn += std::norm(X(i, j));
        nn = std::sqrt(nn);
        for (int i = 0; i < n; ++i) X(i, j) /= nn;
    }
}

// lower Cholesky factor of a Hermitian positive-definite B (returns false if it is not)
inline bool cholesky(const Mat& B, Mat& L) {
    const int n = B.r;
    L = Mat(n, n);
    for (int j = 0; j < n; ++j) {
        double d = B(j, j).real();
        for (int k = 0; k < j; ++k) d -= std::norm(L(j, k));
        if (!(d > 0)) return false;
        L(j, j) = std::sqrt(d);
        for (int i = j + 1; i < n; ++i) {
            cx s = B(i, j);
            for (int k = 0; k < j; ++k) s -= L(i, k) * std::conj(L(j, k));
            L(i, j) = s / L(j, j);
        }
    }
    return true;
}
// X = L^-1 M (forward substitution on each column)
inline Mat lower_solve(const Mat& L, const Mat& M) {
    const int n = L.r;
    Mat X = M;
    for (int c = 0; c < M.c; ++c)
        for (int i = 0; i < n; ++i) {
            cx s = X(i, c);
            for (int k = 0; k < i; ++k) s -= L(i, k) * X(k, c);
            X(i, c) = s / L(i, i);
        }
    return X;
}
// X = L^-H M (back substitution with the adjoint)
inline Mat lower_adj_solve(const Mat& L, const Mat& M) {
    const int n = L.r;
    Mat X = M;
    for (int c = 0; c < M.c; ++c)
        for (int i = n - 1; i >= 0; --i) {
            cx s = X(i, c);
            for (int k = i + 1; k < n; ++k) s -= std::conj(L(k, i)) * X(k, c);
            X(i, c) = s / L(i, i);
        }
    return X;
}

// the k eigenvectors (columns) of A p = lam B p whose eigenvalues are smallest (or largest) in
// modulus; B = nullptr means B = I; hermitian says A is Hermitian.  Returns false on failure.
inline bool select_eigvecs(const Mat& A, const Mat* B, bool hermitian, int k, bool smallest, Mat& P) {
    const int n = A.r;
    Mat C = A, L;
    if (B) {
        if (!cholesky(*B, L)) return false;
        C = lower_solve(L, adj(lower_solve(L, adj(A))));  // L^-1 A L^-H
    }
    std::vector<cx> lam(n);
    Mat Y;
    if (hermitian) {
        // symmetrise against round-off, then Jacobi
        Mat H = C;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) H(i, j) = 0.5 * (C(i, j) + std::conj(C(j, i)));
        std::vector<double> w;
        herm_eig(H, w, Y);
        for (int i = 0; i < n; ++i) lam[i] = w[i];
    } else {
        gen_eig(C, lam, Y);
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        return smallest ? std::abs(lam[a]) < std::abs(lam[b]) : std::abs(lam[a]) > std::abs(lam[b]);
    });
    Mat Ys(n, k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < n; ++i) Ys(i, j) = Y(i, idx[j]);
    P = B ? lower_adj_solve(L, Ys) : Ys;
    for (const cx& z : P.a)
        if (!std::isfinite(z.real()) || !std::isfinite(z.imag())) return false;
    return true;
}

}  // namespace sla
}  // namespace clg
