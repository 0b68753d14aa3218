// clg_smallla_dev.h -- small dense complex FP64 linear algebra on the device, for one workgroup
// (dimensions <= 32): least squares on upper-Hessenberg matrices, thin QR, triangular inverses,
// Cholesky factors and solves, Hermitian and general eigenproblems and the eigenvector selection of
// the deflated solver.  GCRO-DR's small problems run with it next to the solver's other device data,
// so a cycle does not copy its Hessenberg matrix to the host and back.
//
// Matrices are row-major in LDS with leading dimension kLD; every function is called by all threads
// of the workgroup (blockDim.x a multiple of 64, at least 64) and ends with a barrier, so its result
// is visible to all of them.  The algorithms are those of the host reference implementation the
// tests compare against (Givens least squares, Gram-Schmidt QR with a second pass, Jacobi rotations
// for Hermitian matrices, Householder reduction to Hessenberg form and single-shift QR iteration with
// Wilkinson shifts for general ones), arranged so that the work of a step is spread over the threads:
// the Jacobi sweeps rotate n/2 disjoint (p, q) pairs at once in round-robin order.
#pragma once

#include "clg_internal.h"

namespace clgd {
namespace sld {

constexpr int kLD = 32;
typedef C2<double> cz;

__device__ __forceinline__ cz zz() { return cz{0.0, 0.0}; }
__device__ __forceinline__ cz one() { return cz{1.0, 0.0}; }
__device__ __forceinline__ double nrm2(cz a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double mag(cz a) { return hypot(a.x, a.y); }
__device__ __forceinline__ cz cdiv(cz a, cz b) {
    const double d = b.x * b.x + b.y * b.y;
    return cz{(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__device__ __forceinline__ cz csqrt_(cz a) {
    const double r = mag(a);
    if (r == 0.0) return zz();
    double re = sqrt(0.5 * (r + fabs(a.x)));
    double im = a.y / (2.0 * re);
    if (a.x < 0) {
        const double t = re;
        re = fabs(im);
        im = a.y >= 0 ? t : -t;
    }
    return cz{re, im};
}
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ cz& at(cz* A, int i, int j) { return A[i * kLD + j]; }

// block-wide sum of one double per thread (the same order in every call), broadcast to all threads
__device__ inline double block_sum(double v, double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    v = wsum(v);
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += red[q];
    __syncthreads();
    return s;
}

__device__ inline void set_zero(cz* A, int r, int c) {
    for (int e = threadIdx.x; e < r * c; e += blockDim.x) at(A, e / c, e % c) = zz();
    __syncthreads();
}
__device__ inline void set_eye(cz* A, int n) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) at(A, e / n, e % n) = (e / n == e % n) ? one() : zz();
    __syncthreads();
}
__device__ inline void copy(cz* B, const cz* A, int r, int c) {
    for (int e = threadIdx.x; e < r * c; e += blockDim.x) B[(e / c) * kLD + e % c] = A[(e / c) * kLD + e % c];
    __syncthreads();
}
// B (c x r) = A^H for A (r x c)
__device__ inline void adjoint(cz* B, const cz* A, int r, int c) {
    for (int e = threadIdx.x; e < r * c; e += blockDim.x) {
        const int i = e / c, j = e % c;
        B[j * kLD + i] = cconj(A[i * kLD + j]);
    }
    __syncthreads();
}
// C (r x c) = A B with A (r x k), B (k x c); adjA: C = A^H B with A stored (k x r).  C is another matrix.
__device__ inline void matmul(cz* C, const cz* A, const cz* B, int r, int k, int c, bool adjA) {
    for (int e = threadIdx.x; e < r * c; e += blockDim.x) {
        const int i = e / c, j = e % c;
        cz s = zz();
        for (int l = 0; l < k; ++l) {
            const cz a = adjA ? cconj(A[l * kLD + i]) : A[i * kLD + l];
            cfma(s, a, B[l * kLD + j]);
        }
        C[i * kLD + j] = s;
    }
    __syncthreads();
}

// min_y |rhs - H y| for an (m+1) x m upper-Hessenberg H by Givens rotations (H and rhs are overwritten);
// y (m) receives the solution.  Returns |residual| in every thread.
__device__ inline double hessenberg_lsq(cz* H, cz* rhs, cz* y, int m, double* red) {
    __shared__ cz rot[2];
    __shared__ int skip;
    for (int k = 0; k < m; ++k) {
        if (threadIdx.x == 0) {
            const cz a = at(H, k, k), b = at(H, k + 1, k);
            const double den = sqrt(nrm2(a) + nrm2(b));
            skip = den == 0.0;
            if (!skip) {
                rot[0] = cscale(a, 1.0 / den);
                rot[1] = cscale(b, 1.0 / den);
            }
        }
        __syncthreads();
        if (!skip) {
            const cz c = rot[0], s = rot[1];  // G = [[conj c, conj s], [-s, c]]
            for (int j = k + (int)threadIdx.x; j < m; j += blockDim.x) {
                const cz hk = at(H, k, j), h1 = at(H, k + 1, j);
                at(H, k, j) = cadd(cmulc(c, hk), cmulc(s, h1));
                at(H, k + 1, j) = csub(cmul(c, h1), cmul(s, hk));
            }
            if (threadIdx.x == blockDim.x - 1) {
                const cz gk = rhs[k], g1 = rhs[k + 1];
                rhs[k] = cadd(cmulc(c, gk), cmulc(s, g1));
                rhs[k + 1] = csub(cmul(c, g1), cmul(s, gk));
            }
        }
        __syncthreads();
    }
    // back substitution, one wave: y_k = (rhs_k - sum_{j>k} H(k,j) y_j) / H(k,k)
    for (int k = m - 1; k >= 0; --k) {
        if (threadIdx.x < 64) {
            cz a = zz();
            for (int j = k + 1 + (int)threadIdx.x; j < m; j += 64) cfma(a, at(H, k, j), y[j]);
            a.x = wsum(a.x);
            a.y = wsum(a.y);
            if (threadIdx.x == 0) y[k] = cdiv(csub(rhs[k], a), at(H, k, k));
        }
        __syncthreads();
    }
    (void)red;
    return mag(rhs[m]);
}

// T (r x k) = Q R, Q (r x k) with orthonormal columns, R (k x k) upper triangular: classical
// Gram-Schmidt with a second pass.  Q and R are other matrices than T.
__device__ inline void thin_qr(const cz* T, cz* Q, cz* R, int r, int k, double* red) {
    __shared__ cz d[kLD];
    copy(Q, T, r, k);
    set_zero(R, k, k);
    for (int i = 0; i < k; ++i) {
        for (int pass = 0; pass < 2 && i > 0; ++pass) {
            for (int j = threadIdx.x; j < i; j += blockDim.x) {
                cz s = zz();
                for (int l = 0; l < r; ++l) cfma(s, cconj(at(Q, l, j)), at(Q, l, i));
                d[j] = s;
                at(R, j, i) = cadd(at(R, j, i), s);
            }
            __syncthreads();
            for (int l = threadIdx.x; l < r; l += blockDim.x) {
                cz q = at(Q, l, i);
                for (int j = 0; j < i; ++j) q = csub(q, cmul(d[j], at(Q, l, j)));
                at(Q, l, i) = q;
            }
            __syncthreads();
        }
        double part = 0.0;
        for (int l = threadIdx.x; l < r; l += blockDim.x) part += nrm2(at(Q, l, i));
        const double nn = sqrt(block_sum(part, red));
        for (int l = threadIdx.x; l < r; l += blockDim.x) at(Q, l, i) = cscale(at(Q, l, i), 1.0 / nn);
        if (threadIdx.x == 0) at(R, i, i) = cz{nn, 0.0};
        __syncthreads();
    }
}

// X (k x k) = R^-1 for upper-triangular R; thread j builds column j
__device__ inline void upper_inverse(const cz* R, cz* X, int k) {
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        for (int i = 0; i < k; ++i) at(X, i, j) = zz();
        at(X, j, j) = cdiv(one(), R[j * kLD + j]);
        for (int i = j - 1; i >= 0; --i) {
            cz acc = zz();
            for (int l = i + 1; l <= j; ++l) cfma(acc, R[i * kLD + l], at(X, l, j));
            at(X, i, j) = cdiv(cz{-acc.x, -acc.y}, R[i * kLD + i]);
        }
    }
    __syncthreads();
}

// lower Cholesky factor L of a Hermitian positive-definite B (n x n); false (in every thread) if B is not
__device__ inline bool cholesky(const cz* B, cz* L, int n, double* red) {
    __shared__ int bad;
    set_zero(L, n, n);
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        double part = 0.0;
        for (int l = threadIdx.x; l < j; l += blockDim.x) part += nrm2(at(L, j, l));
        const double dd = B[j * kLD + j].x - block_sum(part, red);
        if (!(dd > 0)) return false;  // uniform: every thread computed the same dd
        const double ljj = sqrt(dd);
        if (threadIdx.x == 0) at(L, j, j) = cz{ljj, 0.0};
        for (int i = j + 1 + (int)threadIdx.x; i < n; i += blockDim.x) {
            cz s = B[i * kLD + j];
            for (int l = 0; l < j; ++l) s = csub(s, cmul(at(L, i, l), cconj(at(L, j, l))));
            at(L, i, j) = cscale(s, 1.0 / ljj);
        }
        __syncthreads();
    }
    (void)bad;
    return true;
}
// X = L^-1 M (n x c; X may be M): thread per column, forward substitution
__device__ inline void lower_solve(const cz* L, const cz* M, cz* X, int n, int c) {
    for (int col = threadIdx.x; col < c; col += blockDim.x) {
        for (int i = 0; i < n; ++i) {
            cz s = M[i * kLD + col];
            for (int l = 0; l < i; ++l) s = csub(s, cmul(L[i * kLD + l], X[l * kLD + col]));
            X[i * kLD + col] = cdiv(s, L[i * kLD + i]);
        }
    }
    __syncthreads();
}
// X = L^-H M (n x c; X may be M): thread per column, back substitution with the adjoint
__device__ inline void lower_adj_solve(const cz* L, const cz* M, cz* X, int n, int c) {
    for (int col = threadIdx.x; col < c; col += blockDim.x) {
        for (int i = n - 1; i >= 0; --i) {
            cz s = M[i * kLD + col];
            for (int l = i + 1; l < n; ++l) s = csub(s, cmul(cconj(L[l * kLD + i]), X[l * kLD + col]));
            X[i * kLD + col] = cdiv(s, L[i * kLD + i]);
        }
    }
    __syncthreads();
}

// player of position pos in step s of the round-robin schedule of N (even) players
__device__ __forceinline__ int rr_player(int pos, int s, int N) { return pos == 0 ? 0 : 1 + (pos - 1 + s) % (N - 1); }

// Hermitian A (n x n, overwritten) = V diag(w) V^H by Jacobi rotations, eigenvalues ascending (w in LDS);
// tmp is a scratch matrix
__device__ inline void herm_eig(cz* A, cz* V, double* w, cz* tmp, int n, double* red) {
    struct Rot {
        int p, q;
        cz j00, j01, j10, j11;
    };
    __shared__ Rot rot[kLD / 2];
    set_eye(V, n);
    const int N = n + (n & 1);
    double part = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) part += nrm2(at(A, e / n, e % n));
    const double scale = fmax(sqrt(block_sum(part, red)), 1e-300);
    for (int sweep = 0; sweep < 100; ++sweep) {
        part = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x)
            if (e / n < e % n) part += nrm2(at(A, e / n, e % n));
        if (sqrt(block_sum(part, red)) < 1e-16 * scale) break;
        for (int s = 0; s < N - 1; ++s) {
            for (int i = threadIdx.x; i < N / 2; i += blockDim.x) {
                int p = rr_player(i, s, N), q = rr_player(N - 1 - i, s, N);
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                Rot r;
                r.p = -1;
                r.q = -1;
                if (q < n) {
                    const cz apq = at(A, p, q);
                    const double b = mag(apq);
                    if (b >= 1e-300) {
                        const cz ph = cscale(apq, 1.0 / b);
                        const double app = at(A, p, p).x, aqq = at(A, q, q).x;
                        const double theta = (aqq - app) / (2.0 * b);
                        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                        r.p = p;
                        r.q = q;
                        r.j00 = cz{c, 0.0};
                        r.j01 = cz{sn, 0.0};
                        r.j10 = cscale(cconj(ph), -sn);
                        r.j11 = cscale(cconj(ph), c);
                    }
                }
                rot[i] = r;
            }
            __syncthreads();
            // columns p, q of every pair (A <- A J), and of V
            for (int e = threadIdx.x; e < (N / 2) * n; e += blockDim.x) {
                const Rot& r = rot[e / n];
                const int row = e % n;
                if (r.p < 0) continue;
                const cz arp = at(A, row, r.p), arq = at(A, row, r.q);
                at(A, row, r.p) = cadd(cmul(arp, r.j00), cmul(arq, r.j10));
                at(A, row, r.q) = cadd(cmul(arp, r.j01), cmul(arq, r.j11));
                const cz vrp = at(V, row, r.p), vrq = at(V, row, r.q);
                at(V, row, r.p) = cadd(cmul(vrp, r.j00), cmul(vrq, r.j10));
                at(V, row, r.q) = cadd(cmul(vrp, r.j01), cmul(vrq, r.j11));
            }
            __syncthreads();
            // rows p, q of every pair (A <- J^H A)
            for (int e = threadIdx.x; e < (N / 2) * n; e += blockDim.x) {
                const Rot& r = rot[e / n];
                const int col = e % n;
                if (r.p < 0) continue;
                const cz apc = at(A, r.p, col), aqc = at(A, r.q, col);
                at(A, r.p, col) = cadd(cmulc(r.j00, apc), cmulc(r.j10, aqc));
                at(A, r.q, col) = cadd(cmulc(r.j01, apc), cmulc(r.j11, aqc));
            }
            __syncthreads();
            for (int i = threadIdx.x; i < N / 2; i += blockDim.x) {
                const Rot& r = rot[i];
                if (r.p < 0) continue;
                at(A, r.p, r.q) = zz();
                at(A, r.q, r.p) = zz();
                at(A, r.p, r.p) = cz{at(A, r.p, r.p).x, 0.0};
                at(A, r.q, r.q) = cz{at(A, r.q, r.q).x, 0.0};
            }
            __syncthreads();
        }
    }
    // ascending order: thread i places eigenpair i at its rank (ties by index)
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double wi = at(A, i, i).x;
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double wj = at(A, j, j).x;
            rank += (wj < wi || (wj == wi && j < i)) ? 1 : 0;
        }
        w[rank] = wi;
        for (int r = 0; r < n; ++r) tmp[r * kLD + rank] = at(V, r, i);
    }
    __syncthreads();
    copy(V, tmp, n, n);
}

// general A (n x n, overwritten by its Schur form) : eigenvalues lam, unit eigenvectors in the columns of X.
// Z and Y are scratch matrices.
__device__ inline void gen_eig(cz* A, cz* lam, cz* X, cz* Z, cz* Y, int n, double* red) {
    __shared__ cz v[kLD];
    __shared__ cz cs[kLD], sn[kLD];
    __shared__ int ctl[3];   // l, hi, action (0 deflate/continue, 1 QR step, 2 done)
    __shared__ cz mu_s;
    set_eye(Z, n);
    double part = 0.0;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) part += nrm2(at(A, e / n, e % n));
    const double anorm = fmax(sqrt(block_sum(part, red)), 1e-300);
    // Householder reduction to upper-Hessenberg form
    for (int k = 0; k + 2 < n; ++k) {
        part = 0.0;
        for (int i = k + 1 + (int)threadIdx.x; i < n; i += blockDim.x) part += nrm2(at(A, i, k));
        const double xn = sqrt(block_sum(part, red));
        if (xn < 1e-300) continue;
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = i > k ? at(A, i, k) : zz();
        __syncthreads();
        if (threadIdx.x == 0) {
            const cz x0 = at(A, k + 1, k);
            const double ax = mag(x0);
            const cz ph = ax > 0 ? cscale(x0, 1.0 / ax) : one();
            v[k + 1] = cadd(v[k + 1], cscale(ph, xn));
        }
        __syncthreads();
        part = 0.0;
        for (int i = k + 1 + (int)threadIdx.x; i < n; i += blockDim.x) part += nrm2(v[i]);
        const double vn = block_sum(part, red);
        if (vn < 1e-300) continue;
        // A <- (I - 2 v v^H / |v|^2) A : thread per column
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            cz d = zz();
            for (int i = k + 1; i < n; ++i) cfma(d, cconj(v[i]), at(A, i, j));
            d = cscale(d, 2.0 / vn);
            for (int i = k + 1; i < n; ++i) at(A, i, j) = csub(at(A, i, j), cmul(d, v[i]));
        }
        __syncthreads();
        // A <- A (I - ...), Z <- Z (I - ...) : thread per row
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            cz d = zz(), dz = zz();
            for (int j = k + 1; j < n; ++j) {
                cfma(d, at(A, i, j), v[j]);
                cfma(dz, at(Z, i, j), v[j]);
            }
            d = cscale(d, 2.0 / vn);
            dz = cscale(dz, 2.0 / vn);
            for (int j = k + 1; j < n; ++j) {
                at(A, i, j) = csub(at(A, i, j), cmul(d, cconj(v[j])));
                at(Z, i, j) = csub(at(Z, i, j), cmul(dz, cconj(v[j])));
            }
        }
        __syncthreads();
        for (int i = k + 2 + (int)threadIdx.x; i < n; i += blockDim.x) at(A, i, k) = zz();
        __syncthreads();
    }
    // single-shift QR iteration with Wilkinson shifts and deflation
    int iter = 0, total = 0;
    if (threadIdx.x == 0) {
        ctl[1] = n - 1;
        ctl[2] = 0;
    }
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) {
            int hi = ctl[1];
            if (hi <= 0 || total >= 10000) {
                ctl[2] = 2;
            } else {
                int l = hi;
                while (l > 0) {
                    const double s = mag(at(A, l - 1, l - 1)) + mag(at(A, l, l));
                    if (mag(at(A, l, l - 1)) <= 1e-16 * (s > 0 ? s : anorm)) {
                        at(A, l, l - 1) = zz();
                        break;
                    }
                    --l;
                }
                if (l == hi) {
                    ctl[1] = hi - 1;
                    ctl[2] = 0;
                    iter = 0;
                } else {
                    ++iter;
                    ++total;
                    const cz a = at(A, hi - 1, hi - 1), b = at(A, hi - 1, hi), c = at(A, hi, hi - 1), d = at(A, hi, hi);
                    cz mu;
                    if (iter % 11 == 10) {
                        mu = cadd(d, cz{mag(at(A, hi, hi - 1)), 0.0});  // exceptional shift
                    } else {
                        const cz tr = cadd(a, d), det = csub(cmul(a, d), cmul(b, c));
                        const cz disc = csqrt_(csub(cscale(cmul(tr, tr), 0.25), det));
                        const cz e1 = cadd(cscale(tr, 0.5), disc), e2 = csub(cscale(tr, 0.5), disc);
                        mu = mag(csub(e1, d)) < mag(csub(e2, d)) ? e1 : e2;
                    }
                    mu_s = mu;
                    ctl[0] = l;
                    ctl[2] = 1;
                }
            }
        }
        __syncthreads();
        const int action = ctl[2], l = ctl[0], hi = ctl[1];
        __syncthreads();
        if (action == 2) break;
        if (action == 0) continue;
        const cz mu = mu_s;
        for (int i = l + (int)threadIdx.x; i <= hi; i += blockDim.x) at(A, i, i) = csub(at(A, i, i), mu);
        __syncthreads();
        for (int k = l; k < hi; ++k) {
            if (threadIdx.x == 0) {
                const cz x = at(A, k, k), y = at(A, k + 1, k);
                const double den = sqrt(nrm2(x) + nrm2(y));
                cs[k - l] = den > 0 ? cscale(x, 1.0 / den) : one();
                sn[k - l] = den > 0 ? cscale(y, 1.0 / den) : zz();
            }
            __syncthreads();
            const cz cc = cs[k - l], ss = sn[k - l];
            for (int j = k + (int)threadIdx.x; j < n; j += blockDim.x) {
                const cz p = at(A, k, j), q = at(A, k + 1, j);
                at(A, k, j) = cadd(cmulc(cc, p), cmulc(ss, q));
                at(A, k + 1, j) = csub(cmul(cc, q), cmul(ss, p));
            }
            __syncthreads();
        }
        for (int k = l; k < hi; ++k) {
            const cz cc = cs[k - l], ss = sn[k - l];
            const int top = min(hi, k + 2);
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                if (i <= top) {
                    const cz p = at(A, i, k), q = at(A, i, k + 1);
                    at(A, i, k) = cadd(cmul(p, cc), cmul(q, ss));
                    at(A, i, k + 1) = cadd(cmul(cscale(p, -1.0), cconj(ss)), cmul(q, cconj(cc)));
                }
                const cz pz = at(Z, i, k), qz = at(Z, i, k + 1);
                at(Z, i, k) = cadd(cmul(pz, cc), cmul(qz, ss));
                at(Z, i, k + 1) = cadd(cmul(cscale(pz, -1.0), cconj(ss)), cmul(qz, cconj(cc)));
            }
            __syncthreads();
        }
        for (int i = l + (int)threadIdx.x; i <= hi; i += blockDim.x) at(A, i, i) = cadd(at(A, i, i), mu);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) lam[i] = at(A, i, i);
    __syncthreads();
    // eigenvectors of the triangular T: thread i builds column i of Y by back substitution
    const double small = 1e-15 * anorm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        for (int j = 0; j < n; ++j) at(Y, j, i) = zz();
        at(Y, i, i) = one();
        for (int j = i - 1; j >= 0; --j) {
            cz acc = zz();
            for (int k = j + 1; k <= i; ++k) cfma(acc, at(A, j, k), at(Y, k, i));
            cz den = csub(at(A, j, j), lam[i]);
            if (mag(den) < small) den = cz{small, 0.0};
            at(Y, j, i) = cdiv(cz{-acc.x, -acc.y}, den);
        }
    }
    __syncthreads();
    matmul(X, Z, Y, n, n, n, false);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double nn = 0.0;
        for (int i = 0; i < n; ++i) nn += nrm2(at(X, i, j));
        nn = sqrt(nn);
        for (int i = 0; i < n; ++i) at(X, i, j) = cscale(at(X, i, j), 1.0 / nn);
    }
    __syncthreads();
}

// P (n x k): the k eigenvectors of A p = lam B p whose eigenvalues are smallest (or largest) in modulus
// (B = nullptr: B = I; hermitian: A Hermitian).  A and B are kept; S0..S4 are scratch matrices.
// Returns false (in every thread) when B is not positive definite or a vector is not finite.
__device__ inline bool select_eigvecs(const cz* A, const cz* B, bool hermitian, int n, int k, bool smallest, cz* P,
                                      cz* S0, cz* S1, cz* S2, cz* S3, cz* S4, double* red) {
    __shared__ cz lam[kLD];
    __shared__ double w[kLD];
    __shared__ int order[kLD];
    __shared__ int badv;
    cz* C = S0;
    cz* L = S1;
    if (B) {
        if (!cholesky(B, L, n, red)) return false;
        adjoint(S2, A, n, n);             // A^H
        lower_solve(L, S2, S2, n, n);     // L^-1 A^H
        adjoint(C, S2, n, n);             // A L^-H
        lower_solve(L, C, C, n, n);       // L^-1 A L^-H
    } else {
        copy(C, A, n, n);
    }
    cz* Y = S2;
    if (hermitian) {
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int i = e / n, j = e % n;
            if (i <= j) {
                const cz a = at(C, i, j), b = cconj(at(C, j, i));
                const cz h = cscale(cadd(a, b), 0.5);
                at(S3, i, j) = h;
                at(S3, j, i) = cconj(h);
            }
        }
        __syncthreads();
        herm_eig(S3, Y, w, S4, n, red);
        for (int i = threadIdx.x; i < n; i += blockDim.x) lam[i] = cz{w[i], 0.0};
        __syncthreads();
    } else {
        gen_eig(C, lam, Y, S3, S4, n, red);
    }
    // stable order by |lam|
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double mi = mag(lam[i]);
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double mj = mag(lam[j]);
            const bool before = smallest ? mj < mi : mj > mi;
            rank += (before || (mj == mi && j < i)) ? 1 : 0;
        }
        order[rank] = i;
    }
    if (threadIdx.x == 0) badv = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < n * k; e += blockDim.x) {
        const int i = e / k, j = e % k;
        at(S3, i, j) = at(Y, i, order[j]);
    }
    __syncthreads();
    if (B) lower_adj_solve(L, S3, P, n, k);
    else copy(P, S3, n, k);
    for (int e = threadIdx.x; e < n * k; e += blockDim.x) {
        const cz z = at(P, e / k, e % k);
        if (!isfinite(z.x) || !isfinite(z.y)) badv = 1;
    }
    __syncthreads();
    return badv == 0;
}

}  // namespace sld
}  // namespace clgd
