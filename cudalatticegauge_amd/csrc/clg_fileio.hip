// Configuration files: Bridge++ text/binary gauge configurations and the CLG binary format
// (double, float) for gauge, Wilson and staggered fields, with the MD5 string the reference
// returns from SaveToFile.  Host code only: files are parsed into (or written from) the host
// AoS order that clg_field_upload/download exchange, which for the CLG binary format is the
// file's own order (CField.cpp:32-73, CFieldGaugeSU3.cu:1080-1185, 1692-1760,
// CFieldFermionWilsonSquareSU3.cu:602-633, 1144-1220).
#include "clg_internal.h"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

namespace clg {
namespace {

// ---- MD5 (RFC 1321) -------------------------------------------------------------------------
// The length word follows CLGMD5.h:351-382: the message length in bits modulo 2^32 in the low
// word, zero in the high word.  Below 512 MiB this is standard MD5.
struct Md5 {
    uint32_t h[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};

    static uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }

    void block(const unsigned char* p) {
        // per-round shift amounts and the sine table K[i] = floor(2^32 |sin(i + 1)|) (RFC 1321 3.4)
        static const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
        static const std::array<uint32_t, 64> K = [] {
            std::array<uint32_t, 64> k{};
            for (int i = 0; i < 64; ++i) k[i] = (uint32_t)(std::fabs(std::sin((double)(i + 1))) * 4294967296.0);
            return k;
        }();
        uint32_t M[16];
        for (int i = 0; i < 16; ++i)
            M[i] = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
                   ((uint32_t)p[4 * i + 3] << 24);
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
        for (int i = 0; i < 64; ++i) {
            uint32_t f;
            int g;
            const int r = i >> 4;
            if (r == 0) { f = (b & c) | (~b & d); g = i; }
            else if (r == 1) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
            else if (r == 2) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
            else { f = c ^ (b | ~d); g = (7 * i) & 15; }
            const uint32_t t = d;
            d = c;
            c = b;
            b = b + rotl(a + f + K[i] + M[g], S[r][i & 3]);
            a = t;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d;
    }

    void hex(const unsigned char* data, size_t n, char out[33]) {
        size_t i = 0;
        for (; i + 64 <= n; i += 64) block(data + i);
        unsigned char tail[128] = {0};
        const size_t rem = n - i;
        std::memcpy(tail, data + i, rem);
        tail[rem] = 0x80;
        const size_t tl = rem + 1 + 8 <= 64 ? 64 : 128;
        const uint32_t bits = (uint32_t)((uint64_t)n * 8u);
        for (int k = 0; k < 4; ++k) tail[tl - 8 + k] = (unsigned char)(bits >> (8 * k));
        block(tail);
        if (tl == 128) block(tail + 64);
        static const char* dig = "0123456789ABCDEF";
        for (int w = 0; w < 4; ++w)
            for (int k = 0; k < 4; ++k) {
                const unsigned byte = (h[w] >> (8 * k)) & 0xffu;
                out[8 * w + 2 * k] = dig[byte >> 4];
                out[8 * w + 2 * k + 1] = dig[byte & 15];
            }
        out[32] = 0;
    }
};

long sites_of(const int dims[4]) { return (long)dims[0] * dims[1] * dims[2] * dims[3]; }

int reals_per_site(int kind) {
    if (kind == CLG_WILSON) return 24;
    if (kind == CLG_KS) return 6;
    if (kind == CLG_GAUGE || kind == CLG_GAUGE12) return 72;
    return 0;
}

bool read_all(const char* path, std::vector<unsigned char>& out) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return false;
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    const size_t got = n > 0 ? std::fread(out.data(), 1, (size_t)n, fp) : 0;
    std::fclose(fp);
    return got == out.size();
}

// Bridge++ orders links as ((((t*Lz + z)*Ly + y)*Lx + x)*4 + mu) (CFieldGaugeSU3.cu:574-578); the
// host order here is ((((x*Ly + y)*Lz + z)*Lt + t)*4 + mu).  18 reals per link, row-major 3x3.
void bridge_to_host(const int dims[4], const double* b, double* host) {
    const int Lx = dims[0], Ly = dims[1], Lz = dims[2], Lt = dims[3];
    for (int x = 0; x < Lx; ++x)
        for (int y = 0; y < Ly; ++y)
            for (int z = 0; z < Lz; ++z)
                for (int t = 0; t < Lt; ++t) {
                    const long hs = (((long)x * Ly + y) * Lz + z) * Lt + t;
                    const long bs = (((long)t * Lz + z) * Ly + y) * Lx + x;
                    std::memcpy(host + hs * 72, b + bs * 72, 72 * sizeof(double));
                }
}

// ---- the compressed format: one SU(3) logarithm per link ---------------------------------------
// CFieldGaugeSU3::SaveToCompressedFile / InitialWithByteCompressed (CFieldGaugeSU3.cu:1225-1265,
// 1651-1690) store, per link in the CLG binary link order, 9 reals of A = log U (principal logarithm
// of each eigenvalue, deviceSU3::Log, SU3.h:1948-1966): Re/Im A01, Re/Im A02, Re/Im A12, Im A00,
// Im A11, Im A22.  A is anti-Hermitian (A10 = -conj A01, ...; the real parts of the diagonal are 0)
// and its trace may be +-2 pi i.  Loading rebuilds U = exp(A) exactly (deviceSU3::StrictExp).  Both
// directions here diagonalise a Hermitian matrix with cyclic complex Jacobi rotations.
typedef std::complex<double> cplx;

// H (Hermitian, row-major 3x3, overwritten) = V diag(w) V^H; the columns of V are the eigenvectors
void herm3_eig(cplx H[9], double w[3], cplx V[9]) {
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    double scale = 0.0;
    for (int i = 0; i < 9; ++i) scale += std::norm(H[i]);
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = std::norm(H[1]) + std::norm(H[2]) + std::norm(H[5]);
        if (off <= 1e-30 * scale) break;  // off-diagonal below 1e-15 of the matrix: rounding level
        static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (const auto& pq : PQ) {
            const int p = pq[0], q = pq[1];
            const cplx c = H[3 * p + q];
            const double ac = std::abs(c);
            if (ac == 0.0) continue;
            // J = diag(1, e^{-i phi}) on (p, q), then the real rotation that zeroes |c| (Numerical
            // Recipes' Jacobi step): J_pp = co, J_pq = si, J_qp = -si e^{-i phi}, J_qq = co e^{-i phi}
            const cplx ph = std::conj(c) / ac;
            const double a = H[3 * p + p].real(), b = H[3 * q + q].real();
            const double th = (b - a) / (2.0 * ac);
            const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
            const double co = 1.0 / std::sqrt(t * t + 1.0), si = t * co;
            cplx J[9];
            for (int i = 0; i < 9; ++i) J[i] = (i % 4 == 0) ? 1.0 : 0.0;
            J[3 * p + p] = co;
            J[3 * p + q] = si;
            J[3 * q + p] = -si * ph;
            J[3 * q + q] = co * ph;
            cplx T[9], N[9];
            for (int i = 0; i < 3; ++i)  // T = H J
                for (int j = 0; j < 3; ++j) T[3 * i + j] = H[3 * i] * J[j] + H[3 * i + 1] * J[3 + j] + H[3 * i + 2] * J[6 + j];
            for (int i = 0; i < 3; ++i)  // H = J^H T
                for (int j = 0; j < 3; ++j)
                    N[3 * i + j] = std::conj(J[i]) * T[j] + std::conj(J[3 + i]) * T[3 + j] + std::conj(J[6 + i]) * T[6 + j];
            for (int i = 0; i < 9; ++i) H[i] = N[i];
            for (int i = 0; i < 3; ++i)  // V = V J
                for (int j = 0; j < 3; ++j) T[3 * i + j] = V[3 * i] * J[j] + V[3 * i + 1] * J[3 + j] + V[3 * i + 2] * J[6 + j];
            for (int i = 0; i < 9; ++i) V[i] = T[i];
        }
    }
    for (int k = 0; k < 3; ++k) w[k] = H[4 * k].real();
}

// M = V diag(d) V^H
void from_eig(const cplx V[9], const cplx d[3], cplx M[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cplx s = 0.0;
            for (int k = 0; k < 3; ++k) s += V[3 * i + k] * d[k] * std::conj(V[3 * j + k]);
            M[3 * i + j] = s;
        }
}

// A = log U for unitary U: the eigenvectors of U (a normal matrix) are those of the Hermitian
// (e^{-i f} U + e^{i f} U^H) / 2, whose eigenvalues cos(theta_k - f) coincide when an eigenvalue of U
// sits at e^{-2 i f} (theta_j + theta_k = 2 f while the three sum to 0 mod 2 pi).  Near that point
// the eigenvectors of the pair mix and V^H U V is not diagonal, so up to three angles f are tried and
// the most diagonal V^H U V is kept; no unitary matrix is near the bad point of all three.
void su3_log(const cplx U[9], cplx A[9]) {
    cplx bestV[9], bestB[9];
    double best = -1.0;
    for (double f : {0.7236, 2.1817, -1.3109}) {
        const cplx e = std::polar(1.0, -f);
        cplx H[9], V[9];
        double w[3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[3 * i + j] = 0.5 * (e * U[3 * i + j] + std::conj(e * U[3 * j + i]));
        herm3_eig(H, w, V);
        cplx B[9];  // V^H U V
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                cplx s = 0.0;
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) s += std::conj(V[3 * k + i]) * U[3 * k + l] * V[3 * l + j];
                B[3 * i + j] = s;
            }
        const double offd = std::abs(B[1]) + std::abs(B[2]) + std::abs(B[3]) + std::abs(B[5]) + std::abs(B[6]) +
                            std::abs(B[7]);
        if (best < 0.0 || offd < best) {
            best = offd;
            for (int i = 0; i < 9; ++i) {
                bestV[i] = V[i];
                bestB[i] = B[i];
            }
        }
        if (best < 1e-14) break;
    }
    cplx d[3];
    for (int k = 0; k < 3; ++k) d[k] = cplx(0.0, std::arg(bestB[4 * k]));
    from_eig(bestV, d, A);
}

// U = exp(A) for anti-Hermitian A: -i A is Hermitian with eigenvalues mu_k, U = V diag(e^{i mu}) V^H
void ah3_exp(const cplx A[9], cplx U[9]) {
    cplx H[9], V[9];
    double w[3];
    for (int i = 0; i < 9; ++i) H[i] = cplx(A[i].imag(), -A[i].real());
    herm3_eig(H, w, V);
    cplx d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::polar(1.0, w[k]);
    from_eig(V, d, U);
}

// one link: 18 host reals (row-major complex 3x3) <-> 9 compressed reals
void link_compress(const double* u, double* c) {
    cplx U[9], A[9];
    for (int i = 0; i < 9; ++i) U[i] = cplx(u[2 * i], u[2 * i + 1]);
    su3_log(U, A);
    c[0] = A[1].real(); c[1] = A[1].imag();
    c[2] = A[2].real(); c[3] = A[2].imag();
    c[4] = A[5].real(); c[5] = A[5].imag();
    c[6] = A[0].imag(); c[7] = A[4].imag(); c[8] = A[8].imag();
}
void link_expand(const double* c, double* u) {
    cplx A[9], U[9];
    A[1] = cplx(c[0], c[1]);
    A[3] = cplx(-c[0], c[1]);
    A[2] = cplx(c[2], c[3]);
    A[6] = cplx(-c[2], c[3]);
    A[5] = cplx(c[4], c[5]);
    A[7] = cplx(-c[4], c[5]);
    A[0] = cplx(0.0, c[6]);
    A[4] = cplx(0.0, c[7]);
    A[8] = cplx(0.0, c[8]);
    ah3_exp(A, U);
    for (int i = 0; i < 9; ++i) {
        u[2 * i] = U[i].real();
        u[2 * i + 1] = U[i].imag();
    }
}

// f(l) for l in [0, n), split over up to 16 host threads (the compressed format's per-link
// logarithms and exponentials are independent)
template <class F>
void parallel_links(long n, F f) {
    const long nt = std::max(1L, std::min<long>({16L, (long)std::thread::hardware_concurrency(), n / 4096}));
    if (nt <= 1) {
        for (long l = 0; l < n; ++l) f(l);
        return;
    }
    std::vector<std::thread> th;
    long started = 0;
    try {
        for (long t = 0; t < nt; ++t, ++started)
            th.emplace_back([=]() {
                for (long l = n * t / nt; l < n * (t + 1) / nt; ++l) f(l);
            });
    } catch (const std::system_error&) {
        // no more threads: this thread takes the chunks that were not started
        for (long l = n * started / nt; l < n; ++l) f(l);
    }
    for (auto& x : th) x.join();
}

}  // namespace
}  // namespace clg

using namespace clg;

extern "C" {

int clg_md5(const void* data, unsigned long n, char md5_out[33]) {
    CLG_CHECK_ARG(md5_out && (data || n == 0), "null argument");
    Md5 m;
    m.hex((const unsigned char*)data, (size_t)n, md5_out);
    return CLG_OK;
}

static int clg_file_read_impl(const char* path, int file_type, int kind, const int dims[4], double* host) {
    CLG_CHECK_ARG(path && dims && host, "null argument");
    const int rps = reals_per_site(kind);
    CLG_CHECK_ARG(rps > 0, "unknown field kind %d", kind);
    CLG_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0 && dims[3] > 0, "bad lattice extents");
    const bool gauge = rps == 72;
    const long nreal = sites_of(dims) * rps;
    CLG_CHECK_ARG(file_type >= CLG_FILE_BRIDGEPP_TXT && file_type <= CLG_FILE_CLG_BIN_DOUBLE, "unknown file type %d",
                  file_type);
    // fermion files are the CLG binary format in the build's Real (CFieldFermionWilsonSquareSU3.cu:
    // 604-613, CFieldFermionKSSU3.cu:1056-1078): double, or float for a single-precision build's files
    CLG_CHECK_ARG(gauge || file_type == CLG_FILE_CLG_BIN || file_type == CLG_FILE_CLG_BIN_FLOAT ||
                      file_type == CLG_FILE_CLG_BIN_DOUBLE,
                  "fermion fields are read from CLG binary files only (CFieldFermionWilsonSquareSU3.cu:604-608)");
    if (file_type == CLG_FILE_BRIDGEPP_TXT) {
        FILE* fp = std::fopen(path, "r");
        CLG_CHECK_ARG(fp, "cannot open %s: %s", path, std::strerror(errno));
        std::vector<double> b((size_t)nreal);
        long got = 0;
        char line[256];
        bool extra = false;
        while (std::fgets(line, sizeof(line), fp)) {
            char* s = line;
            while (*s == ' ' || *s == '\t') ++s;
            if (*s == '\n' || *s == '\r' || *s == 0) continue;
            if (got == nreal) { extra = true; break; }
            char* end = nullptr;
            b[(size_t)got] = std::strtod(s, &end);
            if (end == s) break;
            ++got;
        }
        std::fclose(fp);
        CLG_CHECK_ARG(got == nreal && !extra, "%s: expected %ld values (18 per link), read %ld%s", path, nreal, got,
                      extra ? " and more" : "");
        bridge_to_host(dims, b.data(), host);
        return CLG_OK;
    }
    std::vector<unsigned char> bytes;
    CLG_CHECK_ARG(read_all(path, bytes), "cannot read %s", path);
    if (file_type == CLG_FILE_CLG_BIN_COMPRESSED) {
        // 9 doubles per link, rebuilt as U = exp(A) (CFieldGaugeSU3.cu:1225-1265)
        const long links = sites_of(dims) * 4;
        CLG_CHECK_ARG(bytes.size() == (size_t)links * 72, "%s holds %zu bytes, the lattice needs %zu (9 doubles a link)",
                      path, bytes.size(), (size_t)links * 72);
        const double* c = (const double*)bytes.data();
        parallel_links(links, [&](long l) { link_expand(c + 9 * l, host + 18 * l); });
        return CLG_OK;
    }
    const size_t want = (size_t)nreal * (file_type == CLG_FILE_CLG_BIN_FLOAT ? 4 : 8);
    CLG_CHECK_ARG(bytes.size() == want, "%s holds %zu bytes, the lattice needs %zu", path, bytes.size(), want);
    if (file_type == CLG_FILE_BRIDGEPP_BIN) {
        // big-endian doubles (CFieldGaugeSU3.cu:1108-1126)
        std::vector<double> b((size_t)nreal);
        for (long i = 0; i < nreal; ++i) {
            unsigned char le[8];
            for (int j = 0; j < 8; ++j) le[j] = bytes[(size_t)i * 8 + 7 - j];
            std::memcpy(&b[(size_t)i], le, 8);
        }
        bridge_to_host(dims, b.data(), host);
    } else if (file_type == CLG_FILE_CLG_BIN_FLOAT) {
        const float* f = (const float*)bytes.data();
        for (long i = 0; i < nreal; ++i) host[i] = (double)f[i];
    } else {
        std::memcpy(host, bytes.data(), want);
    }
    return CLG_OK;
}

static int clg_file_write_impl(const char* path, int file_type, int kind, const int dims[4], const double* host,
                   char md5_out[33]) {
    CLG_CHECK_ARG(path && dims && host, "null argument");
    const int rps = reals_per_site(kind);
    CLG_CHECK_ARG(rps > 0, "unknown field kind %d", kind);
    CLG_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0 && dims[3] > 0, "bad lattice extents");
    if (file_type != CLG_FILE_CLG_BIN && file_type != CLG_FILE_CLG_BIN_FLOAT && file_type != CLG_FILE_CLG_BIN_DOUBLE &&
        file_type != CLG_FILE_CLG_BIN_COMPRESSED) {
        set_error("saving supports the CLG binary formats only (CField.cpp:32-72)");
        return CLG_UNSUPPORTED;
    }
    CLG_CHECK_ARG(file_type != CLG_FILE_CLG_BIN_COMPRESSED || rps == 72,
                  "the compressed format holds gauge fields only (CFieldGaugeSU3.cu:1651-1690)");
    const long nreal = sites_of(dims) * rps;
    std::vector<unsigned char> bytes;
    if (file_type == CLG_FILE_CLG_BIN_COMPRESSED) {
        // 9 doubles of log U per link (CFieldGaugeSU3.cu:1651-1690)
        const long links = sites_of(dims) * 4;
        bytes.resize((size_t)links * 72);
        double* c = (double*)bytes.data();
        parallel_links(links, [&](long l) { link_compress(host + 18 * l, c + 9 * l); });
    } else if (file_type == CLG_FILE_CLG_BIN_FLOAT) {
        bytes.resize((size_t)nreal * 4);
        float* f = (float*)bytes.data();
        for (long i = 0; i < nreal; ++i) f[i] = (float)host[i];
    } else {
        bytes.resize((size_t)nreal * 8);
        std::memcpy(bytes.data(), host, bytes.size());
    }
    FILE* fp = std::fopen(path, "wb");
    CLG_CHECK_ARG(fp, "cannot open %s: %s", path, std::strerror(errno));
    const size_t put = std::fwrite(bytes.data(), 1, bytes.size(), fp);
    const int cl = std::fclose(fp);
    CLG_CHECK_ARG(put == bytes.size() && cl == 0, "short write to %s", path);
    if (md5_out) {
        Md5 m;
        m.hex(bytes.data(), bytes.size(), md5_out);
    }
    return CLG_OK;
}

static int clg_field_load_file_impl(clg_field f, const char* path, int file_type) {
    CLG_CHECK_ARG(f && path, "null argument");
    CLG_CHECK_ARG(f->subset == CLG_FULL, "files hold full-lattice fields");
    std::vector<double> host((size_t)clg_field_host_doubles(f));
    CLG_TRY(clg_file_read(path, file_type, f->kind, f->dims, host.data()));
    // Bridge++ configurations are re-unitarised after loading (SetByArray -> ElementNormalize,
    // CFieldGaugeSU3.cu:1262-1285); the normalisation runs on an FP64 18-real field, so other
    // gauge storage goes through an FP64 staging field and a conversion.
    const bool bridge = file_type == CLG_FILE_BRIDGEPP_TXT || file_type == CLG_FILE_BRIDGEPP_BIN;
    if (!bridge) return clg_field_upload(f, host.data());
    if (f->kind == CLG_GAUGE && f->prec == CLG_F64) {
        CLG_TRY(clg_field_upload(f, host.data()));
        CLG_TRY(clg_gauge_normalize(f->ctx, f));
        CLG_HIP(hipStreamSynchronize(f->ctx->stream));
        return CLG_OK;
    }
    clg_field tmp = nullptr;
    CLG_TRY(clg_field_create(f->ctx, CLG_GAUGE, CLG_F64, CLG_FULL, f->dims, &tmp));
    int st = clg_field_upload(tmp, host.data());
    if (st == CLG_OK) st = clg_gauge_normalize(f->ctx, tmp);
    if (st == CLG_OK) st = clg_field_copy(f, tmp);
    if (st == CLG_OK && hipStreamSynchronize(f->ctx->stream) != hipSuccess) {
        set_error("hipStreamSynchronize failed after loading %s", path);
        st = CLG_HIP_ERROR;
    }
    clg_field_destroy(tmp);
    return st;
}

static int clg_field_save_file_impl(clg_field f, const char* path, int file_type, char md5_out[33]) {
    CLG_CHECK_ARG(f && path, "null argument");
    CLG_CHECK_ARG(f->subset == CLG_FULL, "files hold full-lattice fields");
    std::vector<double> host((size_t)clg_field_host_doubles(f));
    CLG_TRY(clg_field_download(f, host.data()));
    return clg_file_write(path, file_type, f->kind, f->dims, host.data(), md5_out);
}

// host allocations of a whole field (or file) can fail for very large lattices: reported as
// CLG_OUT_OF_MEMORY instead of an exception crossing the C interface
#define CLG_HOST_GUARD(call)                                                  \
    try {                                                                     \
        return call;                                                          \
    } catch (const std::bad_alloc&) {                                         \
        set_error("host memory exhausted");                                   \
        return CLG_OUT_OF_MEMORY;                                             \
    }

int clg_file_read(const char* path, int file_type, int kind, const int dims[4], double* host) {
    CLG_HOST_GUARD(clg_file_read_impl(path, file_type, kind, dims, host))
}
int clg_file_write(const char* path, int file_type, int kind, const int dims[4], const double* host,
                   char md5_out[33]) {
    CLG_HOST_GUARD(clg_file_write_impl(path, file_type, kind, dims, host, md5_out))
}
int clg_field_load_file(clg_field f, const char* path, int file_type) {
    CLG_HOST_GUARD(clg_field_load_file_impl(f, path, file_type))
}
int clg_field_save_file(clg_field f, const char* path, int file_type, char md5_out[33]) {
    CLG_HOST_GUARD(clg_field_save_file_impl(f, path, file_type, md5_out))
}

}  // extern "C"
