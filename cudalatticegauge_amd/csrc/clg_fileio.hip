// Configuration files: Bridge++ text/binary gauge configurations and the CLG binary format
// (double, float) for gauge, Wilson and staggered fields, with the MD5 string the reference
// returns from SaveToFile.  Host code only: files are parsed into (or written from) the host
// AoS order that clg_field_upload/download exchange, which for the CLG binary format is the
// file's own order (CField.cpp:32-73, CFieldGaugeSU3.cu:1080-1185, 1692-1760,
// CFieldFermionWilsonSquareSU3.cu:602-633, 1144-1220).
#include "clg_internal.h"

#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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

int clg_file_read(const char* path, int file_type, int kind, const int dims[4], double* host) {
    CLG_CHECK_ARG(path && dims && host, "null argument");
    const int rps = reals_per_site(kind);
    CLG_CHECK_ARG(rps > 0, "unknown field kind %d", kind);
    CLG_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0 && dims[3] > 0, "bad lattice extents");
    const bool gauge = rps == 72;
    const long nreal = sites_of(dims) * rps;
    if (file_type == CLG_FILE_CLG_BIN_COMPRESSED) {
        set_error("CLG compressed files (SU(3) logarithms) are not supported");
        return CLG_UNSUPPORTED;
    }
    CLG_CHECK_ARG(file_type >= CLG_FILE_BRIDGEPP_TXT && file_type <= CLG_FILE_CLG_BIN_DOUBLE, "unknown file type %d",
                  file_type);
    CLG_CHECK_ARG(gauge || file_type == CLG_FILE_CLG_BIN,
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

int clg_file_write(const char* path, int file_type, int kind, const int dims[4], const double* host,
                   char md5_out[33]) {
    CLG_CHECK_ARG(path && dims && host, "null argument");
    const int rps = reals_per_site(kind);
    CLG_CHECK_ARG(rps > 0, "unknown field kind %d", kind);
    CLG_CHECK_ARG(dims[0] > 0 && dims[1] > 0 && dims[2] > 0 && dims[3] > 0, "bad lattice extents");
    if (file_type != CLG_FILE_CLG_BIN && file_type != CLG_FILE_CLG_BIN_FLOAT && file_type != CLG_FILE_CLG_BIN_DOUBLE) {
        set_error("saving supports the CLG binary formats only (CField.cpp:32-72)");
        return CLG_UNSUPPORTED;
    }
    const long nreal = sites_of(dims) * rps;
    std::vector<unsigned char> bytes;
    if (file_type == CLG_FILE_CLG_BIN_FLOAT) {
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

int clg_field_load_file(clg_field f, const char* path, int file_type) {
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

int clg_field_save_file(clg_field f, const char* path, int file_type, char md5_out[33]) {
    CLG_CHECK_ARG(f && path, "null argument");
    CLG_CHECK_ARG(f->subset == CLG_FULL, "files hold full-lattice fields");
    std::vector<double> host((size_t)clg_field_host_doubles(f));
    CLG_TRY(clg_field_download(f, host.data()));
    return clg_file_write(path, file_type, f->kind, f->dims, host.data(), md5_out);
}

}  // extern "C"
