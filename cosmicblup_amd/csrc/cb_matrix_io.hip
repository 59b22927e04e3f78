// Relationship-matrix files on the output side of the GRM path:
//   - COSMIC_UPPER_PACKED (src/core/upper_packed_binary.{h,cpp}): a 32-byte
//     magic, version 1, n (i64), index base 1, storage 1 (upper packed),
//     dtype 1 (float64), endian 1 (little), then A(0,0), A(0,1), A(1,1),
//     A(0,2), ... (upper triangle, column-major) as doubles. For a symmetric
//     matrix stored row-major that is the lower triangle row by row: row j,
//     columns 0..j.
//   - text triplets "i<TAB>j<TAB>value" for j <= i, 1-based, the value in
//     the default 6-significant-digit stream format (write_matrix_txt,
//     src/core/matrix_io.cpp:464-472).
// A matrix in device memory is streamed through two pinned panels: the copy
// of panel k+1 runs while panel k is formatted and written; text panels are
// formatted by several host threads.
#include <algorithm>
#include <array>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int32_t kVersion = 1, kIndexBase = 1, kStorageUpperPacked = 1, kDtypeFloat64 = 1, kEndianLittle = 1;

std::array<char, 32> packed_magic() {
    std::array<char, 32> m{};
    const char* t = "COSMIC_UPPER_PACKED";
    std::memcpy(m.data(), t, std::strlen(t));
    return m;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

void put(FILE* f, const void* p, size_t bytes, const std::string& what) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw Error(CB_ERR_IO, what);
}

void write_packed_header(FILE* f, int64_t n, const std::string& what) {
    const auto m = packed_magic();
    put(f, m.data(), m.size(), what);
    put(f, &kVersion, 4, what);
    put(f, &n, 8, what);
    put(f, &kIndexBase, 4, what);
    put(f, &kStorageUpperPacked, 4, what);
    put(f, &kDtypeFloat64, 4, what);
    put(f, &kEndianLittle, 4, what);
}

// rows [r0, r1) of a row-major symmetric matrix (row r at rows + (r - r0) * ld)
void write_rows(FILE* f, int format, const double* rows, int64_t ld, int64_t r0, int64_t r1,
                const std::string& what) {
    if (format == 1) {
        for (int64_t r = r0; r < r1; ++r) put(f, rows + (r - r0) * ld, (size_t)(r + 1) * 8, what);
        return;
    }
    // text: split the panel's lines over threads by equal triangle area
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(16, (r1 - r0) * (r1 + r0) / 2 / 200000));
    std::vector<int64_t> cut(nt + 1, r0);
    cut[nt] = r1;
    {
        const double total = 0.5 * ((double)r1 * r1 - (double)r0 * r0);
        int64_t r = r0;
        double acc = 0;
        for (int t = 1; t < nt; ++t) {
            while (r < r1 && acc < total * t / nt) acc += (double)(++r);
            cut[t] = std::min(r, r1);
        }
    }
    std::vector<std::string> out(nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            std::string& s = out[t];
            char buf[64];
            for (int64_t r = cut[t]; r < cut[t + 1]; ++r) {
                const double* row = rows + (r - r0) * ld;
                for (int64_t c = 0; c <= r; ++c) {
                    const int len = std::snprintf(buf, sizeof buf, "%" PRId64 "\t%" PRId64 "\t%g\n", r + 1, c + 1,
                                                  row[c]);
                    s.append(buf, (size_t)len);
                }
            }
        });
    for (auto& x : th) x.join();
    for (const auto& s : out) put(f, s.data(), s.size(), what);
}

std::string write_error(int format, const char* path) {
    return format == 1 ? std::string("cannot write upper-packed binary matrix file: ") + path
                       : std::string("Cannot write matrix file: ") + path;
}

}  // namespace

namespace cbamd {

void write_symmetric(cb_context* ctx, const double* A, int64_t n, int64_t ld, bool on_device, const char* path,
                     int format) {
    CB_REQUIRE(path && A && n > 0 && ld >= n, "write_symmetric: bad arguments");
    CB_REQUIRE(format == 0 || format == 1, "write_symmetric: format must be 0 (text) or 1 (upper packed)");
    const std::string what = write_error(format, path);
    File f;
    f.f = std::fopen(path, format == 1 ? "wb" : "w");
    if (!f.f) throw Error(CB_ERR_IO, what);
    if (format == 1) write_packed_header(f.f, n, what);
    if (!on_device) {
        write_rows(f.f, format, A, ld, 0, n, what);
    } else {
        // panels of about 32 MB; two pinned buffers, copy of k+1 under the write of k
        hipStream_t st = ctx->stream;
        const int64_t rows = std::max<int64_t>(1, (int64_t)(4u << 20) / n);
        const int64_t npan = ceil_div(n, rows);
        double* pin[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        auto release = [&] {
            for (int b = 0; b < 2; ++b) {
                if (pin[b]) (void)hipHostFree(pin[b]);
                if (ev[b]) (void)hipEventDestroy(ev[b]);
            }
        };
        try {
            for (int b = 0; b < 2; ++b) {
                CB_HIP(hipHostMalloc((void**)&pin[b], (size_t)rows * n * 8, hipHostMallocDefault));
                CB_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
            }
            auto issue = [&](int64_t k) {
                const int64_t r0 = k * rows, r1 = std::min(n, r0 + rows);
                CB_HIP(hipMemcpy2DAsync(pin[k & 1], (size_t)n * 8, A + r0 * ld, (size_t)ld * 8, (size_t)n * 8,
                                        (size_t)(r1 - r0), hipMemcpyDeviceToHost, st));
                CB_HIP(hipEventRecord(ev[k & 1], st));
            };
            issue(0);
            for (int64_t k = 0; k < npan; ++k) {
                if (k + 1 < npan) issue(k + 1);
                CB_HIP(hipEventSynchronize(ev[k & 1]));
                const int64_t r0 = k * rows, r1 = std::min(n, r0 + rows);
                write_rows(f.f, format, pin[k & 1], n, r0, r1, what);
                // the buffer is reused by issue(k + 2) only after this write
            }
        } catch (...) {
            (void)hipStreamSynchronize(ctx->stream);
            release();
            throw;
        }
        release();
    }
    if (std::fflush(f.f) != 0) throw Error(CB_ERR_IO, what);
}

}  // namespace cbamd

extern "C" {

int cb_matrix_write(const char* path, int64_t n, const double* A, int64_t lda, int format) {
    return guarded([&] {
        CB_REQUIRE(path && A && n > 0 && lda >= n, "cb_matrix_write: bad arguments");
        write_symmetric(nullptr, A, n, lda, false, path, format);
    });
}

int cb_upper_packed_read(const char* path, int64_t* n_out, double* out, int64_t ldo) {
    return guarded([&] {
        CB_REQUIRE(path && n_out, "cb_upper_packed_read: NULL argument");
        const std::string p(path);
        File f;
        f.f = std::fopen(path, "rb");
        if (!f.f) throw Error(CB_ERR_IO, "cannot read upper-packed binary matrix file: " + p);
        std::array<char, 32> magic{};
        if (std::fread(magic.data(), 1, 32, f.f) != 32 || magic != packed_magic())
            throw Error(CB_ERR_IO, "unsupported binary matrix header: " + p);
        int32_t version = 0, base = 0, storage = 0, dtype = 0, endian = 0;
        int64_t n = 0;
        auto get = [&](void* dst, size_t bytes) {
            if (std::fread(dst, 1, bytes, f.f) != bytes)
                throw Error(CB_ERR_IO, "truncated upper-packed binary matrix file: " + p);
        };
        get(&version, 4);
        get(&n, 8);
        get(&base, 4);
        get(&storage, 4);
        get(&dtype, 4);
        get(&endian, 4);
        if (version != kVersion || base != kIndexBase || storage != kStorageUpperPacked || dtype != kDtypeFloat64 ||
            endian != kEndianLittle)
            throw Error(CB_ERR_IO, "unsupported upper-packed binary matrix metadata: " + p);
        if (n <= 0 || n > (int64_t)INT32_MAX)
            throw Error(CB_ERR_IO, "invalid upper-packed binary matrix dimension: " + p);
        *n_out = n;
        if (!out) return;
        CB_REQUIRE(ldo >= n, "cb_upper_packed_read: ldo < n");
        std::vector<double> col((size_t)n);
        for (int64_t j = 0; j < n; ++j) {
            get(col.data(), (size_t)(j + 1) * 8);
            for (int64_t i = 0; i <= j; ++i) {
                out[i * ldo + j] = col[i];
                out[j * ldo + i] = col[i];
            }
        }
        char trailing;
        if (std::fread(&trailing, 1, 1, f.f) == 1)
            throw Error(CB_ERR_IO, "upper-packed binary matrix file has trailing bytes: " + p);
    });
}

}  // extern "C"
