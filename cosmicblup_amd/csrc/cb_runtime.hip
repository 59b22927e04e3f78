// Context, error text, per-kernel timing and the device-resident PLINK BED
// (the MI355X counterpart of PlinkReader's mmap, src/core/plink_reader.cpp:66-91).
#include "cb_internal.h"
#include "cb_genotypes.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace cbamd {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// Pinned staging buffer for copies of caller-owned host memory, allocated on
// first use and kept for the life of the process (pinned memory serves every
// device). One copy uses it at a time.
namespace {
constexpr size_t kStageBytes = size_t(64) << 20;
std::mutex g_stage_mu;
char* g_stage = nullptr;

char* staging() {
    if (!g_stage) CB_HIP(hipHostMalloc((void**)&g_stage, kStageBytes, hipHostMallocDefault));
    return g_stage;
}
}  // namespace

void copy_host_2d(hipStream_t st, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                  size_t rows, hipMemcpyKind kind) {
    if (kind == hipMemcpyDeviceToDevice) {
        CB_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, kind, st));
        return;
    }
    CB_REQUIRE(kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToHost, "copy_host: unsupported copy kind");
    if (width == 0 || rows == 0) return;
    const bool up = kind == hipMemcpyHostToDevice;
    std::lock_guard<std::mutex> lock(g_stage_mu);
    char* stage = staging();
    CB_HIP(hipStreamSynchronize(st));  // the work queued before the copy is done
    // (row, byte offset) pieces of at most kStageBytes; whole rows when they fit
    const size_t seg = std::min(width, kStageBytes);
    const size_t per = width <= kStageBytes ? kStageBytes / width : 1;
    for (size_t r0 = 0; r0 < rows; r0 += per) {
        const size_t nr = std::min(per, rows - r0);
        for (size_t off = 0; off < width; off += seg) {
            const size_t len = std::min(seg, width - off);
            char* d = (char*)dst + r0 * dpitch + off;
            const char* s = (const char*)src + r0 * spitch + off;
            if (up) {
                // the previous piece's DMA has finished (synchronised below)
                for (size_t r = 0; r < nr; ++r) std::memcpy(stage + r * len, s + r * spitch, len);
                CB_HIP(hipMemcpy2DAsync(d, dpitch, stage, len, len, nr, kind, st));
                CB_HIP(hipStreamSynchronize(st));
            } else {
                CB_HIP(hipMemcpy2DAsync(stage, len, s, spitch, len, nr, kind, st));
                CB_HIP(hipStreamSynchronize(st));
                for (size_t r = 0; r < nr; ++r) std::memcpy(d + r * dpitch, stage + r * len, len);
            }
        }
    }
}

void check_guards(const void* user, size_t bytes) {
    std::vector<unsigned char> f(GUARD_FRONT), b(GUARD_BACK);
    const char* u = static_cast<const char*>(user);
    if (hipMemcpy(f.data(), u - GUARD_FRONT, GUARD_FRONT, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b.data(), u + bytes, GUARD_BACK, hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "[guard] could not read the guards of a %zu-byte buffer\n", bytes);
        return;
    }
    size_t bad_f = 0, bad_b = 0, first_b = GUARD_BACK;
    for (size_t i = 0; i < GUARD_FRONT; ++i) bad_f += f[i] != 0xA5;
    for (size_t i = 0; i < GUARD_BACK; ++i)
        if (b[i] != 0xA5) {
            ++bad_b;
            if (first_b == GUARD_BACK) first_b = i;
        }
    if (bad_f || bad_b)
        std::fprintf(stderr, "[guard] %zu-byte buffer: %zu bytes written before it, %zu after it (first at +%zu)\n",
                     bytes, bad_f, bad_b, first_b);
}

void copy_host(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (kind == hipMemcpyDeviceToDevice) {
        CB_HIP(hipMemcpyAsync(dst, src, bytes, kind, st));
        return;
    }
    // one row split into pieces, or rows of kStageBytes
    const size_t w = std::min(bytes, kStageBytes);
    if (w == 0) return;
    const size_t rows = bytes / w, tail = bytes - rows * w;
    copy_host_2d(st, dst, w, src, w, w, rows, kind);
    if (tail) copy_host_2d(st, (char*)dst + rows * w, tail, (const char*)src + rows * w, tail, tail, 1, kind);
}
}  // namespace cbamd

using namespace cbamd;

void cb_context::resolve_profile() {
    for (auto& t : pending) {
        float ms = 0.f;
        if (hipEventSynchronize(t.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            auto& tot = totals[t.name];
            tot.first += ms;
            tot.second += 1;
        }
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    pending.clear();
}

extern "C" {

const char* cb_last_error(void) { return g_last_error.c_str(); }

int cb_context_create(int device, cb_context** out) {
    return guarded([&] {
        CB_REQUIRE(out != nullptr, "cb_context_create: out is NULL");
        int count = 0;
        CB_HIP(hipGetDeviceCount(&count));
        CB_REQUIRE(device >= 0 && device < count,
                   "cb_context_create: device " + std::to_string(device) + " not present (" +
                       std::to_string(count) + " visible)");
        CB_HIP(hipSetDevice(device));
        auto* c = new cb_context();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error(CB_ERR_DEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        }
        c->own_stream = true;
        *out = c;
    });
}

void cb_context_destroy(cb_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->resolve_profile();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (tls_stream == ctx->stream) tls_stream = nullptr;
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int cb_context_set_stream(cb_context* ctx, void* hip_stream) {
    return guarded([&] {
        CB_REQUIRE(ctx, "cb_context_set_stream: ctx is NULL");
        set_device(ctx);
        if (ctx->own_stream && ctx->stream) {
            CB_HIP(hipStreamSynchronize(ctx->stream));
            CB_HIP(hipStreamDestroy(ctx->stream));
            ctx->stream = nullptr;
            tls_stream = nullptr;
            ctx->own_stream = false;
        }
        if (hip_stream) {
            ctx->stream = (hipStream_t)hip_stream;
        } else {
            CB_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
        tls_stream = ctx->stream;
    });
}

void* cb_context_stream(cb_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cb_context_synchronize(cb_context* ctx) {
    return guarded([&] {
        CB_REQUIRE(ctx, "cb_context_synchronize: ctx is NULL");
        set_device(ctx);
        CB_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cb_context_set_profiling(cb_context* ctx, int enabled) {
    return guarded([&] {
        CB_REQUIRE(ctx, "cb_context_set_profiling: ctx is NULL");
        ctx->profiling = enabled != 0;
    });
}

int cb_context_set_graphs(cb_context* ctx, int enabled) {
    return guarded([&] {
        CB_REQUIRE(ctx, "cb_context_set_graphs: ctx is NULL");
        ctx->graphs = enabled != 0;
    });
}

int cb_context_reset_profile(cb_context* ctx) {
    return guarded([&] {
        CB_REQUIRE(ctx, "cb_context_reset_profile: ctx is NULL");
        set_device(ctx);
        ctx->resolve_profile();
        ctx->totals.clear();
    });
}

int cb_context_kernel_time(cb_context* ctx, const char* kernel, double* total_ms,
                           int64_t* launches) {
    return guarded([&] {
        CB_REQUIRE(ctx && kernel && total_ms && launches, "cb_context_kernel_time: NULL argument");
        set_device(ctx);
        ctx->resolve_profile();
        auto it = ctx->totals.find(kernel);
        *total_ms = it == ctx->totals.end() ? 0.0 : it->second.first;
        *launches = it == ctx->totals.end() ? 0 : it->second.second;
    });
}

static cb_genotypes* new_genotypes(cb_context* ctx, int64_t n, int64_t m) {
    CB_REQUIRE(n > 0, "No samples found in .fam file");
    CB_REQUIRE(m > 0, "Empty dataset provided to GrmBuilder");
    CB_REQUIRE(n < (int64_t(1) << 31) && m < (int64_t(1) << 31), "genotype dimensions exceed 2^31");
    auto* g = new cb_genotypes();
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    g->bytes_per_snp = ceil_div(n, 4);
    g->pitch = round_up(g->bytes_per_snp, 16);
    try {
        g->bed.alloc((size_t)g->pitch * (size_t)m);
        CB_HIP(hipMemsetAsync(g->bed.p, 0, g->bed.bytes(), ctx->stream));
    } catch (...) {
        delete g;
        throw;
    }
    return g;
}

int cb_genotypes_from_bed(cb_context* ctx, const uint8_t* bed, size_t bed_size, int64_t n_samples,
                          int64_t n_snps, cb_genotypes** out) {
    return guarded([&] {
        CB_REQUIRE(ctx && out, "cb_genotypes_from_bed: NULL argument");
        CB_REQUIRE(bed != nullptr || bed_size == 0, "cb_genotypes_from_bed: bed is NULL");
        CB_REQUIRE(bed_size >= 3, "Invalid .bed file (too small)");
        CB_REQUIRE(bed[0] == 0x6C && bed[1] == 0x1B, "Invalid .bed file magic number");
        CB_REQUIRE(bed[2] == 0x01, "Only SNP-major mode (0x01) .bed files are supported");
        set_device(ctx);
        const int64_t bps = ceil_div(n_samples, 4);
        CB_REQUIRE(n_samples > 0, "No samples found in .fam file");
        CB_REQUIRE(n_snps > 0, "Empty dataset provided to GrmBuilder");
        CB_REQUIRE((int64_t)bed_size >= 3 + n_snps * bps,
                   ".bed file size (" + std::to_string(bed_size) + ") < expected (" +
                       std::to_string(3 + n_snps * bps) + ")");
        cb_genotypes* g = new_genotypes(ctx, n_samples, n_snps);
        try {
            copy_host_2d(ctx->stream, g->bed.p, g->pitch, bed + 3, bps, bps, n_snps, hipMemcpyHostToDevice);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int cb_genotypes_from_device_rows(cb_context* ctx, const void* d_rows, size_t row_pitch,
                                  int64_t n_samples, int64_t n_snps, cb_genotypes** out) {
    return guarded([&] {
        CB_REQUIRE(ctx && out && d_rows, "cb_genotypes_from_device_rows: NULL argument");
        set_device(ctx);
        const int64_t bps = ceil_div(n_samples, 4);
        CB_REQUIRE((int64_t)row_pitch >= bps, "cb_genotypes_from_device_rows: row_pitch < ceil(n/4)");
        cb_genotypes* g = new_genotypes(ctx, n_samples, n_snps);
        hipError_t e = hipMemcpy2DAsync(g->bed.p, g->pitch, d_rows, row_pitch, bps, n_snps,
                                        hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            delete g;
            throw Error(CB_ERR_DEVICE, std::string("BED device copy failed: ") + hipGetErrorString(e));
        }
        *out = g;
    });
}

void cb_genotypes_destroy(cb_genotypes* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    tls_stream = g->ctx->stream;  // its buffers are released behind the context's queued work
    delete g;
}

int cb_genotypes_dims(const cb_genotypes* g, int64_t* n_samples, int64_t* n_snps) {
    return guarded([&] {
        CB_REQUIRE(g, "cb_genotypes_dims: g is NULL");
        if (n_samples) *n_samples = g->n;
        if (n_snps) *n_snps = g->m;
    });
}

}  // extern "C"
