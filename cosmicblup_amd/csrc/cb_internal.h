// Internal runtime shared by the C-ABI translation units: the context
// (device, stream, per-kernel HIP-event timing), thread-local error text,
// and device-buffer helpers. Nothing here crosses the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cosmicblup_amd.h"

namespace cbamd {

void set_error(const std::string& msg);

// Error thrown inside the library; every extern "C" entry point catches it
// and turns it into a status code plus cb_last_error() text.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CB_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::cbamd::Error(CB_ERR_DEVICE,                               \
                                 std::string(#call " failed: ") +             \
                                     hipGetErrorString(e_));                  \
    } while (0)

#define CB_REQUIRE(cond, msg)                                                 \
    do {                                                                      \
        if (!(cond)) throw ::cbamd::Error(CB_ERR_INVALID, msg);               \
    } while (0)

// Wraps an entry point body: returns CB_OK or the error code.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return CB_OK;
    } catch (const Error& e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return CB_ERR_OOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return CB_ERR_INTERNAL;
    }
}

// The stream of the context the calling thread's current library call works
// on (set by set_device and the destroy calls).
inline thread_local hipStream_t tls_stream = nullptr;

// Device buffer owned by a handle. A buffer is released only after the
// calling context's stream has drained, so its memory is never handed to
// another context (another thread's stream) while a kernel queued on this
// one still reads or writes it, whatever hipFree itself waits for.
//
// With CB_DEVICE_GUARDS set in the environment every buffer carries a
// 256-byte guard before it and a 4 KB guard after it, filled with 0xA5 at
// allocation and checked at release; a write outside the buffer is reported
// on stderr with the buffer's size (a debugging aid; off by default).
inline bool device_guards() {
    static const bool on = std::getenv("CB_DEVICE_GUARDS") != nullptr;
    return on;
}
constexpr size_t GUARD_FRONT = 256, GUARD_BACK = 4096;
void check_guards(const void* user, size_t bytes);  // cb_runtime.hip

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (tls_stream) (void)hipStreamSynchronize(tls_stream);
            if (device_guards()) {
                check_guards(p, n * sizeof(T));
                (void)hipFree(reinterpret_cast<char*>(p) - GUARD_FRONT);
            } else {
                (void)hipFree(p);
            }
        }
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        const bool g = device_guards();
        const size_t bytes = count * sizeof(T) + (g ? GUARD_FRONT + GUARD_BACK : 0);
        void* base = nullptr;
        hipError_t e = hipMalloc(&base, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(CB_ERR_OOM, "hipMalloc of " + std::to_string(bytes) +
                                        " bytes failed: " + hipGetErrorString(e));
        }
        if (g) {
            (void)hipMemset(base, 0xA5, GUARD_FRONT);
            (void)hipMemset(static_cast<char*>(base) + GUARD_FRONT + count * sizeof(T), 0xA5, GUARD_BACK);
            (void)hipDeviceSynchronize();
            p = reinterpret_cast<T*>(static_cast<char*>(base) + GUARD_FRONT);
        } else {
            p = static_cast<T*>(base);
        }
        n = count;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct TimedLaunch {
    std::string name;
    hipEvent_t start, stop;
};

}  // namespace cbamd

struct cb_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool profiling = false;
    bool graphs = false;  // replay fixed launch sequences (PCG iterations) as HIP graphs
    std::vector<cbamd::TimedLaunch> pending;
    std::map<std::string, std::pair<double, long>> totals;  // name -> (ms, launches)

    // Resolves pending events into totals (blocks on the last event).
    void resolve_profile();
};

namespace cbamd {

// Run `launch` on ctx->stream; when profiling is on, bracket it with HIP
// events recorded on that same stream and book the time under `name`.
template <class F>
void timed(cb_context* ctx, const char* name, F&& launch) {
    if (!ctx->profiling) {
        launch();
        CB_HIP(hipGetLastError());
        return;
    }
    TimedLaunch t;
    t.name = name;
    CB_HIP(hipEventCreate(&t.start));
    CB_HIP(hipEventCreate(&t.stop));
    CB_HIP(hipEventRecord(t.start, ctx->stream));
    launch();
    CB_HIP(hipGetLastError());
    CB_HIP(hipEventRecord(t.stop, ctx->stream));
    ctx->pending.push_back(t);
}

// Waits for the calling context's stream after a rocBLAS/rocSOLVER call:
// the call's work has finished before the host goes on to read its results
// or release its arguments.
inline void blas_drain() {
    if (tls_stream) CB_HIP(hipStreamSynchronize(tls_stream));
}

inline void set_device(cb_context* ctx) {
    CB_HIP(hipSetDevice(ctx->device));
    tls_stream = ctx->stream;
}

// A fixed sequence of launches on ctx->stream, captured once and replayed
// with hipGraphLaunch. The sequence must not allocate, copy to or from the
// host, or clear memory with hipMemsetAsync (kernels only); its buffers must
// stay in place while the graph lives. Kernel timing is suspended during
// the capture (a replay is not booked per kernel).
struct LaunchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    LaunchGraph() = default;
    LaunchGraph(const LaunchGraph&) = delete;
    LaunchGraph& operator=(const LaunchGraph&) = delete;
    ~LaunchGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    // graphs are used on the context's own or a caller's non-null stream
    // that is not already being captured by the caller
    static bool usable(cb_context* ctx) {
        if (!ctx->graphs || ctx->stream == nullptr) return false;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess) return false;
        return cs == hipStreamCaptureStatusNone;
    }
    template <class F>
    void capture(cb_context* ctx, F&& body) {
        hipStream_t st = ctx->stream;
        const bool prof = ctx->profiling;
        ctx->profiling = false;
        CB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            ctx->profiling = prof;
            throw;
        }
        ctx->profiling = prof;
        CB_HIP(hipStreamEndCapture(st, &graph));
        CB_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    void launch(cb_context* ctx) { CB_HIP(hipGraphLaunch(exec, ctx->stream)); }
};

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Copies between the device and host memory the library does not own
// (pageable), ordered on stream st: the data go through a pinned staging
// buffer of the library's, so that every transfer is a DMA queued on st
// behind the work before it, never a copy the runtime makes on the host at
// the call. Both return once the copy has completed. A device-to-device copy
// is queued as it is.
void copy_host(hipStream_t st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
void copy_host_2d(hipStream_t st, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                  size_t rows, hipMemcpyKind kind);

// Shared between translation units (defined in cb_linalg.hip / cb_pedigree.hip).
// Symmetric n x n FP64 matrix, row-major with leading dimension ld, in device
// (on_device) or host memory, to `path`: format 0 text triplets, 1
// COSMIC_UPPER_PACKED (cb_matrix_io.hip).
void write_symmetric(cb_context* ctx, const double* A, int64_t n, int64_t ld, bool on_device, const char* path,
                     int format);
void invert_psd(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge, double* dOut,
                int64_t ldo, hipMemcpyKind in_kind, hipMemcpyKind out_kind, double* min_eig, double* max_eig);
void pedigree_inbreeding(const std::vector<int32_t>& dam, const std::vector<int32_t>& sire, std::vector<double>& F,
                         std::vector<double>& dii);

}  // namespace cbamd
