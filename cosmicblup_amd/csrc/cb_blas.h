// rocBLAS/rocSOLVER handles for the library's dense steps.
#pragma once
#include <rocblas/rocblas.h>

#include "cb_internal.h"

namespace cbamd {

// A handle on stream st with a device workspace of the library's own. Each
// call is first run in rocBLAS's size-query mode (reserve), that much device
// memory is allocated, and the handle is told to use it: the handle never
// allocates or frees device memory itself, and its workspace is released
// like every DevBuf, behind the context's queued work.
struct BlasHandle {
    rocblas_handle h = nullptr;
    DevBuf<uint8_t> ws;
    explicit BlasHandle(hipStream_t st) {
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            h = nullptr;
            throw Error(CB_ERR_DEVICE, "rocblas_create_handle failed");
        }
        if (rocblas_set_stream(h, st) != rocblas_status_success)
            throw Error(CB_ERR_DEVICE, "rocblas_set_stream failed");
        // no atomics: rocBLAS may otherwise split a product's reduction over
        // workgroups that add into the result atomically, in an order that
        // depends on how the GPU schedules them, and repeated calls would
        // not give bit-identical results.
        if (rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success)
            throw Error(CB_ERR_DEVICE, "rocblas_set_atomics_mode failed");
    }
    BlasHandle(const BlasHandle&) = delete;
    BlasHandle& operator=(const BlasHandle&) = delete;
    ~BlasHandle() {
        if (tls_stream) (void)hipStreamSynchronize(tls_stream);
        if (h) (void)rocblas_destroy_handle(h);
    }
    // calls(): the section's rocBLAS/rocSOLVER calls with their real
    // arguments; in query mode they only report their workspace sizes
    template <class F>
    void reserve(F&& calls) {
        if (rocblas_start_device_memory_size_query(h) != rocblas_status_success)
            throw Error(CB_ERR_DEVICE, "rocblas_start_device_memory_size_query failed");
        calls();
        size_t size = 0;
        if (rocblas_stop_device_memory_size_query(h, &size) != rocblas_status_success)
            throw Error(CB_ERR_DEVICE, "rocblas_stop_device_memory_size_query failed");
        if (size > ws.n) {
            ws.alloc(size);
            if (rocblas_set_workspace(h, ws.p, ws.n) != rocblas_status_success)
                throw Error(CB_ERR_DEVICE, "rocblas_set_workspace failed");
        }
    }
};

}  // namespace cbamd

// One rocBLAS/rocSOLVER call on BlasHandle bh: its workspace reserved first
// (the call run once in size-query mode), then the call, its status checked,
// then the context's stream drained.
#define CB_BLAS(bh, call, what)                                                                         \
    do {                                                                                                \
        (bh).reserve([&] { (void)(call); });                                                            \
        const rocblas_status s_ = (call);                                                               \
        if (s_ != rocblas_status_success)                                                               \
            throw ::cbamd::Error(CB_ERR_DEVICE, std::string(what) + " failed: " + rocblas_status_to_string(s_)); \
        ::cbamd::blas_drain();                                                                          \
    } while (0)
