// Eigen-bent inverse of a symmetric PSD matrix — the counterpart of
// invert_psd_matrix (src/core/matrix_io.cpp:488-501), used for G^-1:
//   M + ridge I = U diag(l) U^T,  eps = max(1e-12 max(l), 1e-6),
//   l <- max(l, eps),  M^-1 = (U diag(1/l)) U^T.
// The symmetric eigensolve is rocSOLVER's divide-and-conquer (dsyevd), the
// two n^3 products a plain rocBLAS dgemm; the small kernels around them
// (ridge, floor and column scaling) are here.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "cb_blas.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

#define CB_ROCBLAS(call)                                                                     \
    do {                                                                                     \
        rocblas_status s_ = (call);                                                          \
        if (s_ != rocblas_status_success)                                                    \
            throw Error(CB_ERR_DEVICE, std::string(#call " failed: ") + rocblas_status_to_string(s_)); \
        blas_drain();                                                                        \
    } while (0)

__global__ void k_add_diag(double* A, int64_t n, int64_t ld, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i * ld + i] += v;
}

// V[:, j] = U[:, j] / max(l_j, eps), eps from the largest eigenvalue (l ascending)
__global__ void k_scale_cols(const double* __restrict__ U, double* __restrict__ V, const double* __restrict__ l,
                             int64_t n, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    const double eps = fmax(1e-12 * l[n - 1], 1e-6);
    const double lj = l[j] < eps ? eps : l[j];
    V[j * ld + i] = U[j * ld + i] * (1.0 / lj);
}


}  // namespace

namespace cbamd {

void invert_psd(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge, double* dOut,
                int64_t ldo, hipMemcpyKind in_kind, hipMemcpyKind out_kind, double* min_eig, double* max_eig) {
    CB_REQUIRE(n > 0 && n < (int64_t)1 << 31 && ldm >= n && ldo >= n, "cb_invert_psd: bad dimensions");
    set_device(ctx);  // also points tls_stream at this context's stream (drains, releases)
    hipStream_t st = ctx->stream;
    DevBuf<double> U, V, l, e;
    DevBuf<int> info;
    U.alloc((size_t)n * n);
    V.alloc((size_t)n * n);
    l.alloc(n);
    e.alloc(n);
    info.alloc(1);
    // column-major n x n copy (M is symmetric, so row- or column-major input is the same matrix)
    copy_host_2d(st, U.p, n * 8, dM, ldm * 8, n * 8, n, in_kind);
    if (ridge > 0.0) k_add_diag<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(U.p, n, n, ridge);
    CB_HIP(hipGetLastError());
    BlasHandle h(st);
    CB_BLAS(h, rocsolver_dsyevd(h.h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, U.p,
                                (rocblas_int)n, l.p, e.p, info.p), "rocsolver_dsyevd");
    int hinfo = 0;
    double lo = 0, hi = 0;
    copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    copy_host(st, &lo, l.p, 8, hipMemcpyDeviceToHost);
    copy_host(st, &hi, l.p + (n - 1), 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    CB_REQUIRE(hinfo == 0, "cb_invert_psd: eigensolver did not converge");
    if (min_eig) *min_eig = lo;
    if (max_eig) *max_eig = hi;
    k_scale_cols<<<dim3((unsigned)ceil_div(n, 256), (unsigned)n), 256, 0, st>>>(U.p, V.p, l.p, n, n);
    CB_HIP(hipGetLastError());
    // out = V U^T (column-major); the result is symmetric up to rounding
    const double one = 1.0, zero = 0.0;
    DevBuf<double> O;
    double* dst = dOut;
    int64_t ldd = ldo;
    if (out_kind != hipMemcpyDeviceToDevice) {
        O.alloc((size_t)n * n);
        dst = O.p;
        ldd = n;
    }
    CB_ROCBLAS(rocblas_set_pointer_mode(h.h, rocblas_pointer_mode_host));
    CB_BLAS(h, rocblas_dgemm(h.h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)n,
                             (rocblas_int)n, (rocblas_int)n, &one, V.p, (rocblas_int)n, U.p, (rocblas_int)n, &zero,
                             dst, (rocblas_int)ldd), "rocblas_dgemm");
    if (out_kind != hipMemcpyDeviceToDevice)
        copy_host_2d(st, dOut, ldo * 8, O.p, n * 8, n * 8, n, out_kind);
    CB_HIP(hipStreamSynchronize(st));
}

}  // namespace cbamd

extern "C" {

int cb_invert_psd(cb_context* ctx, const double* M, int64_t ldm, int64_t n, double ridge, double* out,
                  int64_t ldo, double* min_eig, double* max_eig) {
    return guarded([&] {
        CB_REQUIRE(ctx && M && out, "cb_invert_psd: NULL argument");
        invert_psd(ctx, M, ldm, n, ridge, out, ldo, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, min_eig, max_eig);
    });
}

int cb_invert_psd_device(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge, double* dout,
                         int64_t ldo, double* min_eig, double* max_eig) {
    return guarded([&] {
        CB_REQUIRE(ctx && dM && dout, "cb_invert_psd_device: NULL argument");
        invert_psd(ctx, dM, ldm, n, ridge, dout, ldo, hipMemcpyDeviceToDevice, hipMemcpyDeviceToDevice, min_eig,
                   max_eig);
    });
}

}  // extern "C"
