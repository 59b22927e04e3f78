// Eigen-bent inverse of a symmetric PSD matrix — the counterpart of
// invert_psd_matrix (src/core/matrix_io.cpp:488-501), used for G^-1:
//   M + ridge I = U diag(l) U^T,  eps = max(1e-12 max(l), 1e-6),
//   l <- max(l, eps),  M^-1 = (U diag(1/l)) U^T.
// The symmetric eigensolve is rocSOLVER's divide-and-conquer (dsyevd), the
// two n^3 products a plain rocBLAS dgemm; the small kernels around them
// (ridge, floor and column scaling) are here.
// When every eigenvalue of M + ridge I exceeds 1e-6 and none exceeds 1e6,
// eps is 1e-6 and the floor changes nothing: the bent inverse is the
// inverse. A row-sum bound (<= 1e6) and a successful Cholesky of
// M + (ridge - 1e-6) I certify both, and the inverse then comes from
// Cholesky (potrf + potri, O(n^3) in BLAS3) instead of the eigensolve, whose
// tridiagonalisation is bound by HBM. A VanRaden G has one eigenvalue below
// the floor (its centring direction): that eigenpair (l1, v) comes from
// inverse iteration on a Cholesky factor of M + ridge I + 1e-3 I, a Cholesky
// of M + (ridge - 1e-6) I + v v' certifies that every other eigenvalue
// exceeds 1e-6, and M + ridge I + (1e-6 - l1) v v' - the matrix with that one
// eigenvalue floored - is inverted by Cholesky. Anything else (more than one
// eigenvalue below the floor, no convergence, a failed certificate) takes
// the eigensolve, as do callers that ask for the extreme eigenvalues.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
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


// s[i] = sum_j |A[i, j]| + |shift| [j == i], a wave per row of the symmetric
// column-major A (row i read as column i)
__global__ __launch_bounds__(256) void k_abs_row_sums(const double* __restrict__ A, int64_t n, int64_t ld,
                                                      double shift, double* __restrict__ s) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* a = A + i * ld;
    double t = 0;
    for (int64_t j = lane; j < n; j += 64) t += fabs(j == i ? a[j] + shift : a[j]);
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off);
    if (lane == 0) s[i] = t;
}

// the upper triangle from the lower one (column-major)
__global__ void k_mirror_lower(double* A, int64_t n, int64_t ld) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // row
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y)                // columns j > i
        if (i < j) A[j * ld + i] = A[i * ld + j];
}

// A[i, j] += s v_i v_j (column-major)
__global__ void k_rank1(double* A, int64_t n, int64_t ld, double s, const double* __restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int64_t j = blockIdx.y; j < n; j += gridDim.y) A[j * ld + i] += s * v[i] * v[j];
}

constexpr double FLOOR_MIN = 1e-6;   // eps = max(1e-12 max(l), 1e-6)
constexpr double FLOOR_SHIFT = 1e-3;  // shift of the factor the inverse iteration solves with
constexpr int FLOOR_MAX_IT = 60;
constexpr double FLOOR_ROW_BOUND = 1e6;  // max(l) <= this keeps eps at 1e-6

}  // namespace

namespace cbamd {

// U := the Cholesky inverse of the positive-definite matrix in U (lower
// factor, then mirrored), copied to out
static void cholesky_inverse_out(cb_context* ctx, BlasHandle& h, DevBuf<double>& U, DevBuf<int>& info, int64_t n,
                                 double* dOut, int64_t ldo, hipMemcpyKind out_kind) {
    hipStream_t st = ctx->stream;
    int hinfo = 0;
    CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)n, U.p, (rocblas_int)n, info.p),
            "rocsolver_dpotrf");
    CB_BLAS(h, rocsolver_dpotri(h.h, rocblas_fill_lower, (rocblas_int)n, U.p, (rocblas_int)n, info.p),
            "rocsolver_dpotri");
    copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    CB_REQUIRE(hinfo == 0, "cb_invert_psd: Cholesky inverse failed after the positive-definite check");
    k_mirror_lower<<<dim3((unsigned)ceil_div(n, 256), (unsigned)std::min<int64_t>(n, 65535)), 256, 0, st>>>(U.p, n, n);
    CB_HIP(hipGetLastError());
    copy_host_2d(st, dOut, ldo * 8, U.p, n * 8, n * 8, n, out_kind);
    CB_HIP(hipStreamSynchronize(st));
}

// The one eigenpair below the floor and the bent inverse by Cholesky (see
// the file header); false (nothing written) when that does not apply. U
// holds M + ridge I on entry (it is overwritten).
static bool inverse_with_one_floored(cb_context* ctx, BlasHandle& h, DevBuf<double>& U, DevBuf<int>& info,
                                     int64_t n, double bound, double* dOut, int64_t ldo, hipMemcpyKind out_kind) {
    hipStream_t st = ctx->stream;
    DevBuf<double> B, dv, dw;
    B.alloc((size_t)n * n);
    dv.alloc(n);
    dw.alloc(n);
    CB_HIP(hipMemcpyAsync(B.p, U.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));  // M + ridge I, kept
    k_add_diag<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(U.p, n, n, FLOOR_SHIFT);
    CB_HIP(hipGetLastError());
    int hinfo = 0;
    CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)n, U.p, (rocblas_int)n, info.p),
            "rocsolver_dpotrf");
    copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) return false;
    // inverse iteration from the centring direction, lightly perturbed
    std::vector<double> v(n), w(n), y(n);
    double nv = 0;
    for (int64_t i = 0; i < n; ++i) {
        v[i] = 1.0 + 1e-3 * std::sin(0.7 * (double)i);
        nv += v[i] * v[i];
    }
    for (double& x : v) x /= std::sqrt(nv);
    const double one = 1.0, zero = 0.0;
    CB_ROCBLAS(rocblas_set_pointer_mode(h.h, rocblas_pointer_mode_host));
    double l1 = 0, res = 0;
    // the direction converges by (l1 + shift) / (l2 + shift) per step until
    // its change reaches the rounding level; two more steps polish it
    int polish = -1;
    for (int it = 0; it < FLOOR_MAX_IT && polish != 0; ++it) {
        copy_host(st, dw.p, v.data(), n * 8, hipMemcpyHostToDevice);
        CB_BLAS(h, rocsolver_dpotrs(h.h, rocblas_fill_lower, (rocblas_int)n, 1, U.p, (rocblas_int)n, dw.p,
                                    (rocblas_int)n), "rocsolver_dpotrs");
        copy_host(st, w.data(), dw.p, n * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        double nw = 0, dot = 0;
        for (int64_t i = 0; i < n; ++i) nw += w[i] * w[i];
        nw = std::sqrt(nw);
        if (!(nw > 0)) return false;
        for (int64_t i = 0; i < n; ++i) {
            w[i] /= nw;
            dot += w[i] * v[i];
        }
        // sign-aligned change of the direction
        double dchg = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double d = w[i] - (dot < 0 ? -v[i] : v[i]);
            dchg = std::max(dchg, std::fabs(d));
        }
        v.swap(w);
        if (polish > 0) --polish;
        else if (polish < 0 && it >= 3 && dchg < 1e-13) polish = 2;
    }
    if (polish != 0) return false;  // no convergence: the eigensolve decides
    // Rayleigh quotient and residual against M + ridge I
    copy_host(st, dv.p, v.data(), n * 8, hipMemcpyHostToDevice);
    CB_BLAS(h, rocblas_dgemv(h.h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, &one, B.p, (rocblas_int)n,
                             dv.p, 1, &zero, dw.p, 1), "rocblas_dgemv");
    copy_host(st, y.data(), dw.p, n * 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) l1 += v[i] * y[i];
    for (int64_t i = 0; i < n; ++i) res = std::max(res, std::fabs(y[i] - l1 * v[i]));
    if (!(l1 < FLOOR_MIN) || !(res <= 1e-12 * bound)) return false;
    // every other eigenvalue above the floor: M + ridge I - 1e-6 I + v v' positive definite
    CB_HIP(hipMemcpyAsync(U.p, B.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
    k_add_diag<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(U.p, n, n, -FLOOR_MIN);
    const dim3 g2((unsigned)ceil_div(n, 256), (unsigned)std::min<int64_t>(n, 65535));
    k_rank1<<<g2, 256, 0, st>>>(U.p, n, n, 1.0, dv.p);
    CB_HIP(hipGetLastError());
    CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)n, U.p, (rocblas_int)n, info.p),
            "rocsolver_dpotrf");
    copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    if (hinfo != 0) return false;
    // the bent matrix: that eigenvalue moved to the floor
    CB_HIP(hipMemcpyAsync(U.p, B.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
    k_rank1<<<g2, 256, 0, st>>>(U.p, n, n, FLOOR_MIN - l1, dv.p);
    CB_HIP(hipGetLastError());
    cholesky_inverse_out(ctx, h, U, info, n, dOut, ldo, out_kind);
    return true;
}

// The bent inverse without the eigensolve, when the floor is inactive or
// acts on one eigenvalue only; false (nothing written) otherwise.
static bool inverse_by_cholesky(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge,
                                double* dOut, int64_t ldo, hipMemcpyKind in_kind, hipMemcpyKind out_kind) {
    hipStream_t st = ctx->stream;
    DevBuf<double> U, rs;
    DevBuf<int> info;
    U.alloc((size_t)n * n);
    rs.alloc(n);
    info.alloc(1);
    copy_host_2d(st, U.p, n * 8, dM, ldm * 8, n * 8, n, in_kind);
    // largest eigenvalue <= the largest absolute row sum
    k_abs_row_sums<<<(unsigned)ceil_div(n, 4), 256, 0, st>>>(U.p, n, n, ridge, rs.p);
    CB_HIP(hipGetLastError());
    std::vector<double> hrs(n);
    copy_host(st, hrs.data(), rs.p, n * 8, hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    double bound = 0;
    for (double v : hrs) bound = std::max(bound, v);
    if (!(bound <= FLOOR_ROW_BOUND)) return false;
    // every eigenvalue above 1e-6: M + (ridge - 1e-6) I positive definite
    k_add_diag<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(U.p, n, n, ridge - FLOOR_MIN);
    CB_HIP(hipGetLastError());
    BlasHandle h(st);
    int hinfo = 0;
    CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)n, U.p, (rocblas_int)n, info.p),
            "rocsolver_dpotrf");
    copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
    CB_HIP(hipStreamSynchronize(st));
    // U := M + ridge I again
    copy_host_2d(st, U.p, n * 8, dM, ldm * 8, n * 8, n, in_kind);
    if (ridge != 0.0) k_add_diag<<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(U.p, n, n, ridge);
    CB_HIP(hipGetLastError());
    if (hinfo != 0) return inverse_with_one_floored(ctx, h, U, info, n, bound, dOut, ldo, out_kind);
    // no eigenvalue below the floor: (M + ridge I)^-1 = L^-T L^-1
    cholesky_inverse_out(ctx, h, U, info, n, dOut, ldo, out_kind);
    return true;
}

void invert_psd(cb_context* ctx, const double* dM, int64_t ldm, int64_t n, double ridge, double* dOut,
                int64_t ldo, hipMemcpyKind in_kind, hipMemcpyKind out_kind, double* min_eig, double* max_eig) {
    CB_REQUIRE(n > 0 && n < (int64_t)1 << 31 && ldm >= n && ldo >= n, "cb_invert_psd: bad dimensions");
    set_device(ctx);  // also points tls_stream at this context's stream (drains, releases)
    if (!min_eig && !max_eig && inverse_by_cholesky(ctx, dM, ldm, n, ridge, dOut, ldo, in_kind, out_kind))
        return;
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
