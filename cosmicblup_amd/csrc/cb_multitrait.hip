// Multi-trait GBLUP on one device — the counterparts of MvEigenVce
// (src/core/eigen_vce.cpp:106-262, with the options of
// src/io/variance_estimator.cpp:285-294) and GSSolve::solveMultiTrait
// (src/solve/mixed_model_solver.cpp:425-524).
//
// VCE. Covariate projection: A = the last n - c columns of the full Q of a
// QR of X (rocSOLVER geqrf + orgqr), G_r = A' G A and Y_r = A' Y (rocBLAS),
// G_r = U diag(d) U' (dsyevd), rows y_i of U' Y_r. The EM then needs only
// (d_i, y_i): k_mv_em computes each row's contribution to S_g and S_e
// (t x t, t <= 4) in registers and reduces them over fixed row chunks; the
// chunks are summed in order on the host, so every run gives the same bits.
//
// Solve. The Kronecker MME [Ve^-1 (x) X'X, Ve^-1 (x) X'; ., Ve^-1 (x) I +
// Vg^-1 (x) G^-1] (unknowns trait-major) is assembled densely on the device
// by one kernel and solved by Cholesky (dpotrf + dpotrs).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cb_blas.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int MV_CHUNKS = 256;
constexpr int MV_MAXT = 4;

#define CB_RBS(call, what)                                                                             \
    do {                                                                                               \
        if ((call) != rocblas_status_success) throw Error(CB_ERR_DEVICE, std::string(what " failed")); \
        blas_drain();                                                                                  \
    } while (0)


// in-register t x t inverse of an SPD matrix by Gauss-Jordan (t <= 4)
template <int T>
__device__ void inv_small(const double (&a)[T][T], double (&inv)[T][T]) {
    double m[T][2 * T];
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) {
            m[i][j] = a[i][j];
            m[i][T + j] = i == j ? 1.0 : 0.0;
        }
#pragma unroll
    for (int k = 0; k < T; ++k) {
        const double piv = 1.0 / m[k][k];  // SPD: no pivoting needed
#pragma unroll
        for (int j = 0; j < 2 * T; ++j) m[k][j] *= piv;
#pragma unroll
        for (int i = 0; i < T; ++i) {
            if (i == k) continue;
            const double f = m[i][k];
#pragma unroll
            for (int j = 0; j < 2 * T; ++j) m[i][j] -= f * m[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < T; ++i)
#pragma unroll
        for (int j = 0; j < T; ++j) inv[i][j] = m[i][T + j];
}

// One EM pass (MvEigenVce::updateEM): per row i of the rotated data,
//   V_i = d_i Vg + Ve,  W_g = d_i Vg V_i^-1,  u = W_g y_i,
//   S_g += u u' / d_i + Vg - W_g Vg   (Vg alone when d_i <= 1e-12),
//   W_e = Ve V_i^-1,  e = W_e y_i,  S_e += e e' + Ve - W_e Ve.
// part[chunk][0 .. T*T) = S_g partial, [T*T .. 2 T*T) = S_e partial.
template <int T>
__global__ __launch_bounds__(256) void k_mv_em(const double* __restrict__ d, const double* __restrict__ Y,
                                               int64_t ns, int64_t ldy, const double* __restrict__ vgve,
                                               double* __restrict__ part) {
    __shared__ double red[256];
    double Vg[T][T], Ve[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) {
            Vg[a][b] = vgve[a * T + b];
            Ve[a][b] = vgve[T * T + a * T + b];
        }
    double sg[T][T], se[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) sg[a][b] = se[a][b] = 0.0;
    const int64_t per = (ns + MV_CHUNKS - 1) / MV_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(ns, r0 + per);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
        const double di = d[i];
        double V[T][T], P[T][T], y[T];
#pragma unroll
        for (int a = 0; a < T; ++a) {
            y[a] = Y[(int64_t)a * ldy + i];
#pragma unroll
            for (int b = 0; b < T; ++b) V[a][b] = Vg[a][b] * di + Ve[a][b];
        }
        inv_small<T>(V, P);
        double Wg[T][T], We[T][T], u[T], e[T];
#pragma unroll
        for (int a = 0; a < T; ++a)
#pragma unroll
            for (int b = 0; b < T; ++b) {
                double sgv = 0, sev = 0;
#pragma unroll
                for (int k = 0; k < T; ++k) {
                    sgv += Vg[a][k] * di * P[k][b];
                    sev += Ve[a][k] * P[k][b];
                }
                Wg[a][b] = sgv;
                We[a][b] = sev;
            }
#pragma unroll
        for (int a = 0; a < T; ++a) {
            double su = 0, sv = 0;
#pragma unroll
            for (int k = 0; k < T; ++k) {
                su += Wg[a][k] * y[k];
                sv += We[a][k] * y[k];
            }
            u[a] = su;
            e[a] = sv;
        }
        const bool idf = di > 1e-12;
#pragma unroll
        for (int a = 0; a < T; ++a)
#pragma unroll
            for (int b = 0; b < T; ++b) {
                double wvg = 0, wve = 0;
#pragma unroll
                for (int k = 0; k < T; ++k) {
                    wvg += Wg[a][k] * Vg[k][b];
                    wve += We[a][k] * Ve[k][b];
                }
                sg[a][b] += idf ? (u[a] * u[b] / di + Vg[a][b] - wvg) : Vg[a][b];
                se[a][b] += e[a] * e[b] + Ve[a][b] - wve;
            }
    }
    // block reduction, entry by entry, fixed order
#pragma unroll
    for (int q = 0; q < 2 * T * T; ++q) {
        const int a = (q % (T * T)) / T, b = q % T;
        red[threadIdx.x] = q < T * T ? sg[a][b] : se[a][b];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * 2 * T * T + q] = red[0];
        __syncthreads();
    }
}

// Dense Kronecker MME of solveMultiTrait (column-major, ld = dim), unknowns
// [beta: trait-major blocks of c | u: trait-major blocks of n].
__global__ void k_mv_lhs(const double* __restrict__ XtX, const double* __restrict__ X, const double* __restrict__ Gi,
                         int64_t ldg, const double* __restrict__ vi, int64_t n, int64_t c, int t,
                         double* __restrict__ L, int64_t dim) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = blockIdx.y;
    if (r >= dim) return;
    const int64_t ct = c * t;
    const double* Vei = vi;
    const double* Vgi = vi + t * t;
    double v;
    if (r < ct && s < ct) {
        const int64_t a = r / c, i = r % c, b = s / c, j = s % c;
        v = Vei[a * t + b] * XtX[j * c + i];
    } else if (r < ct) {
        const int64_t a = r / c, i = r % c, b = (s - ct) / n, j = (s - ct) % n;
        v = Vei[a * t + b] * X[i * n + j];
    } else if (s < ct) {
        const int64_t a = (r - ct) / n, i = (r - ct) % n, b = s / c, j = s % c;
        v = Vei[a * t + b] * X[j * n + i];
    } else {
        const int64_t a = (r - ct) / n, i = (r - ct) % n, b = (s - ct) / n, j = (s - ct) % n;
        v = (i == j ? Vei[a * t + b] : 0.0) + Vgi[a * t + b] * Gi[j * ldg + i];
    }
    L[s * dim + r] = v;
}

std::vector<double> inv_host(const std::vector<double>& A, int t) {
    std::vector<double> m(A), inv((size_t)t * t, 0.0);
    for (int i = 0; i < t; ++i) inv[i * t + i] = 1.0;
    for (int k = 0; k < t; ++k) {
        int pr = k;
        for (int i = k + 1; i < t; ++i)
            if (std::fabs(m[i * t + k]) > std::fabs(m[pr * t + k])) pr = i;
        CB_REQUIRE(std::fabs(m[pr * t + k]) > 0, "multi-trait: singular covariance matrix");
        for (int j = 0; j < t; ++j) {
            std::swap(m[k * t + j], m[pr * t + j]);
            std::swap(inv[k * t + j], inv[pr * t + j]);
        }
        const double piv = 1.0 / m[k * t + k];
        for (int j = 0; j < t; ++j) {
            m[k * t + j] *= piv;
            inv[k * t + j] *= piv;
        }
        for (int i = 0; i < t; ++i) {
            if (i == k) continue;
            const double f = m[i * t + k];
            for (int j = 0; j < t; ++j) {
                m[i * t + j] -= f * m[k * t + j];
                inv[i * t + j] -= f * inv[k * t + j];
            }
        }
    }
    return inv;
}

template <int T>
void launch_em(const double* d, const double* Y, int64_t ns, int64_t ldy, const double* vgve, double* part,
               hipStream_t st) {
    k_mv_em<T><<<MV_CHUNKS, 256, 0, st>>>(d, Y, ns, ldy, vgve, part);
}

}  // namespace

extern "C" {

void cb_mv_options_default(cb_mv_options* o) {
    if (!o) return;
    o->max_iter = 500;  // max(500, 25 * ai_maxit) with ai_maxit 20 (variance_estimator.cpp:291)
    o->tol = 1e-4;      // vce_tol (config.h:57)
}

int cb_mv_vce(cb_context* ctx, int64_t n, int64_t t, int64_t c, const double* Y, const double* X, const double* dG,
              int64_t ldg, const cb_mv_options* opts, double* Vg_out, double* Ve_out, int64_t* iterations,
              int* converged) {
    return guarded([&] {
        CB_REQUIRE(ctx && Y && dG && opts && Vg_out && Ve_out, "cb_mv_vce: NULL argument");
        CB_REQUIRE(t >= 1 && t <= MV_MAXT, "cb_mv_vce: 1 to 4 traits");
        CB_REQUIRE(n > c + 1 && c >= 0 && (c == 0 || X) && ldg >= n && n < INT32_MAX, "cb_mv_vce: bad dimensions");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        BlasHandle h(st);
        CB_RBS(rocblas_set_pointer_mode(h.h, rocblas_pointer_mode_host), "pointer mode");
        const int64_t ns = n - c;
        DevBuf<double> Qf, tau, T1, Gr, dY, Yr, Uy, dd, ee;
        DevBuf<int> info;
        info.alloc(1);
        dY.alloc((size_t)n * t);
        copy_host(st, dY.p, Y, (size_t)n * t * 8, hipMemcpyHostToDevice);
        const double one = 1.0, zero = 0.0;
        // A = null-space basis of X' (the last n - c columns of the full Q)
        Qf.alloc((size_t)n * n);
        CB_HIP(hipMemsetAsync(Qf.p, 0, Qf.bytes(), st));
        const double* A;
        if (c > 0) {
            tau.alloc(c);
            copy_host(st, Qf.p, X, (size_t)n * c * 8, hipMemcpyHostToDevice);
            CB_BLAS(h, rocsolver_dgeqrf(h.h, (rocblas_int)n, (rocblas_int)c, Qf.p, (rocblas_int)n, tau.p), "dgeqrf");
            CB_BLAS(h, rocsolver_dorgqr(h.h, (rocblas_int)n, (rocblas_int)n, (rocblas_int)c, Qf.p, (rocblas_int)n, tau.p),
                   "dorgqr");
            A = Qf.p + (size_t)c * n;
        } else {
            std::vector<double> I((size_t)n * n, 0.0);
            for (int64_t i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
            copy_host(st, Qf.p, I.data(), Qf.bytes(), hipMemcpyHostToDevice);
            CB_HIP(hipStreamSynchronize(st));
            A = Qf.p;
        }
        // G_r = A' G A, Y_r = A' Y
        T1.alloc((size_t)n * ns);
        Gr.alloc((size_t)ns * ns);
        CB_BLAS(h, rocblas_dgemm(h.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n, (rocblas_int)ns,
                             (rocblas_int)n, &one, dG, (rocblas_int)ldg, A, (rocblas_int)n, &zero, T1.p,
                             (rocblas_int)n), "dgemm(G A)");
        CB_BLAS(h, rocblas_dgemm(h.h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)ns,
                             (rocblas_int)ns, (rocblas_int)n, &one, A, (rocblas_int)n, T1.p, (rocblas_int)n, &zero,
                             Gr.p, (rocblas_int)ns), "dgemm(A' G A)");
        Yr.alloc((size_t)ns * t);
        CB_BLAS(h, rocblas_dgemm(h.h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)ns,
                             (rocblas_int)t, (rocblas_int)n, &one, A, (rocblas_int)n, dY.p, (rocblas_int)n, &zero,
                             Yr.p, (rocblas_int)ns), "dgemm(A' Y)");
        T1.release();
        dd.alloc(ns);
        ee.alloc(ns);
        CB_BLAS(h, rocsolver_dsyevd(h.h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)ns, Gr.p,
                                (rocblas_int)ns, dd.p, ee.p, info.p), "dsyevd(A' G A)");
        int hinfo = 0;
        copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        CB_REQUIRE(hinfo == 0, "cb_mv_vce: eigensolver did not converge");
        Uy.alloc((size_t)ns * t);
        CB_BLAS(h, rocblas_dgemm(h.h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)ns,
                             (rocblas_int)t, (rocblas_int)ns, &one, Gr.p, (rocblas_int)ns, Yr.p, (rocblas_int)ns,
                             &zero, Uy.p, (rocblas_int)ns), "dgemm(U' Y_r)");
        // eigenvalues clamped at 0 (eigen_vce.cpp:127)
        std::vector<double> hd(ns);
        copy_host(st, hd.data(), dd.p, ns * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        for (double& v : hd)
            if (v < 0) v = 0;
        copy_host(st, dd.p, hd.data(), ns * 8, hipMemcpyHostToDevice);
        // start: Vg = Ve = Cov(Y) / 2 with the divisor n - c - 1 (eigen_vce.cpp:152-156)
        const int T = (int)t;
        std::vector<double> mean(T, 0.0), vg((size_t)T * T), ve((size_t)T * T);
        for (int a = 0; a < T; ++a) {
            for (int64_t i = 0; i < n; ++i) mean[a] += Y[(size_t)a * n + i];
            mean[a] /= (double)n;
        }
        for (int a = 0; a < T; ++a)
            for (int b = 0; b < T; ++b) {
                double sum = 0;
                for (int64_t i = 0; i < n; ++i) sum += (Y[(size_t)a * n + i] - mean[a]) * (Y[(size_t)b * n + i] - mean[b]);
                vg[a * T + b] = ve[a * T + b] = 0.5 * sum / (double)(ns - 1);
            }
        DevBuf<double> vgve, part;
        vgve.alloc(2 * T * T);
        part.alloc((size_t)MV_CHUNKS * 2 * T * T);
        std::vector<double> hp((size_t)MV_CHUNKS * 2 * T * T), sg((size_t)T * T), se((size_t)T * T);
        int64_t it = 0;
        int conv = 0;
        auto fro = [&](const std::vector<double>& a, const std::vector<double>* b) {
            double s = 0;
            for (size_t k = 0; k < a.size(); ++k) {
                const double dlt = b ? a[k] - (*b)[k] : a[k];
                s += dlt * dlt;
            }
            return std::sqrt(s);
        };
        while (it < opts->max_iter) {
            std::vector<double> both(vg);
            both.insert(both.end(), ve.begin(), ve.end());
            copy_host(st, vgve.p, both.data(), both.size() * 8, hipMemcpyHostToDevice);
            switch (T) {
                case 1: launch_em<1>(dd.p, Uy.p, ns, ns, vgve.p, part.p, st); break;
                case 2: launch_em<2>(dd.p, Uy.p, ns, ns, vgve.p, part.p, st); break;
                case 3: launch_em<3>(dd.p, Uy.p, ns, ns, vgve.p, part.p, st); break;
                default: launch_em<4>(dd.p, Uy.p, ns, ns, vgve.p, part.p, st); break;
            }
            CB_HIP(hipGetLastError());
            copy_host(st, hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(st));
            std::fill(sg.begin(), sg.end(), 0.0);
            std::fill(se.begin(), se.end(), 0.0);
            for (int k = 0; k < MV_CHUNKS; ++k)
                for (int q = 0; q < T * T; ++q) {
                    sg[q] += hp[(size_t)k * 2 * T * T + q];
                    se[q] += hp[(size_t)k * 2 * T * T + T * T + q];
                }
            for (int q = 0; q < T * T; ++q) {
                sg[q] /= (double)ns;
                se[q] /= (double)ns;
            }
            // relative Frobenius changes (eigen_vce.cpp:240-246)
            const double dg = fro(sg, &vg) / std::max(fro(vg, nullptr), 1e-12);
            const double de = fro(se, &ve) / std::max(fro(ve, nullptr), 1e-12);
            vg = sg;
            ve = se;
            ++it;
            if (dg < opts->tol && de < opts->tol) {
                conv = 1;
                break;
            }
        }
        std::copy(vg.begin(), vg.end(), Vg_out);
        std::copy(ve.begin(), ve.end(), Ve_out);
        if (iterations) *iterations = it;
        if (converged) *converged = conv;
    });
}

int cb_mv_solve(cb_context* ctx, int64_t n, int64_t t, int64_t c, const double* Y, const double* X,
                const double* dGinv, int64_t ldg, const double* Vg, const double* Ve, double* beta, double* u) {
    return guarded([&] {
        CB_REQUIRE(ctx && Y && dGinv && Vg && Ve && u && (c == 0 || (X && beta)), "cb_mv_solve: NULL argument");
        CB_REQUIRE(t >= 1 && n > 0 && c >= 0 && ldg >= n, "cb_mv_solve: bad dimensions");
        const int64_t ct = c * t, dim = ct + n * t;
        CB_REQUIRE(dim <= 60000, "cb_mv_solve: dense MME limited to (n + c) t <= 60000");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const int T = (int)t;
        std::vector<double> vg(Vg, Vg + T * T), ve(Ve, Ve + T * T);
        std::vector<double> vei = inv_host(ve, T), vgi = inv_host(vg, T);
        // X'X and the RHS on the host (n x c and n x t inputs)
        std::vector<double> xtx((size_t)c * c, 0.0), rhs(dim, 0.0);
        for (int64_t i = 0; i < c; ++i)
            for (int64_t j = 0; j < c; ++j) {
                double sum = 0;
                for (int64_t r = 0; r < n; ++r) sum += X[i * n + r] * X[j * n + r];
                xtx[j * c + i] = sum;
            }
        // RHS1 = vec(X' Y Ve^-1'), RHS2 = vec(Y Ve^-1') (trait-major)
        for (int b = 0; b < T; ++b)
            for (int64_t r = 0; r < n; ++r) {
                double yv = 0;
                for (int a = 0; a < T; ++a) yv += Y[(size_t)a * n + r] * vei[b * T + a];
                rhs[ct + (size_t)b * n + r] = yv;
                for (int64_t i = 0; i < c; ++i) rhs[(size_t)b * c + i] += X[i * n + r] * yv;
            }
        DevBuf<double> L, dX, dxtx, dvi, b;
        DevBuf<int> info;
        info.alloc(1);
        L.alloc((size_t)dim * dim);
        dX.alloc(std::max<int64_t>(1, n * c));
        dxtx.alloc(std::max<int64_t>(1, c * c));
        dvi.alloc(2 * T * T);
        b.alloc(dim);
        if (c > 0) {
            copy_host(st, dX.p, X, (size_t)n * c * 8, hipMemcpyHostToDevice);
            copy_host(st, dxtx.p, xtx.data(), (size_t)c * c * 8, hipMemcpyHostToDevice);
        }
        std::vector<double> vv(vei);
        vv.insert(vv.end(), vgi.begin(), vgi.end());
        copy_host(st, dvi.p, vv.data(), vv.size() * 8, hipMemcpyHostToDevice);
        copy_host(st, b.p, rhs.data(), dim * 8, hipMemcpyHostToDevice);
        k_mv_lhs<<<dim3((unsigned)ceil_div(dim, 256), (unsigned)dim), 256, 0, st>>>(dxtx.p, dX.p, dGinv, ldg, dvi.p, n,
                                                                                    c, T, L.p, dim);
        CB_HIP(hipGetLastError());
        BlasHandle h(st);
        CB_BLAS(h, rocsolver_dpotrf(h.h, rocblas_fill_lower, (rocblas_int)dim, L.p, (rocblas_int)dim, info.p), "dpotrf");
        int hinfo = 0;
        copy_host(st, &hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        CB_REQUIRE(hinfo == 0, "cb_mv_solve: the multi-trait MME is not positive definite");
        CB_BLAS(h, rocsolver_dpotrs(h.h, rocblas_fill_lower, (rocblas_int)dim, 1, L.p, (rocblas_int)dim, b.p,
                                (rocblas_int)dim), "dpotrs");
        std::vector<double> sol(dim);
        copy_host(st, sol.data(), b.p, dim * 8, hipMemcpyDeviceToHost);
        CB_HIP(hipStreamSynchronize(st));
        if (c > 0) std::copy(sol.begin(), sol.begin() + ct, beta);  // c x t column-major
        std::copy(sol.begin() + ct, sol.end(), u);                  // n x t column-major
    });
}

}  // extern "C"
