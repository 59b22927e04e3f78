// Multi-trait GBLUP on one device — the counterparts of MvEigenVce
// (src/core/eigen_vce.cpp:106-262, with the options of
// src/io/variance_estimator.cpp:285-294) and GSSolve::solveMultiTrait
// (src/solve/mixed_model_solver.cpp:425-524).
//
// VCE. Covariate projection: A = the last n - c columns of the full Q of a
// QR of X (rocSOLVER geqrf + orgqr), G_r = A' G A and Y_r = A' Y (rocBLAS),
// G_r = U diag(d) U' (dsyevd), rows y_i of U' Y_r. The EM then needs only
// (d_i, y_i): k_mv_em_run runs every EM iteration in one launch of one
// workgroup, each row's contribution to S_g and S_e (t x t, t <= 4) formed
// in registers and summed in a fixed order, so every run gives the same bits
// and no iteration waits for the host.
//
// Solve. The reference assembles the Kronecker MME [Ve^-1 (x) X'X,
// Ve^-1 (x) X'; ., Ve^-1 (x) I + Vg^-1 (x) G^-1] (unknowns trait-major) and
// factorises it by dense LDLT, O((n t)^3). With complete records and one
// design for every trait that system is t independent single-trait MMEs
// after the canonical transformation of the traits (canonical() below), so
// it is solved here as t single-trait PCG solves on the library's own
// operator kernels (the dense G^-1 product, W'W, block-Jacobi with the stiff
// direction deflated), FP64 to a relative residual of MV_SOLVE_TOL, and
// transformed back: the same solution to rounding, deterministic, and
// O(iterations n^2).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "cb_blas.h"
#include "cb_internal.h"

using namespace cbamd;

namespace {

constexpr int MV_MAXT = 4;
constexpr double MV_SOLVE_TOL = 1e-12;  // relative residual of each transformed-trait solve

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

// One row's EM contribution (MvEigenVce::updateEM): for row i of the
// rotated data,
//   V_i = d_i Vg + Ve,  W_g = d_i Vg V_i^-1,  u = W_g y_i,
//   S_g += u u' / d_i + Vg - W_g Vg   (Vg alone when d_i <= 1e-12),
//   W_e = Ve V_i^-1,  e = W_e y_i,  S_e += e e' + Ve - W_e Ve.
template <int T>
__device__ __forceinline__ void em_row(double di, const double (&y)[T], const double (&Vg)[T][T],
                                       const double (&Ve)[T][T], double (&sg)[T][T], double (&se)[T][T]) {
    double V[T][T], P[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) V[a][b] = Vg[a][b] * di + Ve[a][b];
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

// The whole EM run on the device, one workgroup: each iteration sums the
// rows' contributions (thread t takes rows t, t + 256, ... in order; then a
// butterfly within each wave and the four waves' sums in order, so every
// run gives the same bits), sets Vg = S_g / ns and Ve = S_e / ns, and stops
// when both relative Frobenius changes are below tol (eigen_vce.cpp:240-246)
// or after max_iter iterations. vgve holds Vg then Ve (row-major t x t): the
// start on entry, the estimates on exit; res = (iterations, converged).
template <int T>
__global__ __launch_bounds__(256) void k_mv_em_run(const double* __restrict__ d, const double* __restrict__ Y,
                                                   int64_t ns, int64_t ldy, double* __restrict__ vgve,
                                                   int64_t max_iter, double tol, int64_t* __restrict__ res) {
    constexpr int NQ = 2 * T * T;
    __shared__ double red[4][NQ];
    __shared__ double cur[NQ];  // Vg, then Ve
    __shared__ int state[2];    // converged, stop
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < NQ) cur[tid] = vgve[tid];
    __syncthreads();
    int64_t it = 0;  // counted by thread 0
    for (;;) {
        double Vg[T][T], Ve[T][T], sg[T][T], se[T][T];
#pragma unroll
        for (int a = 0; a < T; ++a)
#pragma unroll
            for (int b = 0; b < T; ++b) {
                Vg[a][b] = cur[a * T + b];
                Ve[a][b] = cur[T * T + a * T + b];
                sg[a][b] = se[a][b] = 0.0;
            }
        for (int64_t i = tid; i < ns; i += 256) {
            double y[T];
#pragma unroll
            for (int a = 0; a < T; ++a) y[a] = Y[(int64_t)a * ldy + i];
            em_row<T>(d[i], y, Vg, Ve, sg, se);
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double v = q < T * T ? sg[q / T][q % T] : se[(q - T * T) / T][(q - T * T) % T];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) red[wave][q] = v;
        }
        __syncthreads();
        if (tid == 0) {
            double S[NQ];
            for (int q = 0; q < NQ; ++q) S[q] = (((red[0][q] + red[1][q]) + red[2][q]) + red[3][q]) / (double)ns;
            double dg = 0, ng = 0, de = 0, ne = 0;
            for (int q = 0; q < T * T; ++q) {
                dg += (S[q] - cur[q]) * (S[q] - cur[q]);
                ng += cur[q] * cur[q];
                de += (S[T * T + q] - cur[T * T + q]) * (S[T * T + q] - cur[T * T + q]);
                ne += cur[T * T + q] * cur[T * T + q];
            }
            const bool conv = sqrt(dg) / fmax(sqrt(ng), 1e-12) < tol && sqrt(de) / fmax(sqrt(ne), 1e-12) < tol;
            for (int q = 0; q < NQ; ++q) cur[q] = S[q];
            ++it;
            state[0] = conv ? 1 : 0;
            state[1] = conv || it >= max_iter ? 1 : 0;
        }
        __syncthreads();
        if (state[1]) break;  // the same for every thread
    }
    if (tid < NQ) vgve[tid] = cur[tid];
    if (tid == 0) {
        res[0] = it;
        res[1] = state[0];
    }
}

// Canonical transformation of a t-trait model with complete records and one
// design for every trait: with Ve = L L' and L^-1 Vg L^-T = W diag(d) W',
// T = L^-T W gives T' Ve T = I and T' Vg T = diag(d), so Y T splits into t
// independent single-trait models with lambda_a = 1 / d_a, and the
// Kronecker MME's solution is the transformed traits' solutions times T^-1
// (T^-1 = W' L'). Matrices row-major t x t.
struct Canonical {
    std::vector<double> d, T, Tinv;
};

Canonical canonical(const std::vector<double>& Vg, const std::vector<double>& Ve, int t) {
    // Cholesky of Ve (row-major lower L)
    std::vector<double> L((size_t)t * t, 0.0);
    for (int j = 0; j < t; ++j) {
        double s = Ve[j * t + j];
        for (int k = 0; k < j; ++k) s -= L[j * t + k] * L[j * t + k];
        CB_REQUIRE(s > 0 && std::isfinite(s), "multi-trait: Ve is not positive definite");
        L[j * t + j] = std::sqrt(s);
        for (int i = j + 1; i < t; ++i) {
            double v = Ve[i * t + j];
            for (int k = 0; k < j; ++k) v -= L[i * t + k] * L[j * t + k];
            L[i * t + j] = v / L[j * t + j];
        }
    }
    // Li = L^-1 (lower, forward substitution on unit columns)
    std::vector<double> Li((size_t)t * t, 0.0);
    for (int cidx = 0; cidx < t; ++cidx)
        for (int i = 0; i < t; ++i) {
            double v = i == cidx ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) v -= L[i * t + k] * Li[k * t + cidx];
            Li[i * t + cidx] = v / L[i * t + i];
        }
    // S = Li Vg Li'
    std::vector<double> A((size_t)t * t, 0.0), S((size_t)t * t, 0.0);
    for (int i = 0; i < t; ++i)
        for (int j = 0; j < t; ++j)
            for (int k = 0; k < t; ++k) A[i * t + j] += Li[i * t + k] * Vg[k * t + j];
    for (int i = 0; i < t; ++i)
        for (int j = 0; j < t; ++j)
            for (int k = 0; k < t; ++k) S[i * t + j] += A[i * t + k] * Li[j * t + k];
    for (int i = 0; i < t; ++i)
        for (int j = 0; j < i; ++j) S[i * t + j] = S[j * t + i] = 0.5 * (S[i * t + j] + S[j * t + i]);
    // cyclic Jacobi eigen-decomposition S = W diag(d) W'
    std::vector<double> W((size_t)t * t, 0.0);
    for (int i = 0; i < t; ++i) W[i * t + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < t; ++i)
            for (int j = 0; j < t; ++j) (i == j ? tot : off) += S[i * t + j] * S[i * t + j];
        if (off <= 1e-30 * tot) break;
        for (int pi = 0; pi < t; ++pi)
            for (int qi = pi + 1; qi < t; ++qi) {
                const double apq = S[pi * t + qi];
                if (apq == 0.0) continue;
                const double th = (S[qi * t + qi] - S[pi * t + pi]) / (2.0 * apq);
                const double tn = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(tn * tn + 1.0), sn = tn * cs;
                for (int k = 0; k < t; ++k) {  // S <- S J (columns p, q)
                    const double skp = S[k * t + pi], skq = S[k * t + qi];
                    S[k * t + pi] = cs * skp - sn * skq;
                    S[k * t + qi] = sn * skp + cs * skq;
                }
                for (int k = 0; k < t; ++k) {  // S <- J' S (rows p, q)
                    const double spk = S[pi * t + k], sqk = S[qi * t + k];
                    S[pi * t + k] = cs * spk - sn * sqk;
                    S[qi * t + k] = sn * spk + cs * sqk;
                }
                for (int k = 0; k < t; ++k) {  // W <- W J
                    const double wkp = W[k * t + pi], wkq = W[k * t + qi];
                    W[k * t + pi] = cs * wkp - sn * wkq;
                    W[k * t + qi] = sn * wkp + cs * wkq;
                }
            }
    }
    Canonical cz;
    cz.d.resize(t);
    double dmax = 0;
    for (int i = 0; i < t; ++i) dmax = std::max(dmax, std::fabs(S[i * t + i]));
    for (int i = 0; i < t; ++i) {
        cz.d[i] = S[i * t + i];
        CB_REQUIRE(cz.d[i] > 1e-12 * dmax && std::isfinite(cz.d[i]),
                   "multi-trait: Vg is not positive definite (the MME needs Vg^-1)");
    }
    // T = Li' W, Tinv = W' L'
    cz.T.assign((size_t)t * t, 0.0);
    cz.Tinv.assign((size_t)t * t, 0.0);
    for (int i = 0; i < t; ++i)
        for (int j = 0; j < t; ++j)
            for (int k = 0; k < t; ++k) {
                cz.T[i * t + j] += Li[k * t + i] * W[k * t + j];
                cz.Tinv[i * t + j] += W[k * t + i] * L[j * t + k];
            }
    return cz;
}

template <int T>
void launch_em(const double* d, const double* Y, int64_t ns, int64_t ldy, double* vgve, int64_t max_iter, double tol,
               int64_t* res, hipStream_t st) {
    k_mv_em_run<T><<<1, 256, 0, st>>>(d, Y, ns, ldy, vgve, max_iter, tol, res);
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
        DevBuf<double> vgve;
        DevBuf<int64_t> res;
        vgve.alloc(2 * T * T);
        res.alloc(2);
        int64_t it = 0;
        int conv = 0;
        if (opts->max_iter > 0) {
            std::vector<double> both(vg);
            both.insert(both.end(), ve.begin(), ve.end());
            copy_host(st, vgve.p, both.data(), both.size() * 8, hipMemcpyHostToDevice);
            switch (T) {
                case 1: launch_em<1>(dd.p, Uy.p, ns, ns, vgve.p, opts->max_iter, opts->tol, res.p, st); break;
                case 2: launch_em<2>(dd.p, Uy.p, ns, ns, vgve.p, opts->max_iter, opts->tol, res.p, st); break;
                case 3: launch_em<3>(dd.p, Uy.p, ns, ns, vgve.p, opts->max_iter, opts->tol, res.p, st); break;
                default: launch_em<4>(dd.p, Uy.p, ns, ns, vgve.p, opts->max_iter, opts->tol, res.p, st); break;
            }
            CB_HIP(hipGetLastError());
            int64_t hres[2];
            copy_host(st, both.data(), vgve.p, both.size() * 8, hipMemcpyDeviceToHost);
            copy_host(st, hres, res.p, sizeof hres, hipMemcpyDeviceToHost);
            CB_HIP(hipStreamSynchronize(st));
            std::copy(both.begin(), both.begin() + T * T, vg.begin());
            std::copy(both.begin() + T * T, both.end(), ve.begin());
            it = hres[0];
            conv = (int)hres[1];
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
        CB_REQUIRE(t >= 1 && t <= MV_MAXT && n > 0 && c >= 0 && ldg >= n && n + c < INT32_MAX,
                   "cb_mv_solve: bad dimensions");
        set_device(ctx);
        const int T = (int)t;
        const Canonical cz = canonical(std::vector<double>(Vg, Vg + T * T), std::vector<double>(Ve, Ve + T * T), T);
        // the single-trait model [X'X X'; X I + lambda G^-1] shared by the
        // transformed traits: one record per individual, fixed row X[i, :]
        std::vector<int64_t> xp(n + 1);
        std::vector<int32_t> xc((size_t)n * c), aid(n);
        std::vector<double> xv((size_t)n * c), y0(n, 0.0);
        for (int64_t i = 0; i < n; ++i) {
            xp[i] = i * c;
            aid[i] = (int32_t)i;
            for (int64_t j = 0; j < c; ++j) {
                xc[(size_t)i * c + j] = (int32_t)j;
                xv[(size_t)i * c + j] = X[j * n + i];
            }
        }
        xp[n] = n * c;
        cb_mme* raw = nullptr;
        if (cb_mme_create(ctx, n, c, n, xp.data(), xc.data(), xv.data(), aid.data(), y0.data(), &raw) != CB_OK)
            throw Error(CB_ERR_DEVICE, std::string("cb_mv_solve: ") + cb_last_error());
        struct Own {
            cb_mme* m;
            ~Own() { cb_mme_destroy(m); }
        } own{raw};
        if (cb_mme_set_qinv_dense_device(raw, dGinv, ldg) != CB_OK)
            throw Error(CB_ERR_DEVICE, std::string("cb_mv_solve: ") + cb_last_error());
        // transformed traits Y* = Y T, each solved with lambda_a = 1 / d_a
        const int64_t dim = c + n;
        std::vector<double> ys((size_t)n), rhs(dim), sol((size_t)dim * T);
        cb_solve_options so;
        cb_solve_options_default(&so);
        so.precision = CB_PREC_FP64;
        so.precond = CB_PRECOND_BLOCK_JACOBI;
        so.stop_rule = CB_STOP_RELATIVE;
        so.tol = MV_SOLVE_TOL;
        so.max_iter = 20000;
        for (int a = 0; a < T; ++a) {
            for (int64_t i = 0; i < n; ++i) {
                double v = 0;
                for (int b = 0; b < T; ++b) v += Y[(size_t)b * n + i] * cz.T[b * T + a];
                ys[i] = v;
            }
            for (int64_t j = 0; j < c; ++j) {
                double v = 0;
                for (int64_t i = 0; i < n; ++i) v += X[j * n + i] * ys[i];
                rhs[j] = v;
            }
            std::copy(ys.begin(), ys.end(), rhs.begin() + c);
            int64_t its = 0;
            double rel = 0;
            if (cb_mme_solve(raw, 1.0 / cz.d[a], &so, 1, rhs.data(), sol.data() + (size_t)a * dim, &its, &rel) != CB_OK)
                throw Error(CB_ERR_NOT_CONVERGED, std::string("cb_mv_solve: ") + cb_last_error());
        }
        // back to the original traits: [beta; u] = [beta*; u*] T^-1
        for (int b = 0; b < T; ++b) {
            for (int64_t r = 0; r < dim; ++r) {
                double v = 0;
                for (int a = 0; a < T; ++a) v += sol[(size_t)a * dim + r] * cz.Tinv[a * T + b];
                if (r < c)
                    beta[(size_t)b * c + r] = v;  // c x t column-major
                else
                    u[(size_t)b * n + (r - c)] = v;  // n x t column-major
            }
        }
    });
}

}  // extern "C"
