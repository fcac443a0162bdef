// Gauss quadrature of the Krylov recurrences on the device, in FP64.
//
// Reference: after each Lanczos run imate calls LAPACK ?stev('V') on the
// tridiagonal T (alpha diagonal, beta off-diagonal), takes the eigenvalues
// theta_i and the first component tau_i of each eigenvector, and forms
// n * sum_i tau_i^2 f(theta_i^p) (imate/_c_trace_estimator/c_trace_estimator.cpp:
// 545-621, diagonalization.cpp:88-118). For Golub-Kahn it calls ?bdsdc on the
// upper bidiagonal B and uses theta = sigma^2 with tau the first component of
// each right singular vector (c_trace_estimator.cpp:505-535,
// diagonalization.cpp:189-231). With B = I the Ritz values of A + t_j I are
// the Ritz values of A + t_0 I shifted by t_j - t_0 (c_trace_estimator.cpp:
// 564-594, c_affine_matrix_function.cpp get_eigenvalue).
//
// Here each thread owns one probe and runs the implicit QL algorithm with
// Wilkinson-type shifts (Bowdler, Martin, Reinsch & Wilkinson 1968; the
// EISPACK tql2 scheme) on its tridiagonal, accumulating only the first row of
// the eigenvector matrix (Golub & Welsch 1969): every plane rotation of QL is
// applied to that row alone, O(k) per rotation. The bidiagonal case is solved
// as the 2k x 2k Golub-Kahan tridiagonal with zero diagonal and off-diagonal
// (alpha_0, beta_0, alpha_1, ..., alpha_{k-1}), whose eigenvalues are
// +-sigma_i; the eigenvector of +sigma_i is (v_0, u_0, v_1, u_1, ...)/sqrt(2),
// so tau_i^2 = 2 z_0^2. The k largest eigenvalues are taken.
//
// Per-thread arrays live in LDS, interleaved by thread.

#include <cfloat>
#include <cmath>

#include "slq_internal.hpp"
#include "ql.hpp"

namespace slq {

namespace {

// Probes per workgroup: one. Each probe's QL is a chain of dependent FP64
// operations; a wavefront per probe (64 of them, on as many compute units)
// runs no lane in lockstep with another probe's different iteration counts:
// measured 434 against 522 us at lanczos_degree 30 and 1.62 against 2.04 ms
// at 60 for 16 probes per workgroup (8 and 4 per workgroup: no better).
constexpr int kQuadThreads = 1;

__global__ __launch_bounds__(64) void quadrature_kernel(QuadArgs a)
{
    extern __shared__ double sh[];
    const int t = threadIdx.x;
    const int p = blockIdx.x * kQuadThreads + t;
    const int K2 = a.gram ? 2 * a.m : a.m;
    double* d = sh;
    double* e = sh + (size_t) K2 * kQuadThreads;
    double* z = sh + (size_t) 2 * K2 * kQuadThreads;
    const int S = kQuadThreads;
    d += t;
    e += t;
    z += t;
    if (p >= a.count) return;
    quadrature_one(a, a.alpha + p, a.beta + p, kBlock, a.sizes[p], d, e, z, S,
                   a.samples + (size_t) p * a.q_stride,
                   a.theta_out ? a.theta_out + (size_t) p * a.m : nullptr,
                   a.tau_out ? a.tau_out + (size_t) p * a.m : nullptr);
}

}  // namespace

int run_quadrature(hipStream_t s, const double* alpha, const double* beta,
                   const int32_t* sizes, int count, int m, int gram, int func,
                   const double* func_params, double exponent, double n,
                   const double* params_dev, int q,
                   int shift, int j_out, int q_stride, double* samples_dev,
                   double* theta_dev, double* tau_dev)
{
    QuadArgs a;
    a.alpha = alpha;
    a.beta = beta;
    a.sizes = sizes;
    a.count = count;
    a.m = m;
    a.gram = gram;
    a.func = func;
    for (int i = 0; i < 4; ++i) a.fp[i] = func_params ? func_params[i] : 0.0;
    a.exponent = exponent;
    a.n = n;
    a.params = params_dev;
    a.q = q;
    a.shift = shift;
    a.j_out = j_out;
    a.q_stride = q_stride;
    a.samples = samples_dev;
    a.theta_out = theta_dev;
    a.tau_out = tau_dev;
    const int K2 = gram ? 2 * m : m;
    const size_t lds = sizeof(double) * 3 * (size_t) K2 * kQuadThreads;
    if (lds > 160 * 1024)
    {
        set_error("lanczos_degree %d too large for the device eigensolver", m);
        return SLQ_E_ARG;
    }
    const int grid = (count + kQuadThreads - 1) / kQuadThreads;
    if (lds > 64 * 1024)
        SLQ_CHECK(hipFuncSetAttribute((const void*) quadrature_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds));
    hipLaunchKernelGGL(quadrature_kernel, dim3(grid), dim3(kQuadThreads), lds, s, a);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

}  // namespace slq
