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

namespace slq {

namespace {

constexpr int kQuadThreads = 16;

struct QuadArgs {
    const double* alpha;      // [m][64]
    const double* beta;       // [m][64]
    const int32_t* sizes;     // [64], 0 = ran all m steps
    int count;
    int m;
    int gram;
    int func;
    double fp[4];             // coefficients of f
    double exponent;
    double n;
    const double* params;     // [q] (device) or nullptr
    int q;
    int shift;                // Ritz shift by params[j] - params[0]
    int j_out;                // column of samples for the unshifted case
    int q_stride;             // row stride of samples
    double* samples;          // [count][q_stride]
    double* theta_out;        // [count][m] or nullptr
    double* tau_out;          // [count][m] or nullptr
};

__device__ __forceinline__ double fpow(double x, double p)
{
    return p == 1.0 ? x : pow(x, p);
}

// f of the matrix functions (imate/functions/*.cpp): logarithm, inverse,
// identity, exponential exp(c x), homographic (a x + b) / (c x + d),
// indicator of [a, b], Gaussian density with mean mu and deviation sigma.
__device__ __forceinline__ double ffunc(int func, const double* fp, double x)
{
    switch (func)
    {
        case SLQ_FUNC_LOG: return log(x);
        case SLQ_FUNC_INV: return 1.0 / x;
        case SLQ_FUNC_EXP: return exp(x * fp[0]);
        case SLQ_FUNC_LINFRAC: return (fp[0] * x + fp[1]) / (fp[2] * x + fp[3]);
        case SLQ_FUNC_INDICATOR: return (x < fp[0] || x > fp[1]) ? 0.0 : 1.0;
        case SLQ_FUNC_GAUSSIAN:
        {
            const double z = (x - fp[0]) / fp[1];
            // 1 / (sigma sqrt(2 pi)) written as imate writes it
            return (0.5 * M_SQRT1_2 * M_2_SQRTPI / fp[1]) * exp(-0.5 * z * z);
        }
        default: return x;
    }
}

// Implicit QL on (d, e) of size K (e[i] couples i and i+1, e[K-1] = 0),
// tracking z = first row of the eigenvector matrix. Arrays are strided by S.
__device__ int ql_first_row(double* __restrict__ d, double* __restrict__ e,
                            double* __restrict__ z, int K, int S)
{
#define D(i) d[(i) * S]
#define E(i) e[(i) * S]
#define Z(i) z[(i) * S]
    for (int l = 0; l < K; ++l)
    {
        int iter = 0;
        while (true)
        {
            int mm = l;
            for (; mm < K - 1; ++mm)
            {
                const double dd = fabs(D(mm)) + fabs(D(mm + 1));
                if (fabs(E(mm)) <= DBL_EPSILON * dd) break;
            }
            if (mm == l) break;
            if (++iter > 64) return -1;
            double g = (D(l + 1) - D(l)) / (2.0 * E(l));
            double r = hypot(g, 1.0);
            g = D(mm) - D(l) + E(l) / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            // The rotation sweep carries D(i + 1) (never written before it
            // is read) and the rotated Z(i + 1) in registers and loads the
            // next E, D, Z ahead, so the dependent chain runs through
            // arithmetic rather than LDS round trips; the arithmetic and
            // its order are unchanged.
            double d1 = D(mm), z1 = Z(mm);
            for (int i = mm - 1; i >= l; --i)
            {
                const double ei = E(i), di = D(i), zi = Z(i);
                const double f = s * ei;
                const double b = c * ei;
                r = hypot(f, g);
                E(i + 1) = r;
                if (r == 0.0)
                {
                    D(i + 1) = d1 - p;
                    Z(i + 1) = z1;
                    E(mm) = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d1 - p;
                r = (di - g) * s + 2.0 * c * b;
                p = s * r;
                D(i + 1) = g + p;
                g = c * r - b;
                Z(i + 1) = s * zi + c * z1;
                z1 = c * zi - s * z1;
                d1 = di;
            }
            if (early) continue;
            Z(l) = z1;
            D(l) = d1 - p;
            E(l) = g;
            E(mm) = 0.0;
        }
    }
    return 0;
#undef D
#undef E
#undef Z
}

__global__ __launch_bounds__(kQuadThreads) void quadrature_kernel(QuadArgs a)
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

    const int k = a.sizes[p] > 0 ? a.sizes[p] : a.m;
    int K;
    if (!a.gram)
    {
        K = k;
        for (int i = 0; i < k; ++i)
        {
            d[i * S] = a.alpha[i * kBlock + p];
            e[i * S] = (i < k - 1) ? a.beta[i * kBlock + p] : 0.0;
            z[i * S] = (i == 0) ? 1.0 : 0.0;
        }
    }
    else
    {
        K = 2 * k;
        for (int i = 0; i < k; ++i)
        {
            d[(2 * i) * S] = 0.0;
            d[(2 * i + 1) * S] = 0.0;
            e[(2 * i) * S] = a.alpha[i * kBlock + p];
            e[(2 * i + 1) * S] = (i < k - 1) ? a.beta[i * kBlock + p] : 0.0;
            z[(2 * i) * S] = (i == 0) ? 1.0 : 0.0;
            z[(2 * i + 1) * S] = 0.0;
        }
    }
    const int info = ql_first_row(d, e, z, K, S);

    if (a.gram)
    {
        // keep the k largest eigenvalues (selection by repeated maximum),
        // compacted into the first k slots as theta = sigma^2, tau^2 = 2 z^2
        for (int i = 0; i < k; ++i)
        {
            int best = i;
            for (int r = i + 1; r < K; ++r)
                if (d[r * S] > d[best * S]) best = r;
            const double db = d[best * S], zb = z[best * S];
            d[best * S] = d[i * S];
            z[best * S] = z[i * S];
            d[i * S] = db;
            z[i * S] = zb;
        }
        for (int i = 0; i < k; ++i)
        {
            const double sg = d[i * S];
            d[i * S] = sg * sg;
            z[i * S] = sqrt(2.0) * z[i * S];
        }
    }
    if (a.theta_out)
        for (int i = 0; i < a.m; ++i)
        {
            a.theta_out[(size_t) p * a.m + i] = i < k ? d[i * S] : 0.0;
            a.tau_out[(size_t) p * a.m + i] = i < k ? z[i * S] : 0.0;
        }

    const int jn = a.shift ? a.q : 1;
    const double t0 = a.shift ? a.params[0] : 0.0;
    for (int j = 0; j < jn; ++j)
    {
        double sum = 0.0;
        const double tj = a.shift ? a.params[j] : 0.0;
        for (int i = 0; i < k; ++i)
        {
            double th = d[i * S];
            if (a.shift) th = (th - t0) + tj;
            const double tau = z[i * S];
            sum += tau * tau * ffunc(a.func, a.fp, fpow(th, a.exponent));
        }
        const int col = a.shift ? j : a.j_out;
        a.samples[(size_t) p * a.q_stride + col] = info == 0 ? a.n * sum : NAN;
    }
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
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int) lds));
    hipLaunchKernelGGL(quadrature_kernel, dim3(grid), dim3(kQuadThreads), lds, s, a);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

}  // namespace slq
