// Device and host implicit-QL eigensolve and Gauss quadrature of a Lanczos
// tridiagonal or Golub-Kahn bidiagonal, and the matrix functions, in FP64.
// Included by quadrature.hip (one probe per GPU thread) and by host_slq.cpp
// (the CPU path), so both run the same arithmetic.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "slq_hip.h"

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define SLQ_HD __host__ __device__
#else
#define SLQ_HD
#endif

namespace slq {

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

SLQ_HD inline double fpow(double x, double p)
{
    return p == 1.0 ? x : pow(x, p);
}

// f of the matrix functions (imate/functions/*.cpp): logarithm, inverse,
// identity, exponential exp(c x), homographic (a x + b) / (c x + d),
// indicator of [a, b], Gaussian density with mean mu and deviation sigma.
SLQ_HD inline double ffunc(int func, const double* fp, double x)
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
SLQ_HD inline int ql_first_row(double* __restrict__ d, double* __restrict__ e,
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

// Eigensolve and Gauss quadrature of one recurrence: alpha[i * ab_stride],
// beta[i * ab_stride] (i < m), its size (0 = all m steps), work arrays d, e,
// z of 2 m entries (m without gram) strided by S; writes out[j] (or
// out[j_out]) for the inquiries and theta / tau when non-null. Shared by
// the device kernel (one probe per thread, arrays in LDS) and the host path.
SLQ_HD inline void quadrature_one(const QuadArgs& a, const double* alpha,
                                               const double* beta, int64_t ab_stride,
                                               int size, double* d, double* e, double* z,
                                               int S, double* out, double* theta_out,
                                               double* tau_out)
{
    const int k = size > 0 ? size : a.m;
    int K;
    if (!a.gram)
    {
        K = k;
        for (int i = 0; i < k; ++i)
        {
            d[i * S] = alpha[i * ab_stride];
            e[i * S] = (i < k - 1) ? beta[i * ab_stride] : 0.0;
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
            e[(2 * i) * S] = alpha[i * ab_stride];
            e[(2 * i + 1) * S] = (i < k - 1) ? beta[i * ab_stride] : 0.0;
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
    if (theta_out)
        for (int i = 0; i < a.m; ++i)
        {
            theta_out[i] = i < k ? d[i * S] : 0.0;
            tau_out[i] = i < k ? z[i * S] : 0.0;
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
        out[col] = info == 0 ? a.n * sum : NAN;
    }
}

}  // namespace slq
