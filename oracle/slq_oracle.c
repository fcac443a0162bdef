/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of imate's stochastic Lanczos quadrature inner loops
 * (reference: imate 0.25.2, CPU path with USE_CBLAS=0, 32-bit indices).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only to check results or as the timed CPU baseline; the
 * product library (imate_amd/csrc) never links or calls it.
 *
 * Contents and the reference lines they restate:
 *   SplitMix64 seeding ............ _random_generator/split_mix_64.cpp:28-64
 *   xoshiro256** next / jump ...... _random_generator/xoshiro_256_star_star.cpp:33-133
 *   per-thread streams ............ _random_generator/random_number_generator.cpp:70-153
 *   Rademacher fill ............... _random_generator/random_array_generator.cpp:62-165
 *   vector ops (long double) ...... _c_basic_algebra/c_vector_operations.cpp:80-422
 *   CSR matvec / transposed ....... _c_basic_algebra/c_matrix_operations.cpp:469-635
 *   affine A + tI / A + tB ........ _c_linear_operator/c_csr_affine_matrix_function.cpp,
 *                                   c_affine_matrix_function.cpp
 *   Gram-Schmidt .................. _c_trace_estimator/c_orthogonalization.cpp:125-224
 *   Lanczos ....................... _c_trace_estimator/c_lanczos_tridiagonalization.cpp:117-235
 *   Golub-Kahn .................... _c_trace_estimator/c_golub_kahn_bidiagonalization.cpp:111-231
 *   erf_inv ....................... _c_trace_estimator/special_functions.cpp:63-104
 *
 * Pinning (how each part is checked; tests/test_oracle_*.py):
 *   - RNG: SplitMix64 known answers, xoshiro jump == 2^128 steps via GF(2)
 *     matrix powers, and bit-exact equality with the reference's own RNG
 *     sources compiled into oracle/_ref (golden fixtures in tests/golden).
 *   - Lanczos / Golub-Kahn / Gram-Schmidt / matvec: bit-exact equality with
 *     the reference's own sources compiled into oracle/_ref on seeded inputs
 *     (fixtures), plus invariants (A = cI stops after one step, Gauss
 *     quadrature exact for x and x^2, GK on A == Lanczos on A^T A).
 *   - erf_inv: scipy.special.erfinv and the reference sources via oracle/_ref.
 */

#include <stdint.h>
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* CSR operand: int64 row pointers, int32 column indices, values in the
 * precision of the entry point that uses it. */
typedef struct
{
    int64_t nrows;
    int64_t ncols;
    const int64_t* indptr;
    const int32_t* indices;
    const void* data;
} orc_csr;

enum { ORC_OP_MATRIX = 0, ORC_OP_AFFINE_I = 1, ORC_OP_AFFINE_B = 2 };

typedef struct
{
    int32_t kind;   /* ORC_OP_* */
    double t;       /* parameter of A + tB (ignored for a plain matrix) */
    orc_csr A;
    orc_csr B;      /* only for ORC_OP_AFFINE_B */
} orc_op;

static int orc_warn_count = 0;

int orc_warnings(void) { return orc_warn_count; }

/* ---------------- RNG ---------------- */

/* SplitMix64::next, split_mix_64.cpp:58-64 (published SplitMix64 constants) */
static uint64_t splitmix_next(uint64_t* state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

/* Xoshiro256StarStar::next, xoshiro_256_star_star.cpp:73-91 */
static uint64_t xoshiro_next(uint64_t* s)
{
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

/* Xoshiro256StarStar::jump, xoshiro_256_star_star.cpp:100-133: advance by 2^128
 * steps with the published jump polynomial. */
static void xoshiro_jump(uint64_t* s)
{
    static const uint64_t J[4] = {
        0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
        0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
    uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int i = 0; i < 4; ++i)
    {
        for (int b = 0; b < 64; ++b)
        {
            if (J[i] & ((uint64_t) 1 << b))
            {
                a0 ^= s[0]; a1 ^= s[1]; a2 ^= s[2]; a3 ^= s[3];
            }
            xoshiro_next(s);
        }
    }
    s[0] = a0; s[1] = a1; s[2] = a2; s[3] = a3;
}

/* State of the generator of `thread_id`: SplitMix64 seeded with
 * (seed << 32) | seed (split_mix_64.cpp:28-50), four draws, then jump()
 * thread_id + 1 times (random_number_generator.cpp:126-136). A negative seed
 * uses the clock, like the reference. */
void orc_thread_state(int64_t seed, int32_t thread_id, uint64_t* out)
{
    uint64_t sd;
    if (seed >= 0) sd = (uint64_t) seed;
    else
    {
        struct timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        sd = (uint64_t) ts.tv_sec * 1000000000ULL + (uint64_t) ts.tv_nsec;
    }
    uint64_t st = (sd << 32) | sd;
    for (int i = 0; i < 4; ++i) out[i] = splitmix_next(&st);
    for (int j = 0; j < thread_id + 1; ++j) xoshiro_jump(out);
}

void orc_splitmix(uint64_t state, int32_t count, uint64_t* out)
{
    for (int32_t i = 0; i < count; ++i) out[i] = splitmix_next(&state);
}

void orc_xoshiro_words(uint64_t* state, int64_t count, uint64_t* out)
{
    for (int64_t i = 0; i < count; ++i)
    {
        uint64_t w = xoshiro_next(state);
        if (out) out[i] = w;
    }
}

void orc_xoshiro_jump(uint64_t* state) { xoshiro_jump(state); }

/* erf_inv, special_functions.cpp:63-104: rational initial guess (published
 * coefficients of the approximation the reference cites) refined by two Newton
 * steps on erf. */
static double sign_d(double x) { return (x > 0) - (x < 0); }

double orc_erf_inv(double x)
{
    double r;
    if ((x == 1.0) || (x == -1.0)) return sign_d(x) * INFINITY;
    double a[4] = {0.886226899, -1.645349621, 0.914624893, -0.140543331};
    double b[5] = {1.0, -2.118377725, 1.442710462, -0.329097515, 0.012229801};
    double c[4] = {-1.970840454, -1.62490649, 3.429567803, 1.641345311};
    double d[3] = {1.0, 3.543889200, 1.637067800};
    double z = sign_d(x) * x;
    if (z <= 0.7)
    {
        double x2 = z * z;
        r = z * (((a[3] * x2 + a[2]) * x2 + a[1]) * x2 + a[0]);
        r /= (((b[4] * x2 + b[3]) * x2 + b[2]) * x2 + b[1]) * x2 + b[0];
    }
    else
    {
        double y = sqrt(-log((1.0 - z) / 2.0));
        r = (((c[3] * y + c[2]) * y + c[1]) * y + c[0]);
        r /= ((d[2] * y + d[1]) * y + d[0]);
    }
    r = r * sign_d(x);
    z = z * sign_d(x);
    r -= (erf(r) - z) / ((2.0 / sqrt(M_PI)) * exp(-r * r));
    r -= (erf(r) - z) / ((2.0 / sqrt(M_PI)) * exp(-r * r));
    return r;
}

/* ---------------- floating-point body, per precision ---------------- */

#define REAL double
#define SFX f64
#define SQRT sqrt
#define EPS DBL_EPSILON
#include "slq_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef EPS

#define REAL float
#define SFX f32
#define SQRT sqrtf
#define EPS FLT_EPSILON
#include "slq_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef EPS
