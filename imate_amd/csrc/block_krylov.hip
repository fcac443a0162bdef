// Block Lanczos form of the estimator: the 64 probes of a block share one
// Krylov basis, so the full re-orthogonalisation against the stored basis
// is a matrix product with a shared operand, run on the matrix cores.
//
// imate runs one Lanczos recurrence per probe (c_lanczos_tridiagonalization.
// cpp:117-235); with full re-orthogonalisation (c_orthogonalization.cpp:
// 125-224) each probe projects against its own basis, which on a block of
// probes is a batched matrix-vector product with nothing to share. This form
// runs block Lanczos instead (an opt-in: its samples differ from imate's,
// since each probe's quadrature uses the whole block's Krylov space):
//
//   Z = Q_0 R_0                      (probe block, Cholesky QR twice)
//   for j = 0 .. m-1:
//     W = (A + tB) Q_j               (the tiled SpMM)
//     H = Q_{0..j}^T W; A_j = sym(H_j); W -= Q_{0..j} H   (pass 1)
//     H = Q_{0..j}^T W; W -= Q_{0..j} H                    (pass 2)
//     W = Q_{j+1} B_j                (Cholesky QR twice)
//   T = block tridiagonal (A_j on the diagonal, B_j below), 64 m x 64 m
//   T = Y diag(theta) Y^T            (rocSOLVER, loaded on first use)
//   sample_p = sum_k f(theta_k^p) (sum_l R_0[l][p] Y[l][k])^2
//
// The eigensolve is a chain of small latency-bound launches (the dense
// tridiagonalisation takes 92 of its 133 ms at 64 m = 3840), so the
// recurrences of a group of blocks keep their T in slots and one batched
// rocsolver_dsyevd_strided_batched call solves them all (80 ms per matrix
// at four, 71 at eight).
//
// which is the block Gauss quadrature of z_p^T f(A) z_p (exact for
// polynomials of degree < 2m, like the scalar rule). The first pass of the
// projection also removes the Q_j A_j + Q_{j-1} B_{j-1}^T terms of the
// three-term block recurrence.
//
// Matrix-core kernels (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32):
//   bk_tn_kernel: partial 64 x 64 blocks of Q_b^T W, one wavefront per
//     (row chunk, basis block) holding all 16 output tiles, the chunk's
//     rows of W shared through LDS by four blocks; partial blocks are added
//     in a fixed order by bk_reduce_kernel (deterministic);
//   bk_nn_kernel: W -= sum_b Q_b H_b (or Y = X H), one wave per 32 rows,
//     H_b staged through LDS, the contraction index permuted so that each
//     lane reads 16 consecutive values of its basis row.
// Every row of the basis block is 64 contiguous values (the probe block
// layout of the SpMM), so both products stream the stored basis once per
// pass: 16 flops per byte in FP64, 32 in FP32, above the HBM ridge of the
// matrix cores.

#include <dlfcn.h>

#include <cmath>
#include <mutex>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ql.hpp"
#include "slq_internal.hpp"

namespace slq {

namespace {

constexpr int kB = kBlock;            // probes per block

// Consecutive 256-byte aligned pieces of a workspace.
struct Carve {
    char* p;
    size_t left;
    template <typename U>
    U* take(size_t n)
    {
        const size_t bytes = (n * sizeof(U) + 255) & ~(size_t) 255;
        if (bytes > left) return nullptr;
        U* r = (U*) p;
        p += bytes;
        left -= bytes;
        return r;
    }
};
constexpr int kTile = kB * kB;        // values of one 64 x 64 block
constexpr int kBkThreads = 256;       // four wavefronts

template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef double V4 __attribute__((ext_vector_type(4)));
    // D element m of lane l: row l / 16 + 4 m, column l % 16
    static __device__ __forceinline__ int row(int kq, int m) { return kq + 4 * m; }
    static __device__ __forceinline__ V4 op(double a, double b, V4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mma<float> {
    typedef float V4 __attribute__((ext_vector_type(4)));
    // D element m of lane l: row 4 (l / 16) + m, column l % 16
    static __device__ __forceinline__ int row(int kq, int m) { return 4 * kq + m; }
    static __device__ __forceinline__ V4 op(float a, float b, V4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// P[c][b] = sum over rows i of chunk c of Q_b[i][p] W[i][q] (64 x 64,
// row-major in p), Q_b = Q + b * bstride. Grid (chunks, ceil(nkb / 4)),
// 256 threads: the four waves of a workgroup take basis blocks b = 4 y + w
// of the same chunk, so the chunk's rows of W are staged once through LDS
// (16 rows at a time, double buffered, the next rows loaded during this
// step's MFMAs) and serve all four; each wave computes the whole 64 x 64
// block (16 accumulator tiles) and reads its rows of Q_b once (FP32: the
// next 16 rows load during this step's MFMAs, 0.46 -> 0.49 of the matrix
// peak; the FP64 prefetch registers would halve the waves per SIMD, which
// measured slower). MFMA
// operands per four rows: A[r][k] = Q_b[i0 + k][16 pb + r],
// B[k][c] = W[i0 + k][16 qb + c], with r = c = lane % 16 and k = lane / 16.
// Chunks are a multiple of 16 rows.
template <typename T>
__global__ __launch_bounds__(kBkThreads) void bk_tn_kernel(const T* __restrict__ Q,
                                                           int64_t bstride,
                                                           const T* __restrict__ W,
                                                           int64_t n, int64_t chunk_rows,
                                                           int nkb, double* __restrict__ P)
{
    using V4 = typename Mma<T>::V4;
    constexpr int kRows = 16;
    constexpr int kPer = kRows * kB / kBkThreads;   // staged values per thread
    __shared__ T Ws[2][kRows * kB];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int c = blockIdx.x;
    const int b = blockIdx.y * 4 + wave;
    const bool active = b < nkb;
    const int r = lane & 15, kq = lane >> 4;
    const T* __restrict__ Qb = Q + (int64_t) (active ? b : 0) * bstride;
    const int64_t r0 = (int64_t) c * chunk_rows;
    const int64_t r1 = r0 + chunk_rows < n ? r0 + chunk_rows : n;
    V4 acc[4][4];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
#pragma unroll
        for (int qb = 0; qb < 4; ++qb) acc[pb][qb] = V4{T(0), T(0), T(0), T(0)};
    auto load_w = [&](int64_t i0, T (&v)[kPer]) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const int e = k * kBkThreads + threadIdx.x;
            const int64_t i = i0 + e / kB;
            v[k] = i < r1 ? W[i * kB + e % kB] : T(0);
        }
    };
    constexpr bool kPrefQ = sizeof(T) == 4;
    auto load_q = [&](int64_t i0, T (&v)[4][4]) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
        {
            const int64_t i = i0 + 4 * g + kq;
            const bool ok = active && i < r1;
#pragma unroll
            for (int x = 0; x < 4; ++x) v[g][x] = ok ? Qb[i * kB + x * 16 + r] : T(0);
        }
    };
    T a[4][4];
    if (kPrefQ) load_q(r0, a);
    T wv[kPer];
    load_w(r0, wv);
#pragma unroll
    for (int k = 0; k < kPer; ++k) Ws[0][k * kBkThreads + threadIdx.x] = wv[k];
    __syncthreads();
    int buf = 0;
    for (int64_t i0 = r0; i0 < r1; i0 += kRows)
    {
        const bool more = i0 + kRows < r1;
        if (more) load_w(i0 + kRows, wv);
        T an[4][4];
        if (kPrefQ && more) load_q(i0 + kRows, an);
        if (!kPrefQ) load_q(i0, a);
        if (active)
        {
            const T* __restrict__ ws = Ws[buf];
#pragma unroll
            for (int g = 0; g < 4; ++g)
            {
                T w[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) w[x] = ws[(4 * g + kq) * kB + x * 16 + r];
#pragma unroll
                for (int pb = 0; pb < 4; ++pb)
#pragma unroll
                    for (int qb = 0; qb < 4; ++qb)
                        acc[pb][qb] = Mma<T>::op(a[g][pb], w[qb], acc[pb][qb]);
            }
        }
        if (more)
        {
#pragma unroll
            for (int k = 0; k < kPer; ++k) Ws[buf ^ 1][k * kBkThreads + threadIdx.x] = wv[k];
            if (kPrefQ)
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int x = 0; x < 4; ++x) a[g][x] = an[g][x];
        }
        buf ^= 1;
        __syncthreads();
    }
    if (!active) return;
    double* __restrict__ Pc = P + ((int64_t) c * nkb + b) * kTile;
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
#pragma unroll
        for (int qb = 0; qb < 4; ++qb)
#pragma unroll
            for (int m = 0; m < 4; ++m)
                Pc[(pb * 16 + Mma<T>::row(kq, m)) * kB + qb * 16 + r] = (double) acc[pb][qb][m];
}

// H[e] = sum_c P[c][e] over the chunks, e < len (= nkb * 4096). A
// workgroup takes 64 consecutive elements; its four wavefronts add the
// chunks c = w, w + 4, ... (four loads in flight per lane), and wave 0 adds
// the four sums in order, so the result is fixed for a given chunk count.
__global__ __launch_bounds__(256) void bk_reduce_kernel(const double* __restrict__ P, int nch,
                                                        int64_t len, double* __restrict__ H)
{
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int64_t e = (int64_t) blockIdx.x * 64 + lane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (e < len)
    {
        int c = w;
        for (; c + 12 < nch; c += 16)
        {
            s0 += P[(int64_t) c * len + e];
            s1 += P[(int64_t) (c + 4) * len + e];
            s2 += P[(int64_t) (c + 8) * len + e];
            s3 += P[(int64_t) (c + 12) * len + e];
        }
        for (; c < nch; c += 4) s0 += P[(int64_t) c * len + e];
    }
    part[w][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (w == 0 && e < len) H[e] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// Y = X H (SUB = false) or Y -= X H (SUB = true) with X = [X_0 .. X_{nkb-1}]
// (X_b = X + b * bstride, n x 64 each) and H = [H_0; ..; H_{nkb-1}] (64 x 64
// each, row-major). Each wave computes NnTiles 16-row tiles of Y (each value
// of H read from LDS serves them all); H_b is staged in LDS per workgroup. In
// k-step s of block b the MFMA contracts over p = 16 k + s (k = lane / 16),
// so lane (r, k) reads X_b[row r][16 k .. 16 k + 15]: 128 B (FP64)
// contiguous per lane. X and Y must not overlap.
// 16-row tiles of Y per wave in bk_nn_kernel: two (each value of H read
// from LDS and each staging of H_b serve twice the rows; FP64 0.46 -> 0.51
// of the matrix peak against one tile, four measured slower for FP32).
template <typename T>
struct NnTiles {
    static constexpr int value = 2;
};

template <typename T, bool SUB>
__global__ __launch_bounds__(kBkThreads) void bk_nn_kernel(const T* __restrict__ X,
                                                           int64_t bstride, int nkb,
                                                           const double* __restrict__ H,
                                                           T* __restrict__ Y, int64_t n)
{
    using V4 = typename Mma<T>::V4;
    constexpr int kPer = kTile / kBkThreads;   // values of H_b per thread
    constexpr int U = NnTiles<T>::value;
    __shared__ T Hs[sizeof(T) == 4 ? 2 : 1][kTile];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int64_t tile0 = ((int64_t) blockIdx.x * 4 + wave) * U;
    V4 acc[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int qb = 0; qb < 4; ++qb) acc[u][qb] = V4{T(0), T(0), T(0), T(0)};
    auto load_a = [&](int b, T (&a)[U][16]) {
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int64_t arow = (tile0 + u) * 16 + r;
            const bool ok = arow < n;
            const T* __restrict__ src = X + (int64_t) b * bstride + arow * kB + 16 * kq;
#pragma unroll
            for (int s = 0; s < 16; ++s) a[u][s] = ok ? src[s] : T(0);
        }
    };
    auto load_h = [&](int b, T (&h)[kPer]) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            h[k] = (T) H[(int64_t) b * kTile + k * kBkThreads + threadIdx.x];
    };
    T a[U][16], hv[kPer];
    // FP32: block b + 1's operands load during block b's MFMAs, with H
    // double-buffered in LDS; FP64: one buffer, loads then MFMAs (the
    // prefetch registers would halve the waves per SIMD, which measured
    // slower)
    constexpr bool kAhead = sizeof(T) == 4;
    if (kAhead)
    {
        load_h(0, hv);
#pragma unroll
        for (int k = 0; k < kPer; ++k) Hs[0][k * kBkThreads + threadIdx.x] = hv[k];
        load_a(0, a);
        __syncthreads();
    }
    for (int b = 0; b < nkb; ++b)
    {
        T an[U][16];
        const bool more = kAhead && b + 1 < nkb;
        if (more)
        {
            load_h(b + 1, hv);
            load_a(b + 1, an);
        }
        if (!kAhead)
        {
            load_h(b, hv);
            load_a(b, a);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kPer; ++k) Hs[0][k * kBkThreads + threadIdx.x] = hv[k];
            __syncthreads();
        }
        const T* __restrict__ hs = Hs[kAhead ? (b & 1) : 0];
#pragma unroll
        for (int s = 0; s < 16; ++s)
        {
            const T* __restrict__ hrow = hs + (16 * kq + s) * kB + r;
#pragma unroll
            for (int qb = 0; qb < 4; ++qb)
            {
                const T h = hrow[qb * 16];
#pragma unroll
                for (int u = 0; u < U; ++u) acc[u][qb] = Mma<T>::op(a[u][s], h, acc[u][qb]);
            }
        }
        if (more)
        {
#pragma unroll
            for (int k = 0; k < kPer; ++k)
                Hs[kAhead ? ((b + 1) & 1) : 0][k * kBkThreads + threadIdx.x] = hv[k];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int s = 0; s < 16; ++s) a[u][s] = an[u][s];
        }
        if (kAhead) __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            const int64_t orow = (tile0 + u) * 16 + Mma<T>::row(kq, m);
            if (orow >= n) continue;
#pragma unroll
            for (int qb = 0; qb < 4; ++qb)
            {
                const int64_t o = orow * kB + qb * 16 + r;
                if constexpr (SUB) Y[o] = Y[o] - acc[u][qb][m];
                else Y[o] = acc[u][qb][m];
            }
        }
}

// Cholesky factor of G (64 x 64, its upper triangle read): G = R^T R with R
// upper triangular. Writes Rinv = R^{-1} and Rout = R Rprev (R when Rprev
// is null), both row-major; a pivot that is not positive (or not finite)
// sets *flag. 256 threads: thread (w, l) holds rows 16 w .. 16 w + 15 of
// column l in registers. At step k the wavefront holding row k publishes
// it through a double-buffered LDS row (one barrier per step: a buffer is
// rewritten two steps later, after the next barrier), every thread scales
// it by 1 / R[k][k] and updates its rows below k; the rows of R are kept in
// LDS for the inverse (one wavefront, its column of x in registers) and
// the product (all four, Rprev staged in LDS). Rows of a column below its
// diagonal are updated too and never read, which keeps the loops free of
// lane masks. (A one-wavefront form with readlane broadcasts, 42k
// instructions, took 110 us.)
__global__ __launch_bounds__(256) void bk_chol_kernel(const double* __restrict__ G,
                                                      const double* __restrict__ Rprev,
                                                      double* __restrict__ Rinv,
                                                      double* __restrict__ Rout,
                                                      int* __restrict__ flag)
{
    __shared__ double raw[2][kB];
    __shared__ double Rs[kB][kB];
    __shared__ double Ps[kB][kB];
    const int l = threadIdx.x & (kB - 1);
    const int w = threadIdx.x >> 6;
    double c[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) c[r] = G[(16 * w + r) * kB + l];
    if (Rprev)
    {
#pragma unroll
        for (int r = 0; r < 16; ++r) Ps[16 * w + r][l] = Rprev[(16 * w + r) * kB + l];
    }
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kB; ++k)
    {
        const int b = k & 1;
        if (w == k / 16) raw[b][l] = c[k % 16];
        __syncthreads();
        double d = raw[b][k];
        if (!(d > 0.0) || !isfinite(d))
        {
            bad = true;
            d = 1.0;
        }
        const double rkk = sqrt(d);
        const double inv = 1.0 / rkk;
        const double rkl = l > k ? raw[b][l] * inv : (l == k ? rkk : 0.0);
        if (w == 0) Rs[k][l] = rkl;
        // rows i = 16 w + r: A[i][l] -= R[k][i] R[k][l]
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] -= (raw[b][16 * w + r] * inv) * rkl;
    }
    if (bad && threadIdx.x == 0) atomicOr(flag, 1);
    __syncthreads();
    if (w == 0)
    {
        // column l of R^{-1}, bottom row first
        double x[kB];
#pragma unroll
        for (int i = kB - 1; i >= 0; --i)
        {
            // four partial sums (k mod 4) shorten the dependent chain
            double acc[4] = {(i == l) ? 1.0 : 0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = i + 1; k < kB; ++k) acc[k & 3] -= Rs[i][k] * x[k];
            x[i] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) / Rs[i][i];
        }
#pragma unroll
        for (int i = 0; i < kB; ++i) Rinv[i * kB + l] = x[i];
    }
    // rows 16 w .. 16 w + 15 of R Rprev (Rprev upper triangular) or of R
    for (int r = 0; r < 16; ++r)
    {
        const int i = 16 * w + r;
        double acc;
        if (Rprev)
        {
            double a0 = 0.0, a1 = 0.0;
            int k = i;
            for (; k + 1 < kB; k += 2)
            {
                a0 += Rs[i][k] * Ps[k][l];
                a1 += Rs[i][k + 1] * Ps[k + 1][l];
            }
            if (k < kB) a0 += Rs[i][k] * Ps[k][l];
            acc = a0 + a1;
        }
        else
            acc = Rs[i][l];
        Rout[i * kB + l] = acc;
    }
}

// A = (H + H^T) / 2 of one 64 x 64 block.
__global__ __launch_bounds__(256) void bk_sym_kernel(const double* __restrict__ H,
                                                     double* __restrict__ A)
{
    for (int e = threadIdx.x; e < kTile; e += 256)
    {
        const int i = e / kB, j = e % kB;
        A[e] = 0.5 * (H[i * kB + j] + H[j * kB + i]);
    }
}

// Column-major block tridiagonal T (N = 64 m): A_j on the diagonal, B_j
// below it and B_j^T above. AB holds A_j at 2 j and B_j at 2 j + 1.
__global__ __launch_bounds__(256) void bk_assemble_kernel(const double* __restrict__ AB,
                                                          int m, double* __restrict__ T)
{
    const int64_t N = (int64_t) m * kB;
    const int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (e >= N * N) return;
    const int64_t col = e / N, row = e % N;
    const int64_t bi = row / kB, bj = col / kB;
    const int ri = (int) (row % kB), cj = (int) (col % kB);
    double v = 0.0;
    if (bi == bj) v = AB[(2 * bi) * kTile + ri * kB + cj];
    else if (bi == bj + 1) v = AB[(2 * bj + 1) * kTile + ri * kB + cj];
    else if (bj == bi + 1) v = AB[(2 * bi + 1) * kTile + cj * kB + ri];
    T[e] = v;
}

// U[p][k] = sum_{l <= p} R0[l][p] Y[l][k] (Y column-major, N x N).
__global__ __launch_bounds__(256) void bk_weights_kernel(const double* __restrict__ R0,
                                                         const double* __restrict__ Y,
                                                         int64_t N, double* __restrict__ U)
{
    const int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (e >= N * kB) return;
    const int p = (int) (e / N);
    const int64_t k = e % N;
    double s = 0.0;
    for (int l = 0; l <= p; ++l) s += R0[l * kB + p] * Y[k * N + l];
    U[e] = s;
}

struct BkQuadArgs {
    int64_t N;
    const double* theta;      // [N]
    const double* U;          // [64][N]
    const int* info;
    int func;
    double fp[4];
    double exponent;
    const double* params;     // device, or nullptr
    int q;
    int shift;
    int j_out;
    int q_stride;
    double* samples;
};

// sample_p = sum_k f((theta_k - t_0 + t_j)^p) U[p][k]^2: one workgroup per
// probe, a fixed-order tree over the threads.
__global__ __launch_bounds__(256) void bk_quadrature_kernel(BkQuadArgs a)
{
    __shared__ double red[256];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int jn = a.shift ? a.q : 1;
    const double t0 = a.shift ? a.params[0] : 0.0;
    for (int j = 0; j < jn; ++j)
    {
        const double tj = a.shift ? a.params[j] : 0.0;
        double s = 0.0;
        for (int64_t k = t; k < a.N; k += 256)
        {
            double th = a.theta[k];
            if (a.shift) th = (th - t0) + tj;
            const double u = a.U[(int64_t) p * a.N + k];
            s += u * u * ffunc(a.func, a.fp, fpow(th, a.exponent));
        }
        red[t] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1)
        {
            if (t < w) red[t] += red[t + w];
            __syncthreads();
        }
        if (t == 0)
        {
            const int col = a.shift ? j : a.j_out;
            a.samples[(int64_t) p * a.q_stride + col] = *a.info == 0 ? red[0] : NAN;
        }
        __syncthreads();
    }
}

__global__ void bk_zero_int_kernel(int* p, int n)
{
    if ((int) threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void bk_zero_kernel(double* p, int64_t n)
{
    const int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x;
    if (e < n) p[e] = 0.0;
}

// ---------------------------------------------------------------- solver

// rocSOLVER / rocBLAS, opened on first use so that the library itself does
// not load them.
struct Solver {
    bool tried = false;
    bool ok = false;
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocsolver_dsyevd_strided_batched) dsyevd = nullptr;
};

std::mutex g_solver_mu;

Solver& solver()
{
    static Solver s;
    std::lock_guard<std::mutex> lock(g_solver_mu);
    if (s.tried) return s;
    s.tried = true;
    void* blas = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
    if (!blas) blas = dlopen("librocblas.so", RTLD_NOW | RTLD_LOCAL);
    void* sol = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!sol) sol = dlopen("librocsolver.so", RTLD_NOW | RTLD_LOCAL);
    if (!blas || !sol) return s;
    s.create = (decltype(s.create)) dlsym(blas, "rocblas_create_handle");
    s.destroy = (decltype(s.destroy)) dlsym(blas, "rocblas_destroy_handle");
    s.set_stream = (decltype(s.set_stream)) dlsym(blas, "rocblas_set_stream");
    s.dsyevd = (decltype(s.dsyevd)) dlsym(sol, "rocsolver_dsyevd_strided_batched");
    s.ok = s.create && s.destroy && s.set_stream && s.dsyevd;
    return s;
}

// ---------------------------------------------------------------- driver

struct BkBuffers {
    double* P = nullptr;      // partial blocks [max_chunks * m][4096]
    double* H = nullptr;      // [m][4096]
    double* AB = nullptr;     // [2 m][4096]
    double* R0 = nullptr;     // R_0 of the recurrence's slot
    double* R1 = nullptr;
    double* Rinv = nullptr;
    double* G = nullptr;
    int* flag = nullptr;      // [0] rank loss, [1 + s] eigensolver info of slot s
};

// The block tridiagonals of a group, one slot each, laid out for the
// batched eigensolve (T, D and E strided by slot), with each slot's R_0.
struct BkSlots {
    int count = 0;
    double* R0 = nullptr;     // [count][4096]
    double* T = nullptr;      // [count][N][N]
    double* D = nullptr;      // [count][N]
    double* E = nullptr;      // [count][N]
    double* U = nullptr;      // [64][N], one slot at a time
};

// Row chunks of a Q^T W product over nkb basis blocks (one workgroup per
// chunk and four blocks): about 512 workgroups (two wavefronts per SIMD,
// the kernel's occupancy), at least 32 chunks, chunks of at least 16 rows.
// Each chunk writes a partial block that bk_reduce_kernel reads back, so
// fewer chunks move fewer bytes: against about 1024 workgroups, GEMM time
// per call 415 -> 407 ms FP64 and 210 -> 205 ms FP32; about 256 were
// slower (462 and 231 ms).
int tn_chunks(int64_t n, int nkb)
{
    const int groups = (nkb + 3) / 4;
    int c = 512 / groups;
    if (c < 32) c = 32;
    const int64_t most = (n + 15) / 16;
    if (c > most) c = (int) most;
    return c < 1 ? 1 : c;
}

int64_t max_partial_blocks(int64_t n, int m)
{
    int64_t best = 0;
    for (int k = 1; k <= m; ++k)
    {
        const int64_t v = (int64_t) tn_chunks(n, k) * k;
        if (v > best) best = v;
    }
    return best;
}

// H = [Q_0 .. Q_{nkb-1}]^T W
template <typename T>
int tn(slq_op* op, const T* Q, int64_t bstride, int nkb, const T* W, int64_t n,
       const BkBuffers& bb, double* H)
{
    const hipStream_t s = op->stream;
    const int nch0 = tn_chunks(n, nkb);
    int64_t chunk_rows = (n + nch0 - 1) / nch0;
    chunk_rows = (chunk_rows + 15) / 16 * 16;
    const int nch = (int) ((n + chunk_rows - 1) / chunk_rows);
    hipEvent_t ev;
    prof_begin(op, &ev);
    hipLaunchKernelGGL(bk_tn_kernel<T>, dim3(nch, (nkb + 3) / 4), dim3(kBkThreads), 0, s, Q,
                       bstride, W, n, chunk_rows, nkb, bb.P);
    prof_end_gemm(op, ev, 2.0 * (double) n * kTile * nkb);
    const int64_t len = (int64_t) nkb * kTile;
    hipLaunchKernelGGL(bk_reduce_kernel, dim3((unsigned) ((len + 63) / 64)), dim3(256), 0, s,
                       (const double*) bb.P, nch, len, H);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

template <typename T, bool SUB>
int nn(slq_op* op, const T* X, int64_t bstride, int nkb, const double* H, T* Y, int64_t n)
{
    const hipStream_t s = op->stream;
    hipEvent_t ev;
    prof_begin(op, &ev);
    const int64_t tiles = (n + 15) / 16;
    const int64_t per_wg = 4 * NnTiles<T>::value;
    hipLaunchKernelGGL((bk_nn_kernel<T, SUB>), dim3((unsigned) ((tiles + per_wg - 1) / per_wg)),
                       dim3(kBkThreads), 0, s, X, bstride, nkb, H, Y, n);
    prof_end_gemm(op, ev, 2.0 * (double) n * kTile * nkb);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

// Y R = X (Cholesky QR twice): Y has orthonormal columns; Rout = R (upper
// triangular). Tmp is an n x 64 buffer; X, Tmp and Y are distinct.
template <typename T>
int cholqr2(slq_op* op, const T* X, T* Tmp, T* Y, int64_t n, const BkBuffers& bb,
            double* Rout)
{
    const hipStream_t s = op->stream;
    SLQ_TRY(tn<T>(op, X, 0, 1, X, n, bb, bb.G));
    hipLaunchKernelGGL(bk_chol_kernel, dim3(1), dim3(256), 0, s, (const double*) bb.G,
                       (const double*) nullptr, bb.Rinv, bb.R1, bb.flag);
    SLQ_TRY((nn<T, false>(op, X, 0, 1, bb.Rinv, Tmp, n)));
    SLQ_TRY(tn<T>(op, Tmp, 0, 1, Tmp, n, bb, bb.G));
    hipLaunchKernelGGL(bk_chol_kernel, dim3(1), dim3(256), 0, s, (const double*) bb.G,
                       (const double*) bb.R1, bb.Rinv, Rout, bb.flag);
    SLQ_TRY((nn<T, false>(op, Tmp, 0, 1, bb.Rinv, Y, n)));
    return SLQ_OK;
}

template <typename T>
int block_lanczos(const BlockKrylovArgs& a, Carve& bump)
{
    slq_op* op = a.op;
    const int64_t n = op->A->csr.nrows;
    const int m = a.m;
    const int64_t N = (int64_t) m * kB;
    const int64_t NN = N * N;
    const hipStream_t s = op->stream;
    const size_t blk = (size_t) n * kB;
    T* Q = bump.take<T>(blk * m);
    T* W = bump.take<T>(blk);
    T* W2 = bump.take<T>(blk);
    BkBuffers bb;
    bb.P = bump.take<double>((size_t) max_partial_blocks(n, m) * kTile);
    bb.H = bump.take<double>((size_t) m * kTile);
    bb.AB = bump.take<double>((size_t) 2 * m * kTile);
    bb.R1 = bump.take<double>(kTile);
    bb.Rinv = bump.take<double>(kTile);
    bb.G = bump.take<double>(kTile);
    bb.flag = bump.take<int>(64);
    BkSlots sl;
    sl.count = block_krylov_slots(m);
    sl.R0 = bump.take<double>((size_t) sl.count * kTile);
    sl.T = bump.take<double>((size_t) sl.count * NN);
    sl.D = bump.take<double>((size_t) sl.count * N);
    sl.E = bump.take<double>((size_t) sl.count * N);
    sl.U = bump.take<double>((size_t) kB * N);
    if (!Q || !W || !W2 || !bb.P || !bb.H || !bb.AB || !bb.R1 || !bb.Rinv || !bb.G ||
        !bb.flag || !sl.R0 || !sl.T || !sl.D || !sl.E || !sl.U)
    {
        set_error("internal: block Krylov workspace too small");
        return SLQ_E_OOM;
    }
    if (a.flags_out) *a.flags_out = bb.flag;
    void* spmm_scratch = bump.p;
    const size_t spmm_bytes = bump.left;
    if (a.phase == 0)
    {
        hipLaunchKernelGGL(bk_zero_int_kernel, dim3(1), dim3(64), 0, s, bb.flag, 1);
        SLQ_CHECK(hipGetLastError());
        return SLQ_OK;
    }
    if (a.phase == 1)
    {
        if (a.slot < 0 || a.slot >= sl.count)
        {
            set_error("internal: block Krylov slot %d out of range", a.slot);
            return SLQ_E_ARG;
        }
        bb.R0 = sl.R0 + (size_t) a.slot * kTile;
        // the B_{m-1} slot is never computed (nor read): keep it defined
        {
            const int64_t len = (int64_t) 2 * m * kTile;
            hipLaunchKernelGGL(bk_zero_kernel, dim3((unsigned) ((len + 255) / 256)),
                               dim3(256), 0, s, bb.AB, len);
        }

        SLQ_TRY(cholqr2<T>(op, (const T*) a.X0, W2, Q, n, bb, bb.R0));
        for (int j = 0; j < m; ++j)
        {
            T* Qj = Q + (size_t) j * blk;
            SLQ_TRY(spmm_block(op, a.t, sizeof(T) == 4 ? SLQ_F32 : SLQ_F64, Qj, W, spmm_scratch,
                               spmm_bytes));
            // pass 1: A_j = sym(Q_j^T W), W -= Q H
            SLQ_TRY(tn<T>(op, Q, (int64_t) blk, j + 1, W, n, bb, bb.H));
            hipLaunchKernelGGL(bk_sym_kernel, dim3(1), dim3(256), 0, s,
                               (const double*) (bb.H + (size_t) j * kTile),
                               bb.AB + (size_t) 2 * j * kTile);
            SLQ_TRY((nn<T, true>(op, Q, (int64_t) blk, j + 1, bb.H, W, n)));
            // pass 2
            SLQ_TRY(tn<T>(op, Q, (int64_t) blk, j + 1, W, n, bb, bb.H));
            SLQ_TRY((nn<T, true>(op, Q, (int64_t) blk, j + 1, bb.H, W, n)));
            if (j + 1 < m)
                SLQ_TRY(cholqr2<T>(op, W, W2, Q + (size_t) (j + 1) * blk, n, bb,
                                   bb.AB + (size_t) (2 * j + 1) * kTile));
        }
        // the block tridiagonal, into the slot
        hipLaunchKernelGGL(bk_assemble_kernel, dim3((unsigned) ((NN + 255) / 256)), dim3(256),
                           0, s, (const double*) bb.AB, m, sl.T + (size_t) a.slot * NN);
        SLQ_CHECK(hipGetLastError());
        return SLQ_OK;
    }
    // phase 2: one batched eigensolve of the group's slots, then the
    // quadrature of each
    if (a.nslots < 1 || a.nslots > sl.count)
    {
        set_error("internal: block Krylov batch of %d slots (at most %d)", a.nslots, sl.count);
        return SLQ_E_ARG;
    }
    Solver& sv = solver();
    if (!sv.ok)
    {
        set_error("block Krylov: rocSOLVER (librocsolver.so.0, librocblas.so.5) could not be "
                  "loaded");
        return SLQ_E_HIP;
    }
    if (!op->rocblas)
    {
        rocblas_handle h = nullptr;
        if (sv.create(&h) != rocblas_status_success)
        {
            set_error("block Krylov: rocblas_create_handle failed");
            return SLQ_E_HIP;
        }
        op->rocblas = h;
    }
    rocblas_handle h = (rocblas_handle) op->rocblas;
    if (sv.set_stream(h, s) != rocblas_status_success ||
        sv.dsyevd(h, rocblas_evect_original, rocblas_fill_lower, (rocblas_int) N, sl.T,
                  (rocblas_int) N, (rocblas_stride) NN, sl.D, (rocblas_stride) N, sl.E,
                  (rocblas_stride) N, bb.flag + 1, (rocblas_int) a.nslots) !=
            rocblas_status_success)
    {
        set_error("block Krylov: rocsolver_dsyevd_strided_batched failed");
        return SLQ_E_HIP;
    }
    for (int k = 0; k < a.nslots; ++k)
    {
        hipLaunchKernelGGL(bk_weights_kernel, dim3((unsigned) ((N * kB + 255) / 256)), dim3(256),
                           0, s, (const double*) (sl.R0 + (size_t) k * kTile),
                           (const double*) (sl.T + (size_t) k * NN), N, sl.U);
        BkQuadArgs qa;
        qa.N = N;
        qa.theta = sl.D + (size_t) k * N;
        qa.U = sl.U;
        qa.info = bb.flag + 1 + k;
        qa.func = a.func;
        for (int e = 0; e < 4; ++e) qa.fp[e] = a.func_params ? a.func_params[e] : 0.0;
        qa.exponent = a.exponent;
        qa.params = a.params_dev;
        qa.q = a.q;
        qa.shift = a.shift;
        qa.j_out = a.j_out[k];
        qa.q_stride = a.q_stride;
        qa.samples = a.samples_dev[k];
        hipLaunchKernelGGL(bk_quadrature_kernel, dim3(kB), dim3(256), 0, s, qa);
    }
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

}  // namespace

int block_krylov_slots(int m)
{
    const size_t tbytes = sizeof(double) * (size_t) m * kB * (size_t) m * kB;
    size_t k = ((size_t) 1 << 30) / (tbytes ? tbytes : 1);
    if (k > (size_t) kBkMaxSlots) k = kBkMaxSlots;
    return k < 1 ? 1 : (int) k;
}

size_t block_krylov_scratch_bytes(const slq_op* op, int m, int value_bytes)
{
    auto r = [](size_t b) { return (b + 255) & ~(size_t) 255; };
    const int64_t n = op->A->csr.nrows;
    const int64_t N = (int64_t) m * kB;
    const size_t S = (size_t) block_krylov_slots(m);
    const size_t blk = (size_t) n * kB * value_bytes;
    size_t b = r(blk * m) + 2 * r(blk);
    b += r(sizeof(double) * (size_t) max_partial_blocks(n, m) * kTile);
    b += r(sizeof(double) * (size_t) m * kTile) + r(sizeof(double) * (size_t) 2 * m * kTile);
    b += 3 * r(sizeof(double) * kTile) + r(sizeof(int) * 64);
    b += r(sizeof(double) * S * kTile) + r(sizeof(double) * S * (size_t) N * N) +
         2 * r(sizeof(double) * S * (size_t) N) + r(sizeof(double) * (size_t) kB * N);
    b += spmm_block_scratch_bytes(op);
    return b;
}

int run_block_krylov(const BlockKrylovArgs& a, void* scratch, size_t scratch_bytes)
{
    Carve bump{(char*) scratch, scratch_bytes};
    if (a.value_dtype == SLQ_F32) return block_lanczos<float>(a, bump);
    return block_lanczos<double>(a, bump);
}

void release_solver(slq_op* op)
{
    if (op->rocblas)
    {
        Solver& sv = solver();
        if (sv.ok) (void) sv.destroy((rocblas_handle) op->rocblas);
        op->rocblas = nullptr;
    }
}

}  // namespace slq
