// Block Krylov recurrences on gfx950: CSR SpMM over a 64-probe block with
// fused Lanczos / Golub-Kahn epilogues, deterministic grid reductions and
// classical Gram-Schmidt (twice) re-orthogonalisation.
//
// Reference algorithms (one vector at a time there):
//   Lanczos     imate/_c_trace_estimator/c_lanczos_tridiagonalization.cpp:117-235
//               (GPU twin cu_lanczos_tridiagonalization.cu:119-248)
//   Golub-Kahn  imate/_c_trace_estimator/c_golub_kahn_bidiagonalization.cpp:111-231
//   MGS         imate/_c_trace_estimator/c_orthogonalization.cpp:125-224
//   operators   imate/_c_linear_operator/c_csr_matrix.cpp:137-190,
//               c_csr_affine_matrix_function.cpp (A x + t x, A x + t B x)
//
// Layout: every vector block is [rows][64] row-major, lane p of a wavefront
// owns probe p. A CSR row is processed by one wavefront: the row's column
// indices and values are read once, coalesced, by the lanes, broadcast with
// v_readlane, and each non-zero gathers the contiguous 64-probe row of the
// operand block (512 B in FP64). Rows are handed out in contiguous runs per
// workgroup, with workgroup ids mapped XCD-major so that neighbouring runs
// (which share columns of a spatially ordered matrix) land on the same XCD's
// L2.
//
// Reductions (alpha, beta, Gram-Schmidt coefficients) are per-probe sums over
// all rows: each workgroup writes one partial row, and either the last
// workgroup to finish (atomic ticket) or a follow-up kernel adds the partials
// in a fixed order, so results are bitwise reproducible run to run.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "slq_internal.hpp"

namespace slq {

namespace {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double readlane(double v, int k)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) u, k);
    const uint32_t hi = (uint32_t) __builtin_amdgcn_readlane((int) (uint32_t) (u >> 32), k);
    return __builtin_bit_cast(double, ((uint64_t) hi << 32) | lo);
}

__device__ __forceinline__ float readlane(float v, int k)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}

__device__ __forceinline__ int readlane(int v, int k)
{
    return __builtin_amdgcn_readlane(v, k);
}

// Contiguous row run [r0, r1) of this workgroup. Workgroup ids are dealt
// round-robin to the 8 XCDs; the remapping gives XCD x the runs
// x*G/8 .. (x+1)*G/8 - 1. G is a multiple of 8.
__device__ __forceinline__ void row_run(int64_t nrows, int64_t& r0, int64_t& r1)
{
    const int64_t G = gridDim.x;
    const int64_t bid = blockIdx.x;
    const int64_t run = (bid % kXcds) * (G / kXcds) + bid / kXcds;
    const int64_t per = (nrows + G - 1) / G;
    r0 = run * per;
    r1 = r0 + per < nrows ? r0 + per : nrows;
}

// sum_k a_ik * (s * X[c_k][lane]) over the CSR row [beg, end).
template <typename T>
__device__ __forceinline__ T row_gather(int64_t beg, int64_t end,
                                        const int32_t* __restrict__ indices,
                                        const T* __restrict__ vals,
                                        const T* __restrict__ X, T s, int lane)
{
    T acc = T(0);
    for (int64_t base = beg; base < end; base += kWave)
    {
        const int64_t rem = end - base;
        const int cnt = rem < kWave ? (int) rem : kWave;
        int cl = 0;
        T vl = T(0);
        if (lane < cnt)
        {
            cl = indices[base + lane];
            vl = vals[base + lane];
        }
        int k = 0;
        for (; k + 4 <= cnt; k += 4)
        {
            const int c0 = readlane(cl, k);
            const int c1 = readlane(cl, k + 1);
            const int c2 = readlane(cl, k + 2);
            const int c3 = readlane(cl, k + 3);
            const T x0 = X[(int64_t) c0 * kBlock + lane];
            const T x1 = X[(int64_t) c1 * kBlock + lane];
            const T x2 = X[(int64_t) c2 * kBlock + lane];
            const T x3 = X[(int64_t) c3 * kBlock + lane];
            const T a0 = readlane(vl, k);
            const T a1 = readlane(vl, k + 1);
            const T a2 = readlane(vl, k + 2);
            const T a3 = readlane(vl, k + 3);
            acc = fma(a0, s * x0, acc);
            acc = fma(a1, s * x1, acc);
            acc = fma(a2, s * x2, acc);
            acc = fma(a3, s * x3, acc);
        }
        for (; k < cnt; ++k)
        {
            const int c = readlane(cl, k);
            const T a = readlane(vl, k);
            acc = fma(a, s * X[(int64_t) c * kBlock + lane], acc);
        }
    }
    return acc;
}

// Workgroup sum of one double per lane: partial[blockIdx.x][lane].
__device__ __forceinline__ void wg_partial(double v, double* __restrict__ partial)
{
    __shared__ double red[kWavesPerWG][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    red[wave][lane] = v;
    __syncthreads();
    if (wave == 0)
    {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWavesPerWG; ++w) s += red[w][lane];
        partial[(int64_t) blockIdx.x * kWave + lane] = s;
    }
}

// After wg_partial: returns true in the last workgroup to finish, which then
// holds the per-lane totals (fixed summation order) in `total` of wave 0.
__device__ __forceinline__ bool last_block_total(const double* __restrict__ partial,
                                                 unsigned int* __restrict__ counter,
                                                 double& total)
{
    __shared__ bool is_last;
    __shared__ double red[kWavesPerWG][kWave];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const unsigned int ticket = atomicAdd(counter, 1u);
        is_last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int G = gridDim.x;
    double s0 = 0.0, s1 = 0.0;
    int g = wave;
    for (; g + kWavesPerWG < G; g += 2 * kWavesPerWG)
    {
        s0 += partial[(int64_t) g * kWave + lane];
        s1 += partial[(int64_t) (g + kWavesPerWG) * kWave + lane];
    }
    if (g < G) s0 += partial[(int64_t) g * kWave + lane];
    red[wave][lane] = s0 + s1;
    __syncthreads();
    total = 0.0;
    if (wave == 0)
    {
#pragma unroll
        for (int w = 0; w < kWavesPerWG; ++w) total += red[w][lane];
        if (lane == 0) *counter = 0u;
    }
    return true;
}

// beta_j = sqrt(total); breakdown test of the reference (beta < tol sqrt(n))
// marks the probe finished with size j + 1; the scale of the next basis
// vector is 1 / beta (0 for finished probes, which then stay zero).
__device__ __forceinline__ void finish_norm(double total, int lane, int j,
                                            double tol_sqrt_n,
                                            double* __restrict__ beta_out,
                                            double* __restrict__ scale_out,
                                            int32_t* __restrict__ sizes)
{
    const double b = sqrt(total);
    beta_out[lane] = b;
    int32_t sz = sizes[lane];
    if (sz == 0 && b < tol_sqrt_n)
    {
        sz = j + 1;
        sizes[lane] = sz;
    }
    scale_out[lane] = (sz == 0) ? 1.0 / b : 0.0;
}

// ---------------------------------------------------------------- kernels

template <typename T>
struct SpmmArgs {
    // A (rows of the output) and optionally B of A + tB
    const int64_t* indptr;
    const int32_t* indices;
    const T* vals;
    const int64_t* b_indptr;
    const int32_t* b_indices;
    const T* b_vals;
    int64_t nrows;
    int64_t shift_len;      // rows that get + t x[i] (identity shift)
    int shift_mode;         // 0 none, 1 + t x (B = I), 2 + t B x
    double t;
    // gathered operand and its per-probe scale
    const T* X;
    const double* xscale;   // nullptr -> 1
    // optional: Vout[i] = xscale * X[i] (square operators)
    T* Vout;
    // optional epilogue subtraction: y -= sub_coef[p] * (sub_scale[p] * Sub[i])
    const T* Sub;
    const double* sub_coef;
    const double* sub_scale;
    T* Y;
    // reduction: 0 none, 1 sum v*y -> red_out (alpha), 2 sum y*y -> norm
    int red_mode;
    double* partial;
    unsigned int* counter;
    double* red_out;        // alpha[j] or norm output
    // norm finishing (red_mode 2)
    int finish_sqrt;        // 1: red_out = sqrt, scale_out = 1/sqrt (+ breakdown)
    double* scale_out;
    int32_t* sizes;
    int j;
    double tol_sqrt_n;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void spmm_kernel(SpmmArgs<T> a)
{
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int64_t r0, r1;
    row_run(a.nrows, r0, r1);
    const T s = a.xscale ? (T) a.xscale[lane] : T(1);
    const T tt = (T) a.t;
    T sub_c = T(0), sub_s = T(0);
    if (a.Sub)
    {
        sub_c = (T) a.sub_coef[lane];
        sub_s = (T) a.sub_scale[lane];
    }
    double red = 0.0;
    for (int64_t row = r0 + wave; row < r1; row += kWavesPerWG)
    {
        const int64_t beg = a.indptr[row];
        const int64_t end = a.indptr[row + 1];
        T y = row_gather<T>(beg, end, a.indices, a.vals, a.X, s, lane);
        const int64_t o = row * kBlock + lane;
        T v = T(0);
        if (a.Vout || a.shift_mode == 1 || a.red_mode == 1)
            v = (row < a.shift_len || a.Vout || a.red_mode == 1) ? s * a.X[o] : T(0);
        if (a.Vout) a.Vout[o] = v;
        if (a.shift_mode == 1)
        {
            // y - (-t) x: imate adds t x through subtract_scaled_vector(-t)
            if (row < a.shift_len) y = y + tt * v;
        }
        else if (a.shift_mode == 2)
        {
            const T yb = row_gather<T>(a.b_indptr[row], a.b_indptr[row + 1],
                                       a.b_indices, a.b_vals, a.X, s, lane);
            y = y + tt * yb;
        }
        if (a.Sub) y = y - sub_c * (sub_s * a.Sub[o]);
        a.Y[o] = y;
        if (a.red_mode == 1) red += (double) v * (double) y;
        else if (a.red_mode == 2) red += (double) y * (double) y;
    }
    if (a.red_mode == 0) return;
    wg_partial(red, a.partial);
    double total;
    if (!last_block_total(a.partial, a.counter, total)) return;
    if (wave != 0) return;
    if (a.red_mode == 1 || !a.finish_sqrt)
    {
        a.red_out[lane] = total;
        return;
    }
    finish_norm(total, lane, a.j, a.tol_sqrt_n, a.red_out, a.scale_out, a.sizes);
}


// Row-tiled variant of spmm_kernel: one wavefront per tile of R rows, R
// accumulators per lane; every gathered operand row X[u] feeds the R values
// of column u (the slab is read with scalar loads since its address is
// uniform across the wavefront). The epilogue is the CSR kernel's, row by row.
template <typename T, int R>
__global__ __launch_bounds__(kThreads) void tiled_spmm_kernel(
    SpmmArgs<T> a, int64_t ntiles, const int64_t* __restrict__ tptr,
    const int32_t* __restrict__ tcols, const T* __restrict__ tvals)
{
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int64_t t0, t1;
    row_run(ntiles, t0, t1);
    const T s = a.xscale ? (T) a.xscale[lane] : T(1);
    const T tt = (T) a.t;
    T sub_c = T(0), sub_s = T(0);
    if (a.Sub)
    {
        sub_c = (T) a.sub_coef[lane];
        sub_s = (T) a.sub_scale[lane];
    }
    const T* __restrict__ X = a.X;
    double red = 0.0;
    for (int64_t t = t0 + wave; t < t1; t += kWavesPerWG)
    {
        T acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = T(0);
        const int64_t beg = tptr[t];
        const int64_t end = tptr[t + 1];
        for (int64_t base = beg; base < end; base += kWave)
        {
            const int64_t rem = end - base;
            const int cnt = rem < kWave ? (int) rem : kWave;
            const int cl = lane < cnt ? tcols[base + lane] : 0;
            const T* __restrict__ vb = tvals + base * R;
            int k = 0;
            for (; k + 4 <= cnt; k += 4)
            {
                const int c0 = readlane(cl, k);
                const int c1 = readlane(cl, k + 1);
                const int c2 = readlane(cl, k + 2);
                const int c3 = readlane(cl, k + 3);
                const T x0 = s * X[(int64_t) c0 * kBlock + lane];
                const T x1 = s * X[(int64_t) c1 * kBlock + lane];
                const T x2 = s * X[(int64_t) c2 * kBlock + lane];
                const T x3 = s * X[(int64_t) c3 * kBlock + lane];
                const T* v = vb + (int64_t) k * R;
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(v[r], x0, acc[r]);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(v[R + r], x1, acc[r]);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(v[2 * R + r], x2, acc[r]);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(v[3 * R + r], x3, acc[r]);
            }
            for (; k < cnt; ++k)
            {
                const int c = readlane(cl, k);
                const T x = s * X[(int64_t) c * kBlock + lane];
                const T* v = vb + (int64_t) k * R;
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(v[r], x, acc[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            const int64_t row = t * R + r;
            if (row >= a.nrows) break;
            const int64_t o = row * kBlock + lane;
            T y = acc[r];
            T v = T(0);
            if (a.Vout || a.red_mode == 1 || (a.shift_mode == 1 && row < a.shift_len))
                v = s * X[o];
            if (a.Vout) a.Vout[o] = v;
            if (a.shift_mode == 1 && row < a.shift_len) y = y + tt * v;
            if (a.Sub) y = y - sub_c * (sub_s * a.Sub[o]);
            a.Y[o] = y;
            if (a.red_mode == 1) red += (double) v * (double) y;
            else if (a.red_mode == 2) red += (double) y * (double) y;
        }
    }
    if (a.red_mode == 0) return;
    wg_partial(red, a.partial);
    double total;
    if (!last_block_total(a.partial, a.counter, total)) return;
    if (wave != 0) return;
    if (a.red_mode == 1 || !a.finish_sqrt)
    {
        a.red_out[lane] = total;
        return;
    }
    finish_norm(total, lane, a.j, a.tol_sqrt_n, a.red_out, a.scale_out, a.sizes);
}

// Lanczos three-term update r = (w - alpha v_j) - beta_{j-1} v_{j-1}, in
// place over w, with the sum of squares of r (the order of the two
// subtractions follows c_lanczos_tridiagonalization.cpp:196-209).
template <typename T>
struct UpdateArgs {
    int64_t nrows;
    T* R;                  // in: w, out: r
    const T* Vj;
    const T* Vjm1;         // nullptr at j = 0
    const double* alpha;   // [64]
    const double* beta;    // [64] beta_{j-1}
    int do_norm;           // 1: reduce |r|^2 and finish beta_j
    double* partial;
    unsigned int* counter;
    double* beta_out;
    double* scale_out;
    int32_t* sizes;
    int j;
    double tol_sqrt_n;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void lanczos_update_kernel(UpdateArgs<T> a)
{
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int64_t r0, r1;
    row_run(a.nrows, r0, r1);
    const double al = a.alpha[lane];
    const double be = a.Vjm1 ? a.beta[lane] : 0.0;
    double red = 0.0;
    for (int64_t row = r0 + wave; row < r1; row += kWavesPerWG)
    {
        const int64_t o = row * kBlock + lane;
        double r = (double) a.R[o] - al * (double) a.Vj[o];
        if (a.Vjm1) r = r - be * (double) a.Vjm1[o];
        const T rt = (T) r;
        a.R[o] = rt;
        red += (double) rt * (double) rt;
    }
    if (!a.do_norm) return;
    wg_partial(red, a.partial);
    double total;
    if (!last_block_total(a.partial, a.counter, total)) return;
    if (wave != 0) return;
    finish_norm(total, lane, a.j, a.tol_sqrt_n, a.beta_out, a.scale_out, a.sizes);
}

// ------------------------------------------------ Gram-Schmidt (CGS2)

constexpr int kGsGroup = 8;     // basis vectors per register group

// Optionally r -= sum_k h_k basis_k (h from a previous pass), then either the
// per-probe dots h'_k = <basis_k, r> (partials [G][K][64]) or |r|^2.
template <typename T>
struct GsArgs {
    int64_t nrows;
    T* R;
    const T* basis;          // [nslots][nrows][64]
    int last;                // slot of the newest vector; vector k is slot last - k
    int K;
    const double* h_sub;     // [K][64] or nullptr (no subtraction)
    int mode;                // 0: dots into partial, 1: norm finish
    double* partial;         // [G][K][64] (mode 0) or [G][64] (mode 1)
    unsigned int* counter;
    double* beta_out;
    double* scale_out;
    int32_t* sizes;
    int j;
    double tol_sqrt_n;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void gs_kernel(GsArgs<T> a)
{
    const int lane = threadIdx.x % kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int64_t r0, r1;
    row_run(a.nrows, r0, r1);
    const int64_t stride = a.nrows * kBlock;
    if (a.h_sub)
    {
        for (int64_t row = r0 + wave; row < r1; row += kWavesPerWG)
        {
            const int64_t o = row * kBlock + lane;
            double r = (double) a.R[o];
            for (int k = 0; k < a.K; ++k)
                r -= a.h_sub[k * kWave + lane] *
                     (double) a.basis[(int64_t) (a.last - k) * stride + o];
            a.R[o] = (T) r;
        }
    }
    if (a.mode == 0)
    {
        __shared__ double red[kWavesPerWG][kGsGroup][kWave];
        for (int k0 = 0; k0 < a.K; k0 += kGsGroup)
        {
            const int kn = (a.K - k0) < kGsGroup ? (a.K - k0) : kGsGroup;
            double acc[kGsGroup];
#pragma unroll
            for (int g = 0; g < kGsGroup; ++g) acc[g] = 0.0;
            for (int64_t row = r0 + wave; row < r1; row += kWavesPerWG)
            {
                const int64_t o = row * kBlock + lane;
                const double r = (double) a.R[o];
#pragma unroll
                for (int g = 0; g < kGsGroup; ++g)
                    if (g < kn)
                        acc[g] += r * (double) a.basis[(int64_t) (a.last - k0 - g) * stride + o];
            }
#pragma unroll
            for (int g = 0; g < kGsGroup; ++g) red[wave][g][lane] = acc[g];
            __syncthreads();
            if (wave < kn)
            {
                double s = 0.0;
#pragma unroll
                for (int w = 0; w < kWavesPerWG; ++w) s += red[w][wave][lane];
                a.partial[((int64_t) blockIdx.x * a.K + k0 + wave) * kWave + lane] = s;
            }
            __syncthreads();
        }
        return;
    }
    double red = 0.0;
    for (int64_t row = r0 + wave; row < r1; row += kWavesPerWG)
    {
        const double r = (double) a.R[row * kBlock + lane];
        red += r * r;
    }
    wg_partial(red, a.partial);
    double total;
    if (!last_block_total(a.partial, a.counter, total)) return;
    if (wave != 0) return;
    finish_norm(total, lane, a.j, a.tol_sqrt_n, a.beta_out, a.scale_out, a.sizes);
}

// out[k][p] = sum_g partial[g][k][p], fixed order; one workgroup per k.
__global__ __launch_bounds__(256) void reduce_partials_kernel(
    const double* __restrict__ partial, int G, int K, double* __restrict__ out)
{
    __shared__ double red[4][kWave];
    const int lane = threadIdx.x % kWave;
    const int grp = threadIdx.x / kWave;
    const int k = blockIdx.x;
    double s = 0.0;
    for (int g = grp; g < G; g += 4)
        s += partial[((int64_t) g * K + k) * kWave + lane];
    red[grp][lane] = s;
    __syncthreads();
    if (grp == 0) out[k * kWave + lane] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
}

// Block set-up: scale[p] = s0 for probes p < count (0 otherwise), sizes = 0.
__global__ void init_block_kernel(double* scale, int32_t* sizes, int count,
                                  double s0)
{
    const int p = threadIdx.x;
    if (p < kBlock)
    {
        scale[p] = p < count ? s0 : 0.0;
        sizes[p] = 0;
    }
}

__global__ void zero_u32_kernel(unsigned int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename S, typename D>
__global__ void convert_kernel(const S* __restrict__ in, D* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (D) in[i];
}

// ---------------------------------------------------------------- host side

int grid_for(int64_t rows)
{
    int64_t g = (rows + 63) / 64;            // aim at >= 64 rows per workgroup
    if (g > kMaxGrid) g = kMaxGrid;
    g = ((g + kXcds - 1) / kXcds) * kXcds;
    if (g < kXcds) g = kXcds;
    return (int) g;
}

struct Bump {
    char* p;
    size_t left;
    template <typename U>
    U* take(size_t n)
    {
        size_t bytes = (n * sizeof(U) + 255) & ~(size_t) 255;
        if (bytes > left) return nullptr;
        U* r = (U*) p;
        p += bytes;
        left -= bytes;
        return r;
    }
};

template <typename T>
const T* values_of(MatrixStore& M, bool trans, bool want_f32)
{
    if (want_f32 && M.csr.dtype == SLQ_F64)
        return (const T*) (trans ? (void*) M.trans_data_f32 : (void*) M.data_f32);
    return (const T*) (trans ? M.trans.data : M.csr.data);
}

template <typename T>
void fill_matrix(SpmmArgs<T>& s, MatrixStore& A, MatrixStore* B, bool trans,
                 bool f32)
{
    const DevCSR& c = trans ? A.trans : A.csr;
    s.indptr = c.indptr;
    s.indices = c.indices;
    s.vals = values_of<T>(A, trans, f32);
    s.nrows = c.nrows;
    if (B)
    {
        const DevCSR& cb = trans ? B->trans : B->csr;
        s.b_indptr = cb.indptr;
        s.b_indices = cb.indices;
        s.b_vals = values_of<T>(*B, trans, f32);
    }
}

template <typename T>
SpmmArgs<T> spmm_defaults()
{
    SpmmArgs<T> s;
    std::memset(&s, 0, sizeof(s));
    return s;
}

// Tiled form of the operand matrix for this launch, or nullptr (CSR path):
// only for A and A + tI (the general A + tB keeps the two CSR streams).
template <typename T>
const DevTiles* tiles_for(slq_op* op, bool trans)
{
    if (op->kind == OP_AFFINE_B) return nullptr;
    const DevTiles& t = trans ? op->A->trans_tiles : op->A->tiles;
    if (t.R == 0) return nullptr;
    if (sizeof(T) == 4 && op->A->csr.dtype == SLQ_F64 && !t.vals_f32) return nullptr;
    return &t;
}

template <typename T>
int launch_spmm(slq_op* op, const SpmmArgs<T>& s, const DevTiles* tl = nullptr)
{
    hipEvent_t ev = nullptr;
    prof_begin(op, &ev);
    if (tl)
    {
        const T* tv = (sizeof(T) == 4 && op->A->csr.dtype == SLQ_F64)
                          ? (const T*) tl->vals_f32 : (const T*) tl->vals;
        if (tl->R == 8)
            hipLaunchKernelGGL((tiled_spmm_kernel<T, 8>), dim3(grid_for(tl->ntiles)),
                               dim3(kThreads), 0, op->stream, s, tl->ntiles, tl->ptr,
                               tl->cols, tv);
        else
            hipLaunchKernelGGL((tiled_spmm_kernel<T, 16>), dim3(grid_for(tl->ntiles)),
                               dim3(kThreads), 0, op->stream, s, tl->ntiles, tl->ptr,
                               tl->cols, tv);
    }
    else
        hipLaunchKernelGGL(spmm_kernel<T>, dim3(grid_for(s.nrows)), dim3(kThreads),
                           0, op->stream, s);
    SLQ_CHECK(hipGetLastError());
    prof_end(op, ev, 0);
    return SLQ_OK;
}

template <typename T>
int launch_update(slq_op* op, const UpdateArgs<T>& u)
{
    hipEvent_t ev = nullptr;
    prof_begin(op, &ev);
    hipLaunchKernelGGL(lanczos_update_kernel<T>, dim3(grid_for(u.nrows)),
                       dim3(kThreads), 0, op->stream, u);
    SLQ_CHECK(hipGetLastError());
    prof_end(op, ev, 1);
    return SLQ_OK;
}

// CGS2 of r against the K newest basis vectors (slots last, last-1, ...);
// the final pass finishes the norm (beta_j, or alpha_j for Golub-Kahn u).
template <typename T>
int reorthogonalize(slq_op* op, int64_t nrows, T* R, const T* basis, int last,
                    int K, double* h1, double* h2, double* partial_k,
                    double* partial, unsigned int* counter, double* norm_out,
                    double* scale_out, int32_t* sizes, int j, double tol_sqrt_n)
{
    const int G = grid_for(nrows);
    GsArgs<T> g;
    std::memset(&g, 0, sizeof(g));
    g.nrows = nrows;
    g.R = R;
    g.basis = basis;
    g.last = last;
    g.K = K;
    g.counter = counter;
    g.beta_out = norm_out;
    g.scale_out = scale_out;
    g.sizes = sizes;
    g.j = j;
    g.tol_sqrt_n = tol_sqrt_n;
    if (K > 0)
    {
        // pass 1: h1 = V^T r
        g.h_sub = nullptr;
        g.mode = 0;
        g.partial = partial_k;
        hipLaunchKernelGGL(gs_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g);
        SLQ_CHECK(hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), dim3(256), 0, op->stream,
                           partial_k, G, K, h1);
        SLQ_CHECK(hipGetLastError());
        // pass 2: r -= V h1; h2 = V^T r
        g.h_sub = h1;
        hipLaunchKernelGGL(gs_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g);
        SLQ_CHECK(hipGetLastError());
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), dim3(256), 0, op->stream,
                           partial_k, G, K, h2);
        SLQ_CHECK(hipGetLastError());
        g.h_sub = h2;
    }
    // pass 3: r -= V h2; norm
    g.mode = 1;
    g.partial = partial;
    hipLaunchKernelGGL(gs_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

// X[i][p] *= scale[p]; then scale[p] = 1 (second launch, after every reader).
template <typename T>
__global__ __launch_bounds__(256) void scale_rows_kernel(T* __restrict__ X, int64_t n,
                                                         const double* __restrict__ scale)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * kBlock) X[i] = (T) ((double) X[i] * scale[i % kBlock]);
}

__global__ void set_ones_kernel(double* s)
{
    if (threadIdx.x < kBlock) s[threadIdx.x] = 1.0;
}

template <typename T>
int normalize_in_place(slq_op* op, T* X, int64_t n, double* scale)
{
    const int64_t g = (n * kBlock + 255) / 256;
    hipLaunchKernelGGL(scale_rows_kernel<T>, dim3((unsigned) g), dim3(256), 0,
                       op->stream, X, n, (const double*) scale);
    hipLaunchKernelGGL(set_ones_kernel, dim3(1), dim3(kBlock), 0, op->stream, scale);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

template <typename T>
__global__ __launch_bounds__(256) void copy_block_kernel(const T* __restrict__ in,
                                                         T* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

template <typename T>
int lanczos(const RecurrenceArgs& ra, Bump& bump)
{
    slq_op* op = ra.op;
    MatrixStore& A = *op->A;
    const int64_t n = A.csr.nrows;
    const int m = ra.m;
    const bool f32 = (sizeof(T) == 4);
    const bool reorth = ra.ortho != 0;
    int buffer = 2;
    if (ra.ortho == 0 || ra.ortho == 1) buffer = 2;
    else if (ra.ortho < 0 || ra.ortho > m) buffer = m;
    else buffer = ra.ortho;
    const int nslots = reorth ? m : 2;
    const size_t blk = (size_t) n * kBlock;

    T* P = bump.take<T>(blk);
    T* Q = bump.take<T>(blk);
    T* V = bump.take<T>(blk * nslots);
    double* scale = bump.take<double>(kBlock);
    double* partial = bump.take<double>((size_t) kMaxGrid * kBlock);
    unsigned int* counter = bump.take<unsigned int>(64);
    double* h1 = nullptr;
    double* h2 = nullptr;
    double* partial_k = nullptr;
    if (reorth)
    {
        h1 = bump.take<double>((size_t) (m + 1) * kBlock);
        h2 = bump.take<double>((size_t) (m + 1) * kBlock);
        partial_k = bump.take<double>((size_t) kMaxGrid * (m + 1) * kBlock);
    }
    if (!P || !Q || !V || !scale || !partial || !counter ||
        (reorth && (!h1 || !h2 || !partial_k)))
    {
        set_error("internal: Lanczos workspace too small");
        return SLQ_E_OOM;
    }
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, op->stream, counter, 64);
    // initial_beta = |probe| = sqrt(n) exactly for a Rademacher probe
    const double s0 = 1.0 / std::sqrt((double) n);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       scale, ra.sizes, kBlock, s0);
    SLQ_CHECK(hipGetLastError());
    const double tol_sqrt_n = ra.tol * std::sqrt((double) n);

    const T* src = (const T*) ra.X0;
    T* bufs[2] = {Q, P};
    for (int j = 0; j < m; ++j)
    {
        T* W = bufs[j & 1];
        T* Vj = V + (size_t) (reorth ? j : (j & 1)) * blk;
        T* Vjm1 = j > 0 ? V + (size_t) (reorth ? j - 1 : ((j - 1) & 1)) * blk : nullptr;

        SpmmArgs<T> s = spmm_defaults<T>();
        fill_matrix<T>(s, A, op->kind == OP_AFFINE_B ? op->B.get() : nullptr, false, f32);
        s.shift_mode = op->kind == OP_AFFINE_I ? 1 : (op->kind == OP_AFFINE_B ? 2 : 0);
        s.shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
        s.t = ra.t;
        s.X = src;
        s.xscale = scale;
        s.Vout = Vj;
        s.Y = W;
        s.red_mode = 1;
        s.partial = partial;
        s.counter = counter;
        s.red_out = ra.alpha + (size_t) j * kBlock;
        SLQ_TRY(launch_spmm<T>(op, s, tiles_for<T>(op, false)));

        UpdateArgs<T> u;
        std::memset(&u, 0, sizeof(u));
        u.nrows = n;
        u.R = W;
        u.Vj = Vj;
        u.Vjm1 = Vjm1;
        u.alpha = ra.alpha + (size_t) j * kBlock;
        u.beta = j > 0 ? ra.beta + (size_t) (j - 1) * kBlock : nullptr;
        u.do_norm = reorth ? 0 : 1;
        u.partial = partial;
        u.counter = counter;
        u.beta_out = ra.beta + (size_t) j * kBlock;
        u.scale_out = scale;
        u.sizes = ra.sizes;
        u.j = j;
        u.tol_sqrt_n = tol_sqrt_n;
        SLQ_TRY(launch_update<T>(op, u));

        if (reorth)
        {
            const int num_ortho = j < buffer ? j + 1 : buffer;
            SLQ_TRY(reorthogonalize<T>(op, n, W, V, j, num_ortho, h1, h2,
                                       partial_k, partial, counter,
                                       ra.beta + (size_t) j * kBlock, scale,
                                       ra.sizes, j, tol_sqrt_n));
        }
        src = W;
    }
    return SLQ_OK;
}

// Golub-Kahn on a (possibly non-square) operator: u vectors have nrows
// entries and v vectors ncols. Without re-orthogonalisation the vectors are
// kept unnormalised with a per-probe scale (1/alpha for u, 1/beta for v)
// applied where they are read; with it they are normalised in place so that
// the stored basis is orthonormal.
template <typename T>
int golub_kahn(const RecurrenceArgs& ra, Bump& bump)
{
    slq_op* op = ra.op;
    MatrixStore& A = *op->A;
    MatrixStore* Bm = op->kind == OP_AFFINE_B ? op->B.get() : nullptr;
    const int64_t nr = A.csr.nrows, nc = A.csr.ncols;
    const int m = ra.m;
    const bool f32 = (sizeof(T) == 4);
    const bool reorth = ra.ortho != 0;
    int buffer = 2;
    if (ra.ortho == 0) buffer = 2;
    else if (ra.ortho < 0 || ra.ortho > m - 1) buffer = m;
    else buffer = ra.ortho + 1;
    const int nslots = reorth ? m + 1 : 3;
    const size_t ublk = (size_t) nr * kBlock, vblk = (size_t) nc * kBlock;

    T* U = bump.take<T>(ublk * nslots);
    T* V = bump.take<T>(vblk * nslots);
    double* su = bump.take<double>((size_t) (m + 1) * kBlock);   // 1/alpha_j
    double* sv = bump.take<double>((size_t) (m + 1) * kBlock);   // 1/beta_{j-1}
    double* partial = bump.take<double>((size_t) kMaxGrid * kBlock);
    unsigned int* counter = bump.take<unsigned int>(64);
    int32_t* usizes = bump.take<int32_t>(kBlock);
    double* h1 = nullptr;
    double* h2 = nullptr;
    double* partial_k = nullptr;
    if (reorth)
    {
        h1 = bump.take<double>((size_t) (m + 1) * kBlock);
        h2 = bump.take<double>((size_t) (m + 1) * kBlock);
        partial_k = bump.take<double>((size_t) kMaxGrid * (m + 1) * kBlock);
    }
    if (!U || !V || !su || !sv || !partial || !counter || !usizes ||
        (reorth && (!h1 || !h2 || !partial_k)))
    {
        set_error("internal: Golub-Kahn workspace too small");
        return SLQ_E_OOM;
    }
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, op->stream, counter, 64);
    // v_0 = probe / |probe|, |probe| = sqrt(ncols) exactly
    const double s0 = 1.0 / std::sqrt((double) nc);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       sv, ra.sizes, kBlock, s0);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       su, usizes, kBlock, 0.0);
    SLQ_CHECK(hipGetLastError());
    auto uslot = [&](int j) { return U + (size_t) (reorth ? j : j % 3) * ublk; };
    auto vslot = [&](int j) { return V + (size_t) (reorth ? j : j % 3) * vblk; };
    {
        const int64_t g = ((int64_t) vblk + 255) / 256;
        hipLaunchKernelGGL(copy_block_kernel<T>, dim3((unsigned) g), dim3(256), 0,
                           op->stream, (const T*) ra.X0, vslot(0), (int64_t) vblk);
        SLQ_CHECK(hipGetLastError());
        if (reorth) SLQ_TRY(normalize_in_place<T>(op, vslot(0), nc, sv));
    }
    const double tol_sqrt_n = ra.tol * std::sqrt((double) nc);
    const int64_t shift_len = nr < nc ? nr : nc;
    const int mode = op->kind == OP_AFFINE_I ? 1 : (op->kind == OP_AFFINE_B ? 2 : 0);

    for (int j = 0; j < m; ++j)
    {
        // u_j = A v_j - beta_{j-1} u_{j-1}; alpha_j = |u_j|
        // (c_golub_kahn_bidiagonalization.cpp:162-195)
        SpmmArgs<T> s = spmm_defaults<T>();
        fill_matrix<T>(s, A, Bm, false, f32);
        s.shift_mode = mode;
        s.shift_len = shift_len;
        s.t = ra.t;
        s.X = vslot(j);
        s.xscale = sv + (size_t) j * kBlock;
        s.Y = uslot(j);
        if (j > 0)
        {
            s.Sub = uslot(j - 1);
            s.sub_coef = ra.beta + (size_t) (j - 1) * kBlock;
            s.sub_scale = su + (size_t) (j - 1) * kBlock;
        }
        s.red_mode = reorth ? 0 : 2;
        s.finish_sqrt = 1;
        s.partial = partial;
        s.counter = counter;
        s.red_out = ra.alpha + (size_t) j * kBlock;
        s.scale_out = su + (size_t) j * kBlock;
        s.sizes = usizes;            // a small alpha never ends the recurrence
        s.j = j;
        s.tol_sqrt_n = -1.0;
        SLQ_TRY(launch_spmm<T>(op, s, tiles_for<T>(op, false)));
        if (reorth)
        {
            const int num_ortho = j < buffer ? j : buffer - 1;
            SLQ_TRY(reorthogonalize<T>(op, nr, uslot(j), U, j - 1,
                                       j > 0 ? num_ortho : 0, h1, h2, partial_k,
                                       partial, counter,
                                       ra.alpha + (size_t) j * kBlock,
                                       su + (size_t) j * kBlock, usizes, j, -1.0));
            SLQ_TRY(normalize_in_place<T>(op, uslot(j), nr, su + (size_t) j * kBlock));
        }

        // v_{j+1} = A^T u_j - alpha_j v_j; beta_j = |v_{j+1}|
        // (c_golub_kahn_bidiagonalization.cpp:197-228)
        SpmmArgs<T> t = spmm_defaults<T>();
        fill_matrix<T>(t, A, Bm, true, f32);
        t.shift_mode = mode;
        t.shift_len = shift_len;
        t.t = ra.t;
        t.X = uslot(j);
        t.xscale = su + (size_t) j * kBlock;
        t.Y = vslot(j + 1);
        t.Sub = vslot(j);
        t.sub_coef = ra.alpha + (size_t) j * kBlock;
        t.sub_scale = sv + (size_t) j * kBlock;
        t.red_mode = reorth ? 0 : 2;
        t.finish_sqrt = 1;
        t.partial = partial;
        t.counter = counter;
        t.red_out = ra.beta + (size_t) j * kBlock;
        t.scale_out = sv + (size_t) (j + 1) * kBlock;
        t.sizes = ra.sizes;
        t.j = j;
        t.tol_sqrt_n = tol_sqrt_n;
        SLQ_TRY(launch_spmm<T>(op, t, tiles_for<T>(op, true)));
        if (reorth)
        {
            const int num_ortho = j < buffer ? j : buffer - 1;
            SLQ_TRY(reorthogonalize<T>(op, nc, vslot(j + 1), V, j, num_ortho, h1,
                                       h2, partial_k, partial, counter,
                                       ra.beta + (size_t) j * kBlock,
                                       sv + (size_t) (j + 1) * kBlock, ra.sizes,
                                       j, tol_sqrt_n));
            SLQ_TRY(normalize_in_place<T>(op, vslot(j + 1), nc,
                                          sv + (size_t) (j + 1) * kBlock));
        }
    }
    return SLQ_OK;
}

}  // namespace

// ---------------------------------------------------------------- exported

size_t recurrence_scratch_bytes(const slq_op* op, int gram, int m, int ortho,
                                int value_bytes)
{
    const int64_t nr = op->A->csr.nrows, nc = op->A->csr.ncols;
    const bool reorth = ortho != 0;
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (bytes + 255) & ~(size_t) 255; };
    if (!gram)
    {
        const size_t blk = (size_t) nr * kBlock * value_bytes;
        add(blk);
        add(blk);
        add(blk * (reorth ? m : 2));
    }
    else
    {
        const int ns = reorth ? m + 1 : 3;
        add((size_t) nr * kBlock * value_bytes * ns);
        add((size_t) nc * kBlock * value_bytes * ns);
        add(sizeof(double) * (m + 1) * kBlock * 2);
        add(sizeof(int32_t) * kBlock);
    }
    add(sizeof(double) * kBlock);
    add(sizeof(double) * kMaxGrid * kBlock);
    add(sizeof(unsigned int) * 64);
    if (reorth)
    {
        add(sizeof(double) * (m + 1) * kBlock * 2);
        add(sizeof(double) * (size_t) kMaxGrid * (m + 1) * kBlock);
        add(sizeof(int32_t) * (m + 1));
    }
    return b + 4096;
}

int run_recurrence(const RecurrenceArgs& ra, void* scratch, size_t scratch_bytes)
{
    Bump bump{(char*) scratch, scratch_bytes};
    const bool f32 = ra.precision == SLQ_PREC_MIXED || ra.op->A->csr.dtype == SLQ_F32;
    if (ra.gram && !ra.op->A->has_trans)
    {
        set_error("internal: transpose not built");
        return SLQ_E_ARG;
    }
    if (f32)
        return ra.gram ? golub_kahn<float>(ra, bump) : lanczos<float>(ra, bump);
    return ra.gram ? golub_kahn<double>(ra, bump) : lanczos<double>(ra, bump);
}

int run_spmm(slq_op* op, double t, int transpose, const void* X, void* Y,
             int value_dtype)
{
    MatrixStore& A = *op->A;
    MatrixStore* Bm = op->kind == OP_AFFINE_B ? op->B.get() : nullptr;
    const bool f32 = value_dtype == SLQ_F32;
    const int64_t shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
    const int mode = op->kind == OP_AFFINE_I ? 1 : (op->kind == OP_AFFINE_B ? 2 : 0);
    if (f32)
    {
        SpmmArgs<float> s = spmm_defaults<float>();
        fill_matrix<float>(s, A, Bm, transpose != 0, A.csr.dtype == SLQ_F64);
        s.shift_mode = mode;
        s.shift_len = shift_len;
        s.t = t;
        s.X = (const float*) X;
        s.Y = (float*) Y;
        return launch_spmm<float>(op, s, tiles_for<float>(op, transpose != 0));
    }
    SpmmArgs<double> s = spmm_defaults<double>();
    fill_matrix<double>(s, A, Bm, transpose != 0, false);
    s.shift_mode = mode;
    s.shift_len = shift_len;
    s.t = t;
    s.X = (const double*) X;
    s.Y = (double*) Y;
    return launch_spmm<double>(op, s, tiles_for<double>(op, transpose != 0));
}

int tiles_f32(hipStream_t s, DevTiles& t, int dtype)
{
    if (dtype != SLQ_F64 || t.R == 0 || t.vals_f32) return SLQ_OK;
    const int64_t n = t.nslabs * t.R;
    SLQ_CHECK(hipMalloc(&t.vals_f32, sizeof(float) * n));
    const int64_t g = (n + 255) / 256;
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned) g), dim3(256), 0,
                       s, (const double*) t.vals, t.vals_f32, n);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

int ensure_f32(hipStream_t s, MatrixStore& M, bool need_trans)
{
    if (M.csr.dtype != SLQ_F64) return SLQ_OK;
    SLQ_TRY(tiles_f32(s, M.tiles, M.csr.dtype));
    if (need_trans) SLQ_TRY(tiles_f32(s, M.trans_tiles, M.csr.dtype));
    if (!M.data_f32)
    {
        SLQ_CHECK(hipMalloc(&M.data_f32, sizeof(float) * (M.csr.nnz > 0 ? M.csr.nnz : 1)));
        const int64_t g = (M.csr.nnz + 255) / 256;
        if (g > 0)
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned) g),
                               dim3(256), 0, s, (const double*) M.csr.data,
                               M.data_f32, M.csr.nnz);
        SLQ_CHECK(hipGetLastError());
    }
    if (need_trans && M.has_trans && !M.trans_data_f32)
    {
        SLQ_CHECK(hipMalloc(&M.trans_data_f32, sizeof(float) * (M.trans.nnz > 0 ? M.trans.nnz : 1)));
        const int64_t g = (M.trans.nnz + 255) / 256;
        if (g > 0)
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned) g),
                               dim3(256), 0, s, (const double*) M.trans.data,
                               M.trans_data_f32, M.trans.nnz);
        SLQ_CHECK(hipGetLastError());
    }
    return SLQ_OK;
}

}  // namespace slq
