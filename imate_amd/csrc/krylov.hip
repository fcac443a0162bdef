// Block Krylov recurrences on gfx950: CSR / row-tiled SpMM over a 64-probe
// block with fused Lanczos / Golub-Kahn epilogues, deterministic grid
// reductions and classical Gram-Schmidt (twice) re-orthogonalisation.
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
// owns probe p, so a row of a block is one coalesced 512 B (FP64) access and
// the gather of column c reads the contiguous row c of the operand block.
// The matrix-core SpMM maps lanes to probes by its MFMA operand layout
// instead (Lanes<T>), keeping every 16 B access on whole cache lines.
//
// Work distribution ("XCD windows"): workgroup ids are dealt round-robin to
// the 8 XCDs (id mod 8). XCD x owns the contiguous eighth x of the rows (or
// row tiles), and its workgroups sweep that range together, one window of
// (workgroups per XCD) x (wavefronts per workgroup) units at a time. Units
// processed concurrently on an XCD are therefore neighbours; on a spatially
// ordered matrix their gathered operand rows overlap and stay in that XCD's
// 4 MB L2. The tiled SpMM hands each XCD's tiles out in order from a
// per-XCD ticket counter instead, for the same locality with dynamic load
// balance.
//
// Reductions (alpha, beta, norms) are per-probe sums over all rows, computed
// deterministically and without fences: each workgroup (or, in the tiled
// SpMM, each row tile) writes one partial row, and a following launch of
// one workgroup adds the rows in a fixed order and finishes the value (the
// kernel boundary orders the two). The result is bitwise reproducible run
// to run.

#include <cmath>
#include <cstdio>
#include <cstdlib>
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

__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// Units [0, N) swept in XCD windows (see the header). G is a multiple of 8.
struct Sweep {
    int64_t first;
    int64_t end;
    int64_t stride;
};

template <int WPG>
__device__ __forceinline__ Sweep sweep_wpg(int64_t N, int wave)
{
    const int64_t G = gridDim.x;
    const int64_t bid = blockIdx.x;
    const int64_t x = bid % kXcds;
    const int64_t local = bid / kXcds;
    const int64_t per_xcd = G / kXcds;
    const int64_t span = (N + kXcds - 1) / kXcds;
    const int64_t b = x * span;
    Sweep s;
    s.end = b + span < N ? b + span : N;
    s.first = b + local * WPG + wave;
    s.stride = per_xcd * WPG;
    return s;
}

__device__ __forceinline__ Sweep sweep(int64_t N, int wave)
{
    const int64_t G = gridDim.x;
    const int64_t bid = blockIdx.x;
    const int64_t x = bid % kXcds;
    const int64_t local = bid / kXcds;
    const int64_t per_xcd = G / kXcds;
    const int64_t span = (N + kXcds - 1) / kXcds;
    const int64_t b = x * span;
    Sweep s;
    s.end = b + span < N ? b + span : N;
    s.first = b + local * kWavesPerWG + wave;
    s.stride = per_xcd * kWavesPerWG;
    return s;
}

// The same units swept grid-wide: wave w of the grid takes units w, w + W,
// ... (W = waves in the grid), so at any moment the whole chip streams one
// contiguous window of memory.
__device__ __forceinline__ Sweep sweep_linear(int64_t N, int wave)
{
    Sweep s;
    s.first = (int64_t) blockIdx.x * kWavesPerWG + wave;
    s.stride = (int64_t) gridDim.x * kWavesPerWG;
    s.end = N;
    return s;
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

// ---------------------------------------------------------------- reduction

struct RedBuf {
    double* partial;            // [G][64], one row per workgroup
    // in-order tile schedule of the tiled SpMM: per XCD a ticket counter
    // (uint 64 * x) and a finished-wave counter (uint 64 * x + 32), zero
    // between launches, and one partial row per tile ([ntiles][64])
    unsigned int* tickets;
    double* tile_partial;
};
constexpr int kTicketWords = kXcds * 64;

// Adds v over the workgroup's wavefronts (lane = probe) and stores the sum
// as this workgroup's partial row rb.partial[blockIdx.x]. A separate
// one-workgroup launch adds the rows in order (ordered_total), so no kernel
// needs a grid-wide fence: each such fence writes back the XCD's L2, which
// cost a streaming kernel over a tenth of its time.
template <int WPG = kWavesPerWG>
__device__ __forceinline__ void wg_partial(double v, const RedBuf& rb)
{
    __shared__ double red[WPG][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    red[wave][lane] = v;
    __syncthreads();
    if (wave != 0) return;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < WPG; ++w) s += red[w][lane];
    rb.partial[(int64_t) blockIdx.x * kWave + lane] = s;
}

// Sum over rows [0, G) of rows[g][lane] for a workgroup of kTotalWaves
// wavefronts: wave w adds the rows of its contiguous range in order, eight
// independent loads at a time (the rows come from another XCD's writes, so
// each load is a long round trip), and the wave sums are added in wave
// order: the result does not depend on timing. Valid in wave 0.
constexpr int kTotalWaves = 16;

__device__ __forceinline__ double ordered_total(const double* __restrict__ rows, int64_t G,
                                                int64_t row_stride = kWave)
{
    __shared__ double red[kTotalWaves][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int64_t per = (G + kTotalWaves - 1) / kTotalWaves;
    const int64_t b0 = wave * per;
    const int64_t b1 = b0 + per < G ? b0 + per : G;
    double s = 0.0;
    int64_t b = b0;
    for (; b + 8 <= b1; b += 8)
    {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = rows[(b + k) * row_stride + lane];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; b < b1; ++b) s += rows[b * row_stride + lane];
    red[wave][lane] = s;
    __syncthreads();
    double t = 0.0;
    if (wave == 0)
        for (int w = 0; w < kTotalWaves; ++w) t += red[w][lane];
    return t;
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

// ---------------------------------------------------------------- SpMM

template <typename T>
struct SpmmArgs {
    // A (rows of the output) and optionally B of A + tB (CSR)
    const int64_t* indptr;
    const int32_t* indices;
    const T* vals;
    const int64_t* b_indptr;
    const int32_t* b_indices;
    const T* b_vals;
    int64_t nrows;
    int64_t ncols;          // columns of A (rows of the gathered operand)
    int sorted;             // A's column indices ascend within every row
    double avg_row_nnz;     // non-zeros of A per row
    const T* dense;         // dense A (row-major, rows of ld values) or nullptr
    int64_t ld;
    int64_t shift_len;      // rows that get + t x[i] (identity shift)
    int shift_mode;         // 0 none, 1 + t x (B = I), 2 + t B x
    double t;
    const double* xscale;   // per-probe scale of the gathered operand (nullptr -> 1)
    // epilogue subtraction: y -= sub_coef[p] * (sub_scale[p] * Sub[i])
    const double* sub_coef;
    const double* sub_scale;
    // reduction: 0 none, 1 sum v*y -> red_out (alpha), 2 sum y*y -> norm
    int red_mode;
    RedBuf rb;
    double* red_out;
    int finish_sqrt;        // red_mode 2: red_out = sqrt, scale_out = 1/sqrt
    double* scale_out;
    int32_t* sizes;
    int j;
    double tol_sqrt_n;
    int nt_store;           // non-temporal stores of Y / Vout
};

// Epilogue of one output row: optional v = s X[row] (written to Vout), the
// affine shift, the subtraction, the store and the reduction term.
template <typename T, bool NT>
__device__ __forceinline__ void spmm_epilogue(const SpmmArgs<T>& a, int64_t row, int lane,
                                              T y, T s, T tt, T sub_c, T sub_s,
                                              const T* __restrict__ X,
                                              T* __restrict__ Y, T* __restrict__ Vout,
                                              const T* __restrict__ Sub, double& red)
{
    const int64_t o = row * kBlock + lane;
    const bool shift_row = a.shift_mode == 1 && row < a.shift_len;
    T v = T(0);
    if (Vout || a.red_mode == 1 || shift_row) v = s * X[o];
    // imate adds t x through subtract_scaled_vector(-t): y - (-t) x
    if (shift_row) y = y + tt * v;
    // the reduction term v . (A + tB) v is taken before the subtraction
    if (a.red_mode == 1) red += (double) v * (double) y;
    if (Sub) y = y - sub_c * (sub_s * Sub[o]);
    if constexpr (NT)
    {
        // streamed outputs: do not let them evict gathered operand rows
        if (Vout) __builtin_nontemporal_store(v, &Vout[o]);
        __builtin_nontemporal_store(y, &Y[o]);
    }
    else
    {
        if (Vout) Vout[o] = v;
        Y[o] = y;
    }
    if (a.red_mode == 2) red += (double) y * (double) y;
}

// Stores the grid total of the SpMM's reduction term (alpha, or |u|^2 /
// its root for Golub-Kahn).
template <typename T>
__device__ __forceinline__ void spmm_store_total(const SpmmArgs<T>& a, double total, int lane)
{
    if (a.red_mode == 1 || !a.finish_sqrt)
    {
        a.red_out[lane] = total;
        return;
    }
    finish_norm(total, lane, a.j, a.tol_sqrt_n, a.red_out, a.scale_out, a.sizes);
}

// Partial row of the SpMM's reduction term; total_kernel finishes it.
template <typename T>
__device__ __forceinline__ void spmm_finish(const SpmmArgs<T>& a, double red)
{
    if (a.red_mode == 0) return;
    wg_partial(red, a.rb);
}

// CSR form: one wavefront per row.
template <typename T, bool NT>
__global__ __launch_bounds__(kThreads) void spmm_kernel(SpmmArgs<T> a,
                                                        const T* __restrict__ X,
                                                        T* __restrict__ Y,
                                                        T* __restrict__ Vout,
                                                        const T* __restrict__ Sub)
{
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const Sweep sw = sweep(a.nrows, wave);
    const T s = a.xscale ? (T) a.xscale[lane] : T(1);
    const T tt = (T) a.t;
    const T sub_c = Sub ? (T) a.sub_coef[lane] : T(0);
    const T sub_s = Sub ? (T) a.sub_scale[lane] : T(0);
    double red = 0.0;
    for (int64_t row = sw.first; row < sw.end; row += sw.stride)
    {
        T y = row_gather<T>(a.indptr[row], a.indptr[row + 1], a.indices, a.vals, X, s, lane);
        if (a.shift_mode == 2)
            y = y + tt * row_gather<T>(a.b_indptr[row], a.b_indptr[row + 1], a.b_indices,
                                       a.b_vals, X, s, lane);
        spmm_epilogue<T, NT>(a, row, lane, y, s, tt, sub_c, sub_s, X, Y, Vout, Sub, red);
    }
    spmm_finish<T>(a, red);
}

// Column-blocked CSR form, for matrices with long rows over many columns
// (the Golub-Kahn operands A and A^T of a wide random matrix): there every
// non-zero gathers a random operand row, and rows drawn from a table larger
// than the L2 arrive at about half the rate of rows drawn from an
// L2-resident table (measured: 14-15 vs 27.7 G rows/s for 512 B rows). A
// wavefront owns kBlkRows rows and sweeps the columns in blocks of `cb`; the
// waves of a workgroup pass the blocks together (a barrier every `sync`
// blocks), the workgroups are as large as allowed (16 wavefronts, one per
// CU) and the grid holds only as many as fit on the chip, so all of them
// start the sweep together. The workgroups of an XCD still drift apart by a few
// blocks, so a block must take well under the 4 MB L2: on the 2^20 x 2^18
// Golub-Kahn case (8-wavefront workgroups) the L2 hit rate was 92 % with
// 256-column blocks and 13-14 % with 2048 or more; with 16-wavefront
// workgroups and a barrier every four blocks 1024 columns (512 KB of
// operand rows) is fastest, 1.7x the row kernel. Each row keeps a window of 64 (column, value) pairs in registers
// with a cursor; the pairs of one block are a prefix of what is left of the
// window (columns ascend), found with a ballot. Per row the products are
// added in CSR order with the same fma chain as row_gather, so the result
// equals spmm_kernel's bit for bit.
constexpr int kBlkRows = 8;

template <typename T, bool NT, int WPG>
__global__ __launch_bounds__(WPG * kWave) void csr_blocked_kernel(SpmmArgs<T> a,
                                                                  const T* __restrict__ X,
                                                                  T* __restrict__ Y,
                                                                  T* __restrict__ Vout,
                                                                  const T* __restrict__ Sub,
                                                                  int cb, int sync)
{
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const int64_t ngroups = (a.nrows + kBlkRows - 1) / kBlkRows;
    const Sweep sw = sweep_wpg<WPG>(ngroups, wave);
    const T s = a.xscale ? (T) a.xscale[lane] : T(1);
    const T tt = (T) a.t;
    const T sub_c = Sub ? (T) a.sub_coef[lane] : T(0);
    const T sub_s = Sub ? (T) a.sub_scale[lane] : T(0);
    const int64_t nb = (a.ncols + cb - 1) / cb;
    // Every wave of the grid runs the same number of rounds (a group index
    // past its XCD's span is an empty group), so the workgroup barrier after
    // each column block is reached by all of its waves: the waves of a
    // workgroup move from block to block together.
    const int64_t span = (ngroups + kXcds - 1) / kXcds;
    const int64_t rounds = (span + sw.stride - 1) / sw.stride;
    double red = 0.0;
    for (int64_t round = 0; round < rounds; ++round)
    {
        const int64_t gi = sw.first + round * sw.stride;
        const int64_t r0 = gi < sw.end ? gi * kBlkRows : a.nrows;
        int64_t cur[kBlkRows], end[kBlkRows];
        int wc[kBlkRows], wpos[kBlkRows], wcnt[kBlkRows];
        T wv[kBlkRows], acc[kBlkRows];
        auto load_window = [&](int k) {
            const int64_t rem = end[k] - cur[k];
            wcnt[k] = rem < kWave ? (int) rem : kWave;
            wpos[k] = 0;
            wc[k] = INT32_MAX;
            wv[k] = T(0);
            if (lane < wcnt[k])
            {
                wc[k] = __builtin_nontemporal_load(a.indices + cur[k] + lane);
                wv[k] = __builtin_nontemporal_load(a.vals + cur[k] + lane);
            }
        };
#pragma unroll
        for (int k = 0; k < kBlkRows; ++k)
        {
            const bool in = r0 + k < a.nrows;
            cur[k] = in ? a.indptr[r0 + k] : 0;
            end[k] = in ? a.indptr[r0 + k + 1] : 0;
            acc[k] = T(0);
            load_window(k);
        }
        for (int64_t b = 0; b < nb; ++b)
        {
            const int64_t cend = (b + 1) * (int64_t) cb;
#pragma unroll
            for (int k = 0; k < kBlkRows; ++k)
            {
                while (wcnt[k] > 0)
                {
                    const bool take = lane >= wpos[k] && lane < wcnt[k] && (int64_t) wc[k] < cend;
                    const int n = __popcll(__ballot(take));
                    int e = wpos[k];
                    const int e1 = e + n;
                    for (; e + 4 <= e1; e += 4)
                    {
                        const int c0 = readlane(wc[k], e);
                        const int c1 = readlane(wc[k], e + 1);
                        const int c2 = readlane(wc[k], e + 2);
                        const int c3 = readlane(wc[k], e + 3);
                        const T x0 = X[(int64_t) c0 * kBlock + lane];
                        const T x1 = X[(int64_t) c1 * kBlock + lane];
                        const T x2 = X[(int64_t) c2 * kBlock + lane];
                        const T x3 = X[(int64_t) c3 * kBlock + lane];
                        acc[k] = fma(readlane(wv[k], e), s * x0, acc[k]);
                        acc[k] = fma(readlane(wv[k], e + 1), s * x1, acc[k]);
                        acc[k] = fma(readlane(wv[k], e + 2), s * x2, acc[k]);
                        acc[k] = fma(readlane(wv[k], e + 3), s * x3, acc[k]);
                    }
                    for (; e < e1; ++e)
                    {
                        const int c = readlane(wc[k], e);
                        acc[k] = fma(readlane(wv[k], e), s * X[(int64_t) c * kBlock + lane],
                                     acc[k]);
                    }
                    wpos[k] = e1;
                    if (e1 < wcnt[k]) break;            // the rest is in later blocks
                    cur[k] += wcnt[k];                  // window used up: next one
                    load_window(k);
                }
            }
            // every `sync` blocks (b is the same in every wave): rows whose
            // entries per block vary (Poisson counts in A^T) then wait for
            // the workgroup's slowest wavefront once per `sync` blocks, and
            // the waves stay within `sync` blocks of each other
            if ((b + 1) % sync == 0 || b + 1 == nb) __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kBlkRows; ++k)
            if (r0 + k < a.nrows)
                spmm_epilogue<T, NT>(a, r0 + k, lane, acc[k], s, tt, sub_c, sub_s, X, Y, Vout,
                                     Sub, red);
    }
    if (a.red_mode != 0) wg_partial<WPG>(red, a.rb);
}

template <typename T, int R>
__device__ __forceinline__ void slabs_scalar(const T* __restrict__ X,
                                             const int32_t* __restrict__ tcols,
                                             const T* __restrict__ tvals, int64_t base, int cnt,
                                             T s, int lane, T (&acc)[R])
{
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
        const T xv = s * X[(int64_t) c * kBlock + lane];
        const T* v = vb + (int64_t) k * R;
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = fma(v[r], xv, acc[r]);
    }
}

// Row-tiled form with an in-order tile schedule. Tiles are handed out to
// wavefronts in increasing order by a per-XCD ticket counter (XCD = workgroup
// id mod 8, each owning a contiguous eighth of the tiles), so the tiles in
// flight on an XCD at any moment are consecutive and their gathered operand
// rows stay within that XCD's L2. (A static interleaved assignment lets fast
// and slow wavefronts drift apart; the set of tiles in flight then spreads
// over several sweeps and the operand rows are fetched from HBM several
// times.) Which wavefront handles which tile varies from run to run, so the
// reduction terms are written per tile (tile_partial[t][probe], rows of the
// tile added in order) and summed in tile order by tile_sum_kernel and
// total_kernel: the result does not depend on the schedule.
template <typename T, int R, bool NT>
__global__ __launch_bounds__(kThreads) void tiled_spmm_sched_kernel(
    SpmmArgs<T> a, const T* __restrict__ X, T* __restrict__ Y, T* __restrict__ Vout,
    const T* __restrict__ Sub, int64_t ntiles, const int64_t* __restrict__ tptr,
    const int32_t* __restrict__ tcols, const T* __restrict__ tvals)
{
    const int lane = threadIdx.x % kWave;
    // V groups of workgroups (id mod V); V = 8 = the XCDs unless the grid is
    // smaller, so that every range has workgroups
    const int V = gridDim.x < (unsigned int) kXcds ? (int) gridDim.x : kXcds;
    const int x = blockIdx.x % V;
    const int64_t span = (ntiles + V - 1) / V;
    const int64_t b = x * span;
    const int64_t e = b + span < ntiles ? b + span : ntiles;
    // wavefronts of the workgroups of this group
    const unsigned int waves = (unsigned int) ((gridDim.x - x + V - 1) / V) * kWavesPerWG;
    unsigned int* ticket = a.rb.tickets + 64 * x;
    unsigned int* done = ticket + 32;
    const T s = a.xscale ? (T) a.xscale[lane] : T(1);
    const T tt = (T) a.t;
    const T sub_c = Sub ? (T) a.sub_coef[lane] : T(0);
    const T sub_s = Sub ? (T) a.sub_scale[lane] : T(0);
    for (;;)
    {
        unsigned int k = 0;
        if (lane == 0) k = atomicAdd(ticket, 1u);
        const int64_t t = b + (int64_t) __builtin_amdgcn_readfirstlane(k);
        if (t >= e) break;
        T acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = T(0);
        const int64_t beg = tptr[t];
        const int64_t end = tptr[t + 1];
        for (int64_t base = beg; base < end; base += kWave)
        {
            const int64_t rem = end - base;
            const int cnt = rem < kWave ? (int) rem : kWave;
            slabs_scalar<T, R>(X, tcols, tvals, base, cnt, s, lane, acc);
        }
        double tred = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            const int64_t row = t * R + r;
            if (row < a.nrows)
                spmm_epilogue<T, NT>(a, row, lane, acc[r], s, tt, sub_c, sub_s, X, Y, Vout,
                                     Sub, tred);
        }
        if (a.red_mode) a.rb.tile_partial[t * kWave + lane] = tred;
    }
    // The last wavefront of this XCD to finish resets its counters for the
    // next launch: every wavefront has drawn its final ticket before it
    // counts itself done, so no ticket is drawn after the reset.
    unsigned int d = 0;
    if (lane == 0) d = atomicAdd(done, 1u);
    d = __builtin_amdgcn_readfirstlane(d);
    if (d == waves - 1 && lane == 0)
    {
        atomicExch(ticket, 0u);
        atomicExch(done, 0u);
    }
}

// ------------------------------------------------ matrix-core SpMM

// v_mfma_{f64,f32}_16x16x4: D[16][16] += A[16][4] B[4][16], one A and one B
// element per lane: A[l % 16][l / 16], B[l / 16][l % 16]. Lane l holds four
// elements of column l % 16 of D: rows 4 (l / 16) + m for the f32 form and
// rows l / 16 + 4 m for the f64 form, m = 0..3 (measured on gfx950).
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    static constexpr int kRowStride = 4;   // row = l / 16 + 4 m
    typedef double V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 op(double a, double b, V4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma16<float> {
    static constexpr int kRowStride = 1;   // row = 4 (l / 16) + m
    typedef float V4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V4 op(float a, float b, V4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename T>
struct Pair2Of;
template <>
struct Pair2Of<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair2Of<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};
template <typename T>
using Pair2 = typename Pair2Of<T>::type;

// Four consecutive values p[0..3] (16 B aligned): one 32 B row segment as
// two 16 B loads (FP64) or one (FP32).
__device__ __forceinline__ void load_row4(const double* __restrict__ p, double (&v)[4])
{
    typedef double D2 __attribute__((ext_vector_type(2)));
    const D2 lo = *reinterpret_cast<const D2*>(p);
    const D2 hi = *reinterpret_cast<const D2*>(p + 2);
    v[0] = lo.x;
    v[1] = lo.y;
    v[2] = hi.x;
    v[3] = hi.y;
}

__device__ __forceinline__ void load_row4(const float* __restrict__ p, float (&v)[4])
{
    typedef float F4 __attribute__((ext_vector_type(4)));
    const F4 w = *reinterpret_cast<const F4*>(p);
    v[0] = w.x;
    v[1] = w.y;
    v[2] = w.z;
    v[3] = w.w;
}

template <bool NT>
__device__ __forceinline__ void store_row4(double* __restrict__ p, const double (&v)[4])
{
    typedef double D2 __attribute__((ext_vector_type(2)));
    const D2 lo = {v[0], v[1]}, hi = {v[2], v[3]};
    if constexpr (NT)
    {
        __builtin_nontemporal_store(lo, reinterpret_cast<D2*>(p));
        __builtin_nontemporal_store(hi, reinterpret_cast<D2*>(p + 2));
    }
    else
    {
        *reinterpret_cast<D2*>(p) = lo;
        *reinterpret_cast<D2*>(p + 2) = hi;
    }
}

template <bool NT>
__device__ __forceinline__ void store_row4(float* __restrict__ p, const float (&v)[4])
{
    typedef float F4 __attribute__((ext_vector_type(4)));
    const F4 w = {v[0], v[1], v[2], v[3]};
    if constexpr (NT) __builtin_nontemporal_store(w, reinterpret_cast<F4*>(p));
    else *reinterpret_cast<F4*>(p) = w;
}

// The four probes of lane column i0 (0..15) of the MFMA kernels, one per
// 16-probe block q, and their accesses in a 64-probe row. FP64: probes
// 2 i0 + q for q < 2 and 30 + 2 i0 + q for q >= 2, so each 16 B access of
// lanes 0..15 covers 256 contiguous bytes of the 512 B row (two whole
// 128 B lines; the earlier probes 4 i0 + q took half of each of four lines
// per access). FP32: probes 4 i0 + q, one 16 B access covering the row.
template <typename T>
struct Lanes;
template <>
struct Lanes<double> {
    static __device__ __forceinline__ int probe(int i0, int q)
    {
        return q < 2 ? 2 * i0 + q : 30 + 2 * i0 + q;
    }
    static __device__ __forceinline__ void load(const double* __restrict__ row, int i0,
                                                double (&v)[4])
    {
        typedef double D2 __attribute__((ext_vector_type(2)));
        const D2 lo = *reinterpret_cast<const D2*>(row + 2 * i0);
        const D2 hi = *reinterpret_cast<const D2*>(row + 32 + 2 * i0);
        v[0] = lo.x;
        v[1] = lo.y;
        v[2] = hi.x;
        v[3] = hi.y;
    }
    template <bool NT>
    static __device__ __forceinline__ void store(double* __restrict__ row, int i0,
                                                 const double (&v)[4])
    {
        typedef double D2 __attribute__((ext_vector_type(2)));
        const D2 lo = {v[0], v[1]}, hi = {v[2], v[3]};
        D2* plo = reinterpret_cast<D2*>(row + 2 * i0);
        D2* phi = reinterpret_cast<D2*>(row + 32 + 2 * i0);
        if constexpr (NT)
        {
            __builtin_nontemporal_store(lo, plo);
            __builtin_nontemporal_store(hi, phi);
        }
        else
        {
            *plo = lo;
            *phi = hi;
        }
    }
};
template <>
struct Lanes<float> {
    static __device__ __forceinline__ int probe(int i0, int q) { return 4 * i0 + q; }
    static __device__ __forceinline__ void load(const float* __restrict__ row, int i0,
                                                float (&v)[4])
    {
        load_row4(row + 4 * i0, v);
    }
    template <bool NT>
    static __device__ __forceinline__ void store(float* __restrict__ row, int i0,
                                                 const float (&v)[4])
    {
        store_row4<NT>(row + 4 * i0, v);
    }
};

// spmm_epilogue for the four probes Lanes<T>::probe(i0, q) of one row, with
// 16 B loads and stores (the same arithmetic per probe).
template <typename T, bool NT>
__device__ __forceinline__ void spmm_epilogue4(const SpmmArgs<T>& a, int64_t row, int i0,
                                               const T (&yin)[4], const T (&s)[4], T tt,
                                               const T (&sub_c)[4], const T (&sub_s)[4],
                                               const T* __restrict__ X, T* __restrict__ Y,
                                               T* __restrict__ Vout, const T* __restrict__ Sub,
                                               double (&red)[4])
{
    const int64_t o = row * kBlock;
    const bool shift_row = a.shift_mode == 1 && row < a.shift_len;
    T v[4] = {T(0), T(0), T(0), T(0)}, y[4], sb[4];
    if (Vout || a.red_mode == 1 || shift_row)
    {
        T xv[4];
        Lanes<T>::load(X + o, i0, xv);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = s[q] * xv[q];
    }
    if (Sub) Lanes<T>::load(Sub + o, i0, sb);
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        y[q] = yin[q];
        // imate adds t x through subtract_scaled_vector(-t): y - (-t) x
        if (shift_row) y[q] = y[q] + tt * v[q];
        // the reduction term v . (A + tB) v is taken before the subtraction
        if (a.red_mode == 1) red[q] += (double) v[q] * (double) y[q];
        if (Sub) y[q] = y[q] - sub_c[q] * (sub_s[q] * sb[q]);
    }
    if (Vout) Lanes<T>::template store<NT>(Vout + o, i0, v);
    Lanes<T>::template store<NT>(Y + o, i0, y);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (a.red_mode == 2) red[q] += (double) y[q] * (double) y[q];
}

// Tiled SpMM on the matrix cores. A 16-row tile times the gathered operand
// rows of its slabs is a dense product (16 x slabs) (slabs x 64 probes):
// four slabs at a time, A = the 16 x 4 slab values (64 consecutive values of
// the slab array, one coalesced load), B = the four gathered operand rows,
// one 16-probe block per MFMA (four per group). The matrix values arrive
// through ordinary vector loads and are broadcast by the matrix core (no
// scalar loads, no readlane per value), so loads stay in flight: the column
// indices of up to 64 slabs are loaded once per chunk (lane l holds slab
// chunk + l) and broadcast with readlane, so a group's gathers need no
// dependent index load, and the loads of the next group of four slabs are
// issued before the MFMAs of the current one (two register sets). Tiles are
// handed out in order as in tiled_spmm_sched_kernel; the per-probe reduction
// term of a tile adds the four rows held by a lane, then the four row groups
// (lanes l, l ^ 16, l ^ 32, l ^ 48) in a fixed order.
template <typename T, bool NT>
__global__ __launch_bounds__(kThreads) void mfma_spmm_kernel(
    SpmmArgs<T> a, const T* __restrict__ X, T* __restrict__ Y, T* __restrict__ Vout,
    const T* __restrict__ Sub, int64_t ntiles, const int64_t* __restrict__ tptr,
    const int32_t* __restrict__ tcols, const T* __restrict__ tvals)
{
    using V4 = typename Mfma16<T>::V4;
    const int lane = threadIdx.x % kWave;
    const int i0 = lane % 16;
    const int kq = lane / 16;
    const int V = gridDim.x < (unsigned int) kXcds ? (int) gridDim.x : kXcds;
    const int x = blockIdx.x % V;
    const int64_t span = (ntiles + V - 1) / V;
    const int64_t b = x * span;
    const int64_t e = b + span < ntiles ? b + span : ntiles;
    const unsigned int waves = (unsigned int) ((gridDim.x - x + V - 1) / V) * kWavesPerWG;
    unsigned int* ticket = a.rb.tickets + 64 * x;
    unsigned int* done = ticket + 32;
    // block q of the four 16-probe MFMA blocks holds probe
    // Lanes<T>::probe(i0, q), so a lane's four B values of one gathered row
    // take two 16 B loads (FP64) that with lanes 0..15 read whole lines
    T sq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sq[q] = a.xscale ? (T) a.xscale[Lanes<T>::probe(i0, q)] : T(1);
    const T tt = (T) a.t;
    for (;;)
    {
        unsigned int k = 0;
        if (lane == 0) k = atomicAdd(ticket, 1u);
        const int64_t t = b + (int64_t) __builtin_amdgcn_readfirstlane(k);
        if (t >= e) break;
        V4 acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = V4{T(0), T(0), T(0), T(0)};
        const int64_t beg = tptr[t];
        const int64_t end = tptr[t + 1];
        for (int64_t base = beg; base < end; base += kWave)
        {
            const int64_t rem = end - base;
            const int cnt = rem < kWave ? (int) rem : kWave;   // multiple of 4
            const int cl = lane < cnt ? tcols[base + lane] : 0;
            auto col_of = [&](int g) {
                const int c0 = readlane(cl, 4 * g);
                const int c1 = readlane(cl, 4 * g + 1);
                const int c2 = readlane(cl, 4 * g + 2);
                const int c3 = readlane(cl, 4 * g + 3);
                return (int64_t) (kq == 0 ? c0 : (kq == 1 ? c1 : (kq == 2 ? c2 : c3)));
            };
            // pairs of MFMA groups (eight slabs): one access loads a lane's
            // A values of both groups
            const int np = cnt / 8;
            const Pair2<T>* __restrict__ ap2 =
                reinterpret_cast<const Pair2<T>*>(tvals + base * 16) + lane;
            Pair2<T> ap = ap2[0], an;
            T b0[4], b1[4];
            Lanes<T>::load(X + col_of(0) * kBlock, i0, b0);
            for (int P = 0; P < np; ++P)
            {
                Lanes<T>::load(X + col_of(2 * P + 1) * kBlock, i0, b1);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = Mfma16<T>::op(ap.x, sq[q] * b0[q], acc[q]);
                if (P + 1 < np)
                {
                    an = ap2[(P + 1) * kWave];
                    Lanes<T>::load(X + col_of(2 * P + 2) * kBlock, i0, b0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = Mfma16<T>::op(ap.y, sq[q] * b1[q], acc[q]);
                ap = an;
            }
        }
        double red[4] = {0.0, 0.0, 0.0, 0.0};
        T sub_c[4], sub_s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            sub_c[q] = Sub ? (T) a.sub_coef[Lanes<T>::probe(i0, q)] : T(0);
            sub_s[q] = Sub ? (T) a.sub_scale[Lanes<T>::probe(i0, q)] : T(0);
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            const int64_t row = t * 16 +
                                (Mfma16<T>::kRowStride == 4 ? kq + 4 * m : 4 * kq + m);
            const T y[4] = {acc[0][m], acc[1][m], acc[2][m], acc[3][m]};
            if (row < a.nrows)
                spmm_epilogue4<T, NT>(a, row, i0, y, sq, tt, sub_c, sub_s, X, Y, Vout, Sub,
                                      red);
        }
        if (a.red_mode)
        {
            double mine = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double r1 = red[q] + __shfl_xor(red[q], 16);
                const double r2 = r1 + __shfl_xor(r1, 32);
                if (q == kq) mine = r2;
            }
            a.rb.tile_partial[t * kWave + Lanes<T>::probe(i0, kq)] = mine;
        }
    }
    unsigned int d = 0;
    if (lane == 0) d = atomicAdd(done, 1u);
    d = __builtin_amdgcn_readfirstlane(d);
    if (d == waves - 1 && lane == 0)
    {
        atomicExch(ticket, 0u);
        atomicExch(done, 0u);
    }
}

// The matrix-core SpMM with S wavefronts per tile (S = 2 or 4): the waves
// of a team split the tile's slab list into S contiguous ranges of eight-slab
// pairs, so the tiles in flight on an XCD are 1 / S as many as with one wave
// per tile and the operand rows they gather take 1 / S of the L2 (measured at
// n = 2^20: HBM reads per launch 2.96 GB with one wave, 2.02 GB with two,
// 1.78 GB with four), while every slab value and every gathered row is still
// loaded once. Each wave writes the partial products of the row slices it
// does not finish to LDS; wave j of a team then adds the S partials of
// slices m = j (4 / S) .. (j + 1) (4 / S) - 1 in the fixed order of the
// waves and runs the epilogue on them, writing the reduction term of its
// rows as partial row t S + j (tile_sum_kernel adds them in order, so
// results are deterministic for a given S). The waves of a team meet at an
// LDS arrival counter rather than a workgroup barrier, so teams never wait
// for each other: wave 0 of a team takes the team's next tile from the
// XCD's in-order ticket one tile ahead and publishes it with its partials,
// and a second counter of waves done reading a tile's partials keeps a wave
// from overwriting them before its whole team has read them.
template <typename T, bool NT, int S>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
void mfma_spmm_split_kernel(
    SpmmArgs<T> a, const T* __restrict__ X, T* __restrict__ Y, T* __restrict__ Vout,
    const T* __restrict__ Sub, int64_t ntiles, const int64_t* __restrict__ tptr,
    const int32_t* __restrict__ tcols, const T* __restrict__ tvals)
{
    static_assert(S == 2 || S == 4, "two or four wavefronts per tile");
    using V4 = typename Mfma16<T>::V4;
    constexpr int kTeams = kWavesPerWG / S;
    constexpr int MS = 4 / S;   // row slices finished per wave
    // per team: for every slice m and every wave j not finishing it, the
    // four values q of each lane
    __shared__ T part[kTeams][4 * (S - 1)][4][kWave];
    __shared__ int64_t s_tile[kTeams][2];
    __shared__ unsigned int s_arrive[kTeams];
    __shared__ unsigned int s_read[kTeams];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int team = wave / S;
    const int j = wave % S;
    const int i0 = lane % 16;
    const int kq = lane / 16;
    const int V = gridDim.x < (unsigned int) kXcds ? (int) gridDim.x : kXcds;
    const int x = blockIdx.x % V;
    const int64_t span = (ntiles + V - 1) / V;
    const int64_t b = x * span;
    const int64_t e = b + span < ntiles ? b + span : ntiles;
    const unsigned int waves = (unsigned int) ((gridDim.x - x + V - 1) / V) * kWavesPerWG;
    unsigned int* ticket = a.rb.tickets + 64 * x;
    unsigned int* done = ticket + 32;
    T sq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sq[q] = a.xscale ? (T) a.xscale[Lanes<T>::probe(i0, q)] : T(1);
    const T tt = (T) a.t;
    auto slot = [](int m, int src) { return m * (S - 1) + (src < m / MS ? src : src - 1); };
    if (j == 0 && lane == 0)
    {
        s_arrive[team] = 0u;
        s_read[team] = 0u;
        s_tile[team][0] = b + (int64_t) atomicAdd(ticket, 1u);
    }
    __syncthreads();   // the only workgroup barrier: counters and first tiles
    for (unsigned int k = 0;; ++k)
    {
        const int par = (int) (k & 1u);
        const int64_t t = s_tile[team][par];
        if (t >= e) break;   // the same value in every wave of the team
        // wave 0 of the team asks for the team's next tile now; the answer
        // is published with its partials below
        unsigned int nxt = 0;
        if (j == 0 && lane == 0) nxt = atomicAdd(ticket, 1u);
        V4 acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = V4{T(0), T(0), T(0), T(0)};
        {
            const int64_t beg0 = tptr[t];
            const int64_t npt = (tptr[t + 1] - beg0) / 8;
            const int64_t beg = beg0 + 8 * ((npt * j) / S);
            const int64_t end = beg0 + 8 * ((npt * (j + 1)) / S);
            for (int64_t base = beg; base < end; base += kWave)
            {
                const int64_t rem = end - base;
                const int cnt = rem < kWave ? (int) rem : kWave;   // multiple of 8
                const int cl = lane < cnt ? tcols[base + lane] : 0;
                auto col_of = [&](int g) {
                    const int c0 = readlane(cl, 4 * g);
                    const int c1 = readlane(cl, 4 * g + 1);
                    const int c2 = readlane(cl, 4 * g + 2);
                    const int c3 = readlane(cl, 4 * g + 3);
                    return (int64_t) (kq == 0 ? c0 : (kq == 1 ? c1 : (kq == 2 ? c2 : c3)));
                };
                const int np = cnt / 8;
                const Pair2<T>* __restrict__ ap2 =
                    reinterpret_cast<const Pair2<T>*>(tvals + base * 16) + lane;
                Pair2<T> ap = ap2[0], an;
                T b0[4], b1[4];
                Lanes<T>::load(X + col_of(0) * kBlock, i0, b0);
                for (int P = 0; P < np; ++P)
                {
                    Lanes<T>::load(X + col_of(2 * P + 1) * kBlock, i0, b1);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc[q] = Mfma16<T>::op(ap.x, sq[q] * b0[q], acc[q]);
                    if (P + 1 < np)
                    {
                        an = ap2[(P + 1) * kWave];
                        Lanes<T>::load(X + col_of(2 * P + 2) * kBlock, i0, b0);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc[q] = Mfma16<T>::op(ap.y, sq[q] * b1[q], acc[q]);
                    ap = an;
                }
            }
        }
        // the team has read the partials of tile k - 1 (its waves run the
        // same code, so this wait ends)
        const unsigned int read_target = (unsigned int) S * k;
        while (__hip_atomic_load(&s_read[team], __ATOMIC_ACQUIRE,
                                 __HIP_MEMORY_SCOPE_WORKGROUP) < read_target)
            __builtin_amdgcn_s_sleep(1);
        // partial products of the slices other waves finish, the next tile,
        // then the arrival of this wave at tile k
#pragma unroll
        for (int m = 0; m < 4; ++m)
            if (m / MS != j)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q) part[team][slot(m, j)][q][lane] = acc[q][m];
            }
        if (j == 0 && lane == 0) s_tile[team][par ^ 1] = b + (int64_t) nxt;
        if (lane == 0)
            __hip_atomic_fetch_add(&s_arrive[team], 1u, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        // every wave of the team has written its partials of tile k (the
        // other waves of the team run the same code, so this wait ends)
        const unsigned int target = (unsigned int) S * (k + 1u);
        while (__hip_atomic_load(&s_arrive[team], __ATOMIC_ACQUIRE,
                                 __HIP_MEMORY_SCOPE_WORKGROUP) < target)
            __builtin_amdgcn_s_sleep(1);
        // the sums of this wave's slices, waves 0 .. S-1 in order (its own
        // partial in place), then the partials are released to the team
        T y[MS][4];
#pragma unroll
        for (int mm = 0; mm < MS; ++mm)
        {
            const int m = j * MS + mm;
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                T v = j == 0 ? acc[q][m] : part[team][slot(m, 0)][q][lane];
#pragma unroll
                for (int src = 1; src < S; ++src)
                    v += src == j ? acc[q][m] : part[team][slot(m, src)][q][lane];
                y[mm][q] = v;
            }
        }
        if (lane == 0)
            __hip_atomic_fetch_add(&s_read[team], 1u, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        double red[4] = {0.0, 0.0, 0.0, 0.0};
        T sub_c[4], sub_s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            sub_c[q] = Sub ? (T) a.sub_coef[Lanes<T>::probe(i0, q)] : T(0);
            sub_s[q] = Sub ? (T) a.sub_scale[Lanes<T>::probe(i0, q)] : T(0);
        }
#pragma unroll
        for (int mm = 0; mm < MS; ++mm)
        {
            const int m = j * MS + mm;
            const int64_t row = t * 16 +
                                (Mfma16<T>::kRowStride == 4 ? kq + 4 * m : 4 * kq + m);
            if (row < a.nrows)
                spmm_epilogue4<T, NT>(a, row, i0, y[mm], sq, tt, sub_c, sub_s, X, Y, Vout,
                                      Sub, red);
        }
        if (a.red_mode)
        {
            double mine = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double r1 = red[q] + __shfl_xor(red[q], 16);
                const double r2 = r1 + __shfl_xor(r1, 32);
                if (q == kq) mine = r2;
            }
            a.rb.tile_partial[(t * S + j) * kWave + Lanes<T>::probe(i0, kq)] = mine;
        }
    }
    unsigned int d = 0;
    if (lane == 0) d = atomicAdd(done, 1u);
    d = __builtin_amdgcn_readfirstlane(d);
    if (d == waves - 1 && lane == 0)
    {
        atomicExch(ticket, 0u);
        atomicExch(done, 0u);
    }
}

// Dense SpMM on the matrix cores: Y = A (s X) for a row-major A padded to
// rows of ld values (ld a multiple of 64, zero padding). A workgroup of 8
// wavefronts takes 128 rows (16 per wavefront, one MFMA row block) and
// sweeps the columns in chunks of 64: the 64 scaled operand rows of a chunk
// are staged in LDS once for the workgroup (double-buffered), and each
// wavefront streams its 16 x 64 slice of A from HBM, four consecutive values
// per lane; the next chunk's A slice and operand rows are loaded into
// registers before the current chunk's MFMAs and the rows written to LDS
// after them, so both loads overlap the matrix-core work. MFMA
// group (c, g) gives k-slot kq to column 16 c + 4 kq + g of the chunk (A
// and B alike, so the product is unchanged); the four 16-probe blocks hold
// probes Lanes<T>::probe(i, q) as in mfma_spmm_kernel, whose epilogue it shares. At 64
// probes the product does 16 flops per FP64 matrix byte: the bound is the
// matrix cores (78.6 TFLOP/s is 4.9 TB/s of A) as much as HBM.
constexpr int kDenseKC = 64;

template <typename T>
__device__ __forceinline__ void load_vals4(const T* __restrict__ p, T (&v)[4])
{
    load_row4(p, v);
}

// Per-probe reduction terms of a workgroup's rows (red[q] of lane (kq, i0)
// holds probe Lanes<T>::probe(i0, q) over that lane's rows) into its
// partial row.
template <typename T>
__device__ __forceinline__ void dense_red_partial(const SpmmArgs<T>& a, const double (&red)[4],
                                                  int wave, int kq, int i0, int lane)
{
    double mine = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const double r1 = red[q] + __shfl_xor(red[q], 16);
        const double r2 = r1 + __shfl_xor(r1, 32);
        if (q == kq) mine = r2;
    }
    // lane (kq, i0) holds probe Lanes<T>::probe(i0, kq): to lane = probe
    // through LDS, then the wavefronts in order
    __shared__ double pr[kWavesPerWG][kWave];
    pr[wave][Lanes<T>::probe(i0, kq)] = mine;
    __syncthreads();
    if (wave == 0)
    {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWavesPerWG; ++w) t += pr[w][lane];
        a.rb.partial[(int64_t) blockIdx.x * kWave + lane] = t;
    }
}

// Work item it of the dense SpMM: row block it / ksplit, column chunks of
// split it % ksplit. With ksplit > 1 (too few row blocks to fill the chip)
// each item stores its partial product in scratch[split][row][64] and
// dense_reduce_kernel adds the splits in order and finishes the rows.
template <typename T, bool NT>
__global__ __launch_bounds__(kThreads) void dense_spmm_kernel(SpmmArgs<T> a,
                                                              const T* __restrict__ X,
                                                              T* __restrict__ Y,
                                                              T* __restrict__ Vout,
                                                              const T* __restrict__ Sub,
                                                              int ksplit, T* __restrict__ scratch)
{
    using V4 = typename Mfma16<T>::V4;
    __shared__ T xs[2][kDenseKC][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const int i0 = lane % 16;
    const int kq = lane / 16;
    // staging: thread -> chunk row kk, probes pp .. pp + 7
    const int kk = (int) threadIdx.x >> 3;
    const int pp = ((int) threadIdx.x & 7) * 8;
    T ssc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ssc[u] = a.xscale ? (T) a.xscale[pp + u] : T(1);
    T sq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sq[q] = a.xscale ? (T) a.xscale[Lanes<T>::probe(i0, q)] : T(1);
    const T tt = (T) a.t;
    T sub_c[4], sub_s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        sub_c[q] = Sub ? (T) a.sub_coef[Lanes<T>::probe(i0, q)] : T(0);
        sub_s[q] = Sub ? (T) a.sub_scale[Lanes<T>::probe(i0, q)] : T(0);
    }
    const int64_t nchunk = (a.ncols + kDenseKC - 1) / kDenseKC;
    const int64_t cps = (nchunk + ksplit - 1) / ksplit;
    const int64_t rblocks = (a.nrows + 127) / 128;
    // operand rows of chunk ci: loaded into registers (xload), written to
    // LDS later (xstore), so that the loads are in flight during the
    // current chunk's MFMAs
    auto xload = [&](int64_t ci, T (&v)[8]) {
        const int64_t k = ci * kDenseKC + kk;
        if (k < a.ncols)
        {
            T w[4];
            load_row4(X + k * kBlock + pp, w);
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = w[u];
            load_row4(X + k * kBlock + pp + 4, w);
#pragma unroll
            for (int u = 0; u < 4; ++u) v[4 + u] = w[u];
        }
        else
        {
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = T(0);
        }
    };
    auto xstore = [&](const T (&v)[8], int buf) {
#pragma unroll
        for (int u = 0; u < 8; ++u) xs[buf][kk][pp + u] = ssc[u] * v[u];
    };
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t it = blockIdx.x; it < rblocks * ksplit; it += gridDim.x)
    {
        const int64_t rb = it / ksplit;
        const int sp = (int) (it - rb * ksplit);
        const int64_t c0 = sp * cps;
        const int64_t c1 = c0 + cps < nchunk ? c0 + cps : nchunk;
        const int64_t r0 = rb * 128 + wave * 16;
        const int64_t myrow = r0 + i0 < a.nrows ? r0 + i0 : a.nrows - 1;
        const T* __restrict__ arow = a.dense + myrow * a.ld + 4 * kq;
        V4 acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = V4{T(0), T(0), T(0), T(0)};
        T av[4][4];
        if (c0 < c1)
        {
            T xv[8];
            xload(c0, xv);
#pragma unroll
            for (int c = 0; c < 4; ++c) load_row4(arow + c0 * kDenseKC + 16 * c, av[c]);
            xstore(xv, 0);
        }
        __syncthreads();
        for (int64_t ci = c0; ci < c1; ++ci)
        {
            const int buf = (int) ((ci - c0) & 1);
            const bool more = ci + 1 < c1;
            // next chunk's A slice and operand rows, in flight during the
            // MFMAs below
            T an[4][4];
            T xv[8];
            if (more)
            {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    load_row4(arow + (ci + 1) * kDenseKC + 16 * c, an[c]);
                xload(ci + 1, xv);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
            {
#pragma unroll
                for (int g = 0; g < 4; ++g)
                {
                    T bv[4];
                    Lanes<T>::load(&xs[buf][16 * c + 4 * kq + g][0], i0, bv);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc[q] = Mfma16<T>::op(av[c][g], bv[q], acc[q]);
                }
            }
            if (more)
            {
                xstore(xv, buf ^ 1);
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int g = 0; g < 4; ++g) av[c][g] = an[c][g];
            }
            __syncthreads();
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            const int64_t row = r0 + (Mfma16<T>::kRowStride == 4 ? kq + 4 * m : 4 * kq + m);
            const T y[4] = {acc[0][m], acc[1][m], acc[2][m], acc[3][m]};
            if (row >= a.nrows) continue;
            if (ksplit > 1)
                Lanes<T>::template store<false>(
                    scratch + ((int64_t) sp * a.nrows + row) * kBlock, i0, y);
            else
                spmm_epilogue4<T, NT>(a, row, i0, y, sq, tt, sub_c, sub_s, X, Y, Vout, Sub,
                                      red);
        }
    }
    if (ksplit == 1 && a.red_mode) dense_red_partial<T>(a, red, wave, kq, i0, lane);
}

// Adds the ksplit partial products of each row in split order and finishes
// the rows as dense_spmm_kernel does without a split.
template <typename T, bool NT>
__global__ __launch_bounds__(kThreads) void dense_reduce_kernel(SpmmArgs<T> a,
                                                                const T* __restrict__ X,
                                                                T* __restrict__ Y,
                                                                T* __restrict__ Vout,
                                                                const T* __restrict__ Sub,
                                                                int ksplit,
                                                                const T* __restrict__ scratch)
{
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const int i0 = lane % 16;
    const int kq = lane / 16;
    T sq[4], sub_c[4], sub_s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        sq[q] = a.xscale ? (T) a.xscale[Lanes<T>::probe(i0, q)] : T(1);
        sub_c[q] = Sub ? (T) a.sub_coef[Lanes<T>::probe(i0, q)] : T(0);
        sub_s[q] = Sub ? (T) a.sub_scale[Lanes<T>::probe(i0, q)] : T(0);
    }
    const T tt = (T) a.t;
    double red[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t rb = blockIdx.x; rb * 128 < a.nrows; rb += gridDim.x)
    {
        const int64_t r0 = rb * 128 + wave * 16;
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            const int64_t row = r0 + (Mfma16<T>::kRowStride == 4 ? kq + 4 * m : 4 * kq + m);
            if (row >= a.nrows) continue;
            T y[4] = {T(0), T(0), T(0), T(0)};
            for (int sp = 0; sp < ksplit; ++sp)
            {
                T v[4];
                Lanes<T>::load(scratch + ((int64_t) sp * a.nrows + row) * kBlock, i0, v);
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] += v[q];
            }
            spmm_epilogue4<T, NT>(a, row, i0, y, sq, tt, sub_c, sub_s, X, Y, Vout, Sub, red);
        }
    }
    if (a.red_mode) dense_red_partial<T>(a, red, wave, kq, i0, lane);
}

// Column splits of the dense SpMM for an nrows x ncols operand: enough work
// items (row blocks x splits) for ~512 workgroups, at most one split per
// 64-column chunk and 64 in all.
int dense_ksplit(int64_t nrows, int64_t ncols)
{
    const int64_t rblocks = (nrows + 127) / 128;
    const int64_t nchunk = (ncols + kDenseKC - 1) / kDenseKC;
    int64_t k = (512 + rblocks - 1) / rblocks;
    if (k > nchunk) k = nchunk;
    if (k > 64) k = 64;
    return k < 1 ? 1 : (int) k;
}

// Sum of the per-tile partial rows in tile order, in two launches (the
// kernel boundary orders them; no fences): tile_sum_kernel gives workgroup g
// a contiguous range of tiles, its wavefronts interleaved with four
// accumulators each, combined in a fixed order into rb.partial[g];
// total_kernel (one workgroup) adds the rows of rb.partial in order and
// stores the total like the SpMM's own reduction.
constexpr int kTileSumGrid = 256;

__global__ __launch_bounds__(kThreads) void tile_sum_kernel(const double* __restrict__ tp,
                                                            int64_t ntiles,
                                                            double* __restrict__ partial)
{
    __shared__ double red[kWavesPerWG][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const int64_t G = gridDim.x;
    const int64_t per = (ntiles + G - 1) / G;
    const int64_t t0 = (int64_t) blockIdx.x * per;
    const int64_t t1 = t0 + per < ntiles ? t0 + per : ntiles;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int64_t t = t0 + wave;
    for (; t + 3 * kWavesPerWG < t1; t += 4 * kWavesPerWG)
    {
        v0 += tp[t * kWave + lane];
        v1 += tp[(t + kWavesPerWG) * kWave + lane];
        v2 += tp[(t + 2 * kWavesPerWG) * kWave + lane];
        v3 += tp[(t + 3 * kWavesPerWG) * kWave + lane];
    }
    for (; t < t1; t += kWavesPerWG) v0 += tp[t * kWave + lane];
    red[wave][lane] = (v0 + v1) + (v2 + v3);
    __syncthreads();
    if (wave == 0)
    {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWavesPerWG; ++w) s += red[w][lane];
        partial[(int64_t) blockIdx.x * kWave + lane] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(kTotalWaves * kWave) void total_kernel(SpmmArgs<T> a, int G)
{
    const double t = ordered_total(a.rb.partial, G);
    if (threadIdx.x >= kWave) return;
    spmm_store_total(a, t, threadIdx.x);
}

// ---------------------------------------------------------------- update

// Lanczos three-term update r = (w - alpha v_j) - beta_{j-1} v_{j-1}, in
// place over w, with the sum of squares of r (the order of the two
// subtractions follows c_lanczos_tridiagonalization.cpp:196-209). Each lane
// owns two consecutive probes of a row (one 16 B access per array for FP64),
// so a wavefront covers two rows per instruction; two such row pairs per
// iteration keep six loads in flight.
struct UpdateArgs {
    int64_t nrows;
    const double* alpha;   // [64]
    const double* beta;    // [64] beta_{j-1} (nullptr at j = 0)
    // per-probe scales of Vj / Vjm1 (nullptr = 1): v = scale * V, rounded
    // in T exactly as the SpMM epilogue rounds the basis vector it stores,
    // so reading an unnormalised residual with its scale gives the same bits
    const double* vscale;
    const double* uscale;
    int do_norm;           // 1: reduce |r|^2 and finish beta_j
    RedBuf rb;
    double* beta_out;
    double* scale_out;
    int32_t* sizes;
    int j;
    double tol_sqrt_n;
};

template <typename T>
struct Pair;
template <>
struct Pair<double> {
    using V = double2;
};
template <>
struct Pair<float> {
    using V = float2;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void lanczos_update_kernel(
    UpdateArgs a, T* __restrict__ R, const T* __restrict__ Vj, const T* __restrict__ Vjm1)
{
    using V2 = typename Pair<T>::V;
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    // unit = pair of rows; lane l covers row 2u + (l >= 32), probes 2(l%32), +1
    const int64_t npairs = (a.nrows + 1) / 2;
    const Sweep sw = sweep_linear(npairs, wave);
    const int half = lane >> 5;
    const int p0 = 2 * (lane & 31);
    const double al0 = a.alpha[p0], al1 = a.alpha[p0 + 1];
    const double be0 = Vjm1 ? a.beta[p0] : 0.0, be1 = Vjm1 ? a.beta[p0 + 1] : 0.0;
    const T sv0 = a.vscale ? (T) a.vscale[p0] : T(1);
    const T sv1 = a.vscale ? (T) a.vscale[p0 + 1] : T(1);
    const T su0 = (Vjm1 && a.uscale) ? (T) a.uscale[p0] : T(1);
    const T su1 = (Vjm1 && a.uscale) ? (T) a.uscale[p0 + 1] : T(1);
    double red0 = 0.0, red1 = 0.0;
    V2* __restrict__ R2 = reinterpret_cast<V2*>(R);
    const V2* __restrict__ Vj2 = reinterpret_cast<const V2*>(Vj);
    const V2* __restrict__ Vm2 = reinterpret_cast<const V2*>(Vjm1);
    for (int64_t u = sw.first; u < sw.end; u += 2 * sw.stride)
    {
        const int64_t u2 = u + sw.stride;
        const int64_t ra = 2 * u + half;
        const int64_t rb = 2 * u2 + half;
        const bool oka = ra < a.nrows;
        const bool okb = u2 < sw.end && rb < a.nrows;
        const int64_t oa = (ra * kBlock + p0) / 2;
        const int64_t ob = (rb * kBlock + p0) / 2;
        V2 wa, va, ua, wb, vb, ub;
        if (oka)
        {
            wa = R2[oa];
            va = Vj2[oa];
            if (Vjm1) ua = Vm2[oa];
        }
        if (okb)
        {
            wb = R2[ob];
            vb = Vj2[ob];
            if (Vjm1) ub = Vm2[ob];
        }
        if (oka)
        {
            double r0 = (double) wa.x - al0 * (double) (sv0 * va.x);
            double r1 = (double) wa.y - al1 * (double) (sv1 * va.y);
            if (Vjm1)
            {
                r0 = r0 - be0 * (double) (su0 * ua.x);
                r1 = r1 - be1 * (double) (su1 * ua.y);
            }
            V2 o;
            o.x = (T) r0;
            o.y = (T) r1;
            R2[oa] = o;
            red0 += (double) o.x * (double) o.x;
            red1 += (double) o.y * (double) o.y;
        }
        if (okb)
        {
            double r0 = (double) wb.x - al0 * (double) (sv0 * vb.x);
            double r1 = (double) wb.y - al1 * (double) (sv1 * vb.y);
            if (Vjm1)
            {
                r0 = r0 - be0 * (double) (su0 * ub.x);
                r1 = r1 - be1 * (double) (su1 * ub.y);
            }
            V2 o;
            o.x = (T) r0;
            o.y = (T) r1;
            R2[ob] = o;
            red0 += (double) o.x * (double) o.x;
            red1 += (double) o.y * (double) o.y;
        }
    }
    if (!a.do_norm) return;
    // lanes l and l + 32 hold the same probes: move to lane-per-probe form
    __shared__ double ex[kWavesPerWG][2][kWave];
    ex[wave][0][lane] = red0;
    ex[wave][1][lane] = red1;
    __syncthreads();
    // probe q = lane: pair slot (q % 2), owner lanes q/2 and q/2 + 32
    const int q = lane;
    const double mine = ex[wave][q & 1][q >> 1] + ex[wave][q & 1][(q >> 1) + 32];
    wg_partial(mine, a.rb);
}

// Finishes a norm from the G partial rows of the kernel before: beta =
// sqrt(sum) and the breakdown test.
__global__ __launch_bounds__(kTotalWaves * kWave) void norm_total_kernel(
    const double* __restrict__ partial, int G, int j, double tol_sqrt_n,
    double* __restrict__ beta_out, double* __restrict__ scale_out,
    int32_t* __restrict__ sizes)
{
    const double t = ordered_total(partial, G);
    if (threadIdx.x >= kWave) return;
    finish_norm(t, threadIdx.x, j, tol_sqrt_n, beta_out, scale_out, sizes);
}

// ------------------------------------------------ Gram-Schmidt (CGS2)

// Classical Gram-Schmidt, twice, against the K newest basis vectors of each
// probe (basis layout [slot][row][64]; vector k of a pass is slot last - k):
// h1 = V^T r, r -= V h1, h2 = V^T r, r -= V h2, |r|. Every probe has its own
// basis, so each product is a GEMV per probe (0.5 flop per byte): the
// kernels are HBM streams, written for loads in flight: a wavefront takes
// one row per iteration and issues the loads of kGsChunk basis vectors at
// once.
constexpr int kGsChunk = 32;     // basis vectors per pass over the rows

// Element stride between the stored basis vectors of a re-orthogonalised
// recurrence. (Padding it away from a power of two was measured and made no
// difference to the Gram-Schmidt kernels.)
size_t slot_stride(int64_t rows) { return (size_t) rows * kBlock; }

struct GsArgs {
    int64_t nrows;
    int64_t slot;            // element stride between basis vectors
    int last;                // slot of the newest vector; vector k is slot last - k
    int K;
    const double* h;         // [K][64] coefficients (gs_sub_kernel)
    double* partial_k;       // [G][K][64] (gs_dots_kernel)
    int do_norm;             // gs_sub_kernel: partial rows of |r|^2 in rb
    RedBuf rb;
};

// The row loops of the two kernels for a chunk of kn <= C vectors starting at
// `base` (vector g at base - g * stride). Vectors past the last are loaded
// again as the last (a cache hit), so every load is unconditional and all C
// are in flight at once; chunks are cut to C = 8, 16 or 32 to keep the
// repeats few.
template <typename T, int C>
__device__ __forceinline__ void dots_rows(const T* __restrict__ base, const T* __restrict__ R,
                                          int64_t stride, int kn, const Sweep& sw, int lane,
                                          double (&acc)[kGsChunk])
{
    for (int64_t row = sw.first; row < sw.end; row += sw.stride)
    {
        const int64_t o = row * kBlock + lane;
        T v[C];
#pragma unroll
        for (int g = 0; g < C; ++g) v[g] = base[o - (int64_t) (g < kn ? g : kn - 1) * stride];
        const double r = (double) R[o];
#pragma unroll
        for (int g = 0; g < C; ++g) acc[g] += r * (double) v[g];
    }
}

// r -= sum_g hs[g] v_g (hs is zero past the last vector: r - 0 * v = r).
template <typename T, int C>
__device__ __forceinline__ double sub_rows(const T* __restrict__ base, T* __restrict__ R,
                                           int64_t stride, int kn, const Sweep& sw, int lane,
                                           const double (*hs)[kWave], bool norm)
{
    double red = 0.0;
    for (int64_t row = sw.first; row < sw.end; row += sw.stride)
    {
        const int64_t o = row * kBlock + lane;
        T v[C];
#pragma unroll
        for (int g = 0; g < C; ++g) v[g] = base[o - (int64_t) (g < kn ? g : kn - 1) * stride];
        double r = (double) R[o];
#pragma unroll
        for (int g = 0; g < C; ++g) r -= hs[g][lane] * (double) v[g];
        const T rt = (T) r;
        R[o] = rt;
        if (norm) red += (double) rt * (double) rt;
    }
    return red;
}

// partial_k[g][k] = sum over the rows of workgroup g of basis_k * r.
template <typename T>
__global__ __launch_bounds__(kThreads) void gs_dots_kernel(GsArgs a, const T* __restrict__ R,
                                                           const T* __restrict__ basis)
{
    __shared__ double red[kWavesPerWG][8][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const Sweep sw = sweep_linear(a.nrows, wave);
    const int64_t stride = a.slot;
    for (int k0 = 0; k0 < a.K; k0 += kGsChunk)
    {
        const int kn = (a.K - k0) < kGsChunk ? (a.K - k0) : kGsChunk;
        const T* __restrict__ base = basis + (int64_t) (a.last - k0) * stride;
        double acc[kGsChunk];
#pragma unroll
        for (int g = 0; g < kGsChunk; ++g) acc[g] = 0.0;
        if (kn > 16) dots_rows<T, 32>(base, R, stride, kn, sw, lane, acc);
        else if (kn > 8) dots_rows<T, 16>(base, R, stride, kn, sw, lane, acc);
        else dots_rows<T, 8>(base, R, stride, kn, sw, lane, acc);
        // workgroup sums, eight vectors at a time through LDS
#pragma unroll
        for (int g0 = 0; g0 < kGsChunk; g0 += 8)
        {
            if (g0 >= kn) break;
#pragma unroll
            for (int g = 0; g < 8; ++g) red[wave][g][lane] = acc[g0 + g];
            __syncthreads();
            if (wave < 8 && g0 + wave < kn)
            {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < kWavesPerWG; ++w) t += red[w][wave][lane];
                a.partial_k[((int64_t) blockIdx.x * a.K + k0 + g0 + wave) * kWave + lane] = t;
            }
            __syncthreads();
        }
    }
}

// r -= sum_k h_k basis_k (in k order), optionally with the partial rows of
// |r|^2 for norm_total_kernel.
template <typename T>
__global__ __launch_bounds__(kThreads) void gs_sub_kernel(GsArgs a, T* __restrict__ R,
                                                          const T* __restrict__ basis)
{
    __shared__ double hs[kGsChunk][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    const Sweep sw = sweep_linear(a.nrows, wave);
    const int64_t stride = a.slot;
    double red = 0.0;
    if (a.K == 0)
    {
        // norm only
        for (int64_t row = sw.first; row < sw.end; row += sw.stride)
        {
            const double r = (double) R[row * kBlock + lane];
            red += r * r;
        }
    }
    for (int k0 = 0; k0 < a.K; k0 += kGsChunk)
    {
        const int kn = (a.K - k0) < kGsChunk ? (a.K - k0) : kGsChunk;
        const bool norm = a.do_norm && k0 + kGsChunk >= a.K;
        __syncthreads();
        for (int i = threadIdx.x; i < kGsChunk * kWave; i += blockDim.x)
            hs[i / kWave][i % kWave] = i < kn * kWave ? a.h[(int64_t) k0 * kWave + i] : 0.0;
        __syncthreads();
        const T* __restrict__ base = basis + (int64_t) (a.last - k0) * stride;
        if (kn > 16) red = sub_rows<T, 32>(base, R, stride, kn, sw, lane, hs, norm);
        else if (kn > 8) red = sub_rows<T, 16>(base, R, stride, kn, sw, lane, hs, norm);
        else red = sub_rows<T, 8>(base, R, stride, kn, sw, lane, hs, norm);
    }
    if (a.do_norm) wg_partial(red, a.rb);
}

// The first subtraction and the second product of CGS2 in one pass over the
// basis: per row r' = r - sum_k h1_k v_k (k order, all K <= C vectors held
// in registers), stored, and partial_k[g][k] = sum over the rows of
// workgroup g of v_k r'. Per row this is the arithmetic of gs_sub_kernel
// followed by gs_dots_kernel (for K > kGsChunk without the rounding of r to
// T between chunks), with one read of the basis instead of two.
// C <= 32 runs in 512-thread workgroups (two wavefronts per SIMD: at C =
// 32 the values, the sums and the addresses take 252 VGPRs); C = 64 in
// 256-thread workgroups, one wavefront per SIMD with up to 512 VGPRs.
template <typename T, int C, int WPG>
__global__ __launch_bounds__(WPG * kWave) void gs_subdots_kernel(GsArgs a, T* __restrict__ R,
                                                                 const T* __restrict__ basis)
{
    __shared__ double hs[C][kWave];
    __shared__ double red[WPG][8][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = wave_id();
    Sweep sw;
    sw.first = (int64_t) blockIdx.x * WPG + wave;
    sw.stride = (int64_t) gridDim.x * WPG;
    sw.end = a.nrows;
    const int64_t stride = a.slot;
    const int kn = a.K;
    for (int i = threadIdx.x; i < C * kWave; i += blockDim.x)
        hs[i / kWave][i % kWave] = i < kn * kWave ? a.h[i] : 0.0;
    __syncthreads();
    const T* __restrict__ base = basis + (int64_t) a.last * stride;
    double acc[C];
#pragma unroll
    for (int g = 0; g < C; ++g) acc[g] = 0.0;
#pragma unroll 1
    for (int64_t row = sw.first; row < sw.end; row += sw.stride)
    {
        // re-read hs from LDS every row rather than hold C coefficients in
        // registers beside the C values and C sums
        asm volatile("" ::: "memory");
        const int64_t o = row * kBlock + lane;
        T v[C];
#pragma unroll
        for (int g = 0; g < C; ++g) v[g] = base[o - (int64_t) (g < kn ? g : kn - 1) * stride];
        double r = (double) R[o];
#pragma unroll
        for (int g = 0; g < C; ++g) r -= hs[g][lane] * (double) v[g];
        const T rt = (T) r;
        R[o] = rt;
        const double rd = (double) rt;
#pragma unroll
        for (int g = 0; g < C; ++g) acc[g] += rd * (double) v[g];
    }
    // workgroup sums, eight vectors at a time through LDS
#pragma unroll
    for (int g0 = 0; g0 < C; g0 += 8)
    {
        if (g0 >= kn) break;
#pragma unroll
        for (int g = 0; g < 8; ++g) red[wave][g][lane] = acc[g0 + g];
        __syncthreads();
        for (int q = wave; q < 8; q += WPG)
            if (g0 + q < kn)
            {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < WPG; ++w) t += red[w][q][lane];
                a.partial_k[((int64_t) blockIdx.x * a.K + g0 + q) * kWave + lane] = t;
            }
        __syncthreads();
    }
}

// out[k][p] = sum_g partial[g][k][p] in a fixed order; one workgroup per k.
__global__ __launch_bounds__(kTotalWaves * kWave) void reduce_partials_kernel(
    const double* __restrict__ partial, int G, int K, double* __restrict__ out)
{
    const int k = blockIdx.x;
    const double t = ordered_total(partial + (int64_t) k * kWave, G, (int64_t) K * kWave);
    if (threadIdx.x < kWave) out[k * kWave + threadIdx.x] = t;
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
__global__ __launch_bounds__(256) void copy_block_kernel(const T* __restrict__ in,
                                                         T* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// ---------------------------------------------------------------- host side

// Grid of a sweep over `units`: at most kMaxGrid workgroups (4 per CU at 8
// wavefronts each, i.e. one resident wave of workgroups), at least one
// wavefront-unit each, a multiple of 8 (one share per XCD).
int max_grid()
{
    static int g = [] {
        const char* e = std::getenv("SLQ_GRID");
        int v = e ? std::atoi(e) : kGrid;
        if (v < kXcds || v > kMaxGrid) v = kMaxGrid;
        return (v / kXcds) * kXcds;
    }();
    return g;
}

}  // namespace

int grid_for(int64_t units, int cap)
{
    // reduction buffers hold kMaxGrid partial rows: never exceed it
    int c = cap > 0 ? ((cap < kXcds ? kXcds : cap) / kXcds) * kXcds : max_grid();
    if (c > kMaxGrid) c = kMaxGrid;
    int64_t g = (units + kWavesPerWG - 1) / kWavesPerWG;
    if (g > c) g = c;
    g = ((g + kXcds - 1) / kXcds) * kXcds;
    if (g < kXcds) g = kXcds;
    return (int) g;
}

namespace {

// Grid of the streaming kernels (update, Gram-Schmidt): by default 256
// workgroups, one per CU, for the Gram-Schmidt kernels (measured to reach
// the copy-kernel bandwidth) and 512 for the update kernel (0.405 ms against
// 0.417 at 256 and 0.426 at 1024 for n = 2^20).
int stream_grid(int64_t units, int cap = 0, int fallback = 256)
{
    if (cap <= 0)
    {
        const char* e = std::getenv("SLQ_STREAM_GRID");
        cap = e ? std::atoi(e) : fallback;
    }
    return grid_for(units, cap);
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

// Wavefronts per tile of the matrix-core SpMM: SLQ_SPMM_SPLIT = 1 (default),
// 2 or 4 (mfma_spmm_split_kernel), unless the handle sets its own.
int spmm_split_default()
{
    static const int v = [] {
        const char* e = std::getenv("SLQ_SPMM_SPLIT");
        const int s = e ? std::atoi(e) : 1;
        return (s == 2 || s == 4) ? s : 1;
    }();
    return v;
}

// The handle's setting (slq_op_set_launch) or the default.
int spmm_split(const slq_op* op)
{
    return op->spmm_split > 0 ? op->spmm_split : spmm_split_default();
}

// Tiles of the largest tiled operand a recurrence on `op` may multiply by
// (A and, for Golub-Kahn, A^T), times the partial rows per tile of the
// matrix-core SpMM: the length of the per-tile partial rows.
int64_t sched_tiles(const slq_op* op)
{
    const MatrixStore& A = *op->A;
    if (A.dense)
    {
        // rows of 64 values of the dense SpMM's split scratch, either way
        const int64_t nr = A.csr.nrows, nc = A.csr.ncols;
        const int64_t a = dense_ksplit(nr, nc) > 1 ? dense_ksplit(nr, nc) * nr : 0;
        const int64_t b = dense_ksplit(nc, nr) > 1 ? dense_ksplit(nc, nr) * nc : 0;
        return a > b ? a : b;
    }
    if (A.tiles.R == 0) return 0;
    const int64_t R = A.tiles.R;
    const int64_t a = (A.csr.nrows + R - 1) / R, b = (A.csr.ncols + R - 1) / R;
    return (a > b ? a : b) * (R == 16 ? spmm_split(op) : 1);
}

RedBuf take_redbuf(Bump& bump, int64_t tiles)
{
    RedBuf rb;
    rb.partial = bump.take<double>((size_t) kMaxGrid * kBlock);
    rb.tickets = bump.take<unsigned int>(kTicketWords);
    rb.tile_partial = tiles > 0 ? bump.take<double>((size_t) tiles * kBlock) : nullptr;
    if (tiles > 0 && !rb.tile_partial) rb.partial = nullptr;   // reported as too small
    return rb;
}

bool redbuf_ok(const RedBuf& rb)
{
    return rb.partial && rb.tickets;
}

void zero_counters(hipStream_t s, const RedBuf& rb)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3(kTicketWords / 64), dim3(64), 0, s, rb.tickets,
                       kTicketWords);
}

size_t redbuf_bytes(int64_t tiles)
{
    auto r = [](size_t b) { return (b + 255) & ~(size_t) 255; };
    return r(sizeof(double) * kMaxGrid * kBlock) + r(sizeof(unsigned int) * kTicketWords) +
           (tiles > 0 ? r(sizeof(double) * (size_t) tiles * kBlock) : 0);
}

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
    if (A.dense)
    {
        const bool cvt = f32 && A.csr.dtype == SLQ_F64;
        const void* d = trans ? (cvt ? (const void*) A.dn.trans_f32 : A.dn.trans)
                              : (cvt ? (const void*) A.dn.data_f32 : A.dn.data);
        s.dense = (const T*) d;
        s.ld = trans ? A.dn.trans_ld : A.dn.ld;
        s.nrows = trans ? A.csr.ncols : A.csr.nrows;
        s.ncols = trans ? A.csr.nrows : A.csr.ncols;
        return;
    }
    const DevCSR& c = trans ? A.trans : A.csr;
    s.indptr = c.indptr;
    s.indices = c.indices;
    s.vals = values_of<T>(A, trans, f32);
    s.nrows = c.nrows;
    s.ncols = c.ncols;
    s.sorted = c.sorted ? 1 : 0;
    s.avg_row_nnz = c.nrows > 0 ? (double) c.nnz / (double) c.nrows : 0.0;
    if (B)
    {
        const DevCSR& cb = trans ? B->trans : B->csr;
        s.b_indptr = cb.indptr;
        s.b_indices = cb.indices;
        s.b_vals = values_of<T>(*B, trans, f32);
    }
}

int nt_default()
{
    static int v = [] {
        const char* e = std::getenv("SLQ_NT_STORE");
        return e ? std::atoi(e) : 1;
    }();
    return v;
}

// Column block of the blocked CSR SpMM, or 0 for the row-per-wavefront
// kernel: blocked when the rows are sorted, long (>= 128 non-zeros on
// average) and the gathered operand is well beyond the L2 (> 16 MiB), and
// no second (B) stream. SLQ_CSR_BLOCK_COLS sets the block (0 disables).
template <typename T>
int csr_blocked_cols(const SpmmArgs<T>& s)
{
    static int cb = [] {
        const char* e = std::getenv("SLQ_CSR_BLOCK_COLS");
        return e ? std::atoi(e) : 1024;
    }();
    if (cb <= 0 || !s.sorted || s.shift_mode == 2 || s.nrows <= 0) return 0;
    if ((double) s.ncols * kBlock * sizeof(T) <= 16.0 * (1 << 20)) return 0;
    if (s.avg_row_nnz < 128.0) return 0;
    return cb;
}

template <typename T>
SpmmArgs<T> spmm_defaults()
{
    SpmmArgs<T> s;
    std::memset(&s, 0, sizeof(s));
    s.nt_store = nt_default();
    return s;
}

// Tiled form of the operand matrix for this launch, or nullptr (CSR path):
// only for A and A + tI (the general A + tB keeps the two CSR streams).
template <typename T>
const DevTiles* tiles_for(slq_op* op, bool trans)
{
    if (op->kind == OP_AFFINE_B || op->tile_mode == 1) return nullptr;
    const DevTiles& t = trans ? op->A->trans_tiles : op->A->tiles;
    if (t.R == 0) return nullptr;
    if (sizeof(T) == 4 && op->A->csr.dtype == SLQ_F64 && !t.vals_f32) return nullptr;
    return &t;
}

// Values of A + tB in the tile layout of the union pattern: out[i] =
// a[src_a[i]] + t b[src_b[i]] (absent terms zero), rounded to D.
template <typename S, typename D>
__global__ __launch_bounds__(256) void pair_values_kernel(const int32_t* __restrict__ src_a,
                                                          const int32_t* __restrict__ src_b,
                                                          const S* __restrict__ a,
                                                          const S* __restrict__ b, S t,
                                                          D* __restrict__ out, int64_t count)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        const int32_t ka = src_a[i], kb = src_b[i];
        const S va = ka >= 0 ? a[ka] : S(0);
        const S vb = kb >= 0 ? b[kb] : S(0);
        out[i] = (D) (va + t * vb);
    }
}

// The tiled A + tB path is on unless SLQ_PAIR_TILES=0 (read per call, so a
// process can time both paths).
bool pair_tiles_enabled()
{
    const char* e = std::getenv("SLQ_PAIR_TILES");
    return !(e && std::atoi(e) == 0);
}

// Tiled form of A + tB (with `trans`, of A^T + tB^T) at t for SpMMs of
// value type T, its values gathered on the operator's stream; *out =
// nullptr keeps the two-stream CSR SpMM.
template <typename T>
int pair_tiles_at(slq_op* op, double t, const DevTiles** out, bool trans = false)
{
    *out = nullptr;
    PairTiles& ab = trans ? op->ab_t : op->ab;
    if (op->kind != OP_AFFINE_B || ab.tiles.R == 0 || !pair_tiles_enabled() ||
        op->tile_mode == 1)
        return SLQ_OK;
    const bool a64 = op->A->csr.dtype == SLQ_F64;
    const void* da = trans ? op->A->trans.data : op->A->csr.data;
    const void* db = trans ? op->B->trans.data : op->B->csr.data;
    if (sizeof(T) == 4 && a64 && !ab.tiles.vals_f32) return SLQ_OK;
    const int64_t count = ab.tiles.nslabs * ab.tiles.R;
    const int64_t g = (count + 255) / 256;
    const dim3 grid((unsigned) (g < 16384 ? g : 16384));
    if (a64 && sizeof(T) == 4)
        hipLaunchKernelGGL((pair_values_kernel<double, float>), grid, dim3(256), 0, op->stream,
                           ab.src_a, ab.src_b, (const double*) da, (const double*) db, t,
                           ab.tiles.vals_f32, count);
    else if (a64)
        hipLaunchKernelGGL((pair_values_kernel<double, double>), grid, dim3(256), 0,
                           op->stream, ab.src_a, ab.src_b, (const double*) da,
                           (const double*) db, t, (double*) ab.tiles.vals, count);
    else
        hipLaunchKernelGGL((pair_values_kernel<float, float>), grid, dim3(256), 0, op->stream,
                           ab.src_a, ab.src_b, (const float*) da, (const float*) db,
                           (float) t, (float*) ab.tiles.vals, count);
    SLQ_CHECK(hipGetLastError());
    *out = &ab.tiles;
    return SLQ_OK;
}

// Operand tiles of an SpMM at t (of the transpose with `trans`): the pair
// tiles for A + tB, else tiles_for. *shift_mode is the epilogue's shift
// (none when the pair tiles already hold A + tB).
template <typename T>
int operand_tiles(slq_op* op, double t, const DevTiles** out, int* shift_mode,
                  bool trans = false)
{
    *shift_mode = op->kind == OP_AFFINE_I ? 1 : (op->kind == OP_AFFINE_B ? 2 : 0);
    if (op->kind != OP_AFFINE_B)
    {
        *out = tiles_for<T>(op, trans);
        return SLQ_OK;
    }
    SLQ_TRY(pair_tiles_at<T>(op, t, out, trans));
    if (*out) *shift_mode = 0;
    return SLQ_OK;
}

// Buffers of one SpMM launch.
template <typename T>
struct SpmmIO {
    const T* X = nullptr;
    T* Y = nullptr;
    T* Vout = nullptr;
    const T* Sub = nullptr;
};

template <typename T, bool NT>
void launch_spmm_nt(slq_op* op, const SpmmArgs<T>& s, const SpmmIO<T>& io, const DevTiles* tl)
{
    if (s.dense)
    {
        const int64_t rblocks = (s.nrows + 127) / 128;
        const int ks = dense_ksplit(s.nrows, s.ncols);
        T* scratch = (T*) s.rb.tile_partial;
        // without a split each workgroup writes a partial row (kMaxGrid rows)
        const int64_t items = rblocks * ks;
        const int64_t cap = ks > 1 ? 2048 : kMaxGrid;
        const int grid = (int) (items < cap ? items : cap);
        hipLaunchKernelGGL((dense_spmm_kernel<T, NT>), dim3(grid), dim3(kThreads), 0,
                           op->stream, s, io.X, io.Y, io.Vout, io.Sub, ks, scratch);
        int rg = grid;
        if (ks > 1)
        {
            rg = (int) (rblocks < kMaxGrid ? rblocks : kMaxGrid);
            hipLaunchKernelGGL((dense_reduce_kernel<T, NT>), dim3(rg), dim3(kThreads), 0,
                               op->stream, s, io.X, io.Y, io.Vout, io.Sub, ks,
                               (const T*) scratch);
        }
        if (s.red_mode)
            hipLaunchKernelGGL((total_kernel<T>), dim3(1), dim3(kTotalWaves * kWave), 0,
                               op->stream, s, rg);
        return;
    }
    if (tl)
    {
        const T* tv = (sizeof(T) == 4 && op->A->csr.dtype == SLQ_F64)
                          ? (const T*) tl->vals_f32 : (const T*) tl->vals;
        const int grid = grid_for(tl->ntiles, op->spmm_grid);
        // tiles are handed out in order from the XCD's ticket counter (16-row
        // tiles on the matrix cores, 8-row tiles on the vector ALUs);
        // launch_spmm has checked the counters and partial rows
        {
            const int split = tl->R == 16 ? spmm_split(op) : 1;
            if (tl->R == 16 && split == 2)
                hipLaunchKernelGGL((mfma_spmm_split_kernel<T, NT, 2>), dim3(grid),
                                   dim3(kThreads), 0, op->stream, s, io.X, io.Y, io.Vout,
                                   io.Sub, tl->ntiles, tl->ptr, tl->cols, tv);
            else if (tl->R == 16 && split == 4)
                hipLaunchKernelGGL((mfma_spmm_split_kernel<T, NT, 4>), dim3(grid),
                                   dim3(kThreads), 0, op->stream, s, io.X, io.Y, io.Vout,
                                   io.Sub, tl->ntiles, tl->ptr, tl->cols, tv);
            else if (tl->R == 16)
                hipLaunchKernelGGL((mfma_spmm_kernel<T, NT>), dim3(grid), dim3(kThreads), 0,
                                   op->stream, s, io.X, io.Y, io.Vout, io.Sub, tl->ntiles,
                                   tl->ptr, tl->cols, tv);
            else
                hipLaunchKernelGGL((tiled_spmm_sched_kernel<T, 8, NT>), dim3(grid),
                                   dim3(kThreads), 0, op->stream, s, io.X, io.Y, io.Vout,
                                   io.Sub, tl->ntiles, tl->ptr, tl->cols, tv);
            if (s.red_mode)
            {
                const int64_t prows = tl->ntiles * split;
                int rg = (int) ((prows + 63) / 64);
                rg = rg < 1 ? 1 : (rg > kTileSumGrid ? kTileSumGrid : rg);
                hipLaunchKernelGGL(tile_sum_kernel, dim3(rg), dim3(kThreads), 0, op->stream,
                                   (const double*) s.rb.tile_partial, prows, s.rb.partial);
                hipLaunchKernelGGL((total_kernel<T>), dim3(1), dim3(kTotalWaves * kWave), 0,
                                   op->stream, s, rg);
            }
        }
    }
    else if (csr_blocked_cols(s) > 0)
    {
        const int64_t ngroups = (s.nrows + kBlkRows - 1) / kBlkRows;
        int grid = grid_for(ngroups, op->spmm_grid);
        // no more workgroups than fit on the chip at once, so that all of
        // them start the column sweep together
        // 1024-thread workgroups (one per CU): half as many workgroups per
        // XCD to drift apart as with 512 threads (SLQ_BLK_WAVES=8)
        static const int wpg = [] {
            const char* e = std::getenv("SLQ_BLK_WAVES");
            return e && std::atoi(e) == 8 ? 8 : 16;
        }();
        auto kern = wpg == 16 ? csr_blocked_kernel<T, NT, 16> : csr_blocked_kernel<T, NT, 8>;
        const int threads = wpg * kWave;
        if (wpg == 16) grid = (int) std::max<int64_t>(kXcds, ((grid / 2) / kXcds) * kXcds);
        static const int resident = [&] {
            int per_cu = 0, cus = 0, dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                    hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) !=
                    hipSuccess)
                return 0;
            return ((per_cu * cus) / kXcds) * kXcds;
        }();
        if (resident >= kXcds && grid > resident) grid = resident;
        static const int sync = [] {
            const char* e = std::getenv("SLQ_BLK_SYNC");
            const int v = e ? std::atoi(e) : 4;
            return v >= 1 ? v : 1;
        }();
        hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, op->stream, s, io.X, io.Y,
                           io.Vout, io.Sub, csr_blocked_cols(s), sync);
        if (s.red_mode)
            hipLaunchKernelGGL((total_kernel<T>), dim3(1), dim3(kTotalWaves * kWave), 0,
                               op->stream, s, grid);
    }
    else
    {
        const int grid = grid_for(s.nrows, op->spmm_grid);
        hipLaunchKernelGGL((spmm_kernel<T, NT>), dim3(grid), dim3(kThreads), 0, op->stream, s,
                           io.X, io.Y, io.Vout, io.Sub);
        if (s.red_mode)
            hipLaunchKernelGGL((total_kernel<T>), dim3(1), dim3(kTotalWaves * kWave), 0,
                               op->stream, s, grid);
    }
}

template <typename T>
int launch_spmm(slq_op* op, const SpmmArgs<T>& s, const SpmmIO<T>& io,
                const DevTiles* tl = nullptr)
{
    if (s.dense && dense_ksplit(s.nrows, s.ncols) > 1 && !s.rb.tile_partial)
    {
        set_error("internal: the dense SpMM needs its split scratch");
        return SLQ_E_ARG;
    }
    if (tl && (!s.rb.tickets || (s.red_mode && !s.rb.tile_partial)))
    {
        set_error("internal: the tiled SpMM needs ticket counters and tile partials");
        return SLQ_E_ARG;
    }
    hipEvent_t ev = nullptr;
    prof_begin(op, &ev);
    if (s.nt_store) launch_spmm_nt<T, true>(op, s, io, tl);
    else launch_spmm_nt<T, false>(op, s, io, tl);
    SLQ_CHECK(hipGetLastError());
    prof_end(op, ev, 0);
    return SLQ_OK;
}

template <typename T>
int launch_update(slq_op* op, const UpdateArgs& u, T* R, const T* Vj, const T* Vjm1)
{
    hipEvent_t ev = nullptr;
    prof_begin(op, &ev);
    const int G = stream_grid((u.nrows + 3) / 4, op->stream_grid, 512);
    hipLaunchKernelGGL(lanczos_update_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, u,
                       R, Vj, Vjm1);
    SLQ_CHECK(hipGetLastError());
    if (u.do_norm)
    {
        hipLaunchKernelGGL(norm_total_kernel, dim3(1), dim3(kTotalWaves * kWave), 0,
                           op->stream, (const double*) u.rb.partial, G, u.j, u.tol_sqrt_n,
                           u.beta_out, u.scale_out, u.sizes);
        SLQ_CHECK(hipGetLastError());
    }
    prof_end(op, ev, 1);
    return SLQ_OK;
}

// Grid of the Gram-Schmidt kernels: fewer workgroups, since every one writes
// K partial rows.
int gs_grid(int64_t rows, int cap) { return stream_grid(rows, cap); }

// The fused subtraction + product pass is on unless SLQ_FUSED_GS=0 (read per
// call, so a process can time both).
bool fused_gs_enabled()
{
    const char* e = std::getenv("SLQ_FUSED_GS");
    return !(e && std::atoi(e) == 0);
}

// CGS2 of r against the K newest basis vectors (slots last, last-1, ...);
// the final pass finishes the norm (beta_j, or alpha_j for Golub-Kahn u).
template <typename T>
int reorthogonalize(slq_op* op, int64_t nrows, T* R, const T* basis, size_t slot,
                    int last, int K, double* h1, double* h2, double* partial_k,
                    const RedBuf& rb, double* norm_out, double* scale_out,
                    int32_t* sizes, int j, double tol_sqrt_n)
{
    const int G = gs_grid(nrows, op->stream_grid);
    GsArgs g;
    std::memset(&g, 0, sizeof(g));
    g.nrows = nrows;
    g.slot = (int64_t) slot;
    g.last = last;
    g.K = K;
    g.partial_k = partial_k;
    g.rb = rb;
    if (K > 0)
    {
        // h1 = V^T r; r -= V h1; h2 = V^T r; then r -= V h2 below
        const dim3 tg(kTotalWaves * kWave);
        hipLaunchKernelGGL(gs_dots_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g, R,
                           basis);
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), tg, 0, op->stream, partial_k, G,
                           K, h1);
        g.h = h1;
        // r -= V h1 and h2 = V^T r in one pass while the K vectors of a row
        // fit in registers without spilling (K <= 64), else two passes
        if (K <= 64 && fused_gs_enabled())
        {
            if (K > 32)
                hipLaunchKernelGGL((gs_subdots_kernel<T, 64, 4>), dim3(G), dim3(4 * kWave), 0,
                                   op->stream, g, R, basis);
            else if (K > 16)
                hipLaunchKernelGGL((gs_subdots_kernel<T, 32, kWavesPerWG>), dim3(G),
                                   dim3(kThreads), 0, op->stream, g, R, basis);
            else if (K > 8)
                hipLaunchKernelGGL((gs_subdots_kernel<T, 16, kWavesPerWG>), dim3(G),
                                   dim3(kThreads), 0, op->stream, g, R, basis);
            else
                hipLaunchKernelGGL((gs_subdots_kernel<T, 8, kWavesPerWG>), dim3(G),
                                   dim3(kThreads), 0, op->stream, g, R, basis);
        }
        else
        {
            hipLaunchKernelGGL(gs_sub_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g, R,
                               basis);
            hipLaunchKernelGGL(gs_dots_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g,
                               R, basis);
        }
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(K), tg, 0, op->stream, partial_k, G,
                           K, h2);
        SLQ_CHECK(hipGetLastError());
        g.h = h2;
    }
    g.do_norm = 1;
    hipLaunchKernelGGL(gs_sub_kernel<T>, dim3(G), dim3(kThreads), 0, op->stream, g, R, basis);
    SLQ_CHECK(hipGetLastError());
    hipLaunchKernelGGL(norm_total_kernel, dim3(1), dim3(kTotalWaves * kWave), 0, op->stream,
                       (const double*) rb.partial, G, j, tol_sqrt_n, norm_out, scale_out,
                       sizes);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
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

// Where Lanczos subtracts beta_{j-1} v_{j-1}: in the SpMM's epilogue
// (SLQ_SPMM_BETA = 1, the default: the SpMM reads v_{j-1}'s own rows while
// it streams its output, and the update kernel then reads two blocks
// instead of three) or in the update kernel (0).
bool spmm_beta_enabled()
{
    static const bool v = [] {
        const char* e = std::getenv("SLQ_SPMM_BETA");
        return !(e && std::atoi(e) == 0);
    }();
    return v;
}

// Lanczos without re-orthogonalisation: the basis is never stored. Step j
// gathers X_j = r_{j-1} (the previous residual, unnormalised) with its
// per-probe scale s_j = 1 / beta_{j-1}; the update reads v_j = s_j X_j and
// v_{j-1} = s_{j-1} X_{j-1} from the two previous residual buffers of a ring
// of three, rounding the products exactly as a stored basis would be
// rounded. One block write per step less than storing v_j.
template <typename T>
int lanczos_ring(const RecurrenceArgs& ra, Bump& bump)
{
    slq_op* op = ra.op;
    MatrixStore& A = *op->A;
    const int64_t n = A.csr.nrows;
    const int m = ra.m;
    const bool f32 = (sizeof(T) == 4);
    const size_t blk = (size_t) n * kBlock;
    T* B[3] = {bump.take<T>(blk), bump.take<T>(blk), bump.take<T>(blk)};
    double* S[3] = {bump.take<double>(kBlock), bump.take<double>(kBlock),
                    bump.take<double>(kBlock)};
    const RedBuf rb = take_redbuf(bump, sched_tiles(op));
    if (!B[0] || !B[1] || !B[2] || !S[0] || !S[1] || !S[2] || !redbuf_ok(rb))
    {
        set_error("internal: Lanczos workspace too small");
        return SLQ_E_OOM;
    }
    zero_counters(op->stream, rb);
    const double s0 = 1.0 / std::sqrt((double) n);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       S[0], ra.sizes, kBlock, s0);
    SLQ_CHECK(hipGetLastError());
    const double tol_sqrt_n = ra.tol * std::sqrt((double) n);
    const DevTiles* tl = nullptr;
    int shift_mode = 0;
    SLQ_TRY(operand_tiles<T>(op, ra.t, &tl, &shift_mode));
    const T* X0 = (const T*) ra.X0;
    const bool beta_in_spmm = spmm_beta_enabled();
    for (int j = 0; j < m; ++j)
    {
        T* W = B[j % 3];
        const T* Xj = j == 0 ? X0 : B[(j + 2) % 3];
        const T* Xjm1 = j == 0 ? nullptr : (j == 1 ? X0 : B[(j + 1) % 3]);

        SpmmArgs<T> s = spmm_defaults<T>();
        fill_matrix<T>(s, A, op->kind == OP_AFFINE_B ? op->B.get() : nullptr, false, f32);
        s.shift_mode = shift_mode;
        s.shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
        s.t = ra.t;
        s.xscale = S[j % 3];
        s.red_mode = 1;
        s.rb = rb;
        s.red_out = ra.alpha + (size_t) j * kBlock;
        SpmmIO<T> io;
        io.X = Xj;
        io.Y = W;
        // r = (w - beta_{j-1} v_{j-1}) - alpha_j v_j when the epilogue takes
        // the beta term (alpha_j = v_j . w is summed before it)
        const bool fused = j > 0 && beta_in_spmm;
        if (fused)
        {
            io.Sub = Xjm1;
            s.sub_coef = ra.beta + (size_t) (j - 1) * kBlock;
            s.sub_scale = S[(j + 2) % 3];
        }
        SLQ_TRY(launch_spmm<T>(op, s, io, tl));

        UpdateArgs u;
        std::memset(&u, 0, sizeof(u));
        u.nrows = n;
        u.alpha = ra.alpha + (size_t) j * kBlock;
        u.beta = j > 0 && !fused ? ra.beta + (size_t) (j - 1) * kBlock : nullptr;
        u.vscale = S[j % 3];
        u.uscale = j > 0 && !fused ? S[(j + 2) % 3] : nullptr;
        u.do_norm = 1;
        u.rb = rb;
        u.beta_out = ra.beta + (size_t) j * kBlock;
        u.scale_out = S[(j + 1) % 3];
        u.sizes = ra.sizes;
        u.j = j;
        u.tol_sqrt_n = tol_sqrt_n;
        SLQ_TRY(launch_update<T>(op, u, W, Xj, fused ? nullptr : Xjm1));
    }
    return SLQ_OK;
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

    if (!reorth) return lanczos_ring<T>(ra, bump);
    T* P = bump.take<T>(blk);
    T* Q = bump.take<T>(blk);
    const size_t vst = slot_stride(n);
    T* V = bump.take<T>(vst * nslots);
    double* scale = bump.take<double>(kBlock);
    double* ones = bump.take<double>(kBlock);   // scale of the stored basis
    const RedBuf rb = take_redbuf(bump, sched_tiles(op));
    double* h1 = nullptr;
    double* h2 = nullptr;
    double* partial_k = nullptr;
    if (reorth)
    {
        h1 = bump.take<double>((size_t) (m + 1) * kBlock);
        h2 = bump.take<double>((size_t) (m + 1) * kBlock);
        partial_k = bump.take<double>((size_t) kMaxGrid * (m + 1) * kBlock);
    }
    if (!P || !Q || !V || !scale || !ones || !redbuf_ok(rb) ||
        (reorth && (!h1 || !h2 || !partial_k)))
    {
        set_error("internal: Lanczos workspace too small");
        return SLQ_E_OOM;
    }
    zero_counters(op->stream, rb);
    hipLaunchKernelGGL(set_ones_kernel, dim3(1), dim3(kBlock), 0, op->stream, ones);
    const bool beta_in_spmm = spmm_beta_enabled();
    // initial_beta = |probe| = sqrt(n) exactly for a Rademacher probe
    const double s0 = 1.0 / std::sqrt((double) n);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       scale, ra.sizes, kBlock, s0);
    SLQ_CHECK(hipGetLastError());
    const double tol_sqrt_n = ra.tol * std::sqrt((double) n);
    const DevTiles* tl = nullptr;
    int shift_mode = 0;
    SLQ_TRY(operand_tiles<T>(op, ra.t, &tl, &shift_mode));

    const T* src = (const T*) ra.X0;
    T* bufs[2] = {Q, P};
    for (int j = 0; j < m; ++j)
    {
        T* W = bufs[j & 1];
        T* Vj = V + (size_t) (reorth ? j : (j & 1)) * vst;
        T* Vjm1 = j > 0 ? V + (size_t) (reorth ? j - 1 : ((j - 1) & 1)) * vst : nullptr;

        SpmmArgs<T> s = spmm_defaults<T>();
        fill_matrix<T>(s, A, op->kind == OP_AFFINE_B ? op->B.get() : nullptr, false, f32);
        s.shift_mode = shift_mode;
        s.shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
        s.t = ra.t;
        s.xscale = scale;
        s.red_mode = 1;
        s.rb = rb;
        s.red_out = ra.alpha + (size_t) j * kBlock;
        SpmmIO<T> io;
        io.X = src;
        io.Y = W;
        io.Vout = Vj;
        // beta_{j-1} v_{j-1} subtracted in the epilogue (see lanczos_ring)
        const bool fused = j > 0 && beta_in_spmm;
        if (fused)
        {
            io.Sub = Vjm1;
            s.sub_coef = ra.beta + (size_t) (j - 1) * kBlock;
            s.sub_scale = ones;
        }
        SLQ_TRY(launch_spmm<T>(op, s, io, tl));

        UpdateArgs u;
        std::memset(&u, 0, sizeof(u));
        u.nrows = n;
        u.alpha = ra.alpha + (size_t) j * kBlock;
        u.beta = j > 0 && !fused ? ra.beta + (size_t) (j - 1) * kBlock : nullptr;
        u.do_norm = reorth ? 0 : 1;
        u.rb = rb;
        u.beta_out = ra.beta + (size_t) j * kBlock;
        u.scale_out = scale;
        u.sizes = ra.sizes;
        u.j = j;
        u.tol_sqrt_n = tol_sqrt_n;
        SLQ_TRY(launch_update<T>(op, u, W, Vj, fused ? nullptr : Vjm1));

        if (reorth)
        {
            const int num_ortho = j < buffer ? j + 1 : buffer;
            SLQ_TRY(reorthogonalize<T>(op, n, W, V, vst, j, num_ortho, h1, h2, partial_k, rb,
                                       ra.beta + (size_t) j * kBlock, scale, ra.sizes, j,
                                       tol_sqrt_n));
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
    const size_t vblk = (size_t) nc * kBlock;

    const size_t ust = slot_stride(nr), vst = slot_stride(nc);
    T* U = bump.take<T>(ust * nslots);
    T* V = bump.take<T>(vst * nslots);
    double* su = bump.take<double>((size_t) (m + 1) * kBlock);   // 1/alpha_j
    double* sv = bump.take<double>((size_t) (m + 1) * kBlock);   // 1/beta_{j-1}
    const RedBuf rb = take_redbuf(bump, sched_tiles(op));
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
    if (!U || !V || !su || !sv || !redbuf_ok(rb) || !usizes ||
        (reorth && (!h1 || !h2 || !partial_k)))
    {
        set_error("internal: Golub-Kahn workspace too small");
        return SLQ_E_OOM;
    }
    zero_counters(op->stream, rb);
    // v_0 = probe / |probe|, |probe| = sqrt(ncols) exactly
    const double s0 = 1.0 / std::sqrt((double) nc);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       sv, ra.sizes, kBlock, s0);
    hipLaunchKernelGGL(init_block_kernel, dim3(1), dim3(kBlock), 0, op->stream,
                       su, usizes, kBlock, 0.0);
    SLQ_CHECK(hipGetLastError());
    auto uslot = [&](int j) { return U + (size_t) (reorth ? j : j % 3) * ust; };
    auto vslot = [&](int j) { return V + (size_t) (reorth ? j : j % 3) * vst; };
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
    const DevTiles* tla = nullptr;
    const DevTiles* tlt = nullptr;
    int mode_a = mode, mode_t = mode;
    SLQ_TRY(operand_tiles<T>(op, ra.t, &tla, &mode_a, false));
    SLQ_TRY(operand_tiles<T>(op, ra.t, &tlt, &mode_t, true));

    for (int j = 0; j < m; ++j)
    {
        // u_j = A v_j - beta_{j-1} u_{j-1}; alpha_j = |u_j|
        // (c_golub_kahn_bidiagonalization.cpp:162-195)
        SpmmArgs<T> s = spmm_defaults<T>();
        fill_matrix<T>(s, A, Bm, false, f32);
        s.shift_mode = mode_a;
        s.shift_len = shift_len;
        s.t = ra.t;
        s.xscale = sv + (size_t) j * kBlock;
        SpmmIO<T> io;
        io.X = vslot(j);
        io.Y = uslot(j);
        if (j > 0)
        {
            io.Sub = uslot(j - 1);
            s.sub_coef = ra.beta + (size_t) (j - 1) * kBlock;
            s.sub_scale = su + (size_t) (j - 1) * kBlock;
        }
        s.red_mode = reorth ? 0 : 2;
        s.finish_sqrt = 1;
        s.rb = rb;
        s.red_out = ra.alpha + (size_t) j * kBlock;
        s.scale_out = su + (size_t) j * kBlock;
        s.sizes = usizes;            // a small alpha never ends the recurrence
        s.j = j;
        s.tol_sqrt_n = -1.0;
        SLQ_TRY(launch_spmm<T>(op, s, io, tla));
        if (reorth)
        {
            const int num_ortho = j < buffer ? j : buffer - 1;
            SLQ_TRY(reorthogonalize<T>(op, nr, uslot(j), U, ust, j - 1, j > 0 ? num_ortho : 0,
                                       h1, h2, partial_k, rb, ra.alpha + (size_t) j * kBlock,
                                       su + (size_t) j * kBlock, usizes, j, -1.0));
            SLQ_TRY(normalize_in_place<T>(op, uslot(j), nr, su + (size_t) j * kBlock));
        }

        // v_{j+1} = A^T u_j - alpha_j v_j; beta_j = |v_{j+1}|
        // (c_golub_kahn_bidiagonalization.cpp:197-228)
        SpmmArgs<T> t = spmm_defaults<T>();
        fill_matrix<T>(t, A, Bm, true, f32);
        t.shift_mode = mode_t;
        t.shift_len = shift_len;
        t.t = ra.t;
        t.xscale = su + (size_t) j * kBlock;
        SpmmIO<T> io2;
        io2.X = uslot(j);
        io2.Y = vslot(j + 1);
        io2.Sub = vslot(j);
        t.sub_coef = ra.alpha + (size_t) j * kBlock;
        t.sub_scale = sv + (size_t) j * kBlock;
        t.red_mode = reorth ? 0 : 2;
        t.finish_sqrt = 1;
        t.rb = rb;
        t.red_out = ra.beta + (size_t) j * kBlock;
        t.scale_out = sv + (size_t) (j + 1) * kBlock;
        t.sizes = ra.sizes;
        t.j = j;
        t.tol_sqrt_n = tol_sqrt_n;
        SLQ_TRY(launch_spmm<T>(op, t, io2, tlt));
        if (reorth)
        {
            const int num_ortho = j < buffer ? j : buffer - 1;
            SLQ_TRY(reorthogonalize<T>(op, nc, vslot(j + 1), V, vst, j, num_ortho, h1, h2,
                                       partial_k, rb, ra.beta + (size_t) j * kBlock,
                                       sv + (size_t) (j + 1) * kBlock, ra.sizes, j,
                                       tol_sqrt_n));
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
        if (reorth)
        {
            add(blk);
            add(blk);
            add(slot_stride(nr) * value_bytes * m);
            add(sizeof(double) * kBlock);
            add(sizeof(double) * kBlock);
        }
        else
        {
            for (int k = 0; k < 3; ++k) add(blk);
            for (int k = 0; k < 3; ++k) add(sizeof(double) * kBlock);
        }
    }
    else
    {
        const int ns = reorth ? m + 1 : 3;
        add(slot_stride(nr) * value_bytes * ns);
        add(slot_stride(nc) * value_bytes * ns);
        add(sizeof(double) * (m + 1) * kBlock);
        add(sizeof(double) * (m + 1) * kBlock);
        add(sizeof(int32_t) * kBlock);
    }
    b += redbuf_bytes(sched_tiles(op));
    if (reorth)
    {
        add(sizeof(double) * (m + 1) * kBlock);
        add(sizeof(double) * (m + 1) * kBlock);
        add(sizeof(double) * (size_t) kMaxGrid * (m + 1) * kBlock);
    }
    return b + 4096;
}

int run_recurrence(const RecurrenceArgs& ra, void* scratch, size_t scratch_bytes)
{
    Bump bump{(char*) scratch, scratch_bytes};
    const bool f32 = ra.precision == SLQ_PREC_MIXED || ra.op->A->csr.dtype == SLQ_F32;
    if (ra.gram && !ra.op->A->has_trans && !ra.op->A->dn.trans)
    {
        set_error("internal: transpose not built");
        return SLQ_E_ARG;
    }
    if (f32)
        return ra.gram ? golub_kahn<float>(ra, bump) : lanczos<float>(ra, bump);
    return ra.gram ? golub_kahn<double>(ra, bump) : lanczos<double>(ra, bump);
}

size_t spmm_block_scratch_bytes(const slq_op* op)
{
    return redbuf_bytes(sched_tiles(op));
}

namespace {

template <typename T>
int spmm_block_t(slq_op* op, double t, const T* X, T* Y, Bump& bump)
{
    MatrixStore& A = *op->A;
    const bool f32 = sizeof(T) == 4;
    const RedBuf rb = take_redbuf(bump, sched_tiles(op));
    if (!redbuf_ok(rb))
    {
        set_error("internal: SpMM workspace too small");
        return SLQ_E_OOM;
    }
    zero_counters(op->stream, rb);
    SpmmArgs<T> s = spmm_defaults<T>();
    fill_matrix<T>(s, A, op->kind == OP_AFFINE_B ? op->B.get() : nullptr, false, f32);
    s.shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
    s.t = t;
    s.rb = rb;
    const DevTiles* tl = nullptr;
    int shift_mode = 0;
    SLQ_TRY(operand_tiles<T>(op, t, &tl, &shift_mode));
    s.shift_mode = shift_mode;
    SpmmIO<T> io;
    io.X = X;
    io.Y = Y;
    return launch_spmm<T>(op, s, io, tl);
}

}  // namespace

int spmm_block(slq_op* op, double t, int value_dtype, const void* X, void* Y, void* scratch,
               size_t scratch_bytes)
{
    Bump bump{(char*) scratch, scratch_bytes};
    if (value_dtype == SLQ_F32)
        return spmm_block_t<float>(op, t, (const float*) X, (float*) Y, bump);
    return spmm_block_t<double>(op, t, (const double*) X, (double*) Y, bump);
}

int run_spmm(slq_op* op, double t, int transpose, const void* X, void* Y,
             int value_dtype)
{
    MatrixStore& A = *op->A;
    MatrixStore* Bm = op->kind == OP_AFFINE_B ? op->B.get() : nullptr;
    const int64_t shift_len = A.csr.nrows < A.csr.ncols ? A.csr.nrows : A.csr.ncols;
    const int mode = op->kind == OP_AFFINE_I ? 1 : (op->kind == OP_AFFINE_B ? 2 : 0);
    if (op->kind == OP_AFFINE_B)
        SLQ_TRY(ensure_pair_tiles(op, value_dtype == SLQ_F32, transpose != 0));
    // ticket counters of the in-order tile schedule (no reduction here)
    unsigned int* tickets = nullptr;
    SLQ_CHECK(hipMalloc(&tickets, sizeof(unsigned int) * kTicketWords));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(kTicketWords / 64), dim3(64), 0, op->stream,
                       tickets, kTicketWords);
    // split scratch of the dense SpMM (no reduction here, so no tile rows)
    double* scratch = nullptr;
    const int64_t srows = A.dense ? sched_tiles(op) : 0;
    if (srows > 0 && hipMalloc(&scratch, sizeof(double) * (size_t) srows * kBlock) != hipSuccess)
    {
        (void) hipFree(tickets);
        set_error("slq_op_spmm: out of device memory for the dense split scratch");
        return SLQ_E_OOM;
    }
    int r;
    if (value_dtype == SLQ_F32)
    {
        SpmmArgs<float> s = spmm_defaults<float>();
        fill_matrix<float>(s, A, Bm, transpose != 0, A.csr.dtype == SLQ_F64);
        s.shift_mode = mode;
        s.shift_len = shift_len;
        s.t = t;
        s.rb.tickets = tickets;
        s.rb.tile_partial = scratch;
        SpmmIO<float> io;
        io.X = (const float*) X;
        io.Y = (float*) Y;
        const DevTiles* tl = nullptr;
        r = operand_tiles<float>(op, t, &tl, &s.shift_mode, transpose != 0);
        if (r == SLQ_OK) r = launch_spmm<float>(op, s, io, tl);
    }
    else
    {
        SpmmArgs<double> s = spmm_defaults<double>();
        fill_matrix<double>(s, A, Bm, transpose != 0, false);
        s.shift_mode = mode;
        s.shift_len = shift_len;
        s.t = t;
        s.rb.tickets = tickets;
        s.rb.tile_partial = scratch;
        SpmmIO<double> io;
        io.X = (const double*) X;
        io.Y = (double*) Y;
        const DevTiles* tl = nullptr;
        r = operand_tiles<double>(op, t, &tl, &s.shift_mode, transpose != 0);
        if (r == SLQ_OK) r = launch_spmm<double>(op, s, io, tl);
    }
    const hipError_t e = hipStreamSynchronize(op->stream);
    (void) hipFree(tickets);
    if (scratch) (void) hipFree(scratch);
    if (r != SLQ_OK) return r;
    SLQ_CHECK(e);
    return SLQ_OK;
}

template <typename T>
__global__ void tiles_refill_kernel(const int32_t* __restrict__ src,
                                    const T* __restrict__ data, T* __restrict__ vals,
                                    int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
    {
        const int32_t k = src[i];
        vals[i] = k >= 0 ? data[k] : T(0);
    }
}

int tiles_refill(hipStream_t s, DevTiles& t, const void* csr_data, int dtype)
{
    if (t.R == 0) return SLQ_OK;
    if (!t.src)
    {
        set_error("the matrix has duplicate entries or too many non-zeros for an "
                  "in-place value refresh; create a new operator instead");
        return SLQ_E_ARG;
    }
    const int64_t n = t.nslabs * t.R;
    const int64_t g = (n + 255) / 256;
    if (dtype == SLQ_F64)
        hipLaunchKernelGGL((tiles_refill_kernel<double>), dim3((unsigned) g), dim3(256), 0, s,
                           t.src, (const double*) csr_data, (double*) t.vals, n);
    else
        hipLaunchKernelGGL((tiles_refill_kernel<float>), dim3((unsigned) g), dim3(256), 0, s,
                           t.src, (const float*) csr_data, (float*) t.vals, n);
    SLQ_CHECK(hipGetLastError());
    if (t.vals_f32)
    {
        (void) hipStreamSynchronize(s);
        (void) hipFree(t.vals_f32);
        t.vals_f32 = nullptr;
    }
    return SLQ_OK;
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

// out[c][r] = in[r][c] for r < nrows, c < ncols; 32 x 32 tiles via LDS.
template <typename T>
__global__ void dense_transpose_kernel(const T* __restrict__ in, int64_t nrows, int64_t ncols,
                                       int64_t ld_in, T* __restrict__ out, int64_t ld_out)
{
    __shared__ T tile[32][33];
    const int64_t c0 = (int64_t) blockIdx.x * 32, r0 = (int64_t) blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y)
    {
        const int64_t r = r0 + j, c = c0 + threadIdx.x;
        tile[j][threadIdx.x] = (r < nrows && c < ncols) ? in[r * ld_in + c] : T(0);
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y)
    {
        const int64_t c = c0 + j, r = r0 + threadIdx.x;
        if (c < ncols && r < nrows) out[c * ld_out + r] = tile[threadIdx.x][j];
    }
}

int dense_transpose(hipStream_t s, MatrixStore& M)
{
    if (M.dn.trans) return SLQ_OK;
    const int64_t nr = M.csr.nrows, nc = M.csr.ncols;
    const size_t vb = M.csr.dtype == SLQ_F64 ? 8 : 4;
    M.dn.trans_ld = (nr + kDenseLd - 1) / kDenseLd * kDenseLd;
    const size_t bytes = (size_t) nc * (size_t) M.dn.trans_ld * vb;
    SLQ_CHECK(hipMalloc(&M.dn.trans, bytes));
    SLQ_CHECK(hipMemsetAsync(M.dn.trans, 0, bytes, s));
    const dim3 grid((unsigned) ((nc + 31) / 32), (unsigned) ((nr + 31) / 32));
    if (M.csr.dtype == SLQ_F64)
        hipLaunchKernelGGL(dense_transpose_kernel<double>, grid, dim3(32, 8), 0, s,
                           (const double*) M.dn.data, nr, nc, M.dn.ld, (double*) M.dn.trans,
                           M.dn.trans_ld);
    else
        hipLaunchKernelGGL(dense_transpose_kernel<float>, grid, dim3(32, 8), 0, s,
                           (const float*) M.dn.data, nr, nc, M.dn.ld, (float*) M.dn.trans,
                           M.dn.trans_ld);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

int ensure_f32(hipStream_t s, MatrixStore& M, bool need_trans)
{
    if (M.csr.dtype != SLQ_F64) return SLQ_OK;
    if (M.dense)
    {
        auto convert = [&](const void* src, float** dst, int64_t n) -> int {
            if (*dst || !src) return SLQ_OK;
            SLQ_CHECK(hipMalloc(dst, sizeof(float) * n));
            hipLaunchKernelGGL((convert_kernel<double, float>), dim3((unsigned) ((n + 255) / 256)),
                               dim3(256), 0, s, (const double*) src, *dst, n);
            SLQ_CHECK(hipGetLastError());
            return SLQ_OK;
        };
        SLQ_TRY(convert(M.dn.data, &M.dn.data_f32, M.csr.nrows * M.dn.ld));
        if (need_trans) SLQ_TRY(convert(M.dn.trans, &M.dn.trans_f32, M.csr.ncols * M.dn.trans_ld));
        return SLQ_OK;
    }
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
