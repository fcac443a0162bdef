// Internal declarations of libslq_hip.so (gfx950). Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "slq_hip.h"
#include "estimator.hpp"

namespace slq {

// ---------------------------------------------------------------- errors

int hip_status(hipError_t e, const char* what, const char* file, int line);

#define SLQ_CHECK(expr)                                                      \
    do {                                                                     \
        hipError_t slq_e_ = (expr);                                          \
        if (slq_e_ != hipSuccess)                                            \
            return ::slq::hip_status(slq_e_, #expr, __FILE__, __LINE__);     \
    } while (0)

#define SLQ_TRY(expr)                                                        \
    do {                                                                     \
        int slq_r_ = (expr);                                                 \
        if (slq_r_ != SLQ_OK) return slq_r_;                                 \
    } while (0)

// ---------------------------------------------------------------- layout

// Probe blocks are row-major [len][kBlock]: entry (i, p) of probe p sits at
// i * kBlock + p, so one wavefront reads or writes a whole row of the block
// (kBlock consecutive values) per instruction and a CSR gather of column c
// fetches the contiguous row c of the block.
constexpr int kBlock = 64;
constexpr int kWave = 64;

// Persistent-style grids: workgroups of kThreads threads (kWavesPerWG
// wavefronts) sweeping the rows in XCD windows. kGrid workgroups (2 per CU)
// was measured best for the SpMM at 2^20 rows: more resident wavefronts
// widen the window of gathered rows past an XCD's L2, fewer cannot hide the
// gather latency. kMaxGrid bounds the reduction buffers.
constexpr int kThreads = 512;
constexpr int kWavesPerWG = kThreads / kWave;
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups round-robin
constexpr int kGrid = 512;
constexpr int kMaxGrid = 1024;

// ---------------------------------------------------------------- matrices

// A CSR matrix in device memory: int64 row pointers, int32 column indices,
// values of type dtype (SLQ_F64 / SLQ_F32). Owned by a MatrixStore.
struct DevCSR {
    int64_t nrows = 0;
    int64_t ncols = 0;
    int64_t nnz = 0;
    int dtype = SLQ_F64;
    int64_t* indptr = nullptr;
    int32_t* indices = nullptr;
    void* data = nullptr;
    bool sorted = false;   // column indices non-decreasing within every row
};

// Row-tiled "column slab" form of a CSR matrix for the SpMM. Rows are grouped
// in tiles of R consecutive rows; tile t lists the sorted union of the column
// indices of its rows (cols[ptr[t] .. ptr[t+1])) and, for each such column u,
// the R values a[t*R + r][u] (zero where absent) stored contiguously
// (vals[u * R + r]). One gathered operand row X[u] then feeds R FMAs instead
// of one, which is what keeps the gather traffic of a 64-probe block out of
// the cache bandwidth limit on spatially ordered matrices.
struct DevTiles {
    int R = 0;                 // 0 = not built / not profitable
    int64_t ntiles = 0;
    int64_t nslabs = 0;        // total unique (tile, column) pairs
    int64_t* ptr = nullptr;    // [ntiles + 1]
    int32_t* cols = nullptr;   // [nslabs]
    void* vals = nullptr;      // [nslabs * R], dtype of the matrix
    float* vals_f32 = nullptr; // FP32 copy (mixed precision)
    // [nslabs * R] position in the CSR value array of each slot (-1: zero
    // fill); null when the CSR has duplicate entries or more than 2^31-1
    // non-zeros. Lets new values of the same pattern be scattered in place.
    int32_t* src = nullptr;
};

// Device copy of a matrix plus derived forms built on demand: the transpose
// (CSR of A^T, used as a gather for A^T u), an FP32 copy of the values of an
// FP64 matrix (for the mixed-precision form) and the row-tiled form.
// Dense row-major form (a dense input matrix): rows padded to `ld`, a
// multiple of 64, with zeros; the transpose and FP32 copies are made on
// first use (Golub-Kahn, mixed precision).
struct DevDense {
    int64_t ld = 0;
    void* data = nullptr;          // [nrows][ld], dtype of the matrix
    int64_t trans_ld = 0;
    void* trans = nullptr;         // [ncols][trans_ld]
    float* data_f32 = nullptr;
    float* trans_f32 = nullptr;
};
constexpr int64_t kDenseLd = 64;

struct MatrixStore {
    bool dense = false;    // csr holds only nrows, ncols, nnz and dtype
    DevDense dn;
    DevCSR csr;
    DevCSR trans;          // valid when has_trans
    bool has_trans = false;
    float* data_f32 = nullptr;        // FP32 values of csr (mixed precision)
    float* trans_data_f32 = nullptr;  // FP32 values of trans
    DevTiles tiles;                   // of csr
    DevTiles trans_tiles;             // of trans
    bool trans_tiles_tried = false;
    // made from CSC arrays: CSC position of each CSR entry (host), so that
    // new values can be given in CSC order; empty otherwise
    std::vector<int64_t> csc_order;
    // value family (slq_op_family_attach): per-entry d, w, c in CSR order,
    // FP64, [3][nnz]; values w exp(-d / scale) + c are made on the device
    double* family = nullptr;
    ~MatrixStore();
};

// Host construction of the tiled form (R in {8, 16}); returns R = 0 in
// `out` when fewer than `min_density` of the slab entries would be non-zero.
// With R = 16 every tile's slab count is padded to a multiple of eight (zero
// slabs repeating the tile's last column) and the values are laid out as the
// matrix-core SpMM's operands (tile_slot in capi.hip).
// `data` may be null (values left zero); `src_host`, when given, receives
// the host copy of the slot -> CSR position map (empty if not kept).
int build_tiles(int64_t nrows, int64_t ncols, const int64_t* indptr,
                const int32_t* indices, const void* data, int dtype, int R,
                double min_density, DevTiles& out,
                std::vector<int32_t>* src_host = nullptr);
int tiles_f32(hipStream_t s, DevTiles& t, int dtype);
// Refill t.vals from new CSR values through t.src.
int tiles_refill(hipStream_t s, DevTiles& t, const void* csr_data, int dtype);

// Device buffers reused across calls on one handle.
struct Workspace {
    size_t bytes = 0;
    void* base = nullptr;
    ~Workspace();
    int reserve(size_t need);
};

struct ProfileState {
    bool enabled = false;
    double spmm_ms = 0.0;
    int64_t spmm_launches = 0;
    double update_ms = 0.0;
    int64_t update_launches = 0;
    // completed Lanczos / Golub-Kahn steps of the processed samples (summed
    // over the recurrences each sample ran) of slq_estimate calls
    int64_t steps = 0;
    std::vector<hipEvent_t> pool;            // free events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spmm_pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> update_pending;
    // matrix-core GEMMs of the block Krylov form (Q^T W and W -= Q H)
    double gemm_ms = 0.0;
    int64_t gemm_launches = 0;
    double gemm_flops = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_pending;
    std::vector<double> gemm_pending_flops;
};

// A + tB in the tiled form (the general affine operator): the tile
// structure of the union of A's and B's patterns, and for each slot the
// position of its value in A's and in B's CSR value arrays (-1: none). The
// values of A + tB for the t of a recurrence are gathered into tiles.vals
// (dtype of A; tiles.vals_f32 for the mixed-precision form of an FP64 pair)
// on the device before it, so new values set on A or B are always seen.
struct PairTiles {
    DevTiles tiles;            // tiles.src unused
    int32_t* src_a = nullptr;  // [nslabs * R]
    int32_t* src_b = nullptr;  // [nslabs * R]
    bool tried = false;
    // a view of another handle's pair tiles (with its own value buffer):
    // owns nothing
    bool borrowed = false;
    ~PairTiles();
};

}  // namespace slq

enum OpKind { OP_MATRIX = 0, OP_AFFINE_I = 1, OP_AFFINE_B = 2 };

struct slq_op {
    int kind = OP_MATRIX;
    std::shared_ptr<slq::MatrixStore> A;
    std::shared_ptr<slq::MatrixStore> B;   // OP_AFFINE_B only
    hipStream_t stream = nullptr;
    // streams of the recurrences of an A + tB estimate that run side by
    // side, one per t of a group (created on first use)
    std::vector<hipStream_t> lane_streams;
    // the per-block quadrature and copy-back of slq_estimate (first use)
    hipStream_t aux_stream = nullptr;
    int device = 0;
    slq::Workspace ws;
    slq::ProfileState prof;
    slq::PairTiles ab;                      // OP_AFFINE_B only: A + tB
    slq::PairTiles ab_t;                    // and A^T + tB^T (Golub-Kahn)
    // launch-grid caps of the kernels (0 = library default)
    int spmm_grid = 0;
    int stream_grid = 0;
    // caps chosen by the autotuner (0 = not tuned); slq_estimate_batch runs
    // with them, every other entry point with the caps above, so a tuned
    // handle gives slq_estimate the same reduction order as before
    int tuned_spmm_grid = 0;
    int tuned_stream_grid = 0;
    // wavefronts per tile of the matrix-core SpMM (1, 2, 4; 0 = default)
    int spmm_split = 0;
    int tuned_spmm_split = 0;
    // operand form of the SpMM: 0 the row tiles where they were built,
    // 1 the CSR rows (tuned: 0 = not tuned, else mode + 1)
    int tile_mode = 0;
    int tuned_tile_mode = 0;
    // rocBLAS handle of the block Krylov form's eigensolver (first use)
    void* rocblas = nullptr;
};

namespace slq {

// ---------------------------------------------------------------- kernels

// Rademacher probes: words [count][W] of the thread-0 stream starting at
// sample first_sample, expanded into X [len][kBlock] (probes >= count zero).
int rng_words(hipStream_t s, const uint64_t base_state[4], int64_t W,
              int64_t first_sample, int32_t count, uint64_t* words);
int rng_expand(hipStream_t s, const uint64_t* words, int64_t W, int64_t len,
               int32_t count, int dtype, void* X);

// Lanczos / Golub-Kahn over one probe block; outputs alpha/beta [m][kBlock]
// (FP64, device) and sizes [kBlock] (device).
struct RecurrenceArgs {
    slq_op* op = nullptr;
    double t = 0.0;             // affine parameter
    int gram = 0;
    int m = 0;
    double tol = 0.0;
    int ortho = 0;
    int precision = SLQ_PREC_NATIVE;
    const void* X0 = nullptr;   // probe block [len][kBlock] (value type of the run)
    double* alpha = nullptr;    // device [m][kBlock]
    double* beta = nullptr;     // device [m][kBlock]
    int32_t* sizes = nullptr;   // device [kBlock]
};
int run_recurrence(const RecurrenceArgs& a, void* scratch, size_t scratch_bytes);
size_t recurrence_scratch_bytes(const slq_op* op, int gram, int m, int ortho,
                                int value_bytes);

// Device eigensolve + quadrature for `count` probes.
int run_quadrature(hipStream_t s, const double* alpha, const double* beta,
                   const int32_t* sizes, int count, int m, int gram, int func,
                   const double* func_params,
                   double exponent, double n, const double* params_dev, int q,
                   int shift, int j_out, int q_stride, double* samples_dev,
                   double* theta_dev, double* tau_dev);

// Y = (A + tB) X on one 64-probe block (value type SLQ_F64 or SLQ_F32; the
// FP32 copy of an FP64 matrix must exist), with the SpMM's counters and
// partial rows taken from `scratch` (spmm_block_scratch_bytes of them).
size_t spmm_block_scratch_bytes(const slq_op* op);
int spmm_block(slq_op* op, double t, int value_dtype, const void* X, void* Y, void* scratch,
               size_t scratch_bytes);

// Block Lanczos with full re-orthogonalisation of the shared basis of a
// 64-probe block (block_krylov.hip), run in groups of blocks:
//   phase 0: clears the group's rank-loss flag;
//   phase 1: m block steps from the probe block X0 ([len][64], value type of
//     the run) for A + tB, the block tridiagonal (64 m x 64 m) kept in slot
//     `slot`;
//   phase 2: the eigensolves of slots 0 .. nslots - 1 in one batched call,
//     then block Gauss quadrature; slot s writes its per-probe samples to
//     samples_dev[s][p * q_stride + j_out[s]] like run_quadrature (every t
//     of the Ritz shift, by params_dev[j] - params_dev[0], when shift).
// All on the handle's stream. *flags_out receives the device address of the
// flags: [0] rank loss of a recurrence since phase 0, [1 + s] eigensolver
// info of slot s.
constexpr int kBkMaxSlots = 8;
struct BlockKrylovArgs {
    slq_op* op = nullptr;
    int phase = 0;
    int m = 0;
    int value_dtype = SLQ_F64;
    int** flags_out = nullptr;
    // phase 1
    int slot = 0;
    double t = 0.0;
    const void* X0 = nullptr;
    // phase 2
    int nslots = 0;
    int j_out[kBkMaxSlots] = {};
    double* samples_dev[kBkMaxSlots] = {};
    int func = 0;
    const double* func_params = nullptr;
    double exponent = 1.0;
    const double* params_dev = nullptr;
    int q = 1;
    int shift = 0;
    int q_stride = 1;
};
// Slots (block tridiagonals held for one batched eigensolve) at this
// lanczos_degree: at most kBkMaxSlots and 1 GiB of them, at least one.
int block_krylov_slots(int m);
size_t block_krylov_scratch_bytes(const slq_op* op, int m, int value_bytes);
int run_block_krylov(const BlockKrylovArgs& a, void* scratch, size_t scratch_bytes);
// Frees the handle's rocBLAS handle (if the block form created one).
void release_solver(slq_op* op);

// Plain block SpMM (tests): Y = (A + tB) X or its transpose.
int run_spmm(slq_op* op, double t, int transpose, const void* X, void* Y,
             int value_dtype);

// Transpose of a device CSR (stable: rows ascending within a column).
int build_transpose(hipStream_t s, const DevCSR& A, DevCSR& T);
int ensure_f32(hipStream_t s, MatrixStore& M, bool need_trans);
// Builds op->ab (or, with `trans`, op->ab_t from the CSR transposes of A and
// B, which must exist) for an A + tB operator (once; R = 0 if not
// profitable) and the FP32 value buffer when `need_f32`.
int ensure_pair_tiles(slq_op* op, bool need_f32, bool trans = false);
// Transposed copy of a dense matrix (for Golub-Kahn).
int dense_transpose(hipStream_t s, MatrixStore& M);

// Profiling hooks around kernel launches.
void prof_begin(slq_op* op, hipEvent_t* ev);
void prof_end(slq_op* op, hipEvent_t ev0, int kind);
// end of a timed block Krylov GEMM of `flops` floating-point operations
void prof_end_gemm(slq_op* op, hipEvent_t ev0, double flops);
void prof_collect(slq_op* op);

int copy_bandwidth(int64_t bytes, int iters, double* ms);

// Grid of a launch over `units` work units with an optional cap (0 = the
// library default cap).
int grid_for(int64_t units, int cap = 0);

}  // namespace slq
