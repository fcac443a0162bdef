/*
 * slq_hip.h — C ABI of libslq_hip.so, the MI355X (gfx950) implementation of
 * stochastic Lanczos quadrature (SLQ) trace estimation behind
 * imate.logdet / traceinv / trace / schatten(A, method='slq', gpu=True).
 *
 * Plain C types only: host pointers and sizes in, status codes out. Every
 * call is blocking; the library copies host arrays during the call and keeps
 * no host pointer. Device memory lives in slq_op handles (one HIP stream per
 * handle). Calls on one handle are not re-entrant; separate handles may be
 * used from separate threads.
 *
 * Each entry point names the reference interface it replaces
 * (imate 0.25.2, paths relative to the reference root).
 */
#ifndef SLQ_HIP_H_
#define SLQ_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (negative = error). The message of the last error of the
 * calling thread is returned by slq_last_error(). Replaces the abort() and
 * assert() exits of imate/_cu_linear_operator/cu_linear_operator.cu:77-85,
 * 251-270 and imate/_c_trace_estimator/c_orthogonalization.cpp. */
#define SLQ_OK 0
#define SLQ_E_ARG -1       /* invalid argument (ValueError / TypeError) */
#define SLQ_E_NODEVICE -2  /* no HIP device */
#define SLQ_E_OOM -3       /* device allocation failed */
#define SLQ_E_HIP -4       /* any other HIP runtime error */
#define SLQ_E_NOTSPD -5    /* a Ritz value left the domain of f (e.g. log) */

/* Value types (the matrix data type; imate accepts float32/float64 on GPU,
 * py_cu_matrix.pyx:136-142). */
#define SLQ_F64 0
#define SLQ_F32 1

/* Matrix functions of imate/functions: applied as f(theta^p) in the
 * quadrature, c_trace_estimator.cpp:616-617. */
#define SLQ_FUNC_LOG 0      /* functions/logarithm.cpp:42-45 (logdet) */
#define SLQ_FUNC_INV 1      /* functions/inverse.cpp:41-44   (traceinv) */
#define SLQ_FUNC_IDENT 2    /* functions/identity.cpp:41-44  (trace) */
/* The functions below take coefficients in func_params[]:              */
#define SLQ_FUNC_EXP 3      /* exp(c x), c = [0]; exponential.cpp (trexp)      */
#define SLQ_FUNC_LINFRAC 4  /* (a x + b) / (c x + d), [0..3]; homographic.cpp
                               (trlinfrac)                                    */
#define SLQ_FUNC_INDICATOR 5 /* 1 on [a, b], else 0, [0..1]; indicator.cpp
                               (eigencount)                                   */
#define SLQ_FUNC_GAUSSIAN 6 /* normal density, mu = [0], sigma = [1];
                               gaussian.cpp (density)                         */

/* Precision modes of the recurrences. */
#define SLQ_PREC_NATIVE 0   /* vectors and SpMM in the matrix dtype, FP64 recurrences */
#define SLQ_PREC_MIXED 1    /* FP32 SpMM / basis with FP64 alpha, beta and eigensolve */

typedef struct slq_op slq_op;

/* Last error message of the calling thread (always NUL-terminated). */
int slq_last_error(char* buf, size_t len);

/* Device count and properties of device 0.
 * Replaces pycuLinearOperator.get_device_properties
 * (imate/_cu_linear_operator/py_cu_linear_operator.pyx:195) and
 * imate/_cuda_utilities/query_device.cu:35-48. */
int slq_device_properties(int* num_devices, int* num_multiprocessors,
                          int* num_threads_per_multiprocessor,
                          char* name, size_t name_len);

/* Upload a CSR matrix (sorted column indices) to the device.
 * Replaces pycuMatrix / cuCSRMatrix<T> construction and the host-to-device
 * upload (imate/_cu_linear_operator/py_cu_matrix.pyx:118-200,365-407;
 * cu_csr_matrix.cu:52-84,164-224), with 64-bit row pointers instead of the
 * reference's int32 (types.h:60). dtype is SLQ_F64 or SLQ_F32 and gives the
 * element type of `data`. */
int slq_op_create_csr(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int64_t* indptr, const int32_t* indices,
                      const void* data, int dtype, slq_op** out);

/* Upload a CSC matrix: indptr[ncols + 1], row indices[nnz] (any order
 * within a column; repeated entries are summed by the products) and
 * data[nnz]. The handle holds the matrix as sorted CSR (a counting sort on
 * the host, once) and from there runs every path of a CSR handle: 16-row
 * tiles on the matrix cores, the device transpose for Golub-Kahn, A + tB.
 * slq_op_set_values on it takes values in this CSC order. Replaces the CSC
 * branch of pycuMatrix and cuCSCMatrix<T> (imate/_cu_linear_operator/
 * py_cu_matrix.pyx:118-200; cu_csc_matrix.cu:52-84,223-459, whose dot is a
 * cuSPARSE CSC SpMV per vector). */
int slq_op_create_csc(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int64_t* indptr, const int32_t* indices,
                      const void* data, int dtype, slq_op** out);

/* Dense matrix handle from host row-major data[nrows][ncols] (dtype
 * SLQ_F64 or SLQ_F32). The SpMM runs on the matrix cores with operand rows
 * staged through LDS; A + tI over it is allowed (slq_op_create_affine with
 * B == NULL), a general B is not (use CSR handles). Replaces the dense
 * branch of pycuMatrix and cuDenseMatrix (imate/_cu_linear_operator/
 * py_cu_matrix.pyx:249,306; cu_dense_matrix.cu:38-218, whose dot and
 * transpose_dot call cuBLAS gemv per vector). */
int slq_op_create_dense(int64_t nrows, int64_t ncols, const void* data, int dtype,
                        slq_op** out);

/* Affine operator A + tB over existing handles (B == NULL means B = I, for
 * which the Ritz values of A + tI follow from one Lanczos run by a shift).
 * The new handle references A's (and B's) device arrays; destroy it before
 * them. With a general B (CSR, same shape and dtype) the first estimate tiles
 * the union of both patterns (for gram = 1 also of their transposes) and
 * each t gathers A + tB into those tiles on the device (SLQ_PAIR_TILES=0
 * keeps two CSR products, A x + t B x). Replaces pycuAffineMatrixFunction / cuCSRAffineMatrixFunction
 * (imate/_cu_linear_operator/py_cu_affine_matrix_function.pyx:31,171;
 * cu_csr_affine_matrix_function.cu:29-211). */
int slq_op_create_affine(slq_op* A, slq_op* B, slq_op** out);

/* Replace the values of a matrix handle, keeping its sparsity pattern:
 * data[nnz] in the handle's dtype and CSR order of slq_op_create_csr (CSC
 * order for a handle made by slq_op_create_csc). The
 * tiled layout is refilled on the device; the transpose and FP32 copies are
 * rebuilt on next use. Affine handles made from this matrix see the new
 * values. Fails with SLQ_E_ARG when the CSR had duplicate entries.
 * No counterpart in the reference (imate builds a new pycuMatrix for new
 * data, py_cu_matrix.pyx:118-200): it serves families of matrices with one
 * pattern, such as a compactly supported kernel over its length scale. */
int slq_op_set_values(slq_op* op, const void* data);

/* A seeded sparse Gaussian matrix generated on the device (the rectangular
 * Golub-Kahn workload): row r has one entry in each of per_row column strata
 * [floor(s ncols / per_row), floor((s + 1) ncols / per_row)), at a hashed
 * column of the stratum (distinct, ascending), with an N(0, 1) value
 * (Box-Muller of hashed uniforms); row pointers r * per_row. No counterpart
 * in the reference (imate's sample matrices are host-built,
 * imate/sample_matrices/toeplitz.py, correlation_matrix.py): a 2^20 x 2^18
 * matrix at 1 % density has 2.75e9 entries, too many to build on the host. */
/* Value family of a CSR matrix handle: per-entry FP64 coefficients d, w, c
 * (in the CSR value order of slq_op_create_csr) are uploaded once; then
 * slq_op_family_values(op, scale) sets every value to
 * w_k exp(-d_k / scale) + c_k on the device (rounded once to the matrix
 * type) and refreshes the handle as slq_op_set_values does, with no host
 * values and no upload. For the kernel-matrix families of the learned
 * surrogate (values exp(-d / l) w(d) of one pattern over the length scale
 * l). No reference counterpart: imate builds every member on the host. */
int slq_op_family_attach(slq_op* op, const double* d, const double* w, const double* c);
int slq_op_family_values(slq_op* op, double scale);

int slq_op_create_sparse_gaussian(int64_t nrows, int64_t ncols, int32_t per_row,
                                  int64_t seed, int dtype, slq_op** out);

/* Copy a matrix handle's CSR back to the host: indptr[nrows + 1], and, when
 * non-NULL, indices[nnz] and data[nnz] (dtype of the handle). The reverse of
 * slq_op_create_csr (no counterpart in the reference, which keeps the host
 * copy: py_cu_matrix.pyx:118-200). */
int slq_op_get_csr(slq_op* op, int64_t* indptr, int32_t* indices, void* data);

/* Shape, non-zeros and dtype of a handle (pycuLinearOperator get_num_rows /
 * get_num_columns, py_cu_linear_operator.pyx:95-125). */
int slq_op_shape(slq_op* op, int64_t* nrows, int64_t* ncols, int64_t* nnz,
                 int32_t* dtype);

/* Free a handle (pycuMatrix.__dealloc__, cu_csr_matrix.cu:94-154). */
int slq_op_destroy(slq_op* op);

/* SLQ trace estimation of f((A + t B)^p) (or of (A^T A)^p with gram != 0,
 * by Golub-Kahn bidiagonalization).
 * Replaces pycu_trace_estimator / cuTraceEstimator<T>::cu_trace_estimator
 * (imate/_cu_trace_estimator/py_cu_trace_estimator.pyx:36-141,263-376;
 * cu_trace_estimator.cu:197-310,419-~670). Arguments have the meaning of
 * imate's trace_estimator (imate/_trace_estimator/trace_estimator.pyx:38-390):
 *   params[num_inquiries]  t values of an affine operator (ignored otherwise)
 *   func, exponent         f and p of f(theta^p)
 *   func_params            coefficients of f (SLQ_FUNC_EXP .. GAUSSIAN;
 *                          may be NULL for log, inverse and identity)
 *   orthogonalize          0 none, <0 or >= lanczos_degree full, else window
 *   seed                   >= 0 reproducible; < 0 uses the clock
 * Probe vectors are the Rademacher vectors of imate's thread-0 xoshiro256**
 * stream (sample i uses words [i*W, (i+1)*W), W = floor(len/64) + 1), i.e. the
 * reference's reproducible configuration with one device.
 * block_size: probes per device block (0 = library default);
 * precision: SLQ_PREC_NATIVE or SLQ_PREC_MIXED.
 * Outputs (caller-allocated host arrays, like trace_estimator.pyx:151-174):
 *   trace[q], error[q], samples[max_num_samples][q] (row-major, processing
 *   order, NaN where not computed), processed_samples_indices[max],
 *   num_samples_used[q], num_outliers[q], converged[q], *all_converged,
 *   *alg_wall_time (seconds, the sampling loop only). */
int slq_estimate(slq_op* op, const double* params, int32_t num_inquiries,
                 int32_t func, const double* func_params, int32_t gram,
                 double exponent,
                 int32_t orthogonalize, int64_t seed, int32_t lanczos_degree,
                 double lanczos_tol, int32_t min_num_samples,
                 int32_t max_num_samples, double error_atol, double error_rtol,
                 double confidence_level, double outlier_significance_level,
                 int32_t block_size, int32_t precision,
                 double* trace, double* error, double* samples,
                 int32_t* processed_samples_indices, int32_t* num_samples_used,
                 int32_t* num_outliers, int32_t* converged,
                 int32_t* all_converged, double* alg_wall_time);

/* slq_estimate without device allocation: same arguments, outputs and
 * kernels (results equal slq_estimate's bit for bit), with the per-call
 * workspace taken from a fixed 256 MiB array in device memory, one per device
 * for the life of the process, instead of the handle's allocated workspace.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops, code that forbids hipMalloc). Limits:
 *   - the workspace of the estimate must fit 256 MiB (plain FP64 Lanczos up
 *     to about 120k rows at 64 probes per block; larger needs, SLQ_E_ARG);
 *   - the derived forms an estimate reads (the transpose for gram, FP32
 *     copies for SLQ_PREC_MIXED, A + tB tiles) must already exist on the
 *     handle, built by an earlier slq_estimate of the same form (else
 *     SLQ_E_ARG); a plain CSR, dense or A + tI handle needs none;
 *   - profiling must be off on the handle (else SLQ_E_ARG).
 * Every call on one device shares the array: calls that overlap on one
 * device (from several host threads) race. */
int slq_estimate_noalloc(slq_op* op, const double* params, int32_t num_inquiries,
                         int32_t func, const double* func_params, int32_t gram,
                         double exponent, int32_t orthogonalize, int64_t seed,
                         int32_t lanczos_degree, double lanczos_tol,
                         int32_t min_num_samples, int32_t max_num_samples,
                         double error_atol, double error_rtol,
                         double confidence_level, double outlier_significance_level,
                         int32_t block_size, int32_t precision,
                         double* trace, double* error, double* samples,
                         int32_t* processed_samples_indices, int32_t* num_samples_used,
                         int32_t* num_outliers, int32_t* converged,
                         int32_t* all_converged, double* alg_wall_time);

/* Batched SLQ over `count` independent matrix handles (distinct handles, one
 * device), each with a fixed number of samples (no early stop): per problem
 * trace[i], error[i] (after the outlier pass) and samples[i][num_samples]
 * (may be NULL). use_graph != 0 captures every kernel of every problem into
 * one hipGraph (one branch per problem, so uneven problems run side by side)
 * and launches it once; use_graph == 0 runs the same kernels problem after
 * problem (a per-problem launch loop). *run_time is the launch-to-completion
 * time (s), *setup_time the preparation and capture time before it.
 * Replaces a Python loop of imate.logdet(..., method='slq', gpu=True) calls
 * over separate matrices; imate has no batched entry point. */
int slq_estimate_batch(slq_op** ops, int32_t count, int32_t func,
                       const double* func_params, int32_t gram, double exponent,
                       int32_t orthogonalize, int64_t seed,
                       int32_t lanczos_degree, double lanczos_tol,
                       int32_t num_samples, double confidence_level,
                       double outlier_significance_level, int32_t precision,
                       int32_t use_graph, double* trace, double* error,
                       double* samples, double* run_time, double* setup_time);

/* Autotuner: times one 64-probe block of Lanczos steps over candidate launch
 * grids for the SpMM and for the streaming kernels, then over the SpMM's
 * operand form (the 16-row tiles with 1, 2 or 4 wavefronts per tile, or the
 * CSR rows), keeps the fastest in the handle for slq_estimate_batch and
 * caches it per (rows, non-zeros per row, value type, tile rows). Outputs
 * the chosen caps when non-NULL. The other
 * entry points keep the library's default grids on a tuned handle, so their
 * reduction order, and with it every result bit, does not depend on tuning.
 * Square matrix or A + tI handles only (Lanczos is what is timed).
 * No reference counterpart (imate has no tunable kernels). */
int slq_autotune(slq_op* op, int32_t precision, int32_t* spmm_grid,
                 int32_t* stream_grid);

/* Launch configuration of a handle's kernels: grid caps of the SpMM and of
 * the streaming kernels (0 .. 1024), wavefronts per tile of the
 * matrix-core SpMM (1, 2 or 4) and the SpMM's operand form (tile_mode 0:
 * the row tiles where they were built, 1: the CSR rows); 0 keeps the
 * library default (environment SLQ_GRID / SLQ_STREAM_GRID / SLQ_SPMM_SPLIT,
 * else the built-in). Every setting gives results within rounding of every
 * other; for a fixed setting they are reproducible bit for bit (the
 * reduction order follows the grid, the split and the form). No reference
 * counterpart (imate has no tunable kernels). */
int slq_op_set_launch(slq_op* op, int32_t spmm_grid, int32_t stream_grid,
                      int32_t spmm_split, int32_t tile_mode);
int slq_op_get_launch(slq_op* op, int32_t* spmm_grid, int32_t* stream_grid,
                      int32_t* spmm_split, int32_t* tile_mode);

/* ---- learned surrogate (no reference counterpart) ----
 * A 2 -> 128 -> 128 -> 1 SiLU MLP used as a surrogate of SLQ estimates
 * tr f(A(l) + tI) over (t, kernel scale l). Parameters are a flat FP32 array
 * of slq_mlp_num_params() values: W1[128][2], b1[128], W2[128][128],
 * b2[128], W3[128], b3. Inputs x[B][2], targets t[B]; loss = mean squared
 * error. imate's own surrogates over t are its interpolators
 * (imate/interpolator, restated on the host in imate_amd/interpolator). */
int slq_mlp_num_params(void);
/* Loss and gradient of all parameters at `params` (one forward + backward). */
int slq_mlp_grad(const float* params, const float* x, const float* target,
                 int32_t B, float* loss, float* grad);
/* `steps` full-batch Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on
 * the device; params updated in place, losses[steps] when non-NULL. */
int slq_mlp_train(float* params, const float* x, const float* target,
                  int32_t B, int32_t steps, float lr, float* losses);
/* y[Q] = MLP(x[Q][2]); bf16 != 0 runs the 128x128 layer on the matrix cores
 * in BF16 with FP32 accumulation. *seconds: kernel time. */
int slq_mlp_predict(const float* params, const float* x, int32_t Q,
                    int32_t bf16, float* y, double* seconds);

/* ---- component entry points (tests, calibration and benchmarks) ---- */

/* Block Lanczos form (an opt-in with no counterpart in imate, which runs
 * one Lanczos recurrence per probe: c_lanczos_tridiagonalization.cpp:117-235
 * with c_orthogonalization.cpp:125-224 as its re-orthogonalisation). The 64
 * probes of a block (consecutive samples of the thread-0 stream) share one
 * Krylov basis, fully re-orthogonalised by matrix-core GEMMs against all
 * stored basis blocks; lanczos_degree = m block steps give a 64 m x 64 m
 * block tridiagonal and per-probe block Gauss quadrature samples, fed to
 * the reference's convergence test like slq_estimate's. The blocks that
 * cannot reach min_num_samples run as a group whose eigensolves are one
 * batched rocSOLVER call (loaded on first use; the block tridiagonals take
 * up to 1 GiB of workspace). Arguments and outputs as
 * slq_estimate, without gram (Lanczos only), orthogonalize (always full),
 * lanczos_tol and block_size (64). Requires 64 m <= the matrix size. */
int slq_estimate_block_krylov(slq_op* op, const double* params, int32_t num_inquiries,
                              int32_t func, const double* func_params, double exponent,
                              int64_t seed, int32_t lanczos_degree,
                              int32_t min_num_samples, int32_t max_num_samples,
                              double error_atol, double error_rtol,
                              double confidence_level, double outlier_significance_level,
                              int32_t precision, double* trace, double* error,
                              double* samples, int32_t* processed_samples_indices,
                              int32_t* num_samples_used, int32_t* num_outliers,
                              int32_t* converged, int32_t* all_converged,
                              double* alg_wall_time);

/* `count` consecutive Rademacher probes of length n, starting at sample
 * `first_sample` of the thread-0 stream of `seed`, written probe-major into
 * host `out[count][n]` (dtype SLQ_F64 / SLQ_F32), and the raw stream words
 * into `words[count][floor(n/64)+1]` when non-NULL.
 * Replaces RandomArrayGenerator<T>::generate_random_array
 * (imate/_random_generator/random_array_generator.cpp:62-165) fed by
 * RandomNumberGenerator (random_number_generator.cpp:70-153). */
int slq_rademacher(int64_t seed, int64_t n, int64_t first_sample,
                   int32_t count, int dtype, void* out, uint64_t* words);

/* Y = op * X (transpose != 0: op^T * X) for a block of b vectors stored
 * row-major X[len][b], in the op's dtype, host buffers.
 * Replaces cuCSRMatrix::dot / transpose_dot and the affine variants
 * (imate/_cu_linear_operator/cu_csr_matrix.cu:277-318,373-414;
 * cu_csr_affine_matrix_function.cu:123-155), one vector at a time there. */
int slq_op_spmm(slq_op* op, double t, int32_t b, int32_t transpose,
                const void* X, void* Y);

/* Recurrence coefficients of `count` consecutive probes (from first_sample
 * of the thread-0 stream of seed) at parameter t: alpha[count][m],
 * beta[count][m] (FP64) and sizes[count].
 * Replaces cu_lanczos_tridiagonalization / cu_golub_kahn_bidiagonalization
 * (imate/_cu_trace_estimator/cu_lanczos_tridiagonalization.cu:119-248;
 * cu_golub_kahn_bidiagonalization.cu:113-242). */
int slq_recurrences(slq_op* op, double t, int32_t gram, int64_t seed,
                    int64_t first_sample, int32_t count, int32_t lanczos_degree,
                    double lanczos_tol, int32_t orthogonalize,
                    int32_t precision, double* alpha, double* beta,
                    int32_t* sizes);

/* Device eigensolve + Gauss quadrature of `count` recurrences:
 * theta/tau of the tridiagonal (gram == 0: alpha diagonal, beta sub-diagonal)
 * or of the bidiagonal's squared singular values (gram != 0), then
 * samples[count][q] = n * sum_i tau_i^2 f((theta_i - params[0] + params[j])^p)
 * (shift only when shift != 0; else one inquiry). Outputs theta/tau
 * [count][m] when non-NULL.
 * Replaces Diagonalization::eigh_tridiagonal / svd_bidiagonal (LAPACK ?stev /
 * ?bdsdc, imate/_c_trace_estimator/diagonalization.cpp:88-118,189-231) and the
 * quadrature of cu_trace_estimator.cu:~555-630. */
int slq_quadrature(const double* alpha, const double* beta,
                   const int32_t* sizes, int32_t count, int32_t m,
                   int32_t gram, int32_t func, const double* func_params,
                   double exponent, double n, const double* params, int32_t q,
                   int32_t shift,
                   double* samples, double* theta, double* tau);

/* Per-kernel timing with HIP events on the handle's stream.
 * enable != 0 starts collecting (and clears the totals). slq_profile_read
 * returns the summed duration (ms) and launch count of the SpMM kernel and of
 * the update kernel since enabling. No reference counterpart: imate times
 * only the whole loop (imate/_cuda_utilities/cuda_timer.cu:28-84). */
int slq_profile_enable(slq_op* op, int32_t enable);
int slq_profile_read(slq_op* op, double* spmm_ms, int64_t* spmm_launches,
                     double* update_ms, int64_t* update_launches);
/* Completed Lanczos / Golub-Kahn steps of the samples processed by
 * slq_estimate since profiling was enabled (each sample's recurrence length,
 * shorter than lanczos_degree where it broke down; summed over the
 * recurrences a sample ran). Like imate's info['convergence'], no reference
 * counterpart for the count itself. */
int slq_profile_steps(slq_op* op, int64_t* steps);

/* Accumulated device time (ms, HIP events around each launch),
 * launches and floating-point operations of the block Krylov form's
 * matrix-core GEMMs (the Q^T W products and the W -= Q H updates of its
 * re-orthogonalisation and Cholesky QR) while profiling is enabled on the
 * handle. No counterpart in the reference (its re-orthogonalisation is not
 * a GEMM: c_orthogonalization.cpp:125-224). */
int slq_profile_gemm(slq_op* op, double* ms, int64_t* launches, double* flops);

/* Device copy of `bytes` bytes (a plain streaming kernel) timed with HIP
 * events over `iters` launches; used to calibrate rocprof byte counters and
 * the achievable HBM bandwidth. No reference counterpart. */
int slq_copy_bandwidth(int64_t bytes, int32_t iters, double* ms_per_launch);

/* ------------------------------------------------------------ CPU path
 *
 * imate's gpu=False branch (trace_estimator.pyx:246-277) on the host cores,
 * in this library's own form: probes in blocks of 8 stored row-interleaved,
 * one multiply of the operator per block and step, rows split over
 * num_threads worker threads with reductions summed in a fixed order (the
 * result does not depend on the thread count). Same probes as slq_estimate
 * (thread-0 stream), same eigensolve and quadrature code as the device path.
 * No HIP call: usable on a machine without a GPU. Host handles copy the
 * caller's arrays. */

#define SLQ_F128 2  /* long double (numpy float128); host handles only */

typedef struct slq_host_op slq_host_op;

/* Host copy of a CSR / CSC matrix (pycMatrix.set_csr_matrix_* /
 * set_csc_matrix_*, imate/_c_linear_operator/py_c_matrix.pyx:119-215; the
 * matrices of c_csr_matrix.cpp / c_csc_matrix.cpp). A CSC matrix is kept as
 * the CSR matrix it equals. dtype SLQ_F64, SLQ_F32 or SLQ_F128. */
int slq_host_op_create_csr(int64_t nrows, int64_t ncols, int64_t nnz,
                           const int64_t* indptr, const int32_t* indices,
                           const void* data, int32_t dtype, slq_host_op** out);
int slq_host_op_create_csc(int64_t nrows, int64_t ncols, int64_t nnz,
                           const int64_t* indptr, const int32_t* indices,
                           const void* data, int32_t dtype, slq_host_op** out);
/* Host copy of a row-major dense matrix (pycMatrix.set_dense_matrix_*,
 * py_c_matrix.pyx:216-380; c_dense_matrix.cpp). */
int slq_host_op_create_dense(int64_t nrows, int64_t ncols, const void* data,
                             int32_t dtype, slq_host_op** out);
/* A + t B; B == NULL means the identity (pycAffineMatrixFunction,
 * imate/_c_linear_operator/py_c_affine_matrix_function.pyx:172-305). B must
 * have A's shape and dtype. */
int slq_host_op_create_affine(slq_host_op* A, slq_host_op* B, slq_host_op** out);
int slq_host_op_destroy(slq_host_op* op);

/* SLQ trace estimation on the host: replaces pyc_trace_estimator /
 * cTraceEstimator<T>::c_trace_estimator
 * (imate/_c_trace_estimator/py_c_trace_estimator.pyx:39-172,
 * c_trace_estimator.cpp:191-307,410-~600). Arguments and outputs as
 * slq_estimate, with num_threads host threads (>= 1) in place of the
 * device's block size and precision. Samples equal those of imate's CPU
 * path with one thread, and those of slq_estimate, to rounding. */
int slq_estimate_host(slq_host_op* op, const double* params, int32_t num_inquiries,
                      int32_t func, const double* func_params, int32_t gram,
                      double exponent, int32_t orthogonalize, int64_t seed,
                      int32_t lanczos_degree, double lanczos_tol,
                      int32_t min_num_samples, int32_t max_num_samples,
                      double error_atol, double error_rtol, double confidence_level,
                      double outlier_significance_level, int32_t num_threads,
                      double* trace, double* error, double* samples,
                      int32_t* processed_samples_indices, int32_t* num_samples_used,
                      int32_t* num_outliers, int32_t* converged,
                      int32_t* all_converged, double* alg_wall_time);

#ifdef __cplusplus
}
#endif

#endif /* SLQ_HIP_H_ */
