/*
 * ORACLE — test infrastructure only. Never linked into the product library.
 *
 * Floating-point body of the CPU restatement of imate's SLQ inner loops,
 * instantiated once per precision by slq_oracle.c with
 *     REAL  = double | float
 *     SFX   = f64    | f32
 *     SQRT  = sqrt   | sqrtf
 *     EPS   = DBL_EPSILON | FLT_EPSILON
 * Every routine states the reference lines it follows; the arithmetic
 * (operand order, long-double accumulators, chunk-5 partial sums, reciprocal
 * scaling vs division) is kept identical so the results reproduce the
 * reference's own CPU code bit for bit when compiled with the same flags
 * (gcc -O3, baseline x86-64, no FMA).
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* cVectorOperations::copy_scaled_vector, c_vector_operations.cpp:80-107 */
static void FN(vcopy_scaled)(const REAL* x, int64_t n, REAL s, REAL* y)
{
    for (int64_t i = 0; i < n; ++i) y[i] = s * x[i];
}

/* cVectorOperations::subtract_scaled_vector, c_vector_operations.cpp:134-166
 * (returns early when the scale is exactly zero) */
static void FN(vsub_scaled)(const REAL* x, int64_t n, REAL s, REAL* y)
{
    if (s == 0.0) return;
    for (int64_t i = 0; i < n; ++i) y[i] -= s * x[i];
}

/* cVectorOperations::inner_product, c_vector_operations.cpp:203-244:
 * chunks of five products summed in REAL, added into a long double. */
static REAL FN(vdot)(const REAL* a, const REAL* b, int64_t n)
{
    long double acc = 0.0;
    int64_t chunk = 5;
    int64_t nc = n - (n % chunk);
    for (int64_t i = 0; i < nc; i += chunk)
    {
        acc += a[i] * b[i] +
               a[i+1] * b[i+1] +
               a[i+2] * b[i+2] +
               a[i+3] * b[i+3] +
               a[i+4] * b[i+4];
    }
    for (int64_t i = nc; i < n; ++i) acc += a[i] * b[i];
    return (REAL) acc;
}

/* cVectorOperations::euclidean_norm, c_vector_operations.cpp:280-321 */
static REAL FN(vnorm)(const REAL* a, int64_t n)
{
    long double acc = 0.0;
    int64_t chunk = 5;
    int64_t nc = n - (n % chunk);
    for (int64_t i = 0; i < nc; i += chunk)
    {
        acc += a[i] * a[i] +
               a[i+1] * a[i+1] +
               a[i+2] * a[i+2] +
               a[i+3] * a[i+3] +
               a[i+4] * a[i+4];
    }
    for (int64_t i = nc; i < n; ++i) acc += a[i] * a[i];
    return SQRT((REAL) acc);
}

/* cVectorOperations::normalize_vector_in_place, c_vector_operations.cpp:337-370
 * (divides by the norm) */
static REAL FN(vnormalize)(REAL* a, int64_t n)
{
    REAL nrm = FN(vnorm)(a, n);
    for (int64_t i = 0; i < n; ++i) a[i] /= nrm;
    return nrm;
}

/* cVectorOperations::normalize_vector_and_copy, c_vector_operations.cpp:388-422 */
static REAL FN(vnormalize_copy)(const REAL* a, int64_t n, REAL* out)
{
    REAL nrm = FN(vnorm)(a, n);
    for (int64_t i = 0; i < n; ++i) out[i] = a[i] / nrm;
    return nrm;
}

/* cMatrixOperations::csr_matvec, c_matrix_operations.cpp:469-494:
 * per-row long-double sum of REAL products. */
static void FN(csr_matvec)(const orc_csr* A, const REAL* x, REAL* y)
{
    const REAL* data = (const REAL*) A->data;
    for (int64_t row = 0; row < A->nrows; ++row)
    {
        long double sum = 0.0;
        for (int64_t k = A->indptr[row]; k < A->indptr[row+1]; ++k)
            sum += data[k] * x[A->indices[k]];
        y[row] = (REAL) sum;
    }
}

/* cMatrixOperations::csr_matvec_plus, c_matrix_operations.cpp:540-571 */
static void FN(csr_matvec_plus)(const orc_csr* A, const REAL* x, REAL alpha,
                               REAL* y)
{
    if (alpha == 0.0) return;
    const REAL* data = (const REAL*) A->data;
    for (int64_t row = 0; row < A->nrows; ++row)
    {
        long double sum = 0.0;
        for (int64_t k = A->indptr[row]; k < A->indptr[row+1]; ++k)
            sum += data[k] * x[A->indices[k]];
        y[row] += alpha * (REAL) sum;
    }
}

/* cMatrixOperations::csr_transposed_matvec, c_matrix_operations.cpp:606-635:
 * zero the output, then scatter row by row in REAL. */
static void FN(csr_tmatvec)(const orc_csr* A, const REAL* x, REAL* y)
{
    const REAL* data = (const REAL*) A->data;
    for (int64_t c = 0; c < A->ncols; ++c) y[c] = 0.0;
    for (int64_t row = 0; row < A->nrows; ++row)
        for (int64_t k = A->indptr[row]; k < A->indptr[row+1]; ++k)
            y[A->indices[k]] += data[k] * x[row];
}

/* cMatrixOperations::csr_transposed_matvec_plus (same file, after :635) */
static void FN(csr_tmatvec_plus)(const orc_csr* A, const REAL* x, REAL alpha,
                                REAL* y)
{
    if (alpha == 0.0) return;
    const REAL* data = (const REAL*) A->data;
    for (int64_t row = 0; row < A->nrows; ++row)
        for (int64_t k = A->indptr[row]; k < A->indptr[row+1]; ++k)
            y[A->indices[k]] += alpha * data[k] * x[row];
}

/* Operator y = (A + t B) x.
 * Plain matrix: cCSRMatrix::dot, c_csr_matrix.cpp:137-149.
 * Affine, B = I: cCSRAffineMatrixFunction::dot, c_csr_affine_matrix_function.cpp
 *   (A.dot, then _add_scaled_vector = subtract_scaled_vector with -t over
 *   min(rows, cols) entries, c_affine_matrix_function.cpp).
 * Affine, general B: A.dot then B.dot_plus(x, t). */
static void FN(op_dot)(const orc_op* op, const REAL* x, REAL* y)
{
    FN(csr_matvec)(&op->A, x, y);
    if (op->kind == ORC_OP_AFFINE_I)
    {
        int64_t mn = op->A.nrows < op->A.ncols ? op->A.nrows : op->A.ncols;
        FN(vsub_scaled)(x, mn, (REAL) -op->t, y);
    }
    else if (op->kind == ORC_OP_AFFINE_B)
    {
        FN(csr_matvec_plus)(&op->B, x, (REAL) op->t, y);
    }
}

/* Operator y = (A + t B)^T x (cCSRAffineMatrixFunction::transpose_dot). */
static void FN(op_tdot)(const orc_op* op, const REAL* x, REAL* y)
{
    FN(csr_tmatvec)(&op->A, x, y);
    if (op->kind == ORC_OP_AFFINE_I)
    {
        int64_t mn = op->A.nrows < op->A.ncols ? op->A.nrows : op->A.ncols;
        FN(vsub_scaled)(x, mn, (REAL) -op->t, y);
    }
    else if (op->kind == ORC_OP_AFFINE_B)
    {
        FN(csr_tmatvec_plus)(&op->B, x, (REAL) op->t, y);
    }
}

/* cOrthogonalization::gram_schmidt_process, c_orthogonalization.cpp:125-224:
 * newest-first modified Gram-Schmidt against the last num_ortho vectors of the
 * ring buffer V[num_vectors][n], dividing by |V_i|^2, skipping zero vectors and
 * (near-)identical vectors. */
static void FN(gram_schmidt)(const REAL* V, int64_t n, int32_t num_vectors,
                             int32_t last_vector, int32_t num_ortho, REAL* v)
{
    int32_t num_steps;
    if ((num_ortho == 0) || (num_vectors < 2)) return;
    else if ((num_ortho < 0) || (num_ortho > num_vectors)) num_steps = num_vectors;
    else num_steps = num_ortho;
    if ((int64_t) num_steps > n) num_steps = (int32_t) n;

    const REAL eps = EPS;
    for (int32_t step = 0; step < num_steps; ++step)
    {
        int32_t i;
        if ((last_vector % num_vectors) >= step) i = (last_vector % num_vectors) - step;
        else i = (last_vector % num_vectors) - step + num_vectors;

        REAL norm = FN(vnorm)(&V[n * i], n);
        if (norm < eps * sqrt((double) n))
        {
            orc_warn_count++;
            continue;
        }
        REAL inner = FN(vdot)(&V[n * i], v, n);
        REAL scale = inner / (norm * norm);
        if (fabs(fabs(scale) - 1.0) <= 2.0 * eps)
        {
            REAL norm_v = FN(vnorm)(v, n);
            REAL distance2 = norm_v * norm_v - 2.0 * inner + norm * norm;
            if (distance2 < 2.0 * eps * n) continue;
        }
        FN(vsub_scaled)(&V[n * i], n, scale, v);
    }
}

/* c_lanczos_tridiagonalization, c_lanczos_tridiagonalization.cpp:117-235.
 * Returns lanczos_size; alpha/beta hold the tridiagonal (beta[j] = T[j+1,j]).
 * `work` must hold n*(m + 1) REALs (ring buffer + r). */
static int32_t FN(lanczos)(const orc_op* op, const REAL* v, int64_t n, int32_t m,
                           REAL tol, int32_t ortho, REAL* alpha, REAL* beta,
                           REAL* work)
{
    int32_t buffer_size;
    if (ortho == 0 || ortho == 1) buffer_size = 2;
    else if ((ortho < 0) || (ortho > m)) buffer_size = m;
    else buffer_size = ortho;

    REAL* V = work;
    REAL* r = work + (int64_t) n * buffer_size;
    memcpy(r, v, sizeof(REAL) * n);
    REAL initial_beta = FN(vnorm)(r, n);
    int32_t size = 0;

    for (int32_t j = 0; j < m; ++j)
    {
        ++size;
        REAL* vj = &V[(int64_t)(j % buffer_size) * n];
        if (j == 0) FN(vcopy_scaled)(r, n, (REAL)(1.0 / initial_beta), vj);
        else FN(vcopy_scaled)(r, n, (REAL)(1.0 / beta[j-1]), vj);

        FN(op_dot)(op, vj, r);
        alpha[j] = FN(vdot)(vj, r, n);
        FN(vsub_scaled)(vj, n, alpha[j], r);
        if (j > 0)
            FN(vsub_scaled)(&V[(int64_t)((j-1) % buffer_size) * n], n, beta[j-1], r);

        if (ortho != 0)
        {
            int32_t num_ortho = (j < buffer_size) ? j + 1 : buffer_size;
            FN(gram_schmidt)(V, n, buffer_size, j % buffer_size, num_ortho, r);
        }

        beta[j] = FN(vnorm)(r, n);
        if (beta[j] < tol * sqrt((double) n)) break;
    }
    return size;
}

/* c_golub_kahn_bidiagonalization, c_golub_kahn_bidiagonalization.cpp:111-231,
 * with the sizes of a non-square operator kept apart: u vectors have nrows
 * entries and v vectors ncols (the reference sizes both by nrows).
 * `work` must hold (nrows + ncols) * m REALs. */
static int32_t FN(golub_kahn)(const orc_op* op, const REAL* v, int32_t m,
                              REAL tol, int32_t ortho, REAL* alpha, REAL* beta,
                              REAL* work)
{
    int64_t nr = op->A.nrows, nc = op->A.ncols;
    int32_t buffer_size;
    if (ortho == 0) buffer_size = 2;
    else if ((ortho < 0) || (ortho > m - 1)) buffer_size = m;
    else buffer_size = ortho + 1;

    REAL* U = work;
    REAL* V = work + (int64_t) nr * buffer_size;
    FN(vnormalize_copy)(v, nc, &V[0]);
    int32_t size = 0;
    int32_t num_ortho = 0;

    for (int32_t j = 0; j < m; ++j)
    {
        ++size;
        REAL* uj = &U[(int64_t)(j % buffer_size) * nr];
        REAL* vj = &V[(int64_t)(j % buffer_size) * nc];
        REAL* vn = &V[(int64_t)((j + 1) % buffer_size) * nc];

        FN(op_dot)(op, vj, uj);
        if (j > 0)
            FN(vsub_scaled)(&U[(int64_t)((j-1) % buffer_size) * nr], nr, beta[j-1], uj);

        if (ortho != 0)
        {
            num_ortho = (j < buffer_size) ? j : buffer_size - 1;
            if (j > 0)
                FN(gram_schmidt)(U, nr, buffer_size, (j - 1) % buffer_size,
                                 num_ortho, uj);
        }

        alpha[j] = FN(vnormalize)(uj, nr);

        FN(op_tdot)(op, uj, vn);
        FN(vsub_scaled)(vj, nc, alpha[j], vn);

        if (ortho != 0)
            FN(gram_schmidt)(V, nc, buffer_size, j % buffer_size, num_ortho, vn);

        beta[j] = FN(vnormalize)(vn, nc);
        if (beta[j] < tol * sqrt((double) nc)) break;
    }
    return size;
}

/* RandomArrayGenerator::generate_random_array, random_array_generator.cpp:62-165,
 * single-thread branch: floor(n/64) words fill 64 entries each (bit j of a word
 * -> entry 64*i + j, 1 -> +1, 0 -> -1), then one more word for the tail, which
 * is drawn even when n is a multiple of 64. */
static void FN(rademacher)(uint64_t* s, int64_t n, REAL* out)
{
    int64_t chunks = n / 64;
    for (int64_t i = 0; i < chunks; ++i)
    {
        uint64_t bits = xoshiro_next(s);
        for (int j = 0; j < 64; ++j)
            out[i * 64 + j] = (bits & ((uint64_t) 1 << j)) ? 1.0 : -1.0;
    }
    uint64_t bits = xoshiro_next(s);
    int j = 0;
    for (int64_t i = chunks * 64; i < n; ++i, ++j)
        out[i] = (bits & ((uint64_t) 1 << j)) ? 1.0 : -1.0;
}

/* ---- exported entry points (precision-suffixed) ---- */

void FN(orc_rademacher)(uint64_t* state, int64_t n, int32_t count, REAL* out)
{
    for (int32_t c = 0; c < count; ++c) FN(rademacher)(state, n, out + (int64_t) c * n);
}

void FN(orc_op_dot)(const orc_op* op, const REAL* x, REAL* y) { FN(op_dot)(op, x, y); }
void FN(orc_op_tdot)(const orc_op* op, const REAL* x, REAL* y) { FN(op_tdot)(op, x, y); }
REAL FN(orc_dot)(const REAL* a, const REAL* b, int64_t n) { return FN(vdot)(a, b, n); }
REAL FN(orc_norm)(const REAL* a, int64_t n) { return FN(vnorm)(a, n); }

void FN(orc_gram_schmidt)(const REAL* V, int64_t n, int32_t num_vectors,
                          int32_t last_vector, int32_t num_ortho, REAL* v)
{
    FN(gram_schmidt)(V, n, num_vectors, last_vector, num_ortho, v);
}

/* One recurrence (Lanczos or Golub-Kahn) from a given start vector. */
int32_t FN(orc_recurrence)(const orc_op* op, int32_t gram, const REAL* v,
                           int32_t m, REAL tol, int32_t ortho, REAL* alpha,
                           REAL* beta)
{
    int64_t len = gram ? (op->A.nrows + op->A.ncols) : (op->A.nrows + 1);
    REAL* work = (REAL*) malloc(sizeof(REAL) * len * (size_t)(m + 2));
    if (!work) return -1;
    int32_t k = gram ? FN(golub_kahn)(op, v, m, tol, ortho, alpha, beta, work)
                     : FN(lanczos)(op, v, op->A.nrows, m, tol, ortho, alpha, beta, work);
    free(work);
    return k;
}

/* Monte-Carlo driver for the recurrences of c_trace_estimator.cpp:191-307 and
 * _c_stochastic_lanczos_quadrature (:410-560): sample i draws its probe from
 * the stream of the thread that runs it (SplitMix64 seed, then jump() thread+1
 * times), of length ncols for Golub-Kahn and nrows for Lanczos. Only the
 * recurrence coefficients are produced here; eigensolve and quadrature run in
 * the Python half of the oracle (LAPACK through SciPy, like the reference).
 * With num_threads == 1 the samples are consecutive draws of stream 0, which is
 * the reproducible configuration. Samples are statically interleaved over the
 * threads (sample i on thread i % T) so the result does not depend on timing.
 * alpha/beta: [num_samples][m]; sizes: [num_samples]. */
int32_t FN(orc_slq_recurrences)(const orc_op* op, int32_t gram, int64_t seed,
                                int32_t num_threads, int32_t num_samples,
                                int32_t m, REAL tol, int32_t ortho,
                                REAL* alpha, REAL* beta, int32_t* sizes)
{
    if (num_threads < 1) num_threads = 1;
    int64_t plen = gram ? op->A.ncols : op->A.nrows;
    int64_t wlen = gram ? (op->A.nrows + op->A.ncols) : (op->A.nrows + 1);
    int failed = 0;

    #pragma omp parallel num_threads(num_threads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        uint64_t s[4];
        orc_thread_state(seed, tid, s);
        REAL* probe = (REAL*) malloc(sizeof(REAL) * plen);
        REAL* work = (REAL*) malloc(sizeof(REAL) * wlen * (size_t)(m + 2));
        if (!probe || !work)
        {
            #pragma omp atomic write
            failed = 1;
        }
        else
        {
            for (int32_t i = tid; i < num_samples; i += num_threads)
            {
                FN(rademacher)(s, plen, probe);
                REAL* a = alpha + (int64_t) i * m;
                REAL* b = beta + (int64_t) i * m;
                for (int32_t j = 0; j < m; ++j) { a[j] = 0.0; b[j] = 0.0; }
                sizes[i] = gram ? FN(golub_kahn)(op, probe, m, tol, ortho, a, b, work)
                                : FN(lanczos)(op, probe, op->A.nrows, m, tol, ortho, a, b, work);
            }
        }
        free(probe);
        free(work);
    }
    return failed ? -1 : 0;
}

#undef FN
#undef CAT
#undef CAT2
