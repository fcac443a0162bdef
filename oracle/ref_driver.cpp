/*
 * ORACLE — test infrastructure only.
 *
 * Thin C driver over the reference's own CPU sources, compiled by
 * oracle/Makefile straight from /root/reference into oracle/_ref/libimate_ref.so.
 * It exists to pin the oracle restatement (slq_oracle.c) bit for bit; it is
 * never shipped and nothing under imate_amd/ loads it.
 *
 * Only reference translation units that build from their own sources are
 * used (RNG, vector/matrix ops, CSR and affine operators, Gram-Schmidt,
 * Lanczos, Golub-Kahn, convergence tools, erf_inv). The LAPACK shim
 * (lapack_api.*, diagonalization.cpp) and the Monte-Carlo driver
 * (c_trace_estimator.cpp) need Cython-generated code and are not built; the
 * per-sample loop below follows c_trace_estimator.cpp:410-560 with one thread.
 */

#include <cstdint>
#include <cstring>
#include "_random_generator/random_number_generator.h"
#include "_random_generator/random_array_generator.h"
#include "_c_linear_operator/c_csr_matrix.h"
#include "_c_linear_operator/c_csr_affine_matrix_function.h"
#include "_c_trace_estimator/c_lanczos_tridiagonalization.h"
#include "_c_trace_estimator/c_golub_kahn_bidiagonalization.h"
#include "_c_trace_estimator/c_orthogonalization.h"
#include "_c_trace_estimator/convergence_tools.h"
#include "_c_trace_estimator/special_functions.h"

namespace {

template <typename T>
int run(const T* data, const int* indices, const int* indptr, int nrows,
        int ncols, int kind, double t, const T* bdata, const int* bindices,
        const int* bindptr, int gram, long long seed, int num_samples, int m,
        T tol, int ortho, T* alpha, T* beta, int* sizes, T* probes)
{
    cLinearOperator<T>* op = nullptr;
    T param = static_cast<T>(t);
    if (kind == 0)
        op = new cCSRMatrix<T>(data, indices, indptr, nrows, ncols);
    else if (kind == 1)
        op = new cCSRAffineMatrixFunction<T>(data, indices, indptr, nrows, ncols);
    else
        op = new cCSRAffineMatrixFunction<T>(data, indices, indptr, nrows, ncols,
                                             bdata, bindices, bindptr);
    if (kind != 0) op->set_parameters(&param);

    // Same probe length as the reference's driver: the number of rows.
    int n = nrows;
    RandomNumberGenerator rng(1, seed);
    T* v = new T[n];
    for (int i = 0; i < num_samples; ++i)
    {
        RandomArrayGenerator<T>::generate_random_array(rng, v, n, 1);
        if (probes) std::memcpy(probes + (long long) i * n, v, sizeof(T) * n);
        T* a = alpha + (long long) i * m;
        T* b = beta + (long long) i * m;
        for (int j = 0; j < m; ++j) { a[j] = 0; b[j] = 0; }
        if (gram)
            sizes[i] = c_golub_kahn_bidiagonalization<T>(op, v, n, m, tol, ortho, a, b);
        else
            sizes[i] = c_lanczos_tridiagonalization<T>(op, v, n, m, tol, ortho, a, b);
    }
    delete[] v;
    delete op;
    return 0;
}

}  // namespace

extern "C" {

int ref_slq_recurrences_f64(const double* data, const int* indices,
                            const int* indptr, int nrows, int ncols, int kind,
                            double t, const double* bdata, const int* bindices,
                            const int* bindptr, int gram, long long seed,
                            int num_samples, int m, double tol, int ortho,
                            double* alpha, double* beta, int* sizes,
                            double* probes)
{
    return run<double>(data, indices, indptr, nrows, ncols, kind, t, bdata,
                       bindices, bindptr, gram, seed, num_samples, m, tol,
                       ortho, alpha, beta, sizes, probes);
}

int ref_slq_recurrences_f32(const float* data, const int* indices,
                            const int* indptr, int nrows, int ncols, int kind,
                            double t, const float* bdata, const int* bindices,
                            const int* bindptr, int gram, long long seed,
                            int num_samples, int m, float tol, int ortho,
                            float* alpha, float* beta, int* sizes,
                            float* probes)
{
    return run<float>(data, indices, indptr, nrows, ncols, kind, t, bdata,
                      bindices, bindptr, gram, seed, num_samples, m, tol,
                      ortho, alpha, beta, sizes, probes);
}

/* Words of the stream of thread `thread_id` among `num_threads`. */
void ref_stream_words(long long seed, int num_threads, int thread_id,
                      long long count, unsigned long long* out)
{
    RandomNumberGenerator rng(num_threads, seed);
    for (long long i = 0; i < count; ++i) out[i] = rng.next(thread_id);
}

double ref_erf_inv(double x) { return erf_inv(x); }

void ref_gram_schmidt_f64(const double* V, int n, int num_vectors,
                          int last_vector, int num_ortho, double* v)
{
    cOrthogonalization<double>::gram_schmidt_process(V, n, num_vectors,
                                                     last_vector, num_ortho, v);
}

/* ConvergenceTools on a row-major samples[max][q] array. */
int ref_check_convergence(double* samples, int max_num_samples, int q,
                          int min_num_samples, const int* processed,
                          int num_processed, double confidence, double atol,
                          double rtol, double* error, int* num_used,
                          int* converged)
{
    double** rows = new double*[max_num_samples];
    for (int i = 0; i < max_num_samples; ++i) rows[i] = samples + (long long) i * q;
    int r = ConvergenceTools<double>::check_convergence(
        rows, min_num_samples, q, processed, num_processed, confidence, atol,
        rtol, error, num_used, converged);
    delete[] rows;
    return r;
}

void ref_average_estimates(double* samples, int max_num_samples, int q,
                           double confidence, double outlier,
                           const int* num_used, const int* processed,
                           int* num_outliers, double* trace, double* error)
{
    double** rows = new double*[max_num_samples];
    for (int i = 0; i < max_num_samples; ++i) rows[i] = samples + (long long) i * q;
    ConvergenceTools<double>::average_estimates(confidence, outlier, q,
                                                max_num_samples, num_used,
                                                processed, rows, num_outliers,
                                                trace, error);
    delete[] rows;
}

}  // extern "C"
