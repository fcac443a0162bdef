// CPU path of the SLQ estimator: imate's gpu=False branch
// (imate/_trace_estimator/trace_estimator.pyx:246-277, run by
// cTraceEstimator<T>::c_trace_estimator, c_trace_estimator.cpp:191-307, and
// _c_stochastic_lanczos_quadrature, :410-~600) on the host cores.
//
// Layout and schedule. imate runs one probe per OpenMP thread, each thread
// streaming the whole matrix for its own probe. Here the probes of a block
// (kLanes of them, consecutive samples) are stored row-interleaved,
// X[i * kLanes + b], so one pass over the matrix multiplies all of them: a
// CSR row gathers kLanes contiguous values per stored entry. The rows of
// every pass are cut into chunks of kChunk rows that the threads of a team
// take round robin; each chunk writes its own partial sums of the block's
// inner products, and those are added chunk by chunk in a fixed order, so
// the result does not depend on the number of threads.
//
// Arithmetic. Vectors are kept in the matrix type; products and sums run in
// double (long double for float128), like the device path; the recurrence
// follows the same steps as the device path (Lanczos with two-pass
// classical Gram-Schmidt against the window of imate's orthogonalize
// option, Golub-Kahn with the reference's separate u and v windows), and the
// eigensolve and quadrature are the device's own code (ql.hpp). Probes are
// the Rademacher vectors of imate's thread-0 xoshiro256** stream, sample i
// at word offset i * (floor(len / 64) + 1): the samples equal those of
// imate's CPU path with one thread and those of the device path, to
// rounding. Samples are processed in order with the reference's convergence
// test after each one.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>
#include <vector>

#include "estimator.hpp"
#include "ql.hpp"

namespace slq {
namespace host {

constexpr int kLanes = 8;            // probes per block
constexpr int64_t kChunk = 2048;     // rows per chunk of a pass

int value_size(int dtype)
{
    switch (dtype)
    {
        case SLQ_F64: return 8;
        case SLQ_F32: return 4;
        case SLQ_F128: return (int) sizeof(long double);
        default: return 0;
    }
}

// A matrix in host memory: CSR (sorted or not) or dense row-major.
struct Matrix {
    int64_t nrows = 0, ncols = 0;
    int dtype = SLQ_F64;
    bool dense = false;
    std::vector<int64_t> indptr;
    std::vector<int32_t> indices;
    std::vector<unsigned char> values;   // nnz values, or nrows * ncols
    std::mutex mu;                       // guards trans
    std::shared_ptr<Matrix> trans;       // transpose, made on first use

    template <typename T>
    const T* data() const
    {
        return reinterpret_cast<const T*>(values.data());
    }
    int64_t nnz() const { return dense ? nrows * ncols : (int64_t) indices.size(); }
};

// Transpose of M (CSR rows of the result ascend in their column index).
std::shared_ptr<Matrix> transpose_of(const Matrix& M)
{
    auto T = std::make_shared<Matrix>();
    T->nrows = M.ncols;
    T->ncols = M.nrows;
    T->dtype = M.dtype;
    T->dense = M.dense;
    const int vs = value_size(M.dtype);
    T->values.resize(M.values.size());
    if (M.dense)
    {
        for (int64_t r = 0; r < M.nrows; ++r)
            for (int64_t c = 0; c < M.ncols; ++c)
                std::memcpy(&T->values[(size_t) (c * M.nrows + r) * vs],
                            &M.values[(size_t) (r * M.ncols + c) * vs], vs);
        return T;
    }
    const int64_t nnz = M.nnz();
    T->indptr.assign(M.ncols + 1, 0);
    T->indices.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) T->indptr[M.indices[k] + 1] += 1;
    for (int64_t c = 0; c < M.ncols; ++c) T->indptr[c + 1] += T->indptr[c];
    std::vector<int64_t> next(T->indptr.begin(), T->indptr.end() - 1);
    for (int64_t r = 0; r < M.nrows; ++r)
        for (int64_t k = M.indptr[r]; k < M.indptr[r + 1]; ++k)
        {
            const int64_t pos = next[M.indices[k]]++;
            T->indices[pos] = (int32_t) r;
            std::memcpy(&T->values[(size_t) pos * vs], &M.values[(size_t) k * vs], vs);
        }
    return T;
}

const Matrix& transposed(Matrix& M)
{
    std::lock_guard<std::mutex> lock(M.mu);
    if (!M.trans) M.trans = transpose_of(M);
    return *M.trans;
}

// A team of threads that run one function together (the calling thread is
// member 0) and return when every member has finished it.
class Team {
public:
    explicit Team(int n) : n_(n)
    {
        try
        {
            for (int t = 1; t < n; ++t) th_.emplace_back([this, t] { loop(t); });
        }
        catch (...)
        {
            // the threads already started must end before the members go
            shutdown();
            throw;
        }
    }
    ~Team() { shutdown(); }
    int size() const { return n_; }

    void run(const std::function<void(int)>& f)
    {
        if (n_ == 1)
        {
            f(0);
            return;
        }
        {
            std::lock_guard<std::mutex> l(mu_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        go_.notify_all();
        f(0);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return pending_ == 0; });
        job_ = nullptr;
    }

    // body(c, r0, r1) for every chunk c of [0, rows), chunks round robin
    template <typename F>
    void chunks(int64_t rows, F body)
    {
        const int64_t nch = (rows + kChunk - 1) / kChunk;
        run([&](int tid) {
            for (int64_t c = tid; c < nch; c += n_)
                body(c, c * kChunk, std::min(rows, (c + 1) * kChunk));
        });
    }

private:
    void shutdown()
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
            ++gen_;
        }
        go_.notify_all();
        for (auto& t : th_)
            if (t.joinable()) t.join();
        th_.clear();
    }

    void loop(int t)
    {
        uint64_t seen = 0;
        for (;;)
        {
            const std::function<void(int)>* f = nullptr;
            {
                std::unique_lock<std::mutex> l(mu_);
                go_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                f = job_;
            }
            (*f)(t);
            std::lock_guard<std::mutex> l(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    const std::function<void(int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

int64_t num_chunks(int64_t rows) { return (rows + kChunk - 1) / kChunk; }

template <typename T>
struct AccOf {
    using type = double;
};
template <>
struct AccOf<long double> {
    using type = long double;
};

// y = (A + t B) x (or its transpose) on a block: the matrix applied (A or
// A^T), B likewise or none, and the identity shift of A + tI.
template <typename T>
struct OpView {
    using Acc = typename AccOf<T>::type;
    const Matrix* A = nullptr;
    const Matrix* B = nullptr;
    bool shift = false;
    int64_t shift_len = 0;
    Acc t = 0;
};

template <typename T, typename Acc>
inline void row_sum(const Matrix& M, int64_t i, const T* X, Acc* s)
{
    const T* v = M.data<T>();
    if (M.dense)
    {
        const T* row = v + i * M.ncols;
        for (int64_t c = 0; c < M.ncols; ++c)
        {
            const Acc a = row[c];
            const T* x = X + c * kLanes;
            for (int b = 0; b < kLanes; ++b) s[b] += a * (Acc) x[b];
        }
        return;
    }
    const int32_t* idx = M.indices.data();
    for (int64_t k = M.indptr[i], e = M.indptr[i + 1]; k < e; ++k)
    {
        const Acc a = v[k];
        const T* x = X + (int64_t) idx[k] * kLanes;
        for (int b = 0; b < kLanes; ++b) s[b] += a * (Acc) x[b];
    }
}

// Y[r0, r1) = op X; adds dot_with . Y per lane into part when dot_with is
// given (part[b] += sum_i dot_with[i][b] Y[i][b]).
template <typename T>
void product_rows(const OpView<T>& op, int64_t r0, int64_t r1, const T* X, T* Y)
{
    using Acc = typename AccOf<T>::type;
    for (int64_t i = r0; i < r1; ++i)
    {
        Acc s[kLanes] = {};
        row_sum(*op.A, i, X, s);
        if (op.B)
        {
            Acc u[kLanes] = {};
            row_sum(*op.B, i, X, u);
            for (int b = 0; b < kLanes; ++b) s[b] += op.t * u[b];
        }
        else if (op.shift && i < op.shift_len)
            for (int b = 0; b < kLanes; ++b) s[b] += op.t * (Acc) X[i * kLanes + b];
        for (int b = 0; b < kLanes; ++b) Y[i * kLanes + b] = (T) s[b];
    }
}

// Per-chunk partial sums of K quantities per lane, reduced in chunk order.
template <typename Acc>
struct Partials {
    std::vector<Acc> v;
    int64_t nch = 0;
    int K = 0;
    void reset(int64_t chunks, int k)
    {
        nch = chunks;
        K = k;
        v.assign((size_t) chunks * k * kLanes, Acc(0));
    }
    Acc* at(int64_t c) { return v.data() + (size_t) c * K * kLanes; }
    void reduce(Acc* out) const   // out[K][kLanes]
    {
        for (int i = 0; i < K * kLanes; ++i) out[i] = 0;
        for (int64_t c = 0; c < nch; ++c)
            for (int i = 0; i < K * kLanes; ++i) out[i] += v[(size_t) c * K * kLanes + i];
    }
};

// Outputs of one recurrence of a block.
struct BlockOut {
    std::vector<double> alpha, beta;   // [m][kLanes]
    int32_t sizes[kLanes];             // 0 = ran all m steps
};

// Vectors and partial sums of one estimate, sized once.
template <typename T>
struct Work {
    using Acc = typename AccOf<T>::type;
    std::vector<T> X0;                 // probe block [len][kLanes]
    std::vector<T> W;                  // product / residual [rows][kLanes]
    std::vector<T> U, V;               // basis slots
    Partials<Acc> part;
    std::vector<Acc> h1, h2;           // Gram-Schmidt coefficients [K][kLanes]
};

// Two-pass classical Gram-Schmidt of R (rows n) against the K basis vectors
// in slots last, last-1, ..., last-K+1 of S (slot stride n * kLanes), then
// the squared norm of R per lane in nrm2.
template <typename T>
void cgs2(Team& team, int64_t n, T* R, const T* S, int last, int K, Work<T>& w,
          typename AccOf<T>::type* nrm2)
{
    using Acc = typename AccOf<T>::type;
    const size_t st = (size_t) n * kLanes;
    const int64_t nch = num_chunks(n);
    auto slot = [&](int k) { return S + (size_t) (last - k) * st; };
    if (K > 0)
    {
        w.h1.assign((size_t) K * kLanes, 0);
        w.h2.assign((size_t) K * kLanes, 0);
        // h1 = S^T R
        w.part.reset(nch, K);
        team.chunks(n, [&](int64_t c, int64_t r0, int64_t r1) {
            Acc* p = w.part.at(c);
            for (int k = 0; k < K; ++k)
            {
                const T* v = slot(k);
                for (int64_t i = r0; i < r1; ++i)
                    for (int b = 0; b < kLanes; ++b)
                        p[k * kLanes + b] += (Acc) v[i * kLanes + b] * (Acc) R[i * kLanes + b];
            }
        });
        w.part.reduce(w.h1.data());
        // R -= S h1; h2 = S^T R
        w.part.reset(nch, K);
        team.chunks(n, [&](int64_t c, int64_t r0, int64_t r1) {
            Acc* p = w.part.at(c);
            for (int64_t i = r0; i < r1; ++i)
            {
                Acc r[kLanes];
                for (int b = 0; b < kLanes; ++b) r[b] = R[i * kLanes + b];
                for (int k = 0; k < K; ++k)
                {
                    const T* v = slot(k) + i * kLanes;
                    for (int b = 0; b < kLanes; ++b) r[b] -= w.h1[k * kLanes + b] * (Acc) v[b];
                }
                for (int b = 0; b < kLanes; ++b) R[i * kLanes + b] = (T) r[b];
            }
            for (int k = 0; k < K; ++k)
            {
                const T* v = slot(k);
                for (int64_t i = r0; i < r1; ++i)
                    for (int b = 0; b < kLanes; ++b)
                        p[k * kLanes + b] += (Acc) v[i * kLanes + b] * (Acc) R[i * kLanes + b];
            }
        });
        w.part.reduce(w.h2.data());
    }
    // R -= S h2; |R|^2
    w.part.reset(nch, 1);
    team.chunks(n, [&](int64_t c, int64_t r0, int64_t r1) {
        Acc* p = w.part.at(c);
        for (int64_t i = r0; i < r1; ++i)
        {
            Acc r[kLanes];
            for (int b = 0; b < kLanes; ++b) r[b] = R[i * kLanes + b];
            for (int k = 0; k < K; ++k)
            {
                const T* v = slot(k) + i * kLanes;
                for (int b = 0; b < kLanes; ++b) r[b] -= w.h2[k * kLanes + b] * (Acc) v[b];
            }
            for (int b = 0; b < kLanes; ++b)
            {
                const T x = (T) r[b];
                R[i * kLanes + b] = x;
                p[b] += (Acc) x * (Acc) x;
            }
        }
    });
    w.part.reduce(nrm2);
}

// Y = s Y per lane (rows n).
template <typename T>
void scale_rows(Team& team, int64_t n, const typename AccOf<T>::type* s, const T* X, T* Y)
{
    using Acc = typename AccOf<T>::type;
    team.chunks(n, [&](int64_t, int64_t r0, int64_t r1) {
        for (int64_t i = r0; i < r1; ++i)
            for (int b = 0; b < kLanes; ++b)
                Y[i * kLanes + b] = (T) (s[b] * (Acc) X[i * kLanes + b]);
    });
}

// Lanczos tridiagonalisation of a block (c_lanczos_tridiagonalization.cpp:
// 117-235): v_0 = probe / sqrt(n); w = A v_j; alpha_j = v_j . w;
// r = w - alpha_j v_j - beta_{j-1} v_{j-1}; with orthogonalize != 0 r is
// re-orthogonalised against the window of the newest basis vectors (two
// passes of classical Gram-Schmidt); beta_j = |r|; a probe stops once
// beta_j < tol sqrt(n) (its size is then j + 1).
template <typename T>
void lanczos(Team& team, const OpView<T>& op, int64_t n, int m, double tol, int ortho,
             int live_lanes, Work<T>& w, BlockOut& out)
{
    using Acc = typename AccOf<T>::type;
    const bool reorth = ortho != 0;
    int buffer = 2;
    if (ortho == 0 || ortho == 1) buffer = 2;
    else if (ortho < 0 || ortho > m) buffer = m;
    else buffer = ortho;
    const int nslots = reorth ? m : 2;
    const size_t st = (size_t) n * kLanes;
    w.V.resize(st * nslots);
    w.W.resize(st);
    auto slot = [&](int j) { return w.V.data() + (size_t) (reorth ? j : (j & 1)) * st; };
    const int64_t nch = num_chunks(n);
    bool live[kLanes];
    Acc ca[kLanes], cb[kLanes], sc[kLanes];
    for (int b = 0; b < kLanes; ++b)
    {
        live[b] = b < live_lanes;
        out.sizes[b] = 0;
        cb[b] = 0;
        sc[b] = live[b] ? Acc(1) / std::sqrt((Acc) n) : Acc(0);
    }
    out.alpha.assign((size_t) m * kLanes, 0.0);
    out.beta.assign((size_t) m * kLanes, 0.0);
    const double tol_sqrt_n = tol * std::sqrt((double) n);
    scale_rows<T>(team, n, sc, w.X0.data(), slot(0));
    T* Wp = w.W.data();
    for (int j = 0; j < m; ++j)
    {
        const T* Vj = slot(j);
        const T* Vp = j > 0 ? slot(j - 1) : nullptr;
        // w = A v_j, alpha_j = v_j . w
        w.part.reset(nch, 1);
        team.chunks(n, [&](int64_t c, int64_t r0, int64_t r1) {
            product_rows<T>(op, r0, r1, Vj, Wp);
            Acc* p = w.part.at(c);
            for (int64_t i = r0; i < r1; ++i)
                for (int b = 0; b < kLanes; ++b)
                    p[b] += (Acc) Vj[i * kLanes + b] * (Acc) Wp[i * kLanes + b];
        });
        Acc a[kLanes];
        w.part.reduce(a);
        for (int b = 0; b < kLanes; ++b)
        {
            ca[b] = live[b] ? a[b] : Acc(0);
            out.alpha[(size_t) j * kLanes + b] = (double) ca[b];
        }
        // r = w - alpha_j v_j - beta_{j-1} v_{j-1}
        w.part.reset(nch, 1);
        team.chunks(n, [&](int64_t c, int64_t r0, int64_t r1) {
            Acc* p = w.part.at(c);
            for (int64_t i = r0; i < r1; ++i)
                for (int b = 0; b < kLanes; ++b)
                {
                    Acc r = (Acc) Wp[i * kLanes + b] - ca[b] * (Acc) Vj[i * kLanes + b];
                    if (Vp) r -= cb[b] * (Acc) Vp[i * kLanes + b];
                    const T x = (T) r;
                    Wp[i * kLanes + b] = x;
                    p[b] += (Acc) x * (Acc) x;
                }
        });
        Acc nrm2[kLanes];
        if (reorth)
        {
            const int K = j < buffer ? j + 1 : buffer;
            cgs2<T>(team, n, Wp, w.V.data(), j, K, w, nrm2);
        }
        else
            w.part.reduce(nrm2);
        for (int b = 0; b < kLanes; ++b)
        {
            const Acc bt = std::sqrt(nrm2[b]);
            out.beta[(size_t) j * kLanes + b] = live[b] ? (double) bt : 0.0;
            if (live[b] && (double) bt < tol_sqrt_n)
            {
                out.sizes[b] = j + 1;
                live[b] = false;
            }
            cb[b] = live[b] ? bt : Acc(0);
            sc[b] = live[b] ? Acc(1) / bt : Acc(0);
        }
        if (j + 1 < m) scale_rows<T>(team, n, sc, Wp, slot(j + 1));
    }
}

// Golub-Kahn bidiagonalisation of a block (c_golub_kahn_bidiagonalization.
// cpp:111-231, u vectors of nrows entries and v vectors of ncols):
// v_0 = probe / sqrt(ncols); u_j = A v_j - beta_{j-1} u_{j-1} (then
// re-orthogonalised against the u window), alpha_j = |u_j|;
// v_{j+1} = A^T u_j - alpha_j v_j (re-orthogonalised against the v window),
// beta_j = |v_{j+1}|; a probe stops once beta_j < tol sqrt(ncols).
template <typename T>
void golub_kahn(Team& team, const OpView<T>& opA, const OpView<T>& opT, int64_t nr,
                int64_t nc, int m, double tol, int ortho, int live_lanes, Work<T>& w,
                BlockOut& out)
{
    using Acc = typename AccOf<T>::type;
    const bool reorth = ortho != 0;
    int buffer = 2;
    if (ortho == 0) buffer = 2;
    else if (ortho < 0 || ortho > m - 1) buffer = m;
    else buffer = ortho + 1;
    const size_t ust = (size_t) nr * kLanes, vst = (size_t) nc * kLanes;
    w.U.resize(ust * (reorth ? m : 2));
    w.V.resize(vst * (reorth ? m + 1 : 2));
    auto uslot = [&](int j) { return w.U.data() + (size_t) (reorth ? j : (j & 1)) * ust; };
    auto vslot = [&](int j) { return w.V.data() + (size_t) (reorth ? j : (j & 1)) * vst; };
    bool live[kLanes];
    Acc ca[kLanes], cb[kLanes], sc[kLanes];
    for (int b = 0; b < kLanes; ++b)
    {
        live[b] = b < live_lanes;
        out.sizes[b] = 0;
        cb[b] = 0;
        sc[b] = live[b] ? Acc(1) / std::sqrt((Acc) nc) : Acc(0);
    }
    out.alpha.assign((size_t) m * kLanes, 0.0);
    out.beta.assign((size_t) m * kLanes, 0.0);
    const double tol_sqrt_n = tol * std::sqrt((double) nc);
    scale_rows<T>(team, nc, sc, w.X0.data(), vslot(0));
    for (int j = 0; j < m; ++j)
    {
        T* Uj = uslot(j);
        const T* Up = j > 0 ? uslot(j - 1) : nullptr;
        const T* Vj = vslot(j);
        T* Vn = vslot(j + 1);
        const int K = j < buffer ? j : buffer - 1;
        // u_j = A v_j - beta_{j-1} u_{j-1}
        w.part.reset(num_chunks(nr), 1);
        team.chunks(nr, [&](int64_t c, int64_t r0, int64_t r1) {
            product_rows<T>(opA, r0, r1, Vj, Uj);
            Acc* p = w.part.at(c);
            for (int64_t i = r0; i < r1; ++i)
                for (int b = 0; b < kLanes; ++b)
                {
                    Acc u = (Acc) Uj[i * kLanes + b];
                    if (Up) u -= cb[b] * (Acc) Up[i * kLanes + b];
                    const T x = (T) u;
                    Uj[i * kLanes + b] = x;
                    p[b] += (Acc) x * (Acc) x;
                }
        });
        Acc nrm2[kLanes];
        if (reorth)
            cgs2<T>(team, nr, Uj, w.U.data(), j - 1, j > 0 ? K : 0, w, nrm2);
        else
            w.part.reduce(nrm2);
        for (int b = 0; b < kLanes; ++b)
        {
            const Acc al = std::sqrt(nrm2[b]);
            ca[b] = live[b] ? al : Acc(0);
            out.alpha[(size_t) j * kLanes + b] = (double) ca[b];
            sc[b] = live[b] ? Acc(1) / al : Acc(0);
        }
        scale_rows<T>(team, nr, sc, Uj, Uj);
        // v_{j+1} = A^T u_j - alpha_j v_j
        w.part.reset(num_chunks(nc), 1);
        team.chunks(nc, [&](int64_t c, int64_t r0, int64_t r1) {
            product_rows<T>(opT, r0, r1, Uj, Vn);
            Acc* p = w.part.at(c);
            for (int64_t i = r0; i < r1; ++i)
                for (int b = 0; b < kLanes; ++b)
                {
                    const Acc v = (Acc) Vn[i * kLanes + b] - ca[b] * (Acc) Vj[i * kLanes + b];
                    const T x = (T) v;
                    Vn[i * kLanes + b] = x;
                    p[b] += (Acc) x * (Acc) x;
                }
        });
        if (reorth)
            cgs2<T>(team, nc, Vn, w.V.data(), j, K, w, nrm2);
        else
            w.part.reduce(nrm2);
        for (int b = 0; b < kLanes; ++b)
        {
            const Acc bt = std::sqrt(nrm2[b]);
            out.beta[(size_t) j * kLanes + b] = live[b] ? (double) bt : 0.0;
            if (live[b] && (double) bt < tol_sqrt_n)
            {
                out.sizes[b] = j + 1;
                live[b] = false;
            }
            cb[b] = live[b] ? bt : Acc(0);
            sc[b] = live[b] ? Acc(1) / bt : Acc(0);
        }
        scale_rows<T>(team, nc, sc, Vn, Vn);
    }
}

}  // namespace host
}  // namespace slq

struct slq_host_op {
    int kind = 0;   // 0 matrix, 1 A + tI, 2 A + tB
    std::shared_ptr<slq::host::Matrix> A, B;
};

namespace slq {
namespace host {

constexpr int kMatrix = 0, kAffineI = 1, kAffineB = 2;

// CSR arrays of an nrows x ncols matrix: indptr[0] = 0, non-decreasing, ends
// at nnz; column indices inside [0, ncols).
int check_csr(const char* who, int64_t nrows, int64_t ncols, int64_t nnz,
              const int64_t* indptr, const int32_t* indices)
{
    if (indptr[0] != 0 || indptr[nrows] != nnz)
    {
        set_error("%s: indptr must start at 0 and end at nnz", who);
        return SLQ_E_ARG;
    }
    for (int64_t r = 0; r < nrows; ++r)
        if (indptr[r + 1] < indptr[r])
        {
            set_error("%s: indptr must be non-decreasing", who);
            return SLQ_E_ARG;
        }
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || indices[k] >= ncols)
        {
            set_error("%s: index %lld of entry %lld is out of range", who,
                      (long long) indices[k], (long long) k);
            return SLQ_E_ARG;
        }
    return SLQ_OK;
}

int make_sparse(const char* who, bool csc, int64_t nrows, int64_t ncols, int64_t nnz,
                const int64_t* indptr, const int32_t* indices, const void* data, int dtype,
                slq_host_op** out)
{
    if (!out || nrows < 1 || ncols < 1 || nnz < 0 || !indptr || (nnz > 0 && (!indices || !data)) ||
        value_size(dtype) == 0 || nrows >= (int64_t) 1 << 31 || ncols >= (int64_t) 1 << 31)
    {
        set_error("%s: invalid argument", who);
        return SLQ_E_ARG;
    }
    // a CSC matrix is the CSR of its transpose
    const int64_t rows = csc ? ncols : nrows, cols = csc ? nrows : ncols;
    const int rc = check_csr(who, rows, cols, nnz, indptr, indices);
    if (rc != SLQ_OK) return rc;
    try
    {
        auto M = std::make_shared<Matrix>();
        M->nrows = rows;
        M->ncols = cols;
        M->dtype = dtype;
        M->indptr.assign(indptr, indptr + rows + 1);
        M->indices.assign(indices, indices + nnz);
        const size_t bytes = (size_t) nnz * value_size(dtype);
        M->values.resize(bytes);
        if (bytes) std::memcpy(M->values.data(), data, bytes);
        auto op = new slq_host_op;
        op->A = csc ? transpose_of(*M) : M;
        *out = op;
    }
    catch (const std::bad_alloc&)
    {
        set_error("%s: out of host memory", who);
        return SLQ_E_OOM;
    }
    return SLQ_OK;
}

}  // namespace host
}  // namespace slq

using slq::set_error;
namespace sh = slq::host;

extern "C" {

int slq_host_op_create_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* indptr,
                           const int32_t* indices, const void* data, int32_t dtype,
                           slq_host_op** out)
{
    return sh::make_sparse("slq_host_op_create_csr", false, nrows, ncols, nnz, indptr, indices,
                           data, dtype, out);
}

int slq_host_op_create_csc(int64_t nrows, int64_t ncols, int64_t nnz, const int64_t* indptr,
                           const int32_t* indices, const void* data, int32_t dtype,
                           slq_host_op** out)
{
    return sh::make_sparse("slq_host_op_create_csc", true, nrows, ncols, nnz, indptr, indices,
                           data, dtype, out);
}

int slq_host_op_create_dense(int64_t nrows, int64_t ncols, const void* data, int32_t dtype,
                             slq_host_op** out)
{
    if (!out || !data || nrows < 1 || ncols < 1 || sh::value_size(dtype) == 0)
    {
        set_error("slq_host_op_create_dense: invalid argument");
        return SLQ_E_ARG;
    }
    try
    {
        auto M = std::make_shared<sh::Matrix>();
        M->nrows = nrows;
        M->ncols = ncols;
        M->dtype = dtype;
        M->dense = true;
        const size_t bytes = (size_t) nrows * ncols * sh::value_size(dtype);
        M->values.resize(bytes);
        std::memcpy(M->values.data(), data, bytes);
        auto op = new slq_host_op;
        op->A = M;
        *out = op;
    }
    catch (const std::bad_alloc&)
    {
        set_error("slq_host_op_create_dense: out of host memory");
        return SLQ_E_OOM;
    }
    return SLQ_OK;
}

int slq_host_op_create_affine(slq_host_op* A, slq_host_op* B, slq_host_op** out)
{
    if (!out || !A || A->kind != sh::kMatrix || (B && B->kind != sh::kMatrix))
    {
        set_error("slq_host_op_create_affine: invalid argument");
        return SLQ_E_ARG;
    }
    if (B && (B->A->nrows != A->A->nrows || B->A->ncols != A->A->ncols ||
              B->A->dtype != A->A->dtype))
    {
        set_error("slq_host_op_create_affine: B must have the shape and dtype of A");
        return SLQ_E_ARG;
    }
    auto op = new (std::nothrow) slq_host_op;
    if (!op)
    {
        set_error("slq_host_op_create_affine: out of host memory");
        return SLQ_E_OOM;
    }
    op->kind = B ? sh::kAffineB : sh::kAffineI;
    op->A = A->A;
    if (B) op->B = B->A;
    *out = op;
    return SLQ_OK;
}

int slq_host_op_destroy(slq_host_op* op)
{
    delete op;
    return SLQ_OK;
}

}  // extern "C"

namespace slq {
namespace host {

struct HostArgs {
    const double* params;
    int q, func;
    const double* fp;
    int gram;
    double exponent;
    int ortho;
    int64_t seed;
    int m;
    double tol;
    int min_s, max_s;
    double atol, rtol, confidence, outlier;
    int threads;
    double *trace, *error, *samples;
    int32_t *processed, *num_used, *num_outliers, *converged, *all_converged;
    double* wall;
};

template <typename T>
int estimate(slq_host_op* op, HostArgs a)
{
    Matrix& A = *op->A;
    const int64_t nr = A.nrows, nc = A.ncols;
    const int64_t len = a.gram ? nc : nr;
    const int64_t W = len / 64 + 1;      // words per probe
    const int q = a.q;
    const bool relation = op->kind == kAffineI;
    const int required = relation ? 1 : q;
    const int T_ = std::max(1, std::min<int>(a.threads, (int) std::max<int64_t>(
                                                            1, num_chunks(std::max(nr, nc)))));
    if (a.seed < 0) a.seed = clock_seed();

    std::unique_ptr<Team> team;
    Work<T> w;
    std::vector<uint64_t> words;
    try
    {
        team.reset(new Team(T_));
        w.X0.resize((size_t) len * kLanes);
        words.resize((size_t) W * kLanes);
    }
    catch (const std::bad_alloc&)
    {
        set_error("slq_estimate_host: out of host memory");
        return SLQ_E_OOM;
    }
    catch (const std::system_error&)
    {
        set_error("slq_estimate_host: cannot start %d threads", T_);
        return SLQ_E_ARG;
    }
    const Matrix* At = nullptr;
    const Matrix* Bt = nullptr;
    if (a.gram)
    {
        At = &transposed(A);
        if (op->B) Bt = &transposed(*op->B);
    }
    auto view = [&](double t, bool trans) {
        OpView<T> v;
        v.A = trans ? At : &A;
        v.B = op->kind == kAffineB ? (trans ? Bt : op->B.get()) : nullptr;
        v.shift = op->kind == kAffineI;
        v.shift_len = std::min(nr, nc);
        v.t = (typename OpView<T>::Acc) t;
        return v;
    };

    uint64_t base[4];
    thread_state(a.seed, 0, base);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (int64_t i = 0; i < (int64_t) a.max_s * q; ++i) a.samples[i] = nan;
    for (int j = 0; j < q; ++j)
    {
        a.converged[j] = 0;
        a.num_used[j] = 0;
        a.num_outliers[j] = 0;
    }
    QuadArgs qa;
    std::memset(&qa, 0, sizeof(qa));
    qa.m = a.m;
    qa.gram = a.gram;
    qa.func = a.func;
    for (int k = 0; k < 4; ++k) qa.fp[k] = a.fp ? a.fp[k] : 0.0;
    qa.exponent = a.exponent;
    qa.n = (double) len;
    qa.params = relation ? a.params : nullptr;
    qa.q = q;
    qa.shift = relation ? 1 : 0;
    std::vector<double> qd(2 * (size_t) a.m), qe(2 * (size_t) a.m), qz(2 * (size_t) a.m);
    std::vector<double> block((size_t) kLanes * q);
    BlockOut out;

    int nproc = 0, all_conv = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t s0 = 0; s0 < a.max_s && !all_conv; s0 += kLanes)
    {
        const int cnt = (int) std::min<int64_t>(kLanes, a.max_s - s0);
        // probes of samples s0 .. s0 + cnt - 1: each lane jumps to its word
        // offset of the thread-0 stream and draws W words
        team->run([&](int tid) {
            for (int b = tid; b < kLanes; b += team->size())
            {
                uint64_t* wb = words.data() + (size_t) b * W;
                if (b >= cnt)
                {
                    std::fill(wb, wb + W, 0);
                    continue;
                }
                uint64_t s[4];
                stream_at(base, (uint64_t) (s0 + b) * (uint64_t) W, s);
                for (int64_t k = 0; k < W; ++k) wb[k] = stream_next(s);
            }
        });
        team->chunks(len, [&](int64_t, int64_t r0, int64_t r1) {
            for (int64_t i = r0; i < r1; ++i)
                for (int b = 0; b < kLanes; ++b)
                {
                    const uint64_t wd = words[(size_t) b * W + (i >> 6)];
                    w.X0[i * kLanes + b] =
                        b < cnt ? (((wd >> (i & 63)) & 1) ? T(1) : T(-1)) : T(0);
                }
        });
        for (int j = 0; j < required; ++j)
        {
            if (!relation && a.converged[j]) continue;
            const double t = op->kind != kMatrix ? a.params[j] : 0.0;
            if (a.gram)
                golub_kahn<T>(*team, view(t, false), view(t, true), nr, nc, a.m, a.tol, a.ortho,
                              cnt, w, out);
            else
                lanczos<T>(*team, view(t, false), nr, a.m, a.tol, a.ortho, cnt, w, out);
            qa.j_out = j;
            for (int b = 0; b < cnt; ++b)
                quadrature_one(qa, out.alpha.data() + b, out.beta.data() + b, kLanes,
                               out.sizes[b], qd.data(), qe.data(), qz.data(), 1,
                               block.data() + (size_t) b * q, nullptr, nullptr);
        }
        for (int b = 0; b < cnt; ++b)
        {
            const int64_t i = s0 + b;
            if (all_conv) break;   // later samples stay NaN
            for (int j = 0; j < q; ++j)
                a.samples[i * q + j] = block[(size_t) b * q + j];
            if (relation && q > 1)
                for (int j = 0; j < q; ++j) a.converged[j] = 0;
            for (int j = 0; j < q; ++j)
                if (a.converged[j]) a.samples[i * q + j] = nan;
            a.processed[nproc++] = (int32_t) i;
            all_conv = check_convergence(a.samples, q, a.min_s, a.processed, nproc, a.confidence,
                                         a.atol, a.rtol, a.error, a.num_used, a.converged);
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (a.wall) *a.wall = std::chrono::duration<double>(t1 - t0).count();
    for (int i = nproc; i < a.max_s; ++i) a.processed[i] = 0;
    average_estimates(a.samples, q, a.confidence, a.outlier, a.num_used, a.processed,
                      a.num_outliers, a.trace, a.error);
    if (a.all_converged) *a.all_converged = all_conv;
    return SLQ_OK;
}

}  // namespace host
}  // namespace slq

extern "C" int slq_estimate_host(
    slq_host_op* op, const double* params, int32_t q, int32_t func, const double* func_params,
    int32_t gram, double exponent, int32_t orthogonalize, int64_t seed, int32_t lanczos_degree,
    double lanczos_tol, int32_t min_num_samples, int32_t max_num_samples, double error_atol,
    double error_rtol, double confidence_level, double outlier_significance_level,
    int32_t num_threads, double* trace, double* error, double* samples,
    int32_t* processed_samples_indices, int32_t* num_samples_used, int32_t* num_outliers,
    int32_t* converged, int32_t* all_converged, double* alg_wall_time)
{
    const char* who = "slq_estimate_host";
    if (!op || q < 1 || lanczos_degree < 1 || min_num_samples < 1 ||
        max_num_samples < min_num_samples || num_threads < 1 || !trace || !error ||
        !samples || !processed_samples_indices || !num_samples_used || !num_outliers ||
        !converged || !func_ok(func, func_params))
    {
        set_error("%s: invalid argument", who);
        return SLQ_E_ARG;
    }
    if (op->kind != sh::kMatrix && !params)
    {
        set_error("%s: an affine operator needs parameters", who);
        return SLQ_E_ARG;
    }
    if (op->kind == sh::kMatrix && q != 1)
    {
        set_error("%s: a matrix operator has exactly one inquiry", who);
        return SLQ_E_ARG;
    }
    if (!gram && op->A->nrows != op->A->ncols)
    {
        set_error("%s: Lanczos needs a square operator", who);
        return SLQ_E_ARG;
    }
    sh::HostArgs a{params, q, func, func_params, gram, exponent, orthogonalize, seed,
                   lanczos_degree, lanczos_tol, min_num_samples, max_num_samples, error_atol,
                   error_rtol, confidence_level, outlier_significance_level, num_threads,
                   trace, error, samples, processed_samples_indices, num_samples_used,
                   num_outliers, converged, all_converged, alg_wall_time};
    try
    {
        switch (op->A->dtype)
        {
            case SLQ_F64: return sh::estimate<double>(op, a);
            case SLQ_F32: return sh::estimate<float>(op, a);
            case SLQ_F128: return sh::estimate<long double>(op, a);
            default: break;
        }
    }
    catch (const std::bad_alloc&)
    {
        set_error("%s: out of host memory", who);
        return SLQ_E_OOM;
    }
    set_error("%s: unknown dtype", who);
    return SLQ_E_ARG;
}
