// C ABI of libslq_hip.so and the host side of the SLQ estimator.
//
// The Monte-Carlo loop follows imate's cTraceEstimator / cuTraceEstimator
// (imate/_c_trace_estimator/c_trace_estimator.cpp:191-307,
// imate/_cu_trace_estimator/cu_trace_estimator.cu:197-310) in its
// reproducible one-device configuration: samples are processed in order, the
// convergence test runs after every sample and sampling stops at the first
// sample after which every inquiry has converged. The device computes a block
// of 64 probes at a time; samples of a block that lie beyond the stopping
// sample are discarded, so every output equals the one-thread reference run.
// ConvergenceTools (convergence_tools.cpp:94-376) and erf_inv
// (special_functions.cpp:63-104) are restated here in C++.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "slq_internal.hpp"

namespace slq {

// ---------------------------------------------------------------- errors

namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int hip_status(hipError_t e, const char* what, const char* file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SLQ_E_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SLQ_E_NODEVICE;
    return SLQ_E_HIP;
}

// ---------------------------------------------------------------- storage

static void free_tiles(DevTiles& t)
{
    if (t.ptr) (void) hipFree(t.ptr);
    if (t.cols) (void) hipFree(t.cols);
    if (t.vals) (void) hipFree(t.vals);
    if (t.vals_f32) (void) hipFree(t.vals_f32);
    if (t.src) (void) hipFree(t.src);
    t = DevTiles();
}

MatrixStore::~MatrixStore()
{
    if (dn.data) (void) hipFree(dn.data);
    if (dn.trans) (void) hipFree(dn.trans);
    if (dn.data_f32) (void) hipFree(dn.data_f32);
    if (dn.trans_f32) (void) hipFree(dn.trans_f32);
    free_tiles(tiles);
    free_tiles(trans_tiles);
    if (csr.indptr) (void) hipFree(csr.indptr);
    if (csr.indices) (void) hipFree(csr.indices);
    if (csr.data) (void) hipFree(csr.data);
    if (trans.indptr) (void) hipFree(trans.indptr);
    if (trans.indices) (void) hipFree(trans.indices);
    if (trans.data) (void) hipFree(trans.data);
    if (data_f32) (void) hipFree(data_f32);
    if (trans_data_f32) (void) hipFree(trans_data_f32);
    if (family) (void) hipFree(family);
}

Workspace::~Workspace()
{
    if (base) (void) hipFree(base);
}

int Workspace::reserve(size_t need)
{
    if (need <= bytes) return SLQ_OK;
    if (base)
    {
        (void) hipFree(base);
        base = nullptr;
        bytes = 0;
    }
    SLQ_CHECK(hipMalloc(&base, need));
    bytes = need;
    return SLQ_OK;
}

// ---------------------------------------------------------------- tiles

// Position of the value (slab, row r) in the tile value array. R = 8: slab
// by slab, R values each. R = 16 (matrix-core layout): a group of four
// slabs is one MFMA A operand, lane l = 16 (slab % 4) + r; the two groups of
// each eight slabs are interleaved per lane ([slab / 8][lane][2]) so that a
// lane loads both of its A values with one access.
size_t tile_slot(int R, int64_t slab, int r)
{
    if (R != 16) return (size_t) slab * R + r;
    const int64_t pair = slab >> 3, half = (slab >> 2) & 1;
    const int lane = 16 * (int) (slab & 3) + r;
    return (size_t) pair * 128 + (size_t) lane * 2 + (size_t) half;
}

int build_tiles(int64_t nrows, int64_t ncols, const int64_t* indptr,
                const int32_t* indices, const void* data, int dtype, int R,
                double min_density, DevTiles& out, std::vector<int32_t>* src_host)
{
    (void) ncols;
    out = DevTiles();
    const int64_t nnz = indptr[nrows];
    if (nnz == 0) return SLQ_OK;
    const int64_t T = (nrows + R - 1) / R;
    std::vector<int64_t> tptr(T + 1, 0);
    std::vector<int32_t> cols;
    cols.reserve((size_t) (nnz / 2 + 16));
    std::vector<int32_t> buf;
    for (int64_t t = 0; t < T; ++t)
    {
        buf.clear();
        const int64_t r1 = std::min<int64_t>(nrows, (t + 1) * R);
        for (int64_t row = t * R; row < r1; ++row)
            buf.insert(buf.end(), indices + indptr[row], indices + indptr[row + 1]);
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        // 16-row tiles feed the matrix-core SpMM eight slabs (two MFMA
        // groups) at a time: pad each tile's slab list to a multiple of
        // eight with zero slabs that repeat its last column (a valid row of
        // the operand)
        if (R == 16 && !buf.empty())
            while (buf.size() % 8) buf.push_back(buf.back());
        cols.insert(cols.end(), buf.begin(), buf.end());
        tptr[t + 1] = (int64_t) cols.size();
    }
    const int64_t U = (int64_t) cols.size();
    const double density = (double) nnz / ((double) U * R);
    if (density < min_density) return SLQ_OK;
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    std::vector<char> vals((size_t) U * R * vb, 0);
    bool keep_src = nnz <= (int64_t) INT32_MAX;
    std::vector<int32_t> src(keep_src ? (size_t) U * R : 0, -1);
    for (int64_t t = 0; t < T; ++t)
    {
        const int32_t* cb = cols.data() + tptr[t];
        const int32_t* ce = cols.data() + tptr[t + 1];
        const int64_t r1 = std::min<int64_t>(nrows, (t + 1) * R);
        for (int64_t row = t * R; row < r1; ++row)
        {
            const int r = (int) (row - t * R);
            const int32_t* p = cb;
            for (int64_t k = indptr[row]; k < indptr[row + 1]; ++k)
            {
                p = std::lower_bound(p, ce, indices[k]);
                const size_t slot = tile_slot(R, p - cols.data(), r);
                if (!data)
                    ;
                else if (dtype == SLQ_F64)
                    ((double*) vals.data())[slot] += ((const double*) data)[k];
                else
                    ((float*) vals.data())[slot] += ((const float*) data)[k];
                if (keep_src)
                {
                    if (src[slot] >= 0) keep_src = false;   // duplicate entry
                    else src[slot] = (int32_t) k;
                }
            }
        }
    }
    DevTiles d;
    d.R = R;
    d.ntiles = T;
    d.nslabs = U;
    hipError_t e;
    if ((e = hipMalloc(&d.ptr, sizeof(int64_t) * (T + 1))) != hipSuccess ||
        (e = hipMalloc(&d.cols, sizeof(int32_t) * U)) != hipSuccess ||
        (e = hipMalloc(&d.vals, vals.size())) != hipSuccess ||
        (keep_src && (e = hipMalloc(&d.src, sizeof(int32_t) * src.size())) != hipSuccess) ||
        (e = hipMemcpy(d.ptr, tptr.data(), sizeof(int64_t) * (T + 1),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d.cols, cols.data(), sizeof(int32_t) * U,
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d.vals, vals.data(), vals.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (keep_src && (e = hipMemcpy(d.src, src.data(), sizeof(int32_t) * src.size(),
                                    hipMemcpyHostToDevice)) != hipSuccess))
    {
        free_tiles(d);
        return hip_status(e, "tile upload", __FILE__, __LINE__);
    }
    out = d;
    if (src_host)
    {
        if (keep_src) src_host->swap(src);
        else src_host->clear();
    }
    return SLQ_OK;
}

PairTiles::~PairTiles()
{
    if (borrowed) return;
    free_tiles(tiles);
    if (src_a) (void) hipFree(src_a);
    if (src_b) (void) hipFree(src_b);
}

namespace {

// Union of the patterns of A and B (host CSR arrays, square, same shape):
// per row the distinct columns in ascending order, and for each entry its
// position in A's and in B's value arrays (-1: absent). False when a matrix
// repeats a column within a row (its values would need summing) or when the
// union holds more than 2^31 - 1 entries.
bool union_pattern(int64_t nrows, const std::vector<int64_t>& ipa,
                   const std::vector<int32_t>& ixa, const std::vector<int64_t>& ipb,
                   const std::vector<int32_t>& ixb, std::vector<int64_t>& ipu,
                   std::vector<int32_t>& ixu, std::vector<int32_t>& ka,
                   std::vector<int32_t>& kb)
{
    if (ipa[nrows] + ipb[nrows] > (int64_t) INT32_MAX) return false;
    ipu.assign(nrows + 1, 0);
    ixu.clear();
    ka.clear();
    kb.clear();
    ixu.reserve(ipa[nrows] + ipb[nrows]);
    struct Entry { int32_t col; int32_t pos; bool in_b; };
    std::vector<Entry> row;
    for (int64_t r = 0; r < nrows; ++r)
    {
        row.clear();
        for (int64_t k = ipa[r]; k < ipa[r + 1]; ++k)
            row.push_back({ixa[k], (int32_t) k, false});
        for (int64_t k = ipb[r]; k < ipb[r + 1]; ++k)
            row.push_back({ixb[k], (int32_t) k, true});
        std::sort(row.begin(), row.end(), [](const Entry& x, const Entry& y) {
            return x.col != y.col ? x.col < y.col : (x.in_b < y.in_b);
        });
        for (size_t i = 0; i < row.size();)
        {
            const int32_t c = row[i].col;
            int32_t pa = -1, pb = -1;
            for (; i < row.size() && row[i].col == c; ++i)
            {
                int32_t& dst = row[i].in_b ? pb : pa;
                if (dst >= 0) return false;   // repeated column
                dst = row[i].pos;
            }
            ixu.push_back(c);
            ka.push_back(pa);
            kb.push_back(pb);
        }
        ipu[r + 1] = (int64_t) ixu.size();
    }
    return true;
}

}  // namespace

// Tiled form of A + tB (R of A's tiles), built once per affine operator on
// the host from device copies of both patterns. Leaves op->ab.tiles.R = 0
// (the two-stream CSR SpMM) when A has no tiled form, a pattern repeats a
// column or the union is too sparse for tiles.
int ensure_pair_tiles(slq_op* op, bool need_f32, bool trans)
{
    PairTiles& ab = trans ? op->ab_t : op->ab;
    if (!ab.tried)
    {
        const MatrixStore& A = *op->A;
        const MatrixStore& B = *op->B;
        if (trans && !(A.has_trans && B.has_trans))
        {
            set_error("internal: the pair tiles of the transposes need both transposes");
            return SLQ_E_ARG;
        }
        ab.tried = true;
        const int R = A.tiles.R;
        if (R == 0 || A.dense || B.dense) return SLQ_OK;
        const DevCSR& ca = trans ? A.trans : A.csr;
        const DevCSR& cb = trans ? B.trans : B.csr;
        const int64_t n = ca.nrows;
        SLQ_CHECK(hipStreamSynchronize(op->stream));   // transposes built on it
        std::vector<int64_t> ipa(n + 1), ipb(n + 1), ipu;
        std::vector<int32_t> ixa(ca.nnz > 0 ? ca.nnz : 1), ixb(cb.nnz > 0 ? cb.nnz : 1), ixu,
            ka, kb;
        SLQ_CHECK(hipMemcpy(ipa.data(), ca.indptr, sizeof(int64_t) * (n + 1),
                            hipMemcpyDeviceToHost));
        SLQ_CHECK(hipMemcpy(ipb.data(), cb.indptr, sizeof(int64_t) * (n + 1),
                            hipMemcpyDeviceToHost));
        if (ca.nnz > 0)
            SLQ_CHECK(hipMemcpy(ixa.data(), ca.indices, sizeof(int32_t) * ca.nnz,
                                hipMemcpyDeviceToHost));
        if (cb.nnz > 0)
            SLQ_CHECK(hipMemcpy(ixb.data(), cb.indices, sizeof(int32_t) * cb.nnz,
                                hipMemcpyDeviceToHost));
        if (!union_pattern(n, ipa, ixa, ipb, ixb, ipu, ixu, ka, kb)) return SLQ_OK;
        std::vector<int32_t> src;
        DevTiles t;
        SLQ_TRY(build_tiles(n, ca.ncols, ipu.data(), ixu.data(), nullptr, ca.dtype, R, 0.25, t,
                            &src));
        if (t.R == 0) return SLQ_OK;
        if (src.empty())
        {
            free_tiles(t);
            return SLQ_OK;
        }
        if (t.src)
        {
            (void) hipFree(t.src);
            t.src = nullptr;
        }
        // slot -> union entry -> position in A's / B's values
        std::vector<int32_t> sa(src.size()), sb(src.size());
        for (size_t i = 0; i < src.size(); ++i)
        {
            sa[i] = src[i] >= 0 ? ka[src[i]] : -1;
            sb[i] = src[i] >= 0 ? kb[src[i]] : -1;
        }
        hipError_t e;
        if ((e = hipMalloc(&ab.src_a, sizeof(int32_t) * sa.size())) != hipSuccess ||
            (e = hipMalloc(&ab.src_b, sizeof(int32_t) * sb.size())) != hipSuccess ||
            (e = hipMemcpy(ab.src_a, sa.data(), sizeof(int32_t) * sa.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(ab.src_b, sb.data(), sizeof(int32_t) * sb.size(),
                           hipMemcpyHostToDevice)) != hipSuccess)
        {
            free_tiles(t);
            if (ab.src_a) (void) hipFree(ab.src_a);
            if (ab.src_b) (void) hipFree(ab.src_b);
            ab.src_a = ab.src_b = nullptr;
            return hip_status(e, "pair tile upload", __FILE__, __LINE__);
        }
        ab.tiles = t;
    }
    if (need_f32 && ab.tiles.R > 0 && op->A->csr.dtype == SLQ_F64 && !ab.tiles.vals_f32)
        SLQ_CHECK(hipMalloc(&ab.tiles.vals_f32,
                            sizeof(float) * (size_t) (ab.tiles.nslabs * ab.tiles.R)));
    return SLQ_OK;
}

// ---------------------------------------------------------------- profiling

void prof_begin(slq_op* op, hipEvent_t* ev)
{
    *ev = nullptr;
    if (!op->prof.enabled) return;
    hipEvent_t e;
    if (!op->prof.pool.empty())
    {
        e = op->prof.pool.back();
        op->prof.pool.pop_back();
    }
    else if (hipEventCreate(&e) != hipSuccess)
        return;
    (void) hipEventRecord(e, op->stream);
    *ev = e;
}

void prof_end(slq_op* op, hipEvent_t ev0, int kind)
{
    if (!ev0) return;
    hipEvent_t e;
    if (!op->prof.pool.empty())
    {
        e = op->prof.pool.back();
        op->prof.pool.pop_back();
    }
    else if (hipEventCreate(&e) != hipSuccess)
        return;
    (void) hipEventRecord(e, op->stream);
    (kind == 0 ? op->prof.spmm_pending : op->prof.update_pending).push_back({ev0, e});
}

void prof_end_gemm(slq_op* op, hipEvent_t ev0, double flops)
{
    if (!ev0) return;
    hipEvent_t e;
    if (!op->prof.pool.empty())
    {
        e = op->prof.pool.back();
        op->prof.pool.pop_back();
    }
    else if (hipEventCreate(&e) != hipSuccess)
        return;
    (void) hipEventRecord(e, op->stream);
    op->prof.gemm_pending.push_back({ev0, e});
    op->prof.gemm_pending_flops.push_back(flops);
}

void prof_collect(slq_op* op)
{
    for (size_t k = 0; k < op->prof.gemm_pending.size(); ++k)
    {
        auto& pr = op->prof.gemm_pending[k];
        float ms = 0.f;
        if (hipEventSynchronize(pr.second) == hipSuccess &&
            hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess)
        {
            op->prof.gemm_ms += ms;
            op->prof.gemm_launches += 1;
            op->prof.gemm_flops += op->prof.gemm_pending_flops[k];
        }
        op->prof.pool.push_back(pr.first);
        op->prof.pool.push_back(pr.second);
    }
    op->prof.gemm_pending.clear();
    op->prof.gemm_pending_flops.clear();
    for (int kind = 0; kind < 2; ++kind)
    {
        auto& pend = kind == 0 ? op->prof.spmm_pending : op->prof.update_pending;
        for (auto& pr : pend)
        {
            float ms = 0.f;
            if (hipEventSynchronize(pr.second) == hipSuccess &&
                hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess)
            {
                if (kind == 0)
                {
                    op->prof.spmm_ms += ms;
                    op->prof.spmm_launches += 1;
                }
                else
                {
                    op->prof.update_ms += ms;
                    op->prof.update_launches += 1;
                }
            }
            op->prof.pool.push_back(pr.first);
            op->prof.pool.push_back(pr.second);
        }
        pend.clear();
    }
}

// ---------------------------------------------------------------- transpose

namespace {

__global__ void expand_rows_kernel(const int64_t* __restrict__ indptr, int64_t nrows,
                                   int32_t* __restrict__ rowidx)
{
    const int64_t row = (int64_t) blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (row >= nrows) return;
    for (int64_t k = indptr[row] + lane; k < indptr[row + 1]; k += 64)
        rowidx[k] = (int32_t) row;
}

__global__ void iota_kernel(int32_t* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t) i;
}

__global__ void col_count_kernel(const int32_t* __restrict__ sorted_cols, int64_t nnz,
                                 int64_t* __restrict__ colptr)
{
    // colptr[c + 1] = number of entries with column c, from the sorted keys:
    // an entry that starts a run of column c records the run end.
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t c = sorted_cols[k];
    if (k == nnz - 1 || sorted_cols[k + 1] != c) colptr[c + 1] = k + 1;
}

__global__ void fill_colptr_kernel(int64_t* __restrict__ colptr, int64_t ncols)
{
    // run ends -> monotone pointers: empty columns inherit the previous end.
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t last = 0;
    colptr[0] = 0;
    for (int64_t c = 1; c <= ncols; ++c)
    {
        if (colptr[c] < last) colptr[c] = last;
        last = colptr[c];
    }
}

template <typename T>
__global__ void gather_trans_kernel(const int32_t* __restrict__ perm,
                                    const int32_t* __restrict__ rowidx,
                                    const T* __restrict__ data, int64_t nnz,
                                    int32_t* __restrict__ tind, T* __restrict__ tdata)
{
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int32_t src = perm[k];
    tind[k] = rowidx[src];
    tdata[k] = data[src];
}

__global__ void set_neg_kernel(int64_t* p, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = -1;
}

__global__ void fill_u32_kernel(unsigned int* p, int64_t n)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Chunked transpose (2^31 non-zeros and more): column counts, then per range
// of columns the entries of each row inside the range (a contiguous run of
// the sorted row), stably sorted by column.
__global__ void count_cols_kernel(const int32_t* __restrict__ cols, int64_t nnz,
                                  unsigned int* __restrict__ counts)
{
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t) gridDim.x * blockDim.x)
        atomicAdd(&counts[cols[k]], 1u);
}

__device__ __forceinline__ int64_t lower_bound_dev(const int32_t* a, int64_t lo, int64_t hi,
                                                   int32_t v)
{
    while (lo < hi)
    {
        const int64_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void row_range_kernel(const int64_t* __restrict__ indptr,
                                 const int32_t* __restrict__ cols, int64_t nrows,
                                 int32_t c0, int32_t c1, int64_t* __restrict__ lo,
                                 int64_t* __restrict__ cnt)
{
    const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int64_t a = lower_bound_dev(cols, indptr[r], indptr[r + 1], c0);
    const int64_t b = lower_bound_dev(cols, a, indptr[r + 1], c1);
    lo[r] = a;
    cnt[r] = b - a;
}

__global__ void chunk_scatter_kernel(const int32_t* __restrict__ cols,
                                     const int64_t* __restrict__ lo,
                                     const int64_t* __restrict__ cnt,
                                     const int64_t* __restrict__ off, int64_t nrows, int32_t c0,
                                     int32_t* __restrict__ keys, int64_t* __restrict__ pos)
{
    const int64_t r = (int64_t) blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (r >= nrows) return;
    const int64_t a = lo[r], o = off[r];
    for (int64_t i = lane; i < cnt[r]; i += 64)
    {
        keys[o + i] = cols[a + i] - c0;
        pos[o + i] = a + i;
    }
}

template <typename T>
__global__ void chunk_gather_kernel(const int64_t* __restrict__ pos, int64_t m,
                                    const int64_t* __restrict__ indptr, int64_t nrows,
                                    const T* __restrict__ data, int32_t* __restrict__ tind,
                                    T* __restrict__ tdata)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t) gridDim.x * blockDim.x)
    {
        const int64_t p = pos[i];
        // row of entry p: the last row whose pointer is <= p
        int64_t lo = 0, hi = nrows;
        while (hi - lo > 1)
        {
            const int64_t mid = lo + (hi - lo) / 2;
            if (indptr[mid] <= p) lo = mid;
            else hi = mid;
        }
        tind[i] = (int32_t) lo;
        tdata[i] = data[p];
    }
}

}  // namespace

// CSR of A^T by a stable radix sort of the column keys: within a column the
// rows stay ascending, so the transpose is deterministic.
// Transpose of a matrix with 2^31 non-zeros or more, in ranges of columns
// holding at most 2^30 entries each (the radix sort counts in int).
static int build_transpose_chunked(hipStream_t s, const DevCSR& A, DevCSR& T)
{
    const int64_t nnz = A.nnz;
    const size_t vb = A.dtype == SLQ_F64 ? 8 : 4;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.nnz = nnz;
    T.dtype = A.dtype;
    T.sorted = true;   // each chunk is sorted stably by column: rows ascend
    SLQ_CHECK(hipMalloc(&T.indptr, sizeof(int64_t) * (T.nrows + 1)));
    SLQ_CHECK(hipMalloc(&T.indices, sizeof(int32_t) * nnz));
    SLQ_CHECK(hipMalloc(&T.data, vb * nnz));
    // column counts -> pointers (host scan)
    unsigned int* dcnt = nullptr;
    SLQ_CHECK(hipMalloc(&dcnt, sizeof(unsigned int) * A.ncols));
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned) ((A.ncols + 255) / 256)), dim3(256),
                       0, s, dcnt, A.ncols);
    hipLaunchKernelGGL(count_cols_kernel, dim3(4096), dim3(256), 0, s, A.indices, nnz, dcnt);
    std::vector<unsigned int> cnt(A.ncols);
    SLQ_CHECK(hipMemcpyAsync(cnt.data(), dcnt, sizeof(unsigned int) * A.ncols,
                             hipMemcpyDeviceToHost, s));
    SLQ_CHECK(hipStreamSynchronize(s));
    (void) hipFree(dcnt);
    std::vector<int64_t> ptr(A.ncols + 1, 0);
    for (int64_t c = 0; c < A.ncols; ++c) ptr[c + 1] = ptr[c] + cnt[c];
    SLQ_CHECK(hipMemcpy(T.indptr, ptr.data(), sizeof(int64_t) * (A.ncols + 1),
                        hipMemcpyHostToDevice));
    // per-row scratch and the largest chunk's sort buffers; SLQ_TRANSPOSE_CHUNK
    // sets a smaller chunk (entries) to exercise the chunked path
    int64_t kChunk = (int64_t) 1 << 30;
    if (const char* env = std::getenv("SLQ_TRANSPOSE_CHUNK"))
        kChunk = std::max<int64_t>(1, std::min<int64_t>(kChunk, std::atoll(env)));
    int64_t *lo = nullptr, *rc = nullptr, *off = nullptr;
    SLQ_CHECK(hipMalloc(&lo, sizeof(int64_t) * A.nrows));
    SLQ_CHECK(hipMalloc(&rc, sizeof(int64_t) * A.nrows));
    SLQ_CHECK(hipMalloc(&off, sizeof(int64_t) * A.nrows));
    int32_t *keys = nullptr, *keys2 = nullptr;
    int64_t *pos = nullptr, *pos2 = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0, scan_bytes = 0;
    int r = SLQ_OK;
    int64_t c0 = 0;
    while (c0 < A.ncols && r == SLQ_OK)
    {
        int64_t c1 = c0 + 1;
        while (c1 < A.ncols && ptr[c1 + 1] - ptr[c0] <= kChunk) ++c1;
        const int64_t m = ptr[c1] - ptr[c0];
        if (m > kChunk)
        {
            set_error("transpose: one column holds more than 2^30 entries");
            r = SLQ_E_ARG;
            break;
        }
        if (m > 0)
        {
            if (!keys)
            {
                // sized for the largest chunk, found by a dry pass of the
                // same greedy split
                int64_t big = 0, a = 0;
                while (a < A.ncols)
                {
                    int64_t b = a + 1;
                    while (b < A.ncols && ptr[b + 1] - ptr[a] <= kChunk) ++b;
                    big = std::max<int64_t>(big, ptr[b] - ptr[a]);
                    a = b;
                }
                hipError_t e;
                if ((e = hipMalloc(&keys, sizeof(int32_t) * big)) != hipSuccess ||
                    (e = hipMalloc(&keys2, sizeof(int32_t) * big)) != hipSuccess ||
                    (e = hipMalloc(&pos, sizeof(int64_t) * big)) != hipSuccess ||
                    (e = hipMalloc(&pos2, sizeof(int64_t) * big)) != hipSuccess)
                {
                    r = hip_status(e, "transpose buffers", __FILE__, __LINE__);
                    break;
                }
                size_t sb = 0;
                (void) hipcub::DeviceRadixSort::SortPairs(nullptr, sb, keys, keys2, pos, pos2,
                                                          (int) big, 0, 32, s);
                (void) hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, rc, off,
                                                        (int) A.nrows, s);
                tmp_bytes = std::max(sb, scan_bytes);
                if ((e = hipMalloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1)) != hipSuccess)
                {
                    r = hip_status(e, "transpose buffers", __FILE__, __LINE__);
                    break;
                }
            }
            const unsigned gr = (unsigned) ((A.nrows + 255) / 256);
            hipLaunchKernelGGL(row_range_kernel, dim3(gr), dim3(256), 0, s, A.indptr,
                               A.indices, A.nrows, (int32_t) c0, (int32_t) c1, lo, rc);
            size_t sb = tmp_bytes;
            hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, sb, rc, off, (int) A.nrows, s);
            if (e != hipSuccess)
            {
                r = hip_status(e, "transpose scan", __FILE__, __LINE__);
                break;
            }
            hipLaunchKernelGGL(chunk_scatter_kernel, dim3((unsigned) ((A.nrows + 3) / 4)),
                               dim3(256), 0, s, A.indices, lo, rc, off, A.nrows,
                               (int32_t) c0, keys, pos);
            int end_bit = 1;
            while (end_bit < 31 && ((int64_t) 1 << end_bit) < c1 - c0) ++end_bit;
            sb = tmp_bytes;
            e = hipcub::DeviceRadixSort::SortPairs(tmp, sb, keys, keys2, pos, pos2, (int) m, 0,
                                                   end_bit, s);
            if (e != hipSuccess)
            {
                r = hip_status(e, "transpose sort", __FILE__, __LINE__);
                break;
            }
            if (A.dtype == SLQ_F64)
                hipLaunchKernelGGL(chunk_gather_kernel<double>, dim3(8192), dim3(256), 0, s,
                                   pos2, m, A.indptr, A.nrows, (const double*) A.data,
                                   T.indices + ptr[c0], (double*) T.data + ptr[c0]);
            else
                hipLaunchKernelGGL(chunk_gather_kernel<float>, dim3(8192), dim3(256), 0, s,
                                   pos2, m, A.indptr, A.nrows, (const float*) A.data,
                                   T.indices + ptr[c0], (float*) T.data + ptr[c0]);
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) r = hip_status(e, "transpose", __FILE__, __LINE__);
        }
        c0 = c1;
    }
    (void) hipFree(lo);
    (void) hipFree(rc);
    (void) hipFree(off);
    (void) hipFree(keys);
    (void) hipFree(keys2);
    (void) hipFree(pos);
    (void) hipFree(pos2);
    (void) hipFree(tmp);
    return r;
}

int build_transpose(hipStream_t s, const DevCSR& A, DevCSR& T)
{
    const int64_t nnz = A.nnz;
    // The chunked transpose finds each row's run of a column range by binary
    // search, so it needs the column indices sorted within every row. An
    // unsorted CSR takes the one-sort path where the sort can count it, and
    // is refused above that.
    const bool chunked = nnz >= (int64_t) 1 << 31 || std::getenv("SLQ_TRANSPOSE_CHUNK");
    if (chunked && A.sorted) return build_transpose_chunked(s, A, T);
    if (nnz >= (int64_t) 1 << 31)
    {
        set_error("transpose: a matrix with 2^31 or more non-zeros needs its column "
                  "indices sorted within every row");
        return SLQ_E_ARG;
    }
    const size_t vb = A.dtype == SLQ_F64 ? 8 : 4;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.nnz = nnz;
    T.dtype = A.dtype;
    T.sorted = true;   // stable sort by column: the rows of A ascend
    SLQ_CHECK(hipMalloc(&T.indptr, sizeof(int64_t) * (T.nrows + 1)));
    SLQ_CHECK(hipMalloc(&T.indices, sizeof(int32_t) * (nnz > 0 ? nnz : 1)));
    SLQ_CHECK(hipMalloc(&T.data, vb * (nnz > 0 ? nnz : 1)));
    const int64_t n1 = T.nrows + 1;
    hipLaunchKernelGGL(set_neg_kernel, dim3((unsigned) ((n1 + 255) / 256)), dim3(256),
                       0, s, T.indptr, n1);
    if (nnz > 0)
    {
        int32_t *rowidx = nullptr, *iota = nullptr, *keys = nullptr, *perm = nullptr;
        void* tmp = nullptr;
        size_t tmp_bytes = 0;
        SLQ_CHECK(hipMalloc(&rowidx, sizeof(int32_t) * nnz));
        SLQ_CHECK(hipMalloc(&iota, sizeof(int32_t) * nnz));
        SLQ_CHECK(hipMalloc(&keys, sizeof(int32_t) * nnz));
        SLQ_CHECK(hipMalloc(&perm, sizeof(int32_t) * nnz));
        const unsigned g = (unsigned) ((nnz + 255) / 256);
        hipLaunchKernelGGL(expand_rows_kernel, dim3((unsigned) ((A.nrows + 3) / 4)),
                           dim3(256), 0, s, A.indptr, A.nrows, rowidx);
        hipLaunchKernelGGL(iota_kernel, dim3(g), dim3(256), 0, s, iota, nnz);
        int end_bit = 1;
        while (end_bit < 31 && ((int64_t) 1 << end_bit) < A.ncols) ++end_bit;
        SLQ_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, A.indices, keys,
                                                     iota, perm, (int) nnz, 0, end_bit, s));
        SLQ_CHECK(hipMalloc(&tmp, tmp_bytes > 0 ? tmp_bytes : 1));
        SLQ_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, A.indices, keys,
                                                     iota, perm, (int) nnz, 0, end_bit, s));
        hipLaunchKernelGGL(col_count_kernel, dim3(g), dim3(256), 0, s, keys, nnz, T.indptr);
        if (A.dtype == SLQ_F64)
            hipLaunchKernelGGL(gather_trans_kernel<double>, dim3(g), dim3(256), 0, s, perm,
                               rowidx, (const double*) A.data, nnz, T.indices,
                               (double*) T.data);
        else
            hipLaunchKernelGGL(gather_trans_kernel<float>, dim3(g), dim3(256), 0, s, perm,
                               rowidx, (const float*) A.data, nnz, T.indices,
                               (float*) T.data);
        SLQ_CHECK(hipGetLastError());
        SLQ_CHECK(hipStreamSynchronize(s));
        (void) hipFree(rowidx);
        (void) hipFree(iota);
        (void) hipFree(keys);
        (void) hipFree(perm);
        (void) hipFree(tmp);
    }
    hipLaunchKernelGGL(fill_colptr_kernel, dim3(1), dim3(1), 0, s, T.indptr, T.nrows);
    SLQ_CHECK(hipGetLastError());
    SLQ_CHECK(hipStreamSynchronize(s));
    return SLQ_OK;
}

// ---------------------------------------------------------------- statistics

namespace {
double sign_d(double x) { return (x > 0) - (x < 0); }
}  // namespace

// erf_inv, special_functions.cpp:63-104: rational initial guess with the
// published coefficients the reference cites, then two Newton steps.
double erf_inv(double x)
{
    if (x == 1.0 || x == -1.0) return sign_d(x) * INFINITY;
    const double a[4] = {0.886226899, -1.645349621, 0.914624893, -0.140543331};
    const double b[5] = {1.0, -2.118377725, 1.442710462, -0.329097515, 0.012229801};
    const double c[4] = {-1.970840454, -1.62490649, 3.429567803, 1.641345311};
    const double d[3] = {1.0, 3.543889200, 1.637067800};
    double z = sign_d(x) * x;
    double r;
    if (z <= 0.7)
    {
        const double x2 = z * z;
        r = z * (((a[3] * x2 + a[2]) * x2 + a[1]) * x2 + a[0]);
        r /= (((b[4] * x2 + b[3]) * x2 + b[2]) * x2 + b[1]) * x2 + b[0];
    }
    else
    {
        const double y = std::sqrt(-std::log((1.0 - z) / 2.0));
        r = (((c[3] * y + c[2]) * y + c[1]) * y + c[0]);
        r /= ((d[2] * y + d[1]) * y + d[0]);
    }
    r = r * sign_d(x);
    z = z * sign_d(x);
    r -= (std::erf(r) - z) / ((2.0 / std::sqrt(M_PI)) * std::exp(-r * r));
    r -= (std::erf(r) - z) / ((2.0 / std::sqrt(M_PI)) * std::exp(-r * r));
    return r;
}

// ConvergenceTools::check_convergence, convergence_tools.cpp:94-194.
int check_convergence(const double* samples, int q, int min_num_samples,
                      const int32_t* processed, int num_processed,
                      double confidence, double atol, double rtol,
                      double* error, int32_t* num_used, int32_t* converged)
{
    if (num_processed < min_num_samples)
    {
        for (int j = 0; j < q; ++j)
        {
            error[j] = INFINITY;
            converged[j] = 0;
            num_used[j] = num_processed;
        }
        return 0;
    }
    const double z = std::sqrt(2) * erf_inv(confidence);
    for (int j = 0; j < q; ++j)
    {
        if (converged[j] != 0) continue;
        double s = 0.0;
        for (int i = 0; i < num_processed; ++i) s += samples[(size_t) processed[i] * q + j];
        const double mean = s / num_processed;
        double sd;
        if (num_processed > 1)
        {
            s = 0.0;
            for (int i = 0; i < num_processed; ++i)
            {
                const double d = samples[(size_t) processed[i] * q + j];
                s += (d - mean) * (d - mean);
            }
            sd = std::sqrt(s / (num_processed - 1.0));
        }
        else
            sd = INFINITY;
        error[j] = z * sd / std::sqrt((double) num_processed);
        if (error[j] < std::max(atol, rtol * mean)) converged[j] = 1;
        num_used[j] = num_processed;
    }
    for (int j = 0; j < q; ++j)
        if (converged[j] == 0) return 0;
    return 1;
}

// ConvergenceTools::average_estimates, convergence_tools.cpp:255-376.
void average_estimates(const double* samples, int q, double confidence,
                       double outlier, const int32_t* num_used,
                       const int32_t* processed, int32_t* num_outliers,
                       double* trace, double* error)
{
    const double ez = std::sqrt(2) * erf_inv(confidence);
    const double oz = std::sqrt(2.0) * erf_inv(1.0 - outlier);
    for (int j = 0; j < q; ++j)
    {
        const int N = num_used[j];
        num_outliers[j] = 0;
        double s = 0.0;
        for (int i = 0; i < N; ++i) s += samples[(size_t) processed[i] * q + j];
        double mean = s / N;
        double sd;
        if (N > 1)
        {
            s = 0.0;
            for (int i = 0; i < N; ++i)
            {
                const double d = samples[(size_t) processed[i] * q + j] - mean;
                s += d * d;
            }
            sd = std::sqrt(s / (N - 1.0));
        }
        else
            sd = INFINITY;
        // outlier test against the first pass's mean and deviation; it is
        // re-evaluated in the passes below instead of being stored
        const double half = oz * sd;
        const double mean0 = mean;
        auto outlier_at = [&](int i) {
            return std::abs(samples[(size_t) processed[i] * q + j] - mean0) > half;
        };
        for (int i = 0; i < N; ++i)
            if (outlier_at(i)) num_outliers[j] += 1;
        const int no = num_outliers[j];
        s = 0.0;
        for (int i = 0; i < N; ++i)
            if (!outlier_at(i)) s += samples[(size_t) processed[i] * q + j];
        mean = s / (N - no);
        if (N > 1 + no)
        {
            s = 0.0;
            for (int i = 0; i < N; ++i)
                if (!outlier_at(i))
                {
                    const double d = samples[(size_t) processed[i] * q + j] - mean;
                    s += d * d;
                }
            sd = std::sqrt(s / (N - no - 1.0));
        }
        else
            sd = INFINITY;
        trace[j] = mean;
        error[j] = ez * sd / std::sqrt((double) (N - no));
    }
}

int64_t clock_seed()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (int64_t) (((uint64_t) ts.tv_sec * 1000000000ULL + (uint64_t) ts.tv_nsec) &
                      0x7fffffffffffffffULL);
}

namespace {

// Workspace carve-up for one estimate / recurrence call.
struct Plan {
    int64_t len = 0;        // probe length
    int64_t W = 0;          // words per probe
    int vbytes = 8;         // value bytes of the run
    int m = 0;
    int q = 1;
    size_t off_words = 0, off_X0 = 0, off_alpha = 0, off_beta = 0, off_sizes = 0,
           off_samples = 0, off_params = 0, off_scratch = 0, total = 0;
    // a second set of per-block outputs, so that the next block's recurrence
    // can run while this block's quadrature and copy-back finish
    size_t off_alpha2 = 0, off_beta2 = 0, off_sizes2 = 0, off_samples2 = 0;
    size_t alpha_of(int par) const { return par ? off_alpha2 : off_alpha; }
    size_t beta_of(int par) const { return par ? off_beta2 : off_beta; }
    size_t sizes_of(int par) const { return par ? off_sizes2 : off_sizes; }
    size_t samples_of(int par) const { return par ? off_samples2 : off_samples; }
    size_t scratch_bytes = 0;
    char* base = nullptr;   // device workspace the offsets are taken from
    // false for slq_estimate_noalloc: no streams or events are created, so
    // blocks run one after the other on the handle's stream
    bool pipeline = true;
    // recurrences of several t of A + tB side by side: per lane alpha, beta,
    // sizes, recurrence scratch and the pair-tile values of its t (both
    // directions for Golub-Kahn), lane c at off_lanes + c * lane_stride
    int lanes = 0;
    size_t off_lanes = 0, lane_stride = 0;
    // block Krylov form: slots of a group and the group's samples [slots][64][q]
    int bk_slots = 0;
    size_t off_bk_samples = 0;
    size_t lane_alpha = 0, lane_beta = 0, lane_sizes = 0, lane_vals = 0, lane_vals_t = 0,
           lane_scratch = 0;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

Plan make_plan(slq_op* op, int gram, int m, int ortho, int q, int vbytes)
{
    Plan p;
    p.len = gram ? op->A->csr.ncols : op->A->csr.nrows;
    p.W = p.len / 64 + 1;
    p.vbytes = vbytes;
    p.m = m;
    p.q = q;
    size_t o = 0;
    p.off_words = o;   o += align256(sizeof(uint64_t) * kBlock * p.W);
    p.off_X0 = o;      o += align256((size_t) vbytes * kBlock * p.len);
    p.off_alpha = o;   o += align256(sizeof(double) * kBlock * m);
    p.off_beta = o;    o += align256(sizeof(double) * kBlock * m);
    p.off_sizes = o;   o += align256(sizeof(int32_t) * kBlock);
    p.off_samples = o; o += align256(sizeof(double) * kBlock * q);
    p.off_alpha2 = o;  o += align256(sizeof(double) * kBlock * m);
    p.off_beta2 = o;   o += align256(sizeof(double) * kBlock * m);
    p.off_sizes2 = o;  o += align256(sizeof(int32_t) * kBlock);
    p.off_samples2 = o; o += align256(sizeof(double) * kBlock * q);
    p.off_params = o;  o += align256(sizeof(double) * q);
    p.off_scratch = o;
    p.scratch_bytes = recurrence_scratch_bytes(op, gram, m, ortho, vbytes);
    o += p.scratch_bytes;
    p.total = o;
    return p;
}

// Lanes for the t of an A + tB estimate with q inquiries: 0 (one recurrence
// after the other) unless the union tiles exist in every direction the
// recurrence multiplies by, q > 1 and SLQ_CONCURRENT_T is not 0; at most 8,
// and no more than fit 8 GB of lane workspace.
int t_lanes(const slq_op* op, int gram, int q, const Plan& p)
{
    const char* env = std::getenv("SLQ_CONCURRENT_T");   // read per estimate
    const bool on = !(env && std::atoi(env) == 0);
    if (!on || op->kind != OP_AFFINE_B || q < 2 || op->ab.tiles.R == 0 ||
        (gram && op->ab_t.tiles.R == 0))
        return 0;
    const size_t per = 2 * align256(sizeof(double) * kBlock * p.m) +
                       align256(sizeof(int32_t) * kBlock) +
                       align256((size_t) p.vbytes * op->ab.tiles.nslabs * op->ab.tiles.R) +
                       (gram ? align256((size_t) p.vbytes * op->ab_t.tiles.nslabs *
                                        op->ab_t.tiles.R) : 0) +
                       p.scratch_bytes;
    int lanes = q < 8 ? q : 8;
    while (lanes > 1 && per * (size_t) lanes > ((size_t) 8 << 30)) --lanes;
    return lanes > 1 ? lanes : 0;
}

// Adds the lane regions to a plan.
void plan_lanes(const slq_op* op, int gram, int lanes, Plan& p)
{
    p.lanes = lanes;
    if (lanes <= 0) return;
    size_t o = 0;
    p.lane_alpha = o;   o += align256(sizeof(double) * kBlock * p.m);
    p.lane_beta = o;    o += align256(sizeof(double) * kBlock * p.m);
    p.lane_sizes = o;   o += align256(sizeof(int32_t) * kBlock);
    p.lane_vals = o;
    o += align256((size_t) p.vbytes * op->ab.tiles.nslabs * op->ab.tiles.R);
    p.lane_vals_t = o;
    if (gram) o += align256((size_t) p.vbytes * op->ab_t.tiles.nslabs * op->ab_t.tiles.R);
    p.lane_scratch = o; o += p.scratch_bytes;
    p.lane_stride = o;
    p.off_lanes = p.total;
    p.total += o * (size_t) lanes;
}

template <typename U>
U* at(const Plan& p, size_t off) { return (U*) (p.base + off); }

int value_dtype(const slq_op* op, int precision)
{
    return (precision == SLQ_PREC_MIXED || op->A->csr.dtype == SLQ_F32) ? SLQ_F32 : SLQ_F64;
}

int prepare(slq_op* op, int gram, int precision)
{
    if (op->A->dense)
    {
        if (gram) SLQ_TRY(dense_transpose(op->stream, *op->A));
        if (precision == SLQ_PREC_MIXED) SLQ_TRY(ensure_f32(op->stream, *op->A, gram != 0));
        return SLQ_OK;
    }
    if (gram)
    {
        if (!op->A->has_trans)
        {
            SLQ_TRY(build_transpose(op->stream, op->A->csr, op->A->trans));
            op->A->has_trans = true;
        }
        if (!op->A->trans_tiles_tried && op->A->tiles.R > 0)
        {
            // tiled form of A^T, built on the host from the device transpose
            op->A->trans_tiles_tried = true;
            const DevCSR& T = op->A->trans;
            const size_t vb = T.dtype == SLQ_F64 ? 8 : 4;
            std::vector<int64_t> ip(T.nrows + 1);
            std::vector<int32_t> ix(T.nnz > 0 ? T.nnz : 1);
            std::vector<char> dv(vb * (T.nnz > 0 ? T.nnz : 1));
            SLQ_CHECK(hipMemcpy(ip.data(), T.indptr, sizeof(int64_t) * (T.nrows + 1),
                                hipMemcpyDeviceToHost));
            if (T.nnz > 0)
            {
                SLQ_CHECK(hipMemcpy(ix.data(), T.indices, sizeof(int32_t) * T.nnz,
                                    hipMemcpyDeviceToHost));
                SLQ_CHECK(hipMemcpy(dv.data(), T.data, vb * T.nnz, hipMemcpyDeviceToHost));
            }
            SLQ_TRY(build_tiles(T.nrows, T.ncols, ip.data(), ix.data(), dv.data(), T.dtype,
                                op->A->tiles.R, 0.25, op->A->trans_tiles));
        }
        if (op->kind == OP_AFFINE_B && !op->B->has_trans)
        {
            SLQ_TRY(build_transpose(op->stream, op->B->csr, op->B->trans));
            op->B->has_trans = true;
        }
    }
    if (op->kind == OP_AFFINE_B)
    {
        SLQ_TRY(ensure_pair_tiles(op, precision == SLQ_PREC_MIXED, false));
        if (gram) SLQ_TRY(ensure_pair_tiles(op, precision == SLQ_PREC_MIXED, true));
    }
    if (precision == SLQ_PREC_MIXED)
    {
        SLQ_TRY(ensure_f32(op->stream, *op->A, gram != 0));
        if (op->kind == OP_AFFINE_B) SLQ_TRY(ensure_f32(op->stream, *op->B, gram != 0));
    }
    return SLQ_OK;
}

// Probes of samples [first, first + count) into the block X0.
int make_probes(slq_op* op, const Plan& p, const uint64_t base[4], int64_t first,
                int count, int vdtype)
{
    SLQ_TRY(rng_words(op->stream, base, p.W, first, count, at<uint64_t>(p, p.off_words)));
    SLQ_TRY(rng_expand(op->stream, at<uint64_t>(p, p.off_words), p.W, p.len, count,
                       vdtype, at<void>(p, p.off_X0)));
    return SLQ_OK;
}

int recurrence_block(slq_op* op, const Plan& p, double t, int gram, double tol,
                     int ortho, int precision, int par = 0)
{
    RecurrenceArgs ra;
    ra.op = op;
    ra.t = t;
    ra.gram = gram;
    ra.m = p.m;
    ra.tol = tol;
    ra.ortho = ortho;
    ra.precision = precision;
    ra.X0 = at<void>(p, p.off_X0);
    ra.alpha = at<double>(p, p.alpha_of(par));
    ra.beta = at<double>(p, p.beta_of(par));
    ra.sizes = at<int32_t>(p, p.sizes_of(par));
    return run_recurrence(ra, at<void>(p, p.off_scratch), p.scratch_bytes);
}

// A view of an A + tB handle for one lane: the same matrices and launch
// settings, the lane's stream, and the handle's pair tiles with the lane's
// own value buffers (its t's values are gathered there), so the recurrences
// of several t run side by side without sharing writable state.
struct LaneOp {
    slq_op op;
    LaneOp(const slq_op* parent, hipStream_t s, void* vals, void* vals_t)
    {
        op.kind = parent->kind;
        op.A = parent->A;
        op.B = parent->B;
        op.stream = s;
        op.device = parent->device;
        op.spmm_grid = parent->spmm_grid;
        op.stream_grid = parent->stream_grid;
        op.spmm_split = parent->spmm_split;
        op.tile_mode = parent->tile_mode;
        borrow(op.ab, parent->ab, vals);
        borrow(op.ab_t, parent->ab_t, vals_t);
    }
    static void borrow(PairTiles& dst, const PairTiles& src, void* vals)
    {
        dst.tiles = src.tiles;
        dst.tiles.vals = vals;
        dst.tiles.vals_f32 = (float*) vals;
        dst.src_a = src.src_a;
        dst.src_b = src.src_b;
        dst.tried = src.tried;
        dst.borrowed = true;
    }
};

// The recurrence of lane c of a plan (on the lane view's stream).
int recurrence_lane(slq_op* lop, const Plan& p, int c, double t, int gram, double tol,
                    int ortho, int precision)
{
    const size_t o = p.off_lanes + (size_t) c * p.lane_stride;
    RecurrenceArgs ra;
    ra.op = lop;
    ra.t = t;
    ra.gram = gram;
    ra.m = p.m;
    ra.tol = tol;
    ra.ortho = ortho;
    ra.precision = precision;
    ra.X0 = at<void>(p, p.off_X0);
    ra.alpha = at<double>(p, o + p.lane_alpha);
    ra.beta = at<double>(p, o + p.lane_beta);
    ra.sizes = at<int32_t>(p, o + p.lane_sizes);
    return run_recurrence(ra, at<void>(p, o + p.lane_scratch), p.scratch_bytes);
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void) hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void) hipSetDevice(prev);
    }
};

}  // namespace

// ---------------------------------------------------------------- copy kernel

namespace {
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ in,
                                                   uint4* __restrict__ out, int64_t n)
{
    // four 16 B loads in flight per thread, then four stores
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride)
    {
        const uint4 a = in[i], b = in[i + stride], c = in[i + 2 * stride],
                    d = in[i + 3 * stride];
        out[i] = a;
        out[i + stride] = b;
        out[i + 2 * stride] = c;
        out[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) out[i] = in[i];
}
}  // namespace

int copy_bandwidth(int64_t bytes, int iters, double* ms)
{
    void *a = nullptr, *b = nullptr;
    const int64_t n = bytes / 16;
    SLQ_CHECK(hipMalloc(&a, n * 16));
    SLQ_CHECK(hipMalloc(&b, n * 16));
    SLQ_CHECK(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    SLQ_CHECK(hipEventCreate(&e0));
    SLQ_CHECK(hipEventCreate(&e1));
    const int grid = 256 * 16;
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, (const uint4*) a, (uint4*) b, n);
    SLQ_CHECK(hipEventRecord(e0, 0));
    for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, (const uint4*) a,
                           (uint4*) b, n);
    SLQ_CHECK(hipEventRecord(e1, 0));
    SLQ_CHECK(hipEventSynchronize(e1));
    float t = 0.f;
    SLQ_CHECK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / (iters > 0 ? iters : 1);
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    (void) hipFree(a);
    (void) hipFree(b);
    return SLQ_OK;
}

}  // namespace slq

using namespace slq;

// ---------------------------------------------------------------- autotuner

namespace {

// Runs one probe block of `steps` Lanczos steps with the op's current grid
// caps and returns the device time (ms), or a negative value on error.
double time_block(slq_op* op, int steps, int precision)
{
    const int vdt = value_dtype(op, precision);
    Plan p = make_plan(op, 0, steps, 0, 1, vdt == SLQ_F64 ? 8 : 4);
    if (op->ws.reserve(p.total) != SLQ_OK) return -1.0;
    p.base = (char*) op->ws.base;
    uint64_t base[4];
    thread_state(1, 0, base);
    if (make_probes(op, p, base, 0, kBlock, vdt) != SLQ_OK) return -1.0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1.0;
    if (hipEventCreate(&e1) != hipSuccess)
    {
        (void) hipEventDestroy(e0);
        return -1.0;
    }
    double best = -1.0;
    for (int rep = 0; rep < 2; ++rep)   // first repetition warms up
    {
        (void) hipEventRecord(e0, op->stream);
        if (recurrence_block(op, p, 0.0, 0, 1e-300, 0, precision) != SLQ_OK) break;
        (void) hipEventRecord(e1, op->stream);
        if (hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.f;
        (void) hipEventElapsedTime(&ms, e0, e1);
        best = ms;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    return best;
}

struct TuneKey {
    int log2_rows;
    int nnz_per_row;
    int vdtype;
    int tile_rows;
    bool operator<(const TuneKey& o) const
    {
        if (log2_rows != o.log2_rows) return log2_rows < o.log2_rows;
        if (nnz_per_row != o.nnz_per_row) return nnz_per_row < o.nnz_per_row;
        if (vdtype != o.vdtype) return vdtype < o.vdtype;
        return tile_rows < o.tile_rows;
    }
};

struct TuneValue {
    int spmm_grid, stream_grid, spmm_split, tile_mode;
};

std::mutex g_tune_mu;
std::map<TuneKey, TuneValue> g_tune_cache;

}  // namespace

extern "C" int slq_autotune(slq_op* op, int32_t precision, int32_t* spmm_grid,
                            int32_t* stream_grid)
{
    if (!op || op->kind == OP_AFFINE_B ||
        op->A->csr.nrows != op->A->csr.ncols)
    {
        set_error("slq_autotune: a square matrix or A + tI handle is required");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    SLQ_TRY(prepare(op, 0, precision));
    const DevCSR& c = op->A->csr;
    TuneKey key;
    key.log2_rows = (int) std::lround(std::log2((double) c.nrows));
    key.nnz_per_row = (int) std::lround((double) c.nnz / (double) c.nrows);
    key.vdtype = value_dtype(op, precision);
    key.tile_rows = op->A->tiles.R;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end())
        {
            op->tuned_spmm_grid = it->second.spmm_grid;
            op->tuned_stream_grid = it->second.stream_grid;
            op->tuned_spmm_split = it->second.spmm_split;
            op->tuned_tile_mode = it->second.tile_mode;
            if (spmm_grid) *spmm_grid = op->tuned_spmm_grid;
            if (stream_grid) *stream_grid = op->tuned_stream_grid;
            return SLQ_OK;
        }
    }
    // candidate caps (multiples of 8 up to the natural grid of the problem),
    // timed on the handle's own grid fields, which are restored afterwards
    const int cands[] = {8, 16, 32, 64, 128, 256, 512, 1024};
    const int steps = 4;
    const int keep_s = op->spmm_grid, keep_u = op->stream_grid;
    int best_s = 0, best_u = 0;
    double best_t = 1e300;
    int r = SLQ_OK;
    op->stream_grid = 0;
    for (int cap : cands)
    {
        if (cap > grid_for(c.nrows, kMaxGrid)) break;
        op->spmm_grid = cap;
        const double t = time_block(op, steps, precision);
        if (t < 0)
        {
            r = SLQ_E_HIP;
            break;
        }
        if (t < best_t)
        {
            best_t = t;
            best_s = cap;
        }
    }
    op->spmm_grid = best_s;
    best_t = 1e300;
    for (int cap : cands)
    {
        if (r != SLQ_OK || cap > grid_for(c.nrows, kMaxGrid)) break;
        op->stream_grid = cap;
        const double t = time_block(op, steps, precision);
        if (t < 0)
        {
            r = SLQ_E_HIP;
            break;
        }
        if (t < best_t)
        {
            best_t = t;
            best_u = cap;
        }
    }
    // operand form at the chosen caps: the 16-row tiles with 1, 2 or 4
    // wavefronts per tile, or the CSR rows
    int best_k = 0, best_m = 0;
    const int keep_k = op->spmm_split, keep_m = op->tile_mode;
    if (r == SLQ_OK && op->A->tiles.R > 0)
    {
        op->spmm_grid = best_s;
        op->stream_grid = best_u;
        best_t = 1e300;
        struct Form {
            int mode, split;
        };
        std::vector<Form> forms;
        if (op->A->tiles.R == 16)
            forms = {{0, 1}, {0, 2}, {0, 4}, {1, 1}};
        else
            forms = {{0, 1}, {1, 1}};
        for (const Form& f : forms)
        {
            op->tile_mode = f.mode;
            op->spmm_split = f.split;
            const double t = time_block(op, steps, precision);
            if (t < 0)
            {
                r = SLQ_E_HIP;
                break;
            }
            if (t < best_t)
            {
                best_t = t;
                best_m = f.mode + 1;
                best_k = f.split;
            }
        }
    }
    op->spmm_grid = keep_s;
    op->stream_grid = keep_u;
    op->spmm_split = keep_k;
    op->tile_mode = keep_m;
    if (r != SLQ_OK) return r;
    op->tuned_spmm_grid = best_s;
    op->tuned_stream_grid = best_u;
    op->tuned_spmm_split = best_k;
    op->tuned_tile_mode = best_m;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune_cache[key] = {best_s, best_u, best_k, best_m};
    }
    if (spmm_grid) *spmm_grid = best_s;
    if (stream_grid) *stream_grid = best_u;
    return SLQ_OK;
}

// ---------------------------------------------------------------- batch

// Swaps the autotuned launch settings into the handles of a batch and back.
struct TunedLaunch {
    struct Saved {
        int spmm_grid, stream_grid, spmm_split, tile_mode;
    };
    slq_op** ops;
    int count;
    std::vector<Saved> saved;
    TunedLaunch(slq_op** o, int c) : ops(o), count(c), saved(c)
    {
        for (int i = 0; i < count; ++i)
        {
            slq_op* op = ops[i];
            saved[i] = {op->spmm_grid, op->stream_grid, op->spmm_split, op->tile_mode};
            if (op->tuned_spmm_grid > 0) op->spmm_grid = op->tuned_spmm_grid;
            if (op->tuned_stream_grid > 0) op->stream_grid = op->tuned_stream_grid;
            if (op->tuned_spmm_split > 0) op->spmm_split = op->tuned_spmm_split;
            if (op->tuned_tile_mode > 0) op->tile_mode = op->tuned_tile_mode - 1;
        }
    }
    ~TunedLaunch()
    {
        for (int i = 0; i < count; ++i)
        {
            ops[i]->spmm_grid = saved[i].spmm_grid;
            ops[i]->stream_grid = saved[i].stream_grid;
            ops[i]->spmm_split = saved[i].spmm_split;
            ops[i]->tile_mode = saved[i].tile_mode;
        }
    }
};

// A valid matrix function: a known code, with coefficients where the code
// takes them (a positive sigma for the Gaussian).
bool func_ok(int func, const double* fp)
{
    if (func < SLQ_FUNC_LOG || func > SLQ_FUNC_GAUSSIAN) return false;
    if (func >= SLQ_FUNC_EXP && !fp) return false;
    if (func == SLQ_FUNC_GAUSSIAN && !(fp[1] > 0.0)) return false;
    return true;
}

extern "C" int slq_estimate_batch(slq_op** ops, int32_t count, int32_t func,
                                  const double* func_params, int32_t gram,
                                  double exponent, int32_t orthogonalize,
                                  int64_t seed, int32_t m, double tol,
                                  int32_t num_samples, double confidence,
                                  double outlier, int32_t precision, int32_t use_graph,
                                  double* trace, double* error, double* samples,
                                  double* run_time, double* setup_time)
{
    if (!ops || count < 1 || m < 1 || num_samples < 1 || !trace || !error ||
        !func_ok(func, func_params) ||
        (precision != SLQ_PREC_NATIVE && precision != SLQ_PREC_MIXED))
    {
        set_error("slq_estimate_batch: invalid argument");
        return SLQ_E_ARG;
    }
    for (int i = 0; i < count; ++i)
    {
        if (!ops[i] || ops[i]->kind != OP_MATRIX ||
            (!gram && ops[i]->A->csr.nrows != ops[i]->A->csr.ncols) ||
            ops[i]->device != ops[0]->device)
        {
            set_error("slq_estimate_batch: operator %d must be a (square unless gram) "
                      "matrix handle on the device of operator 0", i);
            return SLQ_E_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (ops[j] == ops[i])
            {
                set_error("slq_estimate_batch: operator %d repeated (each problem needs "
                          "its own handle)", i);
                return SLQ_E_ARG;
            }
    }
    DeviceGuard guard(ops[0]->device);
    // the autotuned launch settings of each handle apply for this call only
    // (plans, captures and launches); the handle's own are restored on exit
    TunedLaunch tuned(ops, count);
    if (seed < 0) seed = clock_seed();
    uint64_t base[4];
    thread_state(seed, 0, base);
    const auto c0 = std::chrono::steady_clock::now();
    // one-time uploads (jump tables) must not happen inside a capture
    SLQ_TRY(rng_words(ops[0]->stream, base, 1, 0, 0, nullptr));
    std::vector<Plan> plans(count);
    std::vector<size_t> sample_off(count);
    for (int i = 0; i < count; ++i)
    {
        slq_op* op = ops[i];
        op->prof.enabled = false;
        SLQ_TRY(prepare(op, gram, precision));
        const int vdt = value_dtype(op, precision);
        plans[i] = make_plan(op, gram, m, orthogonalize, 1, vdt == SLQ_F64 ? 8 : 4);
        sample_off[i] = plans[i].total;
        SLQ_TRY(op->ws.reserve(plans[i].total + align256(sizeof(double) * num_samples)));
        plans[i].base = (char*) op->ws.base;
    }
    SLQ_CHECK(hipDeviceSynchronize());
    // every block of problem i on its own stream
    auto enqueue_blocks = [&](int i) -> int {
        slq_op* op = ops[i];
        const Plan& p = plans[i];
        const int vdt = value_dtype(op, precision);
        double* dsamp = at<double>(p, sample_off[i]);
        for (int64_t s0 = 0; s0 < num_samples; s0 += kBlock)
        {
            const int cnt = (int) std::min<int64_t>(kBlock, num_samples - s0);
            SLQ_TRY(make_probes(op, p, base, s0, cnt, vdt));
            SLQ_TRY(recurrence_block(op, p, 0.0, gram, tol, orthogonalize, precision));
            SLQ_TRY(run_quadrature(op->stream, at<double>(p, p.off_alpha),
                                   at<double>(p, p.off_beta), at<int32_t>(p, p.off_sizes),
                                   cnt, m, gram, func, func_params, exponent, (double) p.len,
                                   nullptr, 1, 0, 0, 1, dsamp + s0, nullptr, nullptr));
        }
        return SLQ_OK;
    };
    auto enqueue = [&](int i) -> int { return enqueue_blocks(i); };
    double t_run = 0.0;
    if (use_graph)
    {
        hipStream_t main = nullptr;
        SLQ_CHECK(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
        std::vector<hipEvent_t> ev(count + 1, nullptr);
        int r = SLQ_OK;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        for (auto& e : ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                r = SLQ_E_HIP;
        if (r == SLQ_OK &&
            hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal) != hipSuccess)
            r = SLQ_E_HIP;
        if (r == SLQ_OK)
        {
            bool capturing = true;
            if (hipEventRecord(ev[count], main) != hipSuccess) r = SLQ_E_HIP;
            for (int i = 0; i < count && r == SLQ_OK; ++i)
            {
                if (hipStreamWaitEvent(ops[i]->stream, ev[count], 0) != hipSuccess)
                {
                    r = SLQ_E_HIP;
                    break;
                }
                r = enqueue(i);
                if (r == SLQ_OK && (hipEventRecord(ev[i], ops[i]->stream) != hipSuccess ||
                                    hipStreamWaitEvent(main, ev[i], 0) != hipSuccess))
                    r = SLQ_E_HIP;
            }
            hipError_t ee = hipStreamEndCapture(main, &graph);
            capturing = false;
            if (r == SLQ_OK && ee != hipSuccess) r = hip_status(ee, "hipStreamEndCapture", __FILE__, __LINE__);
            (void) capturing;
        }
        if (r == SLQ_OK)
        {
            hipError_t ee = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ee != hipSuccess) r = hip_status(ee, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        if (r == SLQ_OK)
        {
            if (setup_time)
                *setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
            const auto t0 = std::chrono::steady_clock::now();
            hipError_t ee = hipGraphLaunch(exec, main);
            if (ee == hipSuccess) ee = hipStreamSynchronize(main);
            t_run = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (ee != hipSuccess) r = hip_status(ee, "hipGraphLaunch", __FILE__, __LINE__);
        }
        if (exec) (void) hipGraphExecDestroy(exec);
        if (graph) (void) hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void) hipEventDestroy(e);
        (void) hipStreamDestroy(main);
        if (r != SLQ_OK)
        {
            if (g_last_error.empty()) set_error("slq_estimate_batch: graph capture failed");
            return r;
        }
    }
    else
    {
        if (setup_time)
            *setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < count; ++i)
        {
            SLQ_TRY(enqueue(i));
            SLQ_CHECK(hipStreamSynchronize(ops[i]->stream));
        }
        t_run = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (run_time) *run_time = t_run;
    std::vector<double> hs(num_samples);
    std::vector<int32_t> processed(num_samples);
    for (int k = 0; k < num_samples; ++k) processed[k] = k;
    for (int i = 0; i < count; ++i)
    {
        SLQ_CHECK(hipMemcpy(hs.data(), at<double>(plans[i], sample_off[i]),
                            sizeof(double) * num_samples, hipMemcpyDeviceToHost));
        if (samples) std::memcpy(samples + (size_t) i * num_samples, hs.data(),
                                 sizeof(double) * num_samples);
        int32_t used = num_samples, nout = 0;
        average_estimates(hs.data(), 1, confidence, outlier, &used, processed.data(), &nout,
                          trace + i, error + i);
    }
    return SLQ_OK;
}

// ================================================================= C ABI

extern "C" {

int slq_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return SLQ_E_ARG;
    std::snprintf(buf, len, "%s", g_last_error.c_str());
    return SLQ_OK;
}

int slq_device_properties(int* num_devices, int* num_multiprocessors,
                          int* num_threads_per_multiprocessor, char* name,
                          size_t name_len)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1)
    {
        set_error("no HIP device found (%s)", hipGetErrorString(e));
        return SLQ_E_NODEVICE;
    }
    hipDeviceProp_t prop;
    SLQ_CHECK(hipGetDeviceProperties(&prop, 0));
    if (num_devices) *num_devices = n;
    if (num_multiprocessors) *num_multiprocessors = prop.multiProcessorCount;
    if (num_threads_per_multiprocessor)
        *num_threads_per_multiprocessor = prop.maxThreadsPerMultiProcessor;
    if (name && name_len) std::snprintf(name, name_len, "%s", prop.gcnArchName);
    return SLQ_OK;
}

int slq_op_create_csr(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int64_t* indptr, const int32_t* indices,
                      const void* data, int dtype, slq_op** out)
{
    if (!out || nrows < 1 || ncols < 1 || nnz < 0 || !indptr ||
        (nnz > 0 && (!indices || !data)) || (dtype != SLQ_F64 && dtype != SLQ_F32))
    {
        set_error("slq_op_create_csr: invalid argument");
        return SLQ_E_ARG;
    }
    if (ncols > INT32_MAX)
    {
        set_error("slq_op_create_csr: more than 2^31-1 columns");
        return SLQ_E_ARG;
    }
    if (indptr[0] != 0 || indptr[nrows] != nnz)
    {
        set_error("slq_op_create_csr: indptr[0] must be 0 and indptr[nrows] == nnz");
        return SLQ_E_ARG;
    }
    // Every kernel gathers X[column]: reject pointers or columns out of range
    // here rather than let a kernel read outside its buffers.
    for (int64_t r = 0; r < nrows; ++r)
        if (indptr[r + 1] < indptr[r])
        {
            set_error("slq_op_create_csr: indptr decreases at row %lld", (long long) r);
            return SLQ_E_ARG;
        }
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || (int64_t) indices[k] >= ncols)
        {
            set_error("slq_op_create_csr: column index %d out of range at %lld",
                      indices[k], (long long) k);
            return SLQ_E_ARG;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    {
        set_error("no HIP device found");
        return SLQ_E_NODEVICE;
    }
    slq_op* op = new (std::nothrow) slq_op();
    if (!op)
    {
        set_error("host allocation failed");
        return SLQ_E_OOM;
    }
    (void) hipGetDevice(&op->device);
    op->A = std::make_shared<MatrixStore>();
    DevCSR& c = op->A->csr;
    c.nrows = nrows;
    c.ncols = ncols;
    c.nnz = nnz;
    c.dtype = dtype;
    c.sorted = true;
    for (int64_t r = 0; r < nrows && c.sorted; ++r)
        for (int64_t k = indptr[r] + 1; k < indptr[r + 1]; ++k)
            if (indices[k] < indices[k - 1])
            {
                c.sorted = false;
                break;
            }
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    auto fail = [&](int code) {
        delete op;
        return code;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(hip_status(e, "hipStreamCreate", __FILE__, __LINE__));
    if ((e = hipMalloc(&c.indptr, sizeof(int64_t) * (nrows + 1))) != hipSuccess ||
        (e = hipMalloc(&c.indices, sizeof(int32_t) * (nnz > 0 ? nnz : 1))) != hipSuccess ||
        (e = hipMalloc(&c.data, vb * (nnz > 0 ? nnz : 1))) != hipSuccess)
        return fail(hip_status(e, "hipMalloc", __FILE__, __LINE__));
    if ((e = hipMemcpy(c.indptr, indptr, sizeof(int64_t) * (nrows + 1),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (nnz > 0 && (e = hipMemcpy(c.indices, indices, sizeof(int32_t) * nnz,
                                   hipMemcpyHostToDevice)) != hipSuccess) ||
        (nnz > 0 && (e = hipMemcpy(c.data, data, vb * nnz, hipMemcpyHostToDevice)) != hipSuccess))
        return fail(hip_status(e, "hipMemcpy", __FILE__, __LINE__));
    {
        // 16-row tiles (matrix-core SpMM) unless SLQ_TILE_ROWS asks for 8 or
        // 0 (CSR only)
        const char* env = std::getenv("SLQ_TILE_ROWS");
        const int R = env ? std::atoi(env) : 16;
        if (R == 8 || R == 16)
        {
            int r = build_tiles(nrows, ncols, indptr, indices, data, dtype, R, 0.25,
                                op->A->tiles);
            if (r != SLQ_OK) return fail(r);
        }
    }
    *out = op;
    return SLQ_OK;
}

int slq_op_create_csc(int64_t nrows, int64_t ncols, int64_t nnz,
                      const int64_t* indptr, const int32_t* indices,
                      const void* data, int dtype, slq_op** out)
{
    if (!out || nrows < 1 || ncols < 1 || nnz < 0 || !indptr ||
        (nnz > 0 && (!indices || !data)) || (dtype != SLQ_F64 && dtype != SLQ_F32))
    {
        set_error("slq_op_create_csc: invalid argument");
        return SLQ_E_ARG;
    }
    if (nrows > INT32_MAX || ncols > INT32_MAX)
    {
        set_error("slq_op_create_csc: more than 2^31-1 rows or columns");
        return SLQ_E_ARG;
    }
    if (indptr[0] != 0 || indptr[ncols] != nnz)
    {
        set_error("slq_op_create_csc: indptr[0] must be 0 and indptr[ncols] == nnz");
        return SLQ_E_ARG;
    }
    for (int64_t c = 0; c < ncols; ++c)
        if (indptr[c + 1] < indptr[c])
        {
            set_error("slq_op_create_csc: indptr decreases at column %lld", (long long) c);
            return SLQ_E_ARG;
        }
    for (int64_t k = 0; k < nnz; ++k)
        if (indices[k] < 0 || (int64_t) indices[k] >= nrows)
        {
            set_error("slq_op_create_csc: row index %d out of range at %lld", indices[k],
                      (long long) k);
            return SLQ_E_ARG;
        }
    // CSR by a counting sort over rows; columns are visited in ascending
    // order, so every row's column indices come out sorted (and repeated
    // entries stay, in column order). order[j] is the CSC position of CSR
    // entry j, kept for new values given in CSC order.
    std::vector<int64_t> ip(nrows + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++ip[indices[k] + 1];
    for (int64_t r = 0; r < nrows; ++r) ip[r + 1] += ip[r];
    std::vector<int64_t> next(ip.begin(), ip.end() - 1);
    std::vector<int32_t> ix(nnz > 0 ? nnz : 1);
    std::vector<int64_t> order(nnz > 0 ? nnz : 1);
    for (int64_t c = 0; c < ncols; ++c)
        for (int64_t k = indptr[c]; k < indptr[c + 1]; ++k)
        {
            const int64_t j = next[indices[k]]++;
            ix[j] = (int32_t) c;
            order[j] = k;
        }
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    std::vector<char> dv(vb * (nnz > 0 ? nnz : 1));
    for (int64_t j = 0; j < nnz; ++j)
        std::memcpy(dv.data() + vb * j, (const char*) data + vb * order[j], vb);
    SLQ_TRY(slq_op_create_csr(nrows, ncols, nnz, ip.data(), ix.data(), dv.data(), dtype, out));
    (*out)->A->csc_order.swap(order);
    if (nnz == 0) (*out)->A->csc_order.clear();
    return SLQ_OK;
}

}  // extern "C"

namespace {

// Counter-based hash: the SplitMix64 output function of a combination of the
// seed, the row and the stratum (Steele, Lea and Flood 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Row r of the sparse Gaussian matrix has one entry in each of k column
// strata [floor(s n / k), floor((s + 1) n / k)): a hashed column of the
// stratum (so the k columns are distinct and ascending) and an N(0, 1)
// value by the Box-Muller transform of two hashed uniforms.
template <typename T>
__global__ void sparse_gaussian_kernel(int64_t nrows, int64_t ncols, int32_t k, uint64_t seed,
                                       int32_t* __restrict__ indices, T* __restrict__ data)
{
    const int64_t total = nrows * k;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t) gridDim.x * blockDim.x)
    {
        const int64_t r = i / k;
        const int64_t st = i - r * k;
        const int64_t c0 = st * ncols / k;
        const int64_t c1 = (st + 1) * ncols / k;
        const uint64_t key = seed * 0x9e3779b97f4a7c15ull + (uint64_t) r * 0xd1b54a32d192ed03ull +
                             (uint64_t) st;
        const uint64_t h0 = mix64(key);
        const uint64_t h1 = mix64(key ^ 0x5851f42d4c957f2dull);
        const uint64_t h2 = mix64(key ^ 0x14057b7ef767814full);
        indices[i] = (int32_t) (c0 + (int64_t) (h0 % (uint64_t) (c1 - c0)));
        const double u1 = (double) ((h1 >> 11) + 1) * 0x1.0p-53;   // (0, 1]
        const double u2 = (double) (h2 >> 11) * 0x1.0p-53;         // [0, 1)
        data[i] = (T) (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    }
}

__global__ void row_pointers_kernel(int64_t* __restrict__ indptr, int64_t nrows, int32_t k)
{
    const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= nrows) indptr[r] = r * k;
}

}  // namespace

extern "C" {

int slq_op_create_sparse_gaussian(int64_t nrows, int64_t ncols, int32_t per_row, int64_t seed,
                                  int dtype, slq_op** out)
{
    if (!out || nrows < 1 || ncols < 1 || ncols > INT32_MAX || per_row < 1 ||
        per_row > ncols || (dtype != SLQ_F64 && dtype != SLQ_F32))
    {
        set_error("slq_op_create_sparse_gaussian: invalid argument");
        return SLQ_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    {
        set_error("no HIP device found");
        return SLQ_E_NODEVICE;
    }
    slq_op* op = new (std::nothrow) slq_op();
    if (!op)
    {
        set_error("host allocation failed");
        return SLQ_E_OOM;
    }
    (void) hipGetDevice(&op->device);
    op->A = std::make_shared<MatrixStore>();
    DevCSR& c = op->A->csr;
    c.nrows = nrows;
    c.ncols = ncols;
    c.nnz = nrows * per_row;
    c.dtype = dtype;
    c.sorted = true;   // one entry per column stratum, strata in order
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    auto fail = [&](int code) {
        delete op;
        return code;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&c.indptr, sizeof(int64_t) * (nrows + 1))) != hipSuccess ||
        (e = hipMalloc(&c.indices, sizeof(int32_t) * c.nnz)) != hipSuccess ||
        (e = hipMalloc(&c.data, vb * c.nnz)) != hipSuccess)
        return fail(hip_status(e, "hipMalloc", __FILE__, __LINE__));
    hipLaunchKernelGGL(row_pointers_kernel, dim3((unsigned) ((nrows + 256) / 256)), dim3(256),
                       0, op->stream, c.indptr, nrows, per_row);
    if (dtype == SLQ_F64)
        hipLaunchKernelGGL(sparse_gaussian_kernel<double>, dim3(8192), dim3(256), 0, op->stream,
                           nrows, ncols, per_row, (uint64_t) seed, c.indices, (double*) c.data);
    else
        hipLaunchKernelGGL(sparse_gaussian_kernel<float>, dim3(8192), dim3(256), 0, op->stream,
                           nrows, ncols, per_row, (uint64_t) seed, c.indices, (float*) c.data);
    if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(op->stream)) != hipSuccess)
        return fail(hip_status(e, "sparse Gaussian generation", __FILE__, __LINE__));
    // no tiled form: random columns leave 16-row slabs about 1/16 full
    *out = op;
    return SLQ_OK;
}

int slq_op_get_csr(slq_op* op, int64_t* indptr, int32_t* indices, void* data)
{
    if (!op || op->kind != OP_MATRIX || !indptr)
    {
        set_error("slq_op_get_csr: a matrix handle and an indptr buffer are required");
        return SLQ_E_ARG;
    }
    if (op->A->dense)
    {
        set_error("slq_op_get_csr: the handle holds a dense matrix");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    const DevCSR& c = op->A->csr;
    const size_t vb = c.dtype == SLQ_F64 ? 8 : 4;
    SLQ_CHECK(hipStreamSynchronize(op->stream));
    SLQ_CHECK(hipMemcpy(indptr, c.indptr, sizeof(int64_t) * (c.nrows + 1),
                        hipMemcpyDeviceToHost));
    if (c.nnz > 0 && indices)
        SLQ_CHECK(hipMemcpy(indices, c.indices, sizeof(int32_t) * c.nnz, hipMemcpyDeviceToHost));
    if (c.nnz > 0 && data)
        SLQ_CHECK(hipMemcpy(data, c.data, vb * c.nnz, hipMemcpyDeviceToHost));
    return SLQ_OK;
}

int slq_op_shape(slq_op* op, int64_t* nrows, int64_t* ncols, int64_t* nnz, int32_t* dtype)
{
    if (!op)
    {
        set_error("slq_op_shape: null handle");
        return SLQ_E_ARG;
    }
    const DevCSR& c = op->A->csr;
    if (nrows) *nrows = c.nrows;
    if (ncols) *ncols = c.ncols;
    if (nnz) *nnz = c.nnz;
    if (dtype) *dtype = c.dtype;
    return SLQ_OK;
}

int slq_op_create_affine(slq_op* A, slq_op* B, slq_op** out)
{
    if (!A || !out || A->kind != OP_MATRIX || (B && B->kind != OP_MATRIX))
    {
        set_error("slq_op_create_affine: A (and B) must be matrix handles");
        return SLQ_E_ARG;
    }
    if (B && (B->A->csr.nrows != A->A->csr.nrows || B->A->csr.ncols != A->A->csr.ncols ||
              B->A->csr.dtype != A->A->csr.dtype))
    {
        set_error("slq_op_create_affine: B must have the shape and dtype of A");
        return SLQ_E_ARG;
    }
    if (B && (A->A->dense || B->A->dense))
    {
        set_error("slq_op_create_affine: A + tB with a general B needs CSR handles for A "
                  "and B");
        return SLQ_E_ARG;
    }
    slq_op* op = new (std::nothrow) slq_op();
    if (!op)
    {
        set_error("host allocation failed");
        return SLQ_E_OOM;
    }
    op->device = A->device;
    op->kind = B ? OP_AFFINE_B : OP_AFFINE_I;
    op->A = A->A;
    if (B) op->B = B->A;
    hipError_t e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        delete op;
        return hip_status(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    *out = op;
    return SLQ_OK;
}

int slq_op_create_dense(int64_t nrows, int64_t ncols, const void* data, int dtype,
                        slq_op** out)
{
    if (!out || nrows < 1 || ncols < 1 || !data || (dtype != SLQ_F64 && dtype != SLQ_F32))
    {
        set_error("slq_op_create_dense: invalid argument");
        return SLQ_E_ARG;
    }
    if (nrows > INT32_MAX || ncols > INT32_MAX)
    {
        set_error("slq_op_create_dense: more than 2^31-1 rows or columns");
        return SLQ_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    {
        set_error("no HIP device found");
        return SLQ_E_NODEVICE;
    }
    slq_op* op = new (std::nothrow) slq_op();
    if (!op)
    {
        set_error("host allocation failed");
        return SLQ_E_OOM;
    }
    (void) hipGetDevice(&op->device);
    op->A = std::make_shared<MatrixStore>();
    MatrixStore& M = *op->A;
    M.dense = true;
    M.csr.nrows = nrows;
    M.csr.ncols = ncols;
    M.csr.nnz = nrows * ncols;
    M.csr.dtype = dtype;
    M.dn.ld = (ncols + kDenseLd - 1) / kDenseLd * kDenseLd;
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    auto fail = [&](int code) {
        delete op;
        return code;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(hip_status(e, "hipStreamCreate", __FILE__, __LINE__));
    const size_t bytes = (size_t) nrows * (size_t) M.dn.ld * vb;
    if ((e = hipMalloc(&M.dn.data, bytes)) != hipSuccess)
        return fail(hip_status(e, "dense upload", __FILE__, __LINE__));
    if ((e = hipMemset(M.dn.data, 0, bytes)) != hipSuccess ||
        (e = hipMemcpy2D(M.dn.data, (size_t) M.dn.ld * vb, data, (size_t) ncols * vb,
                         (size_t) ncols * vb, (size_t) nrows, hipMemcpyHostToDevice)) !=
            hipSuccess)
        return fail(hip_status(e, "dense upload", __FILE__, __LINE__));
    *out = op;
    return SLQ_OK;
}

}  // extern "C"

namespace {

// value_k = w_k exp(-d_k / scale) + c_k, rounded once to the matrix type.
template <typename T>
__global__ __launch_bounds__(256) void family_kernel(const double* __restrict__ f, int64_t nnz,
                                                     double scale, T* __restrict__ out)
{
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    for (int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride)
        out[k] = (T) (f[nnz + k] * exp(-f[k] / scale) + f[2 * nnz + k]);
}

int family_fill(hipStream_t s, const double* f, int64_t nnz, double scale, int dtype,
                void* out)
{
    if (nnz <= 0) return SLQ_OK;
    const int64_t g = (nnz + 255) / 256;
    const dim3 grid((unsigned) (g < 8192 ? g : 8192));
    if (dtype == SLQ_F64)
        hipLaunchKernelGGL(family_kernel<double>, grid, dim3(256), 0, s, f, nnz, scale,
                           (double*) out);
    else
        hipLaunchKernelGGL(family_kernel<float>, grid, dim3(256), 0, s, f, nnz, scale,
                           (float*) out);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

// After new CSR values of a matrix handle: the tiled layout is refilled on
// the device and the derived forms are rebuilt from the new values on
// their next use.
int values_changed(slq_op* op)
{
    MatrixStore& M = *op->A;
    SLQ_TRY(tiles_refill(op->stream, M.tiles, M.csr.data, M.csr.dtype));
    if (M.data_f32) (void) hipFree(M.data_f32);
    M.data_f32 = nullptr;
    if (M.trans_data_f32) (void) hipFree(M.trans_data_f32);
    M.trans_data_f32 = nullptr;
    if (M.has_trans)
    {
        (void) hipFree(M.trans.indptr);
        (void) hipFree(M.trans.indices);
        (void) hipFree(M.trans.data);
        M.trans = DevCSR();
        M.has_trans = false;
    }
    free_tiles(M.trans_tiles);
    M.trans_tiles_tried = false;
    SLQ_CHECK(hipStreamSynchronize(op->stream));
    return SLQ_OK;
}

}  // namespace

extern "C" {

int slq_op_set_values(slq_op* op, const void* data)
{
    if (!op || op->kind != OP_MATRIX || (!data && op->A->csr.nnz > 0))
    {
        set_error("slq_op_set_values: a matrix handle and its new values are required");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    MatrixStore& M = *op->A;
    if (M.dense)
    {
        // new row-major values: replace the padded copy, drop derived forms
        const size_t vb = M.csr.dtype == SLQ_F64 ? 8 : 4;
        SLQ_CHECK(hipStreamSynchronize(op->stream));
        SLQ_CHECK(hipMemcpy2D(M.dn.data, (size_t) M.dn.ld * vb, data,
                              (size_t) M.csr.ncols * vb, (size_t) M.csr.ncols * vb,
                              (size_t) M.csr.nrows, hipMemcpyHostToDevice));
        for (void** p : {&M.dn.trans, (void**) &M.dn.data_f32, (void**) &M.dn.trans_f32})
            if (*p)
            {
                (void) hipFree(*p);
                *p = nullptr;
            }
        return SLQ_OK;
    }
    if (M.tiles.R > 0 && !M.tiles.src)
    {
        set_error("slq_op_set_values: the matrix has duplicate entries or too many "
                  "non-zeros for an in-place value refresh; create a new operator");
        return SLQ_E_ARG;
    }
    SLQ_CHECK(hipDeviceSynchronize());
    const size_t vb = M.csr.dtype == SLQ_F64 ? 8 : 4;
    if (M.csr.nnz > 0 && !M.csc_order.empty())
    {
        // values in the CSC order of slq_op_create_csc
        std::vector<char> perm(vb * M.csr.nnz);
        for (int64_t j = 0; j < M.csr.nnz; ++j)
            std::memcpy(perm.data() + vb * j, (const char*) data + vb * M.csc_order[j], vb);
        SLQ_CHECK(hipMemcpy(M.csr.data, perm.data(), perm.size(), hipMemcpyHostToDevice));
    }
    else if (M.csr.nnz > 0)
        SLQ_CHECK(hipMemcpy(M.csr.data, data, vb * M.csr.nnz, hipMemcpyHostToDevice));
    return values_changed(op);
}

int slq_op_family_attach(slq_op* op, const double* d, const double* w, const double* c)
{
    if (!op || op->kind != OP_MATRIX || op->A->dense || !op->A->csc_order.empty() ||
        (op->A->csr.nnz > 0 && (!d || !w || !c)))
    {
        set_error("slq_op_family_attach: a CSR matrix handle and its d, w, c arrays are "
                  "required");
        return SLQ_E_ARG;
    }
    MatrixStore& M = *op->A;
    if (M.tiles.R > 0 && !M.tiles.src)
    {
        set_error("slq_op_family_attach: the matrix has duplicate entries or too many "
                  "non-zeros for an in-place value refresh");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    const int64_t nnz = M.csr.nnz;
    if (!M.family)
        SLQ_CHECK(hipMalloc(&M.family, sizeof(double) * 3 * (size_t) (nnz > 0 ? nnz : 1)));
    if (nnz > 0)
    {
        SLQ_CHECK(hipMemcpy(M.family, d, sizeof(double) * nnz, hipMemcpyHostToDevice));
        SLQ_CHECK(hipMemcpy(M.family + nnz, w, sizeof(double) * nnz, hipMemcpyHostToDevice));
        SLQ_CHECK(hipMemcpy(M.family + 2 * nnz, c, sizeof(double) * nnz,
                            hipMemcpyHostToDevice));
    }
    return SLQ_OK;
}

int slq_op_family_values(slq_op* op, double scale)
{
    if (!op || op->kind != OP_MATRIX || !op->A->family || !(scale > 0.0))
    {
        set_error("slq_op_family_values: a matrix handle with an attached family and a "
                  "positive scale are required");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    MatrixStore& M = *op->A;
    // work on other handles sharing the matrix (A + tI) may still read it
    SLQ_CHECK(hipDeviceSynchronize());
    SLQ_TRY(family_fill(op->stream, M.family, M.csr.nnz, scale, M.csr.dtype, M.csr.data));
    return values_changed(op);
}

int slq_op_destroy(slq_op* op)
{
    if (!op) return SLQ_OK;
    if (op->stream) (void) hipStreamSynchronize(op->stream);
    for (hipStream_t ls : op->lane_streams) (void) hipStreamSynchronize(ls);
    if (op->aux_stream) (void) hipStreamSynchronize(op->aux_stream);
    prof_collect(op);
    for (hipEvent_t e : op->prof.pool) (void) hipEventDestroy(e);
    release_solver(op);
    if (op->stream) (void) hipStreamDestroy(op->stream);
    for (hipStream_t ls : op->lane_streams) (void) hipStreamDestroy(ls);
    if (op->aux_stream) (void) hipStreamDestroy(op->aux_stream);
    delete op;
    return SLQ_OK;
}

}  // extern "C"

namespace {

// Arguments of one estimate, shared by slq_estimate and slq_estimate_noalloc.
struct EstimateArgs {
    const double* params;
    int32_t q, func;
    const double* func_params;
    int32_t gram;
    double exponent;
    int32_t orthogonalize;
    int64_t seed;
    int32_t m;
    double tol;
    int32_t min_s, max_s;
    double atol, rtol, confidence, outlier;
    int32_t block_size, precision;
    double *trace, *error, *samples;
    int32_t *processed, *num_used, *num_outliers, *converged, *all_converged;
    double* alg_wall_time;
    // 1: block Lanczos over each 64-probe block (slq_estimate_block_krylov)
    int32_t krylov = 0;
};

int check_estimate(const char* who, const slq_op* op, const EstimateArgs& a)
{
    if (!op || a.q < 1 || a.m < 1 || a.min_s < 1 || a.max_s < a.min_s || !a.trace ||
        !a.error || !a.samples || !a.processed || !a.num_used || !a.num_outliers ||
        !a.converged || !func_ok(a.func, a.func_params) ||
        (a.precision != SLQ_PREC_NATIVE && a.precision != SLQ_PREC_MIXED))
    {
        set_error("%s: invalid argument", who);
        return SLQ_E_ARG;
    }
    if (op->kind != OP_MATRIX && !a.params)
    {
        set_error("%s: an affine operator needs parameters", who);
        return SLQ_E_ARG;
    }
    if (op->kind == OP_MATRIX && a.q != 1)
    {
        set_error("%s: a matrix operator has exactly one inquiry", who);
        return SLQ_E_ARG;
    }
    if (!a.gram && op->A->csr.nrows != op->A->csr.ncols)
    {
        set_error("%s: Lanczos needs a square operator", who);
        return SLQ_E_ARG;
    }
    return SLQ_OK;
}

// The recurrences and quadratures of one probe block for every unconverged
// t of an A + tB estimate, p.lanes at a time side by side: each t of a group
// runs on its own lane (stream, buffers, pair-tile values) after the probes
// are made on the handle's stream, and the handle's stream waits for the
// whole group before the next one and before the samples are read. Every
// t's kernels and inputs are those of the one-after-the-other path, so the
// samples are the same bit for bit.
int lanes_block(slq_op* op, const Plan& p, const EstimateArgs& a, int cnt, double nq,
                bool count_steps, std::vector<int32_t>& block_sizes, std::vector<char>& ran)
{
    while ((int) op->lane_streams.size() < p.lanes)
    {
        hipStream_t ls = nullptr;
        SLQ_CHECK(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
        op->lane_streams.push_back(ls);
    }
    std::vector<int> todo;
    for (int j = 0; j < a.q; ++j)
        if (!a.converged[j]) todo.push_back(j);
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join(p.lanes, nullptr);
    int r = SLQ_OK;
    hipError_t ee = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (auto& e : join)
        if (ee == hipSuccess) ee = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (ee != hipSuccess) r = hip_status(ee, "hipEventCreate", __FILE__, __LINE__);
    double* samples_dev = at<double>(p, p.off_samples);
    for (size_t g = 0; r == SLQ_OK && g < todo.size(); g += p.lanes)
    {
        const int n_in = (int) std::min<size_t>(p.lanes, todo.size() - g);
        if (hipEventRecord(fork, op->stream) != hipSuccess)
        {
            r = SLQ_E_HIP;
            break;
        }
        for (int c = 0; c < n_in && r == SLQ_OK; ++c)
        {
            const int j = todo[g + c];
            hipStream_t ls = op->lane_streams[c];
            const size_t o = p.off_lanes + (size_t) c * p.lane_stride;
            LaneOp lane(op, ls, at<void>(p, o + p.lane_vals), at<void>(p, o + p.lane_vals_t));
            if (hipStreamWaitEvent(ls, fork, 0) != hipSuccess) r = SLQ_E_HIP;
            if (r == SLQ_OK)
                r = recurrence_lane(&lane.op, p, c, a.params[j], a.gram, a.tol,
                                    a.orthogonalize, a.precision);
            if (r == SLQ_OK)
                r = run_quadrature(ls, at<double>(p, o + p.lane_alpha),
                                   at<double>(p, o + p.lane_beta),
                                   at<int32_t>(p, o + p.lane_sizes), cnt, a.m, a.gram, a.func,
                                   a.func_params, a.exponent, nq, nullptr, a.q, 0, j, a.q,
                                   samples_dev, nullptr, nullptr);
            if (r == SLQ_OK && count_steps)
            {
                if (hipMemcpyAsync(block_sizes.data() + (size_t) j * kBlock,
                                   at<int32_t>(p, o + p.lane_sizes), sizeof(int32_t) * kBlock,
                                   hipMemcpyDeviceToHost, ls) != hipSuccess)
                    r = SLQ_E_HIP;
                ran[j] = 1;
            }
            if (r == SLQ_OK && (hipEventRecord(join[c], ls) != hipSuccess ||
                                hipStreamWaitEvent(op->stream, join[c], 0) != hipSuccess))
                r = SLQ_E_HIP;
        }
    }
    if (r != SLQ_OK)
    {
        if (g_last_error.empty()) set_error("A + tB lanes: HIP call failed");
        // no lane may still write the workspace when the caller returns
        for (hipStream_t ls : op->lane_streams) (void) hipStreamSynchronize(ls);
    }
    // otherwise the handle's stream has waited for every lane that ran
    if (fork) (void) hipEventDestroy(fork);
    for (auto& e : join)
        if (e) (void) hipEventDestroy(e);
    return r;
}

// The block Krylov form's loop (slq_estimate_block_krylov): blocks of 64
// probes in groups, each block but the last of a group ending before
// min_num_samples, so the convergence test cannot stop the loop inside it.
// Every recurrence of a group (one per block, or per unconverged inquiry of
// A + tB with a general B) keeps its block tridiagonal in a slot of its own;
// the slots' eigensolves then run as one batched call (a group with more
// recurrences than slots is solved a slotful at a time). The samples of the
// group come back in one copy and are tested in order by `process`.
template <typename Process>
int krylov_groups(slq_op* op, const Plan& p, const EstimateArgs& a, const uint64_t* base,
                  int vdt, const double* params_dev, const int32_t* converged,
                  const int& all_conv, Process&& process)
{
    const int q = a.q;
    const bool relation = op->kind == OP_AFFINE_I;
    const int required = relation ? 1 : q;
    const int S = p.bk_slots;
    double* gsamples = at<double>(p, p.off_bk_samples);   // [S][64][q]
    void* scratch = at<void>(p, p.off_scratch);
    int* flags_dev = nullptr;
    BlockKrylovArgs c;   // fields every phase reads
    c.op = op;
    c.m = a.m;
    c.value_dtype = vdt;
    c.flags_out = &flags_dev;
    int64_t s0 = 0;
    while (!all_conv && s0 < a.max_s)
    {
        std::vector<int> todo;
        for (int j = 0; j < required; ++j)
            if (relation || !converged[j]) todo.push_back(j);
        const int per = (int) todo.size();
        int g = 1;
        for (;;)
        {
            const int64_t end = std::min<int64_t>(s0 + (int64_t) g * kBlock, a.max_s);
            if (end >= a.min_s || end >= a.max_s || g >= S || (g + 1) * per > S) break;
            ++g;
        }
        BlockKrylovArgs z = c;
        z.phase = 0;
        SLQ_TRY(run_block_krylov(z, scratch, p.scratch_bytes));
        BlockKrylovArgs solve = c;
        solve.phase = 2;
        solve.func = a.func;
        solve.func_params = a.func_params;
        solve.exponent = a.exponent;
        solve.params_dev = relation ? params_dev : nullptr;
        solve.q = q;
        solve.shift = relation ? 1 : 0;
        solve.q_stride = q;
        int slot = 0;
        for (int k = 0; k < g; ++k)
        {
            // every probe of the block (also past max_num_samples): the
            // block's basis needs all 64
            SLQ_TRY(make_probes(op, p, base, s0 + (int64_t) k * kBlock, kBlock, vdt));
            for (int j : todo)
            {
                BlockKrylovArgs r = c;
                r.phase = 1;
                r.slot = slot;
                r.t = op->kind != OP_MATRIX ? a.params[j] : 0.0;
                r.X0 = at<void>(p, p.off_X0);
                SLQ_TRY(run_block_krylov(r, scratch, p.scratch_bytes));
                solve.j_out[slot] = j;
                solve.samples_dev[slot] = gsamples + (size_t) k * kBlock * q;
                if (++slot == S)
                {
                    solve.nslots = slot;
                    SLQ_TRY(run_block_krylov(solve, scratch, p.scratch_bytes));
                    slot = 0;
                }
            }
        }
        if (slot > 0)
        {
            solve.nslots = slot;
            SLQ_TRY(run_block_krylov(solve, scratch, p.scratch_bytes));
        }
        const int64_t gend = std::min<int64_t>(s0 + (int64_t) g * kBlock, a.max_s);
        int rank_lost = 0;
        SLQ_CHECK(hipMemcpyAsync(a.samples + s0 * q, gsamples, sizeof(double) * (gend - s0) * q,
                                 hipMemcpyDeviceToHost, op->stream));
        SLQ_CHECK(hipMemcpyAsync(&rank_lost, flags_dev, sizeof(int), hipMemcpyDeviceToHost,
                                 op->stream));
        SLQ_CHECK(hipStreamSynchronize(op->stream));
        if (rank_lost)
        {
            // a basis that lost rank makes the block's samples meaningless
            set_error("block Krylov: the block basis lost rank (lanczos_degree x 64 = %d too "
                      "large for this %lld-row matrix)", a.m * kBlock,
                      (long long) op->A->csr.nrows);
            return SLQ_E_ARG;
        }
        for (int k = 0; k < g; ++k)
        {
            const int64_t b0 = s0 + (int64_t) k * kBlock;
            process(b0, (int) std::min<int64_t>(kBlock, a.max_s - b0), 0);
        }
        s0 = gend;
    }
    return SLQ_OK;
}

// The sampling loop of one estimate on the workspace p.base (sized by p),
// following cu_trace_estimator.cu:197-310 with one device: probes in blocks
// of 64 in sample order, one host synchronisation per block (the block's
// samples are copied straight into the caller's `samples` rows), then the
// reference's per-sample convergence test. Allocates no device memory.
int estimate_loop(slq_op* op, const Plan& p, EstimateArgs a)
{
    const int q = a.q;
    const int bs = (a.block_size <= 0 || a.block_size > kBlock) ? kBlock : a.block_size;
    const bool relation = op->kind == OP_AFFINE_I;
    const int required = relation ? 1 : q;
    const int vdt = value_dtype(op, a.precision);
    if (a.seed < 0) a.seed = clock_seed();
    double* params_dev = at<double>(p, p.off_params);
    if (op->kind != OP_MATRIX)
        SLQ_CHECK(hipMemcpyAsync(params_dev, a.params, sizeof(double) * q,
                                 hipMemcpyHostToDevice, op->stream));
    uint64_t base[4];
    thread_state(a.seed, 0, base);
    const double nq = (double) p.len;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    double* samples = a.samples;
    int32_t* converged = a.converged;

    for (int64_t i = 0; i < (int64_t) a.max_s * q; ++i) samples[i] = nan;
    for (int j = 0; j < q; ++j)
    {
        converged[j] = 0;
        a.num_used[j] = 0;
        a.num_outliers[j] = 0;
    }
    // One recurrence per block (a matrix, or A + tI with its Ritz shifts):
    // the quadrature and copy-back of a block run on a second stream, and
    // when the convergence test cannot stop the loop within a block (fewer
    // than min_num_samples processed by its end) the next block's probes
    // and recurrence are queued before the host waits for it. The kernels
    // and their inputs are those of the plain order, so results are equal.
    const bool pipelined = p.pipeline && p.lanes == 0 && required == 1 && !a.krylov;
    if (pipelined && !op->aux_stream)
        SLQ_CHECK(hipStreamCreateWithFlags(&op->aux_stream, hipStreamNonBlocking));
    hipStream_t qstream = pipelined ? op->aux_stream : op->stream;
    // completed recurrence steps per block, kept only while profiling
    const bool count_steps = op->prof.enabled;
    std::vector<int32_t> block_sizes[2];
    std::vector<char> ran[2];
    for (int k = 0; k < 2; ++k)
    {
        block_sizes[k].assign(count_steps ? (size_t) kBlock * required : 0, 0);
        ran[k].assign(count_steps ? required : 0, 0);
    }
    hipEvent_t rec_done[2] = {nullptr, nullptr};
    struct Events {
        hipEvent_t* e;
        ~Events()
        {
            for (int k = 0; k < 2; ++k)
                if (e[k]) (void) hipEventDestroy(e[k]);
        }
    } events{rec_done};
    if (pipelined)
        for (int k = 0; k < 2; ++k)
            SLQ_CHECK(hipEventCreateWithFlags(&rec_done[k], hipEventDisableTiming));

    // probes, recurrences and quadratures of the block at s0 (outputs of
    // parity par when pipelined)
    auto enqueue = [&](int64_t s0, int par) -> int {
        const int cnt = (int) std::min<int64_t>(bs, a.max_s - s0);
        std::fill(ran[par].begin(), ran[par].end(), 0);
        SLQ_TRY(make_probes(op, p, base, s0, cnt, vdt));
        if (p.lanes > 0)
            return lanes_block(op, p, a, cnt, nq, count_steps, block_sizes[par], ran[par]);
        for (int j = 0; j < required; ++j)
        {
            if (!relation && converged[j]) continue;   // masked below anyway
            const double t = op->kind != OP_MATRIX ? a.params[j] : 0.0;
            SLQ_TRY(recurrence_block(op, p, t, a.gram, a.tol, a.orthogonalize, a.precision,
                                     par));
            if (pipelined)
            {
                SLQ_CHECK(hipEventRecord(rec_done[par], op->stream));
                SLQ_CHECK(hipStreamWaitEvent(qstream, rec_done[par], 0));
            }
            SLQ_TRY(run_quadrature(qstream, at<double>(p, p.alpha_of(par)),
                                   at<double>(p, p.beta_of(par)),
                                   at<int32_t>(p, p.sizes_of(par)), cnt, a.m, a.gram, a.func,
                                   a.func_params, a.exponent, nq,
                                   relation ? params_dev : nullptr, q, relation ? 1 : 0, j, q,
                                   at<double>(p, p.samples_of(par)), nullptr, nullptr));
            if (count_steps && !pipelined)
                SLQ_CHECK(hipMemcpyAsync(block_sizes[par].data() + (size_t) j * kBlock,
                                         at<int32_t>(p, p.sizes_of(par)),
                                         sizeof(int32_t) * kBlock, hipMemcpyDeviceToHost,
                                         qstream));
            if (count_steps) ran[par][j] = 1;
        }
        return SLQ_OK;
    };

    int nproc = 0;
    int all_conv = 0;
    // the convergence test over the samples of the block at s0 (cnt of them,
    // in host memory), in order; samples past the stopping one become NaN
    auto process = [&](int64_t s0, int cnt, int par) {
        for (int pi = 0; pi < cnt; ++pi)
        {
            const int64_t i = s0 + pi;
            if (all_conv)
            {
                // past the stopping sample: not processed
                for (int j = 0; j < q; ++j) samples[i * q + j] = nan;
                continue;
            }
            if (relation && q > 1)
                for (int j = 0; j < q; ++j) converged[j] = 0;
            for (int j = 0; j < q; ++j)
                if (converged[j]) samples[i * q + j] = nan;
            if (count_steps)
                for (int j = 0; j < required; ++j)
                    if (ran[par][j])
                    {
                        const int32_t sz = block_sizes[par][(size_t) j * kBlock + pi];
                        op->prof.steps += sz > 0 ? sz : a.m;
                    }
            a.processed[nproc++] = (int32_t) i;
            all_conv = check_convergence(samples, q, a.min_s, a.processed, nproc, a.confidence,
                                         a.atol, a.rtol, a.error, a.num_used, converged);
        }
    };
    SLQ_CHECK(hipStreamSynchronize(op->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if (a.krylov)
    {
        SLQ_TRY(krylov_groups(op, p, a, base, vdt, params_dev, converged, all_conv, process));
    }
    else
    {
        int64_t s0 = 0;
        int par = 0;
        SLQ_TRY(enqueue(0, 0));
        for (;;)
        {
            const int cnt = (int) std::min<int64_t>(bs, a.max_s - s0);
            const int64_t s1 = s0 + bs;
            const int nxt = pipelined ? par ^ 1 : 0;
            bool ahead = false;
            if (pipelined && s1 < a.max_s && nproc + cnt < a.min_s)
            {
                SLQ_TRY(enqueue(s1, nxt));
                ahead = true;
            }
            // copies to host memory wait for the stream, so they are issued
            // only here, when the host needs the block (a pipelined block's
            // sizes too)
            if (count_steps && pipelined)
                SLQ_CHECK(hipMemcpyAsync(block_sizes[par].data(),
                                         at<int32_t>(p, p.sizes_of(par)),
                                         sizeof(int32_t) * kBlock, hipMemcpyDeviceToHost,
                                         qstream));
            SLQ_CHECK(hipMemcpyAsync(samples + s0 * q, at<double>(p, p.samples_of(par)),
                                     sizeof(double) * cnt * q, hipMemcpyDeviceToHost, qstream));
            SLQ_CHECK(hipStreamSynchronize(qstream));
            process(s0, cnt, par);
            if (all_conv || s1 >= a.max_s) break;
            if (!ahead) SLQ_TRY(enqueue(s1, nxt));
            s0 = s1;
            par = nxt;
        }
    }
    SLQ_CHECK(hipStreamSynchronize(op->stream));
    const auto t1 = std::chrono::steady_clock::now();
    if (a.alg_wall_time) *a.alg_wall_time = std::chrono::duration<double>(t1 - t0).count();
    for (int i = nproc; i < a.max_s; ++i) a.processed[i] = 0;
    average_estimates(samples, q, a.confidence, a.outlier, a.num_used, a.processed,
                      a.num_outliers, a.trace, a.error);
    if (a.all_converged) *a.all_converged = all_conv;
    prof_collect(op);
    return SLQ_OK;
}

// True when every derived form an estimate of this kind reads already exists
// on the handle, i.e. prepare() would allocate nothing. Mirrors prepare().
bool prepared(const slq_op* op, int gram, int precision)
{
    const MatrixStore& A = *op->A;
    const bool mixed = precision == SLQ_PREC_MIXED;
    if (A.dense)
    {
        if (gram && !A.dn.trans) return false;
        if (mixed && A.csr.dtype == SLQ_F64 &&
            (!A.dn.data_f32 || (gram && !A.dn.trans_f32)))
            return false;
        return true;
    }
    auto f32_ready = [gram](const MatrixStore& M) {
        if (M.csr.dtype != SLQ_F64) return true;
        if (M.tiles.R > 0 && !M.tiles.vals_f32) return false;
        if (gram && M.trans_tiles.R > 0 && !M.trans_tiles.vals_f32) return false;
        if (!M.data_f32) return false;
        if (gram && M.has_trans && !M.trans_data_f32) return false;
        return true;
    };
    if (gram)
    {
        if (!A.has_trans) return false;
        if (A.tiles.R > 0 && !A.trans_tiles_tried) return false;
        if (op->kind == OP_AFFINE_B && !op->B->has_trans) return false;
    }
    if (op->kind == OP_AFFINE_B)
    {
        if (!op->ab.tried || (gram && !op->ab_t.tried)) return false;
        if (mixed && A.csr.dtype == SLQ_F64 &&
            ((op->ab.tiles.R > 0 && !op->ab.tiles.vals_f32) ||
             (gram && op->ab_t.tiles.R > 0 && !op->ab_t.tiles.vals_f32)))
            return false;
    }
    if (mixed)
    {
        if (!f32_ready(A)) return false;
        if (op->kind == OP_AFFINE_B && !f32_ready(*op->B)) return false;
    }
    return true;
}

}  // namespace

// The no-allocation entry point's workspace: one array per device, reached
// through the address hipGetSymbolAddress gives on that device (the array's
// name in host code is only a host-side stand-in). Every call on a device
// shares it, so calls that overlap on one device race.
static constexpr size_t kNoallocScratchBytes = (size_t) 256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device index

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_address(void** out)
{
    int dev = 0;
    SLQ_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int) g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev])
    {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol,
                              "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return SLQ_OK;
}

}  // namespace

extern "C" {

int slq_estimate(slq_op* op, const double* params, int32_t q, int32_t func,
                 const double* func_params, int32_t gram, double exponent,
                 int32_t orthogonalize,
                 int64_t seed, int32_t m, double tol, int32_t min_s,
                 int32_t max_s, double atol, double rtol, double confidence,
                 double outlier, int32_t block_size, int32_t precision,
                 double* trace, double* error, double* samples,
                 int32_t* processed, int32_t* num_used, int32_t* num_outliers,
                 int32_t* converged, int32_t* all_converged,
                 double* alg_wall_time)
{
    const EstimateArgs a{params, q, func, func_params, gram, exponent, orthogonalize, seed, m,
                         tol, min_s, max_s, atol, rtol, confidence, outlier, block_size,
                         precision, trace, error, samples, processed, num_used, num_outliers,
                         converged, all_converged, alg_wall_time};
    SLQ_TRY(check_estimate("slq_estimate", op, a));
    DeviceGuard guard(op->device);
    const int vdt = value_dtype(op, precision);
    SLQ_TRY(prepare(op, gram, precision));
    Plan p = make_plan(op, gram, m, orthogonalize, q, vdt == SLQ_F64 ? 8 : 4);
    plan_lanes(op, gram, t_lanes(op, gram, q, p), p);
    SLQ_TRY(op->ws.reserve(p.total));
    p.base = (char*) op->ws.base;
    return estimate_loop(op, p, a);
}

int slq_estimate_noalloc(slq_op* op, const double* params, int32_t q, int32_t func,
                         const double* func_params, int32_t gram, double exponent,
                         int32_t orthogonalize, int64_t seed, int32_t m, double tol,
                         int32_t min_s, int32_t max_s, double atol, double rtol,
                         double confidence, double outlier, int32_t block_size,
                         int32_t precision, double* trace, double* error, double* samples,
                         int32_t* processed, int32_t* num_used, int32_t* num_outliers,
                         int32_t* converged, int32_t* all_converged, double* alg_wall_time)
{
    const EstimateArgs a{params, q, func, func_params, gram, exponent, orthogonalize, seed, m,
                         tol, min_s, max_s, atol, rtol, confidence, outlier, block_size,
                         precision, trace, error, samples, processed, num_used, num_outliers,
                         converged, all_converged, alg_wall_time};
    SLQ_TRY(check_estimate("slq_estimate_noalloc", op, a));
    if (op->prof.enabled)
    {
        set_error("slq_estimate_noalloc: profiling is enabled on this handle (its events "
                  "are created on demand); disable it first");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    if (!prepared(op, gram, precision))
    {
        set_error("slq_estimate_noalloc: the handle lacks a derived form this estimate "
                  "reads (transpose, FP32 copy or A + tB tiles); run slq_estimate of the "
                  "same form on it once first");
        return SLQ_E_ARG;
    }
    const int vdt = value_dtype(op, precision);
    Plan p = make_plan(op, gram, m, orthogonalize, q, vdt == SLQ_F64 ? 8 : 4);
    if (p.total > kNoallocScratchBytes)
    {
        set_error("slq_estimate_noalloc: this estimate needs %zu bytes of workspace, more "
                  "than the %zu of the fixed scratch array; use slq_estimate",
                  p.total, kNoallocScratchBytes);
        return SLQ_E_ARG;
    }
    void* ptr = nullptr;
    SLQ_TRY(scratch_address(&ptr));
    p.base = (char*) ptr;
    p.pipeline = false;
    return estimate_loop(op, p, a);
}

int slq_estimate_block_krylov(slq_op* op, const double* params, int32_t q, int32_t func,
                              const double* func_params, double exponent, int64_t seed,
                              int32_t m, int32_t min_s, int32_t max_s, double atol,
                              double rtol, double confidence, double outlier,
                              int32_t precision, double* trace, double* error,
                              double* samples, int32_t* processed, int32_t* num_used,
                              int32_t* num_outliers, int32_t* converged,
                              int32_t* all_converged, double* alg_wall_time)
{
    EstimateArgs a{params, q, func, func_params, 0, exponent, -1, seed, m,
                   0.0, min_s, max_s, atol, rtol, confidence, outlier, kBlock,
                   precision, trace, error, samples, processed, num_used, num_outliers,
                   converged, all_converged, alg_wall_time};
    a.krylov = 1;
    SLQ_TRY(check_estimate("slq_estimate_block_krylov", op, a));
    const int64_t n = op->A->csr.nrows;
    if ((int64_t) m * kBlock > n || m > 256)
    {
        set_error("slq_estimate_block_krylov: lanczos_degree x 64 = %lld must not exceed the "
                  "matrix size %lld (and lanczos_degree must be at most 256)",
                  (long long) m * kBlock, (long long) n);
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    const int vdt = value_dtype(op, precision);
    SLQ_TRY(prepare(op, 0, precision));
    Plan p = make_plan(op, 0, m, 0, q, vdt == SLQ_F64 ? 8 : 4);
    // the group's samples, then the block form's own scratch, in place of
    // the recurrence scratch
    p.total = p.off_scratch;
    p.bk_slots = block_krylov_slots(m);
    p.off_bk_samples = p.total;
    p.total += align256(sizeof(double) * (size_t) p.bk_slots * kBlock * q);
    p.off_scratch = p.total;
    p.scratch_bytes = block_krylov_scratch_bytes(op, m, vdt == SLQ_F64 ? 8 : 4);
    p.total += p.scratch_bytes;
    SLQ_TRY(op->ws.reserve(p.total));
    p.base = (char*) op->ws.base;
    return estimate_loop(op, p, a);
}

int slq_rademacher(int64_t seed, int64_t n, int64_t first_sample, int32_t count,
                   int dtype, void* out, uint64_t* words)
{
    if (n < 1 || count < 1 || count > kBlock || first_sample < 0 || !out ||
        (dtype != SLQ_F64 && dtype != SLQ_F32))
    {
        set_error("slq_rademacher: invalid argument (count must be 1..64)");
        return SLQ_E_ARG;
    }
    if (seed < 0) seed = clock_seed();
    const int64_t W = n / 64 + 1;
    const size_t vb = dtype == SLQ_F64 ? 8 : 4;
    uint64_t* dw = nullptr;
    void* dx = nullptr;
    SLQ_CHECK(hipMalloc(&dw, sizeof(uint64_t) * W * kBlock));
    SLQ_CHECK(hipMalloc(&dx, vb * n * kBlock));
    uint64_t base[4];
    thread_state(seed, 0, base);
    int r = rng_words(0, base, W, first_sample, count, dw);
    if (r == SLQ_OK) r = rng_expand(0, dw, W, n, count, dtype, dx);
    std::vector<char> host(vb * n * kBlock);
    if (r == SLQ_OK)
    {
        hipError_t e = hipMemcpy(host.data(), dx, vb * n * kBlock, hipMemcpyDeviceToHost);
        if (e == hipSuccess && words)
            e = hipMemcpy(words, dw, sizeof(uint64_t) * W * count, hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = hip_status(e, "hipMemcpy", __FILE__, __LINE__);
    }
    (void) hipFree(dw);
    (void) hipFree(dx);
    if (r != SLQ_OK) return r;
    for (int p = 0; p < count; ++p)
        for (int64_t i = 0; i < n; ++i)
            std::memcpy((char*) out + vb * ((size_t) p * n + i),
                        host.data() + vb * ((size_t) i * kBlock + p), vb);
    return SLQ_OK;
}

int slq_op_spmm(slq_op* op, double t, int32_t b, int32_t transpose, const void* X,
                void* Y)
{
    if (!op || !X || !Y || b < 1 || b > kBlock)
    {
        set_error("slq_op_spmm: invalid argument (b must be 1..64)");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    if (transpose) SLQ_TRY(prepare(op, 1, SLQ_PREC_NATIVE));
    const DevCSR& c = op->A->csr;
    const int64_t in_len = transpose ? c.nrows : c.ncols;
    const int64_t out_len = transpose ? c.ncols : c.nrows;
    const int vdt = c.dtype;
    const size_t vb = vdt == SLQ_F64 ? 8 : 4;
    void *dx = nullptr, *dy = nullptr;
    SLQ_CHECK(hipMalloc(&dx, vb * in_len * kBlock));
    SLQ_CHECK(hipMalloc(&dy, vb * out_len * kBlock));
    std::vector<char> hx(vb * in_len * kBlock, 0);
    for (int64_t i = 0; i < in_len; ++i)
        std::memcpy(hx.data() + vb * i * kBlock, (const char*) X + vb * i * b, vb * b);
    int r = SLQ_OK;
    hipError_t e = hipMemcpy(dx, hx.data(), hx.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) r = hip_status(e, "hipMemcpy", __FILE__, __LINE__);
    if (r == SLQ_OK) r = run_spmm(op, t, transpose, dx, dy, vdt);
    std::vector<char> hy(vb * out_len * kBlock);
    if (r == SLQ_OK)
    {
        e = hipStreamSynchronize(op->stream);
        if (e == hipSuccess) e = hipMemcpy(hy.data(), dy, hy.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = hip_status(e, "hipMemcpy", __FILE__, __LINE__);
    }
    (void) hipFree(dx);
    (void) hipFree(dy);
    prof_collect(op);
    if (r != SLQ_OK) return r;
    for (int64_t i = 0; i < out_len; ++i)
        std::memcpy((char*) Y + vb * i * b, hy.data() + vb * i * kBlock, vb * b);
    return SLQ_OK;
}

int slq_recurrences(slq_op* op, double t, int32_t gram, int64_t seed,
                    int64_t first_sample, int32_t count, int32_t m, double tol,
                    int32_t ortho, int32_t precision, double* alpha, double* beta,
                    int32_t* sizes)
{
    if (!op || count < 1 || count > kBlock || m < 1 || !alpha || !beta || !sizes ||
        first_sample < 0)
    {
        set_error("slq_recurrences: invalid argument (count must be 1..64)");
        return SLQ_E_ARG;
    }
    if (!gram && op->A->csr.nrows != op->A->csr.ncols)
    {
        set_error("slq_recurrences: Lanczos needs a square operator");
        return SLQ_E_ARG;
    }
    DeviceGuard guard(op->device);
    if (seed < 0) seed = clock_seed();
    const int vdt = value_dtype(op, precision);
    SLQ_TRY(prepare(op, gram, precision));
    Plan p = make_plan(op, gram, m, ortho, 1, vdt == SLQ_F64 ? 8 : 4);
    SLQ_TRY(op->ws.reserve(p.total));
    p.base = (char*) op->ws.base;
    uint64_t base[4];
    thread_state(seed, 0, base);
    SLQ_TRY(make_probes(op, p, base, first_sample, count, vdt));
    SLQ_TRY(recurrence_block(op, p, t, gram, tol, ortho, precision));
    std::vector<double> ha((size_t) m * kBlock), hb((size_t) m * kBlock);
    std::vector<int32_t> hs(kBlock);
    SLQ_CHECK(hipMemcpyAsync(ha.data(), at<double>(p, p.off_alpha), sizeof(double) * m * kBlock,
                             hipMemcpyDeviceToHost, op->stream));
    SLQ_CHECK(hipMemcpyAsync(hb.data(), at<double>(p, p.off_beta), sizeof(double) * m * kBlock,
                             hipMemcpyDeviceToHost, op->stream));
    SLQ_CHECK(hipMemcpyAsync(hs.data(), at<int32_t>(p, p.off_sizes), sizeof(int32_t) * kBlock,
                             hipMemcpyDeviceToHost, op->stream));
    SLQ_CHECK(hipStreamSynchronize(op->stream));
    prof_collect(op);
    for (int c = 0; c < count; ++c)
    {
        sizes[c] = hs[c] > 0 ? hs[c] : m;
        for (int j = 0; j < m; ++j)
        {
            alpha[(size_t) c * m + j] = ha[(size_t) j * kBlock + c];
            beta[(size_t) c * m + j] = hb[(size_t) j * kBlock + c];
        }
    }
    return SLQ_OK;
}

int slq_quadrature(const double* alpha, const double* beta, const int32_t* sizes,
                   int32_t count, int32_t m, int32_t gram, int32_t func,
                   const double* func_params, double exponent, double n,
                   const double* params, int32_t q, int32_t shift, double* samples,
                   double* theta, double* tau)
{
    if (!alpha || !beta || !sizes || count < 1 || count > kBlock || m < 1 || q < 1 ||
        !samples || (shift && !params) || !func_ok(func, func_params))
    {
        set_error("slq_quadrature: invalid argument (count must be 1..64)");
        return SLQ_E_ARG;
    }
    std::vector<double> ha((size_t) m * kBlock, 0.0), hb((size_t) m * kBlock, 0.0);
    std::vector<int32_t> hs(kBlock, 1);
    for (int c = 0; c < count; ++c)
    {
        hs[c] = sizes[c];
        for (int j = 0; j < m; ++j)
        {
            ha[(size_t) j * kBlock + c] = alpha[(size_t) c * m + j];
            hb[(size_t) j * kBlock + c] = beta[(size_t) c * m + j];
        }
    }
    const int qs = shift ? q : 1;
    double *da = nullptr, *db = nullptr, *dsmp = nullptr, *dp = nullptr, *dth = nullptr,
           *dta = nullptr;
    int32_t* ds = nullptr;
    SLQ_CHECK(hipMalloc(&da, sizeof(double) * ha.size()));
    SLQ_CHECK(hipMalloc(&db, sizeof(double) * hb.size()));
    SLQ_CHECK(hipMalloc(&ds, sizeof(int32_t) * kBlock));
    SLQ_CHECK(hipMalloc(&dsmp, sizeof(double) * kBlock * qs));
    SLQ_CHECK(hipMalloc(&dp, sizeof(double) * q));
    SLQ_CHECK(hipMalloc(&dth, sizeof(double) * kBlock * m));
    SLQ_CHECK(hipMalloc(&dta, sizeof(double) * kBlock * m));
    SLQ_CHECK(hipMemcpy(da, ha.data(), sizeof(double) * ha.size(), hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(db, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice));
    SLQ_CHECK(hipMemcpy(ds, hs.data(), sizeof(int32_t) * kBlock, hipMemcpyHostToDevice));
    if (params) SLQ_CHECK(hipMemcpy(dp, params, sizeof(double) * q, hipMemcpyHostToDevice));
    int r = run_quadrature(0, da, db, ds, count, m, gram, func, func_params, exponent, n,
                           params ? dp : nullptr, q, shift, 0, qs, dsmp, dth, dta);
    if (r == SLQ_OK)
    {
        hipError_t e = hipDeviceSynchronize();
        if (e == hipSuccess)
            e = hipMemcpy(samples, dsmp, sizeof(double) * count * qs, hipMemcpyDeviceToHost);
        if (e == hipSuccess && theta)
            e = hipMemcpy(theta, dth, sizeof(double) * count * m, hipMemcpyDeviceToHost);
        if (e == hipSuccess && tau)
            e = hipMemcpy(tau, dta, sizeof(double) * count * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) r = hip_status(e, "hipMemcpy", __FILE__, __LINE__);
    }
    (void) hipFree(da);
    (void) hipFree(db);
    (void) hipFree(ds);
    (void) hipFree(dsmp);
    (void) hipFree(dp);
    (void) hipFree(dth);
    (void) hipFree(dta);
    return r;
}

int slq_profile_enable(slq_op* op, int32_t enable)
{
    if (!op) return SLQ_E_ARG;
    prof_collect(op);
    op->prof.enabled = enable != 0;
    op->prof.spmm_ms = 0.0;
    op->prof.spmm_launches = 0;
    op->prof.update_ms = 0.0;
    op->prof.update_launches = 0;
    op->prof.steps = 0;
    op->prof.gemm_ms = 0.0;
    op->prof.gemm_launches = 0;
    op->prof.gemm_flops = 0.0;
    return SLQ_OK;
}

int slq_op_set_launch(slq_op* op, int32_t spmm_grid, int32_t stream_grid,
                      int32_t spmm_split, int32_t tile_mode)
{
    if (!op || spmm_grid < 0 || spmm_grid > kMaxGrid || stream_grid < 0 ||
        stream_grid > kMaxGrid || !(spmm_split == 0 || spmm_split == 1 || spmm_split == 2 ||
                                    spmm_split == 4) ||
        !(tile_mode == 0 || tile_mode == 1))
    {
        set_error("slq_op_set_launch: grid caps must be 0..%d, the split 0, 1, 2 or 4 and "
                  "the tile mode 0 or 1", kMaxGrid);
        return SLQ_E_ARG;
    }
    op->spmm_grid = spmm_grid;
    op->stream_grid = stream_grid;
    op->spmm_split = spmm_split;
    op->tile_mode = tile_mode;
    return SLQ_OK;
}

int slq_op_get_launch(slq_op* op, int32_t* spmm_grid, int32_t* stream_grid,
                      int32_t* spmm_split, int32_t* tile_mode)
{
    if (!op) return SLQ_E_ARG;
    if (spmm_grid) *spmm_grid = op->spmm_grid;
    if (stream_grid) *stream_grid = op->stream_grid;
    if (spmm_split) *spmm_split = op->spmm_split;
    if (tile_mode) *tile_mode = op->tile_mode;
    return SLQ_OK;
}

int slq_profile_steps(slq_op* op, int64_t* steps)
{
    if (!op || !steps) return SLQ_E_ARG;
    *steps = op->prof.steps;
    return SLQ_OK;
}

int slq_profile_read(slq_op* op, double* spmm_ms, int64_t* spmm_launches,
                     double* update_ms, int64_t* update_launches)
{
    if (!op) return SLQ_E_ARG;
    prof_collect(op);
    if (spmm_ms) *spmm_ms = op->prof.spmm_ms;
    if (spmm_launches) *spmm_launches = op->prof.spmm_launches;
    if (update_ms) *update_ms = op->prof.update_ms;
    if (update_launches) *update_launches = op->prof.update_launches;
    return SLQ_OK;
}

int slq_profile_gemm(slq_op* op, double* ms, int64_t* launches, double* flops)
{
    if (!op) return SLQ_E_ARG;
    prof_collect(op);
    if (ms) *ms = op->prof.gemm_ms;
    if (launches) *launches = op->prof.gemm_launches;
    if (flops) *flops = op->prof.gemm_flops;
    return SLQ_OK;
}

int slq_copy_bandwidth(int64_t bytes, int32_t iters, double* ms_per_launch)
{
    if (bytes < 16 || iters < 1 || !ms_per_launch)
    {
        set_error("slq_copy_bandwidth: invalid argument");
        return SLQ_E_ARG;
    }
    return copy_bandwidth(bytes, iters, ms_per_launch);
}

}  // extern "C"
