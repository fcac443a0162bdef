// Rademacher probe generation on the device, bit-exact to imate's per-thread
// xoshiro256** stream.
//
// Stream definition (reference): thread 0's generator is xoshiro256** seeded
// by four SplitMix64 draws from state (seed << 32) | seed and advanced by one
// jump() (imate/_random_generator/split_mix_64.cpp:28-64,
// random_number_generator.cpp:126-136). A probe of length n consumes
// W = floor(n/64) + 1 consecutive words; bit j of word i is entry 64 i + j,
// 1 -> +1 and 0 -> -1 (random_array_generator.cpp:98-164).
//
// Parallel generation: the state transition of xoshiro256** is linear over
// GF(2), so the state k steps ahead is M^k s. The host builds the 256 x 256
// bit matrices M^(2^e), e = 0..63, once; each wavefront jumps to its chunk
// group's offset by multiplying the base state with the matrices of the set
// bits of the offset (each product shared by the 64 lanes), hands each lane
// the state of its chunk, and every lane steps its chunk of words
// sequentially. A second kernel expands the packed words into the [len][64]
// probe block.

#include <mutex>
#include <vector>

#include "slq_internal.hpp"

namespace slq {

namespace {

constexpr int kChunkWords = 64;   // words generated per thread after the jump

inline uint64_t rotl_h(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// State transition of one xoshiro256** step (the output is not needed).
inline void advance_h(uint64_t s[4])
{
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl_h(s[3], 45);
}

inline uint64_t next_h(uint64_t s[4])
{
    const uint64_t result = rotl_h(s[1] * 5, 7) * 9;
    advance_h(s);
    return result;
}

// Column-major bit matrix: col[j] = image of the unit state e_j.
struct BitMatrix {
    uint64_t col[256][4];
};

void apply_h(const BitMatrix& M, const uint64_t x[4], uint64_t y[4])
{
    uint64_t r[4] = {0, 0, 0, 0};
    for (int j = 0; j < 256; ++j)
    {
        if ((x[j >> 6] >> (j & 63)) & 1)
        {
            r[0] ^= M.col[j][0];
            r[1] ^= M.col[j][1];
            r[2] ^= M.col[j][2];
            r[3] ^= M.col[j][3];
        }
    }
    y[0] = r[0]; y[1] = r[1]; y[2] = r[2]; y[3] = r[3];
}

std::vector<BitMatrix>& jump_tables()
{
    static std::vector<BitMatrix> tables;
    static std::once_flag once;
    std::call_once(once, [] {
        tables.resize(64);
        for (int j = 0; j < 256; ++j)
        {
            uint64_t s[4] = {0, 0, 0, 0};
            s[j >> 6] = (uint64_t) 1 << (j & 63);
            advance_h(s);
            for (int w = 0; w < 4; ++w) tables[0].col[j][w] = s[w];
        }
        for (int e = 1; e < 64; ++e)
            for (int j = 0; j < 256; ++j)
                apply_h(tables[e - 1], tables[e - 1].col[j], tables[e].col[j]);
    });
    return tables;
}

}  // namespace
}  // namespace slq

// Device copy of the 64 jump matrices (512 KiB), one per device in a static
// array, so that probe generation never allocates.
static __device__ __align__(256) uint64_t jump_tab[64 * 256 * 4];

namespace slq {
namespace {

// Device address of the tables on the current device, filled from the host
// tables on first use there (looked up with that device current).
uint64_t* device_tables()
{
    static std::mutex mu;
    static std::vector<uint64_t*> per_device;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int) per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev])
    {
        auto& t = jump_tables();
        static_assert(sizeof(BitMatrix) * 64 == sizeof(jump_tab), "jump table size");
        void* d = nullptr;
        if (hipGetSymbolAddress(&d, HIP_SYMBOL(jump_tab)) != hipSuccess || !d) return nullptr;
        if (hipMemcpy(d, t.data(), sizeof(BitMatrix) * t.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return nullptr;
        per_device[dev] = (uint64_t*) d;
    }
    return per_device[dev];
}

__device__ __forceinline__ uint64_t rotl_d(uint64_t x, int k)
{
    return (x << k) | (x >> (64 - k));
}

// y = M x over GF(2) by the 64 lanes of a wavefront together (x uniform
// across the wavefront): lane l takes columns 4 l .. 4 l + 3 of M (column
// major [256][4] in global memory), XORs those selected by its four bits of
// x, and a butterfly of XOR shuffles gives every lane the full product.
__device__ __forceinline__ void apply_wave(const uint64_t* __restrict__ M, const uint64_t (&x)[4],
                                           uint64_t (&y)[4], int lane)
{
    const uint64_t xw = lane < 16 ? x[0] : (lane < 32 ? x[1] : (lane < 48 ? x[2] : x[3]));
    const uint32_t bits = (uint32_t) (xw >> (4 * (lane & 15))) & 0xFu;
    const uint64_t* __restrict__ c = M + (size_t) lane * 16;
    uint64_t r[4] = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const uint64_t mask = 0 - (uint64_t) ((bits >> i) & 1u);
#pragma unroll
        for (int w = 0; w < 4; ++w) r[w] ^= c[4 * i + w] & mask;
    }
#pragma unroll
    for (int sh = 1; sh < kWave; sh <<= 1)
    {
#pragma unroll
        for (int w = 0; w < 4; ++w) r[w] ^= __shfl_xor(r[w], sh);
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) y[w] = r[w];
}

static_assert(kChunkWords == 64, "lane states are spaced by jump tables 6 + k");

// One wavefront per (probe p, group g of 64 chunks): lane l generates words
// [c L, min((c + 1) L, W)) of probe p, c = 64 g + l, i.e. stream words from
// offset (first_sample + p) W + c L. The wavefront jumps the base state to
// the group's first chunk with the matrices of the set bits of that offset,
// then spreads it to its lanes by doubling (lanes s + 2^k take M^(2^k L) of
// lane s, k = 0..5), every product computed by the whole wavefront.
__global__ __launch_bounds__(256) void rng_words_kernel(
    const uint64_t* __restrict__ tables, uint64_t s0, uint64_t s1, uint64_t s2,
    uint64_t s3, int64_t W, int64_t first_sample, int32_t count, int64_t chunks,
    int64_t groups, uint64_t* __restrict__ words)
{
    const int lane = threadIdx.x % kWave;
    const int64_t wv = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t p = wv / groups;
    const int64_t g = wv - p * groups;
    if (p >= count) return;
    const uint64_t offset = (uint64_t) (first_sample + p) * (uint64_t) W +
                            (uint64_t) (g * kWave * kChunkWords);
    uint64_t x[4] = {s0, s1, s2, s3};
    for (int e = 0; e < 64; ++e)
        if ((offset >> e) & 1) apply_wave(tables + (size_t) e * 256 * 4, x, x, lane);
    uint64_t st[4] = {x[0], x[1], x[2], x[3]};
    for (int k = 0; k < 6; ++k)
    {
        const int half = 1 << k;
        for (int src = 0; src < half; ++src)
        {
            uint64_t v[4], r[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) v[w] = __shfl(st[w], src);
            apply_wave(tables + (size_t) (6 + k) * 256 * 4, v, r, lane);
            if (lane == src + half)
            {
#pragma unroll
                for (int w = 0; w < 4; ++w) st[w] = r[w];
            }
        }
    }
    const int64_t c = g * kWave + lane;
    if (c >= chunks) return;
    uint64_t x0 = st[0], x1 = st[1], x2 = st[2], x3 = st[3];
    uint64_t* out = words + p * W;
    const int64_t beg = c * kChunkWords;
    const int64_t end = (beg + kChunkWords < W) ? beg + kChunkWords : W;
    for (int64_t i = beg; i < end; ++i)
    {
        out[i] = rotl_d(x1 * 5, 7) * 9;
        const uint64_t t = x1 << 17;
        x2 ^= x0;
        x3 ^= x1;
        x1 ^= x2;
        x0 ^= x3;
        x2 ^= t;
        x3 = rotl_d(x3, 45);
    }
}

// X[i][p] = +1 / -1 from bit (i mod 64) of word floor(i/64) of probe p; zero
// for p >= count. A thread writes two probes of a row (one access).
template <typename T>
__global__ __launch_bounds__(256) void rng_expand_kernel(
    const uint64_t* __restrict__ words, int64_t W, int64_t len, int32_t count,
    T* __restrict__ X)
{
    typedef T T2 __attribute__((ext_vector_type(2)));
    const int64_t gid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = gid / (kBlock / 2);
    const int p = 2 * (int) (gid - i * (kBlock / 2));
    if (i >= len) return;
    T2 v = {T(0), T(0)};
    if (p < count)
    {
        const uint64_t w = words[(int64_t) p * W + (i >> 6)];
        v.x = ((w >> (i & 63)) & 1) ? T(1) : T(-1);
    }
    if (p + 1 < count)
    {
        const uint64_t w = words[(int64_t) (p + 1) * W + (i >> 6)];
        v.y = ((w >> (i & 63)) & 1) ? T(1) : T(-1);
    }
    *reinterpret_cast<T2*>(X + i * kBlock + p) = v;
}

}  // namespace

void thread_state(int64_t seed, int thread_id, uint64_t out[4])
{
    uint64_t sd = (uint64_t) seed;
    uint64_t st = (sd << 32) | sd;
    for (int i = 0; i < 4; ++i)
    {
        uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        out[i] = z ^ (z >> 31);
    }
    // jump(): 2^128 steps, xoshiro_256_star_star.cpp:100-133 (published
    // jump polynomial of xoshiro256**).
    static const uint64_t J[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                                  0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
    for (int rep = 0; rep < thread_id + 1; ++rep)
    {
        uint64_t a[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int b = 0; b < 64; ++b)
            {
                if (J[i] & ((uint64_t) 1 << b))
                    for (int w = 0; w < 4; ++w) a[w] ^= out[w];
                next_h(out);
            }
        for (int w = 0; w < 4; ++w) out[w] = a[w];
    }
}

void stream_at(const uint64_t base[4], uint64_t k, uint64_t out[4])
{
    const auto& t = jump_tables();
    uint64_t s[4] = {base[0], base[1], base[2], base[3]};
    for (int e = 0; e < 64; ++e)
        if ((k >> e) & 1)
        {
            uint64_t y[4];
            apply_h(t[e], s, y);
            for (int w = 0; w < 4; ++w) s[w] = y[w];
        }
    for (int w = 0; w < 4; ++w) out[w] = s[w];
}

uint64_t stream_next(uint64_t s[4]) { return next_h(s); }

int rng_words(hipStream_t s, const uint64_t base[4], int64_t W,
              int64_t first_sample, int32_t count, uint64_t* words)
{
    uint64_t* tables = device_tables();
    if (!tables)
    {
        set_error("could not upload the xoshiro256** jump tables");
        return SLQ_E_HIP;
    }
    if (count <= 0) return SLQ_OK;
    const int64_t chunks = (W + kChunkWords - 1) / kChunkWords;
    const int64_t groups = (chunks + kWave - 1) / kWave;
    const int64_t threads = groups * count * kWave;
    const int64_t grid = (threads + 255) / 256;
    hipLaunchKernelGGL(rng_words_kernel, dim3((unsigned) grid), dim3(256), 0, s,
                       tables, base[0], base[1], base[2], base[3], W,
                       first_sample, count, chunks, groups, words);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

int rng_expand(hipStream_t s, const uint64_t* words, int64_t W, int64_t len,
               int32_t count, int dtype, void* X)
{
    const int64_t threads = len * (kBlock / 2);
    const int64_t grid = (threads + 255) / 256;
    if (dtype == SLQ_F64)
        hipLaunchKernelGGL(rng_expand_kernel<double>, dim3((unsigned) grid),
                           dim3(256), 0, s, words, W, len, count, (double*) X);
    else
        hipLaunchKernelGGL(rng_expand_kernel<float>, dim3((unsigned) grid),
                           dim3(256), 0, s, words, W, len, count, (float*) X);
    SLQ_CHECK(hipGetLastError());
    return SLQ_OK;
}

}  // namespace slq
