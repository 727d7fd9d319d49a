// sort.hip - device-wide primitives for 64-wide CDNA4 wavefronts:
//  * exclusive prefix scan of uint32 arrays (reduce -> scan of block sums ->
//    downsweep), with wave-level inclusive scans inside each block;
//  * stable LSD radix sort of (key, value) uint32 pairs, 8 bits per pass:
//    per-tile digit histograms, a scan of the digit-major histogram matrix,
//    and a scatter that ranks equal digits inside a wave with ballots, so the
//    order of equal keys is the input order.
#include "kernels.hpp"

namespace wsrt {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanRound = kScanThreads * 4;          // 4 consecutive elements per thread
constexpr int kScanRounds = 16;
constexpr int kScanTile = kScanRound * kScanRounds;  // 16384 elements per block

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t n = __shfl_up(v, off, 64);
        if (lane >= off) v += n;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_reduce(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total,
                                                         uint32_t* smem /* >= 16 */) {
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) smem[wave] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t w = threadIdx.x < (unsigned)nwaves ? smem[threadIdx.x] : 0u;
        uint32_t wi = wave_inclusive_scan(w);
        if (threadIdx.x < (unsigned)nwaves) smem[16 + threadIdx.x] = wi - w;
        if (threadIdx.x == (unsigned)nwaves - 1) smem[32] = wi;
    }
    __syncthreads();
    uint32_t res = smem[16 + wave] + inc - v;
    total = smem[32];
    __syncthreads();
    return res;
}

// Four consecutive elements from i (zero past n). kVec: one 16-byte load
// (n is a multiple of 4 and the array 16-byte aligned).
template <bool kVec>
__device__ __forceinline__ uint4 load4(const uint32_t* __restrict__ in, uint64_t i, uint32_t n) {
    if (kVec) return i < n ? *reinterpret_cast<const uint4*>(in + i) : make_uint4(0u, 0u, 0u, 0u);
    uint4 v;
    v.x = i < n ? in[i] : 0u;
    v.y = i + 1 < n ? in[i + 1] : 0u;
    v.z = i + 2 < n ? in[i + 2] : 0u;
    v.w = i + 3 < n ? in[i + 3] : 0u;
    return v;
}
template <bool kVec>
__device__ __forceinline__ void store4(uint32_t* __restrict__ out, uint64_t i, uint32_t n, uint4 v) {
    if (kVec) {
        if (i < n) *reinterpret_cast<uint4*>(out + i) = v;
        return;
    }
    if (i < n) out[i] = v.x;
    if (i + 1 < n) out[i + 1] = v.y;
    if (i + 2 < n) out[i + 2] = v.z;
    if (i + 3 < n) out[i + 3] = v.w;
}

// Tile sums. A tile is kScanRounds rounds of 1024 elements; in a round,
// thread t holds elements 4t .. 4t+3, so every load is one contiguous 4 KiB.
template <bool kVec>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                              uint32_t* __restrict__ block_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + 4u * threadIdx.x;
    uint32_t s = 0;
#pragma unroll 4
    for (int r = 0; r < kScanRounds; ++r) {
        const uint4 v = load4<kVec>(in, base + (uint64_t)r * kScanRound, n);
        s += (v.x + v.y) + (v.z + v.w);
    }
    s = wave_reduce(s);
    __shared__ uint32_t part[kScanThreads / 64];
    if (__lane_id() == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += part[w];
        block_sums[blockIdx.x] = t;
    }
}

// Single-block exclusive scan of up to 1024*16 block sums, in place.
__global__ __launch_bounds__(1024) void k_scan_block_sums(uint32_t* __restrict__ sums, uint32_t n,
                                                          uint32_t* __restrict__ total_out) {
    __shared__ uint32_t smem[64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < n ? sums[i] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, total, smem);
        uint32_t c = carry;
        if (i < n) sums[i] = ex + c;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Exclusive scan of a tile from its offset in block_sums, round by round.
template <bool kVec>
__global__ __launch_bounds__(kScanThreads) void k_scan_downsweep(const uint32_t* __restrict__ in,
                                                                 uint32_t* __restrict__ out, uint32_t n,
                                                                 const uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t smem[64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + 4u * threadIdx.x;
    uint32_t carry = block_sums[blockIdx.x];
    for (int r = 0; r < kScanRounds; ++r) {
        const uint64_t i = base + (uint64_t)r * kScanRound;
        if (i - 4u * threadIdx.x >= n) break;  // block-uniform: the round starts past the end
        const uint4 v = load4<kVec>(in, i, n);
        uint32_t total;
        const uint32_t ex = block_exclusive_scan((v.x + v.y) + (v.z + v.w), total, smem) + carry;
        uint4 o;
        o.x = ex;
        o.y = o.x + v.x;
        o.z = o.y + v.y;
        o.w = o.z + v.z;
        store4<kVec>(out, i, n, o);
        carry += total;
    }
}

// ---- radix sort ---------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 keys per tile
constexpr int kRadix = 256;

// n_dev (optional): the element count in device memory (at most n), so a
// sort whose size is produced on the device can be launched for its maximum.
// The kernels loop over the tiles that hold elements (a fixed grid); the
// histogram entries of the others stay zero from one fill per sort.
__global__ __launch_bounds__(kSortThreads) void k_radix_histogram(const uint32_t* __restrict__ keys,
                                                                  uint32_t n,
                                                                  const uint32_t* __restrict__ n_dev,
                                                                  int shift, uint32_t num_tiles,
                                                                  uint32_t* __restrict__ hist) {
    if (n_dev) n = min(n, *n_dev);
    const uint32_t used = (uint32_t)(((uint64_t)n + kSortTile - 1) / kSortTile);
    __shared__ uint32_t h[kRadix];
    for (uint32_t tile = blockIdx.x; tile < used; tile += gridDim.x) {
        const uint64_t base = (uint64_t)tile * kSortTile;
        h[threadIdx.x] = 0;
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < kSortRounds; ++r) {
            uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
            if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
        }
        __syncthreads();
        // digit-major layout: hist[digit * num_tiles + tile]
        hist[(uint64_t)threadIdx.x * num_tiles + tile] = h[threadIdx.x];
        __syncthreads();
    }
}

// Stable scatter of one tile. Keys are taken in input order 256 at a time;
// equal digits inside a wave are ranked with 8 ballots (match-any), waves of
// a round are ordered by per-digit counts in LDS, rounds by a running count.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, uint32_t n,
    const uint32_t* __restrict__ n_dev, int shift, uint32_t num_tiles,
    const uint32_t* __restrict__ hist_scanned) {
    if (n_dev) n = min(n, *n_dev);
    const uint32_t used = (uint32_t)(((uint64_t)n + kSortTile - 1) / kSortTile);
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wave_cnt[kWaves][kRadix];
    const int wave = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    for (uint32_t tile = blockIdx.x; tile < used; tile += gridDim.x) {
        run[threadIdx.x] = hist_scanned[(uint64_t)threadIdx.x * num_tiles + tile];
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wave_cnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t base = (uint64_t)tile * kSortTile;
        for (int r = 0; r < kSortRounds; ++r) {
            uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
            bool valid = i < n;
            uint32_t key = valid ? keys_in[i] : 0u;
            uint32_t val = valid ? vals_in[i] : 0u;
            uint32_t dig = (key >> shift) & 0xFFu;
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                uint64_t m = __ballot((dig >> b) & 1u);
                peers &= ((dig >> b) & 1u) ? m : ~m;
            }
            uint32_t rank = (uint32_t)__popcll(peers & lt);
            bool leader = valid && (peers & lt) == 0;
            if (leader) wave_cnt[wave][dig] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (valid) {
                uint32_t pos = run[dig] + rank;
                for (int w = 0; w < wave; ++w) pos += wave_cnt[w][dig];
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
            __syncthreads();
            {
                uint32_t add = 0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w) {
                    add += wave_cnt[w][threadIdx.x];
                    wave_cnt[w][threadIdx.x] = 0;
                }
                run[threadIdx.x] += add;
            }
            __syncthreads();
        }
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint4 v4 = make_uint4(v, v, v, v);
    uint64_t n4 = n / 4;
    uint4* p4 = reinterpret_cast<uint4*>(p);
    for (uint64_t k = i; k < n4; k += stride) p4[k] = v4;
    for (uint64_t k = n4 * 4 + i; k < n; k += stride) p[k] = v;
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

}  // namespace

int device_cus() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

size_t scan_temp_words(uint32_t n) {
    uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    return (size_t)blocks + 1;
}

int scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* temp,
                   uint32_t* total_out, hipStream_t s) {
    if (n == 0) return 0;
    uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks > 1024u * 64u) return -1;
    const bool vec = (n % 4u) == 0 && (reinterpret_cast<uintptr_t>(in) & 15u) == 0 &&
                     (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
    if (vec) hipLaunchKernelGGL(k_scan_reduce<true>, dim3(blocks), dim3(kScanThreads), 0, s, in, n, temp);
    else hipLaunchKernelGGL(k_scan_reduce<false>, dim3(blocks), dim3(kScanThreads), 0, s, in, n, temp);
    hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(1024), 0, s, temp, blocks, total_out);
    if (vec) hipLaunchKernelGGL(k_scan_downsweep<true>, dim3(blocks), dim3(kScanThreads), 0, s, in, out, n, temp);
    else hipLaunchKernelGGL(k_scan_downsweep<false>, dim3(blocks), dim3(kScanThreads), 0, s, in, out, n, temp);
    return 0;
}

size_t radix_temp_words(uint32_t n) {
    uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    size_t hist = (size_t)tiles * kRadix;
    return 2 * (size_t)n + 2 * hist + scan_temp_words((uint32_t)hist);
}

int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int bits, uint32_t* temp,
                     hipStream_t s, bool* result_in_temp, const uint32_t* n_dev) {
    // temp layout: keys_alt[n], vals_alt[n], hist[tiles*256], hist_scan[tiles*256], scan temp
    *result_in_temp = false;
    if (n <= 1) return 0;
    uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t hist_n = tiles * kRadix;
    uint32_t* kalt = temp;
    uint32_t* valt = temp + n;
    uint32_t* hist = temp + 2 * (size_t)n;
    uint32_t* hscan = hist + hist_n;
    uint32_t* stemp = hscan + hist_n;
    uint32_t *ki = keys, *vi = vals, *ko = kalt, *vo = valt;
    int passes = (bits + 7) / 8;
    const uint32_t grid = tiles < 2048u ? tiles : 2048u;
    if (n_dev) fill_u32(hist, hist_n, 0u, s);  // entries of tiles past the device count
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_histogram, dim3(grid), dim3(kSortThreads), 0, s, ki, n, n_dev, shift,
                           tiles, hist);
        if (scan_exclusive(hist, hscan, hist_n, stemp, nullptr, s)) return -1;
        hipLaunchKernelGGL(k_radix_scatter, dim3(grid), dim3(kSortThreads), 0, s, ki, vi, ko, vo,
                           n, n_dev, shift, tiles, hscan);
        uint32_t* t;
        t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    *result_in_temp = (ki != keys);
    return 0;
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    uint64_t blocks = (n / 4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
}

void iota_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

}  // namespace wsrt
