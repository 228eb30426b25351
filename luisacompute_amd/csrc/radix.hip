// radix.hip -- the library's own key sort (LSD radix, 8-bit digits) and exclusive scan (gfx950).
//
// Sort: a stable sort of (uint32 key, uint32 value) pairs by bits [begin_bit, end_bit) of the key,
// one pass per 8-bit digit, each pass three kernels over tiles of 4,096 items (256 threads x 16):
//   k_rs_upsweep   digit counts of each tile -> counts[digit][tile]
//   k_rs_scan      one workgroup per digit: exclusive prefix of its row over the tiles, row total
//   k_rs_downsweep ranks each tile's items stably (per-wave digit counters in LDS; lanes of a wave
//                  holding the same digit find each other with 8 ballots), reorders the tile in
//                  LDS by digit, and writes each digit's run contiguously to its global offset
//                  (exclusive prefix of the digit totals + the tile's prefix in the digit's row).
// Every kernel reads the item count from `count_ptr` when one is given (items at or beyond it are
// neither read nor written), and writes every word it later reads, so a sort needs no memset and
// can be recorded into a graph with the grid sized for the capacity. Values may be the identity
// (`vals_in` null): the first pass then writes each item's input index.
// Traffic per pass and item: 4 B read (upsweep) + 8 B read and 8 B written (downsweep).
//
// Scan: exclusive prefix sum of up to 2^32-1 uint32 words by one workgroup of 1,024 threads
// (16,384 words per round); the lists it serves (per-workgroup counts of the build phases) are
// tens of thousands of words.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lc_hip_internal.h"

namespace lchip {

namespace {

constexpr int kRsBlock = 256;
constexpr int kRsItems = 16;
constexpr uint32_t kRsTile = kRsBlock * kRsItems;
constexpr int kRsWaves = kRsBlock / 64;
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 16;

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

__device__ __forceinline__ uint32_t item_count(const uint32_t *count_ptr, uint32_t cap) {
    return count_ptr ? min(*count_ptr, cap) : cap;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Lanes of `active` whose 8-bit digit equals this lane's.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Exclusive prefix of `v` over the workgroup (BLOCK threads); `total` gets the sum. s_wave holds
// BLOCK / 64 words; the call synchronises the workgroup twice.
template<int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t &total, uint32_t *s_wave) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t incl = wave_inclusive_scan(v, lane);
    if (lane == 63u) s_wave[w] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)(BLOCK / 64); k++) {
        const uint32_t s = s_wave[k];
        before += k < w ? s : 0u;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return before + incl - v;
}

// Tile layout of the ranking: wave w of tile t holds items t*kRsTile + w*64*kRsItems + j*64 + lane,
// j < kRsItems, so the order (wave, j, lane) is the input order.
__global__ __launch_bounds__(kRsBlock) void k_rs_upsweep(const uint32_t *keys, uint32_t cap, const uint32_t *count_ptr,
                                                         uint32_t shift, uint32_t mask, uint32_t n_tiles,
                                                         uint32_t *counts) {
    __shared__ uint32_t s_hist[kRsWaves][256];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    const uint32_t n = item_count(count_ptr, cap);
#pragma unroll
    for (int k = 0; k < kRsWaves; k++) s_hist[k][t] = 0;
    __syncthreads();
    const uint32_t tile0 = blockIdx.x * kRsTile;
    if (tile0 < n) {
        // counts do not depend on the order: 16-B loads, thread t holding words 4 (t + 256 k) + e
        uint32_t d[kRsItems];
#pragma unroll
        for (int k = 0; k < kRsItems / 4; k++) {
            const uint32_t i = tile0 + 4u * (t + kRsBlock * k);
            uint32_t v[4];
            if (i + 3u < n) {
                const uint4 x = *reinterpret_cast<const uint4 *>(keys + i);
                v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = i + e < n ? keys[i + e] : 0u;
            }
#pragma unroll
            for (int e = 0; e < 4; e++) d[4 * k + e] = i + e < n ? (v[e] >> shift) & mask : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int j = 0; j < kRsItems; j++) {
            const bool valid = d[j] != 0xFFFFFFFFu;
            if (__ballot(valid) == 0) continue; // past the count
            const uint32_t first = __builtin_amdgcn_readfirstlane(d[j]);
            if (first != 0xFFFFFFFFu && __ballot(d[j] == first) == ~0ull) { // whole wave valid, one digit
                if (lane == 0) s_hist[w][first] += 64u;
            } else if (valid) {
                atomicAdd(&s_hist[w][d[j]], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kRsWaves; k++) c += s_hist[k][t];
    counts[(size_t)t * n_tiles + blockIdx.x] = c;
}

// Row `blockIdx.x` of counts (one digit over all tiles) -> exclusive prefix in place; row total.
__global__ __launch_bounds__(kRsBlock) void k_rs_scan(uint32_t *counts, uint32_t n_tiles, uint32_t *totals) {
    __shared__ uint32_t s_wave[kRsWaves];
    uint32_t *row = counts + (size_t)blockIdx.x * n_tiles;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n_tiles; c0 += kRsTile) {
        const uint32_t b = c0 + threadIdx.x * kRsItems;
        uint32_t v[kRsItems], sum = 0;
#pragma unroll
        for (int j = 0; j < kRsItems; j++) {
            v[j] = b + j < n_tiles ? row[b + j] : 0u;
            sum += v[j];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan<kRsBlock>(sum, total, s_wave);
#pragma unroll
        for (int j = 0; j < kRsItems; j++) {
            if (b + j < n_tiles) row[b + j] = run;
            run += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template<bool IDENTITY>
__global__ __launch_bounds__(kRsBlock) void k_rs_downsweep(const uint32_t *keys_in, const uint32_t *vals_in,
                                                           uint32_t *keys_out, uint32_t *vals_out, uint32_t cap,
                                                           const uint32_t *count_ptr, uint32_t shift, uint32_t mask,
                                                           uint32_t n_tiles, const uint32_t *counts,
                                                           const uint32_t *totals) {
    __shared__ uint32_t s_hist[kRsWaves][256]; // per-wave digit counters, then offsets in the tile
    __shared__ uint32_t s_gbase[256];          // global offset of the digit's run - its offset in the tile
    __shared__ uint32_t s_buf[kRsTile];        // the tile ordered by digit (keys, then values)
    __shared__ uint32_t s_wave[kRsWaves];
    const uint32_t t = threadIdx.x, w = t >> 6, lane = t & 63u;
    const uint32_t n = item_count(count_ptr, cap);
    const uint32_t tile0 = blockIdx.x * kRsTile;
    if (tile0 >= n) return; // uniform over the workgroup
#pragma unroll
    for (int k = 0; k < kRsWaves; k++) s_hist[k][t] = 0;
    uint32_t all_items;
    const uint32_t dstart = block_exclusive_scan<kRsBlock>(totals[t], all_items, s_wave); // synchronises
    const uint32_t gdigit = dstart + counts[(size_t)t * n_tiles + blockIdx.x];

    const uint32_t base = tile0 + w * 64u * kRsItems + lane;
    uint32_t key[kRsItems], pos[kRsItems];
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        const uint32_t i = base + j * 64u;
        key[j] = i < n ? keys_in[i] : 0u;
    }
    // stable ranks within the wave's digit: counter before this item + lanes below with its digit
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        const bool valid = base + j * 64u < n;
        const uint64_t act = __ballot(valid);
        if (act == 0) break;
        const uint32_t d = (key[j] >> shift) & mask;
        const uint32_t first = __builtin_amdgcn_readfirstlane(d);
        uint32_t before;
        if (act == ~0ull && __ballot(d == first) == ~0ull) { // whole wave valid, one digit
            before = s_hist[w][first];
            pos[j] = before + lane;
            if (lane == 0) s_hist[w][first] = before + 64u;
        } else {
            const uint64_t m = match_digit(d, act);
            before = valid ? s_hist[w][d] : 0u;
            pos[j] = before + lanes_below(m);
            if (valid && lane == (uint32_t)__builtin_ctzll(m)) s_hist[w][d] = before + (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
    // digit t: offsets of the waves' items within the digit, the digit's count in this tile
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kRsWaves; k++) {
        const uint32_t c = s_hist[k][t];
        s_hist[k][t] = run;
        run += c;
    }
    uint32_t tile_items;
    const uint32_t tstart = block_exclusive_scan<kRsBlock>(run, tile_items, s_wave); // synchronises
#pragma unroll
    for (int k = 0; k < kRsWaves; k++) s_hist[k][t] += tstart;
    s_gbase[t] = gdigit - tstart;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        if (base + j * 64u < n) {
            pos[j] += s_hist[w][(key[j] >> shift) & mask];
            s_buf[pos[j]] = key[j];
        }
    }
    __syncthreads();
    uint32_t gpos[kRsItems];
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        const uint32_t li = j * kRsBlock + t;
        gpos[j] = 0xFFFFFFFFu;
        if (li < tile_items) {
            const uint32_t k = s_buf[li];
            gpos[j] = s_gbase[(k >> shift) & mask] + li;
            if (gpos[j] < n) keys_out[gpos[j]] = k; // always true for consistent counts
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        const uint32_t i = base + j * 64u;
        if (i < n) s_buf[pos[j]] = IDENTITY ? i : vals_in[i];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        const uint32_t li = j * kRsBlock + t;
        if (gpos[j] < n) vals_out[gpos[j]] = s_buf[li];
    }
}

// count_ptr != null: n = ceil(*count_ptr / div) (a list's chunk count, known on the device only).
__global__ __launch_bounds__(kScanBlock) void k_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t n,
                                                               const uint32_t *count_ptr, uint32_t div) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    if (count_ptr) n = (uint32_t)(((uint64_t)*count_ptr + div - 1u) / div);
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += (uint32_t)(kScanBlock * kScanItems)) {
        const uint32_t b = c0 + threadIdx.x * kScanItems;
        uint32_t v[kScanItems], sum = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            v[j] = b + j < n ? in[b + j] : 0u;
            sum += v[j];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan<kScanBlock>(sum, total, s_wave);
#pragma unroll
        for (int j = 0; j < kScanItems; j++) {
            if (b + j < n) out[b + j] = run;
            run += v[j];
        }
        carry += total;
        if (c0 + (uint32_t)(kScanBlock * kScanItems) < c0) break; // the next round's start would wrap
    }
}

uint32_t rs_tiles(uint32_t cap) { return (uint32_t)(((uint64_t)cap + kRsTile - 1) / kRsTile); }

} // namespace

size_t radix_sort_count_bytes(uint32_t cap) { return align256(256 * (size_t)rs_tiles(cap) * 4) + align256(256 * 4); }

size_t radix_sort_temp_bytes(uint32_t cap) { return 2 * align256((size_t)cap * 4) + radix_sort_count_bytes(cap); }

hipError_t radix_sort_pairs(hipStream_t stream, void *temp, size_t temp_bytes, const uint32_t *keys_in,
                            uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out, uint32_t cap,
                            const uint32_t *count_ptr, uint32_t begin_bit, uint32_t end_bit) {
    if (cap == 0) return begin_bit < end_bit && end_bit <= 32u ? hipSuccess : hipErrorInvalidValue;
    if (!temp || temp_bytes < radix_sort_temp_bytes(cap) || cap > 0x7FFFFFFFu) return hipErrorInvalidValue;
    char *p = static_cast<char *>(temp);
    const size_t a = align256((size_t)cap * 4);
    return radix_sort_pairs_with(stream, p + 2 * a, radix_sort_count_bytes(cap), reinterpret_cast<uint32_t *>(p),
                                 reinterpret_cast<uint32_t *>(p + a), keys_in, keys_out, vals_in, vals_out, cap,
                                 count_ptr, begin_bit, end_bit);
}

hipError_t radix_sort_pairs_with(hipStream_t stream, void *count_temp, size_t count_bytes, uint32_t *alt_k,
                                 uint32_t *alt_v, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                                 uint32_t *vals_out, uint32_t cap, const uint32_t *count_ptr, uint32_t begin_bit,
                                 uint32_t end_bit) {
    if (begin_bit >= end_bit || end_bit > 32u || cap > 0x7FFFFFFFu) return hipErrorInvalidValue;
    if (cap == 0) return hipSuccess;
    if (!count_temp || count_bytes < radix_sort_count_bytes(cap) || !alt_k || !alt_v || !keys_in || !keys_out ||
        !vals_out)
        return hipErrorInvalidValue;
    // the digit counts read the keys 16 B at a time
    if ((reinterpret_cast<uintptr_t>(keys_in) | reinterpret_cast<uintptr_t>(keys_out) |
         reinterpret_cast<uintptr_t>(alt_k)) % 16 != 0)
        return hipErrorInvalidValue;
    const uint32_t nt = rs_tiles(cap);
    uint32_t *counts = static_cast<uint32_t *>(count_temp);
    uint32_t *totals = reinterpret_cast<uint32_t *>(static_cast<char *>(count_temp) + align256(256 * (size_t)nt * 4));
    const uint32_t passes = (end_bit - begin_bit + 7u) / 8u;
    const uint32_t *kin = keys_in, *vin = vals_in;
    for (uint32_t q = 0; q < passes; q++) {
        const uint32_t shift = begin_bit + 8u * q;
        const uint32_t bits = end_bit - shift < 8u ? end_bit - shift : 8u;
        const uint32_t mask = (1u << bits) - 1u;
        // the last pass writes the outputs; earlier ones alternate so that pass 0 never writes in place
        const bool to_out = ((passes - 1u - q) & 1u) == 0u;
        uint32_t *kout = to_out ? keys_out : alt_k, *vout = to_out ? vals_out : alt_v;
        hipLaunchKernelGGL(k_rs_upsweep, dim3(nt), dim3(kRsBlock), 0, stream, kin, cap, count_ptr, shift, mask, nt,
                           counts);
        hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(kRsBlock), 0, stream, counts, nt, totals);
        if (vin)
            hipLaunchKernelGGL(k_rs_downsweep<false>, dim3(nt), dim3(kRsBlock), 0, stream, kin, vin, kout, vout, cap,
                               count_ptr, shift, mask, nt, (const uint32_t *)counts, (const uint32_t *)totals);
        else
            hipLaunchKernelGGL(k_rs_downsweep<true>, dim3(nt), dim3(kRsBlock), 0, stream, kin, (const uint32_t *)nullptr,
                               kout, vout, cap, count_ptr, shift, mask, nt, (const uint32_t *)counts,
                               (const uint32_t *)totals);
        kin = kout;
        vin = vout;
    }
    return hipGetLastError();
}

hipError_t exclusive_scan_u32(hipStream_t stream, const uint32_t *in, uint32_t *out, uint32_t n) {
    if (n == 0) return hipSuccess;
    if (!in || !out) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(kScanBlock), 0, stream, in, out, n, (const uint32_t *)nullptr,
                       1u);
    return hipGetLastError();
}

hipError_t exclusive_scan_u32_counted(hipStream_t stream, const uint32_t *in, uint32_t *out, const uint32_t *count_ptr,
                                      uint32_t div) {
    if (!in || !out || !count_ptr || div == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(kScanBlock), 0, stream, in, out, 0u, count_ptr, div);
    return hipGetLastError();
}

} // namespace lchip
