// Stable LSD radix sort of (cell key, particle index) pairs for gfx950.
//
// Replaces the reference's 2-bit LSD sort (RadixSort::Run / SortBits,
// src/RadixSort.cpp:124-200; counting.glsl, blockscan.glsl, addblocksum.glsl,
// globalsort.glsl). Both are stable sorts by the key bits [0, key_bits), so the
// permutation is identical (ties in ascending index). Here: 8-bit digits, so
// ceil(key_bits / 8) passes instead of key_bits / 2, and each pass is
//   upsweep   per-tile digit counts (wave64 ballot match, one LDS add per
//             distinct digit per wave), digit-major like the reference's
//             blocksum layout (counting.glsl:96-101);
//   scan      per digit, an exclusive scan of its tile counts (one workgroup
//             per digit); the downsweep adds the scanned digit totals;
//   downsweep stable in-tile ranks (ballot match + per-(item, wave) digit
//             offsets scanned in LDS) and the scatter.
// A tile is 256 threads x 8 items = 2048 keys, element order (item, wave, lane).
#include "pbf_internal.h"

namespace pbf {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;

__device__ __forceinline__ unsigned lane_id() { return threadIdx.x & 63u; }

// Lanes of the wave whose 8-bit digit equals this lane's (valid lanes only).
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, bool valid) {
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ __launch_bounds__(RS_THREADS) void k_upsweep(const uint32_t *__restrict__ keys, int32_t n, int shift,
                                                        int32_t tiles, uint32_t *__restrict__ counts) {
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int32_t base = blockIdx.x * RS_TILE;
    const unsigned long long lt = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const int32_t idx = base + k * RS_THREADS + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & 255u : 0u;
        const unsigned long long m = match_digit(d, valid);
        if (valid && (m & lt) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = hist[threadIdx.x];
}

// Inclusive scan over the 64 lanes of a wave (all active) with DPP moves, no
// LDS: Hillis-Steele within each 16-lane row (row_shr 1, 2, 4, 8; a lane whose
// source lies before its row reads 0), then row 0's total into row 1 and row
// 2's into row 3 (row_bcast 15 on rows 1 and 3), then the first two rows'
// total into rows 2 and 3 (row_bcast 31 on rows 2 and 3); rows outside a row
// mask add 0.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Inclusive scan of one value per thread over a 256-thread block; returns the
// inclusive prefix, and the block total in *total (LDS scratch of 4 words).
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t *wsum, uint32_t *total) {
    const uint32_t inc = wave_scan_dpp(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) {
        const uint32_t t = wsum[k];
        before += k < w ? t : 0u;
        all += t;
    }
    __syncthreads();
    *total = all;
    return before + inc;
}

// Per digit d (one workgroup each): exclusive scan of counts[d][0, tiles) in
// place, row total to totals[d]. The digit offsets (scan of totals) are added
// by the downsweep.
__global__ __launch_bounds__(RS_THREADS) void k_scan_rows(uint32_t *__restrict__ counts, int32_t tiles,
                                                          uint32_t *__restrict__ totals) {
    __shared__ uint32_t wsum[RS_WAVES];
    uint32_t *row = counts + (size_t)blockIdx.x * tiles;
    uint32_t run = 0;
    for (int32_t base = 0; base < tiles; base += RS_THREADS) {
        const int32_t k = base + threadIdx.x;
        const uint32_t v = k < tiles ? row[k] : 0u;
        uint32_t tot;
        const uint32_t inc = block_scan256(v, wsum, &tot);
        if (k < tiles) row[k] = run + inc - v;
        run += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

__global__ __launch_bounds__(RS_THREADS) void k_downsweep(const uint32_t *__restrict__ keys_in,
                                                          const uint32_t *__restrict__ vals_in, int32_t n, int shift,
                                                          int32_t tiles, const uint32_t *__restrict__ offsets,
                                                          const uint32_t *__restrict__ totals,
                                                          uint32_t *__restrict__ keys_out,
                                                          uint32_t *__restrict__ vals_out,
                                                          const float4 *__restrict__ gather_src,
                                                          float4 *__restrict__ gather_dst,
                                                          uint2 *__restrict__ q_dst, float qscale,
                                                          uint32_t *__restrict__ tiny_dst) {
    __shared__ uint32_t cnt[RS_ITEMS * RS_WAVES][256];
    __shared__ uint32_t wsum[RS_WAVES];
    for (int s = 0; s < RS_ITEMS * RS_WAVES; s++) cnt[s][threadIdx.x] = 0;
    // global start of digit d (thread d): exclusive scan of the digit totals
    const uint32_t dt = totals[threadIdx.x];
    uint32_t all;
    const uint32_t digit_base = block_scan256(dt, wsum, &all) - dt;
    const uint32_t tile_base = digit_base + offsets[(size_t)threadIdx.x * tiles + blockIdx.x];
    __syncthreads();
    const int32_t base = blockIdx.x * RS_TILE;
    const int w = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane_id()) - 1ull;
    uint32_t key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const int32_t idx = base + k * RS_THREADS + threadIdx.x;
        const bool valid = idx < n;
        key[k] = valid ? keys_in[idx] : 0u;
        val[k] = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
        const uint32_t d = (key[k] >> shift) & 255u;
        const unsigned long long m = match_digit(d, valid);
        rank[k] = (uint32_t)__popcll(m & lt);
        if (valid && rank[k] == 0) cnt[k * RS_WAVES + w][d] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
        uint32_t run = tile_base;
        for (int s = 0; s < RS_ITEMS * RS_WAVES; s++) {
            const uint32_t t = cnt[s][threadIdx.x];
            cnt[s][threadIdx.x] = run;
            run += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        const int32_t idx = base + k * RS_THREADS + threadIdx.x;
        if (idx >= n) continue;
        const uint32_t d = (key[k] >> shift) & 255u;
        const uint32_t dst = cnt[k * RS_WAVES + w][d] + rank[k];
        if (keys_out) keys_out[dst] = key[k];
        if (vals_out) vals_out[dst] = val[k];
        if (gather_dst) {  // the record itself (last pass)
            const float4 r = gather_src[val[k]];
            gather_dst[dst] = r;
            if (q_dst) q_dst[dst] = quantize_pos(r, qscale);
            note_tiny(tiny_dst, r);
        }
    }
}

}  // namespace

bool radix_alloc(RadixScratch &s, int32_t n) {
    radix_free(s);
    const int32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    if (hipMalloc(&s.keys_alt, nn * 4) != hipSuccess) return false;
    if (hipMalloc(&s.vals, nn * 4) != hipSuccess) return false;
    if (hipMalloc(&s.vals_alt, nn * 4) != hipSuccess) return false;
    if (hipMalloc(&s.counts, (size_t)256 * (tiles > 0 ? tiles : 1) * 4) != hipSuccess) return false;
    if (hipMalloc(&s.totals, 256 * 4) != hipSuccess) return false;
    s.capacity = n;
    s.tiles = tiles;
    return true;
}

void radix_free(RadixScratch &s) {
    if (s.keys_alt) (void)hipFree(s.keys_alt);
    if (s.vals) (void)hipFree(s.vals);
    if (s.vals_alt) (void)hipFree(s.vals_alt);
    if (s.counts) (void)hipFree(s.counts);
    if (s.totals) (void)hipFree(s.totals);
    s = RadixScratch();
}

void radix_sort_perm(RadixScratch &s, uint32_t *keys, int32_t n, int key_bits, uint32_t *perm, hipStream_t st,
                     const float4 *gather_src, float4 *gather_dst, uint2 *q_dst, float qscale, uint32_t *tiny_dst) {
    if (n <= 0) return;
    const int32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    const int passes = key_bits <= 8 ? 1 : (key_bits + 7) / 8;
    uint32_t *kin = keys, *kout = s.keys_alt;
    const uint32_t *vin = nullptr;
    uint32_t *vbuf[2] = {s.vals, s.vals_alt};
    for (int p = 0; p < passes; p++) {
        const bool last = p == passes - 1;
        uint32_t *vout = last ? perm : vbuf[p & 1];
        hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(RS_THREADS), 0, st, kin, n, 8 * p, tiles, s.counts);
        hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(RS_THREADS), 0, st, s.counts, tiles, s.totals);
        hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(RS_THREADS), 0, st, kin, vin, n, 8 * p, tiles,
                           s.counts, s.totals, last ? nullptr : kout, vout, last ? gather_src : nullptr,
                           last ? gather_dst : nullptr, last ? q_dst : nullptr, qscale, last ? tiny_dst : nullptr);
        vin = vout;
        uint32_t *t = kin;
        kin = kout;
        kout = t;
    }
}

}  // namespace pbf
