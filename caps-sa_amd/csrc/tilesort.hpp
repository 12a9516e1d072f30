// LDS radix sort of one tile of (key, idx) records, over only the key bits
// that vary inside the tile, with tie-group emission for tiles that hold a
// whole bucket.
//
// Records are held warp-striped (item i = wave*64*ITEMS + j*64 + lane), so
// processing j in order and ranking lanes with a ballot-based match gives a
// stable rank without any per-item atomics; per-wave digit counters are
// scanned digit-major to place every item.  8-bit digits.
#pragma once
#include "device.hpp"
#include "refine.hpp"

namespace caps {

struct Tile {
    uint64_t start;
    uint32_t len;
    uint32_t whole;  // 1 when the tile is a whole bucket (its tie groups are complete)
};

template <typename I, int T, int B>
__global__ __launch_bounds__(B) void k_tile_radix(uint64_t *__restrict__ keys, I *__restrict__ idx,
                                                  const Tile *__restrict__ tiles, GroupLists gl, int emit) {
    constexpr int WAVES = B / 64;
    constexpr int ITEMS = T / B;
    __shared__ uint64_t sk[T];
    __shared__ uint16_t sl[T];
    __shared__ I si[T];
    __shared__ uint32_t cnt[WAVES * 256];
    __shared__ uint32_t dsum[256];
    __shared__ unsigned long long s_or;
    __shared__ uint32_t gcnt[kGroupClasses], gbase[kGroupClasses];

    const Tile tl = tiles[blockIdx.x];
    const uint32_t len = tl.len;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (len == 0) return;
    const uint64_t first = keys[tl.start];
    uint64_t k[ITEMS];
    uint16_t li[ITEMS];
    uint64_t my_or = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t i = (uint32_t)(w * 64 * ITEMS + j * 64 + lane);
        if (i < len) {
            k[j] = keys[tl.start + i];
            li[j] = (uint16_t)i;
            si[i] = idx[tl.start + i];
            my_or |= k[j] ^ first;
        } else {
            k[j] = ~0ull;
            li[j] = 0xFFFF;
        }
    }
    if (tid == 0) s_or = 0;
    __syncthreads();
    if (my_or) atomicOr(&s_or, (unsigned long long)my_or);
    __syncthreads();
    const uint64_t diff = s_or;
    if (diff) {
        const int lo = __builtin_ctzll(diff);
        const int hi = 64 - __builtin_clzll(diff);
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        for (int shift = lo; shift < hi; shift += 8) {
            for (int d = lane; d < 256; d += 64) cnt[w * 256 + d] = 0;
            __syncthreads();
            uint32_t rank[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
                uint64_t peers = ~0ull;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    uint64_t bm = __ballot((d >> b) & 1u);
                    peers &= ((d >> b) & 1u) ? bm : ~bm;
                }
                uint32_t below = (uint32_t)__popcll(peers & lt);
                uint32_t base = cnt[w * 256 + d];
                rank[j] = base + below;
                if (below == 0) cnt[w * 256 + d] = base + (uint32_t)__popcll(peers);
            }
            __syncthreads();
            // digit totals, exclusive scan over digits, per-wave offsets
            for (int d = tid; d < 256; d += B) {
                uint32_t s = 0;
#pragma unroll
                for (int ww = 0; ww < WAVES; ++ww) s += cnt[ww * 256 + d];
                dsum[d] = s;
            }
            __syncthreads();
            for (int off = 1; off < 256; off <<= 1) {
                uint32_t v[(256 + B - 1) / B];
                int c = 0;
                for (int d = tid; d < 256; d += B, ++c) v[c] = d >= off ? dsum[d - off] : 0u;
                __syncthreads();
                c = 0;
                for (int d = tid; d < 256; d += B, ++c) dsum[d] += v[c];
                __syncthreads();
            }
            for (int d = tid; d < 256; d += B) {
                uint32_t run = d ? dsum[d - 1] : 0u;  // exclusive
#pragma unroll
                for (int ww = 0; ww < WAVES; ++ww) {
                    uint32_t c = cnt[ww * 256 + d];
                    cnt[ww * 256 + d] = run;
                    run += c;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
                uint32_t p = cnt[w * 256 + d] + rank[j];
                sk[p] = k[j];
                sl[p] = li[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                uint32_t i = (uint32_t)(w * 64 * ITEMS + j * 64 + lane);
                k[j] = sk[i];
                li[j] = sl[i];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t i = (uint32_t)(w * 64 * ITEMS + j * 64 + lane);
        if (i < len) {
            keys[tl.start + i] = k[j];
            idx[tl.start + i] = si[li[j]];
        }
        sk[i] = k[j];
    }
    if (!emit || !tl.whole || len < 2) return;
    // tie groups of the bucket: two passes (count, then place)
    if (tid < kGroupClasses) gcnt[tid] = 0;
    __syncthreads();
    for (uint32_t r = tid; r < len; r += B) {
        uint64_t kr = sk[r];
        if ((r == 0 || sk[r - 1] != kr) && r + 1 < len && sk[r + 1] == kr) {
            uint32_t e = r + 2;
            while (e < len && sk[e] == kr) ++e;
            atomicAdd(&gcnt[group_class(e - r)], 1u);
        }
    }
    __syncthreads();
    if (tid < kGroupClasses) {
        gbase[tid] = gcnt[tid] ? atomicAdd(&gl.counts[tid], gcnt[tid]) : 0u;
        gcnt[tid] = 0;
    }
    __syncthreads();
    for (uint32_t r = tid; r < len; r += B) {
        uint64_t kr = sk[r];
        if ((r == 0 || sk[r - 1] != kr) && r + 1 < len && sk[r + 1] == kr) {
            uint32_t e = r + 2;
            while (e < len && sk[e] == kr) ++e;
            int c = group_class(e - r);
            uint32_t slot = atomicAdd(&gcnt[c], 1u);
            gl.list[c][gbase[c] + slot] = Group{tl.start + r, e - r};
        }
    }
}

}  // namespace caps
