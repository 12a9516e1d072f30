// LDS radix sort of one tile of (key, idx) records, over only the key bits
// that vary inside the tile, with tie-group emission for tiles that hold a
// whole bucket.
//
// Items are held blocked (thread t owns items t*ITEMS .. t*ITEMS+ITEMS-1,
// consecutive in the current order).  Each 4-bit digit pass counts digits in
// registers (sixteen 8-bit fields in two words per thread), writes the
// counts to LDS columns cnt[digit][thread], scans them digit-major, and
// scatters: thread order = item order, so the sort is stable without atomics
// or ballots.  LDS rows are padded by one slot per
// thread's run to keep the blocked reads conflict-light.  Tiles are launched
// by size class (1024 / 2048 / 4096 records).
#pragma once
#include "device.hpp"
#include "refine.hpp"

namespace caps {

struct Tile {
    uint64_t start;
    uint32_t len;
    uint32_t whole;  // 1 when the tile is a whole bucket (its tie groups are complete)
};

constexpr int kTileClasses = 3;  // capacities 1024, 2048, 4096
__host__ __device__ __forceinline__ int tile_class(uint32_t len) { return len <= 1024 ? 0 : len <= 2048 ? 1 : 2; }

// Exclusive scan of one u32 per thread across a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < w) before += wsum[q];
    __syncthreads();
    return before + x - v;
}

template <typename I, int T>
__global__ __launch_bounds__(256) void k_tile_radix(uint64_t *__restrict__ keys, I *__restrict__ idx,
                                                    const Tile *__restrict__ tiles, GroupLists gl, int emit) {
    constexpr int B = 256;
    constexpr int ITEMS = T / B;
    constexpr int PADT = T + B;  // one pad slot per thread run
    __shared__ uint64_t sk[PADT];
    __shared__ uint16_t sl[PADT];
    __shared__ I si[T];
    __shared__ uint32_t cnt[16 * B];
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long s_or;
    __shared__ uint32_t gcnt[kGroupClasses], gbase[kGroupClasses];
    auto px = [](uint32_t i) { return i + i / ITEMS; };

    const Tile tl = tiles[blockIdx.x];
    const uint32_t len = tl.len;
    const int tid = threadIdx.x;
    if (len == 0) return;
    const uint64_t first = keys[tl.start];
    uint64_t my_or = 0;
    for (uint32_t i = tid; i < (uint32_t)T; i += B) {
        if (i < len) {
            uint64_t kk = keys[tl.start + i];
            sk[px(i)] = kk;
            sl[px(i)] = (uint16_t)i;
            si[i] = idx[tl.start + i];
            my_or |= kk ^ first;
        } else {
            sk[px(i)] = ~0ull;
            sl[px(i)] = 0xFFFF;
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
        uint64_t k[ITEMS];
        uint16_t li[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            k[j] = sk[px(tid * ITEMS + j)];
            li[j] = sl[px(tid * ITEMS + j)];
        }
        for (int shift = lo; shift < hi; shift += 4) {
            // per-thread digit counts in registers: 16 x 8-bit fields in two words
            uint64_t h0 = 0, h1 = 0;
            uint32_t r[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                uint32_t d = (uint32_t)(k[j] >> shift) & 15u;
                uint32_t sh = (d & 7u) * 8u;
                uint64_t hw = d < 8 ? h0 : h1;
                r[j] = (uint32_t)(hw >> sh) & 0xFFu;
                if (d < 8) h0 += 1ull << sh;
                else h1 += 1ull << sh;
            }
#pragma unroll
            for (int d = 0; d < 16; ++d)
                cnt[d * B + tid] = (uint32_t)(((d < 8 ? h0 : h1) >> ((d & 7) * 8)) & 0xFFu);
            __syncthreads();
            // exclusive scan of the 16*B counters in digit-major order; thread
            // u owns the flat range [16u, 16u + 16)
            uint32_t v[16], s = 0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                v[q] = cnt[16 * tid + q];
                s += v[q];
            }
            uint32_t run = block_excl_scan256(s, wsum);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                cnt[16 * tid + q] = run;
                run += v[q];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                uint32_t d = (uint32_t)(k[j] >> shift) & 15u;
                uint32_t p = cnt[d * B + tid] + r[j];
                sk[px(p)] = k[j];
                sl[px(p)] = li[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                k[j] = sk[px(tid * ITEMS + j)];
                li[j] = sl[px(tid * ITEMS + j)];
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < len; i += B) {
        keys[tl.start + i] = sk[px(i)];
        idx[tl.start + i] = si[sl[px(i)]];
    }
    if (!emit || !tl.whole || len < 2) return;
    // tie groups of the bucket: two passes (count, then place)
    if (tid < kGroupClasses) gcnt[tid] = 0;
    __syncthreads();
    for (uint32_t r = tid; r < len; r += B) {
        uint64_t kr = sk[px(r)];
        if ((r == 0 || sk[px(r - 1)] != kr) && r + 1 < len && sk[px(r + 1)] == kr) {
            uint32_t e = r + 2;
            while (e < len && sk[px(e)] == kr) ++e;
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
        uint64_t kr = sk[px(r)];
        if ((r == 0 || sk[px(r - 1)] != kr) && r + 1 < len && sk[px(r + 1)] == kr) {
            uint32_t e = r + 2;
            while (e < len && sk[px(e)] == kr) ++e;
            int c = group_class(e - r);
            uint32_t slot = gbase[c] + atomicAdd(&gcnt[c], 1u);
            if (c == 0) ((PairRec<I> *)gl.list[0])[slot] = PairRec<I>{tl.start + r, si[sl[px(r)]], si[sl[px(r + 1)]]};
            else gl.list[c][slot] = Group{tl.start + r, e - r};
        }
    }
}

}  // namespace caps

namespace caps {

// ---------------------------------------------------------------------------
// Merge-path block merge sort of one tile (alternative to the radix sort):
// every thread sorts ITEMS = T/256 records in registers (odd-even
// transposition network), then eight merge levels double the sorted runs
// (8 -> 2048 for T = 2048): each thread finds where its ITEMS outputs start on
// the merge path of its two input runs (binary search in LDS) and merges them
// serially.  Records compare by (key, index in the tile), a strict total
// order, so the sort is stable, pads (key ~0, index >= len) stay behind every
// real record, and no two records tie.  LDS holds the keys and tile indices
// (blocked, one pad slot per thread run) and the positions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool rec_less(uint64_t ka, uint32_t la, uint64_t kb, uint32_t lb) {
    return ka < kb || (ka == kb && la < lb);
}

// Block merge sort of B*ITEMS records held blocked in registers (thread t
// holds records t*ITEMS .. t*ITEMS+ITEMS-1), through the padded LDS arrays
// sk/sl.  PACKED: the records are unique u64 values (key bits and tile index
// packed together; li unused); otherwise (key, li) pairs compared as above.
// Orders this wavefront's LDS accesses: every earlier LDS write of the
// wavefront is complete and visible to its lanes before any later LDS access.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int B, int ITEMS, bool PACKED>
__device__ __forceinline__ void block_merge_sort(uint64_t (&k)[ITEMS], uint32_t (&li)[ITEMS], uint64_t *sk,
                                                 uint16_t *sl, uint32_t len) {
    constexpr uint32_t T = B * ITEMS;
    auto px = [](uint32_t i) { return i + i / ITEMS; };
    auto less = [](uint64_t ka, uint32_t la, uint64_t kb, uint32_t lb) {
        return PACKED ? ka < kb : rec_less(ka, la, kb, lb);
    };
    const int tid = threadIdx.x;
    // Records with index >= len are pads, above every real record.  Merge
    // sort keeps index ranges together (the merged run [a0, a0 + 2w) holds
    // exactly the records of indices in that range), so a run's real records
    // come first and their count is known: outputs past it are pads, and a
    // second run made only of pads leaves the first one in place.
    const bool all_pads = (uint32_t)tid * ITEMS >= len;
    // register sort: odd-even transposition network over ITEMS records
#pragma unroll
    for (int round = 0; round < ITEMS && !all_pads; ++round) {
#pragma unroll
        for (int j = round & 1; j + 1 < ITEMS; j += 2) {
            if (less(k[j + 1], li[j + 1], k[j], li[j])) {
                uint64_t tk = k[j];
                k[j] = k[j + 1];
                k[j + 1] = tk;
                if (!PACKED) {
                    uint32_t tli = li[j];
                    li[j] = li[j + 1];
                    li[j + 1] = tli;
                }
            }
        }
    }
    // merge levels: runs of w records -> runs of 2w.  While a merged run
    // (2w records, 2w / ITEMS threads) lies inside one wavefront, only that
    // wavefront reads the LDS it wrote: a wavefront fence replaces the
    // workgroup barrier, and the wavefronts of the block run independently.
    constexpr uint32_t kWaveSpan = 64 * ITEMS;
#pragma unroll 1
    for (uint32_t w = ITEMS; w < T; w *= 2) {
        const bool in_wave = 2 * w <= kWaveSpan;
        if (in_wave) wave_lds_sync();
        else __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            sk[px(tid * ITEMS + j)] = k[j];
            if (!PACKED) sl[px(tid * ITEMS + j)] = (uint16_t)li[j];
        }
        if (in_wave) wave_lds_sync();
        else __syncthreads();
        const uint32_t per = (2 * w) / ITEMS;  // threads per merged run
        const uint32_t a0 = (tid / per) * 2 * w, b0 = a0 + w;
        const uint32_t diag = (tid % per) * ITEMS;
        const uint32_t realc = len > a0 ? (len - a0 < 2 * w ? len - a0 : 2 * w) : 0;
        if (diag >= realc || b0 >= len) {  // pads only, or the second run is pads
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const uint32_t q = diag + j;
                if (q < realc) {
                    k[j] = sk[px(a0 + q)];
                    if (!PACKED) li[j] = sl[px(a0 + q)];
                } else {
                    k[j] = ~0ull;
                    if (!PACKED) li[j] = 0xFFFFu;
                }
            }
            continue;
        }
        // merge path: the first x in [lo, hi) with B[diag-1-x] < A[x] (hi if
        // none), by bit descent over a range of at most w + 1 candidates: the
        // loop runs the same log2(w) + 1 steps on every lane
        const uint32_t lo = diag > w ? diag - w : 0, hi = diag < w ? diag : w;
        uint32_t x = lo;
        for (uint32_t st = w; st; st >>= 1) {
            const uint32_t c = x + st;  // candidate count of A records before the split
            if (c <= hi) {
                const uint32_t bi = b0 + diag - c, ai = a0 + c - 1;
                const bool bl = PACKED ? sk[px(bi)] < sk[px(ai)]
                                       : rec_less(sk[px(bi)], sl[px(bi)], sk[px(ai)], sl[px(ai)]);
                if (!bl) x = c;
            }
        }
        uint32_t a = a0 + x, b = b0 + (diag - x);
        const uint32_t ae = a0 + w, be = b0 + w;
        uint64_t ka = a < ae ? sk[px(a)] : ~0ull, kb = b < be ? sk[px(b)] : ~0ull;
        uint32_t la = 0, lb = 0;
        if (!PACKED) {
            la = a < ae ? sl[px(a)] : 0xFFFFFFFFu;
            lb = b < be ? sl[px(b)] : 0xFFFFFFFFu;
        }
        // serial merge without branches: each step refills only the side it
        // took, with one LDS read
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool take_a = b >= be || (a < ae && !less(kb, lb, ka, la));
            k[j] = take_a ? ka : kb;
            if (!PACKED) li[j] = take_a ? la : lb;
            a += take_a ? 1u : 0u;
            b += take_a ? 0u : 1u;
            const uint32_t nx = take_a ? a : b;
            const bool in = nx < (take_a ? ae : be);
            const uint64_t nk = in ? sk[px(nx)] : ~0ull;
            ka = take_a ? nk : ka;
            kb = take_a ? kb : nk;
            if (!PACKED) {
                const uint32_t nl = in ? sl[px(nx)] : 0xFFFFFFFFu;
                la = take_a ? nl : la;
                lb = take_a ? lb : nl;
            }
        }
    }
    __syncthreads();
}

template <typename I, int T>
__global__ __launch_bounds__(256) void k_tile_merge(uint64_t *__restrict__ keys, I *__restrict__ idx,
                                                    const Tile *__restrict__ tiles, GroupLists gl, int emit) {
    constexpr int B = 256;
    constexpr int ITEMS = T / B;
    constexpr int PADT = T + B;
    constexpr int LB = __builtin_ctz(T);  // tile index bits
    __shared__ uint64_t sk[PADT];
    __shared__ uint16_t sl[PADT];
    __shared__ I si[T];
    __shared__ unsigned long long s_or;
    __shared__ uint32_t gcnt[kGroupClasses], gbase[kGroupClasses];
    auto px = [](uint32_t i) { return i + i / ITEMS; };

    const Tile tl = tiles[blockIdx.x];
    const uint32_t len = tl.len;
    const int tid = threadIdx.x;
    if (len == 0) return;
    const uint64_t first = keys[tl.start];
    uint64_t my_or = 0;
    for (uint32_t i = tid; i < (uint32_t)T; i += B) {
        if (i < len) {
            const uint64_t kk = keys[tl.start + i];
            sk[px(i)] = kk;
            si[i] = idx[tl.start + i];
            my_or |= kk ^ first;
        } else {
            sk[px(i)] = ~0ull;
        }
        sl[px(i)] = (uint16_t)i;
    }
    if (tid == 0) s_or = 0;
    __syncthreads();
    if (my_or) atomicOr(&s_or, (unsigned long long)my_or);
    __syncthreads();
    const uint64_t diff = s_or;  // bits that vary inside the tile
    const int hi = diff ? 64 - __builtin_clzll(diff) : 0;
    uint64_t k[ITEMS];
    uint32_t li[ITEMS];
    if (diff == 0) {
        // one key: already in order
    } else if (hi + LB <= 64) {
        // varying bits below `hi`, tile index below them: unique u64 records
        // (a pad, ~0, is above every real record: when hi + LB = 64 a real
        // record reaches ~0 only with index T-1, and then the tile is full)
        const uint64_t mask = hi == 64 ? ~0ull : ((1ull << hi) - 1);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t i = tid * ITEMS + j;
            k[j] = i < len ? ((sk[px(i)] & mask) << LB) | i : ~0ull;
        }
        block_merge_sort<B, ITEMS, true>(k, li, sk, sl, len);
        const uint64_t top = first & ~mask;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            sk[px(tid * ITEMS + j)] = top | (k[j] >> LB);  // pads (positions >= len) are never read
            sl[px(tid * ITEMS + j)] = (uint16_t)(k[j] & (T - 1));
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            k[j] = sk[px(tid * ITEMS + j)];
            li[j] = sl[px(tid * ITEMS + j)];
        }
        block_merge_sort<B, ITEMS, false>(k, li, sk, sl, len);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            sk[px(tid * ITEMS + j)] = k[j];
            sl[px(tid * ITEMS + j)] = (uint16_t)li[j];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < len; i += B) {
        keys[tl.start + i] = sk[px(i)];
        idx[tl.start + i] = si[sl[px(i)]];
    }
    if (!emit || !tl.whole || len < 2) return;
    // tie groups of the bucket, in one pass over the sorted tile: each group
    // start takes a slot of its size class in LDS and keeps (size, class,
    // slot) in a register (size <= T <= 4096: 13 bits; 8 classes: 3 bits;
    // slot < T / 2: 11 bits); the block then reserves each class's slots in
    // the global lists once and writes the groups
    static_assert(T <= 4096 && kGroupClasses <= 8, "group record packing");
    if (tid < kGroupClasses) gcnt[tid] = 0;
    __syncthreads();
    uint32_t info[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t r = tid + i * B;
        info[i] = 0;
        if (r < len) {
            const uint64_t kr = sk[px(r)];
            if ((r == 0 || sk[px(r - 1)] != kr) && r + 1 < len && sk[px(r + 1)] == kr) {
                uint32_t e = r + 2;
                while (e < len && sk[px(e)] == kr) ++e;
                const uint32_t c = (uint32_t)group_class(e - r);
                info[i] = (e - r) << 14 | c << 11 | atomicAdd(&gcnt[c], 1u);
            }
        }
    }
    __syncthreads();
    if (tid < kGroupClasses) gbase[tid] = gcnt[tid] ? atomicAdd(&gl.counts[tid], gcnt[tid]) : 0u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        if (info[i]) {
            const uint32_t r = tid + i * B;
            const uint32_t size = info[i] >> 14, c = (info[i] >> 11) & 7u;
            const uint32_t slot = gbase[c] + (info[i] & 0x7FFu);
            if (c == 0) ((PairRec<I> *)gl.list[0])[slot] = PairRec<I>{tl.start + r, si[sl[px(r)]], si[sl[px(r + 1)]]};
            else gl.list[c][slot] = Group{tl.start + r, size};
        }
    }
}

}  // namespace caps
