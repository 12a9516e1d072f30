// Sorts of one tile of (key, idx) records in LDS, over only the key bits that
// vary inside the tile, with tie-group emission for tiles that hold a whole
// bucket.  Tiles are launched by size class (1024 / 2048 / 4096 records).
//
// k_tile_radix (CAPS_FLAG_TILE_RADIX): 4-bit digit passes.  Items are held
// blocked (thread t owns items t*ITEMS .. t*ITEMS+ITEMS-1, consecutive in the
// current order).  Each pass counts digits in registers (sixteen 8-bit fields
// in two words per thread), writes the counts to LDS columns
// cnt[digit][thread], scans them digit-major, and scatters: thread order =
// item order, so the sort is stable without atomics or ballots.  LDS rows are
// padded by one slot per thread's run to keep the blocked reads
// conflict-light.
//
// k_tile_merge (the default): a merge sort, described below.
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
                                                    const Tile *__restrict__ tiles, GroupLists gl, int emit,
                                                    int keep_keys) {
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
    const bool wk = keep_keys || !tl.whole;  // sorted keys read again: see k_tile_merge
    for (uint32_t i = tid; i < len; i += B) {
        if (wk) keys[tl.start + i] = sk[px(i)];
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
// Merge sort of one tile (the default tile sort).  Records are unique u64
// values: the key bits that vary inside the tile above the tile index (a
// strict total order, so the sort is stable and no two records tie).  Every
// thread sorts ITEMS = T/256 records in registers (Batcher's odd-even merge
// network), then log2(256) merge levels double the sorted runs.  Runs are
// kept in LDS in a bitonic layout: even runs ascending, odd runs descending.
// The merge of an ascending A and a descending B then reads, past the end of
// either run, the other run's largest records, which are never taken before
// a smaller one: the serial merge needs no bounds tests and no branches, and
// every read stays inside the two runs.  The merge path of each thread is
// found by bit descent over its diagonal.
//
// The sort buffer is the only large LDS array (18.4 KB for 2,048 records):
// eight workgroups per compute unit.  After the sort the same buffer stages
// the tile's index values (original order, then sorted) and then the sorted
// keys, so both outputs are written with coalesced stores.
// ---------------------------------------------------------------------------

// Orders this wavefront's LDS accesses: every earlier LDS write of the
// wavefront is complete and visible to its lanes before any later LDS access.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Batcher's odd-even merge sort network over N registers: cas(a, b) for each
// comparator a < b (19 comparators for N = 8, 63 for N = 16).
template <int N, typename F>
__device__ __forceinline__ void batcher_network(F &&cas) {
#pragma unroll
    for (int p = 1; p < N; p <<= 1)
#pragma unroll
        for (int k = p; k >= 1; k >>= 1)
#pragma unroll
            for (int j = k % p; j + k < N; j += 2 * k)
#pragma unroll
                for (int i = 0; i < (k < N - j - k ? k : N - j - k); ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) cas(i + j, i + j + k);
}

// Exclusive max-scan of one u32 per thread across a 256-thread block (0 for
// thread 0).
__device__ __forceinline__ uint32_t block_excl_max256(uint32_t v, uint32_t *wmax) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x = x > y ? x : y;
    }
    const uint32_t ex = __shfl_up(x, 1);
    if (lane == 63) wmax[w] = x;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < w) before = before > wmax[q] ? before : wmax[q];
    const uint32_t mine = lane ? ex : 0u;
    return before > mine ? before : mine;
}

// Sorts the B*ITEMS unique records held blocked in registers (thread t holds
// positions t*ITEMS .. t*ITEMS+ITEMS-1); positions >= len hold pads (~0),
// above every real record.  sk: the padded LDS buffer (one pad slot per
// ITEMS).  Returns with the sorted records in the same blocked registers.
template <int B, int ITEMS>
__device__ __forceinline__ void block_merge_sort(uint64_t (&k)[ITEMS], uint64_t *sk, uint32_t len) {
    constexpr uint32_t T = B * ITEMS;
    auto px = [](uint32_t i) { return i + i / ITEMS; };
    const uint32_t tid = threadIdx.x;
    const uint32_t p0 = tid * ITEMS;
    if (p0 < len) {
        batcher_network<ITEMS>([&](int a, int b) {
            const uint64_t ta = k[a], tb = k[b];
            const bool sw = tb < ta;
            k[a] = sw ? tb : ta;
            k[b] = sw ? ta : tb;
        });
    }
    // While a merged run (2w records) lies inside one wavefront, only that
    // wavefront reads the LDS it wrote: a wavefront fence replaces the
    // workgroup barrier.
    constexpr uint32_t kWaveSpan = 64 * ITEMS;
#pragma unroll 1
    for (uint32_t w = ITEMS; w < T; w *= 2) {
        const bool in_wave = 2 * w <= kWaveSpan;
        if (in_wave) wave_lds_sync();
        else __syncthreads();
        {
            // this thread's ITEMS records of run r = p0 / w, at value ranks
            // q0 .. q0+ITEMS-1 of the run: an aligned group of ITEMS slots,
            // reversed in an odd (descending) run
            const uint32_t r = p0 / w, q0 = p0 - r * w;
            const bool desc = r & 1u;
            const uint32_t g = px(desc ? (r + 1) * w - ITEMS - q0 : p0);
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) sk[g + (desc ? ITEMS - 1 - j : j)] = k[j];
        }
        if (in_wave) wave_lds_sync();
        else __syncthreads();
        const uint32_t lgper = __builtin_ctz(2 * w / ITEMS);  // threads per merged run: a power of two
        const uint32_t q = tid & ((1u << lgper) - 1);
        const uint32_t a0 = (tid >> lgper) * 2 * w, b0 = a0 + w;
        const uint32_t diag = q * ITEMS;
        const uint32_t realc = len > a0 ? (len - a0 < 2 * w ? len - a0 : 2 * w) : 0;
        // Merge sort keeps index ranges together, so a run's real records come
        // first and their count is known.  Outputs past it are pads; when B
        // holds only pads the merged run is A (all of whose records are real
        // whenever B has one) followed by pads.
        if (diag >= realc || b0 >= len) {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) k[j] = diag + j < realc ? sk[px(a0 + diag + j)] : ~0ull;
            continue;
        }
        // B's i-th smallest record is at slot bt - i.  Merge path: the largest
        // x in [lo, hi] with A[x-1] <= B[diag-x] (bit descent, the same
        // log2(w) + 1 steps on every lane, no branch inside).  A lane with
        // diag > 0 has lo < hi, so the clamped candidate min(c, hi) lies in
        // [lo + 1, hi] and both of its reads stay inside the two runs; a lane
        // with diag = 0 has x = 0 and skips the search.
        const uint32_t bt = b0 + w - 1;
        const uint32_t lo = diag > w ? diag - w : 0, hi = diag < w ? diag : w;
        uint32_t x = lo;
        if (diag) {
            const uint32_t bb = bt - diag;  // B slot of candidate c: bb + c
#pragma unroll 1
            for (uint32_t st = w; st; st >>= 1) {
                const uint32_t c = x + st;
                const uint32_t cc = c < hi ? c : hi;
                const bool take = !(sk[px(bb + cc)] < sk[px(a0 + cc - 1)]);
                x = (c <= hi && take) ? c : x;
            }
        }
        // Serial merge: pa walks A up, pb walks B's slots down.  A exhausted,
        // pa reads B's largest record; B exhausted, pb reads A's largest: the
        // other side is then always taken (a tie is the same slot, read by
        // both, or pads, whose order does not matter).  After j outputs
        // pb = pa + e - j, so only pa is carried, and both heads are read
        // again each step (two independent loads instead of one load and
        // the selects that would keep the untaken head).
        const uint32_t e = bt - diag - a0;
        uint32_t pa = a0 + x;
        uint64_t ka = sk[px(pa)], kb = sk[px(pa + e)];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool ta = ka <= kb;
            k[j] = ta ? ka : kb;
            pa += ta ? 1u : 0u;
            if (j + 1 < ITEMS) {
                ka = sk[px(pa)];
                kb = sk[px(pa + e - (uint32_t)(j + 1))];
            }
        }
    }
}

// One tile per workgroup: sort, write keys and indices in place, and (for a
// tile that holds a whole bucket, when emit is set) append its tie groups to
// the group lists.  Tiles whose keys vary in more than 64 - log2(T) bits sort
// twice, stably: by the low bits, then by the high bits and the rank of the
// first pass.
template <typename I, int T, bool AOS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(T <= 2048 ? 8 : 4))) void k_tile_merge(
    uint64_t *__restrict__ keys, I *__restrict__ idx, const Tile *__restrict__ tiles, GroupLists gl, int emit,
    int keep_keys, const KRec<I> *__restrict__ rec = nullptr) {
    constexpr int B = 256;
    constexpr int ITEMS = T / B;
    constexpr int PADT = T + B;
    constexpr int LB = __builtin_ctz(T);  // tile index bits
    static_assert(sizeof(I) * T <= sizeof(uint64_t) * PADT, "index staging fits the sort buffer");
    static_assert(3 * LB <= 64, "two-pass records: high bits, rank, tile index");
    __shared__ uint64_t sk[PADT];
    __shared__ unsigned long long s_or;
    __shared__ uint32_t gcnt[kGroupClasses], gbase[kGroupClasses], wmax[4];
    __shared__ I wlast[4];
    auto px = [](uint32_t i) { return i + i / ITEMS; };
    I *stage = (I *)sk;

    const Tile tl = tiles[blockIdx.x];
    const uint32_t len = tl.len;
    const uint32_t tid = threadIdx.x;
    if (len == 0) return;
    // AOS: the records come from rec (key, index) and the sorted index
    // values and keys go to idx and keys; else keys and idx hold both
    const uint64_t first = AOS ? rec[tl.start].key : keys[tl.start];
    const uint32_t p0 = tid * ITEMS;
    uint64_t my_or = 0;
    // each thread loads its own ITEMS consecutive keys into registers, and
    // their index values, which go to LDS after the sort: no LDS round trip
    // before the sort, no global one between the sort and the output
    uint64_t kk[ITEMS];
    I ix[ITEMS];
    // a thread whose records all lie inside the tile loads (and later
    // stores) them unconditionally: adjacent accesses merge into 16-byte
    // vector loads and stores (2,048-record class 2.54 -> 2.46 ms)
    if (p0 + ITEMS <= len) {
        // packed 12-byte records: the thread's 3 * ITEMS dwords as 16-byte loads
        uint32_t wd[sizeof(KRec<I>) == 12 ? 3 * ITEMS : 1];
        if constexpr (AOS && sizeof(KRec<I>) == 12) {
            typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
            const u32x4_a4 *src = (const u32x4_a4 *)(rec + tl.start + p0);
#pragma unroll
            for (int q = 0; q < 3 * ITEMS / 4; ++q) {
                const u32x4_a4 v = src[q];
                wd[4 * q] = v.x, wd[4 * q + 1] = v.y, wd[4 * q + 2] = v.z, wd[4 * q + 3] = v.w;
            }
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if constexpr (AOS && sizeof(KRec<I>) == 12) {
                kk[j] = (uint64_t)wd[3 * j] | ((uint64_t)wd[3 * j + 1] << 32);
                ix[j] = (I)wd[3 * j + 2];
            } else if constexpr (AOS) {
                const KRec<I> r = rec[tl.start + p0 + j];
                kk[j] = r.key;
                ix[j] = r.idx;
            } else {
                kk[j] = keys[tl.start + p0 + j];
                ix[j] = idx[tl.start + p0 + j];
            }
            my_or |= kk[j] ^ first;
        }
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t i = p0 + j;
            if constexpr (AOS) {
                const KRec<I> r = i < len ? rec[tl.start + i] : KRec<I>{~0ull, (I)0};
                kk[j] = r.key;
                ix[j] = r.idx;
            } else {
                kk[j] = i < len ? keys[tl.start + i] : ~0ull;
                ix[j] = i < len ? idx[tl.start + i] : (I)0;
            }
            my_or |= i < len ? kk[j] ^ first : 0ull;
        }
    }
    if (tid == 0) s_or = 0;
    if (tid < kGroupClasses) gcnt[tid] = 0;
    __syncthreads();
    if (my_or) atomicOr(&s_or, (unsigned long long)my_or);
    __syncthreads();
    const uint64_t diff = s_or;  // bits that vary inside the tile
    const int hi = diff ? 64 - __builtin_clzll(diff) : 0;
    const bool one_pass = hi + LB <= 64;
    // records: (varying key bits below `hi`, tile index); a pad (~0) is above
    // every real record (a real record reaches ~0 only with index T-1, and
    // then the tile is full)
    const uint64_t mask = one_pass ? ((1ull << hi) - 1) : ((1ull << (64 - LB)) - 1);
    uint64_t k[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t i = p0 + j;
        k[j] = i < len ? ((kk[j] & mask) << LB) | i : ~0ull;
    }
    if (diff) block_merge_sort<B, ITEMS>(k, sk, len);
    if (!one_pass) {
        // second pass: (top LB key bits, rank after the first pass, tile index);
        // its first level writes the buffer every wavefront read in the last
        // level of the first pass
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t p = p0 + j;
            const uint32_t o1 = (uint32_t)(k[j] & (T - 1));
            const uint64_t ko = p < len ? (AOS ? rec[tl.start + o1].key : keys[tl.start + o1]) : 0ull;
            k[j] = p < len ? ((ko >> (64 - LB)) << (2 * LB)) | ((uint64_t)p << LB) | o1 : ~0ull;
        }
        block_merge_sort<B, ITEMS>(k, sk, len);
    }
    __syncthreads();  // the sort buffer is free
    // index values: original order into LDS, gathered in sorted order, each
    // thread's ITEMS consecutive ones written out directly
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        if (p0 + j < len) stage[p0 + j] = ix[j];
    __syncthreads();
    I o[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) o[j] = p0 + j < len ? stage[(uint32_t)(k[j] & (T - 1))] : (I)0;
    if (p0 + ITEMS <= len) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) idx[tl.start + p0 + j] = o[j];
    } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            if (p0 + j < len) idx[tl.start + p0 + j] = o[j];
    }
    if ((tid & 63) == 63) wlast[tid >> 6] = o[ITEMS - 1];
    uint64_t key[ITEMS];
    const uint64_t top = first & ~mask;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        key[j] = one_pass     ? top | (k[j] >> LB)
                 : p0 + j >= len ? ~0ull
                 : AOS         ? rec[tl.start + (uint32_t)(k[j] & (T - 1))].key
                               : keys[tl.start + (uint32_t)(k[j] & (T - 1))];
    __syncthreads();  // every gather from the staged index values is done
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) sk[px(p0 + j)] = key[j];  // pads (positions >= len) are never read
    __syncthreads();
    // sorted keys are read again only by the key merges of oversized buckets
    // (tiles that are not whole) and by the LCP array's bucket boundaries
    if (keep_keys || !tl.whole)
        for (uint32_t i = tid; i < len; i += B) keys[tl.start + i] = sk[px(i)];
    if (!emit || !tl.whole || len < 2) return;
    // Tie groups of the bucket: a run of equal keys starts where the key
    // changes; an exclusive max-scan over the block carries the latest run
    // start into each thread, and the thread holding a run's last record
    // emits it: a slot of its size class in LDS, then one reservation per
    // class in the global lists.  (size, class, slot) packed in a register:
    // size <= T <= 4096: 13 bits; 8 classes: 3 bits; slot < T / 2: 11 bits.
    static_assert(T <= 4096 && kGroupClasses <= 8, "group record packing");
    const uint64_t prev = p0 ? sk[px(p0 - 1)] : 0ull;
    const uint64_t next = p0 + ITEMS < (uint32_t)T ? sk[px(p0 + ITEMS)] : 0ull;
    I oprev = (I)__shfl_up((unsigned long long)o[ITEMS - 1], 1);  // the previous thread's last index
    if ((tid & 63) == 0 && tid) oprev = wlast[(tid >> 6) - 1];
    uint32_t st_mask = 0, end_mask = 0, last_start = 0;  // run starts as position + 1
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t p = p0 + j;
        const bool st = p == 0 || (j ? key[j - 1] : prev) != key[j];
        const bool en = p + 1 >= len || (j + 1 < ITEMS ? key[j + 1] : next) != key[j];
        if (p < len) {
            st_mask |= st ? 1u << j : 0u;
            end_mask |= en ? 1u << j : 0u;
            last_start = st ? p + 1 : last_start;
        }
    }
    uint32_t rs = block_excl_max256(last_start, wmax);
    uint32_t info[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t p = p0 + j;
        rs = (st_mask >> j) & 1u ? p + 1 : rs;
        info[j] = 0;
        const uint32_t size = p + 2 - rs;
        if (((end_mask >> j) & 1u) && size >= 2) {
            const uint32_t c = (uint32_t)group_class(size);
            info[j] = size << 14 | c << 11 | atomicAdd(&gcnt[c], 1u);
        }
    }
    __syncthreads();
    if (tid < kGroupClasses) gbase[tid] = gcnt[tid] ? atomicAdd(&gl.counts[tid], gcnt[tid]) : 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (info[j]) {
            const uint32_t size = info[j] >> 14, c = (info[j] >> 11) & 7u;
            const uint32_t r = p0 + j + 1 - size;
            const uint32_t slot = gbase[c] + (info[j] & 0x7FFu);
            if (c == 0) ((PairRec<I> *)gl.list[0])[slot] = PairRec<I>{tl.start + r, j ? o[j - 1] : oprev, o[j]};
            else gl.list[c][slot] = Group{tl.start + r, size};
        }
    }
}

}  // namespace caps
