// HIP kernels of the suffix-array build (gfx950, wave64).
//
// Pipeline (one text, any position subset):
//   alphabet scan -> sample keys -> splitters -> classify+histogram ->
//   scan -> scatter (key, idx) into buckets -> LDS tile sort by key ->
//   key merges for oversized buckets -> group detection (equal keys) ->
//   pair refinement (wave LCP) / group refinement (rounds of k-symbol keys) /
//   giant groups (chunked refinement + merge-path LCP merge).
#pragma once
#include "device.hpp"
#include "refine.hpp"
#include "router.hpp"
#include "tilesort.hpp"

namespace caps {

// ---------------------------------------------------------------------------
// Alphabet scan (pack.rs:105-122): presence bitmap of the symbol values.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_alphabet(const S *__restrict__ t, uint64_t n, uint32_t *__restrict__ bitmap) {
    constexpr uint32_t WORDS = sizeof(S) == 1 ? 8 : 2048;
    __shared__ uint32_t bm[WORDS];
    for (uint32_t w = threadIdx.x; w < WORDS; w += blockDim.x) bm[w] = 0;
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    auto note = [&](uint32_t c) {
        uint32_t bit = 1u << (c & 31);
        if (!(bm[c >> 5] & bit)) atomicOr(&bm[c >> 5], bit);
    };
    if (sizeof(S) == 1 && ((uintptr_t)t & 15) == 0) {
        // 16 bytes per load over the whole 16-byte words, then the tail
        const uint64_t n16 = n >> 4;
        const uint4 *t16 = (const uint4 *)t;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
            const uint4 v = t16[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int b = 0; b < 4; ++b) note((w[q] >> (8 * b)) & 0xFFu);
        }
        for (uint64_t i = (n16 << 4) + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) note(t[i]);
    } else if (sizeof(S) == 2 && ((uintptr_t)t & 15) == 0) {
        // eight u16 symbols per 16-byte load, then the tail
        const uint64_t n8 = n >> 3;
        const uint4 *t8 = (const uint4 *)t;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride) {
            const uint4 v = t8[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                note(w[q] & 0xFFFFu);
                note(w[q] >> 16);
            }
        }
        for (uint64_t i = (n8 << 3) + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) note(t[i]);
    } else {
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) note(t[i]);
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < WORDS; w += blockDim.x)
        if (bm[w]) atomicOr(&bitmap[w], bm[w]);
}

// ---------------------------------------------------------------------------
// Kept-position bitmaps (bit q & 63 of word q >> 6 set: position q kept).
// ---------------------------------------------------------------------------
// First kept position at or after j, looking at most 32 words ahead; j
// itself when none is found there (a sample position only steers bucket
// sizes, never the output).  Bits at and past n are clear.
__device__ __forceinline__ uint64_t next_kept(const uint64_t *__restrict__ keep, uint64_t j, uint64_t n) {
    uint64_t w = j >> 6;
    uint64_t m = keep[w] & (~0ull << (j & 63));
    const uint64_t nw = (n + 63) >> 6;
    for (int t = 0; t < 32; ++t) {
        if (m) return (w << 6) + (uint64_t)__builtin_ctzll(m);
        if (++w >= nw) break;
        m = keep[w];
    }
    return j;
}

// Clears the bits at and past n (the last word), then counts the kept
// positions into *total.
__global__ __launch_bounds__(256) void k_keep_count(uint64_t *__restrict__ keep, uint64_t n,
                                                    unsigned long long *__restrict__ total) {
    const uint64_t nw = (n + 63) >> 6;
    unsigned long long c = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t v = keep[w];
        if (w == nw - 1 && (n & 63)) {
            v &= (1ull << (n & 63)) - 1;
            keep[w] = v;
        }
        c += (unsigned long long)__popcll(v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += (unsigned long long)__shfl_xor((long long)c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}

// ---------------------------------------------------------------------------
// Keys of evenly spaced sample positions (ext_mem.rs:1291-1305 midpoints).
// With a kept-position bitmap the midpoints are taken over the text and
// moved to the next kept position.
// ---------------------------------------------------------------------------
template <typename S, typename I>
__global__ __launch_bounds__(256) void k_sample_keys(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                     uint64_t max_ctx, const I *__restrict__ pos, uint64_t n_pos,
                                                     uint64_t s, uint64_t *__restrict__ keys, uint32_t *__restrict__ ids,
                                                     const uint64_t *__restrict__ keep, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s) return;
    const uint64_t span = keep ? n : n_pos;
    uint64_t j = ((2 * i + 1) * span) / (2 * s);
    uint64_t p = keep ? next_kept(keep, j, n) : pos ? (uint64_t)pos[j] : j;
    uint64_t seg;
    uint64_t lim = lim_seg(L, p, &seg);
    keys[i] = with_id(g, key_at_seed(t, cm, g, p, lim < max_ctx ? lim : max_ctx), seg);
    ids[i] = (uint32_t)i;
}

// Keys of a seeded pseudo-random sample of the positions (router training).
template <typename S, typename I>
__global__ __launch_bounds__(256) void k_random_keys(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                     uint64_t max_ctx, const I *__restrict__ pos, uint64_t n_pos,
                                                     uint64_t m, uint64_t seed, uint64_t *__restrict__ keys,
                                                     const uint64_t *__restrict__ keep, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;  // SplitMix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const uint64_t span = keep ? n : n_pos;
    uint64_t j = (uint64_t)(((unsigned __int128)z * span) >> 64);
    uint64_t p = keep ? next_kept(keep, j, n) : pos ? (uint64_t)pos[j] : j;
    uint64_t seg;
    uint64_t lim = lim_seg(L, p, &seg);
    keys[i] = with_id(g, key_at_seed(t, cm, g, p, lim < max_ctx ? lim : max_ctx), seg);
}

// Sets bit 0 of *bad when any of the positions is n or more, bit 1 when the
// list is not strictly increasing (one atomic per wavefront that finds one).
// With `keep` (zeroed beforehand) it also writes the list's kept-position
// bitmap, complete whenever the list turns out strictly increasing: each
// wavefront takes 64 consecutive entries; lanes whose positions share a
// bitmap word (a contiguous run of lanes) OR their bits together by a
// segmented suffix scan of shuffles, and the run's first lane ORs the word
// in (64 consecutive positions skip the scan: their bits are one range over
// at most two words).  A wavefront whose entries are out of range or out of
// order writes nothing (its list takes the position-list path).
template <typename I>
__global__ __launch_bounds__(256) void k_positions_check(const I *__restrict__ pos, uint64_t n_pos, uint64_t n,
                                                         unsigned int *__restrict__ bad,
                                                         unsigned long long *__restrict__ keep) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    bool out = false, down = false;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c * 64 < n_pos; c += nw) {
        const uint64_t i = c * 64 + lane;
        const bool have = i < n_pos;
        const uint64_t p = have ? (uint64_t)pos[i] : ~0ull;
        uint64_t prev = shfl64(p, lane ? lane - 1 : 0);
        if (lane == 0 && i > 0) prev = (uint64_t)pos[i - 1];
        const bool o = have && p >= n, d = have && i > 0 && prev >= p;
        out |= o;
        down |= d;
        if (keep && __ballot(o || d) == 0) {
            // 64 consecutive positions (a dense list away from its gaps): two
            // words at most, whose bits are a contiguous range
            const uint64_t p0 = shfl64(p, 0);
            if (__ballot(have && p == p0 + (uint64_t)lane) == ~0ull) {
                const uint32_t sh = (uint32_t)(p0 & 63);
                if (lane == 0) atomicOr(&keep[p0 >> 6], ~0ull << sh);
                if (lane == 1 && sh) atomicOr(&keep[(p0 >> 6) + 1], (1ull << sh) - 1);
                continue;
            }
            const uint64_t w = p >> 6;
            unsigned long long bits = have ? 1ull << (p & 63) : 0ull;
#pragma unroll
            for (int s = 1; s < 64; s <<= 1) {
                const uint64_t wn = shfl64(w, lane + s < 64 ? lane + s : lane);
                const unsigned long long bn = shfl64(bits, lane + s < 64 ? lane + s : lane);
                if (lane + s < 64 && wn == w) bits |= bn;
            }
            const uint64_t wp = shfl64(w, lane ? lane - 1 : 0);
            if (have && (lane == 0 || wp != w)) atomicOr(&keep[w], bits);
        }
    }
    const uint64_t bo = __ballot(out), bd = __ballot(down);
    if (lane == 0 && (bo || bd)) atomicOr(bad, (bo ? 1u : 0u) | (bd ? 2u : 0u));
}

// splitter j-1 = sorted_sample[(j * s) / p], j = 1..p-1 (ext_mem.rs:1341).
__global__ void k_pick_splitters(const uint64_t *__restrict__ sorted, uint64_t s, uint32_t p,
                                 uint64_t *__restrict__ spl) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j >= p) return;
    spl[j - 1] = sorted[((uint64_t)j * s) / p];
}

// ---------------------------------------------------------------------------
// Bucket classification.  Bucket of a key = number of splitters <= key, so
// equal keys always share a bucket.  The p buckets form coarse intervals of
// `stride` consecutive buckets each (the last may be shorter): interval c
// holds buckets [c*stride, c*stride + stride), and coarse splitter c is
// splitter (c+1)*stride - 1, the last splitter below interval c+1.  The
// classify pass finds a key's coarse interval (the number of coarse
// splitters <= key), by binary search or the learned router; the second
// scatter pass places each record of one interval into its bucket by a
// search over the interval's stride-1 inner splitters
// spl[c*stride .. c*stride + stride - 2], staged in LDS.  Both levels are
// equal-depth: every bucket and every interval covers the same share of the
// sorted sample.
// ---------------------------------------------------------------------------
constexpr uint32_t kCoarseMax = 4095;         // coarse splitters (LDS table of the classify pass)
constexpr uint32_t kCoarseBucketsMax = 4096;  // coarse intervals = kCoarseMax + 1
constexpr uint32_t kFineMax = 4096;           // buckets per coarse interval (the stride)

struct Splitters {
    const uint64_t *spl;    // p-1 sorted keys
    uint32_t n_spl;         // p-1
    uint32_t stride;        // buckets per coarse interval
    uint32_t n_coarse;      // coarse splitters: ceil(p / stride) - 1
    uint32_t coarse_steps;  // ceil(log2(n_coarse + 1)): binary-lifting steps over the coarse table
    uint32_t fine_steps;    // ceil(log2(stride)): steps inside one coarse interval
};

// Coarse splitter c (c < n_coarse).
__device__ __forceinline__ uint64_t coarse_splitter(const Splitters &SP, uint32_t c) {
    return SP.spl[(uint64_t)(c + 1) * SP.stride - 1];
}

// Splitter tables searched in LDS are kept in breadth-first (Eytzinger)
// order: node i (1-based) of the complete binary tree of height L holds the
// sorted table's entry in_order(i), entries past the table hold ~0.  The
// nodes of one tree level are contiguous, so the first levels of a search,
// where the lanes of a wavefront probe few distinct entries, read distinct
// LDS banks; in sorted order those entries lie 2^k entries apart and all fall
// in one bank (the classify pass stalled on bank conflicts for most of its
// time).
__host__ __device__ constexpr uint32_t pow2_at_least(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}
__device__ __forceinline__ uint32_t eyt_in_order(uint32_t i, uint32_t L) {
    const uint32_t d = 31 - __builtin_clz(i);
    return ((2 * (i - (1u << d)) + 1) << (L - 1 - d)) - 1;
}
// Number of the table's n entries at or below each key (tree of height L,
// node i at t[i - 1]), for N keys in lockstep.
template <int N>
__device__ __forceinline__ void eyt_upper_rank(const uint64_t *t, uint32_t L, uint32_t n, const uint64_t (&key)[N],
                                               uint32_t (&pos)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) pos[j] = 1;
    for (uint32_t lvl = 0; lvl < L; ++lvl) {
#pragma unroll
        for (int j = 0; j < N; ++j) pos[j] = 2 * pos[j] + (t[pos[j] - 1] <= key[j] ? 1u : 0u);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t r = pos[j] - (1u << L);
        pos[j] = r < n ? r : n;
    }
}

// LDS of the classify kernels, sized by the launch: the coarse table (room
// for the padded breadth-first tree), then the per-interval histogram.
__host__ __device__ constexpr uint32_t classify_lds_bytes(uint32_t n_coarse) {
    return pow2_at_least(n_coarse + 1) * 8 + (n_coarse + 1) * 4;
}

// Packed code text: word w holds the codes of symbols w*F .. w*F+F-1
// (F = 2^lgf = 64 / bits), the first in the most significant bits.
template <typename S>
__global__ void k_pack_codes(const S *__restrict__ t, uint64_t n, CodeMap<S> cm, uint32_t bits, uint32_t lgf,
                             uint64_t *__restrict__ out, uint64_t nwords) {
    const uint32_t F = 1u << lgf;
    // text at a 16-byte aligned address, a word's symbols filling whole
    // 8-byte loads (u8: 8+ codes a word, u16: 4+; aligned: j0 is a multiple
    // of F)
    constexpr uint32_t PER = 8 / sizeof(S);  // symbols per 8-byte load
    const bool vec = F >= PER && ((uintptr_t)t & 15) == 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j0 = w << lgf;
        uint64_t v = 0;
        if (vec && j0 + F <= n) {
            const uint64_t *src = (const uint64_t *)(t + j0);
            for (uint32_t q = 0; q < F / PER; ++q) {
                const uint64_t x = src[q];
#pragma unroll
                for (uint32_t b = 0; b < PER; ++b) {
                    const uint32_t c = cm((S)((x >> (8 * sizeof(S) * b)) & ((1ull << (8 * sizeof(S))) - 1)));
                    v |= (uint64_t)c << (64 - (q * PER + b + 1) * bits);
                }
            }
        } else {
            for (uint32_t f = 0; f < F; ++f) {
                const uint64_t j = j0 + f;
                const uint32_t c = j < n ? cm(t[j]) : 0u;
                v |= (uint64_t)c << (64 - (f + 1) * bits);
            }
        }
        out[w] = v;
    }
}

// Seed-width code text: K codes of `bits` bits per word (bits not a power of
// two), MSB-first, the low 64 - K*bits bits zero.
template <typename S>
__global__ void k_pack_seed(const S *__restrict__ t, uint64_t n, CodeMap<S> cm, uint32_t bits, uint32_t K,
                            uint64_t *__restrict__ out, uint64_t nwords) {
    // u8 text at a 4-byte aligned address, at most 21 codes a word (fields
    // of 3+ bits): the word's bytes from seven aligned dword loads, shifted
    // into place by byte alignment
    const bool vec = sizeof(S) == 1 && K <= 21 && ((uintptr_t)t & 3) == 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j0 = w * K;
        uint64_t v = 0;
        if (vec && j0 + 28 <= n) {
            const uint32_t *src = (const uint32_t *)(t + (j0 & ~3ull));
            const uint32_t sh = (uint32_t)(j0 & 3);
            uint32_t d[7], e[6];
#pragma unroll
            for (int q = 0; q < 7; ++q) d[q] = src[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) e[q] = __builtin_amdgcn_alignbyte(d[q + 1], d[q], sh);
#pragma unroll
            for (uint32_t f = 0; f < 21; ++f) {
                if (f < K) {
                    const uint32_t c = cm((S)((e[f >> 2] >> (8 * (f & 3))) & 0xFFu));
                    v |= (uint64_t)c << (64 - (f + 1) * bits);
                }
            }
            out[w] = v;
            continue;
        }
        for (uint32_t f = 0; f < K; ++f) {
            const uint64_t j = j0 + f;
            const uint32_t c = j < n ? cm(t[j]) : 0u;
            v |= (uint64_t)c << (64 - (f + 1) * bits);
        }
        out[w] = v;
    }
}

// The coarse splitter table (what the learned router is trained to index).
__global__ void k_coarse_table(Splitters SP, uint64_t *__restrict__ ctab) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < SP.n_coarse) ctab[c] = coarse_splitter(SP, c);
}

// Exact position given a predicted one in a sorted table `tab` of n keys (the
// number of entries <= key): verify tab[b-1] <= key < tab[b], otherwise
// gallop up to `window` steps, otherwise binary search (the fallback).
__device__ __forceinline__ uint32_t router_correct(uint32_t window, const uint64_t *tab, uint32_t n, uint64_t key,
                                                   uint32_t b, uint32_t *fallback) {
    if (b > n) b = n;
    bool lo_ok = (b == 0) || tab[b - 1] <= key;
    bool hi_ok = (b == n) || key < tab[b];
    if (lo_ok && hi_ok) return b;
    if (!hi_ok) {  // tab[b] <= key: the answer is above b
        uint32_t lo = b + 1, step = 1;
        for (uint32_t it = 0; it < window; ++it) {
            uint32_t hi = lo + step;
            if (hi > n) hi = n;
            if (hi == n || key < tab[hi]) {
                uint32_t x = lo, e = hi;
                while (x < e) {
                    uint32_t mid = (x + e) >> 1;
                    if (tab[mid] <= key) x = mid + 1;
                    else e = mid;
                }
                return x;
            }
            lo = hi + 1;
            step <<= 1;
        }
    } else {  // tab[b-1] > key: the answer is below b
        uint32_t hi = b - 1, step = 1;
        for (uint32_t it = 0; it < window; ++it) {
            uint32_t lo = hi >= step ? hi - step : 0;
            if (lo == 0 || tab[lo - 1] <= key) {
                uint32_t x = lo, e = hi;
                while (x < e) {
                    uint32_t mid = (x + e) >> 1;
                    if (tab[mid] <= key) x = mid + 1;
                    else e = mid;
                }
                return x;
            }
            hi = lo - 1;
            step <<= 1;
        }
    }
    *fallback = 1;
    uint32_t x = 0, e = n;
    while (x < e) {
        uint32_t mid = (x + e) >> 1;
        if (tab[mid] <= key) x = mid + 1;
        else e = mid;
    }
    return x;
}

// Bucket of a key = the number of splitters <= key, exactly the upper bound
// of ext_mem.rs:1705-1734 on packed keys.  It is found in two steps: the
// position among the coarse splitters (every stride-th splitter, staged in
// LDS), then among the stride-1 splitters between two coarse entries (global
// memory).  The coarse position comes from a binary search or from the
// learned router (prediction, neighbour check in LDS, bounded local search,
// binary search as the fallback).  N keys are handled in lockstep, branch
// free (binary lifting over a fixed number of steps), so every step has N
// independent loads in flight.
template <int N>
__device__ __forceinline__ void coarse_bsearch(const Splitters &SP, const uint64_t *coarse, const uint64_t (&key)[N],
                                               uint32_t (&pos)[N]) {
    eyt_upper_rank<N>(coarse, SP.coarse_steps, SP.n_coarse, key, pos);
}

// The classify kernels' coarse table: breadth-first for the binary search
// (RK == 0), sorted for the routers' local correction.
template <int RK>
__device__ __forceinline__ void load_coarse_table(const Splitters &SP, uint64_t *coarse) {
    if (RK == 0) {
        const uint32_t L = SP.coarse_steps, m = (1u << L) - 1;
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
            const uint32_t c = eyt_in_order(i + 1, L);
            coarse[i] = c < SP.n_coarse ? coarse_splitter(SP, c) : ~0ull;
        }
    } else {
        for (uint32_t c = threadIdx.x; c < SP.n_coarse; c += blockDim.x) coarse[c] = coarse_splitter(SP, c);
    }
}

template <int RK, int N>
__device__ __forceinline__ void coarse_routed(const Splitters &SP, const uint64_t *coarse, const RouterDev &R,
                                              const uint64_t *rknots, const uint16_t *rradix, const float *ra,
                                              const float *rb, const float *rinv, const uint32_t *rlo,
                                              const uint32_t *rhi, const float *nib, const float *w2,
                                              const uint64_t (&key)[N], const bool (&valid)[N], uint32_t (&pos)[N],
                                              unsigned long long &sum, unsigned long long &mx,
                                              unsigned long long &fbs, unsigned long long &cnt) {
    const uint32_t nc = SP.n_coarse;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint32_t pred = RK == 1 ? pwl_predict(R, rknots, rradix, ra, rb, rinv, rlo, rhi, key[j])
                                : mlp_predict(R, nib, w2, key[j]);
        pred = pred > nc ? nc : pred;
        const bool ok = (pred == 0 || coarse[pred - 1] <= key[j]) && (pred == nc || key[j] < coarse[pred]);
        uint32_t fb = 0;
        pos[j] = ok ? pred : router_correct(R.window, coarse, nc, key[j], pred, &fb);
        if (valid[j]) {
            unsigned long long e = pos[j] > pred ? pos[j] - pred : pred - pos[j];
            sum += e;
            mx = e > mx ? e : mx;
            fbs += fb;
            cnt += 1;
        }
    }
}


// The spline router (RK == 3): prediction, then the exact position inside
// the window [pred - eh, pred + el] that the spline's error bound guarantees,
// found by a fixed number of branch-free steps over the sorted table (N keys
// in lockstep) across the window widened by one position on each side,
// [lo, hi].  The steps return the largest c in [lo, hi] with c == lo or
// tab[c-1] <= key: a result strictly inside (lo, hi) is exact (tab[x-1] <=
// key < tab[x]), and a result on a widened end means the key's position lay
// outside its window.  Such a key is counted as a fallback, and the host,
// finding any, classifies the whole input again by binary search.
template <int N>
__device__ __forceinline__ void coarse_spline(const uint64_t *tab, uint32_t n, const RouterDev &R, const uint32_t *kp,
                                              const uint64_t (&key)[N], const bool (&valid)[N], uint32_t (&pos)[N],
                                              unsigned long long &sum, unsigned long long &mx,
                                              unsigned long long &fbs, unsigned long long &cnt) {
    const uint32_t el1 = (uint32_t)R.el + 1, eh1 = (uint32_t)R.eh + 1;
    uint32_t pred[N], x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) pred[j] = spline_predict_pair(kp, key[j]);
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = pred[j] > eh1 ? pred[j] - eh1 : 0u;
    for (uint32_t st = R.steps ? 1u << (R.steps - 1) : 0u; st; st >>= 1) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const uint32_t hi = pred[j] + el1 < n ? pred[j] + el1 : n;
            const uint32_t c = x[j] + st;
            const uint32_t cc = c < hi ? c : hi;
            const bool take = (c <= hi) & (tab[cc ? cc - 1 : 0] <= key[j]);
            x[j] = take ? c : x[j];
        }
    }
    uint32_t esum = 0, emax = 0, nv = 0, nbad = 0;  // errors are below 2^16
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const bool low = pred[j] >= eh1 && x[j] == pred[j] - eh1;
        const bool high = pred[j] + el1 <= n && x[j] == pred[j] + el1;
        nbad += valid[j] && (low || high) ? 1u : 0u;
        pos[j] = x[j];
        const uint32_t e = x[j] > pred[j] ? x[j] - pred[j] : pred[j] - x[j];
        esum += valid[j] ? e : 0u;
        emax = valid[j] && e > emax ? e : emax;
        nv += valid[j] ? 1u : 0u;
    }
    sum += esum;
    mx = emax > mx ? emax : mx;
    fbs += nbad;
    cnt += nv;
}

template <typename S, typename I, int RK>
__global__ __launch_bounds__(256) void k_classify(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                  uint64_t max_ctx, const I *__restrict__ pos, uint64_t n_pos,
                                                  Splitters SP, RouterDev R, uint16_t *__restrict__ cid,
                                                  uint32_t *__restrict__ hblk, uint32_t pc, uint64_t chunk) {
    constexpr uint32_t NS = kPwlSeg;
    extern __shared__ uint64_t classify_lds[];
    uint64_t *coarse = classify_lds;
    uint32_t *chist = (uint32_t *)(classify_lds + pow2_at_least(SP.n_coarse + 1));
    for (uint32_t c = threadIdx.x; c < pc; c += blockDim.x) chist[c] = 0;
    __shared__ uint64_t rknots[RK == 1 ? NS + 1 : 1];
    __shared__ float ra[RK == 1 ? NS : 1], rb[RK == 1 ? NS : 1], rinv[RK == 1 ? NS : 1];
    __shared__ uint32_t rlo[RK == 1 ? NS : 1], rhi[RK == 1 ? NS : 1];
    __shared__ uint16_t rradix[RK == 1 ? (1u << kRouterRadixBits) + 1 : 1];
    __shared__ float nib[RK == 2 ? 16 * 16 * kMlpHid : 1], w2[RK == 2 ? kMlpHid + 1 : 1];
    __shared__ uint32_t rkn[RK == 3 ? kSplineCells : 1];
    __shared__ unsigned long long st_sum, st_max, st_fb, st_cnt;
    load_coarse_table<RK>(SP, coarse);
    if (RK == 1) {
        for (uint32_t c = threadIdx.x; c < R.ns; c += blockDim.x) {
            rknots[c] = R.knots[c];
            ra[c] = R.a[c];
            rb[c] = R.b[c];
            rinv[c] = R.inv[c];
            rlo[c] = R.lo[c];
            rhi[c] = R.hi[c];
        }
        for (uint32_t c = threadIdx.x; c <= (1u << kRouterRadixBits); c += blockDim.x) rradix[c] = R.radix[c];
    }
    if (RK == 2) {
        for (uint32_t c = threadIdx.x; c < 16 * 16 * kMlpHid; c += blockDim.x) nib[c] = R.nib[c];
        for (uint32_t c = threadIdx.x; c <= (uint32_t)kMlpHid; c += blockDim.x) w2[c] = R.w2[c];
    }
    if (RK == 3)
        for (uint32_t c = threadIdx.x; c < kSplineCells; c += blockDim.x)
            rkn[c] = (uint32_t)R.spline[c] | ((uint32_t)R.spline[c + 1] << 16);
    if (threadIdx.x == 0) st_sum = st_max = st_fb = st_cnt = 0;
    __syncthreads();
    unsigned long long my_sum = 0, my_max = 0, my_fb = 0, my_cnt = 0;
    constexpr int NB = RK == 2 ? 1 : 8;  // the MLP's hidden layer does not fit eight keys in registers
    // this block's positions: [c0, c1), the same range the scatter's block owns
    const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t c1 = c0 + chunk < n_pos ? c0 + chunk : n_pos;
    for (uint64_t base = c0; base < c1; base += (uint64_t)blockDim.x * NB) {
        uint64_t key[NB];
        bool valid[NB];
        uint32_t b[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint64_t i = base + threadIdx.x + (uint64_t)j * blockDim.x;
            valid[j] = i < c1;
            key[j] = 0;
            if (valid[j]) {
                uint64_t p = pos ? (uint64_t)pos[i] : i;
                uint64_t seg;
                uint64_t lim = lim_seg(L, p, &seg);
                key[j] = with_id(g, key_at_seed(t, cm, g, p, lim < max_ctx ? lim : max_ctx), seg);
            }
        }
        if (RK == 0) coarse_bsearch<NB>(SP, coarse, key, b);
        else if (RK == 3)
            coarse_spline<NB>(coarse, SP.n_coarse, R, rkn, key, valid, b, my_sum, my_max, my_fb, my_cnt);
        else
            coarse_routed<RK, NB>(SP, coarse, R, rknots, rradix, ra, rb, rinv, rlo, rhi, nib, w2, key, valid, b,
                                  my_sum, my_max, my_fb, my_cnt);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint64_t i = base + threadIdx.x + (uint64_t)j * blockDim.x;
            if (valid[j]) {
                cid[i] = (uint16_t)b[j];
                atomicAdd(&chist[b[j]], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < pc; c += blockDim.x) hblk[(uint64_t)blockIdx.x * pc + c] = chist[c];
    if (RK != 0 && R.stats) {
        atomicAdd(&st_sum, my_sum);
        atomicMax(&st_max, my_max);
        atomicAdd(&st_fb, my_fb);
        atomicAdd(&st_cnt, my_cnt);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(&R.stats[0], st_sum);
            atomicMax(&R.stats[1], st_max);
            atomicAdd(&R.stats[2], st_fb);
            atomicAdd(&R.stats[3], st_cnt);
        }
    }
}

// ---------------------------------------------------------------------------
// Contiguous positions (every suffix of the text): a block owns 4096
// consecutive positions, stages their text window in LDS with coalesced
// loads, and each thread rolls the packed key over 16 consecutive positions
// into an LDS key array; classification / scatter then walk that array in
// coalesced order.
// ---------------------------------------------------------------------------
constexpr uint32_t kSeqItems = 16;
constexpr uint32_t kSeqChunk = 256 * kSeqItems;

template <typename S, int NI = kSeqItems>
__device__ __forceinline__ void thread_keys_seq(const S *__restrict__ t, uint64_t n, CodeMap<S> cm, const KeyGeom &g,
                                                const DevLimits &L, uint64_t max_ctx, uint64_t base, uint32_t cnt,
                                                S *win, uint64_t (&out)[NI]) {
    // window: positions [base, base + cnt + k - 1) clipped to the text
    __syncthreads();  // the previous chunk's readers of win are done
    uint64_t wlen = (uint64_t)cnt + g.k - 1;
    if (wlen > n - base) wlen = n - base;
    const uint32_t wl = (uint32_t)wlen;
    if (sizeof(S) == 1 && (((uintptr_t)(t + base)) & 3) == 0) {
        const uint32_t *src = (const uint32_t *)(t + base);
        uint32_t *dst = (uint32_t *)win;
        for (uint32_t i = threadIdx.x; i < wl / 4; i += blockDim.x) dst[i] = src[i];
        for (uint32_t i = (wl / 4) * 4 + threadIdx.x; i < wl; i += blockDim.x) win[i] = t[base + i];
    } else {
        for (uint32_t i = threadIdx.x; i < wl; i += blockDim.x) win[i] = t[base + i];
    }
    __syncthreads();
    const uint32_t first = threadIdx.x * NI;
    uint64_t caps[NI], keys[NI], segs[NI];
    uint64_t seg_end = 0, seg = 0;
#pragma unroll
    for (uint32_t i = 0; i < NI; ++i) {
        uint64_t q = base + first + i;
        uint64_t lim;
        if (q >= n) {
            lim = 0;
        } else if (L.ends == nullptr) {
            lim = n - q;
        } else {
            if (q >= seg_end) seg_end = q + lim_seg(L, q, &seg);
            lim = seg_end - q;
        }
        caps[i] = lim < max_ctx ? lim : max_ctx;
        segs[i] = seg;
    }
    if (first < cnt) keys_from_window<S, NI>(win, cm, g, first, caps, keys);
#pragma unroll
    for (uint32_t i = 0; i < NI; ++i) out[i] = first + i < cnt ? with_id(g, keys[i], segs[i]) : 0ull;
}

// Each thread classifies CI consecutive positions.  Up to 2,048 intervals
// (a coarse table of at most 16 KB) CI = 8 and the binary-search variant is
// held to 64 VGPRs, eight wavefronts per SIMD (the allocator spills a few
// dwords): 1.08 -> 0.99 ms on the 268M text against sixteen positions at four
// wavefronts.  With 4,096 intervals the LDS allows three workgroups per
// compute unit whatever the registers, and sixteen positions per thread keep
// more searches in flight (1.42 against 1.66 ms).  The spline variant keeps
// four wavefronts (at eight it would spill 32 dwords), the other routers
// their own allocation.
template <typename S, typename I, int RK, bool KEEP, int CI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RK == 0 ? (CI == 8 ? 8 : 4) : RK == 3 ? 4 : 1))) void k_classify_seq(const S *__restrict__ t, uint64_t n, CodeMap<S> cm, KeyGeom g,
                                                      DevLimits L, uint64_t max_ctx, Splitters SP, RouterDev R,
                                                      uint16_t *__restrict__ cid, uint32_t *__restrict__ hblk,
                                                      uint32_t pc, uint64_t chunk, const uint64_t *__restrict__ keep) {
    constexpr uint32_t NS = kPwlSeg;
    extern __shared__ uint64_t classify_lds[];
    uint64_t *coarse = classify_lds;
    uint32_t *chist = (uint32_t *)(classify_lds + pow2_at_least(SP.n_coarse + 1));
    for (uint32_t c = threadIdx.x; c < pc; c += blockDim.x) chist[c] = 0;
    __shared__ S win[(256 * CI) + 72];
    __shared__ uint64_t rknots[RK == 1 ? NS + 1 : 1];
    __shared__ float ra[RK == 1 ? NS : 1], rb[RK == 1 ? NS : 1], rinv[RK == 1 ? NS : 1];
    __shared__ uint32_t rlo[RK == 1 ? NS : 1], rhi[RK == 1 ? NS : 1];
    __shared__ uint16_t rradix[RK == 1 ? (1u << kRouterRadixBits) + 1 : 1];
    __shared__ float nib[RK == 2 ? 16 * 16 * kMlpHid : 1], w2[RK == 2 ? kMlpHid + 1 : 1];
    __shared__ uint32_t rkn[RK == 3 ? kSplineCells : 1];
    __shared__ unsigned long long st_sum, st_max, st_fb, st_cnt;
    load_coarse_table<RK>(SP, coarse);
    if (RK == 1) {
        for (uint32_t c = threadIdx.x; c < R.ns; c += blockDim.x) {
            rknots[c] = R.knots[c];
            ra[c] = R.a[c];
            rb[c] = R.b[c];
            rinv[c] = R.inv[c];
            rlo[c] = R.lo[c];
            rhi[c] = R.hi[c];
        }
        for (uint32_t c = threadIdx.x; c <= (1u << kRouterRadixBits); c += blockDim.x) rradix[c] = R.radix[c];
    }
    if (RK == 2) {
        for (uint32_t c = threadIdx.x; c < 16 * 16 * kMlpHid; c += blockDim.x) nib[c] = R.nib[c];
        for (uint32_t c = threadIdx.x; c <= (uint32_t)kMlpHid; c += blockDim.x) w2[c] = R.w2[c];
    }
    if (RK == 3)
        for (uint32_t c = threadIdx.x; c < kSplineCells; c += blockDim.x)
            rkn[c] = (uint32_t)R.spline[c] | ((uint32_t)R.spline[c + 1] << 16);
    if (threadIdx.x == 0) st_sum = st_max = st_fb = st_cnt = 0;
    unsigned long long my_sum = 0, my_max = 0, my_fb = 0, my_cnt = 0;
    // this block's positions: [c0, c1), the same range the scatter's block owns
    const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    for (uint64_t base = c0; base < c1; base += (256 * CI)) {
        uint32_t cnt = (uint32_t)(c1 - base < (256 * CI) ? c1 - base : (256 * CI));
        uint64_t key[CI];
        bool valid[CI];
        uint32_t b[CI];
        thread_keys_seq<S, CI>(t, n, cm, g, L, max_ctx, base, cnt, win, key);
        const uint32_t first = threadIdx.x * CI;
        // kept positions (bit q of the bitmap; all when keep is null): the
        // thread's positions start at a multiple of their count, inside one word
        constexpr uint32_t kAll = (uint32_t)((1ull << CI) - 1);
        uint32_t kb = kAll;
        if (KEEP && first < cnt) kb = (uint32_t)(keep[(base + first) >> 6] >> ((base + first) & 63)) & kAll;
#pragma unroll
        for (uint32_t j = 0; j < CI; ++j) valid[j] = first + j < cnt && ((kb >> j) & 1u);
        if (RK == 0) coarse_bsearch<CI>(SP, coarse, key, b);
        else if (RK == 3) {
            // groups of eight keys: sixteen keys' windows would spill
#pragma unroll
            for (int h = 0; h < (int)CI / 8; ++h) {
                uint64_t k8[8];
                bool v8[8];
                uint32_t b8[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) k8[j] = key[8 * h + j], v8[j] = valid[8 * h + j];
                coarse_spline<8>(coarse, SP.n_coarse, R, rkn, k8, v8, b8, my_sum, my_max, my_fb, my_cnt);
#pragma unroll
                for (int j = 0; j < 8; ++j) b[8 * h + j] = b8[j];
            }
        }
        else
            coarse_routed<RK, CI>(SP, coarse, R, rknots, rradix, ra, rb, rinv, rlo, rhi, nib, w2, key, valid,
                                         b, my_sum, my_max, my_fb, my_cnt);
        if (first + CI <= cnt) {  // 2 * CI contiguous, aligned bytes per thread
            uint4 *dst = (uint4 *)(cid + base + first);
#pragma unroll
            for (uint32_t q = 0; q < CI / 8; ++q)
                dst[q] = make_uint4(b[8 * q] | (b[8 * q + 1] << 16), b[8 * q + 2] | (b[8 * q + 3] << 16),
                                    b[8 * q + 4] | (b[8 * q + 5] << 16), b[8 * q + 6] | (b[8 * q + 7] << 16));
        } else {
#pragma unroll
            for (uint32_t j = 0; j < CI; ++j)
                if (valid[j]) cid[base + first + j] = (uint16_t)b[j];
        }
#pragma unroll
        for (uint32_t j = 0; j < CI; ++j)
            if (valid[j]) atomicAdd(&chist[b[j]], 1u);
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < pc; c += blockDim.x) hblk[(uint64_t)blockIdx.x * pc + c] = chist[c];
    if (RK != 0 && R.stats) {
        atomicAdd(&st_sum, my_sum);
        atomicMax(&st_max, my_max);
        atomicAdd(&st_fb, my_fb);
        atomicAdd(&st_cnt, my_cnt);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(&R.stats[0], st_sum);
            atomicMax(&R.stats[1], st_max);
            atomicAdd(&R.stats[2], st_fb);
            atomicAdd(&R.stats[3], st_cnt);
        }
    }
}

// Exclusive scan of m u64 counts into m+1 u64 offsets (one workgroup).
__global__ __launch_bounds__(1024) void k_scan_u64(const uint64_t *__restrict__ in, uint64_t m,
                                                   uint64_t *__restrict__ out) {
    __shared__ uint64_t s[1024];
    uint32_t tid = threadIdx.x;
    uint64_t per = (m + 1023) / 1024;
    uint64_t b = tid * per, e = b + per;
    if (b > m) b = m;
    if (e > m) e = m;
    uint64_t sum = 0;
    for (uint64_t i = b; i < e; ++i) sum += in[i];
    s[tid] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t v = tid >= off ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    uint64_t run = tid ? s[tid - 1] : 0;
    for (uint64_t i = b; i < e; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (tid == 1023) out[m] = s[1023];
}

// Per-block histograms (hblk[b * pc + c]) -> exclusive prefix over blocks in
// place (block b's first slot inside interval c, relative to the interval) and
// interval totals in hist[c], summed in 64 bits: an interval of 2^32 or more
// records (which the host then refuses) cannot wrap its total.  A workgroup
// of 1,024 threads takes 16 intervals; its 64 rows of threads each sum a
// contiguous slice of the blocks (independent loads, eight in flight), the
// slice sums are scanned across the rows in LDS, and every row rewrites its
// slice with its running prefix.
constexpr uint32_t kBoCols = 16, kBoRows = 64;
__global__ __launch_bounds__(1024) void k_block_offsets(uint32_t *__restrict__ hblk, uint32_t nblocks, uint32_t pc,
                                                        uint64_t *__restrict__ hist) {
    __shared__ unsigned long long part[kBoRows][kBoCols];
    const uint32_t cl = threadIdx.x % kBoCols, r = threadIdx.x / kBoCols;
    const uint32_t c = blockIdx.x * kBoCols + cl;
    const uint32_t per = (nblocks + kBoRows - 1) / kBoRows;
    const uint32_t b0 = r * per < nblocks ? r * per : nblocks;
    const uint32_t b1 = b0 + per < nblocks ? b0 + per : nblocks;
    unsigned long long sum = 0;
    if (c < pc) {
        uint32_t b = b0;
        for (; b + 8 <= b1; b += 8) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = hblk[(uint64_t)(b + u) * pc + c];
#pragma unroll
            for (int u = 0; u < 8; ++u) sum += v[u];
        }
        for (; b < b1; ++b) sum += hblk[(uint64_t)b * pc + c];
    }
    part[r][cl] = sum;
    __syncthreads();
    unsigned long long run = 0, total = 0;
    for (uint32_t q = 0; q < kBoRows; ++q) {
        const unsigned long long v = part[q][cl];
        run += q < r ? v : 0ull;
        total += v;
    }
    if (c >= pc) return;
    uint32_t b = b0;
    for (; b + 8 <= b1; b += 8) {
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = hblk[(uint64_t)(b + u) * pc + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            hblk[(uint64_t)(b + u) * pc + c] = (uint32_t)run;
            run += v[u];
        }
    }
    for (; b < b1; ++b) {
        const uint32_t v = hblk[(uint64_t)b * pc + c];
        hblk[(uint64_t)b * pc + c] = (uint32_t)run;
        run += v;
    }
    if (r == 0) hist[c] = total;
}

// ---------------------------------------------------------------------------
// Two-level scatter.  Pass 1: a block owns the chunk of positions the
// classify pass counted per interval (so its run in every interval is known
// from the prefix over blocks) and writes (key, position) records into those
// runs.  Pass 2: one
// workgroup per coarse interval stages the interval's inner splitters in LDS,
// finds each record's bucket by binary lifting over them (kept as a u16 for
// the placing pass), histograms and scans the buckets, writes the bucket
// offsets, and places every record with LDS cursors.
// ---------------------------------------------------------------------------

template <typename S, typename I, bool SEQ, bool KEEP>
__global__ __launch_bounds__(256) void k_scatter1(const S *__restrict__ t, uint64_t n, CodeMap<S> cm, KeyGeom g,
                                                  DevLimits L, uint64_t max_ctx, const I *__restrict__ pos,
                                                  uint64_t n_pos, const uint16_t *__restrict__ cid, uint32_t pc,
                                                  const uint64_t *__restrict__ coff,
                                                  const uint32_t *__restrict__ hblk, KRec<I> *__restrict__ tmp,
                                                  uint64_t chunk, const uint64_t *__restrict__ keep, uint32_t c_lo,
                                                  uint32_t c_hi, uint64_t rbase, uint32_t hstride) {
    __shared__ S win[SEQ ? kSeqChunk + 72 : 1];
    // LDS sized by the launch (pc * 12 bytes): run cursors, then counts
    extern __shared__ unsigned long long scatter1_lds[];
    unsigned long long *cbase = scatter1_lds;
    uint32_t *lcnt = (uint32_t *)(scatter1_lds + pc);
    // chunk: a multiple of kSeqChunk, so key windows never straddle two blocks
    const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
    if (c0 >= n_pos) return;
    const uint64_t c1 = c0 + chunk < n_pos ? c0 + chunk : n_pos;
    // this block's runs: the classify pass counted its range per interval
    // (hstride consecutive classify blocks, whose runs in each interval are
    // consecutive); only intervals [c_lo, c_hi) (this batch) are written, at
    // offsets relative to the batch's first record rbase
    for (uint32_t c = threadIdx.x; c < pc; c += blockDim.x) {
        cbase[c] = coff[c] - rbase + hblk[(uint64_t)blockIdx.x * hstride * pc + c];
        lcnt[c] = 0;
    }
    __syncthreads();
    if (SEQ) {
        // the classify pass's layout: a thread owns 16 consecutive positions,
        // keeps their keys in registers and reads their 16 interval ids with
        // two 16-byte loads (32-byte aligned: base and first are multiples of 16)
        constexpr uint32_t kSkip = 0xFFFFFFFFu;  // not a record of this batch
        for (uint64_t base = c0; base < c1; base += kSeqChunk) {
            const uint32_t cnt = (uint32_t)(c1 - base < kSeqChunk ? c1 - base : kSeqChunk);
            uint64_t key[kSeqItems];
            thread_keys_seq<S>(t, n, cm, g, L, max_ctx, base, cnt, win, key);
            const uint32_t first = threadIdx.x * kSeqItems;
            uint32_t cc[kSeqItems];
            if (first + kSeqItems <= cnt) {
                const uint4 *src = (const uint4 *)(cid + base + first);
                const uint4 v0 = src[0], v1 = src[1];
                const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (uint32_t j = 0; j < kSeqItems; ++j) cc[j] = (w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            } else {
#pragma unroll
                for (uint32_t j = 0; j < kSeqItems; ++j) cc[j] = first + j < cnt ? (uint32_t)cid[base + first + j] : kSkip;
            }
            uint32_t kb = 0xFFFFu;
            if (KEEP && first < cnt) kb = (uint32_t)(keep[(base + first) >> 6] >> ((base + first) & 63)) & 0xFFFFu;
#pragma unroll
            for (uint32_t j = 0; j < kSeqItems; ++j) {
                const uint32_t c = cc[j];
                if (((kb >> j) & 1u) && c != kSkip && c - c_lo < c_hi - c_lo) {
                    const unsigned long long slot = cbase[c] + atomicAdd(&lcnt[c], 1u);
                    tmp[slot] = KRec<I>{key[j], (I)(base + first + j)};
                }
            }
        }
    } else {
        for (uint64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
            uint32_t c = cid[i];
            if (c < c_lo || c >= c_hi) continue;
            uint64_t p = pos ? (uint64_t)pos[i] : i;
            unsigned long long slot = cbase[c] + atomicAdd(&lcnt[c], 1u);
            uint64_t seg;
            uint64_t lim = lim_seg(L, p, &seg);
            tmp[slot] = KRec<I>{with_id(g, key_at_seed(t, cm, g, p, lim < max_ctx ? lim : max_ctx), seg), (I)p};
        }
    }
}

// Scatter pass 1 for contiguous positions (every suffix, or the kept ones
// of a bitmap) with coalesced writes: each kStageStep-position step of a
// block counting-sorts its records by interval in LDS, then consecutive
// threads write each interval's run of the step, so a wavefront's store
// covers a few contiguous runs instead of 64 scattered records.  For at most
// kStageIntervals intervals; the block owns the same range of positions as in
// k_scatter1.  One block of 1,024 threads per compute unit (136 KB of LDS
// with u8 text: keys, step-local positions, interval ids, interval tables)
// stages 8,192 positions a step, eight per thread: about eight records per
// interval and step over 1,024 intervals, instead of the four of
// 4,096-position steps at two blocks of 256 (and 1,024 threads hide the
// step's latencies better than 512 with sixteen positions each: 2.30 ->
// 2.11 ms on 268M positions).
constexpr uint32_t kStageIntervals = 2048;
constexpr uint32_t kStageStep = 8192;
constexpr uint32_t kStageNT = 1024;
constexpr uint32_t kStageItems = kStageStep / kStageNT;  // consecutive positions per thread
static_assert(kStageItems % 8 == 0 && kStageItems <= 32, "whole 16-byte interval-id loads, one keep word");
// scatter pass 1 workgroups per compute unit (staged / direct pass), each
// taking consecutive classify blocks
constexpr uint32_t kP1StagedPerCu = 4, kP1DirectPerCu = 2;

// Exclusive scan of one u32 per thread across a block of NT threads; wsum
// holds NT / 64 entries.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan_nt(uint32_t v, uint32_t *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int q = 0; q < NT / 64; ++q)
        if (q < w) before += wsum[q];
    __syncthreads();
    return before + x - v;
}

template <typename S, typename I, bool KEEP>
__global__ __launch_bounds__(kStageNT) void k_scatter1_staged(const S *__restrict__ t, uint64_t n, CodeMap<S> cm,
                                                              KeyGeom g, DevLimits L, uint64_t max_ctx,
                                                              const uint16_t *__restrict__ cid, uint32_t pc,
                                                              const uint64_t *__restrict__ coff,
                                                              const uint32_t *__restrict__ hblk,
                                                              KRec<I> *__restrict__ tmp, uint64_t chunk,
                                                              const uint64_t *__restrict__ keep, uint32_t c_lo,
                                                              uint32_t c_hi, uint64_t rbase, uint32_t hstride) {
    constexpr uint32_t E = kStageIntervals / kStageNT;  // intervals per thread in the scan
    constexpr uint32_t kSkip = 0xFFFFFFFFu;
    __shared__ S win[kStageStep + 72];
    __shared__ uint64_t s_key[kStageStep];
    __shared__ uint16_t s_li[kStageStep];  // position - step base
    __shared__ uint16_t s_c[kStageStep];
    __shared__ unsigned long long cbase[kStageIntervals];
    __shared__ uint32_t icnt[kStageIntervals], ioff[kStageIntervals];
    __shared__ uint32_t wsum[kStageNT / 64];
    const uint64_t c0 = (uint64_t)blockIdx.x * chunk;
    if (c0 >= n) return;
    const uint64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    for (uint32_t c = threadIdx.x; c < kStageIntervals; c += blockDim.x) {
        cbase[c] = c < pc ? coff[c] - rbase + hblk[(uint64_t)blockIdx.x * hstride * pc + c] : 0ull;
        icnt[c] = 0;
    }
    for (uint64_t base = c0; base < c1; base += kStageStep) {
        const uint32_t cnt = (uint32_t)(c1 - base < kStageStep ? c1 - base : kStageStep);
        constexpr uint32_t NI = kStageItems;
        uint64_t key[NI];
        thread_keys_seq<S, NI>(t, n, cm, g, L, max_ctx, base, cnt, win, key);  // begins with a barrier
        const uint32_t first = threadIdx.x * NI;
        uint32_t cc[NI], lr[NI];
        if (first + NI <= cnt) {
            const uint4 *src = (const uint4 *)(cid + base + first);
            uint32_t w[NI / 2];
#pragma unroll
            for (uint32_t q = 0; q < NI / 8; ++q) {
                const uint4 v = src[q];
                w[4 * q] = v.x, w[4 * q + 1] = v.y, w[4 * q + 2] = v.z, w[4 * q + 3] = v.w;
            }
#pragma unroll
            for (uint32_t j = 0; j < NI; ++j) cc[j] = (w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
        } else {
#pragma unroll
            for (uint32_t j = 0; j < NI; ++j) cc[j] = first + j < cnt ? (uint32_t)cid[base + first + j] : kSkip;
        }
        constexpr uint32_t kAll = (uint32_t)((1ull << NI) - 1);
        uint32_t kb = kAll;
        if (KEEP && first < cnt) kb = (uint32_t)(keep[(base + first) >> 6] >> ((base + first) & 63)) & kAll;
        // local rank of each record inside its interval's run of this step
#pragma unroll
        for (uint32_t j = 0; j < NI; ++j) {
            const uint32_t c = cc[j];
            const bool take = ((kb >> j) & 1u) && c != kSkip && c - c_lo < c_hi - c_lo;
            cc[j] = take ? c : kSkip;
            lr[j] = take ? atomicAdd(&icnt[c], 1u) : 0u;
        }
        __syncthreads();
        // interval offsets of the step: exclusive scan of icnt
        uint32_t v[E], sum = 0;
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) {
            v[e] = icnt[threadIdx.x * E + e];
            sum += v[e];
        }
        uint32_t pre = block_excl_scan_nt<kStageNT>(sum, wsum);
        uint32_t total = 0;
#pragma unroll
        for (uint32_t q = 0; q < kStageNT / 64; ++q) total += wsum[q];
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) {
            ioff[threadIdx.x * E + e] = pre;
            pre += v[e];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < NI; ++j)
            if (cc[j] != kSkip) {
                const uint32_t q = ioff[cc[j]] + lr[j];
                s_key[q] = key[j];
                s_li[q] = (uint16_t)(first + j);
                s_c[q] = (uint16_t)cc[j];
            }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < total; q += blockDim.x) {
            const uint32_t c = s_c[q];
            tmp[cbase[c] + (q - ioff[c])] = KRec<I>{s_key[q], (I)(base + s_li[q])};
        }
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < kStageIntervals; c += blockDim.x) {
            cbase[c] += icnt[c];
            icnt[c] = 0;
        }
    }
}

// Exclusive scan of cnt[0..nb) into out[0..nb) (+ base) across the block;
// each thread scans a contiguous slice, the slice sums are scanned with
// wavefront shuffles and one exchange of wavefront totals.  Returns the total
// in every thread.  part: blockDim.x / 64 u64 of LDS.
template <typename O>
__device__ __forceinline__ unsigned long long block_scan_slices(const uint32_t *cnt, uint32_t nb,
                                                                unsigned long long base, O *out,
                                                                unsigned long long *part) {
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    const uint32_t j0 = threadIdx.x * per, j1 = j0 + per < nb ? j0 + per : nb;
    unsigned long long sum = 0;
    for (uint32_t j = j0; j < j1; ++j) sum += cnt[j];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    unsigned long long x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) part[w] = x;
    __syncthreads();
    unsigned long long before = 0, total = 0;
    for (int q = 0; q < nw; ++q) {
        const unsigned long long v = part[q];
        before += q < w ? v : 0ull;
        total += v;
    }
    unsigned long long run = base + before + x - sum;
    for (uint32_t j = j0; j < j1; ++j) {
        out[j] = (O)run;
        run += cnt[j];
    }
    __syncthreads();  // part is reused by the next scan
    return total;
}

// LDS of scatter pass 2, sized by the launch for `stride` buckets per
// interval: the inner splitters as a breadth-first tree, then 32-bit bucket
// cursors (relative to the interval's first record: an interval holds fewer
// than 2^32 records), tile offsets and counts.
__host__ __device__ constexpr uint32_t scatter2_lds_bytes(uint32_t stride) {
    return pow2_at_least(stride) * 8 + stride * (4 + 4 + 4);
}

// NT threads and TILE records staged per placing step: 512 and 4,096 (two
// workgroups per compute unit), or 1,024 and 2,048 when the LDS of a wide
// interval (or a batch of few intervals) leaves one.  The bucket of a record
// is found twice, in the counting pass and again in the placing pass, instead
// of being stored between them: a binary search over the interval's inner
// splitters in LDS costs less than writing and reading back two bytes per
// record (3.18 against 3.48 ms on 268M records in 1,024 intervals; 4,096-record
// tiles instead of 2,048 took 3.82 to 3.48 ms: longer bucket runs per write).
template <typename I, int NT, int TILE, bool AOS = false>
__global__ __launch_bounds__(NT) void k_scatter2(const KRec<I> *__restrict__ tmp, const uint64_t *__restrict__ coff,
                                                  Splitters SP, uint32_t p, uint64_t *__restrict__ off,
                                                  uint64_t *__restrict__ keys, I *__restrict__ idx, uint32_t c_lo,
                                                  uint32_t c_hi, uint64_t rbase, KRec<I> *__restrict__ brec = nullptr) {
    constexpr int U = TILE / NT;
    static_assert(TILE % NT == 0 && NT % 64 == 0, "tile of whole rows");
    // LDS sized by the launch (scatter2_lds_bytes)
    extern __shared__ unsigned long long scatter2_lds[];
    __shared__ unsigned long long part[NT / 64];
    __shared__ uint64_t tkey[TILE];
    __shared__ I tidx[TILE];
    __shared__ uint16_t tf[TILE];
    const uint32_t p2 = pow2_at_least(SP.stride);
    uint64_t *sspl = (uint64_t *)scatter2_lds;
    uint32_t *gcur = (uint32_t *)(scatter2_lds + p2);
    uint32_t *lbase = gcur + SP.stride;
    uint32_t *cnt = lbase + SP.stride;
    // interval c of the batch [c_lo, c_hi); record offsets relative to rbase
    const uint32_t c = c_lo + blockIdx.x;
    const uint64_t s = coff[c] - rbase, e = coff[c + 1] - rbase;
    // buckets of interval c: b0 .. b0 + nb - 1; its inner splitters spl[b0 .. b0 + nb - 1)
    const uint32_t b0 = c * SP.stride;
    const uint32_t nb = SP.stride < p - b0 ? SP.stride : p - b0;
    const uint32_t ninner = nb - 1;
    uint32_t fsteps = 0;
    while ((1u << fsteps) <= ninner) ++fsteps;
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x) cnt[j] = 0;
    for (uint32_t i = threadIdx.x; i + 1 < (1u << fsteps); i += blockDim.x) {
        const uint32_t j = eyt_in_order(i + 1, fsteps);
        sspl[i] = j < ninner ? SP.spl[b0 + j] : ~0ull;
    }
    __syncthreads();
    // counting pass: bucket sizes
    for (uint64_t i0 = s; i0 < e; i0 += 8 * NT) {
        uint64_t kk[8];
        uint32_t f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t i = i0 + u * NT + threadIdx.x;
            kk[u] = i < e ? tmp[i].key : 0ull;
        }
        eyt_upper_rank<8>(sspl, fsteps, ninner, kk, f);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i0 + u * NT + threadIdx.x < e) atomicAdd(&cnt[f[u]], 1u);
    }
    __syncthreads();
    block_scan_slices(cnt, nb, 0, gcur, part);
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x) {
        off[b0 + j] = s + gcur[j];
        cnt[j] = 0;
    }
    if (c + 1 == c_hi && threadIdx.x == 0) off[b0 + nb] = e;  // the batch's end
    __syncthreads();
    // placing pass, one tile of TILE records at a time: bucket of every
    // record, count the tile's buckets, scan, stage the records in LDS
    // grouped by bucket, then write each bucket's run of the tile with
    // consecutive threads
    for (uint64_t t0 = s; t0 < e; t0 += TILE) {
        const uint32_t tl = (uint32_t)(e - t0 < TILE ? e - t0 : TILE);
        uint64_t kk[U];
        I ix[U];
        uint32_t f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t q = u * NT + threadIdx.x;
            const KRec<I> r = q < tl ? tmp[t0 + q] : KRec<I>{0ull, (I)0};
            kk[u] = r.key;
            ix[u] = r.idx;
        }
        eyt_upper_rank<U>(sspl, fsteps, ninner, kk, f);
        // the count's atomic returns each record's rank inside its bucket's
        // run of the tile
        uint32_t rk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) rk[u] = u * NT + threadIdx.x < tl ? atomicAdd(&cnt[f[u]], 1u) : 0u;
        __syncthreads();
        block_scan_slices(cnt, nb, 0, lbase, part);  // ends with a barrier
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (u * NT + threadIdx.x < tl) {
                const uint32_t slot = (uint32_t)lbase[f[u]] + rk[u];
                tkey[slot] = kk[u];
                tidx[slot] = ix[u];
                tf[slot] = (uint16_t)f[u];
            }
        }
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < tl; q += NT) {
            const uint32_t fb = tf[q];
            const uint64_t dst = s + gcur[fb] + (q - lbase[fb]);
            if constexpr (AOS) {
                brec[dst] = KRec<I>{tkey[q], tidx[q]};
            } else {
                keys[dst] = tkey[q];
                idx[dst] = tidx[q];
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x) {
            gcur[j] += cnt[j];
            cnt[j] = 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Tiles: every bucket is cut into tiles of at most T records.
// ---------------------------------------------------------------------------

struct Range {
    uint64_t start;
    uint64_t len;
};

// Cuts every bucket into tiles of at most T records and files each tile
// under its size class (tiles[cls * cap + slot]); buckets larger than T are
// also listed in `big` for the key merges.  trec[c] receives the records of
// class c (the tile sort's work).  A workgroup takes kEmitPer * 256
// consecutive buckets in two identical walks: the first counts each
// wavefront's tiles per class (ballots) and reserves the wavefront's slots in
// LDS, then one global reservation per class and workgroup; the second files
// the tiles into those slots.
constexpr uint32_t kEmitPer = 4;
__global__ __launch_bounds__(256) void k_tile_emit(const uint64_t *__restrict__ off, uint32_t p, uint32_t T,
                                                   Tile *__restrict__ tiles, uint64_t cap,
                                                   unsigned int *__restrict__ tcount, Range *__restrict__ big,
                                                   unsigned int *__restrict__ n_big,
                                                   unsigned int *__restrict__ max_bucket,
                                                   unsigned long long *__restrict__ trec) {
    __shared__ unsigned int s_cnt[kTileClasses], s_base[kTileClasses], s_max;
    __shared__ unsigned long long s_rec[kTileClasses];
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1;
    if (threadIdx.x < kTileClasses) {
        s_cnt[threadIdx.x] = 0;
        s_rec[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) s_max = 0;
    __syncthreads();
    const uint64_t b_first = (uint64_t)blockIdx.x * 256 * kEmitPer;
    uint32_t wcnt[kTileClasses] = {};
    unsigned long long recs[kTileClasses] = {};
    uint32_t mx = 0;
    for (uint32_t k = 0; k < kEmitPer; ++k) {
        const uint64_t b = b_first + k * 256 + threadIdx.x;
        const uint64_t s = b < p ? off[b] : 0, sz = b < p ? off[b + 1] - s : 0;
        const uint32_t szc = (uint32_t)(sz < 0xFFFFFFFFull ? sz : 0xFFFFFFFFull);
        mx = szc > mx ? szc : mx;
        for (uint64_t x = 0; __any(x < sz); x += T) {
            const bool have = x < sz;
            const uint32_t len = have ? (uint32_t)(sz - x < T ? sz - x : T) : 0u;
            const int cls = have ? tile_class(len) : -1;
#pragma unroll
            for (int c = 0; c < kTileClasses; ++c) {
                wcnt[c] += (uint32_t)__popcll(__ballot(cls == c));
                recs[c] += cls == c ? len : 0u;
            }
        }
    }
    // wavefront totals: largest bucket and records per class
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t v = __shfl_xor(mx, o);
        mx = v > mx ? v : mx;
#pragma unroll
        for (int c = 0; c < kTileClasses; ++c) recs[c] += __shfl_xor(recs[c], o);
    }
    uint32_t wbase[kTileClasses];
#pragma unroll
    for (int c = 0; c < kTileClasses; ++c) {
        uint32_t v = 0;
        if (lane == 0 && wcnt[c]) v = atomicAdd(&s_cnt[c], wcnt[c]);
        wbase[c] = __shfl(v, 0);
        if (lane == 0 && recs[c]) atomicAdd(&s_rec[c], recs[c]);
    }
    if (lane == 0 && mx) atomicMax(&s_max, mx);
    __syncthreads();
    if (threadIdx.x < kTileClasses) {
        const uint32_t c = threadIdx.x;
        s_base[c] = s_cnt[c] ? atomicAdd(&tcount[c], s_cnt[c]) : 0u;
        if (s_rec[c]) atomicAdd(&trec[c], s_rec[c]);
    }
    if (threadIdx.x == 0 && s_max) atomicMax(max_bucket, s_max);
    __syncthreads();
    uint32_t cur[kTileClasses];
#pragma unroll
    for (int c = 0; c < kTileClasses; ++c) cur[c] = s_base[c] + wbase[c];
    for (uint32_t k = 0; k < kEmitPer; ++k) {
        const uint64_t b = b_first + k * 256 + threadIdx.x;
        const uint64_t s = b < p ? off[b] : 0, sz = b < p ? off[b + 1] - s : 0;
        for (uint64_t x = 0; __any(x < sz); x += T) {
            const bool have = x < sz;
            Tile tl{};
            int cls = -1;
            if (have) {
                tl.start = s + x;
                tl.len = (uint32_t)(sz - x < T ? sz - x : T);
                tl.whole = sz <= T ? 1u : 0u;
                cls = tile_class(tl.len);
            }
#pragma unroll
            for (int c = 0; c < kTileClasses; ++c) {
                const uint64_t m = __ballot(cls == c);
                if (cls == c) tiles[(uint64_t)c * cap + cur[c] + __popcll(m & below)] = tl;
                cur[c] += (uint32_t)__popcll(m);
            }
        }
        if (b < p && sz > T) {
            const unsigned int q = atomicAdd(n_big, 1u);
            big[q].start = s;
            big[q].len = sz;
        }
    }
}

// ---------------------------------------------------------------------------
// Merge-path key merge of sorted runs of length `run` inside ranges
// (oversized buckets and the splitter sample).  Ties take the left run.
// ---------------------------------------------------------------------------
// Merge-path split of diagonal d between sorted runs A (nA) and B (nB):
// the number of A elements among the first d outputs, ties to A.
template <typename K>
__device__ __forceinline__ uint64_t merge_split(const K *A, uint64_t nA, const K *B, uint64_t nB, uint64_t d) {
    uint64_t lo = d > nB ? d - nB : 0, hi = d < nA ? d : nA;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (A[mid] <= B[d - mid - 1]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One merge level: every block writes kKeyMergeTile consecutive outputs of
// one pair of runs.  Two threads find the block's merge-path splits in global
// memory; the block stages its A and B segments in LDS with coalesced loads,
// each thread merges 8 outputs from LDS, and the outputs go back through LDS
// for coalesced stores.  Requires run to be a power of two >= half a tile, so
// a block's outputs never span two pairs of runs.
constexpr uint32_t kKeyMergeTile = 2048;

template <typename I>
__global__ __launch_bounds__(256) void k_key_merge(const Range *__restrict__ ranges, uint64_t run,
                                                   const uint64_t *__restrict__ ks, const I *__restrict__ is,
                                                   uint64_t *__restrict__ kd, I *__restrict__ id) {
    constexpr uint32_t ITEMS = kKeyMergeTile / 256;
    __shared__ uint64_t sk[kKeyMergeTile];
    __shared__ I si[kKeyMergeTile];
    __shared__ uint64_t split[2];
    const Range r = ranges[blockIdx.y];
    const uint64_t o0 = (uint64_t)blockIdx.x * kKeyMergeTile;
    if (o0 >= r.len) return;
    const uint32_t L = (uint32_t)(r.len - o0 < kKeyMergeTile ? r.len - o0 : kKeyMergeTile);
    const uint64_t base = (o0 / (2 * run)) * (2 * run);
    const uint64_t nA = r.len - base < run ? r.len - base : run;
    const uint64_t restB = r.len - base - nA;
    const uint64_t nB = restB < run ? restB : run;
    const uint64_t *A = ks + r.start + base;
    const uint64_t *B = A + nA;
    const I *Ai = is + r.start + base;
    const I *Bi = Ai + nA;
    const uint64_t d0 = o0 - base;
    if (threadIdx.x < 2) split[threadIdx.x] = merge_split(A, nA, B, nB, d0 + (threadIdx.x ? L : 0));
    __syncthreads();
    const uint64_t a0 = split[0], a1 = split[1];
    const uint32_t la = (uint32_t)(a1 - a0);
    const uint64_t b0 = d0 - a0;
    for (uint32_t q = threadIdx.x; q < L; q += 256) {
        if (q < la) {
            sk[q] = A[a0 + q];
            si[q] = Ai[a0 + q];
        } else {
            sk[q] = B[b0 + (q - la)];
            si[q] = Bi[b0 + (q - la)];
        }
    }
    __syncthreads();
    const uint32_t lb = L - la;
    const uint32_t d = threadIdx.x * ITEMS < L ? threadIdx.x * ITEMS : L;
    uint32_t i = (uint32_t)merge_split(sk, la, sk + la, lb, d), j = d - i;
    uint64_t ok[ITEMS];
    I oi[ITEMS];
#pragma unroll
    for (uint32_t q = 0; q < ITEMS; ++q) {
        const bool takeA = j >= lb || (i < la && sk[i] <= sk[la + j]);
        const uint32_t src = takeA ? i : la + j;
        ok[q] = sk[src < L ? src : 0];
        oi[q] = si[src < L ? src : 0];
        if (takeA) ++i;
        else ++j;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < ITEMS; ++q)
        if (d + q < L) {
            sk[d + q] = ok[q];
            si[d + q] = oi[q];
        }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < L; q += 256) {
        kd[r.start + o0 + q] = sk[q];
        id[r.start + o0 + q] = si[q];
    }
}

template <typename I>
__global__ __launch_bounds__(256) void k_copy_ranges(const Range *__restrict__ ranges, const uint64_t *__restrict__ ks,
                                                     const I *__restrict__ is, uint64_t *__restrict__ kd,
                                                     I *__restrict__ id) {
    Range r = ranges[blockIdx.y];
    for (uint64_t o = (uint64_t)blockIdx.x * 256 + threadIdx.x; o < r.len; o += (uint64_t)gridDim.x * 256) {
        kd[r.start + o] = ks[r.start + o];
        id[r.start + o] = is[r.start + o];
    }
}

// ---------------------------------------------------------------------------
// Group detection over the key-sorted records.  Between different keys the
// LCP follows from the keys (pack.rs:307-317); runs of equal keys are groups
// refined at depth k.
// ---------------------------------------------------------------------------
// LCPs across key boundaries, once refinement has fixed which member of a
// group sits next to the boundary: min(shared fields, both windows)
// (pack.rs:307-317 with max_context).
// has_prev: this is a later batch, whose first suffix follows the previous
// batch's last one (key prev_key, position prev_pos; the keys differ).  The
// previous suffix is passed by value: a per-lane choice between a pointer
// load and an argument leaves no load through an optional pointer.
// Escape geometry: offset of the first code >= esc among the `take` (<= 32)
// codes of suffix p, from the 4-bit packed code text; take when none.
__device__ __forceinline__ uint32_t first_escape(const uint64_t *__restrict__ pk, uint32_t lgf, uint64_t p,
                                                 uint32_t take, uint32_t esc) {
    if (take == 0) return 0;
    uint64_t hi, lo;
    load128_codes(pk, lgf, 4, p, p + take, hi, lo);
    // nibble >= esc (esc = 4: either of the nibble's two high bits)
    const uint64_t m = 0xCCCCCCCCCCCCCCCCull;
    (void)esc;
    const uint64_t xh = hi & m, xl = lo & m;
    const uint32_t f = xh ? (uint32_t)__builtin_clzll(xh) / 4 : xl ? 16 + (uint32_t)__builtin_clzll(xl) / 4 : 32u;
    return f < take ? f : take;
}

template <typename I>
__global__ __launch_bounds__(256) void k_boundary_lcp(const uint64_t *__restrict__ keys, const I *__restrict__ idx,
                                                      uint64_t N, DevLimits L, uint64_t max_ctx, KeyGeom g,
                                                      I *__restrict__ lcp, uint32_t has_prev, uint64_t prev_key,
                                                      uint64_t prev_pos, const uint64_t *__restrict__ pk,
                                                      uint32_t lgf) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        if (i == 0 && !has_prev) {
            lcp[0] = 0;
            continue;
        }
        const uint64_t k0 = i ? keys[i - 1] : prev_key, k1 = keys[i];
        if (k0 == k1) continue;
        const uint64_t p0 = i ? (uint64_t)idx[i - 1] : prev_pos;
        uint64_t v = shared_fields(g, k0, k1);
        const uint64_t p1 = (uint64_t)idx[i];
        uint64_t ca = cap_at(L, p0, max_ctx), cb = cap_at(L, p1, max_ctx);
        v = v < ca ? v : ca;
        v = v < cb ? v : cb;
        if (g.esc && v) {  // escape fields agree with T: only the fields before an escape are symbols
            const uint32_t ea = first_escape(pk, lgf, p0, (uint32_t)v, g.esc);
            const uint32_t eb = first_escape(pk, lgf, p1, (uint32_t)v, g.esc);
            v = ea < v ? ea : v;
            v = eb < v ? eb : v;
        }
        lcp[i] = (I)v;
    }
}

// ---------------------------------------------------------------------------
// Group refinement: a workgroup refines groups of suffixes known to agree on
// their first depth0 symbols (65..2048 members, or chunks of kGiantChunk
// members of larger groups) one at a time, in LDS arrays sized for M
// members: groups of up to 128 members go to 128-thread workgroups of 8.4 KB,
// up to 512 to 33.8 KB, up to 1,024 to 67.6 KB, larger ones to 135 KB
// (k_split_mid files them by size), so small groups keep many workgroups per
// compute unit.  The first kChunkRounds rounds sort each subgroup by the next k
// symbols at its depth (chunk key, plus the boundary value of members whose
// window ends inside the chunk), which splits short common prefixes into many
// classes at once.  Later rounds compare every unresolved member with the first
// member (in the current order) of its subgroup, the subgroup's reference:
// one LCP scan from the subgroup's depth D over the packed codes, as long as
// the windows allow, gives l, the offset where member and reference part.
// The comparator's outcome against the reference then fixes the member's
// class:
//   - the codes differ at D+l: the member lies before or after the reference
//     (by the codes), and ties among members are only possible with the same
//     l and the same code; such ties continue at depth D+l+1;
//   - the member's window ends first (at D+l): before the reference when
//     shorter suffixes come first, after it in the STAR order; settled;
//   - both windows end together: the reference's class, ordered by
//     boundary_then_pos (limit, then position); settled;
//   - the reference's window ends first: after the reference (shorter first)
//     or before it (STAR); such members continue together at depth D+l.
// On the "before" side members order by l ascending, on the "after" side by
// l descending (a member that parts from the reference later is closer to
// it), then by code.  So one round orders whole staircases of nested common
// prefixes (e.g. suffixes before runs of one symbol of different lengths),
// which rounds of k symbols would peel off one member at a time.  Sorted
// neighbours in different classes share D + min(l, l') symbols; the final
// LCP caps that by both windows.
// ---------------------------------------------------------------------------
constexpr uint64_t kLMax = (1ull << 44) - 1;
constexpr uint32_t kChunkRounds = 2;

// LCP of the codes of suffixes a and b from offset 0, at most `limit`.
template <typename S>
__device__ __forceinline__ uint64_t lcp_codes(const S *__restrict__ t, const CodeMap<S> &cm, const KeyGeom &g,
                                              uint64_t a, uint64_t b, uint64_t limit) {
    if (!cm.packed) return thread_lcp(t, a, b, 0, limit);
    const uint32_t bits = g.bits, F = 1u << cm.lgf;
    uint64_t o = 0;
    while (o < limit) {
        uint64_t va[2], vb[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            va[u] = vb[u] = 0;
            if (u == 1 && o + F >= limit) break;  // every word read holds a symbol inside the window
            const uint64_t qa = a + o + (uint64_t)u * F, qb = b + o + (uint64_t)u * F;
            const uint64_t wa = qa >> cm.lgf, wb = qb >> cm.lgf;
            const uint32_t ra = (uint32_t)(qa & (F - 1)) * bits, rb = (uint32_t)(qb & (F - 1)) * bits;
            va[u] = ra ? (cm.packed[wa] << ra) | (cm.packed[wa + 1] >> (64 - ra)) : cm.packed[wa];
            vb[u] = rb ? (cm.packed[wb] << rb) | (cm.packed[wb + 1] >> (64 - rb)) : cm.packed[wb];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint64_t x = va[u] ^ vb[u];
            if (x) {
                const uint64_t r = o + (uint64_t)u * F + (uint32_t)__builtin_clzll(x) / bits;
                return r < limit ? r : limit;
            }
        }
        // the first word matched and repeats one code: recorded runs of that
        // code under both suffixes agree on the shorter rest (run memo)
        if (cm.runs.n_runs && va[0] == ((va[0] << bits) | (va[0] >> (64 - bits)))) {
            const uint64_t ra = run_remaining(cm.runs, a + o), rb = run_remaining(cm.runs, b + o);
            const uint64_t sk = ra < rb ? ra : rb;
            if (sk > 2ull * F) {
                o += sk;
                continue;
            }
        }
        o += 2ull * F;
    }
    return limit;
}

// Run memo construction, pass 1: every first word of a stretch of aligned
// words that all repeat one code (and that the previous word does not
// continue) measures its stretch; stretches of kRunMinWords words or more are
// extended to the exact run (codes of that value in the neighbouring words)
// and appended, up to `cap` runs (a subset is still a correct memo).
__global__ void k_find_runs(const uint64_t *__restrict__ pk, uint64_t nwords, uint32_t lgf, uint64_t n,
                            uint64_t *__restrict__ rs, uint64_t *__restrict__ re, unsigned long long *__restrict__ count,
                            uint64_t cap) {
    const uint32_t bits = 64u >> lgf;
    const uint64_t F = 1ull << lgf;
    auto uniform = [&](uint64_t v) { return v == ((v << bits) | (v >> (64 - bits))); };
    // four consecutive words per thread and step, in two 16-byte loads
    for (uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; w0 < nwords;
         w0 += (uint64_t)gridDim.x * blockDim.x * 4) {
        uint64_t vw[4];
        if (w0 + 4 <= nwords) {
            const uint4 *p4 = (const uint4 *)(pk + w0);
            const uint4 a = p4[0], b = p4[1];
            vw[0] = (uint64_t)a.x | ((uint64_t)a.y << 32), vw[1] = (uint64_t)a.z | ((uint64_t)a.w << 32);
            vw[2] = (uint64_t)b.x | ((uint64_t)b.y << 32), vw[3] = (uint64_t)b.z | ((uint64_t)b.w << 32);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) vw[k] = w0 + k < nwords ? pk[w0 + k] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t w = w0 + (uint64_t)k;
            if (w >= nwords || (w << lgf) >= n) break;
            const uint64_t v = vw[k];
            if (!uniform(v)) continue;
            if (w > 0 && (k ? vw[k - 1] : pk[w - 1]) == v) continue;  // not the stretch's first word
            // the stretch's length, eight independent word loads a step (a
            // 10,000-symbol run is about 2,500 words)
            uint64_t m = 1;
            for (;;) {
                uint32_t eq = 0;
                bool on = true;
    #pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const uint64_t x = w + m + (uint64_t)q;
                    const bool same = ((x << lgf) < n) && pk[x] == v;
                    on = on && same;
                    eq += on ? 1u : 0u;
                }
                m += eq;
                if (eq < 8) break;
            }
            if (m < kRunMinWords) continue;
            const uint64_t code = v >> (64 - bits);
            // codes of the same value just before and just after the stretch
            uint64_t back = 0, fwd = 0;
            if (w > 0) {
                const uint64_t pv = pk[w - 1];
                while (back < F && ((pv >> (back * bits)) & ((1ull << bits) - 1)) == code) ++back;
            }
            if (((w + m) << lgf) < n) {
                const uint64_t nv = pk[w + m];
                while (fwd < F && ((nv >> (64 - (fwd + 1) * bits)) & ((1ull << bits) - 1)) == code) ++fwd;
            }
            const uint64_t s = (w << lgf) - back;
            uint64_t e = ((w + m) << lgf) + fwd;
            if (e > n) e = n;
            // one reservation per wavefront for the lanes that found a run
            // here (one counter for every run would serialize on it)
            const uint64_t here = __ballot(true);
            const int lane = threadIdx.x & 63, first = __ffsll((unsigned long long)here) - 1;
            uint64_t base = 0;
            if (lane == first) base = atomicAdd(count, (unsigned long long)__popcll(here));
            base = shfl64(base, first);
            const uint64_t slot = base + (uint64_t)__popcll(here & ((1ull << lane) - 1));
            if (slot < cap) {
                rs[slot] = s;
                re[slot] = e;
            }
        }
    }
}

// Segment directory: dir[b] = the number of segment ends at or below
// min(b << shift, n) (the first end past block b's start), for b in [0, nb].
__global__ void k_seg_dir(const uint64_t *__restrict__ ends, uint64_t ne, uint64_t n, uint32_t shift, uint64_t nb,
                          uint32_t *__restrict__ dir) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    const uint64_t x = (b << shift) < n ? (b << shift) : n;
    uint64_t lo = 0, hi = ne;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ends[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    dir[b] = (uint32_t)lo;
}

// Run memo construction, pass 2: dir[b] = the first run (sorted) whose end
// lies past block b's start, for b in [0, n_dir].
__global__ void k_run_dir(const uint64_t *__restrict__ re, uint64_t n_runs, uint64_t n_dir,
                          uint32_t *__restrict__ dir) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > n_dir) return;
    const uint64_t x = b << kRunDirShift;
    uint64_t lo = 0, hi = n_runs;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (re[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    dir[b] = (uint32_t)lo;
}

// Class keys that members can share and refine further: a differing code
// (symbol field 1..2^16), or "the reference's window ended first" (field 0 on
// the after side when shorter suffixes come first, field 0x1FFFF on the before
// side in the STAR order).  Window-ended and both-ended members never tie.
__device__ __forceinline__ bool group_key_continues(uint64_t k, bool star) {
    const uint64_t cls = k >> 62, sym = k & 0x1FFFFull;
    if (cls == 1) return false;
    if (sym != 0 && sym != 0x1FFFFull) return true;
    return star ? (cls == 0 && sym == 0x1FFFFull) : (cls == 2 && sym == 0);
}

template <typename S>
__device__ __forceinline__ uint32_t code_at(const S *__restrict__ t, const CodeMap<S> &cm, const KeyGeom &g,
                                            uint64_t q) {
    if (!cm.packed) return cm(t[q]);
    const uint64_t w = cm.packed[q >> cm.lgf];
    const uint32_t f = (uint32_t)(q & ((1ull << cm.lgf) - 1));
    return (uint32_t)(w >> (64 - (f + 1) * g.bits)) & ((1u << g.bits) - 1);
}

// Mid-size groups by LDS class: class c of k_split_mid holds sizes in
// (kMidClassMax[c-1], kMidClassMax[c]].
constexpr uint32_t kMidClasses = 4;
__host__ __device__ __forceinline__ uint32_t mid_class(uint64_t sz) {
    return sz <= 128 ? 0u : sz <= 512 ? 1u : sz <= 1024 ? 2u : 3u;
}

// Files the mid-size groups (65..2048 members) into the lists of their LDS
// class (out + c * cap, counts[c]; counts start at zero).  The order inside a
// list does not matter: every group is refined on its own.
__global__ __launch_bounds__(256) void k_split_mid(const Group *__restrict__ in, uint32_t n, Group *__restrict__ out,
                                                   uint64_t cap, unsigned int *__restrict__ counts) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool have = i < n;
    const Group gr = have ? in[i] : Group{0, 0};
    const uint32_t c = have ? mid_class(gr.size) : kMidClasses;
    // one reservation per class and wavefront (three counters for every
    // group would serialize on them)
#pragma unroll
    for (uint32_t k = 0; k < kMidClasses; ++k) {
        const uint64_t m = __ballot(c == k);
        if (!m) continue;
        const int first = __ffsll((unsigned long long)m) - 1;
        unsigned int base = 0;
        if (lane == first) base = atomicAdd(&counts[k], (unsigned int)__popcll(m));
        base = __shfl(base, first);
        if (c == k) out[k * cap + base + (unsigned int)__popcll(m & ((1ull << lane) - 1))] = gr;
    }
}

// groups[0 .. n), n = *n_dev when n_dev is set (a count written by an earlier
// kernel), else n_host; every group has at most M members.  With `next` (a
// counter that starts at zero) workgroups claim groups one at a time, so
// groups of very different cost balance over the grid; else workgroup b takes
// groups b, b + gridDim.x, ...
template <typename S, typename I, int B, int M>
__global__ __launch_bounds__(B) void k_groups(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                              uint64_t max_ctx, const Group *__restrict__ groups,
                                              const unsigned int *__restrict__ n_dev, uint32_t n_host,
                                              unsigned int *__restrict__ next, uint64_t depth0,
                                              I *__restrict__ idx, I *__restrict__ lcp, int run_mode) {
    static_assert((M & (M - 1)) == 0 && M <= 65536, "power-of-two member slots with u16 ranks");
    // member-indexed: position, window, class key, tie key, subgroup depth, l
    __shared__ uint64_t s_pos[M], s_cap[M], s_key[M], s_aux[M], s_dep[M], s_l[M];
    __shared__ uint64_t s_lcp[M];  // split depth of sorted neighbours r-1, r
    __shared__ uint32_t s_sub[M];  // subgroup (its first sorted position) | settled bit
    __shared__ uint16_t s_perm[M];
    __shared__ uint32_t s_scan[M];
    __shared__ uint32_t s_unres, s_gi;
    const uint32_t ng = n_dev ? *n_dev : n_host;
    const bool star = L.order == 1;
    uint64_t nsym = 0;  // symbols read
    for (uint32_t gi = blockIdx.x;; gi += gridDim.x) {
        __syncthreads();  // the previous group's last reads of the LDS arrays and of s_gi are done
        if (next) {
            if (threadIdx.x == 0) s_gi = atomicAdd(next, 1u);
            __syncthreads();
            gi = s_gi;
        }
        if (gi >= ng) break;
        const Group gr = groups[gi];
        const uint64_t gstart = gr.start;
        const uint32_t m = (uint32_t)gr.size;
        uint32_t P = 1;
        while (P < m) P <<= 1;
        for (uint32_t s = threadIdx.x; s < m; s += B) {
            uint64_t p = idx[gstart + s];
            uint64_t lim = lim_at(L, p);
            s_pos[s] = p;
            s_cap[s] = lim < max_ctx ? lim : max_ctx;
            s_sub[s] = 0;
            s_dep[s] = depth0;
            s_lcp[s] = 0;
        }
        for (uint32_t r = threadIdx.x; r < P; r += B) s_perm[r] = (uint16_t)r;
        __syncthreads();
        const uint64_t cont = star ? 0 : kAuxCont;
        for (uint32_t round = 0;; ++round) {
            // the first rounds split by chunk keys (k symbols at the subgroup's
            // depth: many classes at once when members part early); later rounds
            // by the reference's LCP (one round for nested common prefixes)
            const bool chunk_round = round < kChunkRounds;
            // (1) sort key of every unresolved member
            for (uint32_t s = threadIdx.x; s < m; s += B) {
                if (s_sub[s] & 0x80000000u) continue;  // settled
                const uint64_t D = s_dep[s];
                if (chunk_round) {
                    const uint64_t cap = s_cap[s];
                    s_key[s] = key_at(t, cm, g, s_pos[s], D, cap);
                    const uint64_t rem = cap > D ? cap - D : 0;
                    nsym += rem < g.k ? rem : g.k;
                    uint64_t aux = cont;
                    if (rem < g.k) {
                        const uint64_t lm = lim_at(L, s_pos[s]);
                        aux = star ? (1 + (kAuxCont - 1 - lm)) : lm;
                    }
                    s_aux[s] = aux;
                    continue;
                }
                const uint32_t ref = s_perm[s_sub[s]];
                const uint64_t pa = s_pos[s], pr = s_pos[ref];
                const uint64_t cs = s_cap[s], cref = s_cap[ref];
                const uint64_t cmin = cs < cref ? cs : cref;
                const uint64_t lim2 = cmin > D ? cmin - D : 0;  // common window from depth D
                const uint64_t l = s == ref ? lim2 : lcp_codes(t, cm, g, pa + D, pr + D, lim2);
                if (s != ref) nsym += 2 * (l < lim2 ? l + 1 : l);
                // A: absolute depth where the member parts from the reference (or
                // its window ends, possibly above D); classes order by it
                uint64_t key, aux = 0;
                if (l < lim2) {  // codes differ
                    const uint64_t A = D + l;
                    const uint32_t sa = code_at(t, cm, g, pa + A), sr = code_at(t, cm, g, pr + A);
                    key = sa < sr ? ((A << 17) | (sa + 1)) : ((2ull << 62) | ((kLMax - A) << 17) | (sa + 1));
                } else if (cs < cref) {  // the member's window ends first: settled
                    const uint64_t A = cs;
                    key = star ? ((2ull << 62) | ((kLMax - A) << 17) | 0x1FFFFull) : ((A << 17) | 0u);
                    const uint64_t lm = lim_at(L, pa);
                    aux = star ? ~lm : lm;
                } else if (cs == cref) {  // both end here: the boundary order decides; settled
                    key = 1ull << 62;
                    const uint64_t lm = lim_at(L, pa);
                    aux = star ? ~lm : lm;
                } else {  // the reference's window ends first: these continue together
                    const uint64_t A = D + l;
                    key = star ? ((A << 17) | 0x1FFFFull) : ((2ull << 62) | ((kLMax - A) << 17) | 0u);
                }
                s_key[s] = key;
                s_aux[s] = aux;
                s_l[s] = l;
            }
            __syncthreads();
            // (2) bitonic sort of the permutation by (sub, key, aux, pos)
            for (uint32_t size = 2; size <= P; size <<= 1) {
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t q = threadIdx.x; q < P / 2; q += B) {
                        uint32_t i = 2 * q - (q & (stride - 1));
                        uint32_t j = i + stride;
                        bool up = (i & size) == 0;
                        uint32_t a = s_perm[i], b = s_perm[j];
                        bool gt;
                        if (a >= m) gt = true;
                        else if (b >= m) gt = false;
                        else {
                            uint32_t sa = s_sub[a] & 0x7FFFFFFFu, sb = s_sub[b] & 0x7FFFFFFFu;
                            if (sa != sb) gt = sa > sb;
                            else {
                                // settled members keep their place: they are alone in their subgroup
                                uint64_t ka = (s_sub[a] & 0x80000000u) ? 0 : s_key[a];
                                uint64_t kb = (s_sub[b] & 0x80000000u) ? 0 : s_key[b];
                                uint64_t xa = (s_sub[a] & 0x80000000u) ? 0 : s_aux[a];
                                uint64_t xb = (s_sub[b] & 0x80000000u) ? 0 : s_aux[b];
                                if (ka != kb) gt = ka > kb;
                                else if (xa != xb) gt = xa > xb;
                                else gt = s_pos[a] > s_pos[b];
                            }
                        }
                        if (gt == up) {
                            s_perm[i] = (uint16_t)b;
                            s_perm[j] = (uint16_t)a;
                        }
                    }
                    __syncthreads();
                }
            }
            // (3) new subgroups: a member starts one unless it continues a tie
            // (same old subgroup, same class key, a class that can tie)
            for (uint32_t r = threadIdx.x; r < m; r += B) {
                uint32_t s = s_perm[r];
                bool f = true;
                if (r > 0 && !(s_sub[s] & 0x80000000u)) {
                    uint32_t q = s_perm[r - 1];
                    bool same_old = (s_sub[s] & 0x7FFFFFFFu) == (s_sub[q] & 0x7FFFFFFFu);
                    if (same_old && chunk_round) {
                        s_lcp[r] = s_dep[s] + shared_fields(g, s_key[q], s_key[s]);
                        f = !(s_key[s] == s_key[q] && s_aux[s] == s_aux[q]);
                    } else if (same_old) {
                        const uint64_t ls = s_l[s], lq = s_l[q];
                        s_lcp[r] = s_dep[s] + (ls < lq ? ls : lq);
                        const uint64_t k = s_key[s];
                        f = !(group_key_continues(k, star) && k == s_key[q]);
                    }
                }
                s_scan[r] = f ? r : 0;
            }
            __syncthreads();
            for (uint32_t off = 1; off < m; off <<= 1) {
                uint32_t v[(M + B - 1) / B];
                uint32_t c = 0;
                for (uint32_t r = threadIdx.x; r < m; r += B, ++c)
                    v[c] = (r >= off && s_scan[r - off] > s_scan[r]) ? s_scan[r - off] : s_scan[r];
                __syncthreads();
                c = 0;
                for (uint32_t r = threadIdx.x; r < m; r += B, ++c) s_scan[r] = v[c];
                __syncthreads();
            }
            // (4) settle singletons and ended members; ties move to their new depth
            if (threadIdx.x == 0) s_unres = 0;
            __syncthreads();
            uint32_t unres = 0;
            for (uint32_t r = threadIdx.x; r < m; r += B) {
                uint32_t s = s_perm[r];
                uint32_t ns = s_scan[r];
                bool has_nb = (r + 1 < m && s_scan[r + 1] == ns) || (r > 0 && s_scan[r - 1] == ns);
                bool settled = (s_sub[s] & 0x80000000u) || !has_nb;
                // chunk rounds: members whose window ended inside the chunk are
                // comparator-Equal (position order) with their tie
                if (chunk_round && !settled && s_aux[s] != cont) settled = true;
                if (!settled) {
                    if (chunk_round) {
                        s_dep[s] = s_dep[s] + g.k;
                    } else {  // a tie: past the differing code, or past the reference's end
                        const uint64_t k = s_key[s];
                        const uint64_t sym = k & 0x1FFFFull;
                        s_dep[s] = s_dep[s] + s_l[s] + (sym != 0 && sym != 0x1FFFFull ? 1 : 0);
                    }
                }
                s_sub[s] = ns | (settled ? 0x80000000u : 0u);
                unres += settled ? 0 : 1;
            }
            if (unres) atomicAdd(&s_unres, unres);
            __syncthreads();
            if (s_unres == 0) break;
            __syncthreads();
        }
        for (uint32_t r = threadIdx.x; r < m; r += B) {
            uint32_t s = s_perm[r];
            idx[gstart + r] = (I)s_pos[s];
            if (r > 0) {
                if (lcp) {
                    uint32_t q = s_perm[r - 1];
                    uint64_t v = s_lcp[r];
                    v = v < s_cap[s] ? v : s_cap[s];
                    v = v < s_cap[q] ? v : s_cap[q];
                    lcp[gstart + r] = (I)v;
                }
            } else if (run_mode && lcp) {
                lcp[gstart] = 0;
            }
        }
    }
    count_symbols(2, nsym);
}

// Copies the (idx, lcp) ranges of a list of chunks back from the ping-pong
// buffers (one block per chunk).
template <typename I>
__global__ __launch_bounds__(256) void k_copy_chunks(const Group *__restrict__ chunks, const I *__restrict__ si,
                                                     const I *__restrict__ sl, I *__restrict__ di,
                                                     I *__restrict__ dl) {
    const Group c = chunks[blockIdx.x];
    for (uint64_t q = threadIdx.x; q < c.size; q += 256) {
        di[c.start + q] = si[c.start + q];
        dl[c.start + q] = sl[c.start + q];
    }
}

// ---------------------------------------------------------------------------
// Merge-path LCP merge (sample_sort.rs:397-522 restated for a partitioned
// GPU merge).  Each thread owns W consecutive outputs of one merge job: a
// merge-path search with the full comparator finds its (i, j) split, one
// explicit LCP rebuilds the merge state m = LCP(last output, other head),
// and the three-case LCP merge runs sequentially from there.  Every output
// LCP is the true LCP of neighbours, so the result is identical to the
// sequential merge.
// ---------------------------------------------------------------------------
struct MergeJob {
    uint64_t start;  // x at start, y at start + nx, output at start
    uint64_t nx, ny;
};

template <typename S, typename I, int W>
__global__ __launch_bounds__(128) void k_lcp_merge(const S *__restrict__ t, CodeMap<S> cm, DevLimits L,
                                                   uint64_t max_ctx, uint64_t known,
                                                   const MergeJob *__restrict__ jobs, const I *__restrict__ si,
                                                   const I *__restrict__ sl, I *__restrict__ di,
                                                   I *__restrict__ dl) {
    // suffix compares scan the packed codes when the build has them
    auto cmp = [&](uint64_t a, uint64_t b, uint64_t kn, uint64_t *l) {
        return cm.packed ? thread_cmp_packed(cm, L, max_ctx, a, b, kn, l) : thread_cmp(t, L, max_ctx, a, b, kn, l);
    };
    MergeJob jb = jobs[blockIdx.y];
    uint64_t rlen = jb.nx + jb.ny;
    uint64_t d = ((uint64_t)blockIdx.x * 128 + threadIdx.x) * W;
    if (d >= rlen) return;
    const uint64_t nX = jb.nx, nY = jb.ny;
    const I *X = si + jb.start, *Y = X + nX;
    const I *LX = sl + jb.start, *LY = LX + nX;
    uint64_t cnt = rlen - d < W ? rlen - d : W;
    I *oz = di + jb.start + d;
    I *ol = dl + jb.start + d;
    if (nY == 0 || nX == 0) {  // a lone run is carried over
        const I *Z = nX ? X : Y, *LZ = nX ? LX : LY;
        for (uint64_t q = 0; q < cnt; ++q) {
            oz[q] = Z[d + q];
            ol[q] = LZ[d + q];
        }
        return;
    }
    uint64_t lo = d > nY ? d - nY : 0, hi = d < nX ? d : nX;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        uint64_t l;
        if (cmp((uint64_t)X[mid], (uint64_t)Y[d - mid - 1], known, &l) < 0) lo = mid + 1;
        else hi = mid;
    }
    uint64_t i = lo, j = d - lo;
    // A = stream of the last output, B = the other; m = LCP(last output, B head)
    const I *A = X, *Bs = Y, *LA = LX, *LB = LY;
    uint64_t iA = i, iB = j, nA = nX, nB = nY;
    uint64_t m = 0;
    if (d > 0) {
        bool lastX;
        if (i == 0) lastX = false;
        else if (j == 0) lastX = true;
        else {
            uint64_t l;
            lastX = cmp((uint64_t)X[i - 1], (uint64_t)Y[j - 1], known, &l) > 0;
        }
        if (!lastX) {
            A = Y, Bs = X, LA = LY, LB = LX;
            iA = j, iB = i, nA = nY, nB = nX;
        }
        if (iB < nB) {
            uint64_t l;
            cmp((uint64_t)A[iA - 1], (uint64_t)Bs[iB], known, &l);
            m = l;
        }
    }
    for (uint64_t q = 0; q < cnt; ++q) {
        bool hasA = iA < nA, hasB = iB < nB;
        bool outA;
        uint64_t lo_out, new_m = m;
        if (hasA && hasB) {
            uint64_t la = (d + q == 0) ? 0 : (uint64_t)LA[iA];
            if (la > m) {
                outA = true, lo_out = la;
            } else if (la < m) {
                outA = false, lo_out = m, new_m = la;
            } else {
                uint64_t l;
                int c = cmp((uint64_t)A[iA], (uint64_t)Bs[iB], m > known ? m : known, &l);
                outA = c < 0, lo_out = m, new_m = l;
            }
        } else if (hasA) {
            outA = true, lo_out = (uint64_t)LA[iA];
        } else {
            outA = false, lo_out = m;
        }
        if (outA) {
            oz[q] = A[iA];
            ol[q] = (I)lo_out;
            ++iA;
        } else {
            oz[q] = Bs[iB];
            ol[q] = (I)lo_out;
            ++iB;
            const I *tp;
            tp = A, A = Bs, Bs = tp;
            tp = LA, LA = LB, LB = tp;
            uint64_t u;
            u = iA, iA = iB, iB = u;
            u = nA, nA = nB, nB = u;
        }
        m = new_m;
    }
}

template <typename I>
__global__ __launch_bounds__(256) void k_copy_idx_lcp(uint64_t start, uint64_t len, const I *__restrict__ si,
                                                      const I *__restrict__ sl, I *__restrict__ di,
                                                      I *__restrict__ dl) {
    for (uint64_t o = (uint64_t)blockIdx.x * 256 + threadIdx.x; o < len; o += (uint64_t)gridDim.x * 256) {
        di[start + o] = si[start + o];
        dl[start + o] = sl[start + o];
    }
}

// ---------------------------------------------------------------------------
// Kernels exposed for parity tests.
// ---------------------------------------------------------------------------
// LcpDispatch::lcp (lcp.rs:125-153): window min(n-p, n-q, max_ctx).
template <typename S>
__global__ __launch_bounds__(256) void k_lcp_pairs(const S *__restrict__ t, uint64_t n, const uint64_t *__restrict__ pq,
                                                   uint64_t count, uint64_t max_ctx, uint64_t *__restrict__ out) {
    uint64_t w = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= count) return;
    uint64_t p = pq[2 * w], q = pq[2 * w + 1];
    uint64_t lp = n > p ? n - p : 0, lq = n > q ? n - q : 0;
    uint64_t cap = lp < lq ? lp : lq;
    cap = cap < max_ctx ? cap : max_ctx;
    uint64_t r = wave_lcp(t, p, q, 0, cap);
    if ((threadIdx.x & 63) == 0) out[w] = r;
}

// LcpDispatch::suffix_cmp_with (lcp.rs:184-201), boundary_order without the
// position tie for the default convention.
template <typename S>
__global__ __launch_bounds__(256) void k_cmp_pairs(const S *__restrict__ t, DevLimits L, const uint64_t *__restrict__ pq,
                                                   uint64_t count, uint64_t max_ctx, int32_t *__restrict__ out) {
    uint64_t w = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= count) return;
    uint64_t a = pq[2 * w], b = pq[2 * w + 1];
    uint64_t la = lim_at(L, a), lb = lim_at(L, b);
    uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
    uint64_t cap = ca < cb ? ca : cb;
    uint64_t c = wave_lcp(t, a, b, 0, cap);
    if ((threadIdx.x & 63) != 0) return;
    int r;
    if (c < cap) r = t[a + c] < t[b + c] ? -1 : (t[a + c] > t[b + c] ? 1 : 0);
    else if (L.order == 1) r = la != lb ? (lb < la ? -1 : 1) : (a < b ? -1 : (a > b ? 1 : 0));
    else r = la < lb ? -1 : (la > lb ? 1 : 0);
    out[w] = r;
}

template <typename S>
__global__ __launch_bounds__(256) void k_seed_keys(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                   const uint64_t *__restrict__ pos, uint64_t count,
                                                   uint64_t *__restrict__ keys) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t p = pos[i];
    keys[i] = key_at(t, cm, g, p, 0, lim_at(L, p));
}

// Writes positions 0..n-1 (or a copy of the caller's positions).
template <typename I>
__global__ __launch_bounds__(256) void k_iota(I *__restrict__ out, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (I)i;
}

}  // namespace caps
