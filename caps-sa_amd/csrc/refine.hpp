// Tie-group bookkeeping and the register-resident refinement kernels.
//
// A tie group is a run of records with equal packed keys.  Groups are binned
// by size: two members (one 16-lane LCP compare), 3-4, 5-8, 9-16, 17-32 and
// 33-64 members (one 4..64-lane slice of a wavefront per group, sorted with
// shuffles, no LDS), 65-2048 (one workgroup, LDS) and larger (chunked
// refinement + merge-path LCP merges).
#pragma once
#include "device.hpp"

namespace caps {

struct Group {
    uint64_t start;
    uint64_t size;
};

// A record between the two scatter passes (and, from scatter pass 2 to the
// tile sort, a bucketed record): key and index, 16 bytes with a u64 index;
// with a u32 index packed to 12 bytes (4-byte aligned), so both passes move
// a quarter fewer bytes.
template <typename I>
struct KRec {
    uint64_t key;
    I idx;
};
template <>
struct __attribute__((packed, aligned(4))) KRec<uint32_t> {
    uint64_t key;
    uint32_t idx;
};
static_assert(sizeof(KRec<uint32_t>) == 12 && alignof(KRec<uint32_t>) == 4, "packed u32 record");
static_assert(sizeof(KRec<uint64_t>) == 16, "u64 record");

constexpr int kGroupClasses = 8;
constexpr uint32_t kMidMax = 2048;
// Groups larger than kMidMax are sorted in chunks of kGiantChunk members
// (34 KB of LDS: four workgroups per compute unit), then merged by LCP
// merge levels: whole genome giant-group phase 31.1 -> 26.9 ms against
// chunks of 2,048 (135 KB, one workgroup) and two merge levels fewer.
constexpr uint32_t kGiantChunk = 512;

__host__ __device__ __forceinline__ int group_class(uint64_t sz) {
    return sz <= 2 ? 0 : sz <= 4 ? 1 : sz <= 8 ? 2 : sz <= 16 ? 3 : sz <= 32 ? 4 : sz <= 64 ? 5 : sz <= kMidMax ? 6 : 7;
}

// Smallest group size of each class (list capacities are n / min_size).
__host__ __device__ __forceinline__ uint64_t group_class_min(int c) {
    const uint64_t m[kGroupClasses] = {2, 3, 5, 9, 17, 33, 65, kMidMax + 1};
    return m[c];
}

struct GroupLists {
    Group *list[kGroupClasses];
    unsigned int *counts;  // kGroupClasses counters
};

// ---------------------------------------------------------------------------
// 64-bit shuffles (two 32-bit ds_bpermute lanes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl(lo, src);
    hi = __shfl(hi, src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, m);
    hi = __shfl_xor(hi, m);
    return ((uint64_t)hi << 32) | lo;
}

// Lane moves inside a slot of W lanes (W a power of two, slots aligned).
// A slot of W <= 16 lanes lies inside one row of 16, so DPP (a modifier on a
// VALU move, no round trip through the LDS crossbar) does them: quad
// permutations for xor 1 and xor 2, row shifts for lane -1, +1 and -o.  Other
// moves, and wider slots, use ds_bpermute.  seg_prev / seg_next / seg_up
// return the lane's own value where the source lane is outside the row;
// callers use them only where the source lies inside the slot.
template <int W>
__device__ __forceinline__ uint32_t seg_xor(uint32_t v, int stride) {
    if (W <= 16 && stride == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    if (W <= 16 && stride == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    return (uint32_t)__shfl_xor((int)v, stride);
}
template <int W>
__device__ __forceinline__ uint32_t seg_prev(uint32_t v) {
    if (W <= 16) return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xF, 0xF, false);
    const int lane = threadIdx.x & 63;
    return (uint32_t)__shfl((int)v, lane - 1 < 0 ? 0 : lane - 1);
}
template <int W>
__device__ __forceinline__ uint32_t seg_next(uint32_t v) {
    if (W <= 16) return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x101, 0xF, 0xF, false);
    const int lane = threadIdx.x & 63;
    return (uint32_t)__shfl((int)v, lane + 1 < 64 ? lane + 1 : lane);
}
template <int W>
__device__ __forceinline__ uint32_t seg_up(uint32_t v, int o) {
    if (W <= 16) {
        switch (o) {
            case 1: return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xF, 0xF, false);
            case 2: return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x112, 0xF, 0xF, false);
            case 4: return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x114, 0xF, 0xF, false);
            case 8: return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x118, 0xF, 0xF, false);
            default: break;
        }
    }
    return (uint32_t)__shfl_up((int)v, o, W);
}
template <int W>
__device__ __forceinline__ uint64_t seg_xor64(uint64_t v, int stride) {
    return ((uint64_t)seg_xor<W>((uint32_t)(v >> 32), stride) << 32) | seg_xor<W>((uint32_t)v, stride);
}
template <int W>
__device__ __forceinline__ uint64_t seg_prev64(uint64_t v) {
    return ((uint64_t)seg_prev<W>((uint32_t)(v >> 32)) << 32) | seg_prev<W>((uint32_t)v);
}

// Two-member groups are listed with their positions, so the pair kernel reads
// the text without first gathering the positions from the suffix array.
template <typename I>
struct PairRec {
    uint64_t start;
    I a, b;
};

// ---------------------------------------------------------------------------
// Group detection inside key-sorted ranges (buckets larger than one tile,
// after their tiles were merged).  Appends are aggregated per wavefront.
// ---------------------------------------------------------------------------
template <typename I>
__device__ __forceinline__ void wave_append(GroupLists &gl, bool have, int cls, Group gr, const I *__restrict__ idx) {
    for (int c = 0; c < kGroupClasses; ++c) {
        bool mine = have && cls == c;
        uint64_t mask = __ballot(mine);
        if (!mask) continue;
        uint32_t lane = threadIdx.x & 63;
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&gl.counts[c], (unsigned)__popcll(mask));
        base = __shfl(base, leader);
        if (mine) {
            uint64_t below = mask & ((1ull << lane) - 1);
            uint32_t slot = base + __popcll(below);
            if (c == 0) ((PairRec<I> *)gl.list[0])[slot] = PairRec<I>{gr.start, idx[gr.start], idx[gr.start + 1]};
            else gl.list[c][slot] = gr;
        }
    }
}

struct Range;  // defined in kernels.hpp

template <typename RangeT, typename I>
__global__ __launch_bounds__(256) void k_detect_ranges(const uint64_t *__restrict__ keys, const RangeT *__restrict__ ranges,
                                                       GroupLists gl, const I *__restrict__ idx) {
    RangeT r = ranges[blockIdx.y];
    for (uint64_t base = (uint64_t)blockIdx.x * 256; base < r.len; base += (uint64_t)gridDim.x * 256) {
        uint64_t o = base + threadIdx.x;
        bool have = false;
        Group gr{0, 0};
        if (o < r.len) {
            uint64_t i = r.start + o;
            uint64_t k0 = keys[i];
            bool start = o == 0 || keys[i - 1] != k0;
            if (start && o + 1 < r.len && keys[i + 1] == k0) {
                // the range is sorted, so the group ends at the first larger
                // key: gallop, then bisect (lo: equal, hi: larger or r.len)
                uint64_t lo = o + 1, hi = r.len;
                for (uint64_t step = 1;; step <<= 1) {
                    const uint64_t j = o + 1 + step;
                    if (j >= r.len) break;
                    if (keys[r.start + j] != k0) {
                        hi = j;
                        break;
                    }
                    lo = j;
                }
                while (hi - lo > 1) {
                    const uint64_t mid = lo + ((hi - lo) >> 1);
                    if (keys[r.start + mid] == k0) lo = mid;
                    else hi = mid;
                }
                gr = Group{i, hi - o};
                have = true;
            }
        }
        wave_append<I>(gl, have, have ? group_class(gr.size) : 0, gr, idx);
    }
}

// ---------------------------------------------------------------------------
// Two-member groups: four lanes per pair, sixteen pairs per wavefront.  Each
// lane compares 16 bytes of both suffixes per step (three aligned 8-byte
// loads per suffix, funnel-shifted), so a pair advances 64 bytes per step;
// a ballot over the pair's four lanes and ffs find the first differing byte
// (lcp.rs:292-360 as a lane-parallel compare).  Bytes past the comparison
// window count as a mismatch at the window's end; no load touches a word
// without an in-window byte.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load16(uint64_t x, uint64_t end, uint64_t &lo, uint64_t &hi) {
    // x, end: absolute byte addresses; words are aligned on the address, so a
    // word that holds one in-window byte never crosses into another page
    uint64_t a8 = x & ~7ull;
    uint32_t sh = (uint32_t)(x & 7) * 8;
    uint64_t w0 = a8 < end ? *(const uint64_t *)a8 : 0ull;
    uint64_t w1 = a8 + 8 < end ? *(const uint64_t *)(a8 + 8) : 0ull;
    uint64_t w2 = a8 + 16 < end ? *(const uint64_t *)(a8 + 16) : 0ull;
    lo = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
    hi = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
}

template <typename S, typename I>
__global__ __launch_bounds__(256) void k_pairs4(const S *__restrict__ t, DevLimits L, uint64_t max_ctx,
                                                const PairRec<I> *__restrict__ pairs, uint32_t n_pairs, uint64_t known,
                                                I *__restrict__ idx, I *__restrict__ lcp) {
    constexpr uint64_t SZ = sizeof(S);
    const uint32_t lane = threadIdx.x & 63, grp = lane >> 2, sl = lane & 3;
    const uint32_t pid = blockIdx.x * 64 + (threadIdx.x >> 2);
    const bool valid = pid < n_pairs;
    const uint64_t tb = (uint64_t)(uintptr_t)t;
    uint64_t s = 0, a = 0, b = 0, la = 0, lb = 0, cap = 0;
    if (valid) {
        PairRec<I> pr = pairs[pid];
        s = pr.start;
        a = pr.a;
        b = pr.b;
        la = lim_at(L, a);
        lb = lim_at(L, b);
        uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
        cap = ca < cb ? ca : cb;
    }
    const uint64_t capb = cap * SZ;
    const uint64_t enda = tb + (a + cap) * SZ, endb = tb + (b + cap) * SZ;
    uint64_t ob = (known < cap ? known : cap) * SZ;  // bytes known equal
    bool done = !valid || ob >= capb;
    uint64_t resb = capb;
    // first step: only the pair's first lane compares 16 bytes (one cache
    // line per suffix in most cases; most pairs part there); then all four
    // lanes, 64 bytes per step
    if (!done && sl == 0) {
        uint64_t alo, ahi, blo, bhi;
        load16(tb + a * SZ + ob, enda, alo, ahi);
        load16(tb + b * SZ + ob, endb, blo, bhi);
        uint64_t d0 = alo ^ blo, d1 = ahi ^ bhi;
        uint32_t fb = d0 ? (uint32_t)__builtin_ctzll(d0) >> 3 : d1 ? 8 + ((uint32_t)__builtin_ctzll(d1) >> 3) : 16u;
        uint64_t in_win = capb - ob;
        if (in_win < fb) fb = (uint32_t)in_win;
        if (fb < 16) {
            resb = ob + fb;
            done = true;
        }
    }
    done = __shfl(done ? 1 : 0, (int)(lane & ~3u)) != 0;
    resb = shfl64(resb, (int)(lane & ~3u));
    if (!done) ob += 16;
    while (__any(!done)) {
        uint32_t fb = 16;
        if (!done) {
            uint64_t rel = ob + 16 * sl;
            uint64_t alo, ahi, blo, bhi;
            load16(tb + a * SZ + rel, enda, alo, ahi);
            load16(tb + b * SZ + rel, endb, blo, bhi);
            uint64_t d0 = alo ^ blo, d1 = ahi ^ bhi;
            fb = d0 ? (uint32_t)__builtin_ctzll(d0) >> 3 : d1 ? 8 + ((uint32_t)__builtin_ctzll(d1) >> 3) : 16u;
            uint64_t in_win = capb > rel ? capb - rel : 0;
            if (in_win < fb) fb = (uint32_t)in_win;
        }
        uint64_t mask = __ballot(!done && fb < 16);
        uint32_t gm = (uint32_t)(mask >> (grp * 4)) & 0xFu;
        int fl = gm ? __ffs(gm) - 1 : 0;
        uint32_t fq = __shfl(fb, (int)(grp * 4 + fl));
        if (!done) {
            if (gm) {
                uint64_t r = ob + 16 * (uint64_t)fl + fq;
                resb = r < capb ? r : capb;
                done = true;
            } else {
                ob += 64;
            }
        }
    }
    if (valid && sl == 0) {
        uint64_t res = resb / SZ;
        int c;
        if (res < cap) c = t[a + res] < t[b + res] ? -1 : 1;
        else c = boundary_then_pos(L, a, la, b, lb);
        if (c > 0) {
            idx[s] = (I)b;
            idx[s + 1] = (I)a;
        }
        if (lcp) lcp[s + 1] = (I)res;
    }
}

// ---------------------------------------------------------------------------
// Groups of 3..64 members: W lanes per group (W = 4..64).  Each round every
// unresolved member packs the next k symbols at the current depth into a
// chunk key plus a boundary value for members whose window ends inside the
// chunk; the group is bitonic-sorted across its W lanes by (subgroup, chunk,
// boundary value, position); subgroups split where chunk or boundary value
// differ.  A subgroup is settled when it has one member or its members all
// ended (comparator-Equal: position order).  The depth at which sorted
// neighbours r-1, r part stays with position r (subgroup boundaries never
// move), and the final LCP caps it by both members' windows.
// ---------------------------------------------------------------------------
constexpr uint64_t kAuxCont = ~0ull;

// ---------------------------------------------------------------------------
// Two-member groups on the packed code text: four lanes per pair, each lane
// compares 128 bits of codes (128 / bits symbols) of both suffixes per step,
// loaded as three aligned words and shifted; the first step uses the pair's
// first lane only.  The first differing bit gives the first differing code,
// so the LCP (in symbols) is exact; codes order like symbols, so the two
// windows also give the pair's order without touching the text.  A word is
// loaded only if it holds a symbol inside the comparison window.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load128_codes(const uint64_t *__restrict__ pk, uint32_t lgf, uint32_t bits,
                                              uint64_t q, uint64_t end, uint64_t &hi, uint64_t &lo) {
    const uint64_t w = q >> lgf;
    const uint32_t rb = (uint32_t)(q & ((1ull << lgf) - 1)) * bits;
    const uint64_t w0 = (w << lgf) < end ? pk[w] : 0ull;
    const uint64_t w1 = ((w + 1) << lgf) < end ? pk[w + 1] : 0ull;
    const uint64_t w2 = rb && ((w + 2) << lgf) < end ? pk[w + 2] : 0ull;
    hi = rb ? (w0 << rb) | (w1 >> (64 - rb)) : w0;
    lo = rb ? (w1 << rb) | (w2 >> (64 - rb)) : w1;
}

// first differing code of two 128-bit windows (128 / bits if none)
__device__ __forceinline__ uint32_t first_diff_code(uint64_t ahi, uint64_t alo, uint64_t bhi, uint64_t blo,
                                                    uint32_t bits) {
    const uint64_t x = ahi ^ bhi, y = alo ^ blo;
    if (x) return (uint32_t)__builtin_clzll(x) / bits;
    if (y) return (64 + (uint32_t)__builtin_clzll(y)) / bits;
    return 128 / bits;
}

template <typename S, typename I>
__global__ __launch_bounds__(256) void k_pairs_packed(const uint64_t *__restrict__ pk,
                                                      uint32_t lgf, uint32_t bits, DevLimits L, uint64_t max_ctx,
                                                      const PairRec<I> *__restrict__ pairs, uint32_t n_pairs,
                                                      uint64_t known, I *__restrict__ idx, I *__restrict__ lcp) {
    const uint32_t lane = threadIdx.x & 63, grp = lane >> 2, sl = lane & 3;
    const uint32_t pid = blockIdx.x * 64 + (threadIdx.x >> 2);
    const bool valid = pid < n_pairs;
    const uint32_t FL = 128 / bits;  // codes per lane and step
    uint64_t s = 0, a = 0, b = 0, la = 0, lb = 0, cap = 0;
    if (valid) {
        PairRec<I> pr = pairs[pid];
        s = pr.start;
        a = pr.a;
        b = pr.b;
        la = lim_at(L, a);
        lb = lim_at(L, b);
        uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
        cap = ca < cb ? ca : cb;
    }
    uint64_t o = known < cap ? known : cap;  // symbols known equal
    bool done = !valid || o >= cap;
    uint64_t res = cap;
    // a's suffix orders after b's (meaningful when res < cap): the 128-bit
    // windows that hold the first differing code compare like it, so the
    // order comes from the packed words already loaded, not from the text
    bool gt = false;
    uint64_t nsym = 0;  // symbols read (both members)
    if (!done && sl == 0) {
        uint64_t ahi, alo, bhi, blo;
        load128_codes(pk, lgf, bits, a + o, a + cap, ahi, alo);
        load128_codes(pk, lgf, bits, b + o, b + cap, bhi, blo);
        uint32_t fc = first_diff_code(ahi, alo, bhi, blo, bits);
        const uint64_t in_win = cap - o;
        nsym += 2 * (in_win < FL ? in_win : FL);
        if (in_win < fc) fc = (uint32_t)in_win;
        if (fc < FL) {
            res = o + fc;
            gt = ahi != bhi ? ahi > bhi : alo > blo;
            done = true;
        }
    }
    done = __shfl(done ? 1 : 0, (int)(lane & ~3u)) != 0;
    res = shfl64(res, (int)(lane & ~3u));
    if (!done) o += FL;
    while (__any(!done)) {
        uint32_t fc = FL;
        bool g = false;
        if (!done) {
            const uint64_t rel = o + (uint64_t)sl * FL;
            if (rel < cap) {
                uint64_t ahi, alo, bhi, blo;
                load128_codes(pk, lgf, bits, a + rel, a + cap, ahi, alo);
                load128_codes(pk, lgf, bits, b + rel, b + cap, bhi, blo);
                fc = first_diff_code(ahi, alo, bhi, blo, bits);
                g = ahi != bhi ? ahi > bhi : alo > blo;
                nsym += 2 * (cap - rel < FL ? cap - rel : FL);
            } else {
                fc = 0;
            }
            const uint64_t in_win = cap > rel ? cap - rel : 0;
            if (in_win < fc) fc = (uint32_t)in_win;
        }
        const uint64_t mask = __ballot(!done && fc < FL);
        const uint32_t gm = (uint32_t)(mask >> (grp * 4)) & 0xFu;
        const int fl = gm ? __ffs(gm) - 1 : 0;
        const uint32_t fq = __shfl(fc, (int)(grp * 4 + fl));
        const bool gq = __shfl(g ? 1 : 0, (int)(grp * 4 + fl)) != 0;
        if (!done) {
            if (gm) {
                const uint64_t r = o + (uint64_t)fl * FL + fq;
                res = r < cap ? r : cap;
                gt = gq;
                done = true;
            } else {
                o += 4ull * FL;
            }
        }
    }
    if (valid && sl == 0) {
        int c;
        if (res < cap) c = gt ? 1 : -1;
        else c = boundary_then_pos(L, a, la, b, lb);
        if (c > 0) {
            idx[s] = (I)b;
            idx[s + 1] = (I)a;
        }
        if (lcp) lcp[s + 1] = (I)res;
    }
    count_symbols(0, nsym);
}

// Fast-forward for a slot of W lanes whose members share a long stretch
// (a repeat copy, a run of one symbol): every open member loads the next lim
// (<= J) whole code words (2^lgf codes each) at depth D that lie inside its window
// and compares them with those of its subgroup's first member (lane `sub`
// of the slot).  Returns the number of symbols, uniform across the slot, on
// which every open member agrees with its subgroup's first member, all
// inside every open window: rounds over them would split nothing, settle
// nothing and end no window.  Open and settled members never share a
// subgroup, so a first member of an open member is open.  Loads touch only
// words holding an in-window symbol.  Must be called by all W lanes of the
// slot at once, with at least one open lane.
template <int W, int J>
__device__ __forceinline__ uint64_t ff_skip(const uint64_t *__restrict__ pk, uint32_t lgf, uint32_t bits,
                                            uint64_t pos, uint64_t D, uint64_t cap, bool open, uint32_t sub,
                                            uint32_t lim, uint64_t &nsym) {
    const int lane = threadIdx.x & 63;
    uint32_t jw = 0;
    uint64_t pw[J];
#pragma unroll
    for (int j = 0; j < J; ++j) pw[j] = 0;
    if (open) {
        const uint64_t rem = cap > D ? cap - D : 0;
        const uint64_t fw = rem >> lgf;
        jw = fw < (uint64_t)lim ? (uint32_t)fw : lim;
        const uint64_t q = pos + D;
        const uint64_t w0 = q >> lgf;
        const uint32_t rb = (uint32_t)(q & ((1ull << lgf) - 1)) * bits;
        // word j of the stretch is raw[j] (and raw[j+1] when unaligned); raw[jw]
        // then holds the stretch's last symbol, which is inside the window
        const uint32_t nl = jw ? jw + (rb ? 1u : 0u) : 0u;
        nsym += (uint64_t)jw << lgf;
        uint64_t raw[J + 1];
#pragma unroll
        for (int i = 0; i <= J; ++i) raw[i] = (uint32_t)i < nl ? pk[w0 + i] : 0ull;
#pragma unroll
        for (int j = 0; j < J; ++j) pw[j] = rb ? (raw[j] << rb) | (raw[j + 1] >> (64 - rb)) : raw[j];
    }
    const int ld = (lane & ~(W - 1)) + (int)(sub & (W - 1));
    const uint32_t jl = (uint32_t)__shfl((int)jw, ld);
    uint32_t f = open ? (jw < jl ? jw : jl) : (uint32_t)J;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const uint64_t lw = shfl64(pw[j], ld);
        if (open && (uint32_t)j < f && lw != pw[j]) f = (uint32_t)j;
    }
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
        const uint32_t v = seg_xor<W>(f, o);
        f = v < f ? v : f;
    }
    return (uint64_t)f << lgf;
}

// Probing schedule of the small classes: after the first round a slot
// probes kFFFirst words; while a probe matches in full (every open member
// agrees with its subgroup's first member on every word) it probes again,
// kFFWords words at a time, up to kFFProbes probes per round.  Repeat copies
// with scattered differences part within the first short probe; runs of one
// symbol are crossed kFFWords words per probe without the round's sort.
constexpr uint32_t kFFAfter = 1;
constexpr uint32_t kFFFirst = 4;
constexpr int kFFWords = 4;
constexpr int kFFProbes = 16;

// After a probe that matched in full, with a run memo: when every open
// member's last matched symbol lies in a recorded run of one code, each
// member agrees with its subgroup's first member (whose last matched symbol
// is the same code) on the shorter rest of their runs, so the slot skips the
// least rest over its open members (whole words, inside every open window)
// at once.
template <int W>
__device__ __forceinline__ uint64_t ff_advance(const uint64_t *__restrict__ pk, uint32_t lgf, uint32_t bits,
                                               const RunTable &runs, uint64_t pos, uint64_t D, uint64_t cap,
                                               bool open, uint32_t sub, uint64_t &nsym) {
    const int lane = threadIdx.x & 63;
    const uint64_t smask = (W == 64) ? ~0ull : (((1ull << W) - 1) << (lane & ~(W - 1)));
    if ((__ballot(open) & smask) == 0) return 0;  // a settled slot
    uint64_t adv = 0;
    uint32_t lim = kFFFirst;
    for (int it = 0; it < kFFProbes; ++it) {  // uniform across the slot
        const uint64_t sk = ff_skip<W, kFFWords>(pk, lgf, bits, pos, D + adv, cap, open, sub, lim, nsym);
        adv += sk;
        if (sk != ((uint64_t)lim << lgf)) break;
        lim = kFFWords;
        if (runs.n_runs) {
            uint64_t r = ~0ull;  // symbols this member may skip (settled lanes do not bound the slot)
            if (open) {
                const uint64_t x = pos + D + adv;        // first symbol past the matched words (adv > 0)
                const uint64_t rr = run_remaining(runs, x - 1);  // x - 1 inside a run: rr >= 1
                const uint64_t rem = cap > D + adv ? cap - (D + adv) : 0;
                r = rr ? rr - 1 : 0;
                r = r < rem ? r : rem;
                r = (r >> lgf) << lgf;
            }
#pragma unroll
            for (int o = 1; o < W; o <<= 1) {
                const uint64_t v = seg_xor64<W>(r, o);
                r = v < r ? v : r;
            }
            if (r != ~0ull && r > 0) adv += r;
        }
    }
    return adv;
}

// ---------------------------------------------------------------------------
// Groups of 3..64 members, one W-lane slot per group (grid launch).  The
// round structure is the one described for k_wgroups_wq below; the wavefront
// runs until all of its slots' groups are settled.
// ---------------------------------------------------------------------------
template <typename S, typename I, int W>
__global__ __launch_bounds__(256) void k_wgroups_grid(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                 uint64_t max_ctx, const Group *__restrict__ groups,
                                                 uint32_t n_groups, uint64_t depth0, I *__restrict__ idx,
                                                 I *__restrict__ lcp) {
    const int lane = threadIdx.x & 63;
    const uint32_t r = lane & (W - 1);
    const uint32_t gid = (blockIdx.x * 256 + threadIdx.x) / W;
    Group gr{0, 0};
    if (gid < n_groups) gr = groups[gid];
    const uint32_t m = (uint32_t)gr.size;
    const bool act = r < m;
    const bool star = L.order == 1;
    const uint64_t cont = star ? 0 : kAuxCont;
    uint64_t pos = 0, lim = 0, cap = 0, chunk = ~0ull, aux = ~0ull, h = 0;
    uint32_t sub = act ? 0 : 0xFFFFFFFFu;
    bool settled = !act;
    if (act) {
        pos = idx[gr.start + r];
        lim = lim_at(L, pos);
        cap = lim < max_ctx ? lim : max_ctx;
    }
    uint64_t D = depth0;
    uint64_t nsym = 0;  // symbols read
    for (uint32_t round = 0;; ++round) {
        if (cm.packed && round >= kFFAfter && __any(!settled))
            D += ff_advance<W>(cm.packed, cm.lgf, g.bits, cm.runs, pos, D, cap, !settled, sub, nsym);
        if (!settled) {
            chunk = key_at(t, cm, g, pos, D, cap);
            uint64_t rem = cap > D ? cap - D : 0;
            nsym += rem < g.k ? rem : g.k;
            if (rem >= g.k) aux = cont;
            else aux = star ? (1 + (kAuxCont - 1 - lim)) : lim;
        }
        // bitonic sort of the W-lane segment
#pragma unroll
        for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                uint32_t p_sub = seg_xor<W>(sub, stride);
                uint64_t p_chunk = seg_xor64<W>(chunk, stride);
                uint64_t p_aux = seg_xor64<W>(aux, stride);
                uint64_t p_pos = seg_xor64<W>(pos, stride);
                uint64_t p_lim = seg_xor64<W>(lim, stride);
                uint64_t p_cap = seg_xor64<W>(cap, stride);
                int p_settled = (int)seg_xor<W>((uint32_t)settled, stride);
                bool less;  // me < partner
                if (sub != p_sub) less = sub < p_sub;
                else if (chunk != p_chunk) less = chunk < p_chunk;
                else if (aux != p_aux) less = aux < p_aux;
                else less = pos < p_pos;
                bool up = (r & size) == 0;
                bool lower = (r & stride) == 0;
                bool keep = (lower == up) ? less : !less;
                if (!keep) {
                    sub = p_sub, chunk = p_chunk, aux = p_aux, pos = p_pos, lim = p_lim, cap = p_cap;
                    settled = p_settled != 0;
                }
            }
        }
        // neighbour r-1 after the sort
        uint32_t q_sub = seg_prev<W>(sub);
        uint64_t q_chunk = seg_prev64<W>(chunk);
        uint64_t q_aux = seg_prev64<W>(aux);
        bool same_old = r > 0 && r < m && sub == q_sub;
        bool f = !(same_old && chunk == q_chunk && aux == q_aux);
        if (same_old && !settled) h = D + shared_fields(g, q_chunk, chunk);
        // new subgroup id: inclusive max-scan of (f ? r : 0) within the segment
        uint32_t ns = f ? r : 0;
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            uint32_t v = seg_up<W>(ns, o);
            if ((int)r >= o) ns = v > ns ? v : ns;
        }
        uint32_t ns_prev = seg_prev<W>(ns);
        uint32_t ns_next = seg_next<W>(ns);
        bool nb = (r > 0 && r < m && ns_prev == ns) || (r + 1 < m && ns_next == ns);
        bool ended = aux != cont;
        if (act) {
            settled = settled || !nb || ended;
            sub = ns;
        }
        if (!__any(!settled)) break;
        D += g.k;
    }
    uint64_t q_cap = seg_prev64<W>(cap);
    if (act) {
        idx[gr.start + r] = (I)pos;
        if (lcp && r > 0) {
            uint64_t v = h;
            v = v < cap ? v : cap;
            v = v < q_cap ? v : q_cap;
            lcp[gr.start + r] = (I)v;
        }
    }
    count_symbols(1, nsym);
}


// Work queue shared by the persistent refinement kernels.  A wavefront
// claims `claim` consecutive items with one atomic and hands them to its
// slots of W lanes in order (ballot + popcount, no further atomics); `cb`/`ce`
// hold the wavefront's unclaimed range and are uniform across it.  The host
// sizes `claim` (at least the slots per wavefront) so that the queue holds a
// few claims per wavefront: large enough to keep the counter cold, small
// enough that no wavefront is left with a long tail.  Returns, in every lane
// of a slot that asked, the item it got (>= n: queue empty).
template <int W>
__device__ __forceinline__ uint32_t wq_take(unsigned int *next, uint32_t claim, bool ask, uint32_t &cb,
                                            uint32_t &ce) {
    const uint32_t lane = threadIdx.x & 63;
    const bool leader_asks = ask && (lane & (W - 1)) == 0;
    const uint64_t m = __ballot(leader_asks);
    uint32_t got = 0xFFFFFFFFu;
    if (m) {
        const uint32_t k = (uint32_t)__popcll(m);
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1));
        const uint32_t avail = ce - cb;
        if (k > avail) {
            const uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1;
            uint32_t nb = 0;
            if (lane == first) nb = atomicAdd(next, claim);
            nb = __shfl(nb, (int)first);
            got = rank < avail ? cb + rank : nb + (rank - avail);
            cb = nb + (k - avail);
            ce = nb + claim;
        } else {
            got = cb + rank;
            cb += k;
        }
    }
    return __shfl(got, (int)(lane & ~(uint32_t)(W - 1)));
}

// Groups of 3..64 members: W lanes per group (W = 4..64), persistent.  Each
// slot of W lanes refines one group at a time, round by round, and moves to
// the next as soon as it is done, so a wavefront never waits for its slowest
// group.  The queue is software-pipelined two groups deep: the group record
// of the slot's next-but-one group and the positions of its next group are
// loaded in the same iteration as the current round's text, and used one
// iteration later, so taking a new group adds no load latency of its own.
//
// Each round every unresolved member packs the next k symbols at the slot's
// depth into a chunk key plus a boundary value for members whose window ends
// inside the chunk; the group is bitonic-sorted across its W lanes by
// (subgroup, chunk, boundary value, position); subgroups split where chunk or
// boundary value differ.  A subgroup is settled when it has one member or its
// members all ended (comparator-Equal: position order).  The depth at which
// sorted neighbours r-1, r part stays with position r (subgroup boundaries
// never move), and the final LCP caps it by both members' windows.  Every
// wavefront exits once the queue is empty and its slots are idle; each group
// ends because the depth grows by k per round until every window has ended.
template <typename S, typename I, int W>
__global__ __launch_bounds__(256) void k_wgroups_wq(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
                                                 uint64_t max_ctx, const Group *__restrict__ groups,
                                                 uint32_t n_groups, uint64_t depth0, I *__restrict__ idx,
                                                 I *__restrict__ lcp, unsigned int *next, uint32_t claim) {
    const int lane = threadIdx.x & 63;
    const uint32_t r = lane & (W - 1);
    const bool star = L.order == 1;
    const uint64_t cont = star ? 0 : kAuxCont;
    uint32_t cb = 0, ce = 0;  // this wavefront's claimed, not yet handed out queue range
    bool drained = false;
    // stage 1: group record loaded; stage 2: its member positions loaded
    bool v1 = false, v2 = false;
    uint64_t s1 = 0, s2 = 0, p2 = 0;
    uint32_t m1 = 0, m2 = 0;
    // current group
    bool busy = false;
    uint64_t start = 0, D = 0;
    uint32_t m = 0, rounds = 0;
    uint64_t pos = 0, lim = 0, cap = 0, chunk = ~0ull, aux = ~0ull, h = 0;
    uint32_t sub = 0xFFFFFFFFu;
    bool settled = true;
    uint64_t nsym = 0;  // symbols read
    for (;;) {
        // advance the pipeline (all conditions are uniform within a slot)
        if (!busy && v2) {
            start = s2;
            m = m2;
            busy = true;
            v2 = false;
            D = depth0;
            rounds = 0;
            h = 0;
            chunk = ~0ull;
            aux = ~0ull;
            settled = r >= m;
            sub = settled ? 0xFFFFFFFFu : 0;
            if (!settled) {
                pos = p2;
                lim = lim_at(L, pos);
                cap = lim < max_ctx ? lim : max_ctx;
            }
        }
        if (v1 && !v2) {
            s2 = s1;
            m2 = m1;
            if (r < m1) p2 = idx[s1 + r];
            v2 = true;
            v1 = false;
        }
        const bool ask = !v1 && !drained;
        const uint32_t gid = wq_take<W>(next, claim, ask, cb, ce);
        if (ask) {
            if (gid < n_groups) {
                const Group gr = groups[gid];
                s1 = gr.start;
                m1 = (uint32_t)gr.size;
                v1 = true;
            } else {
                drained = true;
            }
        }
        if (!__any(busy || v1 || v2)) break;
        const bool act = busy && r < m;
        if (cm.packed && busy && rounds >= kFFAfter)
            D += ff_advance<W>(cm.packed, cm.lgf, g.bits, cm.runs, pos, D, cap, !settled, sub, nsym);
        if (!settled) {
            chunk = key_at(t, cm, g, pos, D, cap);
            uint64_t rem = cap > D ? cap - D : 0;
            nsym += rem < g.k ? rem : g.k;
            if (rem >= g.k) aux = cont;
            else aux = star ? (1 + (kAuxCont - 1 - lim)) : lim;
        }
        // bitonic sort of the W-lane segment
#pragma unroll
        for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                uint32_t p_sub = seg_xor<W>(sub, stride);
                uint64_t p_chunk = seg_xor64<W>(chunk, stride);
                uint64_t p_aux = seg_xor64<W>(aux, stride);
                uint64_t p_pos = seg_xor64<W>(pos, stride);
                uint64_t p_lim = seg_xor64<W>(lim, stride);
                uint64_t p_cap = seg_xor64<W>(cap, stride);
                int p_settled = (int)seg_xor<W>((uint32_t)settled, stride);
                bool less;  // me < partner
                if (sub != p_sub) less = sub < p_sub;
                else if (chunk != p_chunk) less = chunk < p_chunk;
                else if (aux != p_aux) less = aux < p_aux;
                else less = pos < p_pos;
                bool up = (r & size) == 0;
                bool lower = (r & stride) == 0;
                bool keep = (lower == up) ? less : !less;
                if (!keep) {
                    sub = p_sub, chunk = p_chunk, aux = p_aux, pos = p_pos, lim = p_lim, cap = p_cap;
                    settled = p_settled != 0;
                }
            }
        }
        // neighbour r-1 after the sort
        uint32_t q_sub = seg_prev<W>(sub);
        uint64_t q_chunk = seg_prev64<W>(chunk);
        uint64_t q_aux = seg_prev64<W>(aux);
        uint64_t q_cap = seg_prev64<W>(cap);
        bool same_old = r > 0 && r < m && sub == q_sub;
        bool f = !(same_old && chunk == q_chunk && aux == q_aux);
        if (act && same_old && !settled) h = D + shared_fields(g, q_chunk, chunk);
        // new subgroup id: inclusive max-scan of (f ? r : 0) within the segment
        uint32_t ns = f ? r : 0;
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            uint32_t v = seg_up<W>(ns, o);
            if ((int)r >= o) ns = v > ns ? v : ns;
        }
        uint32_t ns_prev = seg_prev<W>(ns);
        uint32_t ns_next = seg_next<W>(ns);
        bool nb = (r > 0 && r < m && ns_prev == ns) || (r + 1 < m && ns_next == ns);
        bool ended = aux != cont;
        if (act) {
            settled = settled || !nb || ended;
            sub = ns;
        }
        // a slot is done when none of its lanes is unsettled
        const uint64_t open = __ballot(!settled);
        const uint64_t smask = (W == 64) ? ~0ull : (((1ull << W) - 1) << (lane & ~(W - 1)));
        if (busy && (open & smask) == 0) {
            if (act) {
                idx[start + r] = (I)pos;
                if (lcp && r > 0) {
                    uint64_t v = h;
                    v = v < cap ? v : cap;
                    v = v < q_cap ? v : q_cap;
                    lcp[start + r] = (I)v;
                }
            }
            busy = false;
            m = 0;
            sub = 0xFFFFFFFFu;
            settled = true;
        } else {
            D += g.k;
            ++rounds;
        }
    }
    count_symbols(1, nsym);
}

// ---------------------------------------------------------------------------
// Two-member groups, persistent: the four-lane compare above, with slots that
// move to the next pair as soon as theirs is resolved.  The next pair's
// record is loaded one iteration ahead, beside the current compare's text.
// ---------------------------------------------------------------------------
template <typename S, typename I>
__global__ __launch_bounds__(256) void k_pairs_wq(const S *__restrict__ t, DevLimits L, uint64_t max_ctx,
                                                  const PairRec<I> *__restrict__ pairs, uint32_t n_pairs,
                                                  uint64_t known, I *__restrict__ idx, I *__restrict__ lcp,
                                                  unsigned int *next, uint32_t claim) {
    constexpr uint64_t SZ = sizeof(S);
    const uint32_t lane = threadIdx.x & 63, grp = lane >> 2, sl = lane & 3;
    const uint64_t tb = (uint64_t)(uintptr_t)t;
    bool busy = false, drained = false, v1 = false;
    uint32_t cb = 0, ce = 0;
    PairRec<I> nx{0, 0, 0};
    uint64_t s = 0, a = 0, b = 0, la = 0, lb = 0, cap = 0, capb = 0, enda = 0, endb = 0, ob = 0, resb = 0;
    for (;;) {
        if (!busy && v1) {
            s = nx.start;
            a = nx.a;
            b = nx.b;
            la = lim_at(L, a);
            lb = lim_at(L, b);
            uint64_t ca = la < max_ctx ? la : max_ctx, cb2 = lb < max_ctx ? lb : max_ctx;
            cap = ca < cb2 ? ca : cb2;
            capb = cap * SZ;
            enda = tb + (a + cap) * SZ;
            endb = tb + (b + cap) * SZ;
            ob = (known < cap ? known : cap) * SZ;
            resb = capb;
            busy = true;
            v1 = false;
        }
        const bool ask = !v1 && !drained;
        const uint32_t pid = wq_take<4>(next, claim, ask, cb, ce);
        if (ask) {
            if (pid < n_pairs) {
                nx = pairs[pid];
                v1 = true;
            } else {
                drained = true;
            }
        }
        if (!__any(busy || v1)) break;
        uint32_t fb = 16;
        const bool work = busy && ob < capb;
        if (work) {
            uint64_t rel = ob + 16 * sl;
            uint64_t alo, ahi, blo, bhi;
            load16(tb + a * SZ + rel, enda, alo, ahi);
            load16(tb + b * SZ + rel, endb, blo, bhi);
            uint64_t d0 = alo ^ blo, d1 = ahi ^ bhi;
            fb = d0 ? (uint32_t)__builtin_ctzll(d0) >> 3 : d1 ? 8 + ((uint32_t)__builtin_ctzll(d1) >> 3) : 16u;
            uint64_t in_win = capb > rel ? capb - rel : 0;
            if (in_win < fb) fb = (uint32_t)in_win;
        }
        uint64_t mask = __ballot(work && fb < 16);
        uint32_t gm = (uint32_t)(mask >> (grp * 4)) & 0xFu;
        int fl = gm ? __ffs(gm) - 1 : 0;
        uint32_t fq = __shfl(fb, (int)(grp * 4 + fl));
        bool finish = busy && !work;  // window already exhausted by the known prefix
        if (work) {
            if (gm) {
                uint64_t rr = ob + 16 * (uint64_t)fl + fq;
                resb = rr < capb ? rr : capb;
                finish = true;
            } else {
                ob += 64;
            }
        }
        if (finish) {
            if (sl == 0) {
                uint64_t res = resb / SZ;
                int c;
                if (res < cap) c = t[a + res] < t[b + res] ? -1 : 1;
                else c = boundary_then_pos(L, a, la, b, lb);
                if (c > 0) {
                    idx[s] = (I)b;
                    idx[s + 1] = (I)a;
                }
                if (lcp) lcp[s + 1] = (I)res;
            }
            busy = false;
        }
    }
}

}  // namespace caps
