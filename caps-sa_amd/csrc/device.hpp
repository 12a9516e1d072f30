// Device-side primitives shared by every kernel of the suffix-array path:
// the limit provider, packed prefix keys, the suffix comparator and the
// 64-lane wavefront LCP scan.  gfx950 only (wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace caps {

constexpr uint64_t kMaxCtx = ~0ull;

// Refinement work counters: symbols read from the text by the tie
// refinement (both members of a compare counted), per kernel family:
// [0] pairs, [1] groups of 3..64, [2] groups of 65..2,048 and giant-group
// chunks.  Counted only while the host sets g_count_symbols (a measurement
// build); one atomic per wavefront at the end of a kernel, spread over
// kCountLanes counters per family so that a grid of millions of wavefronts
// does not queue on one address (the host sums them).
constexpr int kCountLanes = 256;
__device__ unsigned long long g_cmp_symbols[3 * kCountLanes];
__device__ unsigned int g_count_symbols;

// Must be reached by every lane of the wavefront.
__device__ __forceinline__ void count_symbols(int slot, uint64_t v) {
    if (!g_count_symbols) return;
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o), hi = (uint32_t)__shfl_xor((int)(v >> 32), o);
        v += ((uint64_t)hi << 32) | lo;
    }
    const uint32_t spread = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kCountLanes - 1);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&g_cmp_symbols[slot * kCountLanes + spread], (unsigned long long)v);
}

// Limit provider descriptor.  PlainText when ends == nullptr
// (limits.rs:130-156); SegmentedText otherwise (limits.rs:182-336), with the
// coarse block directory of limits.rs:194-236 when dir != nullptr.
struct DevLimits {
    uint64_t n;
    const uint64_t *ends;
    uint64_t n_ends;
    const uint32_t *dir;
    uint32_t block_shift;
    uint32_t order;  // 0: shorter suffix first; 1: STAR (longer first, then position)
};

// limits.rs:277-303.
__device__ __forceinline__ uint64_t lim_at(const DevLimits &L, uint64_t p) {
    if (L.ends == nullptr) return L.n - p;
    uint64_t lo = 0, hi = L.n_ends;
    if (L.dir != nullptr && p < L.n) {
        uint64_t b = p >> L.block_shift;
        lo = L.dir[b];
        hi = L.dir[b + 1];
        if (hi < lo + 1) hi = lo + 1;
        if (hi > L.n_ends) hi = L.n_ends;
    }
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (L.ends[mid] <= p) lo = mid + 1;
        else hi = mid;
    }
    // the last end equals n (checked on the host), so p at or past it reads
    // that end; no select between ends[] and L.n, which the compiler would
    // turn into a generic load from a private copy of L
    return L.ends[lo < L.n_ends ? lo : L.n_ends - 1] - p;
}

// Segment index of position p (0 for PlainText) and its limit.
__device__ __forceinline__ uint64_t lim_seg(const DevLimits &L, uint64_t p, uint64_t *seg) {
    if (L.ends == nullptr) {
        *seg = 0;
        return L.n - p;
    }
    uint64_t lo = 0, hi = L.n_ends;
    if (L.dir != nullptr && p < L.n) {
        uint64_t b = p >> L.block_shift;
        lo = L.dir[b];
        hi = L.dir[b + 1];
        if (hi < lo + 1) hi = lo + 1;
        if (hi > L.n_ends) hi = L.n_ends;
    }
    while (lo < hi) {
        uint64_t mid = lo + ((hi - lo) >> 1);
        if (L.ends[mid] <= p) lo = mid + 1;
        else hi = mid;
    }
    *seg = lo;
    return L.ends[lo < L.n_ends ? lo : L.n_ends - 1] - p;
}

// Comparison window of a suffix: its limit, clipped by max_context.
__device__ __forceinline__ uint64_t cap_at(const DevLimits &L, uint64_t p, uint64_t max_ctx) {
    uint64_t l = lim_at(L, p);
    return l < max_ctx ? l : max_ctx;
}

__device__ __forceinline__ int cmp64(uint64_t a, uint64_t b) { return (a > b) - (a < b); }

// limits.rs:66-75 (default) and the STAR convention
// lim_b.cmp(&lim_a).then(p_a.cmp(&p_b)) (sample_sort.rs:1006-1014).  Under
// the default convention, suffixes the provider calls Equal are ordered by
// ascending position (a canonical order; the reference leaves it unspecified).
__device__ __forceinline__ int boundary_then_pos(const DevLimits &L, uint64_t pa, uint64_t la, uint64_t pb,
                                                 uint64_t lb) {
    int c = L.order == 1 ? cmp64(lb, la) : cmp64(la, lb);
    return c ? c : cmp64(pa, pb);
}

// Packed prefix key geometry.  Fields are `bits` wide, `k` per key, MSB
// first, left-aligned by `shift` = 64 - k*bits.  A suffix contributes
// min(k, remaining window) codes (+bias); the rest of the key is `pad`.
//   reference geometry (pack.rs:187-203): bias 1 / pad 0 (shorter first) or
//     bias 0 / pad alphabet (longer first), strictly order-preserving;
//   dense geometry: bias 0, pad 0 (shorter first) or all-ones (longer first);
//     every code is a symbol, so keys are weakly order-preserving: a strict
//     key inequality still implies the same suffix inequality, and equal keys
//     are settled by refinement.
//   batched builds: the top `idbits` bits carry the text (segment) index, so
//     one sort orders many texts text-major; the fields sit below them.
//   escape geometry (esc > 0; 2-bit fields, 32 per key, no shift): codes
//     below esc (the four nucleotides of a genome) are their own fields; a
//     code >= esc (a spacer, an N) and every field after it inside the
//     window become the largest field value, whatever the window does later;
//     a window end without an escape pads as the dense geometry does.  Codes
//     >= esc order above every field value, so a strict key inequality still
//     implies the suffix inequality (weakly order-preserving), and equal keys
//     do not imply equal leading symbols: their refinement starts at depth 0.
struct KeyGeom {
    uint32_t bits, k, bias, pad, shift, idbits;
    uint32_t esc;  // 0: no escape codes
};

// Escape geometry: the fields from the first escape on, given the escape
// indicator key e (a field is all ones where its symbol is an escape).
__device__ __forceinline__ uint64_t escape_tail(uint64_t e) {
    return e ? (~0ull >> ((uint32_t)__builtin_clzll(e) & ~1u)) : 0ull;
}

// Run memo: maximal runs of one code in the packed code text that cover at
// least kRunMinWords aligned words, as [start, end) symbol intervals sorted
// by start, with a directory over blocks of 2^kRunDirShift symbols (dir[b]:
// the first run whose end lies past the block's start).  Each entry proves
// that every suffix starting inside the run begins with end - x copies of
// the run's code, so two suffixes inside runs of the same code agree on the
// shorter remainder: an LCP scan skips it at once.  A subset of the runs is
// as correct as all of them (a missing run only means no skip).
struct RunTable {
    const uint64_t *start = nullptr, *end = nullptr;
    const uint32_t *dir = nullptr;  // n_dir + 1 entries
    uint64_t n_runs = 0;
    uint64_t n_dir = 0;
};
constexpr uint32_t kRunDirShift = 12;
constexpr uint32_t kRunMinWords = 3;

// Symbols from x to the end of the recorded run holding x (0 if none).
__device__ __forceinline__ uint64_t run_remaining(const RunTable &R, uint64_t x) {
    const uint64_t b = x >> kRunDirShift;
    if (b >= R.n_dir) return 0;
    uint64_t i = R.dir[b];
    while (i < R.n_runs && R.end[i] <= x) ++i;
    return (i < R.n_runs && R.start[i] <= x) ? R.end[i] - x : 0;
}

template <typename S>
struct CodeMap;
// Symbol codes.  `packed`, when set, holds the whole text as codes of
// 64 >> lgf bits (a power of two) each, MSB-first in u64 words (2^lgf codes
// a word, zero words past the end): a key of that field width is then two
// aligned loads and a shift.  `seed`, when set, holds the text in the seed
// keys' field width b: sk = 64 / b codes per word, MSB-first, the low
// 64 - sk*b bits zero (the packed text itself when b is a power of two);
// word q / sk holds code q, and q / sk = umulhi(q, srec) with
// srec = ceil(2^64 / sk), exact for q < 2^40.
template <>
struct CodeMap<uint8_t> {
    const uint8_t *code = nullptr;  // nullptr = identity
    const uint64_t *packed = nullptr;
    RunTable runs;
    uint32_t lgf = 0;
    uint32_t sk = 0;
    const uint64_t *seed = nullptr;
    uint64_t srec = 0;
    __device__ __forceinline__ uint32_t operator()(uint8_t c) const { return code ? code[c] : c; }
};
template <>
struct CodeMap<uint16_t> {
    const uint16_t *code = nullptr;  // nullptr = identity
    const uint64_t *packed = nullptr;
    RunTable runs;
    uint32_t lgf = 0;
    uint32_t sk = 0;
    const uint64_t *seed = nullptr;
    uint64_t srec = 0;
    __device__ __forceinline__ uint32_t operator()(uint16_t c) const { return code ? code[c] : c; }
};

// Pad fields for the last `m` of the k fields: pad repeated m times.
__device__ __forceinline__ uint64_t pad_fields(const KeyGeom &g, uint32_t m) {
    if (g.pad == 0 || m == 0) return 0;
    uint64_t r = 0;
    for (uint32_t i = 0; i < m; ++i) r = (r << g.bits) | g.pad;
    return r;
}

// key_at from the packed code text (dense geometry: no bias; power-of-two
// field width, so a word holds exactly 2^lgf codes).
template <typename CM>
__device__ __forceinline__ uint64_t key_at_packed(const CM &cm, const KeyGeom &g, uint64_t p, uint64_t d,
                                                  uint64_t cap) {
    const uint64_t rem = cap > d ? cap - d : 0;
    const uint32_t take = rem < g.k ? (uint32_t)rem : g.k;
    if (take == 0) return pad_fields(g, g.k) << g.shift;
    const uint32_t bits = g.bits;
    const uint64_t q = p + d;
    const uint64_t w = q >> cm.lgf;
    const uint32_t rb = (uint32_t)(q & ((1ull << cm.lgf) - 1)) * bits;
    uint64_t v = cm.packed[w];
    if (rb) v = (v << rb) | (cm.packed[w + 1] >> (64 - rb));
    const uint32_t kb = g.k * bits;
    uint64_t fields = kb >= 64 ? v : (v >> (64 - kb));
    if (take < g.k) {
        const uint32_t drop = (g.k - take) * bits;
        fields = ((fields >> drop) << drop) | pad_fields(g, g.k - take);
    }
    return fields << g.shift;
}

// Key from the text bytes: all aligned 8-byte words covering the window are
// loaded first (each holds at least one in-range byte), then the fields are
// extracted with constant indices, so the loads overlap instead of forming a
// dependent chain.
template <typename S>
__device__ __forceinline__ uint64_t key_at_bytes(const S *__restrict__ t, CodeMap<S> cm, const KeyGeom &g, uint64_t p,
                                                 uint64_t d, uint64_t cap) {
    constexpr uint32_t SZ = sizeof(S);
    constexpr int NW = 9;  // 64 bytes of window + up to 7 bytes of misalignment
    uint64_t rem = cap > d ? cap - d : 0;
    uint32_t take = rem < g.k ? (uint32_t)rem : g.k;
    uint64_t key = 0;
    if (take) {
        // words aligned on the absolute address: each loaded word holds an
        // in-window byte, so none crosses into a page past the text
        uint64_t b0 = (uint64_t)(uintptr_t)t + (p + d) * SZ, b1 = b0 + (uint64_t)take * SZ;
        uint64_t a0 = b0 & ~7ull;
        uint32_t nw = (uint32_t)(((b1 + 7) & ~7ull) - a0) >> 3;
        uint64_t w[NW];
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = i < (int)nw ? *(const uint64_t *)(a0 + 8 * i) : 0ull;
        uint32_t lo = (uint32_t)(b0 - a0), hi = lo + take * SZ;
        uint64_t esc = 0;  // escape geometry: the escape indicator fields
        const uint64_t fmax = (1ull << g.bits) - 1;
#pragma unroll
        for (int j = 0; j < NW * 8; j += SZ) {
            if ((uint32_t)j >= lo && (uint32_t)j < hi) {
                uint32_t sym = SZ == 1 ? (uint32_t)(w[j >> 3] >> (8 * (j & 7))) & 0xFFu
                                       : (uint32_t)(w[j >> 3] >> (8 * (j & 7))) & 0xFFFFu;
                const uint32_t c = cm((S)sym) + g.bias;
                const bool e = g.esc && c >= g.esc;
                key = (key << g.bits) | (e ? fmax : (uint64_t)c);
                esc = (esc << g.bits) | (e ? fmax : 0ull);
            }
        }
        if (g.esc) {  // 2-bit fields, 32 per key: fields after the first escape, past the window too
            const uint64_t tail = take < g.k ? (uint64_t)(g.k - take) * g.bits : 0;
            if (esc) return ((key << tail) | escape_tail(esc << tail));
        }
    }
    if (take == 0) key = pad_fields(g, g.k);
    else if (take < g.k) key = (key << (g.bits * (g.k - take))) | pad_fields(g, g.k - take);
    return key << g.shift;
}

// Key of the `k` symbols starting at depth d of suffix p, whose comparison
// window is `cap` symbols (limit and max_context already applied), in the
// packed text's field width (refinement chunks).
template <typename S>
__device__ __forceinline__ uint64_t key_at(const S *__restrict__ t, CodeMap<S> cm, const KeyGeom &g, uint64_t p,
                                           uint64_t d, uint64_t cap) {
    if (cm.packed) return key_at_packed(cm, g, p, d, cap);
    return key_at_bytes(t, cm, g, p, d, cap);
}

// Seed key of suffix p (depth 0), in the seed geometry, whose fields may be
// narrower than the packed text's: read from the seed-width code text
// (word w holds codes w*sk .. w*sk+sk-1; the next word is loaded only when
// the key's codes reach into it, so it holds an in-window code).
template <typename S>
__device__ __forceinline__ uint64_t key_at_seed(const S *__restrict__ t, CodeMap<S> cm, const KeyGeom &g, uint64_t p,
                                                uint64_t cap) {
    if (!cm.seed) return key_at_bytes(t, cm, g, p, 0, cap);
    const uint32_t take = cap < g.k ? (uint32_t)cap : g.k;
    if (take == 0) return pad_fields(g, g.k) << g.shift;
    const uint32_t bits = g.bits;
    const uint64_t w = __umul64hi(p, cm.srec);
    const uint32_t r = (uint32_t)(p - w * cm.sk);
    uint64_t v = cm.seed[w] << (r * bits);
    if (r + take > cm.sk) v |= cm.seed[w + 1] >> ((cm.sk - r) * bits);
    const uint32_t kb = g.k * bits;
    uint64_t fields = kb >= 64 ? v : (v >> (64 - kb));
    if (take < g.k) {
        const uint32_t drop = (g.k - take) * bits;
        fields = ((fields >> drop) << drop) | pad_fields(g, g.k - take);
    }
    return fields << g.shift;
}

// Keys of the consecutive positions base+first .. base+first+cnt-1 from a
// text window staged in LDS (win[i] = text[base + i]), rolling the key by one
// field per position while the whole window is inside the suffix's limit.
// caps[i] is the comparison window of position base+first+i.
template <typename S, int CNT>
__device__ __forceinline__ void keys_from_window_esc(const S *win, CodeMap<S> cm, const KeyGeom &g, uint32_t first,
                                                     const uint64_t *caps, uint64_t *out) {
    // escape geometry: 2-bit fields, 32 per key (k * bits == 64, no shift);
    // the raw fields and the escape indicator roll together
    const uint32_t k = g.k;
    uint64_t ku = 0, eu = 0;
    bool rolling = false;
#pragma unroll
    for (int i = 0; i < CNT; ++i) {
        const uint32_t q = first + i;
        const uint64_t cap = caps[i];
        if (cap >= k) {
            if (rolling) {
                const uint32_t c = cm(win[q + k - 1]);
                const bool e = c >= g.esc;
                ku = (ku << 2) | (e ? 3ull : (uint64_t)c);
                eu = (eu << 2) | (e ? 3ull : 0ull);
            } else {
                ku = 0, eu = 0;
                for (uint32_t f = 0; f < k; ++f) {
                    const uint32_t c = cm(win[q + f]);
                    const bool e = c >= g.esc;
                    ku = (ku << 2) | (e ? 3ull : (uint64_t)c);
                    eu = (eu << 2) | (e ? 3ull : 0ull);
                }
                rolling = true;
            }
            out[i] = ku | escape_tail(eu);
        } else {
            const uint32_t take = (uint32_t)cap;
            uint64_t key = 0, e = 0;
            for (uint32_t f = 0; f < take; ++f) {
                const uint32_t c = cm(win[q + f]);
                const bool x = c >= g.esc;
                key = (key << 2) | (x ? 3ull : (uint64_t)c);
                e = (e << 2) | (x ? 3ull : 0ull);
            }
            const uint32_t tail = (k - take) * 2;
            if (e) out[i] = (take ? key << tail : 0ull) | escape_tail(take ? e << tail : 0ull);
            else out[i] = take ? ((key << tail) | pad_fields(g, k - take)) : pad_fields(g, k);
            rolling = false;
        }
    }
}

template <typename S, int CNT>
__device__ __forceinline__ void keys_from_window(const S *win, CodeMap<S> cm, const KeyGeom &g, uint32_t first,
                                                 const uint64_t *caps, uint64_t *out) {
    if (g.esc) {
        keys_from_window_esc<S, CNT>(win, cm, g, first, caps, out);
        return;
    }
    const uint32_t k = g.k, bits = g.bits;
    const uint64_t mask = (k * bits >= 64) ? ~0ull : ((1ull << (k * bits)) - 1);
    uint64_t ku = 0;
    bool rolling = false;
#pragma unroll
    for (int i = 0; i < CNT; ++i) {
        uint32_t q = first + i;
        uint64_t cap = caps[i];
        if (cap >= k) {
            if (rolling) {
                ku = ((ku << bits) & mask) | (uint64_t)(cm(win[q + k - 1]) + g.bias);
            } else {
                ku = 0;
                for (uint32_t f = 0; f < k; ++f) ku = (ku << bits) | (uint64_t)(cm(win[q + f]) + g.bias);
                rolling = true;
            }
            out[i] = ku << g.shift;
        } else {
            uint32_t take = (uint32_t)cap;
            uint64_t key = 0;
            for (uint32_t f = 0; f < take; ++f) key = (key << bits) | (uint64_t)(cm(win[q + f]) + g.bias);
            key = take ? ((key << (bits * (k - take))) | pad_fields(g, k - take)) : pad_fields(g, k);
            out[i] = key << g.shift;
            rolling = false;
        }
    }
}

// Symbols two keys share, from the first differing field (pack.rs:207-209);
// keys of different texts (different id bits) share nothing.
__device__ __forceinline__ uint32_t shared_fields(const KeyGeom &g, uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    uint32_t lz = x ? (uint32_t)__builtin_clzll(x) : 64u;
    if (lz < g.idbits) return 0;
    uint32_t f = (lz - g.idbits) / g.bits;
    return f < g.k ? f : g.k;
}

__device__ __forceinline__ uint64_t with_id(const KeyGeom &g, uint64_t key, uint64_t id) {
    return g.idbits ? (key | (id << (64 - g.idbits))) : key;
}

// The 8 bytes starting at byte address x (little-endian), from the aligned
// words that hold them; a word is loaded only when it holds a byte below
// `end` (x < end), so no load leaves the window.  Bytes at and past end read
// as zero or as whatever the loaded word holds: callers mask them.
__device__ __forceinline__ uint64_t load8_upto(uint64_t x, uint64_t end) {
    const uint64_t a8 = x & ~7ull;
    const uint32_t sh = (uint32_t)(x & 7) * 8;
    const uint64_t w0 = *(const uint64_t *)a8;
    if (!sh) return w0;
    const uint64_t w1 = a8 + 8 < end ? *(const uint64_t *)(a8 + 8) : 0ull;
    return (w0 >> sh) | (w1 << (64 - sh));
}

// Scalar LCP of text[p+start..] and text[q+start..], stopping at `cap`
// (lcp.rs:233-246 as eight bytes per step: the first differing byte of the
// two 8-byte stretches is the lowest set byte of their XOR).
template <typename S>
__device__ __forceinline__ uint64_t thread_lcp(const S *__restrict__ t, uint64_t p, uint64_t q, uint64_t start,
                                               uint64_t cap) {
    constexpr uint64_t SZ = sizeof(S);
    const uint64_t tb = (uint64_t)(uintptr_t)t;
    const uint64_t endp = tb + (p + cap) * SZ, endq = tb + (q + cap) * SZ;
    uint64_t o = start;
    while (o < cap) {
        const uint64_t wa = load8_upto(tb + (p + o) * SZ, endp), wb = load8_upto(tb + (q + o) * SZ, endq);
        uint64_t x = wa ^ wb;
        const uint64_t rem = (cap - o) * SZ;  // window bytes left
        if (rem < 8) x |= ~0ull << (rem * 8);  // bytes past the window differ
        if (x) {
            const uint64_t r = o + ((uint64_t)__builtin_ctzll(x) >> 3) / SZ;
            return r < cap ? r : cap;
        }
        o += 8 / SZ;
    }
    return cap;
}

// 64-lane wavefront LCP (the SIMD LCP of lcp.rs:292-360 as a wave-wide
// compare): every lane compares one symbol per step, a ballot collects the
// mismatch mask and ffs picks the first one.  Four steps (256 symbols) per
// iteration.  All 64 lanes must call it with the same arguments.
template <typename S>
__device__ __forceinline__ uint64_t wave_lcp(const S *__restrict__ t, uint64_t p, uint64_t q, uint64_t start,
                                             uint64_t cap) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t off = start;
    while (off < cap) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint64_t o = off + (uint64_t)(j * 64) + lane;
            bool mism = o < cap ? (t[p + o] != t[q + o]) : true;
            uint64_t mask = __ballot(mism);
            if (mask) {
                uint64_t r = off + (uint64_t)(j * 64) + (uint64_t)(__ffsll((unsigned long long)mask) - 1);
                return r < cap ? r : cap;
            }
        }
        off += 256;
    }
    return cap < start ? start : cap;
}

// Full order of two suffixes known to agree on their first `known` symbols
// (clipped to the window): lcp.rs:184-201 followed by the position tie.
// Returns <0, 0 (same position) or >0, and the LCP (window-capped) in *lcp.
template <typename S>
__device__ __forceinline__ int thread_cmp(const S *__restrict__ t, const DevLimits &L, uint64_t max_ctx, uint64_t a,
                                          uint64_t b, uint64_t known, uint64_t *lcp) {
    uint64_t la = lim_at(L, a), lb = lim_at(L, b);
    uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
    uint64_t cap = ca < cb ? ca : cb;
    uint64_t s = known < cap ? known : cap;
    uint64_t c = thread_lcp(t, a, b, s, cap);
    *lcp = c;
    if (c < cap) return t[a + c] < t[b + c] ? -1 : 1;
    return boundary_then_pos(L, a, la, b, lb);
}

// LCP from the packed code text (2^lgf codes of 64 >> lgf bits per word):
// a word of codes per side and step (two aligned loads and a funnel shift),
// the first differing code from the leading zeros of their XOR.  Codes past
// the window count as a difference at its end.  Every load stays inside the
// packed buffer, which holds two words past the last code.  When the result
// is below `cap`, *gt tells whether p's code there is above q's: the words
// agree above their first differing code, so they compare like it.
template <typename CM>
__device__ __forceinline__ uint64_t packed_lcp(const CM &cm, uint64_t p, uint64_t q, uint64_t start, uint64_t cap,
                                               bool *gt) {
    const uint32_t lgf = cm.lgf, bits = 64u >> lgf;
    const uint64_t F = 1ull << lgf;
    auto word = [&](uint64_t x) {
        const uint64_t w = x >> lgf;
        const uint32_t rb = (uint32_t)(x & (F - 1)) * bits;
        uint64_t v = cm.packed[w];
        if (rb) v = (v << rb) | (cm.packed[w + 1] >> (64 - rb));
        return v;
    };
    uint64_t o = start;
    while (o < cap) {
        const uint64_t wp = word(p + o), wq = word(q + o);
        uint64_t x = wp ^ wq;
        const uint64_t rem = cap - o;
        if (rem < F) x |= ~0ull >> (rem * bits);  // codes past the window differ
        if (x) {
            const uint64_t r = o + (uint64_t)__builtin_clzll(x) / bits;
            *gt = wp > wq;
            return r < cap ? r : cap;
        }
        // a matched word of one repeated code: both suffixes may sit in
        // recorded runs of that code, which then agree on the shorter rest
        if (cm.runs.n_runs && wp == ((wp << bits) | (wp >> (64 - bits)))) {
            const uint64_t rp = run_remaining(cm.runs, p + o), rq = run_remaining(cm.runs, q + o);
            const uint64_t sk = rp < rq ? rp : rq;
            if (sk > F) {
                o += sk;
                continue;
            }
        }
        o += F;
    }
    return cap;
}

// thread_cmp with the LCP scan over the packed codes (same result); the
// order at the first differing code comes from the packed words as well.
template <typename S>
__device__ __forceinline__ int thread_cmp_packed(const CodeMap<S> &cm, const DevLimits &L, uint64_t max_ctx,
                                                 uint64_t a, uint64_t b, uint64_t known, uint64_t *lcp) {
    uint64_t la = lim_at(L, a), lb = lim_at(L, b);
    uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
    uint64_t cap = ca < cb ? ca : cb;
    uint64_t s = known < cap ? known : cap;
    bool gt = false;
    uint64_t c = packed_lcp(cm, a, b, s, cap, &gt);
    *lcp = c;
    if (c < cap) return gt ? 1 : -1;
    return boundary_then_pos(L, a, la, b, lb);
}

template <typename S>
__device__ __forceinline__ int wave_cmp(const S *__restrict__ t, const DevLimits &L, uint64_t max_ctx, uint64_t a,
                                        uint64_t b, uint64_t known, uint64_t *lcp) {
    uint64_t la = lim_at(L, a), lb = lim_at(L, b);
    uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
    uint64_t cap = ca < cb ? ca : cb;
    uint64_t s = known < cap ? known : cap;
    uint64_t c = wave_lcp(t, a, b, s, cap);
    *lcp = c;
    if (c < cap) return t[a + c] < t[b + c] ? -1 : 1;
    return boundary_then_pos(L, a, la, b, lb);
}

}  // namespace caps
