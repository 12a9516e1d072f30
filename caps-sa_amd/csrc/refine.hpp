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

constexpr int kGroupClasses = 8;
constexpr uint32_t kMidMax = 2048;

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

// ---------------------------------------------------------------------------
// Group detection inside key-sorted ranges (buckets larger than one tile,
// after their tiles were merged).  Appends are aggregated per wavefront.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_append(GroupLists &gl, bool have, int cls, Group gr) {
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
            gl.list[c][base + __popcll(below)] = gr;
        }
    }
}

struct Range;  // defined in kernels.hpp

template <typename RangeT>
__global__ __launch_bounds__(256) void k_detect_ranges(const uint64_t *__restrict__ keys, const RangeT *__restrict__ ranges,
                                                       GroupLists gl) {
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
                uint64_t j = o + 2;
                while (j < r.len && keys[r.start + j] == k0) ++j;
                gr = Group{i, j - o};
                have = true;
            }
        }
        wave_append(gl, have, have ? group_class(gr.size) : 0, gr);
    }
}

// ---------------------------------------------------------------------------
// Two-member groups: 16 lanes per pair, 4 symbols per lane per step (64
// symbols), four pairs per wavefront; ballot + ffs find the first mismatch.
// ---------------------------------------------------------------------------
template <typename S, typename I>
__global__ __launch_bounds__(256) void k_pairs16(const S *__restrict__ t, DevLimits L, uint64_t max_ctx,
                                                 const Group *__restrict__ pairs, uint32_t n_pairs, uint64_t known,
                                                 I *__restrict__ idx, I *__restrict__ lcp) {
    const uint32_t lane = threadIdx.x & 63, sub = lane >> 4, sl = lane & 15;
    const uint32_t pid = blockIdx.x * 16 + (threadIdx.x >> 4);
    const bool valid = pid < n_pairs;
    uint64_t s = 0, a = 0, b = 0, la = 0, lb = 0, cap = 0;
    if (valid) {
        s = pairs[pid].start;
        a = idx[s];
        b = idx[s + 1];
        la = lim_at(L, a);
        lb = lim_at(L, b);
        uint64_t ca = la < max_ctx ? la : max_ctx, cb = lb < max_ctx ? lb : max_ctx;
        cap = ca < cb ? ca : cb;
    }
    uint64_t o = known < cap ? known : cap;
    bool done = !valid || o >= cap;
    uint64_t res = cap;
    while (__any(!done)) {
        uint32_t firstbad = 4;
        if (!done) {
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                uint64_t oo = o + 4 * sl + q;
                bool mism = oo >= cap || t[a + oo] != t[b + oo];
                if (mism && firstbad == 4) firstbad = q;
            }
        }
        uint64_t mask = __ballot(!done && firstbad < 4);
        uint32_t gm = (uint32_t)(mask >> (sub * 16)) & 0xFFFFu;
        int fl = gm ? __ffs(gm) - 1 : 0;
        uint32_t fq = __shfl(firstbad, (int)(sub * 16 + fl));
        if (!done) {
            if (gm) {
                uint64_t r = o + 4 * (uint64_t)fl + fq;
                res = r < cap ? r : cap;
                done = true;
            } else {
                o += 64;
            }
        }
    }
    if (valid && sl == 0) {
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

template <typename S, typename I, int W>
__global__ __launch_bounds__(256) void k_wgroups(const S *__restrict__ t, CodeMap<S> cm, KeyGeom g, DevLimits L,
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
    for (;;) {
        if (!settled) {
            chunk = key_at(t, cm, g, pos, D, cap);
            uint64_t rem = cap > D ? cap - D : 0;
            if (rem >= g.k) aux = cont;
            else aux = star ? (1 + (kAuxCont - 1 - lim)) : lim;
        }
        // bitonic sort of the W-lane segment
#pragma unroll
        for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                uint32_t p_sub = __shfl_xor(sub, stride);
                uint64_t p_chunk = shfl_xor64(chunk, stride);
                uint64_t p_aux = shfl_xor64(aux, stride);
                uint64_t p_pos = shfl_xor64(pos, stride);
                uint64_t p_lim = shfl_xor64(lim, stride);
                uint64_t p_cap = shfl_xor64(cap, stride);
                int p_settled = __shfl_xor((int)settled, stride);
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
        int prev = lane - 1 < 0 ? 0 : lane - 1;
        uint32_t q_sub = __shfl(sub, prev);
        uint64_t q_chunk = shfl64(chunk, prev);
        uint64_t q_aux = shfl64(aux, prev);
        bool same_old = r > 0 && r < m && sub == q_sub;
        bool f = !(same_old && chunk == q_chunk && aux == q_aux);
        if (same_old && !settled) h = D + shared_fields(g, q_chunk, chunk);
        // new subgroup id: inclusive max-scan of (f ? r : 0) within the segment
        uint32_t ns = f ? r : 0;
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            uint32_t v = __shfl_up(ns, o, W);
            if ((int)r >= o) ns = v > ns ? v : ns;
        }
        uint32_t ns_prev = __shfl(ns, prev);
        uint32_t ns_next = __shfl(ns, lane + 1 < 64 ? lane + 1 : lane);
        bool nb = (r > 0 && r < m && ns_prev == ns) || (r + 1 < m && ns_next == ns);
        bool ended = aux != cont;
        if (act) {
            settled = settled || !nb || ended;
            sub = ns;
        }
        if (!__any(!settled)) break;
        D += g.k;
    }
    uint64_t q_cap = shfl64(cap, lane - 1 < 0 ? 0 : lane - 1);
    if (act) {
        idx[gr.start + r] = (I)pos;
        if (lcp && r > 0) {
            uint64_t v = h;
            v = v < cap ? v : cap;
            v = v < q_cap ? v : q_cap;
            lcp[gr.start + r] = (I)v;
        }
    }
}

}  // namespace caps
