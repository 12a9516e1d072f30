// Learned bucket router: predicts the splitter rank (bucket) of a packed
// prefix key, so classification needs one model evaluation and two neighbour
// checks instead of a binary search over the splitter table.  The exact
// bucket is always restored (bounded gallop, then binary search), so the
// router affects speed only, never the suffix array.
//
//   PWL: a monotone piecewise-linear CDF.  Segment j covers keys in
//        [knot_j, knot_j+1) (knots at equal-count quantiles of the sorted
//        splitter sample); its line y = a_j * x + b_j over
//        x = (key - knot_j) / (knot_j+1 - knot_j) is fit by least squares on
//        the sample keys of the segment, clamped to the bucket range the
//        segment's keys can take.  A radix table on the top 12 key bits
//        narrows the segment search.
//   MLP: 64 key bits -> 32 ReLU -> 1, output the bucket as a fraction of the
//        bucket count.  Trained by SGD (Adam) with the kernels below on a
//        seeded suffix sample; evaluated through nibble tables (16 nibble
//        positions x 16 values x 32 hidden sums) held in LDS.
//   Spline: a monotone piecewise-linear CDF with knots at 4,096 equal-width
//        cells of the key space, each valued at the exact splitter rank of
//        the cell's first key; in between, linear in the next 16 key bits
//        (integer arithmetic).  Its error bound is exact (k_spline_bound),
//        so classification corrects a prediction with a fixed number of
//        branch-free steps over a window of the sorted splitter table.
#pragma once
#include "device.hpp"

namespace caps {

constexpr uint32_t kRouterRadixBits = 12;
constexpr uint32_t kPwlSeg = 256;
constexpr int kMlpIn = 64, kMlpHid = 32;
// parameter layout: W1[64][32], b1[32], w2[32], b2
constexpr int kMlpParams = kMlpIn * kMlpHid + kMlpHid + kMlpHid + 1;

struct RouterDev {
    uint32_t kind;  // 0 none, 1 PWL, 2 MLP, 3 spline
    uint32_t window;
    // spline: knots (kSplineCells + 1), error bound pred - pos in [-el, eh],
    // search steps (bit length of el + eh)
    const uint16_t *spline;
    int32_t el, eh;
    uint32_t steps;
    // PWL
    const uint64_t *knots;  // ns + 1
    const float *a, *b, *inv;
    const uint32_t *lo, *hi;
    const uint16_t *radix;  // 2^12 + 1
    uint32_t ns;
    // MLP
    const float *nib;  // 16 * 16 * 32 nibble tables (b1 folded into nibble position 0)
    const float *w2;   // 32 output weights, then b2
    float scale;  // bucket count - 1
    uint32_t n_buckets;
    unsigned long long *stats;  // [sum |err|, max |err|, fallbacks, count]
};

// ---- PWL ----------------------------------------------------------------
__device__ __forceinline__ uint32_t pwl_segment(const RouterDev &R, const uint64_t *knots, const uint16_t *radix,
                                                uint64_t key) {
    uint32_t t = (uint32_t)(key >> (64 - kRouterRadixBits));
    uint32_t lo = radix[t], hi = radix[t + 1];
    // last knot <= key within [lo, hi]
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (knots[mid] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo < R.ns ? lo : R.ns - 1;
}

__device__ __forceinline__ uint32_t pwl_predict(const RouterDev &R, const uint64_t *knots, const uint16_t *radix,
                                                const float *a, const float *b, const float *inv,
                                                const uint32_t *lo, const uint32_t *hi, uint64_t key) {
    uint32_t j = pwl_segment(R, knots, radix, key);
    uint64_t d = key > knots[j] ? key - knots[j] : 0;
    float x = (float)d * inv[j];
    float y = a[j] * x + b[j];
    int v = (int)(y + 0.5f);
    if (y < 0.f) v = 0;
    uint32_t u = (uint32_t)v;
    if (u < lo[j]) u = lo[j];
    if (u > hi[j]) u = hi[j];
    return u;
}

// ---- spline -------------------------------------------------------------
// Knot t (0 <= t <= 4096) is the number of splitters below t * 2^52; a key
// of cell t = key >> 52 is placed at knot t plus the fraction f / 2^16 of the
// cell's rank span, f = the next 16 key bits.  The prediction never
// decreases as the key grows (inside a cell f does not; across cells it
// stays below knot t+1, the next cell's first prediction), and the exact
// position (the number of splitters <= key) is a non-decreasing step
// function: on any key range where the position is constant and one cell
// holds every key, the error pred - pos is extreme at the range's ends.  Those
// ends are the splitters, the keys one below them, and the cells' first and
// last keys, where k_spline_bound evaluates it.
constexpr uint32_t kSplineBits = 12;
constexpr uint32_t kSplineCells = 1u << kSplineBits;

__device__ __forceinline__ uint32_t spline_predict(const uint16_t *kn, uint64_t key) {
    const uint32_t t = (uint32_t)(key >> (64 - kSplineBits));
    const uint32_t lo = kn[t], hi = kn[t + 1];
    const uint32_t f = (uint32_t)(key >> (48 - kSplineBits)) & 0xFFFFu;
    return lo + ((f * (hi - lo)) >> 16);
}

// The same prediction from knot pairs kp[t] = knot t | knot t+1 << 16 (one
// LDS read per key in classification).
__device__ __forceinline__ uint32_t spline_predict_pair(const uint32_t *kp, uint64_t key) {
    const uint32_t t = (uint32_t)(key >> (64 - kSplineBits));
    const uint32_t pr = kp[t];
    const uint32_t lo = pr & 0xFFFFu, hi = pr >> 16;
    const uint32_t f = (uint32_t)(key >> (48 - kSplineBits)) & 0xFFFFu;
    return lo + ((f * (hi - lo)) >> 16);
}

// Knots from the sorted splitter table ctab[0..n), n < 2^16.
__global__ void k_spline_knots(const uint64_t *__restrict__ ctab, uint32_t n, uint16_t *__restrict__ kn) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > kSplineCells) return;
    uint32_t lo = 0, hi = n;
    if (t < kSplineCells) {
        const uint64_t k0 = (uint64_t)t << (64 - kSplineBits);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ctab[mid] < k0) lo = mid + 1;
            else hi = mid;
        }
    } else {
        lo = n;
    }
    kn[t] = (uint16_t)lo;
}

// Error bound of the spline: bound[0] = min and bound[1] = max of
// pred(q) - pos(q) over the 2n + 2 * 4096 range ends (bound[] starts at 0).
__global__ void k_spline_bound(const uint64_t *__restrict__ ctab, uint32_t n, const uint16_t *__restrict__ kn,
                               int *__restrict__ bound) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n + 2 * kSplineCells) return;
    uint64_t q;
    if (i < 2 * n) {
        const uint64_t sp = ctab[i >> 1];
        q = (i & 1) ? sp : (sp ? sp - 1 : 0ull);
    } else {
        const uint32_t j = i - 2 * n, t = j >> 1;
        const uint64_t c0 = (uint64_t)t << (64 - kSplineBits);
        q = (j & 1) ? c0 + ((1ull << (64 - kSplineBits)) - 1) : c0;
    }
    uint32_t lo = 0, hi = n;  // exact position: splitters <= q
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ctab[mid] <= q) lo = mid + 1;
        else hi = mid;
    }
    const int e = (int)spline_predict(kn, q) - (int)lo;
    atomicMin(&bound[0], e);
    atomicMax(&bound[1], e);
}

// ---- MLP ----------------------------------------------------------------
// Forward pass from nibble tables: h = relu(sum_q T[q][nibble_q(key)]),
// out = w2 . h + b2.
__device__ __forceinline__ float mlp_forward_nib(const float *nib, const float *w2, float b2, uint64_t key) {
    float h[kMlpHid];
#pragma unroll
    for (int j = 0; j < kMlpHid; ++j) h[j] = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        uint32_t v = (uint32_t)(key >> (60 - 4 * q)) & 15u;
        const float *row = nib + (q * 16 + v) * kMlpHid;
#pragma unroll
        for (int j = 0; j < kMlpHid; ++j) h[j] += row[j];
    }
    float out = b2;
#pragma unroll
    for (int j = 0; j < kMlpHid; ++j) out += w2[j] * (h[j] > 0.f ? h[j] : 0.f);
    return out;
}

// w2 holds the 32 output weights followed by b2.
__device__ __forceinline__ uint32_t mlp_predict(const RouterDev &R, const float *nib, const float *w2, uint64_t key) {
    float y = mlp_forward_nib(nib, w2, w2[kMlpHid], key) * R.scale;
    int v = (int)(y + 0.5f);
    if (y < 0.f) v = 0;
    uint32_t u = (uint32_t)v;
    return u < R.n_buckets ? u : R.n_buckets - 1;
}

// Builds the nibble tables from W1 and b1: T[q][v][j] = sum over the set bits
// of v of W1[4q + bit][j] (+ b1[j] for q = 0).  Input bit i is key bit 63-i.
__global__ void k_mlp_nibbles(const float *__restrict__ params, float *__restrict__ nib) {
    int q = blockIdx.x, v = threadIdx.x >> 5, j = threadIdx.x & 31;  // 16 x (16*32)
    const float *W1 = params, *b1 = params + kMlpIn * kMlpHid;
    float s = q == 0 ? b1[j] : 0.f;
    for (int bit = 0; bit < 4; ++bit)
        if (v & (8 >> bit)) s += W1[(4 * q + bit) * kMlpHid + j];
    nib[(q * 16 + v) * kMlpHid + j] = s;
}

// One gradient pass of the mean squared error over (keys, targets):
// forward per sample, then each thread sums whole parameters over the
// block's samples (no per-sample atomics); grad (zeroed by the caller) and
// loss receive the totals.  `count` is the full batch size (the mean's
// denominator); this launch covers samples [0, n_here).
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params, const uint64_t *__restrict__ keys,
                                                  const float *__restrict__ targets, uint32_t n_here, uint32_t count,
                                                  float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float sp[kMlpParams];
    __shared__ float sd[256][kMlpHid + 1];  // dL/dpre (0 where relu is off)
    __shared__ float sh[256][kMlpHid + 1];  // relu(pre)
    __shared__ float so[256];               // dL/dout
    __shared__ uint64_t sk[256];
    __shared__ float sl[256];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
    const float *W1 = sp, *b1 = sp + kMlpIn * kMlpHid, *w2 = b1 + kMlpHid;
    const float b2 = w2[kMlpHid];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int tid = threadIdx.x;
    if (i < n_here) {
        uint64_t key = keys[i];
        float pre[kMlpHid];
#pragma unroll
        for (int j = 0; j < kMlpHid; ++j) pre[j] = b1[j];
        for (int bit = 0; bit < kMlpIn; ++bit)
            if ((key >> (63 - bit)) & 1ull) {
#pragma unroll
                for (int j = 0; j < kMlpHid; ++j) pre[j] += W1[bit * kMlpHid + j];
            }
        float out = b2;
#pragma unroll
        for (int j = 0; j < kMlpHid; ++j) out += w2[j] * (pre[j] > 0.f ? pre[j] : 0.f);
        float err = out - targets[i];
        float d = 2.f * err / (float)count;
        so[tid] = d;
        sk[tid] = key;
        sl[tid] = err * err / (float)count;
#pragma unroll
        for (int j = 0; j < kMlpHid; ++j) {
            sh[tid][j] = pre[j] > 0.f ? pre[j] : 0.f;
            sd[tid][j] = pre[j] > 0.f ? d * w2[j] : 0.f;
        }
    } else {
        so[tid] = 0.f;
        sk[tid] = 0;
        sl[tid] = 0.f;
#pragma unroll
        for (int j = 0; j < kMlpHid; ++j) sh[tid][j] = 0.f, sd[tid][j] = 0.f;
    }
    __syncthreads();
    for (int p = tid; p < kMlpParams; p += blockDim.x) {
        float s = 0.f;
        if (p < kMlpIn * kMlpHid) {  // W1[bit][j]
            int bit = p / kMlpHid, j = p % kMlpHid;
            for (int q = 0; q < 256; ++q)
                if ((sk[q] >> (63 - bit)) & 1ull) s += sd[q][j];
        } else if (p < kMlpIn * kMlpHid + kMlpHid) {  // b1[j]
            int j = p - kMlpIn * kMlpHid;
            for (int q = 0; q < 256; ++q) s += sd[q][j];
        } else if (p < kMlpIn * kMlpHid + 2 * kMlpHid) {  // w2[j]
            int j = p - kMlpIn * kMlpHid - kMlpHid;
            for (int q = 0; q < 256; ++q) s += so[q] * sh[q][j];
        } else {  // b2
            for (int q = 0; q < 256; ++q) s += so[q];
        }
        if (s != 0.f) atomicAdd(&grad[p], s);
    }
    if (loss && tid == 0) {
        float s = 0.f;
        for (int q = 0; q < 256; ++q) s += sl[q];
        atomicAdd(loss, s);
    }
}

// Adam update (Kingma & Ba 2015) of the parameters from grad.
__global__ void k_mlp_adam(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                           float *__restrict__ v, float lr, float b1p, float b2p) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float g = grad[p];
    float mm = beta1 * m[p] + (1.f - beta1) * g;
    float vv = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float mh = mm / (1.f - b1p), vh = vv / (1.f - b2p);
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Exact bucket labels of sorted keys (number of splitters <= key).
__global__ void k_router_labels(const uint64_t *__restrict__ keys, uint64_t count, const uint64_t *__restrict__ spl,
                                uint32_t n_spl, uint32_t *__restrict__ labels, float *__restrict__ ftarget,
                                float inv_scale) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t key = keys[i];
    uint32_t lo = 0, hi = n_spl;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (spl[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    if (labels) labels[i] = lo;
    if (ftarget) ftarget[i] = (float)lo * inv_scale;
}

// PWL fit: one workgroup per segment j over the sorted sample range
// [j*s/ns, (j+1)*s/ns): least squares of label on x.
__global__ __launch_bounds__(256) void k_pwl_fit(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ labels,
                                                 uint64_t s, uint32_t ns, uint32_t n_buckets,
                                                 uint64_t *__restrict__ knots, float *__restrict__ a,
                                                 float *__restrict__ b, float *__restrict__ inv,
                                                 uint32_t *__restrict__ lo, uint32_t *__restrict__ hi) {
    __shared__ double red[5][256];
    uint32_t j = blockIdx.x;
    uint64_t i0 = (uint64_t)j * s / ns, i1 = (uint64_t)(j + 1) * s / ns;
    uint64_t k0 = keys[i0];
    uint64_t k1 = (j + 1 < ns) ? keys[i1] : keys[s - 1];
    double span = k1 > k0 ? (double)(k1 - k0) : 1.0;
    float invf = (float)(1.0 / span);
    double sx = 0, sy = 0, sxx = 0, sxy = 0, cnt = 0;
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        double x = (double)((float)(keys[i] - k0) * invf);
        double y = labels[i];
        sx += x, sy += y, sxx += x * x, sxy += x * y, cnt += 1;
    }
    red[0][threadIdx.x] = sx, red[1][threadIdx.x] = sy, red[2][threadIdx.x] = sxx, red[3][threadIdx.x] = sxy,
    red[4][threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 5; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double n = red[4][0], Sx = red[0][0], Sy = red[1][0], Sxx = red[2][0], Sxy = red[3][0];
        double den = n * Sxx - Sx * Sx;
        double sl = 0, ic = n > 0 ? Sy / n : 0;
        if (n > 1 && den > 1e-12 * n * n) {
            sl = (n * Sxy - Sx * Sy) / den;
            ic = (Sy - sl * Sx) / n;
        }
        if (sl < 0) {  // keep the model monotone
            sl = 0;
            ic = n > 0 ? Sy / n : 0;
        }
        knots[j] = k0;
        if (j + 1 == ns) knots[ns] = ~0ull;
        a[j] = (float)sl;
        b[j] = (float)ic;
        inv[j] = invf;
        lo[j] = labels[i0];
        hi[j] = (j + 1 < ns) ? labels[i1] : n_buckets - 1;
    }
}

// Radix table: segment range for each top-12-bit prefix.
__global__ void k_pwl_radix(const uint64_t *__restrict__ knots, uint32_t ns, uint16_t *__restrict__ radix) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > (1u << kRouterRadixBits)) return;
    uint64_t prefix = t == (1u << kRouterRadixBits) ? ~0ull : ((uint64_t)t << (64 - kRouterRadixBits));
    // last knot <= prefix (segment holding the smallest key with this prefix)
    uint32_t lo = 0, hi = ns - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (knots[mid] <= prefix) lo = mid;
        else hi = mid - 1;
    }
    radix[t] = (uint16_t)lo;
}

// Forward pass of `count` keys through the nibble-table path (tests).
__global__ void k_mlp_predict(const float *__restrict__ nib, const float *__restrict__ w2,
                              const uint64_t *__restrict__ keys, uint32_t count, float *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = mlp_forward_nib(nib, w2, w2[kMlpHid], keys[i]);
}

}  // namespace caps
