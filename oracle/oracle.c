/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h for who may load it).
 *
 * Restates caps-sa 0.7.0's in-memory suffix-array algorithm in plain C.  Every
 * function names the reference file:line it follows.  Parallel parts use
 * OpenMP in place of rayon's fork-join; the arithmetic is unchanged.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }
static inline uint64_t max64(uint64_t a, uint64_t b) { return a > b ? a : b; }

static inline uint32_t sym_at(const orc_text *T, uint64_t i) {
    return T->sym_bytes == 1 ? ((const uint8_t *)T->data)[i] : ((const uint16_t *)T->data)[i];
}

/* ------------------------------------------------------------------------ */
/* limits.rs                                                                 */
/* ------------------------------------------------------------------------ */

/* limits.rs:194-236 (BoundaryDirectory::build). */
int orc_limits_init(orc_limits *L, uint64_t n, const uint64_t *ends, uint64_t n_ends, int order) {
    memset(L, 0, sizeof(*L));
    L->n = n;
    L->ends = ends;
    L->n_ends = n_ends;
    L->order = order;
    if (n_ends < 256 || n_ends > 0xFFFFFFFFull || n == 0) return 0;
    uint64_t target_blocks = n_ends * 2;
    if (target_blocks > 2000000) target_blocks = 2000000;
    if (target_blocks < 1) target_blocks = 1;
    uint64_t min_block = (n + target_blocks - 1) / target_blocks;
    uint64_t bs = 1;
    while (bs < min_block) bs <<= 1;
    uint32_t shift = (uint32_t)__builtin_ctzll(bs);
    uint64_t n_blocks = (n + bs - 1) / bs;
    uint32_t *fa = (uint32_t *)malloc((n_blocks + 1) * sizeof(uint32_t));
    if (!fa) return -1;
    uint64_t e = 0;
    for (uint64_t b = 0; b <= n_blocks; b++) {
        uint64_t start = min64(b * bs, n);
        while (e < n_ends && ends[e] <= start) e++;
        fa[b] = (uint32_t)e;
    }
    L->has_dir = 1;
    L->block_shift = shift;
    L->first_after_block_start = fa;
    L->n_dir = n_blocks + 1;
    return 0;
}

void orc_limits_free(orc_limits *L) {
    free(L->first_after_block_start);
    L->first_after_block_start = NULL;
    L->has_dir = 0;
}

/* First index i in ends[lo..hi) with ends[i] > p (partition_point of b <= p). */
static inline uint64_t pp_le(const uint64_t *ends, uint64_t lo, uint64_t hi, uint64_t p) {
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (ends[mid] <= p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* limits.rs:290-303 (SegmentedText::lim_at fallback) and 144-148 (PlainText). */
uint64_t orc_lim_at_bsearch(const orc_limits *L, uint64_t p) {
    if (L->n_ends == 0) return L->n - p;
    uint64_t i = pp_le(L->ends, 0, L->n_ends, p);
    return i < L->n_ends ? L->ends[i] - p : L->n - p;
}

/* limits.rs:277-303 (SegmentedText::lim_at with the directory). */
uint64_t orc_lim_at(const orc_limits *L, uint64_t p) {
    if (L->n_ends == 0) return L->n - p;
    if (L->has_dir && p < L->n) {
        uint64_t block = p >> L->block_shift;
        uint64_t lo = L->first_after_block_start[block];
        uint64_t hi = L->first_after_block_start[block + 1];
        hi = min64(max64(hi, lo + 1), L->n_ends);
        uint64_t i = pp_le(L->ends, lo, hi, p);
        return L->ends[i] - p;
    }
    return orc_lim_at_bsearch(L, p);
}

static inline int cmp_u64(uint64_t a, uint64_t b) { return (a > b) - (a < b); }

/* limits.rs:66-75 (default boundary_order); STAR: sample_sort.rs:1006-1014. */
int orc_boundary_order(const orc_limits *L, uint64_t pa, uint64_t lim_a, uint64_t pb, uint64_t lim_b) {
    if (L->order == ORC_ORDER_STAR) {
        int c = cmp_u64(lim_b, lim_a);
        return c ? c : cmp_u64(pa, pb);
    }
    return cmp_u64(lim_a, lim_b);
}

/* ------------------------------------------------------------------------ */
/* lcp.rs                                                                    */
/* ------------------------------------------------------------------------ */

/* lcp.rs:125-153: byte-view LCP divided by the symbol width.  Eight bytes per
 * step (xor + count-trailing-zeros) stands in for the AVX2/AVX-512 stride of
 * lcp.rs:292-393; lcp_scalar (lcp.rs:233-246) gives the same answer. */
uint64_t orc_lcp(const orc_text *T, uint64_t p, uint64_t q, uint64_t max_ctx) {
    uint64_t n = T->n;
    uint64_t lim_p = n > p ? n - p : 0, lim_q = n > q ? n - q : 0;
    uint64_t lim = min64(min64(lim_p, lim_q), max_ctx);
    uint64_t k = T->sym_bytes;
    const uint8_t *a = (const uint8_t *)T->data + p * k;
    const uint8_t *b = (const uint8_t *)T->data + q * k;
    uint64_t limb = lim * k, i = 0;
    while (i + 8 <= limb) {
        uint64_t x, y;
        memcpy(&x, a + i, 8);
        memcpy(&y, b + i, 8);
        if (x != y) return (i + (uint64_t)(__builtin_ctzll(x ^ y) >> 3)) / k;
        i += 8;
    }
    while (i < limb && a[i] == b[i]) i++;
    return i / k;
}

/* lcp.rs:184-201. */
int orc_suffix_cmp_with(const orc_text *T, const orc_limits *L, uint64_t p, uint64_t q, uint64_t max_ctx) {
    uint64_t lim_p = orc_lim_at(L, p), lim_q = orc_lim_at(L, q);
    uint64_t lim = min64(min64(lim_p, lim_q), max_ctx);
    uint64_t common = orc_lcp(T, p, q, lim);
    if (common < lim) {
        uint32_t a = sym_at(T, p + common), b = sym_at(T, q + common);
        return (a > b) - (a < b);
    }
    return orc_boundary_order(L, p, lim_p, q, lim_q);
}

/* ------------------------------------------------------------------------ */
/* sample_sort.rs                                                            */
/* ------------------------------------------------------------------------ */

/* sample_sort.rs:614-636. */
static void drain(const uint64_t *arr, const uint64_t *lcp_src, uint64_t i, uint64_t len, uint64_t *z,
                  uint64_t *lcp_z, uint64_t *k, uint64_t boundary_m) {
    int first = 1;
    while (i < len) {
        z[*k] = arr[i];
        lcp_z[*k] = first ? boundary_m : lcp_src[i];
        first = 0;
        i++;
        (*k)++;
    }
}

/* sample_sort.rs:397-502 (merge_body!, direct extension 359-361). */
void orc_merge(const orc_text *T, const orc_limits *L, const uint64_t *x, uint64_t nx, const uint64_t *y,
               uint64_t ny, const uint64_t *lcp_x, const uint64_t *lcp_y, uint64_t *z, uint64_t *lcp_z,
               uint64_t max_ctx) {
    if (nx == 0) {
        memcpy(z, y, ny * 8);
        memcpy(lcp_z, lcp_y, ny * 8);
        return;
    }
    if (ny == 0) {
        memcpy(z, x, nx * 8);
        memcpy(lcp_z, lcp_x, nx * 8);
        return;
    }
    const uint64_t *arr_a = x, *arr_b = y, *lcp_a = lcp_x, *lcp_b = lcp_y;
    uint64_t len_a = nx, len_b = ny, i_a = 0, i_b = 0, m = 0, k = 0;
    while (i_a < len_a && i_b < len_b) {
        uint64_t l_a = lcp_a[i_a];
        int output_a;
        uint64_t lcp_out, new_m;
        if (l_a > m) {
            output_a = 1, lcp_out = l_a, new_m = m;
        } else if (l_a < m) {
            output_a = 0, lcp_out = m, new_m = l_a;
        } else {
            uint64_t p_a = arr_a[i_a], p_b = arr_b[i_b];
            uint64_t lim_a = orc_lim_at(L, p_a), lim_b = orc_lim_at(L, p_b);
            uint64_t cap = min64(min64(lim_a, lim_b), max_ctx);
            uint64_t remaining = cap > m ? cap - m : 0;
            uint64_t ext = orc_lcp(T, p_a + m, p_b + m, remaining);
            uint64_t total = m + ext;
            int a_smaller;
            if (total < cap) a_smaller = sym_at(T, p_a + total) < sym_at(T, p_b + total);
            else a_smaller = orc_boundary_order(L, p_a, lim_a, p_b, lim_b) < 0;
            output_a = a_smaller, lcp_out = m, new_m = total;
        }
        if (output_a) {
            z[k] = arr_a[i_a];
            lcp_z[k] = lcp_out;
            i_a++;
        } else {
            z[k] = arr_b[i_b];
            lcp_z[k] = lcp_out;
            i_b++;
            const uint64_t *t;
            uint64_t u;
            t = arr_a, arr_a = arr_b, arr_b = t;
            t = lcp_a, lcp_a = lcp_b, lcp_b = t;
            u = len_a, len_a = len_b, len_b = u;
            u = i_a, i_a = i_b, i_b = u;
        }
        m = new_m;
        k++;
    }
    drain(arr_a, lcp_a, i_a, len_a, z, lcp_z, &k, m);
    drain(arr_b, lcp_b, i_b, len_b, z, lcp_z, &k, m);
}

/* sample_sort.rs:221-266: top-down split at n/2, merge into the workspace,
 * copy back.  rayon::join becomes an OpenMP task pair above a cutoff. */
static void merge_sort_rec(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t *sa_w, uint64_t *lcp,
                           uint64_t *lcp_w, uint64_t n, uint64_t max_ctx, int par) {
    if (n <= 1) {
        if (n == 1) lcp[0] = 0;
        return;
    }
    uint64_t mid = n / 2;
    if (par && n > 8192) {
#pragma omp task default(shared)
        merge_sort_rec(T, L, sa, sa_w, lcp, lcp_w, mid, max_ctx, par);
#pragma omp task default(shared)
        merge_sort_rec(T, L, sa + mid, sa_w + mid, lcp + mid, lcp_w + mid, n - mid, max_ctx, par);
#pragma omp taskwait
    } else {
        merge_sort_rec(T, L, sa, sa_w, lcp, lcp_w, mid, max_ctx, 0);
        merge_sort_rec(T, L, sa + mid, sa_w + mid, lcp + mid, lcp_w + mid, n - mid, max_ctx, 0);
    }
    orc_merge(T, L, sa, mid, sa + mid, n - mid, lcp, lcp + mid, sa_w, lcp_w, max_ctx);
    memcpy(sa, sa_w, n * 8);
    memcpy(lcp, lcp_w, n * 8);
}

void orc_merge_sort(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t *sa_w, uint64_t *lcp,
                    uint64_t *lcp_w, uint64_t n, uint64_t max_ctx, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 1 && n > 8192) {
#pragma omp parallel num_threads(nthreads)
#pragma omp single
        merge_sort_rec(T, L, sa, sa_w, lcp, lcp_w, n, max_ctx, 1);
        return;
    }
#endif
    (void)nthreads;
    merge_sort_rec(T, L, sa, sa_w, lcp, lcp_w, n, max_ctx, 0);
}

/* sample_sort.rs:168-206. */
void orc_build_in_memory(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx,
                         uint64_t *lcp_out, int nthreads) {
    if (n == 0) return;
    uint64_t *sa_w = (uint64_t *)malloc(n * 8), *lcp = (uint64_t *)malloc(n * 8),
             *lcp_w = (uint64_t *)malloc(n * 8);
    orc_merge_sort(T, L, sa, sa_w, lcp, lcp_w, n, max_ctx, nthreads);
    if (lcp_out) memcpy(lcp_out, lcp, n * 8);
    free(sa_w);
    free(lcp);
    free(lcp_w);
}

/* ------------------------------------------------------------------------ */
/* ext_mem.rs: cascade merge (phase 4)                                       */
/* ------------------------------------------------------------------------ */

/* ext_mem.rs:2194-2243 and 2251-2389. */
void orc_cascade_merge(const orc_text *T, const orc_limits *L, const uint64_t *sa, const uint64_t *lcp,
                       uint64_t n, const uint64_t *boundaries, uint64_t n_boundaries, uint64_t max_ctx,
                       uint64_t *out_sa, uint64_t *out_lcp) {
    if (n == 0) return;
    uint64_t *a_sa = (uint64_t *)malloc(n * 8), *a_lcp = (uint64_t *)malloc(n * 8);
    uint64_t *b_sa = (uint64_t *)malloc(n * 8), *b_lcp = (uint64_t *)malloc(n * 8);
    memcpy(a_sa, sa, n * 8);
    memcpy(a_lcp, lcp, n * 8);
    uint64_t *runs = (uint64_t *)malloc((n_boundaries + 1) * 8), nr = 0;
    for (uint64_t i = 0; i + 1 < n_boundaries; i++) {
        uint64_t l = boundaries[i + 1] - boundaries[i];
        if (l > 0) runs[nr++] = l;
    }
    int src_is_a = 1;
    while (nr > 1) {
        uint64_t *ssa = src_is_a ? a_sa : b_sa, *slcp = src_is_a ? a_lcp : b_lcp;
        uint64_t *dsa = src_is_a ? b_sa : a_sa, *dlcp = src_is_a ? b_lcp : a_lcp;
        uint64_t so = 0, d = 0, nn = 0;
        for (uint64_t i = 0; i < nr;) {
            uint64_t l1 = runs[i];
            if (i + 1 < nr) {
                uint64_t l2 = runs[i + 1];
                orc_merge(T, L, ssa + so, l1, ssa + so + l1, l2, slcp + so, slcp + so + l1, dsa + d, dlcp + d,
                          max_ctx);
                runs[nn++] = l1 + l2;
                so += l1 + l2;
                d += l1 + l2;
                i += 2;
            } else {
                memcpy(dsa + d, ssa + so, l1 * 8);
                memcpy(dlcp + d, slcp + so, l1 * 8);
                runs[nn++] = l1;
                so += l1;
                d += l1;
                i += 1;
            }
        }
        nr = nn;
        src_is_a = !src_is_a;
    }
    memcpy(out_sa, src_is_a ? a_sa : b_sa, n * 8);
    if (out_lcp) memcpy(out_lcp, src_is_a ? a_lcp : b_lcp, n * 8);
    free(a_sa), free(a_lcp), free(b_sa), free(b_lcp), free(runs);
}

/* ------------------------------------------------------------------------ */
/* ext_mem.rs: in-memory sample sort (phases 1-4)                            */
/* ------------------------------------------------------------------------ */

/* ext_mem.rs:1159-1178 with PHASE1_TARGET_CHUNK = 65,536 and
 * PHASE1_MAX_PARTITIONS = 8,192 (ext_mem.rs:1132-1136). */
uint64_t orc_effective_subproblem_count(uint64_t n, uint64_t requested, uint64_t nthreads) {
    if (n == 0) return 0;
    uint64_t raw = requested;
    if (requested == 0) {
        uint64_t t = nthreads ? nthreads : 1;
        uint64_t from_size = (n + 65535) / 65536;
        raw = from_size < t ? t : (from_size > 8192 ? 8192 : from_size);
    }
    if (raw < 1) raw = 1;
    if (raw > n) raw = n;
    return raw;
}

/* ext_mem.rs:1282-1287. */
uint64_t orc_sample_target_total(uint64_t n, uint64_t p) {
    double ln_n = log((double)n);
    if (ln_n < 1.0) ln_n = 1.0;
    uint64_t per = (uint64_t)ceil(4.0 * ln_n);
    uint64_t t = p * per; /* no overflow at the sizes tested */
    if (t < p) t = p;
    if (t > n) t = n;
    return t;
}

/* ext_mem.rs:1291-1305. */
uint64_t orc_evenly_spaced(const uint64_t *xs, uint64_t n, uint64_t count, uint64_t *out) {
    if (count == 0 || n == 0) return 0;
    if (count >= n) {
        memcpy(out, xs, n * 8);
        return n;
    }
    for (uint64_t i = 0; i < count; i++) out[i] = xs[(2 * i + 1) * n / (2 * count)];
    return count;
}

/* ext_mem.rs:1705-1734. */
uint64_t orc_upper_bound_by_pivot(const orc_text *T, const orc_limits *L, const uint64_t *recs, uint64_t n,
                                  uint64_t pivot, uint64_t max_ctx) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (orc_suffix_cmp_with(T, L, recs[mid], pivot, max_ctx) > 0) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

typedef struct {
    uint64_t *sa, *lcp, len;
} run_t;

typedef struct {
    uint64_t *sa, *lcp; /* records appended in distribution order */
    uint64_t len, cap;
    uint64_t *bounds; /* InMemBucket::boundaries, starting [0] */
    uint64_t nb, bcap;
} part_t;

static void part_push(part_t *P, const uint64_t *sa, const uint64_t *lcp, uint64_t len) {
    if (len == 0) return;
    if (P->len + len > P->cap) {
        uint64_t c = P->cap ? P->cap : 64;
        while (c < P->len + len) c *= 2;
        P->sa = (uint64_t *)realloc(P->sa, c * 8);
        P->lcp = (uint64_t *)realloc(P->lcp, c * 8);
        P->cap = c;
    }
    memcpy(P->sa + P->len, sa, len * 8);
    memcpy(P->lcp + P->len, lcp, len * 8);
    P->lcp[P->len] = 0; /* add_slice_reset_first_lcp (ext_bucket.rs:203-211) */
    P->len += len;
    /* mark_boundary (ext_bucket.rs:169-175) */
    if (P->nb + 1 > P->bcap) {
        P->bcap = P->bcap ? P->bcap * 2 : 16;
        P->bounds = (uint64_t *)realloc(P->bounds, P->bcap * 8);
    }
    if (P->bounds[P->nb - 1] != P->len) P->bounds[P->nb++] = P->len;
}

/* ext_mem.rs:651-702; phase 1 1191-1277, phase 2 1309-1342, phase 3
 * 1361-1428, phase 4 1745-1885 / 2046-2108.  Phase 3 appends the pieces of
 * subarray i before those of subarray i+1 (the reference appends under a
 * mutex in completion order; for plain texts the output is the same). */
int orc_build_in_memory_sample_sort(const orc_text *T, const orc_limits *L, const uint64_t *positions,
                                    uint64_t n_pos, uint64_t subproblem_count, uint64_t max_ctx, int nthreads,
                                    orc_emit_fn emit, void *ctx) {
    uint64_t n = n_pos;
    if (n == 0) return 0;
    if (nthreads < 1) nthreads = 1;
    uint64_t p = orc_effective_subproblem_count(n, subproblem_count, (uint64_t)nthreads);
    uint64_t chunk = (n + p - 1) / p;
    uint64_t target = orc_sample_target_total(n, p);
    uint64_t per_sub = (target + p - 1) / p;
    run_t *subs = (run_t *)calloc(p, sizeof(run_t));
    uint64_t *samp_cnt = (uint64_t *)calloc(p, 8);
    uint64_t *samp = (uint64_t *)malloc(p * per_sub * 8 + 8);

    /* phase 1 */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (uint64_t i = 0; i < p; i++) {
        uint64_t start = min64(i * chunk, n), end = min64((i + 1) * chunk, n), len = end - start;
        subs[i].len = len;
        if (len == 0) continue;
        uint64_t *sa = (uint64_t *)malloc(len * 8), *lcp = (uint64_t *)malloc(len * 8);
        for (uint64_t j = 0; j < len; j++) sa[j] = positions ? positions[start + j] : start + j;
        uint64_t *sa_w = (uint64_t *)malloc(len * 8), *lcp_w = (uint64_t *)malloc(len * 8);
        orc_merge_sort(T, L, sa, sa_w, lcp, lcp_w, len, max_ctx, 1);
        free(sa_w), free(lcp_w);
        subs[i].sa = sa, subs[i].lcp = lcp;
        uint64_t spc = min64(per_sub, len);
        samp_cnt[i] = orc_evenly_spaced(sa, len, spc, samp + i * per_sub);
    }
    uint64_t ns = 0;
    for (uint64_t i = 0; i < p; i++) {
        memmove(samp + ns, samp + i * per_sub, samp_cnt[i] * 8);
        ns += samp_cnt[i];
    }

    /* phase 2 */
    uint64_t n_piv = 0;
    uint64_t *piv = NULL;
    if (p > 1 && ns > 0) {
        uint64_t *w = (uint64_t *)malloc(ns * 8), *l = (uint64_t *)malloc(ns * 8), *lw = (uint64_t *)malloc(ns * 8);
        orc_merge_sort(T, L, samp, w, l, lw, ns, max_ctx, nthreads);
        free(w), free(l), free(lw);
        n_piv = p - 1;
        piv = (uint64_t *)malloc(n_piv * 8);
        for (uint64_t j = 1; j < p; j++) piv[j - 1] = samp[(j * ns) / p];
    }

    /* phase 3: splits per subarray (parallel), appended in subarray order */
    uint64_t *splits = (uint64_t *)malloc(p * (n_piv + 2) * 8);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (uint64_t i = 0; i < p; i++) {
        uint64_t *s = splits + i * (n_piv + 2);
        s[0] = 0;
        for (uint64_t j = 0; j < n_piv; j++)
            s[j + 1] = subs[i].len ? orc_upper_bound_by_pivot(T, L, subs[i].sa, subs[i].len, piv[j], max_ctx) : 0;
        s[n_piv + 1] = subs[i].len;
    }
    uint64_t n_parts = n_piv + 1;
    part_t *parts = (part_t *)calloc(n_parts, sizeof(part_t));
    for (uint64_t j = 0; j < n_parts; j++) {
        parts[j].bcap = 16;
        parts[j].bounds = (uint64_t *)malloc(16 * 8);
        parts[j].bounds[0] = 0;
        parts[j].nb = 1;
    }
    for (uint64_t i = 0; i < p; i++) {
        if (subs[i].len == 0) continue;
        uint64_t *s = splits + i * (n_piv + 2);
        for (uint64_t j = 0; j < n_parts; j++) {
            uint64_t lo = s[j], hi = s[j + 1];
            if (lo >= hi) continue;
            part_push(&parts[j], subs[i].sa + lo, subs[i].lcp + lo, hi - lo);
        }
        free(subs[i].sa), free(subs[i].lcp);
    }

    /* phase 4: per-partition cascade merge (parallel), emitted in order */
    uint64_t **out = (uint64_t **)calloc(n_parts, sizeof(uint64_t *));
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (uint64_t j = 0; j < n_parts; j++) {
        if (parts[j].len == 0) continue;
        out[j] = (uint64_t *)malloc(parts[j].len * 8);
        orc_cascade_merge(T, L, parts[j].sa, parts[j].lcp, parts[j].len, parts[j].bounds, parts[j].nb, max_ctx,
                          out[j], NULL);
    }
    int rc = 0;
    for (uint64_t j = 0; j < n_parts; j++) {
        if (!rc && parts[j].len && emit) rc = emit(ctx, out[j], parts[j].len);
        free(out[j]);
        free(parts[j].sa), free(parts[j].lcp), free(parts[j].bounds);
    }
    free(out), free(parts), free(splits), free(piv), free(samp), free(samp_cnt), free(subs);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* pack.rs                                                                   */
/* ------------------------------------------------------------------------ */

/* pack.rs:103-167 (Packer::new) and 217-234 (seed_params for S == u8). */
int orc_seed_params(const uint8_t *text, uint64_t n, int policy, uint64_t remap_budget, orc_packer *P) {
    memset(P, 0, sizeof(*P));
    if (policy == ORC_SEED_DISABLED) return 0;
    int seen[256] = {0};
    for (uint64_t i = 0; i < n; i++) seen[text[i]] = 1;
    uint32_t next = 0;
    int identity = 1;
    for (int b = 0; b < 256; b++)
        if (seen[b]) {
            P->code[b] = (uint8_t)next;
            identity &= (next == (uint32_t)b);
            next++;
        }
    P->bits = next <= 1 ? 1 : next <= 3 ? 2 : next <= 15 ? 4 : 8;
    if (!identity) {
        if (policy != ORC_SEED_REMAP) return 0;
        if (n > remap_budget) return 0;
        P->ranked = (uint8_t *)malloc(n ? n : 1);
        if (!P->ranked) return 0;
        for (uint64_t i = 0; i < n; i++) P->ranked[i] = P->code[text[i]];
    }
    P->k = 64 / P->bits;
    P->alphabet = next;
    /* pack.rs:174-176: a sentinel code must fit beside the alphabet */
    if (!((uint64_t)P->alphabet < (1ull << P->bits))) {
        orc_packer_free(P);
        return 0;
    }
    return 1;
}

void orc_packer_free(orc_packer *P) {
    free(P->ranked);
    P->ranked = NULL;
}

/* pack.rs:187-203. */
uint64_t orc_key_at_bounded(const orc_packer *P, const uint8_t *text, uint64_t n, uint64_t p, uint64_t lim,
                            int rank) {
    const uint8_t *src = P->ranked ? P->ranked : text;
    uint64_t take = min64(min64(P->k, lim), n - p);
    uint64_t bias = rank == ORC_RANK_SHORTER_FIRST ? 1 : 0;
    uint64_t pad = rank == ORC_RANK_SHORTER_FIRST ? 0 : P->alphabet;
    uint64_t key = 0;
    for (uint64_t i = 0; i < take; i++) key = (key << P->bits) | ((uint64_t)src[p + i] + bias);
    for (uint64_t i = take; i < P->k; i++) key = (key << P->bits) | pad;
    return key;
}

/* pack.rs:207-209 (leading_zeros of 0 is 64). */
uint64_t orc_shared_fields(const orc_packer *P, uint64_t a, uint64_t b) {
    uint64_t x = a ^ b;
    uint32_t lz = x ? (uint32_t)__builtin_clzll(x) : 64;
    return lz / P->bits;
}

typedef struct {
    uint64_t key, pos, ord;
} keyed_t;

static int keyed_cmp(const void *a, const void *b) {
    const keyed_t *x = (const keyed_t *)a, *y = (const keyed_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->ord > y->ord) - (x->ord < y->ord);
}

/* pack.rs:241-321.  sort_unstable_by_key (pack.rs:271) leaves equal keys in
 * an unspecified order; this restatement keeps input order, which only
 * matters inside comparator-Equal classes. */
void orc_seed_subarray(const orc_text *T, const orc_limits *L, const orc_packer *P, int rank, uint64_t *sa,
                       uint64_t *lcp, uint64_t len) {
    if (len < 2) {
        if (len == 1) lcp[0] = 0;
        return;
    }
    const uint8_t *bytes = (const uint8_t *)T->data;
    keyed_t *kv = (keyed_t *)malloc(len * sizeof(keyed_t));
    for (uint64_t i = 0; i < len; i++) {
        kv[i].key = orc_key_at_bounded(P, bytes, T->n, sa[i], orc_lim_at(L, sa[i]), rank);
        kv[i].pos = sa[i];
        kv[i].ord = i;
    }
    qsort(kv, len, sizeof(keyed_t), keyed_cmp);
    for (uint64_t i = 0; i < len; i++) sa[i] = kv[i].pos;
    uint64_t *sa_w = (uint64_t *)malloc(len * 8), *lcp_w = (uint64_t *)malloc(len * 8);
    uint64_t i = 0;
    while (i < len) {
        uint64_t j = i + 1;
        while (j < len && kv[j].key == kv[i].key) j++;
        if (j - i > 1) merge_sort_rec(T, L, sa + i, sa_w + i, lcp + i, lcp_w + i, j - i, ORC_MAX_CTX, 0);
        else lcp[i] = 0;
        if (i > 0) {
            uint64_t a = sa[i - 1], b = sa[i];
            uint64_t shared = orc_shared_fields(P, kv[i - 1].key, kv[i].key);
            lcp[i] = min64(min64(shared, orc_lim_at(L, a)), orc_lim_at(L, b));
        }
        i = j;
    }
    lcp[0] = 0;
    free(sa_w), free(lcp_w), free(kv);
}

/* ------------------------------------------------------------------------ */
/* independent checkers                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    const orc_text *T;
    const orc_limits *L;
    uint64_t max_ctx;
} cmp_ctx;

static int brute_cmp(const void *a, const void *b, void *arg) {
    const cmp_ctx *c = (const cmp_ctx *)arg;
    uint64_t p = *(const uint64_t *)a, q = *(const uint64_t *)b;
    /* plain symbol-by-symbol scan, not orc_lcp */
    uint64_t lp = orc_lim_at_bsearch(c->L, p), lq = orc_lim_at_bsearch(c->L, q);
    uint64_t lim = min64(min64(lp, lq), c->max_ctx);
    for (uint64_t i = 0; i < lim; i++) {
        uint32_t x = sym_at(c->T, p + i), y = sym_at(c->T, q + i);
        if (x != y) return x < y ? -1 : 1;
    }
    int o = orc_boundary_order(c->L, p, lp, q, lq);
    if (o) return o;
    return (p > q) - (p < q);
}

void orc_brute_sa(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx) {
    cmp_ctx c = {T, L, max_ctx};
    qsort_r(sa, n, 8, brute_cmp, &c);
}

/* Manber-Myers prefix doubling with two stable counting sorts per round. */
int orc_doubling_sa(const orc_text *T, uint64_t *sa) {
    uint64_t n = T->n;
    if (n == 0) return 0;
    uint64_t R = (n > 65537 ? n : 65537) + 2;
    uint64_t *rk = (uint64_t *)malloc(n * 8), *nrk = (uint64_t *)malloc(n * 8), *tmp = (uint64_t *)malloc(n * 8);
    uint64_t *cnt = (uint64_t *)malloc(R * 8);
    if (!rk || !nrk || !tmp || !cnt) return -1;
    for (uint64_t i = 0; i < n; i++) rk[i] = (uint64_t)sym_at(T, i) + 1;
    for (uint64_t h = 1;; h <<= 1) {
        /* sort by second key rank[i+h] (0 past the end) */
        memset(cnt, 0, R * 8);
        for (uint64_t i = 0; i < n; i++) cnt[i + h < n ? rk[i + h] : 0]++;
        for (uint64_t r = 1; r < R; r++) cnt[r] += cnt[r - 1];
        for (uint64_t i = n; i-- > 0;) tmp[--cnt[i + h < n ? rk[i + h] : 0]] = i;
        /* stable by first key */
        memset(cnt, 0, R * 8);
        for (uint64_t i = 0; i < n; i++) cnt[rk[i]]++;
        for (uint64_t r = 1; r < R; r++) cnt[r] += cnt[r - 1];
        for (uint64_t i = n; i-- > 0;) sa[--cnt[rk[tmp[i]]]] = tmp[i];
        nrk[sa[0]] = 1;
        for (uint64_t i = 1; i < n; i++) {
            uint64_t a = sa[i - 1], b = sa[i];
            int same = rk[a] == rk[b] && (a + h < n ? rk[a + h] : 0) == (b + h < n ? rk[b + h] : 0);
            nrk[b] = nrk[a] + (same ? 0 : 1);
        }
        uint64_t *t = rk;
        rk = nrk, nrk = t;
        if (rk[sa[n - 1]] == n || h >= n) break;
    }
    free(rk), free(nrk), free(tmp), free(cnt);
    return 0;
}

int orc_kasai_lcp(const orc_text *T, const uint64_t *sa, uint64_t *lcp) {
    uint64_t n = T->n;
    if (n == 0) return 0;
    uint64_t *rank = (uint64_t *)malloc(n * 8);
    if (!rank) return -1;
    for (uint64_t i = 0; i < n; i++) rank[sa[i]] = i;
    uint64_t h = 0;
    lcp[0] = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (rank[i] > 0) {
            uint64_t j = sa[rank[i] - 1];
            while (i + h < n && j + h < n && sym_at(T, i + h) == sym_at(T, j + h)) h++;
            lcp[rank[i]] = h;
            if (h > 0) h--;
        } else {
            h = 0;
        }
    }
    free(rank);
    return 0;
}

uint64_t orc_naive_lcp(const orc_text *T, const orc_limits *L, uint64_t a, uint64_t b, uint64_t max_ctx) {
    uint64_t cap = min64(min64(orc_lim_at_bsearch(L, a), orc_lim_at_bsearch(L, b)), max_ctx);
    uint64_t i = 0;
    while (i < cap && sym_at(T, a + i) == sym_at(T, b + i)) i++;
    return i;
}

int64_t orc_verify_sa(const orc_text *T, const orc_limits *L, const uint64_t *sa, uint64_t n,
                      const uint64_t *positions, uint64_t max_ctx, int strict, int canon, int nthreads) {
    uint64_t tn = T->n;
    uint64_t words = tn / 64 + 1;
    uint64_t *want = (uint64_t *)calloc(words, 8), *seen = (uint64_t *)calloc(words, 8);
    if (!want || !seen) return (int64_t)n;
    int64_t bad = -1;
    if (positions) {
        for (uint64_t i = 0; i < n; i++) {
            if (positions[i] >= tn) { bad = (int64_t)n; break; }
            want[positions[i] >> 6] |= 1ull << (positions[i] & 63);
        }
    } else if (n != tn) {
        bad = (int64_t)n;
    }
    for (uint64_t i = 0; i < n && bad < 0; i++) {
        uint64_t v = sa[i];
        if (v >= tn) { bad = (int64_t)n; break; }
        uint64_t w = v >> 6, m = 1ull << (v & 63);
        if (seen[w] & m) { bad = (int64_t)n; break; }
        if (positions && !(want[w] & m)) { bad = (int64_t)n; break; }
        seen[w] |= m;
    }
    free(want), free(seen);
    if (bad >= 0) return bad;
    int64_t first = -1;
#pragma omp parallel for schedule(dynamic, 65536) num_threads(nthreads > 0 ? nthreads : 1)
    for (uint64_t i = 1; i < n; i++) {
        int c = orc_suffix_cmp_with(T, L, sa[i - 1], sa[i], max_ctx);
        if (canon && c == 0) c = cmp_u64(sa[i - 1], sa[i]);
        if (c > 0 || (strict && c == 0)) {
#pragma omp critical
            if (first < 0 || (int64_t)i < first) first = (int64_t)i;
        }
    }
    return first;
}

/* The same checks against a kept-position bitmap (keep_words: bit q & 63 of
 * word q >> 6) instead of a position list: every entry kept and distinct,
 * as many entries as kept positions, then adjacent order.  The permutation
 * part runs in parallel with atomic bitmap updates. */
int64_t orc_verify_sa_keep(const orc_text *T, const orc_limits *L, const uint64_t *sa, uint64_t n,
                           const uint64_t *keep_words, uint64_t max_ctx, int strict, int canon, int nthreads) {
    uint64_t tn = T->n;
    uint64_t words = (tn + 63) / 64;
    uint64_t kept = 0;
    for (uint64_t w = 0; w < words; w++) {
        uint64_t v = keep_words[w];
        if (w == words - 1 && (tn & 63)) v &= (1ull << (tn & 63)) - 1;
        kept += (uint64_t)__builtin_popcountll(v);
    }
    if (kept != n) return (int64_t)n;
    uint64_t *seen = (uint64_t *)calloc(words + 1, 8);
    if (!seen) return (int64_t)n;
    int bad = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : bad)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v = sa[i];
        if (v >= tn) {
            bad = 1;
            continue;
        }
        uint64_t w = v >> 6, m = 1ull << (v & 63);
        if (!(keep_words[w] & m)) bad = 1;
        if (__atomic_fetch_or(&seen[w], m, __ATOMIC_RELAXED) & m) bad = 1;
    }
    free(seen);
    if (bad) return (int64_t)n;
    int64_t first = -1;
#pragma omp parallel for schedule(dynamic, 65536) num_threads(nthreads > 0 ? nthreads : 1)
    for (uint64_t i = 1; i < n; i++) {
        int c = orc_suffix_cmp_with(T, L, sa[i - 1], sa[i], max_ctx);
        if (canon && c == 0) c = cmp_u64(sa[i - 1], sa[i]);
        if (c > 0 || (strict && c == 0)) {
#pragma omp critical
            if (first < 0 || (int64_t)i < first) first = (int64_t)i;
        }
    }
    return first;
}

static int u64_cmp(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

void orc_canonicalize(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx) {
    uint64_t i = 0;
    while (i < n) {
        uint64_t j = i + 1;
        while (j < n && orc_suffix_cmp_with(T, L, sa[j - 1], sa[j], max_ctx) == 0) j++;
        if (j - i > 1) qsort(sa + i, j - i, 8, u64_cmp);
        i = j;
    }
}
