/*
 * CPU oracle for the suffix-array path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of caps-sa 0.7.0's in-memory suffix-array algorithm
 * (src/sample_sort.rs, src/lcp.rs, src/limits.rs, src/pack.rs, src/ext_mem.rs),
 * plus independent checkers (brute force, prefix doubling, Kasai LCP, an
 * order/permutation verifier).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the
 * timed CPU baseline.  The product library (caps-sa_amd/) never links it.
 *
 * Pinning (see tests/test_oracle.py):
 *   - lcp / suffix_cmp_with: the reference's known answers (lcp.rs:438-471,
 *     482-491, 546-564) and a brute-force byte scan;
 *   - limits: the reference's known answers (limits.rs:343-422) and a plain
 *     binary search over ends (the directory must agree with it);
 *   - merge_sort / merge / build_in_memory: brute-force sort by slices,
 *     prefix doubling (an unrelated algorithm), the reference's fixtures
 *     (banana, mississippi, abracadabra, ...), and the LCP postcondition;
 *   - packer: the reference's known answers (pack.rs:484-563);
 *   - sample sort (phases 1-4): equality with build_in_memory, as the
 *     reference's own differential tests (ext_mem.rs:2784-2806).
 *
 * Positions and LCPs are carried as uint64_t; the reference's u32 outputs are
 * the same values narrowed.
 */
#ifndef CAPS_ORACLE_H
#define CAPS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_CTX UINT64_MAX

/* A text of u8 or u16 symbols. */
typedef struct {
    const void *data;
    uint64_t n;
    uint32_t sym_bytes; /* 1 or 2 */
} orc_text;

/* boundary_order conventions (limits.rs:66-75; sample_sort.rs:1006-1014). */
enum {
    ORC_ORDER_SHORTER_FIRST = 0, /* lim_a.cmp(lim_b): the default provider order */
    ORC_ORDER_STAR = 1,          /* lim_b.cmp(lim_a).then(p_a.cmp(p_b))          */
};

/* PlainText (limits.rs:130-156) when n_ends == 0, SegmentedText otherwise
 * (limits.rs:182-336), with the coarse directory of limits.rs:194-236. */
typedef struct {
    uint64_t n;
    const uint64_t *ends;
    uint64_t n_ends;
    int order;
    int has_dir;
    uint32_t block_shift;
    uint32_t *first_after_block_start;
    uint64_t n_dir;
} orc_limits;

int orc_limits_init(orc_limits *L, uint64_t n, const uint64_t *ends, uint64_t n_ends, int order);
void orc_limits_free(orc_limits *L);
uint64_t orc_lim_at(const orc_limits *L, uint64_t p);
uint64_t orc_lim_at_bsearch(const orc_limits *L, uint64_t p); /* no directory */
int orc_boundary_order(const orc_limits *L, uint64_t pa, uint64_t lim_a, uint64_t pb, uint64_t lim_b);

/* lcp.rs:125-153 (LcpDispatch::lcp) and lcp.rs:184-201 (suffix_cmp_with). */
uint64_t orc_lcp(const orc_text *T, uint64_t p, uint64_t q, uint64_t max_ctx);
int orc_suffix_cmp_with(const orc_text *T, const orc_limits *L, uint64_t p, uint64_t q, uint64_t max_ctx);

/* sample_sort.rs:397-522 (merge) and 614-636 (drain). */
void orc_merge(const orc_text *T, const orc_limits *L, const uint64_t *x, uint64_t nx, const uint64_t *y,
               uint64_t ny, const uint64_t *lcp_x, const uint64_t *lcp_y, uint64_t *z, uint64_t *lcp_z,
               uint64_t max_ctx);
/* sample_sort.rs:221-266 (merge_sort). nthreads<=1 runs sequentially. */
void orc_merge_sort(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t *sa_w, uint64_t *lcp,
                    uint64_t *lcp_w, uint64_t n, uint64_t max_ctx, int nthreads);
/* sample_sort.rs:168-206: sorts `sa` (the positions) in place; lcp optional. */
void orc_build_in_memory(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx,
                         uint64_t *lcp_out, int nthreads);

/* ext_mem.rs:2194-2389 (CascadeWorkspace::cascade_merge_impl / merge_one_level).
 * boundaries = [0, e1, e2, ..., n] as InMemBucket::boundaries (ext_bucket.rs:169-175). */
void orc_cascade_merge(const orc_text *T, const orc_limits *L, const uint64_t *sa, const uint64_t *lcp,
                       uint64_t n, const uint64_t *boundaries, uint64_t n_boundaries, uint64_t max_ctx,
                       uint64_t *out_sa, uint64_t *out_lcp);

/* ext_mem.rs:1159-1178, 1282-1287, 1291-1305. */
uint64_t orc_effective_subproblem_count(uint64_t n, uint64_t requested, uint64_t nthreads);
uint64_t orc_sample_target_total(uint64_t n, uint64_t p);
uint64_t orc_evenly_spaced(const uint64_t *xs, uint64_t n, uint64_t count, uint64_t *out);
/* ext_mem.rs:1705-1734. */
uint64_t orc_upper_bound_by_pivot(const orc_text *T, const orc_limits *L, const uint64_t *recs, uint64_t n,
                                  uint64_t pivot, uint64_t max_ctx);

typedef int (*orc_emit_fn)(void *ctx, const uint64_t *pos, uint64_t count);
/* ext_mem.rs:651-702 (build_in_memory_ss_inner), phases 1-4.  positions==NULL
 * means the identity source 0..n.  Returns 0, or the first nonzero emit code. */
int orc_build_in_memory_sample_sort(const orc_text *T, const orc_limits *L, const uint64_t *positions,
                                    uint64_t n_pos, uint64_t subproblem_count, uint64_t max_ctx, int nthreads,
                                    orc_emit_fn emit, void *ctx);

/* pack.rs:84-209: the alphabet map and the packed-prefix key. */
enum { ORC_SEED_DISABLED = 0, ORC_SEED_DENSE_ONLY = 1, ORC_SEED_REMAP = 2 };
enum { ORC_RANK_SHORTER_FIRST = 0, ORC_RANK_LONGER_FIRST = 1 };
typedef struct {
    uint8_t code[256];
    uint8_t *ranked; /* NULL when the identity map is dense */
    uint32_t bits;
    uint32_t k;
    uint32_t alphabet;
} orc_packer;
/* pack.rs:217-234 (seed_params; S == u8 is the caller's job).  Returns 1 and
 * fills *P when a packed key is available, 0 otherwise. */
int orc_seed_params(const uint8_t *text, uint64_t n, int policy, uint64_t remap_budget, orc_packer *P);
void orc_packer_free(orc_packer *P);
uint64_t orc_key_at_bounded(const orc_packer *P, const uint8_t *text, uint64_t n, uint64_t p, uint64_t lim,
                            int rank);
uint64_t orc_shared_fields(const orc_packer *P, uint64_t a, uint64_t b);
/* pack.rs:241-321. */
void orc_seed_subarray(const orc_text *T, const orc_limits *L, const orc_packer *P, int rank, uint64_t *sa,
                       uint64_t *lcp, uint64_t len);

/* ---- independent checkers (not restatements) ---- */
/* Brute force: sort by suffix_cmp_with, ascending position among Equal. */
void orc_brute_sa(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx);
/* Prefix doubling over all positions of a plain text (Manber-Myers ranks). */
int orc_doubling_sa(const orc_text *T, uint64_t *sa);
/* Kasai et al. LCP of a full plain-text SA (lcp[0] = 0). */
int orc_kasai_lcp(const orc_text *T, const uint64_t *sa, uint64_t *lcp);
/* Adjacent-pair LCP bounded by both limits and max_ctx (naive scan). */
uint64_t orc_naive_lcp(const orc_text *T, const orc_limits *L, uint64_t a, uint64_t b, uint64_t max_ctx);
/* Checks sa is a permutation of positions (NULL = 0..n) and adjacent pairs are
 * non-decreasing (strict=0) or increasing (strict=1) under suffix_cmp_with,
 * ascending position among Equal when canon=1.  Returns -1 if ok, else the
 * first failing index (n for a permutation failure). */
int64_t orc_verify_sa_keep(const orc_text *T, const orc_limits *L, const uint64_t *sa, uint64_t n,
                           const uint64_t *keep_words, uint64_t max_ctx, int strict, int canon, int nthreads);
int64_t orc_verify_sa(const orc_text *T, const orc_limits *L, const uint64_t *sa, uint64_t n,
                      const uint64_t *positions, uint64_t max_ctx, int strict, int canon, int nthreads);
/* Reorders each run of comparator-Equal neighbours by ascending position. */
void orc_canonicalize(const orc_text *T, const orc_limits *L, uint64_t *sa, uint64_t n, uint64_t max_ctx);

#ifdef __cplusplus
}
#endif
#endif
