/*
 * caps_sa.h — C ABI of the MI355X suffix-array constructor (gfx950, HIP).
 *
 * Drop-in boundary for the in-memory suffix-array path of caps-sa 0.7.0.
 * Every entry point names the Rust item it replaces.  A Rust facade with the
 * crate's generic signatures monomorphizes onto these symbols (S = u8 | u16,
 * I = u32 | u64, L = PlainText | SegmentedText | the STAR provider); see
 * INTEGRATION.md for the binding.
 *
 * Conventions
 *   - Host-pointer entry points borrow their inputs for the call only and
 *     never retain a pointer after returning.
 *   - *_device entry points take device pointers (HBM-resident inputs and
 *     outputs) and a hipStream_t passed as void*; they return after the work
 *     on that stream has completed.
 *   - Return codes are CAPS_OK or a CAPS_E_* code; caps_last_error() gives a
 *     message (thread-local).  Nothing falls back to a CPU path: an input the
 *     GPU build does not cover returns CAPS_E_UNSUPPORTED.
 *   - Symbols are compared as unsigned integers of sym_bytes bytes.
 *   - Among suffixes the limit provider calls Equal (possible only with
 *     SegmentedText under the default boundary order), output is in
 *     ascending position order.
 */
#ifndef CAPS_SA_H
#define CAPS_SA_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAPS_OK 0
#define CAPS_E_INVALID 1     /* bad arguments: the reference's panics        */
#define CAPS_E_DEVICE 2      /* HIP/device failure: BuildError::Io           */
#define CAPS_E_EMIT 3        /* the emit callback aborted: BuildError::Emit  */
#define CAPS_E_UNSUPPORTED 4 /* symbol/index/provider outside the GPU build  */
#define CAPS_E_NOMEM 5

/* LimitProvider descriptor (limits.rs:34-105).
 *   kind PLAIN:     PlainText::new(n)                    (limits.rs:130-156)
 *   kind SEGMENTED: SegmentedText::from_ends(n, ends)    (limits.rs:261-281)
 *   order SHORTER_FIRST: default boundary_order lim_a.cmp(&lim_b)
 *                        (limits.rs:66-75), BoundaryRank::ShorterFirst
 *   order STAR: lim_b.cmp(&lim_a).then(p_a.cmp(&p_b)), BoundaryRank::LongerFirst
 *               (sample_sort.rs:998-1015, pack.rs:329-351) */
#define CAPS_LIMITS_PLAIN 0u
#define CAPS_LIMITS_SEGMENTED 1u
#define CAPS_ORDER_SHORTER_FIRST 0u
#define CAPS_ORDER_STAR 1u
typedef struct caps_limits {
    uint32_t kind;
    uint32_t order;
    const uint64_t *ends; /* host pointer, strictly increasing, last == n */
    uint64_t n_ends;
} caps_limits_t;

/* Bucket routers for the classification pass. */
#define CAPS_ROUTER_BSEARCH 0u /* binary search over the splitter table   */
#define CAPS_ROUTER_PWL 1u     /* learned piecewise-linear CDF + check     */
#define CAPS_ROUTER_MLP 2u     /* learned 64-32-1 MLP + check              */
#define CAPS_ROUTER_SPLINE 3u  /* spline CDF + bounded-window exact search */

/* Packed key geometry. */
#define CAPS_KEY_DENSE 0u     /* ceil(log2(sigma))-bit fields, weakly order preserving */
#define CAPS_KEY_REFERENCE 1u /* pack.rs geometry (u8 only; dense when ineligible)    */
#define CAPS_KEY_ESCAPE32 2u  /* caps_seed_keys_device only: the CAPS_FLAG_SEED_ESCAPE32
                                 keys (u8 text of 5..16 symbols)                      */

/* Opts (sample_sort.rs:71-85) + ExtMemOpts (ext_mem.rs:58-239) + GPU knobs. */
typedef struct caps_opts {
    uint64_t max_context;      /* Opts::max_context; UINT64_MAX = unbounded            */
    uint64_t subproblem_count; /* ExtMemOpts::subproblem_count (the output never
                                  depends on it; kept for the API)                     */
    uint32_t router;           /* CAPS_ROUTER_*                                         */
    uint32_t key_geometry;     /* CAPS_KEY_*                                            */
    uint32_t gpu_buckets;      /* classification buckets the classify pass (binary
                                  search or router) decides among, 1..4096; each is
                                  split into LDS-sized buckets by the second scatter
                                  pass.  0 = auto (1,024, more for huge inputs)         */
    uint32_t tile_keys;        /* mean records per bucket the splitters aim at
                                  (0 = 1536; 64..4096)                                 */
    uint32_t want_lcp;         /* fill the LCP array where an entry point takes one     */
    uint32_t flags;            /* CAPS_FLAG_* bits, 0 by default                        */
    uint64_t batch_records;    /* records per bucket batch: the build sorts consecutive
                                  coarse intervals of at most this many suffixes at a
                                  time (0 = auto: one batch while its buffers fit in
                                  HBM).  The output never changes.                     */
    uint32_t packed_prefix_seed; /* ExtMemOpts::packed_prefix_seed (pack.rs:37-69):
                                  CAPS_SEED_POLICY_*; with key_geometry
                                  CAPS_KEY_REFERENCE the build seeds with the
                                  reference's packed keys exactly when the
                                  reference would (caps_seed_params_device), else
                                  with dense keys.  The output never changes.     */
    uint32_t reserved;
    uint64_t remap_budget;     /* PackedPrefixSeedPolicy::Remap { max_extra_bytes }  */
} caps_opts_t;
#define CAPS_SEED_POLICY_DISABLED 0u
#define CAPS_SEED_POLICY_DENSE_ALPHABET_ONLY 1u
#define CAPS_SEED_POLICY_REMAP 2u
/* Skip router training (the learned router runs with its initial parameters;
 * used to show what training gains).  The spline router keeps its knots but
 * skips its error bound (a window of one position), so keys outside it take
 * the binary-search fallback.  The output never changes. */
#define CAPS_FLAG_ROUTER_UNTRAINED 1u
/* Refine groups of 2..64 members with persistent kernels fed from device work
 * queues (slots take the next group when theirs is done) instead of one grid
 * slot per group.  The output never changes. */
#define CAPS_FLAG_PERSISTENT_REFINE 2u
/* Sort tiles with the 4-bit LDS radix sort instead of the merge-path block
 * merge sort.  The output never changes. */
#define CAPS_FLAG_TILE_RADIX 4u
/* Merge-path partition width of the giant-group LCP merges: 16 or 64 outputs
 * per thread instead of the default (32, or 64 when the largest giant group
 * has 131,072+ members).  The output never changes. */
#define CAPS_FLAG_MERGE_W16 8u
#define CAPS_FLAG_MERGE_W64 16u
/* Dense seed keys with power-of-two fields (the packed text's width, e.g. 4
 * bits and 16 symbols for a 6-symbol alphabet) instead of ceil(log2(sigma))
 * bits (3 bits, 21 symbols): fewer symbols per key, more ties to refine.
 * The output never changes. */
#define CAPS_FLAG_POW2_SEED 32u
/* Bucket batches: run scatter pass 1 once per batch (records of that batch
 * only) even when the records of all batches fit in HBM at once.  The output
 * never changes. */
#define CAPS_FLAG_SCATTER_PER_BATCH 64u
/* Measurement build: count the symbols the tie refinement reads from the
 * text (stats cmp_symbols).  Costs a few atomics per wavefront; the output
 * never changes. */
#define CAPS_FLAG_COUNT_WORK 128u
/* Scatter pass 1 writes each record straight to its interval run instead of
 * staging every 4,096-position step in LDS first (the staged pass is used for
 * every suffix of one text, u32 index, up to 1,024 classification buckets).
 * The output never changes. */
#define CAPS_FLAG_SCATTER_DIRECT 256u
/* Graph builds: the tile sort, the refinement and the giant-group merges of
 * every bucket batch are each captured into a hipGraph (on a library-owned
 * stream, default priority) and launched on the build's stream; the
 * instantiated graphs are kept and updated in place by the next build whose
 * phase launches the same kernels (hipGraphExecUpdate), else instantiated
 * again.  Their kernels take their work from the device group lists and
 * work queues as in an eager build.  The output never changes. */
#define CAPS_FLAG_GRAPH 512u
/* Genome seed keys: for u8 texts of 5..16 symbols (nucleotides 0..3 plus
 * spacers and N), 32 two-bit fields per seed key instead of ceil(log2 sigma)
 * bits (3 bits, 21 fields for 6 symbols).  A code >= 4 and every field after
 * it become the largest field value; the keys stay weakly order-preserving,
 * and tie groups are refined from depth 0.  Not for batched builds.  The
 * output never changes. */
#define CAPS_FLAG_SEED_ESCAPE32 1024u
/* Build without the run memo: by default, when a build has groups of 65 or
 * more tied suffixes, the maximal runs of one code (48+ symbols) of the
 * packed text are recorded once, and the LCP scans of those groups skip the
 * common part of two runs of the same code at once.  The output never
 * changes. */
#define CAPS_FLAG_NO_RUN_MEMO 2048u
/* Refine the groups of 65..2048 tied suffixes in four launches by size (up
 * to 128, 512, 1,024 and 2,048 members: smaller LDS arrays, so more
 * workgroups per compute unit) whatever their number; by default this
 * happens from 16 such groups per compute unit up, and fewer groups run in
 * one launch of one workgroup per group.  The output never changes. */
#define CAPS_FLAG_MID_BY_SIZE 4096u

/* Fills opts with the reference defaults (max_context unbounded, p auto). */
void caps_opts_default(caps_opts_t *opts);

/* Thread-local message for the last nonzero return code. */
const char *caps_last_error(void);

/* Number of HIP devices; 0 when none is usable. */
int caps_device_count(void);

/* ---- in-memory builders: sample_sort.rs:97-206 --------------------------- */

/* build_in_memory / build_in_memory_with_opts / build_in_memory_with
 * (sample_sort.rs:97-127).  out_sa holds n entries of idx_bytes (4 | 8);
 * out_lcp (nullable, same width) receives lcp[i] = LCP(sa[i-1], sa[i]) capped
 * by both limits and max_context, lcp[0] = 0 — merge_sort's postcondition
 * (sample_sort.rs:208-216). */
int caps_build_in_memory(const void *text, uint64_t n, uint32_t sym_bytes, uint32_t idx_bytes,
                         const caps_limits_t *limits, const caps_opts_t *opts, void *out_sa, void *out_lcp);

/* build_in_memory_for_positions(_with_opts|_with) (sample_sort.rs:143-206):
 * sorts `positions` (n_pos entries of idx_bytes) by their suffixes.  out_sa
 * may alias positions (the reference reuses the moved-in Vec).  A position of
 * n or more, or positions == NULL with n_pos != n, returns CAPS_E_INVALID
 * before any device work (the reference panics on the out-of-range index). */
int caps_build_in_memory_for_positions(const void *text, uint64_t n, uint32_t sym_bytes, const void *positions,
                                       uint64_t n_pos, uint32_t idx_bytes, const caps_limits_t *limits,
                                       const caps_opts_t *opts, void *out_sa, void *out_lcp);

/* ---- streaming builders: ext_mem.rs:289-871 ------------------------------ */

/* Emit callback: called sequentially on the caller's thread, in lexicographic
 * order, with batches of positions widened to u64.  A nonzero return aborts
 * the build; caps_* then returns CAPS_E_EMIT and caps_last_emit_code() gives
 * the callback's value (BuildError::Emit, ext_mem.rs:257-277). */
typedef int (*caps_emit_fn)(void *ctx, const uint64_t *pos, uint64_t count);
int caps_last_emit_code(void);

/* (try_)build_in_memory_sample_sort(_with|_for_positions|_for_positions_with)
 * (ext_mem.rs:711-871).  positions == NULL means every position 0..n-1. */
int caps_build_in_memory_sample_sort(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *positions,
                                     uint64_t n_pos, const caps_limits_t *limits, const caps_opts_t *opts,
                                     caps_emit_fn emit, void *ctx);

/* (try_)build_ext_mem_for_positions(_with) (ext_mem.rs:370-449).  The GPU
 * build keeps text and records in HBM instead of spilling to disk; work_dir
 * and physical_file_count have no counterpart.  Same contract as above. */
int caps_build_ext_mem_for_positions(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *positions,
                                     uint64_t n_pos, const caps_limits_t *limits, const caps_opts_t *opts,
                                     caps_emit_fn emit, void *ctx);

/* (try_)build_ext_mem_for_filter(_with) (ext_mem.rs:472-547).  The crate walks
 * the caller's keep predicate once into a bitmap of kept positions with
 * per-block popcounts (FilteredSource, ext_mem.rs:927-1092) and never calls it
 * again; here the caller passes that bitmap: (n + 63) / 64 words, position q
 * kept iff bit q & 63 of word q >> 6 is set (bits past n are ignored).  No
 * position list is materialised: the GPU walks the bitmap beside the text
 * windows.  Emits every kept position, in order, as above. */
int caps_build_ext_mem_for_filter(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *keep_words,
                                  const caps_limits_t *limits, const caps_opts_t *opts, caps_emit_fn emit,
                                  void *ctx);

/* ---- device-resident builder (inputs and outputs already in HBM) ---------- */

/* Same as caps_build_in_memory_for_positions with device pointers.
 * d_positions == NULL means 0..n-1.  d_out_lcp nullable.  The positions are
 * checked on the device first: one of n or more returns CAPS_E_INVALID. */
int caps_build_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const void *d_positions, uint64_t n_pos,
                      uint32_t idx_bytes, const caps_limits_t *limits, const caps_opts_t *opts, void *d_out_sa,
                      void *d_out_lcp, void *stream);

/* caps_build_ext_mem_for_filter with device pointers: d_keep_words is the
 * kept-position bitmap in HBM (read only).  *n_out receives the kept count;
 * d_out_sa (and d_out_lcp when not NULL) must hold out_capacity >= that many
 * entries, else CAPS_E_INVALID. */
int caps_build_device_filtered(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *d_keep_words,
                               uint32_t idx_bytes, const caps_limits_t *limits, const caps_opts_t *opts,
                               void *d_out_sa, uint64_t out_capacity, uint64_t *n_out, void *d_out_lcp,
                               void *stream);

/* ---- no-alloc builder (no counterpart in the reference) ------------------- */

/* For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as caps_build_device, but every device buffer the
 * build needs beyond its inputs and outputs (packed text, keys, records,
 * splitters, group lists, merge scratch, counters) is carved from one static
 * device array of CAPS_NOALLOC_SCRATCH_MIB MiB, which each GPU the library
 * runs on holds until the process ends.  No hipMalloc, hipFree or allocating
 * library routine is called.  n (and n_pos) above CAPS_NOALLOC_MAX_SYMBOLS,
 * or more than CAPS_NOALLOC_MAX_ENDS segment ends, return CAPS_E_INVALID.
 * CAPS_FLAG_GRAPH is ignored here (instantiating a graph may allocate).
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents (the build clears what must start at zero): calls that overlap on
 * one GPU race.  The library serializes its calls within a process and each
 * call waits for its stream before returning. */
#define CAPS_NOALLOC_SCRATCH_MIB 128u
#define CAPS_NOALLOC_MAX_SYMBOLS (1ull << 20)
#define CAPS_NOALLOC_MAX_ENDS (1ull << 16)
int caps_build_device_noalloc(const void *d_text, uint64_t n, uint32_t sym_bytes, const void *d_positions,
                              uint64_t n_pos, uint32_t idx_bytes, const caps_limits_t *limits,
                              const caps_opts_t *opts, void *d_out_sa, void *d_out_lcp, void *stream);

/* ---- batched builds: many independent texts in one pipeline -------------- */

/* One text of a batch (host pointers).  out_sa receives n entries of
 * idx_bytes: the text's own suffix array (positions local to the text). */
typedef struct caps_text_item {
    const void *text;
    uint64_t n;
    void *out_sa;
} caps_text_item_t;

/* build_in_memory (sample_sort.rs:97-103) for every item, as one GPU build:
 * the texts are laid back to back, each a segment whose end stops every
 * comparison, and the packed keys carry the text index in their top bits so
 * one sort orders all suffixes text-major.  All items share sym_bytes.
 * Empty texts are skipped.  Batches larger than the key's text-index budget
 * are split into several builds. */
int caps_build_batch(const caps_text_item_t *items, uint64_t count, uint32_t sym_bytes, uint32_t idx_bytes,
                     const caps_opts_t *opts);

/* Device variant: d_text holds the texts back to back, ends (host pointer)
 * their n_texts cumulative ends (strictly increasing, last == n).  d_out
 * receives n positions: text after text, each text's suffix array as global
 * positions (subtract the text's start). */
int caps_build_batch_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *ends,
                            uint64_t n_texts, uint32_t idx_bytes, const caps_opts_t *opts, void *d_out, void *stream);

/* Per-phase device times (ms) and counters of the last build on this thread. */
typedef struct caps_stats {
    double ms_total, ms_alphabet, ms_sample, ms_classify, ms_scatter, ms_tile_sort, ms_merge, ms_detect,
        ms_refine, ms_giant;
    /* single-kernel durations (event pairs around one launch on the build's stream) */
    double k_classify, k_scatter, k_tile_sort, k_detect, k_pairs, k_groups_small, k_groups_mid, k_boundary_lcp;
    uint64_t buckets, tiles, oversized, pairs, small_groups, mid_groups, giant_groups;
    uint64_t router_err_sum, router_err_max, router_fallbacks, router_count;
    uint32_t key_bits, key_k;
    uint64_t bucket_max; /* records in the largest bucket (imbalance = bucket_max * buckets / n) */
    double k_scatter1, k_scatter2; /* the two scatter passes (their sum is k_scatter) */
    uint64_t batches;              /* bucket batches the build ran                      */
    uint64_t coarse_buckets;       /* classification buckets (coarse intervals)         */
    uint64_t coarse_max;           /* records in the largest classification bucket      */
    uint64_t scratch_bytes;        /* no-alloc builds: bytes of the scratch array used  */
    double k_tile[3];              /* tile sort launches of the 1,024 / 2,048 / 4,096
                                      record classes (their sum is k_tile_sort)         */
    uint64_t tile_records[3];      /* records sorted by each of those launches          */
    uint64_t cmp_symbols[3];       /* CAPS_FLAG_COUNT_WORK builds: symbols read by the
                                      refinement (pairs; groups of 3..64; larger groups) */
    double k_graph[3];             /* CAPS_FLAG_GRAPH builds: launch-to-end time of the
                                      tile-sort, refinement and giant-group graphs      */
    uint64_t runs;                 /* runs recorded in the run memo (0: not built)      */
    int64_t router_bound_lo, router_bound_hi; /* spline router: the exact range of its
                                      prediction error (predicted - true position)      */
} caps_stats_t;
int caps_last_stats(caps_stats_t *out);

/* ---- autotuner (no counterpart in the reference) ------------------------- */

/* Times builds of the caller's own input over a small grid of knobs (mean
 * bucket size = keys per LDS tile, tile sort, refinement launch style,
 * merge-path width) and writes the fastest into *opts (its other fields are
 * kept).  d_scratch_sa: n entries of idx_bytes, overwritten.  reps: timed
 * builds per configuration after one warm build.  The suffix array does not
 * depend on any of these knobs. */
#define CAPS_TUNE_MAX 16
typedef struct caps_tune_report {
    uint32_t n_tried;
    double best_ms;
    double ms[CAPS_TUNE_MAX];
    uint32_t tile_keys[CAPS_TUNE_MAX];
    uint32_t flags[CAPS_TUNE_MAX];
} caps_tune_report_t;
int caps_autotune_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                         uint32_t idx_bytes, void *d_scratch_sa, uint32_t reps, caps_opts_t *opts, void *stream,
                         caps_tune_report_t *report);
/* The same for a batch (caps_build_batch_device's arguments). */
int caps_autotune_batch_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *ends,
                               uint64_t n_texts, uint32_t idx_bytes, void *d_scratch_sa, uint32_t reps,
                               caps_opts_t *opts, void *stream, caps_tune_report_t *report);

/* ---- kernels exposed for parity tests (device pointers) ------------------ */

/* LcpDispatch::lcp (lcp.rs:125-153) for `count` pairs (p, q): the 64-lane
 * wave compare.  d_pairs: 2*count u64; d_out: count u64. */
int caps_lcp_pairs_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *d_pairs,
                          uint64_t count, uint64_t max_ctx, uint64_t *d_out, void *stream);

/* LcpDispatch::suffix_cmp_with (lcp.rs:184-201) for `count` pairs, without
 * the position tie: d_out[i] in {-1, 0, 1}. */
int caps_suffix_cmp_pairs_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                                 const uint64_t *d_pairs, uint64_t count, uint64_t max_ctx, int32_t *d_out,
                                 void *stream);

/* packed_prefix_seed_params (ext_mem.rs:1495-1513) then seed_params
 * (pack.rs:217-234): whether the reference seeds a build of this text with
 * packed keys, checked in the reference's order, stopping at the first check
 * that fails: the policy is Disabled; max_context is bounded; the symbol type
 * is not u8; (boundary_rank: every descriptor here declares one); the
 * alphabet (one scan of the device text) is not the identity map and the
 * policy is DenseAlphabetOnly, or Remap over its byte budget; no sentinel
 * code fits beside the alphabet (pack.rs:174-176).  out receives the reason
 * (CAPS_SEED_*) and, when eligible, the geometry. */
#define CAPS_SEED_OK 0u
#define CAPS_SEED_DISABLED 1u
#define CAPS_SEED_MAX_CONTEXT 2u
#define CAPS_SEED_NOT_U8 3u
#define CAPS_SEED_GAPPED_ALPHABET 4u
#define CAPS_SEED_REMAP_BUDGET 5u
#define CAPS_SEED_NO_SENTINEL 6u
typedef struct caps_seed_params {
    uint32_t reason;   /* CAPS_SEED_*                                   */
    uint32_t bits, k;  /* field width and symbols per key when eligible */
    uint32_t alphabet; /* distinct symbols (0 when the scan was skipped) */
    uint32_t remap;    /* 1 when the ranked copy of the text is needed  */
} caps_seed_params_t;
int caps_seed_params_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                            const caps_opts_t *opts, caps_seed_params_t *out, void *stream);

/* Packer::key_at_bounded (pack.rs:187-203) for `count` positions, in the
 * geometry `key_geometry` (u8 text).  rank: 0 shorter first, 1 longer first.
 * Geometry out: bits[0], k[1]. Returns CAPS_E_UNSUPPORTED when the reference
 * geometry is ineligible (pack.rs:217-234). */
int caps_seed_keys_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                          uint32_t key_geometry, uint32_t rank, const uint64_t *d_pos, uint64_t count,
                          uint64_t *d_keys, uint32_t *geometry_out, void *stream);

/* merge (sample_sort.rs:505-522): LCP-enhanced merge of two sorted runs with
 * LCP arrays (lcp_*[0] = 0) into z / lcp_z, as a merge-path GPU merge.
 * Host pointers; all arrays u64. */
int caps_merge_lcp(const void *text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits, const uint64_t *x,
                   const uint64_t *lcp_x, uint64_t nx, const uint64_t *y, const uint64_t *lcp_y, uint64_t ny,
                   uint64_t max_ctx, uint64_t *z, uint64_t *lcp_z);

/* CascadeWorkspace::cascade_merge (ext_mem.rs:2144-2243): merge the sorted
 * runs delimited by boundaries = [0, e1, ..., n] level by level.  Host
 * pointers; arrays u64. */
int caps_cascade_merge(const void *text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                       const uint64_t *sa, const uint64_t *lcp, uint64_t n_sa, const uint64_t *boundaries,
                       uint64_t n_boundaries, uint64_t max_ctx, uint64_t *out_sa, uint64_t *out_lcp);

/* ---- learned router kernels (device pointers; for gradient checks) ------- */

/* Mean-squared-error gradient of the 64-32-1 router MLP over `count`
 * (key, target) pairs.  params: W1[64][32] (input i = key bit 63-i), b1[32],
 * w2[32], b2 (2113 floats).  d_grad (2113 floats) and d_loss (1 float) are
 * overwritten. */
int caps_router_mlp_grad_device(const float *d_params, const uint64_t *d_keys, const float *d_targets, uint32_t count,
                                float *d_grad, float *d_loss, void *stream);

/* Network output for `count` keys, evaluated the way classification does
 * (nibble tables). */
int caps_router_mlp_forward_device(const float *d_params, const uint64_t *d_keys, uint32_t count, float *d_out,
                                   void *stream);

/* The piecewise-linear router's least-squares fit, as the build runs it on
 * its sorted sample: ns segments over the sorted keys d_keys[0..s) with
 * their bucket labels d_labels (non-decreasing).  Segment j covers sample
 * entries [j*s/ns, (j+1)*s/ns); its knot is the first key, x = float(key -
 * knot) * inv with inv = float(1 / (next knot - knot)), and (a, b) is the
 * least-squares line label = a*x + b (a clamped at 0 to keep the CDF
 * monotone).  Outputs (host arrays of ns entries): knots (ns + 1, the last
 * = UINT64_MAX), a, b, inv, lo / hi (the labels at the segment's two
 * knots: exact bounds for any key between them). */
int caps_router_pwl_fit_device(const uint64_t *d_keys, const uint32_t *d_labels, uint64_t s, uint32_t ns,
                               uint32_t n_buckets, uint64_t *knots, float *a, float *b, float *inv, uint32_t *lo,
                               uint32_t *hi, void *stream);

/* The spline router's fit, as the build runs it on its sorted splitter table
 * d_splitters[0..n) (n < 65,536): knots (4,097 u16, host) = the number of
 * splitters below t * 2^52 for t < 4,096, and n; bound (2 int32, host) =
 * the minimum and maximum of predicted - true position over every key, where
 * predicted = knots[t] + ((f * (knots[t+1] - knots[t])) >> 16) for
 * t = key >> 52, f = (key >> 36) & 0xFFFF, and true = the number of
 * splitters <= key. */
int caps_router_spline_fit_device(const uint64_t *d_splitters, uint32_t n, uint16_t *knots, int32_t *bound,
                                  void *stream);

#ifdef __cplusplus
}
#endif
#endif
