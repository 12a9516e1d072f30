// Host side of the MI355X suffix-array constructor: device buffers, the
// phase pipeline, and the C ABI declared in include/caps_sa.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/caps_sa.h"
#include "engine.hpp"
#include "kernels.hpp"

namespace caps {

thread_local std::string g_err;
thread_local int g_emit_code = 0;
thread_local caps_stats_t g_stats;
// one build at a time per process (the device context is serialized);
// recursive: the autotuner's entry points run builds through the public ones
std::recursive_mutex g_mu;

struct CapsError : std::runtime_error {
    int code;
    CapsError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                         \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess)                                                                             \
            throw CapsError(e_ == hipErrorOutOfMemory ? CAPS_E_NOMEM : CAPS_E_DEVICE,                     \
                            std::string(#x) + ": " + hipGetErrorString(e_));                              \
    } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

// ---------------------------------------------------------------------------
// No-alloc builds: every device buffer of the build is carved from one
// static device array per GPU instead of hipMalloc.  Its device address is
// looked up once per device with hipGetSymbolAddress (the array's name in
// host code is only a host-side handle, never a device address).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = CAPS_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct Arena {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0;
    std::vector<std::pair<const DevBuf *, std::pair<void *, size_t>>> slots;  // per buffer: grow-only, as DevBuf
    void *take(const DevBuf *owner, size_t bytes) {
        for (auto &sl : slots)
            if (sl.first == owner && sl.second.second >= bytes) return sl.second.first;
        const size_t at = (used + 255) & ~(size_t)255;
        if (at + bytes > cap)
            throw CapsError(CAPS_E_INVALID, "the build needs more than the " + std::to_string(cap >> 20) +
                                                " MiB no-alloc scratch");
        used = at + bytes;
        void *ptr = base + at;
        bool found = false;
        for (auto &sl : slots)
            if (sl.first == owner) sl.second = {ptr, bytes}, found = true;
        if (!found) slots.push_back({owner, {ptr, bytes}});
        return ptr;
    }
};
static Arena *g_arena = nullptr;  // set for the duration of a no-alloc build (under g_mu)

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

void *DevBuf::get(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (g_arena) return g_arena->take(this, bytes);
    if (bytes > cap) {
        release();
        HIPCHK(hipMalloc(&p, bytes));
        cap = bytes;
    }
    return p;
}

Workspace &workspace() {
    static Workspace *w = new Workspace();
    return *w;
}

static inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Blocks of `kernel` that fit on the device at once (one resident wave of
// blocks), for grid-stride kernels that stage tables once per block.
template <typename K>
static unsigned resident_blocks(K kernel, unsigned block, size_t dyn_lds = 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, dyn_lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (unsigned)per_cu * std::max<uint32_t>(workspace().n_cu, 1);
}

static unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 65536u * 4) {
    uint64_t g = ceil_div(n ? n : 1, block);
    return (unsigned)std::min<uint64_t>(g, cap);
}

// ---------------------------------------------------------------------------
// Limits: upload ends + the reference's coarse directory (limits.rs:194-236).
// ---------------------------------------------------------------------------
DevLimits upload_limits(Workspace &W, const caps_limits_t *lim, uint64_t n, hipStream_t st) {
    DevLimits L{};
    L.n = n;
    L.order = lim ? lim->order : CAPS_ORDER_SHORTER_FIRST;
    if (L.order > 1) throw CapsError(CAPS_E_INVALID, "unknown boundary order");
    if (!lim || lim->kind == CAPS_LIMITS_PLAIN) return L;
    if (lim->kind != CAPS_LIMITS_SEGMENTED) throw CapsError(CAPS_E_INVALID, "unknown limits kind");
    const uint64_t *ends = lim->ends;
    uint64_t ne = lim->n_ends;
    // SegmentedText::from_ends checks (limits.rs:261-281)
    for (uint64_t i = 1; i < ne; ++i)
        if (!(ends[i - 1] < ends[i])) throw CapsError(CAPS_E_INVALID, "segment ends must be strictly increasing");
    if (ne == 0 ? n != 0 : ends[ne - 1] != n) throw CapsError(CAPS_E_INVALID, "last segment end must equal n");
    if (ne == 0) return L;
    W.h_ends.assign(ends, ends + ne);
    uint64_t *d_ends = (uint64_t *)W.ends.get(ne * 8);
    HIPCHK(hipMemcpyAsync(d_ends, W.h_ends.data(), ne * 8, hipMemcpyHostToDevice, st));
    L.ends = d_ends;
    L.n_ends = ne;
    if (ne >= 256 && ne <= 0xFFFFFFFFull && n > 0) {
        uint64_t target = std::min<uint64_t>(std::max<uint64_t>(ne * 2, 1), 2000000);
        uint64_t minb = ceil_div(n, target);
        uint64_t bs = 1;
        while (bs < minb) bs <<= 1;
        uint32_t shift = (uint32_t)__builtin_ctzll(bs);
        uint64_t nb = ceil_div(n, bs);
        // built on the device from the uploaded ends (one binary search per
        // block) instead of a host walk and a second upload
        uint32_t *d_dir = (uint32_t *)W.dir.get((nb + 1) * 4);
        k_seg_dir<<<(unsigned)ceil_div(nb + 1, 256), 256, 0, st>>>(d_ends, ne, n, shift, nb, d_dir);
        LAUNCHCHK();
        L.dir = d_dir;
        L.block_shift = shift;
    }
    return L;
}

// ---------------------------------------------------------------------------
// Whether the reference seeds a build with packed keys: the checks of
// packed_prefix_seed_params (ext_mem.rs:1495-1513) and seed_params /
// Packer::new / has_sentinel (pack.rs:103-176, 217-234), in the reference's
// order.  The cheap checks come first; the alphabet scan only after them.
// ---------------------------------------------------------------------------
struct SeedDecision {
    uint32_t reason = CAPS_SEED_OK, bits = 0, k = 0, alphabet = 0, remap = 0;
};

static SeedDecision seed_checks_before_scan(uint32_t policy, uint64_t max_ctx, uint32_t sym_bytes) {
    SeedDecision d;
    if (policy == CAPS_SEED_POLICY_DISABLED) d.reason = CAPS_SEED_DISABLED;
    else if (max_ctx != ~0ull) d.reason = CAPS_SEED_MAX_CONTEXT;
    else if (sym_bytes != 1) d.reason = CAPS_SEED_NOT_U8;
    // boundary_rank(): every limits descriptor declares one (ShorterFirst or
    // LongerFirst), so the provider check always passes here
    return d;
}

static void seed_checks_after_scan(SeedDecision &d, uint32_t policy, uint64_t budget, uint64_t n, uint32_t distinct,
                                   bool identity) {
    d.alphabet = distinct;
    if (!identity) {  // the ranked copy: Remap only, within its byte budget
        if (policy != CAPS_SEED_POLICY_REMAP) {
            d.reason = CAPS_SEED_GAPPED_ALPHABET;
            return;
        }
        if (n > budget) {
            d.reason = CAPS_SEED_REMAP_BUDGET;
            return;
        }
        d.remap = 1;
    }
    const uint32_t bits = distinct <= 1 ? 1 : distinct <= 3 ? 2 : distinct <= 15 ? 4 : 8;
    if ((uint64_t)distinct >= (1ull << bits)) {  // no sentinel code beside the alphabet
        d.reason = CAPS_SEED_NO_SENTINEL;
        d.remap = 0;
        return;
    }
    d.bits = bits;
    d.k = 64 / bits;
}

// Distinct values of the text (one device scan into a presence bitmap) and
// whether they are already the dense range 0..distinct.
template <typename S>
static void scan_alphabet(Workspace &W, const S *d_text, uint64_t n, hipStream_t st, std::vector<uint32_t> &code,
                          uint32_t &distinct, uint32_t &maxv, bool &identity) {
    constexpr uint32_t WORDS = sizeof(S) == 1 ? 8 : 2048;
    uint32_t *bm = (uint32_t *)W.bitmap.get(WORDS * 4);
    HIPCHK(hipMemsetAsync(bm, 0, WORDS * 4, st));
    if (n) {
        k_alphabet<S><<<grid_for(n, 256, 2048), 256, 0, st>>>(d_text, n, bm);
        LAUNCHCHK();
    }
    std::vector<uint32_t> h(WORDS);
    HIPCHK(hipMemcpyAsync(h.data(), bm, WORDS * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    distinct = 0, maxv = 0;
    identity = true;
    code.assign(WORDS * 32, 0);
    for (uint32_t v = 0; v < WORDS * 32; ++v)
        if (h[v >> 5] >> (v & 31) & 1) {
            code[v] = distinct;
            identity &= (distinct == v);
            ++distinct;
            maxv = v;
        }
}

// ---------------------------------------------------------------------------
// Key geometry from the alphabet (pack.rs:103-176 for the reference form).
// mode CAPS_KEY_REFERENCE: the reference's packed keys when the seed checks
// above pass for (policy, budget, max_ctx), dense keys otherwise.
// ---------------------------------------------------------------------------
template <typename S>
Geometry compute_geometry(Workspace &W, const S *d_text, uint64_t n, uint32_t mode, uint32_t order, hipStream_t st,
                          bool pow2_seed = false, uint32_t policy = CAPS_SEED_POLICY_REMAP, uint64_t budget = ~0ull,
                          uint64_t max_ctx = ~0ull, bool escape32 = false) {
    std::vector<uint32_t> code;
    uint32_t distinct = 0, maxv = 0;
    bool identity = true;
    scan_alphabet<S>(W, d_text, n, st, code, distinct, maxv, identity);
    Geometry G{};
    G.distinct = distinct;
    bool longer_first = order == CAPS_ORDER_STAR;
    bool ref = false;
    if (mode == CAPS_KEY_REFERENCE) {
        SeedDecision d = seed_checks_before_scan(policy, max_ctx, sizeof(S));
        if (d.reason == CAPS_SEED_OK) seed_checks_after_scan(d, policy, budget, n, distinct, identity);
        if (d.reason == CAPS_SEED_OK) {
            ref = true;
            G.g.bits = d.bits;
            G.g.k = d.k;
            G.g.shift = 0;
            G.g.bias = longer_first ? 0 : 1;
            G.g.pad = longer_first ? distinct : 0;
            G.remap = !identity;
        }
    }
    if (!ref) {
        // seed keys: ceil(log2(distinct))-bit fields, as many symbols as fit
        // in 64 bits; the packed code text (and the refinement's chunks) use
        // the power of two above (1, 2, 4, 8, 16), so codes tile u64 words
        // exactly, as in the reference's packer
        uint32_t bits = 1;
        while ((1ull << bits) < distinct) ++bits;
        uint32_t pb = 1;
        while (pb < bits) pb <<= 1;
        if (pow2_seed) bits = pb;
        G.remap = (uint64_t)maxv >= (1ull << bits);
        if (escape32 && sizeof(S) == 1 && distinct >= 5 && distinct <= 16) {
            // 32 two-bit fields, codes >= 4 escape (the packed text keeps 4-bit codes)
            bits = 2;
            pb = 4;
            G.g.esc = 4;
            G.remap = (uint64_t)maxv >= 16;
        }
        G.g.bits = bits;
        G.g.k = 64 / bits;
        G.g.shift = 64 - G.g.k * bits;
        G.g.bias = 0;
        G.g.pad = longer_first ? (uint32_t)((1ull << bits) - 1) : 0;
        G.pbits = pb;
    }
    G.reference = ref;
    if (G.remap) {
        if (sizeof(S) == 1) {
            W.h_code8.resize(256);
            for (int v = 0; v < 256; ++v) W.h_code8[v] = (uint8_t)code[v];
            void *d = W.code.get(256);
            HIPCHK(hipMemcpyAsync(d, W.h_code8.data(), 256, hipMemcpyHostToDevice, st));
            G.code = d;
        } else {
            W.h_code16.resize(65536);
            for (int v = 0; v < 65536; ++v) W.h_code16[v] = (uint16_t)code[v];
            void *d = W.code.get(65536 * 2);
            HIPCHK(hipMemcpyAsync(d, W.h_code16.data(), 65536 * 2, hipMemcpyHostToDevice, st));
            G.code = d;
        }
    }
    return G;
}

template <typename S>
static CodeMap<S> codemap(const Geometry &G) {
    CodeMap<S> cm;
    cm.code = (const decltype(cm.code))G.code;
    return cm;
}

// ---------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------
struct PhaseTimer {
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    std::vector<int> tag;
    explicit PhaseTimer(hipStream_t s) : st(s) { mark(-1); }
    void mark(int t) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, st));
        ev.push_back(e);
        tag.push_back(t);
    }
    ~PhaseTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    void collect(caps_stats_t &s) {
        double *slot[] = {&s.ms_alphabet, &s.ms_sample, &s.ms_classify, &s.ms_scatter, &s.ms_tile_sort,
                          &s.ms_merge,    &s.ms_detect, &s.ms_refine,   &s.ms_giant};
        for (size_t i = 1; i < ev.size(); ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
            if (tag[i] >= 0 && tag[i] < 9) *slot[tag[i]] += ms;
        }
        float tot = 0;
        HIPCHK(hipEventElapsedTime(&tot, ev.front(), ev.back()));
        s.ms_total = tot;
    }
};
enum { PH_ALPHA, PH_SAMPLE, PH_CLASSIFY, PH_SCATTER, PH_TILE, PH_MERGE, PH_DETECT, PH_REFINE, PH_GIANT };

// Event pair around one kernel launch.
struct KernelTimer {
    hipStream_t st;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<double *> dst;
    explicit KernelTimer(hipStream_t s) : st(s) {}
    void begin(double *slot) {
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, st));
        ev.push_back({a, b});
        dst.push_back(slot);
    }
    void end() { HIPCHK(hipEventRecord(ev.back().second, st)); }
    void collect() {
        for (size_t i = 0; i < ev.size(); ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[i].first, ev[i].second));
            *dst[i] += ms;
        }
    }
    ~KernelTimer() {
        for (auto &p : ev) (void)hipEventDestroy(p.first), (void)hipEventDestroy(p.second);
    }
};

// The run memo of the packed code text (k_find_runs, then the host sorts the
// runs by start and k_run_dir builds the block directory).  At most kRunCap
// runs are kept; a subset is as correct as all of them.
constexpr uint64_t kRunCap = 1ull << 20;
template <typename S>
static RunTable build_run_memo(Workspace &W, const CodeMap<S> &cm, uint64_t n, hipStream_t st) {
    RunTable R{};
    if (!cm.packed || n == 0) return R;
    const uint64_t nwords = (n >> cm.lgf) + 1;
    const uint64_t cap = std::min<uint64_t>(kRunCap, nwords / kRunMinWords + 1);  // runs the text can hold
    uint64_t *rs = (uint64_t *)W.run_s.get(cap * 8), *re = (uint64_t *)W.run_e.get(cap * 8);
    unsigned long long *cnt = (unsigned long long *)W.run_count.get(8);
    HIPCHK(hipMemsetAsync(cnt, 0, 8, st));
    k_find_runs<<<grid_for(nwords, 256, 8192), 256, 0, st>>>(cm.packed, nwords, cm.lgf, n, rs, re, cnt, cap);
    LAUNCHCHK();
    unsigned long long total = 0;
    HIPCHK(hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint64_t m = std::min<uint64_t>(total, cap);
    if (m == 0) return R;
    std::vector<uint64_t> hs(m), he(m);
    HIPCHK(hipMemcpyAsync(hs.data(), rs, m * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(he.data(), re, m * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    W.h_runs.resize(m);
    for (uint64_t i = 0; i < m; ++i) W.h_runs[i] = {hs[i], he[i]};
    std::sort(W.h_runs.begin(), W.h_runs.end());
    for (uint64_t i = 0; i < m; ++i) hs[i] = W.h_runs[i].first, he[i] = W.h_runs[i].second;
    uint64_t *ss = (uint64_t *)W.run_ss.get(m * 8), *se = (uint64_t *)W.run_se.get(m * 8);
    HIPCHK(hipMemcpyAsync(ss, hs.data(), m * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(se, he.data(), m * 8, hipMemcpyHostToDevice, st));
    const uint64_t n_dir = (n >> kRunDirShift) + 1;
    uint32_t *dir = (uint32_t *)W.run_dir.get((n_dir + 1) * 4);
    k_run_dir<<<(unsigned)ceil_div(n_dir + 1, 256), 256, 0, st>>>(se, m, n_dir, dir);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(st));  // the host vectors are pageable staging for the copies above
    R.start = ss, R.end = se, R.dir = dir, R.n_runs = m, R.n_dir = n_dir;
    return R;
}

// A phase of the build captured into a hipGraph (CAPS_FLAG_GRAPH): while
// the section is open its launches go to a private capture stream (`ls`);
// finish() ends the capture, updates the phase's instantiated graph in place
// when its kernels are the same as last time (only grids and arguments
// differ) or instantiates it again, and launches it on the build's stream.
// Closed, the section launches straight on the build's stream.
struct GraphSection {
    Workspace &W;
    hipStream_t st, ls;
    int slot;
    bool capturing = false;
    GraphSection(Workspace &w, hipStream_t s, bool on, int sl) : W(w), st(s), ls(s), slot(sl) {
        if (!on) return;
        if (!W.cap_stream) HIPCHK(hipStreamCreateWithFlags(&W.cap_stream, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(W.cap_stream, hipStreamCaptureModeRelaxed));
        capturing = true;
        ls = W.cap_stream;
    }
    // returns true when a graph was launched (timed by the caller's event pair)
    void finish(KernelTimer &kt, double *slot_ms) {
        if (!capturing) return;
        capturing = false;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(W.cap_stream, &g));
        hipGraphExec_t &ex = W.graph_exec[slot];
        bool updated = false;
        if (ex) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(ex, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess)
                updated = true;
            else {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(ex);
                ex = nullptr;
            }
        }
        hipError_t e = updated ? hipSuccess : hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        kt.begin(slot_ms);
        HIPCHK(hipGraphLaunch(ex, st));
        kt.end();
    }
    ~GraphSection() {
        if (capturing) {  // an exception left the capture open
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(W.cap_stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
        }
    }
};

// Sort `count` keys (with u32 payload) in place: LDS tiles + key merges.
template <int T>
static void sort_keys_single_range(Workspace &W, uint64_t *keys, uint32_t *ids, uint64_t *keys2, uint32_t *ids2,
                                   uint64_t count, hipStream_t st, uint64_t **out_keys, uint32_t **out_ids) {
    uint64_t nt = ceil_div(count, T);
    W.h_tiles.resize(nt);
    for (uint64_t i = 0; i < nt; ++i) {
        W.h_tiles[i].start = i * T;
        W.h_tiles[i].len = (uint32_t)std::min<uint64_t>(T, count - i * T);
        W.h_tiles[i].whole = 0;
    }
    Tile *dt = (Tile *)W.stiles.get(nt * sizeof(Tile));
    HIPCHK(hipMemcpyAsync(dt, W.h_tiles.data(), nt * sizeof(Tile), hipMemcpyHostToDevice, st));
    GroupLists none{};
    k_tile_merge<uint32_t, T><<<(unsigned)nt, 256, 0, st>>>(keys, ids, dt, none, 0, 1);
    LAUNCHCHK();
    Range r{0, count};
    W.h_range = r;
    Range *dr = (Range *)W.srange.get(sizeof(Range));
    HIPCHK(hipMemcpyAsync(dr, &W.h_range, sizeof(Range), hipMemcpyHostToDevice, st));
    uint64_t *ks = keys, *kd = keys2;
    uint32_t *is = ids, *id = ids2;
    static_assert(T >= kKeyMergeTile / 2 && (T & (T - 1)) == 0, "key merge runs: powers of two >= half a tile");
    for (uint64_t run = T; run < count; run *= 2) {
        dim3 grid((unsigned)ceil_div(count, kKeyMergeTile), 1);
        k_key_merge<uint32_t><<<grid, 256, 0, st>>>(dr, run, ks, is, kd, id);
        LAUNCHCHK();
        std::swap(ks, kd);
        std::swap(is, id);
    }
    *out_keys = ks;
    *out_ids = is;
}

// ---------------------------------------------------------------------------
// Router training (on the GPU, from the sample the build already draws).
// ---------------------------------------------------------------------------
static std::vector<float> mlp_init_params(uint64_t seed) {
    std::vector<float> p(kMlpParams, 0.f);
    uint64_t x = seed;
    auto next = [&]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto unif = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
    for (int i = 0; i < kMlpIn * kMlpHid; ++i) p[i] = (float)(unif() * 0.25);         // W1
    for (int j = 0; j < kMlpHid; ++j) p[kMlpIn * kMlpHid + j] = (float)(unif() * 0.1);  // b1
    for (int j = 0; j < kMlpHid; ++j) p[kMlpIn * kMlpHid + kMlpHid + j] = (float)(unif() * 0.1);  // w2
    p[kMlpParams - 1] = 0.5f;                                                         // b2
    return p;
}

// Spline router: knots from the sorted splitter table ctab[0..nc), then its
// exact error bound into bound[0..2) (zeroed by the caller).
static void spline_fit(const uint64_t *ctab, uint32_t nc, uint16_t *kn, int *bound, hipStream_t st) {
    k_spline_knots<<<(unsigned)ceil_div(kSplineCells + 1, 256), 256, 0, st>>>(ctab, nc, kn);
    LAUNCHCHK();
    k_spline_bound<<<(unsigned)ceil_div(2ull * nc + 2 * kSplineCells, 256), 256, 0, st>>>(ctab, nc, kn, bound);
    LAUNCHCHK();
}

template <typename S, typename I>
static RouterDev train_router(Workspace &W, uint32_t kind, uint32_t flags, const uint64_t *sorted, uint64_t s,
                              const Splitters &SP, const S *d_text, CodeMap<S> cm, KeyGeom g, DevLimits L,
                              uint64_t max_ctx, const I *d_pos, uint64_t n_pos, hipStream_t st,
                              const uint64_t *d_keep, uint64_t n) {
    // the router predicts the position among the coarse splitters (0..n_coarse)
    const uint32_t nc = SP.n_coarse;
    uint64_t *ctab = (uint64_t *)W.r_tmp.get((uint64_t)std::max<uint32_t>(nc, 1) * 8);
    k_coarse_table<<<(unsigned)ceil_div(std::max<uint32_t>(nc, 1), 256), 256, 0, st>>>(SP, ctab);
    LAUNCHCHK();
    const uint64_t *spl = ctab;
    const uint64_t p = (uint64_t)nc + 1;
    RouterDev R{};
    R.kind = kind;
    R.window = 8;
    R.n_buckets = (uint32_t)p;
    R.scale = (float)(p - 1);
    unsigned long long *stats = (unsigned long long *)W.r_stats.get(64);
    HIPCHK(hipMemsetAsync(stats, 0, 64, st));
    R.stats = stats;
    if (kind == CAPS_ROUTER_PWL) {
        uint32_t ns = (uint32_t)std::min<uint64_t>(kPwlSeg, s);
        uint32_t *labels = (uint32_t *)W.r_labels.get(s * 4);
        k_router_labels<<<(unsigned)ceil_div(s, 256), 256, 0, st>>>(sorted, s, spl, (uint32_t)(p - 1), labels,
                                                                     nullptr, 0.f);
        LAUNCHCHK();
        uint64_t *knots = (uint64_t *)W.r_knots.get((ns + 1) * 8);
        float *a = (float *)W.r_a.get(ns * 4), *b = (float *)W.r_b.get(ns * 4), *inv = (float *)W.r_inv.get(ns * 4);
        uint32_t *lo = (uint32_t *)W.r_lo.get(ns * 4), *hi = (uint32_t *)W.r_hi.get(ns * 4);
        uint16_t *radix = (uint16_t *)W.r_radix.get(((1u << kRouterRadixBits) + 1) * 2);
        k_pwl_fit<<<ns, 256, 0, st>>>(sorted, labels, s, ns, (uint32_t)p, knots, a, b, inv, lo, hi);
        LAUNCHCHK();
        k_pwl_radix<<<(unsigned)ceil_div((1u << kRouterRadixBits) + 1, 256), 256, 0, st>>>(knots, ns, radix);
        LAUNCHCHK();
        R.knots = knots, R.a = a, R.b = b, R.inv = inv, R.lo = lo, R.hi = hi, R.radix = radix, R.ns = ns;
    } else if (kind == CAPS_ROUTER_MLP) {
        // seeded 1% suffix sample (at least 64K), labels from the splitters
        uint64_t m = std::max<uint64_t>(65536, n_pos / 100);
        m = std::min<uint64_t>(m, std::max<uint64_t>(n_pos, 1) * 4);
        m = std::max<uint64_t>(m, 8192);  // at least one minibatch (sampling is with replacement)
        m = (m + 255) / 256 * 256;
        uint64_t *keys = (uint64_t *)W.r_x.get(m * 8);
        float *y = (float *)W.r_y.get(m * 4);
        k_random_keys<S, I><<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, m,
                                                                         0x5EEDull, keys, d_keep, n);
        LAUNCHCHK();
        k_router_labels<<<(unsigned)ceil_div(m, 256), 256, 0, st>>>(keys, m, spl, (uint32_t)(p - 1), nullptr, y,
                                                                     1.f / (float)std::max<uint64_t>(p - 1, 1));
        LAUNCHCHK();
        float *params = (float *)W.r_w.get(kMlpParams * 4 * 4);
        float *grad = params + kMlpParams, *am = grad + kMlpParams, *av = am + kMlpParams;
        W.h_mlp = mlp_init_params(0xC0FFEEull);
        HIPCHK(hipMemcpyAsync(params, W.h_mlp.data(), kMlpParams * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(am, 0, 2 * kMlpParams * 4, st));
        const uint32_t B = 8192;
        const int steps = (flags & CAPS_FLAG_ROUTER_UNTRAINED) ? 0 : 300;
        double b1p = 1.0, b2p = 1.0;
        for (int it = 0; it < steps; ++it) {
            uint64_t off = ((uint64_t)it * B) % (m - B + 1);
            b1p *= 0.9, b2p *= 0.999;
            HIPCHK(hipMemsetAsync(grad, 0, kMlpParams * 4, st));
            k_mlp_grad<<<B / 256, 256, 0, st>>>(params, keys + off, y + off, B, B, grad, nullptr);
            k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, st>>>(params, grad, am, av, 3e-3f, (float)b1p,
                                                                 (float)b2p);
        }
        LAUNCHCHK();
        float *nib = (float *)W.r_nib.get(16 * 16 * kMlpHid * 4);
        k_mlp_nibbles<<<16, 16 * kMlpHid, 0, st>>>(params, nib);
        LAUNCHCHK();
        R.nib = nib;
        R.w2 = params + kMlpIn * kMlpHid + kMlpHid;
    } else if (kind == CAPS_ROUTER_SPLINE) {
        uint16_t *kn = (uint16_t *)W.r_radix.get((kSplineCells + 1) * 2);
        int *bound = (int *)(stats + 4);  // zeroed with the statistics
        spline_fit(ctab, nc, kn, bound, st);
        int hb[2];
        HIPCHK(hipMemcpyAsync(hb, bound, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        R.spline = kn;
        R.el = -hb[0];
        R.eh = hb[1];
        if (flags & CAPS_FLAG_ROUTER_UNTRAINED) R.el = R.eh = 0;  // no bound: the fallback's test
        // steps over the window widened by one position on each side: its
        // positions lie within el + eh + 2 of its first
        uint32_t w = (uint32_t)(R.el + R.eh + 2), bl = 0;
        while (w >> bl) ++bl;
        R.steps = bl;
    }
    return R;
}

// ---------------------------------------------------------------------------
// The build.
// ---------------------------------------------------------------------------
// Device addresses of the refinement work counters (looked up once per
// device) and whether the current device's counting flag is set.
struct CountAddr {
    unsigned long long *sym;
    unsigned int *flag;
    bool &on;
};
static CountAddr count_symbols_addr() {
    static unsigned long long *sym[64] = {};
    static unsigned int *flag[64] = {};
    static bool on[64] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw CapsError(CAPS_E_UNSUPPORTED, "device ordinal above 63");
    if (!sym[dev]) {
        void *a = nullptr, *b = nullptr;
        HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(g_cmp_symbols)));
        HIPCHK(hipGetSymbolAddress(&b, HIP_SYMBOL(g_count_symbols)));
        sym[dev] = (unsigned long long *)a;
        flag[dev] = (unsigned int *)b;
    }
    return CountAddr{sym[dev], flag[dev], on[dev]};
}

// idbits > 0 builds many texts at once: `lim` delimits the texts and the top
// idbits bits of every packed key hold the text index, so suffixes sort
// text-major and no comparison crosses a text end.
constexpr uint32_t kBucketTarget = 1536;

// Launch style per refinement size class (pairs, 3-4, 5-8, 9-16, 17-32,
// 33-64 members).  CAPS_FLAG_PERSISTENT_REFINE queues every class; the
// default queues the classes where slots finish at very different rounds.
struct RefinePlan {
    bool queue[6];
};
static RefinePlan refine_plan(const caps_opts_t &o) {
    RefinePlan p{};
    const bool all = (o.flags & CAPS_FLAG_PERSISTENT_REFINE) != 0;
    for (int c = 0; c < 6; ++c) p.queue[c] = all || c == 1 || c == 2;
    return p;
}

// Receives each finished batch of the suffix array (device pointer, count,
// the build's stream) in order, instead of the build writing into d_out.
template <typename I>
struct BatchSink {
    std::function<void(const I *, uint64_t, hipStream_t)> flush;
};

// Bucket-batch plan.  Per record a batch needs at most: 16-byte record,
// 8-byte key, about 20 bytes of group lists (sized for the worst case of
// every class), 8-byte output staging and 32 bytes of merge scratch (keys,
// positions, two LCP arrays) when oversized buckets or giant groups need
// them: 112 bytes with margin; 96 without the record when the
// records of all batches are written at once (pass 1 runs once instead of
// once per batch).  75% of the free HBM is used; buffers the workspace
// already holds count as free, since this build reuses them.
struct BatchPlan {
    uint64_t bmax;     // records per batch
    bool all_records;  // scatter pass 1 writes every record once
};
static BatchPlan plan_batches(const Workspace &W, uint64_t n_pos, uint64_t requested, size_t rec_bytes,
                              bool per_batch_pass1) {
    if (g_arena) return {n_pos, false};  // no-alloc builds: one batch, sized by the scratch
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b == 0) return {requested ? requested : n_pos, false};
    // every buffer this build allocates after planning counts as free, so
    // that a repeated build of the same input plans the same batches (a
    // different plan would regrow multi-GB buffers in the middle of a build)
    size_t held = W.recs.cap + W.keys.cap + W.brec.cap + W.sinkbuf.cap + W.keys2.cap + W.idx2.cap + W.lcp.cap +
                  W.lcp2.cap + W.tiles.cap + W.big.cap + W.chunks.cap + W.jobs.cap + W.run_s.cap + W.run_e.cap +
                  W.run_ss.cap + W.run_se.cap + W.run_dir.cap;
    for (int c = 0; c < kGroupClasses; ++c) held += W.glist[c].cap;
    held += W.mid.cap;
    const double avail = (double)(free_b + held) * 0.75;
    if (!requested && (double)n_pos * 112 <= avail) return {n_pos, false};  // one batch
    const double rest = avail - (double)n_pos * rec_bytes;  // beside all records
    // batch sizes in whole quanta of 2^24 records: small changes in free
    // memory between builds leave the plan as it was
    constexpr uint64_t kQuantum = 1ull << 24;
    const uint64_t b_each =
        requested ? requested : std::max<uint64_t>(kQuantum, (uint64_t)(avail / 112) / kQuantum * kQuantum);
    if (!per_batch_pass1 && rest > 0) {
        if (requested) {
            if ((double)requested * 96 <= rest) return {requested, true};
        } else {
            // one pass 1 for all records (and smaller batches beside them)
            // unless that takes many more batches: a pass 1 over every
            // position costs about as much as nine batches' own overhead
            // (whole genome: 47 ms a pass against about 5 ms a batch; 10
            // batches with one pass 1 build in 464 ms, 5 with a pass 1 each
            // in 580 ms)
            const uint64_t b = (uint64_t)(rest / 96) / kQuantum * kQuantum;
            const uint64_t nb_per = ceil_div(n_pos, b_each);
            if (b && ceil_div(n_pos, b) <= nb_per + 9 * (nb_per - 1)) return {std::min<uint64_t>(b, n_pos), true};
        }
    }
    return {std::min<uint64_t>(b_each, n_pos), false};
}

// Clears the bits past n of a device bitmap and returns its kept count.
static uint64_t device_keep_count(Workspace &W, uint64_t *d_keep, uint64_t n, hipStream_t st) {
    unsigned long long *cnt = (unsigned long long *)W.keep_count.get(8);
    HIPCHK(hipMemsetAsync(cnt, 0, 8, st));
    if (n) {
        k_keep_count<<<grid_for(ceil_div(n, 64), 256, 4096), 256, 0, st>>>(d_keep, n, cnt);
        LAUNCHCHK();
    }
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
}

// Position of the only kept position of a bitmap (n_pos == 1 builds).
template <typename I>
__global__ void k_only_kept(const uint64_t *__restrict__ keep, uint64_t n, I *__restrict__ out) {
    const uint64_t nw = (n + 63) >> 6;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x)
        if (keep[w]) out[0] = (I)((w << 6) + (uint64_t)__builtin_ctzll(keep[w]));
}

// d_keep (optional): a bitmap of kept positions (bit q & 63 of word q >> 6;
// n_pos = its population count, d_pos null).  sink (optional): batches go
// to it instead of d_out.
template <typename S, typename I>
void build_impl(const S *d_text, uint64_t n, const I *d_pos, uint64_t n_pos, const caps_limits_t *lim,
                const caps_opts_t &o, I *d_out, I *d_lcp_out, hipStream_t st, caps_stats_t &stats,
                uint32_t idbits = 0, const uint64_t *d_keep = nullptr, BatchSink<I> *sink = nullptr) {
    Workspace &W = workspace();
    if (!W.n_cu) {
        int dev = 0, cus = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        W.n_cu = cus > 0 ? (uint32_t)cus : 256u;
    }
    memset(&stats, 0, sizeof(stats));
    // refinement work counters: on for a measurement build only
    const bool counting = (o.flags & CAPS_FLAG_COUNT_WORK) != 0;
    const CountAddr count_addr = count_symbols_addr();
    if (counting || count_addr.on) {
        HIPCHK(hipMemsetD32Async((hipDeviceptr_t)count_addr.flag, counting ? 1 : 0, 1, st));
        if (counting) HIPCHK(hipMemsetAsync(count_addr.sym, 0, 3 * kCountLanes * 8, st));
        count_addr.on = counting;
    }
    PhaseTimer tm(st);
    KernelTimer kt(st);
    const uint64_t max_ctx = o.max_context;
    DevLimits L = upload_limits(W, lim, n, st);
    if (n_pos == 0) {
        tm.mark(PH_ALPHA);
        HIPCHK(hipStreamSynchronize(st));
        tm.collect(stats);
        return;
    }
    if (d_pos) {  // every position must address the text: refuse the build before any key gather
        // A strictly increasing list covering at least half of the text is
        // the same set as a kept-position bitmap, whose builds walk the text
        // in LDS-staged windows: about 21 ps per text position for classify
        // and scatter pass 1 against 35 ps per listed position (820M genome
        // share: 48.3 against 59.5 ms a build).  The suffix array is the
        // same.  The check writes the bitmap on the way.
        const bool try_keep = !d_keep && !idbits && !g_arena && n_pos < n && 2 * n_pos >= n;
        unsigned long long *kb = nullptr;
        if (try_keep) {
            kb = (unsigned long long *)W.keep.get(ceil_div(n, 64) * 8);
            HIPCHK(hipMemsetAsync(kb, 0, ceil_div(n, 64) * 8, st));
        }
        unsigned int *bad = (unsigned int *)W.poscheck.get(4);
        unsigned int hbad = 0;
        HIPCHK(hipMemsetAsync(bad, 0, 4, st));
        k_positions_check<I><<<grid_for(n_pos, 256, 8192), 256, 0, st>>>(d_pos, n_pos, n, bad, kb);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hbad & 1u) throw CapsError(CAPS_E_INVALID, "position out of range");
        if (try_keep && !(hbad & 2u)) {
            d_keep = (const uint64_t *)kb;
            d_pos = nullptr;
        }
    }
    if (d_pos && (const void *)d_pos == (const void *)d_out) {
        I *cp = (I *)W.pos_copy.get(n_pos * sizeof(I));
        HIPCHK(hipMemcpyAsync(cp, d_pos, n_pos * sizeof(I), hipMemcpyDeviceToDevice, st));
        d_pos = cp;
    }
    const bool want_lcp = d_lcp_out != nullptr && !sink;
    const bool graph = (o.flags & CAPS_FLAG_GRAPH) != 0;
    if (n_pos == 1) {
        I *o1 = sink ? (I *)W.sinkbuf.get(16) : d_out;
        if (d_pos) HIPCHK(hipMemcpyAsync(o1, d_pos, sizeof(I), hipMemcpyDeviceToDevice, st));
        else if (d_keep) k_only_kept<I><<<grid_for(ceil_div(n, 64), 256, 1024), 256, 0, st>>>(d_keep, n, o1);
        else HIPCHK(hipMemsetAsync(o1, 0, sizeof(I), st));
        if (want_lcp) HIPCHK(hipMemsetAsync(d_lcp_out, 0, sizeof(I), st));
        tm.mark(PH_ALPHA);
        HIPCHK(hipStreamSynchronize(st));
        tm.collect(stats);
        if (sink) sink->flush(o1, 1, st);
        return;
    }

    // ---- alphabet and key geometry ----
    Geometry G = compute_geometry<S>(W, d_text, n, idbits ? CAPS_KEY_DENSE : o.key_geometry, L.order, st,
                                     (o.flags & CAPS_FLAG_POW2_SEED) != 0, o.packed_prefix_seed, o.remap_budget,
                                     o.max_context, !idbits && (o.flags & CAPS_FLAG_SEED_ESCAPE32) != 0);
    KeyGeom gref = G.g;  // refinement chunks: packed-text field width, full width, no text index
    if (!G.reference) {
        gref.bits = G.pbits;
        gref.k = 64 / G.pbits;
        gref.shift = 0;
        gref.pad = L.order == CAPS_ORDER_STAR ? (uint32_t)((1ull << G.pbits) - 1) : 0;
        gref.esc = 0;  // refinement chunks hold exact codes
    }
    KeyGeom g = G.g;  // seed keys
    if (idbits) {
        if (idbits + G.g.bits > 64) throw CapsError(CAPS_E_INVALID, "too many texts in one batch");
        g.idbits = idbits;
        g.k = (64 - idbits) / g.bits;
        g.shift = 64 - idbits - g.k * g.bits;
    }
    CodeMap<S> cm = codemap<S>(G);
    if (!G.reference) {  // packed code text for the key gathers
        uint32_t lgf = 0;
        while ((gref.bits << lgf) < 64) ++lgf;
        const uint64_t nwords = (n >> lgf) + 2;
        uint64_t *pk = (uint64_t *)W.packed.get(nwords * 8);
        k_pack_codes<S><<<grid_for(nwords, 256, 8192), 256, 0, st>>>(d_text, n, cm, gref.bits, lgf, pk, nwords);
        LAUNCHCHK();
        cm.packed = pk;
        cm.lgf = lgf;
        // seed keys: the packed text itself, or a text in their narrower fields
        if (n >= (1ull << 40)) throw CapsError(CAPS_E_INVALID, "text too long for packed seed keys");
        const uint32_t K = 64 / g.bits;
        if (g.esc) {
            cm.seed = nullptr;  // escape keys come from the text bytes (their fields cannot hold every code)
        } else if (g.bits == gref.bits) {
            cm.seed = pk;
        } else {
            const uint64_t swords = n / K + 2;
            uint64_t *sp = (uint64_t *)W.seedpk.get(swords * 8);
            k_pack_seed<S><<<grid_for(swords, 256, 8192), 256, 0, st>>>(d_text, n, cm, g.bits, K, sp, swords);
            LAUNCHCHK();
            cm.seed = sp;
        }
        cm.sk = K;
        cm.srec = (uint64_t)(~0ull / K) + 1;  // ceil(2^64 / K)
    }
    stats.key_bits = g.bits;
    stats.key_k = g.k;
    tm.mark(PH_ALPHA);

    // ---- buckets and splitters ----
    constexpr uint32_t T = 4096;
    const bool every_suffix = sizeof(I) == 4 && !idbits && !d_keep && d_pos == nullptr && n_pos == n;
    // mean bucket size the splitters aim at: most buckets fit the
    // 2048-record tile class (sample-based splitters spread bucket sizes by
    // about a quarter around the mean); 1,536 rather than 1,280 records:
    // fewer buckets for scatter pass 2 against more 4,096-record tiles
    // (268M text: 14.50-14.59 against 14.69-14.76 ms; whole genome 343.5
    // against 346.8 ms)
    const uint64_t target = o.tile_keys ? std::max<uint32_t>(64, std::min<uint32_t>(o.tile_keys, T)) : kBucketTarget;
    // p buckets in C coarse intervals of `stride` buckets.  C is the number of
    // classification buckets (the classify pass decides among them, by binary
    // search or the router): gpu_buckets when set, else 1,024 (more when the
    // intervals would exceed kFineMax buckets).  The fine buckets aim at
    // `target` records each.
    if (o.gpu_buckets > kCoarseBucketsMax) throw CapsError(CAPS_E_INVALID, "gpu_buckets above 4096");
    const uint64_t want = std::max<uint64_t>(1, ceil_div(n_pos, target));
    // 1,024 classification buckets for every suffix of one text with a u32
    // index (longer runs per interval in scatter pass 1), else 2,048:
    // smaller intervals keep scatter pass 2's LDS small enough for several
    // workgroups per compute unit with 16-byte records
    const uint64_t c_default = every_suffix ? 1024 : 2048;
    uint64_t C = o.gpu_buckets ? o.gpu_buckets
                               : std::min<uint64_t>(kCoarseBucketsMax,
                                                    std::max<uint64_t>(std::min<uint64_t>(want, c_default),
                                                                       ceil_div(want, kFineMax)));
    C = std::max<uint64_t>(1, std::min<uint64_t>(C, n_pos));
    uint64_t stride = std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(want, C), n_pos / C));
    stride = std::min<uint64_t>(stride, kFineMax);
    uint64_t p = C * stride;
    // sample: 16 keys per bucket and at least 1% of the positions (the
    // router's training sample), at most all of them
    const uint64_t oversample = 16;
    uint64_t s = p > 1 ? std::min<uint64_t>(n_pos, std::max<uint64_t>(p * oversample, n_pos / 100)) : 0;
    uint64_t *spl = (uint64_t *)W.spl.get(std::max<uint64_t>(p, 1) * 8);
    uint64_t *sorted_sample = nullptr;
    if (p > 1) {
        uint64_t *sk = (uint64_t *)W.skeys.get(s * 8), *sk2 = (uint64_t *)W.skeys2.get(s * 8);
        uint32_t *si = (uint32_t *)W.sids.get(s * 4), *si2 = (uint32_t *)W.sids2.get(s * 4);
        k_sample_keys<S, I><<<grid_for(s, 256, 1u << 30), 256, 0, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, s,
                                                                         sk, si, d_keep, n);
        LAUNCHCHK();
        uint32_t *sorted_ids;
        sort_keys_single_range<T>(W, sk, si, sk2, si2, s, st, &sorted_sample, &sorted_ids);
        k_pick_splitters<<<(unsigned)ceil_div(p, 256), 256, 0, st>>>(sorted_sample, s, (uint32_t)p, spl);
        LAUNCHCHK();
    }
    Splitters SP{};
    SP.spl = spl;
    SP.n_spl = (uint32_t)(p - 1);
    SP.stride = (uint32_t)stride;
    SP.n_coarse = (uint32_t)(C - 1);
    while ((1ull << SP.coarse_steps) <= SP.n_coarse) ++SP.coarse_steps;
    while ((1ull << SP.fine_steps) < SP.stride) ++SP.fine_steps;
    stats.buckets = p;

    // ---- classify (coarse interval) + per-block histograms ----
    if (SP.stride > kFineMax) throw CapsError(CAPS_E_INVALID, "too many buckets");
    const uint32_t pc = SP.n_coarse + 1;
    uint16_t *cid = (uint16_t *)W.bid.get(n_pos * 2);
    uint64_t *hist = (uint64_t *)W.hist.get(pc * 8);
    uint64_t *coff = (uint64_t *)W.toff.get((std::max<uint64_t>(pc, p) + 1) * 8);
    uint64_t *off = (uint64_t *)W.off.get((p + 1) * 8);
    RouterDev R{};
    uint32_t rkind = p > 1 ? o.router : CAPS_ROUTER_BSEARCH;
    if (rkind > CAPS_ROUTER_SPLINE) throw CapsError(CAPS_E_INVALID, "unknown router");
    if (rkind != CAPS_ROUTER_BSEARCH)
        R = train_router<S, I>(W, rkind, o.flags, sorted_sample, s, SP, d_text, cm, g, L, max_ctx, d_pos, n_pos, st,
                               d_keep, n);
    tm.mark(PH_SAMPLE);
    kt.begin(&stats.k_classify);
    // sequential: blocks own ranges of text positions and take keys from
    // LDS-staged text windows (every position, or the kept ones of a bitmap)
    const bool seq = d_keep != nullptr || (d_pos == nullptr && n_pos == n);
    const uint64_t n_iter = seq ? n : n_pos;  // positions the classify blocks walk
    if (seq && n_iter > n_pos) cid = (uint16_t *)W.bid.get(n_iter * 2);
    // Every block owns one contiguous range of positions (a multiple of the
    // key window), in both the classify pass and scatter pass 1: the classify
    // block's histogram is then exactly the scatter block's run sizes.  About
    // one resident wave of classify blocks; the coarse splitter table is
    // staged once per block.
    // The grid follows the binary-search kernel whichever router runs, so
    // the block ranges (and scatter pass 1's grid, which derives from them)
    // do not depend on the router.
    const size_t cls_lds = classify_lds_bytes(SP.n_coarse);
    const bool cls16 = pc > 2048;  // positions per classify thread: 16, else 8
    const unsigned resident = !seq    ? resident_blocks(k_classify<S, I, 0>, 256, cls_lds)
                              : cls16 ? resident_blocks(k_classify_seq<S, I, 0, false, 16>, 256, cls_lds)
                                      : resident_blocks(k_classify_seq<S, I, 0, false, 8>, 256, cls_lds);
    const uint64_t chunk =
        std::max<uint64_t>(kStageStep, ceil_div(ceil_div(n_iter, (uint64_t)resident), kStageStep) * kStageStep);
    const unsigned nblk = (unsigned)ceil_div(n_iter, chunk);
    uint32_t *hblk = (uint32_t *)W.hblk.get((uint64_t)nblk * pc * 4);
    auto classify = [&](uint32_t kind) {
        if (seq) {
            auto classify_seq = [&](auto rk, auto kp) {
                constexpr int RK = decltype(rk)::value;
                constexpr bool KP = decltype(kp)::value;
                if constexpr (RK == 1 || RK == 2)
                    k_classify_seq<S, I, RK, KP, 16><<<nblk, 256, cls_lds, st>>>(d_text, n, cm, g, L, max_ctx, SP, R,
                                                                               cid, hblk, pc, chunk, d_keep);
                else if (cls16)
                    k_classify_seq<S, I, RK, KP, 16><<<nblk, 256, cls_lds, st>>>(d_text, n, cm, g, L, max_ctx, SP, R,
                                                                               cid, hblk, pc, chunk, d_keep);
                else
                    k_classify_seq<S, I, RK, KP, 8><<<nblk, 256, cls_lds, st>>>(d_text, n, cm, g, L, max_ctx, SP, R,
                                                                              cid, hblk, pc, chunk, d_keep);
            };
            using kno = std::false_type;
            using kyes = std::true_type;
            if (kind == CAPS_ROUTER_PWL)
                d_keep ? classify_seq(std::integral_constant<int, 1>{}, kyes{})
                       : classify_seq(std::integral_constant<int, 1>{}, kno{});
            else if (kind == CAPS_ROUTER_MLP)
                d_keep ? classify_seq(std::integral_constant<int, 2>{}, kyes{})
                       : classify_seq(std::integral_constant<int, 2>{}, kno{});
            else if (kind == CAPS_ROUTER_SPLINE)
                d_keep ? classify_seq(std::integral_constant<int, 3>{}, kyes{})
                       : classify_seq(std::integral_constant<int, 3>{}, kno{});
            else
                d_keep ? classify_seq(std::integral_constant<int, 0>{}, kyes{})
                       : classify_seq(std::integral_constant<int, 0>{}, kno{});
        } else if (kind == CAPS_ROUTER_PWL)
            k_classify<S, I, 1><<<nblk, 256, cls_lds, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, SP, R, cid, hblk, pc, chunk);
        else if (kind == CAPS_ROUTER_MLP)
            k_classify<S, I, 2><<<nblk, 256, cls_lds, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, SP, R, cid, hblk, pc, chunk);
        else if (kind == CAPS_ROUTER_SPLINE)
            k_classify<S, I, 3><<<nblk, 256, cls_lds, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, SP, R, cid, hblk, pc, chunk);
        else
            k_classify<S, I, 0><<<nblk, 256, cls_lds, st>>>(d_text, cm, g, L, max_ctx, d_pos, n_pos, SP, R, cid, hblk, pc, chunk);
        LAUNCHCHK();
        if (kind == rkind) kt.end();  // the build's own classify launch (opened above)
        k_block_offsets<<<(unsigned)ceil_div(pc, kBoCols), kBoCols * kBoRows, 0, st>>>(hblk, nblk, pc, hist);
        LAUNCHCHK();
        k_scan_u64<<<1, 1024, 0, st>>>(hist, pc, coff);
        LAUNCHCHK();
    };
    classify(rkind);
    tm.mark(PH_CLASSIFY);

    // ---- bucket batches: consecutive coarse intervals, at most bmax records each ----
    // (an interval larger than bmax is a batch of its own).  Batches follow
    // key order, so the suffix array is their concatenation.
    const BatchPlan plan_b = plan_batches(W, n_pos, o.batch_records, sizeof(KRec<I>),
                                          (o.flags & CAPS_FLAG_SCATTER_PER_BATCH) != 0);
    const uint64_t bmax = plan_b.bmax;
    // One batch of every record, known without the interval totals: binary
    // search (no router fallback to check), all records within one batch and
    // fewer than 2^32 of them (so is every interval).  The batch then starts
    // without waiting for classification; the totals are read at the end,
    // for the statistics.
    const bool one_batch = rkind == CAPS_ROUTER_BSEARCH && bmax >= n_pos && n_pos < (1ull << 32);
    std::vector<uint64_t> hcoff(pc + 1);
    if (!one_batch) {
        HIPCHK(hipMemcpyAsync(hcoff.data(), coff, (pc + 1) * 8, hipMemcpyDeviceToHost, st));
        unsigned long long spline_fb = 0;  // keys outside the spline's window (none while its bound holds)
        if (rkind == CAPS_ROUTER_SPLINE)
            HIPCHK(hipMemcpyAsync(&spline_fb, R.stats + 2, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (spline_fb) {  // binary search is the judge: classify every key again
            classify(CAPS_ROUTER_BSEARCH);
            HIPCHK(hipMemcpyAsync(hcoff.data(), coff, (pc + 1) * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
    } else {
        hcoff[pc] = n_pos;  // the batch: records [0, n_pos)
    }
    // per-interval counts and in-interval offsets are 32-bit in the scatter
    // passes (block prefixes, LDS bucket counts)
    auto interval_max = [&]() {
        uint64_t m = 0;
        for (uint32_t c = 0; c < pc; ++c) m = std::max<uint64_t>(m, hcoff[c + 1] - hcoff[c]);
        return m;
    };
    if (!one_batch && interval_max() >= (1ull << 32))
        throw CapsError(CAPS_E_UNSUPPORTED, "a classification bucket holds 2^32 or more suffixes");
    stats.coarse_buckets = pc;
    std::vector<uint32_t> cuts{0};
    if (!one_batch)
        for (uint32_t c = 0; c < pc; ++c)
            if (c > cuts.back() && hcoff[c + 1] - hcoff[cuts.back()] > bmax) cuts.push_back(c);
    cuts.push_back(pc);
    uint64_t bcap = 0;
    for (size_t bi = 0; bi + 1 < cuts.size(); ++bi) bcap = std::max(bcap, hcoff[cuts[bi + 1]] - hcoff[cuts[bi]]);
    const bool single = cuts.size() == 2;
    const bool all_records = !single && plan_b.all_records;  // pass 1 once, into a full-size record buffer
    stats.batches = cuts.size() - 1;
    auto first_bucket = [&](uint32_t c) -> uint64_t {  // first bucket of coarse interval c
        return std::min<uint64_t>(p, (uint64_t)c * SP.stride);
    };

    // batch-sized buffers
    uint64_t *keys = (uint64_t *)W.keys.get(bcap * 8 + 8);
    KRec<I> *tmp = (KRec<I> *)W.recs.get((all_records ? n_pos : bcap) * sizeof(KRec<I>) + 16);
    // merge tile sort: scatter pass 2 writes bucketed records (key and index
    // side by side, one write stream per bucket run instead of two) and the
    // tile sort reads them, writing the sorted index values to the output and
    // sorted keys where they are read again (oversized buckets, LCP builds:
    // the boundary LCPs and the next batch's first one)
    const bool aos = !(o.flags & CAPS_FLAG_TILE_RADIX);
    KRec<I> *brec = aos ? (KRec<I> *)W.brec.get(bcap * sizeof(KRec<I>) + 16) : nullptr;
    I *sink_buf = sink ? (I *)W.sinkbuf.get(bcap * sizeof(I) + 8) : nullptr;
    // the previous batch's last suffix (key, position), for the next batch's first LCP
    uint64_t prev_key = 0, prev_pos = 0;
    const uint64_t tcap = ceil_div(bcap, 512) + p + 1;  // tiles per class (every tile but the last of a bucket is full)
    Tile *tiles = (Tile *)W.tiles.get(kTileClasses * tcap * sizeof(Tile));
    Range *big = (Range *)W.big.get(p * sizeof(Range));
    // [0] oversized, [1..4) tile classes, [4] largest bucket, [8..16) group counts,
    // [16..24) refinement work queues, [24..30) class records (u64),
    // [32..36) mid-size groups per LDS class, [36..40) their claim counters
    unsigned int *counters = (unsigned int *)W.counts.get(256);
    GroupLists gl{};
    for (int c = 0; c < kGroupClasses; ++c)
        gl.list[c] = (Group *)W.glist[c].get((bcap / group_class_min(c) + 1) *
                                             (c == 0 ? sizeof(PairRec<I>) : sizeof(Group)));
    const uint64_t mid_cap = bcap / group_class_min(6) + 1;
    Group *mid = (Group *)W.mid.get(kMidClasses * mid_cap * sizeof(Group));
    unsigned int *mid_counts = counters + 32;
    gl.counts = counters + 8;
    I *const lcp_all = want_lcp ? d_lcp_out : nullptr;
    // symbols every tie group's members are known to share: the key's, or none
    // for escape keys (equal escape keys may differ from the first escape on)
    const uint64_t depth0 = g.esc ? 0 : g.k;
    bool runs_built = false;
    // refinement of groups of 2..64, per size class: one slot per group (grid
    // launch), or persistent slots fed from a device work queue
    const RefinePlan plan = refine_plan(o);
    unsigned int *queues = counters + 16;
    const uint32_t waves = W.n_cu * 32;  // persistent grid: 8 waves per SIMD
    auto pgrid = [&](uint64_t items, uint32_t per_block) {
        return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(items, per_block), waves / 4));
    };
    auto chunk_for = [&](uint64_t items, uint32_t slots) {  // about four claims per wavefront
        return (uint32_t)std::max<uint64_t>(slots, std::min<uint64_t>(256, items / ((uint64_t)waves * 4)));
    };

    // scatter pass 1 for coarse intervals [c0, c1) into tmp, at offsets relative to rb
    auto pass1 = [&](uint32_t c0, uint32_t c1, uint64_t rb) {
        kt.begin(&stats.k_scatter1);
        // contiguous positions (every suffix, or a bitmap's kept ones), at
        // most 2,048 intervals: the LDS-staged pass with coalesced record
        // writes.  Not for batched builds: their keys are text-major, so a
        // step's records fall in one or two intervals and are written
        // contiguously already.
        const bool staged = seq && !idbits && pc <= kStageIntervals && !(o.flags & CAPS_FLAG_SCATTER_DIRECT);
        // Over contiguous positions a workgroup of pass 1 takes hstride
        // consecutive classify blocks (their runs in every interval are
        // consecutive): about kP1StagedPerCu (staged) or kP1DirectPerCu
        // workgroups per compute unit, so each workgroup's run per interval
        // is long (the 4,096-interval direct pass: 4.59 -> 3.96 ms).  A
        // position list keeps one classify block per workgroup: its keys are
        // scattered text reads, and fewer workgroups leave their latency
        // exposed (820M genome share: 17.5 against 29.7 ms).
        const uint32_t target = (staged ? kP1StagedPerCu : kP1DirectPerCu) * W.n_cu;
        const uint32_t hs = seq ? (uint32_t)std::max<uint64_t>(1, ceil_div(nblk, std::max<uint32_t>(target, 1))) : 1u;
        const unsigned nb1 = (unsigned)ceil_div(nblk, hs);
        const uint64_t ch1 = chunk * hs;
        if (staged) {
            if (d_keep)
                k_scatter1_staged<S, I, true><<<nb1, kStageNT, 0, st>>>(d_text, n_iter, cm, g, L, max_ctx, cid, pc,
                                                                        coff, hblk, tmp, ch1, d_keep, c0, c1, rb, hs);
            else
                k_scatter1_staged<S, I, false><<<nb1, kStageNT, 0, st>>>(d_text, n_iter, cm, g, L, max_ctx, cid, pc,
                                                                         coff, hblk, tmp, ch1, nullptr, c0, c1, rb, hs);
        } else if (seq && d_keep)
            k_scatter1<S, I, true, true><<<nb1, 256, (size_t)pc * 12, st>>>(d_text, n, cm, g, L, max_ctx, d_pos, n_iter, cid, pc,
                                                               coff, hblk, tmp, ch1, d_keep, c0, c1, rb, hs);
        else if (seq)
            k_scatter1<S, I, true, false><<<nb1, 256, (size_t)pc * 12, st>>>(d_text, n, cm, g, L, max_ctx, d_pos, n_iter, cid, pc,
                                                                coff, hblk, tmp, ch1, nullptr, c0, c1, rb, hs);
        else
            k_scatter1<S, I, false, false><<<nb1, 256, (size_t)pc * 12, st>>>(d_text, n, cm, g, L, max_ctx, d_pos, n_pos, cid, pc,
                                                                 coff, hblk, tmp, ch1, nullptr, c0, c1, rb, hs);
        LAUNCHCHK();
        kt.end();
    };
    if (all_records) pass1(0, pc, 0);

    for (size_t bi = 0; bi + 1 < cuts.size(); ++bi) {
        const uint32_t c_lo = cuts[bi], c_hi = cuts[bi + 1];
        const uint64_t rbase = hcoff[c_lo], nb = hcoff[c_hi] - rbase;
        const uint64_t b_lo = first_bucket(c_lo), b_hi = c_hi == pc ? p : first_bucket(c_hi);
        const uint32_t pb = (uint32_t)(b_hi - b_lo);  // buckets of the batch
        uint64_t *boff = off + b_lo;                  // their offsets, relative to rbase
        I *out = sink ? sink_buf : d_out + rbase;
        I *lcp = lcp_all ? lcp_all + rbase : nullptr;
        if (nb == 0) continue;

        // ---- two-level scatter ----
        if (!all_records) pass1(c_lo, c_hi, rbase);
        kt.begin(&stats.k_scatter2);
        // 1,024 threads and 4,096-record tiles when the LDS of a wide
        // interval (up to 80 KB for 4,096 buckets) or a batch of few intervals
        // calls for them; else 1,024 threads with 8,192-record tiles for a u32
        // index (112 KB + about 20 bytes per bucket: one block per compute
        // unit; twice the bucket run per write of 4,096-record tiles at two
        // blocks, 3.08 -> 2.72 ms on 268M records); a u64 index takes
        // 8,192-record tiles (144 KB) when the interval's tables fit in the
        // 12 KB left beside them, else 4,096-record tiles (72 KB)
        const uint32_t s2_lds = scatter2_lds_bytes(SP.stride);
        const bool wide = s2_lds > 22 * 1024 || (c_hi - c_lo) < 2 * W.n_cu;
        KRec<I> *trec = all_records ? tmp + rbase : tmp;
        auto scatter2 = [&](auto kern) {
            kern<<<c_hi - c_lo, 1024, s2_lds, st>>>(trec, coff, SP, (uint32_t)p, off, keys, out, c_lo, c_hi, rbase,
                                                    brec);
        };
        if (wide) {
            if (aos) scatter2(k_scatter2<I, 1024, 4096, true>);
            else scatter2(k_scatter2<I, 1024, 4096, false>);
        } else if (sizeof(I) == 4 || s2_lds <= 12 * 1024) {
            // u64 staging of 8,192 records takes 147.6 KB: room for at most
            // 12 KB of splitters and cursors beside it
            if (aos) scatter2(k_scatter2<I, 1024, 8192, true>);
            else scatter2(k_scatter2<I, 1024, 8192, false>);
        } else {
            if (aos) scatter2(k_scatter2<I, 1024, 4096, true>);
            else scatter2(k_scatter2<I, 1024, 4096, false>);
        }
        LAUNCHCHK();
        kt.end();
        tm.mark(PH_SCATTER);

        // ---- tiles: sort per tile (by size class); whole-bucket tiles emit their tie groups ----
        HIPCHK(hipMemsetAsync(counters, 0, 256, st));
        k_tile_emit<<<(unsigned)ceil_div(pb, 256 * kEmitPer), 256, 0, st>>>(boff, pb, T, tiles, tcap, counters + 1, big,
                                                                  counters, counters + 4,
                                                                  (unsigned long long *)(counters + 24));
        LAUNCHCHK();
        unsigned int hc[32];
        HIPCHK(hipMemcpyAsync(hc, counters, 128, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        stats.bucket_max = std::max<uint64_t>(stats.bucket_max, hc[4]);
        const unsigned int n_big = hc[0];
        stats.tiles += (uint64_t)hc[1] + hc[2] + hc[3];
        stats.oversized += n_big;
        for (int c = 0; c < 3; ++c) {
            unsigned long long r;
            memcpy(&r, hc + 24 + 2 * c, 8);
            stats.tile_records[c] += r;
        }
        // one launch per tile class, each timed on its own (graph builds:
        // the three launches are one captured graph, timed as a whole)
        const bool radix = (o.flags & CAPS_FLAG_TILE_RADIX) != 0;
        {
            GraphSection gs(W, st, graph && (hc[1] || hc[2] || hc[3]), 0);
            auto tile_sort = [&](auto ttag, int c) {
                constexpr uint32_t TT = decltype(ttag)::value;
                if (!hc[1 + c]) return;
                if (!gs.capturing) kt.begin(&stats.k_tile[c]);
                const int keep = lcp ? 1 : 0;  // boundary LCPs read the sorted keys
                if (radix) k_tile_radix<I, TT><<<hc[1 + c], 256, 0, gs.ls>>>(keys, out, tiles + c * tcap, gl, 1, keep);
                else if (aos)
                    k_tile_merge<I, TT, true><<<hc[1 + c], 256, 0, gs.ls>>>(keys, out, tiles + c * tcap, gl, 1, keep,
                                                                           brec);
                else k_tile_merge<I, TT><<<hc[1 + c], 256, 0, gs.ls>>>(keys, out, tiles + c * tcap, gl, 1, keep);
                LAUNCHCHK();
                if (!gs.capturing) kt.end();
            };
            tile_sort(std::integral_constant<uint32_t, 1024>{}, 0);
            tile_sort(std::integral_constant<uint32_t, 2048>{}, 1);
            tile_sort(std::integral_constant<uint32_t, 4096>{}, 2);
            gs.finish(kt, &stats.k_graph[0]);
        }
        tm.mark(PH_TILE);

        // ---- oversized buckets: merge the sorted tiles by key, then find their groups ----
        if (n_big) {
            std::vector<Range> hb(n_big);
            HIPCHK(hipMemcpyAsync(hb.data(), big, n_big * sizeof(Range), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            uint64_t maxlen = 0;
            for (auto &r : hb) maxlen = std::max(maxlen, r.len);
            uint64_t *k2 = (uint64_t *)W.keys2.get(bcap * 8);
            I *i2 = (I *)W.idx2.get(bcap * sizeof(I));
            uint64_t *ks = keys, *kd = k2;
            I *is = out, *id = i2;
            int levels = 0;
            static_assert(T >= kKeyMergeTile / 2 && (T & (T - 1)) == 0,
                          "key merge runs: powers of two >= half a tile");
            for (uint64_t run = T; run < maxlen; run *= 2, ++levels) {
                for (unsigned y0 = 0; y0 < n_big; y0 += 65535) {
                    dim3 gr((unsigned)ceil_div(maxlen, kKeyMergeTile), std::min(65535u, n_big - y0));
                    k_key_merge<I><<<gr, 256, 0, st>>>(big + y0, run, ks, is, kd, id);
                    LAUNCHCHK();
                }
                std::swap(ks, kd);
                std::swap(is, id);
            }
            if (levels & 1) {
                for (unsigned y0 = 0; y0 < n_big; y0 += 65535) {
                    dim3 gr((unsigned)std::min<uint64_t>(ceil_div(maxlen, 256), 4096), std::min(65535u, n_big - y0));
                    k_copy_ranges<I><<<gr, 256, 0, st>>>(big + y0, k2, i2, keys, out);
                    LAUNCHCHK();
                }
            }
            for (unsigned y0 = 0; y0 < n_big; y0 += 65535) {
                dim3 gr((unsigned)std::min<uint64_t>(ceil_div(maxlen, 256), 1024), std::min(65535u, n_big - y0));
                k_detect_ranges<Range, I><<<gr, 256, 0, st>>>(keys, big + y0, gl, out);
                LAUNCHCHK();
            }
        }
        tm.mark(PH_MERGE);

        unsigned int cnt[kGroupClasses];
        HIPCHK(hipMemcpyAsync(cnt, gl.counts, sizeof(cnt), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        stats.pairs += cnt[0];
        stats.small_groups += cnt[1] + cnt[2] + cnt[3] + cnt[4] + cnt[5];
        stats.mid_groups += cnt[6];
        stats.giant_groups += cnt[7];
        tm.mark(PH_DETECT);

        // the run memo, once per build, for the LCP scans of larger groups
        if ((cnt[6] || cnt[7]) && !runs_built && !(o.flags & CAPS_FLAG_NO_RUN_MEMO)) {
            cm.runs = build_run_memo<S>(W, cm, n, st);
            runs_built = true;
            stats.runs = cm.runs.n_runs;
        }

        // ---- refinement ----
        {
            bool any = false;
            for (int c = 0; c < 7; ++c) any |= cnt[c] != 0;
            GraphSection gs(W, st, graph && any, 1);
            const hipStream_t ls = gs.ls;
            const bool timed = !gs.capturing;
            if (cnt[0]) {
                if (timed) kt.begin(&stats.k_pairs);
                const PairRec<I> *pl = (const PairRec<I> *)gl.list[0];
                if (plan.queue[0])
                    k_pairs_wq<S, I><<<pgrid(cnt[0], 64), 256, 0, ls>>>(d_text, L, max_ctx, pl, cnt[0], depth0, out,
                                                                         lcp, queues + 0, chunk_for(cnt[0], 16));
                else if (cm.packed)
                    k_pairs_packed<S, I><<<(unsigned)ceil_div(cnt[0], 64), 256, 0, ls>>>(
                        cm.packed, cm.lgf, gref.bits, L, max_ctx, pl, cnt[0], depth0, out, lcp);
                else
                    k_pairs4<S, I><<<(unsigned)ceil_div(cnt[0], 64), 256, 0, ls>>>(d_text, L, max_ctx, pl, cnt[0],
                                                                                    depth0, out, lcp);
                LAUNCHCHK();
                if (timed) kt.end();
            }
            if (timed) kt.begin(&stats.k_groups_small);
            auto wgroups = [&](auto wtag, int c) {
                constexpr int Wd = decltype(wtag)::value;
                if (!cnt[c]) return;
                constexpr uint32_t per_block = 256 / Wd;
                if (plan.queue[c])
                    k_wgroups_wq<S, I, Wd><<<pgrid(cnt[c], per_block), 256, 0, ls>>>(
                        d_text, cm, gref, L, max_ctx, gl.list[c], cnt[c], depth0, out, lcp, queues + c,
                        chunk_for(cnt[c], 64 / Wd));
                else
                    k_wgroups_grid<S, I, Wd><<<(unsigned)ceil_div(cnt[c], per_block), 256, 0, ls>>>(
                        d_text, cm, gref, L, max_ctx, gl.list[c], cnt[c], depth0, out, lcp);
            };
            wgroups(std::integral_constant<int, 4>{}, 1);
            wgroups(std::integral_constant<int, 8>{}, 2);
            wgroups(std::integral_constant<int, 16>{}, 3);
            wgroups(std::integral_constant<int, 32>{}, 4);
            wgroups(std::integral_constant<int, 64>{}, 5);
            LAUNCHCHK();
            if (timed) kt.end();
            if (cnt[6]) {
                if (timed) kt.begin(&stats.k_groups_mid);
                if (cnt[6] >= 16u * W.n_cu || (o.flags & CAPS_FLAG_MID_BY_SIZE)) {
                    // many groups: by LDS class, so groups of up to 512
                    // members keep 4 to 14 workgroups per compute unit; each
                    // launch claims the groups its split list holds, with a
                    // grid of at most one resident wave of it
                    k_split_mid<<<(unsigned)ceil_div(cnt[6], 256), 256, 0, ls>>>(gl.list[6], cnt[6], mid, mid_cap,
                                                                                 mid_counts);
                    auto mid_class = [&](auto kern, unsigned block, uint32_t c) {
                        const unsigned grid = (unsigned)std::min<uint64_t>(cnt[6], resident_blocks(kern, block));
                        kern<<<grid, block, 0, ls>>>(d_text, cm, gref, L, max_ctx, mid + c * mid_cap, mid_counts + c,
                                                     0, mid_counts + kMidClasses + c, depth0, out, lcp, 0);
                    };
                    mid_class(k_groups<S, I, 128, 128>, 128, 0);
                    mid_class(k_groups<S, I, 256, 512>, 256, 1);
                    mid_class(k_groups<S, I, 256, 1024>, 256, 2);
                    mid_class(k_groups<S, I, 256, kMidMax>, 256, 3);
                } else {
                    // few groups (often long ones): one launch, one workgroup
                    // per group, so no launch waits for another's last groups
                    k_groups<S, I, 256, kMidMax><<<cnt[6], 256, 0, ls>>>(d_text, cm, gref, L, max_ctx, gl.list[6],
                                                                          nullptr, cnt[6], nullptr, depth0, out, lcp,
                                                                          0);
                }
                LAUNCHCHK();
                if (timed) kt.end();
            }
            gs.finish(kt, &stats.k_graph[1]);
        }
        tm.mark(PH_REFINE);

        // ---- giant groups: chunked refinement + merge-path LCP merges ----
        // All giant groups at once: one launch sorts every chunk of every
        // group, then one launch per merge level merges the runs of every group
        // (a group already in one run is carried over by a copy job).  The job
        // lists of all levels are built and uploaded together, so no level
        // waits for the host.
        if (cnt[7]) {
            std::vector<Group> hg(cnt[7]);
            HIPCHK(hipMemcpyAsync(hg.data(), gl.list[7], cnt[7] * sizeof(Group), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            I *lbuf = lcp ? lcp : (I *)W.lcp.get(bcap * sizeof(I));
            I *i2 = (I *)W.idx2.get(bcap * sizeof(I));
            I *l2 = (I *)W.lcp2.get(bcap * sizeof(I));
            std::vector<Group> hch;
            uint64_t maxsize = 0;
            for (const Group &gr : hg) {
                maxsize = std::max<uint64_t>(maxsize, gr.size);
                for (uint64_t c = 0; c < gr.size; c += kGiantChunk)
                    hch.push_back(Group{gr.start + c, std::min<uint64_t>(kGiantChunk, gr.size - c)});
            }
            int levels = 0;
            for (uint64_t run = kGiantChunk; run < maxsize; run *= 2) ++levels;
            std::vector<uint64_t> level_off(levels + 1, 0);
            W.h_jobs.clear();
            for (int lv = 0; lv < levels; ++lv) {
                const uint64_t run = (uint64_t)kGiantChunk << lv;
                level_off[lv] = W.h_jobs.size();
                for (const Group &gr : hg)
                    for (uint64_t b = 0; b < gr.size; b += 2 * run) {
                        MergeJob jb;
                        jb.start = gr.start + b;
                        jb.nx = std::min<uint64_t>(run, gr.size - b);
                        jb.ny = std::min<uint64_t>(run, gr.size - b - jb.nx);
                        W.h_jobs.push_back(jb);
                    }
            }
            level_off[levels] = W.h_jobs.size();
            Group *dc = (Group *)W.chunks.get(hch.size() * sizeof(Group));
            HIPCHK(hipMemcpyAsync(dc, hch.data(), hch.size() * sizeof(Group), hipMemcpyHostToDevice, st));
            MergeJob *dj = (MergeJob *)W.jobs.get(std::max<size_t>(W.h_jobs.size(), 1) * sizeof(MergeJob));
            if (!W.h_jobs.empty())
                HIPCHK(hipMemcpyAsync(dj, W.h_jobs.data(), W.h_jobs.size() * sizeof(MergeJob), hipMemcpyHostToDevice,
                                      st));
            GraphSection gsg(W, st, graph, 2);
            const hipStream_t gst = gsg.ls;
            for (uint64_t c0 = 0; c0 < hch.size(); c0 += 65535u * 16) {
                const unsigned nbk = (unsigned)std::min<uint64_t>(hch.size() - c0, 65535u * 16);
                k_groups<S, I, 256, kGiantChunk><<<nbk, 256, 0, gst>>>(d_text, cm, gref, L, max_ctx, dc + c0, nullptr, nbk,
                                                                  nullptr, depth0, out, lbuf, 1);
                LAUNCHCHK();
            }
            I *is = out, *ls = lbuf, *id = i2, *ld = l2;
            // merge-path partition width: outputs per thread (a tuning knob).
            // Unset, groups of 128K+ members take 64: each partition search
            // compares whole suffixes, and members of huge groups share long
            // stretches, so fewer, longer partitions pay there.
            const int mw = (o.flags & CAPS_FLAG_MERGE_W16)   ? 16
                           : (o.flags & CAPS_FLAG_MERGE_W64) ? 64
                           : maxsize >= (1ull << 17)         ? 64
                                                             : 32;
            auto merge_level = [&](auto wtag, dim3 grid, const MergeJob *jobs) {
                constexpr int MW = decltype(wtag)::value;
                k_lcp_merge<S, I, MW><<<grid, 128, 0, gst>>>(d_text, cm, L, max_ctx, depth0, jobs, is, ls, id, ld);
            };
            for (int lv = 0; lv < levels; ++lv) {
                const uint64_t run = (uint64_t)kGiantChunk << lv;
                const uint64_t span = std::min<uint64_t>(2 * run, maxsize);
                for (uint64_t j0 = level_off[lv]; j0 < level_off[lv + 1]; j0 += 65535) {
                    const unsigned nj = (unsigned)std::min<uint64_t>(level_off[lv + 1] - j0, 65535);
                    dim3 gr2((unsigned)ceil_div(ceil_div(span, mw), 128), nj);
                    if (mw == 16) merge_level(std::integral_constant<int, 16>{}, gr2, dj + j0);
                    else if (mw == 64) merge_level(std::integral_constant<int, 64>{}, gr2, dj + j0);
                    else merge_level(std::integral_constant<int, 32>{}, gr2, dj + j0);
                    LAUNCHCHK();
                }
                std::swap(is, id);
                std::swap(ls, ld);
            }
            if (levels & 1) {
                for (uint64_t c0 = 0; c0 < hch.size(); c0 += 65535u * 16) {
                    const unsigned nbk = (unsigned)std::min<uint64_t>(hch.size() - c0, 65535u * 16);
                    k_copy_chunks<I><<<nbk, 256, 0, gst>>>(dc + c0, i2, l2, out, lbuf);
                    LAUNCHCHK();
                }
            }
            gsg.finish(kt, &stats.k_graph[2]);
        }
        if (lcp) {
            kt.begin(&stats.k_boundary_lcp);
            k_boundary_lcp<I><<<grid_for(nb, 256, 16384), 256, 0, st>>>(keys, out, nb, L, max_ctx, g, lcp,
                                                                         bi ? 1u : 0u, prev_key, prev_pos,
                                                                         cm.packed, cm.lgf);
            LAUNCHCHK();
            kt.end();
            if (!single) {
                I last = 0;
                HIPCHK(hipMemcpyAsync(&prev_key, keys + nb - 1, 8, hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(&last, out + nb - 1, sizeof(I), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                prev_pos = (uint64_t)last;
            }
        }
        tm.mark(PH_GIANT);
        if (sink) sink->flush(out, nb, st);
    }
    if (one_batch) HIPCHK(hipMemcpyAsync(hcoff.data(), coff, (pc + 1) * 8, hipMemcpyDeviceToHost, st));
    if (R.stats) {
        unsigned long long rs[4];
        HIPCHK(hipMemcpyAsync(rs, R.stats, 32, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        stats.router_err_sum = rs[0], stats.router_err_max = rs[1], stats.router_fallbacks = rs[2];
        stats.router_count = rs[3];
        if (R.kind == CAPS_ROUTER_SPLINE) stats.router_bound_lo = -R.el, stats.router_bound_hi = R.eh;
    }
    HIPCHK(hipStreamSynchronize(st));
    stats.coarse_max = interval_max();
    tm.collect(stats);
    kt.collect();
    stats.k_scatter = stats.k_scatter1 + stats.k_scatter2;
    stats.k_tile_sort = stats.k_tile[0] + stats.k_tile[1] + stats.k_tile[2];
    if (counting) {
        std::vector<unsigned long long> cs(3 * kCountLanes);
        HIPCHK(hipMemcpyAsync(cs.data(), count_addr.sym, cs.size() * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int f = 0; f < 3; ++f)
            for (int l = 0; l < kCountLanes; ++l) stats.cmp_symbols[f] += cs[f * kCountLanes + l];
    }
}

// ---------------------------------------------------------------------------
// Small device helpers used by the C ABI.
// ---------------------------------------------------------------------------
struct HostText {
    const void *d;
    uint64_t n;
    uint32_t w;
};

static const void *upload_text(Workspace &W, const void *text, uint64_t n, uint32_t w, hipStream_t st) {
    void *d = W.text.get(n * w + 16);
    if (n) HIPCHK(hipMemcpyAsync(d, text, n * w, hipMemcpyHostToDevice, st));
    return d;
}

template <typename S, typename I>
static void build_host(const void *text, uint64_t n, const void *positions, uint64_t n_pos,
                       const caps_limits_t *lim, const caps_opts_t &o, void *out_sa, void *out_lcp) {
    Workspace &W = workspace();
    hipStream_t st = nullptr;
    const S *dt = (const S *)upload_text(W, text, n, sizeof(S), st);
    const I *dp = nullptr;
    if (positions) {
        I *d = (I *)W.hpos.get(n_pos * sizeof(I) + 8);
        if (n_pos) HIPCHK(hipMemcpyAsync(d, positions, n_pos * sizeof(I), hipMemcpyHostToDevice, st));
        dp = d;
    }
    I *dout = (I *)W.out.get(n_pos * sizeof(I) + 8);
    I *dl = out_lcp ? (I *)W.outlcp.get(n_pos * sizeof(I) + 8) : nullptr;
    build_impl<S, I>(dt, n, dp, n_pos, lim, o, dout, dl, st, g_stats);
    if (n_pos) {
        HIPCHK(hipMemcpyAsync(out_sa, dout, n_pos * sizeof(I), hipMemcpyDeviceToHost, st));
        if (out_lcp) HIPCHK(hipMemcpyAsync(out_lcp, dl, n_pos * sizeof(I), hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
}

template <typename S, typename I>
static void dispatch_positions(const void *text, uint64_t n, const void *positions, uint64_t n_pos,
                               const caps_limits_t *lim, const caps_opts_t &o, void *out_sa, void *out_lcp) {
    build_host<S, I>(text, n, positions, n_pos, lim, o, out_sa, out_lcp);
}

static void check_common(const void *text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *lim) {
    if (!text && n) throw CapsError(CAPS_E_INVALID, "text is NULL");
    if (sym_bytes != 1 && sym_bytes != 2)
        throw CapsError(CAPS_E_UNSUPPORTED, "the GPU build covers 1- and 2-byte symbols only");
    if (lim && lim->kind == CAPS_LIMITS_SEGMENTED && lim->n_ends && !lim->ends)
        throw CapsError(CAPS_E_INVALID, "segment ends pointer is NULL");
}

template <typename F>
static int guarded(F &&f) {
    try {
        std::lock_guard<std::recursive_mutex> lk(g_mu);
        f();
        return CAPS_OK;
    } catch (const CapsError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return CAPS_E_NOMEM;
    } catch (const std::exception &e) {
        g_err = e.what();
        return CAPS_E_DEVICE;
    }
}

static caps_opts_t opts_or_default(const caps_opts_t *o) {
    caps_opts_t d;
    caps_opts_default(&d);
    return o ? *o : d;
}

template <typename S, typename I>
static void run_device(const void *dt, uint64_t n, const void *dp, uint64_t n_pos, const caps_limits_t *lim,
                       const caps_opts_t &o, void *dout, void *dl, hipStream_t st) {
    build_impl<S, I>((const S *)dt, n, (const I *)dp, n_pos, lim, o, (I *)dout, (I *)dl, st, g_stats);
}

}  // namespace caps

using namespace caps;

extern "C" {

void caps_opts_default(caps_opts_t *o) {
    memset(o, 0, sizeof(*o));
    o->max_context = ~0ull;
}

const char *caps_last_error(void) { return g_err.c_str(); }

int caps_last_emit_code(void) { return g_emit_code; }

int caps_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int caps_last_stats(caps_stats_t *out) {
    if (!out) return CAPS_E_INVALID;
    *out = g_stats;
    return CAPS_OK;
}

int caps_build_in_memory_for_positions(const void *text, uint64_t n, uint32_t sym_bytes, const void *positions,
                                       uint64_t n_pos, uint32_t idx_bytes, const caps_limits_t *limits,
                                       const caps_opts_t *opts, void *out_sa, void *out_lcp) {
    return guarded([&] {
        check_common(text, n, sym_bytes, limits);
        if (!out_sa && n_pos) throw CapsError(CAPS_E_INVALID, "out_sa is NULL");
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (idx_bytes == 4 && n > (1ull << 32)) throw CapsError(CAPS_E_INVALID, "u32 index cannot address text");
        if (!positions && n_pos != n) throw CapsError(CAPS_E_INVALID, "positions is NULL but n_pos != n");
        if (positions)
            for (uint64_t i = 0; i < n_pos; ++i) {
                const uint64_t q = idx_bytes == 4 ? ((const uint32_t *)positions)[i] : ((const uint64_t *)positions)[i];
                if (q >= n) throw CapsError(CAPS_E_INVALID, "position out of range");
            }
        caps_opts_t o = opts_or_default(opts);
        if (sym_bytes == 1 && idx_bytes == 4) build_host<uint8_t, uint32_t>(text, n, positions, n_pos, limits, o, out_sa, out_lcp);
        else if (sym_bytes == 1) build_host<uint8_t, uint64_t>(text, n, positions, n_pos, limits, o, out_sa, out_lcp);
        else if (idx_bytes == 4) build_host<uint16_t, uint32_t>(text, n, positions, n_pos, limits, o, out_sa, out_lcp);
        else build_host<uint16_t, uint64_t>(text, n, positions, n_pos, limits, o, out_sa, out_lcp);
    });
}

int caps_build_in_memory(const void *text, uint64_t n, uint32_t sym_bytes, uint32_t idx_bytes,
                         const caps_limits_t *limits, const caps_opts_t *opts, void *out_sa, void *out_lcp) {
    return caps_build_in_memory_for_positions(text, n, sym_bytes, nullptr, n, idx_bytes, limits, opts, out_sa,
                                              out_lcp);
}

}  // extern "C"

namespace caps {

// Uploads a host bitmap of kept positions ((n + 63) / 64 words), clears the
// bits past n and returns the device copy and the kept count.
static const uint64_t *upload_keep(Workspace &W, const uint64_t *keep_words, uint64_t n, uint64_t *count,
                                   hipStream_t st) {
    const uint64_t nw = ceil_div(n, 64);
    uint64_t *d = (uint64_t *)W.keep.get(nw * 8 + 8);
    if (nw) HIPCHK(hipMemcpyAsync(d, keep_words, nw * 8, hipMemcpyHostToDevice, st));
    *count = device_keep_count(W, d, n, st);
    return d;
}

// A streaming build: batches of the suffix array go from HBM through a
// pinned staging buffer to `emit`, in lexicographic order, on the caller's
// thread (ext_mem.rs:1933-1937); the first nonzero emit aborts the build.
template <typename S, typename I>
static void stream_typed(const void *text, uint64_t n, const uint64_t *positions, uint64_t n_pos,
                         const uint64_t *keep_words, const caps_limits_t *limits, const caps_opts_t &o,
                         caps_emit_fn emit, void *ctx) {
    Workspace &W = workspace();
    hipStream_t st = nullptr;
    const S *dt = (const S *)upload_text(W, text, n, sizeof(S), st);
    const I *dp = nullptr;
    const uint64_t *dk = nullptr;
    if (keep_words) {
        dk = upload_keep(W, keep_words, n, &n_pos, st);
    } else if (positions) {
        I *d = (I *)W.hpos.get(n_pos * sizeof(I) + 8);
        if (sizeof(I) == 8) {
            if (n_pos) HIPCHK(hipMemcpyAsync(d, positions, n_pos * 8, hipMemcpyHostToDevice, st));
        } else {
            std::vector<I> narrow(n_pos);
            for (uint64_t i = 0; i < n_pos; ++i) narrow[i] = (I)positions[i];
            if (n_pos) HIPCHK(hipMemcpyAsync(d, narrow.data(), n_pos * sizeof(I), hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        dp = d;
    }
    constexpr uint64_t kStage = 1ull << 22;  // positions per emit call
    if (W.h_stage_bytes < kStage * 16) {
        if (W.h_stage) (void)hipHostFree(W.h_stage);
        W.h_stage = nullptr;
        W.h_stage_bytes = 0;
        HIPCHK(hipHostMalloc(&W.h_stage, kStage * 16, hipHostMallocDefault));
        W.h_stage_bytes = kStage * 16;
    }
    I *stage = (I *)W.h_stage;
    uint64_t *e64 = (uint64_t *)((char *)W.h_stage + kStage * 8);
    BatchSink<I> sink;
    sink.flush = [&](const I *d, uint64_t cnt, hipStream_t s) {
        for (uint64_t off = 0; off < cnt; off += kStage) {
            const uint64_t m = std::min(kStage, cnt - off);
            HIPCHK(hipMemcpyAsync(stage, d + off, m * sizeof(I), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            const uint64_t *src = (const uint64_t *)stage;
            if (sizeof(I) != 8) {
                for (uint64_t i = 0; i < m; ++i) e64[i] = (uint64_t)stage[i];
                src = e64;
            }
            const int r = emit(ctx, src, m);
            if (r != 0) {
                g_emit_code = r;
                throw CapsError(CAPS_E_EMIT, "emit callback aborted the build");
            }
        }
    };
    build_impl<S, I>(dt, n, dp, n_pos, limits, o, nullptr, nullptr, st, g_stats, 0, dk, &sink);
}

static int stream_build(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *positions, uint64_t n_pos,
                        const uint64_t *keep_words, const caps_limits_t *limits, const caps_opts_t *opts,
                        caps_emit_fn emit, void *ctx) {
    g_emit_code = 0;
    return guarded([&] {
        check_common(text, n, sym_bytes, limits);
        if (!emit) throw CapsError(CAPS_E_INVALID, "emit callback is NULL");
        if (keep_words == nullptr && n && !text) throw CapsError(CAPS_E_INVALID, "text is NULL");
        caps_opts_t o = opts_or_default(opts);
        if (positions)
            for (uint64_t i = 0; i < n_pos; ++i)
                if (positions[i] >= n) throw CapsError(CAPS_E_INVALID, "position out of range");
        // index width follows the reference: u32 iff text.len() <= u32::MAX + 1
        const bool narrow = n <= (1ull << 32);
        if (narrow && sym_bytes == 1)
            stream_typed<uint8_t, uint32_t>(text, n, positions, n_pos, keep_words, limits, o, emit, ctx);
        else if (narrow)
            stream_typed<uint16_t, uint32_t>(text, n, positions, n_pos, keep_words, limits, o, emit, ctx);
        else if (sym_bytes == 1)
            stream_typed<uint8_t, uint64_t>(text, n, positions, n_pos, keep_words, limits, o, emit, ctx);
        else
            stream_typed<uint16_t, uint64_t>(text, n, positions, n_pos, keep_words, limits, o, emit, ctx);
    });
}

}  // namespace caps

extern "C" {

int caps_build_in_memory_sample_sort(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *positions,
                                     uint64_t n_pos, const caps_limits_t *limits, const caps_opts_t *opts,
                                     caps_emit_fn emit, void *ctx) {
    return stream_build(text, n, sym_bytes, positions, positions ? n_pos : n, nullptr, limits, opts, emit, ctx);
}

int caps_build_ext_mem_for_positions(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *positions,
                                     uint64_t n_pos, const caps_limits_t *limits, const caps_opts_t *opts,
                                     caps_emit_fn emit, void *ctx) {
    if (!positions && n_pos) {
        g_err = "positions is NULL";
        return CAPS_E_INVALID;
    }
    return stream_build(text, n, sym_bytes, positions, n_pos, nullptr, limits, opts, emit, ctx);
}

int caps_build_ext_mem_for_filter(const void *text, uint64_t n, uint32_t sym_bytes, const uint64_t *keep_words,
                                  const caps_limits_t *limits, const caps_opts_t *opts, caps_emit_fn emit,
                                  void *ctx) {
    if (!keep_words && n) {
        g_err = "keep bitmap is NULL";
        return CAPS_E_INVALID;
    }
    return stream_build(text, n, sym_bytes, nullptr, 0, keep_words, limits, opts, emit, ctx);
}

int caps_build_device_filtered(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *d_keep_words,
                               uint32_t idx_bytes, const caps_limits_t *limits, const caps_opts_t *opts,
                               void *d_out_sa, uint64_t out_capacity, uint64_t *n_out, void *d_out_lcp,
                               void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, limits);
        if (!d_keep_words && n) throw CapsError(CAPS_E_INVALID, "keep bitmap is NULL");
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (idx_bytes == 4 && n > (1ull << 32)) throw CapsError(CAPS_E_INVALID, "u32 index cannot address text");
        caps_opts_t o = opts_or_default(opts);
        hipStream_t st = (hipStream_t)stream;
        Workspace &W = workspace();
        // the caller's bitmap is left as it is: bits past n are cleared in a copy
        const uint64_t nw = ceil_div(n, 64);
        uint64_t *dk = (uint64_t *)W.keep.get(nw * 8 + 8);
        if (nw) HIPCHK(hipMemcpyAsync(dk, d_keep_words, nw * 8, hipMemcpyDeviceToDevice, st));
        const uint64_t kept = device_keep_count(W, dk, n, st);
        if (n_out) *n_out = kept;
        if (kept > out_capacity) throw CapsError(CAPS_E_INVALID, "output capacity below the kept count");
        if (kept && !d_out_sa) throw CapsError(CAPS_E_INVALID, "d_out_sa is NULL");
        if (sym_bytes == 1 && idx_bytes == 4)
            build_impl<uint8_t, uint32_t>((const uint8_t *)d_text, n, nullptr, kept, limits, o, (uint32_t *)d_out_sa,
                                          (uint32_t *)d_out_lcp, st, g_stats, 0, dk);
        else if (sym_bytes == 1)
            build_impl<uint8_t, uint64_t>((const uint8_t *)d_text, n, nullptr, kept, limits, o, (uint64_t *)d_out_sa,
                                          (uint64_t *)d_out_lcp, st, g_stats, 0, dk);
        else if (idx_bytes == 4)
            build_impl<uint16_t, uint32_t>((const uint16_t *)d_text, n, nullptr, kept, limits, o,
                                           (uint32_t *)d_out_sa, (uint32_t *)d_out_lcp, st, g_stats, 0, dk);
        else
            build_impl<uint16_t, uint64_t>((const uint16_t *)d_text, n, nullptr, kept, limits, o,
                                           (uint64_t *)d_out_sa, (uint64_t *)d_out_lcp, st, g_stats, 0, dk);
    });
}

int caps_build_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const void *d_positions, uint64_t n_pos,
                      uint32_t idx_bytes, const caps_limits_t *limits, const caps_opts_t *opts, void *d_out_sa,
                      void *d_out_lcp, void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, limits);
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (idx_bytes == 4 && n > (1ull << 32)) throw CapsError(CAPS_E_INVALID, "u32 index cannot address text");
        if (!d_positions) n_pos = n;
        caps_opts_t o = opts_or_default(opts);
        hipStream_t st = (hipStream_t)stream;
        if (sym_bytes == 1 && idx_bytes == 4) run_device<uint8_t, uint32_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else if (sym_bytes == 1) run_device<uint8_t, uint64_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else if (idx_bytes == 4) run_device<uint16_t, uint32_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else run_device<uint16_t, uint64_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
    });
}

// Device address of the no-alloc scratch array on the current device, looked
// up once per device (each GPU holds its own copy of the array).
static unsigned char *noalloc_scratch() {
    static unsigned char *addr[64] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw CapsError(CAPS_E_UNSUPPORTED, "device ordinal above 63");
    if (!addr[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw CapsError(CAPS_E_DEVICE, "no device address for the no-alloc scratch");
        addr[dev] = (unsigned char *)ptr;
    }
    return addr[dev];
}

int caps_build_device_noalloc(const void *d_text, uint64_t n, uint32_t sym_bytes, const void *d_positions,
                              uint64_t n_pos, uint32_t idx_bytes, const caps_limits_t *limits,
                              const caps_opts_t *opts, void *d_out_sa, void *d_out_lcp, void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, limits);
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (!d_positions) n_pos = n;
        if (n > CAPS_NOALLOC_MAX_SYMBOLS || n_pos > CAPS_NOALLOC_MAX_SYMBOLS)
            throw CapsError(CAPS_E_INVALID, "no-alloc builds take at most 2^20 symbols and positions");
        if (limits && limits->kind == CAPS_LIMITS_SEGMENTED && limits->n_ends > CAPS_NOALLOC_MAX_ENDS)
            throw CapsError(CAPS_E_INVALID, "no-alloc builds take at most 2^16 segment ends");
        caps_opts_t o = opts_or_default(opts);
        o.batch_records = 0;
        o.flags &= ~CAPS_FLAG_GRAPH;  // graph instantiation may allocate device memory
        Arena a;
        a.base = noalloc_scratch();
        a.cap = kNoAllocScratchBytes;
        struct Off {
            ~Off() { g_arena = nullptr; }
        } off;
        g_arena = &a;
        hipStream_t st = (hipStream_t)stream;
        if (sym_bytes == 1 && idx_bytes == 4) run_device<uint8_t, uint32_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else if (sym_bytes == 1) run_device<uint8_t, uint64_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else if (idx_bytes == 4) run_device<uint16_t, uint32_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        else run_device<uint16_t, uint64_t>(d_text, n, d_positions, n_pos, limits, o, d_out_sa, d_out_lcp, st);
        g_stats.scratch_bytes = a.used;
    });
}

static uint32_t id_bits_for(uint64_t n_texts) {
    uint32_t b = 1;
    while ((1ull << b) < n_texts) ++b;
    return b;
}

// Largest number of texts one batched build can index next to at least one
// key field of `sym_bits` bits... resolved per build from the alphabet; the
// host API keeps batches within 2^kBatchIdBitsMax texts.
static constexpr uint32_t kBatchIdBitsMax = 16;

int caps_build_batch_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *ends,
                            uint64_t n_texts, uint32_t idx_bytes, const caps_opts_t *opts, void *d_out, void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, nullptr);
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (idx_bytes == 4 && n > (1ull << 32)) throw CapsError(CAPS_E_INVALID, "u32 index cannot address the batch");
        if (n_texts == 0 || !ends) throw CapsError(CAPS_E_INVALID, "empty batch");
        if (id_bits_for(n_texts) > kBatchIdBitsMax) throw CapsError(CAPS_E_INVALID, "too many texts in one batch");
        caps_limits_t lim{CAPS_LIMITS_SEGMENTED, CAPS_ORDER_SHORTER_FIRST, ends, n_texts};
        caps_opts_t o = opts_or_default(opts);
        o.max_context = ~0ull;
        hipStream_t st = (hipStream_t)stream;
        uint32_t ib = id_bits_for(n_texts);
        if (sym_bytes == 1 && idx_bytes == 4)
            build_impl<uint8_t, uint32_t>((const uint8_t *)d_text, n, nullptr, n, &lim, o, (uint32_t *)d_out, nullptr, st, g_stats, ib);
        else if (sym_bytes == 1)
            build_impl<uint8_t, uint64_t>((const uint8_t *)d_text, n, nullptr, n, &lim, o, (uint64_t *)d_out, nullptr, st, g_stats, ib);
        else if (idx_bytes == 4)
            build_impl<uint16_t, uint32_t>((const uint16_t *)d_text, n, nullptr, n, &lim, o, (uint32_t *)d_out, nullptr, st, g_stats, ib);
        else
            build_impl<uint16_t, uint64_t>((const uint16_t *)d_text, n, nullptr, n, &lim, o, (uint64_t *)d_out, nullptr, st, g_stats, ib);
    });
}

int caps_build_batch(const caps_text_item_t *items, uint64_t count, uint32_t sym_bytes, uint32_t idx_bytes,
                     const caps_opts_t *opts) {
    return guarded([&] {
        if (sym_bytes != 1 && sym_bytes != 2)
            throw CapsError(CAPS_E_UNSUPPORTED, "the GPU build covers 1- and 2-byte symbols only");
        if (idx_bytes != 4 && idx_bytes != 8) throw CapsError(CAPS_E_UNSUPPORTED, "index width must be 4 or 8");
        if (count && !items) throw CapsError(CAPS_E_INVALID, "items is NULL");
        caps_opts_t o = opts_or_default(opts);
        o.max_context = ~0ull;
        Workspace &W = workspace();
        hipStream_t st = nullptr;
        const uint64_t max_texts = 1ull << kBatchIdBitsMax;
        uint64_t i = 0;
        std::vector<uint64_t> ends;
        std::vector<uint64_t> which;
        std::vector<uint8_t> host;
        while (i < count) {
            // one sub-batch: up to max_texts non-empty texts
            ends.clear();
            which.clear();
            uint64_t total = 0;
            for (; i < count && which.size() < max_texts; ++i) {
                if (items[i].n == 0) continue;
                if (!items[i].text || !items[i].out_sa) throw CapsError(CAPS_E_INVALID, "batch item has NULL pointers");
                if (idx_bytes == 4 && items[i].n > (1ull << 32))
                    throw CapsError(CAPS_E_INVALID, "u32 index cannot address a text");
                total += items[i].n;
                ends.push_back(total);
                which.push_back(i);
            }
            if (which.empty()) break;
            host.resize(total * sym_bytes);
            for (size_t t = 0; t < which.size(); ++t) {
                uint64_t start = t ? ends[t - 1] : 0;
                memcpy(host.data() + start * sym_bytes, items[which[t]].text, items[which[t]].n * sym_bytes);
            }
            const void *dt = upload_text(W, host.data(), total, sym_bytes, st);
            caps_limits_t lim{CAPS_LIMITS_SEGMENTED, CAPS_ORDER_SHORTER_FIRST, ends.data(), ends.size()};
            uint32_t ib = id_bits_for(ends.size());
            bool wide = total > (1ull << 32);
            size_t w = wide ? 8 : 4;
            void *dout = W.out.get(total * w + 8);
            if (sym_bytes == 1 && !wide)
                build_impl<uint8_t, uint32_t>((const uint8_t *)dt, total, nullptr, total, &lim, o, (uint32_t *)dout, nullptr, st, g_stats, ib);
            else if (sym_bytes == 1)
                build_impl<uint8_t, uint64_t>((const uint8_t *)dt, total, nullptr, total, &lim, o, (uint64_t *)dout, nullptr, st, g_stats, ib);
            else if (!wide)
                build_impl<uint16_t, uint32_t>((const uint16_t *)dt, total, nullptr, total, &lim, o, (uint32_t *)dout, nullptr, st, g_stats, ib);
            else
                build_impl<uint16_t, uint64_t>((const uint16_t *)dt, total, nullptr, total, &lim, o, (uint64_t *)dout, nullptr, st, g_stats, ib);
            std::vector<uint8_t> sa(total * w);
            HIPCHK(hipMemcpy(sa.data(), dout, total * w, hipMemcpyDeviceToHost));
            for (size_t t = 0; t < which.size(); ++t) {
                uint64_t start = t ? ends[t - 1] : 0, len = ends[t] - start;
                void *out = items[which[t]].out_sa;
                for (uint64_t r = 0; r < len; ++r) {
                    uint64_t v = wide ? ((const uint64_t *)sa.data())[start + r] : ((const uint32_t *)sa.data())[start + r];
                    v -= start;
                    if (idx_bytes == 4) ((uint32_t *)out)[r] = (uint32_t)v;
                    else ((uint64_t *)out)[r] = v;
                }
            }
        }
    });
}

int caps_lcp_pairs_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *d_pairs,
                          uint64_t count, uint64_t max_ctx, uint64_t *d_out, void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, nullptr);
        hipStream_t st = (hipStream_t)stream;
        if (!count) return;
        unsigned gr = (unsigned)ceil_div(count, 4);
        if (sym_bytes == 1) k_lcp_pairs<uint8_t><<<gr, 256, 0, st>>>((const uint8_t *)d_text, n, d_pairs, count, max_ctx, d_out);
        else k_lcp_pairs<uint16_t><<<gr, 256, 0, st>>>((const uint16_t *)d_text, n, d_pairs, count, max_ctx, d_out);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_suffix_cmp_pairs_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                                 const uint64_t *d_pairs, uint64_t count, uint64_t max_ctx, int32_t *d_out,
                                 void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, limits);
        hipStream_t st = (hipStream_t)stream;
        DevLimits L = upload_limits(workspace(), limits, n, st);
        if (!count) return;
        unsigned gr = (unsigned)ceil_div(count, 4);
        if (sym_bytes == 1) k_cmp_pairs<uint8_t><<<gr, 256, 0, st>>>((const uint8_t *)d_text, L, d_pairs, count, max_ctx, d_out);
        else k_cmp_pairs<uint16_t><<<gr, 256, 0, st>>>((const uint16_t *)d_text, L, d_pairs, count, max_ctx, d_out);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_seed_params_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                            const caps_opts_t *opts, caps_seed_params_t *out, void *stream) {
    return guarded([&] {
        if (!out) throw CapsError(CAPS_E_INVALID, "out is NULL");
        check_common(d_text, n, sym_bytes, limits);
        const caps_opts_t o = opts_or_default(opts);
        if (o.packed_prefix_seed > CAPS_SEED_POLICY_REMAP) throw CapsError(CAPS_E_INVALID, "unknown seed policy");
        SeedDecision d = seed_checks_before_scan(o.packed_prefix_seed, o.max_context, sym_bytes);
        if (d.reason == CAPS_SEED_OK) {  // u8 text: scan its alphabet once
            Workspace &W = workspace();
            std::vector<uint32_t> code;
            uint32_t distinct = 0, maxv = 0;
            bool identity = true;
            scan_alphabet<uint8_t>(W, (const uint8_t *)d_text, n, (hipStream_t)stream, code, distinct, maxv,
                                   identity);
            seed_checks_after_scan(d, o.packed_prefix_seed, o.remap_budget, n, distinct, identity);
        }
        out->reason = d.reason, out->bits = d.bits, out->k = d.k, out->alphabet = d.alphabet, out->remap = d.remap;
    });
}

int caps_seed_keys_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                          uint32_t key_geometry, uint32_t rank, const uint64_t *d_pos, uint64_t count,
                          uint64_t *d_keys, uint32_t *geometry_out, void *stream) {
    return guarded([&] {
        check_common(d_text, n, sym_bytes, limits);
        hipStream_t st = (hipStream_t)stream;
        Workspace &W = workspace();
        DevLimits L = upload_limits(W, limits, n, st);
        uint32_t order = rank ? CAPS_ORDER_STAR : CAPS_ORDER_SHORTER_FIRST;
        unsigned gr = (unsigned)ceil_div(count ? count : 1, 256);
        if (sym_bytes == 1) {
            const bool esc = key_geometry == CAPS_KEY_ESCAPE32;
            Geometry G = compute_geometry<uint8_t>(W, (const uint8_t *)d_text, n, esc ? CAPS_KEY_DENSE : key_geometry,
                                                   order, st, false, CAPS_SEED_POLICY_REMAP, ~0ull, ~0ull, esc);
            if (key_geometry == CAPS_KEY_REFERENCE && !G.reference)
                throw CapsError(CAPS_E_UNSUPPORTED, "no sentinel code fits the alphabet (pack.rs:174-176)");
            if (esc && !G.g.esc) throw CapsError(CAPS_E_UNSUPPORTED, "escape keys take 5..16 symbols");
            if (geometry_out) geometry_out[0] = G.g.bits, geometry_out[1] = G.g.k;
            if (count) k_seed_keys<uint8_t><<<gr, 256, 0, st>>>((const uint8_t *)d_text, codemap<uint8_t>(G), G.g, L, d_pos, count, d_keys);
        } else {
            if (key_geometry == CAPS_KEY_REFERENCE)
                throw CapsError(CAPS_E_UNSUPPORTED, "the reference key geometry is u8-only (pack.rs:223-228)");
            if (key_geometry == CAPS_KEY_ESCAPE32) throw CapsError(CAPS_E_UNSUPPORTED, "escape keys are u8-only");
            Geometry G = compute_geometry<uint16_t>(W, (const uint16_t *)d_text, n, key_geometry, order, st);
            if (geometry_out) geometry_out[0] = G.g.bits, geometry_out[1] = G.g.k;
            if (count) k_seed_keys<uint16_t><<<gr, 256, 0, st>>>((const uint16_t *)d_text, codemap<uint16_t>(G), G.g, L, d_pos, count, d_keys);
        }
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(st));
    });
}

}  // extern "C"

namespace caps {
template <typename S>
static void merge_jobs_host(const void *text, uint64_t n, const caps_limits_t *lim, const uint64_t *sa,
                            const uint64_t *lcp, uint64_t n_sa, std::vector<uint64_t> runs, uint64_t max_ctx,
                            uint64_t *out_sa, uint64_t *out_lcp) {
    Workspace &W = workspace();
    hipStream_t st = nullptr;
    const S *dt = (const S *)upload_text(W, text, n, sizeof(S), st);
    DevLimits L = upload_limits(W, lim, n, st);
    uint64_t *a_sa = (uint64_t *)W.m_a.get(n_sa * 8 + 8), *a_l = (uint64_t *)W.m_al.get(n_sa * 8 + 8);
    uint64_t *b_sa = (uint64_t *)W.m_b.get(n_sa * 8 + 8), *b_l = (uint64_t *)W.m_bl.get(n_sa * 8 + 8);
    HIPCHK(hipMemcpyAsync(a_sa, sa, n_sa * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(a_l, lcp, n_sa * 8, hipMemcpyHostToDevice, st));
    uint64_t *s_sa = a_sa, *s_l = a_l, *d_sa = b_sa, *d_l = b_l;
    while (runs.size() > 1) {
        W.h_jobs.clear();
        std::vector<uint64_t> next;
        uint64_t off = 0, maxlen = 0;
        for (size_t i = 0; i < runs.size(); i += 2) {
            MergeJob jb;
            jb.start = off;
            jb.nx = runs[i];
            jb.ny = i + 1 < runs.size() ? runs[i + 1] : 0;
            W.h_jobs.push_back(jb);
            next.push_back(jb.nx + jb.ny);
            off += jb.nx + jb.ny;
            maxlen = std::max(maxlen, jb.nx + jb.ny);
        }
        MergeJob *dj = (MergeJob *)W.jobs.get(W.h_jobs.size() * sizeof(MergeJob));
        HIPCHK(hipMemcpyAsync(dj, W.h_jobs.data(), W.h_jobs.size() * sizeof(MergeJob), hipMemcpyHostToDevice, st));
        constexpr int MW = 32;
        dim3 gr((unsigned)ceil_div(ceil_div(maxlen, MW), 128), (unsigned)W.h_jobs.size());
        k_lcp_merge<S, uint64_t, MW><<<gr, 128, 0, st>>>(dt, CodeMap<S>{}, L, max_ctx, 0, dj, s_sa, s_l, d_sa, d_l);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(st));
        std::swap(s_sa, d_sa);
        std::swap(s_l, d_l);
        runs.swap(next);
    }
    HIPCHK(hipMemcpyAsync(out_sa, s_sa, n_sa * 8, hipMemcpyDeviceToHost, st));
    if (out_lcp) HIPCHK(hipMemcpyAsync(out_lcp, s_l, n_sa * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
}
}  // namespace caps

extern "C" {

int caps_router_mlp_grad_device(const float *d_params, const uint64_t *d_keys, const float *d_targets, uint32_t count,
                                float *d_grad, float *d_loss, void *stream) {
    return guarded([&] {
        hipStream_t st = (hipStream_t)stream;
        HIPCHK(hipMemsetAsync(d_grad, 0, kMlpParams * 4, st));
        HIPCHK(hipMemsetAsync(d_loss, 0, 4, st));
        if (count) {
            k_mlp_grad<<<(unsigned)ceil_div(count, 256), 256, 0, st>>>(d_params, d_keys, d_targets, count, count,
                                                                        d_grad, d_loss);
            LAUNCHCHK();
        }
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_router_mlp_forward_device(const float *d_params, const uint64_t *d_keys, uint32_t count, float *d_out,
                                   void *stream) {
    return guarded([&] {
        hipStream_t st = (hipStream_t)stream;
        float *nib = (float *)workspace().r_tmp.get(16 * 16 * kMlpHid * 4);
        k_mlp_nibbles<<<16, 16 * kMlpHid, 0, st>>>(d_params, nib);
        LAUNCHCHK();
        if (count) {
            k_mlp_predict<<<(unsigned)ceil_div(count, 256), 256, 0, st>>>(nib, d_params + kMlpIn * kMlpHid + kMlpHid,
                                                                           d_keys, count, d_out);
            LAUNCHCHK();
        }
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_router_pwl_fit_device(const uint64_t *d_keys, const uint32_t *d_labels, uint64_t s, uint32_t ns,
                               uint32_t n_buckets, uint64_t *knots, float *a, float *b, float *inv, uint32_t *lo,
                               uint32_t *hi, void *stream) {
    return guarded([&] {
        if (s == 0 || ns == 0 || ns > s) throw CapsError(CAPS_E_INVALID, "need 1 <= ns <= s");
        if (!knots || !a || !b || !inv || !lo || !hi) throw CapsError(CAPS_E_INVALID, "output pointer is NULL");
        hipStream_t st = (hipStream_t)stream;
        Workspace &W = workspace();
        uint64_t *dk = (uint64_t *)W.r_knots.get((ns + 1) * 8);
        float *da = (float *)W.r_a.get(ns * 4), *db = (float *)W.r_b.get(ns * 4), *di = (float *)W.r_inv.get(ns * 4);
        uint32_t *dlo = (uint32_t *)W.r_lo.get(ns * 4), *dhi = (uint32_t *)W.r_hi.get(ns * 4);
        k_pwl_fit<<<ns, 256, 0, st>>>(d_keys, d_labels, s, ns, n_buckets, dk, da, db, di, dlo, dhi);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(knots, dk, (ns + 1) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(a, da, ns * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(b, db, ns * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(inv, di, ns * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(lo, dlo, ns * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hi, dhi, ns * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_router_spline_fit_device(const uint64_t *d_splitters, uint32_t n, uint16_t *knots, int32_t *bound,
                                  void *stream) {
    return guarded([&] {
        if (n >= 65536) throw CapsError(CAPS_E_INVALID, "at most 65,535 splitters");
        if (!knots || !bound) throw CapsError(CAPS_E_INVALID, "output pointer is NULL");
        if (n && !d_splitters) throw CapsError(CAPS_E_INVALID, "splitter pointer is NULL");
        hipStream_t st = (hipStream_t)stream;
        Workspace &W = workspace();
        uint16_t *dk = (uint16_t *)W.r_radix.get((kSplineCells + 1) * 2);
        int *db = (int *)W.r_lo.get(8);
        HIPCHK(hipMemsetAsync(db, 0, 8, st));
        spline_fit(d_splitters, n, dk, db, st);
        HIPCHK(hipMemcpyAsync(knots, dk, (kSplineCells + 1) * 2, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(bound, db, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    });
}

int caps_merge_lcp(const void *text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits, const uint64_t *x,
                   const uint64_t *lcp_x, uint64_t nx, const uint64_t *y, const uint64_t *lcp_y, uint64_t ny,
                   uint64_t max_ctx, uint64_t *z, uint64_t *lcp_z) {
    return guarded([&] {
        check_common(text, n, sym_bytes, limits);
        std::vector<uint64_t> sa(nx + ny), lc(nx + ny);
        std::copy(x, x + nx, sa.begin());
        std::copy(y, y + ny, sa.begin() + nx);
        std::copy(lcp_x, lcp_x + nx, lc.begin());
        std::copy(lcp_y, lcp_y + ny, lc.begin() + nx);
        std::vector<uint64_t> runs;
        runs.push_back(nx);
        runs.push_back(ny);
        if (nx + ny == 0) return;
        if (sym_bytes == 1) merge_jobs_host<uint8_t>(text, n, limits, sa.data(), lc.data(), nx + ny, runs, max_ctx, z, lcp_z);
        else merge_jobs_host<uint16_t>(text, n, limits, sa.data(), lc.data(), nx + ny, runs, max_ctx, z, lcp_z);
    });
}

int caps_cascade_merge(const void *text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                       const uint64_t *sa, const uint64_t *lcp, uint64_t n_sa, const uint64_t *boundaries,
                       uint64_t n_boundaries, uint64_t max_ctx, uint64_t *out_sa, uint64_t *out_lcp) {
    return guarded([&] {
        check_common(text, n, sym_bytes, limits);
        if (n_sa == 0) return;
        std::vector<uint64_t> runs;  // ext_mem.rs:2215-2221: non-empty sub-subarrays
        for (uint64_t i = 0; i + 1 < n_boundaries; ++i)
            if (boundaries[i + 1] > boundaries[i]) runs.push_back(boundaries[i + 1] - boundaries[i]);
        if (sym_bytes == 1) merge_jobs_host<uint8_t>(text, n, limits, sa, lcp, n_sa, runs, max_ctx, out_sa, out_lcp);
        else merge_jobs_host<uint16_t>(text, n, limits, sa, lcp, n_sa, runs, max_ctx, out_sa, out_lcp);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner: times real builds of the caller's input over a small grid of
// knobs and writes the fastest into *opts: the mean bucket size the splitters
// aim at (keys per LDS tile), the tile sort (merge path or 4-bit radix), the
// refinement launch style (per-class plan or persistent work queues for every
// class) and the merge-path partition width of the giant-group merges.  Each
// configuration is built `reps` times after one warm build; the device time
// of the fastest build counts.  The output of every build is the same suffix
// array, so tuning only moves time.
// ---------------------------------------------------------------------------
namespace {
struct TuneCand {
    uint32_t tile_keys, flags;
};
template <typename RUN>
static void tune(caps_opts_t &o, uint32_t reps, caps_tune_report_t *rep, RUN &&run) {
    static const uint32_t kKeys[] = {1280, 1536, 1792};
    static const uint32_t kFlags[] = {0, CAPS_FLAG_TILE_RADIX, CAPS_FLAG_PERSISTENT_REFINE, CAPS_FLAG_MERGE_W16,
                                      CAPS_FLAG_MERGE_W64, CAPS_FLAG_MID_BY_SIZE};
    std::vector<TuneCand> cands;
    const uint32_t keep = o.flags & ~(CAPS_FLAG_TILE_RADIX | CAPS_FLAG_PERSISTENT_REFINE | CAPS_FLAG_MERGE_W16 |
                                      CAPS_FLAG_MERGE_W64 | CAPS_FLAG_MID_BY_SIZE);
    for (uint32_t k : kKeys) cands.push_back({k, keep});
    for (size_t f = 1; f < sizeof(kFlags) / sizeof(kFlags[0]); ++f) cands.push_back({0, keep | kFlags[f]});
    double best = 1e300;
    TuneCand pick{o.tile_keys, o.flags};
    uint32_t tried = 0;
    for (size_t ci = 0; ci < cands.size(); ++ci) {
        // the flag candidates take the best bucket size found so far
        TuneCand c = cands[ci];
        if (c.tile_keys == 0) c.tile_keys = pick.tile_keys;
        caps_opts_t t = o;
        t.tile_keys = c.tile_keys;
        t.flags = c.flags;
        run(t);  // warm: buffers sized, code loaded
        double ms = 1e300;
        for (uint32_t r = 0; r < std::max<uint32_t>(reps, 1); ++r) {
            run(t);
            ms = std::min(ms, g_stats.ms_total);
        }
        if (rep && tried < CAPS_TUNE_MAX) {
            rep->ms[tried] = ms;
            rep->tile_keys[tried] = c.tile_keys;
            rep->flags[tried] = c.flags;
        }
        ++tried;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    o.tile_keys = pick.tile_keys;
    o.flags = pick.flags;
    if (rep) {
        rep->n_tried = std::min<uint32_t>(tried, CAPS_TUNE_MAX);
        rep->best_ms = best;
    }
}
}  // namespace

extern "C" {

int caps_autotune_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const caps_limits_t *limits,
                         uint32_t idx_bytes, void *d_scratch_sa, uint32_t reps, caps_opts_t *opts, void *stream,
                         caps_tune_report_t *report) {
    if (!opts) return guarded([] { throw CapsError(CAPS_E_INVALID, "opts is required"); });
    caps_opts_t o = *opts;
    int rc = guarded([&] {
        if (report) memset(report, 0, sizeof(*report));
        tune(o, reps, report, [&](const caps_opts_t &t) {
            int r = caps_build_device(d_text, n, sym_bytes, nullptr, n, idx_bytes, limits, &t, d_scratch_sa, nullptr,
                                      stream);
            if (r != CAPS_OK) throw CapsError(r, caps_last_error());
        });
    });
    if (rc == CAPS_OK) *opts = o;
    return rc;
}

int caps_autotune_batch_device(const void *d_text, uint64_t n, uint32_t sym_bytes, const uint64_t *ends,
                               uint64_t n_texts, uint32_t idx_bytes, void *d_scratch_sa, uint32_t reps,
                               caps_opts_t *opts, void *stream, caps_tune_report_t *report) {
    if (!opts) return guarded([] { throw CapsError(CAPS_E_INVALID, "opts is required"); });
    caps_opts_t o = *opts;
    int rc = guarded([&] {
        if (report) memset(report, 0, sizeof(*report));
        tune(o, reps, report, [&](const caps_opts_t &t) {
            int r = caps_build_batch_device(d_text, n, sym_bytes, ends, n_texts, idx_bytes, &t, d_scratch_sa, stream);
            if (r != CAPS_OK) throw CapsError(r, caps_last_error());
        });
    });
    if (rc == CAPS_OK) *opts = o;
    return rc;
}

}  // extern "C"
