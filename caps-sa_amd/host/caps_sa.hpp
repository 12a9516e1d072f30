// C++ mirror of caps-sa 0.7.0's public API (lib.rs:28-45) over the C ABI of
// include/caps_sa.h.  Same function names, argument meaning and error
// behaviour as the crate:
//   - in-memory builders return the suffix array (sample_sort.rs:97-206) and
//     throw std::invalid_argument where the crate panics (bad segment ends,
//     limits.rs:247-281);
//   - streaming builders call `emit` sequentially, in lexicographic order,
//     with positions widened to u64 (ext_mem.rs:711-871); the plain forms
//     throw BuildError on failure, the try_ forms return it (ext_mem.rs:
//     257-277): Kind::Emit carries the callback's error, Kind::Io a device
//     or host failure;
//   - a symbol or index type outside the GPU build (anything but u8/u16 and
//     u32/u64) is a compile-time error here, where the crate would take its
//     CPU path.
// Header-only; link against libcaps_sa_hip.so.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <limits>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/caps_sa.h"

namespace caps_sa {

// ---- limit providers (limits.rs) ------------------------------------------

enum class BoundaryRank { ShorterFirst, LongerFirst };  // limits.rs:113-123

// PlainText (limits.rs:130-156): every suffix runs to the end of the text.
struct PlainText {
    uint64_t n = 0;
    explicit PlainText(uint64_t text_len) : n(text_len) {}
    uint64_t lim_at(uint64_t p) const { return n - p; }
    BoundaryRank boundary_rank() const { return BoundaryRank::ShorterFirst; }
    caps_limits_t c_limits() const { return caps_limits_t{CAPS_LIMITS_PLAIN, CAPS_ORDER_SHORTER_FIRST, nullptr, 0}; }
};

// SegmentedText (limits.rs:182-336): suffixes stop at their segment's end.
// `star` selects STAR's boundary order lim_b.cmp(lim_a).then(p_a.cmp(p_b))
// (BoundaryRank::LongerFirst) instead of the default lim_a.cmp(lim_b).
class SegmentedText {
  public:
    static SegmentedText from_ends(uint64_t text_len, std::vector<uint64_t> ends, bool star = false) {
        for (size_t i = 1; i < ends.size(); ++i)
            if (!(ends[i - 1] < ends[i]))
                throw std::invalid_argument("SegmentedText::from_ends: ends must be strictly increasing");
        if (ends.empty() && text_len != 0)
            throw std::invalid_argument("SegmentedText::from_ends: empty ends but text_len != 0");
        if (!ends.empty() && ends.back() != text_len)
            throw std::invalid_argument("SegmentedText::from_ends: last end != text_len");
        return SegmentedText(text_len, std::move(ends), star);
    }
    static SegmentedText from_lengths(uint64_t text_len, const std::vector<uint64_t> &lengths, bool star = false) {
        std::vector<uint64_t> ends;
        uint64_t acc = 0;
        for (uint64_t l : lengths) ends.push_back(acc += l);
        if (acc != text_len)
            throw std::invalid_argument("SegmentedText::from_lengths: per-segment lengths do not sum to text_len");
        return from_ends(text_len, std::move(ends), star);
    }
    uint64_t n_segments() const { return ends_.size(); }
    uint64_t lim_at(uint64_t p) const {
        size_t lo = 0, hi = ends_.size();
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (ends_[mid] <= p) lo = mid + 1;
            else hi = mid;
        }
        return (lo < ends_.size() ? ends_[lo] : n_) - p;
    }
    BoundaryRank boundary_rank() const { return star_ ? BoundaryRank::LongerFirst : BoundaryRank::ShorterFirst; }
    caps_limits_t c_limits() const {
        return caps_limits_t{CAPS_LIMITS_SEGMENTED, star_ ? CAPS_ORDER_STAR : CAPS_ORDER_SHORTER_FIRST, ends_.data(),
                             ends_.size()};
    }

  private:
    SegmentedText(uint64_t n, std::vector<uint64_t> ends, bool star) : n_(n), ends_(std::move(ends)), star_(star) {}
    uint64_t n_;
    std::vector<uint64_t> ends_;
    bool star_;
};

// ---- options -----------------------------------------------------------------

// Opts (sample_sort.rs:71-85).
struct Opts {
    uint64_t max_context = std::numeric_limits<uint64_t>::max();
    Opts &with_max_context(uint64_t m) {
        max_context = m;
        return *this;
    }
};

// GeometricMemoizationConfig (lcp_memo.rs:64-145): tuning of the crate's
// geometric LCP memo; every value non-zero (std::invalid_argument otherwise).
// Memoization never changes the output; the GPU build accepts the policy and
// does not use it (its long-LCP scans read packed code words and fast-forward
// over shared stretches instead).
struct GeometricMemoizationConfig {
    uint64_t probe_symbols = 256;
    uint64_t min_lcp_symbols = 1024;
    uint64_t activate_after_entries = 64;
    uint64_t max_entries_per_partition = 4096;
    GeometricMemoizationConfig with_probe_symbols(uint64_t v) const { return set(&GeometricMemoizationConfig::probe_symbols, v); }
    GeometricMemoizationConfig with_min_lcp_symbols(uint64_t v) const {
        return set(&GeometricMemoizationConfig::min_lcp_symbols, v);
    }
    GeometricMemoizationConfig with_activate_after_entries(uint64_t v) const {
        return set(&GeometricMemoizationConfig::activate_after_entries, v);
    }
    GeometricMemoizationConfig with_max_entries_per_partition(uint64_t v) const {
        return set(&GeometricMemoizationConfig::max_entries_per_partition, v);
    }

  private:
    GeometricMemoizationConfig set(uint64_t GeometricMemoizationConfig::*f, uint64_t v) const {
        if (v == 0) throw std::invalid_argument("memoization settings must be non-zero");
        GeometricMemoizationConfig c = *this;
        c.*f = v;
        return c;
    }
};

namespace detail {
// read_env_usize / read_env_bool (ext_mem.rs:241-253)
inline std::optional<uint64_t> env_usize(const char *name) {
    const char *v = std::getenv(name);
    if (!v) return std::nullopt;
    if (*v == '+') ++v;
    if (!*v) return std::nullopt;
    uint64_t x = 0;
    for (; *v; ++v) {
        if (*v < '0' || *v > '9') return std::nullopt;
        const uint64_t d = (uint64_t)(*v - '0');
        if (x > (std::numeric_limits<uint64_t>::max() - d) / 10) return std::nullopt;
        x = x * 10 + d;
    }
    return x;
}
inline bool env_bool(const char *name) {
    const char *v = std::getenv(name);
    if (!v) return false;
    const std::string s(v);
    return s == "1" || s == "true" || s == "yes" || s == "on";
}
}  // namespace detail

// ExtMemOpts (ext_mem.rs:58-239): the fields with a meaning on the GPU, with
// the crate's builder methods, plus the memoization policy (std::nullopt:
// LcpMemoizationPolicy::Disabled), accepted and unused.  The output never
// depends on them.
struct ExtMemOpts {
    uint64_t max_context = std::numeric_limits<uint64_t>::max();
    uint64_t subproblem_count = 0;
    uint32_t router = CAPS_ROUTER_BSEARCH;
    uint64_t batch_records = 0;  // GPU knob: records per bucket batch (0: auto)
    // PackedPrefixSeedPolicy (pack.rs:37-69): CAPS_SEED_POLICY_*, with the
    // Remap byte budget
    uint32_t packed_prefix_seed = CAPS_SEED_POLICY_DISABLED;
    uint64_t remap_budget = 0;
    std::optional<GeometricMemoizationConfig> lcp_memoization;
    bool collect_lcp_memoization_stats = false;
    // ExtMemOpts::from_env (ext_mem.rs:154-196) for the variables with a GPU
    // meaning and the memoization policy.
    static ExtMemOpts from_env() {
        ExtMemOpts o;
        if (auto v = detail::env_usize("CAPS_SA_SUBPROBLEMS")) o.subproblem_count = *v;
        if (auto v = detail::env_usize("CAPS_SA_MAX_CONTEXT")) o.max_context = *v;
        if (auto v = detail::env_usize("CAPS_SA_PACKED_PREFIX_REMAP_BYTES")) {
            o.packed_prefix_seed = CAPS_SEED_POLICY_REMAP;
            o.remap_budget = *v;
        } else if (detail::env_bool("CAPS_SA_PACKED_PREFIX_SEED")) {
            o.packed_prefix_seed = CAPS_SEED_POLICY_DENSE_ALPHABET_ONLY;
        }
        if (detail::env_bool("CAPS_SA_GEOMETRIC_MEMO")) {
            GeometricMemoizationConfig c;
            if (auto v = detail::env_usize("CAPS_SA_MEMO_PROBE"); v && *v) c = c.with_probe_symbols(*v);
            if (auto v = detail::env_usize("CAPS_SA_MEMO_MIN_LCP"); v && *v) c = c.with_min_lcp_symbols(*v);
            if (auto v = detail::env_usize("CAPS_SA_MEMO_ACTIVATE_ENTRIES"); v && *v)
                c = c.with_activate_after_entries(*v);
            if (auto v = detail::env_usize("CAPS_SA_MEMO_CAPACITY"); v && *v) c = c.with_max_entries_per_partition(*v);
            o.lcp_memoization = c;
        }
        o.collect_lcp_memoization_stats = detail::env_bool("CAPS_SA_MEMO_STATS");
        return o;
    }
    ExtMemOpts &with_lcp_memoization(std::optional<GeometricMemoizationConfig> p) {
        lcp_memoization = p;
        return *this;
    }
    ExtMemOpts &with_max_context(uint64_t m) {
        max_context = m;
        return *this;
    }
    ExtMemOpts &with_subproblem_count(uint64_t p) {
        subproblem_count = p;
        return *this;
    }
    ExtMemOpts &with_router(uint32_t r) {
        router = r;
        return *this;
    }
    ExtMemOpts &with_batch_records(uint64_t b) {
        batch_records = b;
        return *this;
    }
};

// ---- errors ------------------------------------------------------------------

// BuildError (ext_mem.rs:257-277).
template <class E>
struct BuildError : std::runtime_error {
    enum class Kind { Io, Emit };
    Kind kind;
    std::optional<E> emit_error;
    BuildError(Kind k, std::string msg, std::optional<E> e = std::nullopt)
        : std::runtime_error(std::move(msg)), kind(k), emit_error(std::move(e)) {}
};

namespace detail {
template <class S>
constexpr uint32_t sym_bytes() {
    static_assert(std::is_same_v<S, uint8_t> || std::is_same_v<S, uint16_t>,
                  "the GPU build takes u8 or u16 symbols");
    return sizeof(S);
}
template <class I>
constexpr uint32_t idx_bytes() {
    static_assert(std::is_same_v<I, uint32_t> || std::is_same_v<I, uint64_t>,
                  "the GPU build produces u32 or u64 indices");
    return sizeof(I);
}
inline caps_opts_t c_opts(uint64_t max_context) {
    caps_opts_t o;
    caps_opts_default(&o);
    o.max_context = max_context;
    return o;
}
inline caps_opts_t c_opts(const ExtMemOpts &e) {
    caps_opts_t o = c_opts(e.max_context);
    o.subproblem_count = e.subproblem_count;
    o.router = e.router;
    o.batch_records = e.batch_records;
    o.packed_prefix_seed = e.packed_prefix_seed;
    o.remap_budget = e.remap_budget;
    o.key_geometry = e.packed_prefix_seed == CAPS_SEED_POLICY_DISABLED ? CAPS_KEY_DENSE : CAPS_KEY_REFERENCE;
    return o;
}
inline void check(int rc) {
    if (rc == CAPS_OK) return;
    if (rc == CAPS_E_INVALID) throw std::invalid_argument(caps_last_error());
    throw std::runtime_error(caps_last_error());
}
// The emit trampoline: parks the first error of the caller's callback and
// stops the build with a nonzero return.
template <class E>
struct EmitCtx {
    std::function<std::optional<E>(uint64_t)> *f;
    std::optional<E> err;
};
template <class E>
int trampoline(void *ctx, const uint64_t *pos, uint64_t count) {
    auto *c = static_cast<EmitCtx<E> *>(ctx);
    for (uint64_t i = 0; i < count; ++i) {
        std::optional<E> e = (*c->f)(pos[i]);
        if (e) {
            c->err = std::move(e);
            return 1;
        }
    }
    return 0;
}
}  // namespace detail

// ---- in-memory builders (sample_sort.rs:97-206) --------------------------------

template <class S, class I, class L>
std::vector<I> build_in_memory_with(const std::vector<S> &text, const L &lp, const Opts &opts) {
    std::vector<I> sa(text.size());
    caps_limits_t lim = lp.c_limits();
    caps_opts_t o = detail::c_opts(opts.max_context);
    detail::check(caps_build_in_memory(text.data(), text.size(), detail::sym_bytes<S>(), detail::idx_bytes<I>(), &lim,
                                       &o, sa.data(), nullptr));
    return sa;
}
template <class S, class I>
std::vector<I> build_in_memory_with_opts(const std::vector<S> &text, const Opts &opts) {
    return build_in_memory_with<S, I>(text, PlainText(text.size()), opts);
}
template <class S, class I>
std::vector<I> build_in_memory(const std::vector<S> &text) {
    return build_in_memory_with_opts<S, I>(text, Opts{});
}

// The positions are moved in and their buffer becomes the result, as the
// crate reuses the moved-in Vec (sample_sort.rs:184).
template <class S, class I, class L>
std::vector<I> build_in_memory_for_positions_with(const std::vector<S> &text, std::vector<I> positions, const L &lp,
                                                  const Opts &opts) {
    caps_limits_t lim = lp.c_limits();
    caps_opts_t o = detail::c_opts(opts.max_context);
    detail::check(caps_build_in_memory_for_positions(text.data(), text.size(), detail::sym_bytes<S>(),
                                                     positions.data(), positions.size(), detail::idx_bytes<I>(), &lim,
                                                     &o, positions.data(), nullptr));
    return positions;
}
template <class S, class I>
std::vector<I> build_in_memory_for_positions_with_opts(const std::vector<S> &text, std::vector<I> positions,
                                                       const Opts &opts) {
    return build_in_memory_for_positions_with<S, I>(text, std::move(positions), PlainText(text.size()), opts);
}
template <class S, class I>
std::vector<I> build_in_memory_for_positions(const std::vector<S> &text, std::vector<I> positions) {
    return build_in_memory_for_positions_with_opts<S, I>(text, std::move(positions), Opts{});
}

// ---- streaming builders (ext_mem.rs:370-871) -----------------------------------

// try_ forms: emit returns std::nullopt to go on or an error to abort; the
// result is the first error (Kind::Emit) or a device/host failure (Kind::Io).
template <class E, class S, class L>
std::optional<BuildError<E>> try_build_in_memory_sample_sort_for_positions_with(
    const std::vector<S> &text, const std::vector<uint64_t> *positions, const L &lp, const ExtMemOpts &opts,
    std::function<std::optional<E>(uint64_t)> emit) {
    caps_limits_t lim = lp.c_limits();
    caps_opts_t o = detail::c_opts(opts);
    detail::EmitCtx<E> ctx{&emit, std::nullopt};
    int rc = caps_build_in_memory_sample_sort(text.data(), text.size(), detail::sym_bytes<S>(),
                                              positions ? positions->data() : nullptr,
                                              positions ? positions->size() : text.size(), &lim, &o,
                                              &detail::trampoline<E>, &ctx);
    if (rc == CAPS_OK) return std::nullopt;
    if (rc == CAPS_E_EMIT) return BuildError<E>(BuildError<E>::Kind::Emit, "emit failed", std::move(ctx.err));
    if (rc == CAPS_E_INVALID) throw std::invalid_argument(caps_last_error());
    return BuildError<E>(BuildError<E>::Kind::Io, caps_last_error());
}
template <class E, class S, class L>
std::optional<BuildError<E>> try_build_in_memory_sample_sort_with(const std::vector<S> &text, const L &lp,
                                                                  const ExtMemOpts &opts,
                                                                  std::function<std::optional<E>(uint64_t)> emit) {
    return try_build_in_memory_sample_sort_for_positions_with<E>(text, nullptr, lp, opts, std::move(emit));
}
template <class E, class S>
std::optional<BuildError<E>> try_build_in_memory_sample_sort(const std::vector<S> &text, const ExtMemOpts &opts,
                                                             std::function<std::optional<E>(uint64_t)> emit) {
    return try_build_in_memory_sample_sort_with<E>(text, PlainText(text.size()), opts, std::move(emit));
}
template <class E, class S>
std::optional<BuildError<E>> try_build_in_memory_sample_sort_for_positions(
    const std::vector<S> &text, const std::vector<uint64_t> &positions, const ExtMemOpts &opts,
    std::function<std::optional<E>(uint64_t)> emit) {
    return try_build_in_memory_sample_sort_for_positions_with<E>(text, &positions, PlainText(text.size()), opts,
                                                                 std::move(emit));
}

// Plain forms: emit cannot fail; failures throw BuildError<std::string>.
template <class S, class L>
void build_in_memory_sample_sort_with(const std::vector<S> &text, const L &lp, const ExtMemOpts &opts,
                                      std::function<void(uint64_t)> emit) {
    auto r = try_build_in_memory_sample_sort_with<std::string>(
        text, lp, opts, [&](uint64_t p) -> std::optional<std::string> {
            emit(p);
            return std::nullopt;
        });
    if (r) throw *r;
}
template <class S>
void build_in_memory_sample_sort(const std::vector<S> &text, const ExtMemOpts &opts,
                                 std::function<void(uint64_t)> emit) {
    build_in_memory_sample_sort_with(text, PlainText(text.size()), opts, std::move(emit));
}

// build_ext_mem_for_positions(_with) (ext_mem.rs:370-449): the records stay
// in HBM instead of spilling to disk.
template <class E, class S, class L>
std::optional<BuildError<E>> try_build_ext_mem_for_positions_with(const std::vector<S> &text,
                                                                  const std::vector<uint64_t> &positions, const L &lp,
                                                                  const ExtMemOpts &opts,
                                                                  std::function<std::optional<E>(uint64_t)> emit) {
    caps_limits_t lim = lp.c_limits();
    caps_opts_t o = detail::c_opts(opts);
    detail::EmitCtx<E> ctx{&emit, std::nullopt};
    int rc = caps_build_ext_mem_for_positions(text.data(), text.size(), detail::sym_bytes<S>(), positions.data(),
                                              positions.size(), &lim, &o, &detail::trampoline<E>, &ctx);
    if (rc == CAPS_OK) return std::nullopt;
    if (rc == CAPS_E_EMIT) return BuildError<E>(BuildError<E>::Kind::Emit, "emit failed", std::move(ctx.err));
    if (rc == CAPS_E_INVALID) throw std::invalid_argument(caps_last_error());
    return BuildError<E>(BuildError<E>::Kind::Io, caps_last_error());
}
template <class S, class L>
void build_ext_mem_for_positions_with(const std::vector<S> &text, const std::vector<uint64_t> &positions, const L &lp,
                                      const ExtMemOpts &opts, std::function<void(uint64_t)> emit) {
    auto r = try_build_ext_mem_for_positions_with<std::string>(
        text, positions, lp, opts, [&](uint64_t p) -> std::optional<std::string> {
            emit(p);
            return std::nullopt;
        });
    if (r) throw *r;
}

// build_ext_mem(_with) (ext_mem.rs:289-356): every position of the text.  On
// the GPU this is the same build as build_in_memory_sample_sort over all
// positions (the crate's two paths differ only in where buckets are kept).
template <class E, class S, class L>
std::optional<BuildError<E>> try_build_ext_mem_with(const std::vector<S> &text, const L &lp, const ExtMemOpts &opts,
                                                    std::function<std::optional<E>(uint64_t)> emit) {
    return try_build_in_memory_sample_sort_for_positions_with<E>(text, nullptr, lp, opts, std::move(emit));
}
template <class E, class S>
std::optional<BuildError<E>> try_build_ext_mem(const std::vector<S> &text, const ExtMemOpts &opts,
                                               std::function<std::optional<E>(uint64_t)> emit) {
    return try_build_ext_mem_with<E>(text, PlainText(text.size()), opts, std::move(emit));
}
template <class S, class L>
void build_ext_mem_with(const std::vector<S> &text, const L &lp, const ExtMemOpts &opts,
                        std::function<void(uint64_t)> emit) {
    auto r = try_build_ext_mem_with<std::string>(text, lp, opts, [&](uint64_t p) -> std::optional<std::string> {
        emit(p);
        return std::nullopt;
    });
    if (r) throw *r;
}
template <class S>
void build_ext_mem(const std::vector<S> &text, const ExtMemOpts &opts, std::function<void(uint64_t)> emit) {
    build_ext_mem_with(text, PlainText(text.size()), opts, std::move(emit));
}

// build_ext_mem_for_filter(_with) (ext_mem.rs:472-547): the keep predicate is
// walked once into a bitmap of kept positions, as the crate's FilteredSource
// does (ext_mem.rs:940-990), and never called again; no position list is built.
template <class E, class S, class L, class Pred>
std::optional<BuildError<E>> try_build_ext_mem_for_filter_with(const std::vector<S> &text, Pred keep, const L &lp,
                                                               const ExtMemOpts &opts,
                                                               std::function<std::optional<E>(uint64_t)> emit) {
    const uint64_t n = text.size();
    std::vector<uint64_t> bits((n + 63) / 64, 0);
    for (uint64_t q = 0; q < n; ++q)
        if (keep(q)) bits[q >> 6] |= 1ull << (q & 63);
    caps_limits_t lim = lp.c_limits();
    caps_opts_t o = detail::c_opts(opts);
    detail::EmitCtx<E> ctx{&emit, std::nullopt};
    int rc = caps_build_ext_mem_for_filter(text.data(), n, detail::sym_bytes<S>(), bits.data(), &lim, &o,
                                           &detail::trampoline<E>, &ctx);
    if (rc == CAPS_OK) return std::nullopt;
    if (rc == CAPS_E_EMIT) return BuildError<E>(BuildError<E>::Kind::Emit, "emit failed", std::move(ctx.err));
    if (rc == CAPS_E_INVALID) throw std::invalid_argument(caps_last_error());
    return BuildError<E>(BuildError<E>::Kind::Io, caps_last_error());
}
template <class S, class L, class Pred>
void build_ext_mem_for_filter_with(const std::vector<S> &text, Pred keep, const L &lp, const ExtMemOpts &opts,
                                   std::function<void(uint64_t)> emit) {
    auto r = try_build_ext_mem_for_filter_with<std::string>(text, keep, lp, opts,
                                                            [&](uint64_t p) -> std::optional<std::string> {
                                                                emit(p);
                                                                return std::nullopt;
                                                            });
    if (r) throw *r;
}
template <class S, class Pred>
void build_ext_mem_for_filter(const std::vector<S> &text, Pred keep, const ExtMemOpts &opts,
                              std::function<void(uint64_t)> emit) {
    build_ext_mem_for_filter_with(text, keep, PlainText(text.size()), opts, emit);
}

// ---- batched builds (no counterpart in the crate) ------------------------------

// build_in_memory for every text, as one GPU build per call.
template <class S, class I>
std::vector<std::vector<I>> build_batch(const std::vector<std::vector<S>> &texts) {
    std::vector<std::vector<I>> out(texts.size());
    std::vector<caps_text_item_t> items(texts.size());
    for (size_t i = 0; i < texts.size(); ++i) {
        out[i].resize(texts[i].size());
        items[i] = caps_text_item_t{texts[i].empty() ? nullptr : texts[i].data(), texts[i].size(),
                                    texts[i].empty() ? nullptr : out[i].data()};
    }
    caps_opts_t o = detail::c_opts(std::numeric_limits<uint64_t>::max());
    detail::check(caps_build_batch(items.data(), items.size(), detail::sym_bytes<S>(), detail::idx_bytes<I>(), &o));
    return out;
}

}  // namespace caps_sa
