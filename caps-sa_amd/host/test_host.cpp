// The reference crate's own fixture tests, through the C++ mirror of its API
// (caps_sa.hpp) on the GPU library.  Fixtures follow sample_sort.rs:792-821,
// 1060-1070 and ext_mem.rs:2784-2806, 2895-2915; brute force sorts suffixes
// by their slices as sample_sort.rs:643-647 does.  Prints "ok" and exits 0
// when every check holds.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "caps_sa.hpp"

using namespace caps_sa;

static int failures = 0;
#define CHECK(...)                                                                        \
    do {                                                                                  \
        if (!(__VA_ARGS__)) {                                                             \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #__VA_ARGS__); \
            ++failures;                                                                   \
        }                                                                                 \
    } while (0)

template <class S>
static std::vector<uint64_t> brute(const std::vector<S> &t) {
    std::vector<uint64_t> sa(t.size());
    std::iota(sa.begin(), sa.end(), 0);
    std::sort(sa.begin(), sa.end(), [&](uint64_t a, uint64_t b) {
        return std::lexicographical_compare(t.begin() + a, t.end(), t.begin() + b, t.end());
    });
    return sa;
}

template <class I>
static std::vector<uint64_t> widen(const std::vector<I> &v) {
    return std::vector<uint64_t>(v.begin(), v.end());
}

static std::vector<uint8_t> bytes(const char *s) { return std::vector<uint8_t>(s, s + std::char_traits<char>::length(s)); }

static uint64_t lcg_state = 12345;
static uint32_t lcg() {  // a small seeded generator for the random cases
    lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(lcg_state >> 33);
}

int main() {
    // fixtures (sample_sort.rs:792-821, ext_mem.rs:2784-2791)
    CHECK((widen(build_in_memory<uint8_t, uint32_t>(bytes("banana"))) == std::vector<uint64_t>{5, 3, 1, 0, 4, 2}));
    CHECK((widen(build_in_memory<uint8_t, uint32_t>(bytes("mississippi"))) ==
           std::vector<uint64_t>{10, 7, 4, 1, 0, 9, 8, 6, 3, 5, 2}));
    CHECK((widen(build_in_memory<uint8_t, uint64_t>(bytes("abracadabra"))) ==
           std::vector<uint64_t>{10, 7, 0, 3, 5, 8, 1, 4, 6, 9, 2}));
    CHECK((build_in_memory<uint8_t, uint32_t>(std::vector<uint8_t>{7}) == std::vector<uint32_t>{0}));
    CHECK(build_in_memory<uint8_t, uint32_t>(std::vector<uint8_t>{}).empty());
    {
        std::vector<uint8_t> t{0, 1, 2, 0, 1, 5, 0, 2, 1, 6};
        CHECK(widen(build_in_memory<uint8_t, uint32_t>(t)) == brute(t));
    }
    // random u8 and u16 texts against brute force
    for (int rep = 0; rep < 6; ++rep) {
        std::vector<uint8_t> t8(2000 + lcg() % 3000);
        for (auto &c : t8) c = (uint8_t)(lcg() % (rep + 2));
        CHECK(widen(build_in_memory<uint8_t, uint32_t>(t8)) == brute(t8));
        std::vector<uint16_t> t16(1500 + lcg() % 2000);
        for (auto &c : t16) c = (uint16_t)(lcg() % (rep * 700 + 3));
        CHECK(widen(build_in_memory<uint16_t, uint64_t>(t16)) == brute(t16));
    }
    // STAR order on AAAA: longer first, then position (sample_sort.rs:1060-1070)
    {
        std::vector<uint8_t> t = bytes("AAAA");
        auto lp = SegmentedText::from_ends(4, {4}, /*star=*/true);
        CHECK((widen(build_in_memory_with<uint8_t, uint32_t>(t, lp, Opts{})) == std::vector<uint64_t>{0, 1, 2, 3}));
    }
    // segment ends are validated like the crate's constructors (limits.rs:247-281)
    {
        bool threw = false;
        try {
            SegmentedText::from_ends(10, {4, 4, 10});
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK(threw);
        threw = false;
        try {
            SegmentedText::from_lengths(10, {3, 3});
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK(threw);
    }
    // positions subset equals the restriction of the full suffix array
    {
        std::vector<uint8_t> t(5000);
        for (auto &c : t) c = (uint8_t)(lcg() % 3);
        std::vector<uint32_t> pos;
        for (uint32_t p = 0; p < t.size(); p += 3) pos.push_back(p);
        auto sub = build_in_memory_for_positions<uint8_t, uint32_t>(t, pos);
        auto full = brute(t);
        std::vector<uint64_t> want;
        for (uint64_t p : full)
            if (p % 3 == 0) want.push_back(p);
        CHECK(widen(sub) == want);
    }
    // max_context: the compare stops after max_context symbols, then the
    // boundary order (shorter suffix first) and the position decide
    // (lcp.rs:184-201)
    {
        std::vector<uint8_t> t(3000);
        for (auto &c : t) c = (uint8_t)(lcg() % 2);
        const uint64_t mx = 5, n = t.size();
        auto sa = build_in_memory_with_opts<uint8_t, uint32_t>(t, Opts{}.with_max_context(mx));
        std::vector<uint64_t> want(n);
        std::iota(want.begin(), want.end(), 0);
        std::sort(want.begin(), want.end(), [&](uint64_t a, uint64_t b) {
            uint64_t la = n - a, lb = n - b, m = std::min(std::min(la, lb), mx), l = 0;
            while (l < m && t[a + l] == t[b + l]) ++l;
            if (l < m) return t[a + l] < t[b + l];
            return la != lb ? la < lb : a < b;
        });
        CHECK(widen(sa) == want);
    }
    // streaming build: emits the suffix array in order (ext_mem.rs:2784-2806)
    {
        std::vector<uint8_t> t(20000);
        for (auto &c : t) c = (uint8_t)(lcg() % 4);
        std::vector<uint64_t> got;
        build_in_memory_sample_sort(t, ExtMemOpts{}.with_subproblem_count(64), [&](uint64_t p) { got.push_back(p); });
        CHECK(got == brute(t));
    }
    // the first emit error aborts the build with that error (ext_mem.rs:2895-2915)
    {
        std::vector<uint8_t> t = bytes("banana");
        int calls = 0;
        auto err = try_build_in_memory_sample_sort<std::string>(
            t, ExtMemOpts{}, [&](uint64_t) -> std::optional<std::string> {
                if (++calls == 2) return std::string("stop");
                return std::nullopt;
            });
        CHECK(err.has_value());
        CHECK(err && err->kind == BuildError<std::string>::Kind::Emit);
        CHECK(err && err->emit_error && *err->emit_error == "stop");
        CHECK(calls == 2);
    }
    // ext-mem builder over a segmented, filtered position list
    {
        std::vector<uint8_t> t(30000);
        for (auto &c : t) c = (uint8_t)(lcg() % 4);
        std::vector<uint64_t> ends;
        for (uint64_t e = 0; e < t.size();) {
            e = std::min<uint64_t>(t.size(), e + 50 + lcg() % 400);
            ends.push_back(e);
        }
        auto lp = SegmentedText::from_ends(t.size(), ends);
        std::vector<uint64_t> pos;
        for (uint64_t p = 0; p < t.size(); ++p)
            if (t[p] != 3) pos.push_back(p);
        std::vector<uint64_t> got;
        build_ext_mem_for_positions_with(t, pos, lp, ExtMemOpts{}, [&](uint64_t p) { got.push_back(p); });
        // adjacent order under the segmented comparator (shorter first, then position)
        bool ok = got.size() == pos.size();
        for (size_t i = 1; ok && i < got.size(); ++i) {
            uint64_t a = got[i - 1], b = got[i], la = lp.lim_at(a), lb = lp.lim_at(b);
            uint64_t l = 0, m = std::min(la, lb);
            while (l < m && t[a + l] == t[b + l]) ++l;
            ok = l < m ? t[a + l] < t[b + l] : (la < lb || (la == lb && a < b));
        }
        std::vector<uint64_t> s = got;
        std::sort(s.begin(), s.end());
        CHECK(ok && s == pos);
    }
    // filter builder: the kept positions of a predicate, in bucket batches, equal
    // to the positions builder over the same positions (ext_mem.rs:472-547)
    {
        std::vector<uint8_t> t(40000);
        for (auto &c : t) c = (uint8_t)(lcg() % 6);
        auto keep = [&](uint64_t q) { return t[q] < 4; };
        std::vector<uint64_t> pos, a, b;
        for (uint64_t q = 0; q < t.size(); ++q)
            if (keep(q)) pos.push_back(q);
        build_ext_mem_for_positions_with(t, pos, PlainText(t.size()), ExtMemOpts{}, [&](uint64_t p) { a.push_back(p); });
        build_ext_mem_for_filter(t, keep, ExtMemOpts{}.with_batch_records(5000), [&](uint64_t p) { b.push_back(p); });
        CHECK(a.size() == pos.size() && a == b);
    }
    // batched builds: every text's own suffix array
    {
        std::vector<std::vector<uint8_t>> texts;
        for (int i = 0; i < 40; ++i) {
            std::vector<uint8_t> t(lcg() % 3000);
            for (auto &c : t) c = (uint8_t)(lcg() % 4);
            texts.push_back(std::move(t));
        }
        auto out = build_batch<uint8_t, uint32_t>(texts);
        for (size_t i = 0; i < texts.size(); ++i) CHECK(widen(out[i]) == brute(texts[i]));
    }
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
