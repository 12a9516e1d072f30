// On-disk formats of caps-sa 0.7.0, for tools that exchange files with it:
//   - suffix-array files: packed little-endian u32[] or u64[], no header
//     (examples/caps_sa.rs:117-190; the in-memory path writes u32 when the
//     text has at most 2^32 symbols, the sample-sort paths always u64);
//   - SaLcp records: (pos, lcp), each little-endian, 8 bytes for u32 and 16
//     bytes for u64, pos first (ext_bucket.rs:65-106).  A sorted run is
//     stored with its first record's lcp set to 0, as the crate's
//     add_soa_reset_first_lcp / add_slice_reset_first_lcp do
//     (ext_bucket.rs:203-229).
// Header-only, plain C++17; independent of the GPU library.
#pragma once
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace caps_sa::formats {

template <class I>
inline void put_le(uint8_t *out, I v) {
    static_assert(std::is_unsigned_v<I>);
    for (size_t b = 0; b < sizeof(I); ++b) out[b] = (uint8_t)(v >> (8 * b));
}
template <class I>
inline I get_le(const uint8_t *in) {
    static_assert(std::is_unsigned_v<I>);
    I v = 0;
    for (size_t b = 0; b < sizeof(I); ++b) v |= (I)in[b] << (8 * b);
    return v;
}

// SaLcp<I> (ext_bucket.rs:65-106): I is uint32_t (8-byte record) or uint64_t
// (16-byte record).
template <class I>
struct SaLcp {
    static_assert(std::is_same_v<I, uint32_t> || std::is_same_v<I, uint64_t>, "SaLcp holds u32 or u64");
    static constexpr size_t SIZE = 2 * sizeof(I);
    I pos = 0;
    I lcp = 0;
    void write_to(uint8_t *out) const {
        put_le<I>(out, pos);
        put_le<I>(out + sizeof(I), lcp);
    }
    static SaLcp read_from(const uint8_t *in) { return SaLcp{get_le<I>(in), get_le<I>(in + sizeof(I))}; }
    bool operator==(const SaLcp &o) const { return pos == o.pos && lcp == o.lcp; }
};

// Buffered writer of little-endian integers to a file; write errors surface
// as std::runtime_error (the crate's io::Error).
class LeWriter {
   public:
    explicit LeWriter(const std::string &path) : path_(path), f_(std::fopen(path.c_str(), "wb")) {
        if (!f_) throw std::runtime_error("cannot create " + path);
        buf_.reserve(kBuf);
    }
    LeWriter(const LeWriter &) = delete;
    LeWriter &operator=(const LeWriter &) = delete;
    ~LeWriter() {
        if (f_) std::fclose(f_);
    }
    template <class I>
    void put(I v) {
        uint8_t b[sizeof(I)];
        put_le<I>(b, v);
        bytes(b, sizeof(I));
    }
    template <class I>
    void put_array(const I *v, size_t n) {
        if constexpr (__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__) {
            bytes(reinterpret_cast<const uint8_t *>(v), n * sizeof(I));
        } else {
            for (size_t i = 0; i < n; ++i) put<I>(v[i]);
        }
    }
    void bytes(const uint8_t *p, size_t len) {
        if (buf_.size() + len > kBuf) flush();
        if (len >= kBuf) {
            if (std::fwrite(p, 1, len, f_) != len) throw std::runtime_error("write failed: " + path_);
            return;
        }
        buf_.insert(buf_.end(), p, p + len);
    }
    void flush() {
        if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
            throw std::runtime_error("write failed: " + path_);
        buf_.clear();
    }
    void close() {
        flush();
        const int rc = std::fclose(f_);
        f_ = nullptr;
        if (rc != 0) throw std::runtime_error("close failed: " + path_);
    }

   private:
    static constexpr size_t kBuf = 1 << 20;
    std::string path_;
    std::FILE *f_;
    std::vector<uint8_t> buf_;
};

inline std::vector<uint8_t> read_file(const std::string &path) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint8_t> data;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) data.insert(data.end(), chunk, chunk + got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) throw std::runtime_error("read failed: " + path);
    return data;
}

// Suffix-array files (examples/caps_sa.rs:122-190).
template <class I>
void write_sa(const std::string &path, const I *sa, size_t n) {
    LeWriter w(path);
    w.put_array<I>(sa, n);
    w.close();
}
template <class I>
std::vector<I> read_sa(const std::string &path) {
    const std::vector<uint8_t> data = read_file(path);
    if (data.size() % sizeof(I)) throw std::runtime_error(path + ": size is not a multiple of the index width");
    std::vector<I> sa(data.size() / sizeof(I));
    for (size_t i = 0; i < sa.size(); ++i) sa[i] = get_le<I>(data.data() + i * sizeof(I));
    return sa;
}

// One sorted run of SaLcp records: lcp[0] is stored as 0 whatever it holds
// (ext_bucket.rs:203-218).  Appends to `w`.
template <class I>
void write_salcp_run(LeWriter &w, const I *pos, const I *lcp, size_t n) {
    uint8_t rec[SaLcp<I>::SIZE];
    for (size_t i = 0; i < n; ++i) {
        SaLcp<I>{pos[i], i == 0 ? I(0) : lcp[i]}.write_to(rec);
        w.bytes(rec, sizeof rec);
    }
}
// Every record of a SaLcp file, as two arrays (load_all_soa,
// ext_bucket.rs:135).
template <class I>
void read_salcp(const std::string &path, std::vector<I> &pos, std::vector<I> &lcp) {
    const std::vector<uint8_t> data = read_file(path);
    if (data.size() % SaLcp<I>::SIZE) throw std::runtime_error(path + ": size is not a multiple of the record size");
    const size_t n = data.size() / SaLcp<I>::SIZE;
    pos.resize(n);
    lcp.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const SaLcp<I> r = SaLcp<I>::read_from(data.data() + i * SaLcp<I>::SIZE);
        pos[i] = r.pos;
        lcp[i] = r.lcp;
    }
}

}  // namespace caps_sa::formats
