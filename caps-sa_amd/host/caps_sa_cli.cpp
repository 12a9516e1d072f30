// caps_sa: command-line suffix-array builder with the arguments, output file
// format, stderr report and exit codes of caps-sa 0.7.0's example tool
// (examples/caps_sa.rs:1-194):
//
//   caps_sa <input> <output> [--ext-mem | --in-mem-ss] [--subproblem-count N] [--threads N]
//
// <input> is read as raw bytes (u8 symbols).  <output> receives the suffix
// array as packed little-endian u32[] (in-memory path, n <= 2^32) or u64[]
// (sample-sort paths, or n > 2^32), examples/caps_sa.rs:117-190.  The build
// runs on the GPU through the C ABI; the sample-sort paths stream their
// batches of positions straight into the output file from the emit callback,
// inside the timed build as the crate's tool does.
//
// Exit codes: 0 success or --help; 2 wrong positional count
// (examples/caps_sa.rs:77-83); 101 a malformed or missing integer value (the
// crate's parse().expect panics); 1 an I/O or build error (main returning
// io::Error).  --threads sizes the crate's Rayon pool; the GPU build has no
// host worker pool, so the value is checked and otherwise unused.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "caps_sa.hpp"
#include "formats.hpp"

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a) { return std::chrono::duration<double>(Clock::now() - a).count(); }

struct Args {
    std::string input, output;
    bool ext_mem = false, in_mem_ss = false;
    uint64_t subproblem_count = 0;
    bool threads_set = false;
    uint64_t threads = 0;
};

[[noreturn]] void panic(const std::string &msg) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(101);
}

// usize::from_str: an optional '+', then decimal digits, no overflow.
uint64_t parse_usize(const char *flag, int argc, char **argv, int i) {
    if (i + 1 >= argc) panic(std::string("error: ") + flag + " needs a value");
    const char *s = argv[i + 1];
    if (*s == '+') ++s;
    if (!*s) panic(std::string(flag) + " expects a positive integer");
    uint64_t v = 0;
    for (; *s; ++s) {
        if (*s < '0' || *s > '9') panic(std::string(flag) + " expects a positive integer");
        const uint64_t d = (uint64_t)(*s - '0');
        if (v > (UINT64_MAX - d) / 10) panic(std::string(flag) + " expects a positive integer");
        v = v * 10 + d;
    }
    return v;
}

Args parse_args(int argc, char **argv) {
    Args a;
    std::vector<std::string> positional;
    for (int i = 1; i < argc;) {
        const std::string s = argv[i];
        if (s == "--ext-mem") {
            a.ext_mem = true;
            i += 1;
        } else if (s == "--in-mem-ss") {
            a.in_mem_ss = true;
            i += 1;
        } else if (s == "--subproblem-count") {
            a.subproblem_count = parse_usize("--subproblem-count", argc, argv, i);
            i += 2;
        } else if (s == "--threads") {
            a.threads = parse_usize("--threads", argc, argv, i);
            a.threads_set = true;
            i += 2;
        } else if (s == "--help" || s == "-h") {
            std::fprintf(stderr,
                         "usage: caps_sa <input> <output> [--ext-mem | --in-mem-ss] [--subproblem-count N] "
                         "[--threads N]\n");
            std::exit(0);
        } else {
            positional.push_back(s);
            i += 1;
        }
    }
    if (positional.size() != 2) {
        std::fprintf(stderr, "error: expected 2 positional args (input, output), got %zu\n", positional.size());
        std::exit(2);
    }
    a.input = positional[0];
    a.output = positional[1];
    return a;
}

// Emit target of the streaming builds: every batch goes to the output file.
struct StreamOut {
    caps_sa::formats::LeWriter *w;
    uint64_t count = 0;
    std::string error;
};

int emit_to_file(void *ctx, const uint64_t *pos, uint64_t count) {
    auto *o = static_cast<StreamOut *>(ctx);
    try {
        o->w->put_array<uint64_t>(pos, count);
    } catch (const std::exception &e) {
        o->error = e.what();
        return 1;
    }
    o->count += count;
    return 0;
}

template <class I>
double build_in_memory_to(const std::vector<uint8_t> &text, const std::string &path, const char *label,
                          double &write_s) {
    const auto t0 = Clock::now();
    const std::vector<I> sa = caps_sa::build_in_memory<uint8_t, I>(text);
    const double build_s = secs(t0);
    std::fprintf(stderr, "build: mode=in-mem(%s) n=%zu entries in %.3fs\n", label, sa.size(), build_s);
    const auto t1 = Clock::now();
    caps_sa::formats::write_sa<I>(path, sa.data(), sa.size());
    write_s = secs(t1);
    return build_s;
}

int run(const Args &a) {
    const auto t_read = Clock::now();
    const std::vector<uint8_t> text = caps_sa::formats::read_file(a.input);
    std::fprintf(stderr, "read: n=%zu bytes in %.3fs\n", text.size(), secs(t_read));

    const bool use_u32 = text.size() <= (1ull << 32) && !a.ext_mem && !a.in_mem_ss;
    double write_s = 0.0;
    if (use_u32) {
        build_in_memory_to<uint32_t>(text, a.output, "u32", write_s);
    } else if (a.ext_mem || a.in_mem_ss) {
        caps_sa::formats::LeWriter w(a.output);
        StreamOut out{&w};
        caps_opts_t o;
        caps_opts_default(&o);
        o.subproblem_count = a.subproblem_count;
        const caps_limits_t lim = caps_sa::PlainText(text.size()).c_limits();
        const auto t0 = Clock::now();
        // build_ext_mem and build_in_memory_sample_sort over every position
        // are one GPU build (ext_mem.rs:289-356, 717-723).
        const int rc = caps_build_in_memory_sample_sort(text.data(), text.size(), 1, nullptr, text.size(), &lim, &o,
                                                        &emit_to_file, &out);
        const double build_s = secs(t0);
        if (rc == CAPS_E_EMIT) throw std::runtime_error(out.error);
        if (rc != CAPS_OK) throw std::runtime_error(caps_last_error());
        w.close();
        std::fprintf(stderr, "build: mode=%s(u64,stream) n=%llu entries in %.3fs\n",
                     a.ext_mem ? "ext-mem" : "in-mem-ss", (unsigned long long)out.count, build_s);
    } else {
        build_in_memory_to<uint64_t>(text, a.output, "u64", write_s);
    }
    std::fprintf(stderr, "write: %.3fs\n", write_s);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    const Args a = parse_args(argc, argv);
    try {
        return run(a);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Error: %s\n", e.what());
        return 1;
    }
}
