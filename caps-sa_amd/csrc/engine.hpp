// Host-side state of the suffix-array build: grow-only device buffers kept
// across calls so repeated builds do not reallocate.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace caps {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t bytes);
    void release();
};

struct Geometry {
    KeyGeom g;                   // seed keys
    uint32_t pbits = 0;          // packed code text and refinement chunks: power of two >= g.bits
    const void *code = nullptr;  // device code map or nullptr (identity)
    bool remap = false;
    bool reference = false;
    uint32_t distinct = 0;
};

struct Workspace {
    uint32_t n_cu = 0;  // compute units of the current device (sizes persistent grids)
    // inputs staged from host pointers
    DevBuf text, hpos, out, outlcp, pos_copy;
    // limits
    DevBuf ends, dir;
    std::vector<uint64_t> h_ends;
    // key geometry
    DevBuf bitmap, code, packed, seedpk;
    std::vector<uint8_t> h_code8;
    std::vector<uint16_t> h_code16;
    // sample and splitters
    DevBuf skeys, skeys2, sids, sids2, spl, stiles, srange;
    std::vector<Tile> h_tiles;
    Range h_range;
    // classification / scatter
    DevBuf bid, hist, hblk, off, cursor, keys, keys2, idx2, brec;
    // bucket batches: per-record bucket ids, a batch of output; kept-position
    // bitmap, its count
    DevBuf sinkbuf, keep, keep_count, poscheck;
    // tiles
    DevBuf tcnt, toff, toff2, tiles, big, counts, recs;
    // refinement
    DevBuf glist[kGroupClasses], mid, lcp, lcp2, first, jobs, chunks;
    // run memo: raw and sorted run bounds, directory, run count
    DevBuf run_s, run_e, run_ss, run_se, run_dir, run_count;
    std::vector<std::pair<uint64_t, uint64_t>> h_runs;
    std::vector<MergeJob> h_jobs;
    // pinned staging for streamed batches (kept until the process ends)
    void *h_stage = nullptr;
    size_t h_stage_bytes = 0;
    // standalone merges
    DevBuf m_a, m_al, m_b, m_bl;
    // router
    DevBuf r_a, r_b, r_inv, r_lo, r_hi, r_knots, r_radix, r_labels, r_stats, r_x, r_y, r_w, r_nib, r_tmp;
    std::vector<float> h_mlp;
    // graph builds: the capture stream and the instantiated graph of each
    // captured phase (tile sort, refinement, giant groups), updated in place
    // by the next build that captures the same kernels
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
};

Workspace &workspace();

}  // namespace caps
