// Tree and interaction lists built on the GPU (bit-exact with the reference's Tree/Traverser).
#pragma once
#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "dbuf.h"
#include "device.h"
#include "fmm3dpp.h"
#include "structure.h"

namespace fmm3dpp {

struct LevelTable { int ncol; int64_t in_off, out_off; };

struct GpuStructure {
    // sizes (host mirror)
    int64_t n = 0;
    int ntree = 0, ncells = 0, nleaf = 0, maxd = 0, nl = 0, nimg = 0, ntar = 0;
    int sort_levels = 0;   // key levels the next build sorts first (the last tree's depth + 2; 0: all)
    int64_t n_rows = 0, n_entries = 0, lost = 0;
    int max_level_cols = 1;
    std::vector<int32_t> level_start;          // tree levels 0..maxd, plus end
    std::vector<LevelTable> m2m_levels;        // deepest first
    std::vector<LevelTable> l2l_levels;        // root first
    std::vector<std::array<int32_t, 2>> img_m2m, img_l2l;  // (parent, child) image-cell pairs
    double ms_build = 0;

    // bodies in tree order
    DBuf x, y, z, perm;
    // cells in BFS order followed by the image cells
    DBuf prefix, cx, cy, cz, cr, depth, octant, coff, cnum, boff, bnum;
    size_t cell_cap = 0;
    // leaves in body order
    DBuf leaf_cell, leaf_depth, iota;
    // operator-apply column tables
    DBuf m2m_in, m2m_out, l2l_in, l2l_out, img_a, img_b;
    // M2L Hadamard tables (sources = targets = non-leaf cells, then image cells)
    DBuf nl_idx, nl_cells, child_slots, tar_scale, slot_src_T, img_src_T, img_tar, img_scale;
    // staging tables of the 8x8 Hadamard over its targets (the non-leaf cells in mixed precision, all
    // M2L targets in float64): distinct sources and rows per chunk of targets
    DBuf had_uniq, had_cnt, had_loc;
    HadTables had() const { return {had_uniq.as<int32_t>(), had_cnt.as<int32_t>(), had_loc.as<int32_t>()}; }
    // P2P rows (target, source, offset index) and per-leaf CSR (source << 5 | offset index)
    DBuf rows_tc, rows_sc, rows_oi, p2p_off, p2p_ent;
    size_t rows_cap = 0;
    // temporaries
    DBuf key, key_s, idx0, perm1, tkey, tkey_s, ldep, head, rank, nchild, cstart, scan32, packed, scan64, lvl_nl,
        ekey, eval, ekey_s, eval_s, cnt64, flags, lvl_info, small, temp;
    DBuf temp_lists;   // rocPRIM scratch of build_lists, apart from build_tree's `temp` (another stream)
    DBuf f_tc[2], f_sc[2], f_tcp[2], f_scp[2], f_oi[2];
    // pinned host staging for the small host-built tables of a build: their copies stay in stream
    // order without a host wait; the region is reused from the next build on (builds end synced)
    struct PinnedStage {
        char* p = nullptr;
        size_t cap = 0, off = 0;
        PinnedStage() = default;
        PinnedStage(const PinnedStage&) = delete;
        PinnedStage& operator=(const PinnedStage&) = delete;
        ~PinnedStage() { if (p) (void)hipHostFree(p); }
    } stage;
    // mapped pinned words for the small counts a build reads back: a one-wave kernel queued behind the
    // build's kernels copies them there and then writes a sequence number the host polls for
    struct Mailbox {
        uint32_t* h = nullptr;   // host view; word kWords is the sequence number
        uint32_t* d = nullptr;   // device view of the same memory
        uint32_t seq = 0;
        static constexpr int kWords = 63;
        Mailbox() = default;
        Mailbox(const Mailbox&) = delete;
        Mailbox& operator=(const Mailbox&) = delete;
        ~Mailbox() { if (h) (void)hipHostFree(h); }
    } mailbox;
    size_t f_cap[2] = {0, 0};
    // Recorded on the build stream once build_tree has queued its last kernel and copy.  A list
    // build on another stream waits for it (it reads the cells, leaves and ranks of the tree), and
    // the next build_tree waits for it on the host before it reuses the pinned staging region,
    // whose copies the previous build left in stream order.
    struct TreeDone {
        hipEvent_t ev = nullptr;
        bool recorded = false;
        TreeDone() = default;
        TreeDone(const TreeDone&) = delete;
        TreeDone& operator=(const TreeDone&) = delete;
        ~TreeDone() { if (ev) (void)hipEventDestroy(ev); }
    } tree_done;

    // build = build_tree (keys, sort, levels, leaves, operator-apply and M2L tables) + build_lists (the
    // P2P lists); build_lists reads only what build_tree wrote and may use another stream.
    int build(const fmm3dpp_config& cfg, int64_t n, const double* dx, const double* dy, const double* dz,
              hipStream_t st, std::string& err);
    int build_tree(const fmm3dpp_config& cfg, int64_t n, const double* dx, const double* dy, const double* dz,
                   hipStream_t st, std::string& err);
    int build_lists(const fmm3dpp_config& cfg, hipStream_t st, std::string& err);
    std::chrono::steady_clock::time_point t_build;
    // Host copies for the structure queries of the C-ABI.
    int export_structure(const fmm3dpp_config& cfg, Structure& out, hipStream_t st, std::string& err);
};

}  // namespace fmm3dpp
