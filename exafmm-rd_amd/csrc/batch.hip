// Batches of independent periodic systems (no counterpart in the reference, which solves one
// system per LaplaceFMM): every system gets its own tree and lists (the single-system GPU build),
// then the systems of equal P are merged into one set of tables with offset indices, so each FMM
// stage is one launch over all systems of that P.  The whole stage sequence of an evaluation can be
// captured once in a hipGraph and replayed.  An autotuner picks the leaf size (ncrit) per P from
// timed evaluations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dbuf.h"
#include "device.h"
#include "device_ops.h"
#include "fmm3dpp.h"
#include "gpu_structure.h"
#include "structure.h"

using namespace fmm3dpp;

namespace {

// Blocks per system of the dipole-moment reduction: the same partition and summation order as the
// single-system plan, so a system's moments are bit-identical to its own plan's.
constexpr int kDipoleBlocks = 256;
// Host threads (and streams) for the per-system tree builds.
constexpr int kBuildThreads = 8;

// One merged group: the systems with the same P.
struct Group {
    int P = 0;
    std::vector<int> sys;         // system ids, in merge order
    Operators ops;
    bool ops_uploaded = false;
    // merged structure sizes
    int64_t n = 0;
    int nleaf = 0, ncells = 0, nl = 0, nimg = 0, ntar = 0, max_level_cols = 1;
    std::vector<LevelTable> m2m_levels, l2l_levels;   // merged per tree level
    std::vector<std::array<int64_t, 2>> img_m2m_lv, img_l2l_lv;   // (offset, count) into img tables per image level
    // merged device tables
    DBuf x, y, z, perm, cx, cy, cz, cr, depth, boff, bnum, leaf_cell, leaf_depth, iota, p2p_off, p2p_ent, m2m_in,
        m2m_out, l2l_in, l2l_out, img_a, img_b, child_slots, tar_scale, slot_src_T, img_src_T, img_tar, seg_off;
    DBuf roots;   // merged index of every system's root cell
    DBuf had_uniq, had_cnt, had_loc;   // Hadamard staging tables of the merged targets
    HadTables had() const { return {had_uniq.as<int32_t>(), had_cnt.as<int32_t>(), had_loc.as<int32_t>()}; }
    DeviceOperators dop;
    M2LTiles up_tiles, down_tiles;   // transform tiles of the tree parents (0, 0: measured defaults)
    // work buffers
    DBuf q, pot, gx, gy, gz, pcleaf, pcleaf2, qeq, pck, Qk, Pk, Qk_img, Pk_img, rel, m2m_part, partial, moments;
    DBuf qcell;   // per-cell charges of the mixed path's monopole constraint
};

}  // namespace

struct fmm3dpp_batch {
    std::vector<fmm3dpp_config> cfg;   // per system (P, ncrit, centre); the shared fields are equal
    int device = 0;
    std::string err;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> build_streams;   // one per build thread
    std::vector<hipStream_t> group_streams;   // groups 1.. run beside group 0 (fork / join events)
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::unique_ptr<GpuStructure>> gs;
    std::vector<int64_t> counts, offsets;   // bodies per system and their offsets in the caller's arrays
    std::vector<std::unique_ptr<Group>> groups;
    bool particles = false;
    DBuf pos_x, pos_y, pos_z;   // copy of the positions of the last set_particles (autotune rebuilds from it)
    const double *px = nullptr, *py = nullptr, *pz = nullptr;
    std::map<std::string, double> ms;
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    const void* graph_key[3] = {nullptr, nullptr, nullptr};

    bool mixed() const { return cfg[0].precision == FMM3DPP_PREC_MIXED; }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    ~fmm3dpp_batch() {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        groups.clear();   // device buffers, freed before the streams go away
        gs.clear();
        for (hipStream_t bs : build_streams) (void)hipStreamDestroy(bs);
        for (hipStream_t gs_ : group_streams) (void)hipStreamDestroy(gs_);
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define BFAIL(code, msg)   \
    do {                   \
        b->err = (msg);    \
        return (code);     \
    } while (0)
#define BHIP(call)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) BFAIL(FMM3DPP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int upload_group_operators(fmm3dpp_batch* b, Group& g) {
    BHIP(g.dop.upload(g.ops, b->mixed(), true, b->stream));
    g.ops_uploaded = true;
    return FMM3DPP_OK;
}

// dst[r * dst_stride + c] = map(src[r * src_stride + c]) for r < rows, c < cols, where map keeps
// negative values (absent entries), adds add_lo below lim and add_hi from lim on.  Used to merge the
// index tables of several systems: cell, leaf, target and body indices gain their system's offset.
template <class T>
__global__ __launch_bounds__(256) void k_remap(const T* __restrict__ src, int64_t rows, int64_t cols,
                                               int64_t src_stride, T* __restrict__ dst, int64_t dst_stride, T lim,
                                               T add_lo, T add_hi) {
    const int64_t total = rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
        const int64_t r = i / cols, c = i - r * cols;
        const T v = src[r * src_stride + c];
        dst[r * dst_stride + c] = v < 0 ? v : (v < lim ? v + add_lo : v + add_hi);
    }
}

__global__ __launch_bounds__(256) void k_iota32(int32_t* __restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = (int32_t)i;
}

template <class T>
hipError_t remap(const T* src, int64_t rows, int64_t cols, int64_t src_stride, T* dst, int64_t dst_stride, T lim,
                 T add_lo, T add_hi, hipStream_t st) {
    const int64_t total = rows * cols;
    if (total <= 0) return hipSuccess;
    if (!src || !dst || cols > src_stride || cols > dst_stride) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 4096);
    k_remap<T><<<blocks, 256, 0, st>>>(src, rows, cols, src_stride, dst, dst_stride, lim, add_lo, add_hi);
    return hipGetLastError();
}

// Shift of every non-negative value by add.
template <class T>
hipError_t shift(const T* src, int64_t n, T* dst, T add, hipStream_t st) {
    return remap<T>(src, 1, n, n, dst, n, std::numeric_limits<T>::max(), add, add, st);
}

hipError_t copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    if (!dst || !src) return hipErrorInvalidValue;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
}

// The single-system GPU builds of systems `ids`, spread over the batch's build threads, each with
// its own stream (a build synchronises its stream at every tree level and frontier step, so several
// in flight keep the GPU busy).  Returns after every build has finished.
int build_systems(fmm3dpp_batch* b, const std::vector<int>& ids) {
    BHIP(hipStreamSynchronize(b->stream));   // positions copied on the batch stream
    const int nthr = std::min<int>((int)b->build_streams.size(), (int)ids.size());
    std::atomic<int> next{0};
    std::vector<int> rcs(nthr, FMM3DPP_OK);
    std::vector<std::string> errs(nthr);
    auto work = [&](int w) {
        if (hipSetDevice(b->device) != hipSuccess) { rcs[w] = FMM3DPP_ERR_DEVICE; errs[w] = "hipSetDevice failed"; return; }
        for (int i = next++; i < (int)ids.size(); i = next++) {
            const int k = ids[i];
            const int rc = b->gs[k]->build(b->cfg[k], b->counts[k], b->px + b->offsets[k], b->py + b->offsets[k],
                                           b->pz + b->offsets[k], b->build_streams[w], errs[w]);
            if (rc) { rcs[w] = rc; errs[w] = "system " + std::to_string(k) + ": " + errs[w]; return; }
        }
    };
    std::vector<std::thread> threads;
    for (int w = 1; w < nthr; w++) threads.emplace_back(work, w);
    work(0);
    for (auto& t : threads) t.join();
    for (int w = 0; w < nthr; w++)
        if (rcs[w]) BFAIL(rcs[w], errs[w]);
    return FMM3DPP_OK;
}

// Builds the trees of the group's systems (one GPU build each) and merges their tables on the GPU:
// cells, leaves and bodies are concatenated in system order; the M2L targets are the regular targets
// of every system, then the image targets of every system; the operator-apply tables of tree level l
// (and image level m) of all systems form one table per level.
int build_group(fmm3dpp_batch* b, Group& g) {
    hipStream_t s = b->stream;
    const int images = b->cfg[0].images;
    int rc = build_systems(b, g.sys);
    if (rc) return rc;
    // merged sizes and offsets (host mirrors of the per-system sizes)
    int64_t nb = 0, nent = 0, ncell64 = 0;
    int nleaf = 0, maxd = 0, nl = 0, nimg = 0;
    for (int k : g.sys) {
        const GpuStructure& G = *b->gs[k];
        nb += G.n; nent += G.n_entries; ncell64 += G.ncells; nleaf += G.nleaf; nl += G.nl; nimg += G.nimg;
        maxd = std::max(maxd, G.maxd);
    }
    if (ncell64 > (std::numeric_limits<int32_t>::max() >> 5))   // P2P entries hold cell << 5
        BFAIL(FMM3DPP_ERR_ARG, "batch: too many cells in one P group");
    const int ncell = (int)ncell64, ntar = nl + nimg;
    std::vector<int> lvl_cols(maxd, 0);
    for (int k : g.sys)
        for (int l = 0; l < b->gs[k]->maxd; l++) lvl_cols[l] += b->gs[k]->l2l_levels[l].ncol;
    g.l2l_levels.clear();
    g.m2m_levels.clear();
    g.max_level_cols = 1;
    int64_t tab = 0, col = 0;
    for (int l = 0; l < maxd; l++) {
        g.l2l_levels.push_back({lvl_cols[l], tab, col});
        g.max_level_cols = std::max(g.max_level_cols, lvl_cols[l]);
        tab += 8 * (int64_t)lvl_cols[l];
        col += lvl_cols[l];
    }
    for (int l = maxd - 1; l >= 0; l--) g.m2m_levels.push_back(g.l2l_levels[l]);
    const size_t nc = (size_t)std::max(ncell, 1), nlf = (size_t)std::max(nleaf, 1), nt = (size_t)std::max(ntar, 1);
    BHIP(g.x.ensure(nb * 8)); BHIP(g.y.ensure(nb * 8)); BHIP(g.z.ensure(nb * 8)); BHIP(g.perm.ensure(nb * 4));
    BHIP(g.cx.ensure(nc * 8)); BHIP(g.cy.ensure(nc * 8)); BHIP(g.cz.ensure(nc * 8)); BHIP(g.cr.ensure(nc * 8));
    BHIP(g.depth.ensure(nc * 4)); BHIP(g.boff.ensure(nc * 4)); BHIP(g.bnum.ensure(nc * 4));
    BHIP(g.leaf_cell.ensure(nlf * 4)); BHIP(g.leaf_depth.ensure(nlf * 4)); BHIP(g.p2p_off.ensure((nlf + 1) * 8));
    BHIP(g.p2p_ent.ensure((size_t)std::max<int64_t>(nent, 1) * 4));
    BHIP(g.m2m_in.ensure((size_t)std::max<int64_t>(tab, 1) * 4)); BHIP(g.l2l_in.ensure((size_t)std::max<int64_t>(tab, 1) * 4));
    BHIP(g.l2l_out.ensure((size_t)std::max<int64_t>(tab, 1) * 4)); BHIP(g.m2m_out.ensure((size_t)std::max<int64_t>(col, 1) * 4));
    BHIP(g.child_slots.ensure(nt * 32)); BHIP(g.tar_scale.ensure(nt * 8));
    BHIP(g.slot_src_T.ensure(26 * nt * 4)); BHIP(g.img_src_T.ensure(26 * nt * 4));
    const int32_t kMax32 = std::numeric_limits<int32_t>::max();
    std::vector<int> lvl_fill(maxd, 0);
    std::vector<std::vector<std::array<int32_t, 2>>> img_m2m_lv(std::max(images, 1)), img_l2l_lv(std::max(images, 1));
    std::vector<int32_t> img_tar;
    std::vector<int64_t> seg(1, 0);
    std::vector<int32_t> roots;
    int64_t bo = 0, eo = 0;
    int co = 0, lo = 0, tl = 0, ti = nl;
    for (int k : g.sys) {
        GpuStructure& G = *b->gs[k];
        BHIP(copy_d2d(g.x.as<double>() + bo, G.x.p, G.n * 8, s));
        BHIP(copy_d2d(g.y.as<double>() + bo, G.y.p, G.n * 8, s));
        BHIP(copy_d2d(g.z.as<double>() + bo, G.z.p, G.n * 8, s));
        BHIP(shift<int32_t>(G.perm.as<int32_t>(), G.n, g.perm.as<int32_t>() + bo, (int32_t)b->offsets[k], s));
        BHIP(copy_d2d(g.cx.as<double>() + co, G.cx.p, (size_t)G.ncells * 8, s));
        BHIP(copy_d2d(g.cy.as<double>() + co, G.cy.p, (size_t)G.ncells * 8, s));
        BHIP(copy_d2d(g.cz.as<double>() + co, G.cz.p, (size_t)G.ncells * 8, s));
        BHIP(copy_d2d(g.cr.as<double>() + co, G.cr.p, (size_t)G.ncells * 8, s));
        BHIP(copy_d2d(g.depth.as<int32_t>() + co, G.depth.p, (size_t)G.ncells * 4, s));
        BHIP(copy_d2d(g.bnum.as<int32_t>() + co, G.bnum.p, (size_t)G.ncells * 4, s));
        BHIP(shift<int32_t>(G.boff.as<int32_t>(), G.ncells, g.boff.as<int32_t>() + co, (int32_t)bo, s));
        BHIP(shift<int32_t>(G.leaf_cell.as<int32_t>(), G.nleaf, g.leaf_cell.as<int32_t>() + lo, co, s));
        BHIP(copy_d2d(g.leaf_depth.as<int32_t>() + lo, G.leaf_depth.p, (size_t)G.nleaf * 4, s));
        BHIP(shift<int64_t>(G.p2p_off.as<int64_t>(), G.nleaf + 1, g.p2p_off.as<int64_t>() + lo, eo, s));
        // P2P entries are (source cell << 5 | offset index): the cell shift is co << 5
        BHIP(shift<int32_t>(G.p2p_ent.as<int32_t>(), G.n_entries, g.p2p_ent.as<int32_t>() + eo, co << 5, s));
        for (int l = 0; l < G.maxd; l++) {
            const LevelTable& sl = G.l2l_levels[l];
            const LevelTable& dl = g.l2l_levels[l];
            const int64_t c0 = lvl_fill[l];
            BHIP(remap<int32_t>(G.m2m_in.as<int32_t>() + sl.in_off, 8, sl.ncol, sl.ncol,
                                g.m2m_in.as<int32_t>() + dl.in_off + c0, dl.ncol, kMax32, co, co, s));
            BHIP(remap<int32_t>(G.l2l_in.as<int32_t>() + sl.in_off, 8, sl.ncol, sl.ncol,
                                g.l2l_in.as<int32_t>() + dl.in_off + c0, dl.ncol, kMax32, co, co, s));
            BHIP(remap<int32_t>(G.l2l_out.as<int32_t>() + sl.in_off, 8, sl.ncol, sl.ncol,
                                g.l2l_out.as<int32_t>() + dl.in_off + c0, dl.ncol, kMax32, co, co, s));
            BHIP(shift<int32_t>(G.m2m_out.as<int32_t>() + sl.out_off, sl.ncol, g.m2m_out.as<int32_t>() + dl.out_off + c0,
                                co, s));
            lvl_fill[l] += sl.ncol;
        }
        // M2L targets: regular t < G.nl -> tl + t, image t >= G.nl -> ti + (t - G.nl)
        const int32_t hi_add = ti - G.nl;
        if (G.nl > 0) {
            BHIP(remap<int32_t>(G.child_slots.as<int32_t>(), G.nl, 8, 8, g.child_slots.as<int32_t>() + (size_t)tl * 8, 8,
                                kMax32, co, co, s));
            BHIP(copy_d2d(g.tar_scale.as<double>() + tl, G.tar_scale.p, (size_t)G.nl * 8, s));
            BHIP(remap<int32_t>(G.slot_src_T.as<int32_t>(), 26, G.nl, G.ntar, g.slot_src_T.as<int32_t>() + tl, ntar,
                                G.nl, tl, hi_add, s));
            BHIP(remap<int32_t>(G.img_src_T.as<int32_t>(), 26, G.nl, G.ntar, g.img_src_T.as<int32_t>() + tl, ntar,
                                G.nl, tl, hi_add, s));
        }
        if (G.nimg > 0) {
            BHIP(remap<int32_t>(G.child_slots.as<int32_t>() + (size_t)G.nl * 8, G.nimg, 8, 8,
                                g.child_slots.as<int32_t>() + (size_t)ti * 8, 8, kMax32, co, co, s));
            BHIP(copy_d2d(g.tar_scale.as<double>() + ti, G.tar_scale.as<double>() + G.nl, (size_t)G.nimg * 8, s));
            BHIP(remap<int32_t>(G.slot_src_T.as<int32_t>() + G.nl, 26, G.nimg, G.ntar, g.slot_src_T.as<int32_t>() + ti,
                                ntar, G.nl, tl, hi_add, s));
            BHIP(remap<int32_t>(G.img_src_T.as<int32_t>() + G.nl, 26, G.nimg, G.ntar, g.img_src_T.as<int32_t>() + ti,
                                ntar, G.nl, tl, hi_add, s));
            for (int m = 0; m < G.nimg; m++) img_tar.push_back(ti + m);
        }
        // image levels: (parent, child) cell pairs, level i of every system in one launch
        for (size_t i = 0; i < G.img_m2m.size(); i++) img_m2m_lv[i].push_back({G.img_m2m[i][0] + co, G.img_m2m[i][1] + co});
        for (size_t i = 0; i < G.img_l2l.size(); i++) img_l2l_lv[i].push_back({G.img_l2l[i][0] + co, G.img_l2l[i][1] + co});
        roots.push_back(co);
        bo += G.n; eo += G.n_entries; co += G.ncells; lo += G.nleaf; tl += G.nl; ti += G.nimg;
        seg.push_back(bo);
    }
    g.n = nb; g.nleaf = nleaf; g.ncells = ncell; g.nl = nl; g.nimg = nimg; g.ntar = ntar;
    std::vector<int32_t> ia, ib;
    g.img_m2m_lv.clear();
    g.img_l2l_lv.clear();
    for (auto& lv : img_m2m_lv) {
        if (lv.empty()) continue;
        g.img_m2m_lv.push_back({(int64_t)ia.size(), (int64_t)lv.size()});
        for (auto& pr : lv) { ia.push_back(pr[0]); ib.push_back(pr[1]); }
    }
    for (auto& lv : img_l2l_lv) {
        if (lv.empty()) continue;
        g.img_l2l_lv.push_back({(int64_t)ia.size(), (int64_t)lv.size()});
        for (auto& pr : lv) { ia.push_back(pr[0]); ib.push_back(pr[1]); }
    }
    if (ia.empty()) { ia.push_back(-1); ib.push_back(-1); }
    if (img_tar.empty()) img_tar.push_back(0);
    BHIP(upload(g.img_a, ia, s));
    BHIP(upload(g.img_b, ib, s));
    BHIP(upload(g.img_tar, img_tar, s));
    BHIP(upload(g.seg_off, seg, s));
    BHIP(upload(g.roots, roots, s));
    {   // Hadamard staging tables over the merged targets (tree parents only in mixed precision)
        const int ntar_h = b->mixed() ? nl : ntar;
        const size_t nch = (size_t)std::max(had_chunks(ntar_h), 1);
        BHIP(g.had_uniq.ensure(nch * kHadUniqStride * 4)); BHIP(g.had_cnt.ensure(nch * 4));
        BHIP(g.had_loc.ensure(nch * 26 * kHadChunk * 4));
        BHIP(launch_hadamard_tables(ntar_h, ntar, g.slot_src_T.as<int32_t>(), g.had_uniq.as<int32_t>(),
                                    g.had_cnt.as<int32_t>(), g.had_loc.as<int32_t>(), s));
    }
    const int64_t iota_n = std::max<int64_t>(std::max<int64_t>(nleaf, 8 * (int64_t)g.max_level_cols), 1);
    BHIP(g.iota.ensure((size_t)iota_n * 4));
    k_iota32<<<(unsigned)std::min<int64_t>((iota_n + 255) / 256, 4096), 256, 0, s>>>(g.iota.as<int32_t>(), iota_n);
    BHIP(hipGetLastError());
    // work buffers
    const int Sn = surface_count(g.P), N = 2 * g.P - 1;
    const size_t F = (size_t)N * N * (N / 2 + 1);
    BHIP(g.q.ensure(nb * 8)); BHIP(g.pot.ensure(nb * 8));
    BHIP(g.gx.ensure(nb * 8)); BHIP(g.gy.ensure(nb * 8)); BHIP(g.gz.ensure(nb * 8));
    BHIP(g.pcleaf.ensure(nlf * Sn * 8)); BHIP(g.pcleaf2.ensure(nlf * Sn * 8));
    BHIP(g.qeq.ensure(nc * Sn * 8)); BHIP(g.pck.ensure(nc * Sn * 8)); BHIP(g.qcell.ensure(nc * 8));
    BHIP(launch_fill_zero(g.qeq.as<double>(), (int64_t)nc * Sn, s));
    if (b->mixed()) {
        BHIP(g.Qk.ensure(F * std::max(nl, 1) * 64)); BHIP(g.Pk.ensure(F * std::max(nl, 1) * 64));
        BHIP(g.Qk_img.ensure(F * std::max(nimg, 1) * 128)); BHIP(g.Pk_img.ensure(F * std::max(nimg, 1) * 128));
        BHIP(g.rel.ensure(nb * 16));
    } else {
        BHIP(g.Qk.ensure(F * nt * 128)); BHIP(g.Pk.ensure(F * nt * 128));
    }
    BHIP(g.m2m_part.ensure((size_t)8 * g.max_level_cols * Sn * 8));
    BHIP(g.partial.ensure((size_t)g.sys.size() * kDipoleBlocks * 5 * 8));
    BHIP(g.moments.ensure((size_t)g.sys.size() * 5 * 8));
    BHIP(hipStreamSynchronize(s));
    return FMM3DPP_OK;
}

// The stage sequence of one group (the order of LaplaceFMM::solve, src/fmm.cpp:50-80), every stage
// one launch over all systems of the group.  No host synchronisation and no memset: capturable.
int run_group(fmm3dpp_batch* b, Group& g, const double* d_qin, double* d_phi, double* d_grad, hipStream_t s) {
    const Operators& o = g.ops;
    const int Sn = o.S, nleaf = g.nleaf;
    const CellsDev cells{g.cx.as<double>(), g.cy.as<double>(), g.cz.as<double>(), g.cr.as<double>(),
                         g.depth.as<int32_t>(), g.boff.as<int32_t>(), g.bnum.as<int32_t>()};
    const BodiesDev bod{g.x.as<double>(), g.y.as<double>(), g.z.as<double>(), g.q.as<double>()};
    double* qeq = g.qeq.as<double>();
    double* pck = g.pck.as<double>();
    const int32_t* iota = g.iota.as<int32_t>();
    const bool mixed = b->mixed();
    const bool f32 = mixed && f32_apply_orders(o.P);   // float32 tree operator stages
    BHIP(launch_gather(d_qin, g.perm.as<int32_t>(), g.q.as<double>(), (int)g.n, s));
    BHIP(launch_fill_zero(pck, (int64_t)g.ncells * Sn, s));
    // P2M
    BHIP(launch_p2m_check(Sn, g.dop.o_usurf.as<double>(), g.leaf_cell.as<int32_t>(), nleaf, cells, bod,
                          g.pcleaf.as<double>(), s, f32));
    {
        const ApplyCols g1{Sn, nleaf, 1, false, g.pcleaf.as<double>(), iota, iota, 0, g.pcleaf2.as<double>(), 0, nullptr};
        BHIP(launch_apply(g1, g.dop.lo_ut, s));
        const ApplyCols g2{Sn, nleaf, 1, false, g.pcleaf2.as<double>(), iota, g.leaf_cell.as<int32_t>(), 0, qeq, 0,
                           g.leaf_depth.as<int32_t>()};
        BHIP(launch_apply(g2, g.dop.lo_vs, s));
        if (f32)
            BHIP(launch_monopole_leaves(qeq, Sn, g.leaf_cell.as<int32_t>(), nleaf, cells, g.q.as<double>(),
                                        g.qcell.as<double>(), s));
    }
    // M2M per level, then the image levels (level m of every system in one launch)
    for (const auto& lv : g.m2m_levels) {
        const ApplyCols gm{Sn, lv.ncol, 8, false, qeq, g.m2m_in.as<int32_t>() + lv.in_off, iota, lv.ncol,
                           g.m2m_part.as<double>(), 0, nullptr};
        BHIP(launch_apply(gm, g.dop.lo_m2m, s));
        BHIP(launch_sum_partials(Sn, lv.ncol, 8, g.m2m_part.as<double>(), g.m2m_out.as<int32_t>() + lv.out_off, qeq, s));
        if (f32)
            BHIP(launch_monopole_parents(qeq, Sn, g.m2m_out.as<int32_t>() + lv.out_off, lv.ncol,
                                         g.m2m_in.as<int32_t>() + lv.in_off, g.qcell.as<double>(), s));
    }
    // infinite-lattice gauge: every system's root and image cells get their charge as monopole (the
    // system's total charge times 27^m for c(-m)); the per-system moments are those of the finalize
    const bool neutral = b->cfg[0].dipole_mode == FMM3DPP_DIPOLE_PHYSICAL;
    const int nsys_g = (int)g.sys.size();
    if (neutral) {
        BHIP(launch_dipole_seg(bod, g.seg_off.as<int64_t>(), nsys_g, kDipoleBlocks, g.partial.as<double>(),
                               g.moments.as<double>(), s));
        BHIP(launch_neutralize(qeq, Sn, g.roots.as<int32_t>(), 0, nsys_g, g.moments.as<double>() + 4, 5, 1.0, s));
    }
    double img_scale = 1.0;
    for (const auto& lv : g.img_m2m_lv) {
        const ApplyCols gi{Sn, (int)lv[1], 1, false, qeq, g.img_b.as<int32_t>() + lv[0], g.img_a.as<int32_t>() + lv[0], 0,
                           qeq, 0, nullptr};
        BHIP(launch_apply(gi, g.dop.lo_m2m_img, s));
        img_scale *= 27.0;
        if (neutral)
            BHIP(launch_neutralize(qeq, Sn, g.img_a.as<int32_t>() + lv[0], 0, (int)lv[1], g.moments.as<double>() + 4, 5,
                                   img_scale, s));
    }
    // M2L: targets = sources = all systems' non-leaf cells [0, nl), then all image cells [nl, ntar)
    const int ntar = g.ntar, nl = g.nl, nimg = g.nimg;
    const int32_t* slots = g.child_slots.as<int32_t>();
    const int32_t* inv = g.dop.o_invlat.as<int32_t>();
    const double* tw = g.dop.o_tw.as<double>();
    if (!mixed) {
        BHIP(launch_m2l_up(o.P, false, Sn, ntar, slots, inv, tw, qeq, g.Qk.p, s, g.up_tiles));
        BHIP(launch_hadamard8(false, o.F, ntar, ntar, g.slot_src_T.as<int32_t>(), g.had(), g.dop.o_gkA.p, g.Qk.p, g.Pk.p,
                              s));
        BHIP(launch_hadamard27(o.F, nimg, g.img_tar.as<int32_t>(), g.img_src_T.as<int32_t>(), ntar, 0,
                               g.dop.o_gk27.as<double>(), g.Qk.as<double>(), g.Pk.as<double>(), s));
        BHIP(launch_m2l_down(o.P, false, Sn, ntar, slots, g.tar_scale.as<double>(), inv, tw, g.Pk.p, pck, s,
                             g.down_tiles));
    } else {
        BHIP(launch_m2l_up(o.P, true, Sn, nl, slots, inv, tw, qeq, g.Qk.p, s, g.up_tiles));
        BHIP(launch_hadamard8(true, o.F, nl, ntar, g.slot_src_T.as<int32_t>(), g.had(), g.dop.o_gkA.p, g.Qk.p, g.Pk.p,
                              s));
        BHIP(launch_m2l_down(o.P, true, Sn, nl, slots, g.tar_scale.as<double>(), inv, tw, g.Pk.p, pck, s,
                             g.down_tiles));
        if (nimg > 0) {
            BHIP(launch_m2l_up(o.P, false, Sn, nimg, slots + (size_t)nl * 8, inv, tw, qeq, g.Qk_img.p, s));
            BHIP(launch_hadamard27(o.F, nimg, g.img_tar.as<int32_t>(), g.img_src_T.as<int32_t>(), ntar, nl,
                                   g.dop.o_gk27.as<double>(), g.Qk_img.as<double>(), g.Pk_img.as<double>(), s));
            BHIP(launch_m2l_down(o.P, false, Sn, nimg, slots + (size_t)nl * 8, g.tar_scale.as<double>() + nl, inv, tw,
                                 g.Pk_img.p, pck, s));
        }
    }
    // L2L: image levels from the outermost, then tree levels
    for (const auto& lv : g.img_l2l_lv) {
        const ApplyCols gi{Sn, (int)lv[1], 1, true, pck, g.img_a.as<int32_t>() + lv[0], g.img_b.as<int32_t>() + lv[0], 0,
                           pck, 0, nullptr};
        BHIP(launch_apply(gi, g.dop.lo_l2l_img, s));
    }
    for (const auto& lv : g.l2l_levels) {
        const ApplyCols gl{Sn, lv.ncol, 8, true, pck, g.l2l_in.as<int32_t>() + lv.in_off,
                           g.l2l_out.as<int32_t>() + lv.in_off, lv.ncol, pck, 0, nullptr};
        BHIP(launch_apply(gl, g.dop.lo_l2l, s));
    }
    // L2P
    {
        const ApplyCols g1{Sn, nleaf, 1, false, pck, g.leaf_cell.as<int32_t>(), iota, 0, g.pcleaf2.as<double>(), 0, nullptr};
        BHIP(launch_apply(g1, g.dop.lo_utl, s));
        const ApplyCols g2{Sn, nleaf, 1, false, g.pcleaf2.as<double>(), iota, iota, 0, g.pcleaf.as<double>(), 0,
                           g.leaf_depth.as<int32_t>()};
        BHIP(launch_apply(g2, g.dop.lo_vsl, s));
    }
    BHIP(launch_l2p_eval(Sn, g.dop.o_usurf.as<double>(), g.leaf_cell.as<int32_t>(), nleaf, cells, bod, g.pcleaf.as<double>(),
                         g.pot.as<double>(), g.gx.as<double>(), g.gy.as<double>(), g.gz.as<double>(), s, f32));
    // P2P
    const double cyc = b->cfg[0].cycle;
    if (!mixed) {
        BHIP(launch_p2p(g.leaf_cell.as<int32_t>(), nleaf, cells, bod, g.p2p_off.as<int64_t>(), g.p2p_ent.as<int32_t>(),
                        cyc, g.pot.as<double>(), g.gx.as<double>(), g.gy.as<double>(), g.gz.as<double>(), true, s));
    } else {
        BHIP(launch_pack_rel(g.leaf_cell.as<int32_t>(), nleaf, cells, bod, g.rel.as<float4>(), s));
        BHIP(launch_p2p_f32(g.leaf_cell.as<int32_t>(), nleaf, cells, g.rel.as<float4>(), g.p2p_off.as<int64_t>(),
                            g.p2p_ent.as<int32_t>(), cyc, g.pot.as<double>(), g.gx.as<double>(), g.gy.as<double>(),
                            g.gz.as<double>(), true, s));
    }
    // dipole term per system, 1/(4 pi), input order
    const int nseg = (int)g.sys.size();
    BHIP(launch_dipole_seg(bod, g.seg_off.as<int64_t>(), nseg, kDipoleBlocks, g.partial.as<double>(),
                           g.moments.as<double>(), s));
    BHIP(launch_finalize_seg(g.n, g.perm.as<int32_t>(), bod, g.pot.as<double>(), g.gx.as<double>(), g.gy.as<double>(),
                             g.gz.as<double>(), g.seg_off.as<int64_t>(), nseg, g.moments.as<double>(),
                             b->cfg[0].dipole_mode, cyc, b->cfg[0].divide_4pi ? 1.0 / (4 * M_PI) : 1.0, d_phi, d_grad,
                             s));
    return FMM3DPP_OK;
}

// All groups: group 0 on s, the others on their own streams forked from s and joined back into it
// (graph branches under capture).  The groups share only the read-only charges and write disjoint
// outputs (each system's own bodies).
int run_all(fmm3dpp_batch* b, const double* d_q, double* d_phi, double* d_grad, hipStream_t s) {
    const size_t ng = b->groups.size();
    if (ng > 1) BHIP(hipEventRecord(b->fork, s));
    for (size_t i = 1; i < ng; i++) {
        hipStream_t gs_ = b->group_streams[i - 1];
        BHIP(hipStreamWaitEvent(gs_, b->fork, 0));
        int rc = run_group(b, *b->groups[i], d_q, d_phi, d_grad, gs_);
        if (rc) return rc;
        BHIP(hipEventRecord(b->joins[i - 1], gs_));
    }
    int rc = run_group(b, *b->groups[0], d_q, d_phi, d_grad, s);
    if (rc) return rc;
    for (size_t i = 1; i < ng; i++) BHIP(hipStreamWaitEvent(s, b->joins[i - 1], 0));
    return FMM3DPP_OK;
}

int ensure_operators(fmm3dpp_batch* b) {
    for (auto& gp : b->groups) {
        Group& g = *gp;
        if (g.ops.P == 0) {
            double t0 = now_ms();
            int rc = build_operators_cached(g.P, b->cfg[0].images, b->cfg[0].cycle, g.ops, b->err);
            if (rc) return rc;
            b->ms["precompute"] += now_ms() - t0;
        }
        if (!g.ops_uploaded) {
            int rc = upload_group_operators(b, g);
            if (rc) return rc;
        }
    }
    return FMM3DPP_OK;
}

// Time of one evaluation of group g (the second of two runs), milliseconds.
int time_group(fmm3dpp_batch* b, Group& g, const double* d_q, double* d_phi, double* d_grad, float& ms) {
    hipStream_t s = b->stream;
    for (int rep = 0; rep < 2; rep++) {
        BHIP(hipEventRecord(b->ev0, s));
        int rc = run_group(b, g, d_q, d_phi, d_grad, s);
        if (rc) return rc;
        BHIP(hipEventRecord(b->ev1, s));
        BHIP(hipEventSynchronize(b->ev1));
    }
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    return FMM3DPP_OK;
}

// Transform tiles of the group's tree parents: each candidate (slot group, column chunk) of the down
// and then the up transform is timed over a whole group evaluation; the fastest is kept.  The
// results do not depend on the tiles.
int tune_tiles(fmm3dpp_batch* b, Group& g, const double* d_q, double* d_phi, double* d_grad) {
    const bool f32 = b->mixed();
    const M2LTiles down_cand[] = {{0, 0}, {4, 4}, {8, 2}, {4, 2}, {8, 4}, {2, 8}};
    const M2LTiles up_cand[] = {{0, 0}, {8, 4}, {8, 2}, {4, 4}, {8, 3}};
    for (int pass = 0; pass < 2; pass++) {
        M2LTiles& cur = pass == 0 ? g.down_tiles : g.up_tiles;
        const M2LTiles* cand = pass == 0 ? down_cand : up_cand;
        const int ncand = pass == 0 ? 6 : 5;
        M2LTiles best = cur;
        float best_ms = 1e30f;
        for (int c = 0; c < ncand; c++) {
            if (!m2l_tiles_valid(g.P, f32, pass == 0, cand[c])) continue;
            cur = cand[c];
            float ms = 0;
            int rc = time_group(b, g, d_q, d_phi, d_grad, ms);
            if (rc) return rc;
            if (ms < best_ms) { best_ms = ms; best = cand[c]; }
        }
        cur = best;
    }
    return FMM3DPP_OK;
}

}  // namespace

extern "C" {

int fmm3dpp_batch_create(const fmm3dpp_config* cfgs, int32_t nsys, int device, fmm3dpp_batch** out) {
    if (!out || !cfgs || nsys < 1) return FMM3DPP_ERR_ARG;
    *out = nullptr;
    const fmm3dpp_config& c0 = cfgs[0];
    std::string err;
    for (int k = 0; k < nsys; k++) {
        const fmm3dpp_config& c = cfgs[k];
        if (validate_config(&c, err)) return FMM3DPP_ERR_ARG;
        if (c.images != c0.images || c.cycle != c0.cycle || c.precision != c0.precision ||
            c.dipole_mode != c0.dipole_mode || c.divide_4pi != c0.divide_4pi)
            return FMM3DPP_ERR_ARG;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return FMM3DPP_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return FMM3DPP_ERR_DEVICE;
    std::unique_ptr<fmm3dpp_batch> b(new fmm3dpp_batch());
    b->cfg.assign(cfgs, cfgs + nsys);
    b->device = device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return FMM3DPP_ERR_HIP;
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) return FMM3DPP_ERR_HIP;
    for (int w = 0; w < std::min(nsys, kBuildThreads); w++) {
        hipStream_t bs = nullptr;
        if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess) return FMM3DPP_ERR_HIP;
        b->build_streams.push_back(bs);
    }
    for (int k = 0; k < nsys; k++) b->gs.emplace_back(new GpuStructure());
    std::map<int, int> by_p;
    for (int k = 0; k < nsys; k++) {
        auto it = by_p.find(cfgs[k].P);
        if (it == by_p.end()) {
            by_p[cfgs[k].P] = (int)b->groups.size();
            b->groups.emplace_back(new Group());
            b->groups.back()->P = cfgs[k].P;
            it = by_p.find(cfgs[k].P);
        }
        b->groups[it->second]->sys.push_back(k);
    }
    if (b->groups.size() > 1 && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return FMM3DPP_ERR_HIP;
    for (size_t i = 1; i < b->groups.size(); i++) {
        hipStream_t gs_ = nullptr;
        hipEvent_t e = nullptr;
        if (hipStreamCreateWithFlags(&gs_, hipStreamNonBlocking) != hipSuccess) return FMM3DPP_ERR_HIP;
        b->group_streams.push_back(gs_);
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return FMM3DPP_ERR_HIP;
        b->joins.push_back(e);
    }
    *out = b.release();
    return FMM3DPP_OK;
}

int fmm3dpp_batch_destroy(fmm3dpp_batch* b) {
    if (!b) return FMM3DPP_ERR_ARG;
    delete b;
    return FMM3DPP_OK;
}

const char* fmm3dpp_batch_last_error(const fmm3dpp_batch* b) { return b ? b->err.c_str() : "null batch"; }

int fmm3dpp_batch_set_particles_device(fmm3dpp_batch* b, const int64_t* counts, const double* d_x, const double* d_y,
                                       const double* d_z) {
    if (!b) return FMM3DPP_ERR_ARG;
    if (!counts || !d_x || !d_y || !d_z) BFAIL(FMM3DPP_ERR_ARG, "null argument");
    const int nsys = (int)b->cfg.size();
    b->counts.assign(counts, counts + nsys);
    b->offsets.assign(nsys + 1, 0);
    for (int k = 0; k < nsys; k++) {
        if (counts[k] < 1 || counts[k] > (int64_t)INT32_MAX / 2) BFAIL(FMM3DPP_ERR_ARG, "system body count out of range");
        b->offsets[k + 1] = b->offsets[k] + counts[k];
    }
    if (b->offsets[nsys] > (int64_t)INT32_MAX / 2) BFAIL(FMM3DPP_ERR_ARG, "too many bodies in the batch");
    BHIP(hipSetDevice(b->device));
    b->particles = false;
    b->drop_graph();
    const size_t bytes = (size_t)b->offsets[nsys] * sizeof(double);
    BHIP(b->pos_x.ensure(bytes));
    BHIP(b->pos_y.ensure(bytes));
    BHIP(b->pos_z.ensure(bytes));
    BHIP(hipMemcpyAsync(b->pos_x.p, d_x, bytes, hipMemcpyDeviceToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->pos_y.p, d_y, bytes, hipMemcpyDeviceToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->pos_z.p, d_z, bytes, hipMemcpyDeviceToDevice, b->stream));
    b->px = b->pos_x.as<double>();
    b->py = b->pos_y.as<double>();
    b->pz = b->pos_z.as<double>();
    double t0 = now_ms();
    for (auto& gp : b->groups) {
        Group& g = *gp;
        int rc = build_group(b, g);
        if (rc) return rc;
    }
    b->ms["set_particles"] = now_ms() - t0;
    b->particles = true;
    return FMM3DPP_OK;
}

int fmm3dpp_batch_set_graph(fmm3dpp_batch* b, int enable) {
    if (!b) return FMM3DPP_ERR_ARG;
    b->use_graph = enable != 0;
    if (!b->use_graph) b->drop_graph();
    return FMM3DPP_OK;
}

int fmm3dpp_batch_evaluate_device(fmm3dpp_batch* b, const double* d_q, double* d_phi, double* d_grad) {
    if (!b) return FMM3DPP_ERR_ARG;
    if (!b->particles) BFAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate");
    if (!d_q || !d_phi || !d_grad) BFAIL(FMM3DPP_ERR_ARG, "null array");
    BHIP(hipSetDevice(b->device));
    int rc = ensure_operators(b);
    if (rc) return rc;
    hipStream_t s = b->stream;
    const void* key[3] = {d_q, d_phi, d_grad};
    if (b->use_graph && b->graph && !std::equal(key, key + 3, b->graph_key)) b->drop_graph();
    BHIP(hipEventRecord(b->ev0, s));
    if (b->use_graph && b->graph) {
        BHIP(hipGraphLaunch(b->graph, s));
    } else if (b->use_graph) {
        // first run uncaptured (sets kernel attributes), then capture the sequence for the next calls
        rc = run_all(b, d_q, d_phi, d_grad, s);
        if (rc) return rc;
        hipGraph_t graph = nullptr;
        BHIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = run_all(b, d_q, d_phi, d_grad, s);
        hipError_t ec = hipStreamEndCapture(s, &graph);
        if ((rc || ec != hipSuccess) && graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        BHIP(ec);
        hipError_t ei = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BHIP(ei);
        std::copy(key, key + 3, b->graph_key);
    } else {
        rc = run_all(b, d_q, d_phi, d_grad, s);
        if (rc) return rc;
    }
    BHIP(hipEventRecord(b->ev1, s));
    BHIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->ms["evaluate"] = ms;
    return FMM3DPP_OK;
}

int fmm3dpp_batch_autotune(fmm3dpp_batch* b, const int32_t* candidates, int32_t ncand, const double* d_q,
                           double* d_phi, double* d_grad, int32_t* chosen) {
    if (!b) return FMM3DPP_ERR_ARG;
    if (!b->particles) BFAIL(FMM3DPP_ERR_STATE, "set_particles must precede autotune");
    if (!candidates || ncand < 1 || !d_q || !d_phi || !d_grad) BFAIL(FMM3DPP_ERR_ARG, "null argument");
    for (int c = 0; c < ncand; c++)
        if (candidates[c] < 1) BFAIL(FMM3DPP_ERR_ARG, "leaf size must be >= 1");
    BHIP(hipSetDevice(b->device));
    int rc = ensure_operators(b);
    if (rc) return rc;
    const bool graph = b->use_graph;
    b->use_graph = false;
    b->drop_graph();
    for (auto& gp : b->groups) {
        Group& g = *gp;
        double best = 1e300;
        int best_c = b->cfg[g.sys[0]].ncrit;
        for (int c = 0; c < ncand; c++) {
            for (int k : g.sys) b->cfg[k].ncrit = candidates[c];
            double t0 = now_ms();
            rc = build_group(b, g);
            if (rc) return rc;
            const double tb = now_ms() - t0;
            float ms = 0;
            rc = time_group(b, g, d_q, d_phi, d_grad, ms);
            if (rc) return rc;
            const double t = ms + tb;   // one step = tree and lists + evaluation
            if (t < best) { best = t; best_c = candidates[c]; }
        }
        for (int k : g.sys) b->cfg[k].ncrit = best_c;
        rc = build_group(b, g);
        if (rc) return rc;
        rc = tune_tiles(b, g, d_q, d_phi, d_grad);
        if (rc) return rc;
    }
    if (chosen)
        for (size_t k = 0; k < b->cfg.size(); k++) chosen[k] = b->cfg[k].ncrit;
    b->use_graph = graph;
    return FMM3DPP_OK;
}

double fmm3dpp_batch_timing_ms(const fmm3dpp_batch* b, const char* name) {
    if (!b || !name) return -1;
    auto it = b->ms.find(name);
    return it == b->ms.end() ? -1 : it->second;
}

}  // extern "C"
