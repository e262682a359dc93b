// C-ABI of the MI355X periodic KIFMM (include/fmm3dpp.h): plan lifecycle, host tree/list build,
// device buffers and the evaluation stage sequence of LaplaceFMM::solve (3dpp/src/fmm.cpp:13-81).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "device.h"
#include "fmm3dpp.h"
#include "structure.h"

using namespace fmm3dpp;

namespace {

// Device allocation that grows on demand and is freed with the plan.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t ensure(size_t b) {
        if (b <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, std::max<size_t>(b, 256));
        if (e == hipSuccess) bytes = std::max<size_t>(b, 256);
        return e;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

template <class T>
hipError_t upload(DBuf& d, const std::vector<T>& h, hipStream_t st) {
    hipError_t e = d.ensure(h.size() * sizeof(T));
    if (e != hipSuccess || h.empty()) return e;
    return hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

struct GemmLevel { int ncol; size_t in_off, out_off; };

const char* kStageNames[] = {"P2M", "M2M", "M2L_up", "M2L_hadamard", "M2L_down", "L2L", "L2P", "P2P", "finalize"};
constexpr int kStages = 9;

}  // namespace

struct fmm3dpp_plan {
    fmm3dpp_config cfg{};
    int device = FMM3DPP_DEVICE_NONE;
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool ops_ready = false, ops_uploaded = false, particles = false;
    Operators ops;
    Structure st;
    std::map<std::string, double> host_ms;
    hipEvent_t ev[kStages + 1] = {};
    bool ev_recorded = false;

    // host-side index arrays derived from the structure
    std::vector<int32_t> leaf_cell, leaf_depth, iota;
    std::vector<int32_t> m2m_in, m2m_out;          // per level, 8 children per parent
    std::vector<GemmLevel> m2m_levels;
    std::vector<int32_t> l2l_in, l2l_out;
    std::vector<GemmLevel> l2l_levels;
    int max_level_cols = 0;
    std::vector<std::array<int32_t, 2>> img_m2m;   // (parent c(-m), child), deepest first
    std::vector<std::array<int32_t, 2>> img_l2l;   // (parent, child), outermost first
    std::vector<int32_t> src_cells, tar_cells, src_child, tar_child, slot_src_T, img_src_T, img_tar;
    std::vector<double> tar_scale;
    std::vector<int64_t> p2p_off;
    std::vector<int32_t> p2p_src;
    std::vector<double> p2p_shift;

    // device buffers
    DBuf d_x, d_y, d_z, d_perm, d_q, d_qin, d_p, d_gx, d_gy, d_gz, d_phi, d_grad;
    DBuf d_cx, d_cy, d_cz, d_cr, d_depth, d_boff, d_bnum;
    DBuf d_leaf_cell, d_leaf_depth, d_iota, d_iota8, d_pcleaf, d_pcleaf2, d_qeq, d_pck;
    DBuf d_m2m_in, d_m2m_out, d_l2l_in, d_l2l_out, d_img_a, d_img_b;
    DBuf d_src_child, d_tar_child, d_slot_src_T, d_img_src_T, d_img_tar, d_tar_scale, d_Qk, d_Pk;
    DBuf d_p2p_off, d_p2p_src, d_p2p_shift, d_partial, d_m2m_part;
    DBuf o_ut, o_vs, o_utl, o_vsl, o_m2m_all, o_l2l, o_m2m_img, o_l2l_img, o_gk8, o_gk27, o_usurf, o_lattice, o_tw;

    CellsDev cells_dev() const {
        return {d_cx.as<double>(), d_cy.as<double>(), d_cz.as<double>(), d_cr.as<double>(), d_depth.as<int32_t>(),
                d_boff.as<int32_t>(), d_bnum.as<int32_t>()};
    }
    BodiesDev bodies_dev() const { return {d_x.as<double>(), d_y.as<double>(), d_z.as<double>(), d_q.as<double>()}; }
};

#define FAIL(code, msg)          \
    do {                         \
        plan->err = (msg);       \
        return (code);           \
    } while (0)
#define HIPCK(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) FAIL(FMM3DPP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

int validate(const fmm3dpp_config* c, std::string& err) {
    if (!c) { err = "null config"; return FMM3DPP_ERR_ARG; }
    if (c->P < 2 || c->P > 12) { err = "P must be in [2, 12]"; return FMM3DPP_ERR_ARG; }
    if (c->ncrit < 1) { err = "ncrit must be >= 1"; return FMM3DPP_ERR_ARG; }
    if (c->images < 0 || c->images > 20) { err = "images must be in [0, 20]"; return FMM3DPP_ERR_ARG; }
    if (!(c->cycle > 0) || !std::isfinite(c->cycle)) { err = "cycle must be positive and finite"; return FMM3DPP_ERR_ARG; }
    if (c->precision != FMM3DPP_PREC_FP64 && c->precision != FMM3DPP_PREC_MIXED) { err = "unknown precision"; return FMM3DPP_ERR_ARG; }
    if (c->precision == FMM3DPP_PREC_MIXED) { err = "mixed precision is not available in this build"; return FMM3DPP_ERR_ARG; }
    if (c->dipole_mode != FMM3DPP_DIPOLE_OFF && c->dipole_mode != FMM3DPP_DIPOLE_REFERENCE) { err = "unknown dipole mode"; return FMM3DPP_ERR_ARG; }
    return FMM3DPP_OK;
}

// Derive the device-side index arrays from the host tree/lists.
void derive_indices(fmm3dpp_plan* plan) {
    const Structure& S = plan->st;
    const auto& C = S.cells;
    // leaves in body order
    plan->leaf_cell.clear();
    for (int32_t c = 0; c < S.ntree; c++) if (C[c].cnum == 0) plan->leaf_cell.push_back(c);
    std::sort(plan->leaf_cell.begin(), plan->leaf_cell.end(), [&](int a, int b) { return C[a].boff < C[b].boff; });
    int nleaf = (int)plan->leaf_cell.size();
    plan->leaf_depth.resize(nleaf);
    plan->iota.resize(nleaf);
    for (int i = 0; i < nleaf; i++) { plan->leaf_depth[i] = C[plan->leaf_cell[i]].depth; plan->iota[i] = i; }
    // M2M and L2L per level: parent columns, child slots octant-major ([octant][column])
    plan->m2m_in.clear(); plan->m2m_out.clear(); plan->m2m_levels.clear();
    plan->l2l_in.clear(); plan->l2l_out.clear(); plan->l2l_levels.clear();
    std::vector<std::vector<int32_t>> by_depth(S.max_depth + 1);
    for (int32_t c = 0; c < S.ntree; c++) if (C[c].cnum > 0) by_depth[C[c].depth].push_back(c);
    plan->max_level_cols = 0;
    for (int d = S.max_depth - 1; d >= 0; d--) {
        const auto& pars = by_depth[d];
        int ncol = (int)pars.size();
        if (ncol == 0) continue;
        plan->max_level_cols = std::max(plan->max_level_cols, ncol);
        GemmLevel lv{ncol, plan->m2m_in.size(), plan->m2m_out.size()};
        std::vector<int32_t> slots((size_t)8 * ncol, -1);
        for (int c = 0; c < ncol; c++)
            for (int k = 0; k < C[pars[c]].cnum; k++) {
                int32_t ch = C[pars[c]].coff + k;
                slots[(size_t)C[ch].octant * ncol + c] = ch;
            }
        plan->m2m_in.insert(plan->m2m_in.end(), slots.begin(), slots.end());
        plan->m2m_out.insert(plan->m2m_out.end(), pars.begin(), pars.end());
        plan->m2m_levels.push_back(lv);
    }
    for (int d = 0; d < S.max_depth; d++) {
        const auto& pars = by_depth[d];
        int ncol = (int)pars.size();
        if (ncol == 0) continue;
        GemmLevel lv{ncol, plan->l2l_in.size(), plan->l2l_out.size()};
        std::vector<int32_t> slots((size_t)8 * ncol, -1), src((size_t)8 * ncol);
        for (int c = 0; c < ncol; c++) {
            for (int o = 0; o < 8; o++) src[(size_t)o * ncol + c] = pars[c];
            for (int k = 0; k < C[pars[c]].cnum; k++) {
                int32_t ch = C[pars[c]].coff + k;
                slots[(size_t)C[ch].octant * ncol + c] = ch;
            }
        }
        plan->l2l_in.insert(plan->l2l_in.end(), src.begin(), src.end());
        plan->l2l_out.insert(plan->l2l_out.end(), slots.begin(), slots.end());
        plan->l2l_levels.push_back(lv);
    }
    plan->img_m2m.clear(); plan->img_l2l.clear();
    const int images = plan->cfg.images;
    // M2M image levels run for depths -1 .. -(images-1) (fmm.cpp:114: depth > min depth)
    for (int m = 1; m <= images - 1; m++) {
        int32_t par = S.ntree + m - 1;
        plan->img_m2m.push_back({par, C[par].coff});
    }
    // L2L image levels for depths -images .. -1 (fmm.cpp:198)
    for (int m = images; m >= 1; m--) {
        int32_t par = S.ntree + m - 1;
        plan->img_l2l.push_back({par, C[par].coff});
    }
    // M2L parent map -> Hadamard tables (kernel.cpp:1145-1240)
    std::vector<int32_t> tar_local(C.size(), -1), src_local(C.size(), -1);
    plan->tar_cells.clear(); plan->src_cells.clear();
    for (const auto& r : S.m2l) {
        if (tar_local[r.tar] < 0) { tar_local[r.tar] = (int32_t)plan->tar_cells.size(); plan->tar_cells.push_back(r.tar); }
        if (src_local[r.src] < 0) src_local[r.src] = 0;
    }
    for (size_t c = 0; c < C.size(); c++)
        if (src_local[c] >= 0) { src_local[c] = (int32_t)plan->src_cells.size(); plan->src_cells.push_back((int32_t)c); }
    int ntar = (int)plan->tar_cells.size(), nsrc = (int)plan->src_cells.size();
    auto slots_of = [&](int32_t par, int32_t* out) {
        for (int o = 0; o < 8; o++) out[o] = -1;
        for (int k = 0; k < C[par].cnum; k++) {
            int oct = C[C[par].coff + k].octant;
            out[oct == 13 ? 0 : oct] = C[par].coff + k;
        }
    };
    plan->src_child.assign((size_t)nsrc * 8, -1);
    for (int s = 0; s < nsrc; s++) slots_of(plan->src_cells[s], &plan->src_child[(size_t)s * 8]);
    plan->tar_child.assign((size_t)ntar * 8, -1);
    plan->tar_scale.assign(ntar, 0.0);
    for (int t = 0; t < ntar; t++) {
        int32_t c = plan->tar_cells[t];
        slots_of(c, &plan->tar_child[(size_t)t * 8]);
        plan->tar_scale[t] = std::pow(C[c].depth >= -1 ? 2.0 : 3.0, C[c].depth + 1);  // kernel.cpp:1372
    }
    plan->slot_src_T.assign((size_t)26 * ntar, -1);
    plan->img_src_T.assign((size_t)26 * ntar, -1);
    std::vector<char> is_img_tar(ntar, 0);
    for (const auto& r : S.m2l) {
        int t = tar_local[r.tar];
        if (r.img) { plan->img_src_T[(size_t)r.pattern * ntar + t] = src_local[r.src]; is_img_tar[t] = 1; }
        else plan->slot_src_T[(size_t)r.pattern * ntar + t] = src_local[r.src];
    }
    plan->img_tar.clear();
    for (int t = 0; t < ntar; t++) if (is_img_tar[t]) plan->img_tar.push_back(t);
    // P2P rows -> per-leaf source lists (rows whose target is a non-leaf go to every leaf below it)
    std::vector<std::vector<std::pair<int32_t, int64_t>>> per_leaf(nleaf);  // (src cell, row index)
    std::vector<int32_t> leaf_boff(nleaf);
    for (int i = 0; i < nleaf; i++) leaf_boff[i] = C[plan->leaf_cell[i]].boff;
    for (size_t ri = 0; ri < S.p2p.size(); ri++) {
        const auto& r = S.p2p[ri];
        const auto& tc = C[r.tar];
        if (tc.bnum == 0) continue;
        auto lo = std::lower_bound(leaf_boff.begin(), leaf_boff.end(), tc.boff) - leaf_boff.begin();
        for (int64_t l = lo; l < nleaf && leaf_boff[l] < tc.boff + tc.bnum; l++) per_leaf[l].push_back({r.src, (int64_t)ri});
    }
    plan->p2p_off.assign(nleaf + 1, 0);
    plan->p2p_src.clear(); plan->p2p_shift.clear();
    for (int l = 0; l < nleaf; l++) {
        for (auto& e : per_leaf[l]) {
            plan->p2p_src.push_back(e.first);
            const auto& r = S.p2p[e.second];
            plan->p2p_shift.push_back(r.off[0]);
            plan->p2p_shift.push_back(r.off[1]);
            plan->p2p_shift.push_back(r.off[2]);
        }
        plan->p2p_off[l + 1] = (int64_t)plan->p2p_src.size();
    }
}

int upload_operators(fmm3dpp_plan* plan) {
    hipStream_t s = plan->stream;
    const Operators& o = plan->ops;
    std::vector<double> l2l_all, m2m_all;
    for (int k = 0; k < 8; k++) l2l_all.insert(l2l_all.end(), o.l2l[k].begin(), o.l2l[k].end());
    for (int k = 0; k < 8; k++) m2m_all.insert(m2m_all.end(), o.m2m[k].begin(), o.m2m[k].end());
    std::vector<double> zeros((size_t)o.S * o.S, 0.0);
    HIPCK(upload(plan->o_ut, o.UT, s));
    HIPCK(upload(plan->o_vs, o.VSinv, s));
    HIPCK(upload(plan->o_utl, o.UT_l2p, s));
    HIPCK(upload(plan->o_vsl, o.VSinv_l2p, s));
    HIPCK(upload(plan->o_m2m_all, m2m_all, s));
    HIPCK(upload(plan->o_l2l, l2l_all, s));
    HIPCK(upload(plan->o_m2m_img, o.images > 0 ? o.m2m_img_sum : zeros, s));
    HIPCK(upload(plan->o_l2l_img, o.images > 0 ? o.l2l_img : zeros, s));
    HIPCK(upload(plan->o_gk8, o.gk8, s));
    HIPCK(upload(plan->o_gk27, o.gk27_sum, s));
    HIPCK(upload(plan->o_usurf, o.unit_surf, s));
    HIPCK(upload(plan->o_lattice, o.lattice, s));
    HIPCK(upload(plan->o_tw, o.twiddle, s));
    HIPCK(hipStreamSynchronize(s));
    plan->ops_uploaded = true;
    return FMM3DPP_OK;
}

int upload_structure(fmm3dpp_plan* plan) {
    hipStream_t s = plan->stream;
    const Structure& S = plan->st;
    const auto& C = S.cells;
    size_t nc = C.size();
    std::vector<double> cx(nc), cy(nc), cz(nc), cr(nc);
    std::vector<int32_t> dep(nc), boff(nc), bnum(nc);
    for (size_t c = 0; c < nc; c++) {
        cx[c] = C[c].x[0]; cy[c] = C[c].x[1]; cz[c] = C[c].x[2]; cr[c] = C[c].r;
        dep[c] = C[c].depth; boff[c] = C[c].boff; bnum[c] = C[c].bnum;
    }
    HIPCK(upload(plan->d_x, S.x, s));
    HIPCK(upload(plan->d_y, S.y, s));
    HIPCK(upload(plan->d_z, S.z, s));
    HIPCK(upload(plan->d_perm, S.perm, s));
    HIPCK(upload(plan->d_cx, cx, s)); HIPCK(upload(plan->d_cy, cy, s)); HIPCK(upload(plan->d_cz, cz, s));
    HIPCK(upload(plan->d_cr, cr, s)); HIPCK(upload(plan->d_depth, dep, s));
    HIPCK(upload(plan->d_boff, boff, s)); HIPCK(upload(plan->d_bnum, bnum, s));
    HIPCK(upload(plan->d_leaf_cell, plan->leaf_cell, s));
    HIPCK(upload(plan->d_leaf_depth, plan->leaf_depth, s));
    HIPCK(upload(plan->d_iota, plan->iota, s));
    {   // 0 .. 8*widest-1: M2M partial of (octant z, column c) goes to row z*ncol + c of the scratch
        int w = std::max(1, plan->max_level_cols);
        std::vector<int32_t> io8((size_t)8 * w);
        for (size_t i = 0; i < io8.size(); i++) io8[i] = (int32_t)i;
        HIPCK(upload(plan->d_iota8, io8, s));
    }
    HIPCK(upload(plan->d_m2m_in, plan->m2m_in, s)); HIPCK(upload(plan->d_m2m_out, plan->m2m_out, s));
    HIPCK(upload(plan->d_l2l_in, plan->l2l_in, s)); HIPCK(upload(plan->d_l2l_out, plan->l2l_out, s));
    std::vector<int32_t> ia, ib;
    for (auto& p : plan->img_m2m) { ia.push_back(p[0]); ib.push_back(p[1]); }
    for (auto& p : plan->img_l2l) { ia.push_back(p[0]); ib.push_back(p[1]); }
    HIPCK(upload(plan->d_img_a, ia, s)); HIPCK(upload(plan->d_img_b, ib, s));
    HIPCK(upload(plan->d_src_child, plan->src_child, s)); HIPCK(upload(plan->d_tar_child, plan->tar_child, s));
    HIPCK(upload(plan->d_slot_src_T, plan->slot_src_T, s)); HIPCK(upload(plan->d_img_src_T, plan->img_src_T, s));
    HIPCK(upload(plan->d_img_tar, plan->img_tar, s)); HIPCK(upload(plan->d_tar_scale, plan->tar_scale, s));
    HIPCK(upload(plan->d_p2p_off, plan->p2p_off, s)); HIPCK(upload(plan->d_p2p_src, plan->p2p_src, s));
    HIPCK(upload(plan->d_p2p_shift, plan->p2p_shift, s));
    const int Sn = plan->ops.S > 0 ? plan->ops.S : surface_count(plan->cfg.P);
    const int F = (2 * plan->cfg.P - 1) * (2 * plan->cfg.P - 1) * ((2 * plan->cfg.P - 1) / 2 + 1);
    size_t n = (size_t)S.n;
    HIPCK(plan->d_q.ensure(n * 8)); HIPCK(plan->d_qin.ensure(n * 8));
    HIPCK(plan->d_p.ensure(n * 8)); HIPCK(plan->d_gx.ensure(n * 8)); HIPCK(plan->d_gy.ensure(n * 8)); HIPCK(plan->d_gz.ensure(n * 8));
    HIPCK(plan->d_phi.ensure(n * 8)); HIPCK(plan->d_grad.ensure(n * 24));
    HIPCK(plan->d_pcleaf.ensure((size_t)plan->leaf_cell.size() * Sn * 8));
    HIPCK(plan->d_pcleaf2.ensure((size_t)plan->leaf_cell.size() * Sn * 8));
    HIPCK(plan->d_qeq.ensure(nc * Sn * 8)); HIPCK(plan->d_pck.ensure(nc * Sn * 8));
    HIPCK(hipMemsetAsync(plan->d_qeq.p, 0, nc * Sn * 8, s));
    HIPCK(plan->d_Qk.ensure((size_t)F * plan->src_cells.size() * 8 * 16));
    HIPCK(plan->d_Pk.ensure((size_t)F * plan->tar_cells.size() * 8 * 16));
    HIPCK(plan->d_partial.ensure(256 * 3 * 8));
    HIPCK(plan->d_m2m_part.ensure((size_t)8 * std::max(1, plan->max_level_cols) * Sn * 8));
    HIPCK(hipStreamSynchronize(s));
    return FMM3DPP_OK;
}

int ensure_ops(fmm3dpp_plan* plan) {
    if (!plan->ops_ready) {
        double t0 = 0;
        int rc = build_operators(plan->cfg.P, plan->cfg.images, plan->cfg.cycle, plan->ops, plan->err);
        (void)t0;
        if (rc) return rc;
        plan->ops_ready = true;
        plan->ops_uploaded = false;
    }
    if (plan->device >= 0 && !plan->ops_uploaded) return upload_operators(plan);
    return FMM3DPP_OK;
}

}  // namespace

extern "C" {

int fmm3dpp_plan_create(const fmm3dpp_config* cfg, int device, fmm3dpp_plan** out) {
    if (!out) return FMM3DPP_ERR_ARG;
    *out = nullptr;
    std::string err;
    int rc = validate(cfg, err);
    if (rc) return rc;
    std::unique_ptr<fmm3dpp_plan> plan(new fmm3dpp_plan());
    plan->cfg = *cfg;
    plan->device = device;
    if (device != FMM3DPP_DEVICE_NONE) {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FMM3DPP_ERR_DEVICE;
        if (device < 0 || device >= count) return FMM3DPP_ERR_DEVICE;
        if (hipSetDevice(device) != hipSuccess) return FMM3DPP_ERR_DEVICE;
        if (hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking) != hipSuccess) return FMM3DPP_ERR_HIP;
        plan->own_stream = true;
        for (int i = 0; i <= kStages; i++)
            if (hipEventCreate(&plan->ev[i]) != hipSuccess) return FMM3DPP_ERR_HIP;
    }
    *out = plan.release();
    return FMM3DPP_OK;
}

int fmm3dpp_plan_destroy(fmm3dpp_plan* plan) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device >= 0) {
        (void)hipSetDevice(plan->device);
        if (plan->stream) (void)hipStreamSynchronize(plan->stream);
        for (int i = 0; i <= kStages; i++) if (plan->ev[i]) (void)hipEventDestroy(plan->ev[i]);
        if (plan->own_stream && plan->stream) (void)hipStreamDestroy(plan->stream);
    }
    delete plan;
    return FMM3DPP_OK;
}

const char* fmm3dpp_last_error(const fmm3dpp_plan* plan) { return plan ? plan->err.c_str() : "null plan"; }

int fmm3dpp_set_stream(fmm3dpp_plan* plan, void* s) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan has no device");
    if (plan->own_stream && plan->stream) (void)hipStreamDestroy(plan->stream);
    plan->stream = static_cast<hipStream_t>(s);
    plan->own_stream = false;
    return FMM3DPP_OK;
}

int fmm3dpp_precompute(fmm3dpp_plan* plan) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device >= 0) HIPCK(hipSetDevice(plan->device));
    return ensure_ops(plan);
}

int fmm3dpp_set_particles(fmm3dpp_plan* plan, int64_t n, const double* x, const double* y, const double* z) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (!x || !y || !z) FAIL(FMM3DPP_ERR_ARG, "null position array");
    plan->particles = false;
    auto clk = [] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    int rc = build_structure(plan->cfg, n, x, y, z, plan->st, plan->err);
    if (rc) return rc;
    double t0 = clk();
    derive_indices(plan);
    double t1 = clk();
    if (plan->device >= 0) {
        HIPCK(hipSetDevice(plan->device));
        rc = upload_structure(plan);
        if (rc) return rc;
    }
    plan->host_ms["tree"] = plan->st.ms_tree;
    plan->host_ms["lists_p2p"] = plan->st.ms_p2p;
    plan->host_ms["lists_m2l"] = plan->st.ms_m2l;
    plan->host_ms["indices"] = t1 - t0;
    plan->host_ms["upload"] = clk() - t1;
    plan->particles = true;
    return FMM3DPP_OK;
}

int fmm3dpp_evaluate_device(fmm3dpp_plan* plan, const double* d_qin, double* d_phi, double* d_grad) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate");
    HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    hipStream_t s = plan->stream;
    const Operators& o = plan->ops;
    const Structure& S = plan->st;
    const int n = (int)S.n, Sn = o.S, nleaf = (int)plan->leaf_cell.size();
    const CellsDev cells = plan->cells_dev();
    const BodiesDev bod = plan->bodies_dev();
    double* qeq = plan->d_qeq.as<double>();
    double* pck = plan->d_pck.as<double>();
    const size_t nc = S.cells.size();

    HIPCK(hipEventRecord(plan->ev[0], s));
    HIPCK(launch_gather(d_qin, plan->d_perm.as<int32_t>(), plan->d_q.as<double>(), n, s));
    HIPCK(hipMemsetAsync(pck, 0, nc * Sn * sizeof(double), s));
    // P2M (fmm.cpp:83-109)
    HIPCK(launch_p2m_check(Sn, plan->o_usurf.as<double>(), plan->d_leaf_cell.as<int32_t>(), nleaf, cells, bod,
                           plan->d_pcleaf.as<double>(), s));
    {   // q_equiv = 0.5^depth * VSinv * (UT * p_check)   (kernel.cpp:369-370)
        const int32_t* iota = plan->d_iota.as<int32_t>();
        GemmCols g1{Sn, nleaf, 1, 0, plan->o_ut.as<double>(), 0, plan->d_pcleaf.as<double>(), iota, iota, 0,
                    plan->d_pcleaf2.as<double>(), 0, nullptr};
        HIPCK(launch_gemm_cols(g1, s));
        GemmCols g2{Sn, nleaf, 1, 0, plan->o_vs.as<double>(), 0, plan->d_pcleaf2.as<double>(), iota,
                    plan->d_leaf_cell.as<int32_t>(), 0, qeq, 0, plan->d_leaf_depth.as<int32_t>()};
        HIPCK(launch_gemm_cols(g2, s));
    }
    HIPCK(hipEventRecord(plan->ev[1], s));
    // M2M (fmm.cpp:111-150): per level, one GEMM per child octant into partials, then the child sum
    const size_t SS = (size_t)Sn * Sn;
    for (const auto& lv : plan->m2m_levels) {
        GemmCols g{Sn, lv.ncol, 8, 0, plan->o_m2m_all.as<double>(), (int64_t)SS, qeq,
                   plan->d_m2m_in.as<int32_t>() + lv.in_off, plan->d_iota8.as<int32_t>(), lv.ncol,
                   plan->d_m2m_part.as<double>(), 0, nullptr};
        HIPCK(launch_gemm_cols(g, s));
        HIPCK(launch_sum_partials(Sn, lv.ncol, 8, plan->d_m2m_part.as<double>(),
                                  plan->d_m2m_out.as<int32_t>() + lv.out_off, qeq, s));
    }
    for (size_t i = 0; i < plan->img_m2m.size(); i++) {
        GemmCols g{Sn, 1, 1, 0, plan->o_m2m_img.as<double>(), 0, qeq, plan->d_img_b.as<int32_t>() + i,
                   plan->d_img_a.as<int32_t>() + i, 0, qeq, 0, nullptr};
        HIPCK(launch_gemm_cols(g, s));
    }
    HIPCK(hipEventRecord(plan->ev[2], s));
    // M2L (fmm.cpp:152-193, kernel.cpp:1140-1397)
    const int ntar = (int)plan->tar_cells.size(), nsrc = (int)plan->src_cells.size();
    HIPCK(launch_m2l_up(o.P, Sn, nsrc, plan->d_src_child.as<int32_t>(), plan->o_lattice.as<int32_t>(),
                        plan->o_tw.as<double>(), qeq, plan->d_Qk.as<double>(), s));
    HIPCK(hipEventRecord(plan->ev[3], s));
    HIPCK(launch_hadamard8(o.F, ntar, nsrc, plan->d_slot_src_T.as<int32_t>(), plan->o_gk8.as<double>(),
                           plan->d_Qk.as<double>(), plan->d_Pk.as<double>(), s));
    HIPCK(hipEventRecord(plan->ev[4], s));
    HIPCK(launch_hadamard27(o.F, (int)plan->img_tar.size(), plan->d_img_tar.as<int32_t>(),
                            plan->d_img_src_T.as<int32_t>(), ntar, nsrc, plan->o_gk27.as<double>(),
                            plan->d_Qk.as<double>(), plan->d_Pk.as<double>(), s));
    HIPCK(launch_m2l_down(o.P, Sn, ntar, plan->d_tar_child.as<int32_t>(), plan->d_tar_scale.as<double>(),
                          plan->o_lattice.as<int32_t>(), plan->o_tw.as<double>(), plan->d_Pk.as<double>(), pck, s));
    HIPCK(hipEventRecord(plan->ev[5], s));
    // L2L (fmm.cpp:195-229): image levels from the outermost image, then one launch per tree level
    const size_t nimgm2m = plan->img_m2m.size();
    for (size_t i = 0; i < plan->img_l2l.size(); i++) {
        GemmCols g{Sn, 1, 1, 1, plan->o_l2l_img.as<double>(), 0, pck, plan->d_img_a.as<int32_t>() + nimgm2m + i,
                   plan->d_img_b.as<int32_t>() + nimgm2m + i, 0, pck, 0, nullptr};
        HIPCK(launch_gemm_cols(g, s));
    }
    for (const auto& lv : plan->l2l_levels) {
        GemmCols g{Sn, lv.ncol, 8, 1, plan->o_l2l.as<double>(), (int64_t)SS, pck,
                   plan->d_l2l_in.as<int32_t>() + lv.in_off, plan->d_l2l_out.as<int32_t>() + lv.out_off, lv.ncol,
                   pck, 0, nullptr};
        HIPCK(launch_gemm_cols(g, s));
    }
    HIPCK(hipEventRecord(plan->ev[6], s));
    // L2P (fmm.cpp:231-257)
    {   // q_down = 0.5^depth * VSinv_l2p * (UT_l2p * p_check)   (kernel.cpp:1465-1467)
        const int32_t* iota = plan->d_iota.as<int32_t>();
        GemmCols g1{Sn, nleaf, 1, 0, plan->o_utl.as<double>(), 0, pck, plan->d_leaf_cell.as<int32_t>(), iota, 0,
                    plan->d_pcleaf2.as<double>(), 0, nullptr};
        HIPCK(launch_gemm_cols(g1, s));
        GemmCols g2{Sn, nleaf, 1, 0, plan->o_vsl.as<double>(), 0, plan->d_pcleaf2.as<double>(), iota, iota, 0,
                    plan->d_pcleaf.as<double>(), 0, plan->d_leaf_depth.as<int32_t>()};
        HIPCK(launch_gemm_cols(g2, s));
    }
    HIPCK(launch_l2p_eval(Sn, plan->o_usurf.as<double>(), plan->d_leaf_cell.as<int32_t>(), nleaf, cells, bod,
                          plan->d_pcleaf.as<double>(), plan->d_p.as<double>(), plan->d_gx.as<double>(),
                          plan->d_gy.as<double>(), plan->d_gz.as<double>(), s));
    HIPCK(hipEventRecord(plan->ev[7], s));
    // P2P (fmm.cpp:259-294)
    HIPCK(launch_p2p(plan->d_leaf_cell.as<int32_t>(), nleaf, cells, bod, plan->d_p2p_off.as<int64_t>(),
                     plan->d_p2p_src.as<int32_t>(), plan->d_p2p_shift.as<double>(), plan->d_p.as<double>(),
                     plan->d_gx.as<double>(), plan->d_gy.as<double>(), plan->d_gz.as<double>(), s));
    HIPCK(hipEventRecord(plan->ev[8], s));
    // dipole correction, 1/(4 pi), input order (fmm.cpp:75-80)
    const int nblk = 256;
    HIPCK(launch_dipole_partial(bod, n, plan->d_partial.as<double>(), nblk, s));
    HIPCK(launch_finalize(n, plan->d_perm.as<int32_t>(), plan->d_q.as<double>(), plan->d_p.as<double>(),
                          plan->d_gx.as<double>(), plan->d_gy.as<double>(), plan->d_gz.as<double>(),
                          plan->d_partial.as<double>(), nblk, plan->cfg.dipole_mode, plan->cfg.cycle,
                          plan->cfg.divide_4pi ? 1.0 / (4 * M_PI) : 1.0, d_phi, d_grad, s));
    HIPCK(hipEventRecord(plan->ev[9], s));
    plan->ev_recorded = true;
    return FMM3DPP_OK;
}

int fmm3dpp_evaluate(fmm3dpp_plan* plan, const double* q, double* phi, double* grad) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (!q || !phi || !grad) FAIL(FMM3DPP_ERR_ARG, "null array");
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate");
    HIPCK(hipSetDevice(plan->device));
    size_t n = (size_t)plan->st.n;
    HIPCK(hipMemcpyAsync(plan->d_qin.p, q, n * 8, hipMemcpyHostToDevice, plan->stream));
    int rc = fmm3dpp_evaluate_device(plan, plan->d_qin.as<double>(), plan->d_phi.as<double>(), plan->d_grad.as<double>());
    if (rc) return rc;
    HIPCK(hipMemcpyAsync(phi, plan->d_phi.p, n * 8, hipMemcpyDeviceToHost, plan->stream));
    HIPCK(hipMemcpyAsync(grad, plan->d_grad.p, n * 24, hipMemcpyDeviceToHost, plan->stream));
    HIPCK(hipStreamSynchronize(plan->stream));
    return FMM3DPP_OK;
}

int fmm3dpp_get_counts(const fmm3dpp_plan* plan, fmm3dpp_counts* out) {
    if (!plan || !out) return FMM3DPP_ERR_ARG;
    const Structure& S = plan->st;
    out->n_bodies = S.n;
    out->n_cells = (int64_t)S.cells.size();
    out->n_tree_cells = S.ntree;
    out->n_p2p = (int64_t)S.p2p.size();
    out->n_m2l = (int64_t)S.m2l.size();
    out->n_lost = S.lost;
    out->S = surface_count(plan->cfg.P);
    int N = 2 * plan->cfg.P - 1;
    out->n_freq = N * N * (N / 2 + 1);
    return FMM3DPP_OK;
}

int fmm3dpp_get_cells(const fmm3dpp_plan* plan, int32_t* ints, double* dbl) {
    if (!plan || !ints || !dbl) return FMM3DPP_ERR_ARG;
    const auto& C = plan->st.cells;
    for (size_t c = 0; c < C.size(); c++) {
        int32_t* o = ints + 7 * c;
        o[0] = (int32_t)c; o[1] = C[c].octant; o[2] = C[c].depth; o[3] = C[c].coff; o[4] = C[c].cnum;
        o[5] = C[c].boff; o[6] = C[c].bnum;
        double* d = dbl + 4 * c;
        d[0] = C[c].r; d[1] = C[c].x[0]; d[2] = C[c].x[1]; d[3] = C[c].x[2];
    }
    return FMM3DPP_OK;
}

int fmm3dpp_get_body_order(const fmm3dpp_plan* plan, int32_t* perm) {
    if (!plan || !perm) return FMM3DPP_ERR_ARG;
    std::memcpy(perm, plan->st.perm.data(), plan->st.perm.size() * sizeof(int32_t));
    return FMM3DPP_OK;
}

int fmm3dpp_get_p2p(const fmm3dpp_plan* plan, int32_t* tar, int32_t* src, double* off3) {
    if (!plan || !tar || !src || !off3) return FMM3DPP_ERR_ARG;
    const auto& R = plan->st.p2p;
    for (size_t i = 0; i < R.size(); i++) {
        tar[i] = R[i].tar; src[i] = R[i].src;
        off3[3 * i] = R[i].off[0]; off3[3 * i + 1] = R[i].off[1]; off3[3 * i + 2] = R[i].off[2];
    }
    return FMM3DPP_OK;
}

int fmm3dpp_get_m2l(const fmm3dpp_plan* plan, int32_t* tar, int32_t* src, double* off3, int32_t* img) {
    if (!plan || !tar || !src || !off3 || !img) return FMM3DPP_ERR_ARG;
    const auto& R = plan->st.m2l;
    for (size_t i = 0; i < R.size(); i++) {
        tar[i] = R[i].tar; src[i] = R[i].src; img[i] = R[i].img;
        off3[3 * i] = R[i].off[0]; off3[3 * i + 1] = R[i].off[1]; off3[3 * i + 2] = R[i].off[2];
    }
    return FMM3DPP_OK;
}

static std::vector<double>* op_named(fmm3dpp_plan* plan, const char* name) {
    Operators& o = plan->ops;
    std::string s(name ? name : "");
    if (s == "UT") return &o.UT;
    if (s == "VSinv") return &o.VSinv;
    if (s == "UT_l2p") return &o.UT_l2p;
    if (s == "VSinv_l2p") return &o.VSinv_l2p;
    if (s == "L2L_IMG") return &o.l2l_img;
    if (s == "GK8") return &o.gk8;
    if (s == "GK27") return &o.gk27;
    if (s == "M2M_IMG_SUM") return &o.m2m_img_sum;   // derived forms the device applies (read-only)
    if (s == "GK27_SUM") return &o.gk27_sum;
    if (s.rfind("M2M_IMG", 0) == 0) { int k = atoi(s.c_str() + 7); return (k >= 0 && k < 27) ? &o.m2m_img[k] : nullptr; }
    if (s.rfind("M2M", 0) == 0) { int k = atoi(s.c_str() + 3); return (k >= 0 && k < 8) ? &o.m2m[k] : nullptr; }
    if (s.rfind("L2L", 0) == 0) { int k = atoi(s.c_str() + 3); return (k >= 0 && k < 8) ? &o.l2l[k] : nullptr; }
    return nullptr;
}

int64_t fmm3dpp_get_operator(fmm3dpp_plan* plan, const char* name, double* out, int64_t cap) {
    if (!plan) return FMM3DPP_ERR_ARG;
    int rc = ensure_ops(plan);
    if (rc) return rc;
    auto* v = op_named(plan, name);
    if (!v) FAIL(FMM3DPP_ERR_ARG, std::string("unknown operator ") + (name ? name : ""));
    if (out && cap >= (int64_t)v->size()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (int64_t)v->size();
}

int fmm3dpp_set_operator(fmm3dpp_plan* plan, const char* name, const double* in, int64_t n) {
    if (!plan || !in) return FMM3DPP_ERR_ARG;
    int rc = ensure_ops(plan);
    if (rc) return rc;
    auto* v = op_named(plan, name);
    if (!v) FAIL(FMM3DPP_ERR_ARG, std::string("unknown operator ") + (name ? name : ""));
    std::string nm(name);
    if (nm == "M2M_IMG_SUM" || nm == "GK27_SUM")
        FAIL(FMM3DPP_ERR_ARG, "fused operators are derived from the factors and cannot be set");
    if ((int64_t)v->size() != n) FAIL(FMM3DPP_ERR_ARG, "operator size mismatch");
    v->assign(in, in + n);
    derive_fused(plan->ops);
    plan->ops_uploaded = false;
    if (plan->device >= 0) {
        HIPCK(hipSetDevice(plan->device));
        return upload_operators(plan);
    }
    return FMM3DPP_OK;
}

int fmm3dpp_get_stage_vectors(fmm3dpp_plan* plan, double* q_equiv, double* p_check) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "no particles");
    HIPCK(hipSetDevice(plan->device));
    size_t bytes = plan->st.cells.size() * (size_t)surface_count(plan->cfg.P) * 8;
    HIPCK(hipStreamSynchronize(plan->stream));
    if (q_equiv) HIPCK(hipMemcpy(q_equiv, plan->d_qeq.p, bytes, hipMemcpyDeviceToHost));
    if (p_check) HIPCK(hipMemcpy(p_check, plan->d_pck.p, bytes, hipMemcpyDeviceToHost));
    return FMM3DPP_OK;
}

double fmm3dpp_timing_ms(const fmm3dpp_plan* plan, const char* name) {
    if (!plan || !name) return -1.0;
    auto hit = plan->host_ms.find(name);
    if (hit != plan->host_ms.end()) return hit->second;
    if (plan->device < 0 || !plan->ev_recorded) return -1.0;
    auto* p = const_cast<fmm3dpp_plan*>(plan);
    if (hipEventSynchronize(p->ev[kStages]) != hipSuccess) return -1.0;
    std::string s(name);
    auto el = [&](int a, int b) {
        float ms = 0;
        return hipEventElapsedTime(&ms, p->ev[a], p->ev[b]) == hipSuccess ? (double)ms : -1.0;
    };
    if (s == "evaluate") return el(0, kStages);
    for (int i = 0; i < kStages; i++)
        if (s == kStageNames[i]) {
            // events: 0 P2M 1 M2M 2 up 3 hadamard 4 down 5 L2L 6 L2P 7 P2P 8 finalize 9
            return el(i, i + 1);
        }
    return -1.0;
}

}  // extern "C"
