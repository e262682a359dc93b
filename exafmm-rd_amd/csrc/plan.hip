// C-ABI of the MI355X periodic KIFMM (include/fmm3dpp.h): plan lifecycle, operator upload, the
// GPU tree/list build and the evaluation stage sequence of LaplaceFMM::solve (3dpp/src/fmm.cpp:13-81).
// Plans created without a device (FMM3DPP_DEVICE_NONE) build the same tree and lists on the host
// for inspection only; they cannot evaluate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "plan.h"

using namespace fmm3dpp;

namespace fmm3dpp {

int validate_config(const fmm3dpp_config* c, std::string& err) {
    if (!c) { err = "null config"; return FMM3DPP_ERR_ARG; }
    if (c->P < 2 || c->P > 12) { err = "P must be in [2, 12]"; return FMM3DPP_ERR_ARG; }
    if (c->ncrit < 1) { err = "ncrit must be >= 1"; return FMM3DPP_ERR_ARG; }
    if (c->images < 0 || c->images > 20) { err = "images must be in [0, 20]"; return FMM3DPP_ERR_ARG; }
    if (!(c->cycle > 0) || !std::isfinite(c->cycle)) { err = "cycle must be positive and finite"; return FMM3DPP_ERR_ARG; }
    for (int k = 0; k < 3; k++)
        if (!std::isfinite(c->center[k])) { err = "center must be finite"; return FMM3DPP_ERR_ARG; }
    if (c->precision != FMM3DPP_PREC_FP64 && c->precision != FMM3DPP_PREC_MIXED) { err = "unknown precision"; return FMM3DPP_ERR_ARG; }
    if (c->dipole_mode != FMM3DPP_DIPOLE_OFF && c->dipole_mode != FMM3DPP_DIPOLE_REFERENCE &&
        c->dipole_mode != FMM3DPP_DIPOLE_PHYSICAL) { err = "unknown dipole mode"; return FMM3DPP_ERR_ARG; }
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int upload_operators(fmm3dpp_plan* plan) {
    HIPCK(plan->dop.upload(plan->ops, plan->mixed(), plan->apply_f32, plan->stream));
    HIPCK(hipStreamSynchronize(plan->stream));
    plan->ops_uploaded = true;
    return FMM3DPP_OK;
}

}  // namespace

namespace fmm3dpp {

int ensure_ops(fmm3dpp_plan* plan) {
    if (!plan->ops_ready) {
        double t0 = now_ms();
        int rc = build_operators_cached(plan->cfg.P, plan->cfg.images, plan->cfg.cycle, plan->ops, plan->err);
        if (rc) return rc;
        plan->host_ms["precompute"] = now_ms() - t0;
        plan->ops_ready = true;
        plan->ops_uploaded = false;
    }
    if (plan->device >= 0 && !plan->ops_uploaded) return upload_operators(plan);
    return FMM3DPP_OK;
}

int ensure_host_structure(fmm3dpp_plan* plan) {
    if (plan->host_struct_valid) return FMM3DPP_OK;
    if (plan->device < 0 || !plan->particles) FAIL(FMM3DPP_ERR_STATE, "no particles");
    HIPCK(hipSetDevice(plan->device));
    int rc = plan->gs.export_structure(plan->cfg, plan->st, plan->stream, plan->err);
    if (rc) return rc;
    plan->host_struct_valid = true;
    return FMM3DPP_OK;
}

// P2M (fmm.cpp:83-109, kernel.cpp:332-371): check potentials of the leaves' bodies, then
// q_equiv = 0.5^depth * VSinv * (UT * p_check), UT and VSinv at the truncated rank on the mixed path.
int stage_p2m_leaves(fmm3dpp_plan* plan, const int32_t* leaf_cell, const int32_t* leaf_depth, int nleaf) {
    hipStream_t s = plan->stream;
    const int Sn = plan->ops.S;
    const int32_t* iota = plan->gs.iota.as<int32_t>();
    HIPCK(launch_p2m_check(Sn, plan->dop.o_usurf.as<double>(), leaf_cell, nleaf, plan->cells_dev(), plan->bodies_dev(),
                           plan->d_pcleaf.as<double>(), s, plan->f32_applies()));
    const ApplyCols g1{Sn, nleaf, 1, false, plan->d_pcleaf.as<double>(), iota, iota, 0, plan->d_pcleaf2.as<double>(), 0,
                       nullptr};
    HIPCK(launch_apply(g1, plan->dop.lo_ut, s));
    const ApplyCols g2{Sn, nleaf, 1, false, plan->d_pcleaf2.as<double>(), iota, leaf_cell, 0, plan->d_qeq.as<double>(), 0,
                       leaf_depth};
    HIPCK(launch_apply(g2, plan->dop.lo_vs, s));
    if (plan->f32_applies())
        HIPCK(launch_monopole_leaves(plan->d_qeq.as<double>(), Sn, leaf_cell, nleaf, plan->cells_dev(),
                                     plan->d_q.as<double>(), plan->d_qcell.as<double>(), s));
    return FMM3DPP_OK;
}

int total_charge(fmm3dpp_plan* plan, const double** qtot) {
    hipStream_t s = plan->stream;
    const int nblk = 256;
    double* part = plan->d_partial.as<double>();
    HIPCK(launch_dipole_partial(plan->bodies_dev(), (int)plan->gs.n, part, nblk, s));
    HIPCK(launch_moments(part, nblk, part + (size_t)nblk * 5, s));
    *qtot = part + (size_t)nblk * 5 + 4;
    return FMM3DPP_OK;
}

// M2M (fmm.cpp:111-150, kernel.cpp:395-404, 434-442): per tree level from the deepest, one apply per
// child octant into partials and the child sum (and, with float32 applies, each parent's monopole set
// to the sum of its children's charges); then the image levels.  In the infinite-lattice
// gauge the root's and every image cell's monopole are set to their charge (the bodies' total times
// 27^m for c(-m)), removing the rounding-level monopole the pseudo-inverse leaves.
int stage_m2m_all(fmm3dpp_plan* plan) {
    hipStream_t s = plan->stream;
    const GpuStructure& g = plan->gs;
    const int Sn = plan->ops.S;
    double* qeq = plan->d_qeq.as<double>();
    const int32_t* iota = g.iota.as<int32_t>();
    for (const auto& lv : g.m2m_levels) {
        const ApplyCols gm{Sn, lv.ncol, 8, false, qeq, g.m2m_in.as<int32_t>() + lv.in_off, iota, lv.ncol,
                           plan->d_m2m_part.as<double>(), 0, nullptr};
        HIPCK(launch_apply(gm, plan->dop.lo_m2m, s));
        HIPCK(launch_sum_partials(Sn, lv.ncol, 8, plan->d_m2m_part.as<double>(), g.m2m_out.as<int32_t>() + lv.out_off,
                                  qeq, s));
        if (plan->f32_applies())
            HIPCK(launch_monopole_parents(qeq, Sn, g.m2m_out.as<int32_t>() + lv.out_off, lv.ncol,
                                          g.m2m_in.as<int32_t>() + lv.in_off, plan->d_qcell.as<double>(), s));
    }
    const double* qtot = nullptr;   // the bodies' total charge (infinite-lattice gauge)
    if (plan->cfg.dipole_mode == FMM3DPP_DIPOLE_PHYSICAL) {
        int rc = total_charge(plan, &qtot);
        if (rc) return rc;
        HIPCK(launch_neutralize(qeq, Sn, nullptr, 0, 1, qtot, 0, 1.0, s));
    }
    for (size_t i = 0; i < g.img_m2m.size(); i++) {
        const ApplyCols gi{Sn, 1, 1, false, qeq, g.img_b.as<int32_t>() + i, g.img_a.as<int32_t>() + i, 0, qeq, 0, nullptr};
        HIPCK(launch_apply(gi, plan->dop.lo_m2m_img, s));
        if (qtot) HIPCK(launch_neutralize(qeq, Sn, g.img_a.as<int32_t>() + i, 0, 1, qtot, 0, std::pow(27.0, i + 1), s));
    }
    return FMM3DPP_OK;
}

// M2L (fmm.cpp:152-193, kernel.cpp:1140-1397): up transform, 8x8 Hadamard, image lattice 1x27,
// down transform accumulating into p_check.  Targets = sources = the tree's non-leaf cells [0, nl),
// then the image cells [nl, ntar).  Records ev[3] after the up transform and ev[4] after the 8x8
// Hadamard (stage timers).
int stage_m2l_all(fmm3dpp_plan* plan) {
    hipStream_t s = plan->stream;
    const Operators& o = plan->ops;
    const GpuStructure& g = plan->gs;
    const int Sn = o.S;
    double* qeq = plan->d_qeq.as<double>();
    double* pck = plan->d_pck.as<double>();
    const int ntar = g.ntar, nl = g.nl, nimg = g.nimg;
    const int32_t* slots = g.child_slots.as<int32_t>();
    const int32_t* inv = plan->dop.o_invlat.as<int32_t>();
    const double* tw = plan->dop.o_tw.as<double>();
    if (!plan->mixed()) {
        HIPCK(launch_m2l_up(o.P, false, Sn, ntar, slots, inv, tw, qeq, plan->d_Qk.p, s));
        HIPCK(hipEventRecord(plan->ev[3], s));
        HIPCK(launch_hadamard8(false, o.F, ntar, ntar, g.slot_src_T.as<int32_t>(), g.had(), plan->dop.o_gkA.p,
                               plan->d_Qk.p, plan->d_Pk.p, s, plan->had_cap));
        HIPCK(hipEventRecord(plan->ev[4], s));
        HIPCK(launch_hadamard27(o.F, nimg, g.img_tar.as<int32_t>(), g.img_src_T.as<int32_t>(), ntar, 0,
                                plan->dop.o_gk27.as<double>(), plan->d_Qk.as<double>(), plan->d_Pk.as<double>(), s));
        HIPCK(launch_m2l_down(o.P, false, Sn, ntar, slots, g.tar_scale.as<double>(), inv, tw, plan->d_Pk.p, pck, s));
    } else {
        // float32 spectra and Hadamard for the tree; the image lattice stays float64 end to end
        HIPCK(launch_m2l_up(o.P, true, Sn, nl, slots, inv, tw, qeq, plan->d_Qk.p, s));
        HIPCK(hipEventRecord(plan->ev[3], s));
        HIPCK(launch_hadamard8(true, o.F, nl, ntar, g.slot_src_T.as<int32_t>(), g.had(), plan->dop.o_gkA.p,
                               plan->d_Qk.p, plan->d_Pk.p, s, plan->had_cap));
        HIPCK(hipEventRecord(plan->ev[4], s));
        HIPCK(launch_m2l_down(o.P, true, Sn, nl, slots, g.tar_scale.as<double>(), inv, tw, plan->d_Pk.p, pck, s));
        if (nimg > 0) {
            HIPCK(launch_m2l_up(o.P, false, Sn, nimg, slots + (size_t)nl * 8, inv, tw, qeq, plan->d_Qk_img.p, s));
            HIPCK(launch_hadamard27(o.F, nimg, g.img_tar.as<int32_t>(), g.img_src_T.as<int32_t>(), ntar, nl,
                                    plan->dop.o_gk27.as<double>(), plan->d_Qk_img.as<double>(),
                                    plan->d_Pk_img.as<double>(), s));
            HIPCK(launch_m2l_down(o.P, false, Sn, nimg, slots + (size_t)nl * 8, g.tar_scale.as<double>() + nl, inv,
                                  tw, plan->d_Pk_img.p, pck, s));
        }
    }
    return FMM3DPP_OK;
}

// L2L (fmm.cpp:195-229, kernel.cpp:1419-1435): image levels from the outermost image, then one
// launch per tree level, each child's p_check += L2L[octant] * its parent's.
int stage_l2l_all(fmm3dpp_plan* plan) {
    hipStream_t s = plan->stream;
    const GpuStructure& g = plan->gs;
    const int Sn = plan->ops.S;
    double* pck = plan->d_pck.as<double>();
    const size_t nimgm2m = g.img_m2m.size();
    for (size_t i = 0; i < g.img_l2l.size(); i++) {
        const ApplyCols gi{Sn, 1, 1, true, pck, g.img_a.as<int32_t>() + nimgm2m + i, g.img_b.as<int32_t>() + nimgm2m + i,
                           0, pck, 0, nullptr};
        HIPCK(launch_apply(gi, plan->dop.lo_l2l_img, s));
    }
    for (const auto& lv : g.l2l_levels) {
        const ApplyCols gl{Sn, lv.ncol, 8, true, pck, g.l2l_in.as<int32_t>() + lv.in_off,
                           g.l2l_out.as<int32_t>() + lv.in_off, lv.ncol, pck, 0, nullptr};
        HIPCK(launch_apply(gl, plan->dop.lo_l2l, s));
    }
    return FMM3DPP_OK;
}

// L2P (fmm.cpp:231-257, kernel.cpp:1458-1511): q_down = 0.5^depth * VSinv_l2p * (UT_l2p * p_check),
// then potential and gradient at the leaves' bodies (written, not accumulated).
int stage_l2p_leaves(fmm3dpp_plan* plan, const int32_t* leaf_cell, const int32_t* leaf_depth, int nleaf) {
    hipStream_t s = plan->stream;
    const int Sn = plan->ops.S;
    const int32_t* iota = plan->gs.iota.as<int32_t>();
    const ApplyCols g1{Sn, nleaf, 1, false, plan->d_pck.as<double>(), leaf_cell, iota, 0, plan->d_pcleaf2.as<double>(), 0,
                       nullptr};
    HIPCK(launch_apply(g1, plan->dop.lo_utl, s));
    const ApplyCols g2{Sn, nleaf, 1, false, plan->d_pcleaf2.as<double>(), iota, iota, 0, plan->d_pcleaf.as<double>(), 0,
                       leaf_depth};
    HIPCK(launch_apply(g2, plan->dop.lo_vsl, s));
    HIPCK(launch_l2p_eval(Sn, plan->dop.o_usurf.as<double>(), leaf_cell, nleaf, plan->cells_dev(), plan->bodies_dev(),
                          plan->d_pcleaf.as<double>(), plan->d_p.as<double>(), plan->d_gx.as<double>(),
                          plan->d_gy.as<double>(), plan->d_gz.as<double>(), s, plan->f32_applies()));
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp

namespace {

// Device tree/lists plus the per-evaluate work buffers sized for them.
// with_lists = false leaves the near-field lists to GpuStructure::build_lists (fmm3dpp_solve_device).
// sync = false leaves the tree's last kernels queued on the plan's stream (solve_device queues the
// far field behind them); every host-side count the evaluation needs is known on return either way.
int build_device(fmm3dpp_plan* plan, int64_t n, const double* dx, const double* dy, const double* dz,
                 bool with_lists = true, bool sync = true) {
    hipStream_t s = plan->stream;
    int rc = with_lists ? plan->gs.build(plan->cfg, n, dx, dy, dz, s, plan->err)
                        : plan->gs.build_tree(plan->cfg, n, dx, dy, dz, s, plan->err);
    if (rc) return rc;
    const GpuStructure& g = plan->gs;
    const int Sn = surface_count(plan->cfg.P);
    const int N = 2 * plan->cfg.P - 1;
    const size_t F = (size_t)N * N * (N / 2 + 1);
    const size_t nc = (size_t)g.ncells;
    HIPCK(plan->d_q.ensure(n * 8)); HIPCK(plan->d_qin.ensure(n * 8));
    HIPCK(plan->d_p.ensure(n * 8)); HIPCK(plan->d_gx.ensure(n * 8)); HIPCK(plan->d_gy.ensure(n * 8)); HIPCK(plan->d_gz.ensure(n * 8));
    HIPCK(plan->d_pn.ensure(n * 8)); HIPCK(plan->d_gxn.ensure(n * 8)); HIPCK(plan->d_gyn.ensure(n * 8));
    HIPCK(plan->d_gzn.ensure(n * 8));
    HIPCK(plan->d_phi.ensure(n * 8)); HIPCK(plan->d_grad.ensure(n * 24));
    HIPCK(plan->d_pcleaf.ensure((size_t)g.nleaf * Sn * 8));
    HIPCK(plan->d_pcleaf2.ensure((size_t)g.nleaf * Sn * 8));
    HIPCK(plan->d_qeq.ensure(nc * Sn * 8));
    HIPCK(hipMemsetAsync(plan->d_qeq.p, 0, nc * Sn * 8, s));
    HIPCK(plan->d_pck.ensure(nc * Sn * 8));
    HIPCK(plan->d_qcell.ensure(nc * 8));
    if (plan->mixed()) {   // float32 spectra of the tree parents, float64 spectra of the image cells
        HIPCK(plan->d_Qk.ensure(F * std::max(g.nl, 1) * 8 * 8));
        HIPCK(plan->d_Pk.ensure(F * std::max(g.nl, 1) * 8 * 8));
        HIPCK(plan->d_Qk_img.ensure(F * std::max(g.nimg, 1) * 8 * 16));
        HIPCK(plan->d_Pk_img.ensure(F * std::max(g.nimg, 1) * 8 * 16));
        HIPCK(plan->d_rel.ensure((size_t)n * 16));
    } else {
        HIPCK(plan->d_Qk.ensure(F * std::max(g.ntar, 1) * 8 * 16));
        HIPCK(plan->d_Pk.ensure(F * std::max(g.ntar, 1) * 8 * 16));
    }
    HIPCK(plan->d_partial.ensure((256 + 1) * 5 * 8));   // 256 block partials and their sum
    HIPCK(plan->d_m2m_part.ensure((size_t)8 * std::max(1, g.max_level_cols) * Sn * 8));
    if (sync) HIPCK(hipStreamSynchronize(s));
    plan->host_ms["tree_and_lists"] = g.ms_build;
    return FMM3DPP_OK;
}

int launch_near(fmm3dpp_plan* plan);

// Far field of an evaluation (fmm.cpp:50-74 up to L2P): gather, P2M, M2M, M2L, L2L, L2P, on the
// plan's stream.  Needs the tree (build_tree), not the near-field lists.  near_now: the lists are
// built, so the P2P is queued on the side stream right after the M2L.
int eval_far(fmm3dpp_plan* plan, const double* d_qin, bool near_now) {
    hipStream_t s = plan->stream;
    const GpuStructure& g = plan->gs;
    const int Sn = plan->ops.S;
    HIPCK(hipEventRecord(plan->ev[0], s));
    HIPCK(launch_gather(d_qin, g.perm.as<int32_t>(), plan->d_q.as<double>(), (int)g.n, s));
    HIPCK(hipMemsetAsync(plan->d_pck.p, 0, (size_t)g.ncells * Sn * sizeof(double), s));
    int rc = stage_p2m_leaves(plan, g.leaf_cell.as<int32_t>(), g.leaf_depth.as<int32_t>(), g.nleaf);
    if (rc) return rc;
    HIPCK(hipEventRecord(plan->ev[1], s));
    if ((rc = stage_m2m_all(plan))) return rc;
    HIPCK(hipEventRecord(plan->ev[2], s));
    if ((rc = stage_m2l_all(plan))) return rc;
    HIPCK(hipEventRecord(plan->ev[5], s));
    if (near_now && (rc = launch_near(plan))) return rc;
    if ((rc = stage_l2l_all(plan))) return rc;
    HIPCK(hipEventRecord(plan->ev[6], s));
    if ((rc = stage_l2p_leaves(plan, g.leaf_cell.as<int32_t>(), g.leaf_depth.as<int32_t>(), g.nleaf))) return rc;
    HIPCK(hipEventRecord(plan->ev[7], s));
    return FMM3DPP_OK;
}

// Near field (P2P, fmm.cpp:259-294) on the side stream, from the end of the M2L on
// (ev[kEvForkNear], recorded after the charges are in tree order); stored apart from the far field,
// which the finalize adds.  Needs the lists.
int launch_near(fmm3dpp_plan* plan) {
    HIPCK(plan->ensure_near_stream());
    hipStream_t s = plan->near_stream();
    const GpuStructure& g = plan->gs;
    const int nleaf = g.nleaf;
    const CellsDev cells = plan->cells_dev();
    const BodiesDev bod = plan->bodies_dev();
    double *pn = plan->d_pn.as<double>(), *gxn = plan->d_gxn.as<double>(), *gyn = plan->d_gyn.as<double>(),
           *gzn = plan->d_gzn.as<double>();
    HIPCK(hipStreamWaitEvent(s, plan->ev[kEvForkNear], 0));
    HIPCK(hipEventRecord(plan->ev[kEvNear0], s));
    if (!plan->mixed()) {
        HIPCK(launch_p2p(g.leaf_cell.as<int32_t>(), nleaf, cells, bod, g.p2p_off.as<int64_t>(),
                         g.p2p_ent.as<int32_t>(), plan->cfg.cycle, pn, gxn, gyn, gzn, false, s));
    } else {
        HIPCK(launch_pack_rel(g.leaf_cell.as<int32_t>(), nleaf, cells, bod, plan->d_rel.as<float4>(), s));
        HIPCK(launch_p2p_f32(g.leaf_cell.as<int32_t>(), nleaf, cells, plan->d_rel.as<float4>(),
                             g.p2p_off.as<int64_t>(), g.p2p_ent.as<int32_t>(), plan->cfg.cycle, pn, gxn, gyn, gzn,
                             false, s));
    }
    HIPCK(hipEventRecord(plan->ev[kEvNear1], s));
    return FMM3DPP_OK;
}

// Join of the near field, dipole correction, 1/(4 pi), input order (fmm.cpp:75-80).
int eval_final(fmm3dpp_plan* plan, double* d_phi, double* d_grad) {
    hipStream_t s = plan->stream;
    const GpuStructure& g = plan->gs;
    const int n = (int)g.n;
    const BodiesDev bod = plan->bodies_dev();
    HIPCK(hipStreamWaitEvent(s, plan->ev[kEvNear1], 0));
    HIPCK(hipEventRecord(plan->ev[8], s));
    const int nblk = 256;
    const NearDev near{plan->d_pn.as<double>(), plan->d_gxn.as<double>(), plan->d_gyn.as<double>(),
                       plan->d_gzn.as<double>()};
    HIPCK(launch_dipole_partial(bod, n, plan->d_partial.as<double>(), nblk, s));
    HIPCK(launch_finalize(n, g.perm.as<int32_t>(), bod, plan->d_p.as<double>(), plan->d_gx.as<double>(),
                          plan->d_gy.as<double>(), plan->d_gz.as<double>(), near, plan->d_partial.as<double>(), nblk,
                          plan->cfg.dipole_mode, plan->cfg.cycle, plan->cfg.divide_4pi ? 1.0 / (4 * M_PI) : 1.0,
                          d_phi, d_grad, s));
    HIPCK(hipEventRecord(plan->ev[9], s));
    plan->ev_recorded = true;
    return FMM3DPP_OK;
}

}  // namespace

extern "C" {

int fmm3dpp_plan_create(const fmm3dpp_config* cfg, int device, fmm3dpp_plan** out) {
    if (!out) return FMM3DPP_ERR_ARG;
    *out = nullptr;
    std::string err;
    int rc = validate_config(cfg, err);
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
        for (int i = 0; i < kEvents; i++)
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
        if (plan->side_stream) (void)hipStreamSynchronize(plan->side_stream);
        for (int i = 0; i < kEvents; i++) if (plan->ev[i]) (void)hipEventDestroy(plan->ev[i]);
        if (plan->own_stream && plan->stream) (void)hipStreamDestroy(plan->stream);
        if (plan->side_stream) (void)hipStreamDestroy(plan->side_stream);
    }
    delete plan;
    return FMM3DPP_OK;
}

int fmm3dpp_set_near_field_stream(fmm3dpp_plan* plan, int enable) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device >= 0) {   // work already queued keeps its order
        HIPCK(hipSetDevice(plan->device));
        if (plan->side_stream) HIPCK(hipStreamSynchronize(plan->side_stream));
        HIPCK(hipStreamSynchronize(plan->stream));
    }
    plan->near_side = enable != 0;
    return FMM3DPP_OK;
}

int fmm3dpp_set_hadamard_staging(fmm3dpp_plan* plan, int enable) {
    if (!plan) return FMM3DPP_ERR_ARG;
    plan->had_cap = enable ? kHadCap : -1;
    return FMM3DPP_OK;
}

int fmm3dpp_set_mixed_applies(fmm3dpp_plan* plan, int float32) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->apply_f32 != (float32 != 0)) {
        plan->apply_f32 = float32 != 0;
        plan->ops_uploaded = false;   // the operators' device layout depends on it
    }
    return FMM3DPP_OK;
}

const char* fmm3dpp_last_error(const fmm3dpp_plan* plan) { return plan ? plan->err.c_str() : "null plan"; }

int fmm3dpp_set_stream(fmm3dpp_plan* plan, void* s) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan has no device");
    if (plan->own_stream && plan->stream) {
        (void)hipStreamSynchronize(plan->stream);
        (void)hipStreamDestroy(plan->stream);
    }
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
    if (n < 1 || n > (int64_t)INT32_MAX / 2) FAIL(FMM3DPP_ERR_ARG, "number of bodies must be in [1, 2^30)");
    plan->particles = false;
    plan->host_struct_valid = false;
    plan->host_lists_valid = false;
    if (plan->device < 0) {
        int rc = build_structure(plan->cfg, n, x, y, z, plan->st, plan->err);
        if (rc) return rc;
        plan->host_ms["tree_and_lists"] = plan->st.ms_tree + plan->st.ms_p2p + plan->st.ms_m2l;
        plan->host_struct_valid = true;
        plan->particles = true;
        return FMM3DPP_OK;
    }
    HIPCK(hipSetDevice(plan->device));
    hipStream_t s = plan->stream;
    HIPCK(plan->d_inx.ensure(n * 8)); HIPCK(plan->d_iny.ensure(n * 8)); HIPCK(plan->d_inz.ensure(n * 8));
    HIPCK(hipMemcpyAsync(plan->d_inx.p, x, n * 8, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(plan->d_iny.p, y, n * 8, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(plan->d_inz.p, z, n * 8, hipMemcpyHostToDevice, s));
    int rc = build_device(plan, n, plan->d_inx.as<double>(), plan->d_iny.as<double>(), plan->d_inz.as<double>());
    if (rc) return rc;
    plan->particles = true;
    return FMM3DPP_OK;
}

int fmm3dpp_set_particles_device(fmm3dpp_plan* plan, int64_t n, const double* d_x, const double* d_y, const double* d_z) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!d_x || !d_y || !d_z) FAIL(FMM3DPP_ERR_ARG, "null position array");
    if (n < 1 || n > (int64_t)INT32_MAX / 2) FAIL(FMM3DPP_ERR_ARG, "number of bodies must be in [1, 2^30)");
    plan->particles = false;
    plan->host_struct_valid = false;
    plan->host_lists_valid = false;
    HIPCK(hipSetDevice(plan->device));
    int rc = build_device(plan, n, d_x, d_y, d_z);
    if (rc) return rc;
    plan->particles = true;
    return FMM3DPP_OK;
}

int fmm3dpp_evaluate_device(fmm3dpp_plan* plan, const double* d_qin, double* d_phi, double* d_grad) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate");
    if (!d_qin || !d_phi || !d_grad) FAIL(FMM3DPP_ERR_ARG, "null array");
    HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    rc = eval_far(plan, d_qin, true);
    if (rc) return rc;
    return eval_final(plan, d_phi, d_grad);
}

int fmm3dpp_solve_device(fmm3dpp_plan* plan, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                         const double* d_q, double* d_phi, double* d_grad) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!d_x || !d_y || !d_z || !d_q || !d_phi || !d_grad) FAIL(FMM3DPP_ERR_ARG, "null array");
    if (n < 1 || n > (int64_t)INT32_MAX / 2) FAIL(FMM3DPP_ERR_ARG, "number of bodies must be in [1, 2^30)");
    plan->particles = false;
    plan->host_struct_valid = false;
    plan->host_lists_valid = false;
    HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    const double t0 = now_ms();
    rc = build_device(plan, n, d_x, d_y, d_z, false, false);   // tree on the plan's stream
    if (rc) return rc;
    plan->host_ms["tree"] = now_ms() - t0;
    rc = eval_far(plan, d_q, false);                    // queued on the plan's stream
    if (rc) return rc;
    HIPCK(plan->ensure_near_stream());
    rc = plan->gs.build_lists(plan->cfg, plan->near_stream(), plan->err);   // host waits on that stream only
    if (rc) return rc;
    plan->host_ms.erase("tree_and_lists");   // the lists overlap the far field: no serial figure
    plan->particles = true;
    rc = launch_near(plan);
    if (rc) return rc;
    return eval_final(plan, d_phi, d_grad);
}

int fmm3dpp_evaluate(fmm3dpp_plan* plan, const double* q, double* phi, double* grad) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (!q || !phi || !grad) FAIL(FMM3DPP_ERR_ARG, "null array");
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate");
    HIPCK(hipSetDevice(plan->device));
    size_t n = (size_t)plan->gs.n;
    HIPCK(hipMemcpyAsync(plan->d_qin.p, q, n * 8, hipMemcpyHostToDevice, plan->stream));
    int rc = fmm3dpp_evaluate_device(plan, plan->d_qin.as<double>(), plan->d_phi.as<double>(), plan->d_grad.as<double>());
    if (rc) return rc;
    HIPCK(hipMemcpyAsync(phi, plan->d_phi.p, n * 8, hipMemcpyDeviceToHost, plan->stream));
    HIPCK(hipMemcpyAsync(grad, plan->d_grad.p, n * 24, hipMemcpyDeviceToHost, plan->stream));
    HIPCK(hipStreamSynchronize(plan->stream));
    return FMM3DPP_OK;
}

int fmm3dpp_get_counts(const fmm3dpp_plan* cplan, fmm3dpp_counts* out) {
    if (!cplan || !out) return FMM3DPP_ERR_ARG;
    auto* plan = const_cast<fmm3dpp_plan*>(cplan);
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "no particles");
    int rc = plan->device >= 0 ? ensure_host_structure(plan) : FMM3DPP_OK;
    if (rc) return rc;
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

#define NEED_HOST_STRUCTURE(cplan)                                                      \
    auto* plan = const_cast<fmm3dpp_plan*>(cplan);                                      \
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "no particles");                      \
    if (plan->device >= 0) {                                                            \
        int rc_ = ensure_host_structure(plan);                                          \
        if (rc_) return rc_;                                                            \
    }

int fmm3dpp_get_cells(const fmm3dpp_plan* cplan, int32_t* ints, double* dbl) {
    if (!cplan || !ints || !dbl) return FMM3DPP_ERR_ARG;
    NEED_HOST_STRUCTURE(cplan);
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

int fmm3dpp_get_body_order(const fmm3dpp_plan* cplan, int32_t* perm) {
    if (!cplan || !perm) return FMM3DPP_ERR_ARG;
    NEED_HOST_STRUCTURE(cplan);
    std::memcpy(perm, plan->st.perm.data(), plan->st.perm.size() * sizeof(int32_t));
    return FMM3DPP_OK;
}

int fmm3dpp_get_p2p(const fmm3dpp_plan* cplan, int32_t* tar, int32_t* src, double* off3) {
    if (!cplan || !tar || !src || !off3) return FMM3DPP_ERR_ARG;
    NEED_HOST_STRUCTURE(cplan);
    const auto& R = plan->st.p2p;
    for (size_t i = 0; i < R.size(); i++) {
        tar[i] = R[i].tar; src[i] = R[i].src;
        off3[3 * i] = R[i].off[0]; off3[3 * i + 1] = R[i].off[1]; off3[3 * i + 2] = R[i].off[2];
    }
    return FMM3DPP_OK;
}

int fmm3dpp_get_m2l(const fmm3dpp_plan* cplan, int32_t* tar, int32_t* src, double* off3, int32_t* img) {
    if (!cplan || !tar || !src || !off3 || !img) return FMM3DPP_ERR_ARG;
    NEED_HOST_STRUCTURE(cplan);
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
    if (plan->device >= 0) HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    auto* v = op_named(plan, name);
    if (!v) FAIL(FMM3DPP_ERR_ARG, std::string("unknown operator ") + (name ? name : ""));
    if (out && cap >= (int64_t)v->size()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (int64_t)v->size();
}

int fmm3dpp_set_operator(fmm3dpp_plan* plan, const char* name, const double* in, int64_t n) {
    if (!plan || !in) return FMM3DPP_ERR_ARG;
    if (plan->device >= 0) HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    auto* v = op_named(plan, name);
    if (!v) FAIL(FMM3DPP_ERR_ARG, std::string("unknown operator ") + (name ? name : ""));
    std::string nm(name);
    if (nm == "M2M_IMG_SUM" || nm == "GK27_SUM")
        FAIL(FMM3DPP_ERR_ARG, "derived operators are recomputed from the factors and cannot be set");
    if ((int64_t)v->size() != n) FAIL(FMM3DPP_ERR_ARG, "operator size mismatch");
    v->assign(in, in + n);
    derive_fused(plan->ops);
    plan->ops_uploaded = false;
    if (plan->device >= 0) return upload_operators(plan);
    return FMM3DPP_OK;
}

int fmm3dpp_get_stage_vectors(fmm3dpp_plan* plan, double* q_equiv, double* p_check) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "no particles");
    HIPCK(hipSetDevice(plan->device));
    size_t bytes = (size_t)plan->gs.ncells * (size_t)surface_count(plan->cfg.P) * 8;
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
    if (s == "P2P") return el(kEvNear0, kEvNear1);   // on the side stream, beside the far field
    if (s == "near_start") return el(0, kEvNear0);   // far-field start to P2P start
    if (s == "far_end") return el(0, 7);             // far-field start to the end of L2P
    for (int i = 0; i < kStages; i++)
        if (s == kStageNames[i]) return el(i, i + 1);
    return -1.0;
}

int fmm3dpp_save_operators(fmm3dpp_plan* plan, const char* path) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (!path) FAIL(FMM3DPP_ERR_ARG, "null path");
    if (plan->device >= 0) HIPCK(hipSetDevice(plan->device));
    int rc = ensure_ops(plan);
    if (rc) return rc;
    return save_operators(plan->ops, plan->cfg.cycle, path, plan->err);
}

int fmm3dpp_load_operators(fmm3dpp_plan* plan, const char* path) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (!path) FAIL(FMM3DPP_ERR_ARG, "null path");
    double t0 = now_ms();
    Operators o;
    int rc = load_operators(o, plan->cfg.P, plan->cfg.images, plan->cfg.cycle, path, plan->err);
    if (rc) return rc;
    plan->ops = std::move(o);
    plan->ops_ready = true;
    plan->ops_uploaded = false;
    plan->host_ms["precompute"] = now_ms() - t0;
    if (plan->device >= 0) {
        HIPCK(hipSetDevice(plan->device));
        return upload_operators(plan);
    }
    return FMM3DPP_OK;
}

int fmm3dpp_evaluate_adjoint_device(fmm3dpp_plan* plan, const double* d_q, const double* d_w, const double* d_grad,
                                    double* d_dq, double* d_dx) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate_adjoint");
    if (!d_q || !d_w || !d_grad || !d_dq || !d_dx) FAIL(FMM3DPP_ERR_ARG, "null array");
    if (plan->cfg.dipole_mode == FMM3DPP_DIPOLE_REFERENCE)
        FAIL(FMM3DPP_ERR_ARG, "the reference dipole correction (phi -= c |M|^2 / (n q_i)) has no adjoint; "
                              "use FMM3DPP_DIPOLE_OFF or FMM3DPP_DIPOLE_PHYSICAL");
    HIPCK(hipSetDevice(plan->device));
    const int64_t n = plan->gs.n;
    HIPCK(plan->d_adj_phi.ensure(n * 8));
    HIPCK(plan->d_adj_grad.ensure(n * 24));
    // the lattice kernel is symmetric: A^T w is the evaluation with charges w, up to the gauge term
    int rc = fmm3dpp_evaluate_device(plan, d_w, plan->d_adj_phi.as<double>(), plan->d_adj_grad.as<double>());
    if (rc) return rc;
    const double cyc = plan->cfg.cycle;
    const double c = plan->cfg.dipole_mode == FMM3DPP_DIPOLE_PHYSICAL ? 4 * M_PI / (3 * cyc * cyc * cyc) : 0.0;
    const double out_scale = plan->cfg.divide_4pi ? 1.0 / (4 * M_PI) : 1.0;
    HIPCK(launch_adjoint_combine((int)n, plan->gs.perm.as<int32_t>(), plan->bodies_dev(), plan->d_partial.as<double>(),
                                 256, c, out_scale, d_q, d_w, d_grad, plan->d_adj_phi.as<double>(),
                                 plan->d_adj_grad.as<double>(), d_dq, d_dx, plan->stream));
    return FMM3DPP_OK;
}

int fmm3dpp_evaluate_adjoint(fmm3dpp_plan* plan, const double* q, const double* w, const double* grad, double* dq,
                             double* dx) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede evaluate_adjoint");
    if (!q || !w || !grad || !dq || !dx) FAIL(FMM3DPP_ERR_ARG, "null array");
    HIPCK(hipSetDevice(plan->device));
    const int64_t n = plan->gs.n;
    hipStream_t s = plan->stream;
    HIPCK(plan->d_qin.ensure(n * 8));
    HIPCK(plan->d_adj_w.ensure(n * 8));
    HIPCK(plan->d_adj_g.ensure(n * 24));
    HIPCK(plan->d_adj_dq.ensure(n * 8));
    HIPCK(plan->d_adj_dx.ensure(n * 24));
    HIPCK(hipMemcpyAsync(plan->d_qin.p, q, n * 8, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(plan->d_adj_w.p, w, n * 8, hipMemcpyHostToDevice, s));
    HIPCK(hipMemcpyAsync(plan->d_adj_g.p, grad, n * 24, hipMemcpyHostToDevice, s));
    int rc = fmm3dpp_evaluate_adjoint_device(plan, plan->d_qin.as<double>(), plan->d_adj_w.as<double>(),
                                             plan->d_adj_g.as<double>(), plan->d_adj_dq.as<double>(),
                                             plan->d_adj_dx.as<double>());
    if (rc) return rc;
    HIPCK(hipMemcpyAsync(dq, plan->d_adj_dq.p, n * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(dx, plan->d_adj_dx.p, n * 24, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    return FMM3DPP_OK;
}

int fmm3dpp_direct_device(void* hip_stream, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                          const double* d_q, int32_t ntar, const int32_t* d_targets, double cycle, int images,
                          double* d_phi, double* d_grad) {
    if (n < 1 || ntar < 1 || !d_x || !d_y || !d_z || !d_q || !d_targets || !d_phi || !d_grad) return FMM3DPP_ERR_ARG;
    if (images < 0 || images > 10 || !(cycle > 0)) return FMM3DPP_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    DBuf part;
    if (direct_device(n, d_x, d_y, d_z, d_q, ntar, d_targets, cycle, images, d_phi, d_grad, part, st) != hipSuccess)
        return FMM3DPP_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? FMM3DPP_OK : FMM3DPP_ERR_HIP;
}

int fmm3dpp_direct(int device, int64_t n, const double* x, const double* y, const double* z, const double* q,
                   int32_t ntar, const int32_t* targets, double cycle, int images, double* phi, double* grad) {
    if (n < 1 || ntar < 1 || !x || !y || !z || !q || !targets || !phi || !grad) return FMM3DPP_ERR_ARG;
    for (int32_t t = 0; t < ntar; t++)
        if (targets[t] < 0 || targets[t] >= n) return FMM3DPP_ERR_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return FMM3DPP_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return FMM3DPP_ERR_DEVICE;
    DBuf dx, dy, dz, dq, dt, dp, dg;
    const size_t nb = (size_t)n * 8;
    if (dx.ensure(nb) || dy.ensure(nb) || dz.ensure(nb) || dq.ensure(nb) || dt.ensure((size_t)ntar * 4) ||
        dp.ensure((size_t)ntar * 8) || dg.ensure((size_t)ntar * 24))
        return FMM3DPP_ERR_ALLOC;
    if (hipMemcpy(dx.p, x, nb, hipMemcpyHostToDevice) || hipMemcpy(dy.p, y, nb, hipMemcpyHostToDevice) ||
        hipMemcpy(dz.p, z, nb, hipMemcpyHostToDevice) || hipMemcpy(dq.p, q, nb, hipMemcpyHostToDevice) ||
        hipMemcpy(dt.p, targets, (size_t)ntar * 4, hipMemcpyHostToDevice))
        return FMM3DPP_ERR_HIP;
    int rc = fmm3dpp_direct_device(nullptr, n, dx.as<double>(), dy.as<double>(), dz.as<double>(), dq.as<double>(),
                                   ntar, dt.as<int32_t>(), cycle, images, dp.as<double>(), dg.as<double>());
    if (rc) return rc;
    if (hipMemcpy(phi, dp.p, (size_t)ntar * 8, hipMemcpyDeviceToHost) ||
        hipMemcpy(grad, dg.p, (size_t)ntar * 24, hipMemcpyDeviceToHost))
        return FMM3DPP_ERR_HIP;
    return FMM3DPP_OK;
}

int fmm3dpp_ewald_device(void* hip_stream, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                         const double* d_q, int32_t ntar, const int32_t* d_targets, double cycle, int ksize,
                         double alpha, double cutoff, double* d_phi, double* d_grad) {
    if (n < 1 || ntar < 1 || !d_x || !d_y || !d_z || !d_q || !d_targets || !d_phi || !d_grad) return FMM3DPP_ERR_ARG;
    if (ksize < 1 || ksize > 23 || !(cycle > 0) || !(alpha > 0) || !(cutoff > 0) || cutoff > cycle / 2)
        return FMM3DPP_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    EwaldScratch ws;
    if (ewald_device(n, d_x, d_y, d_z, d_q, ntar, d_targets, cycle, ksize, alpha, cutoff, d_phi, d_grad, ws, st) !=
        hipSuccess)
        return FMM3DPP_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? FMM3DPP_OK : FMM3DPP_ERR_HIP;
}

int fmm3dpp_ewald(int device, int64_t n, const double* x, const double* y, const double* z, const double* q,
                  int32_t ntar, const int32_t* targets, double cycle, int ksize, double alpha, double cutoff,
                  double* phi, double* grad) {
    if (n < 1 || ntar < 1 || !x || !y || !z || !q || !targets || !phi || !grad) return FMM3DPP_ERR_ARG;
    for (int32_t t = 0; t < ntar; t++)
        if (targets[t] < 0 || targets[t] >= n) return FMM3DPP_ERR_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return FMM3DPP_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return FMM3DPP_ERR_DEVICE;
    DBuf dx, dy, dz, dq, dt, dp, dg;
    const size_t nb = (size_t)n * 8;
    if (dx.ensure(nb) || dy.ensure(nb) || dz.ensure(nb) || dq.ensure(nb) || dt.ensure((size_t)ntar * 4) ||
        dp.ensure((size_t)ntar * 8) || dg.ensure((size_t)ntar * 24))
        return FMM3DPP_ERR_ALLOC;
    if (hipMemcpy(dx.p, x, nb, hipMemcpyHostToDevice) || hipMemcpy(dy.p, y, nb, hipMemcpyHostToDevice) ||
        hipMemcpy(dz.p, z, nb, hipMemcpyHostToDevice) || hipMemcpy(dq.p, q, nb, hipMemcpyHostToDevice) ||
        hipMemcpy(dt.p, targets, (size_t)ntar * 4, hipMemcpyHostToDevice))
        return FMM3DPP_ERR_HIP;
    int rc = fmm3dpp_ewald_device(nullptr, n, dx.as<double>(), dy.as<double>(), dz.as<double>(), dq.as<double>(), ntar,
                                  dt.as<int32_t>(), cycle, ksize, alpha, cutoff, dp.as<double>(), dg.as<double>());
    if (rc) return rc;
    if (hipMemcpy(phi, dp.p, (size_t)ntar * 8, hipMemcpyDeviceToHost) ||
        hipMemcpy(grad, dg.p, (size_t)ntar * 24, hipMemcpyDeviceToHost))
        return FMM3DPP_ERR_HIP;
    return FMM3DPP_OK;
}

}  // extern "C"
