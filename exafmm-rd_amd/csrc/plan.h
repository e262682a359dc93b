// The plan object behind the C-ABI (include/fmm3dpp.h) and the far-field stage functions its entry
// points share: evaluate runs them in sequence, the per-stage entry points (stages.hip) one at a time.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "dbuf.h"
#include "device.h"
#include "device_ops.h"
#include "fmm3dpp.h"
#include "gpu_structure.h"
#include "structure.h"

namespace fmm3dpp {
inline constexpr const char* kStageNames[] = {"P2M", "M2M", "M2L_up", "M2L_hadamard", "M2L_down", "L2L", "L2P", "P2P", "finalize"};
constexpr int kStages = 9;
// Events: 0..kStages are the stage boundaries on the plan's stream (7: end of L2P, 8: near field
// joined); kEvNear0 / kEvNear1 bracket the P2P on the side stream.  The P2P is forked at ev[5],
// the end of the M2L, so it overlaps L2L and L2P rather than the Hadamard or the down transform
// (overlapping the Hadamard slowed both by about the P2P's own time; after the M2L the step is
// 0.5% shorter than after the Hadamard).
constexpr int kEvNear0 = kStages + 1, kEvNear1 = kStages + 2, kEvents = kStages + 3, kEvForkNear = 5;
}  // namespace fmm3dpp

using namespace fmm3dpp;   // internal header: the library's sources share these names

struct fmm3dpp_plan {
    fmm3dpp_config cfg{};
    int device = FMM3DPP_DEVICE_NONE;
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool ops_ready = false, ops_uploaded = false, particles = false, host_struct_valid = false;
    Operators ops;
    Structure st;              // host tree/lists (structure-only plans; export cache for device plans)
    GpuStructure gs;           // device tree/lists (device plans)
    std::map<std::string, double> host_ms;
    hipEvent_t ev[kEvents] = {};
    // Near field (lists, P2P) beside the far field: created at the first evaluation that uses it, so
    // that a process whose plans keep the near field on their own streams holds one stream per plan
    // (HIP spreads streams over a few hardware queues; fewer streams, more plans on distinct queues).
    hipStream_t side_stream = nullptr;
    bool near_side = true;   // near field on side_stream (else on the plan's stream)
    hipError_t ensure_near_stream() {
        if (!near_side || side_stream) return hipSuccess;
        return hipStreamCreateWithFlags(&side_stream, hipStreamNonBlocking);
    }
    hipStream_t near_stream() const { return near_side ? side_stream : stream; }
    bool ev_recorded = false;
    int had_cap = kHadCap;   // distinct sources a Hadamard chunk may stage through LDS (fmm3dpp_set_hadamard_staging)

    DBuf d_inx, d_iny, d_inz, d_q, d_qin, d_p, d_gx, d_gy, d_gz, d_phi, d_grad;
    DBuf d_pn, d_gxn, d_gyn, d_gzn;   // near field, written by the P2P on the side stream
    DBuf d_pcleaf, d_pcleaf2, d_qeq, d_pck, d_Qk, d_Pk, d_partial, d_m2m_part, d_qcell;
    DBuf d_Qk_img, d_Pk_img, d_rel;   // mixed precision: float64 image-lattice spectra, float32 body records
    DBuf d_adj_phi, d_adj_grad, d_adj_w, d_adj_g, d_adj_dq, d_adj_dx;   // adjoint evaluation
    DeviceOperators dop;

    bool mixed() const { return cfg.precision == FMM3DPP_PREC_MIXED; }
    // Mixed path: tree operator stages (P2M check, P2M / L2P pseudo-inverses, M2M, L2L, L2P
    // evaluation) in float32, the applies on the float32 matrix cores, each cell's equivalent-density
    // monopole set to its charge after P2M and M2M (d_qcell holds the charges), for the orders
    // f32_apply_orders() admits (P = 6..8); off: float64 as in the float64 path (fmm3dpp_set_mixed_applies).
    bool apply_f32 = true;
    bool f32_applies() const { return mixed() && apply_f32 && f32_apply_orders(cfg.P); }

    CellsDev cells_dev() const {
        return {gs.cx.as<double>(), gs.cy.as<double>(), gs.cz.as<double>(), gs.cr.as<double>(), gs.depth.as<int32_t>(),
                gs.boff.as<int32_t>(), gs.bnum.as<int32_t>()};
    }
    BodiesDev bodies_dev() const { return {gs.x.as<double>(), gs.y.as<double>(), gs.z.as<double>(), d_q.as<double>()}; }

    // Per-stage entry points (stages.hip): cell selections and their tables in HBM, and host copies of
    // the leaf list and the P2P lists (read once per set of particles).
    DBuf d_sel, d_sel_depth, d_sel_in, d_sel_out, d_sel_off, d_sel_ent;
    bool host_lists_valid = false;
    std::vector<int32_t> h_leaf_cell, h_p2p_ent;
    std::vector<int64_t> h_p2p_off;
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

namespace fmm3dpp {

// Operators computed (or loaded) and uploaded; the host tree and lists of a device plan exported.
int ensure_ops(fmm3dpp_plan* plan);
int ensure_host_structure(fmm3dpp_plan* plan);

// Far-field stages of LaplaceFMM::solve (3dpp/src/fmm.cpp:50-74) on the plan's stream, over the
// plan's device tree.  p2m / l2p act on the leaves leaf_cell[0..nleaf) (depths leaf_depth) and write
// q_equiv / the per-body fields of those leaves; m2m, m2l and l2l cover every cell as evaluate does.
int stage_p2m_leaves(fmm3dpp_plan* plan, const int32_t* leaf_cell, const int32_t* leaf_depth, int nleaf);
// The bodies' total charge on the device (the moments of launch_dipole_partial, summed into d_partial's
// row 256); *qtot points at it.
int total_charge(fmm3dpp_plan* plan, const double** qtot);
int stage_m2m_all(fmm3dpp_plan* plan);
int stage_m2l_all(fmm3dpp_plan* plan);
int stage_l2l_all(fmm3dpp_plan* plan);
int stage_l2p_leaves(fmm3dpp_plan* plan, const int32_t* leaf_cell, const int32_t* leaf_depth, int nleaf);

}  // namespace fmm3dpp
