// Per-stage entry points of the C-ABI: the operator API of the reference's LaplaceKernel
// (3dpp/include/kernel.h:39-115, 3dpp/src/kernel.cpp:258-311, 332-442, 1140-1511) on the plan's
// device state.  Each stage runs the kernels evaluate runs, over every cell it covers (cells = NULL)
// or over a list of cells; running the stages in evaluate's order with NULL lists gives evaluate's
// results bit for bit.  The per-cell vectors (q_equiv, p_check) and the per-body fields live in HBM;
// get/set entry points move them to and from the caller.  Every entry point returns after its work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "plan.h"

namespace {

int ready(fmm3dpp_plan* plan) {
    if (!plan) return FMM3DPP_ERR_ARG;
    if (plan->device < 0) FAIL(FMM3DPP_ERR_DEVICE, "plan was created without a device");
    if (!plan->particles) FAIL(FMM3DPP_ERR_STATE, "set_particles must precede the stages");
    HIPCK(hipSetDevice(plan->device));
    return ensure_ops(plan);
}

int finish(fmm3dpp_plan* plan) {
    HIPCK(hipStreamSynchronize(plan->stream));
    return FMM3DPP_OK;
}

// Host copies of the leaf list and the per-leaf P2P lists, once per set of particles.
int host_lists(fmm3dpp_plan* plan) {
    if (plan->host_lists_valid) return FMM3DPP_OK;
    const GpuStructure& g = plan->gs;
    plan->h_leaf_cell.resize(g.nleaf);
    plan->h_p2p_off.resize((size_t)g.nleaf + 1);
    plan->h_p2p_ent.resize((size_t)g.n_entries);
    HIPCK(hipStreamSynchronize(plan->stream));
    HIPCK(hipMemcpy(plan->h_leaf_cell.data(), g.leaf_cell.p, (size_t)g.nleaf * 4, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(plan->h_p2p_off.data(), g.p2p_off.p, ((size_t)g.nleaf + 1) * 8, hipMemcpyDeviceToHost));
    if (g.n_entries > 0)
        HIPCK(hipMemcpy(plan->h_p2p_ent.data(), g.p2p_ent.p, (size_t)g.n_entries * 4, hipMemcpyDeviceToHost));
    plan->host_lists_valid = true;
    return FMM3DPP_OK;
}

int check_cells(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    if (ncells < 0 || (ncells > 0 && !cells)) FAIL(FMM3DPP_ERR_ARG, "null or negative cell list");
    const int64_t nc = (int64_t)plan->st.cells.size();
    for (int64_t i = 0; i < ncells; i++)
        if (cells[i] < 0 || cells[i] >= nc) FAIL(FMM3DPP_ERR_ARG, "cell index out of range");
    return FMM3DPP_OK;
}

// The listed cells must be leaves; their cell indices and depths go to d_sel / d_sel_depth.
int select_leaves(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    int rc = ensure_host_structure(plan);
    if (rc || (rc = check_cells(plan, ncells, cells))) return rc;
    if (ncells > plan->gs.nleaf) FAIL(FMM3DPP_ERR_ARG, "more cells listed than the tree has leaves");
    std::vector<int32_t> dep((size_t)ncells);
    for (int64_t i = 0; i < ncells; i++) {
        const CellRec& c = plan->st.cells[cells[i]];
        if (c.cnum != 0 || c.depth < 0) FAIL(FMM3DPP_ERR_ARG, "listed cell is not a leaf");
        dep[i] = c.depth;
    }
    HIPCK(upload(plan->d_sel, std::vector<int32_t>(cells, cells + ncells), plan->stream));
    HIPCK(upload(plan->d_sel_depth, dep, plan->stream));
    return FMM3DPP_OK;
}

// Listed parents grouped by depth: tree parents (depth >= 0, with children) and image cells.
int parents_by_depth(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells, std::map<int, std::vector<int32_t>>& lv) {
    int rc = ensure_host_structure(plan);
    if (rc || (rc = check_cells(plan, ncells, cells))) return rc;
    for (int64_t i = 0; i < ncells; i++) {
        const CellRec& c = plan->st.cells[cells[i]];
        if (c.cnum <= 0) FAIL(FMM3DPP_ERR_ARG, "listed cell has no children");
        if (c.depth < 0 && c.cnum != 1) FAIL(FMM3DPP_ERR_ARG, "image cell without its single child");
        if (c.depth >= 0)   // tree children carry octants 0..7 (the octant-batched GEMM's batch index)
            for (int k = 0; k < c.cnum; k++) {
                const int o = plan->st.cells[c.coff + k].octant;
                if (o < 0 || o > 7) FAIL(FMM3DPP_ERR_ARG, "tree child with an octant outside 0..7");
            }
        lv[c.depth].push_back(cells[i]);
    }
    return FMM3DPP_OK;
}

// Column tables of one level of listed tree parents for the octant-batched GEMM: in[z][col] / out[z][col].
// to_child: in = parent, out = child of octant z (L2L); else in = child of octant z, out = z * ncol + col
// (M2M partial rows).  Absent children are -1.
void octant_tables(const Structure& st, const std::vector<int32_t>& par, size_t c0, size_t nc, bool to_child,
                   std::vector<int32_t>& in, std::vector<int32_t>& out) {
    in.assign(8 * nc, -1);
    out.assign(8 * nc, -1);
    for (size_t j = 0; j < nc; j++) {
        const CellRec& p = st.cells[par[c0 + j]];
        for (int k = 0; k < p.cnum; k++) {
            const int child = p.coff + k, z = st.cells[child].octant;
            if (to_child) {
                in[z * nc + j] = par[c0 + j];
                out[z * nc + j] = child;
            } else {
                in[z * nc + j] = child;
            }
        }
        if (!to_child)
            for (int z = 0; z < 8; z++) out[z * nc + j] = (int32_t)(z * nc + j);
    }
}

}  // namespace

extern "C" {

int fmm3dpp_stage_charges(fmm3dpp_plan* plan, const double* q) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!q) FAIL(FMM3DPP_ERR_ARG, "null charges");
    const size_t n = (size_t)plan->gs.n;
    HIPCK(hipMemcpyAsync(plan->d_qin.p, q, n * 8, hipMemcpyHostToDevice, plan->stream));
    HIPCK(launch_gather(plan->d_qin.as<double>(), plan->gs.perm.as<int32_t>(), plan->d_q.as<double>(), (int)n,
                        plan->stream));
    return finish(plan);
}

int fmm3dpp_stage_p2m(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    int rc = ready(plan);
    if (rc) return rc;
    const GpuStructure& g = plan->gs;
    if (!cells) {   // evaluate's order: p_check cleared, then every leaf
        HIPCK(hipMemsetAsync(plan->d_pck.p, 0, (size_t)g.ncells * plan->ops.S * 8, plan->stream));
        rc = stage_p2m_leaves(plan, g.leaf_cell.as<int32_t>(), g.leaf_depth.as<int32_t>(), g.nleaf);
    } else {
        if ((rc = select_leaves(plan, ncells, cells))) return rc;
        rc = stage_p2m_leaves(plan, plan->d_sel.as<int32_t>(), plan->d_sel_depth.as<int32_t>(), (int)ncells);
    }
    return rc ? rc : finish(plan);
}

int fmm3dpp_stage_m2m(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!cells) return (rc = stage_m2m_all(plan)) ? rc : finish(plan);
    std::map<int, std::vector<int32_t>> lv;
    if ((rc = parents_by_depth(plan, ncells, cells, lv))) return rc;
    const int Sn = plan->ops.S;
    const size_t chunk = (size_t)std::max(1, plan->gs.max_level_cols);
    const double* qtot = nullptr;   // the bodies' total charge (infinite-lattice gauge)
    if (plan->cfg.dipole_mode == FMM3DPP_DIPOLE_PHYSICAL && (rc = total_charge(plan, &qtot))) return rc;
    double* qeq = plan->d_qeq.as<double>();
    hipStream_t s = plan->stream;
    for (auto it = lv.rbegin(); it != lv.rend(); ++it) {   // deepest level first
        const std::vector<int32_t>& par = it->second;
        if (it->first >= 0) {
            for (size_t c0 = 0; c0 < par.size(); c0 += chunk) {
                const size_t nc = std::min(chunk, par.size() - c0);
                std::vector<int32_t> in, out;
                octant_tables(plan->st, par, c0, nc, false, in, out);
                HIPCK(upload(plan->d_sel_in, in, s));
                HIPCK(upload(plan->d_sel_out, out, s));
                HIPCK(upload(plan->d_sel, std::vector<int32_t>(par.begin() + c0, par.begin() + c0 + nc), s));
                const ApplyCols gm{Sn, (int)nc, 8, false, qeq, plan->d_sel_in.as<int32_t>(), plan->d_sel_out.as<int32_t>(),
                                   (int64_t)nc, plan->d_m2m_part.as<double>(), 0, nullptr};
                HIPCK(launch_apply(gm, plan->dop.lo_m2m, s));
                HIPCK(launch_sum_partials(Sn, (int)nc, 8, plan->d_m2m_part.as<double>(), plan->d_sel.as<int32_t>(), qeq,
                                          s));
                if (plan->f32_applies())
                    HIPCK(launch_monopole_parents(qeq, Sn, plan->d_sel.as<int32_t>(), (int)nc, plan->d_sel_in.as<int32_t>(),
                                                  plan->d_qcell.as<double>(), s));
                if (qtot && it->first == 0)
                    HIPCK(launch_neutralize(qeq, Sn, plan->d_sel.as<int32_t>(), 0, (int)nc, qtot, 0, 1.0, s));
                HIPCK(hipStreamSynchronize(s));   // the tables are reused by the next chunk
            }
        } else {
            for (int32_t p : par) {   // image cells: M2M_IMG (the 27 image operators summed) of the single child
                const int32_t child = plan->st.cells[p].coff;
                HIPCK(upload(plan->d_sel_in, std::vector<int32_t>{child}, s));
                HIPCK(upload(plan->d_sel_out, std::vector<int32_t>{p}, s));
                const ApplyCols gi{Sn, 1, 1, false, qeq, plan->d_sel_in.as<int32_t>(), plan->d_sel_out.as<int32_t>(), 0,
                                   qeq, 0, nullptr};
                HIPCK(launch_apply(gi, plan->dop.lo_m2m_img, s));
                if (qtot)
                    HIPCK(launch_neutralize(qeq, Sn, plan->d_sel_out.as<int32_t>(), 0, 1, qtot, 0,
                                            std::pow(27.0, -plan->st.cells[p].depth), s));
                HIPCK(hipStreamSynchronize(s));
            }
        }
    }
    return finish(plan);
}

int fmm3dpp_stage_m2l(fmm3dpp_plan* plan) {
    int rc = ready(plan);
    if (rc) return rc;
    return (rc = stage_m2l_all(plan)) ? rc : finish(plan);
}

int fmm3dpp_stage_l2l(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!cells) return (rc = stage_l2l_all(plan)) ? rc : finish(plan);
    std::map<int, std::vector<int32_t>> lv;
    if ((rc = parents_by_depth(plan, ncells, cells, lv))) return rc;
    const int Sn = plan->ops.S;
    double* pck = plan->d_pck.as<double>();
    hipStream_t s = plan->stream;
    for (auto& kv : lv) {   // outermost image cell first, then the tree levels from the root down
        const std::vector<int32_t>& par = kv.second;
        if (kv.first < 0) {
            for (int32_t p : par) {
                HIPCK(upload(plan->d_sel_in, std::vector<int32_t>{p}, s));
                HIPCK(upload(plan->d_sel_out, std::vector<int32_t>{plan->st.cells[p].coff}, s));
                const ApplyCols gi{Sn, 1, 1, true, pck, plan->d_sel_in.as<int32_t>(), plan->d_sel_out.as<int32_t>(), 0,
                                   pck, 0, nullptr};
                HIPCK(launch_apply(gi, plan->dop.lo_l2l_img, s));
                HIPCK(hipStreamSynchronize(s));
            }
        } else {
            std::vector<int32_t> in, out;
            octant_tables(plan->st, par, 0, par.size(), true, in, out);
            HIPCK(upload(plan->d_sel_in, in, s));
            HIPCK(upload(plan->d_sel_out, out, s));
            const ApplyCols gl{Sn, (int)par.size(), 8, true, pck, plan->d_sel_in.as<int32_t>(),
                               plan->d_sel_out.as<int32_t>(), (int64_t)par.size(), pck, 0, nullptr};
            HIPCK(launch_apply(gl, plan->dop.lo_l2l, s));
            HIPCK(hipStreamSynchronize(s));
        }
    }
    return finish(plan);
}

int fmm3dpp_stage_l2p(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells) {
    int rc = ready(plan);
    if (rc) return rc;
    const GpuStructure& g = plan->gs;
    if (!cells) {
        rc = stage_l2p_leaves(plan, g.leaf_cell.as<int32_t>(), g.leaf_depth.as<int32_t>(), g.nleaf);
    } else {
        if ((rc = select_leaves(plan, ncells, cells))) return rc;
        rc = stage_l2p_leaves(plan, plan->d_sel.as<int32_t>(), plan->d_sel_depth.as<int32_t>(), (int)ncells);
    }
    return rc ? rc : finish(plan);
}

int fmm3dpp_stage_p2p(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells, int accumulate) {
    int rc = ready(plan);
    if (rc) return rc;
    const GpuStructure& g = plan->gs;
    hipStream_t s = plan->stream;
    const int32_t* tl = g.leaf_cell.as<int32_t>();
    const int64_t* off = g.p2p_off.as<int64_t>();
    const int32_t* ent = g.p2p_ent.as<int32_t>();
    int nt = g.nleaf;
    if (cells) {   // the listed target leaves' rows of the P2P lists
        if ((rc = select_leaves(plan, ncells, cells)) || (rc = host_lists(plan))) return rc;
        std::vector<int32_t> leaf_of(plan->st.cells.size(), -1);
        for (int j = 0; j < g.nleaf; j++) leaf_of[plan->h_leaf_cell[j]] = j;
        std::vector<int64_t> so(1, 0);
        std::vector<int32_t> se;
        for (int64_t i = 0; i < ncells; i++) {
            const int j = leaf_of[cells[i]];
            if (j < 0) FAIL(FMM3DPP_ERR_ARG, "listed cell is not in the leaf list");
            se.insert(se.end(), plan->h_p2p_ent.begin() + plan->h_p2p_off[j], plan->h_p2p_ent.begin() + plan->h_p2p_off[j + 1]);
            so.push_back((int64_t)se.size());
        }
        if (se.empty()) se.push_back(0);   // an allocation for the pointer; no row reads it
        HIPCK(upload(plan->d_sel_off, so, s));
        HIPCK(upload(plan->d_sel_ent, se, s));
        tl = plan->d_sel.as<int32_t>();
        off = plan->d_sel_off.as<int64_t>();
        ent = plan->d_sel_ent.as<int32_t>();
        nt = (int)ncells;
    }
    const bool acc = accumulate != 0;
    if (!plan->mixed()) {
        HIPCK(launch_p2p(tl, nt, plan->cells_dev(), plan->bodies_dev(), off, ent, plan->cfg.cycle, plan->d_p.as<double>(),
                         plan->d_gx.as<double>(), plan->d_gy.as<double>(), plan->d_gz.as<double>(), acc, s));
    } else {   // float32 body records of every leaf (sources may be any leaf), then the listed targets
        HIPCK(launch_pack_rel(g.leaf_cell.as<int32_t>(), g.nleaf, plan->cells_dev(), plan->bodies_dev(),
                              plan->d_rel.as<float4>(), s));
        HIPCK(launch_p2p_f32(tl, nt, plan->cells_dev(), plan->d_rel.as<float4>(), off, ent, plan->cfg.cycle,
                             plan->d_p.as<double>(), plan->d_gx.as<double>(), plan->d_gy.as<double>(),
                             plan->d_gz.as<double>(), acc, s));
    }
    return finish(plan);
}

int fmm3dpp_stage_finalize(fmm3dpp_plan* plan, double* phi, double* grad) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!phi || !grad) FAIL(FMM3DPP_ERR_ARG, "null array");
    const GpuStructure& g = plan->gs;
    const int n = (int)g.n, nblk = 256;
    hipStream_t s = plan->stream;
    const BodiesDev bod = plan->bodies_dev();
    HIPCK(launch_dipole_partial(bod, n, plan->d_partial.as<double>(), nblk, s));
    HIPCK(launch_finalize(n, g.perm.as<int32_t>(), bod, plan->d_p.as<double>(), plan->d_gx.as<double>(),
                          plan->d_gy.as<double>(), plan->d_gz.as<double>(), NearDev{}, plan->d_partial.as<double>(),
                          nblk, plan->cfg.dipole_mode, plan->cfg.cycle,
                          plan->cfg.divide_4pi ? 1.0 / (4 * M_PI) : 1.0, plan->d_phi.as<double>(),
                          plan->d_grad.as<double>(), s));
    HIPCK(hipMemcpyAsync(phi, plan->d_phi.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipMemcpyAsync(grad, plan->d_grad.p, (size_t)n * 24, hipMemcpyDeviceToHost, s));
    return finish(plan);
}

int fmm3dpp_get_cell_vectors(fmm3dpp_plan* plan, int which, int64_t ncells, const int32_t* cells, double* out) {
    int rc = ready(plan);
    if (rc) return rc;
    if (which != 0 && which != 1) FAIL(FMM3DPP_ERR_ARG, "which must be 0 (q_equiv) or 1 (p_check)");
    if (!out) FAIL(FMM3DPP_ERR_ARG, "null array");
    const size_t S = (size_t)plan->ops.S;
    const double* src = (which == 0 ? plan->d_qeq : plan->d_pck).as<double>();
    HIPCK(hipStreamSynchronize(plan->stream));
    if (!cells) {
        HIPCK(hipMemcpy(out, src, (size_t)plan->gs.ncells * S * 8, hipMemcpyDeviceToHost));
        return FMM3DPP_OK;
    }
    if (ncells < 0) FAIL(FMM3DPP_ERR_ARG, "negative cell count");
    for (int64_t i = 0; i < ncells; i++) {
        if (cells[i] < 0 || cells[i] >= plan->gs.ncells) FAIL(FMM3DPP_ERR_ARG, "cell index out of range");
        HIPCK(hipMemcpy(out + i * S, src + (size_t)cells[i] * S, S * 8, hipMemcpyDeviceToHost));
    }
    return FMM3DPP_OK;
}

int fmm3dpp_set_cell_vectors(fmm3dpp_plan* plan, int which, int64_t ncells, const int32_t* cells, const double* in) {
    int rc = ready(plan);
    if (rc) return rc;
    if (which != 0 && which != 1) FAIL(FMM3DPP_ERR_ARG, "which must be 0 (q_equiv) or 1 (p_check)");
    if (!in) FAIL(FMM3DPP_ERR_ARG, "null array");
    const size_t S = (size_t)plan->ops.S;
    double* dst = (which == 0 ? plan->d_qeq : plan->d_pck).as<double>();
    HIPCK(hipStreamSynchronize(plan->stream));
    if (!cells) {
        HIPCK(hipMemcpy(dst, in, (size_t)plan->gs.ncells * S * 8, hipMemcpyHostToDevice));
        return FMM3DPP_OK;
    }
    if (ncells < 0) FAIL(FMM3DPP_ERR_ARG, "negative cell count");
    for (int64_t i = 0; i < ncells; i++) {
        if (cells[i] < 0 || cells[i] >= plan->gs.ncells) FAIL(FMM3DPP_ERR_ARG, "cell index out of range");
        HIPCK(hipMemcpy(dst + (size_t)cells[i] * S, in + i * S, S * 8, hipMemcpyHostToDevice));
    }
    return FMM3DPP_OK;
}

int fmm3dpp_get_body_fields(fmm3dpp_plan* plan, int64_t first, int64_t count, double* phi, double* grad) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!phi || !grad || first < 0 || count < 0 || first + count > plan->gs.n) FAIL(FMM3DPP_ERR_ARG, "bad body range");
    std::vector<double> g3(3 * (size_t)count);
    HIPCK(hipStreamSynchronize(plan->stream));
    HIPCK(hipMemcpy(phi, plan->d_p.as<double>() + first, count * 8, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(g3.data(), plan->d_gx.as<double>() + first, count * 8, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(g3.data() + count, plan->d_gy.as<double>() + first, count * 8, hipMemcpyDeviceToHost));
    HIPCK(hipMemcpy(g3.data() + 2 * count, plan->d_gz.as<double>() + first, count * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < count; i++)
        for (int d = 0; d < 3; d++) grad[3 * i + d] = g3[d * count + i];
    return FMM3DPP_OK;
}

int fmm3dpp_set_body_fields(fmm3dpp_plan* plan, int64_t first, int64_t count, const double* phi, const double* grad) {
    int rc = ready(plan);
    if (rc) return rc;
    if (!phi || !grad || first < 0 || count < 0 || first + count > plan->gs.n) FAIL(FMM3DPP_ERR_ARG, "bad body range");
    std::vector<double> g3(3 * (size_t)count);
    for (int64_t i = 0; i < count; i++)
        for (int d = 0; d < 3; d++) g3[d * count + i] = grad[3 * i + d];
    HIPCK(hipStreamSynchronize(plan->stream));
    HIPCK(hipMemcpy(plan->d_p.as<double>() + first, phi, count * 8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(plan->d_gx.as<double>() + first, g3.data(), count * 8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(plan->d_gy.as<double>() + first, g3.data() + count, count * 8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(plan->d_gz.as<double>() + first, g3.data() + 2 * count, count * 8, hipMemcpyHostToDevice));
    return FMM3DPP_OK;
}

}  // extern "C"
