// Device-side views and kernel launchers of the periodic FMM (float64 and mixed-precision paths).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "dbuf.h"

namespace fmm3dpp {

// Structure-of-arrays view of the cells in HBM (BFS order + image cells).
struct CellsDev {
    const double* x;
    const double* y;
    const double* z;
    const double* r;
    const int32_t* depth;
    const int32_t* boff;
    const int32_t* bnum;
};

struct BodiesDev {
    const double* x;
    const double* y;
    const double* z;
    const double* q;
};

// Launchers refuse views with an unallocated array (a null device pointer would fault the kernel).
inline bool valid(const CellsDev& c) { return c.x && c.y && c.z && c.r && c.depth && c.boff && c.bnum; }
inline bool valid(const BodiesDev& b) { return b.x && b.y && b.z && b.q; }

// Columns of an operator apply, batch z = blockIdx.z:
//   Y_z[out_idx[z*idx_zstride + c]][0:M] (+)= scale_c * A_z * X[in_idx[z*idx_zstride + c]][0:K]
// with X and Y_z = Y + z*y_zstride laid out [vector][S]; in_idx < 0 reads zeros, out_idx < 0 skips
// the column; scale_c = 0.5^scale_depth[c] if given.  The operator A_z (M x K) is a LaneOp.
struct ApplyCols {
    int S, ncol, nz;
    bool accumulate;
    const double* X;
    const int32_t* in_idx;
    const int32_t* out_idx;
    int64_t idx_zstride;
    double* Y;
    int64_t y_zstride;
    const int32_t* scale_depth;
};

// An operator, or a batch of nz operators, in the lane layout of the apply kernels: element
// [kc][b][u][lane] = A[16 b + lane % 16][16 kc + 4 u + lane / 16] (zero outside M x K), operator z at
// p + z * zstride elements; float64, or float32 for the float32 matrix cores.
struct LaneOp {
    const void* p = nullptr;
    bool f32 = false;
    int M = 0, K = 0;       // rows and inner dimension
    int nmb = 0, nkc = 0;   // 16-row blocks and 16-k chunks: (M + 15) / 16, (K + 15) / 16
    int64_t zstride = 0;    // nmb * nkc * 256
    int nops = 0;           // operators held (a launch's batch may use up to this many)
};

// Y (+)= A X over the columns (kernels_apply.hip).  mbw > 0 forces the tiled kernel with that many
// 16-row blocks per workgroup (20, 12, 8 or 4); the result is the same bits for every choice.
hipError_t launch_apply(const ApplyCols& g, const LaneOp& op, hipStream_t st, int mbw = 0);
// Once per process before the first apply (and before any graph capture): LDS limits of the kernels.
hipError_t apply_init();
// out[at ...] = the lane layout of the M x K leading block of A (row stride lda).
template <class T>
void lane_layout(const double* A, int lda, int M, int K, std::vector<T>& out, size_t at);
// Per-cell monopole constraint of the mixed path: each listed cell's equivalent density is shifted
// uniformly so that its sum equals the cell's charge, recorded in qcell.  Leaves: the charge of the
// leaf's bodies (tree order).  Parents of a level: the sum of the children's recorded charges, in
// octant order over the level's child table children[o * ncol + col] (-1: no child).
hipError_t launch_monopole_leaves(double* qeq, int S, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                                  const double* q, double* qcell, hipStream_t st);
hipError_t launch_monopole_parents(double* qeq, int S, const int32_t* parents, int ncol, const int32_t* children,
                                   double* qcell, hipStream_t st);

hipError_t launch_gather(const double* qin, const int32_t* perm, double* qt, int n, hipStream_t st);
// f32: float32 form (coordinates relative to the leaf centre, float32 sums) of the mixed path's
// float32 operator applies.
hipError_t launch_p2m_check(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                            BodiesDev b, double* pc, hipStream_t st, bool f32 = false);
// Y[out[c]][r] = sum_{z < nz} part[z][c][r] for out[c] >= 0 (fixed order: deterministic M2M child sum)
hipError_t launch_sum_partials(int S, int ncol, int nz, const double* part, const int32_t* out, double* Y,
                               hipStream_t st);
// M2L (f32 selects the float32 complex layout of Qk/Pk and gk8; the transforms read q_equiv and
// accumulate p_check in float64 either way).  Qk/Pk are [parent][f][8 children] complex;
// inv_lat maps the P^3 corner of the convolution grid ((k * P + j) * P + i) to a surface point or -1.
// Tiles: SG child slots per workgroup (1, 2, 4 or 8) and FC half-spectrum columns per LDS chunk
// (1..P); 0, 0 chooses them from the measured defaults.  Results do not depend on the tiles.
struct M2LTiles { int sg = 0, fc = 0; };
hipError_t launch_m2l_up(int P, bool f32, int S, int nsrc, const int32_t* src_child, const int32_t* inv_lat,
                         const double* twiddle, const double* qeq, void* Qk, hipStream_t st, M2LTiles tiles = {});
// Hadamard staging tables per chunk c of kHadChunk consecutive targets: the chunk's distinct sources
// uniq[c * kHadUniqStride + r] (r < cnt[c], increasing) and each (pattern p, target c * kHadChunk +
// tt)'s row among them, loc[(c * 26 + p) * kHadChunk + tt] (cnt[c] when it has no source).  A chunk
// with more than kHadCap distinct sources is read from the spectra directly.
constexpr int kHadChunk = 256, kHadCap = 767, kHadUniqStride = 26 * kHadChunk;
inline int had_chunks(int ntar) { return (ntar + kHadChunk - 1) / kHadChunk; }
struct HadTables {
    const int32_t *uniq = nullptr, *cnt = nullptr, *loc = nullptr;
};
hipError_t launch_hadamard_tables(int ntar, int ntab, const int32_t* slot_src_T, int32_t* uniq, int32_t* cnt,
                                  int32_t* loc, hipStream_t st);
// the 8x8 kernel spectra gk8 ([26][F][64] complex) as matrix-core A operands in lane order, F x 26 x 64
// vectors of 4 (float32 when f32, else float64)
hipError_t launch_hadamard_lanes(bool f32, int F, const void* gk8, void* gkA, hipStream_t st);
// targets t < ntar, slot_src_T rows of stride ntab (-1: no source), tables of launch_hadamard_tables
// for the same (ntar, ntab, slot_src_T), gkA of launch_hadamard_lanes.  Chunks of at most cap
// (<= kHadCap) distinct sources are staged through LDS, the others read the spectra directly; cap
// -1 stages none.  The results do not depend on cap.
hipError_t launch_hadamard8(bool f32, int F, int ntar, int ntab, const int32_t* slot_src_T, HadTables tab,
                            const void* gkA, const void* Qk, void* Pk, hipStream_t st, int cap = kHadCap);
// float64 1x27 image lattice; img_src_T rows have stride ntab, Qk/Pk hold the parents from index base on
hipError_t launch_hadamard27(int F, int nimg, const int32_t* img_tar, const int32_t* img_src_T, int ntab, int base,
                             const double* gk27_sum, const double* Qk, double* Pk, hipStream_t st);
hipError_t launch_m2l_down(int P, bool f32, int S, int ntar, const int32_t* tar_child, const double* tar_scale,
                           const int32_t* inv_lat, const double* twiddle, const void* Pk, double* pck,
                           hipStream_t st, M2LTiles tiles = {});
// Whether explicit tiles fit (slot group divides 8, chunk within the half spectrum, LDS <= 160 KiB).
bool m2l_tiles_valid(int P, bool f32, bool down, M2LTiles tiles);
hipError_t launch_l2p_eval(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                           BodiesDev b, const double* qdown, double* p, double* gx, double* gy, double* gz,
                           hipStream_t st, bool f32 = false);
// list_ent[e] = source leaf cell << 5 | periodic offset index (offset = ((i%3)-1, (i/3%3)-1, (i/9)-1) * cycle).
// accumulate: p += phi_near, g -= sum q d / r^3 (onto the far field); otherwise the near field alone
// is stored (p = phi_near, g = -sum q d / r^3), for a finalize that adds the two.
hipError_t launch_p2p(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, const int64_t* list_off,
                      const int32_t* list_ent, double cycle, double* p, double* gx, double* gy, double* gz,
                      bool accumulate, hipStream_t st);
// float32 body records (x, y, z relative to the leaf centre, q) and the float32 P2P over them
hipError_t launch_pack_rel(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, float4* rel,
                           hipStream_t st);
hipError_t launch_p2p_f32(const int32_t* leaf_cell, int nleaf, CellsDev cells, const float4* rel,
                          const int64_t* list_off, const int32_t* list_ent, double cycle, double* p, double* gx,
                          double* gy, double* gz, bool accumulate, hipStream_t st);
// Infinite-lattice gauge: shift the equivalent density of cells[0..count) (first..first+count without a
// list) uniformly so that its sum (monopole) is scale * qsum[i * qstride] for the i-th listed cell.
hipError_t launch_neutralize(double* qeq, int S, const int32_t* cells, int first, int count, const double* qsum,
                             int qstride, double scale, hipStream_t st);
// moments[0..5) = the sum of partial's nblocks rows of 5 (launch_dipole_partial's), in row order.
hipError_t launch_moments(const double* partial, int nblocks, double* moments, hipStream_t st);
// partial[5 * block + (0..4)] = sum q x, sum q y, sum q z, sum q |x|^2, sum q
hipError_t launch_dipole_partial(BodiesDev b, int n, double* partial /* 5 * nblocks */, int nblocks, hipStream_t st);
// dL/dq and dL/dx from w = dL/dphi, grad phi, and (psi, grad psi) = the evaluation with charges w
// partial_w holds nblocks + 1 rows of kMoments; row nblocks receives the summed moments.
hipError_t launch_adjoint_combine(int n, const int32_t* perm, BodiesDev b, double* partial_w, int nblocks,
                                  double c, double out_scale, const double* qin, const double* win,
                                  const double* grad_phi, const double* psi, const double* grad_psi, double* dq,
                                  double* dx, hipStream_t st);
// Near-field results stored apart from the far field (P2P with accumulate = false); all null: none.
struct NearDev {
    const double *p = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr;
};
// mode: FMM3DPP_DIPOLE_* (0 off, 1 reference, 2 infinite-lattice gauge); partial holds nblocks + 1
// rows of kMoments (launch_dipole_partial's and one for their sum).  With `near` set, phi =
// p + near.p and grad = g + near.g, the same roundings as the P2P accumulating onto p and g.
hipError_t launch_finalize(int n, const int32_t* perm, BodiesDev b, const double* p, const double* gx,
                           const double* gy, const double* gz, NearDev near, double* partial, int nblocks,
                           int mode, double cycle, double out_scale, double* phi, double* grad, hipStream_t st);

// Batches: moments and the final correction per segment (system) of the tree-ordered bodies;
// partial holds nseg x nblocks x 5 doubles, moments nseg x 5.  A fill kernel instead of
// hipMemsetAsync, so that the stage sequence can be captured in a graph.
hipError_t launch_dipole_seg(BodiesDev b, const int64_t* seg_off, int nseg, int nblocks, double* partial,
                             double* moments, hipStream_t st);
hipError_t launch_finalize_seg(int64_t n, const int32_t* perm, BodiesDev b, const double* p, const double* gx,
                               const double* gy, const double* gz, const int64_t* seg_off, int nseg,
                               const double* moments, int mode, double cycle, double out_scale, double* phi,
                               double* grad, hipStream_t st);
hipError_t launch_fill_zero(double* a, int64_t n, hipStream_t st);

// Tin-foil Ewald sum at ntar target bodies (device arrays), float64; K = max |k_i| <= 23.
struct EwaldScratch {
    DBuf sk_part, val, part;
};
hipError_t ewald_device(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                        const int32_t* tar, double cycle, int K, double alpha, double cutoff, double* phi,
                        double* grad, EwaldScratch& ws, hipStream_t st);

// Direct lattice sum over the (3^images)^3 copies at ntar targets (device arrays), float64; part is
// scratch for the per-chunk partial sums.
hipError_t direct_device(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                         const int32_t* tar, double cycle, int images, double* phi, double* grad, DBuf& part,
                         hipStream_t st);

}  // namespace fmm3dpp
