// Device-side views and kernel launchers of the periodic FMM (float64 path).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

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

// Column-gathered batched GEMM, batch z = blockIdx.z:
//   Y_z[out_idx[z*idx_zstride + c]][0:S] (+)= scale_c * A_z * X[in_idx[z*idx_zstride + c]][0:S]
// with A_z = A + z*a_zstride (S x S row-major), X and Y_z = Y + z*y_zstride laid out [vector][S];
// in_idx < 0 reads zeros, out_idx < 0 skips the column; scale_c = 0.5^scale_depth[c] if given.
struct GemmCols {
    int S, ncol, nz, accumulate;
    const double* A;
    int64_t a_zstride;
    const double* X;
    const int32_t* in_idx;
    const int32_t* out_idx;
    int32_t idx_zstride;
    double* Y;
    int64_t y_zstride;
    const int32_t* scale_depth;
};

hipError_t launch_gather(const double* qin, const int32_t* perm, double* qt, int n, hipStream_t st);
hipError_t launch_p2m_check(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                            BodiesDev b, double* pc, hipStream_t st);
hipError_t launch_gemm_cols(const GemmCols& g, hipStream_t st);
// Y[out[c]][r] = sum_{z < nz} part[z][c][r] for out[c] >= 0 (fixed order: deterministic M2M child sum)
hipError_t launch_sum_partials(int S, int ncol, int nz, const double* part, const int32_t* out, double* Y,
                               hipStream_t st);
hipError_t launch_m2l_up(int P, int S, int nsrc, const int32_t* src_child, const int32_t* lattice,
                         const double* twiddle, const double* qeq, double* Qk, hipStream_t st);
hipError_t launch_hadamard8(int F, int ntar, int nsrc, const int32_t* slot_src_T, const double* gk8, const double* Qk,
                            double* Pk, hipStream_t st);
hipError_t launch_hadamard27(int F, int nimg, const int32_t* img_tar, const int32_t* img_src, int ntar,
                             int nsrc, const double* gk27_sum, const double* Qk, double* Pk, hipStream_t st);
hipError_t launch_m2l_down(int P, int S, int ntar, const int32_t* tar_child, const double* tar_scale,
                           const int32_t* lattice, const double* twiddle, const double* Pk, double* pck,
                           hipStream_t st);
hipError_t launch_l2p_eval(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                           BodiesDev b, const double* qdown, double* p, double* gx, double* gy, double* gz,
                           hipStream_t st);
// list_ent[e] = source cell << 5 | periodic offset index (offset = ((i%3)-1, (i/3%3)-1, (i/9)-1) * cycle)
hipError_t launch_p2p(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, const int64_t* list_off,
                      const int32_t* list_ent, double cycle, double* p, double* gx, double* gy, double* gz,
                      hipStream_t st);
hipError_t launch_dipole_partial(BodiesDev b, int n, double* partial /* 3 * nblocks */, int nblocks, hipStream_t st);
hipError_t launch_finalize(int n, const int32_t* perm, const double* q, const double* p, const double* gx,
                           const double* gy, const double* gz, const double* partial, int nblocks, int mode,
                           double cycle, double out_scale, double* phi, double* grad, hipStream_t st);

}  // namespace fmm3dpp
