// Float64 kernels shared by both precision paths of the periodic KIFMM on CDNA4 (gfx950): charge
// gather, the M2M child sum, the image-cell monopoles, the dipole moment and the final
// correction/reorder.  The operator applies are in kernels_apply.hip.
//
// Data layout in HBM:
//   bodies      SoA x, y, z, q, p, gx, gy, gz in tree order (leaf bodies contiguous)
//   q_equiv     [cell][S], p_check [cell][S]            (S = 6 (P-1)^2 + 2 surface points)
// Each kernel cites the reference loop it replaces (paths under 3dpp/).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"

namespace fmm3dpp {

namespace {

constexpr int kMoments = 5;   // sum q x, sum q y, sum q z, sum q |x|^2, sum q

__global__ void k_gather(const double* __restrict__ qin, const int32_t* __restrict__ perm, double* __restrict__ qt, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) qt[i] = qin[perm[i]];
}

__global__ __launch_bounds__(256) void k_sum_partials(int S, int ncol, int nz, const double* __restrict__ part,
                                                      const int32_t* __restrict__ out, double* __restrict__ Y) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)ncol * S) return;
    const int c = (int)(e / S), r = (int)(e - (int64_t)c * S);
    if (out[c] < 0) return;   // leaf column of an adaptive level: no M2M
    double a = 0;
    for (int z = 0; z < nz; z++) a += part[((int64_t)z * ncol + c) * S + r];
    Y[(int64_t)out[c] * S + r] = a;
}

// Infinite-lattice gauge (no counterpart in the reference): the root's equivalent density carries a
// monopole that differs from the bodies' total charge by ~1e-10 of the density's size (the rounding
// the pseudo-inverse leaves); each image level copies it 27 times and the 1x27 lattice M2L turns it
// into a potential offset that grows 9x per image level.  Each listed cell's density is shifted
// uniformly so that its sum, the cell's monopole, equals its charge: scale * qsum[block * qstride]
// (the total charge of the block's system times 27^m for the image cell c(-m)).  One wave per cell,
// cell = cells[block] (or first + block without a list); lane-strided partial sums, then an xor
// butterfly, after which every lane holds the same sum.
__global__ __launch_bounds__(64) void k_neutralize(double* __restrict__ qeq, int S, const int32_t* __restrict__ cells,
                                                   int first, const double* __restrict__ qsum, int qstride,
                                                   double scale) {
    const int c = cells ? cells[blockIdx.x] : first + (int)blockIdx.x;
    if (c < 0) return;   // uniform over the wave
    double* v = qeq + (int64_t)c * S;
    double a = 0;
    for (int j = threadIdx.x; j < S; j += 64) a += v[j];
    for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m);
    const double shift = (scale * qsum[(int64_t)blockIdx.x * qstride] - a) / S;
    for (int j = threadIdx.x; j < S; j += 64) v[j] += shift;
}

// Dipole moment sum_i q_i x_i (src/kernel.cpp:1587-1591), sum_i q_i |x_i|^2 and sum_i q_i as
// per-block partial sums [block][kMoments].
__global__ __launch_bounds__(256) void k_dipole_partial(BodiesDev b, int n, double* __restrict__ partial) {
    __shared__ double s[kMoments][256];
    const int tid = threadIdx.x;
    double a[kMoments] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        const double x = b.x[i], y = b.y[i], z = b.z[i], q = b.q[i];
        a[0] += x * q;
        a[1] += y * q;
        a[2] += z * q;
        a[3] += (x * x + y * y + z * z) * q;
        a[4] += q;
    }
    for (int k = 0; k < kMoments; k++) s[k][tid] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < kMoments; k++) s[k][tid] += s[k][tid + w];
        __syncthreads();
    }
    if (tid < kMoments) partial[kMoments * blockIdx.x + tid] = s[tid][0];
}


// Moments of each segment (a system; the whole body set for a plan) in a fixed order: wave k sums
// moment k of the segment's per-block partials, lane l the blocks l, l + 64, ..., then an xor
// butterfly over the lanes (one chain of nblocks dependent loads per moment took 31 us).
__global__ __launch_bounds__(64 * kMoments) void k_moments_seg(const double* __restrict__ partial, int nblocks,
                                                               double* __restrict__ moments) {
    const int seg = blockIdx.x, k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double a = 0;
    for (int j = lane; j < nblocks; j += 64) a += partial[((size_t)seg * nblocks + j) * kMoments + k];
    for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m);
    if (lane == 0) moments[seg * kMoments + k] = a;
}

// Dipole correction, 1/(4 pi) scaling (src/body.cpp:88-96) and the return to input order
// (src/body.cpp:177-190).  mode 1: the reference's correction (src/kernel.cpp:1580-1596);
// mode 2: the infinite-lattice gauge  phi -= (4 pi / 3V) M.x - (2 pi / 3V) sum_j q_j |x_j|^2.
// Both subtract (4 pi / 3V) M from the gradient.
__global__ __launch_bounds__(256) void k_finalize(int n, const int32_t* __restrict__ perm, BodiesDev b,
                                                  const double* __restrict__ p, const double* __restrict__ gx,
                                                  const double* __restrict__ gy, const double* __restrict__ gz,
                                                  NearDev near, const double* __restrict__ M, int mode,
                                                  double cycle, double out_scale, double* __restrict__ phi,
                                                  double* __restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pv = p[i], g0 = gx[i], g1 = gy[i], g2 = gz[i];
    if (near.p) {   // far + near, the sums the accumulating P2P forms
        pv += near.p[i];
        g0 += near.gx[i];
        g1 += near.gy[i];
        g2 += near.gz[i];
    }
    if (mode != 0) {
        const double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
        if (mode == 1) {
            const double dn = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
            pv -= coef * dn / n / b.q[i];
        } else {
            pv -= coef * (M[0] * b.x[i] + M[1] * b.y[i] + M[2] * b.z[i]) - 0.5 * coef * M[3];
        }
        g0 -= coef * M[0];
        g1 -= coef * M[1];
        g2 -= coef * M[2];
    }
    const int o = perm[i];
    phi[o] = pv * out_scale;
    grad[3 * (size_t)o] = g0 * out_scale;
    grad[3 * (size_t)o + 1] = g1 * out_scale;
    grad[3 * (size_t)o + 2] = g2 * out_scale;
}

// Backward of a loss L(phi) through phi = s A q (s = 1 or 1/(4 pi)), given w = dL/dphi, the
// forward gradient grad phi and the evaluation (psi, grad psi) of the same plan with charges w.
// The free-space and lattice kernels are symmetric, so A^T w = psi up to the gauge term:
//   dL/dq_k = psi_k + s c/2 (W |x_k|^2 - sum_i w_i |x_i|^2)
//   dL/dx_k = w_k grad phi_k + q_k (grad psi_k + s c W x_k)
// with W = sum w_i, c = 4 pi / 3V in the infinite-lattice gauge and c = 0 without a dipole term.
// Tree order in, input order out; Wm holds the moments of w (k_dipole_partial of the w charges).
__global__ __launch_bounds__(256) void k_adjoint_combine(int n, const int32_t* __restrict__ perm, BodiesDev b,
                                                         const double* __restrict__ Wm, double c,
                                                         double out_scale, const double* __restrict__ qin,
                                                         const double* __restrict__ win,
                                                         const double* __restrict__ grad_phi,
                                                         const double* __restrict__ psi,
                                                         const double* __restrict__ grad_psi, double* __restrict__ dq,
                                                         double* __restrict__ dx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    const double x = b.x[i], y = b.y[i], z = b.z[i];
    const double W = Wm[4], sc = out_scale * c;
    dq[o] = psi[o] + 0.5 * sc * (W * (x * x + y * y + z * z) - Wm[3]);
    const double wk = win[o], qk = qin[o];
    dx[3 * (size_t)o] = wk * grad_phi[3 * (size_t)o] + qk * (grad_psi[3 * (size_t)o] + sc * W * x);
    dx[3 * (size_t)o + 1] = wk * grad_phi[3 * (size_t)o + 1] + qk * (grad_psi[3 * (size_t)o + 1] + sc * W * y);
    dx[3 * (size_t)o + 2] = wk * grad_phi[3 * (size_t)o + 2] + qk * (grad_psi[3 * (size_t)o + 2] + sc * W * z);
}

// Segmented forms for batches of independent systems (one segment = one system's bodies in tree
// order): moments per segment, then the correction of k_finalize with that segment's moments.
__global__ __launch_bounds__(256) void k_dipole_seg(BodiesDev b, const int64_t* __restrict__ seg_off,
                                                    double* __restrict__ partial) {
    __shared__ double sm[kMoments][256];
    const int tid = threadIdx.x, seg = blockIdx.y;
    const int64_t i0 = seg_off[seg], i1 = seg_off[seg + 1];
    double a[kMoments] = {0, 0, 0, 0, 0};
    for (int64_t i = i0 + blockIdx.x * 256 + tid; i < i1; i += (int64_t)gridDim.x * 256) {
        const double x = b.x[i], y = b.y[i], z = b.z[i], q = b.q[i];
        a[0] += x * q;
        a[1] += y * q;
        a[2] += z * q;
        a[3] += (x * x + y * y + z * z) * q;
        a[4] += q;
    }
    for (int k = 0; k < kMoments; k++) sm[k][tid] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < kMoments; k++) sm[k][tid] += sm[k][tid + w];
        __syncthreads();
    }
    if (tid < kMoments) partial[((size_t)seg * gridDim.x + blockIdx.x) * kMoments + tid] = sm[tid][0];
}


__global__ __launch_bounds__(256) void k_finalize_seg(int64_t n, const int32_t* __restrict__ perm, BodiesDev b,
                                                      const double* __restrict__ p, const double* __restrict__ gx,
                                                      const double* __restrict__ gy, const double* __restrict__ gz,
                                                      const int64_t* __restrict__ seg_off, int nseg,
                                                      const double* __restrict__ moments, int mode, double cycle,
                                                      double out_scale, double* __restrict__ phi,
                                                      double* __restrict__ grad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nseg - 1;   // segment of body i: last seg with seg_off[seg] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg_off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const double* M = moments + lo * kMoments;
    double pv = p[i], g0 = gx[i], g1 = gy[i], g2 = gz[i];
    if (mode != 0) {
        const double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
        if (mode == 1) {
            const double dn = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
            pv -= coef * dn / (double)(seg_off[lo + 1] - seg_off[lo]) / b.q[i];
        } else {
            pv -= coef * (M[0] * b.x[i] + M[1] * b.y[i] + M[2] * b.z[i]) - 0.5 * coef * M[3];
        }
        g0 -= coef * M[0];
        g1 -= coef * M[1];
        g2 -= coef * M[2];
    }
    const int64_t o = perm[i];
    phi[o] = pv * out_scale;
    grad[3 * o] = g0 * out_scale;
    grad[3 * o + 1] = g1 * out_scale;
    grad[3 * o + 2] = g2 * out_scale;
}

__global__ __launch_bounds__(256) void k_fill_zero(double* __restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) a[i] = 0.0;
}

}  // namespace

hipError_t launch_gather(const double* qin, const int32_t* perm, double* qt, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!qin || !perm || !qt) return hipErrorInvalidValue;
    k_gather<<<(n + 255) / 256, 256, 0, st>>>(qin, perm, qt, n);
    return hipGetLastError();
}

hipError_t launch_sum_partials(int S, int ncol, int nz, const double* part, const int32_t* out, double* Y,
                               hipStream_t st) {
    if (ncol <= 0) return hipSuccess;
    if (!part || !out || !Y || S <= 0) return hipErrorInvalidValue;
    int64_t tot = (int64_t)ncol * S;
    k_sum_partials<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, ncol, nz, part, out, Y);
    return hipGetLastError();
}

hipError_t launch_neutralize(double* qeq, int S, const int32_t* cells, int first, int count, const double* qsum,
                             int qstride, double scale, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (!qeq || !qsum || S <= 0 || qstride < 0 || (!cells && first < 0)) return hipErrorInvalidValue;
    k_neutralize<<<count, 64, 0, st>>>(qeq, S, cells, first, qsum, qstride, scale);
    return hipGetLastError();
}

hipError_t launch_moments(const double* partial, int nblocks, double* moments, hipStream_t st) {
    if (!partial || !moments || nblocks <= 0) return hipErrorInvalidValue;
    k_moments_seg<<<1, 64 * kMoments, 0, st>>>(partial, nblocks, moments);
    return hipGetLastError();
}

hipError_t launch_dipole_partial(BodiesDev b, int n, double* partial, int nblocks, hipStream_t st) {
    if (!partial || !valid(b) || nblocks <= 0) return hipErrorInvalidValue;
    k_dipole_partial<<<nblocks, 256, 0, st>>>(b, n, partial);
    return hipGetLastError();
}

hipError_t launch_finalize(int n, const int32_t* perm, BodiesDev b, const double* p, const double* gx,
                           const double* gy, const double* gz, NearDev near, double* partial, int nblocks,
                           int mode, double cycle, double out_scale, double* phi, double* grad, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!perm || !p || !gx || !gy || !gz || !partial || !phi || !grad || !valid(b) || nblocks <= 0)
        return hipErrorInvalidValue;
    if (near.p && (!near.gx || !near.gy || !near.gz)) return hipErrorInvalidValue;
    double* M = partial + (size_t)nblocks * kMoments;   // row nblocks: the summed moments
    k_moments_seg<<<1, 64 * kMoments, 0, st>>>(partial, nblocks, M);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_finalize<<<(n + 255) / 256, 256, 0, st>>>(n, perm, b, p, gx, gy, gz, near, M, mode, cycle, out_scale, phi,
                                                grad);
    return hipGetLastError();
}

hipError_t launch_adjoint_combine(int n, const int32_t* perm, BodiesDev b, double* partial_w, int nblocks,
                                  double c, double out_scale, const double* qin, const double* win,
                                  const double* grad_phi, const double* psi, const double* grad_psi, double* dq,
                                  double* dx, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!perm || !partial_w || !qin || !win || !grad_phi || !psi || !grad_psi || !dq || !dx || !valid(b) ||
        nblocks <= 0)
        return hipErrorInvalidValue;
    double* Wm = partial_w + (size_t)nblocks * kMoments;   // row nblocks: the summed moments of w
    k_moments_seg<<<1, 64 * kMoments, 0, st>>>(partial_w, nblocks, Wm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_adjoint_combine<<<(n + 255) / 256, 256, 0, st>>>(n, perm, b, Wm, c, out_scale, qin, win, grad_phi, psi,
                                                        grad_psi, dq, dx);
    return hipGetLastError();
}

hipError_t launch_dipole_seg(BodiesDev b, const int64_t* seg_off, int nseg, int nblocks, double* partial,
                             double* moments, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    if (!seg_off || !partial || !moments || !valid(b) || nblocks <= 0) return hipErrorInvalidValue;
    k_dipole_seg<<<dim3(nblocks, nseg), 256, 0, st>>>(b, seg_off, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_moments_seg<<<nseg, 64 * kMoments, 0, st>>>(partial, nblocks, moments);
    return hipGetLastError();
}

hipError_t launch_finalize_seg(int64_t n, const int32_t* perm, BodiesDev b, const double* p, const double* gx,
                               const double* gy, const double* gz, const int64_t* seg_off, int nseg,
                               const double* moments, int mode, double cycle, double out_scale, double* phi,
                               double* grad, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!perm || !p || !gx || !gy || !gz || !seg_off || !moments || !phi || !grad || !valid(b) || nseg <= 0)
        return hipErrorInvalidValue;
    k_finalize_seg<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, perm, b, p, gx, gy, gz, seg_off, nseg, moments, mode,
                                                               cycle, out_scale, phi, grad);
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* a, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!a) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    k_fill_zero<<<(unsigned)blocks, 256, 0, st>>>(a, n);
    return hipGetLastError();
}

}  // namespace fmm3dpp
