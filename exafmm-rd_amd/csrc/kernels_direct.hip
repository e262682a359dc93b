// Direct lattice sum on the GPU, float64 (the reference's LaplaceKernel::direct,
// 3dpp/src/kernel.cpp:11-77): for each target, every source in the (3^images)^3 copies of the box,
// shifts ordered z, y, x outer, sources inner; pairs at r = 0 are skipped.  No dipole term; the
// caller applies it as the reference's test driver does.  Targets are split over source chunks
// whose partial sums are added in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.h"

namespace fmm3dpp {

namespace {

__global__ __launch_bounds__(256) void k_direct(int ntar, const int32_t* __restrict__ tar, int64_t n,
                                                const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, const double* __restrict__ q, int dm,
                                                double cycle, int64_t chunk_len, double4* __restrict__ part) {
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const bool active = t < ntar;
    const int i = active ? tar[t] : 0;
    const double xi = x[i], yi = y[i], zi = z[i];
    double ap = 0, ax = 0, ay = 0, az = 0;
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len, j1 = min(n, j0 + chunk_len);
    for (int64_t jb = j0; jb < j1; jb += 256) {
        const int nb = (int)min((int64_t)256, j1 - jb);
        __syncthreads();
        if ((int)threadIdx.x < nb) {
            const int64_t j = jb + threadIdx.x;
            sx[threadIdx.x] = x[j];
            sy[threadIdx.x] = y[j];
            sz[threadIdx.x] = z[j];
            sq[threadIdx.x] = q[j];
        }
        __syncthreads();
        for (int pz = -dm; pz <= dm; pz++)
            for (int py = -dm; py <= dm; py++)
                for (int px = -dm; px <= dm; px++) {
                    const double ox = px * cycle, oy = py * cycle, oz = pz * cycle;
                    for (int r = 0; r < nb; r++) {
                        const double dx = xi - (sx[r] + ox), dy = yi - (sy[r] + oy), dz = zi - (sz[r] + oz);
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        const double ir = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
                        double qi = sq[r] * ir;
                        ap += qi;
                        qi *= ir * ir;
                        ax += qi * dx;
                        ay += qi * dy;
                        az += qi * dz;
                    }
                }
    }
    if (active) part[(size_t)blockIdx.y * ntar + t] = make_double4(ap, ax, ay, az);
}

__global__ __launch_bounds__(256) void k_direct_final(int ntar, int nchunk, const double4* __restrict__ part,
                                                      double* __restrict__ phi, double* __restrict__ grad) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntar) return;
    double ap = 0, ax = 0, ay = 0, az = 0;
    for (int c = 0; c < nchunk; c++) {
        const double4 v = part[(size_t)c * ntar + t];
        ap += v.x; ax += v.y; ay += v.z; az += v.w;
    }
    phi[t] = ap;
    grad[3 * (size_t)t] = -ax;
    grad[3 * (size_t)t + 1] = -ay;
    grad[3 * (size_t)t + 2] = -az;
}

}  // namespace

hipError_t direct_device(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                         const int32_t* tar, double cycle, int images, double* phi, double* grad, DBuf& part,
                         hipStream_t st) {
    if (images < 0 || images > 10) return hipErrorInvalidValue;
    int dm = 0;
    for (int i = 0; i < images; i++) dm = 3 * dm + 1;   // (3^images - 1) / 2
    const int tblocks = (ntar + 255) / 256;
    const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(2048 / tblocks, (n + 255) / 256));
    const int64_t len = (n + nchunk - 1) / nchunk;
    const int nch = (int)((n + len - 1) / len);
    hipError_t e = part.ensure((size_t)nch * ntar * 32);
    if (e != hipSuccess) return e;
    k_direct<<<dim3(tblocks, nch), 256, 0, st>>>(ntar, tar, n, x, y, z, q, dm, cycle, len, part.as<double4>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_direct_final<<<tblocks, 256, 0, st>>>(ntar, nch, part.as<double4>(), phi, grad);
    return hipGetLastError();
}

}  // namespace fmm3dpp
