// Direct lattice sum on the GPU, float64 (the reference's LaplaceKernel::direct,
// 3dpp/src/kernel.cpp:11-77): for each target, every source in the (3^images)^3 copies of the box,
// shifts ordered z, y, x outer, sources inner; pairs at r = 0 are skipped.  No dipole term; the
// caller applies it as the reference's test driver does.  Targets are split over source chunks
// whose partial sums are added in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "device.h"
#include "fmm3dpp.h"

namespace fmm3dpp {

namespace {

// Partial sums of fmm3dpp_direct_device_noalloc: one double4 per (source chunk, target).  The chunk
// count is chosen so that chunks x target blocks <= kDirectBlocks, so kDirectScratch entries hold any
// call with at most kDirectMaxTargets targets.  Each GPU holds its own copy for the life of the
// process; its address on the current device comes from hipGetSymbolAddress (the name itself is a
// host-side stand-in, not a device address).
constexpr int kDirectBlocks = 2048;
constexpr int64_t kDirectMaxTargets = (int64_t)kDirectBlocks * 256;   // 524,288 targets
constexpr int64_t kDirectScratch = kDirectMaxTargets;                  // 16 MiB of double4
static __device__ __align__(256) double4 scratch[kDirectScratch];

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

// Source chunks of a direct sum over n sources at ntar targets: enough to fill the GPU with
// chunks x target blocks <= kDirectBlocks (one chunk when the targets alone exceed that).
void direct_chunks(int64_t n, int ntar, int64_t& len, int& nch) {
    const int tblocks = (ntar + 255) / 256;
    const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(kDirectBlocks / tblocks, (n + 255) / 256));
    len = (n + nchunk - 1) / nchunk;
    nch = (int)((n + len - 1) / len);
}

hipError_t direct_run(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                      const int32_t* tar, double cycle, int images, double* phi, double* grad, double4* part,
                      int64_t len, int nch, hipStream_t st) {
    int dm = 0;
    for (int i = 0; i < images; i++) dm = 3 * dm + 1;   // (3^images - 1) / 2
    const int tblocks = (ntar + 255) / 256;
    k_direct<<<dim3(tblocks, nch), 256, 0, st>>>(ntar, tar, n, x, y, z, q, dm, cycle, len, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_direct_final<<<tblocks, 256, 0, st>>>(ntar, nch, part, phi, grad);
    return hipGetLastError();
}

// The address of `scratch` on each device, looked up once per device (the current one at the first
// call on it).
constexpr int kMaxDevices = 64;
double4* scratch_ptr[kMaxDevices] = {};
std::mutex scratch_mutex;

hipError_t scratch_on_current_device(double4** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        double4* p = nullptr;
        e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

hipError_t direct_device(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                         const int32_t* tar, double cycle, int images, double* phi, double* grad, DBuf& part,
                         hipStream_t st) {
    if (images < 0 || images > 10) return hipErrorInvalidValue;
    int64_t len;
    int nch;
    direct_chunks(n, ntar, len, nch);
    hipError_t e = part.ensure((size_t)nch * ntar * 32);
    if (e != hipSuccess) return e;
    return direct_run(n, x, y, z, q, ntar, tar, cycle, images, phi, grad, part.as<double4>(), len, nch, st);
}

}  // namespace fmm3dpp

extern "C" int fmm3dpp_direct_device_noalloc(void* hip_stream, int64_t n, const double* d_x, const double* d_y,
                                             const double* d_z, const double* d_q, int32_t ntar,
                                             const int32_t* d_targets, double cycle, int images, double* d_phi,
                                             double* d_grad) {
    using namespace fmm3dpp;
    if (n < 1 || ntar < 1 || !d_x || !d_y || !d_z || !d_q || !d_targets || !d_phi || !d_grad) return FMM3DPP_ERR_ARG;
    if (images < 0 || images > 10 || !(cycle > 0)) return FMM3DPP_ERR_ARG;
    if (ntar > kDirectMaxTargets) return FMM3DPP_ERR_ARG;   // beyond the static scratch
    int64_t len;
    int nch;
    direct_chunks(n, ntar, len, nch);
    if ((int64_t)nch * ntar > kDirectScratch) return FMM3DPP_ERR_ARG;
    double4* part = nullptr;
    if (scratch_on_current_device(&part) != hipSuccess) return FMM3DPP_ERR_HIP;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    if (direct_run(n, d_x, d_y, d_z, d_q, ntar, d_targets, cycle, images, d_phi, d_grad, part, len, nch, st) !=
        hipSuccess)
        return FMM3DPP_ERR_HIP;
    return hipStreamSynchronize(st) == hipSuccess ? FMM3DPP_OK : FMM3DPP_ERR_HIP;
}
