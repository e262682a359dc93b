// Tin-foil Ewald sum of the periodic system on the GPU, float64 (the reference's EwaldSolver,
// 3dpp/src/ewald.cpp:6-203), evaluated at a chosen set of target bodies.  It is the accuracy
// oracle of the FMM at sizes where the reference's CPU form is out of reach (N = 2^20 with
// |k_i| <= 23).
//
//   phi(x_t) = sum_j q_j erfc(alpha r)/r  (minimum image, 0 < r < cutoff)
//            + sum_{k != 0} (4 pi / L^3) exp(-|k|^2 / (4 alpha^2)) / |k|^2 Re(S(k) exp(i k.x_t))
//            - 2 alpha / sqrt(pi) q_t,                      S(k) = sum_j q_j exp(-i k.x_j)
// with k = 2 pi (a, b, c) / L, |a|, |b|, |c| <= K.  S(-k) = conj S(k), so S is formed on c >= 0
// only: the c > 0 half counts twice, the c = 0 plane (which holds both k and -k) once.
// Every sum is a fixed-order reduction of per-chunk partials: results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.h"

namespace fmm3dpp {

namespace {

constexpr int kSrcBatch = 32;   // sources whose phase tables sit in LDS at once

// S(a, b, c) partials: thread = (a, b) pair, blockIdx.y = source chunk.  Per source, the x and y
// phase tables exp(-i s a x) (a = -K..K) are built in LDS by powers of one sincos; the thread walks c
// = 0..K by repeated multiplication with exp(-i s z).
__global__ __launch_bounds__(256) void k_ewald_sk(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ z, const double* __restrict__ q, int K,
                                                  double s, int64_t chunk_len, double2* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int W = 2 * K + 1, nab = W * W;
    double2* tx = reinterpret_cast<double2*>(lds_raw);   // [kSrcBatch][W]
    double2* ty = tx + kSrcBatch * W;                     // [kSrcBatch][W]
    double2* tz = ty + kSrcBatch * W;                     // [kSrcBatch]  exp(-i s z)
    double* tq = reinterpret_cast<double*>(tz + kSrcBatch);
    const int tid = threadIdx.x;
    const int ab = blockIdx.x * 256 + tid;
    const int a = ab % W, b = ab / W;                     // table columns: a - K, b - K
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len, j1 = min(n, j0 + chunk_len);
    double acc_re[24], acc_im[24];                        // K <= 23
#pragma unroll
    for (int c = 0; c < 24; c++) acc_re[c] = acc_im[c] = 0;
    for (int64_t jb = j0; jb < j1; jb += kSrcBatch) {
        const int nb = (int)min((int64_t)kSrcBatch, j1 - jb);
        __syncthreads();
        if (tid < 3 * nb) {                               // one table row per thread: axis d of source r
            const int r = tid % nb, d = tid / nb;
            const int64_t j = jb + r;
            const double v = d == 0 ? x[j] : d == 1 ? y[j] : z[j];
            double sn, cs;
            sincos(s * v, &sn, &cs);
            const double2 e1 = make_double2(cs, -sn);     // exp(-i s v)
            if (d == 2) {
                tz[r] = e1;
                tq[r] = q[j];
            } else {
                double2* row = (d == 0 ? tx : ty) + r * W;
                double2 e = make_double2(1, 0);
                row[K] = e;
                for (int m = 1; m <= K; m++) {
                    e = make_double2(e.x * e1.x - e.y * e1.y, e.x * e1.y + e.y * e1.x);
                    row[K + m] = e;
                    row[K - m] = make_double2(e.x, -e.y);
                }
            }
        }
        __syncthreads();
        if (ab < nab) {
            for (int r = 0; r < nb; r++) {
                const double2 ex = tx[r * W + a], ey = ty[r * W + b], ez = tz[r];
                const double qq = tq[r];
                double wr = ex.x * ey.x - ex.y * ey.y, wi = ex.x * ey.y + ex.y * ey.x;
#pragma unroll
                for (int c = 0; c < 24; c++) {
                    if (c > K) break;
                    acc_re[c] += qq * wr;
                    acc_im[c] += qq * wi;
                    const double t = wr * ez.x - wi * ez.y;
                    wi = wr * ez.y + wi * ez.x;
                    wr = t;
                }
            }
        }
    }
    if (ab < nab) {
        double2* o = part + ((size_t)blockIdx.y * nab + ab) * (K + 1);
#pragma unroll
        for (int c = 0; c < 24; c++)
            if (c <= K) o[c] = make_double2(acc_re[c], acc_im[c]);
    }
}

// val(k) = weight * (4 pi / L^3) exp(-|k|^2 / (4 alpha^2)) / |k|^2 * sum over chunks of S partials
__global__ __launch_bounds__(256) void k_ewald_val(int K, int nchunk, double s, double alpha, double coef,
                                                   const double2* __restrict__ part, double2* __restrict__ val) {
    const int W = 2 * K + 1, nab = W * W, nk = nab * (K + 1);
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nk) return;
    const int c = e % (K + 1), ab = e / (K + 1), a = ab % W - K, b = ab / W - K;
    double re = 0, im = 0;
    for (int ch = 0; ch < nchunk; ch++) {
        const double2 v = part[(size_t)ch * nk + e];
        re += v.x;
        im += v.y;
    }
    const double k2 = (double)(a * a + b * b + c * c) * s * s;
    const double w = k2 > 0 ? (c > 0 ? 2.0 : 1.0) * coef * exp(-k2 / (4 * alpha * alpha)) / k2 : 0.0;
    val[e] = make_double2(re * w, im * w);
}

// Reciprocal part at the targets: thread = target, blockIdx.y = chunk of (a, b) pairs.
__global__ __launch_bounds__(256) void k_ewald_recip(int ntar, const int32_t* __restrict__ tar,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ z, int K, double s, int ab_len,
                                                     const double2* __restrict__ val, double4* __restrict__ part) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int W = 2 * K + 1, nab = W * W;
    const int ab0 = blockIdx.y * ab_len, ab1 = min(nab, ab0 + ab_len);
    if (t >= ntar) return;
    const int i = tar[t];
    const double xi = x[i], yi = y[i], zi = z[i];
    double sz, cz;
    sincos(s * zi, &sz, &cz);                              // exp(+i s z)
    double ap = 0, ax = 0, ay = 0, az = 0;
    for (int ab = ab0; ab < ab1; ab++) {
        const int a = ab % W - K, b = ab / W - K;
        double sn, cs;
        sincos(s * (a * xi + b * yi), &sn, &cs);
        double wr = cs, wi = sn;
        const double2* v = val + (size_t)ab * (K + 1);
        for (int c = 0; c <= K; c++) {
            const double2 vk = v[c];
            const double pr = vk.x * wr - vk.y * wi, pi = vk.x * wi + vk.y * wr;   // val(k) exp(i k.x)
            ap += pr;
            const double g = -pi * s;                                              // Re(i s k v) / k
            ax += g * a;
            ay += g * b;
            az += g * c;
            const double tt = wr * cz - wi * sz;
            wi = wr * sz + wi * cz;
            wr = tt;
        }
    }
    part[(size_t)blockIdx.y * ntar + t] = make_double4(ap, ax, ay, az);
}

// Real-space part: thread = target, blockIdx.y = source chunk staged through LDS.
__global__ __launch_bounds__(256) void k_ewald_real(int ntar, const int32_t* __restrict__ tar, int64_t n,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ q,
                                                    double L, double alpha, double cutoff, int64_t chunk_len,
                                                    double4* __restrict__ part) {
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const bool active = t < ntar;
    const int i = active ? tar[t] : 0;
    const double xi = x[i], yi = y[i], zi = z[i];
    const double c2 = cutoff * cutoff, a2 = alpha * alpha, ka = 2 * alpha / sqrt(M_PI);
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
        for (int r = 0; r < nb; r++) {
            double dx = xi - sx[r], dy = yi - sy[r], dz = zi - sz[r];
            dx -= L * rint(dx / L);
            dy -= L * rint(dy / L);
            dz -= L * rint(dz / L);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0 && r2 < c2) {
                const double rr = sqrt(r2), ir = 1.0 / rr;
                const double e = erfc(alpha * rr);
                ap += sq[r] * e * ir;
                const double fr = -sq[r] * ir * ir * ir * (e + ka * rr * exp(-a2 * r2));
                ax += fr * dx;
                ay += fr * dy;
                az += fr * dz;
            }
        }
    }
    if (active) part[(size_t)blockIdx.y * ntar + t] = make_double4(ap, ax, ay, az);
}

__global__ __launch_bounds__(256) void k_ewald_final(int ntar, const int32_t* __restrict__ tar,
                                                     const double* __restrict__ q, double alpha, int nreal,
                                                     const double4* __restrict__ preal, int nrec,
                                                     const double4* __restrict__ prec, double* __restrict__ phi,
                                                     double* __restrict__ grad) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ntar) return;
    double ap = 0, ax = 0, ay = 0, az = 0;
    for (int c = 0; c < nreal; c++) {
        const double4 v = preal[(size_t)c * ntar + t];
        ap += v.x; ax += v.y; ay += v.z; az += v.w;
    }
    for (int c = 0; c < nrec; c++) {
        const double4 v = prec[(size_t)c * ntar + t];
        ap += v.x; ax += v.y; ay += v.z; az += v.w;
    }
    ap -= 2 * alpha / sqrt(M_PI) * q[tar[t]];             // self term (ewald.cpp:112-118)
    phi[t] = ap;
    grad[3 * (size_t)t] = ax;
    grad[3 * (size_t)t + 1] = ay;
    grad[3 * (size_t)t + 2] = az;
}

}  // namespace

hipError_t ewald_device(int64_t n, const double* x, const double* y, const double* z, const double* q, int ntar,
                        const int32_t* tar, double cycle, int K, double alpha, double cutoff, double* phi,
                        double* grad, EwaldScratch& ws, hipStream_t st) {
    if (K < 1 || K > 23) return hipErrorInvalidValue;
    const int W = 2 * K + 1, nab = W * W, nk = nab * (K + 1);
    const double L = cycle, s = 2 * M_PI / L, coef = 4 * M_PI / (L * L * L);
    // structure factor: enough source chunks to fill the GPU with the (a, b) blocks
    const int ab_blocks = (nab + 255) / 256;
    int nchunk = (int)std::min<int64_t>(std::max<int64_t>(1, (2048 + ab_blocks - 1) / ab_blocks),
                                        std::max<int64_t>(1, n / 256));
    const int64_t chunk_len = (n + nchunk - 1) / nchunk;
    nchunk = (int)((n + chunk_len - 1) / chunk_len);
    hipError_t e;
    if ((e = ws.sk_part.ensure((size_t)nchunk * nk * 16)) != hipSuccess) return e;
    if ((e = ws.val.ensure((size_t)nk * 16)) != hipSuccess) return e;
    const size_t lds = (size_t)(2 * kSrcBatch * W + kSrcBatch) * 16 + kSrcBatch * 8;
    k_ewald_sk<<<dim3(ab_blocks, nchunk), 256, lds, st>>>(n, x, y, z, q, K, s, chunk_len,
                                                           ws.sk_part.as<double2>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_ewald_val<<<(nk + 255) / 256, 256, 0, st>>>(K, nchunk, s, alpha, coef, ws.sk_part.as<double2>(),
                                                   ws.val.as<double2>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // reciprocal and real parts at the targets, split so that the grid holds >= 1024 blocks
    const int tblocks = (ntar + 255) / 256;
    const int nrec = std::max(1, std::min(nab, 1024 / tblocks));
    const int ab_len = (nab + nrec - 1) / nrec;
    const int nrec_eff = (nab + ab_len - 1) / ab_len;
    const int nreal = (int)std::max<int64_t>(1, std::min<int64_t>(2048 / tblocks, (n + 255) / 256));
    const int64_t rlen = (n + nreal - 1) / nreal;
    const int nreal_eff = (int)((n + rlen - 1) / rlen);
    if ((e = ws.part.ensure((size_t)(nrec_eff + nreal_eff) * ntar * 32)) != hipSuccess) return e;
    double4* prec = ws.part.as<double4>();
    double4* preal = prec + (size_t)nrec_eff * ntar;
    k_ewald_recip<<<dim3(tblocks, nrec_eff), 256, 0, st>>>(ntar, tar, x, y, z, K, s, ab_len, ws.val.as<double2>(),
                                                           prec);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_ewald_real<<<dim3(tblocks, nreal_eff), 256, 0, st>>>(ntar, tar, n, x, y, z, q, L, alpha, cutoff, rlen, preal);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_ewald_final<<<tblocks, 256, 0, st>>>(ntar, tar, q, alpha, nreal_eff, preal, nrec_eff, prec, phi, grad);
    return hipGetLastError();
}

}  // namespace fmm3dpp
