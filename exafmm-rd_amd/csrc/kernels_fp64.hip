// Float64 HIP kernels of the periodic KIFMM evaluation path for CDNA4 (gfx950).
//
// Data layout in HBM:
//   bodies      SoA x, y, z, q, p, gx, gy, gz in tree order (leaf bodies contiguous)
//   q_equiv     [cell][S], p_check [cell][S]            (S = 6 (P-1)^2 + 2 surface points)
//   Qk, Pk      [f][parent][8 children] complex double  (frequency-major so one frequency slice of
//               every parent is contiguous and stays in an XCD's L2 while the Hadamard streams it)
// Each kernel cites the reference loop it replaces (paths under 3dpp/).
#include <hip/hip_runtime.h>

#include "device.h"

namespace fmm3dpp {

namespace {

constexpr int kMaxPts = 6;  // check/equivalent points per thread in the 256-thread surface kernels (P <= 16)

__global__ void k_gather(const double* __restrict__ qin, const int32_t* __restrict__ perm, double* __restrict__ qt, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) qt[i] = qin[perm[i]];
}

// P2M source -> check-surface potentials (src/kernel.cpp:332-368): one workgroup per leaf,
// check points on threads, leaf bodies staged through LDS.
__global__ __launch_bounds__(256) void k_p2m_check(int S, const double* __restrict__ usurf,
                                                   const int32_t* __restrict__ leaf_cell, CellsDev cells,
                                                   BodiesDev b, double* __restrict__ pc) {
    __shared__ double sx[256], sy[256], sz[256], sq[256];
    const int li = blockIdx.x, tid = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    double px[kMaxPts], py[kMaxPts], pz[kMaxPts], acc[kMaxPts];
#pragma unroll
    for (int m = 0; m < kMaxPts; m++) {
        int s = tid + 256 * m;
        acc[m] = 0;
        if (s < S) {
            px[m] = usurf[3 * s] * rr + cx;
            py[m] = usurf[3 * s + 1] * rr + cy;
            pz[m] = usurf[3 * s + 2] * rr + cz;
        } else {
            px[m] = py[m] = pz[m] = 0;
        }
    }
    for (int b0 = 0; b0 < bnum; b0 += 256) {
        const int nb = min(256, bnum - b0);
        __syncthreads();
        if (tid < nb) {
            sx[tid] = b.x[boff + b0 + tid];
            sy[tid] = b.y[boff + b0 + tid];
            sz[tid] = b.z[boff + b0 + tid];
            sq[tid] = b.q[boff + b0 + tid];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kMaxPts; m++) {
            if (tid + 256 * m >= S) continue;
            double a = acc[m];
            for (int j = 0; j < nb; j++) {
                double dx = px[m] - sx[j], dy = py[m] - sy[j], dz = pz[m] - sz[j];
                double r2 = dx * dx + dy * dy + dz * dz;
                a += r2 > 0 ? sq[j] * (1.0 / sqrt(r2)) : 0.0;
            }
            acc[m] = a;
        }
    }
#pragma unroll
    for (int m = 0; m < kMaxPts; m++) {
        int s = tid + 256 * m;
        if (s < S) pc[(size_t)li * S + s] = acc[m];
    }
}

// Column-gathered batched GEMM for the operator applies: P2M/L2P pseudo-inverse factors
// (src/kernel.cpp:369-370, 1465-1467), M2M per child octant (395-404), L2L per child octant
// (1419-1428) and the image-level M2M/L2L (434-442, 1430-1435).  64x64 output tile per workgroup,
// 4x4 per thread, K in steps of 16 with the next tile prefetched into registers.
constexpr int TM = 64, TN = 64, TK = 16;
__global__ __launch_bounds__(256) void k_gemm_cols(GemmCols g) {
    __shared__ double As[TK][TM + 2];
    __shared__ double Bs[TK][TN + 2];
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int row0 = blockIdx.x * TM, col0 = blockIdx.y * TN, z = blockIdx.z;
    const int S = g.S, K = g.S;
    const double* A = g.A + (int64_t)z * g.a_zstride;
    const int32_t* in_idx = g.in_idx + (int64_t)z * g.idx_zstride;
    const int32_t* out_idx = g.out_idx + (int64_t)z * g.idx_zstride;
    double* Y = g.Y + (int64_t)z * g.y_zstride;
    const int lr = tid >> 2, lk = (tid & 3) * 4;
    const int gr = row0 + lr, gc = col0 + lr;
    const int src = gc < g.ncol ? in_idx[gc] : -1;
    const double* arow = gr < S ? A + (int64_t)gr * K : nullptr;
    const double* bcol = src >= 0 ? g.X + (int64_t)src * S : nullptr;
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int gk = k0 + lk + u;
            ra[u] = (arow && gk < K) ? arow[gk] : 0.0;
            rb[u] = (bcol && gk < K) ? bcol[gk] : 0.0;
        }
    };
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += TK) {
#pragma unroll
        for (int u = 0; u < 4; u++) { As[lk + u][lr] = ra[u]; Bs[lk + u][lr] = rb[u]; }
        __syncthreads();
        if (k0 + TK < K) fetch(k0 + TK);
#pragma unroll
        for (int kk = 0; kk < TK; kk++) {
            double a[4], bb[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) bb[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * bb[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int c = col0 + tx * 4 + j;
        if (c >= g.ncol) continue;
        int oi = out_idx[c];
        if (oi < 0) continue;
        double sc = g.scale_depth ? ldexp(1.0, -g.scale_depth[c]) : 1.0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int r = row0 + ty * 4 + i;
            if (r >= S) continue;
            double v = sc * acc[i][j];
            double* y = &Y[(int64_t)oi * S + r];
            *y = g.accumulate ? *y + v : v;
        }
    }
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

// M2L up (src/kernel.cpp:1286-1328): scatter each child's S equivalent charges on the P^3 corner of
// the (2P-1)^3 grid and r2c-transform it as three pruned DFT passes (x, y, z) held in LDS; the
// last pass writes Qk[f][src][slot].  One workgroup per source parent and slot group.
template <int P>
__global__ __launch_bounds__(256) void k_m2l_up(int S, int nsrc, int G, const int32_t* __restrict__ src_child,
                                                const int32_t* __restrict__ lattice, const double* __restrict__ twiddle,
                                                const double* __restrict__ qeq, double2* __restrict__ Qk) {
    extern __shared__ double lds[];
    const int N = 2 * P - 1, Nh = N / 2 + 1, F = N * N * Nh;
    const int src = blockIdx.x, slot0 = blockIdx.y * G, tid = threadIdx.x, nt = blockDim.x;
    double2* tw = reinterpret_cast<double2*>(lds);               // N
    constexpr int Lsz = (P * P * P + 1) & ~1;                   // even, keeps the double2 arrays 16-byte aligned
    double* L = lds + 2 * N;                                     // P^3
    double2* T1 = reinterpret_cast<double2*>(L + Lsz);           // P*P*Nh
    double2* T2 = T1 + P * P * Nh;                               // G * P*N*Nh
    for (int m = tid; m < N; m += nt) tw[m] = make_double2(twiddle[2 * m], twiddle[2 * m + 1]);
    const int t2sz = P * N * Nh;
    for (int sl = 0; sl < G; sl++) {
        const int c = src_child[src * 8 + slot0 + sl];
        double2* T2s = T2 + sl * t2sz;
        __syncthreads();
        if (c < 0) {
            for (int o = tid; o < t2sz; o += nt) T2s[o] = make_double2(0, 0);
            continue;
        }
        for (int o = tid; o < P * P * P; o += nt) L[o] = 0;
        __syncthreads();
        for (int s = tid; s < S; s += nt) {
            int i = lattice[3 * s], j = lattice[3 * s + 1], k = lattice[3 * s + 2];
            L[(k * P + j) * P + i] = qeq[(size_t)c * S + s];
        }
        __syncthreads();
        for (int o = tid; o < P * P * Nh; o += nt) {   // along x: real -> half spectrum
            int fi = o % Nh, kj = o / Nh;
            double re = 0, im = 0;
            for (int i = 0; i < P; i++) {
                double v = L[kj * P + i];
                double2 w = tw[(i * fi) % N];
                re += v * w.x;
                im += v * w.y;
            }
            T1[o] = make_double2(re, im);
        }
        __syncthreads();
        for (int o = tid; o < t2sz; o += nt) {          // along y
            int fi = o % Nh, r = o / Nh, fj = r % N, k = r / N;
            double re = 0, im = 0;
            for (int j = 0; j < P; j++) {
                double2 v = T1[(k * P + j) * Nh + fi];
                double2 w = tw[(j * fj) % N];
                re += v.x * w.x - v.y * w.y;
                im += v.x * w.y + v.y * w.x;
            }
            T2s[o] = make_double2(re, im);
        }
    }
    __syncthreads();
    for (int o = tid; o < F * G; o += nt) {             // along z, written as Qk[f][src][slot]
        int sl = o % G, f = o / G;
        int fi = f % Nh, r = f / Nh, fj = r % N, fk = r / N;
        const double2* T2s = T2 + sl * t2sz;
        double re = 0, im = 0;
        for (int k = 0; k < P; k++) {
            double2 v = T2s[(k * N + fj) * Nh + fi];
            double2 w = tw[(k * fk) % N];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
        }
        Qk[((size_t)f * nsrc + src) * 8 + slot0 + sl] = make_double2(re, im);
    }
}

// Hadamard 8x8 (src/kernel.cpp:1012-1082, include/c2c.h:201-296): for one frequency f and 256
// target parents per workgroup, Pk[t][it] = sum over the 26 neighbour patterns p of
// sum_is Gk[p][f][is][it] * Qk[src(t,p)][is].  The 26 kernels of f sit in LDS and are read as
// broadcasts; blocks are remapped so that consecutive frequency slices run on one XCD.
__global__ __launch_bounds__(256) void k_hadamard8(int F, int ntar, int nsrc, int nchunk, int total,
                                                   const int32_t* __restrict__ slot_src_T,
                                                   const double* __restrict__ gk8, const double2* __restrict__ Qk,
                                                   double2* __restrict__ Pk) {
    __shared__ double2 gs[26 * 64];
    const int per = (total + 7) / 8;
    const int L = (blockIdx.x % 8) * per + blockIdx.x / 8;  // consecutive logical blocks on one XCD
    if (L >= total) return;
    const int f = L / nchunk, chunk = L % nchunk, tid = threadIdx.x;
    for (int e = tid; e < 26 * 64; e += 256) {
        int p = e >> 6, m = e & 63;
        const double* g = gk8 + ((size_t)p * F + f) * 128 + 2 * m;
        gs[e] = make_double2(g[0], g[1]);
    }
    __syncthreads();
    const int t = chunk * 256 + tid;
    if (t >= ntar) return;
    double ar[8], ai[8];
#pragma unroll
    for (int i = 0; i < 8; i++) ar[i] = ai[i] = 0;
    for (int p = 0; p < 26; p++) {
        const int s = slot_src_T[(size_t)p * ntar + t];
        if (s < 0) continue;
        const double2* q = Qk + ((size_t)f * nsrc + s) * 8;
        double2 qv[8];
#pragma unroll
        for (int is = 0; is < 8; is++) qv[is] = q[is];
        const double2* g = gs + p * 64;
#pragma unroll
        for (int is = 0; is < 8; is++) {
#pragma unroll
            for (int it = 0; it < 8; it++) {
                double2 gg = g[is * 8 + it];
                ar[it] += gg.x * qv[is].x - gg.y * qv[is].y;
                ai[it] += gg.x * qv[is].y + gg.y * qv[is].x;
            }
        }
    }
    double2* o = Pk + ((size_t)f * ntar + t) * 8;
#pragma unroll
    for (int it = 0; it < 8; it++) o[it] = make_double2(ar[it], ai[it]);
}

// Hadamard 1x27 (src/kernel.cpp:1084-1138): image-lattice M2L for the image parents c(-1..-(i-1));
// every one of the 27 lattice children of a neighbouring image is the same single child, so the
// 27 kernels of a pattern are pre-summed.  Only child slot 0 exists.
__global__ __launch_bounds__(256) void k_hadamard27(int F, int ntar, int nsrc, const int32_t* __restrict__ img_tar,
                                                    const int32_t* __restrict__ img_src_T, const double* __restrict__ g27,
                                                    const double2* __restrict__ Qk, double2* __restrict__ Pk) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int t = img_tar[blockIdx.y];
    double re = 0, im = 0;
    for (int p = 0; p < 26; p++) {
        const int s = img_src_T[(size_t)p * ntar + t];
        if (s < 0) continue;
        double2 q = Qk[((size_t)f * nsrc + s) * 8];
        double gr = g27[((size_t)p * F + f) * 2], gi = g27[((size_t)p * F + f) * 2 + 1];
        re += gr * q.x - gi * q.y;
        im += gr * q.y + gi * q.x;
    }
    double2* o = Pk + ((size_t)f * ntar + t) * 8;
    double2 v = *o;
    *o = make_double2(v.x + re, v.y + im);
}

// M2L down (src/kernel.cpp:1343-1385): c2r of each child's Pk evaluated only on the P^3 block that
// holds the shifted check surface, then gather, scale (2 or 3)^(depth+1) / N^3 and accumulate into
// p_check.  Pass 1 (z) reads Pk[f][tar][slots] once into LDS; passes 2 (y) and 3 (x, Hermitian) per child.
template <int P>
__global__ __launch_bounds__(256) void k_m2l_down(int S, int ntar, int G, const int32_t* __restrict__ tar_child,
                                                  const double* __restrict__ tar_scale, const int32_t* __restrict__ lattice,
                                                  const double* __restrict__ twiddle, const double2* __restrict__ Pk,
                                                  double* __restrict__ pck) {
    extern __shared__ double lds[];
    const int N = 2 * P - 1, Nh = N / 2 + 1, N3 = N * N * N;
    const int t = blockIdx.x, slot0 = blockIdx.y * G, tid = threadIdx.x, nt = blockDim.x;
    double2* tw = reinterpret_cast<double2*>(lds);          // N
    double2* T1 = tw + N;                                   // G * P*N*Nh   [sl][kk][fj][fi]
    double2* T2 = T1 + G * P * N * Nh;                      // P*P*Nh       [kk][jj][fi]
    double* out = reinterpret_cast<double*>(T2 + P * P * Nh);  // P^3
    for (int m = tid; m < N; m += nt) tw[m] = make_double2(twiddle[2 * m], twiddle[2 * m + 1]);
    __syncthreads();
    // pass 1: for each (fj, fi, slot) read the N values along z once, emit the P outputs k = P-1+kk
    for (int o = tid; o < G * N * Nh; o += nt) {
        int sl = o % G, r = o / G, fi = r % Nh, fj = r / Nh;
        double re[P], im[P];
#pragma unroll
        for (int kk = 0; kk < P; kk++) re[kk] = im[kk] = 0;
        for (int fk = 0; fk < N; fk++) {
            double2 v = Pk[((size_t)((fk * N + fj) * Nh + fi) * ntar + t) * 8 + slot0 + sl];
#pragma unroll
            for (int kk = 0; kk < P; kk++) {
                double2 w = tw[((P - 1 + kk) * fk) % N];   // conj(w): exp(+2 pi i k fk / N)
                re[kk] += v.x * w.x + v.y * w.y;
                im[kk] += v.y * w.x - v.x * w.y;
            }
        }
#pragma unroll
        for (int kk = 0; kk < P; kk++) T1[((sl * P + kk) * N + fj) * Nh + fi] = make_double2(re[kk], im[kk]);
    }
    const double scale = tar_scale[t];
    for (int sl = 0; sl < G; sl++) {
        const int c = tar_child[t * 8 + slot0 + sl];
        __syncthreads();
        if (c < 0) continue;
        for (int o = tid; o < P * P * Nh; o += nt) {     // pass 2 along y
            int fi = o % Nh, r = o / Nh, jj = r % P, kk = r / P;
            double re = 0, im = 0;
            for (int fj = 0; fj < N; fj++) {
                double2 v = T1[((sl * P + kk) * N + fj) * Nh + fi];
                double2 w = tw[((P - 1 + jj) * fj) % N];
                re += v.x * w.x + v.y * w.y;
                im += v.y * w.x - v.x * w.y;
            }
            T2[o] = make_double2(re, im);
        }
        __syncthreads();
        for (int o = tid; o < P * P * P; o += nt) {      // pass 3 along x: half spectrum -> real
            int ii = o % P, r = o / P;
            const double2* row = T2 + r * Nh;
            double a = row[0].x;
            for (int fi = 1; fi < Nh; fi++) {
                double2 w = tw[((P - 1 + ii) * fi) % N];
                a += 2 * (row[fi].x * w.x + row[fi].y * w.y);
            }
            out[o] = a;
        }
        __syncthreads();
        for (int s = tid; s < S; s += nt) {
            int i = lattice[3 * s], j = lattice[3 * s + 1], k = lattice[3 * s + 2];
            pck[(size_t)c * S + s] += out[(k * P + j) * P + i] / N3 * scale;
        }
    }
}

// L2P evaluation (src/kernel.cpp:1471-1510): potential and gradient at the leaf's bodies from the
// S downward equivalent charges on the 2.95 r surface.  Writes p and grad (first contribution).
__global__ __launch_bounds__(256) void k_l2p_eval(int S, const double* __restrict__ usurf,
                                                  const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                                  const double* __restrict__ qdown, double* __restrict__ p,
                                                  double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ gz) {
    extern __shared__ double lds[];
    double* ex = lds;
    double* ey = ex + S;
    double* ez = ey + S;
    double* eq = ez + S;
    const int li = blockIdx.x, tid = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    for (int s = tid; s < S; s += 256) {
        ex[s] = usurf[3 * s] * rr + cx;
        ey[s] = usurf[3 * s + 1] * rr + cy;
        ez[s] = usurf[3 * s + 2] * rr + cz;
        eq[s] = qdown[(size_t)li * S + s];
    }
    __syncthreads();
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    for (int i = tid; i < bnum; i += 256) {
        const int bi = boff + i;
        const double xt = b.x[bi], yt = b.y[bi], zt = b.z[bi];
        double ap = 0, ax = 0, ay = 0, az = 0;
        for (int s = 0; s < S; s++) {
            double dx = xt - ex[s], dy = yt - ey[s], dz = zt - ez[s];
            double r2 = dx * dx + dy * dy + dz * dz;
            double ir = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
            double qi = eq[s] * ir;
            ap += qi;
            qi *= ir * ir;
            ax += qi * dx;
            ay += qi * dy;
            az += qi * dz;
        }
        p[bi] = ap;
        gx[bi] = -ax;
        gy[bi] = -ay;
        gz[bi] = -az;
    }
}

// P2P near field (src/kernel.cpp:258-311): one wave per target leaf, one target body per lane;
// source coordinates are wave-uniform so they arrive through scalar loads.
__global__ __launch_bounds__(64) void k_p2p(const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                            const int64_t* __restrict__ list_off, const int32_t* __restrict__ list_ent,
                                            double cycle, double* __restrict__ p, double* __restrict__ gx,
                                            double* __restrict__ gy, double* __restrict__ gz) {
    const int li = blockIdx.x, tid = threadIdx.x;
    const int c = leaf_cell[li];
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    const int64_t e0 = list_off[li], e1 = list_off[li + 1];
    for (int t0 = 0; t0 < bnum; t0 += 64) {
        const int i = t0 + tid;
        const bool active = i < bnum;
        const int bi = boff + (active ? i : 0);
        const double xt = b.x[bi], yt = b.y[bi], zt = b.z[bi];
        double ap = 0, ax = 0, ay = 0, az = 0;
        for (int64_t e = e0; e < e1; e++) {
            const int ent = list_ent[e], sc = ent >> 5, oi = ent & 31;   // source cell, periodic offset index
            const double ox = (oi % 3 - 1) * cycle, oy = ((oi / 3) % 3 - 1) * cycle, oz = (oi / 9 - 1) * cycle;
            const int sb = cells.boff[sc], sn = cells.bnum[sc];
            for (int j = sb; j < sb + sn; j++) {
                double dx = xt - (b.x[j] + ox), dy = yt - (b.y[j] + oy), dz = zt - (b.z[j] + oz);
                double r2 = dx * dx + dy * dy + dz * dz;
                double ir = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
                double qi = b.q[j] * ir;
                ap += qi;
                qi *= ir * ir;
                ax += qi * dx;
                ay += qi * dy;
                az += qi * dz;
            }
        }
        if (active) {
            p[bi] += ap;
            gx[bi] -= ax;
            gy[bi] -= ay;
            gz[bi] -= az;
        }
    }
}

// Dipole moment sum_i q_i x_i (src/kernel.cpp:1587-1591) as per-block partial sums.
__global__ __launch_bounds__(256) void k_dipole_partial(BodiesDev b, int n, double* __restrict__ partial) {
    __shared__ double s[3][256];
    const int tid = threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int i = blockIdx.x * 256 + tid; i < n; i += gridDim.x * 256) {
        a0 += b.x[i] * b.q[i];
        a1 += b.y[i] * b.q[i];
        a2 += b.z[i] * b.q[i];
    }
    s[0][tid] = a0; s[1][tid] = a1; s[2][tid] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) { s[0][tid] += s[0][tid + w]; s[1][tid] += s[1][tid + w]; s[2][tid] += s[2][tid + w]; }
        __syncthreads();
    }
    if (tid == 0) { partial[3 * blockIdx.x] = s[0][0]; partial[3 * blockIdx.x + 1] = s[1][0]; partial[3 * blockIdx.x + 2] = s[2][0]; }
}

// Dipole correction (src/kernel.cpp:1580-1596), 1/(4 pi) scaling (src/body.cpp:88-96) and the
// return to input order (src/body.cpp:177-190).
__global__ __launch_bounds__(256) void k_finalize(int n, const int32_t* __restrict__ perm, const double* __restrict__ q,
                                                  const double* __restrict__ p, const double* __restrict__ gx,
                                                  const double* __restrict__ gy, const double* __restrict__ gz,
                                                  const double* __restrict__ partial, int nblocks, int mode, double cycle,
                                                  double out_scale, double* __restrict__ phi, double* __restrict__ grad) {
    __shared__ double M[3];
    if (threadIdx.x < 3) {
        double a = 0;
        for (int k = 0; k < nblocks; k++) a += partial[3 * k + threadIdx.x];
        M[threadIdx.x] = a;
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double pv = p[i], g0 = gx[i], g1 = gy[i], g2 = gz[i];
    if (mode == 1) {
        double coef = 4 * M_PI / (3 * cycle * cycle * cycle);
        double dn = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
        pv -= coef * dn / n / q[i];
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

size_t m2l_up_lds(int P, int G) {
    int N = 2 * P - 1, Nh = N / 2 + 1;
    return sizeof(double) * (2 * N + ((P * P * P + 1) & ~1) + 2 * P * P * Nh + 2 * G * P * N * Nh);
}
size_t m2l_down_lds(int P, int G) {
    int N = 2 * P - 1, Nh = N / 2 + 1;
    return sizeof(double) * (2 * N + 2 * G * P * N * Nh + 2 * P * P * Nh + P * P * P);
}
int pick_group(size_t (*lds)(int, int), int P) {
    for (int G = 8; G >= 1; G /= 2)
        if (lds(P, G) <= 160 * 1024) return G;
    return 1;
}

}  // namespace

hipError_t launch_gather(const double* qin, const int32_t* perm, double* qt, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_gather<<<(n + 255) / 256, 256, 0, st>>>(qin, perm, qt, n);
    return hipGetLastError();
}

hipError_t launch_p2m_check(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                            BodiesDev b, double* pc, hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    if (S > 256 * kMaxPts) return hipErrorInvalidValue;
    k_p2m_check<<<nleaf, 256, 0, st>>>(S, usurf, leaf_cell, cells, b, pc);
    return hipGetLastError();
}

hipError_t launch_gemm_cols(const GemmCols& g, hipStream_t st) {
    if (g.ncol <= 0) return hipSuccess;
    dim3 grid((g.S + TM - 1) / TM, (g.ncol + TN - 1) / TN, g.nz);
    k_gemm_cols<<<grid, 256, 0, st>>>(g);
    return hipGetLastError();
}

hipError_t launch_sum_partials(int S, int ncol, int nz, const double* part, const int32_t* out, double* Y,
                               hipStream_t st) {
    if (ncol <= 0) return hipSuccess;
    int64_t tot = (int64_t)ncol * S;
    k_sum_partials<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, ncol, nz, part, out, Y);
    return hipGetLastError();
}

template <int P>
hipError_t up_p(int S, int nsrc, const int32_t* src_child, const int32_t* lattice, const double* twiddle,
                const double* qeq, double* Qk, hipStream_t st) {
    int G = pick_group(m2l_up_lds, P);
    size_t lds = m2l_up_lds(P, G);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_m2l_up<P>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    k_m2l_up<P><<<dim3(nsrc, 8 / G), 256, lds, st>>>(S, nsrc, G, src_child, lattice, twiddle, qeq,
                                                      reinterpret_cast<double2*>(Qk));
    return hipGetLastError();
}

template <int P>
hipError_t down_p(int S, int ntar, const int32_t* tar_child, const double* tar_scale, const int32_t* lattice,
                  const double* twiddle, const double* Pk, double* pck, hipStream_t st) {
    int G = pick_group(m2l_down_lds, P);
    size_t lds = m2l_down_lds(P, G);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_m2l_down<P>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    k_m2l_down<P><<<dim3(ntar, 8 / G), 256, lds, st>>>(S, ntar, G, tar_child, tar_scale, lattice, twiddle,
                                                        reinterpret_cast<const double2*>(Pk), pck);
    return hipGetLastError();
}

#define FMM3DPP_P_CASES(M) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12)

hipError_t launch_m2l_up(int P, int S, int nsrc, const int32_t* src_child, const int32_t* lattice,
                         const double* twiddle, const double* qeq, double* Qk, hipStream_t st) {
    if (nsrc <= 0) return hipSuccess;
    switch (P) {
#define UPCASE(p) case p: return up_p<p>(S, nsrc, src_child, lattice, twiddle, qeq, Qk, st);
        FMM3DPP_P_CASES(UPCASE)
#undef UPCASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_hadamard8(int F, int ntar, int nsrc, const int32_t* slot_src_T, const double* gk8, const double* Qk,
                            double* Pk, hipStream_t st) {
    if (ntar <= 0) return hipSuccess;
    int nchunk = (ntar + 255) / 256;
    int total = nchunk * F;
    int per = (total + 7) / 8;
    k_hadamard8<<<per * 8, 256, 0, st>>>(F, ntar, nsrc, nchunk, total, slot_src_T, gk8,
                                         reinterpret_cast<const double2*>(Qk), reinterpret_cast<double2*>(Pk));
    return hipGetLastError();
}

hipError_t launch_hadamard27(int F, int nimg, const int32_t* img_tar, const int32_t* img_src_T, int ntar, int nsrc,
                             const double* gk27_sum, const double* Qk, double* Pk, hipStream_t st) {
    if (nimg <= 0) return hipSuccess;
    k_hadamard27<<<dim3((F + 255) / 256, nimg), 256, 0, st>>>(F, ntar, nsrc, img_tar, img_src_T, gk27_sum,
                                                              reinterpret_cast<const double2*>(Qk),
                                                              reinterpret_cast<double2*>(Pk));
    return hipGetLastError();
}

hipError_t launch_m2l_down(int P, int S, int ntar, const int32_t* tar_child, const double* tar_scale,
                           const int32_t* lattice, const double* twiddle, const double* Pk, double* pck,
                           hipStream_t st) {
    if (ntar <= 0) return hipSuccess;
    switch (P) {
#define DNCASE(p) case p: return down_p<p>(S, ntar, tar_child, tar_scale, lattice, twiddle, Pk, pck, st);
        FMM3DPP_P_CASES(DNCASE)
#undef DNCASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_l2p_eval(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                           BodiesDev b, const double* qdown, double* p, double* gx, double* gy, double* gz,
                           hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    k_l2p_eval<<<nleaf, 256, 4 * S * sizeof(double), st>>>(S, usurf, leaf_cell, cells, b, qdown, p, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launch_p2p(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, const int64_t* list_off,
                      const int32_t* list_ent, double cycle, double* p, double* gx, double* gy, double* gz,
                      hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    k_p2p<<<nleaf, 64, 0, st>>>(leaf_cell, cells, b, list_off, list_ent, cycle, p, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launch_dipole_partial(BodiesDev b, int n, double* partial, int nblocks, hipStream_t st) {
    k_dipole_partial<<<nblocks, 256, 0, st>>>(b, n, partial);
    return hipGetLastError();
}

hipError_t launch_finalize(int n, const int32_t* perm, const double* q, const double* p, const double* gx,
                           const double* gy, const double* gz, const double* partial, int nblocks, int mode,
                           double cycle, double out_scale, double* phi, double* grad, hipStream_t st) {
    k_finalize<<<(n + 255) / 256, 256, 0, st>>>(n, perm, q, p, gx, gy, gz, partial, nblocks, mode, cycle, out_scale,
                                                phi, grad);
    return hipGetLastError();
}

}  // namespace fmm3dpp
