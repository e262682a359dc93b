// Particle-level kernels of the periodic KIFMM for CDNA4 (gfx950): P2P near field (float64 and
// float32 forms), P2M check potentials and L2P evaluation.
//
// All three are pair sums of 1/r over (target, source) point pairs inside one leaf's wave:
//   lanes = W target slots x G source groups, W = the smallest of 8/16/32/64 that holds the leaf's
//   bodies, G = 64 / W.  Group g takes every G-th source, the G partial sums of a target are
//   folded with cross-lane shuffles.  Uniform trees hold about ncrit/2 bodies per leaf, so a lane
//   per target alone would leave half of every wave idle.
// 1/r uses the hardware reciprocal square root (v_rsq_f64 / v_rsq_f32).  Measured on gfx950 over
// r^2 in [1e-12, 1e2]: v_rsq_f64 is within 5.2e-8 relative, v_rsq_f32 within 1.1e-7.  The float64
// kernels refine v_rsq_f64 by one third-order step of the series (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8
// + ..., e = 1 - r^2 y0^2 formed with one fma: the dropped terms are O(e^3) ~ 1e-22, so the result
// is within 1 ulp of the correctly rounded 1/sqrt (the reference's 1/sqrt(r^2), src/type.h:493-496)
// for any start within 1e-7 (checked in exact arithmetic on the host: 0.95 ulp worst over 30,000
// cases); one Newton step instead leaves up to ~40 ulp from the measured start.
#include <hip/hip_runtime.h>

#include "device.h"

namespace fmm3dpp {

namespace {

// y0 (1 + e/2 + 3e^2/8), e = 1 - r2 y0^2, in the order the host check follows
__device__ __forceinline__ double rsq_refine(double r2, double y0) {
    const double h = r2 * y0;
    const double e = fma(-h, y0, 1.0);
    const double c = e * fma(0.375, e, 0.5);
    return fma(y0, c, y0);
}

// 1/sqrt(r2) for r2 > 0, 0 for r2 == 0 (the reference's r > 0 mask, src/kernel.cpp:283-290)
__device__ __forceinline__ double inv_sqrt(double r2) {
    const double y = rsq_refine(r2, __builtin_amdgcn_rsq(r2));
    return r2 > 0 ? y : 0.0;
}

// For r2 > 0 known in advance: bodies of a leaf against its check / equivalent surface, which lies
// 2.95 r from the centre while the bodies are within sqrt(3) r, so r2 >= (1.95 r)^2.
__device__ __forceinline__ double inv_sqrt_pos(double r2) { return rsq_refine(r2, __builtin_amdgcn_rsq(r2)); }

__device__ __forceinline__ float inv_sqrt(float r2) {
    const float y = __builtin_amdgcn_rsqf(r2);
    return r2 > 0 ? y : 0.0f;
}

// log2 of the target slots W of a leaf with n bodies
__device__ __forceinline__ int slot_bits(int n) { return n <= 8 ? 3 : n <= 16 ? 4 : n <= 32 ? 5 : 6; }

template <typename T>
__device__ __forceinline__ T fold_groups(T v, int wbits) {
    for (int m = 1 << wbits; m < 64; m <<= 1) v += __shfl_xor(v, m);
    return v;
}

// P2P, float64 (src/kernel.cpp:258-311 p2p_1toN_256 over the leaf's P2P list, fmm.cpp:259-294):
// phi_t += q_s / r, grad_t -= q_s (x_t - x_s) / r^3 with x_s shifted by the list entry's periodic
// offset.  One wave per target leaf.  The list is walked in blocks of 64 entries whose source
// ranges go to LDS first; each source leaf is then read in pieces of up to 64 bodies, one body per
// lane into LDS, with the next piece's loads issued before the current piece is summed.
__global__ __launch_bounds__(64) void k_p2p_f64(const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                                const int64_t* __restrict__ list_off,
                                                const int32_t* __restrict__ list_ent, double cycle,
                                                double* __restrict__ p, double* __restrict__ gx,
                                                double* __restrict__ gy, double* __restrict__ gz, bool accumulate) {
    __shared__ int m_sb[64], m_ns[64], m_oi[64];
    __shared__ double4 sbuf[64];
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    const int64_t e0 = list_off[li], e1 = list_off[li + 1];
    const int wb = slot_bits(bnum), W = 1 << wb, G = 64 >> wb;
    const int g = lane >> wb, tl = lane & (W - 1);
    auto piece = [&](int k, int k0) {   // this lane's body of piece (entry k, bodies k0..k0+63)
        const int j = k0 + lane;
        if (j >= m_ns[k]) return make_double4(0, 0, 0, 0);
        const int idx = m_sb[k] + j;
        return make_double4(b.x[idx], b.y[idx], b.z[idx], b.q[idx]);
    };
    for (int t0 = 0; t0 < bnum; t0 += W) {
        const int i = t0 + tl;
        const bool active = i < bnum;
        const int bi = boff + (active ? i : 0);
        const double xt = b.x[bi], yt = b.y[bi], zt = b.z[bi];
        double ap = 0, ax = 0, ay = 0, az = 0;
        for (int64_t eb = e0; eb < e1; eb += 64) {
            const int ne = (int)min((int64_t)64, e1 - eb);
            __syncthreads();
            if (lane < ne) {
                const int ent = list_ent[eb + lane], sc = ent >> 5;   // source cell, periodic offset index
                m_sb[lane] = cells.boff[sc];
                m_ns[lane] = cells.bnum[sc];
                m_oi[lane] = ent & 31;
            }
            __syncthreads();
            int k = 0, k0 = 0;
            double4 r = piece(0, 0);
            while (k < ne) {
                sbuf[lane] = r;
                __syncthreads();
                int nk = k, nk0 = k0 + 64;
                if (nk0 >= m_ns[k]) { nk++; nk0 = 0; }
                if (nk < ne) r = piece(nk, nk0);   // in flight while this piece is summed
                const int oi = m_oi[k], cnt = min(64, m_ns[k] - k0);
                const double xo = xt - (oi % 3 - 1) * cycle, yo = yt - ((oi / 3) % 3 - 1) * cycle,
                             zo = zt - (oi / 9 - 1) * cycle;
#pragma unroll 4   // independent pairs in flight; the sums keep their order
                for (int j = g; j < cnt; j += G) {
                    const double4 sv = sbuf[j];
                    const double dx = xo - sv.x, dy = yo - sv.y, dz = zo - sv.z;
                    const double ir = inv_sqrt(dx * dx + dy * dy + dz * dz);
                    double qi = sv.w * ir;
                    ap += qi;
                    qi *= ir * ir;
                    ax += qi * dx;
                    ay += qi * dy;
                    az += qi * dz;
                }
                __syncthreads();
                k = nk;
                k0 = nk0;
            }
        }
        ap = fold_groups(ap, wb);
        ax = fold_groups(ax, wb);
        ay = fold_groups(ay, wb);
        az = fold_groups(az, wb);
        if (active && g == 0) {
            if (accumulate) {
                p[bi] += ap;
                gx[bi] -= ax;
                gy[bi] -= ay;
                gz[bi] -= az;
            } else {
                p[bi] = ap;
                gx[bi] = -ax;
                gy[bi] = -ay;
                gz[bi] = -az;
            }
        }
    }
}

// float32 body records for the mixed-precision P2P: (x, y, z) relative to the body's own leaf
// centre and q, so that every float32 coordinate is of the order of the leaf size.
__global__ __launch_bounds__(64) void k_pack_rel(const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                                 float4* __restrict__ rel) {
    const int c = leaf_cell[blockIdx.x];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c];
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    for (int i = threadIdx.x; i < bnum; i += 64) {
        const int j = boff + i;
        rel[j] = make_float4((float)(b.x[j] - cx), (float)(b.y[j] - cy), (float)(b.z[j] - cz), (float)b.q[j]);
    }
}

// P2P, float32 with two-level summation: the pairs of one (target, piece of a source leaf) are
// summed in float32, each piece's partial is added to float64 accumulators.  The source leaf's
// displacement (its centre + periodic offset - the target leaf centre) is formed in float64 once
// per list entry, so no float32 coordinate carries an absolute position.  List blocks and source
// pieces are staged through LDS as in k_p2p_f64.
__global__ __launch_bounds__(64) void k_p2p_f32(const int32_t* __restrict__ leaf_cell, CellsDev cells,
                                                const float4* __restrict__ rel, const int64_t* __restrict__ list_off,
                                                const int32_t* __restrict__ list_ent, double cycle,
                                                double* __restrict__ p, double* __restrict__ gx,
                                                double* __restrict__ gy, double* __restrict__ gz, bool accumulate) {
    __shared__ int m_sb[64], m_ns[64];
    __shared__ float m_dx[64], m_dy[64], m_dz[64];
    __shared__ float4 sbuf[64];
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c];
    const int64_t e0 = list_off[li], e1 = list_off[li + 1];
    const int wb = slot_bits(bnum), W = 1 << wb, G = 64 >> wb;
    const int g = lane >> wb, tl = lane & (W - 1);
    auto piece = [&](int k, int k0) {
        const int j = k0 + lane;
        return j < m_ns[k] ? rel[m_sb[k] + j] : make_float4(0, 0, 0, 0);
    };
    for (int t0 = 0; t0 < bnum; t0 += W) {
        const int i = t0 + tl;
        const bool active = i < bnum;
        const int bi = boff + (active ? i : 0);
        const float4 tv = rel[bi];
        double ap = 0, ax = 0, ay = 0, az = 0;
        for (int64_t eb = e0; eb < e1; eb += 64) {
            const int ne = (int)min((int64_t)64, e1 - eb);
            __syncthreads();
            if (lane < ne) {
                const int ent = list_ent[eb + lane], sc = ent >> 5, oi = ent & 31;
                m_sb[lane] = cells.boff[sc];
                m_ns[lane] = cells.bnum[sc];
                m_dx[lane] = (float)(cells.x[sc] + (oi % 3 - 1) * cycle - cx);
                m_dy[lane] = (float)(cells.y[sc] + ((oi / 3) % 3 - 1) * cycle - cy);
                m_dz[lane] = (float)(cells.z[sc] + (oi / 9 - 1) * cycle - cz);
            }
            __syncthreads();
            int k = 0, k0 = 0;
            float4 r = piece(0, 0);
            while (k < ne) {
                sbuf[lane] = r;
                __syncthreads();
                int nk = k, nk0 = k0 + 64;
                if (nk0 >= m_ns[k]) { nk++; nk0 = 0; }
                if (nk < ne) r = piece(nk, nk0);   // in flight while this piece is summed
                const int cnt = min(64, m_ns[k] - k0);
                const float xo = tv.x - m_dx[k], yo = tv.y - m_dy[k], zo = tv.z - m_dz[k];
                float fp = 0, fx = 0, fy = 0, fz = 0;
#pragma unroll 4   // independent pairs in flight; the sums keep their order
                for (int j = g; j < cnt; j += G) {
                    const float4 sv = sbuf[j];
                    const float dx = xo - sv.x, dy = yo - sv.y, dz = zo - sv.z;
                    const float ir = inv_sqrt(dx * dx + dy * dy + dz * dz);
                    float qi = sv.w * ir;
                    fp += qi;
                    qi *= ir * ir;
                    fx += qi * dx;
                    fy += qi * dy;
                    fz += qi * dz;
                }
                ap += fp;
                ax += fx;
                ay += fy;
                az += fz;
                __syncthreads();
                k = nk;
                k0 = nk0;
            }
        }
        ap = fold_groups(ap, wb);
        ax = fold_groups(ax, wb);
        ay = fold_groups(ay, wb);
        az = fold_groups(az, wb);
        if (active && g == 0) {
            if (accumulate) {
                p[bi] += ap;
                gx[bi] -= ax;
                gy[bi] -= ay;
                gz[bi] -= az;
            } else {
                p[bi] = ap;
                gx[bi] = -ax;
                gy[bi] = -ay;
                gz[bi] = -az;
            }
        }
    }
}

// P2M source -> check-surface potentials (src/kernel.cpp:332-368): one wave per leaf, KP check
// points per lane, the leaf's bodies staged through LDS 64 at a time.
template <int KP>
__global__ __launch_bounds__(64) void k_p2m_check(int S, const double* __restrict__ usurf,
                                                  const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                                  double* __restrict__ pc) {
    __shared__ double4 sb[64];   // (x, y, z, q) of 64 bodies
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    double px[KP], py[KP], pz[KP], acc[KP];
#pragma unroll
    for (int m = 0; m < KP; m++) {
        const int s = lane + 64 * m;
        acc[m] = 0;
        const int sr = s < S ? s : 0;   // slots past S evaluate point 0 and are not written
        px[m] = usurf[3 * sr] * rr + cx;
        py[m] = usurf[3 * sr + 1] * rr + cy;
        pz[m] = usurf[3 * sr + 2] * rr + cz;
    }
    for (int b0 = 0; b0 < bnum; b0 += 64) {
        const int nb = min(64, bnum - b0);
        __syncthreads();
        if (lane < nb) {
            const int bj = boff + b0 + lane;
            sb[lane] = make_double4(b.x[bj], b.y[bj], b.z[bj], b.q[bj]);
        }
        __syncthreads();
        for (int j = 0; j < nb; j++) {
            const double4 e = sb[j];
#pragma unroll
            for (int m = 0; m < KP; m++) {
                const double dx = px[m] - e.x, dy = py[m] - e.y, dz = pz[m] - e.z;
                acc[m] += e.w * inv_sqrt_pos(dx * dx + dy * dy + dz * dz);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < KP; m++) {
        const int s = lane + 64 * m;
        if (s < S) pc[(size_t)li * S + s] = acc[m];
    }
}

// L2P evaluation (src/kernel.cpp:1471-1510): potential and gradient at the leaf's bodies from the
// S downward equivalent charges on the 2.95 r surface.  One wave per leaf, equivalent points in
// LDS, lanes split as in P2P.  Writes p and grad (the first contribution of the evaluation).
__global__ __launch_bounds__(64) void k_l2p_eval(int S, const double* __restrict__ usurf,
                                                 const int32_t* __restrict__ leaf_cell, CellsDev cells, BodiesDev b,
                                                 const double* __restrict__ qdown, double* __restrict__ p,
                                                 double* __restrict__ gx, double* __restrict__ gy,
                                                 double* __restrict__ gz) {
    extern __shared__ double4 epts[];   // S equivalent points (x, y, z, density)
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    for (int s = lane; s < S; s += 64)
        epts[s] = make_double4(usurf[3 * s] * rr + cx, usurf[3 * s + 1] * rr + cy, usurf[3 * s + 2] * rr + cz,
                               qdown[(size_t)li * S + s]);
    __syncthreads();
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    const int wb = slot_bits(bnum), W = 1 << wb, G = 64 >> wb;
    const int g = lane >> wb, tl = lane & (W - 1);
    for (int t0 = 0; t0 < bnum; t0 += W) {
        const int i = t0 + tl;
        const bool active = i < bnum;
        const int bi = boff + (active ? i : 0);
        const double xt = b.x[bi], yt = b.y[bi], zt = b.z[bi];
        double ap = 0, ax = 0, ay = 0, az = 0;
#pragma unroll 4   // independent evaluations in flight; the sums keep their order
        for (int s = g; s < S; s += G) {
            const double4 e = epts[s];
            const double dx = xt - e.x, dy = yt - e.y, dz = zt - e.z;
            const double ir = inv_sqrt_pos(dx * dx + dy * dy + dz * dz);
            double qi = e.w * ir;
            ap += qi;
            qi *= ir * ir;
            ax += qi * dx;
            ay += qi * dy;
            az += qi * dz;
        }
        ap = fold_groups(ap, wb);
        ax = fold_groups(ax, wb);
        ay = fold_groups(ay, wb);
        az = fold_groups(az, wb);
        if (active && g == 0) {
            p[bi] = ap;
            gx[bi] = -ax;
            gy[bi] = -ay;
            gz[bi] = -az;
        }
    }
}

// Float32 forms of the two kernels above for the mixed path's float32 operator applies: the same
// waves and lanes, with coordinates taken relative to the leaf's centre (float32 differences of
// values below 3 r), v_rsq_f32, float32 sums, float64 outputs.
template <int KP>
__global__ __launch_bounds__(64) void k_p2m_check_f32(int S, const double* __restrict__ usurf,
                                                      const int32_t* __restrict__ leaf_cell, CellsDev cells,
                                                      BodiesDev b, double* __restrict__ pc) {
    __shared__ float4 sb[64];   // (x, y, z relative to the centre, q) of 64 bodies
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    float px[KP], py[KP], pz[KP], acc[KP];
#pragma unroll
    for (int m = 0; m < KP; m++) {
        const int s = lane + 64 * m;
        acc[m] = 0;
        const int sr = s < S ? s : 0;   // slots past S evaluate point 0 and are not written
        px[m] = (float)(usurf[3 * sr] * rr);
        py[m] = (float)(usurf[3 * sr + 1] * rr);
        pz[m] = (float)(usurf[3 * sr + 2] * rr);
    }
    for (int b0 = 0; b0 < bnum; b0 += 64) {
        const int nb = min(64, bnum - b0);
        __syncthreads();
        if (lane < nb) {
            const int bj = boff + b0 + lane;
            sb[lane] = make_float4((float)(b.x[bj] - cx), (float)(b.y[bj] - cy), (float)(b.z[bj] - cz), (float)b.q[bj]);
        }
        __syncthreads();
        for (int j = 0; j < nb; j++) {
            const float4 e = sb[j];
#pragma unroll
            for (int m = 0; m < KP; m++) {
                const float dx = px[m] - e.x, dy = py[m] - e.y, dz = pz[m] - e.z;
                acc[m] += e.w * __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < KP; m++) {
        const int s = lane + 64 * m;
        if (s < S) pc[(size_t)li * S + s] = (double)acc[m];
    }
}

__global__ __launch_bounds__(64) void k_l2p_eval_f32(int S, const double* __restrict__ usurf,
                                                     const int32_t* __restrict__ leaf_cell, CellsDev cells,
                                                     BodiesDev b, const double* __restrict__ qdown,
                                                     double* __restrict__ p, double* __restrict__ gx,
                                                     double* __restrict__ gy, double* __restrict__ gz) {
    extern __shared__ float4 eptf[];   // S equivalent points relative to the centre, and their densities
    const int li = blockIdx.x, lane = threadIdx.x;
    const int c = leaf_cell[li];
    const double cx = cells.x[c], cy = cells.y[c], cz = cells.z[c], rr = cells.r[c] * 2.95;
    for (int s = lane; s < S; s += 64)
        eptf[s] = make_float4((float)(usurf[3 * s] * rr), (float)(usurf[3 * s + 1] * rr), (float)(usurf[3 * s + 2] * rr),
                              (float)qdown[(size_t)li * S + s]);
    __syncthreads();
    const int boff = cells.boff[c], bnum = cells.bnum[c];
    const int wb = slot_bits(bnum), W = 1 << wb, G = 64 >> wb;
    const int g = lane >> wb, tl = lane & (W - 1);
    for (int t0 = 0; t0 < bnum; t0 += W) {
        const int i = t0 + tl;
        const bool active = i < bnum;
        const int bi = boff + (active ? i : 0);
        const float xt = (float)(b.x[bi] - cx), yt = (float)(b.y[bi] - cy), zt = (float)(b.z[bi] - cz);
        float ap = 0, ax = 0, ay = 0, az = 0;
#pragma unroll 4
        for (int s = g; s < S; s += G) {
            const float4 e = eptf[s];
            const float dx = xt - e.x, dy = yt - e.y, dz = zt - e.z;
            const float ir = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
            float qi = e.w * ir;
            ap += qi;
            qi *= ir * ir;
            ax += qi * dx;
            ay += qi * dy;
            az += qi * dz;
        }
        ap = fold_groups(ap, wb);
        ax = fold_groups(ax, wb);
        ay = fold_groups(ay, wb);
        az = fold_groups(az, wb);
        if (active && g == 0) {
            p[bi] = (double)ap;
            gx[bi] = -(double)ax;
            gy[bi] = -(double)ay;
            gz[bi] = -(double)az;
        }
    }
}

template <int KP>
hipError_t p2m_kp(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b,
                  double* pc, bool f32, hipStream_t st) {
    if (f32)
        k_p2m_check_f32<KP><<<nleaf, 64, 0, st>>>(S, usurf, leaf_cell, cells, b, pc);
    else
        k_p2m_check<KP><<<nleaf, 64, 0, st>>>(S, usurf, leaf_cell, cells, b, pc);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_p2m_check(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                            BodiesDev b, double* pc, hipStream_t st, bool f32) {
    if (nleaf <= 0) return hipSuccess;
    if (!usurf || !leaf_cell || !pc || !valid(cells) || !valid(b)) return hipErrorInvalidValue;
    switch ((S + 63) / 64) {
        case 1: return p2m_kp<1>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 2: return p2m_kp<2>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 3: return p2m_kp<3>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 4: return p2m_kp<4>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 5: return p2m_kp<5>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 6: return p2m_kp<6>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 7: return p2m_kp<7>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 8: return p2m_kp<8>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 9: return p2m_kp<9>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 10: return p2m_kp<10>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 11: return p2m_kp<11>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        case 12: return p2m_kp<12>(S, usurf, leaf_cell, nleaf, cells, b, pc, f32, st);
        default: return hipErrorInvalidValue;   // S <= 728 (P <= 12)
    }
}

hipError_t launch_l2p_eval(int S, const double* usurf, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                           BodiesDev b, const double* qdown, double* p, double* gx, double* gy, double* gz,
                           hipStream_t st, bool f32) {
    if (nleaf <= 0) return hipSuccess;
    if (!usurf || !leaf_cell || !qdown || !p || !gx || !gy || !gz || !valid(cells) || !valid(b) || S <= 0 || S > 728)
        return hipErrorInvalidValue;
    if (f32)
        k_l2p_eval_f32<<<nleaf, 64, (size_t)S * sizeof(float4), st>>>(S, usurf, leaf_cell, cells, b, qdown, p, gx, gy,
                                                                       gz);
    else
        k_l2p_eval<<<nleaf, 64, 4 * S * sizeof(double), st>>>(S, usurf, leaf_cell, cells, b, qdown, p, gx, gy, gz);
    return hipGetLastError();
}

hipError_t launch_p2p(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, const int64_t* list_off,
                      const int32_t* list_ent, double cycle, double* p, double* gx, double* gy, double* gz,
                      bool accumulate, hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    if (!leaf_cell || !list_off || !list_ent || !p || !gx || !gy || !gz || !valid(cells) || !valid(b))
        return hipErrorInvalidValue;
    k_p2p_f64<<<nleaf, 64, 0, st>>>(leaf_cell, cells, b, list_off, list_ent, cycle, p, gx, gy, gz, accumulate);
    return hipGetLastError();
}

hipError_t launch_pack_rel(const int32_t* leaf_cell, int nleaf, CellsDev cells, BodiesDev b, float4* rel,
                           hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    if (!leaf_cell || !rel || !valid(cells) || !valid(b)) return hipErrorInvalidValue;
    k_pack_rel<<<nleaf, 64, 0, st>>>(leaf_cell, cells, b, rel);
    return hipGetLastError();
}

hipError_t launch_p2p_f32(const int32_t* leaf_cell, int nleaf, CellsDev cells, const float4* rel,
                          const int64_t* list_off, const int32_t* list_ent, double cycle, double* p, double* gx,
                          double* gy, double* gz, bool accumulate, hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    if (!leaf_cell || !rel || !list_off || !list_ent || !p || !gx || !gy || !gz || !valid(cells))
        return hipErrorInvalidValue;
    k_p2p_f32<<<nleaf, 64, 0, st>>>(leaf_cell, cells, rel, list_off, list_ent, cycle, p, gx, gy, gz, accumulate);
    return hipGetLastError();
}

}  // namespace fmm3dpp
