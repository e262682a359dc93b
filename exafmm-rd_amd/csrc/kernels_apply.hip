// Operator applies of the periodic KIFMM on the matrix cores of CDNA4 (gfx950): the P2M / L2P
// pseudo-inverse factors (3dpp/src/kernel.cpp:369-370, 1465-1467), M2M per child octant (395-404),
// L2L per child octant (1419-1428) and the image-level M2M / L2L (434-442, 1430-1435), each as
//   Y[out[c]][r] (+)= 2^-depth[c] * sum_k A_z[r][k] X[in[c]][k]
// over gathered columns c, in float64 (v_mfma_f64_16x16x4f64) or float32 (v_mfma_f32_16x16x4f32)
// with float64 vectors in HBM either way.
//
// Operators live in HBM in a lane layout made once at upload: element [kc][b][u][lane] of an M x K
// operator is A[16 b + lane % 16][16 kc + 4 u + lane / 16] (zero outside M x K), i.e. chunk kc of 16 k,
// 16-row block b, matrix-op step u, lane.  One step's B operand of a block is then one contiguous
// 64-element row, so a workgroup copies a chunk of every block it computes into LDS as one
// contiguous range, and a wave reads its operand for a step with one conflict-free LDS load.
//
// Every output element sees the same sequence of matrix ops whatever the tile: K in steps of 4 from
// k = 0, lane kq = lane / 16 feeding k = 4 step + kq.  The tiled and the few-column kernel therefore
// give the same bits, and so do all tile shapes (the batch of merged systems and each system's own
// plan take different tiles for the same column).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.h"

namespace fmm3dpp {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T> struct Mfma;
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator register q of lane (n, kq) holds D[kq + 4 q][n] (rows interleaved)
    static __device__ __forceinline__ int row(int kq, int q) { return kq + 4 * q; }
};
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator register q of lane (n, kq) holds D[4 kq + q][n] (rows blocked)
    static __device__ __forceinline__ int row(int kq, int q) { return 4 * kq + q; }
};

constexpr int kSmallCols = 128;   // column counts up to this go to k_apply_small

// Tiled apply.  Workgroup: WAVES waves, 16 columns each, all computing the same MBW 16-row blocks
// [b0, b0 + MBW) of the output; the blocks' chunk of the operator is double-buffered in LDS (one
// barrier per chunk), the next chunk and each lane's next 4 vector elements are read into registers
// while the current chunk's matrix ops run.  D[v][r] = sum_k X[v][k] A[r][k] is computed with the
// vectors as the row operand, so a lane's accumulators are 16 consecutive components of one output
// vector per 16 lanes (128-byte stores).  blockIdx.x is remapped so that the row tiles of a column
// tile run on one XCD (consecutive logical tiles per XCD) and share its L2 copy of the vectors.
template <class T, int MBW, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_apply(ApplyCols g, LaneOp op, int nmt, int ntiles) {
    typedef typename Mfma<T>::acc_t acc_t;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);   // [2][MBW * 256]
    const int per = (ntiles + 7) / 8;
    const int L = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (L >= ntiles) return;   // whole workgroup, before any barrier
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, n = lane & 15, kq = lane >> 4;
    const int mt = L % nmt, ct = L / nmt, z = blockIdx.y;
    const int b0 = mt * MBW, nb = min(MBW, op.nmb - b0);
    const int M = op.M, K = op.K, nkc = op.nkc;
    const T* A = static_cast<const T*>(op.p) + (int64_t)z * op.zstride + (int64_t)b0 * 256;
    const int64_t achunk = (int64_t)op.nmb * 256;
    const int col0 = ct * 16 * WAVES + wave * 16;
    const int cv = col0 + n;   // the vector this lane feeds
    const int src = cv < g.ncol ? g.in_idx[(int64_t)z * g.idx_zstride + cv] : -1;
    const double* xr = src >= 0 ? g.X + (int64_t)src * g.S : nullptr;

    // chunk copies: 16-byte pieces, nb * 256 elements of the chunk
    constexpr int EPV = 16 / (int)sizeof(T);
    constexpr int NVMAX = MBW * 256 / EPV;
    constexpr int VPT = (NVMAX + 64 * WAVES - 1) / (64 * WAVES);
    typedef T vec_t __attribute__((ext_vector_type(EPV)));
    const int nv = nb * 256 / EPV;
    vec_t ra[VPT];
    auto fetch_a = [&](int kc) {
        const vec_t* s = reinterpret_cast<const vec_t*>(A + (int64_t)kc * achunk);
#pragma unroll
        for (int i = 0; i < VPT; i++) {
            const int v = tid + i * 64 * WAVES;
            if (v < nv) ra[i] = s[v];
        }
    };
    auto store_a = [&](int buf) {
        vec_t* d = reinterpret_cast<vec_t*>(lds + buf * MBW * 256);
#pragma unroll
        for (int i = 0; i < VPT; i++) {
            const int v = tid + i * 64 * WAVES;
            if (v < nv) d[v] = ra[i];
        }
    };
    T rx[4];
    auto fetch_x = [&](int kc) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = 16 * kc + 4 * u + kq;
            rx[u] = (xr && k < K) ? (T)xr[k] : (T)0;
        }
    };

    acc_t acc[MBW];
#pragma unroll
    for (int b = 0; b < MBW; b++) acc[b] = acc_t{0, 0, 0, 0};
    fetch_a(0);
    fetch_x(0);
    store_a(0);
    __syncthreads();
    for (int kc = 0; kc < nkc; kc++) {
        T xv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = rx[u];
        const bool more = kc + 1 < nkc;
        if (more) {
            fetch_a(kc + 1);
            fetch_x(kc + 1);
        }
        const T* as = lds + (kc & 1) * MBW * 256 + lane;
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int b = 0; b < MBW; b++)
                if (b < nb) acc[b] = Mfma<T>::op(xv[u], as[(b * 4 + u) * 64], acc[b]);
        if (more) store_a((kc + 1) & 1);
        __syncthreads();
    }

    double* Y = g.Y + (int64_t)z * g.y_zstride;
    const int32_t* out_idx = g.out_idx + (int64_t)z * g.idx_zstride;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = col0 + Mfma<T>::row(kq, q);
        if (c >= g.ncol) continue;
        const int oi = out_idx[c];
        if (oi < 0) continue;
        const double sc = g.scale_depth ? ldexp(1.0, -g.scale_depth[c]) : 1.0;
#pragma unroll
        for (int b = 0; b < MBW; b++) {
            const int r = 16 * (b0 + b) + n;
            if (b >= nb || r >= M) continue;
            const double v = sc * (double)acc[b][q];
            double* y = &Y[(int64_t)oi * g.S + r];
            *y = g.accumulate ? *y + v : v;
        }
    }
}

// Few columns (the upper tree levels and the image levels): one wave per 16 x 16 output block,
// operands read straight from L2 (the lane layout makes each step's operand one coalesced 64-element
// row) with 32 steps of loads in flight (a launch is one chain of dependent matrix ops per wave, so
// its time is the load latency per batch of steps plus the chain: 8 steps in flight took 15-27 us for
// K = 296), instead of an LDS stage behind a barrier per chunk.  Same sequence of matrix ops per
// output element as k_apply.
template <class T>
__global__ __launch_bounds__(64) void k_apply_small(ApplyCols g, LaneOp op) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x, n = lane & 15, kq = lane >> 4;
    const int rb = blockIdx.x, vb = blockIdx.y, z = blockIdx.z;   // 16 components x 16 vectors
    const int M = op.M, K = op.K;
    const T* A = static_cast<const T*>(op.p) + (int64_t)z * op.zstride + (int64_t)rb * 256 + lane;
    const int64_t achunk = (int64_t)op.nmb * 256;
    const int cv = vb * 16 + n;
    const int src = cv < g.ncol ? g.in_idx[(int64_t)z * g.idx_zstride + cv] : -1;
    const double* xr = src >= 0 ? g.X + (int64_t)src * g.S : nullptr;
    acc_t acc = {0, 0, 0, 0};
    const int nsteps = 4 * op.nkc;
    constexpr int U = 32;
    for (int s0 = 0; s0 < nsteps; s0 += U) {
        T xv[U], av[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int s = s0 + u, k = 4 * s + kq;
            xv[u] = (xr && k < K) ? (T)xr[k] : (T)0;
            av[u] = s < nsteps ? A[(int64_t)(s >> 2) * achunk + (s & 3) * 64] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (s0 + u < nsteps) acc = Mfma<T>::op(xv[u], av[u], acc);
    }
    double* Y = g.Y + (int64_t)z * g.y_zstride;
    const int32_t* out_idx = g.out_idx + (int64_t)z * g.idx_zstride;
    const int r = rb * 16 + n;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = vb * 16 + Mfma<T>::row(kq, q);
        if (c >= g.ncol || r >= M) continue;
        const int oi = out_idx[c];
        if (oi < 0) continue;
        const double sc = g.scale_depth ? ldexp(1.0, -g.scale_depth[c]) : 1.0;
        const double v = sc * (double)acc[q];
        double* y = &Y[(int64_t)oi * g.S + r];
        *y = g.accumulate ? *y + v : v;
    }
}

// Few columns, staged: the launch is one chain of dependent matrix ops per output block, so its time
// is the chain plus the latency of its loads.  Four waves copy every operand of the block's chain --
// the block's rows of all K chunks and the 16 columns' K values, both in step x lane order -- into
// LDS with all loads in flight at once (about 40 per thread at P = 8), then wave 0 runs the chain
// from LDS.  Same sequence of matrix ops per output element as k_apply and k_apply_small.
template <class T>
__global__ __launch_bounds__(256) void k_apply_small_lds(ApplyCols g, LaneOp op) {
    typedef typename Mfma<T>::acc_t acc_t;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int nsteps = 4 * op.nkc;
    T* As = reinterpret_cast<T*>(lds_raw);   // [step][lane]
    T* Xs = As + (size_t)nsteps * 64;        // [step][lane]
    int* src_s = reinterpret_cast<int*>(Xs + (size_t)nsteps * 64);   // the 16 columns' vectors
    const int tid = threadIdx.x;
    const int rb = blockIdx.x, vb = blockIdx.y, z = blockIdx.z;
    const int M = op.M, K = op.K;
    const T* A = static_cast<const T*>(op.p) + (int64_t)z * op.zstride + (int64_t)rb * 256;
    const int64_t achunk = (int64_t)op.nmb * 256;
    if (tid < 16) {
        const int cv = vb * 16 + tid;
        src_s[tid] = cv < g.ncol ? g.in_idx[(int64_t)z * g.idx_zstride + cv] : -1;
    }
    __syncthreads();
    constexpr int UB = 16;   // loads issued before their stores: 2 x 16 per thread in flight
    const int total = nsteps * 64;
    for (int e0 = tid; e0 < total; e0 += 256 * UB) {
        T ra[UB], rx[UB];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int e = e0 + u * 256;
            if (e < total) {
                const int s = e >> 6, l = e & 63, k = 4 * s + (l >> 4), src = src_s[l & 15];
                ra[u] = A[(int64_t)(s >> 2) * achunk + (s & 3) * 64 + l];
                rx[u] = (src >= 0 && k < K) ? (T)g.X[(int64_t)src * g.S + k] : (T)0;
            }
        }
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int e = e0 + u * 256;
            if (e < total) {
                As[e] = ra[u];
                Xs[e] = rx[u];
            }
        }
    }
    __syncthreads();
    if (tid >= 64) return;   // wave 0 runs the chain (no barrier follows)
    const int lane = tid, n = lane & 15, kq = lane >> 4;
    acc_t acc = {0, 0, 0, 0};
    for (int s0 = 0; s0 < nsteps; s0 += 4) {   // nsteps = 4 nkc: the next 4 steps' operands read first
        T xa[4], aa[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            xa[u] = Xs[(s0 + u) * 64 + lane];
            aa[u] = As[(s0 + u) * 64 + lane];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc = Mfma<T>::op(xa[u], aa[u], acc);
    }
    double* Y = g.Y + (int64_t)z * g.y_zstride;
    const int32_t* out_idx = g.out_idx + (int64_t)z * g.idx_zstride;
    const int r = rb * 16 + n;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = vb * 16 + Mfma<T>::row(kq, q);
        if (c >= g.ncol || r >= M) continue;
        const int oi = out_idx[c];
        if (oi < 0) continue;
        const double sc = g.scale_depth ? ldexp(1.0, -g.scale_depth[c]) : 1.0;
        const double v = sc * (double)acc[q];
        double* y = &Y[(int64_t)oi * g.S + r];
        *y = g.accumulate ? *y + v : v;
    }
}

template <class T>
size_t small_lds(const LaneOp& op) { return (size_t)op.nkc * 4 * 64 * 2 * sizeof(T) + 16 * sizeof(int); }
constexpr size_t kMaxLds = 160 * 1024;

// Sum over a wave in a fixed order: lane-strided partial sums, then an xor butterfly after which
// every lane holds the same value.
__device__ __forceinline__ double wave_sum(double a) {
    for (int m = 32; m > 0; m >>= 1) a += __shfl_xor(a, m);
    return a;
}

// Uniform shift of cell c's equivalent density v[0..S) to the sum Q (the cell's charge).
__device__ __forceinline__ void set_monopole(double* __restrict__ v, int S, double Q) {
    double a = 0;
    for (int j = threadIdx.x; j < S; j += 64) a += v[j];
    const double shift = (Q - wave_sum(a)) / S;
    for (int j = threadIdx.x; j < S; j += 64) v[j] += shift;
}

// One wave per leaf: its charge (its bodies in tree order), recorded, and the monopole set to it.
__global__ __launch_bounds__(64) void k_monopole_leaves(double* __restrict__ qeq, int S,
                                                        const int32_t* __restrict__ leaf_cell, CellsDev cells,
                                                        const double* __restrict__ q, double* __restrict__ qcell) {
    const int c = leaf_cell[blockIdx.x];
    const int b0 = cells.boff[c], nb = cells.bnum[c];
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += 64) a += q[b0 + i];
    const double Q = wave_sum(a);
    if (threadIdx.x == 0) qcell[c] = Q;
    set_monopole(qeq + (int64_t)c * S, S, Q);
}

// One wave per parent column of a level (after its M2M): the children's charges summed in octant
// order, recorded, and the monopole set to it.
__global__ __launch_bounds__(64) void k_monopole_parents(double* __restrict__ qeq, int S,
                                                         const int32_t* __restrict__ parents, int ncol,
                                                         const int32_t* __restrict__ children,
                                                         double* __restrict__ qcell) {
    const int col = blockIdx.x, c = parents[col];
    if (c < 0) return;   // uniform over the wave
    double Q = 0;
    for (int o = 0; o < 8; o++) {
        const int ch = children[(int64_t)o * ncol + col];
        if (ch >= 0) Q += qcell[ch];
    }
    if (threadIdx.x == 0) qcell[c] = Q;
    set_monopole(qeq + (int64_t)c * S, S, Q);
}

template <class T, int MBW, int WAVES>
size_t tiled_lds() { return 2 * (size_t)MBW * 256 * sizeof(T); }

template <class T, int MBW, int WAVES>
hipError_t tiled_init() {   // opt-in above 64 KiB of LDS per workgroup
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_apply<T, MBW, WAVES>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)tiled_lds<T, MBW, WAVES>());
}

template <class T, int MBW, int WAVES>
hipError_t run_tiled(const ApplyCols& g, const LaneOp& op, hipStream_t st) {
    const int nmt = (op.nmb + MBW - 1) / MBW;
    const int nct = (g.ncol + 16 * WAVES - 1) / (16 * WAVES);
    const int ntiles = nmt * nct;
    dim3 grid(8 * ((ntiles + 7) / 8), g.nz);
    k_apply<T, MBW, WAVES><<<grid, 64 * WAVES, tiled_lds<T, MBW, WAVES>(), st>>>(g, op, nmt, ntiles);
    return hipGetLastError();
}

template <class T>
hipError_t run_apply(const ApplyCols& g, const LaneOp& op, int mbw, hipStream_t st) {
    if (g.ncol <= kSmallCols && mbw <= 0) {
        dim3 grid(op.nmb, (g.ncol + 15) / 16, g.nz);
        const size_t lds = small_lds<T>(op);
        if (lds <= kMaxLds)
            k_apply_small_lds<T><<<grid, 256, lds, st>>>(g, op);
        else   // operands of more than 40 K chunks (float64 at P >= 11): straight from L2
            k_apply_small<T><<<grid, 64, 0, st>>>(g, op);
        return hipGetLastError();
    }
    if (mbw <= 0) {
        // Row tiles measured on gfx950 at P = 8 (32,768 columns, or 8 x 4,096): float64 8 blocks
        // 41 TFLOP/s (12: 35, 20: 31, 4: 14); float32 12 blocks 53-62 TFLOP/s (20: 48-56, 4: 24, 8:
        // under 6).  A 12-block tile when 8 would leave a ragged tile of at most 4 blocks.
        if (sizeof(T) == 8)
            mbw = (op.nmb > 8 && op.nmb <= 12) ? 12 : 8;
        else
            mbw = 12;
    }
    switch (mbw) {
        case 20: return run_tiled<T, 20, 4>(g, op, st);
        case 12: return run_tiled<T, 12, 4>(g, op, st);
        case 8: return run_tiled<T, 8, 4>(g, op, st);
        case 4: return run_tiled<T, 4, 4>(g, op, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_apply(const ApplyCols& g, const LaneOp& op, hipStream_t st, int mbw) {
    if (g.ncol <= 0) return hipSuccess;
    if (!op.p || !g.X || !g.Y || !g.in_idx || !g.out_idx || g.S <= 0 || g.nz <= 0) return hipErrorInvalidValue;
    if (op.M <= 0 || op.M > g.S || op.K <= 0 || op.K > g.S || op.nmb != (op.M + 15) / 16 || op.nkc != (op.K + 15) / 16)
        return hipErrorInvalidValue;
    if (g.nz > op.nops || op.zstride < (int64_t)op.nmb * op.nkc * 256) return hipErrorInvalidValue;
    return op.f32 ? run_apply<float>(g, op, mbw, st) : run_apply<double>(g, op, mbw, st);
}

hipError_t apply_init() {
    hipError_t e = hipSuccess;
    for (const void* fn : {reinterpret_cast<const void*>(k_apply_small_lds<double>),
                           reinterpret_cast<const void*>(k_apply_small_lds<float>)}) {
        hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
        if (r != hipSuccess) e = r;
    }
    for (hipError_t r : {tiled_init<double, 20, 4>(), tiled_init<double, 12, 4>(), tiled_init<double, 8, 4>(),
                         tiled_init<double, 4, 4>(), tiled_init<float, 20, 4>(), tiled_init<float, 12, 4>(),
                         tiled_init<float, 8, 4>(), tiled_init<float, 4, 4>()})
        if (r != hipSuccess) e = r;
    return e;
}

hipError_t launch_monopole_leaves(double* qeq, int S, const int32_t* leaf_cell, int nleaf, CellsDev cells,
                                  const double* q, double* qcell, hipStream_t st) {
    if (nleaf <= 0) return hipSuccess;
    if (!qeq || !leaf_cell || !q || !qcell || !valid(cells) || S <= 0) return hipErrorInvalidValue;
    k_monopole_leaves<<<nleaf, 64, 0, st>>>(qeq, S, leaf_cell, cells, q, qcell);
    return hipGetLastError();
}

hipError_t launch_monopole_parents(double* qeq, int S, const int32_t* parents, int ncol, const int32_t* children,
                                   double* qcell, hipStream_t st) {
    if (ncol <= 0) return hipSuccess;
    if (!qeq || !parents || !children || !qcell || S <= 0) return hipErrorInvalidValue;
    k_monopole_parents<<<ncol, 64, 0, st>>>(qeq, S, parents, ncol, children, qcell);
    return hipGetLastError();
}

template <class T>
void lane_layout(const double* A, int lda, int M, int K, std::vector<T>& out, size_t at) {
    const int nmb = (M + 15) / 16, nkc = (K + 15) / 16;
    for (int kc = 0; kc < nkc; kc++)
        for (int b = 0; b < nmb; b++)
            for (int u = 0; u < 4; u++)
                for (int lane = 0; lane < 64; lane++) {
                    const int r = 16 * b + (lane & 15), k = 16 * kc + 4 * u + (lane >> 4);
                    out[at + (((size_t)kc * nmb + b) * 4 + u) * 64 + lane] =
                        (r < M && k < K) ? (T)A[(size_t)r * lda + k] : (T)0;
                }
}
template void lane_layout<double>(const double*, int, int, int, std::vector<double>&, size_t);
template void lane_layout<float>(const double*, int, int, int, std::vector<float>&, size_t);

}  // namespace fmm3dpp
