// FFT-based M2L of the periodic KIFMM for CDNA4 (gfx950): up transform, 8x8 Hadamard product on
// the matrix cores, 1x27 image-lattice Hadamard, down transform.
//
// Data layout in HBM (T = double for the float64 path, float for the mixed path):
//   Qk, Pk   [parent][f][8 children] complex T, i.e. one 16-real vector per (parent, f): the
//            Hadamard reads and writes whole 128-byte (float64) or 64-byte (float32) vectors, and the
//            up/down transforms stream each parent's spectrum contiguously.  The Hadamard runs one
//            (frequency, chunk of targets) per workgroup with the chunks of one frequency on one
//            XCD, and copies each chunk's distinct source vectors into LDS once.
//   inv_lat  [k][j][i] -> surface point index s on the P^3 corner of the convolution grid, or -1
//   gk8      [26 patterns][f][is * 8 + it] complex T
//   gkA      [f][26 patterns][64 lanes] 4-vectors of T: gk8 as matrix-core A operands in lane order
// Each kernel cites the reference loop it replaces (paths under 3dpp/).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "device.h"

namespace fmm3dpp {

namespace {

template <typename T> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };
template <typename T> using cplx = typename Cplx<T>::type;

template <typename T>
__device__ __forceinline__ cplx<T> mkc(T re, T im) {
    cplx<T> v;
    v.x = re;
    v.y = im;
    return v;
}

// M2L up (src/kernel.cpp:1286-1328): scatter each child's S equivalent charges on the P^3 corner of
// the (2P-1)^3 grid and r2c-transform it as three pruned DFT passes.  One workgroup per source
// parent and group of SG child slots.  The x and y passes of FC half-spectrum columns fi run fused
// per (slot, k, fi) thread (P inputs -> P x-outputs in registers -> N y-outputs into LDS); the z
// pass reads P values per (slot, fj, fi) and writes its N outputs straight to Qk[src][f][slot], so
// consecutive lanes store consecutive complex values.
template <int P, typename T>
__global__ __launch_bounds__(256) void k_m2l_up(int S, int SG, int FC, const int32_t* __restrict__ src_child,
                                                const int32_t* __restrict__ inv_lat, const double* __restrict__ twiddle,
                                                const double* __restrict__ qeq, cplx<T>* __restrict__ Qk) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int N = 2 * P - 1, Nh = N / 2 + 1, F = N * N * Nh, P3 = P * P * P;
    const int ngrp = 8 / SG;
    const int src = blockIdx.x / ngrp, slot0 = (blockIdx.x % ngrp) * SG, tid = threadIdx.x;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(lds_raw);          // N
    cplx<T>* T2c = tw + N;                                       // [sl][k][fj][fil]  SG*P*N*FC
    T* L = reinterpret_cast<T*>(T2c + SG * P * N * FC);          // [sl][k][j][i]     SG*P^3
    for (int m = tid; m < N; m += 256) tw[m] = mkc<T>((T)twiddle[2 * m], (T)twiddle[2 * m + 1]);
    for (int e = tid; e < SG * P3; e += 256) {
        const int sl = e / P3, g = e - sl * P3;
        const int c = src_child[src * 8 + slot0 + sl], sp = inv_lat[g];
        L[e] = (c >= 0 && sp >= 0) ? (T)qeq[(size_t)c * S + sp] : (T)0;
    }
    for (int fi0 = 0; fi0 < Nh; fi0 += FC) {
        const int fc = min(FC, Nh - fi0);
        __syncthreads();
        for (int it = tid; it < SG * P * fc; it += 256) {   // x then y
            const int fil = it % fc, r = it / fc, k = r % P, sl = r / P, fi = fi0 + fil;
            const T* Lr = L + sl * P3 + k * P * P;
            cplx<T> t1[P];
#pragma unroll
            for (int j = 0; j < P; j++) {
                T re = 0, im = 0;
#pragma unroll
                for (int i = 0; i < P; i++) {
                    const T v = Lr[j * P + i];
                    const cplx<T> w = tw[(i * fi) % N];
                    re += v * w.x;
                    im += v * w.y;
                }
                t1[j] = mkc<T>(re, im);
            }
            cplx<T>* dst = T2c + (size_t)(sl * P + k) * N * fc + fil;
#pragma unroll
            for (int fj = 0; fj < N; fj++) {
                T re = 0, im = 0;
#pragma unroll
                for (int j = 0; j < P; j++) {
                    const cplx<T> w = tw[(j * fj) % N];
                    re += t1[j].x * w.x - t1[j].y * w.y;
                    im += t1[j].x * w.y + t1[j].y * w.x;
                }
                dst[fj * fc] = mkc<T>(re, im);
            }
        }
        __syncthreads();
        for (int it = tid; it < SG * fc * N; it += 256) {   // z, written as Qk[src][f][slot]
            const int sl = it % SG, r = it / SG, fil = r % fc, fj = r / fc, fi = fi0 + fil;
            const cplx<T>* sp = T2c + (size_t)(sl * P * N + fj) * fc + fil;
            cplx<T> v[P];
#pragma unroll
            for (int k = 0; k < P; k++) v[k] = sp[k * N * fc];
            cplx<T>* out = Qk + ((size_t)src * F + fj * Nh + fi) * 8 + slot0 + sl;
#pragma unroll
            for (int fk = 0; fk < N; fk++) {
                T re = 0, im = 0;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const cplx<T> w = tw[(k * fk) % N];
                    re += v[k].x * w.x - v[k].y * w.y;
                    im += v[k].x * w.y + v[k].y * w.x;
                }
                out[(size_t)fk * N * Nh * 8] = mkc<T>(re, im);
            }
        }
    }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename T> struct Vec4;
template <> struct Vec4<double> { using type = f64x4; };
template <> struct Vec4<float> { using type = f32x4; };
template <typename T> using vec4 = typename Vec4<T>::type;

// D(16x16) += A(16x4) B(4x16) on the matrix cores.  Lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; output register i holds D[4 (l / 16) + i][l % 16] for float32 and
// D[l / 16 + 4 i][l % 16] for float64 (both measured on gfx950).  d_row<T>(g, i) is that row.
template <typename T> __device__ __forceinline__ int d_row(int g, int i);
template <> __device__ __forceinline__ int d_row<float>(int g, int i) { return 4 * g + i; }
template <> __device__ __forceinline__ int d_row<double>(int g, int i) { return g + 4 * i; }
__device__ __forceinline__ f64x4 mfma16x16x4(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16x16x4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// A operand of lane l for pattern p at frequency f, laid out as the Hadamard reads it:
// gkA[(f * 26 + p) * 64 + l].  A complex 8x8 block G acts on interleaved (re, im) vectors as the real
// 16x16 matrix M[2it][2is] = M[2it+1][2is+1] = Re G[is][it], M[2it][2is+1] = -Im G, M[2it+1][2is] =
// Im G.  Lane (nl, kl) holds A row nl, K columns 4kl..4kl+3 in the permuted K order of the B operand
// (step k, lane group kl -> real 4kl + k); the rows are permuted so that output register i of lane
// (n, kr) is real 4kr + i of target n whatever the accumulator layout of the type (A row
// d_row(g, i) holds output real 4g + i).
template <typename T>
__global__ __launch_bounds__(256) void k_hadamard_lanes(int F, const T* __restrict__ gk8, vec4<T>* __restrict__ gkA) {
    const int f = blockIdx.x;
    for (int e = threadIdx.x; e < 26 * 64; e += 256) {
        const int p = e >> 6, l = e & 63, nl = l & 15, kl = l >> 4;
        int o = nl;   // output real o = 2 it + (re 0 / im 1) of A row nl
        for (int g = 0; g < 4; g++)
            for (int i = 0; i < 4; i++)
                if (d_row<T>(g, i) == nl) o = 4 * g + i;
        const int it = o >> 1, im = o & 1;
        const T* gp = gk8 + ((size_t)p * F + f) * 128;
        const T g0r = gp[((2 * kl) * 8 + it) * 2], g0i = gp[((2 * kl) * 8 + it) * 2 + 1];
        const T g1r = gp[((2 * kl + 1) * 8 + it) * 2], g1i = gp[((2 * kl + 1) * 8 + it) * 2 + 1];
        gkA[(size_t)f * 26 * 64 + e] = im ? vec4<T>{g0i, g0r, g1i, g1r} : vec4<T>{g0r, -g0i, g1r, -g1i};
    }
}

// Staging tables of the Hadamard, one workgroup per chunk c of kHadChunk consecutive targets: the
// distinct sources of all 26 patterns of the chunk (uniq[c * kHadUniqStride + r], r < cnt[c]) and,
// per (pattern, target), the row of its source among them (loc[(c * 26 + p) * kHadChunk + tt];
// cnt[c] when there is none).  The chunk's source indices are inserted into an open-addressing hash
// set in LDS (linear probing, more slots than entries, so every probe sequence ends), the occupied
// slots are ranked by a block scan, and each entry looks its source up again.  The row order
// depends on the insertion order; the Hadamard's results do not depend on the rows.
constexpr int kHadHash = 8192;   // hash slots, a power of two above 26 * kHadChunk
static_assert(26 * kHadChunk < kHadHash && 26 * kHadChunk < 32768, "hash slots, 16-bit ranks");
__global__ __launch_bounds__(1024) void k_hadamard_tables(int ntar, int ntab, const int32_t* __restrict__ slot_src_T,
                                                          int32_t* __restrict__ uniq, int32_t* __restrict__ cnt,
                                                          int32_t* __restrict__ loc) {
    __shared__ int32_t hkey[kHadHash];
    __shared__ int16_t hrank[kHadHash];   // ranks < 26 * kHadChunk < 2^15
    __shared__ int32_t part[1024];
    const int c = blockIdx.x, tid = threadIdx.x;
    auto source = [&](int i) {   // entry i = p * kHadChunk + tt of the chunk, -1 when none
        const int p = i / kHadChunk, t = c * kHadChunk + i % kHadChunk;
        return t < ntar ? slot_src_T[(size_t)p * ntab + t] : -1;
    };
    auto slot0 = [](int32_t s) { return (int)(((uint32_t)s * 2654435761u) >> 19) & (kHadHash - 1); };
    for (int i = tid; i < kHadHash; i += 1024) hkey[i] = -1;
    __syncthreads();
    for (int i = tid; i < 26 * kHadChunk; i += 1024) {
        const int32_t s = source(i);
        if (s < 0) continue;
        int h = slot0(s);
        for (int probe = 0; probe < kHadHash; probe++) {   // at most kHadHash - 1 other keys
            const int32_t old = atomicCAS(&hkey[h], -1, s);
            if (old == -1 || old == s) break;
            h = (h + 1) & (kHadHash - 1);
        }
    }
    __syncthreads();
    constexpr int PER = kHadHash / 1024;   // consecutive slots per thread
    int mine = 0;
    for (int e = 0; e < PER; e++) mine += hkey[tid * PER + e] >= 0;
    part[tid] = mine;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive scan of the per-thread counts
        const int v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    const int total = part[1023];
    int r = part[tid] - mine;
    int32_t* un = uniq + (size_t)c * kHadUniqStride;
    for (int e = 0; e < PER; e++) {
        const int h = tid * PER + e;
        if (hkey[h] >= 0) {
            hrank[h] = (int16_t)r;
            un[r] = hkey[h];
            r++;
        }
    }
    __syncthreads();
    for (int i = tid; i < 26 * kHadChunk; i += 1024) {
        const int32_t s = source(i);
        int l = total;
        if (s >= 0) {   // s is in the set
            int h = slot0(s);
            for (int probe = 0; probe < kHadHash && hkey[h] != s; probe++) h = (h + 1) & (kHadHash - 1);
            if (hkey[h] == s) l = hrank[h];   // always found; otherwise the zero row
        }
        loc[(size_t)c * 26 * kHadChunk + i] = l;
    }
    if (tid == 0) cnt[c] = total;
}

// Hadamard 8x8 (src/kernel.cpp:1012-1082, include/c2c.h:201-296) on the matrix cores.  Targets
// t < ntar, source table rows slot_src_T[p * ntab + t] (-1: no source; sources < ntab).  For one
// frequency f:  Pk[t][it] = sum over the 26 neighbour patterns p of sum_is G_p[is][it] Qk[src(t,p)][is].
// One workgroup per (frequency, chunk of kHadChunk targets): 8 waves of 2 tiles of 16 targets.  A
// wave multiplies the real 16x16 form of G_p (A operand, lane order of k_hadamard_lanes, staged in
// LDS per frequency) by 16 targets' source vectors (B operand, columns = targets) with four K=4
// steps; lane (n, kr) feeds reals 4kr..4kr+3 of target n's source vector.  The chunk's distinct
// source vectors at f (at most cap <= kHadCap, k_hadamard_tables) are first copied into LDS with one
// zero row for the missing sources, so the 26 x 256 source reads of the chunk come from LDS; a chunk
// with more distinct sources reads them from Qk instead (missing sources read row 0 and are zeroed in
// registers).  Both forms multiply the same values in the same order.  Source rows and table entries
// are read ahead of the matrix ops of the pattern before.  Blocks are remapped so that consecutive
// frequency slices run on one XCD.
template <typename T>
__global__ __launch_bounds__(512) void k_hadamard8_staged(int F, int ntar, int ntab, int nchunk, int total,
                                                          const int32_t* __restrict__ slot_src_T,
                                                          const int32_t* __restrict__ uniq,
                                                          const int32_t* __restrict__ cnt_tab,
                                                          const int32_t* __restrict__ loc,
                                                          const vec4<T>* __restrict__ gkA, const T* __restrict__ Qk,
                                                          T* __restrict__ Pk, int cap) {
    constexpr int TT = 2, NT = 512, CT = kHadChunk;
    static_assert(16 * TT * (NT / 64) == CT, "tiles cover the chunk");
    constexpr int MAXU = (kHadCap + 1) * 4 / NT;   // staged vec4 per thread
    static_assert(MAXU * NT == (kHadCap + 1) * 4, "staging covers the rows");
    extern __shared__ __align__(16) char smem[];
    vec4<T>* as = reinterpret_cast<vec4<T>*>(smem);   // 26 x 64 A operands
    vec4<T>* bs = as + 26 * 64;                        // staged source rows, 4 vec4 each
    const int per = (total + 7) / 8;
    const int Lb = (blockIdx.x % 8) * per + blockIdx.x / 8;   // consecutive logical blocks on one XCD
    if (Lb >= total) return;                                    // uniform over the workgroup
    const int f = Lb / nchunk, c = Lb % nchunk;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < 26 * 64; e += NT) as[e] = gkA[(size_t)f * 26 * 64 + e];
    // 32-bit element offsets: the launcher checks ntab * F * 4 < 2^32
    const vec4<T>* q4 = reinterpret_cast<const vec4<T>*>(Qk) + (size_t)f * 4;
    const uint32_t rowstride = (uint32_t)F * 4u;
    const int cnt = cnt_tab[c];
    const bool staged = cnt <= cap;   // uniform over the workgroup; cap <= kHadCap
    if (staged) {
        const int32_t* un = uniq + (size_t)c * kHadUniqStride;
        int src[MAXU];
        vec4<T> v[MAXU];
#pragma unroll
        for (int u = 0; u < MAXU; u++) {
            const int i = tid + NT * u;
            src[u] = (i >> 2) < cnt ? un[i >> 2] : -1;
        }
#pragma unroll
        for (int u = 0; u < MAXU; u++) {
            const int i = tid + NT * u;
            v[u] = src[u] >= 0 ? q4[(uint32_t)src[u] * rowstride + (i & 3)] : vec4<T>{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < MAXU; u++) {
            const int i = tid + NT * u;
            if ((i >> 2) <= cnt) bs[i] = v[u];   // rows 0..cnt, row cnt zero
        }
    }
    __syncthreads();
    const int n = lane & 15, kr = lane >> 4;
    const int tw = wave * 16 * TT;   // this wave's first target in the chunk
    vec4<T> acc[TT];
#pragma unroll
    for (int j = 0; j < TT; j++) acc[j] = vec4<T>{0, 0, 0, 0};
    if (staged) {
        const int32_t* lc = loc + (size_t)c * 26 * CT + tw + n;
        int l0[TT], l1[TT], l2[TT];
#pragma unroll
        for (int j = 0; j < TT; j++) {
            l0[j] = lc[16 * j];
            l1[j] = lc[CT + 16 * j];
        }
        vec4<T> a0 = as[lane], a1;
        vec4<T> b0[TT], b1[TT];
#pragma unroll
        for (int j = 0; j < TT; j++) b0[j] = bs[l0[j] * 4 + kr];
#pragma unroll   // the buffer rotation below becomes register renaming
        for (int p = 0; p < 26; p++) {
            __builtin_amdgcn_sched_barrier(0);
            if (p + 2 < 26) {
#pragma unroll
                for (int j = 0; j < TT; j++) l2[j] = lc[(p + 2) * CT + 16 * j];
            }
            if (p + 1 < 26) {
                a1 = as[(p + 1) * 64 + lane];
#pragma unroll
                for (int j = 0; j < TT; j++) b1[j] = bs[l1[j] * 4 + kr];
            }
            __builtin_amdgcn_sched_barrier(0);   // issue those reads before this pattern's matrix ops
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < TT; j++) acc[j] = mfma16x16x4(a0[k], b0[j][k], acc[j]);
            a0 = a1;
#pragma unroll
            for (int j = 0; j < TT; j++) {
                b0[j] = b1[j];
                l1[j] = l2[j];
            }
        }
    } else {
        const vec4<T>* qk = q4 + kr;
        int tcl[TT];
#pragma unroll
        for (int j = 0; j < TT; j++) tcl[j] = min(c * CT + tw + 16 * j + n, ntar - 1);   // lanes past ntar store nothing
        auto load_idx = [&](int p, int* sidx) {
#pragma unroll
            for (int j = 0; j < TT; j++) sidx[j] = slot_src_T[(size_t)p * ntab + tcl[j]];
        };
        auto load_vec = [&](const int* sidx, vec4<T>* b) {
#pragma unroll
            for (int j = 0; j < TT; j++) b[j] = qk[(uint32_t)max(sidx[j], 0) * rowstride];
        };
        vec4<T> bv[TT], bn[TT];
        int s0[TT], s1[TT], s2[TT];
        load_idx(0, s0);
        load_vec(s0, bv);
        load_idx(1, s1);
#pragma unroll
        for (int p = 0; p < 26; p++) {
            __builtin_amdgcn_sched_barrier(0);
            if (p + 1 < 26) load_vec(s1, bn);
            if (p + 2 < 26) load_idx(p + 2, s2);
            __builtin_amdgcn_sched_barrier(0);
            const vec4<T> a = as[p * 64 + lane];
#pragma unroll
            for (int j = 0; j < TT; j++)
                if (s0[j] < 0) bv[j] = vec4<T>{0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int j = 0; j < TT; j++) acc[j] = mfma16x16x4(a[k], bv[j][k], acc[j]);
#pragma unroll
            for (int j = 0; j < TT; j++) {
                bv[j] = bn[j];
                s0[j] = s1[j];
                s1[j] = s2[j];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < TT; j++) {
        const int t = c * CT + tw + 16 * j + n;
        if (t < ntar) *reinterpret_cast<vec4<T>*>(Pk + ((size_t)t * F + f) * 16 + 4 * kr) = acc[j];
    }
}

// Hadamard 1x27 (src/kernel.cpp:1084-1138): image-lattice M2L for the image parents c(-1..-(i-1)),
// always in float64.  Every one of the 27 lattice children of a neighbouring image is the same
// single child, so the 27 kernels of a pattern are pre-summed; only child slot 0 exists.
// Target/source indices are global (row stride ntab of img_src_T), minus `base` in Qk/Pk.  All 8
// slots of a target are written (slots 1..7 zero).
__global__ __launch_bounds__(256) void k_hadamard27(int F, int ntab, int base,
                                                    const int32_t* __restrict__ img_tar,
                                                    const int32_t* __restrict__ img_src_T, const double* __restrict__ g27,
                                                    const double2* __restrict__ Qk, double2* __restrict__ Pk) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int t = img_tar[blockIdx.y];
    double re = 0, im = 0;
    for (int p = 0; p < 26; p++) {
        const int s = img_src_T[(size_t)p * ntab + t];
        if (s < 0) continue;
        double2 q = Qk[((size_t)(s - base) * F + f) * 8];
        double gr = g27[((size_t)p * F + f) * 2], gi = g27[((size_t)p * F + f) * 2 + 1];
        re += gr * q.x - gi * q.y;
        im += gr * q.y + gi * q.x;
    }
    double2* o = Pk + ((size_t)(t - base) * F + f) * 8;
    o[0] = make_double2(re, im);
    for (int k = 1; k < 8; k++) o[k] = make_double2(0, 0);
}

// M2L down (src/kernel.cpp:1343-1385): c2r of each child's Pk evaluated only on the P^3 block that
// holds the shifted check surface, then gather, scale (2 or 3)^(depth+1) / N^3 and accumulate into
// p_check (float64).  One workgroup per target parent and group of SG child slots; the half
// spectrum is taken FC columns fi at a time: z pass (N values of Pk[t][f][slot] -> P outputs),
// y pass (N -> P), and the Hermitian x pass accumulates into per-thread registers, so only the two
// column chunks live in LDS.  blockIdx.x is remapped so that the slot groups of a parent run on one XCD
// at about the same time: each reads part of every 64-byte (float32) Pk vector, and the line one
// fetches from HBM is then still in that XCD's L2 for the others.
template <int P, typename T>
__global__ __launch_bounds__(256) void k_m2l_down(int S, int SG, int FC, int nblk, const int32_t* __restrict__ tar_child,
                                                  const double* __restrict__ tar_scale,
                                                  const int32_t* __restrict__ inv_lat,
                                                  const double* __restrict__ twiddle, const cplx<T>* __restrict__ Pk,
                                                  double* __restrict__ pck) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int N = 2 * P - 1, Nh = N / 2 + 1, N3 = N * N * N, F = N * N * Nh;
    constexpr int XI = (8 * P * P + 255) / 256;                  // x-pass items per thread (SG <= 8)
    const int ngrp = 8 / SG;
    const int per = (nblk + 7) / 8;
    const int L = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);   // consecutive on one XCD
    if (L >= nblk) return;   // whole workgroup, before any barrier
    const int t = L / ngrp, slot0 = (L % ngrp) * SG, tid = threadIdx.x;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(lds_raw);          // N
    cplx<T>* T1c = tw + N;                                       // [sl][kk][fj][fil]  SG*P*N*FC
    cplx<T>* T2c = T1c + SG * P * N * FC;                        // [sl][kk][jj][fil]  SG*P*P*FC
    for (int m = tid; m < N; m += 256) tw[m] = mkc<T>((T)twiddle[2 * m], (T)twiddle[2 * m + 1]);
    T acc[XI][P];
#pragma unroll
    for (int m = 0; m < XI; m++)
#pragma unroll
        for (int ii = 0; ii < P; ii++) acc[m][ii] = 0;
    const cplx<T>* pk = Pk + (size_t)t * F * 8 + slot0;
    for (int fi0 = 0; fi0 < Nh; fi0 += FC) {
        const int fc = min(FC, Nh - fi0);
        __syncthreads();
        for (int it = tid; it < SG * fc * N; it += 256) {   // z: k = P-1+kk from all N fk
            const int sl = it % SG, r = it / SG, fil = r % fc, fj = r / fc, fi = fi0 + fil;
            T re[P], im[P];
#pragma unroll
            for (int kk = 0; kk < P; kk++) re[kk] = im[kk] = 0;
#pragma unroll
            for (int fk = 0; fk < N; fk++) {
                const cplx<T> v = pk[(size_t)((fk * N + fj) * Nh + fi) * 8 + sl];
#pragma unroll
                for (int kk = 0; kk < P; kk++) {
                    const cplx<T> w = tw[((P - 1 + kk) * fk) % N];   // conj(w): exp(+2 pi i k fk / N)
                    re[kk] += v.x * w.x + v.y * w.y;
                    im[kk] += v.y * w.x - v.x * w.y;
                }
            }
            cplx<T>* dst = T1c + (size_t)(sl * P * N + fj) * fc + fil;
#pragma unroll
            for (int kk = 0; kk < P; kk++) dst[kk * N * fc] = mkc<T>(re[kk], im[kk]);
        }
        __syncthreads();
        for (int it = tid; it < SG * P * fc; it += 256) {   // y: j = P-1+jj
            const int fil = it % fc, r = it / fc;             // r = sl * P + kk
            const cplx<T>* s1 = T1c + (size_t)r * N * fc + fil;
            cplx<T> v[N];
#pragma unroll
            for (int fj = 0; fj < N; fj++) v[fj] = s1[fj * fc];
            cplx<T>* dst = T2c + (size_t)r * P * fc + fil;
#pragma unroll
            for (int jj = 0; jj < P; jj++) {
                T re = 0, im = 0;
#pragma unroll
                for (int fj = 0; fj < N; fj++) {
                    const cplx<T> w = tw[((P - 1 + jj) * fj) % N];
                    re += v[fj].x * w.x + v[fj].y * w.y;
                    im += v[fj].y * w.x - v[fj].x * w.y;
                }
                dst[jj * fc] = mkc<T>(re, im);
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < XI; m++) {                        // x: Hermitian half spectrum -> real
            const int it = tid + 256 * m;                     // it = (sl * P + kk) * P + jj
            if (it >= SG * P * P) continue;
            const cplx<T>* s2 = T2c + (size_t)it * fc;
            for (int fil = 0; fil < fc; fil++) {
                const int fi = fi0 + fil;
                const cplx<T> v = s2[fil];
#pragma unroll
                for (int ii = 0; ii < P; ii++) {
                    if (fi == 0) {
                        acc[m][ii] += v.x;
                    } else {
                        const cplx<T> w = tw[((P - 1 + ii) * fi) % N];
                        acc[m][ii] += 2 * (v.x * w.x + v.y * w.y);
                    }
                }
            }
        }
    }
    const double scale = tar_scale[t] / N3;
#pragma unroll
    for (int m = 0; m < XI; m++) {
        const int it = tid + 256 * m;
        if (it >= SG * P * P) continue;
        const int sl = it / (P * P), kj = it % (P * P);        // kj = kk * P + jj
        const int c = tar_child[t * 8 + slot0 + sl];
        if (c < 0) continue;
#pragma unroll
        for (int ii = 0; ii < P; ii++) {
            const int sp = inv_lat[kj * P + ii];
            if (sp >= 0) pck[(size_t)c * S + sp] += (double)acc[m][ii] * scale;
        }
    }
}

// LDS of the transforms for slot group SG and column chunk FC (element size esz).
size_t m2l_up_lds(int P, int SG, int FC, size_t esz) {
    const int N = 2 * P - 1;
    return esz * (2 * N + 2 * SG * P * N * FC + SG * P * P * P);
}
size_t m2l_down_lds(int P, int SG, int FC, size_t esz) {
    const int N = 2 * P - 1;
    return esz * (2 * N + 2 * SG * P * N * FC + 2 * SG * P * P * FC);
}
// Largest column chunk (then slot group) that leaves room for two workgroups per CU; one
// workgroup per CU if nothing smaller fits.
void pick_tiles(size_t (*lds)(int, int, int, size_t), int P, size_t esz, int& SG, int& FC) {
    const int Nh = P;   // N / 2 + 1 with N = 2P - 1
    for (size_t budget : {(size_t)80 * 1024, (size_t)160 * 1024})
        for (int sg = 8; sg >= 1; sg /= 2)
            for (int fc = Nh; fc >= 1; fc--)
                if (lds(P, sg, fc, esz) <= budget) { SG = sg; FC = fc; return; }
    SG = 1;
    FC = 1;
}

// Raises a kernel's dynamic LDS limit to 160 KiB once per device (the attribute is per device;
// `done` holds one bit per device id, safe to call from several host threads).
hipError_t allow_large_lds(const void* fn, std::atomic<uint32_t>& done) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint32_t bit = dev < 32 ? 1u << dev : 0u;
    if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess && bit) done.fetch_or(bit, std::memory_order_release);
    return e;
}

// Up transform tiles: all 8 child slots per workgroup and a column chunk of 2 to 4 columns that
// divides the half spectrum, within two workgroups' LDS per CU; else the largest chunk that fits.
// Measured at P=8 on 4681 parents: float32 SG=8, FC=4 0.265 ms against 0.303 ms for FC=8; float64
// SG=8, FC=2 0.454 ms (FC=3 0.446, FC=1 0.626).
void pick_tiles_up(int P, size_t esz, int& SG, int& FC) {
    const int Nh = P;
    for (int fc = std::min(4, Nh); fc >= 2; fc--)
        if (Nh % fc == 0 && m2l_up_lds(P, 8, fc, esz) <= (size_t)80 * 1024) { SG = 8; FC = fc; return; }
    pick_tiles(m2l_up_lds, P, esz, SG, FC);
}

template <int P, typename T>
hipError_t up_p(int S, int nsrc, const int32_t* src_child, const int32_t* inv_lat, const double* twiddle,
                const double* qeq, void* Qk, hipStream_t st, M2LTiles tiles) {
    int SG = tiles.sg, FC = tiles.fc;
    if (SG == 0 || FC == 0) pick_tiles_up(P, sizeof(T), SG, FC);
    const size_t lds = m2l_up_lds(P, SG, FC, sizeof(T));
    static std::atomic<uint32_t> attr{0};
    hipError_t e = allow_large_lds((const void*)k_m2l_up<P, T>, attr);
    if (e != hipSuccess) return e;
    k_m2l_up<P, T><<<nsrc * (8 / SG), 256, lds, st>>>(S, SG, FC, src_child, inv_lat, twiddle, qeq,
                                                       static_cast<cplx<T>*>(Qk));
    return hipGetLastError();
}

// Down transform tiles: 4 child slots per workgroup and a column chunk that divides the half
// spectrum (no ragged last chunk), at most 4 columns and 48 KiB of LDS.  Measured at P=8 on 4681
// parents (float32 / float64): SG=4, FC=4 takes 0.35 / 0.53 ms against 0.58 / 0.75 ms for the
// largest chunk that fits two workgroups per CU (SG=8, FC=6 / 3).
void pick_tiles_down(int P, size_t esz, int& SG, int& FC) {
    const int Nh = P;
    for (int sg : {4, 8, 2, 1})
        for (int fc = std::min(4, Nh); fc >= 1; fc--)
            if (Nh % fc == 0 && m2l_down_lds(P, sg, fc, esz) <= (size_t)48 * 1024) { SG = sg; FC = fc; return; }
    pick_tiles(m2l_down_lds, P, esz, SG, FC);
}

template <int P, typename T>
hipError_t down_p(int S, int ntar, const int32_t* tar_child, const double* tar_scale, const int32_t* inv_lat,
                  const double* twiddle, const void* Pk, double* pck, hipStream_t st, M2LTiles tiles) {
    int SG = tiles.sg, FC = tiles.fc;
    if (SG == 0 || FC == 0) pick_tiles_down(P, sizeof(T), SG, FC);
    const size_t lds = m2l_down_lds(P, SG, FC, sizeof(T));
    static std::atomic<uint32_t> attr{0};
    hipError_t e = allow_large_lds((const void*)k_m2l_down<P, T>, attr);
    if (e != hipSuccess) return e;
    const int nblk = ntar * (8 / SG);
    k_m2l_down<P, T><<<8 * ((nblk + 7) / 8), 256, lds, st>>>(S, SG, FC, nblk, tar_child, tar_scale, inv_lat, twiddle,
                                                              static_cast<const cplx<T>*>(Pk), pck);
    return hipGetLastError();
}

template <typename T>
size_t hadamard8_lds() { return (size_t)(26 * 64 + (kHadCap + 1) * 4) * sizeof(vec4<T>); }

template <typename T>
hipError_t hadamard8_t(int F, int ntar, int ntab, const int32_t* slot_src_T, HadTables tab, const void* gkA,
                       const void* Qk, void* Pk, int cap, hipStream_t st) {
    static std::atomic<uint32_t> attr{0};
    if ((uint64_t)ntab * (uint64_t)F * 4u >= (uint64_t(1) << 32)) return hipErrorInvalidValue;   // 32-bit offsets
    hipError_t e = allow_large_lds((const void*)k_hadamard8_staged<T>, attr);
    if (e != hipSuccess) return e;
    const int nchunk = had_chunks(ntar);
    const int total = nchunk * F;
    const int per = (total + 7) / 8;
    k_hadamard8_staged<T><<<per * 8, 512, hadamard8_lds<T>(), st>>>(F, ntar, ntab, nchunk, total, slot_src_T, tab.uniq,
                                                                     tab.cnt, tab.loc,
                                                                     static_cast<const vec4<T>*>(gkA),
                                                                     static_cast<const T*>(Qk), static_cast<T*>(Pk),
                                                                     std::min(cap, kHadCap));
    return hipGetLastError();
}

#define FMM3DPP_P_CASES(M) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12)

}  // namespace

bool m2l_tiles_valid(int P, bool f32, bool down, M2LTiles t) {
    if (t.sg == 0 && t.fc == 0) return true;
    if ((t.sg != 1 && t.sg != 2 && t.sg != 4 && t.sg != 8) || t.fc < 1 || t.fc > P) return false;
    const size_t esz = f32 ? sizeof(float) : sizeof(double);
    return (down ? m2l_down_lds(P, t.sg, t.fc, esz) : m2l_up_lds(P, t.sg, t.fc, esz)) <= (size_t)160 * 1024;
}

hipError_t launch_m2l_up(int P, bool f32, int S, int nsrc, const int32_t* src_child, const int32_t* inv_lat,
                         const double* twiddle, const double* qeq, void* Qk, hipStream_t st, M2LTiles tiles) {
    if (nsrc <= 0) return hipSuccess;
    if (!src_child || !inv_lat || !twiddle || !qeq || !Qk) return hipErrorInvalidValue;
    if (!m2l_tiles_valid(P, f32, false, tiles)) return hipErrorInvalidValue;
    switch (P) {
#define UPCASE(p)                                                                                  \
    case p:                                                                                        \
        return f32 ? up_p<p, float>(S, nsrc, src_child, inv_lat, twiddle, qeq, Qk, st, tiles)      \
                   : up_p<p, double>(S, nsrc, src_child, inv_lat, twiddle, qeq, Qk, st, tiles);
        FMM3DPP_P_CASES(UPCASE)
#undef UPCASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_hadamard8(bool f32, int F, int ntar, int ntab, const int32_t* slot_src_T, HadTables tab,
                            const void* gkA, const void* Qk, void* Pk, hipStream_t st, int cap) {
    if (ntar <= 0) return hipSuccess;
    if (!slot_src_T || !tab.uniq || !tab.cnt || !tab.loc || !gkA || !Qk || !Pk || ntab < ntar)
        return hipErrorInvalidValue;
    return f32 ? hadamard8_t<float>(F, ntar, ntab, slot_src_T, tab, gkA, Qk, Pk, cap, st)
               : hadamard8_t<double>(F, ntar, ntab, slot_src_T, tab, gkA, Qk, Pk, cap, st);
}

hipError_t launch_hadamard_tables(int ntar, int ntab, const int32_t* slot_src_T, int32_t* uniq, int32_t* cnt,
                                  int32_t* loc, hipStream_t st) {
    if (ntar <= 0) return hipSuccess;
    if (!slot_src_T || !uniq || !cnt || !loc || ntab < ntar) return hipErrorInvalidValue;
    k_hadamard_tables<<<had_chunks(ntar), 1024, 0, st>>>(ntar, ntab, slot_src_T, uniq, cnt, loc);
    return hipGetLastError();
}

hipError_t launch_hadamard_lanes(bool f32, int F, const void* gk8, void* gkA, hipStream_t st) {
    if (F <= 0 || !gk8 || !gkA) return hipErrorInvalidValue;
    if (f32)
        k_hadamard_lanes<float><<<F, 256, 0, st>>>(F, static_cast<const float*>(gk8), static_cast<f32x4*>(gkA));
    else
        k_hadamard_lanes<double><<<F, 256, 0, st>>>(F, static_cast<const double*>(gk8), static_cast<f64x4*>(gkA));
    return hipGetLastError();
}

hipError_t launch_hadamard27(int F, int nimg, const int32_t* img_tar, const int32_t* img_src_T, int ntab, int base,
                             const double* gk27_sum, const double* Qk, double* Pk, hipStream_t st) {
    if (nimg <= 0) return hipSuccess;
    if (!img_tar || !img_src_T || !gk27_sum || !Qk || !Pk) return hipErrorInvalidValue;
    k_hadamard27<<<dim3((F + 255) / 256, nimg), 256, 0, st>>>(F, ntab, base, img_tar, img_src_T, gk27_sum,
                                                              reinterpret_cast<const double2*>(Qk),
                                                              reinterpret_cast<double2*>(Pk));
    return hipGetLastError();
}

hipError_t launch_m2l_down(int P, bool f32, int S, int ntar, const int32_t* tar_child, const double* tar_scale,
                           const int32_t* inv_lat, const double* twiddle, const void* Pk, double* pck,
                           hipStream_t st, M2LTiles tiles) {
    if (ntar <= 0) return hipSuccess;
    if (!tar_child || !tar_scale || !inv_lat || !twiddle || !Pk || !pck) return hipErrorInvalidValue;
    if (!m2l_tiles_valid(P, f32, true, tiles)) return hipErrorInvalidValue;
    switch (P) {
#define DNCASE(p)                                                                                          \
    case p:                                                                                                \
        return f32 ? down_p<p, float>(S, ntar, tar_child, tar_scale, inv_lat, twiddle, Pk, pck, st, tiles) \
                   : down_p<p, double>(S, ntar, tar_child, tar_scale, inv_lat, twiddle, Pk, pck, st, tiles);
        FMM3DPP_P_CASES(DNCASE)
#undef DNCASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fmm3dpp
