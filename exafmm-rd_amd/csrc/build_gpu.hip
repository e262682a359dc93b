// GPU construction of the adaptive octree and the interaction lists, bit-exact with the
// reference's Tree::build_nonuniform_octree (3dpp/src/tree.cpp:102-188) and Traverser
// (3dpp/src/traverser.cpp:19-331).
//
//  1. every body descends 21 levels with the reference's float64 octant test (x > centre,
//     child centre = centre +/- r/2) -> 63-bit key; radix sort by key
//  2. top-down, one kernel pair per level: a cell with more than ncrit bodies splits, its octant
//     ranges are binary-searched in the sorted keys, non-empty children are appended in octant
//     order (BFS numbering: level-major, parent order, octant order)
//  3. radix sort, stable from input order, by the key truncated at each body's leaf depth: bodies
//     of a leaf keep input order, as the reference's stable partitions leave them
//  4. P2P map by a level-synchronous frontier over (target, source, target parent, source parent,
//     periodic offset) with the reference's decision rules; rows expanded to leaves, sorted into CSR
//  5. M2L parent map: the 26 same-depth neighbours of every non-leaf cell (periodic wrap when
//     images >= 1) by binary search in the level's sorted Morton prefixes
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "gpu_structure.h"

namespace fmm3dpp {

namespace {

constexpr int kD = 21;

#define CK(call)                                                                  \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            err = std::string(#call) + ": " + hipGetErrorString(e_);              \
            return FMM3DPP_ERR_HIP;                                               \
        }                                                                         \
    } while (0)

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// Body order inside each leaf = input order (the reference's stable partition keeps it): sort each
// leaf's slice of the key-sorted permutation ascending.  One workgroup per cell (non-leaves exit),
// bitonic sort in LDS; leaves hold <= ncrit <= kLeafSortMax bodies.
constexpr int kLeafSortMax = 1024;
__global__ __launch_bounds__(256) void k_leaf_sort(int ntree, const int32_t* __restrict__ cnum,
                                                   const int32_t* __restrict__ boff, const int32_t* __restrict__ bnum,
                                                   const int32_t* __restrict__ perm1, int32_t* __restrict__ perm) {
    __shared__ int32_t v[kLeafSortMax];
    const int c = blockIdx.x;
    if (cnum[c] != 0) return;                  // uniform over the workgroup
    const int b = boff[c], m = bnum[c];
    if (m > kLeafSortMax) return;              // leaves hold <= ncrit <= kLeafSortMax bodies (checked on the host)
    int w = 1;
    while (w < m) w <<= 1;
    for (int i = threadIdx.x; i < w; i += blockDim.x) v[i] = i < m ? perm1[b + i] : INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= w; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < w; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const int a = v[i], bb = v[l];
                    if ((a > bb) == ((i & k) == 0)) {
                        v[i] = bb;
                        v[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < m; i += blockDim.x) perm[b + i] = v[i];
}

__global__ void k_keys(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, double c0x, double c0y, double c0z, double r0,
                       uint64_t* __restrict__ key, int32_t* __restrict__ idx, int* __restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(flag, 1);
    double cx = c0x, cy = c0y, cz = c0z, r = r0;
    uint64_t k = 0;
    for (int l = 0; l < kD; l++) {
        int oct = ((px > cx) << 2) | ((py > cy) << 1) | (pz > cz);
        k = (k << 3) | (uint64_t)oct;
        double h = r / 2;
        cx = cx + ((oct & 4) ? h : -h);
        cy = cy + ((oct & 2) ? h : -h);
        cz = cz + ((oct & 1) ? h : -h);
        r = r / 2;
    }
    key[i] = k;
    idx[i] = (int32_t)i;
}

// Octant sub-ranges of every splitting cell of level l (cells [ls, le)).
// Split of the cells of level l (tree.cpp:147-186): 8 lanes per cell, lane o finds the end of
// octant o in the cell's key-sorted bodies by its own binary search on digit l, so the search depth
// is log2(bodies) rather than 8 x that.  cstart[9 * lc + o] = first body of octant o, [8] = end.
__global__ void k_split(int ls, int le, int l, int ncrit, int dmax, const uint64_t* __restrict__ K1,
                        const int32_t* __restrict__ boff, const int32_t* __restrict__ bnum, int32_t* __restrict__ nchild,
                        int32_t* __restrict__ cstart, int* __restrict__ flag, int32_t* __restrict__ lvl_nl) {
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int o = (int)(gt & 7);
    const int64_t c64 = ls + (gt >> 3);
    const bool valid = c64 < le;
    const int c = valid ? (int)c64 : ls;   // lanes past the level keep the shuffles below well defined
    const int lc = c - ls, b = boff[c], m = bnum[c];
    const bool split = valid && m > ncrit && l < dmax;   // keys are sorted on their top dmax levels
    if (valid && m > ncrit && l >= dmax && o == 0) atomicOr(flag, dmax < kD ? 4 : 2);
    int ub = b + m;
    if (split) {
        const int shift = 3 * (kD - l - 1);
        int lo = b, hi = b + m;   // first body whose digit exceeds o
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)((K1[mid] >> shift) & 7) <= o) lo = mid + 1; else hi = mid;
        }
        ub = lo;
    }
    // start of octant o = end of octant o - 1 (the cell's first body for o = 0)
    int prev = __shfl_up(ub, 1, 8);
    if (o == 0) prev = b;
    int nonempty = split && ub > prev ? 1 : 0;
    for (int d = 1; d < 8; d <<= 1) nonempty += __shfl_xor(nonempty, d, 8);
    if (!valid) return;
    if (split) {
        cstart[lc * 9 + o] = prev;
        if (o == 7) cstart[lc * 9 + 8] = b + m;
    }
    if (o == 0) nchild[lc] = split ? nonempty : 0;
    if (o == 0 && split) atomicAdd(&lvl_nl[l], 1);   // non-leaf cells of level l (a count: order-free)
}

__global__ void k_emit_children(int ls, int le, const int32_t* __restrict__ nchild, const int32_t* __restrict__ incl,
                                const int32_t* __restrict__ cstart, uint64_t* __restrict__ prefix,
                                double* __restrict__ cx, double* __restrict__ cy, double* __restrict__ cz,
                                double* __restrict__ cr, int32_t* __restrict__ depth, int32_t* __restrict__ octant,
                                int32_t* __restrict__ coff, int32_t* __restrict__ cnum, int32_t* __restrict__ boff,
                                int32_t* __restrict__ bnum) {
    int c = ls + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= le) return;
    int lc = c - ls, nc = nchild[lc];
    if (nc == 0) { coff[c] = 0; cnum[c] = 0; return; }
    int base = le + incl[lc] - nc;
    coff[c] = base;
    cnum[c] = nc;
    const double r = cr[c], x0 = cx[c], y0 = cy[c], z0 = cz[c];
    const uint64_t p = prefix[c];
    const int d = depth[c];
    int k = 0;
    for (int o = 0; o < 8; o++) {
        int s = cstart[lc * 9 + o], e = cstart[lc * 9 + o + 1];
        if (e <= s) continue;
        int ch = base + k++;
        double h = r / 2;
        prefix[ch] = (p << 3) | (uint64_t)o;
        boff[ch] = s;
        bnum[ch] = e - s;
        depth[ch] = d + 1;
        octant[ch] = o;
        cr[ch] = r / 2;
        cx[ch] = x0 + ((o & 4) ? h : -h);
        cy[ch] = y0 + ((o & 2) ? h : -h);
        cz[ch] = z0 + ((o & 1) ? h : -h);
        coff[ch] = 0;
        cnum[ch] = 0;
    }
}

__global__ void k_leaf_depth(int ntree, const int32_t* __restrict__ cnum, const int32_t* __restrict__ boff,
                             const int32_t* __restrict__ bnum, const int32_t* __restrict__ depth,
                             const int32_t* __restrict__ perm1, int8_t* __restrict__ ld, int32_t* __restrict__ head) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ntree || cnum[c] != 0) return;
    int b = boff[c], m = bnum[c];
    for (int j = b; j < b + m; j++) ld[perm1[j]] = (int8_t)depth[c];
    head[b] = 1;
}

__global__ void k_tkey(int64_t n, const uint64_t* __restrict__ key, const int8_t* __restrict__ ld,
                       uint64_t* __restrict__ tkey, int32_t* __restrict__ idx) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int sh = 3 * (kD - ld[i]);
    tkey[i] = sh >= 64 ? 0 : ((key[i] >> sh) << sh);
    idx[i] = (int32_t)i;
}

__global__ void k_gather_pos(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ x,
                             const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ xs,
                             double* __restrict__ ys, double* __restrict__ zs) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int i = perm[j];
    xs[j] = x[i]; ys[j] = y[i]; zs[j] = z[i];
}

__global__ void k_leaf_list(int ntree, const int32_t* __restrict__ cnum, const int32_t* __restrict__ boff,
                            const int32_t* __restrict__ depth, const int32_t* __restrict__ rank_incl,
                            int32_t* __restrict__ leaf_cell, int32_t* __restrict__ leaf_depth) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ntree || cnum[c] != 0) return;
    int r = rank_incl[boff[c]] - 1;
    leaf_cell[r] = c;
    leaf_depth[r] = depth[c];
}

__global__ void k_iota(int n, int32_t* __restrict__ a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// Column tables of the per-level operator applies: M2M reads the 8 children of a parent column
// (octant-major) and writes the parent; L2L reads the parent and writes each child.
// Operator-apply tables of the tree levels, one column per non-leaf cell of the level (leaves have
// no children to gather or scatter).  info = [ls | tab_off | col_off | nls], each [0..maxd]: level
// starts, table offsets, column offsets, and the number of non-leaf cells before each level;
// nl_incl is the inclusive scan of the non-leaf flags.
__global__ void k_level_tables(int ntree, int maxd, const int32_t* __restrict__ info,
                               const int32_t* __restrict__ depth, const int32_t* __restrict__ coff,
                               const int32_t* __restrict__ cnum, const int32_t* __restrict__ octant,
                               const int32_t* __restrict__ nl_incl, int32_t* __restrict__ m2m_in,
                               int32_t* __restrict__ m2m_out, int32_t* __restrict__ l2l_in,
                               int32_t* __restrict__ l2l_out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ntree) return;
    const int nc = cnum[c], l = depth[c];
    if (nc == 0 || l >= maxd) return;
    const int32_t* tab_off = info + (maxd + 1);
    const int32_t* col_off = info + 2 * (maxd + 1);
    const int32_t* nls = info + 3 * (maxd + 1);
    const int col = nl_incl[c] - 1 - nls[l], ncol = nls[l + 1] - nls[l];
    const int64_t tab = tab_off[l];
    for (int o = 0; o < 8; o++) {
        m2m_in[tab + (int64_t)o * ncol + col] = -1;
        l2l_out[tab + (int64_t)o * ncol + col] = -1;
        l2l_in[tab + (int64_t)o * ncol + col] = c;
    }
    for (int k = 0; k < nc; k++) {
        const int ch = coff[c] + k, o = octant[ch];
        m2m_in[tab + (int64_t)o * ncol + col] = ch;
        l2l_out[tab + (int64_t)o * ncol + col] = ch;
    }
    m2m_out[col_off[l] + col] = c;
}

__device__ inline void deinterleave_d(uint64_t prefix, int depth, int64_t c[3]) {
    c[0] = c[1] = c[2] = 0;
    for (int l = 0; l < depth; l++) {
        int oct = (int)((prefix >> (3 * (depth - 1 - l))) & 7);
        c[0] = (c[0] << 1) | ((oct >> 2) & 1);
        c[1] = (c[1] << 1) | ((oct >> 1) & 1);
        c[2] = (c[2] << 1) | (oct & 1);
    }
}
__device__ inline uint64_t interleave_d(const int64_t c[3], int depth) {
    uint64_t p = 0;
    for (int l = depth - 1; l >= 0; l--) {
        int oct = (int)((((c[0] >> l) & 1) << 2) | (((c[1] >> l) & 1) << 1) | ((c[2] >> l) & 1));
        p = (p << 3) | (uint64_t)oct;
    }
    return p;
}

__global__ void k_nonleaf_flag(int ntree, const int32_t* __restrict__ cnum, int32_t* __restrict__ flag) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ntree) flag[c] = cnum[c] > 0 ? 1 : 0;
}

// M2L parent map (traverser.cpp:234-282): for each non-leaf cell the same-depth non-leaf neighbour
// in each of the 26 directions (pattern order of get_relx(1,1): z outer, x inner), periodic wrap
// when images >= 1.  Also the child slots and the M2L down scale 2^(depth+1).
__global__ void k_m2l_tables(int ntree, int ntar, int periodic, const int32_t* __restrict__ info,
                             const uint64_t* __restrict__ prefix, const int32_t* __restrict__ depth,
                             const int32_t* __restrict__ coff, const int32_t* __restrict__ cnum,
                             const int32_t* __restrict__ octant, const int32_t* __restrict__ nl_incl,
                             int32_t* __restrict__ nl_cells, int32_t* __restrict__ child_slots,
                             double* __restrict__ tar_scale, int32_t* __restrict__ slot_src_T,
                             int32_t* __restrict__ img_src_T) {
    // one thread per (cell, item): items 0..25 are the neighbour patterns, item 26 the cell's own
    // entries (target index, child slots, scale)
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t c64 = gt / 27;
    const int p = (int)(gt - c64 * 27);
    if (c64 >= ntree) return;
    const int c = (int)c64;
    if (cnum[c] == 0) return;
    const int32_t* ls = info;
    const int t = nl_incl[c] - 1;
    const int d = depth[c];
    if (p == 26) {
        nl_cells[t] = c;
        for (int o = 0; o < 8; o++) child_slots[t * 8 + o] = -1;
        for (int k = 0; k < cnum[c]; k++) {
            const int ch = coff[c] + k;
            child_slots[t * 8 + octant[ch]] = ch;
        }
        tar_scale[t] = ldexp(1.0, d + 1);
        return;
    }
    int64_t a[3];
    deinterleave_d(prefix[c], d, a);
    const int64_t ns = int64_t(1) << d;
    const int lo0 = ls[d], hi0 = ls[d + 1];
    {
        int lin = p < 13 ? p : p + 1;
        int dl[3] = {lin % 3 - 1, (lin / 3) % 3 - 1, lin / 9 - 1};
        int64_t b[3];
        bool ok = true;
        for (int k = 0; k < 3; k++) {
            int64_t tt = a[k] + dl[k];
            if (periodic) b[k] = ((tt % ns) + ns) % ns;
            else { b[k] = tt; if (tt < 0 || tt >= ns) ok = false; }
        }
        int found = -1;
        if (ok) {
            uint64_t key = interleave_d(b, d);
            int lo = lo0, hi = hi0;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (prefix[mid] < key) lo = mid + 1; else hi = mid;
            }
            if (lo < hi0 && prefix[lo] == key && cnum[lo] > 0) found = nl_incl[lo] - 1;
        }
        slot_src_T[(int64_t)p * ntar + t] = found;
        img_src_T[(int64_t)p * ntar + t] = -1;
    }
}

// One frontier step of Traverser::horizontal_origin (traverser.cpp:70-148).
struct CellView {
    const double *x, *y, *z, *r;
    const int32_t *depth, *coff, *cnum;
};

__device__ inline bool adjacent_d(const CellView& C, int a, int b, double ox, double oy, double oz) {
    double dist = (C.r[a] + C.r[b]) * 1.001;
    return fabs(C.x[a] - C.x[b] - ox) <= dist && fabs(C.y[a] - C.y[b] - oy) <= dist &&
           fabs(C.z[a] - C.z[b] - oz) <= dist;
}

// The traversal's first items: (root, root) with parents `top` (the root, or the outermost image
// cell) under offset 13 (none) alone without images, else under each of the 27 offsets.
__global__ void k_frontier_init(int nitems, int top, int32_t* __restrict__ ftc, int32_t* __restrict__ fsc,
                                int32_t* __restrict__ ftcp, int32_t* __restrict__ fscp, int32_t* __restrict__ foi) {
    const int i = threadIdx.x;
    if (i >= nitems) return;
    ftc[i] = 0;
    fsc[i] = 0;
    ftcp[i] = top;
    fscp[i] = top;
    foi[i] = nitems == 1 ? 13 : i;
}

__global__ void k_frontier_classify(int64_t nitems, const int32_t* __restrict__ ftc, const int32_t* __restrict__ fsc,
                                    const int32_t* __restrict__ ftcp, const int32_t* __restrict__ fscp,
                                    const int32_t* __restrict__ foi, CellView C, double cycle,
                                    int64_t* __restrict__ packed, int8_t* __restrict__ act,
                                    unsigned long long* __restrict__ lost) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    const int tc = ftc[i], sc = fsc[i], tcp = ftcp[i], scp = fscp[i], oi = foi[i];
    const double ox = (oi % 3 - 1) * cycle, oy = ((oi / 3) % 3 - 1) * cycle, oz = (oi / 9 - 1) * cycle;
    const bool lt = C.cnum[tc] == 0, lsrc = C.cnum[sc] == 0;
    int divide = 0, emit = 0;
    if (!adjacent_d(C, tc, sc, ox, oy, oz)) {
        if (adjacent_d(C, tc, scp, ox, oy, oz) && lt && C.depth[scp] >= C.depth[tc]) emit = 1;
        else if (adjacent_d(C, tcp, sc, ox, oy, oz) && lsrc && C.depth[tcp] >= C.depth[sc]) emit = 1;
        else if (adjacent_d(C, tcp, scp, ox, oy, oz) && C.depth[tcp] == C.depth[scp]) { /* M2L pair */ }
        else if (lt && lsrc) atomicAdd(lost, 1ull);
        else if (!lt && lsrc) divide = 1;
        else if (lt && !lsrc) divide = 2;
        else divide = C.r[tc] >= C.r[sc] ? 1 : 2;
    } else {
        if (lt && lsrc) emit = 1;
        else if (!lt && lsrc) divide = 1;
        else if (lt && !lsrc) divide = 2;
        else divide = C.r[tc] >= C.r[sc] ? 1 : 2;
    }
    const int nch = divide == 1 ? C.cnum[tc] : (divide == 2 ? C.cnum[sc] : 0);
    packed[i] = ((int64_t)emit << 32) | (int64_t)nch;   // rows in the high word, child items in the low word
    act[i] = (int8_t)divide;
}

__global__ void k_frontier_emit(int64_t nitems, const int32_t* __restrict__ ftc, const int32_t* __restrict__ fsc,
                                const int32_t* __restrict__ ftcp, const int32_t* __restrict__ fscp,
                                const int32_t* __restrict__ foi, const int64_t* __restrict__ packed,
                                const int8_t* __restrict__ act, const int64_t* __restrict__ incl,
                                const int32_t* __restrict__ coff, int64_t row_base, int32_t* __restrict__ rtc,
                                int32_t* __restrict__ rsc, int32_t* __restrict__ roi, int32_t* __restrict__ ntc,
                                int32_t* __restrict__ nsc, int32_t* __restrict__ ntcp, int32_t* __restrict__ nscp,
                                int32_t* __restrict__ noi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    const int64_t pk = packed[i], in = incl[i];
    const int emit = (int)(pk >> 32), nch = (int)(pk & 0xffffffffll);
    const int tc = ftc[i], sc = fsc[i], tcp = ftcp[i], scp = fscp[i], oi = foi[i];
    if (emit) {
        int64_t r = row_base + (in >> 32) - 1;
        rtc[r] = tc; rsc[r] = sc; roi[r] = oi;
    }
    if (nch > 0) {
        int64_t b = (in & 0xffffffffll) - nch;
        if (act[i] == 1) {
            for (int k = 0; k < nch; k++) { ntc[b + k] = coff[tc] + k; nsc[b + k] = sc; ntcp[b + k] = tc; nscp[b + k] = scp; noi[b + k] = oi; }
        } else {
            for (int k = 0; k < nch; k++) { ntc[b + k] = tc; nsc[b + k] = coff[sc] + k; ntcp[b + k] = tcp; nscp[b + k] = sc; noi[b + k] = oi; }
        }
    }
}

// P2P rows -> CSR entries (target leaf rank, source leaf cell << 5 | offset index): one entry per
// (target leaf, source leaf) under the row's two cells.  Adaptive trees have rows with non-leaf cells on either side (traverser.cpp:75-86);
// expanding both to leaves lets the near-field kernels treat every entry as a leaf pair.
__global__ void k_p2p_count(int64_t nrows, const int32_t* __restrict__ rtc, const int32_t* __restrict__ rsc,
                            const int32_t* __restrict__ boff, const int32_t* __restrict__ bnum,
                            const int32_t* __restrict__ rank_incl, int64_t* __restrict__ cnt) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int t = rtc[r], sc = rsc[r];
    const int64_t nt = rank_incl[boff[t] + bnum[t] - 1] - rank_incl[boff[t]] + 1;
    const int64_t ns = rank_incl[boff[sc] + bnum[sc] - 1] - rank_incl[boff[sc]] + 1;
    cnt[r] = nt * ns;
}

__global__ void k_p2p_expand(int64_t nrows, const int32_t* __restrict__ rtc, const int32_t* __restrict__ rsc,
                             const int32_t* __restrict__ roi, const int32_t* __restrict__ boff,
                             const int32_t* __restrict__ bnum, const int32_t* __restrict__ rank_incl,
                             const int32_t* __restrict__ leaf_cell, const int64_t* __restrict__ cnt,
                             const int64_t* __restrict__ incl, int32_t* __restrict__ ekey, int32_t* __restrict__ eval) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int sc = rsc[r];
    const int lo = rank_incl[boff[rtc[r]]] - 1;
    const int slo = rank_incl[boff[sc]] - 1, ns = rank_incl[boff[sc] + bnum[sc] - 1] - slo;
    const int64_t c = cnt[r], b = incl[r] - c;
    for (int64_t k = 0; k < c; k++) {
        const int tl = (int)(k / ns), sl = (int)(k % ns);
        ekey[b + k] = lo + tl;
        eval[b + k] = (leaf_cell[slo + sl] << 5) | roi[r];
    }
}

__global__ void k_csr_offsets(int nleaf, int64_t ne, const int32_t* __restrict__ keys, int64_t* __restrict__ off) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nleaf) return;
    int64_t lo = 0, hi = ne;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < r) lo = mid + 1; else hi = mid;
    }
    off[r] = lo;
}

// Root cell (tree.cpp:107-117): centre, half size r0, all bodies, octant 13.
__global__ void k_root_init(double c0x, double c0y, double c0z, double r0, int32_t n, uint64_t* __restrict__ prefix,
                            double* __restrict__ cx, double* __restrict__ cy, double* __restrict__ cz,
                            double* __restrict__ cr, int32_t* __restrict__ depth, int32_t* __restrict__ octant,
                            int32_t* __restrict__ boff, int32_t* __restrict__ bnum, int32_t* __restrict__ coff,
                            int32_t* __restrict__ cnum) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    prefix[0] = 0;
    cx[0] = c0x; cy[0] = c0y; cz[0] = c0z; cr[0] = r0;
    depth[0] = 0; octant[0] = 13; boff[0] = 0; bnum[0] = n; coff[0] = 0; cnum[0] = 0;
}

// dst[1] = a[idx] (the total of an inclusive scan next to the error flag dst[0], read back together)
__global__ void k_copy_last(const int32_t* __restrict__ a, int64_t idx, int32_t* __restrict__ dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) dst[1] = a[idx];
}

// M2L tables of the image targets t = nl + m - 1, m = 1..nimg (the parents c(-1..-(images-1))):
// one child (slot 0: the root for m = 1, else c(-(m-1))), scale sc[m-1], no 8x8 sources and the
// 26 self entries of the 1x27 lattice kernel (traverser.cpp:43-64).
__global__ void k_image_targets(int nl, int nimg, int ntree, int ntar, const double* __restrict__ sc,
                                int32_t* __restrict__ nl_cells, int32_t* __restrict__ child_slots,
                                double* __restrict__ tar_scale, int32_t* __restrict__ slot_src_T,
                                int32_t* __restrict__ img_src_T, int32_t* __restrict__ img_tar) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nimg) return;
    const int m = i + 1, t = nl + i;
    img_tar[i] = t;
    nl_cells[t] = ntree + m - 1;
    for (int o = 0; o < 8; o++) child_slots[(size_t)t * 8 + o] = -1;
    child_slots[(size_t)t * 8] = m == 1 ? 0 : ntree + m - 2;
    tar_scale[t] = sc[i];
    for (int p = 0; p < 26; p++) {
        slot_src_T[(size_t)p * ntar + t] = -1;
        img_src_T[(size_t)p * ntar + t] = t;
    }
}

__global__ void k_image_cells(int ntree, int images, int64_t n, double* __restrict__ cx, double* __restrict__ cy,
                              double* __restrict__ cz, double* __restrict__ cr, int32_t* __restrict__ depth,
                              int32_t* __restrict__ octant, int32_t* __restrict__ coff, int32_t* __restrict__ cnum,
                              int32_t* __restrict__ boff, int32_t* __restrict__ bnum, uint64_t* __restrict__ prefix) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int m = 1; m <= images; m++) {   // c(-1..-images), traverser.cpp:150-162, 186-193
        int c = ntree + m - 1, child = m == 1 ? 0 : c - 1;
        cx[c] = cx[child]; cy[c] = cy[child]; cz[c] = cz[child];
        cr[c] = cr[child] * 3;
        depth[c] = -m; octant[c] = 13; coff[c] = child; cnum[c] = 1;
        boff[c] = 0; bnum[c] = m == 1 ? (int32_t)n : 0; prefix[c] = 0;
    }
}

// Inclusive scan of up to kSmallScan items in one workgroup: each thread sums a contiguous run, the
// run totals are scanned in LDS, then each run is rewritten with its offset.  Integer sums, so the
// result equals any other order's.
constexpr int kSmallScan = 4096;
template <class T>
__global__ __launch_bounds__(1024) void k_scan_small(const T* __restrict__ in, T* __restrict__ out, int n) {
    __shared__ T part[1024];
    const int tid = threadIdx.x, per = (n + 1023) / 1024;
    const int b = min(n, tid * per), e = min(n, b + per);
    T acc = 0;
    for (int i = b; i < e; i++) acc += in[i];
    part[tid] = acc;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // Hillis-Steele over the run totals
        const T v = tid >= d ? part[tid - d] : (T)0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    T run = tid > 0 ? part[tid - 1] : (T)0;
    for (int i = b; i < e; i++) {
        run += in[i];
        out[i] = run;
    }
}

template <class T>
hipError_t scan_incl(DBuf& temp, const T* in, T* out, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (n <= kSmallScan) {
        k_scan_small<T><<<1, 1024, 0, st>>>(in, out, (int)n);
        return hipGetLastError();
    }
    size_t bytes = 0;
    hipError_t e = rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, rocprim::plus<T>(), st);
    if (e != hipSuccess) return e;
    if ((e = temp.ensure(bytes)) != hipSuccess) return e;
    return rocprim::inclusive_scan(temp.p, bytes, in, out, (size_t)n, rocprim::plus<T>(), st);
}

// Stable LSD radix sort over bits [begin_bit, bits).  The onesweep algorithm is used at every size
// above one block (rocPRIM's default switches to a merge sort up to 2^20 items, measured ~2x
// slower here for the body keys and the list entries).
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;
template <class K, class V>
hipError_t sort_pairs(DBuf& temp, const K* kin, K* kout, const V* vin, V* vout, int64_t n, int bits, hipStream_t st,
                      int begin_bit = 0) {
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, kin, kout, vin, vout, (size_t)n, begin_bit,
                                                        bits, st);
    if (e != hipSuccess) return e;
    if ((e = temp.ensure(bytes)) != hipSuccess) return e;
    return rocprim::radix_sort_pairs<SortConfig>(temp.p, bytes, kin, kout, vin, vout, (size_t)n, begin_bit, bits, st);
}

// Host table -> device through the build's pinned staging region, in stream order (no host wait).
template <class T>
hipError_t upload_staged(GpuStructure::PinnedStage& ps, DBuf& d, const std::vector<T>& h, hipStream_t st) {
    const size_t bytes = h.size() * sizeof(T), span = (bytes + 255) & ~(size_t)255;
    hipError_t e = d.ensure(bytes);
    if (e != hipSuccess || h.empty()) return e;
    if (ps.off + span > ps.cap) {
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;   // copies in flight read the old region
        if (ps.p) (void)hipHostFree(ps.p);
        ps.p = nullptr;
        ps.cap = std::max(2 * (ps.off + span), (size_t)1 << 16);
        if ((e = hipHostMalloc((void**)&ps.p, ps.cap)) != hipSuccess) { ps.cap = 0; return e; }
        ps.off = 0;
    }
    std::memcpy(ps.p + ps.off, h.data(), bytes);
    e = hipMemcpyAsync(d.p, ps.p + ps.off, bytes, hipMemcpyHostToDevice, st);
    ps.off += span;
    return e;
}

template <class T>
hipError_t read_back(const T* d, T* h, size_t cnt, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(h, d, cnt * sizeof(T), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

__global__ void k_publish(const uint32_t* __restrict__ src, int nwords, uint32_t* dst, uint32_t seq) {
    const int i = threadIdx.x;
    if (i < nwords) dst[i] = src[i];
    __threadfence_system();
    __syncthreads();
    if (i == 0) {
        dst[GpuStructure::Mailbox::kWords] = seq;
        __threadfence_system();
    }
}

// A few values the host needs before it can queue more work: published to mapped pinned memory by a
// kernel queued behind everything before it on st, the host polling for the sequence number (no
// copy engine and no stream synchronisation).  Falls back to read_back when the mailbox is missing
// or too small; after 5 s without the sequence number the stream is synchronised and its error, if
// any, returned.
template <class T>
hipError_t small_read(GpuStructure::Mailbox& mb, const T* d, T* h, size_t cnt, hipStream_t st) {
    const size_t bytes = cnt * sizeof(T);
    if (bytes % 4 != 0 || bytes / 4 > (size_t)GpuStructure::Mailbox::kWords) return read_back(d, h, cnt, st);
    hipError_t e;
    if (!mb.h) {
        e = hipHostMalloc((void**)&mb.h, 64 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) {
            mb.h = nullptr;
            return e;
        }
        if ((e = hipHostGetDevicePointer((void**)&mb.d, mb.h, 0)) != hipSuccess) return e;
        mb.h[GpuStructure::Mailbox::kWords] = mb.seq;
    }
    const uint32_t seq = ++mb.seq;
    k_publish<<<1, 64, 0, st>>>(reinterpret_cast<const uint32_t*>(d), (int)(bytes / 4), mb.d, seq);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    volatile uint32_t* flag = mb.h + GpuStructure::Mailbox::kWords;
    const auto t0 = std::chrono::steady_clock::now();
    while (*flag != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5)) {
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            if (*flag != seq) return hipErrorUnknown;
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    std::memcpy(h, const_cast<const uint32_t*>(static_cast<volatile uint32_t*>(mb.h)), bytes);
    return hipSuccess;
}

int bits_for(int64_t v) {
    int b = 1;
    while ((int64_t(1) << b) <= v) b++;
    return b;
}

}  // namespace

int GpuStructure::build(const fmm3dpp_config& cfg, int64_t nb, const double* dx, const double* dy, const double* dz,
                        hipStream_t st, std::string& err) {
    int rc = build_tree(cfg, nb, dx, dy, dz, st, err);
    if (rc) return rc;
    return build_lists(cfg, st, err);
}

int GpuStructure::build_tree(const fmm3dpp_config& cfg, int64_t nb, const double* dx, const double* dy,
                             const double* dz, hipStream_t st, std::string& err) {
    t_build = std::chrono::steady_clock::now();
    // The pinned staging region is reused from its start: the previous build's copies out of it
    // were queued before its tree_done event (solve_device returns without waiting for them).
    if (tree_done.recorded) CK(hipEventSynchronize(tree_done.ev));
    stage.off = 0;
    if (nb < 1 || nb > (int64_t)INT32_MAX / 2) { err = "number of bodies must be in [1, 2^30)"; return FMM3DPP_ERR_ARG; }
    n = nb;
    const int images = cfg.images;
    const double r0 = cfg.cycle / 2;
    // --- 1. keys and sort
    CK(key.ensure(n * 8)); CK(key_s.ensure(n * 8)); CK(idx0.ensure(n * 4)); CK(perm1.ensure(n * 4));
    CK(small.ensure(64));
    CK(hipMemsetAsync(small.p, 0, 64, st));
    int* flag = small.as<int>();
    k_keys<<<blocks(n), 256, 0, st>>>(n, dx, dy, dz, cfg.center[0], cfg.center[1], cfg.center[2], r0,
                                       key.as<uint64_t>(), idx0.as<int32_t>(), flag);
    CK(hipGetLastError());
    // The keys are sorted on their top D levels only (fewer radix passes), D = the previous tree's
    // depth + 2: the tree needs key bits down to its own depth, and the per-leaf sort below restores
    // input order inside every leaf whatever order the bodies of one leaf have.  A cell that would
    // split below level D sets flag bit 4 and the build restarts with all kD levels sorted.
    auto grow_cells = [&](size_t need) -> hipError_t {
        if (need <= cell_cap) return hipSuccess;
        size_t cap = std::max(need + (size_t)images + 8, cell_cap * 2);
        size_t keep = cell_cap;
        hipError_t e;
        if ((e = prefix.grow_keep(cap * 8, keep * 8, st))) return e;
        if ((e = cx.grow_keep(cap * 8, keep * 8, st))) return e;
        if ((e = cy.grow_keep(cap * 8, keep * 8, st))) return e;
        if ((e = cz.grow_keep(cap * 8, keep * 8, st))) return e;
        if ((e = cr.grow_keep(cap * 8, keep * 8, st))) return e;
        if ((e = depth.grow_keep(cap * 4, keep * 4, st))) return e;
        if ((e = octant.grow_keep(cap * 4, keep * 4, st))) return e;
        if ((e = coff.grow_keep(cap * 4, keep * 4, st))) return e;
        if ((e = cnum.grow_keep(cap * 4, keep * 4, st))) return e;
        if ((e = boff.grow_keep(cap * 4, keep * 4, st))) return e;
        if ((e = bnum.grow_keep(cap * 4, keep * 4, st))) return e;
        cell_cap = cap;
        return hipSuccess;
    };
    int D = sort_levels > 0 && sort_levels < kD ? sort_levels : kD;
    int hflag = 0;
    for (bool retry = false;; retry = true) {
        if (retry) CK(hipMemsetAsync(small.p, 0, 64, st));   // bit 4 of the first attempt
        CK(sort_pairs(temp, key.as<uint64_t>(), key_s.as<uint64_t>(), idx0.as<int32_t>(), perm1.as<int32_t>(), n,
                      3 * kD, st, 3 * (kD - D)));
        // --- 2. top-down levels
        CK(grow_cells(std::max<size_t>(1024, (size_t)(n / std::max(1, cfg.ncrit)) * 3)));
        k_root_init<<<1, 1, 0, st>>>(cfg.center[0], cfg.center[1], cfg.center[2], r0, (int32_t)n, prefix.as<uint64_t>(),
                                     cx.as<double>(), cy.as<double>(), cz.as<double>(), cr.as<double>(),
                                     depth.as<int32_t>(), octant.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>(),
                                     coff.as<int32_t>(), cnum.as<int32_t>());
        CK(hipGetLastError());
        CK(lvl_nl.ensure((kD + 2) * 4));
        CK(hipMemsetAsync(lvl_nl.p, 0, (kD + 2) * 4, st));
        level_start.assign(1, 0);
        level_start.push_back(1);
        bool deeper = false;
        for (int l = 0;; l++) {
            int ls = level_start[l], le = level_start[l + 1], ncl = le - ls;
            CK(nchild.ensure((size_t)ncl * 4)); CK(scan32.ensure((size_t)ncl * 4)); CK(cstart.ensure((size_t)ncl * 9 * 4));
            k_split<<<blocks((int64_t)ncl * 8), 256, 0, st>>>(ls, le, l, cfg.ncrit, D, key_s.as<uint64_t>(),
                                                              boff.as<int32_t>(), bnum.as<int32_t>(), nchild.as<int32_t>(),
                                                              cstart.as<int32_t>(), flag, lvl_nl.as<int32_t>());
            CK(hipGetLastError());
            CK(scan_incl(temp, nchild.as<int32_t>(), scan32.as<int32_t>(), ncl, st));
            k_copy_last<<<1, 1, 0, st>>>(scan32.as<int32_t>(), ncl - 1, flag);
            CK(hipGetLastError());
            int32_t fl[2] = {0, 0};   // error flag, children total
            CK(small_read(mailbox, flag, fl, 2, st));
            hflag = fl[0];
            const int32_t total = fl[1];
            if (hflag & 1) { err = "non-finite body position"; return FMM3DPP_ERR_ARG; }
            if (hflag & 4) { deeper = true; break; }
            if (hflag & 2) { err = "more than ncrit bodies share a cell at depth 21 (coincident bodies?)"; return FMM3DPP_ERR_TREE; }
            CK(grow_cells((size_t)le + total));
            k_emit_children<<<blocks(ncl), 256, 0, st>>>(ls, le, nchild.as<int32_t>(), scan32.as<int32_t>(),
                                                         cstart.as<int32_t>(), prefix.as<uint64_t>(), cx.as<double>(),
                                                         cy.as<double>(), cz.as<double>(), cr.as<double>(),
                                                         depth.as<int32_t>(), octant.as<int32_t>(), coff.as<int32_t>(),
                                                         cnum.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>());
            CK(hipGetLastError());
            if (total == 0) break;
            level_start.push_back(le + total);
        }
        if (!deeper) break;
        D = kD;   // a deeper tree than the sorted levels: sort all levels and build again
    }
    ntree = level_start.back();
    maxd = (int)level_start.size() - 2;
    sort_levels = std::min(kD, maxd + 2);
    ncells = ntree + images;
    CK(grow_cells((size_t)ncells));
    // --- 3. leaf depth, stable re-sort, positions in tree order, leaves in body order
    CK(ldep.ensure(n)); CK(head.ensure(n * 4)); CK(rank.ensure(n * 4));
    CK(hipMemsetAsync(head.p, 0, n * 4, st));
    k_leaf_depth<<<blocks(ntree), 256, 0, st>>>(ntree, cnum.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>(),
                                                 depth.as<int32_t>(), perm1.as<int32_t>(), ldep.as<int8_t>(),
                                                 head.as<int32_t>());
    CK(hipGetLastError());
    CK(perm.ensure(n * 4));
    if (cfg.ncrit <= kLeafSortMax) {   // leaves hold <= ncrit bodies: sort each leaf's slice by input index
        int lsb = 64;   // one wave for leaves of up to 64 bodies, up to 4 for larger ncrit
        while (lsb < cfg.ncrit && lsb < 256) lsb *= 2;
        k_leaf_sort<<<ntree, lsb, 0, st>>>(ntree, cnum.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>(),
                                           perm1.as<int32_t>(), perm.as<int32_t>());
        CK(hipGetLastError());
    } else {                           // stable re-sort by the key truncated at the leaf depth
        CK(tkey.ensure(n * 8)); CK(tkey_s.ensure(n * 8));
        k_tkey<<<blocks(n), 256, 0, st>>>(n, key.as<uint64_t>(), ldep.as<int8_t>(), tkey.as<uint64_t>(),
                                          idx0.as<int32_t>());
        CK(hipGetLastError());
        CK(sort_pairs(temp, tkey.as<uint64_t>(), tkey_s.as<uint64_t>(), idx0.as<int32_t>(), perm.as<int32_t>(), n,
                      3 * kD, st));
    }
    CK(x.ensure(n * 8)); CK(y.ensure(n * 8)); CK(z.ensure(n * 8));
    k_gather_pos<<<blocks(n), 256, 0, st>>>(n, perm.as<int32_t>(), dx, dy, dz, x.as<double>(), y.as<double>(), z.as<double>());
    CK(hipGetLastError());
    CK(scan_incl(temp, head.as<int32_t>(), rank.as<int32_t>(), n, st));
    // one read-back for the leaf count (after lvl_nl[kD]) and the non-leaf cells per level
    k_copy_last<<<1, 1, 0, st>>>(rank.as<int32_t>(), n - 1, lvl_nl.as<int32_t>() + kD);   // -> lvl_nl[kD + 1]
    CK(hipGetLastError());
    std::vector<int32_t> lvl_count(kD + 2, 0);
    CK(small_read(mailbox, lvl_nl.as<int32_t>(), lvl_count.data(), (size_t)kD + 2, st));
    nleaf = lvl_count[kD + 1];
    CK(leaf_cell.ensure((size_t)nleaf * 4)); CK(leaf_depth.ensure((size_t)nleaf * 4));
    k_leaf_list<<<blocks(ntree), 256, 0, st>>>(ntree, cnum.as<int32_t>(), boff.as<int32_t>(), depth.as<int32_t>(),
                                                rank.as<int32_t>(), leaf_cell.as<int32_t>(), leaf_depth.as<int32_t>());
    CK(hipGetLastError());
    if (images >= 1) {
        k_image_cells<<<1, 1, 0, st>>>(ntree, images, n, cx.as<double>(), cy.as<double>(), cz.as<double>(),
                                       cr.as<double>(), depth.as<int32_t>(), octant.as<int32_t>(), coff.as<int32_t>(),
                                       cnum.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>(), prefix.as<uint64_t>());
        CK(hipGetLastError());
    }
    // --- 4. per-level operator-apply tables over the non-leaf cells of each level;
    //        lvl_info = [ls | tab_off | col_off | nls], each [0..maxd]
    CK(flags.ensure((size_t)ntree * 4)); CK(nl_idx.ensure((size_t)ntree * 4));
    k_nonleaf_flag<<<blocks(ntree), 256, 0, st>>>(ntree, cnum.as<int32_t>(), flags.as<int32_t>());
    CK(hipGetLastError());
    CK(scan_incl(temp, flags.as<int32_t>(), nl_idx.as<int32_t>(), ntree, st));
    std::vector<int32_t> info(4 * (maxd + 1), 0);
    m2m_levels.clear(); l2l_levels.clear();
    max_level_cols = 1;
    int64_t tab = 0, col = 0;
    int32_t nls = 0;
    for (int l = 0; l <= maxd; l++) {
        info[l] = level_start[l];
        info[3 * (maxd + 1) + l] = nls;
        nls += lvl_count[l];
    }
    for (int l = 0; l < maxd; l++) {
        const int ncol = lvl_count[l];
        info[(maxd + 1) + l] = (int32_t)tab;
        info[2 * (maxd + 1) + l] = (int32_t)col;
        l2l_levels.push_back({ncol, tab, col});
        max_level_cols = std::max(max_level_cols, ncol);
        tab += 8 * (int64_t)ncol;
        col += ncol;
    }
    for (int l = maxd - 1; l >= 0; l--) m2m_levels.push_back(l2l_levels[l]);
    CK(upload_staged(stage, lvl_info, info, st));
    CK(m2m_in.ensure(std::max<int64_t>(tab, 1) * 4)); CK(l2l_in.ensure(std::max<int64_t>(tab, 1) * 4));
    CK(l2l_out.ensure(std::max<int64_t>(tab, 1) * 4)); CK(m2m_out.ensure(std::max<int64_t>(col, 1) * 4));
    if (maxd > 0) {
        k_level_tables<<<blocks(ntree), 256, 0, st>>>(ntree, maxd, lvl_info.as<int32_t>(), depth.as<int32_t>(),
                                                       coff.as<int32_t>(), cnum.as<int32_t>(), octant.as<int32_t>(),
                                                       nl_idx.as<int32_t>(), m2m_in.as<int32_t>(),
                                                       m2m_out.as<int32_t>(), l2l_in.as<int32_t>(),
                                                       l2l_out.as<int32_t>());
        CK(hipGetLastError());
    }
    // iota covers the leaves and the M2M partial rows (8 per column of the widest level)
    int iota_n = std::max<int>(nleaf, 8 * max_level_cols);
    CK(iota.ensure((size_t)iota_n * 4));
    k_iota<<<blocks(iota_n), 256, 0, st>>>(iota_n, iota.as<int32_t>());
    CK(hipGetLastError());
    img_m2m.clear(); img_l2l.clear();
    for (int m = 1; m <= images - 1; m++) img_m2m.push_back({ntree + m - 1, m == 1 ? 0 : ntree + m - 2});
    for (int m = images; m >= 1; m--) img_l2l.push_back({ntree + m - 1, m == 1 ? 0 : ntree + m - 2});
    {
        std::vector<int32_t> ia, ib;
        for (auto& p : img_m2m) { ia.push_back(p[0]); ib.push_back(p[1]); }
        for (auto& p : img_l2l) { ia.push_back(p[0]); ib.push_back(p[1]); }
        CK(upload_staged(stage, img_a, ia, st)); CK(upload_staged(stage, img_b, ib, st));
    }
    // --- 5. M2L tables (targets = sources = the non-leaf cells, counted per level above)
    nl = nls;
    nimg = std::max(0, images - 1);
    ntar = nl + nimg;
    CK(nl_cells.ensure((size_t)std::max(ntar, 1) * 4)); CK(child_slots.ensure((size_t)std::max(ntar, 1) * 32));
    CK(tar_scale.ensure((size_t)std::max(ntar, 1) * 8));
    CK(slot_src_T.ensure((size_t)26 * std::max(ntar, 1) * 4)); CK(img_src_T.ensure((size_t)26 * std::max(ntar, 1) * 4));
    if (nl > 0) {
        k_m2l_tables<<<blocks((int64_t)ntree * 27), 256, 0, st>>>(ntree, ntar, images >= 1 ? 1 : 0, lvl_info.as<int32_t>(),
                                                     prefix.as<uint64_t>(), depth.as<int32_t>(), coff.as<int32_t>(),
                                                     cnum.as<int32_t>(), octant.as<int32_t>(), nl_idx.as<int32_t>(),
                                                     nl_cells.as<int32_t>(), child_slots.as<int32_t>(),
                                                     tar_scale.as<double>(), slot_src_T.as<int32_t>(),
                                                     img_src_T.as<int32_t>());
        CK(hipGetLastError());
    }
    if (nimg > 0) {   // image targets c(-1..-(images-1)): 26 self entries through the 1x27 lattice kernel
        std::vector<double> sc(nimg);
        for (int m = 1; m <= nimg; m++) sc[m - 1] = std::pow(-m >= -1 ? 2.0 : 3.0, -m + 1);   // kernel.cpp:1372
        CK(upload_staged(stage, img_scale, sc, st));
        CK(img_tar.ensure((size_t)nimg * 4));
        k_image_targets<<<blocks(nimg), 256, 0, st>>>(nl, nimg, ntree, ntar, img_scale.as<double>(),
                                                       nl_cells.as<int32_t>(), child_slots.as<int32_t>(),
                                                       tar_scale.as<double>(), slot_src_T.as<int32_t>(),
                                                       img_src_T.as<int32_t>(), img_tar.as<int32_t>());
        CK(hipGetLastError());
    } else {
        CK(img_tar.ensure(4));
    }
    // --- Hadamard staging tables over the targets of the 8x8 Hadamard
    const int ntar_h = cfg.precision == FMM3DPP_PREC_MIXED ? nl : ntar;
    const size_t nch = (size_t)std::max(had_chunks(ntar_h), 1);
    CK(had_uniq.ensure(nch * kHadUniqStride * 4)); CK(had_cnt.ensure(nch * 4));
    CK(had_loc.ensure(nch * 26 * kHadChunk * 4));
    CK(launch_hadamard_tables(ntar_h, ntar, slot_src_T.as<int32_t>(), had_uniq.as<int32_t>(), had_cnt.as<int32_t>(),
                              had_loc.as<int32_t>(), st));
    if (!tree_done.ev) CK(hipEventCreateWithFlags(&tree_done.ev, hipEventDisableTiming));
    CK(hipEventRecord(tree_done.ev, st));
    tree_done.recorded = true;
    return FMM3DPP_OK;
}

// Steps 6-7, the near-field lists; needs only the tree of build_tree, so it may run on another stream
// while the far-field stages of an evaluation run on the tree's stream.
int GpuStructure::build_lists(const fmm3dpp_config& cfg, hipStream_t st, std::string& err) {
    const int images = cfg.images;
    if (!tree_done.recorded) { err = "build_lists needs a tree"; return FMM3DPP_ERR_STATE; }
    CK(hipStreamWaitEvent(st, tree_done.ev, 0));   // the tree's last kernels may still be queued on its stream
    // --- 6. P2P map by frontier traversal, from (root, root) under each periodic offset
    const int top = images == 0 ? 0 : ntree;
    int cur = 0;
    int64_t nitems = images == 0 ? 1 : 27;
    auto ensure_frontier = [&](int b, int64_t cnt) -> hipError_t {
        if ((size_t)cnt <= f_cap[b]) return hipSuccess;
        size_t cap = std::max<size_t>(cnt, f_cap[b] * 2);
        hipError_t e;
        if ((e = f_tc[b].ensure(cap * 4))) return e;
        if ((e = f_sc[b].ensure(cap * 4))) return e;
        if ((e = f_tcp[b].ensure(cap * 4))) return e;
        if ((e = f_scp[b].ensure(cap * 4))) return e;
        if ((e = f_oi[b].ensure(cap * 4))) return e;
        f_cap[b] = cap;
        return hipSuccess;
    };
    CK(ensure_frontier(0, std::max<int64_t>(nitems, 1 << 16)));
    k_frontier_init<<<1, 32, 0, st>>>((int)nitems, top, f_tc[0].as<int32_t>(), f_sc[0].as<int32_t>(),
                                      f_tcp[0].as<int32_t>(), f_scp[0].as<int32_t>(), f_oi[0].as<int32_t>());
    CK(hipGetLastError());
    unsigned long long* d_lost = reinterpret_cast<unsigned long long*>(small.as<char>() + 16);
    n_rows = 0;
    const CellView cv{cx.as<double>(), cy.as<double>(), cz.as<double>(), cr.as<double>(), depth.as<int32_t>(),
                      coff.as<int32_t>(), cnum.as<int32_t>()};
    while (nitems > 0) {
        CK(packed.ensure(nitems * 8)); CK(scan64.ensure(nitems * 8)); CK(cnt64.ensure(nitems));  // cnt64 holds actions
        k_frontier_classify<<<blocks(nitems), 256, 0, st>>>(nitems, f_tc[cur].as<int32_t>(), f_sc[cur].as<int32_t>(),
                                                            f_tcp[cur].as<int32_t>(), f_scp[cur].as<int32_t>(),
                                                            f_oi[cur].as<int32_t>(), cv, cfg.cycle, packed.as<int64_t>(),
                                                            cnt64.as<int8_t>(), d_lost);
        CK(hipGetLastError());
        CK(scan_incl(temp_lists, packed.as<int64_t>(), scan64.as<int64_t>(), nitems, st));
        int64_t tot = 0;
        CK(small_read(mailbox, scan64.as<int64_t>() + (nitems - 1), &tot, 1, st));
        const int64_t new_rows = tot >> 32, new_items = tot & 0xffffffffll;
        if ((size_t)(n_rows + new_rows) > rows_cap) {
            size_t cap = std::max<size_t>((size_t)(n_rows + new_rows), rows_cap * 2);
            CK(rows_tc.grow_keep(cap * 4, (size_t)n_rows * 4, st));
            CK(rows_sc.grow_keep(cap * 4, (size_t)n_rows * 4, st));
            CK(rows_oi.grow_keep(cap * 4, (size_t)n_rows * 4, st));
            rows_cap = cap;
        }
        const int nxt = cur ^ 1;
        CK(ensure_frontier(nxt, std::max<int64_t>(new_items, 1)));
        k_frontier_emit<<<blocks(nitems), 256, 0, st>>>(nitems, f_tc[cur].as<int32_t>(), f_sc[cur].as<int32_t>(),
                                                        f_tcp[cur].as<int32_t>(), f_scp[cur].as<int32_t>(),
                                                        f_oi[cur].as<int32_t>(), packed.as<int64_t>(), cnt64.as<int8_t>(),
                                                        scan64.as<int64_t>(), coff.as<int32_t>(), n_rows,
                                                        rows_tc.as<int32_t>(), rows_sc.as<int32_t>(), rows_oi.as<int32_t>(),
                                                        f_tc[nxt].as<int32_t>(), f_sc[nxt].as<int32_t>(),
                                                        f_tcp[nxt].as<int32_t>(), f_scp[nxt].as<int32_t>(),
                                                        f_oi[nxt].as<int32_t>());
        CK(hipGetLastError());
        n_rows += new_rows;
        nitems = new_items;
        cur = nxt;
    }
    unsigned long long hl = 0;
    CK(small_read(mailbox, d_lost, &hl, 1, st));
    lost = (int64_t)hl;
    // --- 7. rows -> per-leaf CSR
    CK(p2p_off.ensure((size_t)(nleaf + 1) * 8));
    if (n_rows > 0) {
        CK(cnt64.ensure((size_t)n_rows * 8)); CK(scan64.ensure((size_t)n_rows * 8));
        k_p2p_count<<<blocks(n_rows), 256, 0, st>>>(n_rows, rows_tc.as<int32_t>(), rows_sc.as<int32_t>(),
                                                    boff.as<int32_t>(), bnum.as<int32_t>(), rank.as<int32_t>(),
                                                    cnt64.as<int64_t>());
        CK(hipGetLastError());
        CK(scan_incl(temp_lists, cnt64.as<int64_t>(), scan64.as<int64_t>(), n_rows, st));
        CK(small_read(mailbox, scan64.as<int64_t>() + (n_rows - 1), &n_entries, 1, st));
        CK(ekey.ensure((size_t)n_entries * 4)); CK(eval.ensure((size_t)n_entries * 4));
        CK(ekey_s.ensure((size_t)n_entries * 4)); CK(p2p_ent.ensure((size_t)n_entries * 4));
        k_p2p_expand<<<blocks(n_rows), 256, 0, st>>>(n_rows, rows_tc.as<int32_t>(), rows_sc.as<int32_t>(),
                                                     rows_oi.as<int32_t>(), boff.as<int32_t>(), bnum.as<int32_t>(),
                                                     rank.as<int32_t>(), leaf_cell.as<int32_t>(), cnt64.as<int64_t>(),
                                                     scan64.as<int64_t>(), ekey.as<int32_t>(), eval.as<int32_t>());
        CK(hipGetLastError());
        CK(sort_pairs(temp_lists, ekey.as<uint32_t>(), ekey_s.as<uint32_t>(), eval.as<int32_t>(), p2p_ent.as<int32_t>(),
                      n_entries, bits_for(nleaf), st));
    } else {
        n_entries = 0;
        CK(ekey_s.ensure(4)); CK(p2p_ent.ensure(4));
    }
    k_csr_offsets<<<blocks(nleaf + 1), 256, 0, st>>>(nleaf, n_entries, ekey_s.as<int32_t>(), p2p_off.as<int64_t>());
    CK(hipGetLastError());
    CK(hipStreamSynchronize(st));
    ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    return FMM3DPP_OK;
}

int GpuStructure::export_structure(const fmm3dpp_config& cfg, Structure& out, hipStream_t st, std::string& err) {
    std::vector<uint64_t> hp(ncells);
    std::vector<double> hx(ncells), hy(ncells), hz(ncells), hr(ncells);
    std::vector<int32_t> hd(ncells), ho(ncells), hco(ncells), hcn(ncells), hbo(ncells), hbn(ncells);
    CK(read_back(prefix.as<uint64_t>(), hp.data(), ncells, st));
    CK(read_back(cx.as<double>(), hx.data(), ncells, st));
    CK(read_back(cy.as<double>(), hy.data(), ncells, st));
    CK(read_back(cz.as<double>(), hz.data(), ncells, st));
    CK(read_back(cr.as<double>(), hr.data(), ncells, st));
    CK(read_back(depth.as<int32_t>(), hd.data(), ncells, st));
    CK(read_back(octant.as<int32_t>(), ho.data(), ncells, st));
    CK(read_back(coff.as<int32_t>(), hco.data(), ncells, st));
    CK(read_back(cnum.as<int32_t>(), hcn.data(), ncells, st));
    CK(read_back(boff.as<int32_t>(), hbo.data(), ncells, st));
    CK(read_back(bnum.as<int32_t>(), hbn.data(), ncells, st));
    out = Structure();
    out.n = n;
    out.ntree = ntree;
    out.max_depth = maxd;
    out.min_depth = cfg.images >= 1 ? -cfg.images : 0;
    out.lost = lost;
    out.cells.resize(ncells);
    for (int c = 0; c < ncells; c++) {
        CellRec& r = out.cells[c];
        r.octant = ho[c]; r.depth = hd[c]; r.coff = hco[c]; r.cnum = hcn[c]; r.boff = hbo[c]; r.bnum = hbn[c];
        r.r = hr[c]; r.x[0] = hx[c]; r.x[1] = hy[c]; r.x[2] = hz[c]; r.prefix = hp[c];
    }
    out.perm.resize(n);
    CK(read_back(perm.as<int32_t>(), out.perm.data(), n, st));
    std::vector<int32_t> tc(n_rows), sc(n_rows), oi(n_rows);
    if (n_rows) {
        CK(read_back(rows_tc.as<int32_t>(), tc.data(), n_rows, st));
        CK(read_back(rows_sc.as<int32_t>(), sc.data(), n_rows, st));
        CK(read_back(rows_oi.as<int32_t>(), oi.data(), n_rows, st));
    }
    out.p2p.resize(n_rows);
    for (int64_t r = 0; r < n_rows; r++) {
        int o = oi[r];
        out.p2p[r] = {tc[r], sc[r], {(o % 3 - 1) * cfg.cycle, ((o / 3) % 3 - 1) * cfg.cycle, (o / 9 - 1) * cfg.cycle}};
    }
    std::stable_sort(out.p2p.begin(), out.p2p.end(), [](const P2PRow& a, const P2PRow& b) { return a.tar < b.tar; });
    // M2L rows from the slot tables
    std::vector<int32_t> cellsl(std::max(ntar, 1)), slot((size_t)26 * std::max(ntar, 1)), img((size_t)26 * std::max(ntar, 1));
    if (ntar > 0) {
        CK(read_back(nl_cells.as<int32_t>(), cellsl.data(), ntar, st));
        CK(read_back(slot_src_T.as<int32_t>(), slot.data(), (size_t)26 * ntar, st));
        CK(read_back(img_src_T.as<int32_t>(), img.data(), (size_t)26 * ntar, st));
    }
    for (int t = 0; t < ntar; t++) {
        const int tcell = cellsl[t];
        const CellRec& T = out.cells[tcell];
        for (int p = 0; p < 26; p++) {
            int lin = p < 13 ? p : p + 1;
            int dl[3] = {lin % 3 - 1, (lin / 3) % 3 - 1, lin / 9 - 1};
            int s8 = slot[(size_t)p * ntar + t], si = img[(size_t)p * ntar + t];
            if (s8 >= 0) {
                const CellRec& Sc = out.cells[cellsl[s8]];
                double off[3];
                for (int k = 0; k < 3; k++) {   // source + off - target = delta * 2 r
                    double v = dl[k] * 2 * T.r - (Sc.x[k] - T.x[k]);
                    off[k] = std::round(v / cfg.cycle) * cfg.cycle;
                }
                out.m2l.push_back({tcell, cellsl[s8], {off[0], off[1], off[2]}, 0, p});
            }
            if (si >= 0) {
                double w = T.r * 2;
                out.m2l.push_back({tcell, cellsl[si], {dl[0] * w, dl[1] * w, dl[2] * w}, 1, p});
            }
        }
    }
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp
