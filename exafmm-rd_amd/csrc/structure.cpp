// Tree and interaction lists of the periodic FMM, bit-exact with the reference's
// Tree::build_nonuniform_octree (3dpp/src/tree.cpp:102-188) and Traverser (3dpp/src/traverser.cpp).
//
// The reference builds its adaptive octree with a BFS queue and a stable 8-way partition per split.
// Here the same tree comes from Morton keys: each body descends with the reference's floating-point
// octant test (x > centre, centre_child = centre +/- r/2) so ties resolve identically; a radix sort
// by the key truncated at the body's leaf depth, stable from input order, reproduces the reference's
// body order; BFS numbering is level-major with ascending prefix inside a level.  Lists come from a
// level-synchronous frontier version of the reference's dual-tree recursion and an integer-grid
// neighbour search for the M2L parent map.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "structure.h"

namespace fmm3dpp {

namespace {

constexpr int kMaxDepth = 21;  // 63-bit keys

// LSD radix sort of `idx` by keys[idx] (stable), 11 bits per pass.
void radix_sort_by_key(const std::vector<uint64_t>& keys, std::vector<int32_t>& idx, int key_bits) {
    const int B = 11, R = 1 << B;
    std::vector<int32_t> tmp(idx.size());
    std::vector<int64_t> cnt(R);
    for (int shift = 0; shift < key_bits; shift += B) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int32_t i : idx) cnt[(keys[i] >> shift) & (R - 1)]++;
        int64_t s = 0;
        for (int b = 0; b < R; b++) { int64_t c = cnt[b]; cnt[b] = s; s += c; }
        for (int32_t i : idx) tmp[cnt[(keys[i] >> shift) & (R - 1)]++] = i;
        idx.swap(tmp);
    }
}

// Floating-point child centre exactly as Tree::get_child_cell_x (3dpp/src/tree.cpp:195-210).
inline void child_center(const double* xp, double rp, int oct, double* out) {
    double h = rp / 2;
    out[0] = xp[0] + ((oct & 4) ? h : -h);
    out[1] = xp[1] + ((oct & 2) ? h : -h);
    out[2] = xp[2] + ((oct & 1) ? h : -h);
}

// Integer lattice coordinates of a Morton prefix of `depth` levels (x bit is the high bit of each triple).
inline void deinterleave(uint64_t prefix, int depth, int64_t c[3]) {
    c[0] = c[1] = c[2] = 0;
    for (int l = 0; l < depth; l++) {
        int oct = (int)((prefix >> (3 * (depth - 1 - l))) & 7);
        c[0] = (c[0] << 1) | ((oct >> 2) & 1);
        c[1] = (c[1] << 1) | ((oct >> 1) & 1);
        c[2] = (c[2] << 1) | (oct & 1);
    }
}
inline uint64_t interleave(const int64_t c[3], int depth) {
    uint64_t p = 0;
    for (int l = depth - 1; l >= 0; l--) {
        int oct = (int)((((c[0] >> l) & 1) << 2) | (((c[1] >> l) & 1) << 1) | ((c[2] >> l) & 1));
        p = (p << 3) | (uint64_t)oct;
    }
    return p;
}

struct Frontier { int32_t tc, sc, tcp, scp, oi; };

}  // namespace

int build_structure(const fmm3dpp_config& cfg, int64_t n, const double* X, const double* Y, const double* Z,
                    Structure& st, std::string& err) {
    if (n < 1 || n > (int64_t)INT32_MAX) { err = "number of bodies must be in [1, 2^31)"; return FMM3DPP_ERR_ARG; }
    const double r0 = cfg.cycle / 2;
    const double* c0 = cfg.center;
    auto clk = [] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    const double t_start = clk();
    for (int64_t i = 0; i < n; i++) {
        if (!std::isfinite(X[i]) || !std::isfinite(Y[i]) || !std::isfinite(Z[i])) { err = "non-finite body position"; return FMM3DPP_ERR_ARG; }
    }
    // 1. keys by floating-point descent (tree.cpp:136-140 octant test, 201-208 child centre)
    std::vector<uint64_t> key(n);
    for (int64_t i = 0; i < n; i++) {
        double c[3] = {c0[0], c0[1], c0[2]}, r = r0, nc[3];
        uint64_t k = 0;
        for (int l = 0; l < kMaxDepth; l++) {
            int oct = ((X[i] > c[0]) << 2) | ((Y[i] > c[1]) << 1) | (Z[i] > c[2]);
            k = (k << 3) | (uint64_t)oct;
            child_center(c, r, oct, nc);
            c[0] = nc[0]; c[1] = nc[1]; c[2] = nc[2];
            r = r / 2;
        }
        key[i] = k;
    }
    // 2. sort by full key; 3. leaf depth = first level whose cell holds <= ncrit bodies
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
    radix_sort_by_key(key, order, 3 * kMaxDepth);
    std::vector<int8_t> leaf_depth(n, -1);
    int64_t active = n;
    for (int l = 0; l <= kMaxDepth && active > 0; l++) {
        int shift = 3 * (kMaxDepth - l);
        int64_t s = 0;
        while (s < n) {
            int32_t bi = order[s];
            if (leaf_depth[bi] >= 0) { s++; continue; }
            uint64_t pf = l == 0 ? 0 : (key[bi] >> shift);
            int64_t e = s + 1;
            while (e < n && leaf_depth[order[e]] < 0 && (l == 0 ? 0 : (key[order[e]] >> shift)) == pf) e++;
            if (e - s <= cfg.ncrit) {
                for (int64_t t = s; t < e; t++) leaf_depth[order[t]] = (int8_t)l;
                active -= e - s;
            }
            s = e;
        }
        if (l == kMaxDepth && active > 0) {
            err = "more than ncrit bodies share a cell at depth 21 (coincident bodies?)";
            return FMM3DPP_ERR_TREE;
        }
    }
    // 4. final body order: stable (from input order) by key truncated at the leaf depth
    std::vector<uint64_t> tkey(n);
    for (int64_t i = 0; i < n; i++) {
        int sh = 3 * (kMaxDepth - leaf_depth[i]);
        tkey[i] = sh >= 64 ? 0 : ((key[i] >> sh) << sh);
    }
    for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
    radix_sort_by_key(tkey, order, 3 * kMaxDepth);
    st.n = n;
    st.perm = order;
    st.x.resize(n); st.y.resize(n); st.z.resize(n);
    for (int64_t i = 0; i < n; i++) { st.x[i] = X[order[i]]; st.y[i] = Y[order[i]]; st.z[i] = Z[order[i]]; }
    // 5. cells level by level (BFS numbering)
    st.cells.clear();
    int maxd = 0;
    for (int64_t i = 0; i < n; i++) maxd = std::max(maxd, (int)leaf_depth[i]);
    std::vector<int32_t> level_start(maxd + 2, 0);
    {
        CellRec root{13, 0, 0, 0, 0, (int32_t)n, r0, {c0[0], c0[1], c0[2]}, 0};
        st.cells.push_back(root);
    }
    level_start[0] = 0;
    level_start[1] = 1;
    for (int l = 1; l <= maxd; l++) {
        int shift = 3 * (kMaxDepth - l);
        int32_t pstart = level_start[l - 1], pend = level_start[l];
        int32_t pi = pstart;  // parent pointer (parents at level l-1 in prefix order)
        int64_t s = 0;
        while (s < n) {
            int32_t bi = order[s];
            if (leaf_depth[bi] < l) { s++; continue; }
            uint64_t pf = key[bi] >> shift;
            int64_t e = s + 1;
            while (e < n && leaf_depth[order[e]] >= l && (key[order[e]] >> shift) == pf) e++;
            while (pi < pend && st.cells[pi].prefix != (pf >> 3)) pi++;
            if (pi >= pend) { err = "internal: parent not found"; return FMM3DPP_ERR_TREE; }
            CellRec& par = st.cells[pi];
            CellRec c;
            c.octant = (int32_t)(pf & 7);
            c.depth = l;
            c.coff = 0; c.cnum = 0;
            c.boff = (int32_t)s; c.bnum = (int32_t)(e - s);
            c.r = par.r / 2;
            child_center(par.x, par.r, c.octant, c.x);
            c.prefix = pf;
            if (par.cnum == 0) par.coff = (int32_t)st.cells.size();
            par.cnum++;
            st.cells.push_back(c);
            s = e;
        }
        level_start[l + 1] = (int32_t)st.cells.size();
    }
    st.ntree = (int32_t)st.cells.size();
    st.max_depth = maxd;
    // 6. image cells c(-1..-images) (traverser.cpp:150-162, 186-193)
    const int images = cfg.images;
    if (images >= 1) {
        CellRec c{13, -1, 0, 1, 0, (int32_t)n, st.cells[0].r * 3, {st.cells[0].x[0], st.cells[0].x[1], st.cells[0].x[2]}, 0};
        st.cells.push_back(c);
        for (int m = 1; m < images; m++) {
            int32_t child = (int32_t)st.cells.size() - 1;
            CellRec ci{13, -1 - m, child, 1, 0, 0, st.cells[child].r * 3,
                       {st.cells[child].x[0], st.cells[child].x[1], st.cells[child].x[2]}, 0};
            st.cells.push_back(ci);
        }
    }
    st.min_depth = images >= 1 ? -images : 0;

    const double t_tree = clk();
    st.ms_tree = t_tree - t_start;
    // 7. P2P map: level-synchronous restatement of horizontal_origin (traverser.cpp:70-148)
    const std::vector<CellRec>& C = st.cells;
    std::vector<std::array<double, 3>> offs;
    if (images == 0) offs.push_back({0, 0, 0});
    else
        for (int pz = -1; pz <= 1; pz++)
            for (int py = -1; py <= 1; py++)
                for (int px = -1; px <= 1; px++) offs.push_back({px * cfg.cycle, py * cfg.cycle, pz * cfg.cycle});
    auto adjacent = [&](int a, int b, const double* o) {
        double dist = (C[a].r + C[b].r) * 1.001;
        for (int d = 0; d < 3; d++)
            if (!(std::fabs(C[a].x[d] - C[b].x[d] - o[d]) <= dist)) return false;
        return true;
    };
    auto leaf = [&](int c) { return C[c].cnum == 0; };
    std::vector<Frontier> cur, nxt;
    int32_t top = images == 0 ? 0 : st.ntree;  // c(-1) is the parent of the root for periodic runs
    for (int oi = 0; oi < (int)offs.size(); oi++) cur.push_back({0, 0, top, top, oi});
    st.p2p.clear();
    st.lost = 0;
    while (!cur.empty()) {
        nxt.clear();
        for (const Frontier& f : cur) {
            const double* o = offs[f.oi].data();
            int divide = 0;
            bool emit = false;
            if (!adjacent(f.tc, f.sc, o)) {
                if (adjacent(f.tc, f.scp, o) && leaf(f.tc) && C[f.scp].depth >= C[f.tc].depth) emit = true;
                else if (adjacent(f.tcp, f.sc, o) && leaf(f.sc) && C[f.tcp].depth >= C[f.sc].depth) emit = true;
                else if (adjacent(f.tcp, f.scp, o) && C[f.tcp].depth == C[f.scp].depth) { /* M2L via parent map */ }
                else if (leaf(f.tc) && leaf(f.sc)) st.lost++;
                else if (!leaf(f.tc) && leaf(f.sc)) divide = 1;
                else if (leaf(f.tc) && !leaf(f.sc)) divide = 2;
                else divide = C[f.tc].r >= C[f.sc].r ? 1 : 2;
            } else {
                if (leaf(f.tc) && leaf(f.sc)) emit = true;
                else if (!leaf(f.tc) && leaf(f.sc)) divide = 1;
                else if (leaf(f.tc) && !leaf(f.sc)) divide = 2;
                else divide = C[f.tc].r >= C[f.sc].r ? 1 : 2;
            }
            if (emit) st.p2p.push_back({f.tc, f.sc, {o[0], o[1], o[2]}});
            if (divide == 1)
                for (int i = 0; i < C[f.tc].cnum; i++) nxt.push_back({C[f.tc].coff + i, f.sc, f.tc, f.scp, f.oi});
            else if (divide == 2)
                for (int i = 0; i < C[f.sc].cnum; i++) nxt.push_back({f.tc, C[f.sc].coff + i, f.tcp, f.sc, f.oi});
        }
        cur.swap(nxt);
    }
    std::stable_sort(st.p2p.begin(), st.p2p.end(), [](const P2PRow& a, const P2PRow& b) { return a.tar < b.tar; });

    const double t_p2p = clk();
    st.ms_p2p = t_p2p - t_tree;
    // 8. M2L parent map (traverser.cpp:234-282 and 43-64) on the integer lattice of each level:
    //    source = target + delta (delta in {-1,0,1}^3), wrapped with offset o*cycle when periodic.
    st.m2l.clear();
    std::vector<std::unordered_map<uint64_t, int32_t>> nonleaf(maxd + 1);
    for (int32_t c = 0; c < st.ntree; c++)
        if (!leaf(c)) nonleaf[C[c].depth][C[c].prefix] = c;
    auto pattern_of = [](int dx, int dy, int dz) {  // index in get_relx(1,1) order (kernel.h:176-194)
        int lin = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
        return lin < 13 ? lin : lin - 1;
    };
    for (int32_t t = 0; t < st.ntree; t++) {
        if (leaf(t)) continue;
        int d = C[t].depth;
        int64_t a[3];
        deinterleave(C[t].prefix, d, a);
        int64_t nside = int64_t(1) << d;
        std::vector<M2LRow> np, pp;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (!dx && !dy && !dz) continue;
                    int64_t tt[3] = {a[0] + dx, a[1] + dy, a[2] + dz}, b[3];
                    int o[3];
                    bool inside = true;
                    for (int k = 0; k < 3; k++) {
                        b[k] = ((tt[k] % nside) + nside) % nside;
                        o[k] = (int)((tt[k] - b[k]) / nside);
                        if (o[k] != 0) inside = false;
                    }
                    if (!inside && images == 0) continue;
                    auto it = nonleaf[d].find(interleave(b, d));
                    if (it == nonleaf[d].end()) continue;
                    M2LRow r{t, it->second, {o[0] * cfg.cycle, o[1] * cfg.cycle, o[2] * cfg.cycle}, 0, pattern_of(dx, dy, dz)};
                    (inside ? np : pp).push_back(r);
                }
        auto by_src = [](const M2LRow& p, const M2LRow& q) { return p.src < q.src; };
        std::stable_sort(np.begin(), np.end(), by_src);
        std::stable_sort(pp.begin(), pp.end(), by_src);
        for (auto& r : np) st.m2l.push_back(r);
        for (auto& r : pp) st.m2l.push_back(r);
    }
    for (int m = 1; m < images; m++) {
        int32_t ci = st.ntree + m - 1;  // c(-m)
        double w = C[ci].r * 2;
        for (int pz = -1; pz <= 1; pz++)
            for (int py = -1; py <= 1; py++)
                for (int px = -1; px <= 1; px++)
                    if (px || py || pz) st.m2l.push_back({ci, ci, {px * w, py * w, pz * w}, 1, pattern_of(px, py, pz)});
    }
    st.ms_m2l = clk() - t_p2p;
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp
