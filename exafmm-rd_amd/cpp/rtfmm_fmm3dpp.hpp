// Drop-in C++ front end of the 3dpp evaluate API (namespace rtfmm) over the fmm3dpp C-ABI.
//
// Mirrors the reference's types and call pattern so a caller of
//   rtfmm::LaplaceFMM fmm(bodies, args); rtfmm::Bodies3 out = fmm.solve();
// (3dpp/include/fmm.h:12-17, 3dpp/test/test_pbc.cpp:28-30) compiles against this header unchanged.
// Argument keeps the reference's public fields and defaults (3dpp/include/argument.h:14-44,
// 3dpp/src/argument.cpp:10-69); Body3 keeps idx, q, p, x, f with f = grad(phi) (3dpp/include/body.h:8-15).
// Errors raise std::runtime_error instead of the reference's exit(1) (3dpp/include/type.h:42-49).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fmm3dpp.h"

namespace rtfmm {

using real = double;

struct vec3r {
    std::array<real, 3> d{0, 0, 0};
    vec3r() = default;
    vec3r(real a, real b, real c) : d{a, b, c} {}
    real& operator[](int i) { return d[i]; }
    const real& operator[](int i) const { return d[i]; }
};

struct Body3 {
    int idx = 0;
    real q = 0;
    real p = 0;
    vec3r x;
    vec3r f;   // gradient of the potential, as the reference stores it
};
using Bodies3 = std::vector<Body3>;

inline int verbose = 0;   // the reference's global rtfmm::verbose (src/argument.cpp:5)

class Argument {
public:
    Argument() { finalize(); }
    // Flags and defaults of the reference's command line (src/argument.cpp:7-70), plus --precision
    // (0 float64, 1 mixed) and --device (HIP device).  Long flags take "--name value" or
    // "--name=value"; short flags "-P 8" or "-P8".
    Argument(int argc, char* argv[]) {
        std::string algo = "fde";
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i], key, val;
            bool has_val = false;
            if (a.rfind("--", 0) == 0) {
                key = a.substr(2);
                size_t eq = key.find('=');
                if (eq != std::string::npos) { val = key.substr(eq + 1); key = key.substr(0, eq); has_val = true; }
            } else if (a.size() >= 2 && a[0] == '-') {
                static const char* shorts[][2] = {{"P", "P"}, {"n", "nbody"}, {"m", "ncrit"}, {"t", "timing"},
                                                  {"i", "images"}, {"v", "verbose"}, {"c", "cycle"},
                                                  {"a", "algorithm"}};
                for (auto& sp : shorts)
                    if (a[1] == sp[0][0]) key = sp[1];
                if (key.empty()) throw std::runtime_error("unknown option " + a);
                if (a.size() > 2) { val = a.substr(2); has_val = true; }
            } else {
                throw std::runtime_error("unexpected argument " + a);
            }
            if (!has_val) {
                if (i + 1 >= argc) throw std::runtime_error("option --" + key + " needs a value");
                val = argv[++i];
            }
            auto I = [&]() { return std::atoi(val.c_str()); };
            if (key == "P") P = I();
            else if (key == "nbody") n = I();
            else if (key == "ncrit") ncrit = I();
            else if (key == "timing") timing = I();
            else if (key == "images") images = I();
            else if (key == "verbose") verbose = I();
            else if (key == "cycle") cycle = std::atof(val.c_str());
            else if (key == "num_compare") num_compare = I();
            else if (key == "ewald_ksize") ewald_ksize = I();
            else if (key == "th_num") th_num = I();
            else if (key == "seed") seed = I();
            else if (key == "algorithm") algo = val;
            else if (key == "check_tree") check_tree = I();
            else if (key == "use_fft") use_fft = I();
            else if (key == "use_precompute") use_precompute = I();
            else if (key == "use_simd") use_simd = I();
            else if (key == "dipole_correction") dipole_correction = I();
            else if (key == "zero_netcharge") zero_netcharge = I();
            else if (key == "print_body_number") print_body_number = I();
            else if (key == "divide_4pi") divide_4pi = I();
            else if (key == "setting_t") setting_t = I();
            else if (key == "precision") precision = I();
            else if (key == "device") device = I();
            else throw std::runtime_error("unknown option --" + key);
        }
        if (P < 2 || P > 32) throw std::runtime_error("option value is invalid: --P=" + std::to_string(P));
        if (images < 0 || images > 20) throw std::runtime_error("option value is invalid: --images=" + std::to_string(images));
        if (num_compare == -1) num_compare = n;
        enable_fmm = algo.find('f') != std::string::npos;
        enable_direct = algo.find('d') != std::string::npos;
        enable_ewald = algo.find('e') != std::string::npos;
        finalize();
    }
    // The reference's parameter block (src/argument.cpp:74-101).
    void show() const {
        printf("input parameters=[\n");
        printf("  %-20s = %d\n", "P", P);
        printf("  %-20s = %d\n", "n", n);
        printf("  %-20s = %d\n", "ncrit", ncrit);
        printf("  %-20s = %d\n", "images", images);
        printf("  %-20s = %.4f\n", "cycle", cycle);
        printf("  %-20s = %d\n", "ewald_ksize", ewald_ksize);
        printf("  %-20s = (%d,%d,%d)\n", "(f,d,e)", enable_fmm, enable_direct, enable_ewald);
        printf("  %-20s = %d\n", "num_compare", num_compare);
        printf("  %-20s = %d\n", "th_num", th_num);
        printf("  %-20s = %d\n", "seed", seed);
        printf("  %-20s = %d\n", "check_tree", check_tree);
        printf("  %-20s = %d\n", "timing", timing);
        printf("  %-20s = %d\n", "verbose", verbose);
        printf("  %-20s = %d\n", "use_fft", use_fft);
        printf("  %-20s = %d\n", "use_precompute", use_precompute);
        printf("  %-20s = %d\n", "use_simd", use_simd);
        printf("  %-20s = %d\n", "dipole_correction", dipole_correction);
        printf("  %-20s = %d\n", "zero_netcharge", zero_netcharge);
        printf("  %-20s = %d\n", "print_body_number", print_body_number);
        printf("  %-20s = %d\n", "divide_4pi", divide_4pi);
        printf("  %-20s = %d\n", "setting_t", setting_t);
        printf("]\n\n");
    }
    vec3r x;
    real r = M_PI;
    int n = 1000;
    int P = 4;
    int ncrit = 16;
    int timing = 1;
    int images = 0;
    real cycle = 2 * M_PI;
    int seed = 5;
    int num_compare = -1;
    int ewald_ksize = 11;
    int th_num = 8;
    int dipole_correction = 1;
    int zero_netcharge = 1;
    int print_body_number = 3;
    int divide_4pi = 0;
    int enable_fmm = 1, enable_direct = 0, enable_ewald = 1;
    int check_tree = 0, use_fft = 1, use_precompute = 1, use_simd = 1, setting_t = 0;
    int device = 0;      // HIP device of the plan (no counterpart in the reference)
    int precision = 0;   // FMM3DPP_PREC_* (no counterpart in the reference)
    // argument.cpp:61-69: box centre 0, half size cycle/2; setting_t switches off dipole/zero-net and divides by 4 pi
    void finalize() {
        x = vec3r(0, 0, 0);
        r = cycle / 2;
        if (setting_t) { dipole_correction = 0; zero_netcharge = 0; divide_4pi = 1; }
    }
};

class LaplaceFMM {
public:
    LaplaceFMM(const Bodies3& bs, const Argument& args) : bs_(bs), args_(args) {
        if ((int)bs_.size() != args_.n) throw std::runtime_error("LaplaceFMM init body size error");
        fmm3dpp_config cfg{};
        cfg.P = args_.P;
        cfg.ncrit = args_.ncrit;
        cfg.images = args_.images;
        cfg.cycle = args_.cycle;
        cfg.center[0] = args_.x[0]; cfg.center[1] = args_.x[1]; cfg.center[2] = args_.x[2];
        cfg.precision = args_.precision;
        cfg.dipole_mode = args_.dipole_correction ? FMM3DPP_DIPOLE_REFERENCE : FMM3DPP_DIPOLE_OFF;
        cfg.divide_4pi = args_.divide_4pi;
        int rc = fmm3dpp_plan_create(&cfg, args_.device, &plan_);
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_plan_create failed: " + std::to_string(rc));
    }
    ~LaplaceFMM() { if (plan_) fmm3dpp_plan_destroy(plan_); }
    LaplaceFMM(const LaplaceFMM&) = delete;
    LaplaceFMM& operator=(const LaplaceFMM&) = delete;

    // Returns the bodies in input order with p and f set (fmm.cpp:13-81).
    Bodies3 solve() {
        const size_t n = bs_.size();
        std::vector<real> x(n), y(n), z(n), q(n), phi(n), grad(3 * n);
        for (size_t i = 0; i < n; i++) { x[i] = bs_[i].x[0]; y[i] = bs_[i].x[1]; z[i] = bs_[i].x[2]; q[i] = bs_[i].q; }
        check(fmm3dpp_set_particles(plan_, (int64_t)n, x.data(), y.data(), z.data()));
        check(fmm3dpp_evaluate(plan_, q.data(), phi.data(), grad.data()));
        Bodies3 out = bs_;
        for (size_t i = 0; i < n; i++) {
            out[i].p = phi[i];
            out[i].f = vec3r(grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
        }
        return out;
    }
    fmm3dpp_plan* plan() { return plan_; }

private:
    void check(int rc) {
        if (rc != FMM3DPP_OK) throw std::runtime_error(std::string("fmm3dpp: ") + fmm3dpp_last_error(plan_));
    }
    Bodies3 bs_;
    Argument args_;
    fmm3dpp_plan* plan_ = nullptr;
};

// generate_random_bodies (3dpp/src/body.cpp:98-129): glibc drand48, x in [-r, r)^3 + offset.
inline Bodies3 generate_random_bodies(int num, real r, vec3r offset = vec3r(0, 0, 0), int seed = 0, int zero_netcharge = 1) {
    Bodies3 bodies;
    double q_avg = 0;
    srand48(seed);
    for (int i = 0; i < num; i++) {
        Body3 b;
        for (int d = 0; d < 3; d++) b.x[d] = drand48() * r * 2 - r;
        double q = drand48() - 0.5;
        q_avg += q;
        b.idx = i;
        b.q = q;
        for (int d = 0; d < 3; d++) b.x[d] += offset[d];
        bodies.push_back(b);
    }
    q_avg /= num;
    if (zero_netcharge)
        for (auto& b : bodies) b.q -= q_avg;
    return bodies;
}

inline void title(const std::string& s) { std::cout << "<" << s << ">\n\n"; }

// scale_bodies (src/body.cpp:88-96): p and f divided by 4 pi.
inline void scale_bodies(Bodies3& bs, real fac = 4 * M_PI) {
    for (auto& b : bs) {
        b.p /= fac;
        for (int d = 0; d < 3; d++) b.f[d] /= fac;
    }
}

// dipole_correction (src/kernel.cpp:1580-1596), the reference's form: p -= c |M|^2 / (n q), f -= c M.
inline void dipole_correction(Bodies3& bs, real cycle) {
    real M[3] = {0, 0, 0};
    for (auto& b : bs)
        for (int d = 0; d < 3; d++) M[d] += b.x[d] * b.q;
    const real coef = 4 * M_PI / (3 * cycle * cycle * cycle);
    const real dn = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
    const real n = (real)bs.size();
    for (auto& b : bs) {
        b.p -= coef * dn / n / b.q;
        for (int d = 0; d < 3; d++) b.f[d] -= coef * M[d];
    }
}

namespace detail {
inline void soa(const Bodies3& bs, std::vector<real>& x, std::vector<real>& y, std::vector<real>& z,
                std::vector<real>& q) {
    const size_t n = bs.size();
    x.resize(n); y.resize(n); z.resize(n); q.resize(n);
    for (size_t i = 0; i < n; i++) { x[i] = bs[i].x[0]; y[i] = bs[i].x[1]; z[i] = bs[i].x[2]; q[i] = bs[i].q; }
}
inline void store(Bodies3& out, const std::vector<real>& phi, const std::vector<real>& grad) {
    for (size_t i = 0; i < out.size(); i++) {
        out[i].p += phi[i];
        for (int d = 0; d < 3; d++) out[i].f[d] += grad[3 * i + d];
    }
}
}  // namespace detail

// ---- tree and interaction lists: the reference's Tree and Traverser (include/tree.h:9-75,
//      include/traverser.h:9-56), built on the host by a structure-only plan (no GPU needed), with
//      the same cell numbering (BFS, image cells appended) and list contents as the evaluation ----
struct Range {   // the reference's OffsetAndNumber (include/type.h:550-565)
    int offset = 0;
    int number = 0;
    Range() = default;
    Range(int o, int n) : offset(o), number(n) {}
};

struct Cell3 {   // include/tree.h:9-30
    int idx = 0;
    int octant = 0;   // 0..7 below a parent, 13 for the single child of an image cell
    int depth = 0;
    real r = 0;
    vec3r x;
    Range crange;
    Range brange;
    // Equivalent density and check potential (S values each once an operator has written them; empty
    // reads as zeros).  LaplaceFMM keeps them in HBM; LaplaceKernel's per-cell operators read and
    // write them here, as the reference's Matrix members.
    std::vector<real> q_equiv, p_check;
};
using Cells3 = std::vector<Cell3>;

namespace detail {
// A structure-only plan (FMM3DPP_DEVICE_NONE) for one box and leaf size.
class StructurePlan {
public:
    StructurePlan(int ncrit, int images, real cycle, const vec3r& center) {
        fmm3dpp_config cfg{};
        cfg.P = 4;   // the structure does not depend on P
        cfg.ncrit = ncrit;
        cfg.images = images;
        cfg.cycle = cycle;
        for (int d = 0; d < 3; d++) cfg.center[d] = center[d];
        cfg.precision = FMM3DPP_PREC_FP64;
        cfg.dipole_mode = FMM3DPP_DIPOLE_OFF;
        if (fmm3dpp_plan_create(&cfg, FMM3DPP_DEVICE_NONE, &plan_) != FMM3DPP_OK)
            throw std::runtime_error("tree: invalid box or leaf size");
    }
    ~StructurePlan() { fmm3dpp_plan_destroy(plan_); }
    StructurePlan(const StructurePlan&) = delete;
    StructurePlan& operator=(const StructurePlan&) = delete;

    void set(const Bodies3& bs) {
        std::vector<real> x, y, z, q;
        soa(bs, x, y, z, q);
        check(fmm3dpp_set_particles(plan_, (int64_t)bs.size(), x.data(), y.data(), z.data()));
        check(fmm3dpp_get_counts(plan_, &counts_));
    }
    std::vector<int32_t> order() {
        std::vector<int32_t> perm(counts_.n_bodies);
        check(fmm3dpp_get_body_order(plan_, perm.data()));
        return perm;
    }
    Cells3 cells(bool with_images) {
        std::vector<int32_t> ints(7 * (size_t)counts_.n_cells);
        std::vector<double> dbl(4 * (size_t)counts_.n_cells);
        check(fmm3dpp_get_cells(plan_, ints.data(), dbl.data()));
        const int64_t nc = with_images ? counts_.n_cells : counts_.n_tree_cells;
        Cells3 out((size_t)nc);
        for (int64_t c = 0; c < nc; c++) {
            const int32_t* I = &ints[7 * c];
            const double* D = &dbl[4 * c];
            Cell3& cell = out[c];
            cell.idx = I[0];
            cell.octant = I[1];
            cell.depth = I[2];
            cell.crange = Range(I[3], I[4]);
            cell.brange = Range(I[5], I[6]);
            cell.r = D[0];
            cell.x = vec3r(D[1], D[2], D[3]);
        }
        return out;
    }
    const fmm3dpp_counts& counts() const { return counts_; }
    fmm3dpp_plan* plan() { return plan_; }
    void check(int rc) {
        if (rc != FMM3DPP_OK) throw std::runtime_error(std::string("fmm3dpp: ") + fmm3dpp_last_error(plan_));
    }

private:
    fmm3dpp_plan* plan_ = nullptr;
    fmm3dpp_counts counts_{};
};
}  // namespace detail

class Tree {
public:
    enum class TreeType { uniform, nonuniform };

    // Adaptive octree of box (x, r) with at most m bodies per leaf (src/tree.cpp:102-188); reorders
    // `bodies` into tree order as the reference does.  Only the nonuniform tree is on the
    // evaluation path (src/fmm.cpp:19).
    void build(Bodies3& bodies, vec3r x, real r, int m, TreeType type) {
        if (type != TreeType::nonuniform)
            throw std::runtime_error("tree: only the adaptive (nonuniform) octree of the evaluation path is built");
        detail::StructurePlan sp(m, 0, 2 * r, x);
        sp.set(bodies);
        const std::vector<int32_t> perm = sp.order();
        Bodies3 sorted(bodies.size());
        for (size_t i = 0; i < sorted.size(); i++) sorted[i] = bodies[perm[i]];
        bodies.swap(sorted);
        cells_ = sp.cells(false);
        bodies_ = bodies;
        x_ = x;
        r_ = r;
        ncrit_ = m;
    }
    Cells3 get_cells() { return cells_; }

    // Centre of child `octant` of a cell (src/tree.cpp:195-219): octant bits x = 4, y = 2, z = 1;
    // periodic children 0..26 of an image cell sit on a 3x3x3 grid of spacing 2 r / 3.
    static vec3r get_child_cell_x(vec3r x_par, real r_par, int octant, int is_periodic) {
        vec3r x;
        if (!is_periodic) {
            if (octant < 0 || octant > 7) throw std::runtime_error("octant out of range");
            for (int d = 0; d < 3; d++) x[d] = x_par[d] + ((octant >> (2 - d)) & 1 ? r_par / 2 : -r_par / 2);
        } else {
            if (octant < 0 || octant > 26) throw std::runtime_error("periodic octant out of range");
            const int k[3] = {octant % 3 - 1, (octant % 9) / 3 - 1, octant / 9 - 1};
            for (int d = 0; d < 3; d++) x[d] = x_par[d] + k[d] * (r_par * 2 / 3.0);
        }
        return x;
    }

    const Bodies3& bodies() const { return bodies_; }
    vec3r center() const { return x_; }
    real radius() const { return r_; }
    int ncrit() const { return ncrit_; }

private:
    Cells3 cells_;
    Bodies3 bodies_;
    vec3r x_;
    real r_ = 0;
    int ncrit_ = 0;
};

enum class OperatorType { P2P, M2L, M2P, P2L };
using PeriodicInteractionPair = std::pair<int, std::pair<int, vec3r>>;
using PeriodicInteractionPairs = std::vector<PeriodicInteractionPair>;
using PeriodicInteractionMap = std::map<int, std::vector<std::pair<int, vec3r>>>;

struct PeriodicParentSource {   // include/traverser.h:17-23
    PeriodicParentSource(int idx_, vec3r offset_, int is_single_parent_)
        : idx(idx_), offset(offset_), is_single_parent(is_single_parent_) {}
    int idx;
    vec3r offset;
    int is_single_parent;
};
using PeriodicM2LMap = std::map<int, std::vector<PeriodicParentSource>>;

// Interaction lists of a tree (src/traverser.cpp:19-68): the P2P list of every leaf with periodic
// offsets and the M2L parent map (same-depth adjacent non-leaf parents, image cells' 26 self
// entries).  The image cells c(-1..-images) are appended to the cells.  cycle must equal the tree's
// box size 2 r when images >= 1 (the reference's driver always builds it so, src/argument.cpp:61-62).
class Traverser {
public:
    void traverse(Tree& tree, real cycle = 2 * M_PI, int images = 0) {
        if (images >= 1 && std::fabs(cycle - 2 * tree.radius()) > 1e-12 * cycle)
            throw std::runtime_error("traverser: cycle must equal the tree's box size 2 r");
        detail::StructurePlan sp(tree.ncrit(), images, 2 * tree.radius(), tree.center());
        sp.set(tree.bodies());
        cells_ = sp.cells(true);
        const fmm3dpp_counts& c = sp.counts();
        std::vector<int32_t> tar(c.n_p2p), src(c.n_p2p);
        std::vector<double> off(3 * (size_t)c.n_p2p);
        sp.check(fmm3dpp_get_p2p(sp.plan(), tar.data(), src.data(), off.data()));
        p2p_.clear();
        for (int64_t i = 0; i < c.n_p2p; i++)
            p2p_.push_back({tar[i], {src[i], vec3r(off[3 * i], off[3 * i + 1], off[3 * i + 2])}});
        std::vector<int32_t> mt(c.n_m2l), ms(c.n_m2l), mi(c.n_m2l);
        std::vector<double> mo(3 * (size_t)c.n_m2l);
        sp.check(fmm3dpp_get_m2l(sp.plan(), mt.data(), ms.data(), mo.data(), mi.data()));
        m2l_.clear();
        for (int64_t i = 0; i < c.n_m2l; i++)
            m2l_[mt[i]].push_back(PeriodicParentSource(ms[i], vec3r(mo[3 * i], mo[3 * i + 1], mo[3 * i + 2]), mi[i]));
    }
    PeriodicInteractionPairs get_pairs(OperatorType type) {
        if (type != OperatorType::P2P) throw std::runtime_error("traverser: only P2P pairs are on the evaluation path");
        return p2p_;
    }
    PeriodicInteractionMap get_map(OperatorType type) {
        if (type != OperatorType::P2P) throw std::runtime_error("traverser: only the P2P map is on the evaluation path");
        PeriodicInteractionMap m;
        for (const auto& pr : p2p_) m[pr.first].push_back(pr.second);
        return m;
    }
    PeriodicM2LMap get_M2L_parent_map() { return m2l_; }
    Cells3 get_cells() { return cells_; }

private:
    Cells3 cells_;
    PeriodicInteractionPairs p2p_;
    PeriodicM2LMap m2l_;
};

// LaplaceKernel (include/kernel.h:15-115) on the GPU.
//  * direct (src/kernel.cpp:11-77): bs_tar.p/f += the direct lattice sum of bs_src over (3^images)^3
//    copies; sources and targets are the same bodies, as the reference's test driver calls it.
//  * p2p (src/kernel.cpp:79-100): one cell pair.
//  * the operator API of the evaluation path (precompute, p2m_precompute, m2m_precompute,
//    m2m_img_precompute, m2l_fft_precompute_t, l2l_precompute, l2l_img_precompute, l2p_precompute,
//    p2p_1toN_256): the same signatures and the same effect on the caller's Cell3::q_equiv / p_check
//    and Body3::p / f, computed by the per-stage entry points of the C-ABI on a device plan.  The
//    device needs the tree the cells come from, so a kernel used for these is constructed with the
//    run's Argument (ncrit, box centre, cycle, device); the first call that passes the bodies builds
//    the same tree on the GPU from them (bodies in tree order, as Tree::build leaves them) and checks
//    every cell it is given against it.  Differences to the reference, all in its scratch use:
//    p2m_precompute writes q_equiv (the reference adds the bodies' check potential to the q_equiv it
//    finds, zero on its evaluation path, before the solve), and l2p_precompute leaves p_check
//    unchanged (the reference overwrites it with the downward equivalent density).
class LaplaceKernel {
public:
    explicit LaplaceKernel(int device = 0) : device_(device) {}
    explicit LaplaceKernel(const Argument& args)
        : device_(args.device), ncrit_(args.ncrit), cycle_(args.cycle), center_(args.x), bound_args_(true) {}
    ~LaplaceKernel() { if (plan_) fmm3dpp_plan_destroy(plan_); }
    LaplaceKernel(const LaplaceKernel&) = delete;
    LaplaceKernel& operator=(const LaplaceKernel&) = delete;

    // precompute (kernel.h:113, src/kernel.cpp:1638-1704): the operators of (P, images) for a box of
    // half size r0 (cycle 2 r0), M2L kernel spectra included.
    void precompute(int P, real r0, int images) {
        if (!bound_args_) throw std::runtime_error("LaplaceKernel: construct with the run's Argument for the operator API");
        if (images >= 1 && std::fabs(2 * r0 - cycle_) > 1e-12 * cycle_)
            throw std::runtime_error("LaplaceKernel::precompute: r0 must be half the Argument's cycle");
        if (plan_) fmm3dpp_plan_destroy(plan_);
        plan_ = nullptr;
        fmm3dpp_config cfg{};
        cfg.P = P;
        cfg.ncrit = ncrit_;
        cfg.images = images;
        cfg.cycle = 2 * r0;
        for (int d = 0; d < 3; d++) cfg.center[d] = center_[d];
        cfg.precision = FMM3DPP_PREC_FP64;
        cfg.dipole_mode = FMM3DPP_DIPOLE_OFF;
        int rc = fmm3dpp_plan_create(&cfg, device_, &plan_);
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_plan_create failed: " + std::to_string(rc));
        check(fmm3dpp_precompute(plan_));
        P_ = P;
        bodies_ = nullptr;
        nbodies_ = 0;
        S_ = 6 * (P - 1) * (P - 1) + 2;
    }
    // precompute_m2l (kernel.h:115, src/kernel.cpp:1706-1850): the M2L spectra of every relative
    // position are part of precompute here; this checks the arguments only.
    void precompute_m2l(int P, real r0, Cells3 cs, PeriodicInteractionMap m2l_map, int images) {
        (void)cs; (void)m2l_map; (void)r0; (void)images;
        need(P);
    }

    // p2m_precompute (kernel.h:43, src/kernel.cpp:332-371): cell_src.q_equiv = 0.5^depth VSinv UT
    // (check potential of the leaf's bodies on its 2.95 r surface).
    void p2m_precompute(int P, Bodies3& bs_src, Cell3& cell_src) {
        need(P);
        bind(bs_src, true);
        const int32_t c = leaf(cell_src);
        check(fmm3dpp_stage_p2m(plan_, 1, &c));
        cell_src.q_equiv.resize(S_);
        check(fmm3dpp_get_cell_vectors(plan_, 0, 1, &c, cell_src.q_equiv.data()));
    }
    // m2m_precompute (kernel.h:47, src/kernel.cpp:395-404): parent.q_equiv += sum of M2M[octant]
    // child.q_equiv over the parent's children in cs.
    void m2m_precompute(int P, Cell3& cell_parent, Cells3& cs) {
        need(P);
        upward(cell_parent, cs, false);
    }
    // m2m_img_precompute (kernel.h:51, src/kernel.cpp:434-442): image cell parent.q_equiv += the 27
    // image M2M operators applied to its single child's q_equiv.  `cycle` (the image cell's box, as
    // the reference's driver passes it) does not enter: the operators depend on ratios of radii only.
    void m2m_img_precompute(int P, Cell3& cell_parent, Cells3& cs, real cycle) {
        (void)cycle;
        need(P);
        upward(cell_parent, cs, true);
    }
    // m2l_fft_precompute_t (kernel.h:97, src/kernel.cpp:1140-1397): every cell's p_check += M2L over
    // the M2L parent map of the tree (Traverser::get_M2L_parent_map of the same tree, checked by size).
    void m2l_fft_precompute_t(int P, Cells3& cs, PeriodicM2LMap& m2l_parent_map) {
        need(P);
        need_bound();
        if ((int64_t)cs.size() != counts_.n_cells) throw std::runtime_error("m2l: cs must hold every cell, image cells included");
        int64_t rows = 0;
        for (const auto& kv : m2l_parent_map) rows += (int64_t)kv.second.size();
        if (rows != counts_.n_m2l) throw std::runtime_error("m2l: the M2L parent map is not the tree's");
        std::vector<real> qe((size_t)S_ * cs.size(), 0.0), pc((size_t)S_ * cs.size(), 0.0);
        for (size_t c = 0; c < cs.size(); c++) {
            if (cs[c].q_equiv.size() == (size_t)S_) std::copy(cs[c].q_equiv.begin(), cs[c].q_equiv.end(), &qe[c * S_]);
            if (cs[c].p_check.size() == (size_t)S_) std::copy(cs[c].p_check.begin(), cs[c].p_check.end(), &pc[c * S_]);
        }
        check(fmm3dpp_set_cell_vectors(plan_, 0, 0, nullptr, qe.data()));
        check(fmm3dpp_set_cell_vectors(plan_, 1, 0, nullptr, pc.data()));
        check(fmm3dpp_stage_m2l(plan_));
        check(fmm3dpp_get_cell_vectors(plan_, 1, 0, nullptr, pc.data()));
        for (size_t c = 0; c < cs.size(); c++) cs[c].p_check.assign(&pc[c * S_], &pc[(c + 1) * S_]);
    }
    // l2l_precompute (kernel.h:101, src/kernel.cpp:1419-1428): each child's p_check += L2L[octant]
    // parent.p_check.
    void l2l_precompute(int P, Cell3& cell_parent, Cells3& cs) {
        need(P);
        downward(cell_parent, cs);
    }
    // l2l_img_precompute (kernel.h:103, src/kernel.cpp:1430-1435): the single child's p_check +=
    // L2L_IMG parent.p_check.
    void l2l_img_precompute(int P, Cell3& cell_parent, Cells3& cs) {
        need(P);
        downward(cell_parent, cs);
    }
    // l2p_precompute (kernel.h:107, src/kernel.cpp:1458-1511): bodies of cell_tar get p += and f +=
    // the field of q_down = 0.5^depth VSinv_l2p UT_l2p p_check at their positions.
    void l2p_precompute(int P, Bodies3& bs_tar, Cell3& cell_tar) {
        need(P);
        bind(bs_tar, false);
        const int32_t c = leaf(cell_tar);
        std::vector<real> pc = cell_tar.p_check;
        pc.resize(S_, 0.0);
        check(fmm3dpp_set_cell_vectors(plan_, 1, 1, &c, pc.data()));
        check(fmm3dpp_stage_l2p(plan_, 1, &c));
        add_fields(bs_tar, cell_tar);
    }
    // p2p_1toN_256 (kernel.h:39, src/kernel.cpp:258-311): bodies of cell_tar in bs_tar get p += sum q/r
    // and f += sum -q dx/r^3 over the bodies of every (cell, offset) of p2ps in bs_src, shifted by the
    // offset (r = 0 skipped): one direct sum on the GPU over exactly the listed sources.
    void p2p_1toN_256(Bodies3& bs_src, Bodies3& bs_tar, Cells3& cs, std::vector<std::pair<int, vec3r>>& p2ps,
                      Cell3& cell_tar) {
        const int nt = cell_tar.brange.number;
        if (nt <= 0) return;
        if (cell_tar.brange.offset < 0 || cell_tar.brange.offset + nt > (int)bs_tar.size())
            throw std::runtime_error("p2p_1toN_256: target body range out of range");
        std::vector<real> x, y, z, q;
        for (const auto& e : p2ps) {
            const Cell3& c = cs.at(e.first);
            if (c.brange.offset < 0 || c.brange.offset + c.brange.number > (int)bs_src.size())
                throw std::runtime_error("p2p_1toN_256: source body range out of range");
            for (int i = 0; i < c.brange.number; i++) {
                const Body3& b = bs_src[c.brange.offset + i];
                x.push_back(b.x[0] + e.second[0]);
                y.push_back(b.x[1] + e.second[1]);
                z.push_back(b.x[2] + e.second[2]);
                q.push_back(b.q);
            }
        }
        const int ns = (int)x.size();
        std::vector<int32_t> tar(nt);
        for (int j = 0; j < nt; j++) {   // the targets, with zero charge
            const Body3& b = bs_tar[cell_tar.brange.offset + j];
            x.push_back(b.x[0]); y.push_back(b.x[1]); z.push_back(b.x[2]); q.push_back(0.0);
            tar[j] = ns + j;
        }
        std::vector<real> phi(nt), grad(3 * (size_t)nt);
        int rc = fmm3dpp_direct(device_, (int64_t)x.size(), x.data(), y.data(), z.data(), q.data(), nt, tar.data(), 1.0,
                                0, phi.data(), grad.data());
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_direct failed: " + std::to_string(rc));
        for (int j = 0; j < nt; j++) {
            Body3& b = bs_tar[cell_tar.brange.offset + j];
            b.p += phi[j];
            for (int d = 0; d < 3; d++) b.f[d] += grad[3 * (size_t)j + d];
        }
    }

    void direct(const Bodies3& bs_src, Bodies3& bs_tar, int images, real cycle) {
        if (&bs_src != &bs_tar && bs_src.size() != bs_tar.size())
            throw std::runtime_error("direct: sources and targets must be the same bodies");
        std::vector<real> x, y, z, q;
        detail::soa(bs_src, x, y, z, q);
        const int n = (int)x.size();
        std::vector<int32_t> tar(n);
        for (int i = 0; i < n; i++) tar[i] = i;
        std::vector<real> phi(n), grad(3 * (size_t)n);
        int rc = fmm3dpp_direct(device_, n, x.data(), y.data(), z.data(), q.data(), n, tar.data(), cycle, images,
                                phi.data(), grad.data());
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_direct failed: " + std::to_string(rc));
        detail::store(bs_tar, phi, grad);
    }

    // p2p (src/kernel.cpp:79-100): the bodies of cell_tar in bs_tar get p += sum q/r and
    // f += sum -q dx/r^3 over the bodies of cell_src in bs_src shifted by offset (dx = x_t - x_s -
    // offset, r = 0 skipped), on the GPU.  use_simd only selects the reference's CPU code path.
    void p2p(const Bodies3& bs_src, Bodies3& bs_tar, const Cell3& cell_src, const Cell3& cell_tar, vec3r offset,
             int use_simd = 1) {
        (void)use_simd;
        const int ns = cell_src.brange.number, nt = cell_tar.brange.number;
        if (ns <= 0 || nt <= 0) return;
        if (cell_src.brange.offset < 0 || cell_src.brange.offset + ns > (int)bs_src.size() ||
            cell_tar.brange.offset < 0 || cell_tar.brange.offset + nt > (int)bs_tar.size())
            throw std::runtime_error("p2p: body range out of range");
        // one body set: the shifted sources, then the targets with zero charge
        const int n = ns + nt;
        std::vector<real> x(n), y(n), z(n), q(n, 0.0);
        for (int i = 0; i < ns; i++) {
            const Body3& b = bs_src[cell_src.brange.offset + i];
            x[i] = b.x[0] + offset[0]; y[i] = b.x[1] + offset[1]; z[i] = b.x[2] + offset[2]; q[i] = b.q;
        }
        std::vector<int32_t> tar(nt);
        for (int j = 0; j < nt; j++) {
            const Body3& b = bs_tar[cell_tar.brange.offset + j];
            x[ns + j] = b.x[0]; y[ns + j] = b.x[1]; z[ns + j] = b.x[2];
            tar[j] = ns + j;
        }
        std::vector<real> phi(nt), grad(3 * (size_t)nt);
        int rc = fmm3dpp_direct(device_, n, x.data(), y.data(), z.data(), q.data(), nt, tar.data(), 1.0, 0,
                                phi.data(), grad.data());
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_direct failed: " + std::to_string(rc));
        for (int j = 0; j < nt; j++) {
            Body3& b = bs_tar[cell_tar.brange.offset + j];
            b.p += phi[j];
            for (int d = 0; d < 3; d++) b.f[d] += grad[3 * (size_t)j + d];
        }
    }

private:
    void check(int rc) {
        if (rc != FMM3DPP_OK)
            throw std::runtime_error(std::string("fmm3dpp: ") + (plan_ ? fmm3dpp_last_error(plan_) : "no plan"));
    }
    void need_plan() {
        if (!plan_) throw std::runtime_error("LaplaceKernel: precompute must come first");
    }
    void need(int P) {
        need_plan();
        if (P != P_) throw std::runtime_error("LaplaceKernel: P differs from the precompute's");
    }
    void need_bound() {
        if (!bodies_) throw std::runtime_error("LaplaceKernel: no bodies yet (p2m_precompute binds them)");
    }
    // The first call with a body set builds the device tree and lists from it; a body set of another
    // address or size rebuilds.  charges: upload the charges (they may change between P2M and P2P).
    void bind(const Bodies3& bs, bool charges) {
        if (&bs != bodies_ || bs.size() != nbodies_) {
            std::vector<real> x, y, z, q;
            detail::soa(bs, x, y, z, q);
            check(fmm3dpp_set_particles(plan_, (int64_t)bs.size(), x.data(), y.data(), z.data()));
            check(fmm3dpp_get_counts(plan_, &counts_));
            std::vector<int32_t> perm(bs.size());
            check(fmm3dpp_get_body_order(plan_, perm.data()));
            for (size_t i = 0; i < perm.size(); i++)
                if (perm[i] != (int32_t)i) throw std::runtime_error("LaplaceKernel: bodies must be in tree order (Tree::build)");
            ints_.assign(7 * (size_t)counts_.n_cells, 0);
            std::vector<double> dbl(4 * (size_t)counts_.n_cells);
            check(fmm3dpp_get_cells(plan_, ints_.data(), dbl.data()));
            bodies_ = &bs;
            nbodies_ = bs.size();
            charges = true;
        }
        if (charges) {
            std::vector<real> x, y, z, q;
            detail::soa(bs, x, y, z, q);
            check(fmm3dpp_stage_charges(plan_, q.data()));
        }
    }
    // The cell must be the device tree's cell of the same index (same body range, children).
    int32_t cell_index(const Cell3& c) {
        need_bound();
        if (c.idx < 0 || c.idx >= counts_.n_cells) throw std::runtime_error("LaplaceKernel: cell index out of range");
        const int32_t* I = &ints_[7 * (size_t)c.idx];
        if (I[2] != c.depth || I[3] != c.crange.offset || I[4] != c.crange.number || I[5] != c.brange.offset ||
            I[6] != c.brange.number)
            throw std::runtime_error("LaplaceKernel: cell does not belong to the tree of these bodies");
        return c.idx;
    }
    int32_t leaf(const Cell3& c) {
        const int32_t i = cell_index(c);
        if (c.crange.number != 0) throw std::runtime_error("LaplaceKernel: not a leaf cell");
        return i;
    }
    static std::vector<real> padded(const std::vector<real>& v, int S) {
        std::vector<real> out = v;
        out.resize(S, 0.0);
        return out;
    }
    // M2M of one parent from its children's q_equiv in cs, added to parent.q_equiv.
    void upward(Cell3& parent, Cells3& cs, bool image) {
        const int32_t p = cell_index(parent);
        if (image != (parent.depth < 0)) throw std::runtime_error("m2m: image operator on a tree cell or the reverse");
        for (int k = 0; k < parent.crange.number; k++) {
            const Cell3& ch = cs.at(parent.crange.offset + k);
            const int32_t c = cell_index(ch);
            std::vector<real> v = padded(ch.q_equiv, S_);
            check(fmm3dpp_set_cell_vectors(plan_, 0, 1, &c, v.data()));
        }
        check(fmm3dpp_stage_m2m(plan_, 1, &p));
        std::vector<real> r(S_);
        check(fmm3dpp_get_cell_vectors(plan_, 0, 1, &p, r.data()));
        parent.q_equiv.resize(S_, 0.0);
        for (int i = 0; i < S_; i++) parent.q_equiv[i] += r[i];
    }
    // L2L of one parent: children's p_check (from cs) += L2L parent.p_check.
    void downward(Cell3& parent, Cells3& cs) {
        const int32_t p = cell_index(parent);
        std::vector<real> v = padded(parent.p_check, S_);
        check(fmm3dpp_set_cell_vectors(plan_, 1, 1, &p, v.data()));
        std::vector<int32_t> kids;
        for (int k = 0; k < parent.crange.number; k++) {
            Cell3& ch = cs.at(parent.crange.offset + k);
            const int32_t c = cell_index(ch);
            std::vector<real> w = padded(ch.p_check, S_);
            check(fmm3dpp_set_cell_vectors(plan_, 1, 1, &c, w.data()));
            kids.push_back(c);
        }
        check(fmm3dpp_stage_l2l(plan_, 1, &p));
        for (int k = 0; k < parent.crange.number; k++) {
            Cell3& ch = cs[parent.crange.offset + k];
            ch.p_check.resize(S_);
            check(fmm3dpp_get_cell_vectors(plan_, 1, 1, &kids[k], ch.p_check.data()));
        }
    }
    void add_fields(Bodies3& bs, const Cell3& c) {
        const int n = c.brange.number;
        std::vector<real> phi(n), grad(3 * (size_t)n);
        check(fmm3dpp_get_body_fields(plan_, c.brange.offset, n, phi.data(), grad.data()));
        for (int i = 0; i < n; i++) {
            Body3& b = bs[c.brange.offset + i];
            b.p += phi[i];
            for (int d = 0; d < 3; d++) b.f[d] += grad[3 * (size_t)i + d];
        }
    }

    int device_;
    int ncrit_ = 0;
    real cycle_ = 0;
    vec3r center_;
    bool bound_args_ = false;
    fmm3dpp_plan* plan_ = nullptr;
    int P_ = 0, S_ = 0;
    const Bodies3* bodies_ = nullptr;
    size_t nbodies_ = 0;
    fmm3dpp_counts counts_{};
    std::vector<int32_t> ints_;
};

// EwaldSolver (include/ewald.h:20-21, src/ewald.cpp:6-203) on the GPU with the reference's
// parameters: alpha = ksize / cycle, cutoff = cycle / 2.
class EwaldSolver {
public:
    EwaldSolver(const Bodies3& bs, const Argument& args) : bs_(bs), args_(args) {}
    Bodies3 solve() {
        std::vector<real> x, y, z, q;
        detail::soa(bs_, x, y, z, q);
        const int n = (int)x.size();
        std::vector<int32_t> tar(n);
        for (int i = 0; i < n; i++) tar[i] = i;
        std::vector<real> phi(n), grad(3 * (size_t)n);
        int rc = fmm3dpp_ewald(args_.device, n, x.data(), y.data(), z.data(), q.data(), n, tar.data(), args_.cycle,
                               args_.ewald_ksize, args_.ewald_ksize / args_.cycle, args_.cycle / 2, phi.data(),
                               grad.data());
        if (rc != FMM3DPP_OK) throw std::runtime_error("fmm3dpp_ewald failed: " + std::to_string(rc));
        Bodies3 out = bs_;
        for (auto& b : out) { b.p = 0; b.f = vec3r(0, 0, 0); }
        detail::store(out, phi, grad);
        return out;
    }

private:
    Bodies3 bs_;
    Argument args_;
};

// compare / BodyCompareResult::show (src/body.cpp:131-175, 192-205), same formulas and print format.
struct BodyCompareResult {
    std::string name1, name2;
    int num_compared = 0;
    real l2p = 0, l2f = 0, l2e = 0, rmsp = 0, rmsf = 0, epot1 = 0, epot2 = 0;
    void show() const {
        int bar = 72 - (int)name1.size() - (int)name2.size() - 4, left = bar / 2;
        for (int i = 0; i < left; i++) printf("-");
        printf("%s vs %s", name1.c_str(), name2.c_str());
        for (int i = 0; i < bar - left; i++) printf("-");
        printf("[%d]\n", num_compared);
        printf("%-8s : %8.5e   %-8s : %8.5e   %-8s : %8.5e\n", "L2  (p)", l2p, "L2  (f)", l2f, "L2  (e)", l2e);
        printf("%-8s : %8.5e   %-8s : %8.5e\n", "Rms (p)", rmsp, "Rms (f)", rmsf);
        printf("p-energy1 : %8.12e\n", epot1);
        printf("p-energy2 : %8.12e\n", epot2);
        printf("\n");
    }
};

inline BodyCompareResult compare(const Bodies3& bs1, const Bodies3& bs2, std::string name1, std::string name2,
                                 int num_compare = -1) {
    if (bs1.size() != bs2.size()) throw std::runtime_error("inconsistent size in comparison");
    BodyCompareResult res;
    res.name1 = name1;
    res.name2 = name2;
    const int num = num_compare == -1 ? (int)bs1.size() : std::min(num_compare, (int)bs1.size());
    res.num_compared = num;
    real pdif = 0, pnrm = 0, fdif = 0, fnrm = 0, e1 = 0, e2 = 0;
    for (int i = 0; i < num; i++) {
        const Body3 &b1 = bs1[i], &b2 = bs2[i];
        e1 += b1.p * b1.q;
        e2 += b2.p * b2.q;
        pdif += (b1.p - b2.p) * (b1.p - b2.p);
        pnrm += b2.p * b2.p;
        for (int d = 0; d < 3; d++) {
            fdif += (b1.f[d] - b2.f[d]) * (b1.f[d] - b2.f[d]);
            fnrm += b2.f[d] * b2.f[d];
        }
    }
    printf("pnrm = %.8f\n", pnrm);
    res.rmsp = std::sqrt(pdif / num);
    res.rmsf = std::sqrt(fdif / num);
    res.l2p = std::sqrt(pdif / pnrm);
    res.l2f = std::sqrt(fdif / fnrm);
    res.epot1 = e1;
    res.epot2 = e2;
    res.l2e = std::sqrt((e1 - e2) * (e1 - e2) / e2 / e2);
    return res;
}

// print_bodies (src/body.cpp:16-26)
inline void print_bodies(const Bodies3& bs, int num, int offset, const std::string& name) {
    std::cout << name << ":" << std::endl;
    if (num == -1) num = (int)bs.size();
    const int s = std::min(num, (int)bs.size());
    for (int i = offset; i < offset + s && i < (int)bs.size(); i++) {
        const Body3& b = bs[i];   // print_body (src/body.cpp:5-14)
        printf("[%d],(%.4f,%.4f,%.4f),%.4f,%.12f,(%.8f,%.8f,%.8f)\n", b.idx, b.x[0], b.x[1], b.x[2], b.q, b.p, b.f[0],
               b.f[1], b.f[2]);
    }
    printf("\n");
}

}  // namespace rtfmm
