// Host-side data model of the periodic FMM: tree, interaction lists and precomputed operators.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fmm3dpp.h"

namespace fmm3dpp {

struct CellRec {
    int32_t octant, depth, coff, cnum, boff, bnum;
    double r, x[3];
    uint64_t prefix;  // Morton prefix (3 bits per level, root = 0); image cells: 0
};

struct P2PRow { int32_t tar, src; double off[3]; };
struct M2LRow { int32_t tar, src; double off[3]; int32_t img; int32_t pattern; };

// Tree + lists of one particle set (Tree::build + Traverser::traverse of the reference).
struct Structure {
    int64_t n = 0;
    std::vector<int32_t> perm;       // tree order -> input index
    std::vector<double> x, y, z;     // positions in tree order
    std::vector<CellRec> cells;      // BFS order, then image cells c(-1..-images)
    int32_t ntree = 0;
    int32_t min_depth = 0, max_depth = 0;
    std::vector<P2PRow> p2p;         // P2P map rows (targets may be non-leaf)
    std::vector<M2LRow> m2l;         // M2L parent map rows
    int64_t lost = 0;                // well-separated leaf pairs the traversal drops
    double ms_tree = 0, ms_p2p = 0, ms_m2l = 0;  // host build timings of the last call
};

int build_structure(const fmm3dpp_config& cfg, int64_t n, const double* x, const double* y, const double* z,
                    Structure& out, std::string& err);

// Operators of one (P, images, cycle).  Factors follow the reference's members
// (3dpp/include/kernel.h:197-229); fused forms are what the device applies.
struct Operators {
    int P = 0, S = 0, N = 0, Nh = 0, F = 0, images = 0;
    double r0 = 0;
    std::vector<double> UT, VSinv, UT_l2p, VSinv_l2p;     // S x S row-major
    std::vector<double> m2m[8], m2m_img[27], l2l[8], l2l_img;
    std::vector<double> gk8;       // [26][F][64][2] (src child * 8 + tar child)
    std::vector<double> gk27;      // [26][F][28][2]
    // device forms.  P2M/L2P stay factored (UT then VSinv): a pre-multiplied pseudo-inverse
    // rounded to float64 spreads cond*eps errors into directions the far field sees.
    std::vector<double> m2m_cat;   // [M2M0 | ... | M2M7]   (S x 8S)
    std::vector<double> m2m_img_sum;  // sum over the 27 image M2M matrices (S x S)
    std::vector<double> gk27_sum;  // [26][F][2] sum of the 27 lattice kernels per pattern
    std::vector<double> unit_surf; // S x 3 surface points of the unit cube (radius 1, centre 0)
    std::vector<int32_t> lattice;  // S x 3 surface lattice (i, j, k) in [0, P)
    std::vector<double> twiddle;   // N x 2: exp(-2 pi i m / N)
};

int build_operators(int P, int images, double cycle, Operators& ops, std::string& err);
// Recompute the fused forms after a factor was replaced.
void derive_fused(Operators& ops);

// Operator cache file (every field, bit-exact); load checks the configuration and a checksum.
int save_operators(const Operators& ops, double cycle, const std::string& path, std::string& err);
int load_operators(Operators& ops, int P, int images, double cycle, const std::string& path, std::string& err);
// build_operators through the directory named by the environment variable FMM3DPP_OPERATOR_CACHE
// when it is set: the operators of (P, images, cycle) are loaded from their file there when it is
// present and valid (configuration and checksum), otherwise computed and written there (a directory
// that cannot be written costs only the file).  Without the variable, build_operators.
int build_operators_cached(int P, int images, double cycle, Operators& ops, std::string& err);

inline int surface_count(int P) { return 6 * (P - 1) * (P - 1) + 2; }

// Checks a configuration (ranges of P, ncrit, images, cycle, centre, precision, dipole mode).
int validate_config(const fmm3dpp_config* c, std::string& err);

}  // namespace fmm3dpp
