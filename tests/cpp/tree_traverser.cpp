// Exercises rtfmm::Tree and rtfmm::Traverser of the C++ front end on the host (structure-only
// plans, no GPU): prints the sizes and every list entry so the Python test can compare them with
// the binding's structure queries, and checks the reference's own tree invariants.
#include <cstdio>

#include "rtfmm_fmm3dpp.hpp"

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 3000;
    const int ncrit = argc > 2 ? atoi(argv[2]) : 32;
    const int images = argc > 3 ? atoi(argv[3]) : 2;
    const double cycle = 2 * M_PI;
    rtfmm::Bodies3 bs = rtfmm::generate_random_bodies(n, cycle / 2, rtfmm::vec3r(0, 0, 0), 5, 1);
    rtfmm::Tree tree;
    tree.build(bs, rtfmm::vec3r(0, 0, 0), cycle / 2, ncrit, rtfmm::Tree::TreeType::nonuniform);
    rtfmm::Cells3 cells = tree.get_cells();
    int bad = 0;
    for (const auto& c : cells) {   // children sit at get_child_cell_x of their parent; body ranges nest
        for (int k = 0; k < c.crange.number; k++) {
            const auto& ch = cells[c.crange.offset + k];
            const rtfmm::vec3r x = rtfmm::Tree::get_child_cell_x(c.x, c.r, ch.octant, 0);
            for (int d = 0; d < 3; d++) bad += x[d] != ch.x[d];
            bad += ch.r != c.r / 2 || ch.depth != c.depth + 1;
            bad += ch.brange.offset < c.brange.offset ||
                   ch.brange.offset + ch.brange.number > c.brange.offset + c.brange.number;
        }
        if (c.crange.number == 0) bad += c.brange.number > ncrit;
        for (int b = c.brange.offset; b < c.brange.offset + c.brange.number; b++)
            for (int d = 0; d < 3; d++) bad += std::fabs(bs[b].x[d] - c.x[d]) > c.r;
    }
    printf("tree cells %zu bad %d first_idx %d\n", cells.size(), bad, bs[0].idx);
    rtfmm::Traverser tr;
    tr.traverse(tree, cycle, images);
    rtfmm::Cells3 all = tr.get_cells();
    rtfmm::PeriodicInteractionPairs p2p = tr.get_pairs(rtfmm::OperatorType::P2P);
    rtfmm::PeriodicInteractionMap pm = tr.get_map(rtfmm::OperatorType::P2P);
    rtfmm::PeriodicM2LMap m2l = tr.get_M2L_parent_map();
    size_t nm = 0;
    for (const auto& kv : m2l) nm += kv.second.size();
    printf("cells %zu p2p %zu p2p_targets %zu m2l %zu\n", all.size(), p2p.size(), pm.size(), nm);
    for (const auto& pr : p2p)
        printf("P %d %d %.17g %.17g %.17g\n", pr.first, pr.second.first, pr.second.second[0], pr.second.second[1],
               pr.second.second[2]);
    for (const auto& kv : m2l)
        for (const auto& s : kv.second)
            printf("M %d %d %.17g %.17g %.17g %d\n", kv.first, s.idx, s.offset[0], s.offset[1], s.offset[2],
                   s.is_single_parent);
    for (int i = 0; i < n; i++) printf("B %d\n", bs[i].idx);
    return bad != 0;
}
