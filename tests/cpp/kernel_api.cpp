// LaplaceFMM::solve's stage loops (3dpp/src/fmm.cpp:13-81, 83-300) written against the drop-in
// LaplaceKernel operator API (per-cell p2m_precompute ... p2p_1toN_256), compared with the drop-in
// LaplaceFMM::solve on the same bodies.  Prints "rel_phi <x> rel_grad <y> leaves <n>".
// Arguments: n P ncrit images.
#include <cstdio>

#include "rtfmm_fmm3dpp.hpp"

using namespace rtfmm;

int main(int argc, char* argv[]) {
    try {
        const int n = argc > 1 ? atoi(argv[1]) : 3000, P = argc > 2 ? atoi(argv[2]) : 6;
        const int ncrit = argc > 3 ? atoi(argv[3]) : 32, images = argc > 4 ? atoi(argv[4]) : 2;
        const char* av[] = {"kernel_api"};
        Argument args(1, const_cast<char**>(av));
        args.n = n; args.P = P; args.ncrit = ncrit; args.images = images;
        args.dipole_correction = 1;
        args.finalize();
        Bodies3 input = generate_random_bodies(n, args.r, args.x, 5, 1);

        LaplaceFMM fmm(input, args);
        Bodies3 ref = fmm.solve();

        // the stage loops of fmm.cpp over the kernel's per-cell operators
        Bodies3 bs = input;
        Tree tree;
        tree.build(bs, args.x, args.r, args.ncrit, Tree::TreeType::nonuniform);
        Traverser trav;
        trav.traverse(tree, args.cycle, args.images);
        Cells3 cs = trav.get_cells();
        int mind = 0, maxd = 0;
        for (auto& c : cs) { mind = std::min(mind, c.depth); maxd = std::max(maxd, c.depth); }
        LaplaceKernel kernel(args);
        kernel.precompute(args.P, args.r, args.images);
        kernel.precompute_m2l(args.P, args.r, cs, PeriodicInteractionMap(), args.images);
        int leaves = 0;
        for (auto& c : cs)
            if (c.crange.number == 0) { kernel.p2m_precompute(args.P, bs, c); leaves++; }
        for (int d = maxd; d > mind; d--)
            for (auto& c : cs)
                if (c.depth == d && c.crange.number > 0) {
                    if (d >= 0) kernel.m2m_precompute(args.P, c, cs);
                    else kernel.m2m_img_precompute(args.P, c, cs, args.cycle * std::pow(3, -d - 1));
                }
        PeriodicM2LMap m2l = trav.get_M2L_parent_map();
        kernel.m2l_fft_precompute_t(args.P, cs, m2l);
        for (int d = mind; d < maxd; d++)
            for (auto& c : cs)
                if (c.depth == d && c.crange.number > 0) {
                    if (d >= 0) kernel.l2l_precompute(args.P, c, cs);
                    else kernel.l2l_img_precompute(args.P, c, cs);
                }
        for (auto& c : cs)
            if (c.crange.number == 0) kernel.l2p_precompute(args.P, bs, c);
        PeriodicInteractionMap p2p = trav.get_map(OperatorType::P2P);
        for (auto& kv : p2p) kernel.p2p_1toN_256(bs, bs, cs, kv.second, cs[kv.first]);
        dipole_correction(bs, args.cycle);
        Bodies3 out(bs.size());
        for (auto& b : bs) out[b.idx] = b;   // sort_bodies_by_idx (src/body.cpp:177-190)

        double dp = 0, np = 0, dg = 0, ng = 0;
        for (int i = 0; i < n; i++) {
            dp += (out[i].p - ref[i].p) * (out[i].p - ref[i].p);
            np += ref[i].p * ref[i].p;
            for (int d = 0; d < 3; d++) {
                dg += (out[i].f[d] - ref[i].f[d]) * (out[i].f[d] - ref[i].f[d]);
                ng += ref[i].f[d] * ref[i].f[d];
            }
        }
        printf("rel_phi %.3e rel_grad %.3e leaves %d cells %zu\n", std::sqrt(dp / np), std::sqrt(dg / ng), leaves,
               cs.size());
    } catch (const std::exception& ex) {
        printf("error : %s\n", ex.what());
        return 1;
    }
    return 0;
}
