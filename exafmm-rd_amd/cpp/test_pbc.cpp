// Periodic run through the drop-in C++ API, printing the reference's energy line
// (3dpp/test/test_pbc.cpp:10-79 and the compare() block of 3dpp/src/body.cpp:192-205).
// Usage: test_pbc P n ncrit images [seed]   (cycle 2*pi, drand48 bodies as the reference)
#include <cstdio>
#include <cstdlib>

#include "rtfmm_fmm3dpp.hpp"

int main(int argc, char** argv) {
    rtfmm::Argument args;
    args.P = argc > 1 ? atoi(argv[1]) : 8;
    args.n = argc > 2 ? atoi(argv[2]) : 24000;
    args.ncrit = argc > 3 ? atoi(argv[3]) : 128;
    args.images = argc > 4 ? atoi(argv[4]) : 3;
    args.seed = argc > 5 ? atoi(argv[5]) : 5;
    args.finalize();
    rtfmm::Bodies3 bs = rtfmm::generate_random_bodies(args.n, args.r, args.x, args.seed, args.zero_netcharge);
    try {
        rtfmm::LaplaceFMM fmm(bs, args);
        rtfmm::Bodies3 res = fmm.solve();
        double e = 0;
        for (auto& b : res) e += b.p * b.q;
        printf("[FMM_kernels time measured : %.5f seconds.]\n", fmm3dpp_timing_ms(fmm.plan(), "evaluate") * 1e-3);
        printf("p-energy1 : %8.12e\n", e);
    } catch (const std::exception& ex) {
        fprintf(stderr, "error : %s\n", ex.what());
        return 1;
    }
    return 0;
}
