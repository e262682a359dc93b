// The reference's free-space test driver (3dpp/test/test_fmm.cpp:7-60) over the drop-in API: the
// same flags (src/argument.cpp:7-70), bodies (drand48 seed 5), FMM solve, direct sum through
// LaplaceKernel::p2p over one cell holding every body, and the compare() block of src/body.cpp:192-205.
// FMM and direct run on the GPU; --precision 1 selects the mixed-precision FMM, --device the HIP device.
#include <chrono>
#include <cstdio>

#include "rtfmm_fmm3dpp.hpp"

int main(int argc, char* argv[]) {
    try {
        std::cout << "rtfmm_3dp_test_fmm" << std::endl;
        rtfmm::Argument args(argc, argv);
        args.show();
        rtfmm::Bodies3 bs = rtfmm::generate_random_bodies(args.n, args.r, args.x, 5, args.zero_netcharge);

        rtfmm::LaplaceFMM fmm(bs, args);
        auto t0 = std::chrono::high_resolution_clock::now();
        rtfmm::Bodies3 res_fmm = fmm.solve();
        auto t1 = std::chrono::high_resolution_clock::now();
        if (args.timing)
            printf("[%s time measured : %.5f seconds.]\n", "FMM",
                   std::chrono::duration<double>(t1 - t0).count());

        rtfmm::Bodies3 res_direct = bs;
        rtfmm::LaplaceKernel kernel(args.device);
        rtfmm::Cell3 cell;
        cell.brange = rtfmm::Range(0, args.n);
        t0 = std::chrono::high_resolution_clock::now();
        kernel.p2p(bs, res_direct, cell, cell, rtfmm::vec3r(0, 0, 0), args.use_simd);
        if (args.divide_4pi) rtfmm::scale_bodies(res_direct);
        if (args.dipole_correction) rtfmm::dipole_correction(res_direct, args.cycle);
        t1 = std::chrono::high_resolution_clock::now();
        if (args.timing)
            printf("[%s time measured : %.5f seconds.]\n", "direct",
                   std::chrono::duration<double>(t1 - t0).count());

        if (rtfmm::verbose) {
            rtfmm::print_bodies(res_fmm, args.print_body_number, 0, "res_fmm");
            rtfmm::print_bodies(res_direct, args.print_body_number, 0, "res_direct");
        }
        rtfmm::compare(res_fmm, res_direct, "FMM", "Direct").show();
    } catch (const std::exception& ex) {
        printf("error : %s\n", ex.what());   // the reference's assert_exit format (include/type.h:42-49)
        return 1;
    }
    return 0;
}
