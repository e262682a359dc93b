// The reference's periodic test driver (3dpp/test/test_pbc.cpp:10-79) over the drop-in API: same
// flags (src/argument.cpp:7-70), same printed blocks ("<title>", the input parameter block,
// "[<stage> time measured : ... seconds.]", the compare() block of src/body.cpp:192-205), so output
// parsers written for the reference read it unchanged.  FMM, Ewald and direct run on the GPU;
// --precision 1 selects the mixed-precision FMM, --device the HIP device.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "rtfmm_fmm3dpp.hpp"

namespace {
double seconds_since(std::chrono::high_resolution_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::high_resolution_clock::now() - t0)
               .count() * 1e-9;
}
}  // namespace

int main(int argc, char* argv[]) {
    try {
        rtfmm::title("rtfmm_3dpp_test_pbc");
        rtfmm::Argument args(argc, argv);
        args.show();
        rtfmm::Bodies3 bs = rtfmm::generate_random_bodies(args.n, args.r, args.x, args.seed, args.zero_netcharge);

        rtfmm::Bodies3 res_fmm;
        if (args.enable_fmm) {
            rtfmm::LaplaceFMM fmm(bs, args);
            auto t0 = std::chrono::high_resolution_clock::now();
            res_fmm = fmm.solve();
            if (args.timing) {
                const char* stages[] = {"P2M", "M2M", "M2L_up", "M2L_hadamard", "M2L_down", "L2L", "L2P", "P2P"};
                printf("[%s time measured : %.5f seconds.]\n", "build_tree_and_lists",
                       fmm3dpp_timing_ms(fmm.plan(), "tree_and_lists") * 1e-3);
                printf("[%s time measured : %.5f seconds.]\n", "precompute",
                       fmm3dpp_timing_ms(fmm.plan(), "precompute") * 1e-3);
                for (const char* st : stages)
                    printf("[%s time measured : %.5f seconds.]\n", st, fmm3dpp_timing_ms(fmm.plan(), st) * 1e-3);
                printf("[%s time measured : %.5f seconds.]\n", "FMM", seconds_since(t0));
            }
        }
        rtfmm::Bodies3 res_ewald;
        if (args.enable_ewald) {
            rtfmm::EwaldSolver ewald(bs, args);
            auto t0 = std::chrono::high_resolution_clock::now();
            res_ewald = ewald.solve();
            if (args.divide_4pi) rtfmm::scale_bodies(res_ewald);
            if (args.timing) printf("[%s time measured : %.5f seconds.]\n", "EWALD", seconds_since(t0));
        }
        rtfmm::Bodies3 res_direct = bs;
        if (args.enable_direct) {
            rtfmm::LaplaceKernel kernel(args.device);
            auto t0 = std::chrono::high_resolution_clock::now();
            kernel.direct(res_direct, res_direct, args.images, args.cycle);
            if (args.dipole_correction) rtfmm::dipole_correction(res_direct, args.cycle);
            if (args.divide_4pi) rtfmm::scale_bodies(res_direct);
            if (args.timing) printf("[%s time measured : %.5f seconds.]\n", "DIRECT", seconds_since(t0));
        }
        if (rtfmm::verbose) {
            if (args.enable_fmm) rtfmm::print_bodies(res_fmm, args.print_body_number, 0, "fmm");
            if (args.enable_direct) rtfmm::print_bodies(res_direct, args.print_body_number, 0, "direct");
            if (args.enable_ewald) rtfmm::print_bodies(res_ewald, args.print_body_number, 0, "ewald");
        }
        if (args.enable_fmm && args.enable_direct)
            rtfmm::compare(res_fmm, res_direct, "FMM", "Direct", args.num_compare).show();
        if (args.enable_fmm && args.enable_ewald)
            rtfmm::compare(res_fmm, res_ewald, "FMM", "Ewald", args.num_compare).show();
        if (args.enable_direct && args.enable_ewald)
            rtfmm::compare(res_direct, res_ewald, "Direct", "Ewald", args.num_compare).show();
    } catch (const std::exception& ex) {
        printf("error : %s\n", ex.what());   // the reference's assert_exit format (include/type.h:42-49)
        return 1;
    }
    return 0;
}
