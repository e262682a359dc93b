// The precomputed operators of one (P, images, cycle) in HBM, in the layouts the kernels read.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "dbuf.h"
#include "device.h"
#include "structure.h"

namespace fmm3dpp {

struct DeviceOperators {
    DBuf o_ut, o_vs, o_utl, o_vsl;   // factored pseudo-inverses (P2M and L2P), S x S row-major
    DBuf o_m2m_all, o_l2l;           // the 8 octant operators, concatenated
    DBuf o_m2m_img, o_l2l_img;       // image-level M2M (sum over the 27 children) and L2L; zeros without images
    DBuf o_gk8, o_gk8f;              // 8x8 M2L kernel spectra, float64 and (mixed precision) float32
    DBuf o_gkA;                      // the same as matrix-core A operands in lane order, in the Hadamard's type
    DBuf o_gk27;                     // 1x27 image-lattice kernel spectra, summed over the 27 children
    DBuf o_usurf;                    // unit check/equivalent surface
    DBuf o_invlat;                   // P^3 grid corner (k * P + j) * P + i -> surface point, -1 inside
    DBuf o_tw;                       // FFT twiddles of length N = 2P - 1

    hipError_t upload(const Operators& o, bool mixed, hipStream_t s) {
        std::vector<double> l2l_all, m2m_all;
        for (int k = 0; k < 8; k++) l2l_all.insert(l2l_all.end(), o.l2l[k].begin(), o.l2l[k].end());
        for (int k = 0; k < 8; k++) m2m_all.insert(m2m_all.end(), o.m2m[k].begin(), o.m2m[k].end());
        const std::vector<double> zeros((size_t)o.S * o.S, 0.0);
        std::vector<int32_t> inv((size_t)o.P * o.P * o.P, -1);   // surface.h:28-48 lattice order
        for (int sp = 0; sp < o.S; sp++)
            inv[((size_t)o.lattice[3 * sp + 2] * o.P + o.lattice[3 * sp + 1]) * o.P + o.lattice[3 * sp]] = sp;
        hipError_t e = hipSuccess;
        auto up = [&](DBuf& d, const auto& h) {
            if (e == hipSuccess) e = fmm3dpp::upload(d, h, s);
        };
        up(o_ut, o.UT);
        up(o_vs, o.VSinv);
        up(o_utl, o.UT_l2p);
        up(o_vsl, o.VSinv_l2p);
        up(o_m2m_all, m2m_all);
        up(o_l2l, l2l_all);
        up(o_m2m_img, o.images > 0 ? o.m2m_img_sum : zeros);
        up(o_l2l_img, o.images > 0 ? o.l2l_img : zeros);
        up(o_gk8, o.gk8);
        if (mixed) up(o_gk8f, std::vector<float>(o.gk8.begin(), o.gk8.end()));
        up(o_gk27, o.gk27_sum);
        up(o_usurf, o.unit_surf);
        up(o_invlat, inv);
        up(o_tw, o.twiddle);
        if (e == hipSuccess) e = o_gkA.ensure((size_t)o.F * 26 * 64 * 4 * (mixed ? sizeof(float) : sizeof(double)));
        if (e == hipSuccess) e = launch_hadamard_lanes(mixed, o.F, mixed ? o_gk8f.p : o_gk8.p, o_gkA.p, s);
        return e;
    }
};

}  // namespace fmm3dpp
