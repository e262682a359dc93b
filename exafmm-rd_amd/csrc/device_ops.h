// The precomputed operators of one (P, images, cycle) in HBM, in the layouts the kernels read.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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
    // Rank at which the mixed path applies the P2M / L2P pseudo-inverses (0: full, S): the leading
    // rows of UT and columns of VSinv whose singular values exceed 1e-7 of the largest, rounded up
    // to a multiple of 16.  The dropped directions change the field outside the equivalent surface
    // by ~1e-8 relative (P = 8, 10), three orders below the float32 M2L's own error.
    int pinv_rank = 0;

    // Column j of VSinv = V diag(1 / s) has norm 1 / s_j, with s in descending order.
    static int truncated_rank(const std::vector<double>& vs, int S) {
        std::vector<double> n2(S, 0.0);
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) n2[j] += vs[(size_t)i * S + j] * vs[(size_t)i * S + j];
        double smax = 0;
        for (int j = 0; j < S; j++)
            if (n2[j] > 0) smax = std::max(smax, 1.0 / std::sqrt(n2[j]));
        int r = 0;
        for (int j = 0; j < S; j++)
            if (n2[j] > 0 && 1.0 / std::sqrt(n2[j]) > 1e-7 * smax) r = j + 1;
        return std::min(S, (r + 15) / 16 * 16);
    }

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
        pinv_rank = mixed ? truncated_rank(o.VSinv, o.S) : 0;
        if (e == hipSuccess) e = o_gkA.ensure((size_t)o.F * 26 * 64 * 4 * (mixed ? sizeof(float) : sizeof(double)));
        if (e == hipSuccess) e = launch_hadamard_lanes(mixed, o.F, mixed ? o_gk8f.p : o_gk8.p, o_gkA.p, s);
        return e;
    }
};

}  // namespace fmm3dpp
