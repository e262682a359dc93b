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

// Orders P at which the mixed path runs its tree operator stages in float32 (measured with the
// float32 emulation of the restated reference and on the GPU, against the float64 result): from P = 6
// to 8 the difference stays at the float32 level (<= 1.3e-6 in the potential up to a constant, <= 2e-6
// in the gradient); below, the per-cell monopole constraint moves the potential by the float64
// operators' own monopole truncation error (P = 5: 1.2e-5 in the infinite-lattice gauge, P = 4: 7.7e-5
// without images, towards the direct sum), and at P = 10 the float32 L2L of the image lattice's field
// loses 2e-5 in the gradient.
inline bool f32_apply_orders(int P) { return P >= 6 && P <= 8; }

struct DeviceOperators {
    // Operator applies, in the lane layout of launch_apply: the factored pseudo-inverses of P2M and
    // L2P, the 8 octant operators of M2M and L2L (one batch each), and the image-level M2M (the sum
    // over the 27 children) and L2L (zeros without images).  Float64, except the tree operators of the
    // mixed path with float32 applies (the pseudo-inverses there at the truncated rank); the image
    // levels stay float64.
    DBuf l_ut, l_vs, l_utl, l_vsl, l_m2m, l_l2l, l_m2m_img, l_l2l_img;
    LaneOp lo_ut, lo_vs, lo_utl, lo_vsl, lo_m2m, lo_l2l, lo_m2m_img, lo_l2l_img;
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

    // mixed: the mixed-precision path (truncated-rank pseudo-inverses, float32 M2L); f32_apply: its tree
    // operator applies on the float32 matrix cores.
    hipError_t upload(const Operators& o, bool mixed, bool f32_apply, hipStream_t s) {
        const std::vector<double> zeros((size_t)o.S * o.S, 0.0);
        std::vector<int32_t> inv((size_t)o.P * o.P * o.P, -1);   // surface.h:28-48 lattice order
        for (int sp = 0; sp < o.S; sp++)
            inv[((size_t)o.lattice[3 * sp + 2] * o.P + o.lattice[3 * sp + 1]) * o.P + o.lattice[3 * sp]] = sp;
        hipError_t e = apply_init();
        auto up = [&](DBuf& d, const auto& h) {
            if (e == hipSuccess) e = fmm3dpp::upload(d, h, s);
        };
        auto lanes = [&](DBuf& d, LaneOp& lo, const std::vector<const double*>& mats, int M, int K, bool f32) {
            lo.f32 = f32;
            lo.M = M;
            lo.K = K;
            lo.nmb = (M + 15) / 16;
            lo.nkc = (K + 15) / 16;
            lo.zstride = (int64_t)lo.nmb * lo.nkc * 256;
            lo.nops = (int)mats.size();
            const size_t tot = (size_t)lo.zstride * mats.size();
            if (f32) {
                std::vector<float> v(tot);
                for (size_t i = 0; i < mats.size(); i++) lane_layout<float>(mats[i], o.S, M, K, v, i * lo.zstride);
                up(d, v);
            } else {
                std::vector<double> v(tot);
                for (size_t i = 0; i < mats.size(); i++) lane_layout<double>(mats[i], o.S, M, K, v, i * lo.zstride);
                up(d, v);
            }
            lo.p = d.p;
        };
        pinv_rank = mixed ? truncated_rank(o.VSinv, o.S) : 0;
        const int S = o.S, r = pinv_rank ? pinv_rank : S;
        const bool f32 = mixed && f32_apply && f32_apply_orders(o.P);
        std::vector<const double*> m2m, l2l;
        for (int k = 0; k < 8; k++) m2m.push_back(o.m2m[k].data());
        for (int k = 0; k < 8; k++) l2l.push_back(o.l2l[k].data());
        lanes(l_ut, lo_ut, {o.UT.data()}, r, S, f32);          // leading r rows of UT
        lanes(l_vs, lo_vs, {o.VSinv.data()}, S, r, f32);       // leading r columns of VSinv
        lanes(l_utl, lo_utl, {o.UT_l2p.data()}, r, S, f32);
        lanes(l_vsl, lo_vsl, {o.VSinv_l2p.data()}, S, r, f32);
        lanes(l_m2m, lo_m2m, m2m, S, S, f32);
        lanes(l_l2l, lo_l2l, l2l, S, S, f32);
        lanes(l_m2m_img, lo_m2m_img, {o.images > 0 ? o.m2m_img_sum.data() : zeros.data()}, S, S, false);
        lanes(l_l2l_img, lo_l2l_img, {o.images > 0 ? o.l2l_img.data() : zeros.data()}, S, S, false);
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
