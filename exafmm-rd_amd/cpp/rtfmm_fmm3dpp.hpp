// Drop-in C++ front end of the 3dpp evaluate API (namespace rtfmm) over the fmm3dpp C-ABI.
//
// Mirrors the reference's types and call pattern so a caller of
//   rtfmm::LaplaceFMM fmm(bodies, args); rtfmm::Bodies3 out = fmm.solve();
// (3dpp/include/fmm.h:12-17, 3dpp/test/test_pbc.cpp:28-30) compiles against this header unchanged.
// Argument keeps the reference's public fields and defaults (3dpp/include/argument.h:14-44,
// 3dpp/src/argument.cpp:10-69); Body3 keeps idx, q, p, x, f with f = grad(phi) (3dpp/include/body.h:8-15).
// Errors raise std::runtime_error instead of the reference's exit(1) (3dpp/include/type.h:42-49).
#pragma once
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
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

class Argument {
public:
    Argument() { finalize(); }
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
    int device = 0;   // HIP device of the plan (no counterpart in the reference)
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
        cfg.precision = FMM3DPP_PREC_FP64;
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

}  // namespace rtfmm
