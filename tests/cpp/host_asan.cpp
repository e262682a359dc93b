// Driver of the host-sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer on the host code
// of libfmm3dpp, device code unchanged): exercises the C-ABI's host paths -- argument checks,
// structure-only plans, operator precompute / get / set / cache round trip and a corrupted cache --
// and, when a GPU is present ("gpu" argument), the device paths: solve_device and evaluate in both
// precisions, the per-stage entry points over all and over listed cells, the adjoint, Ewald, direct
// (allocating and not), and a batch with graph replay.  Exits 0 when every call returned what it
// should; a sanitizer report aborts the process with its own exit code.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "fmm3dpp.h"

static int failures = 0;
#define EXPECT(cond, what)                                                  \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("FAIL %s (line %d)\n", what, __LINE__);            \
            failures++;                                                     \
        }                                                                   \
    } while (0)

static fmm3dpp_config make_cfg(int P, int ncrit, int images, int precision, int dipole) {
    fmm3dpp_config c{};
    c.P = P;
    c.ncrit = ncrit;
    c.images = images;
    c.cycle = 1.0;
    c.precision = precision;
    c.dipole_mode = dipole;
    return c;
}

struct Bodies {
    std::vector<double> x, y, z, q;
};

static Bodies uniform(int n, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-0.5, 0.5), c(-1, 1);
    Bodies b;
    b.x.resize(n); b.y.resize(n); b.z.resize(n); b.q.resize(n);
    double s = 0;
    for (int i = 0; i < n; i++) {
        b.x[i] = u(rng); b.y[i] = u(rng); b.z[i] = u(rng); b.q[i] = c(rng); s += b.q[i];
    }
    for (int i = 0; i < n; i++) b.q[i] -= s / n;
    return b;
}

static void host_paths() {
    fmm3dpp_plan* plan = nullptr;
    fmm3dpp_config bad = make_cfg(1, 64, 2, FMM3DPP_PREC_FP64, FMM3DPP_DIPOLE_REFERENCE);
    EXPECT(fmm3dpp_plan_create(&bad, FMM3DPP_DEVICE_NONE, &plan) == FMM3DPP_ERR_ARG, "P=1 refused");
    fmm3dpp_config cfg = make_cfg(4, 32, 2, FMM3DPP_PREC_FP64, FMM3DPP_DIPOLE_REFERENCE);
    EXPECT(fmm3dpp_plan_create(&cfg, FMM3DPP_DEVICE_NONE, &plan) == FMM3DPP_OK, "structure-only plan");
    const int n = 3000;
    Bodies b = uniform(n, 7);
    EXPECT(fmm3dpp_set_particles(plan, 0, b.x.data(), b.y.data(), b.z.data()) == FMM3DPP_ERR_ARG, "n = 0 refused");
    EXPECT(fmm3dpp_set_particles(plan, n, b.x.data(), b.y.data(), b.z.data()) == FMM3DPP_OK, "set_particles");
    fmm3dpp_counts c{};
    EXPECT(fmm3dpp_get_counts(plan, &c) == FMM3DPP_OK, "counts");
    std::vector<int32_t> ints((size_t)c.n_cells * 7), perm(n);
    std::vector<double> dbl((size_t)c.n_cells * 4);
    EXPECT(fmm3dpp_get_cells(plan, ints.data(), dbl.data()) == FMM3DPP_OK, "cells");
    EXPECT(fmm3dpp_get_body_order(plan, perm.data()) == FMM3DPP_OK, "body order");
    std::vector<int32_t> pt(c.n_p2p), ps(c.n_p2p), mt(c.n_m2l), ms(c.n_m2l), mi(c.n_m2l);
    std::vector<double> po((size_t)c.n_p2p * 3), mo((size_t)c.n_m2l * 3);
    EXPECT(fmm3dpp_get_p2p(plan, pt.data(), ps.data(), po.data()) == FMM3DPP_OK, "p2p list");
    EXPECT(fmm3dpp_get_m2l(plan, mt.data(), ms.data(), mo.data(), mi.data()) == FMM3DPP_OK, "m2l list");
    EXPECT(fmm3dpp_precompute(plan) == FMM3DPP_OK, "precompute");
    const int64_t nut = fmm3dpp_get_operator(plan, "UT", nullptr, 0);
    EXPECT(nut == (int64_t)c.S * c.S, "UT size");
    std::vector<double> ut(nut);
    EXPECT(fmm3dpp_get_operator(plan, "UT", ut.data(), nut) == nut, "get UT");
    {   // a buffer one short: the size comes back and nothing is written
        std::vector<double> shortbuf(nut - 1, -7.0);
        EXPECT(fmm3dpp_get_operator(plan, "UT", shortbuf.data(), nut - 1) == nut && shortbuf[0] == -7.0,
               "short buffer left alone");
    }
    EXPECT(fmm3dpp_set_operator(plan, "UT", ut.data(), nut) == FMM3DPP_OK, "set UT");
    EXPECT(fmm3dpp_set_operator(plan, "NOPE", ut.data(), nut) != FMM3DPP_OK, "unknown operator refused");
    const std::string path = "host_asan_ops.bin";
    EXPECT(fmm3dpp_save_operators(plan, path.c_str()) == FMM3DPP_OK, "save operators");
    fmm3dpp_plan* p2 = nullptr;
    EXPECT(fmm3dpp_plan_create(&cfg, FMM3DPP_DEVICE_NONE, &p2) == FMM3DPP_OK, "second plan");
    EXPECT(fmm3dpp_load_operators(p2, path.c_str()) == FMM3DPP_OK, "load operators");
    {   // a corrupted cache: flip bytes in the middle, then truncate
        FILE* f = std::fopen(path.c_str(), "r+b");
        if (f) {
            std::fseek(f, 200, SEEK_SET);
            const char junk[16] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
            std::fwrite(junk, 1, sizeof(junk), f);
            std::fclose(f);
        }
        EXPECT(fmm3dpp_load_operators(p2, path.c_str()) != FMM3DPP_OK, "corrupted cache refused");
        f = std::fopen(path.c_str(), "wb");
        if (f) {
            const char head[10] = {0};
            std::fwrite(head, 1, sizeof(head), f);
            std::fclose(f);
        }
        EXPECT(fmm3dpp_load_operators(p2, path.c_str()) != FMM3DPP_OK, "truncated cache refused");
        std::remove(path.c_str());
    }
    EXPECT(fmm3dpp_plan_destroy(p2) == FMM3DPP_OK, "destroy");
    EXPECT(fmm3dpp_plan_destroy(plan) == FMM3DPP_OK, "destroy");
}

template <class T>
static T* dev_copy(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}

static void device_paths() {
    const int n = 4096;
    Bodies b = uniform(n, 11);
    double *dx = dev_copy(b.x), *dy = dev_copy(b.y), *dz = dev_copy(b.z), *dq = dev_copy(b.q), *dphi = nullptr,
           *dgrad = nullptr;
    EXPECT(dx && dy && dz && dq, "device inputs");
    EXPECT(hipMalloc(&dphi, n * 8) == hipSuccess && hipMalloc(&dgrad, n * 24) == hipSuccess, "device outputs");
    std::vector<double> phi(n), grad(3 * n), phi2(n), grad2(3 * n);
    for (int prec : {FMM3DPP_PREC_FP64, FMM3DPP_PREC_MIXED}) {
        fmm3dpp_config cfg = make_cfg(6, 48, 2, prec, FMM3DPP_DIPOLE_PHYSICAL);
        fmm3dpp_plan* plan = nullptr;
        EXPECT(fmm3dpp_plan_create(&cfg, 0, &plan) == FMM3DPP_OK, "device plan");
        EXPECT(fmm3dpp_solve_device(plan, n, dx, dy, dz, dq, dphi, dgrad) == FMM3DPP_OK, "solve_device");
        EXPECT(fmm3dpp_solve_device(plan, n / 2, dx, dy, dz, dq, dphi, dgrad) == FMM3DPP_OK, "solve_device, fewer");
        EXPECT(fmm3dpp_set_particles(plan, n, b.x.data(), b.y.data(), b.z.data()) == FMM3DPP_OK, "set_particles");
        EXPECT(fmm3dpp_evaluate(plan, b.q.data(), phi.data(), grad.data()) == FMM3DPP_OK, "evaluate");
        // the stages in evaluate's order over every cell, then over listed cells
        EXPECT(fmm3dpp_stage_charges(plan, b.q.data()) == FMM3DPP_OK, "stage charges");
        EXPECT(fmm3dpp_stage_p2m(plan, 0, nullptr) == FMM3DPP_OK, "stage p2m");
        EXPECT(fmm3dpp_stage_m2m(plan, 0, nullptr) == FMM3DPP_OK, "stage m2m");
        EXPECT(fmm3dpp_stage_m2l(plan) == FMM3DPP_OK, "stage m2l");
        EXPECT(fmm3dpp_stage_l2l(plan, 0, nullptr) == FMM3DPP_OK, "stage l2l");
        EXPECT(fmm3dpp_stage_l2p(plan, 0, nullptr) == FMM3DPP_OK, "stage l2p");
        EXPECT(fmm3dpp_stage_p2p(plan, 0, nullptr, 1) == FMM3DPP_OK, "stage p2p");
        EXPECT(fmm3dpp_stage_finalize(plan, phi2.data(), grad2.data()) == FMM3DPP_OK, "stage finalize");
        fmm3dpp_counts c{};
        EXPECT(fmm3dpp_get_counts(plan, &c) == FMM3DPP_OK, "counts");
        std::vector<int32_t> ints((size_t)c.n_cells * 7);
        std::vector<double> dbl((size_t)c.n_cells * 4);
        EXPECT(fmm3dpp_get_cells(plan, ints.data(), dbl.data()) == FMM3DPP_OK, "cells");
        std::vector<int32_t> leaves, parents;
        for (int64_t i = 0; i < c.n_tree_cells; i++) {
            if (ints[i * 7 + 4] == 0) leaves.push_back((int32_t)i);
            else parents.push_back((int32_t)i);
        }
        leaves.resize(std::min<size_t>(leaves.size(), 5));
        parents.resize(std::min<size_t>(parents.size(), 3));
        EXPECT(fmm3dpp_stage_p2m(plan, (int64_t)leaves.size(), leaves.data()) == FMM3DPP_OK, "listed p2m");
        EXPECT(fmm3dpp_stage_m2m(plan, (int64_t)parents.size(), parents.data()) == FMM3DPP_OK, "listed m2m");
        EXPECT(fmm3dpp_stage_l2l(plan, (int64_t)parents.size(), parents.data()) == FMM3DPP_OK, "listed l2l");
        EXPECT(fmm3dpp_stage_l2p(plan, (int64_t)leaves.size(), leaves.data()) == FMM3DPP_OK, "listed l2p");
        EXPECT(fmm3dpp_stage_p2p(plan, (int64_t)leaves.size(), leaves.data(), 0) == FMM3DPP_OK, "listed p2p");
        const int32_t bad_cell = (int32_t)c.n_cells + 5;
        EXPECT(fmm3dpp_stage_p2m(plan, 1, &bad_cell) == FMM3DPP_ERR_ARG, "cell out of range refused");
        EXPECT(fmm3dpp_stage_p2m(plan, 1, parents.data()) == FMM3DPP_ERR_ARG, "non-leaf refused");
        std::vector<double> vec((size_t)leaves.size() * c.S);
        EXPECT(fmm3dpp_get_cell_vectors(plan, 0, (int64_t)leaves.size(), leaves.data(), vec.data()) == FMM3DPP_OK,
               "get cell vectors");
        EXPECT(fmm3dpp_set_cell_vectors(plan, 1, (int64_t)leaves.size(), leaves.data(), vec.data()) == FMM3DPP_OK,
               "set cell vectors");
        EXPECT(fmm3dpp_get_body_fields(plan, 10, 20, phi2.data(), grad2.data()) == FMM3DPP_OK, "get body fields");
        EXPECT(fmm3dpp_get_body_fields(plan, n - 5, 20, phi2.data(), grad2.data()) == FMM3DPP_ERR_ARG,
               "body range past the end refused");
        std::vector<double> dqa(n), dxa(3 * n);
        EXPECT(fmm3dpp_evaluate_adjoint(plan, b.q.data(), b.q.data(), grad.data(), dqa.data(), dxa.data()) ==
                   FMM3DPP_OK,
               "adjoint");
        EXPECT(fmm3dpp_plan_destroy(plan) == FMM3DPP_OK, "destroy");
    }
    std::vector<int32_t> tar(300);
    for (int i = 0; i < 300; i++) tar[i] = 13 * i;
    std::vector<double> pe(300), ge(900);
    EXPECT(fmm3dpp_ewald(0, n, b.x.data(), b.y.data(), b.z.data(), b.q.data(), 300, tar.data(), 1.0, 8, 8.0, 0.5,
                         pe.data(), ge.data()) == FMM3DPP_OK,
           "ewald");
    EXPECT(fmm3dpp_direct(0, n, b.x.data(), b.y.data(), b.z.data(), b.q.data(), 300, tar.data(), 1.0, 1, pe.data(),
                          ge.data()) == FMM3DPP_OK,
           "direct");
    int32_t* dtar = dev_copy(tar);
    EXPECT(fmm3dpp_direct_device_noalloc(nullptr, n, dx, dy, dz, dq, 300, dtar, 1.0, 0, dphi, dgrad) == FMM3DPP_OK,
           "direct, no allocation");
    // a batch of three systems, P 6 and 8, graph replay
    fmm3dpp_config cfgs[3] = {make_cfg(6, 48, 2, FMM3DPP_PREC_MIXED, FMM3DPP_DIPOLE_PHYSICAL),
                              make_cfg(8, 64, 2, FMM3DPP_PREC_MIXED, FMM3DPP_DIPOLE_PHYSICAL),
                              make_cfg(6, 32, 2, FMM3DPP_PREC_MIXED, FMM3DPP_DIPOLE_PHYSICAL)};
    fmm3dpp_batch* batch = nullptr;
    EXPECT(fmm3dpp_batch_create(cfgs, 3, 0, &batch) == FMM3DPP_OK, "batch");
    const int64_t counts[3] = {1500, 2000, 596};
    EXPECT(fmm3dpp_batch_set_particles_device(batch, counts, dx, dy, dz) == FMM3DPP_OK, "batch particles");
    EXPECT(fmm3dpp_batch_set_graph(batch, 1) == FMM3DPP_OK, "batch graph");
    for (int r = 0; r < 3; r++)
        EXPECT(fmm3dpp_batch_evaluate_device(batch, dq, dphi, dgrad) == FMM3DPP_OK, "batch evaluate");
    EXPECT(fmm3dpp_batch_destroy(batch) == FMM3DPP_OK, "batch destroy");
    (void)hipFree(dx); (void)hipFree(dy); (void)hipFree(dz); (void)hipFree(dq); (void)hipFree(dphi);
    (void)hipFree(dgrad); (void)hipFree(dtar);
}

int main(int argc, char** argv) {
    host_paths();
    if (argc > 1 && std::strcmp(argv[1], "gpu") == 0) device_paths();
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
