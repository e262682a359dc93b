// One-time operator precompute of the periodic KIFMM (LaplaceKernel::precompute and
// precompute_m2l, 3dpp/src/kernel.cpp:1638-1850), done on the host in 80-bit long double.
//
// The upward check-to-equivalent matrix (equivalent surface 1.05 r, check surface 2.95 r) has
// condition number ~3.6e12 at P = 8, so a float64 SVD leaves operator errors near 1e-4 in the
// weakest directions.  Computing the SVD (one-sided Jacobi), the pseudo-inverse and the operator
// products in long double and rounding once keeps every operator within a few ulps of the exact
// matrices; the 4e-16 relative pseudo-inverse cutoff of the reference (3dpp/src/mathfunc.cpp:154-169)
// is applied unchanged.  Geometry (surface points, convolution grid) is generated in float64 with
// the reference's expressions so both codes sample the same points.
#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include "structure.h"

namespace fmm3dpp {

namespace {

using ld = long double;

// Minimal persistent pool: parallel_for(n, f) runs f(i) for i in [0, n) on all workers.  The
// precompute issues thousands of short jobs in a row (one per round of Jacobi rotations), so a worker
// that finished a job spins briefly for the next one before sleeping, and the caller spins briefly
// for the workers to finish before waiting on the condition variable.
class Pool {
public:
    Pool() {
        unsigned hw = std::thread::hardware_concurrency();
        nthreads_ = std::max(1u, std::min(hw == 0 ? 1u : hw, 32u));
        for (unsigned t = 1; t < nthreads_; t++) workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }
    void parallel_for(int64_t n, const std::function<void(int64_t)>& f) {
        if (n <= 0) return;
        if (nthreads_ == 1 || n == 1) { for (int64_t i = 0; i < n; i++) f(i); return; }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            n_ = n;
            next_.store(0, std::memory_order_relaxed);
            busy_.store((int)nthreads_ - 1, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        run();
        for (int k = 0; k < kSpin && busy_.load(std::memory_order_acquire) != 0; k++) std::this_thread::yield();
        if (busy_.load(std::memory_order_acquire) != 0) {
            std::unique_lock<std::mutex> lk(m_);
            done_.wait(lk, [this] { return busy_.load(std::memory_order_acquire) == 0; });
        }
        job_ = nullptr;
    }
private:
    static constexpr int kSpin = 2000;
    void run() {
        for (;;) {
            int64_t i = next_.fetch_add(1);
            if (i >= n_) break;
            (*job_)(i);
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            for (int k = 0; k < kSpin && gen_.load(std::memory_order_acquire) == seen; k++) std::this_thread::yield();
            if (gen_.load(std::memory_order_acquire) == seen) {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_.load(std::memory_order_acquire) != seen; });
            }
            seen = gen_.load(std::memory_order_acquire);
            if (stop_) return;
            run();
            if (busy_.fetch_sub(1, std::memory_order_acq_rel) == 1) {
                std::lock_guard<std::mutex> g(m_);
                done_.notify_one();
            }
        }
    }
    unsigned nthreads_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int64_t)>* job_ = nullptr;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
    std::atomic<int> busy_{0};
    std::atomic<uint64_t> gen_{0};
    bool stop_ = false;
};

Pool& pool() {
    static Pool p;
    return p;
}

// Surface points (3dpp/src/surface.cpp:3-47, dir 0), float64 as the reference computes them.
void surface(int P, double r, const double c[3], std::vector<double>& out) {
    out.clear();
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++)
            for (int k = 0; k < P; k++)
                if (i == 0 || i == P - 1 || j == 0 || j == P - 1 || k == 0 || k == P - 1) {
                    out.push_back((-1.0 + i * 2.0 / (P - 1)) * r + c[0]);
                    out.push_back((-1.0 + j * 2.0 / (P - 1)) * r + c[1]);
                    out.push_back((-1.0 + k * 2.0 / (P - 1)) * r + c[2]);
                }
}

// K[t][s] = 1/|x_t - x_s| in long double (0 at coincident points).
void laplace_matrix(const std::vector<double>& src, const std::vector<double>& tar, std::vector<ld>& K) {
    size_t ns = src.size() / 3, nt = tar.size() / 3;
    K.assign(nt * ns, 0.0L);
    for (size_t t = 0; t < nt; t++)
        for (size_t s = 0; s < ns; s++) {
            ld dx = (ld)tar[3 * t] - src[3 * s], dy = (ld)tar[3 * t + 1] - src[3 * s + 1], dz = (ld)tar[3 * t + 2] - src[3 * s + 2];
            ld r = sqrtl(dx * dx + dy * dy + dz * dz);
            K[t * ns + s] = r == 0 ? 0.0L : 1.0L / r;
        }
}

// C = A (m x k) * B (k x n), long double, rows in parallel.  Every element is summed over l in
// order from 0; four elements of a row are held in registers while l runs.
void matmul(const std::vector<ld>& A, const std::vector<ld>& B, int m, int k, int n, std::vector<ld>& C) {
    C.assign((size_t)m * n, 0.0L);
    pool().parallel_for(m, [&](int64_t i) {
        ld* c = &C[(size_t)i * n];
        const ld* a = &A[(size_t)i * k];
        int j = 0;
        for (; j + 4 <= n; j += 4) {
            ld c0 = 0, c1 = 0, c2 = 0, c3 = 0;
            for (int l = 0; l < k; l++) {
                const ld al = a[l];
                const ld* b = &B[(size_t)l * n + j];
                c0 += al * b[0];
                c1 += al * b[1];
                c2 += al * b[2];
                c3 += al * b[3];
            }
            c[j] = c0;
            c[j + 1] = c1;
            c[j + 2] = c2;
            c[j + 3] = c3;
        }
        for (; j < n; j++) {
            ld cj = 0;
            for (int l = 0; l < k; l++) cj += a[l] * B[(size_t)l * n + j];
            c[j] = cj;
        }
    });
}

std::vector<double> to_double(const std::vector<ld>& a) {
    std::vector<double> o(a.size());
    for (size_t i = 0; i < a.size(); i++) o[i] = (double)a[i];
    return o;
}

std::vector<double> transpose_d(const std::vector<double>& a, int m, int n) {
    std::vector<double> t((size_t)m * n);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) t[(size_t)j * m + i] = a[(size_t)i * n + j];
    return t;
}

// One-sided Jacobi SVD, long double: A (n x n) = U diag(s) V^T with s descending.
// Column pairs are rotated in round-robin rounds of n/2 disjoint pairs (parallel).
void svd_ld(const std::vector<ld>& A, int n, std::vector<ld>& U, std::vector<ld>& s, std::vector<ld>& V) {
    int nn = n + (n & 1);
    std::vector<ld> W((size_t)nn * n, 0.0L), Vt((size_t)nn * n, 0.0L);  // column-major columns
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) W[(size_t)j * n + i] = A[(size_t)i * n + j];
    for (int j = 0; j < n; j++) Vt[(size_t)j * n + j] = 1.0L;
    std::vector<int> ring(nn);
    for (int i = 0; i < nn; i++) ring[i] = i;
    const ld tol = 4 * LDBL_EPSILON;
    for (int sweep = 0; sweep < 60; sweep++) {
        std::atomic<int> rotated{0};
        for (int round = 0; round < nn - 1; round++) {
            pool().parallel_for(nn / 2, [&](int64_t pi) {
                int p = ring[pi], q = ring[nn - 1 - pi];
                if (p > q) std::swap(p, q);
                if (q >= n) return;
                ld* wp = &W[(size_t)p * n];
                ld* wq = &W[(size_t)q * n];
                ld al = 0, be = 0, ga = 0;
                for (int i = 0; i < n; i++) { al += wp[i] * wp[i]; be += wq[i] * wq[i]; ga += wp[i] * wq[i]; }
                if (ga == 0 || fabsl(ga) <= tol * sqrtl(al * be)) return;
                rotated++;
                ld zeta = (be - al) / (2 * ga);
                ld t = (zeta >= 0 ? 1.0L : -1.0L) / (fabsl(zeta) + sqrtl(1 + zeta * zeta));
                ld c = 1 / sqrtl(1 + t * t), sn = c * t;
                for (int i = 0; i < n; i++) {
                    ld a = wp[i], b = wq[i];
                    wp[i] = c * a - sn * b;
                    wq[i] = sn * a + c * b;
                }
                ld* vp = &Vt[(size_t)p * n];
                ld* vq = &Vt[(size_t)q * n];
                for (int i = 0; i < n; i++) {
                    ld a = vp[i], b = vq[i];
                    vp[i] = c * a - sn * b;
                    vq[i] = sn * a + c * b;
                }
            });
            int last = ring[nn - 1];
            for (int i = nn - 1; i > 1; i--) ring[i] = ring[i - 1];
            ring[1] = last;
        }
        if (rotated.load() == 0) break;
    }
    std::vector<ld> sv(n);
    for (int j = 0; j < n; j++) {
        ld a = 0;
        for (int i = 0; i < n; i++) a += W[(size_t)j * n + i] * W[(size_t)j * n + i];
        sv[j] = sqrtl(a);
    }
    std::vector<int> ord(n);
    for (int j = 0; j < n; j++) ord[j] = j;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sv[a] > sv[b]; });
    U.assign((size_t)n * n, 0.0L); V.assign((size_t)n * n, 0.0L); s.assign(n, 0.0L);
    for (int jj = 0; jj < n; jj++) {
        int j = ord[jj];
        s[jj] = sv[j];
        for (int i = 0; i < n; i++) {
            U[(size_t)i * n + jj] = sv[j] > 0 ? W[(size_t)j * n + i] / sv[j] : 0.0L;
            V[(size_t)i * n + jj] = Vt[(size_t)j * n + i];
        }
    }
}

// Forward r2c DFT (FFTW sign, unnormalised) of a real N^3 grid g[k][j][i], long double.
void r2c_ld(int N, const std::vector<ld>& g, const std::vector<ld>& wr, const std::vector<ld>& wi,
            std::vector<ld>& outr, std::vector<ld>& outi) {
    int Nh = N / 2 + 1;
    size_t F = (size_t)N * N * Nh;
    std::vector<ld> ar(F), ai(F), br(F), bi(F);
    for (int k = 0; k < N; k++)
        for (int j = 0; j < N; j++)
            for (int f = 0; f < Nh; f++) {
                ld sr = 0, si = 0;
                for (int i = 0; i < N; i++) {
                    int m = (i * f) % N;
                    ld v = g[((size_t)k * N + j) * N + i];
                    sr += v * wr[m]; si += v * wi[m];
                }
                ar[((size_t)k * N + j) * Nh + f] = sr; ai[((size_t)k * N + j) * Nh + f] = si;
            }
    for (int k = 0; k < N; k++)
        for (int fj = 0; fj < N; fj++)
            for (int f = 0; f < Nh; f++) {
                ld sr = 0, si = 0;
                for (int j = 0; j < N; j++) {
                    int m = (j * fj) % N;
                    ld xr = ar[((size_t)k * N + j) * Nh + f], xi = ai[((size_t)k * N + j) * Nh + f];
                    sr += xr * wr[m] - xi * wi[m]; si += xr * wi[m] + xi * wr[m];
                }
                br[((size_t)k * N + fj) * Nh + f] = sr; bi[((size_t)k * N + fj) * Nh + f] = si;
            }
    outr.assign(F, 0); outi.assign(F, 0);
    for (int fk = 0; fk < N; fk++)
        for (int fj = 0; fj < N; fj++)
            for (int f = 0; f < Nh; f++) {
                ld sr = 0, si = 0;
                for (int k = 0; k < N; k++) {
                    int m = (k * fk) % N;
                    ld xr = br[((size_t)k * N + fj) * Nh + f], xi = bi[((size_t)k * N + fj) * Nh + f];
                    sr += xr * wr[m] - xi * wi[m]; si += xr * wi[m] + xi * wr[m];
                }
                outr[((size_t)fk * N + fj) * Nh + f] = sr; outi[((size_t)fk * N + fj) * Nh + f] = si;
            }
}

inline int rel_key(int x, int y, int z) { return ((x + 8) * 17 + (y + 8)) * 17 + (z + 8); }

}  // namespace

void derive_fused(Operators& op) {
    int S = op.S;
    op.m2m_cat.assign((size_t)S * 8 * S, 0.0);
    for (int o = 0; o < 8; o++)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) op.m2m_cat[(size_t)i * 8 * S + (size_t)o * S + j] = op.m2m[o][(size_t)i * S + j];
    op.m2m_img_sum.assign((size_t)S * S, 0.0);
    if (op.images > 0 && !op.m2m_img[0].empty())
        for (size_t e = 0; e < (size_t)S * S; e++) {
            ld a = 0;
            for (int o = 0; o < 27; o++) a += op.m2m_img[o][e];
            op.m2m_img_sum[e] = (double)a;
        }
    op.gk27_sum.assign((size_t)26 * op.F * 2, 0.0);
    if (!op.gk27.empty())
        for (size_t pf = 0; pf < (size_t)26 * op.F; pf++) {
            ld sr = 0, si = 0;
            for (int c = 0; c < 28; c++) { sr += op.gk27[pf * 56 + 2 * c]; si += op.gk27[pf * 56 + 2 * c + 1]; }
            op.gk27_sum[2 * pf] = (double)sr;
            op.gk27_sum[2 * pf + 1] = (double)si;
        }
}

int build_operators(int P, int images, double cycle, Operators& op, std::string& err) {
    if (P < 2 || P > 16) { err = "P must be in [2, 16]"; return FMM3DPP_ERR_ARG; }
    const int S = surface_count(P);
    const double r0 = cycle / 2;
    op.P = P; op.S = S; op.N = 2 * P - 1; op.Nh = op.N / 2 + 1; op.F = op.N * op.N * op.Nh;
    op.images = images; op.r0 = r0;
    const double o[3] = {0, 0, 0};
    // unit surface and its lattice (surface.h:28-48)
    surface(P, 1.0, o, op.unit_surf);
    op.lattice.clear();
    for (int i = 0; i < P; i++)
        for (int j = 0; j < P; j++)
            for (int k = 0; k < P; k++)
                if (i == 0 || i == P - 1 || j == 0 || j == P - 1 || k == 0 || k == P - 1) {
                    op.lattice.push_back(i); op.lattice.push_back(j); op.lattice.push_back(k);
                }
    // SVD of the upward equivalent -> check matrix (kernel.cpp:1642-1651)
    std::vector<double> chk, eqv;
    surface(P, r0 * 2.95, o, chk);
    surface(P, r0 * 1.05, o, eqv);
    std::vector<ld> A, U, s, V;
    laplace_matrix(eqv, chk, A);
    svd_ld(A, S, U, s, V);
    ld smax = 0;
    for (ld v : s) smax = std::max(smax, fabsl(v));
    std::vector<ld> sinv(S);
    for (int i = 0; i < S; i++) sinv[i] = s[i] > smax * 4e-16L ? 1.0L / s[i] : 0.0L;
    std::vector<ld> UT((size_t)S * S), VS((size_t)S * S), VSl((size_t)S * S);
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) {
            UT[(size_t)i * S + j] = U[(size_t)j * S + i];
            VS[(size_t)i * S + j] = V[(size_t)i * S + j] * sinv[j];
            VSl[(size_t)i * S + j] = U[(size_t)i * S + j] * sinv[j];
        }
    op.UT = to_double(UT);
    op.VSinv = to_double(VS);
    op.VSinv_l2p = to_double(VSl);
    {
        std::vector<double> Vd = to_double(V);
        op.UT_l2p = transpose_d(Vd, S, S);   // kernel.cpp:1699: UT_l2p = V^T
    }
    // M2M (kernel.cpp:1660-1668) and image M2M (1671-1683): VSinv * (UT * K(child equiv -> parent check))
    std::vector<std::vector<ld>> mats(35);
    for (int oc = 0; oc < 35; oc++) {
        if (oc >= 8 && images == 0) break;
        double cx[3];
        std::vector<double> ce;
        if (oc < 8) {
            double h = r0 / 2;
            cx[0] = (oc & 4) ? h : -h; cx[1] = (oc & 2) ? h : -h; cx[2] = (oc & 1) ? h : -h;
            surface(P, r0 / 2 * 1.05, cx, ce);
        } else {
            int q = oc - 8, kk = q / 9 - 1, jj = (q % 9) / 3 - 1, ii = q % 3 - 1;
            double w = r0 * 2 / 3.0;
            cx[0] = ii * w; cx[1] = jj * w; cx[2] = kk * w;
            surface(P, r0 / 3 * 1.05, cx, ce);
        }
        std::vector<ld> K, m1;
        laplace_matrix(ce, chk, K);
        matmul(UT, K, S, S, S, m1);
        matmul(VS, m1, S, S, S, mats[oc]);
    }
    for (int oc = 0; oc < 8; oc++) {
        op.m2m[oc] = to_double(mats[oc]);
        op.l2l[oc] = transpose_d(op.m2m[oc], S, S);
    }
    for (int oc = 0; oc < 27; oc++) op.m2m_img[oc] = images > 0 ? to_double(mats[8 + oc]) : std::vector<double>();
    op.l2l_img = images > 0 ? transpose_d(op.m2m_img[13], S, S) : std::vector<double>();

    // M2L kernels on the convolution grid (kernel.cpp:1706-1850)
    const int N = op.N, F = op.F;
    const double delta = 2 * r0 / (P - 1) * 1.05, gsize = r0 * 2 * 1.05;
    std::vector<ld> wr(N), wi(N);
    op.twiddle.resize(2 * N);
    for (int m = 0; m < N; m++) {
        ld a = 2.0L * 3.14159265358979323846264338327950288L * m / N;
        wr[m] = cosl(a); wi[m] = -sinl(a);
        op.twiddle[2 * m] = (double)wr[m];
        op.twiddle[2 * m + 1] = (double)wi[m];
    }
    int range = images >= 2 ? 4 : 3;
    std::vector<std::array<int, 3>> rels;
    for (int k = -range; k <= range; k++)
        for (int j = -range; j <= range; j++)
            for (int i = -range; i <= range; i++)
                if (std::abs(i) >= 2 || std::abs(j) >= 2 || std::abs(k) >= 2) rels.push_back({i, j, k});
    std::vector<int> rel_index(17 * 17 * 17, -1);
    for (size_t t = 0; t < rels.size(); t++) rel_index[rel_key(rels[t][0], rels[t][1], rels[t][2])] = (int)t;
    std::vector<std::vector<double>> gk(rels.size());
    pool().parallel_for((int64_t)rels.size(), [&](int64_t t) {
        std::vector<ld> g((size_t)N * N * N), gr, gi;
        double rp[3] = {rels[t][0] * r0 * 2, rels[t][1] * r0 * 2, rels[t][2] * r0 * 2};
        for (int k = 0; k < N; k++)
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) {
                    double px = -gsize + i * delta - rp[0], py = -gsize + j * delta - rp[1], pz = -gsize + k * delta - rp[2];
                    ld r = sqrtl((ld)px * px + (ld)py * py + (ld)pz * pz);
                    g[((size_t)k * N + j) * N + i] = r == 0 ? 0.0L : 1.0L / r;
                }
        r2c_ld(N, g, wr, wi, gr, gi);
        gk[t].resize(2 * (size_t)F);
        for (int f = 0; f < F; f++) { gk[t][2 * f] = (double)gr[f]; gk[t][2 * f + 1] = (double)gi[f]; }
    });
    std::vector<std::array<int, 3>> pars;
    for (int k = -1; k <= 1; k++)
        for (int j = -1; j <= 1; j++)
            for (int i = -1; i <= 1; i++)
                if (i || j || k) pars.push_back({i, j, k});
    op.gk8.assign((size_t)26 * F * 128, 0.0);
    for (int pi = 0; pi < 26; pi++)
        for (int is = 0; is < 8; is++)
            for (int it = 0; it < 8; it++) {
                // child offsets in child-diameter units: parent at 2*rel, children at +-0.5
                int d[3];
                for (int a = 0; a < 3; a++) {
                    int bs = (is >> (2 - a)) & 1, bt = (it >> (2 - a)) & 1;
                    d[a] = 2 * pars[pi][a] + bs - bt;
                }
                if (std::abs(d[0]) <= 1 && std::abs(d[1]) <= 1 && std::abs(d[2]) <= 1) continue;  // adjacent: P2P
                const std::vector<double>& src = gk[rel_index[rel_key(d[0], d[1], d[2])]];
                for (int f = 0; f < F; f++) {
                    size_t id = ((size_t)pi * F + f) * 128 + (is * 8 + it) * 2;
                    op.gk8[id] = src[2 * f];
                    op.gk8[id + 1] = src[2 * f + 1];
                }
            }
    op.gk27.clear();
    if (images >= 2) {
        op.gk27.assign((size_t)26 * F * 56, 0.0);
        for (int pi = 0; pi < 26; pi++) {
            int cnt = 0;
            for (int pz = -1; pz <= 1; pz++)
                for (int py = -1; py <= 1; py++)
                    for (int px = -1; px <= 1; px++, cnt++) {
                        const std::vector<double>& src =
                            gk[rel_index[rel_key(3 * pars[pi][0] + px, 3 * pars[pi][1] + py, 3 * pars[pi][2] + pz)]];
                        for (int f = 0; f < F; f++) {
                            size_t id = ((size_t)pi * F + f) * 56 + cnt * 2;
                            op.gk27[id] = src[2 * f];
                            op.gk27[id + 1] = src[2 * f + 1];
                        }
                    }
        }
    }
    derive_fused(op);
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp
