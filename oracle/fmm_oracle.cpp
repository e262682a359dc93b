// TEST INFRASTRUCTURE ONLY — never linked into, loaded by, or called from the product path.
//
// CPU restatement (C++17 + OpenMP, FP64) of the `3dpp` periodic kernel-independent FMM
// evaluation path of mumairsadiq/exafmm-rd.  It is the oracle the HIP product's parity tests compare
// against (tests/, __graft_entry__.smoke()) and the reported CPU baseline (bench.py cpu_baseline leg).
// Every function names the reference file:line it follows.  The reference's third-party
// numerics (LAPACK dgesvd, FFTW r2c/c2r, CBLAS dgemv) are absent here and are restated:
//   * SVD      : one-sided Jacobi (Hestenes) — replaces dgesvd_ (3dpp/src/mathfunc.cpp:81-110)
//   * FFT      : separable direct DFT on the (2P-1)^3 grid — replaces fftw r2c/c2r
//                (3dpp/src/kernel.cpp:1290-1295,1347-1352,1725)
//   * GEMV     : plain loops — replaces cblas_dgemv (3dpp/src/mathfunc.cpp:11-23)
// Pinning of each part is documented in the tests that exercise it (tests/test_oracle_*.py).
// It also emulates the mixed-precision path's reduced-precision arithmetic stage by stage (float32,
// split bfloat16; the "emulate*" options, no counterpart in the reference), so that the error each
// stage would add can be measured on the host (tests/test_precision_emulation.py).
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <queue>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>
#include <chrono>
#include <omp.h>

namespace orc {

using cplx = std::complex<double>;

struct Cell {
    int idx, octant, depth;
    double r;
    double x[3];
    int coff, cnum;  // child range (crange)
    int boff, bnum;  // body range  (brange)
};
struct P2PEntry { int src; double off[3]; };
struct M2LEntry { int src; double off[3]; int img; };

static inline int surf_num(int p) { return 6 * (p - 1) * (p - 1) + 2; }   // surface.h:10-13

// surface.cpp:3-47 (dir = 0): i outer, k inner, points on the faces of [-1,1]^3 scaled by r.
static void surface_points(int p, double r, const double c[3], std::vector<double>& out) {
    out.clear();
    for (int i = 0; i < p; i++)
        for (int j = 0; j < p; j++)
            for (int k = 0; k < p; k++)
                if (i == 0 || i == p - 1 || j == 0 || j == p - 1 || k == 0 || k == p - 1) {
                    double a = -1.0 + i * 2.0 / (p - 1);
                    double b = -1.0 + j * 2.0 / (p - 1);
                    double d = -1.0 + k * 2.0 / (p - 1);
                    out.push_back(a * r + c[0]);
                    out.push_back(b * r + c[1]);
                    out.push_back(d * r + c[2]);
                }
}

// surface.h:28-48: surface index -> (i,j,k) on the P^3 lattice (same enumeration as above).
static void surface_lattice(int p, std::vector<int>& ijk) {
    ijk.clear();
    for (int i = 0; i < p; i++)
        for (int j = 0; j < p; j++)
            for (int k = 0; k < p; k++)
                if (i == 0 || i == p - 1 || j == 0 || j == p - 1 || k == 0 || k == p - 1) {
                    ijk.push_back(i); ijk.push_back(j); ijk.push_back(k);
                }
}

// tree.cpp:195-219
static void child_center(const double xp[3], double rp, int octant, int periodic, double out[3]) {
    if (!periodic) {
        double h = rp / 2;
        out[0] = xp[0] + ((octant & 4) ? h : -h);
        out[1] = xp[1] + ((octant & 2) ? h : -h);
        out[2] = xp[2] + ((octant & 1) ? h : -h);
    } else {
        int k = octant / 9 - 1, j = (octant % 9) / 3 - 1, i = octant % 3 - 1;
        double s = rp * 2 / 3.0;
        out[0] = xp[0] + i * s;
        out[1] = xp[1] + j * s;
        out[2] = xp[2] + k * s;
    }
}

// kernel.cpp:1535-1556: M[t*ns + s] = 1/|x_t - x_s| (0 at r == 0).
static void kernel_matrix(const std::vector<double>& src, const std::vector<double>& tar, std::vector<double>& M) {
    int ns = (int)src.size() / 3, nt = (int)tar.size() / 3;
    M.assign((size_t)nt * ns, 0.0);
    for (int j = 0; j < nt; j++)
        for (int i = 0; i < ns; i++) {
            double dx = tar[3 * j] - src[3 * i], dy = tar[3 * j + 1] - src[3 * i + 1], dz = tar[3 * j + 2] - src[3 * i + 2];
            double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            M[(size_t)j * ns + i] = r == 0 ? 0 : 1 / r;
        }
}

// C = A(m x k) * B(k x n), row-major
static void matmul(const std::vector<double>& A, const std::vector<double>& B, int m, int k, int n, std::vector<double>& C) {
    C.assign((size_t)m * n, 0.0);
    for (int i = 0; i < m; i++)
        for (int l = 0; l < k; l++) {
            double a = A[(size_t)i * k + l];
            const double* b = &B[(size_t)l * n];
            double* c = &C[(size_t)i * n];
            for (int j = 0; j < n; j++) c[j] += a * b[j];
        }
}

static void transpose(const std::vector<double>& A, int m, int n, std::vector<double>& T) {
    T.assign((size_t)m * n, 0.0);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) T[(size_t)j * m + i] = A[(size_t)i * n + j];
}

// y = k * A(m x n) x  (cblas_dgemv semantics of mathfunc.cpp:11-23, beta = 0)
static inline void gemv(const double* A, const double* x, double* y, int m, int n, double k) {
    for (int i = 0; i < m; i++) {
        const double* a = A + (size_t)i * n;
        double s = 0;
        for (int j = 0; j < n; j++) s += a[j] * x[j];
        y[i] = k * s;
    }
}

// ---- reduced-precision emulation (test infrastructure for the mixed-precision path) ----
// EMUL_MONOPOLE: every tree cell's equivalent density is shifted uniformly so that its sum (the
// monopole) equals the charge of the cell's bodies, after P2M (leaves) and M2M (parents).
// EMUL_M2L: the M2L of tree parents (depth >= 0) in float32: equivalent densities and twiddles rounded
// to float32, up transform, 8x8 Hadamard (the spectra, the kernel tables and every partial sum in
// float32, the target spectra accumulated in float32 over the patterns) and down transform in float32,
// the check potentials accumulated in float64; the image cells' M2L stays float64.  With a split
// bfloat16 kind the Hadamard's products are formed as that scheme forms them (transforms float32).
enum EmulBits { EMUL_P2M_CHECK = 1, EMUL_P2M_PINV = 2, EMUL_M2M = 4, EMUL_L2L = 8, EMUL_L2P_PINV = 16,
                EMUL_L2P_EVAL = 32, EMUL_MONOPOLE = 64, EMUL_M2L = 128 };

// Uniform shift of v[0..S) that makes its sum equal `charge`.
static void set_monopole(double* v, int S, double charge) {
    double sum = 0;
    for (int s = 0; s < S; s++) sum += v[s];
    const double shift = (charge - sum) / S;
    for (int s = 0; s < S; s++) v[s] += shift;
}
// float32 operands and accumulation; split bfloat16 with 2 pieces and 3 products (hi*hi, hi*lo,
// lo*hi), or 3 pieces and the 6 products of order <= 2, each product exact in float32 and
// accumulated in float32 (the matrix-core split-product schemes).
enum EmulKind { EMUL_F32 = 0, EMUL_BF16X3 = 1, EMUL_BF16X6 = 2 };

// Round to the nearest bfloat16 (ties to even), returned as a float.
static inline float bf16_round(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}

static inline void bf16_split(double v, int pieces, float* out) {
    double r = v;
    for (int i = 0; i < pieces; i++) {
        out[i] = bf16_round((float)r);
        r -= out[i];
    }
}

// a * b as the matrix cores' split-bfloat16 scheme `kind` forms it: both split into 2 (BF16X3) or 3
// (BF16X6) bfloat16 pieces, the products of order < pieces, each exact, summed in float32.
static inline float split_product(double a, double b, int kind) {
    const int pieces = kind == EMUL_BF16X3 ? 2 : 3;
    float as[3], bs[3];
    bf16_split(a, pieces, as);
    bf16_split(b, pieces, bs);
    float s = 0;
    for (int u = 0; u < pieces; u++)
        for (int v = 0; u + v < pieces; v++) s += as[u] * bs[v];
    return s;
}

// y[i] = k * sum_{j<n} A[i*lda + j] x[j] in the arithmetic `kind` (EmulKind), rounded back to double.
static void gemv_emul(const double* A, int lda, const double* x, double* y, int m, int n, double k, int kind) {
    if (kind == EMUL_F32) {
        std::vector<float> xf(n);
        for (int j = 0; j < n; j++) xf[j] = (float)x[j];
        for (int i = 0; i < m; i++) {
            const double* a = A + (size_t)i * lda;
            float s = 0;
            for (int j = 0; j < n; j++) s += (float)a[j] * xf[j];
            y[i] = k * (double)s;
        }
        return;
    }
    const int pieces = kind == EMUL_BF16X3 ? 2 : 3;
    std::vector<float> xs((size_t)n * 3);
    for (int j = 0; j < n; j++) bf16_split(x[j], pieces, &xs[(size_t)j * 3]);
    for (int i = 0; i < m; i++) {
        const double* a = A + (size_t)i * lda;
        float s = 0;
        for (int j = 0; j < n; j++) {
            float as[3];
            bf16_split(a[j], pieces, as);
            const float* xj = &xs[(size_t)j * 3];
            for (int u = 0; u < pieces; u++)
                for (int v = 0; u + v < pieces; v++) s += as[u] * xj[v];
        }
        y[i] = k * (double)s;
    }
}

// One-sided Jacobi SVD of a square n x n row-major A: A = U diag(s) V^T, s descending.
// Stands in for LAPACK dgesvd_ (mathfunc.cpp:81-110).  Rotations are scheduled in
// round-robin rounds of disjoint column pairs so each round runs in parallel.
static void svd_jacobi(const std::vector<double>& A, int n, std::vector<double>& U, std::vector<double>& s, std::vector<double>& V) {
    int m = n;
    int nn = n + (n & 1);
    std::vector<double> W((size_t)nn * m, 0.0), Vc((size_t)nn * nn, 0.0);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) W[(size_t)j * m + i] = A[(size_t)i * n + j];
    for (int j = 0; j < nn; j++) Vc[(size_t)j * nn + j] = 1.0;
    std::vector<int> ring(nn);
    for (int i = 0; i < nn; i++) ring[i] = i;
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotated = 0;
        for (int round = 0; round < nn - 1; round++) {
#pragma omp parallel for reduction(+ : rotated) schedule(static)
            for (int pi = 0; pi < nn / 2; pi++) {
                int p = ring[pi], q = ring[nn - 1 - pi];
                if (p > q) std::swap(p, q);
                if (q >= n) continue;
                double* wp = &W[(size_t)p * m];
                double* wq = &W[(size_t)q * m];
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < m; i++) { al += wp[i] * wp[i]; be += wq[i] * wq[i]; ga += wp[i] * wq[i]; }
                if (ga == 0 || std::fabs(ga) <= tol * std::sqrt(al * be)) continue;
                rotated++;
                double zeta = (be - al) / (2 * ga);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                double c = 1 / std::sqrt(1 + t * t), sn = c * t;
                for (int i = 0; i < m; i++) {
                    double a = wp[i], b = wq[i];
                    wp[i] = c * a - sn * b;
                    wq[i] = sn * a + c * b;
                }
                double* vp = &Vc[(size_t)p * nn];
                double* vq = &Vc[(size_t)q * nn];
                for (int i = 0; i < n; i++) {
                    double a = vp[i], b = vq[i];
                    vp[i] = c * a - sn * b;
                    vq[i] = sn * a + c * b;
                }
            }
            // rotate ring positions 1..nn-1
            int last = ring[nn - 1];
            for (int i = nn - 1; i > 1; i--) ring[i] = ring[i - 1];
            ring[1] = last;
        }
        if (rotated == 0) break;
    }
    std::vector<double> sv(n);
    for (int j = 0; j < n; j++) {
        double a = 0;
        for (int i = 0; i < m; i++) a += W[(size_t)j * m + i] * W[(size_t)j * m + i];
        sv[j] = std::sqrt(a);
    }
    std::vector<int> order(n);
    for (int j = 0; j < n; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sv[a] > sv[b]; });
    U.assign((size_t)m * n, 0.0);
    V.assign((size_t)n * n, 0.0);
    s.assign(n, 0.0);
    for (int jj = 0; jj < n; jj++) {
        int j = order[jj];
        s[jj] = sv[j];
        for (int i = 0; i < m; i++) U[(size_t)i * n + jj] = sv[j] > 0 ? W[(size_t)j * m + i] / sv[j] : 0.0;
        for (int i = 0; i < n; i++) V[(size_t)i * n + jj] = Vc[(size_t)j * nn + i];
    }
}

// ---- separable DFT on the N^3 convolution grid (x fastest), FFTW r2c/c2r conventions ----
struct Dft {
    int N, Nh;
    std::vector<cplx> w;  // w[m] = exp(-2 pi i m / N)
    std::vector<std::complex<float>> wf;  // w rounded to float32 (reduced-precision emulation)
    explicit Dft(int n = 1) : N(n), Nh(n / 2 + 1), w(n), wf(n) {
        for (int m = 0; m < n; m++) w[m] = cplx(std::cos(2 * M_PI * m / n), -std::sin(2 * M_PI * m / n));
        for (int m = 0; m < n; m++) wf[m] = std::complex<float>((float)w[m].real(), (float)w[m].imag());
    }
    // full r2c of a real N^3 grid g[k][j][i] -> G[k][j][i'], i' < Nh  (FFTW forward, unnormalised)
    void r2c(const double* g, cplx* G, int nz_lim = -1) const { r2c_t<double>(g, G, nz_lim, w); }
    // the same in float32 arithmetic (inputs and twiddles rounded to float32)
    void r2c_f32(const double* g, cplx* G, int nz_lim = -1) const { r2c_t<float>(g, G, nz_lim, wf); }
    // c2r (FFTW backward, unnormalised) evaluated only at i,j,k in [lo, N)
    void c2r(const cplx* G, double* g, int lo = 0) const { c2r_t<double>(G, g, lo, w); }
    void c2r_f32(const cplx* G, double* g, int lo = 0) const { c2r_t<float>(G, g, lo, wf); }

    template <class T>
    void r2c_t(const double* g, cplx* G, int nz_lim, const std::vector<std::complex<T>>& tw) const {
        typedef std::complex<T> ct;
        int lim = nz_lim < 0 ? N : nz_lim;  // inputs nonzero only for i,j,k < lim
        std::vector<ct> t1((size_t)N * N * Nh, ct(0, 0)), t2((size_t)N * N * Nh, ct(0, 0));
        for (int k = 0; k < lim; k++)
            for (int j = 0; j < lim; j++)
                for (int f = 0; f < Nh; f++) {
                    ct a(0, 0);
                    for (int i = 0; i < lim; i++) a += (T)g[((size_t)k * N + j) * N + i] * tw[(i * f) % N];
                    t1[((size_t)k * N + j) * Nh + f] = a;
                }
        for (int k = 0; k < lim; k++)
            for (int fj = 0; fj < N; fj++)
                for (int f = 0; f < Nh; f++) {
                    ct a(0, 0);
                    for (int j = 0; j < lim; j++) a += t1[((size_t)k * N + j) * Nh + f] * tw[(j * fj) % N];
                    t2[((size_t)k * N + fj) * Nh + f] = a;
                }
        for (int fk = 0; fk < N; fk++)
            for (int fj = 0; fj < N; fj++)
                for (int f = 0; f < Nh; f++) {
                    ct a(0, 0);
                    for (int k = 0; k < lim; k++) a += t2[((size_t)k * N + fj) * Nh + f] * tw[(k * fk) % N];
                    G[((size_t)fk * N + fj) * Nh + f] = cplx(a.real(), a.imag());
                }
    }
    template <class T>
    void c2r_t(const cplx* G, double* g, int lo, const std::vector<std::complex<T>>& tw) const {
        typedef std::complex<T> ct;
        std::vector<ct> t1((size_t)N * N * Nh), t2((size_t)N * N * Nh);
        for (int k = lo; k < N; k++)
            for (int fj = 0; fj < N; fj++)
                for (int f = 0; f < Nh; f++) {
                    ct a(0, 0);
                    for (int fk = 0; fk < N; fk++) {
                        const cplx& v = G[((size_t)fk * N + fj) * Nh + f];
                        a += ct((T)v.real(), (T)v.imag()) * std::conj(tw[(k * fk) % N]);
                    }
                    t1[((size_t)k * N + fj) * Nh + f] = a;
                }
        for (int k = lo; k < N; k++)
            for (int j = lo; j < N; j++)
                for (int f = 0; f < Nh; f++) {
                    ct a(0, 0);
                    for (int fj = 0; fj < N; fj++) a += t1[((size_t)k * N + fj) * Nh + f] * std::conj(tw[(j * fj) % N]);
                    t2[((size_t)k * N + j) * Nh + f] = a;
                }
        for (int k = lo; k < N; k++)
            for (int j = lo; j < N; j++)
                for (int i = lo; i < N; i++) {
                    const ct* row = &t2[((size_t)k * N + j) * Nh];
                    T a = row[0].real();
                    for (int f = 1; f < Nh; f++) {
                        if (2 * f == N) { a += (row[f] * std::conj(tw[(i * f) % N])).real(); continue; }
                        a += 2 * (row[f] * std::conj(tw[(i * f) % N])).real();
                    }
                    g[((size_t)k * N + j) * N + i] = a;
                }
    }
};

// kernel.h:136-141
static inline int rel_hash(int x, int y, int z) {
    int signs = ((x < 0 ? 1 : 0) << 2) + ((y < 0 ? 1 : 0) << 1) + (z < 0 ? 1 : 0);
    return std::abs(x) * 1000 + std::abs(y) * 100 + std::abs(z) * 10 + signs;
}
// kernel.h:149-168 / 176-194: z outer, x inner over [-range,range]^3 with some |c| >= minv
static std::vector<std::array<int, 3>> relx_list(int range, int minv) {
    std::vector<std::array<int, 3>> out;
    for (int k = -range; k <= range; k++)
        for (int j = -range; j <= range; j++)
            for (int i = -range; i <= range; i++)
                if (std::abs(i) >= minv || std::abs(j) >= minv || std::abs(k) >= minv) out.push_back({i, j, k});
    return out;
}

struct Plan {
    int P, ncrit, images, dipole, div4pi, nthreads;
    double cycle, r0, x0[3];
    int S = 0, N = 0, N3 = 0, Nh = 0, F = 0;
    int m2l_bruteforce = 0;
    // Reduced-precision emulation of the operator applies (no counterpart in the reference; models
    // the error of a mixed-precision evaluation): emul_mask selects the stages (EMUL_* bits),
    // emul_kind the arithmetic (EmulKind), pinv_rank > 0 applies the P2M and L2P pseudo-inverses
    // at that rank (leading rows of UT / columns of VSinv), as the mixed path does.
    int emul_mask = 0, emul_kind = 0, pinv_rank = 0;
    // emul_p2p (EmulP2P): the near field in float32 in one of three forms, see stage_p2p_emul.
    int emul_p2p = 0;
    std::string err;
    // bodies in tree order
    int n = 0;
    std::vector<int> bidx;
    std::vector<double> bx, by, bz;
    // cells and lists
    std::vector<Cell> cells;
    int ntree = 0, lost_pairs = 0;
    std::vector<std::vector<P2PEntry>> p2p;   // indexed by target cell
    std::vector<std::vector<M2LEntry>> m2l;   // indexed by target parent cell
    // operators
    int have_ops = 0;
    std::vector<double> UT, VSinv, UT_l2p, VSinv_l2p, m2m[8], m2m_img[27], l2l[8], l2l_img;
    std::vector<double> cc8, cc27;  // [26][F][64][2], [26][F][28][2]
    std::vector<int> up_idx, down_idx;
    // per-cell vectors
    std::vector<double> qeq, pck;
    // charges and body fields in tree order (kept between the per-stage calls)
    std::vector<double> q, p, gx, gy, gz;
    std::map<std::string, double> timers;
};

// y = k * A(m x n) x with A's leading dimension lda, in float64 or (stage bit set) the emulated arithmetic.
static void gemv_stage(const Plan& pl, int bit, const double* A, int lda, const double* x, double* y, int m, int n,
                       double k) {
    if (pl.emul_mask & bit) { gemv_emul(A, lda, x, y, m, n, k, pl.emul_kind); return; }
    if (lda == n) { gemv(A, x, y, m, n, k); return; }
    for (int i = 0; i < m; i++) {
        const double* a = A + (size_t)i * lda;
        double s = 0;
        for (int j = 0; j < n; j++) s += a[j] * x[j];
        y[i] = k * s;
    }
}

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------- tree: tree.cpp:102-188 (adaptive BFS octree, stable 8-way partition) -------------
static int build_tree(Plan& pl, std::vector<int>& perm, const double* X, const double* Y, const double* Z) {
    int n = pl.n;
    pl.cells.clear();
    Cell root;
    root.idx = 0; root.octant = 13; root.depth = 0; root.r = pl.r0;
    root.x[0] = pl.x0[0]; root.x[1] = pl.x0[1]; root.x[2] = pl.x0[2];
    root.coff = 0; root.cnum = 0; root.boff = 0; root.bnum = n;
    pl.cells.push_back(root);
    perm.resize(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    std::vector<int> buf;
    std::queue<int> q;
    q.push(0);
    while (!q.empty()) {
        int ci = q.front(); q.pop();
        Cell b = pl.cells[ci];
        if (b.bnum > pl.ncrit) {
            if (b.depth >= 60) { pl.err = "tree depth exceeds 60 (more than ncrit coincident bodies?)"; return -1; }
            int begin = b.boff, num = b.bnum;
            int cnt[8] = {0}, off[8] = {0};
            auto oct = [&](int bi) {
                return ((X[bi] > b.x[0]) << 2) + ((Y[bi] > b.x[1]) << 1) + ((Z[bi] > b.x[2]) << 0);
            };
            for (int i = begin; i < begin + num; i++) cnt[oct(perm[i])]++;
            int sum = 0, nchild = 0;
            for (int o = 0; o < 8; o++) { off[o] = sum; sum += cnt[o]; if (cnt[o] > 0) nchild++; }
            buf.resize(num);
            for (int i = begin; i < begin + num; i++) { int o = oct(perm[i]); buf[off[o]++] = perm[i]; }
            for (int i = 0; i < num; i++) perm[begin + i] = buf[i];
            int insert = (int)pl.cells.size();
            pl.cells[ci].coff = insert;
            pl.cells[ci].cnum = nchild;
            for (int o = 0; o < 8; o++) {
                if (cnt[o] == 0) continue;
                Cell c;
                c.idx = (int)pl.cells.size(); c.octant = o; c.depth = b.depth + 1; c.r = b.r / 2;
                c.coff = 0; c.cnum = 0; c.boff = begin + off[o] - cnt[o]; c.bnum = cnt[o];
                child_center(b.x, b.r, o, 0, c.x);
                pl.cells.push_back(c);
                q.push(c.idx);
            }
        }
    }
    return 0;
}

// ---------------- traverser: traverser.cpp:19-331 -------------
struct Trav {
    Plan& pl;
    explicit Trav(Plan& p) : pl(p) {}
    // traverser.cpp:314-321
    bool adjacent(int a, int b, const double off[3]) const {
        const Cell& ca = pl.cells[a];
        const Cell& cb = pl.cells[b];
        double dist = (ca.r + cb.r) * 1.001;
        for (int d = 0; d < 3; d++) {
            double dx = std::fabs(ca.x[d] - cb.x[d] - off[d]);
            if (!(dx <= dist)) return false;
        }
        return true;
    }
    bool neighbour(int a, int b, const double off[3]) const { return adjacent(a, b, off) && pl.cells[a].depth == pl.cells[b].depth; }
    bool leaf(int c) const { return pl.cells[c].cnum == 0; }
    void add_p2p(int tc, int sc, const double off[3]) {
        P2PEntry e; e.src = sc; e.off[0] = off[0]; e.off[1] = off[1]; e.off[2] = off[2];
        pl.p2p[tc].push_back(e);
    }
    // traverser.cpp:70-148
    void horizontal(int tc, int sc, int tcp, int scp, const double off[3]) {
        int divide = 0;
        if (!adjacent(tc, sc, off)) {
            if (adjacent(tc, scp, off) && leaf(tc) && pl.cells[scp].depth >= pl.cells[tc].depth) add_p2p(tc, sc, off);
            else if (adjacent(tcp, sc, off) && leaf(sc) && pl.cells[tcp].depth >= pl.cells[sc].depth) add_p2p(tc, sc, off);
            else if (neighbour(tcp, scp, off)) { /* M2L pair: served by the parent map */ }
            else {
                if (leaf(tc) && leaf(sc)) pl.lost_pairs++;
                else if (!leaf(tc) && leaf(sc)) divide = 1;
                else if (leaf(tc) && !leaf(sc)) divide = 2;
                else divide = pl.cells[tc].r >= pl.cells[sc].r ? 1 : 2;
            }
        } else {
            if (leaf(tc) && leaf(sc)) add_p2p(tc, sc, off);
            else if (!leaf(tc) && leaf(sc)) divide = 1;
            else if (leaf(tc) && !leaf(sc)) divide = 2;
            else divide = pl.cells[tc].r >= pl.cells[sc].r ? 1 : 2;
        }
        if (divide == 1) {
            Cell c = pl.cells[tc];
            for (int i = 0; i < c.cnum; i++) horizontal(c.coff + i, sc, tc, scp, off);
        } else if (divide == 2) {
            Cell c = pl.cells[sc];
            for (int i = 0; i < c.cnum; i++) horizontal(tc, c.coff + i, tcp, sc, off);
        }
    }
    void push_m2l(int tar, int src, const double off[3], int img) {
        M2LEntry e; e.src = src; e.off[0] = off[0]; e.off[1] = off[1]; e.off[2] = off[2]; e.img = img;
        pl.m2l[tar].push_back(e);
    }
    // traverser.cpp:234-282 restated.  Brute force O(C^2 * 27) like the reference, or an
    // equivalent same-depth grid lookup (integer cell coordinates + the same FP adjacency test).
    void parent_map(bool periodic) {
        int C = (int)pl.cells.size();
        const double zero[3] = {0, 0, 0};
        std::vector<std::vector<M2LEntry>> np(C), pp(C);
        if (pl.m2l_bruteforce) {
#pragma omp parallel for schedule(dynamic, 16)
            for (int j = 0; j < C; j++) {
                for (int i = 0; i < C; i++)
                    if (i != j && !leaf(j) && !leaf(i) && neighbour(i, j, zero)) {
                        M2LEntry e; e.src = i; e.off[0] = e.off[1] = e.off[2] = 0; e.img = 0; np[j].push_back(e);
                    }
                if (!periodic || pl.cells[j].depth == -1) continue;
                for (int i = 0; i < C; i++)
                    for (int pz = -1; pz <= 1; pz++)
                        for (int py = -1; py <= 1; py++)
                            for (int px = -1; px <= 1; px++) {
                                if (!(px || py || pz)) continue;
                                double off[3] = {px * pl.cycle, py * pl.cycle, pz * pl.cycle};
                                if (!leaf(j) && !leaf(i) && neighbour(j, i, off)) {
                                    M2LEntry e; e.src = i; e.off[0] = off[0]; e.off[1] = off[1]; e.off[2] = off[2]; e.img = 0;
                                    pp[j].push_back(e);
                                }
                            }
            }
        } else {
            // group non-leaf tree cells by depth, keyed by integer grid coordinates
            int maxd = 0;
            for (int c = 0; c < pl.ntree; c++) maxd = std::max(maxd, pl.cells[c].depth);
            std::vector<std::unordered_map<long long, int>> grid(maxd + 1);
            std::vector<std::array<long long, 3>> coord(C);
            auto key3 = [](long long a, long long b, long long c) { return (a << 42) ^ (b << 21) ^ c; };
            for (int c = 0; c < pl.ntree; c++) {
                const Cell& cc = pl.cells[c];
                double w = 2 * cc.r;
                for (int d = 0; d < 3; d++) coord[c][d] = std::llround((cc.x[d] - (pl.x0[d] - pl.r0)) / w - 0.5);
                if (!leaf(c)) grid[cc.depth][key3(coord[c][0], coord[c][1], coord[c][2])] = c;
            }
#pragma omp parallel for schedule(dynamic, 16)
            for (int j = 0; j < pl.ntree; j++) {
                if (leaf(j)) continue;
                int d = pl.cells[j].depth;
                long long nside = 1LL << d;
                std::set<int> cand;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            long long a = coord[j][0] + dx, b = coord[j][1] + dy, c = coord[j][2] + dz;
                            if (periodic) { a = ((a % nside) + nside) % nside; b = ((b % nside) + nside) % nside; c = ((c % nside) + nside) % nside; }
                            else if (a < 0 || b < 0 || c < 0 || a >= nside || b >= nside || c >= nside) continue;
                            auto it = grid[d].find(key3(a, b, c));
                            if (it != grid[d].end()) cand.insert(it->second);
                        }
                for (int i : cand) {
                    if (i != j && neighbour(i, j, zero)) {
                        M2LEntry e; e.src = i; e.off[0] = e.off[1] = e.off[2] = 0; e.img = 0; np[j].push_back(e);
                    }
                }
                if (!periodic) continue;
                for (int i : cand)
                    for (int pz = -1; pz <= 1; pz++)
                        for (int py = -1; py <= 1; py++)
                            for (int px = -1; px <= 1; px++) {
                                if (!(px || py || pz)) continue;
                                double off[3] = {px * pl.cycle, py * pl.cycle, pz * pl.cycle};
                                if (neighbour(j, i, off)) {
                                    M2LEntry e; e.src = i; e.off[0] = off[0]; e.off[1] = off[1]; e.off[2] = off[2]; e.img = 0;
                                    pp[j].push_back(e);
                                }
                            }
            }
        }
        for (int j = 0; j < C; j++) {
            for (auto& e : np[j]) pl.m2l[j].push_back(e);
            for (auto& e : pp[j]) pl.m2l[j].push_back(e);
        }
    }
    // traverser.cpp:19-68
    void traverse() {
        const double zero[3] = {0, 0, 0};
        pl.p2p.assign(pl.cells.size() + pl.images + 1, {});
        pl.m2l.assign(pl.cells.size() + pl.images + 1, {});
        if (pl.images == 0) {
            horizontal(0, 0, 0, 0, zero);
            parent_map(false);
            return;
        }
        // horizontal_periodic_near (traverser.cpp:150-176)
        Cell c;
        c.idx = (int)pl.cells.size(); c.depth = -1; c.octant = 13; c.r = pl.cells[0].r * 3;
        c.x[0] = pl.cells[0].x[0]; c.x[1] = pl.cells[0].x[1]; c.x[2] = pl.cells[0].x[2];
        c.coff = 0; c.cnum = 1; c.boff = 0; c.bnum = pl.cells[0].bnum;
        pl.cells.push_back(c);
        int cidx = c.idx;
        for (int pz = -1; pz <= 1; pz++)
            for (int py = -1; py <= 1; py++)
                for (int px = -1; px <= 1; px++) {
                    double off[3] = {px * pl.cycle, py * pl.cycle, pz * pl.cycle};
                    horizontal(0, 0, cidx, cidx, off);
                }
        parent_map(true);  // make_M2L_parent_map_i1
        if (pl.images >= 2) {
            // horizontal_periodic_far (traverser.cpp:178-232): image cells c(-2..-images)
            int child = (int)pl.cells.size() - 1;
            for (int m = 1; m < pl.images; m++) {
                Cell ci;
                ci.idx = (int)pl.cells.size(); ci.depth = -1 - m; ci.octant = 13; ci.r = pl.cells[child].r * 3;
                ci.x[0] = pl.cells[child].x[0]; ci.x[1] = pl.cells[child].x[1]; ci.x[2] = pl.cells[child].x[2];
                ci.coff = child; ci.cnum = 1; ci.boff = 0; ci.bnum = 0;
                pl.cells.push_back(ci);
                child = ci.idx;
            }
            // 1x27 image parent entries (traverser.cpp:43-64)
            int idx = (int)pl.cells.size() - 1 - pl.images;
            for (int m = 1; m < pl.images; m++) {
                int ci = idx + m;
                const Cell& cc = pl.cells[ci];
                for (int pz = -1; pz <= 1; pz++)
                    for (int py = -1; py <= 1; py++)
                        for (int px = -1; px <= 1; px++)
                            if (px || py || pz) {
                                double off[3] = {px * cc.r * 2, py * cc.r * 2, pz * cc.r * 2};
                                push_m2l(ci, ci, off, 1);
                            }
            }
        }
    }
};

// ---------------- precompute: kernel.cpp:1638-1704 and 1706-1850 -------------
static void precompute(Plan& pl) {
    int P = pl.P, S = pl.S;
    double r0 = pl.r0;
    const double o[3] = {0, 0, 0};
    std::vector<double> chk, eqv, A;
    surface_points(P, r0 * 2.95, o, chk);
    surface_points(P, r0 * 1.05, o, eqv);
    kernel_matrix(eqv, chk, A);  // rows: check points, cols: equivalent points
    std::vector<double> U, s, V;
    svd_jacobi(A, S, U, s, V);
    // mathfunc.cpp:154-169 pseudo-inverse cutoff
    double smax = 0;
    for (double v : s) smax = std::max(smax, std::fabs(v));
    std::vector<double> sinv(S);
    for (int i = 0; i < S; i++) sinv[i] = s[i] > smax * 4e-16 ? 1.0 / s[i] : 0.0;
    transpose(U, S, S, pl.UT);
    pl.VSinv.assign((size_t)S * S, 0.0);
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) pl.VSinv[(size_t)i * S + j] = V[(size_t)i * S + j] * sinv[j];
    // M2M (kernel.cpp:1660-1668) and image M2M (1671-1683)
#pragma omp parallel for
    for (int oc = 0; oc < 8 + 27; oc++) {
        double cx[3];
        std::vector<double> ce, K, m1;
        if (oc < 8) {
            child_center(o, r0, oc, 0, cx);
            surface_points(P, r0 / 2 * 1.05, cx, ce);
        } else {
            child_center(o, r0, oc - 8, 1, cx);
            surface_points(P, r0 / 3 * 1.05, cx, ce);
        }
        kernel_matrix(ce, chk, K);
        matmul(pl.UT, K, S, S, S, m1);
        if (oc < 8) matmul(pl.VSinv, m1, S, S, S, pl.m2m[oc]);
        else matmul(pl.VSinv, m1, S, S, S, pl.m2m_img[oc - 8]);
    }
    for (int oc = 0; oc < 8; oc++) transpose(pl.m2m[oc], S, S, pl.l2l[oc]);    // 1686-1690
    transpose(pl.m2m_img[13], S, S, pl.l2l_img);                                // 1693-1696
    // L2P factors (1699-1703): UT_l2p = V^T, VSinv_l2p = U * Sinv
    transpose(V, S, S, pl.UT_l2p);
    pl.VSinv_l2p.assign((size_t)S * S, 0.0);
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) pl.VSinv_l2p[(size_t)i * S + j] = U[(size_t)i * S + j] * sinv[j];

    // precompute_m2l: Gk of 1/r on the conv grid for every |rel|_max in [2, range]
    int N = pl.N, F = pl.F;
    double delta = 2 * r0 / (P - 1) * 1.05, gsize = r0 * 2 * 1.05;   // surface.h:58-86
    int range = pl.images >= 2 ? 4 : 3;
    auto rels = relx_list(range, 2);
    std::map<int, int> relidx;
    for (size_t t = 0; t < rels.size(); t++) relidx[rel_hash(rels[t][0], rels[t][1], rels[t][2])] = (int)t;
    std::vector<std::vector<cplx>> Gk(rels.size(), std::vector<cplx>(F));
    Dft dft(N);
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < (int)rels.size(); t++) {
        std::vector<double> G((size_t)N * N * N);
        double rp[3] = {rels[t][0] * r0 * 2, rels[t][1] * r0 * 2, rels[t][2] * r0 * 2};
        for (int k = 0; k < N; k++)        // surface.cpp:49-67 + kernel.cpp:1598-1614
            for (int j = 0; j < N; j++)
                for (int i = 0; i < N; i++) {
                    double px = -gsize + i * delta - rp[0], py = -gsize + j * delta - rp[1], pz = -gsize + k * delta - rp[2];
                    double r = std::sqrt(px * px + py * py + pz * pz);
                    G[((size_t)k * N + j) * N + i] = r == 0 ? 0 : 1 / r;
                }
        dft.r2c(G.data(), Gk[t].data());
    }
    auto pars = relx_list(1, 1);  // 26 parent patterns
    pl.cc8.assign((size_t)26 * F * 128, 0.0);
    for (int pi = 0; pi < 26; pi++) {
        double parx[3] = {pars[pi][0] * 2.0, pars[pi][1] * 2.0, pars[pi][2] * 2.0};
        for (int is = 0; is < 8; is++) {
            double xs[3], xt[3];
            child_center(parx, 1, is, 0, xs);
            for (int it = 0; it < 8; it++) {
                child_center(o, 1, it, 0, xt);
                int d0 = (int)std::round(xs[0] - xt[0]), d1 = (int)std::round(xs[1] - xt[1]), d2 = (int)std::round(xs[2] - xt[2]);
                auto f = relidx.find(rel_hash(d0, d1, d2));
                if (f == relidx.end()) continue;
                for (int fq = 0; fq < F; fq++) {
                    size_t id = ((size_t)pi * F + fq) * 128 + (is * 8 + it) * 2;
                    pl.cc8[id] = Gk[f->second][fq].real();
                    pl.cc8[id + 1] = Gk[f->second][fq].imag();
                }
            }
        }
    }
    pl.cc27.assign((size_t)26 * F * 56, 0.0);
    if (pl.images >= 2) {
        for (int pi = 0; pi < 26; pi++) {
            int cnt = 0;
            for (int pz = -1; pz <= 1; pz++)
                for (int py = -1; py <= 1; py++)
                    for (int px = -1; px <= 1; px++) {
                        int d0 = pars[pi][0] * 3 + px, d1 = pars[pi][1] * 3 + py, d2 = pars[pi][2] * 3 + pz;
                        auto f = relidx.find(rel_hash(d0, d1, d2));
                        for (int fq = 0; fq < F; fq++) {
                            size_t id = ((size_t)pi * F + fq) * 56 + cnt * 2;
                            pl.cc27[id] = Gk[f->second][fq].real();
                            pl.cc27[id + 1] = Gk[f->second][fq].imag();
                        }
                        cnt++;
                    }
        }
    }
    pl.have_ops = 1;
}

static void setup_grid(Plan& pl) {
    pl.S = surf_num(pl.P);
    pl.N = 2 * pl.P - 1;
    pl.N3 = pl.N * pl.N * pl.N;
    pl.Nh = pl.N / 2 + 1;
    pl.F = pl.N * pl.N * pl.Nh;
    std::vector<int> ijk;
    surface_lattice(pl.P, ijk);
    pl.up_idx.resize(pl.S);
    pl.down_idx.resize(pl.S);
    int N = pl.N, P = pl.P;
    for (int s = 0; s < pl.S; s++) {   // kernel.cpp:1246-1259
        int i = ijk[3 * s], j = ijk[3 * s + 1], k = ijk[3 * s + 2];
        pl.up_idx[s] = k * N * N + j * N + i;
        pl.down_idx[s] = (k + P - 1) * N * N + (j + P - 1) * N + (i + P - 1);
    }
}

// ---------------- FMM stages -------------
// kernel.cpp:332-371
static void stage_p2m(Plan& pl, const std::vector<int>& leaves, const double* q) {
    int S = pl.S;
#pragma omp parallel for schedule(dynamic, 8)
    for (int li = 0; li < (int)leaves.size(); li++) {
        const Cell& c = pl.cells[leaves[li]];
        std::vector<double> chk, pc(S, 0.0), buf(S);
        surface_points(pl.P, c.r * 2.95, c.x, chk);
        for (int s = 0; s < S; s++) {
            double acc = 0;
            if (pl.emul_mask & EMUL_P2M_CHECK) {   // float32 cell-relative differences and sums
                float af = 0;
                for (int b = c.boff; b < c.boff + c.bnum; b++) {
                    float dx = (float)(chk[3 * s] - c.x[0]) - (float)(pl.bx[b] - c.x[0]);
                    float dy = (float)(chk[3 * s + 1] - c.x[1]) - (float)(pl.by[b] - c.x[1]);
                    float dz = (float)(chk[3 * s + 2] - c.x[2]) - (float)(pl.bz[b] - c.x[2]);
                    float r2 = dx * dx + dy * dy + dz * dz;
                    af += r2 > 0 ? (float)q[b] * (1.0f / std::sqrt(r2)) : 0.0f;
                }
                acc = af;
            } else {
                for (int b = c.boff; b < c.boff + c.bnum; b++) {
                    double dx = chk[3 * s] - pl.bx[b], dy = chk[3 * s + 1] - pl.by[b], dz = chk[3 * s + 2] - pl.bz[b];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    acc += r2 > 0 ? q[b] * (1.0 / std::sqrt(r2)) : 0.0;
                }
            }
            pc[s] = acc;
        }
        double scale = std::pow(0.5, c.depth);
        const int r = pl.pinv_rank > 0 ? std::min(pl.pinv_rank, S) : S;
        gemv_stage(pl, EMUL_P2M_PINV, pl.UT.data(), S, pc.data(), buf.data(), r, S, 1.0);
        gemv_stage(pl, EMUL_P2M_PINV, pl.VSinv.data(), S, buf.data(), &pl.qeq[(size_t)c.idx * S], S, r, scale);
        if (pl.emul_mask & EMUL_MONOPOLE) {
            double qc = 0;
            for (int b = c.boff; b < c.boff + c.bnum; b++) qc += q[b];
            set_monopole(&pl.qeq[(size_t)c.idx * S], S, qc);
        }
    }
}

// fmm.cpp:111-150 with kernel.cpp:395-404 and 434-442.
// qtot != null selects the infinite-lattice gauge (dipole mode 2, no counterpart in the reference):
// the root's and every image cell's equivalent density are shifted uniformly so that their sum (the
// cell's monopole) equals the cell's charge, 27^m *qtot for the image cell c(-m).  The
// pseudo-inverse leaves a monopole that differs from the charge by ~1e-10 of the density's size;
// each image level copies it 27 times and the 1x27 lattice M2L turns it into a potential offset
// growing 9x per image level.
static void stage_m2m(Plan& pl, int mind, int maxd, const double* qtot = nullptr) {
    int S = pl.S;
    for (int d = maxd; d > mind; d--) {
        std::vector<int> cs;
        for (auto& c : pl.cells) if (c.depth == d && c.cnum > 0) cs.push_back(c.idx);
#pragma omp parallel for schedule(dynamic, 4)
        for (int t = 0; t < (int)cs.size(); t++) {
            const Cell& c = pl.cells[cs[t]];
            std::vector<double> tmp(S);
            double* qp = &pl.qeq[(size_t)c.idx * S];
            if (d >= 0) {
                for (int ch = 0; ch < c.cnum; ch++) {
                    const Cell& cc = pl.cells[c.coff + ch];
                    gemv_stage(pl, EMUL_M2M, pl.m2m[cc.octant].data(), S, &pl.qeq[(size_t)cc.idx * S], tmp.data(), S, S, 1.0);
                    for (int s = 0; s < S; s++) qp[s] += tmp[s];
                }
                if (pl.emul_mask & EMUL_MONOPOLE) {
                    double qcell = 0;
                    for (int b = c.boff; b < c.boff + c.bnum; b++) qcell += pl.q[b];
                    set_monopole(qp, S, qcell);
                }
            } else {
                const double* qc = &pl.qeq[(size_t)c.coff * S];
                for (int o = 0; o < 27; o++) {
                    gemv(pl.m2m_img[o].data(), qc, tmp.data(), S, S, 1.0);
                    for (int s = 0; s < S; s++) qp[s] += tmp[s];
                }
            }
            if (qtot && d <= 0) {
                double sum = 0;
                for (int s = 0; s < S; s++) sum += qp[s];
                const double shift = (*qtot * std::pow(27.0, -d) - sum) / S;
                for (int s = 0; s < S; s++) qp[s] += shift;
            }
        }
    }
}

// kernel.cpp:1140-1397 (m2l_fft_precompute_t) with hadamard_8x8 (1012-1082) and hadamard_1x27 (1084-1138)
// F32: the tree parents' M2L in float32 (EMUL_M2L); the float64 form is the restatement.
template <bool F32>
static void stage_m2l_t(Plan& pl) {
    int S = pl.S, N = pl.N, N3 = pl.N3, F = pl.F, P = pl.P;
    double ts0 = now();
    std::vector<int> tars;
    std::set<int> srcset;
    for (int t = 0; t < (int)pl.m2l.size(); t++) {
        if (pl.m2l[t].empty()) continue;
        tars.push_back(t);
        for (auto& e : pl.m2l[t]) srcset.insert(e.src);
    }
    std::vector<int> srcs(srcset.begin(), srcset.end());
    std::unordered_map<int, int> srcloc;
    for (int i = 0; i < (int)srcs.size(); i++) srcloc[srcs[i]] = i;
    auto pars = relx_list(1, 1);
    std::map<int, int> par_idx;
    for (int i = 0; i < 26; i++) par_idx[rel_hash(pars[i][0], pars[i][1], pars[i][2])] = i;
    // per target: pattern -> (src local, is_image)   (kernel.cpp:1189-1211)
    std::vector<std::array<int, 26>> slot_src(tars.size()), slot_img(tars.size());
    for (size_t t = 0; t < tars.size(); t++) {
        slot_src[t].fill(-1); slot_img[t].fill(-1);
        const Cell& tc = pl.cells[tars[t]];
        for (auto& e : pl.m2l[tars[t]]) {
            const Cell& sc = pl.cells[e.src];
            int rv[3];
            for (int d = 0; d < 3; d++) rv[d] = (int)std::round((sc.x[d] + e.off[d] - tc.x[d]) / (tc.r * 2));
            int pi = par_idx.at(rel_hash(rv[0], rv[1], rv[2]));
            slot_src[t][pi] = srcloc[e.src];
            slot_img[t][pi] = e.img;
        }
    }
    size_t fsz = (size_t)F * 8;  // complex entries per parent
    double ts1 = now();
    // like the reference (kernel.cpp:1263-1283): Qk is fully written by the up pass, Pk is zeroed in parallel
    struct Free { void operator()(cplx* p) const { std::free(p); } };
    std::unique_ptr<cplx[], Free> Qk_mem(static_cast<cplx*>(std::malloc(srcs.size() * fsz * sizeof(cplx) + 64)));
    std::unique_ptr<cplx[], Free> Pk_mem(static_cast<cplx*>(std::malloc(tars.size() * fsz * sizeof(cplx) + 64)));
    cplx* Qk = Qk_mem.get();
    cplx* Pk = Pk_mem.get();
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)tars.size(); t++)
        for (size_t e = 0; e < fsz; e++) Pk[t * fsz + e] = cplx(0, 0);
    Dft dft(N);
    double t0 = now();
    pl.timers["m2l_setup"] = ts1 - ts0;
    pl.timers["m2l_alloc"] = t0 - ts1;
    // up (kernel.cpp:1286-1328)
#pragma omp parallel for schedule(dynamic, 2)
    for (int si = 0; si < (int)srcs.size(); si++) {
        const Cell& cp = pl.cells[srcs[si]];
        std::vector<double> grid(N3);
        std::vector<cplx> out(F);
        cplx* dst = &Qk[si * fsz];
        for (int slot = 0; slot < 8; slot++) for (int f = 0; f < F; f++) dst[(size_t)f * 8 + slot] = 0;
        for (int ch = 0; ch < cp.cnum; ch++) {
            const Cell& cc = pl.cells[cp.coff + ch];
            int slot = cc.octant == 13 ? 0 : cc.octant;
            std::fill(grid.begin(), grid.end(), 0.0);
            for (int s = 0; s < S; s++) grid[pl.up_idx[s]] = pl.qeq[(size_t)cc.idx * S + s];
            if constexpr (F32) {
                if (cp.depth >= 0) dft.r2c_f32(grid.data(), out.data(), P);
                else dft.r2c(grid.data(), out.data(), P);
            } else {
                dft.r2c(grid.data(), out.data(), P);
            }
            for (int f = 0; f < F; f++) dst[(size_t)f * 8 + slot] = out[f];
        }
    }
    double t1 = now();
    // hadamard 8x8 and 1x27 (frequency-parallel like the reference)
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; f++) {
        for (int pi = 0; pi < 26; pi++) {
            const double* g8 = &pl.cc8[((size_t)pi * F + f) * 128];
            for (size_t t = 0; t < tars.size(); t++) {
                int sl = slot_src[t][pi];
                if (sl < 0 || slot_img[t][pi] != 0) continue;
                const cplx* qk = &Qk[sl * fsz + (size_t)f * 8];
                cplx* pk = &Pk[t * fsz + (size_t)f * 8];
                if constexpr (F32) {
                    if (pl.cells[tars[t]].depth >= 0) {   // float32 kernel, spectra and sums
                        for (int it = 0; it < 8; it++) {
                            float ar = 0, ai = 0;
                            for (int is = 0; is < 8; is++) {
                                const double gr = g8[(is * 8 + it) * 2], gi = g8[(is * 8 + it) * 2 + 1];
                                const float qr = (float)qk[is].real(), qi = (float)qk[is].imag();
                                if (pl.emul_kind == EMUL_F32) {
                                    ar += (float)gr * qr - (float)gi * qi;
                                    ai += (float)gr * qi + (float)gi * qr;
                                } else {   // split bfloat16 products (matrix-core schemes), float32 sums
                                    ar += split_product(gr, qr, pl.emul_kind) - split_product(gi, qi, pl.emul_kind);
                                    ai += split_product(gr, qi, pl.emul_kind) + split_product(gi, qr, pl.emul_kind);
                                }
                            }
                            pk[it] = cplx((float)pk[it].real() + ar, (float)pk[it].imag() + ai);
                        }
                        continue;
                    }
                }
                for (int it = 0; it < 8; it++) {
                    double ar = 0, ai = 0;
                    for (int is = 0; is < 8; is++) {
                        double gr = g8[(is * 8 + it) * 2], gi = g8[(is * 8 + it) * 2 + 1];
                        double qr = qk[is].real(), qi = qk[is].imag();
                        ar += gr * qr - gi * qi;
                        ai += gr * qi + gi * qr;
                    }
                    pk[it] += cplx(ar, ai);
                }
            }
        }
        if (pl.images >= 2) {
            for (int pi = 0; pi < 26; pi++) {
                const double* g27 = &pl.cc27[((size_t)pi * F + f) * 56];
                double sr = 0, si = 0;
                for (int c = 0; c < 28; c++) { sr += g27[2 * c]; si += g27[2 * c + 1]; }
                for (size_t t = 0; t < tars.size(); t++) {
                    int sl = slot_src[t][pi];
                    if (sl < 0 || slot_img[t][pi] != 1) continue;
                    cplx qk = Qk[sl * fsz + (size_t)f * 8];
                    double ar = 0, ai = 0;
                    for (int c = 0; c < 28; c++) {
                        double gr = g27[2 * c], gi = g27[2 * c + 1];
                        ar += gr * qk.real() - gi * qk.imag();
                        ai += gr * qk.imag() + gi * qk.real();
                    }
                    Pk[t * fsz + (size_t)f * 8] += cplx(ar, ai);
                }
            }
        }
    }
    double t2 = now();
    // down (kernel.cpp:1343-1385) and store (1387-1396)
#pragma omp parallel for schedule(dynamic, 2)
    for (int ti = 0; ti < (int)tars.size(); ti++) {
        const Cell& ct = pl.cells[tars[ti]];
        std::vector<cplx> in(F);
        std::vector<double> g(N3);
        double scale = std::pow(ct.depth >= -1 ? 2.0 : 3.0, ct.depth + 1);
        for (int ch = 0; ch < ct.cnum; ch++) {
            const Cell& cc = pl.cells[ct.coff + ch];
            int slot = cc.octant == 13 ? 0 : cc.octant;
            for (int f = 0; f < F; f++) in[f] = Pk[ti * fsz + (size_t)f * 8 + slot];
            if constexpr (F32) {
                if (ct.depth >= 0) dft.c2r_f32(in.data(), g.data(), P - 1);
                else dft.c2r(in.data(), g.data(), P - 1);
            } else {
                dft.c2r(in.data(), g.data(), P - 1);
            }
            double* pc = &pl.pck[(size_t)cc.idx * S];
            for (int s = 0; s < S; s++) pc[s] += g[pl.down_idx[s]] / N3 * scale;
        }
    }
    double t3 = now();
    pl.timers["m2l_up"] = t1 - t0;
    pl.timers["m2l_hadamard"] = t2 - t1;
    pl.timers["m2l_down"] = t3 - t2;
}

static void stage_m2l(Plan& pl) {
    if (pl.emul_mask & EMUL_M2L) stage_m2l_t<true>(pl);
    else stage_m2l_t<false>(pl);
}

// fmm.cpp:195-229 with kernel.cpp:1419-1435
static void stage_l2l(Plan& pl, int mind, int maxd) {
    int S = pl.S;
    for (int d = mind; d < maxd; d++) {
        std::vector<int> cs;
        for (auto& c : pl.cells) if (c.depth == d && c.cnum > 0) cs.push_back(c.idx);
#pragma omp parallel for schedule(dynamic, 4)
        for (int t = 0; t < (int)cs.size(); t++) {
            const Cell& c = pl.cells[cs[t]];
            std::vector<double> tmp(S);
            const double* pp = &pl.pck[(size_t)c.idx * S];
            if (d >= 0) {
                for (int ch = 0; ch < c.cnum; ch++) {
                    const Cell& cc = pl.cells[c.coff + ch];
                    gemv_stage(pl, EMUL_L2L, pl.l2l[cc.octant].data(), S, pp, tmp.data(), S, S, 1.0);
                    double* pc = &pl.pck[(size_t)cc.idx * S];
                    for (int s = 0; s < S; s++) pc[s] += tmp[s];
                }
            } else {
                gemv(pl.l2l_img.data(), pp, tmp.data(), S, S, 1.0);
                double* pc = &pl.pck[(size_t)c.coff * S];
                for (int s = 0; s < S; s++) pc[s] += tmp[s];
            }
        }
    }
}

// kernel.cpp:1458-1511
static void stage_l2p(Plan& pl, const std::vector<int>& leaves, double* p, double* gx, double* gy, double* gz) {
    int S = pl.S;
#pragma omp parallel for schedule(dynamic, 8)
    for (int li = 0; li < (int)leaves.size(); li++) {
        const Cell& c = pl.cells[leaves[li]];
        std::vector<double> eq, buf(S), qd(S);
        surface_points(pl.P, c.r * 2.95, c.x, eq);
        double scale = std::pow(0.5, c.depth);
        const int r = pl.pinv_rank > 0 ? std::min(pl.pinv_rank, S) : S;
        gemv_stage(pl, EMUL_L2P_PINV, pl.UT_l2p.data(), S, &pl.pck[(size_t)c.idx * S], buf.data(), r, S, 1.0);
        gemv_stage(pl, EMUL_L2P_PINV, pl.VSinv_l2p.data(), S, buf.data(), qd.data(), S, r, scale);
        if (pl.emul_mask & EMUL_L2P_EVAL) {   // float32 cell-relative differences, products and sums
            for (int b = c.boff; b < c.boff + c.bnum; b++) {
                float ap = 0, ax = 0, ay = 0, az = 0;
                const float bxr = (float)(pl.bx[b] - c.x[0]), byr = (float)(pl.by[b] - c.x[1]),
                            bzr = (float)(pl.bz[b] - c.x[2]);
                for (int s = 0; s < S; s++) {
                    float dx = bxr - (float)(eq[3 * s] - c.x[0]), dy = byr - (float)(eq[3 * s + 1] - c.x[1]),
                          dz = bzr - (float)(eq[3 * s + 2] - c.x[2]);
                    float r2 = dx * dx + dy * dy + dz * dz;
                    float ir = r2 > 0 ? 1.0f / std::sqrt(r2) : 0.0f;
                    float qi = (float)qd[s] * ir;
                    ap += qi;
                    qi *= ir * ir;
                    ax += qi * dx; ay += qi * dy; az += qi * dz;
                }
                p[b] += ap; gx[b] -= ax; gy[b] -= ay; gz[b] -= az;
            }
            continue;
        }
        for (int b = c.boff; b < c.boff + c.bnum; b++) {
            double ap = 0, ax = 0, ay = 0, az = 0;
            for (int s = 0; s < S; s++) {
                double dx = pl.bx[b] - eq[3 * s], dy = pl.by[b] - eq[3 * s + 1], dz = pl.bz[b] - eq[3 * s + 2];
                double r2 = dx * dx + dy * dy + dz * dz;
                double ir = r2 > 0 ? 1.0 / std::sqrt(r2) : 0.0;
                double qi = qd[s] * ir;
                ap += qi;
                qi *= ir * ir;
                ax += qi * dx; ay += qi * dy; az += qi * dz;
            }
            p[b] += ap; gx[b] -= ax; gy[b] -= ay; gz[b] -= az;
        }
    }
}

// Float32 near field (no counterpart in the reference; models the mixed path's P2P).  Forms:
// EMUL_P2P_CELLREL: coordinates relative to the body's own cell centre rounded to float32, the
//   source cell's displacement (its centre + periodic offset - the target cell's centre) formed in
//   float64 and rounded once, charges rounded to float32, float32 arithmetic, and the pairs of
//   one (target, piece of 64 source bodies, lane group j mod G) summed in float32 before a float64
//   accumulator takes them (G = 64 / the target cell's slot count 8, 16, 32 or 64), as the GPU
//   kernel does for leaf cells;
// EMUL_P2P_CELLREL_F64SUM: the same values, every pair added straight into float64 (a perfectly
//   compensated summation);
// EMUL_P2P_DISP: the displacement x_t - x_s - offset formed in float64 and rounded to float32
//   (no coordinate rounding beyond the displacement's own), float32 charges and arithmetic,
//   float64 sums;
// EMUL_P2P_TWOFLOAT: every cell-relative coordinate and the cells' offset held as two float32
//   values (hi + lo); per (target, list entry) the target's position relative to the source cell is
//   formed with an exact float32 two-sum of the hi parts, per pair dx = (xo_hi - s_hi) + (xo_lo -
//   s_lo) in float32 (the hi difference of a close pair is exact), float32 arithmetic after it,
//   sums as in EMUL_P2P_CELLREL.
enum EmulP2P { EMUL_P2P_OFF = 0, EMUL_P2P_CELLREL = 1, EMUL_P2P_CELLREL_F64SUM = 2, EMUL_P2P_DISP = 3,
               EMUL_P2P_TWOFLOAT = 4 };

// a + b = s + e exactly, s = fl(a + b) (Knuth's two-sum in float32)
static inline void two_sum_f(float a, float b, float& s, float& e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

static void stage_p2p_emul(Plan& pl, const double* q, double* p, double* gx, double* gy, double* gz) {
    const int mode = pl.emul_p2p;
    int maxd = 0;
    for (auto& c : pl.cells) maxd = std::max(maxd, c.depth);
    for (int d = 0; d <= maxd; d++) {   // non-leaf targets share bodies with their leaves: a depth at a time
    std::vector<int> ts;
    for (int t = 0; t < (int)pl.p2p.size(); t++)
        if (!pl.p2p[t].empty() && pl.cells[t].depth == d) ts.push_back(t);
#pragma omp parallel for schedule(dynamic, 4)
    for (int ti = 0; ti < (int)ts.size(); ti++) {
        const int t = ts[ti];
        const auto& lst = pl.p2p[t];
        const Cell& ct = pl.cells[t];
        const int W = ct.bnum <= 8 ? 8 : ct.bnum <= 16 ? 16 : ct.bnum <= 32 ? 32 : 64, G = 64 / W;
        for (int b = ct.boff; b < ct.boff + ct.bnum; b++) {
            const float tx = (float)(pl.bx[b] - ct.x[0]), ty = (float)(pl.by[b] - ct.x[1]),
                        tz = (float)(pl.bz[b] - ct.x[2]);
            const float tlx = (float)(pl.bx[b] - ct.x[0] - tx), tly = (float)(pl.by[b] - ct.x[1] - ty),
                        tlz = (float)(pl.bz[b] - ct.x[2] - tz);
            double ap = 0, ax = 0, ay = 0, az = 0;
            for (auto& e : lst) {
                const Cell& cs = pl.cells[e.src];
                const float ox = (float)(cs.x[0] + e.off[0] - ct.x[0]), oy = (float)(cs.x[1] + e.off[1] - ct.x[1]),
                            oz = (float)(cs.x[2] + e.off[2] - ct.x[2]);
                const float xo = tx - ox, yo = ty - oy, zo = tz - oz;
                float xh = 0, yh = 0, zh = 0, xl = 0, yl = 0, zl = 0;   // EMUL_P2P_TWOFLOAT
                if (mode == EMUL_P2P_TWOFLOAT) {
                    const double odx = cs.x[0] + e.off[0] - ct.x[0], ody = cs.x[1] + e.off[1] - ct.x[1],
                                 odz = cs.x[2] + e.off[2] - ct.x[2];
                    const float olx = (float)(odx - ox), oly = (float)(ody - oy), olz = (float)(odz - oz);
                    float ex, ey, ez;
                    two_sum_f(tx, -ox, xh, ex);
                    two_sum_f(ty, -oy, yh, ey);
                    two_sum_f(tz, -oz, zh, ez);
                    xl = (tlx - olx) + ex;
                    yl = (tly - oly) + ey;
                    zl = (tlz - olz) + ez;
                }
                for (int k0 = 0; k0 < cs.bnum; k0 += 64) {
                    const int cnt = std::min(64, cs.bnum - k0);
                    for (int g = 0; g < G; g++) {
                        float fp = 0, fx = 0, fy = 0, fz = 0;
                        for (int j = g; j < cnt; j += G) {
                            const int s = cs.boff + k0 + j;
                            float dx, dy, dz;
                            if (mode == EMUL_P2P_DISP) {
                                dx = (float)(pl.bx[b] - (pl.bx[s] + e.off[0]));
                                dy = (float)(pl.by[b] - (pl.by[s] + e.off[1]));
                                dz = (float)(pl.bz[b] - (pl.bz[s] + e.off[2]));
                            } else if (mode == EMUL_P2P_TWOFLOAT) {
                                const double rx = pl.bx[s] - cs.x[0], ry = pl.by[s] - cs.x[1], rz = pl.bz[s] - cs.x[2];
                                const float shx = (float)rx, shy = (float)ry, shz = (float)rz;
                                const float slx = (float)(rx - shx), sly = (float)(ry - shy), slz = (float)(rz - shz);
                                dx = (xh - shx) + (xl - slx);
                                dy = (yh - shy) + (yl - sly);
                                dz = (zh - shz) + (zl - slz);
                            } else {
                                dx = xo - (float)(pl.bx[s] - cs.x[0]);
                                dy = yo - (float)(pl.by[s] - cs.x[1]);
                                dz = zo - (float)(pl.bz[s] - cs.x[2]);
                            }
                            const float r2 = dx * dx + dy * dy + dz * dz;
                            const float ir = r2 > 0 ? 1.0f / std::sqrt(r2) : 0.0f;
                            float qi = (float)q[s] * ir;
                            const float qi3 = qi * (ir * ir);
                            if (mode == EMUL_P2P_CELLREL || mode == EMUL_P2P_TWOFLOAT) {
                                fp += qi; fx += qi3 * dx; fy += qi3 * dy; fz += qi3 * dz;
                            } else {
                                ap += qi; ax += (double)(qi3 * dx); ay += (double)(qi3 * dy); az += (double)(qi3 * dz);
                            }
                        }
                        ap += fp; ax += fx; ay += fy; az += fz;
                    }
                }
            }
            p[b] += ap; gx[b] -= ax; gy[b] -= ay; gz[b] -= az;
        }
    }
    }
}

// fmm.cpp:259-294 with kernel.cpp:258-311 (targets may be non-leaf: process one depth at a time)
static void stage_p2p(Plan& pl, const double* q, double* p, double* gx, double* gy, double* gz) {
    if (pl.emul_p2p) { stage_p2p_emul(pl, q, p, gx, gy, gz); return; }
    int maxd = 0;
    for (auto& c : pl.cells) maxd = std::max(maxd, c.depth);
    for (int d = 0; d <= maxd; d++) {
        std::vector<int> ts;
        for (int t = 0; t < (int)pl.p2p.size(); t++)
            if (!pl.p2p[t].empty() && pl.cells[t].depth == d) ts.push_back(t);
#pragma omp parallel for schedule(dynamic, 4)
        for (int ti = 0; ti < (int)ts.size(); ti++) {
            const Cell& ct = pl.cells[ts[ti]];
            const auto& lst = pl.p2p[ts[ti]];
            for (int b = ct.boff; b < ct.boff + ct.bnum; b++) {
                double xb = pl.bx[b], yb = pl.by[b], zb = pl.bz[b];
                double ap = 0, ax = 0, ay = 0, az = 0;
                for (auto& e : lst) {
                    const Cell& cs = pl.cells[e.src];
#pragma omp simd reduction(+ : ap, ax, ay, az)
                    for (int s = cs.boff; s < cs.boff + cs.bnum; s++) {
                        double dx = xb - (pl.bx[s] + e.off[0]);
                        double dy = yb - (pl.by[s] + e.off[1]);
                        double dz = zb - (pl.bz[s] + e.off[2]);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        double ir = r2 > 0 ? 1.0 / std::sqrt(r2) : 0.0;
                        double qi = q[s] * ir;
                        ap += qi;
                        qi *= ir * ir;
                        ax += qi * dx; ay += qi * dy; az += qi * dz;
                    }
                }
                p[b] += ap; gx[b] -= ax; gy[b] -= ay; gz[b] -= az;
            }
        }
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

// body.cpp:98-129 (glibc srand48/drand48; x in [-r, r) + offset, q = U[0,1) - 0.5, optional zero net charge)
int orc_drand48_bodies(int n, double r, double ox, double oy, double oz, int seed, int zero_netcharge,
                       double* x, double* y, double* z, double* q) {
    srand48(seed);
    double qavg = 0;
    for (int i = 0; i < n; i++) {
        double v[3];
        for (int d = 0; d < 3; d++) v[d] = drand48() * r * 2 - r;
        double qq = drand48() - 0.5;
        qavg += qq;
        x[i] = v[0] + ox; y[i] = v[1] + oy; z[i] = v[2] + oz; q[i] = qq;
    }
    qavg /= n;
    if (zero_netcharge)
        for (int i = 0; i < n; i++) q[i] -= qavg;
    return 0;
}

void* orc_plan_create(int P, int ncrit, int images, double cycle, int dipole, int div4pi, int nthreads) {
    if (P < 2 || P > 32 || images < 0 || images > 20 || ncrit < 1 || !(cycle > 0)) return nullptr;
    Plan* pl = new Plan();
    pl->P = P; pl->ncrit = ncrit; pl->images = images; pl->cycle = cycle;
    pl->dipole = dipole; pl->div4pi = div4pi; pl->nthreads = nthreads;
    pl->r0 = cycle / 2;                      // argument.cpp:61-62
    pl->x0[0] = pl->x0[1] = pl->x0[2] = 0;
    setup_grid(*pl);
    return pl;
}

void orc_plan_destroy(void* h) { delete (Plan*)h; }

int orc_set_option(void* h, const char* name, int v) {
    Plan* pl = (Plan*)h;
    if (!strcmp(name, "m2l_bruteforce")) { pl->m2l_bruteforce = v; return 0; }
    if (!strcmp(name, "threads")) { pl->nthreads = v; return 0; }
    if (!strcmp(name, "emulate")) { pl->emul_mask = v; return 0; }
    if (!strcmp(name, "emulate_kind")) { pl->emul_kind = v; return v >= EMUL_F32 && v <= EMUL_BF16X6 ? 0 : -1; }
    if (!strcmp(name, "pinv_rank")) { pl->pinv_rank = v; return 0; }
    if (!strcmp(name, "emulate_p2p")) { pl->emul_p2p = v; return v >= EMUL_P2P_OFF && v <= EMUL_P2P_TWOFLOAT ? 0 : -1; }
    return -1;
}

const char* orc_last_error(void* h) { return ((Plan*)h)->err.c_str(); }

int orc_precompute(void* h) {
    Plan* pl = (Plan*)h;
    if (pl->nthreads > 0) omp_set_num_threads(pl->nthreads);
    double t0 = now();
    precompute(*pl);
    pl->timers["precompute"] = now() - t0;
    return 0;
}

// Builds the tree and the interaction lists (fmm.cpp:15-30). x,y,z in the reference frame: box
// centred at the origin with half size cycle/2.
int orc_set_particles(void* h, int n, const double* x, const double* y, const double* z) {
    Plan* pl = (Plan*)h;
    if (pl->nthreads > 0) omp_set_num_threads(pl->nthreads);
    double t0 = now();
    pl->n = n;
    std::vector<int> perm;
    if (build_tree(*pl, perm, x, y, z)) return -1;
    pl->ntree = (int)pl->cells.size();
    pl->bidx = perm;
    pl->bx.resize(n); pl->by.resize(n); pl->bz.resize(n);
    for (int i = 0; i < n; i++) { pl->bx[i] = x[perm[i]]; pl->by[i] = y[perm[i]]; pl->bz[i] = z[perm[i]]; }
    double t1 = now();
    pl->lost_pairs = 0;
    Trav tr(*pl);
    tr.traverse();
    pl->p2p.resize(pl->cells.size());
    pl->m2l.resize(pl->cells.size());
    pl->timers["build_tree"] = t1 - t0;
    pl->timers["traverse"] = now() - t1;
    return 0;
}

// LaplaceFMM::solve stage sequence (fmm.cpp:50-80).  Outputs in input order.
static void depth_range(const Plan& pl, int& mind, int& maxd) {
    mind = 0; maxd = 0;
    for (auto& c : pl.cells) { mind = std::min(mind, c.depth); maxd = std::max(maxd, c.depth); }
}

static std::vector<int> leaf_cells(const Plan& pl) {
    std::vector<int> leaves;
    for (auto& c : pl.cells) if (c.cnum == 0) leaves.push_back(c.idx);
    return leaves;
}

// dipole_correction (kernel.cpp:1580-1596) or the infinite-lattice gauge, scale_bodies (body.cpp:88-96),
// and input order (fmm.cpp:80), from the body fields in tree order.
static void finalize(Plan* pl, double* pout, double* gxo, double* gyo, double* gzo) {
    const int n = pl->n;
    std::vector<double> p = pl->p, gx = pl->gx, gy = pl->gy, gz = pl->gz;
    const std::vector<double>& q = pl->q;
    if (pl->dipole == 2) {   // infinite-lattice gauge (no counterpart in the reference)
        const double coef = 4 * M_PI / (3 * pl->cycle * pl->cycle * pl->cycle);
        double M[4] = {0, 0, 0, 0};
        for (int i = 0; i < n; i++) {
            M[0] += pl->bx[i] * q[i]; M[1] += pl->by[i] * q[i]; M[2] += pl->bz[i] * q[i];
            M[3] += (pl->bx[i] * pl->bx[i] + pl->by[i] * pl->by[i] + pl->bz[i] * pl->bz[i]) * q[i];
        }
        for (int i = 0; i < n; i++) {
            p[i] -= coef * (M[0] * pl->bx[i] + M[1] * pl->by[i] + M[2] * pl->bz[i]) - 0.5 * coef * M[3];
            gx[i] -= coef * M[0]; gy[i] -= coef * M[1]; gz[i] -= coef * M[2];
        }
    } else if (pl->dipole) {   // kernel.cpp:1580-1596 (reference, energy-only potential term)
        double coef = 4 * M_PI / (3 * pl->cycle * pl->cycle * pl->cycle);
        double M[3] = {0, 0, 0};
        for (int i = 0; i < n; i++) { M[0] += pl->bx[i] * q[i]; M[1] += pl->by[i] * q[i]; M[2] += pl->bz[i] * q[i]; }
        double dn = M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
        for (int i = 0; i < n; i++) {
            p[i] -= coef * dn / n / q[i];
            gx[i] -= coef * M[0]; gy[i] -= coef * M[1]; gz[i] -= coef * M[2];
        }
    }
    if (pl->div4pi) {   // body.cpp:88-96
        double s = 1.0 / (4 * M_PI);
        for (int i = 0; i < n; i++) { p[i] *= s; gx[i] *= s; gy[i] *= s; gz[i] *= s; }
    }
    for (int i = 0; i < n; i++) {
        int o = pl->bidx[i];
        pout[o] = p[i]; gxo[o] = gx[i]; gyo[o] = gy[i]; gzo[o] = gz[i];
    }
}

// Charges in tree order; per-cell vectors and body fields cleared.
static void set_charges(Plan* pl, const double* qin) {
    const int n = pl->n;
    const size_t C = pl->cells.size();
    pl->q.resize(n);
    for (int i = 0; i < n; i++) pl->q[i] = qin[pl->bidx[i]];
    pl->p.assign(n, 0.0); pl->gx.assign(n, 0.0); pl->gy.assign(n, 0.0); pl->gz.assign(n, 0.0);
    pl->qeq.assign(C * pl->S, 0.0);
    pl->pck.assign(C * pl->S, 0.0);
}

// LaplaceFMM::solve stage sequence (fmm.cpp:50-80).  Outputs in input order.
int orc_evaluate(void* h, const double* qin, double* pout, double* gxo, double* gyo, double* gzo) {
    Plan* pl = (Plan*)h;
    if (pl->nthreads > 0) omp_set_num_threads(pl->nthreads);
    if (!pl->have_ops) { pl->err = "operators not computed"; return -1; }
    set_charges(pl, qin);
    int mind = 0, maxd = 0;
    depth_range(*pl, mind, maxd);
    const std::vector<int> leaves = leaf_cells(*pl);
    double t0 = now();
    stage_p2m(*pl, leaves, pl->q.data());
    double t1 = now();
    double qsum = 0;
    for (int i = 0; i < pl->n; i++) qsum += pl->q[i];
    stage_m2m(*pl, mind, maxd, pl->dipole == 2 ? &qsum : nullptr);
    double t2 = now();
    stage_m2l(*pl);
    double t3 = now();
    stage_l2l(*pl, mind, maxd);
    double t4 = now();
    stage_l2p(*pl, leaves, pl->p.data(), pl->gx.data(), pl->gy.data(), pl->gz.data());
    double t5 = now();
    stage_p2p(*pl, pl->q.data(), pl->p.data(), pl->gx.data(), pl->gy.data(), pl->gz.data());
    double t6 = now();
    finalize(pl, pout, gxo, gyo, gzo);
    pl->timers["P2M"] = t1 - t0; pl->timers["M2M"] = t2 - t1; pl->timers["M2L"] = t3 - t2;
    pl->timers["L2L"] = t4 - t3; pl->timers["L2P"] = t5 - t4; pl->timers["P2P"] = t6 - t5;
    pl->timers["FMM_kernels"] = now() - t0;
    return 0;
}

// One stage of orc_evaluate at a time (what the library's per-stage entry points are compared with):
// "charges" (qin, input order; clears the per-cell vectors and body fields), "p2m", "m2m", "m2l",
// "l2l", "l2p" (body fields = far field), "p2p" (body fields += near field), "p2p_only" (body
// fields = near field).  Each stage reads the state the previous calls left.
int orc_stage(void* h, const char* name, const double* qin) {
    Plan* pl = (Plan*)h;
    if (pl->nthreads > 0) omp_set_num_threads(pl->nthreads);
    if (!pl->have_ops) { pl->err = "operators not computed"; return -1; }
    const std::string s(name ? name : "");
    if (s == "charges") {
        if (!qin) return -1;
        set_charges(pl, qin);
        return 0;
    }
    if ((int)pl->q.size() != pl->n) { pl->err = "charges not set"; return -1; }
    int mind = 0, maxd = 0;
    depth_range(*pl, mind, maxd);
    if (s == "p2m") stage_p2m(*pl, leaf_cells(*pl), pl->q.data());
    else if (s == "m2m") {
        double qsum = 0;
        for (int i = 0; i < pl->n; i++) qsum += pl->q[i];
        stage_m2m(*pl, mind, maxd, pl->dipole == 2 ? &qsum : nullptr);
    }
    else if (s == "m2l") stage_m2l(*pl);
    else if (s == "l2l") stage_l2l(*pl, mind, maxd);
    else if (s == "l2p" || s == "p2p_only") {
        const int n = pl->n;
        pl->p.assign(n, 0.0); pl->gx.assign(n, 0.0); pl->gy.assign(n, 0.0); pl->gz.assign(n, 0.0);
        if (s == "l2p") stage_l2p(*pl, leaf_cells(*pl), pl->p.data(), pl->gx.data(), pl->gy.data(), pl->gz.data());
        else stage_p2p(*pl, pl->q.data(), pl->p.data(), pl->gx.data(), pl->gy.data(), pl->gz.data());
    } else if (s == "p2p") stage_p2p(*pl, pl->q.data(), pl->p.data(), pl->gx.data(), pl->gy.data(), pl->gz.data());
    else { pl->err = "unknown stage " + s; return -1; }
    return 0;
}

int orc_stage_finalize(void* h, double* pout, double* gxo, double* gyo, double* gzo) {
    Plan* pl = (Plan*)h;
    if ((int)pl->q.size() != pl->n) return -1;
    finalize(pl, pout, gxo, gyo, gzo);
    return 0;
}

// Body fields in tree order.
int orc_get_body_fields(void* h, double* p, double* gx, double* gy, double* gz) {
    Plan* pl = (Plan*)h;
    if ((int)pl->p.size() != pl->n) return -1;
    std::memcpy(p, pl->p.data(), pl->n * sizeof(double));
    std::memcpy(gx, pl->gx.data(), pl->n * sizeof(double));
    std::memcpy(gy, pl->gy.data(), pl->n * sizeof(double));
    std::memcpy(gz, pl->gz.data(), pl->n * sizeof(double));
    return 0;
}

int orc_set_stage_vectors(void* h, const double* qeq, const double* pck) {
    Plan* pl = (Plan*)h;
    const size_t m = pl->cells.size() * pl->S;
    if (qeq) pl->qeq.assign(qeq, qeq + m);
    if (pck) pl->pck.assign(pck, pck + m);
    return 0;
}

double orc_timer(void* h, const char* name) {
    Plan* pl = (Plan*)h;
    auto it = pl->timers.find(name);
    return it == pl->timers.end() ? -1.0 : it->second;
}

int orc_num_cells(void* h) { return (int)((Plan*)h)->cells.size(); }
int orc_num_tree_cells(void* h) { return ((Plan*)h)->ntree; }
int orc_lost_pairs(void* h) { return ((Plan*)h)->lost_pairs; }
int orc_surface_num(void* h) { return ((Plan*)h)->S; }
int orc_num_freq(void* h) { return ((Plan*)h)->F; }

// ints: 7 per cell (idx, octant, depth, coff, cnum, boff, bnum); dbl: 4 per cell (r, x, y, z)
int orc_get_cells(void* h, int* ints, double* dbl) {
    Plan* pl = (Plan*)h;
    for (size_t c = 0; c < pl->cells.size(); c++) {
        const Cell& cc = pl->cells[c];
        int* o = ints + 7 * c;
        o[0] = cc.idx; o[1] = cc.octant; o[2] = cc.depth; o[3] = cc.coff; o[4] = cc.cnum; o[5] = cc.boff; o[6] = cc.bnum;
        double* d = dbl + 4 * c;
        d[0] = cc.r; d[1] = cc.x[0]; d[2] = cc.x[1]; d[3] = cc.x[2];
    }
    return 0;
}

int orc_get_body_order(void* h, int* perm) {
    Plan* pl = (Plan*)h;
    for (int i = 0; i < pl->n; i++) perm[i] = pl->bidx[i];
    return 0;
}

long orc_p2p_size(void* h) {
    Plan* pl = (Plan*)h;
    long s = 0;
    for (auto& v : pl->p2p) s += (long)v.size();
    return s;
}
// rows in map order (target ascending, then insertion order)
int orc_get_p2p(void* h, int* tar, int* src, double* off) {
    Plan* pl = (Plan*)h;
    long k = 0;
    for (size_t t = 0; t < pl->p2p.size(); t++)
        for (auto& e : pl->p2p[t]) {
            tar[k] = (int)t; src[k] = e.src;
            off[3 * k] = e.off[0]; off[3 * k + 1] = e.off[1]; off[3 * k + 2] = e.off[2];
            k++;
        }
    return 0;
}
long orc_m2l_size(void* h) {
    Plan* pl = (Plan*)h;
    long s = 0;
    for (auto& v : pl->m2l) s += (long)v.size();
    return s;
}
int orc_get_m2l(void* h, int* tar, int* src, double* off, int* img) {
    Plan* pl = (Plan*)h;
    long k = 0;
    for (size_t t = 0; t < pl->m2l.size(); t++)
        for (auto& e : pl->m2l[t]) {
            tar[k] = (int)t; src[k] = e.src; img[k] = e.img;
            off[3 * k] = e.off[0]; off[3 * k + 1] = e.off[1]; off[3 * k + 2] = e.off[2];
            k++;
        }
    return 0;
}

static std::vector<double>* op_by_name(Plan* pl, const char* name) {
    std::string s(name);
    if (s == "UT") return &pl->UT;
    if (s == "VSinv") return &pl->VSinv;
    if (s == "UT_l2p") return &pl->UT_l2p;
    if (s == "VSinv_l2p") return &pl->VSinv_l2p;
    if (s == "L2L_IMG") return &pl->l2l_img;
    if (s == "GK8") return &pl->cc8;
    if (s == "GK27") return &pl->cc27;
    if (s.rfind("M2M_IMG", 0) == 0) { int o = atoi(s.c_str() + 7); return (o >= 0 && o < 27) ? &pl->m2m_img[o] : nullptr; }
    if (s.rfind("M2M", 0) == 0) { int o = atoi(s.c_str() + 3); return (o >= 0 && o < 8) ? &pl->m2m[o] : nullptr; }
    if (s.rfind("L2L", 0) == 0) { int o = atoi(s.c_str() + 3); return (o >= 0 && o < 8) ? &pl->l2l[o] : nullptr; }
    return nullptr;
}
long orc_get_operator(void* h, const char* name, double* out, long cap) {
    auto* v = op_by_name((Plan*)h, name);
    if (!v) return -1;
    if (out && cap >= (long)v->size()) std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (long)v->size();
}
// Overwrite one operator (tests use this to evaluate the restatement with another operator set).
int orc_set_operator(void* h, const char* name, const double* in, long n) {
    auto* v = op_by_name((Plan*)h, name);
    if (!v) return -1;
    v->assign(in, in + n);
    return 0;
}

int orc_get_stage_vectors(void* h, double* qeq, double* pck) {
    Plan* pl = (Plan*)h;
    if (qeq) std::memcpy(qeq, pl->qeq.data(), pl->qeq.size() * sizeof(double));
    if (pck) std::memcpy(pck, pl->pck.data(), pl->pck.size() * sizeof(double));
    return 0;
}

// kernel.cpp:11-77: lattice brute force over (3^images)^3 copies, evaluated at the listed targets.
int orc_direct(int n, const double* x, const double* y, const double* z, const double* q, int ntar, const int* tar,
               int images, double cycle, double* p, double* gx, double* gy, double* gz, int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int dm = (int)((std::pow(3, images) - 1) / 2);
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < ntar; t++) {
        int j = tar[t];
        double ap = 0, ax = 0, ay = 0, az = 0;
        for (int pz = -dm; pz <= dm; pz++)
            for (int py = -dm; py <= dm; py++)
                for (int px = -dm; px <= dm; px++) {
                    double ox = px * cycle, oy = py * cycle, oz = pz * cycle;
#pragma omp simd reduction(+ : ap, ax, ay, az)
                    for (int i = 0; i < n; i++) {
                        double dx = x[j] - (x[i] + ox), dy = y[j] - (y[i] + oy), dz = z[j] - (z[i] + oz);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        double ir = r2 > 0 ? 1.0 / std::sqrt(r2) : 0.0;
                        double qi = q[i] * ir;
                        ap += qi;
                        qi *= ir * ir;
                        ax += qi * dx; ay += qi * dy; az += qi * dz;
                    }
                }
        p[t] = ap; gx[t] = -ax; gy[t] = -ay; gz[t] = -az;
    }
    return 0;
}

// ewald.cpp:6-203 restated per target: real space (erfc, minimum image, r < cutoff),
// reciprocal space over the full cube |k_i| <= ksize without k = 0 (tin-foil), self term.
// Returns phi and grad(phi) (the reference's `f` is grad phi) at the listed targets.
// ref_ncrit > 0 selects the reference's real-space schedule instead of the per-body minimum image:
// its own adaptive tree with that ncrit, one periodic offset per (target leaf, source cell) pair from
// the cell centres, cell pruning by (|dx - offset| - r_a - r_b < sqrt(3) * cutoff) (ewald.cpp:120-174).
// In that mode every body is a target (ntar must equal n and tar the identity).
static void ewald_real_cells(Plan& pl, const double* q, double alpha, double cutoff, std::vector<double>& rp,
                             std::vector<double>& rg) {
    double L = pl.cycle;
    int nleaf = 0;
    std::vector<int> leaves;
    for (auto& c : pl.cells) if (c.cnum == 0) leaves.push_back(c.idx);
    nleaf = (int)leaves.size();
    const float s3 = sqrtf(3.0f);
#pragma omp parallel for schedule(dynamic, 1)
    for (int li = 0; li < nleaf; li++) {
        const Cell& tc = pl.cells[leaves[li]];
        std::vector<int> stack{0};
        while (!stack.empty()) {
            int si = stack.back(); stack.pop_back();
            const Cell& sc = pl.cells[si];
            double off[3] = {0, 0, 0}, dx[3];
            for (int d = 0; d < 3; d++) {
                dx[d] = tc.x[d] - sc.x[d];
                if (dx[d] < -L / 2) off[d] -= L;
                else if (dx[d] > L / 2) off[d] += L;
            }
            double e0 = dx[0] - off[0], e1 = dx[1] - off[1], e2 = dx[2] - off[2];
            if (!(std::sqrt(e0 * e0 + e1 * e1 + e2 * e2) - tc.r - sc.r < s3 * cutoff)) continue;
            if (sc.cnum > 0) {
                for (int j = sc.cnum - 1; j >= 0; j--) stack.push_back(sc.coff + j);
                continue;
            }
            for (int i = tc.boff; i < tc.boff + tc.bnum; i++)
                for (int j = sc.boff; j < sc.boff + sc.bnum; j++) {
                    double d0 = pl.bx[i] - pl.bx[j] - off[0], d1 = pl.by[i] - pl.by[j] - off[1], d2 = pl.bz[i] - pl.bz[j] - off[2];
                    double r = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    if (r > 0 && r < cutoff) {
                        double e = std::erfc(alpha * r);
                        rp[i] += q[j] * e / r;
                        double fr = -q[j] / std::pow(r, 3) * (e + (2 * alpha * r * std::pow(M_E, -alpha * alpha * r * r)) / std::sqrt(M_PI));
                        rg[3 * i] += fr * d0; rg[3 * i + 1] += fr * d1; rg[3 * i + 2] += fr * d2;
                    }
                }
        }
    }
}

int orc_ewald(int n, const double* x, const double* y, const double* z, const double* q, int ntar, const int* tar,
              double cycle, int ksize, double alpha, double cutoff, double* p, double* gx, double* gy, double* gz,
              int nthreads, int ref_ncrit) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double L = cycle, scale = 2 * M_PI / L;
    int K = ksize, W = 2 * K + 1;
    // structure factor S(k) = sum_j q_j exp(-i k.x_j scale), over the full cube (ewald.cpp:177-188)
    size_t nw = (size_t)W * W * W;
    std::vector<cplx> Sk(nw, cplx(0, 0));
    int nth = std::max(1, omp_get_max_threads());
    std::vector<std::vector<cplx>> part(nth, std::vector<cplx>(nw, cplx(0, 0)));
#pragma omp parallel
    {
        int tid = omp_get_thread_num();
        std::vector<cplx>& acc = part[tid];
        std::vector<cplx> ex(W), ey(W), ez(W);
#pragma omp for schedule(static)
        for (int j = 0; j < n; j++) {
            for (int k = -K; k <= K; k++) {
                ex[k + K] = std::exp(cplx(0, -scale * x[j] * k));
                ey[k + K] = std::exp(cplx(0, -scale * y[j] * k));
                ez[k + K] = std::exp(cplx(0, -scale * z[j] * k));
            }
            for (int a = 0; a < W; a++)
                for (int b = 0; b < W; b++) {
                    cplx eab = ex[a] * ey[b] * q[j];
                    cplx* row = &acc[((size_t)a * W + b) * W];
                    for (int c = 0; c < W; c++) row[c] += eab * ez[c];
                }
        }
    }
    for (int t = 0; t < nth; t++)
        for (size_t w = 0; w < nw; w++) Sk[w] += part[t][w];
    double coef = 4 * M_PI / (L * L * L);
    std::vector<cplx> val(nw, cplx(0, 0));
    for (int a = 0; a < W; a++)
        for (int b = 0; b < W; b++)
            for (int c = 0; c < W; c++) {
                int ka = a - K, kb = b - K, kc = c - K;
                if (!ka && !kb && !kc) continue;
                double k2 = (double)(ka * ka + kb * kb + kc * kc) * scale * scale;
                size_t w = ((size_t)a * W + b) * W + c;
                val[w] = Sk[w] * (coef * std::exp(-k2 / (4 * alpha * alpha)) / k2);
            }
    std::vector<double> rp, rg;
    std::vector<int> tree_pos;
    if (ref_ncrit > 0) {
        if (ntar != n) return -1;
        Plan pl;
        pl.ncrit = ref_ncrit; pl.cycle = cycle; pl.r0 = cycle / 2; pl.x0[0] = pl.x0[1] = pl.x0[2] = 0; pl.n = n;
        std::vector<int> perm;
        if (build_tree(pl, perm, x, y, z)) return -1;
        pl.bx.resize(n); pl.by.resize(n); pl.bz.resize(n);
        std::vector<double> qt(n), tp(n, 0.0), tg(3 * (size_t)n, 0.0);
        for (int i = 0; i < n; i++) { pl.bx[i] = x[perm[i]]; pl.by[i] = y[perm[i]]; pl.bz[i] = z[perm[i]]; qt[i] = q[perm[i]]; }
        ewald_real_cells(pl, qt.data(), alpha, cutoff, tp, tg);
        rp.assign(n, 0.0); rg.assign(3 * (size_t)n, 0.0);
        for (int i = 0; i < n; i++) {
            rp[perm[i]] = tp[i];
            for (int d = 0; d < 3; d++) rg[3 * (size_t)perm[i] + d] = tg[3 * (size_t)i + d];
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < ntar; t++) {
        int i = tar[t];
        double ap = 0, ax = 0, ay = 0, az = 0;
        if (ref_ncrit > 0) { ap = rp[i]; ax = rg[3 * (size_t)i]; ay = rg[3 * (size_t)i + 1]; az = rg[3 * (size_t)i + 2]; }
        // real space with minimum image (ewald.cpp:120-174 restated per body pair)
        for (int j = 0; j < n && ref_ncrit <= 0; j++) {
            double d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            for (int k = 0; k < 3; k++) {
                if (d[k] < -L / 2) d[k] += L;
                else if (d[k] > L / 2) d[k] -= L;
            }
            double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r > 0 && r < cutoff) {
                double e = std::erfc(alpha * r);
                ap += q[j] * e / r;
                double fr = -q[j] / (r * r * r) * (e + 2 * alpha * r * std::exp(-alpha * alpha * r * r) / std::sqrt(M_PI));
                ax += fr * d[0]; ay += fr * d[1]; az += fr * d[2];
            }
        }
        // reciprocal (ewald.cpp:191-203)
        std::vector<cplx> ex(W), ey(W), ez(W);
        for (int k = -K; k <= K; k++) {
            ex[k + K] = std::exp(cplx(0, scale * x[i] * k));
            ey[k + K] = std::exp(cplx(0, scale * y[i] * k));
            ez[k + K] = std::exp(cplx(0, scale * z[i] * k));
        }
        for (int a = 0; a < W; a++)
            for (int b = 0; b < W; b++) {
                cplx eab = ex[a] * ey[b];
                for (int c = 0; c < W; c++) {
                    size_t w = ((size_t)a * W + b) * W + c;
                    cplx v = val[w] * eab * ez[c];
                    ap += v.real();
                    double im = -v.imag() * scale;   // Re(v * i * scale)
                    ax += im * (a - K); ay += im * (b - K); az += im * (c - K);
                }
            }
        ap -= M_2_SQRTPI * q[i] * alpha;   // ewald.cpp:112-118
        p[t] = ap; gx[t] = ax; gy[t] = ay; gz[t] = az;
    }
    return 0;
}

}  // extern "C"

// ---- test probes for the restated numerics (used only by tests/) ----
extern "C" {
int orc_test_r2c(int N, const double* g, double* out /* 2 * N*N*(N/2+1) */) {
    Dft d(N);
    std::vector<cplx> G((size_t)N * N * (N / 2 + 1));
    d.r2c(g, G.data());
    std::memcpy(out, G.data(), G.size() * sizeof(cplx));
    return 0;
}
int orc_test_c2r(int N, const double* in, double* g) {
    Dft d(N);
    d.c2r(reinterpret_cast<const cplx*>(in), g, 0);
    return 0;
}
int orc_test_svd(int n, const double* A, double* U, double* s, double* V) {
    std::vector<double> a(A, A + (size_t)n * n), u, sv, v;
    svd_jacobi(a, n, u, sv, v);
    std::memcpy(U, u.data(), u.size() * sizeof(double));
    std::memcpy(s, sv.data(), sv.size() * sizeof(double));
    std::memcpy(V, v.data(), v.size() * sizeof(double));
    return 0;
}
}
