/*
 * fmm3dpp — MI355X-native periodic kernel-independent FMM for the 3-D Laplace kernel.
 *
 * C-ABI of the evaluation path that replaces the `3dpp` C++ class API of mumairsadiq/exafmm-rd.
 * Every entry point returns an int status (0 = OK, negative = error, see FMM3DPP_ERR_*) and never
 * exits the process; fmm3dpp_last_error() gives the message.  The caller owns every host array
 * (structure-of-arrays, float64); the plan owns every device allocation.  One plan is not
 * thread-safe; distinct plans may be used concurrently.
 *
 * Reference interfaces replaced (paths under the reference's 3dpp/ directory):
 *   fmm3dpp_plan_create      <- rtfmm::LaplaceFMM::LaplaceFMM(const Bodies3&, const Argument&)
 *                               (include/fmm.h:15; src/fmm.cpp:7-11) and Argument (src/argument.cpp:7-70)
 *   fmm3dpp_precompute       <- LaplaceKernel::precompute + precompute_m2l (src/kernel.cpp:1638-1850)
 *   fmm3dpp_set_particles    <- Tree::build (src/tree.cpp:10-20,102-188) + Traverser::traverse
 *                               (src/traverser.cpp:19-68)
 *   fmm3dpp_evaluate         <- LaplaceFMM::solve stages P2M..P2P, dipole_correction, scale_bodies,
 *                               sort_bodies_by_idx (src/fmm.cpp:50-80)
 *   fmm3dpp_get_cells/_p2p/_m2l <- Tree::get_cells, Traverser::get_map(P2P), get_M2L_parent_map
 *                               (include/tree.h:69; include/traverser.h:50-54)
 *   fmm3dpp_get/set_operator <- the private operator members of LaplaceKernel (include/kernel.h:197-229)
 */
#ifndef FMM3DPP_H
#define FMM3DPP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM3DPP_OK 0
#define FMM3DPP_ERR_ARG (-1)      /* invalid argument or configuration                          */
#define FMM3DPP_ERR_DEVICE (-2)   /* no usable GPU, or the plan was created without one          */
#define FMM3DPP_ERR_STATE (-3)    /* call out of order (e.g. evaluate before set_particles)      */
#define FMM3DPP_ERR_TREE (-4)     /* tree cannot be built (e.g. > ncrit coincident bodies)       */
#define FMM3DPP_ERR_ALLOC (-5)    /* host or device allocation failed                            */
#define FMM3DPP_ERR_HIP (-6)      /* a HIP runtime call failed                                   */

#define FMM3DPP_PREC_FP64 0       /* every stage in float64                                       */
#define FMM3DPP_PREC_MIXED 1      /* float32 near field and M2L, float64 operator applies/lattice */

#define FMM3DPP_DIPOLE_OFF 0      /* no periodic dipole correction                                */
#define FMM3DPP_DIPOLE_REFERENCE 1 /* the reference's correction (src/kernel.cpp:1580-1596)      */
#define FMM3DPP_DIPOLE_PHYSICAL 2 /* infinite-lattice (tin-foil Ewald) gauge, no counterpart in the
                                   * reference: with M = sum q_j x_j, V = cycle^3,
                                   *   phi_i  -= (4 pi / 3V) M.x_i - (2 pi / 3V) sum_j q_j |x_j|^2
                                   *   grad_i -= (4 pi / 3V) M
                                   * which turns the lattice sum over the (3^images)^3 cube into the
                                   * Ewald potential of a neutral system (the cube's dipole sum is
                                   * conditionally convergent; de Leeuw, Perram and Smith 1980).
                                   * The equivalent densities of the root and of each image cell
                                   * c(-m) get their charge, 27^m sum_j q_j, as monopole: the
                                   * pseudo-inverse leaves ~1e-10 of their size, which the image
                                   * levels would multiply by 9 per level into a potential offset. */

#define FMM3DPP_DEVICE_NONE (-1)  /* structure-only plan: tree, lists and operators on the host; */
                                  /* evaluate returns FMM3DPP_ERR_DEVICE                          */

typedef struct fmm3dpp_config {
    int P;              /* points per surface edge, 2..12  (Argument::P, -P)                 */
    int ncrit;          /* max bodies per leaf, >= 1       (Argument::ncrit, -m)             */
    int images;         /* periodic image levels, 0..20    (Argument::images, -i); 0 = free space */
    double cycle;       /* period (box edge)               (Argument::cycle, -c)             */
    double center[3];   /* box centre                      (Argument::x; reference uses 0)   */
    int precision;      /* FMM3DPP_PREC_*                                                     */
    int dipole_mode;    /* FMM3DPP_DIPOLE_*                (Argument::dipole_correction)     */
    int divide_4pi;     /* scale outputs by 1/(4 pi)       (Argument::divide_4pi)            */
} fmm3dpp_config;

typedef struct fmm3dpp_plan fmm3dpp_plan;

/* Plan creation binds the calling thread's HIP device `device` (or FMM3DPP_DEVICE_NONE). */
int fmm3dpp_plan_create(const fmm3dpp_config* cfg, int device, fmm3dpp_plan** out);
int fmm3dpp_plan_destroy(fmm3dpp_plan* plan);
const char* fmm3dpp_last_error(const fmm3dpp_plan* plan);

/* One-time operator precompute for (P, images, cycle); implied by the first evaluate.  When the
 * environment variable FMM3DPP_OPERATOR_CACHE names a directory, the operators are loaded from the
 * file of this (P, images, cycle) there if it is present and valid, and otherwise computed and
 * written there, so that processes after the first skip the computation (results bit-identical). */
int fmm3dpp_precompute(fmm3dpp_plan* plan);

/* Builds the adaptive octree and the interaction lists for n bodies (host arrays, input order).
 * Device plans build on the GPU; structure-only plans (FMM3DPP_DEVICE_NONE) on the host. */
int fmm3dpp_set_particles(fmm3dpp_plan* plan, int64_t n, const double* x, const double* y, const double* z);

/* Same as fmm3dpp_set_particles with device arrays (positions already in HBM); the whole tree and
 * list build runs on the GPU.  Ordered on the plan's stream; returns after the build completed. */
int fmm3dpp_set_particles_device(fmm3dpp_plan* plan, int64_t n, const double* d_x, const double* d_y,
                                 const double* d_z);

/* Potential phi and gradient grad(phi) (the reference's `f`) for charges q, all in input order.
 * grad is n x 3, row-major (gx, gy, gz per body).  Host arrays. */
int fmm3dpp_evaluate(fmm3dpp_plan* plan, const double* q, double* phi, double* grad);

/* Same as fmm3dpp_evaluate with device pointers, ordered on the plan's stream (no host sync
 * other than what the tree build needs).  d_grad is n x 3 row-major.  The P2P runs on a second
 * stream of the plan, forked from the plan's stream once the charges are read and joined back
 * before the final correction, so it overlaps the far field; callers see stream order. */
int fmm3dpp_evaluate_device(fmm3dpp_plan* plan, const double* d_q, double* d_phi, double* d_grad);

/* One whole step on device arrays, LaplaceFMM::solve (src/fmm.cpp:13-81) without host copies:
 * set_particles_device then evaluate_device, with the near-field list build (the frontier traversal
 * and its CSR) running on a second stream of the plan while the far-field stages run on the plan's
 * stream.  Results are those of the two calls.  Returns once the lists are built; the stages are
 * ordered on the plan's stream as in evaluate_device. */
int fmm3dpp_solve_device(fmm3dpp_plan* plan, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                         const double* d_q, double* d_phi, double* d_grad);

/* Sets the HIP stream (hipStream_t) the plan launches on; NULL = the default stream. */
int fmm3dpp_set_stream(fmm3dpp_plan* plan, void* hip_stream);

/* No counterpart in the reference.  enable = 1 (default): the near field (the P2P, and in
 * solve_device the list build) runs on a second stream of the plan beside the far field, forked
 * from and joined back to the plan's stream.  enable = 0: it runs on the plan's stream after the far
 * field.  Results are bit-identical.  The overlap needs the two streams on different hardware queues
 * (GPU_MAX_HW_QUEUES, 4 by default); with many live plans in one process, 0 can be faster. */
int fmm3dpp_set_near_field_stream(fmm3dpp_plan* plan, int enable);

/* No counterpart in the reference (a measurement switch of this implementation).  The 8x8 M2L
 * Hadamard copies the distinct source spectra of each chunk of 256 targets into LDS before the
 * matrix ops (enable = 1, the default); enable = 0 reads every source vector from HBM / L2 instead.
 * Results are bit-identical either way. */
int fmm3dpp_set_hadamard_staging(fmm3dpp_plan* plan, int enable);

/* No counterpart in the reference (a precision switch of the mixed path).  float32 = 1 (the
 * default): the mixed path's tree operator stages run in float32 -- P2M check potentials, the P2M and
 * L2P pseudo-inverse applies (truncated rank), M2M, L2L on the float32 matrix cores, and the L2P
 * evaluation -- with every tree cell's equivalent-density monopole set to the cell's charge after P2M
 * and M2M, at orders P = 6..8 (other orders keep them in float64); the image levels stay float64.
 * float32 = 0: those stages in float64 as in the float64 path.  No effect on FMM3DPP_PREC_FP64 plans. */
int fmm3dpp_set_mixed_applies(fmm3dpp_plan* plan, int float32);

/* ---- structure, for bit-exact comparison with the reference's tree and lists ---- */
typedef struct fmm3dpp_counts {
    int64_t n_bodies;
    int64_t n_cells;        /* tree cells + image cells                                   */
    int64_t n_tree_cells;   /* cells produced by Tree::build                              */
    int64_t n_p2p;          /* rows of the P2P map (target, source, offset)               */
    int64_t n_m2l;          /* rows of the M2L parent map (target, source, offset, image) */
    int64_t n_lost;         /* well-separated leaf pairs with no operator (reference prints "here") */
    int32_t S;              /* surface points 6(P-1)^2+2                                  */
    int32_t n_freq;         /* (2P-1)^2 (P) complex frequencies per child                 */
} fmm3dpp_counts;

int fmm3dpp_get_counts(const fmm3dpp_plan* plan, fmm3dpp_counts* out);
/* ints: 7 per cell (idx, octant, depth, child_offset, child_count, body_offset, body_count);
 * dbl: 4 per cell (r, x, y, z).  BFS order with the image cells appended, as Tree/Traverser. */
int fmm3dpp_get_cells(const fmm3dpp_plan* plan, int32_t* ints, double* dbl);
/* Tree order -> input index of every body. */
int fmm3dpp_get_body_order(const fmm3dpp_plan* plan, int32_t* perm);
int fmm3dpp_get_p2p(const fmm3dpp_plan* plan, int32_t* tar, int32_t* src, double* off3);
int fmm3dpp_get_m2l(const fmm3dpp_plan* plan, int32_t* tar, int32_t* src, double* off3, int32_t* is_image);

/* ---- operators (names as the reference members: UT, VSinv, UT_l2p, VSinv_l2p, M2M0..7,
 *      M2M_IMG0..26, L2L0..7, L2L_IMG, GK8 [26][F][64][2], GK27 [26][F][28][2]) ----
 * get returns the element count (call with out = NULL to query); set replaces one operator
 * and re-derives the fused device forms from the current factors. */
int64_t fmm3dpp_get_operator(fmm3dpp_plan* plan, const char* name, double* out, int64_t cap);
int fmm3dpp_set_operator(fmm3dpp_plan* plan, const char* name, const double* in, int64_t n);

/* Per-cell equivalent densities and check potentials [n_cells][S] after the last evaluate. */
int fmm3dpp_get_stage_vectors(fmm3dpp_plan* plan, double* q_equiv, double* p_check);

/* ---- per-stage entry points: the operator API of the reference's LaplaceKernel ----
 * Stages act on the plan's state in HBM after set_particles: per-cell vectors q_equiv and p_check
 * ([n_cells][S], cells numbered as fmm3dpp_get_cells) and per-body fields (potential and gradient,
 * before the dipole term and scaling).  `cells` lists the cells a stage acts on; NULL = every cell
 * it covers, as evaluate.  fmm3dpp_stage_charges, then p2m, m2m, m2l, l2l, l2p, p2p (accumulate = 1)
 * and finalize with NULL lists give evaluate's results bit for bit.  Listed cells of one tree level
 * run at once; levels run in evaluate's order (M2M from the deepest, L2L from the outermost image
 * cell down).  Host arrays; every call returns after its work.  Replaced interfaces (include/kernel.h,
 * src/kernel.cpp of the reference's 3dpp/):
 *   fmm3dpp_stage_charges   <- the charges of Bodies3 the operators read
 *   fmm3dpp_stage_p2m       <- p2m_precompute (kernel.h:43, kernel.cpp:332-371): q_equiv of the listed
 *                              leaves = 0.5^depth VSinv UT p_check(bodies) (written; with cells = NULL
 *                              every p_check is also cleared, as evaluate does first)
 *   fmm3dpp_stage_m2m       <- m2m_precompute / m2m_img_precompute (kernel.h:47-51, kernel.cpp:395-442):
 *                              q_equiv of the listed parents = sum over children of M2M[octant] q_equiv
 *                              (image cells: the 27 image operators summed), written; in the
 *                              infinite-lattice gauge the root's and image cells' monopole is set to their charge
 *   fmm3dpp_stage_m2l       <- m2l_fft_precompute_t (kernel.h:97, kernel.cpp:1140-1397): p_check +=
 *                              M2L over the plan's M2L parent map (= Traverser::get_M2L_parent_map)
 *   fmm3dpp_stage_l2l       <- l2l_precompute / l2l_img_precompute (kernel.h:101-103,
 *                              kernel.cpp:1419-1435): children's p_check += L2L of the listed parents
 *   fmm3dpp_stage_l2p       <- l2p_precompute (kernel.h:107, kernel.cpp:1458-1511): the body fields of
 *                              the listed leaves = the field of 0.5^depth VSinv_l2p UT_l2p p_check
 *   fmm3dpp_stage_p2p       <- p2p_1toN_256 (kernel.h:39, kernel.cpp:258-311) over the plan's P2P list
 *                              of each listed target leaf (= Traverser::get_map(P2P)); accumulate = 0
 *                              writes the near field, 1 adds it to the body fields
 *   fmm3dpp_stage_finalize  <- dipole_correction + scale_bodies + sort_bodies_by_idx (fmm.cpp:75-80):
 *                              phi[n], grad[n][3] in input order from the body fields
 * The precompute of kernel.h:113-115 is fmm3dpp_precompute. */
int fmm3dpp_stage_charges(fmm3dpp_plan* plan, const double* q);
int fmm3dpp_stage_p2m(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells);
int fmm3dpp_stage_m2m(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells);
int fmm3dpp_stage_m2l(fmm3dpp_plan* plan);
int fmm3dpp_stage_l2l(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells);
int fmm3dpp_stage_l2p(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells);
int fmm3dpp_stage_p2p(fmm3dpp_plan* plan, int64_t ncells, const int32_t* cells, int accumulate);
int fmm3dpp_stage_finalize(fmm3dpp_plan* plan, double* phi, double* grad);
/* Per-cell vectors of the listed cells (NULL: all n_cells), which = 0 q_equiv, 1 p_check; out / in
 * hold ncells x S values (the Cell3::q_equiv / p_check members of the reference, include/tree.h:9-30). */
int fmm3dpp_get_cell_vectors(fmm3dpp_plan* plan, int which, int64_t ncells, const int32_t* cells, double* out);
int fmm3dpp_set_cell_vectors(fmm3dpp_plan* plan, int which, int64_t ncells, const int32_t* cells, const double* in);
/* Body fields of bodies [first, first + count) in tree order (fmm3dpp_get_body_order; a cell's body
 * range from fmm3dpp_get_cells): phi[count], grad[count][3]. */
int fmm3dpp_get_body_fields(fmm3dpp_plan* plan, int64_t first, int64_t count, double* phi, double* grad);
int fmm3dpp_set_body_fields(fmm3dpp_plan* plan, int64_t first, int64_t count, const double* phi, const double* grad);

/* ---- operator cache (the reference recomputes its operators on every solve,
 *      src/kernel.cpp:1638-1850) ----
 * save writes the plan's operators (computing them first if needed) to `path`; load replaces the
 * precompute with the contents of `path`, which must come from a plan with the same P, images and
 * cycle (checked, as is a checksum; FMM3DPP_ERR_ARG otherwise).  Loaded operators are bit-identical
 * to the saved ones. */
int fmm3dpp_save_operators(fmm3dpp_plan* plan, const char* path);
int fmm3dpp_load_operators(fmm3dpp_plan* plan, const char* path);

/* ---- adjoint (no counterpart in the reference: the backward of phi = evaluate(q)) ----
 * For a loss L(phi) with w = dL/dphi, and grad = the forward gradient grad(phi) of the same plan and
 * particles, returns dq = dL/dq and dx = dL/dx (n x 3, input order).  The Laplace lattice kernel is
 * symmetric, so A^T w is one more evaluation with charges w (plus the gauge term in
 * FMM3DPP_DIPOLE_PHYSICAL); dx combines w grad(phi) with q grad(A w).  The energy E = 1/2 q.phi has
 * dE/dq = phi and dE/dx = q grad(phi).  Losses of grad(phi) (forces) are not covered.  Refused
 * (FMM3DPP_ERR_ARG) for FMM3DPP_DIPOLE_REFERENCE, whose potential term -c |M|^2 / (n q_i) has none.
 * The _device form takes device arrays, ordered on the plan's stream. */
int fmm3dpp_evaluate_adjoint(fmm3dpp_plan* plan, const double* q, const double* w, const double* grad, double* dq,
                             double* dx);
int fmm3dpp_evaluate_adjoint_device(fmm3dpp_plan* plan, const double* d_q, const double* d_w, const double* d_grad,
                                    double* d_dq, double* d_dx);

/* ---- Ewald summation (the reference's EwaldSolver(bodies, args).solve(), include/ewald.h:20-21,
 *      src/ewald.cpp:6-203), float64 on the GPU, at ntar target bodies ----
 * Tin-foil boundary:  phi_t = sum_j q_j erfc(alpha r)/r (minimum image, 0 < r < cutoff)
 *   + sum_{0 < max|k_i| <= ksize} (4 pi / L^3) exp(-k^2 / 4 alpha^2) / k^2 Re(S(k) e^{i k.x_t})
 *   - 2 alpha / sqrt(pi) q_t,  k = 2 pi (a, b, c) / cycle.  grad as phi.  ksize <= 23.
 * The reference fixes alpha = ksize / cycle and cutoff = cycle / 2 (ewald.cpp:9-14).
 * fmm3dpp_ewald takes host arrays (targets: indices into the n bodies; phi[ntar], grad[ntar][3])
 * and runs on `device`; fmm3dpp_ewald_device takes device arrays on the current device and is
 * ordered on hip_stream (it returns after the stream has finished). */
int fmm3dpp_ewald(int device, int64_t n, const double* x, const double* y, const double* z, const double* q,
                  int32_t ntar, const int32_t* targets, double cycle, int ksize, double alpha, double cutoff,
                  double* phi, double* grad);
int fmm3dpp_ewald_device(void* hip_stream, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                         const double* d_q, int32_t ntar, const int32_t* d_targets, double cycle, int ksize,
                         double alpha, double cutoff, double* d_phi, double* d_grad);

/* ---- direct lattice sum (the reference's LaplaceKernel::direct, src/kernel.cpp:11-77) ----
 * phi_t = sum over the (3^images)^3 copies of the box and all bodies of q / r (r = 0 skipped),
 * grad likewise; no dipole term.  float64 on the GPU, cost ntar x n x 27^images.  Host arrays
 * (fmm3dpp_direct, on `device`) or device arrays on the current device (fmm3dpp_direct_device,
 * ordered on hip_stream, returns after the stream has finished).  images <= 10. */
int fmm3dpp_direct(int device, int64_t n, const double* x, const double* y, const double* z, const double* q,
                   int32_t ntar, const int32_t* targets, double cycle, int images, double* phi, double* grad);
int fmm3dpp_direct_device(void* hip_stream, int64_t n, const double* d_x, const double* d_y, const double* d_z,
                          const double* d_q, int32_t ntar, const int32_t* d_targets, double cycle, int images,
                          double* d_phi, double* d_grad);

/* No counterpart in the reference: fmm3dpp_direct_device for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results (the same kernels), but the per-chunk partial sums live in a static 16 MiB
 * device array of the library instead of a fresh allocation: ntar <= 524,288 (FMM3DPP_ERR_ARG
 * beyond), no hipMalloc / hipFree.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other host threads, or other streams before this call has returned) race. */
int fmm3dpp_direct_device_noalloc(void* hip_stream, int64_t n, const double* d_x, const double* d_y,
                                  const double* d_z, const double* d_q, int32_t ntar, const int32_t* d_targets,
                                  double cycle, int images, double* d_phi, double* d_grad);

/* ---- batches of independent periodic systems (no counterpart in the reference, whose LaplaceFMM
 *      solves one system; for callers with many systems of uneven size) ----
 * nsys configurations: images, cycle, precision, dipole_mode and divide_4pi must agree; P, ncrit and
 * center may differ.  Each system gets its own tree and lists; the systems of equal P are merged,
 * so every FMM stage is one launch over all of them.  Results equal those of one plan per system.
 * Positions, charges, potentials and gradients of all systems are concatenated in system order
 * (counts[k] bodies for system k), device arrays on the batch's device; calls return after the
 * batch's stream has finished. */
typedef struct fmm3dpp_batch fmm3dpp_batch;
int fmm3dpp_batch_create(const fmm3dpp_config* cfgs, int32_t nsys, int device, fmm3dpp_batch** out);
int fmm3dpp_batch_destroy(fmm3dpp_batch* batch);
const char* fmm3dpp_batch_last_error(const fmm3dpp_batch* batch);
int fmm3dpp_batch_set_particles_device(fmm3dpp_batch* batch, const int64_t* counts, const double* d_x,
                                       const double* d_y, const double* d_z);
int fmm3dpp_batch_evaluate_device(fmm3dpp_batch* batch, const double* d_q, double* d_phi, double* d_grad);
/* enable != 0: the evaluation's launches are captured once into a HIP graph (per set of d_q, d_phi,
 * d_grad pointers) and replayed; set_particles and autotune drop the graph. */
int fmm3dpp_batch_set_graph(fmm3dpp_batch* batch, int enable);
/* Leaf size per P group: for each candidate ncrit, rebuilds the group's trees and lists and times one
 * evaluation with charges d_q (outputs land in d_phi, d_grad); keeps the fastest (tree build +
 * evaluation).  Then the M2L transform tiles (child slots per workgroup, spectrum columns per LDS
 * chunk) of each group are timed over a few candidates and the fastest kept; results do not depend
 * on the tiles.  chosen[nsys] (may be NULL) receives the ncrit of every system. */
int fmm3dpp_batch_autotune(fmm3dpp_batch* batch, const int32_t* candidates, int32_t ncand, const double* d_q,
                           double* d_phi, double* d_grad, int32_t* chosen);
/* Milliseconds of the last evaluate ("evaluate", HIP events), set_particles ("set_particles", host
 * wall) and operator precompute ("precompute"); -1 if unknown. */
double fmm3dpp_batch_timing_ms(const fmm3dpp_batch* batch, const char* name);

/* Stage timings of the last evaluate in milliseconds (HIP events on the plan's stream); names:
 * P2M, M2M, M2L_up, M2L_hadamard (the 8x8 kernel only), M2L_down (image-lattice 1x27 + down),
 * L2L, L2P, P2P (on the second stream, overlapping the far-field stages), finalize (from the join
 * of the near field), evaluate (for solve_device: from the first far-field launch, so it includes
 * the list build that overlaps it), near_start / far_end (from the first far-field launch to the
 * start of the P2P / the end of L2P); host wall times: tree_and_lists (last
 * set_particles), tree (last solve_device: host time of the tree build until the far field is queued),
 * precompute.  -1 if unknown. */
double fmm3dpp_timing_ms(const fmm3dpp_plan* plan, const char* name);

#ifdef __cplusplus
}
#endif
#endif /* FMM3DPP_H */
