/*
 * tenes_amd.h -- C ABI of the MI355X-native CTMRG hot path.
 *
 * The library replaces TeNeS's CTM environment path (mptensor + LAPACK on
 * the CPU) with FP64 HIP kernels on one gfx950 GPU.  TeNeS has no FFI for
 * this path (its only C entry point is the whole program,
 * src/iTPS/main.cpp:33-39 tenes_itps_main), so each entry point below
 * replaces one templated C++ function of tenes::itps::core; the file:line of
 * the function replaced is given with each declaration (paths relative to
 * the TeNeS source tree).
 *
 * Conventions
 *  - Every function returns 0 on success and a negative status on failure;
 *    tenes_ctx_last_error() returns the message.  No C++ exception crosses
 *    the boundary.
 *  - Tensors are dense, row-major (last leg fastest), real FP64, resident in
 *    device memory and owned by the library unless created with
 *    tenes_tensor_from_device (borrowed, caller keeps ownership).
 *  - Leg orders are those of TeNeS: site tensor (left, up, right, down,
 *    phys); corners C1..C4 (chi, chi); edges eT (chi, chi, D, D) with the
 *    leg conventions of src/iTPS/core/ctm.cpp:984-1192.
 *  - All work of a context is issued on its HIP stream.  Functions returning
 *    host scalars synchronise that stream.  A context must not be used from
 *    two host threads at once.
 */
#ifndef TENES_AMD_H
#define TENES_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tenes_ctx tenes_ctx;
typedef struct tenes_tensor tenes_tensor;
typedef struct tenes_env tenes_env;

#define TENES_OK 0
#define TENES_ERR_GENERIC (-1)
#define TENES_ERR_HIP (-2)
#define TENES_ERR_ARGUMENT (-3)
#define TENES_ERR_ALLOC (-4)
#define TENES_ERR_UNSUPPORTED (-5)

/* Version of this ABI (major*10000 + minor*100 + patch). */
int tenes_abi_version(void);

/* ---- context ----------------------------------------------------------- */
/* hip_stream may be NULL (the library creates a non-blocking stream). */
int tenes_ctx_create(int device, void *hip_stream, tenes_ctx **out);
int tenes_ctx_destroy(tenes_ctx *ctx);
const char *tenes_ctx_last_error(const tenes_ctx *ctx);
int tenes_ctx_synchronize(tenes_ctx *ctx);
int tenes_ctx_set_seed(tenes_ctx *ctx, uint64_t seed);
/* Accumulated GEMM work issued through this context (2*M*N*K per GEMM). */
int tenes_ctx_stats(tenes_ctx *ctx, double *gemm_flops, int64_t *gemm_calls);
int tenes_ctx_reset_stats(tenes_ctx *ctx);
/* The L projector computations of a CTM move (L = rows or columns of the
 * unit cell) and its renormalisation steps are independent; they run
 * overlapped on up to n side streams (default 8; n = 1: all work in issue
 * order on the context stream).  No counterpart in the reference, whose
 * moves loop over the rows (src/iTPS/core/ctm.cpp:577-615). */
int tenes_ctx_set_max_lanes(tenes_ctx *ctx, int n);
/* Sweep graphs captured and replayed through this context so far
 * (tenes_peps_parameters.graph_sweeps).  No counterpart in the reference. */
int tenes_ctx_graph_stats(tenes_ctx *ctx, int64_t *captures, int64_t *replays);
/* Host run-ahead bound of the CTM driver and moves: after issuing a move
 * (or the pair of moves the driver overlaps) the host waits until at most n
 * issued moves are unfinished on the device; 0 (default): no bound.  No
 * counterpart in the reference. */
int tenes_ctx_set_max_moves_in_flight(tenes_ctx *ctx, int n);
/* Diagnostics: *forked = 1 while the context's work is split over its side
 * streams (inside a CTM move), *lane = the stream new work goes to (-1: the
 * context stream).  Between library calls they are 0 and -1, also after a
 * call that failed inside a move.  No counterpart in the reference. */
int tenes_ctx_lane_state(tenes_ctx *ctx, int *forked, int *lane);
/* Contractions issued from now on (tenes_tensordot and the rest) run on the
 * FP32 MFMA when on == 1 (FP64 data, FP32 operands and accumulation), on
 * the bf16 MFMA as three products of hi/lo bf16 operand halves when on == 2
 * (FP32 accumulation), in FP64 when on == 0; the CTM moves set it
 * themselves from tenes_peps_parameters.contraction_fp32.  No counterpart
 * in the reference. */
int tenes_ctx_set_fp32_contractions(tenes_ctx *ctx, int on);
/* The two contractions of a block tensor with the ket and the bra site
 * tensor in the enlarged corners and in Calc_Next_eT (ctm.cpp:209-212,
 * 224-227, 532-536) run as one kernel that keeps their intermediate on chip
 * (on != 0, the default; FP64 double-layer networks within its size limits)
 * or as the reference's two contractions (on == 0).  The results agree to
 * rounding.  No counterpart in the reference. */
int tenes_ctx_set_fused_site(tenes_ctx *ctx, int on);
/* GEMM autotuning of a context (on != 0): the first contraction of every
 * plan (operand shapes and legs) with at least 1e8 flops times each candidate
 * launch shape of the gather-GEMM (64 or 128 tiles, split-K counts) twice on
 * the context's stream, synchronising, and keeps the fastest for that plan;
 * later contractions of the plan launch it directly.  Results change only at
 * rounding (split-K sums in another order).  Off by default.  No
 * counterpart in the reference. */
int tenes_ctx_set_autotune(tenes_ctx *ctx, int on);
/* Stream-K launches (on != 0) for the FP64 products whose 128 x 128 tile
 * grid leaves at least half a round of 512 workgroups after its whole rounds
 * (2500 x 2500: 400 tiles) with K >= 1024: the K iterations are split
 * evenly over 512 workgroups and a split tile is summed, in a fixed order,
 * by the last workgroup of its share to finish (results change only at
 * rounding, identical from launch to launch).  Off by default; the
 * autotuner tries it as one candidate.  No counterpart in the reference. */
int tenes_ctx_set_stream_k(tenes_ctx *ctx, int on);
/* Truncated SVDs (tenes_svd_top and the projectors) since the last reset and
 * the Jacobi rounds they took (a round orthogonalises p/2 disjoint column
 * pairs of the p x p projected matrix; a sweep is p - 1 rounds). */
int tenes_ctx_svd_stats(tenes_ctx *ctx, int64_t *svd_calls,
                        int64_t *jacobi_rounds);
/* Per-launch GEMM timing with HIP events on the context stream (on != 0).
 * tenes_ctx_kernel_report writes one line per kernel instantiation and
 * problem shape, "name[MxNxK] launches total_ms total_flops", into buf
 * (NUL terminated). */
int tenes_ctx_set_profiling(tenes_ctx *ctx, int on);
int tenes_ctx_kernel_report(tenes_ctx *ctx, char *buf, int buflen);
int tenes_ctx_reset_kernel_report(tenes_ctx *ctx);

/* ---- tensors (mptensor::Tensor replacement) ------------------------------ */
int tenes_tensor_from_host(tenes_ctx *ctx, int rank, const int64_t *shape,
                           const double *data, tenes_tensor **out);
int tenes_tensor_from_device(tenes_ctx *ctx, int rank, const int64_t *shape,
                             double *dev_ptr, tenes_tensor **out);
int tenes_tensor_zeros(tenes_ctx *ctx, int rank, const int64_t *shape,
                       tenes_tensor **out);
int tenes_tensor_rank(const tenes_tensor *t, int *rank);
int tenes_tensor_shape(const tenes_tensor *t, int64_t *shape);
int tenes_tensor_to_host(const tenes_tensor *t, double *out);
int tenes_tensor_device_ptr(const tenes_tensor *t, double **out);
int tenes_tensor_destroy(tenes_tensor *t);

/* mptensor tensordot(A, B, axA, axB) (used throughout src/iTPS/core), with an
 * optional transpose of the result fused into the store (perm == NULL: none).
 * Result legs: free legs of A in order, then free legs of B. */
int tenes_tensordot(const tenes_tensor *A, const tenes_tensor *B, int nax,
                    const int *axA, const int *axB, const int *perm,
                    tenes_tensor **out);
/* mptensor transpose(T, axes): new leg k = old leg axes[k]. */
int tenes_transpose(const tenes_tensor *T, const int *axes, tenes_tensor **out);
/* Leading k singular triplets of a 2-D tensor (mptensor svd at
 * src/iTPS/core/ctm.cpp:273).  U: (m, k), V: (n, k), s: k values, descending.
 * *k_out receives the number returned. */
int tenes_svd_top(const tenes_tensor *A, int k, int power_iters,
                  tenes_tensor **U, tenes_tensor **V, double *s, int *k_out);
/* tenes_svd_top for a matrix of at most 128 x 128 (its exact one-sided
 * Jacobi path) that allocates no device memory.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  A: the m x n
 * row-major device matrix; U (m x k'), V (n x k') and s (k') are
 * caller-owned device buffers, k' = min(k, m, n) returned in *k_out.  Work
 * is issued on the context's stream without synchronising.  The scratch is
 * a fixed per-device array of 4166400 bytes set up when the first context
 * is created on the device; m or n above 128 returns TENES_ERR_UNSUPPORTED.
 * Calls that overlap on one device (several streams or contexts) share the
 * array and race: order them on one stream. */
int tenes_svd_top_noalloc(tenes_ctx *ctx, const double *A, int m, int n,
                          int k, double *U, double *V, double *s, int *k_out);

/* ---- parameters (PEPS_Parameters.hpp:29-94, CTM fields) ------------------ */
typedef struct tenes_peps_parameters {
  int CHI;
  double Inverse_projector_cut;
  double CTM_Convergence_Epsilon;
  int Max_CTM_Iteration;
  int CTM_Projector_corner;
  int Use_RSVD;
  double RSVD_Oversampling_factor;
  int print_level;
  int svd_power_iters;  /* GPU range-finder iterations from a Gaussian start
                         * block (full-SVD mode) */
  int rsvd_power_iters; /* GPU range-finder iterations (Use_RSVD mode)  */
  /* Warm start (no counterpart in the reference): each projector's range
   * finder starts from the right Ritz basis the same projector (environment,
   * move, coordinate, row) produced in its previous sweep, kept across
   * tenes_calc_ctm_environment calls that continue an environment
   * (initialize = 0), with svd_warm_power_iters iterations; start columns
   * that are exactly zero (rank-deficient earlier sweeps) are replaced by
   * Gaussian ones.  Defaults: 1, 8 (svd_power_iters: 8). */
  int svd_warm_start;
  int svd_warm_power_iters;
  /* Mixed precision (no counterpart in the reference): the contractions of
   * the moves on the FP32 MFMA (FP64 data, FP32 operands and accumulation):
   * 1 = all but the projector matrix R = LT LB, 2 = R as well; 3 = all of
   * them, R included, on the bf16 MFMA as three products of hi/lo bf16
   * operand halves (split precision, FP32 accumulation); projector SVDs,
   * QRs, the convergence check, Arnoldi and observables stay FP64.
   * Default 0. */
  int contraction_fp32;
  /* Where the projector cut of ctm.cpp:275-288 is decided inside the
   * environment driver and the moves.  0 (default): on the host, the
   * reference's shapes, one synchronisation per projector.  1 (no
   * counterpart in the reference): on the device -- every projector keeps
   * min(rows, cols, CHI) columns and a dropped direction gets a zero
   * column, so a sweep has no host round trip and fixed shapes; the
   * environment tensors stay CHI-sized where the reference's shrink (zero
   * rows and columns instead: the same network, the same observables).  The
   * standalone projector entry points always cut on the host. */
  int projector_cut_on_device;
  /* Sweeps of tenes_calc_ctm_environment with corner projectors captured
   * into a hipGraph and replayed (no counterpart in the reference; implies
   * the projector cut on the device).  The first sweep in which every
   * projector has its warm-start basis is captured; later sweeps replay it
   * while the environment's buffers, shapes and range-finder decisions stay
   * the same (else it is captured again) and while the environment alone
   * holds its tensors (a replay overwrites them in place); other sweeps run
   * as usual.  Default 0. */
  int graph_sweeps;
} tenes_peps_parameters;
/* Defaults of PEPS_Parameters.cpp:29-77. */
void tenes_peps_parameters_default(tenes_peps_parameters *p);

/* ---- CTM core (src/iTPS/core/ctm.hpp) ------------------------------------- */
/* Calc_projector_left_block, ctm.cpp:174-318 */
int tenes_calc_projector_left_block(
    const tenes_tensor *C1, const tenes_tensor *C4, const tenes_tensor *eT1,
    const tenes_tensor *eT6, const tenes_tensor *eT7, const tenes_tensor *eT8,
    const tenes_tensor *Tn1, const tenes_tensor *Tn4,
    const tenes_peps_parameters *p, tenes_tensor **PU, tenes_tensor **PL);
/* Calc_projector_updown_blocks, ctm.cpp:320-491; T[16] =
 * {C1,C2,C3,C4, eT1..eT8, Tn1..Tn4} */
int tenes_calc_projector_updown_blocks(const tenes_tensor *const *T,
                                       const tenes_peps_parameters *p,
                                       tenes_tensor **PU, tenes_tensor **PL);
/* Calc_Next_CTM, ctm.cpp:493-518 */
int tenes_calc_next_ctm(const tenes_tensor *C1, const tenes_tensor *C4,
                        const tenes_tensor *eT1, const tenes_tensor *eT6,
                        const tenes_tensor *PU, const tenes_tensor *PL,
                        tenes_tensor **C1_out, tenes_tensor **C4_out);
/* Calc_Next_eT, ctm.cpp:520-552 */
int tenes_calc_next_et(const tenes_tensor *eT8, const tenes_tensor *Tn1,
                       const tenes_tensor *PU, const tenes_tensor *PL,
                       tenes_tensor **eT_out);

/* ---- environment object: the iTPS arrays Tn, C1..C4, eTt/eTr/eTb/eTl ------
 * (iTPS.hpp:293-315) on a SquareLattice (SquareLattice.hpp:54-163). */
#define TENES_C1 0
#define TENES_C2 1
#define TENES_C3 2
#define TENES_C4 3
#define TENES_ETT 4
#define TENES_ETR 5
#define TENES_ETB 6
#define TENES_ETL 7
#define TENES_TN 8
int tenes_env_create(tenes_ctx *ctx, int LX, int LY, int skew,
                     tenes_env **out);
int tenes_env_destroy(tenes_env *env);
/* Set/get one tensor of the environment (kind = TENES_C1 .. TENES_TN).
 * Getting returns a new handle sharing the data (destroy it when done). */
int tenes_env_set(tenes_env *env, int kind, int site, const tenes_tensor *t);
int tenes_env_get(tenes_env *env, int kind, int site, tenes_tensor **out);
/* Calc_CTM_Environment, ctm.cpp:971-1253 (initialize != 0: build C/eT
 * from Tn first).  *count = sweeps done, *sig_max = last convergence
 * measure (either may be NULL). */
int tenes_calc_ctm_environment(tenes_env *env, const tenes_peps_parameters *p,
                               int initialize, int *count, double *sig_max);
/* Calc_CTM_Environment_density, ctm_single.cpp:737-1057: the single-layer
 * CTM environment of an iTPO (finite-temperature density operator).  The
 * environment's site tensors (TENES_TN) are TPOs (D_l, D_u, D_r, D_d, d, d)
 * (physical ket leg, then bra leg); the CTM runs on their traces over the
 * physical legs (Make_single_tensor_density, ctm_single.cpp:43-50).  The
 * initial corners and edges keep the site bond dimensions (no padding to
 * CHI, ctm_single.cpp:751-995); edges are stored as (chi, chi, D, 1).
 * Corner and up-down projectors (CTM_Projector_corner = 0:
 * Calc_projector_updown_blocks_single, ctm_single.cpp:267-436, whose
 * full-SVD branch truncates with its loop over the first e78 singular
 * values, ctm_single.cpp:396-415; with the cut on the device the shapes stay
 * CHI-sized with zero columns where that loop shrinks them).  Once the site
 * tensors are TPOs, the contraction entry points below (one-site, op12,
 * cluster, multisite, correlation, transfer matrix) evaluate TeNeS's
 * density versions (contract_density_ctm/): Tr(rho O) with each one-site
 * operator O[a][b] traced against the TPO's (ket a, bra b) legs, and an op12
 * operator split into one-site products by its SVD, as twosite_obs.cpp:
 * 223-243 does for longer bonds. */
int tenes_calc_ctm_environment_density(tenes_env *env,
                                       const tenes_peps_parameters *p,
                                       int initialize, int *count,
                                       double *sig_max);
/* Calc_CTM_Environment (ctm.cpp:971-1253) on n independent environments
 * at once.  Environments that belong to one context run in order on one host
 * thread; different contexts run concurrently on their own streams, so
 * spreading the problems over several contexts fills the device with
 * several latency-bound projector chains.  params: 1 (shared) or n entries;
 * count / sig_max: n entries each (may be NULL).  No counterpart in the
 * reference (one problem per run); on error the message is available from
 * tenes_ctx_last_error(NULL) and from the failing context. */
int tenes_batch_calc_ctm_environment(tenes_env *const *envs, int n,
                                     const tenes_peps_parameters *params,
                                     int nparams, int initialize, int *count,
                                     double *sig_max);
/* Left_move / Right_move / Top_move / Bottom_move, ctm.cpp:562-813.
 * dir: 0 left, 1 right, 2 top, 3 bottom; coord = ix (left/right) or iy. */
int tenes_env_move(tenes_env *env, int dir, int coord,
                   const tenes_peps_parameters *p);
/* Check_Convergence_CTM, ctm.cpp:815-969, against corners saved with
 * tenes_env_save_corners.  *converged = 0/1. */
int tenes_env_save_corners(tenes_env *env);
int tenes_env_check_convergence(tenes_env *env, const tenes_peps_parameters *p,
                                int *converged, double *sig_max);
/* L2-normalised singular spectrum of one corner (diagnostics):
 * writes the corner's min(rows, cols) <= 256 values, *n_out receives the
 * count. */
int tenes_env_corner_spectrum(tenes_env *env, int kind, int site, double *out,
                              int *n_out);

/* ---- environment contractions (contract_itps_ctm.hpp:45-160) -------------
 * Operators are host arrays, row-major: op (d, d); op12 (d1, d2, d1, d2). */
/* Contract_one_site_iTPS_CTM, contract_itps_ctm/ctm.cpp:89-113, on the
 * environment of `site` (onesite_obs.cpp:86-97). */
int tenes_env_contract_one_site(tenes_env *env, int site, const double *op,
                                double *out);
/* Contract_two_sites_{horizontal,vertical}_op12_iTPS_CTM
 * (contract_itps_ctm/ctm.cpp:213-319) for the nearest-neighbour bond
 * (source, dx, dy) with |dx| + |dy| == 1, using the tensor selection of
 * twosite_obs.cpp:146-222 (operator legs swapped when source is the right
 * or bottom site).  Longer bonds (|dx|, |dy| <= 3; TENES_ERR_UNSUPPORTED
 * beyond) take twosite_obs.cpp:223-243: the operator is split by its SVD
 * over (source legs | target legs) and each product of one-site operators
 * is contracted on the bond's cluster (Contract_iTPS_CTM). */
int tenes_env_contract_two_sites_op12(tenes_env *env, int source, int dx,
                                      int dy, const double *op12,
                                      double *out);
/* Contract_iTPS_CTM (contract_itps_ctm/ctm.cpp:45-87, ctm_NxM.cpp) for the
 * (|dy|+1) x (|dx|+1) cluster spanned by the bond (source, dx, dy), with a
 * one-site operator on the source and one on the target (NULL = identity),
 * twosite_obs.cpp:146-260.  With both operators NULL this is the bond norm of
 * twosite_obs.cpp:168.  |dx|, |dy| <= 3 (TENES_ERR_UNSUPPORTED beyond, where
 * twosite_obs.cpp:53-60 skips the bond). */
int tenes_env_contract_cluster(tenes_env *env, int source, int dx, int dy,
                               const double *op_source,
                               const double *op_target, double *out);
/* Bond energy with the environment held fixed, and its gradient.
 * E = (1/N_UNIT) sum_b <op12_b>, where <op12_b> is the op12 contraction of
 * Contract_two_sites_{horizontal,vertical}_op12_iTPS_CTM
 * (contract_itps_ctm/ctm.cpp:213-319) divided by the bond norm, i.e. the
 * bond values of measure_twosite (twosite_obs.cpp:146-262) summed as the
 * bond_hamiltonian energy of save_density (density.cpp:39-76); bond b is
 * (source[b], dx[b], dy[b]) with |dx| + |dy| == 1 and its operator is the
 * next (d_source d_target)^2 doubles of op12, row-major (source, target,
 * source', target').  grad: NULL for the energy only, else N_UNIT tensors
 * shaped like the site tensors, overwritten with dE/dTn_i (real tensors:
 * the ket and the bra occurrence both counted).  The gradient has no
 * counterpart in the reference (TeNeS optimises by imaginary-time
 * evolution); it is the adjoint of the two-site reduced density matrices. */
int tenes_env_bond_energy(tenes_env *env, int nbonds, const int *source,
                          const int *dx, const int *dy, const double *op12,
                          double *energy, tenes_tensor *const *grad);
/* Copy a tensor into caller-owned device memory (size elements, the
 * tensor's row-major order) and wait for the copy. */
int tenes_tensor_to_device(const tenes_tensor *t, double *dst);
/* Multisite observable, multisite_obs.cpp:46-220: <prod_i O_i> / <1> on the
 * smallest cluster holding the source and the nothers sites (dx[i], dy[i])
 * from it (Contract_iTPS_CTM: clusters up to 4x4); ops holds the
 * (d, d) host operators of the source, then of each other site. */
int tenes_env_contract_multisite(tenes_env *env, int source, int nothers,
                                 const int *dx, const int *dy,
                                 const double *ops, double *out);
/* Correlation functions, correlation_function.cpp:44-220 with the
 * contractions of contract_itps_ctm/correlation.cpp:34-105:
 * out[r * nright + j] = <O_left(site) O_j(site + r+1)> / <1 1> for
 * r = 0..r_max-1 along +x (direction 0) or +y (direction 1); left_op and
 * the nright right_ops are host (d, d) arrays (one d for all sites). */
int tenes_env_correlation(tenes_env *env, int left_site, const double *left_op,
                          int nright, const double *right_ops, int r_max,
                          int direction, double *out);
/* Create_Environment_two_sites, core/full_update.cpp:51-83 */
int tenes_create_environment_two_sites(const tenes_tensor *const *T,
                                       tenes_tensor **out);
/* T[12] = {C1,C2,C3,C4, eT1..eT6, Q1, Q2} */

/* ---- simple update (input preparation for the CTM path) ------------------ */
/* Simple_update_bond, core/simple_update.cpp:50-198.  lambda1/lambda2: the
 * four bond-weight vectors of each site concatenated in leg order (lengths
 * = the site tensor's virtual dimensions).  lambda_c: Tn1.shape[connect1]
 * values. */
int tenes_simple_update_bond(const tenes_tensor *Tn1, const tenes_tensor *Tn2,
                             const double *lambda1, const double *lambda2,
                             const tenes_tensor *op12, int connect1,
                             double inverse_lambda_cut, tenes_tensor **Tn1_new,
                             tenes_tensor **Tn2_new, double *lambda_c);
/* simple_update.cpp:103-138: nsteps passes over the two-site evolution
 * operators (source site, source leg, op (d1,d2,d1,d2)) applied to the
 * environment's site tensors; bond weights start at 1 and persist in env. */
int tenes_env_simple_update(tenes_env *env, int nops, const int *source_sites,
                            const int *source_legs,
                            const tenes_tensor *const *ops, int nsteps,
                            double inverse_lambda_cut);
/* Bond weights of (site, leg); *n_out receives the length. */
int tenes_env_get_lambda(tenes_env *env, int site, int leg, double *out,
                         int *n_out);

/* Checkpoint of the environment: iTPS::save_tensors / load_tensors
 * (src/iTPS/saveload_tensors.cpp:38-230).  save writes dir/params.dat
 * (format version 1: N_UNIT, CHI -- chi, or C1[0]'s dimension when chi <= 0
 * -- and every site tensor's shape), one file per tensor (T, Et, Er, Eb,
 * El, C1..C4 _<site>.dat; float64 with a rank/shape header, as mptensor's
 * own format is not available) and lambda_<site>.dat (the bond weights, one
 * per line); the directory is created if needed and every tensor must be
 * set.  load reads such a directory, resizing each tensor to the shape the
 * environment holds for it (zero padding, then cutting: resize_tensor,
 * src/tensor.cpp:114-148; chi > 0 sets the environment legs of C and eT
 * the environment does not hold yet), and fails without changing the
 * environment when the directory or params.dat is missing, the format
 * version is not 1, N_UNIT or a physical dimension differs, or a tensor has
 * another rank. */
int tenes_env_save_tensors(tenes_env *env, const char *dir, int chi);
int tenes_env_load_tensors(tenes_env *env, const char *dir, int chi);

/* Full-update fields of PEPS_Parameters (PEPS_Parameters.hpp:29-94,
 * defaults PEPS_Parameters.cpp:55-64). */
typedef struct tenes_fu_parameters {
  double Inverse_Env_cut;
  double Full_Inverse_precision;
  double Full_Convergence_Epsilon;
  int Full_max_iteration;
  int Full_Gauge_Fix;
  int Full_Use_FastFullUpdate;
} tenes_fu_parameters;
void tenes_fu_parameters_default(tenes_fu_parameters *p);
/* Full_update_bond, core/full_update.cpp:395-441 (horizontal kernel 85-393):
 * T = {C1, C2, C3, C4, eT1, ..., eT6, Tn1, Tn2} (12 handles) as TeNeS passes
 * them, op12 a host array (d1, d2, d1, d2), connect1 the leg of Tn1 on the
 * bond.  The bond environment is built on the GPU; the ALS on tensors of at
 * most (D d)^4 entries runs in FP64 on the host. */
int tenes_full_update_bond(const tenes_tensor *const *T, const double *op12,
                           int connect1, const tenes_fu_parameters *fp,
                           tenes_tensor **Tn1_new, tenes_tensor **Tn2_new);
/* The full-update driver, src/iTPS/full_update.cpp:154-186: a fresh CTM
 * environment, then nsteps passes over the nops two-site operators (bond
 * (source[k], source_leg[k]), operator k the next doubles of ops, row-major
 * (source, target, source', target') with the dimensions op_dims[4k..4k+3];
 * op_dims == NULL: (d_s, d_t, d_s, d_t) from the site tensors.  An operator
 * whose output dimensions differ from its input ones (TeNeS's
 * path-decomposed longer-range gates) changes the sites' physical legs, as
 * Full_update_bond does), each followed by the
 * fast-full-update re-moves (full_update.cpp:114-151) or, with
 * Full_Use_FastFullUpdate = 0, a fresh environment. */
int tenes_env_full_update(tenes_env *env, int nops, const int *source,
                          const int *source_leg, const int *op_dims,
                          const double *ops, int nsteps,
                          const tenes_peps_parameters *pp,
                          const tenes_fu_parameters *fp);

/* ---- transfer matrix (transfer_matrix.hpp:49-139, arnoldi.hpp:32-71) ---- */
typedef struct tenes_tm_parameters {
  int num_eigvals;
  int maxdim_dense_eigensolver;
  int arnoldi_maxdim;
  int arnoldi_restartdim;
  int arnoldi_maxiter;
  double arnoldi_rtol;
} tenes_tm_parameters;
void tenes_tm_parameters_default(tenes_tm_parameters *p);
/* TransferMatrix_ctm::eigenvalues, transfer_matrix.cpp:84-142.
 * On the Arnoldi path (dimension above maxdim_dense_eigensolver) up to
 * arnoldi_maxiter passes over a basis of arnoldi_maxdim vectors, restarted
 * between passes as arnoldi.cpp:117-167 does (shifts: the largest
 * arnoldi_maxdim - arnoldi_restartdim Ritz values); num_eigvals above
 * arnoldi_maxdim there returns TENES_ERR_ARGUMENT.
 * eig: 2*num_eigvals doubles (re, im), sorted by |lambda| descending;
 * *n_out receives the number of eigenvalues written. */
int tenes_env_transfer_matrix_eigenvalues(tenes_env *env, int dir,
                                          int fixed_coord,
                                          const tenes_tm_parameters *p,
                                          double *eig, int *n_out);

#ifdef __cplusplus
}
#endif

#endif /* TENES_AMD_H */
