/*
 * acetn_hip.h -- C-ABI of the MI355X (gfx950) engine for ace-tn's CTMRG step,
 * full-update bond solve and reduced density matrices.
 *
 * Conventions (every entry point):
 *   - returns 0 on success, nonzero on failure; acetn_hip_last_error() gives
 *     the message of the last failure on the calling thread;
 *   - pointers are device pointers (FP64, contiguous row-major) unless the
 *     parameter name says host;
 *   - outputs and the scratch ``workspace`` are caller-allocated (the Python
 *     host side allocates them with torch), sized by the matching
 *     *_workspace() query;
 *   - ``stream`` is a hipStream_t passed as void* (torch's current stream);
 *     work is enqueued on it.  Functions that return a truncation rank to
 *     the host synchronise that stream, as the reference does at
 *     acetn/renormalization/projectors.py:164.
 *
 * Index conventions follow the reference: a site tensor A is (l,u,r,d,p),
 * direction k: 0=left 1=up 2=right 3=down (acetn/ipeps/bond.py:11),
 * corners C[k] are (chi, chi), edges E[k] are (chi, chi, D_ket, D_bra)
 * (acetn/ipeps/site_tensor.py:152-202).
 */
#ifndef ACETN_HIP_H
#define ACETN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One site of the unit cell as the kernels see it: the tensors of
 * acetn.ipeps.SiteTensor (acetn/ipeps/site_tensor.py:4-202) with their
 * current (possibly still growing) boundary dimensions. */
typedef struct acetn_site {
  const double* A;      /* (D, D, D, D, d) */
  const double* C[4];   /* C[k]: (C_dim[k][0], C_dim[k][1]) */
  const double* E[4];   /* E[k]: (E_dim[k][0], E_dim[k][1], D, D) */
  int64_t C_dim[4][2];
  int64_t E_dim[4][2];
  int64_t D;            /* bond dimension */
  int64_t d;            /* physical dimension */
} acetn_site;

/* Arithmetic of the large contractions (the ``precision`` arguments).
 *   ACETN_HIP_FP64:  FP64 matrix cores throughout (the plain method).
 *   ACETN_HIP_MIXED: enlarged-corner, range-finder and renormalisation
 *     products on the bf16 matrix cores with FP32 accumulation, operands split
 *     into bf16 pieces (bf16x3 for the sketch and the power iterations, bf16x6
 *     after the final orthonormalisation); the double-layer absorption on the
 *     FP32 matrix cores; the bases of the power iterations take one shifted
 *     CholeskyQR pass.  The QR of the final basis, the projector SVD and the
 *     normalisations stay FP64, and so does every tensor that crosses this
 *     interface (the half-system corners are held in FP32 inside the
 *     workspace).
 * No reference counterpart: the reference runs one dtype end to end
 * (acetn/ipeps/ipeps_config.py:15, ``dtype``). */
enum { ACETN_HIP_FP64 = 0, ACETN_HIP_MIXED = 1 };
/* OR-ed into the precision argument of a projector entry point: the range
 * finder's orthonormalisations refill the columns their CholeskyQR rank
 * guard drops with random vectors instead of re-deriving them (some 30 fewer
 * launches per orthonormalisation).  Identical results whenever no column is
 * dropped, which holds when the spectrum is not graded: a caller can tell
 * from the previous spectrum of the same pair (s_{q-1} / s_0 >= 1e-3 leaves a
 * margin of 100).  On graded spectra (a product-state start) the refilled
 * columns lose the small singular directions, so the default recovers them. */
enum { ACETN_HIP_RF_FAST = 0x100 };

/* Message of the last failed call on this thread ("" if none). */
const char* acetn_hip_last_error(void);

/* Tile-shape autotuner of the matrix-core products (no reference
 * counterpart).  Mode 0: the built-in thresholds pick each product's tile and
 * split-K count; 1: the first product of every new shape times the candidate
 * tiles and split counts on the caller's stream (HIP events, a host wait per
 * tuned shape) and the fastest is kept in a per-process table that later
 * products use (shapes under stream capture and accumulating products are
 * not tuned); 2: the table is used, nothing new is tuned.  _save/_load write
 * and read the table as text (one shape per line). */
int acetn_hip_autotune(int mode);
int acetn_hip_autotune_entries(int64_t* host_n);
int acetn_hip_autotune_clear(void);
int acetn_hip_autotune_save(const char* path);
int acetn_hip_autotune_load(const char* path, int64_t* host_n);

/* Library / device self-description: writes up to ``len`` bytes. */
int acetn_hip_device_info(char* host_buf, size_t len);

/* Measurement hook (no reference counterpart; it serves the timing
 * convention of acetn/utils/benchmarking.py:45-104): when enabled, every
 * launch of the matrix-core kernels is bracketed by HIP events on its
 * stream; _read() synchronises and returns, for one kernel class, the summed
 * kernel time (ms), the summed algorithmic flops and the launch count.
 * Kernel classes: ACETN_HIP_K_GEMM_F64 (2*M*N*K per launch),
 * ACETN_HIP_K_GEMM_SPLIT (2*M*N*K per launch, not counting the bf16 piece
 * products), ACETN_HIP_K_DOUBLE_LAYER (4*d*D^6 per boundary block),
 * ACETN_HIP_K_JACOBI (the one-workgroup projector eigensolver, rotations and
 * their replay; no flop count: latency-bound), ACETN_HIP_K_CHOLESKY (the
 * one-workgroup CholeskyQR factor and inverse, q^3/2 per launch).
 * Enabling (or disabling) resets the counters. */
enum {
  ACETN_HIP_K_GEMM_F64 = 0,
  ACETN_HIP_K_GEMM_SPLIT = 1,
  ACETN_HIP_K_DOUBLE_LAYER = 2,
  ACETN_HIP_K_JACOBI = 3,
  ACETN_HIP_K_CHOLESKY = 4
};
int acetn_hip_kernel_timing(int enable);
int acetn_hip_kernel_timing_read(int kernel, double* host_ms, double* host_flops,
                                 int64_t* host_launches);
/* Summed algorithmic HBM bytes of one kernel class since the last reset:
 * operands read once and the result written once, in the formats the class
 * moves them in (FP64, or the bf16 planes of ACETN_HIP_K_GEMM_SPLIT's
 * products from planes).  No counterpart in the reference (measurement). */
int acetn_hip_kernel_timing_bytes(int kernel, double* host_bytes);

/* Enlarged corner ("quarter tensor") Q_k of one site, as an
 * (C_dim[k][1]*D^2) x (E_dim[(3+k)%4][0]*D^2) matrix.
 * Replaces ProjectorCalculator.make_quarter_tensor,
 * acetn/renormalization/projectors.py:35-60 (the max-abs scaling of :59 is a
 * scale that cancels in the Frobenius-normalised C/E and is not applied). */
size_t acetn_hip_quarter_tensor_workspace(const acetn_site* s, int k);
int acetn_hip_quarter_tensor(const acetn_site* s, int k, int precision, double* Q,
                             void* workspace, size_t ws_bytes, void* stream);

/* Half-system projector pair from the randomized SVD of Q1*Q4 (Q1 of site
 * s1 with rotation k, Q4 of site s4 with rotation k+3).
 * Replaces calculate_half_system (rsvd branch),
 * acetn/renormalization/projectors.py:138-174, and fused_matmul_svd_lowrank,
 * acetn/linalg/fused_matmul_svd_lowrank.py:4-52.  The sketch is the
 * Philox4x32-10 Gaussian of stream ``sketch_stream`` under ``seed``.
 * P1, P2: (rows x chi) buffers, written as rows x cD_new (row-major, leading
 * dimension cD_new).  host_cD_new receives cD_new; spectrum (device, q
 * doubles) receives the singular values (descending, unnormalised). */
size_t acetn_hip_half_projectors_workspace(const acetn_site* s1, const acetn_site* s4, int k,
                                           int64_t chi, int64_t oversampling);
int acetn_hip_half_projectors(const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                              int64_t oversampling, int niter, double svd_cutoff, uint64_t seed,
                              uint64_t sketch_stream, int precision, double* P1, double* P2,
                              int64_t* host_cD_new, double* spectrum, void* workspace,
                              size_t ws_bytes, void* stream);

/* The randomized SVD of a product M = M_0 M_1 ... M_{nops-1} of dense
 * row-major FP64 device matrices without forming it.  Replaces the three
 * functions of acetn/linalg: svd_lowrank (nops = 1, linalg/svd_lowrank.py:4-44),
 * fused_matmul_svd_lowrank (nops = 2, linalg/fused_matmul_svd_lowrank.py:4-52)
 * and fused_3matmul_svd_lowrank (nops = 4 with mid_qr = 1: an orthonormalisation
 * between the two halves of each power iteration,
 * linalg/fused_3matmul_svd_lowrank.py:4-56).  ops[i] is dims[i] x dims[i+1]
 * (leading dimension dims[i+1]); q is the caller's q already clamped to
 * min(dims[0], dims[nops]) (at most 510).  The sketch is the Philox Gaussian of
 * (seed, sketch_stream) in place of torch.randn.  Outputs (device, caller
 * allocated): U (dims[0] x q), S (q, descending), V (dims[nops] x q) with
 * M ~ U diag(S) V^T, U and V with orthonormal columns. */
size_t acetn_hip_rsvd_product_workspace(int nops, const int64_t* dims, int64_t q, int niter,
                                        int mid_qr);
int acetn_hip_rsvd_product(int nops, const double* const* ops, const int64_t* dims, int64_t q,
                           int niter, int mid_qr, uint64_t seed, uint64_t sketch_stream, double* U,
                           double* S, double* V, void* workspace, size_t ws_bytes, void* stream);

/* The same computation split at the host readback of cD_new, so that the
 * independent projector pairs of a stage (the left/right or up/down move
 * pair of acetn/renormalization/ctmrg.py:26-31) can run concurrently on
 * separate streams with one host synchronisation per stage:
 *   _begin enqueues everything through the truncation count and an async
 *   copy of cD_new into ``pinned_cD_new`` (pinned host memory);
 *   _end (after the host has synchronised the stream and read cD_new)
 *   enqueues the projector products into P1/P2.  Both calls must receive
 *   the same sites, dimensions and workspace (its contents carry over). */
int acetn_hip_half_projectors_begin(const acetn_site* s1, const acetn_site* s4, int k,
                                    int64_t chi, int64_t oversampling, int niter,
                                    double svd_cutoff, uint64_t seed, uint64_t sketch_stream,
                                    int precision, int32_t* pinned_cD_new, double* spectrum,
                                    void* workspace, size_t ws_bytes, void* stream);
int acetn_hip_half_projectors_end(const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                                  int64_t oversampling, int niter, int precision, int64_t cD_new,
                                  double* P1, double* P2, void* workspace, size_t ws_bytes,
                                  void* stream);

/* All projector branches of ProjectorCalculator (projectors.py:21-236):
 * system ACETN_HIP_HALF_SYSTEM (calculate_half_system, :138-174; sites[0] and
 * sites[3] are used) or ACETN_HIP_FULL_SYSTEM (calculate_full_system,
 * :176-217, with fused_3matmul_svd_lowrank, linalg/fused_3matmul_svd_lowrank.py:4-56);
 * svd_type ACETN_HIP_RSVD (randomised, the default) or ACETN_HIP_FULL_RANK
 * (contract_half_system/contract_full_system + svd, :62-136; on the device
 * for matrices with at most 510 rows).  sites[0..3] are the plaquette s1..s4
 * of acetn/renormalization/directional_mover.py:99-181.  Split at the host
 * readback of cD_new like acetn_hip_half_projectors_begin/_end; the
 * full-rank branch draws no sketch.  P1, P2: (rows x chi) buffers written
 * as rows x cD_new. */
enum { ACETN_HIP_HALF_SYSTEM = 0, ACETN_HIP_FULL_SYSTEM = 1 };
enum { ACETN_HIP_RSVD = 0, ACETN_HIP_FULL_RANK = 1 };
size_t acetn_hip_projectors_workspace(const acetn_site* const sites[4], int k, int64_t chi,
                                      int64_t oversampling, int system, int svd_type);
int acetn_hip_projectors_begin(const acetn_site* const sites[4], int k, int64_t chi,
                               int64_t oversampling, int niter, double svd_cutoff, uint64_t seed,
                               uint64_t sketch_stream, int system, int svd_type, int precision,
                               int32_t* pinned_cD_new, double* spectrum, void* workspace,
                               size_t ws_bytes, void* stream);
int acetn_hip_projectors_end(const acetn_site* const sites[4], int k, int64_t chi,
                             int64_t oversampling, int niter, int system, int svd_type,
                             int precision, int64_t cD_new, double* P1, double* P2,
                             void* workspace, size_t ws_bytes, void* stream);

/* acetn_hip_projectors_begin that also records ``mark_event`` (a hipEvent_t)
 * on the stream once the range finder's matrix-core products are enqueued,
 * ahead of the one-workgroup eigensolver (no counterpart in the reference):
 * a caller running independent projector pairs can start another pair's
 * products when this one reaches its latency-bound part. */
int acetn_hip_projectors_begin_mark(const acetn_site* const sites[4], int k, int64_t chi,
                                    int64_t oversampling, int niter, double svd_cutoff,
                                    uint64_t seed, uint64_t sketch_stream, int system,
                                    int svd_type, int precision, int32_t* pinned_cD_new,
                                    double* spectrum, void* workspace, size_t ws_bytes,
                                    void* stream, void* mark_event);

/* The same two phases in a form that HIP stream capture can record (no
 * counterpart in the reference, whose loop reads cD_new on the host,
 * projectors.py:163-164; for callers that capture a whole CTMRG iteration
 * once the boundary dimension has saturated at chi):
 *   _begin_graph draws the sketch from Philox stream sketch_stream +
 *   dev_stream_base[0], read on the device when the sketch kernel runs (a
 *   caller that advances that counter between replays draws fresh sketches),
 *   and copies cD_new to dev_cD_new (device memory) instead of host memory;
 *   _end_padded writes cD_cols = min(chi, q) projector columns whatever
 *   cD_new is (q = chi + oversampling, capped by the matrix sizes): the
 *   columns at and past cD_new are zero, an exact embedding of the reference's
 *   truncated projectors in fixed shapes.  Neither call synchronises.
 * Both must receive the same sites, dimensions and workspace. */
int acetn_hip_projectors_begin_graph(const acetn_site* const sites[4], int k, int64_t chi,
                                     int64_t oversampling, int niter, double svd_cutoff,
                                     uint64_t seed, uint64_t sketch_stream,
                                     const uint64_t* dev_stream_base, int system, int svd_type,
                                     int precision, int32_t* dev_cD_new, double* spectrum,
                                     void* workspace, size_t ws_bytes, void* stream);
int acetn_hip_projectors_end_padded(const acetn_site* const sites[4], int k, int64_t chi,
                                    int64_t oversampling, int niter, int system, int svd_type,
                                    int precision, int64_t cD_cols, double* P1, double* P2,
                                    void* workspace, size_t ws_bytes, void* stream);

/* Boundary renormalisation of one (s1 -> s2) target.
 * Replaces DirectionalMover.renormalize_boundary / renormalize_cj1 /
 * renormalize_cj2 / renormalize_ej, acetn/renormalization/directional_mover.py:273-366.
 * Outputs (each Frobenius-normalised):
 *   C_a = C'[(3+k)%4] of s2: (E_dim[(2+k)%4][0] of s1) x cD_i
 *   C_b = C'[k] of s2:       cD_j x (E_dim[k][1] of s1)
 *   E_o = E'[(3+k)%4] of s2: (cD_i, cD_j, D, D)
 * P1_i/P2_i are the projectors of row/column i, P1_j/P2_j those of j, each
 * (rows x cD) row-major. */
size_t acetn_hip_renormalize_workspace(const acetn_site* s1, int k, int64_t cD_i, int64_t cD_j);
int acetn_hip_renormalize_boundary(const acetn_site* s1, int k, const double* P1_i,
                                   const double* P2_i, int64_t cD_i, const double* P1_j,
                                   const double* P2_j, int64_t cD_j, int precision, double* C_a,
                                   double* C_b, double* E_o, void* workspace, size_t ws_bytes,
                                   void* stream);
/* The same renormalisation in two calls on one stream and one workspace, so
 * that the work one projector pair feeds on its own starts before the other
 * pair's projectors exist (the three functions of
 * directional_mover.py:273-366, regrouped by pair).  ``first`` names the pair
 * of part 1:
 *   first = 0 (pair j): part 1 = C_b (renormalize_cj2) and E_o's
 *     intermediate up to its double layer (P1_j contracted first, as
 *     renormalize_ej does); part 2 = C_a (renormalize_cj1) and E_o (P2_i).
 *   first = 1 (pair i): part 1 = C_a and E_o's intermediate with P2_i
 *     contracted first; part 2 = C_b and E_o (P1_j last).
 * Part 1 reads only its pair's projectors; the other count is then unknown
 * and cD_layout (>= it, e.g. chi) fixes the workspace layout for both calls.
 * Size the workspace with acetn_hip_renormalize_part_workspace, passing
 * cD_layout in place of the count not yet known, and pass the same workspace
 * to both parts.  Part 2 takes every projector and both counts. */
size_t acetn_hip_renormalize_part_workspace(const acetn_site* s1, int k, int first, int64_t cD_i,
                                            int64_t cD_j);
int acetn_hip_renormalize_boundary_part(const acetn_site* s1, int k, int part, int first,
                                        const double* P1_i, const double* P2_i, int64_t cD_i,
                                        const double* P1_j, const double* P2_j, int64_t cD_j,
                                        int64_t cD_layout, int precision, double* C_a,
                                        double* C_b, double* E_o, void* workspace,
                                        size_t ws_bytes, void* stream);

/* One-site reduced density matrix rho[P][p] (d x d) of a site from its
 * C[0..3], E[0..3] and A.  Replaces RDM.build_site_rdm,
 * acetn/measurement/rdm.py:35-67. */
size_t acetn_hip_site_rdm_workspace(const acetn_site* s);
int acetn_hip_site_rdm(const acetn_site* s, double* rho, void* workspace, size_t ws_bytes,
                       void* stream);

/* Bond reduced density matrix rho[P][Q][p][q] (d,d,d,d) of the bond (s1, s2, k)
 * in s1's bond frame.  Replaces RDM.build_bond_rdm / build_bond_rdm_core(_blocked),
 * acetn/measurement/rdm.py:69-154. */
size_t acetn_hip_bond_rdm_workspace(const acetn_site* s1, const acetn_site* s2, int k);
int acetn_hip_bond_rdm(const acetn_site* s1, const acetn_site* s2, int k, double* rho,
                       void* workspace, size_t ws_bytes, void* stream);

/* ---- full update (one bond) ---------------------------------------------
 * Reduced bond dimension nD = d*D (min(D^3, dD) with D^2 >= d).  Bond-frame
 * conventions of acetn/evolution/tensor_update.py:18-75: the bond (s1, s2, k)
 * is rotated so that s2 sits on leg 2 ("right") of s1. */

/* decompose_site_tensors, tensor_update.py:52-68: A1, A2 are the site
 * tensors (D,D,D,D,d) in the site frame; outputs a1q (D,D,D,nD) (r,d,u,y),
 * a1r (nD,D,d), a2q (D,D,D,nD) (d,l,u,x), a2r (nD,D,d).  Householder QR with
 * LAPACK's sign conventions (those of the reference's torch.linalg.qr). */
size_t acetn_hip_bond_decompose_workspace(int64_t D, int64_t d);
int acetn_hip_bond_decompose(const double* A1, const double* A2, int k, int64_t D, int64_t d,
                             double* a1q, double* a1r, double* a2q, double* a2r, void* workspace,
                             size_t ws_bytes, void* stream);

/* build_norm_tensor / build_norm_tensor_core, full_update.py:163-227:
 * n12 (nD,nD,nD,nD) as [y][x][Y][X] from the environments of s1 and s2. */
size_t acetn_hip_norm_tensor_workspace(const acetn_site* s1, const acetn_site* s2, int k);
int acetn_hip_norm_tensor(const acetn_site* s1, const acetn_site* s2, int k, const double* a1q,
                          const double* a2q, double* n12, void* workspace, size_t ws_bytes,
                          void* stream);

/* The bond solver's entry points below take the reduced tensors' shape
 * (nD, D, d) as ALSSolver's ar_shape = (nD, bD, pD) gives it
 * (als_solver.py:10-32), with any reduced dimension 1 <= nD <= 32: a
 * full-update step has nD = min(D^3, D d).
 *
 * Gate application of FullUpdater.update_reduced_tensors,
 * full_update.py:77-86: a12g[y][x][r][s] = sum a1r[y,u,p] a2r[x,u,q] gate[p,q,r,s]. */
size_t acetn_hip_apply_gate_workspace(int64_t nD, int64_t D, int64_t d);
int acetn_hip_apply_gate(const double* a1r, const double* a2r, const double* gate, int64_t nD,
                         int64_t D, int64_t d, double* a12g, void* workspace, size_t ws_bytes,
                         void* stream);

/* positive_approx, full_update.py:262-297: nz (nD, nD, nD^2) from n12
 * (nD, nD, nD, nD), nD^2 <= 510.  Synchronises the stream once per
 * eigen-decomposition (the reference's loop condition); host_shifts
 * receives the number of diagonal shifts applied. */
size_t acetn_hip_positive_approx_workspace(int64_t nD);
int acetn_hip_positive_approx(const double* n12, int64_t nD, double cutoff, double* nz,
                              int32_t* host_shifts, void* workspace, size_t ws_bytes,
                              void* stream);

/* gauge_fix, full_update.py:300-343: from nz and a12g, the gauge-fixed n12
 * and a12g and the pseudo-inverses (absolute cutoff atol) of the R factors. */
size_t acetn_hip_gauge_fix_workspace(int64_t nD, int64_t d);
int acetn_hip_gauge_fix(const double* nz, const double* a12g_in, int64_t nD, int64_t d,
                        double atol, double* n12, double* a12g, double* nzxr_inv,
                        double* nzyr_inv, void* workspace, size_t ws_bytes, void* stream);

/* ALS bond solve.  _initialize is the plain start of
 * ALSSolver.initialize_tensors / initialize_reduced_tensors, als_solver.py:112-146.
 * _solve replaces the extension's als_solve (csrc/evolution/als_solve.cpp:107-116,
 * bound at acetn/evolution/als_solver.py:84-110) and solve_torch
 * (als_solver.py:55-82): a1r/a2r carry the start in and the result out; the
 * sweep loop and its stopping rule run on the device (one workgroup);
 * host_sweeps/host_cost/host_info receive the sweep count, the final cost
 * and flags (1: a Cholesky factorisation failed and LU with partial pivoting
 * was used, as the reference's fallback; 2: a singular system).  The final
 * cost is ALSSolver.calculate_cost (als_solver.py:251-261) of the result;
 * with niter = 0, of the start.  The plain start needs nD d <= 32 and
 * D <= nD d. */
enum { ACETN_HIP_ALS_CHOLESKY = 0, ACETN_HIP_ALS_PINV = 1 };
size_t acetn_hip_als_workspace(int64_t nD, int64_t D, int64_t d);
int acetn_hip_als_initialize(const double* a12g, int64_t nD, int64_t D, int64_t d, double* a1r,
                             double* a2r, void* workspace, size_t ws_bytes, void* stream);
int acetn_hip_als_solve(const double* n12, const double* a12g, double* a1r, double* a2r,
                        int64_t nD, int64_t D, int64_t d, int niter, double tol, int method,
                        double epsilon, int64_t* host_sweeps, double* host_cost,
                        int32_t* host_info, void* workspace, size_t ws_bytes, void* stream);

/* ``count`` independent ALS problems in one launch (one workgroup per
 * problem; the ALS loop of als_solver.py:55-82 in each), for throughput work
 * that has many bonds' problems at hand -- the learned-start training data,
 * its evaluation, parameter sweeps.  No counterpart in the reference, which
 * solves one bond at a time.  Problem b's n12 (nD^4), a12g (nD^2 d^2), a1r
 * and a2r (nD D d) follow problem b-1's contiguously.  With plain_start != 0,
 * a1r/a2r are first overwritten with the plain start (als_solver.py:117-146);
 * else they hold the starts.  Results stay on the device: dev_cost[b] the
 * final cost, dev_info[2b] the sweep count, dev_info[2b+1] the flags of
 * acetn_hip_als_solve.  Nothing synchronises. */
size_t acetn_hip_als_batch_workspace(int64_t count, int64_t nD, int64_t D, int64_t d);
int acetn_hip_als_solve_batch(const double* n12, const double* a12g, double* a1r, double* a2r,
                              int64_t count, int64_t nD, int64_t D, int64_t d, int niter,
                              double tol, int method, double epsilon, int plain_start,
                              double* dev_cost, int32_t* dev_info, void* workspace,
                              size_t ws_bytes, void* stream);

/* acetn_hip_als_solve without a workspace argument, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * scratch is a fixed device array of the library (about 133 KB per device,
 * held until the process ends) sized for nD * D <= 128 (D = 8 at d = 2);
 * larger bonds return an error.  Calls that overlap on one device share that
 * array and race: serialise them. */
int acetn_hip_als_solve_noalloc(const double* n12, const double* a12g, double* a1r, double* a2r,
                                int64_t nD, int64_t D, int64_t d, int niter, double tol,
                                int method, double epsilon, int64_t* host_sweeps,
                                double* host_cost, int32_t* host_info, void* stream);

/* finalize_reduced_tensors (full_update.py:123-161; nzyr_inv/nzxr_inv null
 * when the gauge fix is off) + recompose_site_tensors (tensor_update.py:70-75),
 * each tensor divided by its norm and rotated back to the site frame
 * (tensor_update.py:26-38): A1_out, A2_out (D,D,D,D,d). */
size_t acetn_hip_finalize_bond_workspace(int64_t D, int64_t d);
int acetn_hip_finalize_bond(const double* a1r, const double* a2r, const double* nzyr_inv,
                            const double* nzxr_inv, const double* a1q, const double* a2q, int k,
                            int64_t D, int64_t d, double* A1_out, double* A2_out, void* workspace,
                            size_t ws_bytes, void* stream);

/* The two halves of acetn_hip_finalize_bond on their own.  _finalize_reduced
 * is FullUpdater.finalize_reduced_tensors (full_update.py:123-161) on
 * reduced tensors (nD, D, d), any nD with D <= nD d and D <= 32: a1r_out,
 * a2r_out (nD, D, d), with the bond-index sign rule of _finalize_bond.
 * _recompose is TensorUpdater.recompose_site_tensors (tensor_update.py:70-75)
 * from a1q/a2q (D,D,D,D d) and a1r/a2r (D d, D, d); with normalize != 0 each
 * tensor is divided by its norm (full_update.py:61-63); k rotates the result
 * from the bond frame back to the site frame (tensor_update.py:30-32; k = 0
 * leaves it in the bond frame). */
size_t acetn_hip_finalize_reduced_workspace(int64_t nD, int64_t D, int64_t d);
int acetn_hip_finalize_reduced(const double* a1r, const double* a2r, const double* nzyr_inv,
                               const double* nzxr_inv, int64_t nD, int64_t D, int64_t d,
                               double* a1r_out, double* a2r_out, void* workspace,
                               size_t ws_bytes, void* stream);
size_t acetn_hip_recompose_workspace(int64_t D, int64_t d);
int acetn_hip_recompose(const double* a1r, const double* a2r, const double* a1q,
                        const double* a2q, int k, int normalize, int64_t D, int64_t d,
                        double* A1_out, double* A2_out, void* workspace, size_t ws_bytes,
                        void* stream);

/* SimpleUpdater.bond_update, acetn/evolution/simple_update.py:28-76, with
 * LambdaTensors.absorb/strip (lambda_tensors.py:34-49): lam1[leg]/lam2[leg]
 * are the D-vectors on the four legs of s1/s2 (null for the bond's own leg,
 * i.e. all ones); A1_out/A2_out (site frame) and lam_out (D) receive the
 * update.  The bond index's sign is fixed by the rule of
 * acetn_hip_finalize_bond (a gauge choice; the reference leaves it to LAPACK). */
size_t acetn_hip_simple_update_workspace(int64_t D, int64_t d);
int acetn_hip_simple_update_bond(const double* A1, const double* A2, int k,
                                 const double* const lam1[4], const double* const lam2[4],
                                 const double* gate, int64_t D, int64_t d, double cutoff,
                                 double* A1_out, double* A2_out, double* lam_out, void* workspace,
                                 size_t ws_bytes, void* stream);

/* ---- learned ALS start (no reference counterpart) ------------------------
 * A 3-layer perceptron Fin -> hidden -> hidden -> Fout with ReLU, weights in
 * one flat FP32 buffer [W1 (Fin x H), b1, W2 (H x H), b2, W3 (H x Fout), b3]
 * (matrices row-major [in][out]); every product runs on the bf16 matrix
 * cores with FP32 accumulation.  X (batch x Fin) and Y (batch x Fout) are
 * FP32 row-major.  _backward runs the forward pass, the mean-squared loss
 * against target (written to host_loss), and writes the gradient (same
 * layout as params) and dL/dY over Y.  _adam_step is Adam with decoupled
 * bias correction (step >= 1) and L2 weight decay folded into the gradient. */
int64_t acetn_hip_mlp_param_count(int64_t fin, int64_t hidden, int64_t fout);
size_t acetn_hip_mlp_workspace(int64_t fin, int64_t hidden, int64_t fout, int64_t batch);
int acetn_hip_mlp_forward(const float* params, int64_t fin, int64_t hidden, int64_t fout,
                          const float* X, int64_t batch, float* Y, void* workspace,
                          size_t ws_bytes, void* stream);
int acetn_hip_mlp_backward(const float* params, int64_t fin, int64_t hidden, int64_t fout,
                           const float* X, int64_t batch, const float* target, float* Y,
                           float* grad, float* host_loss, void* workspace, size_t ws_bytes,
                           void* stream);
int acetn_hip_mlp_adam_step(float* params, const float* grad, float* m, float* v, int64_t n,
                            int64_t step, float lr, float beta1, float beta2, float eps,
                            float weight_decay, void* stream);

/* ---- building blocks, exported for the parity tests ---------------------- */

/* Gaussian sketch, n x q row-major (the device side of the generator the
 * oracle restates; replaces torch.randn at fused_matmul_svd_lowrank.py:32). */
int acetn_hip_gaussian_sketch(double* out, int64_t n, int64_t q, uint64_t seed,
                              uint64_t sketch_stream, void* stream);

/* C = alpha * op(A) op(B) + beta * C, dense row-major FP64 on the matrix
 * cores (op = transpose when ta/tb != 0).  No reference counterpart: the
 * reference calls torch matmul/einsum. */
size_t acetn_hip_gemm_workspace(int64_t M, int64_t N, int64_t K);
int acetn_hip_gemm(int64_t M, int64_t N, int64_t K, int ta, const double* A, int64_t lda,
                   int tb, const double* B, int64_t ldb, double alpha, double beta, double* C,
                   int64_t ldc, void* workspace, size_t ws_bytes, void* stream);

/* The same product on the bf16 matrix cores with FP32 accumulation, each
 * operand element rounded to FP32 and split into ``pieces`` bf16 terms
 * (2: bf16x3, 3: bf16x6); the workspace query above covers it.  No reference
 * counterpart (the reference's products run in the tensor dtype). */
int acetn_hip_gemm_split(int64_t M, int64_t N, int64_t K, int ta, const double* A, int64_t lda,
                         int tb, const double* B, int64_t ldb, double alpha, double beta, double* C,
                         int64_t ldc, int pieces, void* workspace, size_t ws_bytes, void* stream);

/* One-sided (Hestenes) Jacobi SVD of the rows x cols matrix M whose column j
 * is contiguous at M + j * ldc; max(rows, cols) <= 2048.  Writes the ``rows``
 * leading singular values, descending (zero past the rank), to s and the
 * left singular vectors as the columns of U (rows x rows, row-major);
 * *host_sweeps receives the number of sweeps (negative when the sweep limit
 * stopped it).  Synchronises the stream.  The projector SVD runs the same
 * kernels (torch.linalg.svd at acetn/linalg/fused_matmul_svd_lowrank.py:48
 * and acetn/renormalization/projectors.py:219-236). */
size_t acetn_hip_svd_workspace(int64_t rows, int64_t cols);
int acetn_hip_svd(const double* M, int64_t rows, int64_t cols, int64_t ldc, double* s, double* U,
                  int32_t* host_sweeps, void* workspace, size_t ws_bytes, void* stream);

/* Symmetric eigen-decomposition (descending) of a q x q matrix (G is
 * overwritten).  Replaces torch.linalg.eigh at acetn/evolution/full_update.py:283. */
size_t acetn_hip_eigh_workspace(int64_t q);
int acetn_hip_eigh(double* G, int64_t q, double* w, double* V, void* workspace,
                   size_t ws_bytes, void* stream);

/* Orthonormal basis of the column space of Y (m x q, row-major, in place),
 * shifted Cholesky-QR3 with rank guarding.  Replaces torch.linalg.qr at
 * acetn/linalg/fused_matmul_svd_lowrank.py:35,40. */
size_t acetn_hip_orthonormalize_workspace(int64_t m, int64_t q);
int acetn_hip_orthonormalize(double* Y, int64_t m, int64_t q, void* workspace, size_t ws_bytes,
                             void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ACETN_HIP_H */
