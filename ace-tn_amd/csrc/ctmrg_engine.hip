// CTMRG engine: host-side orchestration of the HIP kernels for one projector
// pair / one boundary renormalisation, plus the extern "C" boundary.
//
// Every operation is written once against a planning context ``Ctx``: a first
// pass in "measure" mode only sizes the scratch buffers (the *_workspace
// queries), the second pass carves them out of the caller's workspace and
// enqueues the kernels on the caller's stream.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace acetn {
namespace {

// ---------------------------------------------------------------------------
// make_quarter_tensor (acetn/renormalization/projectors.py:35-60)
//   qk = C[k] . E[k]          "ab,bcuU->acuU"
//   qk = qk . E[(3+k)%4]      "acuU,ealL->cuUelL"
//   double layer with A_k, conj(A_k) -> "crRedD"
// ---------------------------------------------------------------------------
struct QuarterDims {
  int64_t ca, cb, cc, ce, D, N1, N2;
};
QuarterDims quarter_dims(const acetn_site* s, int k) {
  k &= 3;
  const int64_t D = s->D;
  QuarterDims q{};
  q.ca = s->C_dim[k][0];
  q.cb = s->C_dim[k][1];
  ACETN_REQUIRE(s->E_dim[k][0] == q.cb, "quarter: C[k] / E[k] dimension mismatch");
  q.cc = s->E_dim[k][1];
  const int k1 = (3 + k) & 3;
  ACETN_REQUIRE(s->E_dim[k1][1] == q.ca, "quarter: C[k] / E[k+3] dimension mismatch");
  q.ce = s->E_dim[k1][0];
  q.D = D;
  q.N1 = q.cc * D * D;
  q.N2 = q.ce * D * D;
  return q;
}

// f32 (mixed precision only): T2 and Q are held in FP32, Q as a float array
// of the same element count (the caller's split of Q into bf16 planes reads
// FP32 either way)
void quarter_tensor(Ctx& c, const acetn_site* s, int k, double* Q, bool f32 = false) {
  SplitScope sc(c, 3);
  f32 = f32 && c.mixed;
  const QuarterDims qd = quarter_dims(s, k);
  const int kk = k & 3;
  const int64_t D = qd.D, D2 = D * D;
  double* T1 = c.take((size_t)qd.ca * qd.cc * D2);
  double* T2 = c.take((size_t)qd.N1 * qd.N2);
  // T1[a, (c,u,U)] = sum_b C[a,b] E[k][b,(c,u,U)]
  mm(c, (int)qd.ca, (int)(qd.cc * D2), (int)qd.cb, false, s->C[kk], qd.cb, false, s->E[kk],
     qd.cc * D2, T1, qd.cc * D2);
  // T2[(c,u,U),(e,l,L)] = sum_a T1[a,(c,u,U)] E[k+3][e,a,(l,L)]
  GemmDesc g{};
  g.M = (int)qd.N1; g.N = (int)qd.N2; g.K = (int)qd.ca;
  g.am = plain(1); g.ak = plain(qd.N1);
  g.bk = plain(D2); g.bn = grouped((int)D2, 1, qd.ca * D2);
  g.cm = plain(qd.N2); g.cn = plain(1);
  g.alpha = 1.0; g.beta = 0.0;
  if (f32 && c.launch) {
    GemmBatch b{};
    b.count = 1;
    b.A[0] = T1; b.B[0] = s->E[(3 + kk) & 3]; b.C[0] = T2;
    gemm_split(g, b, c.split, DType::F64, DType::F64, DType::F32, c.gwork, c.gwork_elems, c.st);
  } else {
    gemm(c, g, T1, s->E[(3 + kk) & 3], T2);
  }
  if (!c.launch) return;
  int64_t ps[4];
  permuted_strides(s, kk, ps);
  DLDesc dl{};
  dl.D = (int)D; dl.d = (int)s->d; dl.n1 = (int)qd.cc; dl.n2 = (int)qd.ce;
  // X = T2 viewed as [c][u][U][e][l][L]
  dl.x_o1 = D2 * qd.N2; dl.x_k1 = D * qd.N2; dl.x_b1 = qd.N2;
  dl.x_o2 = D2; dl.x_k2 = D; dl.x_b2 = 1;
  // legs: 1 = u (axis 1 of A_k), 2 = l (axis 0), 3 = r (axis 2), 4 = d (axis 3)
  dl.a_leg[0] = ps[1]; dl.a_leg[1] = ps[0]; dl.a_leg[2] = ps[2]; dl.a_leg[3] = ps[3];
  dl.a_p = 1;
  // Y = Q viewed as [c][r][R][e][d][D]
  dl.y_o1 = D2 * qd.N2; dl.y_k3 = D * qd.N2; dl.y_b3 = qd.N2;
  dl.y_o2 = D2; dl.y_k4 = D; dl.y_b4 = 1;
  dl.x_f32 = dl.y_f32 = f32 ? 1 : 0;
  // mixed precision: the absorption in FP32 on the matrix cores
  if (c.mixed) double_layer_f32(dl, T2, s->A, Q, c.st);
  else double_layer_f64(dl, T2, s->A, Q, c.st);
}

// ---------------------------------------------------------------------------
// Orthonormal basis of range(Y), Y (m x q) row-major: shifted CholeskyQR3
// (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa 2020) with rank guard.
// Result is written back into Y.
//
// A power-iterated sketch is graded far beyond what a Gram matrix resolves:
// its columns carry the small singular directions only some 1e-7 below the
// large ones or further.  The rank guard of the plain pass drops such
// columns (pivot below 1e-14 of the largest); they are then re-derived from
// the shifted pass's output: projected onto the complement of the kept
// columns (G^+ from the kept columns' Gram matrix, so the kept ones need not
// be orthonormal yet), factorised again, and added back.  Each such level
// recovers some seven decades; what is still dropped after kRecoverLevels
// of them (a sketch of a matrix of lower rank) is refilled at random.  Two
// plain passes over the whole basis finish it.  Like
// Householder QR, and unlike refilling every dropped column at random, this
// keeps the directions the rounding-level content of those columns holds,
// which after a power iteration are the next singular directions.  All of it
// is skipped on the device when no column was dropped.
// ---------------------------------------------------------------------------
// Returns the buffer that holds the basis: Y, or, for a basis that only steers
// the power iteration in mixed precision (steer), the output of the single
// shifted CholeskyQR pass (well conditioned, not orthonormal to working
// precision: the next orthonormalisation restores that).
constexpr int kRecoverLevels = 3;

const double* orthonormalize(Ctx& c, double* Y, int64_t m, int64_t q, bool steer = false) {
  SplitScope sc(c, 0);
  const bool single = steer && c.mixed;
  double* G = c.take((size_t)q * q);
  double* Rinv = c.take((size_t)q * q);
  double* Yt = c.take((size_t)m * q);
  int* dead = c.take_int((size_t)q + 2);  // dead[0..q-1] flags, dead[q] = any dead
  int* flags_a = single ? nullptr : c.take_int((size_t)q + 2);
  int* flags_b = single ? nullptr : c.take_int((size_t)q + 2);
  double* T = single ? nullptr : c.take((size_t)q * q);
  double* T2 = single ? nullptr : c.take((size_t)q * q);
  const size_t pb = (size_t)q * (q + 1) / 2;
  double* scr = (pb * 8 > 150 * 1024) ? c.take(pb) : nullptr;
  const double u = 1.1102230246251565e-16;
  const int qi = (int)q, mi = (int)m;
  const int* any = c.launch ? dead + q : nullptr;
  // Gram matrix and inverse factor of src, then dst = src Rinv (or += with acc)
  auto pass = [&](const double* src, double* dst, double shift, double rank_tol, int* dead_out,
                  const int* only_if, const int* run_if, bool acc) {
    GemmDesc gg = dense_desc(qi, qi, mi, true, q, false, q, q);
    gg.run_if = run_if;
    gemm(c, gg, src, src, G);
    if (c.launch) chol_rinv(G, Rinv, qi, shift, rank_tol, c.st, scr, dead_out, only_if);
    GemmDesc gr = dense_desc(mi, qi, qi, false, q, false, q, q, 1.0, acc ? 1.0 : 0.0);
    gr.run_if = run_if;
    gemm(c, gr, src, Rinv, dst);
  };
  // pass 0: shifted (Y -> Yt)
  pass(Y, Yt, 11.0 * ((double)m * q + (double)q * (q + 1)) * u, 0.0, nullptr, nullptr, nullptr,
       false);
  if (single) return Yt;
  // pass 1: plain with rank guard (Yt -> Y); dropped columns come out zero
  pass(Yt, Y, 0.0, 1e-14, dead, nullptr, nullptr, false);
  if (!c.recover) {
    // ACETN_HIP_RF_FAST: dropped columns refilled at random, then the two
    // plain passes (all skipped on the device when none was dropped)
    if (c.launch) refill_dead_columns(Y, m, qi, dead, 0x243F6A8885A308D3ull, c.st);
    pass(Y, Yt, 0.0, 1e-14, nullptr, dead, any, false);
    pass(Yt, Y, 0.0, 1e-14, nullptr, dead, any, false);
    return Y;
  }
  // recovery levels; ``prev`` flags the columns still dropped (level 0: the
  // pass-1 ones), ``next`` receives those a level leaves dropped
  int* prev = dead;
  int* next = flags_a;
  for (int level = 0; level < kRecoverLevels; ++level) {
    const int* on = c.launch ? prev + q : nullptr;
    // the dropped columns (of the pass-0 output, then of the previous
    // level's residual), projected onto the complement of the columns Y
    // holds: Yt <- (I - Y G^+ Y^T) Yt_dropped, G^+ = Rinv Rinv^T
    if (c.launch) keep_dead_columns(Yt, m, qi, prev, c.st);
    GemmDesc gg = dense_desc(qi, qi, mi, true, q, false, q, q);
    gg.run_if = on;
    gemm(c, gg, Y, Y, G);
    if (c.launch) chol_rinv(G, Rinv, qi, 0.0, 1e-14, c.st, scr, nullptr, prev);
    GemmDesc gt = dense_desc(qi, qi, mi, true, q, false, q, q);
    gt.run_if = on;
    gemm(c, gt, Y, Yt, T);                                   // Y^T Yt
    GemmDesc g1 = dense_desc(qi, qi, qi, true, q, false, q, q);
    g1.run_if = on;
    gemm(c, g1, Rinv, T, T2);                                // Rinv^T (.)
    GemmDesc g2 = dense_desc(qi, qi, qi, false, q, false, q, q);
    g2.run_if = on;
    gemm(c, g2, Rinv, T2, T);                                // Rinv (.)
    GemmDesc g3 = dense_desc(mi, qi, qi, false, q, false, q, q, -1.0, 1.0);
    g3.run_if = on;
    gemm(c, g3, Y, T, Yt);                                   // Yt -= Y (.)
    // factorise the projected columns and add the ones with rank back: the
    // other columns of Yt are zero (their Rinv columns too), the dropped
    // ones of Y are zero, so Y += Yt Rinv fills exactly the recovered ones
    pass(Yt, Y, 0.0, 1e-14, next, prev, on, true);
    if (c.launch) still_dead(next, prev, qi, c.st);
    prev = next;
    next = (next == flags_a) ? flags_b : flags_a;
  }
  if (c.launch) refill_dead_columns(Y, m, qi, prev, 0x243F6A8885A308D3ull, c.st);
  // two plain passes over the whole basis (Y -> Yt -> Y)
  pass(Y, Yt, 0.0, 1e-14, nullptr, dead, any, false);
  pass(Yt, Y, 0.0, 1e-14, nullptr, dead, any, false);
  return Y;
}

// ---------------------------------------------------------------------------
// Singular values / left vectors of B (q x n) given as BT = B^T (n x q):
// two passes of Gram + Jacobi (the second on the rotated, nearly diagonal
// Gram matrix restores absolute accuracy eps*s0 for the small values).
// Outputs: w (q, descending eigenvalues of B B^T), UB (q x q).
// ---------------------------------------------------------------------------
void small_svd(Ctx& c, const double* BT, int64_t n, int64_t q, double* w, double* UB) {
  SplitScope sc(c, 0);
  double* G = c.take((size_t)q * q);
  double* W1 = c.take((size_t)q * q);
  double* W2 = c.take((size_t)q * q);
  double* w1 = c.take((size_t)q);
  double* B2T = c.take((size_t)n * q);
  double* rot = c.take(jacobi_rot_elems((int)q));
  mm(c, (int)q, (int)q, (int)n, true, BT, q, false, BT, q, G, q);
  // pass 1 only preconditions (pass 2 converges quadratically from ~1e-9)
  if (c.launch) jacobi_eigh(G, w1, W1, rot, (int)q, c.st, 1e-9);
  mm(c, (int)n, (int)q, (int)q, false, BT, q, false, W1, q, B2T, q);
  mm(c, (int)q, (int)q, (int)n, true, B2T, q, false, B2T, q, G, q);
  if (c.launch) jacobi_eigh(G, w, W2, rot, (int)q, c.st);
  mm(c, (int)q, (int)q, (int)q, false, W1, q, false, W2, q, UB, q);
}

// ---------------------------------------------------------------------------
// The randomized SVD of a product M = M_0 M_1 ... M_{r-1} of dense row-major
// FP64 matrices (M_i: dims[i] x dims[i+1]) without forming it: svd_lowrank
// (r = 1, linalg/svd_lowrank.py:4-44), fused_matmul_svd_lowrank (r = 2,
// linalg/fused_matmul_svd_lowrank.py:4-52) and fused_3matmul_svd_lowrank
// (r = 4 with an orthonormalisation between the two halves of every power
// iteration, mid_qr, linalg/fused_3matmul_svd_lowrank.py:4-56).  The sketch
// is the Philox Gaussian of (seed, sstream); QR is CholeskyQR3 (the same
// spans as Householder QR); the SVD of the projected q x n matrix is the
// two-pass Gram + Jacobi of small_svd.  Outputs U (m x q), S (q, descending)
// and V (n x q) with M ~ U diag(S) V^T; V = orthonormalised M^T U, whose
// columns are V diag(S), so directions of numerically zero singular values
// still come out orthonormal.
// ---------------------------------------------------------------------------
void rsvd_product(Ctx& c, int r, const double* const* ops, const int64_t* dims, int64_t q,
                  int niter, bool mid_qr, uint64_t seed, uint64_t sstream, double* U, double* S,
                  double* V) {
  SplitScope sc(c, 0);
  const int64_t m = dims[0], n = dims[r];
  int64_t wmax = 0;
  for (int i = 0; i <= r; ++i) wmax = std::max(wmax, dims[i]);
  double* Om = c.take((size_t)n * q);
  double* T0 = c.take((size_t)wmax * q);
  double* T1 = c.take((size_t)wmax * q);
  double* Y = c.take((size_t)m * q);
  double* Z = c.take((size_t)n * q);
  double* w = c.take((size_t)q);
  double* UB = c.take((size_t)q * q);
  int* cd = c.take_int(2);
  // X (dims[r] x q) -> M X (dims[0] x q), right to left
  auto apply = [&](const double* X, double* out) {
    const double* cur = X;
    for (int i = r - 1; i >= 0; --i) {
      double* dst = (i == 0) ? out : (((r - 1 - i) & 1) ? T1 : T0);
      mm(c, (int)dims[i], (int)q, (int)dims[i + 1], false, ops[i], dims[i + 1], false, cur, q, dst,
         q);
      cur = dst;
    }
  };
  // X (dims[0] x q) -> M^T X (dims[r] x q), left to right
  auto apply_t = [&](const double* X, double* out) {
    const double* cur = X;
    for (int i = 0; i < r; ++i) {
      double* dst = (i == r - 1) ? out : ((i & 1) ? T1 : T0);
      mm(c, (int)dims[i + 1], (int)q, (int)dims[i], true, ops[i], dims[i + 1], false, cur, q, dst,
         q);
      cur = dst;
    }
  };
  if (c.launch) gaussian_sketch(Om, n, q, seed, sstream, c.st);
  apply(Om, Y);
  for (int it = 0; it < niter; ++it) {
    const double* Yb = orthonormalize(c, Y, m, q);
    apply_t(Yb, Z);
    if (mid_qr) {
      const double* Zb = orthonormalize(c, Z, n, q);
      apply(Zb, Y);
    } else {
      apply(Z, Y);
    }
  }
  const double* Qm = orthonormalize(c, Y, m, q);
  apply_t(Qm, Z);                              // Z = M^T Q = B^T
  small_svd(c, Z, n, q, w, UB);
  if (c.launch) count_kept(w, (int)q, (int)q, 0.0, S, cd, c.st);  // S = sqrt(max(w, 0))
  mm(c, (int)m, (int)q, (int)q, false, Qm, q, false, UB, q, U, q);
  mm(c, (int)n, (int)q, (int)q, false, Z, q, false, UB, q, V, q);
  const double* Vo = orthonormalize(c, V, n, q);
  if (Vo != V && c.launch)
    ACETN_HIP_CHECK(hipMemcpyAsync(V, Vo, sizeof(double) * n * q, hipMemcpyDeviceToDevice, c.st));
}

// ---------------------------------------------------------------------------
// calculate_half_system, rsvd branch (projectors.py:138-174) with
// fused_matmul_svd_lowrank (linalg/fused_matmul_svd_lowrank.py:4-52).
// ---------------------------------------------------------------------------
struct HalfOut {
  double* P1;
  double* P2;
  int64_t* host_cD;   // phase 0: written after a stream sync
  int* pinned_cD;     // phase 1: async copy target (pinned host memory), or null
  int cD_in;          // phase 2: cD_new read by the host after phase 1 (padded: min(chi, q))
  double* spectrum;
  // graph-capturable form: the sketch's Philox stream is sstream +
  // stream_base[0] (device); phase 1 also copies cD_new to dev_cD (device);
  // phase 2 writes cD_in columns, those at and past the device count zero
  const uint64_t* stream_base = nullptr;
  int32_t* dev_cD = nullptr;
  bool padded = false;
  // phase 1: recorded on the stream once the matrix-core work of the range
  // finder is enqueued, ahead of the one-workgroup eigensolver
  hipEvent_t mark = nullptr;
};

// A product with an enlarged corner as its A operand: from the corner's bf16
// planes in mixed precision, from the FP64 matrix otherwise.  In mixed
// precision a product whose result only feeds the next corner product as its
// B operand hands it over as planes (Cout, then that product's Bin): the
// split-K reduction writes the pieces directly and C is not stored.
void mm_corner(Ctx& c, const SplitPlanes& P, int M, int N, int K, bool ta, const double* A,
               int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
               const SplitPlanes* Bin = nullptr, SplitPlanes* Cout = nullptr) {
  const GemmDesc g = dense_desc(M, N, K, ta, lda, false, ldb, ldc);
  if (c.run && c.launch && c.split > 0 && P.p != nullptr) {
    if (Cout) {
      Cout->rows = N;
      Cout->kp = split_planes_kp(M);
      Cout->pieces = c.split;
    }
    gemm_split_planar(g, P, B, C, c.split, c.gwork, c.gwork_elems, c.st,
                      Bin && Bin->p ? Bin : nullptr, Cout && Cout->p ? Cout : nullptr);
    return;
  }
  gemm(c, g, A, B, C);
}

// phase 0: everything with one host sync; phase 1: through the truncation
// count (async copy of cD_new to pinned host memory); phase 2: the projector
// products, given cD_new.  Phases 1 and 2 carve identical workspace layouts.
// The workspace holds the bf16 planes of both corners in both orientations
// whatever the precision, so one workspace query serves both.
void half_projectors(Ctx& c, const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                     int64_t oversampling, int niter, double cutoff, uint64_t seed,
                     uint64_t sstream, const HalfOut& out, int phase = 0) {
  const QuarterDims d1 = quarter_dims(s1, k);
  const QuarterDims d4 = quarter_dims(s4, k + 3);
  ACETN_REQUIRE(d1.N2 == d4.N1, "half system: Q1 columns != Q4 rows");
  const int64_t m = d1.N1, kk = d1.N2, n = d4.N2;
  int64_t q = std::min<int64_t>(chi + oversampling, std::min(m, n));
  ACETN_REQUIRE(q >= 1 && q <= 510, "half system: sketch width out of range");
  double* Q1 = c.take((size_t)m * kk);
  double* Q4 = c.take((size_t)kk * n);
  double* Om = c.take((size_t)n * q);
  double* Y = c.take((size_t)m * q);
  double* W = c.take((size_t)kk * q);
  double* Z = c.take((size_t)n * q);
  double* Wq = c.take((size_t)kk * q);
  double* BT = c.take((size_t)n * q);
  double* w = c.take((size_t)q);
  double* UB = c.take((size_t)q * q);
  double* s = c.take((size_t)q);
  int* cD_dev = c.take_int(2);
  double* X = c.take((size_t)q * q);
  double* Vp = c.take((size_t)n * q);
  SplitPlanes q1n, q1t, q4n, q4t;
  q1n.p = c.take(split_planes_elems((int)m, (int)kk));
  q1t.p = c.take(split_planes_elems((int)kk, (int)m));
  q4n.p = c.take(split_planes_elems((int)kk, (int)n));
  q4t.p = c.take(split_planes_elems((int)n, (int)kk));
  // the range finder's intermediates W (kk x q) and Z (n x q) as B planes
  SplitPlanes wp, zp;
  wp.p = c.take(split_planes_elems((int)q, (int)kk));
  zp.p = c.take(split_planes_elems((int)q, (int)n));

  const bool replay = c.run && phase == 2;
  if (replay) c.launch = false;
  // mixed precision: Q1 and Q4 only feed the bf16 plane split, so they stay FP32
  quarter_tensor(c, s1, k, Q1, true);
  quarter_tensor(c, s4, k + 3, Q4, true);
  if (c.mixed) {
    if (c.launch) {
      split_transpose(Q1, (int)m, (int)kk, kk, q1n, q1t, c.st, true);
      split_transpose(Q4, (int)kk, (int)n, n, q4n, q4t, c.st, true);
    } else if (c.run) {  // phase 2 replay: the planes are in the workspace already
      q1n.rows = (int)m; q1n.kp = (int)((kk + 31) / 32 * 32); q1n.pieces = 3;
      q1t.rows = (int)kk; q1t.kp = (int)((m + 31) / 32 * 32); q1t.pieces = 3;
      q4n.rows = (int)kk; q4n.kp = (int)((n + 31) / 32 * 32); q4n.pieces = 3;
      q4t.rows = (int)n; q4t.kp = (int)((kk + 31) / 32 * 32); q4t.pieces = 3;
    }
  } else {
    q1n.p = q1t.p = q4n.p = q4t.p = nullptr;
  }
  if (c.launch) gaussian_sketch(Om, n, q, seed, sstream, c.st, out.stream_base);
  // mixed precision: the sketch and the power iterations only steer the range
  // finder (bf16x3); the products after the final orthonormalisation carry the
  // basis itself (bf16x6)
  SplitScope sc(c, 2);
  // Y = Q1 (Q4 Omega)
  mm_corner(c, q4n, (int)kk, (int)q, (int)n, false, Q4, n, Om, q, W, q, nullptr, &wp);
  mm_corner(c, q1n, (int)m, (int)q, (int)kk, false, Q1, kk, W, q, Y, q, &wp);
  for (int it = 0; it < niter; ++it) {
    const double* Yb = orthonormalize(c, Y, m, q, true);
    mm_corner(c, q1t, (int)kk, (int)q, (int)m, true, Q1, kk, Yb, q, W, q, nullptr, &wp);  // Q1^H Y
    mm_corner(c, q4t, (int)n, (int)q, (int)kk, true, Q4, n, W, q, Z, q, &wp, &zp);       // Q4^H (.)
    mm_corner(c, q4n, (int)kk, (int)q, (int)n, false, Q4, n, Z, q, W, q, &zp, &wp);      // Q4 (.)
    mm_corner(c, q1n, (int)m, (int)q, (int)kk, false, Q1, kk, W, q, Y, q, &wp);          // Q1 (.)
  }
  c.split = c.mixed ? 3 : 0;
  orthonormalize(c, Y, m, q);
  // B^T = Q4^H (Q1^H Qm): n x q ; keep Wq = Q1^H Qm for P1
  mm_corner(c, q1t, (int)kk, (int)q, (int)m, true, Q1, kk, Y, q, Wq, q);
  mm_corner(c, q4t, (int)n, (int)q, (int)kk, true, Q4, n, Wq, q, BT, q);
  if (out.mark && c.launch) ACETN_HIP_CHECK(hipEventRecord(out.mark, c.st));
  small_svd(c, BT, n, q, w, UB);
  if (!c.run) {
    // size the split-K scratch of the post-truncation products at cD = chi
    const int cDp = (int)std::min<int64_t>(chi, q);
    for (int cd = 1; cd <= cDp; cd = (cd == cDp) ? cDp + 1 : std::min(2 * cd, cDp)) {
      mm(c, (int)kk, cd, (int)q, false, nullptr, q, false, nullptr, cd, nullptr, cd);
      mm(c, (int)n, cd, (int)q, false, nullptr, q, false, nullptr, cd, nullptr, cd);
      mm(c, (int)kk, cd, (int)n, false, nullptr, n, false, nullptr, cd, nullptr, cd);
    }
    return;
  }
  int cD = 0;
  if (phase == 2) {
    c.launch = true;
    cD = out.cD_in;
  } else {
    count_kept(w, (int)q, (int)chi, cutoff, s, cD_dev, c.st);
    if (out.spectrum) ACETN_HIP_CHECK(hipMemcpyAsync(out.spectrum, s, sizeof(double) * q, hipMemcpyDeviceToDevice, c.st));
    if (phase == 1) {
      if (out.pinned_cD)
        ACETN_HIP_CHECK(hipMemcpyAsync(out.pinned_cD, cD_dev, sizeof(int), hipMemcpyDeviceToHost, c.st));
      if (out.dev_cD)
        ACETN_HIP_CHECK(hipMemcpyAsync(out.dev_cD, cD_dev, sizeof(int), hipMemcpyDeviceToDevice, c.st));
      return;
    }
    ACETN_HIP_CHECK(hipMemcpyAsync(&cD, cD_dev, sizeof(int), hipMemcpyDeviceToHost, c.st));
    ACETN_HIP_CHECK(hipStreamSynchronize(c.st));
    *out.host_cD = cD;
  }
  ACETN_REQUIRE(cD >= 1 && cD <= std::min<int64_t>(chi, q), "half system: no singular value above the cutoff");
  ACETN_REQUIRE(!out.padded || cD == std::min<int64_t>(chi, q),
                "projectors_end_padded: the column count must be min(chi, q)");
  // P1 = Q1^H U S_n^-1/2 = Wq (UB S_n^-1/2)
  c.split = 0;
  const int* kept = out.padded ? cD_dev : nullptr;
  scale_columns(UB, (int)q, s, cD, 0, X, c.st, kept);
  mm(c, (int)kk, cD, (int)q, false, Wq, q, false, X, cD, out.P1, cD);
  // P2 = Q4 V S_n^-1/2,  V = B^T UB S^-1
  scale_columns(UB, (int)q, s, cD, 1, X, c.st, kept);
  mm(c, (int)n, cD, (int)q, false, BT, q, false, X, cD, Vp, cD);
  c.split = c.mixed ? 3 : 0;
  mm_corner(c, q4n, (int)kk, cD, (int)n, false, Q4, n, Vp, cD, out.P2, cD);
  fix_projector_gauge(out.P1, out.P2, kk, kk, cD, c.st);
}

// ---------------------------------------------------------------------------
// The other projector branches of ProjectorCalculator (projectors.py:62-136,
// :176-217): the full-system rSVD of Q2 Q1 Q4 Q3 (fused_3matmul_svd_lowrank,
// linalg/fused_3matmul_svd_lowrank.py:4-56, with its QR between the two
// halves of each power iteration) and the full-rank SVD of the half-system
// (Q1 Q4) or full-system (Q2 Q1)(Q4 Q3) matrix.  Written against the left
// operator L = Q1 or Q2 Q1 (a x c) and the right operator R = Q4 or Q4 Q3
// (c x f): P1 = L^T U S^-1/2, P2 = R V S^-1/2, V = M^T U S^-1 with M = L R.
// The full-rank branch uses the two-pass Jacobi SVD, so min(a, f) <= 510.
// ---------------------------------------------------------------------------
enum { kSysHalf = 0, kSysFull = 1 };
enum { kSvdRsvd = 0, kSvdFullRank = 1 };

void general_projectors(Ctx& c, const acetn_site* const* st, int k, int system, int svd,
                        int64_t chi, int64_t oversampling, int niter, double cutoff,
                        uint64_t seed, uint64_t sstream, const HalfOut& out, int phase = 0) {
  const bool full = system == kSysFull;
  const acetn_site* s1 = st[0];
  const acetn_site* s4 = st[3];
  const QuarterDims d1 = quarter_dims(s1, k);
  const QuarterDims d4 = quarter_dims(s4, k + 3);
  ACETN_REQUIRE(d1.N2 == d4.N1, "projectors: Q1 columns != Q4 rows");
  QuarterDims d2{}, d3{};
  if (full) {
    d2 = quarter_dims(st[1], k + 1);
    d3 = quarter_dims(st[2], k + 2);
    ACETN_REQUIRE(d2.N2 == d1.N1, "full system: Q2 columns != Q1 rows");
    ACETN_REQUIRE(d4.N2 == d3.N1, "full system: Q4 columns != Q3 rows");
  }
  // L = [Q2] Q1 : a x c ; R = Q4 [Q3] : c x f ; b = Q1 rows, e = Q4 columns
  const int64_t b = d1.N1, cc = d1.N2, e = d4.N2;
  const int64_t a = full ? d2.N1 : b, f = full ? d3.N2 : e;
  int64_t q = (svd == kSvdRsvd) ? std::min<int64_t>(chi + oversampling, std::min(a, f))
                                : std::min(a, f);
  ACETN_REQUIRE(q >= 1 && q <= 510, svd == kSvdRsvd
                                        ? "projectors: sketch width out of range"
                                        : "projectors: the full-rank SVD on the device needs "
                                          "min(rows, cols) <= 510 (use svd_type='rsvd')");
  if (svd == kSvdFullRank) q = a;  // all left singular vectors
  ACETN_REQUIRE(q <= 510, "projectors: the full-rank SVD on the device needs rows <= 510");
  double* Q1 = c.take((size_t)b * cc);
  double* Q4 = c.take((size_t)cc * e);
  double* Q2 = full ? c.take((size_t)a * b) : nullptr;
  double* Q3 = full ? c.take((size_t)e * f) : nullptr;
  const int64_t wid = std::max<int64_t>(q, chi);
  double* Om = c.take((size_t)f * q);
  double* Y = c.take((size_t)a * q);
  double* Z = c.take((size_t)f * q);
  double* W = c.take((size_t)cc * wid);
  double* W2 = c.take((size_t)cc * wid);
  double* tb = c.take((size_t)b * wid);
  double* te = c.take((size_t)e * wid);
  double* BT = c.take((size_t)f * q);
  double* w = c.take((size_t)q);
  double* UB = c.take((size_t)q * q);
  double* sv = c.take((size_t)q);
  int* cD_dev = c.take_int(2);
  double* X = c.take((size_t)q * wid);
  double* Vp = c.take((size_t)f * wid);
  double* Lm = (svd == kSvdFullRank && full) ? c.take((size_t)a * cc) : nullptr;
  double* Rm = (svd == kSvdFullRank && full) ? c.take((size_t)cc * f) : nullptr;

  const bool replay = c.run && phase == 2;
  if (replay) c.launch = false;
  quarter_tensor(c, s1, k, Q1);
  quarter_tensor(c, s4, k + 3, Q4);
  if (full) {
    quarter_tensor(c, st[1], k + 1, Q2);
    quarter_tensor(c, st[2], k + 2, Q3);
  }
  // the four operator applications on n-column blocks
  auto apply_R = [&](const double* Xin, int n, double* o) {   // f x n -> c x n
    if (full) {
      mm(c, (int)e, n, (int)f, false, Q3, f, false, Xin, n, te, n);
      mm(c, (int)cc, n, (int)e, false, Q4, e, false, te, n, o, n);
    } else {
      mm(c, (int)cc, n, (int)e, false, Q4, e, false, Xin, n, o, n);
    }
  };
  auto apply_L = [&](const double* Xin, int n, double* o) {   // c x n -> a x n
    if (full) {
      mm(c, (int)b, n, (int)cc, false, Q1, cc, false, Xin, n, tb, n);
      mm(c, (int)a, n, (int)b, false, Q2, b, false, tb, n, o, n);
    } else {
      mm(c, (int)b, n, (int)cc, false, Q1, cc, false, Xin, n, o, n);
    }
  };
  auto apply_LT = [&](const double* Xin, int n, double* o) {  // a x n -> c x n
    if (full) {
      mm(c, (int)b, n, (int)a, true, Q2, b, false, Xin, n, tb, n);
      mm(c, (int)cc, n, (int)b, true, Q1, cc, false, tb, n, o, n);
    } else {
      mm(c, (int)cc, n, (int)b, true, Q1, cc, false, Xin, n, o, n);
    }
  };
  auto apply_RT = [&](const double* Xin, int n, double* o) {  // c x n -> f x n
    if (full) {
      mm(c, (int)e, n, (int)cc, true, Q4, e, false, Xin, n, te, n);
      mm(c, (int)f, n, (int)e, true, Q3, f, false, te, n, o, n);
    } else {
      mm(c, (int)e, n, (int)cc, true, Q4, e, false, Xin, n, o, n);
    }
  };
  const int qi = (int)q;
  SplitScope sc(c, 2);
  if (svd == kSvdRsvd) {
    if (c.launch) gaussian_sketch(Om, f, q, seed, sstream, c.st, out.stream_base);
    apply_R(Om, qi, W);
    apply_L(W, qi, Y);
    for (int it = 0; it < niter; ++it) {
      if (it >= 1) c.split = c.mixed ? 3 : 0;
      orthonormalize(c, Y, a, q);
      apply_LT(Y, qi, W);
      apply_RT(W, qi, Z);
      if (full) orthonormalize(c, Z, f, q);
      apply_R(Z, qi, W2);
      apply_L(W2, qi, Y);
    }
    c.split = c.mixed ? 3 : 0;
    orthonormalize(c, Y, a, q);
    apply_LT(Y, qi, W);   // W = L^T Qm, kept for P1
    apply_RT(W, qi, BT);
  } else {
    // BT = (L R)^T = R^T L^T
    c.split = c.mixed ? 3 : 0;
    const double* Lp = Q1;
    const double* Rp = Q4;
    if (full) {
      mm(c, (int)a, (int)cc, (int)b, false, Q2, b, false, Q1, cc, Lm, cc);
      mm(c, (int)cc, (int)f, (int)e, false, Q4, e, false, Q3, f, Rm, f);
      Lp = Lm;
      Rp = Rm;
    }
    mm(c, (int)f, (int)a, (int)cc, true, Rp, f, true, Lp, cc, BT, a);
  }
  if (out.mark && c.launch) ACETN_HIP_CHECK(hipEventRecord(out.mark, c.st));
  small_svd(c, BT, f, q, w, UB);
  if (!c.run) {
    const int cDp = (int)std::min<int64_t>(chi, q);
    for (int cd = 1; cd <= cDp; cd = (cd == cDp) ? cDp + 1 : std::min(2 * cd, cDp)) {
      mm(c, (int)cc, cd, qi, false, nullptr, q, false, nullptr, cd, nullptr, cd);
      apply_LT(nullptr, cd, nullptr);
      mm(c, (int)f, cd, qi, false, nullptr, q, false, nullptr, cd, nullptr, cd);
      apply_R(nullptr, cd, nullptr);
    }
    return;
  }
  int cD = 0;
  if (phase == 2) {
    c.launch = true;
    cD = out.cD_in;
  } else {
    count_kept(w, qi, (int)chi, cutoff, sv, cD_dev, c.st);
    if (out.spectrum) ACETN_HIP_CHECK(hipMemcpyAsync(out.spectrum, sv, sizeof(double) * q, hipMemcpyDeviceToDevice, c.st));
    if (phase == 1) {
      if (out.pinned_cD)
        ACETN_HIP_CHECK(hipMemcpyAsync(out.pinned_cD, cD_dev, sizeof(int), hipMemcpyDeviceToHost, c.st));
      if (out.dev_cD)
        ACETN_HIP_CHECK(hipMemcpyAsync(out.dev_cD, cD_dev, sizeof(int), hipMemcpyDeviceToDevice, c.st));
      return;
    }
    ACETN_HIP_CHECK(hipMemcpyAsync(&cD, cD_dev, sizeof(int), hipMemcpyDeviceToHost, c.st));
    ACETN_HIP_CHECK(hipStreamSynchronize(c.st));
    *out.host_cD = cD;
  }
  ACETN_REQUIRE(cD >= 1 && cD <= std::min<int64_t>(chi, q), "projectors: no singular value above the cutoff");
  ACETN_REQUIRE(!out.padded || cD == std::min<int64_t>(chi, q),
                "projectors_end_padded: the column count must be min(chi, q)");
  c.split = 0;
  // P1 = L^T U S^-1/2 (U = Qm UB in the rSVD branch, UB in the full-rank one)
  const int* kept = out.padded ? cD_dev : nullptr;
  scale_columns(UB, qi, sv, cD, 0, X, c.st, kept);
  if (svd == kSvdRsvd) {
    mm(c, (int)cc, cD, qi, false, W, q, false, X, cD, out.P1, cD);
  } else {
    c.split = c.mixed ? 3 : 0;
    apply_LT(X, cD, out.P1);
    c.split = 0;
  }
  // P2 = R V S^-1/2, V = B^T UB S^-1
  scale_columns(UB, qi, sv, cD, 1, X, c.st, kept);
  mm(c, (int)f, cD, qi, false, BT, q, false, X, cD, Vp, cD);
  c.split = c.mixed ? 3 : 0;
  apply_R(Vp, cD, out.P2);
  fix_projector_gauge(out.P1, out.P2, cc, cc, cD, c.st);
}

// ---------------------------------------------------------------------------
// renormalize_boundary (directional_mover.py:273-366)
// ---------------------------------------------------------------------------
// part 0: everything in one call.  Parts 1 and 2 split it by projector pair:
// part 1 is what the pair named by ``first`` feeds on its own (0: pair j --
// cj2 and ej up to its double layer; 1: pair i -- cj1 and ej up to its double
// layer, ej then contracted with P2_i first), part 2 the rest.  Both parts
// carve the same workspace, sized with cD_layout (>= the other pair's count)
// for the count not yet known at part 1, and part 2 finds ej's intermediate
// where part 1 left it.
void renormalize(Ctx& c, const acetn_site* s1, int k, const double* P1i, const double* P2i,
                 int64_t cDi, const double* P1j, const double* P2j, int64_t cDj, double* Ca,
                 double* Cb, double* Eo, int part = 0, int64_t cD_layout = 0, int first = 0) {
  SplitScope sc(c, 3);
  k &= 3;
  const int64_t D = s1->D, D2 = D * D;
  if (!c.run && cD_layout > 0) (first == 0 ? cDi : cDj) = cD_layout;
  const bool launch = c.launch;
  // whether a block that depends on pair i (dep_i) or on pair j runs in this call
  auto runs = [&](bool dep_i) { return launch && (part == 0 || (part == 1) == (dep_i == (first == 1))); };
  const int route = part == 0 ? 0 : first;
  double* nrm = c.take(1024);
  // ---- cj1: C[(3+k)] with E[(2+k)] and P1_i
  c.launch = runs(true);
  {
    const int ke = (2 + k) & 3, kc = (3 + k) & 3;
    const int64_t xa = s1->E_dim[ke][0], xb = s1->E_dim[ke][1];
    ACETN_REQUIRE(s1->C_dim[kc][0] == xb, "cj1: E/C dimension mismatch");
    const int64_t xc = s1->C_dim[kc][1];
    double* T = c.take((size_t)xa * D2 * xc);
    GemmDesc g{};
    g.M = (int)(xa * D2); g.N = (int)xc; g.K = (int)xb;
    g.am = grouped((int)D2, 1, xb * D2); g.ak = plain(D2);
    g.bk = plain(xc); g.bn = plain(1);
    g.cm = plain(xc); g.cn = plain(1);
    g.alpha = 1.0; g.beta = 0.0;
    gemm(c, g, s1->E[ke], s1->C[kc], T);
    GemmDesc h{};
    h.M = (int)xa; h.N = (int)cDi; h.K = (int)(xc * D2);
    h.am = plain(D2 * xc); h.ak = grouped((int)D2, xc, 1);
    h.bk = plain(cDi); h.bn = plain(1);
    h.cm = plain(cDi); h.cn = plain(1);
    h.alpha = 1.0; h.beta = 0.0;
    gemm(c, h, T, P1i, Ca);
    if (c.launch) frobenius_normalize(Ca, xa * cDi, nrm, c.st);
  }
  // ---- cj2: C[k] with E[k] and P2_j
  c.launch = runs(false);
  {
    const int64_t xa = s1->C_dim[k][0], xb = s1->C_dim[k][1];
    ACETN_REQUIRE(s1->E_dim[k][0] == xb, "cj2: C/E dimension mismatch");
    const int64_t xc = s1->E_dim[k][1];
    double* T = c.take((size_t)xa * xc * D2);
    mm(c, (int)xa, (int)(xc * D2), (int)xb, false, s1->C[k], xb, false, s1->E[k], xc * D2, T,
       xc * D2);
    GemmDesc h{};
    h.M = (int)cDj; h.N = (int)xc; h.K = (int)(xa * D2);
    h.am = plain(1); h.ak = plain(cDj);
    h.bk = grouped((int)D2, 1, xc * D2); h.bn = plain(D2);
    h.cm = plain(xc); h.cn = plain(1);
    h.alpha = 1.0; h.beta = 0.0;
    gemm(c, h, P2j, T, Cb);
    if (c.launch) frobenius_normalize(Cb, cDj * xc, nrm, c.st);
  }
  // ---- ej: E[(3+k)] with A_k, P1_j, P2_i.  Route 0 (the reference's order)
  // contracts E with P1_j first, route 1 with P2_i first; the double layer
  // runs on that first intermediate and the other projector comes last.
  c.launch = launch && part != 2;
  {
    const int ke = (3 + k) & 3;
    const int64_t xa = s1->E_dim[ke][0], xb = s1->E_dim[ke][1];
    const bool r1 = route == 1;
    const int64_t xf = r1 ? xb : xa;    // E index that stays through the double layer
    const int64_t xs = r1 ? xa : xb;    // E index the first projector contracts
    const int64_t cf = r1 ? cDi : cDj;  // columns of the first projector
    const int64_t cs = r1 ? cDj : cDi;  // columns of the last
    const double* Pf = r1 ? P2i : P1j;
    const double* Ps = r1 ? P1j : P2i;
    double* T = c.take((size_t)xf * D2 * D2 * cf);
    double* Zt = c.take((size_t)cf * D2 * xf * D2);
    // T[f][(e,E)][c][(p,P)] (E's own D legs (e,E), the first projector's
    // (p,P)): the columns run (c, (p,P)), so each block of the double layer
    // below, (f, c), holds its (p,P) pairs contiguously and the blocks of a
    // row follow each other (coalesced X loads)
    GemmDesc g{};
    g.M = (int)(xf * D2); g.N = (int)(D2 * cf); g.K = (int)xs;
    g.am = r1 ? plain(1) : grouped((int)D2, 1, xb * D2);
    g.ak = r1 ? plain(xb * D2) : plain(D2);
    g.bk = plain(D2 * cf); g.bn = grouped((int)D2, cf, 1);
    g.cm = plain(D2 * cf); g.cn = plain(1);
    g.alpha = 1.0; g.beta = 0.0;
    // E_o[i][j][r][R] = Z[(c,r,R)][(f,q,Q)] . Ps[(f,q,Q)][last column]
    GemmDesc h{};
    h.M = (int)(cf * D2); h.N = (int)cs; h.K = (int)(xf * D2);
    h.am = plain(xf * D2); h.ak = plain(1);
    h.bk = plain(cs); h.bn = plain(1);
    h.cm = r1 ? grouped((int)D2, 1, cDj * D2) : plain(1);
    h.cn = r1 ? plain(D2) : plain(cDj * D2);
    h.alpha = 1.0; h.beta = 0.0;
    // mixed precision: T only feeds the FP32 double layer, so it is stored in
    // FP32, and so is Z when the last product splits it into planes.  Z is
    // written by part 1 and read by part 2, which sees the real count where
    // part 1 saw its layout bound: the path (and with it Z's element type)
    // is chosen from the bound in both parts.
    GemmDesc hp = h;
    if (part != 0) hp.N = (int)cD_layout;
    const bool t32 = c.mixed, z32 = c.mixed && gemm_split_is_planar(hp);
    h.planar = z32 ? 1 : 0;
    h.a_f32 = z32;
    if (t32 && c.launch) {
      GemmBatch b{};
      b.count = 1;
      b.A[0] = s1->E[ke]; b.B[0] = Pf; b.C[0] = T;
      gemm_split(g, b, c.split, DType::F64, DType::F64, DType::F32, c.gwork, c.gwork_elems, c.st);
    } else {
      gemm(c, g, s1->E[ke], Pf, T);
    }
    if (c.launch) {
      int64_t ps[4];
      permuted_strides(s1, k, ps);
      DLDesc dl{};
      dl.D = (int)D; dl.d = (int)s1->d; dl.n1 = (int)xf; dl.n2 = (int)cf;
      const int64_t ld = D2 * cf;  // row length of T
      dl.x_o1 = D2 * ld; dl.x_k1 = D * ld; dl.x_b1 = ld;
      dl.x_o2 = D2; dl.x_k2 = D; dl.x_b2 = 1;
      // A legs: E's D leg (axis 0), the first projector's (axis 1 for P1_j,
      // axis 3 for P2_i), the open r (axis 2), the last projector's
      dl.a_leg[0] = ps[0]; dl.a_leg[1] = r1 ? ps[3] : ps[1];
      dl.a_leg[2] = ps[2]; dl.a_leg[3] = r1 ? ps[1] : ps[3];
      dl.a_p = 1;
      // Z[c][r][R][f][q][Q]
      const int64_t rowz = xf * D2;
      dl.y_o2 = D2 * rowz; dl.y_k3 = D * rowz; dl.y_b3 = rowz;
      dl.y_o1 = D2; dl.y_k4 = D; dl.y_b4 = 1;
      dl.x_f32 = t32 ? 1 : 0;
      dl.y_f32 = z32 ? 1 : 0;
      if (c.mixed) double_layer_f32(dl, T, s1->A, Zt, c.st);
      else double_layer_f64(dl, T, s1->A, Zt, c.st);
    }
    c.launch = launch && part != 1;
    gemm(c, h, Zt, Ps, Eo);
    if (c.launch) frobenius_normalize(Eo, cDi * cDj * D2, nrm, c.st);
  }
  c.launch = launch;
}

// ---------------------------------------------------------------------------
// Reduced density matrices (acetn/measurement/rdm.py)
// ---------------------------------------------------------------------------

// build_site_rdm, rdm.py:35-67.  rho (d x d) = "Pp".
void site_rdm(Ctx& c, const acetn_site* s, double* rho) {
  const int64_t D = s->D, d = s->d, D2 = D * D;
  const int64_t a0 = s->C_dim[3][0], b0 = s->C_dim[3][1];           // c4 (a,b)
  ACETN_REQUIRE(s->E_dim[3][0] == b0, "site rdm: c4/e4 mismatch");
  const int64_t c4 = s->E_dim[3][1];                                // e4 (b,c,l,L)
  ACETN_REQUIRE(s->E_dim[2][1] == a0, "site rdm: c4/e3 mismatch");
  const int64_t e3 = s->E_dim[2][0];                                // e3 (e,a,d,D)
  // t1a[a][c][l][L] = c4 . e4
  double* t1a = c.take((size_t)a0 * c4 * D2);
  mm(c, (int)a0, (int)(c4 * D2), (int)b0, false, s->C[3], b0, false, s->E[3], c4 * D2, t1a, c4 * D2);
  // t1b[c][l][L][e][d][D] = sum_a t1a[a][(c,l,L)] e3[e][a][(d,D)]
  const int64_t N1 = c4 * D2, Ne = e3 * D2;
  double* t1b = c.take((size_t)N1 * Ne);
  gemm(c, gdesc((int)N1, (int)Ne, (int)a0, plain(1), plain(N1), plain(D2),
                grouped((int)D2, 1, a0 * D2), plain(Ne), plain(1)), t1a, s->E[2], t1b);
  // t1c[c][l][e][d][U][R][P] = sum_{L,D} t1b[c][l][L][e][d][D] abar[L][U][R][D][P]
  const int64_t Mc = c4 * D * e3 * D, Nc = D2 * d;
  double* t1c = c.take((size_t)Mc * Nc);
  gemm(c, gdesc((int)Mc, (int)Nc, (int)D2,
                grouped4((int)D, D, (int)e3, D2, (int)D, D * Ne, D2 * Ne),  // m = (c,l,e,d)
                grouped((int)D, 1, Ne),                                   // k = (L,D)
                grouped((int)D, d, D * D2 * d),                           // A: D -> d, L -> D^3 d
                grouped3((int)d, 1, (int)D, D * d, D2 * d),               // n = (U,R,P)
                plain(Nc), plain(1)), t1b, s->A, t1c);
  // t2[a][c][u][U] = c1 . e1
  const int64_t a1 = s->C_dim[0][0], b1 = s->C_dim[0][1];
  ACETN_REQUIRE(s->E_dim[0][0] == b1, "site rdm: c1/e1 mismatch");
  const int64_t c1 = s->E_dim[0][1];
  double* t2 = c.take((size_t)a1 * c1 * D2);
  mm(c, (int)a1, (int)(c1 * D2), (int)b1, false, s->C[0], b1, false, s->E[0], c1 * D2, t2, c1 * D2);
  // t3a[b][c][r][R] = sum_a c3[a][b] e2[c][a][r][R]
  const int64_t a3 = s->C_dim[2][0], b3 = s->C_dim[2][1];
  ACETN_REQUIRE(s->E_dim[1][1] == a3, "site rdm: c3/e2 mismatch");
  const int64_t c2 = s->E_dim[1][0];
  double* t3a = c.take((size_t)b3 * c2 * D2);
  gemm(c, gdesc((int)b3, (int)(c2 * D2), (int)a3, plain(1), plain(b3), plain(D2),
                grouped((int)D2, 1, a3 * D2), plain(c2 * D2), plain(1)), s->C[2], s->E[1], t3a);
  // t3b[e][b][r][R] = sum_c c2[e][c] t3a[b][c][r][R]
  const int64_t e2 = s->C_dim[1][0];
  ACETN_REQUIRE(s->C_dim[1][1] == c2, "site rdm: c2/e2 mismatch");
  double* t3b = c.take((size_t)e2 * b3 * D2);
  gemm(c, gdesc((int)e2, (int)(b3 * D2), (int)c2, plain(c2), plain(1), plain(D2),
                grouped((int)D2, 1, c2 * D2), plain(b3 * D2), plain(1)), s->C[1], t3a, t3b);
  // t3c[b][r][R][a][u][U] = sum_e t3b[e][(b,r,R)] t2[a][e][(u,U)]
  ACETN_REQUIRE(c1 == e2, "site rdm: e1/c2 mismatch");
  const int64_t Mb = b3 * D2, Na = a1 * D2;
  double* t3c = c.take((size_t)Mb * Na);
  gemm(c, gdesc((int)Mb, (int)Na, (int)e2, plain(1), plain(Mb), plain(D2),
                grouped((int)D2, 1, c1 * D2), plain(Na), plain(1)), t3b, t2, t3c);
  // t4[r][u][l][d][P] = sum_{e,R,c,U} t3c[e][r][R][c][u][U] t1c[c][l][e][d][U][R][P]
  //   (t3c's leading index b plays e, its a plays c)
  ACETN_REQUIRE(b3 == e3 && a1 == c4, "site rdm: ring dimension mismatch");
  const int64_t ne = b3, nc = a1, Kt = ne * D * nc * D;
  double* t4 = c.take((size_t)D2 * D2 * d);
  gemm(c, gdesc((int)D2, (int)(D2 * d), (int)Kt,
                grouped((int)D, D, D * nc * D2),                               // m = (r,u)
                grouped4((int)D, 1, (int)nc, D2, (int)D, nc * D2, D2 * nc * D2),  // k = (e,R,c,U)
                grouped4((int)D, D * d, (int)nc, D * e3 * D * D2 * d, (int)D, d, D * D2 * d),
                grouped3((int)d, 1, (int)D, D2 * d, e3 * D * D2 * d),           // n = (l,d,P)
                plain(D2 * d), plain(1)), t3c, t1c, t4);
  // rho[P][p] = sum_{r,u,l,d} t4[(r,u,l,d)][P] a[l][u][r][d][p]
  gemm(c, gdesc((int)d, (int)d, (int)(D2 * D2), plain(1), plain(d),
                grouped4((int)D, d, (int)D, D * D2 * d, (int)D, D2 * d, D * d),
                plain(1), plain(d), plain(1)), t4, s->A, rho);
}

// build_bond_rdm + build_bond_rdm_core, rdm.py:69-120 (the blocked form of
// :123-154 computes the same tensor).  rho (d,d,d,d) = "PQpq".
}  // namespace

// The two halves of the bond network (rdm.py:102-120; the same network
// builds the full-update norm tensor, full_update.py:209-227):
//   r1[f][c][L][P][l][p]  (site s1 with its C12, E12, E11, C13, E13)
//   r2[f][c][R][Q][r][q]  (site s2 with its C21, E21, E24, C24, E23)
// A1/A2 are addressed through their (l, u, r, d) leg strides ps1/ps2 in the
// bond frame and a unit physical stride, so a reduced tensor a1q (r,d,u,y)
// with y = (l, p) enters exactly like a site tensor.
void bond_halves(Ctx& c, const acetn_site* s1, const acetn_site* s2, int k, const double* A1,
                 const int64_t ps1[4], const double* A2, const int64_t ps2[4], BondHalves& out) {
  k &= 3;
  const int64_t D = s1->D, d = s1->d, D2 = D * D, d2 = d * d;
  ACETN_REQUIRE(s2->D == D && s2->d == d, "bond network: site dimensions differ");
  const int k1 = (k + 1) & 3, k2 = (k + 2) & 3, k3 = (k + 3) & 3;
  // ---------------- right half (site s1)
  const int64_t c12a = s1->C_dim[k1][0], c12b = s1->C_dim[k1][1];
  ACETN_REQUIRE(s1->E_dim[k1][0] == c12b, "bond rdm: c12/e12 mismatch");
  const int64_t nc = s1->E_dim[k1][1];                 // c range
  ACETN_REQUIRE(s1->E_dim[k][1] == c12a, "bond rdm: c12/e11 mismatch");
  const int64_t ne = s1->E_dim[k][0];                  // e range
  const int64_t Mr = nc * D2, Nr = ne * D2;
  double* rt1 = c.take((size_t)c12a * Mr);
  mm(c, (int)c12a, (int)Mr, (int)c12b, false, s1->C[k1], c12b, false, s1->E[k1], Mr, rt1, Mr);
  double* rt2 = c.take((size_t)Mr * Nr);
  gemm(c, gdesc((int)Mr, (int)Nr, (int)c12a, plain(1), plain(Mr), plain(D2),
                grouped((int)D2, 1, c12a * D2), plain(Nr), plain(1)), rt1, s1->E[k], rt2);
  const int64_t Lr = ne * D2 * d2;                     // row length of t: (c,L,P,l,p)
  double* t = c.take((size_t)nc * D2 * Lr);
  if (c.launch) {
    DLDesc dl{};
    dl.D = (int)D; dl.d = (int)d; dl.n1 = (int)nc; dl.n2 = (int)ne;
    dl.x_o1 = D2 * Nr; dl.x_k2 = D * Nr; dl.x_b2 = Nr;   // (c, r, R)
    dl.x_o2 = D2; dl.x_k1 = D; dl.x_b1 = 1;              // (e, u, U)
    dl.a_leg[0] = ps1[1]; dl.a_leg[1] = ps1[2]; dl.a_leg[2] = ps1[0]; dl.a_leg[3] = ps1[3];
    dl.a_p = 1;
    dl.open_p = 1;
    // t[a=c][d][D][c=e][L][P][l][p]
    dl.y_o1 = D2 * Lr; dl.y_k4 = D * Lr; dl.y_b4 = Lr;
    dl.y_o2 = D2 * d2; dl.y_b3 = D * d2; dl.y_pb = D * d; dl.y_k3 = d; dl.y_p = 1;
    double_layer_f64(dl, rt2, A1, t, c.st);
  }
  const int64_t c13a = s1->C_dim[k2][0], c13b = s1->C_dim[k2][1];
  ACETN_REQUIRE(s1->E_dim[k2][0] == c13b, "bond rdm: c13/e13 mismatch");
  ACETN_REQUIRE(c13a == nc, "bond rdm: c13/e12 mismatch");
  const int64_t F = s1->E_dim[k2][1];
  double* r1a = c.take((size_t)c13a * F * D2);
  mm(c, (int)c13a, (int)(F * D2), (int)c13b, false, s1->C[k2], c13b, false, s1->E[k2], F * D2, r1a,
     F * D2);
  double* r1 = c.take((size_t)F * Lr);
  gemm(c, gdesc((int)F, (int)Lr, (int)(nc * D2), plain(D2), grouped((int)D2, 1, F * D2), plain(Lr),
                plain(1), plain(Lr), plain(1)), r1a, t, r1);
  // ---------------- left half (site s2)
  const int64_t c21a = s2->C_dim[k][0], c21b = s2->C_dim[k][1];
  ACETN_REQUIRE(s2->E_dim[k][0] == c21b, "bond rdm: c21/e21 mismatch");
  const int64_t mc = s2->E_dim[k][1];
  ACETN_REQUIRE(s2->E_dim[k3][1] == c21a, "bond rdm: c21/e24 mismatch");
  const int64_t me = s2->E_dim[k3][0];
  const int64_t Ml = mc * D2, Nl = me * D2;
  double* lt1 = c.take((size_t)c21a * Ml);
  mm(c, (int)c21a, (int)Ml, (int)c21b, false, s2->C[k], c21b, false, s2->E[k], Ml, lt1, Ml);
  double* lt2 = c.take((size_t)Ml * Nl);
  gemm(c, gdesc((int)Ml, (int)Nl, (int)c21a, plain(1), plain(Ml), plain(D2),
                grouped((int)D2, 1, c21a * D2), plain(Nl), plain(1)), lt1, s2->E[k3], lt2);
  const int64_t Ll = mc * D2 * d2;                     // row length of t2: (c,R,Q,r,q)
  double* t2 = c.take((size_t)me * D2 * Ll);
  if (c.launch) {
    DLDesc dl{};
    dl.D = (int)D; dl.d = (int)d; dl.n1 = (int)mc; dl.n2 = (int)me;
    dl.x_o1 = D2 * Nl; dl.x_k1 = D * Nl; dl.x_b1 = Nl;   // (c, u, U)
    dl.x_o2 = D2; dl.x_k2 = D; dl.x_b2 = 1;              // (e, l, L)
    dl.a_leg[0] = ps2[1]; dl.a_leg[1] = ps2[0]; dl.a_leg[2] = ps2[2]; dl.a_leg[3] = ps2[3];
    dl.a_p = 1;
    dl.open_p = 1;
    // t2[e][d][D][c][R][Q][r][q]
    dl.y_o2 = D2 * Ll; dl.y_k4 = D * Ll; dl.y_b4 = Ll;
    dl.y_o1 = D2 * d2; dl.y_b3 = D * d2; dl.y_pb = D * d; dl.y_k3 = d; dl.y_p = 1;
    double_layer_f64(dl, lt2, A2, t2, c.st);
  }
  const int64_t c24a = s2->C_dim[k3][0], c24b = s2->C_dim[k3][1];
  ACETN_REQUIRE(c24b == me, "bond rdm: c24/e24 mismatch");
  ACETN_REQUIRE(s2->E_dim[k2][1] == c24a, "bond rdm: c24/e23 mismatch");
  const int64_t F2 = s2->E_dim[k2][0];
  double* r2a = c.take((size_t)me * F2 * D2);
  gemm(c, gdesc((int)me, (int)(F2 * D2), (int)c24a, plain(1), plain(c24b), plain(D2),
                grouped((int)D2, 1, c24a * D2), plain(F2 * D2), plain(1)), s2->C[k3], s2->E[k2], r2a);
  double* r2 = c.take((size_t)F2 * Ll);
  gemm(c, gdesc((int)F2, (int)Ll, (int)(me * D2), plain(D2), grouped((int)D2, 1, F2 * D2),
                plain(Ll), plain(1), plain(Ll), plain(1)), r2a, t2, r2);
  ACETN_REQUIRE(F == F2 && ne == mc, "bond network: halves do not close");
  out.r1 = r1;
  out.r2 = r2;
  out.F = F;
  out.nc = ne;
}

namespace {

void bond_rdm(Ctx& c, const acetn_site* s1, const acetn_site* s2, int k, double* rho) {
  int64_t ps1[4], ps2[4];
  permuted_strides(s1, k, ps1);
  permuted_strides(s2, k, ps2);
  BondHalves h{};
  bond_halves(c, s1, s2, k, s1->A, ps1, s2->A, ps2, h);
  const int64_t D = s1->D, d = s1->d, D2 = D * D, d2 = d * d;
  const double* r1 = h.r1;
  const double* r2 = h.r2;
  // ---------------- rho[P][Q][p][q] = sum_{f,c,R,r} r1[f][c][R][P][r][p] r2[f][c][R][Q][r][q]
  const int64_t K = h.F * h.nc * D2;
  gemm(c, gdesc((int)d2, (int)d2, (int)K, grouped((int)d, 1, D * d),
                grouped3((int)D, d, (int)D, D * d2, D2 * d2), grouped3((int)D, d, (int)D, D * d2, D2 * d2),
                grouped((int)d, 1, D * d), grouped((int)d, d, d2 * d), grouped((int)d, 1, d2)),
       r1, r2, rho);
}

void check_projector_args(const acetn_site* const* sites, int system, int svd_type) {
  ACETN_REQUIRE(sites != nullptr, "projectors: null site array");
  ACETN_REQUIRE(system == kSysHalf || system == kSysFull, "projectors: unknown system");
  ACETN_REQUIRE(svd_type == kSvdRsvd || svd_type == kSvdFullRank, "projectors: unknown svd type");
  check_site(sites[0]);
  check_site(sites[3]);
  if (system == kSysFull) {
    check_site(sites[1]);
    check_site(sites[2]);
  }
}

}  // namespace
}  // namespace acetn

using namespace acetn;

extern "C" {

const char* acetn_hip_last_error(void) { return g_last_error.c_str(); }

int acetn_hip_autotune(int mode) {
  return guarded([&] { autotune_set_mode(mode); });
}

int acetn_hip_autotune_entries(int64_t* host_n) {
  return guarded([&] {
    ACETN_REQUIRE(host_n, "autotune_entries: null output");
    *host_n = autotune_entries();
  });
}

int acetn_hip_autotune_clear(void) {
  return guarded([&] { autotune_clear(); });
}

int acetn_hip_autotune_save(const char* path) {
  return guarded([&] { autotune_save(path); });
}

int acetn_hip_autotune_load(const char* path, int64_t* host_n) {
  return guarded([&] {
    const int64_t n = autotune_load(path);
    if (host_n) *host_n = n;
  });
}

int acetn_hip_kernel_timing(int enable) {
  return guarded([&] { kernel_timing_enable(enable != 0); });
}

int acetn_hip_kernel_timing_bytes(int kernel, double* host_bytes) {
  return guarded([&] { *host_bytes = kernel_timing_bytes(kernel); });
}

int acetn_hip_kernel_timing_read(int kernel, double* host_ms, double* host_flops,
                                 int64_t* host_launches) {
  return guarded([&] { kernel_timing_read(kernel, host_ms, host_flops, host_launches); });
}

int acetn_hip_device_info(char* buf, size_t len) {
  return guarded([&] {
    int dev = 0;
    ACETN_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    ACETN_HIP_CHECK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s|%s|cu=%d|lds=%zu|clock_khz=%d", p.name, p.gcnArchName,
             p.multiProcessorCount, (size_t)p.sharedMemPerBlock, p.clockRate);
  });
}

size_t acetn_hip_quarter_tensor_workspace(const acetn_site* s, int k) {
  size_t r = 0;
  guarded([&] { check_site(s); r = measure([&](Ctx& c) { quarter_tensor(c, s, k, nullptr); }); });
  return r;
}

int acetn_hip_quarter_tensor(const acetn_site* s, int k, int precision, double* Q, void* ws,
                             size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_site(s);
    check_precision(precision);
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      quarter_tensor(c, s, k, Q);
    });
  });
}

size_t acetn_hip_half_projectors_workspace(const acetn_site* s1, const acetn_site* s4, int k,
                                           int64_t chi, int64_t oversampling) {
  size_t r = 0;
  guarded([&] {
    check_site(s1);
    check_site(s4);
    HalfOut o{nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    r = measure([&](Ctx& c) {
      half_projectors(c, s1, s4, k, chi, oversampling, 2, 0.0, 0, 0, o);
    });
  });
  return r;
}

int acetn_hip_half_projectors(const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                              int64_t oversampling, int niter, double svd_cutoff, uint64_t seed,
                              uint64_t sketch_stream, int precision, double* P1, double* P2,
                              int64_t* host_cD_new, double* spectrum, void* ws, size_t ws_bytes,
                              void* stream) {
  return guarded([&] {
    check_site(s1);
    check_site(s4);
    check_precision(precision);
    ACETN_REQUIRE(P1 && P2 && host_cD_new, "half_projectors: null output");
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "half_projectors: niter out of range");
    HalfOut o{P1, P2, host_cD_new, nullptr, 0, spectrum};
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      half_projectors(c, s1, s4, k, chi, oversampling, niter, svd_cutoff, seed, sketch_stream, o);
    });
  });
}

static void check_rsvd_args(int nops, const int64_t* dims, int64_t q) {
  ACETN_REQUIRE(nops >= 1 && nops <= 8 && dims != nullptr, "rsvd_product: 1 to 8 factors");
  for (int i = 0; i <= nops; ++i)
    ACETN_REQUIRE(dims[i] >= 1 && dims[i] < (int64_t(1) << 31), "rsvd_product: bad dimension");
  ACETN_REQUIRE(q >= 1 && q <= std::min(dims[0], dims[nops]) && q <= 510,
                "rsvd_product: q must be in [1, min(rows, cols)] and at most 510");
}

size_t acetn_hip_rsvd_product_workspace(int nops, const int64_t* dims, int64_t q, int niter,
                                        int mid_qr) {
  size_t r = 0;
  guarded([&] {
    check_rsvd_args(nops, dims, q);
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "rsvd_product: niter out of range");
    const double* none[8] = {};
    r = measure([&](Ctx& c) {
      rsvd_product(c, nops, none, dims, q, niter, mid_qr != 0, 0, 0, nullptr, nullptr, nullptr);
    });
  });
  return r;
}

int acetn_hip_rsvd_product(int nops, const double* const* ops, const int64_t* dims, int64_t q,
                           int niter, int mid_qr, uint64_t seed, uint64_t sketch_stream, double* U,
                           double* S, double* V, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_rsvd_args(nops, dims, q);
    ACETN_REQUIRE(ops != nullptr && U && S && V, "rsvd_product: null pointer");
    for (int i = 0; i < nops; ++i) ACETN_REQUIRE(ops[i] != nullptr, "rsvd_product: null factor");
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "rsvd_product: niter out of range");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      rsvd_product(c, nops, ops, dims, q, niter, mid_qr != 0, seed, sketch_stream, U, S, V);
    });
  });
}

int acetn_hip_half_projectors_begin(const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                                    int64_t oversampling, int niter, double svd_cutoff,
                                    uint64_t seed, uint64_t sketch_stream, int precision,
                                    int32_t* pinned_cD_new, double* spectrum, void* ws,
                                    size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_site(s1);
    check_site(s4);
    check_precision(precision);
    ACETN_REQUIRE(pinned_cD_new, "half_projectors_begin: null cD target");
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "half_projectors: niter out of range");
    HalfOut o{nullptr, nullptr, nullptr, pinned_cD_new, 0, spectrum};
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      half_projectors(c, s1, s4, k, chi, oversampling, niter, svd_cutoff, seed, sketch_stream, o, 1);
    });
  });
}

int acetn_hip_half_projectors_end(const acetn_site* s1, const acetn_site* s4, int k, int64_t chi,
                                  int64_t oversampling, int niter, int precision, int64_t cD_new,
                                  double* P1, double* P2, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_site(s1);
    check_site(s4);
    check_precision(precision);
    ACETN_REQUIRE(P1 && P2, "half_projectors_end: null output");
    HalfOut o{P1, P2, nullptr, nullptr, (int)cD_new, nullptr};
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      half_projectors(c, s1, s4, k, chi, oversampling, niter, 0.0, 0, 0, o, 2);
    });
  });
}

size_t acetn_hip_projectors_workspace(const acetn_site* const sites[4], int k, int64_t chi,
                                      int64_t oversampling, int system, int svd_type) {
  size_t r = 0;
  guarded([&] {
    check_projector_args(sites, system, svd_type);
    HalfOut o{nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    r = measure([&](Ctx& c) {
      if (system == kSysHalf && svd_type == kSvdRsvd)
        half_projectors(c, sites[0], sites[3], k, chi, oversampling, 2, 0.0, 0, 0, o);
      else
        general_projectors(c, sites, k, system, svd_type, chi, oversampling, 2, 0.0, 0, 0, o);
    });
  });
  return r;
}

int acetn_hip_projectors_begin(const acetn_site* const sites[4], int k, int64_t chi,
                               int64_t oversampling, int niter, double svd_cutoff, uint64_t seed,
                               uint64_t sketch_stream, int system, int svd_type, int precision,
                               int32_t* pinned_cD_new, double* spectrum, void* ws,
                               size_t ws_bytes, void* stream) {
  return acetn_hip_projectors_begin_mark(sites, k, chi, oversampling, niter, svd_cutoff, seed,
                                         sketch_stream, system, svd_type, precision,
                                         pinned_cD_new, spectrum, ws, ws_bytes, stream, nullptr);
}

int acetn_hip_projectors_begin_mark(const acetn_site* const sites[4], int k, int64_t chi,
                                    int64_t oversampling, int niter, double svd_cutoff,
                                    uint64_t seed, uint64_t sketch_stream, int system,
                                    int svd_type, int precision, int32_t* pinned_cD_new,
                                    double* spectrum, void* ws, size_t ws_bytes, void* stream,
                                    void* mark_event) {
  return guarded([&] {
    check_projector_args(sites, system, svd_type);
    check_precision(precision);
    ACETN_REQUIRE(pinned_cD_new, "projectors_begin: null cD target");
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "projectors: niter out of range");
    HalfOut o{nullptr, nullptr, nullptr, pinned_cD_new, 0, spectrum};
    o.mark = reinterpret_cast<hipEvent_t>(mark_event);
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      if (system == kSysHalf && svd_type == kSvdRsvd)
        half_projectors(c, sites[0], sites[3], k, chi, oversampling, niter, svd_cutoff, seed,
                        sketch_stream, o, 1);
      else
        general_projectors(c, sites, k, system, svd_type, chi, oversampling, niter, svd_cutoff,
                           seed, sketch_stream, o, 1);
    });
  });
}

int acetn_hip_projectors_end(const acetn_site* const sites[4], int k, int64_t chi,
                             int64_t oversampling, int niter, int system, int svd_type,
                             int precision, int64_t cD_new, double* P1, double* P2, void* ws,
                             size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_projector_args(sites, system, svd_type);
    check_precision(precision);
    ACETN_REQUIRE(P1 && P2, "projectors_end: null output");
    HalfOut o{P1, P2, nullptr, nullptr, (int)cD_new, nullptr};
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      if (system == kSysHalf && svd_type == kSvdRsvd)
        half_projectors(c, sites[0], sites[3], k, chi, oversampling, niter, 0.0, 0, 0, o, 2);
      else
        general_projectors(c, sites, k, system, svd_type, chi, oversampling, niter, 0.0, 0, 0, o,
                           2);
    });
  });
}

int acetn_hip_projectors_begin_graph(const acetn_site* const sites[4], int k, int64_t chi,
                                     int64_t oversampling, int niter, double svd_cutoff,
                                     uint64_t seed, uint64_t sketch_stream,
                                     const uint64_t* dev_stream_base, int system, int svd_type,
                                     int precision, int32_t* dev_cD_new, double* spectrum,
                                     void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_projector_args(sites, system, svd_type);
    check_precision(precision);
    ACETN_REQUIRE(dev_stream_base && dev_cD_new, "projectors_begin_graph: null device pointer");
    ACETN_REQUIRE(niter >= 0 && niter <= 16, "projectors: niter out of range");
    HalfOut o{nullptr, nullptr, nullptr, nullptr, 0, spectrum};
    o.stream_base = dev_stream_base;
    o.dev_cD = dev_cD_new;
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      if (system == kSysHalf && svd_type == kSvdRsvd)
        half_projectors(c, sites[0], sites[3], k, chi, oversampling, niter, svd_cutoff, seed,
                        sketch_stream, o, 1);
      else
        general_projectors(c, sites, k, system, svd_type, chi, oversampling, niter, svd_cutoff,
                           seed, sketch_stream, o, 1);
    });
  });
}

int acetn_hip_projectors_end_padded(const acetn_site* const sites[4], int k, int64_t chi,
                                    int64_t oversampling, int niter, int system, int svd_type,
                                    int precision, int64_t cD_cols, double* P1, double* P2,
                                    void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_projector_args(sites, system, svd_type);
    check_precision(precision);
    ACETN_REQUIRE(P1 && P2, "projectors_end_padded: null output");
    HalfOut o{P1, P2, nullptr, nullptr, (int)cD_cols, nullptr};
    o.padded = true;
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      if (system == kSysHalf && svd_type == kSvdRsvd)
        half_projectors(c, sites[0], sites[3], k, chi, oversampling, niter, 0.0, 0, 0, o, 2);
      else
        general_projectors(c, sites, k, system, svd_type, chi, oversampling, niter, 0.0, 0, 0, o,
                           2);
    });
  });
}

size_t acetn_hip_renormalize_workspace(const acetn_site* s1, int k, int64_t cD_i, int64_t cD_j) {
  size_t r = 0;
  guarded([&] {
    check_site(s1);
    r = measure([&](Ctx& c) {
      renormalize(c, s1, k, nullptr, nullptr, cD_i, nullptr, nullptr, cD_j, nullptr, nullptr,
                  nullptr);
    });
  });
  return r;
}

int acetn_hip_renormalize_boundary(const acetn_site* s1, int k, const double* P1_i,
                                   const double* P2_i, int64_t cD_i, const double* P1_j,
                                   const double* P2_j, int64_t cD_j, int precision, double* C_a,
                                   double* C_b, double* E_o, void* ws, size_t ws_bytes,
                                   void* stream) {
  return guarded([&] {
    check_site(s1);
    check_precision(precision);
    ACETN_REQUIRE(cD_i >= 1 && cD_j >= 1, "renormalize: empty projector");
    ACETN_REQUIRE(P1_i && P2_i && P1_j && P2_j && C_a && C_b && E_o, "renormalize: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      renormalize(c, s1, k, P1_i, P2_i, cD_i, P1_j, P2_j, cD_j, C_a, C_b, E_o);
    });
  });
}

size_t acetn_hip_renormalize_part_workspace(const acetn_site* s1, int k, int first, int64_t cD_i,
                                            int64_t cD_j) {
  size_t r = 0;
  guarded([&] {
    check_site(s1);
    ACETN_REQUIRE(first == 0 || first == 1, "renormalize_part: first must be 0 or 1");
    r = measure([&](Ctx& c) {
      renormalize(c, s1, k, nullptr, nullptr, cD_i, nullptr, nullptr, cD_j, nullptr, nullptr,
                  nullptr, 1, first == 0 ? cD_i : cD_j, first);
    });
  });
  return r;
}

int acetn_hip_renormalize_boundary_part(const acetn_site* s1, int k, int part, int first,
                                        const double* P1_i, const double* P2_i, int64_t cD_i,
                                        const double* P1_j, const double* P2_j, int64_t cD_j,
                                        int64_t cD_layout, int precision, double* C_a,
                                        double* C_b, double* E_o, void* ws, size_t ws_bytes,
                                        void* stream) {
  return guarded([&] {
    check_site(s1);
    check_precision(precision);
    ACETN_REQUIRE(part == 1 || part == 2, "renormalize_part: part must be 1 or 2");
    ACETN_REQUIRE(first == 0 || first == 1, "renormalize_part: first must be 0 or 1");
    ACETN_REQUIRE(cD_layout >= 1, "renormalize_part: empty projector");
    // the pair of part 1 and the one whose count is only bounded there
    int64_t& later = first == 0 ? cD_i : cD_j;
    const int64_t known = first == 0 ? cD_j : cD_i;
    ACETN_REQUIRE(known >= 1, "renormalize_part: empty projector");
    const bool ok1 = first == 0 ? (P1_j && P2_j && C_b) : (P1_i && P2_i && C_a);
    const bool ok2 = P1_i && P2_i && P1_j && P2_j && E_o && (first == 0 ? C_a : C_b);
    if (part == 1) {
      ACETN_REQUIRE(ok1, "renormalize_part: null pointer");
      later = cD_layout;
    } else {
      ACETN_REQUIRE(later >= 1 && later <= cD_layout, "renormalize_part: count above its layout bound");
      ACETN_REQUIRE(ok2, "renormalize_part: null pointer");
    }
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      set_precision(c, precision);
      renormalize(c, s1, k, P1_i, P2_i, cD_i, P1_j, P2_j, cD_j, C_a, C_b, E_o, part, cD_layout,
                  first);
    });
  });
}

size_t acetn_hip_site_rdm_workspace(const acetn_site* s) {
  size_t r = 0;
  guarded([&] { check_site(s); r = measure([&](Ctx& c) { site_rdm(c, s, nullptr); }); });
  return r;
}

int acetn_hip_site_rdm(const acetn_site* s, double* rho, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_site(s);
    ACETN_REQUIRE(rho, "site_rdm: null output");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { site_rdm(c, s, rho); });
  });
}

size_t acetn_hip_bond_rdm_workspace(const acetn_site* s1, const acetn_site* s2, int k) {
  size_t r = 0;
  guarded([&] {
    check_site(s1);
    check_site(s2);
    r = measure([&](Ctx& c) { bond_rdm(c, s1, s2, k, nullptr); });
  });
  return r;
}

int acetn_hip_bond_rdm(const acetn_site* s1, const acetn_site* s2, int k, double* rho, void* ws,
                       size_t ws_bytes, void* stream) {
  return guarded([&] {
    check_site(s1);
    check_site(s2);
    ACETN_REQUIRE(rho, "bond_rdm: null output");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { bond_rdm(c, s1, s2, k, rho); });
  });
}

int acetn_hip_gaussian_sketch(double* out, int64_t n, int64_t q, uint64_t seed,
                              uint64_t sketch_stream, void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(out && n >= 0 && q >= 0, "sketch: bad arguments");
    gaussian_sketch(out, n, q, seed, sketch_stream, reinterpret_cast<hipStream_t>(stream));
  });
}

size_t acetn_hip_gemm_workspace(int64_t M, int64_t N, int64_t K) {
  size_t r = 0;
  guarded([&] {
    r = measure([&](Ctx& c) {
      mm(c, (int)M, (int)N, (int)K, false, nullptr, K, false, nullptr, N, nullptr, N);
    });
  });
  return r;
}

int acetn_hip_gemm(int64_t M, int64_t N, int64_t K, int ta, const double* A, int64_t lda, int tb,
                   const double* B, int64_t ldb, double alpha, double beta, double* C,
                   int64_t ldc, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) && N < (1LL << 31) &&
                      K < (1LL << 31), "gemm: bad shape");
    ACETN_REQUIRE(A && B && C, "gemm: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      mm(c, (int)M, (int)N, (int)K, ta != 0, A, lda, tb != 0, B, ldb, C, ldc, alpha, beta);
    });
  });
}

int acetn_hip_gemm_split(int64_t M, int64_t N, int64_t K, int ta, const double* A, int64_t lda,
                         int tb, const double* B, int64_t ldb, double alpha, double beta, double* C,
                         int64_t ldc, int pieces, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(M >= 0 && N >= 0 && K >= 0 && M < (1LL << 31) && N < (1LL << 31) &&
                      K < (1LL << 31), "gemm_split: bad shape");
    ACETN_REQUIRE(A && B && C, "gemm_split: null pointer");
    ACETN_REQUIRE(pieces == 2 || pieces == 3, "gemm_split: pieces must be 2 or 3");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      c.mixed = true;
      c.split = pieces;
      mm(c, (int)M, (int)N, (int)K, ta != 0, A, lda, tb != 0, B, ldb, C, ldc, alpha, beta);
    });
  });
}

size_t acetn_hip_svd_workspace(int64_t rows, int64_t cols) {
  if (rows < 1 || cols < 1 || rows > 2048 || cols > 2048) return 0;
  return sizeof(double) * jacobi_svd_scratch_elems((int)rows, (int)cols) + 512;
}

int acetn_hip_svd(const double* M, int64_t rows, int64_t cols, int64_t ldc, double* s, double* U,
                  int32_t* host_sweeps, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(M && s && U && ws, "svd: null pointer");
    ACETN_REQUIRE(rows >= 1 && cols >= 1 && rows <= 2048 && cols <= 2048 && ldc >= rows,
                  "svd: bad shape");
    ACETN_REQUIRE(ws_bytes + 512 >= acetn_hip_svd_workspace(rows, cols), "svd: workspace too small");
    double* scr = reinterpret_cast<double*>(((uintptr_t)ws + 255) & ~uintptr_t(255));
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    jacobi_svd(M, (int)rows, (int)cols, ldc, s, U, scr, st);
    int info[2] = {0, 0};
    ACETN_HIP_CHECK(hipMemcpyAsync(info, scr, sizeof(info), hipMemcpyDeviceToHost, st));
    ACETN_HIP_CHECK(hipStreamSynchronize(st));
    if (host_sweeps) *host_sweeps = info[1] ? info[0] : -info[0];
  });
}

size_t acetn_hip_eigh_workspace(int64_t q) {
  return sizeof(double) * (jacobi_rot_elems((int)q) + 64) + 512;
}

int acetn_hip_eigh(double* G, int64_t q, double* w, double* V, void* ws, size_t ws_bytes,
                   void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(q >= 1 && q <= 510, "eigh: q out of range");
    ACETN_REQUIRE(ws_bytes >= acetn_hip_eigh_workspace(q) - 512, "eigh: workspace too small");
    jacobi_eigh(G, w, V, reinterpret_cast<double*>(ws), (int)q,
                reinterpret_cast<hipStream_t>(stream));
  });
}

size_t acetn_hip_orthonormalize_workspace(int64_t m, int64_t q) {
  size_t r = 0;
  guarded([&] { r = measure([&](Ctx& c) { orthonormalize(c, nullptr, m, q); }); });
  return r;
}

int acetn_hip_orthonormalize(double* Y, int64_t m, int64_t q, void* ws, size_t ws_bytes,
                             void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(Y && m >= 1 && q >= 1 && q <= 510, "orthonormalize: bad arguments");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { orthonormalize(c, Y, m, q); });
  });
}

}  // extern "C"
