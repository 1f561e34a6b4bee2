// Full-update bond step: host-side orchestration of the HIP kernels for
// tensor_update.py (decompose / recompose), full_update.py (norm tensor,
// gate, positive_approx, gauge_fix, finalize) and als_solver.py (plain
// start, ALS), plus their extern "C" boundary.  Index names follow the
// reference's einsum strings.
#include "engine.h"
#include "fu_internal.h"

#include <mutex>
#include <vector>

namespace acetn {
namespace {

struct FuDims {
  int64_t D, d, nD;
};
FuDims fu_dims(int64_t D, int64_t d) {
  ACETN_REQUIRE(D >= 1 && D <= 8 && d >= 1 && d <= 4, "full update: unsupported D/d");
  ACETN_REQUIRE(D * D >= d, "full update: reduced bond dimension min(D^3, dD) must equal dD");
  return FuDims{D, d, D * d};
}

// Dimensions of the bond solver (gate, positive_approx, gauge_fix, ALS):
// reduced tensors (nD, D, d) with any reduced dimension nD, as ALSSolver
// takes ar_shape = (nD, bD, pD) (als_solver.py:10-32); a full-update step has
// nD = min(D^3, D d).
FuDims solver_dims(int64_t nD, int64_t D, int64_t d) {
  ACETN_REQUIRE(nD >= 1 && nD <= kQrMaxN && D >= 1 && D <= 64 && d >= 1 && d <= 8,
                "bond solver: unsupported reduced tensor shape (nD, D, d)");
  return FuDims{D, d, nD};
}

// site-frame strides of A (l,u,r,d,p) contiguous
void site_strides(int64_t D, int64_t d, int64_t st[4]) {
  st[0] = D * D * D * d; st[1] = D * D * d; st[2] = D * d; st[3] = d;
}

// ---------------------------------------------------------------------------
// decompose_site_tensors (tensor_update.py:52-68): the bond-frame tensors
// a1 = "lurdp->rdulp", a2 = "lurdp->dlurp" reshaped (D^3, dD) and QR'd.
// ---------------------------------------------------------------------------
void decompose(Ctx& c, const double* A1, const double* A2, int k, const FuDims& f, double* a1q,
               double* a1r, double* a2q, double* a2r) {
  k &= 3;
  const int64_t D = f.D, d = f.d, nD = f.nD, m = D * D * D;
  double* M = c.take((size_t)m * nD);
  double* W = c.take((size_t)m * nD);
  if (!c.launch) return;
  int64_t st[4];
  site_strides(D, d, st);
  PermDesc p{};
  p.n = 5; p.total = m * nD;
  p.od[0] = D; p.od[1] = D; p.od[2] = D; p.od[3] = D; p.od[4] = d;
  // a1: (r, d, u, l, p) of the bond frame; bond axis j is site axis (j + k) % 4
  p.is[0] = st[(2 + k) & 3]; p.is[1] = st[(3 + k) & 3]; p.is[2] = st[(1 + k) & 3];
  p.is[3] = st[k]; p.is[4] = 1;
  permute_copy(A1, M, p, 1.0, c.st);
  householder_qr(M, (int)m, (int)nD, W, a1q, a1r, c.st);
  // a2: (d, l, u, r, p)
  p.is[0] = st[(3 + k) & 3]; p.is[1] = st[k]; p.is[2] = st[(1 + k) & 3];
  p.is[3] = st[(2 + k) & 3]; p.is[4] = 1;
  permute_copy(A2, M, p, 1.0, c.st);
  householder_qr(M, (int)m, (int)nD, W, a2q, a2r, c.st);
}

// ---------------------------------------------------------------------------
// build_norm_tensor (full_update.py:163-227): the bond network of the reduced
// density matrix with a1q/a2q in place of the site tensors, the reduced leg
// y = (l, p) of a1q (x = (r, p) of a2q) left open; n12[y][x][Y][X].
// ---------------------------------------------------------------------------
void norm_tensor(Ctx& c, const acetn_site* s1, const acetn_site* s2, int k, const double* a1q,
                 const double* a2q, double* n12) {
  const FuDims f = fu_dims(s1->D, s1->d);
  const int64_t D = f.D, d = f.d, nD = f.nD;
  // a1q (r, d, u, y) with y = l*d + p: leg strides (l, u, r, d)
  const int64_t ps1[4] = {d, nD, D * D * nD, D * nD};
  // a2q (d, l, u, x) with x = r*d + p
  const int64_t ps2[4] = {D * nD, nD, d, D * D * nD};
  BondHalves h{};
  bond_halves(c, s1, s2, k, a1q, ps1, a2q, ps2, h);
  // n12[y][x][Y][X] = sum_{f,c} r1[f][c][Y=(L,P)][y=(l,p)] r2[f][c][X=(R,Q)][x=(r,q)]
  const int64_t nD2 = nD * nD;
  const IdxMap open = grouped3((int)d, D * d, (int)D, D * d * d, 1);
  gemm(c, gdesc((int)nD2, (int)nD2, (int)(h.F * h.nc), open, grouped((int)h.nc, nD2, h.nc * nD2),
                grouped((int)h.nc, nD2, h.nc * nD2), open, grouped((int)nD, nD, nD2 * nD),
                grouped((int)nD, 1, nD2)),
       h.r1, h.r2, n12);
}

// ---------------------------------------------------------------------------
// gate application (full_update.py:77-86):
//   a12g = "yup,xuq->yxpq" (a1r, a2r); a12g = "yxpq,pqrs->yxrs" (a12g, gate)
// ---------------------------------------------------------------------------
void apply_gate(Ctx& c, const double* a1r, const double* a2r, const double* gate,
                const FuDims& f, double* a12g) {
  const int64_t D = f.D, d = f.d, nD = f.nD;
  double* t = c.take((size_t)nD * nD * d * d);
  gemm(c, gdesc((int)(nD * d), (int)(nD * d), (int)D, grouped((int)d, 1, D * d), plain(d),
                plain(d), grouped((int)d, 1, D * d), grouped((int)d, d, nD * d * d),
                grouped((int)d, 1, d * d)),
       a1r, a2r, t);
  mm(c, (int)(nD * nD), (int)(d * d), (int)(d * d), false, t, d * d, false, gate, d * d, a12g,
     d * d);
}

// ---------------------------------------------------------------------------
// positive_approx (full_update.py:262-297): eigh of N = n12 as (nD^2, nD^2);
// while the smallest eigenvalue is below the cutoff, N += 2 max(cutoff,
// |lambda_min|) I and again; nz = Z sqrt(lambda).  The eigenvalue check reads
// one number back per decomposition (a host synchronisation, as the
// reference's loop condition is).
// ---------------------------------------------------------------------------
void positive_approx(Ctx& c, const double* n12, const FuDims& f, double cutoff, double* nz,
                     int* host_shifts) {
  const int q = (int)(f.nD * f.nD);
  ACETN_REQUIRE(q <= 510, "positive_approx: nD^2 must be at most 510");
  double* N = c.take((size_t)q * q);
  double* G = c.take((size_t)q * q);
  double* w = c.take((size_t)q);
  double* V = c.take((size_t)q * q);
  double* rot = c.take(jacobi_rot_elems(q));
  if (!c.launch) return;
  ACETN_HIP_CHECK(hipMemcpyAsync(N, n12, sizeof(double) * q * q, hipMemcpyDeviceToDevice, c.st));
  // torch.linalg.eigh reads the lower triangle; n12 is symmetric only up to
  // the environment's own asymmetry
  lower_to_full(N, q, c.st);
  int shifts = 0;
  for (;;) {
    ACETN_HIP_CHECK(hipMemcpyAsync(G, N, sizeof(double) * q * q, hipMemcpyDeviceToDevice, c.st));
    jacobi_eigh(G, w, V, rot, q, c.st);
    double wmin = 0.0;
    ACETN_HIP_CHECK(hipMemcpyAsync(&wmin, w + (q - 1), sizeof(double), hipMemcpyDeviceToHost, c.st));
    ACETN_HIP_CHECK(hipStreamSynchronize(c.st));
    ACETN_REQUIRE(std::isfinite(wmin), "positive_approx: non-finite eigenvalue");
    if (!(wmin < cutoff)) break;
    ACETN_REQUIRE(++shifts <= 64, "positive_approx: no positive shift found");
    add_diag(N, q, 2.0 * std::max(cutoff, std::fabs(wmin)), c.st);
  }
  scale_sqrt_columns(V, w, q, q, nz, c.st);
  if (host_shifts) *host_shifts = shifts;
}

// ---------------------------------------------------------------------------
// gauge_fix (full_update.py:300-343)
// ---------------------------------------------------------------------------
void gauge_fix(Ctx& c, const double* nz, const double* a12g_in, const FuDims& f, double atol,
               double* n12, double* a12g, double* nzxr_inv, double* nzyr_inv) {
  const int64_t d = f.d, nD = f.nD, nD2 = nD * nD, nD3 = nD2 * nD, d2 = d * d;
  double* M = c.take((size_t)nD3 * nD);
  double* W = c.take((size_t)nD3 * nD);
  double* nzxr = c.take((size_t)nD2);
  double* nzyr = c.take((size_t)nD2);
  double* U = c.take((size_t)nD2);
  double* sv = c.take((size_t)nD);
  double* V = c.take((size_t)nD2);
  double* T = c.take((size_t)nD3 * nD);
  double* nz2 = c.take((size_t)nD3 * nD);
  double* U1 = c.take((size_t)nD2 * d2);
  if (c.launch) {
    PermDesc p{};
    p.n = 3; p.total = nD3 * nD;
    p.od[0] = nD2; p.od[1] = nD; p.od[2] = nD;
    // nzxr: R of "yxz->zyx" as (nD^3, nD)
    p.is[0] = 1; p.is[1] = nD * nD2; p.is[2] = nD2;
    permute_copy(nz, M, p, 1.0, c.st);
    householder_qr(M, (int)nD3, (int)nD, W, nullptr, nzxr, c.st);
    // nzyr: R of "yxz->zxy"
    p.is[0] = 1; p.is[1] = nD2; p.is[2] = nD * nD2;
    permute_copy(nz, M, p, 1.0, c.st);
    householder_qr(M, (int)nD3, (int)nD, W, nullptr, nzyr, c.st);
    // pseudo-inverses with absolute cutoff (torch.linalg.pinv(atol=...))
    svd_small(nzxr, (int)nD, (int)nD, U, sv, V, c.st);
    pinv_finish(U, sv, V, (int)nD, (int)nD, atol, nzxr_inv, c.st);
    svd_small(nzyr, (int)nD, (int)nD, U, sv, V, c.st);
    pinv_finish(U, sv, V, (int)nD, (int)nD, atol, nzyr_inv, c.st);
  }
  // T[y][z][w] = sum_x nz[y][x][z] nzxr_inv[x][w]
  gemm(c, gdesc((int)(nD * nD2), (int)nD, (int)nD, grouped((int)nD2, 1, nD3), plain(nD2),
                plain(nD), plain(1), plain(nD), plain(1)),
       nz, nzxr_inv, T);
  // nz2[z][v][w] = sum_y nzyr_inv[y][v] T[y][z][w]
  gemm(c, gdesc((int)nD, (int)(nD2 * nD), (int)nD, plain(1), plain(nD), plain(nD2 * nD), plain(1),
                plain(nD), grouped((int)nD, 1, nD2)),
       nzyr_inv, T, nz2);
  // n12[(v,w)][(V,W)] = sum_z nz2[z][(v,w)] nz2[z][(V,W)]
  mm(c, (int)nD2, (int)nD2, (int)nD2, true, nz2, nD2, false, nz2, nD2, n12, nD2);
  // U1[y][z][pq] = sum_x nzxr[z][x] a12g[y][x][pq]
  gemm(c, gdesc((int)nD, (int)(nD * d2), (int)nD, plain(nD), plain(1), plain(d2),
                grouped((int)d2, 1, nD * d2), plain(d2), grouped((int)d2, 1, nD * d2)),
       nzxr, a12g_in, U1);
  // a12g[w][z][pq] = sum_y nzyr[w][y] U1[y][z][pq]
  mm(c, (int)nD, (int)(nD * d2), (int)nD, false, nzyr, nD, false, U1, nD * d2, a12g, nD * d2);
}

// ---------------------------------------------------------------------------
// initialize_reduced_tensors (als_solver.py:117-146), the plain start
// ---------------------------------------------------------------------------
__global__ void plain_start_kernel(const double* __restrict__ U, const double* __restrict__ s,
                                   const double* __restrict__ V, int nD, int bD, int pD,
                                   double* __restrict__ a1r, double* __restrict__ a2r) {
  const int n = nD * pD;
  for (int e = threadIdx.x; e < nD * bD * pD; e += blockDim.x) {
    const int y = e / (bD * pD), u = (e / pD) % bD, p = e % pD;
    const double sc = sqrt(s[u] / s[0]);
    a1r[e] = U[(y * pD + p) * n + u] * sc;
    a2r[e] = V[(y * pD + p) * n + u] * sc;
  }
}

// batched form: workgroup b reads a12g of problem b and writes M, then the
// plain start's a1r/a2r from that problem's U, s, V
__global__ void plain_matrix_batch_kernel(const double* __restrict__ a12g, int nD, int pD,
                                          double* __restrict__ M) {
  const int n = nD * pD;
  const int64_t b = blockIdx.x;
  a12g += b * n * n;
  M += b * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    // M[(y,p)][(x,q)] = a12g[y][x][p][q]  ("yxpq->ypxq")
    const int y = e / (pD * n), p = (e / n) % pD, x = (e % n) / pD, q = e % pD;
    M[e] = a12g[((y * nD + x) * pD + p) * pD + q];
  }
}

__global__ void plain_start_batch_kernel(const double* __restrict__ U,
                                         const double* __restrict__ s,
                                         const double* __restrict__ V, int nD, int bD, int pD,
                                         double* __restrict__ a1r, double* __restrict__ a2r) {
  const int n = nD * pD;
  const int64_t b = blockIdx.x;
  U += b * n * n;
  V += b * n * n;
  s += b * n;
  a1r += b * nD * bD * pD;
  a2r += b * nD * bD * pD;
  for (int e = threadIdx.x; e < nD * bD * pD; e += blockDim.x) {
    const int y = e / (bD * pD), u = (e / pD) % bD, p = e % pD;
    const double sc = sqrt(s[u] / s[0]);
    a1r[e] = U[(y * pD + p) * n + u] * sc;
    a2r[e] = V[(y * pD + p) * n + u] * sc;
  }
}

void plain_start(Ctx& c, const double* a12g, const FuDims& f, double* a1r, double* a2r) {
  const int64_t D = f.D, d = f.d, nD = f.nD, n = nD * d;
  ACETN_REQUIRE(n <= kSvdMaxN && D <= n,
                "ALS plain start: needs nD d <= 32 and D <= nD d (the SVD of a12g as (nD d, nD d))");
  double* M = c.take((size_t)n * n);
  double* U = c.take((size_t)n * n);
  double* sv = c.take((size_t)n);
  double* V = c.take((size_t)n * n);
  if (!c.launch) return;
  PermDesc p{};
  p.n = 4; p.total = n * n;
  p.od[0] = nD; p.od[1] = d; p.od[2] = nD; p.od[3] = d;   // "yxpq->ypxq"
  p.is[0] = nD * d * d; p.is[1] = d; p.is[2] = d * d; p.is[3] = 1;
  permute_copy(a12g, M, p, 1.0, c.st);
  svd_small(M, (int)n, (int)n, U, sv, V, c.st);
  hipLaunchKernelGGL(plain_start_kernel, dim3(1), dim3(256), 0, c.st, U, sv, V, (int)nD, (int)D,
                     (int)d, a1r, a2r);
  ACETN_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// recompose_site_tensors (tensor_update.py:70-75), optionally each tensor
// divided by its norm (full_update.py:61-63), rotated back to the site frame
// (tensor_update.py:26-38).
// ---------------------------------------------------------------------------
void recompose(Ctx& c, const double* b1, const double* b2, const double* a1q, const double* a2q,
               int k, const FuDims& f, bool normalize, double* A1_out, double* A2_out) {
  k &= 3;
  const int64_t D = f.D, d = f.d, nD = f.nD, Dd = D * d;
  double* T = c.take((size_t)D * D * D * Dd);
  double* nrm = c.take(1024);
  int64_t bs1[4], bs2[4];
  // a1 = "rdux,xlp->lurdp": T laid out (r, d, u, l, p)
  mm(c, (int)(D * D * D), (int)Dd, (int)nD, false, a1q, nD, false, b1, Dd, T, Dd);
  bs1[0] = d; bs1[1] = D * d; bs1[2] = D * D * D * d; bs1[3] = D * D * d;
  if (c.launch) {
    if (normalize) frobenius_normalize(T, D * D * D * Dd, nrm, c.st);
    PermDesc p{};
    p.n = 5; p.total = D * D * D * Dd;
    for (int i = 0; i < 4; ++i) { p.od[i] = D; p.is[i] = bs1[(i - k + 4) & 3]; }
    p.od[4] = d; p.is[4] = 1;
    permute_copy(T, A1_out, p, 1.0, c.st);
  }
  // a2 = "dlux,xrp->lurdp": T laid out (d, l, u, r, p)
  mm(c, (int)(D * D * D), (int)Dd, (int)nD, false, a2q, nD, false, b2, Dd, T, Dd);
  bs2[0] = D * D * d; bs2[1] = D * d; bs2[2] = d; bs2[3] = D * D * D * d;
  if (c.launch) {
    if (normalize) frobenius_normalize(T, D * D * D * Dd, nrm, c.st);
    PermDesc p{};
    p.n = 5; p.total = D * D * D * Dd;
    for (int i = 0; i < 4; ++i) { p.od[i] = D; p.is[i] = bs2[(i - k + 4) & 3]; }
    p.od[4] = d; p.is[4] = 1;
    permute_copy(T, A2_out, p, 1.0, c.st);
  }
}

// ---------------------------------------------------------------------------
// finalize_reduced_tensors (full_update.py:123-161): the gauge inverses (when
// given), then the QR-QR-SVD rebalancing with sqrt(s / |s|), then the
// bond-index sign gauge; b1/b2 (nD, D, d) receive the result.
// ---------------------------------------------------------------------------
void finalize_reduced(Ctx& c, const double* a1r_in, const double* a2r_in, const double* nzyr_inv,
                      const double* nzxr_inv, const FuDims& f, double* b1, double* b2) {
  const int64_t D = f.D, d = f.d, nD = f.nD, Dd = D * d;
  ACETN_REQUIRE(D <= kSvdMaxN && D <= nD * d,
                "finalize_reduced: needs D <= 32 and D <= nD d (QR of (nD d, D))");
  double* a1r = c.take((size_t)nD * Dd);
  double* a2r = c.take((size_t)nD * Dd);
  double* P = c.take((size_t)nD * Dd);
  double* W = c.take((size_t)nD * Dd);
  double* q1 = c.take((size_t)nD * Dd);
  double* q2 = c.take((size_t)nD * Dd);
  double* r1 = c.take((size_t)D * D);
  double* r2 = c.take((size_t)D * D);
  double* Mr = c.take((size_t)D * D);
  double* U = c.take((size_t)D * D);
  double* sv = c.take((size_t)D);
  double* V = c.take((size_t)D * D);
  double* r1n = c.take((size_t)D * D);
  double* r2n = c.take((size_t)D * D);
  if (nzyr_inv != nullptr || !c.run) {
    mm(c, (int)nD, (int)Dd, (int)nD, false, nzyr_inv, nD, false, a1r_in, Dd, a1r, Dd);
    mm(c, (int)nD, (int)Dd, (int)nD, false, nzxr_inv, nD, false, a2r_in, Dd, a2r, Dd);
  } else if (c.launch) {
    ACETN_HIP_CHECK(hipMemcpyAsync(a1r, a1r_in, sizeof(double) * nD * Dd, hipMemcpyDeviceToDevice, c.st));
    ACETN_HIP_CHECK(hipMemcpyAsync(a2r, a2r_in, sizeof(double) * nD * Dd, hipMemcpyDeviceToDevice, c.st));
  }
  if (c.launch) {
    PermDesc p{};
    p.n = 3; p.total = nD * Dd;
    p.od[0] = nD; p.od[1] = d; p.od[2] = D;   // "yup->ypu"
    p.is[0] = Dd; p.is[1] = 1; p.is[2] = d;
    permute_copy(a1r, P, p, 1.0, c.st);
    householder_qr(P, (int)(nD * d), (int)D, W, q1, r1, c.st);
    permute_copy(a2r, P, p, 1.0, c.st);
    householder_qr(P, (int)(nD * d), (int)D, W, q2, r2, c.st);
  }
  // "au,bu->ab"
  mm(c, (int)D, (int)D, (int)D, false, r1, D, true, r2, D, Mr, D);
  if (c.launch) {
    svd_small(Mr, (int)D, (int)D, U, sv, V, c.st);
    finalize_scale(U, sv, V, (int)D, (int)D, r1n, r2n, c.st);
  }
  // a1r[y][u][p] = sum_a q1[(y,p)][a] r1n[a][u] ; a2r[x][v][q] = sum_b q2[(x,q)][b] r2n[v][b]
  gemm(c, gdesc((int)(nD * d), (int)D, (int)D, plain(D), plain(1), plain(D), plain(1),
                grouped((int)d, 1, Dd), plain(d)),
       q1, r1n, b1);
  gemm(c, gdesc((int)(nD * d), (int)D, (int)D, plain(D), plain(1), plain(1), plain(D),
                grouped((int)d, 1, Dd), plain(d)),
       q2, r2n, b2);
  if (c.launch) bond_gauge(b1, b2, (int)nD, (int)nD, (int)D, (int)d, c.st);
}

// ---------------------------------------------------------------------------
// finalize_reduced + recompose_site_tensors (tensor_update.py:70-75) +
// normalisation + the inverse bond permutation (tensor_update.py:26-38).
// A1_out/A2_out are in the site frame.
// ---------------------------------------------------------------------------
void finalize_bond(Ctx& c, const double* a1r_in, const double* a2r_in, const double* nzyr_inv,
                   const double* nzxr_inv, const double* a1q, const double* a2q, int k,
                   const FuDims& f, double* A1_out, double* A2_out) {
  const int64_t Dd = f.D * f.d;
  double* b1 = c.take((size_t)f.nD * Dd);
  double* b2 = c.take((size_t)f.nD * Dd);
  finalize_reduced(c, a1r_in, a2r_in, nzyr_inv, nzxr_inv, f, b1, b2);
  recompose(c, b1, b2, a1q, a2q, k & 3, f, true, A1_out, A2_out);
}

// ---------------------------------------------------------------------------
// SimpleUpdater.bond_update (simple_update.py:28-76): lambdas absorbed on the
// non-bond legs, QR split, gate, SVD truncated to D, new bond lambda, sign
// gauge of the bond index, recomposition, lambdas stripped.
// ---------------------------------------------------------------------------
void simple_update_bond(Ctx& c, const double* A1, const double* A2, int k,
                        const LambdaLegs& w1, const LambdaLegs& w2, const double* gate,
                        const FuDims& f, double cutoff, double* A1_out, double* A2_out,
                        double* lam_out) {
  const int64_t D = f.D, d = f.d, nD = f.nD, n = nD * d, nA = D * D * D * D * d;
  double* B1 = c.take((size_t)nA);
  double* B2 = c.take((size_t)nA);
  double* a1q = c.take((size_t)D * D * D * nD);
  double* a2q = c.take((size_t)D * D * D * nD);
  double* a1r = c.take((size_t)nD * D * d);
  double* a2r = c.take((size_t)nD * D * d);
  double* a12g = c.take((size_t)n * n);
  double* M = c.take((size_t)n * n);
  double* U = c.take((size_t)n * n);
  double* sv = c.take((size_t)n);
  double* V = c.take((size_t)n * n);
  if (c.launch) {
    lambda_scale(A1, B1, w1, (int)D, (int)d, 0, cutoff, c.st);
    lambda_scale(A2, B2, w2, (int)D, (int)d, 0, cutoff, c.st);
  }
  decompose(c, B1, B2, k, f, a1q, a1r, a2q, a2r);
  apply_gate(c, a1r, a2r, gate, f, a12g);
  if (c.launch) {
    PermDesc p{};
    p.n = 4; p.total = n * n;
    p.od[0] = nD; p.od[1] = d; p.od[2] = nD; p.od[3] = d;   // "yxrs->yrxs"
    p.is[0] = nD * d * d; p.is[1] = d; p.is[2] = d * d; p.is[3] = 1;
    permute_copy(a12g, M, p, 1.0, c.st);
    svd_small(M, (int)n, (int)n, U, sv, V, c.st);
    su_split(U, sv, V, (int)nD, (int)D, (int)d, a1r, a2r, lam_out, c.st);
    bond_gauge(a1r, a2r, (int)nD, (int)nD, (int)D, (int)d, c.st);
  }
  recompose(c, a1r, a2r, a1q, a2q, k, f, false, B1, B2);
  if (c.launch) {
    lambda_scale(B1, A1_out, w1, (int)D, (int)d, 1, cutoff, c.st);
    lambda_scale(B2, A2_out, w2, (int)D, (int)d, 1, cutoff, c.st);
  }
}

}  // namespace
}  // namespace acetn

using namespace acetn;

namespace {

// Scratch of acetn_hip_als_solve_noalloc: the ALS system matrix and
// right-hand sides of the largest problem it accepts (nD * D <= 128 rows,
// d <= 4), then the final cost and the sweep / flag words.  One copy per
// device, held until the process ends; the kernel writes every element it
// reads, so nothing has to start at zero.
constexpr int kNoAllocMaxNR = 128;
constexpr size_t kNoAllocWork = (size_t)kNoAllocMaxNR * kNoAllocMaxNR + (size_t)kNoAllocMaxNR * 4;
constexpr size_t kNoAllocElems = kNoAllocWork + 4 + 4;
__device__ __align__(256) double scratch[kNoAllocElems];

// the device address of ``scratch`` on each device, looked up once per
// device with that device current (the host-side name is only a stand-in)
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;

double* noalloc_scratch() {
  int dev = 0;
  ACETN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    ACETN_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ACETN_REQUIRE(p != nullptr, "als_solve_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = reinterpret_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

size_t acetn_hip_bond_decompose_workspace(int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = fu_dims(D, d);
    r = measure([&](Ctx& c) { decompose(c, nullptr, nullptr, 0, f, nullptr, nullptr, nullptr, nullptr); });
  });
  return r;
}

int acetn_hip_bond_decompose(const double* A1, const double* A2, int k, int64_t D, int64_t d,
                             double* a1q, double* a1r, double* a2q, double* a2r, void* ws,
                             size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = fu_dims(D, d);
    ACETN_REQUIRE(A1 && A2 && a1q && a1r && a2q && a2r, "bond_decompose: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { decompose(c, A1, A2, k, f, a1q, a1r, a2q, a2r); });
  });
}

size_t acetn_hip_norm_tensor_workspace(const acetn_site* s1, const acetn_site* s2, int k) {
  size_t r = 0;
  guarded([&] {
    check_site(s1);
    check_site(s2);
    r = measure([&](Ctx& c) { norm_tensor(c, s1, s2, k, nullptr, nullptr, nullptr); });
  });
  return r;
}

int acetn_hip_norm_tensor(const acetn_site* s1, const acetn_site* s2, int k, const double* a1q,
                          const double* a2q, double* n12, void* ws, size_t ws_bytes,
                          void* stream) {
  return guarded([&] {
    check_site(s1);
    check_site(s2);
    ACETN_REQUIRE(a1q && a2q && n12, "norm_tensor: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { norm_tensor(c, s1, s2, k, a1q, a2q, n12); });
  });
}

size_t acetn_hip_apply_gate_workspace(int64_t nD, int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    r = measure([&](Ctx& c) { apply_gate(c, nullptr, nullptr, nullptr, f, nullptr); });
  });
  return r;
}

int acetn_hip_apply_gate(const double* a1r, const double* a2r, const double* gate, int64_t nD,
                         int64_t D, int64_t d, double* a12g, void* ws, size_t ws_bytes,
                         void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(a1r && a2r && gate && a12g, "apply_gate: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { apply_gate(c, a1r, a2r, gate, f, a12g); });
  });
}

size_t acetn_hip_positive_approx_workspace(int64_t nD) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, 1, 1);  // only nD enters
    r = measure([&](Ctx& c) { positive_approx(c, nullptr, f, 0.0, nullptr, nullptr); });
  });
  return r;
}

int acetn_hip_positive_approx(const double* n12, int64_t nD, double cutoff, double* nz,
                              int32_t* host_shifts, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, 1, 1);
    ACETN_REQUIRE(n12 && nz, "positive_approx: null pointer");
    int shifts = 0;
    execute(ws, ws_bytes, stream, [&](Ctx& c) { positive_approx(c, n12, f, cutoff, nz, &shifts); });
    if (host_shifts) *host_shifts = shifts;
  });
}

size_t acetn_hip_gauge_fix_workspace(int64_t nD, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, 1, d);  // D does not enter
    r = measure([&](Ctx& c) {
      gauge_fix(c, nullptr, nullptr, f, 0.0, nullptr, nullptr, nullptr, nullptr);
    });
  });
  return r;
}

int acetn_hip_gauge_fix(const double* nz, const double* a12g_in, int64_t nD, int64_t d,
                        double atol, double* n12, double* a12g, double* nzxr_inv,
                        double* nzyr_inv, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, 1, d);
    ACETN_REQUIRE(nz && a12g_in && n12 && a12g && nzxr_inv && nzyr_inv, "gauge_fix: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      gauge_fix(c, nz, a12g_in, f, atol, n12, a12g, nzxr_inv, nzyr_inv);
    });
  });
}

size_t acetn_hip_als_workspace(int64_t nD, int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    als_check_shape((int)f.nD, (int)f.D, (int)f.d, 0);
    r = measure([&](Ctx& c) {
      // the plain start where it exists (acetn_hip_als_initialize reports
      // the shapes it cannot start from)
      if (f.nD * f.d <= kSvdMaxN && f.D <= f.nD * f.d) plain_start(c, nullptr, f, nullptr, nullptr);
      c.take(als_work_elems((int)f.nD, (int)f.D, (int)f.d));
      c.take(2);
      c.take_int(4);
    });
  });
  return r;
}

int acetn_hip_als_initialize(const double* a12g, int64_t nD, int64_t D, int64_t d, double* a1r,
                             double* a2r, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(a12g && a1r && a2r, "als_initialize: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) { plain_start(c, a12g, f, a1r, a2r); });
  });
}

int acetn_hip_als_solve(const double* n12, const double* a12g, double* a1r, double* a2r,
                        int64_t nD, int64_t D, int64_t d, int niter, double tol, int method,
                        double epsilon, int64_t* host_sweeps, double* host_cost,
                        int32_t* host_info, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(n12 && a12g && a1r && a2r, "als_solve: null pointer");
    ACETN_REQUIRE(method == ACETN_HIP_ALS_CHOLESKY || method == ACETN_HIP_ALS_PINV,
                  "als_solve: unknown method");
    ACETN_REQUIRE(niter >= 0, "als_solve: negative niter");
    als_check_shape((int)f.nD, (int)f.D, (int)f.d, method == ACETN_HIP_ALS_PINV ? 1 : 0);
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      double* gw = c.take(als_work_elems((int)f.nD, (int)f.D, (int)f.d));
      double* cost = c.take(2);
      int* info = c.take_int(4);
      if (!c.launch) return;
      als_solve_device(n12, a12g, a1r, a2r, (int)f.nD, (int)f.D, (int)f.d, niter, tol, epsilon,
                       gw, cost, info, c.st, method == ACETN_HIP_ALS_PINV ? 1 : 0);
      double hc = 0.0;
      int hi[2] = {0, 0};
      ACETN_HIP_CHECK(hipMemcpyAsync(&hc, cost, sizeof(double), hipMemcpyDeviceToHost, c.st));
      ACETN_HIP_CHECK(hipMemcpyAsync(hi, info, 2 * sizeof(int), hipMemcpyDeviceToHost, c.st));
      ACETN_HIP_CHECK(hipStreamSynchronize(c.st));
      if (host_sweeps) *host_sweeps = hi[0];
      if (host_cost) *host_cost = hc;
      if (host_info) *host_info = hi[1];
    });
  });
}

size_t acetn_hip_als_batch_workspace(int64_t count, int64_t nD, int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(count >= 1 && count <= (int64_t(1) << 24), "als_batch: count out of range");
    const int64_t n = f.nD * f.d;
    r = measure([&](Ctx& c) {
      c.take((size_t)count * (3 * n * n + n));
      c.take((size_t)count * als_work_elems((int)f.nD, (int)f.D, (int)f.d));
    });
  });
  return r;
}

int acetn_hip_als_solve_batch(const double* n12, const double* a12g, double* a1r, double* a2r,
                              int64_t count, int64_t nD, int64_t D, int64_t d, int niter,
                              double tol, int method, double epsilon, int plain_start,
                              double* dev_cost, int32_t* dev_info, void* ws, size_t ws_bytes,
                              void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(n12 && a12g && a1r && a2r && dev_cost && dev_info, "als_batch: null pointer");
    ACETN_REQUIRE(count >= 1 && count <= (int64_t(1) << 24), "als_batch: count out of range");
    ACETN_REQUIRE(method == ACETN_HIP_ALS_CHOLESKY || method == ACETN_HIP_ALS_PINV,
                  "als_batch: unknown method");
    ACETN_REQUIRE(niter >= 0, "als_batch: negative niter");
    const int64_t n = f.nD * f.d;
    ACETN_REQUIRE(!plain_start || (n <= kSvdMaxN && f.D <= n),
                  "als_batch: the plain start needs nD d <= 32 and D <= nD d");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      double* M = c.take((size_t)count * n * n);
      double* U = c.take((size_t)count * n * n);
      double* V = c.take((size_t)count * n * n);
      double* sv = c.take((size_t)count * n);
      double* gw = c.take((size_t)count * als_work_elems((int)f.nD, (int)f.D, (int)f.d));
      if (!c.launch) return;
      if (plain_start) {
        hipLaunchKernelGGL(plain_matrix_batch_kernel, dim3(count), dim3(256), 0, c.st, a12g,
                           (int)f.nD, (int)f.d, M);
        ACETN_HIP_CHECK(hipGetLastError());
        // U and V share one stride so that svd_small_batch can offset both
        svd_small_batch(M, (int)n, (int)n, U, sv, V, (int)count, n * n, c.st);
        hipLaunchKernelGGL(plain_start_batch_kernel, dim3(count), dim3(256), 0, c.st, U, sv, V,
                           (int)f.nD, (int)f.D, (int)f.d, a1r, a2r);
        ACETN_HIP_CHECK(hipGetLastError());
      }
      als_solve_device(n12, a12g, a1r, a2r, (int)f.nD, (int)f.D, (int)f.d, niter, tol, epsilon,
                       gw, dev_cost, dev_info, c.st, method == ACETN_HIP_ALS_PINV ? 1 : 0,
                       (int)count);
    });
  });
}

int acetn_hip_als_solve_noalloc(const double* n12, const double* a12g, double* a1r, double* a2r,
                                int64_t nD, int64_t D, int64_t d, int niter, double tol,
                                int method, double epsilon, int64_t* host_sweeps,
                                double* host_cost, int32_t* host_info, void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(n12 && a12g && a1r && a2r, "als_solve_noalloc: null pointer");
    ACETN_REQUIRE(method == ACETN_HIP_ALS_CHOLESKY,
                  "als_solve_noalloc: only the 'cholesky' method (with its LU fallback) runs on "
                  "the device");
    ACETN_REQUIRE(niter >= 0, "als_solve_noalloc: negative niter");
    ACETN_REQUIRE(f.nD * f.D <= kNoAllocMaxNR && f.d <= 4,
                  "als_solve_noalloc: nD * D must be at most 128 (the fixed scratch holds the "
                  "system of D = 8, d = 2); use acetn_hip_als_solve for larger bonds");
    ACETN_REQUIRE(als_work_elems((int)f.nD, (int)f.D, (int)f.d) <= kNoAllocWork,
                  "als_solve_noalloc: problem exceeds the fixed scratch");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    double* base = noalloc_scratch();
    double* cost = base + kNoAllocWork;
    int* info = reinterpret_cast<int*>(base + kNoAllocWork + 4);
    als_solve_device(n12, a12g, a1r, a2r, (int)f.nD, (int)f.D, (int)f.d, niter, tol, epsilon, base,
                     cost, info, st);
    double hc = 0.0;
    int hi[2] = {0, 0};
    ACETN_HIP_CHECK(hipMemcpyAsync(&hc, cost, sizeof(double), hipMemcpyDeviceToHost, st));
    ACETN_HIP_CHECK(hipMemcpyAsync(hi, info, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    ACETN_HIP_CHECK(hipStreamSynchronize(st));
    if (host_sweeps) *host_sweeps = hi[0];
    if (host_cost) *host_cost = hc;
    if (host_info) *host_info = hi[1];
  });
}

size_t acetn_hip_finalize_bond_workspace(int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = fu_dims(D, d);
    r = measure([&](Ctx& c) {
      finalize_bond(c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, f, nullptr,
                    nullptr);
    });
  });
  return r;
}

int acetn_hip_finalize_bond(const double* a1r, const double* a2r, const double* nzyr_inv,
                            const double* nzxr_inv, const double* a1q, const double* a2q, int k,
                            int64_t D, int64_t d, double* A1_out, double* A2_out, void* ws,
                            size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = fu_dims(D, d);
    ACETN_REQUIRE(a1r && a2r && a1q && a2q && A1_out && A2_out, "finalize_bond: null pointer");
    ACETN_REQUIRE((nzyr_inv == nullptr) == (nzxr_inv == nullptr),
                  "finalize_bond: give both gauge inverses or neither");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      finalize_bond(c, a1r, a2r, nzyr_inv, nzxr_inv, a1q, a2q, k, f, A1_out, A2_out);
    });
  });
}

size_t acetn_hip_finalize_reduced_workspace(int64_t nD, int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    r = measure([&](Ctx& c) {
      finalize_reduced(c, nullptr, nullptr, nullptr, nullptr, f, nullptr, nullptr);
    });
  });
  return r;
}

int acetn_hip_finalize_reduced(const double* a1r, const double* a2r, const double* nzyr_inv,
                               const double* nzxr_inv, int64_t nD, int64_t D, int64_t d,
                               double* a1r_out, double* a2r_out, void* ws, size_t ws_bytes,
                               void* stream) {
  return guarded([&] {
    const FuDims f = solver_dims(nD, D, d);
    ACETN_REQUIRE(a1r && a2r && a1r_out && a2r_out, "finalize_reduced: null pointer");
    ACETN_REQUIRE((nzyr_inv == nullptr) == (nzxr_inv == nullptr),
                  "finalize_reduced: give both gauge inverses or neither");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      finalize_reduced(c, a1r, a2r, nzyr_inv, nzxr_inv, f, a1r_out, a2r_out);
    });
  });
}

size_t acetn_hip_recompose_workspace(int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = fu_dims(D, d);
    r = measure([&](Ctx& c) {
      recompose(c, nullptr, nullptr, nullptr, nullptr, 0, f, false, nullptr, nullptr);
    });
  });
  return r;
}

int acetn_hip_recompose(const double* a1r, const double* a2r, const double* a1q,
                        const double* a2q, int k, int normalize, int64_t D, int64_t d,
                        double* A1_out, double* A2_out, void* ws, size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = fu_dims(D, d);
    ACETN_REQUIRE(a1r && a2r && a1q && a2q && A1_out && A2_out, "recompose: null pointer");
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      recompose(c, a1r, a2r, a1q, a2q, k & 3, f, normalize != 0, A1_out, A2_out);
    });
  });
}

size_t acetn_hip_simple_update_workspace(int64_t D, int64_t d) {
  size_t r = 0;
  guarded([&] {
    const FuDims f = fu_dims(D, d);
    LambdaLegs w{};
    r = measure([&](Ctx& c) {
      simple_update_bond(c, nullptr, nullptr, 0, w, w, nullptr, f, 0.0, nullptr, nullptr, nullptr);
    });
  });
  return r;
}

int acetn_hip_simple_update_bond(const double* A1, const double* A2, int k,
                                 const double* const lam1[4], const double* const lam2[4],
                                 const double* gate, int64_t D, int64_t d, double cutoff,
                                 double* A1_out, double* A2_out, double* lam_out, void* ws,
                                 size_t ws_bytes, void* stream) {
  return guarded([&] {
    const FuDims f = fu_dims(D, d);
    ACETN_REQUIRE(A1 && A2 && gate && A1_out && A2_out && lam_out, "simple_update: null pointer");
    LambdaLegs w1{}, w2{};
    for (int i = 0; i < 4; ++i) {
      w1.v[i] = lam1 ? lam1[i] : nullptr;
      w2.v[i] = lam2 ? lam2[i] : nullptr;
    }
    execute(ws, ws_bytes, stream, [&](Ctx& c) {
      simple_update_bond(c, A1, A2, k, w1, w2, gate, f, cutoff, A1_out, A2_out, lam_out);
    });
  });
}

}  // extern "C"
