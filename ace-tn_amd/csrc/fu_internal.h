// Internal declarations of the full-update kernels (fu_kernels.hip).
#pragma once
#include "internal.h"

namespace acetn {

constexpr int kQrMaxN = 32;   // Householder QR: at most 32 columns
constexpr int kSvdMaxN = 32;  // one-sided Jacobi SVD: at most 64 x 32

// out (contiguous, dims od[0..n-1]) = scale * in[sum_i idx_i * is[i]]
struct PermDesc {
  int n;
  int64_t od[6];
  int64_t is[6];
  int64_t total;
};
void permute_copy(const double* in, double* out, const PermDesc& p, double scale, hipStream_t s);

// Householder QR (LAPACK sign conventions) of a row-major m x n matrix:
// W (m x n) scratch, Q (m x n, thin; may be null), R (n x n).
void householder_qr(const double* A, int m, int n, double* W, double* Q, double* R, hipStream_t s);

// One-sided Jacobi SVD, A (m x n, m <= 64, n <= 32) = U diag(s) V^T,
// U (m x n), V (n x n), s descending.
void svd_small(const double* A, int m, int n, double* U, double* sv, double* V, hipStream_t s);
// ``count`` problems, matrices (A, U, V) ``stride`` doubles apart, sv n apart.
void svd_small_batch(const double* A, int m, int n, double* U, double* sv, double* V, int count,
                     int64_t stride, hipStream_t s);

void lower_to_full(double* M, int n, hipStream_t s);
void add_diag(double* M, int n, double shift, hipStream_t s);
void scale_sqrt_columns(const double* Z, const double* w, int rows, int cols, double* X,
                        hipStream_t s);
// P (n x m) = V diag(1/s_k for s_k > atol) U^T
void pinv_finish(const double* U, const double* sv, const double* V, int m, int n, double atol,
                 double* P, hipStream_t s);
void finalize_scale(const double* U, const double* sv, const double* V, int n, int bD,
                    double* r1n, double* r2n, hipStream_t s);
void bond_gauge(double* a1r, double* a2r, int nD1, int nD2, int bD, int pD, hipStream_t s);

// Simple-update helpers: leg weights (null = ones) and the SVD split.
struct LambdaLegs {
  const double* v[4];
};
void lambda_scale(const double* A, double* out, const LambdaLegs& w, int D, int d, int invert,
                  double cutoff, hipStream_t s);
void su_split(const double* U, const double* S, const double* V, int nD, int bD, int pD,
              double* a1r, double* a2r, double* lt, hipStream_t s);

// Persistent ALS (one workgroup).  out_info[0] = sweeps, out_info[1] = flags
// (1: a Cholesky factorisation failed and LU was used; 2: singular system).
size_t als_work_elems(int nD, int D, int d);
size_t als_lds_bytes(int nD, int D, int d);
// Raises unless the one-workgroup ALS holds the problem (method 0 cholesky, 1 pinv).
void als_check_shape(int nD, int D, int d, int method);
// With count > 1, problem b's n12, a12g, a1r, a2r, gwork, out_cost and
// out_info follow problem b-1's contiguously (one workgroup per problem).
void als_solve_device(const double* n12, const double* a12g, double* a1r, double* a2r, int nD,
                      int D, int d, int niter, double tol, double eps, double* gwork,
                      double* out_cost, int* out_info, hipStream_t s, int method = 0,
                      int count = 1);

}  // namespace acetn
