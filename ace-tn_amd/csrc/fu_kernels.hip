// Kernels of the full-update bond step (acetn/evolution/full_update.py,
// als_solver.py, tensor_update.py): tiny dense linear algebra that runs in
// one workgroup per problem (latency-bound; the matrices are 6x6 .. 1728x12),
// the persistent ALS solver, and a strided permute-copy.
#include "internal.h"
#include "fu_internal.h"

#include <algorithm>
#include <cmath>

namespace acetn {
namespace {

// ---------------------------------------------------------------------------
// permute-copy: out (contiguous, dims od[0..n-1]) = scale * in[sum_i idx_i * is[i]]
// ---------------------------------------------------------------------------
__global__ void permute_kernel(const double* __restrict__ in, double* __restrict__ out,
                               PermDesc p, double scale) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < p.total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e, off = 0;
    for (int i = p.n - 1; i >= 0; --i) {
      const int64_t x = r % p.od[i];
      r /= p.od[i];
      off += x * p.is[i];
    }
    out[e] = scale * in[off];
  }
}

// ---------------------------------------------------------------------------
// Householder QR of a dense row-major m x n matrix (n <= kQrMaxN), one
// workgroup.  Conventions of LAPACK dgeqrf/dorgqr (and so of numpy.linalg.qr):
// beta = -sign(alpha) * ||x||, R[j][j] = beta, thin Q formed by backward
// accumulation.  W (m x n) is scratch and receives the reflectors.
// ---------------------------------------------------------------------------
constexpr int QR_NT = 256;

__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;
}

__global__ __launch_bounds__(QR_NT) void householder_qr_kernel(const double* __restrict__ A,
                                                               int m, int n,
                                                               double* __restrict__ W,
                                                               double* __restrict__ Q,
                                                               double* __restrict__ R) {
  __shared__ double red[QR_NT / 64];
  __shared__ double part[QR_NT / 64][kQrMaxN];
  __shared__ double tau[kQrMaxN], wv[kQrMaxN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int64_t e = tid; e < (int64_t)m * n; e += QR_NT) W[e] = A[e];
  __syncthreads();
  const int kmax = min(m, n);
  for (int j = 0; j < kmax; ++j) {
    double xs = 0.0;
    for (int i = j + 1 + tid; i < m; i += QR_NT) xs += W[(int64_t)i * n + j] * W[(int64_t)i * n + j];
    const double xnorm2 = block_sum(xs, red);
    const double alpha = W[(int64_t)j * n + j];
    double beta = alpha, t = 0.0, scal = 0.0;
    if (xnorm2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    if (tid == 0) { tau[j] = t; W[(int64_t)j * n + j] = beta; }
    for (int i = j + 1 + tid; i < m; i += QR_NT) W[(int64_t)i * n + j] *= scal;
    __syncthreads();
    if (t != 0.0 && j + 1 < n) {
      // w_c = v^T W[:, c] for c > j  (v_j = 1)
      double acc[kQrMaxN];
#pragma unroll
      for (int c = 0; c < kQrMaxN; ++c) acc[c] = 0.0;
      for (int i = j + tid; i < m; i += QR_NT) {
        const double vi = (i == j) ? 1.0 : W[(int64_t)i * n + j];
#pragma unroll
        for (int c = 0; c < kQrMaxN; ++c)
          if (c > j && c < n) acc[c] += vi * W[(int64_t)i * n + c];
      }
#pragma unroll
      for (int c = 0; c < kQrMaxN; ++c) {
        double v = acc[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) part[wave][c] = v;
      }
      __syncthreads();
      if (tid < n) {
        double s = 0.0;
        for (int w = 0; w < QR_NT / 64; ++w) s += part[w][tid];
        wv[tid] = s;
      }
      __syncthreads();
      for (int i = j + tid; i < m; i += QR_NT) {
        const double vi = (i == j) ? 1.0 : W[(int64_t)i * n + j];
        for (int c = j + 1; c < n; ++c) W[(int64_t)i * n + c] -= t * vi * wv[c];
      }
      __syncthreads();
    }
  }
  // R (n x n, upper; rows beyond kmax are zero)
  for (int e = tid; e < n * n; e += QR_NT) {
    const int r = e / n, c = e % n;
    R[e] = (r <= c && r < kmax) ? W[(int64_t)r * n + c] : 0.0;
  }
  if (Q == nullptr) return;
  // thin Q = H_0 ... H_{kmax-1} I[:, :kmax]  (m x kmax, leading dimension n)
  for (int64_t e = tid; e < (int64_t)m * n; e += QR_NT) {
    const int r = (int)(e / n), c = (int)(e % n);
    Q[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = kmax - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;
    double acc[kQrMaxN];
#pragma unroll
    for (int c = 0; c < kQrMaxN; ++c) acc[c] = 0.0;
    for (int i = j + tid; i < m; i += QR_NT) {
      const double vi = (i == j) ? 1.0 : W[(int64_t)i * n + j];
#pragma unroll
      for (int c = 0; c < kQrMaxN; ++c)
        if (c >= j && c < kmax) acc[c] += vi * Q[(int64_t)i * n + c];
    }
#pragma unroll
    for (int c = 0; c < kQrMaxN; ++c) {
      double v = acc[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) part[wave][c] = v;
    }
    __syncthreads();
    if (tid < n) {
      double s = 0.0;
      for (int w = 0; w < QR_NT / 64; ++w) s += part[w][tid];
      wv[tid] = s;
    }
    __syncthreads();
    for (int i = j + tid; i < m; i += QR_NT) {
      const double vi = (i == j) ? 1.0 : W[(int64_t)i * n + j];
      for (int c = j; c < kmax; ++c) Q[(int64_t)i * n + c] -= t * vi * wv[c];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD of a tiny dense m x n matrix (m <= 64,
// n <= 32): A V = U S with S descending.  One wave per column pair (lane i
// holds row i), round-robin pair ordering; a pair is rotated while
// |a_p . a_q| > 1e-15 * ||a_p|| ||a_q||.
// ---------------------------------------------------------------------------
constexpr int SVD_NT = 1024;

__global__ __launch_bounds__(SVD_NT) void svd_small_kernel(const double* __restrict__ A, int m,
                                                           int n, double* __restrict__ U,
                                                           double* __restrict__ s,
                                                           double* __restrict__ V,
                                                           int64_t batch_stride) {
  // workgroup b decomposes problem b: matrices batch_stride doubles apart,
  // singular values n apart
  A += blockIdx.x * batch_stride;
  U += blockIdx.x * batch_stride;
  V += blockIdx.x * batch_stride;
  s += (int64_t)blockIdx.x * n;
  __shared__ double a[kSvdMaxN + 1][64];   // columns, padded to even count
  __shared__ double v[kSvdMaxN + 1][kSvdMaxN + 1];
  __shared__ double nrm[kSvdMaxN + 1];
  __shared__ int order[kSvdMaxN + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nn = n + (n & 1);
  for (int e = tid; e < (kSvdMaxN + 1) * 64; e += SVD_NT) {
    const int c = e / 64, r = e % 64;
    a[c][r] = (c < n && r < m) ? A[r * n + c] : 0.0;
  }
  for (int e = tid; e < (kSvdMaxN + 1) * (kSvdMaxN + 1); e += SVD_NT) {
    const int r = e / (kSvdMaxN + 1), c = e % (kSvdMaxN + 1);
    v[r][c] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rot = 0;
    for (int r = 0; r < nn - 1; ++r) {
      const int pr = wave;
      if (pr < nn / 2) {
        int x, y;
        if (pr == 0) { x = nn - 1; y = r; }
        else { x = (r + pr) % (nn - 1); y = (r - pr + nn - 1) % (nn - 1); }
        const int p = min(x, y), q = max(x, y);
        if (q < n) {
          const double ap = a[p][lane], aq = a[q][lane];
          double al = ap * ap, be = aq * aq, ga = ap * aq;
          for (int o = 32; o > 0; o >>= 1) {
            al += __shfl_xor(al, o);
            be += __shfl_xor(be, o);
            ga += __shfl_xor(ga, o);
          }
          if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
            const double zeta = (be - al) / (2.0 * ga);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
            a[p][lane] = c * ap - sn * aq;
            a[q][lane] = sn * ap + c * aq;
            if (lane < n) {
              const double vp = v[lane][p], vq = v[lane][q];
              v[lane][p] = c * vp - sn * vq;
              v[lane][q] = sn * vp + c * vq;
            }
            rot = 1;
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(rot)) break;
  }
  for (int c = wave; c < n; c += SVD_NT / 64) {
    const double x = a[c][lane];
    double ss = x * x;
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    if (lane == 0) nrm[c] = sqrt(ss);
  }
  __syncthreads();
  if (tid < n) {
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (nrm[j] > nrm[tid]) || (nrm[j] == nrm[tid] && j < tid);
    order[rank] = tid;
  }
  __syncthreads();
  for (int e = tid; e < m * n; e += SVD_NT) {
    const int r = e / n, k = e % n, c = order[k];
    U[e] = nrm[c] > 0.0 ? a[c][r] / nrm[c] : 0.0;
  }
  for (int e = tid; e < n * n; e += SVD_NT) {
    const int r = e / n, k = e % n;
    V[e] = v[r][order[k]];
  }
  if (tid < n) s[tid] = nrm[order[tid]];
}

// ---------------------------------------------------------------------------
// small elementwise helpers
// ---------------------------------------------------------------------------
// M[i][j] = M[j][i] for i < j: the symmetric matrix that an eigensolver
// reading only the lower triangle sees (LAPACK's and torch's UPLO='L').
__global__ void lower_to_full_kernel(double* M, int n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (i < j) M[e] = M[(int64_t)j * n + i];
  }
}

__global__ void add_diag_kernel(double* M, int n, double shift) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    M[(int64_t)i * n + i] += shift;
}

// X[r][z] = Z[r][z] * sqrt(w[z])
__global__ void scale_sqrt_kernel(const double* __restrict__ Z, const double* __restrict__ w,
                                  int rows, int cols, double* __restrict__ X) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)rows * cols;
       e += (int64_t)gridDim.x * blockDim.x)
    X[e] = Z[e] * sqrt(w[e % cols]);
}

// pinv = V diag(1/s, s > atol) U^T  (U m x n, V n x n; result n x m)
__global__ void pinv_finish_kernel(const double* __restrict__ U, const double* __restrict__ s,
                                   const double* __restrict__ V, int m, int n, double atol,
                                   double* __restrict__ P) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * m; e += gridDim.x * blockDim.x) {
    const int i = e / m, j = e % m;
    double acc = 0.0;
    for (int k = 0; k < n; ++k)
      if (s[k] > atol) acc += V[i * n + k] * (1.0 / s[k]) * U[j * n + k];
    P[e] = acc;
  }
}

// finalize_reduced_tensors scaling: s' = sqrt(s[:bD] / ||s||);
// r1n[a][u] = U[a][u] s'_u (U n x n), r2n[v][b] = s'_v Vt[v][b] = s'_v V[b][v]
__global__ void finalize_scale_kernel(const double* __restrict__ U, const double* __restrict__ s,
                                      const double* __restrict__ V, int n, int bD,
                                      double* __restrict__ r1n, double* __restrict__ r2n) {
  __shared__ double nrm;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += s[k] * s[k];
    nrm = sqrt(t);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * bD; e += blockDim.x) {
    const int a = e / bD, u = e % bD;
    const double sp = sqrt(s[u] / nrm);
    r1n[a * bD + u] = U[a * n + u] * sp;
    r2n[u * n + a] = sp * V[a * n + u];
  }
}

// Bond-index sign gauge shared with the oracle: sign(sum_i w_i f[u][i]) > 0
// with f[u][(y, p)] = a1r[y][u][p] and w_i = 1 + frac(i * 0.6180339887498949).
__global__ void bond_gauge_kernel(double* __restrict__ a1r, double* __restrict__ a2r, int nD1,
                                  int nD2, int bD, int pD) {
  __shared__ double red[4];
  __shared__ double sgn;
  const int u = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nD1 * pD; i += blockDim.x) {
    const int y = i / pD, p = i % pD;
    const double x = (double)i * 0.6180339887498949;
    acc += (1.0 + (x - floor(x))) * a1r[(y * bD + u) * pD + p];
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) sgn = (t < 0.0) ? -1.0 : 1.0;
  __syncthreads();
  for (int i = threadIdx.x; i < nD1 * pD; i += blockDim.x) {
    const int y = i / pD, p = i % pD;
    a1r[(y * bD + u) * pD + p] *= sgn;
  }
  for (int i = threadIdx.x; i < nD2 * pD; i += blockDim.x) {
    const int x = i / pD, q = i % pD;
    a2r[(x * bD + u) * pD + q] *= sgn;
  }
}

// Simple update: out = A * prod_legs w_leg (or 1 / max(w_leg, cutoff) when
// invert) over the four bond legs; a null weight vector is all ones.
__global__ void lambda_scale_kernel(const double* __restrict__ A, double* __restrict__ out,
                                    LambdaLegs w, int D, int d, int invert, double cutoff) {
  const int64_t n = (int64_t)D * D * D * D * d;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e / d;
    double f = 1.0;
    for (int leg = 3; leg >= 0; --leg) {
      const int i = (int)(r % D);
      r /= D;
      if (w.v[leg] != nullptr) {
        const double x = w.v[leg][i];
        f *= invert ? 1.0 / fmax(x, cutoff) : x;
      }
    }
    out[e] = A[e] * f;
  }
}

// Simple-update split of the gated bond matrix (simple_update.py:53-61):
// lt = sqrt(S[:bD] / ||S[:bD]||), a1r[y][a][p] = U[(y,p)][a] lt[a],
// a2r[x][a][q] = V[(x,q)][a] lt[a]  (U, V n x n with n = nD pD).
__global__ void su_split_kernel(const double* __restrict__ U, const double* __restrict__ S,
                                const double* __restrict__ V, int nD, int bD, int pD,
                                double* __restrict__ a1r, double* __restrict__ a2r,
                                double* __restrict__ lt) {
  __shared__ double nrm;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int a = 0; a < bD; ++a) t += S[a] * S[a];
    nrm = sqrt(t);
  }
  __syncthreads();
  const int n = nD * pD;
  for (int e = threadIdx.x; e < nD * bD * pD; e += blockDim.x) {
    const int y = e / (bD * pD), a = (e / pD) % bD, p = e % pD;
    const double l = sqrt(S[a] / nrm);
    a1r[e] = U[(y * pD + p) * n + a] * l;
    a2r[e] = V[(y * pD + p) * n + a] * l;
  }
  for (int a = threadIdx.x; a < bD; a += blockDim.x) lt[a] = sqrt(S[a] / nrm);
}

// ---------------------------------------------------------------------------
// Persistent ALS solver (ALSSolver.solve / solve_torch, als_solver.py:34-82,
// with build_s1/s2/r1/r2, solve_ar and calculate_cost, :148-260), one
// workgroup, the whole sweep loop and its stopping rule on the device.
//
// Index names follow the reference: n12[y][x][Y][X], a12g[y][x][p][q],
// a1r[y][u][p], a2r[x][v][q].  Per sweep:
//   R1[(Y,U),(y,u)] = sum_{x,X,Q} n12[y,x,Y,X] a2r[x,u,Q] a2r[X,U,Q]
//   S1[(Y,U),p]     = sum_{X,Q}   n12g[Y,X,p,Q] a2r[X,U,Q]
//   a1r = solve(sym(R1) + eps max|R1| I, S1)   (Cholesky; LU fallback)
// and the mirror image for a2r; then the cost
//   d2 - 2 d3,  d2 = <a12n|N|a12n>, d3 = <a12n|N|a12g>, a12n = a1r a2r.
// Stops after sweep i (0-based) when |d2 - d1| / |d1| < tol and i > 1.
// ---------------------------------------------------------------------------
constexpr int ALS_NT = 256;

struct AlsShared {
  double* R;     // nR x nR
  double* S;     // nR x d
  double* a1;    // nD x D x d
  double* a2;
  double* T;     // nD x nD x D x d (one outer index of the R build)
  double* n12g;  // nD^2 d^2
  double* a12n;  // nD^2 d^2
  double* red;   // ALS_NT/64
};

// N(a, b, A, B): side 0 -> n12[a][b][A][B], side 1 -> n12[b][a][B][A]
__device__ __forceinline__ int64_t nidx(int side, int nD, int a, int b, int A, int B) {
  return side == 0 ? (((int64_t)a * nD + b) * nD + A) * nD + B
                   : (((int64_t)b * nD + a) * nD + B) * nD + A;
}

// Build R and S for the tensor of side ``side`` from ``other`` (the fixed one).
__device__ void als_build(const AlsShared& sh, const double* __restrict__ n12, int side, int nD,
                          int D, int d, const double* other) {
  const int tid = threadIdx.x;
  const int nR = nD * D, Dd = D * d;
  for (int a = 0; a < nD; ++a) {
    // T[A][B][u][Q] = sum_b N(a,b,A,B) other[b][u][Q]
    for (int e = tid; e < nD * nD * Dd; e += ALS_NT) {
      const int A = e / (nD * Dd), r = e % (nD * Dd), B = r / Dd, uQ = r % Dd;
      double acc = 0.0;
      for (int b = 0; b < nD; ++b) acc += n12[nidx(side, nD, a, b, A, B)] * other[b * Dd + uQ];
      sh.T[e] = acc;
    }
    __syncthreads();
    // R[(A,U),(a,u)] = sum_{B,Q} T[A][B][u][Q] other[B][U][Q]
    for (int e = tid; e < nD * D * D; e += ALS_NT) {
      const int A = e / (D * D), U = (e / D) % D, u = e % D;
      double acc = 0.0;
      for (int B = 0; B < nD; ++B)
        for (int Q = 0; Q < d; ++Q)
          acc += sh.T[((A * nD + B) * D + u) * d + Q] * other[(B * D + U) * d + Q];
      sh.R[(int64_t)(A * D + U) * nR + a * D + u] = acc;
    }
    __syncthreads();
  }
  // S[(A,U)][p] = sum_{B,Q} G(A,B,p,Q) other[B][U][Q]
  const int d2 = d * d;
  for (int e = tid; e < nR * d; e += ALS_NT) {
    const int A = e / (D * d), U = (e / d) % D, p = e % d;
    double acc = 0.0;
    for (int B = 0; B < nD; ++B)
      for (int Q = 0; Q < d; ++Q) {
        const double g = side == 0 ? sh.n12g[(A * nD + B) * d2 + p * d + Q]
                                   : sh.n12g[(B * nD + A) * d2 + Q * d + p];
        acc += g * other[(B * D + U) * d + Q];
      }
    sh.S[e] = acc;
  }
  __syncthreads();
}

__device__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = fmax(t, red[w]);
  return t;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Solve (sym(R) + eps max|R| I) X = S in place (X -> S), wave 0 doing the
// factorisation.  Returns (block-uniform) 1 if the Cholesky factorisation
// failed and partial-pivoting elimination was used, 2 if that found R singular.
// ALS system through the pseudo-inverse (als_solver.py:228-230:
// pinv(R, hermitian=True, rcond=eps) @ S): R <- (R + R^T)/2 (no shift),
// R = V diag(lam) V^T by cyclic two-sided Jacobi with round-robin parallel
// pairs (R in LDS, V in the global scratch), then S <- V diag(w) V^T S with
// w_i = 1/lam_i for |lam_i| > eps * max|lam|, else 0.
__device__ void als_pinv_solve(const AlsShared& sh, int nR, int d, double eps,
                               double* __restrict__ V) {
  const int tid = threadIdx.x;
  double* R = sh.R;
  double* S = sh.S;
  __shared__ double rc[128], rs[128];
  __shared__ int rp[128], rq[128];
  __shared__ int s_rot;
  for (int e = tid; e < nR * nR; e += ALS_NT) {
    const int i = e / nR, j = e % nR;
    if (i < j) {
      const double v = 0.5 * (R[i * nR + j] + R[j * nR + i]);
      R[i * nR + j] = v;
      R[j * nR + i] = v;
    }
    V[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int n = nR + (nR & 1);  // padding index n - 1 (odd nR) never rotates
  const int half = n / 2;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (tid == 0) s_rot = 0;
    __syncthreads();
    for (int round = 0; round < n - 1; ++round) {
      // circle method: index n - 1 fixed, the others rotate
      for (int a = tid; a < half; a += ALS_NT) {
        int p = (a == 0) ? n - 1 : (round + a) % (n - 1);
        int q = (round - a + (n - 1)) % (n - 1);
        if (a == 0) q = round;
        if (p > q) { const int t = p; p = q; q = t; }
        double c = 1.0, sn = 0.0;
        if (q < nR) {
          const double app = R[p * nR + p], aqq = R[q * nR + q], apq = R[p * nR + q];
          if (fabs(apq) > 1e-15 * sqrt(fabs(app * aqq)) && apq != 0.0) {
            const double zeta = (aqq - app) / (2.0 * apq);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = c * t;
            s_rot = 1;
          }
        }
        rc[a] = c; rs[a] = sn; rp[a] = p; rq[a] = q;
      }
      __syncthreads();
      // rows: R[p,:], R[q,:] <- c R[p,:] - s R[q,:], s R[p,:] + c R[q,:]
      for (int e = tid; e < half * nR; e += ALS_NT) {
        const int a = e / nR, k = e % nR;
        const int p = rp[a], q = rq[a];
        if (q >= nR || rs[a] == 0.0) continue;
        const double c = rc[a], sn = rs[a];
        const double x = R[p * nR + k], y = R[q * nR + k];
        R[p * nR + k] = c * x - sn * y;
        R[q * nR + k] = sn * x + c * y;
      }
      __syncthreads();
      // columns of R and V
      for (int e = tid; e < half * nR; e += ALS_NT) {
        const int a = e / nR, k = e % nR;
        const int p = rp[a], q = rq[a];
        if (q >= nR || rs[a] == 0.0) continue;
        const double c = rc[a], sn = rs[a];
        const double x = R[k * nR + p], y = R[k * nR + q];
        R[k * nR + p] = c * x - sn * y;
        R[k * nR + q] = sn * x + c * y;
        const double vx = V[k * nR + p], vy = V[k * nR + q];
        V[k * nR + p] = c * vx - sn * vy;
        V[k * nR + q] = sn * vx + c * vy;
      }
      __syncthreads();
    }
    if (!s_rot) break;
    __syncthreads();
  }
  // S <- V diag(w) V^T S; T = V^T S staged in R's off-diagonal-free copy
  double mx = 0.0;
  for (int i = tid; i < nR; i += ALS_NT) mx = fmax(mx, fabs(R[i * nR + i]));
  mx = block_max(mx, sh.red);
  double* T = R;  // the eigenvalues are read once below, before T overwrites row 0
  __shared__ double lam[256];
  for (int i = tid; i < nR; i += ALS_NT) lam[i] = R[i * nR + i];
  __syncthreads();
  for (int e = tid; e < nR * d; e += ALS_NT) {
    const int i = e / d, c = e % d;
    double acc = 0.0;
    for (int k = 0; k < nR; ++k) acc += V[k * nR + i] * S[k * d + c];
    const double li = lam[i];
    T[e] = (fabs(li) > eps * mx) ? acc / li : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < nR * d; e += ALS_NT) {
    const int i = e / d, c = e % d;
    double acc = 0.0;
    for (int k = 0; k < nR; ++k) acc += V[i * nR + k] * T[k * d + c];
    S[e] = acc;
  }
  __syncthreads();
}

__device__ int als_solve_system(const AlsShared& sh, int nR, int d, double eps,
                                double* __restrict__ Rsave, int* flag) {
  const int tid = threadIdx.x, lane = tid & 63;
  double* R = sh.R;
  double* S = sh.S;
  // symmetrise and regularise
  for (int e = tid; e < nR * nR; e += ALS_NT) {
    const int i = e / nR, j = e % nR;
    if (i < j) {
      const double v = 0.5 * (R[i * nR + j] + R[j * nR + i]);
      R[i * nR + j] = v;
      R[j * nR + i] = v;
    }
  }
  __syncthreads();
  double mx = 0.0;
  for (int e = tid; e < nR * nR; e += ALS_NT) mx = fmax(mx, fabs(R[e]));
  mx = block_max(mx, sh.red);
  for (int i = tid; i < nR; i += ALS_NT) R[i * nR + i] += eps * mx;
  __syncthreads();
  for (int e = tid; e < nR * nR; e += ALS_NT) Rsave[e] = R[e];
  for (int e = tid; e < nR * d; e += ALS_NT) Rsave[nR * nR + e] = S[e];
  if (tid == 0) *flag = 0;
  __syncthreads();
  if (tid < 64) {
    // Cholesky R = L L^T (lower, in place)
    int fail = 0;
    for (int j = 0; j < nR && !fail; ++j) {
      const double piv = R[j * nR + j];
      if (!(piv > 0.0)) { fail = 1; break; }
      const double ljj = sqrt(piv);
      wave_sync();
      if (lane == 0) R[j * nR + j] = ljj;
      for (int i = j + 1 + lane; i < nR; i += 64) R[i * nR + j] /= ljj;
      wave_sync();
      for (int i = j + 1 + lane; i < nR; i += 64) {
        const double lij = R[i * nR + j];
        for (int k = j + 1; k <= i; ++k) R[i * nR + k] -= lij * R[k * nR + j];
      }
      wave_sync();
    }
    if (!fail) {
      // forward: L y = S ; backward: L^T x = y   (all right-hand sides)
      for (int k = 0; k < nR; ++k) {
        const double lkk = R[k * nR + k];
        wave_sync();
        if (lane < d) S[k * d + lane] /= lkk;
        wave_sync();
        for (int i = k + 1 + lane; i < nR; i += 64) {
          const double lik = R[i * nR + k];
          for (int c = 0; c < d; ++c) S[i * d + c] -= lik * S[k * d + c];
        }
        wave_sync();
      }
      for (int k = nR - 1; k >= 0; --k) {
        const double lkk = R[k * nR + k];
        wave_sync();
        if (lane < d) S[k * d + lane] /= lkk;
        wave_sync();
        for (int i = lane; i < k; i += 64) {
          const double lki = R[k * nR + i];
          for (int c = 0; c < d; ++c) S[i * d + c] -= lki * S[k * d + c];
        }
        wave_sync();
      }
    } else if (lane == 0) {
      *flag = 1;
    }
  }
  __syncthreads();
  int status = *flag;
  if (status == 0) return 0;
  // LU with partial pivoting on the saved (regularised) system
  for (int e = tid; e < nR * nR; e += ALS_NT) R[e] = Rsave[e];
  for (int e = tid; e < nR * d; e += ALS_NT) S[e] = Rsave[nR * nR + e];
  __syncthreads();
  if (tid < 64) {
    int singular = 0;
    for (int j = 0; j < nR; ++j) {
      // pivot search over rows j..nR-1
      double best = -1.0;
      int bi = j;
      for (int i = j + lane; i < nR; i += 64) {
        const double v = fabs(R[i * nR + j]);
        if (v > best) { best = v; bi = i; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (!(best > 0.0)) { singular = 1; break; }
      wave_sync();
      if (bi != j) {
        for (int c = lane; c < nR; c += 64) {
          const double t = R[j * nR + c];
          R[j * nR + c] = R[bi * nR + c];
          R[bi * nR + c] = t;
        }
        if (lane < d) {
          const double t = S[j * d + lane];
          S[j * d + lane] = S[bi * d + lane];
          S[bi * d + lane] = t;
        }
      }
      wave_sync();
      const double pjj = R[j * nR + j];
      for (int i = j + 1 + lane; i < nR; i += 64) {
        const double f = R[i * nR + j] / pjj;
        for (int c = j + 1; c < nR; ++c) R[i * nR + c] -= f * R[j * nR + c];
        for (int c = 0; c < d; ++c) S[i * d + c] -= f * S[j * d + c];
      }
      wave_sync();
    }
    if (!singular) {
      for (int k = nR - 1; k >= 0; --k) {
        const double ukk = R[k * nR + k];
        wave_sync();
        if (lane < d) S[k * d + lane] /= ukk;
        wave_sync();
        for (int i = lane; i < k; i += 64) {
          const double uik = R[i * nR + k];
          for (int c = 0; c < d; ++c) S[i * d + c] -= uik * S[k * d + c];
        }
        wave_sync();
      }
    } else if (lane == 0) {
      *flag = 2;
    }
  }
  __syncthreads();
  return *flag;
}

// cost = d2 - 2 d3 of the current (a1, a2)
__device__ double als_cost(const AlsShared& sh, const double* __restrict__ n12, int nD, int D,
                           int d) {
  const int tid = threadIdx.x;
  const int d2 = d * d, nE = nD * nD * d2;
  for (int e = tid; e < nE; e += ALS_NT) {
    const int y = e / (nD * d2), x = (e / d2) % nD, p = (e / d) % d, q = e % d;
    double acc = 0.0;
    for (int u = 0; u < D; ++u) acc += sh.a1[(y * D + u) * d + p] * sh.a2[(x * D + u) * d + q];
    sh.a12n[e] = acc;
  }
  __syncthreads();
  double part = 0.0;
  for (int e = tid; e < nE; e += ALS_NT) {
    const int Y = e / (nD * d2), X = (e / d2) % nD, pq = e % d2;
    double na = 0.0;
    for (int y = 0; y < nD; ++y)
      for (int x = 0; x < nD; ++x)
        na += n12[(((int64_t)y * nD + x) * nD + Y) * nD + X] * sh.a12n[(y * nD + x) * d2 + pq];
    part += na * sh.a12n[e] - 2.0 * sh.n12g[e] * sh.a12n[e];
  }
  return block_sum(part, sh.red);
}

__global__ __launch_bounds__(ALS_NT) void als_kernel(const double* __restrict__ n12,
                                                     const double* __restrict__ a12g,
                                                     double* __restrict__ a1r,
                                                     double* __restrict__ a2r, int nD, int D,
                                                     int d, int niter, double tol, double eps,
                                                     int method, double* __restrict__ gwork,
                                                     double* __restrict__ out_cost,
                                                     int* __restrict__ out_info,
                                                     int64_t n12_stride, int64_t gwork_stride) {
  // workgroup b solves problem b of a batch: n12 n12_stride doubles apart,
  // a12g nE apart, a1r/a2r nR*d apart, scratch gwork_stride apart, one cost
  // and two info words each
  {
    const int64_t b = blockIdx.x;
    const int64_t nE_ = (int64_t)nD * nD * d * d, nRd = (int64_t)nD * D * d;
    n12 += b * n12_stride;
    a12g += b * nE_;
    a1r += b * nRd;
    a2r += b * nRd;
    gwork += b * gwork_stride;
    out_cost += b;
    out_info += 2 * b;
  }
  extern __shared__ double lds[];
  __shared__ int flag;
  __shared__ double s_d1, s_d2;
  __shared__ int s_stop;
  const int tid = threadIdx.x;
  const int nR = nD * D, d2 = d * d, nE = nD * nD * d2;
  AlsShared sh;
  sh.R = lds;
  sh.S = sh.R + (size_t)nR * nR;
  sh.a1 = sh.S + (size_t)nR * d;
  sh.a2 = sh.a1 + (size_t)nR * d;
  sh.T = sh.a2 + (size_t)nR * d;
  sh.n12g = sh.T + (size_t)nD * nD * D * d;
  sh.a12n = sh.n12g + nE;
  sh.red = sh.a12n + nE;
  for (int e = tid; e < nR * d; e += ALS_NT) { sh.a1[e] = a1r[e]; sh.a2[e] = a2r[e]; }
  // n12g[Y][X][p][q] = sum_{y,x} n12[y][x][Y][X] a12g[y][x][p][q]
  for (int e = tid; e < nE; e += ALS_NT) {
    const int Y = e / (nD * d2), X = (e / d2) % nD, pq = e % d2;
    double acc = 0.0;
    for (int y = 0; y < nD; ++y)
      for (int x = 0; x < nD; ++x)
        acc += n12[(((int64_t)y * nD + x) * nD + Y) * nD + X] * a12g[(y * nD + x) * d2 + pq];
    sh.n12g[e] = acc;
  }
  __syncthreads();
  const double c0 = als_cost(sh, n12, nD, D, d);
  if (tid == 0) { s_d1 = fabs(c0); s_d2 = c0; s_stop = 0; }
  __syncthreads();
  int sweeps = 0, info = 0;
  for (int i = 0; i < niter; ++i) {
    als_build(sh, n12, 0, nD, D, d, sh.a2);
    int st = 0;
    if (method == 1) als_pinv_solve(sh, nR, d, eps, gwork);
    else st = als_solve_system(sh, nR, d, eps, gwork, &flag);
    info |= st;
    for (int e = tid; e < nR * d; e += ALS_NT) sh.a1[e] = sh.S[e];
    __syncthreads();
    als_build(sh, n12, 1, nD, D, d, sh.a1);
    if (method == 1) als_pinv_solve(sh, nR, d, eps, gwork);
    else st = als_solve_system(sh, nR, d, eps, gwork, &flag);
    info |= st;
    for (int e = tid; e < nR * d; e += ALS_NT) sh.a2[e] = sh.S[e];
    __syncthreads();
    sweeps = i + 1;
    const double c = als_cost(sh, n12, nD, D, d);
    if (tid == 0) {
      s_d2 = c;
      const double err = fabs(c - s_d1) / fabs(s_d1);
      s_stop = (err < tol && i > 1) ? 1 : 0;
      if (!s_stop) s_d1 = c;
    }
    __syncthreads();
    if (s_stop) break;
  }
  for (int e = tid; e < nR * d; e += ALS_NT) { a1r[e] = sh.a1[e]; a2r[e] = sh.a2[e]; }
  if (tid == 0) {
    out_cost[0] = s_d2;
    out_info[0] = sweeps;
    out_info[1] = info;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
void permute_copy(const double* in, double* out, const PermDesc& p, double scale, hipStream_t s) {
  if (p.total <= 0) return;
  const int blocks = (int)std::min<int64_t>((p.total + 255) / 256, 2048);
  hipLaunchKernelGGL(permute_kernel, dim3(blocks), dim3(256), 0, s, in, out, p, scale);
  ACETN_HIP_CHECK(hipGetLastError());
}

void householder_qr(const double* A, int m, int n, double* W, double* Q, double* R,
                    hipStream_t s) {
  ACETN_REQUIRE(m >= 1 && n >= 1 && n <= kQrMaxN, "householder_qr: unsupported shape");
  hipLaunchKernelGGL(householder_qr_kernel, dim3(1), dim3(QR_NT), 0, s, A, m, n, W, Q, R);
  ACETN_HIP_CHECK(hipGetLastError());
}

void svd_small(const double* A, int m, int n, double* U, double* sv, double* V, hipStream_t s) {
  ACETN_REQUIRE(m >= 1 && m <= 64 && n >= 1 && n <= kSvdMaxN, "svd_small: unsupported shape");
  hipLaunchKernelGGL(svd_small_kernel, dim3(1), dim3(SVD_NT), 0, s, A, m, n, U, sv, V,
                     (int64_t)0);
  ACETN_HIP_CHECK(hipGetLastError());
}

void svd_small_batch(const double* A, int m, int n, double* U, double* sv, double* V, int count,
                     int64_t stride, hipStream_t s) {
  ACETN_REQUIRE(m >= 1 && m <= 64 && n >= 1 && n <= kSvdMaxN, "svd_small: unsupported shape");
  ACETN_REQUIRE(count >= 1 && stride >= (int64_t)m * n, "svd_small_batch: bad batch layout");
  hipLaunchKernelGGL(svd_small_kernel, dim3(count), dim3(SVD_NT), 0, s, A, m, n, U, sv, V,
                     stride);
  ACETN_HIP_CHECK(hipGetLastError());
}

void lower_to_full(double* M, int n, hipStream_t s) {
  const int blocks = std::min((n * n + 255) / 256, 1024);
  hipLaunchKernelGGL(lower_to_full_kernel, dim3(blocks), dim3(256), 0, s, M, n);
  ACETN_HIP_CHECK(hipGetLastError());
}

void add_diag(double* M, int n, double shift, hipStream_t s) {
  hipLaunchKernelGGL(add_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, M, n, shift);
  ACETN_HIP_CHECK(hipGetLastError());
}

void scale_sqrt_columns(const double* Z, const double* w, int rows, int cols, double* X,
                        hipStream_t s) {
  const int blocks = std::min((rows * cols + 255) / 256, 1024);
  hipLaunchKernelGGL(scale_sqrt_kernel, dim3(blocks), dim3(256), 0, s, Z, w, rows, cols, X);
  ACETN_HIP_CHECK(hipGetLastError());
}

void pinv_finish(const double* U, const double* sv, const double* V, int m, int n, double atol,
                 double* P, hipStream_t s) {
  hipLaunchKernelGGL(pinv_finish_kernel, dim3((n * m + 255) / 256), dim3(256), 0, s, U, sv, V, m,
                     n, atol, P);
  ACETN_HIP_CHECK(hipGetLastError());
}

void finalize_scale(const double* U, const double* sv, const double* V, int n, int bD,
                    double* r1n, double* r2n, hipStream_t s) {
  hipLaunchKernelGGL(finalize_scale_kernel, dim3(1), dim3(256), 0, s, U, sv, V, n, bD, r1n, r2n);
  ACETN_HIP_CHECK(hipGetLastError());
}

void bond_gauge(double* a1r, double* a2r, int nD1, int nD2, int bD, int pD, hipStream_t s) {
  hipLaunchKernelGGL(bond_gauge_kernel, dim3(bD), dim3(256), 0, s, a1r, a2r, nD1, nD2, bD, pD);
  ACETN_HIP_CHECK(hipGetLastError());
}

void lambda_scale(const double* A, double* out, const LambdaLegs& w, int D, int d, int invert,
                  double cutoff, hipStream_t s) {
  const int64_t n = (int64_t)D * D * D * D * d;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(lambda_scale_kernel, dim3(blocks), dim3(256), 0, s, A, out, w, D, d, invert,
                     cutoff);
  ACETN_HIP_CHECK(hipGetLastError());
}

void su_split(const double* U, const double* S, const double* V, int nD, int bD, int pD,
              double* a1r, double* a2r, double* lt, hipStream_t s) {
  hipLaunchKernelGGL(su_split_kernel, dim3(1), dim3(256), 0, s, U, S, V, nD, bD, pD, a1r, a2r, lt);
  ACETN_HIP_CHECK(hipGetLastError());
}

size_t als_lds_bytes(int nD, int D, int d) {
  const size_t nR = (size_t)nD * D;
  const size_t nE = (size_t)nD * nD * d * d;
  return sizeof(double) * (nR * nR + 3 * nR * d + (size_t)nD * nD * D * d + 2 * nE + ALS_NT / 64);
}

size_t als_work_elems(int nD, int D, int d) {
  const size_t nR = (size_t)nD * D;
  return nR * nR + nR * d;
}

void als_check_shape(int nD, int D, int d, int method) {
  ACETN_REQUIRE(method == 0 || method == 1, "als: unknown method");
  ACETN_REQUIRE(method == 0 || nD * D <= 256, "als: pinv needs nD * D <= 256");
  ACETN_REQUIRE(als_lds_bytes(nD, D, d) <= 150 * 1024,
                "als: reduced tensors too large for one workgroup");
}

void als_solve_device(const double* n12, const double* a12g, double* a1r, double* a2r, int nD,
                      int D, int d, int niter, double tol, double eps, double* gwork,
                      double* out_cost, int* out_info, hipStream_t s, int method, int count) {
  ACETN_REQUIRE(count >= 1, "als: empty batch");
  als_check_shape(nD, D, d, method);
  const size_t lds = als_lds_bytes(nD, D, d);
  static bool attr = false;
  if (!attr) {
    ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)als_kernel,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
    attr = true;
  }
  const int64_t n12_stride = (int64_t)nD * nD * nD * nD;
  hipLaunchKernelGGL(als_kernel, dim3(count), dim3(ALS_NT), lds, s, n12, a12g, a1r, a2r, nD, D, d,
                     niter, tol, eps, method, gwork, out_cost, out_info, n12_stride,
                     (int64_t)als_work_elems(nD, D, d));
  ACETN_HIP_CHECK(hipGetLastError());
}

}  // namespace acetn
