// Small dense FP64 linear algebra for projector construction, one workgroup
// per problem:
//   * Gaussian sketch (Philox4x32-10 + Box-Muller) for the range finder,
//   * the triangular factor of a (shifted) Cholesky QR, with rank guarding,
//   * a cyclic parallel-ordered Jacobi symmetric eigensolver whose rotations
//     are recorded and replayed row-parallel onto the eigenvector matrix,
//   * projector finishing: cutoff count, singular-value column scaling, the
//     sign gauge and Frobenius normalisation.
//
// The q x q working matrices are kept as packed upper triangles (q(q+1)/2
// doubles), in LDS when they fit and in a global scratch area otherwise.
#include "internal.h"

#include <cmath>

namespace acetn {
namespace {

constexpr int LA_NT = 256;     // Cholesky: latency-bound steps, cheap barriers
constexpr int JAC_NT = 1024;   // Jacobi: LDS-latency-bound, needs many waves
constexpr size_t kLdsBudget = 150 * 1024;

__device__ __forceinline__ int pk(int i, int j) {  // packed upper, i <= j required
  return (j * (j + 1)) / 2 + i;
}
__device__ __forceinline__ int pks(int i, int j) { return i <= j ? pk(i, j) : pk(j, i); }

// --------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) + Box-Muller.
// --------------------------------------------------------------------------
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

__global__ void sketch_kernel(double* out, int64_t total, uint64_t seed, uint64_t stream0,
                              const uint64_t* __restrict__ stream_base) {
  // the stream number: a constant, plus a device counter when one is given
  // (read when the kernel runs, so a captured launch draws a new sketch on
  // every replay once the counter has moved)
  const uint64_t stream = stream0 + (stream_base ? stream_base[0] : 0);
  const int64_t pairs = (total + 1) / 2;
  for (int64_t call = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; call < pairs;
       call += (int64_t)gridDim.x * blockDim.x) {
    uint32_t c[4] = {(uint32_t)call, (uint32_t)((uint64_t)call >> 32), (uint32_t)stream,
                     (uint32_t)(stream >> 32)};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = (uint64_t)c[0] | ((uint64_t)c[1] << 32);
    const uint64_t b = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
    const double ua = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double ub = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(ua));
    const double th = (2.0 * M_PI) * ub;
    const int64_t e0 = 2 * call;
    out[e0] = r * cos(th);
    if (e0 + 1 < total) out[e0 + 1] = r * sin(th);
  }
}

// --------------------------------------------------------------------------
// Cholesky factor inverse: G + shift*I = R^T R (R upper), Rinv = R^-1.
// --------------------------------------------------------------------------
// kLds: the packed factor lives in LDS (else in ``gscratch``); a template
// parameter so that every access compiles to LDS (not generic) instructions.
template <bool kLds>
__global__ __launch_bounds__(LA_NT) void chol_rinv_kernel(const double* __restrict__ G,
                                                          double* __restrict__ Rinv, int q,
                                                          double shift_rel, double rank_tol,
                                                          double* gscratch, int use_lds,
                                                          int* dead_out, const int* only_if_dead) {
  extern __shared__ double lds[];
  double* S = kLds ? lds : gscratch;
  if (only_if_dead) {
    // completion pass: nothing to do unless an earlier pass refilled a column
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < q; i += LA_NT)
      if (only_if_dead[i]) any = 1;
    __syncthreads();
    if (!any) {
      for (int e = threadIdx.x; e < q * q; e += LA_NT) Rinv[e] = (e / q == e % q) ? 1.0 : 0.0;
      return;
    }
  }
  __shared__ double s_red[LA_NT / 64];
  __shared__ double s_trace, s_dmax;
  __shared__ int dead[512];
  __shared__ double tmp[512];
  const int tid = threadIdx.x;

  // trace and max diagonal
  double tr = 0.0, dm = 0.0;
  for (int i = tid; i < q; i += LA_NT) { tr += G[(int64_t)i * q + i]; dm = fmax(dm, G[(int64_t)i * q + i]); }
  for (int o = 32; o > 0; o >>= 1) { tr += __shfl_down(tr, o); dm = fmax(dm, __shfl_down(dm, o)); }
  if ((tid & 63) == 0) s_red[tid >> 6] = tr;
  __syncthreads();
  if (tid == 0) { double t = 0; for (int w = 0; w < LA_NT / 64; ++w) t += s_red[w]; s_trace = t; }
  __syncthreads();
  if ((tid & 63) == 0) s_red[tid >> 6] = dm;
  __syncthreads();
  if (tid == 0) { double t = 0; for (int w = 0; w < LA_NT / 64; ++w) t = fmax(t, s_red[w]); s_dmax = t; }
  __syncthreads();
  const double shift = shift_rel * s_trace;
  const double thresh = rank_tol * (s_dmax + shift);

  // packed upper triangle of G + shift*I
  for (int e = tid; e < q * q; e += LA_NT) {
    const int i = e / q, j = e % q;
    if (i <= j) S[pk(i, j)] = G[(int64_t)i * q + j] + (i == j ? shift : 0.0);
  }
  __syncthreads();

  // right-looking Cholesky, G = R^T R (R upper, stored in place)
  for (int j = 0; j < q; ++j) {
    const double piv = S[pk(j, j)];
    const bool is_dead = !(piv > thresh) || !isfinite(piv);
    const double rjj = is_dead ? 0.0 : sqrt(piv);
    __syncthreads();
    if (tid == 0) { S[pk(j, j)] = rjj; dead[j] = is_dead ? 1 : 0; }
    for (int k = j + 1 + tid; k < q; k += LA_NT) S[pk(j, k)] = is_dead ? 0.0 : S[pk(j, k)] / rjj;
    __syncthreads();
    if (!is_dead) {
      // rows ii by thread groups of 32, columns kk >= ii strided within the group
      const int m = q - j - 1;
      for (int ii = tid >> 5; ii < m; ii += LA_NT / 32) {
        const int i = j + 1 + ii;
        const double rji = S[pk(j, i)];
        for (int kk = ii + (tid & 31); kk < m; kk += 32) {
          const int k = j + 1 + kk;
          S[pk(i, k)] -= rji * S[pk(j, k)];
        }
      }
    }
    __syncthreads();
  }
  if (dead_out) {
    int anyd = 0;
    for (int i = tid; i < q; i += LA_NT) {
      dead_out[i] = dead[i];
      anyd |= dead[i];
    }
    anyd = __syncthreads_or(anyd);
    if (tid == 0) dead_out[q] = anyd;
  }

  // in-place inverse of the packed upper factor, column by column:
  // Rinv[0:j, j] = -Rinv[0:j, 0:j] R[0:j, j] / R[j, j].  Dead columns/rows are 0.
  for (int j = 0; j < q; ++j) {
    const bool dj = dead[j] != 0;
    for (int i = tid; i < j; i += LA_NT) tmp[i] = S[pk(i, j)];
    __syncthreads();
    if (dj) {
      for (int i = tid; i <= j; i += LA_NT) S[pk(i, j)] = 0.0;
    } else {
      const double inv = 1.0 / S[pk(j, j)];
      // 8 lanes per row i, each summing a strided part of k in [i, j)
      const int g = tid & 7;
      for (int i0 = 0; i0 < j; i0 += LA_NT / 8) {
        const int i = i0 + (tid >> 3);
        double y = 0.0;
        if (i < j && !dead[i])
          for (int k = i + g; k < j; k += 8) y += S[pk(i, k)] * tmp[k];
        y += __shfl_xor(y, 1);
        y += __shfl_xor(y, 2);
        y += __shfl_xor(y, 4);
        if (i < j && g == 0) S[pk(i, j)] = -y * inv;
      }
      if (tid == 0) S[pk(j, j)] = inv;
    }
    __syncthreads();
  }
  for (int e = tid; e < q * q; e += LA_NT) {
    const int i = e / q, j = e % q;
    Rinv[e] = (i <= j) ? S[pk(i, j)] : 0.0;
  }
}

// Blocked variant on the FP64 matrix cores for q <= kCbMaxQ: the matrix is
// held as its upper 16 x 16 tiles in LDS (rows padded to 17 doubles), and
//   for each tile column k:  one wave factors the diagonal tile (element
//       Cholesky with the same rank guard) and replaces it by its inverse
//       X_kk = R_kk^-1; the panel becomes R_kj = X_kk^T A_kj and the trailing
//       tiles A_IJ -= R_kI^T R_kJ (16 x 16 x 16 products, v_mfma_f64_16x16x4);
//   inverse, tile rows i from the bottom:  X_ij = -X_ii sum_{l=i+1..j} R_il X_lj.
// Three barriers per tile column and two per inverse row replace the one per
// matrix row of the element algorithms.  Dead pivots give zero rows of R and
// zero rows and columns of X (the dead rows of X_kk vanish, so they vanish in
// every product), as in chol_rinv_kernel.
typedef double cb_d4 __attribute__((ext_vector_type(4)));
constexpr int CB_WAVES = 8;
constexpr int CB_NT = 64 * CB_WAVES;
constexpr int CB_LD = 17;             // tile row stride (doubles)
constexpr int CB_TS = 16 * CB_LD;     // tile size (doubles)
constexpr int kCbMaxQ = 176;          // 11 tile rows: 66 tiles, 144 KB
constexpr int kCbMaxT = kCbMaxQ / 16;

__device__ __forceinline__ int cb_tile(int I, int J) { return (J * (J + 1) / 2 + I) * CB_TS; }

__device__ __forceinline__ void cb_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc += s * op(A) B over one 16-deep tile product, op(A) = A^T (TA) or A.
// Operand lanes: a = op(A)[lane & 15][4 ks + (lane >> 4)], b = B[4 ks + (lane >> 4)][lane & 15];
// acc[r] = C[(lane >> 4) + 4 r][lane & 15].
template <bool TA>
__device__ __forceinline__ void cb_mma(cb_d4& acc, const double* A, const double* B, int lane,
                                       double sgn) {
  const int m = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const int k = 4 * ks + kq;
    const double a = TA ? A[k * CB_LD + m] : A[m * CB_LD + k];
    const double b = B[k * CB_LD + m];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sgn * a, b, acc, 0, 0, 0);
  }
}
__device__ __forceinline__ cb_d4 cb_load_acc(const double* T, int lane) {
  cb_d4 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = T[((lane >> 4) + 4 * r) * CB_LD + (lane & 15)];
  return v;
}
__device__ __forceinline__ void cb_store_acc(double* T, int lane, const cb_d4& v) {
#pragma unroll
  for (int r = 0; r < 4; ++r) T[((lane >> 4) + 4 * r) * CB_LD + (lane & 15)] = v[r];
}

// One wave: Cholesky of the diagonal tile D (upper factor, rank guard), then
// D <- R^-1 (upper, lower triangle zero).  Lane l owns column j = l & 15 and
// rows i = (l >> 4) + 4 t.  dead[c] flags the dead pivots of this tile.
__device__ __forceinline__ void cb_factor_diag(double* D, double* xbuf, int lane, double thresh,
                                               int* dead) {
  const int j = lane & 15, i0 = lane >> 4;
  for (int c = 0; c < 16; ++c) {
    const double d = D[c * CB_LD + c];
    const bool dd = !(d > thresh) || !isfinite(d);
    const double rcj = D[c * CB_LD + j];
    double rci[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) rci[t] = D[c * CB_LD + i0 + 4 * t];
    cb_wave_sync();
    const double invd = dd ? 0.0 : 1.0 / d;
    const double rcc = dd ? 0.0 : sqrt(d);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = i0 + 4 * t;
      if (i > c && j >= i) D[i * CB_LD + j] -= rci[t] * invd * rcj;
      else if (i == c && j >= c) D[c * CB_LD + j] = (j == c) ? rcc : (dd ? 0.0 : rcj / rcc);
    }
    if (lane == 0) dead[c] = dd ? 1 : 0;
    cb_wave_sync();
  }
  // X = R^-1 row by row from the bottom, X in registers, rows published in xbuf
  double x[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) x[t] = (i0 + 4 * t == j) ? 1.0 : 0.0;
  for (int r = 15; r >= 0; --r) {
    double* xr = xbuf + 16 * (r & 1);
    const double rr = D[r * CB_LD + r];
    const double inv = (dead[r] || rr == 0.0) ? 0.0 : 1.0 / rr;
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (i0 + 4 * t == r) {
        x[t] *= inv;
        xr[j] = x[t];
      }
    cb_wave_sync();
    const double xrj = xr[j];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = i0 + 4 * t;
      if (i < r) x[t] -= D[i * CB_LD + r] * xrj;
    }
  }
  cb_wave_sync();
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = i0 + 4 * t;
    D[i * CB_LD + j] = (j >= i) ? x[t] : 0.0;
  }
  cb_wave_sync();
}

__global__ __launch_bounds__(CB_NT) void chol_rinv_blk_kernel(const double* __restrict__ G,
                                                              double* __restrict__ Rinv, int q,
                                                              double shift_rel, double rank_tol,
                                                              int* dead_out,
                                                              const int* only_if_dead) {
  extern __shared__ double T[];  // upper tiles
  __shared__ int dead[kCbMaxQ];
  __shared__ double xbuf[32];
  __shared__ double s_red[CB_WAVES][2];
  __shared__ int any;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (only_if_dead) {
    if (tid == 0) any = 0;
    __syncthreads();
    for (int i = tid; i < q; i += CB_NT)
      if (only_if_dead[i]) any = 1;
    __syncthreads();
    if (!any) {
      for (int e = tid; e < q * q; e += CB_NT) Rinv[e] = (e / q == e % q) ? 1.0 : 0.0;
      return;
    }
  }
  double tr = 0.0, dm = 0.0;
  for (int i = tid; i < q; i += CB_NT) {
    const double gii = G[(int64_t)i * q + i];
    tr += gii;
    dm = fmax(dm, gii);
  }
  for (int o = 32; o > 0; o >>= 1) { tr += __shfl_down(tr, o); dm = fmax(dm, __shfl_down(dm, o)); }
  if (lane == 0) { s_red[wave][0] = tr; s_red[wave][1] = dm; }
  __syncthreads();
  if (tid == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < CB_WAVES; ++w) { t0 += s_red[w][0]; t1 = fmax(t1, s_red[w][1]); }
    s_red[0][0] = t0; s_red[0][1] = t1;
  }
  __syncthreads();
  const double shift = shift_rel * s_red[0][0];
  const double thresh = rank_tol * (s_red[0][1] + shift);
  const int Q = (q + 15) / 16;
  const int ntiles = Q * (Q + 1) / 2;

  // tiles of G + shift I (upper tiles; diagonal tiles symmetric; padding = identity)
  for (int t = wave; t < ntiles; t += CB_WAVES) {
    int J = 0;
    while ((J + 1) * (J + 2) / 2 <= t) ++J;
    const int I = t - J * (J + 1) / 2;
    for (int e = lane; e < 256; e += 64) {
      const int a = 16 * I + (e >> 4), b = 16 * J + (e & 15);
      double x;
      if (a < q && b < q) {
        const int lo = min(a, b), hi = max(a, b);
        x = G[(int64_t)lo * q + hi] + (a == b ? shift : 0.0);
      } else {
        x = (a == b) ? 1.0 : 0.0;
      }
      T[t * CB_TS + (e >> 4) * CB_LD + (e & 15)] = x;
    }
  }
  __syncthreads();

  // Look-ahead: wave 0 brings the next diagonal tile up to date first and
  // factors it while waves 1.. update the rest of the trailing matrix, so a
  // tile column costs one panel pass and one trailing pass (two barriers).
  if (wave == 0) cb_factor_diag(T + cb_tile(0, 0), xbuf, lane, thresh, dead);
  __syncthreads();
  for (int k = 0; k < Q; ++k) {
    for (int jj = k + 1 + wave; jj < Q; jj += CB_WAVES) {
      cb_d4 acc = {0.0, 0.0, 0.0, 0.0};
      cb_mma<true>(acc, T + cb_tile(k, k), T + cb_tile(k, jj), lane, 1.0);
      cb_wave_sync();
      cb_store_acc(T + cb_tile(k, jj), lane, acc);
    }
    __syncthreads();
    if (k + 1 == Q) break;
    const int nrem = Q - k - 1;
    const int ntr = nrem * (nrem + 1) / 2;
    if (wave == 0) {
      // trailing tile u = 0 is the next diagonal tile (k + 1, k + 1)
      cb_d4 acc = cb_load_acc(T + cb_tile(k + 1, k + 1), lane);
      cb_mma<true>(acc, T + cb_tile(k, k + 1), T + cb_tile(k, k + 1), lane, -1.0);
      cb_wave_sync();
      cb_store_acc(T + cb_tile(k + 1, k + 1), lane, acc);
      cb_wave_sync();
      cb_factor_diag(T + cb_tile(k + 1, k + 1), xbuf, lane, thresh, dead + 16 * (k + 1));
    } else {
      for (int u = wave; u < ntr; u += CB_WAVES - 1) {
        int Jp = 0;
        while ((Jp + 1) * (Jp + 2) / 2 <= u) ++Jp;
        const int I = k + 1 + (u - Jp * (Jp + 1) / 2), J = k + 1 + Jp;
        cb_d4 acc = cb_load_acc(T + cb_tile(I, J), lane);
        cb_mma<true>(acc, T + cb_tile(k, I), T + cb_tile(k, J), lane, -1.0);
        cb_store_acc(T + cb_tile(I, J), lane, acc);
      }
    }
    __syncthreads();
  }
  if (dead_out) {
    int anyd = 0;
    for (int i = tid; i < q; i += CB_NT) {
      dead_out[i] = dead[i];
      anyd |= dead[i];
    }
    anyd = __syncthreads_or(anyd);
    if (tid == 0) dead_out[q] = anyd;
  }

  // X = R^-1 by tile rows from the bottom: X_ij = -X_ii sum_{l=i+1..j} R_il X_lj
  constexpr int kJPer = (kCbMaxT + CB_WAVES - 1) / CB_WAVES;
  for (int i = Q - 2; i >= 0; --i) {
    cb_d4 res[kJPer];
#pragma unroll
    for (int m = 0; m < kJPer; ++m) {
      const int jj = i + 1 + wave + CB_WAVES * m;
      res[m] = cb_d4{0.0, 0.0, 0.0, 0.0};
      if (jj < Q) {
        cb_d4 S = {0.0, 0.0, 0.0, 0.0};
        for (int l = i + 1; l <= jj; ++l) cb_mma<false>(S, T + cb_tile(i, l), T + cb_tile(l, jj), lane, 1.0);
        // the accumulator layout of S is the B-operand layout (row 4 ks + (lane >> 4) in S[ks])
        const double* Xii = T + cb_tile(i, i);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const double a = Xii[(lane & 15) * CB_LD + 4 * ks + (lane >> 4)];
          res[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a, S[ks], res[m], 0, 0, 0);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kJPer; ++m) {
      const int jj = i + 1 + wave + CB_WAVES * m;
      if (jj < Q) cb_store_acc(T + cb_tile(i, jj), lane, res[m]);
    }
    __syncthreads();
  }
  for (int e = tid; e < q * q; e += CB_NT) {
    const int a = e / q, b = e % q;
    Rinv[e] = (a <= b) ? T[cb_tile(a >> 4, b >> 4) + (a & 15) * CB_LD + (b & 15)] : 0.0;
  }
}

// Dead (numerically rank-deficient) columns of an orthonormalisation are
// replaced by Gaussian vectors, which the following Cholesky-QR passes
// orthogonalise against the live ones: like Householder QR, the result always
// holds q orthonormal columns, so a sketch as wide as the space completes it.
__global__ void refill_kernel(double* __restrict__ Y, int64_t m, int q,
                              const int* __restrict__ dead, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    for (int j = 0; j < q; ++j) {
      if (!dead[j]) continue;
      const uint64_t e = (uint64_t)i * q + j;
      uint32_t c[4] = {(uint32_t)e, (uint32_t)(e >> 32), 0x0C0DEu, 0x5EEDu};
      philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
      const uint64_t a = (uint64_t)c[0] | ((uint64_t)c[1] << 32);
      const uint64_t b = (uint64_t)c[2] | ((uint64_t)c[3] << 32);
      const double ua = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
      const double ub = ((double)(b >> 11) + 0.5) * 0x1.0p-53;
      Y[i * q + j] = sqrt(-2.0 * log(ua)) * cos((2.0 * M_PI) * ub);
    }
  }
}

// Zero the columns of Y (m x q) that ``dead`` does not flag; nothing at all
// when dead[q] (any column flagged) is zero.
__global__ void keep_dead_columns_kernel(double* __restrict__ Y, int64_t m, int q,
                                         const int* __restrict__ dead) {
  if (!dead[q]) return;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m * q;
       e += (int64_t)gridDim.x * blockDim.x)
    if (!dead[e % q]) Y[e] = 0.0;
}

// still[j] <- still[j] && dead[j] (all zero when dead[q] is zero: ``still``
// may then hold stale flags of a skipped factorisation); still[q] <- any.
__global__ void still_dead_kernel(int* __restrict__ still, const int* __restrict__ dead, int q) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const int on = dead[q];
  for (int j = threadIdx.x; j < q; j += blockDim.x) {
    const int v = (on && dead[j] && still[j]) ? 1 : 0;
    still[j] = v;
    if (v) any = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) still[q] = any;
}

// --------------------------------------------------------------------------
// Jacobi eigensolver.  Round-robin ("circle method") parallel ordering over
// n = q rounded up to even indices; the padding index, if any, never rotates.
// rot layout: [(sweep*(n-1) + round) * (n/2) + pair] * 2 -> (c, s);
// info[0] = number of sweeps executed.
// --------------------------------------------------------------------------
// r in [0, n-2], a in [0, n/2): (r + a) mod (n-1) and (r - a) mod (n-1)
// need one correction each
__device__ __forceinline__ void round_pair(int n, int r, int a, int& p, int& qq) {
  int x, y;
  if (a == 0) { x = n - 1; y = r; }
  else {
    x = r + a;
    if (x >= n - 1) x -= n - 1;
    y = r - a;
    if (y < 0) y += n - 1;
  }
  p = min(x, y); qq = max(x, y);
}

// Jacobi rotation (c, s) annihilating apq of [[app, apq], [apq, aqq]]:
// t = tan(theta) is the smaller root of t^2 + 2 tau t - 1 = 0,
// tau = (aqq - app) / (2 apq), written as t = 2 apq sgn(d) / (|d| + sqrt(d^2 + 4 apq^2))
// with d = aqq - app (sgn(0) = 1): one square root, one division and one
// reciprocal square root.  The tau form covers the case where the squares
// underflow.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double& c,
                                                double& s) {
  const double d = aqq - app;
  const double h = sqrt(d * d + 4.0 * apq * apq);
  double t;
  if (h > 0.0 && isfinite(h)) {
    t = (d >= 0.0 ? 2.0 * apq : -2.0 * apq) / (fabs(d) + h);
  } else {
    const double tau = d / (2.0 * apq);
    t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
  }
  c = rsqrt(1.0 + t * t);
  s = t * c;
}

constexpr int kMaxSweeps = 40;
// 2x2 blocks cached per thread when n/2 (n/2 + 1)/2 <= kJacList * JAC_NT
// (q <= 180); larger problems walk the triangular block order each round.
constexpr int kJacList = 4;

// R^T S R on the 2x2 block (pair a, pair b) of the packed matrix, with the
// round's rotations R = [[c, s], [-s, c]] of both pairs.
__device__ __forceinline__ void jacobi_block(double* S, const short* pp, const short* qq_,
                                             const double* cs_c, const double* cs_s, int a, int b) {
  const double ca = cs_c[a], sa = cs_s[a], cb = cs_c[b], sb = cs_s[b];
  if (a == b) {
    if (sa == 0.0) return;
    const int pa = pp[a], qa = qq_[a];
    const double app = S[pk(pa, pa)], aqq = S[pk(qa, qa)], apq = S[pk(pa, qa)];
    const double t = sa / ca;
    S[pk(pa, pa)] = app - t * apq;
    S[pk(qa, qa)] = aqq + t * apq;
    S[pk(pa, qa)] = 0.0;
    return;
  }
  if (sa == 0.0 && sb == 0.0) return;
  const int pa = pp[a], qa = qq_[a], pb = pp[b], qb = qq_[b];
  const int i11 = pks(pa, pb), i12 = pks(pa, qb), i21 = pks(qa, pb), i22 = pks(qa, qb);
  const double x11 = S[i11], x12 = S[i12], x21 = S[i21], x22 = S[i22];
  const double y11 = x11 * cb - x12 * sb, y12 = x11 * sb + x12 * cb;
  const double y21 = x21 * cb - x22 * sb, y22 = x21 * sb + x22 * cb;
  S[i11] = ca * y11 - sa * y21;
  S[i12] = ca * y12 - sa * y22;
  S[i21] = sa * y11 + ca * y21;
  S[i22] = sa * y12 + ca * y22;
}

// One workgroup.  Each sweep starts with a scan of every pair against the
// rotation threshold (a sweep that would rotate nothing is never run), then
// runs the n-1 rounds: the round's n/2 rotations are computed from the packed
// matrix, then every 2x2 block (a <= b) is transformed.
template <bool kList, bool kLds>
__global__ __launch_bounds__(JAC_NT) void jacobi_kernel(const double* __restrict__ Gin, int q,
                                                        double* __restrict__ w,
                                                        double* __restrict__ rot, int* info,
                                                        double* gscratch, int use_lds,
                                                        double tol) {
  extern __shared__ double lds[];
  double* S = kLds ? lds : gscratch;
  __shared__ double cs_c[256], cs_s[256];
  __shared__ short pp[256], qq_[256];
  const int tid = threadIdx.x;
  const int n = q + (q & 1);
  const int half = n / 2;
  const int np = n * (n + 1) / 2;

  for (int e = tid; e < np; e += JAC_NT) {
    int j = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while ((j + 1) * (j + 2) / 2 <= e) ++j;
    while (j * (j + 1) / 2 > e) --j;
    const int i = e - j * (j + 1) / 2;
    S[e] = (i < q && j < q) ? 0.5 * (Gin[(int64_t)i * q + j] + Gin[(int64_t)j * q + i]) : 0.0;
  }

  const int nblocks = half * (half + 1) / 2;
  // this thread's share of the 2x2 blocks (a <= b) in triangular order
  // blk = b(b+1)/2 + a: start at blk = tid, advance by JAC_NT per block
  int a_start, b_start;
  {
    int b = (int)((sqrt(8.0 * tid + 1.0) - 1.0) * 0.5);
    while ((b + 1) * (b + 2) / 2 <= tid) ++b;
    while (b * (b + 1) / 2 > tid) --b;
    b_start = b;
    a_start = tid - b * (b + 1) / 2;
  }
  int my_a[kJacList], my_b[kJacList];
  if (kList) {
    int a = a_start, b = b_start;
#pragma unroll
    for (int i = 0; i < kJacList; ++i) {
      if (i > 0) {
        a += JAC_NT;
        while (a > b) { a -= b + 1; ++b; }
      }
      my_a[i] = a;
      my_b[i] = b;
    }
  }
  __syncthreads();

  int sweep = 0;
  for (; sweep < kMaxSweeps; ++sweep) {
    // pairs (i < j < q) split over 8 lanes per column
    int need = 0;
    for (int j = tid >> 3; j < q; j += JAC_NT / 8) {
      const double ajj = S[pk(j, j)];
      for (int i = tid & 7; i < j && !need; i += 8) {
        const double aij = S[pk(i, j)];
        need = aij != 0.0 && fabs(aij) > tol * sqrt(fabs(S[pk(i, i)] * ajj));
      }
    }
    if (!__syncthreads_or(need)) break;
    for (int r = 0; r < n - 1; ++r) {
      // rotation parameters (and this round's pair table)
      for (int a = tid; a < half; a += JAC_NT) {
        int p, qq;
        round_pair(n, r, a, p, qq);
        pp[a] = (short)p;
        qq_[a] = (short)qq;
        double c = 1.0, s = 0.0;
        if (qq < q) {
          const double app = S[pk(p, p)], aqq = S[pk(qq, qq)], apq = S[pk(p, qq)];
          if (apq != 0.0 && fabs(apq) > tol * sqrt(fabs(app * aqq))) jacobi_rotation(app, aqq, apq, c, s);
        }
        cs_c[a] = c; cs_s[a] = s;
        const int64_t ro = ((int64_t)(sweep * (n - 1) + r) * half + a) * 2;
        rot[ro] = c; rot[ro + 1] = s;
      }
      __syncthreads();
      if (kList) {
#pragma unroll
        for (int i = 0; i < kJacList; ++i)
          if (tid + i * JAC_NT < nblocks) jacobi_block(S, pp, qq_, cs_c, cs_s, my_a[i], my_b[i]);
      } else {
        int a = a_start, b = b_start;
        for (int blk = tid; blk < nblocks; blk += JAC_NT) {
          if (blk != tid) {
            a += JAC_NT;
            while (a > b) { a -= b + 1; ++b; }
          }
          jacobi_block(S, pp, qq_, cs_c, cs_s, a, b);
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) info[0] = sweep;
  for (int i = tid; i < q; i += JAC_NT) w[i] = S[pk(i, i)];
}

// Replay the recorded rotations on the rows of V (V starts as identity).
// Rows of V are independent under column rotations, so each wave owns one
// row in LDS and applies round after round with wave-level ordering only;
// the workgroup meets once per chunk of rotations, staged through LDS from
// registers loaded during the previous chunk.  Afterwards the eigenpairs
// are sorted by descending eigenvalue (ties: lower index first).
constexpr int kRepWaves = 4;
constexpr int kRepPrefetch = 16;  // doubles of the next chunk per thread

__device__ __forceinline__ void rep_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * kRepWaves) void jacobi_vectors_kernel(
    const double* __restrict__ rot, const int* __restrict__ info, int q, double* __restrict__ w,
    double* __restrict__ V, const double* __restrict__ wraw) {
  extern __shared__ double lds[];  // kRepWaves rows x n, then one rotation chunk
  __shared__ int perm[512];
  constexpr int NT = 64 * kRepWaves;
  const int n = q + (q & 1);
  const int half = n / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = blockIdx.x * kRepWaves + wave;  // rows >= q idle
  double* v = lds + wave * n;
  double* rs = lds + kRepWaves * n;
  for (int c = lane; c < n; c += 64) v[c] = (c == row) ? 1.0 : 0.0;
  const int rpc = max(1, kRepPrefetch * NT / (2 * half));  // rounds per chunk
  const int chunk = rpc * half * 2;
  const int total = info[0] * (n - 1);
  const int64_t rot_elems = (int64_t)total * half * 2;
  double pre[kRepPrefetch];
  auto prefetch = [&](int st0) {
#pragma unroll
    for (int i = 0; i < kRepPrefetch; ++i) {
      const int e = tid + i * NT;
      const int64_t g = (int64_t)st0 * half * 2 + e;
      pre[i] = (e < chunk && g < rot_elems) ? rot[g] : 0.0;
    }
  };
  prefetch(0);
  int rnd0 = 0;  // round index (mod n - 1) of the chunk's first round
  for (int st0 = 0; st0 < total; st0 += rpc) {
    const int nrounds = min(rpc, total - st0);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRepPrefetch; ++i)
      if (tid + i * NT < chunk) rs[tid + i * NT] = pre[i];
    __syncthreads();
    if (st0 + rpc < total) prefetch(st0 + rpc);
    if (row < q) {
      int rnd = rnd0;
      for (int rr = 0; rr < nrounds; ++rr) {
        for (int a = lane; a < half; a += 64) {
          const double c = rs[(rr * half + a) * 2], sn = rs[(rr * half + a) * 2 + 1];
          if (sn == 0.0) continue;
          int p, qv;
          round_pair(n, rnd, a, p, qv);
          const double vp = v[p], vq = v[qv];
          v[p] = c * vp - sn * vq;
          v[qv] = sn * vp + c * vq;
        }
        rep_wave_sync();
        if (++rnd == n - 1) rnd = 0;
      }
    }
    rnd0 = (rnd0 + nrounds) % (n - 1);
  }
  // descending order permutation from the raw eigenvalues
  for (int i = tid; i < q; i += NT) {
    const double wi = wraw[i];
    int rank = 0;
    for (int j = 0; j < q; ++j) {
      const double wj = wraw[j];
      rank += (wj > wi) || (wj == wi && j < i);
    }
    perm[rank] = i;
  }
  __syncthreads();
  if (row < q)
    for (int cpos = lane; cpos < q; cpos += 64) V[(int64_t)row * q + cpos] = v[perm[cpos]];
  if (blockIdx.x == 0)
    for (int i = tid; i < q; i += NT) w[i] = wraw[perm[i]];
}

// --------------------------------------------------------------------------
// Block Jacobi for q <= kBjMaxQ.  The indices are grouped in blocks of 8
// (N blocks, N even, padding indices never rotate); a block round pairs the
// blocks by the same round-robin order, and for each block pair (I, J) one
// wave rotates the 16 x 16 subproblem [[S_II, S_IJ], [S_JI, S_JJ]] in place
// (the first block round of a sweep runs all 15 inner rounds, the others only
// the 8 rounds of cross pairs (p in I, q in J), so every index pair is
// rotated once or more per sweep), accumulating V_a = product of its
// rotations.  The rest of the matrix follows on the FP64 matrix cores: every
// off-diagonal 16 x 16 tile of two block pairs (a, c) becomes
// V_a^T S_ac V_c.  A sweep is (N - 1) block rounds with two barriers each,
// against n - 1 element rounds; the V_a of every block round are recorded
// and replayed on V by row tiles with the same matrix-core products.
// S holds the upper triangle as 8 x 8 blocks (diagonal blocks use their upper
// half) in LDS.
// --------------------------------------------------------------------------
constexpr int kBjMaxQ = 160;
constexpr int kBjMaxPairs = kBjMaxQ / 16;   // block pairs per block round (N / 2)
constexpr int BJ_NT = 64 * kBjMaxPairs;
// the eigensolver kernel itself runs 16 waves: phase 1 uses one wave per
// block pair, phase 2 spreads the off-diagonal tiles over all 16
constexpr int BJE_NT = 1024;
constexpr int BJ_LD = 17;
constexpr int kBjMaxSweeps = 40;

__device__ __forceinline__ int bj_addr(int i, int j) {
  if (i > j) { const int t = i; i = j; j = t; }
  const int I = i >> 3, J = j >> 3;
  return (J * (J + 1) / 2 + I) * 64 + (i & 7) * 8 + (j & 7);
}
__device__ __forceinline__ int bj_global(int li, int I, int J) {
  return li < 8 ? 8 * I + li : 8 * J + li - 8;
}

__global__ __launch_bounds__(BJE_NT) void jacobi_blk_kernel(const double* __restrict__ Gin, int q,
                                                           double* __restrict__ w,
                                                           double* __restrict__ vrec, int* info,
                                                           double tol) {
  extern __shared__ double S[];  // 8 x 8 blocks of the upper triangle
  __shared__ double Vs[kBjMaxPairs][16 * BJ_LD];
  __shared__ double Ws[kBjMaxPairs][16 * BJ_LD];
  __shared__ double csw[kBjMaxPairs][8][2];
  __shared__ int pqw[kBjMaxPairs][8];
  __shared__ int rotw[kBjMaxPairs];  // block pair rotated in this block round
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = (q + 7) / 8;
  const int N = nb + (nb & 1);
  const int NP = N / 2;
  const int nstore = N * (N + 1) / 2 * 64;
  double* vflag = vrec + (int64_t)kBjMaxSweeps * (N - 1) * NP * 256;
  for (int e = tid; e < nstore; e += BJE_NT) S[e] = 0.0;
  __syncthreads();
  for (int e = tid; e < q * q; e += BJE_NT) {
    const int i = e / q, j = e % q;
    if (i <= j) S[bj_addr(i, j)] = 0.5 * (Gin[(int64_t)i * q + j] + Gin[(int64_t)j * q + i]);
  }
  // this lane's 2 x 2 block (ka <= kb) of the 8 inner pairs (lanes < 36)
  int ka = 0, kb = 0;
  {
    int k = 0;
    while ((k + 1) * (k + 2) / 2 <= lane) ++k;
    kb = k;
    ka = lane - k * (k + 1) / 2;
  }
  const int vr = lane >> 2, vk = 2 * (lane & 3);  // V_a row and first pair of this lane
  const int ntile = NP * (NP - 1) / 2;
  __syncthreads();

  int sweep = 0;
  for (; sweep < kBjMaxSweeps; ++sweep) {
    int need = 0;
    for (int j = tid >> 3; j < q; j += BJE_NT / 8) {
      const double ajj = S[bj_addr(j, j)];
      for (int i = tid & 7; i < j && !need; i += 8) {
        const double aij = S[bj_addr(i, j)];
        need = aij != 0.0 && aij * aij > tol * tol * fabs(S[bj_addr(i, i)] * ajj);
      }
    }
    if (!__syncthreads_or(need)) break;
    for (int r = 0; r < N - 1; ++r) {
      // ---- phase 1: wave a rotates the subproblem of block pair a ----
      // (on a dense symmetric copy W, element (x, y) at min * BJ_LD + max)
      if (wave < NP) {
        int I, J;
        round_pair(N, r, wave, I, J);
        double* V = Vs[wave];
        double* W = Ws[wave];
        for (int e = lane; e < 256; e += 64) {
          const int x = e >> 4, y = e & 15;
          if (x <= y) W[x * BJ_LD + y] = S[bj_addr(bj_global(x, I, J), bj_global(y, I, J))];
          V[x * BJ_LD + y] = (x == y) ? 1.0 : 0.0;
        }
        // local indices >= nI (block I) or >= 8 + nJ (block J) are padding
        const int nI = min(8, max(0, q - 8 * I)), nJ = min(8, max(0, q - 8 * J));
        rep_wave_sync();
        // Does any pair the inner rounds visit (all pairs in the first block
        // round of a sweep, the cross pairs otherwise) pass the threshold?  If
        // none does, no rotation happens (W only changes by rotations), so
        // the rounds are skipped and V_a stays the identity.
        bool want = false;
        for (int e = lane; e < 256; e += 64) {
          const int x = e >> 4, y = e & 15;
          if (x < y && (r == 0 || (x < 8 && y >= 8))) {
            const double apq = W[x * BJ_LD + y];
            want |= apq != 0.0 &&
                    apq * apq > tol * tol * fabs(W[x * (BJ_LD + 1)] * W[y * (BJ_LD + 1)]);
          }
        }
        const bool rotate = __any(want);
        const int inner = !rotate ? 0 : (r == 0) ? 15 : 8;
        for (int rr = 0; rr < inner; ++rr) {
          if (lane < 8) {
            int p, qq;
            if (r == 0) round_pair(16, rr, lane, p, qq);
            else { p = lane; qq = 8 + ((lane + rr) & 7); }
            double c = 1.0, sn = 0.0;
            const bool live = (p < 8 ? p < nI : p - 8 < nJ) && (qq < 8 ? qq < nI : qq - 8 < nJ);
            if (live) {
              // the three operands in one LDS round trip
              const double app = W[p * (BJ_LD + 1)], aqq = W[qq * (BJ_LD + 1)], apq = W[p * BJ_LD + qq];
              if (apq != 0.0 && apq * apq > tol * tol * fabs(app * aqq)) jacobi_rotation(app, aqq, apq, c, sn);
            }
            csw[wave][lane][0] = c;
            csw[wave][lane][1] = sn;
            pqw[wave][lane] = p | (qq << 8);
          }
          rep_wave_sync();
          // Every rotation parameter this lane needs (its W block pair and its
          // two V pairs) is read at once, and the V operands are loaded next to
          // the W operands: the round costs a few LDS round trips instead of
          // one per dependent read.  Unrotated pairs hold (1, 0) and valid
          // indices, so the loads need no branch.
          const double c0 = csw[wave][vk][0], s0 = csw[wave][vk][1];
          const double c1 = csw[wave][vk + 1][0], s1 = csw[wave][vk + 1][1];
          const int pq0 = pqw[wave][vk], pq1 = pqw[wave][vk + 1];
          double* vrow = V + vr * BJ_LD;
          const int vp0 = pq0 & 255, vq0 = pq0 >> 8, vp1 = pq1 & 255, vq1 = pq1 >> 8;
          const double v0p = vrow[vp0], v0q = vrow[vq0], v1p = vrow[vp1], v1q = vrow[vq1];
          if (lane < 36) {
            const double ca = csw[wave][ka][0], sa = csw[wave][ka][1];
            const double cb = csw[wave][kb][0], sb = csw[wave][kb][1];
            const int pqa = pqw[wave][ka], pqb = pqw[wave][kb];
            const int pa = pqa & 255, qa = pqa >> 8, pb = pqb & 255, qb = pqb >> 8;
            if (ka == kb) {
              if (sa != 0.0) {
                const int ipp = pa * (BJ_LD + 1), iqq = qa * (BJ_LD + 1), ipq = pa * BJ_LD + qa;
                const double app = W[ipp], aqq = W[iqq], apq = W[ipq];
                const double t = sa / ca;
                W[ipp] = app - t * apq;
                W[iqq] = aqq + t * apq;
                W[ipq] = 0.0;
              }
            } else if (sa != 0.0 || sb != 0.0) {
              const int i11 = min(pa, pb) * BJ_LD + max(pa, pb), i12 = min(pa, qb) * BJ_LD + max(pa, qb);
              const int i21 = min(qa, pb) * BJ_LD + max(qa, pb), i22 = min(qa, qb) * BJ_LD + max(qa, qb);
              const double x11 = W[i11], x12 = W[i12], x21 = W[i21], x22 = W[i22];
              const double y11 = x11 * cb - x12 * sb, y12 = x11 * sb + x12 * cb;
              const double y21 = x21 * cb - x22 * sb, y22 = x21 * sb + x22 * cb;
              W[i11] = ca * y11 - sa * y21;
              W[i12] = ca * y12 - sa * y22;
              W[i21] = sa * y11 + ca * y21;
              W[i22] = sa * y12 + ca * y22;
            }
          }
          // the two pairs of this lane's V row are disjoint
          if (s0 != 0.0) {
            vrow[vp0] = c0 * v0p - s0 * v0q;
            vrow[vq0] = s0 * v0p + c0 * v0q;
          }
          if (s1 != 0.0) {
            vrow[vp1] = c1 * v1p - s1 * v1q;
            vrow[vq1] = s1 * v1p + c1 * v1q;
          }
          rep_wave_sync();
        }
        if (rotate) {
          for (int e = lane; e < 256; e += 64) {
            const int x = e >> 4, y = e & 15;
            if (x <= y) S[bj_addr(bj_global(x, I, J), bj_global(y, I, J))] = W[x * BJ_LD + y];
          }
          double* out = vrec + ((int64_t)(sweep * (N - 1) + r) * NP + wave) * 256;
          for (int e = lane; e < 256; e += 64) out[e] = V[(e >> 4) * BJ_LD + (e & 15)];
        }
        if (lane == 0) {
          rotw[wave] = rotate ? 1 : 0;
          vflag[(int64_t)(sweep * (N - 1) + r) * NP + wave] = rotate ? 1.0 : 0.0;
        }
      }
      __syncthreads();
      // ---- phase 2: S_ac <- V_a^T S_ac V_c for every pair of block pairs a < c ----
      for (int u = wave; u < ntile; u += BJE_NT / 64) {
        int c = 1;
        while (c * (c + 1) / 2 <= u) ++c;
        const int a = u - c * (c - 1) / 2;
        // V_a^T S_ac V_c = S_ac when neither block pair rotated
        if (!rotw[a] && !rotw[c]) continue;
        int Ia, Ja, Ic, Jc;
        round_pair(N, r, a, Ia, Ja);
        round_pair(N, r, c, Ic, Jc);
        const int m = lane & 15, kq = lane >> 4;
        // element (x, y) of the tile lives in block (x < 8 ? Ia : Ja, y < 8 ? Ic : Jc);
        // per lane, the two blocks of its row m (loads) and of its column m (stores)
        int lb[2], sb[2];
        bool lt[2], st[2];
        {
          const int bx = m < 8 ? Ia : Ja, by = m < 8 ? Ic : Jc;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int cy = h ? Jc : Ic, rx = h ? Ja : Ia;
            lt[h] = bx > cy;
            lb[h] = lt[h] ? (bx * (bx + 1) / 2 + cy) * 64 + (m & 7) : (cy * (cy + 1) / 2 + bx) * 64 + (m & 7) * 8;
            st[h] = rx > by;
            sb[h] = st[h] ? (rx * (rx + 1) / 2 + by) * 64 + (m & 7) * 8 : (by * (by + 1) / 2 + rx) * 64 + (m & 7);
          }
        }
        cb_d4 T = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const int k = 4 * ks + kq, h = ks >> 1, yi = k & 7;
          const double av = S[lb[h] + (lt[h] ? yi * 8 : yi)];
          const double bv = Vs[c][k * BJ_LD + m];
          T = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, T, 0, 0, 0);
        }
        cb_d4 R = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const double av = Vs[a][(4 * ks + kq) * BJ_LD + m];
          R = __builtin_amdgcn_mfma_f64_16x16x4f64(av, T[ks], R, 0, 0, 0);
        }
        rep_wave_sync();
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
          const int xi = (kq + 4 * r4) & 7, h = r4 >> 1;
          S[sb[h] + (st[h] ? xi : xi * 8)] = R[r4];
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) info[0] = sweep;
  for (int i = tid; i < q; i += BJE_NT) w[i] = S[bj_addr(i, i)];
}

// Replay of the recorded block rotations on V, one workgroup per 16-row tile
// (rows held in LDS); the V_a of the next block round are loaded into
// registers while the current one is applied.  Then the eigenpairs are
// sorted by descending eigenvalue (ties: lower index first).
__global__ __launch_bounds__(BJ_NT) void jacobi_blk_vectors_kernel(
    const double* __restrict__ vrec, const int* __restrict__ info, int q, double* __restrict__ w,
    double* __restrict__ V, const double* __restrict__ wraw) {
  __shared__ double Vt[16][kBjMaxQ + 1];
  __shared__ int perm[kBjMaxQ];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = (q + 7) / 8;
  const int N = nb + (nb & 1);
  const int NP = N / 2;
  const int P = 8 * N;
  const int row0 = blockIdx.x * 16;
  for (int e = tid; e < 16 * P; e += BJ_NT) {
    const int i = e / P, j = e % P;
    Vt[i][j] = (row0 + i == j) ? 1.0 : 0.0;
  }
  const int nrounds = info[0] * (N - 1);
  const double* vflag = vrec + (int64_t)kBjMaxSweeps * (N - 1) * NP * 256;
  const int m = lane & 15, kq = lane >> 4;
  double bnext[4] = {0.0, 0.0, 0.0, 0.0};
  if (wave < NP && nrounds > 0) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) bnext[ks] = vrec[(int64_t)wave * 256 + (4 * ks + kq) * 16 + m];
  }
  for (int t = 0; t < nrounds; ++t) {
    __syncthreads();
    if (wave < NP) {
      double b[4];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) b[ks] = bnext[ks];
      if (t + 1 < nrounds) {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          bnext[ks] = vrec[((int64_t)(t + 1) * NP + wave) * 256 + (4 * ks + kq) * 16 + m];
      }
      // block pairs that did not rotate recorded no V_a (identity)
      if (vflag[(int64_t)t * NP + wave] == 0.0) continue;
      int I, J;
      round_pair(N, t % (N - 1), wave, I, J);
      cb_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Vt[m][bj_global(4 * ks + kq, I, J)], b[ks], acc, 0, 0, 0);
      rep_wave_sync();
      const int gn = bj_global(m, I, J);
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) Vt[kq + 4 * r4][gn] = acc[r4];
    }
  }
  for (int i = tid; i < q; i += BJ_NT) {
    const double wi = wraw[i];
    int rank = 0;
    for (int j = 0; j < q; ++j) {
      const double wj = wraw[j];
      rank += (wj > wi) || (wj == wi && j < i);
    }
    perm[rank] = i;
  }
  __syncthreads();
  for (int e = tid; e < 16 * q; e += BJ_NT) {
    const int i = e / q, pos = e % q;
    if (row0 + i < q) V[(int64_t)(row0 + i) * q + pos] = Vt[i][perm[pos]];
  }
  if (blockIdx.x == 0)
    for (int i = tid; i < q; i += BJ_NT) w[i] = wraw[perm[i]];
}

// --------------------------------------------------------------------------
__global__ void count_kept_kernel(const double* __restrict__ w, int q, int chi, double cutoff,
                                  double* __restrict__ s_out, int* cD_out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const double s0 = sqrt(fmax(w[0], 0.0));
  for (int i = threadIdx.x; i < q; i += blockDim.x) {
    const double si = sqrt(fmax(w[i], 0.0));
    s_out[i] = si;
    if (s0 > 0.0 && si / s0 > cutoff) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) cD_out[0] = min(chi, cnt);
}

__global__ void scale_columns_kernel(const double* __restrict__ Ub, int q,
                                     const double* __restrict__ s, int cD, int mode,
                                     const int* __restrict__ kept, double* __restrict__ X) {
  const double s0 = s[0];
  // with ``kept`` (a count on the device), columns z >= kept[0] are zero
  const int nk = kept ? kept[0] : cD;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < q * cD; e += gridDim.x * blockDim.x) {
    const int i = e / cD, z = e % cD;
    if (z >= nk) {
      X[e] = 0.0;
      continue;
    }
    const double sz = s[z];
    const double sn = sz / s0;
    const double f = (mode == 0) ? 1.0 / sqrt(sn) : 1.0 / (sz * sqrt(sn));
    X[e] = Ub[(int64_t)i * q + z] * f;
  }
}

// one block per column z: sign of sum_i w_i P1[i, z], w_i = 1 + frac(i*phi)
// (fixed asymmetric weights: no symmetry-forced ties).
__global__ void gauge_kernel(double* __restrict__ P1, double* __restrict__ P2, int64_t rows1,
                             int64_t rows2, int cD) {
  const int z = blockIdx.x;
  __shared__ double red[256];
  double acc = 0.0;
  // frac(x) = x - floor(x), exact for 0 <= x < 2^52 (the value fmod(x, 1) gives)
  for (int64_t i = threadIdx.x; i < rows1; i += blockDim.x) {
    const double x = (double)i * 0.6180339887498949;
    acc += (1.0 + (x - floor(x))) * P1[i * cD + z];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (!(red[0] < 0.0)) return;
  for (int64_t i = threadIdx.x; i < rows1; i += blockDim.x) P1[i * cD + z] = -P1[i * cD + z];
  for (int64_t i = threadIdx.x; i < rows2; i += blockDim.x) P2[i * cD + z] = -P2[i * cD + z];
}

constexpr int kNormBlocks = 512;
__global__ void sumsq_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void scale_by_norm_kernel(double* __restrict__ x, int64_t n,
                                     const double* __restrict__ part, int nparts) {
  // every block sums the partials the same way: strided per thread, then a
  // fixed tree (blockDim.x = 256)
  __shared__ double red[256];
  double t = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) t += part[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double inv = 1.0 / sqrt(red[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= inv;
}

__global__ void zero_kernel(double* x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

size_t packed_bytes(int n) { return sizeof(double) * (size_t)n * (n + 1) / 2; }

}  // namespace

void gaussian_sketch(double* out, int64_t n, int64_t q, uint64_t seed, uint64_t stream,
                     hipStream_t s, const uint64_t* stream_base) {
  const int64_t total = n * q;
  if (total <= 0) return;
  const int64_t pairs = (total + 1) / 2;
  const int blocks = (int)std::min<int64_t>((pairs + 255) / 256, 4096);
  hipLaunchKernelGGL(sketch_kernel, dim3(blocks), dim3(256), 0, s, out, total, seed, stream,
                     stream_base);
  ACETN_HIP_CHECK(hipGetLastError());
}

void chol_rinv(const double* G, double* Rinv, int q, double shift_rel, double rank_tol,
               hipStream_t s, double* gscratch, int* dead_out, const int* only_if_dead) {
  ACETN_REQUIRE(q >= 1 && q <= 512, "chol_rinv: q out of range");
  // q^3/3 (factor) + q^3/6 (inverse)
  hipEvent_t e0 = kernel_timing_begin(s);
  struct End {
    hipEvent_t e0; hipStream_t s; double f;
    ~End() { if (e0) kernel_timing_end(e0, s, kKCholesky, f, 0.0); }
  } timing_end{e0, s, 0.5 * (double)q * q * q};
  const size_t pb = packed_bytes(q);
  if (q <= kCbMaxQ) {
    const int Q = (q + 15) / 16;
    const size_t tb = sizeof(double) * (size_t)CB_TS * Q * (Q + 1) / 2;
    static bool attr = false;
    if (!attr) {
      ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)chol_rinv_blk_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
      attr = true;
    }
    hipLaunchKernelGGL(chol_rinv_blk_kernel, dim3(1), dim3(CB_NT), tb, s, G, Rinv, q, shift_rel,
                       rank_tol, dead_out, only_if_dead);
    ACETN_HIP_CHECK(hipGetLastError());
    return;
  }
  const int use_lds = pb <= kLdsBudget;
  ACETN_REQUIRE(use_lds || gscratch != nullptr, "chol_rinv: scratch required");
  if (use_lds)
    hipLaunchKernelGGL(chol_rinv_kernel<true>, dim3(1), dim3(LA_NT), pb, s, G, Rinv, q, shift_rel,
                       rank_tol, gscratch, use_lds, dead_out, only_if_dead);
  else
    hipLaunchKernelGGL(chol_rinv_kernel<false>, dim3(1), dim3(LA_NT), 0, s, G, Rinv, q, shift_rel,
                       rank_tol, gscratch, use_lds, dead_out, only_if_dead);
  ACETN_HIP_CHECK(hipGetLastError());
}

void keep_dead_columns(double* Y, int64_t m, int q, const int* dead, hipStream_t s) {
  const int blocks = (int)std::min<int64_t>((m * q + 255) / 256, 2048);
  hipLaunchKernelGGL(keep_dead_columns_kernel, dim3(blocks), dim3(256), 0, s, Y, m, q, dead);
  ACETN_HIP_CHECK(hipGetLastError());
}

void still_dead(int* still, const int* dead, int q, hipStream_t s) {
  hipLaunchKernelGGL(still_dead_kernel, dim3(1), dim3(256), 0, s, still, dead, q);
  ACETN_HIP_CHECK(hipGetLastError());
}

void refill_dead_columns(double* Y, int64_t m, int q, const int* dead, uint64_t seed,
                         hipStream_t s) {
  const int blocks = (int)std::min<int64_t>((m + 255) / 256, 1024);
  hipLaunchKernelGGL(refill_kernel, dim3(blocks), dim3(256), 0, s, Y, m, q, dead, seed);
  ACETN_HIP_CHECK(hipGetLastError());
}

// recorded rotations: element rounds ((c, s) per pair) or, for q <= kBjMaxQ,
// block rounds (a 16 x 16 V_a per block pair)
size_t jacobi_record_elems(int q) {
  const int n = q + (q & 1);
  size_t e = (size_t)kMaxSweeps * (n - 1) * (n / 2) * 2;
  if (q <= kBjMaxQ) {
    // the 16 x 16 V_a of every (sweep, block round, pair), then one flag each
    const int nb = (q + 7) / 8, N = nb + (nb & 1);
    e = std::max(e, (size_t)kBjMaxSweeps * (N - 1) * (N / 2) * 257);
  }
  return e;
}

size_t jacobi_rot_elems(int q) {
  const int n = q + (q & 1);
  // rotations + raw eigenvalues (q) + info (as doubles) + packed scratch
  return jacobi_record_elems(q) + (size_t)q + 8 + (size_t)n * (n + 1) / 2;
}

void jacobi_eigh(double* G, double* w, double* V, double* rot, int q, hipStream_t s, double tol) {
  ACETN_REQUIRE(q >= 1 && q <= 510, "jacobi_eigh: q out of range");
  hipEvent_t e0 = kernel_timing_begin(s);
  struct End {
    hipEvent_t e0; hipStream_t s;
    ~End() { if (e0) kernel_timing_end(e0, s, kKJacobi, 0.0, 0.0); }
  } timing_end{e0, s};
  const int n = q + (q & 1);
  double* wraw = rot + jacobi_record_elems(q);
  int* info = reinterpret_cast<int*>(wraw + q);
  double* gscr = wraw + q + 8;
  if (q <= kBjMaxQ) {
    const int nb = (q + 7) / 8, N = nb + (nb & 1);
    const size_t lds = sizeof(double) * (size_t)N * (N + 1) / 2 * 64;
    static bool attr = false;
    if (!attr) {
      // the largest block store (N = kBjMaxQ / 8 blocks) next to the static LDS
      constexpr int kNmax = kBjMaxQ / 8;
      ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)jacobi_blk_kernel,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)(sizeof(double) * kNmax * (kNmax + 1) / 2 * 64)));
      attr = true;
    }
    hipLaunchKernelGGL(jacobi_blk_kernel, dim3(1), dim3(BJE_NT), lds, s, G, q, wraw, rot, info, tol);
    ACETN_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(jacobi_blk_vectors_kernel, dim3((8 * N + 15) / 16), dim3(BJ_NT), 0, s, rot,
                       info, q, w, V, wraw);
    ACETN_HIP_CHECK(hipGetLastError());
    return;
  }
  const size_t pb = packed_bytes(n);
  const int use_lds = pb <= kLdsBudget;
  const int half = n / 2;
  const bool list = half * (half + 1) / 2 <= kJacList * JAC_NT;
  if (list && use_lds)
    hipLaunchKernelGGL((jacobi_kernel<true, true>), dim3(1), dim3(JAC_NT), pb, s, G, q, wraw, rot,
                       info, gscr, use_lds, tol);
  else if (use_lds)
    hipLaunchKernelGGL((jacobi_kernel<false, true>), dim3(1), dim3(JAC_NT), pb, s, G, q, wraw, rot,
                       info, gscr, use_lds, tol);
  else
    hipLaunchKernelGGL((jacobi_kernel<false, false>), dim3(1), dim3(JAC_NT), 0, s, G, q, wraw, rot,
                       info, gscr, use_lds, tol);
  ACETN_HIP_CHECK(hipGetLastError());
  const int nblk = (q + kRepWaves - 1) / kRepWaves;
  const size_t lds = sizeof(double) * ((size_t)kRepWaves * n + (size_t)kRepPrefetch * 64 * kRepWaves);
  hipLaunchKernelGGL(jacobi_vectors_kernel, dim3(nblk), dim3(64 * kRepWaves), lds, s, rot, info, q,
                     w, V, wraw);
  ACETN_HIP_CHECK(hipGetLastError());
}

void scale_columns(const double* Ub, int q, const double* s, int cD, int mode, double* X,
                   hipStream_t st, const int* kept) {
  if (cD <= 0) return;
  const int blocks = std::min((q * cD + 255) / 256, 1024);
  hipLaunchKernelGGL(scale_columns_kernel, dim3(blocks), dim3(256), 0, st, Ub, q, s, cD, mode,
                     kept, X);
  ACETN_HIP_CHECK(hipGetLastError());
}

void count_kept(const double* w, int q, int chi, double cutoff, double* s_out, int* cD_out,
                hipStream_t st) {
  hipLaunchKernelGGL(count_kept_kernel, dim3(1), dim3(256), 0, st, w, q, chi, cutoff, s_out, cD_out);
  ACETN_HIP_CHECK(hipGetLastError());
}

void fix_projector_gauge(double* P1, double* P2, int64_t rows1, int64_t rows2, int cD,
                         hipStream_t st) {
  if (cD <= 0) return;
  hipLaunchKernelGGL(gauge_kernel, dim3(cD), dim3(256), 0, st, P1, P2, rows1, rows2, cD);
  ACETN_HIP_CHECK(hipGetLastError());
}

void frobenius_normalize(double* x, int64_t n, double* scratch, hipStream_t st) {
  if (n <= 0) return;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, kNormBlocks);
  hipLaunchKernelGGL(sumsq_kernel, dim3(blocks), dim3(256), 0, st, x, n, scratch);
  ACETN_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(scale_by_norm_kernel, dim3(blocks), dim3(256), 0, st, x, n, scratch, blocks);
  ACETN_HIP_CHECK(hipGetLastError());
}

void zero_fill(double* x, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, st, x, n);
  ACETN_HIP_CHECK(hipGetLastError());
}

}  // namespace acetn
