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

__global__ void sketch_kernel(double* out, int64_t total, uint64_t seed, uint64_t stream) {
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

// Register-resident variant for q <= kCrMaxQ: the same factor and inverse,
// with the upper triangle held in registers.  Thread t owns a run of E
// consecutive rows i0 <= i < i0 + E of one column k (i <= k), so a step reads
// one value for the column and E consecutive values for the rows.  Both
// phases are right-looking with one barrier per step:
//   Cholesky, step j:  row j (raw) is published in a double-buffered row,
//                      R[j][k] = S[j][k] / sqrt(S[j][j]) goes to the packed
//                      factor in LDS, S[i][k] -= S[j][i] S[j][k] / S[j][j]
//                      for j < i <= k in registers;
//   inverse, step r (descending):  X[r][c] = T[r][c] / R[r][r] is published,
//                      T[i][c] -= R[i][r] X[r][c] for i < r <= c,
// with T = I initially, so X = R^-1 row by row.  Dead pivots give zero rows
// of R and zero rows and columns of X, as in chol_rinv_kernel.
constexpr int CR_NT = 1024;
constexpr int kCrMaxQ = 173;  // packed factor (<= 120 KB) in LDS, runs of <= 16 rows

// column runs of length E covering the upper triangle of a q x q matrix
int cr_runs(int q, int E) {
  int n = 0;
  for (int k = 0; k < q; ++k) n += (k + E) / E;
  return n;
}

template <int E>
__global__ __launch_bounds__(CR_NT) void chol_rinv_reg_kernel(const double* __restrict__ G,
                                                              double* __restrict__ Rinv, int q,
                                                              double shift_rel, double rank_tol,
                                                              int* dead_out,
                                                              const int* only_if_dead) {
  extern __shared__ double RS[];  // packed upper factor, pk(i, k)
  __shared__ double buf[2][kCrMaxQ + 4];
  __shared__ int dead[kCrMaxQ + 4];
  __shared__ double s_red[CR_NT / 64][2];
  __shared__ int any;
  const int tid = threadIdx.x;
  if (only_if_dead) {
    if (tid == 0) any = 0;
    __syncthreads();
    for (int i = tid; i < q; i += CR_NT)
      if (only_if_dead[i]) any = 1;
    __syncthreads();
    if (!any) {
      for (int e = tid; e < q * q; e += CR_NT) Rinv[e] = (e / q == e % q) ? 1.0 : 0.0;
      return;
    }
  }
  double tr = 0.0, dm = 0.0;
  for (int i = tid; i < q; i += CR_NT) {
    const double gii = G[(int64_t)i * q + i];
    tr += gii;
    dm = fmax(dm, gii);
  }
  for (int o = 32; o > 0; o >>= 1) { tr += __shfl_down(tr, o); dm = fmax(dm, __shfl_down(dm, o)); }
  if ((tid & 63) == 0) { s_red[tid >> 6][0] = tr; s_red[tid >> 6][1] = dm; }
  __syncthreads();
  if (tid == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < CR_NT / 64; ++w) { t0 += s_red[w][0]; t1 = fmax(t1, s_red[w][1]); }
    s_red[0][0] = t0; s_red[0][1] = t1;
  }
  __syncthreads();
  const double shift = shift_rel * s_red[0][0];
  const double thresh = rank_tol * (s_red[0][1] + shift);

  // this thread's run: column kc, rows i0 .. min(i0 + E, kc + 1) - 1 (kc = -1: none)
  int kc = -1, i0 = 0;
  {
    int acc = 0;
    for (int k = 0; k < q; ++k) {
      const int cnt = (k + E) / E;
      if (tid < acc + cnt) { kc = k; i0 = (tid - acc) * E; break; }
      acc += cnt;
    }
  }
  const int iend = kc + 1;  // rows i < iend are in the triangle
  double v[E];
#pragma unroll
  for (int t = 0; t < E; ++t) {
    const int i = i0 + t;
    v[t] = (i < iend) ? G[(int64_t)i * q + kc] + (i == kc ? shift : 0.0) : 0.0;
  }
  const int colbase = kc * (kc + 1) / 2;  // pk(i, kc) = colbase + i

  for (int j = 0; j < q; ++j) {
    double* rb = buf[j & 1];
#pragma unroll
    for (int t = 0; t < E; ++t)
      if (i0 + t == j && j < iend) rb[kc] = v[t];
    __syncthreads();
    const double d = rb[j];
    const bool is_dead = !(d > thresh) || !isfinite(d);
    if (tid == 0) dead[j] = is_dead ? 1 : 0;
    if (kc < j) continue;  // column finished (uniform per thread)
    if (is_dead) {
      if (j >= i0 && j < i0 + E) RS[colbase + j] = 0.0;
      continue;
    }
    const double rk = rb[kc] * (1.0 / d);
#pragma unroll
    for (int t = 0; t < E; ++t) {
      const int i = i0 + t;
      if (i == j) RS[colbase + j] = (kc == j) ? sqrt(d) : v[t] / sqrt(d);
      else if (i > j && i < iend) v[t] -= rb[i] * rk;
    }
  }
  __syncthreads();
  if (dead_out) {
    int anyd = 0;
    for (int i = tid; i < q; i += CR_NT) {
      dead_out[i] = dead[i];
      anyd |= dead[i];
    }
    anyd = __syncthreads_or(anyd);
    if (tid == 0) dead_out[q] = anyd;
  }

#pragma unroll
  for (int t = 0; t < E; ++t) v[t] = (i0 + t == kc) ? 1.0 : 0.0;
  for (int r = q - 1; r >= 0; --r) {
    double* xb = buf[r & 1];
    const bool dr = dead[r] != 0;
    const double inv = dr ? 0.0 : 1.0 / RS[r * (r + 1) / 2 + r];
#pragma unroll
    for (int t = 0; t < E; ++t)
      if (i0 + t == r && r < iend) {
        v[t] *= inv;
        xb[kc] = v[t];
      }
    __syncthreads();
    if (dr || kc < r) continue;
    const double x = xb[kc];
    const double* rcol = RS + r * (r + 1) / 2;  // R[i][r], i <= r
#pragma unroll
    for (int t = 0; t < E; ++t)
      if (i0 + t < r) v[t] -= rcol[i0 + t] * x;
  }
#pragma unroll
  for (int t = 0; t < E; ++t)
    if (i0 + t < iend) Rinv[(int64_t)(i0 + t) * q + kc] = v[t];
  for (int e = tid; e < q * q; e += CR_NT)
    if (e / q > e % q) Rinv[e] = 0.0;
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

// --------------------------------------------------------------------------
// Jacobi eigensolver.  Round-robin ("circle method") parallel ordering over
// n = q rounded up to even indices; the padding index, if any, never rotates.
// rot layout: [(sweep*(n-1) + round) * (n/2) + pair] * 2 -> (c, s);
// info[0] = number of sweeps executed.
// --------------------------------------------------------------------------
__device__ __forceinline__ void round_pair(int n, int r, int a, int& p, int& qq) {
  int x, y;
  if (a == 0) { x = n - 1; y = r; }
  else { x = (r + a) % (n - 1); y = (r - a + n - 1) % (n - 1); }
  p = min(x, y); qq = max(x, y);
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
          if (apq != 0.0 && fabs(apq) > tol * sqrt(fabs(app * aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
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
// Each workgroup owns ``rows`` rows in LDS.  Afterwards the eigenpairs are
// sorted by descending eigenvalue (ties: lower index first).
constexpr int kReplayRounds = 32;
constexpr int kRepList = 8;  // (row, pair) items per thread: rows * n/2 <= 8 * 256

__global__ __launch_bounds__(256) void jacobi_vectors_kernel(const double* __restrict__ rot,
                                                             const int* __restrict__ info, int q,
                                                             int rows, double* __restrict__ w,
                                                             double* __restrict__ V,
                                                             const double* __restrict__ wraw) {
  extern __shared__ double vr[];  // rows x n, then kReplayRounds x half x 2 rotations
  __shared__ int perm[512];
  __shared__ short pp[kReplayRounds][256], qq_[kReplayRounds][256];
  const int n = q + (q & 1);
  const int half = n / 2;
  double* rs = vr + rows * n;
  const int r0 = blockIdx.x * rows;
  const int nr = min(rows, q - r0);
  const int tid = threadIdx.x;
  for (int e = tid; e < rows * n; e += blockDim.x) {
    const int rr = e / n, cc = e % n;
    vr[e] = (rr < nr && r0 + rr == cc) ? 1.0 : 0.0;
  }
  // (row, pair) work items of this thread, identical in every round
  int my_row[kRepList], my_pair[kRepList];
#pragma unroll
  for (int i = 0; i < kRepList; ++i) {
    const int e = tid + i * 256;
    my_pair[i] = e % half;
    my_row[i] = e / half;
  }
  const int total = info[0] * (n - 1);
  for (int st0 = 0; st0 < total; st0 += kReplayRounds) {
    const int nrounds = min(kReplayRounds, total - st0);
    __syncthreads();
    for (int e = tid; e < nrounds * half * 2; e += blockDim.x)
      rs[e] = rot[(int64_t)st0 * half * 2 + e];
    for (int e = tid; e < nrounds * half; e += blockDim.x) {
      const int rr = e / half, a = e % half;
      int p, qv;
      round_pair(n, (st0 + rr) % (n - 1), a, p, qv);
      pp[rr][a] = (short)p;
      qq_[rr][a] = (short)qv;
    }
    __syncthreads();
    for (int rr = 0; rr < nrounds; ++rr) {
#pragma unroll
      for (int i = 0; i < kRepList; ++i) {
        if (tid + i * 256 >= half * nr) break;
        const int a = my_pair[i];
        const double c = rs[(rr * half + a) * 2], sn = rs[(rr * half + a) * 2 + 1];
        if (sn == 0.0) continue;
        double* v = vr + my_row[i] * n;
        const int p = pp[rr][a], qv = qq_[rr][a];
        const double vp = v[p], vq = v[qv];
        v[p] = c * vp - sn * vq;
        v[qv] = sn * vp + c * vq;
      }
      __syncthreads();
    }
  }
  // descending order permutation from the raw eigenvalues
  for (int i = tid; i < q; i += blockDim.x) {
    const double wi = wraw[i];
    int rank = 0;
    for (int j = 0; j < q; ++j) {
      const double wj = wraw[j];
      rank += (wj > wi) || (wj == wi && j < i);
    }
    perm[rank] = i;
  }
  __syncthreads();
  for (int e = tid; e < nr * q; e += blockDim.x) {
    const int rr = e / q, cpos = e % q;
    V[(int64_t)(r0 + rr) * q + cpos] = vr[rr * n + perm[cpos]];
  }
  if (blockIdx.x == 0)
    for (int i = tid; i < q; i += blockDim.x) w[i] = wraw[perm[i]];
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
                                     double* __restrict__ X) {
  const double s0 = s[0];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < q * cD; e += gridDim.x * blockDim.x) {
    const int i = e / cD, z = e % cD;
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
  for (int64_t i = threadIdx.x; i < rows1; i += blockDim.x)
    acc += (1.0 + fmod((double)i * 0.6180339887498949, 1.0)) * P1[i * cD + z];
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
  __shared__ double inv;
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < nparts; ++i) t += part[i];
    inv = 1.0 / sqrt(t);
  }
  __syncthreads();
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
                     hipStream_t s) {
  const int64_t total = n * q;
  if (total <= 0) return;
  const int64_t pairs = (total + 1) / 2;
  const int blocks = (int)std::min<int64_t>((pairs + 255) / 256, 4096);
  hipLaunchKernelGGL(sketch_kernel, dim3(blocks), dim3(256), 0, s, out, total, seed, stream);
  ACETN_HIP_CHECK(hipGetLastError());
}

void chol_rinv(const double* G, double* Rinv, int q, double shift_rel, double rank_tol,
               hipStream_t s, double* gscratch, int* dead_out, const int* only_if_dead) {
  ACETN_REQUIRE(q >= 1 && q <= 512, "chol_rinv: q out of range");
  const size_t pb = packed_bytes(q);
  if (q <= kCrMaxQ) {
    ACETN_REQUIRE(cr_runs(q, 16) <= CR_NT, "chol_rinv: register layout too small");
    if (cr_runs(q, 8) <= CR_NT)
      hipLaunchKernelGGL(chol_rinv_reg_kernel<8>, dim3(1), dim3(CR_NT), pb, s, G, Rinv, q,
                         shift_rel, rank_tol, dead_out, only_if_dead);
    else if (cr_runs(q, 12) <= CR_NT)
      hipLaunchKernelGGL(chol_rinv_reg_kernel<12>, dim3(1), dim3(CR_NT), pb, s, G, Rinv, q,
                         shift_rel, rank_tol, dead_out, only_if_dead);
    else
      hipLaunchKernelGGL(chol_rinv_reg_kernel<16>, dim3(1), dim3(CR_NT), pb, s, G, Rinv, q,
                         shift_rel, rank_tol, dead_out, only_if_dead);
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

void refill_dead_columns(double* Y, int64_t m, int q, const int* dead, uint64_t seed,
                         hipStream_t s) {
  const int blocks = (int)std::min<int64_t>((m + 255) / 256, 1024);
  hipLaunchKernelGGL(refill_kernel, dim3(blocks), dim3(256), 0, s, Y, m, q, dead, seed);
  ACETN_HIP_CHECK(hipGetLastError());
}

size_t jacobi_rot_elems(int q) {
  const int n = q + (q & 1);
  // rotations + raw eigenvalues (q) + info (as doubles) + packed scratch
  return (size_t)kMaxSweeps * (n - 1) * (n / 2) * 2 + (size_t)q + 8 + (size_t)n * (n + 1) / 2;
}

void jacobi_eigh(double* G, double* w, double* V, double* rot, int q, hipStream_t s, double tol) {
  ACETN_REQUIRE(q >= 1 && q <= 510, "jacobi_eigh: q out of range");
  const int n = q + (q & 1);
  double* wraw = rot + (size_t)kMaxSweeps * (n - 1) * (n / 2) * 2;
  int* info = reinterpret_cast<int*>(wraw + q);
  double* gscr = wraw + q + 8;
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
  const int rows = 8;
  const int nblk = (q + rows - 1) / rows;
  const size_t lds = sizeof(double) * ((size_t)rows * n + (size_t)kReplayRounds * (n / 2) * 2);
  hipLaunchKernelGGL(jacobi_vectors_kernel, dim3(nblk), dim3(256), lds, s, rot, info, q, rows, w,
                     V, wraw);
  ACETN_HIP_CHECK(hipGetLastError());
}

void scale_columns(const double* Ub, int q, const double* s, int cD, int mode, double* X,
                   hipStream_t st) {
  if (cD <= 0) return;
  const int blocks = std::min((q * cD + 255) / 256, 1024);
  hipLaunchKernelGGL(scale_columns_kernel, dim3(blocks), dim3(256), 0, st, Ub, q, s, cD, mode, X);
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
