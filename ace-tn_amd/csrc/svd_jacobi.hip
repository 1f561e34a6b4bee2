// One-sided (Hestenes) Jacobi SVD of a small square FP64 matrix M, one
// workgroup per problem.  Pairs of columns of M are rotated until every pair
// is orthogonal to working precision relative to the pair's norms:
//   rotate (x, y) while |x.y| > tol * |x| |y|,
// with the rotation of Rutishauser / de Rijk computed from the three dot
// products; the column norms are then the singular values and the
// normalised columns the left singular vectors (no rotation is
// accumulated).  Nothing is squared: each rotation is applied to the columns
// themselves, so the singular values carry the absolute accuracy
// eps * |M| of a backward-stable SVD.
//
// Parallel ordering: the "circle" round robin over n columns (n even) --
// n/2 slots hold a (top, bottom) pair each; after every round the tops move
// one slot to the right, the bottoms one to the left, slot 0's top stays and
// the ends wrap (slot 0's bottom becomes slot 1's top, the last top the last
// bottom).  Every pair of columns meets once in the n - 1 rounds of a sweep.
//
// hj_reg_kernel (q <= kRegMaxQ): the columns live in registers.  A 16-lane
// DPP row ("group") owns two slots, every lane R rows of each of the four
// columns (row = lane + 16 r); dot products are row reductions by DPP
// butterflies, and between rounds a group passes one top column to the next
// group and one bottom column to the previous one through LDS.
// hj_glb_kernel (larger q): the columns live in global memory (L2), one
// wave rotates one pair at a time with the two columns in registers.
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace acetn {
namespace {

constexpr int kRegMaxR = 12;                // rows per lane of the register kernel
constexpr int kRegMaxQ = 16 * kRegMaxR;     // 192
constexpr int kGlbNT = 1024;
constexpr int kMaxSweepsSvd = 40;

// double moved across lanes by a DPP control (two 32-bit halves)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a DPP row, the same bits in every lane (each step
// adds the two operands in either order, and addition commutes).
__device__ __forceinline__ double row_sum(double x) {
  x += dpp_mov<0xB1>(x);   // quad_perm [1,0,3,2]: lane ^ 1
  x += dpp_mov<0x4E>(x);   // quad_perm [2,3,0,1]: lane ^ 2
  x += dpp_mov<0x141>(x);  // row_half_mirror: the other quad of the 8
  x += dpp_mov<0x140>(x);  // row_mirror: the other half of the row
  return x;
}

// Sum over the 64 lanes of the wave, the same bits in every lane.
__device__ __forceinline__ double wave_sum(double x) {
  x = row_sum(x);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

// 1/x and 1/sqrt(x) from the hardware approximations with two Newton steps
// each (a rotation only needs c^2 + s^2 = 1 to working precision, which the
// refined c carries; t itself may be off by a few ulps)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

// Rotation that makes x' = c x - s y and y' = s x + c y orthogonal, given
// a = |x|^2, b = |y|^2, g = x.y; false (no rotation) when the pair already
// satisfies |g| <= tol |x| |y|, or when either column is negligible
// (|x|^2 or |y|^2 <= small2: below the roundoff of the whole matrix, where
// relative orthogonality means nothing -- a matrix with more columns than
// rank leaves such columns behind).  t = tan(theta) is the smaller root of
// t^2 + 2 zeta t - 1 = 0, zeta = (b - a) / (2 g).
__device__ __forceinline__ bool jacobi_rotation(double a, double b, double g, double tol,
                                                double small2, double& c, double& s) {
  if (!(a > small2 && b > small2)) return false;
  if (!(fabs(g) > tol * (sqrt(a) * sqrt(b)))) return false;
  const double zeta = (b - a) / (2.0 * g);
  const double az = fabs(zeta);
  double t;
  if (az > 1e150) {
    t = 0.5 * rcp_nr(zeta);
  } else {
    const double w = fma(zeta, zeta, 1.0);
    t = copysign(rcp_nr(az + w * rsqrt_nr(w)), zeta);
  }
  c = rsqrt_nr(fma(t, t, 1.0));
  s = c * t;
  return true;
}

// Rank of value v at position p among n values in LDS (descending; ties
// broken by position), counted by the 16 lanes of a row.
__device__ __forceinline__ int rank_in_row(const double* sv, int n, double v, int p, int l16) {
  int cnt = 0;
  for (int i = l16; i < n; i += 16) {
    const double u = sv[i];
    cnt += (u > v || (u == v && i < p)) ? 1 : 0;
  }
  return (int)row_sum((double)cnt);
}

template <int R>
__global__ __launch_bounds__(64 * R) void hj_reg_kernel(const double* __restrict__ Mc, int q,
                                                        int cols, int64_t ldc, double tol,
                                                        int max_sweeps, double* __restrict__ sig,
                                                        double* __restrict__ U, int* info) {
  // q rows (and outputs), cols columns; both <= 16 R
  // groups G = 4R, slots 2G, columns n = 4G, rows 16R
  constexpr int G = 4 * R, n = 4 * G;
  extern __shared__ double xch[];  // xtop, xbot: [G + 1][R][16] each
  double* xtop = xch;
  double* xbot = xch + (G + 1) * R * 16;
  const int tid = threadIdx.x;
  const int g = tid >> 4, l = tid & 15;
  // register vectors, not arrays: a select between an array element and an
  // LDS value would otherwise be turned into a load through a selected
  // pointer, which sends the array to scratch memory
  typedef double colv __attribute__((ext_vector_type(R)));
  colv t0, b0, t1, b1;
  // slot 2g holds columns (4g, 4g + 1), slot 2g + 1 columns (4g + 2, 4g + 3)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = l + 16 * r;
    const bool ir = i < q;
    const int j = 4 * g;
    t0[r] = (ir && j < cols) ? Mc[(int64_t)j * ldc + i] : 0.0;
    b0[r] = (ir && j + 1 < cols) ? Mc[(int64_t)(j + 1) * ldc + i] : 0.0;
    t1[r] = (ir && j + 2 < cols) ? Mc[(int64_t)(j + 2) * ldc + i] : 0.0;
    b1[r] = (ir && j + 3 < cols) ? Mc[(int64_t)(j + 3) * ldc + i] : 0.0;
  }
  // |M|_F^2, summed in the same order by every thread
  double small2;
  {
    double a = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      a = fma(t0[r], t0[r], a);
      a = fma(b0[r], b0[r], a);
      a = fma(t1[r], t1[r], a);
      a = fma(b1[r], b1[r], a);
    }
    a = row_sum(a);
    if (l == 0) xch[g] = a;
    __syncthreads();
    double fro = 0.0;
    for (int e = 0; e < G; ++e) fro += xch[e];
    __syncthreads();
    small2 = tol * tol * fro;
  }
  int sweeps = 0, converged = 0;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    int rotated = 0;
    for (int round = 0; round < n - 1; ++round) {
      double aa0 = 0.0, bb0 = 0.0, gg0 = 0.0, aa1 = 0.0, bb1 = 0.0, gg1 = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        aa0 = fma(t0[r], t0[r], aa0);
        bb0 = fma(b0[r], b0[r], bb0);
        gg0 = fma(t0[r], b0[r], gg0);
        aa1 = fma(t1[r], t1[r], aa1);
        bb1 = fma(b1[r], b1[r], bb1);
        gg1 = fma(t1[r], b1[r], gg1);
      }
      aa0 = row_sum(aa0); bb0 = row_sum(bb0); gg0 = row_sum(gg0);
      aa1 = row_sum(aa1); bb1 = row_sum(bb1); gg1 = row_sum(gg1);
      double c, s;
      if (jacobi_rotation(aa0, bb0, gg0, tol, small2, c, s)) {
        rotated = 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double x = t0[r], y = b0[r];
          t0[r] = c * x - s * y;
          b0[r] = s * x + c * y;
        }
      }
      if (jacobi_rotation(aa1, bb1, gg1, tol, small2, c, s)) {
        rotated = 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double x = t1[r], y = b1[r];
          t1[r] = c * x - s * y;
          b1[r] = s * x + c * y;
        }
      }
      // ring step through LDS, every incoming column one read without a
      // select: xtop[e] (e = 0..G) is what slot 2e's top receives -- the
      // second top of group e - 1, and for e = 0 group 0's own first top,
      // which stays; xbot[e] is what slot 2e - 1's bottom receives -- the
      // first bottom of group e, and for e = G the last group's second
      // top, which turns round into the last bottom
#pragma unroll
      for (int r = 0; r < R; ++r) {
        xtop[((g + 1) * R + r) * 16 + l] = t1[r];
        xbot[(g * R + r) * 16 + l] = b0[r];
        if (g == 0) xtop[r * 16 + l] = t0[r];
        if (g == G - 1) xbot[(G * R + r) * 16 + l] = t1[r];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        t1[r] = g == 0 ? b0[r] : t0[r];
        t0[r] = xtop[(g * R + r) * 16 + l];
        b0[r] = b1[r];
        b1[r] = xbot[((g + 1) * R + r) * 16 + l];
      }
      __syncthreads();
    }
    ++sweeps;
    if (!__syncthreads_or(rotated)) {
      converged = 1;
      break;
    }
  }
  // singular values: the column norms; columns ranked by descending norm
  double nrm[4];
  {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      a0 = fma(t0[r], t0[r], a0);
      a1 = fma(b0[r], b0[r], a1);
      a2 = fma(t1[r], t1[r], a2);
      a3 = fma(b1[r], b1[r], a3);
    }
    nrm[0] = sqrt(row_sum(a0));
    nrm[1] = sqrt(row_sum(a1));
    nrm[2] = sqrt(row_sum(a2));
    nrm[3] = sqrt(row_sum(a3));
  }
  double* sv = xch;  // n norms
  if (l < 4) sv[4 * g + l] = l == 0 ? nrm[0] : l == 1 ? nrm[1] : l == 2 ? nrm[2] : nrm[3];
  __syncthreads();
  int rk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) rk[k] = rank_in_row(sv, n, nrm[k], 4 * g + k, l);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = rk[k];
    if (j >= q) continue;
    const double inv = nrm[k] > 0.0 ? 1.0 / nrm[k] : 0.0;
    if (l == 0) sig[j] = nrm[k];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = l + 16 * r;
      const double v = k == 0 ? t0[r] : k == 1 ? b0[r] : k == 2 ? t1[r] : b1[r];
      if (i < q) U[(int64_t)i * q + j] = v * inv;
    }
  }
  if (tid == 0) {
    info[0] = sweeps;
    info[1] = converged;
  }
}

// Larger matrices: columns in global memory W (column j at W + j * ldw, rows
// >= q zero), slot -> column maps in LDS.  Each wave rotates its slots'
// pairs one after another, the pair's two columns in registers in RR
// chunks of 64 rows (or streamed from memory twice when RR == 0).
template <int RR>
__device__ __forceinline__ bool glb_pair(double* __restrict__ x, double* __restrict__ y, int q,
                                         int lane, double tol, double small2) {
  if constexpr (RR > 0) {
    double xv[RR], yv[RR];
    double a = 0.0, b = 0.0, gg = 0.0;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int i = lane + 64 * r;
      xv[r] = i < q ? x[i] : 0.0;
      yv[r] = i < q ? y[i] : 0.0;
      a = fma(xv[r], xv[r], a);
      b = fma(yv[r], yv[r], b);
      gg = fma(xv[r], yv[r], gg);
    }
    a = wave_sum(a); b = wave_sum(b); gg = wave_sum(gg);
    double c, s;
    if (!jacobi_rotation(a, b, gg, tol, small2, c, s)) return false;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int i = lane + 64 * r;
      if (i < q) {
        x[i] = c * xv[r] - s * yv[r];
        y[i] = s * xv[r] + c * yv[r];
      }
    }
    return true;
  } else {
    double a = 0.0, b = 0.0, gg = 0.0;
    for (int i = lane; i < q; i += 64) {
      const double xi = x[i], yi = y[i];
      a = fma(xi, xi, a);
      b = fma(yi, yi, b);
      gg = fma(xi, yi, gg);
    }
    a = wave_sum(a); b = wave_sum(b); gg = wave_sum(gg);
    double c, s;
    if (!jacobi_rotation(a, b, gg, tol, small2, c, s)) return false;
    for (int i = lane; i < q; i += 64) {
      const double xi = x[i], yi = y[i];
      x[i] = c * xi - s * yi;
      y[i] = s * xi + c * yi;
    }
    return true;
  }
}

template <int RR>
__global__ __launch_bounds__(kGlbNT) void hj_glb_kernel(double* __restrict__ W, int q, int cols,
                                                         int n, int64_t ldw, double tol,
                                                         int max_sweeps, double* __restrict__ sig,
                                                         double* __restrict__ U, int* info) {
  // q rows (and outputs), cols columns, n >= cols even
  extern __shared__ int ids[];  // top[n/2], bot[n/2]; later n norms (doubles)
  const int ns = n / 2;
  int* top = ids;
  int* bot = ids + ns;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = kGlbNT / 64;
  for (int s = tid; s < ns; s += kGlbNT) {
    top[s] = 2 * s;
    bot[s] = 2 * s + 1;
  }
  // |M|_F^2: per-wave partial sums, then the same order in every thread
  __shared__ double fro_part[NW];
  {
    double a = 0.0;
    for (int j = wave; j < cols; j += NW) {
      const double* x = W + (int64_t)j * ldw;
      for (int i = lane; i < q; i += 64) a = fma(x[i], x[i], a);
    }
    a = wave_sum(a);
    if (lane == 0) fro_part[wave] = a;
  }
  __syncthreads();
  double fro = 0.0;
  for (int e = 0; e < NW; ++e) fro += fro_part[e];
  const double small2 = tol * tol * fro;
  int sweeps = 0, converged = 0;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    int rotated = 0;
    for (int round = 0; round < n - 1; ++round) {
      for (int s = wave; s < ns; s += NW) {
        const int a = top[s], b = bot[s];
        if (a < cols && b < cols &&
            glb_pair<RR>(W + (int64_t)a * ldw, W + (int64_t)b * ldw, q, lane, tol, small2))
          rotated = 1;
      }
      // the next round's slot map (one slot per thread: n / 2 <= kGlbNT)
      int nt = 0, nb = 0;
      if (tid < ns) {
        nt = tid == 0 ? top[0] : tid == 1 ? bot[0] : top[tid - 1];
        nb = tid == ns - 1 ? top[ns - 1] : bot[tid + 1];
      }
      __syncthreads();  // columns written and slot maps read
      if (tid < ns) {
        top[tid] = nt;
        bot[tid] = nb;
      }
      __syncthreads();
    }
    ++sweeps;
    if (!__syncthreads_or(rotated)) {
      converged = 1;
      break;
    }
  }
  // norms of the n columns (zero past cols), ranks, the q leading outputs
  double* sv = reinterpret_cast<double*>(ids);
  for (int j = wave; j < n; j += NW) {
    double a = 0.0;
    if (j < cols) {
      const double* x = W + (int64_t)j * ldw;
      for (int i = lane; i < q; i += 64) a = fma(x[i], x[i], a);
      a = wave_sum(a);
    }
    if (lane == 0) sv[j] = sqrt(a);
  }
  __syncthreads();
  for (int j = wave; j < n; j += NW) {
    const double v = sv[j];
    int cnt = 0;
    for (int i = lane; i < n; i += 64) {
      const double u = sv[i];
      cnt += (u > v || (u == v && i < j)) ? 1 : 0;
    }
    const int rk = (int)wave_sum((double)cnt);
    if (rk >= q) continue;
    const double inv = v > 0.0 ? 1.0 / v : 0.0;
    if (lane == 0) sig[rk] = v;
    const double* x = W + (int64_t)j * ldw;
    for (int i = lane; i < q; i += 64) U[(int64_t)i * q + rk] = j < cols ? x[i] * inv : 0.0;
  }
  if (tid == 0) {
    info[0] = sweeps;
    info[1] = converged;
  }
}

__global__ void copy_columns_kernel(const double* __restrict__ Mc, int q, int cols, int64_t ldc,
                                    double* __restrict__ W, int64_t ldw) {
  const int64_t total = (int64_t)cols * ldw;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / ldw, i = e % ldw;
    W[e] = i < q ? Mc[j * ldc + i] : 0.0;
  }
}

template <int R>
void launch_reg(const double* Mc, int q, int cols, int64_t ldc, double tol, double* sig,
                double* U, int* info, hipStream_t st) {
  constexpr int G = 4 * R;
  const size_t lds = sizeof(double) * (size_t)2 * (G + 1) * R * 16;
  static bool attr = false;
  if (!attr) {
    ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)hj_reg_kernel<R>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  hipLaunchKernelGGL(hj_reg_kernel<R>, dim3(1), dim3(64 * R), lds, st, Mc, q, cols, ldc, tol,
                     kMaxSweepsSvd, sig, U, info);
  ACETN_HIP_CHECK(hipGetLastError());
}

template <int RR>
void launch_glb(double* W, int q, int cols, int n, int64_t ldw, double tol, double* sig,
                double* U, int* info, hipStream_t st) {
  const size_t lds = sizeof(double) * (size_t)n;  // n <= 2048: 16 KB
  hipLaunchKernelGGL(hj_glb_kernel<RR>, dim3(1), dim3(kGlbNT), lds, st, W, q, cols, n, ldw, tol,
                     kMaxSweepsSvd, sig, U, info);
  ACETN_HIP_CHECK(hipGetLastError());
}

int64_t glb_ld(int q) { return ((int64_t)q + 63) / 64 * 64; }

}  // namespace

size_t jacobi_svd_scratch_elems(int rows, int cols) {
  const int qq = std::max(rows, cols);
  if (qq <= kRegMaxQ) return 8;
  const int n = qq + (qq & 1);
  return (size_t)n * glb_ld(rows) + 8;
}

void jacobi_svd(const double* Mc, int rows, int cols, int64_t ldc, double* sig, double* U,
                double* scratch, hipStream_t st) {
  ACETN_REQUIRE(rows >= 1 && cols >= 1, "jacobi_svd: empty matrix");
  ACETN_REQUIRE(scratch != nullptr, "jacobi_svd: scratch required");
  int* info = reinterpret_cast<int*>(scratch);
  const int q = rows, qq = std::max(rows, cols);
  // tolerance on |x.y| / (|x| |y|): qq units of roundoff, the accuracy to
  // which a dot product of two columns is computed
  const double tol = (double)qq * 1.1102230246251565e-16;
  if (qq <= kRegMaxQ) {
    const int R = (qq + 15) / 16;
    switch (R) {
      case 1: launch_reg<1>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 2: launch_reg<2>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 3: launch_reg<3>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 4: launch_reg<4>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 5: launch_reg<5>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 6: launch_reg<6>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 7: launch_reg<7>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 8: launch_reg<8>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 9: launch_reg<9>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 10: launch_reg<10>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      case 11: launch_reg<11>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
      default: launch_reg<12>(Mc, q, cols, ldc, tol, sig, U, info, st); break;
    }
    return;
  }
  const int n = qq + (qq & 1);
  ACETN_REQUIRE(n / 2 <= kGlbNT, "jacobi_svd: matrices wider than 2048 columns are not supported");
  const int64_t ldw = glb_ld(q);
  double* W = scratch + 8;
  const int blocks = (int)std::min<int64_t>(((int64_t)cols * ldw + 255) / 256, 2048);
  hipLaunchKernelGGL(copy_columns_kernel, dim3(blocks), dim3(256), 0, st, Mc, q, cols, ldc, W, ldw);
  ACETN_HIP_CHECK(hipGetLastError());
  switch ((q + 63) / 64) {
    case 1: launch_glb<1>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 2: launch_glb<2>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 3: launch_glb<3>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 4: launch_glb<4>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 5: launch_glb<5>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 6: launch_glb<6>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 7: launch_glb<7>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    case 8: launch_glb<8>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
    default: launch_glb<0>(W, q, cols, n, ldw, tol, sig, U, info, st); break;
  }
}

}  // namespace acetn
