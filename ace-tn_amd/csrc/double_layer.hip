// Double-layer absorption of one site tensor A and its conjugate into a
// block tensor X, block by block over the two outer (boundary) indices:
//
//   Y[o1,o2; k3,b3,k4,b4] = sum_{k1,b1,k2,b2,p} X[o1,o2; k1,b1,k2,b2]
//                             * A[k1,k2,k3,k4,p] * conj(A)[b1,b2,b3,b4,p]
//
// This is the pair of einsums "cuUelL,LURDP->cuelRDP" + "lurdp,cuelRDp->crRedD"
// of make_quarter_tensor (acetn/renormalization/projectors.py:54-55) and the
// pair "LURDP,alLuUx->RDPalux" + "lurdp,RDpalux->rdRDax" of renormalize_ej
// (acetn/renormalization/directional_mover.py:363-364): for a fixed boundary
// pair (o1,o2) the work is two (D^2 x D^2 x D^2) products per physical index,
// done in LDS, so the 7-index chi^2 D^4 d intermediate never reaches HBM.
// A is re-read from LDS for every block (it is loaded once per workgroup);
// X and Y are read/written exactly once.
#include "internal.h"

namespace acetn {
namespace {
constexpr int NT = 256;
constexpr int MAXPER = 16;  // D^4 / NT for D <= 8

__global__ __launch_bounds__(NT) void double_layer_kernel(DLDesc g, const double* __restrict__ X,
                                                          const double* __restrict__ A,
                                                          double* __restrict__ Y) {
  extern __shared__ double sm[];
  const int D = g.D, d = g.d, D2 = D * D, D4 = D2 * D2;
  double* Ac = sm;           // [l1][l2][l3][l4][p]
  double* Xs = Ac + D4 * d;  // [(k1,k2)][(b1,b2)]
  double* Ts = Xs + D4;      // [(k1,k2)][(b3,b4)] for one p
  const int tid = threadIdx.x;

  for (int e = tid; e < D4 * d; e += NT) {
    int r = e;
    const int p = r % d; r /= d;
    const int l4 = r % D; r /= D;
    const int l3 = r % D; r /= D;
    const int l2 = r % D;
    const int l1 = r / D;
    Ac[e] = A[l1 * g.a_leg[0] + l2 * g.a_leg[1] + l3 * g.a_leg[2] + l4 * g.a_leg[3] + p * g.a_p];
  }

  const int nblk = g.n1 * g.n2;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int o1 = blk / g.n2, o2 = blk % g.n2;
    const int64_t xb = (int64_t)o1 * g.x_o1 + (int64_t)o2 * g.x_o2;
    __syncthreads();
    for (int e = tid; e < D4; e += NT) {
      const int row = e / D2, col = e % D2;
      const int k1 = row / D, k2 = row % D, b1 = col / D, b2 = col % D;
      Xs[e] = X[xb + k1 * g.x_k1 + b1 * g.x_b1 + k2 * g.x_k2 + b2 * g.x_b2];
    }
    double yacc[MAXPER];
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) yacc[i] = 0.0;

    for (int p = 0; p < d; ++p) {
      __syncthreads();
      // Ts[(k1,k2),(b3,b4)] = sum_{(b1,b2)} Xs[(k1,k2),(b1,b2)] * A[(b1,b2),(b3,b4),p]
      for (int e = tid; e < D4; e += NT) {
        const int row = e / D2, col = e % D2;
        const double* xr = Xs + row * D2;
        double s = 0.0;
        for (int kk = 0; kk < D2; ++kk) s += xr[kk] * Ac[(kk * D2 + col) * d + p];
        Ts[e] = s;
      }
      __syncthreads();
      // Y[(k3,k4),(b3,b4)] += sum_{(k1,k2)} A[(k1,k2),(k3,k4),p] * Ts[(k1,k2),(b3,b4)]
#pragma unroll
      for (int i = 0; i < MAXPER; ++i) {
        const int e = tid + i * NT;
        if (e < D4) {
          const int row = e / D2, col = e % D2;
          double s = 0.0;
          for (int kk = 0; kk < D2; ++kk) s += Ac[(kk * D2 + row) * d + p] * Ts[kk * D2 + col];
          yacc[i] += s;
        }
      }
    }
    const int64_t yb = (int64_t)o1 * g.y_o1 + (int64_t)o2 * g.y_o2;
#pragma unroll
    for (int i = 0; i < MAXPER; ++i) {
      const int e = tid + i * NT;
      if (e < D4) {
        const int row = e / D2, col = e % D2;
        const int k3 = row / D, k4 = row % D, b3 = col / D, b4 = col % D;
        Y[yb + k3 * g.y_k3 + b3 * g.y_b3 + k4 * g.y_k4 + b4 * g.y_b4] = yacc[i];
      }
    }
  }
}
}  // namespace

void double_layer_f64(const DLDesc& g, const double* X, const double* A, double* Y,
                      hipStream_t s) {
  ACETN_REQUIRE(g.D >= 1 && g.D <= 8, "double_layer: D must be in [1, 8]");
  ACETN_REQUIRE(g.d >= 1 && g.d <= 8, "double_layer: d must be in [1, 8]");
  const int D4 = g.D * g.D * g.D * g.D;
  const size_t lds = sizeof(double) * ((size_t)D4 * g.d + 2 * (size_t)D4);
  ACETN_REQUIRE(lds <= 160 * 1024, "double_layer: LDS budget exceeded");
  const int64_t nblk = (int64_t)g.n1 * g.n2;
  if (nblk <= 0) return;
  const int grid = (int)std::min<int64_t>(nblk, 2048);
  hipLaunchKernelGGL(double_layer_kernel, dim3(grid), dim3(NT), lds, s, g, X, A, Y);
  ACETN_HIP_CHECK(hipGetLastError());
}

}  // namespace acetn
