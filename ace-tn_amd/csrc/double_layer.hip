// Double-layer absorption of one site tensor A and its conjugate into a
// block tensor X, block by block over the two outer (boundary) indices:
//
//   Y[o1,o2; k3,b3,k4,b4] = sum_{k1,b1,k2,b2,p} X[o1,o2; k1,b1,k2,b2]
//                             * A[k1,k2,k3,k4,p] * conj(A)[b1,b2,b3,b4,p]
//
// This is the pair of einsums "cuUelL,LURDP->cuelRDP" + "lurdp,cuelRDp->crRedD"
// of make_quarter_tensor (acetn/renormalization/projectors.py:54-55) and the
// pair "LURDP,alLuUx->RDPalux" + "lurdp,RDpalux->rdRDax" of renormalize_ej
// (acetn/renormalization/directional_mover.py:363-364).
//
// For a fixed boundary pair the work is two (D^2 x D^2 x D^2) products per
// physical index p:  T = X . Abar_p  and  Y += A_p^T . T.  NB blocks are
// stacked so that the row count is a multiple of the 16-row MFMA tile
// (NB = 4 at D = 6 gives exactly 144 rows), and both products run on
// v_mfma_f64_16x16x4_f64 from LDS.  Y stays in registers across p; the
// 7-index chi^2 D^4 d intermediate never reaches HBM, X and Y are read and
// written once.
#include "internal.h"

namespace acetn {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Compute type of the double-layer products: FP64 (v_mfma_f64_16x16x4f64)
// or FP32 (v_mfma_f32_16x16x4f32, the mixed-precision path).  Operands of
// both: A(16x4) lane l holds A[l%16][l/16], B(4x16) lane l holds
// B[l/16][l%16].  Results: the FP64 16x16 tile holds D[l/16 + 4r][l%16] in
// element r of lane l, the FP32 one D[4(l/16) + r][l%16].
template <typename T> struct DLMath;
template <> struct DLMath<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct DLMath<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

namespace {
// 8 waves per workgroup: the LDS footprint (A, the stacked X tile and the
// intermediate T) allows one workgroup per CU at D = 6, so the waves of that
// one workgroup are what hides latency.
constexpr int NT = 512;
constexpr int NW = NT / 64;

template <int D>
struct DLShape {
  static constexpr int D2 = D * D;
  static constexpr int NB = (D == 6) ? 4 : ((64 / D2) > 0 ? (64 / D2) : 1);
  static constexpr int ROWS = NB * D2;            // stacked rows
  static constexpr int MP = (ROWS + 15) / 16 * 16; // padded rows (step c M, step d N)
  static constexpr int KP = (D2 + 3) / 4 * 4;      // padded contraction length
  static constexpr int XLD = KP + 1;               // odd LDS row stride of the X tile (no bank conflicts)
  static constexpr int NP = (D2 + 15) / 16 * 16;   // padded (b3,b4) columns / step d rows
  static constexpr int TC = (MP / 16) * (NP / 16); // tiles per product
  static constexpr int TPW = (TC + NW - 1) / NW;   // tiles per wave
  static constexpr int XPT = (MP * KP + NT - 1) / NT; // X elements per thread
};

template <int D, typename T>
__global__ __launch_bounds__(NT) void dl_mfma_kernel(DLDesc g, const double* __restrict__ X,
                                                     const double* __restrict__ A,
                                                     double* __restrict__ Y) {
  using S = DLShape<D>;
  using M = DLMath<T>;
  typedef typename M::acc_t acc_t;
  extern __shared__ __align__(16) char smem[];
  T* sm = reinterpret_cast<T*>(smem);
  const int d = g.d;
  T* Ac = sm;                               // [p][KP][NP]: A[(l1,l2),(l3,l4),p], zero padded
  T* Xs = Ac + (size_t)d * S::KP * S::NP;   // [MP][XLD]: stacked blocks, rows (blk,k1,k2)
  T* Ts = Xs + S::MP * S::XLD;              // [MP][NP]: rows (blk,k1,k2), cols (b3,b4)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int e = tid; e < d * S::KP * S::NP; e += NT) {
    const int p = e / (S::KP * S::NP), r = e % (S::KP * S::NP);
    const int kk = r / S::NP, nn = r % S::NP;
    double v = 0.0;
    if (kk < S::D2 && nn < S::D2) {
      const int l1 = kk / D, l2 = kk % D, l3 = nn / D, l4 = nn % D;
      v = A[l1 * g.a_leg[0] + l2 * g.a_leg[1] + l3 * g.a_leg[2] + l4 * g.a_leg[3] + p * g.a_p];
    }
    Ac[e] = (T)v;
  }

  const int nblk = g.n1 * g.n2;
  const int ngroups = (nblk + S::NB - 1) / S::NB;
  // X tile of a group, element e = (row, col) of the stacked [MP][KP] tile;
  // loaded into registers one group ahead of its use.  The in-block part of
  // each element's offset is fixed, only the block base changes per group.
  T xr[S::XPT];
  int x_loff[S::XPT];
  int x_blk[S::XPT];
#pragma unroll
  for (int t = 0; t < S::XPT; ++t) {
    const int e = tid + t * NT;
    const int row = e / S::KP, col = e % S::KP;
    const int blk = row / S::D2, kk = row % S::D2;
    const bool ok = e < S::MP * S::KP && row < S::ROWS && col < S::D2;
    const int k1 = kk / D, k2 = kk % D, b1 = col / D, b2 = col % D;
    x_loff[t] = ok ? (int)(k1 * g.x_k1 + b1 * g.x_b1 + k2 * g.x_k2 + b2 * g.x_b2) : 0;
    x_blk[t] = ok ? blk : -1;
  }
  auto load_x = [&](int grp) {
    const int b0 = grp * S::NB;
    int64_t base[S::NB];
#pragma unroll
    for (int bb = 0; bb < S::NB; ++bb) {
      const int b = b0 + bb;
      base[bb] = (b < nblk) ? (int64_t)(b / g.n2) * g.x_o1 + (int64_t)(b % g.n2) * g.x_o2 : -1;
    }
#pragma unroll
    for (int t = 0; t < S::XPT; ++t) {
      int64_t bo = -1;
#pragma unroll
      for (int bb = 0; bb < S::NB; ++bb)
        if (x_blk[t] == bb) bo = base[bb];
      if (bo < 0)
        xr[t] = (T)0;
      else if (g.x_f32)
        xr[t] = (T)reinterpret_cast<const float*>(X)[bo + x_loff[t]];
      else
        xr[t] = (T)X[bo + x_loff[t]];
    }
  };
  if ((int)blockIdx.x < ngroups) load_x(blockIdx.x);
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int b0 = grp * S::NB;
    __syncthreads();
    // ---- stacked X tile (from registers), then start loading the next group
#pragma unroll
    for (int t = 0; t < S::XPT; ++t) {
      const int e = tid + t * NT;
      if (e < S::MP * S::KP) Xs[(e / S::KP) * S::XLD + e % S::KP] = (T)xr[t];
    }
    if (grp + (int)gridDim.x < ngroups) load_x(grp + gridDim.x);
    acc_t yacc[S::TPW];
#pragma unroll
    for (int t = 0; t < S::TPW; ++t) yacc[t] = acc_t{0, 0, 0, 0};

    // closed: sum_p A_p^T (X Abar_p);  open: every (p, pb) pair written out
    for (int pb = 0; pb < d; ++pb) {
      const T* Apb = Ac + (size_t)pb * S::KP * S::NP;
      __syncthreads();
      // ---- step c: Ts[MP x NP] = Xs[MP x KP] . Abar_pb[KP x NP]
#pragma unroll
      for (int t = 0; t < S::TPW; ++t) {
        const int tile = t * NW + wave;  // round-robin over the waves (even SIMD load)
        if (tile < S::TC) {
          const int tm = tile / (S::NP / 16), tn = tile % (S::NP / 16);
          acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll 3
          for (int ks = 0; ks < S::KP; ks += 4) {
            const T a = Xs[(tm * 16 + (lane & 15)) * S::XLD + ks + (lane >> 4)];
            const T b = Apb[(ks + (lane >> 4)) * S::NP + tn * 16 + (lane & 15)];
            acc = M::mfma(a, b, acc);
          }
          for (int r = 0; r < 4; ++r)
            Ts[(tm * 16 + M::row(lane, r)) * S::NP + tn * 16 + (lane & 15)] = acc[r];
        }
      }
      __syncthreads();
      const int p_lo = g.open_p ? 0 : pb, p_hi = g.open_p ? d : pb + 1;
      for (int p = p_lo; p < p_hi; ++p) {
        const T* Ap = Ac + (size_t)p * S::KP * S::NP;
        if (g.open_p) {
#pragma unroll
          for (int t = 0; t < S::TPW; ++t) yacc[t] = acc_t{0, 0, 0, 0};
        }
        // ---- step d: Y[(k3,k4), (blk,b3,b4)] += A_p^T[(k3,k4),(k1,k2)] . T[(blk,k1,k2),(b3,b4)]
#pragma unroll
        for (int t = 0; t < S::TPW; ++t) {
          const int tile = t * NW + wave;  // round-robin over the waves (even SIMD load)
          if (tile < S::TC) {
            const int tm = tile / (S::MP / 16), tn = tile % (S::MP / 16);
            const int nn = tn * 16 + (lane & 15);
            const int blk = nn / S::D2, b34 = nn % S::D2;
#pragma unroll 3
            for (int ks = 0; ks < S::KP; ks += 4) {
              const int kk = ks + (lane >> 4);
              const T a = Ap[kk * S::NP + tm * 16 + (lane & 15)];
              const T b = (kk < S::D2 && nn < S::ROWS) ? Ts[(blk * S::D2 + kk) * S::NP + b34] : T(0);
              yacc[t] = M::mfma(a, b, yacc[t]);
            }
          }
        }
        if (g.open_p || (p == d - 1)) {
          const int64_t poff = g.open_p ? (int64_t)p * g.y_p + (int64_t)pb * g.y_pb : 0;
          // ---- write Y (after the last p when closed, per (p, pb) when open)
#pragma unroll
          for (int t = 0; t < S::TPW; ++t) {
            const int tile = t * NW + wave;  // round-robin over the waves (even SIMD load)
            if (tile < S::TC) {
              const int tm = tile / (S::MP / 16), tn = tile % (S::MP / 16);
              const int nn = tn * 16 + (lane & 15);
              const int blk = nn / S::D2, b34 = nn % S::D2;
              if (nn < S::ROWS && b0 + blk < nblk) {
                const int o1 = (b0 + blk) / g.n2, o2 = (b0 + blk) % g.n2;
                const int64_t yb = (int64_t)o1 * g.y_o1 + (int64_t)o2 * g.y_o2 + poff;
                const int b3 = b34 / D, b4 = b34 % D;
                for (int r = 0; r < 4; ++r) {
                  const int m = tm * 16 + M::row(lane, r);
                  if (m < S::D2) {
                    const int k3 = m / D, k4 = m % D;
                    const int64_t yo = yb + k3 * g.y_k3 + b3 * g.y_b3 + k4 * g.y_k4 + b4 * g.y_b4;
                    if (g.y_f32)
                      reinterpret_cast<float*>(Y)[yo] = (float)yacc[t][r];
                    else
                      Y[yo] = (double)yacc[t][r];
                  }
                }
              }
            }
          }
        }
      }
    }
  }
}

template <int D, typename T>
void launch_dl(const DLDesc& g, const double* X, const double* A, double* Y, hipStream_t s) {
  using S = DLShape<D>;
  static_assert(S::TC == (S::NP / 16) * (S::MP / 16), "tile counts of the two products differ");
  const size_t lds = sizeof(T) * ((size_t)g.d * S::KP * S::NP + (size_t)S::MP * S::XLD +
                                  (size_t)S::MP * S::NP);
  ACETN_REQUIRE(lds <= 160 * 1024, "double_layer: LDS budget exceeded");
  ACETN_REQUIRE((int64_t)(D - 1) * (g.x_k1 + g.x_b1 + g.x_k2 + g.x_b2) < (int64_t(1) << 31),
                "double_layer: block strides exceed 32-bit offsets");
  static bool attr = false;
  if (!attr) {
    ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)dl_mfma_kernel<D, T>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  const int64_t nblk = (int64_t)g.n1 * g.n2;
  const int64_t ngroups = (nblk + S::NB - 1) / S::NB;
  // one round of resident workgroups (LDS-limited, 256 CUs), each looping
  // over its share of the groups with the next group's X prefetched: the A
  // tensor is staged once per workgroup instead of once per group or two
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024) / lds));
  const int grid = (int)std::min<int64_t>(ngroups, 256 * per_cu);
  const double d6 = (double)S::D2 * S::D2 * S::D2;
  const double flops = (double)nblk * (2.0 * g.d * d6 + 2.0 * (g.open_p ? g.d * g.d : g.d) * d6);
  hipEvent_t e0 = kernel_timing_begin(s);
  hipLaunchKernelGGL((dl_mfma_kernel<D, T>), dim3(grid), dim3(NT), lds, s, g, X, A, Y);
  ACETN_HIP_CHECK(hipGetLastError());
  // algorithmic bytes: the X blocks read once, the Y blocks written once
  const double yblk = g.open_p ? (double)g.d * g.d : 1.0;
  const double xb = g.x_f32 ? 4.0 : 8.0, yb = g.y_f32 ? 4.0 : 8.0;
  kernel_timing_end(e0, s, kKDoubleLayer, flops, (double)nblk * S::D2 * S::D2 * (xb + yb * yblk));
}
}  // namespace

namespace {
template <typename T>
void double_layer_t(const DLDesc& g, const double* X, const double* A, double* Y, hipStream_t s) {
  ACETN_REQUIRE(g.d >= 1 && g.d <= 4, "double_layer: d must be in [1, 4]");
  if ((int64_t)g.n1 * g.n2 <= 0) return;
  switch (g.D) {
    case 1: launch_dl<1, T>(g, X, A, Y, s); break;
    case 2: launch_dl<2, T>(g, X, A, Y, s); break;
    case 3: launch_dl<3, T>(g, X, A, Y, s); break;
    case 4: launch_dl<4, T>(g, X, A, Y, s); break;
    case 5: launch_dl<5, T>(g, X, A, Y, s); break;
    case 6: launch_dl<6, T>(g, X, A, Y, s); break;
    case 7: launch_dl<7, T>(g, X, A, Y, s); break;
    case 8: launch_dl<8, T>(g, X, A, Y, s); break;
    default: ACETN_REQUIRE(false, "double_layer: D must be in [1, 8]");
  }
}
}  // namespace

void double_layer_f64(const DLDesc& g, const double* X, const double* A, double* Y,
                      hipStream_t s) {
  ACETN_REQUIRE(!g.x_f32 && !g.y_f32, "double_layer_f64: FP64 X and Y only");
  double_layer_t<double>(g, X, A, Y, s);
}

void double_layer_f32(const DLDesc& g, const double* X, const double* A, double* Y,
                      hipStream_t s) {
  double_layer_t<float>(g, X, A, Y, s);
}

}  // namespace acetn
