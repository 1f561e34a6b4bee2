// Split-bf16 GEMM on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16) for
// the mixed-precision CTMRG path.
//
// Each operand element x (FP32 or FP64 in HBM) is rounded to FP32 and split
// into SPLIT bf16 pieces x = x0 + x1 (+ x2) with x_{i+1} = bf16(x - sum_{j<=i} x_j).
// The product keeps every piece pair (i, j) with i + j < SPLIT:
//   SPLIT = 2 ("bf16x3"): 3 products, ~16-bit operands (rel. error ~1e-5)
//   SPLIT = 3 ("bf16x6"): 6 products, ~24-bit operands (FP32-equivalent)
// accumulated in FP32 by the MFMA.  bf16 products are exact in FP32, so the
// only errors are the operand truncation and the FP32 accumulation.
//
// Tiles: 4 waves in a 2 x 2 grid, each owning WM x WN MFMA blocks of 16 x 16,
// BK = 32 (one MFMA k-step).  LDS holds the pieces k-contiguous ([m][k] and
// [n][k]) so every lane fetches its 8 bf16 of a fragment with one 16-byte
// read: A lane l holds A[l%16][8(l/16)..+7], B lane l holds B[8(l/16)..+7][l%16],
// and the f32 result lane l holds D[4(l/16)+r][l%16] (checked on gfx950).
// Operand addressing uses the same multi-level index maps as the FP64 GEMM,
// so einsum permutations are folded into the loads here as well.
#include "internal.h"

namespace acetn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int BK = 32, NT = 256, KPAD = 8;

__device__ __forceinline__ bool contiguous(const IdxMap& m) { return m.s0 == 1; }

template <typename T>
__device__ __forceinline__ float ld(const T* p, int64_t off) { return (float)p[off]; }

template <int SPLIT>
__device__ __forceinline__ void split(float x, __bf16 (&o)[SPLIT]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < SPLIT; ++i) {
    o[i] = (__bf16)r;
    r -= (float)o[i];
  }
}

template <int WM, int WN, int SPLIT, typename TA, typename TB, typename TC>
__global__ __launch_bounds__(NT) void gemm_split_kernel(GemmDesc g, GemmBatch bt, int ksplit,
                                                        double* __restrict__ work) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int AE = BM * BK / NT, BE = BN * BK / NT;
  constexpr int LK = BK + KPAD;
  extern __shared__ __align__(16) char smem[];
  __bf16* As = reinterpret_cast<__bf16*>(smem);       // [SPLIT][BM][LK]
  __bf16* Bs = As + SPLIT * BM * LK;                   // [SPLIT][BN][LK]
  if (g.run_if && *g.run_if == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int prob = blockIdx.z / ksplit, kpart = blockIdx.z % ksplit;
  const TA* __restrict__ A = reinterpret_cast<const TA*>(bt.A[prob]);
  const TB* __restrict__ B = reinterpret_cast<const TB*>(bt.B[prob]);
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  int kchunk = (g.K + ksplit - 1) / ksplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  const int kbeg = kpart * kchunk;
  const int kend = min(g.K, kbeg + kchunk);

  const bool a_mfast = contiguous(g.am) || !contiguous(g.ak);
  int a_ml[AE], a_kl[AE];
  int64_t a_moff[AE];
  bool a_mok[AE];
#pragma unroll
  for (int i = 0; i < AE; ++i) {
    const int e = tid + i * NT;
    if (a_mfast) { a_ml[i] = e % BM; a_kl[i] = e / BM; }
    else         { a_kl[i] = e % BK; a_ml[i] = e / BK; }
    const int m = m0 + a_ml[i];
    a_mok[i] = m < g.M;
    a_moff[i] = a_mok[i] ? imap(g.am, m) : 0;
  }
  const bool b_nfast = contiguous(g.bn) || !contiguous(g.bk);
  int b_nl[BE], b_kl[BE];
  int64_t b_noff[BE];
  bool b_nok[BE];
#pragma unroll
  for (int i = 0; i < BE; ++i) {
    const int e = tid + i * NT;
    if (b_nfast) { b_nl[i] = e % BN; b_kl[i] = e / BN; }
    else         { b_kl[i] = e % BK; b_nl[i] = e / BK; }
    const int n = n0 + b_nl[i];
    b_nok[i] = n < g.N;
    b_noff[i] = b_nok[i] ? imap(g.bn, n) : 0;
  }

  f4 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  float ra[AE], rb[BE];
  auto load_tile = [&](int kt) {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int k = kt + a_kl[i];
      ra[i] = (a_mok[i] && k < kend) ? ld(A, a_moff[i] + imap(g.ak, k)) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int k = kt + b_kl[i];
      rb[i] = (b_nok[i] && k < kend) ? ld(B, imap(g.bk, k) + b_noff[i]) : 0.f;
    }
  };

  if (kbeg < kend) load_tile(kbeg);
  for (int kt = kbeg; kt < kend; kt += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      __bf16 pc[SPLIT];
      split<SPLIT>(ra[i], pc);
#pragma unroll
      for (int s = 0; s < SPLIT; ++s) As[(s * BM + a_ml[i]) * LK + a_kl[i]] = pc[s];
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      __bf16 pc[SPLIT];
      split<SPLIT>(rb[i], pc);
#pragma unroll
      for (int s = 0; s < SPLIT; ++s) Bs[(s * BN + b_nl[i]) * LK + b_kl[i]] = pc[s];
    }
    __syncthreads();
    if (kt + BK < kend) load_tile(kt + BK);
    const int k0 = 8 * (lane >> 4);
    bf8 a[SPLIT][WM], b[SPLIT][WN];
#pragma unroll
    for (int s = 0; s < SPLIT; ++s) {
#pragma unroll
      for (int i = 0; i < WM; ++i)
        a[s][i] = *reinterpret_cast<const bf8*>(
            &As[(s * BM + wm * 16 * WM + i * 16 + (lane & 15)) * LK + k0]);
#pragma unroll
      for (int j = 0; j < WN; ++j)
        b[s][j] = *reinterpret_cast<const bf8*>(
            &Bs[(s * BN + wn * 16 * WN + j * 16 + (lane & 15)) * LK + k0]);
    }
    // smallest piece products first
#pragma unroll
    for (int ord = SPLIT - 1; ord >= 0; --ord)
#pragma unroll
      for (int sa = 0; sa <= ord; ++sa) {
        const int sb = ord - sa;
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[sb][j], acc[i][j], 0, 0, 0);
      }
  }

#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 16 * WM + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 16 * WN + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const double v = (double)acc[i][j][r];
        if (ksplit == 1) {
          TC* C = reinterpret_cast<TC*>(bt.C[prob]);
          const int64_t off = imap(g.cm, m) + imap(g.cn, n);
          C[off] = (TC)((g.beta != 0.0) ? g.alpha * v + g.beta * (double)C[off] : g.alpha * v);
        } else {
          work[(((int64_t)prob * ksplit + kpart) * g.M + m) * g.N + n] = v;
        }
      }
}

template <typename TC>
__global__ void splitk_reduce_typed(GemmDesc g, GemmBatch bt, int ksplit,
                                    const double* __restrict__ work) {
  if (g.run_if && *g.run_if == 0) return;
  const int64_t mn = (int64_t)g.M * g.N;
  const int prob = blockIdx.y;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < ksplit; ++p) s += work[((int64_t)prob * ksplit + p) * mn + e];
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    TC* C = reinterpret_cast<TC*>(bt.C[prob]);
    const int64_t off = imap(g.cm, m) + imap(g.cn, n);
    C[off] = (TC)((g.beta != 0.0) ? g.alpha * s + g.beta * (double)C[off] : g.alpha * s);
  }
}

struct Tile {
  int wm, wn;
};
Tile choose_tile(const GemmDesc& g) {
  if (g.N > 96 && g.N <= 160) return g.M >= 2048 ? Tile{2, 5} : Tile{1, 5};
  if (g.M >= 1024 && g.N >= 1024) return Tile{4, 4};
  return Tile{2, 2};
}

int choose_ksplit(const GemmDesc& g, int count) {
  const Tile t = choose_tile(g);
  const int BM = 32 * t.wm, BN = 32 * t.wn;
  const int64_t tiles = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN) * count;
  if (tiles >= 512 || g.K < 8 * BK * 4) return 1;
  int ks = (int)((1024 + tiles - 1) / tiles);
  ks = std::min(ks, g.K / (BK * 4));
  ks = std::min(ks, 64);
  return std::max(ks, 1);
}

template <int WM, int WN, int SPLIT, typename TA, typename TB, typename TC>
void launch(const GemmDesc& g, const GemmBatch& b, int ks, double* work, hipStream_t s) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const size_t lds = sizeof(__bf16) * (size_t)SPLIT * (BM + BN) * (BK + KPAD);
  static bool attr = false;
  if (!attr) {
    ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)gemm_split_kernel<WM, WN, SPLIT, TA, TB, TC>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr = true;
  }
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, b.count * ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_split: grid too large");
  hipEvent_t e0 = kernel_timing_begin(s);
  hipLaunchKernelGGL((gemm_split_kernel<WM, WN, SPLIT, TA, TB, TC>), grid, dim3(NT), lds, s, g, b,
                     ks, work);
  ACETN_HIP_CHECK(hipGetLastError());
  kernel_timing_end(e0, s, kKGemmSplit, 2.0 * g.M * g.N * (double)g.K * b.count);
  if (ks > 1) {
    const int64_t mn = (int64_t)g.M * g.N;
    const int blocks = (int)std::min<int64_t>((mn + 255) / 256, 2048);
    hipLaunchKernelGGL(splitk_reduce_typed<TC>, dim3(blocks, b.count), dim3(256), 0, s, g, b, ks,
                       work);
    ACETN_HIP_CHECK(hipGetLastError());
  }
}

template <int SPLIT, typename TA, typename TB, typename TC>
void dispatch(const GemmDesc& g, const GemmBatch& b, int ks, double* work, hipStream_t s) {
  const Tile t = choose_tile(g);
  if (t.wm == 2 && t.wn == 5) launch<2, 5, SPLIT, TA, TB, TC>(g, b, ks, work, s);
  else if (t.wm == 1 && t.wn == 5) launch<1, 5, SPLIT, TA, TB, TC>(g, b, ks, work, s);
  else if (t.wm == 4 && t.wn == 4) launch<4, 4, SPLIT, TA, TB, TC>(g, b, ks, work, s);
  else launch<2, 2, SPLIT, TA, TB, TC>(g, b, ks, work, s);
}
}  // namespace

size_t gemm_split_work_elems(const GemmDesc& g, int count) {
  const int ks = choose_ksplit(g, count);
  return ks > 1 ? (size_t)ks * count * g.M * g.N : 0;
}

void gemm_split(const GemmDesc& g, const GemmBatch& b, int split, DType ta, DType tb, DType tc,
                double* work, size_t work_elems, hipStream_t s) {
  ACETN_REQUIRE(b.count >= 1 && b.count <= kMaxBatch, "gemm_split: bad batch count");
  ACETN_REQUIRE(split == 2 || split == 3, "gemm_split: split must be 2 or 3");
  if (g.M <= 0 || g.N <= 0) return;
  int ks = choose_ksplit(g, b.count);
  if (ks > 1 && (work == nullptr || work_elems < (size_t)ks * b.count * g.M * g.N)) ks = 1;
  const int key = (ta == DType::F32) * 4 + (tb == DType::F32) * 2 + (tc == DType::F32);
#define ACETN_SPLIT_CASE(K, A_, B_, C_)                                          \
  case K:                                                                        \
    if (split == 2) dispatch<2, A_, B_, C_>(g, b, ks, work, s);                  \
    else dispatch<3, A_, B_, C_>(g, b, ks, work, s);                             \
    break;
  switch (key) {
    ACETN_SPLIT_CASE(0, double, double, double)
    default: ACETN_REQUIRE(false, "gemm_split: unsupported operand types");
  }
#undef ACETN_SPLIT_CASE
}

}  // namespace acetn
