// Split-bf16 GEMM on the CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16) for
// the mixed-precision CTMRG path.
//
// Each operand element x (FP64 in HBM) is rounded to FP32 and split into
// SPLIT bf16 pieces x = x0 + x1 (+ x2) with x_{i+1} = bf16(x - sum_{j<=i} x_j).
// The product keeps every piece pair (i, j) with i + j < SPLIT:
//   SPLIT = 2 ("bf16x3"): 3 products, ~16-bit operands (rel. error ~1e-5)
//   SPLIT = 3 ("bf16x6"): 6 products, ~24-bit operands (FP32-equivalent)
// accumulated in FP32 by the MFMA.  bf16 products are exact in FP32, so the
// only errors are the operand truncation and the FP32 accumulation.
//
// Two kernels:
//  * products with an enlarged corner as A (the range finder): the corner is
//    split once into planes [plane][m][k] of both orientations
//    (split_transpose), B is split once per call into planes [plane][n][k]
//    (k padded to 32 with zeros), and gemm_split_kernel (4 waves in a 2 x 2
//    grid, 64 x 160 tiles) moves both with 16-byte loads straight into LDS;
//  * every other product: gemm_general_kernel (8 waves, 4 x 2, 128 x 160
//    tiles) reads both operands as FP64 through the same multi-level index
//    maps as the FP64 GEMM (einsum permutations folded into the loads) and
//    splits them in registers on the way to LDS.
// LDS holds the pieces k-contiguous ([m][k] and [n][k]) so every lane fetches
// its 8 bf16 of a fragment with one 16-byte read: A lane l holds
// A[l%16][8(l/16)..+7], B lane l holds B[8(l/16)..+7][l%16], and the f32 result
// lane l holds D[4(l/16)+r][l%16] (checked on gfx950).  BK = 32 (one MFMA
// k-step); long K is split into partial sums reduced in a fixed order.
#include "internal.h"

namespace acetn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int BK = 32, NT = 256, KPAD = 8;
constexpr int kCUs = 256;  // MI355X compute units

__device__ __forceinline__ bool contiguous(const IdxMap& m) { return m.s0 == 1; }

// C = alpha v + beta C at one element, C in FP64 or (g.c_f32) FP32
__device__ __forceinline__ void store_c(const GemmDesc& g, double* C, int64_t off, double v) {
  if (g.c_f32) {
    float* Cf = reinterpret_cast<float*>(C);
    Cf[off] = (float)((g.beta != 0.0) ? g.alpha * v + g.beta * (double)Cf[off] : g.alpha * v);
  } else {
    C[off] = (g.beta != 0.0) ? g.alpha * v + g.beta * C[off] : g.alpha * v;
  }
}

template <int SPLIT>
__device__ __forceinline__ void split(float x, __bf16 (&o)[SPLIT]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < SPLIT; ++i) {
    o[i] = (__bf16)r;
    r -= (float)o[i];
  }
}

// planes[s][n][k] (k padded to Kp) of B(k, n) = B[bk(k) + bn(n)]
template <int SPLIT>
__global__ void split_planes_kernel(const double* __restrict__ B, IdxMap bk, IdxMap bn, int K,
                                    int N, int Kp, __bf16* __restrict__ planes, int b_f32) {
  const int64_t total = (int64_t)N * Kp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int n = (int)(e / Kp), k = (int)(e % Kp);
    float x = 0.f;
    if (k < K) {
      const int64_t o = imap(bk, k) + imap(bn, n);
      x = b_f32 ? reinterpret_cast<const float*>(B)[o] : (float)B[o];
    }
    __bf16 pc[SPLIT];
    split<SPLIT>(x, pc);
#pragma unroll
    for (int s = 0; s < SPLIT; ++s) planes[(int64_t)s * total + e] = pc[s];
  }
}

// The same planes when n is the unit-stride index of X (k strided): 64 x 64
// tiles transposed through LDS, so reads run along n and writes along k.
template <int SPLIT>
__global__ __launch_bounds__(256) void split_planes_t_kernel(const double* __restrict__ B, IdxMap bk,
                                                             IdxMap bn, int K, int N, int Kp,
                                                             __bf16* __restrict__ planes, int b_f32) {
  __shared__ float tile[64][65];
  const int k0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int64_t total = (int64_t)N * Kp;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int kk = e / 64, nn = e % 64;
    const int k = k0 + kk, n = n0 + nn;
    float x = 0.f;
    if (k < K && n < N) {
      const int64_t o = imap(bk, k) + imap(bn, n);
      x = b_f32 ? reinterpret_cast<const float*>(B)[o] : (float)B[o];
    }
    tile[kk][nn] = x;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int nn = e / 64, kk = e % 64;
    const int k = k0 + kk, n = n0 + nn;
    if (n < N && k < Kp) {
      __bf16 pc[SPLIT];
      split<SPLIT>(tile[kk][nn], pc);
#pragma unroll
      for (int s = 0; s < SPLIT; ++s) planes[(int64_t)s * total + (int64_t)n * Kp + k] = pc[s];
    }
  }
}

// planes[s][n][k] of X(k, n) = X[km(k) + nm(n)], the kernel chosen by which
// index is unit-stride
template <int SPLIT>
void split_into_planes(const double* X, const IdxMap& km, const IdxMap& nm, int K, int N, int Kp,
                       __bf16* planes, hipStream_t s, bool x_f32 = false) {
  if (nm.s0 == 1 && km.s0 != 1) {
    dim3 grid((Kp + 63) / 64, (N + 63) / 64);
    ACETN_REQUIRE(grid.y <= 65535, "split_planes: too many rows");
    hipLaunchKernelGGL(split_planes_t_kernel<SPLIT>, grid, dim3(256), 0, s, X, km, nm, K, N, Kp,
                       planes, x_f32 ? 1 : 0);
  } else {
    const int64_t total = (int64_t)N * Kp;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(split_planes_kernel<SPLIT>, dim3(blocks), dim3(256), 0, s, X, km, nm, K, N,
                       Kp, planes, x_f32 ? 1 : 0);
  }
  ACETN_HIP_CHECK(hipGetLastError());
}

// X (R x Cc, row-major, leading dimension ld) -> planes of both orientations:
// PN[s][r][c] (Cp columns, zero padded) and PT[s][c][r] (Rp columns).  One
// 64 x 64 tile per block, transposed through LDS.
constexpr int TT = 64;
__global__ __launch_bounds__(256) void split_transpose_kernel(const double* __restrict__ X, int R,
                                                              int Cc, int64_t ld, int Rp, int Cp,
                                                              __bf16* __restrict__ PN,
                                                              __bf16* __restrict__ PT, int x_f32) {
  __shared__ __bf16 tile[3][TT][TT + 8];
  const int r0 = blockIdx.y * TT, c0 = blockIdx.x * TT;
  const int64_t pn = (int64_t)R * Cp, pt = (int64_t)Cc * Rp;
  for (int e = threadIdx.x; e < TT * TT; e += 256) {
    const int rr = e / TT, cc = e % TT;
    const int r = r0 + rr, c = c0 + cc;
    float x = 0.f;
    if (r < R && c < Cc)
      x = x_f32 ? reinterpret_cast<const float*>(X)[(int64_t)r * ld + c] : (float)X[(int64_t)r * ld + c];
    __bf16 pc[3];
    split<3>(x, pc);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      tile[s][rr][cc] = pc[s];
      if (r < R && c < Cp) PN[s * pn + (int64_t)r * Cp + c] = pc[s];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TT * TT; e += 256) {
    const int cc = e / TT, rr = e % TT;
    const int r = r0 + rr, c = c0 + cc;
    if (c < Cc && r < Rp) {
#pragma unroll
      for (int s = 0; s < 3; ++s) PT[s * pt + (int64_t)c * Rp + r] = tile[s][rr][cc];
    }
  }
}

// Planar product: A and B both arrive as bf16 planes, A [s][m][Kp] and B
// [s][n][Kp] (k padded to Kp with zeros), workgroup tile 32*WM x 32*WN,
// 4 waves in a 2 x 2 grid.  Every lane moves whole 16-byte octets; rows past
// the matrix edge are clamped to the last row (their results are never
// stored), so the loads carry no predicates.  The next k-tile is in flight
// in registers while the MFMAs run on the LDS tile (a second A tile in
// flight measured no faster: the kernel is not load-latency bound).  BKT = 64
// (two-piece products, tuner variant 1) stages two MFMA k-steps per barrier
// pair: 44.0 against 47.5 us on the 5184 x 146 x 5184 range-finder product.
// Split-K partials
// are FP32, the accumulator type, so they lose nothing.  Offsets are 32-bit
// (the host checks that every plane set stays below 2^31 elements).
template <int WM, int WN, int SPLIT, int BKT = BK>
__global__ __launch_bounds__(NT, 2) void gemm_split_kernel(GemmDesc g, const __bf16* __restrict__ Ap,
                                                        int64_t aplane,
                                                        const __bf16* __restrict__ Bp, int Kp,
                                                        double* __restrict__ C, int ksplit,
                                                        float* __restrict__ work) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  constexpr int LK = BKT + KPAD;
  constexpr int BOCT = SPLIT * BN * (BKT / 8);  // 16-byte octets of B per k-tile
  constexpr int BV = (BOCT + NT - 1) / NT;
  constexpr int AOCT = SPLIT * BM * (BKT / 8);  // same for A
  constexpr int AV = (AOCT + NT - 1) / NT;
  extern __shared__ __align__(16) char smem[];
  __bf16* As = reinterpret_cast<__bf16*>(smem);  // [SPLIT][BM][LK]
  __bf16* Bs = As + SPLIT * BM * LK;              // [SPLIT][BN][LK]
  if (g.run_if && *g.run_if == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int kpart = blockIdx.z;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  int kchunk = (g.K + ksplit - 1) / ksplit;
  kchunk = (kchunk + BKT - 1) / BKT * BKT;
  const int kbeg = kpart * kchunk;
  const int kend = min(g.K, kbeg + kchunk);
  const int64_t bplane = (int64_t)g.N * Kp;

  // octet o -> (piece s, row, k-octet q); octets past the tile (o >= AOCT)
  // re-read the last one and are not stored
  int pa[AV], sa_off[AV];
#pragma unroll
  for (int i = 0; i < AV; ++i) {
    const int o = min(tid + i * NT, AOCT - 1);
    const int q = o % (BKT / 8), ml = (o / (BKT / 8)) % BM, s = o / (BKT / 8 * BM);
    pa[i] = s * (int)aplane + min(m0 + ml, g.M - 1) * Kp + 8 * q;
    sa_off[i] = (s * BM + ml) * LK + 8 * q;
  }
  int pb[BV], sb_off[BV], b_k8[BV];
#pragma unroll
  for (int i = 0; i < BV; ++i) {
    const int o = min(tid + i * NT, BOCT - 1);
    const int q = o % (BKT / 8), nl = (o / (BKT / 8)) % BN, s = o / (BKT / 8 * BN);
    pb[i] = s * (int)bplane + min(n0 + nl, g.N - 1) * Kp + 8 * q;
    sb_off[i] = (s * BN + nl) * LK + 8 * q;
    b_k8[i] = 8 * q;
  }

  f4 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  auto load_a = [&](u4 (&ra)[AV], int kt) {
#pragma unroll
    for (int i = 0; i < AV; ++i) ra[i] = *reinterpret_cast<const u4*>(Ap + (pa[i] + kt));
  };
  auto load_b = [&](u4 (&rb)[BV], int kt) {
#pragma unroll
    for (int i = 0; i < BV; ++i) rb[i] = *reinterpret_cast<const u4*>(Bp + (pb[i] + kt));
  };
  // registers -> LDS; inside K, k >= kend belongs to another split-K part and
  // is zeroed in B only (one zero factor per product suffices; both planes
  // are zero beyond K)
  auto stage = [&](const u4 (&ra)[AV], const u4 (&rb)[BV], int kt) {
#pragma unroll
    for (int i = 0; i < AV; ++i)
      if (AOCT % NT == 0 || tid + i * NT < AOCT)
        *reinterpret_cast<u4*>(&As[sa_off[i]]) = ra[i];
#pragma unroll
    for (int i = 0; i < BV; ++i) {
      if (BOCT % NT == 0 || tid + i * NT < BOCT) {
        u4 v = rb[i];
        const int k0 = kt + b_k8[i];
        if (k0 + 8 > kend) {
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int ka = k0 + 2 * t;
            if (ka >= kend) v[t] = 0u;
            else if (ka + 1 >= kend) v[t] &= 0x0000FFFFu;
          }
        }
        *reinterpret_cast<u4*>(&Bs[sb_off[i]]) = v;
      }
    }
  };
  // smallest piece products first; the A fragments of every piece stay in
  // registers, B fragments are re-read from LDS per product (fewer live
  // registers, more resident waves)
  auto compute = [&](int kk) {
    const int k0 = kk + 8 * (lane >> 4);
    bf8 a[SPLIT][WM];
#pragma unroll
    for (int s = 0; s < SPLIT; ++s)
#pragma unroll
      for (int i = 0; i < WM; ++i)
        a[s][i] = *reinterpret_cast<const bf8*>(
            &As[(s * BM + wm * 16 * WM + i * 16 + (lane & 15)) * LK + k0]);
#pragma unroll
    for (int ord = SPLIT - 1; ord >= 0; --ord)
#pragma unroll
      for (int sa = 0; sa <= ord; ++sa) {
        const int sb = ord - sa;
        bf8 b[WN];
#pragma unroll
        for (int j = 0; j < WN; ++j)
          b[j] = *reinterpret_cast<const bf8*>(
              &Bs[(sb * BN + wn * 16 * WN + j * 16 + (lane & 15)) * LK + k0]);
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[j], acc[i][j], 0, 0, 0);
      }
  };

  u4 ra0[AV], rb[BV];
  if (kbeg < kend) {
    load_b(rb, kbeg);
    load_a(ra0, kbeg);
  }
  for (int kt = kbeg; kt < kend; kt += BKT) {
    __syncthreads();
    stage(ra0, rb, kt);
    __syncthreads();
    if (kt + BKT < kend) {
      load_b(rb, kt + BKT);
      load_a(ra0, kt + BKT);
    }
    // one MFMA k-step per 32 of the tile, in k order (the accumulation order
    // of 32-deep tiles)
#pragma unroll
    for (int kk = 0; kk < BKT; kk += BK) compute(kk);
  }

  // one loop per destination, the choice made once (uniform per launch)
  auto each = [&](auto&& put) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm * 16 * WM + i * 16 + 4 * (lane >> 4) + r;
          const int n = n0 + wn * 16 * WN + j * 16 + (lane & 15);
          if (m < g.M && n < g.N) put(m, n, acc[i][j][r]);
        }
  };
  if (ksplit != 1) {
    each([&](int m, int n, float v) { work[((int64_t)kpart * g.M + m) * g.N + n] = v; });
  } else if (g.c_f32) {
    float* Cf = reinterpret_cast<float*>(C);
    each([&](int m, int n, float v) {
      const int64_t off = imap(g.cm, m) + imap(g.cn, n);
      Cf[off] = (float)((g.beta != 0.0) ? g.alpha * v + g.beta * (double)Cf[off] : g.alpha * v);
    });
  } else {
    each([&](int m, int n, float v) {
      const int64_t off = imap(g.cm, m) + imap(g.cn, n);
      C[off] = (g.beta != 0.0) ? g.alpha * (double)v + g.beta * C[off] : g.alpha * (double)v;
    });
  }
}

// General products (operands FP64 through their index maps, split in
// registers on the way to LDS): 8 waves, 4 along M x 2 along N, each owning
// WM x WN blocks of 16 x 16; the 128 x 160 tile keeps two workgroups per CU.
// Split-K partials are FP32 (the accumulator type), reduced in a fixed order.
constexpr int NT8 = 512;

template <int SPLIT, int WM, int WN>
__global__ __launch_bounds__(NT8) void gemm_general_kernel(GemmDesc g, const double* __restrict__ A,
                                                           const double* __restrict__ B,
                                                           double* __restrict__ C, int ksplit,
                                                           float* __restrict__ work) {
  constexpr int BM = 64 * WM, BN = 32 * WN;
  constexpr int LK = BK + KPAD;
  constexpr int AE = BM * BK / NT8;               // A elements per thread and k-tile
  constexpr int BE = (BN * BK + NT8 - 1) / NT8;   // B elements per thread and k-tile
  extern __shared__ __align__(16) char smem[];
  __bf16* As = reinterpret_cast<__bf16*>(smem);  // [SPLIT][BM][LK]
  __bf16* Bs = As + SPLIT * BM * LK;              // [SPLIT][BN][LK]
  if (g.run_if && *g.run_if == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int kpart = blockIdx.z;
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
    const int e = tid + i * NT8;
    if (a_mfast) { a_ml[i] = e % BM; a_kl[i] = e / BM; }
    else         { a_kl[i] = e % BK; a_ml[i] = e / BK; }
    const int m = m0 + a_ml[i];
    a_mok[i] = m < g.M;
    a_moff[i] = a_mok[i] ? imap(g.am, m) : 0;
  }
  const bool b_nfast = contiguous(g.bn) || !contiguous(g.bk);
  int b_nl[BE], b_kl[BE];
  int64_t b_noff[BE];
  bool b_ok[BE];
#pragma unroll
  for (int i = 0; i < BE; ++i) {
    const int e = tid + i * NT8;
    if (b_nfast) { b_nl[i] = e % BN; b_kl[i] = e / BN; }
    else         { b_kl[i] = e % BK; b_nl[i] = e / BK; }
    const int n = n0 + b_nl[i];
    b_ok[i] = e < BN * BK && n < g.N;
    b_noff[i] = b_ok[i] ? imap(g.bn, n) : 0;
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
      ra[i] = (a_mok[i] && k < kend) ? (float)A[a_moff[i] + imap(g.ak, k)] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int k = kt + b_kl[i];
      rb[i] = (b_ok[i] && k < kend) ? (float)B[imap(g.bk, k) + b_noff[i]] : 0.f;
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      __bf16 pc[SPLIT];
      split<SPLIT>(ra[i], pc);
#pragma unroll
      for (int s = 0; s < SPLIT; ++s) As[(s * BM + a_ml[i]) * LK + a_kl[i]] = pc[s];
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      if (tid + i * NT8 < BN * BK) {
        __bf16 pc[SPLIT];
        split<SPLIT>(rb[i], pc);
#pragma unroll
        for (int s = 0; s < SPLIT; ++s) Bs[(s * BN + b_nl[i]) * LK + b_kl[i]] = pc[s];
      }
    }
  };

  if (kbeg < kend) load_tile(kbeg);
  const int k0 = 8 * (lane >> 4);
  for (int kt = kbeg; kt < kend; kt += BK) {
    __syncthreads();
    store_tile();
    __syncthreads();
    if (kt + BK < kend) load_tile(kt + BK);  // global loads in flight during the MFMAs
    bf8 a[SPLIT][WM];
#pragma unroll
    for (int s = 0; s < SPLIT; ++s)
#pragma unroll
      for (int i = 0; i < WM; ++i)
        a[s][i] = *reinterpret_cast<const bf8*>(
            &As[(s * BM + wm * 16 * WM + i * 16 + (lane & 15)) * LK + k0]);
    // B pieces one at a time, the smaller piece products first
#pragma unroll
    for (int sb = SPLIT - 1; sb >= 0; --sb) {
      bf8 b[WN];
#pragma unroll
      for (int j = 0; j < WN; ++j)
        b[j] = *reinterpret_cast<const bf8*>(
            &Bs[(sb * BN + wn * 16 * WN + j * 16 + (lane & 15)) * LK + k0]);
#pragma unroll
      for (int sa = SPLIT - 1 - sb; sa >= 0; --sa)
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[j], acc[i][j], 0, 0, 0);
    }
  }

  // one loop per destination, the choice made once (uniform per launch)
  auto each = [&](auto&& put) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm * 16 * WM + i * 16 + 4 * (lane >> 4) + r;
          const int n = n0 + wn * 16 * WN + j * 16 + (lane & 15);
          if (m < g.M && n < g.N) put(m, n, acc[i][j][r]);
        }
  };
  if (ksplit != 1) {
    each([&](int m, int n, float v) { work[((int64_t)kpart * g.M + m) * g.N + n] = v; });
  } else if (g.c_f32) {
    float* Cf = reinterpret_cast<float*>(C);
    each([&](int m, int n, float v) {
      const int64_t off = imap(g.cm, m) + imap(g.cn, n);
      Cf[off] = (float)((g.beta != 0.0) ? g.alpha * v + g.beta * (double)Cf[off] : g.alpha * v);
    });
  } else {
    each([&](int m, int n, float v) {
      const int64_t off = imap(g.cm, m) + imap(g.cn, n);
      C[off] = (g.beta != 0.0) ? g.alpha * (double)v + g.beta * C[off] : g.alpha * (double)v;
    });
  }
}

__global__ void splitk_reduce_f32(GemmDesc g, double* __restrict__ C, int ksplit,
                                  const float* __restrict__ work) {
  if (g.run_if && *g.run_if == 0) return;
  const int64_t mn = (int64_t)g.M * g.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    // eight partial loads in flight per step; the sum keeps the part order
    double s = 0.0;
    int p = 0;
    for (; p + 8 <= ksplit; p += 8) {
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = work[(int64_t)(p + t) * mn + e];
#pragma unroll
      for (int t = 0; t < 8; ++t) s += (double)v[t];
    }
    for (; p < ksplit; ++p) s += (double)work[(int64_t)p * mn + e];
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    const int64_t off = imap(g.cm, m) + imap(g.cn, n);
    store_c(g, C, off, s);
  }
}

// The FP32 split-K partials of C (M x N, row-major) reduced in the order of
// splitk_reduce_f32 and written as the bf16 planes [piece][n][kp] of a
// following product's B operand (B(k, n) = alpha C(k, n), zero for
// M <= k < kp): the same pieces the FP64 C would be split into, without
// storing it.  64 x 64 tiles transposed through LDS (partials read along n,
// planes written along k).
template <int SPLIT>
__global__ __launch_bounds__(256) void splitk_reduce_planes_kernel(int M, int N, int ksplit,
                                                                   double alpha,
                                                                   const float* __restrict__ work,
                                                                   int Kp, __bf16* __restrict__ planes) {
  constexpr int EPT = 64 * 64 / 256;  // elements per thread
  __shared__ float tile[64][65];
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int64_t mn = (int64_t)M * N;
  // all EPT elements' loads of one part in flight together; each element
  // still sums its parts in ascending order
  double acc[EPT];
  int64_t off[EPT];
  bool ok[EPT];
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int e = threadIdx.x + j * 256;
    const int m = m0 + e / 64, n = n0 + e % 64;
    ok[j] = m < M && n < N;
    off[j] = ok[j] ? (int64_t)m * N + n : 0;
    acc[j] = 0.0;
  }
  for (int p = 0; p < ksplit; ++p) {
    float x[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) x[j] = work[(int64_t)p * mn + off[j]];
#pragma unroll
    for (int j = 0; j < EPT; ++j) acc[j] += (double)x[j];
  }
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int e = threadIdx.x + j * 256;
    tile[e / 64][e % 64] = ok[j] ? (float)(alpha * acc[j]) : 0.f;
  }
  __syncthreads();
  const int64_t plane = (int64_t)N * Kp;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int ni = e / 64, mi = e % 64;
    const int m = m0 + mi, n = n0 + ni;
    if (n < N && m < Kp) {
      __bf16 pc[SPLIT];
      split<SPLIT>(tile[mi][ni], pc);
#pragma unroll
      for (int s2 = 0; s2 < SPLIT; ++s2) planes[s2 * plane + (int64_t)n * Kp + m] = pc[s2];
    }
  }
}

// general-path tiles: 128 x 160 for large products, 64 x 64 otherwise
struct GTile {
  int wm, wn;
};
GTile general_tile(const GemmDesc& g) {
  if (g.M >= 1024 && g.N >= 96) return GTile{2, 5};
  return GTile{1, 2};
}
int general_ksplit(const GemmDesc& g) {
  const GTile t = general_tile(g);
  const int BM = 64 * t.wm, BN = 32 * t.wn;
  const int64_t tiles = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  if (tiles >= 384 || g.K < 8 * BK) return 1;
  int ks = (int)((512 + tiles - 1) / tiles);
  ks = std::min(ks, g.K / (4 * BK));
  ks = std::min(ks, 32);
  return std::max(ks, 1);
}
size_t general_partial_elems(const GemmDesc& g, int ks) {  // doubles holding FP32 partials
  return ks > 1 ? ((size_t)ks * g.M * g.N + 1) / 2 : 0;
}

template <int SPLIT, int WM, int WN>
void launch_general(const GemmDesc& g, const double* A, const double* B, double* C, int ks,
                    double* work, hipStream_t s) {
  constexpr int BM = 64 * WM, BN = 32 * WN;
  const size_t lds = sizeof(__bf16) * (size_t)SPLIT * (BM + BN) * (BK + KPAD);
  static bool attr = false;
  if (!attr) {
    ACETN_HIP_CHECK(hipFuncSetAttribute((const void*)gemm_general_kernel<SPLIT, WM, WN>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_split: grid too large");
  float* wf = reinterpret_cast<float*>(work);
  hipEvent_t e0 = kernel_timing_begin(s);
  hipLaunchKernelGGL((gemm_general_kernel<SPLIT, WM, WN>), grid, dim3(NT8), lds, s, g, A, B, C, ks,
                     wf);
  ACETN_HIP_CHECK(hipGetLastError());
  // algorithmic bytes: FP64 operands read once, C written once
  kernel_timing_end(e0, s, kKGemmSplit, 2.0 * g.M * g.N * (double)g.K,
                    8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N));
  if (ks > 1) {
    const int64_t mn = (int64_t)g.M * g.N;
    const int blocks = (int)std::min<int64_t>((mn + 255) / 256, 2048);
    hipLaunchKernelGGL(splitk_reduce_f32, dim3(blocks), dim3(256), 0, s, g, C, ks, wf);
    ACETN_HIP_CHECK(hipGetLastError());
  }
}

template <int SPLIT>
void run_general(const GemmDesc& g, const double* A, const double* B, double* C, int ks,
                 double* work, hipStream_t s) {
  if (general_tile(g).wm == 2) launch_general<SPLIT, 2, 5>(g, A, B, C, ks, work, s);
  else launch_general<SPLIT, 1, 2>(g, A, B, C, ks, work, s);
}

struct Tile {
  int wm, wn;
};
Tile choose_tile(const GemmDesc& g) {
  if (g.N > 96 && g.N <= 160) return g.M >= 2048 ? Tile{2, 5} : Tile{1, 5};
  // short k loops (the quarter tensor's K = chi) finish a tile in a few steps:
  // smaller tiles keep more of them in flight
  if (g.M >= 1024 && g.N >= 1024) return g.K <= 256 ? Tile{2, 2} : Tile{4, 4};
  return Tile{2, 2};
}

// Workgroups of one planar kernel instance resident per CU (registers and
// LDS), from the runtime's occupancy calculator; 1 if it cannot say.
template <int WM, int WN, int SPLIT, int BKT = BK>
int planar_wgs_per_cu() {
  static int n = 0;
  if (n == 0) {
    constexpr int BM = 32 * WM, BN = 32 * WN;
    const size_t lds = sizeof(__bf16) * (size_t)SPLIT * (BM + BN) * (BKT + KPAD);
    const void* f = (const void*)gemm_split_kernel<WM, WN, SPLIT, BKT>;
    ACETN_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, f, NT, lds) != hipSuccess || v < 1) v = 1;
    n = v;
  }
  return n;
}

template <int SPLIT>
int planar_wgs_per_cu(const Tile& t, int var = 0) {
  if constexpr (SPLIT == 2) {
    if (var == 1) {
      if (t.wm == 2 && t.wn == 5) return planar_wgs_per_cu<2, 5, 2, 2 * BK>();
      if (t.wm == 1 && t.wn == 5) return planar_wgs_per_cu<1, 5, 2, 2 * BK>();
      if (t.wm == 4 && t.wn == 4) return planar_wgs_per_cu<4, 4, 2, 2 * BK>();
      return planar_wgs_per_cu<2, 2, 2, 2 * BK>();
    }
  }
  if (t.wm == 2 && t.wn == 5) return planar_wgs_per_cu<2, 5, SPLIT>();
  if (t.wm == 1 && t.wn == 5) return planar_wgs_per_cu<1, 5, SPLIT>();
  if (t.wm == 4 && t.wn == 4) return planar_wgs_per_cu<4, 4, SPLIT>();
  return planar_wgs_per_cu<2, 2, SPLIT>();
}

// Split-K count: the largest that keeps tiles x parts within one round of
// resident workgroups on the 256 CUs (a second, partly filled round costs as
// much as a full one), each part at least 4 k-tiles long.  per_cu = 0 gives
// the upper bound used to size the scratch (LDS limit only, 32-deep tiles).
int choose_ksplit_for(const GemmDesc& g, int split, int per_cu, const Tile& t) {
  const int BM = 32 * t.wm, BN = 32 * t.wn;
  const int64_t tiles = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  if (per_cu <= 0) {
    const size_t lds = sizeof(__bf16) * (size_t)split * (BM + BN) * (BK + KPAD);
    per_cu = std::max(1, (int)(160 * 1024 / lds));
  }
  const int64_t cap = (int64_t)kCUs * per_cu;
  if (2 * tiles > cap || g.K < 8 * BK) return 1;
  int ks = (int)(cap / tiles);
  ks = std::min(ks, g.K / (BK * 4));
  ks = std::min(ks, 64);
  return std::max(ks, 1);
}
int choose_ksplit(const GemmDesc& g, int split, int per_cu) {
  return choose_ksplit_for(g, split, per_cu, choose_tile(g));
}
int choose_ksplit_run(const GemmDesc& g, int split) {
  const Tile t = choose_tile(g);
  return choose_ksplit(g, split, split == 2 ? planar_wgs_per_cu<2>(t) : planar_wgs_per_cu<3>(t));
}
int ksplit_bound(const GemmDesc& g) {
  return std::max(choose_ksplit(g, 2, 0), choose_ksplit(g, 3, 0));
}

int padded_k(int K) { return (std::max(K, 1) + BK - 1) / BK * BK; }

// scratch (in doubles): FP32 split-K partials, then the B planes
// (rounded to 32 doubles so that the planes start 256-byte aligned)
size_t partial_elems(const GemmDesc& g, int ks) {
  return ks > 1 ? (((size_t)ks * g.M * g.N + 1) / 2 + 31) / 32 * 32 : 0;
}
size_t plane_elems(const GemmDesc& g, int split) {
  const size_t bytes = (size_t)split * g.N * padded_k(g.K) * sizeof(__bf16);
  return (bytes + 255) / 256 * 256 / sizeof(double);
}

template <int WM, int WN, int SPLIT, int BKT = BK>
void launch(const GemmDesc& g, const __bf16* Ap, int64_t aplane, const __bf16* Bp, int Kp,
            double* C, int ks, double* work, hipStream_t s, SplitPlanes* Cp = nullptr) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  const size_t lds = sizeof(__bf16) * (size_t)SPLIT * (BM + BN) * (BKT + KPAD);
  // the k loop reads whole tiles: the padded length must be a multiple
  ACETN_REQUIRE(Kp % BKT == 0, "gemm_split: padded K is not a multiple of the k-tile");
  planar_wgs_per_cu<WM, WN, SPLIT, BKT>();  // sets the dynamic-LDS attribute once
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_split: grid too large");
  float* wf = reinterpret_cast<float*>(work);
  hipEvent_t e0 = kernel_timing_begin(s);
  hipLaunchKernelGGL((gemm_split_kernel<WM, WN, SPLIT, BKT>), grid, dim3(NT), lds, s, g, Ap, aplane,
                     Bp, Kp, C, ks, wf);
  ACETN_HIP_CHECK(hipGetLastError());
  // algorithmic bytes: both operands as SPLIT bf16 planes, C once in FP64
  kernel_timing_end(e0, s, kKGemmSplit, 2.0 * g.M * g.N * (double)g.K,
                    2.0 * SPLIT * ((double)g.M + g.N) * Kp + (g.c_f32 ? 4.0 : 8.0) * g.M * g.N);
  if (ks > 1 && Cp != nullptr) {
    const dim3 grid((Cp->kp + 63) / 64, (g.N + 63) / 64);
    __bf16* out = reinterpret_cast<__bf16*>(Cp->p);
    if (Cp->pieces == 2)
      hipLaunchKernelGGL(splitk_reduce_planes_kernel<2>, grid, dim3(256), 0, s, g.M, g.N, ks,
                         g.alpha, wf, Cp->kp, out);
    else
      hipLaunchKernelGGL(splitk_reduce_planes_kernel<3>, grid, dim3(256), 0, s, g.M, g.N, ks,
                         g.alpha, wf, Cp->kp, out);
    ACETN_HIP_CHECK(hipGetLastError());
  } else if (ks > 1) {
    const int64_t mn = (int64_t)g.M * g.N;
    const int blocks = (int)std::min<int64_t>((mn + 255) / 256, 2048);
    hipLaunchKernelGGL(splitk_reduce_f32, dim3(blocks), dim3(256), 0, s, g, C, ks, wf);
    ACETN_HIP_CHECK(hipGetLastError());
  }
}

// C = A B with A given as planes; B (FP64 through its index maps) is split
// into planes after the partials in the scratch, unless Bin already holds
// them.  With Cp the result goes to planes (see gemm_split_planar).
// ``work_elems`` (when nonzero) lets the autotuner try other tiles and split
// counts that fit the scratch; the caller's ks is the threshold choice.
template <int SPLIT>
void run(const GemmDesc& g, const __bf16* Ap, int64_t aplane, const double* B, double* C, int ks,
         double* work, hipStream_t s, const SplitPlanes* Bin = nullptr, SplitPlanes* Cp = nullptr,
         size_t work_elems = 0, int flags = 0) {
  const int Kp = padded_k(g.K);
  ACETN_REQUIRE((int64_t)SPLIT * aplane < (int64_t(1) << 31) &&
                    (int64_t)SPLIT * g.N * Kp < (int64_t(1) << 31),
                "gemm_split: operand planes exceed 2^31 elements");
  if (Bin != nullptr)
    ACETN_REQUIRE(Bin->p && Bin->rows == g.N && Bin->kp == Kp && Bin->pieces >= SPLIT,
                  "gemm_split: B planes do not match the product");
  // the B planes follow the split-K partials in the scratch, so their place
  // depends on the split count; each launch splits B again
  auto launch_choice = [&](const TuneChoice& c) {
    __bf16* planes = reinterpret_cast<__bf16*>(work + partial_elems(g, c.ks));
    if (Bin != nullptr) planes = reinterpret_cast<__bf16*>(Bin->p);
    else split_into_planes<SPLIT>(B, g.bk, g.bn, g.K, g.N, Kp, planes, s);
    if constexpr (SPLIT == 2) {
      // variant 1: 64-deep k-tiles (two MFMA k-steps per barrier pair)
      if (c.var == 1) {
        if (c.wm == 2 && c.wn == 5) launch<2, 5, 2, 2 * BK>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
        else if (c.wm == 1 && c.wn == 5) launch<1, 5, 2, 2 * BK>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
        else if (c.wm == 4 && c.wn == 4) launch<4, 4, 2, 2 * BK>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
        else launch<2, 2, 2, 2 * BK>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
        return;
      }
    }
    if (c.wm == 2 && c.wn == 5) launch<2, 5, SPLIT>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
    else if (c.wm == 1 && c.wn == 5) launch<1, 5, SPLIT>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
    else if (c.wm == 4 && c.wn == 4) launch<4, 4, SPLIT>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
    else launch<2, 2, SPLIT>(g, Ap, aplane, planes, Kp, C, c.ks, work, s, Cp);
  };
  const Tile t0 = choose_tile(g);
  TuneChoice ch{t0.wm, t0.wn, ks};
  // products that write C planes reduce their split-K partials straight into
  // the planes; a single part would leave the planes unwritten, so their
  // candidates keep at least two parts (and the threshold choice when that
  // is a single part)
  if ((Cp == nullptr || ks > 1) && work_elems > 0) {
    auto fits = [&](int k) {
      return partial_elems(g, k) + (Bin ? 0 : plane_elems(g, SPLIT)) <= work_elems;
    };
    const TuneKey key{SPLIT, g.M, g.N, g.K,
                      flags | (Bin ? 4 : 0) | (g.c_f32 ? 8 : 0) | (g.a_f32 ? 16 : 0) |
                          (Cp ? 32 : 0)};
    const int kmin = Cp ? 2 : 1;
    if (autotune_should_tune(key, g, s)) {
      std::vector<TuneChoice> cands;
      const Tile tiles[4] = {{1, 5}, {2, 5}, {2, 2}, {4, 4}};
      // two-piece products also try 64-deep k-tiles when the padded K allows
      const int nvar = (SPLIT == 2 && Kp % (2 * BK) == 0) ? 2 : 1;
      for (int var = 0; var < nvar; ++var)
        for (const Tile& t : tiles) {
          const int h = choose_ksplit_for(g, SPLIT, planar_wgs_per_cu<SPLIT>(t, var), t);
          const int opts[4] = {1, h, 2 * h, h / 2};
          for (int k : opts) {
            if (k < kmin || k > 64 || !fits(k) || (k > 1 && g.K / k < (var ? 2 * BK : BK)))
              continue;
            bool dup = false;
            for (const TuneChoice& c : cands)
              dup |= c.wm == t.wm && c.wn == t.wn && c.ks == k && c.var == var;
            if (!dup) {
              TuneChoice c{t.wm, t.wn, k};
              c.var = var;
              cands.push_back(c);
            }
          }
        }
      autotune_run(key, cands, s, launch_choice);
      return;
    }
    TuneChoice tuned;
    if (autotune_lookup(key, &tuned) && fits(tuned.ks) && tuned.ks >= kmin &&
        (tuned.var == 0 || (SPLIT == 2 && Kp % (2 * BK) == 0)))
      ch = tuned;
  }
  launch_choice(ch);
}
}  // namespace

// Large general products split A into planes once too (one extra launch,
// no per-tile conversion); small ones convert in the 8-wave kernel.
constexpr double kPlanarFlops = 1e8;

size_t gemm_split_work_elems(const GemmDesc& g, int count) {
  ACETN_REQUIRE(count == 1, "gemm_split: batched products are not supported");
  return std::max(partial_elems(g, ksplit_bound(g)) + plane_elems(g, 3) +
                      split_planes_elems(g.M, g.K),
                  general_partial_elems(g, general_ksplit(g)));
}

void gemm_split(const GemmDesc& g, const GemmBatch& b, int split, DType ta, DType tb, DType tc,
                double* work, size_t work_elems, hipStream_t s) {
  ACETN_REQUIRE(b.count == 1, "gemm_split: batched products are not supported");
  ACETN_REQUIRE(split == 2 || split == 3, "gemm_split: split must be 2 or 3");
  ACETN_REQUIRE(ta == DType::F64 && tb == DType::F64 && (tc == DType::F64 || tc == DType::F32),
                "gemm_split: FP64 operands, FP64 or FP32 result");
  if (tc == DType::F32 && !g.c_f32) {
    GemmDesc g2 = g;
    g2.c_f32 = true;
    gemm_split(g2, b, split, ta, tb, tc, work, work_elems, s);
    return;
  }
  if (g.M <= 0 || g.N <= 0) return;
  ACETN_REQUIRE(g.K > 0 || g.beta == 0.0, "gemm_split: K == 0 with beta != 0 unsupported");
  if (gemm_split_is_planar(g)) {
    const size_t apl = split_planes_elems(g.M, g.K);
    int ks = choose_ksplit_run(g, split);
    if (work_elems < partial_elems(g, ks) + plane_elems(g, 3) + apl) ks = 1;
    ACETN_REQUIRE(work != nullptr && work_elems >= partial_elems(g, ks) + plane_elems(g, 3) + apl,
                  "gemm_split: scratch too small");
    const int Kp = padded_k(g.K);
    __bf16* Ap = reinterpret_cast<__bf16*>(work + partial_elems(g, ks) + plane_elems(g, 3));
    const int64_t total = (int64_t)g.M * Kp;
    // planes[s][m][k] of A(m, k) = A[am(m) + ak(k)]: the B-plane split with the maps swapped
    if (split == 2) split_into_planes<2>(b.A[0], g.ak, g.am, g.K, g.M, Kp, Ap, s, g.a_f32);
    else split_into_planes<3>(b.A[0], g.ak, g.am, g.K, g.M, Kp, Ap, s, g.a_f32);
    // the tuner may use the scratch up to the A planes
    const size_t avail = partial_elems(g, ks) + plane_elems(g, 3);
    if (split == 2) run<2>(g, Ap, total, b.B[0], b.C[0], ks, work, s, nullptr, nullptr, avail, 1);
    else run<3>(g, Ap, total, b.B[0], b.C[0], ks, work, s, nullptr, nullptr, avail, 1);
    return;
  }
  ACETN_REQUIRE(!g.a_f32, "gemm_split: an FP32 A operand needs the plane path");
  int ks = general_ksplit(g);
  if (work == nullptr || work_elems < general_partial_elems(g, ks)) ks = 1;
  if (split == 2) run_general<2>(g, b.A[0], b.B[0], b.C[0], ks, work, s);
  else run_general<3>(g, b.A[0], b.B[0], b.C[0], ks, work, s);
}

void gemm_split_planar(const GemmDesc& g, const SplitPlanes& A, const double* B, double* C,
                       int split, double* work, size_t work_elems, hipStream_t s,
                       const SplitPlanes* Bp, SplitPlanes* Cp) {
  ACETN_REQUIRE(split == 2 || split == 3, "gemm_split: split must be 2 or 3");
  ACETN_REQUIRE(A.p != nullptr && A.rows == g.M && A.kp == padded_k(g.K) && A.pieces >= split,
                "gemm_split: A planes do not match the product");
  if (g.M <= 0 || g.N <= 0) return;
  ACETN_REQUIRE(!(g.c_f32 && Cp != nullptr), "gemm_split: C planes come from an FP64 product");
  ACETN_REQUIRE(work != nullptr && work_elems >= partial_elems(g, 1) + plane_elems(g, split),
                "gemm_split: scratch too small");
  int ks = choose_ksplit_run(g, split);
  if (work_elems < partial_elems(g, ks) + plane_elems(g, split)) ks = 1;
  if (Cp != nullptr) {
    ACETN_REQUIRE(g.beta == 0.0 && Cp->p && Cp->rows == g.N && Cp->kp == padded_k(g.M) &&
                      (Cp->pieces == 2 || Cp->pieces == 3),
                  "gemm_split: C planes do not match the product");
    // without split-K partials the product writes C: split it afterwards
    if (ks == 1) {
      const int64_t aplane = (int64_t)A.rows * A.kp;
      const __bf16* Ap = reinterpret_cast<const __bf16*>(A.p);
      if (split == 2) run<2>(g, Ap, aplane, B, C, ks, work, s, Bp);
      else run<3>(g, Ap, aplane, B, C, ks, work, s, Bp);
      // planes of C(k = m, n) = C[cm(m) + cn(n)]
      __bf16* out = reinterpret_cast<__bf16*>(Cp->p);
      if (Cp->pieces == 2) split_into_planes<2>(C, g.cm, g.cn, g.M, g.N, Cp->kp, out, s);
      else split_into_planes<3>(C, g.cm, g.cn, g.M, g.N, Cp->kp, out, s);
      return;
    }
  }
  const int64_t aplane = (int64_t)A.rows * A.kp;
  const __bf16* Ap = reinterpret_cast<const __bf16*>(A.p);
  if (split == 2) run<2>(g, Ap, aplane, B, C, ks, work, s, Bp, Cp, work_elems, 2);
  else run<3>(g, Ap, aplane, B, C, ks, work, s, Bp, Cp, work_elems, 2);
}

int split_planes_kp(int K) { return padded_k(K); }

bool gemm_split_is_planar(const GemmDesc& g) {
  if (g.planar >= 0) return g.planar != 0;
  return 2.0 * g.M * g.N * (double)g.K >= kPlanarFlops;
}

size_t split_planes_elems(int rows, int K) {
  const size_t bytes = (size_t)3 * rows * padded_k(K) * sizeof(__bf16);
  return (bytes + 255) / 256 * 256 / sizeof(double);
}

void split_transpose(const double* X, int R, int Cc, int64_t ld, SplitPlanes& N, SplitPlanes& T,
                     hipStream_t s, bool x_f32) {
  ACETN_REQUIRE(N.p && T.p, "split_transpose: null planes");
  N.rows = R; N.kp = padded_k(Cc); N.pieces = 3;
  T.rows = Cc; T.kp = padded_k(R); T.pieces = 3;
  dim3 grid((std::max(N.kp, Cc) + TT - 1) / TT, (std::max(T.kp, R) + TT - 1) / TT);
  hipLaunchKernelGGL(split_transpose_kernel, grid, dim3(256), 0, s, X, R, Cc, ld, T.kp, N.kp,
                     reinterpret_cast<__bf16*>(N.p), reinterpret_cast<__bf16*>(T.p), x_f32 ? 1 : 0);
  ACETN_HIP_CHECK(hipGetLastError());
}

}  // namespace acetn
