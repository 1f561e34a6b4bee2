// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One kernel serves every contraction of the CTMRG / full-update path: each
// operand is addressed through two-level index maps, so the index
// permutations the reference performs with torch.einsum transposes are folded
// into the global->LDS tile loads instead of being materialised in HBM.
//
// Tile: 64 x 64 per workgroup, BK = 16, 4 waves in a 2 x 2 arrangement, each
// wave owning a 32 x 32 sub-tile = 2 x 2 MFMA blocks of 16 x 16.  Operand
// fragments follow the MFMA register layout: A(16x4) lane l holds
// A[l%16][l/16], B(4x16) lane l holds B[l/16][l%16], and the 16x16 f64
// result lane l holds D[l/16 + 4r][l%16], r = 0..3 (checked on gfx950).
// Long-K / few-tile problems are split along K into a deterministic partial
// buffer that a second kernel reduces (no atomics, bitwise reproducible).
#include "internal.h"

namespace acetn {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int BM = 64, BN = 64, BK = 16, NT = 256;
constexpr int LDS_A = BM + 2;
constexpr int LDS_B = BN + 2;

__device__ __forceinline__ bool contiguous(const IdxMap& m) { return m.s_in == 1; }

__global__ __launch_bounds__(NT) void gemm_f64_kernel(GemmDesc g, GemmBatch bt, int ksplit,
                                                      double* __restrict__ work) {
  __shared__ double As[BK][LDS_A];
  __shared__ double Bs[BK][LDS_B];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int prob = blockIdx.z / ksplit, kpart = blockIdx.z % ksplit;
  const double* __restrict__ A = bt.A[prob];
  const double* __restrict__ B = bt.B[prob];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;

  int kchunk = (g.K + ksplit - 1) / ksplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  const int kbeg = kpart * kchunk;
  const int kend = min(g.K, kbeg + kchunk);

  // ---- load pattern for A (64 x 16 tile, 4 elements per thread)
  const bool a_mfast = contiguous(g.am) || !contiguous(g.ak);
  int a_ml[4], a_kl[4];
  int64_t a_moff[4];
  bool a_mok[4];
  for (int i = 0; i < 4; ++i) {
    if (a_mfast) { a_ml[i] = tid % BM; a_kl[i] = tid / BM + 4 * i; }
    else         { a_kl[i] = tid % BK; a_ml[i] = tid / BK + 16 * i; }
    const int m = m0 + a_ml[i];
    a_mok[i] = m < g.M;
    a_moff[i] = a_mok[i] ? imap(g.am, m) : 0;
  }
  // ---- load pattern for B (16 x 64 tile)
  const bool b_nfast = contiguous(g.bn) || !contiguous(g.bk);
  int b_nl[4], b_kl[4];
  int64_t b_noff[4];
  bool b_nok[4];
  for (int i = 0; i < 4; ++i) {
    if (b_nfast) { b_nl[i] = tid % BN; b_kl[i] = tid / BN + 4 * i; }
    else         { b_kl[i] = tid % BK; b_nl[i] = tid / BK + 16 * i; }
    const int n = n0 + b_nl[i];
    b_nok[i] = n < g.N;
    b_noff[i] = b_nok[i] ? imap(g.bn, n) : 0;
  }

  d4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[4], rb[4];
  auto load_tile = [&](int kt) {
    for (int i = 0; i < 4; ++i) {
      const int k = kt + a_kl[i];
      ra[i] = (a_mok[i] && k < kend) ? A[a_moff[i] + imap(g.ak, k)] : 0.0;
    }
    for (int i = 0; i < 4; ++i) {
      const int k = kt + b_kl[i];
      rb[i] = (b_nok[i] && k < kend) ? B[imap(g.bk, k) + b_noff[i]] : 0.0;
    }
  };

  if (kbeg < kend) load_tile(kbeg);
  for (int kt = kbeg; kt < kend; kt += BK) {
    __syncthreads();
    for (int i = 0; i < 4; ++i) As[a_kl[i]][a_ml[i]] = ra[i];
    for (int i = 0; i < 4; ++i) Bs[b_kl[i]][b_nl[i]] = rb[i];
    __syncthreads();
    if (kt + BK < kend) load_tile(kt + BK);  // prefetch next tile into registers
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double a0 = As[kr][wm * 32 + (lane & 15)];
      double a1 = As[kr][wm * 32 + 16 + (lane & 15)];
      double b0 = Bs[kr][wn * 32 + (lane & 15)];
      double b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
  }

  // ---- epilogue
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const double v = acc[i][j][r];
        if (ksplit == 1) {
          double* C = bt.C[prob];
          const int64_t off = imap(g.cm, m) + imap(g.cn, n);
          C[off] = (g.beta != 0.0) ? g.alpha * v + g.beta * C[off] : g.alpha * v;
        } else {
          work[(((int64_t)prob * ksplit + kpart) * g.M + m) * g.N + n] = v;
        }
      }
}

__global__ void splitk_reduce_kernel(GemmDesc g, GemmBatch bt, int ksplit,
                                     const double* __restrict__ work) {
  const int64_t mn = (int64_t)g.M * g.N;
  const int prob = blockIdx.y;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < ksplit; ++p) s += work[((int64_t)prob * ksplit + p) * mn + e];
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    double* C = bt.C[prob];
    const int64_t off = imap(g.cm, m) + imap(g.cn, n);
    C[off] = (g.beta != 0.0) ? g.alpha * s + g.beta * C[off] : g.alpha * s;
  }
}

int choose_ksplit(const GemmDesc& g, int count) {
  const int64_t tiles = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN) * count;
  if (tiles >= 512 || g.K < 4 * BK * 8) return 1;
  int ks = (int)((1024 + tiles - 1) / tiles);
  ks = std::min(ks, g.K / (BK * 8));
  ks = std::min(ks, 64);
  return std::max(ks, 1);
}
}  // namespace

size_t gemm_f64_work_elems(const GemmDesc& g, int count) {
  const int ks = choose_ksplit(g, count);
  return ks > 1 ? (size_t)ks * count * g.M * g.N : 0;
}

void gemm_f64(const GemmDesc& g, const GemmBatch& b, double* work, size_t work_elems,
              hipStream_t s) {
  ACETN_REQUIRE(b.count >= 1 && b.count <= kMaxBatch, "gemm_f64: bad batch count");
  if (g.M <= 0 || g.N <= 0) return;
  if (g.K <= 0) {
    // C = beta*C (or zero): handled by a K=1 product with zero contribution
    ACETN_REQUIRE(g.beta == 0.0, "gemm_f64: K == 0 with beta != 0 unsupported");
  }
  int ks = choose_ksplit(g, b.count);
  if (ks > 1 && (work == nullptr || work_elems < (size_t)ks * b.count * g.M * g.N)) ks = 1;
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, b.count * ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_f64: grid too large");
  hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(NT), 0, s, g, b, ks, work);
  ACETN_HIP_CHECK(hipGetLastError());
  if (ks > 1) {
    const int64_t mn = (int64_t)g.M * g.N;
    int blocks = (int)std::min<int64_t>((mn + 255) / 256, 2048);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(blocks, b.count), dim3(256), 0, s, g, b, ks,
                       work);
    ACETN_HIP_CHECK(hipGetLastError());
  }
}

GemmDesc dense_desc(int M, int N, int K, bool ta, int64_t lda, bool tb, int64_t ldb,
                    int64_t ldc, double alpha, double beta) {
  GemmDesc g{};
  g.M = M; g.N = N; g.K = K;
  g.am = ta ? plain(1) : plain(lda);
  g.ak = ta ? plain(lda) : plain(1);
  g.bk = tb ? plain(1) : plain(ldb);
  g.bn = tb ? plain(ldb) : plain(1);
  g.cm = plain(ldc);
  g.cn = plain(1);
  g.alpha = alpha;
  g.beta = beta;
  return g;
}

}  // namespace acetn
