// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// One kernel serves every contraction of the CTMRG / full-update path: each
// operand is addressed through two-level index maps, so the index
// permutations the reference performs with torch.einsum transposes are folded
// into the global->LDS tile loads instead of being materialised in HBM.
//
// Tiles: 4 waves in a 2 x 2 arrangement, each owning WM x WN MFMA blocks of
// 16 x 16 (workgroup tile 32*WM x 32*WN, chosen per shape), BK = 16.  Operand
// fragments follow the MFMA register layout: A(16x4) lane l holds
// A[l%16][l/16], B(4x16) lane l holds B[l/16][l%16], and the 16x16 f64
// result lane l holds D[l/16 + 4r][l%16], r = 0..3 (checked on gfx950).
// Long-K / few-tile problems are split along K into a deterministic partial
// buffer that a second kernel reduces (no atomics, bitwise reproducible).
#include "internal.h"

#include <mutex>
#include <vector>

namespace acetn {

typedef double d4 __attribute__((ext_vector_type(4)));

// ---- optional live timing of the matrix-core kernel launches (HIP events on
// the launch stream), per kernel class, used by the benchmark for the
// roofline figures.
namespace {
struct KTimer {
  std::mutex mu;
  bool on = false;
  struct Rec { hipEvent_t e0, e1; int cls; };
  std::vector<Rec> ev;
  double flops[kKClasses] = {};
  double bytes[kKClasses] = {};
  int64_t launches[kKClasses] = {};
};
KTimer& ktimer() {
  static KTimer t;
  return t;
}
}  // namespace

void kernel_timing_enable(bool on) {
  KTimer& t = ktimer();
  std::lock_guard<std::mutex> g(t.mu);
  for (auto& e : t.ev) { (void)hipEventDestroy(e.e0); (void)hipEventDestroy(e.e1); }
  t.ev.clear();
  for (int c = 0; c < kKClasses; ++c) { t.flops[c] = 0.0; t.bytes[c] = 0.0; t.launches[c] = 0; }
  t.on = on;
}

void kernel_timing_read(int cls, double* ms, double* flops, int64_t* launches) {
  ACETN_REQUIRE(cls >= 0 && cls < kKClasses, "kernel_timing_read: unknown kernel class");
  KTimer& t = ktimer();
  std::lock_guard<std::mutex> g(t.mu);
  double tot = 0.0;
  for (auto& e : t.ev) {
    if (e.cls != cls) continue;
    ACETN_HIP_CHECK(hipEventSynchronize(e.e1));
    float x = 0.f;
    ACETN_HIP_CHECK(hipEventElapsedTime(&x, e.e0, e.e1));
    tot += x;
  }
  *ms = tot;
  *flops = t.flops[cls];
  *launches = t.launches[cls];
}

hipEvent_t kernel_timing_begin(hipStream_t s) {
  KTimer& t = ktimer();
  if (!t.on) return nullptr;
  hipEvent_t e0 = nullptr;
  ACETN_HIP_CHECK(hipEventCreate(&e0));
  ACETN_HIP_CHECK(hipEventRecord(e0, s));
  return e0;
}

double kernel_timing_bytes(int cls) {
  ACETN_REQUIRE(cls >= 0 && cls < kKClasses, "kernel_timing_bytes: unknown kernel class");
  KTimer& t = ktimer();
  std::lock_guard<std::mutex> g(t.mu);
  return t.bytes[cls];
}

void kernel_timing_end(hipEvent_t e0, hipStream_t s, int cls, double flops, double bytes) {
  if (e0 == nullptr) return;
  KTimer& t = ktimer();
  hipEvent_t e1 = nullptr;
  ACETN_HIP_CHECK(hipEventCreate(&e1));
  ACETN_HIP_CHECK(hipEventRecord(e1, s));
  std::lock_guard<std::mutex> lg(t.mu);
  t.ev.push_back({e0, e1, cls});
  t.flops[cls] += flops;
  t.bytes[cls] += bytes;
  t.launches[cls] += 1;
}

namespace {
constexpr int BK = 16, NT = 256;

__device__ __forceinline__ bool contiguous(const IdxMap& m) { return m.s0 == 1; }

// Workgroup tile BM x BN with BM = 16*WM*WGM, BN = 32*WN: WGM x 2 waves
// (WGM = 2 or 4), each wave owning WM x WN MFMA blocks of 16 x 16, k-tiles
// of BKT (16 or 32).  Double-buffered LDS (one barrier per k-tile) with the
// next tile prefetched into registers; more waves or longer k-tiles give the
// global loads more MFMA work to hide behind.
// LIN: every operand map is a single stride (no grouped digits), so an
// element's address is base + row*stride + k*stride: rows past the edge are
// clamped to the last row (their results are never stored) and only the K
// edge needs a predicate, which keeps the tile loads free of branches.
// Otherwise the k part of the address (a mixed-radix map) is evaluated once
// per k-tile and operand by 2*BK threads into a double-buffered LDS table
// that every element load reads, instead of once per element.
// Four-wave tiles up to 2 x 5 blocks per wave are held to 256 VGPRs, so two
// workgroups share a CU (their LDS fits twice); larger tiles and the
// eight-wave variants take the registers they need.
template <int WM, int WN, bool LIN, int WGM = 2, int BKT = BK>
__global__ __launch_bounds__(128 * WGM, (WGM == 2 && WM < 4 && BKT == 16) ? 2 : 1) void
gemm_f64_kernel(GemmDesc g, GemmBatch bt, int ksplit, double* __restrict__ work) {
  constexpr int NT = 128 * WGM;
  constexpr int BK = BKT;
  constexpr int BM = 16 * WM * WGM, BN = 32 * WN;
  constexpr int LA = BM + 2, LB = BN + 2;
  constexpr int AE = BM * BK / NT, BE = BN * BK / NT;  // elements per thread
  static_assert(AE * NT == BM * BK && BE * NT == BN * BK, "tile not divisible by the threads");
  __shared__ double As[2][BK][LA];
  __shared__ double Bs[2][BK][LB];
  __shared__ int64_t koff[2][2][BK];  // [table][operand A / B][k within the tile]
  if (g.run_if && *g.run_if == 0) return;
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

  // ---- load patterns (coalesced along whichever index is unit-stride)
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
    if (LIN) a_moff[i] = (int64_t)min(m, g.M - 1) * g.am.s0;
    else a_moff[i] = a_mok[i] ? imap(g.am, m) : 0;
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
    if (LIN) b_noff[i] = (int64_t)min(n, g.N - 1) * g.bn.s0;
    else b_noff[i] = b_nok[i] ? imap(g.bn, n) : 0;
  }

  d4 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[AE], rb[BE];
  auto load_tile = [&](int kt) {
    if (LIN) {
      const int64_t sak = g.ak.s0, sbk = g.bk.s0;
#pragma unroll
      for (int i = 0; i < AE; ++i) {
        const int k = kt + a_kl[i];
        const double v = A[a_moff[i] + (int64_t)min(k, kend - 1) * sak];
        ra[i] = k < kend ? v : 0.0;
      }
#pragma unroll
      for (int i = 0; i < BE; ++i) {
        const int k = kt + b_kl[i];
        const double v = B[(int64_t)min(k, kend - 1) * sbk + b_noff[i]];
        rb[i] = k < kend ? v : 0.0;
      }
      return;
    }
    const int tb = ((kt - kbeg) / BK) & 1;
#pragma unroll
    for (int i = 0; i < AE; ++i) {
      const int k = kt + a_kl[i];
      ra[i] = (a_mok[i] && k < kend) ? A[a_moff[i] + koff[tb][0][a_kl[i]]] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < BE; ++i) {
      const int k = kt + b_kl[i];
      rb[i] = (b_nok[i] && k < kend) ? B[koff[tb][1][b_kl[i]] + b_noff[i]] : 0.0;
    }
  };
  // k-offset table of the k-tile at kt (general maps only)
  auto fill_koff = [&](int kt) {
    if (LIN || tid >= 2 * BK) return;
    const int tb = ((kt - kbeg) / BK) & 1, op = tid / BK, j = tid % BK;
    const int k = min(kt + j, kend - 1);
    koff[tb][op][j] = op == 0 ? imap(g.ak, k) : imap(g.bk, k);
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < AE; ++i) As[buf][a_kl[i]][a_ml[i]] = ra[i];
#pragma unroll
    for (int i = 0; i < BE; ++i) Bs[buf][b_kl[i]][b_nl[i]] = rb[i];
  };

  int buf = 0;
  if (kbeg < kend) {
    if (!LIN) {
      fill_koff(kbeg);
      __syncthreads();
    }
    load_tile(kbeg);
    store_tile(0);
    if (kbeg + BK < kend) fill_koff(kbeg + BK);
  }
  __syncthreads();
  for (int kt = kbeg; kt < kend; kt += BK) {
    const bool more = kt + BK < kend;
    if (more) load_tile(kt + BK);  // global loads in flight during the MFMAs
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      double a[WM], b[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i) a[i] = As[buf][kr][wm * 16 * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < WN; ++j) b[j] = Bs[buf][kr][wn * 16 * WN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) store_tile(buf ^ 1);
    // the table of k-tile kt + 2 BK replaces that of kt, read before the
    // previous barrier
    if (kt + 2 * BK < kend) fill_koff(kt + 2 * BK);
    __syncthreads();
    buf ^= 1;
  }

  // ---- epilogue
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 16 * WM + i * 16 + (lane >> 4) + 4 * r;
        const int n = n0 + wn * 16 * WN + j * 16 + (lane & 15);
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
  if (g.run_if && *g.run_if == 0) return;
  const int64_t mn = (int64_t)g.M * g.N;
  const int prob = blockIdx.y;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn;
       e += (int64_t)gridDim.x * blockDim.x) {
    // eight partial loads in flight per step; the sum keeps the part order
    const double* wp = work + (int64_t)prob * ksplit * mn + e;
    double s = 0.0;
    int p = 0;
    for (; p + 8 <= ksplit; p += 8) {
      double v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = wp[(int64_t)(p + t) * mn];
#pragma unroll
      for (int t = 0; t < 8; ++t) s += v[t];
    }
    for (; p < ksplit; ++p) s += wp[(int64_t)p * mn];
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    double* C = bt.C[prob];
    const int64_t off = imap(g.cm, m) + imap(g.cn, n);
    C[off] = (g.beta != 0.0) ? g.alpha * s + g.beta * C[off] : g.alpha * s;
  }
}

// Small-output, long-K products (M, N <= 8, e.g. the d^2 x d^2 reduced
// density matrix summed over the whole environment): a 16 x 16 matrix-core
// tile would be almost all padding, so each lane walks its own k indices and
// accumulates all M x N dot products in registers; each workgroup reduces
// its lanes (shuffles, then LDS across waves) into one partial, and a second
// kernel sums the partials in a fixed order (bitwise reproducible).
constexpr int kDotMax = 8, kDotNT = 256, kDotWGs = 512, kDotMinK = 4096;

bool dot_small_fits(const GemmDesc& g) {
  return g.M <= kDotMax && g.N <= kDotMax && g.K >= kDotMinK;
}
int dot_small_wgs(const GemmDesc& g) {
  return (int)std::min<int64_t>(kDotWGs, ((int64_t)g.K + kDotNT - 1) / kDotNT);
}

__global__ __launch_bounds__(kDotNT) void dot_small_kernel(GemmDesc g, GemmBatch bt,
                                                           double* __restrict__ part) {
  constexpr int MN = kDotMax * kDotMax;
  __shared__ double red[kDotNT / 64][MN];
  if (g.run_if && *g.run_if == 0) return;
  const int prob = blockIdx.y;
  const double* __restrict__ A = bt.A[prob];
  const double* __restrict__ B = bt.B[prob];
  const int M = g.M, N = g.N;
  int64_t aoff[kDotMax], boff[kDotMax];
#pragma unroll
  for (int m = 0; m < kDotMax; ++m) aoff[m] = m < M ? imap(g.am, m) : 0;
#pragma unroll
  for (int n = 0; n < kDotMax; ++n) boff[n] = n < N ? imap(g.bn, n) : 0;
  double acc[kDotMax][kDotMax];
#pragma unroll
  for (int m = 0; m < kDotMax; ++m)
#pragma unroll
    for (int n = 0; n < kDotMax; ++n) acc[m][n] = 0.0;
  for (int k = blockIdx.x * kDotNT + threadIdx.x; k < g.K; k += gridDim.x * kDotNT) {
    const int64_t ka = imap(g.ak, k), kb = imap(g.bk, k);
    double a[kDotMax], b[kDotMax];
#pragma unroll
    for (int m = 0; m < kDotMax; ++m) a[m] = m < M ? A[aoff[m] + ka] : 0.0;
#pragma unroll
    for (int n = 0; n < kDotMax; ++n) b[n] = n < N ? B[kb + boff[n]] : 0.0;
#pragma unroll
    for (int m = 0; m < kDotMax; ++m)
#pragma unroll
      for (int n = 0; n < kDotMax; ++n) acc[m][n] = fma(a[m], b[n], acc[m][n]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < kDotMax; ++m)
#pragma unroll
    for (int n = 0; n < kDotMax; ++n) {
      double v = acc[m][n];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
      if (lane == 0) red[wave][m * kDotMax + n] = v;
    }
  __syncthreads();
  if (threadIdx.x < MN) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kDotNT / 64; ++w) v += red[w][threadIdx.x];
    part[((int64_t)prob * gridDim.x + blockIdx.x) * MN + threadIdx.x] = v;
  }
}

// one workgroup per output element: strided partial sums, then a fixed
// binary tree in LDS (deterministic)
__global__ __launch_bounds__(kDotNT) void dot_small_reduce(GemmDesc g, GemmBatch bt,
                                                           const double* __restrict__ part,
                                                           int nwg) {
  constexpr int MN = kDotMax * kDotMax;
  __shared__ double red[kDotNT];
  if (g.run_if && *g.run_if == 0) return;
  const int prob = blockIdx.y, t = threadIdx.x;
  const int m = blockIdx.x / g.N, n = blockIdx.x % g.N;
  const double* p = part + (int64_t)prob * nwg * MN + m * kDotMax + n;
  double s = 0.0;
  for (int w = t; w < nwg; w += kDotNT) s += p[(int64_t)w * MN];
  red[t] = s;
  __syncthreads();
  for (int o = kDotNT / 2; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t != 0) return;
  s = red[0];
  double* C = bt.C[prob];
  const int64_t off = imap(g.cm, m) + imap(g.cn, n);
  C[off] = (g.beta != 0.0) ? g.alpha * s + g.beta * C[off] : g.alpha * s;
}

struct Tile {
  int wm, wn;
  int var = 0;  // TuneChoice::var
};
int tile_bm(const Tile& t) { return 16 * t.wm * ((t.var & 1) ? 4 : 2); }
// per-shape tile: skinny range-finder products (N <= 160) use 64 x 160 / 32 x 160,
// large products 128 x 128, small ones 64 x 64
Tile choose_tile(const GemmDesc& g) {
  // skinny products: 64-row tiles only when K is long enough to split
  // (short-K products such as Y R^-1 need the tile count)
  if (g.N > 96 && g.N <= 160) return (g.M >= 2048 && g.K >= 1024) ? Tile{2, 5} : Tile{1, 5};
  if (g.M >= 1024 && g.N >= 1024) return Tile{4, 4};
  return Tile{2, 2};
}

int choose_ksplit_for(const GemmDesc& g, int count, const Tile& t) {
  const int BM = tile_bm(t), BN = 32 * t.wn;
  const int64_t tiles = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN) * count;
  if (tiles >= 512 || g.K < 4 * BK * 8) return 1;
  int ks = (int)((1024 + tiles - 1) / tiles);
  ks = std::min(ks, g.K / (BK * 8));
  ks = std::min(ks, 64);
  return std::max(ks, 1);
}

int choose_ksplit(const GemmDesc& g, int count) {
  return choose_ksplit_for(g, count, choose_tile(g));
}

template <int WM, int WN>
void launch_gemm(const GemmDesc& g, const GemmBatch& b, int ks, double* work, hipStream_t s) {
  constexpr int BM = 32 * WM, BN = 32 * WN;
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, b.count * ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_f64: grid too large");
  const bool lin = g.am.n0 <= 0 && g.ak.n0 <= 0 && g.bk.n0 <= 0 && g.bn.n0 <= 0;
  if (lin) hipLaunchKernelGGL((gemm_f64_kernel<WM, WN, true>), grid, dim3(NT), 0, s, g, b, ks, work);
  else hipLaunchKernelGGL((gemm_f64_kernel<WM, WN, false>), grid, dim3(NT), 0, s, g, b, ks, work);
}

// the wider WM x 5 variants (single-stride operand maps only)
template <int WM, int WGM, int BKT>
void launch_gemm_wide(const GemmDesc& g, const GemmBatch& b, int ks, double* work,
                      hipStream_t s) {
  constexpr int BM = 16 * WM * WGM, BN = 32 * 5;
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, b.count * ks);
  ACETN_REQUIRE(grid.y <= 65535 && grid.z <= 65535, "gemm_f64: grid too large");
  hipLaunchKernelGGL((gemm_f64_kernel<WM, 5, true, WGM, BKT>), grid, dim3(128 * WGM), 0, s, g, b,
                     ks, work);
}
}  // namespace

size_t gemm_f64_work_elems(const GemmDesc& g, int count) {
  if (dot_small_fits(g)) return (size_t)dot_small_wgs(g) * count * kDotMax * kDotMax;
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
  if (dot_small_fits(g) && work != nullptr &&
      work_elems >= (size_t)dot_small_wgs(g) * b.count * kDotMax * kDotMax) {
    const int nwg = dot_small_wgs(g);
    hipEvent_t e0 = kernel_timing_begin(s);
    hipLaunchKernelGGL(dot_small_kernel, dim3(nwg, b.count), dim3(kDotNT), 0, s, g, b, work);
    ACETN_HIP_CHECK(hipGetLastError());
    kernel_timing_end(e0, s, kKGemmF64, 2.0 * g.M * g.N * (double)g.K * b.count,
                      8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N) * b.count);
    hipLaunchKernelGGL(dot_small_reduce, dim3(g.M * g.N, b.count), dim3(kDotNT), 0, s, g, b, work,
                       nwg);
    ACETN_HIP_CHECK(hipGetLastError());
    return;
  }
  auto fits = [&](int k) {
    return k == 1 || (work != nullptr && work_elems >= (size_t)k * b.count * g.M * g.N);
  };
  const bool lin = g.am.n0 <= 0 && g.ak.n0 <= 0 && g.bk.n0 <= 0 && g.bn.n0 <= 0;
  auto launch_choice = [&](const TuneChoice& c) {
    hipEvent_t e0 = kernel_timing_begin(s);
    if (c.wm == 2 && c.wn == 5 && lin && c.var == 1) launch_gemm_wide<2, 4, 16>(g, b, c.ks, work, s);
    else if (c.wm == 2 && c.wn == 5 && lin && c.var == 2) launch_gemm_wide<2, 2, 32>(g, b, c.ks, work, s);
    else if (c.wm == 2 && c.wn == 5 && lin && c.var == 3) launch_gemm_wide<2, 4, 32>(g, b, c.ks, work, s);
    else if (c.wm == 1 && c.wn == 5 && lin && c.var == 3) launch_gemm_wide<1, 4, 32>(g, b, c.ks, work, s);
    else if (c.wm == 2 && c.wn == 5) launch_gemm<2, 5>(g, b, c.ks, work, s);
    else if (c.wm == 1 && c.wn == 5) launch_gemm<1, 5>(g, b, c.ks, work, s);
    else if (c.wm == 4 && c.wn == 4) launch_gemm<4, 4>(g, b, c.ks, work, s);
    else launch_gemm<2, 2>(g, b, c.ks, work, s);
    ACETN_HIP_CHECK(hipGetLastError());
    kernel_timing_end(e0, s, kKGemmF64, 2.0 * g.M * g.N * (double)g.K * b.count,
                      8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N) * b.count);
    if (c.ks > 1) {
      const int64_t mn = (int64_t)g.M * g.N;
      int blocks = (int)std::min<int64_t>((mn + 255) / 256, 2048);
      hipLaunchKernelGGL(splitk_reduce_kernel, dim3(blocks, b.count), dim3(256), 0, s, g, b, c.ks,
                         work);
      ACETN_HIP_CHECK(hipGetLastError());
    }
  };
  const TuneKey key{0, g.M, g.N, g.K, (lin ? 1 : 0) | (b.count << 1)};
  if (autotune_should_tune(key, g, s)) {
    std::vector<TuneChoice> cands;
    const Tile tiles[8] = {{1, 5},    {2, 5},    {2, 2},    {4, 4},
                           {2, 5, 1}, {2, 5, 2}, {2, 5, 3}, {1, 5, 3}};
    for (const Tile& t : tiles) {
      if (t.var != 0 && !lin) continue;
      const int h = choose_ksplit_for(g, b.count, t);
      const int opts[4] = {1, h, 2 * h, h / 2};
      for (int k : opts) {
        if (k < 1 || k > 64 || !fits(k) || (k > 1 && g.K / k < 2 * BK)) continue;
        bool dup = false;
        for (const TuneChoice& c : cands)
          dup |= c.wm == t.wm && c.wn == t.wn && c.var == t.var && c.ks == k;
        if (dup) continue;
        TuneChoice c{t.wm, t.wn, k};
        c.var = t.var;
        cands.push_back(c);
      }
    }
    autotune_run(key, cands, s, launch_choice);
    return;
  }
  int ks = choose_ksplit(g, b.count);
  if (!fits(ks)) ks = 1;
  const Tile t = choose_tile(g);
  TuneChoice ch{t.wm, t.wn, ks};
  TuneChoice tuned;
  if (autotune_lookup(key, &tuned) && fits(tuned.ks) && (tuned.var == 0 || lin)) ch = tuned;
  launch_choice(ch);
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
