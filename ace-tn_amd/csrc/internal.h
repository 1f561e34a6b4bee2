// Internal declarations shared by the acetn HIP translation units.
// Nothing here crosses the C-ABI; see include/acetn_hip.h for that.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace acetn {

// ---------------------------------------------------------------------------
// error handling: every HIP call is checked and turned into an exception that
// the C-ABI layer converts to a status code + message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define ACETN_HIP_CHECK(expr)                                                   \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      throw ::acetn::Error(2, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define ACETN_REQUIRE(cond, msg)                                                \
  do {                                                                          \
    if (!(cond)) throw ::acetn::Error(1, std::string(msg));                     \
  } while (0)

// ---------------------------------------------------------------------------
// Multi-level index map: a flat GEMM index i is split into up to four mixed-
// radix digits (innermost first, radices n0, n1, n2; the last digit takes the
// rest), each with its own stride:
//   offset(i) = (i % n0)*s0 + ((i/n0) % n1)*s1 + ((i/(n0 n1)) % n2)*s2
//               + (i/(n0 n1 n2))*s3
// n0 <= 0 means a plain stride: offset(i) = i*s0.  This lets a GEMM operand
// index a grouped einsum mode (e.g. (e, l, L) of an edge tensor laid out as
// [e][a][l][L]) without materialising the permutation.
// ---------------------------------------------------------------------------
struct IdxMap {
  int32_t n0, n1, n2;
  int64_t s0, s1, s2, s3;
};
inline IdxMap plain(int64_t s) { return IdxMap{0, 0, 0, s, 0, 0, 0}; }
// two digits: (i % inner)*s_in + (i / inner)*s_out
inline IdxMap grouped(int32_t inner, int64_t s_in, int64_t s_out) {
  return IdxMap{inner, 0, 0, s_in, s_out, 0, 0};
}
// three digits
inline IdxMap grouped3(int32_t n0, int64_t s0, int32_t n1, int64_t s1, int64_t s2) {
  return IdxMap{n0, n1, 0, s0, s1, s2, 0};
}
// four digits
inline IdxMap grouped4(int32_t n0, int64_t s0, int32_t n1, int64_t s1, int32_t n2, int64_t s2,
                       int64_t s3) {
  return IdxMap{n0, n1, n2, s0, s1, s2, s3};
}

__device__ __forceinline__ int64_t imap(const IdxMap& m, int32_t i) {
  if (m.n0 <= 0) return (int64_t)i * m.s0;
  int64_t off = (int64_t)(i % m.n0) * m.s0;
  int32_t r = i / m.n0;
  if (m.n1 <= 0) return off + (int64_t)r * m.s1;
  off += (int64_t)(r % m.n1) * m.s1;
  r /= m.n1;
  if (m.n2 <= 0) return off + (int64_t)r * m.s2;
  off += (int64_t)(r % m.n2) * m.s2;
  return off + (int64_t)(r / m.n2) * m.s3;
}

constexpr int kMaxBatch = 16;

// C[b](m,n) = alpha * sum_k A[b](m,k) B[b](k,n) + beta * C[b](m,n)
struct GemmDesc {
  int32_t M, N, K;
  IdxMap am, ak;  // A(m,k) = A[am(m) + ak(k)]
  IdxMap bk, bn;  // B(k,n) = B[bk(k) + bn(n)]
  IdxMap cm, cn;  // C(m,n) = C[cm(m) + cn(n)]
  double alpha, beta;
  const int* run_if = nullptr;  // device flag: when given and zero the product is skipped
  bool c_f32 = false;           // split-bf16 products only: C holds FP32 (the pointer is a float*)
  bool a_f32 = false;           // split-bf16 products on the plane path only: A holds FP32
  int8_t planar = -1;           // split-bf16 products: 1 / 0 forces the plane / general path,
                                // -1 picks by size (gemm_split_is_planar)
};

struct GemmBatch {
  int32_t count;
  const double* A[kMaxBatch];
  const double* B[kMaxBatch];
  double* C[kMaxBatch];
};

// FP64 MFMA GEMM.  ``work``/``work_elems`` is an optional scratch area for
// split-K partial sums (ksplit chosen internally when K is long and the tile
// grid is small).
void gemm_f64(const GemmDesc& g, const GemmBatch& b, double* work, size_t work_elems,
              hipStream_t s);
size_t gemm_f64_work_elems(const GemmDesc& g, int count);

// Split-bf16 GEMM (v_mfma_f32_16x16x32_bf16, FP32 accumulation): operands
// are rounded to FP32 and split into ``split`` bf16 pieces (2: "bf16x3",
// 3: "bf16x6" ~ FP32 accuracy).  Operand / result element types per DType;
// the GemmBatch pointers are reinterpreted accordingly.
enum class DType : int { F64 = 0, F32 = 1 };
void gemm_split(const GemmDesc& g, const GemmBatch& b, int split, DType ta, DType tb, DType tc,
                double* work, size_t work_elems, hipStream_t s);
size_t gemm_split_work_elems(const GemmDesc& g, int count);
// A matrix already split into 3 bf16 planes [piece][rows][kp] (k contiguous,
// kp = K rounded up to 32, zero padded).  p points at device memory of
// split_planes_elems(rows, K) doubles.
struct SplitPlanes {
  void* p = nullptr;
  int rows = 0, kp = 0, pieces = 0;
};
size_t split_planes_elems(int rows, int K);
int split_planes_kp(int K);  // kp of K
// whether gemm_split runs this product from bf16 planes (the path that accepts a_f32)
bool gemm_split_is_planar(const GemmDesc& g);
// Both orientations of X (R x Cc, row-major, leading dimension ld): N holds X
// itself (rows R, k = columns), T holds X^T (rows Cc, k = rows of X).
void split_transpose(const double* X, int R, int Cc, int64_t ld, SplitPlanes& N, SplitPlanes& T,
                     hipStream_t s, bool x_f32 = false);  // x_f32: X holds FP32
// C = alpha A B + beta C with A given as planes (g.am/g.ak unused; A(m, k) =
// the planes' row m, column k).  Bp, when given, holds B already split into
// planes of the next product's layout (rows N, k = K; g.bk/g.bn unused).
// Cp, when given (beta = 0), receives C as such planes (rows N, k = M, for a
// following product that takes C as its B operand, at that product's split
// level Cp->pieces) in place of the FP64 C.
void gemm_split_planar(const GemmDesc& g, const SplitPlanes& A, const double* B, double* C,
                       int split, double* work, size_t work_elems, hipStream_t s,
                       const SplitPlanes* Bp = nullptr, SplitPlanes* Cp = nullptr);

// Tile-shape autotuner (autotune.hip).  kind: 0 FP64 product, 2 / 3 split
// product with that many bf16 pieces; flags: operand-layout bits of the
// caller.  A choice is a tile (wm x wn MFMA blocks per wave) and a split-K
// count; ms is its measured time.
struct TuneKey {
  int kind, M, N, K, flags;
};
struct TuneChoice {
  int wm, wn, ks;
  double ms = 0.0;
  // FP64 kernel variant: 0 four waves (2 x 2) with k-tiles of 16, 1 eight
  // waves (4 x 2) with k-tiles of 16, 2 four waves with k-tiles of 32,
  // 3 eight waves with k-tiles of 32 (the split-bf16 kernels have one);
  // the workgroup tile is (16 wm waves_m) x (32 wn)
  int var = 0;
};
void autotune_set_mode(int mode);  // 0 thresholds, 1 tune new shapes, 2 table only
int autotune_mode();
bool autotune_lookup(const TuneKey& k, TuneChoice* out);
// whether this product should be tuned now (mode 1, shape not in the table,
// beta == 0, stream not capturing)
bool autotune_should_tune(const TuneKey& k, const GemmDesc& g, hipStream_t s);
// times every candidate through ``launch`` on s, stores and re-runs the fastest
TuneChoice autotune_run(const TuneKey& k, const std::vector<TuneChoice>& cands, hipStream_t s,
                        const std::function<void(const TuneChoice&)>& launch);
int64_t autotune_entries();
void autotune_clear();
void autotune_save(const char* path);
int64_t autotune_load(const char* path);

// Live timing of the matrix-core kernels (HIP events on the launch stream),
// per kernel class; the benchmark's roofline figures come from here.
enum KernelClass : int {
  kKGemmF64 = 0,
  kKGemmSplit = 1,
  kKDoubleLayer = 2,
  kKJacobi = 3,    // projector eigensolver (one workgroup): rotations + their replay on V
  kKCholesky = 4,  // CholeskyQR factor and inverse (one workgroup)
  kKClasses = 5
};
void kernel_timing_enable(bool on);
void kernel_timing_read(int cls, double* ms, double* flops, int64_t* launches);
// Returns an event recorded on ``s`` (nullptr when timing is off); pass it to
// kernel_timing_end after the timed launch.
hipEvent_t kernel_timing_begin(hipStream_t s);
void kernel_timing_end(hipEvent_t e0, hipStream_t s, int cls, double flops, double bytes = 0.0);
// summed algorithmic HBM bytes (operands read once, result written once) of a class
double kernel_timing_bytes(int cls);

// Convenience: dense row-major C = alpha*op(A) op(B) + beta*C.
// A is (M x K) if !ta else stored (K x M); lda/ldb/ldc are row strides.
GemmDesc dense_desc(int M, int N, int K, bool ta, int64_t lda, bool tb, int64_t ldb,
                    int64_t ldc, double alpha = 1.0, double beta = 0.0);

// ---------------------------------------------------------------------------
// Double-layer absorption of a site tensor into a block tensor:
// Y[o1,o2; k3,b3,k4,b4] = sum_{k1,b1,k2,b2,p} X[o1,o2; k1,b1,k2,b2]
//                           * A[k1,k2,k3,k4,p] * conj(A)[b1,b2,b3,b4,p]
// A's legs are addressed through strides, which folds bond_permute in.
// ---------------------------------------------------------------------------
// With open_p != 0 the physical indices stay open: Y gets the two extra
// indices p (ket, stride y_p) and pb (bra, stride y_pb) and nothing is summed
// over them (the bond reduced density matrix, acetn/measurement/rdm.py:102-120).
struct DLDesc {
  int32_t D, d, n1, n2;
  int64_t x_o1, x_o2, x_k1, x_b1, x_k2, x_b2;
  int64_t y_o1, y_o2, y_k3, y_b3, y_k4, y_b4;
  int64_t a_leg[4];
  int64_t a_p;
  int32_t open_p;
  int64_t y_p, y_pb;
  int32_t x_f32 = 0, y_f32 = 0;  // FP32 path only: X / Y hold FP32 (the pointers are float*)
};
void double_layer_f64(const DLDesc& g, const double* X, const double* A, double* Y,
                      hipStream_t s);
// The same contraction with FP32 operands, products and accumulation on the
// matrix cores (v_mfma_f32_16x16x4f32): the mixed-precision path.  X and Y
// are FP64 unless g.x_f32 / g.y_f32 say FP32.
void double_layer_f32(const DLDesc& g, const double* X, const double* A, double* Y,
                      hipStream_t s);

// ---------------------------------------------------------------------------
// small dense linear algebra (one workgroup per problem)
// ---------------------------------------------------------------------------
// Gaussian sketch (Philox4x32-10 + Box-Muller), n x q row-major, Philox
// stream ``stream`` (+ stream_base[0], read on the device, when given).
void gaussian_sketch(double* out, int64_t n, int64_t q, uint64_t seed, uint64_t stream,
                     hipStream_t s, const uint64_t* stream_base = nullptr);

// Cholesky-based orthonormalisation helper: given the q x q Gram matrix G
// (row-major, ld=q), write the upper-triangular inverse factor Rinv (q x q,
// row-major) with G + shift*I = R^T R.  Columns whose pivot falls below
// rank_tol * max(diag G) are marked dead: their rows/cols of Rinv are zero.
// ``only_if_dead`` (device, q ints): when given and all zero, the kernel
// skips the factorisation and writes Rinv = I (completion passes).
void chol_rinv(const double* G, double* Rinv, int q, double shift_rel, double rank_tol,
               hipStream_t s, double* gscratch, int* dead_out = nullptr,
               const int* only_if_dead = nullptr);
// Replace the columns flagged in ``dead`` (device, q ints) of Y (m x q,
// row-major) by N(0,1) Philox vectors.
void refill_dead_columns(double* Y, int64_t m, int q, const int* dead, uint64_t seed,
                         hipStream_t s);
// Zero the columns of Y (m x q) not flagged in ``dead`` (device, q + 1 ints,
// dead[q] = any); a no-op when none is flagged.
void keep_dead_columns(double* Y, int64_t m, int q, const int* dead, hipStream_t s);
// still[j] <- still[j] && dead[j], still[q] <- any (all zero when dead[q] is 0).
void still_dead(int* still, const int* dead, int q, hipStream_t s);

// Symmetric eigen-decomposition by cyclic parallel Jacobi.  G (q x q,
// row-major) is overwritten; eigenvalues (descending) go to ``w`` and
// eigenvectors (columns, row-major q x q) to ``V``.  ``rot`` is scratch for
// the recorded rotations (see jacobi_rot_elems).
// ``tol``: a pair is rotated while |g_pq| > tol * sqrt(g_pp g_qq); 2.2e-16
// gives a converged decomposition, a looser value a preconditioning one.
void jacobi_eigh(double* G, double* w, double* V, double* rot, int q, hipStream_t s,
                 double tol = 2.2e-16);
size_t jacobi_rot_elems(int q);

// One-sided (Hestenes) Jacobi SVD of the rows x cols matrix M whose column j
// is stored contiguously at Mc + j * ldc (M(i, j) = Mc[j * ldc + i]),
// max(rows, cols) <= 2048.  Writes the rows leading singular values,
// descending, to sig (zeros past the rank) and the matching left singular
// vectors as the columns of U (row-major rows x rows, U[i * rows + k]).
// ``scratch`` holds jacobi_svd_scratch_elems(rows, cols) doubles; its first
// two ints receive the sweep count and a converged flag.
void jacobi_svd(const double* Mc, int rows, int cols, int64_t ldc, double* sig, double* U,
                double* scratch, hipStream_t st);
size_t jacobi_svd_scratch_elems(int rows, int cols);

// Column scaling / projector finishing kernels.
// X[i, z] = Ub[i, z] * f(s[z]) for z < cD, with f selected by ``mode``:
//   mode 0: (s[z]/s0)^-1/2          (P1 factor)
//   mode 1: 1 / (s[z] * sqrt(s[z]/s0)) (P2 factor through V = B^T U S^-1)
// With ``kept`` (a device count), the columns z >= kept[0] are written as zero.
void scale_columns(const double* Ub, int q, const double* s, int cD, int mode, double* X,
                   hipStream_t st, const int* kept = nullptr);
// cD_new = min(chi, #{s/s0 > cutoff}) with s = sqrt(max(w, 0)).
void count_kept(const double* w, int q, int chi, double cutoff, double* s_out, int* cD_out,
                hipStream_t st);
// Sign gauge of the projector pair (shared with the oracle).
void fix_projector_gauge(double* P1, double* P2, int64_t rows1, int64_t rows2, int cD,
                         hipStream_t st);
// x /= ||x||_F over n elements (two-pass, deterministic).
void frobenius_normalize(double* x, int64_t n, double* scratch, hipStream_t st);
// out[i] = sum of squares partials; used by tests for norms.
void zero_fill(double* x, int64_t n, hipStream_t st);

}  // namespace acetn
