// Host-side plumbing shared by the engine's translation units: the planning
// context (workspace sizing and carving), GEMM dispatch by precision, site
// checks and the C-ABI error guard.  Nothing here crosses the C-ABI.
#pragma once
#include "internal.h"
#include "../../include/acetn_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace acetn {

inline thread_local std::string g_last_error;

struct Ctx {
  char* base = nullptr;
  size_t cap = 0, off = 0;
  bool run = false;     // buffers are carved out of the workspace
  bool launch = false;  // kernels are enqueued (false while replaying an earlier phase)
  hipStream_t st = nullptr;
  double* gwork = nullptr;
  size_t gwork_elems = 0;
  size_t gwork_need = 0;
  bool mixed = false;  // ACETN_HIP_MIXED: large products on split-bf16 matrix cores
  bool recover = true; // range finder re-derives the columns CholeskyQR drops
  int split = 0;       // split level of the products enqueued now (0 = FP64)

  double* take(size_t n) {
    const size_t bytes = ((n * sizeof(double) + 255) / 256) * 256;
    if (!run) { off += bytes; return nullptr; }
    ACETN_REQUIRE(off + bytes <= cap, "workspace too small");
    double* p = reinterpret_cast<double*>(base + off);
    off += bytes;
    return p;
  }
  int* take_int(size_t n) { return reinterpret_cast<int*>(take((n + 1) / 2)); }
};

// Run ``body`` in measure mode and return the bytes it needs (gemm split-K
// scratch first, then the carved buffers).
template <class F>
inline size_t measure(F&& body) {
  Ctx c;
  body(c);
  return c.off + ((c.gwork_need * sizeof(double) + 255) / 256) * 256 + 256;
}

template <class F>
inline void execute(void* ws, size_t ws_bytes, void* stream, F&& body) {
  Ctx m;
  body(m);  // sizes
  Ctx c;
  c.run = true;
  c.launch = true;
  c.st = reinterpret_cast<hipStream_t>(stream);
  uintptr_t b = reinterpret_cast<uintptr_t>(ws);
  uintptr_t aligned = (b + 255) & ~uintptr_t(255);
  const size_t lost = aligned - b;
  ACETN_REQUIRE(ws != nullptr || (m.off == 0 && m.gwork_need == 0), "workspace is null");
  ACETN_REQUIRE(ws_bytes >= lost, "workspace too small");
  c.base = reinterpret_cast<char*>(aligned);
  c.cap = ws_bytes - lost;
  c.gwork_elems = m.gwork_need;
  if (m.gwork_need) c.gwork = c.take(m.gwork_need);
  body(c);
}

// Products issued inside a SplitScope run on the split-bf16 matrix cores at
// the given level when the context is in mixed precision, else in FP64.
struct SplitScope {
  Ctx& c;
  int saved;
  SplitScope(Ctx& ctx, int level) : c(ctx), saved(ctx.split) { c.split = c.mixed ? level : 0; }
  ~SplitScope() { c.split = saved; }
};

inline void gemm(Ctx& c, const GemmDesc& g, const double* A, const double* B, double* C) {
  if (!c.run) {
    c.gwork_need = std::max({c.gwork_need, gemm_f64_work_elems(g, 1), gemm_split_work_elems(g, 1)});
    return;
  }
  if (!c.launch) return;
  GemmBatch b{};
  b.count = 1;
  b.A[0] = A; b.B[0] = B; b.C[0] = C;
  if (c.split > 0)
    gemm_split(g, b, c.split, DType::F64, DType::F64, DType::F64, c.gwork, c.gwork_elems, c.st);
  else
    gemm_f64(g, b, c.gwork, c.gwork_elems, c.st);
}

// dense row-major helpers
inline void mm(Ctx& c, int M, int N, int K, bool ta, const double* A, int64_t lda, bool tb,
        const double* B, int64_t ldb, double* C, int64_t ldc, double alpha = 1.0,
        double beta = 0.0) {
  gemm(c, dense_desc(M, N, K, ta, lda, tb, ldb, ldc, alpha, beta), A, B, C);
}

inline void check_site(const acetn_site* s) {
  ACETN_REQUIRE(s != nullptr && s->A != nullptr, "site is null");
  ACETN_REQUIRE(s->D >= 1 && s->D <= 8 && s->d >= 1 && s->d <= 8, "unsupported D/d");
  for (int k = 0; k < 4; ++k) {
    ACETN_REQUIRE(s->C[k] && s->E[k], "site tensor pointer is null");
    ACETN_REQUIRE(s->C_dim[k][0] > 0 && s->C_dim[k][1] > 0 && s->E_dim[k][0] > 0 &&
                      s->E_dim[k][1] > 0, "non-positive boundary dimension");
  }
}

// strides of bond_permute(A, k) axis j (A contiguous (D,D,D,D,d))
inline void permuted_strides(const acetn_site* s, int k, int64_t out[4]) {
  const int64_t D = s->D, d = s->d;
  const int64_t as[4] = {D * D * D * d, D * D * d, D * d, d};
  for (int j = 0; j < 4; ++j) out[j] = as[(j + k) % 4];
}


inline GemmDesc gdesc(int M, int N, int K, IdxMap am, IdxMap ak, IdxMap bk, IdxMap bn, IdxMap cm,
               IdxMap cn) {
  GemmDesc g{};
  g.M = M; g.N = N; g.K = K;
  g.am = am; g.ak = ak; g.bk = bk; g.bn = bn; g.cm = cm; g.cn = cn;
  g.alpha = 1.0; g.beta = 0.0;
  return g;
}

inline void check_precision(int p) {
  const int base = p & ~ACETN_HIP_RF_FAST;
  ACETN_REQUIRE(base == ACETN_HIP_FP64 || base == ACETN_HIP_MIXED, "unknown precision mode");
}
// precision argument -> context: the arithmetic of the large products and
// whether the range finder recovers dropped columns (ACETN_HIP_RF_FAST off)
inline void set_precision(Ctx& c, int p) {
  c.mixed = (p & ~ACETN_HIP_RF_FAST) == ACETN_HIP_MIXED;
  c.recover = (p & ACETN_HIP_RF_FAST) == 0;
}

template <class F>
inline int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 3;
  } catch (...) {
    g_last_error = "unknown error";
    return 3;
  }
}


// Halves of the bond network shared by the bond reduced density matrix and
// the full-update norm tensor (ctmrg_engine.hip).
struct BondHalves {
  double* r1;   // [F][nc][L][P][l][p]
  double* r2;   // [F][nc][R][Q][r][q]
  int64_t F;    // closing boundary index range
  int64_t nc;   // shared boundary index range
};
void bond_halves(Ctx& c, const acetn_site* s1, const acetn_site* s2, int k, const double* A1,
                 const int64_t ps1[4], const double* A2, const int64_t ps2[4], BondHalves& out);

}  // namespace acetn
