// The learned ALS start: a 3-layer perceptron (Fin -> H -> H -> Fout, ReLU)
// with its own forward, backward and Adam kernels.  Weights are kept in FP32
// (master copy) and enter every product as bf16 (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation); activations are stored in FP32.
//
// Parameter layout (one flat FP32 buffer, also used for gradients and the
// two Adam moments): W1 (Fin x H), b1 (H), W2 (H x H), b2 (H), W3 (H x Fout),
// b3 (Fout); every matrix row-major [in][out].
#include "engine.h"

namespace acetn {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int MBM = 64, MBN = 64, MBK = 32, MNT = 256, MLK = MBK + 8;

// C (M x N, ldc) = op(A) op(B) [+ C] [+ bias[n]] [ReLU], FP32 in and out,
// operands rounded to bf16 on their way into LDS.
//   op(A)(m,k) = TA ? A[k*lda + m] : A[m*lda + k]
//   op(B)(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]
// mode: 0 store, 1 accumulate into C, 2 + bias, 3 + bias then ReLU.
template <bool TA, bool TB>
__global__ __launch_bounds__(MNT) void mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A,
                                                       int64_t lda, const float* __restrict__ B,
                                                       int64_t ldb, float* __restrict__ C, int64_t ldc,
                                                       const float* __restrict__ bias, int mode) {
  __shared__ __align__(16) __bf16 As[MBM][MLK];
  __shared__ __align__(16) __bf16 Bs[MBN][MLK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * MBM, n0 = blockIdx.x * MBN;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += MBK) {
    __syncthreads();
    for (int e = tid; e < MBM * MBK; e += MNT) {
      int ml, kl;
      if (TA) { ml = e % MBM; kl = e / MBM; } else { kl = e % MBK; ml = e / MBK; }
      const int m = m0 + ml, k = k0 + kl;
      float v = 0.f;
      if (m < M && k < K) v = TA ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
      As[ml][kl] = (__bf16)v;
    }
    for (int e = tid; e < MBN * MBK; e += MNT) {
      int nl, kl;
      if (TB) { kl = e % MBK; nl = e / MBK; } else { nl = e % MBN; kl = e / MBN; }
      const int n = n0 + nl, k = k0 + kl;
      float v = 0.f;
      if (n < N && k < K) v = TB ? B[(int64_t)n * ldb + k] : B[(int64_t)k * ldb + n];
      Bs[nl][kl] = (__bf16)v;
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    bf8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      a[i] = *reinterpret_cast<const bf8*>(&As[wm * 32 + i * 16 + (lane & 15)][kq]);
#pragma unroll
    for (int j = 0; j < 2; ++j)
      b[j] = *reinterpret_cast<const bf8*>(&Bs[wn * 32 + j * 16 + (lane & 15)][kq]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[i][j][r];
        float* c = C + (int64_t)m * ldc + n;
        if (mode == 1) v += *c;
        if (mode >= 2) v += bias[n];
        if (mode == 3) v = fmaxf(v, 0.f);
        *c = v;
      }
}

// dH *= (H > 0)
__global__ void relu_mask_kernel(float* __restrict__ dH, const float* __restrict__ H, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    if (!(H[e] > 0.f)) dH[e] = 0.f;
}

// db[n] = sum_b dY[b][n]
__global__ void colsum_kernel(const float* __restrict__ dY, int Bn, int N, float* __restrict__ db) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < Bn; ++b) s += dY[(int64_t)b * N + n];
    db[n] = s;
  }
}

// loss partials and dY = 2 (Y - T) / (B*Fout) over rows with weight 1
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, int64_t n,
                           float scale, float* __restrict__ dY, float* __restrict__ part) {
  __shared__ float red[4];
  float s = 0.f;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float r = Y[e] - T[e];
    s += r * r;
    dY[e] = 2.f * r * scale;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1] + red[2] + red[3]) * scale;
}

__global__ void sum_kernel(const float* __restrict__ part, int n, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += part[i];
    out[0] = s;
  }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float wd, float c1, float c2) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[e] + wd * p[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

struct MlpShape {
  int64_t fin, h, fout;
  int64_t w1() const { return 0; }
  int64_t b1() const { return fin * h; }
  int64_t w2() const { return b1() + h; }
  int64_t b2() const { return w2() + h * h; }
  int64_t w3() const { return b2() + h; }
  int64_t b3() const { return w3() + h * fout; }
  int64_t total() const { return b3() + fout; }
};

void mgemm(bool ta, bool tb, int M, int N, int K, const float* A, int64_t lda, const float* B,
           int64_t ldb, float* C, int64_t ldc, const float* bias, int mode, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + MBN - 1) / MBN, (M + MBM - 1) / MBM);
  if (!ta && !tb)
    hipLaunchKernelGGL((mlp_gemm_kernel<false, false>), grid, dim3(MNT), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, mode);
  else if (ta && !tb)
    hipLaunchKernelGGL((mlp_gemm_kernel<true, false>), grid, dim3(MNT), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, mode);
  else if (!ta && tb)
    hipLaunchKernelGGL((mlp_gemm_kernel<false, true>), grid, dim3(MNT), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, mode);
  else
    hipLaunchKernelGGL((mlp_gemm_kernel<true, true>), grid, dim3(MNT), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, mode);
  ACETN_HIP_CHECK(hipGetLastError());
}

MlpShape mlp_shape(int64_t fin, int64_t h, int64_t fout) {
  ACETN_REQUIRE(fin >= 1 && h >= 1 && fout >= 1 && fin <= (1 << 24) && h <= 4096 && fout <= 65536,
                "mlp: unsupported layer sizes");
  return MlpShape{fin, h, fout};
}

}  // namespace
}  // namespace acetn

using namespace acetn;

extern "C" {

int64_t acetn_hip_mlp_param_count(int64_t fin, int64_t hidden, int64_t fout) {
  int64_t r = -1;
  guarded([&] { r = mlp_shape(fin, hidden, fout).total(); });
  return r;
}

size_t acetn_hip_mlp_workspace(int64_t fin, int64_t hidden, int64_t fout, int64_t batch) {
  (void)fin;
  (void)fout;
  // H1, H2, dH1, dH2 (batch x hidden each) + loss partials
  return sizeof(float) * (4 * (size_t)batch * hidden + 1024) + 256;
}

int acetn_hip_mlp_forward(const float* params, int64_t fin, int64_t hidden, int64_t fout,
                          const float* X, int64_t batch, float* Y, void* ws, size_t ws_bytes,
                          void* stream) {
  return guarded([&] {
    const MlpShape sh = mlp_shape(fin, hidden, fout);
    ACETN_REQUIRE(params && X && Y && ws, "mlp_forward: null pointer");
    ACETN_REQUIRE(ws_bytes >= acetn_hip_mlp_workspace(fin, hidden, fout, batch),
                  "mlp_forward: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    float* H1 = reinterpret_cast<float*>(ws);
    float* H2 = H1 + batch * hidden;
    const int B = (int)batch, H = (int)hidden;
    mgemm(false, false, B, H, (int)fin, X, fin, params + sh.w1(), hidden, H1, hidden,
          params + sh.b1(), 3, s);
    mgemm(false, false, B, H, H, H1, hidden, params + sh.w2(), hidden, H2, hidden,
          params + sh.b2(), 3, s);
    mgemm(false, false, B, (int)fout, H, H2, hidden, params + sh.w3(), fout, Y, fout,
          params + sh.b3(), 2, s);
  });
}

int acetn_hip_mlp_backward(const float* params, int64_t fin, int64_t hidden, int64_t fout,
                           const float* X, int64_t batch, const float* target, float* Y,
                           float* grad, float* host_loss, void* ws, size_t ws_bytes,
                           void* stream) {
  return guarded([&] {
    const MlpShape sh = mlp_shape(fin, hidden, fout);
    ACETN_REQUIRE(params && X && target && Y && grad && ws, "mlp_backward: null pointer");
    ACETN_REQUIRE(ws_bytes >= acetn_hip_mlp_workspace(fin, hidden, fout, batch),
                  "mlp_backward: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int B = (int)batch, H = (int)hidden, Fo = (int)fout, Fi = (int)fin;
    float* H1 = reinterpret_cast<float*>(ws);
    float* H2 = H1 + batch * hidden;
    float* dH2 = H2 + batch * hidden;
    float* dH1 = dH2 + batch * hidden;
    float* part = dH1 + batch * hidden;
    // forward (keeps H1, H2)
    mgemm(false, false, B, H, Fi, X, fin, params + sh.w1(), hidden, H1, hidden, params + sh.b1(), 3, s);
    mgemm(false, false, B, H, H, H1, hidden, params + sh.w2(), hidden, H2, hidden, params + sh.b2(), 3, s);
    mgemm(false, false, B, Fo, H, H2, hidden, params + sh.w3(), fout, Y, fout, params + sh.b3(), 2, s);
    // loss = mean (Y - T)^2, dY = 2 (Y - T) / (B Fout), written over Y
    const int64_t ny = batch * fout;
    const int blocks = (int)std::min<int64_t>((ny + 255) / 256, 1000);
    hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, s, Y, target, ny,
                       1.f / (float)ny, Y, part);
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, part, blocks, part + 1000);
    const float* dY = Y;
    // layer 3
    mgemm(true, false, H, Fo, B, H2, hidden, dY, fout, grad + sh.w3(), fout, nullptr, 0, s);
    hipLaunchKernelGGL(colsum_kernel, dim3((Fo + 255) / 256), dim3(256), 0, s, dY, B, Fo, grad + sh.b3());
    mgemm(false, true, B, H, Fo, dY, fout, params + sh.w3(), fout, dH2, hidden, nullptr, 0, s);
    hipLaunchKernelGGL(relu_mask_kernel, dim3(std::min(B * H / 256 + 1, 1024)), dim3(256), 0, s, dH2, H2,
                       batch * hidden);
    // layer 2
    mgemm(true, false, H, H, B, H1, hidden, dH2, hidden, grad + sh.w2(), hidden, nullptr, 0, s);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, s, dH2, B, H, grad + sh.b2());
    mgemm(false, true, B, H, H, dH2, hidden, params + sh.w2(), hidden, dH1, hidden, nullptr, 0, s);
    hipLaunchKernelGGL(relu_mask_kernel, dim3(std::min(B * H / 256 + 1, 1024)), dim3(256), 0, s, dH1, H1,
                       batch * hidden);
    // layer 1
    mgemm(true, false, Fi, H, B, X, fin, dH1, hidden, grad + sh.w1(), hidden, nullptr, 0, s);
    hipLaunchKernelGGL(colsum_kernel, dim3((H + 255) / 256), dim3(256), 0, s, dH1, B, H, grad + sh.b1());
    ACETN_HIP_CHECK(hipGetLastError());
    if (host_loss) {
      ACETN_HIP_CHECK(hipMemcpyAsync(host_loss, part + 1000, sizeof(float), hipMemcpyDeviceToHost, s));
      ACETN_HIP_CHECK(hipStreamSynchronize(s));
    }
  });
}

int acetn_hip_mlp_adam_step(float* params, const float* grad, float* m, float* v, int64_t n,
                            int64_t step, float lr, float beta1, float beta2, float eps,
                            float weight_decay, void* stream) {
  return guarded([&] {
    ACETN_REQUIRE(params && grad && m && v && n > 0 && step >= 1, "mlp_adam_step: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, s, params, grad, m, v, n, lr,
                       beta1, beta2, eps, weight_decay, c1, c2);
    ACETN_HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
