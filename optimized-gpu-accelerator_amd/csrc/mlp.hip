// Learned warm start for parametric LP families: a small MLP maps the family
// parameter theta to coefficients on a POD basis of PDHG solutions, and the
// basis decode z = mean + U a gives the starting point of a warm-started
// solve.  Forward, backward, Adam and decode are hand-written HIP for gfx950:
// every matrix product runs on gfx950 BF16 MFMA tiles
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; master weights, activations and optimiser state
// are FP32.  No reference counterpart: the reference has no warm start
// (x0, y0 are hard-wired, src/pdhg/pdhg.cpp:509,519).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pdhg_hip.h"
#include "core.h"

namespace pdhgamd {

int pod_basis(const double* Z, int k, int d, int rank, double* mean, double* basis, double* coeffs, double* energy);

#define MLP_HIP_OK(expr)                                                                            \
  do {                                                                                              \
    hipError_t _e = (expr);                                                                         \
    if (_e != hipSuccess) throw ::pdhgamd::Error(std::string("HIP error '") + hipGetErrorString(_e) + "' at " #expr); \
  } while (0)

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round-to-nearest-even on the bit pattern (NaN kept quiet).
__device__ inline __bf16 to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
    return __builtin_bit_cast(__bf16, static_cast<unsigned short>((u >> 16) | 0x40));
  u += 0x7fffu + ((u >> 16) & 1u);
  return __builtin_bit_cast(__bf16, static_cast<unsigned short>(u >> 16));
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
// d/dz [z Phi(z)] = Phi(z) + z phi(z).
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.3989422804014327f * expf(-0.5f * z * z);
}

struct Gemm {
  int M, N, K;
  const float* A;  // A(m, k) = A[m * sam + k * sak]
  long sam, sak;
  const float* B;  // B(k, n) = B[k * sbk + n * sbn]
  long sbk, sbn;
  float* C;        // C[m * ldc + n]
  long ldc;
  const float* bias;  // + bias[n]
  float* zOut;        // pre-activation copy (ldc layout)
  const float* gradOf;  // multiply by gelu'(gradOf[m * ldc + n])
  int act;              // 1: GELU
};

// One wavefront per 16 x 16 output tile; K consumed 32 at a time by one
// gfx950 MFMA (v_mfma_f32_16x16x32_bf16).  Lane l feeds A row (l % 16) and
// B column (l % 16) at k = 8 (l / 16) + i, i < 8, and receives C rows
// 4 (l / 16) + i, column l % 16.  Out-of-range operands load as zero, so any
// M, N, K works.
__global__ __launch_bounds__(64) void k_gemm_bf16(Gemm g) {
  const int lane = threadIdx.x;
  const int n0 = blockIdx.x * 16, m0 = blockIdx.y * 16;
  const int r = lane & 15, q = lane >> 4;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + 8 * q + i;
      const int m = m0 + r, n = n0 + r;
      const float av = (m < g.M && k < g.K) ? g.A[m * g.sam + k * g.sak] : 0.f;
      const float bv = (n < g.N && k < g.K) ? g.B[k * g.sbk + n * g.sbn] : 0.f;
      a[i] = to_bf16(av);
      b[i] = to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * q + i, n = n0 + r;
    if (m >= g.M || n >= g.N) continue;
    float v = acc[i];
    if (g.bias) v += g.bias[n];
    const long at = m * g.ldc + n;
    if (g.zOut) g.zOut[at] = v;
    if (g.gradOf) v *= gelu_grad(g.gradOf[at]);
    if (g.act == 1) v = gelu(v);
    g.C[at] = v;
  }
}

// dY = 2 (Y - T) / (N out), and sum (Y - T)^2 / (N out) into *loss (one block).
__global__ __launch_bounds__(256) void k_mse(const float* Y, const float* T, float* dY, int len, float* loss) {
  __shared__ float sh[256];
  float s = 0;
  const float inv = 1.0f / static_cast<float>(len);
  for (int t = threadIdx.x; t < len; t += 256) {
    const float d = Y[t] - T[t];
    dY[t] = 2.0f * d * inv;
    s += d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = sh[0] * inv;
}

// out[n] = sum_m D[m * ld + n] (bias gradients).
__global__ void k_colsum(const float* D, int M, int N, long ld, float* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0;
  for (int m = 0; m < M; ++m) s += D[m * ld + n];
  out[n] = s;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* p, const float* g, float* m1, float* m2, int len, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= len) return;
  const float gt = g[t];
  const float a = b1 * m1[t] + (1.0f - b1) * gt;
  const float v = b2 * m2[t] + (1.0f - b2) * gt * gt;
  m1[t] = a;
  m2[t] = v;
  p[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

template <typename U>
U* dalloc(size_t n) {
  void* p = nullptr;
  MLP_HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(U)));
  return static_cast<U*>(p);
}

}  // namespace

// in -> hid (GELU) -> hid (GELU) -> out.  Parameters in one FP32 buffer:
// W1[hid x in], b1[hid], W2[hid x hid], b2[hid], W3[out x hid], b3[out].
class Mlp {
 public:
  Mlp(int in, int hid, int out, uint64_t seed, int device) : in_(in), hid_(hid), out_(out), dev_(device) {
    PDHG_REQUIRE(in >= 1 && hid >= 1 && out >= 1, "mlp: bad sizes");
    MLP_HIP_OK(hipSetDevice(device));
    MLP_HIP_OK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    off_[0] = 0;
    off_[1] = off_[0] + hid * in;
    off_[2] = off_[1] + hid;
    off_[3] = off_[2] + hid * hid;
    off_[4] = off_[3] + hid;
    off_[5] = off_[4] + out * hid;
    np_ = off_[5] + out;
    p_ = dalloc<float>(np_);
    g_ = dalloc<float>(np_);
    m1_ = dalloc<float>(np_);
    m2_ = dalloc<float>(np_);
    loss_ = dalloc<float>(1);
    // Glorot-uniform weights, zero biases (SplitMix64 stream).
    Rng rng(seed);
    std::vector<float> h(np_, 0.0f);
    auto fill = [&](int off, int fo, int fi) {
      const double a = std::sqrt(6.0 / (fi + fo));
      for (int t = 0; t < fo * fi; ++t) h[off + t] = static_cast<float>(a * (2.0 * rng.uniform() - 1.0));
    };
    fill(off_[0], hid, in);
    fill(off_[2], hid, hid);
    fill(off_[4], out, hid);
    MLP_HIP_OK(hipMemcpyAsync(p_, h.data(), np_ * sizeof(float), hipMemcpyHostToDevice, s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
    reset_adam();
  }
  ~Mlp() {
    for (float* q : {p_, g_, m1_, m2_, loss_}) (void)hipFree(q);
    free_batch();
    (void)hipStreamDestroy(s_);
  }
  int params() const { return np_; }
  // Parameter copies are ordered on the network's own stream, behind any
  // training step still queued there.
  void get(double* out) {
    std::vector<float> h(np_);
    MLP_HIP_OK(hipMemcpyAsync(h.data(), p_, np_ * sizeof(float), hipMemcpyDeviceToHost, s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
    for (int t = 0; t < np_; ++t) out[t] = h[t];
  }
  void set(const double* in) {
    std::vector<float> h(in, in + np_);
    MLP_HIP_OK(hipMemcpyAsync(p_, h.data(), np_ * sizeof(float), hipMemcpyHostToDevice, s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
  }
  void forward(const double* X, int N, double* Y) {
    MLP_HIP_OK(hipSetDevice(dev_));
    load_inputs(X, nullptr, N);
    run_forward(N);
    std::vector<float> h(static_cast<size_t>(N) * out_);
    MLP_HIP_OK(hipMemcpyAsync(h.data(), Y_, h.size() * sizeof(float), hipMemcpyDeviceToHost, s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
    for (size_t t = 0; t < h.size(); ++t) Y[t] = h[t];
  }
  double loss_grad(const double* X, const double* T, int N, double* grad) {
    MLP_HIP_OK(hipSetDevice(dev_));
    load_inputs(X, T, N);
    run_forward(N);
    run_backward(N);
    float l = 0;
    MLP_HIP_OK(hipMemcpyAsync(&l, loss_, sizeof(float), hipMemcpyDeviceToHost, s_));
    std::vector<float> h(np_);
    MLP_HIP_OK(hipMemcpyAsync(h.data(), g_, np_ * sizeof(float), hipMemcpyDeviceToHost, s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
    if (grad)
      for (int t = 0; t < np_; ++t) grad[t] = h[t];
    return l;
  }
  void train(const double* X, const double* T, int N, int steps, double lr, double* losses) {
    MLP_HIP_OK(hipSetDevice(dev_));
    load_inputs(X, T, N);
    const float b1 = 0.9f, b2 = 0.999f;
    std::vector<float> hl(steps);
    for (int it = 0; it < steps; ++it) {
      run_forward(N);
      run_backward(N);
      ++adamStep_;
      const float c1 = 1.0f - std::pow(b1, static_cast<float>(adamStep_));
      const float c2 = 1.0f - std::pow(b2, static_cast<float>(adamStep_));
      hipLaunchKernelGGL(k_adam, dim3((np_ + 255) / 256), dim3(256), 0, s_, p_, g_, m1_, m2_, np_,
                         static_cast<float>(lr), b1, b2, c1, c2, 1e-8f);
      if (losses) MLP_HIP_OK(hipMemcpyAsync(&hl[it], loss_, sizeof(float), hipMemcpyDeviceToHost, s_));
    }
    MLP_HIP_OK(hipStreamSynchronize(s_));
    MLP_HIP_OK(hipGetLastError());
    if (losses)
      for (int it = 0; it < steps; ++it) losses[it] = hl[it];
  }
  void reset_adam() {
    MLP_HIP_OK(hipMemsetAsync(m1_, 0, np_ * sizeof(float), s_));
    MLP_HIP_OK(hipMemsetAsync(m2_, 0, np_ * sizeof(float), s_));
    MLP_HIP_OK(hipStreamSynchronize(s_));
    adamStep_ = 0;
  }

 private:
  void gemm(int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk, long sbn, float* C,
            long ldc, const float* bias, float* zOut, const float* gradOf, int act) {
    Gemm g{M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, zOut, gradOf, act};
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + 15) / 16, (M + 15) / 16), dim3(64), 0, s_, g);
  }
  void free_batch() {
    for (float* q : {X_, T_, Z1_, H1_, Z2_, H2_, Y_, dY_, dH_, dH1_}) (void)hipFree(q);
    X_ = T_ = Z1_ = H1_ = Z2_ = H2_ = Y_ = dY_ = dH_ = dH1_ = nullptr;
    cap_ = 0;
  }
  void ensure(int N) {
    if (N <= cap_) return;
    free_batch();
    cap_ = N;
    X_ = dalloc<float>(static_cast<size_t>(N) * in_);
    T_ = dalloc<float>(static_cast<size_t>(N) * out_);
    Z1_ = dalloc<float>(static_cast<size_t>(N) * hid_);
    H1_ = dalloc<float>(static_cast<size_t>(N) * hid_);
    Z2_ = dalloc<float>(static_cast<size_t>(N) * hid_);
    H2_ = dalloc<float>(static_cast<size_t>(N) * hid_);
    Y_ = dalloc<float>(static_cast<size_t>(N) * out_);
    dY_ = dalloc<float>(static_cast<size_t>(N) * out_);
    dH_ = dalloc<float>(static_cast<size_t>(N) * hid_);
    dH1_ = dalloc<float>(static_cast<size_t>(N) * hid_);
  }
  void load_inputs(const double* X, const double* T, int N) {
    PDHG_REQUIRE(N >= 1, "mlp: empty batch");
    ensure(N);
    std::vector<float> hx(X, X + static_cast<size_t>(N) * in_);
    MLP_HIP_OK(hipMemcpyAsync(X_, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice, s_));
    if (T) {
      std::vector<float> ht(T, T + static_cast<size_t>(N) * out_);
      MLP_HIP_OK(hipMemcpyAsync(T_, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice, s_));
      MLP_HIP_OK(hipStreamSynchronize(s_));
    } else {
      MLP_HIP_OK(hipStreamSynchronize(s_));
    }
  }
  void run_forward(int N) {
    const float *W1 = p_ + off_[0], *b1 = p_ + off_[1], *W2 = p_ + off_[2], *b2 = p_ + off_[3];
    const float *W3 = p_ + off_[4], *b3 = p_ + off_[5];
    // Z1 = X W1^T + b1, H1 = gelu(Z1)
    gemm(N, hid_, in_, X_, in_, 1, W1, 1, in_, H1_, hid_, b1, Z1_, nullptr, 1);
    gemm(N, hid_, hid_, H1_, hid_, 1, W2, 1, hid_, H2_, hid_, b2, Z2_, nullptr, 1);
    gemm(N, out_, hid_, H2_, hid_, 1, W3, 1, hid_, Y_, out_, b3, nullptr, nullptr, 0);
  }
  void run_backward(int N) {
    const float *W2 = p_ + off_[2], *W3 = p_ + off_[4];
    float *gW1 = g_ + off_[0], *gb1 = g_ + off_[1], *gW2 = g_ + off_[2], *gb2 = g_ + off_[3];
    float *gW3 = g_ + off_[4], *gb3 = g_ + off_[5];
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, s_, Y_, T_, dY_, N * out_, loss_);
    // dW3 = dY^T H2, db3 = colsum(dY)
    gemm(out_, hid_, N, dY_, 1, out_, H2_, hid_, 1, gW3, hid_, nullptr, nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3((out_ + 63) / 64), dim3(64), 0, s_, dY_, N, out_, static_cast<long>(out_), gb3);
    // dZ2 = (dY W3) * gelu'(Z2)
    gemm(N, hid_, out_, dY_, out_, 1, W3, hid_, 1, dH_, hid_, nullptr, nullptr, Z2_, 0);
    gemm(hid_, hid_, N, dH_, 1, hid_, H1_, hid_, 1, gW2, hid_, nullptr, nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3((hid_ + 63) / 64), dim3(64), 0, s_, dH_, N, hid_, static_cast<long>(hid_), gb2);
    // dZ1 = (dZ2 W2) * gelu'(Z1)
    gemm(N, hid_, hid_, dH_, hid_, 1, W2, hid_, 1, dH1_, hid_, nullptr, nullptr, Z1_, 0);
    gemm(hid_, in_, N, dH1_, 1, hid_, X_, in_, 1, gW1, in_, nullptr, nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3((hid_ + 63) / 64), dim3(64), 0, s_, dH1_, N, hid_, static_cast<long>(hid_), gb1);
  }

  int in_, hid_, out_, dev_, np_ = 0;
  int off_[6];
  hipStream_t s_ = nullptr;
  float *p_ = nullptr, *g_ = nullptr, *m1_ = nullptr, *m2_ = nullptr, *loss_ = nullptr;
  float *X_ = nullptr, *T_ = nullptr, *Z1_ = nullptr, *H1_ = nullptr, *Z2_ = nullptr, *H2_ = nullptr;
  float *Y_ = nullptr, *dY_ = nullptr, *dH_ = nullptr, *dH1_ = nullptr;
  int cap_ = 0;
  long long adamStep_ = 0;
};

// out (N x d) = mean + coeffs (N x r) U^T with U (d x r): the basis decode,
// BF16 MFMA with FP32 accumulation.
void decode_bf16(const double* basis, const double* mean, int d, int r, const double* coeffs, int N, double* out,
                 int device) {
  MLP_HIP_OK(hipSetDevice(device));
  std::vector<float> hU(basis, basis + static_cast<size_t>(d) * r), hM(mean, mean + d),
      hA(coeffs, coeffs + static_cast<size_t>(N) * r);
  float *U = dalloc<float>(hU.size()), *M = dalloc<float>(d), *A = dalloc<float>(hA.size());
  float* C = dalloc<float>(static_cast<size_t>(N) * d);
  MLP_HIP_OK(hipMemcpy(U, hU.data(), hU.size() * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP_OK(hipMemcpy(M, hM.data(), hM.size() * sizeof(float), hipMemcpyHostToDevice));
  MLP_HIP_OK(hipMemcpy(A, hA.data(), hA.size() * sizeof(float), hipMemcpyHostToDevice));
  Gemm g{N, d, r, A, r, 1, U, 1, r, C, d, M, nullptr, nullptr, 0};
  hipLaunchKernelGGL(k_gemm_bf16, dim3((d + 15) / 16, (N + 15) / 16), dim3(64), 0, 0, g);
  MLP_HIP_OK(hipGetLastError());
  std::vector<float> hC(static_cast<size_t>(N) * d);
  MLP_HIP_OK(hipMemcpy(hC.data(), C, hC.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t t = 0; t < hC.size(); ++t) out[t] = hC[t];
  for (float* q : {U, M, A, C}) (void)hipFree(q);
}

}  // namespace pdhgamd

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
struct pdhg_mlp {
  std::unique_ptr<pdhgamd::Mlp> impl;
};

extern "C" {

void pdhg_set_last_error(const char* msg);

#define MLP_GUARD(body)                          \
  try {                                          \
    body;                                        \
    return 0;                                    \
  } catch (const std::exception& e) {            \
    pdhg_set_last_error(e.what());               \
    return -1;                                   \
  }

pdhg_mlp* pdhg_mlp_create(int in, int hidden, int out, uint64_t seed, int device) {
  try {
    if (!pdhg_gpu_available()) throw pdhgamd::Error("GPU unavailable: the MLP runs on the GPU only");
    auto* h = new pdhg_mlp;
    h->impl = std::make_unique<pdhgamd::Mlp>(in, hidden, out, seed, device);
    return h;
  } catch (const std::exception& e) {
    pdhg_set_last_error(e.what());
    return nullptr;
  }
}
void pdhg_mlp_destroy(pdhg_mlp* h) { delete h; }
int pdhg_mlp_param_count(const pdhg_mlp* h) { return h ? h->impl->params() : -1; }
int pdhg_mlp_get_params(pdhg_mlp* h, double* p) { MLP_GUARD(h->impl->get(p)); }
int pdhg_mlp_set_params(pdhg_mlp* h, const double* p) { MLP_GUARD(h->impl->set(p)); }
int pdhg_mlp_forward(pdhg_mlp* h, const double* X, int N, double* Y) { MLP_GUARD(h->impl->forward(X, N, Y)); }
int pdhg_mlp_loss_grad(pdhg_mlp* h, const double* X, const double* T, int N, double* loss, double* grad) {
  MLP_GUARD(*loss = h->impl->loss_grad(X, T, N, grad));
}
int pdhg_mlp_train(pdhg_mlp* h, const double* X, const double* T, int N, int steps, double lr, double* losses) {
  MLP_GUARD(h->impl->train(X, T, N, steps, lr, losses));
}
int pdhg_pod_basis(const double* Z, int k, int d, int rank, double* mean, double* basis, double* coeffs,
                   double* energy) {
  MLP_GUARD(return pdhgamd::pod_basis(Z, k, d, rank, mean, basis, coeffs, energy));
}
int pdhg_decode_bf16(const double* basis, const double* mean, int d, int rank, const double* coeffs, int N,
                     double* out) {
  MLP_GUARD(pdhgamd::decode_bf16(basis, mean, d, rank, coeffs, N, out, 0));
}

}  // extern "C"
