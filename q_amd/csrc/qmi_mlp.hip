// Overlap predictor of the learned lambda schedule: a small MLP trained and
// evaluated on the GPU by the kernels in this file (no reference
// counterpart; the reference schedules lambda windows by hand).
//
//   x[n_in] -> Linear(H) -> SiLU -> Linear(H) -> SiLU -> Linear(1) -> sigmoid
//
// Weights are kept in BF16 for the forward and backward passes, with an FP32
// master copy for the optimiser (Adam); every product is accumulated in FP32.
// The matrix products are one LDS-tiled kernel (C = op(A) op(B), 16x16 tiles,
// one thread per output element); the batch is small, so the products are
// VALU work.  Reductions run in a fixed order: results are reproducible run
// to run.
#include <hip/hip_bfloat16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/qmi.h"

namespace {

thread_local std::string g_mlp_err;
int mlp_err(int code, const std::string& m) {
  g_mlp_err = m;
  return code;
}
#define MLP_OK(call)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (call);                                                                           \
    if (e_ != hipSuccess) return mlp_err(-100, std::string("HIP error: ") + hipGetErrorString(e_)); \
  } while (0)

// C[M x N] = alpha * op(A) * op(B) + beta * C, op(A) is M x K, op(B) is K x N.
// A and B are row-major FP32 or BF16 (TA / TB = 1: stored transposed).
template <typename TA_, typename TB_, bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, float alpha, const TA_* A, int lda, const TB_* B,
                                              int ldb, float beta, float* C, int ldc) {
  __shared__ float sa[16][17];
  __shared__ float sb[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  float acc = 0.f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    // A tile: rows blockIdx.y*16 + ty, columns k0 + tx
    {
      const int r = blockIdx.y * 16 + ty, c = k0 + tx;
      float v = 0.f;
      if (r < M && c < K) v = TA ? (float)A[(size_t)c * lda + r] : (float)A[(size_t)r * lda + c];
      sa[ty][tx] = v;
    }
    // B tile: rows k0 + ty, columns blockIdx.x*16 + tx
    {
      const int r = k0 + ty, c = blockIdx.x * 16 + tx;
      float v = 0.f;
      if (r < K && c < N) v = TB ? (float)B[(size_t)c * ldb + r] : (float)B[(size_t)r * ldb + c];
      sb[ty][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = fmaf(sa[ty][k], sb[k][tx], acc);
    __syncthreads();
  }
  if (row < M && col < N) {
    float* c = &C[(size_t)row * ldc + col];
    *c = alpha * acc + (beta != 0.f ? beta * *c : 0.f);
  }
}

// z (B x H) += bias; a = silu(z)
__global__ void k_bias_silu(int B, int H, float* z, const float* bias, float* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * H) return;
  const float v = z[i] + bias[i % H];
  z[i] = v;
  a[i] = v / (1.f + expf(-v));
}

// output layer: y = sigmoid(z3 + b3); loss terms and dz3 for MSE
__global__ void k_out(int B, float* z3, const float* b3, float* y, const float* t, float* dz3, float* lossv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const float v = z3[i] + b3[0];
  z3[i] = v;
  const float s = 1.f / (1.f + expf(-v));
  y[i] = s;
  if (t) {
    const float e = s - t[i];
    lossv[i] = e * e;
    dz3[i] = 2.f * e / (float)B * s * (1.f - s);
  }
}

// dz = da * silu'(z), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_back(int n, const float* z, const float* da, float* dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float s = 1.f / (1.f + expf(-z[i]));
  dz[i] = da[i] * s * (1.f + z[i] * (1.f - s));
}

// column sums of dz (B x H) -> db (H), fixed order over the batch
__global__ void k_colsum(int B, int H, const float* dz, float* db) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= H) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += dz[(size_t)b * H + h];
  db[h] = s;
}

// sum of per-sample squared errors, one block, fixed order
__global__ void k_sum(int n, const float* v, float* out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Adam on the FP32 master parameters, then the BF16 working copy
__global__ void k_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, hip_bfloat16* pb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  pb[i] = hip_bfloat16(pi);
}

__global__ void k_to_bf16(long long n, const float* p, hip_bfloat16* pb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = hip_bfloat16(p[i]);
}

// rows idx[0..B) of X (n_in) and T into the batch buffers
__global__ void k_gather(int B, int n_in, const int* idx, const float* X, const float* T, float* xb, float* tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * n_in) return;
  const int b = i / n_in, c = i % n_in;
  xb[i] = X[(size_t)idx[b] * n_in + c];
  if (c == 0) tb[b] = T[idx[b]];
}

inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

struct qmi_mlp {
  int device = 0;
  int n_in = 0, H = 0;
  hipStream_t s = nullptr;
  // parameters, in the order W1 (H x n_in), b1 (H), W2 (H x H), b2 (H), W3 (1 x H), b3 (1)
  long long np = 0;
  long long o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, o_w3 = 0, o_b3 = 0;
  float* p = nullptr;           // FP32 master
  hip_bfloat16* pb = nullptr;   // BF16 working copy
  float *g = nullptr, *m = nullptr, *v = nullptr;
  long long step = 0;
  // activations for a batch of up to cap rows
  int cap = 0;
  float *x = nullptr, *t = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *z3 = nullptr,
        *y = nullptr, *dz3 = nullptr, *da2 = nullptr, *dz2 = nullptr, *da1 = nullptr, *dz1 = nullptr,
        *lossv = nullptr, *loss = nullptr;
  std::vector<void*> bufs;
  ~qmi_mlp() {
    for (void* b : bufs) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
  }
  template <typename T>
  int alloc(T** out, size_t n) {
    void* ptr = nullptr;
    if (hipMalloc(&ptr, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) return mlp_err(-101, "out of device memory");
    bufs.push_back(ptr);
    *out = (T*)ptr;
    return 0;
  }
};

namespace {

int ensure_batch(qmi_mlp* M, int B) {
  if (B <= M->cap) return 0;
  const int H = M->H;
  int rc;
  float** f[] = {&M->z1, &M->a1, &M->z2, &M->a2, &M->dz2, &M->da2, &M->da1, &M->dz1};
  for (float** q : f)
    if ((rc = M->alloc(q, (size_t)B * H)) != 0) return rc;
  if ((rc = M->alloc(&M->x, (size_t)B * M->n_in)) != 0) return rc;
  for (float** q : {&M->t, &M->z3, &M->y, &M->dz3, &M->lossv})
    if ((rc = M->alloc(q, (size_t)B)) != 0) return rc;
  if ((rc = M->alloc(&M->loss, 1)) != 0) return rc;
  M->cap = B;
  return 0;
}

template <typename TA_, typename TB_, bool TA, bool TB>
void gemm(hipStream_t s, int Mr, int N, int K, float alpha, const TA_* A, int lda, const TB_* B, int ldb, float beta,
          float* C, int ldc) {
  dim3 grid(nblk(N, 16), nblk(Mr, 16));
  k_gemm<TA_, TB_, TA, TB><<<grid, 256, 0, s>>>(Mr, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// activations of one forward pass: B x H (z1, a1, z2, a2) and B (z3)
struct Act {
  float *z1, *a1, *z2, *a2, *z3;
};

// forward pass of B rows x[B x n_in] into y[B] on stream s, activations in
// `a`; with targets t also the loss terms and dz3
void forward_on(qmi_mlp* M, hipStream_t s, int B, const float* x, const Act& a, float* y, const float* t,
                float* dz3, float* lossv) {
  const int H = M->H, I = M->n_in;
  const hip_bfloat16* W1 = M->pb + M->o_w1;
  const hip_bfloat16* W2 = M->pb + M->o_w2;
  const hip_bfloat16* W3 = M->pb + M->o_w3;
  gemm<float, hip_bfloat16, false, true>(s, B, H, I, 1.f, x, I, W1, I, 0.f, a.z1, H);   // x W1^T
  k_bias_silu<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, H, a.z1, M->p + M->o_b1, a.a1);
  gemm<float, hip_bfloat16, false, true>(s, B, H, H, 1.f, a.a1, H, W2, H, 0.f, a.z2, H);  // a1 W2^T
  k_bias_silu<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, H, a.z2, M->p + M->o_b2, a.a2);
  gemm<float, hip_bfloat16, false, true>(s, B, 1, H, 1.f, a.a2, H, W3, H, 0.f, a.z3, 1);  // a2 W3^T
  k_out<<<nblk(B, 256), 256, 0, s>>>(B, a.z3, M->p + M->o_b3, y, t, dz3, lossv);
}

// forward pass of B rows in M->x; with targets (M->t) also the loss and dz3
void forward(qmi_mlp* M, int B, bool with_t) {
  forward_on(M, M->s, B, M->x, Act{M->z1, M->a1, M->z2, M->a2, M->z3}, M->y, with_t ? M->t : nullptr, M->dz3,
             M->lossv);
}

// Scratch of qmi_mlp_forward_noalloc: the activations of one forward pass,
// B x (4H + 1) floats, in a fixed device array (4 MiB on every device the
// library runs on, held until the process ends).  Host code never passes the
// array's name to a kernel: it uses the address hipGetSymbolAddress returns
// for the current device, kept per device in scratch_ptr.
constexpr long long kScratchFloats = 1LL << 20;
__device__ __align__(256) float scratch[kScratchFloats];
constexpr int kMaxDevices = 64;
float* scratch_ptr[kMaxDevices] = {};

// set up scratch_ptr for `device`, which must be current
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return mlp_err(-1, "device number beyond the scratch table");
  if (scratch_ptr[device]) return 0;
  void* ptr = nullptr;
  MLP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return mlp_err(-100, "no address for the scratch array");
  scratch_ptr[device] = (float*)ptr;
  return 0;
}

// gradient of the mean squared error into M->g (after forward with targets)
void backward(qmi_mlp* M, int B) {
  const int H = M->H, I = M->n_in;
  hipStream_t s = M->s;
  const hip_bfloat16* W2 = M->pb + M->o_w2;
  const hip_bfloat16* W3 = M->pb + M->o_w3;
  float* g = M->g;
  // output layer
  gemm<float, float, true, false>(s, 1, H, B, 1.f, M->dz3, 1, M->a2, H, 0.f, g + M->o_w3, H);  // dz3^T a2
  k_colsum<<<1, 64, 0, s>>>(B, 1, M->dz3, g + M->o_b3);
  gemm<float, hip_bfloat16, false, false>(s, B, H, 1, 1.f, M->dz3, 1, W3, H, 0.f, M->da2, H);  // dz3 W3
  k_silu_back<<<nblk((long long)B * H, 256), 256, 0, s>>>(B * H, M->z2, M->da2, M->dz2);
  // hidden layer 2
  gemm<float, float, true, false>(s, H, H, B, 1.f, M->dz2, H, M->a1, H, 0.f, g + M->o_w2, H);  // dz2^T a1
  k_colsum<<<nblk(H, 128), 128, 0, s>>>(B, H, M->dz2, g + M->o_b2);
  gemm<float, hip_bfloat16, false, false>(s, B, H, H, 1.f, M->dz2, H, W2, H, 0.f, M->da1, H);  // dz2 W2
  k_silu_back<<<nblk((long long)B * H, 256), 256, 0, s>>>(B * H, M->z1, M->da1, M->dz1);
  // hidden layer 1
  gemm<float, float, true, false>(s, H, I, B, 1.f, M->dz1, H, M->x, I, 0.f, g + M->o_w1, I);  // dz1^T x
  k_colsum<<<nblk(H, 128), 128, 0, s>>>(B, H, M->dz1, g + M->o_b1);
}

}  // namespace

extern "C" {

const char* qmi_mlp_last_error(void) { return g_mlp_err.c_str(); }

int qmi_mlp_create(int hip_device, int n_in, int n_hidden, unsigned long long seed, qmi_mlp** out) {
  if (!out || n_in < 1 || n_in > 4096 || n_hidden < 1 || n_hidden > 4096) return mlp_err(-1, "invalid MLP shape");
  int ndev = 0;
  MLP_OK(hipGetDeviceCount(&ndev));
  if (hip_device < 0 || hip_device >= ndev) return mlp_err(-1, "invalid HIP device");
  MLP_OK(hipSetDevice(hip_device));
  auto* M = new qmi_mlp();
  M->device = hip_device;
  M->n_in = n_in;
  M->H = n_hidden;
  const long long H = n_hidden, I = n_in;
  M->o_w1 = 0;
  M->o_b1 = H * I;
  M->o_w2 = M->o_b1 + H;
  M->o_b2 = M->o_w2 + H * H;
  M->o_w3 = M->o_b2 + H;
  M->o_b3 = M->o_w3 + H;
  M->np = M->o_b3 + 1;
  int rc;
  if (hipStreamCreate(&M->s) != hipSuccess) { delete M; return mlp_err(-100, "stream"); }
  if ((rc = M->alloc(&M->p, M->np)) || (rc = M->alloc(&M->pb, M->np)) || (rc = M->alloc(&M->g, M->np)) ||
      (rc = M->alloc(&M->m, M->np)) || (rc = M->alloc(&M->v, M->np))) {
    delete M;
    return rc;
  }
  // He-style uniform initialisation from the seed, biases zero
  std::vector<float> h(M->np, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](long long off, long long n, long long fan_in) {
    const float a = std::sqrt(6.0f / (float)fan_in);
    std::uniform_real_distribution<float> u(-a, a);
    for (long long k = 0; k < n; ++k) h[off + k] = u(rng);
  };
  fill(M->o_w1, H * I, I);
  fill(M->o_w2, H * H, H);
  fill(M->o_w3, H, H);
  std::vector<float> zero(M->np, 0.f);
  if (hipMemcpy(M->p, h.data(), M->np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M->m, zero.data(), M->np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(M->v, zero.data(), M->np * 4, hipMemcpyHostToDevice) != hipSuccess) {
    delete M;
    return mlp_err(-100, "upload");
  }
  k_to_bf16<<<nblk(M->np, 256), 256, 0, M->s>>>(M->np, M->p, M->pb);
  if (hipStreamSynchronize(M->s) != hipSuccess) { delete M; return mlp_err(-100, "init"); }
  {
    // the fixed scratch array of qmi_mlp_forward_noalloc on this device
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if ((rc = scratch_init(hip_device)) != 0) { delete M; return rc; }
  }
  *out = M;
  return 0;
}

long long qmi_mlp_num_params(qmi_mlp* M) { return M ? M->np : -1; }

int qmi_mlp_get_params(qmi_mlp* M, float* params, long long cap) {
  if (!M || !params) return mlp_err(-1, "null argument");
  if (cap < M->np) return mlp_err(-3, "buffer too small");
  MLP_OK(hipSetDevice(M->device));
  MLP_OK(hipMemcpy(params, M->p, M->np * 4, hipMemcpyDeviceToHost));
  return 0;
}

int qmi_mlp_set_params(qmi_mlp* M, const float* params, long long n) {
  if (!M || !params) return mlp_err(-1, "null argument");
  if (n != M->np) return mlp_err(-3, "wrong number of parameters");
  MLP_OK(hipSetDevice(M->device));
  MLP_OK(hipMemcpy(M->p, params, n * 4, hipMemcpyHostToDevice));
  k_to_bf16<<<nblk(M->np, 256), 256, 0, M->s>>>(M->np, M->p, M->pb);
  MLP_OK(hipStreamSynchronize(M->s));
  return 0;
}

int qmi_mlp_forward(qmi_mlp* M, const float* x, int batch, float* y) {
  if (!M || !x || !y || batch < 1) return mlp_err(-1, "invalid argument");
  MLP_OK(hipSetDevice(M->device));
  int rc;
  if ((rc = ensure_batch(M, batch)) != 0) return rc;
  MLP_OK(hipMemcpyAsync(M->x, x, (size_t)batch * M->n_in * 4, hipMemcpyHostToDevice, M->s));
  forward(M, batch, false);
  MLP_OK(hipGetLastError());
  MLP_OK(hipMemcpyAsync(y, M->y, (size_t)batch * 4, hipMemcpyDeviceToHost, M->s));
  MLP_OK(hipStreamSynchronize(M->s));
  return 0;
}

int qmi_mlp_forward_noalloc(qmi_mlp* M, const float* d_x, int batch, float* d_y, void* stream) {
  if (!M || !d_x || !d_y || batch < 1) return mlp_err(-1, "invalid argument");
  const long long per_row = 4LL * M->H + 1;
  if ((long long)batch * per_row > kScratchFloats)
    return mlp_err(-3, "batch of " + std::to_string(batch) + " rows exceeds the scratch of " +
                           std::to_string(kScratchFloats / per_row) + " rows for this network");
  float* base = M->device >= 0 && M->device < kMaxDevices ? scratch_ptr[M->device] : nullptr;
  if (!base) return mlp_err(-2, "scratch not set up for this device");
  MLP_OK(hipSetDevice(M->device));
  const long long BH = (long long)batch * M->H;
  const Act a{base, base + BH, base + 2 * BH, base + 3 * BH, base + 4 * BH};
  forward_on(M, stream ? (hipStream_t)stream : M->s, batch, d_x, a, d_y, nullptr, nullptr, nullptr);
  MLP_OK(hipGetLastError());
  return 0;
}

int qmi_mlp_grad(qmi_mlp* M, const float* x, const float* t, int batch, float* loss, float* grad) {
  if (!M || !x || !t || batch < 1) return mlp_err(-1, "invalid argument");
  MLP_OK(hipSetDevice(M->device));
  int rc;
  if ((rc = ensure_batch(M, batch)) != 0) return rc;
  MLP_OK(hipMemcpyAsync(M->x, x, (size_t)batch * M->n_in * 4, hipMemcpyHostToDevice, M->s));
  MLP_OK(hipMemcpyAsync(M->t, t, (size_t)batch * 4, hipMemcpyHostToDevice, M->s));
  forward(M, batch, true);
  backward(M, batch);
  k_sum<<<1, 256, 0, M->s>>>(batch, M->lossv, M->loss);
  MLP_OK(hipGetLastError());
  float l = 0.f;
  MLP_OK(hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, M->s));
  if (grad) MLP_OK(hipMemcpyAsync(grad, M->g, M->np * 4, hipMemcpyDeviceToHost, M->s));
  MLP_OK(hipStreamSynchronize(M->s));
  if (loss) *loss = l / (float)batch;
  return 0;
}

int qmi_mlp_train(qmi_mlp* M, const float* X, const float* T, int n, int epochs, int batch, float lr,
                  unsigned long long seed, float* final_loss) {
  if (!M || !X || !T || n < 1 || epochs < 0 || batch < 1) return mlp_err(-1, "invalid argument");
  MLP_OK(hipSetDevice(M->device));
  batch = std::min(batch, n);
  int rc;
  if ((rc = ensure_batch(M, batch)) != 0) return rc;
  // the training set stays on the device for the whole call
  struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
  } bX, bT, bI;
  if (hipMalloc(&bX.p, (size_t)n * M->n_in * 4) != hipSuccess || hipMalloc(&bT.p, (size_t)n * 4) != hipSuccess ||
      hipMalloc(&bI.p, (size_t)n * 4) != hipSuccess)
    return mlp_err(-101, "out of device memory");
  float* dX = (float*)bX.p;
  float* dT = (float*)bT.p;
  int* didx = (int*)bI.p;
  MLP_OK(hipMemcpy(dX, X, (size_t)n * M->n_in * 4, hipMemcpyHostToDevice));
  MLP_OK(hipMemcpy(dT, T, (size_t)n * 4, hipMemcpyHostToDevice));
  std::vector<int> order(n);
  for (int k = 0; k < n; ++k) order[k] = k;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    MLP_OK(hipMemcpyAsync(didx, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, M->s));
    double sum = 0.0;
    std::vector<float> part;
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      k_gather<<<nblk((long long)batch * M->n_in, 256), 256, 0, M->s>>>(batch, M->n_in, didx + b0, dX, dT, M->x, M->t);
      forward(M, batch, true);
      backward(M, batch);
      ++M->step;
      const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
      k_adam<<<nblk(M->np, 256), 256, 0, M->s>>>(M->np, M->p, M->g, M->m, M->v, lr, b1, b2, eps, c1, c2, M->pb);
      if (ep == epochs - 1) {
        k_sum<<<1, 256, 0, M->s>>>(batch, M->lossv, M->loss);
        float l = 0.f;
        MLP_OK(hipMemcpyAsync(&l, M->loss, 4, hipMemcpyDeviceToHost, M->s));
        MLP_OK(hipStreamSynchronize(M->s));
        sum += l;
      }
    }
    MLP_OK(hipGetLastError());
    if (ep == epochs - 1) last = sum / (double)((n / batch) * batch);
  }
  MLP_OK(hipStreamSynchronize(M->s));
  if (final_loss) *final_loss = (float)last;
  return 0;
}

void qmi_mlp_destroy(qmi_mlp* M) { delete M; }

}  // extern "C"
