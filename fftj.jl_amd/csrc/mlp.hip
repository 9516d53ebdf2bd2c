// The learned planner's network: a 3-layer MLP (in -> H -> H -> 1, ReLU,
// FP32) predicting a plan's log run time from its features, with forward,
// backward and Adam kernels written here.  Every reduction is one thread
// looping over its index range (no atomics), so training is bitwise
// reproducible run to run.
//
// Parameters, one flat FP32 vector: W1[H][in], b1[H], W2[H][H], b2[H],
// W3[1][H], b3[1].  Loss: L = 1/(2B) * sum_r (out_r - t_r)^2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fftjh.h"

namespace fftjh {
namespace {

#define GRID_STRIDE(i, count)                                                                 \
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < (count); \
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)

dim3 grid_for(int64_t count) {
  int64_t g = (count + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return dim3(static_cast<unsigned>(g));
}

// Y[r][o] = act(sum_i X[r][i] * W[o][i] + b[o])
__global__ void k_linear_fwd(const float* X, const float* W, const float* b, float* Y, int B, int in, int out,
                             int relu) {
  GRID_STRIDE(idx, static_cast<int64_t>(B) * out) {
    const int r = static_cast<int>(idx / out), o = static_cast<int>(idx % out);
    float acc = b[o];
    for (int i = 0; i < in; ++i) acc = fmaf(X[static_cast<int64_t>(r) * in + i], W[static_cast<int64_t>(o) * in + i], acc);
    Y[idx] = relu ? fmaxf(acc, 0.f) : acc;
  }
}

// dX[r][i] = sum_o dY[r][o] * W[o][i]; then (mask) dX *= (H[r][i] > 0)
__global__ void k_linear_bwd_data(const float* dY, const float* W, const float* H, float* dX, int B, int in,
                                  int out) {
  GRID_STRIDE(idx, static_cast<int64_t>(B) * in) {
    const int r = static_cast<int>(idx / in), i = static_cast<int>(idx % in);
    float acc = 0.f;
    for (int o = 0; o < out; ++o) acc = fmaf(dY[static_cast<int64_t>(r) * out + o], W[static_cast<int64_t>(o) * in + i], acc);
    if (H) acc = H[idx] > 0.f ? acc : 0.f;
    dX[idx] = acc;
  }
}

// dW[o][i] = sum_r dY[r][o] * X[r][i];  db[o] = sum_r dY[r][o]
__global__ void k_linear_bwd_weight(const float* dY, const float* X, float* dW, float* db, int B, int in, int out) {
  GRID_STRIDE(idx, static_cast<int64_t>(out) * (in + 1)) {
    const int o = static_cast<int>(idx / (in + 1)), i = static_cast<int>(idx % (in + 1));
    float acc = 0.f;
    if (i < in) {
      for (int r = 0; r < B; ++r) acc = fmaf(dY[static_cast<int64_t>(r) * out + o], X[static_cast<int64_t>(r) * in + i], acc);
      dW[static_cast<int64_t>(o) * in + i] = acc;
    } else {
      for (int r = 0; r < B; ++r) acc += dY[static_cast<int64_t>(r) * out + o];
      db[o] = acc;
    }
  }
}

// dout = (out - t) / B; loss partial sums per block of rows (summed on the host)
__global__ void k_mse_grad(const float* out, const float* t, float* dout, float* sq, int B) {
  GRID_STRIDE(r, B) {
    const float e = out[r] - t[r];
    dout[r] = e / B;
    sq[r] = e * e;
  }
}

// Adam with bias correction
__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  GRID_STRIDE(i, count) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

}  // namespace

struct Mlp {
  int in = 0, H = 0, cap = 0;  // cap: largest batch the activations hold
  int64_t np = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float *h1 = nullptr, *h2 = nullptr, *out = nullptr, *d1 = nullptr, *d2 = nullptr, *dout = nullptr, *sq = nullptr;
  int64_t step = 0;
  float* W1() const { return params; }
  float* b1() const { return params + static_cast<int64_t>(H) * in; }
  float* W2() const { return b1() + H; }
  float* b2() const { return W2() + static_cast<int64_t>(H) * H; }
  float* W3() const { return b2() + H; }
  float* b3() const { return W3() + H; }
  float* gW1() const { return grads; }
  float* gb1() const { return grads + static_cast<int64_t>(H) * in; }
  float* gW2() const { return gb1() + H; }
  float* gb2() const { return gW2() + static_cast<int64_t>(H) * H; }
  float* gW3() const { return gb2() + H; }
  float* gb3() const { return gW3() + H; }
};

static thread_local std::string g_mlp_err;

static int mlp_fail(int code, const std::string& m) {
  g_mlp_err = m;
  return code;
}

#define HIPX(call)                                                       \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) return fftjh::mlp_fail(FFTJH_ERR_HIP, hipGetErrorString(e_)); \
  } while (0)

static int ensure_cap(Mlp* M, int B) {
  if (B <= M->cap) return 0;
  for (float** p : {&M->h1, &M->h2, &M->out, &M->d1, &M->d2, &M->dout, &M->sq}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  M->cap = 0;
  const int64_t H = M->H;
  HIPX(hipMalloc(&M->h1, sizeof(float) * B * H));
  HIPX(hipMalloc(&M->h2, sizeof(float) * B * H));
  HIPX(hipMalloc(&M->d1, sizeof(float) * B * H));
  HIPX(hipMalloc(&M->d2, sizeof(float) * B * H));
  HIPX(hipMalloc(&M->out, sizeof(float) * B));
  HIPX(hipMalloc(&M->dout, sizeof(float) * B));
  HIPX(hipMalloc(&M->sq, sizeof(float) * B));
  M->cap = B;
  return 0;
}

static int forward(Mlp* M, const float* X, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_linear_fwd, grid_for(int64_t(B) * M->H), dim3(256), 0, st, X, M->W1(), M->b1(), M->h1, B,
                     M->in, M->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, grid_for(int64_t(B) * M->H), dim3(256), 0, st, M->h1, M->W2(), M->b2(), M->h2,
                     B, M->H, M->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, grid_for(B), dim3(256), 0, st, M->h2, M->W3(), M->b3(), M->out, B, M->H, 1, 0);
  HIPX(hipGetLastError());
  return 0;
}

static int backward(Mlp* M, const float* X, const float* t, int B, hipStream_t st) {
  hipLaunchKernelGGL(k_mse_grad, grid_for(B), dim3(256), 0, st, M->out, t, M->dout, M->sq, B);
  hipLaunchKernelGGL(k_linear_bwd_weight, grid_for(int64_t(M->H) + 1), dim3(256), 0, st, M->dout, M->h2, M->gW3(),
                     M->gb3(), B, M->H, 1);
  hipLaunchKernelGGL(k_linear_bwd_data, grid_for(int64_t(B) * M->H), dim3(256), 0, st, M->dout, M->W3(), M->h2,
                     M->d2, B, M->H, 1);
  hipLaunchKernelGGL(k_linear_bwd_weight, grid_for(int64_t(M->H) * (M->H + 1)), dim3(256), 0, st, M->d2, M->h1,
                     M->gW2(), M->gb2(), B, M->H, M->H);
  hipLaunchKernelGGL(k_linear_bwd_data, grid_for(int64_t(B) * M->H), dim3(256), 0, st, M->d2, M->W2(), M->h1,
                     M->d1, B, M->H, M->H);
  hipLaunchKernelGGL(k_linear_bwd_weight, grid_for(int64_t(M->H) * (M->in + 1)), dim3(256), 0, st, M->d1, X,
                     M->gW1(), M->gb1(), B, M->in, M->H);
  HIPX(hipGetLastError());
  return 0;
}

static float host_loss(Mlp* M, int B, hipStream_t st, int* rc) {
  std::vector<float> sq(B);
  if (hipMemcpyAsync(sq.data(), M->sq, sizeof(float) * B, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    *rc = mlp_fail(FFTJH_ERR_HIP, "loss copy");
    return 0.f;
  }
  double s = 0;
  for (float e : sq) s += e;
  *rc = 0;
  return static_cast<float>(s / (2.0 * B));
}

}  // namespace fftjh

using fftjh::Mlp;

extern "C" {

int fftjh_mlp_create(fftjh_mlp* model, int in_features, int hidden) {
  if (!model) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "model pointer is NULL");
  *model = nullptr;
  if (in_features < 1 || hidden < 1) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "sizes must be positive");
  auto* M = new Mlp();
  M->in = in_features;
  M->H = hidden;
  M->np = int64_t(hidden) * in_features + hidden + int64_t(hidden) * hidden + hidden + hidden + 1;
  for (float** p : {&M->params, &M->grads, &M->m, &M->v}) {
    if (hipMalloc(p, sizeof(float) * M->np) != hipSuccess) {
      delete M;
      return fftjh::mlp_fail(FFTJH_ERR_HIP, "hipMalloc");
    }
  }
  (void)hipMemset(M->m, 0, sizeof(float) * M->np);
  (void)hipMemset(M->v, 0, sizeof(float) * M->np);
  (void)hipMemset(M->grads, 0, sizeof(float) * M->np);
  *model = reinterpret_cast<fftjh_mlp>(M);
  return 0;
}

int fftjh_mlp_destroy(fftjh_mlp model) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M) return 0;
  for (float* p : {M->params, M->grads, M->m, M->v, M->h1, M->h2, M->out, M->d1, M->d2, M->dout, M->sq})
    if (p) (void)hipFree(p);
  delete M;
  return 0;
}

int64_t fftjh_mlp_param_count(fftjh_mlp model) {
  auto* M = reinterpret_cast<Mlp*>(model);
  return M ? M->np : -1;
}

int fftjh_mlp_set_params(fftjh_mlp model, const float* host_params) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M || !host_params) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "NULL argument");
  HIPX(hipMemcpy(M->params, host_params, sizeof(float) * M->np, hipMemcpyHostToDevice));
  HIPX(hipMemset(M->m, 0, sizeof(float) * M->np));
  HIPX(hipMemset(M->v, 0, sizeof(float) * M->np));
  M->step = 0;
  return 0;
}

int fftjh_mlp_get_params(fftjh_mlp model, float* host_params) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M || !host_params) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "NULL argument");
  HIPX(hipMemcpy(host_params, M->params, sizeof(float) * M->np, hipMemcpyDeviceToHost));
  return 0;
}

int fftjh_mlp_predict(fftjh_mlp model, const float* X, float* out, int batch, fftjh_stream stream) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M || !X || !out || batch < 1) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "bad predict arguments");
  if (int rc = fftjh::ensure_cap(M, batch)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int rc = fftjh::forward(M, X, batch, st)) return rc;
  HIPX(hipMemcpyAsync(out, M->out, sizeof(float) * batch, hipMemcpyDeviceToDevice, st));
  return 0;
}

int fftjh_mlp_gradients(fftjh_mlp model, const float* X, const float* target, int batch, float* host_grads,
                        float* loss, fftjh_stream stream) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M || !X || !target || batch < 1) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "bad gradient arguments");
  if (int rc = fftjh::ensure_cap(M, batch)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int rc = fftjh::forward(M, X, batch, st)) return rc;
  if (int rc = fftjh::backward(M, X, target, batch, st)) return rc;
  int rc = 0;
  const float l = fftjh::host_loss(M, batch, st, &rc);
  if (rc) return rc;
  if (loss) *loss = l;
  if (host_grads) HIPX(hipMemcpy(host_grads, M->grads, sizeof(float) * M->np, hipMemcpyDeviceToHost));
  return 0;
}

int fftjh_mlp_train_step(fftjh_mlp model, const float* X, const float* target, int batch, float lr, float* loss,
                         fftjh_stream stream) {
  auto* M = reinterpret_cast<Mlp*>(model);
  if (!M || !X || !target || batch < 1) return fftjh::mlp_fail(FFTJH_ERR_ARGUMENT, "bad training arguments");
  if (int rc = fftjh::ensure_cap(M, batch)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int rc = fftjh::forward(M, X, batch, st)) return rc;
  if (int rc = fftjh::backward(M, X, target, batch, st)) return rc;
  M->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, static_cast<float>(M->step));
  const float c2 = 1.f - std::pow(b2, static_cast<float>(M->step));
  hipLaunchKernelGGL(fftjh::k_adam, fftjh::grid_for(M->np), dim3(256), 0, st, M->params, M->grads, M->m, M->v,
                     M->np, lr, b1, b2, 1e-8f, c1, c2);
  HIPX(hipGetLastError());
  if (loss) {
    int rc = 0;
    *loss = fftjh::host_loss(M, batch, st, &rc);
    if (rc) return rc;
  }
  return 0;
}

const char* fftjh_mlp_last_error(void) { return fftjh::g_mlp_err.c_str(); }

}  // extern "C"
