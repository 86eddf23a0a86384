// mlp.hip -- the learned cost model: a 3-layer MLP (din -> H -> H -> dout,
// ReLU) with hand-written HIP forward, backward, loss and Adam kernels.
//
// Used by the tile selector: trained on the autotuner's timing database
// (features of a shape + a kernel configuration -> log2 of the measured call
// time) and evaluated on held-out shapes.  Parameters are one flat float
// array: W0[din][H], b0[H], W1[H][H], b1[H], W2[H][dout], b2[dout].
// Every reduction runs in a fixed order (no atomics in the gradients), so a
// training run is reproducible for a given seed.
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

// Y[i][j] = act(b[j] + sum_k X[i][k] W[k][j])
__global__ void k_linear_fwd(const float *X, const float *W, const float *b, float *Y, int B, int din, int dout,
                             int relu) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= B * dout) return;
  int i = t / dout, j = t % dout;
  float s = b[j];
  const float *x = X + (long long)i * din;
  for (int k = 0; k < din; k++) s = fmaf(x[k], W[k * dout + j], s);
  Y[t] = relu ? fmaxf(s, 0.f) : s;
}

// dY = mask * (Y - T) / denom; loss partial per row (summed later in order)
__global__ void k_mse_grad(const float *Y, const float *T, const float *Mk, float *dY, float *row_loss, int B,
                           int dout, float inv_denom) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B) return;
  float l = 0.f;
  for (int j = 0; j < dout; j++) {
    int t = i * dout + j;
    float m = Mk ? Mk[t] : 1.f;
    float d = (Y[t] - T[t]) * m;
    dY[t] = d * inv_denom;
    l += 0.5f * d * d;
  }
  row_loss[i] = l * inv_denom;
}

// dW[k][j] = sum_i X[i][k] dY[i][j]; db[j] = sum_i dY[i][j]  (batch order)
__global__ void k_linear_bwd_w(const float *X, const float *dY, float *dW, float *db, int B, int din, int dout) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= (din + 1) * dout) return;
  int k = t / dout, j = t % dout;
  float s = 0.f;
  if (k < din) {
    for (int i = 0; i < B; i++) s = fmaf(X[(long long)i * din + k], dY[(long long)i * dout + j], s);
    dW[k * dout + j] = s;
  } else {
    for (int i = 0; i < B; i++) s += dY[(long long)i * dout + j];
    db[j] = s;
  }
}

// dX[i][k] = (sum_j dY[i][j] W[k][j]) * [Xact[i][k] > 0]
__global__ void k_linear_bwd_x(const float *dY, const float *W, const float *Xact, float *dX, int B, int din,
                               int dout) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= B * din) return;
  int i = t / din, k = t % din;
  float s = 0.f;
  for (int j = 0; j < dout; j++) s = fmaf(dY[(long long)i * dout + j], W[k * dout + j], s);
  dX[t] = Xact[t] > 0.f ? s : 0.f;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_gather_rows(const float *src, const int *idx, float *dst, int B, int width) {
  int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= B * width) return;
  int i = t / width, c = t % width;
  dst[t] = src[(long long)idx[i] * width + c];
}

// dst[0] = sum of v[0..n) in a fixed order (one block)
__global__ void k_sum_ordered(const float *v, int n, float scale, double *dst) {
  __shared__ double part[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += (double)v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst = part[0] * (double)scale;
}

inline unsigned nb(long long n) { return (unsigned)((n + 255) / 256); }
} // namespace

struct NumcMlp {
  NumcCtx *ctx;
  int din, H, dout;
  int cap_batch = 0;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr, *Mk = nullptr;
  float *dY = nullptr, *dH2 = nullptr, *dH1 = nullptr, *rowloss = nullptr;
  int *idx = nullptr;
  long long step = 0;
  size_t nparam() const { return (size_t)din * H + H + (size_t)H * H + H + (size_t)H * dout + dout; }
  float *W0(float *p) const { return p; }
  float *b0(float *p) const { return p + (size_t)din * H; }
  float *W1(float *p) const { return b0(p) + H; }
  float *b1(float *p) const { return W1(p) + (size_t)H * H; }
  float *W2(float *p) const { return b1(p) + H; }
  float *b2(float *p) const { return W2(p) + (size_t)H * dout; }
};

static int mlp_reserve(NumcMlp *n, int B) {
  if (B <= n->cap_batch) return 0;
  float **bufs[] = {&n->X, &n->H1, &n->H2, &n->Y, &n->T, &n->Mk, &n->dY, &n->dH2, &n->dH1, &n->rowloss};
  HIP_TRY(hipStreamSynchronize(n->ctx->stream), "hipStreamSynchronize");
  for (auto b : bufs)
    if (*b) (void)hipFree(*b), *b = nullptr;
  if (n->idx) (void)hipFree(n->idx), n->idx = nullptr;
  const size_t w = (size_t)std::max(std::max(n->din, n->H), n->dout);
  for (auto b : bufs) HIP_TRY(hipMalloc((void **)b, (size_t)B * w * sizeof(float)), "hipMalloc(mlp)");
  HIP_TRY(hipMalloc((void **)&n->idx, (size_t)B * sizeof(int)), "hipMalloc(mlp)");
  n->cap_batch = B;
  return 0;
}

// forward of a batch already in n->X (B rows)
static void mlp_forward_dev(NumcMlp *n, int B) {
  hipStream_t s = n->ctx->stream;
  hipLaunchKernelGGL(k_linear_fwd, dim3(nb((long long)B * n->H)), dim3(256), 0, s, n->X, n->W0(n->params),
                     n->b0(n->params), n->H1, B, n->din, n->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, dim3(nb((long long)B * n->H)), dim3(256), 0, s, n->H1, n->W1(n->params),
                     n->b1(n->params), n->H2, B, n->H, n->H, 1);
  hipLaunchKernelGGL(k_linear_fwd, dim3(nb((long long)B * n->dout)), dim3(256), 0, s, n->H2, n->W2(n->params),
                     n->b2(n->params), n->Y, B, n->H, n->dout, 0);
}

// loss + gradients of a batch in n->X / n->T / n->Mk (mask optional)
static void mlp_backward_dev(NumcMlp *n, int B, bool masked, float inv_denom) {
  hipStream_t s = n->ctx->stream;
  hipLaunchKernelGGL(k_mse_grad, dim3(nb(B)), dim3(256), 0, s, n->Y, n->T, masked ? n->Mk : nullptr, n->dY,
                     n->rowloss, B, n->dout, inv_denom);
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(nb((long long)(n->H + 1) * n->dout)), dim3(256), 0, s, n->H2, n->dY,
                     n->W2(n->grads), n->b2(n->grads), B, n->H, n->dout);
  hipLaunchKernelGGL(k_linear_bwd_x, dim3(nb((long long)B * n->H)), dim3(256), 0, s, n->dY, n->W2(n->params), n->H2,
                     n->dH2, B, n->H, n->dout);
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(nb((long long)(n->H + 1) * n->H)), dim3(256), 0, s, n->H1, n->dH2,
                     n->W1(n->grads), n->b1(n->grads), B, n->H, n->H);
  hipLaunchKernelGGL(k_linear_bwd_x, dim3(nb((long long)B * n->H)), dim3(256), 0, s, n->dH2, n->W1(n->params), n->H1,
                     n->dH1, B, n->H, n->H);
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(nb((long long)(n->din + 1) * n->H)), dim3(256), 0, s, n->X, n->dH1,
                     n->W0(n->grads), n->b0(n->grads), B, n->din, n->H);
}

extern "C" NUMC_API NumcMlp *numc_mlp_create(NumcCtx *ctx, int din, int hidden, int dout, uint64_t seed) {
  if (!ctx || din <= 0 || hidden <= 0 || dout <= 0) {
    NUMC_SET_ERROR(NUMC_ERR_VALUE, "mlp_create: bad arguments");
    return nullptr;
  }
  if (numc_gpu_init(ctx)) return nullptr;
  NumcMlp *n = new NumcMlp();
  n->ctx = ctx;
  n->din = din;
  n->H = hidden;
  n->dout = dout;
  const size_t np = n->nparam();
  float **bufs[] = {&n->params, &n->grads, &n->m, &n->v};
  for (auto b : bufs)
    if (hipMalloc((void **)b, np * sizeof(float)) != hipSuccess) {
      NUMC_SET_ERROR(NUMC_ERR_MALLOC, "mlp_create: device allocation failed");
      return nullptr;
    }
  // He-uniform initial weights, zero biases (host, seeded)
  std::vector<float> h(np, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](float *w, int fan_in, size_t count) {
    std::uniform_real_distribution<float> u(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
    for (size_t i = 0; i < count; i++) w[i] = u(rng);
  };
  init(n->W0(h.data()), din, (size_t)din * hidden);
  init(n->W1(h.data()), hidden, (size_t)hidden * hidden);
  init(n->W2(h.data()), hidden, (size_t)hidden * dout);
  (void)hipMemcpy(n->params, h.data(), np * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> z(np, 0.f);
  (void)hipMemcpy(n->m, z.data(), np * sizeof(float), hipMemcpyHostToDevice);
  (void)hipMemcpy(n->v, z.data(), np * sizeof(float), hipMemcpyHostToDevice);
  return n;
}

extern "C" NUMC_API void numc_mlp_free(NumcMlp *n) {
  if (!n) return;
  (void)hipStreamSynchronize(n->ctx->stream);
  float *bufs[] = {n->params, n->grads, n->m, n->v, n->X, n->H1, n->H2, n->Y, n->T, n->Mk, n->dY, n->dH2, n->dH1,
                   n->rowloss};
  for (auto b : bufs)
    if (b) (void)hipFree(b);
  if (n->idx) (void)hipFree(n->idx);
  delete n;
}

extern "C" NUMC_API size_t numc_mlp_param_count(const NumcMlp *n) { return n ? n->nparam() : 0; }

extern "C" NUMC_API int numc_mlp_set_params(NumcMlp *n, const float *p, size_t count) {
  if (!n || !p) return NUMC_SET_ERROR(NUMC_ERR_NULL, "mlp_set_params: NULL argument");
  if (count != n->nparam()) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "mlp_set_params: expected %zu values", n->nparam());
  HIP_TRY(hipMemcpy(n->params, p, count * sizeof(float), hipMemcpyHostToDevice), "H2D");
  return 0;
}

extern "C" NUMC_API int numc_mlp_get_params(NumcMlp *n, float *p, size_t count) {
  if (!n || !p) return NUMC_SET_ERROR(NUMC_ERR_NULL, "mlp_get_params: NULL argument");
  if (count != n->nparam()) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "mlp_get_params: expected %zu values", n->nparam());
  HIP_TRY(hipStreamSynchronize(n->ctx->stream), "hipStreamSynchronize");
  HIP_TRY(hipMemcpy(p, n->params, count * sizeof(float), hipMemcpyDeviceToHost), "D2H");
  return 0;
}

// inference kernel path: y[rows][dout] = mlp(x[rows][din]), in chunks
extern "C" NUMC_API int numc_mlp_forward(NumcMlp *n, const float *x, long long rows, float *y) {
  if (!n || !x || !y) return NUMC_SET_ERROR(NUMC_ERR_NULL, "mlp_forward: NULL argument");
  const int chunk = 8192;
  int err = mlp_reserve(n, chunk);
  if (err) return err;
  hipStream_t s = n->ctx->stream;
  for (long long r0 = 0; r0 < rows; r0 += chunk) {
    int B = (int)std::min<long long>(chunk, rows - r0);
    HIP_TRY(hipMemcpyAsync(n->X, x + r0 * n->din, (size_t)B * n->din * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
    mlp_forward_dev(n, B);
    HIP_TRY(hipMemcpyAsync(y + r0 * n->dout, n->Y, (size_t)B * n->dout * sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
  }
  HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  return 0;
}

// loss and gradient of one batch (for checking against a CPU backprop):
// loss = sum_i sum_j 0.5 mask (y - t)^2 / B, grads in the parameter layout
extern "C" NUMC_API int numc_mlp_loss_grad(NumcMlp *n, const float *x, const float *t, const float *mask, int B,
                                           double *loss, float *grads) {
  if (!n || !x || !t || B <= 0) return NUMC_SET_ERROR(NUMC_ERR_NULL, "mlp_loss_grad: bad arguments");
  int err = mlp_reserve(n, B);
  if (err) return err;
  hipStream_t s = n->ctx->stream;
  HIP_TRY(hipMemcpyAsync(n->X, x, (size_t)B * n->din * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
  HIP_TRY(hipMemcpyAsync(n->T, t, (size_t)B * n->dout * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
  if (mask) HIP_TRY(hipMemcpyAsync(n->Mk, mask, (size_t)B * n->dout * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
  mlp_forward_dev(n, B);
  mlp_backward_dev(n, B, mask != nullptr, 1.f / (float)B);
  std::vector<float> rl(B);
  HIP_TRY(hipMemcpyAsync(rl.data(), n->rowloss, (size_t)B * sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
  if (grads) HIP_TRY(hipMemcpyAsync(grads, n->grads, n->nparam() * sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
  HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (loss) {
    double l = 0;
    for (float v : rl) l += v;
    *loss = l;
  }
  return 0;
}

// Adam training on (x[rows][din], t[rows][dout], optional mask) for `epochs`
// passes of shuffled minibatches; the data stays on the device.  loss_out
// (optional, `epochs` entries) receives the mean loss of each epoch.
extern "C" NUMC_API int numc_mlp_train(NumcMlp *n, const float *x, const float *t, const float *mask, long long rows,
                                       int epochs, int batch, float lr, float lr_final, uint64_t seed,
                                       double *loss_out) {
  if (!n || !x || !t || rows <= 0 || epochs <= 0 || batch <= 0)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "mlp_train: bad arguments");
  int err = mlp_reserve(n, batch);
  if (err) return err;
  hipStream_t s = n->ctx->stream;
  float *dx = nullptr, *dt = nullptr, *dm = nullptr;
  int *dperm = nullptr;
  auto cleanup = [&]() {
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (dm) (void)hipFree(dm);
    if (dperm) (void)hipFree(dperm);
  };
  if (hipMalloc((void **)&dx, (size_t)rows * n->din * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&dt, (size_t)rows * n->dout * sizeof(float)) != hipSuccess ||
      (mask && hipMalloc((void **)&dm, (size_t)rows * n->dout * sizeof(float)) != hipSuccess) ||
      hipMalloc((void **)&dperm, (size_t)rows * sizeof(int)) != hipSuccess) {
    cleanup();
    return NUMC_SET_ERROR(NUMC_ERR_MALLOC, "mlp_train: device allocation failed");
  }
  (void)hipMemcpyAsync(dx, x, (size_t)rows * n->din * sizeof(float), hipMemcpyHostToDevice, s);
  (void)hipMemcpyAsync(dt, t, (size_t)rows * n->dout * sizeof(float), hipMemcpyHostToDevice, s);
  if (mask) (void)hipMemcpyAsync(dm, mask, (size_t)rows * n->dout * sizeof(float), hipMemcpyHostToDevice, s);
  std::vector<int> perm(rows);
  std::mt19937_64 rng(seed);
  const long long nbatch = (rows + batch - 1) / batch;
  double *dloss = nullptr;
  if (hipMalloc((void **)&dloss, (size_t)nbatch * sizeof(double)) != hipSuccess) {
    cleanup();
    return NUMC_SET_ERROR(NUMC_ERR_MALLOC, "mlp_train: device allocation failed");
  }
  std::vector<double> hl(nbatch);
  const long long steps_total = (long long)epochs * ((rows + batch - 1) / batch);
  long long step_i = 0;
  for (int ep = 0; ep < epochs; ep++) {
    for (long long i = 0; i < rows; i++) perm[i] = (int)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    (void)hipMemcpyAsync(dperm, perm.data(), (size_t)rows * sizeof(int), hipMemcpyHostToDevice, s);
    double ep_loss = 0;
    for (long long r0 = 0; r0 < rows; r0 += batch) {
      const int B = (int)std::min<long long>(batch, rows - r0);
      hipLaunchKernelGGL(k_gather_rows, dim3(nb((long long)B * n->din)), dim3(256), 0, s, dx, dperm + r0, n->X, B, n->din);
      hipLaunchKernelGGL(k_gather_rows, dim3(nb((long long)B * n->dout)), dim3(256), 0, s, dt, dperm + r0, n->T, B, n->dout);
      if (mask)
        hipLaunchKernelGGL(k_gather_rows, dim3(nb((long long)B * n->dout)), dim3(256), 0, s, dm, dperm + r0, n->Mk, B, n->dout);
      mlp_forward_dev(n, B);
      mlp_backward_dev(n, B, mask != nullptr, 1.f / (float)B);
      // cosine-decayed learning rate
      const double frac = (double)step_i / (double)std::max<long long>(1, steps_total - 1);
      const float lr_now = (float)(lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(3.14159265358979 * frac)));
      n->step++;
      step_i++;
      const float c1 = 1.f - std::pow(0.9f, (float)n->step), c2 = 1.f - std::pow(0.999f, (float)n->step);
      hipLaunchKernelGGL(k_adam, dim3(nb((long long)n->nparam())), dim3(256), 0, s, n->params, n->grads, n->m, n->v,
                         (int)n->nparam(), lr_now, 0.9f, 0.999f, 1e-8f, c1, c2);
      if (loss_out)
        hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(256), 0, s, n->rowloss, B, (float)B, dloss + r0 / batch);
    }
    if (loss_out) {
      (void)hipMemcpyAsync(hl.data(), dloss, (size_t)nbatch * sizeof(double), hipMemcpyDeviceToHost, s);
      (void)hipStreamSynchronize(s);
      for (double v : hl) ep_loss += v;
      loss_out[ep] = ep_loss / (double)rows;
    }
  }
  hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(dloss);
  cleanup();
  if (e != hipSuccess) return numc_hip_fail(e, "mlp_train");
  e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "mlp_train launch");
  return 0;
}
