// Learned warm start of CTMRG: a fully connected network from the bulk tensor
// to a starting environment, trained and evaluated on the GPU.
//
// The reference starts every CTMRG from _initializect_square (src/ctmrg.jl:62-86,
// @nograd per src/autodiff.jl:5).  This file adds a predicted starting point;
// CTMRG still runs to the caller's tolerance, so the fixed point and its
// gradient (which ignores the initial environment) are unchanged.
//
//  * features of a bulk A: the double layer a = sum_s A (x) conj(A) of the
//    symmetrised bulk, normalised (the environment depends on A only through a);
//  * canonical gauge of an environment: C = O diag(lam) O^T (|lam| descending,
//    Jacobi), T~ = O^T T O, columns of O signed by a weighted coupling rule,
//    global signs of C and T fixed; features (lam, T~) in FP32;
//  * decoding: C0 = diag(lam), T0 = (T~ + T~^{kji}) / 2, normalised;
//  * network: tanh hidden layers, linear output, loss mean_b 0.5 |y_b - t_b|^2,
//    Adam.  Every product is the FP32 matrix-core contraction kernel; bias,
//    activation, loss and optimiser are elementwise kernels here.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "api_internal.h"
#include "ctmrg.h"

namespace tnad {
namespace {

constexpr int NT = 256;

inline unsigned nb(size_t n) {
  size_t b = (n + NT - 1) / NT;
  if (b > 65535ull * 8) b = 65535ull * 8;
  return (unsigned)(b ? b : 1);
}

// Z[:, b] += bias ; H = tanh(Z)  (rows x B)
__global__ void bias_tanh_kernel(const float* __restrict__ Z, const float* __restrict__ bias, int rows, int B,
                                 float* __restrict__ H) {
  size_t tot = (size_t)rows * B;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x)
    H[e] = tanhf(Z[e] + bias[e % rows]);
}

// Y[:, b] += bias
__global__ void bias_add_kernel(float* __restrict__ Y, const float* __restrict__ bias, int rows, int B) {
  size_t tot = (size_t)rows * B;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x)
    Y[e] += bias[e % rows];
}

// dY = (Y - T) / B ; part[block] = sum of (Y - T)^2 over the block's elements (FP64 accumulation)
__global__ void resid_loss_kernel(const float* __restrict__ Y, const float* __restrict__ T, size_t n, float invB,
                                  float* __restrict__ dY, double* __restrict__ part) {
  __shared__ double red[NT / 64];
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    float r = Y[e] - T[e];
    dY[e] = r * invB;
    acc += (double)r * (double)r;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

__global__ void sum_parts_kernel(const double* __restrict__ part, int np, double scale, double* __restrict__ out) {
  __shared__ double red[NT / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += NT) acc += part[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    *out = t * scale;
  }
}

// db[r] = sum_b dZ[r, b]
__global__ void rowsum_kernel(const float* __restrict__ dZ, int rows, int B, float* __restrict__ db) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float acc = 0.0f;
  for (int b = 0; b < B; ++b) acc += dZ[r + (size_t)rows * b];
  db[r] = acc;
}

// dZ = dH * (1 - H^2)
__global__ void tanh_back_kernel(const float* __restrict__ dH, const float* __restrict__ H, size_t n,
                                 float* __restrict__ dZ) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    float h = H[e];
    dZ[e] = dH[e] * (1.0f - h * h);
  }
}

// Adam: m = b1 m + (1-b1) g ; v = b2 v + (1-b2) g^2 ; theta -= lr mh / (sqrt(vh) + eps)
__global__ void adam_kernel(float* __restrict__ theta, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, float lr, float b1, float b2, float c1, float c2,
                            float eps) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    float ge = g[e];
    float me = b1 * m[e] + (1.0f - b1) * ge;
    float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    theta[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
  }
}

// uniform(-a, a) from a counter hash (splitmix64), a per block
__global__ void init_uniform_kernel(float* __restrict__ x, size_t n, float a, unsigned long long seed) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = seed + 0x9e3779b97f4a7c15ull * (e + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    float u = (float)((z >> 40) * (1.0 / 16777216.0));   // [0, 1)
    x[e] = a * (2.0f * u - 1.0f);
  }
}

__global__ void fill_f_kernel(float* __restrict__ x, size_t n, float v) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    x[e] = v;
}

// ---- environment codec ------------------------------------------------------------------
// Signs of the canonical gauge (one workgroup).  First the global signs of C and T (a fixed point
// is one only up to C -> -C, T -> -T): sgn[chi] = sign(lam[0]), sgn[chi+1] = g_T =
// sign(sum_a (a + 1) T~[0, a, 0]) (column signs leave T~[0, :, 0] alone).  Then column signs on
// g_T T~: s_0 = +1 and, for i = 1..chi-1 in order, s_i making
// sum_{r<i, a} (a + 1) / (r + 1) s_r g_T T~[r, a, i] >= 0 (sgn[0:chi]).
__global__ void canon_signs_kernel(const double* __restrict__ Tt, const double* __restrict__ lam, int chi, int Dd,
                                   double* __restrict__ sgn) {
  __shared__ double s_sh[JACOBI_MAX];
  __shared__ double red[NT / 64];
  __shared__ double gT;
  const int t = threadIdx.x;
  if (t == 0) {
    double w = 0.0;
    for (int a = 0; a < Dd; ++a) w += (double)(a + 1) * Tt[(size_t)chi * a];
    gT = w < 0.0 ? -1.0 : 1.0;
    s_sh[0] = 1.0;
    sgn[chi] = lam[0] < 0.0 ? -1.0 : 1.0;
    sgn[chi + 1] = gT;
  }
  __syncthreads();
  for (int i = 1; i < chi; ++i) {
    double acc = 0.0;
    for (int e = t; e < i * Dd; e += NT) {
      int r = e % i, a = e / i;
      acc += (double)(a + 1) / (double)(r + 1) * s_sh[r] * Tt[r + (size_t)chi * (a + (size_t)Dd * i)];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
      for (int w = 0; w < NT / 64; ++w) tot += red[w];
      s_sh[i] = gT * tot < 0.0 ? -1.0 : 1.0;
    }
    __syncthreads();
  }
  for (int i = t; i < chi; i += NT) sgn[i] = s_sh[i];
}

// feat = [g_C lam (chi) ; g_T s_i s_k T~[i, a, k] (chi*Dd*chi)] in FP32
__global__ void env_feat_kernel(const double* __restrict__ lam, const double* __restrict__ Tt,
                                const double* __restrict__ sgn, int chi, int Dd, float* __restrict__ feat) {
  size_t nT = (size_t)chi * Dd * chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nT + chi; e += (size_t)gridDim.x * blockDim.x) {
    if (e < (size_t)chi) {
      feat[e] = (float)(sgn[chi] * lam[e]);
    } else {
      size_t f = e - chi;
      int i = f % chi, k = f / ((size_t)chi * Dd);
      feat[e] = (float)(sgn[chi + 1] * sgn[i] * sgn[k] * Tt[f]);
    }
  }
}

// C0 = diag(feat[0:chi]) ; T0 = (T~ + T~^{kji}) / 2   (unnormalised, FP64)
__global__ void env_decode_kernel(const float* __restrict__ feat, int chi, int Dd, double* __restrict__ C0,
                                  double* __restrict__ T0) {
  size_t nC = (size_t)chi * chi, nT = (size_t)chi * Dd * chi;
  const float* Tt = feat + chi;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nC + nT; e += (size_t)gridDim.x * blockDim.x) {
    if (e < nC) {
      int i = e % chi, j = e / chi;
      C0[e] = (i == j) ? (double)feat[i] : 0.0;
    } else {
      size_t f = e - nC;
      int i = f % chi, a = (f / chi) % Dd, k = f / ((size_t)chi * Dd);
      double x = Tt[f], y = Tt[k + (size_t)chi * (a + (size_t)Dd * i)];
      T0[f] = 0.5 * (x + y);
    }
  }
}

// out = x / sqrt(*sumsq) in FP32
__global__ void scale_cast_kernel(const double* __restrict__ x, size_t n, const double* __restrict__ sumsq,
                                  float* __restrict__ out) {
  const double inv = 1.0 / sqrt(*sumsq);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    out[e] = (float)(x[e] * inv);
}

}  // namespace

// ---------------------------------------------------------------------------
// network state
// ---------------------------------------------------------------------------
struct Mlp {
  int n_in = 0, hidden = 0, layers = 0, n_out = 0;
  std::vector<size_t> off;      // offsets of W1, b1, ..., W_out, b_out in theta
  size_t n_params = 0;
  DevBuf theta, m, v, grad;
  long long t = 0;              // Adam step count
  int rows(int l) const { return l < layers ? hidden : n_out; }
  int cols(int l) const { return l == 0 ? n_in : hidden; }
};

static void mlp_layout(Mlp& M) {
  M.off.clear();
  size_t o = 0;
  for (int l = 0; l <= M.layers; ++l) {
    M.off.push_back(o);
    o += (size_t)M.rows(l) * M.cols(l);
    M.off.push_back(o);
    o += (size_t)M.rows(l);
  }
  M.n_params = o;
}

// Forward pass on library buffers; keeps the activations (acts[l], l = 0..layers, acts[0] = X).
static void mlp_forward(Ctx* c, const Mlp& M, const float* X, int B, std::vector<DevBuf>& acts, float* Y) {
  hipStream_t s = c->stream;
  acts.clear();
  const float* h = X;
  const float* th = M.theta.as<float>();
  for (int l = 0; l <= M.layers; ++l) {
    const int r = M.rows(l), k = M.cols(l);
    const float* W = th + M.off[2 * l];
    const float* bias = th + M.off[2 * l + 1];
    if (l < M.layers) {
      DevBuf Z(c, (size_t)r * B * 4), H(c, (size_t)r * B * 4);
      tc_launch(c, Prec::F32, gemm_desc('N', 'N', r, B, k, r, k, r), W, h, Z.ptr, s);
      hipLaunchKernelGGL(bias_tanh_kernel, dim3(nb((size_t)r * B)), dim3(NT), 0, s, Z.as<float>(), bias, r, B,
                         H.as<float>());
      TNAD_HIP_TRY(hipGetLastError());
      acts.push_back(std::move(H));
      h = acts.back().as<float>();
    } else {
      tc_launch(c, Prec::F32, gemm_desc('N', 'N', r, B, k, r, k, r), W, h, Y, s);
      hipLaunchKernelGGL(bias_add_kernel, dim3(nb((size_t)r * B)), dim3(NT), 0, s, Y, bias, r, B);
      TNAD_HIP_TRY(hipGetLastError());
    }
  }
}

// loss = mean_b 0.5 |y_b - t_b|^2 into *loss_dev; gradient into M.grad
static void mlp_loss_grad(Ctx* c, Mlp& M, const float* X, const float* T, int B, double* loss_dev) {
  hipStream_t s = c->stream;
  std::vector<DevBuf> acts;
  DevBuf Y(c, (size_t)M.n_out * B * 4), dY(c, (size_t)M.n_out * B * 4);
  mlp_forward(c, M, X, B, acts, Y.as<float>());
  const size_t ny = (size_t)M.n_out * B;
  const unsigned np = std::min<unsigned>(nb(ny), 1024);
  DevBuf part(c, (size_t)np * 8);
  hipLaunchKernelGGL(resid_loss_kernel, dim3(np), dim3(NT), 0, s, Y.as<float>(), T, ny, 1.0f / (float)B,
                     dY.as<float>(), part.as<double>());
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(NT), 0, s, part.as<double>(), (int)np, 0.5 / (double)B,
                     loss_dev);
  TNAD_HIP_TRY(hipGetLastError());
  float* g = M.grad.as<float>();
  const float* th = M.theta.as<float>();
  DevBuf dcur = std::move(dY);
  for (int l = M.layers; l >= 0; --l) {
    const int r = M.rows(l), k = M.cols(l);
    const float* hin = (l == 0) ? X : acts[l - 1].as<float>();
    // dW = dZ h^T ; db = rowsum dZ
    tc_launch(c, Prec::F32, gemm_desc('N', 'T', r, k, B, r, k, r), dcur.ptr, hin, g + M.off[2 * l], s);
    hipLaunchKernelGGL(rowsum_kernel, dim3((r + NT - 1) / NT), dim3(NT), 0, s, dcur.as<float>(), r, B,
                       g + M.off[2 * l + 1]);
    TNAD_HIP_TRY(hipGetLastError());
    if (l == 0) break;
    // dh = W^T dZ ; dZ_prev = dh (1 - h^2)
    DevBuf dh(c, (size_t)k * B * 4), dz(c, (size_t)k * B * 4);
    tc_launch(c, Prec::F32, gemm_desc('T', 'N', k, B, r, r, r, k), th + M.off[2 * l], dcur.ptr, dh.ptr, s);
    hipLaunchKernelGGL(tanh_back_kernel, dim3(nb((size_t)k * B)), dim3(NT), 0, s, dh.as<float>(), hin,
                       (size_t)k * B, dz.as<float>());
    TNAD_HIP_TRY(hipGetLastError());
    dcur = std::move(dz);
  }
}

}  // namespace tnad

using namespace tnad;

struct tnad_mlp {
  tnad_ctx* ctx = nullptr;
  Mlp m;
};

extern "C" {

int tnad_mlp_create(tnad_ctx* ctx, int n_in, int hidden, int n_layers, int n_out, unsigned long long seed,
                    tnad_mlp** out) {
  if (!out) return TNAD_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<tnad_mlp> net(new tnad_mlp());
  int rc = guarded(ctx, [&] {
    if (n_in < 1 || hidden < 1 || n_layers < 1 || n_out < 1) throw Error(TNAD_ERR_ARG, "mlp: bad sizes");
    Ctx* c = &ctx->c;
    Mlp& M = net->m;
    M.n_in = n_in; M.hidden = hidden; M.layers = n_layers; M.n_out = n_out;
    mlp_layout(M);
    M.theta = DevBuf(c, M.n_params * 4);
    M.m = DevBuf(c, M.n_params * 4);
    M.v = DevBuf(c, M.n_params * 4);
    M.grad = DevBuf(c, M.n_params * 4);
    // weights uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0
    for (int l = 0; l <= n_layers; ++l) {
      const size_t nw = (size_t)M.rows(l) * M.cols(l);
      const float a = 1.0f / std::sqrt((float)M.cols(l));
      hipLaunchKernelGGL(init_uniform_kernel, dim3(nb(nw)), dim3(NT), 0, c->stream, M.theta.as<float>() + M.off[2 * l],
                         nw, a, (unsigned long long)(seed * 1000003ull + l));
      hipLaunchKernelGGL(fill_f_kernel, dim3(nb(M.rows(l))), dim3(NT), 0, c->stream,
                         M.theta.as<float>() + M.off[2 * l + 1], (size_t)M.rows(l), 0.0f);
    }
    hipLaunchKernelGGL(fill_f_kernel, dim3(nb(M.n_params)), dim3(NT), 0, c->stream, M.m.as<float>(), M.n_params, 0.0f);
    hipLaunchKernelGGL(fill_f_kernel, dim3(nb(M.n_params)), dim3(NT), 0, c->stream, M.v.as<float>(), M.n_params, 0.0f);
    TNAD_HIP_TRY(hipGetLastError());
  });
  if (rc == TNAD_OK) {
    net->ctx = ctx;
    *out = net.release();
  }
  return rc;
}

int tnad_mlp_free(tnad_mlp* net) {
  if (!net) return TNAD_ERR_ARG;
  tnad_ctx* ctx = net->ctx;
  if (ctx) (void)hipSetDevice(ctx->c.device);
  delete net;
  return TNAD_OK;
}

int tnad_mlp_num_params(const tnad_mlp* net, long long* n) {
  if (!net || !n) return TNAD_ERR_ARG;
  *n = (long long)net->m.n_params;
  return TNAD_OK;
}

int tnad_mlp_get_params(tnad_ctx* ctx, const tnad_mlp* net, float* theta) {
  return guarded(ctx, [&] {
    if (!net || net->ctx != ctx) throw Error(TNAD_ERR_ARG, "network does not belong to this context");
    copy_out(&ctx->c, theta, net->m.theta.ptr, net->m.n_params * 4);
  });
}

int tnad_mlp_set_params(tnad_ctx* ctx, tnad_mlp* net, const float* theta) {
  return guarded(ctx, [&] {
    if (!net || net->ctx != ctx) throw Error(TNAD_ERR_ARG, "network does not belong to this context");
    Ctx* c = &ctx->c;
    Mlp& M = net->m;
    copy_in(c, M.theta.ptr, theta, M.n_params * 4);
    hipLaunchKernelGGL(fill_f_kernel, dim3(nb(M.n_params)), dim3(NT), 0, c->stream, M.m.as<float>(), M.n_params, 0.0f);
    hipLaunchKernelGGL(fill_f_kernel, dim3(nb(M.n_params)), dim3(NT), 0, c->stream, M.v.as<float>(), M.n_params, 0.0f);
    TNAD_HIP_TRY(hipGetLastError());
    M.t = 0;
  });
}

int tnad_mlp_forward(tnad_ctx* ctx, const tnad_mlp* net, const float* X, int B, float* Y) {
  return guarded(ctx, [&] {
    if (!net || net->ctx != ctx) throw Error(TNAD_ERR_ARG, "network does not belong to this context");
    if (B < 1) throw Error(TNAD_ERR_ARG, "mlp: batch must be positive");
    Ctx* c = &ctx->c;
    const Mlp& M = net->m;
    DevBuf bx(c, (size_t)M.n_in * B * 4), by(c, (size_t)M.n_out * B * 4);
    copy_in(c, bx.ptr, X, (size_t)M.n_in * B * 4);
    std::vector<DevBuf> acts;
    mlp_forward(c, M, bx.as<float>(), B, acts, by.as<float>());
    copy_out(c, Y, by.ptr, (size_t)M.n_out * B * 4);
  });
}

int tnad_mlp_loss_grad(tnad_ctx* ctx, tnad_mlp* net, const float* X, const float* T, int B, double* loss,
                       float* grad) {
  return guarded(ctx, [&] {
    if (!net || net->ctx != ctx) throw Error(TNAD_ERR_ARG, "network does not belong to this context");
    if (B < 1) throw Error(TNAD_ERR_ARG, "mlp: batch must be positive");
    Ctx* c = &ctx->c;
    Mlp& M = net->m;
    DevBuf bx(c, (size_t)M.n_in * B * 4), bt(c, (size_t)M.n_out * B * 4), bl(c, 16);
    copy_in(c, bx.ptr, X, (size_t)M.n_in * B * 4);
    copy_in(c, bt.ptr, T, (size_t)M.n_out * B * 4);
    mlp_loss_grad(c, M, bx.as<float>(), bt.as<float>(), B, bl.as<double>());
    if (loss) copy_out(c, loss, bl.ptr, 8);
    if (grad) copy_out(c, grad, M.grad.ptr, M.n_params * 4);
  });
}

int tnad_mlp_train_step(tnad_ctx* ctx, tnad_mlp* net, const float* X, const float* T, int B, float lr,
                        double* loss) {
  return guarded(ctx, [&] {
    if (!net || net->ctx != ctx) throw Error(TNAD_ERR_ARG, "network does not belong to this context");
    if (B < 1) throw Error(TNAD_ERR_ARG, "mlp: batch must be positive");
    Ctx* c = &ctx->c;
    Mlp& M = net->m;
    DevBuf bx(c, (size_t)M.n_in * B * 4), bt(c, (size_t)M.n_out * B * 4), bl(c, 16);
    copy_in(c, bx.ptr, X, (size_t)M.n_in * B * 4);
    copy_in(c, bt.ptr, T, (size_t)M.n_out * B * 4);
    mlp_loss_grad(c, M, bx.as<float>(), bt.as<float>(), B, bl.as<double>());
    M.t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)M.t), c2 = 1.0f - std::pow(b2, (float)M.t);
    hipLaunchKernelGGL(adam_kernel, dim3(nb(M.n_params)), dim3(NT), 0, c->stream, M.theta.as<float>(),
                       M.grad.as<float>(), M.m.as<float>(), M.v.as<float>(), M.n_params, lr, b1, b2, c1, c2, 1e-8f);
    TNAD_HIP_TRY(hipGetLastError());
    if (loss) copy_out(c, loss, bl.ptr, 8);
  });
}

int tnad_bulk_features(tnad_ctx* ctx, const double* bulk, int d, int s, float* feat) {
  return guarded(ctx, [&] {
    if (d < 1 || s < 1) throw Error(TNAD_ERR_SHAPE, "bulk dimensions must be positive");
    Ctx* c = &ctx->c;
    hipStream_t st = c->stream;
    const int N = d * d * d * d * s;
    const int D = d * d;
    const size_t nA = (size_t)D * D * D * D;
    DevBuf bb(c, (size_t)N * 8), sym(c, (size_t)N * 8), a(c, nA * 8), red(c, 300 * 8), out(c, nA * 4);
    copy_in(c, bb.ptr, bulk, (size_t)N * 8);
    k_symmetrize_bulk(bb.as<double>(), d, s, sym.as<double>(), st);
    k_double_layer(sym.as<double>(), d, s, nullptr, a.as<double>(), st);
    k_sumsq(a.as<double>(), nA, red.as<double>(), st);
    hipLaunchKernelGGL(scale_cast_kernel, dim3(nb(nA)), dim3(NT), 0, st, a.as<double>(), nA, red.as<double>(),
                       out.as<float>());
    TNAD_HIP_TRY(hipGetLastError());
    copy_out(c, feat, out.ptr, nA * 4);
  });
}

int tnad_env_features(tnad_ctx* ctx, const double* C, const double* T, int chi, int D, float* feat) {
  return guarded(ctx, [&] {
    if (chi < 1 || D < 1) throw Error(TNAD_ERR_SHAPE, "bond dimensions must be positive");
    if (chi > JACOBI_MAX) throw Error(TNAD_ERR_UNSUPPORTED, "env_features: chi above 192");
    Ctx* c = &ctx->c;
    hipStream_t s = c->stream;
    const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf bC(c, nC * 8), bT(c, nT * 8), lam(c, (size_t)chi * 8), Ov(c, nC * 8),
        Vw(c, jacobi_work_doubles(chi, 60) * 8), Y(c, nT * 8), Tt(c, nT * 8), sg(c, (size_t)(chi + 2) * 8),
        out(c, (nT + chi) * 4);
    copy_in(c, bC.ptr, C, nC * 8);
    copy_in(c, bT.ptr, T, nT * 8);
    jacobi_eig_ws(bC.as<double>(), chi, chi, lam.as<double>(), Ov.as<double>(), Vw.as<double>(), 60, nullptr, s);
    const long long lc = chi, lD = D;
    // T~[i,a,k] = sum_{x,y} O[x,i] T[x,a,y] O[y,k]
    tc(c, Prec::F64, cm("xi", {lc, lc}), Ov.ptr, cm("xay", {lc, lD, lc}), bT.ptr, cm("iay", {lc, lD, lc}), Y.ptr);
    tc(c, Prec::F64, cm("iay", {lc, lD, lc}), Y.ptr, cm("yk", {lc, lc}), Ov.ptr, cm("iak", {lc, lD, lc}), Tt.ptr);
    hipLaunchKernelGGL(canon_signs_kernel, dim3(1), dim3(NT), 0, s, Tt.as<double>(), lam.as<double>(), chi, D,
                       sg.as<double>());
    hipLaunchKernelGGL(env_feat_kernel, dim3(nb(nT + chi)), dim3(NT), 0, s, lam.as<double>(), Tt.as<double>(),
                       sg.as<double>(), chi, D, out.as<float>());
    TNAD_HIP_TRY(hipGetLastError());
    copy_out(c, feat, out.ptr, (nT + chi) * 4);
  });
}

int tnad_env_from_features(tnad_ctx* ctx, const float* feat, int chi, int D, double* C0, double* T0) {
  return guarded(ctx, [&] {
    if (chi < 1 || D < 1) throw Error(TNAD_ERR_SHAPE, "bond dimensions must be positive");
    Ctx* c = &ctx->c;
    hipStream_t s = c->stream;
    const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf bf(c, (nT + chi) * 4), bC(c, nC * 8), bT(c, nT * 8), red(c, 300 * 8), nrm(c, 2 * 8);
    copy_in(c, bf.ptr, feat, (nT + chi) * 4);
    hipLaunchKernelGGL(env_decode_kernel, dim3(nb(nC + nT)), dim3(NT), 0, s, bf.as<float>(), chi, D, bC.as<double>(),
                       bT.as<double>());
    TNAD_HIP_TRY(hipGetLastError());
    // C0 /= |C0| ; T0 /= |T0|
    k_sumsq(bC.as<double>(), nC, red.as<double>(), s);
    k_copy(red.as<double>(), nrm.as<double>(), 1, s);
    k_sumsq(bT.as<double>(), nT, red.as<double>(), s);
    k_copy(red.as<double>(), nrm.as<double>() + 1, 1, s);
    std::vector<double> h(2);
    c->read_async(h.data(), nrm.ptr, 16);
    c->sync();
    if (!(h[0] > 0.0) || !(h[1] > 0.0)) throw Error(TNAD_ERR_NAN, "env_from_features: zero or NaN environment");
    k_axpby(1.0 / std::sqrt(h[0]), bC.as<double>(), 0.0, bC.as<double>(), nC, s);
    k_axpby(1.0 / std::sqrt(h[1]), bT.as<double>(), 0.0, bT.as<double>(), nT, s);
    copy_out(c, C0, bC.ptr, nC * 8);
    copy_out(c, T0, bT.ptr, nT * 8);
  });
}

}  // extern "C"
