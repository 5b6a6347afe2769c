// Learned scheduling and paging: per-vertex features, a 16 -> 64 -> 64 -> 2
// MLP with BF16-MFMA inference and FP32 training kernels.
//
// Nothing here changes a count.  The network's two outputs per vertex are
//   y0: log2(1 + work of the jobs rooted at v)  -> job order (heaviest roots
//       first) in the edge phase;
//   y1: log2(1 + rounds in which v's package was used by a paged run)
//       -> package scores for eviction and prefetch.
// Both targets are traces the matcher records itself (pm_opts.root_work,
// pm_opts.package_use).  The reference has no learned component; its
// scheduler takes jobs in CSR order (src/job_queue.h:39-89) and its cache
// is a plain LRU (src/buffer.cu:581-796).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "pm_internal.hpp"

namespace pm {
namespace {

constexpr int kF = PM_FEATURES;  // 16
constexpr int kH = 64;
constexpr int kO = 2;
// parameter layout (floats): W1[f][h], b1[h], W2[i][j], b2[j], W3[j][o], b3[o]
constexpr int oW1 = 0, oB1 = oW1 + kF * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH * kO, kP = oB3 + kO;
static_assert(kP == PM_MODEL_PARAMS, "parameter count");

#define HIP_OK(call)                                                \
  do {                                                              \
    hipError_t _e = (call);                                         \
    if (_e != hipSuccess) {                                         \
      set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return PM_ERR_DEVICE;                                         \
    }                                                               \
  } while (0)

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------

struct FeatArgs {
  uint64_t n;
  const uint64_t* rowptr;
  const uint32_t* rows;
  const uint8_t* lab;  // null: unlabeled
  const uint32_t* maxnbr;
  float* out;          // n x kF
};

__device__ __forceinline__ uint32_t lower_idx(const uint32_t* a, uint32_t n, uint32_t x) {
  uint32_t lo = 0;
  while (n > 0) {
    const uint32_t half = n >> 1;
    if (a[lo + half] < x) {
      lo += half + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// One thread per vertex.  Features: log degree, log 2-hop degree sum, log
// max neighbour degree, log triangle estimate (|N(v) n N(w)| over up to 4
// evenly spaced neighbours w, first 64 entries of the shorter list, scaled),
// package position (4096-vertex granules over n), log id position, label
// one-hot (8), log same-label neighbours, bias.
__global__ void k_features(FeatArgs A) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < A.n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r0 = A.rowptr[v], r1 = A.rowptr[v + 1];
    const uint32_t d = (uint32_t)(r1 - r0);
    const uint32_t* nv = A.rows + r0;
    double two = 0;
    uint32_t same = 0;
    const uint8_t lv = A.lab ? A.lab[v] : 0;
    for (uint32_t i = 0; i < d; i++) {
      const uint32_t w = nv[i];
      two += (double)(A.rowptr[w + 1] - A.rowptr[w]);
      if (A.lab) same += A.lab[w] == lv;
    }
    double tri = 0;
    if (d >= 2) {
      const uint32_t ns = min(d, 4u);
      double hit = 0, tried = 0;
      for (uint32_t s = 0; s < ns; s++) {
        const uint32_t w = nv[(uint64_t)s * d / ns];
        const uint64_t w0 = A.rowptr[w];
        const uint32_t dw = (uint32_t)(A.rowptr[w + 1] - w0);
        const uint32_t* a = d <= dw ? nv : A.rows + w0;
        const uint32_t na = min(d, dw);
        const uint32_t* b = d <= dw ? A.rows + w0 : nv;
        const uint32_t nb = max(d, dw);
        const uint32_t take = min(na, 64u);
        uint32_t c = 0;
        for (uint32_t i = 0; i < take; i++) {
          const uint32_t x = a[(uint64_t)i * na / take];
          const uint32_t p = lower_idx(b, nb, x);
          c += p < nb && b[p] == x;
        }
        hit += (double)c * na / take;
        tried += 1;
      }
      tri = hit / tried * d / 2.0;
    }
    float* o = A.out + v * kF;
    o[0] = log2f(1.0f + d);
    o[1] = (float)log2(1.0 + two);
    o[2] = log2f(1.0f + (float)A.maxnbr[v]);
    o[3] = (float)log2(1.0 + tri);
    o[4] = (float)((double)(v & ~4095ull) / (double)A.n);
    o[5] = (float)(log2(1.0 + (double)v) / log2(1.0 + (double)A.n));
    for (int l = 0; l < 8; l++) o[6 + l] = (A.lab && (lv & 7) == l) ? 1.0f : 0.0f;
    o[14] = log2f(1.0f + same);
    o[15] = 1.0f;
  }
}

// ---------------------------------------------------------------------------
// inference: BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulate
// ---------------------------------------------------------------------------

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct InferArgs {
  const float* X;   // n x kF
  uint64_t n;
  const __bf16* w;  // kP parameters as bf16 (biases unused)
  const float* wf;  // kP parameters as fp32 (biases)
  float* Y;         // n x kO
};

// One wave per 16-sample tile.  Operand layout of the 16x16x32 MFMA: lane l
// supplies A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16] and receives
// C[4*(l/16) .. +3][l%16].  Activations pass between layers through LDS.
__global__ __launch_bounds__(64) void k_mlp_infer(InferArgs A) {
  __shared__ float H[16][kH + 1];
  const int lane = threadIdx.x;
  const int rc = lane & 15, kq = lane >> 4;
  const __bf16 z = (__bf16)0.0f;
  bf16x8 w1f[4], w2f[2][4], w3f[2];
  for (int t = 0; t < 4; t++)
    for (int j = 0; j < 8; j++) {
      const int k = 8 * kq + j;
      w1f[t][j] = k < kF ? A.w[oW1 + k * kH + 16 * t + rc] : z;
    }
  for (int c = 0; c < 2; c++)
    for (int t = 0; t < 4; t++)
      for (int j = 0; j < 8; j++) w2f[c][t][j] = A.w[oW2 + (32 * c + 8 * kq + j) * kH + 16 * t + rc];
  for (int c = 0; c < 2; c++)
    for (int j = 0; j < 8; j++)
      w3f[c][j] = rc < kO ? A.w[oW3 + (32 * c + 8 * kq + j) * kO + rc] : z;
  float b1[4], b2[4];
  for (int t = 0; t < 4; t++) {
    b1[t] = A.wf[oB1 + 16 * t + rc];
    b2[t] = A.wf[oB2 + 16 * t + rc];
  }
  const float b3 = rc < kO ? A.wf[oB3 + rc] : 0.0f;
  for (uint64_t tile = blockIdx.x; tile * 16 < A.n; tile += gridDim.x) {
    const uint64_t base = tile * 16;
    bf16x8 xa;
    for (int j = 0; j < 8; j++) {
      const int k = 8 * kq + j;
      const uint64_t s = base + rc;
      xa[j] = (k < kF && s < A.n) ? (__bf16)A.X[s * kF + k] : z;
    }
    for (int t = 0; t < 4; t++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, w1f[t], acc, 0, 0, 0);
      for (int i = 0; i < 4; i++) H[4 * kq + i][16 * t + rc] = fmaxf(acc[i] + b1[t], 0.0f);
    }
    __syncthreads();
    bf16x8 ha[2];
    for (int c = 0; c < 2; c++)
      for (int j = 0; j < 8; j++) ha[c][j] = (__bf16)H[rc][32 * c + 8 * kq + j];
    __syncthreads();
    for (int t = 0; t < 4; t++) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      for (int c = 0; c < 2; c++)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[c], w2f[c][t], acc, 0, 0, 0);
      for (int i = 0; i < 4; i++) H[4 * kq + i][16 * t + rc] = fmaxf(acc[i] + b2[t], 0.0f);
    }
    __syncthreads();
    for (int c = 0; c < 2; c++)
      for (int j = 0; j < 8; j++) ha[c][j] = (__bf16)H[rc][32 * c + 8 * kq + j];
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 2; c++)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[c], w3f[c], acc, 0, 0, 0);
    if (rc < kO)
      for (int i = 0; i < 4; i++) {
        const uint64_t s = base + 4 * kq + i;
        if (s < A.n) A.Y[s * kO + rc] = acc[i] + b3;
      }
    __syncthreads();
  }
}

__global__ void k_to_bf16(const float* a, __bf16* b, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = (__bf16)a[i];
}

// ---------------------------------------------------------------------------
// training: FP32 forward/backward over 64-sample tiles, deterministic
// reduction of per-tile gradients, Adam
// ---------------------------------------------------------------------------

struct GradArgs {
  const float* X;    // N x kF
  const float* Y;    // N x kO targets
  const float* Wt;   // N x kO loss weights
  const uint32_t* idx;  // samples of this step (or null: 0..nsamp-1)
  uint32_t nsamp;
  float scale;       // 1 / sum of the weights of this step
  const float* w;    // kP
  float* partial;    // tiles x kP
  float* lossp;      // tiles
};

// loss = scale * sum_s sum_o Wt (yhat - y)^2
__global__ __launch_bounds__(64) void k_mlp_grad(GradArgs A) {
  __shared__ float sw[kP];
  __shared__ float sX[64][kF + 1];
  __shared__ float sH1[64][kH + 1];
  __shared__ float sH2[64][kH + 1];
  __shared__ float sD2[64][kH + 1];
  __shared__ float sD1[64][kH + 1];
  __shared__ float sD3[64][kO];
  const int t = threadIdx.x;
  for (int i = t; i < kP; i += 64) sw[i] = A.w[i];
  const uint32_t slot = blockIdx.x * 64u + t;
  const bool live = slot < A.nsamp;
  const uint32_t s = live ? (A.idx ? A.idx[slot] : slot) : 0;
  for (int f = 0; f < kF; f++) sX[t][f] = live ? A.X[(uint64_t)s * kF + f] : 0.0f;
  __syncthreads();
  for (int h = 0; h < kH; h++) {
    float a = sw[oB1 + h];
    for (int f = 0; f < kF; f++) a += sX[t][f] * sw[oW1 + f * kH + h];
    sH1[t][h] = fmaxf(a, 0.0f);
  }
  for (int j = 0; j < kH; j++) {
    float a = sw[oB2 + j];
    for (int i = 0; i < kH; i++) a += sH1[t][i] * sw[oW2 + i * kH + j];
    sH2[t][j] = fmaxf(a, 0.0f);
  }
  float loss = 0;
  for (int o = 0; o < kO; o++) {
    float y = sw[oB3 + o];
    for (int j = 0; j < kH; j++) y += sH2[t][j] * sw[oW3 + j * kO + o];
    float d3 = 0;
    if (live) {
      const float wt = A.Wt[(uint64_t)s * kO + o];
      const float r = y - A.Y[(uint64_t)s * kO + o];
      loss += A.scale * wt * r * r;
      d3 = 2.0f * A.scale * wt * r;
    }
    sD3[t][o] = d3;
  }
  for (int j = 0; j < kH; j++) {
    float a = 0;
    for (int o = 0; o < kO; o++) a += sw[oW3 + j * kO + o] * sD3[t][o];
    sD2[t][j] = sH2[t][j] > 0.0f ? a : 0.0f;
  }
  for (int i = 0; i < kH; i++) {
    float a = 0;
    for (int j = 0; j < kH; j++) a += sw[oW2 + i * kH + j] * sD2[t][j];
    sD1[t][i] = sH1[t][i] > 0.0f ? a : 0.0f;
  }
  __syncthreads();
  float* P = A.partial + (size_t)blockIdx.x * kP;
  for (int e = t; e < kP; e += 64) {
    float g = 0;
    if (e < oB1) {
      const int f = e / kH, h = e % kH;
      for (int q = 0; q < 64; q++) g += sX[q][f] * sD1[q][h];
    } else if (e < oW2) {
      const int h = e - oB1;
      for (int q = 0; q < 64; q++) g += sD1[q][h];
    } else if (e < oB2) {
      const int i = (e - oW2) / kH, j = (e - oW2) % kH;
      for (int q = 0; q < 64; q++) g += sH1[q][i] * sD2[q][j];
    } else if (e < oW3) {
      const int j = e - oB2;
      for (int q = 0; q < 64; q++) g += sD2[q][j];
    } else if (e < oB3) {
      const int j = (e - oW3) / kO, o = (e - oW3) % kO;
      for (int q = 0; q < 64; q++) g += sH2[q][j] * sD3[q][o];
    } else {
      const int o = e - oB3;
      for (int q = 0; q < 64; q++) g += sD3[q][o];
    }
    P[e] = g;
  }
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
  if (t == 0) A.lossp[blockIdx.x] = loss;
}

// grad[e] = sum over tiles in order (deterministic)
__global__ void k_reduce_tiles(const float* partial, int tiles, float* grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kP) return;
  float g = 0;
  for (int b = 0; b < tiles; b++) g += partial[(size_t)b * kP + e];
  grad[e] = g;
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kP) return;
  m[e] = b1 * m[e] + (1.0f - b1) * g[e];
  v[e] = b2 * v[e] + (1.0f - b2) * g[e] * g[e];
  w[e] -= lr * (m[e] / c1) / (sqrtf(v[e] / c2) + eps);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

uint64_t next_model_version() {
  static std::atomic<uint64_t> v{0};
  return ++v;
}

void model_init(pm_model* m, uint64_t seed) {
  m->w.assign(kP, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int count) {
    const double a = std::sqrt(6.0 / fan_in);  // He-uniform for ReLU layers
    std::uniform_real_distribution<double> u(-a, a);
    for (int i = 0; i < count; i++) m->w[off + i] = (float)u(rng);
  };
  fill(oW1, kF, kF * kH);
  fill(oW2, kH, kH * kH);
  fill(oW3, kH, kH * kO);
  m->adam_m.assign(kP, 0.0f);
  m->adam_v.assign(kP, 0.0f);
  m->step = 0;
  m->version = next_model_version();
}

static int model_upload(const pm_model* m, float** d_w, __bf16** d_wb) {
  HIP_OK(hipMalloc(d_w, sizeof(float) * kP));
  HIP_OK(hipMalloc(d_wb, sizeof(__bf16) * kP));
  HIP_OK(hipMemcpy(*d_w, m->w.data(), sizeof(float) * kP, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3(8), dim3(256), 0, 0, *d_w, *d_wb, kP);
  HIP_OK(hipGetLastError());
  return PM_OK;
}

int model_predict_device(const pm_model* m, const float* d_X, uint64_t n, float* d_Y,
                         hipStream_t st) {
  if (n == 0) return PM_OK;
  float* d_w = nullptr;
  __bf16* d_wb = nullptr;
  int rc = model_upload(m, &d_w, &d_wb);
  if (rc) return rc;
  InferArgs A{d_X, n, d_wb, d_w, d_Y};
  const uint64_t tiles = (n + 15) / 16;
  const int blocks = (int)std::min<uint64_t>(tiles, 8192);
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64), 0, st, A);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(st));
  HIP_OK(hipFree(d_w));
  HIP_OK(hipFree(d_wb));
  return PM_OK;
}

int model_predict_host(const pm_model* m, const float* X, uint64_t n, float* Y) {
  if (n == 0) return PM_OK;
  float *d_X = nullptr, *d_Y = nullptr;
  HIP_OK(hipMalloc(&d_X, sizeof(float) * kF * n));
  HIP_OK(hipMalloc(&d_Y, sizeof(float) * kO * n));
  HIP_OK(hipMemcpy(d_X, X, sizeof(float) * kF * n, hipMemcpyHostToDevice));
  int rc = model_predict_device(m, d_X, n, d_Y, 0);
  if (rc) return rc;
  HIP_OK(hipMemcpy(Y, d_Y, sizeof(float) * kO * n, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(d_X));
  HIP_OK(hipFree(d_Y));
  return PM_OK;
}

namespace {
struct TrainBuffers {
  float *X = nullptr, *Y = nullptr, *Wt = nullptr, *w = nullptr, *grad = nullptr;
  float *partial = nullptr, *lossp = nullptr, *am = nullptr, *av = nullptr;
  uint32_t* idx = nullptr;
  int release() {
    float* fs[] = {X, Y, Wt, w, grad, partial, lossp, am, av};
    for (float* f : fs)
      if (f) HIP_OK(hipFree(f));
    if (idx) HIP_OK(hipFree(idx));
    return PM_OK;
  }
};

int train_setup(TrainBuffers& B, const pm_model* m, const float* X, const float* Y,
                const float* Wt, uint64_t n, uint32_t batch) {
  const uint32_t tiles = (batch + 63) / 64;
  HIP_OK(hipMalloc(&B.X, sizeof(float) * kF * std::max<uint64_t>(n, 1)));
  HIP_OK(hipMalloc(&B.Y, sizeof(float) * kO * std::max<uint64_t>(n, 1)));
  HIP_OK(hipMalloc(&B.Wt, sizeof(float) * kO * std::max<uint64_t>(n, 1)));
  HIP_OK(hipMalloc(&B.w, sizeof(float) * kP));
  HIP_OK(hipMalloc(&B.grad, sizeof(float) * kP));
  HIP_OK(hipMalloc(&B.am, sizeof(float) * kP));
  HIP_OK(hipMalloc(&B.av, sizeof(float) * kP));
  HIP_OK(hipMalloc(&B.partial, sizeof(float) * kP * tiles));
  HIP_OK(hipMalloc(&B.lossp, sizeof(float) * tiles));
  HIP_OK(hipMalloc(&B.idx, sizeof(uint32_t) * std::max<uint32_t>(batch, 1)));
  if (n) {
    HIP_OK(hipMemcpy(B.X, X, sizeof(float) * kF * n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(B.Y, Y, sizeof(float) * kO * n, hipMemcpyHostToDevice));
    if (Wt) {
      HIP_OK(hipMemcpy(B.Wt, Wt, sizeof(float) * kO * n, hipMemcpyHostToDevice));
    } else {
      std::vector<float> ones(kO * n, 1.0f);
      HIP_OK(hipMemcpy(B.Wt, ones.data(), sizeof(float) * kO * n, hipMemcpyHostToDevice));
    }
  }
  HIP_OK(hipMemcpy(B.w, m->w.data(), sizeof(float) * kP, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.am, m->adam_m.data(), sizeof(float) * kP, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B.av, m->adam_v.data(), sizeof(float) * kP, hipMemcpyHostToDevice));
  return PM_OK;
}

// gradient of the weighted mean squared error over the samples idx (or
// 0..count-1), into B.grad; returns the loss
int grad_step(TrainBuffers& B, const uint32_t* d_idx, uint32_t count, float scale, double* loss) {
  const uint32_t tiles = std::max<uint32_t>((count + 63) / 64, 1);
  GradArgs A{B.X, B.Y, B.Wt, d_idx, count, scale, B.w, B.partial, B.lossp};
  hipLaunchKernelGGL(k_mlp_grad, dim3(tiles), dim3(64), 0, 0, A);
  hipLaunchKernelGGL(k_reduce_tiles, dim3((kP + 255) / 256), dim3(256), 0, 0, B.partial,
                     (int)tiles, B.grad);
  HIP_OK(hipGetLastError());
  if (loss) {
    std::vector<float> lp(tiles);
    HIP_OK(hipMemcpy(lp.data(), B.lossp, sizeof(float) * tiles, hipMemcpyDeviceToHost));
    double s = 0;
    for (float x : lp) s += x;
    *loss = s;
  }
  return PM_OK;
}
}  // namespace

int model_grad(const pm_model* m, const float* X, const float* Y, const float* Wt, uint64_t n,
               float* grad, double* loss) {
  if (n > (1u << 24)) {
    set_error("pm_model_grad: at most 2^24 samples per call");
    return PM_ERR_ARG;
  }
  TrainBuffers B;
  int rc = train_setup(B, m, X, Y, Wt, n, (uint32_t)n);
  if (rc) return rc;
  double wsum = 0;
  for (uint64_t i = 0; i < kO * n; i++) wsum += Wt ? Wt[i] : 1.0;
  rc = grad_step(B, nullptr, (uint32_t)n, wsum > 0 ? (float)(1.0 / wsum) : 0.0f, loss);
  if (rc) return rc;
  HIP_OK(hipMemcpy(grad, B.grad, sizeof(float) * kP, hipMemcpyDeviceToHost));
  return B.release();
}

int model_train(pm_model* m, const float* X, const float* Y, const float* Wt, uint64_t n,
                const pm_train_opts* o, double* final_loss) {
  pm_train_opts t;
  std::memset(&t, 0, sizeof(t));
  if (o) t = *o;
  const int epochs = t.epochs > 0 ? t.epochs : 20;
  const uint32_t batch = (uint32_t)std::min<uint64_t>(t.batch > 0 ? t.batch : 4096, std::max<uint64_t>(n, 1));
  const float lr = t.lr > 0 ? t.lr : 3e-3f;
  if (n == 0 || n > (1ull << 31)) {
    set_error("pm_model_train: need 1..2^31 samples");
    return PM_ERR_ARG;
  }
  TrainBuffers B;
  int rc = train_setup(B, m, X, Y, Wt, n, batch);
  if (rc) return rc;
  std::vector<float> wrow(n, 0.0f);
  for (uint64_t i = 0; i < n; i++)
    for (int q = 0; q < kO; q++) wrow[i] += Wt ? Wt[i * kO + q] : 1.0f;
  std::mt19937_64 rng(t.seed ? t.seed : 1);
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  double last = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    double ep_loss = 0, ep_w = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, n - b0);
      double ws = 0;
      for (uint32_t i = 0; i < cnt; i++) ws += wrow[perm[b0 + i]];
      if (ws <= 0) continue;
      HIP_OK(hipMemcpy(B.idx, perm.data() + b0, sizeof(uint32_t) * cnt, hipMemcpyHostToDevice));
      double l = 0;
      rc = grad_step(B, B.idx, cnt, (float)(1.0 / ws), &l);
      if (rc) return rc;
      m->step++;
      const float c1 = 1.0f - std::pow(0.9f, (float)m->step);
      const float c2 = 1.0f - std::pow(0.999f, (float)m->step);
      hipLaunchKernelGGL(k_adam, dim3((kP + 255) / 256), dim3(256), 0, 0, B.w, B.grad, B.am, B.av,
                         lr, 0.9f, 0.999f, 1e-8f, c1, c2);
      HIP_OK(hipGetLastError());
      ep_loss += l * ws;
      ep_w += ws;
    }
    last = ep_w > 0 ? ep_loss / ep_w : 0;
  }
  HIP_OK(hipMemcpy(m->w.data(), B.w, sizeof(float) * kP, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(m->adam_m.data(), B.am, sizeof(float) * kP, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(m->adam_v.data(), B.av, sizeof(float) * kP, hipMemcpyDeviceToHost));
  m->version = next_model_version();
  if (final_loss) *final_loss = last;
  return B.release();
}

// Features of every vertex on the device (d_out: n x kF).  Rows not resident
// in HBM are uploaded for the pass and freed after it.
int device_features(pm_graph* g, float* d_out, hipStream_t st) {
  int rc = device_upload_rows(g, false);
  if (rc) return rc;
  const HostGraph& h = g->h;
  if (h.n == 0) return PM_OK;
  uint32_t* rows = g->d.colidx;
  uint32_t* tmp = nullptr;
  if (!rows) {
    HIP_OK(hipMalloc(&tmp, sizeof(uint32_t) * std::max<uint64_t>(h.m, 1)));
    if (h.m) HIP_OK(hipMemcpy(tmp, h.colidx.data(), sizeof(uint32_t) * h.m, hipMemcpyHostToDevice));
    rows = tmp;
  }
  FeatArgs A{h.n, g->d.rowptr, rows, h.labels.empty() ? nullptr : g->d.label8, g->d.maxnbr, d_out};
  const int blocks = (int)std::min<uint64_t>((h.n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_features, dim3(blocks), dim3(256), 0, st, A);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(st));
  if (tmp) HIP_OK(hipFree(tmp));
  return PM_OK;
}

int graph_features_host(pm_graph* g, float* out) {
  const uint64_t n = g->h.n;
  if (n == 0) return PM_OK;
  float* d = nullptr;
  HIP_OK(hipMalloc(&d, sizeof(float) * kF * n));
  int rc = device_features(g, d, 0);
  if (rc) return rc;
  HIP_OK(hipMemcpy(out, d, sizeof(float) * kF * n, hipMemcpyDeviceToHost));
  HIP_OK(hipFree(d));
  return PM_OK;
}

// Predictions of model m for every vertex of g (host n x kO), cached per
// (graph, model version).
int graph_predictions(pm_graph* g, const pm_model* m, const std::vector<float>** out) {
  if (g->pred_model == m && g->pred_version == m->version && g->pred.size() == kO * g->h.n) {
    *out = &g->pred;
    return PM_OK;
  }
  const uint64_t n = g->h.n;
  g->pred.assign(kO * n, 0.0f);
  if (n) {
    float *d_X = nullptr, *d_Y = nullptr;
    HIP_OK(hipMalloc(&d_X, sizeof(float) * kF * n));
    HIP_OK(hipMalloc(&d_Y, sizeof(float) * kO * n));
    int rc = device_features(g, d_X, 0);
    if (rc) return rc;
    rc = model_predict_device(m, d_X, n, d_Y, 0);
    if (rc) return rc;
    HIP_OK(hipMemcpy(g->pred.data(), d_Y, sizeof(float) * kO * n, hipMemcpyDeviceToHost));
    HIP_OK(hipFree(d_X));
    HIP_OK(hipFree(d_Y));
  }
  g->pred_model = m;
  g->pred_version = m->version;
  *out = &g->pred;
  return PM_OK;
}

}  // namespace pm
