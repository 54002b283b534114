// costmodel.hip -- the autotuner and its learned cost model.
//
// Autotuner: for every size class of a job set (same transform length 2^L,
// same kind), time candidate plans on the class's own jobs
// (sba_class_time) and store the fastest in the plan table for (L, kind,
// bucket of the job count). Candidates: every valid plan (exhaustive), or
// the k plans the cost model ranks fastest (k = 1: no timing at all).
//
// Cost model: a multilayer perceptron over CM_IN plan features,
//   f(x) = w3 . relu(W2 relu(W1 x + b1) + b2) + b3,
// predicting the standardised log launch time of a plan relative to the
// other plans of its class, from standardised features. Training (mean squared error, Adam) runs on the GPU
// through the kernels below: k_cm_forward (activations), k_cm_delta
// (back-propagated errors per sample), k_cm_grad (parameter gradients, one
// thread per parameter, samples summed in a fixed order) and k_cm_adam.
// All arithmetic is fp32.
#include "simd_bigint_amd.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

#define CM_IN SBA_CM_IN
#define CM_H1 32
#define CM_H2 32
// parameter layout (floats): W1 [H1][IN], b1 [H1], W2 [H2][H1], b2 [H2], w3 [H2], b3
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + CM_H1 * CM_IN)
#define OFF_W2 (OFF_B1 + CM_H1)
#define OFF_B2 (OFF_W2 + CM_H2 * CM_H1)
#define OFF_W3 (OFF_B2 + CM_H2)
#define OFF_B3 (OFF_W3 + CM_H2)
#define CM_NPARAM (OFF_B3 + 1)
#define CM_AS (CM_H1 + CM_H2)            // activations kept per sample
#define CM_DS (1 + CM_H1 + CM_H2)        // errors kept per sample
#define CM_NT 256

static_assert(CM_NPARAM == SBA_CM_NPARAM, "parameter count");

__device__ __forceinline__ void cm_stage(float *sp, const float *P) {
  for (int i = threadIdx.x; i < CM_NPARAM; i += blockDim.x) sp[i] = P[i];
  __syncthreads();
}

// one thread per sample: y = f(x); A (if not null) keeps a1 and a2
__global__ __launch_bounds__(CM_NT) void k_cm_forward(const float *__restrict__ P,
                                                      const float *__restrict__ X, int n,
                                                      float *__restrict__ Y,
                                                      float *__restrict__ A) {
  __shared__ float sp[CM_NPARAM];
  cm_stage(sp, P);
  const int s = blockIdx.x * CM_NT + threadIdx.x;
  if (s >= n) return;
  float x[CM_IN], a1[CM_H1];
#pragma unroll
  for (int i = 0; i < CM_IN; ++i) x[i] = X[(size_t)s * CM_IN + i];
#pragma unroll
  for (int j = 0; j < CM_H1; ++j) {
    float z = sp[OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < CM_IN; ++i) z += sp[OFF_W1 + j * CM_IN + i] * x[i];
    a1[j] = z > 0.f ? z : 0.f;
  }
  float y = sp[OFF_B3];
#pragma unroll 4
  for (int k = 0; k < CM_H2; ++k) {
    float z = sp[OFF_B2 + k];
#pragma unroll
    for (int j = 0; j < CM_H1; ++j) z += sp[OFF_W2 + k * CM_H1 + j] * a1[j];
    const float a2 = z > 0.f ? z : 0.f;
    y += sp[OFF_W3 + k] * a2;
    if (A) A[(size_t)s * CM_AS + CM_H1 + k] = a2;
  }
  if (A) {
#pragma unroll
    for (int j = 0; j < CM_H1; ++j) A[(size_t)s * CM_AS + j] = a1[j];
  }
  Y[s] = y;
}

// one thread per sample: errors dLoss/dy, dLoss/dz2, dLoss/dz1 for the
// loss (1/n) sum (y - t)^2
__global__ __launch_bounds__(CM_NT) void k_cm_delta(const float *__restrict__ P,
                                                    const float *__restrict__ Y,
                                                    const float *__restrict__ T,
                                                    const float *__restrict__ A, int n,
                                                    float *__restrict__ D) {
  __shared__ float sp[CM_NPARAM];
  cm_stage(sp, P);
  const int s = blockIdx.x * CM_NT + threadIdx.x;
  if (s >= n) return;
  const float e = 2.f * (Y[s] - T[s]) / (float)n;
  float d2[CM_H2];
  float *ds = D + (size_t)s * CM_DS;
  const float *as = A + (size_t)s * CM_AS;
  ds[0] = e;
#pragma unroll
  for (int k = 0; k < CM_H2; ++k) {
    d2[k] = as[CM_H1 + k] > 0.f ? e * sp[OFF_W3 + k] : 0.f;
    ds[1 + CM_H1 + k] = d2[k];
  }
#pragma unroll 4
  for (int j = 0; j < CM_H1; ++j) {
    float g = 0.f;
#pragma unroll
    for (int k = 0; k < CM_H2; ++k) g += sp[OFF_W2 + k * CM_H1 + j] * d2[k];
    ds[1 + j] = as[j] > 0.f ? g : 0.f;
  }
}

// one thread per parameter: gradient summed over the samples in order
__global__ __launch_bounds__(CM_NT) void k_cm_grad(const float *__restrict__ X,
                                                   const float *__restrict__ A,
                                                   const float *__restrict__ D, int n,
                                                   float *__restrict__ G) {
  const int i = blockIdx.x * CM_NT + threadIdx.x;
  if (i >= CM_NPARAM) return;
  float g = 0.f;
  if (i < OFF_B1) {
    const int j = i / CM_IN, c = i % CM_IN;
    for (int s = 0; s < n; ++s) g += D[(size_t)s * CM_DS + 1 + j] * X[(size_t)s * CM_IN + c];
  } else if (i < OFF_W2) {
    const int j = i - OFF_B1;
    for (int s = 0; s < n; ++s) g += D[(size_t)s * CM_DS + 1 + j];
  } else if (i < OFF_B2) {
    const int k = (i - OFF_W2) / CM_H1, j = (i - OFF_W2) % CM_H1;
    for (int s = 0; s < n; ++s)
      g += D[(size_t)s * CM_DS + 1 + CM_H1 + k] * A[(size_t)s * CM_AS + j];
  } else if (i < OFF_W3) {
    const int k = i - OFF_B2;
    for (int s = 0; s < n; ++s) g += D[(size_t)s * CM_DS + 1 + CM_H1 + k];
  } else if (i < OFF_B3) {
    const int k = i - OFF_W3;
    for (int s = 0; s < n; ++s) g += D[(size_t)s * CM_DS] * A[(size_t)s * CM_AS + CM_H1 + k];
  } else {
    for (int s = 0; s < n; ++s) g += D[(size_t)s * CM_DS];
  }
  G[i] = g;
}

// Adam (Kingma & Ba 2015); c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
__global__ __launch_bounds__(CM_NT) void k_cm_adam(float *__restrict__ P,
                                                   const float *__restrict__ G,
                                                   float *__restrict__ M, float *__restrict__ V,
                                                   float lr, float b1, float b2, float eps,
                                                   float c1, float c2) {
  const int i = blockIdx.x * CM_NT + threadIdx.x;
  if (i >= CM_NPARAM) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.f - b1) * g;
  const float v = b2 * V[i] + (1.f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  P[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// ---------------------------------------------------------------------------
// C ABI: training steps on caller-owned device buffers
// ---------------------------------------------------------------------------
#define CMCHK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "simd_bigint_amd: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return SBA_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

static unsigned cm_blocks(int n) { return (unsigned)((n + CM_NT - 1) / CM_NT); }

extern "C" int sba_cm_nparams(void) { return CM_NPARAM; }

extern "C" size_t sba_cm_ws_floats(int n) { return n <= 0 ? 0 : (size_t)n * (CM_AS + CM_DS + 1); }

extern "C" int sba_cm_forward(const float *params, const float *x, int n, float *y, void *stream) {
  if (n <= 0 || !params || !x || !y) return SBA_BAND;
  hipLaunchKernelGGL(k_cm_forward, dim3(cm_blocks(n)), dim3(CM_NT), 0, (hipStream_t)stream,
                     params, x, n, y, (float *)nullptr);
  CMCHK(hipGetLastError());
  return SBA_OK;
}

// forward with activations, errors and the gradient of the mean squared
// error over the n samples into grad[CM_NPARAM]; ws holds sba_cm_ws_floats(n)
// floats (the predictions come first: ws[0..n))
extern "C" int sba_cm_backward(const float *params, const float *x, const float *t, int n,
                               float *grad, float *ws, void *stream) {
  if (n <= 0 || !params || !x || !t || !grad || !ws) return SBA_BAND;
  hipStream_t st = (hipStream_t)stream;
  float *Y = ws, *A = ws + n, *D = A + (size_t)n * CM_AS;
  hipLaunchKernelGGL(k_cm_forward, dim3(cm_blocks(n)), dim3(CM_NT), 0, st, params, x, n, Y, A);
  hipLaunchKernelGGL(k_cm_delta, dim3(cm_blocks(n)), dim3(CM_NT), 0, st, params, Y, t, A, n, D);
  hipLaunchKernelGGL(k_cm_grad, dim3(cm_blocks(CM_NPARAM)), dim3(CM_NT), 0, st, x, A, D, n, grad);
  CMCHK(hipGetLastError());
  return SBA_OK;
}

extern "C" int sba_cm_adam(float *params, const float *grad, float *m, float *v, int step,
                           float lr, void *stream) {
  if (step < 1 || !params || !grad || !m || !v) return SBA_BAND;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = (float)(1.0 / (1.0 - pow(b1, step)));
  const float c2 = (float)(1.0 / (1.0 - pow(b2, step)));
  hipLaunchKernelGGL(k_cm_adam, dim3(cm_blocks(CM_NPARAM)), dim3(CM_NT), 0, (hipStream_t)stream,
                     params, grad, m, v, lr, b1, b2, eps, c1, c2);
  CMCHK(hipGetLastError());
  return SBA_OK;
}

// ---------------------------------------------------------------------------
// plan features (host): what the model sees of a (size class, plan) pair
// ---------------------------------------------------------------------------
extern "C" int sba_cm_features(int L, int sqr, int njobs, double fill, int last, int bm,
                               int split, float *f) {
  if (L < 10 || L > sba_gpu_max_lg() || njobs < 1 || !f) return SBA_BAND;
  const double lj = log2((double)njobs);
  const int single = last == L;
  const int ell1 = single ? 0 : L - last;
  const double k5 = log2(5.0);
  f[0] = (float)L;
  f[1] = (float)(sqr ? 1 : 0);
  f[2] = (float)lj;
  f[3] = (float)fill;
  f[4] = (float)(L + lj);                                   // log2 residues per prime
  f[5] = (float)single;
  f[6] = (float)ell1;                                       // first-pass levels
  f[7] = (float)last;
  f[8] = (float)(bm == 3);
  f[9] = (float)(bm == 4);
  f[10] = (float)(bm == 5);
  f[11] = (float)(split ? 1 : 0);
  f[12] = (float)((single ? 0 : L - 13) + lj + (split ? k5 : 0.0));   // log2 first-pass blocks
  f[13] = (float)(L - last + lj + k5);                                // log2 last-pass blocks
  f[14] = (float)(single ? 0 : 13 - ell1);                            // log2 first-pass columns
  f[15] = (float)bm;
  return SBA_OK;
}

// ---------------------------------------------------------------------------
// the loaded model (one per process; parameters and scratch on the device
// current at load time)
// ---------------------------------------------------------------------------
#define CM_MAXCAND 64
struct CostModel {
  int loaded = 0;
  float fmean[CM_IN], fstd[CM_IN];
  float ymean = 0.f, ystd = 1.f;
  float *dparams = nullptr, *dx = nullptr, *dy = nullptr;
};
static CostModel g_cm;
static std::mutex g_cm_mu;

extern "C" int sba_cm_load(const float *params, const float *fmean, const float *fstd,
                           float ymean, float ystd) {
  if (!params || !fmean || !fstd || !(ystd > 0.f)) return SBA_BAND;
  std::lock_guard<std::mutex> lk(g_cm_mu);
  if (!g_cm.dparams) {
    CMCHK(hipMalloc(&g_cm.dparams, CM_NPARAM * sizeof(float)));
    CMCHK(hipMalloc(&g_cm.dx, CM_MAXCAND * CM_IN * sizeof(float)));
    CMCHK(hipMalloc(&g_cm.dy, CM_MAXCAND * sizeof(float)));
  }
  CMCHK(hipMemcpy(g_cm.dparams, params, CM_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  for (int i = 0; i < CM_IN; ++i) {
    g_cm.fmean[i] = fmean[i];
    g_cm.fstd[i] = fstd[i] > 0.f ? fstd[i] : 1.f;
  }
  g_cm.ymean = ymean;
  g_cm.ystd = ystd;
  g_cm.loaded = 1;
  return SBA_OK;
}

extern "C" int sba_cm_loaded(void) { return g_cm.loaded; }

// relative predicted cost of n (class, plan) rows given as raw features
// (n <= 64), evaluated by the forward kernel on `stream` (synchronizes)
static int cm_predict(const float *feat, int n, double *ms, hipStream_t st) {
  if (!g_cm.loaded) return SBA_ERR_WORKSPACE;
  if (n < 1 || n > CM_MAXCAND) return SBA_BAND;
  float xs[CM_MAXCAND * CM_IN], ys[CM_MAXCAND];
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < CM_IN; ++i)
      xs[r * CM_IN + i] = (feat[r * CM_IN + i] - g_cm.fmean[i]) / g_cm.fstd[i];
  CMCHK(hipMemcpyAsync(g_cm.dx, xs, (size_t)n * CM_IN * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_cm_forward, dim3(1), dim3(CM_NT), 0, st, g_cm.dparams, g_cm.dx, n, g_cm.dy,
                     (float *)nullptr);
  CMCHK(hipGetLastError());
  CMCHK(hipMemcpyAsync(ys, g_cm.dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st));
  CMCHK(hipStreamSynchronize(st));
  for (int r = 0; r < n; ++r) ms[r] = exp((double)ys[r] * g_cm.ystd + g_cm.ymean);
  return SBA_OK;
}

extern "C" int sba_cm_predict(int L, int sqr, int njobs, double fill, int last, int bm, int split,
                              void *stream, double *ms) {
  float f[CM_IN];
  int rc = sba_cm_features(L, sqr, njobs, fill, last, bm, split, f);
  if (rc != SBA_OK) return rc;
  std::lock_guard<std::mutex> lk(g_cm_mu);
  return cm_predict(f, 1, ms, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
extern "C" int sba_autotune(const sba_job *jobs, int njobs, int mode, int topk, int reps,
                            void *ws, size_t ws_bytes, void *stream, int *measured) {
  int dummy;
  if (!measured) measured = &dummy;
  *measured = 0;
  if (mode < SBA_TUNE_DEFAULT || mode > SBA_TUNE_MODEL || reps < 1) return SBA_BAND;
  const size_t need = sba_batch_ws_bytes(jobs, njobs);
  if (!need) return SBA_BAND;
  if (mode == SBA_TUNE_DEFAULT) return SBA_OK;
  if (!ws || ws_bytes < need) return SBA_ERR_WORKSPACE;
  if (mode == SBA_TUNE_MODEL && (!g_cm.loaded || topk < 1)) return SBA_ERR_WORKSPACE;
  const int nlg = sba_gpu_max_lg() + 1;
  // classes by (kind = square? | radix-3? << 1, L)
  std::vector<std::vector<sba_job>> cls(4 * (size_t)nlg);
  for (int i = 0; i < njobs; ++i) {
    int L, kind;
    if (sba_job_class(&jobs[i], &L, &kind) != SBA_OK) return SBA_BAND;
    cls[(size_t)kind * nlg + L].push_back(jobs[i]);
  }
  int plans[3 * CM_MAXCAND];
  for (int kind = 0; kind < 4; ++kind)
    for (int L = 0; L < nlg; ++L) {
      const std::vector<sba_job> &cj = cls[(size_t)kind * nlg + L];
      if (cj.empty()) continue;
      const int sq = kind & 1, r3 = (kind >> 1) & 1;
      // the cost model describes power-of-two classes; radix-3 classes keep
      // their plan in model mode and are timed in exhaustive mode
      if (r3 && mode == SBA_TUNE_MODEL) continue;
      const int n = (int)cj.size();
      double zs = 0;
      for (const sba_job &j : cj) zs += (double)j.an + (double)(j.bp ? j.bn : j.an);
      const double fill = zs / ((double)n * (double)((r3 ? 3u : 1u) << L));
      int nc = sba_plan_candidates(L, kind, plans, CM_MAXCAND);
      if (nc > CM_MAXCAND) nc = CM_MAXCAND;
      if (nc < 1) continue;
      std::vector<int> order(nc);
      for (int c = 0; c < nc; ++c) order[c] = c;
      int ntime = nc;
      if (mode == SBA_TUNE_MODEL) {
        float feat[CM_MAXCAND * CM_IN];
        double pred[CM_MAXCAND];
        for (int c = 0; c < nc; ++c)
          sba_cm_features(L, sq, n, fill, plans[3 * c], plans[3 * c + 1], plans[3 * c + 2],
                          feat + c * CM_IN);
        int rc;
        {
          std::lock_guard<std::mutex> lk(g_cm_mu);
          rc = cm_predict(feat, nc, pred, (hipStream_t)stream);
        }
        if (rc != SBA_OK) return rc;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return pred[a] < pred[b]; });
        ntime = topk < nc ? topk : nc;
      }
      int best = order[0];
      if (ntime > 1) {
        float tbest = 0.f;
        for (int r = 0; r < ntime; ++r) {
          const int c = order[r];
          float t = 0.f;
          const int rc = sba_class_time(cj.data(), n, plans[3 * c], plans[3 * c + 1],
                                        plans[3 * c + 2], reps, ws, ws_bytes, stream, &t);
          if (rc != SBA_OK) return rc;
          ++*measured;
          if (r == 0 || t < tbest) {
            tbest = t;
            best = c;
          }
        }
      }
      const int rc = sba_plan_set(L, kind, sba_plan_bucket(n), plans[3 * best],
                                  plans[3 * best + 1], plans[3 * best + 2]);
      if (rc != SBA_OK) return rc;
    }
  return SBA_OK;
}
