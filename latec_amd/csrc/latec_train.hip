// latec_train.hip -- forward and backward kernels of the GNN trainer (FP32).
//
// The trainer (scripts/train_gnn.py:373-459 run_stages, 195-230 the loss)
// composes the model from these ops, each a torch.autograd.Function whose
// forward and backward are the kernels below:
//   gather / scatter-add rows ...... neighbour reads from the embedding field
//   write rows ..................... the finalize write E[q] = mix(...)
//   GEMM (+bias, split-K) .......... every Linear, forward and both grads
//   column sums .................... bias gradients
//   GELU ........................... exact erf GELU and its derivative
//   segmented attention pool ....... AttnPool over a CSR row segment + null
//   Laplace head loss .............. sigmoid / clamp / discretized-Laplace
//                                    bits and their gradients
// Parameter-gradient reductions that cross rows use fixed-order partial sums
// (split-K GEMM, column sums); the scatter-add of neighbour gradients uses
// atomics (training only; inference never calls these kernels).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "latec_plan.h"

namespace latec {
namespace train {

static int hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_hip_error(what, (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

static int grid1(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g > (1 << 20)) g = 1 << 20;
  return (int)(g < 1 ? 1 : g);
}

__global__ void gather_rows_kernel(const float *__restrict__ src, int64_t w,
                                   const int64_t *__restrict__ idx, int64_t n,
                                   float *__restrict__ out) {
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / w, c = i - r * w;
    out[i] = src[idx[r] * w + c];
  }
}

__global__ void scatter_add_rows_kernel(const float *__restrict__ dout, int64_t w,
                                        const int64_t *__restrict__ idx, int64_t n,
                                        float *__restrict__ dsrc) {
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / w, c = i - r * w;
    atomicAdd(dsrc + idx[r] * w + c, dout[i]);
  }
}

__global__ void write_rows_kernel(float *__restrict__ E, int64_t w,
                                  const int64_t *__restrict__ idx, int64_t n,
                                  const float *__restrict__ vals) {
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / w, c = i - r * w;
    E[idx[r] * w + c] = vals ? vals[i] : 0.0f;
  }
}

// C[z] = A (M x K) * B (K x N) over the K range of split z, generic strides:
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn].
constexpr int BM = 64, BN = 64, BK = 16;
__global__ __launch_bounds__(256) void gemm_kernel(int64_t M, int64_t N, int64_t K,
                                                   const float *__restrict__ A, int64_t sam,
                                                   int64_t sak, const float *__restrict__ B,
                                                   int64_t sbk, int64_t sbn,
                                                   float *__restrict__ C, int64_t kchunk,
                                                   int direct, const float *__restrict__ bias,
                                                   int accumulate, int64_t ldc) {
  __shared__ float sA[BK][BM + 4];
  __shared__ float sB[BK][BN + 4];
  const int tid = threadIdx.x;
  const int tm = (tid / 16) * 4, tn = (tid % 16) * 4;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const int64_t kb = (int64_t)blockIdx.z * kchunk;
  const int64_t ke = kb + kchunk < K ? kb + kchunk : K;
  float acc[4][4] = {};
  for (int64_t k0 = kb; k0 < ke; k0 += BK) {
    for (int e = tid; e < BK * BM; e += 256) {
      int kk, mm;
      if (sak == 1) {  // k fastest: coalesced along k
        kk = e % BK;
        mm = e / BK;
      } else {
        mm = e % BM;
        kk = e / BM;
      }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      sA[kk][mm] = (gm < M && gk < ke) ? A[gm * sam + gk * sak] : 0.0f;
    }
    for (int e = tid; e < BK * BN; e += 256) {
      int kk, nn;
      if (sbn == 1) {
        nn = e % BN;
        kk = e / BN;
      } else {
        kk = e % BK;
        nn = e / BK;
      }
      const int64_t gn = n0 + nn, gk = k0 + kk;
      sB[kk][nn] = (gn < N && gk < ke) ? B[gk * sbk + gn * sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = sA[kk][tm + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = sB[kk][tn + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  if (direct) {  // one split: finish in place
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t gm = m0 + tm + i, gn = n0 + tn + j;
        if (gm < M && gn < N) {
          float v = acc[i][j];
          if (bias) v += bias[gn];
          float *o = C + gm * ldc + gn;
          *o = accumulate ? *o + v : v;
        }
      }
    return;
  }
  float *Cz = C + (int64_t)blockIdx.z * M * N;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t gm = m0 + tm + i, gn = n0 + tn + j;
      if (gm < M && gn < N) Cz[gm * N + gn] = acc[i][j];
    }
}

// out(m, n) = sum_z part[z](m, n) (+ bias[n]) (+ out(m, n) if accumulate),
// written with row stride ldc.
__global__ void gemm_finish_kernel(int64_t M, int64_t N, int splits, const float *__restrict__ part,
                                   const float *__restrict__ bias, int accumulate,
                                   float *__restrict__ out, int64_t ldc) {
  const int64_t total = M * N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    float v = 0.0f;
    for (int z = 0; z < splits; z++) v += part[(int64_t)z * total + i];
    const int64_t m = i / N, n = i - m * N;
    if (bias) v += bias[n];
    float *o = out + m * ldc + n;
    *o = accumulate ? *o + v : v;
  }
}

// column sums of X (M x N): fixed-order partials per block of rows
__global__ void colsum_part_kernel(int64_t M, int64_t N, const float *__restrict__ X,
                                   int64_t rows_per, float *__restrict__ part) {
  const int64_t r0 = (int64_t)blockIdx.x * rows_per;
  const int64_t r1 = r0 + rows_per < M ? r0 + rows_per : M;
  for (int64_t n = threadIdx.x; n < N; n += blockDim.x) {
    float v = 0.0f;
    for (int64_t r = r0; r < r1; r++) v += X[r * N + n];
    part[(int64_t)blockIdx.x * N + n] = v;
  }
}

__global__ void colsum_finish_kernel(int64_t N, int parts, const float *__restrict__ part,
                                     int accumulate, float *__restrict__ out) {
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N;
       n += (int64_t)gridDim.x * blockDim.x) {
    float v = 0.0f;
    for (int p = 0; p < parts; p++) v += part[(int64_t)p * N + n];
    out[n] = accumulate ? out[n] + v : v;
  }
}

__device__ __forceinline__ float gelu_f(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

__global__ void gelu_fwd_kernel(const float *__restrict__ x, float *__restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = gelu_f(x[i]);
}
__global__ void gelu_bwd_kernel(const float *__restrict__ x, const float *__restrict__ dy,
                                float *__restrict__ dx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dx[i] = dy[i] * gelu_d(x[i]);
}

// Segmented attention pool (AttnPool, gnn_predictor.py:720-747) over rows
// [off[i], off[i+1]) plus a null token: out_i = sum_r w_r v_r + w_null null_v.
// One wave per item; lanes over the d features (d <= 64).
__global__ __launch_bounds__(256) void attn_fwd_kernel(int64_t n_items,
                                                       const int64_t *__restrict__ off,
                                                       const float *__restrict__ s,
                                                       const float *__restrict__ v, int d,
                                                       float null_s,
                                                       const float *__restrict__ null_v,
                                                       float *__restrict__ out,
                                                       float *__restrict__ w_rows,
                                                       float *__restrict__ w_null) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t i = wave; i < n_items; i += nw) {
    const int64_t a = off[i], b = off[i + 1];
    float mx = null_s;
    for (int64_t r = a; r < b; r++) mx = fmaxf(mx, s[r]);
    float den = expf(null_s - mx);
    for (int64_t r = a; r < b; r++) den += expf(s[r] - mx);
    const float inv = 1.0f / den;
    const float wn = expf(null_s - mx) * inv;
    float acc = 0.0f;
    for (int64_t r = a; r < b; r++) {
      const float wr = expf(s[r] - mx) * inv;
      if (lane == 0) w_rows[r] = wr;
      if (lane < d) acc += wr * v[r * d + lane];
    }
    if (lane < d) out[i * d + lane] = acc + wn * null_v[lane];
    if (lane == 0) w_null[i] = wn;
  }
}

// Backward: g_r = dout_i . v_r, g_null = dout_i . null_v, gbar = sum w g;
// ds_r = w_r (g_r - gbar), dv_r = w_r dout_i; null grads accumulated.
__global__ __launch_bounds__(256) void attn_bwd_kernel(
    int64_t n_items, const int64_t *__restrict__ off, const float *__restrict__ v, int d,
    const float *__restrict__ null_v, const float *__restrict__ w_rows,
    const float *__restrict__ w_null, const float *__restrict__ dout, float *__restrict__ ds,
    float *__restrict__ dv, float *__restrict__ dnull_s, float *__restrict__ dnull_v) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t i = wave; i < n_items; i += nw) {
    const int64_t a = off[i], b = off[i + 1];
    const float go = lane < d ? dout[i * d + lane] : 0.0f;
    float gn = lane < d ? go * null_v[lane] : 0.0f;
    for (int o = 32; o >= 1; o >>= 1) gn += __shfl_xor(gn, o);
    const float wn = w_null[i];
    float gbar = wn * gn;
    for (int64_t r = a; r < b; r++) {
      float g = lane < d ? go * v[r * d + lane] : 0.0f;
      for (int o = 32; o >= 1; o >>= 1) g += __shfl_xor(g, o);
      gbar += w_rows[r] * g;
    }
    for (int64_t r = a; r < b; r++) {
      const float wr = w_rows[r];
      float g = lane < d ? go * v[r * d + lane] : 0.0f;
      for (int o = 32; o >= 1; o >>= 1) g += __shfl_xor(g, o);
      if (lane == 0) ds[r] = wr * (g - gbar);
      if (lane < d) dv[r * d + lane] = wr * go;
    }
    if (lane == 0) atomicAdd(dnull_s, wn * (gn - gbar));
    if (lane < d) atomicAdd(dnull_v + lane, wn * go);
  }
}

// Head outputs o = (o0, o1) -> mu = sigmoid(o0), log_b = log_eb + clamp(o1,
// -4, 12); bits = discretized Laplace code length of the target
// (train_gnn.py:195-230) + mse_w * (mu - t)^2; writes bits and d bits / d o.
__global__ void nll_kernel(int64_t M, const float *__restrict__ o, const float *__restrict__ t,
                           const float *__restrict__ ebv, float mse_w, float *__restrict__ bits,
                           float *__restrict__ dO, float *__restrict__ mu_out) {
  const float LN2 = 0.69314718055994531f, INV_LN2 = 1.4426950408889634f;
  const float TMAX = 32.0f * LN2, EPS = 1.1920928955078125e-07f, TINY = 1.17549435e-38f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float o0 = o[2 * i], o1 = o[2 * i + 1];
    const float eb = ebv[i];
    const float mu = 1.0f / (1.0f + expf(-o0));
    const float log_eb = log2f(fmaxf(eb, TINY));
    const float o1c = fminf(fmaxf(o1, -4.0f), 12.0f);
    const float log_b = log_eb + o1c;
    const float b = fmaxf(exp2f(log_b), TINY);
    const float r = t[i] - mu, a = fabsf(r);
    const float rho = a / b, e = eb / b;
    float logp, dl_drho, dl_de;
    if (a >= eb) {
      const float q = fminf(-2.0f * e, -EPS);
      const float em = expf(q);
      const float tail = log1pf(-em);
      float tt = rho - e, c = 1.0f;
      if (tt > TMAX) {
        c = 1.0f / (1.0f + (tt - TMAX));
        tt = TMAX + log1pf(tt - TMAX);
      }
      logp = -LN2 - tt + tail;
      dl_drho = -c;
      const float dtail = (-2.0f * e < -EPS) ? 2.0f * em / (1.0f - em) : 0.0f;
      dl_de = c + dtail;
    } else {
      const float rc = fminf(rho, e);
      const float e1 = expf(-(e - rc)), e2 = expf(-(e + rc));
      const float f = 1.0f - 0.5f * e1 - 0.5f * e2;
      if (f > 1e-30f) {
        logp = logf(f);
        dl_drho = (-0.5f * e1 + 0.5f * e2) / f;
        dl_de = (0.5f * e1 + 0.5f * e2) / f;
      } else {
        logp = logf(1e-30f);
        dl_drho = dl_de = 0.0f;
      }
    }
    const float sgn = r > 0.f ? 1.0f : (r < 0.f ? -1.0f : 0.0f);
    // bits = -logp / ln2 ; rho = |t - mu| / b ; e = eb / b ; b = 2^log_b
    const float dbits_dmu = -INV_LN2 * dl_drho * (-sgn / b);
    const float dbits_dlogb = -INV_LN2 * (dl_drho * (-rho * LN2) + dl_de * (-e * LN2));
    float bb = -logp * INV_LN2;
    float dmu = dbits_dmu;
    if (mse_w != 0.0f) {
      bb += mse_w * r * r;
      dmu += mse_w * (-2.0f * r);
    }
    bits[i] = bb;
    dO[2 * i] = dmu * mu * (1.0f - mu);
    dO[2 * i + 1] = (o1 > -4.0f && o1 < 12.0f) ? dbits_dlogb : 0.0f;
    if (mu_out) mu_out[i] = mu;
  }
}

}  // namespace train
}  // namespace latec

using namespace latec::train;

extern "C" {

int latec_train_gather(const float *src, int64_t w, const int64_t *idx, int64_t n, float *out,
                       void *stream) {
  if (n <= 0 || w <= 0) return LATEC_OK;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(grid1(n * w, 256)), dim3(256), 0,
                     (hipStream_t)stream, src, w, idx, n, out);
  return hip_ok(hipGetLastError(), "gather_rows_kernel");
}

int latec_train_scatter_add(const float *dout, int64_t w, const int64_t *idx, int64_t n,
                            float *dsrc, void *stream) {
  if (n <= 0 || w <= 0) return LATEC_OK;
  hipLaunchKernelGGL(scatter_add_rows_kernel, dim3(grid1(n * w, 256)), dim3(256), 0,
                     (hipStream_t)stream, dout, w, idx, n, dsrc);
  return hip_ok(hipGetLastError(), "scatter_add_rows_kernel");
}

int latec_train_write_rows(float *E, int64_t w, const int64_t *idx, int64_t n,
                           const float *vals, void *stream) {
  if (n <= 0 || w <= 0) return LATEC_OK;
  hipLaunchKernelGGL(write_rows_kernel, dim3(grid1(n * w, 256)), dim3(256), 0,
                     (hipStream_t)stream, E, w, idx, n, vals);
  return hip_ok(hipGetLastError(), "write_rows_kernel");
}

// C (M x N, row stride ldc) [+]= A * B (+ bias); work holds split partials.
int latec_train_gemm(int64_t M, int64_t N, int64_t K, const float *A, int64_t sam, int64_t sak,
                     const float *B, int64_t sbk, int64_t sbn, float *C, int64_t ldc,
                     const float *bias, int accumulate, float *work, int64_t work_floats,
                     void *stream) {
  if (M <= 0 || N <= 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int splits = 1;
  if (K > 4096 && tiles < 512) {
    int64_t want = (512 + tiles - 1) / tiles;
    const int64_t maxs = (K + 4095) / 4096;
    if (want > maxs) want = maxs;
    if (want > 256) want = 256;
    splits = (int)want;
  }
  while (splits > 1 && (int64_t)splits * M * N > work_floats) splits--;
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  if (kchunk < BK) kchunk = BK;
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN), (unsigned)splits);
  if (splits == 1) {
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C,
                       kchunk, 1, bias, accumulate, ldc);
    return hip_ok(hipGetLastError(), "gemm_kernel");
  }
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                     work, kchunk, 0, (const float *)nullptr, 0, (int64_t)0);
  int rc = hip_ok(hipGetLastError(), "gemm_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(gemm_finish_kernel, dim3(grid1(M * N, 256)), dim3(256), 0, s, M, N, splits,
                     work, bias, accumulate, C, ldc);
  return hip_ok(hipGetLastError(), "gemm_finish_kernel");
}

int latec_train_colsum(int64_t M, int64_t N, const float *X, float *out, int accumulate,
                       float *work, int64_t work_floats, void *stream) {
  if (N <= 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  int parts = (int)((M + 4095) / 4096);
  if (parts < 1) parts = 1;
  if (parts > 1024) parts = 1024;
  while (parts > 1 && (int64_t)parts * N > work_floats) parts--;
  if ((int64_t)parts * N > work_floats) {
    latec::set_error("latec_train_colsum: workspace too small");
    return LATEC_E_CAPACITY;
  }
  const int64_t rows_per = (M + parts - 1) / parts;
  hipLaunchKernelGGL(colsum_part_kernel, dim3(parts), dim3(256), 0, s, M, N, X,
                     rows_per < 1 ? 1 : rows_per, work);
  int rc = hip_ok(hipGetLastError(), "colsum_part_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(colsum_finish_kernel, dim3(grid1(N, 256)), dim3(256), 0, s, N, parts, work,
                     accumulate, out);
  return hip_ok(hipGetLastError(), "colsum_finish_kernel");
}

int latec_train_gelu(const float *x, float *y, int64_t n, void *stream) {
  if (n <= 0) return LATEC_OK;
  hipLaunchKernelGGL(gelu_fwd_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream, x,
                     y, n);
  return hip_ok(hipGetLastError(), "gelu_fwd_kernel");
}

int latec_train_gelu_bwd(const float *x, const float *dy, float *dx, int64_t n, void *stream) {
  if (n <= 0) return LATEC_OK;
  hipLaunchKernelGGL(gelu_bwd_kernel, dim3(grid1(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     x, dy, dx, n);
  return hip_ok(hipGetLastError(), "gelu_bwd_kernel");
}

int latec_train_attn(int64_t n_items, const int64_t *off, const float *s, const float *v, int d,
                     float null_s, const float *null_v, float *out, float *w_rows, float *w_null,
                     void *stream) {
  if (n_items <= 0) return LATEC_OK;
  if (d > 64) {
    latec::set_error("attention pool supports d <= 64");
    return LATEC_E_ARG;
  }
  hipLaunchKernelGGL(attn_fwd_kernel, dim3(grid1(n_items * 64, 256)), dim3(256), 0,
                     (hipStream_t)stream, n_items, off, s, v, d, null_s, null_v, out, w_rows,
                     w_null);
  return hip_ok(hipGetLastError(), "attn_fwd_kernel");
}

int latec_train_attn_bwd(int64_t n_items, const int64_t *off, const float *v, int d,
                         const float *null_v, const float *w_rows, const float *w_null,
                         const float *dout, float *ds, float *dv, float *dnull_s,
                         float *dnull_v, void *stream) {
  if (n_items <= 0) return LATEC_OK;
  if (d > 64) {
    latec::set_error("attention pool supports d <= 64");
    return LATEC_E_ARG;
  }
  hipLaunchKernelGGL(attn_bwd_kernel, dim3(grid1(n_items * 64, 256)), dim3(256), 0,
                     (hipStream_t)stream, n_items, off, v, d, null_v, w_rows, w_null, dout, ds,
                     dv, dnull_s, dnull_v);
  return hip_ok(hipGetLastError(), "attn_bwd_kernel");
}

int latec_train_nll(int64_t M, const float *o, const float *t, const float *eb, float mse_w,
                    float *bits, float *dO, float *mu, void *stream) {
  if (M <= 0) return LATEC_OK;
  hipLaunchKernelGGL(nll_kernel, dim3(grid1(M, 256)), dim3(256), 0, (hipStream_t)stream, M, o, t,
                     eb, mse_w, bits, dO, mu);
  return hip_ok(hipGetLastError(), "nll_kernel");
}

}  // extern "C"
