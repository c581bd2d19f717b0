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
// Every reduction is deterministic: parameter-gradient reductions that cross
// rows use fixed-order partial sums (split-K GEMM, column sums, the null
// token's gradients), and the scatter-add of neighbour gradients accumulates
// in 64-bit fixed point, whose integer sums do not depend on the order of
// the adds.  GEMMs run in FP32 or, for BF16 training, as v_mfma_f32_16x16x32_bf16
// tiles (operands rounded to bf16, FP32 accumulation; the FP32 master weights
// stay in the optimizer).
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

// Deterministic scatter-add: amax = max |dout| (float bits, atomicMax, which
// is order-independent); each term becomes llrint(dout * 2^sh) with
// 2^sh * amax < 2^(62 - h) and n < 2^h, so the 64-bit integer sum of up to n
// terms cannot overflow and is exact whatever order the adds land in; the
// float result is that sum times 2^-sh (one rounding).
__global__ void fill_kernel(uint32_t *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ void absmax_kernel(const float *__restrict__ x, int64_t n,
                              uint32_t *__restrict__ amax_bits) {
  uint32_t m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t b = __float_as_uint(x[i]) & 0x7FFFFFFFu;  // |x| (NaN stays above inf)
    m = b > m ? b : m;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const uint32_t t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicMax(amax_bits, m);
}

__device__ __forceinline__ int fixed_shift(uint32_t amax_bits, int headroom) {
  int e;
  (void)frexpf(__uint_as_float(amax_bits), &e);  // amax < 2^e
  return 62 - headroom - e;
}

__global__ void scatter_fixed_kernel(const float *__restrict__ dout, int64_t w,
                                     const int64_t *__restrict__ idx, int64_t n,
                                     const uint32_t *__restrict__ amax_bits, int headroom,
                                     unsigned long long *__restrict__ acc) {
  const uint32_t ab = *amax_bits;
  if (ab == 0u || ab >= 0x7F800000u) return;  // all zero, or non-finite (see finish)
  const double sc = ldexp(1.0, fixed_shift(ab, headroom));
  const int64_t total = n * w;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / w, c = i - r * w;
    const long long q = llrint((double)dout[i] * sc);
    if (q) atomicAdd(acc + idx[r] * w + c, (unsigned long long)q);
  }
}

__global__ void fixed_finish_kernel(const unsigned long long *__restrict__ acc, int64_t n,
                                    const uint32_t *__restrict__ amax_bits, int headroom,
                                    float *__restrict__ dsrc) {
  const uint32_t ab = *amax_bits;
  const double inv = ab && ab < 0x7F800000u ? ldexp(1.0, -fixed_shift(ab, headroom)) : 0.0;
  const float bad = ab >= 0x7F800000u ? __uint_as_float(0x7FC00000u) : 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dsrc[i] = ab >= 0x7F800000u ? bad : (float)((double)(long long)acc[i] * inv);
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

// BF16 MFMA GEMM with the FP32 path's interface and split-K: 64 x 64 tiles,
// K in steps of 32; the operands are rounded to bf16 (round to nearest even)
// while staged to LDS, each wave computes a 32 x 32 quarter as 2 x 2
// v_mfma_f32_16x16x32_bf16 tiles with FP32 accumulation.  Operand layout of
// the instruction: lane l holds row l % 16 of A and column l % 16 of B at
// k = 8 (l / 16) .. 8 (l / 16) + 7; result element j of lane l is
// (row 4 (l / 16) + j, column l % 16).
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;
constexpr int TBM = 64, TBN = 64, TBK = 32, TPAD = TBK + 8;

__device__ __forceinline__ uint16_t f2bf(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(int64_t M, int64_t N, int64_t K,
                                                        const float *__restrict__ A, int64_t sam,
                                                        int64_t sak, const float *__restrict__ B,
                                                        int64_t sbk, int64_t sbn,
                                                        float *__restrict__ C, int64_t kchunk,
                                                        int direct,
                                                        const float *__restrict__ bias,
                                                        int accumulate, int64_t ldc) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[TBM * TPAD];  // [m][k]
  __shared__ __attribute__((aligned(16))) uint16_t sB[TBN * TPAD];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const int64_t m0 = (int64_t)blockIdx.x * TBM, n0 = (int64_t)blockIdx.y * TBN;
  const int64_t kb = (int64_t)blockIdx.z * kchunk;
  const int64_t ke = kb + kchunk < K ? kb + kchunk : K;
  f32x4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, kq = (lane >> 4) * 8;
  for (int64_t k0 = kb; k0 < ke; k0 += TBK) {
    for (int e = tid; e < TBM * TBK; e += 256) {
      int kk, mm;
      if (sak == 1) {
        kk = e % TBK;
        mm = e / TBK;
      } else {
        mm = e % TBM;
        kk = e / TBM;
      }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      sA[mm * TPAD + kk] = f2bf((gm < M && gk < ke) ? A[gm * sam + gk * sak] : 0.0f);
    }
    for (int e = tid; e < TBN * TBK; e += 256) {
      int kk, nn;
      if (sbn == 1) {
        nn = e % TBN;
        kk = e / TBN;
      } else {
        kk = e % TBK;
        nn = e / TBK;
      }
      const int64_t gn = n0 + nn, gk = k0 + kk;
      sB[nn * TPAD + kk] = f2bf((gn < N && gk < ke) ? B[gk * sbk + gn * sbn] : 0.0f);
    }
    __syncthreads();
    bf16x8_t a[2], b[2];
    for (int i = 0; i < 2; i++) {
      a[i] = *(const bf16x8_t *)(sA + (wm + i * 16 + r) * TPAD + kq);
      b[i] = *(const bf16x8_t *)(sB + (wn + i * 16 + r) * TPAD + kq);
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int t = 0; t < 4; t++) {
        const int64_t gm = m0 + wm + i * 16 + (lane >> 4) * 4 + t;
        const int64_t gn = n0 + wn + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float v = acc[i][j][t];
        if (direct) {
          if (bias) v += bias[gn];
          float *o = C + gm * ldc + gn;
          *o = accumulate ? *o + v : v;
        } else {
          C[(int64_t)blockIdx.z * M * N + gm * N + gn] = v;
        }
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
    // this item's null-token gradients: row i of a (n_items x (d + 1)) array
    // summed in a fixed order afterwards
    if (lane == 0) dnull_s[i * (d + 1)] = wn * (gn - gbar);
    if (lane < d) dnull_s[i * (d + 1) + 1 + lane] = wn * go;
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
                            float *dsrc, int64_t n_src, uint64_t *acc, void *stream) {
  if (w <= 0 || n_src <= 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  // acc: n_src * w words of 64 bits, then one word holding amax
  const int64_t nacc = n_src * w;
  uint32_t *amax = (uint32_t *)(acc + nacc);
  int headroom = 1;
  while (headroom < 62 && ((int64_t)1 << headroom) <= n) headroom++;
  hipLaunchKernelGGL(fill_kernel, dim3(grid1(2 * nacc + 2, 256)), dim3(256), 0, s,
                     (uint32_t *)acc, 2 * nacc + 2);
  if (n > 0) {
    hipLaunchKernelGGL(absmax_kernel, dim3(grid1(n * w, 256)), dim3(256), 0, s, dout, n * w,
                       amax);
    hipLaunchKernelGGL(scatter_fixed_kernel, dim3(grid1(n * w, 256)), dim3(256), 0, s, dout, w,
                       idx, n, amax, headroom, (unsigned long long *)acc);
  }
  hipLaunchKernelGGL(fixed_finish_kernel, dim3(grid1(nacc, 256)), dim3(256), 0, s,
                     (const unsigned long long *)acc, nacc, amax, headroom, dsrc);
  return hip_ok(hipGetLastError(), "scatter_fixed_kernel");
}

int latec_train_write_rows(float *E, int64_t w, const int64_t *idx, int64_t n,
                           const float *vals, void *stream) {
  if (n <= 0 || w <= 0) return LATEC_OK;
  hipLaunchKernelGGL(write_rows_kernel, dim3(grid1(n * w, 256)), dim3(256), 0,
                     (hipStream_t)stream, E, w, idx, n, vals);
  return hip_ok(hipGetLastError(), "write_rows_kernel");
}

// C (M x N, row stride ldc) [+]= A * B (+ bias); work holds split partials.
static int train_gemm(bool bf16, int64_t M, int64_t N, int64_t K, const float *A, int64_t sam,
                      int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                      int64_t ldc, const float *bias, int accumulate, float *work,
                      int64_t work_floats, void *stream) {
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
  auto kern = bf16 ? gemm_bf16_kernel : gemm_kernel;
  if (splits == 1) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C,
                       kchunk, 1, bias, accumulate, ldc);
    return hip_ok(hipGetLastError(), "gemm_kernel");
  }
  hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk, sbn,
                     work, kchunk, 0, (const float *)nullptr, 0, (int64_t)0);
  int rc = hip_ok(hipGetLastError(), "gemm_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(gemm_finish_kernel, dim3(grid1(M * N, 256)), dim3(256), 0, s, M, N, splits,
                     work, bias, accumulate, C, ldc);
  return hip_ok(hipGetLastError(), "gemm_finish_kernel");
}

int latec_train_gemm(int64_t M, int64_t N, int64_t K, const float *A, int64_t sam, int64_t sak,
                     const float *B, int64_t sbk, int64_t sbn, float *C, int64_t ldc,
                     const float *bias, int accumulate, float *work, int64_t work_floats,
                     void *stream) {
  return train_gemm(false, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, accumulate, work,
                    work_floats, stream);
}

int latec_train_gemm_bf16(int64_t M, int64_t N, int64_t K, const float *A, int64_t sam,
                          int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                          int64_t ldc, const float *bias, int accumulate, float *work,
                          int64_t work_floats, void *stream) {
  return train_gemm(true, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, accumulate, work,
                    work_floats, stream);
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
                         const float *dout, float *ds, float *dv, float *dnull,
                         float *work, int64_t work_floats, void *stream) {
  if (n_items <= 0) return LATEC_OK;
  if (d > 64) {
    latec::set_error("attention pool supports d <= 64");
    return LATEC_E_ARG;
  }
  const int64_t items = n_items * (d + 1);
  if (work_floats < items + d + 1) {
    latec::set_error("latec_train_attn_bwd: workspace too small");
    return LATEC_E_CAPACITY;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(attn_bwd_kernel, dim3(grid1(n_items * 64, 256)), dim3(256), 0, s, n_items,
                     off, v, d, null_v, w_rows, w_null, dout, ds, dv, work, nullptr);
  int rc = hip_ok(hipGetLastError(), "attn_bwd_kernel");
  if (rc) return rc;
  // dnull[0] = the null score's gradient, dnull[1 + k] = the null value's
  return latec_train_colsum(n_items, d + 1, work, dnull, 0, work + items, work_floats - items,
                            stream);
}

int latec_train_nll(int64_t M, const float *o, const float *t, const float *eb, float mse_w,
                    float *bits, float *dO, float *mu, void *stream) {
  if (M <= 0) return LATEC_OK;
  hipLaunchKernelGGL(nll_kernel, dim3(grid1(M, 256)), dim3(256), 0, (hipStream_t)stream, M, o, t,
                     eb, mse_w, bits, dO, mu);
  return hip_ok(hipGetLastError(), "nll_kernel");
}

}  // extern "C"
