// latec_tables.hip -- Laplace-mixture tables and the per-stage level fit on
// the GPU (gfx950).
//
// Tables follow latec/rans.py bit for bit:
//   _level_weights :247-267  -- Laplace bin masses on a 64-level log grid,
//                               marker weight max*1e-9, rows normalized by
//                               NumPy's pairwise sum
//   _shaped_weights :377-384, _with_outlier_weight :172-180
//   _neg_log2_weights :387-392
//   _quantize_pmf_rows :676-722 -- floor at 1, surplus to the largest
//                               fractional parts / deficit from the smallest,
//                               ties by index (NumPy's stable argsort)
// exp is latec_exp (ddmath.h), which reproduces the host libm bit for bit;
// the stable-sort tie-breaking is evaluated as an exact threshold search over
// the IEEE bit patterns of the fractional parts, which orders them exactly
// as the sort does.
//
// Fit: choose_kexp :70-80 and choose_const_level :395-430.  The histogram is
// rebuilt with the reference's float32 per-2^22-tile accumulation; each
// level's cost is summed in float64 in a fixed order (ascending symbols in
// segments of 256), then rounded to float32 and compared as the reference
// compares its float32 sgemv results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "ddmath.h"
#include "latec_plan.h"

namespace latec {

static int hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_hip_error(what, (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

#define NLEV 64
static const int MIX_D[4] = {4, 8, 16, 24};
static const double MIX_W[5] = {0.02, 0.05, 0.12, 0.25, 0.45};
static const double OUTLIER_W[8] = {0.0,     0x1p-13, 0x1p-10, 0x1p-8,
                                    0x1p-6,  0x1p-4,  0x1p-2,  0.5};

static void shape_params(int shape, int *d, double *w) {
  if (shape <= 0) {
    *d = 0;
    *w = 0.0;
  } else {
    *d = MIX_D[(shape - 1) / 5];
    *w = MIX_W[(shape - 1) % 5];
  }
}

int shape_ids(int kexp, int *ids) {  // rans.py:195-201
  if (kexp < 13) {
    for (int i = 0; i < 21; i++) ids[i] = i;
    return 21;
  }
  int n = 0;
  ids[n++] = 0;
  for (int i = 1; i < 21; i++)
    if (MIX_W[(i - 1) % 5] == 0.02) ids[n++] = i;
  return n;
}

// NumPy's pairwise summation of a contiguous float64 run (add.reduce).
__device__ double pairwise_sum_dev(const double *a, int64_t n) {
  // explicit stack of (ptr, n) nodes: depth <= 2*log2(n)
  const double *sp[64];
  int64_t sn[64];
  double sv[64];
  int state[64];
  int top = 0;
  sp[0] = a; sn[0] = n; state[0] = 0;
  double ret = 0.0;
  while (top >= 0) {
    const double *p = sp[top];
    int64_t m = sn[top];
    if (m <= 128) {
      double r;
      if (m < 8) {
        r = 0.0;
        for (int64_t i = 0; i < m; i++) r = i == 0 ? p[0] : r + p[i];
      } else {
        double acc[8];
        for (int j = 0; j < 8; j++) acc[j] = p[j];
        int64_t i;
        for (i = 8; i < m - (m % 8); i += 8)
          for (int j = 0; j < 8; j++) acc[j] += p[i + j];
        r = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
        for (; i < m; i++) r += p[i];
      }
      top--;
      // deliver r to the parent
      while (true) {
        if (top < 0) { ret = r; break; }
        if (state[top] == 1) {  // left done: store, descend right
          sv[top] = r;
          state[top] = 2;
          int64_t n2 = sn[top] / 2;
          n2 -= n2 % 8;
          top++;
          sp[top] = sp[top - 1] + n2; sn[top] = sn[top - 1] - n2; state[top] = 0;
          break;
        }
        // state 2: both done
        r = sv[top] + r;
        top--;
      }
      continue;
    }
    // internal node: descend left
    state[top] = 1;
    int64_t n2 = m / 2;
    n2 -= n2 % 8;
    top++;
    sp[top] = p; sn[top] = n2; state[top] = 0;
  }
  return ret;
}

// NumPy's add.reduce along a row of a 2-D array: 8192-element iterator
// buffers, each summed pairwise, accumulated in order.
__device__ double row_sum_dev(const double *a, int64_t n) {
  double r = 0.0;
  for (int64_t i = 0; i < n; i += 8192) {
    const int64_t m = n - i < 8192 ? n - i : 8192;
    const double c = pairwise_sum_dev(a + i, m);
    r = i == 0 ? c : r + c;
  }
  return r;
}

__device__ __forceinline__ double laplace_cdf(double x, double b) {
  double z = x / b;
  z = z < -700.0 ? -700.0 : (z > 700.0 ? 700.0 : z);
  return x < 0 ? 0.5 * latec_exp(z) : 1.0 - 0.5 * latec_exp(-z);
}

// One block per level row.
__global__ void __launch_bounds__(256) level_weights_kernel(int kexp, double lstart,
                                                            double lstop, double *out) {
  const int lvl = blockIdx.x;
  const int64_t span = ((int64_t)1 << kexp) - 1, A = 2 * span + 2;
  double *row = out + lvl * A;
  const double step = (lstop - lstart) / 63.0;
  double y = (double)lvl * step;
  y = y + lstart;
  if (lvl == 63) y = lstop;
  const double b = latec_exp(y);
  double mx = -INFINITY;
  for (int64_t k = threadIdx.x; k < 2 * span + 1; k += blockDim.x) {
    const double q = (double)(k - span);
    const double v = laplace_cdf(q * 2.0 + 1.0, b) - laplace_cdf(q * 2.0 - 1.0, b);
    row[k + 1] = v;
    mx = v > mx ? v : mx;
  }
  __shared__ double red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  __shared__ double rs;
  if (threadIdx.x == 0) {
    row[0] = red[0] * 1e-9;
    rs = row_sum_dev(row, A);
  }
  __syncthreads();
  for (int64_t k = threadIdx.x; k < A; k += blockDim.x) row[k] = row[k] / rs;
}

__global__ void __launch_bounds__(256) model_weights_kernel(int kexp, int d, double w,
                                                            double ow, const double *base,
                                                            double *out) {
  const int lvl = blockIdx.x;
  const int64_t A = (int64_t)2 << kexp;
  const int j = lvl + d < NLEV - 1 ? lvl + d : NLEV - 1;
  double *row = out + lvl * A;
  for (int64_t k = threadIdx.x; k < A; k += blockDim.x)
    row[k] = w == 0.0 ? base[lvl * A + k] : (1.0 - w) * base[lvl * A + k] + w * base[j * A + k];
  if (ow == 0.0) return;
  __syncthreads();
  __shared__ double f;
  if (threadIdx.x == 0) f = (1.0 - ow) / row_sum_dev(row + 1, A - 1);
  __syncthreads();
  for (int64_t k = 1 + threadIdx.x; k < A; k += blockDim.x) row[k] = row[k] * f;
  if (threadIdx.x == 0) row[0] = ow;
}

__global__ void nl2_kernel(const double *w, int64_t n, double floor_, float *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = w[i] > floor_ ? w[i] : floor_;
    out[i] = (float)(-cr_log2(v));
  }
}

// ---- PMF quantization of one row per block (1024 threads) ----------------

template <typename T>
__device__ T block_sum(T v, T *scratch) {
  const int t = threadIdx.x;
  scratch[t] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (t < o) scratch[t] += scratch[t + o];
    __syncthreads();
  }
  T r = scratch[0];
  __syncthreads();
  return r;
}

// exclusive scan of per-thread values (block order = index order)
__device__ int64_t block_excl_scan(int64_t v, int64_t *scratch) {
  const int t = threadIdx.x;
  scratch[t] = v;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int64_t add = t >= o ? scratch[t - o] : 0;
    __syncthreads();
    scratch[t] += add;
    __syncthreads();
  }
  int64_t incl = scratch[t];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(1024) pmf_rows_kernel(const double *w, int64_t A,
                                                        int64_t total, int l0,
                                                        int64_t *pmf_s, double *frac_s,
                                                        uint32_t *out, int *err) {
  __shared__ int64_t si[1024];
  __shared__ double rs_sh;
  const int row = l0 + blockIdx.x;
  const double *wr = w + row * A;
  int64_t *pmf = pmf_s + blockIdx.x * A;
  double *frac = frac_s + blockIdx.x * A;
  if (threadIdx.x == 0) rs_sh = row_sum_dev(wr, A);
  __syncthreads();
  const double rs = rs_sh;
  // contiguous ownership so that thread order == index order
  const int64_t per = (A + blockDim.x - 1) / blockDim.x;
  const int64_t k0 = threadIdx.x * per, k1 = k0 + per < A ? k0 + per : A;
  int64_t part = 0;
  for (int64_t k = k0; k < k1; k++) {
    const double sc = wr[k] / rs * (double)total;
    const double fl = floor(sc);
    const int64_t f = (int64_t)fl;
    pmf[k] = f > 1 ? f : 1;
    frac[k] = sc - fl;
    part += pmf[k];
  }
  const int64_t diff = total - block_sum<int64_t>(part, si);
  if (diff > 0) {
    const int64_t bulk = diff / A, rem = diff % A;
    // threshold T: largest bit pattern with count(frac >= T) >= rem
    uint64_t T = 0;
    if (rem > 0) {
      for (int bit = 63; bit >= 0; bit--) {
        const uint64_t cand = T | (1ull << bit);
        int64_t c = 0;
        for (int64_t k = k0; k < k1; k++) c += (uint64_t)__double_as_longlong(frac[k]) >= cand;
        if (block_sum<int64_t>(c, si) >= rem) T = cand;
      }
    }
    int64_t gt = 0, eq = 0;
    for (int64_t k = k0; k < k1; k++) {
      const uint64_t fb = (uint64_t)__double_as_longlong(frac[k]);
      gt += rem > 0 && fb > T;
      eq += rem > 0 && fb == T;
    }
    const int64_t n_gt = block_sum<int64_t>(gt, si);
    int64_t before = block_excl_scan(eq, si);
    const int64_t take_eq = rem - n_gt;
    for (int64_t k = k0; k < k1; k++) {
      const uint64_t fb = (uint64_t)__double_as_longlong(frac[k]);
      int64_t add = bulk;
      if (rem > 0 && fb > T) add += 1;
      if (rem > 0 && fb == T) {
        if (before < take_eq) add += 1;
        before++;
      }
      pmf[k] += add;
    }
  } else if (diff < 0) {
    const int64_t need = -diff;
    int64_t tot_av = 0;
    for (int64_t k = k0; k < k1; k++) tot_av += pmf[k] - 1;
    if (block_sum<int64_t>(tot_av, si) < need) {
      if (threadIdx.x == 0) *err = 1;
      return;
    }
    // smallest T with sum(avail over frac <= T) >= need
    uint64_t T = 0;
    for (int bit = 63; bit >= 0; bit--) {
      // test: sum over frac <= (T | ((1<<bit)-1)) (prefix with this bit 0)
      const uint64_t cand = T | ((bit ? (1ull << bit) : 1ull) - 1);
      int64_t c = 0;
      for (int64_t k = k0; k < k1; k++)
        if ((uint64_t)__double_as_longlong(frac[k]) <= cand) c += pmf[k] - 1;
      if (block_sum<int64_t>(c, si) < need) T |= (1ull << bit);
    }
    int64_t lt = 0, eqv = 0;
    for (int64_t k = k0; k < k1; k++) {
      const uint64_t fb = (uint64_t)__double_as_longlong(frac[k]);
      if (fb < T) lt += pmf[k] - 1;
      if (fb == T) eqv += pmf[k] - 1;
    }
    const int64_t sum_lt = block_sum<int64_t>(lt, si);
    int64_t before = block_excl_scan(eqv, si);
    const int64_t remaining = need - sum_lt;
    for (int64_t k = k0; k < k1; k++) {
      const uint64_t fb = (uint64_t)__double_as_longlong(frac[k]);
      const int64_t av = pmf[k] - 1;
      if (fb < T) {
        pmf[k] -= av;
      } else if (fb == T) {
        int64_t take = remaining - before;
        take = take < 0 ? 0 : (take > av ? av : take);
        pmf[k] -= take;
        before += av;
      }
    }
  }
  __syncthreads();
  // CDF
  int64_t loc = 0;
  for (int64_t k = k0; k < k1; k++) loc += pmf[k];
  int64_t c = block_excl_scan(loc, si);
  uint32_t *orow = out + (int64_t)blockIdx.x * (A + 1);
  if (threadIdx.x == 0) orow[0] = 0;
  for (int64_t k = k0; k < k1; k++) {
    c += pmf[k];
    orow[k + 1] = (uint32_t)c;
  }
}

// ---------------------------------------------------------------------------
// host side: cached device tables
// ---------------------------------------------------------------------------

struct TableCache {
  std::mutex mu;
  int device = -1;
  std::map<int, double *> level_w;                       // kexp
  std::map<std::tuple<int, int, int>, double *> model_w;  // kexp, shape, owx
  std::map<std::tuple<int, int, int, int>, float *> nl2;  // kexp, shape, owx, prec
  size_t bytes = 0;
};
static TableCache g_cache;

static int ensure_level_weights(int kexp, hipStream_t st, double **out) {
  auto it = g_cache.level_w.find(kexp);
  if (it != g_cache.level_w.end()) {
    *out = it->second;
    return LATEC_OK;
  }
  const int64_t A = (int64_t)2 << kexp;
  double *d;
  int rc = hip_ok(hipMalloc(&d, sizeof(double) * NLEV * A), "hipMalloc");
  if (rc) return rc;
  hipLaunchKernelGGL(level_weights_kernel, dim3(NLEV), dim3(256), 0, st, kexp,
                     std::log(1.0 / 16.0), std::log(4096.0), d);
  if ((rc = hip_ok(hipGetLastError(), "level_weights_kernel"))) return rc;
  g_cache.level_w[kexp] = d;
  g_cache.bytes += sizeof(double) * NLEV * A;
  *out = d;
  return LATEC_OK;
}

static int ensure_model_weights(int kexp, int shape, int owx, hipStream_t st,
                                double **out) {
  auto key = std::make_tuple(kexp, shape, owx);
  auto it = g_cache.model_w.find(key);
  if (it != g_cache.model_w.end()) {
    *out = it->second;
    return LATEC_OK;
  }
  double *base;
  int rc = ensure_level_weights(kexp, st, &base);
  if (rc) return rc;
  const int64_t A = (int64_t)2 << kexp;
  double *d;
  if ((rc = hip_ok(hipMalloc(&d, sizeof(double) * NLEV * A), "hipMalloc"))) return rc;
  int dd;
  double w;
  shape_params(shape, &dd, &w);
  hipLaunchKernelGGL(model_weights_kernel, dim3(NLEV), dim3(256), 0, st, kexp, dd, w,
                     OUTLIER_W[owx], base, d);
  if ((rc = hip_ok(hipGetLastError(), "model_weights_kernel"))) return rc;
  g_cache.model_w[key] = d;
  g_cache.bytes += sizeof(double) * NLEV * A;
  *out = d;
  return LATEC_OK;
}

static int ensure_nl2(int kexp, int shape, int owx, int precision, hipStream_t st,
                      float **out) {
  auto key = std::make_tuple(kexp, shape, owx, precision);
  auto it = g_cache.nl2.find(key);
  if (it != g_cache.nl2.end()) {
    *out = it->second;
    return LATEC_OK;
  }
  double *mw;
  int rc = ensure_model_weights(kexp, shape, owx, st, &mw);
  if (rc) return rc;
  const int64_t n = NLEV * ((int64_t)2 << kexp);
  float *d;
  if ((rc = hip_ok(hipMalloc(&d, sizeof(float) * n), "hipMalloc"))) return rc;
  hipLaunchKernelGGL(nl2_kernel, dim3((unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)),
                     dim3(256), 0, st, mw, n, 1.0 / (double)((int64_t)1 << precision), d);
  if ((rc = hip_ok(hipGetLastError(), "nl2_kernel"))) return rc;
  g_cache.nl2[key] = d;
  g_cache.bytes += sizeof(float) * n;
  *out = d;
  return LATEC_OK;
}

static int check_table_args(int kexp, int shape, int owx, int precision) {
  if (kexp < 1 || kexp > 15 || shape < 0 || shape > 20 || owx < 0 || owx > 7 ||
      precision < 2 || precision > 30) {
    set_error("table parameters out of range");
    return LATEC_E_ARG;
  }
  if (((int64_t)2 << kexp) > ((int64_t)1 << precision)) {
    set_error("precision is too small for the quantizer alphabet");
    return LATEC_E_ARG;
  }
  return LATEC_OK;
}

int level_weights(int kexp, double *d_out, hipStream_t st) {
  if (kexp < 1 || kexp > 15) {
    set_error("kexp out of range");
    return LATEC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_cache.mu);
  double *d;
  int rc = ensure_level_weights(kexp, st, &d);
  if (rc) return rc;
  return hip_ok(hipMemcpyAsync(d_out, d, sizeof(double) * NLEV * ((int64_t)2 << kexp),
                               hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
}

int model_weights(int kexp, int shape, int owx, double *d_out, hipStream_t st) {
  int rc = check_table_args(kexp, shape, owx, 24);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_cache.mu);
  double *d;
  if ((rc = ensure_model_weights(kexp, shape, owx, st, &d))) return rc;
  return hip_ok(hipMemcpyAsync(d_out, d, sizeof(double) * NLEV * ((int64_t)2 << kexp),
                               hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
}

int neg_log2_weights(int kexp, int shape, int owx, int precision, float *d_out,
                     hipStream_t st) {
  int rc = check_table_args(kexp, shape, owx, precision);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_cache.mu);
  float *d;
  if ((rc = ensure_nl2(kexp, shape, owx, precision, st, &d))) return rc;
  return hip_ok(hipMemcpyAsync(d_out, d, sizeof(float) * NLEV * ((int64_t)2 << kexp),
                               hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
}

int build_cdfs(int kexp, int shape, int precision, int owx, int l0, int l1,
               uint32_t *d_out, hipStream_t st) {
  int rc = check_table_args(kexp, shape, owx, precision);
  if (rc) return rc;
  if (l0 < 0 || l1 > NLEV || l0 >= l1) {
    set_error("level range out of bounds");
    return LATEC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_cache.mu);
  double *mw;
  if ((rc = ensure_model_weights(kexp, shape, owx, st, &mw))) return rc;
  const int64_t A = (int64_t)2 << kexp, rows = l1 - l0;
  int64_t *pmf;
  double *frac;
  int *err;
  if ((rc = hip_ok(hipMallocAsync((void **)&pmf, sizeof(int64_t) * rows * A, st), "hipMallocAsync")))
    return rc;
  if ((rc = hip_ok(hipMallocAsync((void **)&frac, sizeof(double) * rows * A, st), "hipMallocAsync")))
    return rc;
  if ((rc = hip_ok(hipMallocAsync((void **)&err, sizeof(int), st), "hipMallocAsync"))) return rc;
  (void)hipMemsetAsync(err, 0, sizeof(int), st);
  hipLaunchKernelGGL(pmf_rows_kernel, dim3((unsigned)rows), dim3(1024), 0, st, mw, A,
                     (int64_t)1 << precision, l0, pmf, frac, d_out, err);
  rc = hip_ok(hipGetLastError(), "pmf_rows_kernel");
  int herr = 0;
  if (!rc) rc = hip_ok(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st), "copy");
  if (!rc) rc = hip_ok(hipStreamSynchronize(st), "sync");
  (void)hipFreeAsync(pmf, st);
  (void)hipFreeAsync(frac, st);
  (void)hipFreeAsync(err, st);
  if (!rc && herr) {
    set_error("could not normalize PMF at requested precision");
    return LATEC_E_ARG;
  }
  return rc;
}

// ---------------------------------------------------------------------------
// Fit
// ---------------------------------------------------------------------------

struct StageScan {  // per stage, from the full-alphabet histogram
  int32_t kexp;
  uint32_t hist0;
  uint32_t maxbin;    // largest bin count
  uint32_t pad;
};

int64_t hist_words(uint32_t R);

// Fold the striped window copies into the main bins (one block per stage).
__global__ void __launch_bounds__(1024) fold_hist_kernel(uint32_t *hist, int64_t stride,
                                                         uint32_t R, int copies, int win) {
  uint32_t *h = hist + blockIdx.x * stride;
  const uint32_t win_lo = R > (uint32_t)win / 2 ? R - win / 2 : 0;
  for (int i = threadIdx.x; i < win; i += blockDim.x) {
    uint32_t acc = 0;
    for (int c = 0; c < copies; c++) {
      uint32_t *cp = h + 2 * (int64_t)R + (int64_t)c * win + i;
      acc += *cp;
      *cp = 0;
    }
    if (acc && win_lo + i < 2 * R) h[win_lo + i] += acc;
  }
}

// one block per stage: kexp and the largest bin
__global__ void __launch_bounds__(256) scan_hist_kernel(const uint32_t *hist, int64_t stride,
                                                        uint32_t R, StageScan *out) {
  const uint32_t *h = hist + blockIdx.x * stride;
  const int64_t twoR = 2 * (int64_t)R;
  uint32_t peak = 0, mx = 0;
  bool any = false;
  for (int64_t c = 1 + threadIdx.x; c < twoR; c += blockDim.x) {
    const uint32_t v = h[c];
    if (v) {
      const int64_t q = c - (int64_t)R;
      const uint32_t a = (uint32_t)(q < 0 ? -q : q);
      peak = a > peak ? a : peak;
      any = true;
    }
    mx = v > mx ? v : mx;
  }
  __shared__ uint32_t s_peak[256], s_mx[256];
  __shared__ int s_any[256];
  s_peak[threadIdx.x] = peak;
  s_mx[threadIdx.x] = mx;
  s_any[threadIdx.x] = any;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_peak[threadIdx.x] = max(s_peak[threadIdx.x], s_peak[threadIdx.x + o]);
      s_mx[threadIdx.x] = max(s_mx[threadIdx.x], s_mx[threadIdx.x + o]);
      s_any[threadIdx.x] |= s_any[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int k = 1;
    if (s_any[0]) {
      int b = 0;
      while (b < 32 && (s_peak[0] >> b)) b++;
      k = b > 1 ? b : 1;
    }
    out[blockIdx.x].kexp = k;
    out[blockIdx.x].hist0 = h[0];
    out[blockIdx.x].maxbin = max(s_mx[0], h[0]);
  }
}

// Per-tile counts of up to 4 "big" codes (bins reaching 2^24) in a stage's
// code array; counts[tile * 4 + j].
template <typename CodeT>
__global__ void __launch_bounds__(256) tile_count_kernel(const CodeT *codes, int64_t n,
                                                         uint4 big, int nbig,
                                                         unsigned long long *counts) {
  __shared__ unsigned int lc[64 * 4];
  for (int i = threadIdx.x; i < 64 * 4; i += blockDim.x) lc[i] = 0;
  __syncthreads();
  const uint32_t bv[4] = {big.x, big.y, big.z, big.w};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t c = (uint32_t)codes[i];
    const int tile = (int)(i >> 22);
    for (int j = 0; j < nbig; j++)
      if (c == bv[j]) atomicAdd(&lc[tile * 4 + j], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 4; i += blockDim.x)
    if (lc[i]) atomicAdd(&counts[i], (unsigned long long)lc[i]);
}

static int choose_owx(int64_t n_out, int64_t n_codes) {  // rans.py:162-169
  if (n_codes <= 0 || n_out <= 0) return 0;
  const double rate = (double)n_out / (double)n_codes, lr = std::log(rate);
  int best = 0;
  double bd = INFINITY;
  for (int i = 0; i < 8; i++) {
    const double w = OUTLIER_W[i] > 1e-12 ? OUTLIER_W[i] : 1e-12;
    const double d = std::fabs(std::log(w) - lr);
    if (d < bd) {
      bd = d;
      best = i;
    }
  }
  return best;
}

// cost[job][level] for a batch of (stage, shape) jobs: fixed-order float64
// sums of f64(w_f32) * f64(hist_f32) (ascending symbols in segments of 256),
// rounded to float32; one workgroup per job, one thread per level.
struct CostJob {
  const float *nl2;   // (64, A) -log2 weights of the job's shape
  int64_t h_off, A;   // the stage's float32 histogram in the batch array
};
__global__ void __launch_bounds__(64) fit_cost_batch_kernel(const CostJob *jobs, int njobs,
                                                            const float *hist32, float *cost) {
  const int jb = blockIdx.x, lvl = threadIdx.x;
  if (jb >= njobs) return;
  const CostJob J = jobs[jb];
  const float *row = J.nl2 + lvl * J.A;
  const float *h = hist32 + J.h_off;
  double tot = 0.0;
  for (int64_t a = 0; a < J.A; a += 256) {
    double seg = 0.0;
    const int64_t b = a + 256 < J.A ? a + 256 : J.A;
    for (int64_t k = a; k < b; k++) seg += (double)row[k] * (double)h[k];
    tot += seg;
  }
  cost[(int64_t)jb * NLEV + lvl] = (float)tot;
}

// Fit stages given their full histograms (device), code arrays (device) and
// per-stage code counts.  All stages move together: one fold + scan launch,
// one batch of window copies, one cost launch over every (stage, shape) job,
// one copy of the costs back (plus the per-tile recount of the rare stages
// whose bins reach 2^24).
int fit_hists(int nst, const uint32_t *d_hist, const void *d_codes,
              const int64_t *code_off, const int64_t *n_codes, uint32_t R,
              int code_bytes, int precision, latec_fit *fits, hipStream_t st) {
  if (nst <= 0) return LATEC_OK;
  const int64_t twoR = 2 * (int64_t)R;
  const int64_t HS = hist_words(R);
  int rc;
  StageScan *d_scan;
  hipLaunchKernelGGL(fold_hist_kernel, dim3(nst), dim3(1024), 0, st, (uint32_t *)d_hist, HS, R,
                     256, 1024);
  if ((rc = hip_ok(hipMallocAsync((void **)&d_scan, sizeof(StageScan) * nst, st), "hipMallocAsync")))
    return rc;
  hipLaunchKernelGGL(scan_hist_kernel, dim3(nst), dim3(256), 0, st, d_hist, HS, R, d_scan);
  std::vector<StageScan> scan(nst);
  rc = hip_ok(hipMemcpyAsync(scan.data(), d_scan, sizeof(StageScan) * nst, hipMemcpyDeviceToHost, st),
              "copy");
  if (!rc) rc = hip_ok(hipStreamSynchronize(st), "sync");
  (void)hipFreeAsync(d_scan, st);
  if (rc) return rc;

  // window copies of every stage, one synchronization
  std::vector<std::vector<uint32_t>> win(nst);
  for (int s = 0; s < nst; s++) {
    latec_fit &f = fits[s];
    memset(&f, 0, sizeof f);
    f.n_codes = (uint32_t)n_codes[s];
    f.kexp = n_codes[s] ? scan[s].kexp : 1;
    if (!n_codes[s]) continue;
    const int64_t span = ((int64_t)1 << f.kexp) - 1;
    win[s].assign((size_t)(2 * span + 1), 0u);
    const int64_t c_lo = std::max<int64_t>((int64_t)R - span, 1);
    const int64_t c_hi = std::min<int64_t>((int64_t)R + span, twoR - 1);
    if (c_hi >= c_lo &&
        (rc = hip_ok(hipMemcpyAsync(win[s].data() + (c_lo - ((int64_t)R - span)),
                                    d_hist + s * HS + c_lo, sizeof(uint32_t) * (c_hi - c_lo + 1),
                                    hipMemcpyDeviceToHost, st), "copy")))
      return rc;
  }
  if ((rc = hip_ok(hipStreamSynchronize(st), "sync"))) return rc;

  std::lock_guard<std::mutex> lk(g_cache.mu);
  std::vector<float> h32all;
  std::vector<int64_t> h_off(nst, 0);
  std::vector<CostJob> jobs;
  std::vector<int> job0(nst, 0), njob(nst, 0);
  std::vector<std::array<int, 21>> sids(nst);
  for (int s = 0; s < nst; s++) {
    latec_fit &f = fits[s];
    if (!n_codes[s]) continue;
    const int kexp = f.kexp;
    const int64_t A = (int64_t)2 << kexp, span = ((int64_t)1 << kexp) - 1;
    // float32 histogram over the reduced alphabet (rans.py:413-420)
    std::vector<float> h32(A);
    h32[0] = (float)scan[s].hist0;
    for (int64_t k = 1; k < A; k++) h32[k] = (float)win[s][k - 1];
    if (scan[s].maxbin >= (1u << 24) && n_codes[s] > (1 << 22)) {
      // bins >= 2^24 need the per-tile float32 accumulation
      std::vector<uint32_t> big;
      std::vector<int64_t> bigk;
      if (scan[s].hist0 >= (1u << 24)) { big.push_back(0); bigk.push_back(0); }
      for (int64_t k = 1; k < A && big.size() < 4; k++)
        if (win[s][k - 1] >= (1u << 24)) {
          big.push_back((uint32_t)(R - span + k - 1));
          bigk.push_back(k);
        }
      const int64_t ntiles = (n_codes[s] + (1 << 22) - 1) >> 22;
      if (ntiles > 64) {
        set_error("stage too large for the tile histogram (>2^28 codes)");
        return LATEC_E_ARG;
      }
      unsigned long long *d_cnt;
      if ((rc = hip_ok(hipMallocAsync((void **)&d_cnt, sizeof(unsigned long long) * 256, st),
                       "hipMallocAsync")))
        return rc;
      (void)hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long) * 256, st);
      uint4 bv = make_uint4(big.size() > 0 ? big[0] : 0, big.size() > 1 ? big[1] : 0,
                            big.size() > 2 ? big[2] : 0, big.size() > 3 ? big[3] : 0);
      const int64_t n = n_codes[s];
      unsigned blocks = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
      if (code_bytes == 2)
        hipLaunchKernelGGL(tile_count_kernel<uint16_t>, dim3(blocks), dim3(256), 0, st,
                           (const uint16_t *)d_codes + code_off[s], n, bv, (int)big.size(), d_cnt);
      else
        hipLaunchKernelGGL(tile_count_kernel<uint32_t>, dim3(blocks), dim3(256), 0, st,
                           (const uint32_t *)d_codes + code_off[s], n, bv, (int)big.size(), d_cnt);
      std::vector<unsigned long long> cnt(256);
      rc = hip_ok(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(unsigned long long) * 256,
                                 hipMemcpyDeviceToHost, st), "copy");
      if (!rc) rc = hip_ok(hipStreamSynchronize(st), "sync");
      (void)hipFreeAsync(d_cnt, st);
      if (rc) return rc;
      for (size_t j = 0; j < big.size(); j++) {
        float acc = 0.0f;
        for (int64_t t = 0; t < ntiles; t++)
          acc = (float)((double)acc + (double)cnt[t * 4 + j]);
        h32[bigk[j]] = acc;
      }
    }
    f.n_out = scan[s].hist0;
    f.owx = choose_owx((int64_t)h32[0], n_codes[s]);
    const int ns = shape_ids(kexp, sids[s].data());
    h_off[s] = (int64_t)h32all.size();
    h32all.insert(h32all.end(), h32.begin(), h32.end());
    job0[s] = (int)jobs.size();
    njob[s] = ns;
    for (int i = 0; i < ns; i++) {
      float *t;
      if ((rc = ensure_nl2(kexp, sids[s][i], f.owx, precision, st, &t))) return rc;
      jobs.push_back({t, h_off[s], A});
    }
  }
  if (jobs.empty()) return LATEC_OK;
  const int nj = (int)jobs.size();
  CostJob *d_jobs;
  float *d_h32, *d_cost;
  if ((rc = hip_ok(hipMallocAsync((void **)&d_jobs, sizeof(CostJob) * nj, st), "hipMallocAsync")))
    return rc;
  if ((rc = hip_ok(hipMallocAsync((void **)&d_h32, sizeof(float) * h32all.size(), st),
                   "hipMallocAsync")))
    return rc;
  if ((rc = hip_ok(hipMallocAsync((void **)&d_cost, sizeof(float) * (size_t)nj * NLEV, st),
                   "hipMallocAsync")))
    return rc;
  (void)hipMemcpyAsync(d_jobs, jobs.data(), sizeof(CostJob) * nj, hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(d_h32, h32all.data(), sizeof(float) * h32all.size(), hipMemcpyHostToDevice,
                       st);
  hipLaunchKernelGGL(fit_cost_batch_kernel, dim3(nj), dim3(NLEV), 0, st, d_jobs, nj, d_h32,
                     d_cost);
  std::vector<float> cost((size_t)nj * NLEV);
  rc = hip_ok(hipMemcpyAsync(cost.data(), d_cost, sizeof(float) * cost.size(),
                             hipMemcpyDeviceToHost, st), "copy");
  if (!rc) rc = hip_ok(hipStreamSynchronize(st), "sync");
  (void)hipFreeAsync(d_jobs, st);
  (void)hipFreeAsync(d_h32, st);
  (void)hipFreeAsync(d_cost, st);
  if (rc) return rc;
  for (int s = 0; s < nst; s++) {
    if (!n_codes[s]) continue;
    latec_fit &f = fits[s];
    float best = INFINITY;
    for (int i = 0; i < njob[s]; i++) {
      const float *c = cost.data() + (size_t)(job0[s] + i) * NLEV;
      int bl = 0;
      for (int l = 1; l < NLEV; l++)
        if (c[l] < c[bl]) bl = l;
      if (c[bl] < best) {
        best = c[bl];
        f.level = bl;
        f.shape = sids[s][i];
      }
    }
    f.bits = best;
  }
  return LATEC_OK;
}

// scale_to_level (rans.py:270-278) on device by comparisons only: the level of
// v is the number of the 63 float32 thresholds at or below it (thr[k] = the
// smallest float32 the host's NumPy scale_to_level puts at level k + 1), so the
// device reproduces the host's float32 log rounding without evaluating a log
// (the bucketization of latec/gnn_codec.py:827-861).  A NaN scale compares
// below every threshold: level 0.
struct ScaleThr {
  float t[64];  // t[63] = +inf (a sentinel above every finite scale)
};
__global__ void scale_to_level_kernel(const float *s, int64_t n, ScaleThr thr, uint8_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = s[i];
    // branch-free search: largest c in [0, 63] with thr[c - 1] <= v
    int c = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
      c += (v >= thr.t[c + step - 1]) ? step : 0;
    out[i] = (uint8_t)c;
  }
}

int scale_to_level(const float *d_scale, int64_t n, const float *thresholds, uint8_t *d_levels,
                   hipStream_t st) {
  if (!thresholds) {
    set_error("scale_to_level: thresholds required");
    return LATEC_E_ARG;
  }
  ScaleThr thr;
  for (int k = 0; k < 63; k++) {
    thr.t[k] = thresholds[k];
    if (!(thr.t[k] > 0.0f) || (k && thr.t[k] < thr.t[k - 1])) {
      set_error("scale level thresholds must be positive and ascending");
      return LATEC_E_ARG;
    }
  }
  thr.t[63] = INFINITY;
  if (!n) return LATEC_OK;
  unsigned blocks = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  hipLaunchKernelGGL(scale_to_level_kernel, dim3(blocks), dim3(256), 0, st, d_scale, n, thr,
                     d_levels);
  return hip_ok(hipGetLastError(), "scale_to_level_kernel");
}

}  // namespace latec
