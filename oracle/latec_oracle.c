/*
 * latec_oracle.c -- CPU oracle for the LATEC plain-interpolation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference's algorithm, used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  The
 * product (latec_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Compiled with
 * -ffp-contract=off so every FP64 expression is evaluated as written (one
 * IEEE round-to-nearest-even per operation, no FMA), matching NumPy.
 *
 * Parity status (details in DESIGN.md):
 *   schedule, interp, quantizer, kexp/narrow, record framing, header : pinned
 *     (closed forms, brute force, the reference tests' whole-grid oracle).
 *   Laplace tables / -log2 weights / const-level fit : pinned by golden
 *     vectors written from the reference's own rans.py (tests/golden/).
 *   stack rANS word stream (constriction AnsCoder + Categorical(perfect=False))
 *     and zstd frame bytes : parity unpinned (third-party, absent here).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zstd.h>

#define ORC_MAXDIM 8
#define N_LEVELS 64
#define PRECISION 24

/* ------------------------------------------------------------------------ */
/* Schedule: latec/levels.py:100-147 (iter_stage_plan), masks in flat order   */
/* ------------------------------------------------------------------------ */

int orc_n_stages(int ndim, int levels) { return 1 + levels * (2 * ndim - 1); }

/* Fill mask (uint8, N points) for stage `stage` given `covered` (updated in
 * place).  Returns the stage stride; *weight_out = 0 for anchors, 1 for a
 * single-axis sub-stage (axis in *axis_out), w>=2 for a fused sub-stage.
 * levels.py:100-147. */
int orc_stage_mask(const int64_t *shape, int ndim, int levels, int stride,
                   int block, int stage, uint8_t *covered, uint8_t *mask,
                   int *weight_out, int *axis_out) {
  int64_t N = 1;
  for (int j = 0; j < ndim; j++) N *= shape[j];
  int64_t coords[ORC_MAXDIM];
  if (stage == 0) {
    memset(coords, 0, sizeof coords);
    for (int64_t p = 0; p < N; p++) {
      int an = 1;
      for (int j = 0; j < ndim; j++) an &= (coords[j] % stride) < block;
      mask[p] = (uint8_t)an;
      covered[p] |= mask[p];
      for (int j = ndim - 1; j >= 0; j--) {
        if (++coords[j] < shape[j]) break;
        coords[j] = 0;
      }
    }
    *weight_out = 0;
    *axis_out = -1;
    return stride;
  }
  int per = 2 * ndim - 1;
  int k = 1 + (stage - 1) / per;
  int sub = (stage - 1) % per;
  int s = stride >> k;
  if (s < 1) s = 1;
  int single_axis = sub < ndim ? sub : -1;
  int w = sub < ndim ? 1 : sub - ndim + 2;
  int absorb = (k == levels) && ((ndim == 1 && single_axis >= 0) ||
                                 (single_axis < 0 && w == ndim));
  memset(coords, 0, sizeof coords);
  for (int64_t p = 0; p < N; p++) {
    int odd = 0, ok = 1, odd_axis = -1;
    for (int j = 0; j < ndim; j++) {
      int64_t r = coords[j] % (2 * s);
      if (r == s) {
        odd++;
        odd_axis = j;
      } else if (r != 0) {
        ok = 0;
      }
    }
    uint8_t m;
    if (single_axis >= 0)
      m = ok && odd == 1 && odd_axis == single_axis;
    else
      m = ok && odd == w;
    m = m && !covered[p];
    if (absorb && !covered[p]) m = 1;
    mask[p] = m;
    covered[p] |= m;
    for (int j = ndim - 1; j >= 0; j--) {
      if (++coords[j] < shape[j]) break;
      coords[j] = 0;
    }
  }
  *weight_out = w;
  *axis_out = single_axis;
  return s;
}

/* stage_ebs: levels.py:360-386 -- eb * ratio ** log2(stride / finest). */
void orc_stage_ebs(int ndim, int levels, int stride, double eb, double ratio,
                   double *out) {
  int n = orc_n_stages(ndim, levels);
  int finest = stride >> levels;
  if (finest < 1) finest = 1;
  out[0] = eb * pow(ratio, log2((double)stride / (double)finest));
  for (int i = 1; i < n; i++) {
    int k = 1 + (i - 1) / (2 * ndim - 1);
    int s = stride >> k;
    if (s < 1) s = 1;
    out[i] = eb * pow(ratio, log2((double)s / (double)finest));
  }
}

/* ------------------------------------------------------------------------ */
/* Interpolation: latec/predictor.py:75-156 (_flat_gatherer, _interp_axis_at) */
/* and predictor.py:255-317 (predict, centre combine)                         */
/* ------------------------------------------------------------------------ */

#define END_QUAD 1
#define END_EXTRAP 2

static double interp_axis(const float *W, const int64_t *coords,
                          const int64_t *shape, const int64_t *gstr,
                          int64_t flat, int axis, int s, int cubic,
                          int end_mode) {
  int64_t n = shape[axis], ca = coords[axis];
#define TAP(off, val, valid)                                    \
  do {                                                          \
    int64_t t = ca + (off);                                     \
    valid = (t >= 0) && (t < n);                                \
    int64_t tc = t < 0 ? 0 : (t > n - 1 ? n - 1 : t);           \
    val = (double)W[flat + (tc - ca) * gstr[axis]];             \
  } while (0)
  double Lm1, Lp1, Lm3, Lp3;
  int vm1, vp1, vm3, vp3;
  TAP(-s, Lm1, vm1);
  TAP(+s, Lp1, vp1);
  TAP(-3 * s, Lm3, vm3);
  TAP(+3 * s, Lp3, vp3);
#undef TAP
  double lin = 0.5 * (Lm1 + Lp1);
  double pred = lin;
  if (cubic) {
    double cub = (((-Lm3 + 9.0 * Lm1) + 9.0 * Lp1) - Lp3) / 16.0;
    double one_far = lin;
    if (end_mode & END_QUAD) {
      if (vp3 && !vm3)
        one_far = ((3.0 * Lm1 + 6.0 * Lp1) - Lp3) / 8.0;
      else if (vm3 && !vp3)
        one_far = ((-Lm3 + 6.0 * Lm1) + 3.0 * Lp1) / 8.0;
    }
    pred = (vm3 && vp3) ? cub : one_far;
  }
  double left = Lm1, right = Lp1;
  if (end_mode & END_EXTRAP) {
    left = vm3 ? 1.5 * Lm1 - 0.5 * Lm3 : Lm1;
    right = vp3 ? 1.5 * Lp1 - 0.5 * Lp3 : Lp1;
  }
  if (vm1 && vp1) return pred;
  return (vm1 && !vp1) ? left : right;
}

/* Predict every point of `mask` (flat order) from recon (C, N), output (C, n)
 * float32.  weight/axis as returned by orc_stage_mask.  predictor.py:255-317.
 * Masked (even-axis) terms are evaluated exactly like the reference. */
void orc_predict(const float *recon, int C, const int64_t *shape, int ndim,
                 const uint8_t *mask, int s, int weight, int axis, int cubic,
                 int center, int end_mode, float *out, int64_t n) {
  int64_t N = 1, gstr[ORC_MAXDIM], coords[ORC_MAXDIM];
  for (int j = ndim - 1; j >= 0; j--) {
    gstr[j] = N;
    N *= shape[j];
  }
  int axes[ORC_MAXDIM], na = 0;
  if (weight == 1)
    axes[na++] = axis;
  else
    for (int j = 0; j < ndim; j++) axes[na++] = j;
  memset(coords, 0, sizeof coords);
  int64_t at = 0;
  for (int64_t p = 0; p < N; p++) {
    if (mask[p]) {
      for (int c = 0; c < C; c++) {
        const float *W = recon + (int64_t)c * N;
        double res;
        if (center == 0 || na == 1) {
          double total = 0.0, count = 0.0;
          for (int i = 0; i < na; i++) {
            int a = axes[i];
            double v = interp_axis(W, coords, shape, gstr, p, a, s, cubic,
                                   end_mode);
            double m = (coords[a] & s) != 0 ? 1.0 : 0.0;
            if (i == 0) {
              total = v * m;
              count = m;
            } else {
              total = total + v * m;
              count = count + m;
            }
          }
          res = total / count;
        } else {
          res = 0.0;
          for (int i = 0; i < na; i++) {
            int a = center == 1 ? axes[na - 1 - i] : axes[i];
            double v = interp_axis(W, coords, shape, gstr, p, a, s, cubic,
                                   end_mode);
            int take = (coords[a] & s) != 0;
            if (i == 0 || take) res = v;
          }
        }
        out[(int64_t)c * n + at] = (float)res;
      }
      at++;
    }
    for (int j = ndim - 1; j >= 0; j--) {
      if (++coords[j] < shape[j]) break;
      coords[j] = 0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Quantizer: latec/quantizer.py:28-120                                       */
/* ------------------------------------------------------------------------ */

static float recon_from_code(float p, uint32_t code, double eb, int64_t R) {
  int64_t q = (int64_t)code - R;
  return (float)((double)p + (2.0 * eb) * (double)q); /* quantizer.py:114-120 */
}

/* round_mode: 0 none, 1 rint (integer source), 2 (span, offset) rounding.
 * Returns number of outliers written to outl (scan order). */
int64_t orc_quantize(const float *x, const float *pred, int64_t n, double eb,
                     int64_t R, int round_mode, double span, double offset,
                     uint32_t *codes, float *outl) {
  double w = 2.0 * eb;
  float feb = (float)eb;
  int64_t no = 0;
  for (int64_t i = 0; i < n; i++) {
    double qd = rint(((double)x[i] - (double)pred[i]) / w);
    uint32_t code = 0;
    int inr = fabs(qd) < (double)R; /* NaN/inf/huge -> demoted (see header) */
    if (inr) code = (uint32_t)((int64_t)qd + R);
    float rec = recon_from_code(pred[i], code, eb, R);
    float chk = rec;
    if (round_mode == 1)
      chk = rintf(rec);
    else if (round_mode == 2)
      chk = (float)((rint((double)rec * span + offset) - offset) / span);
    int ok = inr && (fabsf(x[i] - chk) <= feb);
    if (!ok) {
      code = 0;
      outl[no++] = x[i];
    }
    codes[i] = code;
  }
  return no;
}

/* dequantize: quantizer.py:90-111.  Returns -1 on outlier-count mismatch. */
int orc_dequantize(const float *pred, const uint32_t *codes,
                   const float *outl, int64_t n_out, int64_t n, double eb,
                   int64_t R, float *recon) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) {
    if (codes[i] == 0) {
      if (k >= n_out) return -1;
      recon[i] = outl[k++];
    } else {
      recon[i] = recon_from_code(pred[i], codes[i], eb, R);
    }
  }
  return k == n_out ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* Laplace-mixture tables: latec/rans.py:119-180, 247-267, 377-392, 676-722   */
/* ------------------------------------------------------------------------ */

static const int MIX_D[4] = {4, 8, 16, 24};
static const double MIX_W[5] = {0.02, 0.05, 0.12, 0.25, 0.45};
static const double OUTLIER_W[8] = {0.0,         0x1p-13, 0x1p-10, 0x1p-8,
                                    0x1p-6,      0x1p-4,  0x1p-2,  0.5};

static void shape_params(int shape, int *d, double *w) {
  if (shape == 0) {
    *d = 0;
    *w = 0.0;
    return;
  }
  *d = MIX_D[(shape - 1) / 5];
  *w = MIX_W[(shape - 1) % 5];
}

/* NumPy's pairwise summation of a contiguous float64 run (the add.reduce inner
 * loop): blocks of <= 128 use 8 accumulators, longer runs split in halves at a
 * multiple of 8. */
static double pairwise_sum(const double *a, int64_t n) {
  if (n < 8) {
    double r = 0.0; /* numpy starts from -0.0 in newer versions; same for +vals */
    for (int64_t i = 0; i < n; i++) r = i == 0 ? a[0] : r + a[i];
    if (n == 0) r = 0.0;
    return r;
  } else if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  } else {
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
  }
}

/* NumPy's add.reduce along a row of a 2-D array: the iterator hands the
 * inner loop 8192-element buffers, each summed pairwise, accumulated in
 * order. */
static double row_sum(const double *a, int64_t n) {
  double r = 0.0;
  for (int64_t i = 0; i < n; i += 8192) {
    int64_t m = n - i < 8192 ? n - i : 8192;
    double c = pairwise_sum(a + i, m);
    r = i == 0 ? c : r + c;
  }
  return r;
}

static double laplace_cdf(double x, double b) {
  double z = x / b;
  if (z < -700.0) z = -700.0;
  if (z > 700.0) z = 700.0;
  return x < 0 ? 0.5 * exp(z) : 1.0 - 0.5 * exp(-z);
}

/* numpy.linspace(start, stop, num) followed by exp: rans.py:257, 297 */
static void scale_grid(double lo, double hi, int n, double *grid) {
  double start = log(lo), stop = log(hi);
  double step = (stop - start) / (double)(n - 1);
  for (int i = 0; i < n; i++) {
    double y = (double)i * step;
    y = y + start;
    if (i == n - 1) y = stop;
    grid[i] = exp(y);
  }
}

/* _level_weights(kexp, 64): rans.py:247-267.  out: (64, A) row-normalized. */
void orc_level_weights(int kexp, double *out) {
  int64_t span = ((int64_t)1 << kexp) - 1, A = 2 * span + 2;
  double grid[N_LEVELS];
  scale_grid(1.0 / 16.0, 4096.0, N_LEVELS, grid);
  for (int i = 0; i < N_LEVELS; i++) {
    double *row = out + (int64_t)i * A;
    double b = grid[i], mx = -INFINITY;
    for (int64_t k = 0; k < 2 * span + 1; k++) {
      double q = (double)(k - span);
      double left = q * 2.0 - 1.0, right = q * 2.0 + 1.0;
      double v = laplace_cdf(right, b) - laplace_cdf(left, b);
      row[k + 1] = v;
      if (v > mx) mx = v;
    }
    row[0] = mx * 1e-9;
    double s = row_sum(row, A);
    for (int64_t k = 0; k < A; k++) row[k] = row[k] / s;
  }
}

/* _shaped_weights + _with_outlier_weight: rans.py:377-384, 172-180 */
void orc_model_weights(int kexp, int shape, int owx, double *out) {
  int64_t A = ((int64_t)2 << kexp);
  double *base = (double *)malloc(sizeof(double) * N_LEVELS * A);
  orc_level_weights(kexp, base);
  int d;
  double w;
  shape_params(shape, &d, &w);
  for (int i = 0; i < N_LEVELS; i++) {
    int j = i + d < N_LEVELS - 1 ? i + d : N_LEVELS - 1;
    for (int64_t k = 0; k < A; k++) {
      if (w == 0.0)
        out[i * A + k] = base[i * A + k];
      else
        out[i * A + k] = (1.0 - w) * base[i * A + k] + w * base[j * A + k];
    }
  }
  free(base);
  double ow = OUTLIER_W[owx];
  if (ow != 0.0) {
    for (int i = 0; i < N_LEVELS; i++) {
      double *row = out + i * A;
      double f = (1.0 - ow) / row_sum(row + 1, A - 1);
      for (int64_t k = 1; k < A; k++) row[k] = row[k] * f;
      row[0] = ow;
    }
  }
}

/* _neg_log2_weights: rans.py:387-392 (float32 result) */
void orc_neg_log2_weights(int kexp, int shape, int owx, int precision,
                          float *out) {
  int64_t A = ((int64_t)2 << kexp);
  double *w = (double *)malloc(sizeof(double) * N_LEVELS * A);
  orc_model_weights(kexp, shape, owx, w);
  double floor_ = 1.0 / (double)((int64_t)1 << precision);
  for (int64_t i = 0; i < N_LEVELS * A; i++)
    out[i] = (float)(-log2(w[i] > floor_ ? w[i] : floor_));
  free(w);
}

typedef struct {
  double frac;
  int64_t idx;
} fr_t;
static int cmp_desc(const void *a, const void *b) { /* -frac asc, stable */
  const fr_t *x = a, *y = b;
  double nx = -x->frac, ny = -y->frac;
  if (nx < ny) return -1;
  if (nx > ny) return 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static int cmp_asc(const void *a, const void *b) { /* frac asc, stable */
  const fr_t *x = a, *y = b;
  if (x->frac < y->frac) return -1;
  if (x->frac > y->frac) return 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* _quantize_pmf_rows for one row: rans.py:676-722.  Returns 0 / -1. */
static int quantize_pmf_row(const double *w, int64_t A, int64_t total,
                            int64_t *pmf, fr_t *scratch) {
  double rs = row_sum(w, A);
  int64_t sum = 0;
  for (int64_t k = 0; k < A; k++) {
    double sc = w[k] / rs * (double)total;
    double fl = floor(sc);
    int64_t f = (int64_t)fl;
    pmf[k] = f > 1 ? f : 1;
    scratch[k].frac = sc - fl;
    scratch[k].idx = k;
    sum += pmf[k];
  }
  int64_t diff = total - sum;
  if (diff > 0) {
    int64_t bulk = diff / A, rem = diff % A;
    for (int64_t k = 0; k < A; k++) pmf[k] += bulk;
    qsort(scratch, A, sizeof(fr_t), cmp_desc);
    for (int64_t r = 0; r < rem; r++) pmf[scratch[r].idx] += 1;
  } else if (diff < 0) {
    int64_t need = -diff, cum = 0;
    qsort(scratch, A, sizeof(fr_t), cmp_asc);
    for (int64_t r = 0; r < A; r++) {
      int64_t k = scratch[r].idx, avail = pmf[k] - 1;
      int64_t before = cum;
      cum += avail;
      int64_t take = need - before;
      if (take < 0) take = 0;
      if (take > avail) take = avail;
      pmf[k] -= take;
    }
    if (cum < need) return -1;
  }
  return 0;
}

/* _build_cdfs_cached: rans.py:305-329.  out: (64, A+1) uint32 CDFs. */
int orc_build_cdfs(int kexp, int shape, int precision, int owx,
                   uint32_t *out) {
  int64_t A = ((int64_t)2 << kexp), total = (int64_t)1 << precision;
  if (A > total) return -1;
  double *w = (double *)malloc(sizeof(double) * N_LEVELS * A);
  int64_t *pmf = (int64_t *)malloc(sizeof(int64_t) * A);
  fr_t *scr = (fr_t *)malloc(sizeof(fr_t) * A);
  orc_model_weights(kexp, shape, owx, w);
  int rc = 0;
  for (int i = 0; i < N_LEVELS && !rc; i++) {
    rc = quantize_pmf_row(w + i * A, A, total, pmf, scr);
    uint32_t *row = out + i * (A + 1);
    uint64_t c = 0;
    row[0] = 0;
    for (int64_t k = 0; k < A; k++) {
      c += (uint64_t)pmf[k];
      row[k + 1] = (uint32_t)c;
    }
  }
  free(w);
  free(pmf);
  free(scr);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Per-stage fit: rans.py:65-97, 162-169, 195-201, 395-430                    */
/* ------------------------------------------------------------------------ */

int orc_choose_kexp(const uint32_t *codes, int64_t n, int64_t R) {
  int64_t peak = -1;
  for (int64_t i = 0; i < n; i++)
    if (codes[i]) {
      int64_t q = (int64_t)codes[i] - R;
      if (q < 0) q = -q;
      if (q > peak) peak = q;
    }
  if (peak < 0) return 1;
  int b = 0;
  while (b < 63 && (peak >> b)) b++;
  return b > 1 ? b : 1;
}

int orc_choose_outlier_weight(int64_t n_out, int64_t n_codes) {
  if (n_codes <= 0 || n_out <= 0) return 0;
  double rate = (double)n_out / (double)n_codes, lr = log(rate);
  int best = 0;
  double bd = INFINITY;
  for (int i = 0; i < 8; i++) {
    double w = OUTLIER_W[i] > 1e-12 ? OUTLIER_W[i] : 1e-12;
    double dd = fabs(log(w) - lr);
    if (dd < bd) {
      bd = dd;
      best = i;
    }
  }
  return best;
}

int orc_n_shape_ids(int kexp, int *ids) {
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

/* Fixed-order cost of one model row against the float32 histogram: products
 * of two floats are exact in float64; they are summed in ascending symbol
 * order within segments of 256 symbols, and segment sums in ascending order.
 * The reference computes this with a float32 BLAS sgemv (rans.py:424-426);
 * the two agree except on near-ties (recorded in DESIGN.md). */
static double fit_cost(const float *wrow, const float *hist, int64_t A) {
  double tot = 0.0;
  for (int64_t a = 0; a < A; a += 256) {
    double seg = 0.0;
    int64_t b = a + 256 < A ? a + 256 : A;
    for (int64_t k = a; k < b; k++) seg += (double)wrow[k] * (double)hist[k];
    tot += seg;
  }
  return tot;
}

/* Histogram with the reference's per-tile float32 accumulation
 * (rans.py:413-420: hist_f32 += bincount(tile) per 2**22 codes). */
void orc_const_hist(const uint32_t *codes, int64_t n, int64_t R, int kexp,
                    float *hist) {
  int64_t A = (int64_t)2 << kexp, span = ((int64_t)1 << kexp) - 1;
  int64_t *cnt = (int64_t *)calloc(A, sizeof(int64_t));
  for (int64_t k = 0; k < A; k++) hist[k] = 0.0f;
  for (int64_t a = 0; a < n; a += (1 << 22)) {
    int64_t b = a + (1 << 22) < n ? a + (1 << 22) : n;
    memset(cnt, 0, sizeof(int64_t) * A);
    for (int64_t i = a; i < b; i++) {
      int64_t s = codes[i] == 0 ? 0 : (int64_t)codes[i] - R + span + 1;
      cnt[s]++;
    }
    for (int64_t k = 0; k < A; k++)
      hist[k] = (float)((double)hist[k] + (double)cnt[k]);
  }
  free(cnt);
}

/* choose_const_level: rans.py:395-430.  Returns bits (as float32). */
float orc_choose_const_level(const float *hist, int64_t n_codes, int kexp,
                             int precision, int *level, int *shape,
                             int *owx) {
  int64_t A = (int64_t)2 << kexp;
  *level = *shape = *owx = 0;
  if (n_codes == 0) return 0.0f;
  *owx = orc_choose_outlier_weight((int64_t)hist[0], n_codes);
  int ids[21], ns = orc_n_shape_ids(kexp, ids);
  float *wt = (float *)malloc(sizeof(float) * N_LEVELS * A);
  double best = INFINITY;
  for (int si = 0; si < ns; si++) {
    orc_neg_log2_weights(kexp, ids[si], *owx, precision, wt);
    int bl = 0;
    float bc = INFINITY; /* the reference's argmin runs over float32 costs */
    for (int l = 0; l < N_LEVELS; l++) {
      float c = (float)fit_cost(wt + l * A, hist, A);
      if (c < bc) {
        bc = c;
        bl = l;
      }
    }
    if (bc < best) {
      best = bc;
      *level = bl;
      *shape = ids[si];
    }
  }
  free(wt);
  return (float)best;
}

/* scale_to_level: rans.py:270-278 (float32 log promoted to float64). */
void orc_scale_to_level(const float *scale, int64_t n, double eb,
                        uint8_t *out) {
  double lo = eb / 16.0, hi = eb * 4096.0;
  float flo = (float)lo, fhi = (float)hi;
  double llo = log(lo), den = log(hi) - llo;
  for (int64_t i = 0; i < n; i++) {
    float s = scale[i];
    if (s < flo) s = flo;
    if (s > fhi) s = fhi;
    double t = ((double)logf(s) - llo) / den;
    double r = rint(t * 63.0);
    if (r < 0) r = 0;
    if (r > 63) r = 63;
    out[i] = (uint8_t)r;
  }
}

/* ------------------------------------------------------------------------ */
/* Stack rANS -- rans.py:466-535, 629-657.  The reference calls the Rust    */
/* crate constriction ~= 0.4.2 (pyproject.toml:16), absent here.  Restated  */
/* from its published design:                                                */
/*  - the model: Python Categorical(probabilities, perfect=False) builds    */
/*    ContiguousCategoricalEntropyModel::from_floating_point_probabilities_ */
/*    fast (module stream::model): scale = f32((2^24 - A) / sum), left[k] = */
/*    floor(f32(cum_k * scale)) + k, every symbol keeping >= 1;             */
/*  - the coder: AnsCoder<u32, u64> (module stream::stack), a stack: the    */
/*    reference encodes symbols in reverse so they decode in order; a push  */
/*    renormalizes (state >> 40 >= p: the low u32 word out, state >>= 32)   */
/*    before state = (state / p) << 24 | (left + state % p);                */
/*  - the seal (into_compressed): the final state's low word, then its      */
/*    high word when nonzero.                                               */
/* Pinned by hand-derived word arrays (tests/test_capi.py RANS_KATS) and     */
/* round trips; the bytes against constriction itself stay UNPINNED.         */
/* ------------------------------------------------------------------------ */

/* Float probabilities (float32 of the table's pmf, rans.py:646-649) to
 * left-sided cumulatives at 24 bits: every symbol keeps >= 1. */
static void categorical_fast(const uint32_t *cdf, int64_t A, uint32_t *left) {
  float sum = 0.0f;
  for (int64_t k = 0; k < A; k++) sum += (float)(cdf[k + 1] - cdf[k]);
  float scale = (float)((1 << PRECISION) - A) / sum;
  float cum = 0.0f;
  for (int64_t k = 0; k < A; k++) {
    left[k] = (uint32_t)(cum * scale) + (uint32_t)k;
    cum += (float)(cdf[k + 1] - cdf[k]);
  }
  left[A] = 1u << PRECISION;
}

/* syms: reduced-alphabet symbols.  Returns number of u32 words written. */
int64_t orc_rans_encode(const uint32_t *syms, int64_t n, const uint32_t *cdf,
                        int64_t A, uint32_t *words, int64_t cap) {
  uint32_t *left = (uint32_t *)malloc(sizeof(uint32_t) * (A + 1));
  categorical_fast(cdf, A, left);
  uint64_t state = 0;
  int64_t nw = 0;
  for (int64_t i = n - 1; i >= 0; i--) {
    uint32_t s = syms[i];
    uint64_t lo = left[s], pr = left[s + 1] - left[s];
    if ((state >> (64 - PRECISION)) >= pr) {
      if (nw >= cap) {
        free(left);
        return -1;
      }
      words[nw++] = (uint32_t)state;
      state >>= 32;
    }
    uint64_t pre = state / pr, rem = state % pr;
    state = (pre << PRECISION) | (lo + rem);
  }
  free(left);
  if (state) {
    if (nw + 2 > cap) return -1;
    words[nw++] = (uint32_t)state;
    if (state >> 32) words[nw++] = (uint32_t)(state >> 32);
  }
  return nw;
}

/* Per-point levels (latec/rans.py:546-578): one stack coder; for level
 * 63, 62, ..., 0 the symbols of that level, in position order, are pushed with
 * encode_reverse (last first) under that level's CDF row.  Returns the word
 * count or -1. */
int64_t orc_rans_encode_levels(const uint32_t *syms, const uint8_t *levels, int64_t n,
                               const uint32_t *cdfs, int64_t A, uint32_t *words,
                               int64_t cap) {
  uint32_t *left = (uint32_t *)malloc(sizeof(uint32_t) * (A + 1));
  uint64_t state = 0;
  int64_t nw = 0;
  for (int level = 63; level >= 0; level--) {
    categorical_fast(cdfs + (int64_t)level * (A + 1), A, left);
    for (int64_t i = n - 1; i >= 0; i--) {
      if (levels[i] != level) continue;
      uint32_t s = syms[i];
      uint64_t lo = left[s], pr = left[s + 1] - left[s];
      if ((state >> (64 - PRECISION)) >= pr) {
        if (nw >= cap) {
          free(left);
          return -1;
        }
        words[nw++] = (uint32_t)state;
        state >>= 32;
      }
      uint64_t pre = state / pr, rem = state % pr;
      state = (pre << PRECISION) | (lo + rem);
    }
  }
  free(left);
  if (state) {
    if (nw + 2 > cap) return -1;
    words[nw++] = (uint32_t)state;
    if (state >> 32) words[nw++] = (uint32_t)(state >> 32);
  }
  return nw;
}

/* Returns 0 ok, -1 trailing data / malformed. */
int orc_rans_decode(const uint32_t *words, int64_t nw, const uint32_t *cdf,
                    int64_t A, uint32_t *syms, int64_t n) {
  uint32_t *left = (uint32_t *)malloc(sizeof(uint32_t) * (A + 1));
  categorical_fast(cdf, A, left);
  int64_t top = nw;
  uint64_t state = 0;
  if (top > 0) {
    uint32_t w0 = words[--top];
    if (w0 == 0) {
      free(left);
      return -1;
    }
    state = w0;
    if (top > 0) state = (state << 32) | words[--top];
  }
  for (int64_t i = 0; i < n; i++) {
    uint32_t qt = (uint32_t)(state & ((1u << PRECISION) - 1));
    int64_t lo = 0, hi = A; /* largest k with left[k] <= qt */
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) / 2;
      if (left[mid] <= qt)
        lo = mid;
      else
        hi = mid;
    }
    syms[i] = (uint32_t)lo;
    uint64_t pr = left[lo + 1] - left[lo];
    state = (state >> PRECISION) * pr + (qt - left[lo]);
    if (state < ((uint64_t)1 << 32) && top > 0)
      state = (state << 32) | words[--top];
  }
  free(left);
  return (top == 0 && state == 0) ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* Stage record: latec/bitstream.py:178-344                                   */
/* ------------------------------------------------------------------------ */

static void put_u32(uint8_t *p, uint32_t v) { memcpy(p, &v, 4); }
static void put_u64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }

/* pack_stage (const-level path).  Writes the record into out (capacity cap);
 * returns its length or -1.  bitstream.py:178-243. */
int64_t orc_pack_stage(const uint32_t *codes, int64_t n, const float *outl,
                       int64_t n_out, int64_t R, int zstd_level, int precision,
                       uint8_t *out, int64_t cap) {
  int kexp = orc_choose_kexp(codes, n, R);
  int64_t A = (int64_t)2 << kexp, span = ((int64_t)1 << kexp) - 1;
  float *hist = (float *)malloc(sizeof(float) * A);
  orc_const_hist(codes, n, R, kexp, hist);
  int level = 0, shape = 0, owx = 0;
  float bits = orc_choose_const_level(hist, n, kexp, precision, &level, &shape,
                                      &owx);
  free(hist);
  int64_t hlen = -1;
  int literal = 0;
  uint8_t *blob = out + 12;
  int64_t blob_cap = cap - 16 - 4 * n_out;
  if (blob_cap < 0) return -1;
  if (n >= 4096) {
    int wide = kexp > 7;
    size_t sb = (size_t)n * (wide ? 2 : 1);
    uint8_t *syms = (uint8_t *)malloc(sb);
    for (int64_t i = 0; i < n; i++) {
      uint32_t s = codes[i] == 0 ? 0 : (uint32_t)((int64_t)codes[i] - R + span + 1);
      if (wide)
        ((uint16_t *)syms)[i] = (uint16_t)s;
      else
        syms[i] = (uint8_t)s;
    }
    size_t bound = ZSTD_compressBound(sb);
    uint8_t *tmp = (uint8_t *)malloc(bound);
    size_t scout = ZSTD_compress(tmp, bound, syms, sb, 1);
    if (!ZSTD_isError(scout) && (double)scout * 0.85 < (double)bits / 8.0) {
      size_t full = ZSTD_compress(tmp, bound, syms, sb, zstd_level);
      if (!ZSTD_isError(full) && (double)full * 8.0 < (double)bits) {
        if ((int64_t)full > blob_cap) {
          free(tmp);
          free(syms);
          return -1;
        }
        memcpy(blob, tmp, full);
        hlen = (int64_t)full;
        literal = 1;
      }
    }
    free(tmp);
    free(syms);
  }
  if (hlen < 0) {
    uint32_t *cdfs = (uint32_t *)malloc(sizeof(uint32_t) * N_LEVELS * (A + 1));
    orc_build_cdfs(kexp, shape, precision, owx, cdfs);
    uint32_t *syms = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int64_t i = 0; i < n; i++)
      syms[i] = codes[i] == 0 ? 0 : (uint32_t)((int64_t)codes[i] - R + span + 1);
    int64_t nw = orc_rans_encode(syms, n, cdfs + (int64_t)level * (A + 1), A,
                                 (uint32_t *)blob, blob_cap / 4);
    free(syms);
    free(cdfs);
    if (nw < 0) return -1;
    hlen = nw * 4;
  }
  uint64_t framed = (uint64_t)hlen | ((uint64_t)kexp << 45) |
                    ((uint64_t)shape << 49) | ((uint64_t)level << 54) |
                    ((uint64_t)owx << 60) | ((uint64_t)literal << 63);
  put_u32(out, (uint32_t)n);
  put_u64(out + 4, framed);
  uint8_t *p = out + 12 + hlen;
  put_u32(p, (uint32_t)n_out);
  memcpy(p + 4, outl, 4 * n_out);
  return 12 + hlen + 4 + 4 * n_out;
}

/* unpack_stage (const path): returns bytes consumed or -1 (error code in
 * *err: 1 no window, 2 count mismatch, 3 misaligned, 4 trailing, 5 other). */
int64_t orc_unpack_stage(const uint8_t *buf, int64_t len, int64_t R,
                         int precision, uint32_t *codes, int64_t n_expect,
                         float *outl, int64_t out_cap, int64_t *n_out,
                         int *err) {
  *err = 0;
  if (len < 12) {
    *err = 5;
    return -1;
  }
  uint32_t n;
  uint64_t framed;
  memcpy(&n, buf, 4);
  memcpy(&framed, buf + 4, 8);
  int64_t hlen = (int64_t)(framed & (((uint64_t)1 << 45) - 1));
  int kexp = (int)((framed >> 45) & 15), shape = (int)((framed >> 49) & 31);
  int level = (int)((framed >> 54) & 63), owx = (int)((framed >> 60) & 7);
  int literal = (int)(framed >> 63);
  if (!kexp) {
    *err = 1;
    return -1;
  }
  if ((int64_t)n != n_expect) {
    *err = 2;
    return -1;
  }
  const uint8_t *blob = buf + 12;
  int64_t A = (int64_t)2 << kexp, span = ((int64_t)1 << kexp) - 1;
  if (literal) {
    int wide = kexp > 7;
    size_t sb = (size_t)n * (wide ? 2 : 1);
    uint8_t *syms = (uint8_t *)malloc(sb ? sb : 1);
    size_t got = ZSTD_decompress(syms, sb, blob, hlen);
    if (ZSTD_isError(got) || got < sb) {
      free(syms);
      *err = 5;
      return -1;
    }
    for (int64_t i = 0; i < n; i++) {
      uint32_t s = wide ? ((uint16_t *)syms)[i] : syms[i];
      codes[i] = s == 0 ? 0 : (uint32_t)((int64_t)s - span - 1 + R);
    }
    free(syms);
  } else if (n == 0) {
    if (hlen) {
      *err = 4;
      return -1;
    }
  } else {
    if (hlen % 4) {
      *err = 3;
      return -1;
    }
    uint32_t *cdfs = (uint32_t *)malloc(sizeof(uint32_t) * N_LEVELS * (A + 1));
    orc_build_cdfs(kexp, shape, precision, owx, cdfs);
    int rc = orc_rans_decode((const uint32_t *)blob, hlen / 4,
                             cdfs + (int64_t)level * (A + 1), A, codes, n);
    free(cdfs);
    if (rc) {
      *err = 4;
      return -1;
    }
    for (int64_t i = 0; i < n; i++)
      codes[i] = codes[i] == 0 ? 0 : (uint32_t)((int64_t)codes[i] - span - 1 + R);
  }
  const uint8_t *p = blob + hlen;
  uint32_t no;
  memcpy(&no, p, 4);
  if ((int64_t)no > out_cap) {
    *err = 5;
    return -1;
  }
  memcpy(outl, p + 4, 4 * (size_t)no);
  *n_out = no;
  return 12 + hlen + 4 + 4 * (int64_t)no;
}

/* ------------------------------------------------------------------------ */
/* Whole plain codec: latec/codec.py:384-510 (_compress/_decompress_region)   */
/* ------------------------------------------------------------------------ */

/* Encode the (C, *shape) field.  Writes the concatenated stage records into
 * payload (capacity cap) and the reconstruction into recon.  Returns payload
 * length or -1.  stage_bytes (optional) receives per-stage record sizes. */
int64_t orc_compress_region(const float *field, int C, const int64_t *shape,
                            int ndim, int levels, int stride, int block,
                            const double *ebs, int cubic, int center,
                            int end_mode, int64_t R, int round_mode,
                            int zstd_level, uint8_t *payload, int64_t cap,
                            float *recon, int64_t *stage_bytes) {
  int64_t N = 1;
  for (int j = 0; j < ndim; j++) N *= shape[j];
  memset(recon, 0, sizeof(float) * C * N);
  uint8_t *covered = (uint8_t *)calloc(N, 1);
  uint8_t *mask = (uint8_t *)malloc(N);
  float *vals = (float *)malloc(sizeof(float) * C * N);
  float *pred = (float *)malloc(sizeof(float) * C * N);
  uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * C * N);
  float *outl = (float *)malloc(sizeof(float) * C * N);
  float *srec = (float *)malloc(sizeof(float) * C * N);
  int ns = orc_n_stages(ndim, levels);
  int64_t off = 0;
  for (int st = 0; st < ns; st++) {
    int weight, axis;
    int s = orc_stage_mask(shape, ndim, levels, stride, block, st, covered, mask,
                           &weight, &axis);
    int64_t n = 0;
    for (int64_t p = 0; p < N; p++) n += mask[p];
    if (n == 0) {
      int64_t r = orc_pack_stage(codes, 0, outl, 0, R, zstd_level, PRECISION,
                                 payload + off, cap - off);
      if (r < 0) goto fail;
      if (stage_bytes) stage_bytes[st] = r;
      off += r;
      continue;
    }
    if (st == 0)
      memset(pred, 0, sizeof(float) * C * n);
    else
      orc_predict(recon, C, shape, ndim, mask, s, weight, axis, cubic, center,
                  end_mode, pred, n);
    int64_t at = 0;
    for (int64_t p = 0; p < N; p++)
      if (mask[p]) {
        for (int c = 0; c < C; c++) vals[c * n + at] = field[c * N + p];
        at++;
      }
    int64_t no = orc_quantize(vals, pred, C * n, ebs[st], R, round_mode, 0.0,
                              0.0, codes, outl);
    orc_dequantize(pred, codes, outl, no, C * n, ebs[st], R, srec);
    at = 0;
    for (int64_t p = 0; p < N; p++)
      if (mask[p]) {
        for (int c = 0; c < C; c++) recon[c * N + p] = srec[c * n + at];
        at++;
      }
    int64_t r = orc_pack_stage(codes, C * n, outl, no, R, zstd_level,
                               PRECISION, payload + off, cap - off);
    if (r < 0) goto fail;
    if (stage_bytes) stage_bytes[st] = r;
    off += r;
  }
  free(covered); free(mask); free(vals); free(pred); free(codes); free(outl);
  free(srec);
  return off;
fail:
  free(covered); free(mask); free(vals); free(pred); free(codes); free(outl);
  free(srec);
  return -1;
}

/* Returns 0 or a negative error (-1 - stage index on a bad record). */
int orc_decompress_region(const uint8_t *payload, int64_t len, int C,
                          const int64_t *shape, int ndim, int levels,
                          int stride, int block, const double *ebs, int cubic,
                          int center, int end_mode, int64_t R, float *recon) {
  int64_t N = 1;
  for (int j = 0; j < ndim; j++) N *= shape[j];
  memset(recon, 0, sizeof(float) * C * N);
  uint8_t *covered = (uint8_t *)calloc(N, 1);
  uint8_t *mask = (uint8_t *)malloc(N);
  float *pred = (float *)malloc(sizeof(float) * C * N);
  uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * C * N);
  float *outl = (float *)malloc(sizeof(float) * C * N);
  float *srec = (float *)malloc(sizeof(float) * C * N);
  int ns = orc_n_stages(ndim, levels), rc = 0;
  int64_t off = 0;
  for (int st = 0; st < ns && !rc; st++) {
    int weight, axis, err;
    int s = orc_stage_mask(shape, ndim, levels, stride, block, st, covered, mask,
                           &weight, &axis);
    int64_t n = 0, no = 0;
    for (int64_t p = 0; p < N; p++) n += mask[p];
    if (n && st) {
      orc_predict(recon, C, shape, ndim, mask, s, weight, axis, cubic, center,
                  end_mode, pred, n);
    } else {
      memset(pred, 0, sizeof(float) * C * (n ? n : 1));
    }
    int64_t r = orc_unpack_stage(payload + off, len - off, R, PRECISION, codes,
                                 C * n, outl, C * n, &no, &err);
    if (r < 0) {
      rc = -1 - st;
      break;
    }
    off += r;
    if (!n) continue;
    if (orc_dequantize(pred, codes, outl, no, C * n, ebs[st], R, srec)) {
      rc = -1 - st;
      break;
    }
    int64_t at = 0;
    for (int64_t p = 0; p < N; p++)
      if (mask[p]) {
        for (int c = 0; c < C; c++) recon[c * N + p] = srec[c * n + at];
        at++;
      }
  }
  free(covered); free(mask); free(pred); free(codes); free(outl); free(srec);
  return rc;
}

/* zstd helpers for the outer frame (bitstream.py:155-164). */
int64_t orc_zstd_compress(const uint8_t *src, int64_t n, int level,
                          uint8_t *dst, int64_t cap) {
  size_t r = ZSTD_compress(dst, cap, src, n, level);
  return ZSTD_isError(r) ? -1 : (int64_t)r;
}
int64_t orc_zstd_bound(int64_t n) { return (int64_t)ZSTD_compressBound(n); }
int64_t orc_zstd_content_size(const uint8_t *src, int64_t n) {
  unsigned long long r = ZSTD_getFrameContentSize(src, n);
  return (r == ZSTD_CONTENTSIZE_ERROR || r == ZSTD_CONTENTSIZE_UNKNOWN) ? -1 : (int64_t)r;
}
int64_t orc_zstd_decompress(const uint8_t *src, int64_t n, uint8_t *dst,
                            int64_t cap) {
  size_t r = ZSTD_decompress(dst, cap, src, n);
  return ZSTD_isError(r) ? -1 : (int64_t)r;
}

/* ------------------------------------------------------------------------ */
/* Multi-lane rANS stage blobs (LATEC001 version 5).  Version 5 is this      */
/* project's own stream version beside v4 (no reference counterpart); this   */
/* is its CPU specification, written independently of the GPU kernels:      */
/*   lanes L = clamp(ceil(n / 4096), 1, 65536); symbol i -> lane i % L;      */
/*   per lane a range-ANS coder (Duda 2013) with 64-bit state in            */
/*   [2^31, 2^63), 32-bit words, and the stage's CDF row (total 2^24,       */
/*   every symbol >= 1) as the model;                                        */
/*   blob = [L][L + 1 word offsets][lane words], a lane's words being its   */
/*   final state (low, high) then its words in decoder order.               */
/* ------------------------------------------------------------------------ */
int64_t orc_mlrans_lanes(int64_t n) {
  if (n <= 0) return 0;
  int64_t L = (n + 4095) / 4096;
  return L > 65536 ? 65536 : L;
}

/* Returns the blob length in words, or -1 (capacity / symbol outside A). */
int64_t orc_mlrans_encode(const uint32_t *syms, int64_t n, const uint32_t *cdf, int64_t A,
                          uint32_t *out, int64_t cap) {
  if (n <= 0) return 0;
  const int64_t L = orc_mlrans_lanes(n);
  if (cap < 2 + L) return -1;
  out[0] = (uint32_t)L;
  uint32_t *offs = out + 1;
  uint32_t *words = out + 2 + L;
  int64_t room = cap - 2 - L;
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)((n + L - 1) / L + 1));
  int64_t at = 0;
  for (int64_t l = 0; l < L; l++) {
    offs[l] = (uint32_t)at;
    int64_t cnt = n > l ? (n - l + L - 1) / L : 0;
    uint64_t x = 1ull << 31;
    int64_t k = 0;
    for (int64_t t = cnt - 1; t >= 0; t--) {
      uint32_t s = syms[l + t * L];
      if ((int64_t)s >= A) {
        free(tmp);
        return -1;
      }
      uint64_t c = cdf[s], f = cdf[s + 1] - cdf[s];
      if (x >= (f << 39)) {
        tmp[k++] = (uint32_t)x;
        x >>= 32;
      }
      x = ((x / f) << 24) + (x % f) + c;
    }
    if (at + 2 + k > room) {
      free(tmp);
      return -1;
    }
    words[at++] = (uint32_t)x;
    words[at++] = (uint32_t)(x >> 32);
    for (int64_t j = k - 1; j >= 0; j--) words[at++] = tmp[j];
  }
  offs[L] = (uint32_t)at;
  free(tmp);
  return 2 + L + at;
}

/* Returns 0, or -1 for a malformed blob. */
int orc_mlrans_decode(const uint32_t *blob, int64_t nw, int64_t n, const uint32_t *cdf,
                      int64_t A, uint32_t *syms) {
  if (n <= 0) return nw == 0 ? 0 : -1;
  const int64_t L = orc_mlrans_lanes(n);
  if (nw < 2 + L || blob[0] != (uint32_t)L) return -1;
  const uint32_t *offs = blob + 1, *words = blob + 2 + L;
  if (offs[0] != 0 || 2 + L + (int64_t)offs[L] != nw) return -1;
  for (int64_t l = 0; l < L; l++) {
    if (offs[l + 1] < offs[l] + 2) return -1;
    const uint32_t *w = words + offs[l];
    int64_t m = offs[l + 1] - offs[l], k = 2;
    uint64_t x = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    int64_t cnt = n > l ? (n - l + L - 1) / L : 0;
    for (int64_t t = 0; t < cnt; t++) {
      uint32_t slot = (uint32_t)(x & ((1u << 24) - 1));
      int64_t lo = 0, hi = A - 1; /* largest s with cdf[s] <= slot */
      while (lo < hi) {
        int64_t mid = (lo + hi + 1) / 2;
        if (cdf[mid] <= slot) lo = mid; else hi = mid - 1;
      }
      uint64_t c = cdf[lo], f = cdf[lo + 1] - cdf[lo];
      x = f * (x >> 24) + (slot - c);
      if (x < (1ull << 31)) {
        if (k >= m) return -1;
        x = (x << 32) | w[k++];
      }
      syms[l + t * L] = (uint32_t)lo;
    }
    if (k != m || x != (1ull << 31)) return -1;
  }
  return 0;
}
