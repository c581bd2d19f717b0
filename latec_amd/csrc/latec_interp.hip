// latec_interp.hip -- fused interpolation + quantization stage kernels (gfx950).
//
// One launch per dyadic sub-stage.  The grid is (x: work, y: parity class);
// each parity class is a box of lattice points with an affine rank into the
// stage array (levels.py:389-512), so a point's grid address and its slot in
// the stage array are both a few integer ops away from its class index.
//
// Per point and channel (FP64, no contraction -- built with
// -ffp-contract=off, every expression below is one IEEE op per operator):
//   prediction ....... predictor.py:97-156 (_interp_axis_at) along each odd
//                      axis, combined by `center` as predictor.py:286-315;
//                      cast to float32 (predictor.py:313)
//   quantization ..... quantizer.py:57-83: q = rint((x - p) / (2 eb)),
//                      recon = f32(p + 2eb * q), float32 |x - chk| <= f32(eb)
//   write-back ....... recon in place (codec.py:340-347), code into the stage
//                      array, outliers appended as (rank, value), histogram.
//
// Masked even-axis terms of the reference's centre-0 sum (v * 0.0) are not
// evaluated: with anchor_block == 1 they read only not-yet-revealed points,
// which hold 0, so they add a signed zero, and a signed-zero prediction
// yields the same code and reconstruction (q = 0 => recon = f32(p + 0.0)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "latec_level.h"
#include "latec_plan.h"
#include "latec_sheet.h"

namespace latec {

struct DeviceBufs {
  ClassDesc *classes = nullptr;     // all stages' classes, concatenated
  uint32_t *excluded = nullptr;     // all stages' excluded ranks
  std::vector<int64_t> class_off;   // per stage: index of first class
  std::vector<int64_t> excl_off;    // per stage: index of first excluded rank
  int device = -1;
  // cached graphs
  struct GraphKey {
    const void *a[8];
    bool operator<(const GraphKey &o) const { return memcmp(a, o.a, sizeof a) < 0; }
  };
  std::map<GraphKey, hipGraphExec_t> graphs;
  // lattice pyramid (anchor_block == 1): level k runs on the compact grid
  // x[::s_k] at stride 1; sub[k] is its one-level plan, X/R its buffers
  // (index K = the caller's full-resolution arrays).
  int K = 0;
  std::vector<latec_plan *> sub;
  std::vector<float *> X, R;
  std::vector<int64_t> npts;
};

static int hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_hip_error(what, (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

// Device copy of a plan's class tables, made on first use; callers on
// several threads may share a plan, so the first use is serialized.
static std::mutex g_upload_mu;

int plan_upload(latec_plan *p) {
  if (__atomic_load_n(&p->dev, __ATOMIC_ACQUIRE)) return LATEC_OK;
  std::lock_guard<std::mutex> lk(g_upload_mu);
  if (p->dev) return LATEC_OK;
  auto *d = new DeviceBufs();
  std::vector<ClassDesc> cls;
  std::vector<uint32_t> ex;
  for (auto &st : p->stages) {
    d->class_off.push_back((int64_t)cls.size());
    d->excl_off.push_back((int64_t)ex.size());
    cls.insert(cls.end(), st.classes.begin(), st.classes.end());
    ex.insert(ex.end(), st.excluded.begin(), st.excluded.end());
  }
  int rc = hip_check(hipGetDevice(&d->device), "hipGetDevice");
  if (!rc && !cls.empty())
    rc = hip_check(hipMalloc(&d->classes, cls.size() * sizeof(ClassDesc)), "hipMalloc");
  if (!rc && !cls.empty())
    rc = hip_check(hipMemcpy(d->classes, cls.data(), cls.size() * sizeof(ClassDesc),
                             hipMemcpyHostToDevice), "hipMemcpy");
  if (!rc && !ex.empty())
    rc = hip_check(hipMalloc(&d->excluded, ex.size() * sizeof(uint32_t)), "hipMalloc");
  if (!rc && !ex.empty())
    rc = hip_check(hipMemcpy(d->excluded, ex.data(), ex.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice), "hipMemcpy");
  if (rc) {
    if (d->classes) (void)hipFree(d->classes);
    if (d->excluded) (void)hipFree(d->excluded);
    delete d;
    return rc;
  }
  __atomic_store_n(&p->dev, d, __ATOMIC_RELEASE);
  return LATEC_OK;
}

void plan_release(latec_plan *p);

void plan_release(latec_plan *p) {
  if (!p->dev) return;
  for (auto *q : p->dev->sub)
    if (q) plan_release(q), delete q;
  for (size_t k = 0; k < p->dev->X.size(); k++) {
    if (p->dev->X[k]) (void)hipFree(p->dev->X[k]);
    if (p->dev->R[k]) (void)hipFree(p->dev->R[k]);
  }
  for (auto &kv : p->dev->graphs) (void)hipGraphExecDestroy(kv.second);
  if (p->dev->classes) (void)hipFree(p->dev->classes);
  if (p->dev->excluded) (void)hipFree(p->dev->excluded);
  delete p->dev;
  p->dev = nullptr;
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

struct GridArgs {
  int32_t ndim, channels;
  int32_t s, stage_stride, block;
  int32_t cubic, center, end_mode, round_mode;
  uint32_t R;
  int64_t N;        // points per channel
  int64_t n;        // stage points per channel
  int64_t code_base;  // C * stage offset
  int64_t shape[LATEC_MAX_DIM];
  int64_t gstr[LATEC_MAX_DIM];
  double eb, w, inv_w, span, offset;
  float feb;
  int32_t pad;
  int64_t n_excl;
  int64_t slab_lo, slab_hi;  // axis-0 grid coordinate range processed
};

__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t magic, uint32_t shift) {
  return (uint32_t)(((uint64_t)__umulhi(n, magic) + n) >> shift);
}

// 1-D interpolation along one odd axis at coordinate c (c >= s always):
// predictor.py:97-156.  `flat` and `gs` in the index type I.
template <typename I>
__device__ __forceinline__ double interp_axis(const float *__restrict__ W, I flat,
                                              int64_t c, int64_t n, I gs, int64_t s,
                                              int cubic, int end_mode) {
  const bool vp1 = c + s < n, vm3 = c >= 3 * s, vp3 = c + 3 * s < n;
  const I o1 = (I)s * gs, o3 = (I)(3 * s) * gs;
  const double B = (double)W[flat - o1];
  if (!vp1) {
    if ((end_mode & 2) && vm3) {
      const double A = (double)W[flat - o3];
      return 1.5 * B - 0.5 * A;
    }
    return B;
  }
  const double C = (double)W[flat + o1];
  if (cubic) {
    if (vm3 && vp3) {
      const double A = (double)W[flat - o3];
      const double D = (double)W[flat + o3];
      return (((-A + 9.0 * B) + 9.0 * C) - D) * 0.0625;
    }
    if (end_mode & 1) {
      if (vp3) {  // !vm3: interp_quad_1
        const double D = (double)W[flat + o3];
        return ((3.0 * B + 6.0 * C) - D) * 0.125;
      }
      if (vm3) {  // !vp3: interp_quad_2
        const double A = (double)W[flat - o3];
        return ((-A + 6.0 * B) + 3.0 * C) * 0.125;
      }
    }
  }
  return 0.5 * (B + C);
}

// Centre combine over the class's odd axes (predictor.py:286-315), cast to
// float32.  ND > 0: compile-time rank; ND == 0: runtime g.ndim.
template <int ND, typename I>
__device__ __forceinline__ float predict_point(const float *__restrict__ W,
                                               const GridArgs &g, int parity,
                                               I flat, const int64_t *coord) {
  double total = 0.0, v_first = 0.0, v_last = 0.0;
  int count = 0;
  constexpr int NMAX = ND > 0 ? ND : LATEC_MAX_DIM;
#pragma unroll
  for (int a = 0; a < NMAX; a++) {
    if (ND == 0 && a >= g.ndim) continue;
    if (!((parity >> a) & 1)) continue;
    const double v = interp_axis<I>(W, flat, coord[a], g.shape[a], (I)g.gstr[a], g.s,
                                    g.cubic, g.end_mode);
    total = count ? total + v : v;
    if (!count) v_first = v;
    v_last = v;
    count++;
  }
  double res;
  if (g.center == 0 || count == 1)
    res = count == 1 ? total : total / (double)count;
  else
    res = g.center == 1 ? v_first : v_last;
  return (float)res;
}

// Quantize one value: quantizer.py:57-83 + _recon_from_codes :114-120.
// q = rint(fl((x - p) / w)) is obtained from the product with 1/w and
// re-derived by the IEEE division whenever the product lies within 1e-12
// (relative) of a rounding boundary, so the code is always the division's.
__device__ __forceinline__ uint32_t quantize_one(float x, float p, const GridArgs &g,
                                                 float *rec) {
  const double d = (double)x - (double)p;
  const double t = d * g.inv_w;
  double qd = rint(t);
  const double dist = 0.5 - fabs(t - qd);
  if (!(dist > 1e-12 * fmax(fabs(t), 1.0))) qd = rint(d / g.w);
  const bool inr = fabs(qd) < (double)g.R;
  uint32_t code = 0;
  float r = x;
  if (inr) {
    const int q = (int)qd;
    code = (uint32_t)(q + (int)g.R);
    const float cand = (float)((double)p + g.w * (double)q);
    float chk = cand;
    if (g.round_mode == 1)
      chk = rintf(cand);
    else if (g.round_mode == 2)
      chk = (float)((rint((double)cand * g.span + g.offset) - g.offset) / g.span);
    if (fabsf(x - chk) <= g.feb)
      r = cand;
    else
      code = 0;
  }
  *rec = r;
  return code;
}

// (histogram windows, level stencils, quantizer and emission helpers: latec_level.h)

template <typename CodeT>
__device__ __forceinline__ void emit(CodeT *__restrict__ codes, int64_t slot,
                                     uint32_t code, float x, uint32_t rank,
                                     uint32_t *__restrict__ out_rank,
                                     float *__restrict__ out_val,
                                     uint32_t *__restrict__ out_count,
                                     uint32_t *lhist, uint32_t *__restrict__ ghist,
                                     uint32_t win_lo, bool active) {
  if (active) {
    codes[slot] = (CodeT)code;
    const uint32_t b = code - win_lo;
    if (code != 0 && b < HIST_WIN)
      lds_add(&lhist[b], 1u);
    else
      gbl_add(&ghist[code], 1u);
  }
  // wave-aggregated outlier append
  const bool out = active && code == 0;
  const uint64_t m = __ballot(out);
  if (m) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(out_count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (out) {
      const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      out_rank[k] = rank;
      out_val[k] = x;
    }
  }
}

// Row walker shared by the encode / decode / predict kernels.  Threads of a
// 256-thread block are split into 256/tr rows of tr lanes; a row is one run
// of the class along the last axis, whose stage ranks are consecutive.  The
// prefix (all but the last axis) is decomposed once per row.
//   ND    compile-time rank (0 = runtime)
//   I     grid index type (uint32_t when C*N < 2^32, else int64_t)
//   BLK1  anchor_block == 1 (no anchor exclusion, no anchor-block coords)
template <int ND, typename I, bool BLK1, typename F>
__device__ __forceinline__ void walk_rows(const ClassDesc &c, const GridArgs &g,
                                          const uint32_t *__restrict__ excl, F &&fn) {
  constexpr int NMAX = ND > 0 ? ND : LATEC_MAX_DIM;
  const int nd = ND > 0 ? ND : g.ndim;
  const int tr = c.tr, rb = 256 / tr;
  const int lrow = threadIdx.x / tr, lcol = threadIdx.x % tr;
  const uint32_t L = c.row_len;
  const int64_t start_last = c.starts[nd - 1];
  // rows of this class whose axis-0 coordinate lies in [slab_lo, slab_hi)
  uint32_t r0 = 0, nrows = c.nrows;
  if (nd > 1 && (g.slab_lo > 0 || g.slab_hi < g.shape[0])) {
    const int64_t st = c.starts[0], stp = c.step, n0 = c.sizes[0];
    int64_t m_lo = g.slab_lo <= st ? 0 : (g.slab_lo - st + stp - 1) / stp;
    int64_t m_hi = g.slab_hi <= st ? 0 : (g.slab_hi - st + stp - 1) / stp;
    m_lo = m_lo < n0 ? m_lo : n0;
    m_hi = m_hi < n0 ? m_hi : n0;
    const uint32_t rpp = c.nrows / (uint32_t)n0;
    r0 = (uint32_t)m_lo * rpp;
    nrows = m_hi > m_lo ? (uint32_t)(m_hi - m_lo) * rpp : 0;
  }
  // XCD-aware order: the hardware deals blocks round-robin to the 8 XCDs, so
  // block b runs on XCD b % 8; renumbering makes each XCD walk one contiguous
  // band of rows, whose shared tap lines then hit in that XCD's L2.
  const uint32_t G = gridDim.x;
  const uint32_t lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  for (uint32_t base = lb * rb; base < nrows; base += G * rb) {
    const uint32_t row = r0 + base + lrow;
    const bool rv = base + lrow < nrows;
    int64_t coord[NMAX];
    I fbase = 0;
    int64_t rbase = c.offset;
    bool anc_pref = true;
    uint32_t idx = rv ? row : 0;
#pragma unroll
    for (int j = NMAX - 2; j >= 0; j--) {
      if (ND == 0 && j > nd - 2) continue;
      const uint32_t q = fdiv(idx, c.magic[j], c.shift[j]);
      const uint32_t m = idx - q * c.sizes[j];
      idx = q;
      int64_t x;
      if (!BLK1 && (c.flags & 1))
        x = (int64_t)(m / (uint32_t)g.block) * g.stage_stride + (m % (uint32_t)g.block);
      else
        x = c.starts[j] + (int64_t)c.step * m;
      coord[j] = x;
      fbase += (I)x * (I)g.gstr[j];
      rbase += (int64_t)m * c.rstr[j];
      if (!BLK1) anc_pref = anc_pref && ((x % g.stage_stride) < g.block);
    }
    for (uint32_t col0 = 0; col0 < L; col0 += tr) {
      const uint32_t col = col0 + lcol;
      bool active = rv && col < L;
      int64_t xl;
      if (!BLK1 && (c.flags & 1))
        xl = (int64_t)(col / (uint32_t)g.block) * g.stage_stride + (col % (uint32_t)g.block);
      else
        xl = start_last + (int64_t)c.step * col;
      coord[nd - 1] = xl;
      const I flat = fbase + (I)xl;
      int64_t rank = rbase + col;
      if (!BLK1 && g.n_excl) {
        if (anc_pref && (xl % g.stage_stride) < g.block) active = false;
        int64_t lo = 0, hi = g.n_excl;  // rebase past the excluded anchors
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if ((int64_t)excl[mid] < rank) lo = mid + 1; else hi = mid;
        }
        rank -= lo;
      }
      fn(active, flat, rank, (const int64_t *)coord);
    }
  }
}

template <int ND, typename I, bool BLK1, typename CodeT>
__global__ void __launch_bounds__(256) encode_stage_kernel(
    GridArgs g, const ClassDesc *__restrict__ classes,
    const uint32_t *__restrict__ excl, const float *__restrict__ x,
    float *__restrict__ recon, CodeT *__restrict__ codes,
    uint32_t *__restrict__ ghist, uint32_t *__restrict__ out_rank,
    float *__restrict__ out_val, uint32_t *__restrict__ out_count, int anchors) {
  __shared__ uint32_t lhist[HIST_WIN];
  const ClassDesc &c = classes[blockIdx.y];
  const uint32_t win_lo = g.R > HIST_WIN / 2 ? g.R - HIST_WIN / 2 : 0;
  for (int i = threadIdx.x; i < HIST_WIN; i += blockDim.x) lhist[i] = 0;
  __syncthreads();
  const int parity = c.parity;
  walk_rows<ND, I, BLK1>(c, g, excl, [&](bool active, I flat, int64_t rank,
                                          const int64_t *coord) {
    for (int ch = 0; ch < g.channels; ch++) {
      uint32_t code = 0;
      float xv = 0.f;
      const I cofs = (I)ch * (I)g.N;
      if (active) {
        const float p = anchors ? 0.0f
                                : predict_point<ND, I>(recon + cofs, g, parity, flat, coord);
        xv = x[cofs + flat];
        float rec;
        code = quantize_one(xv, p, g, &rec);
        recon[cofs + flat] = rec;
      }
      const int64_t srank = (int64_t)ch * g.n + rank;
      emit<CodeT>(codes, g.code_base + srank, code, xv, (uint32_t)srank, out_rank, out_val,
                  out_count, lhist, ghist, win_lo, active);
    }
  });
  __syncthreads();
  flush_hist(lhist, ghist, g.R);
}

// Decode: predict, then recon = f32(p + 2eb q) or the outlier value found by
// binary search of the stage's sorted outlier ranks.
template <int ND, typename I, bool BLK1, typename CodeT>
__global__ void __launch_bounds__(256) decode_stage_kernel(
    GridArgs g, const ClassDesc *__restrict__ classes,
    const uint32_t *__restrict__ excl, const CodeT *__restrict__ codes,
    const uint32_t *__restrict__ out_rank, const float *__restrict__ out_val,
    const uint32_t *__restrict__ out_count, float *__restrict__ recon, int anchors) {
  const ClassDesc &c = classes[blockIdx.y];
  const int parity = c.parity;
  walk_rows<ND, I, BLK1>(c, g, excl, [&](bool active, I flat, int64_t rank,
                                          const int64_t *coord) {
    if (!active) return;
    for (int ch = 0; ch < g.channels; ch++) {
      const I cofs = (I)ch * (I)g.N;
      const float p = anchors ? 0.0f
                              : predict_point<ND, I>(recon + cofs, g, parity, flat, coord);
      const int64_t srank = (int64_t)ch * g.n + rank;
      const uint32_t code = (uint32_t)codes[g.code_base + srank];
      float r;
      if (code) {
        r = (float)((double)p + g.w * (double)((int)code - (int)g.R));
      } else {
        const int64_t cnt = *out_count;
        int64_t lo = 0, hi = cnt;
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if ((int64_t)out_rank[mid] < srank) lo = mid + 1; else hi = mid;
        }
        r = (lo < cnt && (int64_t)out_rank[lo] == srank) ? out_val[lo]
                                                         : __int_as_float(0x7fc00000);
      }
      recon[cofs + flat] = r;
    }
  });
}

// ---------------------------------------------------------------------------
// Level-fused 3-D tile kernel (one dyadic level of the lattice pyramid, all
// five sub-stages in one launch).
//
// On a level's compact grid (stride 1) the coarse points have even
// coordinates on every axis; the new points form 7 parity classes: three
// weight-1 classes (sub-stages 0..2, interpolated along their odd axis from
// coarse points), three weight-2 classes (sub-stage 3, from weight-1 points)
// and the weight-3 class (sub-stage 4, from weight-2 points).  A workgroup
// owns a tile B of the grid; its LDS holds the recon window B +- 3.  Every
// tap of an owned point lies within 3 cells of B along one axis and each
// step of the dependency chain displaces along a different axis, so a class
// is computed over B extended by 2 cells (its own even coordinates) along its
// even axes: weight-1 classes over two extended axes, weight-2 over one,
// weight-3 over B.  Points computed outside B (halo) are recomputed
// identically by their owner; only owned points are written to HBM.  The
// arithmetic is the per-stage kernels' (interp_axis / predict_point /
// quantize_one), so codes, outliers and recon are bit-identical.
// ---------------------------------------------------------------------------

struct L3Class {
  int64_t offset;
  int64_t start[3], rstr[3];
  int32_t step, pad;
};
struct Level3Args {
  int64_t n[3];
  int32_t cubic, center, end_mode, round_mode;
  double span, offset;
  int32_t T[3], tiles_total;
  int64_t tiles[3];
  int64_t hist_words;
  L3Stage st[5];
  L3Class cl[8];  // by parity 1..7
};

#define L3_HBINS 256
#define L3_THREADS 512

__device__ __forceinline__ int l3_stage(int parity) {
  return parity == 1 ? 0 : parity == 2 ? 1 : parity == 4 ? 2 : parity == 7 ? 4 : 3;
}

// The seven classes in dependency order (weight 1, 2, 3).
__constant__ int l3_order[7] = {1, 2, 4, 3, 5, 6, 7};

// Region of class P computed for the tile [blo, bhi): odd axes over the tile,
// even axes over the tile +- 2, clipped to the grid; first coordinate and count.
__device__ __forceinline__ void l3_region(int P, const int64_t *blo, const int64_t *bhi,
                                          const int64_t *n, int64_t *first, int *cnt) {
  for (int j = 0; j < 3; j++) {
    const bool odd = (P >> j) & 1;
    int64_t lo = odd ? blo[j] : blo[j] - 2;
    int64_t hi = odd ? bhi[j] - 1 : bhi[j] + 2;  // inclusive
    if (lo < 0) lo = 0;
    if (hi > n[j] - 1) hi = n[j] - 1;
    if ((lo & 1) != (odd ? 1 : 0)) lo++;
    first[j] = lo;
    cnt[j] = hi >= lo ? (int)((hi - lo) / 2 + 1) : 0;
  }
}

// Per tile: (1) one bulk phase loads the coarse window and, for every point
// the tile computes, its input (x to encode, the code to decode) into LDS,
// with all loads of a thread independent; (2) three compute phases (weight
// 1, 2, 3) read only LDS and write owned results to HBM.
template <bool ENC>
__global__ void __launch_bounds__(L3_THREADS) level3_kernel(
    Level3Args A, const float *__restrict__ x, float *__restrict__ recon,
    uint16_t *__restrict__ codes, uint32_t *__restrict__ hist, uint32_t *__restrict__ out_rank,
    float *__restrict__ out_val, uint32_t *__restrict__ out_count) {
  extern __shared__ __align__(16) unsigned char l3_smem[];
  uint32_t *lhist = (uint32_t *)l3_smem;  // 5 x L3_HBINS
  const int W0 = A.T[0] + 6, W1 = A.T[1] + 6, W2 = A.T[2] + 6;
  float *win = (float *)(l3_smem + 5 * L3_HBINS * sizeof(uint32_t));
  float *xs = win + W0 * W1 * W2;  // per computed point: x (encode) or code (decode)
  const int tid = threadIdx.x;
  const uint32_t G = gridDim.x;
  const uint32_t tb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  const int lstr0 = W1 * W2, lstr1 = W2;
  GridArgs gp;
  gp.ndim = 3;
  gp.s = 1;
  gp.cubic = A.cubic;
  gp.center = A.center;
  gp.end_mode = A.end_mode;
  for (int j = 0; j < 3; j++) gp.shape[j] = A.n[j];
  gp.gstr[0] = lstr0;
  gp.gstr[1] = lstr1;
  gp.gstr[2] = 1;
  for (uint32_t t = tb; t < (uint32_t)A.tiles_total; t += G) {
    const int64_t t2 = t % A.tiles[2], t01 = t / A.tiles[2];
    const int64_t t1 = t01 % A.tiles[1], t0 = t01 / A.tiles[1];
    int64_t blo[3], bhi[3], wlo[3];
    const int64_t tt[3] = {t0, t1, t2};
    for (int j = 0; j < 3; j++) {
      blo[j] = tt[j] * A.T[j];
      bhi[j] = blo[j] + A.T[j] < A.n[j] ? blo[j] + A.T[j] : A.n[j];
      wlo[j] = blo[j] - 3;
    }
    for (int i = tid; i < 5 * L3_HBINS; i += L3_THREADS) lhist[i] = 0;
    // --- bulk load: coarse cells of the window (even coordinates, tile +- 2)
    {
      int64_t c0lo[3];
      int cnt[3];
      for (int j = 0; j < 3; j++) {
        int64_t lo = blo[j] - 2 < 0 ? 0 : blo[j] - 2;
        int64_t hi = bhi[j] + 2 < A.n[j] - 1 ? bhi[j] + 2 : A.n[j] - 1;
        lo += lo & 1;
        c0lo[j] = lo;
        cnt[j] = hi >= lo ? (int)((hi - lo) / 2 + 1) : 0;
      }
      const int total = cnt[0] * cnt[1] * cnt[2];
      for (int i = tid; i < total; i += L3_THREADS) {
        const int i2 = i % cnt[2], i01 = i / cnt[2];
        const int i1 = i01 % cnt[1], i0 = i01 / cnt[1];
        const int64_t c0 = c0lo[0] + 2 * i0, c1 = c0lo[1] + 2 * i1, c2 = c0lo[2] + 2 * i2;
        win[(c0 - wlo[0]) * lstr0 + (c1 - wlo[1]) * lstr1 + (c2 - wlo[2])] =
            recon[(uint32_t)((c0 * A.n[1] + c1) * A.n[2] + c2)];
      }
    }
    // --- bulk load: the input of every point the tile computes
    {
      int off = 0;
      for (int oi = 0; oi < 7; oi++) {
        const int P = l3_order[oi];
        const L3Stage &S = A.st[l3_stage(P)];
        const L3Class &K = A.cl[P];
        int64_t first[3];
        int cnt[3];
        l3_region(P, blo, bhi, A.n, first, cnt);
        const int total = cnt[0] * cnt[1] * cnt[2];
        for (int i = tid; i < total; i += L3_THREADS) {
          const int i2 = i % cnt[2], i01 = i / cnt[2];
          const int i1 = i01 % cnt[1], i0 = i01 / cnt[1];
          const int64_t c0 = first[0] + 2 * i0, c1 = first[1] + 2 * i1, c2 = first[2] + 2 * i2;
          if (ENC) {
            xs[off + i] = x[(uint32_t)((c0 * A.n[1] + c1) * A.n[2] + c2)];
          } else {
            const int64_t rank = K.offset + ((c0 - K.start[0]) >> 1) * K.rstr[0] +
                                 ((c1 - K.start[1]) >> 1) * K.rstr[1] +
                                 ((c2 - K.start[2]) >> 1) * K.rstr[2];
            xs[off + i] = __uint_as_float((uint32_t)codes[S.code_base + rank]);
          }
        }
        off += total;
      }
    }
    __syncthreads();
    // --- compute phases
    int off = 0;
    for (int phase = 0; phase < 3; phase++) {
      const int oi0 = phase == 0 ? 0 : phase == 1 ? 3 : 6;
      const int oi1 = phase == 0 ? 3 : phase == 1 ? 6 : 7;
      for (int oi = oi0; oi < oi1; oi++) {
        const int P = l3_order[oi];
        const int si = l3_stage(P);
        const L3Stage &S = A.st[si];
        const L3Class &K = A.cl[P];
        int64_t first[3];
        int cnt[3];
        l3_region(P, blo, bhi, A.n, first, cnt);
        const int total = cnt[0] * cnt[1] * cnt[2];
        GridArgs gq;
        gq.w = S.w;
        gq.inv_w = S.inv_w;
        gq.R = S.R;
        gq.round_mode = A.round_mode;
        gq.span = A.span;
        gq.offset = A.offset;
        gq.feb = S.feb;
        uint32_t *lh = lhist + si * L3_HBINS;
        const uint32_t win_lo = S.R > L3_HBINS / 2 ? S.R - L3_HBINS / 2 : 0;
        uint32_t *ghist = hist + (int64_t)S.gs * A.hist_words;
        for (int base = 0; base < total; base += L3_THREADS) {
          const int i = base + tid;
          const bool active = i < total;
          bool owned = false;
          int64_t rank = 0;
          uint32_t code = 0;
          float xv = 0.f;
          if (active) {
            const int i2 = i % cnt[2], i01 = i / cnt[2];
            const int i1 = i01 % cnt[1], i0 = i01 / cnt[1];
            int64_t coord[3];
            coord[0] = first[0] + 2 * i0;
            coord[1] = first[1] + 2 * i1;
            coord[2] = first[2] + 2 * i2;
            owned = coord[0] >= blo[0] && coord[0] < bhi[0] && coord[1] >= blo[1] &&
                    coord[1] < bhi[1] && coord[2] >= blo[2] && coord[2] < bhi[2];
            const int l = (int)((coord[0] - wlo[0]) * lstr0 + (coord[1] - wlo[1]) * lstr1 +
                                (coord[2] - wlo[2]));
            rank = K.offset + ((coord[0] - K.start[0]) >> 1) * K.rstr[0] +
                   ((coord[1] - K.start[1]) >> 1) * K.rstr[1] +
                   ((coord[2] - K.start[2]) >> 1) * K.rstr[2];
            const float p = predict_point<3, int>(win, gp, P, l, coord);
            float rec;
            if (ENC) {
              xv = xs[off + i];
              code = quantize_one(xv, p, gq, &rec);
            } else {
              code = __float_as_uint(xs[off + i]);
              if (code) {
                rec = (float)((double)p + S.w * (double)((int)code - (int)S.R));
              } else {
                const int64_t nout = out_count[S.gs];
                const uint32_t *orr = out_rank + S.out_lo;
                int64_t lo = 0, hi = nout;
                while (lo < hi) {
                  const int64_t mid = (lo + hi) >> 1;
                  if ((int64_t)orr[mid] < rank) lo = mid + 1; else hi = mid;
                }
                rec = (lo < nout && (int64_t)orr[lo] == rank) ? out_val[S.out_lo + lo]
                                                               : __int_as_float(0x7fc00000);
              }
            }
            win[l] = rec;
            if (owned) recon[(uint32_t)((coord[0] * A.n[1] + coord[1]) * A.n[2] + coord[2])] = rec;
          }
          if (ENC) {
            const bool emit_ok = active && owned;
            if (emit_ok) {
              codes[S.code_base + rank] = (uint16_t)code;
              const uint32_t b = code - win_lo;
              if (code != 0 && b < L3_HBINS)
                lds_add(&lh[b], 1u);
              else
                gbl_add(&ghist[code], 1u);
            }
            const bool out = emit_ok && code == 0;
            const uint64_t m = __ballot(out);
            if (m) {
              const int lane = __lane_id();
              const int leader = __ffsll((unsigned long long)m) - 1;
              uint32_t obase = 0;
              if (lane == leader) obase = atomicAdd(out_count + S.gs, (uint32_t)__popcll(m));
              obase = __shfl(obase, leader);
              if (out) {
                const uint32_t k = obase + (uint32_t)__popcll(m & ((1ull << lane) - 1));
                out_rank[S.out_lo + k] = (uint32_t)rank;
                out_val[S.out_lo + k] = xv;
              }
            }
          }
        }
        off += total;
      }
      __syncthreads();
    }
    if (ENC) {  // flush the stage windows into the striped copies
      for (int si = 0; si < 5; si++) {
        const L3Stage &S = A.st[si];
        const uint32_t win_lo = S.R > L3_HBINS / 2 ? S.R - L3_HBINS / 2 : 0;
        const uint32_t glo = S.R > HIST_WIN / 2 ? S.R - HIST_WIN / 2 : 0;
        uint32_t *gwin = hist + (int64_t)S.gs * A.hist_words + 2 * (int64_t)S.R +
                         (int64_t)(t % HIST_COPIES) * HIST_WIN;
        for (int i = tid; i < L3_HBINS; i += L3_THREADS) {
          const uint32_t v = lhist[si * L3_HBINS + i];
          const uint32_t code = win_lo + i;
          if (!v) continue;
          if (code >= glo && code - glo < HIST_WIN)
            atomicAdd(&gwin[code - glo], v);
          else
            atomicAdd(&hist[(int64_t)S.gs * A.hist_words + code], v);
        }
      }
    }
    __syncthreads();
  }
}

// Generic quantizer over compact arrays (foreign predictors, GNN stages).
template <typename CodeT>
__global__ void __launch_bounds__(256) quantize_kernel(
    GridArgs g, const float *__restrict__ x, const float *__restrict__ pred,
    int64_t n, CodeT *__restrict__ codes, uint32_t *__restrict__ out_rank,
    float *__restrict__ out_val, uint32_t *__restrict__ out_count,
    uint32_t *__restrict__ ghist) {
  __shared__ uint32_t lhist[HIST_WIN];
  const uint32_t win_lo = g.R > HIST_WIN / 2 ? g.R - HIST_WIN / 2 : 0;
  for (int i = threadIdx.x; i < HIST_WIN; i += blockDim.x) lhist[i] = 0;
  __syncthreads();
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += nthreads) {
    const int64_t i = base + threadIdx.x;
    const bool active = i < n;
    uint32_t code = 0;
    float xv = 0.f, rec;
    if (active) {
      xv = x[i];
      code = quantize_one(xv, pred[i], g, &rec);
    }
    emit<CodeT>(codes, i, code, xv, (uint32_t)i, out_rank, out_val, out_count, lhist, ghist,
                win_lo, active);
  }
  __syncthreads();
  flush_hist(lhist, ghist, g.R);
}

template <typename CodeT>
__global__ void __launch_bounds__(256) dequantize_kernel(
    GridArgs g, const float *__restrict__ pred, const CodeT *__restrict__ codes,
    int64_t n, const uint32_t *__restrict__ out_rank,
    const float *__restrict__ out_val, const uint32_t *__restrict__ out_count,
    float *__restrict__ recon) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t code = (uint32_t)codes[i];
    float r;
    if (code) {
      r = (float)((double)pred[i] + g.w * (double)((int)code - (int)g.R));
    } else {
      const int64_t cnt = *out_count;
      int64_t lo = 0, hi = cnt;
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)out_rank[mid] < i) lo = mid + 1; else hi = mid;
      }
      r = (lo < cnt && (int64_t)out_rank[lo] == i) ? out_val[lo] : __int_as_float(0x7fc00000);
    }
    recon[i] = r;
  }
}

// Prediction only, written in stage order (C, n): InterpPredictor.predict.
template <int ND, typename I, bool BLK1>
__global__ void __launch_bounds__(256) predict_stage_kernel(
    GridArgs g, const ClassDesc *__restrict__ classes,
    const uint32_t *__restrict__ excl, const float *__restrict__ recon,
    float *__restrict__ pred) {
  const ClassDesc &c = classes[blockIdx.y];
  const int parity = c.parity;
  walk_rows<ND, I, BLK1>(c, g, excl, [&](bool active, I flat, int64_t rank,
                                          const int64_t *coord) {
    if (!active) return;
    for (int ch = 0; ch < g.channels; ch++)
      pred[(int64_t)ch * g.n + rank] =
          predict_point<ND, I>(recon + (I)ch * (I)g.N, g, parity, flat, coord);
  });
}

// Per-stage error moments of a finished encode (codec.py:353-381): every
// point's prediction is recomputed from the final reconstruction (each point
// is written once, at its own stage, so the earlier stages' values the
// encoder predicted from are still there: the decoder's prediction, bit for
// bit) and the stage's sums are taken in float64 over float32 differences:
//   pred_sae = sum |x - p|, pred_sse = sum (x - p)^2,
//   recon_sae = sum |x - r|, recon_sse = sum (x - r)^2, recon_max = max |x - r|
// (NaN-sticky, as NumPy's max).  Each workgroup writes its five partials
// (a fixed-shape LDS tree); stats_finish_kernel combines them in a fixed order.
constexpr int NSTAT = 5;
__device__ __forceinline__ double nan_max(double a, double b) {
  return (b > a || b != b) && a == a ? b : a;
}

template <int ND, typename I, bool BLK1>
__global__ void __launch_bounds__(256) stats_stage_kernel(
    GridArgs g, const ClassDesc *__restrict__ classes, const uint32_t *__restrict__ excl,
    const float *__restrict__ x, const float *__restrict__ recon, int anchors,
    double *__restrict__ partial) {
  __shared__ double red[NSTAT][256];
  const ClassDesc &c = classes[blockIdx.y];
  const int parity = c.parity;
  double acc[NSTAT] = {0.0, 0.0, 0.0, 0.0, 0.0};
  walk_rows<ND, I, BLK1>(c, g, excl, [&](bool active, I flat, int64_t rank,
                                          const int64_t *coord) {
    if (!active) return;
    for (int ch = 0; ch < g.channels; ch++) {
      const I cofs = (I)ch * (I)g.N;
      const float p = anchors ? 0.0f
                              : predict_point<ND, I>(recon + cofs, g, parity, flat, coord);
      const float xv = x[cofs + flat], rv = recon[cofs + flat];
      const float pe = xv - p, re = fabsf(xv - rv);
      acc[0] += (double)fabsf(pe);
      acc[1] += (double)pe * (double)pe;
      acc[2] += (double)re;
      acc[3] += (double)re * (double)re;
      acc[4] = nan_max(acc[4], (double)re);
    }
  });
  const int t = threadIdx.x;
  for (int k = 0; k < NSTAT; k++) red[k][t] = acc[k];
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (t < h) {
      for (int k = 0; k < 4; k++) red[k][t] += red[k][t + h];
      red[4][t] = nan_max(red[4][t], red[4][t + h]);
    }
    __syncthreads();
  }
  if (t < NSTAT)
    partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NSTAT + t] = red[t][0];
}

// out[s * 5 + k] = the stage's partials [off[s], off[s + 1]) combined in a
// fixed order: one workgroup per stage, STATS_SPLIT threads per moment k each
// walk one contiguous run of the partials in order, then the run totals are
// added in run order.
constexpr int STATS_SPLIT = 64;
__global__ void __launch_bounds__(NSTAT * STATS_SPLIT) stats_finish_kernel(
    const double *__restrict__ partial, const int64_t *__restrict__ off,
    double *__restrict__ out) {
  __shared__ double run[NSTAT][STATS_SPLIT];
  const int st = blockIdx.x, k = threadIdx.x / STATS_SPLIT, j = threadIdx.x % STATS_SPLIT;
  const int64_t b0 = off[st], n = off[st + 1] - b0;
  const int64_t per = (n + STATS_SPLIT - 1) / STATS_SPLIT;
  const int64_t lo = b0 + j * per, hi = lo + per < b0 + n ? lo + per : b0 + n;
  double a = 0.0;
  for (int64_t b = lo; b < hi; b++) {
    const double v = partial[b * NSTAT + k];
    a = k == 4 ? nan_max(a, v) : a + v;
  }
  run[k][j] = a;
  __syncthreads();
  if (j == 0) {
    double t = 0.0;
    for (int r = 0; r < STATS_SPLIT; r++) t = k == 4 ? nan_max(t, run[k][r]) : t + run[k][r];
    out[st * NSTAT + k] = t;
  }
}

// Segment min / max of float32 values as NumPy's min() / max() give them: a
// NaN anywhere makes both NaN.  Pass 1: block b of segment g reduces its
// slice to a partial pair; pass 2 folds a segment's partials in order.
__device__ __forceinline__ float nan_min_f(float a, float b) {
  return (b < a || b != b) && a == a ? b : a;
}
__device__ __forceinline__ float nan_max_f(float a, float b) {
  return (b > a || b != b) && a == a ? b : a;
}
constexpr int MM_BLOCKS = 256;  // partial pairs per segment
__global__ void __launch_bounds__(256) minmax_kernel(const float *__restrict__ x,
                                                     const int64_t *__restrict__ seg,
                                                     float *__restrict__ part) {
  const int g = blockIdx.y;
  const int64_t lo = seg[2 * g], n = seg[2 * g + 1];
  float mn = INFINITY, mx = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)MM_BLOCKS * 256) {
    const float v = x[lo + i];
    mn = nan_min_f(mn, v);
    mx = nan_max_f(mx, v);
  }
  __shared__ float smn[256], smx[256];
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) {
      smn[threadIdx.x] = nan_min_f(smn[threadIdx.x], smn[threadIdx.x + h]);
      smx[threadIdx.x] = nan_max_f(smx[threadIdx.x], smx[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ((int64_t)g * MM_BLOCKS + blockIdx.x)] = smn[0];
    part[2 * ((int64_t)g * MM_BLOCKS + blockIdx.x) + 1] = smx[0];
  }
}
__global__ void minmax_finish_kernel(const float *__restrict__ part, int nseg,
                                     float *__restrict__ out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  float mn = INFINITY, mx = -INFINITY;
  for (int b = 0; b < MM_BLOCKS; b++) {
    mn = nan_min_f(mn, part[2 * ((int64_t)g * MM_BLOCKS + b)]);
    mx = nan_max_f(mx, part[2 * ((int64_t)g * MM_BLOCKS + b) + 1]);
  }
  out[2 * g] = mn;
  out[2 * g + 1] = mx;
}

template <typename CodeT>
__global__ void __launch_bounds__(256) code_hist_kernel(const CodeT *__restrict__ codes,
                                                        int64_t n, uint32_t R,
                                                        uint32_t *__restrict__ ghist) {
  __shared__ uint32_t lhist[HIST_WIN];
  const uint32_t win_lo = R > HIST_WIN / 2 ? R - HIST_WIN / 2 : 0;
  for (int i = threadIdx.x; i < HIST_WIN; i += blockDim.x) lhist[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t code = (uint32_t)codes[i];
    const uint32_t b = code - win_lo;
    if (code != 0 && b < HIST_WIN)
      lds_add(&lhist[b], 1u);
    else if (code < 2 * R)
      gbl_add(&ghist[code], 1u);
  }
  __syncthreads();
  flush_hist(lhist, ghist, R);
}

// ---------------------------------------------------------------------------
// Marching class kernels (3-D level grids of the lattice pyramid, one
// channel, 16-bit codes).
//
// On a level's compact grid (stride 1) parity class P holds the points whose
// coordinate is odd exactly on the axes of P.  A thread owns a column of
// MARCH_M consecutive class points along its marching axis (axis 0 when odd
// in P, else axis 1 when odd, else axis 0) and walks it, carrying the taps
// c-3, c-1, c+1, c+3 of that axis in registers: one new tap load per point
// instead of four.  Lanes run along the last axis, so a wave's x, recon and
// code accesses are row segments.  One launch takes up to three classes,
// blockIdx.y picking one: the three weight-1 stages of a level are
// independent of each other and share a launch.  P is a template argument, so
// the axis loop, the centre combine and the tap validity are compiled per
// class.  The arithmetic is interp_axis / predict_point / quantize_one's,
// operation for operation, so codes, outliers and recon are bit-identical to
// the stage kernels'.
// ---------------------------------------------------------------------------
#define MARCH_M 8  // most points per thread (MArgs::steps <= MARCH_M)

struct MClass {
  int32_t parity, stage;  // stage: index into MArgs::st
  uint32_t m[3];          // class extents
  uint32_t e[3];          // thread extents (the marching axis divided by MArgs::steps)
  uint32_t nthreads, pad;
  int64_t offset;         // stage rank of the class's first point
  int64_t rstr[3];        // stage-rank strides per class index
};

struct MArgs {
  int32_t n[3];
  int32_t cubic, center, end_mode, round_mode, ncls, peel;
  int32_t steps, pad;  // points per thread along the march (1 .. MARCH_M)
  double span, offset;
  int64_t hist_words;
  const float *coarse;   // weight-1 launches: the compact grid of the level below, where
                         // every weight-1 tap (a coarse point) is read; else nullptr
  int32_t cn[3];         // its shape
  int32_t write_coarse;  // class 100 stores each c-1 tap (a coarse point) into recon
  L3Stage st[3];
  MClass cl[3];
};


__device__ __forceinline__ void hist_add(uint32_t *lhist, uint32_t *__restrict__ ghist,
                                         uint32_t win_lo, uint32_t code, uint32_t k) {
  const uint32_t b = code - win_lo;
  if (code != 0 && b < HIST_WIN)
    lds_add(lhist + b, k);
  else
    gbl_add(ghist + code, k);
}

// emit() with the histogram counted per wave: the wave's most common codes
// (up to `peel` distinct values, found by ballot) take one atomic each, the
// remaining lanes one atomic per lane.  Stores through 32-bit byte offsets.
__device__ __forceinline__ void emit_wave(uint16_t *__restrict__ codes, uint32_t slot,
                                          uint32_t code, float x, uint32_t rank,
                                          uint32_t *__restrict__ out_rank,
                                          float *__restrict__ out_val,
                                          uint32_t *__restrict__ out_count, uint32_t *lhist,
                                          uint32_t *__restrict__ ghist, uint32_t win_lo,
                                          bool active, int peel) {
  const int lane = __lane_id();
  if (active) *(uint16_t *)((char *)codes + 2u * slot) = (uint16_t)code;
  bool todo = active;
  for (int it = 0; it < peel; it++) {  // peel is uniform: a scalar loop
    const uint64_t m = __ballot(todo);
    if (!m) break;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const uint32_t c0 = (uint32_t)__builtin_amdgcn_readlane((int)code, leader);
    const bool mine = todo && code == c0;
    const uint64_t mm = __ballot(mine);
    if (lane == leader) hist_add(lhist, ghist, win_lo, c0, (uint32_t)__popcll(mm));
    todo = todo && !mine;
  }
  if (todo) hist_add(lhist, ghist, win_lo, code, 1u);
  const bool out = active && code == 0;
  const uint64_t m = __ballot(out);
  if (m) {
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(out_count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (out) {
      const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      out_rank[k] = rank;
      out_val[k] = x;
    }
  }
}


template <int P, bool ENC>
__device__ __forceinline__ void march_class(const MArgs &A, const MClass &K,
                                            const float *__restrict__ x,
                                            float *__restrict__ rec,
                                            uint16_t *__restrict__ codes,
                                            uint32_t *__restrict__ hist, uint32_t *lhist,
                                            uint32_t *__restrict__ out_rank,
                                            float *__restrict__ out_val,
                                            uint32_t *__restrict__ out_count) {
  constexpr int M = MARCH_M;
  const uint32_t NS = (uint32_t)A.steps;
  constexpr int MA = (P & 1) ? 0 : (P & 2) ? 1 : 0;
  constexpr bool WIN = (P >> MA) & 1;
  constexpr int NODD = (P & 1) + ((P >> 1) & 1) + ((P >> 2) & 1);
  constexpr int NDIR = NODD - (WIN ? 1 : 0);  // odd axes read directly (not the window)
  // the direct axes, ascending
  constexpr int DA0 = (P & 1) && (!WIN || MA != 0) ? 0
                      : ((P >> 1) & 1) && (!WIN || MA != 1) ? 1 : 2;
  constexpr int DA1 = 2;
  constexpr bool W1C = P == 1 || P == 2 || P == 4;  // weight-1 classes
  const L3Stage &S = A.st[K.stage];
  const uint32_t gstr[3] = {(uint32_t)A.n[1] * (uint32_t)A.n[2], (uint32_t)A.n[2], 1u};
  const uint32_t sMA = gstr[MA];
  const int nMA = A.n[MA];
  const int cubic = A.cubic, end_mode = A.end_mode, center = A.center;
  const int round_mode = A.round_mode;
  const double w = S.w, inv_w = S.inv_w, Rd = (double)S.R;
  const float feb = S.feb;
  const uint32_t win_lo = S.R > HIST_WIN / 2 ? S.R - HIST_WIN / 2 : 0;
  uint32_t *ghist = ENC ? hist + (int64_t)S.gs * A.hist_words : nullptr;
  uint32_t *orank = out_rank + S.out_lo;
  float *oval = out_val + S.out_lo;
  uint32_t *ocnt = out_count + S.gs;
  const int64_t nout = ENC ? 0 : (int64_t)out_count[S.gs];
  const int peel = A.peel;
  // stage ranks and code slots fit 32 bits (the pyramid requires C*N < 2^32)
  uint16_t *scodes = codes + S.code_base;
  const uint32_t rMA = (uint32_t)K.rstr[MA];
  const uint32_t G = gridDim.x;
  const uint32_t lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  for (uint32_t base = lb * 256u; base < K.nthreads; base += G * 256u) {
    const uint32_t tid = base + threadIdx.x;
    const bool tv = tid < K.nthreads;
    uint32_t u = tv ? tid : 0u;
    uint32_t ci[3];
    ci[2] = u % K.e[2];
    u /= K.e[2];
    ci[1] = u % K.e[1];
    ci[0] = u / K.e[1];
    ci[MA] *= NS;
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; a++) c[a] = ((P >> a) & 1) + 2 * (int)ci[a];
    const uint32_t f0 = (uint32_t)c[0] * gstr[0] + (uint32_t)c[1] * gstr[1] + (uint32_t)c[2];
    const uint32_t rk0 = (uint32_t)(K.offset + (int64_t)ci[0] * K.rstr[0] +
                                    (int64_t)ci[1] * K.rstr[1] + (int64_t)ci[2] * K.rstr[2]);
    const uint32_t cnt = tv ? min(NS, K.m[MA] - ci[MA]) : 0u;
    // tap source: the level's own recon, or, for the weight-1 classes when it
    // is given, the compact grid of the level below (every weight-1 tap is a
    // coarse point there: fine c -+ 1, -+ 3 is coarse (c >> 1) + {0, 1, -1, 2}).
    // tb = the index of the point's c-1 tap along an odd axis; an invalid tap
    // reads tb itself (in bounds, value unused)
    const bool ct = W1C && A.coarse != nullptr;
    const float *tsrc = ct ? A.coarse : (const float *)rec;
    const uint32_t cs[3] = {(uint32_t)A.cn[1] * (uint32_t)A.cn[2], (uint32_t)A.cn[2], 1u};
    const uint32_t tb0 = ct ? (uint32_t)(c[0] >> 1) * cs[0] + (uint32_t)(c[1] >> 1) * cs[1] +
                                  (uint32_t)(c[2] >> 1)
                            : f0;
    const uint32_t tstep = ct ? cs[MA] : 2 * sMA;  // per point along the march
    // direct axes: validity is constant along the march
    bool dv[2][3];       // vm3, vp1, vp3
    int32_t doff[2][4];  // offsets of taps A (c-3), B (c-1), C (c+1), D (c+3) from tb
    bool dir_in = true;  // every direct axis has all four taps
#pragma unroll
    for (int j = 0; j < NDIR; j++) {
      const int a = j == 0 ? DA0 : DA1;
      const int ca = c[a], na = A.n[a];
      const int32_t sa = ct ? (int32_t)cs[a] : (int32_t)gstr[a];
      dv[j][0] = ca >= 3;
      dv[j][1] = ca + 1 < na;
      dv[j][2] = ca + 3 < na;
      dir_in = dir_in && dv[j][0] && dv[j][2];
      if (ct) {
        doff[j][0] = dv[j][0] ? -sa : 0;
        doff[j][1] = 0;
        doff[j][2] = dv[j][1] ? sa : 0;
        doff[j][3] = dv[j][2] ? 2 * sa : 0;
      } else {
        doff[j][0] = dv[j][0] ? -3 * sa : 0;
        doff[j][1] = -sa;
        doff[j][2] = dv[j][1] ? sa : 0;
        doff[j][3] = dv[j][2] ? 3 * sa : 0;
      }
    }
    // window offsets along MA from tb
    const int32_t sw = ct ? (int32_t)cs[MA] : (int32_t)sMA;
    const int32_t wo0 = ct ? -sw : -3 * sw, wo1 = ct ? 0 : -sw, wo2 = sw, wo3 = ct ? 2 * sw : 3 * sw;
    // the loads of point t: x (or its code), the window's new tap c+3, and
    // the direct taps -- all unconditional, issued one point ahead
    struct Loads {
      float xin;
      uint32_t cin;
      float w3;
      float dt[NDIR > 0 ? NDIR : 1][4];
    };
    auto load = [&](uint32_t t, Loads &L) {
      const bool has = t < cnt;
      const uint32_t f = has ? f0 + 2 * t * sMA : f0;
      const uint32_t tb = has ? tb0 + t * tstep : tb0;
      if (ENC)
        L.xin = ldo(x, f);
      else
        L.cin = ldo(scodes, has ? rk0 + t * rMA : rk0);
      if (WIN)
        L.w3 = ldo(tsrc, has && c[MA] + 2 * (int)t + 3 < nMA ? (uint32_t)((int32_t)tb + wo3) : tb);
#pragma unroll
      for (int j = 0; j < NDIR; j++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          L.dt[j][k] = ldo(tsrc, (uint32_t)((int32_t)tb + doff[j][k]));
    };
    double w0 = 0.0, w1 = 0.0, w2 = 0.0;  // window taps c-3, c-1, c+1
    if (WIN) {
      const int cm = c[MA];
      w0 = (double)ldo(tsrc, cm >= 3 ? (uint32_t)((int32_t)tb0 + wo0) : tb0);
      w1 = (double)ldo(tsrc, (uint32_t)((int32_t)tb0 + wo1));
      w2 = (double)ldo(tsrc, cm + 1 < nMA ? (uint32_t)((int32_t)tb0 + wo2) : tb0);
    }
    Loads cur, nxt;
    load(0, cur);
#pragma unroll
    for (uint32_t t = 0; t < (uint32_t)M; t++) {
      if (t >= NS) break;  // uniform
      if (t + 1 < NS) load(t + 1, nxt);
      const bool active = t < cnt;
      const uint32_t f = f0 + 2 * t * sMA;
      const uint32_t rank = rk0 + t * rMA;
      const double w3 = WIN ? (double)cur.w3 : 0.0;
      const int cm = c[MA] + 2 * (int)t;
      // --- prediction: the cubic interior stencil when the whole wave is
      // interior, else interp4's general end handling (same values)
      double v[3];
      const bool win_in = !WIN || (cm >= 3 && cm + 3 < nMA);
      if (cubic && __all(win_in && dir_in)) {
#pragma unroll
        for (int a = 0, i = 0; a < 3; a++) {
          if (!((P >> a) & 1)) continue;
          if (WIN && a == MA) {
            v[i++] = cubic4(w0, w1, w2, w3);
          } else {
            const int j = a == DA0 ? 0 : 1;
            v[i++] = cubic4((double)cur.dt[j][0], (double)cur.dt[j][1], (double)cur.dt[j][2],
                            (double)cur.dt[j][3]);
          }
        }
      } else {
#pragma unroll
        for (int a = 0, i = 0; a < 3; a++) {
          if (!((P >> a) & 1)) continue;
          if (WIN && a == MA) {
            v[i++] = interp4(w0, w1, w2, w3, cm >= 3, cm + 1 < nMA, cm + 3 < nMA, cubic, end_mode);
          } else {
            const int j = a == DA0 ? 0 : 1;
            v[i++] = interp4((double)cur.dt[j][0], (double)cur.dt[j][1], (double)cur.dt[j][2],
                             (double)cur.dt[j][3], dv[j][0], dv[j][1], dv[j][2], cubic, end_mode);
          }
        }
      }
      double res;
      if (NODD == 1) {
        res = v[0];
      } else if (center == 0) {
        double total = v[0] + v[1];
        if (NODD == 3) total = total + v[2];
        res = NODD == 3 ? div3(total) : total * 0.5;
      } else {
        res = center == 1 ? v[0] : v[NODD - 1];
      }
      const float p = (float)res;
      uint32_t code = 0;
      float xv = 0.f, r = 0.f;
      if (ENC) {
        // quantize_one, branch-free but for the rare exact-division fallback
        xv = cur.xin;
        const double pd = (double)p;
        const double d = (double)xv - pd;
        const double tq = d * inv_w;
        double qd = rint(tq);
        const bool fb = !(0.5 - fabs(tq - qd) > 1e-12 * fmax(fabs(tq), 1.0));
        if (__any(fb)) {
          if (fb) qd = rint(d / w);
        }
        const bool inr = fabs(qd) < Rd;
        const int qi = inr ? (int)qd : 0;
        const float cand = (float)(pd + w * (double)qi);
        float chk = cand;
        if (round_mode == 1)
          chk = rintf(cand);
        else if (round_mode == 2)
          chk = (float)((rint((double)cand * A.span + A.offset) - A.offset) / A.span);
        const bool ok = inr && fabsf(xv - chk) <= feb;
        code = ok ? (uint32_t)(qi + (int)S.R) : 0u;
        r = ok ? cand : xv;
      } else {
        code = cur.cin;
        r = (float)((double)p + w * (double)((int)code - (int)S.R));
        if (__any(active && code == 0)) {
          if (active && code == 0) {
            int64_t lo = 0, hi = nout;
            while (lo < hi) {
              const int64_t mid = (lo + hi) >> 1;
              if ((int64_t)orank[mid] < (int64_t)rank) lo = mid + 1; else hi = mid;
            }
            r = (lo < nout && (int64_t)orank[lo] == (int64_t)rank) ? oval[lo]
                                                                   : __int_as_float(0x7fc00000);
          }
        }
      }
      if (active) sto(rec, f, r);
      // the c-1 tap of a class-100 point is the coarse point below it: with
      // write_coarse every coarse point of the level lands in recon here
      if (P == 1 && ct && A.write_coarse && active) sto(rec, f - sMA, (float)w1);
      if (ENC)
        emit_wave(scodes, rank, code, xv, rank, orank, oval, ocnt, lhist, ghist, win_lo, active,
                  peel);
      if (WIN) {
        w0 = w1;
        w1 = w2;
        w2 = w3;
      }
      cur = nxt;
    }
  }
}

template <bool ENC>
__global__ void __launch_bounds__(256) march_kernel(MArgs A, const float *__restrict__ x,
                                                    float *__restrict__ rec,
                                                    uint16_t *__restrict__ codes,
                                                    uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ out_rank,
                                                    float *__restrict__ out_val,
                                                    uint32_t *__restrict__ out_count) {
  __shared__ uint32_t lhist[ENC ? HIST_WIN : 1];
  if (ENC) {
    for (int i = threadIdx.x; i < HIST_WIN; i += blockDim.x) lhist[i] = 0;
    __syncthreads();
  }
  const MClass &K = A.cl[blockIdx.y];
  switch (K.parity) {
#define LATEC_MARCH_CASE(PP)                                                                 \
  case PP:                                                                                   \
    march_class<PP, ENC>(A, K, x, rec, codes, hist, lhist, out_rank, out_val, out_count); \
    break;
    LATEC_MARCH_CASE(1)
    LATEC_MARCH_CASE(2)
    LATEC_MARCH_CASE(3)
    LATEC_MARCH_CASE(4)
    LATEC_MARCH_CASE(5)
    LATEC_MARCH_CASE(6)
    LATEC_MARCH_CASE(7)
#undef LATEC_MARCH_CASE
    default: break;
  }
  if (ENC) {
    __syncthreads();
    flush_hist(lhist, hist + (int64_t)A.st[K.stage].gs * A.hist_words, A.st[K.stage].R);
  }
}

// ---------------------------------------------------------------------------
// Line-pair kernels (3-D levels above the first, one channel, 16-bit codes).
//
// A level's lines along axis 2 split by the parity (i & 1, j & 1) of their
// position into four types, and each type holds two classes alternating
// along k:
//   (even, even): the coarse points (k even) and class 001 (k odd);
//   (odd, even):  class 100 and class 101;   (even, odd): class 010 and 011;
//   (odd, odd):   class 110 and class 111.
// One thread codes the pair (k = 2m, 2m + 1) of a line and marches along
// axis 0 or 1 through the lines of its type, so x is read and recon written
// as whole lines.  Three launches follow the dependencies:
//   phase 1, (e,e) lines: the coarse points (copied from the level below's
//            compact grid) and 001 (taps: coarse points along k);
//   phase 2, (o,e) lines marching along axis 0 and (e,o) lines marching
//            along axis 1: 100 / 010 (taps: the (e,e) lines' even points
//            along the march), then 101 / 011 (the (e,e) lines' odd points
//            along the march, and the pair's own even points along k);
//   phase 3, (o,o) lines marching along axis 0: 110 (taps: the (e,o) lines'
//            even points along the march, the (o,e) lines' even points
//            along axis 1), then 111 (the odd points of those lines, and
//            110 along k).
// The k-neighbours of an odd point are the even points of lanes l-1, l+1
// and l+2 at the same step.  When a line has at most 256 pairs, phases 2 and
// 3 give each workgroup whole lines (thread = pair, 256 / npairs lines per
// workgroup) and pass the even points through LDS, one barrier per step:
// every lane owns its pair and every wave stores aligned runs.  Longer lines
// use wave runs: a wave owns the 61 pairs of lanes 1..61 of a run of pairs
// numbered across the lines of the type, and its lanes 0, 62 and 63
// recompute the even points of the neighbouring runs (same inputs, same
// arithmetic, so the same bits) and store and emit nothing.  Arithmetic is
// interp4 / the stage quantizer in double, exactly as the other level
// kernels.
// ---------------------------------------------------------------------------
#define PAIR_M 8    // most march steps per thread (PArgs::steps <= PAIR_M)

#define PAIR_OWN 61  // pairs a wave owns in phases 2 and 3

struct PClass {
  int32_t stage, pad;  // index into PArgs::st
  int64_t offset;      // stage rank of the class's first point
  int64_t rstr[3];     // stage-rank strides per class index
};
struct PSet {
  int32_t ma, ra;       // marching axis and row axis (0 / 1)
  int32_t pm, pr;       // parity of the set's lines along ma and ra
  uint32_t nmarch;      // line positions along ma of that parity
  uint32_t nrows;       // line positions along ra of that parity
  uint32_t nwaves;      // wave runs: waves per march chunk (runs of pairs)
  uint32_t nthreads;    // nchunks * (nwaves * 64 or bpc * 256), nchunks = ceil(nmarch / steps)
  uint32_t lpb, bpc;    // line blocks (lpb > 0): whole lines per workgroup, workgroups per chunk
  PClass ev, od;        // the classes at k = 2m and k = 2m + 1
};
struct PArgs {
  int32_t n[3], cn[3];
  const float *coarse;  // phase 1: the level below's compact grid
  int32_t cubic, center, end_mode, round_mode, peel, pad;
  double span, offset;
  int64_t hist_words;
  uint32_t npairs, own;  // pairs per line; pairs a wave owns
  uint32_t steps, pad2;  // march steps per thread (1 .. PAIR_M): fewer on small levels
  uint32_t xs[3];        // element strides of x along the level's axes (the input
                         // grid's strides times the level's lattice step)
  uint32_t nquads;       // quads per line (line-quad launches: n2 / 4)
  L3Stage st[3];
  PSet set[2];
};


template <int PH, bool ENC, bool AL, bool RND, bool BM, bool XD>
__device__ __forceinline__ void pair_set(const PArgs &A, const PSet &T,
                                         const float *__restrict__ x, float *__restrict__ rec,
                                         uint16_t *__restrict__ codes,
                                         uint32_t *__restrict__ hist, uint32_t *lh, float *xch,
                                         uint32_t *__restrict__ out_rank,
                                         float *__restrict__ out_val,
                                         uint32_t *__restrict__ out_count) {
  constexpr int M = PAIR_M;
  const uint32_t S = A.steps;
  constexpr int NE = PH == 1 ? 0 : PH == 2 ? 1 : 2;  // odd axes of the even class
  constexpr int NO = NE + 1;                         // odd axes of the odd class
  const int ma = T.ma, ra = T.ra;
  const uint32_t gstr[3] = {(uint32_t)A.n[1] * (uint32_t)A.n[2], (uint32_t)A.n[2], 1u};
  const uint32_t sma = gstr[ma], sra = gstr[ra];
  const int n2 = A.n[2], nma = A.n[ma], nra = A.n[ra];
  constexpr bool al = AL;  // every pair starts 8-byte aligned (an even line length)
  const uint32_t xsma = A.xs[ma], xs2 = A.xs[2];
  constexpr bool xdirect = XD;
  const int cubic = A.cubic, end_mode = A.end_mode, center = A.center;
  const int round_mode = A.round_mode;
  const L3Stage &SE = A.st[T.ev.stage], &SO = A.st[T.od.stage];
  const PQ QE = make_pq(SE);
  const PQ QO = make_pq(SO);
  const uint32_t wlo_e = SE.R > HIST_WIN / 2 ? SE.R - HIST_WIN / 2 : 0;
  const uint32_t wlo_o = SO.R > HIST_WIN / 2 ? SO.R - HIST_WIN / 2 : 0;
  uint32_t *gh_e = ENC ? hist + (int64_t)SE.gs * A.hist_words : nullptr;
  uint32_t *gh_o = ENC ? hist + (int64_t)SO.gs * A.hist_words : nullptr;
  uint32_t *or_e = out_rank + SE.out_lo, *or_o = out_rank + SO.out_lo;
  float *ov_e = out_val + SE.out_lo, *ov_o = out_val + SO.out_lo;
  const int64_t nout_e = ENC ? 0 : (int64_t)out_count[SE.gs];
  const int64_t nout_o = ENC ? 0 : (int64_t)out_count[SO.gs];
  uint16_t *sc_e = codes + SE.code_base, *sc_o = codes + SO.code_base;
  // one code resource for both classes: each class's stage starts at a byte
  // offset (the code array is far below 2^31 bytes)
  const __amdgpu_buffer_rsrc_t rrec = buf_rsrc(rec), rcodes = buf_rsrc(codes);
  const uint32_t cb_e = 2u * (uint32_t)SE.code_base, cb_o = 2u * (uint32_t)SO.code_base;
  const uint32_t rma_e = (uint32_t)T.ev.rstr[ma], rma_o = (uint32_t)T.od.rstr[ma];
  const int lane = __lane_id();
  const uint32_t own = A.own, halo = PH == 1 ? 0u : 1u;
  const uint32_t total = T.nrows * A.npairs;
  const uint32_t cs0 = (uint32_t)A.cn[1] * (uint32_t)A.cn[2], cs1 = (uint32_t)A.cn[2];
  const uint32_t G = gridDim.x;
  const uint32_t lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  constexpr bool bm = PH > 1 && BM;
  uint32_t xstep = 0;  // steps taken by the workgroup: picks the exchange buffer
  for (uint32_t base = lb * 256u; base < T.nthreads; base += G * 256u) {
    bool wv, owned;
    uint32_t chunk, row, m, xpos = 0;
    if (bm) {  // nthreads is a multiple of 256: the whole workgroup is in range
      const uint32_t blk = base >> 8;
      chunk = blk / T.bpc;
      // pairs whose odd point has all four k taps (1 .. npairs-3) first, the
      // three end pairs of every line last: only the last wave meets the end
      // rules along k
      const uint32_t np = A.npairs, nin = np >= 3 ? np - 3 : 0, nb = np - nin;
      const uint32_t tl = threadIdx.x;
      uint32_t line;
      if (tl < T.lpb * nin) {
        line = tl / nin;
        m = 1 + tl - line * nin;
      } else {
        const uint32_t b = tl - T.lpb * nin;
        line = b / nb;
        const uint32_t j = b - line * nb;
        m = nin ? (j == 0 ? 0 : np - 3 + j) : j;
      }
      xpos = line < T.lpb ? line * np + m : tl;  // exchange slot (past the lines: tl)
      row = (blk - chunk * T.bpc) * T.lpb + line;
      wv = true;
      owned = line < T.lpb && row < T.nrows;
      row = min(row, T.nrows - 1);
      m = min(m, np - 1);
    } else {
      const uint32_t gw = (base + threadIdx.x) >> 6;  // uniform in the wave
      wv = base + threadIdx.x < T.nthreads;
      chunk = wv ? gw / T.nwaves : 0u;
      const uint32_t run = wv ? gw - chunk * T.nwaves : 0u;
      const int64_t P = (int64_t)run * own + lane - (int64_t)halo;
      owned = wv && lane >= (int)halo && lane < (int)(halo + own) && P < (int64_t)total;
      const uint32_t Pc = (uint32_t)min(max(P, (int64_t)0), (int64_t)total - 1);
      row = Pc / A.npairs;
      m = Pc - row * A.npairs;
    }
    // every lane of a wave (of a workgroup, with line blocks) has the same
    // chunk, so the step count is uniform
    const uint32_t cnt = wv ? min(S, T.nmarch - chunk * S) : 0u;
    const uint32_t ucnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)cnt);
    if (ucnt == 0) continue;
    int c[3];
    c[ma] = T.pm + 2 * (int)(chunk * S);
    c[ra] = T.pr + 2 * (int)row;
    c[2] = 2 * (int)m;
    const bool odd_ok = 2 * (int)m + 1 < n2;
    const uint32_t f0 = (uint32_t)c[0] * gstr[0] + (uint32_t)c[1] * gstr[1] + (uint32_t)c[2];
    // x: the level's own grid (xs == gstr) or the input grid at the level's
    // lattice step (pairs are then two strided loads)
    const uint32_t fx0 = (uint32_t)c[0] * A.xs[0] + (uint32_t)c[1] * A.xs[1] +
                         (uint32_t)c[2] * A.xs[2];
    uint32_t ci[3];
    ci[ma] = chunk * S;
    ci[ra] = row;
    ci[2] = m;
    const uint32_t rk_e0 = (uint32_t)(T.ev.offset + (int64_t)ci[0] * T.ev.rstr[0] +
                                      (int64_t)ci[1] * T.ev.rstr[1] + (int64_t)m * T.ev.rstr[2]);
    const uint32_t rk_o0 = (uint32_t)(T.od.offset + (int64_t)ci[0] * T.od.rstr[0] +
                                      (int64_t)ci[1] * T.od.rstr[1] + (int64_t)m * T.od.rstr[2]);
    // taps along k of the odd point k = 2m + 1 (lanes l-1, l, l+1, l+2)
    const bool kv0 = m >= 1, kv1 = 2 * (int)m + 2 < n2, kv2 = 2 * (int)m + 4 < n2;
    // direct taps along ra (phase 3): validity is constant along the march
    const int cr = c[ra];
    const bool dv0 = cr >= 3, dv1 = cr + 1 < nra, dv2 = cr + 3 < nra;
    int32_t doff[4] = {0, 0, 0, 0};
    if (PH == 3) {
      doff[0] = dv0 ? -3 * (int32_t)sra : -(int32_t)sra;
      doff[1] = -(int32_t)sra;
      doff[2] = dv1 ? (int32_t)sra : -(int32_t)sra;
      doff[3] = dv2 ? 3 * (int32_t)sra : -(int32_t)sra;
    }
    // phase 1: the coarse row of the pair, m-1 .. m+2 (invalid taps read m)
    const uint32_t cb0 = PH == 1 ? (uint32_t)(c[0] >> 1) * cs0 + (uint32_t)(c[1] >> 1) * cs1 + m
                                 : 0u;
    const int32_t co0 = kv0 ? -1 : 0, co2 = kv1 ? 1 : 0, co3 = kv2 ? 2 : 0;
    struct Loads {
      float2 xx;
      uint32_t ce, co;
      float2 w3;
      float2 d[4];
      float ct[4];
    };
    auto load = [&](uint32_t t, Loads &L) {
      const bool has = t < cnt;
      const uint32_t f = has ? f0 + 2 * t * sma : f0;
      if (ENC) {
        const uint32_t fx = has ? fx0 + 2 * t * xsma : fx0;
        if (xdirect) {
          if (PH == 1)
            L.xx.y = ldo(x, odd_ok ? f + 1 : f);
          else
            L.xx = pair_ld(x, f, al, odd_ok);
        } else {
          if (PH > 1) L.xx.x = ldo(x, fx);
          L.xx.y = ldo(x, odd_ok ? fx + xs2 : fx);
        }
      } else {
        if (PH > 1) L.ce = ldo(sc_e, has ? rk_e0 + t * rma_e : rk_e0);
        // a pair without an odd point has no odd rank: its formula can run
        // past the stage (and the code array)
        L.co = has && odd_ok ? ldo(sc_o, rk_o0 + t * rma_o) : 0u;
      }
      if (PH == 1) {
        const uint32_t cb = has ? cb0 + t * cs0 : cb0;
        L.ct[0] = ldo(A.coarse, (uint32_t)((int32_t)cb + co0));
        L.ct[1] = ldo(A.coarse, cb);
        L.ct[2] = ldo(A.coarse, (uint32_t)((int32_t)cb + co2));
        L.ct[3] = ldo(A.coarse, (uint32_t)((int32_t)cb + co3));
      } else {
        const int cm = c[ma] + 2 * (int)t;
        L.w3 = pair_ld(rec, has && cm + 3 < nma ? f + 3 * sma : f - sma, al, odd_ok);
        if (PH == 3) {
#pragma unroll
          for (int q = 0; q < 4; q++)
            L.d[q] = pair_ld(rec, (uint32_t)((int32_t)f + doff[q]), al, odd_ok);
        }
      }
    };
    // the window along ma (phases 2, 3): taps c-3, c-1, c+1 of the first step
    double we0 = 0.0, we1 = 0.0, we2 = 0.0, wo0 = 0.0, wo1 = 0.0, wo2 = 0.0;
    if (PH > 1) {
      const int cm = c[ma];
      const float2 a = pair_ld(rec, cm >= 3 ? f0 - 3 * sma : f0 - sma, al, odd_ok);
      const float2 b = pair_ld(rec, f0 - sma, al, odd_ok);
      const float2 d = pair_ld(rec, cm + 1 < nma ? f0 + sma : f0 - sma, al, odd_ok);
      we0 = a.x; wo0 = a.y; we1 = b.x; wo1 = b.y; we2 = d.x; wo2 = d.y;
    }
    // loads run PAIR_DEPTH steps ahead of the arithmetic
    // every step of the wave's chunk interior on every axis: the cubic
    // stencil runs without per-step checks
    const int cm0 = c[ma];
    const bool chunk_in = PH == 1 || (cm0 >= 3 && cm0 + 2 * ((int)ucnt - 1) + 3 < nma);
    const bool all_in = cubic && __all(chunk_in && (PH < 3 || (dv0 && dv2)) &&
                                       (!odd_ok || (kv0 && kv2)));
    // The loads run one step ahead of the arithmetic: step t's loads were
    // issued at the start of step t-1 and are moved into `cur` at the start of
    // step t, before step t+1's are issued.  (Issuing the next loads first and
    // moving at the end of the step lets the compiler's wait at the loop head
    // cover the just-issued loads too, which serialises every step on them.)
    Loads cur, nxt;
    load(0, nxt);
#pragma unroll
    for (uint32_t t = 0; t < (uint32_t)M; t++) {
      if (t >= ucnt) break;
      cur = nxt;
      if (t + 1 < ucnt) load(t + 1, nxt);
      const bool active = owned;
      const uint32_t f = f0 + 2 * t * sma;
      const uint32_t rke = rk_e0 + t * rma_e, rko = rk_o0 + t * rma_o;
      const int cm = c[ma] + 2 * (int)t;
      const bool mv0 = cm >= 3, mv1 = cm + 1 < nma, mv2 = cm + 3 < nma;
      const double we3 = PH > 1 ? (double)cur.w3.x : 0.0, wo3 = PH > 1 ? (double)cur.w3.y : 0.0;
      // --- the even point
      float re = 0.f, xe = 0.f;
      uint32_t cde = 0;
      if (PH == 1) {
        re = cur.ct[1];  // the coarse point itself
      } else {
        double v[2];
        if (all_in || (cubic && __all(mv0 && mv2 && (PH == 2 || (dv0 && dv2))))) {
          v[0] = cubic4(we0, we1, we2, we3);
          if (PH == 3)
            v[1] = cubic4((double)cur.d[0].x, (double)cur.d[1].x, (double)cur.d[2].x,
                          (double)cur.d[3].x);
        } else {
          v[0] = interp4(we0, we1, we2, we3, mv0, mv1, mv2, cubic, end_mode);
          if (PH == 3)
            v[1] = interp4((double)cur.d[0].x, (double)cur.d[1].x, (double)cur.d[2].x,
                           (double)cur.d[3].x, dv0, dv1, dv2, cubic, end_mode);
        }
        const double pd = (double)(float)pair_combine(v, NE, center);
        if (ENC) {
          xe = cur.xx.x;
          pair_quant<RND>(QE, round_mode, A.span, A.offset, pd, xe, cde, re);
        } else {
          // every lane with a point decodes it: lanes 0, 62, 63 feed the k taps
          cde = cur.ce;
          re = pair_dequant(QE, pd, cde, wv, rke, or_e, ov_e, nout_e);
        }
      }
      // --- the odd point: k taps from the even points of lanes l-1 .. l+2
      double k0, k1, k2, k3;
      if (PH == 1) {
        k0 = cur.ct[0]; k1 = cur.ct[1]; k2 = cur.ct[2]; k3 = cur.ct[3];
      } else if (bm) {
        // double-buffered across steps (and workgroup iterations): a buffer is
        // written again only after the next step's barrier, which every
        // thread reaches after its reads of this one
        float *xb = xch + (xstep++ & 1) * 256;
        xb[xpos] = re;
        __syncthreads();
        // neighbours in the same line (invalid taps read the own slot, unused)
        k0 = (double)xb[kv0 ? xpos - 1 : xpos];
        k1 = (double)re;
        k2 = (double)xb[m + 1 < A.npairs ? xpos + 1 : xpos];
        k3 = (double)xb[m + 2 < A.npairs ? xpos + 2 : xpos];
      } else {
        k0 = (double)__shfl_up(re, 1);
        k1 = (double)re;
        k2 = (double)__shfl_down(re, 1);
        k3 = (double)__shfl_down(re, 2);
      }
      double v[3];
      if (all_in || (cubic && __all(!odd_ok || (kv0 && kv2 && (PH == 1 || (mv0 && mv2)) &&
                                                (PH < 3 || (dv0 && dv2)))))) {
        int i = 0;
        if (PH > 1) v[i++] = cubic4(wo0, wo1, wo2, wo3);
        if (PH == 3)
          v[i++] = cubic4((double)cur.d[0].y, (double)cur.d[1].y, (double)cur.d[2].y,
                          (double)cur.d[3].y);
        v[i++] = cubic4(k0, k1, k2, k3);
      } else {
        int i = 0;
        if (PH > 1) v[i++] = interp4(wo0, wo1, wo2, wo3, mv0, mv1, mv2, cubic, end_mode);
        if (PH == 3)
          v[i++] = interp4((double)cur.d[0].y, (double)cur.d[1].y, (double)cur.d[2].y,
                           (double)cur.d[3].y, dv0, dv1, dv2, cubic, end_mode);
        v[i++] = interp4(k0, k1, k2, k3, kv0, kv1, kv2, cubic, end_mode);
      }
      const double po = (double)(float)pair_combine(v, NO, center);
      float ro, xo = 0.f;
      uint32_t cdo;
      if (ENC) {
        xo = cur.xx.y;
        pair_quant<RND>(QO, round_mode, A.span, A.offset, po, xo, cdo, ro);
      } else {
        cdo = cur.co;
        ro = pair_dequant(QO, po, cdo, active && odd_ok, rko, or_o, ov_o, nout_o);
      }
      if (al) {
        buf_st64(rrec, 4u * f, active, re, ro);
      } else {
        buf_st32(rrec, 4u * f, active, re);
        buf_st32(rrec, 4u * (f + 1), active && odd_ok, ro);
      }
      if (ENC) {
        const bool rare_e =
            PH > 1 && emit_common(rcodes, cb_e, rke, cde, lh, wlo_e, active);
        const bool rare_o =
            emit_common(rcodes, cb_o, rko, cdo, lh + HIST_WIN + 1, wlo_o, active && odd_ok);
        if (__any(rare_e || rare_o)) {
          if (PH > 1) emit_rare(rare_e, rke, cde, xe, or_e, ov_e, out_count + SE.gs, gh_e);
          emit_rare(rare_o, rko, cdo, xo, or_o, ov_o, out_count + SO.gs, gh_o);
        }
      }
      if (PH > 1) {
        we0 = we1; we1 = we2; we2 = we3;
        wo0 = wo1; wo1 = wo2; wo2 = wo3;
      }
    }
  }
}

#ifndef PAIR_WAVES_PER_SIMD
#define PAIR_WAVES_PER_SIMD 1  // occupancy floor handed to the register allocator
#endif
// WPE: the occupancy floor in waves per SIMD (the register budget the
// allocator works to); the "pair_waves" knob picks one wave per SIMD more than
// the compiler's own choice for the finest level's encode phases 2 and 3
// (phase 2: 6 -> 7 waves, 72 VGPRs; phase 3: 4 -> 5 waves, 96 VGPRs).
template <int PH>
struct PairWavesAlt {
  static constexpr int value = PH == 3 ? 5 : 7;
};
template <int PH, bool ENC, bool AL, bool RND, bool BM, bool XD, int WPE = PAIR_WAVES_PER_SIMD>
__global__ void __launch_bounds__(256, WPE) pair_kernel(PArgs A, const float *__restrict__ x,
                                                   float *__restrict__ rec,
                                                   uint16_t *__restrict__ codes,
                                                   uint32_t *__restrict__ hist,
                                                   uint32_t *__restrict__ out_rank,
                                                   float *__restrict__ out_val,
                                                   uint32_t *__restrict__ out_count) {
  // two histogram windows of HIST_WIN bins and a spare word each
  __shared__ uint32_t lh[ENC ? 2 * (HIST_WIN + 1) : 1];
  __shared__ float xch[PH > 1 ? 2 * 256 : 1];
  if (ENC) {
    for (int i = threadIdx.x; i < 2 * (HIST_WIN + 1); i += blockDim.x) lh[i] = 0;
    __syncthreads();
  }
  const PSet &T = A.set[blockIdx.y];
  pair_set<PH, ENC, AL, RND, BM, XD>(A, T, x, rec, codes, hist, lh, xch, out_rank, out_val,
                                     out_count);
  if (ENC) {
    __syncthreads();
    const L3Stage &SE = A.st[T.ev.stage], &SO = A.st[T.od.stage];
    if (PH > 1) flush_hist(lh, hist + (int64_t)SE.gs * A.hist_words, SE.R);
    flush_hist(lh + HIST_WIN + 1, hist + (int64_t)SO.gs * A.hist_words, SO.R);
  }
}

// ---------------------------------------------------------------------------
// Line-quad kernels: phases 2 and 3 of the pair scheme with one thread per
// four points of a line (k = 4q .. 4q + 3: even, odd, even, odd), for lines
// whose length is a multiple of 4 and at most 1024.  A workgroup takes whole
// lines (256 / nq of them, nq = n2 / 4); each load and store moves the four
// points, the two even points go through LDS for the k taps of the odd ones:
//   odd 4q+1: 4q-2 (thread q-1), 4q, 4q+2, 4q+4 (thread q+1);
//   odd 4q+3: 4q, 4q+2, 4q+4, 4q+6 (thread q+1).
// Threads take the lines' interior quads first and the two end quads of every
// line last, so only the workgroup's last wave meets the k end rules and the
// others run the cubic stencil without per-lane cases.  The arithmetic, tap
// order and emission are those of the pair kernels: the same codes, outliers
// and reconstruction.  Per point it issues about a third
// fewer instructions (address arithmetic, loads, stores, exchange and loop
// control are shared by four points instead of two), which is what bounds
// the pair kernels.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void buf_st128(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on,
                                          float a, float b, float c, float d) {
  typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
  const u32x4_t v = {__float_as_uint(a), __float_as_uint(b), __float_as_uint(c),
                     __float_as_uint(d)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, on ? byte_off : BUF_OFF_NONE, 0, 0);
}

// one axis's prediction from four taps (the cubic stencil when `in`)
__device__ __forceinline__ double quad_axis(bool in, double a, double b, double c, double d,
                                            bool v0, bool v1, bool v2, int cubic, int end_mode) {
  return in ? cubic4(a, b, c, d) : interp4(a, b, c, d, v0, v1, v2, cubic, end_mode);
}

template <int PH, bool ENC, bool RND, bool XD>
__device__ __forceinline__ void quad_set(const PArgs &A, const PSet &T,
                                         const float *__restrict__ x, float *__restrict__ rec,
                                         uint16_t *__restrict__ codes,
                                         uint32_t *__restrict__ hist, uint32_t *lh, float *xch,
                                         uint32_t *__restrict__ out_rank,
                                         float *__restrict__ out_val,
                                         uint32_t *__restrict__ out_count) {
  static_assert(PH == 2 || PH == 3, "quads run phases 2 and 3");
  constexpr int M = PAIR_M;
  constexpr int NE = PH == 2 ? 1 : 2;  // odd axes of the even class
  constexpr int NO = NE + 1;           // odd axes of the odd class
  const uint32_t S = A.steps;
  const int ma = T.ma, ra = T.ra;
  const uint32_t gstr[3] = {(uint32_t)A.n[1] * (uint32_t)A.n[2], (uint32_t)A.n[2], 1u};
  const uint32_t sma = gstr[ma], sra = gstr[ra];
  const int nma = A.n[ma], nra = A.n[ra];
  const uint32_t xsma = A.xs[ma], xs2 = A.xs[2];
  const int cubic = A.cubic, end_mode = A.end_mode, center = A.center;
  const int round_mode = A.round_mode;
  const L3Stage &SE = A.st[T.ev.stage], &SO = A.st[T.od.stage];
  const PQ QE = make_pq(SE);
  const PQ QO = make_pq(SO);
  const uint32_t wlo_e = SE.R > HIST_WIN / 2 ? SE.R - HIST_WIN / 2 : 0;
  const uint32_t wlo_o = SO.R > HIST_WIN / 2 ? SO.R - HIST_WIN / 2 : 0;
  uint32_t *gh_e = ENC ? hist + (int64_t)SE.gs * A.hist_words : nullptr;
  uint32_t *gh_o = ENC ? hist + (int64_t)SO.gs * A.hist_words : nullptr;
  uint32_t *or_e = out_rank + SE.out_lo, *or_o = out_rank + SO.out_lo;
  float *ov_e = out_val + SE.out_lo, *ov_o = out_val + SO.out_lo;
  const int64_t nout_e = ENC ? 0 : (int64_t)out_count[SE.gs];
  const int64_t nout_o = ENC ? 0 : (int64_t)out_count[SO.gs];
  const uint16_t *sc_e = codes + SE.code_base, *sc_o = codes + SO.code_base;
  const __amdgpu_buffer_rsrc_t rrec = buf_rsrc(rec), rcodes = buf_rsrc(codes);
  const uint32_t cb_e = 2u * (uint32_t)SE.code_base, cb_o = 2u * (uint32_t)SO.code_base;
  const uint32_t rma_e = (uint32_t)T.ev.rstr[ma], rma_o = (uint32_t)T.od.rstr[ma];
  const uint32_t nq = A.nquads, nin = nq >= 2 ? nq - 2 : 0;
  const int tl = threadIdx.x;
  const uint32_t G = gridDim.x;
  const uint32_t lb = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  uint32_t xstep = 0;  // steps taken by the workgroup: picks the exchange buffer
  for (uint32_t base = lb * 256u; base < T.nthreads; base += G * 256u) {
    // nthreads is a multiple of 256: the whole workgroup is in range
    const uint32_t blk = base >> 8;
    const uint32_t chunk = blk / T.bpc;
    // interior quads 1 .. nq-2 of the workgroup's lines, then the end quads
    uint32_t line, q;
    if ((uint32_t)tl < T.lpb * nin) {
      line = (uint32_t)tl / nin;
      q = 1 + (uint32_t)tl - line * nin;
    } else if (nq == 1) {
      line = (uint32_t)tl - T.lpb * nin;
      q = 0;
    } else {
      const uint32_t b = (uint32_t)tl - T.lpb * nin;
      line = b >> 1;
      q = (b & 1) ? nq - 1 : 0;
    }
    // a thread past the workgroup's lines holds exchange slot tl (>= lpb * nq)
    const uint32_t pos = line < T.lpb ? line * nq + q : (uint32_t)tl;
    uint32_t row = (blk - chunk * T.bpc) * T.lpb + line;
    const bool owned = line < T.lpb && row < T.nrows;
    row = min(row, T.nrows - 1);
    q = min(q, nq - 1);
    const uint32_t ucnt = (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)min(S, T.nmarch - chunk * S));  // the same for the whole workgroup
    int c[3];
    c[ma] = T.pm + 2 * (int)(chunk * S);
    c[ra] = T.pr + 2 * (int)row;
    c[2] = 4 * (int)q;
    const uint32_t f0 = (uint32_t)c[0] * gstr[0] + (uint32_t)c[1] * gstr[1] + (uint32_t)c[2];
    const uint32_t fx0 = (uint32_t)c[0] * A.xs[0] + (uint32_t)c[1] * A.xs[1] +
                         (uint32_t)c[2] * A.xs[2];
    uint32_t ci[3];
    ci[ma] = chunk * S;
    ci[ra] = row;
    ci[2] = 2 * q;  // class index of the quad's first even and first odd point
    const uint32_t rk_e0 = (uint32_t)(T.ev.offset + (int64_t)ci[0] * T.ev.rstr[0] +
                                      (int64_t)ci[1] * T.ev.rstr[1] +
                                      (int64_t)ci[2] * T.ev.rstr[2]);
    const uint32_t rk_o0 = (uint32_t)(T.od.offset + (int64_t)ci[0] * T.od.rstr[0] +
                                      (int64_t)ci[1] * T.od.rstr[1] +
                                      (int64_t)ci[2] * T.od.rstr[2]);
    const uint32_t rs2_e = (uint32_t)T.ev.rstr[2], rs2_o = (uint32_t)T.od.rstr[2];
    // k validity: odd 4q+1 has k-3 iff q >= 1; both odd points have their
    // k+1 / k+3 taps iff q + 1 < nq (4q+1's k+1 tap always exists)
    const bool qv0 = q >= 1, qvr = q + 1 < nq;
    // direct taps along ra (phase 3): validity is constant along the march
    const int cr = c[ra];
    const bool dv0 = cr >= 3, dv1 = cr + 1 < nra, dv2 = cr + 3 < nra;
    int32_t doff[4] = {0, 0, 0, 0};
    if (PH == 3) {
      doff[0] = dv0 ? -3 * (int32_t)sra : -(int32_t)sra;
      doff[1] = -(int32_t)sra;
      doff[2] = dv1 ? (int32_t)sra : -(int32_t)sra;
      doff[3] = dv2 ? 3 * (int32_t)sra : -(int32_t)sra;
    }
    struct Loads {
      float4 xx;
      uint32_t ce0, ce1, co0, co1;
      float4 w3;
      float4 d[PH == 3 ? 4 : 1];
    };
    auto load = [&](uint32_t t, Loads &L) {
      const uint32_t f = f0 + 2 * t * sma;
      if (ENC) {
        if (XD) {
          L.xx = ldo((const float4 *)x, f >> 2);
        } else {
          const uint32_t fx = fx0 + 2 * t * xsma;
          L.xx = make_float4(ldo(x, fx), ldo(x, fx + xs2), ldo(x, fx + 2 * xs2),
                             ldo(x, fx + 3 * xs2));
        }
      } else {
        const uint32_t re = rk_e0 + t * rma_e, ro = rk_o0 + t * rma_o;
        L.ce0 = ldo(sc_e, re);
        L.ce1 = ldo(sc_e, re + rs2_e);
        L.co0 = ldo(sc_o, ro);
        L.co1 = ldo(sc_o, ro + rs2_o);
      }
      const int cm = c[ma] + 2 * (int)t;
      L.w3 = ldo((const float4 *)rec, (cm + 3 < nma ? f + 3 * sma : f - sma) >> 2);
      if (PH == 3) {
#pragma unroll
        for (int k = 0; k < 4; k++)
          L.d[k] = ldo((const float4 *)rec, (uint32_t)((int32_t)f + doff[k]) >> 2);
      }
    };
    // the window along ma: taps c-3, c-1, c+1 of the first step, four columns
    double wa[3], wb[3], wc[3], wd[3];  // columns k = 4q, 4q+1, 4q+2, 4q+3
    {
      const int cm = c[ma];
      const float4 r0 = ldo((const float4 *)rec, (cm >= 3 ? f0 - 3 * sma : f0 - sma) >> 2);
      const float4 r1 = ldo((const float4 *)rec, (f0 - sma) >> 2);
      const float4 r2 = ldo((const float4 *)rec, (cm + 1 < nma ? f0 + sma : f0 - sma) >> 2);
      wa[0] = r0.x; wb[0] = r0.y; wc[0] = r0.z; wd[0] = r0.w;
      wa[1] = r1.x; wb[1] = r1.y; wc[1] = r1.z; wd[1] = r1.w;
      wa[2] = r2.x; wb[2] = r2.y; wc[2] = r2.z; wd[2] = r2.w;
    }
    const int cm0 = c[ma];
    const bool chunk_in = cm0 >= 3 && cm0 + 2 * ((int)ucnt - 1) + 3 < nma;
    const bool all_in = cubic && __all(chunk_in && (PH < 3 || (dv0 && dv2)) && qv0 && qvr);
    Loads cur, nxt;
    load(0, nxt);
#pragma unroll
    for (uint32_t t = 0; t < (uint32_t)M; t++) {
      if (t >= ucnt) break;
      cur = nxt;
      if (t + 1 < ucnt) load(t + 1, nxt);
      const uint32_t f = f0 + 2 * t * sma;
      const uint32_t rke = rk_e0 + t * rma_e, rko = rk_o0 + t * rma_o;
      const int cm = c[ma] + 2 * (int)t;
      const bool mv0 = cm >= 3, mv1 = cm + 1 < nma, mv2 = cm + 3 < nma;
      const double wa3 = cur.w3.x, wb3 = cur.w3.y, wc3 = cur.w3.z, wd3 = cur.w3.w;
      // --- the even points 4q, 4q+2
      const bool ein = all_in || (cubic && __all(mv0 && mv2 && (PH == 2 || (dv0 && dv2))));
      double ve0[2], ve1[2];
      ve0[0] = quad_axis(ein, wa[0], wa[1], wa[2], wa3, mv0, mv1, mv2, cubic, end_mode);
      ve1[0] = quad_axis(ein, wc[0], wc[1], wc[2], wc3, mv0, mv1, mv2, cubic, end_mode);
      if (PH == 3) {
        ve0[1] = quad_axis(ein, cur.d[0].x, cur.d[1].x, cur.d[2].x, cur.d[3].x, dv0, dv1, dv2,
                           cubic, end_mode);
        ve1[1] = quad_axis(ein, cur.d[0].z, cur.d[1].z, cur.d[2].z, cur.d[3].z, dv0, dv1, dv2,
                           cubic, end_mode);
      }
      const double pe0 = (double)(float)pair_combine(ve0, NE, center);
      const double pe1 = (double)(float)pair_combine(ve1, NE, center);
      float re0, re1, xe0 = 0.f, xe1 = 0.f;
      uint32_t cde0, cde1;
      if (ENC) {
        xe0 = cur.xx.x;
        xe1 = cur.xx.z;
        pair_quant2<RND>(QE, round_mode, A.span, A.offset, pe0, xe0, pe1, xe1, cde0, re0, cde1,
                         re1);
      } else {
        // every thread decodes its even points: its neighbours' odd points need them
        cde0 = cur.ce0;
        cde1 = cur.ce1;
        re0 = pair_dequant(QE, pe0, cde0, true, rke, or_e, ov_e, nout_e);
        re1 = pair_dequant(QE, pe1, cde1, true, rke + rs2_e, or_e, ov_e, nout_e);
      }
      // --- exchange: 4q-2 from thread q-1, 4q+4 and 4q+6 from thread q+1
      // (double-buffered across steps: a buffer is written again only after
      // the next step's barrier, which every thread reaches after its reads)
      float *xb = xch + (xstep++ & 1) * 512;
      xb[pos] = re0;
      xb[256 + pos] = re1;
      __syncthreads();
      // invalid taps (line ends) read the thread's own slot, value unused
      const uint32_t pl = qv0 ? pos - 1 : pos, pr = qvr ? pos + 1 : pos;
      const double kl = xb[256 + pl];
      const double kr0 = xb[pr];
      const double kr1 = xb[256 + pr];
      const double k0 = re0, k1 = re1;
      // --- the odd points 4q+1, 4q+3
      const bool oin = all_in || (cubic && __all(qv0 && qvr && mv0 && mv2 &&
                                                 (PH < 3 || (dv0 && dv2))));
      double vo0[3], vo1[3];
      vo0[0] = quad_axis(oin, wb[0], wb[1], wb[2], wb3, mv0, mv1, mv2, cubic, end_mode);
      vo1[0] = quad_axis(oin, wd[0], wd[1], wd[2], wd3, mv0, mv1, mv2, cubic, end_mode);
      if (PH == 3) {
        vo0[1] = quad_axis(oin, cur.d[0].y, cur.d[1].y, cur.d[2].y, cur.d[3].y, dv0, dv1, dv2,
                           cubic, end_mode);
        vo1[1] = quad_axis(oin, cur.d[0].w, cur.d[1].w, cur.d[2].w, cur.d[3].w, dv0, dv1, dv2,
                           cubic, end_mode);
      }
      vo0[NO - 1] = quad_axis(oin, kl, k0, k1, kr0, qv0, true, qvr, cubic, end_mode);
      vo1[NO - 1] = quad_axis(oin, k0, k1, kr0, kr1, true, qvr, qvr, cubic, end_mode);
      const double po0 = (double)(float)pair_combine(vo0, NO, center);
      const double po1 = (double)(float)pair_combine(vo1, NO, center);
      float ro0, ro1, xo0 = 0.f, xo1 = 0.f;
      uint32_t cdo0, cdo1;
      if (ENC) {
        xo0 = cur.xx.y;
        xo1 = cur.xx.w;
        pair_quant2<RND>(QO, round_mode, A.span, A.offset, po0, xo0, po1, xo1, cdo0, ro0, cdo1,
                         ro1);
      } else {
        cdo0 = cur.co0;
        cdo1 = cur.co1;
        ro0 = pair_dequant(QO, po0, cdo0, owned, rko, or_o, ov_o, nout_o);
        ro1 = pair_dequant(QO, po1, cdo1, owned, rko + rs2_o, or_o, ov_o, nout_o);
      }
      buf_st128(rrec, 4u * f, owned, re0, ro0, re1, ro1);
      if (ENC) {
        const bool r0 = emit_common(rcodes, cb_e, rke, cde0, lh, wlo_e, owned);
        const bool r1 = emit_common(rcodes, cb_e, rke + rs2_e, cde1, lh, wlo_e, owned);
        const bool r2 = emit_common(rcodes, cb_o, rko, cdo0, lh + HIST_WIN + 1, wlo_o, owned);
        const bool r3 =
            emit_common(rcodes, cb_o, rko + rs2_o, cdo1, lh + HIST_WIN + 1, wlo_o, owned);
        if (__any(r0 || r1 || r2 || r3)) {
          emit_rare(r0, rke, cde0, xe0, or_e, ov_e, out_count + SE.gs, gh_e);
          emit_rare(r1, rke + rs2_e, cde1, xe1, or_e, ov_e, out_count + SE.gs, gh_e);
          emit_rare(r2, rko, cdo0, xo0, or_o, ov_o, out_count + SO.gs, gh_o);
          emit_rare(r3, rko + rs2_o, cdo1, xo1, or_o, ov_o, out_count + SO.gs, gh_o);
        }
      }
      wa[0] = wa[1]; wa[1] = wa[2]; wa[2] = wa3;
      wb[0] = wb[1]; wb[1] = wb[2]; wb[2] = wb3;
      wc[0] = wc[1]; wc[1] = wc[2]; wc[2] = wc3;
      wd[0] = wd[1]; wd[1] = wd[2]; wd[2] = wd3;
    }
  }
}

template <int PH, bool ENC, bool RND, bool XD>
__global__ void __launch_bounds__(256) quad_kernel(PArgs A, const float *__restrict__ x,
                                                   float *__restrict__ rec,
                                                   uint16_t *__restrict__ codes,
                                                   uint32_t *__restrict__ hist,
                                                   uint32_t *__restrict__ out_rank,
                                                   float *__restrict__ out_val,
                                                   uint32_t *__restrict__ out_count) {
  __shared__ uint32_t lh[ENC ? 2 * (HIST_WIN + 1) : 1];
  __shared__ float xch[2 * 512];
  if (ENC) {
    for (int i = threadIdx.x; i < 2 * (HIST_WIN + 1); i += blockDim.x) lh[i] = 0;
    __syncthreads();
  }
  const PSet &T = A.set[blockIdx.y];
  quad_set<PH, ENC, RND, XD>(A, T, x, rec, codes, hist, lh, xch, out_rank, out_val, out_count);
  if (ENC) {
    __syncthreads();
    const L3Stage &SE = A.st[T.ev.stage], &SO = A.st[T.od.stage];
    flush_hist(lh, hist + (int64_t)SE.gs * A.hist_words, SE.R);
    flush_hist(lh + HIST_WIN + 1, hist + (int64_t)SO.gs * A.hist_words, SO.R);
  }
}

// ---------------------------------------------------------------------------
// Level-fused slab kernel (3-D level grids, one channel, 16-bit codes).
//
// A workgroup owns a tile of SL_TJ x SL_TK columns (axes 1, 2) of the level
// grid and a segment of SL_SI planes along axis 0, and marches through the
// segment plane by plane.  Its LDS holds the tile plus a halo of 3 columns
// on each side for: a ring of the last four even planes (axis-0 taps of the
// odd planes), the current odd plane, and both planes' input x.  Each step
// takes even plane e and odd plane o = e - 3 (whose axis-0 taps o +- 1,
// o +- 3 are then in the ring) through three compute phases: {010, 001 on e;
// 100 on o}, {011 on e; 110, 101 on o}, {111 on o}; the next step's coarse
// points and x are loaded into registers while these run.  A class is computed on the halo only where a later
// class of the tile reads it (100 on the whole region, 001 / 101 on the
// halo rows, 010 / 110 on the halo columns); owned points emit their code,
// histogram count and outliers; every owned point of a finished plane --
// including the coarse points -- is written back to recon as full rows.  x
// is read once per plane as full region rows.  Arithmetic is interp4 /
// quantize_one's, so codes, outliers and recon equal the stage kernels'.
// ---------------------------------------------------------------------------
#define SL_TJ 32
#define SL_TK 64
#define SL_SI 32
#define SL_THREADS 512
#define SL_RJ (SL_TJ + 6)
#define SL_RK (SL_TK + 6)
#define SL_PLANE (SL_RJ * SL_RK)
#define SL_HB 256

struct SlabStage {
  double w, inv_w;
  float feb;
  uint32_t R;
  int64_t code_base;
  int32_t gs, pad;
};
struct SlabClass {
  int64_t offset;
  int64_t rstr[3];
};
struct SlabArgs {
  int32_t n[3], cn[3];  // level grid and its coarse (even-point) grid
  int32_t tiles_j, tiles_k, segs, nblocks;
  int32_t cubic, center, end_mode, round_mode;
  double span, offset;
  int64_t hist_words;
  SlabStage st[5];
  SlabClass cl[8];  // by parity 1..7
};

__device__ __forceinline__ int slab_stage(int P) {
  return P == 1 ? 0 : P == 2 ? 1 : P == 4 ? 2 : P == 7 ? 4 : 3;
}

template <bool ENC>
__global__ void __launch_bounds__(SL_THREADS) slab_kernel(SlabArgs A, const float *__restrict__ x,
                                                          const float *__restrict__ coarse,
                                                          float *__restrict__ rec,
                                                          uint16_t *__restrict__ codes,
                                                          uint32_t *__restrict__ hist,
                                                          uint32_t *__restrict__ out_rank,
                                                          float *__restrict__ out_val,
                                                          uint32_t *__restrict__ out_count) {
  extern __shared__ __align__(16) unsigned char sl_smem[];
  float(*ring)[SL_PLANE] = (float(*)[SL_PLANE])sl_smem;
  float *odd = (float *)sl_smem + 4 * SL_PLANE;
  float *xe = odd + SL_PLANE;             // x of the even plane (encode)
  float *xo = xe + SL_PLANE;              // x of the odd plane (encode)
  uint32_t *lh = (uint32_t *)(xo + SL_PLANE);
  const int tid = threadIdx.x;
  const int n0 = A.n[0], n1 = A.n[1], n2 = A.n[2];
  // XCD-aware block order: each XCD walks a contiguous range of tiles
  const uint32_t G = gridDim.x;
  const uint32_t b = (G % 8 == 0) ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  if ((int)b >= A.nblocks) return;
  const int tk = b % A.tiles_k, tj = (b / A.tiles_k) % A.tiles_j;
  const int seg = b / (A.tiles_k * A.tiles_j);
  const int j0 = tj * SL_TJ, k0 = tk * SL_TK, I0 = seg * SL_SI;
  const int I1 = min(I0 + SL_SI, n0);
  const int j1 = min(j0 + SL_TJ, n1), k1 = min(k0 + SL_TK, n2);      // owned
  const int rj0 = max(j0 - 3, 0), rj1 = min(j0 + SL_TJ + 3, n1);     // region
  const int rk0 = max(k0 - 3, 0), rk1 = min(k0 + SL_TK + 3, n2);
  if (ENC) {
    for (int i = tid; i < 5 * SL_HB; i += SL_THREADS) lh[i] = 0;
  }
  auto L = [&](int j, int k) { return (j - j0 + 3) * SL_RK + (k - k0 + 3); };
  // --- one class on one plane over rows [ja, jb) x cols [ka, kb) of parity P
  auto run_class = [&](auto PC, int i, float *self, const float *xin, int ja, int jb, int ka,
                       int kb) {
    constexpr int P = decltype(PC)::value;  // the class parity, compile-time
    const int pj = (P >> 1) & 1, pk = (P >> 2) & 1;
    ja += (ja & 1) != pj;
    ka += (ka & 1) != pk;
    const int nr = jb > ja ? (jb - ja + 1) / 2 : 0, nc = kb > ka ? (kb - ka + 1) / 2 : 0;
    const int cnt = nr * nc;
    const int si = slab_stage(P);
    const SlabStage &S = A.st[si];
    const SlabClass &K = A.cl[P];
    const bool own_i = i >= I0 && i < I1;
    const uint32_t win_lo = S.R > SL_HB / 2 ? S.R - SL_HB / 2 : 0;
    const int64_t nout = ENC ? 0 : (int64_t)out_count[S.gs];
    constexpr int nodd = (P & 1) + ((P >> 1) & 1) + ((P >> 2) & 1);
    for (int base = 0; base < cnt; base += SL_THREADS) {
      const int q = base + tid;
      const bool act = q < cnt;
      const int r = act ? q / nc : 0, c = act ? q - r * nc : 0;
      const int j = ja + 2 * r, k = ka + 2 * c;
      const int l = L(j, k);
      // taps per odd axis (always inside the LDS planes; invalid ones unused)
      double v[3];
      bool interior = true;
      int na = 0;
      float tap[3][4];
      bool vm3[3], vp1[3], vp3[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (!((P >> a) & 1)) continue;
        const int ca = a == 0 ? i : a == 1 ? j : k;
        const int nn = a == 0 ? n0 : a == 1 ? n1 : n2;
        vm3[na] = ca >= 3;
        vp1[na] = ca + 1 < nn;
        vp3[na] = ca + 3 < nn;
        interior = interior && vm3[na] && vp3[na];
        if (a == 0) {
          tap[na][0] = ring[(((i - 3) >> 1) & 3)][l];
          tap[na][1] = ring[(((i - 1) >> 1) & 3)][l];
          tap[na][2] = ring[(((i + 1) >> 1) & 3)][l];
          tap[na][3] = ring[(((i + 3) >> 1) & 3)][l];
        } else {
          const int s = a == 1 ? SL_RK : 1;
          tap[na][0] = self[l - 3 * s];
          tap[na][1] = self[l - s];
          tap[na][2] = self[l + s];
          tap[na][3] = self[l + 3 * s];
        }
        na++;
      }
      if (A.cubic && __all(interior || !act)) {
#pragma unroll
        for (int m = 0; m < nodd; m++)
          v[m] = cubic4((double)tap[m][0], (double)tap[m][1], (double)tap[m][2],
                        (double)tap[m][3]);
      } else {
#pragma unroll
        for (int m = 0; m < nodd; m++)
          v[m] = interp4((double)tap[m][0], (double)tap[m][1], (double)tap[m][2],
                         (double)tap[m][3], vm3[m], vp1[m], vp3[m], A.cubic, A.end_mode);
      }
      double res;
      if (nodd == 1) {
        res = v[0];
      } else if (A.center == 0) {
        double total = v[0] + v[1];
        if (nodd == 3) total = total + v[2];
        res = total / (double)nodd;
      } else {
        res = A.center == 1 ? v[0] : v[nodd - 1];
      }
      const float p = (float)res;
      const bool owned = act && own_i && j >= j0 && j < j1 && k >= k0 && k < k1;
      const uint32_t rank = (uint32_t)(K.offset + (int64_t)(i >> 1) * K.rstr[0] +
                                       (int64_t)(j >> 1) * K.rstr[1] +
                                       (int64_t)(k >> 1) * K.rstr[2]);
      uint16_t *scodes = codes + S.code_base;
      uint32_t code = 0;
      float xv = 0.f, rv = 0.f;
      if (ENC) {
        xv = xin[l];
        const double pd = (double)p;
        const double d = (double)xv - pd;
        const double tq = d * S.inv_w;
        double qd = rint(tq);
        const bool fb = !(0.5 - fabs(tq - qd) > 1e-12 * fmax(fabs(tq), 1.0));
        if (__any(fb && act)) {
          if (fb) qd = rint(d / S.w);
        }
        const bool inr = fabs(qd) < (double)S.R;
        const int qi = inr ? (int)qd : 0;
        const float cand = (float)(pd + S.w * (double)qi);
        float chk = cand;
        if (A.round_mode == 1)
          chk = rintf(cand);
        else if (A.round_mode == 2)
          chk = (float)((rint((double)cand * A.span + A.offset) - A.offset) / A.span);
        const bool ok = inr && fabsf(xv - chk) <= S.feb;
        code = ok ? (uint32_t)(qi + (int)S.R) : 0u;
        rv = ok ? cand : xv;
      } else {
        code = act ? (uint32_t)scodes[rank] : 1u;
        rv = (float)((double)p + S.w * (double)((int)code - (int)S.R));
        if (__any(act && code == 0)) {
          if (act && code == 0) {
            const uint32_t *orr = out_rank + S.code_base;
            int64_t lo = 0, hi = nout;
            while (lo < hi) {
              const int64_t mid = (lo + hi) >> 1;
              if ((int64_t)orr[mid] < (int64_t)rank) lo = mid + 1; else hi = mid;
            }
            rv = (lo < nout && (int64_t)orr[lo] == (int64_t)rank) ? out_val[S.code_base + lo]
                                                                 : __int_as_float(0x7fc00000);
          }
        }
      }
      if (act) self[l] = rv;
      if (ENC) {
        if (owned) {
          scodes[rank] = (uint16_t)code;
          const uint32_t bb = code - win_lo;
          if (code != 0 && bb < SL_HB)
            lds_add(&lh[si * SL_HB + bb], 1u);
          else
            gbl_add(&hist[(int64_t)S.gs * A.hist_words + code], 1u);
        }
        const bool out = owned && code == 0;
        const uint64_t m = __ballot(out);
        if (m) {
          const int lane = __lane_id();
          const int leader = __ffsll((unsigned long long)m) - 1;
          uint32_t ob = 0;
          if (lane == leader) ob = atomicAdd(out_count + S.gs, (uint32_t)__popcll(m));
          ob = __shfl(ob, leader);
          if (out) {
            const uint32_t kk = ob + (uint32_t)__popcll(m & ((1ull << lane) - 1));
            out_rank[S.code_base + kk] = rank;
            out_val[S.code_base + kk] = xv;
          }
        }
      }
    }
  };
  // --- write the owned part of a finished plane back as full rows
  auto write_plane = [&](int i, const float *self) {
    const int w = k1 - k0;
    for (int q = tid; q < (j1 - j0) * w; q += SL_THREADS) {
      const int j = j0 + q / w, k = k0 + q % w;
      rec[((uint32_t)i * (uint32_t)n1 + (uint32_t)j) * (uint32_t)n2 + (uint32_t)k] = self[L(j, k)];
    }
  };
  // --- the loads of one step (coarse points of even plane e, x of e and of
  // o = e - 3), held in registers until the step's LDS buffers are free
  constexpr int PFC = (((SL_RJ + 1) / 2) * ((SL_RK + 1) / 2) + SL_THREADS - 1) / SL_THREADS;
  constexpr int PFX = (SL_PLANE + SL_THREADS - 1) / SL_THREADS;
  struct Pref {
    float c[PFC], xe[ENC ? PFX : 1], xo[ENC ? PFX : 1];
  };
  const int cn1 = A.cn[1], cn2 = A.cn[2];
  const int cj0 = (rj0 + 1) >> 1, cj1 = (rj1 + 1) >> 1;
  const int ck0 = (rk0 + 1) >> 1, ck1 = (rk1 + 1) >> 1;
  const int cw = ck1 - ck0, ccount = (cj1 - cj0) * cw;
  const int xw = rk1 - rk0, xcount = (rj1 - rj0) * xw;
  auto odd_of = [&](int e) { const int o = e - 3; return o >= I0 + 1 && o < I1 && o < n0; };
  auto issue = [&](int e, Pref &P) {
    if (e >= 0 && e < n0) {
#pragma unroll
      for (int m = 0; m < PFC; m++) {
        const int q = tid + m * SL_THREADS;
        if (q < ccount) {
          const int cj = cj0 + q / cw, ck = ck0 + q % cw;
          P.c[m] = coarse[((uint32_t)(e >> 1) * (uint32_t)cn1 + (uint32_t)cj) * (uint32_t)cn2 +
                          (uint32_t)ck];
        }
      }
      if (ENC) {
#pragma unroll
        for (int m = 0; m < PFX; m++) {
          const int q = tid + m * SL_THREADS;
          if (q < xcount) {
            const int j = rj0 + q / xw, k = rk0 + q % xw;
            P.xe[m] = x[((uint32_t)e * (uint32_t)n1 + (uint32_t)j) * (uint32_t)n2 + (uint32_t)k];
          }
        }
      }
    }
    if (ENC && odd_of(e)) {
      const int o = e - 3;
#pragma unroll
      for (int m = 0; m < PFX; m++) {
        const int q = tid + m * SL_THREADS;
        if (q < xcount) {
          const int j = rj0 + q / xw, k = rk0 + q % xw;
          P.xo[m] = x[((uint32_t)o * (uint32_t)n1 + (uint32_t)j) * (uint32_t)n2 + (uint32_t)k];
        }
      }
    }
  };
  auto commit = [&](int e, const Pref &P) {
    if (e >= 0 && e < n0) {
      float *slot = ring[(e >> 1) & 3];
#pragma unroll
      for (int m = 0; m < PFC; m++) {
        const int q = tid + m * SL_THREADS;
        if (q < ccount) slot[L(2 * (cj0 + q / cw), 2 * (ck0 + q % cw))] = P.c[m];
      }
      if (ENC) {
#pragma unroll
        for (int m = 0; m < PFX; m++) {
          const int q = tid + m * SL_THREADS;
          if (q < xcount) xe[L(rj0 + q / xw, rk0 + q % xw)] = P.xe[m];
        }
      }
    }
    if (ENC && odd_of(e)) {
#pragma unroll
      for (int m = 0; m < PFX; m++) {
        const int q = tid + m * SL_THREADS;
        if (q < xcount) xo[L(rj0 + q / xw, rk0 + q % xw)] = P.xo[m];
      }
    }
  };
  __syncthreads();
  Pref cur;
  issue(I0 - 2, cur);
  int prev_o = -1;
  for (int e = I0 - 2; e <= I1 + 2; e += 2) {
    const bool doE = e >= 0 && e < n0;
    const int o = e - 3;
    const bool doO = odd_of(e);
    float *slot = ring[((e >> 1) & 3)];
    // load phase: the previous odd plane goes out, this step's loads go in
    if (prev_o >= 0) write_plane(prev_o, odd);
    commit(e, cur);
    __syncthreads();
    Pref nxt;
    if (e + 2 <= I1 + 2) issue(e + 2, nxt);
    using C1 = std::integral_constant<int, 1>;
    using C2 = std::integral_constant<int, 2>;
    using C3 = std::integral_constant<int, 3>;
    using C4 = std::integral_constant<int, 4>;
    using C5 = std::integral_constant<int, 5>;
    using C6 = std::integral_constant<int, 6>;
    using C7 = std::integral_constant<int, 7>;
    if (doE) {
      run_class(C2{}, e, slot, xe, j0, j1, rk0, rk1);    // 010: owned rows, all columns
      run_class(C4{}, e, slot, xe, rj0, rj1, k0, k1);    // 001: all rows, owned columns
    }
    if (doO) run_class(C1{}, o, odd, xo, rj0, rj1, rk0, rk1);  // 100: the whole region
    __syncthreads();
    if (doE) run_class(C6{}, e, slot, xe, j0, j1, k0, k1);      // 011
    if (doO) {
      run_class(C3{}, o, odd, xo, j0, j1, rk0, rk1);     // 110: owned rows, all columns
      run_class(C5{}, o, odd, xo, rj0, rj1, k0, k1);     // 101: all rows, owned columns
    }
    __syncthreads();
    if (doO) run_class(C7{}, o, odd, xo, j0, j1, k0, k1);       // 111
    if (doE && e >= I0 && e < I1) write_plane(e, slot);
    __syncthreads();
    prev_o = doO ? o : -1;
    cur = nxt;
  }
  if (prev_o >= 0) write_plane(prev_o, odd);
  if (ENC) {  // flush the stage windows into the striped histogram copies
    __syncthreads();
    for (int si = 0; si < 5; si++) {
      const SlabStage &S = A.st[si];
      const uint32_t win_lo = S.R > SL_HB / 2 ? S.R - SL_HB / 2 : 0;
      const uint32_t glo = S.R > HIST_WIN / 2 ? S.R - HIST_WIN / 2 : 0;
      uint32_t *gh = hist + (int64_t)S.gs * A.hist_words;
      uint32_t *gwin = gh + 2 * (int64_t)S.R + (int64_t)(b % HIST_COPIES) * HIST_WIN;
      for (int i = tid; i < SL_HB; i += SL_THREADS) {
        const uint32_t v = lh[si * SL_HB + i];
        const uint32_t code = win_lo + i;
        if (!v) continue;
        if (code >= glo && code - glo < HIST_WIN)
          atomicAdd(&gwin[code - glo], v);
        else
          atomicAdd(&gh[code], v);
      }
    }
  }
}

// Lattice gather / scatter between a grid (shape `fine`) and its every-other
// point lattice (shape ceil(fine / 2)): gather coarse[m] = fine[2m],
// scatter fine[2m] = coarse[m].  One thread per coarse point, C channels.
struct LatticeArgs {
  int32_t ndim, channels;
  int32_t step, pad;  // coarse point m is fine point step * m on every axis
  int64_t fine[LATEC_MAX_DIM], coarse[LATEC_MAX_DIM];
  int64_t nfine, ncoarse;
};

template <bool GATHER>
__global__ void __launch_bounds__(256) lattice_kernel(LatticeArgs a, float *__restrict__ fine,
                                                      float *__restrict__ coarse) {
  // rows of the coarse grid along the last axis; tr lanes per row
  const int nd = a.ndim;
  const uint32_t L = (uint32_t)a.coarse[nd - 1];
  int tr = 1;
  while (tr < 256 && (uint32_t)tr < L) tr <<= 1;
  const int rb = 256 / tr, lrow = threadIdx.x / tr, lcol = threadIdx.x % tr;
  const uint32_t nrows = (uint32_t)(a.ncoarse / L);
  for (uint32_t base = blockIdx.x * rb; base < nrows; base += gridDim.x * rb) {
    const uint32_t row = base + lrow;
    if (row >= nrows) continue;
    uint32_t rem = row;
    uint32_t fbase = 0, fs = (uint32_t)a.fine[nd - 1];
    for (int j = nd - 2; j >= 0; j--) {
      const uint32_t cj = (uint32_t)a.coarse[j];
      const uint32_t m = rem % cj;
      rem /= cj;
      fbase += (uint32_t)a.step * m * fs;
      fs *= (uint32_t)a.fine[j];
    }
    const uint32_t cbase = row * L;
    for (uint32_t col = lcol; col < L; col += tr)
      for (int ch = 0; ch < a.channels; ch++) {
        const uint32_t fi = (uint32_t)(ch * a.nfine) + fbase + (uint32_t)a.step * col;
        const uint32_t ci = (uint32_t)(ch * a.ncoarse) + cbase + col;
        if (GATHER)
          coarse[ci] = fine[fi];
        else
          fine[fi] = coarse[ci];
      }
  }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------

int64_t hist_words(uint32_t R) { return 2 * (int64_t)R + (int64_t)HIST_COPIES * HIST_WIN; }

static int grid_blocks(int64_t work) {
  int64_t b = (work + 255) / 256;
  const int64_t cap = 256 * 16;  // enough waves to fill 256 CUs several times
  if (b > cap) b = cap;
  return (int)(b < 1 ? 1 : b);
}

// Rows of class c whose axis-0 coordinate lies in [lo, hi) (as walk_rows).
static int64_t slab_rows(const ClassDesc &c, int ndim, int64_t n0, int64_t lo, int64_t hi) {
  if (ndim < 2 || (lo <= 0 && hi >= n0)) return c.nrows;
  const int64_t st = c.starts[0], stp = c.step, s0 = c.sizes[0];
  int64_t m_lo = lo <= st ? 0 : (lo - st + stp - 1) / stp;
  int64_t m_hi = hi <= st ? 0 : (hi - st + stp - 1) / stp;
  m_lo = std::min(m_lo, s0);
  m_hi = std::min(m_hi, s0);
  return m_hi > m_lo ? (m_hi - m_lo) * (c.nrows / s0) : 0;
}

// Blocks along x for a stage (or an axis-0 slab of it): enough rows of every
// class, capped, rounded to a multiple of the 8 XCDs.
static int row_blocks(const StageDesc &st, int ndim = 1, int64_t n0 = 0, int64_t lo = 0,
                      int64_t hi = 0) {
  int64_t b = 1;
  for (auto &c : st.classes) {
    const int64_t rb = 256 / c.tr;
    const int64_t rows = hi > lo ? slab_rows(c, ndim, n0, lo, hi) : (int64_t)c.nrows;
    b = std::max<int64_t>(b, (rows + rb - 1) / rb);
  }
  b = std::min<int64_t>(b, 8192);
  return (int)((b + 7) / 8 * 8);  // a multiple of the 8 XCDs
}

static int check_opts(const latec_interp_opts *o) {
  if (!o || (o->code_bytes != 2 && o->code_bytes != 4)) {
    set_error("code_bytes must be 2 or 4");
    return LATEC_E_ARG;
  }
  if (o->radius < 1 || (o->code_bytes == 2 && o->radius > 32768) || o->radius > (1u << 30)) {
    set_error("radius does not fit the code width");
    return LATEC_E_ARG;
  }
  if (o->round_mode == 2 && !(o->span > 0)) {
    set_error("round_output span must be > 0");
    return LATEC_E_ARG;
  }
  return LATEC_OK;
}

static GridArgs make_args(const latec_plan *p, int stage, double eb,
                          const latec_interp_opts *o) {
  GridArgs g;
  memset(&g, 0, sizeof g);
  const StageDesc &st = p->stages[stage];
  g.ndim = p->ndim;
  g.channels = p->channels;
  g.s = st.stride;
  g.stage_stride = p->stride;
  g.block = p->block;
  g.cubic = o->cubic;
  g.center = o->center;
  g.end_mode = o->end_mode;
  g.round_mode = o->round_mode;
  g.R = o->radius;
  g.N = p->npoints;
  g.n = st.count;
  g.code_base = (int64_t)p->channels * st.offset;
  for (int j = 0; j < p->ndim; j++) {
    g.shape[j] = p->shape[j];
    g.gstr[j] = p->gstr[j];
  }
  g.eb = eb;
  g.w = 2.0 * eb;
  g.inv_w = 1.0 / g.w;
  g.feb = (float)eb;
  g.span = o->span;
  g.offset = o->offset;
  g.n_excl = (int64_t)st.excluded.size();
  g.slab_lo = 0;
  g.slab_hi = p->shape[0];
  return g;
}

// Kernel selection: a compile-time rank with 32-bit grid indices for fields
// below 2^32 values and ranks 1..4, the generic kernel otherwise.
enum class KSel { Fast, Generic };
static KSel pick(const latec_plan *p) {
  const bool small = (int64_t)p->channels * p->npoints < 0xFFFFFFFFll;
  return (small && p->ndim <= 4) ? KSel::Fast : KSel::Generic;
}

#define LATEC_DISPATCH_ND(KERNEL, BLK1, CODET, ...)                                      \
  switch (p->ndim) {                                                                     \
    case 1: hipLaunchKernelGGL((KERNEL<1, uint32_t, BLK1, CODET>), __VA_ARGS__); break;  \
    case 2: hipLaunchKernelGGL((KERNEL<2, uint32_t, BLK1, CODET>), __VA_ARGS__); break;  \
    case 3: hipLaunchKernelGGL((KERNEL<3, uint32_t, BLK1, CODET>), __VA_ARGS__); break;  \
    default: hipLaunchKernelGGL((KERNEL<4, uint32_t, BLK1, CODET>), __VA_ARGS__); break; \
  }

static bool pyramid_ok(const latec_plan *p);
static int pyramid_run(latec_plan *p, bool encode, const float *x, float *recon, void *codes,
                       uint32_t *hist, uint32_t *out_rank, float *out_val,
                       uint32_t *out_count, const double *ebs, const latec_interp_opts *o,
                       hipStream_t stream);

int interp_encode(latec_plan *p, int s0, int s1, const float *x, float *recon,
                  void *codes, uint32_t *hist, uint32_t *out_rank,
                  float *out_val, uint32_t *out_count, const double *ebs,
                  const latec_interp_opts *o, hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if (s0 < 0 || s1 > (int)p->stages.size() || s0 > s1) {
    set_error("stage range out of bounds");
    return LATEC_E_ARG;
  }
  if ((rc = plan_upload(p))) return rc;
  if (s0 == 0 && s1 == (int)p->stages.size() && pyramid_ok(p) && !getenv("LATEC_NO_PYRAMID")) {
    for (int s = 0; s < s1; s++)
      if (p->stages[s].count && !(ebs[s] > 0)) {
        set_error("error bound must be > 0");
        return LATEC_E_ARG;
      }
    return pyramid_run(p, true, x, recon, codes, hist, out_rank, out_val, out_count, ebs, o,
                       stream);
  }
  const int64_t twoR = hist_words(o->radius);
  const KSel sel = pick(p);
  for (int s = s0; s < s1; s++) {
    const StageDesc &st = p->stages[s];
    if (!st.count) continue;
    if (!(ebs[s] > 0)) {
      set_error("error bound must be > 0");
      return LATEC_E_ARG;
    }
    GridArgs g = make_args(p, s, ebs[s], o);
    dim3 grid(row_blocks(st), (unsigned)st.classes.size());
    const ClassDesc *cls = p->dev->classes + p->dev->class_off[s];
    const uint32_t *ex = p->dev->excluded ? p->dev->excluded + p->dev->excl_off[s] : nullptr;
    const int64_t lo = (int64_t)p->channels * st.offset;
    uint32_t *hs = hist + s * twoR;
    if (sel == KSel::Fast && o->code_bytes == 2) {
      if (p->block == 1) {
        LATEC_DISPATCH_ND(encode_stage_kernel, true, uint16_t, grid, dim3(256), 0, stream, g,
                          cls, ex, x, recon, (uint16_t *)codes, hs, out_rank + lo,
                          out_val + lo, out_count + s, s == 0);
      } else {
        LATEC_DISPATCH_ND(encode_stage_kernel, false, uint16_t, grid, dim3(256), 0, stream, g,
                          cls, ex, x, recon, (uint16_t *)codes, hs, out_rank + lo,
                          out_val + lo, out_count + s, s == 0);
      }
    } else if (o->code_bytes == 2) {
      hipLaunchKernelGGL((encode_stage_kernel<0, int64_t, false, uint16_t>), grid, dim3(256), 0,
                         stream, g, cls, ex, x, recon, (uint16_t *)codes, hs, out_rank + lo,
                         out_val + lo, out_count + s, s == 0);
    } else {
      hipLaunchKernelGGL((encode_stage_kernel<0, int64_t, false, uint32_t>), grid, dim3(256), 0,
                         stream, g, cls, ex, x, recon, (uint32_t *)codes, hs, out_rank + lo,
                         out_val + lo, out_count + s, s == 0);
    }
    if ((rc = hip_check(hipGetLastError(), "encode_stage_kernel"))) return rc;
  }
  return LATEC_OK;
}

int interp_decode(latec_plan *p, int s0, int s1, const void *codes,
                  const uint32_t *out_rank, const float *out_val,
                  const uint32_t *out_count, float *recon, const double *ebs,
                  const latec_interp_opts *o, hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if (s0 < 0 || s1 > (int)p->stages.size() || s0 > s1) {
    set_error("stage range out of bounds");
    return LATEC_E_ARG;
  }
  if ((rc = plan_upload(p))) return rc;
  if (s0 == 0 && s1 == (int)p->stages.size() && pyramid_ok(p) && !getenv("LATEC_NO_PYRAMID"))
    return pyramid_run(p, false, nullptr, recon, (void *)codes, nullptr, (uint32_t *)out_rank,
                       (float *)out_val, (uint32_t *)out_count, ebs, o, stream);
  const KSel sel = pick(p);
  for (int s = s0; s < s1; s++) {
    const StageDesc &st = p->stages[s];
    if (!st.count) continue;
    GridArgs g = make_args(p, s, ebs[s], o);
    dim3 grid(row_blocks(st), (unsigned)st.classes.size());
    const ClassDesc *cls = p->dev->classes + p->dev->class_off[s];
    const uint32_t *ex = p->dev->excluded ? p->dev->excluded + p->dev->excl_off[s] : nullptr;
    const int64_t lo = (int64_t)p->channels * st.offset;
    if (sel == KSel::Fast && o->code_bytes == 2) {
      if (p->block == 1) {
        LATEC_DISPATCH_ND(decode_stage_kernel, true, uint16_t, grid, dim3(256), 0, stream, g,
                          cls, ex, (const uint16_t *)codes, out_rank + lo, out_val + lo,
                          out_count + s, recon, s == 0);
      } else {
        LATEC_DISPATCH_ND(decode_stage_kernel, false, uint16_t, grid, dim3(256), 0, stream, g,
                          cls, ex, (const uint16_t *)codes, out_rank + lo, out_val + lo,
                          out_count + s, recon, s == 0);
      }
    } else if (o->code_bytes == 2) {
      hipLaunchKernelGGL((decode_stage_kernel<0, int64_t, false, uint16_t>), grid, dim3(256), 0,
                         stream, g, cls, ex, (const uint16_t *)codes, out_rank + lo,
                         out_val + lo, out_count + s, recon, s == 0);
    } else {
      hipLaunchKernelGGL((decode_stage_kernel<0, int64_t, false, uint32_t>), grid, dim3(256), 0,
                         stream, g, cls, ex, (const uint32_t *)codes, out_rank + lo,
                         out_val + lo, out_count + s, recon, s == 0);
    }
    if ((rc = hip_check(hipGetLastError(), "decode_stage_kernel"))) return rc;
  }
  return LATEC_OK;
}


// ---------------------------------------------------------------------------
// lattice pyramid + slab wavefront
// ---------------------------------------------------------------------------

static bool pyramid_ok(const latec_plan *p) {
  return p->block == 1 && p->ndim <= 4 && (int64_t)p->channels * p->npoints < 0xFFFFFFFFll;
}

static size_t level3_smem();
static size_t slab_smem(bool encode);

static int pyramid_setup(latec_plan *p) {
  int rc = plan_upload(p);
  if (rc) return rc;
  DeviceBufs *d = p->dev;
  if (d->K) return LATEC_OK;
  int K = 0;
  while ((1 << (K + 1)) <= p->stride) K++;
  K = std::min(K, p->levels);
  d->sub.assign(K + 1, nullptr);
  d->X.assign(K + 1, nullptr);
  d->R.assign(K + 1, nullptr);
  d->npts.assign(K + 1, 0);
  for (int k = 1; k <= K; k++) {
    const int64_t sk = (int64_t)p->stride >> k;
    auto *q = new latec_plan();
    q->ndim = p->ndim;
    q->channels = p->channels;
    q->levels = 1;
    q->stride = 2;
    q->block = 1;
    for (int j = 0; j < p->ndim; j++) q->shape[j] = (p->shape[j] + sk - 1) / sk;
    if ((rc = plan_build(q)) || (rc = plan_upload(q))) {
      delete q;
      return rc;
    }
    d->sub[k] = q;
    d->npts[k] = q->npoints;
    if (k < K) {
      const size_t bytes = sizeof(float) * (size_t)p->channels * q->npoints;
      if ((rc = hip_check(hipMalloc(&d->X[k], bytes), "hipMalloc")) ||
          (rc = hip_check(hipMalloc(&d->R[k], bytes), "hipMalloc")))
        return rc;
    }
  }
  d->K = K;
  if (p->ndim == 3) {  // the slab kernel's planes exceed the default LDS limit
    if ((rc = hip_check(hipFuncSetAttribute((const void *)slab_kernel<true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)slab_smem(true)), "hipFuncSetAttribute")) ||
        (rc = hip_check(hipFuncSetAttribute((const void *)slab_kernel<false>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)slab_smem(false)), "hipFuncSetAttribute")))
      return rc;
  }
  if (p->ndim == 3) {  // the level kernel's window exceeds the default LDS limit
    if ((rc = hip_check(hipFuncSetAttribute((const void *)level3_kernel<true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)level3_smem()), "hipFuncSetAttribute")) ||
        (rc = hip_check(hipFuncSetAttribute((const void *)level3_kernel<false>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)level3_smem()), "hipFuncSetAttribute")))
      return rc;
  }
  return LATEC_OK;
}

static int lattice(const latec_plan *fine_p, const latec_plan *coarse_p, float *fine,
                   float *coarse, bool gather, hipStream_t stream, int64_t step = 2) {
  LatticeArgs a;
  memset(&a, 0, sizeof a);
  a.ndim = fine_p->ndim;
  a.channels = fine_p->channels;
  a.step = (int32_t)step;
  for (int j = 0; j < a.ndim; j++) {
    a.fine[j] = fine_p->shape[j];
    a.coarse[j] = coarse_p->shape[j];
  }
  a.nfine = fine_p->npoints;
  a.ncoarse = coarse_p->npoints;
  const int64_t L = std::max<int64_t>(coarse_p->shape[a.ndim - 1], 1);
  int tr = 1;
  while (tr < 256 && tr < L) tr <<= 1;
  const int64_t rows = a.ncoarse / L, rb = 256 / tr;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((rows + rb - 1) / rb, 8192));
  if (gather)
    hipLaunchKernelGGL(lattice_kernel<true>, dim3(blocks), dim3(256), 0, stream, a, fine, coarse);
  else
    hipLaunchKernelGGL(lattice_kernel<false>, dim3(blocks), dim3(256), 0, stream, a, fine, coarse);
  return hip_check(hipGetLastError(), "lattice_kernel");
}

// One stage of a sub-plan addressed into the global plan's arrays.
static int launch_stage(const latec_plan *gp, const latec_plan *q, int js, int gs, bool encode,
                        const float *x, float *recon, void *codes, uint32_t *hist,
                        uint32_t *out_rank, float *out_val, uint32_t *out_count,
                        const double *ebs, const latec_interp_opts *o, int64_t slab_lo,
                        int64_t slab_hi, hipStream_t stream) {
  const StageDesc &st = q->stages[js];
  if (!st.count) return LATEC_OK;
  GridArgs g = make_args(q, js, ebs[gs], o);
  const StageDesc &gst = gp->stages[gs];
  g.code_base = (int64_t)gp->channels * gst.offset;
  g.slab_lo = slab_lo;
  g.slab_hi = slab_hi;
  dim3 grid(row_blocks(st, q->ndim, q->shape[0], slab_lo, slab_hi),
            (unsigned)st.classes.size());
  const ClassDesc *cls = q->dev->classes + q->dev->class_off[js];
  const int64_t lo = (int64_t)gp->channels * gst.offset;
  const int64_t twoR = hist_words(o->radius);
  const latec_plan *p = q;  // for LATEC_DISPATCH_ND
  if (encode) {
    if (o->code_bytes == 2) {
      LATEC_DISPATCH_ND(encode_stage_kernel, true, uint16_t, grid, dim3(256), 0, stream, g, cls,
                        nullptr, x, recon, (uint16_t *)codes, hist + gs * twoR, out_rank + lo,
                        out_val + lo, out_count + gs, gs == 0);
    } else {
      hipLaunchKernelGGL((encode_stage_kernel<0, int64_t, false, uint32_t>), grid, dim3(256), 0,
                         stream, g, cls, nullptr, x, recon, (uint32_t *)codes, hist + gs * twoR,
                         out_rank + lo, out_val + lo, out_count + gs, gs == 0);
    }
  } else {
    if (o->code_bytes == 2) {
      LATEC_DISPATCH_ND(decode_stage_kernel, true, uint16_t, grid, dim3(256), 0, stream, g, cls,
                        nullptr, (const uint16_t *)codes, out_rank + lo, out_val + lo,
                        out_count + gs, recon, gs == 0);
    } else {
      hipLaunchKernelGGL((decode_stage_kernel<0, int64_t, false, uint32_t>), grid, dim3(256), 0,
                         stream, g, cls, nullptr, (const uint32_t *)codes, out_rank + lo,
                         out_val + lo, out_count + gs, recon, gs == 0);
    }
  }
  return hip_check(hipGetLastError(), encode ? "encode_stage_kernel" : "decode_stage_kernel");
}

// Level-fused launch for 3-D, one channel, u16 codes: all five sub-stages of
// level k in one kernel (level3_kernel).  Returns 1 if not applicable.
static const int L3_T[3] = {16, 16, 16};
// largest number of points one tile computes (weight-1 classes over two
// extended axes, weight-2 over one, weight-3 over the tile)
static size_t level3_points() {
  const size_t h0 = L3_T[0] / 2, h1 = L3_T[1] / 2, h2 = L3_T[2] / 2;
  const size_t e0 = h0 + 3, e1 = h1 + 3, e2 = h2 + 3;
  return h0 * e1 * e2 + e0 * h1 * e2 + e0 * e1 * h2 + h0 * h1 * e2 + h0 * e1 * h2 +
         e0 * h1 * h2 + h0 * h1 * h2;
}
static size_t level3_smem() {
  return 5 * L3_HBINS * sizeof(uint32_t) +
         sizeof(float) * ((size_t)(L3_T[0] + 6) * (L3_T[1] + 6) * (L3_T[2] + 6) + level3_points());
}
static bool level3_ok(const latec_plan *gp, const latec_plan *q, const latec_interp_opts *o) {
  static const bool on = getenv("LATEC_LEVEL3") != nullptr;  // opt-in until it beats the stage kernels
  if (!on) return false;
  return gp->ndim == 3 && gp->channels == 1 && o->code_bytes == 2 && q->stages.size() == 6;
}
static int launch_level3(const latec_plan *gp, const latec_plan *q, int g0, bool encode,
                         const float *x, float *recon, void *codes, uint32_t *hist,
                         uint32_t *out_rank, float *out_val, uint32_t *out_count,
                         const double *ebs, const latec_interp_opts *o, hipStream_t stream) {
  Level3Args A;
  memset(&A, 0, sizeof A);
  for (int j = 0; j < 3; j++) {
    A.n[j] = q->shape[j];
    A.T[j] = L3_T[j];
    A.tiles[j] = (q->shape[j] + L3_T[j] - 1) / L3_T[j];
  }
  const int64_t tiles = A.tiles[0] * A.tiles[1] * A.tiles[2];
  if (tiles > 0x7fffffffll) return 1;
  A.tiles_total = (int32_t)tiles;
  A.cubic = o->cubic;
  A.center = o->center;
  A.end_mode = o->end_mode;
  A.round_mode = o->round_mode;
  A.span = o->span;
  A.offset = o->offset;
  A.hist_words = hist_words(o->radius);
  for (int js = 1; js <= 5; js++) {
    const int gs = g0 + js - 1;
    const GridArgs g = make_args(q, js, ebs[gs], o);
    L3Stage &S = A.st[js - 1];
    S.w = g.w;
    S.inv_w = g.inv_w;
    S.feb = g.feb;
    S.R = g.R;
    S.code_base = (int64_t)gp->channels * gp->stages[gs].offset;
    S.out_lo = S.code_base;
    S.gs = gs;
    for (const ClassDesc &c : q->stages[js].classes) {
      if (c.parity < 1 || c.parity > 7 || c.ndim != 3 || c.step != 2) return 1;
      L3Class &K = A.cl[c.parity];
      K.offset = c.offset;
      K.step = c.step;
      for (int j = 0; j < 3; j++) {
        K.start[j] = c.starts[j];
        K.rstr[j] = c.rstr[j];
      }
    }
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  int64_t grid = std::min<int64_t>(tiles, (int64_t)ncu * 2);
  grid = std::max<int64_t>(8, (grid + 7) / 8 * 8);
  if (encode)
    hipLaunchKernelGGL(level3_kernel<true>, dim3((unsigned)grid), dim3(L3_THREADS), level3_smem(),
                       stream, A, x, recon, (uint16_t *)codes, hist, out_rank, out_val,
                       out_count);
  else
    hipLaunchKernelGGL(level3_kernel<false>, dim3((unsigned)grid), dim3(L3_THREADS),
                       level3_smem(), stream, A, nullptr, recon, (uint16_t *)codes, hist,
                       out_rank, out_val, out_count);
  return hip_check(hipGetLastError(), "level3_kernel");
}

// Launch-shape knobs of the march kernels (set by the autotuner through
// latec_tune_set): at most march_blocks workgroups per class (a multiple of
// 8, so the XCD remap stays exact; the threads walk the rest of the class),
// and march_hist_peel, the distinct codes per wave counted with one atomic.
static int64_t g_march_blocks = 2048;
static int64_t g_slab = 0;  // "slab": 1 runs large levels as one level-fused launch
static int64_t g_pair = 1;  // "pair": 1 runs 3-D levels above the first as line-pair launches
// "pair_lines": 1 gives phases 2 and 3 whole lines per workgroup when a line
// has at most 256 pairs (LDS exchange), 0 always uses wave runs
static int64_t g_pair_lines = 1;
// "pair_quads": bit 0 (encode) and bit 1 (decode) run phases 2 and 3 with
// four points per thread when the lines allow it (length a multiple of 4, at
// most 1024); bits 2 (encode) and 3 (decode) do so for phase 2 alone.
// Measured at 512^3: decode 15 % faster, encode 4 % slower (its larger
// register set costs a wave per SIMD), so by default decode only; quads in
// phase 2 alone: encode 1 % slower, decode 3 % slower than quads in both.
static int64_t g_pair_quads = 2;
// "min_threads": a march or pair launch shortens its per-thread march (8, 4,
// 2, 1 steps) until it has at least this many threads
static int64_t g_min_threads = 1 << 19;
// "pair_waves": 1 runs the finest level's encode phases 2 and 3 (the launches
// that read x at its own grid) as the variants compiled for one more wave per
// SIMD than the compiler's own choice (a smaller register budget: more waves
// in flight, more spills); 0 the compiler's own occupancy
static int64_t g_pair_waves = 0;
static int64_t g_march_peel = 0;  // histogram: distinct codes per wave counted by ballot
// "sheet": bit 0 (encode) and bit 1 (decode) run a 3-D level above the first
// as two sheet launches (even sheets, then odd sheets; latec_sheet.hip) when
// its lines have an even length of at most 512 points
static int64_t g_sheet = 3;
// "sheet_blocks": a sheet launch halves its march segment (16, 8, 4 odd
// rows) while it would have fewer workgroups than this
static int64_t g_sheet_blocks = 1024;
// "sheet_min_points": the smallest level (points of its compact grid) that
// runs as sheet launches; smaller levels keep the pair launches, which
// measure faster there (more parallelism per point)
static int64_t g_sheet_min = 1ll << 25;
// "sheet_launches" (read, or set to reset): sheet launches made so far
static std::atomic<int64_t> g_sheet_launches{0};

static int64_t g_tune_epoch = 0;  // bumped by every knob change
static int64_t tune_epoch() { return g_tune_epoch; }
// "quad_launches" (read, or set to reset): line-quad launches made so far, so a
// caller can tell whether a level ran as quads (launches recorded into a graph
// count once, when recorded)
static std::atomic<int64_t> g_quad_launches{0};

int gnn_tune_set(const char *name, int64_t v);  // latec_gnn.hip
int64_t gnn_tune_get(const char *name);
int tune_set(const char *name, int64_t v) {
  if (!strcmp(name, "quad_launches")) {
    g_quad_launches = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "sheet_launches")) {
    g_sheet_launches = v;
    return LATEC_OK;
  }
  g_tune_epoch++;
  if (!strcmp(name, "sheet")) {
    if (v < 0 || v > 3) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_sheet = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "sheet_min_points")) {
    if (v < 0 || v > (1ll << 40)) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_sheet_min = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "sheet_blocks")) {
    if (v < 1 || v > (1 << 20)) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_sheet_blocks = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "march_blocks")) {
    if (v < 8 || v > (1 << 20)) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_march_blocks = (v + 7) / 8 * 8;
    return LATEC_OK;
  }
  if (!strcmp(name, "slab")) {
    if (v < 0 || v > 1) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_slab = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "pair")) {
    if (v < 0 || v > 1) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_pair = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "pair_quads")) {
    if (v < 0 || v > 15) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_pair_quads = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "pair_waves")) {
    if (v < 0 || v > 1) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_pair_waves = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "pair_lines")) {
    if (v < 0 || v > 1) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_pair_lines = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "min_threads")) {
    if (v < 0 || v > (1 << 26)) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_min_threads = v;
    return LATEC_OK;
  }
  if (!strcmp(name, "march_hist_peel")) {
    if (v < 0 || v > 8) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_march_peel = v;
    return LATEC_OK;
  }
  if (!strncmp(name, "gnn_", 4)) return gnn_tune_set(name, v);
  set_error(std::string("unknown tuning knob ") + name);
  return LATEC_E_ARG;
}

int64_t tune_get(const char *name) {
  if (!strcmp(name, "march_blocks")) return g_march_blocks;
  if (!strcmp(name, "march_hist_peel")) return g_march_peel;
  if (!strcmp(name, "slab")) return g_slab;
  if (!strcmp(name, "pair")) return g_pair;
  if (!strcmp(name, "min_threads")) return g_min_threads;
  if (!strcmp(name, "pair_lines")) return g_pair_lines;
  if (!strcmp(name, "pair_quads")) return g_pair_quads;
  if (!strcmp(name, "quad_launches")) return g_quad_launches;
  if (!strcmp(name, "pair_waves")) return g_pair_waves;
  if (!strcmp(name, "sheet")) return g_sheet;
  if (!strcmp(name, "sheet_blocks")) return g_sheet_blocks;
  if (!strcmp(name, "sheet_min_points")) return g_sheet_min;
  if (!strcmp(name, "sheet_launches")) return g_sheet_launches;
  if (!strncmp(name, "gnn_", 4)) return gnn_tune_get(name);
  set_error(std::string("unknown tuning knob ") + name);
  return LATEC_E_ARG;
}

// Marching class launches for a 3-D level (one channel, u16 codes): the
// three weight-1 stages together, then the weight-2 stage, then weight 3.
// Returns 1 if the level's layout is not the one the kernel assumes.
static bool march_ok(const latec_plan *gp, const latec_plan *q, const latec_interp_opts *o) {
  static const bool off = getenv("LATEC_NO_MARCH") != nullptr;
  if (off) return false;
  return gp->ndim == 3 && gp->channels == 1 && o->code_bytes == 2 && q->stages.size() == 6 &&
         q->npoints < (1ll << 30);  // byte offsets of floats fit 32 bits
}

static int launch_march(const latec_plan *gp, const latec_plan *q, int g0, bool encode,
                        const float *x, float *recon, void *codes, uint32_t *hist,
                        uint32_t *out_rank, float *out_val, uint32_t *out_count,
                        const double *ebs, const latec_interp_opts *o, hipStream_t stream,
                        const latec_plan *qc = nullptr, const float *coarse = nullptr,
                        bool write_coarse = false) {
  static const int groups[3][3] = {{1, 2, 3}, {4, 0, 0}, {5, 0, 0}};  // sub-plan stages
  MArgs A[3];
  for (int gi = 0; gi < 3; gi++) {
    MArgs &a = A[gi];
    memset(&a, 0, sizeof a);
    for (int j = 0; j < 3; j++) a.n[j] = (int32_t)q->shape[j];
    a.cubic = o->cubic;
    a.center = o->center;
    a.end_mode = o->end_mode;
    a.round_mode = o->round_mode;
    a.span = o->span;
    a.offset = o->offset;
    a.hist_words = hist_words(o->radius);
    a.peel = (int32_t)g_march_peel;
    a.steps = MARCH_M;
    if (gi == 0 && coarse) {  // the weight-1 stages read their taps from the level below
      a.coarse = coarse;
      for (int j = 0; j < 3; j++) {
        a.cn[j] = (int32_t)qc->shape[j];
        if (qc->shape[j] != (q->shape[j] + 1) / 2) return 1;
      }
      a.write_coarse = write_coarse ? 1 : 0;
    }
    int ns = 0;
    for (int si = 0; si < 3 && groups[gi][si]; si++) {
      const int js = groups[gi][si], gs = g0 + js - 1;
      const StageDesc &st = q->stages[js];
      const GridArgs g = make_args(q, js, ebs[gs], o);
      L3Stage &S = a.st[ns];
      S.w = g.w;
      S.inv_w = g.inv_w;
      S.feb = g.feb;
      S.R = g.R;
      S.code_base = (int64_t)gp->channels * gp->stages[gs].offset;
      S.out_lo = S.code_base;
      S.gs = gs;
      for (const ClassDesc &c : st.classes) {
        if (c.parity < 1 || c.parity > 7 || c.ndim != 3 || c.step != 2 || (c.flags & 1))
          return 1;
        for (int j = 0; j < 3; j++)
          if (c.starts[j] != ((c.parity >> j) & 1)) return 1;
        if (!c.size) continue;
        if (a.ncls == 3) return 1;
        MClass &K = a.cl[a.ncls++];
        K.parity = c.parity;
        K.stage = ns;
        const int ma = (c.parity & 1) ? 0 : (c.parity & 2) ? 1 : 0;
        for (int j = 0; j < 3; j++) {
          K.m[j] = c.sizes[j];
          K.e[j] = c.sizes[j];
          K.rstr[j] = c.rstr[j];
        }

        K.offset = c.offset;
      }
      ns++;
    }
  }
  // points per thread: MARCH_M, halved while the launch would have fewer
  // than g_min_threads threads
  for (int gi = 0; gi < 3; gi++) {
    MArgs &a = A[gi];
    auto threads = [&](int64_t st) {
      int64_t t = 0;
      for (int i = 0; i < a.ncls; i++) {
        const MClass &K = a.cl[i];
        const int ma = (K.parity & 1) ? 0 : (K.parity & 2) ? 1 : 0;
        int64_t nt = 1;
        for (int j = 0; j < 3; j++) nt *= j == ma ? ((int64_t)K.m[j] + st - 1) / st : K.m[j];
        t += nt;
      }
      return t;
    };
    int64_t S = MARCH_M;
    while (S > 1 && threads(S) < g_min_threads) S >>= 1;
    a.steps = (int32_t)S;
    for (int i = 0; i < a.ncls; i++) {
      MClass &K = a.cl[i];
      const int ma = (K.parity & 1) ? 0 : (K.parity & 2) ? 1 : 0;
      K.e[ma] = (uint32_t)(((int64_t)K.m[ma] + S - 1) / S);
      const int64_t nt = (int64_t)K.e[0] * K.e[1] * K.e[2];
      if (nt >= 0x7fffffffll) return 1;
      K.nthreads = (uint32_t)nt;
    }
  }
  for (int gi = 0; gi < 3; gi++) {
    const MArgs &a = A[gi];
    if (!a.ncls) continue;
    int64_t blocks = 1;
    for (int i = 0; i < a.ncls; i++)
      blocks = std::max<int64_t>(blocks, ((int64_t)a.cl[i].nthreads + 255) / 256);
    blocks = std::min<int64_t>(blocks, g_march_blocks);
    blocks = (blocks + 7) / 8 * 8;
    const dim3 grid((unsigned)blocks, (unsigned)a.ncls);
    if (encode)
      hipLaunchKernelGGL(march_kernel<true>, grid, dim3(256), 0, stream, a, x, recon,
                         (uint16_t *)codes, hist, out_rank, out_val, out_count);
    else
      hipLaunchKernelGGL(march_kernel<false>, grid, dim3(256), 0, stream, a, nullptr, recon,
                         (uint16_t *)codes, nullptr, out_rank, out_val, out_count);
    const int rc = hip_check(hipGetLastError(), "march_kernel");
    if (rc) return rc;
  }
  return LATEC_OK;
}

// LDS of the slab kernel: ring of 4 even planes + the odd plane (+ the x of
// both planes and the stage histogram windows when encoding).
static size_t slab_smem(bool encode) {
  return sizeof(float) * (size_t)SL_PLANE * (encode ? 7 : 5) +
         (encode ? sizeof(uint32_t) * 5 * SL_HB : 0);
}

// Level-fused slab launch for a large 3-D level (one channel, u16 codes):
// all five sub-stages of level k in one launch, coarse points read from the
// compact grid of level k-1.  Returns 1 if not applicable.
static bool slab_ok(const latec_plan *gp, const latec_plan *q, int k,
                    const latec_interp_opts *o) {
  if (!g_slab || k < 2 || gp->ndim != 3 || gp->channels != 1 || o->code_bytes != 2 ||
      q->stages.size() != 6 || q->npoints >= 0x7fffffffll)
    return false;
  return q->shape[0] >= 2 * SL_SI && q->shape[1] >= 2 * SL_TJ && q->shape[2] >= 2 * SL_TK;
}

static int launch_slab(const latec_plan *gp, const latec_plan *q, const latec_plan *qc, int g0,
                       bool encode, const float *x, const float *coarse, float *recon,
                       void *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                       uint32_t *out_count, const double *ebs, const latec_interp_opts *o,
                       hipStream_t stream) {
  SlabArgs A;
  memset(&A, 0, sizeof A);
  for (int j = 0; j < 3; j++) {
    A.n[j] = (int32_t)q->shape[j];
    A.cn[j] = (int32_t)qc->shape[j];
    if (qc->shape[j] != (q->shape[j] + 1) / 2) return 1;
  }
  A.tiles_j = (A.n[1] + SL_TJ - 1) / SL_TJ;
  A.tiles_k = (A.n[2] + SL_TK - 1) / SL_TK;
  A.segs = (A.n[0] + SL_SI - 1) / SL_SI;
  const int64_t nb = (int64_t)A.tiles_j * A.tiles_k * A.segs;
  if (nb > 0x7fffffffll) return 1;
  A.nblocks = (int32_t)nb;
  A.cubic = o->cubic;
  A.center = o->center;
  A.end_mode = o->end_mode;
  A.round_mode = o->round_mode;
  A.span = o->span;
  A.offset = o->offset;
  A.hist_words = hist_words(o->radius);
  for (int js = 1; js <= 5; js++) {
    const int gs = g0 + js - 1;
    const GridArgs g = make_args(q, js, ebs[gs], o);
    SlabStage &S = A.st[js - 1];
    S.w = g.w;
    S.inv_w = g.inv_w;
    S.feb = g.feb;
    S.R = g.R;
    S.code_base = (int64_t)gp->channels * gp->stages[gs].offset;
    S.gs = gs;
    for (const ClassDesc &c : q->stages[js].classes) {
      if (c.parity < 1 || c.parity > 7 || c.ndim != 3 || c.step != 2 || (c.flags & 1)) return 1;
      for (int j = 0; j < 3; j++)
        if (c.starts[j] != ((c.parity >> j) & 1)) return 1;
      SlabClass &K = A.cl[c.parity];
      K.offset = c.offset;
      for (int j = 0; j < 3; j++) K.rstr[j] = c.rstr[j];
    }
  }
  const unsigned grid = (unsigned)((nb + 7) / 8 * 8);
  const size_t smem = slab_smem(encode);
  if (encode)
    hipLaunchKernelGGL(slab_kernel<true>, dim3(grid), dim3(SL_THREADS), smem, stream, A, x,
                       coarse, recon, (uint16_t *)codes, hist, out_rank, out_val, out_count);
  else
    hipLaunchKernelGGL(slab_kernel<false>, dim3(grid), dim3(SL_THREADS), smem, stream, A,
                       nullptr, coarse, recon, (uint16_t *)codes, hist, out_rank, out_val,
                       out_count);
  return hip_check(hipGetLastError(), "slab_kernel");
}

// Slab wavefront (opt-in, LATEC_SLAB_MB=<MiB per slab>): a level whose grid
// exceeds four slabs runs its sub-stages slab by slab along axis 0 (weight-w
// stages trail by w-1 slabs) so that each slab stays cache resident.

// Line-pair launches for a 3-D level above the first (one channel, u16
// codes): phase 1 on the (e,e) lines, phase 2 on the (o,e) and (e,o) lines,
// phase 3 on the (o,o) lines.  Reads the coarse points from the level
// below's compact grid `coarse` (shape qc), so the level needs no lattice
// scatter.  Returns 1, before launching anything, if the level's layout is
// not the one the kernels assume.
static bool pair_ok(const latec_plan *gp, const latec_plan *q, int k,
                    const latec_interp_opts *o) {
  // byte offsets of recon (buffer stores) stay below 2^31
  return g_pair && k >= 2 && gp->ndim == 3 && gp->channels == 1 && o->code_bytes == 2 &&
         q->stages.size() == 6 && q->npoints < (1ll << 29);
}

// Launches the pair kernel variant of a phase and layout bits (1 encode,
// 2 aligned pairs, 4 rounding check, 8 line blocks, 16 x is the level grid);
// only the combinations that occur are instantiated.
template <int B>
struct PairLaunch {
  static constexpr bool E = B & 1, AL = B & 2, RND = B & 4, BM = B & 8, XD = B & 16;
  static constexpr bool valid = E || (!RND && XD);
  template <int PH>
  static void one(dim3 grid, hipStream_t stream, const PArgs &a, const float *x, float *recon,
                  uint16_t *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                  uint32_t *out_count) {
    if constexpr (valid && E && XD && PH > 1) {
      if (g_pair_waves) {
        hipLaunchKernelGGL((pair_kernel<PH, E, AL, RND, BM, XD, PairWavesAlt<PH>::value>), grid, dim3(256),
                           0, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
        return;
      }
    }
    if constexpr (valid && (PH > 1 || !BM))
      hipLaunchKernelGGL((pair_kernel<PH, E, AL, RND, BM, XD>), grid, dim3(256), 0, stream, a,
                         x, recon, codes, hist, out_rank, out_val, out_count);
  }
  static void go(int ph, int bits, dim3 grid, hipStream_t stream, const PArgs &a,
                 const float *x, float *recon, uint16_t *codes, uint32_t *hist,
                 uint32_t *out_rank, float *out_val, uint32_t *out_count) {
    if (bits != B) {
      PairLaunch<B - 1>::go(ph, bits, grid, stream, a, x, recon, codes, hist, out_rank,
                            out_val, out_count);
    } else if (ph == 1) {
      one<1>(grid, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
    } else if (ph == 2) {
      one<2>(grid, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
    } else {
      one<3>(grid, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
    }
  }
};
template <>
struct PairLaunch<-1> {
  static void go(int, int, dim3, hipStream_t, const PArgs &, const float *, float *, uint16_t *,
                 uint32_t *, uint32_t *, float *, uint32_t *) {}
};

// The line-quad kernel variant of a phase (2 or 3) and bits 1 encode,
// 2 rounding check, 4 x is the level grid.
template <int B>
struct QuadLaunch {
  static constexpr bool E = B & 1, RND = B & 2, XD = B & 4;
  static constexpr bool valid = E || (!RND && XD);
  template <int PH>
  static void one(dim3 grid, hipStream_t stream, const PArgs &a, const float *x, float *recon,
                  uint16_t *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                  uint32_t *out_count) {
    if constexpr (valid)
      hipLaunchKernelGGL((quad_kernel<PH, E, RND, XD>), grid, dim3(256), 0, stream, a, x,
                         recon, codes, hist, out_rank, out_val, out_count);
  }
  static void go(int ph, int bits, dim3 grid, hipStream_t stream, const PArgs &a,
                 const float *x, float *recon, uint16_t *codes, uint32_t *hist,
                 uint32_t *out_rank, float *out_val, uint32_t *out_count) {
    if (bits != B)
      QuadLaunch<B - 1>::go(ph, bits, grid, stream, a, x, recon, codes, hist, out_rank,
                            out_val, out_count);
    else if (ph == 2)
      one<2>(grid, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
    else
      one<3>(grid, stream, a, x, recon, codes, hist, out_rank, out_val, out_count);
  }
};
template <>
struct QuadLaunch<-1> {
  static void go(int, int, dim3, hipStream_t, const PArgs &, const float *, float *, uint16_t *,
                 uint32_t *, uint32_t *, float *, uint32_t *) {}
};

static int launch_pair(const latec_plan *gp, const latec_plan *q, const latec_plan *qc, int g0,
                       bool encode, const float *x, int64_t xstep, const float *coarse,
                       float *recon, void *codes, uint32_t *hist, uint32_t *out_rank,
                       float *out_val, uint32_t *out_count, const double *ebs,
                       const latec_interp_opts *o, hipStream_t stream, bool launch = true) {
  for (int j = 0; j < 3; j++)
    if (qc->shape[j] != (q->shape[j] + 1) / 2) return 1;
  // the sub-plan's classes by parity, in their stages: 100, 010, 001 alone,
  // then the weight-2 stage, then 111
  static const int want_stage[8] = {0, 1, 2, 4, 3, 4, 4, 5};
  const ClassDesc *cls[8] = {};
  for (int js = 1; js <= 5; js++)
    for (const ClassDesc &c : q->stages[js].classes) {
      if (c.parity < 1 || c.parity > 7 || c.ndim != 3 || c.step != 2 || (c.flags & 1) ||
          want_stage[c.parity] != js || cls[c.parity])
        return 1;
      for (int j = 0; j < 3; j++)
        if (c.starts[j] != ((c.parity >> j) & 1)) return 1;
      cls[c.parity] = &c;
    }
  for (int pb = 1; pb < 8; pb++)
    if (!cls[pb]) return 1;
  const int64_t n[3] = {q->shape[0], q->shape[1], q->shape[2]};
  const uint32_t npairs = (uint32_t)((n[2] + 1) / 2);
  const uint32_t nq = (uint32_t)(n[2] / 4);
  const bool quad_lines = g_pair_lines && n[2] % 4 == 0 && nq >= 1 && nq <= 256;
  const int qb = encode ? 0 : 1;
  // phases 2 (ph 1) and 3 (ph 2) as line quads
  const bool quads_ph[3] = {false,
                            quad_lines && ((g_pair_quads >> qb) & 1 || (g_pair_quads >> (qb + 2)) & 1),
                            quad_lines && ((g_pair_quads >> qb) & 1)};
  auto half = [&](int a, int par) { return (uint32_t)((n[a] - par + 1) / 2); };
  PArgs A[3];
  for (int ph = 0; ph < 3; ph++) {
    PArgs &a = A[ph];
    memset(&a, 0, sizeof a);
    for (int j = 0; j < 3; j++) {
      a.n[j] = (int32_t)n[j];
      a.cn[j] = (int32_t)qc->shape[j];
    }
    a.coarse = coarse;
    a.cubic = o->cubic;
    a.center = o->center;
    a.end_mode = o->end_mode;
    a.round_mode = o->round_mode;
    a.peel = (int32_t)g_march_peel;
    a.span = o->span;
    a.offset = o->offset;
    a.hist_words = hist_words(o->radius);
    a.npairs = npairs;
    a.own = ph == 0 ? 64u : (uint32_t)PAIR_OWN;
    a.nquads = nq;
    // x is the input grid read at the level's lattice step (xstep 1: the level
    // grid); its element offsets stay below 2^30 (32-bit byte offsets)
    if (gp->npoints >= (1ll << 30)) return 1;
    const int64_t xst[3] = {xstep * gp->shape[1] * gp->shape[2], xstep * gp->shape[2], xstep};
    for (int j = 0; j < 3; j++) {
      if (xstep == 1 && gp->shape[j] != n[j]) return 1;
      if ((n[j] - 1) * xstep >= gp->shape[j] || xst[j] >= (1ll << 30)) return 1;
      a.xs[j] = (uint32_t)xst[j];
    }
  }
  int nst[3] = {0, 0, 0};
  // the stage slot of sub-plan stage js in phase ph's argument block
  auto stage_slot = [&](int ph, int js) -> int32_t {
    PArgs &a = A[ph];
    for (int i = 0; i < nst[ph]; i++)
      if (a.st[i].gs == g0 + js - 1) return i;
    const GridArgs g = make_args(q, js, ebs[g0 + js - 1], o);
    L3Stage &S = a.st[nst[ph]];
    S.w = g.w;
    S.inv_w = g.inv_w;
    S.feb = g.feb;
    S.R = g.R;
    S.code_base = (int64_t)gp->channels * gp->stages[g0 + js - 1].offset;
    S.out_lo = S.code_base;
    S.gs = g0 + js - 1;
    return nst[ph]++;
  };
  auto set_class = [&](int ph, PClass &pc, int parity) {
    const ClassDesc &c = *cls[parity];
    pc.stage = stage_slot(ph, want_stage[parity]);
    pc.offset = c.offset;
    for (int j = 0; j < 3; j++) pc.rstr[j] = c.rstr[j];
  };
  // (ma, ra, pm, pr, even class, odd class) per set; phase 1's even class
  // (the coarse points) codes nothing: its slot repeats the odd class
  static const int sets[4][6] = {{0, 1, 0, 0, 4, 4}, {0, 1, 1, 0, 1, 5}, {1, 0, 1, 0, 2, 6},
                                 {0, 1, 1, 1, 3, 7}};
  static const int phase_of[4] = {0, 1, 1, 2};
  int nset[3] = {0, 0, 0};
  for (int si = 0; si < 4; si++) {
    const int ph = phase_of[si];
    const int *d = sets[si];
    PSet &T = A[ph].set[nset[ph]++];
    T.ma = d[0];
    T.ra = d[1];
    T.pm = d[2];
    T.pr = d[3];
    T.nmarch = half(T.ma, T.pm);
    T.nrows = half(T.ra, T.pr);
    set_class(ph, T.od, d[5]);
    set_class(ph, T.ev, d[4]);
    const int64_t pairs = (int64_t)T.nrows * npairs;
    const int64_t own = A[ph].own;
    const int64_t nw = (pairs + own - 1) / own;
    T.nwaves = (uint32_t)nw;
    if (quads_ph[ph]) {  // line quads: whole lines of quads per workgroup
      T.lpb = 256 / nq;
      T.bpc = (T.nrows + T.lpb - 1) / T.lpb;
    } else if (ph > 0 && g_pair_lines && npairs <= 256) {  // line blocks
      T.lpb = 256 / npairs;
      T.bpc = (T.nrows + T.lpb - 1) / T.lpb;
    }
  }
  // threads of a set for s steps per thread
  auto set_threads = [&](const PSet &T, int64_t st) -> int64_t {
    const int64_t nchunks = ((int64_t)T.nmarch + st - 1) / st;
    return T.lpb ? nchunks * T.bpc * 256 : nchunks * T.nwaves * 64;
  };
  // march steps per thread: PAIR_M, halved while a phase would have fewer
  // than g_min_threads threads (a short level is latency-bound: more threads
  // with shorter marches)
  for (int ph = 0; ph < 3; ph++) {
    PArgs &a = A[ph];
    uint32_t S = PAIR_M;
    auto threads = [&](uint32_t st) {
      int64_t t = 0;
      for (int si = 0; si < nset[ph]; si++) t += set_threads(a.set[si], st);
      return t;
    };
    while (S > 1 && threads(S) < g_min_threads) S >>= 1;
    a.steps = S;
    for (int si = 0; si < nset[ph]; si++) {
      PSet &T = a.set[si];
      const int64_t nt = set_threads(T, S);
      if (nt >= 0x7fffffffll) return 1;
      T.nthreads = (uint32_t)nt;
    }
  }
  if (!launch) return 0;  // the layout check alone
  for (int ph = 0; ph < 3; ph++) {
    const PArgs &a = A[ph];
    int64_t blocks = 1;
    for (int si = 0; si < nset[ph]; si++)
      blocks = std::max<int64_t>(blocks, ((int64_t)a.set[si].nthreads + 255) / 256);
    blocks = std::min<int64_t>(blocks, g_march_blocks);
    blocks = (blocks + 7) / 8 * 8;
    const dim3 grid((unsigned)blocks, (unsigned)nset[ph]);
    // layout bits of the kernel variant (decode: no rounding check, x unused)
    const int bits = (encode ? 1 : 0) | (n[2] % 2 == 0 ? 2 : 0) |
                     (encode && o->round_mode != 0 ? 4 : 0) | (ph > 0 && a.set[0].lpb ? 8 : 0) |
                     (!encode || xstep == 1 ? 16 : 0);
    if (quads_ph[ph]) g_quad_launches++;
    if (quads_ph[ph])
      QuadLaunch<7>::go(ph + 1, bits & 1 | (bits & 4) >> 1 | (bits & 16) >> 2, grid, stream,
                        a, encode ? x : nullptr, recon, (uint16_t *)codes, hist, out_rank,
                        out_val, out_count);
    else
      PairLaunch<31>::go(ph + 1, bits, grid, stream, a, encode ? x : nullptr, recon,
                         (uint16_t *)codes, hist, out_rank, out_val, out_count);
  }
  return hip_check(hipGetLastError(), "pair_kernel");
}

// Sheet launches for a 3-D level above the first (one channel, u16 codes,
// lines of an even length of at most 512): the even sheets, then the odd
// sheets (latec_sheet.hip).  Reads the coarse points from the level below's
// compact grid `coarse` (shape qc).  Returns 1, before launching anything, if
// the level's layout is not the one the kernels assume.
static bool sheet_ok(const latec_plan *gp, const latec_plan *q, int k,
                     const latec_interp_opts *o, bool encode) {
  return ((g_sheet >> (encode ? 0 : 1)) & 1) && q->npoints >= g_sheet_min && o->cubic == 1 &&
         o->center >= 0 &&
         o->center <= 2 && k >= 2 && gp->ndim == 3 && gp->channels == 1 &&
         o->code_bytes == 2 && q->stages.size() == 6 && q->npoints < (1ll << 29) &&
         q->shape[2] % 2 == 0 && q->shape[2] <= 512 && q->shape[0] >= 2 && q->shape[1] >= 2;
}

static int launch_sheets(const latec_plan *gp, const latec_plan *q, const latec_plan *qc, int g0,
                         bool encode, const float *x, int64_t xstep, const float *coarse,
                         float *recon, void *codes, uint32_t *hist, uint32_t *out_rank,
                         float *out_val, uint32_t *out_count, const double *ebs,
                         const latec_interp_opts *o, hipStream_t stream) {
  for (int j = 0; j < 3; j++)
    if (qc->shape[j] != (q->shape[j] + 1) / 2) return 1;
  // the sub-plan's classes by parity, in their stages: 100, 010, 001 alone,
  // then the weight-2 stage, then 111
  static const int want_stage[8] = {0, 1, 2, 4, 3, 4, 4, 5};
  const ClassDesc *cls[8] = {};
  for (int js = 1; js <= 5; js++)
    for (const ClassDesc &c : q->stages[js].classes) {
      if (c.parity < 1 || c.parity > 7 || c.ndim != 3 || c.step != 2 || (c.flags & 1) ||
          want_stage[c.parity] != js || cls[c.parity])
        return 1;
      for (int j = 0; j < 3; j++)
        if (c.starts[j] != ((c.parity >> j) & 1)) return 1;
      cls[c.parity] = &c;
    }
  for (int pb = 1; pb < 8; pb++)
    if (!cls[pb]) return 1;
  const int64_t n[3] = {q->shape[0], q->shape[1], q->shape[2]};
  if (n[2] % 2 || n[2] > 512 || n[0] < 2 || n[1] < 2) return 1;
  if (gp->npoints >= (1ll << 30)) return 1;
  SheetArgs A[2];
  // stage slots (fixed: the kernels index them at compile time) and classes by
  // (row parity, k parity); the even sheets' (e,e) class is the coarse copy
  static const int slot_js[2][3] = {{2, 3, 4}, {1, 4, 5}};
  static const int cls_par[2][4] = {{0, 4, 2, 6}, {1, 5, 3, 7}};
  static const int cls_slot[2][4] = {{0, 1, 0, 2}, {0, 1, 1, 2}};
  const uint32_t npairs = (uint32_t)(n[2] / 2);
  const int64_t xst[3] = {xstep * gp->shape[1] * gp->shape[2], xstep * gp->shape[2], xstep};
  for (int j = 0; j < 3; j++) {
    if (xstep == 1 && gp->shape[j] != n[j]) return 1;
    if ((n[j] - 1) * xstep >= gp->shape[j] || xst[j] >= (1ll << 30)) return 1;
  }
  for (int od = 0; od < 2; od++) {
    SheetArgs &a = A[od];
    memset(&a, 0, sizeof a);
    for (int j = 0; j < 3; j++) {
      a.n[j] = (int32_t)n[j];
      a.cn[j] = (int32_t)qc->shape[j];
      a.xs[j] = (uint32_t)xst[j];
    }
    a.coarse = coarse;
    a.cubic = o->cubic;
    a.center = o->center;
    a.end_mode = o->end_mode;
    a.round_mode = o->round_mode;
    a.span = o->span;
    a.offset = o->offset;
    a.hist_words = hist_words(o->radius);
    a.npairs = npairs;
    a.nsheets = (uint32_t)(od ? n[0] / 2 : (n[0] + 1) / 2);
    // wave runs of 61 pairs over the sheets' lines, four waves per workgroup
    a.nruns = (uint32_t)(((int64_t)a.nsheets * npairs + 60) / 61);
    a.rgroups = (a.nruns + 3) / 4;
    const uint32_t no = (uint32_t)(n[1] / 2);
    uint32_t seg = SHEET_SEG;
    while (seg > 4 && (int64_t)a.rgroups * ((no + seg - 1) / seg) < g_sheet_blocks) seg >>= 1;
    a.seg = seg;
    a.nseg = (no + seg - 1) / seg;
    for (int sl = 0; sl < 3; sl++) {
      const int js = slot_js[od][sl];
      const GridArgs g = make_args(q, js, ebs[g0 + js - 1], o);
      L3Stage &S = a.st[sl];
      S.w = g.w;
      S.inv_w = g.inv_w;
      S.feb = g.feb;
      S.R = g.R;
      S.code_base = (int64_t)gp->channels * gp->stages[g0 + js - 1].offset;
      S.out_lo = S.code_base;
      S.gs = g0 + js - 1;
    }
    for (int c = 0; c < 4; c++) {
      SheetClass &K = a.cl[c];
      K.stage = cls_slot[od][c];
      const int par = cls_par[od][c];
      if (!par) continue;
      if (slot_js[od][K.stage] != want_stage[par]) return 1;
      K.offset = cls[par]->offset;
      for (int j = 0; j < 3; j++) K.rstr[j] = cls[par]->rstr[j];
    }
    if ((int64_t)a.rgroups * a.nseg >= (1ll << 31)) return 1;
  }
  for (int od = 0; od < 2; od++) {
    SheetArgs &a = A[od];
    for (int sl = 0; sl < 3; sl++) a.hoff[sl] = (int64_t)a.st[sl].gs * a.hist_words;
  }
  // layout bits of the kernel variant (decode: no rounding check, x unused)
  const int bits = (encode ? 1 : 0) | (encode && o->round_mode != 0 ? 2 : 0) |
                   (!encode || xstep == 1 ? 4 : 0);
  for (int od = 0; od < 2; od++) {
    g_sheet_launches++;
    const int rc = sheet_launch(A[od], od == 1, bits, encode ? x : nullptr, recon,
                                (uint16_t *)codes, hist, out_rank, out_val, out_count, stream);
    if (rc) return rc < 0 ? rc : LATEC_E_HIP;
  }
  return 0;
}

static int pyramid_run(latec_plan *p, bool encode, const float *x, float *recon, void *codes,
                       uint32_t *hist, uint32_t *out_rank, float *out_val,
                       uint32_t *out_count, const double *ebs, const latec_interp_opts *o,
                       hipStream_t stream) {
  int rc = pyramid_setup(p);
  if (rc) return rc;
  DeviceBufs *d = p->dev;
  const int K = d->K, nd = p->ndim, per = 2 * nd - 1;
  auto X = [&](int k) { return k == K ? x : (const float *)d->X[k]; };
  auto R = [&](int k) { return k == K ? recon : d->R[k]; };
  // levels whose pair launches read x straight from the input grid at their
  // lattice step (no compact copy of x); the others get theirs gathered
  uint64_t direct = 0;
  auto xstep = [&](int k) { return (int64_t)1 << (K - k); };
  for (int k = 2; k < K && k < 64; k++)
    if (!slab_ok(p, d->sub[k], k, o) && pair_ok(p, d->sub[k], k, o) &&
        launch_pair(p, d->sub[k], d->sub[k - 1], 1 + (k - 1) * per, true, x, xstep(k), nullptr, nullptr,
                    nullptr, nullptr, nullptr, nullptr, nullptr, ebs, o, stream, false) == 0)
      direct |= 1ull << k;
  if (encode) {  // lattice pyramid of the input: each gathered level straight from x
    for (int k = K - 1; k >= 1; k--)
      if (!((direct >> k) & 1) &&
          (rc = lattice(p, d->sub[k], (float *)x, d->X[k], true, stream, xstep(k))))
        return rc;
  }
  // anchors: stage 0 of the first level's plan (its even lattice)
  if ((rc = launch_stage(p, d->sub[1], 0, 0, encode, X(1), R(1), codes, hist, out_rank, out_val,
                         out_count, ebs, o, 0, d->sub[1]->shape[0], stream)))
    return rc;
  for (int k = 1; k <= K; k++) {
    const latec_plan *q = d->sub[k];
    const int g0 = 1 + (k - 1) * per;
    if (slab_ok(p, q, k, o)) {  // reads level k-1 itself: no lattice scatter
      rc = launch_slab(p, q, d->sub[k - 1], g0, encode, X(k), d->R[k - 1], R(k), codes, hist,
                       out_rank, out_val, out_count, ebs, o, stream);
      if (rc < 0) return rc;
      if (rc == 0) continue;
    }
    if (sheet_ok(p, q, k, o, encode)) {  // reads level k-1 itself: no lattice scatter
      const bool dx = (direct >> k) & 1;
      rc = launch_sheets(p, q, d->sub[k - 1], g0, encode, dx ? x : X(k), dx ? xstep(k) : 1,
                         d->R[k - 1], R(k), codes, hist, out_rank, out_val, out_count, ebs, o,
                         stream);
      if (rc < 0) return rc;
      if (rc == 0) continue;
    }
    if (pair_ok(p, q, k, o)) {  // reads level k-1 itself: no lattice scatter
      const bool dx = (direct >> k) & 1;
      rc = launch_pair(p, q, d->sub[k - 1], g0, encode, dx ? x : X(k), dx ? xstep(k) : 1,
                       d->R[k - 1], R(k), codes, hist, out_rank, out_val, out_count, ebs, o,
                       stream);
      if (rc < 0) return rc;
      if (rc == 0) continue;
    }
    const int64_t n0 = q->shape[0];
    // march levels read the coarse points from the level below and, with an
    // even number of planes, store them into this level themselves
    bool march_direct = k > 1 && !level3_ok(p, q, o) && !getenv("LATEC_SLAB_MB") &&
                        march_ok(p, q, o);
    for (int j = 0; march_direct && j < 3; j++)
      march_direct = d->sub[k - 1]->shape[j] == (q->shape[j] + 1) / 2;
    const bool write_coarse = march_direct && n0 % 2 == 0;
    if (k > 1 && !write_coarse &&
        (rc = lattice(q, d->sub[k - 1], R(k), d->R[k - 1], false, stream)))
      return rc;
    const int64_t plane = (int64_t)p->channels * (q->npoints / std::max<int64_t>(n0, 1)) * 8;
    int64_t T = n0;
    static const int64_t slab_bytes = getenv("LATEC_SLAB_MB") ? atoll(getenv("LATEC_SLAB_MB")) << 20 : 0;
    if (slab_bytes && nd > 1 && (int64_t)p->channels * q->npoints * 8 > 4 * slab_bytes)
      T = std::max<int64_t>(8, (slab_bytes / std::max<int64_t>(plane, 1)) & ~1ll);
    const int64_t nslab = (n0 + T - 1) / T;
    if (nslab == 1 && level3_ok(p, q, o)) {
      rc = launch_level3(p, q, g0, encode, X(k), R(k), codes, hist, out_rank, out_val, out_count,
                         ebs, o, stream);
      if (rc < 0) return rc;
      if (rc == 0) continue;
    }
    if (nslab == 1 && march_ok(p, q, o)) {
      rc = launch_march(p, q, g0, encode, X(k), R(k), codes, hist, out_rank, out_val, out_count,
                        ebs, o, stream, march_direct ? d->sub[k - 1] : nullptr,
                        march_direct ? d->R[k - 1] : nullptr, write_coarse);
      if (rc < 0) return rc;
      if (rc == 0) continue;
    }
    // the march refused the layout: the stage kernels need the scattered coarse points
    if (write_coarse && (rc = lattice(q, d->sub[k - 1], R(k), d->R[k - 1], false, stream)))
      return rc;
    for (int64_t t = 0; t < nslab + nd - 1; t++) {
      for (int w = 1; w <= nd; w++) {
        const int64_t sl = t - (w - 1);
        if (sl < 0 || sl >= nslab) continue;
        const int64_t lo = sl * T, hi = std::min(n0, lo + T);
        if (w == 1) {
          for (int a = 0; a < nd; a++)
            if ((rc = launch_stage(p, q, 1 + a, g0 + a, encode, X(k), R(k), codes, hist, out_rank,
                                   out_val, out_count, ebs, o, lo, hi, stream)))
              return rc;
        } else if ((rc = launch_stage(p, q, 1 + nd + (w - 2), g0 + nd + (w - 2), encode, X(k),
                                      R(k), codes, hist, out_rank, out_val, out_count, ebs, o,
                                      lo, hi, stream))) {
          return rc;
        }
      }
    }
  }
  return LATEC_OK;
}

static int64_t tune_epoch();

int interp_encode_graph(latec_plan *p, const float *x, float *recon, void *codes,
                        uint32_t *hist, uint32_t *out_rank, float *out_val,
                        uint32_t *out_count, const double *ebs,
                        const latec_interp_opts *o, hipStream_t stream) {
  int rc = plan_upload(p);
  if (rc) return rc;
  if (pyramid_ok(p) && (rc = pyramid_setup(p))) return rc;  // no allocation inside capture
  // The graph bakes in pointers, ebs and options: key on all of them.
  DeviceBufs::GraphKey key;
  memset(&key, 0, sizeof key);
  key.a[0] = x; key.a[1] = recon; key.a[2] = codes; key.a[3] = hist;
  key.a[4] = out_rank; key.a[5] = out_val; key.a[6] = out_count;
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void *b, size_t n) {
    for (size_t i = 0; i < n; i++) h = (h ^ ((const uint8_t *)b)[i]) * 1099511628211ull;
  };
  mix(ebs, sizeof(double) * p->stages.size());
  mix(o, sizeof *o);
  const int64_t knobs = tune_epoch();  // launch shapes are baked into the graph too
  mix(&knobs, sizeof knobs);
  key.a[7] = (const void *)(uintptr_t)h;
  auto it = p->dev->graphs.find(key);
  if (it == p->dev->graphs.end()) {
    hipStream_t cap;
    if ((rc = hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate")))
      return rc;
    hipGraph_t graph;
    rc = hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "capture");
    if (!rc) {
      int erc = interp_encode(p, 0, (int)p->stages.size(), x, recon, codes, hist, out_rank,
                              out_val, out_count, ebs, o, cap);
      int crc = hip_check(hipStreamEndCapture(cap, &graph), "end capture");
      rc = erc ? erc : crc;
    }
    hipGraphExec_t exec = nullptr;
    if (!rc) rc = hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    if (!rc) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cap);
    if (rc) return rc;
    it = p->dev->graphs.emplace(key, exec).first;
  }
  return hip_check(hipGraphLaunch(it->second, stream), "hipGraphLaunch");
}

int interp_predict(latec_plan *p, int stage, const float *recon, float *pred,
                   const latec_interp_opts *o, hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if (stage < 1 || stage >= (int)p->stages.size()) {
    set_error("known mask does not match the interp schedule");
    return LATEC_E_ARG;
  }
  if ((rc = plan_upload(p))) return rc;
  const StageDesc &st = p->stages[stage];
  if (!st.count) return LATEC_OK;
  GridArgs g = make_args(p, stage, 1.0, o);
  dim3 grid(row_blocks(st), (unsigned)st.classes.size());
  const uint32_t *ex = p->dev->excluded ? p->dev->excluded + p->dev->excl_off[stage] : nullptr;
  hipLaunchKernelGGL((predict_stage_kernel<0, int64_t, false>), grid, dim3(256), 0, stream, g,
                     p->dev->classes + p->dev->class_off[stage], ex, recon, pred);
  return hip_check(hipGetLastError(), "predict_stage_kernel");
}

// Stage error moments of an encode (stats_stage_kernel): out[s * 5 + k] for
// every stage of the plan (device, float64); scratch holds the partials
// (interp_stats_words(p) float64 words, the first n_stages + 1 of them the
// partial offsets).
int64_t interp_stats_words(latec_plan *p) {
  int64_t nb = 0;
  for (const StageDesc &st : p->stages)
    if (st.count) nb += (int64_t)row_blocks(st) * (int64_t)st.classes.size();
  return (int64_t)p->stages.size() + 1 + NSTAT * nb;
}

int interp_stats(latec_plan *p, const float *x, const float *recon,
                 const latec_interp_opts *o, double *out, double *scratch,
                 int64_t scratch_words, hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if ((rc = plan_upload(p))) return rc;
  const int ns = (int)p->stages.size();
  if (scratch_words < interp_stats_words(p)) {
    set_error("latec_interp_stats: scratch too small");
    return LATEC_E_CAPACITY;
  }
  std::vector<int64_t> off(ns + 1, 0);
  for (int s = 0; s < ns; s++) {
    const StageDesc &st = p->stages[s];
    off[s + 1] = off[s] + (st.count ? (int64_t)row_blocks(st) * (int64_t)st.classes.size() : 0);
  }
  int64_t *d_off = (int64_t *)scratch;
  double *partial = scratch + ns + 1;
  if ((rc = hip_check(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * (ns + 1),
                                     hipMemcpyHostToDevice, stream), "hipMemcpyAsync")))
    return rc;
  const KSel sel = pick(p);
  for (int s = 0; s < ns; s++) {
    const StageDesc &st = p->stages[s];
    if (!st.count) continue;
    GridArgs g = make_args(p, s, 1.0, o);
    dim3 grid(row_blocks(st), (unsigned)st.classes.size());
    const ClassDesc *cls = p->dev->classes + p->dev->class_off[s];
    const uint32_t *ex = p->dev->excluded ? p->dev->excluded + p->dev->excl_off[s] : nullptr;
    double *ps = partial + NSTAT * off[s];
    if (sel == KSel::Fast && p->block == 1) {
      switch (p->ndim) {
        case 1: hipLaunchKernelGGL((stats_stage_kernel<1, uint32_t, true>), grid, dim3(256), 0, stream, g, cls, ex, x, recon, s == 0, ps); break;
        case 2: hipLaunchKernelGGL((stats_stage_kernel<2, uint32_t, true>), grid, dim3(256), 0, stream, g, cls, ex, x, recon, s == 0, ps); break;
        case 3: hipLaunchKernelGGL((stats_stage_kernel<3, uint32_t, true>), grid, dim3(256), 0, stream, g, cls, ex, x, recon, s == 0, ps); break;
        default: hipLaunchKernelGGL((stats_stage_kernel<4, uint32_t, true>), grid, dim3(256), 0, stream, g, cls, ex, x, recon, s == 0, ps); break;
      }
    } else {
      hipLaunchKernelGGL((stats_stage_kernel<0, int64_t, false>), grid, dim3(256), 0, stream, g,
                         cls, ex, x, recon, s == 0, ps);
    }
    if ((rc = hip_check(hipGetLastError(), "stats_stage_kernel"))) return rc;
  }
  hipLaunchKernelGGL(stats_finish_kernel, dim3(ns), dim3(NSTAT * STATS_SPLIT), 0, stream,
                     partial, d_off, out);
  if ((rc = hip_check(hipGetLastError(), "stats_finish_kernel"))) return rc;
  // the host offsets are read by the copy above: finish before they go
  return hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// (min, max) of nseg segments [off, off + len) of device float32 x into
// out[2 g], out[2 g + 1] (device); scratch: minmax_words(nseg) float32 words.
int64_t minmax_words(int nseg) { return 2 * (int64_t)nseg * MM_BLOCKS + 4 * (int64_t)nseg + 4; }

int minmax(int nseg, const float *x, const int64_t *off, const int64_t *len, float *out,
           float *scratch, int64_t scratch_words, hipStream_t stream) {
  if (nseg <= 0) return LATEC_OK;
  if (scratch_words < minmax_words(nseg)) {
    set_error("latec_minmax: scratch too small");
    return LATEC_E_CAPACITY;
  }
  std::vector<int64_t> seg(2 * (size_t)nseg);
  for (int g = 0; g < nseg; g++) {
    if (len[g] <= 0) {
      set_error("latec_minmax: empty segment");
      return LATEC_E_ARG;
    }
    seg[2 * g] = off[g];
    seg[2 * g + 1] = len[g];
  }
  float *part = scratch;
  int64_t *d_seg = (int64_t *)(scratch + 2 * (int64_t)nseg * MM_BLOCKS);
  int rc;
  if ((rc = hip_check(hipMemcpyAsync(d_seg, seg.data(), sizeof(int64_t) * seg.size(),
                                     hipMemcpyHostToDevice, stream), "hipMemcpyAsync")))
    return rc;
  hipLaunchKernelGGL(minmax_kernel, dim3(MM_BLOCKS, (unsigned)nseg), dim3(256), 0, stream, x,
                     d_seg, part);
  hipLaunchKernelGGL(minmax_finish_kernel, dim3((nseg + 255) / 256), dim3(256), 0, stream,
                     part, nseg, out);
  if ((rc = hip_check(hipGetLastError(), "minmax_kernel"))) return rc;
  return hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");  // seg ends here
}

int code_hist(const void *codes, int64_t n, int cb, uint32_t R, uint32_t *hist,
              hipStream_t stream) {
  if (!n) return LATEC_OK;
  if (cb == 2)
    hipLaunchKernelGGL(code_hist_kernel<uint16_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       (const uint16_t *)codes, n, R, hist);
  else
    hipLaunchKernelGGL(code_hist_kernel<uint32_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       (const uint32_t *)codes, n, R, hist);
  return hip_check(hipGetLastError(), "code_hist_kernel");
}

int quantize(const float *x, const float *pred, int64_t n, double eb,
             const latec_interp_opts *o, void *codes, uint32_t *out_rank,
             float *out_val, uint32_t *out_count, uint32_t *hist,
             hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if (!(eb > 0)) {
    set_error("error bound must be > 0, got " + std::to_string(eb));
    return LATEC_E_ARG;
  }
  if (n > 0xFFFFFFFFll) {
    set_error("too many values for 32-bit outlier ranks");
    return LATEC_E_ARG;
  }
  if (!n) return LATEC_OK;
  GridArgs g;
  memset(&g, 0, sizeof g);
  g.R = o->radius; g.eb = eb; g.round_mode = o->round_mode; g.span = o->span;
  g.offset = o->offset; g.w = 2.0 * eb; g.inv_w = 1.0 / g.w; g.feb = (float)eb;
  if (o->code_bytes == 2)
    hipLaunchKernelGGL(quantize_kernel<uint16_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       g, x, pred, n, (uint16_t *)codes, out_rank, out_val, out_count, hist);
  else
    hipLaunchKernelGGL(quantize_kernel<uint32_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       g, x, pred, n, (uint32_t *)codes, out_rank, out_val, out_count, hist);
  return hip_check(hipGetLastError(), "quantize_kernel");
}

int dequantize(const float *pred, const void *codes, int64_t n,
               const uint32_t *out_rank, const float *out_val,
               const uint32_t *out_count, double eb, const latec_interp_opts *o,
               float *recon, hipStream_t stream) {
  int rc = check_opts(o);
  if (rc) return rc;
  if (!n) return LATEC_OK;
  GridArgs g;
  memset(&g, 0, sizeof g);
  g.R = o->radius; g.eb = eb; g.w = 2.0 * eb;
  if (o->code_bytes == 2)
    hipLaunchKernelGGL(dequantize_kernel<uint16_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       g, pred, (const uint16_t *)codes, n, out_rank, out_val, out_count, recon);
  else
    hipLaunchKernelGGL(dequantize_kernel<uint32_t>, dim3(grid_blocks(n)), dim3(256), 0, stream,
                       g, pred, (const uint32_t *)codes, n, out_rank, out_val, out_count, recon);
  return hip_check(hipGetLastError(), "dequantize_kernel");
}

}  // namespace latec

// ---------------------------------------------------------------------------
// Batched uneven workload: many independent fields (any rank <= 4, any
// shape, anchor_block 1) coded together.  Step s of the batch runs stage s
// of every field that has one, as ONE persistent launch over a work list of
// (field-stage, parity class, row range) items of about BATCH_ITEM points;
// workgroups take items round-robin.  The whole step sequence is recorded
// once into a hipGraph and replayed.  The per-point arithmetic is the stage
// kernels' (predict_point / quantize_one / emit), so each field's codes,
// outliers and reconstruction equal its own single-field encode.
// ---------------------------------------------------------------------------
#define BATCH_ITEM 8192

namespace latec {

struct BDesc {  // one (field, stage)
  GridArgs g;   // code_base = the stage's first code in the batch arrays
  const ClassDesc *classes;
  const float *x;
  float *recon;
  int64_t hist_off;  // words: the stage's histogram in the batch array
  int64_t cnt_idx;   // its outlier counter
  int32_t anchors, pad;
};
struct BItem {
  int32_t d, cls;
  uint32_t row0, nrows;
};

// Rows [row0, row0 + nrows) of class c by one workgroup (anchor_block 1).
template <int ND, typename F>
__device__ __forceinline__ void walk_item(const ClassDesc &c, const GridArgs &g, uint32_t row0,
                                          uint32_t nrows, F &&fn) {
  constexpr int NMAX = ND > 0 ? ND : LATEC_MAX_DIM;
  const int nd = ND > 0 ? ND : g.ndim;
  const int tr = c.tr, rb = 256 / tr;
  const int lrow = threadIdx.x / tr, lcol = threadIdx.x % tr;
  const uint32_t L = c.row_len;
  const int64_t start_last = c.starts[nd - 1];
  for (uint32_t base = 0; base < nrows; base += rb) {
    const uint32_t row = row0 + base + lrow;
    const bool rv = base + lrow < nrows;
    int64_t coord[NMAX];
    uint32_t fbase = 0;
    int64_t rbase = c.offset;
    uint32_t idx = rv ? row : row0;
#pragma unroll
    for (int j = NMAX - 2; j >= 0; j--) {
      if (ND == 0 && j > nd - 2) continue;
      const uint32_t q = fdiv(idx, c.magic[j], c.shift[j]);
      const uint32_t m = idx - q * c.sizes[j];
      idx = q;
      const int64_t xj = c.starts[j] + (int64_t)c.step * m;
      coord[j] = xj;
      fbase += (uint32_t)xj * (uint32_t)g.gstr[j];
      rbase += (int64_t)m * c.rstr[j];
    }
    for (uint32_t col0 = 0; col0 < L; col0 += tr) {
      const uint32_t col = col0 + lcol;
      const bool active = rv && col < L;
      const int64_t xl = start_last + (int64_t)c.step * col;
      coord[nd - 1] = xl;
      fn(active, fbase + (uint32_t)xl, rbase + col, (const int64_t *)coord);
    }
  }
}

template <int ND, bool ENC>
__device__ __forceinline__ void batch_item(const BDesc &D, const ClassDesc &c, const BItem &it,
                                           uint16_t *__restrict__ codes,
                                           uint32_t *__restrict__ hist,
                                           uint32_t *__restrict__ out_rank,
                                           float *__restrict__ out_val,
                                           uint32_t *__restrict__ out_count, uint32_t *lhist) {
  const GridArgs &g = D.g;
  const int parity = c.parity;
  const uint32_t win_lo = g.R > HIST_WIN / 2 ? g.R - HIST_WIN / 2 : 0;
  uint32_t *ghist = hist + D.hist_off;
  const int64_t nout = ENC ? 0 : (int64_t)out_count[D.cnt_idx];
  walk_item<ND>(c, g, it.row0, it.nrows, [&](bool active, uint32_t flat, int64_t rank,
                                             const int64_t *coord) {
    uint32_t code = 0;
    float xv = 0.f;
    if (active) {
      const float p = D.anchors ? 0.0f
                                : predict_point<ND, uint32_t>(D.recon, g, parity, flat, coord);
      float r;
      if (ENC) {
        xv = D.x[flat];
        code = quantize_one(xv, p, g, &r);
      } else {
        code = codes[g.code_base + rank];
        if (code) {
          r = (float)((double)p + g.w * (double)((int)code - (int)g.R));
        } else {
          const uint32_t *orr = out_rank + g.code_base;
          int64_t lo = 0, hi = nout;
          while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)orr[mid] < rank) lo = mid + 1; else hi = mid;
          }
          r = (lo < nout && (int64_t)orr[lo] == rank) ? out_val[g.code_base + lo]
                                                      : __int_as_float(0x7fc00000);
        }
      }
      D.recon[flat] = r;
    }
    if (ENC)
      emit<uint16_t>(codes, g.code_base + rank, code, xv, (uint32_t)rank,
                     out_rank + g.code_base, out_val + g.code_base, out_count + D.cnt_idx, lhist,
                     ghist, win_lo, active);
  });
}

template <bool ENC>
__global__ void __launch_bounds__(256) batch_kernel(const BDesc *__restrict__ descs,
                                                    const BItem *__restrict__ items,
                                                    int32_t n_items, uint16_t *__restrict__ codes,
                                                    uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ out_rank,
                                                    float *__restrict__ out_val,
                                                    uint32_t *__restrict__ out_count) {
  __shared__ uint32_t lhist[ENC ? HIST_WIN : 1];
  for (int32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
    const BItem it = items[i];
    const BDesc &D = descs[it.d];
    const ClassDesc &c = D.classes[it.cls];
    if (ENC) {
      for (int k = threadIdx.x; k < HIST_WIN; k += blockDim.x) lhist[k] = 0;
      __syncthreads();
    }
    switch (D.g.ndim) {
      case 1: batch_item<1, ENC>(D, c, it, codes, hist, out_rank, out_val, out_count, lhist); break;
      case 2: batch_item<2, ENC>(D, c, it, codes, hist, out_rank, out_val, out_count, lhist); break;
      case 3: batch_item<3, ENC>(D, c, it, codes, hist, out_rank, out_val, out_count, lhist); break;
      default: batch_item<4, ENC>(D, c, it, codes, hist, out_rank, out_val, out_count, lhist); break;
    }
    if (ENC) {
      __syncthreads();
      const uint32_t R = D.g.R;
      uint32_t *gwin = hist + D.hist_off + 2 * (size_t)R + (size_t)(i % HIST_COPIES) * HIST_WIN;
      for (int k = threadIdx.x; k < HIST_WIN; k += blockDim.x)
        if (lhist[k]) atomicAdd(&gwin[k], lhist[k]);
      __syncthreads();
    }
  }
}

}  // namespace latec

struct latec_batch {
  bool encode = true;
  int n_steps = 0;
  std::vector<int32_t> step_items;  // items per step
  std::vector<int64_t> step_first;  // first item of each step
  latec::BDesc *d_desc = nullptr;
  latec::BItem *d_items = nullptr;
  uint16_t *codes = nullptr;
  uint32_t *hist = nullptr, *out_rank = nullptr, *out_count = nullptr;
  float *out_val = nullptr;
  hipGraphExec_t graph = nullptr;
  hipStream_t graph_stream = nullptr;
};

using namespace latec;

static int batch_launch_steps(latec_batch *b, hipStream_t s) {
  for (int st = 0; st < b->n_steps; st++) {
    const int32_t n = b->step_items[st];
    if (!n) continue;
    const unsigned grid = (unsigned)std::min<int64_t>(n, 4096);
    const BItem *items = b->d_items + b->step_first[st];
    if (b->encode)
      hipLaunchKernelGGL(batch_kernel<true>, dim3(grid), dim3(256), 0, s, b->d_desc, items, n,
                         b->codes, b->hist, b->out_rank, b->out_val, b->out_count);
    else
      hipLaunchKernelGGL(batch_kernel<false>, dim3(grid), dim3(256), 0, s, b->d_desc, items, n,
                         b->codes, b->hist, b->out_rank, b->out_val, b->out_count);
    const int rc = hip_check(hipGetLastError(), "batch_kernel");
    if (rc) return rc;
  }
  return LATEC_OK;
}

extern "C" {

int latec_batch_create(int nf, latec_plan *const *plans, const float *const *d_x,
                       float *const *d_recon, const int64_t *code_base, const int64_t *hist_base,
                       const int64_t *cnt_base, const double *const *ebs,
                       const latec_interp_opts *o, int encode, void *d_codes, uint32_t *d_hist,
                       uint32_t *d_out_rank, float *d_out_val, uint32_t *d_out_count,
                       latec_batch **out) {
  if (!out || nf < 0 || (nf && !o)) {
    set_error("latec_batch_create: bad arguments");
    return LATEC_E_ARG;
  }
  int rc;
  for (int f = 0; f < nf; f++) {
    if ((rc = check_opts(&o[f]))) return rc;
    if (o[f].code_bytes != 2 || o[f].radius != o[0].radius) {
      set_error("a batch codes one radius <= 32768 (16-bit codes)");
      return LATEC_E_ARG;
    }
  }
  const int64_t twoR = nf ? hist_words(o[0].radius) : 0;
  std::vector<BDesc> descs;
  std::vector<std::vector<BItem>> steps;
  for (int f = 0; f < nf; f++) {
    latec_plan *p = plans[f];
    if (!p || p->block != 1 || p->channels != 1 || p->npoints >= 0x7fffffffll) {
      set_error("batch fields need anchor_block 1, one channel and < 2^31 points");
      return LATEC_E_ARG;
    }
    if ((rc = plan_upload(p))) return rc;
    for (int s = 0; s < (int)p->stages.size(); s++) {
      const StageDesc &st = p->stages[s];
      if (!st.count) continue;
      if (encode && !(ebs[f][s] > 0)) {
        set_error("error bound must be > 0");
        return LATEC_E_ARG;
      }
      BDesc D;
      memset(&D, 0, sizeof D);
      D.g = make_args(p, s, ebs[f][s], &o[f]);
      D.g.code_base = code_base[f] + st.offset;
      D.classes = p->dev->classes + p->dev->class_off[s];
      D.x = encode ? d_x[f] : nullptr;
      D.recon = d_recon[f];
      D.hist_off = hist_base[f] + (int64_t)s * twoR;
      D.cnt_idx = cnt_base[f] + s;
      D.anchors = s == 0;
      const int32_t di = (int32_t)descs.size();
      descs.push_back(D);
      if ((int)steps.size() <= s) steps.resize(s + 1);
      for (int ci = 0; ci < (int)st.classes.size(); ci++) {
        const ClassDesc &c = st.classes[ci];
        if (!c.size) continue;
        const uint32_t per = (uint32_t)std::max<int64_t>(1, BATCH_ITEM / std::max<uint32_t>(c.row_len, 1));
        for (uint32_t r0 = 0; r0 < c.nrows; r0 += per)
          steps[s].push_back({di, ci, r0, std::min(per, c.nrows - r0)});
      }
    }
  }
  auto *b = new latec_batch();
  b->encode = encode != 0;
  b->n_steps = (int)steps.size();
  std::vector<BItem> all;
  for (auto &v : steps) {
    b->step_first.push_back((int64_t)all.size());
    b->step_items.push_back((int32_t)v.size());
    all.insert(all.end(), v.begin(), v.end());
  }
  b->codes = (uint16_t *)d_codes;
  b->hist = d_hist;
  b->out_rank = d_out_rank;
  b->out_val = d_out_val;
  b->out_count = d_out_count;
  if ((rc = hip_check(hipMalloc(&b->d_desc, sizeof(BDesc) * std::max<size_t>(descs.size(), 1)),
                      "hipMalloc")) ||
      (rc = hip_check(hipMalloc(&b->d_items, sizeof(BItem) * std::max<size_t>(all.size(), 1)),
                      "hipMalloc")) ||
      (rc = hip_check(hipMemcpy(b->d_desc, descs.data(), sizeof(BDesc) * descs.size(),
                                hipMemcpyHostToDevice), "hipMemcpy")) ||
      (rc = hip_check(hipMemcpy(b->d_items, all.data(), sizeof(BItem) * all.size(),
                                hipMemcpyHostToDevice), "hipMemcpy"))) {
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_items) (void)hipFree(b->d_items);
    delete b;
    return rc;
  }
  *out = b;
  return LATEC_OK;
}

int latec_batch_n_steps(const latec_batch *b) { return b ? b->n_steps : LATEC_E_ARG; }

int latec_batch_run(latec_batch *b, int use_graph, void *stream) {
  if (!b) {
    set_error("latec_batch_run: null batch");
    return LATEC_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) return batch_launch_steps(b, s);
  int rc;
  if (!b->graph) {
    // capture on a private stream (capturing the legacy default stream is not allowed)
    hipStream_t cs;
    if ((rc = hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking),
                        "hipStreamCreateWithFlags")))
      return rc;
    hipGraph_t gr;
    if ((rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                        "hipStreamBeginCapture"))) {
      (void)hipStreamDestroy(cs);
      return rc;
    }
    rc = batch_launch_steps(b, cs);
    const hipError_t e = hipStreamEndCapture(cs, &gr);
    if (!rc) rc = hip_check(e, "hipStreamEndCapture");
    if (!rc) {
      rc = hip_check(hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0),
                     "hipGraphInstantiate");
      (void)hipGraphDestroy(gr);
    }
    (void)hipStreamDestroy(cs);
    if (rc) return rc;
  }
  return hip_check(hipGraphLaunch(b->graph, s), "hipGraphLaunch");
}

void latec_batch_destroy(latec_batch *b) {
  if (!b) return;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_items) (void)hipFree(b->d_items);
  delete b;
}

}  // extern "C"
