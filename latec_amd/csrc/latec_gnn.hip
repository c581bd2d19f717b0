// latec_gnn.hip -- the GNN predictor on gfx950: geometry, message pass, pools,
// head, finalize, and the interpolation gate's rate proxy.
//
// Model (latec/gnn_predictor.py:593-933), per stage of the codec schedule:
//   geometry ...... nearest known sample per half-direction and side, within
//                   `stride` steps (gnn_predictor.py:117-443), evaluated per
//                   query point from the closed-form stage index of a cell;
//   messages ...... per live (line, point) pair and axis k: BiDirEmbed of the
//                   two neighbours' axis-k embeddings (+ direction state,
//                   + log distances), DirEmbed for one-sided pairs
//                   (gnn_predictor.py:686-718, 818-878);
//   line pool ..... single-query attention over the lines (+ null token)
//                   (AttnPool, gnn_predictor.py:720-747);
//   head .......... attention over the axes, then PredHead (mu, log_b)
//                   (gnn_predictor.py:749-789, 922-931);
//   finalize ...... E[q] = MixEmbed(ctx, InitEmbed(value))
//                   (gnn_predictor.py:791-803, 908-920).
//
// Numerics: the three wide MLPs (bidir/dir message, mix) run as MFMA GEMMs --
// BF16 operands with FP32 accumulation (v_mfma_f32_16x16x32_bf16) in the bf16
// mode, FP32 operands (v_mfma_f32_16x16x4f32) in the fp32 mode -- in a fixed
// order with no atomics; pools and head run in FP32.  Every reduction has a
// fixed order, so an encoder and a decoder that feed the same inputs get the
// same bits.  Two algebraic rewrites of the reference's form:
//   * the direction-state columns of the first message layer act on a table
//     row picked by (line, axis, side), so W[:, d:de] @ table[idx] is folded
//     into a per-state bias vector;
//   * the line pool's value projection is linear, so
//     sum_l w_l (Wv m_l + bv) = Wv (sum_l w_l m_l) + (1 - w_null) bv,
//     applied once per (point, axis) instead of once per line.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "latec_plan.h"

namespace latec {
namespace gnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { NUM_FP32 = 0, NUM_BF16 = 1 };
enum { ROW_BI = 0, ROW_DIR = 1, ROW_FIN = 2 };

static int hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_hip_error(what, (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

// float -> bfloat16, round to nearest even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

template <typename T>
__device__ __forceinline__ float ld(const T *p);
template <>
__device__ __forceinline__ float ld<float>(const float *p) { return *p; }
template <>
__device__ __forceinline__ float ld<uint16_t>(const uint16_t *p) { return from_bf16(*p); }
template <typename T>
__device__ __forceinline__ T st(float v);
template <>
__device__ __forceinline__ float st<float>(float v) { return v; }
template <>
__device__ __forceinline__ uint16_t st<uint16_t>(float v) { return to_bf16(v); }

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

// ---------------------------------------------------------------------------
// model parameters on the device
// ---------------------------------------------------------------------------

// Offsets (in floats) into the packed f32 parameter block, in the reference
// state_dict order (gnn_predictor.py:806-816 module order).
struct Layout {
  int d, h, ds, de, ns;
  int64_t init0w, init0b, init2w, init2b, table;
  int64_t dir0w, dir0b, dir2w, dir2b;
  int64_t bi0w, bi0b, bi2w, bi2b;
  int64_t lp[8], ap[8];  // wk0w wk0b wk2w wk2b wvw wvb null_k null_v
  int64_t hd0w, hd0b, hd2w, hd2b;
  int64_t mx0w, mx0b, mx2w, mx2b;
  int64_t total;
};

static Layout make_layout(int d, int agg) {
  Layout L{};
  L.d = d;
  L.h = 2 * d;
  L.ds = d / 4 > 1 ? d / 4 : 1;
  L.de = d + L.ds;
  L.ns = 2 * agg + 1;
  int64_t o = 0;
  auto take = [&](int64_t n) {
    int64_t r = o;
    o += n;
    return r;
  };
  L.init0w = take(d);
  L.init0b = take(d);
  L.init2w = take((int64_t)d * d);
  L.init2b = take(d);
  L.table = take((int64_t)L.ns * L.ds);
  L.dir0w = take((int64_t)L.h * (L.de + 2));
  L.dir0b = take(L.h);
  L.dir2w = take((int64_t)d * L.h);
  L.dir2b = take(d);
  L.bi0w = take((int64_t)L.h * (2 * L.de + 2));
  L.bi0b = take(L.h);
  L.bi2w = take((int64_t)d * L.h);
  L.bi2b = take(d);
  for (int pool = 0; pool < 2; pool++) {
    int64_t *P = pool ? L.ap : L.lp;
    P[0] = take((int64_t)d * d);
    P[1] = take(d);
    P[2] = take(d);
    P[3] = take(1);
    P[4] = take((int64_t)d * d);
    P[5] = take(d);
    P[6] = take(1);
    P[7] = take(d);
  }
  L.hd0w = take((int64_t)L.h * (d + 1));
  L.hd0b = take(L.h);
  L.hd2w = take(2 * (int64_t)L.h);
  L.hd2b = take(2);
  L.mx0w = take((int64_t)L.h * 2 * d);
  L.mx0b = take(L.h);
  L.mx2w = take((int64_t)d * L.h);
  L.mx2b = take(d);
  L.total = o;
  return L;
}

}  // namespace gnn
}  // namespace latec

struct latec_gnn {
  int d = 0, agg = 0, ndim = 0, numerics = 0;
  latec::gnn::Layout lay{};
  int L = 0;                       // half-directions
  std::vector<int8_t> dirs;        // (L, ndim)
  std::vector<float> lognnz;       // (L)
  std::vector<int8_t> sidx;        // (L, ndim, 2): state index (side 0 = -, 1 = +)
  void *dev = nullptr;             // one device block holding every array below
  size_t dev_bytes = 0;
  // device pointers into `dev`
  void *bi_w1 = nullptr, *bi_w2 = nullptr, *dir_w1 = nullptr, *dir_w2 = nullptr;
  void *mx_w1 = nullptr, *mx_w2 = nullptr, *key_w = nullptr, *init_w2 = nullptr;
  // bf16 mode: head weights as high and low bf16 parts, natural order
  void *hd_wkv_hi = nullptr, *hd_wkv_lo = nullptr, *hd_w1_hi = nullptr, *hd_w1_lo = nullptr;
  void *lk_wv_hi = nullptr, *lk_wv_lo = nullptr;  // line pool value projection
  float *f32 = nullptr;            // f32 arrays (biases, state rows, pools, head, init)
  std::vector<int64_t> f32_off;    // named offsets into f32 (see make_device)
  int8_t *d_dirs = nullptr, *d_sidx = nullptr;
  float *d_lognnz = nullptr;
  float scal[4] = {0, 0, 0, 0};    // line pool bk2, null_k; axis pool bk2, null_k
  int n_cu = 256;
  int64_t *d_line_off = nullptr;   // frame offsets of the lines (cached per frame)
  // geometry lookup tables (geo_table_kernel), per (stage, log2 stride)
  std::vector<std::pair<int64_t, uint8_t *>> geo_tabs;
  // gate tables (thresholds, per-level constants) per eb
  std::vector<std::pair<uint64_t, void *>> gate_tabs;
  // a stage's parity classes and their start indices on the device, per
  // (plan shape, levels, stride, stage): the plan fixes them
  struct ClassTab {
    std::vector<int64_t> key;
    void *dev;
  };
  std::vector<ClassTab> class_tabs;
  // latec_gnn_code_stage's / latec_gnn_decode_stage_*'s device temporaries
  void *stage_work = nullptr;
  size_t stage_work_bytes = 0;
  // the decoder's pinned staging of a stage's codes and outliers
  void *stage_host = nullptr;
  size_t stage_host_bytes = 0;
  int64_t dec_n = -1;  // coded points of the stage latec_gnn_decode_stage_begin set up
  int64_t line_off_key[LATEC_MAX_DIM] = {0};
  // workspace
  void *work = nullptr;
  size_t work_bytes = 0;
};

namespace latec {
namespace gnn {

// f32 array slots inside latec_gnn::f32
enum F32Slot {
  F_BI_B2, F_DIR_B2, F_MX_B2,
  F_LK_B1, F_LK_W2, F_LK_WV, F_LK_BV, F_LK_NULLV,
  F_AP_WKV, F_AP_B1, F_AP_W2, F_AP_BV, F_AP_NULLV,
  F_HD_W1, F_HD_COND, F_HD_B1, F_HD_W2, F_HD_B2,
  F_IN_W0, F_IN_B0, F_IN_B2,
  F_SCAL,  // lk_bk2, lk_null_k, ap_bk2, ap_null_k
  F_NSLOTS
};

// ---------------------------------------------------------------------------
// stage index of a cell and the geometry kernel
// ---------------------------------------------------------------------------

struct GeomArgs {
  int32_t ndim, L, stage, levels_log2;  // levels_log2 = log2(stride)
  int32_t stride, limit, pad_lo, has_halo;
  int64_t cshape[LATEC_MAX_DIM];   // chunk shape (local schedule)
  int64_t fstr[LATEC_MAX_DIM];     // frame strides (frame = chunk + pad on every side)
  int64_t origin[LATEC_MAX_DIM];   // chunk origin in the tensor
  int64_t tshape[LATEC_MAX_DIM];   // tensor shape
  int64_t tstr[LATEC_MAX_DIM];     // tensor strides
  int64_t edges[LATEC_MAX_DIM];    // chunk edges (chunk-grid id of a halo cell)
  int64_t grid[LATEC_MAX_DIM];
  const uint8_t *coded;            // per chunk: already coded
  int32_t n_classes, pad2;
};

// Stage of a cell in the block-1 schedule with anchor stride 2^S
// (latec/levels.py:100-147): anchors are stage 0; a point whose coordinates
// are multiples of s = 2^t with odd multiples on the axis set A belongs to
// level (S - 1 - t), sub-stage axis(A) if |A| == 1, else ndim + |A| - 2.
// Coordinates may be negative (halo); only their residues mod 2^S matter.
__device__ __forceinline__ int stage_of(const int64_t *c, int ndim, int S) {
  const int64_t m = ((int64_t)1 << S) - 1;
  int64_t orv = 0;
  for (int j = 0; j < ndim; j++) orv |= (c[j] & m);
  if (orv == 0) return 0;
  const int t = __builtin_ctzll((unsigned long long)orv);
  int w = 0, ax = 0;
  for (int j = 0; j < ndim; j++)
    if ((c[j] >> t) & 1) {
      w++;
      ax = j;
    }
  const int li = S - 1 - t;
  const int sub = w == 1 ? ax : ndim + w - 2;
  return 1 + li * (2 * ndim - 1) + sub;
}

// Usable known sample at chunk-local coordinates c before stage g.stage.
// A cell of the frame is known before the stage when its stage in the
// schedule's periodic pattern is earlier (stage_of works on residues mod the
// stride, so halo cells take the pattern too).  A known cell is usable when
// it lies in the chunk, or, in a halo frame, inside the tensor and the frame
// and in a coded chunk or on the anchor lattice (gnn_predictor.py:1241-1275,
// 1338-1356).  Returns 0 (not known), 1 (known, unusable), 2 (usable).
__device__ __forceinline__ int cell_state(const GeomArgs &g, const int64_t *c) {
  const int st = stage_of(c, g.ndim, g.levels_log2);
  if (st >= g.stage) return 0;
  bool inside = true;
  for (int j = 0; j < g.ndim; j++) inside &= (c[j] >= 0 && c[j] < g.cshape[j]);
  if (inside) return 2;
  if (!g.has_halo) return 1;
  int64_t cid = 0;
  for (int j = 0; j < g.ndim; j++) {
    const int64_t gc = g.origin[j] + c[j];
    if (gc < 0 || gc >= g.tshape[j]) return 1;
    if (c[j] < -g.pad_lo || c[j] >= g.cshape[j] + g.pad_lo) return 1;
    cid = cid * g.grid[j] + gc / g.edges[j];
  }
  return (st == 0 || g.coded[cid]) ? 2 : 1;
}

// Whether a cell is known before a stage depends only on its coordinates'
// residues mod the stride (stage_of), so the first known cell along a line
// from a point is a function of the point's residues, the line and the side:
// tab[(r * L + l) * 2 + side] = the first t in 1..limit with a known cell at
// r + sg t dir_l (0 if none), r the residues packed S bits per axis (axis 0
// highest).  One thread per (residues, line, side).
__global__ void geo_table_kernel(GeomArgs g, const int8_t *__restrict__ dirs, int64_t n,
                                 uint8_t *__restrict__ tab) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int side = (int)(i & 1), l = (int)((i >> 1) % g.L);
  const int64_t r = (i >> 1) / g.L;
  const int S = g.levels_log2;
  const int64_t m = ((int64_t)1 << S) - 1;
  int64_t c[LATEC_MAX_DIM], nb[LATEC_MAX_DIM];
  for (int j = 0; j < g.ndim; j++) c[j] = (r >> (S * (g.ndim - 1 - j))) & m;
  const int sg = side ? -1 : 1;
  int hit = 0;
  for (int t = 1; t <= g.limit; t++) {
    for (int j = 0; j < g.ndim; j++) nb[j] = c[j] + (int64_t)sg * t * dirs[l * g.ndim + j];
    if (stage_of(nb, g.ndim, S) < g.stage) {
      hit = t;
      break;
    }
  }
  tab[i] = (uint8_t)hit;
}

// One thread per query point of the stage (enumerated class by class).
// Writes qidx[rank] (frame flat index), gidx[rank] (tensor flat index) and
// steps[rank * L + l] = sp | sn << 16 (0 = no usable sample on that side).
// With `tab` (geo_table_kernel) each line and side costs one lookup and the
// usability of the cell found; without it, a scan of up to `limit` cells.
__global__ void geometry_kernel(const ClassDesc *__restrict__ cls, GeomArgs g,
                                const int8_t *__restrict__ dirs, int64_t total,
                                const int64_t *__restrict__ cls_start,
                                int64_t *__restrict__ qidx, int64_t *__restrict__ gidx,
                                int32_t *__restrict__ steps, const uint8_t *__restrict__ tab) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int ci = 0;
  while (ci + 1 < g.n_classes && cls_start[ci + 1] <= i) ci++;
  const ClassDesc &c = cls[ci];
  int64_t m = i - cls_start[ci];
  int64_t coord[LATEC_MAX_DIM], rank = c.offset;
  for (int j = c.ndim - 1; j >= 0; j--) {
    const int64_t sz = c.sizes[j];
    const int64_t mj = m % sz;
    m /= sz;
    coord[j] = c.starts[j] + (int64_t)c.step * mj;
    rank += mj * c.rstr[j];
  }
  int64_t qf = 0, gf = 0;
  for (int j = 0; j < g.ndim; j++) {
    qf += (coord[j] + g.pad_lo) * g.fstr[j];
    gf += (g.origin[j] + coord[j]) * g.tstr[j];
  }
  qidx[rank] = qf;
  gidx[rank] = gf;
  int64_t nb[LATEC_MAX_DIM];
  if (tab) {
    const int S = g.levels_log2;
    const int64_t mk = ((int64_t)1 << S) - 1;
    int64_t r = 0;
    for (int j = 0; j < g.ndim; j++) r = (r << S) | (coord[j] & mk);
    const uint8_t *tr = tab + r * g.L * 2;
    for (int l = 0; l < g.L; l++) {
      int res[2];
      for (int side = 0; side < 2; side++) {
        const int t = tr[2 * l + side];
        res[side] = 0;
        if (t) {
          const int sg = side ? -1 : 1;
          for (int j = 0; j < g.ndim; j++) nb[j] = coord[j] + (int64_t)sg * t * dirs[l * g.ndim + j];
          res[side] = cell_state(g, nb) == 2 ? t : 0;
        }
      }
      steps[rank * g.L + l] = res[0] | (res[1] << 16);
    }
    return;
  }
  for (int l = 0; l < g.L; l++) {
    int res[2];
    for (int side = 0; side < 2; side++) {
      const int sg = side ? -1 : 1;
      // the nearest known cell along the line (gnn_predictor.py:139-176); the
      // line is live only when that cell is usable
      int hit = 0;
      for (int t = 1; t <= g.limit; t++) {
        for (int j = 0; j < g.ndim; j++) nb[j] = coord[j] + (int64_t)sg * t * dirs[l * g.ndim + j];
        const int cs = cell_state(g, nb);
        if (cs) {
          hit = cs == 2 ? t : 0;
          break;
        }
      }
      res[side] = hit;
    }
    steps[rank * g.L + l] = res[0] | (res[1] << 16);
  }
}

// ---------------------------------------------------------------------------
// live-pair compaction (deterministic, pair order)
// ---------------------------------------------------------------------------

// Halo cells of a chunk frame (the frame is the chunk grown by `pad` cells on
// every side): one thread per frame cell; a cell outside the chunk is usable
// when it lies inside the tensor and in a coded chunk or on the anchor
// lattice (gnn_predictor.py:1338-1356).  Usable cells are appended to
// (frame row, tensor index) lists in any order: each row is written once.
struct HaloArgs {
  int32_t ndim, stride;
  int64_t pad, n_frame;
  int64_t fshape[LATEC_MAX_DIM], cshape[LATEC_MAX_DIM], origin[LATEC_MAX_DIM];
  int64_t tshape[LATEC_MAX_DIM], edges[LATEC_MAX_DIM], grid[LATEC_MAX_DIM];
  const uint8_t *coded;
};

__global__ __launch_bounds__(256) void halo_list_kernel(HaloArgs h, int64_t *__restrict__ frows,
                                                         int64_t *__restrict__ gflat,
                                                         uint32_t *__restrict__ count) {
  const int64_t gstep = (int64_t)gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < h.n_frame; i0 += gstep) {
    const int64_t i = i0 + threadIdx.x;
    bool ok = i < h.n_frame;
    int64_t r = ok ? i : 0, gf = 0, cid = 0, tstr = 1;
    bool inside = true, anchor = true;
    int64_t gc[LATEC_MAX_DIM];
    for (int j = h.ndim - 1; j >= 0; j--) {
      const int64_t c = r % h.fshape[j] - h.pad;
      r /= h.fshape[j];
      inside &= c >= 0 && c < h.cshape[j];
      gc[j] = c + h.origin[j];
      ok &= gc[j] >= 0 && gc[j] < h.tshape[j];
    }
    ok &= !inside;
    if (ok) {
      for (int j = 0; j < h.ndim; j++) {
        cid = cid * h.grid[j] + gc[j] / h.edges[j];
        anchor &= gc[j] % h.stride == 0;
      }
      for (int j = h.ndim - 1; j >= 0; j--) {
        gf += gc[j] * tstr;
        tstr *= h.tshape[j];
      }
      ok = anchor || h.coded[cid];
    }
    const uint64_t m = __ballot(ok);
    if (!m) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (ok) {
      const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      frows[k] = i;
      gflat[k] = gf;
    }
  }
}

constexpr int PAIR_BLOCK = 1024;

__device__ __forceinline__ int pair_kind(int32_t st) {
  const int sp = st & 0xffff, sn = st >> 16;
  return (sp > 0) + (sn > 0);  // 2 bidir, 1 one-sided, 0 dead
}

// counts[b] = (n_bidir, n_dir) of pair block b
__global__ void pair_count_kernel(const int32_t *__restrict__ steps, int64_t n_pairs,
                                  int32_t *__restrict__ counts) {
  __shared__ int sb[PAIR_BLOCK / 64], sd[PAIR_BLOCK / 64];
  const int64_t p = (int64_t)blockIdx.x * PAIR_BLOCK + threadIdx.x;
  const int k = p < n_pairs ? pair_kind(steps[p]) : 0;
  const uint64_t mb = __ballot(k == 2), md = __ballot(k == 1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sb[w] = __popcll(mb);
    sd[w] = __popcll(md);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int i = 0; i < PAIR_BLOCK / 64; i++) a += sb[i], b += sd[i];
    counts[2 * blockIdx.x] = a;
    counts[2 * blockIdx.x + 1] = b;
  }
}

// Exclusive scan of the block counts (one workgroup, sequential chunks).
// totals[0] = n_bidir, totals[1] = n_dir.
__global__ void pair_scan_kernel(int32_t *__restrict__ counts, int64_t nblocks,
                                 int32_t *__restrict__ totals) {
  __shared__ int sa[1024], sb[1024];
  __shared__ int carry_a, carry_b;
  if (threadIdx.x == 0) carry_a = carry_b = 0;
  __syncthreads();
  for (int64_t base = 0; base < nblocks; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int a = i < nblocks ? counts[2 * i] : 0;
    const int b = i < nblocks ? counts[2 * i + 1] : 0;
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
      const int va = threadIdx.x >= off ? sa[threadIdx.x - off] : 0;
      const int vb = threadIdx.x >= off ? sb[threadIdx.x - off] : 0;
      __syncthreads();
      sa[threadIdx.x] += va;
      sb[threadIdx.x] += vb;
      __syncthreads();
    }
    if (i < nblocks) {
      counts[2 * i] = carry_a + sa[threadIdx.x] - a;
      counts[2 * i + 1] = carry_b + sb[threadIdx.x] - b;
    }
    __syncthreads();
    if (threadIdx.x == 1023) {
      carry_a += sa[1023];
      carry_b += sb[1023];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals[0] = carry_a;
    totals[1] = carry_b;
  }
}

// Per live pair, what the message rows read: the neighbour frame rows and the
// two per-row scalars of the first layer (two-sided: log distances of the
// negative and positive neighbour; one-sided: sign and log distance).
struct __align__(16) PairRec {
  int32_t ra, rb;
  float za, zb;
};

// Writes the bidir and one-sided pair lists, their records (+ the line index,
// with the one-sided side in bit 14), and slot[p] (message slot: bidir
// position, or n_bidir + one-sided position; -1 for a dead pair).
__global__ void pair_write_kernel(const int32_t *__restrict__ steps, int64_t n_pairs, int L,
                                  const int64_t *__restrict__ qidx,
                                  const int64_t *__restrict__ line_off,
                                  const float *__restrict__ lognnz,
                                  const int32_t *__restrict__ offs,
                                  const int32_t *__restrict__ totals,
                                  PairRec *__restrict__ rec_bi, int16_t *__restrict__ lw_bi,
                                  PairRec *__restrict__ rec_dir, int16_t *__restrict__ lw_dir,
                                  int32_t *__restrict__ slot) {
  __shared__ int sb[PAIR_BLOCK / 64], sd[PAIR_BLOCK / 64];
  const int64_t p = (int64_t)blockIdx.x * PAIR_BLOCK + threadIdx.x;
  const int32_t stp = p < n_pairs ? steps[p] : 0;
  const int k = pair_kind(stp);
  const uint64_t mb = __ballot(k == 2), md = __ballot(k == 1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sb[w] = __popcll(mb);
    sd[w] = __popcll(md);
  }
  __syncthreads();
  int pb = offs[2 * blockIdx.x], pd = offs[2 * blockIdx.x + 1];
  for (int i = 0; i < w; i++) pb += sb[i], pd += sd[i];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  pb += __popcll(mb & below);
  pd += __popcll(md & below);
  if (p < n_pairs) {
    const int m = (int)(p / L), l = (int)(p % L);
    const int sp = stp & 0xffff, sn = stp >> 16;
    if (k) {
      const int64_t q = qidx[m], off = line_off[l];
      const float lnz = lognnz[l];
      PairRec r;
      if (k == 2) {
        r.ra = (int32_t)(q - (int64_t)sn * off);
        r.rb = (int32_t)(q + (int64_t)sp * off);
        r.za = log2f((float)sn) + lnz;
        r.zb = log2f((float)sp) + lnz;
        rec_bi[pb] = r;
        lw_bi[pb] = (int16_t)l;
        slot[p] = pb;
      } else {
        const bool pos = sp > 0;
        const int t = pos ? sp : sn;
        r.ra = (int32_t)(q + (pos ? (int64_t)t : -(int64_t)t) * off);
        r.rb = 0;
        r.za = pos ? 1.0f : -1.0f;
        r.zb = log2f((float)t) + lnz;
        rec_dir[pd] = r;
        lw_dir[pd] = (int16_t)(l | (pos ? 1 << 14 : 0));
        slot[p] = totals[0] + pd;
      }
    } else {
      slot[p] = -1;
    }
  }
}

// ---------------------------------------------------------------------------
// MFMA tiles in the transposed form
// ---------------------------------------------------------------------------
//
// A 16x16 MFMA tile computes D = A B + C.  Every layer below is evaluated
// transposed, Y^T = W X^T: A is a 16-row block of the weight matrix (16
// output units), B holds 16 data rows as its columns.  Lane l = 16 g + r then
// always works on data row r: it supplies its row's inputs for the k-set of
// lane group g and receives units 16 ob + 4 g + v (v = 0..3) of its row.  The
// next layer's B operand is made from those registers directly, so hidden
// activations never leave the register file; a layer fed from registers has
// its weights stored with the input index in that register order:
//   bf16 (v_mfma_f32_16x16x32_bf16): k-step of 32, group g supplies 8
//        consecutive k; register position 32 s + 8 g + i holds unit
//        16 (2 s + i / 4) + 4 g + i % 4;
//   f32  (v_mfma_f32_16x16x4f32): k-step of 4, group g supplies one k;
//        register position 4 t + g holds unit 16 (t / 4) + 4 g + t % 4.
// Rows gathered from memory in f32 use the same f32 order (a lane loads four
// consecutive values at 16 q + 4 g and feeds them to steps 4 q .. 4 q + 3);
// bf16 rows gathered from memory are in natural order.

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<uint16_t> {
  static constexpr int KS = 32, KL = 8, NX = 1, PE = 8;  // NX: k-steps of additive terms
  typedef bf16x8 frag;
  typedef bf16x4 quad;
};
template <>
struct Mf<float> {
  static constexpr int KS = 4, KL = 1, NX = 4, PE = 4;
  typedef float frag;
  typedef f32x4 quad;
};

// register position -> unit of a hidden vector (see above)
__host__ __device__ inline int reg_unit(bool bf, int kk) {
  if (bf) {
    const int s = kk / 32, g = (kk % 32) / 8, i = kk % 8;
    return 16 * (2 * s + i / 4) + 4 * g + i % 4;
  }
  const int t = kk / 4, g = kk % 4;
  return 16 * (t / 4) + 4 * g + t % 4;
}

__device__ __forceinline__ void mfma(f32x4 &acc, bf16x8 a, bf16x8 b) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ void mfma(f32x4 &acc, float a, float b) {
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ typename Mf<T>::frag ldfrag(const T *p) {
  if constexpr (sizeof(T) == 2)
    return *(const bf16x8 *)p;
  else
    return *p;
}

// B operand of k-step s from the quads (16-unit blocks) of one row
template <typename T, int NQ>
__device__ __forceinline__ typename Mf<T>::frag bfrag(const typename Mf<T>::quad (&q)[NQ],
                                                      int s) {
  if constexpr (sizeof(T) == 2)
    return __builtin_shufflevector(q[2 * s], q[2 * s + 1], 0, 1, 2, 3, 4, 5, 6, 7);
  else
    return q[s / 4][s % 4];
}

template <typename T>
__device__ __forceinline__ typename Mf<T>::quad to_quad(f32x4 v) {
  if constexpr (sizeof(T) == 2)
    return __builtin_convertvector(v, bf16x4);
  else
    return v;
}

// GELU of the bf16 path: x Phi(x), Phi(x) = 1/2 + xc P(xc^2) with xc = x
// clamped to [-4, 4] and P of degree 7 from a least-squares fit of the normal
// CDF on [-4, 4] with Phi(+-4) pinned to 1 and 0.  Evaluated in float32 the
// error is below 1.5e-4 absolute and 5.7e-4 relative over the whole line
// (the result is rounded to bf16, half an ulp of which is 2e-3).
__device__ __forceinline__ float gelu_fast(float x) {
  const float xc = __builtin_amdgcn_fmed3f(x, -4.0f, 4.0f);
  const float t = xc * xc;
  float p = -1.491892165e-09f;
  p = __builtin_fmaf(p, t, 1.169115293e-07f);
  p = __builtin_fmaf(p, t, -3.999352884e-06f);
  p = __builtin_fmaf(p, t, 7.960923540e-05f);
  p = __builtin_fmaf(p, t, -1.042630989e-03f);
  p = __builtin_fmaf(p, t, 9.651600383e-03f);
  p = __builtin_fmaf(p, t, -6.616644561e-02f);
  p = __builtin_fmaf(p, t, 3.988512158e-01f);
  return x * __builtin_fmaf(xc, p, 0.5f);
}

template <typename T>
__device__ __forceinline__ float gelu_t(float x) {
  if constexpr (sizeof(T) == 2)
    return gelu_fast(x);
  else
    return gelu(x);
}

// Wave-local LDS hand-off: every lane's LDS writes are complete before any
// lane reads (one wave; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// block -> logical block so that each XCD (blocks are dispatched round-robin
// over 8 of them) works on one contiguous range
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// message / mix rows on MFMA
// ---------------------------------------------------------------------------

struct RowArgs {
  int32_t K, L;
  const int32_t *n_items_dev;  // message items (pairs), read on device
  int64_t n_rows_host;         // finalize rows (M * K)
  const PairRec *rec;          // messages: per list entry
  const int16_t *lw;           // messages: line (| side << 14)
  const int8_t *sidx;          // (L, K, 2) state index per line, axis, side
  const void *E;               // (frame rows, K, D)
  void *out;                   // msg (slots, K, D) / E for finalize
  float *score;                // (slots, K) for messages
  int64_t slot_base;           // slot of list[0] (0 for bidir)
  const int32_t *slot_base_dev;
  const void *ctx;             // finalize: (M, K, D)
  const void *vemb;            // finalize: (M, D)
  const int64_t *qout;         // finalize: frame index of each finalized point
  float score_scale;
  uint32_t kmagic;              // ceil(2^16 / K): n / K = (n * kmagic) >> 16 for n < 2^10
};

// Weights of one row MLP, as staged: W1 (H, K1X) with the first-layer input
// in register order and NX k-steps of additive terms appended; W2 (D, H + KS)
// and the key's Wk (D, D + KS) with their inputs in register order and one
// bias k-step appended (bf16: high and low parts in its first two columns).
template <typename T>
struct RowW {
  const T *w1, *w2, *wk;
  const float *wk2;
  float bk2;
};

template <int D, typename T, int MODE>
struct RowShape {
  typedef Mf<T> F;
  static constexpr int H = 2 * D;
  static constexpr int K1 = MODE == ROW_DIR ? D : 2 * D;
  static constexpr int K1X = K1 + F::NX * F::KS;  // + additive terms
  static constexpr int K2X = H + F::KS, K3X = D + F::KS;  // + one bias step
  static constexpr int P1 = K1X + F::PE, P2 = K2X + F::PE, PK = K3X + F::PE;
  static constexpr bool KEY = MODE != ROW_FIN;
  static size_t smem(int L, int K) {
    return sizeof(T) * ((size_t)H * P1 + (size_t)D * P2 + (KEY ? (size_t)D * PK : 0)) +
           sizeof(float) * D + (size_t)((L * K * 2 + 15) & ~15);
  }
};

// Row r of the layer-1 input: its two source rows (or one), the two per-row
// scalars and the state index of the additive terms, and where the output
// row goes.
template <typename T, int NK1>
struct RowIn {
  typename Mf<T>::frag x[NK1];
  float z0, z1;
  int st;
  int64_t dst;  // output row (element offset / D); -1 if the row is past the end
};

// two floats -> two bf16 (round to nearest even) in one word, a in the low half
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// Additive-term k-step(s) of layer 1 (the B side; W1's matching columns hold
// b1, the scalar coefficients and the state bias rows):
//   bf16, one step of 32: group 0 [1, 1, z0h, z0h, z0l, z1h, z1h, z1l] against
//     [b1h, b1l, c0h, c0l, c0h, c1h, c1l, c1h]; groups 1 and 2 one-hot(st) over
//     8 states against the state rows' high and low bf16 parts; group 3 zero
//     (z = zh + zl and every weight = hi + lo split, so the terms keep ~16 bits);
//   f32, four steps of 4 (position e = 4 t + g): [1, z0, z1, one-hot(st) x 13].
template <typename T>
__device__ __forceinline__ typename Mf<T>::frag extra_frag(float z0, float z1, int st, int g,
                                                           int t) {
  if constexpr (sizeof(T) == 2) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const uint32_t h0 = pk_bf16(z0, z0), h1 = pk_bf16(z1, z1);
    const float l0 = z0 - __builtin_bit_cast(float, h0 << 16);
    const float l1 = z1 - __builtin_bit_cast(float, h1 << 16);
    // words (element 0 in the low half): [1 1] [z0h z0h] [z0l z1h] [z1h z1l]
    const uint32_t w2 = pk_bf16(l0, z1), w3 = pk_bf16(z1, l1);
    const uint32_t one = 0x3F80u << ((st & 1) * 16), pr = (uint32_t)st >> 1;
    u32x4 v;
    if (g == 0) {
      v = u32x4{0x3F803F80u, h0, w2, w3};
    } else if (g < 3) {
      v = u32x4{pr == 0 ? one : 0u, pr == 1 ? one : 0u, pr == 2 ? one : 0u, pr == 3 ? one : 0u};
    } else {
      v = u32x4{0u, 0u, 0u, 0u};
    }
    return __builtin_bit_cast(bf16x8, v);
  } else {
    const int e = 4 * t + g;
    return e == 0 ? 1.0f : e == 1 ? z0 : e == 2 ? z1 : (e - 3 == st ? 1.0f : 0.0f);
  }
}

// B side of a bias k-step: 1 at the two positions of group 0 (bf16: the bias'
// high and low parts) or the first (f32), 0 elsewhere.
template <typename T>
__device__ __forceinline__ typename Mf<T>::frag bias_frag(int g) {
  if constexpr (sizeof(T) == 2) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {g == 0 ? 0x3F803F80u : 0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, v);
  } else {
    return g == 0 ? 1.0f : 0.0f;
  }
}

// Row MLP: for each row r (item i = r / K, axis k = r % K)
//   h = gelu(W1 [x0 | x1] + b1 + S[st] + c0 z0 + c1 z1)   (H)
//   y = W2 h + b2                                         (D)
// messages also get score = (wk2 . gelu(Wk y + bk1) + bk2) * scale.  The
// biases ride on MFMA k-steps.  Each wave walks groups of NT 16-row tiles; the
// next group's inputs are gathered while this group's epilogue and key run.
template <int D, typename T, int MODE, int NT, int WV, int MINW = 1>
__global__ __launch_bounds__(64 * WV) __attribute__((amdgpu_waves_per_eu(MINW))) void msg_kernel(
    RowArgs a, RowW<T> w) {
  typedef RowShape<D, T, MODE> S;
  typedef Mf<T> F;
  typedef typename F::frag frag;
  typedef typename F::quad quad;
  constexpr bool BF = sizeof(T) == 2;
  constexpr int KS = F::KS, KL = F::KL, H = S::H, K1 = S::K1, K1X = S::K1X;
  constexpr int K2X = S::K2X, K3X = S::K3X;
  constexpr int NK1 = K1 / KS, NKX = F::NX, NK3 = D / KS;
  constexpr int P1 = S::P1, P2 = S::P2, PK = S::PK;
  constexpr int NB1 = H / 16, NB2 = D / 16;
  constexpr bool KEY = S::KEY;
  constexpr int ROWS = 16 * NT;
  extern __shared__ __align__(16) unsigned char smem[];
  T *sW1 = (T *)smem;
  T *sW2 = sW1 + H * P1;
  T *sWk = sW2 + D * P2;
  float *twk2 = (float *)(sWk + (KEY ? D * PK : 0));
  int8_t *tsx = (int8_t *)(twk2 + D);
  const int tid = threadIdx.x;
  constexpr int NTHR = 64 * WV;
  for (int i = tid; i < H * K1X; i += NTHR) sW1[(i / K1X) * P1 + i % K1X] = w.w1[i];
  for (int i = tid; i < D * K2X; i += NTHR) sW2[(i / K2X) * P2 + i % K2X] = w.w2[i];
  if constexpr (KEY) {
    for (int i = tid; i < D * K3X; i += NTHR) sWk[(i / K3X) * PK + i % K3X] = w.wk[i];
    for (int i = tid; i < D; i += NTHR) twk2[i] = w.wk2[i];
    for (int i = tid; i < a.L * a.K * 2; i += NTHR) tsx[i] = a.sidx[i];
  }
  __syncthreads();

  const int lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  const int K = a.K;
  const int64_t n_rows = MODE == ROW_FIN ? a.n_rows_host : (int64_t)a.n_items_dev[0] * K;
  const int64_t slot_base = a.slot_base_dev ? (int64_t)a.slot_base_dev[0] : a.slot_base;
  const T *E = (const T *)a.E;
  const T *CTX = (const T *)a.ctx;
  const T *VEMB = (const T *)a.vemb;
  const int64_t gstep = (int64_t)gridDim.x * WV;
  int64_t grp = (int64_t)xcd_block() * WV + wv;
  const frag bfr = bias_frag<T>(g);

  // gather the layer-1 inputs of group gi
  auto load = [&](int64_t gi, RowIn<T, NK1> (&in)[NT]) {
    // row = base + o (o < ROWS): item = ib + (kb + o) / K, k = (kb + o) % K
    const int64_t base = gi * ROWS;
    const int32_t ib = (int32_t)(base / K), kb = (int32_t)(base - (int64_t)ib * K);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int64_t row = base + 16 * t + r;
      const T *s0, *s1;
      float z0 = 0.f, z1 = 0.f;
      int st = 0;
      int64_t dst = -1;
      if (row < n_rows) {
        const int32_t o = kb + 16 * t + r;  // < K + ROWS
        const int32_t qo = (int32_t)(((uint32_t)o * a.kmagic) >> 16);
        const int32_t item = ib + qo, k = o - qo * K;
        if constexpr (MODE == ROW_FIN) {
          s0 = CTX + row * D;
          s1 = VEMB + (int64_t)item * D;
          dst = a.qout[item] * K + k;
        } else {
          const PairRec rc = a.rec[item];
          const int16_t lw = a.lw[item];
          const int l = lw & 0x3fff;
          s0 = E + ((int64_t)rc.ra * K + k) * D;
          s1 = MODE == ROW_BI ? E + ((int64_t)rc.rb * K + k) * D : s0;
          st = tsx[(l * K + k) * 2 + (MODE == ROW_DIR ? (lw >> 14) & 1 : 0)];
          z0 = rc.za;
          z1 = rc.zb;
          dst = slot_base * K + row;
        }
      } else {
        s0 = s1 = MODE == ROW_FIN ? CTX : E;  // row 0: a valid address, result unused
      }
      if constexpr (BF) {
#pragma unroll
        for (int s = 0; s < NK1; s++) {
          const bool h1 = MODE != ROW_DIR && s * KS >= D;
          const T *src = h1 ? s1 : s0;
          in[t].x[s] = *(const bf16x8 *)(src + (s * KS - (h1 ? D : 0)) + 8 * g);
        }
      } else {
#pragma unroll
        for (int q = 0; q < K1 / 16; q++) {
          const bool h1 = MODE != ROW_DIR && q * 16 >= D;
          const T *src = h1 ? s1 : s0;
          const f32x4 v = *(const f32x4 *)(src + (q * 16 - (h1 ? D : 0)) + 4 * g);
#pragma unroll
          for (int e = 0; e < 4; e++) in[t].x[4 * q + e] = v[e];
        }
      }
      in[t].z0 = z0;
      in[t].z1 = z1;
      in[t].st = st;
      in[t].dst = dst;
    }
  };

  RowIn<T, NK1> in[NT];
  if (grp * ROWS < n_rows) load(grp, in);
  for (; grp * ROWS < n_rows; grp += gstep) {
    frag xe[NT][NKX];
    int64_t dst[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
      dst[t] = in[t].dst;
#pragma unroll
      for (int s = 0; s < NKX; s++) xe[t][s] = extra_frag<T>(in[t].z0, in[t].z1, in[t].st, g, s);
    }
    // layer-1 blocks in groups that form one layer-2 k-step group each (bf16:
    // two blocks = one step of 32; f32: one block = four steps of 4), so the
    // hidden vector is consumed as it is made; layer 2 starts from its bias step
    constexpr int GB = BF ? 2 : 1, SPG = BF ? 1 : 4;
    f32x4 acc2[NT][NB2];
#pragma unroll
    for (int ob = 0; ob < NB2; ob++) {
      const frag af = ldfrag<T>(sW2 + (16 * ob + r) * P2 + g * KL + H);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        acc2[t][ob] = f32x4{0.f, 0.f, 0.f, 0.f};
        mfma(acc2[t][ob], af, bfr);
      }
    }
#pragma unroll 1
    for (int j = 0; j < NB1 / GB; j++) {
      quad hq[NT][GB];
#pragma unroll
      for (int b = 0; b < GB; b++) {
        f32x4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        const T *wr = sW1 + (16 * (j * GB + b) + r) * P1 + g * KL;
#pragma unroll
        for (int s = 0; s < NK1 + NKX; s++) {
          const frag af = ldfrag<T>(wr + s * KS);
#pragma unroll
          for (int t = 0; t < NT; t++) mfma(acc[t], af, s < NK1 ? in[t].x[s] : xe[t][s - NK1]);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
          f32x4 hv;
#pragma unroll
          for (int v = 0; v < 4; v++) hv[v] = gelu_t<T>(acc[t][v]);
          hq[t][b] = to_quad<T>(hv);
        }
      }
#pragma unroll
      for (int ob = 0; ob < NB2; ob++) {
        const T *wr = sW2 + (16 * ob + r) * P2 + g * KL + j * SPG * KS;
#pragma unroll
        for (int ss = 0; ss < SPG; ss++) {
          const frag af = ldfrag<T>(wr + ss * KS);
#pragma unroll
          for (int t = 0; t < NT; t++) mfma(acc2[t][ob], af, bfrag<T, GB>(hq[t], ss));
        }
      }
    }
    // the next group's gathers fly during the epilogue and the key
    if ((grp + gstep) * ROWS < n_rows) load(grp + gstep, in);
    // --- the output rows
    quad mq[NT][NB2];
    T *out = (T *)a.out;
#pragma unroll
    for (int ob = 0; ob < NB2; ob++)
#pragma unroll
      for (int t = 0; t < NT; t++) {
        mq[t][ob] = to_quad<T>(acc2[t][ob]);
        if (dst[t] >= 0) *(quad *)(out + dst[t] * D + 16 * ob + 4 * g) = mq[t][ob];
      }
    // --- attention key of each message row
    if constexpr (KEY) {
      float part[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) part[t] = 0.f;
#pragma unroll 1
      for (int ob = 0; ob < NB2; ob++) {
        f32x4 acc[NT];
        const T *wr = sWk + (16 * ob + r) * PK + g * KL;
        {
          const frag af = ldfrag<T>(wr + D);
#pragma unroll
          for (int t = 0; t < NT; t++) {
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            mfma(acc[t], af, bfr);
          }
        }
#pragma unroll
        for (int s = 0; s < NK3; s++) {
          const frag af = ldfrag<T>(wr + s * KS);
#pragma unroll
          for (int t = 0; t < NT; t++) mfma(acc[t], af, bfrag<T, NB2>(mq[t], s));
        }
        const f32x4 wk = *(const f32x4 *)(twk2 + 16 * ob + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int v = 0; v < 4; v++) part[t] = __builtin_fmaf(wk[v], gelu_t<T>(acc[t][v]), part[t]);
      }
#pragma unroll
      for (int t = 0; t < NT; t++) {
        float v = part[t];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (g == 0 && dst[t] >= 0) a.score[dst[t]] = (v + w.bk2) * a.score_scale;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// line pool: softmax over each (point, axis) item's live lines + null token,
// value projection on an f32 MFMA tile
// ---------------------------------------------------------------------------

struct PoolArgs {
  const float *wv;      // (D, D) value projection, natural order
  const float *bv, *null_v;
  float null_score;     // null_k * scale
  const uint16_t *wv_hi, *wv_lo;  // bf16 mode: wv as high + low bf16 parts
};

constexpr int POOL_WAVES = 4;

// LDS of the pool kernels: the value projection (bf16 mode: high and low
// bf16 parts, rows of D + 8; fp32: rows of D + 4 floats), bv, null_v, then
// per wave 16 pooled rows and 16 null weights.
template <int D, typename T>
struct PoolLds {
  static constexpr bool BF = sizeof(T) == 2;
  static constexpr int PW = D + 4, PB = D + 8;
  static constexpr size_t WBYTES = BF ? 2 * sizeof(uint16_t) * D * PB : sizeof(float) * D * PW;
  static size_t bytes() {
    return WBYTES + sizeof(float) * (2 * (size_t)D + POOL_WAVES * (16 * (size_t)PW + 16));
  }
};

template <int D, typename T>
__device__ __forceinline__ void pool_load_weights(unsigned char *smem, const PoolArgs &pa,
                                                  int tid) {
  typedef PoolLds<D, T> S;
  float *sbv = (float *)(smem + S::WBYTES), *snv = sbv + D;
  if constexpr (S::BF) {
    uint16_t *sWh = (uint16_t *)smem, *sWl = sWh + D * S::PB;
    for (int i = tid; i < D * D; i += 64 * POOL_WAVES) {
      sWh[(i / D) * S::PB + i % D] = pa.wv_hi[i];
      sWl[(i / D) * S::PB + i % D] = pa.wv_lo[i];
    }
  } else {
    float *sWv = (float *)smem;
    for (int i = tid; i < D * D; i += 64 * POOL_WAVES) sWv[(i / D) * S::PW + i % D] = pa.wv[i];
  }
  for (int i = tid; i < D; i += 64 * POOL_WAVES) {
    sbv[i] = pa.bv[i];
    snv[i] = pa.null_v[i];
  }
}

// split x = hi + lo into two bf16 fragments (hi = RNE(x), lo = RNE(x - hi))
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const __bf16 h = (__bf16)x[i];
    hi[i] = h;
    lo[i] = (__bf16)(x[i] - (float)h);
  }
}

// Value projection of a wave's 16 pooled items (rows of sP, null weights in
// sWn): ctx^T = Wv P^T + (1 - w_null) bv + w_null null_v.  fp32 mode on
// v_mfma_f32_16x16x4f32; bf16 mode on v_mfma_f32_16x16x32_bf16 with split
// operands, Wv_hi P_hi + Wv_lo P_hi + Wv_hi P_lo (about 16 bits of each
// operand; the result is stored as bf16).
template <int D, typename T>
__device__ __forceinline__ void pool_project(const unsigned char *smem, const float *sP,
                                             const float *sWn, int lane, int64_t it0,
                                             int64_t n_items, int64_t item_base, int K,
                                             T *__restrict__ ctx) {
  typedef PoolLds<D, T> S;
  constexpr int PW = S::PW, PB = S::PB;
  const float *sbv = (const float *)(smem + S::WBYTES), *snv = sbv + D;
  const int r = lane & 15, g = lane >> 4;  // item, k group
  const float wn = sWn[r];
  const int64_t item = it0 + r;
  [[maybe_unused]] bf16x8 ph[D / 32], pl[D / 32];
  if constexpr (S::BF) {
#pragma unroll
    for (int q = 0; q < D / 32; q++) {
      float pv[8];
#pragma unroll
      for (int i = 0; i < 8; i++) pv[i] = sP[r * PW + 32 * q + 8 * g + i];
      split_bf16(pv, ph[q], pl[q]);
    }
  }
#pragma unroll
  for (int ob = 0; ob < D / 16; ob++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if constexpr (S::BF) {
      const __bf16 *wr = (const __bf16 *)smem + (16 * ob + r) * PB + 8 * g;
#pragma unroll
      for (int q = 0; q < D / 32; q++) {
        const bf16x8 a_hi = *(const bf16x8 *)(wr + 32 * q);
        const bf16x8 a_lo = *(const bf16x8 *)(wr + D * PB + 32 * q);
        mfma(acc, a_hi, ph[q]);
        mfma(acc, a_lo, ph[q]);
        mfma(acc, a_hi, pl[q]);
      }
    } else {
      const float *wr = (const float *)smem + (16 * ob + r) * PW + g;
      const float *pr = sP + r * PW + g;
#pragma unroll
      for (int t = 0; t < D / 4; t++) mfma(acc, wr[4 * t], pr[4 * t]);
    }
    const int o = 16 * ob + 4 * g;
    f32x4 y;
#pragma unroll
    for (int v = 0; v < 4; v++) y[v] = acc[v] + (1.0f - wn) * sbv[o + v] + wn * snv[o + v];
    if (item < n_items)
      *(typename Mf<T>::quad *)(ctx + (item_base * K + item) * D + o) = to_quad<T>(y);
  }
}

// 16 items per wave: 4 lanes per item for the softmax-weighted sum of its
// messages (D / 4 columns each), line by line, then the value projection.
template <int D, typename T>
__global__ __launch_bounds__(64 * POOL_WAVES) void pool_kernel(
    int64_t n_items, int K, int L, const int32_t *__restrict__ slot, const T *__restrict__ msg,
    const float *__restrict__ score, PoolArgs pa, T *__restrict__ ctx, int64_t item_base) {
  constexpr int PW = D + 4, DC = D / 4;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  float *sP = (float *)(smem + PoolLds<D, T>::WBYTES) + 2 * D + wv * (16 * PW + 16);
  float *sWn = sP + 16 * PW;
  pool_load_weights<D, T>(smem, pa, tid);
  __syncthreads();
  const int ii = lane >> 2, c4 = lane & 3;  // softmax phase: item, column quarter
  const int64_t gstep = (int64_t)gridDim.x * POOL_WAVES * 16;
  for (int64_t it0 = ((int64_t)xcd_block() * POOL_WAVES + wv) * 16; it0 < n_items; it0 += gstep) {
    {
      int64_t it = it0 + ii;
      if (it >= n_items) it = n_items - 1;
      const int64_t m = it / K;
      const int k = (int)(it - m * K);
      const int32_t *sl = slot + m * L;
      float mx = pa.null_score;
      for (int l = 0; l < L; l++) {
        const int32_t s = sl[l];
        if (s >= 0) mx = fmaxf(mx, score[(int64_t)s * K + k]);
      }
      float den = 0.f, acc[DC];
#pragma unroll
      for (int j = 0; j < DC; j++) acc[j] = 0.f;
      for (int l = 0; l < L; l++) {
        const int32_t s = sl[l];
        if (s < 0) continue;
        const float e = __expf(score[(int64_t)s * K + k] - mx);
        den += e;
        const T *mp = msg + ((int64_t)s * K + k) * D + c4 * DC;
        if constexpr (sizeof(T) == 2) {
#pragma unroll
          for (int j = 0; j < DC; j += 8) {
            const bf16x8 v = *(const bf16x8 *)(mp + j);
#pragma unroll
            for (int e2 = 0; e2 < 8; e2++) acc[j + e2] = __builtin_fmaf(e, (float)v[e2], acc[j + e2]);
          }
        } else {
#pragma unroll
          for (int j = 0; j < DC; j += 4) {
            const f32x4 v = *(const f32x4 *)(mp + j);
#pragma unroll
            for (int e2 = 0; e2 < 4; e2++) acc[j + e2] = __builtin_fmaf(e, v[e2], acc[j + e2]);
          }
        }
      }
      const float en = __expf(pa.null_score - mx);
      den += en;
      const float inv = 1.0f / den;
#pragma unroll
      for (int j = 0; j < DC; j++) sP[ii * PW + c4 * DC + j] = acc[j] * inv;
      if (c4 == 0) sWn[ii] = en * inv;
    }
    wave_lds_sync();
    pool_project<D, T>(smem, sP, sWn, lane, it0, n_items, item_base, K, ctx);
    wave_lds_sync();
  }
}

// The same pool for items of at most POOL_LMAX lines, with every load issued
// ahead of its use: the item's slots, then the scores of all its lines, then
// the messages of LB lines at a time (dead lines load slot 0's and select
// zeros); bf16 with LB = 4 holds its registers to four waves per SIMD.  Each live line adds e * m in line order exactly as in
// pool_kernel (a dead line adds +0, which leaves a sum that starts at +0
// unchanged), so both kernels give the same bits.
constexpr int POOL_LMAX = 16;

template <int D, typename T, int LB>
__global__ __launch_bounds__(64 * POOL_WAVES)
__attribute__((amdgpu_waves_per_eu(LB == 4 && sizeof(T) == 2 ? 4 : 1))) void pool_lines_kernel(
    int64_t n_items, int K, int L, const int32_t *__restrict__ slot, const T *__restrict__ msg,
    const float *__restrict__ score, PoolArgs pa, T *__restrict__ ctx, int64_t item_base) {
  constexpr int PW = D + 4, DC = D / 4;
  constexpr bool BF = sizeof(T) == 2;
  static_assert(POOL_LMAX % LB == 0, "lines per batch of message loads");
  constexpr int NV = BF ? DC / 8 : DC / 4;          // vector loads per line
  typedef typename std::conditional<BF, bf16x8, f32x4>::type vec;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  float *sP = (float *)(smem + PoolLds<D, T>::WBYTES) + 2 * D + wv * (16 * PW + 16);
  float *sWn = sP + 16 * PW;
  pool_load_weights<D, T>(smem, pa, tid);
  __syncthreads();
  const int ii = lane >> 2, c4 = lane & 3;
  const int64_t gstep = (int64_t)gridDim.x * POOL_WAVES * 16;
  for (int64_t it0 = ((int64_t)xcd_block() * POOL_WAVES + wv) * 16; it0 < n_items; it0 += gstep) {
    {
      int64_t it = it0 + ii;
      if (it >= n_items) it = n_items - 1;
      const int64_t m = it / K;
      const int k = (int)(it - m * K);
      const int32_t *sl = slot + m * L;
      int32_t s[POOL_LMAX];
#pragma unroll
      for (int l = 0; l < POOL_LMAX; l++) s[l] = l < L ? sl[l] : -1;
      float sc[POOL_LMAX];
#pragma unroll
      for (int l = 0; l < POOL_LMAX; l++) sc[l] = score[(int64_t)(s[l] < 0 ? 0 : s[l]) * K + k];
      float mx = pa.null_score;
#pragma unroll
      for (int l = 0; l < POOL_LMAX; l++)
        if (s[l] >= 0) mx = fmaxf(mx, sc[l]);
      float den = 0.f, acc[DC];
#pragma unroll
      for (int j = 0; j < DC; j++) acc[j] = 0.f;
#pragma unroll
      for (int l0 = 0; l0 < POOL_LMAX; l0 += LB) {
        vec v[LB][NV];
#pragma unroll
        for (int b = 0; b < LB; b++) {
          const int32_t q = s[l0 + b] < 0 ? 0 : s[l0 + b];
          const T *mp = msg + ((int64_t)q * K + k) * D + c4 * DC;
#pragma unroll
          for (int u = 0; u < NV; u++) v[b][u] = *(const vec *)(mp + u * (BF ? 8 : 4));
        }
#pragma unroll
        for (int b = 0; b < LB; b++) {
          const bool live = s[l0 + b] >= 0;
          const float e = live ? __expf(sc[l0 + b] - mx) : 0.f;
          den += e;
#pragma unroll
          for (int u = 0; u < NV; u++) {
            if constexpr (BF) {
              typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
              u32x4 w = __builtin_bit_cast(u32x4, v[b][u]);
#pragma unroll
              for (int z = 0; z < 4; z++) w[z] = live ? w[z] : 0u;
              const bf16x8 vv = __builtin_bit_cast(bf16x8, w);
#pragma unroll
              for (int e2 = 0; e2 < 8; e2++)
                acc[8 * u + e2] = __builtin_fmaf(e, (float)vv[e2], acc[8 * u + e2]);
            } else {
#pragma unroll
              for (int e2 = 0; e2 < 4; e2++)
                acc[4 * u + e2] = __builtin_fmaf(e, live ? v[b][u][e2] : 0.f, acc[4 * u + e2]);
            }
          }
        }
      }
      const float en = __expf(pa.null_score - mx);
      den += en;
      const float inv = 1.0f / den;
#pragma unroll
      for (int j = 0; j < DC; j++) sP[ii * PW + c4 * DC + j] = acc[j] * inv;
      if (c4 == 0) sWn[ii] = en * inv;
    }
    wave_lds_sync();
    pool_project<D, T>(smem, sP, sWn, lane, it0, n_items, item_base, K, ctx);
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// head: axis pool + PredHead on f32 MFMA tiles, 16 points per wave
// ---------------------------------------------------------------------------

struct HeadArgs {
  int64_t M;
  int K;
  float scale, log_eb, vmin, span, vmax;
  const float *wkv;     // (2D, D): axis-pool key layer 1 over value projection, input in f32 order
  const float *bk1, *wk2, *bv, *null_v;
  float bk2, null_k;
  const float *w1;      // (H, D) head.net.0 weight without its log-eb column
  const float *cond;    // (H) head.net.0 weight of the log-eb column
  const float *b1;      // (H)
  const float *w2;      // (2, H)
  const float *b2;      // (2)
  // bf16 mode (head_split_kernel): the two GEMM weights as high + low bf16
  // parts, inputs in natural order
  const uint16_t *wkv_hi, *wkv_lo;  // (2D, D)
  const uint16_t *w1_hi, *w1_lo;    // (H, D)
};

constexpr int HEAD_WAVES = 8;

template <int D>
static size_t head_smem() {
  constexpr int H = 2 * D, PW = D + 4;
  return sizeof(float) * ((size_t)2 * D * PW + (size_t)H * PW + 4 * (size_t)D + 3 * (size_t)H +
                          HEAD_WAVES * (2 * 16 * (size_t)PW + 32));
}

// Phase A per 16-row tile of (point, axis) rows (16 / K points): keys and
// values of the axis pool in one (2D x D) GEMM, the softmax over the point's
// axes + null through LDS, the pooled context of each point into LDS.
// Phase B over the 16 pooled points: z = gelu(W1 p + b1 + cond log_eb) and
// (mu, log_b) = W2 z + b2.
template <int D, typename T>
__global__ __launch_bounds__(64 * HEAD_WAVES) void head2_kernel(const T *__restrict__ ctx,
                                                                HeadArgs h,
                                                                float *__restrict__ pred,
                                                                float *__restrict__ scl,
                                                                float *__restrict__ mu_out,
                                                                float *__restrict__ lb_out) {
  constexpr int H = 2 * D, PW = D + 4;
  extern __shared__ __align__(16) unsigned char smem[];
  float *sWkv = (float *)smem;
  float *sW1 = sWkv + 2 * D * PW;
  float *sbk1 = sW1 + H * PW, *swk2 = sbk1 + D, *sbv = swk2 + D, *snv = sbv + D;
  float *sb1 = snv + D, *sw20 = sb1 + H, *sw21 = sw20 + H;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  float *sV = sw21 + H + wv * (2 * 16 * PW + 32);
  float *sPool = sV + 16 * PW;
  float *sS = sPool + 16 * PW;  // 16 scores, 16 null weights
  for (int i = tid; i < 2 * D * D; i += 64 * HEAD_WAVES) sWkv[(i / D) * PW + i % D] = h.wkv[i];
  for (int i = tid; i < H * D; i += 64 * HEAD_WAVES) sW1[(i / D) * PW + i % D] = h.w1[i];
  for (int i = tid; i < D; i += 64 * HEAD_WAVES) {
    sbk1[i] = h.bk1[i];
    swk2[i] = h.wk2[i];
    sbv[i] = h.bv[i];
    snv[i] = h.null_v[i];
  }
  for (int i = tid; i < H; i += 64 * HEAD_WAVES) {
    sb1[i] = h.b1[i] + h.cond[i] * h.log_eb;
    sw20[i] = h.w2[i];
    sw21[i] = h.w2[H + i];
  }
  __syncthreads();
  const int K = h.K, PPT = 16 / K, NTA = (16 + PPT - 1) / PPT;
  const float sn = h.null_k * h.scale;
  const int64_t gstep = (int64_t)gridDim.x * HEAD_WAVES * 16;
  for (int64_t m0 = ((int64_t)xcd_block() * HEAD_WAVES + wv) * 16; m0 < h.M; m0 += gstep) {
    for (int ta = 0; ta < NTA; ta++) {
      const int pl = ta * PPT + r / K, k = r % K;  // this lane's row: local point, axis
      const bool rv = r < PPT * K && pl < 16 && m0 + pl < h.M;
      const int64_t row = rv ? (m0 + pl) * K + k : 0;
      f32x4 x[D / 16];
      const T *src = ctx + row * D + 4 * g;
#pragma unroll
      for (int q = 0; q < D / 16; q++) {
        if constexpr (sizeof(T) == 2) {
          const bf16x4 v = *(const bf16x4 *)(src + 16 * q);
#pragma unroll
          for (int e = 0; e < 4; e++) x[q][e] = (float)v[e];
        } else {
          x[q] = *(const f32x4 *)(src + 16 * q);
        }
      }
      float part = 0.f;
#pragma unroll 1
      for (int ob = 0; ob < D / 16; ob++) {  // key hidden units
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const float *wr = sWkv + (16 * ob + r) * PW + g;
#pragma unroll
        for (int t = 0; t < D / 4; t++) mfma(acc, wr[4 * t], x[t / 4][t % 4]);
        const int o = 16 * ob + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; v++) part += swk2[o + v] * gelu(acc[v] + sbk1[o + v]);
      }
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      if (g == 0) sS[r] = (part + h.bk2) * h.scale;
      wave_lds_sync();
      // softmax over the point's axes + null
      float mx = sn;
      const int p0 = (r / K) * K;
      for (int kk = 0; kk < K; kk++) mx = fmaxf(mx, sS[p0 + kk < 16 ? p0 + kk : 15]);
      float den = __expf(sn - mx);
      for (int kk = 0; kk < K; kk++) den += __expf(sS[p0 + kk < 16 ? p0 + kk : 15] - mx);
      const float inv = 1.0f / den;
      const float wr_ = __expf(sS[r] - mx) * inv;
#pragma unroll 1
      for (int q = 0; q < D / 16; q++) {  // value units, weighted
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const float *wr = sWkv + (16 * (D / 16 + q) + r) * PW + g;
#pragma unroll
        for (int t = 0; t < D / 4; t++) mfma(acc, wr[4 * t], x[t / 4][t % 4]);
        const int o = 16 * q + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; v++) sV[r * PW + o + v] = wr_ * (acc[v] + sbv[o + v]);
      }
      if (g == 0 && k == 0) sS[16 + r / K] = __expf(sn - mx) * inv;
      wave_lds_sync();
      for (int idx = lane; idx < PPT * D; idx += 64) {
        const int pp = idx / D, o = idx % D, plo = ta * PPT + pp;
        if (plo >= 16) continue;
        float s = sS[16 + pp] * snv[o];
        for (int kk = 0; kk < K; kk++) s += sV[(pp * K + kk) * PW + o];
        sPool[plo * PW + o] = s;
      }
      wave_lds_sync();
    }
    // phase B
    float o0 = 0.f, o1 = 0.f;
#pragma unroll 1
    for (int ob = 0; ob < H / 16; ob++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const float *wr = sW1 + (16 * ob + r) * PW + g;
      const float *pr = sPool + r * PW + g;
#pragma unroll
      for (int t = 0; t < D / 4; t++) mfma(acc, wr[4 * t], pr[4 * t]);
      const int o = 16 * ob + 4 * g;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float z = gelu(acc[v] + sb1[o + v]);
        o0 += sw20[o + v] * z;
        o1 += sw21[o + v] * z;
      }
    }
    o0 += __shfl_xor(o0, 16);
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 16);
    o1 += __shfl_xor(o1, 32);
    const int64_t m = m0 + r;
    if (g == 0 && m < h.M) {
      o0 += h.b2[0];
      o1 += h.b2[1];
      const float mu = 1.0f / (1.0f + expf(-o0));
      const float lb = h.log_eb + fminf(fmaxf(o1, -4.0f), 12.0f);
      float p = mu * h.span + h.vmin;
      p = fminf(fmaxf(p, h.vmin), h.vmax);
      pred[m] = p;
      scl[m] = exp2f(lb) * h.span;
      if (mu_out) mu_out[m] = mu;
      if (lb_out) lb_out[m] = lb;
    }
    wave_lds_sync();
  }
}

template <int D>
static size_t head_split_smem() {
  constexpr int H = 2 * D, PB = D + 8, PW = D + 4;
  return sizeof(uint16_t) * 2 * ((size_t)2 * D * PB + (size_t)H * PB) +
         sizeof(float) * (4 * (size_t)D + 3 * (size_t)H + HEAD_WAVES * (2 * 16 * (size_t)PW + 32));
}

// head2_kernel for the bf16 mode, both GEMMs on v_mfma_f32_16x16x32_bf16 with
// split operands: the axis pool's (2D x D) product takes the bf16 context rows
// as they are against W = W_hi + W_lo (two products), PredHead's first layer
// takes p = p_hi + p_lo against W1 = W1_hi + W1_lo (three products: the
// lo x lo term is below 2^-16 of the result).  The products keep about 16
// bits of each operand; the GELUs are gelu_fast.  Same phases, softmax and
// pooling order as head2_kernel.
template <int D>
__global__ __launch_bounds__(64 * HEAD_WAVES) void head_split_kernel(
    const uint16_t *__restrict__ ctx, HeadArgs h, float *__restrict__ pred,
    float *__restrict__ scl, float *__restrict__ mu_out, float *__restrict__ lb_out) {
  constexpr int H = 2 * D, PB = D + 8, PW = D + 4, NS = D / 32;
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16 *sKh = (__bf16 *)smem, *sKl = sKh + 2 * D * PB;
  __bf16 *s1h = sKl + 2 * D * PB, *s1l = s1h + H * PB;
  float *sbk1 = (float *)(s1l + H * PB);
  float *swk2 = sbk1 + D, *sbv = swk2 + D, *snv = sbv + D;
  float *sb1 = snv + D, *sw20 = sb1 + H, *sw21 = sw20 + H;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  float *sV = sw21 + H + wv * (2 * 16 * PW + 32);
  float *sPool = sV + 16 * PW;
  float *sS = sPool + 16 * PW;  // 16 scores, 16 null weights
  const __bf16 *wkh = (const __bf16 *)h.wkv_hi, *wkl = (const __bf16 *)h.wkv_lo;
  const __bf16 *w1h = (const __bf16 *)h.w1_hi, *w1l = (const __bf16 *)h.w1_lo;
  for (int i = tid; i < 2 * D * D; i += 64 * HEAD_WAVES) {
    sKh[(i / D) * PB + i % D] = wkh[i];
    sKl[(i / D) * PB + i % D] = wkl[i];
  }
  for (int i = tid; i < H * D; i += 64 * HEAD_WAVES) {
    s1h[(i / D) * PB + i % D] = w1h[i];
    s1l[(i / D) * PB + i % D] = w1l[i];
  }
  for (int i = tid; i < D; i += 64 * HEAD_WAVES) {
    sbk1[i] = h.bk1[i];
    swk2[i] = h.wk2[i];
    sbv[i] = h.bv[i];
    snv[i] = h.null_v[i];
  }
  for (int i = tid; i < H; i += 64 * HEAD_WAVES) {
    sb1[i] = h.b1[i] + h.cond[i] * h.log_eb;
    sw20[i] = h.w2[i];
    sw21[i] = h.w2[H + i];
  }
  __syncthreads();
  const int K = h.K, PPT = 16 / K, NTA = (16 + PPT - 1) / PPT;
  const float sn = h.null_k * h.scale;
  const int64_t gstep = (int64_t)gridDim.x * HEAD_WAVES * 16;
  for (int64_t m0 = ((int64_t)xcd_block() * HEAD_WAVES + wv) * 16; m0 < h.M; m0 += gstep) {
    for (int ta = 0; ta < NTA; ta++) {
      const int pl = ta * PPT + r / K, k = r % K;  // this lane's row: local point, axis
      const bool rv = r < PPT * K && pl < 16 && m0 + pl < h.M;
      const int64_t row = rv ? (m0 + pl) * K + k : 0;
      bf16x8 x[NS];
#pragma unroll
      for (int q = 0; q < NS; q++) x[q] = *(const bf16x8 *)(ctx + row * D + 32 * q + 8 * g);
      float part = 0.f;
#pragma unroll 1
      for (int ob = 0; ob < D / 16; ob++) {  // key hidden units
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const __bf16 *wr = sKh + (16 * ob + r) * PB + 8 * g;
#pragma unroll
        for (int q = 0; q < NS; q++) {
          mfma(acc, *(const bf16x8 *)(wr + 32 * q), x[q]);
          mfma(acc, *(const bf16x8 *)(wr + 2 * D * PB + 32 * q), x[q]);
        }
        const int o = 16 * ob + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; v++) part += swk2[o + v] * gelu_fast(acc[v] + sbk1[o + v]);
      }
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      if (g == 0) sS[r] = (part + h.bk2) * h.scale;
      wave_lds_sync();
      // softmax over the point's axes + null
      float mx = sn;
      const int p0 = (r / K) * K;
      for (int kk = 0; kk < K; kk++) mx = fmaxf(mx, sS[p0 + kk < 16 ? p0 + kk : 15]);
      float den = __expf(sn - mx);
      for (int kk = 0; kk < K; kk++) den += __expf(sS[p0 + kk < 16 ? p0 + kk : 15] - mx);
      const float inv = 1.0f / den;
      const float wr_ = __expf(sS[r] - mx) * inv;
#pragma unroll 1
      for (int q = 0; q < D / 16; q++) {  // value units, weighted
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const __bf16 *wr = sKh + (16 * (D / 16 + q) + r) * PB + 8 * g;
#pragma unroll
        for (int t = 0; t < NS; t++) {
          mfma(acc, *(const bf16x8 *)(wr + 32 * t), x[t]);
          mfma(acc, *(const bf16x8 *)(wr + 2 * D * PB + 32 * t), x[t]);
        }
        const int o = 16 * q + 4 * g;
#pragma unroll
        for (int v = 0; v < 4; v++) sV[r * PW + o + v] = wr_ * (acc[v] + sbv[o + v]);
      }
      if (g == 0 && k == 0) sS[16 + r / K] = __expf(sn - mx) * inv;
      wave_lds_sync();
      for (int idx = lane; idx < PPT * D; idx += 64) {
        const int pp = idx / D, o = idx % D, plo = ta * PPT + pp;
        if (plo >= 16) continue;
        float sacc = sS[16 + pp] * snv[o];
        for (int kk = 0; kk < K; kk++) sacc += sV[(pp * K + kk) * PW + o];
        sPool[plo * PW + o] = sacc;
      }
      wave_lds_sync();
    }
    // phase B: p of pooled point r, units 32 q + 8 g + i, split once
    bf16x8 ph[NS], plo[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
      float pv[8];
      const float *pr = sPool + r * PW + 32 * q + 8 * g;
#pragma unroll
      for (int i = 0; i < 8; i++) pv[i] = pr[i];
      split_bf16(pv, ph[q], plo[q]);
    }
    float o0 = 0.f, o1 = 0.f;
#pragma unroll 1
    for (int ob = 0; ob < H / 16; ob++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const __bf16 *wr = s1h + (16 * ob + r) * PB + 8 * g;
#pragma unroll
      for (int q = 0; q < NS; q++) {
        const bf16x8 a_hi = *(const bf16x8 *)(wr + 32 * q);
        const bf16x8 a_lo = *(const bf16x8 *)(wr + H * PB + 32 * q);
        mfma(acc, a_hi, ph[q]);
        mfma(acc, a_lo, ph[q]);
        mfma(acc, a_hi, plo[q]);
      }
      const int o = 16 * ob + 4 * g;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const float z = gelu_fast(acc[v] + sb1[o + v]);
        o0 += sw20[o + v] * z;
        o1 += sw21[o + v] * z;
      }
    }
    o0 += __shfl_xor(o0, 16);
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 16);
    o1 += __shfl_xor(o1, 32);
    const int64_t m = m0 + r;
    if (g == 0 && m < h.M) {
      o0 += h.b2[0];
      o1 += h.b2[1];
      const float mu = 1.0f / (1.0f + expf(-o0));
      const float lb = h.log_eb + fminf(fmaxf(o1, -4.0f), 12.0f);
      float p = mu * h.span + h.vmin;
      p = fminf(fmaxf(p, h.vmin), h.vmax);
      pred[m] = p;
      scl[m] = exp2f(lb) * h.span;
      if (mu_out) mu_out[m] = mu;
      if (lb_out) lb_out[m] = lb;
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// InitEmbed of the clamped, normalized recon value on an MFMA tile
// ---------------------------------------------------------------------------

template <typename T>
struct InitW {
  const float *w0, *b0;  // (D) first layer (one input)
  const T *w2;           // (D, D) natural order
  const float *b2;       // (D)
};

template <int D, typename T>
static size_t init_smem() {
  return sizeof(T) * (size_t)D * (D + Mf<T>::PE) + sizeof(float) * 3 * (size_t)D;
}

// 16 points per wave: lane (r, g) evaluates gelu(w0 v + b0) for the units of
// its k-set (natural order), then out^T = W2 h^T; out rows go to out[m] (the
// value embedding) or, with rows != nullptr, to E[rows[m], k] for every axis
// k < K (halo cells).
template <int D, typename T>
__global__ __launch_bounds__(256) void init_embed_kernel(int64_t M, int K,
                                                         const float *__restrict__ recon,
                                                         const int64_t *__restrict__ gidx,
                                                         const int64_t *__restrict__ rows,
                                                         float vmin, float vmax, float span,
                                                         InitW<T> w, T *__restrict__ out) {
  typedef Mf<T> F;
  constexpr int KS = F::KS, KL = F::KL, P = D + F::PE;
  extern __shared__ __align__(16) unsigned char smem[];
  T *sW = (T *)smem;
  float *sw0 = (float *)(sW + D * P), *sb0 = sw0 + D, *sb2 = sb0 + D;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  for (int i = tid; i < D * D; i += 256) sW[(i / D) * P + i % D] = w.w2[i];
  for (int i = tid; i < D; i += 256) {
    sw0[i] = w.w0[i];
    sb0[i] = w.b0[i];
    sb2[i] = w.b2[i];
  }
  __syncthreads();
  const int64_t gstep = (int64_t)gridDim.x * 4 * 16;
  for (int64_t m0 = ((int64_t)blockIdx.x * 4 + wv) * 16; m0 < M; m0 += gstep) {
    const int64_t m = m0 + r;
    float val = 0.f;
    if (m < M) {
      const float x = recon[gidx[m]];
      val = (fminf(fmaxf(x, vmin), vmax) - vmin) / span;
    }
    typename F::frag hb[D / KS];
#pragma unroll
    for (int s = 0; s < D / KS; s++) {
      if constexpr (sizeof(T) == 2) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int u = s * KS + 8 * g + i;
          hb[s][i] = (__bf16)gelu(__builtin_fmaf(sw0[u], val, sb0[u]));
        }
      } else {
        const int u = s * KS + g;
        hb[s] = gelu(__builtin_fmaf(sw0[u], val, sb0[u]));
      }
    }
#pragma unroll
    for (int ob = 0; ob < D / 16; ob++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const T *wr = sW + (16 * ob + r) * P + g * KL;
#pragma unroll
      for (int s = 0; s < D / KS; s++) mfma(acc, ldfrag<T>(wr + s * KS), hb[s]);
      const int o = 16 * ob + 4 * g;
      const f32x4 b2 = *(const f32x4 *)(sb2 + o);
      const typename F::quad y = to_quad<T>(acc + b2);
      if (m < M) {
        if (rows) {
          for (int k = 0; k < K; k++) *(typename F::quad *)(out + (rows[m] * K + k) * D + o) = y;
        } else {
          *(typename F::quad *)(out + m * D + o) = y;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// gate: rate proxy per candidate and bucket (gnn_codec.py:777-971)
// ---------------------------------------------------------------------------

struct GateArgs {
  int64_t n;
  double eb, two_eb;
  uint32_t radius, n_cand;     // candidates = 1 + 4 * n_interp
  const float *x, *pred, *scale;
  const float *ip[3];
  const float *thr;            // 63 f32 thresholds of the 64-level scale grid
  const double *lvk;           // (3, 64) per grid level: bits0, eb / (b ln 2), bits2
  double bound[12];            // eb * 2^T, T = 2..13
  double *partial;             // (blocks, n_cand, 13)
};

// The reference's rate proxy (gnn_codec.py:777-797) at a snapped scale b:
// with k = rint(|r| / 2eb),
//   k = 0: bits = -log2(-expm1(-eb / b))
//   k > 0: bits = -log2(0.5 exp(-(2k - 1) eb / b) (-expm1(-2 eb / b)))
//               = 1 + (2k - 1) eb / (b ln 2) - log2(-expm1(-2 eb / b)),
// clamped to 24 bits, 56 past the radius.  b takes one of 64 grid values, so
// the two logarithms and eb / (b ln 2) are per-level constants (lvk, built on
// the host in float64); the rewrite moves each point's bits by about an ulp
// (the sums stay within 1e-12 of the reference's).
__device__ __forceinline__ double gate_bits(double absr, int lev, const GateArgs &g,
                                            const double *lvk) {
  const double k = rint(absr / g.two_eb);
  double bits = k == 0.0 ? lvk[lev] : 1.0 + (2.0 * k - 1.0) * lvk[64 + lev] + lvk[128 + lev];
  if (bits > 24.0) bits = 24.0;
  if (k >= (double)g.radius) bits = 56.0;
  return bits;
}

// bucketize(right=True) of b against the 63 thresholds (in LDS): the grid level
__device__ __forceinline__ int snap_level(float b, const float *thr) {
  int lo = 0, hi = 63;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (thr[mid] <= b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

constexpr int GATE_BLOCK = 256;
constexpr int GATE_NB = 13;
constexpr int GATE_NC = 13;  // 1 + 4 * 3 candidates at most

// partial[b][c][bucket]: sums over the points of block b.  Each thread sums
// its points in order, a butterfly of lane shuffles sums each wave, and the
// four wave sums are added in wave order: a fixed order for a given n.
__global__ __launch_bounds__(GATE_BLOCK) void gate_cost_kernel(GateArgs g) {
  __shared__ double wsum[GATE_BLOCK / 64][GATE_NC][GATE_NB];
  __shared__ double lvk[3 * 64];
  __shared__ float thr[64];
  for (int i = threadIdx.x; i < 3 * 64; i += GATE_BLOCK) lvk[i] = g.lvk[i];
  if (threadIdx.x < 63) thr[threadIdx.x] = g.thr[threadIdx.x];
  __syncthreads();
  const int64_t per_block = (g.n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < g.n ? lo + per_block : g.n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t c = 0; c < g.n_cand; c++) {
    double acc[GATE_NB];
#pragma unroll
    for (int b = 0; b < GATE_NB; b++) acc[b] = 0.0;
    const int kind = c ? (int)(c - 1) / 4 : 0, sh = c ? (int)((c - 1) % 4) * 2 : 0;
    const float *ref = c ? g.ip[kind] : g.pred;
    for (int64_t i = lo + threadIdx.x; i < hi; i += GATE_BLOCK) {
      const float sc = g.scale[i];
      const double b = (double)sc;
      int bucket = 0;
      while (bucket < 12 && g.bound[bucket] < b) bucket++;
      const float bs = c ? sc * exp2f(-(float)sh) : sc;
      const double bits =
          gate_bits(fabs((double)g.x[i] - (double)ref[i]), snap_level(bs, thr), g, lvk);
#pragma unroll
      for (int bb = 0; bb < GATE_NB; bb++)
        if (bb == bucket) acc[bb] += bits;
    }
#pragma unroll
    for (int b = 0; b < GATE_NB; b++) {
      double v = acc[b];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) wsum[wv][c][b] = v;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < (int)g.n_cand * GATE_NB; t += GATE_BLOCK) {
    const int c = t / GATE_NB, b = t % GATE_NB;
    double v = wsum[0][c][b];
    for (int w = 1; w < GATE_BLOCK / 64; w++) v += wsum[w][c][b];
    g.partial[((int64_t)blockIdx.x * g.n_cand + c) * GATE_NB + b] = v;
  }
}

// out[i] = sum over blocks b, in block order, of partial[b][i]
__global__ void gate_sum_kernel(const double *__restrict__ partial, int blocks, int width,
                                double *__restrict__ out) {
  for (int i = threadIdx.x; i < width; i += blockDim.x) {
    double v = 0.0;
    for (int b = 0; b < blocks; b++) v += partial[(int64_t)b * width + i];
    out[i] = v;
  }
}

// p = m ? ip : pred ; s = m ? scale * 2^-sh : scale, m by the side of eb*2^T
__global__ void gate_apply_kernel(int64_t n, float *__restrict__ pred,
                                  float *__restrict__ scale, const float *__restrict__ ip,
                                  double thr, float mult, int dir) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float s = scale[i];
  const bool below = (double)s < thr;
  const bool m = dir ? !below : below;
  if (m) {
    pred[i] = ip[i];
    scale[i] = s * mult;
  }
}

__global__ void gather_kernel(int64_t n, const float *__restrict__ src,
                              const int64_t *__restrict__ idx, float *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void scatter_kernel(int64_t n, const float *__restrict__ src,
                               const int64_t *__restrict__ idx, float *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}

__global__ void gather_i64_kernel(int64_t n, const int64_t *__restrict__ src,
                                  const int64_t *__restrict__ idx, int64_t *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// rec[rank[j]] = val[j] for j < *count: the quantizer's outliers (listed in
// any order) over a stage reconstruction
__global__ void outlier_fix_kernel(const uint32_t *__restrict__ rank,
                                   const float *__restrict__ val,
                                   const uint32_t *__restrict__ count, int64_t cap,
                                   float *__restrict__ rec) {
  int64_t k = *count;
  if (k > cap) k = cap;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < k;
       j += (int64_t)gridDim.x * blockDim.x)
    rec[rank[j]] = val[j];
}

__global__ void fill_f32_kernel(int64_t n, float v, float *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void zero_kernel(int64_t n, T *p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

template <typename T>
static void *alloc_part(std::vector<unsigned char> &host, size_t &off, const std::vector<T> &v) {
  off = (off + 255) & ~(size_t)255;
  const size_t at = off;
  host.resize(at + v.size() * sizeof(T));
  memcpy(host.data() + at, v.data(), v.size() * sizeof(T));
  off = at + v.size() * sizeof(T);
  return (void *)at;
}

}  // namespace gnn
}  // namespace latec

using namespace latec::gnn;
using latec::ClassDesc;

namespace {

// Packs the model for the device.  Message / mix MLPs: W1 with its input in
// the kernel's register order and the additive-term columns appended (b1,
// the two scalar coefficients, the state bias rows; bf16 as hi + lo parts),
// W2 and the key's Wk with their inputs in register order, all in the
// storage type.  Pools and head in f32.  Everything in one device block.
int make_device(latec_gnn *g, const float *P) {
  const Layout &Ly = g->lay;
  const int d = Ly.d, h = Ly.h, de = Ly.de, ds = Ly.ds, ns = Ly.ns, agg = g->agg;
  const bool bf = g->numerics == NUM_BF16;
  const int nxc = bf ? 32 : 16;  // additive-term columns
  if (ns > (bf ? 8 : 13)) {
    latec::set_error("GNN kernels support at most " + std::to_string(bf ? 8 : 13) +
                     " direction states, got " + std::to_string(ns));
    return LATEC_E_ARG;
  }
  // a two-sided pair's states are mirror images: st1 = 2 agg - st0
  for (size_t i = 0; i + 1 < g->sidx.size(); i += 2)
    if (g->sidx[i + 1] != 2 * agg - g->sidx[i]) {
      latec::set_error("direction states of a line are not mirrored");
      return LATEC_E_ARG;
    }
  std::vector<unsigned char> host;
  size_t off = 0;
  auto mat = [&](const std::vector<float> &v) -> void * {
    if (bf) {
      std::vector<uint16_t> b(v.size());
      for (size_t i = 0; i < v.size(); i++) b[i] = to_bf16(v[i]);
      return alloc_part(host, off, b);
    }
    return alloc_part(host, off, v);
  };
  auto rows = [&](int64_t base, int n_out, int n_in, int col0, int ncol) {
    std::vector<float> v((size_t)n_out * ncol);
    for (int o = 0; o < n_out; o++)
      for (int c = 0; c < ncol; c++) v[(size_t)o * ncol + c] = P[base + (int64_t)o * n_in + col0 + c];
    return v;
  };
  auto cat2 = [&](const std::vector<float> &a, int na, const std::vector<float> &b, int nb,
                  int n_out) {
    std::vector<float> v((size_t)n_out * (na + nb));
    for (int o = 0; o < n_out; o++) {
      for (int c = 0; c < na; c++) v[(size_t)o * (na + nb) + c] = a[(size_t)o * na + c];
      for (int c = 0; c < nb; c++) v[(size_t)o * (na + nb) + na + c] = b[(size_t)o * nb + c];
    }
    return v;
  };
  // input index -> register order; `gathered` inputs come from memory rows
  // (natural order in bf16, the f32 order in f32), the others from a
  // transposed layer's registers
  auto reg_in = [&](const std::vector<float> &W, int n_out, int n_in, bool gathered, bool bfm) {
    std::vector<float> v(W.size());
    for (int o = 0; o < n_out; o++)
      for (int kk = 0; kk < n_in; kk++)
        v[(size_t)o * n_in + kk] = W[(size_t)o * n_in + ((bfm && gathered) ? kk : reg_unit(bfm, kk))];
    return v;
  };
  auto vec = [&](int64_t base, int n) { return std::vector<float>(P + base, P + base + n); };
  auto col = [&](int64_t wbase, int n_out, int n_in, int c) {
    std::vector<float> v(n_out);
    for (int o = 0; o < n_out; o++) v[o] = P[wbase + (int64_t)o * n_in + c];
    return v;
  };
  // DirState rows folded through the first layer: (ns, h), f64
  auto state_rows = [&](int64_t wbase, int n_in, int col0) {
    std::vector<double> v((size_t)ns * h);
    for (int s = 0; s < ns; s++)
      for (int o = 0; o < h; o++) {
        double acc = 0.0;
        for (int j = 0; j < ds; j++)
          acc += (double)P[wbase + (int64_t)o * n_in + col0 + j] * (double)P[Ly.table + s * ds + j];
        v[(size_t)s * h + o] = acc;
      }
    return v;
  };
  auto hi = [](float x) { return from_bf16(to_bf16(x)); };
  // (n_out, n_in + ks): W then one bias k-step (bf16: high and low parts)
  const int ks = bf ? 32 : 4;
  auto with_bias = [&](const std::vector<float> &W, int n_out, int n_in, const std::vector<float> &b) {
    std::vector<float> v((size_t)n_out * (n_in + ks), 0.f);
    for (int o = 0; o < n_out; o++) {
      for (int c = 0; c < n_in; c++) v[(size_t)o * (n_in + ks) + c] = W[(size_t)o * n_in + c];
      v[(size_t)o * (n_in + ks) + n_in] = bf ? hi(b[o]) : b[o];
      if (bf) v[(size_t)o * (n_in + ks) + n_in + 1] = b[o] - hi(b[o]);
    }
    return v;
  };
  // (h, K1 + nxc): W1 (register order) then the additive-term columns
  auto w1x = [&](const std::vector<float> &W1, int K1, const std::vector<float> &b1,
                 const std::vector<float> *c0, const std::vector<float> *c1,
                 const std::vector<double> *S) {
    const int n = K1 + nxc;
    std::vector<float> v((size_t)h * n, 0.f);
    const std::vector<float> Wr = reg_in(W1, h, K1, true, bf);
    for (int o = 0; o < h; o++) {
      for (int c = 0; c < K1; c++) v[(size_t)o * n + c] = Wr[(size_t)o * K1 + c];
      float *e = &v[(size_t)o * n + K1];
      if (bf) {
        e[0] = hi(b1[o]);
        e[1] = b1[o] - hi(b1[o]);
        if (c0) {
          e[2] = hi((*c0)[o]);
          e[3] = (*c0)[o] - hi((*c0)[o]);
          e[4] = hi((*c0)[o]);
        }
        if (c1) {
          e[5] = hi((*c1)[o]);
          e[6] = (*c1)[o] - hi((*c1)[o]);
          e[7] = hi((*c1)[o]);
        }
        if (S)
          for (int st = 0; st < ns; st++) {
            const float sv = (float)(*S)[(size_t)st * h + o];
            e[8 + st] = hi(sv);
            e[16 + st] = sv - hi(sv);
          }
      } else {
        e[0] = b1[o];
        if (c0) e[1] = (*c0)[o];
        if (c1) e[2] = (*c1)[o];
        if (S)
          for (int st = 0; st < ns; st++) e[3 + st] = (float)(*S)[(size_t)st * h + o];
      }
    }
    return v;
  };
  // message layer 1 (bidir): [E_neg | E_pos] columns, scalars log d_neg, log d_pos
  const int bi_in = 2 * de + 2;
  const std::vector<float> w_bi =
      cat2(rows(Ly.bi0w, h, bi_in, 0, d), d, rows(Ly.bi0w, h, bi_in, de, d), d, h);
  const std::vector<double> s0 = state_rows(Ly.bi0w, bi_in, d), s1 = state_rows(Ly.bi0w, bi_in, de + d);
  std::vector<double> s_bi((size_t)ns * h);
  for (int st = 0; st < ns; st++)
    for (int o = 0; o < h; o++)
      s_bi[(size_t)st * h + o] = s0[(size_t)st * h + o] + s1[(size_t)(2 * agg - st) * h + o];
  const std::vector<float> bi_c0 = col(Ly.bi0w, h, bi_in, 2 * de), bi_c1 = col(Ly.bi0w, h, bi_in, 2 * de + 1);
  void *o_bi_w1 = mat(w1x(w_bi, 2 * d, vec(Ly.bi0b, h), &bi_c0, &bi_c1, &s_bi));
  void *o_bi_w2 = mat(with_bias(reg_in(rows(Ly.bi2w, d, h, 0, h), d, h, false, bf), d, h, vec(Ly.bi2b, d)));
  // one-sided: E_side, scalars sign and log d
  const int dir_in = de + 2;
  const std::vector<double> s_dir = state_rows(Ly.dir0w, dir_in, d);
  const std::vector<float> dir_c0 = col(Ly.dir0w, h, dir_in, de), dir_c1 = col(Ly.dir0w, h, dir_in, de + 1);
  void *o_dir_w1 = mat(w1x(rows(Ly.dir0w, h, dir_in, 0, d), d, vec(Ly.dir0b, h), &dir_c0, &dir_c1, &s_dir));
  void *o_dir_w2 = mat(with_bias(reg_in(rows(Ly.dir2w, d, h, 0, h), d, h, false, bf), d, h, vec(Ly.dir2b, d)));
  // mix: [ctx | value embedding]
  void *o_mx_w1 = mat(w1x(rows(Ly.mx0w, h, 2 * d, 0, 2 * d), 2 * d, vec(Ly.mx0b, h), nullptr, nullptr, nullptr));
  void *o_mx_w2 = mat(with_bias(reg_in(rows(Ly.mx2w, d, h, 0, h), d, h, false, bf), d, h, vec(Ly.mx2b, d)));
  // line-pool key layer 1 on the message registers
  void *o_key_w = mat(with_bias(reg_in(rows(Ly.lp[0], d, d, 0, d), d, d, false, bf), d, d, vec(Ly.lp[1], d)));
  // InitEmbed layer 2, natural order
  void *o_init_w2 = mat(rows(Ly.init2w, d, d, 0, d));
  // bf16 mode: axis-pool key + value layers (2d, d), PredHead's first layer
  // (h, d) without its log-eb column and the line pool's value projection
  // (d, d), high and low parts
  void *o_hd[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (bf) {
    std::vector<float> kv = rows(Ly.ap[0], d, d, 0, d), v2 = rows(Ly.ap[4], d, d, 0, d);
    kv.insert(kv.end(), v2.begin(), v2.end());
    const std::vector<float> w1 = rows(Ly.hd0w, h, d + 1, 0, d);
    auto lo = [&](const std::vector<float> &W) {
      std::vector<float> v(W.size());
      for (size_t i = 0; i < W.size(); i++) v[i] = W[i] - hi(W[i]);
      return v;
    };
    o_hd[0] = mat(kv);
    o_hd[1] = mat(lo(kv));
    o_hd[2] = mat(w1);
    o_hd[3] = mat(lo(w1));
    const std::vector<float> wv = rows(Ly.lp[4], d, d, 0, d);
    o_hd[4] = mat(wv);
    o_hd[5] = mat(lo(wv));
  }
  // f32 arrays
  std::vector<std::vector<float>> f(F_NSLOTS);
  f[F_BI_B2] = vec(Ly.bi2b, d);
  f[F_DIR_B2] = vec(Ly.dir2b, d);
  f[F_MX_B2] = vec(Ly.mx2b, d);
  f[F_LK_B1] = vec(Ly.lp[1], d);
  f[F_LK_W2] = vec(Ly.lp[2], d);
  f[F_LK_WV] = rows(Ly.lp[4], d, d, 0, d);
  f[F_LK_BV] = vec(Ly.lp[5], d);
  f[F_LK_NULLV] = vec(Ly.lp[7], d);
  {
    std::vector<float> kv = rows(Ly.ap[0], d, d, 0, d), v2 = rows(Ly.ap[4], d, d, 0, d);
    kv.insert(kv.end(), v2.begin(), v2.end());
    f[F_AP_WKV] = reg_in(kv, 2 * d, d, true, false);
  }
  f[F_AP_B1] = vec(Ly.ap[1], d);
  f[F_AP_W2] = vec(Ly.ap[2], d);
  f[F_AP_BV] = vec(Ly.ap[5], d);
  f[F_AP_NULLV] = vec(Ly.ap[7], d);
  f[F_HD_W1] = rows(Ly.hd0w, h, d + 1, 0, d);
  f[F_HD_COND] = col(Ly.hd0w, h, d + 1, d);
  f[F_HD_B1] = vec(Ly.hd0b, h);
  f[F_HD_W2] = vec(Ly.hd2w, 2 * h);
  f[F_HD_B2] = vec(Ly.hd2b, 2);
  f[F_IN_W0] = vec(Ly.init0w, d);
  f[F_IN_B0] = vec(Ly.init0b, d);
  f[F_IN_B2] = vec(Ly.init2b, d);
  f[F_SCAL] = {P[Ly.lp[3]], P[Ly.lp[6]], P[Ly.ap[3]], P[Ly.ap[6]]};
  for (int i = 0; i < 4; i++) g->scal[i] = f[F_SCAL][i];
  std::vector<float> flat;
  g->f32_off.assign(F_NSLOTS, 0);
  for (int s = 0; s < F_NSLOTS; s++) {
    while (flat.size() % 64) flat.push_back(0.f);
    g->f32_off[s] = (int64_t)flat.size();
    flat.insert(flat.end(), f[s].begin(), f[s].end());
  }
  void *o_f32 = alloc_part(host, off, flat);
  void *o_dirs = alloc_part(host, off, g->dirs);
  void *o_sidx = alloc_part(host, off, g->sidx);
  void *o_lnz = alloc_part(host, off, g->lognnz);
  void *dev = nullptr;
  int rc = hip_ok(hipMalloc(&dev, host.size()), "hipMalloc(gnn weights)");
  if (!rc) rc = hip_ok(hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice),
                       "hipMemcpy(gnn weights)");
  if (rc) {
    if (dev) (void)hipFree(dev);
    return rc;
  }
  auto at = [&](void *o) { return (void *)((unsigned char *)dev + (size_t)o); };
  g->dev = dev;
  g->dev_bytes = host.size();
  g->bi_w1 = at(o_bi_w1);
  g->bi_w2 = at(o_bi_w2);
  g->dir_w1 = at(o_dir_w1);
  g->dir_w2 = at(o_dir_w2);
  g->mx_w1 = at(o_mx_w1);
  g->mx_w2 = at(o_mx_w2);
  g->key_w = at(o_key_w);
  g->init_w2 = at(o_init_w2);
  if (bf) {
    g->hd_wkv_hi = at(o_hd[0]);
    g->hd_wkv_lo = at(o_hd[1]);
    g->hd_w1_hi = at(o_hd[2]);
    g->hd_w1_lo = at(o_hd[3]);
    g->lk_wv_hi = at(o_hd[4]);
    g->lk_wv_lo = at(o_hd[5]);
  }
  g->f32 = (float *)at(o_f32);
  g->d_dirs = (int8_t *)at(o_dirs);
  g->d_sidx = (int8_t *)at(o_sidx);
  g->d_lognnz = (float *)at(o_lnz);
  return LATEC_OK;
}

const float *fs(const latec_gnn *g, int slot) { return g->f32 + g->f32_off[slot]; }

int ensure_work(latec_gnn *g, size_t bytes) {
  if (g->work_bytes >= bytes) return LATEC_OK;
  if (g->work) {
    // earlier launches on any stream may still read the old block
    int rc = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    (void)hipFree(g->work);
  }
  g->work = nullptr;
  g->work_bytes = 0;
  int rc = hip_ok(hipMalloc(&g->work, bytes), "hipMalloc(gnn workspace)");
  if (!rc) g->work_bytes = bytes;
  return rc;
}

// dynamic LDS attribute + resident workgroups per CU (1..8)
int lds_setup(const void *fn, int threads, size_t smem, int *per_cu) {
  int rc = hip_ok(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem),
                  "hipFuncSetAttribute");
  if (rc) return rc;
  int n = 0;
  rc = hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, smem),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  *per_cu = n < 1 ? 1 : n > 8 ? 8 : n;
  return rc;
}

// launch shape of the message / mix rows (knob gnn_rows): 16-row tiles per
// wave group and waves per workgroup
int g_gnn_rows = 3;
constexpr int GNN_ROW_SHAPES = 8;
// line pool kernel (knob gnn_pool): 0 line by line; 1 and 2 loads issued
// ahead, message loads of 8 (1) or 4 (2) lines at a time (items of at most
// POOL_LMAX lines); all give the same bits
int g_gnn_pool = 2;
// geometry through per-stage lookup tables (knob gnn_geo_table; same steps)
int g_gnn_geo_table = 1;

template <int D, typename T, int MODE, int NT, int WV, int MINW = 1>
int launch_rows_s(latec_gnn *g, RowArgs &a, const RowW<T> &w, hipStream_t s) {
  const size_t smem = RowShape<D, T, MODE>::smem(a.L, a.K);
  const void *fn = (const void *)msg_kernel<D, T, MODE, NT, WV, MINW>;
  int per_cu = 1;
  int rc = lds_setup(fn, 64 * WV, smem, &per_cu);
  if (rc) return rc;
  hipLaunchKernelGGL((msg_kernel<D, T, MODE, NT, WV, MINW>), dim3(g->n_cu * per_cu),
                     dim3(64 * WV), smem, s, a, w);
  return hip_ok(hipGetLastError(), "msg_kernel launch");
}

template <int D, typename T, int MODE>
int launch_rows_m(latec_gnn *g, RowArgs &a, const RowW<T> &w, hipStream_t s) {
  if (sizeof(T) == 4) return launch_rows_s<D, T, MODE, 1, 4>(g, a, w, s);
  switch (g_gnn_rows) {
    case 1: return launch_rows_s<D, T, MODE, 1, 4>(g, a, w, s);
    case 2: return launch_rows_s<D, T, MODE, 1, 8>(g, a, w, s);
    case 3: return launch_rows_s<D, T, MODE, 2, 8>(g, a, w, s);
    case 4: return launch_rows_s<D, T, MODE, 2, 6>(g, a, w, s);
    case 5: return launch_rows_s<D, T, MODE, 1, 6>(g, a, w, s);
    case 6: return launch_rows_s<D, T, MODE, 2, 8, 4>(g, a, w, s);
    case 7: return launch_rows_s<D, T, MODE, 1, 8, 4>(g, a, w, s);
    default: return launch_rows_s<D, T, MODE, 2, 4>(g, a, w, s);
  }
}

template <int D, typename T>
int launch_rows(latec_gnn *g, int mode, RowArgs &a, hipStream_t s) {
  RowW<T> w{};
  w.wk = (const T *)g->key_w;
  w.wk2 = fs(g, F_LK_W2);
  w.bk2 = g->scal[0];
  a.L = g->L;
  a.sidx = g->d_sidx;
  a.kmagic = (65536u + (uint32_t)a.K - 1) / (uint32_t)a.K;
  if (mode == ROW_BI) {
    w.w1 = (const T *)g->bi_w1;
    w.w2 = (const T *)g->bi_w2;
    return launch_rows_m<D, T, ROW_BI>(g, a, w, s);
  }
  if (mode == ROW_DIR) {
    w.w1 = (const T *)g->dir_w1;
    w.w2 = (const T *)g->dir_w2;
    return launch_rows_m<D, T, ROW_DIR>(g, a, w, s);
  }
  w.w1 = (const T *)g->mx_w1;
  w.w2 = (const T *)g->mx_w2;
  return launch_rows_m<D, T, ROW_FIN>(g, a, w, s);
}

}  // namespace

namespace latec {
int gnn_tune_set(const char *name, int64_t v) {
  if (!strcmp(name, "gnn_rows")) {
    if (v < 0 || v >= GNN_ROW_SHAPES) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_gnn_rows = (int)v;
    return LATEC_OK;
  }
  if (!strcmp(name, "gnn_geo_table")) {
    if (v < 0 || v > 1) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_gnn_geo_table = (int)v;
    return LATEC_OK;
  }
  if (!strcmp(name, "gnn_pool")) {
    if (v < 0 || v > 2) {
      set_error(std::string("tuning value out of range for ") + name);
      return LATEC_E_ARG;
    }
    g_gnn_pool = (int)v;
    return LATEC_OK;
  }
  set_error(std::string("unknown tuning knob ") + name);
  return LATEC_E_ARG;
}
int64_t gnn_tune_get(const char *name) {
  if (!strcmp(name, "gnn_rows")) return g_gnn_rows;
  if (!strcmp(name, "gnn_pool")) return g_gnn_pool;
  if (!strcmp(name, "gnn_geo_table")) return g_gnn_geo_table;
  set_error(std::string("unknown tuning knob ") + name);
  return LATEC_E_ARG;
}
}  // namespace latec

namespace {

bool d_supported(int d) { return d == 32 || d == 64; }

// (L, ndim) half-directions in itertools.product order (gnn_predictor.py:117-136),
// their 0.5*log2(nnz) and the DirState index per (line, axis, side)
// (gnn_predictor.py:657-666).
void build_lines(latec_gnn *g) {
  const int nd = g->ndim, agg = g->agg;
  int64_t total = 1;
  for (int j = 0; j < nd; j++) total *= 3;
  g->dirs.clear();
  g->lognnz.clear();
  g->sidx.clear();
  for (int64_t code = 0; code < total; code++) {
    int v[LATEC_MAX_DIM];
    int64_t c = code;
    for (int j = nd - 1; j >= 0; j--) {
      v[j] = (int)(c % 3) - 1;
      c /= 3;
    }
    int first = 0, nnz = 0;
    for (int j = 0; j < nd; j++)
      if (v[j]) {
        if (!first) first = v[j];
        nnz++;
      }
    if (first <= 0 || nnz > agg) continue;
    for (int j = 0; j < nd; j++) g->dirs.push_back((int8_t)v[j]);
    g->lognnz.push_back((float)(0.5 * std::log2((double)nnz)));
    for (int k = 0; k < nd; k++)
      for (int side = 0; side < 2; side++) {
        const int sg = side ? 1 : -1;
        const int sv = sg * v[k];  // sign of the cosine; |cos| = 1/sqrt(nnz)
        int idx = sv > 0 ? 2 * agg - nnz + 1 : sv < 0 ? nnz - 1 : agg;
        idx = idx < 0 ? 0 : idx > 2 * agg ? 2 * agg : idx;
        g->sidx.push_back((int8_t)idx);
      }
  }
  g->L = (int)g->lognnz.size();
}

int line_offsets(latec_gnn *g, const int64_t *frame_shape, hipStream_t s) {
  bool same = g->d_line_off != nullptr;
  for (int j = 0; j < g->ndim; j++) same &= g->line_off_key[j] == frame_shape[j];
  if (same) return LATEC_OK;
  std::vector<int64_t> fstr(g->ndim), off(g->L);
  int64_t acc = 1;
  for (int j = g->ndim - 1; j >= 0; j--) fstr[j] = acc, acc *= frame_shape[j];
  for (int l = 0; l < g->L; l++) {
    int64_t o = 0;
    for (int j = 0; j < g->ndim; j++) o += (int64_t)g->dirs[l * g->ndim + j] * fstr[j];
    off[l] = o;
  }
  int rc = LATEC_OK;
  if (!g->d_line_off)
    rc = hip_ok(hipMalloc(&g->d_line_off, sizeof(int64_t) * LATEC_MAX_DIM * 64), "hipMalloc");
  if (!rc) rc = hip_ok(hipMemcpyAsync(g->d_line_off, off.data(), sizeof(int64_t) * g->L,
                                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (!rc)
    for (int j = 0; j < g->ndim; j++) g->line_off_key[j] = frame_shape[j];
  return rc;
}

template <int D, typename T>
int embed_t(latec_gnn *g, const void *E, const int64_t *frame_shape, const int32_t *steps,
            const int64_t *qidx, int64_t M, void *ctx, hipStream_t s) {
  int rc = line_offsets(g, frame_shape, s);
  if (rc) return rc;
  const int L = g->L, K = g->ndim;
  const int64_t per_q = (int64_t)L * K * D * (int64_t)sizeof(T);
  int64_t MT = ((int64_t)1 << 31) / per_q;
  MT = MT < 1024 ? 1024 : MT > ((int64_t)1 << 20) ? ((int64_t)1 << 20) : MT;
  if (MT > M) MT = M < 1 ? 1 : M;
  const int64_t np = MT * L;
  const int64_t nblk = (np + PAIR_BLOCK - 1) / PAIR_BLOCK;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_counts = al(sizeof(int32_t) * 2 * nblk), b_tot = al(16);
  const size_t b_slot = al(sizeof(int32_t) * np);
  const size_t b_rec = al(sizeof(PairRec) * np), b_lw = al(sizeof(int16_t) * np);
  const size_t b_msg = al((size_t)np * K * D * sizeof(T)), b_score = al(sizeof(float) * np * K);
  rc = ensure_work(g, b_counts + b_tot + b_slot + 2 * (b_rec + b_lw) + b_msg + b_score);
  if (rc) return rc;
  unsigned char *w = (unsigned char *)g->work;
  int32_t *counts = (int32_t *)w;
  int32_t *totals = (int32_t *)(w + b_counts);
  int32_t *slot = (int32_t *)(w + b_counts + b_tot);
  PairRec *rec_bi = (PairRec *)((unsigned char *)slot + b_slot);
  PairRec *rec_dir = (PairRec *)((unsigned char *)rec_bi + b_rec);
  int16_t *lw_bi = (int16_t *)((unsigned char *)rec_dir + b_rec);
  int16_t *lw_dir = (int16_t *)((unsigned char *)lw_bi + b_lw);
  T *msg = (T *)((unsigned char *)lw_dir + b_lw);
  float *score = (float *)((unsigned char *)msg + b_msg);
  const float scale = 1.0f / std::sqrt((float)D);
  const PoolArgs pa{fs(g, F_LK_WV), fs(g, F_LK_BV), fs(g, F_LK_NULLV), g->scal[1] * scale,
                    (const uint16_t *)g->lk_wv_hi, (const uint16_t *)g->lk_wv_lo};
  const size_t psm = PoolLds<D, T>::bytes();
  const int pool_lines = L <= POOL_LMAX ? g_gnn_pool : 0;
  const void *pool_fn = pool_lines == 1   ? (const void *)pool_lines_kernel<D, T, 8>
                        : pool_lines == 2 ? (const void *)pool_lines_kernel<D, T, 4>
                                          : (const void *)pool_kernel<D, T>;
  int pool_per_cu = 1;
  rc = lds_setup(pool_fn, 64 * POOL_WAVES, psm, &pool_per_cu);
  if (rc) return rc;
  for (int64_t m0 = 0; m0 < M; m0 += MT) {
    const int64_t mt = M - m0 < MT ? M - m0 : MT;
    const int64_t n_pairs = mt * L;
    const int64_t nb = (n_pairs + PAIR_BLOCK - 1) / PAIR_BLOCK;
    const int32_t *st_t = steps + m0 * L;
    hipLaunchKernelGGL(pair_count_kernel, dim3((unsigned)nb), dim3(PAIR_BLOCK), 0, s, st_t,
                       n_pairs, counts);
    hipLaunchKernelGGL(pair_scan_kernel, dim3(1), dim3(1024), 0, s, counts, nb, totals);
    hipLaunchKernelGGL(pair_write_kernel, dim3((unsigned)nb), dim3(PAIR_BLOCK), 0, s, st_t,
                       n_pairs, L, qidx + m0, (const int64_t *)g->d_line_off,
                       (const float *)g->d_lognnz, counts, totals, rec_bi, lw_bi, rec_dir,
                       lw_dir, slot);
    rc = hip_ok(hipGetLastError(), "pair kernels");
    if (rc) return rc;
    RowArgs a{};
    a.K = K;
    a.E = E;
    a.out = msg;
    a.score = score;
    a.score_scale = scale;
    a.rec = rec_bi;
    a.lw = lw_bi;
    a.n_items_dev = totals;
    a.slot_base = 0;
    rc = launch_rows<D, T>(g, ROW_BI, a, s);
    if (rc) return rc;
    a.rec = rec_dir;
    a.lw = lw_dir;
    a.n_items_dev = totals + 1;
    a.slot_base_dev = totals;
    rc = launch_rows<D, T>(g, ROW_DIR, a, s);
    if (rc) return rc;
    const int64_t items = mt * K;
    const int64_t waves = (items + 15) / 16;
    int64_t grid = (waves + POOL_WAVES - 1) / POOL_WAVES;
    if (grid > (int64_t)g->n_cu * pool_per_cu * 4) grid = (int64_t)g->n_cu * pool_per_cu * 4;
    if (pool_lines == 1)
      hipLaunchKernelGGL((pool_lines_kernel<D, T, 8>), dim3((unsigned)grid),
                         dim3(64 * POOL_WAVES), psm, s, items, K, L, slot, msg, score, pa,
                         (T *)ctx, m0);
    else if (pool_lines == 2)
      hipLaunchKernelGGL((pool_lines_kernel<D, T, 4>), dim3((unsigned)grid),
                         dim3(64 * POOL_WAVES), psm, s, items, K, L, slot, msg, score, pa,
                         (T *)ctx, m0);
    else
      hipLaunchKernelGGL((pool_kernel<D, T>), dim3((unsigned)grid), dim3(64 * POOL_WAVES), psm,
                         s, items, K, L, slot, msg, score, pa, (T *)ctx, m0);
    rc = hip_ok(hipGetLastError(), "pool_kernel launch");
    if (rc) return rc;
  }
  return LATEC_OK;
}

template <int D, typename T>
int head_t(latec_gnn *g, const void *ctx, int64_t M, double eb, float vmin, float vmax,
           float *pred, float *scale, float *mu, float *logb, hipStream_t s) {
  HeadArgs h{};
  h.M = M;
  h.K = g->ndim;
  h.scale = 1.0f / std::sqrt((float)D);
  float ebf = (float)eb;
  if (!(ebf > 1.17549435e-38f)) ebf = 1.17549435e-38f;
  h.log_eb = std::log2(ebf);
  h.vmin = vmin;
  h.vmax = vmax;
  h.span = vmax - vmin;
  h.wkv = fs(g, F_AP_WKV);
  h.bk1 = fs(g, F_AP_B1);
  h.wk2 = fs(g, F_AP_W2);
  h.bv = fs(g, F_AP_BV);
  h.null_v = fs(g, F_AP_NULLV);
  h.bk2 = g->scal[2];
  h.null_k = g->scal[3];
  h.w1 = fs(g, F_HD_W1);
  h.cond = fs(g, F_HD_COND);
  h.b1 = fs(g, F_HD_B1);
  h.w2 = fs(g, F_HD_W2);
  h.b2 = fs(g, F_HD_B2);
  constexpr bool BF = sizeof(T) == 2;
  h.wkv_hi = (const uint16_t *)g->hd_wkv_hi;
  h.wkv_lo = (const uint16_t *)g->hd_wkv_lo;
  h.w1_hi = (const uint16_t *)g->hd_w1_hi;
  h.w1_lo = (const uint16_t *)g->hd_w1_lo;
  const size_t smem = BF ? head_split_smem<D>() : head_smem<D>();
  const void *fn;
  if constexpr (BF)
    fn = (const void *)head_split_kernel<D>;
  else
    fn = (const void *)head2_kernel<D, T>;
  int per_cu = 1;
  int rc = lds_setup(fn, 64 * HEAD_WAVES, smem, &per_cu);
  if (rc) return rc;
  const int64_t waves = (M + 15) / 16;
  int64_t grid = (waves + HEAD_WAVES - 1) / HEAD_WAVES;
  if (grid > (int64_t)g->n_cu * per_cu) grid = (int64_t)g->n_cu * per_cu;
  if constexpr (BF)
    hipLaunchKernelGGL((head_split_kernel<D>), dim3((unsigned)grid), dim3(64 * HEAD_WAVES), smem,
                       s, (const uint16_t *)ctx, h, pred, scale, mu, logb);
  else
    hipLaunchKernelGGL((head2_kernel<D, T>), dim3((unsigned)grid), dim3(64 * HEAD_WAVES), smem,
                       s, (const T *)ctx, h, pred, scale, mu, logb);
  return hip_ok(hipGetLastError(), "head kernel launch");
}

template <int D, typename T>
int init_embed(latec_gnn *g, int64_t M, const float *recon, const int64_t *gidx,
               const int64_t *rows, float vmin, float vmax, T *out, hipStream_t s) {
  const InitW<T> w{fs(g, F_IN_W0), fs(g, F_IN_B0), (const T *)g->init_w2, fs(g, F_IN_B2)};
  const size_t smem = init_smem<D, T>();
  int per_cu = 1;
  int rc = lds_setup((const void *)init_embed_kernel<D, T>, 256, smem, &per_cu);
  if (rc) return rc;
  int64_t grid = (M + 63) / 64;
  if (grid > (int64_t)g->n_cu * per_cu) grid = (int64_t)g->n_cu * per_cu;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((init_embed_kernel<D, T>), dim3((unsigned)grid), dim3(256), smem, s, M,
                     g->ndim, recon, gidx, rows, vmin, vmax, vmax - vmin, w, out);
  return hip_ok(hipGetLastError(), "init_embed_kernel launch");
}

template <int D, typename T>
int finalize_t(latec_gnn *g, void *E, const void *ctx, const int64_t *qidx,
               const int64_t *gidx, const float *recon, int64_t M, float vmin, float vmax,
               hipStream_t s) {
  int rc = ensure_work(g, (size_t)M * D * sizeof(T) + 256);
  if (rc) return rc;
  T *vemb = (T *)g->work;
  rc = init_embed<D, T>(g, M, recon, gidx, nullptr, vmin, vmax, vemb, s);
  if (rc) return rc;
  RowArgs a{};
  a.K = g->ndim;
  a.n_rows_host = M * g->ndim;
  a.ctx = ctx;
  a.vemb = vemb;
  a.qout = qidx;
  a.out = E;
  return launch_rows<D, T>(g, ROW_FIN, a, s);
}

template <int D, typename T>
int halo_t(latec_gnn *g, void *E, const int64_t *rows, const int64_t *gidx,
           const float *recon, int64_t M, float vmin, float vmax, hipStream_t s) {
  return init_embed<D, T>(g, M, recon, gidx, rows, vmin, vmax, (T *)E, s);
}

#define GNN_DISPATCH(FN, ...)                                                      \
  (g->d == 64 ? (g->numerics == NUM_BF16 ? FN<64, uint16_t>(g, __VA_ARGS__)        \
                                         : FN<64, float>(g, __VA_ARGS__))          \
              : (g->numerics == NUM_BF16 ? FN<32, uint16_t>(g, __VA_ARGS__)        \
                                         : FN<32, float>(g, __VA_ARGS__)))

}  // namespace

extern "C" {

int64_t latec_gnn_param_count(int d, int agg_level) {
  if (d < 2 || agg_level < 1) return LATEC_E_ARG;
  return make_layout(d, agg_level).total;
}

int latec_gnn_create(int d, int agg_level, int ndim, int numerics, const float *params,
                     int64_t n_params, latec_gnn **out) {
  if (!out || !params) {
    latec::set_error("latec_gnn_create: null argument");
    return LATEC_E_ARG;
  }
  *out = nullptr;
  if (!d_supported(d)) {
    latec::set_error("GNN kernels are built for d = 32 or 64, got d = " + std::to_string(d));
    return LATEC_E_ARG;
  }
  if (agg_level < 1) {
    latec::set_error("agg_level must be >= 1");
    return LATEC_E_ARG;
  }
  if (ndim < 1 || ndim > LATEC_MAX_DIM) {
    latec::set_error("GNN rank must be 1.." + std::to_string(LATEC_MAX_DIM));
    return LATEC_E_ARG;
  }
  if (numerics != NUM_FP32 && numerics != NUM_BF16) {
    latec::set_error("numerics must be 0 (fp32) or 1 (bf16)");
    return LATEC_E_ARG;
  }
  const Layout ly = make_layout(d, agg_level);
  if (n_params != ly.total) {
    latec::set_error("parameter block has " + std::to_string(n_params) + " floats, expected " +
                     std::to_string(ly.total));
    return LATEC_E_ARG;
  }
  auto *g = new latec_gnn();
  g->d = d;
  g->agg = agg_level;
  g->ndim = ndim;
  g->numerics = numerics;
  g->lay = ly;
  build_lines(g);
  if (g->L * ndim * 2 > 1024) {
    latec::set_error("too many (line, axis) state entries for this rank and aggregation level");
    delete g;
    return LATEC_E_ARG;
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&g->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (g->n_cu < 1) g->n_cu = 256;
  int rc = make_device(g, params);
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return LATEC_OK;
}

void latec_gnn_destroy(latec_gnn *g) {
  if (!g) return;
  if (g->dev) (void)hipFree(g->dev);
  if (g->work) (void)hipFree(g->work);
  if (g->d_line_off) (void)hipFree(g->d_line_off);
  for (auto &t : g->geo_tabs) (void)hipFree(t.second);
  for (auto &t : g->gate_tabs) (void)hipFree(t.second);
  for (auto &t : g->class_tabs) (void)hipFree(t.dev);
  if (g->stage_work) (void)hipFree(g->stage_work);
  if (g->stage_host) (void)hipHostFree(g->stage_host);
  delete g;
}

int latec_gnn_n_lines(const latec_gnn *g) { return g ? g->L : LATEC_E_ARG; }

int latec_gnn_geometry(latec_gnn *g, latec_plan *p, int stage, const latec_gnn_frame *fr,
                       int64_t *d_qidx, int64_t *d_gidx, int32_t *d_steps, void *stream) {
  if (!g || !p || !fr) {
    latec::set_error("latec_gnn_geometry: null argument");
    return LATEC_E_ARG;
  }
  if (p->ndim != g->ndim || p->block != 1 || (1 << p->levels) != p->stride) {
    latec::set_error("GNN geometry needs a block-1 plan of the model's rank with stride 2^levels");
    return LATEC_E_ARG;
  }
  if (stage < 0 || stage >= (int)p->stages.size()) {
    latec::set_error("stage out of range");
    return LATEC_E_ARG;
  }
  const auto &st = p->stages[stage];
  if (st.count == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  GeomArgs a{};
  a.ndim = p->ndim;
  a.L = g->L;
  a.stage = stage;
  a.levels_log2 = p->levels;
  a.stride = p->stride;
  a.limit = p->stride;
  a.pad_lo = fr->pad;
  a.has_halo = fr->has_halo;
  int64_t acc = 1;
  for (int j = p->ndim - 1; j >= 0; j--) {
    a.cshape[j] = p->shape[j];
    a.fstr[j] = acc;
    acc *= p->shape[j] + 2 * fr->pad;
  }
  acc = 1;
  for (int j = p->ndim - 1; j >= 0; j--) {
    a.origin[j] = fr->origin[j];
    a.tshape[j] = fr->tshape[j];
    a.tstr[j] = acc;
    acc *= fr->tshape[j];
    a.edges[j] = fr->edges[j] > 0 ? fr->edges[j] : 1;
    a.grid[j] = fr->grid[j] > 0 ? fr->grid[j] : 1;
  }
  a.coded = fr->d_coded;
  if (a.has_halo && !a.coded) {
    latec::set_error("a halo frame needs the coded-chunk flags");
    return LATEC_E_ARG;
  }
  // classes of the stage and their start indices on the device, uploaded
  // once per (plan shape, levels, stride, stage) and handle
  const int nc = (int)st.classes.size();
  a.n_classes = nc;
  std::vector<int64_t> starts(nc + 1, 0);
  for (int c = 0; c < nc; c++) starts[c + 1] = starts[c] + st.classes[c].size;
  const size_t bcls = ((sizeof(ClassDesc) * nc + 255) / 256) * 256;
  std::vector<int64_t> key = {p->ndim, p->levels, p->stride, p->block, stage};
  key.insert(key.end(), p->shape, p->shape + p->ndim);
  void *ctab = nullptr;
  for (auto &t : g->class_tabs)
    if (t.key == key) ctab = t.dev;
  int rc = LATEC_OK;
  if (!ctab) {
    if (g->class_tabs.size() >= 1024) {  // many shapes: start over
      rc = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
      if (rc) return rc;
      for (auto &t : g->class_tabs) (void)hipFree(t.dev);
      g->class_tabs.clear();
    }
    std::vector<unsigned char> h(bcls + sizeof(int64_t) * (nc + 1), 0);
    memcpy(h.data(), st.classes.data(), sizeof(ClassDesc) * nc);
    memcpy(h.data() + bcls, starts.data(), sizeof(int64_t) * (nc + 1));
    rc = hip_ok(hipMalloc(&ctab, h.size()), "hipMalloc(class tables)");
    if (rc) return rc;
    rc = hip_ok(hipMemcpy(ctab, h.data(), h.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) {
      (void)hipFree(ctab);
      return rc;
    }
    g->class_tabs.push_back({key, ctab});
  }
  const ClassDesc *dcls = (const ClassDesc *)ctab;
  const int64_t *dstart = (const int64_t *)((unsigned char *)ctab + bcls);
  const int64_t total = starts[nc];
  // lookup table of the stage: (2^levels)^ndim residue tuples x L lines x 2
  // sides, built once per handle (kept up to 2^16 residue tuples)
  uint8_t *tab = nullptr;
  if (g_gnn_geo_table && p->levels * p->ndim <= 16 && a.limit <= 255) {
    const int64_t key = ((int64_t)stage << 8) | p->levels;
    for (auto &t : g->geo_tabs)
      if (t.first == key) tab = t.second;
    if (!tab) {
      const int64_t n = ((int64_t)1 << (p->levels * p->ndim)) * g->L * 2;
      rc = hip_ok(hipMalloc(&tab, (size_t)n), "hipMalloc(geometry table)");
      if (rc) return rc;
      g->geo_tabs.push_back({key, tab});
      hipLaunchKernelGGL(geo_table_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, a,
                         (const int8_t *)g->d_dirs, n, tab);
      rc = hip_ok(hipGetLastError(), "geo_table_kernel launch");
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(geometry_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, dcls, a,
                     (const int8_t *)g->d_dirs, total, dstart, d_qidx, d_gidx, d_steps,
                     (const uint8_t *)tab);
  return hip_ok(hipGetLastError(), "geometry_kernel launch");
}

int latec_gnn_embed(latec_gnn *g, const void *d_E, const int64_t *frame_shape,
                    const int32_t *d_steps, const int64_t *d_qidx, int64_t M, void *d_ctx,
                    void *stream) {
  if (!g || !frame_shape || M < 0) {
    latec::set_error("latec_gnn_embed: bad argument");
    return LATEC_E_ARG;
  }
  if (M == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  return GNN_DISPATCH(embed_t, d_E, frame_shape, d_steps, d_qidx, M, d_ctx, s);
}

int latec_gnn_head(latec_gnn *g, const void *d_ctx, int64_t M, double eb, float vmin,
                   float vmax, float *d_pred, float *d_scale, float *d_mu, float *d_logb,
                   void *stream) {
  if (!g || M < 0 || !(vmax > vmin)) {
    latec::set_error("latec_gnn_head: bad argument");
    return LATEC_E_ARG;
  }
  if (M == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  return GNN_DISPATCH(head_t, d_ctx, M, eb, vmin, vmax, d_pred, d_scale, d_mu, d_logb, s);
}

int latec_gnn_finalize(latec_gnn *g, void *d_E, const void *d_ctx, const int64_t *d_qidx,
                       const int64_t *d_gidx, const float *d_recon, int64_t M, float vmin,
                       float vmax, void *stream) {
  if (!g || M < 0 || !(vmax > vmin)) {
    latec::set_error("latec_gnn_finalize: bad argument");
    return LATEC_E_ARG;
  }
  if (M == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  return GNN_DISPATCH(finalize_t, d_E, d_ctx, d_qidx, d_gidx, d_recon, M, vmin, vmax, s);
}

int latec_gnn_halo(latec_gnn *g, void *d_E, const int64_t *d_rows, const int64_t *d_gidx,
                   const float *d_recon, int64_t M, float vmin, float vmax, void *stream) {
  if (!g || M < 0 || !(vmax > vmin)) {
    latec::set_error("latec_gnn_halo: bad argument");
    return LATEC_E_ARG;
  }
  if (M == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  return GNN_DISPATCH(halo_t, d_E, d_rows, d_gidx, d_recon, M, vmin, vmax, s);
}

int latec_gnn_halo_frame(latec_gnn *g, latec_plan *p, const latec_gnn_frame *fr, void *d_E,
                         const float *d_recon, float vmin, float vmax, void *stream) {
  if (!g || !p || !fr || !d_E || !d_recon || !(vmax > vmin)) {
    latec::set_error("latec_gnn_halo_frame: bad argument");
    return LATEC_E_ARG;
  }
  if (p->ndim != g->ndim || p->block != 1 || (1 << p->levels) != p->stride) {
    latec::set_error("GNN halo needs a block-1 plan of the model's rank with stride 2^levels");
    return LATEC_E_ARG;
  }
  if (!fr->has_halo || fr->pad <= 0) return LATEC_OK;
  if (!fr->d_coded) {
    latec::set_error("a halo frame needs the coded-chunk flags");
    return LATEC_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  HaloArgs h{};
  h.ndim = p->ndim;
  h.stride = p->stride;
  h.pad = fr->pad;
  h.n_frame = 1;
  for (int j = 0; j < p->ndim; j++) {
    h.cshape[j] = p->shape[j];
    h.fshape[j] = p->shape[j] + 2 * (int64_t)fr->pad;
    h.n_frame *= h.fshape[j];
    h.origin[j] = fr->origin[j];
    h.tshape[j] = fr->tshape[j];
    h.edges[j] = fr->edges[j] > 0 ? fr->edges[j] : 1;
    h.grid[j] = fr->grid[j] > 0 ? fr->grid[j] : 1;
  }
  h.coded = fr->d_coded;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t bl = al(sizeof(int64_t) * (size_t)h.n_frame);
  int rc = ensure_work(g, 2 * bl + 256);
  if (rc) return rc;
  int64_t *frows = (int64_t *)g->work;
  int64_t *gflat = (int64_t *)((unsigned char *)g->work + bl);
  uint32_t *cnt = (uint32_t *)((unsigned char *)g->work + 2 * bl);
  hipLaunchKernelGGL((zero_kernel<uint32_t>), dim3(1), dim3(64), 0, s, (int64_t)1, cnt);
  int64_t blocks = (h.n_frame + 255) / 256;
  if (blocks > (int64_t)g->n_cu * 16) blocks = (int64_t)g->n_cu * 16;
  hipLaunchKernelGGL(halo_list_kernel, dim3((unsigned)blocks), dim3(256), 0, s, h, frows, gflat,
                     cnt);
  rc = hip_ok(hipGetLastError(), "halo_list_kernel launch");
  uint32_t n = 0;
  if (!rc) rc = hip_ok(hipMemcpyAsync(&n, cnt, sizeof(n), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (rc || n == 0) return rc;
  return GNN_DISPATCH(halo_t, d_E, frows, gflat, d_recon, (int64_t)n, vmin, vmax, s);
}

}  // extern "C"

namespace {

// Device copies of a bound's gate tables (63 thresholds, then the per-level
// constants of gate_bits), uploaded once per distinct eb and handle: the
// thresholds and grid a caller passes are functions of eb
// (_gate_scale_thresholds).
int gate_tables(latec_gnn *g, double eb, const float *thr63, const double *grid64,
                const float **d_thr, const double **d_lvk) {
  uint64_t key;
  memcpy(&key, &eb, 8);
  for (auto &t : g->gate_tabs)
    if (t.first == key) {
      *d_thr = (const float *)t.second;
      *d_lvk = (const double *)((unsigned char *)t.second + 256);
      return LATEC_OK;
    }
  if (g->gate_tabs.size() >= 64) {  // many bounds: start over (kernels may still read them)
    int rc = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    for (auto &t : g->gate_tabs) (void)hipFree(t.second);
    g->gate_tabs.clear();
  }
  std::vector<unsigned char> h(256 + sizeof(double) * 3 * 64, 0);
  memcpy(h.data(), thr63, sizeof(float) * 63);
  double *lvk = (double *)(h.data() + 256);
  // per grid level: -log2(-expm1(-eb/b)), eb / (b ln 2), -log2(-expm1(-2 eb/b))
  for (int l = 0; l < 64; l++) {
    const double b = grid64[l];
    lvk[l] = -std::log2(-std::expm1(-eb / b));
    lvk[64 + l] = eb / b / std::log(2.0);
    lvk[128 + l] = -std::log2(-std::expm1(-2.0 * eb / b));
  }
  void *d = nullptr;
  int rc = hip_ok(hipMalloc(&d, h.size()), "hipMalloc(gate tables)");
  if (rc) return rc;
  rc = hip_ok(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice), "hipMemcpy(gate tables)");
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  g->gate_tabs.push_back({key, d});
  *d_thr = (const float *)d;
  *d_lvk = (const double *)((unsigned char *)d + 256);
  return LATEC_OK;
}

// gate_cost_kernel + gate_sum_kernel: the bucket sums ((1 + 4 n_interp) * 13
// doubles) at byte sums_off of the handle's workspace, the block partials
// before them.
int gate_costs_launch(latec_gnn *g, const float *d_x, const float *d_pred, const float *d_scale,
                      const float *const *d_ip, int n_interp, int64_t n, double eb,
                      uint32_t radius, const float *thr63, const double *grid64,
                      size_t sums_off, hipStream_t s) {
  const int nc = 1 + 4 * n_interp;
  int blocks = (int)((n + 1023) / 1024);
  if (blocks > 1024) blocks = 1024;
  const size_t bpart = sizeof(double) * (size_t)blocks * nc * GATE_NB;
  const float *dthr;
  const double *dlvk;
  int rc = gate_tables(g, eb, thr63, grid64, &dthr, &dlvk);
  if (rc) return rc;
  rc = ensure_work(g, (bpart > sums_off ? bpart : sums_off) + sizeof(double) * nc * GATE_NB);
  if (rc) return rc;
  double *part = (double *)g->work;
  double *d_sums = (double *)((unsigned char *)g->work + sums_off);
  GateArgs a{};
  a.n = n;
  a.eb = eb;
  a.two_eb = 2.0 * eb;
  a.radius = radius;
  a.n_cand = (uint32_t)nc;
  a.x = d_x;
  a.pred = d_pred;
  a.scale = d_scale;
  for (int i = 0; i < n_interp; i++) a.ip[i] = d_ip[i];
  a.thr = dthr;
  a.lvk = dlvk;
  for (int t = 0; t < 12; t++) a.bound[t] = eb * std::ldexp(1.0, t + 2);
  a.partial = part;
  hipLaunchKernelGGL(gate_cost_kernel, dim3(blocks), dim3(GATE_BLOCK), 0, s, a);
  rc = hip_ok(hipGetLastError(), "gate_cost_kernel launch");
  if (rc) return rc;
  hipLaunchKernelGGL(gate_sum_kernel, dim3(1), dim3(256), 0, s, (const double *)part, blocks,
                     nc * GATE_NB, d_sums);
  return hip_ok(hipGetLastError(), "gate_sum_kernel launch");
}

}  // namespace

extern "C" {

int latec_gnn_gate_costs(latec_gnn *g, const float *d_x, const float *d_pred,
                         const float *d_scale, const float *const *d_ip, int n_interp,
                         int64_t n, double eb, uint32_t radius, const float *thr63,
                         const double *grid64, double *costs, void *stream) {
  if (!g || n_interp < 0 || n_interp > 3 || !thr63 || !grid64 || !costs) {
    latec::set_error("latec_gnn_gate_costs: bad argument");
    return LATEC_E_ARG;
  }
  const int nc = 1 + 4 * n_interp;
  for (int i = 0; i < nc * GATE_NB; i++) costs[i] = 0.0;
  if (n == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  const size_t sums_off = ((sizeof(double) * 1025 * nc * GATE_NB + 255) / 256) * 256;
  int rc = gate_costs_launch(g, d_x, d_pred, d_scale, d_ip, n_interp, n, eb, radius, thr63,
                             grid64, sums_off, s);
  if (rc) return rc;
  rc = hip_ok(hipMemcpyAsync(costs, (unsigned char *)g->work + sums_off,
                             sizeof(double) * nc * GATE_NB, hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  return rc;
}

}  // extern "C"

namespace {

// _gate_select (gnn_codec.py:920-971) on the host from the bucket sums W
// ((1 + 4 n_interp) rows of 13): the same float64 operations in the same
// order as the NumPy function (running sums along the buckets; the cost of
// taking candidate row r below bucket j is (tot - base[j]) + ci[j], above it
// base[j] + (ci[12] - ci[j]); the first minimum over [low rows; high rows]
// wins if it is below tot; a NaN cost, which NumPy's argmin would return,
// never gates).  out = {kind, dir, T, shift}.
void gate_select_host(const double *W, int n_interp, int32_t out[4]) {
  const int nrow = 4 * n_interp, NB = GATE_NB;
  double base[GATE_NB], acc = 0.0;
  for (int b = 0; b < NB; b++) base[b] = (acc += W[b]);
  const double tot = base[NB - 1];
  double best = 0.0;
  int fi = -1;
  bool nan = false;
  for (int half = 0; half < 2; half++)
    for (int r = 0; r < nrow; r++) {
      double ci[GATE_NB], a2 = 0.0;
      for (int b = 0; b < NB; b++) ci[b] = (a2 += W[(1 + r) * NB + b]);
      for (int j = 0; j < NB - 1; j++) {
        const double c = half == 0 ? (tot - base[j]) + ci[j] : base[j] + (ci[NB - 1] - ci[j]);
        nan |= c != c;
        if (fi < 0 || c < best) {
          best = c;
          fi = (half * nrow + r) * (NB - 1) + j;
        }
      }
    }
  out[0] = out[1] = out[2] = out[3] = 0;
  if (fi < 0 || nan || !(best < tot)) return;
  const int row = fi / (NB - 1), side_row = row % nrow;
  out[0] = side_row / 4 + 1;
  out[1] = row >= nrow;
  out[2] = 2 + fi % (NB - 1);
  out[3] = 2 * (side_row % 4);
}

int stage_work(latec_gnn *g, size_t bytes) {
  if (g->stage_work_bytes >= bytes) return LATEC_OK;
  if (g->stage_work) {
    int rc = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    (void)hipFree(g->stage_work);
    g->stage_work = nullptr;
    g->stage_work_bytes = 0;
  }
  bytes = (bytes * 5) / 4 + (1 << 20);
  int rc = hip_ok(hipMalloc(&g->stage_work, bytes), "hipMalloc(stage workspace)");
  if (!rc) g->stage_work_bytes = bytes;
  return rc;
}

}  // namespace

extern "C" {

int latec_gnn_code_stage(latec_gnn *g, latec_plan *p, int stage, const latec_gnn_stage_io *io,
                         void *stream) {
  if (!g || !p || !io || !io->d_pred || !io->d_scale || !io->d_gidx || !io->d_x ||
      !io->d_recon || !io->opts || !io->level_thr || !io->codes || !io->levels ||
      !io->outliers || !io->n_out) {
    latec::set_error("latec_gnn_code_stage: null argument");
    return LATEC_E_ARG;
  }
  if (stage < 1 || stage >= (int)p->stages.size() || p->channels != 1 ||
      io->M != p->stages[stage].count || io->n < 0 || io->n > io->M ||
      (io->n < io->M && !io->d_keep)) {
    latec::set_error("latec_gnn_code_stage: bad stage or counts");
    return LATEC_E_ARG;
  }
  const int64_t M = io->M, n = io->n;
  *io->n_out = 0;
  if (io->gate_word) *io->gate_word = 0;
  if (n == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  const latec_interp_opts &o = *io->opts;
  const int cb = o.code_bytes;
  const bool gate = io->gate && !io->interp_kind;
  const int ncand = gate ? 3 : (io->interp_kind ? 1 : 0);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  // p, sc, x, rec (n f32), gidx (n i64), codes (n), rank, val (n), cnt, levels (n),
  // candidates (ncand x M full, ncand x n kept)
  const size_t b_f = al(4 * (size_t)n), b_i = al(8 * (size_t)n), b_c = al((size_t)cb * n);
  const size_t b_l = al((size_t)n), b_cf = al(4 * (size_t)M);
  const size_t need = 6 * b_f + b_i + b_c + b_l + 256 + ncand * (b_cf + b_f);
  int rc = stage_work(g, need);
  if (rc) return rc;
  unsigned char *w = (unsigned char *)g->stage_work;
  float *dp = (float *)w, *dsc = (float *)(w + b_f), *dx = (float *)(w + 2 * b_f);
  float *drec = (float *)(w + 3 * b_f);
  uint32_t *drank = (uint32_t *)(w + 4 * b_f);
  float *dval = (float *)(w + 5 * b_f);
  int64_t *dgidx = (int64_t *)(w + 6 * b_f);
  void *dcodes = w + 6 * b_f + b_i;
  uint8_t *dlv = (uint8_t *)(w + 6 * b_f + b_i + b_c);
  uint32_t *dcnt = (uint32_t *)(w + 6 * b_f + b_i + b_c + b_l);
  float *dcf[3] = {nullptr, nullptr, nullptr}, *dck[3] = {nullptr, nullptr, nullptr};
  for (int c = 0; c < ncand; c++) {
    dcf[c] = (float *)(w + 6 * b_f + b_i + b_c + b_l + 256 + c * (b_cf + b_f));
    dck[c] = (float *)((unsigned char *)dcf[c] + b_cf);
  }
  const int gb = grid_for(n, 256);
  // the coded rows of the stage (the chunk's column split) and their values
  if (io->d_keep) {
    hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, io->d_pred, io->d_keep, dp);
    hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, io->d_scale, io->d_keep,
                       dsc);
    hipLaunchKernelGGL(gather_i64_kernel, dim3(gb), dim3(256), 0, s, n, io->d_gidx, io->d_keep,
                       dgidx);
  } else {
    rc = hip_ok(hipMemcpyAsync(dp, io->d_pred, 4 * n, hipMemcpyDeviceToDevice, s), "copy");
    if (!rc) rc = hip_ok(hipMemcpyAsync(dsc, io->d_scale, 4 * n, hipMemcpyDeviceToDevice, s),
                         "copy");
    if (!rc) rc = hip_ok(hipMemcpyAsync(dgidx, io->d_gidx, 8 * n, hipMemcpyDeviceToDevice, s),
                         "copy");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, io->d_x, dgidx, dx);
  // interpolation candidates of the stage on the chunk's block (kinds 1..3, or
  // the one kind of an interpolation-only encode)
  for (int c = 0; c < ncand; c++) {
    const int kind = gate ? c + 1 : io->interp_kind;
    rc = latec_interp_predict(p, stage, io->d_block, dcf[c], &io->cand_opts[kind - 1], stream);
    if (rc) return rc;
    if (io->d_keep)
      hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, (const float *)dcf[c],
                         io->d_keep, dck[c]);
    else
      dck[c] = dcf[c];
  }
  if (io->interp_kind) {
    rc = hip_ok(hipMemcpyAsync(dp, dck[0], 4 * n, hipMemcpyDeviceToDevice, s), "copy");
    if (rc) return rc;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(gb), dim3(256), 0, s, n, (float)io->eb, dsc);
  } else if (gate) {
    int32_t *ch = io->choice;
    if (!io->reuse_choice) {
      const int nc = 1 + 4 * 3;
      const size_t sums_off = ((sizeof(double) * 1025 * nc * GATE_NB + 255) / 256) * 256;
      const float *ips[3] = {dck[0], dck[1], dck[2]};
      rc = gate_costs_launch(g, dx, dp, dsc, ips, 3, n, io->eb, io->radius, io->gate_thr,
                             io->gate_grid, sums_off, s);
      if (rc) return rc;
      double W[13 * GATE_NB];
      rc = hip_ok(hipMemcpyAsync(W, (unsigned char *)g->work + sums_off, sizeof(W),
                                 hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
      if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
      if (rc) return rc;
      gate_select_host(W, 3, ch);
      if (io->gate_word)
        *io->gate_word = (uint32_t)((ch[1] << 10) + ch[0] * 256 + ch[2] * 16 + ch[3]);
    }
    if (ch[2] > 0) {
      const float *ip = dck[ch[0] - 1];
      hipLaunchKernelGGL(gate_apply_kernel, dim3(gb), dim3(256), 0, s, n, dp, dsc, ip,
                         io->eb * std::ldexp(1.0, ch[2]), (float)std::ldexp(1.0, -ch[3]),
                         (int)ch[1]);
    }
  }
  // quantize, the reconstruction (outliers written over it), into recon
  rc = hip_ok(hipMemsetAsync(dcnt, 0, 4, s), "hipMemsetAsync");
  if (rc) return rc;
  rc = latec_quantize(dx, dp, n, io->eb, &o, dcodes, drank, dval, dcnt, io->d_hist, stream);
  if (rc) return rc;
  rc = latec_dequantize(dp, dcodes, n, drank, dval, dcnt, io->eb, &o, drec, stream);
  if (rc) return rc;
  int ob = gb < 1024 ? gb : 1024;
  hipLaunchKernelGGL(outlier_fix_kernel, dim3(ob), dim3(256), 0, s, (const uint32_t *)drank,
                     (const float *)dval, (const uint32_t *)dcnt, n, drec);
  hipLaunchKernelGGL(scatter_kernel, dim3(gb), dim3(256), 0, s, n, (const float *)drec,
                     (const int64_t *)dgidx, io->d_recon);
  rc = latec_scale_to_level(dsc, n, io->level_thr, dlv, stream);
  if (rc) return rc;
  // codes, levels and the outliers by stage position to the host
  uint32_t k = 0;
  rc = hip_ok(hipMemcpyAsync(&k, dcnt, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipMemcpyAsync(io->codes, dcodes, (size_t)cb * n, hipMemcpyDeviceToHost,
                                      s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipMemcpyAsync(io->levels, dlv, (size_t)n, hipMemcpyDeviceToHost, s),
                       "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (rc) return rc;
  if (k > (uint32_t)n) {
    latec::set_error("latec_gnn_code_stage: outlier count past the stage");
    return LATEC_E_COUNT;
  }
  if (k) {
    std::vector<uint32_t> r(k);
    std::vector<float> v(k);
    rc = hip_ok(hipMemcpy(r.data(), drank, 4 * (size_t)k, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = hip_ok(hipMemcpy(v.data(), dval, 4 * (size_t)k, hipMemcpyDeviceToHost),
                         "hipMemcpy");
    if (rc) return rc;
    std::vector<uint32_t> ord(k);
    for (uint32_t j = 0; j < k; j++) ord[j] = j;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return r[a] < r[b]; });
    for (uint32_t j = 0; j < k; j++) io->outliers[j] = v[ord[j]];
  }
  *io->n_out = k;
  return LATEC_OK;
}

// Decoder stage, first half: the coded rows' prediction and scale (gate
// applied from the stream's word), their scale levels to the host.  The rows'
// prediction and tensor indices stay in the stage workspace for
// latec_gnn_decode_stage_end.
int latec_gnn_decode_stage_begin(latec_gnn *g, latec_plan *p, int stage,
                                 const latec_gnn_dstage_io *io, void *stream) {
  if (!g || !p || !io || !io->d_pred || !io->d_scale || !io->d_gidx || !io->level_thr ||
      !io->levels || !io->cand_opts) {
    latec::set_error("latec_gnn_decode_stage_begin: null argument");
    return LATEC_E_ARG;
  }
  if (stage < 1 || stage >= (int)p->stages.size() || p->channels != 1 ||
      io->M != p->stages[stage].count || io->n < 0 || io->n > io->M ||
      (io->n < io->M && !io->d_keep)) {
    latec::set_error("latec_gnn_decode_stage_begin: bad stage or counts");
    return LATEC_E_ARG;
  }
  const int64_t M = io->M, n = io->n;
  g->dec_n = -1;
  if (n == 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t word = io->gate_word;
  const int gk = (int)((word >> 8) & 3), gt = (int)((word >> 4) & 15), gs = (int)(word & 15);
  const int gd = (int)((word >> 10) & 1);
  const int kind = io->interp_kind ? io->interp_kind : (gt > 0 ? gk : 0);
  if (kind < 0 || kind > 3 || (io->interp_kind == 0 && gt > 0 && gk == 0)) {
    latec::set_error("latec_gnn_decode_stage_begin: bad gate word");
    return LATEC_E_ARG;
  }
  if (kind && !io->d_block) {
    latec::set_error("latec_gnn_decode_stage_begin: the gate needs the chunk block");
    return LATEC_E_ARG;
  }
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_f = al(4 * (size_t)n), b_i = al(8 * (size_t)n), b_cf = al(4 * (size_t)M);
  // p, sc, rec (n f32), gidx (n i64), codes (4n), ranks (4n), vals (4n), cnt,
  // levels (n), one candidate (M full, n kept)
  const size_t need = 3 * b_f + b_i + 3 * b_f + 256 + al((size_t)n) + b_cf + b_f;
  int rc = stage_work(g, need);
  if (rc) return rc;
  unsigned char *w = (unsigned char *)g->stage_work;
  float *dp = (float *)w, *dsc = (float *)(w + b_f);
  int64_t *dgidx = (int64_t *)(w + 3 * b_f);
  uint8_t *dlv = (uint8_t *)(w + 3 * b_f + b_i + 3 * b_f + 256);
  float *dcf = (float *)(w + 3 * b_f + b_i + 3 * b_f + 256 + al((size_t)n));
  float *dck = (float *)((unsigned char *)dcf + b_cf);
  const int gb = grid_for(n, 256);
  if (io->d_keep) {
    hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, io->d_pred, io->d_keep, dp);
    hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, io->d_scale, io->d_keep,
                       dsc);
    hipLaunchKernelGGL(gather_i64_kernel, dim3(gb), dim3(256), 0, s, n, io->d_gidx, io->d_keep,
                       dgidx);
  } else {
    rc = hip_ok(hipMemcpyAsync(dp, io->d_pred, 4 * n, hipMemcpyDeviceToDevice, s), "copy");
    if (!rc) rc = hip_ok(hipMemcpyAsync(dsc, io->d_scale, 4 * n, hipMemcpyDeviceToDevice, s),
                         "copy");
    if (!rc) rc = hip_ok(hipMemcpyAsync(dgidx, io->d_gidx, 8 * n, hipMemcpyDeviceToDevice, s),
                         "copy");
    if (rc) return rc;
  }
  if (kind) {
    rc = latec_interp_predict(p, stage, io->d_block, dcf, &io->cand_opts[kind - 1], stream);
    if (rc) return rc;
    if (io->d_keep)
      hipLaunchKernelGGL(gather_kernel, dim3(gb), dim3(256), 0, s, n, (const float *)dcf,
                         io->d_keep, dck);
    else
      dck = dcf;
    if (io->interp_kind) {
      rc = hip_ok(hipMemcpyAsync(dp, dck, 4 * n, hipMemcpyDeviceToDevice, s), "copy");
      if (rc) return rc;
      hipLaunchKernelGGL(fill_f32_kernel, dim3(gb), dim3(256), 0, s, n, (float)io->eb, dsc);
    } else {
      hipLaunchKernelGGL(gate_apply_kernel, dim3(gb), dim3(256), 0, s, n, dp, dsc,
                         (const float *)dck, io->eb * std::ldexp(1.0, gt),
                         (float)std::ldexp(1.0, -gs), gd);
    }
  }
  rc = latec_scale_to_level(dsc, n, io->level_thr, dlv, stream);
  if (!rc) rc = hip_ok(hipMemcpyAsync(io->levels, dlv, (size_t)n, hipMemcpyDeviceToHost, s),
                       "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (!rc) g->dec_n = n;
  return rc;
}

// Decoder stage, second half: the decoded codes (host, n of opts->code_bytes)
// and outlier values (host, n_out, by stage position) dequantized against
// the prediction latec_gnn_decode_stage_begin left, written into d_recon at
// the rows' tensor indices.  Stream-ordered; the host arrays are copied to
// pinned staging before return.
int latec_gnn_decode_stage_end(latec_gnn *g, const void *codes, const float *outliers,
                               int64_t n_out, double eb, const latec_interp_opts *o,
                               float *d_recon, void *stream) {
  if (!g || !o || !d_recon || (g->dec_n > 0 && !codes) || n_out < 0 ||
      (n_out && !outliers)) {
    latec::set_error("latec_gnn_decode_stage_end: bad argument");
    return LATEC_E_ARG;
  }
  const int64_t n = g->dec_n;
  if (n < 0) {
    latec::set_error("latec_gnn_decode_stage_end without latec_gnn_decode_stage_begin");
    return LATEC_E_ARG;
  }
  g->dec_n = -1;
  if (n == 0) return LATEC_OK;
  const int cb = o->code_bytes;
  if (cb != 2 && cb != 4) {
    latec::set_error("latec_gnn_decode_stage_end: code_bytes must be 2 or 4");
    return LATEC_E_ARG;
  }
  // ranks of the zero codes, ascending (the order the outliers are stored in)
  std::vector<uint32_t> zr;
  for (int64_t i = 0; i < n; i++) {
    const uint32_t c = cb == 2 ? ((const uint16_t *)codes)[i] : ((const uint32_t *)codes)[i];
    if (c == 0) zr.push_back((uint32_t)i);
  }
  if ((int64_t)zr.size() != n_out) {
    latec::set_error("outlier count mismatch: " + std::to_string(zr.size()) + " zeros vs " +
                     std::to_string(n_out) + " values");
    return LATEC_E_COUNT;
  }
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t k = zr.size();
  const size_t hb_c = al((size_t)cb * n), hb_k = al(4 * (k ? k : 1));
  const size_t hbytes = hb_c + 2 * hb_k + 256;
  hipStream_t s = (hipStream_t)stream;
  if (g->stage_host_bytes < hbytes) {
    // every copy out of the old staging has landed (whatever stream an
    // earlier call used) before it is freed
    int rc = hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (rc) return rc;
    if (g->stage_host) (void)hipHostFree(g->stage_host);
    g->stage_host = nullptr;
    g->stage_host_bytes = 0;
    const size_t want = hbytes * 5 / 4 + (1 << 20);
    rc = hip_ok(hipHostMalloc(&g->stage_host, want, hipHostMallocDefault), "hipHostMalloc");
    if (rc) return rc;
    g->stage_host_bytes = want;
  }
  unsigned char *h = (unsigned char *)g->stage_host;
  memcpy(h, codes, (size_t)cb * n);
  if (k) {
    memcpy(h + hb_c, zr.data(), 4 * k);
    memcpy(h + hb_c + hb_k, outliers, 4 * k);
  }
  const uint32_t k32 = (uint32_t)k;
  memcpy(h + hb_c + 2 * hb_k, &k32, 4);
  auto alw = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_f = alw(4 * (size_t)n), b_i = alw(8 * (size_t)n);
  unsigned char *w = (unsigned char *)g->stage_work;
  const float *dp = (const float *)w;
  float *drec = (float *)(w + 2 * b_f);
  const int64_t *dgidx = (const int64_t *)(w + 3 * b_f);
  unsigned char *dcodes = w + 3 * b_f + b_i;
  uint32_t *drank = (uint32_t *)(dcodes + b_f);
  float *dval = (float *)(dcodes + 2 * b_f);
  uint32_t *dcnt = (uint32_t *)(dcodes + 3 * b_f);
  int rc = hip_ok(hipMemcpyAsync(dcodes, h, (size_t)cb * n, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync");
  if (!rc && k) rc = hip_ok(hipMemcpyAsync(drank, h + hb_c, 4 * k, hipMemcpyHostToDevice, s),
                            "hipMemcpyAsync");
  if (!rc && k) rc = hip_ok(hipMemcpyAsync(dval, h + hb_c + hb_k, 4 * k, hipMemcpyHostToDevice,
                                           s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipMemcpyAsync(dcnt, h + hb_c + 2 * hb_k, 4, hipMemcpyHostToDevice, s),
                       "hipMemcpyAsync");
  if (rc) return rc;
  rc = latec_dequantize(dp, dcodes, n, drank, dval, dcnt, eb, o, drec, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                     (const float *)drec, dgidx, d_recon);
  return hip_ok(hipGetLastError(), "scatter_kernel launch");
}

int latec_gnn_gate_apply(float *d_pred, float *d_scale, const float *d_ip, int64_t n,
                         double threshold, float mult, int dir, void *stream) {
  if (n <= 0) return LATEC_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(gate_apply_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, d_pred,
                     d_scale, d_ip, threshold, mult, dir);
  return hip_ok(hipGetLastError(), "gate_apply_kernel launch");
}

int latec_gather_f32(const float *d_src, const int64_t *d_idx, int64_t n, float *d_dst,
                     void *stream) {
  if (n <= 0) return LATEC_OK;
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     n, d_src, d_idx, d_dst);
  return hip_ok(hipGetLastError(), "gather_kernel launch");
}

int latec_scatter_f32(const float *d_src, const int64_t *d_idx, int64_t n, float *d_dst,
                      void *stream) {
  if (n <= 0) return LATEC_OK;
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     n, d_src, d_idx, d_dst);
  return hip_ok(hipGetLastError(), "scatter_kernel launch");
}

}  // extern "C"
