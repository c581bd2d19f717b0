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

#include <cmath>
#include <cstring>
#include <string>
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

// One two-layer row MLP (message or mix): y = W2 gelu(W1 [a | b] + add) + b2.
// W1 is stored (H, K1) with K1 = 2D (two D-wide input halves) or D; the
// additive terms are per-state bias rows (S0/S1: ns x H), per-row scalar
// coefficient columns (c0/c1: H) and the bias b1 (H).
template <typename T>
struct MlpW {
  const T *w1;        // (H, K1) row-major
  const T *w2;        // (D, H)
  const float *b1;    // (H)
  const float *b2;    // (D)
  const float *s0;    // (ns, H) state bias of half 0 (or nullptr)
  const float *s1;    // (ns, H) state bias of half 1 (or nullptr)
  const float *c0;    // (H) coefficient of row scalar 0 (or nullptr)
  const float *c1;    // (H) coefficient of row scalar 1 (or nullptr)
  int k1;
};

struct PoolW {  // AttnPool in f32: WkT/WvT are (D, D) transposed (in, out)
  const float *wk1t, *bk1, *wk2, *wvt, *bv, *null_v;
  float bk2, null_k;
};

template <typename T>
struct KeyW {  // line-pool key MLP for the message rows: Wk1 (out, in) in T
  const T *wk1;
  const float *bk1, *wk2;
  float bk2;
};

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
  void *mx_w1 = nullptr, *mx_w2 = nullptr, *lk_w1 = nullptr;
  float *f32 = nullptr;            // f32 arrays (biases, state rows, pools, head, init)
  std::vector<int64_t> f32_off;    // named offsets into f32 (see make_device)
  int8_t *d_dirs = nullptr, *d_sidx = nullptr;
  float *d_lognnz = nullptr;
  float scal[4] = {0, 0, 0, 0};    // line pool bk2, null_k; axis pool bk2, null_k
  int n_cu = 256;
  int64_t *d_line_off = nullptr;   // frame offsets of the lines (cached per frame)
  int64_t line_off_key[LATEC_MAX_DIM] = {0};
  // workspace
  void *work = nullptr;
  size_t work_bytes = 0;
};

namespace latec {
namespace gnn {

// f32 array slots inside latec_gnn::f32
enum F32Slot {
  F_BI_B1, F_BI_B2, F_BI_S0, F_BI_S1, F_BI_C0, F_BI_C1,
  F_DIR_B1, F_DIR_B2, F_DIR_S, F_DIR_CSIGN, F_DIR_CLOGD,
  F_MX_B1, F_MX_B2,
  F_LK_B1, F_LK_W2, F_LK_WVT, F_LK_BV, F_LK_NULLV, F_LK_W1T,
  F_AP_W1T, F_AP_B1, F_AP_W2, F_AP_WVT, F_AP_BV, F_AP_NULLV,
  F_HD_W1T, F_HD_COND, F_HD_B1, F_HD_W2, F_HD_B2,
  F_IN_W0, F_IN_B0, F_IN_W2T, F_IN_B2,
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

// One thread per query point of the stage (enumerated class by class).
// Writes qidx[rank] (frame flat index), gidx[rank] (tensor flat index) and
// steps[rank * L + l] = sp | sn << 16 (0 = no usable sample on that side).
__global__ void geometry_kernel(const ClassDesc *__restrict__ cls, GeomArgs g,
                                const int8_t *__restrict__ dirs, int64_t total,
                                const int64_t *__restrict__ cls_start,
                                int64_t *__restrict__ qidx, int64_t *__restrict__ gidx,
                                int32_t *__restrict__ steps) {
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
// row MLP kernel (message / mix) on MFMA
// ---------------------------------------------------------------------------

struct RowArgs {
  int32_t K, L, ns, pad;
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
  // finalize
  const void *ctx;             // (M, K, D)
  const void *vemb;            // (M, D)
  const int64_t *qout;         // (M) frame index of each finalized point
  float score_scale;
  int32_t pad2;
};

template <int D, typename T>
struct RowCfg {
  static constexpr int H = 2 * D;
  static constexpr bool BF = sizeof(T) == 2;
  static constexpr int KS = BF ? 32 : 4;            // MFMA K per instruction
  static constexpr int W1P = 2 * D + (BF ? 8 : 4);  // padded LDS rows (elements)
  static constexpr int W2P = H + (BF ? 8 : 4);
  static constexpr int WKP = D + (BF ? 8 : 4);
  static constexpr int HP = H + (BF ? 8 : 4);
  static constexpr int MP = D + (BF ? 8 : 4);
  static constexpr int WAVES = 8;
  static constexpr int ROWS = 16;                   // rows per wave per iteration
  static constexpr int THREADS = 64 * WAVES;
};

template <typename T>
__device__ __forceinline__ void mfma_step(f32x4 &acc, const T *a, const T *b);

// bf16: a, b point to 8 consecutive bf16 each
template <>
__device__ __forceinline__ void mfma_step<uint16_t>(f32x4 &acc, const uint16_t *a,
                                                    const uint16_t *b) {
  bf16x8 va, vb;
  memcpy(&va, a, 16);
  memcpy(&vb, b, 16);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc, 0, 0, 0);
}
template <>
__device__ __forceinline__ void mfma_step<float>(f32x4 &acc, const float *a, const float *b) {
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(*a, *b, acc, 0, 0, 0);
}

// GELU of the bf16 path: x * Phi(x) with erf from Abramowitz & Stegun
// 7.1.26 (|error| < 1.5e-7, far below the bf16 rounding of the result).
__device__ __forceinline__ float gelu_bf(float x) {
  const float z = fabsf(x) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * z);
  float y = 1.061405429f;
  y = y * t - 1.453152027f;
  y = y * t + 1.421413741f;
  y = y * t - 0.284496736f;
  y = y * t + 0.254829592f;
  const float e = 1.0f - y * t * __expf(-z * z);  // erf(|x| / sqrt 2)
  return 0.5f * x * (1.0f + copysignf(e, x));
}

template <typename T>
__device__ __forceinline__ float gelu_t(float x) {
  return sizeof(T) == 2 ? gelu_bf(x) : gelu(x);
}

// Wave-local LDS hand-off: every lane's LDS writes are complete before any
// lane reads (one wave; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row MLP: for each row r (item i = r / K, axis k = r % K):
//   x = [half0 | half1]  (D or 2D wide, T)
//   h = gelu(W1 x + b1 + S0[st0] + S1[st1] + c0 * z0 + c1 * z1)   (H)
//   y = W2 h + b2                                                  (D)
// messages also get score = (wk2 . gelu(Wk1 y + bk1) + bk2) * scale.
// Each wave walks its own 16-row groups (no workgroup barrier after the
// weights are staged) and prefetches the next group's pair records.
// MFMA fragment layouts (16x16 tiles, lane l): A row l%16, k = KS/4*(l/16)..;
// B col l%16, same k; D rows 4*(l/16)..+3, col l%16.
template <int D, typename T, int MODE>
__global__ __launch_bounds__(512) void row_mlp_kernel(RowArgs a, MlpW<T> w, KeyW<T> pk) {
  using C = RowCfg<D, T>;
  constexpr int H = C::H, KS = C::KS, KL = KS / 4;
  constexpr int K1 = MODE == ROW_DIR ? D : 2 * D;
  constexpr int NW = C::BF ? H * C::W1P + D * C::W2P + D * C::WKP : 0;
  extern __shared__ __align__(16) unsigned char smem[];
  T *sW1 = (T *)smem;
  T *sW2 = sW1 + H * C::W1P;
  T *sWk = sW2 + D * C::W2P;
  float *tb = (float *)(smem + sizeof(T) * (size_t)NW);
  float *tb1 = tb, *tc0 = tb1 + H, *tc1 = tc0 + H, *tb2 = tc1 + H, *tbk1 = tb2 + D;
  float *twk2 = tbk1 + D, *ts0 = twk2 + D, *ts1 = ts0 + a.ns * H;
  int8_t *tsx = (int8_t *)(ts1 + a.ns * H);                      // L*K*2, padded to 1024
  unsigned char *pw = (unsigned char *)tsx + 1024;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr size_t PER_WAVE = sizeof(T) * (C::ROWS * C::HP + C::ROWS * C::MP) + 16 * 4 * 4;
  T *sH = (T *)(pw + wv * PER_WAVE);
  T *sM = sH + C::ROWS * C::HP;
  float *add = (float *)(sM + C::ROWS * C::MP);
  if constexpr (C::BF) {
    for (int i = tid; i < H * K1; i += C::THREADS) sW1[(i / K1) * C::W1P + i % K1] = w.w1[i];
    for (int i = tid; i < D * H; i += C::THREADS) sW2[(i / H) * C::W2P + i % H] = w.w2[i];
    if constexpr (MODE != ROW_FIN)
      for (int i = tid; i < D * D; i += C::THREADS) sWk[(i / D) * C::WKP + i % D] = pk.wk1[i];
  }
  for (int i = tid; i < H; i += C::THREADS) {
    tb1[i] = w.b1[i];
    tc0[i] = w.c0 ? w.c0[i] : 0.f;
    tc1[i] = w.c1 ? w.c1[i] : 0.f;
  }
  for (int i = tid; i < D; i += C::THREADS) {
    tb2[i] = w.b2[i];
    if constexpr (MODE != ROW_FIN) {
      tbk1[i] = pk.bk1[i];
      twk2[i] = pk.wk2[i];
    }
  }
  for (int i = tid; i < a.ns * H; i += C::THREADS) {
    ts0[i] = w.s0 ? w.s0[i] : 0.f;
    ts1[i] = w.s1 ? w.s1[i] : 0.f;
  }
  if constexpr (MODE != ROW_FIN)
    for (int i = tid; i < a.L * a.K * 2; i += C::THREADS) tsx[i] = a.sidx[i];
  __syncthreads();

  const T *gW1 = C::BF ? sW1 : w.w1;
  const T *gW2 = C::BF ? sW2 : w.w2;
  const T *gWk = C::BF ? sWk : pk.wk1;
  constexpr int P1 = C::BF ? C::W1P : K1, P2 = C::BF ? C::W2P : H, PK = C::BF ? C::WKP : D;
  const int64_t n_rows = MODE == ROW_FIN ? a.n_rows_host : (int64_t)a.n_items_dev[0] * a.K;
  const int64_t slot_base = a.slot_base_dev ? (int64_t)a.slot_base_dev[0] : a.slot_base;
  const T *E = (const T *)a.E;
  const int64_t gstep = (int64_t)gridDim.x * C::WAVES;
  const int ri = lane & 15, kq = (lane >> 4) * KL;
  constexpr int NK1 = K1 / KS, NK2 = H / KS, NK3 = D / KS;

  int64_t grp = (int64_t)blockIdx.x * C::WAVES + wv;
  // metadata of this lane's row in group g
  PairRec rc{0, 0, 0.f, 0.f};
  int16_t lc = 0;
  auto fetch = [&](int64_t g, PairRec &r, int16_t &lw) {
    const int64_t row = g * C::ROWS + ri;
    if constexpr (MODE != ROW_FIN) {
      if (row < n_rows) {
        const int64_t item = row / a.K;
        r = a.rec[item];
        lw = a.lw[item];
      }
    }
  };
  if (grp * C::ROWS < n_rows) fetch(grp, rc, lc);
  for (; grp * C::ROWS < n_rows; grp += gstep) {
    // prefetch the next group's records
    PairRec rn{0, 0, 0.f, 0.f};
    int16_t ln = 0;
    if ((grp + gstep) * C::ROWS < n_rows) fetch(grp + gstep, rn, ln);
    const int64_t r0 = grp * C::ROWS;
    const int64_t r = r0 + ri;
    const bool live = r < n_rows;
    const int64_t item = live ? r / a.K : 0;
    const int k = live ? (int)(r % a.K) : 0;
    const T *src0 = nullptr, *src1 = nullptr;
    float z0 = 0.f, z1 = 0.f;
    int st0 = 0, st1 = 0;
    if (live) {
      if constexpr (MODE == ROW_FIN) {
        src0 = (const T *)a.ctx + (item * a.K + k) * D;
        src1 = (const T *)a.vemb + item * D;
      } else if constexpr (MODE == ROW_BI) {
        src0 = E + ((int64_t)rc.ra * a.K + k) * D;  // negative side first
        src1 = E + ((int64_t)rc.rb * a.K + k) * D;
        const int l = lc & 0x3fff;
        st0 = tsx[(l * a.K + k) * 2 + 0];
        st1 = tsx[(l * a.K + k) * 2 + 1];
        z0 = rc.za;
        z1 = rc.zb;
      } else {
        src0 = E + ((int64_t)rc.ra * a.K + k) * D;
        const int l = lc & 0x3fff, side = (lc >> 14) & 1;
        st0 = tsx[(l * a.K + k) * 2 + side];
        z0 = rc.za;
        z1 = rc.zb;
      }
    }
    if (lane < 16) {
      add[ri * 4 + 0] = z0;
      add[ri * 4 + 1] = z1;
      add[ri * 4 + 2] = __int_as_float(st0);
      add[ri * 4 + 3] = __int_as_float(st1);
    }
    T afr[NK1][KL];
#pragma unroll
    for (int ks = 0; ks < NK1; ks++) {
      const int kk = ks * KS + kq;
      const T *sp_ = (MODE == ROW_DIR || kk < D) ? src0 : src1;
      const int ko = (MODE == ROW_DIR || kk < D) ? kk : kk - D;
      if (live) {
        if constexpr (C::BF) {
          memcpy(&afr[ks][0], sp_ + ko, 16);
        } else {
#pragma unroll
          for (int e = 0; e < KL; e++) afr[ks][e] = sp_[ko + e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < KL; e++) afr[ks][e] = T(0);
      }
    }
    wave_lds_sync();
    // --- layer 1
#pragma unroll 1
    for (int cb = 0; cb < H / 16; cb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const T *bw = gW1 + (cb * 16 + (lane & 15)) * P1 + kq;
#pragma unroll
      for (int ks = 0; ks < NK1; ks++) mfma_step<T>(acc, afr[ks], bw + ks * KS);
      const int col = cb * 16 + (lane & 15);
      const float b1 = tb1[col], c0 = tc0[col], c1 = tc1[col];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int row = (lane >> 4) * 4 + j;
        const float *ad = add + row * 4;
        float v = acc[j] + b1;
        if (MODE != ROW_FIN) v += ts0[__float_as_int(ad[2]) * H + col];
        if (MODE == ROW_BI) v += ts1[__float_as_int(ad[3]) * H + col];
        v += c0 * ad[0];
        v += c1 * ad[1];
        sH[row * C::HP + col] = st<T>(gelu_t<T>(v));
      }
    }
    wave_lds_sync();
    // --- layer 2
    T hfr[NK2][KL];
    const T *hsrc = sH + ri * C::HP + kq;
#pragma unroll
    for (int ks = 0; ks < NK2; ks++)
#pragma unroll
      for (int e = 0; e < KL; e++) hfr[ks][e] = hsrc[ks * KS + e];
#pragma unroll 1
    for (int cb = 0; cb < D / 16; cb++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const T *bw = gW2 + (cb * 16 + (lane & 15)) * P2 + kq;
#pragma unroll
      for (int ks = 0; ks < NK2; ks++) mfma_step<T>(acc, hfr[ks], bw + ks * KS);
      const int col = cb * 16 + (lane & 15);
      const float b2 = tb2[col];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int row = (lane >> 4) * 4 + j;
        sM[row * C::MP + col] = st<T>(acc[j] + b2);
      }
    }
    wave_lds_sync();
    // --- coalesced store of the 16 output rows
    {
      T *out = (T *)a.out;
      for (int e = lane; e < C::ROWS * D; e += 64) {
        const int row = e / D, col = e % D;
        const int64_t rr = r0 + row;
        if (rr >= n_rows) continue;
        int64_t dst;
        if constexpr (MODE == ROW_FIN) {
          const int64_t it = rr / a.K;
          dst = (a.qout[it] * a.K + rr % a.K) * D + col;
        } else {
          dst = (slot_base * a.K + rr) * D + col;
        }
        out[dst] = sM[row * C::MP + col];
      }
    }
    // --- attention key of each message row
    if constexpr (MODE != ROW_FIN) {
      T mfr[NK3][KL];
      const T *msrc = sM + ri * C::MP + kq;
#pragma unroll
      for (int ks = 0; ks < NK3; ks++)
#pragma unroll
        for (int e = 0; e < KL; e++) mfr[ks][e] = msrc[ks * KS + e];
      float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
      for (int cb = 0; cb < D / 16; cb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const T *bw = gWk + (cb * 16 + (lane & 15)) * PK + kq;
#pragma unroll
        for (int ks = 0; ks < NK3; ks++) mfma_step<T>(acc, mfr[ks], bw + ks * KS);
        const int col = cb * 16 + (lane & 15);
        const float bk1 = tbk1[col], wk2 = twk2[col];
#pragma unroll
        for (int j = 0; j < 4; j++) part[j] += wk2 * gelu_t<T>(acc[j] + bk1);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        float v = part[j];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        part[j] = v;
      }
      if ((lane & 15) == 0) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int64_t rr = r0 + (lane >> 4) * 4 + j;
          if (rr < n_rows) a.score[slot_base * a.K + rr] = (part[j] + pk.bk2) * a.score_scale;
        }
      }
    }
    wave_lds_sync();
    rc = rn;
    lc = ln;
  }
}

template <int D, typename T>
static size_t row_mlp_smem(int ns) {
  using C = RowCfg<D, T>;
  const size_t nw = C::BF ? (size_t)C::H * C::W1P + (size_t)D * C::W2P + (size_t)D * C::WKP : 0;
  const size_t tables = sizeof(float) * (3 * (size_t)C::H + 3 * (size_t)D + 2 * (size_t)ns * C::H);
  const size_t per_wave = sizeof(T) * (C::ROWS * C::HP + C::ROWS * C::MP) + 16 * 4 * 4;
  return sizeof(T) * nw + tables + 1024 + C::WAVES * per_wave;
}

// ---------------------------------------------------------------------------
// line pool: per (query, axis) softmax over live lines + null, value projection
// ---------------------------------------------------------------------------

// One wave per (m, k) item; lane f < D holds feature f.
template <int D, typename T>
__global__ __launch_bounds__(256) void line_pool_kernel(
    int64_t n_items, int K, int L, const int32_t *__restrict__ slot,
    const T *__restrict__ msg, const float *__restrict__ score, PoolW pw, float scale,
    T *__restrict__ ctx, int64_t item_base) {
  __shared__ float sWvT[D * D];
  __shared__ float sP[4][64];
  for (int i = threadIdx.x; i < D * D; i += 256) sWvT[i] = pw.wvt[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (int64_t)gridDim.x * 4;
  for (int64_t b0 = (int64_t)blockIdx.x * 4; b0 < n_items; b0 += per) {
    const int64_t it0 = b0 + wv;
    const bool act = it0 < n_items;
    const int64_t it = act ? it0 : n_items - 1;
    const int64_t m = it / K;
    const int k = (int)(it % K);
    const int32_t *sl = slot + m * L;
    const float sn = pw.null_k * scale;
    float mx = sn;
    for (int l = 0; l < L; l++) {
      const int32_t s = sl[l];
      if (s >= 0) mx = fmaxf(mx, score[(int64_t)s * K + k]);
    }
    float den = 0.f, acc = 0.f;
    for (int l = 0; l < L; l++) {
      const int32_t s = sl[l];
      if (s < 0) continue;
      const float e = expf(score[(int64_t)s * K + k] - mx);
      den += e;
      if (lane < D) acc += e * ld<T>(msg + ((int64_t)s * K + k) * D + lane);
    }
    const float en = expf(sn - mx);
    den += en;
    const float inv = 1.0f / den;
    const float wnull = en * inv;
    if (lane < D) sP[wv][lane] = acc * inv;
    __syncthreads();
    if (act && lane < D) {
      float o = 0.f;
      for (int j = 0; j < D; j++) o += sWvT[j * D + lane] * sP[wv][j];
      o += (1.0f - wnull) * pw.bv[lane];
      o += wnull * pw.null_v[lane];
      ctx[((item_base + m) * K + k) * D + lane] = st<T>(o);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// head: axis pool + PredHead, one wave per point (FP32)
// ---------------------------------------------------------------------------

struct HeadArgs {
  int64_t M;
  int K;
  float scale, log_eb, vmin, span, vmax;
  const float *w1t;   // (D, H) head.net.0 weight transposed (in, out), cond column apart
  const float *cond;  // (H) head.net.0 weight of the log-eb column
  const float *b1;    // (H)
  const float *w2;    // (2, H)
  const float *b2;    // (2)
};

template <int D, typename T>
__global__ __launch_bounds__(256) void head_kernel(const T *__restrict__ ctx, PoolW pw,
                                                   HeadArgs h, float *__restrict__ pred,
                                                   float *__restrict__ scl,
                                                   float *__restrict__ mu_out,
                                                   float *__restrict__ lb_out) {
  constexpr int H = 2 * D;
  __shared__ float sC[4][LATEC_MAX_DIM][D];
  __shared__ float sP[4][D];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (int64_t)gridDim.x * 4;
  for (int64_t m0 = (int64_t)blockIdx.x * 4; m0 < h.M; m0 += per) {
    const int64_t m = m0 + wv;
    const bool act = m < h.M;
    for (int k = 0; k < h.K; k++)
      if (lane < D) sC[wv][k][lane] = act ? ld<T>(ctx + (m * h.K + k) * D + lane) : 0.f;
    __syncthreads();
    float s[LATEC_MAX_DIM], v[LATEC_MAX_DIM];
    for (int k = 0; k < h.K; k++) {
      float hk = 0.f, vk = 0.f;
      if (lane < D) {
        for (int i = 0; i < D; i++) {
          const float c = sC[wv][k][i];
          hk += pw.wk1t[i * D + lane] * c;
          vk += pw.wvt[i * D + lane] * c;
        }
        hk = gelu(hk + pw.bk1[lane]) * pw.wk2[lane];
        vk += pw.bv[lane];
      }
      for (int o = 32; o >= 1; o >>= 1) hk += __shfl_xor(hk, o);
      s[k] = (hk + pw.bk2) * h.scale;
      v[k] = vk;
    }
    const float sn = pw.null_k * h.scale;
    float mx = sn;
    for (int k = 0; k < h.K; k++) mx = fmaxf(mx, s[k]);
    float den = 0.f, e[LATEC_MAX_DIM];
    for (int k = 0; k < h.K; k++) {
      e[k] = expf(s[k] - mx);
      den += e[k];
    }
    const float en = expf(sn - mx);
    den += en;
    float pooled = 0.f;
    if (lane < D) {
      for (int k = 0; k < h.K; k++) pooled += (e[k] / den) * v[k];
      pooled += (en / den) * pw.null_v[lane];
      sP[wv][lane] = pooled;
    }
    __syncthreads();
    float o0 = 0.f, o1 = 0.f;
    for (int jj = lane; jj < H; jj += 64) {
      float z = 0.f;
      for (int i = 0; i < D; i++) z += h.w1t[i * H + jj] * sP[wv][i];
      z += h.cond[jj] * h.log_eb;
      z = gelu(z + h.b1[jj]);
      o0 += h.w2[jj] * z;
      o1 += h.w2[H + jj] * z;
    }
    for (int o = 32; o >= 1; o >>= 1) {
      o0 += __shfl_xor(o0, o);
      o1 += __shfl_xor(o1, o);
    }
    if (act && lane == 0) {
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
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// value embedding (InitEmbed of the clamped, normalized recon value)
// ---------------------------------------------------------------------------

template <int D, typename T>
__global__ __launch_bounds__(256) void value_embed_kernel(
    int64_t M, const float *__restrict__ recon, const int64_t *__restrict__ gidx,
    float vmin, float vmax, float span, const float *__restrict__ w0,
    const float *__restrict__ b0, const float *__restrict__ w2t,
    const float *__restrict__ b2, T *__restrict__ out) {
  __shared__ float sh[4][D];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (int64_t)gridDim.x * 4;
  for (int64_t m0 = (int64_t)blockIdx.x * 4; m0 < M; m0 += per) {
    const int64_t m = m0 + wv;
    const bool act = m < M;
    float val = 0.f;
    if (act) {
      const float r = recon[gidx[m]];
      val = (fminf(fmaxf(r, vmin), vmax) - vmin) / span;
    }
    if (lane < D) sh[wv][lane] = gelu(w0[lane] * val + b0[lane]);
    __syncthreads();
    if (act && lane < D) {
      float o = 0.f;
      for (int j = 0; j < D; j++) o += w2t[j * D + lane] * sh[wv][j];
      out[m * D + lane] = st<T>(o + b2[lane]);
    }
    __syncthreads();
  }
}

// halo cells: E[row, k, :] = InitEmbed(value) for every axis k
template <int D, typename T>
__global__ __launch_bounds__(256) void halo_embed_kernel(
    int64_t M, int K, const float *__restrict__ recon, const int64_t *__restrict__ gidx,
    const int64_t *__restrict__ rows, float vmin, float vmax, float span,
    const float *__restrict__ w0, const float *__restrict__ b0,
    const float *__restrict__ w2t, const float *__restrict__ b2, T *__restrict__ E) {
  __shared__ float sh[4][D];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t per = (int64_t)gridDim.x * 4;
  for (int64_t m0 = (int64_t)blockIdx.x * 4; m0 < M; m0 += per) {
    const int64_t m = m0 + wv;
    const bool act = m < M;
    float val = 0.f;
    if (act) {
      const float r = recon[gidx[m]];
      val = (fminf(fmaxf(r, vmin), vmax) - vmin) / span;
    }
    if (lane < D) sh[wv][lane] = gelu(w0[lane] * val + b0[lane]);
    __syncthreads();
    if (act && lane < D) {
      float o = 0.f;
      for (int j = 0; j < D; j++) o += w2t[j * D + lane] * sh[wv][j];
      const T v = st<T>(o + b2[lane]);
      for (int k = 0; k < K; k++) E[(rows[m] * K + k) * D + lane] = v;
    }
    __syncthreads();
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
  const double *grid;          // 64 grid scales
  double bound[12];            // eb * 2^T, T = 2..13
  double *partial;             // (blocks, n_cand, 13)
};

__device__ __forceinline__ double gate_bits(double absr, double b, const GateArgs &g) {
  const double k = rint(absr / g.two_eb);
  double p;
  if (k == 0.0)
    p = -expm1(-g.eb / b);
  else
    p = 0.5 * exp(-((2.0 * k - 1.0) * g.eb / b)) * -expm1(-2.0 * g.eb / b);
  double bits = -log2(p);
  if (bits > 24.0) bits = 24.0;
  if (k >= (double)g.radius) bits = 56.0;
  return bits;
}

__device__ __forceinline__ double snap_scale(float b, const GateArgs &g) {
  int lo = 0, hi = 63;  // bucketize(right=True): count of thresholds <= b
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (g.thr[mid] <= b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return g.grid[lo];
}

constexpr int GATE_BLOCK = 256;
constexpr int GATE_NB = 13;

// partial[b][c][bucket]: sums over the points of block b (fixed tree order).
__global__ __launch_bounds__(GATE_BLOCK) void gate_cost_kernel(GateArgs g) {
  __shared__ double red[GATE_BLOCK];
  const int64_t per_block = (g.n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per_block;
  const int64_t hi = lo + per_block < g.n ? lo + per_block : g.n;
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
      const double bits = gate_bits(fabs((double)g.x[i] - (double)ref[i]), snap_scale(bs, g), g);
#pragma unroll
      for (int bb = 0; bb < GATE_NB; bb++)
        if (bb == bucket) acc[bb] += bits;
    }
    for (int b = 0; b < GATE_NB; b++) {
      red[threadIdx.x] = acc[b];
      __syncthreads();
      for (int o = GATE_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
      }
      if (threadIdx.x == 0) g.partial[((int64_t)blockIdx.x * g.n_cand + c) * GATE_NB + b] = red[0];
      __syncthreads();
    }
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

// Packs the model for the device: GEMM operands in the storage type, state
// rows folded into bias rows, pool matrices transposed, all in one block.
int make_device(latec_gnn *g, const float *P) {
  const Layout &Ly = g->lay;
  const int d = Ly.d, h = Ly.h, de = Ly.de, ds = Ly.ds, ns = Ly.ns;
  const bool bf = g->numerics == NUM_BF16;
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
  // message layer 1 (bidir): [E_neg | E_pos] columns
  const int bi_in = 2 * de + 2;
  std::vector<float> w_bi = cat2(rows(Ly.bi0w, h, bi_in, 0, d), d, rows(Ly.bi0w, h, bi_in, de, d), d, h);
  void *o_bi_w1 = mat(w_bi);
  void *o_bi_w2 = mat(rows(Ly.bi2w, d, h, 0, h));
  const int dir_in = de + 2;
  void *o_dir_w1 = mat(rows(Ly.dir0w, h, dir_in, 0, d));
  void *o_dir_w2 = mat(rows(Ly.dir2w, d, h, 0, h));
  void *o_mx_w1 = mat(rows(Ly.mx0w, h, 2 * d, 0, 2 * d));
  void *o_mx_w2 = mat(rows(Ly.mx2w, d, h, 0, h));
  void *o_lk_w1 = mat(rows(Ly.lp[0], d, d, 0, d));
  // f32 arrays
  std::vector<std::vector<float>> f(F_NSLOTS);
  auto state_rows = [&](int64_t wbase, int n_in, int col0) {
    std::vector<float> v((size_t)ns * h);
    for (int s = 0; s < ns; s++)
      for (int o = 0; o < h; o++) {
        double acc = 0.0;
        for (int j = 0; j < ds; j++)
          acc += (double)P[wbase + (int64_t)o * n_in + col0 + j] * (double)P[Ly.table + s * ds + j];
        v[(size_t)s * h + o] = (float)acc;
      }
    return v;
  };
  auto vec = [&](int64_t base, int n) { return std::vector<float>(P + base, P + base + n); };
  auto col = [&](int64_t wbase, int n_out, int n_in, int c) {
    std::vector<float> v(n_out);
    for (int o = 0; o < n_out; o++) v[o] = P[wbase + (int64_t)o * n_in + c];
    return v;
  };
  auto transpose = [&](int64_t base, int n_out, int n_in, int in_used) {
    std::vector<float> v((size_t)in_used * n_out);
    for (int o = 0; o < n_out; o++)
      for (int i = 0; i < in_used; i++) v[(size_t)i * n_out + o] = P[base + (int64_t)o * n_in + i];
    return v;
  };
  f[F_BI_B1] = vec(Ly.bi0b, h);
  f[F_BI_B2] = vec(Ly.bi2b, d);
  f[F_BI_S0] = state_rows(Ly.bi0w, bi_in, d);
  f[F_BI_S1] = state_rows(Ly.bi0w, bi_in, de + d);
  f[F_BI_C0] = col(Ly.bi0w, h, bi_in, 2 * de);
  f[F_BI_C1] = col(Ly.bi0w, h, bi_in, 2 * de + 1);
  f[F_DIR_B1] = vec(Ly.dir0b, h);
  f[F_DIR_B2] = vec(Ly.dir2b, d);
  f[F_DIR_S] = state_rows(Ly.dir0w, dir_in, d);
  f[F_DIR_CSIGN] = col(Ly.dir0w, h, dir_in, de);
  f[F_DIR_CLOGD] = col(Ly.dir0w, h, dir_in, de + 1);
  f[F_MX_B1] = vec(Ly.mx0b, h);
  f[F_MX_B2] = vec(Ly.mx2b, d);
  f[F_LK_B1] = vec(Ly.lp[1], d);
  f[F_LK_W2] = vec(Ly.lp[2], d);
  f[F_LK_WVT] = transpose(Ly.lp[4], d, d, d);
  f[F_LK_BV] = vec(Ly.lp[5], d);
  f[F_LK_NULLV] = vec(Ly.lp[7], d);
  f[F_LK_W1T] = transpose(Ly.lp[0], d, d, d);
  f[F_AP_W1T] = transpose(Ly.ap[0], d, d, d);
  f[F_AP_B1] = vec(Ly.ap[1], d);
  f[F_AP_W2] = vec(Ly.ap[2], d);
  f[F_AP_WVT] = transpose(Ly.ap[4], d, d, d);
  f[F_AP_BV] = vec(Ly.ap[5], d);
  f[F_AP_NULLV] = vec(Ly.ap[7], d);
  f[F_HD_W1T] = transpose(Ly.hd0w, h, d + 1, d);
  f[F_HD_COND] = col(Ly.hd0w, h, d + 1, d);
  f[F_HD_B1] = vec(Ly.hd0b, h);
  f[F_HD_W2] = vec(Ly.hd2w, 2 * h);
  f[F_HD_B2] = vec(Ly.hd2b, 2);
  f[F_IN_W0] = vec(Ly.init0w, d);
  f[F_IN_B0] = vec(Ly.init0b, d);
  f[F_IN_W2T] = transpose(Ly.init2w, d, d, d);
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
  g->lk_w1 = at(o_lk_w1);
  g->f32 = (float *)at(o_f32);
  g->d_dirs = (int8_t *)at(o_dirs);
  g->d_sidx = (int8_t *)at(o_sidx);
  g->d_lognnz = (float *)at(o_lnz);
  return LATEC_OK;
}

const float *fs(const latec_gnn *g, int slot) { return g->f32 + g->f32_off[slot]; }

PoolW pool_w(const latec_gnn *g, bool line) {
  PoolW p{};
  if (line) {
    p.wk1t = fs(g, F_LK_W1T);
    p.bk1 = fs(g, F_LK_B1);
    p.wk2 = fs(g, F_LK_W2);
    p.wvt = fs(g, F_LK_WVT);
    p.bv = fs(g, F_LK_BV);
    p.null_v = fs(g, F_LK_NULLV);
    p.bk2 = g->scal[0];
    p.null_k = g->scal[1];
  } else {
    p.wk1t = fs(g, F_AP_W1T);
    p.bk1 = fs(g, F_AP_B1);
    p.wk2 = fs(g, F_AP_W2);
    p.wvt = fs(g, F_AP_WVT);
    p.bv = fs(g, F_AP_BV);
    p.null_v = fs(g, F_AP_NULLV);
    p.bk2 = g->scal[2];
    p.null_k = g->scal[3];
  }
  return p;
}

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

template <int D, typename T>
int launch_rows(latec_gnn *g, int mode, RowArgs &a, hipStream_t s) {
  MlpW<T> w{};
  KeyW<T> pk{(const T *)g->lk_w1, fs(g, F_LK_B1), fs(g, F_LK_W2), g->scal[0]};
  if (mode == ROW_BI) {
    w = {(const T *)g->bi_w1, (const T *)g->bi_w2, fs(g, F_BI_B1), fs(g, F_BI_B2),
         fs(g, F_BI_S0), fs(g, F_BI_S1), fs(g, F_BI_C0), fs(g, F_BI_C1), 2 * D};
  } else if (mode == ROW_DIR) {
    w = {(const T *)g->dir_w1, (const T *)g->dir_w2, fs(g, F_DIR_B1), fs(g, F_DIR_B2),
         fs(g, F_DIR_S), nullptr, fs(g, F_DIR_CSIGN), fs(g, F_DIR_CLOGD), D};
  } else {
    w = {(const T *)g->mx_w1, (const T *)g->mx_w2, fs(g, F_MX_B1), fs(g, F_MX_B2),
         nullptr, nullptr, nullptr, nullptr, 2 * D};
  }
  a.ns = g->lay.ns;
  a.L = g->L;
  a.sidx = g->d_sidx;
  const size_t smem = row_mlp_smem<D, T>(a.ns);
  const int grid = g->n_cu;
  const void *fn = mode == ROW_BI    ? (const void *)row_mlp_kernel<D, T, ROW_BI>
                   : mode == ROW_DIR ? (const void *)row_mlp_kernel<D, T, ROW_DIR>
                                     : (const void *)row_mlp_kernel<D, T, ROW_FIN>;
  int rc = hip_ok(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem),
                  "hipFuncSetAttribute(row_mlp_kernel)");
  if (rc) return rc;
  using C = RowCfg<D, T>;
  if (mode == ROW_BI)
    hipLaunchKernelGGL((row_mlp_kernel<D, T, ROW_BI>), dim3(grid), dim3(C::THREADS), smem, s, a,
                       w, pk);
  else if (mode == ROW_DIR)
    hipLaunchKernelGGL((row_mlp_kernel<D, T, ROW_DIR>), dim3(grid), dim3(C::THREADS), smem, s, a,
                       w, pk);
  else
    hipLaunchKernelGGL((row_mlp_kernel<D, T, ROW_FIN>), dim3(grid), dim3(C::THREADS), smem, s, a,
                       w, pk);
  return hip_ok(hipGetLastError(), "row_mlp_kernel launch");
}

}  // namespace

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
  const PoolW pw = pool_w(g, true);
  const float scale = 1.0f / std::sqrt((float)D);
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
    int grid = (int)((items + 3) / 4);
    if (grid > 8 * g->n_cu) grid = 8 * g->n_cu;
    hipLaunchKernelGGL((line_pool_kernel<D, T>), dim3(grid), dim3(256), 0, s, items, K, L,
                       slot, msg, score, pw, scale, (T *)ctx, m0);
    rc = hip_ok(hipGetLastError(), "line_pool_kernel launch");
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
  h.w1t = fs(g, F_HD_W1T);
  h.cond = fs(g, F_HD_COND);
  h.b1 = fs(g, F_HD_B1);
  h.w2 = fs(g, F_HD_W2);
  h.b2 = fs(g, F_HD_B2);
  int grid = (int)((M + 3) / 4);
  if (grid > 8 * g->n_cu) grid = 8 * g->n_cu;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((head_kernel<D, T>), dim3(grid), dim3(256), 0, s, (const T *)ctx,
                     pool_w(g, false), h, pred, scale, mu, logb);
  return hip_ok(hipGetLastError(), "head_kernel launch");
}

template <int D, typename T>
int finalize_t(latec_gnn *g, void *E, const void *ctx, const int64_t *qidx,
               const int64_t *gidx, const float *recon, int64_t M, float vmin, float vmax,
               hipStream_t s) {
  int rc = ensure_work(g, (size_t)M * D * sizeof(T) + 256);
  if (rc) return rc;
  T *vemb = (T *)g->work;
  int grid = (int)((M + 3) / 4);
  if (grid > 8 * g->n_cu) grid = 8 * g->n_cu;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((value_embed_kernel<D, T>), dim3(grid), dim3(256), 0, s, M, recon, gidx,
                     vmin, vmax, vmax - vmin, fs(g, F_IN_W0), fs(g, F_IN_B0), fs(g, F_IN_W2T),
                     fs(g, F_IN_B2), vemb);
  rc = hip_ok(hipGetLastError(), "value_embed_kernel launch");
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
  int grid = (int)((M + 3) / 4);
  if (grid > 8 * g->n_cu) grid = 8 * g->n_cu;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((halo_embed_kernel<D, T>), dim3(grid), dim3(256), 0, s, M, g->ndim,
                     recon, gidx, rows, vmin, vmax, vmax - vmin, fs(g, F_IN_W0),
                     fs(g, F_IN_B0), fs(g, F_IN_W2T), fs(g, F_IN_B2), (T *)E);
  return hip_ok(hipGetLastError(), "halo_embed_kernel launch");
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
  // classes of the stage and their start indices, uploaded once per call
  const int nc = (int)st.classes.size();
  a.n_classes = nc;
  std::vector<int64_t> starts(nc + 1, 0);
  for (int c = 0; c < nc; c++) starts[c + 1] = starts[c] + st.classes[c].size;
  const size_t bcls = ((sizeof(ClassDesc) * nc + 255) / 256) * 256;
  int rc = ensure_work(g, bcls + sizeof(int64_t) * (nc + 1));
  if (rc) return rc;
  ClassDesc *dcls = (ClassDesc *)g->work;
  int64_t *dstart = (int64_t *)((unsigned char *)g->work + bcls);
  rc = hip_ok(hipMemcpyAsync(dcls, st.classes.data(), sizeof(ClassDesc) * nc,
                             hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipMemcpyAsync(dstart, starts.data(), sizeof(int64_t) * (nc + 1),
                                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  if (rc) return rc;
  const int64_t total = starts[nc];
  hipLaunchKernelGGL(geometry_kernel, dim3(grid_for(total, 256)), dim3(256), 0, s, dcls, a,
                     (const int8_t *)g->d_dirs, total, dstart, d_qidx, d_gidx, d_steps);
  rc = hip_ok(hipGetLastError(), "geometry_kernel launch");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors die here
  return rc;
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
  int blocks = (int)((n + 4095) / 4096);
  if (blocks > 1024) blocks = 1024;
  const size_t bpart = sizeof(double) * (size_t)blocks * nc * GATE_NB;
  int rc = ensure_work(g, bpart + 1024);
  if (rc) return rc;
  double *part = (double *)g->work;
  float *dthr = (float *)((unsigned char *)g->work + ((bpart + 255) & ~(size_t)255));
  double *dgrid = (double *)((unsigned char *)dthr + 256);
  rc = hip_ok(hipMemcpyAsync(dthr, thr63, sizeof(float) * 63, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipMemcpyAsync(dgrid, grid64, sizeof(double) * 64,
                                      hipMemcpyHostToDevice, s), "hipMemcpyAsync");
  if (rc) return rc;
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
  a.grid = dgrid;
  for (int t = 0; t < 12; t++) a.bound[t] = eb * std::ldexp(1.0, t + 2);
  a.partial = part;
  hipLaunchKernelGGL(gate_cost_kernel, dim3(blocks), dim3(GATE_BLOCK), 0, s, a);
  rc = hip_ok(hipGetLastError(), "gate_cost_kernel launch");
  if (rc) return rc;
  std::vector<double> h((size_t)blocks * nc * GATE_NB);
  rc = hip_ok(hipMemcpyAsync(h.data(), part, bpart, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  if (!rc) rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (rc) return rc;
  for (int b = 0; b < blocks; b++)
    for (int i = 0; i < nc * GATE_NB; i++) costs[i] += h[(size_t)b * nc * GATE_NB + i];
  return LATEC_OK;
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
