// latec_plan.cpp -- closed-form stage schedule (host side of the plan).
//
// Follows latec/levels.py: _check_plan :54-76 (validation and messages),
// iter_stage_plan :100-147 (stage order), stage_counts :221-282,
// _fused_classes / stage_layouts :469-574 (parity classes and their affine
// rank into the flat-C-order stage array).  anchor_block > 1 removes the
// anchor points from every dyadic stage (the `mask &= ~covered` step); the
// removed points are kept as sorted block-1 ranks so a kernel can rebase a
// point's rank by a binary search.
#include <algorithm>
#include <cstring>
#include <string>

#include "latec_plan.h"

namespace latec {

static int64_t slice_count(int64_t n, int64_t start, int64_t step) {
  return n <= start ? 0 : (n - start + step - 1) / step;
}

void fastdiv_magic(uint32_t d, uint32_t *magic, uint32_t *shift) {
  // q = (umulhi(n, magic) + n) >> shift, exact for all 32-bit n (d >= 1)
  uint32_t l = 0;
  while (((uint64_t)1 << l) < d) l++;
  *shift = l;
  *magic = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
}

static ClassDesc make_class(int ndim, int parity, const int64_t *starts,
                            int64_t step, const int64_t *sizes,
                            const int64_t *rstr, int64_t offset) {
  ClassDesc c;
  memset(&c, 0, sizeof c);
  c.parity = parity;
  c.step = (int32_t)step;
  c.ndim = ndim;
  c.offset = offset;
  c.size = 1;
  for (int j = 0; j < ndim; j++) {
    c.starts[j] = starts[j];
    c.sizes[j] = (uint32_t)sizes[j];
    c.rstr[j] = rstr[j];
    c.size *= sizes[j];
    fastdiv_magic(sizes[j] ? (uint32_t)sizes[j] : 1u, &c.magic[j], &c.shift[j]);
  }
  c.row_len = (uint32_t)sizes[ndim - 1];
  c.nrows = c.row_len ? (uint32_t)(c.size / c.row_len) : 0;
  c.tr = 1;
  while (c.tr < 256 && (uint32_t)c.tr < c.row_len) c.tr <<= 1;
  return c;
}

static ClassDesc single_class(int ndim, int parity, const int64_t *starts,
                              int64_t step, const int64_t *sizes) {
  int64_t rstr[LATEC_MAX_DIM];
  rstr[ndim - 1] = 1;
  for (int j = ndim - 2; j >= 0; j--) rstr[j] = rstr[j + 1] * sizes[j + 1];
  return make_class(ndim, parity, starts, step, sizes, rstr, 0);
}

// Every weight-w parity class of a fused sub-stage at stride s
// (levels.py:469-512 with low_axes = ()): F[j][r] counts stage points in axes
// j.. with r odd axes left to place, G[j][r] those for one fixed lattice
// index on axis j.
static std::vector<ClassDesc> fused_classes(int ndim, const int64_t *shape,
                                            int64_t s, int w, int64_t *total) {
  int64_t cnt[LATEC_MAX_DIM][2];
  for (int j = 0; j < ndim; j++) {
    cnt[j][0] = slice_count(shape[j], 0, 2 * s);
    cnt[j][1] = slice_count(shape[j], s, 2 * s);
  }
  std::vector<std::vector<int64_t>> F(ndim + 1, std::vector<int64_t>(w + 1, 0)),
      G(ndim + 1, std::vector<int64_t>(w + 1, 0));
  F[ndim][0] = 1;
  for (int j = ndim - 1; j >= 0; j--)
    for (int r = 0; r <= w; r++) {
      int64_t odd = r ? F[j + 1][r - 1] : 0;
      F[j][r] = cnt[j][0] * F[j + 1][r] + cnt[j][1] * odd;
      G[j][r] = F[j + 1][r] + odd;
    }
  std::vector<ClassDesc> out;
  // combinations of w axes in lexicographic order (itertools.combinations)
  std::vector<int> comb(w);
  for (int i = 0; i < w; i++) comb[i] = i;
  while (true) {
    int parity = 0;
    for (int i = 0; i < w; i++) parity |= 1 << comb[i];
    int64_t starts[LATEC_MAX_DIM], sizes[LATEC_MAX_DIM], rstr[LATEC_MAX_DIM];
    int64_t offset = 0;
    int r = w;
    for (int j = 0; j < ndim; j++) {
      int p = (parity >> j) & 1;
      rstr[j] = G[j][r];
      starts[j] = p ? s : 0;
      sizes[j] = cnt[j][p];
      if (p) offset += F[j + 1][r];
      r -= p;
    }
    out.push_back(make_class(ndim, parity, starts, 2 * s, sizes, rstr, offset));
    int i = w - 1;
    while (i >= 0 && comb[i] == ndim - w + i) i--;
    if (i < 0) break;
    comb[i]++;
    for (int k = i + 1; k < w; k++) comb[k] = comb[k - 1] + 1;
  }
  *total = F[0][w];
  return out;
}

// Block-1 ranks of the anchor points (all coords % stride < block) lying in
// the classes of one stage, sorted.
static std::vector<uint32_t> anchor_ranks(const latec_plan *p,
                                          const StageDesc &st) {
  std::vector<uint32_t> out;
  for (const ClassDesc &c : st.classes) {
    std::vector<std::vector<int64_t>> ms(p->ndim);
    bool empty = false;
    for (int j = 0; j < p->ndim; j++) {
      for (int64_t m = 0; m < c.sizes[j]; m++) {
        int64_t coord = c.starts[j] + (int64_t)c.step * m;
        if (coord % p->stride < p->block) ms[j].push_back(m);
      }
      if (ms[j].empty()) empty = true;
    }
    if (empty) continue;
    std::vector<size_t> it(p->ndim, 0);
    while (true) {
      int64_t rank = c.offset;
      for (int j = 0; j < p->ndim; j++) rank += ms[j][it[j]] * c.rstr[j];
      out.push_back((uint32_t)rank);
      int j = p->ndim - 1;
      while (j >= 0 && ++it[j] == ms[j].size()) it[j--] = 0;
      if (j < 0) break;
    }
  }
  std::sort(out.begin(), out.end());
  return out;
}

int plan_build(latec_plan *p) {
  int nd = p->ndim;
  if (nd < 1 || nd > LATEC_MAX_DIM) {
    set_error(nd < 1 ? "shape must have at least one axis"
                     : "at most 8 spatial axes are supported");
    return LATEC_E_ARG;
  }
  if (p->levels < 1) {
    set_error("levels must be >= 1");
    return LATEC_E_ARG;
  }
  if (p->stride < 2 || (p->stride & (p->stride - 1))) {
    set_error("anchor_stride must be a power of two >= 2");
    return LATEC_E_ARG;
  }
  if (p->block < 1 || p->block > p->stride) {
    set_error("anchor_block must be in [1, anchor_stride]");
    return LATEC_E_ARG;
  }
  if (p->levels < 62 && ((int64_t)1 << p->levels) < p->stride) {
    int lg = 0;
    while ((1 << (lg + 1)) <= p->stride) lg++;
    set_error("levels=" + std::to_string(p->levels) + " too small for anchor_stride=" +
              std::to_string(p->stride) + ": need levels >= log2(anchor_stride) = " +
              std::to_string(lg) + " to densify to stride 1");
    return LATEC_E_ARG;
  }
  int64_t N = 1;
  for (int j = nd - 1; j >= 0; j--) {
    if (p->shape[j] < 0 || p->shape[j] > 0xFFFFFFFFll) {
      set_error("axis extent out of range");
      return LATEC_E_ARG;
    }
    p->gstr[j] = N;
    N *= p->shape[j];
  }
  p->npoints = N;
  p->stages.clear();
  // anchors
  {
    StageDesc st;
    st.stride = p->stride;
    st.weight = 0;
    int64_t starts[LATEC_MAX_DIM] = {0}, sizes[LATEC_MAX_DIM];
    for (int j = 0; j < nd; j++) {
      if (p->block == 1) {
        sizes[j] = slice_count(p->shape[j], 0, p->stride);
      } else {
        int64_t full = p->shape[j] / p->stride, rem = p->shape[j] % p->stride;
        sizes[j] = full * p->block + std::min<int64_t>(rem, p->block);
      }
    }
    ClassDesc c = single_class(nd, 0, starts, p->stride, sizes);
    if (p->block > 1) c.flags = 1;
    st.classes.push_back(c);
    st.count = st.count1 = c.size;
    p->stages.push_back(st);
  }
  int64_t prev_s = -1;
  for (int k = 1; k <= p->levels; k++) {
    int64_t s = std::max<int64_t>((int64_t)p->stride >> std::min(k, 62), 1);
    bool repeat = s == prev_s;
    prev_s = s;
    for (int axis = 0; axis < nd; axis++) {
      StageDesc st;
      st.stride = (int)s;
      st.weight = 1;
      if (!repeat) {
        int64_t starts[LATEC_MAX_DIM], sizes[LATEC_MAX_DIM];
        for (int j = 0; j < nd; j++) {
          starts[j] = j == axis ? s : 0;
          sizes[j] = slice_count(p->shape[j], starts[j], 2 * s);
        }
        st.classes.push_back(single_class(nd, 1 << axis, starts, 2 * s, sizes));
        st.count1 = st.classes[0].size;
        if (st.count1 == 0) st.classes.clear();
      }
      p->stages.push_back(st);
    }
    for (int w = 2; w <= nd; w++) {
      StageDesc st;
      st.stride = (int)s;
      st.weight = w;
      if (!repeat) {
        int64_t total = 0;
        st.classes = fused_classes(nd, p->shape, s, w, &total);
        st.count1 = total;
        std::vector<ClassDesc> keep;
        for (auto &c : st.classes)
          if (c.size) keep.push_back(c);
        st.classes = keep;
      }
      p->stages.push_back(st);
    }
  }
  int64_t off = 0;
  for (auto &st : p->stages) {
    if (st.count1 > 0xFFFFFFFFll) {
      set_error("stage too large for 32-bit ranks");
      return LATEC_E_ARG;
    }
    if (p->block > 1 && st.weight > 0 && st.count1) st.excluded = anchor_ranks(p, st);
    if (st.weight > 0) st.count = st.count1 - (int64_t)st.excluded.size();
    st.offset = off;
    off += st.count;
  }
  if (off != N) {
    set_error("internal: stage counts do not partition the grid");
    return LATEC_E_ARG;
  }
  return LATEC_OK;
}

}  // namespace latec
