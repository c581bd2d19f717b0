// latec_sheet.hip -- sheet kernels: a 3-D level above the first in two launches
// (gfx950).
//
// The level's compact grid (n0 x n1 x n2, stride 1; coarse points at even
// coordinates on every axis, known from the level below) has seven parity
// classes, coded in the reference's sub-stage order 100 | 010 | 001 |
// {011, 101, 110} | 111 (latec/levels.py:410-574; a class's bits name its odd
// axes 0, 1, 2).  Split by the parity of i (axis 0):
//
//   even sheets (i even): the coarse copy, 001 (taps along k: coarse points),
//     010 (along j: coarse points), 011 (along j: 001; along k: 010);
//   odd sheets (i odd): 100 (along i: coarse points), 101 (along i: 001;
//     along k: 100), 110 (along i: 010; along j: 100), 111 (along i: 011;
//     along j: 101; along k: 110).
//
// Every tap of an even sheet lies on the same sheet, and every tap of an odd
// sheet lies on the even sheets or on the same sheet, so the level is two
// launches: the even sheets, then the odd sheets.  A lane owns one pair
// (k = 2m, 2m + 1) of a line (a row j of a sheet) and marches along the rows
// of a segment.  Step t codes even row E = t + 2 (j = 2E) and odd row
// J = t - 1 (j = 2J + 1).  The odd row's taps along j are even rows J - 1 ..
// J + 2, which the lane computed in earlier steps and keeps in registers; the
// taps along k are the even points of pairs m - 1, m + 1, m + 2 of the same
// row, which are the neighbouring lanes' (ds_bpermute); the taps along i (odd
// sheets) are four loads of the even sheets.  So within a step the two
// points before the exchange (even sheets: 010 of row J; odd sheets: 100 of
// row E and 110 of row J) are independent of each other, and so are the two
// after it (001 and 011; 101 and 111): each pair is quantized with one wave
// vote.
//
// A wave owns a run of 61 consecutive pairs of the concatenated lines of the
// segment's sheets (lanes 1 .. 61); lanes 0, 62 and 63 compute the even
// points of the neighbouring pairs (same inputs, same arithmetic, same bits)
// and store nothing, so no wave waits for another.  Where a run crosses a
// line end the neighbouring lanes belong to the next sheet, which is exactly
// where the end rules leave those taps unused.  A segment of `seg` odd rows
// starts four steps early to fill the row windows; those rows are recomputed
// and stored by the segment that owns them.
//
// Addresses are a per-lane buffer offset (sheet, pair) fixed for the march
// plus a uniform offset per row (buffer instructions' scalar offset), so a
// step spends no vector instructions on addressing.  Lanes that own nothing
// store to an offset past the buffer's range, which the hardware drops.
//
// The arithmetic is interp4 / cubic4 / pair_combine and the stage quantizer
// of the other level kernels, operation for operation, so codes, outliers
// and the reconstruction are bit-identical to theirs.  Cubic interpolation
// only (the centre rule is a template argument); a linear level runs as pair
// launches.
#include <type_traits>

#include "latec_common.h"
#include "latec_sheet.h"

namespace latec {

#define SHEET_OWN 61  // pairs a wave owns (lanes 1 .. 61)

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), inlined
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

typedef unsigned int sh_u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bld32(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ __forceinline__ float2 bld64(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  const sh_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0);
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ __forceinline__ uint32_t bld16(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
}
__device__ __forceinline__ void bst64(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so, float a,
                                      float b) {
  const sh_u32x2 v = {__float_as_uint(a), __float_as_uint(b)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, vo, so, 0);
}
__device__ __forceinline__ void bst16(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so,
                                      uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b16((uint16_t)v, r, vo, so, 0);
}
// the value of v in the lane whose byte address (4 * lane) is `addr`
__device__ __forceinline__ float lane_get(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// One stage's quantizer constants (the radius, and with it the range bound
// and the fallback margin, is the same for every stage of a level: QR).
struct SQ {
  double w, inv_w;
  float feb;
};
struct QR {
  double Rd;   // R
  double hlo;  // 0.5 - 1e-12 (R + 1), as PQ::hlo
  uint32_t R;
};

// The quantizer's tail (quantize_one after the rounded quotient qd): range,
// candidate reconstruction, the float32 bound check (and the output rounding
// check when RND).
template <bool RND>
__device__ __forceinline__ void quant_tail(const SQ &Q, const QR &G, int round_mode, double span,
                                           double offset, double pd, float xv, double qd,
                                           uint32_t &code, float &r) {
  const bool inr = fabs(qd) < G.Rd;
  const int qi = (int)qd;  // used only when in range
  // qd + 0.0: the integer's double (+0 for a -0 quotient), as (double)qi
  const float cand = (float)(pd + Q.w * (qd + 0.0));
  float chk = cand;
  if (RND) {
    if (round_mode == 1)
      chk = rintf(cand);
    else if (round_mode == 2)
      chk = (float)((rint((double)cand * span + offset) - offset) / span);
  }
  const bool ok = inr && fabsf(xv - chk) <= Q.feb;
  code = ok ? (uint32_t)(qi + (int)G.R) : 0u;
  r = ok ? cand : xv;
}

// quantize_one: q = rint(fl((x - p) / w)), through the product with 1/w and
// the division itself when the product lies within the margin of a rounding
// boundary (a wave vote).
template <bool RND>
__device__ __forceinline__ void quant1(const SQ &Q, const QR &G, int round_mode, double span,
                                       double offset, double pd, float xv, uint32_t &code,
                                       float &r) {
  const double d = (double)xv - pd;
  const double t = d * Q.inv_w;
  double q = rint(t);
  const bool f = !(fabs(t - q) < G.hlo);
  if (__any(f)) {
    if (f) q = rint(d / Q.w);
  }
  quant_tail<RND>(Q, G, round_mode, span, offset, pd, xv, q, code, r);
}

// Two points of two stages quantized with one wave vote for the fallback.
template <bool RND>
__device__ __forceinline__ void quant2(const SQ &Qa, const SQ &Qb, const QR &G, int round_mode,
                                       double span, double offset, double pa, float xa,
                                       double pb, float xb, uint32_t &ca, float &ra,
                                       uint32_t &cb, float &rb) {
  const double da = (double)xa - pa, db = (double)xb - pb;
  const double ta = da * Qa.inv_w, tb = db * Qb.inv_w;
  double qa = rint(ta), qb = rint(tb);
  const bool fa = !(fabs(ta - qa) < G.hlo), fb = !(fabs(tb - qb) < G.hlo);
  if (__any(fa || fb)) {
    if (fa) qa = rint(da / Qa.w);
    if (fb) qb = rint(db / Qb.w);
  }
  quant_tail<RND>(Qa, G, round_mode, span, offset, pa, xa, qa, ca, ra);
  quant_tail<RND>(Qb, G, round_mode, span, offset, pb, xb, qb, cb, rb);
}

// Decode: reconstruction from a code (pair_dequant's arithmetic); code 0
// looks the value up among the stage's outliers (sorted by rank), a missing
// one gives NaN.
__device__ __forceinline__ float dequant(const SQ &Q, const QR &G, double pd, uint32_t code,
                                         bool active, uint32_t rank,
                                         const uint32_t *__restrict__ orank,
                                         const float *__restrict__ oval, int64_t nout) {
  float r = (float)(pd + Q.w * (double)((int)code - (int)G.R));
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
  return r;
}

// Store a code (vo: the lane's offset, or past the range when the lane owns
// nothing) and count it in the workgroup's LDS window (the window's spare
// word HIST_WIN takes the counts of other lanes); true when the point also
// needs the rare path (a code outside the window: global count; an outlier).
__device__ __forceinline__ bool sheet_emit(__amdgpu_buffer_rsrc_t rc, uint32_t vo, uint32_t so,
                                           uint32_t code, uint32_t *lh, uint32_t wlo,
                                           bool active) {
  bst16(rc, vo, so, code);
  const uint32_t b = code - wlo;
  const bool inwin = code != 0 && b < HIST_WIN;
  lds_add(lh + (inwin && active ? b : (uint32_t)HIST_WIN), 1u);
  return active && !inwin;
}

// the centre rule (pair_combine) over two / three per-axis predictions in
// ascending axis order, for a centre fixed at compile time
template <int CENTER>
__device__ __forceinline__ double combine2(double v0, double v1) {
  if (CENTER == 0) return (v0 + v1) * 0.5;
  return CENTER == 1 ? v0 : v1;
}
template <int CENTER>
__device__ __forceinline__ double combine3(double v0, double v1, double v2) {
  if (CENTER == 0) return div3((v0 + v1) + v2);
  return CENTER == 1 ? v0 : v2;
}

template <bool ODD, bool ENC>
struct SheetLd {
  // even sheets: the coarse point of row E, x at (2E, 2m + 1), x of the odd
  // row's pair (encode) or the codes (decode); odd sheets: the even sheets'
  // pairs at rows 2E and 2J + 1 (taps along i), x of both pairs (encode) or
  // the codes (decode)
  float c;
  float2 te[ODD ? 4 : 1], to[ODD ? 4 : 1];
  float2 xe, xo;
  uint32_t k0, k1, k2, k3;
};

template <bool ODD, bool ENC, bool RND, bool XD, int CENTER>
__global__ void __launch_bounds__(256) sheet_kernel(SheetArgs A, const float *__restrict__ x,
                                                    float *__restrict__ rec,
                                                    uint16_t *__restrict__ codes,
                                                    uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ out_rank,
                                                    float *__restrict__ out_val,
                                                    uint32_t *__restrict__ out_count) {
  __shared__ uint32_t lh[ENC ? 3 * (HIST_WIN + 1) : 1];
  const uint32_t nb = gridDim.x;
  const uint32_t lb = (nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const uint32_t sgi = lb / A.rgroups, rg = lb - sgi * A.rgroups;
  if (sgi >= A.nseg) return;  // padding of the grid to a multiple of 8
  if (ENC) {
    for (int i = threadIdx.x; i < 3 * (HIST_WIN + 1); i += blockDim.x) lh[i] = 0;
    __syncthreads();
  }
  const uint32_t np = A.npairs, total = A.nsheets * np;
  const int lane = (int)(threadIdx.x & 63);
  const uint32_t run = rg * 4 + (threadIdx.x >> 6);
  const int64_t P = (int64_t)run * SHEET_OWN + lane - 1;
  const bool live = (int64_t)run * SHEET_OWN < (int64_t)total;  // uniform in the wave
  const bool own = lane >= 1 && lane <= SHEET_OWN && P < (int64_t)total;
  const bool pvalid = P >= 0 && P < (int64_t)total;
  const uint32_t Pc = (uint32_t)min(max(P, (int64_t)0), (int64_t)total - 1);
  const uint32_t a = Pc / np, m = Pc - a * np;
  const int n0 = A.n[0], n1 = A.n[1], n2 = A.n[2];
  const uint32_t gs1 = (uint32_t)n2, gs0 = (uint32_t)n1 * (uint32_t)n2;
  const int ne = (n1 + 1) / 2, no = n1 / 2;  // even / odd rows
  const int J0 = (int)(sgi * A.seg), J1 = min(J0 + (int)A.seg, no);
  const int E1 = sgi + 1 == A.nseg ? ne : J0 + (int)A.seg;  // owned even rows [J0, E1)
  const uint32_t i = 2 * a + (ODD ? 1u : 0u);
  const int end_mode = A.end_mode;
  // k taps of the odd point 2m + 1: pairs m - 1, m, m + 1, m + 2 (lanes l - 1,
  // l, l + 1, l + 2); kfix: an owned lane without all four (its end rules are
  // applied after the cubic stencil, under a wave vote)
  const bool kfix = own && !(m >= 1 && 2 * (int)m + 4 < n2);
  const bool kany = __any(kfix);  // uniform in the wave
  const int am1 = 4 * max(lane - 1, 0), ap1 = 4 * min(lane + 1, 63), ap2 = 4 * min(lane + 2, 63);
  // taps along i (odd sheets): sheets i - 3, i - 1, i + 1, i + 3; an invalid
  // tap reads sheet i - 1 (value unused)
  const bool ifast = !ODD || __all(!pvalid || (i >= 3 && (int)i + 3 < n0));
  const uint32_t fs = i * gs0 + 2 * m;  // the lane's pair at row 0 of its sheet
  const uint32_t NONE = BUF_OFF_NONE;
  const uint32_t vrec_own = own ? 4u * fs : NONE;
  uint32_t vt[4] = {0, 0, 0, 0};
  if constexpr (ODD) {
    const int32_t g0 = (int32_t)gs0;
    const int32_t ioff[4] = {(i >= 3 ? -3 : -1) * g0, -g0, ((int)i + 1 < n0 ? 1 : -1) * g0,
                             ((int)i + 3 < n0 ? 3 : -1) * g0};
#pragma unroll
    for (int q = 0; q < 4; q++) vt[q] = 4u * (uint32_t)((int32_t)fs + ioff[q]);
  }
  const uint32_t fxs = XD ? fs : i * A.xs[0] + 2 * m * A.xs[2];
  const uint32_t xs1 = XD ? gs1 : A.xs[1];
  const uint32_t vx0 = 4u * fxs, vx1 = 4u * (fxs + (XD ? 1u : A.xs[2]));
  const uint32_t vc = 4u * (a * (uint32_t)A.cn[1] * (uint32_t)A.cn[2] + m);
  // classes: [0] (e,e) [1] (e,o) [2] (o,e) [3] (o,o); stage rank of row r:
  // rb + r * rs; code byte offsets: lane part 2 rb, row part 2 (base + r rs)
  uint32_t rb[4], vk_own[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const SheetClass &K = A.cl[c];
    rb[c] = (uint32_t)(K.offset + (int64_t)a * K.rstr[0] + (int64_t)m * K.rstr[2]);
    vk_own[c] = own ? 2u * rb[c] : NONE;
  }
  // stage slots of the classes (the launcher fills A.st in this order): even
  // sheets 010 | 001 | 011, odd sheets 100 | 101 and 110 | 111
  constexpr int st0 = 0, st1 = 1, st2 = ODD ? 1 : 0, st3 = 2;
  auto slot = [](int c) __attribute__((always_inline)) { return c == 0 ? st0 : c == 1 ? st1 : c == 2 ? st2 : st3; };
  auto rs = [&](int c) __attribute__((always_inline)) { return (uint32_t)A.cl[c].rstr[1]; };
  auto cbs = [&](int c) __attribute__((always_inline)) { return (uint32_t)A.st[slot(c)].code_base; };
  SQ Q[3];
#pragma unroll
  for (int s = 0; s < 3; s++) Q[s] = SQ{A.st[s].w, A.st[s].inv_w, A.st[s].feb};
  const uint32_t R = A.st[0].R;
  const QR G{(double)R, 0.5 - 1e-12 * ((double)R + 1.0), R};
  const uint32_t wlo = R > HIST_WIN / 2 ? R - HIST_WIN / 2 : 0;
  const int round_mode = A.round_mode;
  const __amdgpu_buffer_rsrc_t rrec = buf_rsrc(rec), rcodes = buf_rsrc(codes);
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(ENC ? x : rec), rco = buf_rsrc(ODD ? rec : A.coarse);

  // row windows along j (f64), slot = row mod 4: even sheets Wa = the coarse
  // points, Wb = 001; odd sheets Wa = 100, Wb = 101
  double Wa[4] = {0.0, 0.0, 0.0, 0.0}, Wb[4] = {0.0, 0.0, 0.0, 0.0};
  // the loads run one step ahead of the arithmetic (step t + 1's are issued at
  // the start of step t)
  SheetLd<ODD, ENC> Lb[2];

  auto load = [&](int t, SheetLd<ODD, ENC> &D) __attribute__((always_inline)) {
    const uint32_t E = (uint32_t)min(max(t + 2, 0), ne - 1);
    const uint32_t J = (uint32_t)min(max(t - 1, 0), no - 1);
    const uint32_t sE = 8u * E * gs1, sJ = 4u * (2 * J + 1) * gs1;
    const uint32_t xE = 8u * E * xs1, xJ = 4u * (2 * J + 1) * xs1;
    if constexpr (!ODD) {
      D.c = bld32(rco, vc, 4u * E * (uint32_t)A.cn[2]);
      if constexpr (ENC) {
        D.xe.y = bld32(rx, vx1, xE);
        if (XD)
          D.xo = bld64(rx, vx0, xJ);
        else
          D.xo = make_float2(bld32(rx, vx0, xJ), bld32(rx, vx1, xJ));
      } else {
        D.k1 = bld16(rcodes, 2u * rb[1], 2u * (cbs(1) + E * rs(1)));
        D.k2 = bld16(rcodes, 2u * rb[2], 2u * (cbs(2) + J * rs(2)));
        D.k3 = bld16(rcodes, 2u * rb[3], 2u * (cbs(3) + J * rs(3)));
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) D.te[q] = bld64(rrec, vt[q], sE);
#pragma unroll
      for (int q = 0; q < 4; q++) D.to[q] = bld64(rrec, vt[q], sJ);
      if constexpr (ENC) {
        if (XD) {
          D.xe = bld64(rx, vx0, xE);
          D.xo = bld64(rx, vx0, xJ);
        } else {
          D.xe = make_float2(bld32(rx, vx0, xE), bld32(rx, vx1, xE));
          D.xo = make_float2(bld32(rx, vx0, xJ), bld32(rx, vx1, xJ));
        }
      } else {
        D.k0 = bld16(rcodes, 2u * rb[0], 2u * (cbs(0) + E * rs(0)));
        D.k1 = bld16(rcodes, 2u * rb[1], 2u * (cbs(1) + E * rs(1)));
        D.k2 = bld16(rcodes, 2u * rb[2], 2u * (cbs(2) + J * rs(2)));
        D.k3 = bld16(rcodes, 2u * rb[3], 2u * (cbs(3) + J * rs(3)));
      }
    }
  };

  // along k: the cubic stencil on every lane, then the end rules on the owned
  // lanes that need them (under a wave vote)
  // (the end-rule branches take the tap validity from an opaque copy of the
  // lane's coordinate, so that it is not kept as masks across the march)
  auto kaxis = [&](double A_, double B_, double C_, double D_) __attribute__((always_inline)) {
    double v = cubic4(A_, B_, C_, D_);
    if (kany) {
      if (kfix) {
        int mm = (int)m;
        asm volatile("" : "+v"(mm));
        v = interp4(A_, B_, C_, D_, mm >= 1, 2 * mm + 2 < n2, 2 * mm + 4 < n2, 1, end_mode);
      }
    }
    return v;
  };
  auto iaxis = [&](float a_, float b_, float c_, float d_) __attribute__((always_inline)) {
    if (ifast) return cubic4((double)a_, (double)b_, (double)c_, (double)d_);
    int ii = (int)i;
    asm volatile("" : "+v"(ii));
    return interp4((double)a_, (double)b_, (double)c_, (double)d_, ii >= 3, ii + 1 < n0,
                   ii + 3 < n0, 1, end_mode);
  };
  // decode of one point (code 0: the outlier's value, by rank)
  auto dq = [&](int s, double pd, uint32_t code, bool active, int c, uint32_t row) __attribute__((always_inline)) {
    const L3Stage &S = A.st[s];
    return dequant(Q[s], G, pd, code, active, rb[c] + row * rs(c), out_rank + S.out_lo,
                   out_val + S.out_lo, (int64_t)out_count[S.gs]);
  };
  // the rare emission path of up to two points (global count, outliers)
  auto rare2 = [&](bool ra, bool rb2, int sa, int sb, int ca, int cb, uint32_t row_a,
                   uint32_t row_b, uint32_t cda, uint32_t cdb, float xa, float xb, int npt) __attribute__((always_inline)) {
    if (__any(ra || (npt > 1 && rb2))) {
      // the per-stage offsets this path needs are read here, through an
      // opaque copy of the argument segment's (constant address space)
      // address, so that they are not kept live across the march; the arrays
      // are the kernel's own (global) arguments
      typedef const __attribute__((address_space(4))) SheetArgs *KArgs;
      KArgs pa = (KArgs)__builtin_amdgcn_kernarg_segment_ptr();
      asm volatile("" : "+s"(pa));
      const int64_t lo_a = pa->st[sa].out_lo, ho_a = pa->hoff[sa];
      const int32_t gs_a = pa->st[sa].gs;
      emit_rare(ra, rb[ca] + row_a * (uint32_t)pa->cl[ca].rstr[1], cda, xa, out_rank + lo_a,
                out_val + lo_a, out_count + gs_a, hist + ho_a);
      if (npt > 1) {
        const int64_t lo_b = pa->st[sb].out_lo, ho_b = pa->hoff[sb];
        const int32_t gs_b = pa->st[sb].gs;
        emit_rare(rb2, rb[cb] + row_b * (uint32_t)pa->cl[cb].rstr[1], cdb, xb, out_rank + lo_b,
                  out_val + lo_b, out_count + gs_b, hist + ho_b);
      }
    }
  };

  // One step t (U = t mod 4, known at compile time: J0 is a multiple of 4).
  // MAIN: the odd row J = t - 1 is coded (false in the four lead-in steps).
  auto step = [&](auto Uc, auto Mc, int t) __attribute__((always_inline)) {
    constexpr int U = decltype(Uc)::value;
    constexpr bool MAIN = decltype(Mc)::value;
    // every step issues the same loads (rows clamped into the grid), so the
    // number of memory operations between a load and its use is fixed
    SheetLd<ODD, ENC> &cur = Lb[U & 1], &nxt = Lb[(U + 1) & 1];
    load(t + 1, nxt);
    const int E = t + 2, J = t - 1;
    // stores are issued on every step (an unowned row or lane stores past the
    // buffer's range, which the hardware drops): with a fixed number of memory
    // operations per step, the wait for the next step's loads need not wait
    // for this step's stores
    const bool eown = E >= J0 && E < E1;  // uniform
    const uint32_t vrE = eown ? vrec_own : NONE;
    const bool evalid = pvalid && E >= 0 && E < ne;
    const uint32_t uE = (uint32_t)max(E, 0), uJ = (uint32_t)max(J, 0);
    // taps along j of odd row J: rows J - 1 .. J + 2 = t - 2 .. t + 1
    constexpr int r0 = (U + 2) & 3, r1 = (U + 3) & 3, r2 = U, r3 = (U + 1) & 3;
    const bool jfast = J >= 1 && 2 * J + 4 < n1;  // uniform
    auto jaxis = [&](const double *W) __attribute__((always_inline)) {
      if (jfast) return cubic4(W[r0], W[r1], W[r2], W[r3]);
      return interp4(W[r0], W[r1], W[r2], W[r3], J >= 1, 2 * J + 2 < n1, 2 * J + 4 < n1, 1,
                     end_mode);
    };
    const uint32_t soE = 8u * uE * gs1, soJ = 4u * (2 * uJ + 1) * gs1;
    if constexpr (!ODD) {
      // ---- even sheet: 010 of row J (taps: the coarse points of rows J-1 .. J+2)
      const float cE = cur.c;
      float r010 = 0.f;
      uint32_t c010 = 0;
      if constexpr (MAIN) {
        const double pd = (double)(float)jaxis(Wa);
        if constexpr (ENC)
          quant1<RND>(Q[st2], G, round_mode, A.span, A.offset, pd, cur.xo.x, c010, r010);
        else
          r010 = dq(st2, pd, c010 = cur.k2, pvalid, 2, uJ);
      }
      Wa[r0] = (double)cE;  // row E (the slot of row t - 2, read above)
      const double cm1 = (double)lane_get(cE, am1), cp1 = (double)lane_get(cE, ap1),
                   cp2 = (double)lane_get(cE, ap2);
      // 001 of row E from the coarse points along k
      const double p001 = (double)(float)kaxis(cm1, (double)cE, cp1, cp2);
      float r001, r011 = 0.f;
      uint32_t c001, c011 = 0;
      if constexpr (MAIN) {
        // 011 of row J: 001 along j (rows J-1 .. J+2), 010 along k
        const double gm1 = (double)lane_get(r010, am1), gp1 = (double)lane_get(r010, ap1),
                     gp2 = (double)lane_get(r010, ap2);
        const double p011 =
            (double)(float)combine2<CENTER>(jaxis(Wb), kaxis(gm1, (double)r010, gp1, gp2));
        if constexpr (ENC) {
          quant2<RND>(Q[st1], Q[st3], G, round_mode, A.span, A.offset, p001, cur.xe.y, p011,
                      cur.xo.y, c001, r001, c011, r011);
        } else {
          r001 = dq(st1, p001, c001 = cur.k1, evalid, 1, uE);
          r011 = dq(st3, p011, c011 = cur.k3, pvalid, 3, uJ);
        }
      } else {
        if constexpr (ENC)
          quant1<RND>(Q[st1], G, round_mode, A.span, A.offset, p001, cur.xe.y, c001, r001);
        else
          r001 = dq(st1, p001, c001 = cur.k1, evalid, 1, uE);
      }
      Wb[r0] = (double)r001;
      bst64(rrec, vrE, soE, cE, r001);
      if constexpr (ENC) {
        const bool ra = sheet_emit(rcodes, eown ? vk_own[1] : NONE, 2u * (cbs(1) + uE * rs(1)),
                                   c001, lh + st1 * (HIST_WIN + 1), wlo, own && eown);
        rare2(ra, false, st1, st1, 1, 1, uE, uE, c001, c001, cur.xe.y, cur.xe.y, 1);
      }
      if constexpr (MAIN) {
        bst64(rrec, vrec_own, soJ, r010, r011);
        if constexpr (ENC) {
          const bool ra = sheet_emit(rcodes, vk_own[2], 2u * (cbs(2) + uJ * rs(2)), c010,
                                     lh + st2 * (HIST_WIN + 1), wlo, own);
          const bool rb2 = sheet_emit(rcodes, vk_own[3], 2u * (cbs(3) + uJ * rs(3)), c011,
                                      lh + st3 * (HIST_WIN + 1), wlo, own);
          rare2(ra, rb2, st2, st3, 2, 3, uJ, uJ, c010, c011, cur.xo.x, cur.xo.y, 2);
        }
      }
    } else {
      // ---- odd sheet: 100 of row E (taps along i) and 110 of row J (along
      // i, and 100 of rows J-1 .. J+2 along j)
      const double p100 = (double)(float)iaxis(cur.te[0].x, cur.te[1].x, cur.te[2].x,
                                               cur.te[3].x);
      float r100, r110 = 0.f;
      uint32_t c100, c110 = 0;
      if constexpr (MAIN) {
        const double p110 = (double)(float)combine2<CENTER>(
            iaxis(cur.to[0].x, cur.to[1].x, cur.to[2].x, cur.to[3].x), jaxis(Wa));
        if constexpr (ENC) {
          quant2<RND>(Q[st0], Q[st2], G, round_mode, A.span, A.offset, p100, cur.xe.x, p110,
                      cur.xo.x, c100, r100, c110, r110);
        } else {
          r100 = dq(st0, p100, c100 = cur.k0, evalid, 0, uE);
          r110 = dq(st2, p110, c110 = cur.k2, pvalid, 2, uJ);
        }
      } else {
        if constexpr (ENC)
          quant1<RND>(Q[st0], G, round_mode, A.span, A.offset, p100, cur.xe.x, c100, r100);
        else
          r100 = dq(st0, p100, c100 = cur.k0, evalid, 0, uE);
      }
      Wa[r0] = (double)r100;
      // 101 of row E: 001 along i, 100 along k
      const double hm1 = (double)lane_get(r100, am1), hp1 = (double)lane_get(r100, ap1),
                   hp2 = (double)lane_get(r100, ap2);
      const double p101 = (double)(float)combine2<CENTER>(
          iaxis(cur.te[0].y, cur.te[1].y, cur.te[2].y, cur.te[3].y),
          kaxis(hm1, (double)r100, hp1, hp2));
      float r101, r111 = 0.f;
      uint32_t c101, c111 = 0;
      if constexpr (MAIN) {
        // 111 of row J: 011 along i, 101 along j (rows J-1 .. J+2), 110 along k
        const double gm1 = (double)lane_get(r110, am1), gp1 = (double)lane_get(r110, ap1),
                     gp2 = (double)lane_get(r110, ap2);
        const double p111 = (double)(float)combine3<CENTER>(
            iaxis(cur.to[0].y, cur.to[1].y, cur.to[2].y, cur.to[3].y), jaxis(Wb),
            kaxis(gm1, (double)r110, gp1, gp2));
        if constexpr (ENC) {
          quant2<RND>(Q[st1], Q[st3], G, round_mode, A.span, A.offset, p101, cur.xe.y, p111,
                      cur.xo.y, c101, r101, c111, r111);
        } else {
          r101 = dq(st1, p101, c101 = cur.k1, evalid, 1, uE);
          r111 = dq(st3, p111, c111 = cur.k3, pvalid, 3, uJ);
        }
      } else {
        if constexpr (ENC)
          quant1<RND>(Q[st1], G, round_mode, A.span, A.offset, p101, cur.xe.y, c101, r101);
        else
          r101 = dq(st1, p101, c101 = cur.k1, evalid, 1, uE);
      }
      Wb[r0] = (double)r101;
      bst64(rrec, vrE, soE, r100, r101);
      if constexpr (ENC) {
        const bool ra = sheet_emit(rcodes, eown ? vk_own[0] : NONE, 2u * (cbs(0) + uE * rs(0)),
                                   c100, lh + st0 * (HIST_WIN + 1), wlo, own && eown);
        const bool rb2 = sheet_emit(rcodes, eown ? vk_own[1] : NONE, 2u * (cbs(1) + uE * rs(1)),
                                    c101, lh + st1 * (HIST_WIN + 1), wlo, own && eown);
        rare2(ra, rb2, st0, st1, 0, 1, uE, uE, c100, c101, cur.xe.x, cur.xe.y, 2);
      }
      if constexpr (MAIN) {
        bst64(rrec, vrec_own, soJ, r110, r111);
        if constexpr (ENC) {
          const bool ra = sheet_emit(rcodes, vk_own[2], 2u * (cbs(2) + uJ * rs(2)), c110,
                                     lh + st2 * (HIST_WIN + 1), wlo, own);
          const bool rb2 = sheet_emit(rcodes, vk_own[3], 2u * (cbs(3) + uJ * rs(3)), c111,
                                      lh + st3 * (HIST_WIN + 1), wlo, own);
          rare2(ra, rb2, st2, st3, 2, 3, uJ, uJ, c110, c111, cur.xo.x, cur.xo.y, 2);
        }
      }
    }
  };

  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using BF = std::integral_constant<bool, false>;
  using BT = std::integral_constant<bool, true>;
  auto march = [&]() __attribute__((always_inline)) {
    // lead-in: even rows J0 - 1 .. J0 + 2 (steps J0 - 3 .. J0: t mod 4 = 1, 2, 3, 0)
    load(J0 - 3, Lb[1]);
    step(I1(), BF(), J0 - 3);
    step(I2(), BF(), J0 - 2);
    step(I3(), BF(), J0 - 1);
    step(I0(), BF(), J0);
    for (int t = J0 + 1; t <= J1; t += 4) {
      step(I1(), BT(), t);
      if (t + 1 > J1) break;
      step(I2(), BT(), t + 1);
      if (t + 2 > J1) break;
      step(I3(), BT(), t + 2);
      if (t + 3 > J1) break;
      step(I0(), BT(), t + 3);
    }
  };
  if (live) march();
  if (ENC) {
    __syncthreads();
    for (int s = 0; s < 3; s++)
      flush_hist(lh + s * (HIST_WIN + 1), hist + (int64_t)A.st[s].gs * A.hist_words, A.st[s].R);
  }
}

template <bool ODD, int B>
static void sheet_go(int bits, int center, dim3 grid, hipStream_t stream, const SheetArgs &a,
                     const float *x, float *rec, uint16_t *codes, uint32_t *hist,
                     uint32_t *out_rank, float *out_val, uint32_t *out_count) {
  if constexpr (B < 0) {
    return;
  } else {
    constexpr bool E = B & 1, RND = B & 2, XD = B & 4;
    if (bits != B) {
      sheet_go<ODD, B - 1>(bits, center, grid, stream, a, x, rec, codes, hist, out_rank,
                           out_val, out_count);
    } else if constexpr (E || (!RND && XD)) {
      if (center == 0)
        hipLaunchKernelGGL((sheet_kernel<ODD, E, RND, XD, 0>), grid, dim3(256), 0, stream, a, x,
                           rec, codes, hist, out_rank, out_val, out_count);
      else if (center == 1)
        hipLaunchKernelGGL((sheet_kernel<ODD, E, RND, XD, 1>), grid, dim3(256), 0, stream, a, x,
                           rec, codes, hist, out_rank, out_val, out_count);
      else
        hipLaunchKernelGGL((sheet_kernel<ODD, E, RND, XD, 2>), grid, dim3(256), 0, stream, a, x,
                           rec, codes, hist, out_rank, out_val, out_count);
    }
  }
}

int sheet_launch(const SheetArgs &a, bool odd, int bits, const float *x, float *rec,
                 uint16_t *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                 uint32_t *out_count, hipStream_t stream) {
  const int64_t nblk = ((int64_t)a.rgroups * a.nseg + 7) / 8 * 8;
  const dim3 grid((unsigned)nblk);
  if (a.cubic != 1 || a.center < 0 || a.center > 2) {
    set_error("sheet kernels: cubic interpolation and centre 0, 1 or 2 only");
    return LATEC_E_ARG;
  }
  if (odd)
    sheet_go<true, 7>(bits, a.center, grid, stream, a, x, rec, codes, hist, out_rank, out_val,
                      out_count);
  else
    sheet_go<false, 7>(bits, a.center, grid, stream, a, x, rec, codes, hist, out_rank, out_val,
                       out_count);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_hip_error("sheet_kernel", (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

}  // namespace latec
