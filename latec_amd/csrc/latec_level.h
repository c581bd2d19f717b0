// latec_level.h -- device helpers shared by the 3-D level kernels
// (latec_interp.hip: march / pair / quad kernels; latec_sheet.hip: sheet
// kernels): histogram windows, the stage quantizer of the level kernels, the
// interpolation stencils on taps already read, buffer stores and the code /
// outlier emission.  Every function is the arithmetic of
// latec/predictor.py:97-156 (stencils) and latec/quantizer.py:57-83
// (quantizer), operation for operation (built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace latec {

#define HIST_WIN 1024
#define HIST_COPIES 256
// Per-stage histogram layout: [2R main bins][HIST_COPIES x HIST_WIN copies of
// the central window].  A block flushes its LDS window into copy
// (block id % HIST_COPIES), so thousands of blocks do not serialize their
// atomics on the same few bins; the fit folds the copies into the main bins.
__device__ __forceinline__ void flush_hist(const uint32_t *lhist, uint32_t *__restrict__ ghist,
                                           uint32_t R) {
  const uint32_t bid = blockIdx.y * gridDim.x + blockIdx.x;
  uint32_t *gwin = ghist + 2 * (size_t)R + (size_t)(bid % HIST_COPIES) * HIST_WIN;
  for (int i = threadIdx.x; i < HIST_WIN; i += blockDim.x)
    if (lhist[i]) atomicAdd(&gwin[i], lhist[i]);
}

// Histogram counts into a block's LDS window (workgroup scope) or straight
// into the global bins (device scope).  The distinct scopes keep the two
// atomics of a window-or-global branch apart: merged, they would become one
// flat atomic on a selected pointer, which is slower and makes every later
// wait on vector-memory loads wait for it too.
__device__ __forceinline__ void lds_add(uint32_t *p, uint32_t k) {
  __hip_atomic_fetch_add(p, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void gbl_add(uint32_t *p, uint32_t k) {
  __hip_atomic_fetch_add(p, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct L3Stage {
  double w, inv_w;
  float feb;
  uint32_t R;
  int64_t code_base;  // codes + code_base + rank
  int64_t out_lo;     // outlier arrays offset of the stage
  int32_t gs, pad;    // global stage index (hist, out_count)
};

// The taps are float32 values held in double, so a product of a tap and a
// small integer or half-integer coefficient (1.5, 3, 6, 9: at most 4 more
// significant bits than the tap's 24) is exact in double.  Then the
// reference's RN(u + RN(c t)) equals RN(u + c t), one FMA: each stencil below
// keeps the reference's operation order and rounding, bit for bit, with its
// products folded into the following addition.

// interp_axis on taps already read (A = c-3s, B = c-s, C = c+s, D = c+3s).
__device__ __forceinline__ double interp4(double A, double B, double C, double D, bool vm3,
                                          bool vp1, bool vp3, int cubic, int end_mode) {
  if (!vp1) {
    // 1.5 B - 0.5 A: 0.5 A is exact too
    if ((end_mode & 2) && vm3) return fma(1.5, B, -(0.5 * A));
    return B;
  }
  if (cubic) {
    // (((-A + 9B) + 9C) - D) / 16
    if (vm3 && vp3) return (fma(9.0, C, fma(9.0, B, -A)) - D) * 0.0625;
    if (end_mode & 1) {
      // ((3B + 6C) - D) / 8 and ((-A + 6B) + 3C) / 8
      if (vp3) return (fma(6.0, C, 3.0 * B) - D) * 0.125;
      if (vm3) return fma(3.0, C, fma(6.0, B, -A)) * 0.125;
    }
  }
  return 0.5 * (B + C);
}

// Loads and stores at a 32-bit byte offset from a uniform base (SGPR base +
// VGPR offset addressing; every level grid here is below 2^30 values).
template <typename T>
__device__ __forceinline__ T ldo(const T *__restrict__ base, uint32_t idx) {
  return *(const T *)((const char *)base + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void sto(T *__restrict__ base, uint32_t idx, T v) {
  *(T *)((char *)base + idx * (uint32_t)sizeof(T)) = v;
}

// RN(x / 3) without the division sequence.  Let c = RN(1/3) = (1/3)(1 - 2^-54)
// (1/3 = 0.0101...b, and the double below it is the nearest), q0 = RN(x c) and
// q = RN(q0 + r c) with r = x - 3 q0 from one FMA.  Then
//   |q0 - x/3| <= |q0 - x c| + |x c - x/3| <= ulp(x/3)/2 + 2^-54 |x/3| < ulp(x/3),
// because ulp(t) > 2^-53 |t| for every normal t: q0 is within one ulp of x/3.
// A quotient within one ulp leaves a remainder x - 3 q0 that is a double, so
// the FMA computes r exactly, and with c within half an ulp of 1/3 Markstein's
// theorem (P. Markstein, IBM J. Res. Dev. 34(1), 1990, Theorem 1; restated in
// Muller et al., Handbook of Floating-Point Arithmetic, the division chapter)
// gives q = RN(x / 3) exactly.  Both steps assume that q0 and r are normal
// and that nothing overflows: for |x| >= 2^-968, q0 >= 2^-970 and r (a
// multiple of ulp(q0) >= 2^-1022) is zero or normal; |x| <= 2^1020 keeps 3 q0
// finite; the correction is one FMA, so r c is never rounded on its own.
// Other values (tiny, huge, non-finite) take the division itself.  (Also
// checked bit for bit against the division on 4.6e8 operands across the
// range, including multiples of 3 and their neighbours.)  Every lane of the
// wave calls it.
__device__ __forceinline__ double div3(double x) {
  const double c = 1.0 / 3.0;
  double q = x * c;
  q = fma(fma(-q, 3.0, x), c, q);
  const bool ok = fabs(x) >= 0x1p-968 && fabs(x) <= 0x1p1020;
  if (__any(!ok)) {
    if (!ok) q = x / 3.0;
  }
  return q;
}

// Cubic interior stencil of interp4 (all four taps valid, float32 values):
// (((-A + 9B) + 9C) - D) / 16 with the exact products folded into FMAs.
__device__ __forceinline__ double cubic4(double A, double B, double C, double D) {
  return (fma(9.0, C, fma(9.0, B, -A)) - D) * 0.0625;
}

// the centre rule over the per-axis predictions v[0..nodd) (ascending axes)
__device__ __forceinline__ double pair_combine(const double *v, int nodd, int center) {
  if (nodd == 1) return v[0];
  if (center == 0) {
    double total = v[0] + v[1];
    if (nodd == 3) return div3(total + v[2]);
    return total * 0.5;
  }
  return center == 1 ? v[0] : v[nodd - 1];
}

// One stage's quantizer constants.  hlo = 0.5 - 1e-12 (R + 1): a product
// t = d / w rounded through 1/w whose distance to the nearest integer is
// below hlo rounds like the division (the margin is the general rule's
// 1e-12 max(|t|, 1) at its largest |t| that can give an in-range code; a
// |t| above R + 1 is out of range either way), so the fallback needs one
// compare.
struct PQ {
  double w, inv_w, Rd;
  float feb;
  uint32_t R;
  double hlo;
};
__device__ __forceinline__ PQ make_pq(const L3Stage &S) {
  return PQ{S.w, S.inv_w, (double)S.R, S.feb, S.R, 0.5 - 1e-12 * ((double)S.R + 1.0)};
}

// quantize_one (encode): code and reconstruction of x against prediction p.
// Every lane of the wave calls it (the exact-division fallback is a vote).
template <bool RND>  // RND false: round_mode is 0 (no output rounding check)
__device__ __forceinline__ void pair_quant(const PQ &Q, int round_mode, double span,
                                           double offset, double pd, float xv, uint32_t &code,
                                           float &r) {
  const double d = (double)xv - pd;
  const double tq = d * Q.inv_w;
  double qd = rint(tq);
  const bool fb = !(fabs(tq - qd) < Q.hlo);
  if (__any(fb)) {
    if (fb) qd = rint(d / Q.w);
  }
  const bool inr = fabs(qd) < Q.Rd;
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
  code = ok ? (uint32_t)(qi + (int)Q.R) : 0u;
  r = ok ? cand : xv;
}

// pair_quant of two points with one wave vote for the exact-division fallback.
template <bool RND>
__device__ __forceinline__ void pair_quant2(const PQ &Q, int round_mode, double span,
                                            double offset, double pd0, float x0, double pd1,
                                            float x1, uint32_t &code0, float &r0,
                                            uint32_t &code1, float &r1) {
  const double d0 = (double)x0 - pd0, d1 = (double)x1 - pd1;
  const double t0 = d0 * Q.inv_w, t1 = d1 * Q.inv_w;
  double q0 = rint(t0), q1 = rint(t1);
  const bool fb0 = !(fabs(t0 - q0) < Q.hlo);
  const bool fb1 = !(fabs(t1 - q1) < Q.hlo);
  if (__any(fb0 || fb1)) {
    if (fb0) q0 = rint(d0 / Q.w);
    if (fb1) q1 = rint(d1 / Q.w);
  }
  auto finish = [&](double pd, float xv, double qd, uint32_t &code, float &r) {
    const bool inr = fabs(qd) < Q.Rd;
    const int qi = (int)qd;  // used only when in range
    const float cand = (float)(pd + Q.w * (qd + 0.0));
    float chk = cand;
    if (RND) {
      if (round_mode == 1)
        chk = rintf(cand);
      else if (round_mode == 2)
        chk = (float)((rint((double)cand * span + offset) - offset) / span);
    }
    const bool ok = inr && fabsf(xv - chk) <= Q.feb;
    code = ok ? (uint32_t)(qi + (int)Q.R) : 0u;
    r = ok ? cand : xv;
  };
  finish(pd0, x0, q0, code0, r0);
  finish(pd1, x1, q1, code1, r1);
}

// Decode: reconstruction from a code; code 0 looks the value up among the
// stage's outliers (sorted by rank), a missing one gives NaN.
__device__ __forceinline__ float pair_dequant(const PQ &Q, double pd, uint32_t code, bool active,
                                              uint32_t rank, const uint32_t *__restrict__ orank,
                                              const float *__restrict__ oval, int64_t nout) {
  float r = (float)(pd + Q.w * (double)((int)code - (int)Q.R));
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

// a pair (k, k + 1) at element idx; when the pair has no odd point the odd
// lane reads idx again (in bounds, value unused)
__device__ __forceinline__ float2 pair_ld(const float *__restrict__ b, uint32_t idx, bool al,
                                          bool odd_ok) {
  if (al) return ldo((const float2 *)b, idx >> 1);
  return make_float2(ldo(b, idx), ldo(b, odd_ok ? idx + 1 : idx));
}

// Masked stores as one unconditional buffer instruction: a lane that must
// not store writes to byte offset BUF_OFF_NONE, beyond the resource's range,
// which the hardware drops.  With no branch around them the number of memory
// operations a step issues is fixed, so the wait for the next step's loads
// need not wait for this step's stores.  (gfx9 buffer resource: 32-bit data
// format in the last word; 2^31 - 1 byte range, every array here is smaller.)
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
#define BUF_OFF_NONE 0x7fffffffu
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, 0x7fffffff,
                                           0x00020000);
}
__device__ __forceinline__ void buf_st32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on,
                                         float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, on ? byte_off : BUF_OFF_NONE, 0,
                                        0);
}
__device__ __forceinline__ void buf_st64(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on,
                                         float a, float b) {
  const u32x2_t v = {__float_as_uint(a), __float_as_uint(b)};
  __builtin_amdgcn_raw_buffer_store_b64(v, r, on ? byte_off : BUF_OFF_NONE, 0, 0);
}
__device__ __forceinline__ void buf_st16(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, bool on,
                                         uint16_t v) {
  __builtin_amdgcn_raw_buffer_store_b16(v, r, on ? byte_off : BUF_OFF_NONE, 0, 0);
}

// emit() of the pair kernels in two halves, so that a step votes once for all
// its points: emit_common stores the code and counts an in-window code in LDS
// (branch-free: a lane whose code is outside the window, or that is inactive,
// adds 0 to the window's spare word HIST_WIN) and says whether the point also
// needs emit_rare (a code outside the window, counted globally, or an
// outlier, appended).  emit_rare runs under a wave vote.
__device__ __forceinline__ bool emit_common(__amdgpu_buffer_rsrc_t codes, uint32_t cbase,
                                            uint32_t slot, uint32_t code, uint32_t *lhist,
                                            uint32_t win_lo, bool active) {
  buf_st16(codes, cbase + 2u * slot, active, (uint16_t)code);
  const uint32_t b = code - win_lo;
  const bool inwin = code != 0 && b < HIST_WIN;
  lds_add(lhist + (inwin ? b : (uint32_t)HIST_WIN), active && inwin ? 1u : 0u);
  return active && !inwin;
}
__device__ __forceinline__ void emit_rare(bool rare, uint32_t slot, uint32_t code, float x,
                                          uint32_t *__restrict__ out_rank,
                                          float *__restrict__ out_val,
                                          uint32_t *__restrict__ out_count,
                                          uint32_t *__restrict__ ghist) {
  if (rare) gbl_add(ghist + code, 1u);
  const bool out = rare && code == 0;
  const uint64_t m = __ballot(out);
  if (m) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(out_count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (out) {
      const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      out_rank[k] = slot;
      out_val[k] = x;
    }
  }
}

}  // namespace latec
