// latec_entropy.cpp -- host entropy stage: stack rANS, stage records, zstd.
//
// Stack rANS: the design of constriction 0.4's AnsCoder<u32, u64> driven
// by Categorical(probabilities_f32, perfect=False) at 24-bit precision, as
// bound by latec/rans.py:466-535 and 629-657 (model = the float32 pmf of one
// CDF row; symbols pushed last to first so they decode first to last):
//   model ..... left[k] = trunc(cum_f32(k) * f32((2^24 - A) / sum)) + k
//   encode .... if state >> 40 >= prob: emit low word, state >>= 32;
//               state = (state / prob) << 24 | (left + state % prob)
//   flush ..... the final state as its low word then (if nonzero) high word
//   decode .... the inverse, reading words from the end of the array.
// constriction is absent from this environment, so the word stream is a
// restatement of that published design, not a byte-checked one.
//
// Stage record and literal back end: latec/bitstream.py:178-344.
#include <stdint.h>
#include <string.h>
#define ZSTD_STATIC_LINKING_ONLY  // ZSTD_compressBegin_advanced / ZSTD_compressContinue
#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "latec_common.h"

namespace latec {

static const int PREC = 24;

// Start fn on a new thread, or run it here when the system cannot start one
// (std::system_error): the result is the same, only later.  The caller joins
// `t` (through a guard built before the first start, so an exception never
// destroys a joinable thread).
template <typename F>
static void start_or_run(std::thread &t, F &&fn) {
  try {
    t = std::thread(fn);
  } catch (const std::system_error &) {
    fn();
  }
}

// Joins every joinable thread it holds when it goes out of scope.
struct JoinAll {
  std::thread *t;
  int n;
  ~JoinAll() {
    for (int i = 0; i < n; i++)
      if (t[i].joinable()) t[i].join();
  }
};

struct SymInfo {
  uint32_t left, prob;
  uint64_t rcp;  // floor(2^64 / prob), 0 for prob == 1
};

// Decoder lookup: one 64-bit entry per 2^12-slot bucket of the 2^24 quantile
// range holds the symbol at the bucket's first slot as prob | left << 24 |
// k << 48 (left, prob < 2^24, k < 2^16; the probability in the low bits is
// ready for the state update with one mask), so a quantile inside that symbol
// (every quantile of a bucket with no symbol boundary, >= 94% of buckets for
// A <= 256) costs one 32 KiB-table load; the rest step forward through left[].
static const int DLUT_SHIFT = 12;

struct Model {
  int64_t A = 0;
  std::vector<uint32_t> left;  // A + 1 entries
  std::vector<SymInfo> sym;
  std::vector<uint64_t> dlut;  // 2^(PREC - DLUT_SHIFT) entries
};

// floor(2^64 / p) for 1 < p < 2^32 (0 for p <= 1): a power of two exactly,
// otherwise floor((2^64 - 1) / p), which is the same when p does not divide 2^64
static inline uint64_t rcp64(uint32_t p) {
  if (p <= 1) return 0;
  if ((p & (p - 1)) == 0) return (uint64_t)1 << (64 - __builtin_ctz(p));
  return ~(uint64_t)0 / p;
}

static void build_model(const uint32_t *cdf, int64_t A, Model &m, bool for_decode) {
  m.A = A;
  m.left.resize(A + 1);
  float sum = 0.0f;
  for (int64_t k = 0; k < A; k++) sum += (float)(cdf[k + 1] - cdf[k]);
  const float scale = (float)((1 << PREC) - A) / sum;
  float cum = 0.0f;
  for (int64_t k = 0; k < A; k++) {
    m.left[k] = (uint32_t)(cum * scale) + (uint32_t)k;
    cum += (float)(cdf[k + 1] - cdf[k]);
  }
  m.left[A] = 1u << PREC;
  m.sym.resize(A);
  for (int64_t k = 0; k < A; k++) {
    SymInfo &s = m.sym[k];
    s.left = m.left[k];
    s.prob = m.left[k + 1] - m.left[k];
    s.rcp = rcp64(s.prob);
  }
  if (for_decode) {
    m.dlut.assign((size_t)1 << (PREC - DLUT_SHIFT), 0);
    int64_t k = 0;
    for (size_t b = 0; b < m.dlut.size(); b++) {
      const uint32_t q = (uint32_t)(b << DLUT_SHIFT);
      while (k + 1 < A && m.left[k + 1] <= q) k++;
      m.dlut[b] = (uint64_t)m.sym[k].prob | ((uint64_t)m.sym[k].left << 24) |
                  ((uint64_t)k << 48);
    }
  }
}

static inline void enc_put(uint64_t &state, const SymInfo &s, uint32_t *words,
                           int64_t &nw) {
  if ((state >> (64 - PREC)) >= s.prob) {
    words[nw++] = (uint32_t)state;
    state >>= 32;
  }
  uint64_t q, r;
  if (s.prob == 1) {
    q = state;
    r = 0;
  } else {
    q = (uint64_t)(((unsigned __int128)state * s.rcp) >> 64);
    r = state - q * s.prob;
    if (r >= s.prob) {
      q++;
      r -= s.prob;
    }
  }
  state = (q << PREC) | (s.left + r);
}

static inline int64_t enc_flush(uint64_t state, uint32_t *words, int64_t nw) {
  if (state) {
    words[nw++] = (uint32_t)state;
    if (state >> 32) words[nw++] = (uint32_t)(state >> 32);
  }
  return nw;
}

// Worst case: one word per symbol plus the two-word flush.
static inline int64_t enc_cap(int64_t n) { return n + 2; }

static inline uint32_t dec_get(uint64_t &state, const Model &m, const uint32_t *words,
                               int64_t &top) {
  const uint32_t qt = (uint32_t)(state & ((1u << PREC) - 1));
  const uint64_t e = m.dlut[qt >> DLUT_SHIFT];
  uint32_t prob = (uint32_t)e & 0xffffffu, left = (uint32_t)(e >> 24) & 0xffffffu;
  uint32_t k = (uint32_t)(e >> 48);
  if (qt - left >= prob) {  // past the bucket's first symbol
    do k++;
    while (m.left[k + 1] <= qt);
    left = m.sym[k].left;
    prob = m.sym[k].prob;
  }
  state = (state >> PREC) * prob + (qt - left);
  if (state < ((uint64_t)1 << 32) && top > 0) state = (state << 32) | words[--top];
  return k;
}

static inline bool dec_init(uint64_t &state, const uint32_t *words, int64_t nw,
                            int64_t &top) {
  top = nw;
  state = 0;
  if (top > 0) {
    const uint32_t w0 = words[--top];
    if (w0 == 0) return false;
    state = w0;
    if (top > 0) state = (state << 32) | words[--top];
  }
  return true;
}

int64_t rans_encode(const uint32_t *syms, int64_t n, const uint32_t *cdf, int64_t A,
                    uint32_t *words, int64_t cap) {
  Model m;
  build_model(cdf, A, m, false);
  std::vector<uint32_t> tmp;
  uint32_t *out = words;
  if (cap < enc_cap(n)) {
    tmp.resize(enc_cap(n));
    out = tmp.data();
  }
  uint64_t state = 0;
  int64_t nw = 0;
  for (int64_t i = n - 1; i >= 0; i--) {
    if (syms[i] >= (uint64_t)A) {
      set_error("code outside the narrowed rANS alphabet");
      return LATEC_E_ARG;
    }
    enc_put(state, m.sym[syms[i]], out, nw);
  }
  nw = enc_flush(state, out, nw);
  if (out != words) {
    if (nw > cap) {
      set_error("rANS output capacity");
      return LATEC_E_CAPACITY;
    }
    memcpy(words, out, sizeof(uint32_t) * nw);
  }
  return nw;
}

// Level grouping of positions 0..n-1 (stable): level l's positions take
// [start[l], start[l + 1]) in increasing order.  False on a level >= 64.
static bool group_levels(const uint8_t *levels, int64_t n, int64_t start[65]) {
  int64_t cnt[256] = {0};
  for (int64_t i = 0; i < n; i++) cnt[levels[i]]++;
  for (int l = 64; l < 256; l++)
    if (cnt[l]) return false;
  start[0] = 0;
  for (int l = 0; l < 64; l++) start[l + 1] = start[l] + cnt[l];
  return true;
}

// Per-point models: the symbols sym(i) grouped by level (stable), levels
// pushed 63..0, each group last to first (encode_reverse).
template <class Sym>
static int64_t encode_levels_from(Sym sym, const uint8_t *levels, int64_t n,
                                  const uint32_t *cdfs, int64_t A, uint32_t *words,
                                  int64_t cap) {
  if (n > (int64_t)UINT32_MAX) {
    set_error("stage too large for the rANS coder");
    return LATEC_E_ARG;
  }
  int64_t start[65];
  if (!group_levels(levels, n, start)) {
    set_error("scale level exceeds table count");
    return LATEC_E_ARG;
  }
  std::unique_ptr<uint32_t[]> sorted(new uint32_t[(size_t)std::max<int64_t>(n, 1)]);
  {
    int64_t next[64];
    for (int l = 0; l < 64; l++) next[l] = start[l];
    for (int64_t i = 0; i < n; i++) sorted[next[levels[i]]++] = sym(i);
  }
  // words in place when the caller's array holds the worst case
  std::unique_ptr<uint32_t[]> tmp;
  uint32_t *out = words;
  if (cap < enc_cap(n)) {
    tmp.reset(new uint32_t[(size_t)enc_cap(n)]);
    out = tmp.get();
  }
  uint64_t state = 0;
  int64_t nw = 0;
  for (int l = 63; l >= 0; l--) {
    if (start[l + 1] == start[l]) continue;
    Model m;
    build_model(cdfs + (int64_t)l * (A + 1), A, m, false);
    const SymInfo *si = m.sym.data();
    for (int64_t j = start[l + 1] - 1; j >= start[l]; j--) {
      const uint32_t s = sorted[j];
      if (s >= (uint64_t)A) {
        set_error("code outside the narrowed rANS alphabet");
        return LATEC_E_ARG;
      }
      enc_put(state, si[s], out, nw);
    }
  }
  nw = enc_flush(state, out, nw);
  if (nw > cap) {
    set_error("rANS output capacity");
    return LATEC_E_CAPACITY;
  }
  if (out != words) memcpy(words, out, sizeof(uint32_t) * nw);
  return nw;
}

int64_t rans_encode_levels(const uint32_t *syms, const uint8_t *levels, int64_t n,
                           const uint32_t *cdfs, int64_t A, uint32_t *words, int64_t cap) {
  return encode_levels_from([syms](int64_t i) { return syms[i]; }, levels, n, cdfs, A, words,
                            cap);
}

// The same from the quantizer's codes (u16 or u32): a code c != 0 is the
// symbol c - radius + 2^kexp of the narrowed alphabet (rans.py:83-89); a
// code outside it wraps past A in uint32 arithmetic and is rejected there.
int64_t rans_encode_codes_levels(const void *codes, int cb, const uint8_t *levels, int64_t n,
                                 uint32_t radius, int kexp, const uint32_t *cdfs, int64_t A,
                                 uint32_t *words, int64_t cap) {
  if ((cb != 2 && cb != 4) || kexp < 1 || kexp > 30 || A != ((int64_t)2 << kexp)) {
    set_error("rans_encode_codes_levels: code_bytes 2 or 4 and A = 2 << kexp");
    return LATEC_E_ARG;
  }
  const uint32_t off = (1u << kexp) - radius;  // span + 1 - radius, mod 2^32
  if (cb == 2) {
    const uint16_t *c = (const uint16_t *)codes;
    return encode_levels_from(
        [c, off](int64_t i) { return c[i] ? (uint32_t)c[i] + off : 0u; }, levels, n, cdfs, A,
        words, cap);
  }
  const uint32_t *c = (const uint32_t *)codes;
  return encode_levels_from([c, off](int64_t i) { return c[i] ? c[i] + off : 0u; }, levels, n,
                            cdfs, A, words, cap);
}

// choose_kexp (rans.py:70-80): the bit length of the largest |c - radius|
// over the codes c != 0, at least 1.
int code_kexp(const void *codes, int cb, int64_t n, uint32_t radius) {
  uint32_t peak = 0;
  for (int64_t i = 0; i < n; i++) {
    const uint32_t c = cb == 2 ? ((const uint16_t *)codes)[i] : ((const uint32_t *)codes)[i];
    if (!c) continue;
    const uint32_t d = c >= radius ? c - radius : radius - c;
    peak = std::max(peak, d);
  }
  int b = 0;
  while (b < 32 && (peak >> b)) b++;
  return std::max(1, b);
}

// n symbols of one CDF row, each handed to sink(i, symbol) as it is decoded.
template <class Sink>
static int rans_decode_to(const uint32_t *words, int64_t nw, const uint32_t *cdf, int64_t A,
                          int64_t n, Sink sink) {
  Model m;
  build_model(cdf, A, m, true);
  uint64_t state;
  int64_t top;
  if (!dec_init(state, words, nw, top)) {
    set_error("malformed rANS payload");
    return LATEC_E_FORMAT;
  }
  for (int64_t i = 0; i < n; i++) sink(i, dec_get(state, m, words, top));
  if (top != 0 || state != 0) {
    set_error("trailing data in rANS payload");
    return LATEC_E_FORMAT;
  }
  return LATEC_OK;
}

int rans_decode(const uint32_t *words, int64_t nw, const uint32_t *cdf, int64_t A,
                uint32_t *syms, int64_t n) {
  return rans_decode_to(words, nw, cdf, A, n, [syms](int64_t i, uint32_t k) { syms[i] = k; });
}

// Per-point models: the level groups decoded into contiguous runs of a
// scratch array, then one pass in position order hands every position's
// symbol (from the front of its level's run) to put(i, symbol).
template <class Put>
static int decode_levels_to(const uint32_t *words, int64_t nw, const uint8_t *levels,
                            const uint32_t *cdfs, int64_t A, int64_t n, Put put) {
  uint64_t state;
  int64_t top;
  if (!dec_init(state, words, nw, top)) {
    set_error("malformed rANS payload");
    return LATEC_E_FORMAT;
  }
  if (n > (int64_t)UINT32_MAX) {
    set_error("stage too large for the rANS coder");
    return LATEC_E_ARG;
  }
  int64_t start[65];
  if (!group_levels(levels, n, start)) {
    set_error("scale level exceeds table count");
    return LATEC_E_ARG;
  }
  // the calling thread keeps its scratch between calls up to 2^24 symbols
  static thread_local std::vector<uint32_t> keep;
  std::vector<uint32_t> own;
  std::vector<uint32_t> &tmp = n <= ((int64_t)1 << 24) ? keep : own;
  if ((int64_t)tmp.size() < n) tmp.resize((size_t)n);
  uint32_t *t = tmp.data();
  for (int l = 0; l < 64; l++) {
    if (start[l + 1] == start[l]) continue;
    Model m;
    build_model(cdfs + (int64_t)l * (A + 1), A, m, true);
    for (int64_t j = start[l]; j < start[l + 1]; j++) t[j] = dec_get(state, m, words, top);
  }
  if (top != 0 || state != 0) {
    set_error("trailing data in rANS payload");
    return LATEC_E_FORMAT;
  }
  int64_t next[64];
  for (int l = 0; l < 64; l++) next[l] = start[l];
  for (int64_t i = 0; i < n; i++) put(i, t[next[levels[i]]++]);
  return LATEC_OK;
}

int rans_decode_levels(const uint32_t *words, int64_t nw, const uint8_t *levels,
                       const uint32_t *cdfs, int64_t A, uint32_t *syms, int64_t n) {
  return decode_levels_to(words, nw, levels, cdfs, A, n,
                          [syms](int64_t i, uint32_t k) { syms[i] = k; });
}

// The same, widened back to the quantizer's codes (u16 or u32): symbol
// k != 0 is the code k + radius - 2^kexp in uint32 arithmetic, 0 stays 0
// (rans.py:83-89 inverted); a u16 code past 0xffff is a format error.
int rans_decode_codes_levels(const uint32_t *words, int64_t nw, const uint8_t *levels,
                             const uint32_t *cdfs, int64_t A, uint32_t radius, int kexp,
                             void *codes, int cb, int64_t n) {
  if ((cb != 2 && cb != 4) || kexp < 1 || kexp > 30 || A != ((int64_t)2 << kexp)) {
    set_error("rans_decode_codes_levels: code_bytes 2 or 4 and A = 2 << kexp");
    return LATEC_E_ARG;
  }
  const uint32_t off = radius - (1u << kexp);
  if (cb == 4) {
    uint32_t *c = (uint32_t *)codes;
    return decode_levels_to(words, nw, levels, cdfs, A, n,
                            [c, off](int64_t i, uint32_t k) { c[i] = k ? k + off : 0u; });
  }
  uint16_t *c = (uint16_t *)codes;
  uint32_t over = 0;
  const int rc = decode_levels_to(words, nw, levels, cdfs, A, n, [c, off, &over](int64_t i, uint32_t k) {
    const uint32_t v = k ? k + off : 0u;
    over |= v >> 16;
    c[i] = (uint16_t)v;
  });
  if (rc == LATEC_OK && over) {
    set_error("decoded code exceeds 16 bits");
    return LATEC_E_FORMAT;
  }
  return rc;
}

// ---------------------------------------------------------------------------
// records
// ---------------------------------------------------------------------------

static inline uint32_t code_at(const void *codes, int cb, int64_t i) {
  return cb == 2 ? ((const uint16_t *)codes)[i] : ((const uint32_t *)codes)[i];
}

static const int64_t LITERAL_MIN = 1 << 12;  // bitstream.py:172
static const double SCOUT_SLACK = 0.85;      // bitstream.py:175
static const int64_t PARALLEL_RANS_MIN = 1 << 20;  // stages coded beside their literal scout

// One zstd frame of src at `level`, fed to the compressor in runs of whole
// 128 KiB blocks over the one contiguous source buffer, with the parameters
// ZSTD_compress derives from (level, n) and the content size pledged: the
// frame is ZSTD_compress's one-shot frame byte for byte (same block
// boundaries, same window), but the run can stop between runs of blocks.
// Returns the frame size; ZB_OVER once `over(bytes written so far)` holds
// (the final frame is at least that long, so whatever `over` rules out is
// ruled out for the whole frame), ZB_STOPPED when *stop is raised, ZB_ERROR
// on a zstd error.
static const int64_t ZB_OVER = -1, ZB_STOPPED = -2, ZB_ERROR = -3;
static const size_t ZB_RUN = (size_t)1 << 20;  // eight blocks per call

template <class Over>
static int64_t zstd_bounded(const uint8_t *src, size_t n, int level, uint8_t *dst, size_t cap,
                            Over over, const std::atomic<bool> *stop) {
  ZSTD_CCtx *c = ZSTD_createCCtx();
  if (!c) return ZB_ERROR;
  struct Free {
    ZSTD_CCtx *c;
    ~Free() { ZSTD_freeCCtx(c); }
  } free_c{c};
  if (ZSTD_isError(ZSTD_compressBegin_advanced(c, nullptr, 0, ZSTD_getParams(level, n, 0), n)))
    return ZB_ERROR;
  size_t at = 0, off = 0;
  while (true) {
    if (stop && stop->load(std::memory_order_relaxed)) return ZB_STOPPED;
    const size_t k = std::min(n - off, ZB_RUN);
    const bool last = off + k == n;
    const size_t w = last ? ZSTD_compressEnd(c, dst + at, cap - at, src + off, k)
                          : ZSTD_compressContinue(c, dst + at, cap - at, src + off, k);
    if (ZSTD_isError(w)) return ZB_ERROR;
    at += w;
    off += k;
    if (over(at)) return ZB_OVER;
    if (last) return (int64_t)at;
  }
}

// Header and tail of a stage record around a blob already at out + 12:
// [n u32][hlen u45 | kexp u4 | shape u5 | level u6 | owx u3 | literal u1]
// [blob][n_out u32][outliers f32].  Returns the record length.
static int64_t put_record(uint8_t *out, int64_t n, int64_t hlen, int kexp, const latec_fit *fit,
                          int literal, const float *outl, int64_t n_out) {
  const uint64_t framed = (uint64_t)hlen | ((uint64_t)kexp << 45) |
                          ((uint64_t)fit->shape << 49) | ((uint64_t)fit->level << 54) |
                          ((uint64_t)fit->owx << 60) | ((uint64_t)literal << 63);
  const uint32_t nc = (uint32_t)n;
  memcpy(out, &nc, 4);
  memcpy(out + 4, &framed, 8);
  uint8_t *p = out + 12 + hlen;
  const uint32_t no = (uint32_t)n_out;
  memcpy(p, &no, 4);
  if (n_out) memcpy(p + 4, outl, 4 * (size_t)n_out);
  return 12 + hlen + 4 + 4 * n_out;
}

int64_t pack_stage(const void *codes, int cb, int64_t n, const float *outl, int64_t n_out,
                   uint32_t R, int zstd_level, const latec_fit *fit, const uint32_t *cdf_row,
                   uint8_t *out, int64_t cap, const RecordHook *rans_record) {
  const int kexp = fit->kexp;
  if (kexp < 1 || kexp > 15) {
    set_error("rANS coder window does not fit the stage frame");
    return LATEC_E_ARG;
  }
  const int64_t A = (int64_t)2 << kexp;
  const int64_t head = 12, tail = 4 + 4 * n_out;
  if (cap < head + tail) {
    set_error("record capacity");
    return LATEC_E_CAPACITY;
  }
  uint8_t *blob = out + head;
  const int64_t blob_cap = cap - head - tail;
  int64_t hlen = -1;
  int literal = 0;
  const bool wide = kexp > 7;
  const bool big = n >= LITERAL_MIN && n >= PARALLEL_RANS_MIN;
  const size_t lit_bytes = n >= LITERAL_MIN ? (size_t)n * (wide ? 2 : 1) : 0;
  // buffers without value-initialisation: every byte used is written first
  std::unique_ptr<uint32_t[]> syms(new uint32_t[(size_t)std::max<int64_t>(n, 1)]);
  std::unique_ptr<uint8_t[]> lit(new uint8_t[std::max<size_t>(lit_bytes, 1)]);
  // narrow [lo, hi) into syms (and the literal bytes); false on a code outside
  // the window
  auto narrow = [&](int64_t lo, int64_t hi) {
    bool ok = true;
    for (int64_t i = lo; i < hi; i++) {
      const uint32_t s = narrow_sym(code_at(codes, cb, i), R, kexp);
      ok &= s < (uint64_t)A;
      syms[i] = s;
      if (lit_bytes) {
        if (wide)
          ((uint16_t *)lit.get())[i] = (uint16_t)s;
        else
          lit[i] = (uint8_t)s;
      }
    }
    return ok;
  };
  bool in_window = true;
  if (big) {  // four threads over four runs
    const int T = 4;
    bool oks[T];
    std::thread nt[T - 1];
    {
      JoinAll guard{nt, T - 1};
      for (int t = 1; t < T; t++)
        start_or_run(nt[t - 1], [&, t] { oks[t] = narrow(n * t / T, n * (t + 1) / T); });
      oks[0] = narrow(0, n / T);
    }
    for (int t = 0; t < T; t++) in_window &= oks[t];
  } else {
    in_window = narrow(0, n);
  }
  if (!in_window) {
    set_error("code outside the narrowed rANS alphabet");
    return LATEC_E_ARG;
  }
  // A large stage also runs the full-level literal frame beside its scout
  // from the start: the scout stops as soon as its output rules the literal
  // out (scout * slack >= bits / 8) and then stops the full run; the full run
  // stops as soon as its own output does (full * 8 >= bits).  Both are the
  // one-shot frames when they finish, so the record is the one the serial
  // scout-then-full order writes.
  std::unique_ptr<uint8_t[]> full_buf;
  int64_t full_len = ZB_STOPPED;
  std::atomic<bool> stop_full(false);
  std::thread full_thread, rans_thread;
  // joined on every return path, and built before either thread starts
  JoinAll join_full{&full_thread, 1}, join_rans{&rans_thread, 1};
  const double bits = (double)fit->bits;
  auto full_over = [bits](size_t at) { return (double)at * 8.0 >= bits; };
  if (big)
    start_or_run(full_thread, [&] {
      const size_t bound = ZSTD_compressBound(lit_bytes);
      full_buf.reset(new uint8_t[bound]);
      full_len = zstd_bounded(lit.get(), lit_bytes, zstd_level, full_buf.get(), bound, full_over,
                              &stop_full);
    });
  // A large stage codes its rANS words on a further thread while the literal
  // frames run: the record is the same, the stage's critical path is the
  // longest of the three instead of their sum.
  std::unique_ptr<uint32_t[]> words;
  int64_t nw = -1;
  if (big && fit->level >= 0 && fit->level < 64)
    start_or_run(rans_thread, [&] {
      words.reset(new uint32_t[(size_t)enc_cap(n)]);
      nw = rans_encode(syms.get(), n, cdf_row, A, words.get(), enc_cap(n));
      // the record this stage writes unless its literal frame wins, handed
      // out before the literal frames end
      if (rans_record && nw >= 0 && nw * 4 <= blob_cap) {
        std::unique_ptr<uint8_t[]> rec(new uint8_t[(size_t)(head + nw * 4 + tail)]);
        memcpy(rec.get() + 12, words.get(), (size_t)nw * 4);
        (*rans_record)(rec.get(), put_record(rec.get(), n, nw * 4, kexp, fit, 0, outl, n_out));
      }
    });
  if (big) {
    const size_t bound = ZSTD_compressBound(lit_bytes);
    std::unique_ptr<uint8_t[]> tmp(new uint8_t[bound]);
    const int64_t scout = zstd_bounded(
        lit.get(), lit_bytes, 1, tmp.get(), bound,
        [bits](size_t at) { return (double)at * SCOUT_SLACK >= bits / 8.0; }, nullptr);
    if (scout < 0) stop_full.store(true, std::memory_order_relaxed);
    if (full_thread.joinable()) full_thread.join();
    if (scout >= 0 && full_len >= 0) {
      if (full_len > blob_cap) {
        set_error("record capacity");
        return LATEC_E_CAPACITY;
      }
      memcpy(blob, full_buf.get(), (size_t)full_len);
      hlen = full_len;
      literal = 1;
    }
  } else if (n >= LITERAL_MIN) {
    std::vector<uint8_t> tmp(ZSTD_compressBound(lit_bytes));
    const size_t scout = ZSTD_compress(tmp.data(), tmp.size(), lit.get(), lit_bytes, 1);
    if (!ZSTD_isError(scout) && (double)scout * SCOUT_SLACK < (double)fit->bits / 8.0) {
      const size_t full =
          ZSTD_compress(tmp.data(), tmp.size(), lit.get(), lit_bytes, zstd_level);
      if (!ZSTD_isError(full) && (double)full * 8.0 < (double)fit->bits) {
        if ((int64_t)full > blob_cap) {
          set_error("record capacity");
          return LATEC_E_CAPACITY;
        }
        memcpy(blob, tmp.data(), full);
        hlen = (int64_t)full;
        literal = 1;
      }
    }
  }
  if (hlen < 0 && n == 0) hlen = 0;  // empty stage: empty word stream
  if (hlen < 0) {
    if (fit->level < 0 || fit->level >= 64) {
      set_error("scale level exceeds table count");
      return LATEC_E_ARG;
    }
    if (rans_thread.joinable()) rans_thread.join();
    if (nw < 0) {  // not coded yet, or failed on the helper thread: code (or fail) here
      words.reset(new uint32_t[(size_t)enc_cap(n)]);
      nw = rans_encode(syms.get(), n, cdf_row, A, words.get(), enc_cap(n));
    }
    if (nw < 0) return nw;
    if (nw * 4 > blob_cap) {
      set_error("record capacity");
      return LATEC_E_CAPACITY;
    }
    memcpy(blob, words.get(), (size_t)nw * 4);
    hlen = nw * 4;
  }
  return put_record(out, n, hlen, kexp, fit, literal, outl, n_out);
}

int stage_header(const uint8_t *buf, int64_t len, uint32_t *n_codes, int64_t *hlen,
                 int *kexp, int *shape, int *level, int *owx, int *literal) {
  if (len < 12) {
    set_error("truncated stage record");
    return LATEC_E_FORMAT;
  }
  uint64_t framed;
  memcpy(n_codes, buf, 4);
  memcpy(&framed, buf + 4, 8);
  *hlen = (int64_t)(framed & (((uint64_t)1 << 45) - 1));
  *kexp = (int)((framed >> 45) & 15);
  *shape = (int)((framed >> 49) & 31);
  *level = (int)((framed >> 54) & 63);
  *owx = (int)((framed >> 60) & 7);
  *literal = (int)(framed >> 63);
  return 12;
}

int64_t unpack_stage(const uint8_t *buf, int64_t len, uint32_t R, const uint32_t *cdf_row,
                     int64_t n_expect, void *codes, int cb, uint32_t *out_rank,
                     float *out_val, int64_t out_cap, int64_t *n_out) {
  uint32_t n;
  int64_t hlen;
  int kexp, shape, level, owx, literal;
  int rc = stage_header(buf, len, &n, &hlen, &kexp, &shape, &level, &owx, &literal);
  if (rc < 0) return rc;
  if (!kexp) {
    set_error("rANS stage record has no coder window");
    return LATEC_E_FORMAT;
  }
  if ((int64_t)n != n_expect) {
    set_error("stage code count mismatch");
    return LATEC_E_COUNT;
  }
  if (12 + hlen + 4 > len) {
    set_error("truncated stage record");
    return LATEC_E_FORMAT;
  }
  const uint8_t *blob = buf + 12;
  const int64_t A = (int64_t)2 << kexp;
  // each decoded symbol is widened into `codes` at once; the ranks of the
  // outlier codes (0) are kept up to out_cap and counted
  int64_t k = 0;
  auto emit = [&](int64_t i, uint32_t sym) {
    const uint32_t c = widen_sym(sym, R, kexp);
    if (cb == 2)
      ((uint16_t *)codes)[i] = (uint16_t)c;
    else
      ((uint32_t *)codes)[i] = c;
    if (c == 0) {
      if (k < out_cap) out_rank[k] = (uint32_t)i;
      k++;
    }
  };
  if (literal) {
    const bool wide = kexp > 7;
    const size_t sb = (size_t)n * (wide ? 2 : 1);
    std::unique_ptr<uint8_t[]> lit(new uint8_t[sb ? sb : 1]);
    const size_t got = ZSTD_decompress(lit.get(), sb, blob, (size_t)hlen);
    if (ZSTD_isError(got) || got < sb) {
      set_error("bad literal stage frame");
      return LATEC_E_FORMAT;
    }
    for (uint32_t i = 0; i < n; i++) emit(i, wide ? ((const uint16_t *)lit.get())[i] : lit[i]);
  } else if (n == 0) {
    if (hlen) {
      set_error("non-empty rANS blob for an empty stage");
      return LATEC_E_FORMAT;
    }
  } else {
    if (hlen % 4) {
      set_error("rANS payload is not aligned to 32-bit words");
      return LATEC_E_FORMAT;
    }
    std::unique_ptr<uint32_t[]> words(new uint32_t[(size_t)hlen / 4]);
    memcpy(words.get(), blob, (size_t)hlen);
    rc = rans_decode_to(words.get(), hlen / 4, cdf_row, A, n, emit);
    if (rc) return rc;
  }
  uint32_t no;
  memcpy(&no, blob + hlen, 4);
  if (12 + hlen + 4 + 4 * (int64_t)no > len || (int64_t)no > out_cap) {
    set_error("truncated stage record");
    return LATEC_E_FORMAT;
  }
  if (k != (int64_t)no) {
    set_error("outlier count mismatch: " + std::to_string(k) + " zeros vs " +
              std::to_string(no) + " values");
    return LATEC_E_COUNT;
  }
  memcpy(out_val, blob + hlen + 4, 4 * (size_t)no);
  *n_out = no;
  return 12 + hlen + 4 + 4 * (int64_t)no;
}

// Every stage record of a payload, stages on a pool of `threads` workers
// (largest first), concatenated in stage order.  With frame_level >= 0 the
// calling thread also writes the zstd frame of the payload at that level
// (the one-shot frame, see zstd_bounded) and starts it early: once every
// stage has either its final record or its rANS record (a large stage whose
// literal frames are still running), it frames that payload; if a large
// stage's final record then differs (its literal frame won), the payload is
// rebuilt from the final records and framed again.  The outputs are those of
// packing everything first and framing after.
int64_t pack_stages(int ns, const void *codes, int cb, const int64_t *code_off,
                    const int64_t *n_codes, const float *outl, const int64_t *out_off,
                    const int64_t *n_out, uint32_t R, int zstd_level, const latec_fit *fits,
                    const uint32_t *cdf_rows, const int64_t *row_off, uint8_t *out,
                    int64_t cap, int64_t *rec_len, int threads, int frame_level,
                    uint8_t *frame, int64_t frame_cap, int64_t *frame_len) {
  const bool framing = frame_level >= 0;
  std::vector<std::unique_ptr<uint8_t[]>> recs(ns), spec(ns);
  std::vector<int64_t> rcs(ns, 0), spec_len(ns, -1);
  std::vector<char> done(ns, 0);
  std::vector<std::string> errs(ns);
  std::vector<int> order(ns);
  for (int i = 0; i < ns; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return n_codes[a] > n_codes[b]; });
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<bool> spec_broken(false);
  std::atomic<int> next(0);
  auto work = [&]() {
    while (true) {
      const int j = next.fetch_add(1);
      if (j >= ns) break;
      const int s = order[j];
      const int64_t rc_cap = 16 + 4 * n_out[s] + 4 * (n_codes[s] + 2) + 64;
      recs[s].reset(new uint8_t[(size_t)rc_cap]);
      const void *cs = (const uint8_t *)codes + (size_t)cb * code_off[s];
      RecordHook hook = [&, s](const uint8_t *rec, int64_t len) {
        std::unique_ptr<uint8_t[]> copy(new uint8_t[(size_t)len]);
        memcpy(copy.get(), rec, (size_t)len);
        std::lock_guard<std::mutex> g(mu);
        spec[s] = std::move(copy);
        spec_len[s] = len;
        cv.notify_all();
      };
      const int64_t rc = pack_stage(cs, cb, n_codes[s], outl + out_off[s], n_out[s], R,
                                    zstd_level, &fits[s], cdf_rows + row_off[s], recs[s].get(),
                                    rc_cap, framing ? &hook : nullptr);
      // the message was set on this worker's thread: keep it for the caller
      if (rc < 0) errs[s] = thread_error();
      // spec[s] was set (if at all) on a thread pack_stage joined
      const bool differs = spec_len[s] >= 0 &&
                           (rc != spec_len[s] || memcmp(recs[s].get(), spec[s].get(), rc));
      std::lock_guard<std::mutex> g(mu);
      rcs[s] = rc;
      done[s] = 1;
      if (differs) spec_broken.store(true);
      cv.notify_all();
    }
  };
  if (threads < 1) threads = 1;
  threads = std::min(threads, ns);
  std::vector<std::thread> pool(threads);
  int started = 0;
  JoinAll join_pool{pool.data(), threads};
  try {
    for (int t = framing ? 0 : 1; t < threads; t++, started++) pool[t] = std::thread(work);
  } catch (const std::system_error &) {
    // fewer workers: they take stages from the shared counter, so the
    // ones that started (or this thread, below) finish the rest
  }
  if (!framing || !started) work();
  // payload of stage records in stage order: the final record of every
  // stage in `fin`, else its rANS record; false on a stage error or capacity
  int64_t size = 0;
  auto assemble = [&](const std::vector<char> &fin) -> int64_t {
    int64_t off = 0;
    for (int s = 0; s < ns; s++) {
      const int64_t len = fin[s] ? rcs[s] : spec_len[s];
      if (len < 0) {
        set_error("stage " + std::to_string(s) + ": " +
                  (errs[s].empty() ? std::string("pack failed") : errs[s]));
        return len;
      }
      if (off + len > cap) {
        set_error("payload capacity");
        return LATEC_E_CAPACITY;
      }
      memcpy(out + off, fin[s] ? recs[s].get() : spec[s].get(), (size_t)len);
      rec_len[s] = len;
      off += len;
    }
    return off;
  };
  int64_t flen = ZB_STOPPED;
  if (framing) {
    std::vector<char> fin(ns);
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] {
        for (int s = 0; s < ns; s++)
          if (!done[s] && spec_len[s] < 0) return false;
        return true;
      });
      fin = done;  // records fixed now; the others are rANS records
    }
    size = spec_broken.load() ? -1 : assemble(fin);
    if (size >= 0)
      flen = zstd_bounded(out, (size_t)size, frame_level, frame, (size_t)frame_cap,
                          [](size_t) { return false; }, &spec_broken);
  }
  for (auto &t : pool)
    if (t.joinable()) t.join();
  if (framing && (flen < 0 || spec_broken.load())) {  // frame again from the final records
    size = assemble(std::vector<char>(ns, 1));
    if (size < 0) return size;
    flen = zstd_bounded(out, (size_t)size, frame_level, frame, (size_t)frame_cap,
                        [](size_t) { return false; }, nullptr);
    if (flen < 0) {
      set_error("payload frame capacity");
      return LATEC_E_CAPACITY;
    }
  } else if (!framing) {
    size = assemble(std::vector<char>(ns, 1));
  }
  if (framing) *frame_len = flen;
  return size;
}

int64_t zstd_compress(const uint8_t *src, int64_t n, int level, uint8_t *dst, int64_t cap) {
  const size_t r = ZSTD_compress(dst, (size_t)cap, src, (size_t)n, level);
  if (ZSTD_isError(r)) {
    set_error(std::string("zstd: ") + ZSTD_getErrorName(r));
    return LATEC_E_CAPACITY;
  }
  return (int64_t)r;
}

int64_t zstd_content_size(const uint8_t *src, int64_t n) {
  const unsigned long long r = ZSTD_getFrameContentSize(src, (size_t)n);
  if (r == ZSTD_CONTENTSIZE_ERROR || r == ZSTD_CONTENTSIZE_UNKNOWN) {
    set_error("not a zstd frame with a content size");
    return LATEC_E_FORMAT;
  }
  return (int64_t)r;
}

int64_t zstd_decompress(const uint8_t *src, int64_t n, uint8_t *dst, int64_t cap) {
  const size_t r = ZSTD_decompress(dst, (size_t)cap, src, (size_t)n);
  if (ZSTD_isError(r)) {
    set_error(std::string("zstd: ") + ZSTD_getErrorName(r));
    return LATEC_E_FORMAT;
  }
  return (int64_t)r;
}

}  // namespace latec
