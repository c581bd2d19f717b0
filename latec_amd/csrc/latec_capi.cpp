// latec_capi.cpp -- the extern "C" boundary declared in include/latec.h.
#include <hip/hip_runtime.h>
#include <zstd.h>

#include <cstring>
#include <string>

#include "latec_plan.h"

namespace latec {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
const std::string &thread_error() { return g_err; }
void set_hip_error(const char *what, int code) {
  g_err = std::string(what) + ": " + hipGetErrorString((hipError_t)code);
}

// latec_interp.hip
int plan_upload(latec_plan *p);
void plan_release(latec_plan *p);
int interp_encode(latec_plan *, int, int, const float *, float *, void *, uint32_t *,
                  uint32_t *, float *, uint32_t *, const double *, const latec_interp_opts *,
                  hipStream_t);
int interp_decode(latec_plan *, int, int, const void *, const uint32_t *, const float *,
                  const uint32_t *, float *, const double *, const latec_interp_opts *,
                  hipStream_t);
int interp_encode_graph(latec_plan *, const float *, float *, void *, uint32_t *, uint32_t *,
                        float *, uint32_t *, const double *, const latec_interp_opts *,
                        hipStream_t);
int64_t interp_stats_words(latec_plan *);
int64_t minmax_words(int);
int minmax(int, const float *, const int64_t *, const int64_t *, float *, float *, int64_t,
           hipStream_t);
int interp_stats(latec_plan *, const float *, const float *, const latec_interp_opts *, double *,
                 double *, int64_t, hipStream_t);
int interp_predict(latec_plan *, int, const float *, float *, const latec_interp_opts *,
                   hipStream_t);
int code_hist(const void *, int64_t, int, uint32_t, uint32_t *, hipStream_t);
int quantize(const float *, const float *, int64_t, double, const latec_interp_opts *, void *,
             uint32_t *, float *, uint32_t *, uint32_t *, hipStream_t);
int dequantize(const float *, const void *, int64_t, const uint32_t *, const float *,
               const uint32_t *, double, const latec_interp_opts *, float *, hipStream_t);
// latec_tables.hip
int level_weights(int, double *, hipStream_t);
int model_weights(int, int, int, double *, hipStream_t);
int neg_log2_weights(int, int, int, int, float *, hipStream_t);
int build_cdfs(int, int, int, int, int, int, uint32_t *, hipStream_t);
int fit_hists(int, const uint32_t *, const void *, const int64_t *, const int64_t *, uint32_t,
              int, int, latec_fit *, hipStream_t);
int scale_to_level(const float *, int64_t, const float *, uint8_t *, hipStream_t);
int64_t hist_words(uint32_t R);
// latec_entropy.cpp
int64_t rans_encode(const uint32_t *, int64_t, const uint32_t *, int64_t, uint32_t *, int64_t);
int64_t rans_encode_levels(const uint32_t *, const uint8_t *, int64_t, const uint32_t *,
                           int64_t, uint32_t *, int64_t);
int64_t rans_encode_codes_levels(const void *, int, const uint8_t *, int64_t, uint32_t, int,
                                 const uint32_t *, int64_t, uint32_t *, int64_t);
int code_kexp(const void *, int, int64_t, uint32_t);
int rans_decode_codes_levels(const uint32_t *, int64_t, const uint8_t *, const uint32_t *, int64_t,
                             uint32_t, int, void *, int, int64_t);
int rans_decode(const uint32_t *, int64_t, const uint32_t *, int64_t, uint32_t *, int64_t);
int rans_decode_levels(const uint32_t *, int64_t, const uint8_t *, const uint32_t *, int64_t,
                       uint32_t *, int64_t);
int64_t pack_stage(const void *, int, int64_t, const float *, int64_t, uint32_t, int,
                   const latec_fit *, const uint32_t *, uint8_t *, int64_t, const RecordHook *);
int stage_header(const uint8_t *, int64_t, uint32_t *, int64_t *, int *, int *, int *, int *,
                 int *);
int64_t unpack_stage(const uint8_t *, int64_t, uint32_t, const uint32_t *, int64_t, void *,
                     int, uint32_t *, float *, int64_t, int64_t *);
int64_t pack_stages(int, const void *, int, const int64_t *, const int64_t *, const float *,
                    const int64_t *, const int64_t *, uint32_t, int, const latec_fit *,
                    const uint32_t *, const int64_t *, uint8_t *, int64_t, int64_t *, int, int,
                    uint8_t *, int64_t, int64_t *);
int64_t zstd_compress(const uint8_t *, int64_t, int, uint8_t *, int64_t);
int64_t zstd_content_size(const uint8_t *, int64_t);
int64_t zstd_decompress(const uint8_t *, int64_t, uint8_t *, int64_t);
int tune_set(const char *, int64_t);
int64_t tune_get(const char *);

}  // namespace latec

using namespace latec;

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

extern "C" {

const char *latec_last_error(void) { return g_err.c_str(); }
int latec_abi_version(void) { return 1; }

int latec_tune_set(const char *name, int64_t value) {
  if (!name) {
    set_error("latec_tune_set: null name");
    return LATEC_E_ARG;
  }
  return tune_set(name, value);
}

int64_t latec_tune_get(const char *name) {
  if (!name) {
    set_error("latec_tune_get: null name");
    return LATEC_E_ARG;
  }
  return tune_get(name);
}

int latec_plan_create(int ndim, const int64_t *shape, int channels, int levels,
                      int anchor_stride, int anchor_block, latec_plan **out) {
  if (!out || !shape || ndim < 0 || ndim > LATEC_MAX_DIM || channels < 1) {
    set_error(ndim < 1 ? "shape must have at least one axis" : "bad plan arguments");
    return LATEC_E_ARG;
  }
  auto *p = new latec_plan();
  p->ndim = ndim;
  p->channels = channels;
  p->levels = levels;
  p->stride = anchor_stride;
  p->block = anchor_block;
  for (int j = 0; j < ndim; j++) p->shape[j] = shape[j];
  int rc = plan_build(p);
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return LATEC_OK;
}

void latec_plan_destroy(latec_plan *p) {
  if (!p) return;
  plan_release(p);
  delete p;
}

int latec_plan_n_stages(const latec_plan *p) { return (int)p->stages.size(); }

int latec_plan_stage_info(const latec_plan *p, int stage, int64_t *count, int *stride,
                          int *weight, int *n_classes) {
  if (stage < 0 || stage >= (int)p->stages.size()) {
    set_error("stage out of range");
    return LATEC_E_ARG;
  }
  const auto &st = p->stages[stage];
  *count = st.count;
  *stride = st.stride;
  *weight = st.weight;
  *n_classes = (int)st.classes.size();
  return LATEC_OK;
}

int latec_plan_class_info(const latec_plan *p, int stage, int cls, int64_t *out) {
  if (stage < 0 || stage >= (int)p->stages.size() || cls < 0 ||
      cls >= (int)p->stages[stage].classes.size()) {
    set_error("class out of range");
    return LATEC_E_ARG;
  }
  const ClassDesc &c = p->stages[stage].classes[cls];
  out[0] = c.parity;
  out[1] = c.step;
  out[2] = c.offset;
  out[3] = c.size;
  for (int j = 0; j < p->ndim; j++) {
    out[4 + j] = c.starts[j];
    out[4 + p->ndim + j] = c.sizes[j];
    out[4 + 2 * p->ndim + j] = c.rstr[j];
  }
  return LATEC_OK;
}

int64_t latec_plan_stage_offset(const latec_plan *p, int stage) {
  if (stage < 0 || stage > (int)p->stages.size()) {
    set_error("stage out of range");
    return LATEC_E_ARG;
  }
  if (stage == (int)p->stages.size()) return p->npoints;
  return p->stages[stage].offset;
}

int latec_interp_encode(latec_plan *p, int s0, int s1, const float *x, float *recon,
                        void *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                        uint32_t *out_count, const double *ebs, const latec_interp_opts *o,
                        void *stream) {
  return interp_encode(p, s0, s1, x, recon, codes, hist, out_rank, out_val, out_count, ebs, o,
                       S(stream));
}

int latec_interp_decode(latec_plan *p, int s0, int s1, const void *codes,
                        const uint32_t *out_rank, const float *out_val,
                        const uint32_t *out_count, float *recon, const double *ebs,
                        const latec_interp_opts *o, void *stream) {
  return interp_decode(p, s0, s1, codes, out_rank, out_val, out_count, recon, ebs, o,
                       S(stream));
}

int latec_interp_encode_graph(latec_plan *p, const float *x, float *recon, void *codes,
                              uint32_t *hist, uint32_t *out_rank, float *out_val,
                              uint32_t *out_count, const double *ebs,
                              const latec_interp_opts *o, void *stream) {
  return interp_encode_graph(p, x, recon, codes, hist, out_rank, out_val, out_count, ebs, o,
                             S(stream));
}

int latec_interp_predict(latec_plan *p, int stage, const float *d_recon, float *d_pred,
                         const latec_interp_opts *o, void *stream) {
  return interp_predict(p, stage, d_recon, d_pred, o, S(stream));
}

int64_t latec_interp_stats_words(latec_plan *p) { return interp_stats_words(p); }

int64_t latec_minmax_words(int nseg) { return minmax_words(nseg); }

int latec_minmax(int nseg, const float *d_x, const int64_t *off, const int64_t *len,
                 float *d_out, float *d_scratch, int64_t scratch_words, void *stream) {
  return minmax(nseg, d_x, off, len, d_out, d_scratch, scratch_words, S(stream));
}

int latec_interp_stats(latec_plan *p, const float *d_x, const float *d_recon,
                       const latec_interp_opts *o, double *d_out, double *d_scratch,
                       int64_t scratch_words, void *stream) {
  return interp_stats(p, d_x, d_recon, o, d_out, d_scratch, scratch_words, S(stream));
}

int latec_code_hist(const void *d_codes, int64_t n, int code_bytes, uint32_t radius,
                    uint32_t *d_hist, void *stream) {
  return code_hist(d_codes, n, code_bytes, radius, d_hist, S(stream));
}

int latec_quantize(const float *x, const float *pred, int64_t n, double eb,
                   const latec_interp_opts *o, void *codes, uint32_t *out_rank,
                   float *out_val, uint32_t *out_count, uint32_t *hist, void *stream) {
  return quantize(x, pred, n, eb, o, codes, out_rank, out_val, out_count, hist, S(stream));
}

int latec_dequantize(const float *pred, const void *codes, int64_t n, const uint32_t *out_rank,
                     const float *out_val, const uint32_t *out_count, double eb,
                     const latec_interp_opts *o, float *recon, void *stream) {
  return dequantize(pred, codes, n, out_rank, out_val, out_count, eb, o, recon, S(stream));
}

int latec_level_weights(int kexp, double *d_out, void *stream) {
  return level_weights(kexp, d_out, S(stream));
}
int latec_model_weights(int kexp, int shape, int owx, double *d_out, void *stream) {
  return model_weights(kexp, shape, owx, d_out, S(stream));
}
int latec_neg_log2_weights(int kexp, int shape, int owx, int precision, float *d_out,
                           void *stream) {
  return neg_log2_weights(kexp, shape, owx, precision, d_out, S(stream));
}
int latec_build_cdfs(int kexp, int shape, int precision, int owx, int l0, int l1,
                     uint32_t *d_out, void *stream) {
  return build_cdfs(kexp, shape, precision, owx, l0, l1, d_out, S(stream));
}

int latec_fit_const(latec_plan *p, int s0, int s1, const uint32_t *d_hist, const void *d_codes,
                    uint32_t radius, int code_bytes, int precision, latec_fit *fits,
                    void *stream) {
  if (s0 < 0 || s1 > (int)p->stages.size() || s0 > s1) {
    set_error("stage range out of bounds");
    return LATEC_E_ARG;
  }
  std::vector<int64_t> off, n;
  for (int s = s0; s < s1; s++) {
    off.push_back((int64_t)p->channels * p->stages[s].offset);
    n.push_back((int64_t)p->channels * p->stages[s].count);
  }
  return fit_hists(s1 - s0, d_hist + (int64_t)s0 * hist_words(radius), d_codes, off.data(), n.data(),
                   radius, code_bytes, precision, fits, S(stream));
}

int latec_fit_stages(int nst, const uint32_t *d_hist, const void *d_codes,
                     const int64_t *code_off, const int64_t *n_codes, uint32_t radius,
                     int code_bytes, int precision, latec_fit *fits, void *stream) {
  if (nst < 0 || (nst && (!d_hist || !code_off || !n_codes || !fits))) {
    set_error("latec_fit_stages: bad arguments");
    return LATEC_E_ARG;
  }
  return fit_hists(nst, d_hist, d_codes, code_off, n_codes, radius, code_bytes, precision, fits,
                   S(stream));
}

int latec_fit_codes(const uint32_t *d_hist, const void *d_codes, int64_t n, uint32_t radius,
                    int code_bytes, int precision, latec_fit *fit, void *stream) {
  int64_t off = 0;
  return fit_hists(1, d_hist, d_codes, &off, &n, radius, code_bytes, precision, fit,
                   S(stream));
}

int64_t latec_hist_words(uint32_t radius) { return hist_words(radius); }

int latec_scale_to_level(const float *d_scale, int64_t n, const float *thresholds,
                         uint8_t *d_levels, void *stream) {
  return scale_to_level(d_scale, n, thresholds, d_levels, S(stream));
}

int64_t latec_rans_encode(const uint32_t *syms, int64_t n, const uint32_t *cdf, int64_t A,
                          uint32_t *words, int64_t cap) {
  return rans_encode(syms, n, cdf, A, words, cap);
}
int64_t latec_rans_encode_levels(const uint32_t *syms, const uint8_t *levels, int64_t n,
                                 const uint32_t *cdfs, int64_t A, uint32_t *words,
                                 int64_t cap) {
  return rans_encode_levels(syms, levels, n, cdfs, A, words, cap);
}
int64_t latec_rans_encode_codes_levels(const void *codes, int code_bytes,
                                       const uint8_t *levels, int64_t n, uint32_t radius,
                                       int kexp, const uint32_t *cdfs, int64_t A,
                                       uint32_t *words, int64_t cap) {
  return rans_encode_codes_levels(codes, code_bytes, levels, n, radius, kexp, cdfs, A, words,
                                  cap);
}
int latec_code_kexp(const void *codes, int code_bytes, int64_t n, uint32_t radius) {
  if (code_bytes != 2 && code_bytes != 4) {
    set_error("latec_code_kexp: code_bytes must be 2 or 4");
    return LATEC_E_ARG;
  }
  return code_kexp(codes, code_bytes, n, radius);
}
int latec_rans_decode_codes_levels(const uint32_t *words, int64_t nw, const uint8_t *levels,
                                   const uint32_t *cdfs, int64_t A, uint32_t radius, int kexp,
                                   void *codes, int code_bytes, int64_t n) {
  return rans_decode_codes_levels(words, nw, levels, cdfs, A, radius, kexp, codes, code_bytes,
                                  n);
}
int latec_rans_decode(const uint32_t *words, int64_t nw, const uint32_t *cdf, int64_t A,
                      uint32_t *syms, int64_t n) {
  return rans_decode(words, nw, cdf, A, syms, n);
}
int latec_rans_decode_levels(const uint32_t *words, int64_t nw, const uint8_t *levels,
                             const uint32_t *cdfs, int64_t A, uint32_t *syms, int64_t n) {
  return rans_decode_levels(words, nw, levels, cdfs, A, syms, n);
}

int64_t latec_pack_stage(const void *codes, int code_bytes, int64_t n, const float *outl,
                         int64_t n_out, uint32_t radius, int zstd_level, const latec_fit *fit,
                         const uint32_t *cdf_row, uint8_t *out, int64_t cap) {
  return pack_stage(codes, code_bytes, n, outl, n_out, radius, zstd_level, fit, cdf_row, out,
                    cap, nullptr);
}
int latec_stage_header(const uint8_t *buf, int64_t len, uint32_t *n_codes, int64_t *hlen,
                       int *kexp, int *shape, int *level, int *owx, int *literal) {
  return stage_header(buf, len, n_codes, hlen, kexp, shape, level, owx, literal);
}
int64_t latec_unpack_stage(const uint8_t *buf, int64_t len, uint32_t radius,
                           const uint32_t *cdf_row, int64_t n_expect, void *codes,
                           int code_bytes, uint32_t *out_rank, float *out_val, int64_t out_cap,
                           int64_t *n_out) {
  return unpack_stage(buf, len, radius, cdf_row, n_expect, codes, code_bytes, out_rank,
                      out_val, out_cap, n_out);
}
int64_t latec_pack_stages(int n_stages, const void *codes, int code_bytes,
                          const int64_t *code_off, const int64_t *n_codes, const float *outl,
                          const int64_t *out_off, const int64_t *n_out, uint32_t radius,
                          int zstd_level, const latec_fit *fits, const uint32_t *cdf_rows,
                          const int64_t *row_off, uint8_t *out, int64_t cap, int64_t *rec_len,
                          int threads) {
  return pack_stages(n_stages, codes, code_bytes, code_off, n_codes, outl, out_off, n_out,
                     radius, zstd_level, fits, cdf_rows, row_off, out, cap, rec_len, threads, -1,
                     nullptr, 0, nullptr);
}
int64_t latec_pack_stages_framed(int n_stages, const void *codes, int code_bytes,
                                 const int64_t *code_off, const int64_t *n_codes,
                                 const float *outl, const int64_t *out_off, const int64_t *n_out,
                                 uint32_t radius, int zstd_level, const latec_fit *fits,
                                 const uint32_t *cdf_rows, const int64_t *row_off, uint8_t *out,
                                 int64_t cap, int64_t *rec_len, int threads, int frame_level,
                                 uint8_t *frame, int64_t frame_cap, int64_t *frame_len) {
  if (frame_level < 1 || frame_level > 22 || !frame || !frame_len) {
    set_error("pack_stages_framed: frame level 1..22 and a frame buffer are required");
    return LATEC_E_ARG;
  }
  return pack_stages(n_stages, codes, code_bytes, code_off, n_codes, outl, out_off, n_out,
                     radius, zstd_level, fits, cdf_rows, row_off, out, cap, rec_len, threads,
                     frame_level, frame, frame_cap, frame_len);
}
int64_t latec_zstd_bound(int64_t n) { return (int64_t)ZSTD_compressBound((size_t)n); }
int64_t latec_zstd_compress(const uint8_t *src, int64_t n, int level, uint8_t *dst,
                            int64_t cap) {
  return zstd_compress(src, n, level, dst, cap);
}
int64_t latec_zstd_content_size(const uint8_t *src, int64_t n) {
  return zstd_content_size(src, n);
}
int64_t latec_zstd_decompress(const uint8_t *src, int64_t n, uint8_t *dst, int64_t cap) {
  return zstd_decompress(src, n, dst, cap);
}

}  // extern "C"
