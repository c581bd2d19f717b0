/*
 * latec.h -- C ABI of the MI355X-native LATEC codec hot path (liblatec.so).
 *
 * Plain pointers and sizes only.  "d_" pointers are device (HBM) pointers,
 * all others are host pointers.  `stream` is a hipStream_t (NULL = default
 * stream).  Every function returns 0 (or a non-negative size) on success and
 * a negative LATEC_E_* code on failure; latec_last_error() then holds a
 * message (thread-local).  Callers own every buffer they pass in; the library
 * never frees caller memory.  Device work is asynchronous on `stream`.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   plan / stage geometry ..... latec/levels.py:100-147, 221-282, 297-386, 410-574
 *   interp encode/decode ...... latec/predictor.py:75-317 fused with
 *                               latec/quantizer.py:28-120 and the stage loop
 *                               latec/codec.py:384-510
 *   quantize / dequantize ..... latec/quantizer.py:28-120
 *   tables / fit .............. latec/rans.py:162-180, 247-267, 270-278,
 *                               305-329, 377-430, 676-722
 *   rANS coder ................ constriction 0.4 stack AnsCoder +
 *                               Categorical(perfect=False) as bound at
 *                               latec/rans.py:466-535, 629-657
 *   stage record / frame ...... latec/bitstream.py:155-164, 178-344
 *                               (zstandard bound at bitstream.py:156-163,
 *                               210-212, 312)
 */
#ifndef LATEC_H
#define LATEC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LATEC_OK 0
#define LATEC_E_ARG -1
#define LATEC_E_HIP -2
#define LATEC_E_CAPACITY -3
#define LATEC_E_FORMAT -4
#define LATEC_E_COUNT -5
#define LATEC_MAX_DIM 8

/* Last error message of the calling thread ("" if none). */
const char *latec_last_error(void);
/* ABI version (bumped on any signature change). */
int latec_abi_version(void);

/* Launch-shape knobs of the plain-path kernels, read and set by the
 * autotuner (latec_amd/autotune.py).  No reference counterpart: the
 * reference has no GPU kernels to shape.  Knobs of the 3-D level kernels:
 * "march_blocks" (workgroups per parity class, 8..2^20, rounded up to a
 * multiple of 8), "march_hist_peel" (0..8 distinct codes per wave counted
 * with one histogram atomic each), "slab" (1: large levels run as one
 * level-fused launch; 0: three class launches per level), "pair" (1: levels
 * above the first as line-pair launches), "pair_lines" (1: whole lines per
 * workgroup), "pair_quads" (bits: line quads for encode / decode, and for
 * phase 2 alone), "min_threads" (march shortening threshold) and
 * "pair_waves" (1: the finest level's encode phases 2 and 3 compiled for
 * at least 8 waves per SIMD), "sheet" (bits: encode / decode run large
 * cubic 3-D levels as even-sheet and odd-sheet launches), "sheet_blocks"
 * (workgroup target of a sheet launch, 1..2^20) and "sheet_min_points"
 * (smallest level that runs as sheets).  "quad_launches" and
 * "sheet_launches" read (or, set, reset) counts of line-quad and sheet
 * launches.  GNN kernels: "gnn_rows" (0..7: 16-row tiles per wave group,
 * waves per workgroup and minimum waves per SIMD of the message / mix
 * rows), "gnn_pool" (0: line pool line by line; 1, 2: loads issued ahead,
 * 8 or 4 lines per batch) and
 * "gnn_geo_table" (1: stage geometry through per-stage lookup tables of the
 * first known cell along each line; 0: a scan of the cells).
 * Process-wide; no knob changes a code, a reconstructed value or a GNN
 * prediction. */
int latec_tune_set(const char *name, int64_t value);
int64_t latec_tune_get(const char *name);

/* ------------------------------------------------------------------ */
/* Schedule plan: the dyadic stage schedule of one region shape.       */
/* latec/levels.py:100-147 (iter_stage_plan), 221-282 (stage_counts),  */
/* 410-574 (parity-class layouts with affine stage rank).              */
/* ------------------------------------------------------------------ */
typedef struct latec_plan latec_plan;

/* Validation and messages follow levels.py:54-76 (_check_plan). */
int latec_plan_create(int ndim, const int64_t *shape, int channels, int levels,
                      int anchor_stride, int anchor_block, latec_plan **out);
void latec_plan_destroy(latec_plan *p);
int latec_plan_n_stages(const latec_plan *p);
/* Points per channel of `stage`, its lattice stride, weight (0 anchors, 1
 * single axis, >=2 fused) and number of parity classes. */
int latec_plan_stage_info(const latec_plan *p, int stage, int64_t *count,
                          int *stride, int *weight, int *n_classes);
/* One parity class: out[0]=parity bits, [1]=grid step, [2]=rank offset,
 * [3]=size, then starts[ndim], sizes[ndim], rank strides[ndim]. */
int latec_plan_class_info(const latec_plan *p, int stage, int cls,
                          int64_t *out);
/* Offset of `stage` in the concatenated per-channel stage arrays. */
int64_t latec_plan_stage_offset(const latec_plan *p, int stage);

/* ------------------------------------------------------------------ */
/* Fused interpolation + quantization stage kernels                    */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t cubic;      /* 1 cubic, 0 linear (predictor.py:187-217) */
  int32_t center;     /* 0 average odd axes, 1 first, 2 last */
  int32_t end_mode;   /* END_QUAD=1 | END_EXTRAP=2 (predictor.py:46-48) */
  int32_t round_mode; /* 0 none, 1 rint (integer source), 2 (span, offset) */
  double span, offset;
  uint32_t radius;    /* quantizer radius R, codes in [0, 2R) */
  int32_t code_bytes; /* 2 (R <= 32768) or 4 */
} latec_interp_opts;

/* Encode stages [s0, s1) of the plan in place (codec.py:384-459):
 *  d_x, d_recon ... (C, N) float32; d_recon must hold the reconstruction of
 *                   stages < s0 and zeros elsewhere.
 *  d_codes ........ C * N codes (u16/u32); stage i occupies
 *                   [C*off_i, C*off_i + C*n_i) channel-major, stage order.
 *  d_hist ......... n_stages * latec_hist_words(R) u32, zeroed by the caller:
 *                   per stage the 2R code bins then striped copies of the
 *                   central window (folded by latec_fit_const).
 *  d_out_rank/val . outlier (index in the stage array, value) pairs in
 *                   arbitrary order; stage i's list starts at C*off_i.
 *  d_out_count .... n_stages u32 counters, zeroed by the caller.
 *  ebs ............ host array of per-stage absolute error bounds.
 * Stage 0 (anchors) is quantized against a zero prediction. */
int latec_interp_encode(latec_plan *p, int s0, int s1, const float *d_x,
                        float *d_recon, void *d_codes, uint32_t *d_hist,
                        uint32_t *d_out_rank, float *d_out_val,
                        uint32_t *d_out_count, const double *ebs,
                        const latec_interp_opts *o, void *stream);

/* Decode stages [s0, s1) in place (codec.py:462-510).  d_out_rank must be
 * sorted ascending within each stage; d_out_count (device) per stage. */
int latec_interp_decode(latec_plan *p, int s0, int s1, const void *d_codes,
                        const uint32_t *d_out_rank, const float *d_out_val,
                        const uint32_t *d_out_count, float *d_recon,
                        const double *ebs, const latec_interp_opts *o,
                        void *stream);

/* Capture the whole encode (or decode) of the plan as a hipGraph on first
 * use and replay it afterwards; same arguments as the calls above with
 * s0 = 0, s1 = n_stages.  The graph is cached per plan and argument set. */
int latec_interp_encode_graph(latec_plan *p, const float *d_x, float *d_recon,
                              void *d_codes, uint32_t *d_hist,
                              uint32_t *d_out_rank, float *d_out_val,
                              uint32_t *d_out_count, const double *ebs,
                              const latec_interp_opts *o, void *stream);

/* Prediction only (InterpPredictor.predict, predictor.py:255-317): the
 * stage's (C, n) float32 predictions in stage order from d_recon. */
int latec_interp_predict(latec_plan *p, int stage, const float *d_recon,
                         float *d_pred, const latec_interp_opts *o,
                         void *stream);

/* u32 words of one stage histogram (2R bins + striped window copies). */
/* Per-stage error moments of a finished encode (_stage_error_stats,
 * codec.py:353-381): d_out (device, n_stages x 5 float64) gets for each
 * stage pred_sae, pred_sse, recon_sae, recon_sse, recon_max over its C * n
 * points, the prediction recomputed from d_recon (the encoder's and the
 * decoder's prediction, bit for bit), differences in float32, sums in
 * float64 in a fixed order (identical on every run).  d_scratch holds
 * latec_interp_stats_words(p) float64 words.  Synchronizes the stream. */
int64_t latec_interp_stats_words(latec_plan *p);
int latec_interp_stats(latec_plan *p, const float *d_x, const float *d_recon,
                       const latec_interp_opts *o, double *d_out, double *d_scratch,
                       int64_t scratch_words, void *stream);

/* float32 min and max of nseg segments [off[g], off[g] + len[g]) of device
 * d_x (host off / len, len > 0) into d_out[2 g], d_out[2 g + 1] (device), as
 * NumPy's min() / max() (a NaN makes both NaN): the value range compress
 * takes from its input (codec.py:93-97).  d_scratch holds
 * latec_minmax_words(nseg) float32 words.  Synchronizes the stream. */
int64_t latec_minmax_words(int nseg);
int latec_minmax(int nseg, const float *d_x, const int64_t *off, const int64_t *len,
                 float *d_out, float *d_scratch, int64_t scratch_words, void *stream);

int64_t latec_hist_words(uint32_t radius);

/* Add the histogram of n codes (u16/u32) into d_hist (latec_hist_words). */
int latec_code_hist(const void *d_codes, int64_t n, int code_bytes,
                    uint32_t radius, uint32_t *d_hist, void *stream);

/* Generic quantizer on device buffers (quantizer.py:28-87): codes and the
 * outlier (index, value) pairs as above; *d_out_count is one counter. */
int latec_quantize(const float *d_x, const float *d_pred, int64_t n, double eb,
                   const latec_interp_opts *o, void *d_codes,
                   uint32_t *d_out_rank, float *d_out_val,
                   uint32_t *d_out_count, uint32_t *d_hist, void *stream);
/* dequantize (quantizer.py:90-120) with sorted outlier positions. */
int latec_dequantize(const float *d_pred, const void *d_codes, int64_t n,
                     const uint32_t *d_out_rank, const float *d_out_val,
                     const uint32_t *d_out_count, double eb,
                     const latec_interp_opts *o, float *d_recon, void *stream);

/* ------------------------------------------------------------------ */
/* Laplace-mixture tables and the per-stage fit, built on the GPU       */
/* ------------------------------------------------------------------ */
/* _level_weights (rans.py:247-267): d_out (64, 2^(kexp+1)) float64. */
int latec_level_weights(int kexp, double *d_out, void *stream);
/* _shaped_weights + _with_outlier_weight (rans.py:377-384, 172-180). */
int latec_model_weights(int kexp, int shape, int owx, double *d_out,
                        void *stream);
/* _neg_log2_weights (rans.py:387-392): (64, A) float32. */
int latec_neg_log2_weights(int kexp, int shape, int owx, int precision,
                           float *d_out, void *stream);
/* _build_cdfs_cached rows (rans.py:305-329, 676-722): rows [l0, l1) of the
 * (64, A+1) uint32 CDF table into d_out ((l1-l0) * (A+1)). */
int latec_build_cdfs(int kexp, int shape, int precision, int owx, int l0,
                     int l1, uint32_t *d_out, void *stream);

typedef struct {
  int32_t kexp, level, shape, owx;
  float bits;        /* model bits of the fitted level (rans.py:430) */
  uint32_t n_codes;  /* C * n */
  uint32_t n_out;    /* outliers (histogram bin 0) */
  uint32_t pad;
} latec_fit;

/* choose_kexp + choose_const_level (rans.py:70-80, 395-430) for stages
 * [s0, s1) from the code histograms d_hist (n_stages * latec_hist_words);
 * the striped window copies are folded (and cleared) first; d_codes is read
 * only for stages whose largest bin reaches 2^24, to reproduce the
 * reference's per-2^22-tile float32 histogram accumulation.  Result: host
 * array fits[s1 - s0] (the call synchronizes `stream`). */
int latec_fit_const(latec_plan *p, int s0, int s1, const uint32_t *d_hist,
                    const void *d_codes, uint32_t radius, int code_bytes,
                    int precision, latec_fit *fits, void *stream);
/* The same fit for any list of stages whose histograms are consecutive
 * (stage s at d_hist + s * latec_hist_words), codes at code_off[s] (n_codes[s]
 * of them): e.g. every stage of every field of a batch in one call. */
int latec_fit_stages(int nst, const uint32_t *d_hist, const void *d_codes,
                     const int64_t *code_off, const int64_t *n_codes, uint32_t radius,
                     int code_bytes, int precision, latec_fit *fits, void *stream);
/* The same fit for one code array (no plan): n codes. */
int latec_fit_codes(const uint32_t *d_hist, const void *d_codes, int64_t n,
                    uint32_t radius, int code_bytes, int precision,
                    latec_fit *fit, void *stream);

/* scale_to_level (rans.py:270-278) on device: levels u8, by comparison with
 * 63 ascending positive float32 thresholds (host array), the smallest float32
 * at each level 1..63 under the host's NumPy scale_to_level for this eb, found
 * by bisection over float32 bit patterns as latec/gnn_codec.py:827-861 does
 * (latec_amd.rans.scale_thresholds).  Level = thresholds at or below the
 * scale, so the device levels equal the host's NumPy levels exactly. */
int latec_scale_to_level(const float *d_scale, int64_t n, const float *thresholds,
                         uint8_t *d_levels, void *stream);

/* ------------------------------------------------------------------ */
/* Host entropy coding and records                                      */
/* ------------------------------------------------------------------ */
/* Stack rANS with a 24-bit categorical model quantized from the float32
 * pmf of `cdf` (A+1 entries): the design of constriction's AnsCoder<u32,u64>
 * + Categorical(perfect=False).  Symbols are pushed last-to-first so that
 * decoding yields them first-to-last (rans.py:490-498).  Returns words
 * written or LATEC_E_CAPACITY. */
int64_t latec_rans_encode(const uint32_t *syms, int64_t n,
                          const uint32_t *cdf, int64_t A, uint32_t *words,
                          int64_t cap);
/* Per-point models: symbols grouped by level, levels pushed 63..0
 * (rans.py:546-578); cdfs is (64, A+1). */
int64_t latec_rans_encode_levels(const uint32_t *syms, const uint8_t *levels,
                                 int64_t n, const uint32_t *cdfs, int64_t A,
                                 uint32_t *words, int64_t cap);
/* The same from the quantizer's codes (u16/u32, code_bytes), narrowed to
 * the alphabet of 2^kexp on each side of `radius` on the fly (rans.py:83-89;
 * cdfs is (64, A+1) with A = 2 << kexp): rans_encode's narrowing and coding
 * (rans.py:546-578) in one call. */
int64_t latec_rans_encode_codes_levels(const void *codes, int code_bytes,
                                       const uint8_t *levels, int64_t n,
                                       uint32_t radius, int kexp,
                                       const uint32_t *cdfs, int64_t A,
                                       uint32_t *words, int64_t cap);
/* choose_kexp (rans.py:70-80) over u16/u32 codes: the bit length of the
 * largest |code - radius| of the nonzero codes, at least 1. */
int latec_code_kexp(const void *codes, int code_bytes, int64_t n,
                    uint32_t radius);
/* Returns 0, or LATEC_E_FORMAT for trailing / malformed data. */
int latec_rans_decode(const uint32_t *words, int64_t nw, const uint32_t *cdf,
                      int64_t A, uint32_t *syms, int64_t n);
int latec_rans_decode_levels(const uint32_t *words, int64_t nw,
                             const uint8_t *levels, const uint32_t *cdfs,
                             int64_t A, uint32_t *syms, int64_t n);
/* The same, widened back to the quantizer's codes as latec_rans_encode_
 * codes_levels narrowed them (symbol 0 -> code 0), written as u16 or u32
 * (code_bytes); rans_decode + the widening of rans.py:581-616 in one call.
 * LATEC_E_FORMAT also for a u16 code past 0xffff. */
int latec_rans_decode_codes_levels(const uint32_t *words, int64_t nw,
                                   const uint8_t *levels, const uint32_t *cdfs,
                                   int64_t A, uint32_t radius, int kexp,
                                   void *codes, int code_bytes, int64_t n);

/* pack_stage, const-level path (bitstream.py:178-243): codes are u16/u32
 * (code_bytes), `fit` from latec_fit_*, `cdf_row` the (A+1) CDF of
 * (fit->kexp, fit->shape, fit->owx, fit->level).  Returns bytes written. */
int64_t latec_pack_stage(const void *codes, int code_bytes, int64_t n,
                         const float *outl, int64_t n_out, uint32_t radius,
                         int zstd_level, const latec_fit *fit,
                         const uint32_t *cdf_row, uint8_t *out, int64_t cap);
/* Parse a stage record header (bitstream.py:288-300): n_codes and framing
 * fields; returns 12 or LATEC_E_FORMAT. */
int latec_stage_header(const uint8_t *buf, int64_t len, uint32_t *n_codes,
                       int64_t *hlen, int *kexp, int *shape, int *level,
                       int *owx, int *literal);
/* unpack_stage, const-level path (bitstream.py:288-344).  cdf_row is the row
 * named by the header (ignored for literal records).  Writes codes (u16/u32)
 * and the sorted positions + values of the outliers (code 0).  Returns bytes
 * consumed. */
int64_t latec_unpack_stage(const uint8_t *buf, int64_t len, uint32_t radius,
                           const uint32_t *cdf_row, int64_t n_expect,
                           void *codes, int code_bytes, uint32_t *out_rank,
                           float *out_val, int64_t out_cap, int64_t *n_out);

/* Multi-threaded pack of many stages into one payload buffer: stage i's
 * codes at codes + code_bytes * code_off[i], outliers at outl + out_off[i]
 * (already in stage order), cdf rows at cdf_rows + row_off[i].  Writes the
 * concatenated records; rec_len[i] receives each record's size. */
int64_t latec_pack_stages(int n_stages, const void *codes, int code_bytes,
                          const int64_t *code_off, const int64_t *n_codes,
                          const float *outl, const int64_t *out_off,
                          const int64_t *n_out, uint32_t radius,
                          int zstd_level, const latec_fit *fits,
                          const uint32_t *cdf_rows, const int64_t *row_off,
                          uint8_t *out, int64_t cap, int64_t *rec_len,
                          int threads);

/* latec_pack_stages plus the outer zstd frame of the payload at frame_level
 * (1..22), i.e. the body write_stream adds after the header
 * (latec/bitstream.py:155-157: ZstdCompressor(level).compress(payload)).
 * frame_cap >= latec_zstd_bound(cap); *frame_len receives the frame size.
 * The frame is started once every stage has its final record or its rANS
 * record, and redone if a large stage's literal frame wins: the outputs
 * equal latec_pack_stages followed by latec_zstd_compress. */
int64_t latec_pack_stages_framed(int n_stages, const void *codes, int code_bytes,
                                 const int64_t *code_off, const int64_t *n_codes,
                                 const float *outl, const int64_t *out_off,
                                 const int64_t *n_out, uint32_t radius,
                                 int zstd_level, const latec_fit *fits,
                                 const uint32_t *cdf_rows, const int64_t *row_off,
                                 uint8_t *out, int64_t cap, int64_t *rec_len,
                                 int threads, int frame_level, uint8_t *frame,
                                 int64_t frame_cap, int64_t *frame_len);

/* zstd one-shot frames (the reference's zstandard.ZstdCompressor(level)
 * .compress / ZstdDecompressor().decompress). */
int64_t latec_zstd_bound(int64_t n);
int64_t latec_zstd_compress(const uint8_t *src, int64_t n, int level,
                            uint8_t *dst, int64_t cap);
int64_t latec_zstd_content_size(const uint8_t *src, int64_t n);
int64_t latec_zstd_decompress(const uint8_t *src, int64_t n, uint8_t *dst,
                              int64_t cap);

/* ------------------------------------------------------------------ */
/* GNN predictor (latec/gnn_predictor.py, latec/gnn_codec.py).         */
/* ------------------------------------------------------------------ */

typedef struct latec_gnn latec_gnn;

#define LATEC_GNN_FP32 0 /* FP32 MFMA GEMMs, FP32 pools and head */
#define LATEC_GNN_BF16 1 /* BF16 MFMA GEMMs (FP32 accumulate); pools and head \
                            in FP32, their GEMMs on split-bf16 MFMA */

/* Floats in the packed parameter block of a (d, agg_level) model: the
 * reference checkpoint's state_dict tensors flattened in module order
 * (build_model, gnn_predictor.py:593-933; init, dir_state, dir, bidir,
 * line_pool, axis_pool, head, mix; weight before bias; AttnPool's wk, wv,
 * null_k, null_v). */
int64_t latec_gnn_param_count(int d, int agg_level);

/* Load a model for rank-`ndim` fields (_load_inference_model,
 * gnn_predictor.py:945-981, after the checkpoint is read on the host).
 * d must be 32 or 64.  The handle owns its device weights and a workspace;
 * calls on one handle must not overlap. */
int latec_gnn_create(int d, int agg_level, int ndim, int numerics,
                     const float *params, int64_t n_params, latec_gnn **out);
void latec_gnn_destroy(latec_gnn *g);
/* Half-directions (lines) of the model at its rank (half_directions,
 * gnn_predictor.py:117-136). */
int latec_gnn_n_lines(const latec_gnn *g);

/* Where a chunk's embedding frame sits (ChunkedGNNPredictor frames,
 * gnn_predictor.py:1010-1356): the frame is the chunk grown by `pad` cells on
 * every side; halo cells are usable when inside the tensor and in a coded
 * chunk (d_coded, raster chunk order) or anchors. */
typedef struct latec_gnn_frame {
  int32_t pad;       /* 0 (one chunk, no halo) or the anchor stride */
  int32_t has_halo;
  int64_t origin[LATEC_MAX_DIM];  /* chunk origin in the tensor */
  int64_t tshape[LATEC_MAX_DIM];  /* tensor shape */
  int64_t edges[LATEC_MAX_DIM];   /* chunk edges */
  int64_t grid[LATEC_MAX_DIM];    /* chunks per axis */
  const uint8_t *d_coded;         /* per chunk: already coded (halo only) */
} latec_gnn_frame;

/* Geometry of one stage of a chunk's local schedule (`p`: block 1, stride
 * 2^levels, shape = chunk shape).  Replaces _StageGeom / _ChunkGeoms /
 * _CompactGeom (gnn_predictor.py:270-443, 1010-1310): per stage point in
 * stage order, d_qidx = frame flat index, d_gidx = tensor flat index and
 * d_steps[m * L + l] = sp | sn << 16, the step to the nearest usable known
 * sample along +/- line l within `stride` steps (0 = none). */
int latec_gnn_geometry(latec_gnn *g, latec_plan *p, int stage,
                       const latec_gnn_frame *fr, int64_t *d_qidx,
                       int64_t *d_gidx, int32_t *d_steps, void *stream);

/* Message pass + line pool of M query points (GNN.embed,
 * gnn_predictor.py:818-906): d_E is (frame rows, ndim, d) in the model's
 * storage type (bf16 or f32), frame_shape the host frame dims; writes the
 * per-axis contexts d_ctx (M, ndim, d) in the storage type. */
int latec_gnn_embed(latec_gnn *g, const void *d_E, const int64_t *frame_shape,
                    const int32_t *d_steps, const int64_t *d_qidx, int64_t M,
                    void *d_ctx, void *stream);

/* Axis pool + PredHead (head_of, gnn_predictor.py:749-789, 922-931) and the
 * predictor's output map (predict_wave_stage, gnn_predictor.py:1737-1774):
 * pred = clamp(mu * (vmax - vmin) + vmin), scale = 2^log_b * (vmax - vmin);
 * eb is the normalized bound.  d_mu / d_logb may be NULL. */
int latec_gnn_head(latec_gnn *g, const void *d_ctx, int64_t M, double eb,
                   float vmin, float vmax, float *d_pred, float *d_scale,
                   float *d_mu, float *d_logb, void *stream);

/* finalize (gnn_predictor.py:908-920, stage_forward :984-1000):
 * E[qidx[m], k] = mix(ctx[m, k], init(clamp(recon[gidx[m]]))) . */
int latec_gnn_finalize(latec_gnn *g, void *d_E, const void *d_ctx,
                       const int64_t *d_qidx, const int64_t *d_gidx,
                       const float *d_recon, int64_t M, float vmin, float vmax,
                       void *stream);

/* Halo rows: E[rows[m], k] = init(clamp(recon[gidx[m]])) for every axis k
 * (value_halo_embed, gnn_predictor.py:1357-1366). */
int latec_gnn_halo(latec_gnn *g, void *d_E, const int64_t *d_rows,
                   const int64_t *d_gidx, const float *d_recon, int64_t M,
                   float vmin, float vmax, void *stream);

/* All halo rows of a chunk frame at once (_wave_fill_halo,
 * gnn_predictor.py:1716-1735, with the usable cells of chunk_halo_info
 * :1338-1356): every frame cell outside the chunk (`p`: the chunk's local
 * plan) that lies inside the tensor and in a coded chunk or on the anchor
 * lattice gets E[row, k] = init(clamp(recon[cell])) for every axis k.  No-op
 * for a frame without a halo. */
int latec_gnn_halo_frame(latec_gnn *g, latec_plan *p, const latec_gnn_frame *fr,
                         void *d_E, const float *d_recon, float vmin, float vmax,
                         void *stream);

/* Gate rate proxy (_laplace_bits_t + the bucket sums of _gate_select_t,
 * gnn_codec.py:777-797, 920-971): costs[c * 13 + b] = summed bits of
 * candidate c in scale bucket b (c = 0: the model; c = 1 + 4 * kind + s:
 * interpolation candidate `kind` with scale shift 2s).  thr63 / grid64 are
 * the gate's scale snapping tables (gnn_codec.py:880-917).  Sums are in a
 * fixed order. */
int latec_gnn_gate_costs(latec_gnn *g, const float *d_x, const float *d_pred,
                         const float *d_scale, const float *const *d_ip,
                         int n_interp, int64_t n, double eb, uint32_t radius,
                         const float *thr63, const double *grid64,
                         double *costs, void *stream);

/* One coded stage of the GNN encoder after the model's prediction (the body
 * of the chunk loop, gnn_codec.py:1215-1300): the coded rows' prediction,
 * scale and values; with `gate`, the three interpolation candidates
 * (latec_interp_predict on d_block with cand_opts), the gate's bucket sums,
 * its choice (_gate_select, gnn_codec.py:920-971, on the host in float64 in
 * NumPy's order) and its application; the quantizer, the reconstruction
 * written into d_recon at the points' tensor indices, the scale levels, and
 * the codes, levels and outliers (by stage position) copied to the host.
 * interp_kind 1..3 codes the stage with that interpolation candidate at scale
 * eb instead of the model.  Two host waits: the gate's sums, the outputs. */
typedef struct latec_gnn_stage_io {
  int64_t M;                 /* points of the full local stage */
  int64_t n;                 /* coded points (M, or the chunk's column split) */
  const int64_t *d_keep;     /* n positions in the full stage; NULL if n == M */
  const float *d_pred;       /* (M) model prediction */
  const float *d_scale;      /* (M) model scale */
  const int64_t *d_gidx;     /* (M) tensor index of each stage point */
  const float *d_x;          /* normalized tensor, flat */
  float *d_recon;            /* reconstruction, flat; the coded points written */
  const float *d_block;      /* the chunk's block of the reconstruction */
  const latec_interp_opts *opts;       /* quantizer options */
  const latec_interp_opts *cand_opts;  /* 3: interpolation kinds 1, 2, 3 */
  int32_t gate, interp_kind, reuse_choice, pad0;
  int32_t *choice;           /* host int32[4] {kind, dir, T, shift}: written,
                              * or read when reuse_choice */
  uint32_t *gate_word;       /* host: the stage's gate word (0 none) or NULL */
  double eb;
  uint32_t radius, pad1;
  const float *gate_thr;     /* host: the gate's 63 snapping thresholds of eb */
  const double *gate_grid;   /* host: the gate's 64 grid scales of eb */
  const float *level_thr;    /* host: the 63 scale-level thresholds of eb */
  uint32_t *d_hist;          /* device histogram scratch of latec_quantize */
  void *codes;               /* host out: n codes (opts->code_bytes each) */
  uint8_t *levels;           /* host out: n scale levels */
  float *outliers;           /* host out: outlier values by position (cap n) */
  int64_t *n_out;            /* host out: outlier count */
} latec_gnn_stage_io;
int latec_gnn_code_stage(latec_gnn *g, latec_plan *p, int stage,
                         const latec_gnn_stage_io *io, void *stream);

/* The decoder's stage after the model, in two calls around the host's rANS
 * decode (gnn_codec.py:1430-1490).  begin: the coded rows' prediction and
 * scale, the stream's gate word applied (or interpolation kind interp_kind
 * at scale eb), and their scale levels copied to `levels` (host, n);
 * end: the decoded codes and outlier values (host) dequantized against that
 * prediction and written into d_recon at the rows' tensor indices
 * ("outlier count mismatch" when the zero codes and the values disagree).
 * The prediction and indices stay in the handle between the two calls. */
typedef struct latec_gnn_dstage_io {
  int64_t M;                 /* points of the full local stage */
  int64_t n;                 /* coded points */
  const int64_t *d_keep;     /* n positions in the full stage; NULL if n == M */
  const float *d_pred;       /* (M) model prediction */
  const float *d_scale;      /* (M) model scale */
  const int64_t *d_gidx;     /* (M) tensor index of each stage point */
  const float *d_block;      /* the chunk's block of the reconstruction */
  const latec_interp_opts *cand_opts;  /* 3: interpolation kinds 1, 2, 3 */
  int32_t interp_kind;       /* 0: the model (+ gate_word), 1..3: that kind */
  uint32_t gate_word;        /* the stage's gate word, 0 for none */
  double eb;
  const float *level_thr;    /* host: the 63 scale-level thresholds of eb */
  uint8_t *levels;           /* host out: n scale levels */
} latec_gnn_dstage_io;
int latec_gnn_decode_stage_begin(latec_gnn *g, latec_plan *p, int stage,
                                 const latec_gnn_dstage_io *io, void *stream);
int latec_gnn_decode_stage_end(latec_gnn *g, const void *codes,
                               const float *outliers, int64_t n_out, double eb,
                               const latec_interp_opts *o, float *d_recon,
                               void *stream);

/* _gate_apply_t (gnn_codec.py:974-995): where the scale is below (dir 0) or
 * at/above (dir 1) `threshold`, pred = ip and scale *= mult. */
int latec_gnn_gate_apply(float *d_pred, float *d_scale, const float *d_ip,
                         int64_t n, double threshold, float mult, int dir,
                         void *stream);

/* dst[i] = src[idx[i]] / dst[idx[i]] = src[i] on device (the stage views of
 * _stage_values / _write_stage, gnn_codec.py:685-701). */
int latec_gather_f32(const float *d_src, const int64_t *d_idx, int64_t n,
                     float *d_dst, void *stream);
int latec_scatter_f32(const float *d_src, const int64_t *d_idx, int64_t n,
                      float *d_dst, void *stream);

/* ------------------------------------------------------------------ */
/* GNN trainer ops (scripts/train_gnn.py:195-230, 373-459): FP32       */
/* forward/backward kernels composed by torch.autograd.Functions.      */
/* ------------------------------------------------------------------ */

/* out[r, :] = src[idx[r], :] (rows of width w) / E[idx[r], :] = vals[r, :]
 * (vals NULL: zeros). */
int latec_train_gather(const float *d_src, int64_t w, const int64_t *d_idx,
                       int64_t n, float *d_out, void *stream);
/* dsrc (n_src rows of width w) = sum over r of dout[r, :] into row idx[r],
 * deterministic: 64-bit fixed-point sums scaled by max |dout| (exact in any
 * order), one rounding to float32.  d_acc: n_src * w + 1 words of 64 bits
 * (scratch, cleared here). */
int latec_train_scatter_add(const float *d_dout, int64_t w,
                            const int64_t *d_idx, int64_t n, float *d_dsrc,
                            int64_t n_src, uint64_t *d_acc, void *stream);
int latec_train_write_rows(float *d_E, int64_t w, const int64_t *d_idx,
                           int64_t n, const float *d_vals, void *stream);
/* C (M x N, row stride ldc) [+]= A * B (+ bias[N]) with generic strides
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; split-K partials
 * in d_work (work_floats floats), summed in a fixed order. */
int latec_train_gemm(int64_t M, int64_t N, int64_t K, const float *d_A,
                     int64_t sam, int64_t sak, const float *d_B, int64_t sbk,
                     int64_t sbn, float *d_C, int64_t ldc, const float *d_bias,
                     int accumulate, float *d_work, int64_t work_floats,
                     void *stream);
/* The same product with both operands rounded to bf16 and FP32 accumulation
 * on v_mfma_f32_16x16x32_bf16 tiles (BF16 training; FP32 master weights). */
int latec_train_gemm_bf16(int64_t M, int64_t N, int64_t K, const float *d_A,
                          int64_t sam, int64_t sak, const float *d_B, int64_t sbk,
                          int64_t sbn, float *d_C, int64_t ldc, const float *d_bias,
                          int accumulate, float *d_work, int64_t work_floats,
                          void *stream);
/* out[n] [+]= sum_m X[m, n] in a fixed order. */
int latec_train_colsum(int64_t M, int64_t N, const float *d_X, float *d_out,
                       int accumulate, float *d_work, int64_t work_floats,
                       void *stream);
int latec_train_gelu(const float *d_x, float *d_y, int64_t n, void *stream);
int latec_train_gelu_bwd(const float *d_x, const float *d_dy, float *d_dx,
                         int64_t n, void *stream);
/* AttnPool over CSR row segments + a null token (gnn_predictor.py:720-747):
 * forward saves the softmax weights (w_rows, w_null) for the backward. */
int latec_train_attn(int64_t n_items, const int64_t *d_off, const float *d_s,
                     const float *d_v, int d, float null_s,
                     const float *d_null_v, float *d_out, float *d_w_rows,
                     float *d_w_null, void *stream);
/* Backward: d_dnull (d + 1 floats) gets the null score's gradient then the
 * null value's, summed over items in a fixed order through d_work (at least
 * n_items * (d + 1) + 1024 * (d + 1) floats). */
int latec_train_attn_bwd(int64_t n_items, const int64_t *d_off,
                         const float *d_v, int d, const float *d_null_v,
                         const float *d_w_rows, const float *d_w_null,
                         const float *d_dout, float *d_ds, float *d_dv,
                         float *d_dnull, float *d_work, int64_t work_floats,
                         void *stream);
/* Head loss: mu = sigmoid(o0), log_b = log2(eb) + clamp(o1, -4, 12),
 * bits = discretized-Laplace code length (+ mse_w (mu - t)^2); writes bits,
 * d bits / d o (M x 2) and mu (may be NULL). */
int latec_train_nll(int64_t M, const float *d_o, const float *d_t,
                    const float *d_eb, float mse_w, float *d_bits, float *d_dO,
                    float *d_mu, void *stream);

/* ------------------------------------------------------------------ */
/* Multi-lane rANS stage blobs (LATEC001 version 5).  No reference     */
/* counterpart: version 5 is added beside version 4 (whose serial      */
/* stack-rANS blobs latec/rans.py:466-535 writes) so that entropy      */
/* coding runs on the GPU.  Records keep version 4's framing           */
/* (latec/bitstream.py:178-285) and stage-constant model fit; a blob   */
/* is [L][L + 1 lane word offsets][lane words], symbol i in lane i % L,*/
/* each lane a range-ANS coder with 64-bit state in [2^31, 2^63),      */
/* 32-bit words and the stage's 2^24 CDF row as the model.             */
/* ------------------------------------------------------------------ */
/* Lanes of an n-symbol stage: clamp(ceil(n / 4096), 1, 65536). */
int64_t latec_mlrans_lanes(int64_t n);
/* Device scratch (u32 words) latec_mlrans_encode needs for these stages. */
int64_t latec_mlrans_scratch_words(int ns, const int64_t *n);
/* Encode ns stages of device codes (u16/u32) with per-stage kexp and CDF
 * rows (device, row_off into d_rows, host arrays for the rest); writes the
 * blobs back to back into d_out (4-byte aligned) and their host-side
 * offsets / lengths in bytes.  Synchronizes the stream. */
int latec_mlrans_encode(int ns, const void *d_codes, int code_bytes,
                        const int64_t *code_off, const int64_t *n,
                        const int32_t *kexp, uint32_t radius,
                        const uint32_t *d_rows, const int64_t *row_off,
                        uint32_t *d_scratch, int64_t scratch_words,
                        uint8_t *d_out, int64_t out_cap, int64_t *blob_off,
                        int64_t *blob_len, void *stream);
/* The same encode with no workspace argument and no device allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its workspace is a fixed device array of
 * latec_mlrans_noalloc_words() = 2^23 u32 words (32 MiB) per GPU, held
 * until the process ends; stages whose latec_mlrans_scratch_words exceeds
 * it are refused with LATEC_E_CAPACITY before anything runs.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialized by the caller.  Output and
 * errors as latec_mlrans_encode; synchronizes the stream. */
int64_t latec_mlrans_noalloc_words(void);
int latec_mlrans_encode_noalloc(int ns, const void *d_codes, int code_bytes,
                                const int64_t *code_off, const int64_t *n,
                                const int32_t *kexp, uint32_t radius,
                                const uint32_t *d_rows, const int64_t *row_off,
                                uint8_t *d_out, int64_t out_cap, int64_t *blob_off,
                                int64_t *blob_len, void *stream);
/* Decode the blobs at blob_off / blob_len (bytes, in the device payload)
 * into d_codes at code_off; LATEC_E_FORMAT for a malformed blob.
 * Synchronizes the stream. */
int latec_mlrans_decode(int ns, const uint8_t *d_payload, int64_t payload_len,
                        const int64_t *blob_off, const int64_t *blob_len,
                        const int64_t *n, const int32_t *kexp, uint32_t radius,
                        const uint32_t *d_rows, const int64_t *row_off,
                        void *d_codes, int code_bytes, const int64_t *code_off,
                        uint32_t *d_scratch, int64_t scratch_words, void *stream);
/* Ordered ranks of the zero (outlier) codes of each stage (the positions
 * latec/bitstream.py:288-344 recovers on the host) into d_out_rank at
 * code_off, counts into d_out_count[ns]. */
int latec_zero_ranks(int ns, const void *d_codes, int code_bytes,
                     const int64_t *code_off, const int64_t *n,
                     uint32_t *d_out_rank, uint32_t *d_out_count,
                     uint32_t *d_scratch, int64_t scratch_words, void *stream);

/* latec_zero_ranks plus each outlier's value: d_out_val[code_off + k] =
 * d_x[code_off + d_out_rank[code_off + k]] (d_x in stage order like the
 * codes), so the outliers come out sorted by stage rank, the order
 * latec/quantizer.py:28-87 lists them in. */
int latec_order_outliers(int ns, const void *d_codes, int code_bytes,
                         const int64_t *code_off, const int64_t *n, const float *d_x,
                         uint32_t *d_out_rank, float *d_out_val, uint32_t *d_out_count,
                         uint32_t *d_scratch, int64_t scratch_words, void *stream);
/* Scratch (u32 words) latec_zero_ranks / latec_order_outliers need. */
int64_t latec_zero_ranks_words(int ns, const int64_t *n);

/* ------------------------------------------------------------------ */
/* Batched uneven workload: many independent fields (anchor_block 1,   */
/* one channel, radius <= 32768) coded together.  No reference         */
/* counterpart as one call: the reference codes tensors one at a time  */
/* (latec/codec.py:45-322 per tensor, scripts/eval_tensor.py:145-147   */
/* casting float64 to float32).  Step s runs stage s of every field    */
/* that has one as ONE persistent launch over (field-stage, class,     */
/* rows) items; the steps are recorded once into a hipGraph.  Each     */
/* field's codes/outliers/recon equal its single-field encode.         */
/* ------------------------------------------------------------------ */
typedef struct latec_batch latec_batch;
/* plans[f], d_x[f] / d_recon[f] (device, N_f floats), code_base[f] (first
 * code of field f in d_codes / d_out_rank / d_out_val), hist_base[f]
 * (words, n_stages_f * latec_hist_words each), cnt_base[f] (first of its
 * n_stages_f outlier counters), ebs[f] (host, per stage).  encode = 0
 * builds the decoder (d_x unused; codes, sorted outlier ranks, values and
 * counts are inputs).  o[f]: field f's options (one radius for all). */
int latec_batch_create(int nf, latec_plan *const *plans, const float *const *d_x,
                       float *const *d_recon, const int64_t *code_base,
                       const int64_t *hist_base, const int64_t *cnt_base,
                       const double *const *ebs, const latec_interp_opts *o,
                       int encode, void *d_codes, uint32_t *d_hist,
                       uint32_t *d_out_rank, float *d_out_val,
                       uint32_t *d_out_count, latec_batch **out);
int latec_batch_n_steps(const latec_batch *b);
/* Run every step (use_graph: record once, then replay the hipGraph).
 * Histograms and outlier counters must be zero before an encode run. */
int latec_batch_run(latec_batch *b, int use_graph, void *stream);
void latec_batch_destroy(latec_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* LATEC_H */
