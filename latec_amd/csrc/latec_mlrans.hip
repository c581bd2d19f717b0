// latec_mlrans.hip -- multi-lane rANS stage blobs (LATEC001 version 5) on the
// GPU (gfx950).
//
// Version 5 keeps version 4's stage record and its stage-constant
// Laplace-mixture model (kexp window, shape, scale level and outlier weight
// from the same fit, CDF rows quantized to 2^24), and replaces the single
// serial stack-rANS word stream by L independent lanes that code in
// parallel.  Symbol i of a stage (stage order, narrowed to the kexp window)
// belongs to lane i % L.
//
// Coder (per lane): 64-bit state x in [2^31, 2^63), 32-bit renormalization
// words, probability precision 2^24 (the range-variant ANS of Duda,
// "Asymmetric numeral systems", 2013, with the 64-bit state / 32-bit word
// parameters of ryg_rans' rans64).  Encode walks the lane's symbols
// backwards:
//     if x >= f << 39: emit low 32 bits of x, x >>= 32
//     x = ((x / f) << 24) + x % f + c          (f = freq(s), c = cdf(s))
// starting from x = 2^31.  Decode walks forwards:
//     slot = x & (2^24 - 1);  s = the symbol with c_s <= slot < c_{s+1}
//     x = f_s * (x >> 24) + slot - c_s;  if x < 2^31: x = x << 32 | next word
//
// Blob of a stage with n symbols (all u32, little endian):
//     [L] [off_0 = 0, off_1, ..., off_L] [lane words]
// lane l owns words [off_l, off_{l+1}): its final encoder state (low word,
// high word) followed by its renormalization words in the order the decoder
// reads them.  L = lanes_for(n) = clamp(ceil(n / 4096), 1, 65536); an empty
// stage has an empty blob.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "latec_common.h"

namespace latec {
namespace mlr {

static int hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    set_hip_error(what, (int)e);
    return LATEC_E_HIP;
  }
  return LATEC_OK;
}

constexpr int PREC = 24;
constexpr uint32_t SLOT_MASK = (1u << PREC) - 1;
constexpr uint64_t X_LO = 1ull << 31;
constexpr int SYMS_PER_LANE = 4096;
constexpr int MAX_LANES = 65536;
constexpr int BUCKET_BITS = 12;  // slot -> symbol search start, 2^12 buckets
constexpr int NBUCKET = 1 << BUCKET_BITS;
constexpr int LDS_CDF = 8193;    // CDF rows up to kexp 12 are staged in LDS

inline int64_t lanes_for(int64_t n) {
  if (n <= 0) return 0;
  return std::min<int64_t>(std::max<int64_t>((n + SYMS_PER_LANE - 1) / SYMS_PER_LANE, 1),
                           MAX_LANES);
}

// Per-stage descriptor shared by the kernels.
struct Stage {
  int64_t code_off;   // first code of the stage in the codes array
  int64_t n;          // symbols
  int64_t row_off;    // CDF row (A + 1 entries) in the rows array
  int64_t scratch;    // encode: first scratch word of lane 0; decode: blob offset (bytes)
  int64_t word_base;  // encode: first word of the stage's lane words in the output
  int32_t lanes, kexp;
  int32_t lane_base;  // index of lane 0 among all lanes of the launch
  int32_t pad;
};

__device__ __forceinline__ uint32_t narrow(uint32_t code, uint32_t R, int kexp) {
  return code == 0 ? 0u : code - R + (1u << kexp);
}
__device__ __forceinline__ uint32_t widen(uint32_t sym, uint32_t R, int kexp) {
  return sym == 0 ? 0u : sym - (1u << kexp) + R;
}

// Lane l of a stage holds symbols l, l + L, ...: ceil((n - l) / L) of them.
__device__ __forceinline__ int64_t lane_count(int64_t n, int64_t L, int64_t l) {
  return n > l ? (n - l + L - 1) / L : 0;
}

// ---------------------------------------------------------------------------
// encode: one thread per lane; words are written to the lane's scratch
// region (capacity = its symbol count) in emission order.
// ---------------------------------------------------------------------------
template <typename CodeT>
__global__ void __launch_bounds__(256) enc_kernel(const Stage *__restrict__ st,
                                                  const int32_t *__restrict__ lane_stage,
                                                  int32_t total_lanes,
                                                  const CodeT *__restrict__ codes,
                                                  uint32_t R, const uint32_t *__restrict__ rows,
                                                  uint32_t *__restrict__ scratch,
                                                  uint32_t *__restrict__ emitted,
                                                  uint64_t *__restrict__ state,
                                                  int32_t *__restrict__ status) {
  const int32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
  if (gl >= total_lanes) return;
  const Stage S = st[lane_stage[gl]];
  const int64_t l = gl - S.lane_base, L = S.lanes;
  const int64_t T = lane_count(S.n, L, l);
  const CodeT *cs = codes + S.code_off;
  const uint32_t *cdf = rows + S.row_off;
  const uint32_t A = 2u << S.kexp;
  // lane l's scratch region follows the regions of lanes 0..l-1, each sized
  // by its symbol count: sum_{j<l} lane_count(n, L, j)
  const int64_t q = S.n / L, r = S.n % L;
  uint32_t *out = scratch + S.scratch + l * q + (l < r ? l : r);
  uint64_t x = X_LO;
  uint32_t k = 0;
  for (int64_t t = T - 1; t >= 0; t--) {
    uint32_t s = narrow((uint32_t)cs[l + t * L], R, S.kexp);
    if (s >= A) {  // outside the stage's window: the fit and the codes disagree
      atomicOr(status, 1);
      s = 0;
    }
    const uint32_t c = cdf[s], f = cdf[s + 1] - c;
    if (x >= ((uint64_t)f << 39)) {
      out[k++] = (uint32_t)x;
      x >>= 32;
    }
    x = ((x / f) << PREC) + (x % f) + c;
  }
  emitted[gl] = k;
  state[gl] = x;
}

// Exclusive scan of the per-lane word counts (2 + emitted) within each stage,
// then stage word bases; one block, sequential over stages.
__global__ void __launch_bounds__(1024) scan_kernel(Stage *__restrict__ st, int ns,
                                                    const uint32_t *__restrict__ emitted,
                                                    uint32_t *__restrict__ lane_off,
                                                    int64_t *__restrict__ stage_words) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t carry;
  int64_t base = 0;  // words before this stage (its blob header excluded)
  for (int s = 0; s < ns; s++) {
    const int L = st[s].lanes;
    const int lb = st[s].lane_base;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < L; c0 += 1024) {
      const int l = c0 + threadIdx.x;
      const uint32_t v = l < L ? 2u + emitted[lb + l] : 0u;
      part[threadIdx.x] = v;
      __syncthreads();
      for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele inclusive scan
        const uint32_t a = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += a;
        __syncthreads();
      }
      if (l < L) lane_off[lb + s + l + 1] = carry + part[threadIdx.x];
      __syncthreads();
      if (threadIdx.x == 1023) carry += part[1023];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      lane_off[lb + s] = 0;
      st[s].word_base = base;
      stage_words[s] = carry;
      base += carry;
    }
    __syncthreads();
  }
}

// Blob assembly: lane words at out + 4 * (header words of the stages before
// + word_base); each stage's blob is [L][L + 1 offsets][words].
__global__ void __launch_bounds__(256) assemble_kernel(const Stage *__restrict__ st,
                                                       const int32_t *__restrict__ lane_stage,
                                                       int32_t total_lanes,
                                                       const uint32_t *__restrict__ scratch,
                                                       const uint32_t *__restrict__ emitted,
                                                       const uint64_t *__restrict__ state,
                                                       const uint32_t *__restrict__ lane_off,
                                                       const int64_t *__restrict__ blob_word0,
                                                       uint32_t *__restrict__ out) {
  const int32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
  if (gl >= total_lanes) return;
  const int si = lane_stage[gl];
  const Stage S = st[si];
  const int64_t l = gl - S.lane_base, L = S.lanes;
  uint32_t *blob = out + blob_word0[si];
  const uint32_t *offs = lane_off + S.lane_base + si;
  if (l == 0) blob[0] = (uint32_t)L;
  blob[1 + l] = offs[l];
  if (l == L - 1) blob[1 + L] = offs[L];
  uint32_t *w = blob + 2 + L + offs[l];
  const uint64_t x = state[gl];
  w[0] = (uint32_t)x;
  w[1] = (uint32_t)(x >> 32);
  const uint32_t k = emitted[gl];
  const int64_t q = S.n / L, r = S.n % L;
  const uint32_t *src = scratch + S.scratch + l * q + (l < r ? l : r);
  for (uint32_t j = 0; j < k; j++) w[2 + j] = src[k - 1 - j];
}

// ---------------------------------------------------------------------------
// decode: one thread per lane; blocks never mix stages, so a stage's CDF
// (kexp <= 12) and its slot buckets are staged in LDS.
// ---------------------------------------------------------------------------
struct DecBlock {
  int32_t stage, lane0;
};

template <typename CodeT>
__global__ void __launch_bounds__(256) dec_kernel(const Stage *__restrict__ st,
                                                  const DecBlock *__restrict__ blocks,
                                                  const uint8_t *__restrict__ payload,
                                                  uint32_t R, const uint32_t *__restrict__ rows,
                                                  CodeT *__restrict__ codes,
                                                  int32_t *__restrict__ status) {
  __shared__ uint32_t s_cdf[LDS_CDF];
  __shared__ uint16_t s_bucket[NBUCKET + 1];
  const DecBlock B = blocks[blockIdx.x];
  const Stage S = st[B.stage];
  const int A = 2 << S.kexp;
  const uint32_t *gcdf = rows + S.row_off;
  const bool lds = A + 1 <= LDS_CDF;
  if (lds)
    for (int i = threadIdx.x; i <= A; i += blockDim.x) s_cdf[i] = gcdf[i];
  __syncthreads();
  const uint32_t *cdf = lds ? s_cdf : gcdf;
  // bucket b: the symbol holding slot b << (24 - 12)
  for (int b = threadIdx.x; b <= NBUCKET; b += blockDim.x) {
    const uint32_t slot = b < NBUCKET ? (uint32_t)b << (PREC - BUCKET_BITS) : SLOT_MASK;
    int lo = 0, hi = A - 1;  // largest s with cdf[s] <= slot
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cdf[mid] <= slot) lo = mid; else hi = mid - 1;
    }
    s_bucket[b] = (uint16_t)(lo > 65535 ? 65535 : lo);
  }
  __syncthreads();
  const int64_t L = S.lanes;
  const int64_t l = B.lane0 + threadIdx.x;
  if (l >= L) return;
  const uint32_t *blob = (const uint32_t *)(payload + S.scratch);
  if (l == 0 && blob[0] != (uint32_t)L) atomicOr(status, 8);
  const uint32_t w0 = blob[1 + l], w1 = blob[2 + l];
  // S.word_base holds the blob's length in words: the lane's words must lie in
  // it (compared in 64 bits, so an offset near 2^32 cannot wrap the test)
  if ((uint64_t)w1 < (uint64_t)w0 + 2 || 2 + L + (int64_t)w1 > S.word_base) {
    atomicOr(status, 1);
    return;
  }
  const uint32_t *w = blob + 2 + L + w0;
  const uint32_t nw = w1 - w0;
  const int64_t T = lane_count(S.n, L, l);
  uint64_t x = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  uint32_t k = 2;
  CodeT *dst = codes + S.code_off;
  for (int64_t t = 0; t < T; t++) {
    const uint32_t slot = (uint32_t)x & SLOT_MASK;
    uint32_t s = s_bucket[slot >> (PREC - BUCKET_BITS)];
    while (s + 1 < (uint32_t)A && cdf[s + 1] <= slot) s++;
    const uint32_t c = cdf[s], f = cdf[s + 1] - c;
    x = (uint64_t)f * (x >> PREC) + (slot - c);
    if (x < X_LO) {
      if (k >= nw) {
        atomicOr(status, 2);
        return;
      }
      x = (x << 32) | w[k++];
    }
    dst[l + t * L] = (CodeT)widen(s, R, S.kexp);
  }
  if (k != nw || x != X_LO) atomicOr(status, 4);
}

// ---------------------------------------------------------------------------
// Ordered positions of the zero codes (outliers) of each stage: per chunk of
// 4096 codes a count, a per-stage exclusive scan, then an ordered write.
// ---------------------------------------------------------------------------
constexpr int ZCHUNK = 4096;

struct ZStage {
  int64_t code_off, n;
  int32_t chunk0, nchunk;
};

template <typename CodeT>
__global__ void __launch_bounds__(256) zcount_kernel(const ZStage *__restrict__ zs, int ns,
                                                     const int32_t *__restrict__ chunk_stage,
                                                     const CodeT *__restrict__ codes,
                                                     uint32_t *__restrict__ chunk_cnt) {
  const int gc = blockIdx.x;
  const ZStage Z = zs[chunk_stage[gc]];
  const int64_t c0 = (int64_t)(gc - Z.chunk0) * ZCHUNK;
  uint32_t cnt = 0;
  for (int i = threadIdx.x; i < ZCHUNK; i += 256) {
    const int64_t j = c0 + i;
    if (j < Z.n && codes[Z.code_off + j] == 0) cnt++;
  }
  __shared__ uint32_t red[256];
  red[threadIdx.x] = cnt;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < (unsigned)d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) chunk_cnt[gc] = red[0];
}

// Exclusive scan of the chunk counts of stage blockIdx.x (in place) and its
// outlier total.
__global__ void __launch_bounds__(1024) zscan_kernel(const ZStage *__restrict__ zs,
                                                     uint32_t *__restrict__ chunk_cnt,
                                                     uint32_t *__restrict__ out_count) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t carry;
  const ZStage Z = zs[blockIdx.x];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int c0 = 0; c0 < Z.nchunk; c0 += 1024) {
    const int c = c0 + threadIdx.x;
    const uint32_t v = c < Z.nchunk ? chunk_cnt[Z.chunk0 + c] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const uint32_t a = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
      __syncthreads();
      part[threadIdx.x] += a;
      __syncthreads();
    }
    if (c < Z.nchunk) chunk_cnt[Z.chunk0 + c] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out_count[blockIdx.x] = carry;
}

// With x / out_val: each outlier's value x[rank] (in stage order) goes to
// out_val beside its rank.
template <typename CodeT>
__global__ void __launch_bounds__(256) zwrite_kernel(const ZStage *__restrict__ zs,
                                                     const int32_t *__restrict__ chunk_stage,
                                                     const CodeT *__restrict__ codes,
                                                     const uint32_t *__restrict__ chunk_cnt,
                                                     uint32_t *__restrict__ out_rank,
                                                     const float *__restrict__ x = nullptr,
                                                     float *__restrict__ out_val = nullptr) {
  const int gc = blockIdx.x;
  const int si = chunk_stage[gc];
  const ZStage Z = zs[si];
  const int64_t c0 = (int64_t)(gc - Z.chunk0) * ZCHUNK;
  __shared__ uint32_t wave_cnt[4];
  __shared__ uint32_t base_sh;
  if (threadIdx.x == 0) base_sh = chunk_cnt[gc];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t *dst = out_rank + Z.code_off;
  for (int i0 = 0; i0 < ZCHUNK; i0 += 256) {
    const int64_t j = c0 + i0 + threadIdx.x;
    const bool z = j < Z.n && codes[Z.code_off + j] == 0;
    const uint64_t m = __ballot(z);
    if (lane == 0) wave_cnt[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pre = base_sh;
    for (int v = 0; v < wv; v++) pre += wave_cnt[v];
    if (z) {
      const uint32_t k = pre + (uint32_t)__popcll(m & ((1ull << lane) - 1));
      dst[k] = (uint32_t)j;
      if (out_val) out_val[Z.code_off + k] = x[Z.code_off + j];
    }
    __syncthreads();
    if (threadIdx.x == 0) base_sh += wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    __syncthreads();
  }
}

}  // namespace mlr
}  // namespace latec

using namespace latec;
using namespace latec::mlr;

static int64_t al4(int64_t w) { return (w + 3) & ~(int64_t)3; }

// Encoder scratch (u32 words): every lane's words (<= its symbols), then per
// lane the emitted count (u32) and state (u64), the lane offsets (lanes +
// ns + 1), per-stage word totals (i64), stage descriptors, lane -> stage, and
// a status word.
static int64_t enc_scratch_words(int ns, int64_t need, int64_t lanes) {
  return al4(need) + al4(lanes) + al4(2 * lanes) + al4(lanes + ns + 1) + al4(2 * (int64_t)ns) +
         al4((int64_t)(sizeof(Stage) / 4) * ns) + al4(lanes) + 4;
}


// Workspace of latec_mlrans_encode_noalloc: a fixed device array, one copy per
// GPU for the life of the process (2^23 words = 32 MiB).  Host code never
// passes the array's name to a kernel or a copy (in host code it is only a
// stand-in for the symbol): it uses the device address hipGetSymbolAddress
// returns, looked up once per device with that device current.
static constexpr int64_t NOALLOC_WORDS = (int64_t)1 << 23;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static uint32_t *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// The current device's address of `scratch` (looked up on its first use).
static int noalloc_scratch(uint32_t **out) {
  int dev = 0;
  int rc = hip_ok(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    set_error("latec_mlrans_encode_noalloc: device index out of range");
    return LATEC_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    if ((rc = hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")))
      return rc;
    g_scratch_ptr[dev] = (uint32_t *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return LATEC_OK;
}

extern "C" {

int64_t latec_mlrans_lanes(int64_t n) { return lanes_for(n); }

int64_t latec_mlrans_scratch_words(int ns, const int64_t *n) {
  int64_t need = 0, lanes = 0;
  for (int s = 0; s < ns; s++) {
    need += std::max<int64_t>(n[s], 0);
    lanes += lanes_for(n[s]);
  }
  return enc_scratch_words(ns, need, lanes);
}

int latec_mlrans_encode(int ns, const void *d_codes, int code_bytes, const int64_t *code_off,
                        const int64_t *n, const int32_t *kexp, uint32_t radius,
                        const uint32_t *d_rows, const int64_t *row_off, uint32_t *d_scratch,
                        int64_t scratch_words, uint8_t *d_out, int64_t out_cap,
                        int64_t *blob_off, int64_t *blob_len, void *stream) {
  if (ns < 0 || (code_bytes != 2 && code_bytes != 4) || !n || !kexp || !code_off ||
      !row_off || !blob_off || !blob_len) {
    set_error("latec_mlrans_encode: bad arguments");
    return LATEC_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<Stage> st(ns);
  std::vector<int32_t> lane_stage;
  int64_t need = 0;
  for (int i = 0; i < ns; i++) {
    Stage &S = st[i];
    memset(&S, 0, sizeof S);
    if (kexp[i] < 1 || kexp[i] > 15) {
      set_error("rANS coder window does not fit the stage frame");
      return LATEC_E_ARG;
    }
    S.code_off = code_off[i];
    S.n = n[i];
    S.row_off = row_off[i];
    S.kexp = kexp[i];
    S.lanes = (int32_t)lanes_for(n[i]);
    S.lane_base = (int32_t)lane_stage.size();
    S.scratch = need;
    need += n[i];
    for (int l = 0; l < S.lanes; l++) lane_stage.push_back(i);
  }
  const int64_t total_lanes = (int64_t)lane_stage.size();
  const int64_t w_words = al4(need), w_emit = al4(total_lanes), w_state = al4(2 * total_lanes);
  const int64_t w_off = al4(total_lanes + ns + 1), w_sw = al4(2 * (int64_t)ns);
  const int64_t w_st = al4((int64_t)(sizeof(Stage) / 4) * ns), w_ls = al4(total_lanes);
  const int64_t total = enc_scratch_words(ns, need, total_lanes);
  if (total > scratch_words) {
    set_error("latec_mlrans_encode: scratch too small (" + std::to_string(total) + " words)");
    return LATEC_E_CAPACITY;
  }
  uint32_t *emitted = d_scratch + w_words;
  uint64_t *state = (uint64_t *)(emitted + w_emit);
  uint32_t *lane_off = (uint32_t *)(state) + w_state;
  int64_t *stage_words = (int64_t *)(lane_off + w_off);
  Stage *d_st = (Stage *)((uint32_t *)stage_words + w_sw);
  int32_t *d_ls = (int32_t *)((uint32_t *)d_st + w_st);
  int32_t *d_status = (int32_t *)(d_ls + w_ls);
  int rc;
  if (ns == 0) return LATEC_OK;
  const int32_t zero = 0;
  if ((rc = hip_ok(hipMemcpyAsync(d_st, st.data(), sizeof(Stage) * ns, hipMemcpyHostToDevice,
                                  s), "hipMemcpyAsync")) ||
      (rc = hip_ok(hipMemcpyAsync(d_status, &zero, sizeof zero, hipMemcpyHostToDevice, s),
                   "hipMemcpyAsync")))
    return rc;
  if (total_lanes &&
      (rc = hip_ok(hipMemcpyAsync(d_ls, lane_stage.data(), sizeof(int32_t) * total_lanes,
                                  hipMemcpyHostToDevice, s), "hipMemcpyAsync")))
    return rc;
  const unsigned nb = (unsigned)((total_lanes + 255) / 256);
  if (total_lanes) {
    if (code_bytes == 2)
      hipLaunchKernelGGL(enc_kernel<uint16_t>, dim3(nb), dim3(256), 0, s, d_st, d_ls,
                         (int32_t)total_lanes, (const uint16_t *)d_codes, radius, d_rows,
                         d_scratch, emitted, state, d_status);
    else
      hipLaunchKernelGGL(enc_kernel<uint32_t>, dim3(nb), dim3(256), 0, s, d_st, d_ls,
                         (int32_t)total_lanes, (const uint32_t *)d_codes, radius, d_rows,
                         d_scratch, emitted, state, d_status);
    if ((rc = hip_ok(hipGetLastError(), "mlrans enc_kernel"))) return rc;
  }
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, d_st, ns, emitted, lane_off,
                     stage_words);
  if ((rc = hip_ok(hipGetLastError(), "mlrans scan_kernel"))) return rc;
  std::vector<int64_t> sw(ns);
  int32_t status = 0;
  if ((rc = hip_ok(hipMemcpyAsync(sw.data(), stage_words, sizeof(int64_t) * ns,
                                  hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) ||
      (rc = hip_ok(hipMemcpyAsync(&status, d_status, sizeof status, hipMemcpyDeviceToHost, s),
                   "hipMemcpyAsync")) ||
      (rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize")))
    return rc;
  if (status) {
    set_error("code outside the narrowed rANS alphabet");
    return LATEC_E_ARG;
  }
  // blob layout in d_out: stage blobs back to back (4-byte aligned)
  std::vector<int64_t> word0(ns);
  int64_t at = 0;
  for (int i = 0; i < ns; i++) {
    word0[i] = at;
    const int64_t L = st[i].lanes;
    const int64_t words = L ? 1 + (L + 1) + sw[i] : 0;
    blob_off[i] = 4 * at;
    blob_len[i] = 4 * words;
    at += words;
  }
  if (4 * at > out_cap) {
    set_error("latec_mlrans_encode: output capacity " + std::to_string(out_cap) +
              " < " + std::to_string(4 * at));
    return LATEC_E_CAPACITY;
  }
  // stage word bases live in the scratch tail; reuse stage_words for word0
  if ((rc = hip_ok(hipMemcpyAsync(stage_words, word0.data(), sizeof(int64_t) * ns,
                                  hipMemcpyHostToDevice, s), "hipMemcpyAsync")))
    return rc;
  if (total_lanes) {
    hipLaunchKernelGGL(assemble_kernel, dim3(nb), dim3(256), 0, s, d_st, d_ls,
                       (int32_t)total_lanes, d_scratch, emitted, state, lane_off, stage_words,
                       (uint32_t *)d_out);
    if ((rc = hip_ok(hipGetLastError(), "mlrans assemble_kernel"))) return rc;
  }
  return hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors end here
}

int64_t latec_mlrans_noalloc_words(void) { return NOALLOC_WORDS; }

int latec_mlrans_encode_noalloc(int ns, const void *d_codes, int code_bytes,
                                const int64_t *code_off, const int64_t *n, const int32_t *kexp,
                                uint32_t radius, const uint32_t *d_rows, const int64_t *row_off,
                                uint8_t *d_out, int64_t out_cap, int64_t *blob_off,
                                int64_t *blob_len, void *stream) {
  if (ns < 0 || !n) {
    set_error("latec_mlrans_encode_noalloc: bad arguments");
    return LATEC_E_ARG;
  }
  for (int i = 0; i < ns; i++)
    if (n[i] < 0 || n[i] > NOALLOC_WORDS) {
      set_error("latec_mlrans_encode_noalloc: bad stage size");
      return LATEC_E_ARG;
    }
  const int64_t need = latec_mlrans_scratch_words(ns, n);
  if (need > NOALLOC_WORDS) {
    set_error("latec_mlrans_encode_noalloc: these stages need " + std::to_string(need) +
              " workspace words, more than its fixed " + std::to_string(NOALLOC_WORDS) +
              "; use latec_mlrans_encode with a workspace");
    return LATEC_E_CAPACITY;
  }
  uint32_t *ws = nullptr;
  const int rc = noalloc_scratch(&ws);
  if (rc) return rc;
  return latec_mlrans_encode(ns, d_codes, code_bytes, code_off, n, kexp, radius, d_rows, row_off,
                             ws, NOALLOC_WORDS, d_out, out_cap, blob_off, blob_len, stream);
}

int latec_mlrans_decode(int ns, const uint8_t *d_payload, int64_t payload_len,
                        const int64_t *blob_off, const int64_t *blob_len, const int64_t *n,
                        const int32_t *kexp, uint32_t radius, const uint32_t *d_rows,
                        const int64_t *row_off, void *d_codes, int code_bytes,
                        const int64_t *code_off, uint32_t *d_scratch, int64_t scratch_words,
                        void *stream) {
  if (ns < 0 || (code_bytes != 2 && code_bytes != 4)) {
    set_error("latec_mlrans_decode: bad arguments");
    return LATEC_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<Stage> st(ns);
  std::vector<DecBlock> blocks;
  for (int i = 0; i < ns; i++) {
    Stage &S = st[i];
    memset(&S, 0, sizeof S);
    S.code_off = code_off[i];
    S.n = n[i];
    S.row_off = row_off[i];
    S.kexp = kexp[i];
    S.lanes = (int32_t)lanes_for(n[i]);
    S.scratch = blob_off[i];
    S.word_base = blob_len[i] / 4;
    if (!n[i]) {
      if (blob_len[i]) {
        set_error("multi-lane blob of an empty stage is not empty");
        return LATEC_E_FORMAT;
      }
      continue;
    }
    if (kexp[i] < 1 || kexp[i] > 15) {
      set_error("rANS stage record has no coder window");
      return LATEC_E_FORMAT;
    }
    if (blob_off[i] % 4 || blob_off[i] < 0 || blob_off[i] + blob_len[i] > payload_len ||
        blob_len[i] < 4 * (2 + (int64_t)S.lanes + 2 * (int64_t)S.lanes)) {
      set_error("multi-lane blob is truncated or misaligned");
      return LATEC_E_FORMAT;
    }
    for (int l0 = 0; l0 < S.lanes; l0 += 256) blocks.push_back({i, l0});
  }
  const int64_t nbk = (int64_t)blocks.size();
  const int64_t w_st = ((int64_t)(sizeof(Stage) / 4) * ns + 3) & ~3ll;
  const int64_t need = w_st + 2 * nbk + 4;
  if (need > scratch_words) {
    set_error("latec_mlrans_decode: scratch too small");
    return LATEC_E_CAPACITY;
  }
  if (!nbk) return LATEC_OK;
  Stage *d_st = (Stage *)d_scratch;
  DecBlock *d_bk = (DecBlock *)(d_scratch + w_st);
  int32_t *d_status = (int32_t *)(d_scratch + w_st + 2 * nbk);
  int rc;
  const int32_t zero = 0;
  if ((rc = hip_ok(hipMemcpyAsync(d_st, st.data(), sizeof(Stage) * ns, hipMemcpyHostToDevice,
                                  s), "hipMemcpyAsync")) ||
      (rc = hip_ok(hipMemcpyAsync(d_bk, blocks.data(), sizeof(DecBlock) * nbk,
                                  hipMemcpyHostToDevice, s), "hipMemcpyAsync")) ||
      (rc = hip_ok(hipMemcpyAsync(d_status, &zero, sizeof zero, hipMemcpyHostToDevice, s),
                   "hipMemcpyAsync")))
    return rc;
  // the kernel checks each lane's offsets against its blob before reading
  // words (latec_amd/mlrans.py also validates every blob header on the host
  // before this call) and flags malformed lanes in status
  if (code_bytes == 2)
    hipLaunchKernelGGL(dec_kernel<uint16_t>, dim3((unsigned)nbk), dim3(256), 0, s, d_st, d_bk,
                       d_payload, radius, d_rows, (uint16_t *)d_codes, d_status);
  else
    hipLaunchKernelGGL(dec_kernel<uint32_t>, dim3((unsigned)nbk), dim3(256), 0, s, d_st, d_bk,
                       d_payload, radius, d_rows, (uint32_t *)d_codes, d_status);
  if ((rc = hip_ok(hipGetLastError(), "mlrans dec_kernel"))) return rc;
  int32_t status = 0;
  if ((rc = hip_ok(hipMemcpyAsync(&status, d_status, sizeof status, hipMemcpyDeviceToHost, s),
                   "hipMemcpyAsync")) ||
      (rc = hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize")))
    return rc;
  if (status) {
    set_error("corrupt multi-lane rANS blob (status " + std::to_string(status) + ")");
    return LATEC_E_FORMAT;
  }
  return LATEC_OK;
}

static int zero_ranks(int ns, const void *d_codes, int code_bytes, const int64_t *code_off,
                      const int64_t *n, const float *d_x, uint32_t *d_out_rank,
                      float *d_out_val, uint32_t *d_out_count, uint32_t *d_scratch,
                      int64_t scratch_words, void *stream) {
  if (ns < 0 || (code_bytes != 2 && code_bytes != 4)) {
    set_error("latec_zero_ranks: bad arguments");
    return LATEC_E_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<ZStage> zs(ns);
  std::vector<int32_t> chunk_stage;
  for (int i = 0; i < ns; i++) {
    zs[i].code_off = code_off[i];
    zs[i].n = n[i];
    zs[i].chunk0 = (int32_t)chunk_stage.size();
    zs[i].nchunk = (int32_t)((n[i] + ZCHUNK - 1) / ZCHUNK);
    for (int c = 0; c < zs[i].nchunk; c++) chunk_stage.push_back(i);
  }
  const int64_t nc = (int64_t)chunk_stage.size();
  const int64_t w_zs = ((int64_t)(sizeof(ZStage) / 4) * ns + 3) & ~3ll;
  const int64_t need = w_zs + 2 * nc + 8;
  if (need > scratch_words) {
    set_error("latec_zero_ranks: scratch too small");
    return LATEC_E_CAPACITY;
  }
  ZStage *d_zs = (ZStage *)d_scratch;
  int32_t *d_cs = (int32_t *)(d_scratch + w_zs);
  uint32_t *d_cnt = d_scratch + w_zs + nc;
  int rc;
  if (ns && (rc = hip_ok(hipMemcpyAsync(d_zs, zs.data(), sizeof(ZStage) * ns,
                                        hipMemcpyHostToDevice, s), "hipMemcpyAsync")))
    return rc;
  if (nc && (rc = hip_ok(hipMemcpyAsync(d_cs, chunk_stage.data(), sizeof(int32_t) * nc,
                                        hipMemcpyHostToDevice, s), "hipMemcpyAsync")))
    return rc;
  if (nc) {
    if (code_bytes == 2)
      hipLaunchKernelGGL(zcount_kernel<uint16_t>, dim3((unsigned)nc), dim3(256), 0, s, d_zs, ns,
                         d_cs, (const uint16_t *)d_codes, d_cnt);
    else
      hipLaunchKernelGGL(zcount_kernel<uint32_t>, dim3((unsigned)nc), dim3(256), 0, s, d_zs, ns,
                         d_cs, (const uint32_t *)d_codes, d_cnt);
  }
  if (ns)
    hipLaunchKernelGGL(zscan_kernel, dim3((unsigned)ns), dim3(1024), 0, s, d_zs, d_cnt,
                       d_out_count);
  if (nc) {
    if (code_bytes == 2)
      hipLaunchKernelGGL(zwrite_kernel<uint16_t>, dim3((unsigned)nc), dim3(256), 0, s, d_zs, d_cs,
                         (const uint16_t *)d_codes, d_cnt, d_out_rank, d_x, d_out_val);
    else
      hipLaunchKernelGGL(zwrite_kernel<uint32_t>, dim3((unsigned)nc), dim3(256), 0, s, d_zs, d_cs,
                         (const uint32_t *)d_codes, d_cnt, d_out_rank, d_x, d_out_val);
  }
  if ((rc = hip_ok(hipGetLastError(), "zero-rank kernels"))) return rc;
  return hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize");  // host vectors end here
}

int latec_zero_ranks(int ns, const void *d_codes, int code_bytes, const int64_t *code_off,
                     const int64_t *n, uint32_t *d_out_rank, uint32_t *d_out_count,
                     uint32_t *d_scratch, int64_t scratch_words, void *stream) {
  return zero_ranks(ns, d_codes, code_bytes, code_off, n, nullptr, d_out_rank, nullptr,
                    d_out_count, d_scratch, scratch_words, stream);
}

int latec_order_outliers(int ns, const void *d_codes, int code_bytes, const int64_t *code_off,
                         const int64_t *n, const float *d_x, uint32_t *d_out_rank,
                         float *d_out_val, uint32_t *d_out_count, uint32_t *d_scratch,
                         int64_t scratch_words, void *stream) {
  if (!d_x || !d_out_val) {
    set_error("latec_order_outliers: values required");
    return LATEC_E_ARG;
  }
  return zero_ranks(ns, d_codes, code_bytes, code_off, n, d_x, d_out_rank, d_out_val,
                    d_out_count, d_scratch, scratch_words, stream);
}

int64_t latec_zero_ranks_words(int ns, const int64_t *n) {
  int64_t nc = 0;
  for (int i = 0; i < ns; i++) nc += (n[i] + ZCHUNK - 1) / ZCHUNK;
  return (((int64_t)(sizeof(ZStage) / 4) * ns + 3) & ~3ll) + 2 * nc + 8;
}

}  // extern "C"
