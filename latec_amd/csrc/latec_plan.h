// latec_plan.h -- the dyadic stage schedule as parity-class boxes.
#pragma once
#include <stdint.h>

#include <vector>

#include "latec_common.h"

namespace latec {

// One parity class of a stage: a box of sizes[] lattice points at
// starts[j] + step * m[j] on the grid, with stage-array rank
// offset + sum_j m[j] * rstr[j] (levels.py:389-512).  For the anchor stage of
// an anchor_block > 1 schedule, coordinate m maps to
// (m / block) * stride + m % block instead (flag ANCHOR_BLOCK).
struct ClassDesc {
  int32_t parity;  // bit j: axis j is a midpoint axis of this class
  int32_t step;
  int32_t flags;   // bit 0: anchor-block coordinate map
  int32_t ndim;
  int64_t offset;
  int64_t size;
  int64_t starts[LATEC_MAX_DIM];
  int64_t rstr[LATEC_MAX_DIM];
  uint32_t sizes[LATEC_MAX_DIM];
  uint32_t magic[LATEC_MAX_DIM];  // fast division by sizes[j]
  uint32_t shift[LATEC_MAX_DIM];
  uint32_t row_len;  // sizes[ndim-1]: points per row along the last axis
  uint32_t nrows;    // size / row_len
  int32_t tr;        // threads per row in a 256-thread block (power of two)
  int32_t pad;
};

struct StageDesc {
  int stride = 0, weight = 0;
  int64_t count = 0;      // points per channel
  int64_t offset = 0;     // in the concatenated per-channel stage arrays
  int64_t count1 = 0;     // anchor_block == 1 layout size (before exclusion)
  std::vector<ClassDesc> classes;
  std::vector<uint32_t> excluded;  // sorted block-1 ranks of anchor points
};

struct DeviceBufs;  // plan-owned device copies (class tables, exclusion lists)

}  // namespace latec

struct latec_plan {
  int ndim = 0, channels = 1, levels = 0, stride = 0, block = 1;
  int64_t shape[LATEC_MAX_DIM] = {0};
  int64_t gstr[LATEC_MAX_DIM] = {0};  // row-major grid strides
  int64_t npoints = 0;
  std::vector<latec::StageDesc> stages;
  latec::DeviceBufs *dev = nullptr;
};

namespace latec {
int plan_build(latec_plan *p);
void fastdiv_magic(uint32_t d, uint32_t *magic, uint32_t *shift);
}  // namespace latec
