// latec_sheet.h -- arguments of the sheet kernels (latec_sheet.hip), built by
// the level launcher in latec_interp.hip.
//
// A 3-D level above the first (compact grid n0 x n1 x n2, stride 1, the
// coarse points at even coordinates) runs as two launches: the even sheets
// (i even: the coarse copy and classes 001, 010, 011) and then the odd sheets
// (i odd: classes 100, 101, 110, 111), which read the even sheets as their
// axis-0 taps.  Within a sheet every class's taps are either on other sheets
// or on rows and pairs of the same sheet, so a workgroup marches along the
// rows (axis 1), a wave owning a run of 61 pairs of the sheets' lines and
// taking the k neighbours of a row from its neighbouring lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "latec_level.h"

namespace latec {

#define SHEET_SEG 16  // most odd rows per march segment

struct SheetClass {
  int32_t stage, pad;  // index into SheetArgs::st
  int64_t offset;      // stage rank of the class's first point
  int64_t rstr[3];     // stage-rank strides per class index
};

struct SheetArgs {
  int32_t n[3];          // the level's compact grid
  int32_t cn[3];         // the level below's compact grid (even sheets)
  const float *coarse;   // even sheets: the level below's reconstruction
  uint32_t xs[3];        // element strides of x along the level's axes
  int32_t cubic, center, end_mode, round_mode;
  double span, offset;
  int64_t hist_words;
  uint32_t npairs;       // pairs per line, n2 / 2 (n2 even, at most 512 points)
  uint32_t nsheets;      // sheets of the launch's parity
  uint32_t nruns;        // wave runs per segment: ceil(nsheets * npairs / 61)
  uint32_t rgroups;      // workgroups per segment: ceil(nruns / 4)
  uint32_t seg;          // odd rows per march segment (4, 8 or SHEET_SEG)
  uint32_t nseg;         // segments per sheet
  L3Stage st[3];
  // classes by (row parity, k parity): [0] (e,e) [1] (e,o) [2] (o,e) [3] (o,o);
  // the even launch's (e,e) class is the coarse copy and codes nothing
  SheetClass cl[4];
  // each stage's histogram offset (gs * hist_words), read by the rare
  // emission path
  int64_t hoff[3];
};

// Launches one sheet parity (odd = false: even sheets) of a level.  bits: 1
// encode, 2 rounding check, 4 x is the level grid.
int sheet_launch(const SheetArgs &a, bool odd, int bits, const float *x, float *rec,
                 uint16_t *codes, uint32_t *hist, uint32_t *out_rank, float *out_val,
                 uint32_t *out_count, hipStream_t stream);

}  // namespace latec
