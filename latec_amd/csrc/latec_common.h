// latec_common.h -- shared declarations of the LATEC MI355X library.
#pragma once
#include <stdint.h>
#include <functional>
#include <string>

#include "../../include/latec.h"

namespace latec {

void set_error(const std::string &msg);
// the calling thread's last error message (what latec_last_error returns there)
const std::string &thread_error();

// Receives a stage record (bytes, length) on the thread that wrote it.
using RecordHook = std::function<void(const uint8_t *, int64_t)>;
void set_hip_error(const char *what, int code);

// Narrowed-symbol helpers (latec/rans.py:83-97).
inline uint32_t narrow_sym(uint32_t code, int64_t R, int kexp) {
  int64_t span = ((int64_t)1 << kexp) - 1;
  return code == 0 ? 0u : (uint32_t)((int64_t)code - R + span + 1);
}
inline uint32_t widen_sym(uint32_t sym, int64_t R, int kexp) {
  int64_t span = ((int64_t)1 << kexp) - 1;
  return sym == 0 ? 0u : (uint32_t)((int64_t)sym - span - 1 + R);
}

}  // namespace latec
