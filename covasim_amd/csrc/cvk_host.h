// Host-side error plumbing shared by the covasim_amd translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/covasim_amd.h"

namespace cvk {

// Thread-local message of the last failure (returned by cvk_last_error()).
char* last_error_buffer();

// CVK_TUNE_OVERLAP_RESAMPLE as set by cvk_set_tuning (read by the batched day loop too).
int overlap_resample_enabled();
// CVK_TUNE_GENERATE_DYNAMIC applied to a run with dyn_edges dynamic-layer edges in all (diag: the run
// tests and traces, which reads each day's contacts, so they are written).
bool generate_dynamic_for(int64_t dyn_edges, bool diag);
// CVK_TUNE_SOURCE_LISTS (0 never, 1 always, 2 by the day's infectious share) and CVK_TUNE_SOURCE_PERMILLE.
int source_lists_mode();
int source_permille();

// Testing and tracing step of the day loop (cvk_diag.hip): the ops of `imports`, then
// update_states_post; nothing when imports->diag is null.
int diag_day(const cvk_people* P, const cvk_layer* layers, const cvk_params* prm, cvk_work* W,
             const cvk_imports* imports, int32_t t, hipStream_t st);
int check_diag_schedule(const cvk_people* P, const cvk_layer* layers, const cvk_params* prm, const cvk_work* W,
                        const cvk_imports* imports);

inline int fail(int code, const char* msg) {
  snprintf(last_error_buffer(), 512, "%s", msg);
  return code;
}

// fail() with a printf-style message.
inline int failf(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(last_error_buffer(), 512, fmt, ap);
  va_end(ap);
  return code;
}

inline int check_launch(const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    snprintf(last_error_buffer(), 512, "%s: %s", where, hipGetErrorString(e));
    return CVK_E_HIP;
  }
  return CVK_OK;
}

}  // namespace cvk
