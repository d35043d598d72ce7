// Shared host-side plumbing of libtpf: error state and HIP checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/tpf.h"

namespace tpf {

void set_error(const std::string& msg);

// Records a failure and returns its status; used as `return fail(...)`.
inline tpf_status fail(tpf_status st, const std::string& msg) {
  set_error(msg);
  return st;
}

// Whether a CCD plan carries a grid metric (a non-uniform axis).
bool ccd_plan_has_metric(const tpf_ccd_plan* plan);

}  // namespace tpf

#define TPF_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::tpf::fail(TPF_ERR_RUNTIME,                                    \
                         std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TPF_TRY(call)                 \
  do {                                \
    tpf_status s_ = (call);           \
    if (s_ != TPF_OK) return s_;      \
  } while (0)
