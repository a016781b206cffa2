// bs_common.hpp -- error reporting shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace winds {
// Record the message of a failing call on this thread and return code.
int set_error(int code, const std::string &msg);
void clear_error();
// Compute units of the current device (cached; 256 if unknown).
int device_cu_count();
// Every device allocation of the library.  When the environment variable
// WINDS_FAIL_ALLOC_BYTES holds a positive number, requests of at least that
// many bytes fail with hipErrorOutOfMemory instead (tests of the error paths).
hipError_t dev_alloc(void **p, size_t bytes);
}  // namespace winds

#include "bs_kernels.hpp"
#include "winds_bs.h"

namespace winds {
// Validate public options and convert them to the kernel configuration.
int opts_to_cfg(const winds_bs_opts *o, EvalCfg &c);
// Enqueue one evaluation (device layout of winds_bs_eval_dev) on st.
int eval_dev(const double *t3, int64_t nt, const Seg64 *seg, int64_t nseg, const EvalCfg &c,
             double *u3, void *ws, size_t wsb, hipStream_t st);
// Workspace bytes eval_dev needs for these sizes.
size_t eval_ws_bytes(int64_t nt, int64_t nseg, const EvalCfg &c);
}  // namespace winds
