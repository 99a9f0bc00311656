// abi_util.hpp - host-side helpers of the C ABI: thread-local error text,
// status returns, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "mivegas.h"

namespace mv {

std::string& last_error_slot();

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error_slot() = buf;
  return code;
}

inline int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MV_ERR_LAUNCH, "%s: launch failed: %s", what, hipGetErrorString(e));
  return MV_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int validate_step(const mv_step* s, const char* what, int max_dim) {
  if (!s) return fail(MV_ERR_ARG, "%s: step is NULL", what);
  if (s->n_dim < 1) return fail(MV_ERR_ARG, "%s: n_dim must be >= 1 (got %d)", what, s->n_dim);
  if (max_dim > 0 && s->n_dim > max_dim)
    return fail(MV_ERR_UNSUPPORTED, "%s: n_dim %d exceeds %d", what, s->n_dim, max_dim);
  if (s->precision != MV_F64 && s->precision != MV_MIXED)
    return fail(MV_ERR_ARG, "%s: unknown precision %d", what, s->precision);
  if (s->n_events < 0) return fail(MV_ERR_ARG, "%s: negative n_events", what);
  if ((s->xmin == nullptr) != (s->xdelta == nullptr))
    return fail(MV_ERR_ARG, "%s: xmin and xdelta must both be given or both be NULL", what);
  return MV_OK;
}

inline int validate_fn(const mv_integrand* fn, const char* what) {
  if (!fn) return fail(MV_ERR_ARG, "%s: integrand is NULL", what);
  if (fn->family < MV_FN_CONSTANT || fn->family > MV_FN_BWPOWER)
    return fail(MV_ERR_UNSUPPORTED, "%s: unknown integrand family %d", what, fn->family);
  if (fn->family == MV_FN_GAUSSMIX && (fn->n_comp < 1 || fn->n_comp > MV_MAX_COMP))
    return fail(MV_ERR_ARG, "%s: mixture needs 1..%d components", what, MV_MAX_COMP);
  return MV_OK;
}

}  // namespace mv
