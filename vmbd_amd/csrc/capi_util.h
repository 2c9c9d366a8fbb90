// Error reporting shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/vmbd.h"

// thread-local message buffer behind vmbd_last_error() and the status code
// behind vmbd_last_status()
char* vmbd_err_buf();
int& vmbd_err_code();

inline void vmbd_clear_error() {
  vmbd_err_buf()[0] = '\0';
  vmbd_err_code() = VMBD_OK;
}

inline int vmbd_fail(int code, const char* fmt, ...) {
  vmbd_err_code() = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(vmbd_err_buf(), 512, fmt, ap);
  va_end(ap);
  return code;
}

inline int vmbd_fail_hip(const char* what, hipError_t e) {
  return vmbd_fail(VMBD_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define VMBD_CHECK_LAUNCH(expr)                                      \
  do {                                                               \
    hipError_t e__ = (expr);                                         \
    if (e__ != hipSuccess) return vmbd_fail_hip(#expr, e__);         \
  } while (0)
