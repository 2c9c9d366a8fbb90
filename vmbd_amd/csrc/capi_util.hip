// Thread-local error buffer of the C ABI.
#include "capi_util.h"

char* vmbd_err_buf() {
  static thread_local char buf[512] = {0};
  return buf;
}

int& vmbd_err_code() {
  static thread_local int code = VMBD_OK;
  return code;
}

extern "C" const char* vmbd_last_error(void) { return vmbd_err_buf(); }

extern "C" int vmbd_last_status(void) { return vmbd_err_code(); }
