// Thread-local error buffer of the C ABI.
#include "capi_util.h"

char* vmbd_err_buf() {
  static thread_local char buf[512] = {0};
  return buf;
}

extern "C" const char* vmbd_last_error(void) { return vmbd_err_buf(); }
