// Status strings and the per-thread last error of the C ABI.
#include "status.hpp"

namespace af {

thread_local std::string g_last_error;

af_status fail(af_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

}  // namespace af

extern "C" {

const char* af_status_str(af_status s) {
  switch (s) {
    case AF_OK: return "ok";
    case AF_EINVAL: return "invalid argument";
    case AF_ENOMEM: return "out of memory";
    case AF_EDEVICE: return "device error";
    case AF_ENAN: return "NaN in the solution";
    case AF_EIO: return "I/O error";
    case AF_EUNSUPPORTED: return "unsupported on this backend";
  }
  return "unknown status";
}

const char* af_last_error(void) { return af::g_last_error.c_str(); }

}  // extern "C"
