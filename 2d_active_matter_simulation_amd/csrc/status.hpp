// Status and last-error plumbing of the C ABI, shared by the HIP stepper and
// the host-only code (file formats, initial conditions), which needs no HIP
// headers.
#pragma once
#include <string>

#include "../../include/actflow.h"

namespace af {

extern thread_local std::string g_last_error;
// Records msg as this thread's last error and returns s.
af_status fail(af_status s, const std::string& msg);

}  // namespace af
