// errors.h -- internal error type of libhipdvr.so.  Thrown inside the
// library, caught at the C-ABI (api.hip) and turned into a status code plus
// a per-context message; never crosses the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/hipdvr.h"

namespace hipdvr {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(e == hipErrorOutOfMemory ? HIPDVR_ERR_NOMEM : HIPDVR_ERR_HIP,
             std::string(what) + ": " + hipGetErrorString(e));
    }
}

}  // namespace hipdvr
