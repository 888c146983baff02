// Thread-local error reporting for the C ABI: entry points return an MM_*
// code and leave a message retrievable with mm_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/molmot_hip.h"

inline thread_local char mm_err_buf[512] = {0};

inline void mm_clear_error() { mm_err_buf[0] = 0; }

inline int mm_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(mm_err_buf, sizeof(mm_err_buf), fmt, ap);
    va_end(ap);
    return code;
}

inline int mm_hip_fail(const char* what) {
    hipError_t e = hipGetLastError();
    return mm_fail(MM_EHIP, "%s: %s", what, hipGetErrorString(e));
}

inline int mm_check_launch(const char* kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mm_fail(MM_EHIP, "launch %s: %s", kernel, hipGetErrorString(e));
    return MM_OK;
}
