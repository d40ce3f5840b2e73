// Helpers of the runtime (ps_runtime.hip) used by the other device translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

int ps_internal_fail(int code, const std::string& msg);  // sets ps_last_error, returns code
int ps_internal_require_init();                          // PS_ENOTINIT before ps_init
hipStream_t ps_internal_stream(void* s);                 // NULL -> the library's stream 0
