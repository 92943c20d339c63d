// rtm_host_internal.h -- shared internals of librtm_host.so.
#pragma once
#include <cstdarg>
#include <string>

namespace rtmh {

enum { RTMH_OK = 0, RTMH_E_ARG = -1, RTMH_E_IO = -10, RTMH_E_FORMAT = -11, RTMH_E_DEVICE = -12,
       RTMH_E_DESIGN = -13, RTMH_E_REFUSED = -14, RTMH_E_UNSTABLE = -15 };

// The value of pi the reference uses throughout (kernel.cu:16,
// LSMOrCon_rec_2D.cpp:4); the Ricker wavelet and the coefficient design
// depend on it at the last bit.
constexpr double kRefPi = 3.1415926535898;

extern thread_local std::string g_err;
int set_error(int code, const char* fmt, ...);

}  // namespace rtmh
