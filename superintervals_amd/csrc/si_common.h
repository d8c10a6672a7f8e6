// si_common.h -- internal helpers shared by the library's translation units:
// export macro, status codes and the per-thread error message.
#pragma once

#include <cstdarg>
#include <cstdio>

#include "../../include/si_device.h"

#define SI_EXPORT extern "C" __attribute__((visibility("default")))

// Records a formatted message for siGetLastError() and returns code.
int si_set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
