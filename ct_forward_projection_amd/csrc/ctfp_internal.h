// ctfp_internal.h -- helpers shared by the device and host halves of libctfp.
#pragma once
#include <stdint.h>

#include "ctfp.h"

namespace ctfp {

// Records `fmt` as the calling thread's last error and returns `code`.
int set_error(int code, const char *fmt, ...);
const char *last_error();
void clear_error();

// The reference's FP32 view angle in radians,
// `rotatedDirection * i * thetaStep / 180 * PI` (CBCTPolyProjGPU.cu:51, PI =
// acosf(-1)).  For rotatedDirection = -1 the reference's int*size_t product
// wraps to a huge unsigned value; the intended angle -i*thetaStep is used.
float view_angle(const ctfp_scan_params *scan, const ctfp_system_info *sys, int64_t view);

} // namespace ctfp
