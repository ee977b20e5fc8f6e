// sanitize_stubs.cpp -- the device entry points of include/ctfp.h for the
// CPU-only sanitizer build (AddressSanitizer + UndefinedBehaviorSanitizer) of
// the host code: ctfp_host.cpp, ctfp_errors.cpp and the reference-shaped
// classes under refapi/.  That build runs the host-only paths (TestI0, the
// INI and table readers, the grid and scintillator models); every entry
// point that needs a device reports that this build has none.  Not part of
// libctfp.so.
#include "ctfp.h"
#include "ctfp_internal.h"

namespace {
int no_device() { return ctfp::set_error(CTFP_ERR_UNSUPPORTED, "sanitizer build: no device"); }
} // namespace

extern "C" {

int ctfp_create(const ctfp_scan_params *, int, ctfp_context **out)
{
    if (out) *out = nullptr;
    return no_device();
}
int ctfp_destroy(ctfp_context *) { return CTFP_OK; }
int ctfp_set_labels(ctfp_context *, const uint8_t *, int, int, const float *, int, int) { return no_device(); }
int ctfp_set_spectral(ctfp_context *, const ctfp_spectral *) { return no_device(); }
int ctfp_set_focal_spot(ctfp_context *, const float *, const float *) { return no_device(); }
int ctfp_forward_host(ctfp_context *, float *, float *, float *) { return no_device(); }

} // extern "C"
