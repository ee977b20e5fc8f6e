// ctfp_errors.cpp -- per-thread error strings and the library identity of
// libctfp (host code shared by the device and host halves).
#include <cstdarg>
#include <cstdio>
#include <string>

#include "ctfp.h"
#include "ctfp_internal.h"

namespace ctfp {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char *last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

} // namespace ctfp

extern "C" {

const char *ctfp_last_error(void) { return ctfp::last_error(); }
const char *ctfp_version(void) { return "ctfp 0.2.0 (gfx950)"; }

} // extern "C"
