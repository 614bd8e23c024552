// tg_last_error / tg_version and the thread-local error slot behind tg::fail.
#include <cstdarg>
#include <cstdio>

#include "tensors_amd.h"

namespace tg {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace tg

extern "C" const char* tg_last_error(void) { return tg::g_err; }
extern "C" const char* tg_version(void) { return "tensors_amd 0.1 (gfx950)"; }
