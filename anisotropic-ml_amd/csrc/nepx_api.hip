// Library-level entry points: error text, version, struct-size checks.
#include "nepx_common.h"

namespace nepx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace nepx

extern "C" const char* nepx_last_error(void) { return nepx::g_err; }

extern "C" int nepx_version(void) { return 1; }

extern "C" int nepx_struct_sizes(int32_t* out) {
    if (!out) return NEPX_EINVAL;
    out[0] = (int32_t)sizeof(nepx_pair_model);
    out[1] = (int32_t)sizeof(nepx_net);
    out[2] = (int32_t)sizeof(nepx_sysparam);
    out[3] = (int32_t)sizeof(nepx_thermo);
    out[4] = (int32_t)sizeof(nepx_state);
    out[5] = (int32_t)sizeof(nepx_nlist);
    return NEPX_OK;
}
