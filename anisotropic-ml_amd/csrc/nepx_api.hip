// Library-level entry points: error text, version, struct-size checks, launch tuning.
#include <atomic>

#include "nepx_common.h"

namespace nepx {

static std::atomic<int32_t> g_tune[5] = {{0}, {0}, {0}, {0}, {0}};
static std::atomic<int32_t> g_cus[64] = {};

nepx_tuning tuning() {
    nepx_tuning t;
    t.fused_blocks_per_cu = g_tune[0].load() ? g_tune[0].load() : 1;
    t.pair_blocks_per_cu = g_tune[1].load() ? g_tune[1].load() : 16;
    t.gather_lanes = g_tune[2].load() ? g_tune[2].load() : 16;
    t.net_blocks_per_cu = g_tune[3].load() ? g_tune[3].load() : 2;
    t.pair_lanes = g_tune[4].load() ? g_tune[4].load() : 4;
    return t;
}

int device_cus() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = g_cus[dev].load();
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        g_cus[dev].store(c);
    }
    return c;
}

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
    out[6] = (int32_t)sizeof(nepx_tuning);
    return NEPX_OK;
}

extern "C" int nepx_set_tuning(const nepx_tuning* t) {
    NEPX_REQUIRE(t, NEPX_EINVAL, "null tuning");
    const int gl = t->gather_lanes;
    NEPX_REQUIRE(t->fused_blocks_per_cu >= 0 && t->fused_blocks_per_cu <= 8, NEPX_EINVAL,
                 "fused_blocks_per_cu %d not in 0..8", t->fused_blocks_per_cu);
    NEPX_REQUIRE(t->pair_blocks_per_cu >= 0 && t->pair_blocks_per_cu <= 64, NEPX_EINVAL,
                 "pair_blocks_per_cu %d not in 0..64", t->pair_blocks_per_cu);
    NEPX_REQUIRE(gl == 0 || gl == 4 || gl == 8 || gl == 16 || gl == 32 || gl == 64, NEPX_EINVAL,
                 "gather_lanes %d not in {4, 8, 16, 32, 64}", gl);
    NEPX_REQUIRE(t->net_blocks_per_cu >= 0 && t->net_blocks_per_cu <= 16, NEPX_EINVAL,
                 "net_blocks_per_cu %d not in 0..16", t->net_blocks_per_cu);
    NEPX_REQUIRE(t->pair_lanes == 0 || t->pair_lanes == 1 || t->pair_lanes == 4, NEPX_EINVAL,
                 "pair_lanes %d not in {1, 4}", t->pair_lanes);
    nepx::g_tune[0].store(t->fused_blocks_per_cu);
    nepx::g_tune[1].store(t->pair_blocks_per_cu);
    nepx::g_tune[2].store(gl);
    nepx::g_tune[3].store(t->net_blocks_per_cu);
    nepx::g_tune[4].store(t->pair_lanes);
    return NEPX_OK;
}

extern "C" int nepx_get_tuning(nepx_tuning* t) {
    NEPX_REQUIRE(t, NEPX_EINVAL, "null tuning");
    *t = nepx::tuning();
    return NEPX_OK;
}
