// Four-lanes-per-pair fused kernel (nepx_quad.h) and its launcher.  Built in
// its own translation unit so that its matrix-core accumulators can be kept in
// VGPRs (-mllvm -amdgpu-mfma-vgpr-form), next to the VALU code that reads them.
#include "nepx_pairk.h"
#include "nepx_quad.h"

namespace nepx {

template <class Cfg, class N, int PREC>
static void launch_quad_p(const nepx_pair_model& m, const nepx_net* net, int grid, float scale, const float* qf,
                          const int32_t* pair_i, const int32_t* pair_j, const float* pair_t, const int32_t* np_dev,
                          int64_t maxp, float clip, float* out, RevOut rv, void* stream) {
    using Q = QuadCfg<Cfg>;
#define NEPX_LAUNCH_QUAD(A)                                                                                      \
    hipLaunchKernelGGL((k_fused_quad<Cfg, N, PREC, A>), dim3(grid), dim3(Q::BLOCK), 0, as_stream(stream), m,      \
                       net->w_packed, net->w_half, net->en_min, scale, (const float4*)qf, pair_i, pair_j,          \
                       (const float4*)pair_t, np_dev, maxp, clip, out, rv)
    if (net->act == NEPX_ACT_TANH) NEPX_LAUNCH_QUAD(NEPX_ACT_TANH);
    else if constexpr (PREC == NEPX_NET_F32 || PREC == NEPX_NET_F16X3) NEPX_LAUNCH_QUAD(NEPX_ACT_RELU);
#undef NEPX_LAUNCH_QUAD
}

template <class Cfg, class N>
int launch_fused_quad(const nepx_pair_model& m, const nepx_net* net, int prec, int max_blocks, float scale,
                      const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                      const int32_t* np_dev, int64_t maxp, float clip, float* out, RevOut rv, void* stream) {
    if constexpr (!QuadCfg<Cfg>::OK) {
        return NEPX_EUNSUP;
    } else {
        using Q = QuadCfg<Cfg>;
        int64_t g = (maxp + Q::BLOCK / 4 - 1) / (Q::BLOCK / 4);
        const int grid = (int)(g < 1 ? 1 : g > max_blocks ? max_blocks : g);
        switch (prec) {
            case NEPX_NET_F16X3:
                launch_quad_p<Cfg, N, NEPX_NET_F16X3>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                       clip, out, rv, stream);
                break;
            case NEPX_NET_BF16:
                launch_quad_p<Cfg, N, NEPX_NET_BF16>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                      clip, out, rv, stream);
                break;
            case NEPX_NET_BF16X3:
                launch_quad_p<Cfg, N, NEPX_NET_BF16X3>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev,
                                                        maxp, clip, out, rv, stream);
                break;
            default:
                launch_quad_p<Cfg, N, NEPX_NET_F32>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                     clip, out, rv, stream);
        }
        NEPX_LAUNCH_CHECK();
        return NEPX_OK;
    }
}

#define NEPX_QUAD_INSTANCE(T, N)                                                                              \
    template int launch_fused_quad<T, N>(const nepx_pair_model&, const nepx_net*, int, int, float, const float*, \
                                         const int32_t*, const int32_t*, const float*, const int32_t*, int64_t,  \
                                         float, float*, RevOut, void*);
NEPX_QUAD_INSTANCE(CfgCube8, Net40x64)
NEPX_QUAD_INSTANCE(CfgTet10, Net40x64)
NEPX_QUAD_INSTANCE(CfgCyl12, Net60x128)
#undef NEPX_QUAD_INSTANCE

}  // namespace nepx
