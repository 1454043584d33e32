// Soft bead-bead repulsion between rigid bodies: the label potential of the
// reference's training-data generator (01/SamplerUtils.py:555-631, calc_en_unred):
//   E_ij = sum_{a in i, b in j} eps ((r0 - |p_b - p_a|) / sigma)^2   for |p_b - p_a| < r0
// with body-frame beads b_a (the same table for both bodies), p_a = X_i + R(q_i) b_a and
// p_b = X_i + t + R(q_j) b_b.  Forces and torques are analytic:
//   g = dE/dp_b = -2 eps (r0 - d) / sigma^2 * (p_b - p_a) / d
//   F_i = sum g,  tau_i = sum (R_i b_a) x g,  tau_j = -sum (R_j b_b) x g   (F_j = -F_i).
// Output per pair in the layout of the NEP pair kernels: {F_i, E}{tau_i, 0}{tau_j, 0},
// so the deterministic gather and the integrator take it unchanged.
//
// One wave per pair: the rotated beads of both bodies are staged in LDS, each
// lane takes bead b = lane + 64 k of body j against every bead a of body i,
// FP32 accumulation, fixed-order butterfly reduction over the wave.
#include "nepx_common.h"

namespace nepx {

constexpr int kBeadBlock = 256;                  // 4 waves, one pair each per iteration
constexpr int kBeadWaves = kBeadBlock / 64;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kBeadBlock) void k_bead_pairs(const float* __restrict__ beads, int nb, float eps,
                                                            float sigma, float r0, const float4* __restrict__ qf,
                                                            const int32_t* __restrict__ pi,
                                                            const int32_t* __restrict__ pj,
                                                            const float4* __restrict__ pt,
                                                            const int32_t* __restrict__ np_dev, int64_t maxp,
                                                            float clip, float* __restrict__ out) {
    __shared__ float4 arm_i[kBeadWaves][NEPX_MAX_BEADS], arm_j[kBeadWaves][NEPX_MAX_BEADS];
    const int64_t P = np_dev ? ((int64_t)*np_dev < maxp ? (int64_t)*np_dev : maxp) : maxp;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const float r02 = r0 * r0;
    const float c = -2.f * eps / (sigma * sigma);
    // every block runs the same number of iterations: barriers stay uniform
    for (int64_t base = (int64_t)blockIdx.x * kBeadWaves; base < P; base += (int64_t)gridDim.x * kBeadWaves) {
        const int64_t p = base + wv;
        const bool valid = p < P;
        float4 qi = make_float4(1.f, 0.f, 0.f, 0.f), qj = qi, t = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            qi = qf[pi[p]];
            qj = qf[pj[p]];
            t = pt[p];
        }
        const Mat3 Ri = quat_matrix(qi.x, qi.y, qi.z, qi.w), Rj = quat_matrix(qj.x, qj.y, qj.z, qj.w);
        for (int a = lane; a < nb; a += 64) {
            float x, y, z;
            mat_vec(Ri, beads[3 * a], beads[3 * a + 1], beads[3 * a + 2], x, y, z);
            arm_i[wv][a] = make_float4(x, y, z, 0.f);
            mat_vec(Rj, beads[3 * a], beads[3 * a + 1], beads[3 * a + 2], x, y, z);
            arm_j[wv][a] = make_float4(x, y, z, 0.f);
        }
        __syncthreads();
        float e = 0.f, fx = 0.f, fy = 0.f, fz = 0.f, tix = 0.f, tiy = 0.f, tiz = 0.f, tjx = 0.f, tjy = 0.f,
              tjz = 0.f;
        if (valid) {
            for (int b = lane; b < nb; b += 64) {
                const float4 ab = arm_j[wv][b];
                const float bx = t.x + ab.x, by = t.y + ab.y, bz = t.z + ab.z;
                float gbx = 0.f, gby = 0.f, gbz = 0.f;  // sum of g over a for this bead b
                for (int a = 0; a < nb; ++a) {
                    const float4 aa = arm_i[wv][a];
                    const float dx = bx - aa.x, dy = by - aa.y, dz = bz - aa.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < r02) {
                        const float d = sqrtf(d2);
                        const float w = r0 - d;
                        e += eps * (w * w) / (sigma * sigma);
                        const float s = c * w / d;  // dE/dd / d
                        const float gx = s * dx, gy = s * dy, gz = s * dz;
                        gbx += gx;
                        gby += gy;
                        gbz += gz;
                        tix += aa.y * gz - aa.z * gy;
                        tiy += aa.z * gx - aa.x * gz;
                        tiz += aa.x * gy - aa.y * gx;
                    }
                }
                fx += gbx;
                fy += gby;
                fz += gbz;
                tjx -= ab.y * gbz - ab.z * gby;
                tjy -= ab.z * gbx - ab.x * gbz;
                tjz -= ab.x * gby - ab.y * gbx;
            }
        }
        e = wave_sum(e);
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        fz = wave_sum(fz);
        tix = wave_sum(tix);
        tiy = wave_sum(tiy);
        tiz = wave_sum(tiz);
        tjx = wave_sum(tjx);
        tjy = wave_sum(tjy);
        tjz = wave_sum(tjz);
        if (valid && lane == 0) {
            auto cl = [clip](float v) { return clipf(v, clip); };
            float4* o4 = reinterpret_cast<float4*>(out) + 3 * p;
            o4[0] = make_float4(cl(fx), cl(fy), cl(fz), e);
            o4[1] = make_float4(cl(tix), cl(tiy), cl(tiz), 0.f);
            o4[2] = make_float4(cl(tjx), cl(tjy), cl(tjz), 0.f);
        }
        __syncthreads();  // the next iteration overwrites the staged beads
    }
}

}  // namespace nepx

using namespace nepx;

extern "C" int nepx_bead_pair_forces(const float* beads, int32_t n_beads, float eps, float sigma, float r0,
                                     const float* qf, const int32_t* pair_i, const int32_t* pair_j,
                                     const float* pair_t, const int32_t* n_pairs_dev, int64_t max_pairs,
                                     float clip, float* out, void* stream) {
    NEPX_REQUIRE(n_beads > 0 && n_beads <= NEPX_MAX_BEADS, NEPX_EINVAL, "n_beads %d not in 1..%d", n_beads,
                 NEPX_MAX_BEADS);
    NEPX_REQUIRE(sigma > 0.f && r0 > 0.f, NEPX_EINVAL, "sigma and r0 must be positive");
    if (max_pairs <= 0) return NEPX_OK;
    NEPX_REQUIRE(beads && qf && pair_i && pair_j && pair_t && out, NEPX_EINVAL, "null buffer");
    int64_t grid = (max_pairs + kBeadWaves - 1) / kBeadWaves;
    const int64_t cap = (int64_t)device_cus() * 8;
    if (grid > cap) grid = cap;
    hipLaunchKernelGGL(k_bead_pairs, dim3((unsigned)grid), dim3(kBeadBlock), 0, as_stream(stream), beads, n_beads,
                       eps, sigma, r0, (const float4*)qf, pair_i, pair_j, (const float4*)pair_t, n_pairs_dev,
                       max_pairs, clip, out);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}
