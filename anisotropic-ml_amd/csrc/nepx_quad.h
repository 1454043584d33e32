// Fused pair step with the per-pair site contractions on matrix cores: four lanes per pair.
//
// Per pair, the descriptor forward and its adjoint hold three small dense
// contractions (nepx_desc.h has the math):
//   A_nlm   = sum_a f_n(a) Z_lm(a)                        forward accumulators
//   c_lm(a) = sum_n B_nlm f_n(a),  d_lm(a) = sum_n B_nlm f_n'(a)   adjoint, per site
// with a different small matrix pair for every pair of bodies: a batched
// product, the shape of v_mfma_f32_4x4x1_16b_f32.  That instruction runs
// sixteen independent 4x4 outer-product accumulations per wave; block b lives
// on lanes 4b..4b+3 and
//   D_b[i][j] += a(lane 4b+i) * b(lane 4b+j),   D_b[i][j] held by lane 4b+j, register i.
// A quad of lanes owns one pair and each lane evaluates every fourth site of
// the cloud (geometry, radial functions, harmonics and the per-site gradient
// terms stay on the VALU, one site per lane):
//  * forward: the quad swaps its four sites' f_n and Z_lm through LDS so that
//    lane i holds rows n = 4c+i and columns lm = 4d+i of all four sites; four
//    K=1 products per (c, d) leave A[4c+i][4d+j] in lane j, register i;
//  * adjoint: that layout is the A operand of c^T = B^T f (lane i holds
//    B[n][4d+i]), and each lane's own f_n is the B operand, so the product leaves
//    c_lm of lane j's own site, lm = 4d+i, in lane j, register i: every lane
//    receives its own site's c and d with no further exchange.
// Sums over the quad (radial sums, g_ang, forces and torques) use DPP quad
// permutes in a fixed order, so results do not depend on where a pair sits in
// the list.  The energy network is the same MFMA network as the one-lane-per-pair
// kernel, one 16-pair tile per wave.
#pragma once
#include "nepx_pairk.h"

#include <type_traits>

namespace nepx {

// l and m of accumulator slot lm (inverse of lm_slot).
__host__ __device__ constexpr int l_of_slot(int lm) {
    int l = 1;
    while ((l + 1) * (l + 1) - 1 <= lm) ++l;
    return l;
}
__host__ __device__ constexpr int m_of_slot(int lm) {
    return (lm - (l_of_slot(lm) * l_of_slot(lm) - 1) + 1) / 2;
}
// A 4-column chunk spans at most two l (every l has 2l+1 >= 3 slots): the
// first slot of the upper one, or past the chunk when it holds a single l.
template <int NLM>
__host__ __device__ constexpr int chunk_split(int d) {
    const int lo = l_of_slot(4 * d);
    const int last = 4 * d + 3 < NLM ? 4 * d + 3 : NLM - 1;
    const int hi = l_of_slot(last);
    return hi == lo ? 4 * d + 4 : hi * hi - 1;
}

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
// Sum over the four lanes of a quad (quad_perm 1,0,3,2 then 2,3,0,1), the same in every lane.
__device__ __forceinline__ float quad_sum(float v) {
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    return v;
}

// Orders LDS accesses of the lanes of one wave: the hardware executes a wave's
// LDS instructions in order, the fences keep the compiler from moving them.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sel4(int q, float v0, float v1, float v2, float v3) {
    return q == 0 ? v0 : q == 1 ? v1 : q == 2 ? v2 : v3;
}

// Each block takes BLOCK/4 consecutive pairs per iteration (16 per wave);
// waves share nothing but the network weights, so the loop has no block barriers.
template <class Cfg, class N, int PREC, int ACT>
__global__ __launch_bounds__(QuadCfg<Cfg>::BLOCK) __attribute__((
    amdgpu_waves_per_eu(QuadCfg<Cfg>::WAVES_PER_SIMD, QuadCfg<Cfg>::WAVES_PER_SIMD))) void k_fused_quad(
    nepx_pair_model m, const float* __restrict__ Wg, const void* __restrict__ Whg, float en_min, float scale,
    const float4* __restrict__ qf, const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
    const float4* __restrict__ pt, const int32_t* __restrict__ np_dev, int64_t maxp, float clip,
    float* __restrict__ out, RevOut rv) {
    using Q = QuadCfg<Cfg>;
    static_assert(Q::OK, "shape not supported by the four-lane kernel");
    constexpr int NC = Q::NC, LC = Q::LC, NA = Cfg::NANG + 1, NLM = Cfg::NLM, LMAX = Cfg::LMAX;
    constexpr int GA = Cfg::NRAD + 1;              // first angular entry of a descriptor row
    // descriptor row stride: whole 16-entry chunks of the network's dE/dG tiles (written without
    // a per-lane bound) plus 4 floats
    constexpr int LD = N::MD * 16 + 4;
    constexpr int PW = 16;                          // pairs per wave
    static_assert(LD <= Q::TS, "descriptor rows fit in the wave's exchange area");
    __shared__ __attribute__((aligned(16))) float xch[Q::BLOCK / 64][PW * Q::TS];
    __shared__ float4 stab[Cfg::ND];
    using R = NetRun<N, 1, PREC>;
    __shared__ __attribute__((aligned(16))) float wl[weight_lds_floats<N, PREC>()];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < Cfg::ND; ++a) {
            const int k = a % Cfg::NS;
            stab[a] = make_float4(m.sites[3 * k], m.sites[3 * k + 1], m.sites[3 * k + 2], 0.f);
        }
    }
    const float* W;
    const void* WH;
    stage_net<N, PREC>(wl, Wg, Whg, W, WH);
    __syncthreads();

    const int P = n_pairs(np_dev, maxp);
    const int tag = rev_tag(rv);
    const int lane = threadIdx.x & 63, q = lane & 3, wv = threadIdx.x >> 6, slot = lane >> 2;
    float* const xw = xch[wv];
    float* const xp = xw + slot * Q::TS;  // this pair's exchange area
    float* const grow = xw + slot * LD;   // its descriptor row (reuses the exchange area after the forward)
    const float inv_rn = 1.f / m.r_nep;
    constexpr K2Tab<LMAX> k2{};
    // this lane's accumulator column lm = 4d + q: K_lm^2 as the weight of the chunk's lower l
    // (klo) or of its upper l (khi), the other one zero
    float klo[LC], khi[LC];
#pragma unroll
    for (int d = 0; d < LC; ++d) {
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lm = 4 * d + i;
            v[i] = lm < NLM ? k2(l_of_slot(lm), m_of_slot(lm)) : 0.f;
        }
        const float kk = sel4(q, v[0], v[1], v[2], v[3]);
        const bool up = 4 * d + q >= chunk_split<NLM>(d);
        klo[d] = up ? 0.f : kk;
        khi[d] = up ? kk : 0.f;
    }

    // Two waves per SIMD: each iteration loads the next one's pair records (1 % faster at tet10;
    // with one wave per SIMD, cyl12, the registers they hold cost 4 %).
    constexpr bool AHEAD = Q::WAVES_PER_SIMD > 1;
    const int stride = gridDim.x * (Q::BLOCK / 4);
    float4 nqi = make_float4(1.f, 0.f, 0.f, 0.f), nqj = nqi, nt = make_float4(1.f, 0.f, 0.f, 0.f);
    if constexpr (AHEAD) {
        const int idx0 = blockIdx.x * (Q::BLOCK / 4) + (threadIdx.x >> 2);
        if (idx0 < P) load_pair(qf, pi, pj, pt, idx0, nqi, nqj, nt);
    }
    for (int base = blockIdx.x * (Q::BLOCK / 4); base < P; base += stride) {
        const int idx = base + (threadIdx.x >> 2);
        const bool valid = idx < P;
        float4 qi = nqi, qj = nqj, t = nt;  // lanes past the end keep finite stand-in records
        if constexpr (AHEAD) {
            if (idx + stride < P) load_pair(qf, pi, pj, pt, idx + stride, nqi, nqj, nt);
        } else {
            if (valid) load_pair(qf, pi, pj, pt, idx, qi, qj, t);
        }
        const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
        const Mat3 Ri = quat_matrix(qi.x, qi.y, qi.z, qi.w), Rj = quat_matrix(qj.x, qj.y, qj.z, qj.w);
        // geometry of this lane's site a = 4s + q (sites of body i first, then body j)
        auto site = [&](int s, bool& bj) {
            const int a = 4 * s + q;
            bj = a >= Cfg::NS;
            const float4 b = stab[a];
            Mat3 Rb;
#pragma unroll
            for (int k = 0; k < 9; ++k) Rb.m[k] = bj ? Rj.m[k] : Ri.m[k];
            return site_geo_b(Rb, b.x, b.y, b.z, (bj ? g.tx : 0.f) - g.cx, (bj ? g.ty : 0.f) - g.cy,
                              (bj ? g.tz : 0.f) - g.cz, inv_rn);
        };

        // ---- forward: A on matrix cores, radial sums per lane ----
        // The first trip starts the sums (the accumulation's first products take a zero C operand)
        // instead of zeroing 60 registers ahead of the loop.
        v4f Af[NC][LC];
        float gr[Cfg::NRAD + 1];
        auto fwd_trip = [&](int s, auto first_trip) {
            constexpr bool FIRST = decltype(first_trip)::value;
            bool bj;
            const SiteGeo sg = site(s, bj);
            float fv[4 * NC];
#pragma unroll
            for (int n = NA; n < 4 * NC; ++n) fv[n] = 0.f;
            radial_values<Cfg::NRAD>(sg.rho, 1.f, [&](int n, float f) {
                if constexpr (FIRST) gr[n] = f;
                else gr[n] += f;
                if (n < NA) fv[n] = f;
            });
            Harm<LMAX> h;
            h.eval(sg.ux, sg.uy, sg.uz);
            wave_lds_order();  // after the previous step's exchange reads
            // value 4*ch + i of this lane's site goes to xp[16 ch + 4 i + q]
#pragma unroll
            for (int n = 0; n < 4 * NC; ++n) xp[16 * (n / 4) + 4 * (n % 4) + q] = fv[n];
#pragma unroll
            for (int l = 1; l <= LMAX; ++l)
#pragma unroll
                for (int mm = 0; mm <= l; ++mm)
#pragma unroll
                    for (int part = 0; part < (mm ? 2 : 1); ++part) {
                        const int lm = lm_slot(l, mm, part);
                        xp[16 * (NC + lm / 4) + 4 * (lm % 4) + q] = h.z(l, mm, part);
                    }
#pragma unroll
            for (int lm = NLM; lm < 4 * LC; ++lm) xp[16 * (NC + lm / 4) + 4 * (lm % 4) + q] = 0.f;
            wave_lds_order();
            // element q of every chunk for the quad's four sites (register k = site 4s + k)
            v4f xf[NC], xz[LC];
#pragma unroll
            for (int c = 0; c < NC; ++c) xf[c] = *reinterpret_cast<const v4f*>(xp + 16 * c + 4 * q);
#pragma unroll
            for (int d = 0; d < LC; ++d) xz[d] = *reinterpret_cast<const v4f*>(xp + 16 * (NC + d) + 4 * q);
            // the accumulation issues as one block after the exchange (as in the adjoint below), at
            // the lower wave priority
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < LC; ++d)
                        Af[c][d] = __builtin_amdgcn_mfma_f32_4x4x1f32(
                            xf[c][k], xz[d][k], (FIRST && k == 0) ? v4f{0.f, 0.f, 0.f, 0.f} : Af[c][d], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(1);
        };
        // Both site loops are unrolled: on the trips whose four sites lie on one body the compiler
        // drops the per-lane body selection and the other body's sums (3 % faster than loops).
        fwd_trip(0, std::true_type{});
#pragma unroll
        for (int s = 1; s < Q::STEPS; ++s) fwd_trip(s, std::false_type{});
        // ---- descriptor row: g_rad, g_ang[n][l] = sum_m K^2 A^2 ----
#pragma unroll
        for (int n = 0; n <= Cfg::NRAD; ++n) gr[n] = quad_sum(gr[n]);
        float ga[NA][LMAX];
#pragma unroll
        for (int n = 0; n < NA; ++n)
#pragma unroll
            for (int l = 0; l < LMAX; ++l) ga[n][l] = 0.f;
#pragma unroll
        for (int d = 0; d < LC; ++d) {
            const int lo = l_of_slot(4 * d), hi = l_of_slot(4 * d + 3 < NLM ? 4 * d + 3 : NLM - 1);
#pragma unroll
            for (int n = 0; n < NA; ++n) {
                const float a = Af[n / 4][d][n % 4];
                const float a2 = a * a;
                ga[n][lo - 1] += klo[d] * a2;
                if (lo != hi) ga[n][hi - 1] += khi[d] * a2;
            }
        }
#pragma unroll
        for (int n = 0; n < NA; ++n)
#pragma unroll
            for (int l = 0; l < LMAX; ++l) ga[n][l] = quad_sum(ga[n][l]);
        wave_lds_order();  // the exchange reads are done before the rows overwrite the area
        {  // every lane of the quad holds the same sums and writes them (no branch); rows of pairs
           // past the end hold finite values of the stand-in pair and are unused
#pragma unroll
            for (int n = 0; n <= Cfg::NRAD; ++n) grow[n] = gr[n];
#pragma unroll
            for (int n = 0; n < NA; ++n)
#pragma unroll
                for (int l = 0; l < LMAX; ++l) grow[GA + n * LMAX + l] = ga[n][l];
        }
        wave_lds_order();
        // ---- energy network on the wave's 16 pairs: rows in, dE/dG rows out ----
        float epair;
        {
            float* row = xw + (lane & 15) * LD;
            v4f gin[1][R::IC];
            gather_chunks<R::IC>(row, true, Cfg::D, gin[0]);
            v4f dg[1][N::MD];
            float y[1];
            R::run(W, WH, ACT, gin, scale, dg, y);
            epair = __shfl(en_min + scale * y[0], slot);  // pair p's energy is in lanes with (lane & 15) == p
            scatter_chunks<N::MD>(row, true, 16 * N::MD, dg[0]);  // padding units' dE/dG are zero
        }
        wave_lds_order();
        // ---- adjoint ----
        float dGr[Cfg::NRAD + 1], ndGr[Cfg::NRAD + 1];
        float dGr_hi = 0.f;
#pragma unroll
        for (int n = 0; n <= Cfg::NRAD; ++n) {
            dGr[n] = grow[n];
            ndGr[n] = (float)n * dGr[n];
            if (n >= NA) dGr_hi += dGr[n];
        }
        const float E = epair;
        // B = dE/dA = 2 K^2 dE/dg_ang[n][l] A, in place (rows n >= NA stay zero)
#pragma unroll
        for (int d = 0; d < LC; ++d) {
            const int lo = l_of_slot(4 * d), hi = l_of_slot(4 * d + 3 < NLM ? 4 * d + 3 : NLM - 1);
#pragma unroll
            for (int n = 0; n < NA; ++n) {
                const float w = klo[d] * grow[GA + n * LMAX + lo - 1] +
                                (lo != hi ? khi[d] * grow[GA + n * LMAX + hi - 1] : 0.f);
                Af[n / 4][d][n % 4] *= 2.f * w;
            }
        }
        // per-lane sums over its sites on body i and on body j (the arm sums are the pair's s_i, s_j)
        float ig_x = 0.f, ig_y = 0.f, ig_z = 0.f, it_x = 0.f, it_y = 0.f, it_z = 0.f;
        float jg_x = 0.f, jg_y = 0.f, jg_z = 0.f, jt_x = 0.f, jt_y = 0.f, jt_z = 0.f;
#pragma unroll
        for (int s = 0; s < Q::STEPS; ++s) {
            bool bj;
            const SiteGeo sg = site(s, bj);
            float fa[NA], da[NA];
            float dEdrho = radial_adjoint<Cfg::NRAD, NA>(sg.rho, dGr, ndGr, dGr_hi, fa, da);
            // the c/d products issue as one block, with no other VALU work than the harmonics below
            // scheduled between them, while the other wave on the SIMD issues its VALU work (4 %
            // faster at tet10 and cyl12 than letting the compiler interleave them).  The wave drops
            // to the lower priority for the block and raises it after, so that the VALU work, the
            // bound, issues first when both waves of a SIMD are ready (4 % faster at tet10)
            v4f cc[LC], dd[LC];
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int d = 0; d < LC; ++d) {
                cc[d] = v4f{0.f, 0.f, 0.f, 0.f};
                dd[d] = v4f{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
            for (int n = 0; n < NA; ++n)
#pragma unroll
                for (int d = 0; d < LC; ++d) {
                    const float b = Af[n / 4][d][n % 4];
                    cc[d] = __builtin_amdgcn_mfma_f32_4x4x1f32(b, fa[n], cc[d], 0, 0, 0);
                    dd[d] = __builtin_amdgcn_mfma_f32_4x4x1f32(b, da[n], dd[d], 0, 0, 0);
                }
            Harm<LMAX> h;
            if constexpr (Q::WAVES_PER_SIMD > 1) {
                // two waves per SIMD: the harmonics (independent of the products) fill one slot
                // after every four products of the block (1.4 % faster at tet10; with one wave per
                // SIMD, cyl12, 2 % slower, so there the block stays pure)
                h.eval(sg.ux, sg.uy, sg.uz);
#pragma unroll
                for (int k = 0; k < 2 * NA * LC / 4; ++k) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);  // four matrix products
                    __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);  // one VALU instruction
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(1);
            if constexpr (Q::WAVES_PER_SIMD == 1) h.eval(sg.ux, sg.uy, sg.uz);
            float gx, gy, gz;
            angular_site_grad<LRange<1, LMAX>, true>(
                h, sg.ux, sg.uy,
                [&](int o, float& c, float& dv) {
                    c = cc[o / 4][o % 4];
                    dv = dd[o / 4][o % 4];
                },
                dEdrho, gx, gy, gz);
            float ex, ey, ez;
            site_dEdr(sg, dEdrho, inv_rn, gx, gy, gz, ex, ey, ez);
            const float tx = sg.ay * ez - sg.az * ey, ty = sg.az * ex - sg.ax * ez, tz = sg.ax * ey - sg.ay * ex;
            const float wj = bj ? 1.f : 0.f, wi = 1.f - wj;
            ig_x += wi * ex; ig_y += wi * ey; ig_z += wi * ez;
            it_x += wi * tx; it_y += wi * ty; it_z += wi * tz;
            jg_x += wj * ex; jg_y += wj * ey; jg_z += wj * ez;
            jt_x += wj * tx; jt_y += wj * ty; jt_z += wj * tz;
        }
        AdjAcc acc;
        acc.G1x = quad_sum(jg_x); acc.G1y = quad_sum(jg_y); acc.G1z = quad_sum(jg_z);
        acc.T1x = quad_sum(jt_x); acc.T1y = quad_sum(jt_y); acc.T1z = quad_sum(jt_z);
        acc.A1x = g.sjx; acc.A1y = g.sjy; acc.A1z = g.sjz;
        acc.G0x = quad_sum(ig_x); acc.G0y = quad_sum(ig_y); acc.G0z = quad_sum(ig_z);
        acc.T0x = quad_sum(it_x); acc.T0y = quad_sum(it_y); acc.T0z = quad_sum(it_z);
        acc.A0x = g.six; acc.A0y = g.siy; acc.A0z = g.siz;
        if (valid && q == 0) store_out(out, idx, adj_finish<Cfg>(acc), E, clip, rv, tag);
        wave_lds_order();  // the row reads are done before the next iteration's exchange writes
    }
}

}  // namespace nepx
