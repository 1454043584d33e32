// Building blocks shared by the pair kernels (nepx_pair.hip, nepx_quad.hip):
// pair records, descriptor rows, the network's operand staging, and the
// descriptor / network shapes that have kernels.
#pragma once
#include "nepx_desc.h"
#include "nepx_ffnn.h"

namespace nepx {

constexpr int kBlock = 256;

// Occupancy target of the register-heavy adjoint/fused kernels (waves per SIMD).
#ifndef NEPX_ADJ_WAVES
#define NEPX_ADJ_WAVES 1
#endif
#define NEPX_ADJ_ATTR __attribute__((amdgpu_waves_per_eu(NEPX_ADJ_WAVES, NEPX_ADJ_WAVES)))

__device__ __forceinline__ int n_pairs(const int32_t* np_dev, int64_t maxp) {
    const int p = *np_dev;
    return p < maxp ? p : (int)maxp;
}

__device__ __forceinline__ void load_pair(const float4* __restrict__ qf, const int32_t* __restrict__ pi,
                                          const int32_t* __restrict__ pj, const float4* __restrict__ pt,
                                          int idx, float4& qi, float4& qj, float4& t) {
    const int i = pi[idx];
    const int j = pj[idx];
    t = pt[idx];
    qi = qf[i];
    qj = qf[j];
}

template <class Cfg>
__device__ __forceinline__ void store_row(float* __restrict__ dst, const float (&v)[Cfg::D]) {
    if constexpr (Cfg::D % 4 == 0) {
        float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
        for (int k = 0; k < Cfg::D / 4; ++k) d4[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < Cfg::D; ++k) dst[k] = v[k];
    }
}

template <class Cfg>
__device__ __forceinline__ void load_row(const float* __restrict__ src, float (&v)[Cfg::D]) {
    if constexpr (Cfg::D % 4 == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int k = 0; k < Cfg::D / 4; ++k) {
            const float4 x = s4[k];
            v[4 * k] = x.x;
            v[4 * k + 1] = x.y;
            v[4 * k + 2] = x.z;
            v[4 * k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < Cfg::D; ++k) v[k] = src[k];
    }
}

// Optional second output of the fused pair kernels: the j side of pair idx, {F_i, tag}{tau_j, 0},
// also goes to rrec at the pair's reverse-row entry pe[idx], so that the gather reads it
// contiguously along each particle's reverse row.  tag = flags[F_TAG] of this pass, which the
// kernel records in flags[F_RTAG].  All null: not written.
struct RevOut {
    const int32_t* pe = nullptr;
    float* rrec = nullptr;
    int32_t* flags = nullptr;
};

__device__ __forceinline__ int rev_tag(const RevOut& rv) {
    if (!rv.flags) return 0;
    const int tag = rv.flags[F_TAG];
    if (blockIdx.x == 0 && threadIdx.x == 0) rv.flags[F_RTAG] = tag;
    return tag;
}

__device__ __forceinline__ void store_out(float* __restrict__ out, int idx, const PairOut& o, float E,
                                          float clip, const RevOut& rv = RevOut{}, int tag = 0) {
    float4* o4 = reinterpret_cast<float4*>(out) + 3 * (size_t)idx;
    const float4 f = make_float4(clipf(o.fx, clip), clipf(o.fy, clip), clipf(o.fz, clip), E);
    const float4 tj = make_float4(clipf(o.tjx, clip), clipf(o.tjy, clip), clipf(o.tjz, clip), 0.f);
    o4[0] = f;
    o4[1] = make_float4(clipf(o.tix, clip), clipf(o.tiy, clip), clipf(o.tiz, clip), 0.f);
    o4[2] = tj;
    if (rv.rrec) {
        float4* r4 = reinterpret_cast<float4*>(rv.rrec) + 2 * (size_t)rv.pe[idx];
        r4[0] = make_float4(f.x, f.y, f.z, __int_as_float(tag));
        r4[1] = tj;
    }
}

// Input chunks of one pair for the MFMA tile: chunk c holds inputs
// 16c + 4*(lane/16) + 0..3 (zero past D).
template <int MC>
__device__ __forceinline__ void gather_chunks(const float* __restrict__ row, bool valid, int D, v4f (&gin)[MC]) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int c = 0; c < MC; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 16 * c + 4 * g + r;
            gin[c][r] = (valid && k < D) ? row[k] : 0.f;
        }
    }
}

template <int MC>
__device__ __forceinline__ void scatter_chunks(float* __restrict__ row, bool valid, int D, const v4f (&v)[MC]) {
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int c = 0; c < MC; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 16 * c + 4 * g + r;
            if (valid && k < D) row[k] = v[c][r];
        }
    }
}

// Networks whose packed operands fit this budget are staged in LDS once per block.
constexpr int kWeightLdsBytes = 64 * 1024;
template <int PREC>
constexpr int lowp_parts() {
    return PREC == NEPX_NET_BF16 ? 1 : 2;
}
template <class N, int PREC>
constexpr int operand_floats() {
    return PREC == NEPX_NET_F32 ? N::layout().total() : N::layout_h(lowp_parts<PREC>()).total() / 2;
}
template <class N, int PREC>
constexpr bool weights_in_lds() {
    return operand_floats<N, PREC>() * 4 <= kWeightLdsBytes;
}
template <class N, int PREC>
constexpr int weight_lds_floats() {
    return weights_in_lds<N, PREC>() ? operand_floats<N, PREC>() : 4;
}

constexpr int kNetTiles = 2;  // 16-pair tiles sharing each A-operand read

// Network over NT tiles of 16 pairs with the operands chosen at compile time.
template <class N, int NT, int PREC>
struct NetRun {
    static constexpr int IC = PREC == NEPX_NET_F32 ? N::MD : N::MDH;  // input chunks
    __device__ __forceinline__ static void run(const float* W, const void* WH, int act, const v4f (&gin)[NT][IC],
                                               float scale, v4f (&dg)[NT][N::MD], float (&y)[NT]) {
        if constexpr (PREC == NEPX_NET_F32) ffnn_tiles<N, NT>(W, act, gin, scale, dg, y);
        else ffnn_tiles_h<N, NT, PREC>(W, WH, act, gin, scale, dg, y);
    }
};

// Stage the operands of this precision into LDS when they fit; returns the pointers to use.
template <class N, int PREC>
__device__ __forceinline__ void stage_net(float* wl, const float* Wg, const void* Whg, const float*& W,
                                          const void*& WH) {
    W = Wg;
    WH = Whg;
    if constexpr (weights_in_lds<N, PREC>()) {
        if constexpr (PREC != NEPX_NET_F32) {
            stage_copy(wl, reinterpret_cast<const float*>(Whg), operand_floats<N, PREC>());
            WH = wl;
        } else {
            stage_copy(wl, Wg, operand_floats<N, PREC>());
            W = wl;
        }
        __syncthreads();
    }
}

// Shapes with kernels.
using CfgCube8 = DescCfg<8, 11, 6, 4>;     // cube, 8 vertices, D = 40
using CfgTet10 = DescCfg<10, 11, 6, 4>;    // tetrahedron, 4 vertices + 6 mid-edges, D = 40
using CfgCyl12 = DescCfg<12, 17, 6, 6>;    // twisted cylinder, 12 helix sites, D = 60
using CfgCubeV2 = DescCfg<6, 10, 8, 4>;    // reference cube_v2 with hypers 10 8 4, D = 47
using CfgThV1 = DescCfg<4, 10, 8, 4>;      // reference th_v1 with hypers 10 8 4, D = 47
using CfgCubeV2L10 = DescCfg<6, 10, 10, 10>;  // reference cube_v2 with hypers 10 10 10, D = 121
using CfgThV1L10 = DescCfg<4, 10, 10, 10>;    // reference th_v1 with hypers 10 10 10, D = 121

using Net40x64 = NetCfg<48, 64, 2>;        // 40 -> 64 -> 64 -> 1
using Net60x128 = NetCfg<64, 128, 2>;      // 60 -> 128 -> 128 -> 1
using Net47x150 = NetCfg<48, 160, 4>;      // reference TheEnergyNet(150, 3, 47): 47 -> 150 x4 -> 1
using Net121x150 = NetCfg<128, 160, 4>;    // reference TheEnergyNet(150, 3, 121): 121 -> 150 x4 -> 1

// Layout of the four-lanes-per-pair fused kernel (nepx_quad.h).
template <class Cfg>
struct QuadCfg {
    static constexpr int NC = (Cfg::NANG + 1 + 3) / 4;  // 4-row chunks of n
    static constexpr int LC = (Cfg::NLM + 3) / 4;       // 4-column chunks of lm
    static constexpr int STEPS = Cfg::ND / 4;           // sites per lane
    static constexpr int CH = NC + LC;                  // exchanged chunks per site step
    // LDS floats per pair for the exchange; TS/4 odd puts the 64 scalar writes of
    // one wave instruction on 64 distinct banks
    static constexpr int TS = CH * 16 + 4;
    // forward and adjoint accumulators of up to 2 x 96 registers per lane
    static constexpr bool OK = Cfg::ND % 4 == 0 && NC * LC <= 24;
    // Shapes whose lane state fits 256 registers run two waves per SIMD (512-thread
    // blocks sharing one copy of the network weights), so one wave's LDS and
    // matrix-core latencies overlap the other's work; larger ones run one.
    static constexpr int WAVES_PER_SIMD = NC * LC <= 12 ? 2 : 1;
    static constexpr int BLOCK = 256 * WAVES_PER_SIMD;
};

// Fused descriptor -> network -> adjoint with four lanes per pair and the site
// sums on matrix cores (nepx_quad.hip), on at most max_blocks blocks.  Returns
// NEPX_EUNSUP for shapes it does not cover.
template <class Cfg, class N>
int launch_fused_quad(const nepx_pair_model& m, const nepx_net* net, int prec, int max_blocks, float scale,
                      const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                      const int32_t* np_dev, int64_t maxp, float clip, float* out, RevOut rv, void* stream);

}  // namespace nepx
