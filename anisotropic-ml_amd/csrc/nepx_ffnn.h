// Energy network on MFMA: forward and backward-to-input for tiles of 16 pairs.
//
// TheEnergyNet (nep_simulator/ForceEvaluator.py:6-26): D -> H -> (H)*NH -> 1 with an
// activation after every layer but the last.  Each wave computes the
// transposed products  Z^T (H x 16) = W (H x K) . X^T (K x 16)  with
// v_mfma_f32_16x16x4_f32, so hidden units run along M and the 16 pairs along N.
// In the 16x16 accumulator a lane holds pair (lane % 16) and units
// 16*mt + 4*(lane/16) + r, r = 0..3.  The K order of every product is
// permuted so that k-step s takes unit 16*(s/4) + 4*(lane/16) + s%4 from each
// lane: the accumulator of one layer is then the B operand of the next with
// no cross-lane movement.  The host packs every A operand in that order
// (nepx_net_pack), one coalesced 256-byte read per MFMA.
#pragma once
#include "nepx_common.h"

namespace nepx {

typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int round16(int v) { return (v + 15) / 16 * 16; }

// Offsets (in floats) of the packed operands; identical on host and device.
struct NetLayout {
    int DP, HP, NL;  // padded input, padded width, layers that produce hidden units
    __host__ __device__ constexpr int fwd_size(int k) const {  // layer k = 0..NL-1
        return (HP / 16) * ((k == 0 ? DP : HP) / 4) * 64;
    }
    __host__ __device__ constexpr int bwd_size(int k) const {
        return ((k == 0 ? DP : HP) / 16) * (HP / 4) * 64;
    }
    __host__ __device__ constexpr int fwd_off(int k) const {
        return k == 0 ? 0 : fwd_off(k - 1) + fwd_size(k - 1);
    }
    __host__ __device__ constexpr int bwd_off(int k) const {
        return k == 0 ? fwd_off(NL) : bwd_off(k - 1) + bwd_size(k - 1);
    }
    __host__ __device__ constexpr int bias_off(int k) const { return bwd_off(NL) + k * HP; }
    __host__ __device__ constexpr int wout_off() const { return bias_off(NL); }
    __host__ __device__ constexpr int bout_off() const { return wout_off() + HP; }
    __host__ __device__ constexpr int total() const { return bout_off() + 4; }
};

__host__ __device__ constexpr int round32(int v) { return (v + 31) / 32 * 32; }

// Offsets (in 16-bit elements) of the operands of the 16x16x32 paths (fp16 or
// bf16).  Each (M tile, 32-deep k-step) holds 64 lanes x 8 elements; with two
// parts (hi/lo split) the lo part of a matrix follows its hi part.  Input rows
// are padded to a multiple of 32.
struct NetLayoutH {
    int DPH, DP, HP, NL, PARTS;
    __host__ __device__ constexpr int fwd_size(int k) const {
        return PARTS * (HP / 16) * ((k == 0 ? DPH : HP) / 32) * 512;
    }
    __host__ __device__ constexpr int bwd_size(int k) const {
        return PARTS * ((k == 0 ? DP : HP) / 16) * (HP / 32) * 512;
    }
    __host__ __device__ constexpr int fwd_off(int k) const {
        return k == 0 ? 0 : fwd_off(k - 1) + fwd_size(k - 1);
    }
    __host__ __device__ constexpr int bwd_off(int k) const {
        return k == 0 ? fwd_off(NL) : bwd_off(k - 1) + bwd_size(k - 1);
    }
    __host__ __device__ constexpr int total() const { return bwd_off(NL); }
};

template <int DP_, int HP_, int NL_>
struct NetCfg {
    static constexpr int DP = DP_, HP = HP_, NL = NL_;
    static constexpr int MT = HP / 16;   // M tiles of a hidden layer
    static constexpr int KD = DP / 4;    // k-steps over the input
    static constexpr int KH = HP / 4;    // k-steps over a hidden layer
    static constexpr int MD = DP / 16;   // M tiles of dE/dG
    static constexpr int DPH = round32(DP);
    static constexpr int MDH = DPH / 16; // input chunks of the fp16 path
    static_assert(MD <= MT, "dE/dG tiles reuse the hidden-layer accumulators");
    static constexpr NetLayout layout() { return NetLayout{DP, HP, NL}; }
    static constexpr NetLayoutH layout_h(int parts) { return NetLayoutH{DPH, DP, HP, NL, parts}; }
};

// tanh from its definition, branch-free: 1 - 2/(1 + e^{2z}) (hardware exp2 and
// reciprocal; e^{2z} overflowing to inf gives 1, underflowing to 0 gives -1).
// The absolute error stays below 2e-7 for all z (relative error grows as
// |z| -> 0, where the subtraction cancels); the network consumes activations
// only through products with O(1) weights, so the absolute error is the one
// that matters.
__device__ __forceinline__ float tanh_act(float z) {
    const float t = __builtin_amdgcn_exp2f(2.8853900817779268f * z);  // 2 log2(e) z
    return __builtin_fmaf(-2.f, __builtin_amdgcn_rcpf(1.f + t), 1.f);
}

__device__ __forceinline__ float act_fwd(float z, int act) {
    return act == NEPX_ACT_TANH ? tanh_act(z) : fmaxf(z, 0.f);
}
__device__ __forceinline__ float act_der(float h, int act) {  // in terms of the output
    return act == NEPX_ACT_TANH ? 1.f - h * h : (h > 0.f ? 1.f : 0.f);
}

// NT tiles of 16 pairs each, sharing every A-operand read.  gin[t]: lane's
// 4-float chunks of its pair's input in tile t (chunk c holds inputs
// 16c + 4*(lane/16) + 0..3).  Outputs: y[t] (network output, valid in every
// lane for pair lane%16 of tile t) and dgo[t] (dE/dG chunks in the same
// layout), both scaled by `scale`.  W may point to global memory or to a copy
// in LDS.
template <class N, int NT>
__device__ __forceinline__ void ffnn_tiles(const float* __restrict__ W, int act, const v4f (&gin)[NT][N::MD],
                                           float scale, v4f (&dgo)[NT][N::MD], float (&y)[NT]) {
    constexpr NetLayout Lay = N::layout();
    // An opaque lane index per call keeps the (loop-invariant) weight reads
    // inside the caller's loop; hoisted, they would occupy hundreds of registers.
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int g = lane >> 4;
    v4f h[N::NL][NT][N::MT];  // activations of every hidden layer (D layout)
    // ---- forward ----
#pragma unroll
    for (int k = 0; k < N::NL; ++k) {
        const float* Wf = W + Lay.fwd_off(k);
        const float* bk = W + Lay.bias_off(k);
        const int ks = (k == 0) ? N::KD : N::KH;
        v4f acc[NT][N::MT];
#pragma unroll
        for (int mt = 0; mt < N::MT; ++mt) {
            const float4 b4 = *reinterpret_cast<const float4*>(bk + 16 * mt + 4 * g);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][mt] = v4f{b4.x, b4.y, b4.z, b4.w};
        }
#pragma unroll
        for (int s = 0; s < ks; ++s) {
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) {
                const float aop = Wf[(mt * ks + s) * 64 + lane];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const float bop = (k == 0) ? gin[t][s >> 2][s & 3] : h[k - 1][t][s >> 2][s & 3];
                    acc[t][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, bop, acc[t][mt], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) h[k][t][mt][r] = act_fwd(acc[t][mt][r], act);
    }
    // ---- output layer (dot over units, reduced across the 4 lane groups) ----
    const float* wo = W + Lay.wout_off();
    v4f delta[NT][N::MT];
#pragma unroll
    for (int t = 0; t < NT; ++t) y[t] = 0.f;
#pragma unroll
    for (int mt = 0; mt < N::MT; ++mt) {
        const float4 w4 = *reinterpret_cast<const float4*>(wo + 16 * mt + 4 * g);
        const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                y[t] += wv[r] * h[N::NL - 1][t][mt][r];
                delta[t][mt][r] = scale * wv[r] * act_der(h[N::NL - 1][t][mt][r], act);
            }
    }
    const float bo = W[Lay.bout_off()];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        y[t] += __shfl_xor(y[t], 16);
        y[t] += __shfl_xor(y[t], 32);
        y[t] += bo;
    }
    // ---- backward through the hidden layers ----
#pragma unroll
    for (int k = N::NL - 1; k >= 1; --k) {
        const float* Wb = W + Lay.bwd_off(k);
        v4f acc[NT][N::MT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) acc[t][mt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < N::KH; ++s) {
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) {
                const float aop = Wb[(mt * N::KH + s) * 64 + lane];
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[t][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, delta[t][s >> 2][s & 3], acc[t][mt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) delta[t][mt][r] = acc[t][mt][r] * act_der(h[k - 1][t][mt][r], act);
    }
    // ---- dE/dG = W_1^T delta_1 ----
    {
        const float* Wb = W + Lay.bwd_off(0);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MD; ++mt) dgo[t][mt] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < N::KH; ++s) {
#pragma unroll
            for (int mt = 0; mt < N::MD; ++mt) {
                const float aop = Wb[(mt * N::KH + s) * 64 + lane];
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    dgo[t][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop, delta[t][s >> 2][s & 3], dgo[t][mt], 0, 0, 0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 16-bit operand networks on the 16x16x32 matrix-core instructions, FP32
// accumulation.  K-step s of 32 takes, in lane (p, g), units
// 16*(2s + j/4) + 4g + j%4 for j = 0..7: the two 16-unit accumulator tiles 2s,
// 2s+1 of the previous layer, so there is still no cross-lane movement.
//  * NEPX_NET_F16X3: every product x*w is x_hi w_hi + x_hi w_lo + x_lo w_hi with
//    fp16 parts (x = x_hi + x_lo to ~22 bits).  The backward deltas are carried
//    without the energy range (so their size does not grow with en_max - en_min)
//    and scaled by 2^10 before the split (exact) to stay clear of fp16
//    underflow; the range and 2^-10 are applied once to dE/dG.  Valid while
//    |G|, |h| < 65504 and the range-free deltas stay below 64 (tanh hidden
//    layers are bounded by 1).
//  * NEPX_NET_BF16X3: the same three products with bf16 parts (~16 bits); bf16
//    has the FP32 exponent range, so no scaling and no range limit.
//  * NEPX_NET_BF16: one bf16 product per k-step (8-bit mantissas, round to
//    nearest even on both operands), the fastest and least accurate form.
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int PREC>
struct Lowp;
template <>
struct Lowp<NEPX_NET_F16X3> {
    using T = _Float16;
    using V = v8h;
    static constexpr int PARTS = 2;
    static constexpr float UP = 1024.f;
    __device__ __forceinline__ static v4f mma(const V& a, const V& b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Lowp<NEPX_NET_BF16X3> {
    using T = __bf16;
    using V = v8bf;
    static constexpr int PARTS = 2;
    static constexpr float UP = 1.f;
    __device__ __forceinline__ static v4f mma(const V& a, const V& b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Lowp<NEPX_NET_BF16> {
    using T = __bf16;
    using V = v8bf;
    static constexpr int PARTS = 1;
    static constexpr float UP = 1.f;
    __device__ __forceinline__ static v4f mma(const V& a, const V& b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// Eight consecutive K inputs (two 4-chunks) as the B operand: hi (and lo) parts.
template <int PREC>
__device__ __forceinline__ void split8(const v4f& a, const v4f& b, float sc, typename Lowp<PREC>::V& hi,
                                       typename Lowp<PREC>::V& lo) {
    using T = typename Lowp<PREC>::T;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float x = (j < 4 ? a[j] : b[j - 4]) * sc;
        const T h = (T)x;
        hi[j] = h;
        if constexpr (Lowp<PREC>::PARTS == 2) lo[j] = (T)(x - (float)h);
    }
}

template <int PREC>
__device__ __forceinline__ v4f mma_split(const typename Lowp<PREC>::V* __restrict__ A, int idx, int lo_off,
                                         const typename Lowp<PREC>::V& bh, const typename Lowp<PREC>::V& bl,
                                         v4f acc) {
    using L = Lowp<PREC>;
    const typename L::V ah = A[idx];
    if constexpr (L::PARTS == 2) {
        const typename L::V al = A[idx + lo_off];
        acc = L::mma(al, bh, acc);
        acc = L::mma(ah, bl, acc);
    }
    return L::mma(ah, bh, acc);
}

// Same contract as ffnn_tiles; gin has MDH chunks (zero past D).  W32 supplies
// biases and the output layer, WH the 16-bit operands (global or LDS).
template <class N, int NT, int PREC>
__device__ __forceinline__ void ffnn_tiles_h(const float* __restrict__ W32, const void* __restrict__ WH, int act,
                                             const v4f (&gin)[NT][N::MDH], float scale, v4f (&dgo)[NT][N::MD],
                                             float (&y)[NT]) {
    using L = Lowp<PREC>;
    using V = typename L::V;
    constexpr NetLayout Lay = N::layout();
    constexpr NetLayoutH LH = N::layout_h(L::PARTS);
    constexpr float kUp = L::UP, kDown = 1.f / L::UP;
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int g = lane >> 4;
    const V* A8 = reinterpret_cast<const V*>(WH);
    v4f h[N::NL][NT][N::MT];
#pragma unroll
    for (int k = 0; k < N::NL; ++k) {
        const int base = LH.fwd_off(k) / 8;
        const int ks = (k == 0 ? N::DPH : N::HP) / 32;
        const int lo_off = (N::MT * ks * 512) / 8;
        const float* bk = W32 + Lay.bias_off(k);
        v4f acc[NT][N::MT];
#pragma unroll
        for (int mt = 0; mt < N::MT; ++mt) {
            const float4 b4 = *reinterpret_cast<const float4*>(bk + 16 * mt + 4 * g);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][mt] = v4f{b4.x, b4.y, b4.z, b4.w};
        }
        // every k-step's operand split first, then the layer's products as one block of matrix
        // instructions (the split VALU work of one layer does not interleave with its products;
        // 8 % faster for the 128-wide cylinder network, neutral for 64-wide ones)
        constexpr int KSMAX = (N::DPH > N::HP ? N::DPH : N::HP) / 32;
        V bh[KSMAX][NT], bl[KSMAX][NT];
#pragma unroll
        for (int s = 0; s < ks; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                if (k == 0) split8<PREC>(gin[t][2 * s], gin[t][2 * s + 1], 1.f, bh[s][t], bl[s][t]);
                else split8<PREC>(h[k - 1][t][2 * s], h[k - 1][t][2 * s + 1], 1.f, bh[s][t], bl[s][t]);
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < ks; ++s) {
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) {
                const int idx = base + (mt * ks + s) * 64 + lane;
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][mt] = mma_split<PREC>(A8, idx, lo_off, bh[s][t], bl[s][t], acc[t][mt]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) h[k][t][mt][r] = act_fwd(acc[t][mt][r], act);
    }
    const float* wo = W32 + Lay.wout_off();
    v4f delta[NT][N::MT];
#pragma unroll
    for (int t = 0; t < NT; ++t) y[t] = 0.f;
#pragma unroll
    for (int mt = 0; mt < N::MT; ++mt) {
        const float4 w4 = *reinterpret_cast<const float4*>(wo + 16 * mt + 4 * g);
        const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                y[t] += wv[r] * h[N::NL - 1][t][mt][r];
                delta[t][mt][r] = wv[r] * act_der(h[N::NL - 1][t][mt][r], act);
            }
    }
    const float bo = W32[Lay.bout_off()];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        y[t] += __shfl_xor(y[t], 16);
        y[t] += __shfl_xor(y[t], 32);
        y[t] += bo;
    }
#pragma unroll
    for (int k = N::NL - 1; k >= 0; --k) {
        const int base = LH.bwd_off(k) / 8;
        constexpr int ks = N::HP / 32;
        const int mtn = (k == 0 ? N::MD : N::MT);
        const int lo_off = (mtn * ks * 512) / 8;
        v4f acc[NT][N::MT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) acc[t][mt] = v4f{0.f, 0.f, 0.f, 0.f};
        V bh[ks][NT], bl[ks][NT];
#pragma unroll
        for (int s = 0; s < ks; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) split8<PREC>(delta[t][2 * s], delta[t][2 * s + 1], kUp, bh[s][t], bl[s][t]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < ks; ++s) {
#pragma unroll
            for (int mt = 0; mt < N::MT; ++mt) {
                if (mt >= mtn) break;
                const int idx = base + (mt * ks + s) * 64 + lane;
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[t][mt] = mma_split<PREC>(A8, idx, lo_off, bh[s][t], bl[s][t], acc[t][mt]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (k > 0) {
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int mt = 0; mt < N::MT; ++mt)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        delta[t][mt][r] = acc[t][mt][r] * kDown * act_der(h[k - 1][t][mt][r], act);
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int mt = 0; mt < N::MD; ++mt) dgo[t][mt] = acc[t][mt] * (kDown * scale);
        }
    }
}

// Copy `floats` floats (a multiple of 4) into LDS (all threads of the block cooperate).
__device__ __forceinline__ void stage_copy(float* __restrict__ lds, const float* __restrict__ src, int floats) {
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(lds);
    for (int k = threadIdx.x; k < floats / 4; k += blockDim.x) d4[k] = s4[k];
}

template <class N>
__device__ __forceinline__ void stage_weights(float* __restrict__ lds, const float* __restrict__ W) {
    stage_copy(lds, W, N::layout().total());
    __syncthreads();
}

}  // namespace nepx
