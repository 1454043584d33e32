// Pair kernels: descriptor forward, MFMA energy network, adjoint, and the
// fused step kernel.  See nepx_desc.h and nepx_ffnn.h for the math and layouts.
#include "nepx_pairk.h"

#include <string.h>

namespace nepx {

// Descriptor row of one pair (or cloud), one pass over the sites per l range.
template <class Cfg, class Sites, class Row>
__device__ __forceinline__ void desc_forward_sites(const Sites& sites, Row&& G) {
    float grad[Cfg::NRAD + 1];
    for_passes<Cfg>([&](auto range, auto first) {
        using LR = decltype(range);
        float A[Cfg::NANG + 1][LR::N];
        accumulate_sites<Cfg, LR, decltype(first)::value>(sites, grad, A);
        desc_finish<Cfg, LR>(A, G);
        if constexpr (decltype(first)::value) {
#pragma unroll
            for (int n = 0; n <= Cfg::NRAD; ++n) G[n] = grad[n];
        }
    });
}

template <class Cfg, class Row>
__device__ __forceinline__ void desc_forward(const nepx_pair_model& m, const PairIn& g, Row&& G) {
    desc_forward_sites<Cfg>(PairSites<Cfg>{m, g, 1.f / m.r_nep}, G);
}

// Adjoint over every l range, recomputing each range's forward accumulators.
template <class Cfg>
__device__ __forceinline__ void adjoint_all(const nepx_pair_model& m, const PairIn& g, const float (&dG)[Cfg::D],
                                            AdjAcc& acc) {
    for_passes<Cfg>([&](auto range, auto first) {
        using LR = decltype(range);
        float grad[Cfg::NRAD + 1];
        float A[Cfg::NANG + 1][LR::N];
        desc_accumulate<Cfg, LR, false>(m, g, grad, A);
        desc_adjoint<Cfg, LR, decltype(first)::value>(m, g, dG, A, acc);
    });
}

// ---------------------------------------------------------------------------
template <class Cfg>
__global__ __launch_bounds__(kBlock) void k_desc_fwd(nepx_pair_model m, const float4* __restrict__ qf,
                                                     const int32_t* __restrict__ pi,
                                                     const int32_t* __restrict__ pj,
                                                     const float4* __restrict__ pt,
                                                     const int32_t* __restrict__ np_dev, int64_t maxp,
                                                     float* __restrict__ G) {
    const int P = n_pairs(np_dev, maxp);
    for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < P; idx += gridDim.x * kBlock) {
        float4 qi, qj, t;
        load_pair(qf, pi, pj, pt, idx, qi, qj, t);
        const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
        if constexpr (Cfg::ONE_PASS) {
            float Gr[Cfg::D];
            desc_forward<Cfg>(m, g, Gr);
            store_row<Cfg>(G + (size_t)idx * Cfg::D, Gr);
        } else {  // several passes: each writes its entries (no D-float row held in registers)
            desc_forward<Cfg>(m, g, G + (size_t)idx * Cfg::D);
        }
    }
}

template <class Cfg>
__global__ __launch_bounds__(kBlock) NEPX_ADJ_ATTR void k_adjoint(nepx_pair_model m, const float4* __restrict__ qf,
                                                    const int32_t* __restrict__ pi,
                                                    const int32_t* __restrict__ pj,
                                                    const float4* __restrict__ pt,
                                                    const int32_t* __restrict__ np_dev, int64_t maxp,
                                                    const float* __restrict__ dEdG,
                                                    const float* __restrict__ E, float clip,
                                                    float* __restrict__ out) {
    const int P = n_pairs(np_dev, maxp);
    for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < P; idx += gridDim.x * kBlock) {
        float4 qi, qj, t;
        load_pair(qf, pi, pj, pt, idx, qi, qj, t);
        const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
        float dG[Cfg::D];
        load_row<Cfg>(dEdG + (size_t)idx * Cfg::D, dG);
        AdjAcc acc;
        adjoint_all<Cfg>(m, g, dG, acc);
        store_out(out, idx, adj_finish<Cfg>(acc), E ? E[idx] : 0.f, clip);
    }
}

// Multi-pass shapes with few pairs: the l-range passes of a pair run on different waves of a block
// (four waves, one per SIMD, as the register-heavy passes need; 64 pairs per block).  Wave w takes
// passes w and NP - 1 - w (the cheap low-l passes with the dear high-l ones).  The forward writes
// each pass's descriptor entries as k_desc_fwd does; the adjoint leaves each pass's per-body sums
// in LDS and sums them in pass order, so both give k_desc_fwd's and k_adjoint's results bit for
// bit (every pass's sums start from zero, as k_adjoint's running sums do not become -0).
template <class Cfg>
constexpr int pass_waves() {
    constexpr int np = n_passes(Cfg::NANG, Cfg::LMAX);
    return np < 4 ? np : 4;
}
template <class Cfg>
__device__ __forceinline__ int pass_owner(int p) {
    constexpr int np = n_passes(Cfg::NANG, Cfg::LMAX), w = pass_waves<Cfg>();
    return p < w ? p : (np - 1 - p < w ? np - 1 - p : p % w);
}

template <class Cfg>
__global__ __launch_bounds__(64 * pass_waves<Cfg>()) NEPX_ADJ_ATTR void k_desc_fwd_passes(
    nepx_pair_model m, const float4* __restrict__ qf, const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
    const float4* __restrict__ pt, const int32_t* __restrict__ np_dev, int64_t maxp, float* __restrict__ G) {
    const int P = n_pairs(np_dev, maxp);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int idx = blockIdx.x * 64 + (threadIdx.x & 63); idx < P; idx += gridDim.x * 64) {
        float4 qi, qj, t;
        load_pair(qf, pi, pj, pt, idx, qi, qj, t);
        const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
        const PairSites<Cfg> sites{m, g, 1.f / m.r_nep};
        float* row = G + (size_t)idx * Cfg::D;
        int k = 0;
        for_passes<Cfg>([&](auto range, auto first) {
            if (pass_owner<Cfg>(k++) == wave) {
                using LR = decltype(range);
                float grad[Cfg::NRAD + 1];
                float A[Cfg::NANG + 1][LR::N];
                accumulate_sites<Cfg, LR, decltype(first)::value>(sites, grad, A);
                desc_finish<Cfg, LR>(A, row);
                if constexpr (decltype(first)::value) {
#pragma unroll
                    for (int n = 0; n <= Cfg::NRAD; ++n) row[n] = grad[n];
                }
            }
        });
    }
}

template <class Cfg>
__global__ __launch_bounds__(64 * pass_waves<Cfg>()) NEPX_ADJ_ATTR void k_adjoint_passes(
    nepx_pair_model m, const float4* __restrict__ qf, const int32_t* __restrict__ pi, const int32_t* __restrict__ pj,
    const float4* __restrict__ pt, const int32_t* __restrict__ np_dev, int64_t maxp,
    const float* __restrict__ dEdG, const float* __restrict__ E, float clip, float* __restrict__ out) {
    constexpr int NP = n_passes(Cfg::NANG, Cfg::LMAX);
    __shared__ float part[NP][12][64];  // per pass: G0, T0, G1, T1 (xyz) of the block's 64 pairs
    const int P = n_pairs(np_dev, maxp);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int base = blockIdx.x * 64; base < P; base += gridDim.x * 64) {  // uniform over the block
        const int idx = base + lane;
        const bool valid = idx < P;
        AdjAcc arms;  // pass 0 sets the pair's arm sums
        if (valid) {
            float4 qi, qj, t;
            load_pair(qf, pi, pj, pt, idx, qi, qj, t);
            const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
            const float(&dG)[Cfg::D] = *reinterpret_cast<const float(*)[Cfg::D]>(dEdG + (size_t)idx * Cfg::D);
            int k = 0;
            for_passes<Cfg>([&](auto range, auto first) {
                const int p = k++;
                if (pass_owner<Cfg>(p) == wave) {
                    using LR = decltype(range);
                    float grad[Cfg::NRAD + 1];
                    float A[Cfg::NANG + 1][LR::N];
                    AdjAcc a;
                    desc_accumulate<Cfg, LR, false>(m, g, grad, A);
                    desc_adjoint<Cfg, LR, decltype(first)::value>(m, g, dG, A, a);
                    const float v[12] = {a.G0x, a.G0y, a.G0z, a.T0x, a.T0y, a.T0z,
                                         a.G1x, a.G1y, a.G1z, a.T1x, a.T1y, a.T1z};
#pragma unroll
                    for (int c = 0; c < 12; ++c) part[p][c][lane] = v[c];
                    if constexpr (decltype(first)::value) arms = a;
                }
            });
        }
        __syncthreads();
        if (wave == 0 && valid) {  // pass 0 runs on wave 0, so `arms` is set here
            AdjAcc acc;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                acc.G0x += part[p][0][lane]; acc.G0y += part[p][1][lane]; acc.G0z += part[p][2][lane];
                acc.T0x += part[p][3][lane]; acc.T0y += part[p][4][lane]; acc.T0z += part[p][5][lane];
                acc.G1x += part[p][6][lane]; acc.G1y += part[p][7][lane]; acc.G1z += part[p][8][lane];
                acc.T1x += part[p][9][lane]; acc.T1y += part[p][10][lane]; acc.T1z += part[p][11][lane];
            }
            acc.A0x = arms.A0x; acc.A0y = arms.A0y; acc.A0z = arms.A0z;
            acc.A1x = arms.A1x; acc.A1y = arms.A1y; acc.A1z = arms.A1z;
            store_out(out, idx, adj_finish<Cfg>(acc), E ? E[idx] : 0.f, clip);
        }
        __syncthreads();  // the sums are read before the next pairs' passes write them
    }
}

template <class N, int NT, int PREC, int ACT>
__global__ __launch_bounds__(kBlock) void k_ffnn(const float* __restrict__ Wg, const void* __restrict__ Whg,
                                                 float en_min, float scale, int D, const float* __restrict__ G,
                                                 const int32_t* __restrict__ np_dev, int64_t maxp,
                                                 float* __restrict__ E, float* __restrict__ dEdG) {
    using R = NetRun<N, NT, PREC>;
    __shared__ __attribute__((aligned(16))) float wl[weight_lds_floats<N, PREC>()];
    const float* W;
    const void* WH;
    stage_net<N, PREC>(wl, Wg, Whg, W, WH);
    const int P = n_pairs(np_dev, maxp);
    const int lane = threadIdx.x & 63;
    const int nt = (P + 16 * NT - 1) / (16 * NT);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kBlock) >> 6;
    for (int tile = wave; tile < nt; tile += nwaves) {
        v4f gin[NT][R::IC];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int p = (tile * NT + t) * 16 + (lane & 15);
            gather_chunks<R::IC>(G + (size_t)p * D, p < P, D, gin[t]);
        }
        v4f dg[NT][N::MD];
        float y[NT];
        R::run(W, WH, ACT, gin, scale, dg, y);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int p = (tile * NT + t) * 16 + (lane & 15);
            if (p < P && lane < 16) E[p] = en_min + scale * y[t];
            scatter_chunks<N::MD>(dEdG + (size_t)p * D, p < P, D, dg[t]);
        }
    }
}

// FP32 network for few pairs: one 16-pair tile per block, and the block's waves take the 16-unit
// blocks of every layer between them (wave w computes blocks MPW w .. MPW w + MPW - 1 of every
// hidden layer and of every backward delta, over all k-steps), so that a tile's chain of layers is
// spread over MT / MPW waves instead of running on one.  The layer outputs go through LDS to the
// next layer's B operands.  Every accumulator takes the same products in the same order as
// ffnn_tiles, and the output sum is formed the same way, so E and dE/dG equal k_ffnn's bit for
// bit.  One block per wave (MPW = 1): the 640-thread block of a 160-wide network needs 131 VGPRs
// per lane, within the 168 that three waves per SIMD leave; each layer's weight reads are kept in
// that layer, or the compiler hoists them all and spills.
template <class N>
constexpr int coop_mpw() { return 1; }

template <class N, int ACT>
__global__ __launch_bounds__(N::MT / coop_mpw<N>() * 64) void k_ffnn_coop(
    const float* __restrict__ W, float en_min, float scale, int D, const float* __restrict__ G,
    const int32_t* __restrict__ np_dev, int64_t maxp, float* __restrict__ E, float* __restrict__ dEdG) {
    constexpr NetLayout Lay = N::layout();
    constexpr int MT = N::MT, MPW = coop_mpw<N>();
    static_assert(MT % MPW == 0, "unit blocks split evenly over the waves");
    __shared__ v4f xs[2][MT][64];  // one layer's outputs: block c of lane l at [c][l]
    const int P = n_pairs(np_dev, maxp);
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));  // keeps the weight reads inside the tile loop
    const int g = lane >> 4;
    const int m0 = (threadIdx.x >> 6) * MPW;  // this wave's first 16-unit block
    const int nt = (P + 15) / 16;
    for (int tile = blockIdx.x; tile < nt; tile += gridDim.x) {
        const int p = tile * 16 + (lane & 15);
        v4f gin[N::MD];
        gather_chunks<N::MD>(G + (size_t)p * D, p < P, D, gin);
        v4f hown[N::NL][MPW];  // this wave's blocks of every hidden layer
        int buf = 0;
#pragma unroll
        for (int k = 0; k < N::NL; ++k) {
            asm volatile("" : "+v"(lane));  // each layer's weight reads stay in that layer
            const float* Wf = W + Lay.fwd_off(k);
            const int ks = (k == 0) ? N::KD : N::KH;
            v4f acc[MPW];
#pragma unroll
            for (int i = 0; i < MPW; ++i) {
                const float4 b4 = *reinterpret_cast<const float4*>(W + Lay.bias_off(k) + 16 * (m0 + i) + 4 * g);
                acc[i] = v4f{b4.x, b4.y, b4.z, b4.w};
            }
#pragma unroll
            for (int c = 0; c < ks / 4; ++c) {
                const v4f x = (k == 0) ? gin[c < N::MD ? c : 0] : xs[buf ^ 1][c][lane];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int i = 0; i < MPW; ++i)
                        acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(Wf[((m0 + i) * ks + 4 * c + r) * 64 + lane],
                                                                     x[r], acc[i], 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < MPW; ++i) {
#pragma unroll
                for (int r = 0; r < 4; ++r) hown[k][i][r] = act_fwd(acc[i][r], ACT);
                xs[buf][m0 + i][lane] = hown[k][i];
            }
            __syncthreads();
            buf ^= 1;
        }
        // output layer: every wave forms y from the whole last layer, in ffnn_tiles' order
        const float* wo = W + Lay.wout_off();
        float y = 0.f;
#pragma unroll
        for (int c = 0; c < MT; ++c) {
            const float4 w4 = *reinterpret_cast<const float4*>(wo + 16 * c + 4 * g);
            const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
            const v4f h = xs[buf ^ 1][c][lane];
#pragma unroll
            for (int r = 0; r < 4; ++r) y += wv[r] * h[r];
        }
        y += __shfl_xor(y, 16);
        y += __shfl_xor(y, 32);
        y += W[Lay.bout_off()];
        if (m0 == 0 && p < P && lane < 16) E[p] = en_min + scale * y;
        v4f delta[MPW];
#pragma unroll
        for (int i = 0; i < MPW; ++i) {
            const float4 w4 = *reinterpret_cast<const float4*>(wo + 16 * (m0 + i) + 4 * g);
            const float wv[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int r = 0; r < 4; ++r) delta[i][r] = scale * wv[r] * act_der(hown[N::NL - 1][i][r], ACT);
        }
        // backward: layer k's deltas through LDS, then this wave's blocks of layer k-1 (k = 0: dE/dG)
#pragma unroll
        for (int k = N::NL - 1; k >= 0; --k) {
#pragma unroll
            for (int i = 0; i < MPW; ++i) xs[buf][m0 + i][lane] = delta[i];
            __syncthreads();
            buf ^= 1;
            if (k == 0 && m0 >= N::MD) break;
            asm volatile("" : "+v"(lane));
            const float* Wb = W + Lay.bwd_off(k);
            const int mtn = (k == 0) ? N::MD : MT;  // output blocks of this product
            v4f acc[MPW];
#pragma unroll
            for (int i = 0; i < MPW; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < N::KH / 4; ++c) {
                const v4f x = xs[buf ^ 1][c][lane];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int i = 0; i < MPW; ++i)
                        if (m0 + i < mtn)
                            acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                Wb[((m0 + i) * N::KH + 4 * c + r) * 64 + lane], x[r], acc[i], 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < MPW; ++i) {
                if (k > 0) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) delta[i][r] = acc[i][r] * act_der(hown[k - 1][i][r], ACT);
                } else if (m0 + i < N::MD) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int e = 16 * (m0 + i) + 4 * g + r;
                        if (p < P && e < D) dEdG[(size_t)p * D + e] = acc[i][r];
                    }
                }
            }
        }
        __syncthreads();  // the last reads of xs are done before the next tile writes it
    }
}

// Fused: descriptor -> LDS -> MFMA network -> LDS -> adjoint.  Each block
// takes 256 consecutive pairs per iteration (64 per wave), so every wave of a
// block runs the same number of iterations and the barriers are uniform.
// The forward accumulators A of each lane's pair stay in registers through
// the network phase and feed the adjoint directly.
template <class Cfg, class N, int PREC, int ACT>
__global__ __launch_bounds__(kBlock) NEPX_ADJ_ATTR void k_fused(nepx_pair_model m, const float* __restrict__ Wg,
                                                                const void* __restrict__ Whg,
                                                                float en_min, float scale,
                                                                const float4* __restrict__ qf,
                                                                const int32_t* __restrict__ pi,
                                                                const int32_t* __restrict__ pj,
                                                                const float4* __restrict__ pt,
                                                                const int32_t* __restrict__ np_dev, int64_t maxp,
                                                                float clip, float* __restrict__ out, RevOut rv) {
    constexpr int LD = (Cfg::D + 3) / 4 * 4 + 4;  // row stride in LDS (16-byte aligned)
    __shared__ __attribute__((aligned(16))) float tile[kBlock * LD];
    __shared__ float etile[kBlock];
    using R = NetRun<N, kNetTiles, PREC>;
    __shared__ __attribute__((aligned(16))) float wl[weight_lds_floats<N, PREC>()];
    const float* W;
    const void* WH;
    stage_net<N, PREC>(wl, Wg, Whg, W, WH);
    const int P = n_pairs(np_dev, maxp);
    const int tag = rev_tag(rv);
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    float* myrow = tile + threadIdx.x * LD;
    for (int base = blockIdx.x * kBlock; base < P; base += gridDim.x * kBlock) {
        const int idx = base + threadIdx.x;
        const bool valid = idx < P;
        float4 qi = make_float4(1.f, 0.f, 0.f, 0.f), qj = qi, t = make_float4(1.f, 0.f, 0.f, 0.f);
        if (valid) load_pair(qf, pi, pj, pt, idx, qi, qj, t);
        const PairIn g = pair_in<Cfg>(m, qi, qj, t.x, t.y, t.z);
        // one-pass shapes keep the forward accumulators in registers for the adjoint; multi-pass
        // shapes recompute every range there (keeping the last range's through the network phase
        // spills: measured slower on the 262,144-cylinder step, 14.8 vs 12.9 ms)
        using LRL = LastPass<Cfg>;
        constexpr bool KEEP = Cfg::ONE_PASS;
        float A[KEEP ? Cfg::NANG + 1 : 1][KEEP ? LRL::N : 1];
        {
            if constexpr (Cfg::ONE_PASS) {
                float Gr[Cfg::D];
                float grad[Cfg::NRAD + 1];
                desc_accumulate<Cfg, LRL, true>(m, g, grad, A);
                desc_finish<Cfg, LRL>(A, Gr);
#pragma unroll
                for (int n = 0; n <= Cfg::NRAD; ++n) Gr[n] = grad[n];
#pragma unroll
                for (int k = 0; k < Cfg::D; ++k) myrow[k] = valid ? Gr[k] : 0.f;
            } else {  // several passes: each writes its entries straight into the LDS row
                float grad[Cfg::NRAD + 1];
                for_passes<Cfg>([&](auto range, auto first) {
                    using LR = decltype(range);
                    float Ap[Cfg::NANG + 1][LR::N];
                    desc_accumulate<Cfg, LR, decltype(first)::value>(m, g, grad, Ap);
                    desc_finish<Cfg, LR>(Ap, myrow);
                    if constexpr (decltype(first)::value) {
#pragma unroll
                        for (int n = 0; n <= Cfg::NRAD; ++n) myrow[n] = grad[n];
                    }
                });
                if (!valid)
                    for (int k = 0; k < Cfg::D; ++k) myrow[k] = 0.f;
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int sub = 0; sub < 64 / (16 * kNetTiles); ++sub) {
            v4f gin[kNetTiles][R::IC];
#pragma unroll
            for (int tt = 0; tt < kNetTiles; ++tt) {
                const int local = wv * 64 + (sub * kNetTiles + tt) * 16 + (lane & 15);
                gather_chunks<R::IC>(tile + local * LD, true, Cfg::D, gin[tt]);
            }
            v4f dg[kNetTiles][N::MD];
            float y[kNetTiles];
            R::run(W, WH, ACT, gin, scale, dg, y);
#pragma unroll
            for (int tt = 0; tt < kNetTiles; ++tt) {
                const int local = wv * 64 + (sub * kNetTiles + tt) * 16 + (lane & 15);
                if (lane < 16) etile[local] = en_min + scale * y[tt];
                scatter_chunks<N::MD>(tile + local * LD, true, Cfg::D, dg[tt]);
            }
        }
        __syncthreads();
        if (valid) {
            float dG[Cfg::D];
#pragma unroll
            for (int k = 0; k < Cfg::D; ++k) dG[k] = myrow[k];
            AdjAcc acc;
            for_passes<Cfg>([&](auto range, auto first) {
                using LR = decltype(range);
                if constexpr (KEEP) {
                    desc_adjoint<Cfg, LR, decltype(first)::value>(m, g, dG, A, acc);
                } else {
                    float grad[Cfg::NRAD + 1];
                    float Ap[Cfg::NANG + 1][LR::N];
                    desc_accumulate<Cfg, LR, false>(m, g, grad, Ap);
                    desc_adjoint<Cfg, LR, decltype(first)::value>(m, g, dG, Ap, acc);
                }
            });
            store_out(out, idx, adj_finish<Cfg>(acc), etile[threadIdx.x], clip, rv, tag);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Reference-compat finite-difference stencil (PosToPtsKernels.py:153-446).
// Translating i's sites by +dx e equals translating j by -dx e (the
// descriptor depends only on the cloud up to translation); rotating i's sites
// about X_i by dteta about e equals q_i -> dq q_i with dq = (cos dteta/2,
// sin dteta/2 e); the rows in the frame of j rotate q_j the same way.
__device__ __forceinline__ float4 qmul4(float4 a, float4 b) {  // (w, x, y, z) in (x, y, z, w) slots
    return make_float4(a.x * b.x - (a.y * b.y + a.z * b.z + a.w * b.w),
                       a.x * b.y + b.x * a.y + (a.z * b.w - a.w * b.z),
                       a.x * b.z + b.x * a.z + (a.w * b.y - a.y * b.w),
                       a.x * b.w + b.x * a.w + (a.y * b.z - a.z * b.y));
}

__global__ __launch_bounds__(kBlock) void k_stencil_rows(const float4* __restrict__ qf, const int32_t* __restrict__ pi,
                                                         const int32_t* __restrict__ pj,
                                                         const float4* __restrict__ pt, int P, float dx,
                                                         float dteta, float4* __restrict__ qrows,
                                                         int32_t* __restrict__ ri, int32_t* __restrict__ rj,
                                                         float4* __restrict__ trows) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= 10 * P) return;
    const int o = r / P, p = r % P;
    float4 qi = qf[pi[p]], qj = qf[pj[p]], t = pt[p];
    const float c = cosf(0.5f * dteta), s = sinf(0.5f * dteta);
    if (o >= 1 && o <= 3) {
        if (o == 1) t.x -= dx;
        if (o == 2) t.y -= dx;
        if (o == 3) t.z -= dx;
    } else if (o >= 4) {
        const int ax = (o - 4) % 3;
        const float4 dq = make_float4(c, ax == 0 ? s : 0.f, ax == 1 ? s : 0.f, ax == 2 ? s : 0.f);
        if (o <= 6) qi = qmul4(dq, qi);
        else qj = qmul4(dq, qj);
    }
    qrows[2 * (size_t)r] = qi;
    qrows[2 * (size_t)r + 1] = qj;
    ri[r] = 2 * r;
    rj[r] = 2 * r + 1;
    trows[r] = t;
}

__global__ __launch_bounds__(kBlock) void k_fd_forces(const float* __restrict__ E, int P, float dx, float dteta,
                                                      float clip, float* __restrict__ out) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const float e0 = E[p];
    PairOut o;
    o.fx = -(E[1 * (size_t)P + p] - e0) / dx;
    o.fy = -(E[2 * (size_t)P + p] - e0) / dx;
    o.fz = -(E[3 * (size_t)P + p] - e0) / dx;
    o.tix = -(E[4 * (size_t)P + p] - e0) / dteta;
    o.tiy = -(E[5 * (size_t)P + p] - e0) / dteta;
    o.tiz = -(E[6 * (size_t)P + p] - e0) / dteta;
    o.tjx = -(E[7 * (size_t)P + p] - e0) / dteta;
    o.tjy = -(E[8 * (size_t)P + p] - e0) / dteta;
    o.tjz = -(E[9 * (size_t)P + p] - e0) / dteta;
    store_out(out, p, o, e0, clip);
}

// ---------------------------------------------------------------------------
// Dispatch tables.

static int grid_for_cap(int64_t units, int per_block, int64_t cap) {
    int64_t g = (units + per_block - 1) / per_block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

static int grid_for(int64_t maxp, int per_block) {
    return grid_for_cap(maxp, per_block, (int64_t)device_cus() * tuning().pair_blocks_per_cu);
}

#define NEPX_DESC_CASES(X)                          \
    X(8, 11, 6, 4, CfgCube8)                        \
    X(10, 11, 6, 4, CfgTet10)                       \
    X(12, 17, 6, 6, CfgCyl12)                       \
    X(6, 10, 8, 4, CfgCubeV2)                       \
    X(4, 10, 8, 4, CfgThV1)                         \
    X(6, 10, 10, 10, CfgCubeV2L10)                  \
    X(4, 10, 10, 10, CfgThV1L10)

// Multi-pass shapes with at most two blocks of 64 pairs per CU run their passes on separate waves
// (k_desc_fwd_passes, k_adjoint_passes; same results): one lane per pair would leave most SIMDs idle.
template <class Cfg>
static bool by_passes(int64_t maxp) {
    return !Cfg::ONE_PASS && (maxp + 63) / 64 <= 2 * (int64_t)device_cus();
}
static int pass_grid(int64_t maxp) { return (int)((maxp + 63) / 64); }

template <class T>
static void launch_desc_fwd(const nepx_pair_model& m, const float* qf, const int32_t* pair_i, const int32_t* pair_j,
                            const float* pair_t, const int32_t* np_dev, int64_t maxp, float* G, void* stream) {
    if constexpr (!T::ONE_PASS) {
        if (by_passes<T>(maxp)) {
            hipLaunchKernelGGL(k_desc_fwd_passes<T>, dim3(pass_grid(maxp)), dim3(64 * pass_waves<T>()), 0,
                               as_stream(stream), m, (const float4*)qf, pair_i, pair_j, (const float4*)pair_t,
                               np_dev, maxp, G);
            return;
        }
    }
    hipLaunchKernelGGL(k_desc_fwd<T>, dim3(grid_for(maxp, kBlock)), dim3(kBlock), 0, as_stream(stream), m,
                       (const float4*)qf, pair_i, pair_j, (const float4*)pair_t, np_dev, maxp, G);
}

template <class T>
static void launch_adjoint(const nepx_pair_model& m, const float* qf, const int32_t* pair_i, const int32_t* pair_j,
                           const float* pair_t, const int32_t* np_dev, int64_t maxp, const float* dEdG,
                           const float* E, float clip, float* out, void* stream) {
    if constexpr (!T::ONE_PASS) {
        if (by_passes<T>(maxp)) {
            hipLaunchKernelGGL(k_adjoint_passes<T>, dim3(pass_grid(maxp)), dim3(64 * pass_waves<T>()), 0,
                               as_stream(stream), m, (const float4*)qf, pair_i, pair_j, (const float4*)pair_t,
                               np_dev, maxp, dEdG, E, clip, out);
            return;
        }
    }
    hipLaunchKernelGGL(k_adjoint<T>, dim3(grid_for(maxp, kBlock)), dim3(kBlock), 0, as_stream(stream), m,
                       (const float4*)qf, pair_i, pair_j, (const float4*)pair_t, np_dev, maxp, dEdG, E, clip, out);
}

static bool desc_supported(const nepx_pair_model* m) {
#define X(ns, a, b, c, T) \
    if (m->n_sites == ns && m->nrad == a && m->nang == b && m->lmax == c) return true;
    NEPX_DESC_CASES(X)
#undef X
    return false;
}

static int check_model(const nepx_pair_model* m) {
    NEPX_REQUIRE(m != nullptr, NEPX_EINVAL, "null pair model");
    NEPX_REQUIRE(m->n_sites > 0 && m->n_sites <= NEPX_MAX_SITES, NEPX_EINVAL, "n_sites %d out of range",
                 m->n_sites);
    NEPX_REQUIRE(m->nang <= m->nrad, NEPX_EINVAL, "nang (%d) must not exceed nrad (%d)", m->nang, m->nrad);
    NEPX_REQUIRE(m->r_nep > 0.f, NEPX_EINVAL, "r_nep must be positive");
    NEPX_REQUIRE(desc_supported(m), NEPX_EUNSUP,
                 "no descriptor kernel for n_sites=%d hypers=(%d,%d,%d)", m->n_sites, m->nrad, m->nang,
                 m->lmax);
    return NEPX_OK;
}

static int check_net(const nepx_net* n, int D) {
    NEPX_REQUIRE(n != nullptr && n->w_packed != nullptr, NEPX_EINVAL, "null network");
    NEPX_REQUIRE(n->d_in == D, NEPX_EINVAL, "network input %d != descriptor size %d", n->d_in, D);
    NEPX_REQUIRE(n->act == NEPX_ACT_TANH || n->act == NEPX_ACT_RELU, NEPX_EINVAL, "bad activation %d",
                 n->act);
    return NEPX_OK;
}

static int model_D(const nepx_pair_model* m) { return (m->nrad + 1) + (m->nang + 1) * m->lmax; }

}  // namespace nepx

using namespace nepx;

extern "C" int64_t nepx_net_packed_floats(int32_t d_in, int32_t width, int32_t n_hidden) {
    if (d_in <= 0 || width <= 0 || n_hidden < 0) return -1;
    NetLayout L{round16(d_in), round16(width), n_hidden + 1};
    return L.total();
}

extern "C" int nepx_net_pack(int32_t d_in, int32_t width, int32_t n_hidden, const float* const* ws,
                             const float* const* bs, float* out) {
    NEPX_REQUIRE(d_in > 0 && width > 0 && n_hidden >= 0, NEPX_EINVAL, "bad network dims");
    NEPX_REQUIRE(ws && bs && out, NEPX_EINVAL, "null pointer");
    const int DP = round16(d_in), HP = round16(width), NL = n_hidden + 1;
    NetLayout L{DP, HP, NL};
    memset(out, 0, sizeof(float) * L.total());
    for (int k = 0; k < NL; ++k) {
        const int kin = k == 0 ? d_in : width, kinp = k == 0 ? DP : HP;
        const float* W = ws[k];  // (width x kin)
        NEPX_REQUIRE(W && bs[k], NEPX_EINVAL, "null layer %d", k);
        // forward operand: [mt][s][lane] = W[16mt + lane%16][16(s/4) + 4(lane/16) + s%4]
        float* F = out + L.fwd_off(k);
        for (int mt = 0; mt < HP / 16; ++mt)
            for (int s = 0; s < kinp / 4; ++s)
                for (int lane = 0; lane < 64; ++lane) {
                    const int row = 16 * mt + lane % 16;
                    const int col = 16 * (s / 4) + 4 * (lane / 16) + s % 4;
                    F[(mt * (kinp / 4) + s) * 64 + lane] = (row < width && col < kin) ? W[row * kin + col] : 0.f;
                }
        // backward operand (W^T): [mt][s][lane] = W[16(s/4) + 4(lane/16) + s%4][16mt + lane%16]
        float* B = out + L.bwd_off(k);
        for (int mt = 0; mt < kinp / 16; ++mt)
            for (int s = 0; s < HP / 4; ++s)
                for (int lane = 0; lane < 64; ++lane) {
                    const int row = 16 * (s / 4) + 4 * (lane / 16) + s % 4;
                    const int col = 16 * mt + lane % 16;
                    B[(mt * (HP / 4) + s) * 64 + lane] = (row < width && col < kin) ? W[row * kin + col] : 0.f;
                }
        for (int u = 0; u < width; ++u) out[L.bias_off(k) + u] = bs[k][u];
    }
    NEPX_REQUIRE(ws[NL] && bs[NL], NEPX_EINVAL, "null output layer");
    for (int u = 0; u < width; ++u) out[L.wout_off() + u] = ws[NL][u];
    out[L.bout_off()] = bs[NL][0];
    return NEPX_OK;
}

static bool lowp_precision(int p) {
    return p == NEPX_NET_F16X3 || p == NEPX_NET_BF16 || p == NEPX_NET_BF16X3;
}

// float -> bf16 bits, round to nearest even (NaN stays NaN).
static uint16_t bf16_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

extern "C" int64_t nepx_net_packed_lowp(int32_t precision, int32_t d_in, int32_t width, int32_t n_hidden) {
    if (!lowp_precision(precision) || d_in <= 0 || width <= 0 || n_hidden < 0 || round16(width) % 32) return -1;
    NetLayoutH L{round32(round16(d_in)), round16(d_in), round16(width), n_hidden + 1,
                 precision == NEPX_NET_BF16 ? 1 : 2};
    return L.total();
}

extern "C" int nepx_net_pack_lowp(int32_t precision, int32_t d_in, int32_t width, int32_t n_hidden,
                                  const float* const* ws, uint16_t* out) {
    NEPX_REQUIRE(lowp_precision(precision), NEPX_EINVAL, "precision %d has no 16-bit operands", precision);
    NEPX_REQUIRE(d_in > 0 && width > 0 && n_hidden >= 0, NEPX_EINVAL, "bad network dims");
    NEPX_REQUIRE(round16(width) % 32 == 0, NEPX_EUNSUP,
                 "16-bit operand paths need a padded width that is a multiple of 32");
    NEPX_REQUIRE(ws && out, NEPX_EINVAL, "null pointer");
    const int parts = precision == NEPX_NET_BF16 ? 1 : 2;
    const int DP = round16(d_in), DPH = round32(DP), HP = round16(width), NL = n_hidden + 1;
    NetLayoutH L{DPH, DP, HP, NL, parts};
    for (int64_t k = 0; k < L.total(); ++k) out[k] = 0;
    // hi = w rounded to the 16-bit type, lo = the remainder rounded (split forms only)
    auto put = [&](int base, int lo_off, int idx, float w) {
        if (precision == NEPX_NET_F16X3) {
            const _Float16 h = (_Float16)w;
            const _Float16 l = (_Float16)(w - (float)h);
            memcpy(&out[base + idx], &h, 2);
            memcpy(&out[base + lo_off + idx], &l, 2);
        } else {
            const uint16_t h = bf16_bits(w);
            out[base + idx] = h;
            if (parts == 2) out[base + lo_off + idx] = bf16_bits(w - bf16_value(h));
        }
    };
    for (int k = 0; k < NL; ++k) {
        const int kin = k == 0 ? d_in : width;
        const float* W = ws[k];
        NEPX_REQUIRE(W, NEPX_EINVAL, "null layer %d", k);
        const int ksf = (k == 0 ? DPH : HP) / 32;
        const int lof = (HP / 16) * ksf * 512;
        for (int mt = 0; mt < HP / 16; ++mt)
            for (int s = 0; s < ksf; ++s)
                for (int lane = 0; lane < 64; ++lane)
                    for (int j = 0; j < 8; ++j) {
                        const int row = 16 * mt + lane % 16;
                        const int col = 16 * (2 * s + j / 4) + 4 * (lane / 16) + j % 4;
                        const float w = (row < width && col < kin) ? W[row * kin + col] : 0.f;
                        put(L.fwd_off(k), lof, ((mt * ksf + s) * 64 + lane) * 8 + j, w);
                    }
        const int mtb = (k == 0 ? DP : HP) / 16, ksb = HP / 32;
        const int lob = mtb * ksb * 512;
        for (int mt = 0; mt < mtb; ++mt)
            for (int s = 0; s < ksb; ++s)
                for (int lane = 0; lane < 64; ++lane)
                    for (int j = 0; j < 8; ++j) {
                        const int unit = 16 * (2 * s + j / 4) + 4 * (lane / 16) + j % 4;   // output unit of layer k
                        const int col = 16 * mt + lane % 16;                               // input unit of layer k
                        const float w = (unit < width && col < kin) ? W[unit * kin + col] : 0.f;
                        put(L.bwd_off(k), lob, ((mt * ksb + s) * 64 + lane) * 8 + j, w);
                    }
    }
    return NEPX_OK;
}

extern "C" int64_t nepx_net_packed_halves(int32_t d_in, int32_t width, int32_t n_hidden) {
    return nepx_net_packed_lowp(NEPX_NET_F16X3, d_in, width, n_hidden);
}

extern "C" int nepx_net_pack_half(int32_t d_in, int32_t width, int32_t n_hidden, const float* const* ws,
                                  uint16_t* out) {
    return nepx_net_pack_lowp(NEPX_NET_F16X3, d_in, width, n_hidden, ws, out);
}

extern "C" int nepx_descriptor_fwd(const nepx_pair_model* m, const float* qf, const int32_t* pair_i,
                                   const int32_t* pair_j, const float* pair_t, const int32_t* np_dev,
                                   int64_t maxp, float* G, void* stream) {
    int rc = check_model(m);
    if (rc) return rc;
    NEPX_REQUIRE(qf && pair_i && pair_j && pair_t && np_dev && G, NEPX_EINVAL, "null buffer");
    if (maxp <= 0) return NEPX_OK;
#define X(ns, a, b, c, T)                                                                           \
    if (m->n_sites == ns && m->nrad == a && m->nang == b && m->lmax == c) {                        \
        launch_desc_fwd<T>(*m, qf, pair_i, pair_j, pair_t, np_dev, maxp, G, stream);                \
        NEPX_LAUNCH_CHECK();                                                                        \
        return NEPX_OK;                                                                             \
    }
    NEPX_DESC_CASES(X)
#undef X
    return NEPX_EUNSUP;
}

#define NEPX_NET_CASES(X)              \
    X(40, 64, 1, Net40x64, kNetTiles)  \
    X(60, 128, 1, Net60x128, 1)        \
    X(47, 150, 3, Net47x150, 1)        \
    X(121, 150, 3, Net121x150, 1)

static bool net_supported(const nepx_net* n) {
#define X(d, w, h, T, NT) if (n->d_in == d && n->width == w && n->n_hidden == h) return true;
    NEPX_NET_CASES(X)
#undef X
    return false;
}

static int check_precision(const nepx_net* net) {
    const int p = net->precision;
    NEPX_REQUIRE(p == NEPX_NET_F32 || lowp_precision(p), NEPX_EINVAL, "bad network precision %d", p);
    NEPX_REQUIRE(p == NEPX_NET_F32 || net->w_half, NEPX_EINVAL, "precision %d without packed 16-bit operands", p);
    NEPX_REQUIRE(p == NEPX_NET_F32 || round16(net->width) % 32 == 0, NEPX_EUNSUP,
                 "no 16-bit network kernel for width %d", net->width);
    // bf16 rounding moves pre-activations across zero, where ReLU's derivative jumps
    NEPX_REQUIRE((p != NEPX_NET_BF16 && p != NEPX_NET_BF16X3) || net->act == NEPX_ACT_TANH, NEPX_EUNSUP,
                 "the bf16 network forms are for tanh networks only");
    return NEPX_OK;
}

template <class T, int NT, int PREC>
static void launch_ffnn_p(const nepx_net* net, int grid, float scale, int d, const float* G, const int32_t* np_dev,
                          int64_t maxp, float* E, float* dEdG, void* stream) {
#define NEPX_LAUNCH_FFNN(A)                                                                                   \
    hipLaunchKernelGGL((k_ffnn<T, NT, PREC, A>), dim3(grid), dim3(kBlock), 0, as_stream(stream), net->w_packed, \
                       net->w_half, net->en_min, scale, d, G, np_dev, maxp, E, dEdG)
    if (net->act == NEPX_ACT_TANH) NEPX_LAUNCH_FFNN(NEPX_ACT_TANH);
    else if constexpr (PREC == NEPX_NET_F32 || PREC == NEPX_NET_F16X3) NEPX_LAUNCH_FFNN(NEPX_ACT_RELU);
#undef NEPX_LAUNCH_FFNN
}

template <class T, int NT>
static int launch_ffnn(const nepx_net* net, int grid, float scale, int d, const float* G, const int32_t* np_dev,
                       int64_t maxp, float* E, float* dEdG, void* stream) {
    if constexpr (T::HP % 32 == 0) {
        switch (net->precision) {
            case NEPX_NET_F16X3:
                launch_ffnn_p<T, NT, NEPX_NET_F16X3>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);
                break;
            case NEPX_NET_BF16:
                launch_ffnn_p<T, NT, NEPX_NET_BF16>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);
                break;
            case NEPX_NET_BF16X3:
                launch_ffnn_p<T, NT, NEPX_NET_BF16X3>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);
                break;
            default:
                launch_ffnn_p<T, NT, NEPX_NET_F32>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);
        }
    } else {
        launch_ffnn_p<T, NT, NEPX_NET_F32>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);
    }
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

// Few tiles (up to four per CU): the FP32 network runs one tile per block on k_ffnn_coop, which gives
// the same results as k_ffnn.
template <class T>
static bool launch_ffnn_coop(const nepx_net* net, float scale, int d, const float* G, const int32_t* np_dev,
                             int64_t maxp, float* E, float* dEdG, void* stream) {
    const bool f32 = net->precision == NEPX_NET_F32 || T::HP % 32 != 0;
    const int64_t tiles = (maxp + 15) / 16;
    if (!f32 || tiles > 4 * (int64_t)device_cus()) return false;
    const dim3 grid((unsigned)tiles), block(T::MT / coop_mpw<T>() * 64);
    if (net->act == NEPX_ACT_TANH)
        hipLaunchKernelGGL((k_ffnn_coop<T, NEPX_ACT_TANH>), grid, block, 0, as_stream(stream), net->w_packed,
                           net->en_min, scale, d, G, np_dev, maxp, E, dEdG);
    else
        hipLaunchKernelGGL((k_ffnn_coop<T, NEPX_ACT_RELU>), grid, block, 0, as_stream(stream), net->w_packed,
                           net->en_min, scale, d, G, np_dev, maxp, E, dEdG);
    return true;
}

extern "C" int nepx_ffnn_fwd_bwd(const nepx_net* net, const float* G, const int32_t* np_dev, int64_t maxp,
                                 float* E, float* dEdG, void* stream) {
    NEPX_REQUIRE(net && net->w_packed, NEPX_EINVAL, "null network");
    NEPX_REQUIRE(G && np_dev && E && dEdG, NEPX_EINVAL, "null buffer");
    NEPX_REQUIRE(net_supported(net), NEPX_EUNSUP, "no network kernel for %d -> %d x%d", net->d_in,
                 net->width, net->n_hidden + 1);
    int rc = check_precision(net);
    if (rc) return rc;
    if (maxp <= 0) return NEPX_OK;
    const float scale = net->en_max - net->en_min;
#define X(d, w, h, T, NT)                                                                        \
    if (net->d_in == d && net->width == w && net->n_hidden == h) {                               \
        if (launch_ffnn_coop<T>(net, scale, d, G, np_dev, maxp, E, dEdG, stream)) {              \
            NEPX_LAUNCH_CHECK();                                                                 \
            return NEPX_OK;                                                                      \
        }                                                                                        \
        const int grid = grid_for_cap((maxp + 16 * NT - 1) / (16 * NT), kBlock / 64,            \
                                      (int64_t)device_cus() * tuning().net_blocks_per_cu);       \
        return launch_ffnn<T, NT>(net, grid, scale, d, G, np_dev, maxp, E, dEdG, stream);       \
    }
    NEPX_NET_CASES(X)
#undef X
    return NEPX_EUNSUP;
}

extern "C" int nepx_descriptor_adjoint(const nepx_pair_model* m, const float* qf, const int32_t* pair_i,
                                       const int32_t* pair_j, const float* pair_t, const int32_t* np_dev,
                                       int64_t maxp, const float* dEdG, const float* E, float clip,
                                       float* out, void* stream) {
    int rc = check_model(m);
    if (rc) return rc;
    NEPX_REQUIRE(qf && pair_i && pair_j && pair_t && np_dev && dEdG && out, NEPX_EINVAL, "null buffer");
    if (maxp <= 0) return NEPX_OK;
#define X(ns, a, b, c, T)                                                                            \
    if (m->n_sites == ns && m->nrad == a && m->nang == b && m->lmax == c) {                         \
        launch_adjoint<T>(*m, qf, pair_i, pair_j, pair_t, np_dev, maxp, dEdG, E, clip, out, stream);  \
        NEPX_LAUNCH_CHECK();                                                                         \
        return NEPX_OK;                                                                              \
    }
    NEPX_DESC_CASES(X)
#undef X
    return NEPX_EUNSUP;
}

#define NEPX_FUSED_CASES(X)                      \
    X(8, 11, 6, 4, CfgCube8, 40, 64, 1, Net40x64) \
    X(10, 11, 6, 4, CfgTet10, 40, 64, 1, Net40x64) \
    X(12, 17, 6, 6, CfgCyl12, 60, 128, 1, Net60x128)

// The bf16 forms exist for tanh networks only (check_precision).
static bool fused_precision_ok(const nepx_net* net) {
    return net->precision == NEPX_NET_F32 || net->precision == NEPX_NET_F16X3 || net->act == NEPX_ACT_TANH;
}

extern "C" int nepx_pair_fused_supported(const nepx_pair_model* m, const nepx_net* net) {
    if (!m || !net) return 0;
#define X(ns, a, b, c, T, d, w, h, NT)                                                            \
    if (m->n_sites == ns && m->nrad == a && m->nang == b && m->lmax == c && net->d_in == d &&    \
        net->width == w && net->n_hidden == h && fused_precision_ok(net))                        \
        return QuadCfg<T>::OK && tuning().pair_lanes == 4 ? 4 : 1;
    NEPX_FUSED_CASES(X)
#undef X
    return 0;
}

template <class Cfg, class N, int PREC>
static void launch_fused_p(const nepx_pair_model& m, const nepx_net* net, int grid, float scale, const float* qf,
                           const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                           const int32_t* np_dev, int64_t maxp, float clip, float* out, RevOut rv, void* stream) {
#define NEPX_LAUNCH_FUSED(A)                                                                                     \
    hipLaunchKernelGGL((k_fused<Cfg, N, PREC, A>), dim3(grid), dim3(kBlock), 0, as_stream(stream), m,            \
                       net->w_packed, net->w_half, net->en_min, scale, (const float4*)qf, pair_i, pair_j,          \
                       (const float4*)pair_t, np_dev, maxp, clip, out, rv)
    if (net->act == NEPX_ACT_TANH) NEPX_LAUNCH_FUSED(NEPX_ACT_TANH);
    else if constexpr (PREC == NEPX_NET_F32 || PREC == NEPX_NET_F16X3) NEPX_LAUNCH_FUSED(NEPX_ACT_RELU);
#undef NEPX_LAUNCH_FUSED
}

template <class Cfg, class N>
static int launch_fused(const nepx_pair_model& m, const nepx_net* net, float scale,
                        const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                        const int32_t* np_dev, int64_t maxp, float clip, float* out, RevOut rv, void* stream) {
    const int64_t cap = (int64_t)device_cus() * tuning().fused_blocks_per_cu;
    // four lanes per pair (site sums on MFMA) where the shape allows it, unless tuned off
    if (tuning().pair_lanes == 4) {
        const int rc = launch_fused_quad<Cfg, N>(m, net, net->precision, (int)cap, scale, qf, pair_i, pair_j,
                                                 pair_t, np_dev, maxp, clip, out, rv, stream);
        if (rc != NEPX_EUNSUP) return rc;
    }
    const int grid = grid_for_cap(maxp, kBlock, cap);
    switch (net->precision) {
        case NEPX_NET_F16X3:
            launch_fused_p<Cfg, N, NEPX_NET_F16X3>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                    clip, out, rv, stream);
            break;
        case NEPX_NET_BF16:
            launch_fused_p<Cfg, N, NEPX_NET_BF16>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                   clip, out, rv, stream);
            break;
        case NEPX_NET_BF16X3:
            launch_fused_p<Cfg, N, NEPX_NET_BF16X3>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev,
                                                     maxp, clip, out, rv, stream);
            break;
        default:
            launch_fused_p<Cfg, N, NEPX_NET_F32>(m, net, grid, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp,
                                                  clip, out, rv, stream);
    }
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

static int pair_fused(const nepx_pair_model* m, const nepx_net* net, const float* qf, const int32_t* pair_i,
                      const int32_t* pair_j, const float* pair_t, const int32_t* np_dev, int64_t maxp, float clip,
                      float* out, RevOut rv, void* stream) {
    int rc = check_model(m);
    if (rc) return rc;
    rc = check_net(net, model_D(m));
    if (rc) return rc;
    rc = check_precision(net);
    if (rc) return rc;
    NEPX_REQUIRE(qf && pair_i && pair_j && pair_t && np_dev && out, NEPX_EINVAL, "null buffer");
    if (maxp <= 0) return NEPX_OK;
    const float scale = net->en_max - net->en_min;
#define X(ns, a, b, c, T, d, w, h, NT)                                                               \
    if (m->n_sites == ns && m->nrad == a && m->nang == b && m->lmax == c && net->d_in == d &&        \
        net->width == w && net->n_hidden == h)                                                       \
        return launch_fused<T, NT>(*m, net, scale, qf, pair_i, pair_j, pair_t, np_dev, maxp, clip, out, rv, stream);
    NEPX_FUSED_CASES(X)
#undef X
    set_error("no fused kernel for n_sites=%d hypers=(%d,%d,%d) net %d->%dx%d", m->n_sites, m->nrad,
              m->nang, m->lmax, net->d_in, net->width, net->n_hidden + 1);
    return NEPX_EUNSUP;
}

extern "C" int nepx_pair_fused(const nepx_pair_model* m, const nepx_net* net, const float* qf,
                               const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                               const int32_t* np_dev, int64_t maxp, float clip, float* out, void* stream) {
    return pair_fused(m, net, qf, pair_i, pair_j, pair_t, np_dev, maxp, clip, out, RevOut{}, stream);
}

extern "C" int nepx_pair_fused_list(const nepx_pair_model* m, const nepx_net* net, const nepx_state* st,
                                    nepx_nlist* nl, float clip, float* out, void* stream) {
    NEPX_REQUIRE(st && nl && st->qf && nl->flags, NEPX_EINVAL, "null state or list");
    NEPX_REQUIRE(!(nl->pair_e || nl->rrec) || (nl->rpos && nl->pair_e && nl->rrec), NEPX_EINVAL,
                 "reverse-order records need rpos, pair_e and rrec together");
    RevOut rv;
    if (nl->rrec) {
        rv.pe = nl->pair_e;
        rv.rrec = nl->rrec;
        rv.flags = nl->flags;
    }
    return pair_fused(m, net, st->qf, nl->pair_i, nl->pair_j, nl->pair_t, nl->flags + F_P, nl->cap, clip, out, rv,
                      stream);
}

extern "C" int nepx_stencil_rows(const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                                 int64_t P, float dx, float dteta, float* q_rows, int32_t* row_i, int32_t* row_j,
                                 float* t_rows, void* stream) {
    NEPX_REQUIRE(P >= 0 && P <= (int64_t)1 << 27, NEPX_ECAP, "pair count %lld out of range", (long long)P);
    if (P == 0) return NEPX_OK;
    NEPX_REQUIRE(qf && pair_i && pair_j && pair_t && q_rows && row_i && row_j && t_rows, NEPX_EINVAL, "null buffer");
    const int rows = (int)(10 * P);
    hipLaunchKernelGGL(k_stencil_rows, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, as_stream(stream),
                       (const float4*)qf, pair_i, pair_j, (const float4*)pair_t, (int)P, dx, dteta,
                       (float4*)q_rows, row_i, row_j, (float4*)t_rows);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_fd_forces(const float* E_rows, int64_t P, float dx, float dteta, float clip, float* out,
                              void* stream) {
    NEPX_REQUIRE(P >= 0 && P <= (int64_t)1 << 27, NEPX_ECAP, "pair count %lld out of range", (long long)P);
    if (P == 0) return NEPX_OK;
    NEPX_REQUIRE(E_rows && out, NEPX_EINVAL, "null buffer");
    NEPX_REQUIRE(dx != 0.f && dteta != 0.f, NEPX_EINVAL, "zero finite-difference step");
    hipLaunchKernelGGL(k_fd_forces, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, as_stream(stream), E_rows,
                       (int)P, dx, dteta, clip, out);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

// ---------------------------------------------------------------------------
// Descriptors of arbitrary site clouds (training-set featurisation,
// 05/NEP.py:171-219 get_g_fast and 04/Processors.py:48-88): one cloud per
// lane, sites relative to the cloud centroid, optional per-site weights.
template <class Cfg>
__global__ __launch_bounds__(kBlock) void k_cloud_desc(const float* __restrict__ clouds, int64_t P, int nd,
                                                       const float* __restrict__ w, float r_nep,
                                                       float* __restrict__ G) {
    for (int64_t idx = blockIdx.x * (int64_t)kBlock + threadIdx.x; idx < P; idx += (int64_t)gridDim.x * kBlock) {
        const CloudSites sites(clouds + idx * 3 * (int64_t)nd, w, nd, 1.f / r_nep);
        if constexpr (Cfg::ONE_PASS) {
            float Gr[Cfg::D];
            desc_forward_sites<Cfg>(sites, Gr);
            store_row<Cfg>(G + idx * Cfg::D, Gr);
        } else {
            desc_forward_sites<Cfg>(sites, G + idx * Cfg::D);
        }
    }
}

#define NEPX_CLOUD_CASES(X)     \
    X(11, 6, 4)                 \
    X(10, 8, 4)                 \
    X(17, 6, 6)                 \
    X(10, 10, 10)

extern "C" int nepx_cloud_descriptor(const float* clouds, int64_t n_clouds, int32_t n_sites, const float* site_w,
                                     int32_t nrad, int32_t nang, int32_t lmax, float r_nep, float* G,
                                     void* stream) {
    NEPX_REQUIRE(n_clouds >= 0 && n_sites > 0, NEPX_EINVAL, "bad cloud count %lld or size %d",
                 (long long)n_clouds, n_sites);
    NEPX_REQUIRE(r_nep > 0.f, NEPX_EINVAL, "r_nep must be positive");
    if (n_clouds == 0) return NEPX_OK;
    NEPX_REQUIRE(clouds && G, NEPX_EINVAL, "null buffer");
    const int grid = grid_for(n_clouds, kBlock);
#define X(a, b, c)                                                                                     \
    if (nrad == a && nang == b && lmax == c) {                                                         \
        hipLaunchKernelGGL((k_cloud_desc<DescCfg<1, a, b, c>>), dim3(grid), dim3(kBlock), 0, as_stream(stream), \
                           clouds, n_clouds, (int)n_sites, site_w, r_nep, G);                           \
        NEPX_LAUNCH_CHECK();                                                                           \
        return NEPX_OK;                                                                                \
    }
    NEPX_CLOUD_CASES(X)
#undef X
    set_error("no cloud descriptor kernel for hypers (%d, %d, %d)", nrad, nang, lmax);
    return NEPX_EUNSUP;
}
