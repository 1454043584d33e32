// NEP pair descriptor: forward and hand-written adjoint, one pair per lane.
//
// Math (restated from nep_simulator/cudakernels/NepDescriptorKernels.py:3-144 and
// nep_simulator/DescriptorGenerator.py:296-432):
//   cloud p_a: i's sites R(q_i) b_a, j's sites R(q_j) b_b + t   (PosToPtsKernels.py:3-79)
//   r_a = p_a - centroid, rho_a = |r_a| / r_nep
//   f_n(a) = (T_n(x_a) + 1)/2 * fc(rho_a),  x = 2(rho-1)^2 - 1,  fc = (1 + cos(pi rho))/2
//   g_rad[n] = sum_a f_n(a)
//   g_ang[n*lmax + l-1] = sum_a sum_b f_n(a) f_n(b) P_l(u_a . u_b)
//
// The angular double sum is evaluated through the addition theorem
//   sum_ab f(a) f(b) P_l(u_a.u_b) = sum_m A_lm^2,  A_lm = sum_a f(a) Yhat_lm(u_a),
// with real harmonics normalised so that sum_m Yhat_lm(u)Yhat_lm(v) = P_l(u.v).
// That is O(Nd (lmax+1)^2) per n instead of O(Nd^2 lmax) and has a simple
// gradient.  Yhat_lm is built from the Cartesian recurrence of the
// homogeneous solid harmonics (Q_l^m(z) times Re/Im (x+iy)^m), whose
// derivatives are  dQ_l^m/dx = x Q_{l-1}^{m+1},  dQ_l^m/dy = y Q_{l-1}^{m+1},
// dQ_l^m/dz = (l+m) Q_{l-1}^m  (A. Bigi et al., "sphericart", J. Chem. Phys.
// 159, 064802 (2023)).
#pragma once
#include <type_traits>

#include "nepx_common.h"

namespace nepx {

template <int LMAX>
struct HarmTables {
    static constexpr int NLM = LMAX * (LMAX + 2);  // l = 1..LMAX
    static constexpr int NQ = (LMAX + 1) * (LMAX + 2) / 2;
    __host__ __device__ static constexpr int qi(int l, int m) { return l * (l + 1) / 2 + m; }
};

// Legendre parts Q_l^m(z) at |u| = 1 (Condon-Shortley phase) and
// C_m + i S_m = (x + iy)^m.  The unnormalised solid harmonics are
//   Z_l0 = Q_l^0,  Z_lm^c = Q_l^m C_m,  Z_lm^s = Q_l^m S_m   (m >= 1)
// and Yhat_lm = K_lm Z_lm.  The descriptor kernels accumulate A' = sum f Z
// and weight by K_lm^2 where the squares are formed, so no per-site
// multiplication by K is needed.
template <int LMAX>
struct Harm {
    using T = HarmTables<LMAX>;
    float Q[T::NQ];
    float C[LMAX + 1], S[LMAX + 1];

    __device__ __forceinline__ float q(int l, int m) const { return Q[T::qi(l, m)]; }
    // Z_l0 = Q_l^0; Z_lm = Q_l^m C_m (part 0) or Q_l^m S_m (part 1) for m >= 1.
    __device__ __forceinline__ float z(int l, int m, int part) const {
        return m == 0 ? q(l, 0) : q(l, m) * (part ? S[m] : C[m]);
    }

    __device__ __forceinline__ void eval(float x, float y, float z) {
        Q[0] = 1.f;
#pragma unroll
        for (int m = 1; m <= LMAX; ++m) Q[T::qi(m, m)] = -(2.f * m - 1.f) * Q[T::qi(m - 1, m - 1)];
#pragma unroll
        for (int m = 0; m < LMAX; ++m) Q[T::qi(m + 1, m)] = (2.f * m + 1.f) * z * Q[T::qi(m, m)];
#pragma unroll
        for (int m = 0; m + 2 <= LMAX; ++m) {
#pragma unroll
            for (int l = m + 2; l <= LMAX; ++l) {
                const float a = (2.f * l - 1.f) / (float)(l - m);
                const float b = (float)(l + m - 1) / (float)(l - m);
                Q[T::qi(l, m)] = a * z * Q[T::qi(l - 1, m)] - b * Q[T::qi(l - 2, m)];
            }
        }
        C[0] = 1.f;
        S[0] = 0.f;
#pragma unroll
        for (int m = 1; m <= LMAX; ++m) {
            C[m] = x * C[m - 1] - y * S[m - 1];
            S[m] = x * S[m - 1] + y * C[m - 1];
        }
    }
};

// K_lm^2 = 2 (l-m)!/(l+m)! (m > 0), 1 (m = 0), at compile time.
template <int LMAX>
struct K2Tab {
    float v[(LMAX + 1) * (LMAX + 1)];
    constexpr K2Tab() : v() {
        for (int l = 0; l <= LMAX; ++l)
            for (int m = 0; m <= l; ++m)
                v[l * (LMAX + 1) + m] = m == 0 ? 1.0f : (float)(2.0 * cfact(l - m) / cfact(l + m));
    }
    __host__ __device__ constexpr float operator()(int l, int m) const { return v[l * (LMAX + 1) + m]; }
};

// Floats of angular accumulators one lane holds in a pass over the sites.
// The l range 1..LMAX is cut greedily into passes of at most this many
// accumulators ((nang+1) x lm count); an l that alone exceeds it gets a pass
// of its own.
constexpr int kAccBudget = 176;
constexpr int pass_end(int nang, int lmax, int l0) {
    int l1 = l0;
    while (l1 < lmax && (nang + 1) * ((l1 + 2) * (l1 + 2) - l0 * l0) <= kAccBudget) ++l1;
    return l1;
}

// Slot of (l, m, component) in the NLM accumulators: l*l-1 + {0: m = 0, 2m-1: cos, 2m: sin}.
__host__ __device__ constexpr int lm_slot(int l, int m, int sin_part) {
    return l * l - 1 + (m == 0 ? 0 : 2 * m - 1 + sin_part);
}

template <int NS_, int NRAD_, int NANG_, int LMAX_>
struct DescCfg {
    static constexpr int NS = NS_, NRAD = NRAD_, NANG = NANG_, LMAX = LMAX_;
    static constexpr int ND = 2 * NS;
    static constexpr int NLM = LMAX * (LMAX + 2);
    static constexpr int D = (NRAD + 1) + (NANG + 1) * LMAX;
    static_assert(NANG <= NRAD, "nang must not exceed nrad");
    // One pass over the sites holds every angular accumulator.
    static constexpr bool ONE_PASS = pass_end(NANG, LMAX, 1) >= LMAX;
};

// Number of passes over Cfg's l range.
constexpr int n_passes(int nang, int lmax) {
    int n = 1;
    for (int l0 = 1; pass_end(nang, lmax, l0) < lmax; l0 = pass_end(nang, lmax, l0) + 1) ++n;
    return n;
}

// A contiguous range of l: accumulator slots lm_slot(l, m, part) - OFF.
template <int L0_, int L1_>
struct LRange {
    static constexpr int L0 = L0_, L1 = L1_;
    static constexpr int OFF = L0 * L0 - 1;
    static constexpr int N = (L1 + 1) * (L1 + 1) - L0 * L0;
};

// First l of the last pass.
constexpr int last_pass_start(int nang, int lmax) {
    int l0 = 1;
    while (pass_end(nang, lmax, l0) < lmax) l0 = pass_end(nang, lmax, l0) + 1;
    return l0;
}
template <class Cfg>
using LastPass = LRange<last_pass_start(Cfg::NANG, Cfg::LMAX), Cfg::LMAX>;

// Calls fn(LRange<...>{}, first) for each pass of Cfg's l range, in order.
template <class Cfg, int L0 = 1, class Fn>
__device__ __forceinline__ void for_passes(Fn&& fn) {
    constexpr int L1 = pass_end(Cfg::NANG, Cfg::LMAX, L0);
    fn(LRange<L0, L1>{}, std::integral_constant<bool, L0 == 1>{});
    if constexpr (L1 < Cfg::LMAX) for_passes<Cfg, L1 + 1>(fn);
}

// One pair in the frame centred on particle i: orientations, t = X_j - X_i,
// the sums of the lever arms of each body's sites, s_i = R(q_i) sum_k b_k and
// s_j = R(q_j) sum_k b_k, and the centroid of the 2*Ns sites,
//   c = (s_i + s_j + Ns t) / (2 Ns).
struct PairIn {
    float4 qi, qj;
    float tx, ty, tz;
    float cx, cy, cz;
    float six, siy, siz, sjx, sjy, sjz;
};

template <class Cfg>
__device__ __forceinline__ PairIn pair_in(const nepx_pair_model& m, float4 qi, float4 qj, float tx, float ty,
                                          float tz) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int k = 0; k < Cfg::NS; ++k) {
        sx += m.sites[3 * k];
        sy += m.sites[3 * k + 1];
        sz += m.sites[3 * k + 2];
    }
    PairIn p;
    p.qi = qi;
    p.qj = qj;
    p.tx = tx;
    p.ty = ty;
    p.tz = tz;
    mat_vec(quat_matrix(qi.x, qi.y, qi.z, qi.w), sx, sy, sz, p.six, p.siy, p.siz);
    mat_vec(quat_matrix(qj.x, qj.y, qj.z, qj.w), sx, sy, sz, p.sjx, p.sjy, p.sjz);
    const float inv = 1.f / (float)Cfg::ND;
    p.cx = (p.six + p.sjx) * inv + 0.5f * tx;
    p.cy = (p.siy + p.sjy) * inv + 0.5f * ty;
    p.cz = (p.siz + p.sjz) * inv + 0.5f * tz;
    return p;
}

// Site geometry shared by forward and adjoint.
struct SiteGeo {
    float ax, ay, az;  // lever arm about the owning body's centre: R b_k
    float ux, uy, uz;  // unit vector from the centroid
    float r, ir, rho;  // |r_a|, 1/|r_a|, |r_a|/r_nep
};

// Body-frame site b, body rotation R, body centre minus cloud centroid o.
__device__ __forceinline__ SiteGeo site_geo_b(const Mat3& R, float bx, float by, float bz, float ox, float oy,
                                              float oz, float inv_rn) {
    SiteGeo s;
    mat_vec(R, bx, by, bz, s.ax, s.ay, s.az);
    const float rx = s.ax + ox, ry = s.ay + oy, rz = s.az + oz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    s.ir = __builtin_amdgcn_rsqf(r2);  // hardware 1/sqrt (about 1 ulp)
    s.r = r2 * s.ir;
    s.ux = rx * s.ir;
    s.uy = ry * s.ir;
    s.uz = rz * s.ir;
    s.rho = s.r * inv_rn;
    return s;
}

__device__ __forceinline__ SiteGeo site_geo(const nepx_pair_model& m, const Mat3& R, int k, float ox, float oy,
                                            float oz, float inv_rn) {
    return site_geo_b(R, m.sites[3 * k], m.sites[3 * k + 1], m.sites[3 * k + 2], ox, oy, oz, inv_rn);
}

// cos(pi x) and sin(pi x) on the hardware units (v_cos_f32 / v_sin_f32 take the
// angle in revolutions, valid for |x| < 512); rho is non-negative and small here.
__device__ __forceinline__ float cos_pi(float x) { return __builtin_amdgcn_cosf(0.5f * x); }
__device__ __forceinline__ float sin_pi(float x) { return __builtin_amdgcn_sinf(0.5f * x); }

// Radial functions at rho: fn(n, f_n) for n = 0..NMAX with
// f_n = (T_n(x) + 1) w fc/2,  x = 2(rho-1)^2 - 1,  fc = (1 + cos(pi rho))/2.
template <int NMAX, class Fn>
__device__ __forceinline__ void radial_values(float rho, float w, Fn&& fn) {
    const float x = 2.f * (rho - 1.f) * (rho - 1.f) - 1.f;
    const float hf = 0.25f * (1.f + cos_pi(rho)) * w;  // w fc / 2
    float t0 = 1.f, t1 = x;
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        float tn;
        if (n == 0) tn = 1.f;
        else if (n == 1) tn = x;
        else {
            tn = 2.f * x * t1 - t0;
            t0 = t1;
            t1 = tn;
        }
        fn(n, tn * hf + hf);
    }
}

// Radial functions and their rho-derivatives: fn(n, f_n, f_n') for n = 0..NMAX, with
// f_n' = T_n'(x) x' fc/2 + (T_n(x) + 1) fc'/2 (unit weight).
template <int NMAX, class Fn>
__device__ __forceinline__ void radial_derivs(float rho, Fn&& fn) {
    const float xm = rho - 1.f;
    const float x = 2.f * xm * xm - 1.f;
    const float sp = sin_pi(rho), cp = cos_pi(rho);
    const float hf = 0.25f * (1.f + cp);                // fc / 2
    const float hdx = 2.f * xm * (0.5f * (1.f + cp));    // (dx/drho) fc / 2
    const float hdf = -0.25f * (float)M_PI * sp;         // fc' / 2
    float t0 = 1.f, t1 = x, d0 = 0.f, d1 = 1.f;
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        float tn, dn;
        if (n == 0) {
            tn = 1.f;
            dn = 0.f;
        } else if (n == 1) {
            tn = x;
            dn = 1.f;
        } else {
            tn = 2.f * x * t1 - t0;
            dn = 2.f * t1 + 2.f * x * d1 - d0;
            t0 = t1;
            t1 = tn;
            d0 = d1;
            d1 = dn;
        }
        fn(n, tn * hf + hf, dn * hdx + (tn * hdf + hdf));
    }
}

// Radial part of the adjoint at one site: f_n and f_n' for n < NA (the operands of the c and d
// products) and the return value sum_n w_n f_n'(rho) over n = 0..NMAX, w_n = dE/dg_rad[n].  The
// derivative is taken through T_n' = n U_{n-1} (U of the second kind), and the terms n >= NA, which
// enter only the sum, are gathered as hdx sum (n w_n) U_{n-1} + hdf (sum w_n T_n + sum w_n): four
// instructions each instead of seven.  nw[n] = n w_n and w_hi = sum_{n >= NA} w_n are per pair.
template <int NMAX, int NA>
__device__ __forceinline__ float radial_adjoint(float rho, const float (&w)[NMAX + 1], const float (&nw)[NMAX + 1],
                                                float w_hi, float (&fa)[NA], float (&da)[NA]) {
    const float xm = rho - 1.f;
    const float x = 2.f * xm * xm - 1.f;
    const float sp = sin_pi(rho), cp = cos_pi(rho);
    const float hf = 0.25f * (1.f + cp);                // fc / 2
    const float hdx = 2.f * xm * (0.5f * (1.f + cp));    // (dx/drho) fc / 2
    const float hdf = -0.25f * (float)M_PI * sp;         // fc' / 2
    const float x2 = 2.f * x;
    float t0 = 1.f, t1 = x;   // T_{n-2}, T_{n-1}
    float u0 = 1.f, u1 = x2;  // U_{n-3}, U_{n-2} (from n = 3)
    float s_lo = 0.f, s_t = 0.f, s_u = 0.f;
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        float tn, un;  // T_n, U_{n-1}
        if (n == 0) {
            tn = 1.f;
            un = 0.f;
        } else if (n == 1) {
            tn = x;
            un = 1.f;
        } else {
            tn = x2 * t1 - t0;
            t0 = t1;
            t1 = tn;
            if (n == 2) {
                un = x2;
            } else {
                un = x2 * u1 - u0;
                u0 = u1;
                u1 = un;
            }
        }
        if (n < NA) {
            const float dn = (float)n * un;
            fa[n] = tn * hf + hf;
            da[n] = dn * hdx + (tn * hdf + hdf);
            s_lo += w[n] * da[n];
        } else {
            s_t += w[n] * tn;
            s_u += nw[n] * un;
        }
    }
    return s_lo + (hdx * s_u + hdf * (s_t + w_hi));
}

// Site visitors.  visit(site, body_done) calls site(s, w) for every site
// (geometry relative to the cloud centroid, per-site weight w) and
// body_done(b) after the sites of body b.
template <class Cfg>
struct PairSites {  // the 2 Ns sites of one pair, bodies i (0) and j (1), weight 1
    const nepx_pair_model& m;
    const PairIn& p;
    float inv_rn;
    template <class S, class B>
    __device__ __forceinline__ void visit(S&& site, B&& body_done) const {
#pragma unroll 1
        for (int body = 0; body < 2; ++body) {
            const float4 q = body ? p.qj : p.qi;
            const Mat3 R = quat_matrix(q.x, q.y, q.z, q.w);
            const float ox = (body ? p.tx : 0.f) - p.cx, oy = (body ? p.ty : 0.f) - p.cy,
                        oz = (body ? p.tz : 0.f) - p.cz;
#pragma unroll 1
            for (int k = 0; k < Cfg::NS; ++k) site(site_geo(m, R, k, ox, oy, oz, inv_rn), 1.f);
            body_done(body);
        }
    }
};

struct CloudSites {  // nd points of a site cloud (xyz interleaved), optional per-point weights
    const float* __restrict__ pts;
    const float* __restrict__ w;
    int nd;
    float cx, cy, cz, inv_rn;
    __device__ __forceinline__ CloudSites(const float* pts_, const float* w_, int nd_, float inv_rn_)
        : pts(pts_), w(w_), nd(nd_), inv_rn(inv_rn_) {
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int a = 0; a < nd; ++a) {
            sx += pts[3 * a];
            sy += pts[3 * a + 1];
            sz += pts[3 * a + 2];
        }
        const float inv = 1.f / (float)nd;
        cx = sx * inv;
        cy = sy * inv;
        cz = sz * inv;
    }
    template <class S, class B>
    __device__ __forceinline__ void visit(S&& site, B&& body_done) const {
#pragma unroll 1
        for (int a = 0; a < nd; ++a) {
            SiteGeo s;
            s.ax = s.ay = s.az = 0.f;
            const float rx = pts[3 * a] - cx, ry = pts[3 * a + 1] - cy, rz = pts[3 * a + 2] - cz;
            const float r2 = rx * rx + ry * ry + rz * rz;
            s.ir = __builtin_amdgcn_rsqf(r2);
            s.r = r2 * s.ir;
            s.ux = rx * s.ir;
            s.uy = ry * s.ir;
            s.uz = rz * s.ir;
            s.rho = s.r * inv_rn;
            site(s, w ? w[a] : 1.f);
        }
        body_done(0);
    }
};

// Forward accumulators over the l of range LR: A[n][lm - LR::OFF] = sum_a w_a f_n(a) Z_lm(u_a),
// and (RAD) g_rad[n] = sum_a w_a f_n(a).
template <class Cfg, class LR, bool RAD, class Sites>
__device__ __forceinline__ void accumulate_sites(const Sites& sites, float (&grad)[Cfg::NRAD + 1],
                                                 float (&A)[Cfg::NANG + 1][LR::N]) {
    if constexpr (RAD) {
#pragma unroll
        for (int n = 0; n <= Cfg::NRAD; ++n) grad[n] = 0.f;
    }
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n)
#pragma unroll
        for (int k = 0; k < LR::N; ++k) A[n][k] = 0.f;
    sites.visit(
        [&](const SiteGeo& s, float w) {
            float fa[Cfg::NANG + 1];
            radial_values<RAD ? Cfg::NRAD : Cfg::NANG>(s.rho, w, [&](int n, float f) {
                if constexpr (RAD) grad[n] += f;
                if (n <= Cfg::NANG) fa[n] = f;
            });
            Harm<LR::L1> h;
            h.eval(s.ux, s.uy, s.uz);
#pragma unroll
            for (int l = LR::L0; l <= LR::L1; ++l) {
#pragma unroll
                for (int mm = 0; mm <= l; ++mm) {
#pragma unroll
                    for (int part = 0; part < (mm ? 2 : 1); ++part) {
                        const float z = h.z(l, mm, part);
                        const int o = lm_slot(l, mm, part) - LR::OFF;
#pragma unroll
                        for (int n = 0; n <= Cfg::NANG; ++n) A[n][o] += fa[n] * z;
                    }
                }
            }
        },
        [](int) {});
}

template <class Cfg, class LR, bool RAD>
__device__ __forceinline__ void desc_accumulate(const nepx_pair_model& m, const PairIn& p,
                                                float (&grad)[Cfg::NRAD + 1],
                                                float (&A)[Cfg::NANG + 1][LR::N]) {
    accumulate_sites<Cfg, LR, RAD>(PairSites<Cfg>{m, p, 1.f / m.r_nep}, grad, A);
}

// G entries of range LR: g_ang[n][l] = sum_m K_lm^2 A[n][lm]^2, at G[NRAD+1 + n*LMAX + l-1].
template <class Cfg, class LR, class Row>
__device__ __forceinline__ void desc_finish(const float (&A)[Cfg::NANG + 1][LR::N], Row&& G) {
    constexpr K2Tab<Cfg::LMAX> k2{};
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n) {
#pragma unroll
        for (int l = LR::L0; l <= LR::L1; ++l) {
            const float a0 = A[n][lm_slot(l, 0, 0) - LR::OFF];
            float s = a0 * a0;
#pragma unroll
            for (int mm = 1; mm <= l; ++mm) {
                const float ac = A[n][lm_slot(l, mm, 0) - LR::OFF], as = A[n][lm_slot(l, mm, 1) - LR::OFF];
                s += k2(l, mm) * (ac * ac + as * as);
            }
            G[Cfg::NRAD + 1 + n * Cfg::LMAX + (l - 1)] = s;
        }
    }
}

// Angular part of the adjoint at one site.  cd(o, c, d) supplies, for slot o of
// range LR, c_lm = sum_n B_nlm f_n(rho_a) and d_lm = sum_n B_nlm f_n'(rho_a).
// Adds sum_lm d_lm Z_lm to dEdrho and returns
//   (gx, gy, gz) = grad sum_lm c_lm Z_lm  (Z_lm extended to homogeneous polynomials of degree l)
// using  d/dx Q_l^m = x Q_{l-1}^{m+1},  d/dy Q_l^m = y Q_{l-1}^{m+1},  d/dz Q_l^m = (l+m) Q_{l-1}^m,
//        d/dx (C_m, S_m) = m (C_{m-1}, S_{m-1}),  d/dy (C_m, S_m) = m (-S_{m-1}, C_{m-1}),
// with the sums over l taken first for each m, so each (l, m) costs a few FMAs.  The radial
// projection sum_lm l c_lm Z_lm(u_a) is not summed here: by Euler's theorem on the degree-l
// terms it equals u_a . grad at |u_a| = 1 (site_dEdr).  PER_M: sum_lm d_lm Z_lm is formed per m as
// C_m sum_l d_lm Q_l^m + S_m sum_l ... (fewer multiplies; the four-lane kernel), else per slot (fewer
// live registers; the one-lane kernels, whose hypers 10 10 10 adjoint would spill more).
template <class LR, bool PER_M = false, class CD>
__device__ __forceinline__ void angular_site_grad(const Harm<LR::L1>& h, float ux, float uy, CD&& cd,
                                                  float& dEdrho, float& gx, float& gy, float& gz) {
    float E1 = 0.f, Gz = 0.f, X1 = 0.f, Y1 = 0.f;
#pragma unroll
    for (int mm = 0; mm <= LR::L1; ++mm) {
        float U[2] = {0.f, 0.f}, V[2] = {0.f, 0.f}, W[2] = {0.f, 0.f}, Dq[2] = {0.f, 0.f};
#pragma unroll
        for (int l = (mm > LR::L0 ? mm : LR::L0); l <= LR::L1; ++l) {
            const float ql = h.q(l, mm);
            const float qa = (mm + 1 <= l - 1) ? h.q(l - 1, mm + 1) : 0.f;
            const float qc = (mm <= l - 1) ? (float)(l + mm) * h.q(l - 1, mm) : 0.f;
#pragma unroll
            for (int part = 0; part < (mm ? 2 : 1); ++part) {
                const int o = lm_slot(l, mm, part) - LR::OFF;
                float c, d;
                cd(o, c, d);
                if constexpr (PER_M) {
                    Dq[part] += d * ql;  // sum_l d_lm Q_l^m; times C_m or S_m below
                } else {
                    dEdrho += d * (mm == 0 ? ql : ql * (part ? h.S[mm] : h.C[mm]));
                }
                if (mm + 1 <= l - 1) U[part] += qa * c;
                if (mm <= l - 1) V[part] += qc * c;
                if (mm >= 1) W[part] += ql * c;
            }
        }
        if constexpr (PER_M) dEdrho += mm == 0 ? Dq[0] : h.C[mm] * Dq[0] + h.S[mm] * Dq[1];
        if (mm == 0) {
            E1 += U[0];
            Gz += V[0];
        } else {
            E1 += h.C[mm] * U[0] + h.S[mm] * U[1];
            Gz += h.C[mm] * V[0] + h.S[mm] * V[1];
            const float fm = (float)mm;
            X1 += fm * (h.C[mm - 1] * W[0] + h.S[mm - 1] * W[1]);
            Y1 += fm * (h.C[mm - 1] * W[1] - h.S[mm - 1] * W[0]);
        }
    }
    gx = ux * E1 + X1;
    gy = uy * E1 + Y1;
    gz = Gz;
}

// dE/dr_a = dE/drho_a u_a / r_nep + (grad - (u_a . grad) u_a) / |r_a|, where u_a . grad =
// sum_lm l c_lm Z_lm(u_a) (Euler's theorem on the degree-l terms of the homogeneous extension).
__device__ __forceinline__ void site_dEdr(const SiteGeo& s, float dEdrho, float inv_rn, float gx, float gy, float gz,
                                          float& ox, float& oy, float& oz) {
    const float proj = s.ux * gx + s.uy * gy + s.uz * gz;
    const float rad = dEdrho * inv_rn;
    ox = rad * s.ux + s.ir * (gx - proj * s.ux);
    oy = rad * s.uy + s.ir * (gy - proj * s.uy);
    oz = rad * s.uz + s.ir * (gz - proj * s.uz);
}

// Per-pair result of the adjoint.
struct PairOut {
    float fx, fy, fz;     // F on i (F on j is -F)
    float tix, tiy, tiz;  // torque on i about X_i
    float tjx, tjy, tjz;  // torque on j about X_j
};

// Per-body sums of dE/dr_a, arm x dE/dr_a and arms (body 0 = i, 1 = j), summed over passes;
// the arm sums are the pair's s_i and s_j (PairIn).
struct AdjAcc {
    float G0x = 0.f, G0y = 0.f, G0z = 0.f, T0x = 0.f, T0y = 0.f, T0z = 0.f, A0x = 0.f, A0y = 0.f, A0z = 0.f;
    float G1x = 0.f, G1y = 0.f, G1z = 0.f, T1x = 0.f, T1y = 0.f, T1z = 0.f, A1x = 0.f, A1y = 0.f, A1z = 0.f;
};

// Adjoint.  dG = dE/dG = [dE/dg_rad | dE/dg_ang] (already in energy units).  A is
// the forward accumulator, overwritten by B = dE/dA = 2 K_lm^2 dE/dg_ang[n,l] A.
// Per site a (r_a relative to the centroid, u_a = r_a/|r_a|), with
//   c_lm(a) = sum_n B_nlm f_n(a),   d_lm(a) = sum_n B_nlm f_n'(rho_a):
//   dE/drho_a = sum_n dE/dg_rad[n] f_n'(rho_a) + sum_lm d_lm Z_lm(u_a)
//   dE/dr_a   = dE/drho_a u_a / r_nep
//             + (1/|r_a|) [ grad sum_lm c_lm Z_lm (u_a) - u_a sum_lm l c_lm Z_lm(u_a) ]
// where Z_lm is extended to the homogeneous polynomial of degree l (angular_site_grad).
// Finally dE/dp_a = dE/dr_a - mean_b dE/dr_b through the centroid (adj_finish).
// One pass handles the l of range LR; the radial terms and the arms go with RAD.
template <class Cfg, class LR, bool RAD>
__device__ __forceinline__ void desc_adjoint(const nepx_pair_model& m, const PairIn& p,
                                             const float (&dG)[Cfg::D],
                                             float (&A)[Cfg::NANG + 1][LR::N], AdjAcc& acc) {
    constexpr K2Tab<Cfg::LMAX> k2{};
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n) {
#pragma unroll
        for (int l = LR::L0; l <= LR::L1; ++l) {
            const float w = 2.f * dG[Cfg::NRAD + 1 + n * Cfg::LMAX + (l - 1)];
            A[n][lm_slot(l, 0, 0) - LR::OFF] *= w;
#pragma unroll
            for (int mm = 1; mm <= l; ++mm) {
                const float wk = w * k2(l, mm);
                A[n][lm_slot(l, mm, 0) - LR::OFF] *= wk;
                A[n][lm_slot(l, mm, 1) - LR::OFF] *= wk;
            }
        }
    }
    const float inv_rn = 1.f / m.r_nep;
    float gx_s = 0.f, gy_s = 0.f, gz_s = 0.f, tx_s = 0.f, ty_s = 0.f, tz_s = 0.f;
    if constexpr (RAD) {
        acc.A0x = p.six; acc.A0y = p.siy; acc.A0z = p.siz;
        acc.A1x = p.sjx; acc.A1y = p.sjy; acc.A1z = p.sjz;
    }
    PairSites<Cfg>{m, p, inv_rn}.visit(
        [&](const SiteGeo& s, float) {
            float fa[Cfg::NANG + 1], da[Cfg::NANG + 1];
            float dEdrho = 0.f;
            radial_derivs<RAD ? Cfg::NRAD : Cfg::NANG>(s.rho, [&](int n, float f, float df) {
                if constexpr (RAD) dEdrho += dG[n] * df;
                if (n <= Cfg::NANG) {
                    fa[n] = f;
                    da[n] = df;
                }
            });
            Harm<LR::L1> h;
            h.eval(s.ux, s.uy, s.uz);
            float gx, gy, gz;
            angular_site_grad<LR>(h, s.ux, s.uy,
                                  [&](int o, float& c, float& d) {
                                      c = 0.f;
                                      d = 0.f;
#pragma unroll
                                      for (int n = 0; n <= Cfg::NANG; ++n) {
                                          c += A[n][o] * fa[n];
                                          d += A[n][o] * da[n];
                                      }
                                  },
                                  dEdrho, gx, gy, gz);
            float ga_x, ga_y, ga_z;
            site_dEdr(s, dEdrho, inv_rn, gx, gy, gz, ga_x, ga_y, ga_z);
            gx_s += ga_x;
            gy_s += ga_y;
            gz_s += ga_z;
            tx_s += s.ay * ga_z - s.az * ga_y;
            ty_s += s.az * ga_x - s.ax * ga_z;
            tz_s += s.ax * ga_y - s.ay * ga_x;
        },
        [&](int body) {
            if (body == 0) {
                acc.G0x += gx_s; acc.G0y += gy_s; acc.G0z += gz_s;
                acc.T0x += tx_s; acc.T0y += ty_s; acc.T0z += tz_s;
            } else {
                acc.G1x += gx_s; acc.G1y += gy_s; acc.G1z += gz_s;
                acc.T1x += tx_s; acc.T1y += ty_s; acc.T1z += tz_s;
            }
            gx_s = gy_s = gz_s = tx_s = ty_s = tz_s = 0.f;
        });
}

// dE/dp_a = g_a - gbar ; F_i = -sum_{a in i} dE/dp_a ; tau = -sum arm x dE/dp
template <class Cfg>
__device__ __forceinline__ PairOut adj_finish(const AdjAcc& a) {
    const float inv_nd = 1.f / (float)Cfg::ND;
    const float bx = (a.G0x + a.G1x) * inv_nd, by = (a.G0y + a.G1y) * inv_nd, bz = (a.G0z + a.G1z) * inv_nd;
    PairOut o;
    o.fx = -(a.G0x - (float)Cfg::NS * bx);
    o.fy = -(a.G0y - (float)Cfg::NS * by);
    o.fz = -(a.G0z - (float)Cfg::NS * bz);
    o.tix = -(a.T0x - (a.A0y * bz - a.A0z * by));
    o.tiy = -(a.T0y - (a.A0z * bx - a.A0x * bz));
    o.tiz = -(a.T0z - (a.A0x * by - a.A0y * bx));
    o.tjx = -(a.T1x - (a.A1y * bz - a.A1z * by));
    o.tjy = -(a.T1y - (a.A1z * bx - a.A1x * bz));
    o.tjz = -(a.T1z - (a.A1x * by - a.A1y * bx));
    return o;
}

}  // namespace nepx
