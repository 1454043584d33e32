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
#include "nepx_common.h"

namespace nepx {

template <int LMAX>
struct HarmTables {
    static constexpr int NLM = LMAX * (LMAX + 2);  // l = 1..LMAX
    static constexpr int NQ = (LMAX + 1) * (LMAX + 2) / 2;
    __host__ __device__ static constexpr int qi(int l, int m) { return l * (l + 1) / 2 + m; }
};

// K_lm = sqrt(2 (l-m)!/(l+m)!) for m > 0 (1 for m = 0), evaluated at compile
// time; index l*(LMAX+1)+m.  Used through a constexpr local so every lookup
// folds to an immediate once the (l, m) loops are unrolled.
template <int LMAX>
struct KTab {
    float v[(LMAX + 1) * (LMAX + 1)];
    constexpr KTab() : v() {
        for (int l = 0; l <= LMAX; ++l)
            for (int m = 0; m <= l; ++m)
                v[l * (LMAX + 1) + m] = m == 0 ? 1.0f : (float)csqrt(2.0 * cfact(l - m) / cfact(l + m));
    }
    __host__ __device__ constexpr float operator()(int l, int m) const { return v[l * (LMAX + 1) + m]; }
};

// Q_l^m(z) at |u| = 1, cos/sin of m phi times sin^m theta, and Yhat.
template <int LMAX>
struct Harm {
    using T = HarmTables<LMAX>;
    float Q[T::NQ];
    float C[LMAX + 1], S[LMAX + 1];

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
    // Yhat for l >= 1, index l*l-1 + {0: m=0, 2m-1: cos m, 2m: sin m}
    __device__ __forceinline__ void yhat(float* Y) const {
        constexpr KTab<LMAX> kfac{};
#pragma unroll
        for (int l = 1; l <= LMAX; ++l) {
            const int o = l * l - 1;
            Y[o] = Q[T::qi(l, 0)];
#pragma unroll
            for (int m = 1; m <= l; ++m) {
                const float kq = kfac(l, m) * Q[T::qi(l, m)];
                Y[o + 2 * m - 1] = kq * C[m];
                Y[o + 2 * m] = kq * S[m];
            }
        }
    }
};

template <int NS_, int NRAD_, int NANG_, int LMAX_>
struct DescCfg {
    static constexpr int NS = NS_, NRAD = NRAD_, NANG = NANG_, LMAX = LMAX_;
    static constexpr int ND = 2 * NS;
    static constexpr int NLM = LMAX * (LMAX + 2);
    static constexpr int D = (NRAD + 1) + (NANG + 1) * LMAX;
    static_assert(NANG <= NRAD, "nang must not exceed nrad");
};

// One pair in the frame centred on particle i: orientations, t = X_j - X_i,
// and the centroid of the 2*Ns sites,
//   c = (R(q_i) sum_k b_k + R(q_j) sum_k b_k + Ns t) / (2 Ns).
struct PairIn {
    float4 qi, qj;
    float tx, ty, tz;
    float cx, cy, cz;
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
    float ax, ay, az, bx, by, bz;
    mat_vec(quat_matrix(qi.x, qi.y, qi.z, qi.w), sx, sy, sz, ax, ay, az);
    mat_vec(quat_matrix(qj.x, qj.y, qj.z, qj.w), sx, sy, sz, bx, by, bz);
    const float inv = 1.f / (float)Cfg::ND;
    p.cx = (ax + bx) * inv + 0.5f * tx;
    p.cy = (ay + by) * inv + 0.5f * ty;
    p.cz = (az + bz) * inv + 0.5f * tz;
    return p;
}

// Site geometry shared by forward and adjoint.
struct SiteGeo {
    float ax, ay, az;  // lever arm about the owning body's centre: R b_k
    float ux, uy, uz;  // unit vector from the centroid
    float r, ir, rho;  // |r_a|, 1/|r_a|, |r_a|/r_nep
};

__device__ __forceinline__ SiteGeo site_geo(const nepx_pair_model& m, const Mat3& R, int k, float ox, float oy,
                                            float oz, float inv_rn) {
    SiteGeo s;
    mat_vec(R, m.sites[3 * k], m.sites[3 * k + 1], m.sites[3 * k + 2], s.ax, s.ay, s.az);
    const float rx = s.ax + ox, ry = s.ay + oy, rz = s.az + oz;
    s.r = sqrtf(rx * rx + ry * ry + rz * rz);
    s.ir = 1.f / s.r;
    s.ux = rx * s.ir;
    s.uy = ry * s.ir;
    s.uz = rz * s.ir;
    s.rho = s.r * inv_rn;
    return s;
}

// Forward accumulators: g_rad[n] and A[n][lm] = sum_a f_n(a) Yhat_lm(u_a).
template <class Cfg>
__device__ __forceinline__ void desc_accumulate(const nepx_pair_model& m, const PairIn& p,
                                                float (&grad)[Cfg::NRAD + 1],
                                                float (&A)[Cfg::NANG + 1][Cfg::NLM]) {
#pragma unroll
    for (int n = 0; n <= Cfg::NRAD; ++n) grad[n] = 0.f;
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n)
#pragma unroll
        for (int k = 0; k < Cfg::NLM; ++k) A[n][k] = 0.f;
    const float inv_rn = 1.f / m.r_nep;
#pragma unroll 1
    for (int body = 0; body < 2; ++body) {
        const float4 q = body ? p.qj : p.qi;
        const Mat3 R = quat_matrix(q.x, q.y, q.z, q.w);
        const float ox = (body ? p.tx : 0.f) - p.cx, oy = (body ? p.ty : 0.f) - p.cy,
                    oz = (body ? p.tz : 0.f) - p.cz;
#pragma unroll 1
        for (int k = 0; k < Cfg::NS; ++k) {
            const SiteGeo s = site_geo(m, R, k, ox, oy, oz, inv_rn);
            const float x = 2.f * (s.rho - 1.f) * (s.rho - 1.f) - 1.f;
            const float fc = 0.5f * (1.f + cospif(s.rho));
            float Y[Cfg::NLM];
            {
                Harm<Cfg::LMAX> h;
                h.eval(s.ux, s.uy, s.uz);
                h.yhat(Y);
            }
            float t0 = 1.f, t1 = x;
#pragma unroll
            for (int n = 0; n <= Cfg::NRAD; ++n) {
                float tn;
                if (n == 0) tn = 1.f;
                else if (n == 1) tn = x;
                else {
                    tn = 2.f * x * t1 - t0;
                    t0 = t1;
                    t1 = tn;
                }
                const float f = 0.5f * (tn + 1.f) * fc;
                grad[n] += f;
                if (n <= Cfg::NANG) {
#pragma unroll
                    for (int kk = 0; kk < Cfg::NLM; ++kk) A[n][kk] += f * Y[kk];
                }
            }
        }
    }
}

template <class Cfg>
__device__ __forceinline__ void desc_finish(const float (&grad)[Cfg::NRAD + 1],
                                            const float (&A)[Cfg::NANG + 1][Cfg::NLM],
                                            float (&G)[Cfg::D]) {
#pragma unroll
    for (int n = 0; n <= Cfg::NRAD; ++n) G[n] = grad[n];
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n) {
#pragma unroll
        for (int l = 1; l <= Cfg::LMAX; ++l) {
            float s = 0.f;
#pragma unroll
            for (int k = l * l - 1; k < (l + 1) * (l + 1) - 1; ++k) s += A[n][k] * A[n][k];
            G[Cfg::NRAD + 1 + n * Cfg::LMAX + (l - 1)] = s;
        }
    }
}

// Per-pair result of the adjoint.
struct PairOut {
    float fx, fy, fz;     // F on i (F on j is -F)
    float tix, tiy, tiz;  // torque on i about X_i
    float tjx, tjy, tjz;  // torque on j about X_j
};

// Adjoint.  dG = dE/dG = [dE/dg_rad | dE/dg_ang] (already in energy units).  A is
// the forward accumulator, overwritten by B = dE/dA = 2 dE/dg_ang[n,l] A.
// Per site a (r_a relative to the centroid, u_a = r_a/|r_a|):
//   dE/df_n(a) = dgr[n] + sum_lm B_nlm Yhat_lm(u_a)            (n <= nang)
//   c_lm(a)    = sum_n B_nlm f_n(a)
//   dE/dr_a    = (sum_n dE/df_n f_n'(rho)) u_a / r_nep
//              + (1/|r_a|) [ sum_lm c_lm grad S_lm(u_a) - u_a sum_l l sum_m c_lm Yhat_lm(u_a) ]
// (S_lm: the homogeneous polynomial with S_lm(u) = Yhat_lm(u) on |u| = 1), and
// dE/dp_a = dE/dr_a - mean_b dE/dr_b through the centroid.
template <class Cfg>
__device__ __forceinline__ PairOut desc_adjoint(const nepx_pair_model& m, const PairIn& p,
                                                const float (&dG)[Cfg::D],
                                                float (&A)[Cfg::NANG + 1][Cfg::NLM]) {
#pragma unroll
    for (int n = 0; n <= Cfg::NANG; ++n) {
#pragma unroll
        for (int l = 1; l <= Cfg::LMAX; ++l) {
            const float w = 2.f * dG[Cfg::NRAD + 1 + n * Cfg::LMAX + (l - 1)];
#pragma unroll
            for (int k = l * l - 1; k < (l + 1) * (l + 1) - 1; ++k) A[n][k] *= w;
        }
    }
    using HT = HarmTables<Cfg::LMAX>;
    constexpr KTab<Cfg::LMAX> kfac{};
    const float inv_rn = 1.f / m.r_nep;
    // per-body sums of dE/dr_a, arm x dE/dr_a and arms (body 0 = i, 1 = j)
    float G0x = 0.f, G0y = 0.f, G0z = 0.f, T0x = 0.f, T0y = 0.f, T0z = 0.f, A0x = 0.f, A0y = 0.f, A0z = 0.f;
    float G1x = 0.f, G1y = 0.f, G1z = 0.f, T1x = 0.f, T1y = 0.f, T1z = 0.f, A1x = 0.f, A1y = 0.f, A1z = 0.f;
#pragma unroll 1
    for (int body = 0; body < 2; ++body) {
        const float4 q = body ? p.qj : p.qi;
        const Mat3 R = quat_matrix(q.x, q.y, q.z, q.w);
        const float ox = (body ? p.tx : 0.f) - p.cx, oy = (body ? p.ty : 0.f) - p.cy,
                    oz = (body ? p.tz : 0.f) - p.cz;
        float gx_s = 0.f, gy_s = 0.f, gz_s = 0.f, tx_s = 0.f, ty_s = 0.f, tz_s = 0.f;
        float ax_s = 0.f, ay_s = 0.f, az_s = 0.f;
#pragma unroll 1
        for (int k = 0; k < Cfg::NS; ++k) {
            const SiteGeo s = site_geo(m, R, k, ox, oy, oz, inv_rn);
            const float xm = s.rho - 1.f;
            const float x = 2.f * xm * xm - 1.f;
            const float dx = 4.f * xm;
            float sp, cp;
            sincospif(s.rho, &sp, &cp);
            const float fc = 0.5f * (1.f + cp);
            const float dfc = -0.5f * (float)M_PI * sp;
            float c[Cfg::NLM];
            float dEdrho = 0.f, proj = 0.f;
            {
                float Y[Cfg::NLM];
                {
                    Harm<Cfg::LMAX> h;
                    h.eval(s.ux, s.uy, s.uz);
                    h.yhat(Y);
                }
#pragma unroll
                for (int kk = 0; kk < Cfg::NLM; ++kk) c[kk] = 0.f;
                float t0 = 1.f, t1 = x, d0 = 0.f, d1 = 1.f;
#pragma unroll
                for (int n = 0; n <= Cfg::NRAD; ++n) {
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
                    const float f = 0.5f * (tn + 1.f) * fc;
                    const float df = 0.5f * dn * dx * fc + 0.5f * (tn + 1.f) * dfc;
                    float sn = dG[n];
                    if (n <= Cfg::NANG) {
#pragma unroll
                        for (int kk = 0; kk < Cfg::NLM; ++kk) {
                            sn += A[n][kk] * Y[kk];
                            c[kk] += A[n][kk] * f;
                        }
                    }
                    dEdrho += sn * df;
                }
#pragma unroll
                for (int l = 1; l <= Cfg::LMAX; ++l)
#pragma unroll
                    for (int kk = l * l - 1; kk < (l + 1) * (l + 1) - 1; ++kk) proj += (float)l * c[kk] * Y[kk];
            }
            // gradient of sum_lm c_lm S_lm at u (harmonic recurrences recomputed)
            Harm<Cfg::LMAX> h;
            h.eval(s.ux, s.uy, s.uz);
            float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
            for (int l = 1; l <= Cfg::LMAX; ++l) {
                const int o = l * l - 1;
                {  // m = 0
                    const float cc = c[o];
                    const float qa = (l - 1 >= 1) ? h.Q[HT::qi(l - 1, 1)] : 0.f;
                    gx += cc * s.ux * qa;
                    gy += cc * s.uy * qa;
                    gz += cc * (float)l * h.Q[HT::qi(l - 1, 0)];
                }
#pragma unroll
                for (int mm = 1; mm <= l; ++mm) {
                    const float K = kfac(l, mm);
                    const float cc = c[o + 2 * mm - 1] * K, cs = c[o + 2 * mm] * K;
                    const float qa = (mm + 1 <= l - 1) ? h.Q[HT::qi(l - 1, mm + 1)] : 0.f;
                    const float qb = h.Q[HT::qi(l, mm)];
                    const float qc = (mm <= l - 1) ? h.Q[HT::qi(l - 1, mm)] : 0.f;
                    const float e = cc * h.C[mm] + cs * h.S[mm];
                    gx += s.ux * qa * e + (float)mm * qb * (cc * h.C[mm - 1] + cs * h.S[mm - 1]);
                    gy += s.uy * qa * e + (float)mm * qb * (cs * h.C[mm - 1] - cc * h.S[mm - 1]);
                    gz += (float)(l + mm) * qc * e;
                }
            }
            const float rad = dEdrho * inv_rn;
            const float ga_x = rad * s.ux + s.ir * (gx - proj * s.ux);
            const float ga_y = rad * s.uy + s.ir * (gy - proj * s.uy);
            const float ga_z = rad * s.uz + s.ir * (gz - proj * s.uz);
            gx_s += ga_x;
            gy_s += ga_y;
            gz_s += ga_z;
            tx_s += s.ay * ga_z - s.az * ga_y;
            ty_s += s.az * ga_x - s.ax * ga_z;
            tz_s += s.ax * ga_y - s.ay * ga_x;
            ax_s += s.ax;
            ay_s += s.ay;
            az_s += s.az;
        }
        if (body == 0) {
            G0x = gx_s; G0y = gy_s; G0z = gz_s; T0x = tx_s; T0y = ty_s; T0z = tz_s; A0x = ax_s; A0y = ay_s; A0z = az_s;
        } else {
            G1x = gx_s; G1y = gy_s; G1z = gz_s; T1x = tx_s; T1y = ty_s; T1z = tz_s; A1x = ax_s; A1y = ay_s; A1z = az_s;
        }
    }
    // dE/dp_a = g_a - gbar ; F_i = -sum_{a in i} dE/dp_a ; tau = -sum arm x dE/dp
    const float inv_nd = 1.f / (float)Cfg::ND;
    const float bx = (G0x + G1x) * inv_nd, by = (G0y + G1y) * inv_nd, bz = (G0z + G1z) * inv_nd;
    PairOut o;
    o.fx = -(G0x - (float)Cfg::NS * bx);
    o.fy = -(G0y - (float)Cfg::NS * by);
    o.fz = -(G0z - (float)Cfg::NS * bz);
    o.tix = -(T0x - (A0y * bz - A0z * by));
    o.tiy = -(T0y - (A0z * bx - A0x * bz));
    o.tiz = -(T0z - (A0x * by - A0y * bx));
    o.tjx = -(T1x - (A1y * bz - A1z * by));
    o.tjy = -(T1y - (A1z * bx - A1x * bz));
    o.tjz = -(T1z - (A1x * by - A1y * bx));
    return o;
}

__device__ __forceinline__ float clipf(float v, float c) {
    return c > 0.f ? fminf(fmaxf(v, -c), c) : v;
}

}  // namespace nepx
