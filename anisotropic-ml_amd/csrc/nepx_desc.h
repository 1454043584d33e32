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
    // K_lm = sqrt(2 (l-m)!/(l+m)!) for m > 0
    __host__ __device__ static constexpr float kfac(int l, int m) {
        return m == 0 ? 1.0f : (float)csqrt(2.0 * cfact(l - m) / cfact(l + m));
    }
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
#pragma unroll
        for (int l = 1; l <= LMAX; ++l) {
            const int o = l * l - 1;
            Y[o] = Q[T::qi(l, 0)];
#pragma unroll
            for (int m = 1; m <= l; ++m) {
                const float kq = T::kfac(l, m) * Q[T::qi(l, m)];
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

// Geometry of one pair in the frame centred on particle i.
struct PairGeom {
    Mat3 Ri, Rj;
    float tx, ty, tz;
    float cx, cy, cz;  // centroid of the 2*Ns sites
};

template <class Cfg>
__device__ __forceinline__ PairGeom pair_geom(const nepx_pair_model& m, float4 qi, float4 qj,
                                              float tx, float ty, float tz) {
    PairGeom g;
    g.Ri = quat_matrix(qi.x, qi.y, qi.z, qi.w);
    g.Rj = quat_matrix(qj.x, qj.y, qj.z, qj.w);
    g.tx = tx;
    g.ty = ty;
    g.tz = tz;
    // centroid = (R_i sum b + R_j sum b + Ns t) / (2 Ns), accumulated site by
    // site in the same order as the cloud.
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int a = 0; a < Cfg::NS; ++a) {
        float px, py, pz;
        mat_vec(g.Ri, m.sites[3 * a], m.sites[3 * a + 1], m.sites[3 * a + 2], px, py, pz);
        sx += px;
        sy += py;
        sz += pz;
    }
    for (int a = 0; a < Cfg::NS; ++a) {
        float px, py, pz;
        mat_vec(g.Rj, m.sites[3 * a], m.sites[3 * a + 1], m.sites[3 * a + 2], px, py, pz);
        sx += px + tx;
        sy += py + ty;
        sz += pz + tz;
    }
    const float inv = 1.f / (float)Cfg::ND;
    g.cx = sx * inv;
    g.cy = sy * inv;
    g.cz = sz * inv;
    return g;
}

// Site a of the cloud (a < Ns: body i, else body j), relative to the centroid.
// Also returns the lever arm about the owning body's centre.
template <class Cfg>
__device__ __forceinline__ void site_rel(const nepx_pair_model& m, const PairGeom& g, int a,
                                         float& rx, float& ry, float& rz, float& ax, float& ay,
                                         float& az) {
    const int k = a < Cfg::NS ? a : a - Cfg::NS;
    const float bx = m.sites[3 * k], by = m.sites[3 * k + 1], bz = m.sites[3 * k + 2];
    if (a < Cfg::NS) {
        mat_vec(g.Ri, bx, by, bz, ax, ay, az);
        rx = ax - g.cx;
        ry = ay - g.cy;
        rz = az - g.cz;
    } else {
        mat_vec(g.Rj, bx, by, bz, ax, ay, az);
        rx = (ax + g.tx) - g.cx;
        ry = (ay + g.ty) - g.cy;
        rz = (az + g.tz) - g.cz;
    }
}

// Forward: A[n][lm] and g_rad; returns G (D floats).
template <class Cfg>
__device__ __forceinline__ void desc_accumulate(const nepx_pair_model& m, const PairGeom& g,
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
    for (int a = 0; a < Cfg::ND; ++a) {
        float rx, ry, rz, ax, ay, az;
        site_rel<Cfg>(m, g, a, rx, ry, rz, ax, ay, az);
        const float r = sqrtf(rx * rx + ry * ry + rz * rz);
        const float ir = 1.f / r;
        const float ux = rx * ir, uy = ry * ir, uz = rz * ir;
        const float rho = r * inv_rn;
        const float x = 2.f * (rho - 1.f) * (rho - 1.f) - 1.f;
        const float fc = 0.5f * (1.f + cospif(rho));
        Harm<Cfg::LMAX> h;
        h.eval(ux, uy, uz);
        float Y[Cfg::NLM];
        h.yhat(Y);
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
                for (int k = 0; k < Cfg::NLM; ++k) A[n][k] += f * Y[k];
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

// Adjoint: given dE/dG (already scaled to energy units), per-pair force and
// torques.  A is the forward accumulator (overwritten by dE/dA).
template <class Cfg>
__device__ __forceinline__ PairOut desc_adjoint(const nepx_pair_model& m, const PairGeom& g,
                                                const float (&dG)[Cfg::D],
                                                float (&A)[Cfg::NANG + 1][Cfg::NLM]) {
    // dE/dA_nlm = 2 dE/dg_ang[n,l] A_nlm
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
    const float inv_rn = 1.f / m.r_nep;
    float Gix = 0.f, Giy = 0.f, Giz = 0.f, Gjx = 0.f, Gjy = 0.f, Gjz = 0.f;
    float Tix = 0.f, Tiy = 0.f, Tiz = 0.f, Tjx = 0.f, Tjy = 0.f, Tjz = 0.f;
    float Aix = 0.f, Aiy = 0.f, Aiz = 0.f, Ajx = 0.f, Ajy = 0.f, Ajz = 0.f;
#pragma unroll 1
    for (int a = 0; a < Cfg::ND; ++a) {
        float rx, ry, rz, ax, ay, az;
        site_rel<Cfg>(m, g, a, rx, ry, rz, ax, ay, az);
        const float r = sqrtf(rx * rx + ry * ry + rz * rz);
        const float ir = 1.f / r;
        const float ux = rx * ir, uy = ry * ir, uz = rz * ir;
        const float rho = r * inv_rn;
        const float xm = rho - 1.f;
        const float x = 2.f * xm * xm - 1.f;
        const float dx = 4.f * xm;
        float sp, cp;
        sincospif(rho, &sp, &cp);
        const float fc = 0.5f * (1.f + cp);
        const float dfc = -0.5f * (float)M_PI * sp;
        Harm<Cfg::LMAX> h;
        h.eval(ux, uy, uz);
        float Y[Cfg::NLM];
        h.yhat(Y);
        float c[Cfg::NLM];
#pragma unroll
        for (int k = 0; k < Cfg::NLM; ++k) c[k] = 0.f;
        float dEdrho = 0.f;
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
            float s = dG[n];
            if (n <= Cfg::NANG) {
#pragma unroll
                for (int k = 0; k < Cfg::NLM; ++k) {
                    s += A[n][k] * Y[k];
                    c[k] += A[n][k] * f;
                }
            }
            dEdrho += s * df;
        }
        // gradient of sum_lm c_lm Yhat_lm(r/|r|) with respect to r
        float gx = 0.f, gy = 0.f, gz = 0.f, proj = 0.f;
#pragma unroll
        for (int l = 1; l <= Cfg::LMAX; ++l) {
            const int o = l * l - 1;
            {  // m = 0
                const float cc = c[o];
                const float qa = (l - 1 >= 1) ? h.Q[HT::qi(l - 1, 1)] : 0.f;
                gx += cc * ux * qa;
                gy += cc * uy * qa;
                gz += cc * (float)l * h.Q[HT::qi(l - 1, 0)];
                proj += (float)l * cc * Y[o];
            }
#pragma unroll
            for (int mm = 1; mm <= l; ++mm) {
                const float K = HT::kfac(l, mm);
                const float cc = c[o + 2 * mm - 1] * K, cs = c[o + 2 * mm] * K;
                const float qa = (mm + 1 <= l - 1) ? h.Q[HT::qi(l - 1, mm + 1)] : 0.f;
                const float qb = h.Q[HT::qi(l, mm)];
                const float qc = (mm <= l - 1) ? h.Q[HT::qi(l - 1, mm)] : 0.f;
                const float e = cc * h.C[mm] + cs * h.S[mm];
                gx += ux * qa * e + (float)mm * qb * (cc * h.C[mm - 1] + cs * h.S[mm - 1]);
                gy += uy * qa * e + (float)mm * qb * (cs * h.C[mm - 1] - cc * h.S[mm - 1]);
                gz += (float)(l + mm) * qc * e;
                proj += (float)l * (c[o + 2 * mm - 1] * Y[o + 2 * mm - 1] + c[o + 2 * mm] * Y[o + 2 * mm]);
            }
        }
        const float rad = dEdrho * inv_rn;
        const float ga_x = rad * ux + ir * (gx - proj * ux);
        const float ga_y = rad * uy + ir * (gy - proj * uy);
        const float ga_z = rad * uz + ir * (gz - proj * uz);
        if (a < Cfg::NS) {
            Gix += ga_x; Giy += ga_y; Giz += ga_z;
            Tix += ay * ga_z - az * ga_y;
            Tiy += az * ga_x - ax * ga_z;
            Tiz += ax * ga_y - ay * ga_x;
            Aix += ax; Aiy += ay; Aiz += az;
        } else {
            Gjx += ga_x; Gjy += ga_y; Gjz += ga_z;
            Tjx += ay * ga_z - az * ga_y;
            Tjy += az * ga_x - ax * ga_z;
            Tjz += ax * ga_y - ay * ga_x;
            Ajx += ax; Ajy += ay; Ajz += az;
        }
    }
    // dE/dp_a = g_a - gbar ; F_i = -sum_{a in i} dE/dp_a ; tau = -sum arm x dE/dp
    const float inv_nd = 1.f / (float)Cfg::ND;
    const float bx = (Gix + Gjx) * inv_nd, by = (Giy + Gjy) * inv_nd, bz = (Giz + Gjz) * inv_nd;
    PairOut o;
    o.fx = -(Gix - (float)Cfg::NS * bx);
    o.fy = -(Giy - (float)Cfg::NS * by);
    o.fz = -(Giz - (float)Cfg::NS * bz);
    o.tix = -(Tix - (Aiy * bz - Aiz * by));
    o.tiy = -(Tiy - (Aiz * bx - Aix * bz));
    o.tiz = -(Tiz - (Aix * by - Aiy * bx));
    o.tjx = -(Tjx - (Ajy * bz - Ajz * by));
    o.tjy = -(Tjy - (Ajz * bx - Ajx * bz));
    o.tjz = -(Tjz - (Ajx * by - Ajy * bx));
    return o;
}

__device__ __forceinline__ float clipf(float v, float c) {
    return c > 0.f ? fminf(fmaxf(v, -c), c) : v;
}

}  // namespace nepx
