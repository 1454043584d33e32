// FP64 rigid-body NVT integrator (two-step Nose-Hoover translation + NO_SQUISH
// rotation), per-system kinetic-energy reduction and thermostat update.
//
// Per step, after F and tau are known (nep_simulator/NVT.py:17-129):
//   v <- v*theta + dt/2 F/m ; dp = dt q (x) (0, R(q*) tau) ; p <- p e_r + dp   (step two)
//   v <- (v + dt/2 F/m) theta ; x <- wrap(x + v dt)                          (step one)
//   p <- (p + dp) e_r ; free rotor 3(dt/2) 2(dt/2) 1(dt) 2(dt/2) 3(dt/2) ; q <- q/|q|
// followed by K_t = sum m v^2/2, K_r = sum_k s_k^2/(2 I_k), s = q* (x) p / 2, and
// the thermostat update that the reference performs at the start of the next
// step (IntegrateStepTwoKernels.py:93-125), done here once per system at the
// tail of the step from the same inputs.
#include "nepx_common.h"

// No fused multiply-adds are formed from separate multiplies and adds in this file: every
// operation rounds as written, so the per-particle step gives the same bits in every kernel it is
// inlined into (k_nvt with the step part chosen at run time, k_nvt_one with it fixed), whatever
// basic blocks the compiler forms around it.
#pragma clang fp contract(off)

namespace nepx {

constexpr int kIB = 256;  // particles per integrator block (system ranges are block aligned)

struct Q4 {
    double w, x, y, z;
};

__device__ __forceinline__ Q4 qmul(const Q4& a, const Q4& b) {
    return Q4{a.w * b.w - (a.x * b.x + a.y * b.y + a.z * b.z),
              a.w * b.x + b.w * a.x + (a.y * b.z - a.z * b.y),
              a.w * b.y + b.w * a.y + (a.z * b.x - a.x * b.z),
              a.w * b.z + b.w * a.z + (a.x * b.y - a.y * b.x)};
}

// P_k v of the free-rotor sub-steps (IntegrateStepOneKernels.py:45-157)
__device__ __forceinline__ Q4 perm(int k, const Q4& v) {
    if (k == 1) return Q4{-v.x, v.w, v.z, -v.y};
    if (k == 2) return Q4{-v.y, -v.z, v.w, v.x};
    return Q4{-v.z, v.y, -v.x, v.w};
}

// sin and cos of a rotor angle.  The angles are dt times a rotation rate, so small: below
// 0.125 the Taylor series to x^13 (sin) and x^14 (cos) leaves a truncation error under 1e-25
// relative, far below FP64 rounding; larger angles take the library routine.
__device__ __forceinline__ void sincos_rotor(double x, double* s, double* c) {
    if (fabs(x) < 0.125) {
        const double x2 = x * x;
        *s = x * (1.0 + x2 * (-1.0 / 6.0 + x2 * (1.0 / 120.0 + x2 * (-1.0 / 5040.0 + x2 * (1.0 / 362880.0 +
                 x2 * (-1.0 / 39916800.0 + x2 * (1.0 / 6227020800.0)))))));
        *c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24.0 + x2 * (-1.0 / 720.0 + x2 * (1.0 / 40320.0 +
                 x2 * (-1.0 / 3628800.0 + x2 * (1.0 / 479001600.0 + x2 * (-1.0 / 87178291200.0)))))));
    } else {
        sincos(x, s, c);
    }
}

__device__ __forceinline__ void rotor(int k, double h, double Ik, Q4& p, Q4& q) {
    const Q4 pp = perm(k, p), qq = perm(k, q);
    const double phi = (0.25 / Ik) * (p.w * qq.w + p.x * qq.x + p.y * qq.y + p.z * qq.z);
    double s, c;
    sincos_rotor(h * phi, &s, &c);
    p = Q4{c * p.w + s * pp.w, c * p.x + s * pp.x, c * p.y + s * pp.y, c * p.z + s * pp.z};
    q = Q4{c * q.w + s * qq.w, c * q.x + s * qq.x, c * q.y + s * qq.y, c * q.z + s * qq.z};
}

__device__ __forceinline__ double rot_ke(const Q4& q, const Q4& p, const double* moi) {
    const Q4 s = qmul(Q4{q.w, -q.x, -q.y, -q.z}, p);
    const double sx = 0.5 * s.x, sy = 0.5 * s.y, sz = 0.5 * s.z;
    return 0.5 * (sx * sx / moi[0] + sy * sy / moi[1] + sz * sz / moi[2]);
}

// Maximum that keeps a NaN (fmax would drop it): a non-finite displacement marks
// the system as faulted in finish_system.
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || a > b) ? a : b; }

// Sums of a, b, c and the maximum of dmax (NaN kept) over the block, written to out[0..3] by
// thread 0.  Butterfly sums within each 64-lane wave (every lane ends with the same,
// order-fixed value: a + b == b + a exactly), then the four wave results in wave
// order: one barrier, and the same bits on every run.
__device__ void block_reduce4(double a, double b, double c, double dmax, double* out) {
    static_assert(kIB == 256, "four waves per block");
    __shared__ double red[4][kIB / 64];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
        dmax = nan_max(dmax, __shfl_xor(dmax, o, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
        red[3][w] = dmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        out[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        out[2] = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
        out[3] = nan_max(nan_max(red[3][0], red[3][1]), nan_max(red[3][2], red[3][3]));
    }
}

__device__ __forceinline__ double disp2(const nepx_state& st, int i, double L) {
    double dx = st.x[i] - st.xb[i], dy = st.y[i] - st.yb[i], dz = st.z[i] - st.zb[i];
    dx -= L * rint(dx / L);
    dy -= L * rint(dy / L);
    dz -= L * rint(dz / L);
    return dx * dx + dy * dy + dz * dz;
}

// Parts of the step (mode): 0 = whole step; 1 = kinetic energies only
// (initialisation); 2 = step two (half kicks, torque kick stored in dp);
// 3 = translation of step one; 4 = rotation of step one (uses dp);
// 5 = kinetic energies + displacement for the thermostat tail.
enum { M_STEP = 0, M_KE = 1, M_TWO = 2, M_TRANS = 3, M_ROT = 4, M_THERMO = 5 };

// One particle's part of the step (mode as above); kt, kr, u, d2 are its contributions to the
// energy and displacement sums (left at zero when they are not summed in this mode).
__device__ __forceinline__ void nvt_particle(nepx_state& st, int mode, int i, double& kt, double& kr, double& u,
                                             double& d2) {
    if (!(i < st.n && st.sys[i] >= 0 && st.sp[st.sys[i]].active)) return;
    const int sidx = st.sys[i];
    const nepx_sysparam& sp = st.sp[sidx];
    const nepx_thermo& th = st.th[sidx];
    const double dt = sp.dt, L = sp.L;
    const double m = st.mass[i];
    const double theta = th.theta, er = th.e_r;
    double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
    Q4 q{st.q0[i], st.q1[i], st.q2[i], st.q3[i]};
    Q4 p{st.p0[i], st.p1[i], st.p2[i], st.p3[i]};
    const double ax = st.fx[i] / m, ay = st.fy[i] / m, az = st.fz[i] / m;
    Q4 dp{0.0, 0.0, 0.0, 0.0};
    if (mode == M_STEP || mode == M_TWO) {
        // step two: translation and torque half kicks
        vx = vx * theta + 0.5 * dt * ax;
        vy = vy * theta + 0.5 * dt * ay;
        vz = vz * theta + 0.5 * dt * az;
        const Q4 qc{q.w, -q.x, -q.y, -q.z};
        const Q4 tb = qmul(qmul(qc, Q4{0.0, st.tx[i], st.ty[i], st.tz[i]}), q);  // R(q*) tau
        const Q4 dq = qmul(q, Q4{0.0, tb.x, tb.y, tb.z});
        dp = Q4{dt * dq.w, dt * dq.x, dt * dq.y, dt * dq.z};
        p = Q4{p.w * er + dp.w, p.x * er + dp.x, p.y * er + dp.y, p.z * er + dp.z};
        if (mode == M_TWO) {
            st.vx[i] = vx;
            st.vy[i] = vy;
            st.vz[i] = vz;
            st.p0[i] = p.w;
            st.p1[i] = p.x;
            st.p2[i] = p.y;
            st.p3[i] = p.z;
            st.dp[4 * (size_t)i] = dp.w;
            st.dp[4 * (size_t)i + 1] = dp.x;
            st.dp[4 * (size_t)i + 2] = dp.y;
            st.dp[4 * (size_t)i + 3] = dp.z;
        }
    }
    if (mode == M_STEP || mode == M_TRANS) {
        vx = (vx + 0.5 * dt * ax) * theta;
        vy = (vy + 0.5 * dt * ay) * theta;
        vz = (vz + 0.5 * dt * az) * theta;
        double x = st.x[i] + vx * dt, y = st.y[i] + vy * dt, z = st.z[i] + vz * dt;
        x = x > 0.5 * L ? x - L : x;
        x = x < -0.5 * L ? x + L : x;
        y = y > 0.5 * L ? y - L : y;
        y = y < -0.5 * L ? y + L : y;
        z = z > 0.5 * L ? z - L : z;
        z = z < -0.5 * L ? z + L : z;
        st.x[i] = x;
        st.y[i] = y;
        st.z[i] = z;
        st.vx[i] = vx;
        st.vy[i] = vy;
        st.vz[i] = vz;
    }
    if (mode == M_STEP || mode == M_ROT) {
        if (mode == M_ROT)
            dp = Q4{st.dp[4 * (size_t)i], st.dp[4 * (size_t)i + 1], st.dp[4 * (size_t)i + 2],
                    st.dp[4 * (size_t)i + 3]};
        p = Q4{(p.w + dp.w) * er, (p.x + dp.x) * er, (p.y + dp.y) * er, (p.z + dp.z) * er};
        rotor(3, 0.5 * dt, sp.moi[2], p, q);
        rotor(2, 0.5 * dt, sp.moi[1], p, q);
        rotor(1, dt, sp.moi[0], p, q);
        rotor(2, 0.5 * dt, sp.moi[1], p, q);
        rotor(3, 0.5 * dt, sp.moi[2], p, q);
        const double nq = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        q = Q4{q.w / nq, q.x / nq, q.y / nq, q.z / nq};
        st.q0[i] = q.w;
        st.q1[i] = q.x;
        st.q2[i] = q.y;
        st.q3[i] = q.z;
        st.p0[i] = p.w;
        st.p1[i] = p.x;
        st.p2[i] = p.y;
        st.p3[i] = p.z;
    }
    if (mode == M_STEP || mode == M_ROT || mode == M_KE)
        reinterpret_cast<float4*>(st.qf)[i] = make_float4((float)q.w, (float)q.x, (float)q.y, (float)q.z);
    if (mode == M_STEP || mode == M_THERMO || mode == M_KE) {
        kt = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        kr = rot_ke(q, p, sp.moi);
        u = st.upot ? st.upot[i] : 0.0;
        if (mode != M_KE) d2 = disp2(st, i, L);
    }
}

// With `flags` set and the list overflowed or the step frozen (flags[6]: the list
// went stale inside a multi-step graph), the step is skipped as a whole
// (uniformly: every thread reads the same flag) so the state stays at the last
// good step until the list is grown or rebuilt.  Each block writes its partial
// sums and k_reduce_thermo, a second launch, sums them per system (folding that
// into the last block of each system needs a device-scope fence in every block,
// which made this launch three times slower at 32,768 particles).
__global__ __launch_bounds__(kIB) void k_nvt(nepx_state st, int mode, const int32_t* __restrict__ flags) {
    if (flags && (flags[1] || flags[6])) return;  // list overflow, or a frozen step (stale list)
    double kt = 0.0, kr = 0.0, u = 0.0, d2 = 0.0;
    nvt_particle(st, mode, blockIdx.x * kIB + threadIdx.x, kt, kr, u, d2);
    if (mode == M_STEP || mode == M_THERMO || mode == M_KE)  // uniform per launch
        block_reduce4(kt, kr, u, d2, st.partials + 4 * (size_t)blockIdx.x);
}

// System s's sums out = (K_t, K_r, U, max displacement^2) -> thermostat update, list rebuild test,
// step count, fault stop; one thread.
__device__ void finish_system(nepx_state& st, int32_t* flags, int s, const double* out, int update,
                              int count_step) {
    nepx_sysparam& sp = st.sp[s];
    {
        nepx_thermo& th = st.th[s];
        const double Kt = out[0], Kr = out[1];
        th.Kt = Kt;
        th.Kr = Kr;
        th.U = out[2];
        th.max_disp2 = out[3];
        th.pending = 0.0;
        if (!(isfinite(Kt) && isfinite(Kr) && isfinite(out[2]) && isfinite(out[3]))) {
            // a non-finite state (a blown-up pair, an injected NaN): this system stops here and
            // every kernel skips it from the next launch on; the other systems are untouched
            th.fault = 1.0;
            sp.active = 0;
            return;
        }
        if (update) {
            const double dt = sp.dt, tau = sp.tau, kT = sp.kT;
            const double cfac = 0.5 * ((dt / tau) / tau);
            const double Tt = 2.0 * Kt / sp.ndof;
            const double xi_p = th.xi + cfac * (Tt / kT - 1.0);
            const double xi = xi_p + cfac * (Tt / kT - 1.0);
            th.eta += xi_p * dt;
            th.xi = xi;
            th.theta = exp(-0.5 * xi * dt);
            const double Tr = 2.0 * Kr / sp.rndof;
            const double xr_p = th.xi_r + cfac * (Tr / kT - 1.0);
            const double xr = xr_p + cfac * (Tr / kT - 1.0);
            th.eta_r += xr_p * dt;
            th.xi_r = xr;
            th.e_r = exp(-0.5 * dt * xr);
        } else {  // restored state: the factors follow from the saved xi, xi_r
            th.theta = exp(-0.5 * th.xi * sp.dt);
            th.e_r = exp(-0.5 * sp.dt * th.xi_r);
        }
        // the list is stale once a particle has moved half the skin; with an early fraction f > 0
        // (sp.pad[0]) a rebuild is requested at the next build point from f times that on
        const double half = 0.5 * sp.skin, early = sp.pad[0] * half;
        if (flags && out[3] > half * half) flags[0] = 1;
        else if (flags && early > 0.0 && out[3] > early * early) flags[7] = 1;
        if (count_step) {
            th.steps += 1.0;
            if (th.budget > 0.0 && th.steps >= th.budget) sp.active = 0;  // finished: frozen until harvested
        }
    }
}

// Fixed-order sums of system s's block partials, then finish_system; run by one whole block.
__device__ void reduce_system(nepx_state& st, int32_t* flags, int s, int update, int count_step) {
    const nepx_sysparam& sp = st.sp[s];
    double a = 0, b = 0, c = 0, d = 0;
    for (int k = sp.blk_start + threadIdx.x; k < sp.blk_end; k += kIB) {
        const double* pp = st.partials + 4 * (size_t)k;
        a += pp[0];
        b += pp[1];
        c += pp[2];
        d = nan_max(d, pp[3]);
    }
    __shared__ double out[4];
    block_reduce4(a, b, c, d, out);
    __syncthreads();
    if (threadIdx.x == 0) finish_system(st, flags, s, out, update, count_step);
}

// The whole step of one system of at most kOneBlock particles in a single block: k_nvt's
// per-particle work and k_reduce_thermo's sums in one launch (a small system's step is bound by
// launch latency).  The sums are taken in the two-launch order, bit for bit: each 256-particle
// chunk as block_reduce4 sums a block, then the chunk sums through the same 64-lane butterfly
// and wave combination that reduce_system applies to the block partials.
constexpr int kOneBlock = 1024;
__global__ __launch_bounds__(kOneBlock) void k_nvt_one(nepx_state st, int32_t* __restrict__ flags) {
    if (flags && (flags[1] || flags[6])) return;  // list overflow, or a frozen step (stale list)
    if (!st.sp[0].active) return;                 // uniform: the step is skipped as a whole
    double kt = 0.0, kr = 0.0, u = 0.0, d2 = 0.0;
    nvt_particle(st, M_STEP, threadIdx.x, kt, kr, u, d2);
    __shared__ double red[4][kOneBlock / 64];
    __shared__ double chunk[4][kOneBlock / kIB];
    for (int o = 32; o > 0; o >>= 1) {
        kt += __shfl_xor(kt, o, 64);
        kr += __shfl_xor(kr, o, 64);
        u += __shfl_xor(u, o, 64);
        d2 = nan_max(d2, __shfl_xor(d2, o, 64));
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][w] = kt;
        red[1][w] = kr;
        red[2][w] = u;
        red[3][w] = d2;
    }
    __syncthreads();
    const nepx_sysparam& sp = st.sp[0];
    // the system's 256-particle blocks (from block 0; at most four, as st.n <= kOneBlock)
    const int nchunk = (sp.blk_end < kOneBlock / kIB ? sp.blk_end : kOneBlock / kIB) - sp.blk_start;
    if (threadIdx.x < nchunk) {
        const int c = sp.blk_start + threadIdx.x, w0 = 4 * c;
        for (int k = 0; k < 3; ++k) chunk[k][threadIdx.x] = (red[k][w0] + red[k][w0 + 1]) + (red[k][w0 + 2] + red[k][w0 + 3]);
        chunk[3][threadIdx.x] = nan_max(nan_max(red[3][w0], red[3][w0 + 1]), nan_max(red[3][w0 + 2], red[3][w0 + 3]));
    }
    __syncthreads();
    if (w == 0) {
        // reduce_system: lane t loads partial t (0 past the end), butterflies, then its four-wave
        // combination with the other (all-zero) waves
        double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
        if (lane < nchunk) {
            a += chunk[0][lane];
            b += chunk[1][lane];
            c += chunk[2][lane];
            d = nan_max(d, chunk[3][lane]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_xor(a, o, 64);
            b += __shfl_xor(b, o, 64);
            c += __shfl_xor(c, o, 64);
            d = nan_max(d, __shfl_xor(d, o, 64));
        }
        if (lane == 0) {
            double out[4];
            out[0] = (a + 0.0) + (0.0 + 0.0);
            out[1] = (b + 0.0) + (0.0 + 0.0);
            out[2] = (c + 0.0) + (0.0 + 0.0);
            out[3] = nan_max(nan_max(d, 0.0), nan_max(0.0, 0.0));
            finish_system(st, flags, 0, out, 1, 1);
        }
    }
}

// One block per system: reduce_system after every k_nvt that sums energies.
__global__ __launch_bounds__(kIB) void k_reduce_thermo(nepx_state st, int32_t* flags, int update, int count_step) {
    const int s = blockIdx.x;
    if (!st.sp[s].active) return;  // uniform per block
    if (flags && (flags[1] || flags[6])) return;  // list overflow or frozen step: the step was skipped
    if (update == 2 && st.th[s].pending == 0.0) return;  // only newly admitted systems
    reduce_system(st, flags, s, update, count_step);
}

}  // namespace nepx

using namespace nepx;

static int check_nvt(const nepx_state* st) {
    NEPX_REQUIRE(st && st->n > 0 && st->n_sys > 0, NEPX_EINVAL, "empty state");
    NEPX_REQUIRE(st->x && st->vx && st->q0 && st->p0 && st->mass && st->fx && st->tx && st->sys && st->sp &&
                     st->th && st->partials && st->qf && st->xb,
                 NEPX_EINVAL, "null state buffer");
    return NEPX_OK;
}

extern "C" int nepx_nvt_step(nepx_state* st, nepx_nlist* nl, void* stream) {
    int rc = check_nvt(st);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int nb = (st->n + kIB - 1) / kIB;
    if (st->n_sys == 1 && st->n <= kOneBlock) {  // one small system: one launch
        hipLaunchKernelGGL(k_nvt_one, dim3(1), dim3(kOneBlock), 0, s, *st, nl ? nl->flags : nullptr);
    } else {
        hipLaunchKernelGGL(k_nvt, dim3(nb), dim3(kIB), 0, s, *st, 0, nl ? (const int32_t*)nl->flags : nullptr);
        hipLaunchKernelGGL(k_reduce_thermo, dim3(st->n_sys), dim3(kIB), 0, s, *st, nl ? nl->flags : nullptr, 1, 1);
    }
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_nvt_init(nepx_state* st, nepx_nlist* nl, int32_t update_thermostat, void* stream) {
    int rc = check_nvt(st);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int nb = (st->n + kIB - 1) / kIB;
    hipLaunchKernelGGL(k_nvt, dim3(nb), dim3(kIB), 0, s, *st, 1, (const int32_t*)nullptr);
    hipLaunchKernelGGL(k_reduce_thermo, dim3(st->n_sys), dim3(kIB), 0, s, *st, (int32_t*)nullptr,
                       update_thermostat == 2 ? 2 : (update_thermostat ? 1 : 0), 0);
    NEPX_LAUNCH_CHECK();
    (void)nl;
    return NEPX_OK;
}

extern "C" int nepx_nvt_part(nepx_state* st, nepx_nlist* nl, int32_t part, void* stream) {
    int rc = check_nvt(st);
    if (rc) return rc;
    NEPX_REQUIRE(part >= NEPX_NVT_STEP_TWO && part <= NEPX_NVT_THERMO, NEPX_EINVAL, "bad integrator part %d", part);
    NEPX_REQUIRE(st->dp, NEPX_EINVAL, "null dp buffer");
    hipStream_t s = as_stream(stream);
    const int nb = (st->n + kIB - 1) / kIB;
    hipLaunchKernelGGL(k_nvt, dim3(nb), dim3(kIB), 0, s, *st, (int)part, nl ? (const int32_t*)nl->flags : nullptr);
    if (part == NEPX_NVT_THERMO)
        hipLaunchKernelGGL(k_reduce_thermo, dim3(st->n_sys), dim3(kIB), 0, s, *st, nl ? nl->flags : nullptr, 1, 1);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_md_step(nepx_state* st, nepx_nlist* nl, const nepx_pair_model* m, const nepx_net* net,
                            float clip, float* pair_out, void* stream) {
    int rc = nepx_nlist_build(st, nl, 0, stream);
    if (rc) return rc;
    rc = nepx_pairs_filter(st, nl, stream);
    if (rc) return rc;
    rc = nepx_pair_fused_list(m, net, st, nl, clip, pair_out, stream);
    if (rc) return rc;
    rc = nepx_forces_gather(st, nl, pair_out, stream);
    if (rc) return rc;
    return nepx_nvt_step(st, nl, stream);
}
