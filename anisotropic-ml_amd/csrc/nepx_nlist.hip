// Neighbour list on the GPU: cell binning, half Verlet list within r_list
// (rows i<j sorted by j, reverse rows sorted by i), per-step COM filter with
// an FP64 predicate and deterministic compaction, and the deterministic
// gather of per-pair forces.
//
// Replaces the host cKDTree list (nep_simulator/Sim.py:85-92,192-198), the
// per-step COM filter (nep_simulator/DescriptorGenerator.py:168-185) and the
// index_add_ scatter (nep_simulator/ForceEvaluator.py:135-142).
//
// Pair predicate (all paths, bit-exact with the oracle): t = X_j - X_i per
// component, one wrap (t > L/2 -> t - L, t < -L/2 -> t + L), then
// d2 = (tx*tx + ty*ty) + tz*tz in FP64 with no fused multiply-add, d2 < r*r.
#include <mutex>

#include "nepx_common.h"

namespace nepx {

constexpr int kNB = 256;        // threads per block
constexpr int kScanBlock = 1024;
constexpr int kSortMax = 2048;  // longest row the in-LDS sort accepts
constexpr int kSlotsPerBlock = 1024;

__device__ __forceinline__ double wrap1(double t, double L) {
    t = t > 0.5 * L ? t - L : t;
    t = t < -0.5 * L ? L + t : t;
    return t;
}

// Squared minimum-image distance from (xi, yi, zi) to (xj, yj, zj); t = X_j - X_i.
__device__ __forceinline__ double pos_d2(double xi, double yi, double zi, double xj, double yj, double zj, double L,
                                         double& tx, double& ty, double& tz) {
    tx = wrap1(__dsub_rn(xj, xi), L);
    ty = wrap1(__dsub_rn(yj, yi), L);
    tz = wrap1(__dsub_rn(zj, zi), L);
    return __dadd_rn(__dadd_rn(__dmul_rn(tx, tx), __dmul_rn(ty, ty)), __dmul_rn(tz, tz));
}

__device__ __forceinline__ double pair_d2(const nepx_state& st, double L, int i, int j, double& tx,
                                          double& ty, double& tz) {
    return pos_d2(st.x[i], st.y[i], st.z[i], st.x[j], st.y[j], st.z[j], L, tx, ty, tz);
}

// Cell of coordinate x along one side, clamped to [0, nc): NaN and infinities (a faulted
// system's particles, between the step that went non-finite and its removal) land in the
// edge cells; the comparisons run in floating point so no out-of-range value is converted.
__device__ __forceinline__ int cell_coord(double x, double L, int nc) {
    const double f = floor((x + 0.5 * L) / L * (double)nc);
    if (!(f >= 0.0)) return 0;
    return f >= (double)nc ? nc - 1 : (int)f;
}

// A build runs when the list is stale (F_REBUILD) or due soon (F_SOON, set while it is still valid).
__device__ __forceinline__ bool want_build(const int32_t* flags) { return flags[F_REBUILD] != 0 || flags[F_SOON] != 0; }

// A particle takes part in the list when it belongs to an active system.
__device__ __forceinline__ bool live(const nepx_state& st, int i) {
    const int s = st.sys[i];
    return s >= 0 && st.sp[s].active;
}

__global__ void k_set_flag(int32_t* flags, int idx, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) flags[idx] = v;
}

// Start of a build step: clear the longest-row flag and, when a rebuild is due, the cell counters.
__global__ void k_build_begin(int32_t* flags, int32_t* cell_count, int32_t* cell_fill, int64_t nct) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        flags[F_MAXROW] = 0;
        flags[F_LONGSYS] = 0;
    }
    if (!want_build(flags)) return;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nct; k += (int64_t)gridDim.x * blockDim.x) {
        cell_count[k] = 0;
        cell_fill[k] = 0;
    }
}

__global__ void k_zero_i32(int32_t* a, int64_t n, const int32_t* flags) {
    if (flags && !want_build(flags)) return;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        a[k] = 0;
}

__global__ void k_cells_count(nepx_state st, nepx_nlist nl) {
    if (!want_build(nl.flags)) return;
    const int i = blockIdx.x * kNB + threadIdx.x;
    if (i >= st.n) return;
    // positions at this build: the reference for the displacement test that requests the next one
    st.xb[i] = st.x[i];
    st.yb[i] = st.y[i];
    st.zb[i] = st.z[i];
    if (!live(st, i)) return;
    const int s = st.sys[i];
    const nepx_sysparam& sp = st.sp[s];
    if (sp.nc <= 0) return;
    const int cx = cell_coord(st.x[i], sp.L, sp.nc), cy = cell_coord(st.y[i], sp.L, sp.nc),
              cz = cell_coord(st.z[i], sp.L, sp.nc);
    const int c = sp.cell_off + (cx * sp.nc + cy) * sp.nc + cz;
    nl.cell_of[i] = c;
    atomicAdd(&nl.cell_count[c], 1);
}

// Exclusive scan of a[0..n) into out[0..n]; out[n] = total.  One block.
__device__ void block_exclusive_scan(const int32_t* a, int32_t* out, int n) {
    __shared__ int32_t buf[kScanBlock];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += kScanBlock) {
        const int k = base + threadIdx.x;
        const int32_t v = k < n ? a[k] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kScanBlock; off <<= 1) {
            int32_t add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (k < n) out[k] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry += buf[kScanBlock - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_cells(nepx_nlist nl) {
    if (!want_build(nl.flags)) return;
    block_exclusive_scan(nl.cell_count, nl.cell_start, nl.n_cells_total);
}

__global__ void k_cells_fill(nepx_state st, nepx_nlist nl) {
    if (!want_build(nl.flags)) return;
    const int i = blockIdx.x * kNB + threadIdx.x;
    if (i >= st.n || !live(st, i)) return;
    const int s = st.sys[i];
    if (st.sp[s].nc <= 0) return;
    const int c = nl.cell_of[i];
    const int pos = nl.cell_start[c] + atomicAdd(&nl.cell_fill[c], 1);
    nl.cell_list[pos] = i;
    // positions in cell order: the candidate walk reads a cell's positions contiguously
    reinterpret_cast<double4*>(nl.cell_pos)[pos] = make_double4(st.x[i], st.y[i], st.z[i], 0.0);
}

// The list build gives each particle a group of kLG lanes.  The lanes walk
// the candidates j of particle i (27 cells, or the whole system when the
// box is too small for three cells a side) in rounds of kLG: round r of a
// cell covers its entries r*kLG + sub.  Every lane of a group sees the same
// cells and rounds (same i), so ballots inside a round are uniform over the
// group; f(j, valid, d2) is called for each round slot with the squared
// minimum-image distance d2, valid = false for slots past the end of a cell
// and for j == i.  Cell candidates' positions come from
// the cell-ordered copy (contiguous per cell), the same FP64 values as st.x.
constexpr int kLG = 16;

__device__ __forceinline__ uint32_t group_mask(uint64_t ballot) {
    const int gbase = (threadIdx.x & 63) & ~(kLG - 1);
    return (uint32_t)((ballot >> gbase) & ((1ull << kLG) - 1ull));
}

template <class F>
__device__ __forceinline__ void for_candidates(const nepx_state& st, const nepx_nlist& nl, int i, int sub, F&& f) {
    const int s = st.sys[i];
    const nepx_sysparam& sp = st.sp[s];
    const double L = sp.L;
    const double xi = st.x[i], yi = st.y[i], zi = st.z[i];
    double tx, ty, tz;
    if (sp.nc <= 0) {
        for (int base = sp.p_start; base < sp.p_end; base += kLG) {
            const int j = base + sub;
            const bool valid = j < sp.p_end && j != i && st.sys[j] == s;
            const int jj = valid ? j : i;
            const double d2 = pos_d2(xi, yi, zi, st.x[jj], st.y[jj], st.z[jj], L, tx, ty, tz);
            f(jj, valid, d2);
        }
        return;
    }
    const double4* cpos = reinterpret_cast<const double4*>(nl.cell_pos);
    const int nc = sp.nc;
    const int cx = cell_coord(st.x[i], L, nc), cy = cell_coord(st.y[i], L, nc), cz = cell_coord(st.z[i], L, nc);
    for (int dx = -1; dx <= 1; ++dx) {
        const int ax = (cx + dx + nc) % nc;
        for (int dy = -1; dy <= 1; ++dy) {
            const int ay = (cy + dy + nc) % nc;
            for (int dz = -1; dz <= 1; ++dz) {
                const int az = (cz + dz + nc) % nc;
                const int c = sp.cell_off + (ax * nc + ay) * nc + az;
                const int b = nl.cell_start[c], e = nl.cell_start[c + 1];
                for (int base = b; base < e; base += kLG) {
                    const int k = base + sub;
                    const bool in = k < e;
                    const int j = in ? nl.cell_list[k] : i;
                    const double4 pj = in ? cpos[k] : make_double4(xi, yi, zi, 0.0);
                    const double d2 = pos_d2(xi, yi, zi, pj.x, pj.y, pj.z, L, tx, ty, tz);
                    f(j, j != i, d2);
                }
            }
        }
    }
}

template <int W>
__device__ __forceinline__ int group_sum(int v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

__global__ __launch_bounds__(kNB) void k_nbr_count(nepx_state st, nepx_nlist nl) {
    if (!want_build(nl.flags)) return;
    const int i = (blockIdx.x * kNB + threadIdx.x) / kLG;
    const int sub = threadIdx.x % kLG;
    if (i >= st.n) return;  // whole groups exit together
    int nf = 0, nr = 0;
    if (live(st, i)) {
        const nepx_sysparam& sp = st.sp[st.sys[i]];
        const double r2 = __dmul_rn(sp.r_list, sp.r_list);
        for_candidates(st, nl, i, sub, [&](int j, bool valid, double d2) {
            if (valid && d2 < r2) {
                if (j > i) ++nf;
                else ++nr;
            }
        });
    }
    nf = group_sum<kLG>(nf);
    nr = group_sum<kLG>(nr);
    if (sub == 0) {
        nl.row_count[i] = nf;
        nl.rrow_count[i] = nr;
        if (nf > kSortMax || nr > kSortMax) {
            atomicMax(&nl.flags[F_MAXROW], nf > nr ? nf : nr);
            nl.flags[F_LONGSYS] = st.sys[i] + 1;  // one system holding such a row (any, if several)
        }
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_rows(nepx_nlist nl, int n) {
    if (!want_build(nl.flags)) return;
    if (blockIdx.x == 0) {
        block_exclusive_scan(nl.row_count, nl.row_start, n);
        __syncthreads();
        if (threadIdx.x == 0) {
            nl.flags[F_PL] = nl.row_start[n];
            if ((int64_t)nl.row_start[n] > nl.cap || nl.flags[F_MAXROW] > kSortMax) nl.flags[F_OVERFLOW] = 1;
        }
    } else {
        block_exclusive_scan(nl.rrow_count, nl.rrow_start, n);
    }
}

// Same walk as k_nbr_count; each round's accepted candidates are placed by a
// ballot prefix within the group (no atomics).  The row order this gives
// depends on the cell lists; k_sort_rows makes it canonical.
__global__ __launch_bounds__(kNB) void k_nbr_fill(nepx_state st, nepx_nlist nl) {
    if (!want_build(nl.flags) || nl.flags[F_OVERFLOW]) return;
    const int i = (blockIdx.x * kNB + threadIdx.x) / kLG;
    const int sub = threadIdx.x % kLG;
    if (i >= st.n || !live(st, i)) return;  // uniform per group
    const nepx_sysparam& sp = st.sp[st.sys[i]];
    const double r2 = __dmul_rn(sp.r_list, sp.r_list);
    const uint32_t below = (1u << sub) - 1u;
    int pf = nl.row_start[i], pr = nl.rrow_start[i];
    for_candidates(st, nl, i, sub, [&](int j, bool valid, double d2) {
        const bool in = valid && d2 < r2;
        const bool fwd = in && j > i, rev = in && j < i;
        const uint32_t mf = group_mask(__ballot(fwd)), mr = group_mask(__ballot(rev));
        if (fwd) {
            const int slot = pf + __popc(mf & below);
            nl.nbr[slot] = j;
            nl.slot_row[slot] = i;
        }
        if (rev) nl.rnbr[pr + __popc(mr & below)] = j;
        pf += __popc(mf);
        pr += __popc(mr);
    });
}

// Sort each row ascending: one wave per row (blocks of 64 stride over the
// rows; grid y selects the forward or the reverse rows), bitonic in LDS.  The
// grid is capped, so the launch costs little when no rebuild is due.
constexpr int kSortThreads = 64;
__global__ __launch_bounds__(kSortThreads) void k_sort_rows(nepx_nlist nl, int n) {
    if (!want_build(nl.flags) || nl.flags[F_OVERFLOW]) return;
    const int reverse = blockIdx.y;  // y = 0: forward rows, 1: reverse rows
    __shared__ int32_t buf[kSortMax];
    const int32_t* start = reverse ? nl.rrow_start : nl.row_start;
    int32_t* data = reverse ? nl.rnbr : nl.nbr;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int b = start[i], len = start[i + 1] - b;
        if (len <= 1) continue;  // uniform per block
        int p2 = 1;
        while (p2 < len) p2 <<= 1;
        for (int k = threadIdx.x; k < p2; k += kSortThreads) buf[k] = k < len ? data[b + k] : 0x7fffffff;
        __syncthreads();
        for (int size = 2; size <= p2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int k = threadIdx.x; k < p2; k += kSortThreads) {
                    const int partner = k ^ stride;
                    if (partner > k) {
                        const bool up = (k & size) == 0;
                        const int32_t a = buf[k], c = buf[partner];
                        if ((a > c) == up) {
                            buf[k] = c;
                            buf[partner] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int k = threadIdx.x; k < len; k += kSortThreads) data[b + k] = buf[k];
        __syncthreads();
    }
}

// Reverse entries hold the partner i; replace it by the slot of (i, k) in row i
// (binary search in the sorted row).  kLG lanes per particle k.
__global__ __launch_bounds__(kNB) void k_rev_slots(nepx_nlist nl, int n) {
    if (!want_build(nl.flags) || nl.flags[F_OVERFLOW]) return;
    const int k = (blockIdx.x * kNB + threadIdx.x) / kLG;
    const int sub = threadIdx.x % kLG;
    if (k >= n) return;
    for (int e = nl.rrow_start[k] + sub; e < nl.rrow_start[k + 1]; e += kLG) {
        const int i = nl.rnbr[e];
        int lo = nl.row_start[i], hi = nl.row_start[i + 1] - 1, slot = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int v = nl.nbr[mid];
            if (v == k) {
                slot = mid;
                break;
            }
            if (v < k) lo = mid + 1;
            else hi = mid - 1;
        }
        if (slot < 0) nl.flags[F_OVERFLOW] = 2;  // inconsistent list: never expected
        nl.rnbr[e] = slot;
        if (nl.rpos && slot >= 0) nl.rpos[slot] = e;  // where the reverse row of k holds this slot
    }
}

__global__ void k_build_finish(nepx_nlist nl) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && want_build(nl.flags)) {
        nl.flags[F_REBUILD] = 0;
        nl.flags[F_SOON] = 0;
        nl.flags[F_BUILDS] += 1;
    }
}

// ---- explicit pair list (API path) ----------------------------------------
__global__ void k_pairs_count(nepx_nlist nl, const int64_t* pairs, int64_t m, int n) {
    const int64_t k = blockIdx.x * (int64_t)kNB + threadIdx.x;
    if (k >= m) return;
    int a = (int)pairs[2 * k], b = (int)pairs[2 * k + 1];
    if (a == b || a < 0 || b < 0 || a >= n || b >= n) {
        nl.flags[F_OVERFLOW] = 3;
        return;
    }
    if (a > b) {
        const int t = a;
        a = b;
        b = t;
    }
    atomicAdd(&nl.row_count[a], 1);
    atomicAdd(&nl.rrow_count[b], 1);
}

__global__ void k_pairs_fill(nepx_nlist nl, const int64_t* pairs, int64_t m, int32_t* fill_f, int32_t* fill_r) {
    if (nl.flags[F_OVERFLOW]) return;
    const int64_t k = blockIdx.x * (int64_t)kNB + threadIdx.x;
    if (k >= m) return;
    int a = (int)pairs[2 * k], b = (int)pairs[2 * k + 1];
    if (a > b) {
        const int t = a;
        a = b;
        b = t;
    }
    const int pf = nl.row_start[a] + atomicAdd(&fill_f[a], 1);
    nl.nbr[pf] = b;
    nl.slot_row[pf] = a;
    const int pr = nl.rrow_start[b] + atomicAdd(&fill_r[b], 1);
    nl.rnbr[pr] = a;
}

__global__ void k_rows_maxlen(nepx_nlist nl, int n) {
    const int i = blockIdx.x * kNB + threadIdx.x;
    if (i >= n) return;
    const int nf = nl.row_count[i], nr = nl.rrow_count[i];
    if (nf > kSortMax || nr > kSortMax) atomicMax(&nl.flags[F_MAXROW], nf > nr ? nf : nr);
}

// ---- per-step filter and compaction ---------------------------------------
// Each thread evaluates kSlotsPerThread list slots (stride kNB within the
// block's kSlotsPerBlock), issuing every load of all its slots before the
// predicates so the gathers overlap.
constexpr int kSlotsPerThread = kSlotsPerBlock / kNB;

struct SlotEval {
    bool act[kSlotsPerThread];
    int i[kSlotsPerThread], j[kSlotsPerThread];
    double tx[kSlotsPerThread], ty[kSlotsPerThread], tz[kSlotsPerThread];
};

__device__ __forceinline__ void eval_slots(const nepx_state& st, const nepx_nlist& nl, int base, int PL,
                                           SlotEval& r) {
    bool in[kSlotsPerThread];
    int si[kSlotsPerThread], sj[kSlotsPerThread];
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; ++k) {
        const int s = base + k * kNB + threadIdx.x;
        in[k] = s < PL;
        r.i[k] = in[k] ? nl.slot_row[s] : 0;
        r.j[k] = in[k] ? nl.nbr[s] : 0;
    }
    if (st.n_sys == 1) {  // one system: its particle range stands in for the system-id reads
        const int p0 = st.sp[0].p_start, p1 = st.sp[0].p_end;
#pragma unroll
        for (int k = 0; k < kSlotsPerThread; ++k) {
            si[k] = r.i[k] >= p0 && r.i[k] < p1 ? 0 : -1;
            sj[k] = r.j[k] >= p0 && r.j[k] < p1 ? 0 : -1;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kSlotsPerThread; ++k) {
            si[k] = st.sys[r.i[k]];
            sj[k] = st.sys[r.j[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; ++k) {
        // slots left over from a freed system, or of an inactive one, are not pairs
        const bool same = in[k] && si[k] >= 0 && si[k] == sj[k];
        const nepx_sysparam& sp = st.sp[same ? si[k] : 0];
        const bool live_sys = same && sp.active;
        const double d2 = pair_d2(st, sp.L, r.i[k], r.j[k], r.tx[k], r.ty[k], r.tz[k]);
        r.act[k] = live_sys && d2 < __dmul_rn(sp.rc, sp.rc);
    }
}

// A step whose list is stale (F_REBUILD still set: a step of a multi-step graph after the skin ran
// out) is frozen: the filter selects no pairs and flags F_FROZEN, so the gather and the integrator
// leave the state alone until the next build.
__global__ __launch_bounds__(kNB) void k_filter_count(nepx_state st, nepx_nlist nl) {
    const int PL = nl.flags[F_PL];
    const int base = blockIdx.x * kSlotsPerBlock;
    if (base >= PL || nl.flags[F_OVERFLOW] || nl.flags[F_REBUILD]) {
        if (threadIdx.x == 0) nl.blk_count[blockIdx.x] = 0;
        return;
    }
    SlotEval r;
    eval_slots(st, nl, base, PL, r);
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; ++k) cnt += r.act[k] ? 1 : 0;
    __shared__ int red[kNB];
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = kNB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) nl.blk_count[blockIdx.x] = red[0];
}

// Sum of cnt[0..b) by one block (integer, so the order does not matter).
__device__ __forceinline__ int block_prefix(const int32_t* __restrict__ cnt, int b) {
    __shared__ int wpart[kNB / 64];
    int v = 0;
    for (int k = threadIdx.x; k < b; k += kNB) v += cnt[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < kNB / 64; ++w) t += wpart[w];
    return t;
}

// Each block finds its output offset by summing the counts of the blocks before it (one count per
// kSlotsPerBlock list slots, so few), and the block holding the last slot stores the pair count;
// no separate scan launch.
__global__ __launch_bounds__(kNB) void k_filter_write(nepx_state st, nepx_nlist nl) {
    const int PL = nl.flags[F_PL];
    const int base = blockIdx.x * kSlotsPerBlock;
    const int frozen = nl.flags[F_REBUILD] != 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) nl.flags[F_FROZEN] = frozen;
    if (nl.flags[F_OVERFLOW] || frozen || PL == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) nl.flags[F_P] = 0;
        return;
    }
    // a new pass: the pair kernels tag the reverse-order records they write with it
    if (blockIdx.x == 0 && threadIdx.x == 0) nl.flags[F_TAG] = nl.flags[F_TAG] + 1;
    if (base >= PL) return;
    const int first = block_prefix(nl.blk_count, blockIdx.x);
    if (base + kSlotsPerBlock >= PL && threadIdx.x == 0) nl.flags[F_P] = first + nl.blk_count[blockIdx.x];
    __shared__ int wsum[kSlotsPerThread][kNB / 64];
    SlotEval r;
    eval_slots(st, nl, base, PL, r);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t below[kSlotsPerThread];
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; ++k) {
        const uint64_t mask = __ballot(r.act[k]);
        below[k] = mask & ((1ull << lane) - 1ull);
        if (lane == 0) wsum[k][wv] = __popcll(mask);
    }
    __syncthreads();
    int out = first;
#pragma unroll
    for (int k = 0; k < kSlotsPerThread; ++k) {
        const int s = base + k * kNB + threadIdx.x;
        int woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kNB / 64; ++w) {
            woff += w < wv ? wsum[k][w] : 0;
            tot += wsum[k][w];
        }
        if (s < PL) {
            if (r.act[k]) {
                const int c = out + woff + __popcll(below[k]);
                nl.cidx[s] = c;
                nl.pair_i[c] = r.i[k];
                nl.pair_j[c] = r.j[k];
                if (nl.pair_e) nl.pair_e[c] = nl.rpos[s];
                reinterpret_cast<float4*>(nl.pair_t)[c] =
                    make_float4((float)r.tx[k], (float)r.ty[k], (float)r.tz[k], 0.f);
            } else {
                nl.cidx[s] = -1;
            }
        }
        out += tot;
    }
}

// ---- gather ----------------------------------------------------------------
// GL lanes per particle (tuned; 16 by default): lanes stride over the
// particle's forward row (i-role: +F, +tau_i, +E) and reverse row (j-role:
// -F, +tau_j), accumulate in FP64 and reduce with a fixed butterfly, so the
// result is independent of timing (and, for one GL, of the launch).  The j side
// comes from the reverse-order records when this pass's pair kernel wrote them
// (the same values, so the same sums bit for bit), else through the slot ->
// compacted-index map from the per-pair records.
template <int GL>
__device__ __forceinline__ double red_lanes(double v) {
#pragma unroll
    for (int o = GL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, GL);
    return v;
}

template <int GL>
__global__ __launch_bounds__(kNB) void k_gather(nepx_state st, nepx_nlist nl, const float4* __restrict__ out) {
    // particles of one XCD are contiguous, so the records that neighbouring rows share
    // (forward rows of i, reverse rows of its neighbours) are fetched into one L2
    const int k = (xcd_contiguous_block(blockIdx.x, gridDim.x) * kNB + threadIdx.x) / GL;
    const int sub = threadIdx.x % GL;
    if (k >= st.n) return;  // whole lane groups exit together
    if (nl.flags[F_OVERFLOW] || nl.flags[F_FROZEN]) return;  // skipped step: forces stay as they are
    // the pair kernel of this pass wrote reverse-order records (uniform: every thread reads the same flags)
    const int tag = nl.flags[F_TAG];
    const bool rev = nl.rrec != nullptr && nl.flags[F_RTAG] == tag;
    double fx = 0, fy = 0, fz = 0, tx = 0, ty = 0, tz = 0, u = 0;
    if (live(st, k)) {
        for (int s = nl.row_start[k] + sub; s < nl.row_start[k + 1]; s += GL) {
            const int c = nl.cidx[s];
            if (c < 0) continue;
            const float4 a = out[3 * (size_t)c], b = out[3 * (size_t)c + 1];
            fx += a.x;
            fy += a.y;
            fz += a.z;
            u += a.w;
            tx += b.x;
            ty += b.y;
            tz += b.z;
        }
        if (rev) {  // this pass's j-side records, in reverse-row order: read contiguously
            const float4* rr = reinterpret_cast<const float4*>(nl.rrec);
            for (int e = nl.rrow_start[k] + sub; e < nl.rrow_start[k + 1]; e += GL) {
                const float4 a = rr[2 * (size_t)e];
                if (__float_as_int(a.w) != tag) continue;  // no active pair at this entry this step
                const float4 b = rr[2 * (size_t)e + 1];
                fx -= a.x;
                fy -= a.y;
                fz -= a.z;
                tx += b.x;
                ty += b.y;
                tz += b.z;
            }
        } else {
            for (int e = nl.rrow_start[k] + sub; e < nl.rrow_start[k + 1]; e += GL) {
                const int c = nl.cidx[nl.rnbr[e]];
                if (c < 0) continue;
                const float4 a = out[3 * (size_t)c], b = out[3 * (size_t)c + 2];
                fx -= a.x;
                fy -= a.y;
                fz -= a.z;
                tx += b.x;
                ty += b.y;
                tz += b.z;
            }
        }
    }
    fx = red_lanes<GL>(fx);
    fy = red_lanes<GL>(fy);
    fz = red_lanes<GL>(fz);
    tx = red_lanes<GL>(tx);
    ty = red_lanes<GL>(ty);
    tz = red_lanes<GL>(tz);
    u = red_lanes<GL>(u);
    if (sub == 0) {
        st.fx[k] = fx;
        st.fy[k] = fy;
        st.fz[k] = fz;
        st.tx[k] = tx;
        st.ty[k] = ty;
        st.tz[k] = tz;
        st.upot[k] = u;
    }
}

// State and list buffers every entry point reads; blk_count (the compaction's
// per-block partials) is checked by the entry points that take it from the list.
static int check_state_core(const nepx_state* st, const nepx_nlist* nl) {
    NEPX_REQUIRE(st && nl, NEPX_EINVAL, "null state or list");
    NEPX_REQUIRE(st->n > 0 && st->n_sys > 0, NEPX_EINVAL, "empty state");
    NEPX_REQUIRE(st->x && st->y && st->z && st->sys && st->sp, NEPX_EINVAL, "null state buffer");
    NEPX_REQUIRE(nl->flags && nl->nbr && nl->slot_row && nl->rnbr && nl->row_start && nl->rrow_start &&
                     nl->row_count && nl->rrow_count && nl->cidx && nl->pair_i && nl->pair_j && nl->pair_t,
                 NEPX_EINVAL, "null list buffer");
    NEPX_REQUIRE(nl->cap > 0 && nl->cap < (int64_t)1 << 31, NEPX_ECAP, "list capacity %lld out of range",
                 (long long)nl->cap);
    NEPX_REQUIRE(!(nl->pair_e || nl->rrec) || (nl->rpos && nl->pair_e && nl->rrec), NEPX_EINVAL,
                 "reverse-order records need rpos, pair_e and rrec together");
    return NEPX_OK;
}

static int check_state(const nepx_state* st, const nepx_nlist* nl) {
    const int rc = check_state_core(st, nl);
    if (rc) return rc;
    NEPX_REQUIRE(nl->blk_count, NEPX_EINVAL, "null list buffer (blk_count)");
    return NEPX_OK;
}

static int blocks_for(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace nepx

using namespace nepx;

extern "C" int nepx_nlist_build(nepx_state* st, nepx_nlist* nl, int32_t force, void* stream) {
    int rc = check_state(st, nl);
    if (rc) return rc;
    NEPX_REQUIRE(nl->cell_count && nl->cell_start && nl->cell_fill && nl->cell_list && nl->cell_of && nl->cell_pos,
                 NEPX_EINVAL, "null cell buffer");
    NEPX_REQUIRE(((uintptr_t)nl->cell_pos & 31) == 0, NEPX_EINVAL, "cell_pos must be 32-byte aligned");
    hipStream_t s = as_stream(stream);
    const int n = st->n;
    const int gb = blocks_for(n, kNB);
    if (force) {  // a forced build starts clean: request it and clear a stale overflow
        hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_REBUILD, 1);
        hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_OVERFLOW, 0);
        NEPX_LAUNCH_CHECK();
    }
    const int nct = nl->n_cells_total > 0 ? nl->n_cells_total : 1;
    hipLaunchKernelGGL(k_build_begin, dim3(blocks_for(nct, kNB)), dim3(kNB), 0, s, nl->flags, nl->cell_count,
                       nl->cell_fill, (int64_t)nct);
    hipLaunchKernelGGL(k_cells_count, dim3(gb), dim3(kNB), 0, s, *st, *nl);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(kScanBlock), 0, s, *nl);
    hipLaunchKernelGGL(k_cells_fill, dim3(gb), dim3(kNB), 0, s, *st, *nl);
    const int gg = blocks_for((int64_t)n * kLG, kNB);
    hipLaunchKernelGGL(k_nbr_count, dim3(gg), dim3(kNB), 0, s, *st, *nl);
    hipLaunchKernelGGL(k_scan_rows, dim3(2), dim3(kScanBlock), 0, s, *nl, n);
    hipLaunchKernelGGL(k_nbr_fill, dim3(gg), dim3(kNB), 0, s, *st, *nl);
    const int gs = n < 16384 ? n : 16384;
    hipLaunchKernelGGL(k_sort_rows, dim3(gs, 2), dim3(kSortThreads), 0, s, *nl, n);
    hipLaunchKernelGGL(k_rev_slots, dim3(gg), dim3(kNB), 0, s, *nl, n);
    hipLaunchKernelGGL(k_build_finish, dim3(1), dim3(64), 0, s, *nl);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_nlist_from_pairs(nepx_state* st, nepx_nlist* nl, const int64_t* pairs, int64_t m,
                                     void* stream) {
    int rc = check_state(st, nl);
    if (rc) return rc;
    NEPX_REQUIRE(m >= 0 && (m == 0 || pairs), NEPX_EINVAL, "bad pair array");
    NEPX_REQUIRE(m <= nl->cap, NEPX_ECAP, "%lld pairs exceed list capacity %lld", (long long)m,
                 (long long)nl->cap);
    NEPX_REQUIRE(nl->cell_fill && nl->cell_count, NEPX_EINVAL, "null scratch buffer");
    hipStream_t s = as_stream(stream);
    const int n = st->n;
    // flags: request a "build" so the shared kernels run, clear overflow
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_REBUILD, 1);
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_OVERFLOW, 0);
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_MAXROW, 0);
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, nl->flags, F_LONGSYS, 0);
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(n + 1, kNB)), dim3(kNB), 0, s, nl->row_count,
                       (int64_t)(n + 1), nullptr);
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(n + 1, kNB)), dim3(kNB), 0, s, nl->rrow_count,
                       (int64_t)(n + 1), nullptr);
    if (m > 0)
        hipLaunchKernelGGL(k_pairs_count, dim3(blocks_for(m, kNB)), dim3(kNB), 0, s, *nl, pairs, m, n);
    hipLaunchKernelGGL(k_rows_maxlen, dim3(blocks_for(n, kNB)), dim3(kNB), 0, s, *nl, n);
    hipLaunchKernelGGL(k_scan_rows, dim3(2), dim3(kScanBlock), 0, s, *nl, n);
    // per-row fill counters reuse the cell scratch when it is large enough
    NEPX_REQUIRE(nl->n_cells_total >= 2 * n, NEPX_ECAP, "cell scratch (%d) smaller than 2*n (%d)",
                 nl->n_cells_total, 2 * n);
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(2 * n, kNB)), dim3(kNB), 0, s, nl->cell_fill,
                       (int64_t)(2 * n), nullptr);
    if (m > 0)
        hipLaunchKernelGGL(k_pairs_fill, dim3(blocks_for(m, kNB)), dim3(kNB), 0, s, *nl, pairs, m,
                           nl->cell_fill, nl->cell_fill + n);
    const int gs = n < 16384 ? n : 16384;
    hipLaunchKernelGGL(k_sort_rows, dim3(gs, 2), dim3(kSortThreads), 0, s, *nl, n);
    hipLaunchKernelGGL(k_rev_slots, dim3(blocks_for((int64_t)n * kLG, kNB)), dim3(kNB), 0, s, *nl, n);
    hipLaunchKernelGGL(k_build_finish, dim3(1), dim3(64), 0, s, *nl);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_pairs_filter(nepx_state* st, nepx_nlist* nl, void* stream) {
    int rc = check_state(st, nl);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int nblk = blocks_for(nl->cap, kSlotsPerBlock);
    hipLaunchKernelGGL(k_filter_count, dim3(nblk), dim3(kNB), 0, s, *st, *nl);
    hipLaunchKernelGGL(k_filter_write, dim3(nblk), dim3(kNB), 0, s, *st, *nl);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

// ---- allocation-free filter --------------------------------------------------
// Scratch of nepx_pairs_filter_noalloc: the compaction's per-block partial counts
// (one int per kSlotsPerBlock list slots), in place of the list's blk_count.
// A file-scope device array: each GPU the library runs on holds its own copy
// (256 KiB) until the process ends.  Host code never passes the array's name to
// a kernel; it uses the address hipGetSymbolAddress returns for the current
// device, looked up once per device on the first call there.
constexpr int kNoallocMaxBlocks = 1 << 16;   // list capacity up to 64 Mi slots
static __device__ __align__(256) int32_t scratch[kNoallocMaxBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_dev[kMaxDevices] = {};

int scratch_for_current_device(int32_t** out) {
    int dev = -1;
    NEPX_HIP_CHECK(hipGetDevice(&dev));
    NEPX_REQUIRE(dev >= 0 && dev < kMaxDevices, NEPX_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev] == nullptr) {
        void* p = nullptr;
        NEPX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NEPX_REQUIRE(p != nullptr, NEPX_EHIP, "no device address for the filter scratch");
        g_scratch_dev[dev] = static_cast<int32_t*>(p);
    }
    *out = g_scratch_dev[dev];
    return NEPX_OK;
}
}  // namespace

extern "C" int nepx_pairs_filter_noalloc(nepx_state* st, nepx_nlist* nl, void* stream) {
    int rc = check_state_core(st, nl);
    if (rc) return rc;
    const int nblk = blocks_for(nl->cap, kSlotsPerBlock);
    NEPX_REQUIRE(nblk <= kNoallocMaxBlocks, NEPX_ECAP,
                 "list capacity %lld exceeds the %lld slots of the allocation-free filter's scratch",
                 (long long)nl->cap, (long long)kNoallocMaxBlocks * kSlotsPerBlock);
    int32_t* scr = nullptr;
    rc = scratch_for_current_device(&scr);
    if (rc) return rc;
    // k_filter_count writes every partial that k_filter_write reads (zero for empty blocks), so
    // what an earlier call left in the scratch is never read.
    nepx_nlist l = *nl;
    l.blk_count = scr;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_filter_count, dim3(nblk), dim3(kNB), 0, s, *st, l);
    hipLaunchKernelGGL(k_filter_write, dim3(nblk), dim3(kNB), 0, s, *st, l);
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}

extern "C" int nepx_forces_gather(nepx_state* st, nepx_nlist* nl, const float* pair_out, void* stream) {
    int rc = check_state(st, nl);
    if (rc) return rc;
    NEPX_REQUIRE(pair_out && st->fx && st->fy && st->fz && st->tx && st->ty && st->tz && st->upot, NEPX_EINVAL,
                 "null force buffer");
    const int gl = tuning().gather_lanes;
#define NEPX_GATHER(GL)                                                                                   \
    hipLaunchKernelGGL(k_gather<GL>, dim3(blocks_for((int64_t)st->n * GL, kNB)), dim3(kNB), 0,             \
                       as_stream(stream), *st, *nl, (const float4*)pair_out)
    switch (gl) {
        case 4: NEPX_GATHER(4); break;
        case 8: NEPX_GATHER(8); break;
        case 32: NEPX_GATHER(32); break;
        case 64: NEPX_GATHER(64); break;
        default: NEPX_GATHER(16); break;
    }
#undef NEPX_GATHER
    NEPX_LAUNCH_CHECK();
    return NEPX_OK;
}
