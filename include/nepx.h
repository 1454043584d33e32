/*
 * nepx -- MI355X-native rigid-body NEP pair step (C ABI).
 *
 * One shared library, libnepx.so, built for gfx950 with hipcc.  Every entry
 * point is asynchronous on the HIP stream passed in (NULL = the default
 * stream), takes caller-owned device buffers as plain pointers, returns 0 on
 * success or a negative NEPX_E* code, and never throws.  nepx_last_error()
 * returns the message of the last failure on the calling host thread.
 *
 * The reference has no C FFI: its boundary is a set of Python class methods
 * that launch CuPy raw kernels.  Each entry below names the reference
 * function (file:line, paths relative to the reference checkout) whose work
 * it replaces; the Python package anisotropic-ml_amd/nepx keeps the
 * reference method names on top of these calls.
 */
#ifndef NEPX_H
#define NEPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEPX_OK 0
#define NEPX_EINVAL (-1)    /* bad argument (shape, size, null pointer)        */
#define NEPX_EUNSUP (-2)    /* no kernel instance for these hypers / widths    */
#define NEPX_EHIP (-3)      /* a HIP runtime call failed                       */
#define NEPX_ECAP (-4)      /* a size exceeds a documented maximum             */

#define NEPX_MAX_SITES 16   /* body-frame sites per particle                   */
#define NEPX_ACT_TANH 0
#define NEPX_ACT_RELU 1

/* Pair model: rigid-body sites + NEP hypers (DescriptorGenerator.py:37-43,58-125). */
typedef struct nepx_pair_model {
    int32_t n_sites;                 /* Ns sites per body; a pair cloud has 2*Ns */
    int32_t nrad, nang, lmax;        /* D = (nrad+1) + (nang+1)*lmax             */
    float r_nep;                     /* descriptor cutoff (nep_cutoff)           */
    float sites[3 * NEPX_MAX_SITES]; /* body-frame sites, xyz interleaved        */
} nepx_pair_model;

/* Energy network (ForceEvaluator.py:6-26): D -> width (-> width)*n_hidden -> 1.
 * w_packed is a device buffer written by nepx_net_pack(). */
typedef struct nepx_net {
    int32_t d_in, width, n_hidden, act;
    float en_min, en_max;            /* E = en_min + (en_max-en_min) * NN(G)    */
    const float* w_packed;           /* nepx_net_pack output (FP32 operands, biases, output layer) */
    const void* w_half;              /* nepx_net_pack_half output, or NULL        */
    int32_t precision;               /* NEPX_NET_F32, _F16X3, _BF16 or _BF16X3   */
    int32_t pad;
} nepx_net;

#define NEPX_NET_F32 0               /* v_mfma_f32_16x16x4_f32                    */
#define NEPX_NET_F16X3 1             /* hi/lo fp16 split, 3 x v_mfma_f32_16x16x32_f16, FP32 accumulate */
#define NEPX_NET_BF16 2              /* bf16 operands, 1 x v_mfma_f32_16x16x32_bf16, FP32 accumulate */
#define NEPX_NET_BF16X3 3            /* hi/lo bf16 split, 3 x v_mfma_f32_16x16x32_bf16, FP32 accumulate */
/* The bf16 forms take tanh networks only (NEPX_EUNSUP for ReLU). */

/* Per-system parameters (device array of n_sys records, 16 doubles + 16 ints). */
typedef struct nepx_sysparam {
    double L, kT, dt, tau, ndof, rndof;  /* box, temperature, step, NH tau, N_dof, RN_dof */
    double moi[3];                       /* diagonal inertia (Sim.py:43-44)      */
    double r_list, rc, skin;             /* list radius, COM cutoff, skin        */
    double pad[4];                       /* pad[0]: early-rebuild fraction f of the half skin (0 = off):
                                            a rebuild is requested for the next build point once a
                                            particle has moved f * skin / 2; others zero */
    int32_t p_start, p_end;              /* particle range (block aligned start) */
    int32_t blk_start, blk_end;          /* integrator block range               */
    int32_t nc, cell_off;                /* cells per side (0 = all pairs), offset */
    int32_t active;                      /* 0: the system is skipped by every kernel */
    int32_t ipad[9];                     /* zero */
} nepx_sysparam;

/* Thermostat record per system (Sim.py:99-111, IntegrateStepTwoKernels.py:74-130). */
typedef struct nepx_thermo {
    double xi, eta, xi_r, eta_r;     /* Nose-Hoover variables                    */
    double theta, e_r;               /* exp(-xi dt/2), exp(-xi_r dt/2)            */
    double Kt, Kr, U;                /* kinetic (trans, rot) and potential energy */
    double max_disp2;                /* largest squared displacement since build */
    double steps, budget;            /* steps done; the system deactivates itself at budget (> 0) */
    double pending;                  /* != 0: initial kinetic energies + first update still to do */
    double fault;                    /* 1: energies or displacement went non-finite; the system was
                                        deactivated (active = 0) at that step, its steps not counted */
    double pad[2];
} nepx_thermo;

/* Particle arena, SoA, device pointers (Sim.py:147-169 holds the same state AoS f32). */
typedef struct nepx_state {
    int32_t n;                       /* particles incl. padding (mass 0)          */
    int32_t n_sys;
    double *x, *y, *z, *vx, *vy, *vz;
    double *q0, *q1, *q2, *q3;       /* orientation, scalar first                */
    double *p0, *p1, *p2, *p3;       /* angular-momentum quaternion (angmom)     */
    double *mass;
    double *fx, *fy, *fz, *tx, *ty, *tz, *upot;
    double *xb, *yb, *zb;            /* positions at the last list build          */
    float* qf;                       /* float4 copy of q read by the pair kernels */
    int32_t* sys;                    /* system of each particle                   */
    nepx_sysparam* sp;               /* n_sys records                             */
    nepx_thermo* th;                 /* n_sys records                             */
    double* partials;                /* 4 doubles per integrator block            */
    double* dp;                      /* 4 doubles per particle: torque kick of step two */
} nepx_state;

/* Verlet half list (i<j, rows sorted by j) + reverse rows + per-step compaction. */
typedef struct nepx_nlist {
    int64_t cap;                     /* slots in nbr/slot_row                     */
    int32_t n_cells_total;
    int32_t* cell_count;             /* n_cells_total+1 */
    int32_t* cell_start;             /* n_cells_total+1 */
    int32_t* cell_fill;              /* n_cells_total   */
    int32_t* cell_list;              /* n               */
    double* cell_pos;                /* 4n: x, y, z of cell_list[k] at 4k..4k+2 (written by the build) */
    int32_t* cell_of;                /* n               */
    int32_t* row_count;              /* n+1 (forward), then scanned to row_start  */
    int32_t* row_start;              /* n+1 */
    int32_t* rrow_count;             /* n+1 */
    int32_t* rrow_start;             /* n+1 */
    int32_t* nbr;                    /* cap: j of slot s                          */
    int32_t* slot_row;               /* cap: i of slot s                          */
    int32_t* rnbr;                   /* cap: reverse rows, then slot ids          */
    int32_t* cidx;                   /* cap: compacted index of slot or -1        */
    int32_t* blk_count;              /* compaction partials (cap/1024 rounded up)  */
    int32_t* flags;                  /* 16 ints: [0] list stale (rebuild; steps wait), [1] overflow,
                                        [2] P_L, [3] P, [4] builds, [5] longest row (when over 2048),
                                        [6] step frozen (set by the filter while [0] is set), [7] rebuild
                                        at the next build point, [8] 1 + a system with a row over 2048
                                        (0: none), [9] filter pass count (the tag of this step's
                                        reverse-order records), [10] tag of the last pass whose pair
                                        kernel wrote them; others zero */
    int32_t* pair_i;                 /* cap: compacted pairs                       */
    int32_t* pair_j;
    float* pair_t;                   /* cap*4: minimum-image t = X_j - X_i (f32)  */
    /* Optional (NULL: unused) reverse-order pair records, so that the gather reads the j side of
       each pair contiguously along the reverse rows: */
    int32_t* rpos;                   /* cap: reverse-row entry of list slot s (written by the build) */
    int32_t* pair_e;                 /* cap: reverse-row entry of compacted pair c (by the filter)   */
    float* rrec;                     /* cap*8, zero-initialised: at entry e, {F_i, tag}{tau_j, 0} of
                                        its pair (by nepx_pair_fused_list)                          */
} nepx_nlist;

/* ---- library ---------------------------------------------------------- */
const char* nepx_last_error(void);
int nepx_version(void);
/* Sizes of the shared structs, for binding checks from the host language. */
int nepx_struct_sizes(int32_t* out7);

/* ---- launch tuning ------------------------------------------------------ */
/* Launch parameters chosen by the autotuner (nepx/tune.py).  They are
 * process-wide and read at launch time, so a captured graph keeps the values
 * that were in force when it was captured.  Zero fields select the defaults.
 * No counterpart in the reference (its launch shapes are fixed in
 * nep_simulator/NVT.py:46-111 and DescriptorGenerator.py:349-397). */
typedef struct nepx_tuning {
    int32_t fused_blocks_per_cu;   /* grid of the fused pair kernel = CUs x this (1..8; default 1)      */
    int32_t pair_blocks_per_cu;    /* grid cap of the descriptor / adjoint kernels (1..64; default 16)  */
    int32_t gather_lanes;          /* lanes per particle in the force gather: 4..64, power of 2 (16)   */
    int32_t net_blocks_per_cu;     /* grid cap of the split network kernel (1..16; default 2)          */
    int32_t pair_lanes;            /* lanes per pair in the fused kernel: 4 = site sums on MFMA where the
                                      shape allows (default), 1 = one lane per pair (site sums on VALU) */
} nepx_tuning;
int nepx_set_tuning(const nepx_tuning* t);
int nepx_get_tuning(nepx_tuning* t);

/* ---- network ----------------------------------------------------------- */
/* Bytes of the packed weight buffer for these widths. */
int64_t nepx_net_packed_floats(int32_t d_in, int32_t width, int32_t n_hidden);
/* Pack host float32 weights (PyTorch nn.Linear layout, W[out][in]) into the
 * MFMA operand order.  Replaces TheEnergyNet.load_state_dict + .to(cuda)
 * (ForceEvaluator.py:55-61).  ws/bs: n_hidden+2 host pointers each. */
int nepx_net_pack(int32_t d_in, int32_t width, int32_t n_hidden, const float* const* ws,
                  const float* const* bs, float* host_out);
/* fp16 hi/lo operands for NEPX_NET_F16X3 (width a multiple of 32); count in halves. */
int64_t nepx_net_packed_halves(int32_t d_in, int32_t width, int32_t n_hidden);
int nepx_net_pack_half(int32_t d_in, int32_t width, int32_t n_hidden, const float* const* ws,
                       uint16_t* host_out);
/* 16-bit operands for any precision other than NEPX_NET_F32 (fp16 or bf16, hi only or hi/lo;
 * bf16 rounded to nearest even), in the same operand order; count in 16-bit elements.  Passed as
 * nepx_net.w_half.  Same replacement as nepx_net_pack (ForceEvaluator.py:55-61). */
int64_t nepx_net_packed_lowp(int32_t precision, int32_t d_in, int32_t width, int32_t n_hidden);
int nepx_net_pack_lowp(int32_t precision, int32_t d_in, int32_t width, int32_t n_hidden, const float* const* ws,
                       uint16_t* host_out);

/* ---- descriptor, energy network, adjoint ------------------------------- */
/* P = min(*n_pairs_dev, max_pairs) is read on the device; max_pairs sizes the launch.  Up to about
 * 2 x 64 pairs per CU (descriptor, adjoint of multi-pass shapes) or 4 x 16 per CU (FP32 network)
 * the work of one pair is spread over several waves; the results are the same bit for bit. */
/* G (P x D, f32) of compacted pairs.  Replaces pos_to_pts + pts_to_nep_descriptors
 * (DescriptorGenerator.py:163-218,296-432); one row per pair (no FD stencil). */
int nepx_descriptor_fwd(const nepx_pair_model* m, const float* qf, const int32_t* pair_i,
                        const int32_t* pair_j, const float* pair_t, const int32_t* n_pairs_dev,
                        int64_t max_pairs, float* G, void* stream);
/* E (P) and dE/dG (P x D) by MFMA, f32 in / f32 accumulate.  Replaces
 * TheEnergyNet.forward + the forward-difference gradient (ForceEvaluator.py:76-113). */
int nepx_ffnn_fwd_bwd(const nepx_net* net, const float* G, const int32_t* n_pairs_dev,
                      int64_t max_pairs, float* E, float* dEdG, void* stream);
/* Hand-written adjoint: dE/dG -> per-pair (F_i, tau_i, tau_j), each clipped to
 * +-clip when clip > 0.  Replaces evaluate_gradients/evaluate_forces_torques
 * (ForceEvaluator.py:76-132).  out: P x 12 floats = {F,E}{tau_i,0}{tau_j,0}. */
int nepx_descriptor_adjoint(const nepx_pair_model* m, const float* qf, const int32_t* pair_i,
                            const int32_t* pair_j, const float* pair_t, const int32_t* n_pairs_dev,
                            int64_t max_pairs, const float* dEdG, const float* E, float clip,
                            float* out, void* stream);
/* Reference-compat finite differences (DescriptorGenerator.py:220-255,
 * PosToPtsKernels.py:153-446): expand P pairs into the 10 stencil rows
 * (row o*P+p: o=0 unperturbed, 1-3 i's sites + dx e_x/e_y/e_z, 4-6 i's sites
 * rotated by dteta about x/y/z through X_i, 7-9 j's sites rotated about X_j).
 * Writes row orientations q_rows (20P x 4: i then j of each row), row pair
 * indices (2r, 2r+1) and row translations t_rows (10P x 4), for nepx_descriptor_fwd. */
int nepx_stencil_rows(const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                      int64_t n_pairs, float dx, float dteta, float* q_rows, int32_t* row_i, int32_t* row_j,
                      float* t_rows, void* stream);
/* Forward-difference forces/torques from the 10P row energies, clipped at
 * +-clip (ForceEvaluator.py:86-132).  out: P x 12 floats as the adjoint writes. */
int nepx_fd_forces(const float* E_rows, int64_t n_pairs, float dx, float dteta, float clip, float* out,
                   void* stream);

/* Lanes per pair of the kernel nepx_pair_fused runs for this pair model and network under the
 * current tuning: 4 (site sums on matrix cores) or 1; 0 when there is no fused kernel (callers
 * then run nepx_descriptor_fwd -> nepx_ffnn_fwd_bwd -> nepx_descriptor_adjoint). */
int nepx_pair_fused_supported(const nepx_pair_model* m, const nepx_net* net);
/* Fused descriptor -> MFMA network -> adjoint for the step (no G in HBM). */
int nepx_pair_fused(const nepx_pair_model* m, const nepx_net* net, const float* qf,
                    const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                    const int32_t* n_pairs_dev, int64_t max_pairs, float clip, float* out,
                    void* stream);
/* nepx_pair_fused on the list's compacted pairs (pair_i, pair_j, pair_t, count flags[3]) and the
 * state's orientations; when the list carries rpos / pair_e / rrec it also writes each pair's j side
 * in reverse-row order (tagged with this filter pass), which nepx_forces_gather then reads
 * contiguously.  Same per-pair results and the same gathered sums as nepx_pair_fused.  Replaces
 * the descriptor + network + finite-difference part of ForceEvaluator.evaluate_interactions
 * (ForceEvaluator.py:71-146) as nepx_pair_fused does. */
int nepx_pair_fused_list(const nepx_pair_model* m, const nepx_net* net, const nepx_state* st, nepx_nlist* nl,
                         float clip, float* out, void* stream);

/* Descriptors of arbitrary site clouds, for training-set featurisation.
 * Replaces NEP.get_g_fast (05/NEP.py:171-219, 285-293) and the per-chunk loop of
 * pts_to_nep (04/Processors.py:48-88, 05/Processors.py:46-92).
 * clouds: n_clouds x n_sites x 3 floats (xyz interleaved), sites taken relative
 * to each cloud's centroid; site_w: n_sites per-site weights (the type
 * weights of NEP.set_pts_types) or NULL for 1; G: n_clouds x D floats with
 * D = (nrad+1) + (nang+1)*lmax.  Hypers with a kernel: (11,6,4), (10,8,4),
 * (17,6,6), (10,10,10); others return NEPX_EUNSUP. */
int nepx_cloud_descriptor(const float* clouds, int64_t n_clouds, int32_t n_sites, const float* site_w,
                          int32_t nrad, int32_t nang, int32_t lmax, float r_nep, float* G, void* stream);

/* ---- bead (label) potential -------------------------------------------- */
/* Soft bead-bead repulsion between rigid bodies, the label potential of the
 * reference's data generator (01/SamplerUtils.py:555-631, calc_en_unred):
 * E_ij = sum_{a in i, b in j} eps ((r0 - d_ab)/sigma)^2 for d_ab < r0, with
 * body-frame beads (n_beads x 3 floats, the same table for both bodies) placed
 * by the pair's orientations and minimum-image offset.  Pair inputs and the
 * per-pair output {F_i, E}{tau_i, 0}{tau_j, 0} are those of nepx_pair_fused, so
 * nepx_forces_gather and the integrator take the result unchanged.
 * The reference values are eps = 5, sigma = 1.1, r0 = 2^(1/6). */
#define NEPX_MAX_BEADS 256
int nepx_bead_pair_forces(const float* beads, int32_t n_beads, float eps, float sigma, float r0,
                          const float* qf, const int32_t* pair_i, const int32_t* pair_j, const float* pair_t,
                          const int32_t* n_pairs_dev, int64_t max_pairs, float clip, float* out, void* stream);

/* ---- neighbour list ---------------------------------------------------- */
/* Rebuild the half list within r_list of each system when flags[0] != 0
 * (always when force != 0).  Replaces the host cKDTree query
 * (Sim.py:85-92,192-198). */
int nepx_nlist_build(nepx_state* st, nepx_nlist* nl, int32_t force, void* stream);
/* Load an explicit (M,2) candidate list (i<j not required) into the list rows.
 * Replaces passing Nlist to generate_nep_descriptors (DescriptorGenerator.py:136-185). */
int nepx_nlist_from_pairs(nepx_state* st, nepx_nlist* nl, const int64_t* pairs, int64_t m,
                          void* stream);
/* Per-step COM filter d2 < rc^2 (FP64) and compaction into pair_i/pair_j/pair_t;
 * count in flags[3].  Replaces DescriptorGenerator.py:168-185. */
int nepx_pairs_filter(nepx_state* st, nepx_nlist* nl, void* stream);
/* nepx_pairs_filter without a workspace from the caller: the compaction's per-block
 * partial counts live in a static device array of the library (one per GPU, 256 KiB,
 * held until the process ends) instead of nl->blk_count, which may be NULL; every
 * other buffer is the caller's, as for nepx_pairs_filter, and the results are the
 * same.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc); it calls no allocating routine.  Limit: nl->cap <= 64 Mi slots
 * (NEPX_ECAP beyond).  Calls that overlap on one GPU share the array and race. */
int nepx_pairs_filter_noalloc(nepx_state* st, nepx_nlist* nl, void* stream);
/* Deterministic gather of per-pair outputs into F, T, upot (FP64).
 * Replaces the index_add_ scatter (ForceEvaluator.py:135-142). */
int nepx_forces_gather(nepx_state* st, nepx_nlist* nl, const float* pair_out, void* stream);

/* ---- integrator -------------------------------------------------------- */
/* One NVT step given F, T: integrate_step_two + integrate_step_one
 * (NVT.py:17-129) in FP64, then per-system kinetic energies, thermostat
 * update for the next step (update_variables) and the list rebuild test.
 * While nl->flags reports a list overflow the step does nothing, so the state
 * stays at the last good step until the caller has grown the list. */
int nepx_nvt_step(nepx_state* st, nepx_nlist* nl, void* stream);
/* Kinetic energies of the current state and, when update_thermostat != 0, the
 * thermostat update that the reference performs at the start of the first step
 * (Sim.py:147-169).  A run resumed from a saved thermostat state passes 0.
 * update_thermostat == 2 touches only systems whose record has pending != 0 (newly
 * admitted systems joining running ones) and clears the flag. */
int nepx_nvt_init(nepx_state* st, nepx_nlist* nl, int32_t update_thermostat, void* stream);
/* One part of the step, for callers that drive the reference's methods one by one:
 * NEPX_NVT_STEP_TWO  = integrate_step_two minus the thermostat update (NVT.py:17-64)
 * NEPX_NVT_TRANSLATE = translate1 (NVT.py:67-85)
 * NEPX_NVT_ROTATE    = rotate1 (NVT.py:87-106), using the kick stored by step two
 * NEPX_NVT_THERMO    = advance_thermostat1 (NVT.py:108-129) + the thermostat update
 * Running the four in order equals nepx_nvt_step. */
#define NEPX_NVT_STEP_TWO 2
#define NEPX_NVT_TRANSLATE 3
#define NEPX_NVT_ROTATE 4
#define NEPX_NVT_THERMO 5
int nepx_nvt_part(nepx_state* st, nepx_nlist* nl, int32_t part, void* stream);

/* One full MD step: [list rebuild if needed] -> filter -> fused pair kernel ->
 * gather -> NVT.  Replaces Sim.step (Sim.py:257-285).  pair_out: cap x 12 floats. */
int nepx_md_step(nepx_state* st, nepx_nlist* nl, const nepx_pair_model* m, const nepx_net* net,
                 float clip, float* pair_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NEPX_H */
