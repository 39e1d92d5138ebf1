// Internal definitions shared by the translation units of liblpm_mi355x.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lpm_mi355x.h"

// Compiled limits, checked on the host before any launch.
#define CG_VEC_WG 1024  // most workgroups of the CG vector kernels (grid-stride beyond)
#define LPM_MAXB 32    // bonds per particle (hex 12, SC/FCC 18, BCC 14)
#define LPM_BOND_GP 8  // consecutive particles per workgroup of the bond kernels (their LDS is sized by
                       // the most bonds of any such group, gp_bonds)
#define LPM_MAXC 512   // 2-hop candidates per particle before de-duplication
#define LPM_NSV 7      // state variables per particle (J2 uses 7)
#define LPM_NMAT 10    // model parameters per particle

#define LPM_EPS 1e-6   // include/lpm.h:37

// Device scalar slots (double) used by the CG and the reductions.
enum {
    SC_RHS2 = 0,   // ||b||^2
    SC_THR,        // max(tol^2 ||b||^2, tiny)
    SC_RR,         // ||r||^2 at the last stopping test
    SC_ITER,       // CG updates done
    SC_DONE,       // 0 running, 1 converged, 2 iteration cap, 3 zero rhs
    SC_MAXIT,
    SC_NORM_R,     // residual ||.||^2
    SC_NORM_REAC,  // reaction ||.||^2
    SC_DIAGNORM,   // ||diag K||^2
    SC_FLAG,       // any-flag (bond broken / damage changed) as a double 0/1
    SC_GAMMA_P,    // gamma = (r, M^-1 r) of iterations with parity 0 / 1 (two slots)
    SC_ALPHA_P = SC_GAMMA_P + 2,  // alpha of iterations with parity 0 / 1 (two slots)
    SC_IT0 = SC_ALPHA_P + 2,      // first iteration of the current batch of CG iterations
    SC_TOL,                       // relative tolerance of the current solve
    SC_COUNT = 32
};

struct Group {
    int32_t *ids = nullptr;
    int64_t n = 0;
    bool dup = false;  // duplicate ids: applied by one thread in list order
};

// One event of a recorded fatigue load path (lpm_fatigue_record): a state-variable update at a
// converged state (a checkpoint of the state, and s0, s1, s2 of every particle before and after
// the update: 6 per particle) or a fatigue damage update (its step, excluded type and the damage
// it left: 1 per particle)
struct FatEvent {
    enum { STATE = 1, UPDATE = 2 };
    int kind = 0;
    double dt = 0.0;
    int undamaged = -1;
    void *ckpt = nullptr;
    double *data = nullptr;
};

struct lpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // lattice
    int lattice = 0, dim = 3, model = 0, nneighbors = 1;
    double radius = 0, volume = 0, cutoff[2] = {0, 0};
    int64_t n = 0, nbond = 0, nconn = 0, nnl = 0;
    int maxconn = 0, maxbond = 0;
    int gp_bonds = 0;  // most bonds of any group of LPM_BOND_GP consecutive particles (LDS of the bond kernels)
    int asm_hops = 0;  // most 2-hop bond records of any particle (LDS of the assembly)
    // Assembly plan (mode "asm_plan", k_asm_plan): per row, its blocks by kind and each block's
    // 2-hop paths; built from the pattern on the first planned assembly
    int asm_plan = 1;
    bool asm_unplanned = false;       // the pattern cannot be planned (a partner repeats in a layer)
    // planned-assembly launch: wavefronts (rows) per workgroup, 1, 2 or 4 as the LDS allows; with
    // asm_tune the fastest is measured on the first planned assembly of each precision and kept
    int asm_tune = 1, asm_wpb = 0, asm_tune_key = -1;
    double asm_tuned_ms = 0;          // mean time of the planned assembly in the kept configuration
    int32_t *pl_row = nullptr, *pl_slot = nullptr, *pl_rng = nullptr, *pl_b3 = nullptr;
    uint8_t *pl_own = nullptr;
    // modes
    int diag10_literal = 1, cg_transpose = 1, dirichlet_static = 0, rollback_pex = 0;
    int precision = 0, precond = 0, verbose = 0, max_NR_iter = 30, cg_batch = 16;
    // cg_batch_auto: the first batch of a Newton solve covers the iterations the solve at the same
    // Newton index took last time (+2), and later batches hold cg_batch_next (0: cg_batch);
    // cg_slot is that index (0..3, -1 outside a Newton loop) and cg_hist the iterations recorded
    // per index (-1: none yet)
    int cg_batch_auto = 1, cg_batch_next = 6, cg_slot = -1;
    int cg_hist[4] = {-1, -1, -1, -1};
    int spmv_nt = 1;       // 1: the SpMV reads the matrix stream with non-temporal loads (0: plain loads)
    int nl_cells = 1;      // 1: the nonlocal search bins particles on a grid (0: every pair is tested)
    double cg_tol = 2.220446049250313e-16, tol_NR = 1e-6;
    int cg_forcing = 0;  // 1: per-Newton-iteration CG tolerance from the remaining reduction
    int64_t cg_maxit = -1;
    // topology (device)
    int32_t *bptr = nullptr, *bsplit = nullptr, *p2 = nullptr, *opp = nullptr, *nbarr = nullptr;
    int32_t *cptr = nullptr, *cidx = nullptr, *tslot = nullptr;
    // SELL-64 copy of the pattern used by the CG matrix (see reduce.h)
    int64_t *sp = nullptr;
    int32_t *scol = nullptr;
    int16_t *scol16 = nullptr;        // the same columns as offsets j - i when all fit in int16 (else null)
    int spmv_idx16 = 1;               // mode: the SpMV reads scol16 when it exists
    int64_t nslice = 0, nslots = 0;
    int64_t k_version = 0;            // assembly generation
    int64_t precond_version = -1;     // K generation the Jacobi inverse was formed from
    int64_t coarse_check_version = -1;  // K generation whose eliminated-DoF mask coarse_check formed
    bool coarse_check_diff = false;     // and whether it queued the comparison with the last build's
    int precond_key = -1;             // and its precision / preconditioner kind
    // SpMV launch configuration: slots unrolled per lane, wavefronts per workgroup, XCD remap;
    // with spmv_tune the fastest of all configurations is measured on the first solve of each
    // matrix kind (store, precision) and kept (the pattern never changes)
    int spmv_unr = 2, spmv_wpb = 4, spmv_remap = 1, spmv_tune = 1, tune_key = -1;
    int spmv_coop = 0, spmv_spw = 1;  // cooperative slices: wavefronts per slice (0: one wavefront each), slices per workgroup
    double tune_best_ms = 0;  // mean time of the fused SpMV in the configuration the tuner kept
    // Two-level preconditioner (precond == 2): Jacobi plus an additive coarse correction on
    // rigid-body modes of cubic aggregates of particles, M^-1 = D^-1 + P Ac^-1 P^T, Ac = P^T K P.
    int coarse_cells = 0;             // aggregate edge in lattice spacings (0: automatic)
    int coarse_budget = 768;          // automatic size: the smallest cells whose coarse DoFs per group fit this
    int nagg = 0, nmode = 0, ncoarse = 0;
    int agg_max = 0;                  // members of the largest aggregate (sizes the restriction's workgroups)
    int32_t *agg_of = nullptr, *agg_ptr = nullptr, *agg_idx = nullptr;  // particle -> aggregate; members
    double *agg_d = nullptr;          // x0_i - centre of its aggregate (dim entries per particle)
    int32_t *ra_ptr = nullptr, *ra_id = nullptr;  // distinct aggregates of the columns of each row
    int64_t nra = 0;
    int32_t *cq_ptr = nullptr, *cq_f = nullptr, *cq_gh = nullptr;  // coarse blocks: contributing row entries
    int64_t ncq = 0;
    double *cvbuf = nullptr;          // per (row, aggregate) contributions P_i^T K_ij P_j (nmode^2 each)
    // Ac is block diagonal, one dense block per coarse group (a specimen of an ensemble, or the
    // whole system): block b covers coarse DoFs [cgrp_cb[b], cgrp_cb[b+1]) and is stored row-major
    // at cgrp_off[b] in each of the two Gauss-Jordan buffers.
    int ncgrp = 1, cgrp_nbmax = 0;
    int64_t ac_elems = 0;             // sum of the squared block sizes
    int32_t *cgrp_cb = nullptr, *agg_grp = nullptr;  // block bases (ncgrp + 1) / group of each aggregate
    int64_t *cgrp_off = nullptr;      // block offsets (ncgrp + 1)
    double *Ac = nullptr;             // Gauss-Jordan ping-pong buffers, block inverses, row panels, diagonal, flags
    double *Ainv = nullptr;           // the buffer of Ac that holds Ac^-1 (inside Ac)
    double *cz = nullptr, *cy = nullptr;  // restriction P^T r / correction Ac^-1 P^T r (ncoarse each)
    double *agg_dm = nullptr;         // agg_d in member order (agg_idx)
    hipGraphExec_t gj_graph = nullptr;  // the inversion's launches, captured once (coarse_graph = 1)
    // CG batch graph (mode "cg_graph"): one batch of CG iterations captured once and replayed;
    // cg_key lists every value its launches captured (grids, launch configuration, buffers)
    int cg_graph = 0;
    hipGraphExec_t cg_exec = nullptr;
    std::vector<uintptr_t> cg_key;
    int coarse_graph = 0;             // 1: replay the inversion from a HIP graph, 0: launch directly
    // Load-step predictor (mode "predictor"): free DoFs start each step from the converged
    // position extrapolated from the previous converged steps: 1 linearly (plus the previous
    // increment), 2 quadratically (through the last three converged states) once they exist
    int predictor = 0;
    int pred_hist = 0;                // converged states held before the last: 1 xyz_prev, 2 also xyz_prev2
    double *xyz_prev = nullptr, *xyz_prev2 = nullptr;
    // Newton tolerance multipliers max(||r0||, ||R0||) evaluated at the reference's start (the
    // converged positions with the step's boundary conditions, before the prediction), one per
    // specimen (one for a single system); used by the next Newton call only, then cleared
    std::vector<double> tolm_ref;
    bool tolm_pending = false;  // one specimen: the reference-start norms are on their way to h_scal[SC_COUNT..+1]
    uint8_t *dofmask = nullptr;       // per particle: bit k set when DoF k is eliminated
    int64_t coarse_version = -1;
    // mode "coarse_reuse": keep the coarse space and Ac^-1 across assemblies while the eliminated
    // DoFs stay the same and the CG iteration count stays within 1.25x (+4) of the first solve
    // after the last build
    int coarse_reuse = 0;
    bool coarse_built = false, coarse_stale = false;
    int coarse_ref_iters = -1;
    int64_t st_coarse_builds = 0;
    uint8_t *dofmask_prev = nullptr;
    // Asynchronous dump writer (io.hip): device snapshot -> pinned host copy on a side stream ->
    // formatted and appended by a writer thread; one write in flight per context
    struct DumpState *dump = nullptr;
    // Load-step checkpoints (lpm_checkpoint_save / restore): one device buffer per slot with the
    // state a converged step leaves (positions, loads, constraint flags, particle and bond state
    // and their last-converged copies); null where the slot is empty
    std::vector<void *> ckpt;
    // per-bond scratch of the force JVP / VJP (4 doubles per bond), allocated on first use
    double *adj_scr = nullptr;
    // the last fatigue damage update (SolverFatigue::updateFatigueDamage): positions its nonlocal
    // weights used and each particle's energy range s2 - s1 (the update resets both), its cycle
    // step and excluded type, for the fatigue-parameter sensitivity
    double *fat_xyz = nullptr, *fat_range = nullptr;
    double fat_dt = 0.0;
    int fat_undamaged = -1;
    bool fat_valid = false;
    // fatigue load path record (lpm_fatigue_record)
    bool fat_rec = false;
    std::vector<FatEvent> fat_events;
    // Specimen ensemble: independent particle systems (contiguous id ranges, no bond between
    // them) solved in the same launches, each with its own Newton test.
    int nspec = 0;                    // 0: one system
    std::vector<int64_t> spec_first;  // nspec + 1 particle offsets
    int32_t *spec_of = nullptr;       // specimen of each particle (device)
    double *spec_w = nullptr;         // per-specimen factor (device): rhs scale / displacement scale
    int32_t *spec_blk = nullptr;      // first specimen-residual block of each specimen (device, nspec + 1)
    int64_t *spec_first_d = nullptr;  // spec_first on the device
    int spec_nblk = 0;
    double *spec_part = nullptr;      // per-block partials of the specimen residual norms (2 x spec_nblk)
    std::vector<int> spec_newton;     // Newton iterations of each specimen in the last step
    std::vector<char> spec_ok;        // converged in the last step
    int32_t *nlptr = nullptr, *nlidx = nullptr;
    // particle state (device, FP64 unless noted)
    double *xyz = nullptr, *xyz0 = nullptr, *xyz_last = nullptr, *Pin = nullptr, *Pex = nullptr;
    double *stress = nullptr, *damage = nullptr, *damage_last = nullptr, *damage_visual = nullptr;
    double *Ddot_local = nullptr, *Ddot_nonlocal = nullptr, *nonlocal_L = nullptr;
    double *sv = nullptr, *sv_last = nullptr, *mat = nullptr;
    double *dLe_total = nullptr, *TdLe_total = nullptr, *cs_sum = nullptr;
    int32_t *type = nullptr, *cons = nullptr;
    // bond state (device)
    double *dis0 = nullptr, *dis = nullptr, *dis_last = nullptr, *dLe = nullptr, *dLp = nullptr, *dLp_last = nullptr;
    double *bstrain = nullptr, *Kn = nullptr, *Tv = nullptr, *cs = nullptr, *bforce = nullptr, *bforce_last = nullptr;
    double *bdamage = nullptr, *bdamage_last = nullptr;
    // linear system
    double *residual = nullptr, *disp = nullptr, *diag64 = nullptr;
    void *aval = nullptr;        // BSR values of the CG matrix (float or double), row-SoA per particle row
    int aval_prec = -1;          // precision aval was allocated for
    void *invdiag = nullptr;     // Jacobi (ndof) or block Jacobi (9 per particle)
    void *v_x = nullptr, *v_r = nullptr, *v_u = nullptr, *v_p = nullptr, *v_s = nullptr, *v_w = nullptr;  // padded vectors
    int vpad = 3;                // entries per particle in the CG vectors (3 in 3D, 2 in 2D)
    double *scal = nullptr;      // SC_COUNT device scalars
    double *partial = nullptr;   // per-block partial sums (3 x partial_cap)
    int partial_cap = 0;
    unsigned int *counter = nullptr;  // last-block counters (4)
    int32_t *iflag = nullptr;    // device int flags (4)
    double *h_scal = nullptr;    // pinned host mirror of scal (SC_COUNT + 2: the reference-start norm pair)
    int32_t *h_iflag = nullptr;
    bool k_ready = false;
    // Finite-difference stiffness (mode "stiffness_fd", StiffnessMode::FiniteDifference): a
    // colouring in which no two particles of a class share a conns list (so one perturbed pass
    // per class and axis gives one block column per row), the classes' members, and the saved
    // positions and base forces of the current assembly; built on the first FD assembly
    int stiffness_fd = 0;
    int fd_ncolor = 0;
    uint8_t *elim = nullptr;  // per particle, bit k: DoF k is eliminated by the Dirichlet pass; then n row marks
    int32_t *fd_color = nullptr, *fd_ids = nullptr;  // class of each particle / members by class
    std::vector<int64_t> fd_ptr;                     // class offsets into fd_ids (host)
    double *fd_x0 = nullptr, *fd_P0 = nullptr;
    // groups
    std::vector<Group> groups;
    // statistics
    int64_t st_newton = 0, st_cg_iters = 0, st_cg_solves = 0, st_asm = 0, st_spmv = 0, st_spmv_timed = 0;
    double st_spmv_ms = 0, last_cg_err = 0;
    int last_cg_iters = 0, timing = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

int lpm_fail(lpm_ctx *c, int code, const std::string &msg);

#define LPM_HIP(c, call)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return lpm_fail((c), LPM_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---- launchers (kernels_phys.hip)
int launch_geometry_force(lpm_ctx *c, bool force);
int launch_bond_force(lpm_ctx *c);
int launch_particle_force(lpm_ctx *c);
int launch_residual(lpm_ctx *c, double *nr2, double *nR2);
int launch_assemble(lpm_ctx *c);
int launch_assemble_fd(lpm_ctx *c);  // finite-difference blocks (mode "stiffness_fd")
void fd_free(lpm_ctx *c);
int launch_dirichlet(lpm_ctx *c);
int launch_state_var(lpm_ctx *c);
int launch_nonlocal_ddot(lpm_ctx *c, int undamaged);
int launch_static_damage(lpm_ctx *c, int undamaged, int *any);  // any == nullptr: flag copied, read by flag_wait
int flag_wait(lpm_ctx *c, int *any);  // waits for the stream, then reads the flag launch_static_damage copied
int launch_fatigue_damage(lpm_ctx *c, double dNdt, int undamaged, int *any);
int launch_pattern(lpm_ctx *c);
// joins a dump write still in flight and frees the writer's buffers (lpm_destroy)
void dump_free(lpm_ctx *c);
// bond search, reverse bonds and initial bond geometry on the device (lpm_create without bonds)
int device_topology(lpm_ctx *c, const double *xyz_host, const lpm_cut *cuts, int ncut);
int launch_tslot(lpm_ctx *c);
int launch_nonlocal_search(lpm_ctx *c, double ratio);
int launch_residual_spec(lpm_ctx *c, std::vector<double> &r2, std::vector<double> &R2);
int launch_spec_scale(lpm_ctx *c, double *v);  // v[dof] *= spec_w[specimen]
// ---- launchers (kernels_cg.hip)
int launch_finalize(lpm_ctx *c, int kind, int na);
// k_finalize on caller-given buffers (no context): partial [2 x cap], scal [SC_COUNT]
hipError_t launch_finalize_raw(hipStream_t stream, int kind, int na, const double *partial, int cap, double *scal);
int cg_alloc(lpm_ctx *c);
int cg_solve(lpm_ctx *c, int *iters, double *error, int *success);
int launch_move(lpm_ctx *c);
int launch_scale_x(lpm_ctx *c);  // x_i *= spec_w[specimen of i]
int spmv_bench(lpm_ctx *c, int reps, double *ms);
int launch_precond_setup(lpm_ctx *c);
int coarse_check(lpm_ctx *c);  // queues the coarse-reuse mask comparison (kernels_coarse.hip)
// precond 1 smooths with the inverse 3x3 (2x2) diagonal blocks, 0 and 2 with the inverse
// diagonal; 2 adds the coarse correction
static inline bool block_jacobi(const lpm_ctx *c) { return c->precond == 1; }
static inline bool two_level(const lpm_ctx *c) { return c->precond == 2; }
// ---- two-level preconditioner (kernels_coarse.hip)
int coarse_prepare(lpm_ctx *c);
int launch_coarse_apply(lpm_ctx *c);  // cz = P^T r, cy = Ac^-1 cz (skipped once the CG is done)
void coarse_free(lpm_ctx *c);
int copy_bench(lpm_ctx *c, int64_t bytes, int reps, double *ms);
// ---- host phases (ctx.hip)
int geometry_force_state(lpm_ctx *c);
int calc_stiffness(lpm_ctx *c);
int update_stiffness_disp_bc(lpm_ctx *c);
// ---- tangent and adjoint (kernels_adj.hip)
int force_jvp(lpm_ctx *c, const double *v, double *jv);
int force_vjp(lpm_ctx *c, const double *w, double *y);
int adjoint_reaction(lpm_ctx *c, int gid, int axis, double *gKn, double *gTv, double *F, int *refinements,
                     double *rel_residual);
int adjoint_linear(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gKn, double *gTv, double *gBd, double *G,
                   int *refinements, double *rel_residual, double *gBd_dev = nullptr);
int adjoint_fatigue_params(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gA, double *gB, double *gd,
                           double *G, int *refinements, double *rel_residual);
int adjoint_fatigue_path(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gA, double *gB, double *gd,
                         double *G, int *solves, double *max_rr);
int fat_record_state(lpm_ctx *c, bool after);
int fat_record_update(lpm_ctx *c, double dt, int undamaged);
void fat_record_clear(lpm_ctx *c);
// checkpoints into a buffer of the context's own (allocated by save when *buf is null)
int ckpt_save_buf(lpm_ctx *c, void **buf);
int ckpt_restore_buf(lpm_ctx *c, void *buf);

// helpers
static inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }
