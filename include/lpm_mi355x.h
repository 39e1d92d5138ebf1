/*
 * lpm_mi355x.h -- C ABI of the MI355X (gfx950) implementation of the LPM-MAC
 * static / fatigue Newton load-step path.
 *
 * The reference (longfish/LPM-MAC) has no FFI: its only seam is the header-only
 * C++ template API (the headers under include/). Every entry point below replaces one method
 * of that API; the citation names the reference file:line whose behaviour it
 * provides. The C++ drop-in headers (lpm-mac_amd/host/) call these functions
 * from the reference's own class names, and tests/bench call them through
 * ctypes.
 *
 * Conventions
 *  - Plain pointers and sizes; no C++ or torch types cross the ABI.
 *  - Every function returns 0 on success and a negative LPM_E* code on error;
 *    lpm_last_error() gives a message. No exception crosses the ABI.
 *  - Host arrays are borrowed for the duration of the call only.
 *  - One context = one particle system (one reference Assembly) on one GPU.
 *    Calls on a context are ordered on its own HIP stream; functions that
 *    return host values synchronise that stream.
 *  - Particle ids are 0..n-1 (the reference's DoF index = particle id,
 *    solver.h:102,117). Positions are xyz[3*i+k] (FP64) also in 2D (z kept).
 */
#ifndef LPM_MI355X_H
#define LPM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes */
#define LPM_OK 0
#define LPM_E_ARG (-1)      /* invalid argument / inconsistent sizes */
#define LPM_E_TOPO (-2)     /* topology rejected (asymmetric bond, duplicate, out of range) */
#define LPM_E_HIP (-3)      /* HIP runtime error */
#define LPM_E_STATE (-4)    /* call out of order (e.g. solve before assemble) */
#define LPM_E_NOMEM (-5)
#define LPM_E_LIMIT (-6)    /* a row or list exceeds a compiled limit */

/* ---- enums (same values as the reference's enum classes, include/lpm.h:56-96) */
enum { LPM_SQUARE2D = 0, LPM_HEXAGON2D = 1, LPM_SC3D = 2, LPM_FCC3D = 3, LPM_BCC3D = 4 };
enum { LPM_ELASTIC = 0, LPM_ELASTIC_DAMAGE = 1, LPM_J2PLASTICITY = 2, LPM_FATIGUE_HCF = 3 };
enum { LPM_RELATIVE = 0, LPM_ABSOLUTE = 1 };                 /* LoadMode */
enum { LPM_TIME_LINEAR = 0, LPM_TIME_EXPONENTIAL = 1 };       /* TimeMapMode */
enum { LPM_LOAD_UNIAXIAL_DISP = 0, LPM_LOAD_UNIAXIAL_FORCE = 1,
       LPM_LOAD_DOGBONE_DISP = 2, LPM_LOAD_CT_FORCE = 3 };    /* FatigueLoadType, load_step_fatigue.h:13-19 */

typedef struct lpm_ctx lpm_ctx;

/* ------------------------------------------------------------------ setup (host C++ inside the library)
 * Lattice points; replaces createPlateSQ2D / createPlateHEX2D / createCuboidSC3D /
 * createCuboidFCC3D with createRMatrix (include/utilities.h:8-301).
 * box[6] = {xmin,xmax,ymin,ymax,zmin,zmax}; angles in radians. Writes at most cap points
 * into xyz (3 doubles each) and returns the number of points in *n. */
int lpm_generate_lattice(int lattice, const double *box, double radius, int eulerflag, const double *angles,
                         double *xyz, int64_t cap, int64_t *n);
/* Same with the rotation given as R[3*row+col] (the reference passes createRMatrix's result). */
int lpm_generate_lattice_r(int lattice, const double *box, double radius, const double *R, double *xyz, int64_t cap,
                           int64_t *n);
/* createRMatrix (include/utilities.h:8-82): eulerflag 0 X1-Y2-Z3, 1 Kocks Z1-Y2-Z3, 2 Bunge Z1-X2-Z3. */
int lpm_rotation_matrix(int eulerflag, const double *angles, double *R);

/* Directed bonds with the reference's ordering; replaces Assembly::createBonds
 * (include/assembly.h:212-234): for each particle, layer-0 bonds then layer-1 bonds,
 * each in increasing partner id; dis < 1.01*cutoff[1], layer 0 if dis < 1.01*cutoff[0].
 * Pass bond_p2 == NULL to get the count in *nbonds first. bond_ptr has n+1 entries,
 * bond_split (first layer-1 bond of each particle) n entries. */
int lpm_build_bonds(int lattice, double radius, int64_t n, const double *xyz, int32_t *bond_ptr,
                    int32_t *bond_split, int32_t *bond_p2, int64_t cap, int64_t *nbonds);

/* Unit-cell constants; replaces UnitCell (include/unit_cell.h:7-127).
 * out[0]=dim out[1]=nneighbors out[2]=nneighbors1 out[3]=nneighbors2 out[4]=nneighbors_AFEM
 * out[5]=cutoff0 out[6]=cutoff1 out[7]=particle_volume; map[9] = el_mapping (row r, col c at
 * map[3r+c]) as the reference reads it when literal != 0 (column-major Eigen::Map of the
 * row-major source array, unit_cell.h:49) or as written when literal == 0. */
int lpm_unit_cell(int lattice, double radius, int literal, double *out, double *map);

/* ------------------------------------------------------------------ context
 * Particle system on the device; replaces Assembly's per-particle / per-bond objects
 * (particle.h:13-66, bond.h:12-73) with a structure-of-arrays layout in HBM. The bond
 * list must be symmetric (every bond i->j of layer l has a reverse j->i of layer l; the
 * reference dereferences a missing one, particle.h:207-212) and duplicate-free. The
 * 2-hop sparsity pattern (Assembly::updateConnections, assembly.h:236-266) is built on
 * the device from the bonds. */
typedef struct {
    int32_t lattice;
    double radius;
    int32_t model;             /* LPM_ELASTIC .. LPM_FATIGUE_HCF (Assembly ptype) */
    int64_t n;                 /* particles */
    const double *xyz;         /* [3n] initial positions (also xyz_initial / xyz_last) */
    const int32_t *bond_ptr;   /* [n+1], or NULL: search the bonds on the device (see lpm_create) */
    const int32_t *bond_split; /* [n] first layer-1 bond of each particle */
    const int32_t *bond_p2;    /* [bond_ptr[n]] partner ids */
    int32_t device;            /* HIP device ordinal */
} lpm_topology;

/* With bond lists: validated on the host (Assembly(xyz, bonds), assembly.h:88-106 with
 * readBond). With bond_ptr == NULL: Assembly::createBonds (assembly.h:212-234) runs on the
 * device with the same criterion, layer test and order as lpm_build_bonds (a cell grid, each
 * particle's lists sorted by partner id), followed by the reverse bonds and the initial bond
 * geometry, also on the device. The device search bonds every pair within the cutoff, so
 * ensembles whose specimens overlap in space pass their own lists. */
int lpm_create(const lpm_topology *topo, lpm_ctx **out);

/* ------------------------------------------------------------------ geometry preprocessing
 * The reference's geometry tools (geometry/geo1_CT_configuration.cpp, geo2_dogbone_config2d.cpp)
 * cut holes and notches out of a lattice plate: isValid (:137-171) keeps the points in no cut,
 * and searchNeighbor (:79-134) keeps a bond only if the midpoints of its 10 equal sub-segments
 * are all valid. A cut (z ignored):
 *   LPM_CUT_CIRCLE         (x-cx)^2 + (y-cy)^2 <= size^2                (inCircle, :45-49)
 *   LPM_CUT_NOTCH          |y-ty| <= size, between the lines of slope -+1/sqrt(3) through the
 *                          tip on the side x > tx                        (inNotch, :52-62)
 *   LPM_CUT_REVERSE_NOTCH  the same opening to x < tx                   (inReverseNotch, :64-74) */
enum { LPM_CUT_CIRCLE = 0, LPM_CUT_NOTCH = 1, LPM_CUT_REVERSE_NOTCH = 2 };
typedef struct {
    int32_t kind;
    double x, y;  /* centre or notch tip */
    double size;  /* radius or notch half width */
} lpm_cut;
/* keep[i] = 1 if particle i lies in no cut (the tools' particle filter) */
int lpm_cut_particles(const double *xyz, int64_t n, const lpm_cut *cuts, int32_t ncut, uint8_t *keep);
/* lpm_create with the device bond search (topo->bond_ptr must be NULL) dropping every bond the
 * tools' searchNeighbor drops; the kept lists are in the same order. */
int lpm_create_cut(const lpm_topology *topo, const lpm_cut *cuts, int32_t ncut, lpm_ctx **out);
/* benchmark/AM_Ti64/generate_dump.ipynb createPlate: one particle per pixel equal to 255 of a
 * binary image (row-major, nrows x ncols, row 0 at the top), pixel edge `pixel`:
 * x = pixel/2 + pixel*col, y = nrows*pixel - (pixel/2 + pixel*row), z = 0, in row-major order.
 * Pass xyz == NULL to get the count in *n first. */
int lpm_image_particles(const uint8_t *img, int64_t nrows, int64_t ncols, double pixel, double *xyz, int64_t cap,
                        int64_t *n);
int lpm_destroy(lpm_ctx *ctx);
const char *lpm_last_error(const lpm_ctx *ctx); /* ctx may be NULL: last error of lpm_create */
/* sizes: "n","dim","nbonds","nconn","nnl","ndof","knnz","maxconn","maxbond","nl_ptr",
 * "nslots" (SELL-64 slots of the matrix) */
int64_t lpm_size(const lpm_ctx *ctx, const char *what);

/* Mode switches (each literal reference behaviour has value shown first).
 *  "diag10_literal"   1: diagonal block K[1][0] left 0 (derivative.h:38) | 0: = K[0][1]
 *  "cg_transpose"     1: solve K^T d = r like Eigen CG<Lower|Upper> on column-major K
 *                        (solver.h:195) | 0: solve K d = r
 *  "dirichlet_static" 0: static step does not eliminate constrained DoFs (solver_static.h:113)
 *                      | 1: eliminate (as the fatigue step does, solver_fatigue.h:288)
 *  "rollback_pex"     0: Pex kept on a step restart (assembly.h:150-159) | 1: restored
 *  "stiffness_fd"     0: analytic blocks (StiffnessMode::Analytical, stiffness.h:67-101) |
 *                        1: forward-difference blocks (StiffnessMode::FiniteDifference,
 *                        stiffness.h:103-154, particle.h:183-195): the members of one class of a
 *                        colouring (no two in one conns list) are moved by EPS*radius along one
 *                        axis together, every bond is re-evaluated, and each row takes
 *                        (Pin - Pin_base)/EPS/radius into the block of its moved partner; then
 *                        damage, state variables, bdamage and dLp return to their last converged
 *                        copies as the reference's restore leaves them. Positions are restored
 *                        exactly (the reference's (x + h) - h can leave them one ulp off).
 *                        LPM_E_ARG for J2Plasticity, whose reference FD pass carries the whole
 *                        return-mapping increment. get_mode("fd_colors") = classes once built.
 *  "precision"        0: FP32 K, SpMV and CG (FP64 residual, positions, state) | 1: FP64 everywhere
 *  "cg_tol"           relative CG tolerance (Eigen default 2.22e-16; FP32 runs use ~1e-5)
 *  "cg_maxit"         CG iteration cap (-1: 2*ndof, Eigen default)
 *  "cg_forcing"       0: every Newton iteration solves to cg_tol | 1: inexact Newton, the
 *                        solve of iteration k stops at max(cg_tol, min(0.1, tol_NR*tol_mult/
 *                        (2 ||r_k||))), i.e. it asks only for the reduction still needed to pass
 *                        the Newton test; the Newton test itself is unchanged
 *  "precond"          0: Jacobi (Eigen DiagonalPreconditioner) | 1: 3x3 block Jacobi |
 *                        2: Jacobi + additive coarse correction on the rigid-body modes of cubic
 *                        aggregates of particles (two-level; iteration count no longer grows
 *                        with the lattice width). Changes the CG's rate, not the equilibrium.
 *  "predictor"        1: a static step starts its free DoFs from the converged positions plus the
 *                        previous converged step's increment (after the step's boundary
 *                        conditions); Newton then runs to the same test. 2: from the quadratic
 *                        through the last three converged states (the increment plus its change),
 *                        linear while only two exist. 0 (default): the reference's start, the
 *                        converged positions. Reset by a halving restart and by uploading
 *                        positions.
 *  "coarse_graph"     1: replay the coarse inversion's launches from a HIP graph captured once |
 *                        0 (default): launch them directly
 *  "coarse_reuse"     1: keep the two-level coarse space (Ac and its inverse) across assemblies
 *                        while the eliminated DoFs are unchanged and the CG needs at most 1.25x
 *                        (+4) the iterations of the first solve after the last build; rebuilt
 *                        otherwise (stat "coarse_builds" counts builds) | 0 (default): rebuilt with
 *                        every assembly. K itself is reassembled either way.
 *  "coarse_rebuild"   1: the next assembly rebuilds the coarse space even under coarse_reuse (a
 *                        cold start); write-only
 *  "cg_graph"         1: capture one batch of CG iterations (cg_batch iterations: vector updates,
 *                        coarse correction, SpMV) into a HIP graph and replay it for every batch
 *                        of every solve; re-captured when sizes, launch configuration or buffers
 *                        change. Per-SpMV timing is off in this mode | 0 (default): launch directly
 *  "cg_batch"         CG iterations issued before the host reads the device's stop flag (rounded
 *                        up to even; default 16); iterations issued past the stop are no-ops, so
 *                        the batch sizes never change results
 *  "cg_batch_auto"    1 (default): a Newton solve starts with one batch covering the iterations
 *                        that the solve at the same Newton index of the previous Newton loop took,
 *                        plus 2, then continues with two batches of cg_batch_next and batches of
 *                        cg_batch after those | 0: batches of cg_batch throughout. Not used under
 *                        cg_graph. Every lpm_set_mode call forgets the recorded counts
 *  "cg_batch_next"    size of the two batches after such a sized first batch (rounded up to
 *                        even; default 6; 0: cg_batch)
 *  "coarse_cells"     aggregate edge in lattice spacings for precond 2 (0: smallest of 8, 9,
 *                        10, ... giving at most "coarse_budget" coarse DoFs per specimen);
 *                        get_mode("ncoarse") reads the coarse dimension once built
 *  "coarse_budget"    coarse DoFs per specimen the automatic size stays within (1..1024,
 *                        default 768); like coarse_cells, fixed once the coarse space exists
 *  "spmv_tune"        1: on the first solve of each matrix kind time every SpMV launch
 *                        configuration (two passes, the faster timing of each) and keep the
 *                        fastest one-wavefront-per-slice shape unless a cooperative shape is more
 *                        than 3% faster | 0: use the values below
 *  "spmv_unr" (1,2,4), "spmv_wpb" (wavefronts per workgroup 1,2,4,8), "spmv_remap" (0/1:
 *                        XCD-contiguous workgroup order), "spmv_coop" (0: one wavefront per
 *                        64-row slice | 8, 16: that many wavefronts share each slice),
 *                        "spmv_spw" (consecutive slices per cooperative workgroup) -- the SpMV
 *                        launch configuration
 *  "asm_plan"         1 (default): analytic assembly from a per-pattern plan (each row's blocks
 *                        grouped by kind with their 2-hop paths, built on the first assembly;
 *                        one block row per lane) | 0: the row-list kernel, which rebuilds each
 *                        row's path lists in LDS on every assembly. Both give the same blocks bit
 *                        for bit; patterns with a partner repeated within a layer use the row-list
 *                        kernel either way
 *  "asm_tune"         1 (default): on the first planned assembly of each precision time 1, 2 and 4
 *                        rows (wavefronts) per workgroup, as the LDS allows, and keep the fastest
 *                        (every configuration writes the same blocks) | 0: use "asm_wpb"
 *  "asm_wpb"          rows per workgroup of the planned assembly (1, 2, 4; 0 until tuned: the
 *                        most the LDS allows)
 *  "nl_cells"         1 (default): lpm_search_nonlocal bins the particles (of each specimen) on a
 *                        grid of cells no narrower than the largest cutoff and tests the 27 (9
 *                        in 2D) surrounding cells, then sorts each list by id | 0: every pair is
 *                        tested. Both give the same lists; the grid falls back to every pair
 *                        when a list exceeds 4096 entries
 *  "spmv_nt"          1 (default): the SpMV reads the matrix values and column indices with
 *                        non-temporal loads (read once per SpMV; the input vector keeps the
 *                        caches) | 0: ordinary loads
 *  "spmv_idx16"       1 (default): the SpMV reads its column indices as int16 offsets j - i when
 *                        every offset of the pattern fits (get_mode("spmv_index_bytes") = 2) |
 *                        0: int32 column ids (4)
 *  "max_NR_iter", "tol_NR"  Newton limits (Solver ctor arguments, solver.h:39-50)
 *  "verbose"          print per-iteration norms like the reference
 */
int lpm_set_mode(lpm_ctx *ctx, const char *name, double value);
double lpm_get_mode(const lpm_ctx *ctx, const char *name);

/* ------------------------------------------------------------------ state transfer
 * Field names follow the reference members: particle fields "xyz","xyz_initial","xyz_last",
 * "Pin","Pex","stress"(6/particle),"damage","damage_last","damage_visual","Ddot_local",
 * "Ddot_nonlocal","nonlocal_L","state_var","state_var_last"(7/particle),"mat"(10/particle:
 * model parameters, see lpm_set_material), "dLe_total","TdLe_total"(2/particle),
 * "cs_sum"(6/particle); int32 particle fields "type","disp_constraint"(3/particle),"nb";
 * bond fields "dis_initial","dis","dis_last","dLe","dLp","dLp_last","bstrain","Kn","Tv",
 * "cs"(3/bond),"bforce","bforce_last","bdamage","bdamage_last"; int32 "bond_ptr",
 * "bond_split","bond_p2","opp","conn_ptr","conn_idx","nl_ptr","nl_idx";
 * solver vectors "residual","disp"(dim*n),"reaction" (host NULL and count 0 return the reaction
 * count instead of 0). count = number of elements. */
int lpm_upload(lpm_ctx *ctx, const char *field, const void *host, int64_t count);
int lpm_download(lpm_ctx *ctx, const char *field, void *host, int64_t count);
/* Stiffness matrix actually handed to CG (K, or K^T when cg_transpose) as CSR over scalar
 * rows in FP64 (values converted from the stored precision). Sizes via lpm_size("knnz"). */
int lpm_download_matrix(lpm_ctx *ctx, int32_t *rowptr, int32_t *col, double *val, int64_t cap);

/* ------------------------------------------------------------------ material
 * Per-particle model parameters and per-bond Kn/Tv; replaces the setParticleProperty
 * overloads (particle_elastic.h:70-143, particle_elastic_damage.h:175-254,
 * particle_j2plasticity.h:154-201, particle_fatigue_hcf.h:162-209). id < 0: all particles.
 * params layout (mat, 10 doubles per particle):
 *   Elastic        [0]=critical_bstrain
 *   ElasticDamage  [0]=k0 [1]=k1 [2]=damage_threshold [3]=comp_tensile_ratio
 *   J2Plasticity   [0]=E [1]=mu [2]=sigmay [3]=xi [4]=H [5]=A [6]=damage_threshold [7]=critical_bstrain
 *   FatigueHCF     [0]=A [1]=B [2]=d [3]=E [4]=mu [5]=damage_threshold [6]=fatigue_limit_ratio
 * Ce = (C11, C12, C44); KnTv = el_mapping * Ce (x radius in 3D); elmap_literal selects the
 * reference's column-major read (1) or the row-major intent (0). */
int lpm_set_material(lpm_ctx *ctx, int64_t id, double nonlocal_L, const double *Ce, int elmap_literal,
                     const double *params, int32_t nparams);
/* Ce from (E, mu, plane_stress) exactly as particle_elastic.h:77-88. */
int lpm_ce_from_e(int plane_stress, double E, double mu, double *Ce);

/* Nonlocal neighbours within cutoff_ratio * nonlocal_L (Assembly::searchNonlocalNeighbors,
 * assembly.h:117-130), built on the device. */
int lpm_search_nonlocal(lpm_ctx *ctx, double cutoff_ratio);

/* ------------------------------------------------------------------ boundary conditions
 * Particle groups live on the device; replaces the std::vector<Particle*> groups of DispBC /
 * ForceBC (boundary.h:9-44). */
int lpm_group_create(lpm_ctx *ctx, int64_t n, const int32_t *ids, int32_t *gid);

typedef struct { int32_t group; int32_t mode; int32_t flag; double step; } lpm_disp_bc;   /* flag 'x','y','z' */
typedef struct { int32_t group; int32_t mode; double fx, fy, fz; } lpm_force_bc;
typedef struct {
    int32_t n_disp;
    const lpm_disp_bc *disp;
    int32_t n_force;
    const lpm_force_bc *force;
} lpm_load_step;

/* Solver::updateForceBC then Solver::updateDisplacementBC (solver.h:130-190) */
int lpm_apply_bc(lpm_ctx *ctx, const lpm_load_step *step);

/* ------------------------------------------------------------------ per-phase entry points
 * Assembly::updateGeometry (assembly.h:132-138; particle.h:163-181; bond.h:30-39) */
int lpm_update_geometry(lpm_ctx *ctx);
/* Assembly::updateForceState (assembly.h:178-191): bond forces of the model
 * (particle_elastic.h:38-50 and the damage variants), Pin, stress, damage_visual */
int lpm_update_force_state(lpm_ctx *ctx);
/* Solver::updateRR (solver.h:108-128) plus the two norms NewtonIteration uses (:61-63) */
int lpm_update_rr(lpm_ctx *ctx, double *norm_residual, double *norm_reaction);
/* Stiffness::reset + calcStiffness (stiffness.h:43-48, 228-263) with localStiffnessANA /
 * fdu2dxyz* (stiffness.h:67-101, derivative.h:7-239), into BSR on the device pattern */
int lpm_calc_stiffness(lpm_ctx *ctx);
/* Stiffness::updateStiffnessDispBC (stiffness.h:265-294) */
int lpm_update_stiffness_disp_bc(lpm_ctx *ctx);
/* Solver::solveLinearSystem (solver.h:91-106): LPM_CG (solver.h:192-217) then move every
 * particle by its DoFs of disp. iters/error as Eigen reports them (may be NULL). */
int lpm_solve_linear_system(lpm_ctx *ctx, int32_t *iters, double *error);
/* Solver::NewtonIteration (solver.h:53-89); *n_iter = return value of the reference */
int lpm_newton_iteration(lpm_ctx *ctx, int32_t *n_iter);
/* The computation of lpm_update_rr (Solver::updateRR with the two norms NewtonIteration takes)
 * on caller-owned device arrays, allocating nothing. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).
 * Pex, Pin: FP64 [3n]; cons: int32 [3n] (disp_constraint); residual: FP64 [dim*n], written;
 * norms: host double[2] = ||residual||, ||reaction|| (may be NULL: no host synchronisation);
 * stream: hipStream_t (NULL: the default stream). The per-block partial sums live in a device
 * array the library holds once per GPU (512 KB): n*dim above LPM_NOALLOC_MAX_DOFS returns
 * LPM_E_LIMIT. Calls that overlap on one GPU race on that array. Errors: lpm_last_error(NULL). */
#define LPM_NOALLOC_MAX_DOFS 8388608
int lpm_update_rr_noalloc(int64_t n, int32_t dim, const double *Pex, const double *Pin, const int32_t *cons,
                          double *residual, double *norms, void *stream);
/* Assembly::updateStateVar (assembly.h:161-166) */
int lpm_update_state_var(lpm_ctx *ctx);
/* SolverStatic::updateStaticDamage (solver_static.h:31-62) */
int lpm_update_static_damage(lpm_ctx *ctx, int32_t undamaged_type, int32_t *any_damaged);
/* SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-126) */
int lpm_update_fatigue_damage(lpm_ctx *ctx, double dNdt, int32_t undamaged_type, int32_t *any_damaged);
/* Assembly::storeStateVar / resetStateVar (assembly.h:140-159) */
int lpm_store_state(lpm_ctx *ctx);
int lpm_reset_state(lpm_ctx *ctx, int32_t reset_xyz);

/* ------------------------------------------------------------------ solver drivers
 * SolverStatic::solveProblemStep (solver_static.h:97-149); *converged = its return */
int lpm_solve_step_static(lpm_ctx *ctx, const lpm_load_step *step, int32_t undamaged_type, int32_t *converged);
/* SolverStatic::solveProblem (solver_static.h:64-95) without the dump writer: halving and
 * re-insertion of a failed step; *n_steps = steps run including inserted halves */
int lpm_solve_problem_static(lpm_ctx *ctx, const lpm_load_step *steps, int32_t n_steps_in,
                             int32_t undamaged_type, int32_t *n_steps);
/* SolverFatigue::solveProblemStep (solver_fatigue.h:277-298) */
int lpm_solve_step_fatigue(lpm_ctx *ctx, const lpm_load_step *step, int32_t *converged);
/* SolverFatigue::solveProblemStatic (solver_fatigue.h:243-275) without the dump writer */
int lpm_solve_problem_fatigue_static(lpm_ctx *ctx, const lpm_load_step *steps, int32_t n_steps_in,
                                     int32_t *n_steps);
/* SolverFatigue::solveProblemCyclic + generateLoadCycle + solveProblemOneCycle
 * (solver_fatigue.h:128-241) with the load templates of load_step_fatigue.h:22-85.
 * groups: the template's particle groups (3, or 2 for LPM_LOAD_DOGBONE_DISP).
 * max_cycles_run > 0 stops after that many simulated cycles. *cycles = final N. */
int lpm_solve_problem_cyclic(lpm_ctx *ctx, int32_t load_type, int32_t n_interval, const double *spectrum,
                             int64_t n_spectrum, const int32_t *groups, int32_t n_groups, int32_t time_mode,
                             double tau, int32_t undamaged_type, int32_t max_cycles_run, int32_t *cycles);

/* ------------------------------------------------------------------ output
 * Assembly::writeDump (assembly.h:309-335) without stopping the solver: the fields the dump
 * prints (positions, displacements, stress, damage, damage_visual, state_var[0], type) are
 * copied on the device in stream order, copied to pinned host memory on a side stream, and a
 * writer thread formats and appends the frame to `path` (text byte-identical to the
 * reference's fprintf, box[6] as Assembly::box). Returns at once; the context may step on.
 * One write is in flight per context: a second call first waits for the previous one.
 * lpm_dump_wait joins the write in flight and returns its status (no reference counterpart). */
int lpm_dump_async(lpm_ctx *ctx, const char *path, int32_t step, const double *box);
int lpm_dump_wait(lpm_ctx *ctx);
/* The dump writer's number formatting, exposed for testing: x[0..n) as printf("%.4e") would
 * print them, one per line, NUL-terminated, into out[cap] (LPM_E_LIMIT if it does not fit). */
int lpm_format_e4(const double *x, int64_t n, char *out, int64_t cap);

/* ------------------------------------------------------------------ tangent and adjoint
 * (no reference counterpart: sensitivities of the load-step result, frozen damage / broken
 * bonds / plastic bond lengths, for parameter calibration)
 * Exact directional derivative of the internal force of Assembly::updateGeometry +
 * updateForceState (particle.h:163-215, bond.h:30-39, particle_elastic.h:38-50) at the current
 * state: jv = (dPin/dx) v, both dim*n host arrays in DoF order. */
int lpm_force_jvp(lpm_ctx *ctx, const double *v, double *jv);
/* The transposed product y = (dPin/dx)^T w of the same Jacobian (not symmetric when the bonds
 * of a particle's layer carry different damage, or a bond is broken and its reverse is not). */
int lpm_force_vjp(lpm_ctx *ctx, const double *w, double *y);
/* Adjoint sensitivity of F = sum over the particles of `group` of Pin[axis] (the group's
 * reaction, Solver::updateRR, solver.h:108-128) at the current converged state, with respect
 * to every bond's Kn and Tv (nbonds doubles each; either may be NULL). One transposed solve
 * J_ff^T lambda = (J^T e)_free on the exact force Jacobian J, by iterative refinement whose
 * residuals use the exact product J^T (lpm_force_vjp) and whose inner solves run the CG on
 * the assembled, Dirichlet-eliminated K; then, with w = e_group,axis - lambda and
 * t_b = (cs_b.w_i + cs_opp(b).w_j)/2 (= (w_i - w_j).cs_b / 2 for an unbroken pair),
 * dF/dKn_b = 2 t_b dLe_b (1 - d_b), dF/dTv_b = 2 t_b dLe_total(i,l_b) (1 - d_b).
 * Reassembles K at the current state. */
int lpm_adjoint_reaction(lpm_ctx *ctx, int32_t group, int32_t axis, double *dF_dKn, double *dF_dTv, double *F,
                         int32_t *refinements, double *rel_residual);
/* Adjoint sensitivity of a linear objective of the current converged state,
 *   G = sum_d w_pin[d] Pin[d] + sum_d w_x[d] x[d]
 * (dim*n host arrays in DoF order, d = dim*id + axis; either may be NULL): reaction sums (w_pin
 * on constrained DoFs) and displacement fields (w_x; x - x0 has the same derivative as x).
 * Returns dG/dKn, dG/dTv and dG/dbdamage per bond (nbonds doubles each, any may be NULL) and G.
 * One transposed solve J_ff^T lambda = (J^T w_pin + w_x)_free as in lpm_adjoint_reaction, then
 * with w = w_pin - lambda: dG/dKn_b and dG/dTv_b as there, and
 * dG/dbdamage_b = -2 t_b (Kn_b dLe_b + Tv_b dLe_total(i,l_b)) (zero on broken bonds). */
int lpm_adjoint_linear(lpm_ctx *ctx, const double *w_pin, const double *w_x, double *dG_dKn, double *dG_dTv,
                       double *dG_dbdamage, double *G, int32_t *refinements, double *rel_residual);

/* Sensitivity of the same linear objective, at the current equilibrium, to each particle's HCF
 * parameters A, B, d (mat rows 0-2; n doubles each, any may be NULL) through the last fatigue
 * damage update (lpm_update_fatigue_damage: SolverFatigue::updateFatigueDamage,
 * solver_fatigue.h:75-126, with ParticleFatigueHCF::updateParticleStateVariables and
 * updateBondsForce, particle_fatigue_hcf.h:65-80,127-145):
 *   D'_i = D_i + dt sum_j w_ij A_j s0_j^B_j (s2 - s1)_j^d_j,  bd_b = max(bd_b, max(D'_i, D'_j)),
 * with the energies, the step dt and the earlier history held fixed: one transposed solve gives
 * dG/dbdamage, which is carried back through the maxima (std::max order), the nonlocal weights
 * w_ij at the update's positions, and the rate law. Particles whose damage was clamped to 1 and
 * the excluded type contribute nothing. Needs the HCF model, an update since the context was
 * created, and one nonlocal length (symmetric neighbour lists). */
int lpm_adjoint_fatigue_params(lpm_ctx *ctx, const double *w_pin, const double *w_x, double *dG_dA, double *dG_dB,
                               double *dG_dd, double *G, int32_t *refinements, double *rel_residual);
/* Record a fatigue load path for lpm_adjoint_fatigue_path (no reference counterpart): while on,
 * every HCF state-variable update (ParticleFatigueHCF::updateParticleStateVariables,
 * particle_fatigue_hcf.h:127-145, whether called directly or by the static, fatigue-step and
 * cycle drivers) keeps a checkpoint of its converged state and the energy trackers s0, s1, s2 of
 * every particle before and after it, and every fatigue damage update
 * (SolverFatigue::updateFatigueDamage, solver_fatigue.h:75-126) its cycle step, excluded type
 * and the damage it left. Device memory per recorded state is about one lpm_checkpoint_save
 * slot. on = 1 starts a new record (dropping the old one), 0 stops and releases it. */
int lpm_fatigue_record(lpm_ctx *ctx, int32_t on);
/* Sensitivity of the linear objective G = w_pin.Pin + w_x.x (as in lpm_adjoint_linear) at the
 * current equilibrium to every particle's HCF parameters A, B, d (mat rows 0-2; n doubles each,
 * any may be NULL) through the whole recorded path: each fatigue update's rate law and nonlocal
 * average (with the weights' dependence on the positions of the state the update was taken in),
 * the bond damage each update raises (bd = max(bd, D_i, D_j), std::max order), and each recorded
 * state's energy as the cycle's minimum, maximum or the running maximum s0 carries it to later
 * updates. One transposed solve at the current state and one at every recorded state whose
 * energy reaches a later update (*solves counts them; *max_rel_residual is the largest final
 * relative residual of their refinements). Thresholds are held at their recorded branches:
 * clamped damage and min/max ties contribute nothing. Restores the current state afterwards.
 * Needs the HCF model, a recorded update, one nonlocal length. */
int lpm_adjoint_fatigue_path(lpm_ctx *ctx, const double *w_pin, const double *w_x, double *dG_dA, double *dG_dB,
                             double *dG_dd, double *G, int32_t *solves, double *max_rel_residual);

/* ------------------------------------------------------------------ load-step checkpoints
 * (no reference counterpart: the reference keeps only the last converged copy,
 * storeStateVar / resetStateVar, particle.h:108-148, assembly.h:140-159)
 * Save the state a converged step leaves in device slot `slot` (0 .. LPM_MAX_CHECKPOINTS-1):
 * positions, external loads, constraint flags, damage, state variables and the bond state
 * (dis, dLp, bforce, bdamage), each with its last-converged copy. Restore puts it back,
 * recomputes geometry and forces, and leaves the matrix to be reassembled by the next step
 * or adjoint. Slots are allocated on first save and kept until released (slot -1: all) or
 * lpm_destroy. Used to run an adjoint at each step of a load path after the forward pass. */
#define LPM_MAX_CHECKPOINTS 4096
int lpm_checkpoint_save(lpm_ctx *ctx, int32_t slot);
int lpm_checkpoint_restore(lpm_ctx *ctx, int32_t slot);
int lpm_checkpoint_release(lpm_ctx *ctx, int32_t slot);

/* ------------------------------------------------------------------ specimen ensembles
 * (no reference counterpart: the reference runs one specimen per process, particle.h:17)
 * Declare the context's particles to be nspec independent specimens, specimen s owning ids
 * first[s] .. first[s+1]-1 (first[0] = 0, first[nspec] = n; no bond may join two specimens).
 * Every kernel then runs all specimens in the same launches; Solver::NewtonIteration
 * (solver.h:53-89) keeps each specimen's own tolerance multiplier, residual and iteration
 * count, and the nonlocal search and the two-level preconditioner's aggregates stay within a
 * specimen. nspec = 0 returns to a single system. Set before the first two-level solve. */
int lpm_set_specimens(lpm_ctx *ctx, int32_t nspec, const int64_t *first);
/* lpm_solve_problem_cyclic for an ensemble: the load template is instantiated once per
 * specimen, on that specimen's own groups (groups[need*s .. need*s + need-1], need = 3, or 2
 * for LPM_LOAD_DOGBONE_DISP) with the spectrum's increments times scale[s] (NULL: 1), so that a
 * force splits over each specimen's own group and every specimen gets its own amplitude. All
 * specimens run the same cycles and cycle jumps (the time map does not depend on the state),
 * each fatigue step in the same launches with per-specimen Newton tests. */
int lpm_solve_problem_cyclic_specimens(lpm_ctx *ctx, int32_t load_type, int32_t n_interval, const double *spectrum,
                                       int64_t n_spectrum, const int32_t *groups, const double *scale,
                                       int32_t time_mode, double tau, int32_t undamaged_type, int32_t max_cycles_run,
                                       int32_t *cycles);
/* Per-specimen results of the last static step: "newton_iters", "converged" (nspec doubles). */
int lpm_specimen_stat(const lpm_ctx *ctx, const char *name, double *out);

/* ------------------------------------------------------------------ statistics / measurement
 * Counters since the last reset: "newton_iters","cg_iters","cg_solves","assemblies",
 * "spmv_launches" (SpMVs that did work); "spmv_ms" / "spmv_launches_timed" (sum of SpMV kernel
 * time from HIP events and the number of launches it covers, when timing is on);
 * "last_cg_iters","last_cg_error"; "spmv_tuned_ms" (mean time of the fused CG SpMV in the
 * configuration the tuner kept); "asm_tuned_ms" (the same for the planned assembly). */
double lpm_stat(const lpm_ctx *ctx, const char *name);
int lpm_reset_stats(lpm_ctx *ctx);
/* every > 0: time the CG SpMV of every `every`-th iteration with HIP events on the context
 * stream (1 = every launch; each event record idles the GPU for a few microseconds, so a
 * sparse sample keeps the measurement out of the solve time); 0: off */
int lpm_set_timing(lpm_ctx *ctx, int every);
/* Run the CG SpMV kernel `reps` times on the stored matrix (x = current p vector) and return
 * the mean kernel time in milliseconds (HIP events on the context stream). */
int lpm_spmv_bench(lpm_ctx *ctx, int32_t reps, double *ms);
/* Copy `bytes` between two scratch device buffers `reps` times with a plain streaming kernel
 * (k_copy_calib) and return the mean kernel time: the practical HBM ceiling, and a kernel of
 * known traffic for scaling hardware counters. No reference counterpart (measurement only). */
int lpm_copy_bench(lpm_ctx *ctx, int64_t bytes, int32_t reps, double *ms);
/* Stream handle (hipStream_t) of the context, for callers that order their own work. */
void *lpm_stream(lpm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* LPM_MI355X_H */
