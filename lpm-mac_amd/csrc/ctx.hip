// Context, state transfer, boundary conditions and the solver drivers of the C ABI
// (include/lpm_mi355x.h). The drivers are host code that sequences device phases exactly
// as the reference's Solver / SolverStatic / SolverFatigue methods sequence theirs; every
// floating-point operation on particle or bond data happens on the device.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lpm_internal.h"

static std::string g_create_error;

int lpm_fail(lpm_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

namespace {

// ---------------------------------------------------------------- BC kernels
// Solver::updateDisplacementBC (solver.h:130-166) for one DispBC
__global__ void k_disp_bc(int64_t m, const int32_t *__restrict__ ids, int serial, int mode, int axis, double step,
                          double *__restrict__ xyz, int32_t *__restrict__ cons) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lo = serial ? 0 : t, hi = serial ? (t == 0 ? m : 0) : (t < m ? t + 1 : t);
    for (int64_t q = lo; q < hi; ++q) {
        const int64_t i = ids[q];
        double d[3] = {0, 0, 0};
        d[axis] = step;
        if (mode == LPM_RELATIVE) {
            xyz[3 * i] = xyz[3 * i] + d[0];
            xyz[3 * i + 1] = xyz[3 * i + 1] + d[1];
            xyz[3 * i + 2] = xyz[3 * i + 2] + d[2];
        } else {  // Absolute: moveTo((step,0,0)) etc., solver.h:142-143
            xyz[3 * i] = d[0];
            xyz[3 * i + 1] = d[1];
            xyz[3 * i + 2] = d[2];
        }
        cons[3 * i + axis] = 1;
    }
}

// Solver::updateForceBC (solver.h:168-190) for one ForceBC
// Load-step predictor: every DoF that the step does not prescribe starts from its converged
// position plus an extrapolated increment: order 1 the previous step's converged increment
// (xyz_last - xyz_prev), order 2 that plus its change from the step before (the quadratic through
// the last three converged states); order 0 moves nothing. xyz_prev (and xyz_prev2 when given)
// then take the converged positions the next step extrapolates from.
__global__ void k_predict(int64_t n3, const int32_t *__restrict__ cons, const double *__restrict__ xyz_last,
                          double *__restrict__ xyz_prev, double *__restrict__ xyz_prev2, double *__restrict__ xyz,
                          int order) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n3) return;
    const double last = xyz_last[q], p1 = xyz_prev[q];
    if (order > 0 && cons[q] != 1) {
        double inc = last - p1;
        if (order == 2) inc = inc + (inc - (p1 - xyz_prev2[q]));
        xyz[q] = xyz[q] + inc;
    }
    if (xyz_prev2) xyz_prev2[q] = p1;
    xyz_prev[q] = last;
}

__global__ void k_force_bc(int64_t m, const int32_t *__restrict__ ids, int serial, int mode, double fx, double fy,
                           double fz, double *__restrict__ Pex) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lo = serial ? 0 : t, hi = serial ? (t == 0 ? m : 0) : (t < m ? t + 1 : t);
    const int num = (int)m;
    for (int64_t q = lo; q < hi; ++q) {
        const int64_t i = ids[q];
        if (mode == LPM_RELATIVE) {
            Pex[3 * i] += fx / num;
            Pex[3 * i + 1] += fy / num;
            Pex[3 * i + 2] += fz / num;
        } else {
            Pex[3 * i] = fx / num;
            Pex[3 * i + 1] = fy / num;
            Pex[3 * i + 2] = fz / num;
        }
    }
}

// A load step's BCs in one launch: one workgroup applies them in order (force BCs first, as
// above), each with the per-BC kernels' arithmetic, with a barrier between consecutive BCs so that
// a particle named by several sees them in sequence. The descriptors travel as the kernel argument.
struct BcItem {
    const int32_t *ids;
    int64_t m;
    double v[3];                     // fx, fy, fz or (step at v[0])
    int32_t kind, mode, axis, serial;  // kind 0 force, 1 displacement
};
constexpr int BC_BATCH = 64;        // 64 x 56 B of descriptors: within the 4 KB kernel argument limit
constexpr int64_t BC_BATCH_MAX_IDS = 65536;  // larger groups keep their own grid-wide launch
struct BcBatch {
    BcItem it[BC_BATCH];
    int32_t count;
};

__global__ __launch_bounds__(1024) void k_bc_batch(BcBatch b, double *__restrict__ xyz, int32_t *__restrict__ cons,
                                                    double *__restrict__ Pex) {
    for (int k = 0; k < b.count; ++k) {
        const BcItem &e = b.it[k];
        const int64_t lo = e.serial ? 0 : threadIdx.x, hi = e.serial ? (threadIdx.x == 0 ? e.m : 0) : e.m;
        const int64_t st = e.serial ? 1 : blockDim.x;
        if (e.kind == 0) {
            const int num = (int)e.m;
            for (int64_t q = lo; q < hi; q += st) {
                const int64_t i = e.ids[q];
                if (e.mode == LPM_RELATIVE) {
                    Pex[3 * i] += e.v[0] / num;
                    Pex[3 * i + 1] += e.v[1] / num;
                    Pex[3 * i + 2] += e.v[2] / num;
                } else {
                    Pex[3 * i] = e.v[0] / num;
                    Pex[3 * i + 1] = e.v[1] / num;
                    Pex[3 * i + 2] = e.v[2] / num;
                }
            }
        } else {
            for (int64_t q = lo; q < hi; q += st) {
                const int64_t i = e.ids[q];
                double d[3] = {0, 0, 0};
                d[e.axis] = e.v[0];
                if (e.mode == LPM_RELATIVE) {
                    xyz[3 * i] = xyz[3 * i] + d[0];
                    xyz[3 * i + 1] = xyz[3 * i + 1] + d[1];
                    xyz[3 * i + 2] = xyz[3 * i + 2] + d[2];
                } else {
                    xyz[3 * i] = d[0];
                    xyz[3 * i + 1] = d[1];
                    xyz[3 * i + 2] = d[2];
                }
                cons[3 * i + e.axis] = 1;
            }
        }
        __syncthreads();
    }
}

struct Field {
    void *p;
    size_t es;
    int64_t count;
    bool isint;
};

Field field(lpm_ctx *c, const std::string &w) {
    const int64_t n = c->n, nb = c->nbond, nd = c->n * c->dim;
    const size_t D = sizeof(double), I = sizeof(int32_t);
    if (w == "xyz") return {c->xyz, D, 3 * n, false};
    if (w == "xyz_initial") return {c->xyz0, D, 3 * n, false};
    if (w == "xyz_last") return {c->xyz_last, D, 3 * n, false};
    if (w == "Pin") return {c->Pin, D, 3 * n, false};
    if (w == "Pex") return {c->Pex, D, 3 * n, false};
    if (w == "stress") return {c->stress, D, 6 * n, false};
    if (w == "damage") return {c->damage, D, n, false};
    if (w == "damage_last") return {c->damage_last, D, n, false};
    if (w == "damage_visual") return {c->damage_visual, D, n, false};
    if (w == "Ddot_local") return {c->Ddot_local, D, n, false};
    if (w == "Ddot_nonlocal") return {c->Ddot_nonlocal, D, n, false};
    if (w == "nonlocal_L") return {c->nonlocal_L, D, n, false};
    if (w == "state_var") return {c->sv, D, LPM_NSV * n, false};
    if (w == "state_var_last") return {c->sv_last, D, LPM_NSV * n, false};
    if (w == "mat") return {c->mat, D, LPM_NMAT * n, false};
    if (w == "dLe_total") return {c->dLe_total, D, 2 * n, false};
    if (w == "TdLe_total") return {c->TdLe_total, D, 2 * n, false};
    if (w == "cs_sum") return {c->cs_sum, D, 6 * n, false};
    if (w == "type") return {c->type, I, n, true};
    if (w == "disp_constraint") return {c->cons, I, 3 * n, true};
    if (w == "nb") return {c->nbarr, I, n, true};
    if (w == "dis_initial") return {c->dis0, D, nb, false};
    if (w == "dis") return {c->dis, D, nb, false};
    if (w == "dis_last") return {c->dis_last, D, nb, false};
    if (w == "dLe") return {c->dLe, D, nb, false};
    if (w == "dLp") return {c->dLp, D, nb, false};
    if (w == "dLp_last") return {c->dLp_last, D, nb, false};
    if (w == "bstrain") return {c->bstrain, D, nb, false};
    if (w == "Kn") return {c->Kn, D, nb, false};
    if (w == "Tv") return {c->Tv, D, nb, false};
    if (w == "cs") return {c->cs, D, 3 * nb, false};
    if (w == "bforce") return {c->bforce, D, nb, false};
    if (w == "bforce_last") return {c->bforce_last, D, nb, false};
    if (w == "bdamage") return {c->bdamage, D, nb, false};
    if (w == "bdamage_last") return {c->bdamage_last, D, nb, false};
    if (w == "bond_ptr") return {c->bptr, I, n + 1, true};
    if (w == "bond_split") return {c->bsplit, I, n, true};
    if (w == "bond_p2") return {c->p2, I, nb, true};
    if (w == "opp") return {c->opp, I, nb, true};
    if (w == "conn_ptr") return {c->cptr, I, n + 1, true};
    if (w == "conn_idx") return {c->cidx, I, c->nconn, true};
    if (w == "nl_ptr") return {c->nlptr, I, c->nlptr ? n + 1 : 0, true};
    if (w == "nl_idx") return {c->nlidx, I, c->nnl, true};
    if (w == "residual") return {c->residual, D, nd, false};
    if (w == "disp") return {c->disp, D, nd, false};
    return {nullptr, 0, -1, false};
}

template <typename T>
int dalloc(lpm_ctx *c, T **p, int64_t count) {
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (count > 0 ? count : 1));
    if (e != hipSuccess) return lpm_fail(c, LPM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, sizeof(T) * (count > 0 ? count : 1), c->stream);
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return LPM_OK;
}

int copy_d2d(lpm_ctx *c, void *dst, const void *src, size_t bytes) {
    LPM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return LPM_OK;
}

namespace {
// Several device-to-device copies of double arrays in one launch (the state checkpoint's seven
// arrays): 16-byte pieces in a grid-stride loop per array, an odd last element by thread 0.
// Every array comes from hipMalloc, so 16-byte alignment holds.
constexpr int COPY_JOBS = 8;
struct CopyJobs {
    const double *src[COPY_JOBS];
    double *dst[COPY_JOBS];
    int64_t n[COPY_JOBS];
    int count;
};

__global__ __launch_bounds__(256) void k_copy_many(CopyJobs jobs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int j = 0; j < jobs.count; ++j) {
        const double2 *s = (const double2 *)jobs.src[j];
        double2 *d = (double2 *)jobs.dst[j];
        const int64_t n2 = jobs.n[j] / 2;
        for (int64_t i = t; i < n2; i += stride) d[i] = s[i];
        if (t == 0 && (jobs.n[j] & 1)) jobs.dst[j][jobs.n[j] - 1] = jobs.src[j][jobs.n[j] - 1];
    }
}

int copy_many(lpm_ctx *c, const CopyJobs &jobs) {
    int64_t most = 0;
    for (int j = 0; j < jobs.count; ++j) most = std::max(most, jobs.n[j] / 2);
    if (most == 0) {
        for (int j = 0; j < jobs.count; ++j)
            if (jobs.n[j] > 0) {
                const int r = copy_d2d(c, jobs.dst[j], jobs.src[j], sizeof(double) * jobs.n[j]);
                if (r) return r;
            }
        return LPM_OK;
    }
    const int g = (int)std::min<int64_t>(2048, (most + 255) / 256);
    hipLaunchKernelGGL(k_copy_many, dim3(g), dim3(256), 0, c->stream, jobs);
    LPM_HIP(c, hipGetLastError());
    return LPM_OK;
}
}  // namespace

#define RC(x)                \
    do {                     \
        int rc_ = (x);       \
        if (rc_) return rc_; \
    } while (0)

int check_group(lpm_ctx *c, int32_t g) {
    if (g < 0 || g >= (int32_t)c->groups.size()) return lpm_fail(c, LPM_E_ARG, "unknown particle group " + std::to_string(g));
    return LPM_OK;
}

int apply_bc(lpm_ctx *c, const lpm_load_step *s) {
    if (!s) return lpm_fail(c, LPM_E_ARG, "null load step");
    if ((s->n_force > 0 && !s->force) || (s->n_disp > 0 && !s->disp) || s->n_force < 0 || s->n_disp < 0)
        return lpm_fail(c, LPM_E_ARG, "inconsistent load step");
    BcBatch batch;
    batch.count = 0;
    auto flush = [&]() -> int {
        if (batch.count == 0) return LPM_OK;
        hipLaunchKernelGGL(k_bc_batch, dim3(1), dim3(1024), 0, c->stream, batch, c->xyz, c->cons, c->Pex);
        batch.count = 0;
        LPM_HIP(c, hipGetLastError());
        return LPM_OK;
    };
    auto add = [&](const BcItem &e) -> int {
        batch.it[batch.count++] = e;
        return batch.count == BC_BATCH ? flush() : LPM_OK;
    };
    for (int k = 0; k < s->n_force; ++k) {  // updateForceBC first (solver_static.h:103)
        const lpm_force_bc &f = s->force[k];
        RC(check_group(c, f.group));
        const Group &g = c->groups[f.group];
        if (g.n == 0) continue;
        if (g.n <= BC_BATCH_MAX_IDS) {
            RC(add({g.ids, g.n, {f.fx, f.fy, f.fz}, 0, f.mode, 0, g.dup ? 1 : 0}));
            continue;
        }
        RC(flush());
        const int blocks = g.dup ? 1 : blocks_for(g.n, 256);
        hipLaunchKernelGGL(k_force_bc, dim3(blocks), dim3(256), 0, c->stream, g.n, g.ids, g.dup ? 1 : 0, f.mode, f.fx, f.fy,
                           f.fz, c->Pex);
        LPM_HIP(c, hipGetLastError());
    }
    for (int k = 0; k < s->n_disp; ++k) {
        const lpm_disp_bc &d = s->disp[k];
        RC(check_group(c, d.group));
        const int axis = d.flag == 'x' ? 0 : (d.flag == 'y' ? 1 : (d.flag == 'z' ? 2 : -1));
        if (axis < 0) continue;  // the reference ignores other flags
        const Group &g = c->groups[d.group];
        if (g.n == 0) continue;
        if (g.n <= BC_BATCH_MAX_IDS) {
            RC(add({g.ids, g.n, {d.step, 0.0, 0.0}, 1, d.mode, axis, g.dup ? 1 : 0}));
            continue;
        }
        RC(flush());
        const int blocks = g.dup ? 1 : blocks_for(g.n, 256);
        hipLaunchKernelGGL(k_disp_bc, dim3(blocks), dim3(256), 0, c->stream, g.n, g.ids, g.dup ? 1 : 0, d.mode, axis, d.step,
                           c->xyz, c->cons);
        LPM_HIP(c, hipGetLastError());
    }
    return flush();
}

}  // namespace

// phases shared with the adjoint (kernels_adj.hip)
int geometry_force_state(lpm_ctx *c) {
    RC(launch_geometry_force(c, true));
    return launch_particle_force(c);
}

int calc_stiffness(lpm_ctx *c) {
    RC(cg_alloc(c));
    if (c->cg_transpose) RC(launch_tslot(c));
    if (c->stiffness_fd) RC(launch_assemble_fd(c));
    else RC(launch_assemble(c));
    c->k_ready = true;
    c->k_version++;
    c->st_asm++;
    return LPM_OK;
}

int update_stiffness_disp_bc(lpm_ctx *c) {
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "updateStiffnessDispBC before calcStiffness");
    c->k_version++;
    return launch_dirichlet(c);
}

namespace {

int solve_linear_system(lpm_ctx *c, int *iters, double *error) {
    int it = 0, ok = 0;
    double err = 0;
    RC(cg_solve(c, &it, &err, &ok));
    c->st_cg_iters += it;
    c->st_cg_solves++;
    c->last_cg_iters = it;
    c->last_cg_err = err;
    if (c->verbose) {
        if (!ok) printf("Solving failed.\n");
        else printf("The system has been solved after %d iterations.\nThe estimated error is %g\n", it, err);
    }
    RC(launch_move(c));
    if (iters) *iters = it;
    if (error) *error = err;
    return LPM_OK;
}

// Solver::NewtonIteration (solver.h:53-89) for every specimen of an ensemble at once: each
// specimen keeps its own tolerance multiplier, residual norm and iteration count. One CG solves
// all specimens still iterating; each one's right-hand side is scaled to unit norm, so the CG
// tolerance means the same for all of them, and converged or failed specimens get a zero
// right-hand side (no displacement). *n_iter = the largest count, or max_NR_iter if a specimen failed.
static int newton_iteration_spec(lpm_ctx *c, int *n_iter) {
    const int S = c->nspec;
    std::vector<double> r2, R2;
    RC(geometry_force_state(c));
    RC(coarse_check(c));  // its flag arrives with the residual
    RC(launch_residual_spec(c, r2, R2));
    std::vector<double> tolm(S), nr(S), w(S);
    std::vector<char> active(S), failed(S, 0);
    c->spec_newton.assign(S, 0);
    const bool ref_start = (int)c->tolm_ref.size() == S;
    for (int s = 0; s < S; ++s) {
        nr[s] = std::sqrt(r2[s]);
        tolm[s] = ref_start ? c->tolm_ref[s] : std::max(nr[s], std::sqrt(R2[s]));
        active[s] = nr[s] > c->tol_NR * tolm[s];
        if (!std::isfinite(nr[s]) || !std::isfinite(tolm[s])) failed[s] = 1, active[s] = 0;  // see newton_iteration
    }
    c->tolm_ref.clear();
    for (int newton_round = 0;; ++newton_round) {
        int nact = 0;
        double eta = 0.1;
        for (int s = 0; s < S; ++s) {
            if (!active[s]) continue;
            if (++c->spec_newton[s] > c->max_NR_iter) {
                active[s] = 0;
                failed[s] = 1;
                c->spec_newton[s] = c->max_NR_iter;
                continue;
            }
            ++nact;
            eta = std::min(eta, 0.5 * c->tol_NR * tolm[s] / nr[s]);
        }
        if (nact == 0) break;
        for (int s = 0; s < S; ++s) w[s] = active[s] ? 1.0 / nr[s] : 0.0;
        LPM_HIP(c, hipMemcpyAsync(c->spec_w, w.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
        RC(launch_spec_scale(c, c->residual));
        int it = 0, ok = 0;
        double err = 0;
        // Every active specimen's right-hand side has norm 1, and the CG tests the whole
        // residual: a tolerance of tau / sqrt(active) bounds each specimen's own relative
        // residual by tau. tau is the set cg_tol or, with forcing, the inexact-Newton reduction
        // the strictest active specimen needs.
        const double cg_tol_set = c->cg_tol;
        const double tau = c->cg_forcing ? std::max(cg_tol_set, eta) : cg_tol_set;
        c->cg_tol = tau / std::sqrt((double)nact);
        c->cg_slot = std::min(newton_round, 3);
        const int rc_cg = cg_solve(c, &it, &err, &ok);
        c->cg_slot = -1;
        c->cg_tol = cg_tol_set;
        RC(rc_cg);
        c->st_cg_iters += it;
        c->st_cg_solves++;
        c->last_cg_iters = it;
        c->last_cg_err = err;
        for (int s = 0; s < S; ++s) w[s] = active[s] ? nr[s] : 0.0;
        LPM_HIP(c, hipMemcpyAsync(c->spec_w, w.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
        RC(launch_scale_x(c));
        RC(launch_move(c));
        RC(geometry_force_state(c));
        RC(launch_residual_spec(c, r2, R2));
        c->st_newton++;
        int nleft = 0, worst = -1;
        double wrel = 0.0;
        for (int s = 0; s < S; ++s) {
            nr[s] = std::sqrt(r2[s]);
            if (active[s] && !std::isfinite(nr[s])) failed[s] = 1, active[s] = 0;
            if (active[s] && !(nr[s] > c->tol_NR * tolm[s])) active[s] = 0;
            if (active[s]) {
                ++nleft;
                if (nr[s] / tolm[s] > wrel) wrel = nr[s] / tolm[s], worst = s;
            }
        }
        if (c->verbose)
            printf("|  |  specimens: CG %d iterations (error %.3e), %d still active, worst relative residual %.3e (specimen %d)\n",
                   it, err, nleft, wrel, worst);
    }
    int mx = 0;
    bool anyfail = false;
    c->spec_ok.assign(S, 1);
    for (int s = 0; s < S; ++s) {
        mx = std::max(mx, c->spec_newton[s]);
        if (failed[s]) {
            anyfail = true;
            c->spec_ok[s] = 0;
        }
    }
    *n_iter = anyfail ? c->max_NR_iter : mx;
    return LPM_OK;
}

int newton_iteration(lpm_ctx *c, int *n_iter) {
    // the Jacobi inverse of K runs while the host waits for the first residual
    if (c->k_ready) RC(launch_precond_setup(c));
    if (c->nspec > 0) return newton_iteration_spec(c, n_iter);
    double r2 = 0, R2 = 0;
    RC(geometry_force_state(c));
    RC(coarse_check(c));  // its flag arrives with the residual
    RC(launch_residual(c, &r2, &R2));
    double norm_residual = std::sqrt(r2);
    // the reference's multiplier: at its own start, or at the reference's start before a prediction
    // (copied to h_scal[SC_COUNT..] in stream order before the wait above)
    const double tol_multiplier =
        c->tolm_pending ? std::max(std::sqrt(c->h_scal[SC_COUNT]), std::sqrt(c->h_scal[SC_COUNT + 1]))
                        : (c->tolm_ref.size() == 1 ? c->tolm_ref[0] : std::max(norm_residual, std::sqrt(R2)));
    c->tolm_ref.clear();
    c->tolm_pending = false;
    if (c->verbose)
        printf("|  Norm of residual is %.5e, norm of reaction is %.5e, tolerance criterion is based on %s force\n",
               norm_residual, std::sqrt(R2), norm_residual > std::sqrt(R2) ? "residual" : "reaction");
    int ni = 0;
    // A residual that is no longer finite (a solve that diverged once bonds broke through the
    // specimen) fails the iteration; the reference's test `norm > tol * mult` is false for NaN
    // and would report convergence.
    if (!std::isfinite(norm_residual) || !std::isfinite(tol_multiplier)) {
        *n_iter = c->max_NR_iter;
        return LPM_OK;
    }
    while (norm_residual > c->tol_NR * tol_multiplier) {
        if (++ni > c->max_NR_iter) {
            *n_iter = c->max_NR_iter;
            return LPM_OK;
        }
        if (c->verbose) printf("|  |  Iteration-%d: ", ni);
        const double cg_tol_set = c->cg_tol;
        if (c->cg_forcing) {
            // inexact Newton: ask CG only for the reduction that brings the residual under the
            // Newton target (with a factor 2 margin), never less than the set cg_tol, at most 0.1
            const double eta = 0.5 * c->tol_NR * tol_multiplier / norm_residual;
            c->cg_tol = std::max(cg_tol_set, std::min(0.1, eta));
        }
        c->cg_slot = std::min(ni - 1, 3);
        const int rc_solve = solve_linear_system(c, nullptr, nullptr);
        c->cg_slot = -1;
        c->cg_tol = cg_tol_set;
        RC(rc_solve);
        RC(geometry_force_state(c));
        RC(launch_residual(c, &r2, nullptr));
        norm_residual = std::sqrt(r2);
        c->st_newton++;
        if (c->verbose)
            printf("|  |  Norm of residual is %.3e, residual ratio is %.3e\n", norm_residual, norm_residual / tol_multiplier);
        if (!std::isfinite(norm_residual)) {
            *n_iter = c->max_NR_iter;
            return LPM_OK;
        }
    }
    *n_iter = ni;
    return LPM_OK;
}

int store_state(lpm_ctx *c) {
    const int64_t n = c->n, nb = c->nbond;
    const CopyJobs jobs = {{c->xyz, c->damage, c->sv, c->dis, c->bdamage, c->bforce, c->dLp},
                           {c->xyz_last, c->damage_last, c->sv_last, c->dis_last, c->bdamage_last, c->bforce_last,
                            c->dLp_last},
                           {3 * n, n, LPM_NSV * n, nb, nb, nb, nb},
                           7};
    return copy_many(c, jobs);
}

int reset_state(lpm_ctx *c, bool reset_xyz) {
    const size_t n = c->n, nb = c->nbond, D = sizeof(double);
    c->pred_hist = 0;  // the step is re-run (halved): no increment to extrapolate
    if (reset_xyz) RC(copy_d2d(c, c->xyz, c->xyz_last, 3 * n * D));
    RC(copy_d2d(c, c->damage, c->damage_last, n * D));
    RC(copy_d2d(c, c->sv, c->sv_last, LPM_NSV * n * D));
    RC(copy_d2d(c, c->dis, c->dis_last, nb * D));
    RC(copy_d2d(c, c->bdamage, c->bdamage_last, nb * D));
    RC(copy_d2d(c, c->bforce, c->bforce_last, nb * D));
    RC(copy_d2d(c, c->dLp, c->dLp_last, nb * D));
    return LPM_OK;
}

// SolverStatic::solveProblemStep (solver_static.h:97-149)
// predictor (mode "predictor", static steps only: fatigue substeps alternate loading and
// unloading): after the step's boundary conditions, move the free DoFs by the previous converged
// step's increment; Newton then corrects from there to the same test
int predict(lpm_ctx *c) {
    c->tolm_ref.clear();
    c->tolm_pending = false;
    if (!c->predictor) return LPM_OK;
    const int64_t n3 = 3 * c->n;
    if (!c->xyz_prev) {
        LPM_HIP(c, hipMalloc(&c->xyz_prev, sizeof(double) * n3));
        c->pred_hist = 0;
    }
    if (c->predictor == 2 && !c->xyz_prev2) {
        LPM_HIP(c, hipMalloc(&c->xyz_prev2, sizeof(double) * n3));
        c->pred_hist = std::min(c->pred_hist, 1);
    }
    const int order = std::min(c->pred_hist, c->predictor);
    if (order > 0) {
        // Newton's test keeps the reference's multiplier: the norms at the reference's start
        RC(geometry_force_state(c));
        if (c->nspec > 0) {
            std::vector<double> r2, R2;
            RC(launch_residual_spec(c, r2, R2));
            c->tolm_ref.resize(c->nspec);
            for (int s = 0; s < c->nspec; ++s) c->tolm_ref[s] = std::max(std::sqrt(r2[s]), std::sqrt(R2[s]));
        } else {
            // the two norms follow to the host behind the prediction, the assembly and the first
            // force pass; newton_iteration reads them after its own first wait
            static_assert(SC_NORM_REAC == SC_NORM_R + 1, "the reference-start norms are copied as a pair");
            RC(launch_residual(c, nullptr, nullptr));
            LPM_HIP(c, hipMemcpyAsync(c->h_scal + SC_COUNT, c->scal + SC_NORM_R, sizeof(double) * 2,
                                      hipMemcpyDeviceToHost, c->stream));
            c->tolm_pending = true;
        }
    }
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(n3, 256)), dim3(256), 0, c->stream, n3, (const int32_t *)c->cons,
                       (const double *)c->xyz_last, c->xyz_prev, c->predictor == 2 ? c->xyz_prev2 : (double *)nullptr,
                       c->xyz, order);
    LPM_HIP(c, hipGetLastError());
    c->pred_hist = std::min(c->pred_hist + 1, c->predictor == 2 ? 2 : 1);
    return LPM_OK;
}

int solve_step_static(lpm_ctx *c, const lpm_load_step *s, int undamaged, int *converged) {
    RC(apply_bc(c, s));
    RC(predict(c));
    int n_newton = 0, any = 0;
    bool new_damaged = false;
    do {
        RC(calc_stiffness(c));
        if (c->dirichlet_static) RC(update_stiffness_disp_bc(c));
        RC(newton_iteration(c, &n_newton));
        if (n_newton >= c->max_NR_iter) break;
        RC(launch_state_var(c));
        // the damage flag comes back behind the state checkpoint, which is stored either way
        RC(launch_static_damage(c, undamaged, nullptr));
        RC(store_state(c));
        RC(flag_wait(c, &any));
        new_damaged = any != 0;
        if (new_damaged) RC(geometry_force_state(c));
    } while (new_damaged);
    *converged = n_newton < c->max_NR_iter;
    return LPM_OK;
}

// SolverFatigue::solveProblemStep (solver_fatigue.h:277-298)
int solve_step_fatigue(lpm_ctx *c, const lpm_load_step *s, int *converged) {
    RC(apply_bc(c, s));
    RC(calc_stiffness(c));
    RC(update_stiffness_disp_bc(c));
    int n_newton = 0;
    RC(newton_iteration(c, &n_newton));
    *converged = n_newton < c->max_NR_iter;
    return LPM_OK;
}

// owning copy of a load step, so that halving does not touch the caller's arrays
struct OwnedStep {
    std::vector<lpm_disp_bc> disp;
    std::vector<lpm_force_bc> force;
    explicit OwnedStep(const lpm_load_step &s) : disp(s.disp, s.disp + s.n_disp), force(s.force, s.force + s.n_force) {}
    lpm_load_step view() const { return {(int32_t)disp.size(), disp.data(), (int32_t)force.size(), force.data()}; }
    // LoadStep::loadCutHalf (load_step.h:21-28)
    void cut_half() {
        for (auto &d : disp) d.step /= 2.0;
        for (auto &f : force) { f.fx /= 2.0; f.fy /= 2.0; f.fz /= 2.0; }
    }
};

int check_steps(lpm_ctx *c, const lpm_load_step *steps, int32_t n) {
    if (n < 0 || (n > 0 && !steps)) return lpm_fail(c, LPM_E_ARG, "bad load step list");
    for (int32_t k = 0; k < n; ++k)
        if ((steps[k].n_disp > 0 && !steps[k].disp) || (steps[k].n_force > 0 && !steps[k].force))
            return lpm_fail(c, LPM_E_ARG, "inconsistent load step");
    return LPM_OK;
}

// solveProblem / solveProblemStatic loop with halving (solver_static.h:64-95; solver_fatigue.h:243-275)
int solve_problem(lpm_ctx *c, const lpm_load_step *steps, int32_t n_in, int undamaged, bool fatigue, int32_t *n_out) {
    RC(check_steps(c, steps, n_in));
    std::vector<OwnedStep> load;
    for (int32_t k = 0; k < n_in; ++k) load.emplace_back(steps[k]);
    std::vector<double> pex;
    int n_step = (int)load.size();
    for (int i = 0; i < n_step; i++) {
        int halvings = 0;
        for (;;) {
            if (c->rollback_pex) {
                pex.resize(3 * c->n);
                LPM_HIP(c, hipMemcpyAsync(pex.data(), c->Pex, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost, c->stream));
                LPM_HIP(c, hipStreamSynchronize(c->stream));
            }
            int ok = 0;
            lpm_load_step v = load[i].view();
            if (fatigue) RC(solve_step_fatigue(c, &v, &ok));
            else RC(solve_step_static(c, &v, undamaged, &ok));
            if (ok) break;
            if (++halvings > 30)  // the reference halves without end
                return lpm_fail(c, LPM_E_STATE, "load step " + std::to_string(i) + " does not converge after 30 halvings");
            RC(reset_state(c, true));
            if (c->rollback_pex)
                LPM_HIP(c, hipMemcpyAsync(c->Pex, pex.data(), sizeof(double) * 3 * c->n, hipMemcpyHostToDevice, c->stream));
            load[i].cut_half();
            load.insert(load.begin() + i, load[i]);
            ++n_step;
            if (c->verbose) printf("Step-%d not converging\n\n", i);
        }
        if (fatigue) {
            RC(launch_state_var(c));
            RC(store_state(c));
        }
    }
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    if (n_out) *n_out = n_step;
    return LPM_OK;
}

// load templates of load_step_fatigue.h:22-85
OwnedStep fatigue_template(int type, double v, const int32_t *g) {
    lpm_load_step empty{0, nullptr, 0, nullptr};
    OwnedStep s(empty);
    auto D = [&](int32_t grp, int flag, double step) { s.disp.push_back({grp, LPM_RELATIVE, flag, step}); };
    auto F = [&](int32_t grp, double fy) { s.force.push_back({grp, LPM_RELATIVE, 0.0, fy, 0.0}); };
    if (type == LPM_LOAD_UNIAXIAL_DISP) { D(g[0], 'x', 0.0); D(g[1], 'y', v); D(g[2], 'y', 0.0); }
    else if (type == LPM_LOAD_UNIAXIAL_FORCE) { D(g[0], 'x', 0.0); F(g[1], v); D(g[2], 'y', 0.0); }
    else if (type == LPM_LOAD_DOGBONE_DISP) {
        D(g[0], 'x', 0); D(g[0], 'y', v); D(g[0], 'z', 0); D(g[1], 'x', 0.0); D(g[1], 'y', 0.0); D(g[1], 'z', 0.0);
    } else { D(g[0], 'y', 0); D(g[1], 'x', 0.0); D(g[2], 'x', 0.0); F(g[1], v); F(g[2], -v); }
    return s;
}

// SolverFatigue::solveProblemCyclic + generateLoadCycle + solveProblemOneCycle
// (solver_fatigue.h:128-241). nset template instances run side by side: instance k applies the
// template to groups[need k .. need k + need - 1] with the spectrum's increments times scale[k]
// (one instance for a single system, one per specimen for an ensemble, which then advances all
// specimens through the same cycles and cycle jumps in the same launches).
int solve_cyclic(lpm_ctx *c, int32_t load_type, int32_t n_interval, const double *spec, int64_t nspec,
                 const int32_t *groups, const double *scale, int nset, int32_t tmode, double tau, int32_t undamaged,
                 int32_t max_cycles_run, int32_t *cycles) {
    if (!spec || n_interval <= 0 || tau <= 0 || load_type < 0 || load_type > 3) return lpm_fail(c, LPM_E_ARG, "bad cyclic load");
    const int need = load_type == LPM_LOAD_DOGBONE_DISP ? 2 : 3;
    for (int k = 0; k < need * nset; ++k) RC(check_group(c, groups[k]));
    const double n_cycles = 0.5 * (nspec - 1);
    double t = 0, dNdt = 1;
    int N = 0, ran = 0;
    auto half = [&](double df) {
        lpm_load_step empty{0, nullptr, 0, nullptr};
        OwnedStep st(empty);
        for (int k = 0; k < nset; ++k) {
            const OwnedStep one = fatigue_template(load_type, df * scale[k], groups + need * k);
            st.disp.insert(st.disp.end(), one.disp.begin(), one.disp.end());
            st.force.insert(st.force.end(), one.force.begin(), one.force.end());
        }
        return st;
    };
    do {
        if (2 * (int64_t)N + 2 >= nspec) return lpm_fail(c, LPM_E_ARG, "load spectrum too short for the cycle count");
        const double f0 = spec[2 * N], f1 = spec[2 * N + 1], f2 = spec[2 * N + 2];
        std::vector<OwnedStep> cyc;
        for (int k = 0; k < n_interval; ++k) cyc.push_back(half((f1 - f0) / n_interval));
        for (int k = 0; k < n_interval; ++k) cyc.push_back(half((f2 - f1) / n_interval));
        dNdt = (tmode == LPM_TIME_LINEAR) ? (int)(1 / tau) : (int)(1 / tau * std::exp(t++ / tau));
        for (auto &st : cyc) {
            int ok = 0;
            lpm_load_step v = st.view();
            RC(solve_step_fatigue(c, &v, &ok));
            RC(launch_state_var(c));
        }
        int any = 0;
        RC(launch_fatigue_damage(c, dNdt, undamaged, &any));
        RC(store_state(c));
        RC(geometry_force_state(c));
        N += (int)dNdt;
        ++ran;
        if (max_cycles_run > 0 && ran >= max_cycles_run) break;
    } while (N < n_cycles);
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    if (cycles) *cycles = N;
    return LPM_OK;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char *lpm_last_error(const lpm_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

static int create(const lpm_topology *t, const lpm_cut *cuts, int32_t ncut, lpm_ctx **out) {
    if (!t || !out) return lpm_fail(nullptr, LPM_E_ARG, "null argument");
    *out = nullptr;
    // bond_ptr == NULL: search the bonds on the device; else the caller's lists, validated here
    const bool search = t->bond_ptr == nullptr;
    if (ncut < 0 || (ncut > 0 && (!cuts || !search)))
        return lpm_fail(nullptr, LPM_E_ARG, "cuts apply to the device bond search (bond_ptr NULL)");
    for (int32_t k = 0; k < ncut; ++k)
        if (cuts[k].kind < LPM_CUT_CIRCLE || cuts[k].kind > LPM_CUT_REVERSE_NOTCH)
            return lpm_fail(nullptr, LPM_E_ARG, "unknown cut kind");
    if (t->n <= 0 || !t->xyz || (!search && (!t->bond_split || (!t->bond_p2 && t->bond_ptr[t->n] > 0))))
        return lpm_fail(nullptr, LPM_E_ARG, "empty or incomplete topology");
    if (t->model < 0 || t->model > 3) return lpm_fail(nullptr, LPM_E_ARG, "unknown particle model");
    if (t->n > 0x7ffffffeLL) return lpm_fail(nullptr, LPM_E_LIMIT, "too many particles");
    double uc[8];
    if (lpm_unit_cell(t->lattice, t->radius, 0, uc, nullptr) != LPM_OK) return lpm_fail(nullptr, LPM_E_ARG, "unknown lattice");
    const int64_t n = t->n;
    const int32_t *bp = t->bond_ptr, *bs = t->bond_split, *p2 = t->bond_p2;
    int maxb = 0;
    int64_t nbond = 0;
    std::vector<int32_t> opp;
    if (!search) {
        // ---- host validation of the topology (every index a kernel will dereference)
        if (bp[0] != 0) return lpm_fail(nullptr, LPM_E_TOPO, "bond_ptr[0] must be 0");
        for (int64_t i = 0; i < n; ++i) {
            if (bp[i + 1] < bp[i] || bs[i] < bp[i] || bs[i] > bp[i + 1])
                return lpm_fail(nullptr, LPM_E_TOPO, "bond_ptr / bond_split not monotone at particle " + std::to_string(i));
            maxb = std::max(maxb, bp[i + 1] - bp[i]);
        }
        if (maxb > LPM_MAXB) return lpm_fail(nullptr, LPM_E_LIMIT, "a particle has more than 32 bonds");
        nbond = bp[n];
        opp.assign(nbond, -1);
        for (int64_t i = 0; i < n; ++i) {
            for (int b = bp[i]; b < bp[i + 1]; ++b) {
                const int64_t j = p2[b];
                if (j < 0 || j >= n || j == i) return lpm_fail(nullptr, LPM_E_TOPO, "bond partner out of range or self bond");
                const bool l0 = b < bs[i];
                for (int b2 = (l0 ? bp[i] : bs[i]); b2 < b; ++b2)
                    if (p2[b2] == j) return lpm_fail(nullptr, LPM_E_TOPO, "duplicate bond in one layer");
                const int lo = l0 ? bp[j] : bs[j], hi = l0 ? bs[j] : bp[j + 1];
                for (int q = lo; q < hi; ++q)
                    if (p2[q] == i) { opp[b] = q; break; }
                if (opp[b] < 0)
                    return lpm_fail(nullptr, LPM_E_TOPO,
                                    "bond " + std::to_string(i) + "->" + std::to_string(j) + " has no reverse bond");
            }
        }
    }
    lpm_ctx *c = new lpm_ctx();
    c->device = t->device;
    c->lattice = t->lattice;
    c->radius = t->radius;
    c->model = t->model;
    c->dim = (int)uc[0];
    c->nneighbors = (int)uc[1];
    c->cutoff[0] = uc[5];
    c->cutoff[1] = uc[6];
    c->volume = uc[7];
    c->n = n;
    c->vpad = c->dim == 3 ? 3 : 2;
    c->precision = 0;
    c->cg_tol = 1e-6;
    auto fail = [&](int rc) {
        std::string m = c->err;
        lpm_destroy(c);
        return lpm_fail(nullptr, rc, m);
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) { c->err = hipGetErrorString(e); return fail(LPM_E_HIP); }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { c->err = hipGetErrorString(e); return fail(LPM_E_HIP); }
    int rc = 0;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess;
    };
    const double *x = t->xyz;
#define AL(ptr, cnt) if ((rc = dalloc(c, &c->ptr, (cnt)))) return fail(rc)
    // ---- per-particle arrays
    AL(xyz, 3 * n); AL(xyz0, 3 * n); AL(xyz_last, 3 * n); AL(Pin, 3 * n); AL(Pex, 3 * n); AL(stress, 6 * n);
    AL(damage, n); AL(damage_last, n); AL(damage_visual, n); AL(Ddot_local, n); AL(Ddot_nonlocal, n); AL(nonlocal_L, n);
    AL(sv, LPM_NSV * n); AL(sv_last, LPM_NSV * n); AL(mat, LPM_NMAT * n);
    AL(dLe_total, 2 * n); AL(TdLe_total, 2 * n); AL(cs_sum, 6 * n); AL(type, n); AL(cons, 3 * n);
    AL(residual, c->dim * n); AL(disp, c->dim * n); AL(diag64, c->dim * n);
    AL(scal, SC_COUNT); AL(counter, 8); AL(iflag, 8);
    if (!(up(c->xyz, x, sizeof(double) * 3 * n) && up(c->xyz0, x, sizeof(double) * 3 * n) &&
          up(c->xyz_last, x, sizeof(double) * 3 * n))) {
        c->err = "position upload failed";
        return fail(LPM_E_HIP);
    }
    // ---- bonds: Bond(p1, p2, layer, dis) (bond.h:61-73), reverse bonds, 2-hop bookkeeping
    if (search) {
        if ((rc = device_topology(c, x, cuts, ncut))) return fail(rc);
        nbond = c->nbond;
    } else {
        c->nbond = nbond;
        c->maxbond = maxb;
        for (int64_t p0 = 0; p0 < n; p0 += LPM_BOND_GP)
            c->gp_bonds = std::max<int>(c->gp_bonds, bp[std::min<int64_t>(n, p0 + LPM_BOND_GP)] - bp[p0]);
        // 2-hop records of a row: for each bond i->k, the bonds of k in the same layer
        for (int64_t i = 0; i < n; ++i) {
            int hops = 0;
            for (int b = bp[i]; b < bp[i + 1]; ++b) {
                const int k = p2[b];
                hops += (b < bs[i]) ? bs[k] - bp[k] : bp[k + 1] - bs[k];
            }
            c->asm_hops = std::max(c->asm_hops, hops);
        }
        AL(bptr, n + 1); AL(bsplit, n); AL(p2, nbond); AL(opp, nbond); AL(nbarr, n);
        AL(dis0, nbond); AL(dis, nbond); AL(dis_last, nbond); AL(cs, 3 * nbond);
        std::vector<double> dis0(nbond), cs(3 * nbond);
        std::vector<int32_t> nb(n);
        for (int64_t i = 0; i < n; ++i) {
            nb[i] = bp[i + 1] - bp[i];
            for (int b = bp[i]; b < bp[i + 1]; ++b) {
                const double *a = &x[3 * i], *q = &x[3 * (int64_t)p2[b]];
                const double d = std::sqrt(std::pow((a[0] - q[0]), 2) + std::pow((a[1] - q[1]), 2) + std::pow((a[2] - q[2]), 2));
                dis0[b] = d;
                cs[3 * b] = (a[0] - q[0]) / d;
                cs[3 * b + 1] = (a[1] - q[1]) / d;
                cs[3 * b + 2] = (a[2] - q[2]) / d;
            }
        }
        const bool ok = up(c->bptr, bp, sizeof(int32_t) * (n + 1)) && up(c->bsplit, bs, sizeof(int32_t) * n) &&
                        up(c->p2, p2, sizeof(int32_t) * nbond) && up(c->opp, opp.data(), sizeof(int32_t) * nbond) &&
                        up(c->nbarr, nb.data(), sizeof(int32_t) * n) && up(c->dis0, dis0.data(), sizeof(double) * nbond) &&
                        up(c->dis, dis0.data(), sizeof(double) * nbond) &&
                        up(c->dis_last, dis0.data(), sizeof(double) * nbond) &&
                        up(c->cs, cs.data(), sizeof(double) * 3 * nbond) && hipStreamSynchronize(c->stream) == hipSuccess;
        if (!ok) { c->err = "topology upload failed"; return fail(LPM_E_HIP); }
    }
    // ---- per-bond state
    AL(dLe, nbond); AL(dLp, nbond); AL(dLp_last, nbond);
    AL(bstrain, nbond); AL(Kn, nbond); AL(Tv, nbond); AL(bforce, nbond); AL(bforce_last, nbond);
    AL(bdamage, nbond); AL(bdamage_last, nbond);
#undef AL
    c->partial_cap = std::max({blocks_for(n * c->dim, 256), blocks_for(n, 4), blocks_for(n, 256), 1});
    // CG: (||r||^2, gamma) partials in two parity slots + the SpMV's delta partials
    if ((rc = dalloc(c, &c->partial, 5 * (int64_t)c->partial_cap))) return fail(rc);
    if (hipHostMalloc((void **)&c->h_scal, sizeof(double) * (SC_COUNT + 2)) != hipSuccess ||
        hipHostMalloc((void **)&c->h_iflag, sizeof(int32_t) * 8) != hipSuccess) {
        c->err = "hipHostMalloc failed";
        return fail(LPM_E_NOMEM);
    }
    std::vector<double> sv(LPM_NSV * n, 0.0), mat(LPM_NMAT * n, 0.0);
    if (c->model == LPM_FATIGUE_HCF)  // particle_fatigue_hcf.h:24-25, 229-236
        for (int64_t i = 0; i < n; ++i) {
            sv[LPM_NSV * i + 1] = (double)INT_MAX;
            mat[LPM_NMAT * i + 5] = 1.0;
            mat[LPM_NMAT * i + 6] = 1.108;
        }
    const bool ok = up(c->sv, sv.data(), sizeof(double) * LPM_NSV * n) &&
                    up(c->sv_last, sv.data(), sizeof(double) * LPM_NSV * n) &&
                    up(c->mat, mat.data(), sizeof(double) * LPM_NMAT * n);
    if (!ok) { c->err = "state upload failed"; return fail(LPM_E_HIP); }
    if ((rc = launch_pattern(c))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { c->err = "sync failed"; return fail(LPM_E_HIP); }
    *out = c;
    return LPM_OK;
}

int lpm_create(const lpm_topology *t, lpm_ctx **out) { return create(t, nullptr, 0, out); }

int lpm_create_cut(const lpm_topology *t, const lpm_cut *cuts, int32_t ncut, lpm_ctx **out) {
    return create(t, cuts, ncut, out);
}

int lpm_destroy(lpm_ctx *c) {
    if (!c) return LPM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->bptr, c->bsplit, c->p2, c->opp, c->nbarr, c->cptr, c->cidx, c->tslot, c->nlptr, c->nlidx, c->xyz,
                    c->xyz0, c->xyz_last, c->Pin, c->Pex, c->stress, c->damage, c->damage_last, c->damage_visual,
                    c->Ddot_local, c->Ddot_nonlocal, c->nonlocal_L, c->sv, c->sv_last, c->mat, c->dLe_total, c->TdLe_total,
                    c->cs_sum, c->type, c->cons, c->dis0, c->dis, c->dis_last, c->dLe, c->dLp, c->dLp_last, c->bstrain, c->Kn,
                    c->Tv, c->cs, c->bforce, c->bforce_last, c->bdamage, c->bdamage_last, c->residual, c->disp, c->diag64,
                    c->aval, c->invdiag, c->v_x, c->v_r, c->v_u, c->v_p, c->v_s, c->v_w, c->scal, c->partial, c->counter, c->iflag, c->sp, c->scol,
                    c->scol16, c->pl_row, c->pl_slot, c->pl_rng, c->pl_b3, c->pl_own,
                    c->spec_of, c->spec_w, c->spec_blk, c->spec_first_d, c->spec_part, c->xyz_prev, c->xyz_prev2,
                    c->elim};
    for (void *p : ptrs)
        if (p) hipFree(p);
    dump_free(c);
    coarse_free(c);
    fd_free(c);
    fat_record_clear(c);
    for (void *p : c->ckpt)
        if (p) hipFree(p);
    if (c->cg_exec) hipGraphExecDestroy(c->cg_exec);
    if (c->adj_scr) hipFree(c->adj_scr);
    if (c->fat_xyz) hipFree(c->fat_xyz);
    if (c->fat_range) hipFree(c->fat_range);
    for (auto &g : c->groups)
        if (g.ids) hipFree(g.ids);
    if (c->h_scal) hipHostFree(c->h_scal);
    if (c->h_iflag) hipHostFree(c->h_iflag);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return LPM_OK;
}

int64_t lpm_size(const lpm_ctx *c, const char *what) {
    if (!c || !what) return -1;
    const std::string w(what);
    if (w == "n") return c->n;
    if (w == "dim") return c->dim;
    if (w == "nbonds") return c->nbond;
    if (w == "nconn") return c->nconn;
    if (w == "nnl") return c->nnl;
    if (w == "ndof") return c->n * c->dim;
    if (w == "knnz") return c->nconn * c->dim * c->dim;
    if (w == "maxconn") return c->maxconn;
    if (w == "maxbond") return c->maxbond;
    if (w == "nl_ptr") return c->nlptr ? c->n + 1 : 0;
    if (w == "nslots") return c->nslots;
    return -1;
}

int lpm_set_mode(lpm_ctx *c, const char *name, double v) {
    if (!c || !name) return LPM_E_ARG;
    const std::string w(name);
    // a changed setting (precision, tolerances, preconditioner ...) may change every solve's
    // length: the recorded counts that size the first CG batches start over
    for (int &h : c->cg_hist) h = -1;
    if (w == "diag10_literal") c->diag10_literal = (int)v;
    else if (w == "cg_transpose") c->cg_transpose = (int)v;
    else if (w == "dirichlet_static") c->dirichlet_static = (int)v;
    else if (w == "rollback_pex") c->rollback_pex = (int)v;
    else if (w == "precision") {
        if ((int)v != c->precision) c->k_ready = false;
        c->precision = (int)v ? 1 : 0;
    } else if (w == "cg_tol") c->cg_tol = v;
    else if (w == "cg_maxit") c->cg_maxit = (int64_t)v;
    else if (w == "precond") {
        if (v != 0 && v != 1 && v != 2) return lpm_fail(c, LPM_E_ARG, "precond must be 0, 1 or 2");
        c->precond = (int)v;
    } else if (w == "coarse_cells") {
        if (c->agg_of) return lpm_fail(c, LPM_E_STATE, "coarse_cells is fixed once the coarse space exists");
        c->coarse_cells = std::max(0, (int)v);
    } else if (w == "coarse_budget") {
        if (c->agg_of) return lpm_fail(c, LPM_E_STATE, "coarse_budget is fixed once the coarse space exists");
        if (v < 1 || v > 1024) return lpm_fail(c, LPM_E_ARG, "coarse_budget must lie in 1..1024");
        c->coarse_budget = (int)v;
    }
    else if (w == "max_NR_iter") c->max_NR_iter = (int)v;
    else if (w == "tol_NR") c->tol_NR = v;
    else if (w == "verbose") c->verbose = (int)v;
    else if (w == "cg_batch") c->cg_batch = std::max(1, (int)v);
    else if (w == "cg_batch_next") c->cg_batch_next = std::max(0, (int)v);
    else if (w == "cg_batch_auto") c->cg_batch_auto = (int)v ? 1 : 0;
    else if (w == "asm_plan") c->asm_plan = (int)v ? 1 : 0;
    else if (w == "asm_tune") {
        c->asm_tune = (int)v ? 1 : 0;
        c->asm_tune_key = -1;
    } else if (w == "asm_wpb") {
        if (v != 1 && v != 2 && v != 4) return lpm_fail(c, LPM_E_ARG, "asm_wpb must be 1, 2 or 4");
        c->asm_wpb = (int)v;
    }
    else if (w == "nl_cells") c->nl_cells = (int)v ? 1 : 0;
    else if (w == "spmv_nt") {
        c->spmv_nt = (int)v ? 1 : 0;
        c->tune_key = -1;
    }
    else if (w == "cg_forcing") c->cg_forcing = (int)v ? 1 : 0;
    else if (w == "coarse_graph") c->coarse_graph = (int)v ? 1 : 0;
    else if (w == "cg_graph") c->cg_graph = (int)v ? 1 : 0;
    else if (w == "coarse_reuse") c->coarse_reuse = (int)v ? 1 : 0;
    else if (w == "coarse_rebuild") {
        if ((int)v) c->coarse_stale = true;  // the next assembly rebuilds the coarse space
    }
    else if (w == "predictor") {
        if (v != 0 && v != 1 && v != 2) return lpm_fail(c, LPM_E_ARG, "predictor must be 0, 1 or 2");
        c->predictor = (int)v;
        c->pred_hist = 0;
    }
    else if (w == "spmv_tune") {
        c->spmv_tune = (int)v ? 1 : 0;
        c->tune_key = -1;
    } else if (w == "spmv_unr") {
        if (v != 1 && v != 2 && v != 4) return lpm_fail(c, LPM_E_ARG, "spmv_unr must be 1, 2 or 4");
        c->spmv_unr = (int)v;
    } else if (w == "spmv_wpb") {
        if (v != 1 && v != 2 && v != 4 && v != 8) return lpm_fail(c, LPM_E_ARG, "spmv_wpb must be 1, 2, 4 or 8");
        c->spmv_wpb = (int)v;
    } else if (w == "spmv_coop") {
        if (v != 0 && v != 8 && v != 16) return lpm_fail(c, LPM_E_ARG, "spmv_coop must be 0, 8 or 16");
        c->spmv_coop = (int)v;
    } else if (w == "spmv_spw") {
        if (v < 1 || v > 64) return lpm_fail(c, LPM_E_ARG, "spmv_spw must be 1..64");
        c->spmv_spw = (int)v;
    } else if (w == "spmv_remap") c->spmv_remap = (int)v ? 1 : 0;
    else if (w == "spmv_idx16") c->spmv_idx16 = (int)v ? 1 : 0;
    else if (w == "stiffness_fd") {
        if ((int)v && c->model == LPM_J2PLASTICITY)
            return lpm_fail(c, LPM_E_ARG, "finite-difference stiffness is not provided for J2Plasticity");
        c->stiffness_fd = (int)v ? 1 : 0;
    } else return lpm_fail(c, LPM_E_ARG, "unknown mode " + w);
    return LPM_OK;
}

double lpm_get_mode(const lpm_ctx *c, const char *name) {
    if (!c || !name) return NAN;
    const std::string w(name);
    if (w == "diag10_literal") return c->diag10_literal;
    if (w == "cg_transpose") return c->cg_transpose;
    if (w == "dirichlet_static") return c->dirichlet_static;
    if (w == "rollback_pex") return c->rollback_pex;
    if (w == "precision") return c->precision;
    if (w == "stiffness_fd") return c->stiffness_fd;
    if (w == "fd_colors") return c->fd_ncolor;
    if (w == "cg_tol") return c->cg_tol;
    if (w == "cg_maxit") return (double)c->cg_maxit;
    if (w == "precond") return c->precond;
    if (w == "max_NR_iter") return c->max_NR_iter;
    if (w == "tol_NR") return c->tol_NR;
    if (w == "verbose") return c->verbose;
    if (w == "cg_batch") return c->cg_batch;
    if (w == "cg_batch_next") return c->cg_batch_next;
    if (w == "cg_batch_auto") return c->cg_batch_auto;
    if (w == "spmv_nt") return c->spmv_nt;
    if (w == "asm_plan") return c->asm_plan;
    if (w == "asm_tune") return c->asm_tune;
    if (w == "asm_wpb") return c->asm_wpb;
    if (w == "nl_cells") return c->nl_cells;
    if (w == "spmv_tune") return c->spmv_tune;
    if (w == "cg_forcing") return c->cg_forcing;
    if (w == "coarse_graph") return c->coarse_graph;
    if (w == "cg_graph") return c->cg_graph;
    if (w == "coarse_reuse") return c->coarse_reuse;
    if (w == "predictor") return c->predictor;
    if (w == "coarse_cells") return c->coarse_cells;
    if (w == "coarse_budget") return c->coarse_budget;
    if (w == "ncoarse") return c->ncoarse;
    if (w == "spmv_unr") return c->spmv_unr;
    if (w == "spmv_wpb") return c->spmv_wpb;
    if (w == "spmv_remap") return c->spmv_remap;
    if (w == "spmv_idx16") return c->spmv_idx16;
    if (w == "spmv_index_bytes") return (c->scol16 && c->spmv_idx16) ? 2 : 4;
    if (w == "spmv_coop") return c->spmv_coop;
    if (w == "spmv_spw") return c->spmv_spw;
    return NAN;
}

int lpm_upload(lpm_ctx *c, const char *name, const void *host, int64_t count) {
    if (!c || !name || (!host && count > 0)) return LPM_E_ARG;
    Field f = field(c, name);
    if (!f.p && f.count < 0) return lpm_fail(c, LPM_E_ARG, std::string("unknown field ") + name);
    const std::string w(name);
    if (w == "bond_ptr" || w == "bond_split" || w == "bond_p2" || w == "opp" || w == "conn_ptr" || w == "conn_idx" ||
        w == "nl_ptr" || w == "nl_idx" || w == "nb")
        return lpm_fail(c, LPM_E_ARG, std::string("topology field is read-only: ") + name);
    if (count != f.count) return lpm_fail(c, LPM_E_ARG, std::string("size mismatch for ") + name);
    if (count == 0) return LPM_OK;
    LPM_HIP(c, hipMemcpyAsync(f.p, host, f.es * count, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    if (w == "xyz" || w == "xyz_last") c->pred_hist = 0;  // positions set by hand: nothing to extrapolate
    return LPM_OK;
}

int lpm_download(lpm_ctx *c, const char *name, void *host, int64_t count) {
    if (!c || !name || (!host && count > 0)) return LPM_E_ARG;
    const std::string w(name);
    if (w == "reaction") {  // Solver::reaction_force (solver.h:119-127): Pin at constrained DoFs in (particle, axis) order
        std::vector<double> P(3 * c->n);
        std::vector<int32_t> cons(3 * c->n);
        LPM_HIP(c, hipMemcpyAsync(P.data(), c->Pin, sizeof(double) * 3 * c->n, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipMemcpyAsync(cons.data(), c->cons, sizeof(int32_t) * 3 * c->n, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        std::vector<double> R;
        for (int64_t i = 0; i < c->n; ++i)
            for (int k = 0; k < c->dim; ++k)
                if (cons[3 * i + k] == 1) R.push_back(P[3 * i + k]);
        if (host == nullptr) return (int)R.size();
        if (count != (int64_t)R.size()) return lpm_fail(c, LPM_E_ARG, "size mismatch for reaction");
        std::memcpy(host, R.data(), sizeof(double) * R.size());
        return LPM_OK;
    }
    Field f = field(c, name);
    if (!f.p && f.count < 0) return lpm_fail(c, LPM_E_ARG, std::string("unknown field ") + name);
    if (count != f.count) return lpm_fail(c, LPM_E_ARG, std::string("size mismatch for ") + name);
    if (count == 0) return LPM_OK;
    LPM_HIP(c, hipMemcpyAsync(host, f.p, f.es * count, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_download_matrix(lpm_ctx *c, int32_t *rowptr, int32_t *col, double *val, int64_t cap) {
    if (!c || !rowptr || !col || !val) return LPM_E_ARG;
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "no assembled matrix");
    const int D = c->dim;
    const int64_t nnz = c->nconn * D * D;
    if (cap < nnz) return lpm_fail(c, LPM_E_ARG, "matrix buffer too small");
    std::vector<int32_t> cp(c->n + 1), ci(c->nconn);
    LPM_HIP(c, hipMemcpyAsync(cp.data(), c->cptr, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(ci.data(), c->cidx, sizeof(int32_t) * c->nconn, hipMemcpyDeviceToHost, c->stream));
    const int64_t ebl = (int64_t)D * D;
    std::vector<int64_t> sp(c->nslice + 1);
    LPM_HIP(c, hipMemcpyAsync(sp.data(), c->sp, sizeof(int64_t) * (c->nslice + 1), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> a(ebl * c->nslots);
    if (c->precision == 0) {
        std::vector<float> af(ebl * c->nslots);
        LPM_HIP(c, hipMemcpyAsync(af.data(), c->aval, sizeof(float) * af.size(), hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < af.size(); ++k) a[k] = af[k];
    } else {
        LPM_HIP(c, hipMemcpyAsync(a.data(), c->aval, sizeof(double) * a.size(), hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
    }
    int64_t o = 0;
    for (int64_t i = 0; i < c->n; ++i) {
        const int c0 = cp[i], nc = cp[i + 1] - c0;
        const int64_t base = ebl * sp[i >> 6] + (i & 63);
        for (int r = 0; r < D; ++r) {
            rowptr[D * i + r] = (int32_t)o;
            for (int s = 0; s < nc; ++s)
                for (int q = 0; q < D; ++q) {
                    col[o] = ci[c0 + s] * D + q;
                    val[o] = a[base + ((int64_t)s * ebl + r * D + q) * 64];
                    ++o;
                }
        }
    }
    rowptr[D * c->n] = (int32_t)o;
    return LPM_OK;
}

int lpm_set_material(lpm_ctx *c, int64_t id, double L, const double *Ce, int literal, const double *params, int32_t np) {
    if (!c || !Ce || np < 0 || np > LPM_NMAT || (np > 0 && !params)) return LPM_E_ARG;
    if (id >= c->n) return lpm_fail(c, LPM_E_ARG, "particle id out of range");
    double map[9], uc[8];
    lpm_unit_cell(c->lattice, c->radius, literal, uc, map);
    // KnTv = el_mapping * Ce (x radius in 3D), particle_elastic.h:90-93, accumulated like Eigen's GEMV
    double KnTv[3];
    for (int r = 0; r < 3; ++r) {
        double s = map[3 * r] * Ce[0];
        s = s + map[3 * r + 1] * Ce[1];
        s = s + map[3 * r + 2] * Ce[2];
        KnTv[r] = c->dim == 2 ? s : s * c->radius;
    }
    const int64_t i0 = id < 0 ? 0 : id, i1 = id < 0 ? c->n : id + 1;
    std::vector<int32_t> bp(i1 - i0 + 1), bs(i1 - i0);
    LPM_HIP(c, hipMemcpyAsync(bp.data(), c->bptr + i0, sizeof(int32_t) * (i1 - i0 + 1), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(bs.data(), c->bsplit + i0, sizeof(int32_t) * (i1 - i0), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t b0 = bp[0], b1 = bp[i1 - i0];
    std::vector<double> kn(b1 - b0), tv(b1 - b0), mat(LPM_NMAT * (i1 - i0), 0.0), Ls(i1 - i0, L);
    for (int64_t i = i0; i < i1; ++i) {
        for (int64_t b = bp[i - i0]; b < bp[i - i0 + 1]; ++b) {
            const int layer = b < bs[i - i0] ? 0 : 1;
            if (c->lattice == LPM_HEXAGON2D) { kn[b - b0] = KnTv[0]; tv[b - b0] = KnTv[1]; }
            else { kn[b - b0] = KnTv[layer]; tv[b - b0] = KnTv[2]; }
        }
        double *m = &mat[LPM_NMAT * (i - i0)];
        if (c->model == LPM_FATIGUE_HCF) { m[5] = 1.0; m[6] = 1.108; }
        for (int k = 0; k < np; ++k) m[k] = params[k];
    }
    if (b1 > b0) {
        LPM_HIP(c, hipMemcpyAsync(c->Kn + b0, kn.data(), sizeof(double) * (b1 - b0), hipMemcpyHostToDevice, c->stream));
        LPM_HIP(c, hipMemcpyAsync(c->Tv + b0, tv.data(), sizeof(double) * (b1 - b0), hipMemcpyHostToDevice, c->stream));
    }
    LPM_HIP(c, hipMemcpyAsync(c->mat + LPM_NMAT * i0, mat.data(), sizeof(double) * mat.size(), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->nonlocal_L + i0, Ls.data(), sizeof(double) * Ls.size(), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_search_nonlocal(lpm_ctx *c, double ratio) {
    if (!c) return LPM_E_ARG;
    return launch_nonlocal_search(c, ratio);
}

int lpm_group_create(lpm_ctx *c, int64_t m, const int32_t *ids, int32_t *gid) {
    if (!c || !gid || m < 0 || (m > 0 && !ids)) return LPM_E_ARG;
    std::vector<int32_t> v(ids, ids + m);
    for (int32_t i : v)
        if (i < 0 || i >= c->n) return lpm_fail(c, LPM_E_ARG, "group id out of range");
    std::vector<int32_t> s(v);
    std::sort(s.begin(), s.end());
    Group g;
    g.n = m;
    g.dup = std::adjacent_find(s.begin(), s.end()) != s.end();
    LPM_HIP(c, hipMalloc(&g.ids, sizeof(int32_t) * (m > 0 ? m : 1)));
    if (m > 0) LPM_HIP(c, hipMemcpyAsync(g.ids, v.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    c->groups.push_back(g);
    *gid = (int32_t)c->groups.size() - 1;
    return LPM_OK;
}

int lpm_apply_bc(lpm_ctx *c, const lpm_load_step *s) {
    if (!c) return LPM_E_ARG;
    RC(apply_bc(c, s));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_update_geometry(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    return launch_geometry_force(c, false);
}

int lpm_update_force_state(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    RC(launch_bond_force(c));
    return launch_particle_force(c);
}

int lpm_update_rr(lpm_ctx *c, double *nr, double *nR) {
    if (!c) return LPM_E_ARG;
    double r2 = 0, R2 = 0;
    RC(launch_residual(c, &r2, &R2));
    if (nr) *nr = std::sqrt(r2);
    if (nR) *nR = std::sqrt(R2);
    return LPM_OK;
}

int lpm_calc_stiffness(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    return calc_stiffness(c);
}

int lpm_update_stiffness_disp_bc(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    return update_stiffness_disp_bc(c);
}

int lpm_solve_linear_system(lpm_ctx *c, int32_t *iters, double *error) {
    if (!c) return LPM_E_ARG;
    int it = 0;
    RC(solve_linear_system(c, &it, error));
    if (iters) *iters = it;
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_newton_iteration(lpm_ctx *c, int32_t *n_iter) {
    if (!c || !n_iter) return LPM_E_ARG;
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "NewtonIteration before calcStiffness");
    int ni = 0;
    RC(newton_iteration(c, &ni));
    *n_iter = ni;
    return LPM_OK;
}

int lpm_update_state_var(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    return launch_state_var(c);
}

int lpm_update_static_damage(lpm_ctx *c, int32_t undamaged, int32_t *any) {
    if (!c) return LPM_E_ARG;
    int a = 0;
    RC(launch_static_damage(c, undamaged, &a));
    if (any) *any = a;
    return LPM_OK;
}

int lpm_update_fatigue_damage(lpm_ctx *c, double dNdt, int32_t undamaged, int32_t *any) {
    if (!c) return LPM_E_ARG;
    int a = 0;
    RC(launch_fatigue_damage(c, dNdt, undamaged, &a));
    if (any) *any = a;
    return LPM_OK;
}

int lpm_store_state(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    return store_state(c);
}

int lpm_reset_state(lpm_ctx *c, int32_t reset_xyz) {
    if (!c) return LPM_E_ARG;
    return reset_state(c, reset_xyz != 0);
}

int lpm_solve_step_static(lpm_ctx *c, const lpm_load_step *s, int32_t undamaged, int32_t *converged) {
    if (!c || !converged) return LPM_E_ARG;
    RC(check_steps(c, s, 1));
    int ok = 0;
    RC(solve_step_static(c, s, undamaged, &ok));
    *converged = ok;
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_solve_problem_static(lpm_ctx *c, const lpm_load_step *steps, int32_t n_in, int32_t undamaged, int32_t *n_steps) {
    if (!c) return LPM_E_ARG;
    return solve_problem(c, steps, n_in, undamaged, false, n_steps);
}

int lpm_solve_step_fatigue(lpm_ctx *c, const lpm_load_step *s, int32_t *converged) {
    if (!c || !converged) return LPM_E_ARG;
    RC(check_steps(c, s, 1));
    int ok = 0;
    RC(solve_step_fatigue(c, s, &ok));
    *converged = ok;
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

int lpm_solve_problem_fatigue_static(lpm_ctx *c, const lpm_load_step *steps, int32_t n_in, int32_t *n_steps) {
    if (!c) return LPM_E_ARG;
    return solve_problem(c, steps, n_in, -1, true, n_steps);
}

int lpm_solve_problem_cyclic(lpm_ctx *c, int32_t load_type, int32_t n_interval, const double *spec, int64_t nspec,
                             const int32_t *groups, int32_t ngroups, int32_t tmode, double tau, int32_t undamaged,
                             int32_t max_cycles_run, int32_t *cycles) {
    if (!c || !groups) return LPM_E_ARG;
    const int need = load_type == LPM_LOAD_DOGBONE_DISP ? 2 : 3;
    if (ngroups < need) return lpm_fail(c, LPM_E_ARG, "load template needs more groups");
    const double one = 1.0;
    return solve_cyclic(c, load_type, n_interval, spec, nspec, groups, &one, 1, tmode, tau, undamaged, max_cycles_run,
                        cycles);
}

int lpm_solve_problem_cyclic_specimens(lpm_ctx *c, int32_t load_type, int32_t n_interval, const double *spec,
                                       int64_t nspec, const int32_t *groups, const double *scale, int32_t tmode,
                                       double tau, int32_t undamaged, int32_t max_cycles_run, int32_t *cycles) {
    if (!c || !groups) return LPM_E_ARG;
    if (c->nspec <= 0) return lpm_fail(c, LPM_E_STATE, "no specimens declared (lpm_set_specimens)");
    std::vector<double> ones(c->nspec, 1.0);
    return solve_cyclic(c, load_type, n_interval, spec, nspec, groups, scale ? scale : ones.data(), c->nspec, tmode, tau,
                        undamaged, max_cycles_run, cycles);
}

double lpm_stat(const lpm_ctx *c, const char *name) {
    if (!c || !name) return NAN;
    const std::string w(name);
    if (w == "newton_iters") return (double)c->st_newton;
    if (w == "cg_iters") return (double)c->st_cg_iters;
    if (w == "coarse_builds") return (double)c->st_coarse_builds;
    if (w == "cg_solves") return (double)c->st_cg_solves;
    if (w == "assemblies") return (double)c->st_asm;
    if (w == "spmv_launches") return (double)c->st_spmv;
    if (w == "spmv_ms") return c->st_spmv_ms;
    if (w == "spmv_launches_timed") return (double)c->st_spmv_timed;
    if (w == "last_cg_iters") return c->last_cg_iters;
    if (w == "last_cg_error") return c->last_cg_err;
    if (w == "spmv_tuned_ms") return c->tune_best_ms;
    if (w == "asm_tuned_ms") return c->asm_tuned_ms;
    return NAN;
}

int lpm_reset_stats(lpm_ctx *c) {
    if (!c) return LPM_E_ARG;
    c->st_newton = c->st_cg_iters = c->st_cg_solves = c->st_asm = c->st_spmv = c->st_spmv_timed = 0;
    c->st_coarse_builds = 0;
    c->st_spmv_ms = 0;
    return LPM_OK;
}

int lpm_set_timing(lpm_ctx *c, int on) {
    if (!c) return LPM_E_ARG;
    c->timing = on > 0 ? on : 0;
    return LPM_OK;
}

int lpm_spmv_bench(lpm_ctx *c, int32_t reps, double *ms) {
    if (!c || !ms) return LPM_E_ARG;
    return spmv_bench(c, reps, ms);
}

int lpm_force_jvp(lpm_ctx *c, const double *v, double *jv) {
    if (!c || !v || !jv) return LPM_E_ARG;
    const int64_t nd = c->n * c->dim;
    double *d = nullptr;
    LPM_HIP(c, hipMalloc(&d, sizeof(double) * 2 * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (hipMemcpyAsync(d, v, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "jvp upload");
    if (!rc) rc = force_jvp(c, d, d + nd);
    if (!rc && hipMemcpyAsync(jv, d + nd, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "jvp download");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(d);
    return rc;
}

int lpm_force_vjp(lpm_ctx *c, const double *w, double *y) {
    if (!c || !w || !y) return LPM_E_ARG;
    const int64_t nd = c->n * c->dim;
    double *d = nullptr;
    LPM_HIP(c, hipMalloc(&d, sizeof(double) * 2 * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (hipMemcpyAsync(d, w, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "vjp upload");
    if (!rc) rc = force_vjp(c, d, d + nd);
    if (!rc && hipMemcpyAsync(y, d + nd, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "vjp download");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(d);
    return rc;
}

int lpm_adjoint_reaction(lpm_ctx *c, int32_t group, int32_t axis, double *dF_dKn, double *dF_dTv, double *F,
                         int32_t *refinements, double *rel_residual) {
    if (!c) return LPM_E_ARG;
    if (axis < 0 || axis >= c->dim) return lpm_fail(c, LPM_E_ARG, "axis out of range");
    RC(check_group(c, group));
    int it = 0;
    double rr = 0;
    RC(adjoint_reaction(c, group, axis, dF_dKn, dF_dTv, F, &it, &rr));
    if (refinements) *refinements = it;
    if (rel_residual) *rel_residual = rr;
    return LPM_OK;
}

int lpm_adjoint_linear(lpm_ctx *c, const double *w_pin, const double *w_x, double *dG_dKn, double *dG_dTv,
                       double *dG_dbdamage, double *G, int32_t *refinements, double *rel_residual) {
    if (!c) return LPM_E_ARG;
    if (!w_pin && !w_x) return lpm_fail(c, LPM_E_ARG, "no objective weights");
    const int64_t nd = c->n * c->dim;
    double *d = nullptr;
    LPM_HIP(c, hipMalloc(&d, sizeof(double) * 2 * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (w_pin && hipMemcpyAsync(d, w_pin, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    if (!rc && w_x && hipMemcpyAsync(d + nd, w_x, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    int it = 0;
    double rr = 0;
    if (!rc) rc = adjoint_linear(c, w_pin ? d : nullptr, w_x ? d + nd : nullptr, dG_dKn, dG_dTv, dG_dbdamage, G, &it, &rr);
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(d);
    if (rc) return rc;
    if (refinements) *refinements = it;
    if (rel_residual) *rel_residual = rr;
    return LPM_OK;
}

int lpm_adjoint_fatigue_params(lpm_ctx *c, const double *w_pin, const double *w_x, double *dG_dA, double *dG_dB,
                               double *dG_dd, double *G, int32_t *refinements, double *rel_residual) {
    if (!c) return LPM_E_ARG;
    if (!w_pin && !w_x) return lpm_fail(c, LPM_E_ARG, "no objective weights");
    if (c->nonlocal_L && c->n > 0) {
        // the gather over each particle's own nonlocal list needs symmetric lists: one length
        std::vector<double> L(c->n);
        LPM_HIP(c, hipMemcpyAsync(L.data(), c->nonlocal_L, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        for (double v : L)
            if (v != L[0]) return lpm_fail(c, LPM_E_STATE, "fatigue-parameter sensitivity needs one nonlocal length");
    }
    const int64_t nd = c->n * c->dim;
    double *d = nullptr;
    LPM_HIP(c, hipMalloc(&d, sizeof(double) * 2 * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (w_pin && hipMemcpyAsync(d, w_pin, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    if (!rc && w_x && hipMemcpyAsync(d + nd, w_x, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    int it = 0;
    double rr = 0;
    if (!rc)
        rc = adjoint_fatigue_params(c, w_pin ? d : nullptr, w_x ? d + nd : nullptr, dG_dA, dG_dB, dG_dd, G, &it, &rr);
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(d);
    if (rc) return rc;
    if (refinements) *refinements = it;
    if (rel_residual) *rel_residual = rr;
    return LPM_OK;
}

int lpm_fatigue_record(lpm_ctx *c, int32_t on) {
    if (!c) return LPM_E_ARG;
    fat_record_clear(c);
    c->fat_rec = on != 0;
    return LPM_OK;
}

int lpm_adjoint_fatigue_path(lpm_ctx *c, const double *w_pin, const double *w_x, double *dG_dA, double *dG_dB,
                             double *dG_dd, double *G, int32_t *solves, double *max_rel_residual) {
    if (!c) return LPM_E_ARG;
    if (!w_pin && !w_x) return lpm_fail(c, LPM_E_ARG, "no objective weights");
    if (c->nonlocal_L && c->n > 0) {
        std::vector<double> L(c->n);
        LPM_HIP(c, hipMemcpyAsync(L.data(), c->nonlocal_L, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        for (double v : L)
            if (v != L[0]) return lpm_fail(c, LPM_E_STATE, "fatigue-path sensitivity needs one nonlocal length");
    }
    const int64_t nd = c->n * c->dim;
    double *d = nullptr;
    LPM_HIP(c, hipMalloc(&d, sizeof(double) * 2 * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (w_pin && hipMemcpyAsync(d, w_pin, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    if (!rc && w_x && hipMemcpyAsync(d + nd, w_x, sizeof(double) * nd, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint weights upload");
    int ns = 0;
    double rr = 0;
    const bool rec = c->fat_rec;
    c->fat_rec = false;  // the reverse pass restores states; nothing it does is part of the path
    if (!rc) rc = adjoint_fatigue_path(c, w_pin ? d : nullptr, w_x ? d + nd : nullptr, dG_dA, dG_dB, dG_dd, G, &ns, &rr);
    c->fat_rec = rec;
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(d);
    if (rc) return rc;
    if (solves) *solves = ns;
    if (max_rel_residual) *max_rel_residual = rr;
    return LPM_OK;
}

// The arrays a checkpoint holds, in buffer order: {pointer, doubles} (cons, int32, is stored
// after them, padded to whole doubles).
static std::vector<std::pair<double *, int64_t>> ckpt_fields(lpm_ctx *c) {
    const int64_t n = c->n, nb = c->nbond;
    return {{c->xyz, 3 * n},       {c->xyz_last, 3 * n}, {c->Pex, 3 * n},           {c->damage, n},
            {c->damage_last, n},   {c->sv, LPM_NSV * n}, {c->sv_last, LPM_NSV * n}, {c->dis, nb},
            {c->dis_last, nb},     {c->dLp, nb},         {c->dLp_last, nb},         {c->bforce, nb},
            {c->bforce_last, nb},  {c->bdamage, nb},     {c->bdamage_last, nb}};
}

static int ckpt_copy_buf(lpm_ctx *c, void **pbuf, bool save) {
    const auto f = ckpt_fields(c);
    int64_t tot = 0;
    for (auto &x : f) tot += x.second;
    const int64_t ncons = (3 * c->n + 1) / 2;  // int32 flags in doubles
    if (!*pbuf) LPM_HIP(c, hipMalloc(pbuf, sizeof(double) * (tot + ncons)));
    double *buf = (double *)*pbuf;
    int64_t off = 0;
    for (auto &x : f) {
        if (x.second > 0) {
            if (save) LPM_HIP(c, hipMemcpyAsync(buf + off, x.first, sizeof(double) * x.second, hipMemcpyDeviceToDevice, c->stream));
            else LPM_HIP(c, hipMemcpyAsync(x.first, buf + off, sizeof(double) * x.second, hipMemcpyDeviceToDevice, c->stream));
        }
        off += x.second;
    }
    if (save) LPM_HIP(c, hipMemcpyAsync(buf + off, c->cons, sizeof(int32_t) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    else LPM_HIP(c, hipMemcpyAsync(c->cons, buf + off, sizeof(int32_t) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
    if (!save) {
        c->pred_hist = 0;  // the increments before the restored state are not known
        c->k_ready = false;
        c->k_version++;         // constraint flags changed: the coarse space is rebuilt with K
        RC(geometry_force_state(c));
    }
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

extern "C++" int ckpt_save_buf(lpm_ctx *c, void **buf) { return ckpt_copy_buf(c, buf, true); }

extern "C++" int ckpt_restore_buf(lpm_ctx *c, void *buf) { return ckpt_copy_buf(c, &buf, false); }

static int checkpoint_copy(lpm_ctx *c, int32_t slot, bool save) {
    if (slot < 0 || slot >= LPM_MAX_CHECKPOINTS) return lpm_fail(c, LPM_E_ARG, "checkpoint slot out of range");
    if ((int64_t)c->ckpt.size() <= slot) c->ckpt.resize(slot + 1, nullptr);
    if (!save && !c->ckpt[slot]) return lpm_fail(c, LPM_E_STATE, "checkpoint slot " + std::to_string(slot) + " is empty");
    return ckpt_copy_buf(c, &c->ckpt[slot], save);
}

int lpm_checkpoint_save(lpm_ctx *c, int32_t slot) {
    if (!c) return LPM_E_ARG;
    return checkpoint_copy(c, slot, true);
}

int lpm_checkpoint_restore(lpm_ctx *c, int32_t slot) {
    if (!c) return LPM_E_ARG;
    return checkpoint_copy(c, slot, false);
}

int lpm_checkpoint_release(lpm_ctx *c, int32_t slot) {
    if (!c) return LPM_E_ARG;
    if (slot < -1 || slot >= LPM_MAX_CHECKPOINTS) return lpm_fail(c, LPM_E_ARG, "checkpoint slot out of range");
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    for (int32_t s = 0; s < (int32_t)c->ckpt.size(); ++s)
        if ((slot == -1 || s == slot) && c->ckpt[s]) {
            hipFree(c->ckpt[s]);
            c->ckpt[s] = nullptr;
        }
    return LPM_OK;
}

int lpm_set_specimens(lpm_ctx *c, int32_t nspec, const int64_t *first) {
    if (!c || nspec < 0 || (nspec > 0 && !first)) return LPM_E_ARG;
    if (nspec == 0) {
        c->nspec = 0;
        return LPM_OK;
    }
    if (first[0] != 0 || first[nspec] != c->n) return lpm_fail(c, LPM_E_ARG, "specimen offsets must run from 0 to n");
    for (int s = 0; s < nspec; ++s)
        if (first[s + 1] <= first[s]) return lpm_fail(c, LPM_E_ARG, "specimen offsets must increase");
    // no bond may join two specimens
    std::vector<int32_t> bp(c->n + 1), p2(c->nbond);
    LPM_HIP(c, hipMemcpyAsync(bp.data(), c->bptr, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(p2.data(), c->p2, sizeof(int32_t) * c->nbond, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> sof(c->n);
    for (int s = 0; s < nspec; ++s)
        for (int64_t i = first[s]; i < first[s + 1]; ++i) sof[i] = s;
    for (int64_t i = 0; i < c->n; ++i)
        for (int b = bp[i]; b < bp[i + 1]; ++b)
            if (sof[p2[b]] != sof[i]) return lpm_fail(c, LPM_E_TOPO, "a bond joins two specimens");
    if (c->agg_of) return lpm_fail(c, LPM_E_STATE, "set the specimens before the first two-level solve");
    std::vector<int32_t> blk(nspec + 1, 0);
    for (int s = 0; s < nspec; ++s) blk[s + 1] = blk[s] + blocks_for((first[s + 1] - first[s]) * c->dim, 256);
    void *old[] = {c->spec_of, c->spec_w, c->spec_blk, c->spec_first_d, c->spec_part};
    for (void *p : old)
        if (p) hipFree(p);
    c->spec_of = nullptr;
    c->spec_w = c->spec_part = nullptr;
    c->spec_blk = nullptr;
    c->spec_first_d = nullptr;
    LPM_HIP(c, hipMalloc(&c->spec_of, sizeof(int32_t) * c->n));
    LPM_HIP(c, hipMalloc(&c->spec_w, sizeof(double) * nspec));
    LPM_HIP(c, hipMalloc(&c->spec_blk, sizeof(int32_t) * (nspec + 1)));
    LPM_HIP(c, hipMalloc(&c->spec_first_d, sizeof(int64_t) * (nspec + 1)));
    LPM_HIP(c, hipMalloc(&c->spec_part, sizeof(double) * 2 * std::max(blk[nspec], 1)));
    LPM_HIP(c, hipMemcpyAsync(c->spec_of, sof.data(), sizeof(int32_t) * c->n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->spec_blk, blk.data(), sizeof(int32_t) * (nspec + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->spec_first_d, first, sizeof(int64_t) * (nspec + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    c->nspec = nspec;
    c->spec_nblk = blk[nspec];
    c->spec_first.assign(first, first + nspec + 1);
    c->spec_newton.assign(nspec, 0);
    c->spec_ok.assign(nspec, 1);
    return LPM_OK;
}

int lpm_specimen_stat(const lpm_ctx *c, const char *name, double *out) {
    if (!c || !name || !out) return LPM_E_ARG;
    const std::string w(name);
    for (int s = 0; s < c->nspec; ++s) {
        if (w == "newton_iters") out[s] = c->spec_newton.empty() ? 0.0 : c->spec_newton[s];
        else if (w == "converged") out[s] = c->spec_ok.empty() ? 1.0 : c->spec_ok[s];
        else return LPM_E_ARG;
    }
    return LPM_OK;
}

int lpm_copy_bench(lpm_ctx *c, int64_t bytes, int32_t reps, double *ms) {
    if (!c || !ms) return LPM_E_ARG;
    return copy_bench(c, bytes, reps, ms);
}

void *lpm_stream(lpm_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"
