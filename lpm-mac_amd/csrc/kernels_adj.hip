// Tangent (Jacobian-vector product) of the internal force and the adjoint sensitivity of a
// linear objective of the converged state, G = a.Pin + b.x (reaction sums and displacement
// fields), with respect to the bond parameters Kn, Tv and the bond damage.
//
// The internal force Pin(x) is the one of Assembly::updateGeometry + updateForceState
// (particle.h:163-215, bond.h:30-39, particle_elastic.h:38-50): with d_b = x_i - x_j,
// dis_b = |d_b|, cs_b = d_b / dis_b, dLe_b = dis_b - dis0_b - dLp_b,
// dLe_total(i, l) = sum of dLe_b over the layer-l bonds of i,
// f_b = (2 Kn_b dLe_b + 2 Tv_b dLe_total(i, l_b)) (1 - bdamage_b) and
// Pin_i = sum_b cs_b (f_b + f_opp(b)) / 2.
// Its exact directional derivative (damage, broken bonds and plastic lengths dLp frozen):
//   d dis_b = cs_b . (v_i - v_j),  d cs_b = ((v_i - v_j) - cs_b d dis_b) / dis_b,
//   d f_b = (2 Kn_b d dis_b + 2 Tv_b sum_{layer} d dis) (1 - bdamage_b),
//   d Pin_i = sum_b [ d cs_b (f_b + f_opp)/2 + cs_b (d f_b + d f_opp)/2 ].
// The analytic stiffness of the reference (derivative.h) drops small terms (derivative.h:5),
// so it differs from this Jacobian by O(strain); it serves as the inner operator of the
// adjoint solve, whose residual comes from the exact transposed product J^T w (k_vjp_*).
#include <algorithm>
#include <vector>

#include "lpm_internal.h"
#include "reduce.h"

namespace {

// pass 1, one thread per particle: d dis, d cs and d f of each of its bonds
__global__ __launch_bounds__(256) void k_jvp_bond(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                   const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                   const double *__restrict__ dis, const double *__restrict__ cs,
                                                   const double *__restrict__ Kn, const double *__restrict__ Tv,
                                                   const double *__restrict__ bdamage, const double *__restrict__ v,
                                                   double *__restrict__ dcs, double *__restrict__ dfb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    double vi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) vi[k] = v[(int64_t)dim * i + k];
    double tot[2] = {0.0, 0.0};
    for (int b = b0; b < b1; ++b) {
        const int64_t j = p2[b];
        const bool broken = fabs(bdamage[b] - 1.0) < LPM_EPS;
        double dd[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) dd[k] = vi[k] - v[(int64_t)dim * j + k];
        const double c0 = cs[3 * b], c1 = cs[3 * b + 1], c2 = cs[3 * b + 2];
        const double ddis = broken ? 0.0 : (c0 * dd[0] + c1 * dd[1] + c2 * dd[2]);
        const double inv = broken ? 0.0 : 1.0 / dis[b];
        dcs[3 * b] = (dd[0] - c0 * ddis) * inv;
        dcs[3 * b + 1] = (dd[1] - c1 * ddis) * inv;
        dcs[3 * b + 2] = (dd[2] - c2 * ddis) * inv;
        dfb[b] = ddis;  // d dLe for now
        tot[b < bm ? 0 : 1] += ddis;
    }
    for (int b = b0; b < b1; ++b)
        dfb[b] = (2.0 * Kn[b] * dfb[b] + 2.0 * Tv[b] * tot[b < bm ? 0 : 1]) * (1.0 - bdamage[b]);
}

// pass 2, one thread per particle: d Pin_i (dim entries per particle)
__global__ __launch_bounds__(256) void k_jvp_particle(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                       const int32_t *__restrict__ opp, const double *__restrict__ cs,
                                                       const double *__restrict__ bforce, const double *__restrict__ dcs,
                                                       const double *__restrict__ dfb, double *__restrict__ jv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double P[3] = {0, 0, 0};
    for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
        const int o = opp[b];
        const double f = 0.5 * (bforce[b] + bforce[o]), df = 0.5 * (dfb[b] + dfb[o]);
        for (int k = 0; k < 3; ++k) P[k] += dcs[3 * b + k] * f + cs[3 * b + k] * df;
    }
    for (int k = 0; k < dim; ++k) jv[(int64_t)dim * i + k] = P[k];
}

// Transposed product y = J^T w of the same Jacobian. J is not symmetric once the bonds of a
// particle's layer carry different damage (the dilatation term of f_b is scaled by that bond's
// own 1 - d_b), or when a bond is broken and its reverse is not, so the adjoint needs J^T
// itself. Transposing bond by bond, with o = opp(b), i the owner of b and j its partner
// (cs, dis and the weights a vanish on a broken bond):
//   t_b = (cs_b . w_i + cs_o . w_j) / 2,
//   T(i, l) = sum over the layer-l bonds b of i of 2 Tv_b t_b (1 - d_b),
//   c_b = 2 Kn_b t_b (1 - d_b) + T(i, l_b)   (0 for a broken bond),
//   a_b = (f_b + f_o) / (2 dis_b),
//   y_i = sum over the bonds b of i of
//         a_b (I - cs_b cs_b^T) w_i - a_o (I - cs_o cs_o^T) w_j + c_b cs_b - c_o cs_o.
// pass 1, one thread per particle: c_b of each of its bonds
__global__ __launch_bounds__(256) void k_vjp_bond(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                   const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                   const int32_t *__restrict__ opp, const double *__restrict__ cs,
                                                   const double *__restrict__ Kn, const double *__restrict__ Tv,
                                                   const double *__restrict__ bdamage, const double *__restrict__ w,
                                                   double *__restrict__ cb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    double wi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) wi[k] = w[(int64_t)dim * i + k];
    double T[2] = {0.0, 0.0};
    for (int b = b0; b < b1; ++b) {
        const int64_t j = p2[b];
        const int o = opp[b];
        double t = 0.0;
        for (int k = 0; k < dim; ++k) t += cs[3 * b + k] * wi[k] + cs[3 * (int64_t)o + k] * w[(int64_t)dim * j + k];
        t *= 0.5;
        cb[b] = t;  // t_b for now
        T[b < bm ? 0 : 1] += 2.0 * Tv[b] * t * (1.0 - bdamage[b]);
    }
    for (int b = b0; b < b1; ++b) {
        const bool broken = fabs(bdamage[b] - 1.0) < LPM_EPS;
        cb[b] = broken ? 0.0 : 2.0 * Kn[b] * cb[b] * (1.0 - bdamage[b]) + T[b < bm ? 0 : 1];
    }
}

// pass 2, one thread per particle: y_i
__global__ __launch_bounds__(256) void k_vjp_particle(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                       const int32_t *__restrict__ p2, const int32_t *__restrict__ opp,
                                                       const double *__restrict__ dis, const double *__restrict__ cs,
                                                       const double *__restrict__ bforce, const double *__restrict__ bdamage,
                                                       const double *__restrict__ cb, const double *__restrict__ w,
                                                       double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) wi[k] = w[(int64_t)dim * i + k];
    double Y[3] = {0, 0, 0};
    for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
        const int64_t j = p2[b];
        const int o = opp[b];
        double wj[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) wj[k] = w[(int64_t)dim * j + k];
        const double cb_[3] = {cs[3 * b], cs[3 * b + 1], cs[3 * b + 2]};
        const double co[3] = {cs[3 * (int64_t)o], cs[3 * (int64_t)o + 1], cs[3 * (int64_t)o + 2]};
        const double f = 0.5 * (bforce[b] + bforce[o]);
        const double ab = fabs(bdamage[b] - 1.0) < LPM_EPS ? 0.0 : f / dis[b];
        const double ao = fabs(bdamage[o] - 1.0) < LPM_EPS ? 0.0 : f / dis[o];
        const double pb = cb_[0] * wi[0] + cb_[1] * wi[1] + cb_[2] * wi[2];
        const double po = co[0] * wj[0] + co[1] * wj[1] + co[2] * wj[2];
        for (int k = 0; k < 3; ++k)
            Y[k] += ab * (wi[k] - cb_[k] * pb) - ao * (wj[k] - co[k] * po) + cb[b] * cb_[k] - cb[o] * co[k];
    }
    for (int k = 0; k < dim; ++k) y[(int64_t)dim * i + k] = Y[k];
}

// Zero the entries of DoFs that the Dirichlet elimination removes (constrained, or of a fully
// damaged particle; stiffness.h:272-293).
__global__ __launch_bounds__(256) void k_mask_free(int64_t nd, int dim, const int32_t *__restrict__ cons,
                                                    const double *__restrict__ damage_visual, double *__restrict__ v) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    const int64_t i = d / dim;
    const int k = (int)(d - i * dim);
    if (cons[3 * i + k] == 1 || fabs(damage_visual[i] - 1.0) < LPM_EPS) v[d] = 0.0;
}

// indicator of (group particle, axis)
__global__ __launch_bounds__(256) void k_indicator(int64_t m, const int32_t *__restrict__ ids, int dim, int axis,
                                                    double *__restrict__ e) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) e[(int64_t)dim * ids[t] + axis] = 1.0;
}

// r = g - v ; partial ||r||^2
__global__ __launch_bounds__(256) void k_sub_norm(int64_t nd, const double *__restrict__ g, const double *__restrict__ v,
                                                   double *__restrict__ r, double *partial, int cap) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double x[1] = {0.0};
    if (d < nd) {
        const double t = g[d] - v[d];
        r[d] = t;
        x[0] = t * t;
    }
    block_partial<1>(x, partial, cap);
}

// lambda += CG solution (the CG vector is padded to VP entries per particle)
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_add_x(int64_t n, const T *__restrict__ x, double *__restrict__ lam) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < D; ++k) lam[D * i + k] += (double)x[VP * i + k];
}

// r = r + v (the displacement part of the adjoint right-hand side)
__global__ __launch_bounds__(256) void k_axpy1(int64_t nd, const double *__restrict__ v, double *__restrict__ r) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d < nd) r[d] += v[d];
}

// per-bond sensitivities with w = e - lambda: f_b enters Pin_i as cs_b f_b / 2 and Pin_j as
// cs_o f_b / 2 (o = opp(b), the reverse bond, cs_o = -cs_b unless one of the two is broken), so
// with t_b = (cs_b . w_i + cs_o . w_j) / 2:
//   dG/dKn_b = 2 t_b dLe_b (1 - bdamage_b)
//   dG/dTv_b = 2 t_b dLe_total(i, l_b) (1 - bdamage_b)
//   dG/dbdamage_b = -2 t_b (Kn_b dLe_b + Tv_b dLe_total(i, l_b))
// (a broken bond has dLe = 0, so all three vanish there)
__global__ __launch_bounds__(256) void k_param_grad(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                     const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                     const int32_t *__restrict__ opp,
                                                     const double *__restrict__ cs, const double *__restrict__ dLe,
                                                     const double *__restrict__ dLe_total, const double *__restrict__ bdamage,
                                                     const double *__restrict__ Kn, const double *__restrict__ Tv,
                                                     const double *__restrict__ e, const double *__restrict__ lam,
                                                     double *__restrict__ gKn, double *__restrict__ gTv,
                                                     double *__restrict__ gBd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) wi[k] = e[(int64_t)dim * i + k] - lam[(int64_t)dim * i + k];
    for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
        const int64_t j = p2[b], o = opp[b];
        double s = 0.0;  // 2 t_b
        for (int k = 0; k < dim; ++k)
            s += wi[k] * cs[3 * b + k] + (e[(int64_t)dim * j + k] - lam[(int64_t)dim * j + k]) * cs[3 * o + k];
        const double w = 1.0 - bdamage[b], lt = dLe_total[2 * i + (b < bsplit[i] ? 0 : 1)];
        gKn[b] = s * dLe[b] * w;
        gTv[b] = s * lt * w;
        gBd[b] = -s * (Kn[b] * dLe[b] + Tv[b] * lt);
    }
}

// ---- sensitivity of the next equilibrium to the HCF law's parameters through one fatigue update
// (SolverFatigue::updateFatigueDamage, solver_fatigue.h:75-126; ParticleFatigueHCF,
// particle_fatigue_hcf.h:127-160): D'_i = D_i + dt * Ddot_nl,i, Ddot_nl,i = sum_j w_ij Ddot_j with
// w_ij = phi_ij V_m,j / sum_k phi_ik V_m,k, Ddot_j = A_j s0_j^B_j (s2 - s1)_j^d_j, and the bond
// damage of the next force evaluation bd_b = max(bd_b, max(D_i, D_j)) (std::max order, as
// updateBondsForce, particle_fatigue_hcf.h:65-80). The energies s, the update's step dt and the
// earlier history are held fixed.

// dG/dD'_i: the bond damages that D'_i sets. Bond b = (i -> j) takes D_i when D_i >= D_j, its
// reverse (j -> i) when D_i > D_j, each only where the maximum did not stay at the old bond
// damage (bd_b = D_i). Zero where D'_i was clamped to 1 and for the excluded type.
__global__ __launch_bounds__(256) void k_fat_gd(int64_t n, int undamaged, const int32_t *__restrict__ type,
                                                 const int32_t *__restrict__ bptr, const int32_t *__restrict__ p2,
                                                 const int32_t *__restrict__ opp, const double *__restrict__ damage,
                                                 const double *__restrict__ bdamage, const double *__restrict__ gbd,
                                                 double *__restrict__ gD) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Di = damage[i];
    double s = 0.0;
    if (type[i] != undamaged && Di != 1.0) {
        for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
            const double Dj = damage[p2[b]];
            const int o = opp[b];
            if (Di >= Dj && bdamage[b] == Di) s += gbd[b];
            if (Di > Dj && bdamage[o] == Di) s += gbd[o];
        }
    }
    gD[i] = s;
}

// the nonlocal weights' normaliser sum_k phi_ik V_m,k at the update's positions
__global__ __launch_bounds__(256) void k_fat_norm(int64_t n, int undamaged, double volume, int nneighbors,
                                                   const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                   const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                   const double *__restrict__ xyz, const double *__restrict__ L,
                                                   double *__restrict__ norm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double A = 0.0;
    if (type[i] != undamaged) {
        for (int q = nlptr[i]; q < nlptr[i + 1]; ++q) {
            const int64_t j = nlidx[q];
            if (type[j] == undamaged) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double u = 1 / (1 + pow(sqrt(dx * dx + dy * dy + dz * dz) / L[i], 8.0));
            A += u * u * volume * (bptr[j + 1] - bptr[j]) / nneighbors;
        }
    }
    norm[i] = A;
}

// per particle j: S_j = sum over i with j in NL(i) of dG/dD'_i w_ij (gathered over NL(j), the
// lists being symmetric), then dG/dA_j = dt S_j s0^B r^d, dG/dB_j = dt S_j Ddot_j ln s0,
// dG/dd_j = dt S_j Ddot_j ln r (r = s2 - s1 before the update)
__global__ __launch_bounds__(256) void k_fat_params(int64_t n, int undamaged, double volume, int nneighbors, double dt,
                                                     const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                     const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                     const double *__restrict__ xyz, const double *__restrict__ L,
                                                     const double *__restrict__ norm, const double *__restrict__ gD,
                                                     const double *__restrict__ mat, const double *__restrict__ sv,
                                                     const double *__restrict__ range, double *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double gA = 0.0, gB = 0.0, gd = 0.0;
    if (type[j] != undamaged) {
        const double Vj = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
        double S = 0.0;
        for (int q = nlptr[j]; q < nlptr[j + 1]; ++q) {
            const int64_t i = nlidx[q];
            if (type[i] == undamaged || norm[i] == 0.0) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double u = 1 / (1 + pow(sqrt(dx * dx + dy * dy + dz * dz) / L[i], 8.0));
            S += gD[i] * u * u * Vj / norm[i];
        }
        const double A = mat[LPM_NMAT * j], B = mat[LPM_NMAT * j + 1], d = mat[LPM_NMAT * j + 2];
        const double s0 = sv[LPM_NSV * j], r = range[j];
        if (r > 0) {
            const double base = pow(s0, B) * pow(r, d);
            gA = dt * S * base;
            gB = s0 > 0 ? dt * S * A * base * log(s0) : 0.0;
            gd = dt * S * A * base * log(r);
        }
    }
    out[j] = gA;
    out[n + j] = gB;
    out[2 * n + j] = gd;
}

}  // namespace

static int check(lpm_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LPM_OK;
}

// the context's per-bond scratch (4 doubles per bond; kernels on the context stream use it in
// turn, so one buffer serves every product)
static int adj_scratch(lpm_ctx *c, double **p) {
    if (!c->adj_scr) LPM_HIP(c, hipMalloc(&c->adj_scr, sizeof(double) * 4 * std::max<int64_t>(c->nbond, 1)));
    *p = c->adj_scr;
    return LPM_OK;
}

// jv = dPin/dx . v at the current geometry and forces (both dim*n, device)
int force_jvp(lpm_ctx *c, const double *v, double *jv) {
    const int g = blocks_for(c->n, 256);
    double *dcs = nullptr;
    int rc = adj_scratch(c, &dcs);
    if (rc) return rc;
    double *dfb = dcs + 3 * std::max<int64_t>(c->nbond, 1);
    hipLaunchKernelGGL(k_jvp_bond, dim3(g), dim3(256), 0, c->stream, c->n, c->dim, c->bptr, c->bsplit, c->p2, c->dis, c->cs,
                       c->Kn, c->Tv, c->bdamage, v, dcs, dfb);
    rc = check(c, "k_jvp_bond");
    if (!rc) {
        hipLaunchKernelGGL(k_jvp_particle, dim3(g), dim3(256), 0, c->stream, c->n, c->dim, c->bptr, c->opp, c->cs, c->bforce,
                           dcs, dfb, jv);
        rc = check(c, "k_jvp_particle");
    }
    return rc;
}

// y = (dPin/dx)^T w at the current geometry and forces (both dim*n, device)
int force_vjp(lpm_ctx *c, const double *w, double *y) {
    const int g = blocks_for(c->n, 256);
    double *cb = nullptr;
    int rc = adj_scratch(c, &cb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vjp_bond, dim3(g), dim3(256), 0, c->stream, c->n, c->dim, c->bptr, c->bsplit, c->p2, c->opp, c->cs,
                       c->Kn, c->Tv, c->bdamage, w, cb);
    rc = check(c, "k_vjp_bond");
    if (!rc) {
        hipLaunchKernelGGL(k_vjp_particle, dim3(g), dim3(256), 0, c->stream, c->n, c->dim, c->bptr, c->p2, c->opp, c->dis, c->cs,
                           c->bforce, c->bdamage, cb, w, y);
        rc = check(c, "k_vjp_particle");
    }
    return rc;
}

template <typename T, int D, int VP>
static int add_x_t(lpm_ctx *c, double *lam) {
    hipLaunchKernelGGL((k_add_x<T, D, VP>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, (const T *)c->v_x, lam);
    return check(c, "k_add_x");
}

static int add_x(lpm_ctx *c, double *lam) {
    if (c->precision == 0) return c->dim == 3 ? add_x_t<float, 3, 3>(c, lam) : add_x_t<float, 2, 2>(c, lam);
    return c->dim == 3 ? add_x_t<double, 3, 3>(c, lam) : add_x_t<double, 2, 2>(c, lam);
}

// Adjoint of the linear objective G = a.Pin + b.x at the current state (a, b: dim*n device
// arrays in DoF order; either may be null). On return gKn, gTv, gBd (host, nbond each, any may
// be null) hold dG/dKn, dG/dTv and dG/dbdamage, and G its value.
int adjoint_linear(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gKn, double *gTv, double *gBd, double *G,
                   int *refinements, double *rel_residual, double *gBd_dev) {
    const int64_t n = c->n, nd = n * c->dim;
    int rc = geometry_force_state(c);
    if (rc) return rc;
    // K with constrained DoFs eliminated: the preconditioner of the exact-Jacobian solve
    if ((rc = calc_stiffness(c))) return rc;
    if ((rc = update_stiffness_disp_bc(c))) return rc;
    double *buf = nullptr;  // e, g, lambda, r, jv
    LPM_HIP(c, hipMalloc(&buf, sizeof(double) * 5 * nd));
    double *e = buf, *g = buf + nd, *lam = buf + 2 * nd, *r = buf + 3 * nd, *jv = buf + 4 * nd;
    double *saved_res = c->residual;
    const double saved_tol = c->cg_tol;
    auto done = [&](int code) {
        c->residual = saved_res;
        c->cg_tol = saved_tol;
        if (hipStreamSynchronize(c->stream) != hipSuccess && !code) code = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
        hipFree(buf);
        return code;
    };
    const int gd = blocks_for(nd, 256);
    if (gd > c->partial_cap) return done(lpm_fail(c, LPM_E_LIMIT, "adjoint grid exceeds partial buffer"));
    if (hipMemsetAsync(buf, 0, sizeof(double) * 5 * nd, c->stream) != hipSuccess)
        return done(lpm_fail(c, LPM_E_HIP, "adjoint clear"));
    if (a_dev && hipMemcpyAsync(e, a_dev, sizeof(double) * nd, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return done(lpm_fail(c, LPM_E_HIP, "adjoint weights"));
    // g = (dG/dx) restricted to the free DoFs = (J^T a + b)_free
    if (a_dev && (rc = force_vjp(c, e, g))) return done(rc);
    if (b_dev) {
        hipLaunchKernelGGL(k_axpy1, dim3(gd), dim3(256), 0, c->stream, nd, b_dev, g);
        if ((rc = check(c, "k_axpy1"))) return done(rc);
    }
    hipLaunchKernelGGL(k_mask_free, dim3(gd), dim3(256), 0, c->stream, nd, c->dim, c->cons, c->damage_visual, g);
    if ((rc = check(c, "k_mask_free"))) return done(rc);
    // J_ff^T lambda = g by iterative refinement: r = g - J_ff^T lambda (exact, FP64), K_ff d = r (CG
    // on the assembled stiffness, which approximates J), lambda += d
    double g2 = -1.0, r2 = 0.0;
    int it = 0;
    const int max_ref = 40;
    c->cg_tol = std::max(saved_tol, c->precision == 0 ? 1e-6 : 1e-12);
    for (;; ++it) {
        if ((rc = force_vjp(c, lam, jv))) return done(rc);
        hipLaunchKernelGGL(k_mask_free, dim3(gd), dim3(256), 0, c->stream, nd, c->dim, c->cons, c->damage_visual, jv);
        hipLaunchKernelGGL(k_sub_norm, dim3(gd), dim3(256), 0, c->stream, nd, g, jv, r, c->partial, c->partial_cap);
        if ((rc = check(c, "k_sub_norm"))) return done(rc);
        if ((rc = launch_finalize(c, FIN_DIAG, gd))) return done(rc);
        LPM_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        r2 = c->h_scal[SC_DIAGNORM];
        if (g2 < 0) g2 = r2;
        if (g2 == 0.0 || r2 <= 1e-26 * g2 || it >= max_ref) break;
        c->residual = r;
        int cg_it = 0, ok = 0;
        double err = 0;
        rc = cg_solve(c, &cg_it, &err, &ok);
        c->residual = saved_res;
        if (rc) return done(rc);
        if ((rc = add_x(c, lam))) return done(rc);
        hipLaunchKernelGGL(k_mask_free, dim3(gd), dim3(256), 0, c->stream, nd, c->dim, c->cons, c->damage_visual, lam);
        if ((rc = check(c, "k_mask_free"))) return done(rc);
    }
    if (refinements) *refinements = it;
    if (rel_residual) *rel_residual = g2 > 0 ? sqrt(r2 / g2) : 0.0;
    // per-bond sensitivities, copied into the caller's host arrays
    const int64_t nb = std::max<int64_t>(c->nbond, 1);
    double *gb = nullptr;
    if (hipMalloc(&gb, sizeof(double) * 3 * nb) != hipSuccess) return done(lpm_fail(c, LPM_E_NOMEM, "adjoint output"));
    hipLaunchKernelGGL(k_param_grad, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->dim, c->bptr, c->bsplit, c->p2,
                       c->opp, c->cs, c->dLe, c->dLe_total, c->bdamage, c->Kn, c->Tv, e, lam, gb, gb + nb, gb + 2 * nb);
    rc = check(c, "k_param_grad");
    if (!rc && gBd_dev &&
        hipMemcpyAsync(gBd_dev, gb + 2 * nb, sizeof(double) * c->nbond, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        rc = lpm_fail(c, LPM_E_HIP, "adjoint copy");
    double *outs[3] = {gKn, gTv, gBd};
    for (int k = 0; k < 3 && !rc; ++k)
        if (outs[k] && hipMemcpyAsync(outs[k], gb + k * nb, sizeof(double) * c->nbond, hipMemcpyDeviceToHost, c->stream) !=
                           hipSuccess)
            rc = lpm_fail(c, LPM_E_HIP, "adjoint copy");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(gb);
    if (rc) return done(rc);
    if (G) {
        // G from the current Pin and positions, summed in DoF order
        std::vector<double> pin(3 * n), x(3 * n), wa(a_dev ? nd : 0), wb(b_dev ? nd : 0);
        LPM_HIP(c, hipMemcpyAsync(pin.data(), c->Pin, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipMemcpyAsync(x.data(), c->xyz, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        if (a_dev) LPM_HIP(c, hipMemcpyAsync(wa.data(), a_dev, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
        if (b_dev) LPM_HIP(c, hipMemcpyAsync(wb.data(), b_dev, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        double s = 0;
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < c->dim; ++k) {
                const int64_t d = (int64_t)c->dim * i + k;
                if (a_dev && wa[d] != 0.0) s += wa[d] * pin[3 * i + k];
                if (b_dev && wb[d] != 0.0) s += wb[d] * x[3 * i + k];
            }
        *G = s;
    }
    return done(LPM_OK);
}

// Adjoint of F = sum over the group's particles of Pin[axis] at the current state: the linear
// objective with a = the indicator of (group particle, axis), each particle counted once.
int adjoint_reaction(lpm_ctx *c, int gid, int axis, double *gKn, double *gTv, double *F, int *refinements,
                     double *rel_residual) {
    const int64_t nd = c->n * c->dim;
    const Group &grp = c->groups[gid];
    double *a = nullptr;
    LPM_HIP(c, hipMalloc(&a, sizeof(double) * std::max<int64_t>(nd, 1)));
    int rc = LPM_OK;
    if (hipMemsetAsync(a, 0, sizeof(double) * nd, c->stream) != hipSuccess) rc = lpm_fail(c, LPM_E_HIP, "adjoint clear");
    if (!rc && grp.n > 0) {
        hipLaunchKernelGGL(k_indicator, dim3(blocks_for(grp.n, 256)), dim3(256), 0, c->stream, grp.n, grp.ids, c->dim, axis, a);
        rc = check(c, "k_indicator");
    }
    if (!rc) rc = adjoint_linear(c, a, nullptr, gKn, gTv, nullptr, F, refinements, rel_residual);
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(a);
    return rc;
}

// dG/dA, dG/dB, dG/dd per particle (host arrays, n each, any may be null) for the linear objective
// G = a.Pin + b.x at the current equilibrium, through the last fatigue damage update
int adjoint_fatigue_params(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gA, double *gB, double *gd,
                           double *G, int *refinements, double *rel_residual) {
    if (c->model != 3) return lpm_fail(c, LPM_E_STATE, "fatigue-parameter sensitivity needs the HCF model");
    if (!c->fat_valid) return lpm_fail(c, LPM_E_STATE, "no fatigue damage update to differentiate");
    if (!c->nlptr) return lpm_fail(c, LPM_E_STATE, "no nonlocal neighbour lists");
    const int64_t n = c->n, nb = std::max<int64_t>(c->nbond, 1);
    double *buf = nullptr;  // gbd (nbond), gD (n), norm (n), out (3n)
    LPM_HIP(c, hipMalloc(&buf, sizeof(double) * (nb + 5 * std::max<int64_t>(n, 1))));
    double *gbd = buf, *gD = buf + nb, *norm = gD + n, *out = norm + n;
    int rc = adjoint_linear(c, a_dev, b_dev, nullptr, nullptr, nullptr, G, refinements, rel_residual, gbd);
    const int g = blocks_for(n, 256);
    if (!rc) {
        hipLaunchKernelGGL(k_fat_gd, dim3(g), dim3(256), 0, c->stream, n, c->fat_undamaged, c->type, c->bptr, c->p2, c->opp,
                           c->damage, c->bdamage, gbd, gD);
        rc = check(c, "k_fat_gd");
    }
    if (!rc) {
        hipLaunchKernelGGL(k_fat_norm, dim3(g), dim3(256), 0, c->stream, n, c->fat_undamaged, c->volume, c->nneighbors,
                           c->type, c->nlptr, c->nlidx, c->bptr, c->fat_xyz, c->nonlocal_L, norm);
        rc = check(c, "k_fat_norm");
    }
    if (!rc) {
        hipLaunchKernelGGL(k_fat_params, dim3(g), dim3(256), 0, c->stream, n, c->fat_undamaged, c->volume, c->nneighbors,
                           c->fat_dt, c->type, c->nlptr, c->nlidx, c->bptr, c->fat_xyz, c->nonlocal_L, norm, gD, c->mat,
                           c->sv, c->fat_range, out);
        rc = check(c, "k_fat_params");
    }
    double *outs[3] = {gA, gB, gd};
    for (int k = 0; k < 3 && !rc; ++k)
        if (outs[k] && hipMemcpyAsync(outs[k], out + k * n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = lpm_fail(c, LPM_E_HIP, "fatigue sensitivity copy");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(buf);
    return rc;
}
