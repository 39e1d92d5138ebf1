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


// ---------------------------------------------------------------- adjoint over a fatigue load path
// Recorded path (lpm_fatigue_record): every HCF state-variable update at a converged state and
// every fatigue damage update, in order. Going backwards, the adjoints held are those of the
// current epoch's bond damage (mu_bd), of the particle damage (mu_D) and of the three energy
// trackers s0 = max over the history of s2, s1 / s2 = min / max of the cycle's energies
// (particle_fatigue_hcf.h:127-145): mu_s[3i + 0..2].

// s0, s1, s2 of every particle (3 per particle)
__global__ __launch_bounds__(256) void k_sv_snap(int64_t n, const double *__restrict__ sv, double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) out[3 * i + k] = sv[LPM_NSV * i + k];
}

// Back through one state-variable update: s2' = max(E, s2), s1' = min(E, s1), s0' = max(s2', s0).
// A tracker's adjoint goes to E where the update changed it, else to its value before; the
// adjoint of E is written to muE.
__global__ __launch_bounds__(256) void k_fat_track_rev(int64_t n, const double *__restrict__ sb, const double *__restrict__ sa,
                                                        double *__restrict__ mu_s, double *__restrict__ muE) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s0b = sb[3 * i], s1b = sb[3 * i + 1], s2b = sb[3 * i + 2];
    const double s0a = sa[3 * i], s1a = sa[3 * i + 1], s2a = sa[3 * i + 2];
    double m0 = mu_s[3 * i], m1 = mu_s[3 * i + 1], m2 = mu_s[3 * i + 2];
    if (s0a != s0b) {  // s0' = s2'
        m2 += m0;
        m0 = 0.0;
    }
    double mE = 0.0;
    if (s2a != s2b) {
        mE += m2;
        m2 = 0.0;
    }
    if (s1a != s1b) {
        mE += m1;
        m1 = 0.0;
    }
    mu_s[3 * i] = m0;
    mu_s[3 * i + 1] = m1;
    mu_s[3 * i + 2] = m2;
    muE[i] = mE;
}

// Objective sum_i muE_i E_i with the HCF strain energy (particle_fatigue_hcf.h:82-125)
//   E_i = sum_l [ sum_{b in l} Kn_b dLe_b^2 / 2 + TdLe_total(i,l) dLe_total(i,l) / 2 ] / V_m,i:
// first the per-bond coefficient c_b = muE_i dE_i/d dLe_b ...
__global__ __launch_bounds__(256) void k_energy_coef(int64_t n, double volume, int nneighbors, const int32_t *__restrict__ bptr,
                                                      const int32_t *__restrict__ bsplit, const double *__restrict__ Kn,
                                                      const double *__restrict__ Tv, const double *__restrict__ dLe,
                                                      const double *__restrict__ dLe_total,
                                                      const double *__restrict__ TdLe_total, const double *__restrict__ muE,
                                                      double *__restrict__ coef) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    const double V = volume * (b1 - b0) / nneighbors;
    const double m = muE[i];
    for (int b = b0; b < b1; ++b) {
        const int l = b < bm ? 0 : 1;
        coef[b] = m == 0.0 ? 0.0 : m * (Kn[b] * dLe[b] + 0.5 * (Tv[b] * dLe_total[2 * i + l] + TdLe_total[2 * i + l])) / V;
    }
}

// ... then gx_p = sum over p's bonds b of cs_b (c_b + c_opp(b)): dLe_b and dLe_opp(b) both
// move with the bond length, whose gradient at p is cs_b (zero for a broken bond). Adds to gx.
__global__ __launch_bounds__(256) void k_energy_gx(int64_t n, int dim, const int32_t *__restrict__ bptr,
                                                    const int32_t *__restrict__ opp, const double *__restrict__ cs,
                                                    const double *__restrict__ coef, double *__restrict__ gx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double g[3] = {0.0, 0.0, 0.0};
    for (int b = bptr[p]; b < bptr[p + 1]; ++b) {
        const double w = coef[b] + coef[opp[b]];
        for (int k = 0; k < 3; ++k) g[k] += w * cs[3 * b + k];
    }
    for (int k = 0; k < dim; ++k) gx[(int64_t)dim * p + k] += g[k];
}

// Back through bd' = max(bd, max(D'_i, D'_j)) at an update: the adjoint of a bond whose damage
// the update raised goes to the particle whose D' won std::max (the bond's owner on a tie), the
// others stay with the bond damage before the update. Gathered per particle over its bonds and
// their reverses; the routed bonds are cleared by k_fat_route_clear afterwards.
__global__ __launch_bounds__(256) void k_fat_route(int64_t n, const int32_t *__restrict__ bptr, const int32_t *__restrict__ p2,
                                                    const int32_t *__restrict__ opp, const double *__restrict__ Dn,
                                                    const double *__restrict__ bd_prev, const double *__restrict__ bd_next,
                                                    const double *__restrict__ mu_bd, double *__restrict__ mu_D) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double Di = Dn[i];
    double s = 0.0;
    for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
        const double Dj = Dn[p2[b]];
        const int o = opp[b];
        if (bd_next[b] != bd_prev[b] && Di >= Dj) s += mu_bd[b];  // bond i->j: max(D_i, D_j), D_i on a tie
        if (bd_next[o] != bd_prev[o] && Di > Dj) s += mu_bd[o];   // bond j->i: max(D_j, D_i)
    }
    mu_D[i] += s;
}

__global__ __launch_bounds__(256) void k_fat_route_clear(int64_t nb, const double *__restrict__ bd_prev,
                                                          const double *__restrict__ bd_next, double *__restrict__ mu_bd) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb && bd_next[b] != bd_prev[b]) mu_bd[b] = 0.0;
}

// Back through D' = D + dt NL(Ddot) (clamped to 1 at the threshold): the adjoint of the nonlocal
// rate mu_nl = dt mu_D' where D' was neither clamped nor excluded; mu_D passes to D unclamped.
__global__ __launch_bounds__(256) void k_fat_update_rev(int64_t n, int undamaged, double dt, const int32_t *__restrict__ type,
                                                         const double *__restrict__ Dn, double *__restrict__ mu_D,
                                                         double *__restrict__ mu_nl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (type[i] == undamaged) {
        mu_nl[i] = 0.0;
        return;
    }
    const bool clamped = Dn[i] == 1.0;
    mu_nl[i] = clamped ? 0.0 : dt * mu_D[i];
    if (clamped) mu_D[i] = 0.0;
}

// HCF rate of particle j from its trackers after the cycle (s = s0, s1, s2)
__device__ __forceinline__ double hcf_rate(const double *m, const double *s) {
    const double r = s[2] - s[1];
    return r > 0 ? m[0] * pow(s[0], m[1]) * pow(r, m[2]) : 0.0;
}

// The update's nonlocal normaliser W_i and rate NL_i = sum_j phi_ij V_j Ddot_j / W_i
// (SolverFatigue::updateFatigueDamage, solver_fatigue.h:75-114) at the update's positions.
__global__ __launch_bounds__(256) void k_fat_nl(int64_t n, int undamaged, double volume, int nneighbors,
                                                 const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                 const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                 const double *__restrict__ xyz, const double *__restrict__ L,
                                                 const double *__restrict__ mat, const double *__restrict__ sa,
                                                 double *__restrict__ W, double *__restrict__ NL) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double A = 0.0, dn = 0.0;
    if (type[i] != undamaged) {
        for (int q = nlptr[i]; q < nlptr[i + 1]; ++q) {
            const int64_t j = nlidx[q];
            if (type[j] == undamaged) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double u = 1 / (1 + pow(sqrt(dx * dx + dy * dy + dz * dz) / L[i], 8.0));
            const double V = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
            dn += hcf_rate(&mat[LPM_NMAT * j], &sa[3 * j]) * u * u * V;
            A += u * u * V;
        }
    }
    W[i] = A;
    NL[i] = A > 0.0 ? dn / A : 0.0;
}

// Back through the rate law and the nonlocal weights' values: per particle j,
// S_j = sum_{i: j in NL(i)} mu_nl_i phi_ij V_j / W_i (gathered over NL(j), the lists being
// symmetric); then dG/dA_j += S_j s0^B r^d, dG/dB_j += S_j Ddot_j ln s0, dG/dd_j += S_j Ddot_j ln r,
// and the trackers' adjoints (the cycle's s1, s2 restart from these; s0 accumulates):
// mu_s0 += S_j B Ddot_j / s0, mu_s2 = +S_j d Ddot_j / r, mu_s1 = -S_j d Ddot_j / r.
__global__ __launch_bounds__(256) void k_fat_rate_rev(int64_t n, int undamaged, double volume, int nneighbors,
                                                       const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                       const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                       const double *__restrict__ xyz, const double *__restrict__ L,
                                                       const double *__restrict__ W, const double *__restrict__ mu_nl,
                                                       const double *__restrict__ mat, const double *__restrict__ sa,
                                                       double *__restrict__ gth, double *__restrict__ mu_s) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double m1 = 0.0, m2 = 0.0, m0 = mu_s[3 * j];
    if (type[j] != undamaged) {
        const double Vj = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
        double S = 0.0;
        for (int q = nlptr[j]; q < nlptr[j + 1]; ++q) {
            const int64_t i = nlidx[q];
            if (type[i] == undamaged || W[i] == 0.0 || mu_nl[i] == 0.0) continue;
            const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double u = 1 / (1 + pow(sqrt(dx * dx + dy * dy + dz * dz) / L[i], 8.0));
            S += mu_nl[i] * u * u * Vj / W[i];
        }
        const double *m = &mat[LPM_NMAT * j];
        const double s0 = sa[3 * j], r = sa[3 * j + 2] - sa[3 * j + 1];
        if (r > 0 && S != 0.0) {
            const double base = pow(s0, m[1]) * pow(r, m[2]), rate = m[0] * base;
            gth[j] += S * base;
            if (s0 > 0) {
                gth[n + j] += S * rate * log(s0);
                m0 += S * m[1] * rate / s0;
            }
            gth[2 * n + j] += S * rate * log(r);
            m2 = S * m[2] * rate / r;
            m1 = -m2;
        }
    }
    mu_s[3 * j] = m0;
    mu_s[3 * j + 1] = m1;
    mu_s[3 * j + 2] = m2;
}

// Back through the nonlocal weights' positions: the gradient of sum_i mu_nl_i NL_i with respect
// to the update's positions, phi(r) = (1 + (r/L)^8)^-2, dphi/dr = -16 u^3 (r/L)^7 / L:
//   gx_p = sum_{j in NL(p)} [c(p,j) + c(j,p)] (x_p - x_j) / r_pj,
//   c(i,j) = mu_nl_i V_j (Ddot_j - NL_i) / W_i dphi/dr(r_ij, L_i). Adds to gx.
__global__ __launch_bounds__(256) void k_fat_weights_gx(int64_t n, int dim, int undamaged, double volume, int nneighbors,
                                                         const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                         const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                         const double *__restrict__ xyz, const double *__restrict__ L,
                                                         const double *__restrict__ W, const double *__restrict__ NL,
                                                         const double *__restrict__ mu_nl, const double *__restrict__ mat,
                                                         const double *__restrict__ sa, double *__restrict__ gx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || type[p] == undamaged) return;
    const double Vp = volume * (bptr[p + 1] - bptr[p]) / nneighbors;
    const double Rp = hcf_rate(&mat[LPM_NMAT * p], &sa[3 * p]);
    double g[3] = {0.0, 0.0, 0.0};
    for (int q = nlptr[p]; q < nlptr[p + 1]; ++q) {
        const int64_t j = nlidx[q];
        if (j == p || type[j] == undamaged) continue;
        const double e[3] = {xyz[3 * p] - xyz[3 * j], xyz[3 * p + 1] - xyz[3 * j + 1], xyz[3 * p + 2] - xyz[3 * j + 2]};
        const double r = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        if (r == 0.0) continue;
        double c = 0.0;
        if (W[p] > 0.0 && mu_nl[p] != 0.0) {
            const double t = r / L[p], u = 1 / (1 + pow(t, 8.0));
            const double dphi = -16.0 * u * u * u * pow(t, 7.0) / L[p];
            const double Vj = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
            c += mu_nl[p] * Vj * (hcf_rate(&mat[LPM_NMAT * j], &sa[3 * j]) - NL[p]) / W[p] * dphi;
        }
        if (W[j] > 0.0 && mu_nl[j] != 0.0) {
            const double t = r / L[j], u = 1 / (1 + pow(t, 8.0));
            const double dphi = -16.0 * u * u * u * pow(t, 7.0) / L[j];
            c += mu_nl[j] * Vp * (Rp - NL[j]) / W[j] * dphi;
        }
        for (int k = 0; k < 3; ++k) g[k] += c * e[k] / r;
    }
    for (int k = 0; k < dim; ++k) gx[(int64_t)dim * p + k] += g[k];
}

__global__ __launch_bounds__(256) void k_add(int64_t m, const double *__restrict__ a, double *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) b[i] += a[i];
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

// ---------------------------------------------------------------- fatigue path recording
void fat_record_clear(lpm_ctx *c) {
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &e : c->fat_events) {
        if (e.ckpt) hipFree(e.ckpt);
        if (e.data) hipFree(e.data);
    }
    c->fat_events.clear();
}

// A state-variable update at a converged state: before it (after = false) the state and the
// trackers, after it the trackers again.
int fat_record_state(lpm_ctx *c, bool after) {
    const int64_t n = c->n;
    const int g = blocks_for(n, 256);
    if (!after) {
        FatEvent e;
        e.kind = FatEvent::STATE;
        c->fat_events.push_back(e);
        FatEvent &b = c->fat_events.back();
        LPM_HIP(c, hipMalloc(&b.data, sizeof(double) * 6 * std::max<int64_t>(n, 1)));
        int rc = ckpt_save_buf(c, &b.ckpt);
        if (rc) return rc;
        hipLaunchKernelGGL(k_sv_snap, dim3(g), dim3(256), 0, c->stream, n, (const double *)c->sv, b.data);
        return check(c, "k_sv_snap");
    }
    if (c->fat_events.empty() || c->fat_events.back().kind != FatEvent::STATE)
        return lpm_fail(c, LPM_E_STATE, "fatigue path record out of order");
    hipLaunchKernelGGL(k_sv_snap, dim3(g), dim3(256), 0, c->stream, n, (const double *)c->sv, c->fat_events.back().data + 3 * n);
    return check(c, "k_sv_snap");
}

// A fatigue damage update: its step, excluded type and the damage it left.
int fat_record_update(lpm_ctx *c, double dt, int undamaged) {
    FatEvent e;
    e.kind = FatEvent::UPDATE;
    e.dt = dt;
    e.undamaged = undamaged;
    c->fat_events.push_back(e);
    FatEvent &b = c->fat_events.back();
    LPM_HIP(c, hipMalloc(&b.data, sizeof(double) * std::max<int64_t>(c->n, 1)));
    LPM_HIP(c, hipMemcpyAsync(b.data, c->damage, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
    return LPM_OK;
}

// dG/dA, dG/dB, dG/dd per particle (host arrays, n each, any may be null) for G = a.Pin + b.x at
// the current equilibrium, through every recorded state and fatigue update: one transposed solve
// at the current state, then, going back over the record, one per recorded state whose energy
// reaches a later update (through the trackers), each with the objective sum_i muE_i E_i plus,
// at the state an update was taken in, the nonlocal weights' dependence on its positions.
int adjoint_fatigue_path(lpm_ctx *c, const double *a_dev, const double *b_dev, double *gA, double *gB, double *gd,
                         double *G, int *solves, double *max_rr) {
    if (c->model != 3) return lpm_fail(c, LPM_E_STATE, "fatigue-path sensitivity needs the HCF model");
    if (!c->nlptr) return lpm_fail(c, LPM_E_STATE, "no nonlocal neighbour lists");
    auto &ev = c->fat_events;
    bool any_update = false;
    for (size_t k = 0; k < ev.size(); ++k)
        if (ev[k].kind == FatEvent::UPDATE) {
            any_update = true;
            if (k == 0 || ev[k - 1].kind != FatEvent::STATE)
                return lpm_fail(c, LPM_E_STATE, "a recorded fatigue update must follow a recorded state update");
        }
    if (!any_update) return lpm_fail(c, LPM_E_STATE, "no recorded fatigue damage update");
    const int64_t n = c->n, nb = std::max<int64_t>(c->nbond, 1), nd = n * c->dim;
    const int gn = blocks_for(n, 256), gb = blocks_for(nb, 256), gdx = blocks_for(nd, 256);
    // mu_bd, bd_next, coef (nb each); mu_D, muE, mu_nl, W, NL (n); mu_s, gth (3n); gx, gxw (nd)
    double *buf = nullptr;
    const size_t tot = 3 * (size_t)nb + 5 * (size_t)n + 6 * (size_t)n + 2 * (size_t)nd;
    LPM_HIP(c, hipMalloc(&buf, sizeof(double) * std::max<size_t>(tot, 1)));
    double *mu_bd = buf, *bd_next = mu_bd + nb, *coef = bd_next + nb, *mu_D = coef + nb, *muE = mu_D + n;
    double *mu_nl = muE + n, *W = mu_nl + n, *NL = W + n, *mu_s = NL + n, *gth = mu_s + 3 * n, *gx = gth + 3 * n;
    double *gxw = gx + nd;
    void *final_state = nullptr;
    auto done = [&](int code) {
        if (final_state) {
            if (!code) code = ckpt_restore_buf(c, final_state);
            hipFree(final_state);
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess && !code) code = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
        hipFree(buf);
        return code;
    };
    if (hipMemsetAsync(buf, 0, sizeof(double) * tot, c->stream) != hipSuccess) return done(lpm_fail(c, LPM_E_HIP, "clear"));
    int rc = ckpt_save_buf(c, &final_state);
    if (rc) return done(rc);
    int nsolve = 0, it = 0;
    double rr = 0.0, rmax = 0.0;
    // the objective at the current state
    rc = adjoint_linear(c, a_dev, b_dev, nullptr, nullptr, nullptr, G, &it, &rr, mu_bd);
    if (rc) return done(rc);
    ++nsolve;
    rmax = rr;
    if (hipMemcpyAsync(bd_next, c->bdamage, sizeof(double) * c->nbond, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return done(lpm_fail(c, LPM_E_HIP, "copy"));
    bool pending_w = false;    // gxw holds the weights' term for the next state back
    int64_t restored = -1;     // event whose state the context holds
    for (int64_t k = (int64_t)ev.size() - 1; k >= 0; --k) {
        const FatEvent &e = ev[k];
        if (e.kind == FatEvent::UPDATE) {
            // the state the update was taken in: its bond damage, positions and trackers
            const FatEvent &s = ev[k - 1];
            if ((rc = ckpt_restore_buf(c, s.ckpt))) return done(rc);
            restored = k - 1;
            const double *sa = s.data + 3 * n;
            hipLaunchKernelGGL(k_fat_route, dim3(gn), dim3(256), 0, c->stream, n, c->bptr, c->p2, c->opp, e.data,
                               (const double *)c->bdamage, (const double *)bd_next, (const double *)mu_bd, mu_D);
            hipLaunchKernelGGL(k_fat_route_clear, dim3(gb), dim3(256), 0, c->stream, c->nbond, (const double *)c->bdamage,
                               (const double *)bd_next, mu_bd);
            hipLaunchKernelGGL(k_fat_update_rev, dim3(gn), dim3(256), 0, c->stream, n, e.undamaged, e.dt, c->type, e.data, mu_D,
                               mu_nl);
            hipLaunchKernelGGL(k_fat_nl, dim3(gn), dim3(256), 0, c->stream, n, e.undamaged, c->volume, c->nneighbors, c->type,
                               c->nlptr, c->nlidx, c->bptr, (const double *)c->xyz, (const double *)c->nonlocal_L,
                               (const double *)c->mat, sa, W, NL);
            hipLaunchKernelGGL(k_fat_rate_rev, dim3(gn), dim3(256), 0, c->stream, n, e.undamaged, c->volume, c->nneighbors,
                               c->type, c->nlptr, c->nlidx, c->bptr, (const double *)c->xyz, (const double *)c->nonlocal_L,
                               (const double *)W, (const double *)mu_nl, (const double *)c->mat, sa, gth, mu_s);
            if (hipMemsetAsync(gxw, 0, sizeof(double) * nd, c->stream) != hipSuccess) return done(lpm_fail(c, LPM_E_HIP, "clear"));
            hipLaunchKernelGGL(k_fat_weights_gx, dim3(gn), dim3(256), 0, c->stream, n, c->dim, e.undamaged, c->volume,
                               c->nneighbors, c->type, c->nlptr, c->nlidx, c->bptr, (const double *)c->xyz,
                               (const double *)c->nonlocal_L, (const double *)W, (const double *)NL, (const double *)mu_nl,
                               (const double *)c->mat, sa, gxw);
            if ((rc = check(c, "fatigue update adjoint"))) return done(rc);
            pending_w = true;
            continue;
        }
        // a state-variable update: the trackers back to the state's energy
        if (restored != k && (rc = ckpt_restore_buf(c, e.ckpt))) return done(rc);
        restored = k;
        hipLaunchKernelGGL(k_fat_track_rev, dim3(gn), dim3(256), 0, c->stream, n, (const double *)e.data,
                           (const double *)(e.data + 3 * n), mu_s, muE);
        if (hipMemsetAsync(gx, 0, sizeof(double) * nd, c->stream) != hipSuccess) return done(lpm_fail(c, LPM_E_HIP, "clear"));
        hipLaunchKernelGGL(k_energy_coef, dim3(gn), dim3(256), 0, c->stream, n, c->volume, c->nneighbors, c->bptr, c->bsplit,
                           (const double *)c->Kn, (const double *)c->Tv, (const double *)c->dLe, (const double *)c->dLe_total,
                           (const double *)c->TdLe_total, (const double *)muE, coef);
        hipLaunchKernelGGL(k_energy_gx, dim3(gn), dim3(256), 0, c->stream, n, c->dim, c->bptr, c->opp, (const double *)c->cs,
                           (const double *)coef, gx);
        if (pending_w) {
            hipLaunchKernelGGL(k_add, dim3(gdx), dim3(256), 0, c->stream, nd, (const double *)gxw, gx);
            pending_w = false;
        }
        if ((rc = check(c, "state adjoint objective"))) return done(rc);
        // skip the solve when the objective has no weight here
        double g2 = 0.0;
        {
            std::vector<double> h(nd);
            LPM_HIP(c, hipMemcpyAsync(h.data(), gx, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
            LPM_HIP(c, hipStreamSynchronize(c->stream));
            for (double v : h) g2 += v * v;
        }
        if (g2 > 0.0) {
            double *dbd = coef;  // coef is free again
            if ((rc = adjoint_linear(c, nullptr, gx, nullptr, nullptr, nullptr, nullptr, &it, &rr, dbd))) return done(rc);
            ++nsolve;
            rmax = std::max(rmax, rr);
            hipLaunchKernelGGL(k_add, dim3(gb), dim3(256), 0, c->stream, c->nbond, (const double *)dbd, mu_bd);
            if ((rc = check(c, "k_add"))) return done(rc);
        }
        if (hipMemcpyAsync(bd_next, c->bdamage, sizeof(double) * c->nbond, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return done(lpm_fail(c, LPM_E_HIP, "copy"));
    }
    double *outs[3] = {gA, gB, gd};
    for (int k = 0; k < 3 && !rc; ++k)
        if (outs[k] && hipMemcpyAsync(outs[k], gth + k * n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = lpm_fail(c, LPM_E_HIP, "fatigue-path sensitivity copy");
    if (solves) *solves = nsolve;
    if (max_rr) *max_rr = rmax;
    return done(rc);
}
