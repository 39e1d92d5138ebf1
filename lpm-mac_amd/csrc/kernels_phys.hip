// Per-bond / per-particle physics kernels of the Newton load step (FP64) and the
// block-CSR stiffness assembly. This translation unit is compiled with
// -ffp-contract=off: every kernel performs the reference's floating-point
// operations in the reference's order, so FP64 results are reproducible
// against a CPU restatement.
//
// Data layout in HBM (all structure-of-arrays):
//   particles  xyz[3i+k], Pin[3i+k], Pex[3i+k], stress[6i+k], state_var[7i+k], mat[10i+k],
//              dLe_total/TdLe_total[2i+layer], cs_sum[6i+3*layer+k]
//   bonds      CSR by particle: bonds of i are [bptr[i], bptr[i+1]), layer-0 bonds first,
//              layer 1 from bsplit[i]; per bond p2, opp (reverse bond), dis0, dis, dLe, dLp,
//              bstrain, Kn, Tv, cs[3b+k], bforce, bdamage
//   pattern    conns CSR (cptr, cidx) = block rows of the stiffness matrix
//   matrix     block-sparse with (dim x dim) blocks on the conns pattern, stored SELL-64
//              (slices of 64 rows, lane-major; see sell_off in reduce.h)
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <vector>

#include "lpm_internal.h"
#include "reduce.h"

namespace {

constexpr int ELASTIC = 0, ELASTIC_DAMAGE = 1, J2 = 2, HCF = 3;

__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min

// ---------------------------------------------------------------- geometry + bond force
// Particle::updateBondsGeometry + Bond::updatebGeometry (particle.h:163-181, bond.h:30-39),
// then updateBondsForce of the model (particle_elastic.h:38-50, particle_elastic_damage.h:74-87,
// particle_j2plasticity.h:63-76, particle_fatigue_hcf.h:65-80).
// A workgroup of one wavefront owns GP consecutive particles and their contiguous bond range.
// Per-bond work runs lane-per-bond (coalesced loads and stores of the bond arrays); the
// per-particle sums then run thread-per-particle over the bond values staged in LDS, in bond
// order, so every sum is formed in the reference's order.
constexpr int GP = LPM_BOND_GP;  // particles per workgroup of the bond kernels

__global__ __launch_bounds__(64) void k_geometry_force(
    int64_t n, int model, int do_force, const int32_t *__restrict__ bptr, const int32_t *__restrict__ bsplit,
    const int32_t *__restrict__ p2, const double *__restrict__ xyz, const double *__restrict__ dis0,
    const double *__restrict__ dLp, const double *__restrict__ Kn, const double *__restrict__ Tv,
    const double *__restrict__ damage, double *__restrict__ dis, double *__restrict__ dLe,
    double *__restrict__ bstrain, double *__restrict__ cs, double *__restrict__ dLe_total,
    double *__restrict__ TdLe_total, double *__restrict__ cs_sum, double *__restrict__ bforce,
    double *__restrict__ bdamage) {
    extern __shared__ double lds[];
    __shared__ int sbp[GP + 1], ssp[GP];
    __shared__ double stot[GP][2];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * GP;
    const int np = (int)((n - p0) < GP ? (n - p0) : GP);
    if (tid <= np) sbp[tid] = bptr[p0 + tid];
    if (tid < np) ssp[tid] = bsplit[p0 + tid];
    __syncthreads();
    const int B0 = sbp[0], nbk = sbp[np] - B0;
    double *sle = lds, *stle = sle + nbk, *scx = stle + nbk, *scy = scx + nbk, *scz = scy + nbk;
    int *own = (int *)(scz + nbk);
    if (tid < np)
        for (int b = sbp[tid]; b < sbp[tid + 1]; ++b) own[b - B0] = tid;
    __syncthreads();
    for (int t = tid; t < nbk; t += blockDim.x) {
        const int b = B0 + t;
        const int64_t i = p0 + own[t];
        const int64_t j = p2[b];
        const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1], dz = xyz[3 * i + 2] - xyz[3 * j + 2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const bool broken = fabs(bdamage[b] - 1.0) < LPM_EPS;
        const double d0 = dis0[b];
        const double dl = d - d0;
        const double bs = broken ? 0.0 : (dl / d);
        const double le = broken ? 0.0 : (d - d0 - dLp[b]);
        const double cx = broken ? 0.0 : (dx / d), cy = broken ? 0.0 : (dy / d), cz = broken ? 0.0 : (dz / d);
        dis[b] = d;
        bstrain[b] = bs;
        dLe[b] = le;
        cs[3 * b] = cx;
        cs[3 * b + 1] = cy;
        cs[3 * b + 2] = cz;
        sle[t] = le;
        stle[t] = Tv[b] * le;
        scx[t] = cx;
        scy[t] = cy;
        scz[t] = cz;
    }
    __syncthreads();
    if (tid < np) {
        const int64_t i = p0 + tid;
        for (int l = 0; l < 2; ++l) {
            const int lo = (l == 0 ? sbp[tid] : ssp[tid]) - B0, hi = (l == 0 ? ssp[tid] : sbp[tid + 1]) - B0;
            double dt = 0, tdt = 0, sx = 0, sy = 0, sz = 0;
            for (int t = lo; t < hi; ++t) {
                dt += sle[t];
                tdt += stle[t];
                sx += scx[t];
                sy += scy[t];
                sz += scz[t];
            }
            dLe_total[2 * i + l] = dt;
            TdLe_total[2 * i + l] = tdt;
            cs_sum[6 * i + 3 * l] = sx;
            cs_sum[6 * i + 3 * l + 1] = sy;
            cs_sum[6 * i + 3 * l + 2] = sz;
            stot[tid][l] = dt;
        }
    }
    __syncthreads();
    if (!do_force) return;
    for (int t = tid; t < nbk; t += blockDim.x) {
        const int b = B0 + t, p = own[t];
        double f = 2. * Kn[b] * sle[t] + 2. * Tv[b] * stot[p][b < ssp[p] ? 0 : 1];
        double bd = bdamage[b];
        if (model != ELASTIC) {
            bd = dmax(bd, dmax(damage[p0 + p], damage[p2[b]]));
            bdamage[b] = bd;
        }
        f *= (1.0 - bd);
        bforce[b] = f;
    }
}

// Bond forces only (geometry already current); used after a damage update.
__global__ __launch_bounds__(256) void k_bond_force(int64_t n, int model, const int32_t *__restrict__ bptr,
                                                     const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                     const double *__restrict__ Kn, const double *__restrict__ Tv,
                                                     const double *__restrict__ dLe, const double *__restrict__ dLe_total,
                                                     const double *__restrict__ damage, double *__restrict__ bforce,
                                                     double *__restrict__ bdamage) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    const double Di = damage[i];
    for (int b = b0; b < b1; ++b) {
        double f = 2. * Kn[b] * dLe[b] + 2. * Tv[b] * dLe_total[2 * i + (b < bm ? 0 : 1)];
        double bd = bdamage[b];
        if (model != ELASTIC) {
            bd = dmax(bd, dmax(Di, damage[p2[b]]));
            bdamage[b] = bd;
        }
        f *= (1.0 - bd);
        bforce[b] = f;
    }
}

// Particle::updateParticleForce / updateParticleStress / updateParticleDamageVisual
// (particle.h:150-161, 197-240); the reverse bond comes from the precomputed opp[]. Same
// workgroup layout as k_geometry_force: lane-per-bond loads into LDS, then per-particle sums in
// bond order.
__global__ __launch_bounds__(64) void k_particle_force(int64_t n, double volume, int nneighbors,
                                                        const int32_t *__restrict__ bptr, const int32_t *__restrict__ opp,
                                                        const double *__restrict__ cs, const double *__restrict__ dis,
                                                        const double *__restrict__ bforce, const double *__restrict__ bdamage,
                                                        double *__restrict__ Pin, double *__restrict__ stress,
                                                        double *__restrict__ damage_visual) {
    extern __shared__ double lds[];
    __shared__ int sbp[GP + 1];
    const int tid = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * GP;
    const int np = (int)((n - p0) < GP ? (n - p0) : GP);
    if (tid <= np) sbp[tid] = bptr[p0 + tid];
    __syncthreads();
    const int B0 = sbp[0], nbk = sbp[np] - B0;
    double *sff = lds, *scx = sff + nbk, *scy = scx + nbk, *scz = scy + nbk, *sdis = scz + nbk, *sbd = sdis + nbk;
    for (int t = tid; t < nbk; t += blockDim.x) {
        const int b = B0 + t;
        sff[t] = bforce[b] + bforce[opp[b]];
        scx[t] = cs[3 * b];
        scy[t] = cs[3 * b + 1];
        scz[t] = cs[3 * b + 2];
        sdis[t] = dis[b];
        sbd[t] = bdamage[b];
    }
    __syncthreads();
    if (tid >= np) return;
    const int64_t i = p0 + tid;
    const int lo = sbp[tid] - B0, hi = sbp[tid + 1] - B0;
    const int nb = hi - lo;
    const double V_m = volume * nb / nneighbors;
    double P0 = 0., P1 = 0., P2 = 0.;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, dv = 0;
    for (int t = lo; t < hi; ++t) {
        const double ff = sff[t];
        const double cx = scx[t], cy = scy[t], cz = scz[t];
        P0 += cx * 0.5 * ff;
        P1 += cy * 0.5 * ff;
        P2 += cz * 0.5 * ff;
        const double w = 0.5 / V_m * sdis[t] * 0.5 * ff;
        s0 += w * cx * cx;
        s1 += w * cy * cy;
        s2 += w * cz * cz;
        s3 += w * cy * cz;
        s4 += w * cx * cz;
        s5 += w * cx * cy;
        dv += sbd[t];
    }
    Pin[3 * i] = P0;
    Pin[3 * i + 1] = P1;
    Pin[3 * i + 2] = P2;
    stress[6 * i] = s0;
    stress[6 * i + 1] = s1;
    stress[6 * i + 2] = s2;
    stress[6 * i + 3] = s3;
    stress[6 * i + 4] = s4;
    stress[6 * i + 5] = s5;
    damage_visual[i] = dv / (double)nb;
}

// Solver::updateRR (solver.h:108-128) with the squared norms of residual and reaction.
__global__ __launch_bounds__(256) void k_residual(int64_t n, int dim, const int32_t *__restrict__ cons,
                                                   const double *__restrict__ Pex, const double *__restrict__ Pin,
                                                   double *__restrict__ residual, double *partial, int cap) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (d < n * dim) {
        const int64_t i = d / dim;
        const int k = (int)(d - i * dim);
        const int c = cons[3 * i + k];
        const double r = (1 - c) * (Pex[3 * i + k] - Pin[3 * i + k]);
        residual[d] = r;
        v[0] = r * r;
        if (c == 1) v[1] = Pin[3 * i + k] * Pin[3 * i + k];
    }
    block_partial<2>(v, partial, cap);
}

// Solver::updateRR per specimen of an ensemble: the same residual and, per block of one
// specimen's DoFs, the partial squared norms of residual and reaction (blocks of specimen s
// are spec_blk[s] .. spec_blk[s+1]-1, so the host sums each specimen's partials in order).
__global__ __launch_bounds__(256) void k_residual_spec(int nspec, int dim, const int64_t *__restrict__ first,
                                                        const int32_t *__restrict__ spec_blk, const int32_t *__restrict__ cons,
                                                        const double *__restrict__ Pex, const double *__restrict__ Pin,
                                                        double *__restrict__ residual, double *partial, int cap) {
    int lo = 0, hi = nspec - 1;
    while (lo < hi) {  // specimen of this block
        const int mid = (lo + hi + 1) >> 1;
        if (spec_blk[mid] <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const int64_t d = first[lo] * dim + (int64_t)(blockIdx.x - spec_blk[lo]) * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (d < first[lo + 1] * dim) {
        const int64_t i = d / dim;
        const int k = (int)(d - i * dim);
        const int c = cons[3 * i + k];
        const double r = (1 - c) * (Pex[3 * i + k] - Pin[3 * i + k]);
        residual[d] = r;
        v[0] = r * r;
        if (c == 1) v[1] = Pin[3 * i + k] * Pin[3 * i + k];
    }
    block_partial<2>(v, partial, cap);
}

__global__ __launch_bounds__(256) void k_spec_scale(int64_t nd, int dim, const int32_t *__restrict__ spec_of,
                                                     const double *__restrict__ w, double *__restrict__ v) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d < nd) v[d] *= w[spec_of[d / dim]];
}

// ---------------------------------------------------------------- stiffness assembly
// Stiffness::calcStiffness with localStiffnessANA (stiffness.h:67-101, 228-263) and
// fdu2dxyz / fdu2dxyz1 / fdu2dxyz2 (derivative.h:7-239). One wavefront per particle row i,
// one lane per block (i, j in conns(i)).
// The wavefront first stages in LDS, with all lanes loading in parallel, what the row needs:
// the bonds of i (partner, cosines, Kn, Tv, bdamage), for each of them the same-layer bonds of
// the partner k (the 2-hop paths i -> k -> j2, in bond order), and, for each block of the row,
// the list of 2-hop paths that end on its column (bucketed by a counting sort and then sorted,
// so each list is in path order). Every lane then adds, in the reference's loop order, only
// the terms of its own block, so each FP64 block equals the reference's block bit for bit
// before it is rounded to the storage precision.
struct AsmBond {
    double c[3], tv, kn, bd;
    int p2, pad;
};
struct AsmHop {
    double c[3], tv, bd;
    int p2, owner;  // owner: index of the bond of i the path starts with
};

__host__ __device__ inline size_t asm_lds_bytes(int hops, int maxconn) {
    return sizeof(AsmBond) * LPM_MAXB + sizeof(int) * 2 * (LPM_MAXB + 1) + sizeof(AsmHop) * (size_t)hops +
           sizeof(int) * 2 * (size_t)hops + sizeof(int) * 4 * (size_t)(maxconn + 1);
}

// LDS of one row's 2-hop bookkeeping, carved from `base` (asm_lds_bytes(hops, maxconn) bytes)
struct AsmLds {
    AsmBond *B;   // bonds of i
    int *soff;    // 2-hop segment offsets of each bond of i (+ end)
    int *slo;     // first same-layer bond of the partner
    AsmHop *H;    // the 2-hop paths, in path order
    int *hslot;   // block slot each path ends on (-1: none)
    int *hlist;   // paths grouped by slot, in path order
    int *rcol;    // columns of row i
    int *lcnt;    // paths per slot
    int *lbeg;    // start of each slot's list
    int *lfil;    // fill cursor
};

__device__ __forceinline__ AsmLds asm_lds(char *base, int hops, int maxconn) {
    AsmLds s;
    s.B = (AsmBond *)base;
    s.soff = (int *)(s.B + LPM_MAXB);
    s.slo = s.soff + LPM_MAXB + 1;
    s.H = (AsmHop *)(s.slo + LPM_MAXB + 1);
    s.hslot = (int *)(s.H + hops);
    s.hlist = s.hslot + hops;
    s.rcol = s.hlist + hops;
    s.lcnt = s.rcol + (maxconn + 1);
    s.lbeg = s.lcnt + (maxconn + 1);
    s.lfil = s.lbeg + (maxconn + 1);
    return s;
}

// Stages row i's bonds, its 2-hop paths and its columns in LDS and groups the paths by the
// block they end on, each group in path order. Every thread of the workgroup calls it (it
// synchronises the workgroup); rows past n stage nothing.
__device__ void asm_row_lists(const AsmLds &s, int lane, int ib0, int nbi, int tm, int c0, int nci,
                              const int32_t *__restrict__ bptr, const int32_t *__restrict__ bsplit,
                              const int32_t *__restrict__ p2, const int32_t *__restrict__ cidx,
                              const double *__restrict__ cs, const double *__restrict__ Kn,
                              const double *__restrict__ Tv, const double *__restrict__ bdamage) {
    AsmBond *B = s.B;
    int *soff = s.soff, *slo = s.slo, *hslot = s.hslot, *hlist = s.hlist, *rcol = s.rcol;
    int *lcnt = s.lcnt, *lbeg = s.lbeg, *lfil = s.lfil;
    AsmHop *H = s.H;
    // bonds of i and the length of each bond's 2-hop list; the row's columns
    if (lane < nbi) {
        const int b = ib0 + lane;
        const int k = p2[b];
        AsmBond r;
        r.c[0] = cs[3 * b];
        r.c[1] = cs[3 * b + 1];
        r.c[2] = cs[3 * b + 2];
        r.tv = Tv[b];
        r.kn = Kn[b];
        r.bd = bdamage[b];
        r.p2 = k;
        r.pad = 0;
        B[lane] = r;
        const int klo = lane < tm ? bptr[k] : bsplit[k], khi = lane < tm ? bsplit[k] : bptr[k + 1];
        slo[lane] = klo;
        soff[lane + 1] = khi - klo;
    }
    for (int q = lane; q < nci; q += 64) {
        rcol[q] = cidx[c0 + q];
        lcnt[q] = 0;
        lfil[q] = 0;
    }
    __syncthreads();
    if (lane == 0) {
        soff[0] = 0;
        for (int t = 0; t < nbi; ++t) soff[t + 1] += soff[t];
    }
    __syncthreads();
    const int E = soff[nbi];
    // the paths, and the slot of the column each one ends on (binary search in the sorted row)
    for (int e = lane; e < E; e += 64) {
        int t = 0;
        while (soff[t + 1] <= e) ++t;
        const int b3 = slo[t] + (e - soff[t]);
        AsmHop h;
        h.c[0] = cs[3 * b3];
        h.c[1] = cs[3 * b3 + 1];
        h.c[2] = cs[3 * b3 + 2];
        h.tv = Tv[b3];
        h.bd = bdamage[b3];
        h.p2 = p2[b3];
        h.owner = t;
        H[e] = h;
        int lo = 0, hi = nci - 1, sl = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int v = rcol[mid];
            if (v == h.p2) { sl = mid; break; }
            if (v < h.p2) lo = mid + 1;
            else hi = mid - 1;
        }
        hslot[e] = sl;
        if (sl >= 0) atomicAdd(&lcnt[sl], 1);
    }
    __syncthreads();
    if (lane == 0) {
        int acc = 0;
        for (int q = 0; q < nci; ++q) {
            lbeg[q] = acc;
            acc += lcnt[q];
        }
    }
    __syncthreads();
    for (int e = lane; e < E; e += 64) {
        const int sl = hslot[e];
        if (sl >= 0) hlist[lbeg[sl] + atomicAdd(&lfil[sl], 1)] = e;
    }
    __syncthreads();
    for (int q = lane; q < nci; q += 64) {  // restore path order within each list
        int *L = hlist + lbeg[q];
        for (int a = 1; a < lcnt[q]; ++a) {
            const int v = L[a];
            int b = a - 1;
            while (b >= 0 && L[b] > v) {
                L[b + 1] = L[b];
                --b;
            }
            L[b + 1] = v;
        }
    }
    __syncthreads();
}

template <typename T, int D>
__global__ __launch_bounds__(256) void k_assemble(
    int64_t n, int hops, int maxconn, int diag10_literal, int transpose, const int32_t *__restrict__ bptr,
    const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2, const int32_t *__restrict__ cptr,
    const int32_t *__restrict__ cidx, const int32_t *__restrict__ tslot, const int64_t *__restrict__ sp,
    const double *__restrict__ cs, const double *__restrict__ cs_sum, const double *__restrict__ Kn,
    const double *__restrict__ Tv, const double *__restrict__ bdamage, T *__restrict__ aval, double *__restrict__ diag64) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const AsmLds s = asm_lds((char *)lds + (size_t)wid * asm_lds_bytes(hops, maxconn), hops, maxconn);
    const AsmBond *B = s.B;
    const int *soff = s.soff, *hlist = s.hlist, *rcol = s.rcol, *lcnt = s.lcnt, *lbeg = s.lbeg;
    const AsmHop *H = s.H;
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    const bool row = i < n;
    const int ib0 = row ? bptr[i] : 0, ibm = row ? bsplit[i] : 0, ib1 = row ? bptr[i + 1] : 0;
    const int nbi = ib1 - ib0, tm = ibm - ib0;
    const int c0 = row ? cptr[i] : 0, nci = row ? cptr[i + 1] - c0 : 0;
    asm_row_lists(s, lane, ib0, nbi, tm, c0, nci, bptr, bsplit, p2, cidx, cs, Kn, Tv, bdamage);
    if (!row) return;
    for (int slot = lane; slot < nci; slot += 64) {
        const int j = rcol[slot];
        double K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (j == i) {
            for (int l = 0; l < 2; ++l) {
                const int lo = l == 0 ? 0 : tm, hi = l == 0 ? tm : nbi;
                const double S0 = cs_sum[6 * i + 3 * l], S1 = cs_sum[6 * i + 3 * l + 1], S2 = cs_sum[6 * i + 3 * l + 2];
                for (int b = lo; b < hi; ++b) {
                    const double w = (1 - B[b].bd), kn = B[b].kn, tv = B[b].tv;
                    const double c0_ = B[b].c[0], c1 = B[b].c[1], c2 = B[b].c[2];
                    K[0][0] += w * ((kn * c0_ + tv * S0) * c0_ + (kn + tv) * (c0_ * c0_));
                    K[0][1] += w * ((kn * c1 + tv * S1) * c0_ + (kn + tv) * c1 * c0_);
                    K[1][1] += w * ((kn * c1 + tv * S1) * c1 + (kn + tv) * (c1 * c1));
                    if (D == 3) {
                        K[0][2] += w * ((kn * c2 + tv * S2) * c0_ + (kn + tv) * c2 * c0_);
                        K[1][2] += w * ((kn * c2 + tv * S2) * c1 + (kn + tv) * c2 * c1);
                        K[2][2] += w * ((kn * c2 + tv * S2) * c2 + (kn + tv) * (c2 * c2));
                    }
                }
            }
            if (!diag10_literal) K[1][0] = K[0][1];
            K[2][0] = K[0][2];
            K[2][1] = K[1][2];
            for (int r = 0; r < D; ++r) diag64[(int64_t)D * i + r] = K[r][r];
        } else {
            const int *L = hlist + lbeg[slot];
            const int nL = lcnt[slot];
            int q = 0;  // cursor in this block's path list (layer-0 paths come first)
            double Kl[2][3][3];
            bool flag[2];
            for (int l = 0; l < 2; ++l) {
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) Kl[l][r][c] = 0.0;
                const int lo = l == 0 ? 0 : tm, hi = l == 0 ? tm : nbi;
                const int eend = soff[hi];  // paths of this layer end here
                int bj = -1;  // first bond i->j of this layer
                for (int b = lo; b < hi; ++b)
                    if (bj < 0 && B[b].p2 == j) bj = b;
                bool touched = bj >= 0;
                if (bj >= 0) {  // derivative.h:53-119
                    const double n1[3] = {B[bj].c[0], B[bj].c[1], B[bj].c[2]};
                    const double wT1 = (1 - B[bj].bd) * B[bj].tv;
                    const double S[3] = {cs_sum[6 * j + 3 * l], cs_sum[6 * j + 3 * l + 1], cs_sum[6 * j + 3 * l + 2]};
                    for (int b2 = lo; b2 < hi; ++b2) {
                        const int k = B[b2].p2;
                        const double n2[3] = {B[b2].c[0], B[b2].c[1], B[b2].c[2]};
                        const double wT2 = (1 - B[b2].bd) * B[b2].tv;
                        if (k == j) {
                            const double w = (1 - B[b2].bd), kn = B[b2].kn, kt = B[b2].kn + B[b2].tv, tv = B[b2].tv;
                            for (int r = 0; r < D; ++r)
                                for (int c = 0; c < D; ++c) {
                                    const double pr = (r == c) ? kt * (n2[r] * n2[r])
                                                               : kt * n2[r > c ? r : c] * n2[r > c ? c : r];
                                    Kl[l][r][c] += w * ((-kn * n2[c] + tv * S[c]) * n2[r] - pr);
                                }
                            // a path through j itself cannot end on j (no self bonds)
                        } else {
                            const double a = 0.5 * (wT1 + wT2);
                            for (int r = 0; r < D; ++r)
                                for (int c = 0; c < D; ++c) Kl[l][r][c] -= a * n2[r] * n1[c];
                            const int e1 = soff[b2 + 1];
                            for (; q < nL && L[q] < e1; ++q) {
                                const int e = L[q];
                                const double n3[3] = {H[e].c[0], H[e].c[1], H[e].c[2]};
                                const double a3 = 0.5 * ((1 - H[e].bd) * H[e].tv + wT2);
                                for (int r = 0; r < D; ++r)
                                    for (int c = 0; c < D; ++c) Kl[l][r][c] -= a3 * n2[r] * n3[c];
                            }
                        }
                        // paths of b2 that the reference skips (k == j) are not in j's list
                        const int e1 = soff[b2 + 1];
                        while (q < nL && L[q] < e1) ++q;
                    }
                } else {  // derivative.h:122-139
                    for (; q < nL && L[q] < eend; ++q) {
                        const int e = L[q];
                        const int b1 = H[e].owner;
                        touched = true;
                        const double n1[3] = {B[b1].c[0], B[b1].c[1], B[b1].c[2]};
                        const double n2[3] = {H[e].c[0], H[e].c[1], H[e].c[2]};
                        const double a = 0.5 * ((1 - H[e].bd) * H[e].tv + (1 - B[b1].bd) * B[b1].tv);
                        for (int r = 0; r < D; ++r)
                            for (int c = 0; c < D; ++c) Kl[l][r][c] -= a * n1[r] * n2[c];
                    }
                }
                while (q < nL && L[q] < eend) ++q;
                flag[l] = touched;
            }
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c) {
                    if (flag[0] && flag[1]) K[r][c] = Kl[0][r][c] + Kl[1][r][c];
                    else if (flag[0]) K[r][c] = Kl[0][r][c];
                    else if (flag[1]) K[r][c] = Kl[1][r][c];
                }
        }
        if (!transpose) {
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c) aval[sell_off(sp, i, slot, D * D, r * D + c)] = (T)K[r][c];
        } else {
            const int ts = tslot[c0 + slot];
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c) aval[sell_off(sp, j, ts, D * D, r * D + c)] = (T)K[c][r];
        }
    }
}

// ---- planned assembly
// The row lists above depend on the topology only. k_asm_plan records them once per pattern:
// per row, its blocks ordered direct (a bond i->j in some layer) / 2-hop only / diagonal, and
// for each non-diagonal block its 2-hop paths in path order as (bond of the partner, bond of i
// it starts with). k_assemble_plan then computes one row of one block per lane, so the lanes of
// a loop take the same branch, and sums every entry in the same order as k_assemble.
//   pl_row[2i]   = direct blocks | 2-hop blocks << 16,  pl_row[2i+1] = paths of the row
//   pl_slot[c]   = slot | (first bond to j in layer 0, +1) << 16 | (layer 1, +1) << 24
//   pl_rng[c]    = first path of the block (within the row) | its path count << 16
//   pl_b3/pl_own[i*hops + e] = path e's bond of the partner (global) / its bond of i (row rank)
// with c = cptr[i] + the block's rank in that order.
__host__ __device__ inline size_t asm_plan_build_lds_bytes(int hops, int maxconn) {  // 16-byte multiple
    return (asm_lds_bytes(hops, maxconn) + sizeof(int) * 3 * (size_t)(maxconn + 1) + 15) & ~(size_t)15;
}

__global__ __launch_bounds__(256) void k_asm_plan(int64_t n, int hops, int maxconn, const int32_t *__restrict__ bptr,
                                                   const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                   const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                                   const double *__restrict__ cs, const double *__restrict__ Kn,
                                                   const double *__restrict__ Tv, const double *__restrict__ bdamage,
                                                   int32_t *__restrict__ pl_row, int32_t *__restrict__ pl_slot,
                                                   int32_t *__restrict__ pl_rng, int32_t *__restrict__ pl_b3,
                                                   uint8_t *__restrict__ pl_own, int32_t *dup) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    char *base = (char *)lds + (size_t)wid * asm_plan_build_lds_bytes(hops, maxconn);
    const AsmLds s = asm_lds(base, hops, maxconn);
    int *bjp = (int *)(base + asm_lds_bytes(hops, maxconn));
    int *rank = bjp + (maxconn + 1);
    int *ibeg = rank + (maxconn + 1);
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    const bool row = i < n;
    const int ib0 = row ? bptr[i] : 0, ibm = row ? bsplit[i] : 0, ib1 = row ? bptr[i + 1] : 0;
    const int nbi = ib1 - ib0, tm = ibm - ib0;
    const int c0 = row ? cptr[i] : 0, nci = row ? cptr[i + 1] - c0 : 0;
    asm_row_lists(s, lane, ib0, nbi, tm, c0, nci, bptr, bsplit, p2, cidx, cs, Kn, Tv, bdamage);
    // the planned kernel adds the k == j term once per layer: no partner may repeat in a layer
    if (lane < nbi) {
        const int lo = lane < tm ? 0 : tm;
        for (int b = lo; b < lane; ++b)
            if (s.B[b].p2 == s.B[lane].p2) atomicOr(dup, 1);
    }
    // per block: the first bond i->j of each layer (-1 for the diagonal block)
    for (int q = lane; q < nci; q += 64) {
        const int j = s.rcol[q];
        int b0 = -1, b1 = -1;
        for (int b = 0; b < tm; ++b)
            if (b0 < 0 && s.B[b].p2 == j) b0 = b;
        for (int b = tm; b < nbi; ++b)
            if (b1 < 0 && s.B[b].p2 == j) b1 = b;
        bjp[q] = (j == i) ? -1 : ((b0 + 1) | ((b1 + 1) << 8));
    }
    __syncthreads();
    if (row && lane == 0) {
        int k = 0, off = 0, nd = 0, n2 = 0;
        for (int cls = 0; cls < 3; ++cls)
            for (int q = 0; q < nci; ++q) {
                const int v = bjp[q];
                const int c = v < 0 ? 2 : (v != 0 ? 0 : 1);
                if (c != cls) continue;
                rank[q] = k++;
                ibeg[q] = off;
                if (c != 2) off += s.lcnt[q];
                nd += c == 0;
                n2 += c == 1;
            }
        pl_row[2 * i] = nd | (n2 << 16);
        pl_row[2 * i + 1] = off;
    }
    __syncthreads();
    if (!row) return;
    for (int q = lane; q < nci; q += 64) {
        const int v = bjp[q];
        const int64_t p = c0 + rank[q];
        const int cnt = v < 0 ? 0 : s.lcnt[q];
        pl_slot[p] = q | (v < 0 ? 0 : (v << 16));
        pl_rng[p] = ibeg[q] | (cnt << 16);
        for (int m = 0; m < cnt; ++m) {
            const int e = s.hlist[s.lbeg[q] + m];
            const int t = s.H[e].owner;
            const int64_t o = i * hops + ibeg[q] + m;
            pl_b3[o] = s.slo[t] + (e - s.soff[t]);
            pl_own[o] = (uint8_t)t;
        }
    }
}

// A 2-hop path as the planned assembly stages it: cosines of the partner's bond and its
// (1 - bdamage) * Tv; the bond of i the path starts with is kept in a byte array beside.
struct AsmPath {
    double c[3], wt;
};

// LDS of one row of the planned assembly: the row's bonds (maxb of them), its paths, their owners
__host__ __device__ inline size_t asm_plan_lds_bytes(int hops, int maxb) {
    const size_t h = hops > 0 ? hops : 1;
    return (sizeof(AsmBond) * (size_t)maxb + sizeof(AsmPath) * h + h + 15) & ~(size_t)15;
}

constexpr int ASM_NO_PATH = 0x7fffffff;

// Row r of the off-diagonal block of row i in column j (the else-branch of k_assemble, same
// operations in the same order). DIRECT: some layer has a bond i->j (bjl[l] >= 0). `own` is the
// owner of path q (ASM_NO_PATH past q1), read one path ahead of its use.
template <int D, bool DIRECT>
__device__ __forceinline__ void asm_plan_row(const AsmBond *__restrict__ B, const AsmPath *__restrict__ H,
                                             const uint8_t *__restrict__ Ho, int tm, int nbi, int j, int r,
                                             const int (&bjl)[2], int q, int q1, const double *__restrict__ cs_sum,
                                             double (&K)[D]) {
    double Kl[2][D];
    bool flag[2];
    int own = q < q1 ? Ho[q] : ASM_NO_PATH;
    for (int l = 0; l < 2; ++l) {
        for (int c = 0; c < D; ++c) Kl[l][c] = 0.0;
        const int lo = l == 0 ? 0 : tm, hi = l == 0 ? tm : nbi;
        bool touched = false;
        if (DIRECT && bjl[l] >= 0) {  // derivative.h:53-119
            touched = true;
            const int bj = bjl[l];
            const double n1[3] = {B[bj].c[0], B[bj].c[1], B[bj].c[2]};
            const double wT1 = (1 - B[bj].bd) * B[bj].tv;
            const double S[3] = {cs_sum[6 * j + 3 * l], cs_sum[6 * j + 3 * l + 1], cs_sum[6 * j + 3 * l + 2]};
            // the term of the bond i->j itself (k == j), added when the loop reaches b2 = bj: the
            // layer has no other bond to j (k_asm_plan checks that), and no 2-hop path through j
            // ends on j
            double tA[D];
            {
                const double w = (1 - B[bj].bd), kn = B[bj].kn, kt = B[bj].kn + B[bj].tv, tv = B[bj].tv;
                for (int c = 0; c < D; ++c) {
                    const double pr = (r == c) ? kt * (n1[r] * n1[r]) : kt * n1[r > c ? r : c] * n1[r > c ? c : r];
                    tA[c] = w * ((-kn * n1[c] + tv * S[c]) * n1[r] - pr);
                }
            }
            for (int b2 = lo; b2 < hi; ++b2) {
                const double n2r = B[b2].c[r];
                const double wT2 = (1 - B[b2].bd) * B[b2].tv;
                const double a = 0.5 * (wT1 + wT2);
                const double t2 = a * n2r;
                for (int c = 0; c < D; ++c) Kl[l][c] = (b2 == bj) ? Kl[l][c] + tA[c] : Kl[l][c] - t2 * n1[c];
                while (own <= b2) {
                    const double a3 = 0.5 * (H[q].wt + wT2);
                    for (int c = 0; c < D; ++c) Kl[l][c] -= a3 * n2r * H[q].c[c];
                    ++q;
                    own = q < q1 ? Ho[q] : ASM_NO_PATH;
                }
            }
        } else {  // derivative.h:122-139
            while (own < hi) {
                touched = true;
                const int b1 = own;
                const double a = 0.5 * (H[q].wt + (1 - B[b1].bd) * B[b1].tv);
                for (int c = 0; c < D; ++c) Kl[l][c] -= a * B[b1].c[r] * H[q].c[c];
                ++q;
                own = q < q1 ? Ho[q] : ASM_NO_PATH;
            }
        }
        while (own < hi) {
            ++q;
            own = q < q1 ? Ho[q] : ASM_NO_PATH;
        }
        flag[l] = touched;
    }
    for (int c = 0; c < D; ++c) {
        K[c] = 0.0;
        if (flag[0] && flag[1]) K[c] = Kl[0][c] + Kl[1][c];
        else if (flag[0]) K[c] = Kl[0][c];
        else if (flag[1]) K[c] = Kl[1][c];
    }
}

template <typename T, int D>
__device__ __forceinline__ void asm_plan_store(T *__restrict__ aval, const int64_t *__restrict__ sp, int transpose,
                                               int64_t i, int slot, int64_t j, int ts, int r, const double (&K)[D]) {
    if (!transpose) {
        for (int c = 0; c < D; ++c) aval[sell_off(sp, i, slot, D * D, r * D + c)] = (T)K[c];
    } else {
        for (int c = 0; c < D; ++c) aval[sell_off(sp, j, ts, D * D, c * D + r)] = (T)K[c];
    }
}

template <typename T, int D>
__global__ __launch_bounds__(256) void k_assemble_plan(
    int64_t n, int hops, int maxb, int diag10_literal, int transpose, const int32_t *__restrict__ bptr,
    const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2, const int32_t *__restrict__ cptr,
    const int32_t *__restrict__ cidx, const int32_t *__restrict__ tslot, const int64_t *__restrict__ sp,
    const int32_t *__restrict__ pl_row, const int32_t *__restrict__ pl_slot, const int32_t *__restrict__ pl_rng,
    const int32_t *__restrict__ pl_b3, const uint8_t *__restrict__ pl_own, const double *__restrict__ cs,
    const double *__restrict__ cs_sum, const double *__restrict__ Kn, const double *__restrict__ Tv,
    const double *__restrict__ bdamage, T *__restrict__ aval, double *__restrict__ diag64) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t hb = hops > 0 ? hops : 1;
    AsmBond *B = (AsmBond *)((char *)lds + (size_t)wid * asm_plan_lds_bytes(hops, maxb));
    AsmPath *H = (AsmPath *)(B + maxb);
    uint8_t *Ho = (uint8_t *)(H + hb);
    const int64_t i = xcd_major_block() * (blockDim.x >> 6) + wid;
    const bool row = i < n;
    // row-uniform values (one wavefront per row)
    const int ib0 = __builtin_amdgcn_readfirstlane(row ? bptr[i] : 0);
    const int ibm = __builtin_amdgcn_readfirstlane(row ? bsplit[i] : 0);
    const int ib1 = __builtin_amdgcn_readfirstlane(row ? bptr[i + 1] : 0);
    const int nbi = ib1 - ib0, tm = ibm - ib0;
    const int c0 = __builtin_amdgcn_readfirstlane(row ? cptr[i] : 0);
    const int nci = __builtin_amdgcn_readfirstlane(row ? cptr[i + 1] : 0) - c0;
    const int hdr = __builtin_amdgcn_readfirstlane(row ? pl_row[2 * i] : 0);
    const int E = __builtin_amdgcn_readfirstlane(row ? pl_row[2 * i + 1] : 0);
    const int nd = hdr & 0xffff, n2 = hdr >> 16;
    if (lane < nbi) {
        const int b = ib0 + lane;
        AsmBond r;
        r.c[0] = cs[3 * b];
        r.c[1] = cs[3 * b + 1];
        r.c[2] = cs[3 * b + 2];
        r.tv = Tv[b];
        r.kn = Kn[b];
        r.bd = bdamage[b];
        r.p2 = p2[b];
        r.pad = 0;
        B[lane] = r;
    }
    // the row's paths, four per lane per round: every load of a round is issued before its stores
    const int64_t pb = i * hops;
    for (int e0 = 0; e0 < E; e0 += 256) {
        int b3[4], ow[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + 64 * u + lane;
            b3[u] = e < E ? pl_b3[pb + e] : -1;
            ow[u] = e < E ? pl_own[pb + e] : 0;
        }
        AsmPath h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (b3[u] >= 0) {
                h[u].c[0] = cs[3 * (int64_t)b3[u]];
                h[u].c[1] = cs[3 * (int64_t)b3[u] + 1];
                h[u].c[2] = cs[3 * (int64_t)b3[u] + 2];
                h[u].wt = (1 - bdamage[b3[u]]) * Tv[b3[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + 64 * u + lane;
            if (b3[u] >= 0) {
                H[e] = h[u];
                Ho[e] = (uint8_t)ow[u];
            }
        }
    }
    __syncthreads();
    if (!row) return;
    // blocks with a bond i->j, then blocks reached by 2-hop paths only: lane = (block, row r)
    for (int t = lane; t < nd * D; t += 64) {
        const int k = t / D, r = t - k * D;
        const int ps = pl_slot[c0 + k], rg = pl_rng[c0 + k];
        const int slot = ps & 0xffff;
        const int bjl[2] = {((ps >> 16) & 0xff) - 1, ((ps >> 24) & 0xff) - 1};
        const int j = cidx[c0 + slot];
        const int ts = transpose ? tslot[c0 + slot] : 0;
        double K[D];
        asm_plan_row<D, true>(B, H, Ho, tm, nbi, j, r, bjl, rg & 0xffff, (rg & 0xffff) + (rg >> 16), cs_sum, K);
        asm_plan_store<T, D>(aval, sp, transpose, i, slot, j, ts, r, K);
    }
    for (int t = lane; t < n2 * D; t += 64) {
        const int k = nd + t / D, r = t - (t / D) * D;
        const int ps = pl_slot[c0 + k], rg = pl_rng[c0 + k];
        const int slot = ps & 0xffff;
        const int bjl[2] = {-1, -1};
        const int j = cidx[c0 + slot];
        const int ts = transpose ? tslot[c0 + slot] : 0;
        double K[D];
        asm_plan_row<D, false>(B, H, Ho, tm, nbi, j, r, bjl, rg & 0xffff, (rg & 0xffff) + (rg >> 16), cs_sum, K);
        asm_plan_store<T, D>(aval, sp, transpose, i, slot, j, ts, r, K);
    }
    // the diagonal block: one lane per upper entry (a, b), summed over both layers' bonds
    constexpr int NU = D == 3 ? 6 : 3;
    if (nd + n2 < nci && lane < NU) {
        const int slot = pl_slot[c0 + nd + n2] & 0xffff;
        const int a = D == 3 ? (lane == 0 || lane == 1 || lane == 3 ? 0 : (lane == 2 || lane == 4 ? 1 : 2))
                             : (lane == 2 ? 1 : 0);
        const int b = D == 3 ? (lane == 0 ? 0 : (lane == 1 || lane == 2 ? 1 : 2)) : (lane == 0 ? 0 : 1);
        double K = 0.0;
        for (int l = 0; l < 2; ++l) {
            const int lo = l == 0 ? 0 : tm, hi = l == 0 ? tm : nbi;
            const double Sb = cs_sum[6 * i + 3 * l + b];
            for (int q = lo; q < hi; ++q) {
                const double w = (1 - B[q].bd), kn = B[q].kn, tv = B[q].tv;
                const double ca = B[q].c[a], cb = B[q].c[b];
                if (a == b) K += w * ((kn * cb + tv * Sb) * ca + (kn + tv) * (ca * ca));
                else K += w * ((kn * cb + tv * Sb) * ca + (kn + tv) * cb * ca);
            }
        }
        const double Kba = (a == 0 && b == 1 && diag10_literal) ? 0.0 : K;  // K[b][a]
        const int ts = transpose ? tslot[c0 + slot] : slot;
        // K[a][b] goes to (a, b), or to (b, a) of the transposed store; K[b][a] the other way round
        aval[sell_off(sp, i, ts, D * D, transpose ? b * D + a : a * D + b)] = (T)K;
        if (a != b) aval[sell_off(sp, i, ts, D * D, transpose ? a * D + b : b * D + a)] = (T)Kba;
        else diag64[(int64_t)D * i + a] = K;
    }
}

// Squared norm of diag(K), Stiffness::updateStiffnessDispBC (stiffness.h:269-270). Beside it, the
// DoFs the elimination zeroes, one byte per particle: bit k of elim[p] is set when DoF k of p is
// constrained or p is fully damaged (the reference's test, stiffness.h:277); the row marks that
// k_elim_rows sets next are cleared.
__global__ __launch_bounds__(256) void k_diag_norm(int64_t nd, int D, const double *__restrict__ diag64,
                                                    const int32_t *__restrict__ cons,
                                                    const double *__restrict__ damage_visual, uint8_t *__restrict__ elim,
                                                    uint8_t *__restrict__ touched, double *partial, int cap) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {d < nd ? diag64[d] * diag64[d] : 0.0};
    if (d < nd && d % D == 0) {
        const int64_t p = d / D;
        const bool dead = fabs(damage_visual[p] - 1.0) < LPM_EPS;
        int m = 0;
        for (int k = 0; k < D; ++k)
            if (dead || cons[3 * p + k] == 1) m |= 1 << k;
        elim[p] = (uint8_t)m;
        touched[p] = 0;
    }
    block_partial<1>(v, partial, cap);
}

// The rows the elimination writes: a row of an eliminated DoF, and every row with a block in the
// column of one. The pattern is symmetric (j in row i exactly when i in row j), so these are the
// columns of the rows of particles with an eliminated DoF (the diagonal block included).
__global__ __launch_bounds__(256) void k_elim_rows(int64_t n, const int32_t *__restrict__ cptr,
                                                    const int32_t *__restrict__ cidx, const uint8_t *__restrict__ elim,
                                                    uint8_t *__restrict__ touched) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || elim[p] == 0) return;
    for (int s = cptr[p]; s < cptr[p + 1]; ++s) touched[cidx[s]] = 1;
}

// Stiffness::updateStiffnessDispBC (stiffness.h:272-293): zero the row and the column of every
// eliminated DoF (elim, above) and put ||diag K||_2 on its diagonal. One wavefront per row, a
// lane per block; rows that k_elim_rows left unmarked return at once, and a block whose row and
// column particles have no eliminated DoF is left alone.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_dirichlet(int64_t n, const int32_t *__restrict__ cptr,
                                                    const int32_t *__restrict__ cidx, const int64_t *__restrict__ sp,
                                                    const uint8_t *__restrict__ elim, const uint8_t *__restrict__ touched,
                                                    const double *scal, T *__restrict__ aval) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= n || touched[i] == 0) return;
    const int c0 = cptr[i], nci = cptr[i + 1] - c0;
    const int mi = elim[i];
    for (int slot = lane; slot < nci; slot += 64) {
        const int64_t j = cidx[c0 + slot];
        const int mj = elim[j];
        if ((mi | mj) == 0) continue;
        const T nd = (T)sqrt(scal[SC_DIAGNORM]);
        for (int q = 0; q < D; ++q) {
            const bool ej = (mj >> q) & 1;
            for (int r = 0; r < D; ++r) {
                const bool ei = (mi >> r) & 1;
                if (ei || ej) {
                    T v = (T)0;
                    if (ei && j == i && r == q) v = nd;
                    aval[sell_off(sp, i, slot, D * D, r * D + q)] = v;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- constitutive updates
// calcStrainEnergyDensity (particle_elastic_damage.h:89-104); equals the energy_total that
// ParticleFatigueHCF::calcEqEnergy returns (particle_fatigue_hcf.h:82-125).
__device__ double strain_energy(int64_t i, double volume, int nneighbors, const int32_t *bptr, const int32_t *bsplit,
                                const double *Kn, const double *dLe, const double *dLe_total,
                                const double *TdLe_total) {
    const int nb = bptr[i + 1] - bptr[i];
    const double V_m = volume * nb / nneighbors;
    double e = 0;
    for (int l = 0; l < 2; ++l) {
        const int lo = l == 0 ? bptr[i] : bsplit[i], hi = l == 0 ? bsplit[i] : bptr[i + 1];
        for (int b = lo; b < hi; ++b) e += 0.5 * (Kn[b] * dLe[b] * dLe[b]) / V_m;
        e += 0.5 * TdLe_total[2 * i + l] * dLe_total[2 * i + l] / V_m;
    }
    return e;
}

// updateParticleStateVariables: ElasticDamage (particle_elastic_damage.h:150-161),
// J2Plasticity (particle_j2plasticity.h:78-140), FatigueHCF (particle_fatigue_hcf.h:127-145)
__global__ __launch_bounds__(256) void k_state_var(int64_t n, int model, double volume, int nneighbors,
                                                    const int32_t *__restrict__ bptr, const int32_t *__restrict__ bsplit,
                                                    const double *__restrict__ Kn, const double *__restrict__ dLe,
                                                    const double *__restrict__ dLe_total,
                                                    const double *__restrict__ TdLe_total,
                                                    const double *__restrict__ stress, const double *__restrict__ cs,
                                                    const double *__restrict__ dis, const double *__restrict__ mat,
                                                    const double *__restrict__ damage, const double *__restrict__ sv_last,
                                                    double *__restrict__ sv, double *__restrict__ dLp,
                                                    double *__restrict__ Ddot_local) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *m = mat + LPM_NMAT * i;
    double *s = sv + LPM_NSV * i;
    if (model == ELASTIC_DAMAGE) {
        const double k0 = m[0], k1 = m[1];
        s[0] = strain_energy(i, volume, nneighbors, bptr, bsplit, Kn, dLe, dLe_total, TdLe_total);
        const double delta = s[0] - sv_last[LPM_NSV * i];
        const double f = s[0] - k1 * damage[i] - k0;
        Ddot_local[i] = (f > 0 && delta > 0) ? (1 / k1) * delta : 0;
    } else if (model == J2) {
        const double E = m[0], mu = m[1], sigmay = m[2], xi = m[3], H = m[4], A = m[5];
        double st[6], dpl[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 6; ++k) st[k] = stress[6 * i + k];
        const double sigma_m = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
        for (int k = 0; k < 3; ++k) st[k] -= sigma_m;
        for (int k = 0; k < 6; ++k) st[k] -= s[k + 1];
        double sigma_eq = 0.0;
        for (int k = 0; k < 6; ++k) {
            if (k < 3) sigma_eq += st[k] * st[k];
            else sigma_eq += 2.0 * st[k] * st[k];
        }
        sigma_eq = sqrt(3.0 / 2.0 * sigma_eq);
        const double yield_func = sigma_eq - (sigmay + (1.0 - xi) * H * s[0]);
        const double G = E / 2 / (1 + mu);
        double dlambda = 0.0;
        if (yield_func > 0.0) dlambda = yield_func / (3 * G + H);
        s[0] += dlambda;
        for (int k = 0; k < 6; ++k) {
            if (fabs(sigma_eq) > LPM_EPS) {
                dpl[k] = dlambda * 1.5 * st[k] / sigma_eq;
                s[k + 1] += 2. / 3. * xi * H * dpl[k];
            }
        }
        for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
            const double cx = cs[3 * b], cy = cs[3 * b + 1], cz = cs[3 * b + 2];
            const double ddLp = dis[b] * (dpl[0] * cx * cx + dpl[1] * cy * cy + dpl[2] * cz * cz + 2 * dpl[3] * cy * cz +
                                          2 * dpl[4] * cx * cz + 2 * dpl[5] * cx * cy);
            dLp[b] += ddLp;
        }
        double triax = 0.0;
        if (sigma_eq > LPM_EPS) triax = sigma_m / sigma_eq;
        Ddot_local[i] = dlambda * (1.0 + A * triax);
    } else if (model == HCF) {
        const double A = m[0], B = m[1], dd = m[2];
        const double curr = strain_energy(i, volume, nneighbors, bptr, bsplit, Kn, dLe, dLe_total, TdLe_total);
        s[1] = dmin(curr, s[1]);
        s[2] = dmax(curr, s[2]);
        s[0] = dmax(s[2], s[0]);
        Ddot_local[i] = (s[2] - s[1] > 0) ? A * pow(s[0], B) * pow(s[2] - s[1], dd) : 0;
    }
}

// nonlocal damage-rate average (solver_static.h:34-52; solver_fatigue.h:78-114), with
// func_phi (lpm.h:45-49)
__global__ __launch_bounds__(256) void k_nonlocal_ddot(int64_t n, int undamaged, double volume, int nneighbors,
                                                        const int32_t *__restrict__ type, const int32_t *__restrict__ nlptr,
                                                        const int32_t *__restrict__ nlidx, const int32_t *__restrict__ bptr,
                                                        const double *__restrict__ xyz, const double *__restrict__ L,
                                                        const double *__restrict__ Ddot_local,
                                                        double *__restrict__ Ddot_nonlocal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || type[i] == undamaged) return;
    const double x0 = xyz[3 * i], y0 = xyz[3 * i + 1], z0 = xyz[3 * i + 2], Li = L[i];
    double A = 0, dn = 0;
    const int lo = nlptr ? nlptr[i] : 0, hi = nlptr ? nlptr[i + 1] : 0;
    for (int q = lo; q < hi; ++q) {
        const int64_t j = nlidx[q];
        if (type[j] == undamaged) continue;
        const double dx = x0 - xyz[3 * j], dy = y0 - xyz[3 * j + 1], dz = z0 - xyz[3 * j + 2];
        const double dis = sqrt(dx * dx + dy * dy + dz * dz);
        const double V_m = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
        const double u = 1 / (1 + pow(dis / Li, 8.0));
        const double phi = u * u;
        dn += Ddot_local[j] * phi * V_m;
        A += phi * V_m;
    }
    Ddot_nonlocal[i] = dn / A;
}

__device__ __forceinline__ void set_flag(bool f, int32_t *flag) {
    // one store per wavefront that saw the event
    const unsigned long long m = __ballot(f);
    if (m && (threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicOr(flag, 1);
}

// updateParticleStaticDamage: Elastic (particle_elastic.h:52-68), ElasticDamage
// (particle_elastic_damage.h:163-173), J2Plasticity (particle_j2plasticity.h:142-152)
__global__ __launch_bounds__(256) void k_static_damage(int64_t n, int model, int undamaged,
                                                        const int32_t *__restrict__ type, const int32_t *__restrict__ bptr,
                                                        const double *__restrict__ mat, const double *__restrict__ bstrain,
                                                        double *__restrict__ bdamage, const double *__restrict__ Ddot_nl,
                                                        const double *__restrict__ damage_last,
                                                        double *__restrict__ damage, int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (i < n && type[i] != undamaged) {
        const double *m = mat + LPM_NMAT * i;
        if (model == ELASTIC) {
            const double crit = m[0];
            for (int b = bptr[i]; b < bptr[i + 1]; ++b)
                if (bstrain[b] >= crit && fabs(bdamage[b] - 1.0) > LPM_EPS) {
                    any = true;
                    bdamage[b] = 1;
                }
        } else if (model == ELASTIC_DAMAGE || model == J2) {
            const double thr = model == ELASTIC_DAMAGE ? m[2] : m[6];
            double D = damage[i] + Ddot_nl[i];
            if (D >= thr) D = 1;
            damage[i] = D;
            any = fabs(D - damage_last[i]) > 1e-3;
        }
    }
    set_flag(any, flag);
}

// ParticleFatigueHCF::updateParticleFatigueDamage (particle_fatigue_hcf.h:147-160)
__global__ __launch_bounds__(256) void k_fatigue_damage(int64_t n, int model, int undamaged, double dt,
                                                         const int32_t *__restrict__ type, const double *__restrict__ mat,
                                                         const double *__restrict__ Ddot_nl,
                                                         const double *__restrict__ damage_last,
                                                         double *__restrict__ damage, double *__restrict__ sv,
                                                         int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (i < n && type[i] != undamaged && model == HCF) {
        double D = damage[i] + Ddot_nl[i] * dt;
        if (D >= mat[LPM_NMAT * i + 5]) D = 1;
        damage[i] = D;
        sv[LPM_NSV * i + 1] = (double)INT_MAX;
        sv[LPM_NSV * i + 2] = 0;
        any = fabs(D - damage_last[i]) > 1e-3;
    }
    set_flag(any, flag);
}

// energy range s2 - s1 of each particle before a fatigue update resets it (0 when not positive)
__global__ __launch_bounds__(256) void k_energy_range(int64_t n, const double *__restrict__ sv, double *__restrict__ range) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = sv[LPM_NSV * i + 2] - sv[LPM_NSV * i + 1];
    range[i] = d > 0 ? d : 0.0;
}

// ---------------------------------------------------------------- pattern (2-hop conns)
// Assembly::updateConnections (assembly.h:236-266): sorted unique {bond partners and their
// same-layer partners}, self included. One workgroup per particle: candidates are gathered in
// LDS, sorted with a bitonic network and de-duplicated.
__global__ __launch_bounds__(256) void k_pattern(int64_t n, const int32_t *__restrict__ bptr,
                                                  const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                  int32_t *__restrict__ counts, const int32_t *__restrict__ cptr,
                                                  int32_t *__restrict__ cidx, int32_t *overflow) {
    __shared__ int32_t buf[LPM_MAXC];
    __shared__ int32_t off[LPM_MAXB + 1];
    __shared__ int32_t scan[LPM_MAXC];
    const int64_t i = blockIdx.x;
    const int t = threadIdx.x;
    const int b0 = bptr[i], bm = bsplit[i], nb = bptr[i + 1] - b0;
    if (t == 0) {
        int o = 0;
        for (int q = 0; q < nb; ++q) {
            off[q] = o;
            const int k = p2[b0 + q];
            const bool l0 = (b0 + q) < bm;
            o += 1 + (l0 ? bsplit[k] - bptr[k] : bptr[k + 1] - bsplit[k]);
        }
        off[nb] = o;
    }
    for (int q = t; q < LPM_MAXC; q += blockDim.x) buf[q] = INT_MAX;
    __syncthreads();
    const int total = off[nb];
    if (total > LPM_MAXC) {
        if (t == 0) atomicOr(overflow, 1);
        return;
    }
    if (t < nb) {
        const int k = p2[b0 + t];
        const bool l0 = (b0 + t) < bm;
        int o = off[t];
        buf[o++] = k;
        const int lo = l0 ? bptr[k] : bsplit[k], hi = l0 ? bsplit[k] : bptr[k + 1];
        for (int b = lo; b < hi; ++b) buf[o++] = p2[b];
    }
    __syncthreads();
    // bitonic sort of LPM_MAXC entries, 2 per thread per stage
    for (int size = 2; size <= LPM_MAXC; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = t; q < LPM_MAXC / 2; q += blockDim.x) {
                const int pos = 2 * q - (q & (stride - 1));
                const int a = pos, b = pos + stride;
                const bool up = ((a & size) == 0);
                const int va = buf[a], vb = buf[b];
                if ((va > vb) == up) {
                    buf[a] = vb;
                    buf[b] = va;
                }
            }
            __syncthreads();
        }
    }
    // unique flags and inclusive scan (Hillis-Steele over LPM_MAXC entries)
    for (int q = t; q < LPM_MAXC; q += blockDim.x)
        scan[q] = (buf[q] != INT_MAX && (q == 0 || buf[q] != buf[q - 1])) ? 1 : 0;
    __syncthreads();
    for (int s = 1; s < LPM_MAXC; s <<= 1) {
        int v0 = 0, v1 = 0;
        const int q0 = t, q1 = t + blockDim.x;
        if (q0 < LPM_MAXC) v0 = scan[q0] + (q0 >= s ? scan[q0 - s] : 0);
        if (q1 < LPM_MAXC) v1 = scan[q1] + (q1 >= s ? scan[q1 - s] : 0);
        __syncthreads();
        if (q0 < LPM_MAXC) scan[q0] = v0;
        if (q1 < LPM_MAXC) scan[q1] = v1;
        __syncthreads();
    }
    if (cidx == nullptr) {
        if (t == 0) counts[i] = scan[LPM_MAXC - 1];
        return;
    }
    const int base = cptr[i];
    for (int q = t; q < LPM_MAXC; q += blockDim.x) {
        const bool first = buf[q] != INT_MAX && (q == 0 || buf[q] != buf[q - 1]);
        if (first) cidx[base + scan[q] - 1] = buf[q];
    }
}

// SELL-64 column array: block (row i, slot c) at sp[i>>6] + c*64 + (i&63); padding slots keep
// the row's own index so that every stored column is a valid particle id.
__global__ __launch_bounds__(256) void k_sell_cols(int64_t n, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                                    const int64_t *__restrict__ sp, int32_t *__restrict__ scol) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k = i >> 6, l = i & 63;
    const int W = (int)((sp[k + 1] - sp[k]) >> 6);
    const int c0 = cptr[i], nc = cptr[i + 1] - c0;
    for (int c = 0; c < W; ++c) scol[sp[k] + (int64_t)c * 64 + l] = c < nc ? cidx[c0 + c] : (int32_t)i;
}

// The SELL columns as int16 offsets j - i (padding slots 0: the row itself); *bad is set when an
// offset does not fit, and the int16 copy is then not used.
__global__ __launch_bounds__(256) void k_sell_cols16(int64_t n, const int32_t *__restrict__ cptr,
                                                      const int32_t *__restrict__ cidx, const int64_t *__restrict__ sp,
                                                      int16_t *__restrict__ scol16, int32_t *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k = i >> 6, l = i & 63;
    const int W = (int)((sp[k + 1] - sp[k]) >> 6);
    const int c0 = cptr[i], nc = cptr[i + 1] - c0;
    for (int c = 0; c < W; ++c) {
        int d = c < nc ? cidx[c0 + c] - (int)i : 0;
        if (d < -32768 || d > 32767) {
            *bad = 1;
            d = 0;
        }
        scol16[sp[k] + (int64_t)c * 64 + l] = (int16_t)d;
    }
}

// position of block (j, i) in block row j, for the transposed store
__global__ __launch_bounds__(256) void k_tslot(int64_t nconn, int64_t n, const int32_t *__restrict__ cptr,
                                                const int32_t *__restrict__ cidx, const int32_t *__restrict__ rowof,
                                                int32_t *__restrict__ tslot, int32_t *bad) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nconn) return;
    const int i = rowof[e], j = cidx[e];
    int lo = cptr[j], hi = cptr[j + 1] - 1, pos = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = cidx[mid];
        if (v == i) { pos = mid; break; }
        if (v < i) lo = mid + 1;
        else hi = mid - 1;
    }
    if (pos < 0) {
        atomicOr(bad, 1);
        tslot[e] = 0;
    } else {
        tslot[e] = pos - cptr[j];
    }
}

__global__ __launch_bounds__(256) void k_rowof(int64_t n, const int32_t *__restrict__ cptr, int32_t *__restrict__ rowof) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int e = cptr[i]; e < cptr[i + 1]; ++e) rowof[e] = (int32_t)i;
}

// Assembly::searchNonlocalNeighbors (assembly.h:117-130): all j (self included) with
// |x_i - x_j| < ratio * L_i, in id order. Positions are staged through LDS tiles.
__global__ __launch_bounds__(256) void k_nonlocal_search(int64_t n, double ratio, const double *__restrict__ xyz,
                                                          const double *__restrict__ L, const int32_t *__restrict__ spec_of,
                                                          int32_t *__restrict__ counts, const int32_t *__restrict__ nlptr,
                                                          int32_t *__restrict__ nlidx) {
    __shared__ double tile[3 * 256];
    __shared__ int32_t tspec[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n;
    double x0 = 0, y0 = 0, z0 = 0, cut = 0;
    if (act) { x0 = xyz[3 * i]; y0 = xyz[3 * i + 1]; z0 = xyz[3 * i + 2]; cut = ratio * L[i]; }
    int cnt = 0;
    int64_t o = (act && nlidx) ? nlptr[i] : 0;
    const int32_t si = (act && spec_of) ? spec_of[i] : 0;  // an ensemble's specimens do not see each other
    for (int64_t base = 0; base < n; base += 256) {
        const int64_t j = base + threadIdx.x;
        __syncthreads();
        if (j < n) {
            tile[3 * threadIdx.x] = xyz[3 * j];
            tile[3 * threadIdx.x + 1] = xyz[3 * j + 1];
            tile[3 * threadIdx.x + 2] = xyz[3 * j + 2];
            tspec[threadIdx.x] = spec_of ? spec_of[j] : 0;
        }
        __syncthreads();
        const int m = (int)((n - base) < 256 ? (n - base) : 256);
        if (act) {
            for (int q = 0; q < m; ++q) {
                const double dx = x0 - tile[3 * q], dy = y0 - tile[3 * q + 1], dz = z0 - tile[3 * q + 2];
                if (tspec[q] == si && sqrt(dx * dx + dy * dy + dz * dz) < cut) {
                    if (nlidx) nlidx[o++] = (int32_t)(base + q);
                    ++cnt;
                }
            }
        }
    }
    if (act && !nlidx) counts[i] = cnt;
}

// The same lists from a grid of cells no narrower than the largest cutoff, one layer of cells per
// specimen (cell = (specimen * nc2 + z) * nc1 + y) * nc0 + x): each particle tests the members of
// the 27 surrounding cells of its own layer with the expression above, then k_nl_sort orders
// every list by id, the order of the all-pairs loop.
__global__ __launch_bounds__(256) void k_nl_bin(int64_t n, const double *__restrict__ xyz, double lo0, double lo1,
                                                 double lo2, double h, int64_t nc0, int64_t nc1, int64_t nc2,
                                                 const int32_t *__restrict__ spec_of, int32_t *__restrict__ pcell,
                                                 int32_t *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = {lo0, lo1, lo2};
    const int64_t nc[3] = {nc0, nc1, nc2};
    int64_t q[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = (int64_t)floor((xyz[3 * i + k] - lo[k]) / h);
        q[k] = q[k] < 0 ? 0 : (q[k] > nc[k] - 1 ? nc[k] - 1 : q[k]);
    }
    const int64_t layer = spec_of ? spec_of[i] : 0;
    const int64_t cell = ((layer * nc2 + q[2]) * nc1 + q[1]) * nc0 + q[0];
    pcell[i] = (int32_t)cell;
    atomicAdd(&count[cell], 1);
}

template <bool FILL>
__global__ __launch_bounds__(256) void k_nl_cells(int64_t n, double ratio, const double *__restrict__ xyz,
                                                   const double *__restrict__ L, const int32_t *__restrict__ pcell,
                                                   int64_t nc0, int64_t nc1, int64_t nc2, const int32_t *__restrict__ head,
                                                   const int32_t *__restrict__ items, int32_t *__restrict__ counts,
                                                   const int32_t *__restrict__ nlptr, int32_t *__restrict__ nlidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x0 = xyz[3 * i], y0 = xyz[3 * i + 1], z0 = xyz[3 * i + 2], cut = ratio * L[i];
    const int64_t ncs = nc0 * nc1 * nc2, cell = pcell[i];
    const int64_t layer = cell / ncs, loc = cell - layer * ncs;
    const int64_t cx = loc % nc0, cy = (loc / nc0) % nc1, cz = loc / (nc0 * nc1);
    int cnt = 0;
    int64_t o = FILL ? nlptr[i] : 0;
    for (int64_t ez = -1; ez <= 1; ++ez)
        for (int64_t ey = -1; ey <= 1; ++ey)
            for (int64_t ex = -1; ex <= 1; ++ex) {
                const int64_t x = cx + ex, y = cy + ey, z = cz + ez;
                if (x < 0 || y < 0 || z < 0 || x >= nc0 || y >= nc1 || z >= nc2) continue;
                const int64_t cc = layer * ncs + (z * nc1 + y) * nc0 + x;
                for (int t = head[cc]; t < head[cc + 1]; ++t) {
                    const int64_t j = items[t];
                    const double dx = x0 - xyz[3 * j], dy = y0 - xyz[3 * j + 1], dz = z0 - xyz[3 * j + 2];
                    if (sqrt(dx * dx + dy * dy + dz * dz) < cut) {
                        if (FILL) nlidx[o++] = (int32_t)j;
                        ++cnt;
                    }
                }
            }
    if (!FILL) counts[i] = cnt;
}

// Sorts list i (workgroup i, 64 threads) ascending: a bitonic network in LDS over the list
// padded with INT32_MAX to P, the power of two the launch sizes the LDS for (P >= every length).
__global__ __launch_bounds__(64) void k_nl_sort(const int32_t *__restrict__ nlptr, int32_t *__restrict__ nlidx) {
    extern __shared__ int32_t srt[];
    const int64_t i = blockIdx.x;
    const int o = nlptr[i], k = nlptr[i + 1] - o;
    if (k <= 1) return;  // uniform over the workgroup
    int P = 2;
    while (P < k) P <<= 1;
    for (int e = threadIdx.x; e < P; e += 64) srt[e] = e < k ? nlidx[o + e] : INT32_MAX;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < (P >> 1); t += 64) {
                const int a = 2 * t - (t & (stride - 1)), b = a + stride;  // pair t of this stage
                const bool up = (a & size) == 0;
                const int32_t u = srt[a], v = srt[b];
                if ((u > v) == up) {
                    srt[a] = v;
                    srt[b] = u;
                }
            }
            __syncthreads();
        }
    for (int e = threadIdx.x; e < k; e += 64) nlidx[o + e] = srt[e];
}


// ---------------------------------------------------------------- bond search on the device
// Assembly::createBonds (assembly.h:212-234) tests every pair; here particles are binned on a
// grid whose cell edge exceeds the bond cutoff, and each particle examines the 27 (9 in 2D)
// surrounding cells. The reference's test and distance expression (Particle::distanceTo,
// particle.h:276-280) decide acceptance and layer, and each particle's lists are kept sorted by
// partner id, the order of the reference's all-pairs loop, so the result is bit-identical to it.

__global__ __launch_bounds__(256) void k_cell_count(int64_t n, const double *__restrict__ xyz, double lo0, double lo1,
                                                     double lo2, double h, int64_t nc0, int64_t nc1, int64_t nc2,
                                                     int32_t *__restrict__ pcell, int32_t *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = {lo0, lo1, lo2};
    const int64_t nc[3] = {nc0, nc1, nc2};
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = (int64_t)floor((xyz[3 * i + k] - lo[k]) / h);
        c[k] = c[k] < 0 ? 0 : (c[k] > nc[k] - 1 ? nc[k] - 1 : c[k]);
    }
    const int64_t cell = (c[2] * nc1 + c[1]) * nc0 + c[0];
    pcell[i] = (int32_t)cell;
    atomicAdd(&count[cell], 1);
}

// members of each cell (their order inside a cell does not matter: every list is sorted below)
__global__ __launch_bounds__(256) void k_cell_fill(int64_t n, const int32_t *__restrict__ pcell,
                                                    const int32_t *__restrict__ head, int32_t *__restrict__ fill,
                                                    int32_t *__restrict__ items) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t cell = pcell[i];
    items[head[cell] + atomicAdd(&fill[cell], 1)] = (int32_t)i;
}

// Cut regions of the reference's geometry tools (geometry/geo1_CT_configuration.cpp:45-74,
// isValid at :137-171; the dogbone tools repeat them): a point is valid if it lies in no cut.
// inCircle: (x-cx)^2 + (y-cy)^2 <= r^2. inNotch: |y - ty| <= w and the point lies between the
// lines through the tip of slope -+1/sqrt(3) on the side x > tx; the reverse notch opens to x < tx.
// z is ignored. Written in the tools' operation order (file compiled without FMA contraction).
__host__ __device__ inline bool cut_point_valid(double px, double py, const lpm_cut *cuts, int ncut) {
    const double s3 = 1. / sqrt(3.);
    for (int k = 0; k < ncut; ++k) {
        const lpm_cut &c = cuts[k];
        if (c.kind == LPM_CUT_CIRCLE) {
            const double dx = px - c.x, dy = py - c.y;
            if (dx * dx + dy * dy <= c.size * c.size) return false;
        } else if (fabs(py - c.y) <= c.size) {
            const double up = py - c.y;
            if (c.kind == LPM_CUT_NOTCH) {
                if (up >= s3 * (c.x - px) && up <= s3 * (-c.x + px)) return false;
            } else if (up <= s3 * (c.x - px) && up >= s3 * (-c.x + px)) {
                return false;
            }
        }
    }
    return true;
}

// searchNeighbor's bond test (geometry/geo1_CT_configuration.cpp:95-108): the bond i->j is cut
// if the midpoint of any of its 10 equal sub-segments is not valid.
__host__ __device__ inline bool cut_bond_valid(const double *a, const double *b, const lpm_cut *cuts, int ncut) {
    const int n_interval = 10;
    const double dx = (b[0] - a[0]) / n_interval, dy = (b[1] - a[1]) / n_interval;
    bool valid = true;
    for (int k = 0; k < n_interval; ++k) {
        const double x1 = a[0] + k * dx, y1 = a[1] + k * dy, x2 = a[0] + (k + 1) * dx, y2 = a[1] + (k + 1) * dy;
        valid = cut_point_valid((x1 + x2) / 2.0, (y1 + y2) / 2.0, cuts, ncut) && valid;
    }
    return valid;
}

// insert j into the ascending row[0..m) and return m + 1
__device__ __forceinline__ int sorted_insert(int32_t *row, int m, int32_t j) {
    int pos = m;
    while (pos > 0 && row[pos - 1] > j) {
        row[pos] = row[pos - 1];
        --pos;
    }
    row[pos] = j;
    return m + 1;
}

// One thread per particle: accepted partners are insertion-sorted by id into its layer-0 and
// layer-1 rows of tmp0 / tmp1 (LPM_MAXB entries each); flag[0] = 1 if a particle has more than
// LPM_MAXB bonds (the rows are then incomplete and the caller fails).
__global__ __launch_bounds__(256) void k_bond_search(int64_t n, const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ pcell, int64_t nc0, int64_t nc1,
                                                      int64_t nc2, const int32_t *__restrict__ head,
                                                      const int32_t *__restrict__ items, double c0, double c1,
                                                      const lpm_cut *__restrict__ cuts, int ncut,
                                                      int32_t *__restrict__ tmp0, int32_t *__restrict__ tmp1,
                                                      int32_t *__restrict__ cnt0, int32_t *__restrict__ cnt1,
                                                      int32_t *__restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t cell = pcell[i];
    const int64_t cx = cell % nc0, cy = (cell / nc0) % nc1, cz = cell / (nc0 * nc1);
    const double ax = xyz[3 * i], ay = xyz[3 * i + 1], az = xyz[3 * i + 2];
    int32_t *r0 = tmp0 + i * LPM_MAXB, *r1 = tmp1 + i * LPM_MAXB;
    int n0 = 0, n1 = 0;
    bool over = false;
    for (int64_t dz = -1; dz <= 1; ++dz)
        for (int64_t dy = -1; dy <= 1; ++dy)
            for (int64_t dx = -1; dx <= 1; ++dx) {
                const int64_t x = cx + dx, y = cy + dy, z = cz + dz;
                if (x < 0 || y < 0 || z < 0 || x >= nc0 || y >= nc1 || z >= nc2) continue;
                const int64_t cc = (z * nc1 + y) * nc0 + x;
                for (int t = head[cc]; t < head[cc + 1]; ++t) {
                    const int32_t j = items[t];
                    if (j == (int32_t)i) continue;
                    const double ex = ax - xyz[3 * (int64_t)j], ey = ay - xyz[3 * (int64_t)j + 1],
                                 ez = az - xyz[3 * (int64_t)j + 2];
                    const double d = sqrt(ex * ex + ey * ey + ez * ez);
                    if (!(d < c1)) continue;
                    if (ncut > 0 && !cut_bond_valid(&xyz[3 * i], &xyz[3 * (int64_t)j], cuts, ncut)) continue;
                    if (n0 + n1 >= LPM_MAXB) {
                        over = true;
                        continue;
                    }
                    if (d < c0) n0 = sorted_insert(r0, n0, j);
                    else n1 = sorted_insert(r1, n1, j);
                }
            }
    cnt0[i] = n0;
    cnt1[i] = n1;
    if (over) flag[0] = 1;
}

__global__ __launch_bounds__(256) void k_bond_compact(int64_t n, const int32_t *__restrict__ bptr,
                                                       const int32_t *__restrict__ bsplit, const int32_t *__restrict__ tmp0,
                                                       const int32_t *__restrict__ tmp1, int32_t *__restrict__ p2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    for (int k = 0; k < bm - b0; ++k) p2[b0 + k] = tmp0[i * LPM_MAXB + k];
    for (int k = 0; k < b1 - bm; ++k) p2[bm + k] = tmp1[i * LPM_MAXB + k];
}

// Per particle: bond count, Bond(p1, p2, layer, dis) initial geometry (bond.h:61-73), the reverse
// bond (found in the partner's list of the same layer, which exists since the acceptance test is
// symmetric) and the number of 2-hop records of the assembly (max over particles in hops_max).
__global__ __launch_bounds__(256) void k_bond_init(int64_t n, const int32_t *__restrict__ bptr,
                                                    const int32_t *__restrict__ bsplit, const int32_t *__restrict__ p2,
                                                    const double *__restrict__ xyz, int32_t *__restrict__ nb,
                                                    int32_t *__restrict__ opp, double *__restrict__ dis0,
                                                    double *__restrict__ dis, double *__restrict__ dis_last,
                                                    double *__restrict__ cs, int32_t *__restrict__ hops_max,
                                                    int32_t *__restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b0 = bptr[i], bm = bsplit[i], b1 = bptr[i + 1];
    nb[i] = b1 - b0;
    int hops = 0;
    for (int b = b0; b < b1; ++b) {
        const int64_t j = p2[b];
        const double ex = xyz[3 * i] - xyz[3 * j], ey = xyz[3 * i + 1] - xyz[3 * j + 1], ez = xyz[3 * i + 2] - xyz[3 * j + 2];
        const double d = sqrt(ex * ex + ey * ey + ez * ez);
        dis0[b] = d;
        dis[b] = d;
        dis_last[b] = d;
        cs[3 * b] = ex / d;
        cs[3 * b + 1] = ey / d;
        cs[3 * b + 2] = ez / d;
        const bool l0 = b < bm;
        const int lo = l0 ? bptr[j] : bsplit[j], hi = l0 ? bsplit[j] : bptr[j + 1];
        int o = -1;
        for (int q = lo; q < hi; ++q)
            if (p2[q] == (int32_t)i) {
                o = q;
                break;
            }
        opp[b] = o;
        if (o < 0) flag[1] = 1;
        hops += hi - lo;
    }
    atomicMax(hops_max, hops);
}
}  // namespace

// =================================================================== launchers
static int check_launch(lpm_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LPM_OK;
}

int launch_geometry_force(lpm_ctx *c, bool force) {
    const size_t lds = (size_t)c->gp_bonds * (5 * sizeof(double) + sizeof(int));
    hipLaunchKernelGGL(k_geometry_force, dim3(blocks_for(c->n, GP)), dim3(64), lds, c->stream, c->n, c->model, force ? 1 : 0,
                       c->bptr, c->bsplit, c->p2, c->xyz, c->dis0, c->dLp, c->Kn, c->Tv, c->damage, c->dis, c->dLe,
                       c->bstrain, c->cs, c->dLe_total, c->TdLe_total, c->cs_sum, c->bforce, c->bdamage);
    return check_launch(c, "k_geometry_force");
}

int launch_bond_force(lpm_ctx *c) {
    hipLaunchKernelGGL(k_bond_force, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->model, c->bptr, c->bsplit,
                       c->p2, c->Kn, c->Tv, c->dLe, c->dLe_total, c->damage, c->bforce, c->bdamage);
    return check_launch(c, "k_bond_force");
}

int launch_particle_force(lpm_ctx *c) {
    const size_t lds = (size_t)c->gp_bonds * 6 * sizeof(double);
    hipLaunchKernelGGL(k_particle_force, dim3(blocks_for(c->n, GP)), dim3(64), lds, c->stream, c->n, c->volume,
                       c->nneighbors, c->bptr, c->opp, c->cs, c->dis, c->bforce, c->bdamage, c->Pin, c->stress,
                       c->damage_visual);
    return check_launch(c, "k_particle_force");
}

int launch_residual(lpm_ctx *c, double *nr2, double *nR2) {
    const int64_t nd = c->n * c->dim;
    const int g = blocks_for(nd, 256);
    if (g > c->partial_cap) return lpm_fail(c, LPM_E_LIMIT, "residual grid exceeds partial buffer");
    hipLaunchKernelGGL(k_residual, dim3(g), dim3(256), 0, c->stream, c->n, c->dim, c->cons, c->Pex, c->Pin, c->residual,
                       c->partial, c->partial_cap);
    int rc = check_launch(c, "k_residual");
    if (rc) return rc;
    rc = launch_finalize(c, FIN_NORMS, g);
    if (rc) return rc;
    if (nr2 || nR2) {
        LPM_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        if (nr2) *nr2 = c->h_scal[SC_NORM_R];
        if (nR2) *nR2 = c->h_scal[SC_NORM_REAC];
    }
    return LPM_OK;
}

// ---------------------------------------------------------------- updateRR without allocation
// The residual and its two norms (lpm_update_rr's kernels and grid) on caller-owned device
// arrays. The per-block partial sums and the finalize's scalars live in `scratch`, a file-scope
// device array that every GPU holds once for the life of the process; its address on each
// device is looked up on that device's first call. Every slot the finalize reads is written by
// the same call (one partial pair per block, then the two norms), so nothing needs clearing;
// calls that overlap on one device race on the array.
static constexpr int NOALLOC_CAP = LPM_NOALLOC_MAX_DOFS / 256;  // partial pairs (one per block of 256 DoFs)
static __device__ __align__(256) double scratch[SC_COUNT + 2 * NOALLOC_CAP];
static constexpr int NOALLOC_DEVICES = 64;
static std::atomic<double *> g_scratch[NOALLOC_DEVICES];

extern "C" int lpm_update_rr_noalloc(int64_t n, int32_t dim, const double *Pex, const double *Pin, const int32_t *cons,
                                     double *residual, double *norms, void *stream) {
    if (n <= 0 || (dim != 2 && dim != 3) || !Pex || !Pin || !cons || !residual)
        return lpm_fail(nullptr, LPM_E_ARG, "lpm_update_rr_noalloc: bad arguments");
    const int64_t nd = n * dim;
    if (nd > LPM_NOALLOC_MAX_DOFS)
        return lpm_fail(nullptr, LPM_E_LIMIT, "lpm_update_rr_noalloc: " + std::to_string(nd) + " DoFs exceed the limit of " +
                                                  std::to_string(LPM_NOALLOC_MAX_DOFS));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_DEVICES)
        return lpm_fail(nullptr, LPM_E_HIP, "lpm_update_rr_noalloc: no current device");
    double *scal = g_scratch[dev].load();
    if (!scal) {
        const hipError_t e = hipGetSymbolAddress((void **)&scal, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !scal)
            return lpm_fail(nullptr, LPM_E_HIP, std::string("lpm_update_rr_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_scratch[dev].store(scal);
    }
    double *partial = scal + SC_COUNT;
    hipStream_t st = (hipStream_t)stream;
    const int g = blocks_for(nd, 256);
    hipLaunchKernelGGL(k_residual, dim3(g), dim3(256), 0, st, n, dim, cons, Pex, Pin, residual, partial, NOALLOC_CAP);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_finalize_raw(st, FIN_NORMS, g, partial, NOALLOC_CAP, scal);
    if (e == hipSuccess && norms) {
        double r[SC_COUNT];
        e = hipMemcpyAsync(r, scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess) {
            norms[0] = std::sqrt(r[SC_NORM_R]);
            norms[1] = std::sqrt(r[SC_NORM_REAC]);
        }
    }
    if (e != hipSuccess) return lpm_fail(nullptr, LPM_E_HIP, std::string("lpm_update_rr_noalloc: ") + hipGetErrorString(e));
    return LPM_OK;
}

int launch_residual_spec(lpm_ctx *c, std::vector<double> &r2, std::vector<double> &R2) {
    const int g = c->spec_nblk;
    hipLaunchKernelGGL(k_residual_spec, dim3(g), dim3(256), 0, c->stream, c->nspec, c->dim, (const int64_t *)c->spec_first_d,
                       (const int32_t *)c->spec_blk, c->cons, c->Pex, c->Pin, c->residual, c->spec_part, g);
    int rc = check_launch(c, "k_residual_spec");
    if (rc) return rc;
    std::vector<double> h(2 * (size_t)g);
    LPM_HIP(c, hipMemcpyAsync(h.data(), c->spec_part, sizeof(double) * 2 * g, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    r2.assign(c->nspec, 0.0);
    R2.assign(c->nspec, 0.0);
    // the blocks of specimen s cover DoFs of s only; sum them in block order
    int64_t b0 = 0;
    for (int sp = 0; sp < c->nspec; ++sp) {
        const int64_t nb = blocks_for((c->spec_first[sp + 1] - c->spec_first[sp]) * c->dim, 256);
        for (int64_t b = b0; b < b0 + nb; ++b) {
            r2[sp] += h[b];
            R2[sp] += h[g + b];
        }
        b0 += nb;
    }
    return LPM_OK;
}

int launch_spec_scale(lpm_ctx *c, double *v) {
    const int64_t nd = c->n * c->dim;
    hipLaunchKernelGGL(k_spec_scale, dim3(blocks_for(nd, 256)), dim3(256), 0, c->stream, nd, c->dim, (const int32_t *)c->spec_of,
                       (const double *)c->spec_w, v);
    return check_launch(c, "k_spec_scale");
}

static int read_flag(lpm_ctx *c, int *any);

// Builds the assembly plan once per pattern (k_asm_plan). Returns false, with nothing built,
// where the plan's 16-bit fields cannot hold the pattern.
static int asm_plan_build(lpm_ctx *c, bool *ok) {
    *ok = false;
    if (c->asm_hops > 0xffff || c->maxconn > 0xffff || c->asm_unplanned) return LPM_OK;
    if (c->pl_row) {
        *ok = true;
        return LPM_OK;
    }
    const int64_t hops = std::max(1, c->asm_hops);
    LPM_HIP(c, hipMalloc(&c->pl_row, sizeof(int32_t) * 2 * std::max<int64_t>(1, c->n)));
    LPM_HIP(c, hipMalloc(&c->pl_slot, sizeof(int32_t) * std::max<int64_t>(1, c->nconn)));
    LPM_HIP(c, hipMalloc(&c->pl_rng, sizeof(int32_t) * std::max<int64_t>(1, c->nconn)));
    LPM_HIP(c, hipMalloc(&c->pl_b3, sizeof(int32_t) * hops * std::max<int64_t>(1, c->n)));
    LPM_HIP(c, hipMalloc(&c->pl_own, sizeof(uint8_t) * hops * std::max<int64_t>(1, c->n)));
    const size_t per = asm_plan_build_lds_bytes(c->asm_hops, c->maxconn);
    if (per > 64 * 1024) return lpm_fail(c, LPM_E_LIMIT, "assembly plan row exceeds 64 KB of LDS");
    const int W = (int)std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per));
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    if (c->n > 0)
        hipLaunchKernelGGL(k_asm_plan, dim3(blocks_for(c->n, W)), dim3(64 * W), per * W, c->stream, c->n, c->asm_hops,
                           c->maxconn, c->bptr, c->bsplit, c->p2, c->cptr, c->cidx, c->cs, c->Kn, c->Tv, c->bdamage,
                           c->pl_row, c->pl_slot, c->pl_rng, c->pl_b3, c->pl_own, c->iflag);
    int rc = check_launch(c, "k_asm_plan");
    int dup = 0;
    if (!rc) rc = read_flag(c, &dup);
    if (rc) return rc;
    if (dup) {  // a partner repeats within a layer: assemble with the row-list kernel
        for (void *p : {(void *)c->pl_row, (void *)c->pl_slot, (void *)c->pl_rng, (void *)c->pl_b3, (void *)c->pl_own})
            hipFree(p);
        c->pl_row = c->pl_slot = c->pl_rng = c->pl_b3 = nullptr;
        c->pl_own = nullptr;
        c->asm_unplanned = true;
        return LPM_OK;
    }
    *ok = true;
    return LPM_OK;
}

int launch_assemble(lpm_ctx *c) {
    if (c->cg_transpose && !c->tslot) return lpm_fail(c, LPM_E_STATE, "transpose map missing");
    bool planned = false;
    if (c->asm_plan) {
        const int rc = asm_plan_build(c, &planned);
        if (rc) return rc;
    }
    if (planned) {
        const int maxb = std::max(1, c->maxbond);
        const size_t per = asm_plan_lds_bytes(c->asm_hops, maxb);
        // rows (wavefronts) per workgroup: at most 4 and at most 64 KB of LDS
        const int wmax = (int)std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per));
        auto launch = [&](int W) -> int {
            const int g = blocks_for(c->n, W);
#define ASM_PLAN_LAUNCH(T, D)                                                                                          \
    hipLaunchKernelGGL((k_assemble_plan<T, D>), dim3(g), dim3(64 * W), per * W, c->stream, c->n, c->asm_hops, maxb,       \
                       c->diag10_literal, c->cg_transpose, c->bptr, c->bsplit, c->p2, c->cptr, c->cidx, c->tslot, c->sp,   \
                       c->pl_row, c->pl_slot, c->pl_rng, c->pl_b3, c->pl_own, c->cs, c->cs_sum, c->Kn, c->Tv,             \
                       c->bdamage, (T *)c->aval, c->diag64)
            if (c->n > 0) {
                if (c->precision == 0) {
                    if (c->dim == 3) ASM_PLAN_LAUNCH(float, 3);
                    else ASM_PLAN_LAUNCH(float, 2);
                } else {
                    if (c->dim == 3) ASM_PLAN_LAUNCH(double, 3);
                    else ASM_PLAN_LAUNCH(double, 2);
                }
            }
#undef ASM_PLAN_LAUNCH
            return check_launch(c, "k_assemble_plan");
        };
        // Launch tuner: every row is computed by one wavefront alone, so each configuration
        // writes the same blocks; the candidates are timed on this assembly (the last launch
        // leaves its result) and the fastest is kept for the context.
        if (c->asm_tune && c->asm_tune_key != c->precision && c->n > 0) {
            hipEvent_t a, b;
            LPM_HIP(c, hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
            if (hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess) {
                hipEventDestroy(a);
                return lpm_fail(c, LPM_E_HIP, "assembly tuner events");
            }
            const int reps = 2;
            double best = 1e300;
            int bw = wmax, rc = LPM_OK;
            for (int W = 1; W <= wmax && !rc; W *= 2) {
                rc = launch(W);  // warm
                if (!rc && hipEventRecord(a, c->stream) != hipSuccess) rc = lpm_fail(c, LPM_E_HIP, "assembly tuner");
                for (int k = 0; k < reps && !rc; ++k) rc = launch(W);
                if (!rc && (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess))
                    rc = lpm_fail(c, LPM_E_HIP, "assembly tuner");
                float ms = 0;
                if (!rc && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = lpm_fail(c, LPM_E_HIP, "assembly tuner");
                if (!rc && ms < best) {
                    best = ms;
                    bw = W;
                }
            }
            hipEventDestroy(a);
            hipEventDestroy(b);
            if (rc) return rc;
            c->asm_wpb = bw;
            c->asm_tuned_ms = best / reps;
            c->asm_tune_key = c->precision;
            return launch(bw);
        }
        return launch(c->asm_wpb > 0 ? std::min(c->asm_wpb, wmax) : wmax);
    }
    // wavefronts per workgroup: as many as fit 64 KB of LDS, at most 4
    const size_t per = asm_lds_bytes(c->asm_hops, c->maxconn);
    const int W = (int)std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per));
    const int g = blocks_for(c->n, W);
    const size_t lds = per * W;
#define ASM_LAUNCH(T, D)                                                                                               \
    hipLaunchKernelGGL((k_assemble<T, D>), dim3(g), dim3(64 * W), lds, c->stream, c->n, c->asm_hops, c->maxconn,          \
                       c->diag10_literal,                                                                                 \
                       c->cg_transpose, c->bptr, c->bsplit, c->p2, c->cptr, c->cidx, c->tslot, c->sp, c->cs, c->cs_sum,  \
                       c->Kn, c->Tv, c->bdamage, (T *)c->aval, c->diag64)
    if (c->precision == 0) {
        if (c->dim == 3) ASM_LAUNCH(float, 3);
        else ASM_LAUNCH(float, 2);
    } else {
        if (c->dim == 3) ASM_LAUNCH(double, 3);
        else ASM_LAUNCH(double, 2);
    }
#undef ASM_LAUNCH
    return check_launch(c, "k_assemble");
}

int launch_dirichlet(lpm_ctx *c) {
    const int64_t nd = c->n * c->dim;
    const int g = blocks_for(nd, 256);
    if (g > c->partial_cap) return lpm_fail(c, LPM_E_LIMIT, "diag grid exceeds partial buffer");
    // elim: n bytes of DoF masks, then n bytes of row marks
    if (!c->elim) LPM_HIP(c, hipMalloc(&c->elim, sizeof(uint8_t) * 2 * std::max<int64_t>(1, c->n)));
    uint8_t *touched = c->elim + std::max<int64_t>(1, c->n);
    hipLaunchKernelGGL(k_diag_norm, dim3(g), dim3(256), 0, c->stream, nd, c->dim, c->diag64, c->cons, c->damage_visual,
                       c->elim, touched, c->partial, c->partial_cap);
    int rc = check_launch(c, "k_diag_norm");
    if (rc) return rc;
    hipLaunchKernelGGL(k_elim_rows, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->elim,
                       touched);
    if ((rc = check_launch(c, "k_elim_rows"))) return rc;
    rc = launch_finalize(c, FIN_DIAG, g);
    if (rc) return rc;
    const int g2 = blocks_for(c->n, 4);
    if (c->precision == 0) {
        float *a = (float *)c->aval;
        if (c->dim == 3)
            hipLaunchKernelGGL((k_dirichlet<float, 3>), dim3(g2), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp, c->elim,
                               touched, c->scal, a);
        else
            hipLaunchKernelGGL((k_dirichlet<float, 2>), dim3(g2), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp, c->elim,
                               touched, c->scal, a);
    } else {
        double *a = (double *)c->aval;
        if (c->dim == 3)
            hipLaunchKernelGGL((k_dirichlet<double, 3>), dim3(g2), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp, c->elim,
                               touched, c->scal, a);
        else
            hipLaunchKernelGGL((k_dirichlet<double, 2>), dim3(g2), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp, c->elim,
                               touched, c->scal, a);
    }
    return check_launch(c, "k_dirichlet");
}

// ---------------------------------------------------------------- finite-difference stiffness
// Stiffness::localStiffnessFD (stiffness.h:103-154): column s of block (i, j) is
// (Pin_i(x + h e_js) - Pin_i(x)) / EPS / radius with h = EPS * radius, every bond's geometry and
// force evaluated at the perturbed positions. Pin_i depends only on the positions of conns(i)
// (1-hop partners and same-layer 2-hop partners: bond geometry, the layer sums and the reverse
// bonds' forces), so the members of one colour class -- no two of them in one conns list -- are
// perturbed together, and one pass per class and axis gives, for every row, column s of the one
// block whose partner moved. Positions are restored exactly after each pass.
__global__ __launch_bounds__(256) void k_fd_move(int64_t m, const int32_t *__restrict__ ids, int s, double h,
                                                  const double *__restrict__ x0, double *__restrict__ xyz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int64_t j = ids[t];
    xyz[3 * j + s] = x0[3 * j + s] + h;  // xyz_temp[r] += EPS * radius (stiffness.h:120)
}

__global__ __launch_bounds__(256) void k_fd_restore(int64_t m, const int32_t *__restrict__ ids, int s,
                                                     const double *__restrict__ x0, double *__restrict__ xyz) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const int64_t j = ids[t];
    xyz[3 * j + s] = x0[3 * j + s];
}

// Row i: the slot whose partner belongs to class `col` receives column s of its block,
// K[r][s] = (Pin_i[r] - P0_i[r]) / EPS / radius (stiffness.h:134-135, 147-150), stored like
// k_assemble stores the analytic block (in the transposed position when `transpose`).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_fd_column(int64_t n, int col, int s, int transpose, double radius,
                                                    const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                                    const int32_t *__restrict__ color, const int32_t *__restrict__ tslot,
                                                    const int64_t *__restrict__ sp, const double *__restrict__ Pin,
                                                    const double *__restrict__ P0, T *__restrict__ aval,
                                                    double *__restrict__ diag64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c0 = cptr[i], nci = cptr[i + 1] - c0;
    int slot = -1;
    for (int q = 0; q < nci; ++q)
        if (color[cidx[c0 + q]] == col) {
            slot = q;
            break;
        }
    if (slot < 0) return;
    const int64_t j = cidx[c0 + slot];
    double K[3];
    for (int r = 0; r < D; ++r) K[r] = (Pin[3 * i + r] - P0[3 * i + r]) / LPM_EPS / radius;
    if (!transpose) {
        for (int r = 0; r < D; ++r) aval[sell_off(sp, i, slot, D * D, r * D + s)] = (T)K[r];
    } else {
        const int ts = tslot[c0 + slot];
        for (int r = 0; r < D; ++r) aval[sell_off(sp, j, ts, D * D, s * D + r)] = (T)K[r];
    }
    if (j == i) diag64[(int64_t)D * i + s] = K[s];
}

void fd_free(lpm_ctx *c) {
    for (void *p : {(void *)c->fd_color, (void *)c->fd_ids, (void *)c->fd_x0, (void *)c->fd_P0})
        if (p) hipFree(p);
    c->fd_color = c->fd_ids = nullptr;
    c->fd_x0 = c->fd_P0 = nullptr;
    c->fd_ptr.clear();
    c->fd_ncolor = 0;
}

// Greedy colouring in id order: particle k takes the smallest class not used by any particle
// that shares a conns list with it (the members of conns(i) for every i in conns(k); conns is
// symmetric). Built once: the pattern never changes.
static int fd_prepare(lpm_ctx *c) {
    if (c->fd_color) return LPM_OK;
    const int64_t n = c->n;
    std::vector<int32_t> cp(n + 1), ci(c->nconn);
    LPM_HIP(c, hipMemcpyAsync(cp.data(), c->cptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(ci.data(), c->cidx, sizeof(int32_t) * c->nconn, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> color(n, -1);
    std::vector<int64_t> seen;  // seen[class] == k: class used next to particle k
    int ncol = 0;
    for (int64_t k = 0; k < n; ++k) {
        for (int32_t a = cp[k]; a < cp[k + 1]; ++a) {
            const int64_t i = ci[a];
            for (int32_t b = cp[i]; b < cp[i + 1]; ++b) {
                const int32_t q = color[ci[b]];
                if (q >= 0) seen[q] = k;
            }
        }
        int32_t q = 0;
        while (q < ncol && seen[q] == k) ++q;
        if (q == ncol) {
            ++ncol;
            seen.push_back(-1);
        }
        color[k] = q;
    }
    std::vector<int64_t> ptr(ncol + 1, 0);
    for (int64_t k = 0; k < n; ++k) ++ptr[color[k] + 1];
    for (int q = 0; q < ncol; ++q) ptr[q + 1] += ptr[q];
    std::vector<int32_t> ids(n);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < n; ++k) ids[fill[color[k]]++] = (int32_t)k;
    LPM_HIP(c, hipMalloc(&c->fd_color, sizeof(int32_t) * n));
    LPM_HIP(c, hipMalloc(&c->fd_ids, sizeof(int32_t) * n));
    LPM_HIP(c, hipMalloc(&c->fd_x0, sizeof(double) * 3 * n));
    LPM_HIP(c, hipMalloc(&c->fd_P0, sizeof(double) * 3 * n));
    LPM_HIP(c, hipMemcpyAsync(c->fd_color, color.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->fd_ids, ids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));  // the host vectors go out of scope
    c->fd_ptr = ptr;
    c->fd_ncolor = ncol;
    return LPM_OK;
}

int launch_assemble_fd(lpm_ctx *c) {
    // ParticleJ2Plasticity's FD pass runs the return mapping (particle_j2plasticity.h:78-140) on
    // the stress of the previous force evaluation while the bond geometry is perturbed, so its
    // quotient carries the whole plastic increment; not provided.
    if (c->model == LPM_J2PLASTICITY)
        return lpm_fail(c, LPM_E_ARG, "finite-difference stiffness is not provided for J2Plasticity");
    if (c->cg_transpose && !c->tslot) return lpm_fail(c, LPM_E_STATE, "transpose map missing");
    int rc = fd_prepare(c);
    if (rc) return rc;
    const int64_t n = c->n;
    const size_t n3 = sizeof(double) * 3 * n;
    LPM_HIP(c, hipMemcpyAsync(c->fd_x0, c->xyz, n3, hipMemcpyDeviceToDevice, c->stream));
    // base forces at the current positions (Pin_temp, stiffness.h:114-116)
    rc = launch_geometry_force(c, true);
    if (!rc) rc = launch_particle_force(c);
    if (rc) return rc;
    LPM_HIP(c, hipMemcpyAsync(c->fd_P0, c->Pin, n3, hipMemcpyDeviceToDevice, c->stream));
    const double h = LPM_EPS * c->radius;
    for (int q = 0; q < c->fd_ncolor; ++q) {
        const int64_t m = c->fd_ptr[q + 1] - c->fd_ptr[q];
        const int32_t *ids = c->fd_ids + c->fd_ptr[q];
        for (int s = 0; s < c->dim; ++s) {
            hipLaunchKernelGGL(k_fd_move, dim3(blocks_for(m, 256)), dim3(256), 0, c->stream, m, ids, s, h, c->fd_x0, c->xyz);
            if ((rc = check_launch(c, "k_fd_move"))) return rc;
            // updateBondsGeometry, updateBondsForce and updateParticleForce (stiffness.h:122-132);
            // the state update between them changes no force of these models
            if ((rc = launch_geometry_force(c, true))) return rc;
            if ((rc = launch_particle_force(c))) return rc;
#define FD_COL(T, D)                                                                                                   \
    hipLaunchKernelGGL((k_fd_column<T, D>), dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, q, s, c->cg_transpose, \
                       c->radius, c->cptr, c->cidx, c->fd_color, c->tslot, c->sp, c->Pin, c->fd_P0, (T *)c->aval, c->diag64)
            if (c->precision == 0) {
                if (c->dim == 3) FD_COL(float, 3);
                else FD_COL(float, 2);
            } else {
                if (c->dim == 3) FD_COL(double, 3);
                else FD_COL(double, 2);
            }
#undef FD_COL
            if ((rc = check_launch(c, "k_fd_column"))) return rc;
            hipLaunchKernelGGL(k_fd_restore, dim3(blocks_for(m, 256)), dim3(256), 0, c->stream, m, ids, s, c->fd_x0, c->xyz);
            if ((rc = check_launch(c, "k_fd_restore"))) return rc;
        }
    }
    // resumeParticleState (particle.h:183-195) leaves every particle it visits -- all of them --
    // with damage, state variables and bond state reset to the last converged copies and the
    // bond forces re-evaluated
    const size_t D = sizeof(double), nb = c->nbond;
    LPM_HIP(c, hipMemcpyAsync(c->damage, c->damage_last, n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->sv, c->sv_last, LPM_NSV * n * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->bdamage, c->bdamage_last, nb * D, hipMemcpyDeviceToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->dLp, c->dLp_last, nb * D, hipMemcpyDeviceToDevice, c->stream));
    rc = launch_geometry_force(c, true);
    if (!rc) rc = launch_particle_force(c);
    return rc;
}

int launch_state_var(lpm_ctx *c) {
    if (c->model == 0) return LPM_OK;  // ParticleElastic has no state update (particle.h:62)
    const bool rec = c->fat_rec && c->model == HCF;
    if (rec) {
        const int rc = fat_record_state(c, false);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_state_var, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->model, c->volume,
                       c->nneighbors, c->bptr, c->bsplit, c->Kn, c->dLe, c->dLe_total, c->TdLe_total, c->stress, c->cs,
                       c->dis, c->mat, c->damage, c->sv_last, c->sv, c->dLp, c->Ddot_local);
    const int rc = check_launch(c, "k_state_var");
    return (rc || !rec) ? rc : fat_record_state(c, true);
}

int launch_nonlocal_ddot(lpm_ctx *c, int undamaged) {
    hipLaunchKernelGGL(k_nonlocal_ddot, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, undamaged, c->volume,
                       c->nneighbors, c->type, c->nlptr, c->nlidx, c->bptr, c->xyz, c->nonlocal_L, c->Ddot_local,
                       c->Ddot_nonlocal);
    return check_launch(c, "k_nonlocal_ddot");
}

static int read_flag(lpm_ctx *c, int *any) {
    LPM_HIP(c, hipMemcpyAsync(c->h_iflag, c->iflag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    *any = c->h_iflag[0] != 0;
    return LPM_OK;
}

int launch_static_damage(lpm_ctx *c, int undamaged, int *any) {
    int rc = launch_nonlocal_ddot(c, undamaged);
    if (rc) return rc;
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_static_damage, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->model, undamaged,
                       c->type, c->bptr, c->mat, c->bstrain, c->bdamage, c->Ddot_nonlocal, c->damage_last, c->damage,
                       c->iflag);
    rc = check_launch(c, "k_static_damage");
    if (rc) return rc;
    if (any) return read_flag(c, any);
    // the caller reads the flag later (flag_wait), after queueing more work behind it
    LPM_HIP(c, hipMemcpyAsync(c->h_iflag, c->iflag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    return LPM_OK;
}

int flag_wait(lpm_ctx *c, int *any) {
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    *any = c->h_iflag[0] != 0;
    return LPM_OK;
}

int launch_fatigue_damage(lpm_ctx *c, double dNdt, int undamaged, int *any) {
    int rc = launch_nonlocal_ddot(c, undamaged);
    if (rc) return rc;
    if (c->model == HCF) {
        // what the fatigue-parameter sensitivity needs of this update
        if (!c->fat_xyz) LPM_HIP(c, hipMalloc(&c->fat_xyz, sizeof(double) * 3 * c->n));
        if (!c->fat_range) LPM_HIP(c, hipMalloc(&c->fat_range, sizeof(double) * c->n));
        LPM_HIP(c, hipMemcpyAsync(c->fat_xyz, c->xyz, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL(k_energy_range, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->sv, c->fat_range);
        if ((rc = check_launch(c, "k_energy_range"))) return rc;
        c->fat_dt = dNdt;
        c->fat_undamaged = undamaged;
        c->fat_valid = true;
    }
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_fatigue_damage, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->model, undamaged,
                       dNdt, c->type, c->mat, c->Ddot_nonlocal, c->damage_last, c->damage, c->sv, c->iflag);
    rc = check_launch(c, "k_fatigue_damage");
    if (rc) return rc;
    if (c->fat_rec && c->model == HCF && (rc = fat_record_update(c, dNdt, undamaged))) return rc;
    return read_flag(c, any);
}

// Device bond search and initial bond state for lpm_create (positions already in c->xyz).
// Sets c->nbond, c->maxbond, c->gp_bonds, c->asm_hops and allocates the topology arrays.
int device_topology(lpm_ctx *c, const double *xyz_host, const lpm_cut *cuts_host, int ncut) {
    const int64_t n = c->n;
    const double c1 = 1.01 * c->cutoff[1], c0 = 1.01 * c->cutoff[0];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], xyz_host[3 * i + k]);
            hi[k] = std::max(hi[k], xyz_host[3 * i + k]);
        }
    const double h = c1 * 1.0001;  // a cell edge above the cutoff: partners lie in adjacent cells
    int64_t nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = std::max<int64_t>(1, (int64_t)std::floor((hi[k] - lo[k]) / h) + 1);
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    if (ncell >= 0x7fffffffLL) return lpm_fail(c, LPM_E_LIMIT, "bond search grid too large");
    int32_t *pcell = nullptr, *count = nullptr, *head = nullptr, *items = nullptr, *tmp0 = nullptr, *tmp1 = nullptr,
            *cnt = nullptr, *flag = nullptr;
    lpm_cut *dcuts = nullptr;
    auto cleanup = [&]() {
        void *ps[] = {pcell, count, head, items, tmp0, tmp1, cnt, flag, dcuts};
        for (void *q : ps)
            if (q) hipFree(q);
    };
    auto al = [&](int32_t **p, int64_t cntv) { return hipMalloc(p, sizeof(int32_t) * std::max<int64_t>(cntv, 1)) == hipSuccess; };
    if (!al(&pcell, n) || !al(&count, ncell) || !al(&head, ncell + 1) || !al(&items, n) || !al(&tmp0, n * LPM_MAXB) ||
        !al(&tmp1, n * LPM_MAXB) || !al(&cnt, 2 * n) || !al(&flag, 4)) {
        cleanup();
        return lpm_fail(c, LPM_E_NOMEM, "bond search scratch");
    }
    if (ncut > 0) {
        if (hipMalloc(&dcuts, sizeof(lpm_cut) * ncut) != hipSuccess) { cleanup(); return lpm_fail(c, LPM_E_NOMEM, "cuts"); }
        hipMemcpyAsync(dcuts, cuts_host, sizeof(lpm_cut) * ncut, hipMemcpyHostToDevice, c->stream);
    }
    int rc = LPM_OK;
    std::vector<int32_t> h_count(ncell), h_head(ncell + 1, 0), h_cnt(2 * n);
    hipMemsetAsync(count, 0, sizeof(int32_t) * ncell, c->stream);
    hipMemsetAsync(flag, 0, sizeof(int32_t) * 4, c->stream);
    hipLaunchKernelGGL(k_cell_count, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const double *)c->xyz, lo[0],
                       lo[1], lo[2], h, nc[0], nc[1], nc[2], pcell, count);
    if ((rc = check_launch(c, "k_cell_count"))) { cleanup(); return rc; }
    hipMemcpyAsync(h_count.data(), count, sizeof(int32_t) * ncell, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { cleanup(); return lpm_fail(c, LPM_E_HIP, "bond search sync"); }
    for (int64_t q = 0; q < ncell; ++q) h_head[q + 1] = h_head[q] + h_count[q];
    hipMemcpyAsync(head, h_head.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice, c->stream);
    hipMemsetAsync(count, 0, sizeof(int32_t) * ncell, c->stream);
    hipLaunchKernelGGL(k_cell_fill, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const int32_t *)pcell,
                       (const int32_t *)head, count, items);
    if ((rc = check_launch(c, "k_cell_fill"))) { cleanup(); return rc; }
    hipLaunchKernelGGL(k_bond_search, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const double *)c->xyz,
                       (const int32_t *)pcell, nc[0], nc[1], nc[2], (const int32_t *)head, (const int32_t *)items, c0, c1,
                       (const lpm_cut *)dcuts, ncut, tmp0, tmp1, cnt, cnt + n, flag);
    if ((rc = check_launch(c, "k_bond_search"))) { cleanup(); return rc; }
    int32_t h_flag[4] = {0, 0, 0, 0};
    hipMemcpyAsync(h_cnt.data(), cnt, sizeof(int32_t) * 2 * n, hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(h_flag, flag, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { cleanup(); return lpm_fail(c, LPM_E_HIP, "bond search sync"); }
    if (h_flag[0]) { cleanup(); return lpm_fail(c, LPM_E_LIMIT, "a particle has more than 32 bonds"); }
    std::vector<int32_t> bp(n + 1, 0), bs(n);
    int64_t tot = 0;
    int maxb = 0;
    for (int64_t i = 0; i < n; ++i) {
        bp[i] = (int32_t)tot;
        bs[i] = (int32_t)(tot + h_cnt[i]);
        tot += h_cnt[i] + h_cnt[n + i];
        maxb = std::max(maxb, h_cnt[i] + h_cnt[n + i]);
        if (tot > 0x7fffffffLL) { cleanup(); return lpm_fail(c, LPM_E_LIMIT, "more than 2^31 bonds"); }
    }
    bp[n] = (int32_t)tot;
    c->nbond = tot;
    c->maxbond = maxb;
    c->gp_bonds = 0;
    for (int64_t p0 = 0; p0 < n; p0 += LPM_BOND_GP)
        c->gp_bonds = std::max<int>(c->gp_bonds, bp[std::min<int64_t>(n, p0 + LPM_BOND_GP)] - bp[p0]);
#define ALT(ptr, cntv)                                                                                          \
    if (hipMalloc(&c->ptr, sizeof(*c->ptr) * std::max<int64_t>((cntv), 1)) != hipSuccess) {                     \
        cleanup();                                                                                               \
        return lpm_fail(c, LPM_E_NOMEM, "topology arrays");                                                      \
    }
    ALT(bptr, n + 1); ALT(bsplit, n); ALT(p2, tot); ALT(opp, tot); ALT(nbarr, n);
    ALT(dis0, tot); ALT(dis, tot); ALT(dis_last, tot); ALT(cs, 3 * tot);
#undef ALT
    hipMemcpyAsync(c->bptr, bp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(c->bsplit, bs.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream);
    hipLaunchKernelGGL(k_bond_compact, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const int32_t *)c->bptr,
                       (const int32_t *)c->bsplit, (const int32_t *)tmp0, (const int32_t *)tmp1, c->p2);
    if ((rc = check_launch(c, "k_bond_compact"))) { cleanup(); return rc; }
    hipMemsetAsync(flag, 0, sizeof(int32_t) * 4, c->stream);
    hipLaunchKernelGGL(k_bond_init, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const int32_t *)c->bptr,
                       (const int32_t *)c->bsplit, (const int32_t *)c->p2, (const double *)c->xyz, c->nbarr, c->opp,
                       c->dis0, c->dis, c->dis_last, c->cs, flag + 2, flag);
    if ((rc = check_launch(c, "k_bond_init"))) { cleanup(); return rc; }
    hipMemcpyAsync(h_flag, flag, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { cleanup(); return lpm_fail(c, LPM_E_HIP, "bond init sync"); }
    cleanup();
    if (h_flag[1]) return lpm_fail(c, LPM_E_TOPO, "a bond has no reverse bond");
    c->asm_hops = h_flag[2];
    return LPM_OK;
}

int launch_pattern(lpm_ctx *c) {
    int32_t *counts = nullptr;
    LPM_HIP(c, hipMalloc(&counts, sizeof(int32_t) * (c->n + 1)));
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    if (c->n > 0x7fffffff) return lpm_fail(c, LPM_E_LIMIT, "too many particles for the pattern grid");
    hipLaunchKernelGGL(k_pattern, dim3((unsigned)c->n), dim3(256), 0, c->stream, c->n, c->bptr, c->bsplit, c->p2, counts,
                       (const int32_t *)nullptr, (int32_t *)nullptr, c->iflag);
    int rc = check_launch(c, "k_pattern(count)");
    if (rc) { hipFree(counts); return rc; }
    std::vector<int32_t> h(c->n);
    LPM_HIP(c, hipMemcpyAsync(h.data(), counts, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    int any = 0;
    rc = read_flag(c, &any);
    if (rc) { hipFree(counts); return rc; }
    if (any) { hipFree(counts); return lpm_fail(c, LPM_E_LIMIT, "a particle has more than 512 two-hop candidates"); }
    std::vector<int32_t> cp(c->n + 1, 0);
    int maxc = 0;
    for (int64_t i = 0; i < c->n; ++i) {
        cp[i + 1] = cp[i] + h[i];
        if (h[i] > maxc) maxc = h[i];
    }
    c->nconn = cp[c->n];
    c->maxconn = maxc;
    LPM_HIP(c, hipMalloc(&c->cptr, sizeof(int32_t) * (c->n + 1)));
    LPM_HIP(c, hipMalloc(&c->cidx, sizeof(int32_t) * (c->nconn > 0 ? c->nconn : 1)));
    LPM_HIP(c, hipMemcpyAsync(c->cptr, cp.data(), sizeof(int32_t) * (c->n + 1), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_pattern, dim3((unsigned)c->n), dim3(256), 0, c->stream, c->n, c->bptr, c->bsplit, c->p2, counts,
                       (const int32_t *)c->cptr, c->cidx, c->iflag);
    rc = check_launch(c, "k_pattern(fill)");
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(counts);
    if (rc) return rc;
    // SELL-64 slice offsets: slice k covers rows 64k..64k+63 and is as wide as its longest row
    c->nslice = (c->n + 63) / 64;
    std::vector<int64_t> sp(c->nslice + 1, 0);
    for (int64_t k = 0; k < c->nslice; ++k) {
        int W = 0;
        for (int64_t i = 64 * k; i < std::min<int64_t>(c->n, 64 * k + 64); ++i) W = std::max(W, h[i]);
        sp[k + 1] = sp[k] + 64LL * W;
    }
    c->nslots = sp[c->nslice];
    LPM_HIP(c, hipMalloc(&c->sp, sizeof(int64_t) * (c->nslice + 1)));
    LPM_HIP(c, hipMalloc(&c->scol, sizeof(int32_t) * (c->nslots > 0 ? c->nslots : 1)));
    LPM_HIP(c, hipMemcpyAsync(c->sp, sp.data(), sizeof(int64_t) * (c->nslice + 1), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_sell_cols, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp, c->scol);
    rc = check_launch(c, "k_sell_cols");
    if (rc) return rc;
    // int16 offsets for the SpMV when every column lies within 32k rows of its row
    LPM_HIP(c, hipMalloc(&c->scol16, sizeof(int16_t) * (c->nslots > 0 ? c->nslots : 1)));
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_sell_cols16, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp,
                       c->scol16, c->iflag);
    rc = check_launch(c, "k_sell_cols16");
    if (rc) return rc;
    int32_t bad = 0;
    LPM_HIP(c, hipMemcpyAsync(&bad, c->iflag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    if (bad) {
        hipFree(c->scol16);
        c->scol16 = nullptr;
    }
    return LPM_OK;
}

int launch_tslot(lpm_ctx *c) {
    if (c->tslot) return LPM_OK;
    int32_t *rowof = nullptr;
    LPM_HIP(c, hipMalloc(&rowof, sizeof(int32_t) * c->nconn));
    LPM_HIP(c, hipMalloc(&c->tslot, sizeof(int32_t) * c->nconn));
    hipLaunchKernelGGL(k_rowof, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, rowof);
    LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_tslot, dim3(blocks_for(c->nconn, 256)), dim3(256), 0, c->stream, c->nconn, c->n, c->cptr, c->cidx,
                       rowof, c->tslot, c->iflag);
    int rc = check_launch(c, "k_tslot");
    int bad = 0;
    if (!rc) rc = read_flag(c, &bad);
    hipFree(rowof);
    if (rc) return rc;
    if (bad) return lpm_fail(c, LPM_E_TOPO, "sparsity pattern is not symmetric; K^T cannot be stored");
    return LPM_OK;
}

// Cell-grid nonlocal search (mode nl_cells). Sets *done when it built the lists; leaves them
// to the all-pairs search (*done false, nothing changed) when a coordinate or the largest cutoff
// is not finite, the grid would need 2^31 cells, or a list exceeds the sort's NL_SORT_MAX.
constexpr int NL_SORT_MAX = 4096;

static int nonlocal_search_cells(lpm_ctx *c, double ratio, bool *done) {
    *done = false;
    const int64_t n = c->n;
    std::vector<double> hx(3 * n), hL(n);
    LPM_HIP(c, hipMemcpyAsync(hx.data(), c->xyz, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(hL.data(), c->nonlocal_L, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, cutmax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            const double v = hx[3 * i + k];
            if (!std::isfinite(v)) return LPM_OK;
            lo[k] = std::min(lo[k], v);
            hi[k] = std::max(hi[k], v);
        }
        const double cut = ratio * hL[i];  // a NaN cutoff accepts no partner and does not widen the grid
        if (cut > cutmax) cutmax = cut;
    }
    if (!std::isfinite(cutmax)) return LPM_OK;
    // a cell edge above every cutoff: partners lie in adjacent cells; at most max(4 n, 65536)
    // cells (coarser cells only add candidates)
    double h = cutmax > 0.0 ? cutmax * (1.0 + 1e-6) : 1.0;
    const int64_t layers = c->nspec > 0 ? c->nspec : 1, cap = std::max<int64_t>(4 * n, 65536);
    int64_t nc[3], ncell = 0;
    for (;;) {
        double cells = (double)layers;
        for (int k = 0; k < 3; ++k) {
            nc[k] = std::max<int64_t>(1, (int64_t)std::floor((hi[k] - lo[k]) / h) + 1);
            cells *= (double)nc[k];
        }
        if (cells <= (double)cap) {
            ncell = layers * nc[0] * nc[1] * nc[2];
            break;
        }
        h *= 1.5;
    }
    if (ncell >= 0x7fffffffLL) return LPM_OK;
    int32_t *pcell = nullptr, *count = nullptr, *head = nullptr, *items = nullptr, *counts = nullptr;
    auto cleanup = [&]() {
        void *ps[] = {pcell, count, head, items, counts};
        for (void *q : ps)
            if (q) hipFree(q);
    };
    auto al = [&](int32_t **q, int64_t m) { return hipMalloc(q, sizeof(int32_t) * std::max<int64_t>(m, 1)) == hipSuccess; };
    if (!al(&pcell, n) || !al(&count, ncell) || !al(&head, ncell + 1) || !al(&items, n) || !al(&counts, n)) {
        cleanup();
        return lpm_fail(c, LPM_E_NOMEM, "nonlocal search scratch");
    }
    int rc = LPM_OK;
    auto hip_ok = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        rc = lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
        return false;
    };
    std::vector<int32_t> h_count(ncell), h_head(ncell + 1, 0), h_cnt(n);
    std::vector<int32_t> p(n + 1, 0);
    int maxlen = 0;
    if (!hip_ok(hipMemsetAsync(count, 0, sizeof(int32_t) * ncell, c->stream), "nonlocal search")) { cleanup(); return rc; }
    hipLaunchKernelGGL(k_nl_bin, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const double *)c->xyz, lo[0], lo[1],
                       lo[2], h, nc[0], nc[1], nc[2], (const int32_t *)(c->nspec > 0 ? c->spec_of : nullptr), pcell, count);
    if ((rc = check_launch(c, "k_nl_bin"))) { cleanup(); return rc; }
    if (!hip_ok(hipMemcpyAsync(h_count.data(), count, sizeof(int32_t) * ncell, hipMemcpyDeviceToHost, c->stream), "nonlocal search") ||
        !hip_ok(hipStreamSynchronize(c->stream), "nonlocal search")) { cleanup(); return rc; }
    for (int64_t q = 0; q < ncell; ++q) h_head[q + 1] = h_head[q] + h_count[q];
    if (!hip_ok(hipMemcpyAsync(head, h_head.data(), sizeof(int32_t) * (ncell + 1), hipMemcpyHostToDevice, c->stream), "nonlocal search") ||
        !hip_ok(hipMemsetAsync(count, 0, sizeof(int32_t) * ncell, c->stream), "nonlocal search")) { cleanup(); return rc; }
    hipLaunchKernelGGL(k_cell_fill, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const int32_t *)pcell,
                       (const int32_t *)head, count, items);
    if ((rc = check_launch(c, "k_cell_fill"))) { cleanup(); return rc; }
    hipLaunchKernelGGL((k_nl_cells<false>), dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, ratio, (const double *)c->xyz,
                       (const double *)c->nonlocal_L, (const int32_t *)pcell, nc[0], nc[1], nc[2], (const int32_t *)head,
                       (const int32_t *)items, counts, (const int32_t *)nullptr, (int32_t *)nullptr);
    if ((rc = check_launch(c, "k_nl_cells(count)"))) { cleanup(); return rc; }
    if (!hip_ok(hipMemcpyAsync(h_cnt.data(), counts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream), "nonlocal search") ||
        !hip_ok(hipStreamSynchronize(c->stream), "nonlocal search")) { cleanup(); return rc; }
    for (int64_t i = 0; i < n; ++i) {
        if ((int64_t)p[i] + h_cnt[i] > 0x7fffffff) { cleanup(); return lpm_fail(c, LPM_E_LIMIT, "nonlocal list too long"); }
        p[i + 1] = p[i] + h_cnt[i];
        maxlen = std::max(maxlen, h_cnt[i]);
    }
    if (maxlen > NL_SORT_MAX) {
        cleanup();
        return LPM_OK;
    }
    if (c->nlptr) hipFree(c->nlptr);
    if (c->nlidx) hipFree(c->nlidx);
    c->nlptr = c->nlidx = nullptr;
    c->nnl = p[n];
    if (!hip_ok(hipMalloc(&c->nlptr, sizeof(int32_t) * (n + 1)), "nonlocal lists") ||
        !hip_ok(hipMalloc(&c->nlidx, sizeof(int32_t) * (c->nnl > 0 ? c->nnl : 1)), "nonlocal lists") ||
        !hip_ok(hipMemcpyAsync(c->nlptr, p.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, c->stream), "nonlocal lists")) {
        cleanup();
        return rc;
    }
    hipLaunchKernelGGL((k_nl_cells<true>), dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, ratio, (const double *)c->xyz,
                       (const double *)c->nonlocal_L, (const int32_t *)pcell, nc[0], nc[1], nc[2], (const int32_t *)head,
                       (const int32_t *)items, counts, (const int32_t *)c->nlptr, c->nlidx);
    if ((rc = check_launch(c, "k_nl_cells(fill)"))) { cleanup(); return rc; }
    int P = 2;
    while (P < maxlen) P <<= 1;
    hipLaunchKernelGGL(k_nl_sort, dim3(n), dim3(64), sizeof(int32_t) * P, c->stream, (const int32_t *)c->nlptr, c->nlidx);
    if ((rc = check_launch(c, "k_nl_sort"))) { cleanup(); return rc; }
    hip_ok(hipStreamSynchronize(c->stream), "nonlocal search");
    cleanup();
    if (!rc) *done = true;
    return rc;
}

int launch_nonlocal_search(lpm_ctx *c, double ratio) {
    if (c->nl_cells && c->n > 0) {
        bool done = false;
        const int rc = nonlocal_search_cells(c, ratio, &done);
        if (rc || done) return rc;
    }
    int32_t *counts = nullptr;
    LPM_HIP(c, hipMalloc(&counts, sizeof(int32_t) * (c->n + 1)));
    hipLaunchKernelGGL(k_nonlocal_search, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, ratio, c->xyz,
                       c->nonlocal_L, (const int32_t *)c->spec_of, counts, (const int32_t *)nullptr, (int32_t *)nullptr);
    int rc = check_launch(c, "k_nonlocal_search(count)");
    if (rc) { hipFree(counts); return rc; }
    std::vector<int32_t> h(c->n);
    LPM_HIP(c, hipMemcpyAsync(h.data(), counts, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> p(c->n + 1, 0);
    for (int64_t i = 0; i < c->n; ++i) {
        if ((int64_t)p[i] + h[i] > 0x7fffffff) { hipFree(counts); return lpm_fail(c, LPM_E_LIMIT, "nonlocal list too long"); }
        p[i + 1] = p[i] + h[i];
    }
    if (c->nlptr) hipFree(c->nlptr);
    if (c->nlidx) hipFree(c->nlidx);
    c->nlptr = c->nlidx = nullptr;
    c->nnl = p[c->n];
    LPM_HIP(c, hipMalloc(&c->nlptr, sizeof(int32_t) * (c->n + 1)));
    LPM_HIP(c, hipMalloc(&c->nlidx, sizeof(int32_t) * (c->nnl > 0 ? c->nnl : 1)));
    LPM_HIP(c, hipMemcpyAsync(c->nlptr, p.data(), sizeof(int32_t) * (c->n + 1), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_nonlocal_search, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, ratio, c->xyz,
                       c->nonlocal_L, (const int32_t *)c->spec_of, counts, (const int32_t *)c->nlptr, c->nlidx);
    rc = check_launch(c, "k_nonlocal_search(fill)");
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(counts);
    return rc;
}

extern "C" int lpm_cut_particles(const double *xyz, int64_t n, const lpm_cut *cuts, int32_t ncut, uint8_t *keep) {
    if (n < 0 || (n > 0 && (!xyz || !keep)) || ncut < 0 || (ncut > 0 && !cuts)) return LPM_E_ARG;
    for (int32_t k = 0; k < ncut; ++k)
        if (cuts[k].kind < LPM_CUT_CIRCLE || cuts[k].kind > LPM_CUT_REVERSE_NOTCH) return LPM_E_ARG;
    for (int64_t i = 0; i < n; ++i) keep[i] = cut_point_valid(xyz[3 * i], xyz[3 * i + 1], cuts, ncut) ? 1 : 0;
    return LPM_OK;
}
