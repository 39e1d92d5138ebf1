// Two-level preconditioner for the CG: Jacobi plus an additive coarse-space correction,
//   M^-1 r = D^-1 r + P Ac^-1 P^T r,   Ac = P^T K P,
// where the columns of P are the rigid-body modes (translations and infinitesimal rotations
// about the aggregate centre) of cubic aggregates of particles, with the rows of eliminated
// DoFs set to zero. The Jacobi part alone lets the CG iteration count grow with the lattice
// width; the coarse space removes the smooth, nearly rigid components that cause that growth.
// It changes only the rate of the inner linear solve: the Newton test and the equilibrium it
// converges to are the same as with Jacobi.
//
// Setup (once per context, the pattern never changes): aggregates from the initial positions,
// and for every coarse block (g, h) the list of row entries that contribute to it.
// Per assembly: contributions P_i^T K_ij P_j summed per (row i, aggregate h) in slot order,
// then per coarse block in row order (no atomics: the result does not depend on scheduling),
// then a blocked Gauss-Jordan inversion of Ac (32 pivots per step, captured in a graph). Ac is
// block diagonal with one block per specimen of an ensemble (aggregates never span two
// specimens), so every specimen gets a coarse space of its own size and all blocks are inverted
// in the same launches.
// Per CG iteration: one workgroup per aggregate forms P^T r, one small kernel forms Ac^-1 P^T r;
// the prolongation is fused into the CG's preconditioner step (kernels_cg.hip).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "lpm_internal.h"
#include "reduce.h"

namespace {

__global__ __launch_bounds__(256) void k_dofmask(int64_t n, int dim, const int32_t *__restrict__ cons,
                                                  const double *__restrict__ damage_visual, uint8_t *__restrict__ mask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned m = 0;
    const bool dead = fabs(damage_visual[i] - 1.0) < LPM_EPS;
    for (int k = 0; k < dim; ++k)
        if (cons[3 * i + k] == 1 || dead) m |= 1u << k;
    mask[i] = (uint8_t)m;
}

// Stage 1, one thread per row i: for each aggregate h among the columns of row i (in the
// row's aggregate list), V = P_i^T (sum over slots with agg(j) = h, in slot order, of K_ij P_j).
// With P_j = [I | R(d_j)], K_ij P_j = [K_ij | K_ij R(d_j)]. Column indices come from the SELL
// copy (scol: slot c of row i at sp[i/64] + 64 c + i%64), which the lanes of a wavefront (the
// 64 rows of a slice) read coalesced, where the CSR list would put every lane on its own line.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_coarse_rows(int64_t n, const int32_t *__restrict__ cptr,
                                                      const int32_t *__restrict__ scol, const int64_t *__restrict__ sp,
                                                      const T *__restrict__ aval, const int32_t *__restrict__ agg_of,
                                                      const double *__restrict__ agg_d, const uint8_t *__restrict__ mask,
                                                      const int32_t *__restrict__ ra_ptr, const int32_t *__restrict__ ra_id,
                                                      double *__restrict__ vbuf) {
    constexpr int EB = D * D, M = D == 3 ? 6 : 3;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nc = cptr[i + 1] - cptr[i];
    const int32_t *ci = scol + sp[i >> 6] + (i & 63);
    double Pi[3][6];
    rigid_modes<D>(&agg_d[D * i], mask[i], Pi);
    for (int f = ra_ptr[i]; f < ra_ptr[i + 1]; ++f) {
        const int h = ra_id[f];
        double W[D][M];
        for (int r = 0; r < D; ++r)
            for (int m = 0; m < M; ++m) W[r][m] = 0.0;
        for (int c = 0; c < nc; ++c) {
            const int64_t j = ci[(int64_t)c * 64];
            if (agg_of[j] != h) continue;
            double Pj[3][6];
            rigid_modes<D>(&agg_d[D * j], mask[j], Pj);
            double Kb[D][D];
            for (int r = 0; r < D; ++r)
                for (int q = 0; q < D; ++q) Kb[r][q] = (double)aval[sell_off(sp, i, c, EB, r * D + q)];
            for (int r = 0; r < D; ++r)
                for (int m = 0; m < M; ++m) {
                    double t = 0.0;
                    for (int q = 0; q < D; ++q) t += Kb[r][q] * Pj[q][m];
                    W[r][m] += t;
                }
        }
        double *out = vbuf + (int64_t)f * M * M;
        for (int a = 0; a < M; ++a)
            for (int b = 0; b < M; ++b) {
                double t = 0.0;
                for (int r = 0; r < D; ++r) t += Pi[r][a] * W[r][b];
                out[a * M + b] = t;
            }
    }
}

// Stage 2, one wavefront per coarse block q = (g, h): lanes take the block's row entries in a
// fixed interleave, each sums its entries' M x M contributions, and a fixed shuffle tree adds
// the lanes; Ac[g M + a][h M + b] is then the sum over the block's entries.
template <int M>
__global__ __launch_bounds__(256) void k_coarse_blocks(int64_t ncq, const int32_t *__restrict__ cq_ptr,
                                                        const int32_t *__restrict__ cq_f, const int32_t *__restrict__ cq_gh,
                                                        const double *__restrict__ vbuf, const int32_t *__restrict__ agg_grp,
                                                        const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                        double *__restrict__ Ac) {
    constexpr int MM = M * M;
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= ncq) return;
    double acc[MM];
    for (int e = 0; e < MM; ++e) acc[e] = 0.0;
    for (int k = cq_ptr[q] + lane; k < cq_ptr[q + 1]; k += 64) {
        const double *v = vbuf + (int64_t)cq_f[k] * MM;
        for (int e = 0; e < MM; ++e) acc[e] += v[e];
    }
    for (int e = 0; e < MM; ++e)
        for (int o = 32; o > 0; o >>= 1) acc[e] += __shfl_xor(acc[e], o, 64);
    // g and h belong to the same group: aggregates never span two specimens, nor do matrix rows
    const int g = cq_gh[2 * q], h = cq_gh[2 * q + 1];
    const int b = agg_grp[g], base = cb[b], nb = cb[b + 1] - base;
    double *A = Ac + off[b];
    for (int e = lane; e < MM; e += 64) {
        double val = 0.0;
        for (int t = 0; t < MM; ++t)
            if (t == e) val = acc[t];
        A[(int64_t)(g * M + e / M - base) * nb + (h * M + e % M - base)] = val;
    }
}

// Blocked Gauss-Jordan inversion of the symmetrised Ac (SPD, so no pivoting), NB pivots per
// block step, two launches per step (reading A_in, writing A_out, so no element is read after
// another thread wrote it). A coarse DoF whose row is zero (an aggregate whose DoFs are all
// eliminated) is set to the identity before and to zero after, so it receives no correction.
// A coarse DoF whose pivot (its Schur complement against the pivots before it) falls to
// GJ_DROP times its own diagonal is a combination of earlier modes (an aggregate whose
// unmasked DoFs cannot tell a rotation from a translation): it is skipped as a pivot and its row
// and column of the inverse are zero, which inverts Ac on the span of the other modes.
constexpr int GJ_NB = 32;
constexpr double GJ_DROP = 1e-9;

// The kernels below take the group from blockIdx.y (the Gauss-Jordan steps of all groups run in
// the same launches) unless they say otherwise; nc is then the group's block size.
__global__ __launch_bounds__(256) void k_coarse_sym(const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                    double *__restrict__ A) {
    const int b = blockIdx.y, nc = cb[b + 1] - cb[b];
    A += off[b];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < (int64_t)nc * nc) {
        const int i = (int)(idx / nc), j = (int)(idx % nc);
        if (i < j) {
            const double s = 0.5 * (A[(int64_t)i * nc + j] + A[(int64_t)j * nc + i]);
            A[(int64_t)i * nc + j] = s;
            A[(int64_t)j * nc + i] = s;
        }
    }
}

// zero rows -> identity (and flag them), keep every diagonal for the pivot test; runs after
// k_coarse_sym, one wavefront per coarse DoF r of the whole system
__global__ __launch_bounds__(256) void k_coarse_dead(int ncoarse, int M, const int32_t *__restrict__ agg_grp,
                                                     const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                     double *__restrict__ Ac, double *__restrict__ odiag,
                                                     uint8_t *__restrict__ dead) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= ncoarse) return;
    const int b = agg_grp[r / M], nc = cb[b + 1] - cb[b], i = r - cb[b];
    double *A = Ac + off[b];
    double m = 0.0;
    for (int j = lane; j < nc; j += 64) m = fmax(m, fabs(A[(int64_t)i * nc + j]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) {
        dead[r] = m == 0.0;
        if (m == 0.0) A[(int64_t)i * nc + i] = 1.0;
        odiag[r] = A[(int64_t)i * nc + i];
    }
}

// Block step kb, part 1 (one workgroup per group): Dinv = A[K,K]^-1 by Gauss-Jordan in LDS, with
// the pivot drop described above. Dinv is stored NB x NB (row stride NB), zero outside nb x nb
// and in the rows and columns of dropped pivots.
__global__ __launch_bounds__(1024) void k_gj_diag(int kb, const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                   const double *__restrict__ Ain, const double *__restrict__ odiag_all,
                                                   uint8_t *__restrict__ dead_all, double *__restrict__ Dinv_all) {
    __shared__ double D[GJ_NB][GJ_NB + 1];
    __shared__ uint8_t skip[GJ_NB];
    const int b = blockIdx.y, base = cb[b], nc = cb[b + 1] - base;
    if (kb >= nc) return;  // this group's inversion has finished
    const double *A = Ain + off[b];
    const int nb = min(GJ_NB, nc - kb);
    const int tid = threadIdx.x, i = tid / GJ_NB, j = tid % GJ_NB;  // blockDim = NB * NB
    const bool in = i < nb && j < nb;
    D[i][j] = in ? A[(int64_t)(kb + i) * nc + kb + j] : 0.0;
    if (tid < GJ_NB) skip[tid] = 0;
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        // every thread reads the same pivot, so the skip branch is uniform across the workgroup
        const double pk = D[k][k];
        if (!(pk > GJ_DROP * odiag_all[base + kb + k])) {
            if (tid == 0) {
                skip[k] = 1;
                dead_all[base + kb + k] = 1;
            }
            continue;
        }
        const double rk = (j == k) ? 1.0 : D[k][j], ck = D[i][k];
        __syncthreads();
        if (in) {
            const double p = 1.0 / pk;
            if (i == k) D[i][j] = rk * p;
            else D[i][j] = ((j == k) ? 0.0 : D[i][j]) - ck * rk * p;
        }
        __syncthreads();
    }
    __syncthreads();
    Dinv_all[(int64_t)b * GJ_NB * GJ_NB + tid] = (in && !skip[i] && !skip[j]) ? D[i][j] : 0.0;
}

// Block step kb, part 2: the row panel Rt[j][k] = sum_m Dinv[k][m] A[kb+m][j] for j outside K
// (zero for j in K), row stride NB; four threads per column j, eight k each, Dinv staged in LDS.
__global__ __launch_bounds__(256) void k_gj_panel(int kb, const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                   const double *__restrict__ Ain, const double *__restrict__ Dinv_all,
                                                   double *__restrict__ Rt_all) {
    __shared__ double Ds[GJ_NB][GJ_NB + 1];
    const int b = blockIdx.y, base = cb[b], nc = cb[b + 1] - base;
    if (kb >= nc || (int)(blockIdx.x * (blockDim.x >> 2)) >= nc) return;  // uniform over the workgroup
    const double *Dinv = Dinv_all + (int64_t)b * GJ_NB * GJ_NB;
    for (int e = threadIdx.x; e < GJ_NB * GJ_NB; e += blockDim.x) Ds[e / GJ_NB][e % GJ_NB] = Dinv[e];
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x, j = t >> 2, k0 = (t & 3) * 8;
    if (j >= nc) return;
    const double *A = Ain + off[b];
    const int nb = min(GJ_NB, nc - kb);
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (j < kb || j >= kb + nb)
        for (int m = 0; m < nb; ++m) {
            const double a = A[(int64_t)(kb + m) * nc + j];
            for (int q = 0; q < 8; ++q) acc[q] += Ds[k0 + q][m] * a;
        }
    double *Rt = Rt_all + (int64_t)(base + j) * GJ_NB + k0;
    for (int q = 0; q < 8; ++q) Rt[q] = acc[q];
}

// Block step kb, part 3: the rank-NB update of the group's block, one 32 x 32 tile per
// workgroup (tiles are aligned with the pivot blocks, so a tile lies wholly inside or outside K
// in each direction):
//   tile in K x K:      Dinv
//   row tile in K:      Rt^T
//   column tile in K:   -A[:,K] Dinv
//   else:               A - A[:,K] Rt^T
// A group whose inversion has finished is copied, so that every block ends in the same buffer.
__global__ __launch_bounds__(256) void k_gj_update(int kb, int T, const int32_t *__restrict__ cb,
                                                    const int64_t *__restrict__ off, const double *__restrict__ Ain_all,
                                                    double *__restrict__ Aout_all, const double *__restrict__ Dinv_all,
                                                    const double *__restrict__ Rt_all) {
    static_assert(GJ_NB == 32, "tiles are 32 x 32 pivot blocks");
    __shared__ double L[32][33], R[32][33];
    const int b = blockIdx.y, base = cb[b], nc = cb[b + 1] - base;
    const int i0 = (blockIdx.x / T) * 32, j0 = (blockIdx.x % T) * 32;
    if (i0 >= nc || j0 >= nc) return;  // uniform over the workgroup
    const double *Ain = Ain_all + off[b];
    double *Aout = Aout_all + off[b];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (kb >= nc) {
        for (int q = 0; q < 4; ++q) {
            const int i = i0 + ty + 8 * q, j = j0 + tx;
            if (i < nc && j < nc) Aout[(int64_t)i * nc + j] = Ain[(int64_t)i * nc + j];
        }
        return;
    }
    const double *Dinv = Dinv_all + (int64_t)b * GJ_NB * GJ_NB, *Rt = Rt_all + (int64_t)base * GJ_NB;
    const int nb = min(GJ_NB, nc - kb);
    const bool ik = i0 == kb, jk = j0 == kb;
    if (ik && jk) {
        for (int q = 0; q < 4; ++q) {
            const int ii = ty + 8 * q, i = i0 + ii, j = j0 + tx;
            if (i < nc && j < nc) Aout[(int64_t)i * nc + j] = Dinv[ii * GJ_NB + tx];
        }
        return;
    }
    for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        const int r = e >> 5, m = e & 31;
        // R[jj][m]: Rt rows j0.. (row tile in K and general case) or Dinv^T (column tile in K)
        R[r][m] = jk ? Dinv[m * GJ_NB + r] : ((j0 + r < nc) ? Rt[(int64_t)(j0 + r) * GJ_NB + m] : 0.0);
        L[r][m] = (!ik && i0 + r < nc && m < nb) ? Ain[(int64_t)(i0 + r) * nc + kb + m] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < 4; ++q) {
        const int ii = ty + 8 * q, i = i0 + ii, j = j0 + tx;
        if (i >= nc || j >= nc) continue;
        double v;
        if (ik) {
            v = R[tx][ii];
        } else {
            double s = 0.0;
            for (int m = 0; m < 32; ++m) s += L[ii][m] * R[tx][m];
            v = jk ? -s : Ain[(int64_t)i * nc + j] - s;
        }
        Aout[(int64_t)i * nc + j] = v;
    }
}

__global__ __launch_bounds__(256) void k_coarse_undead(const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                       double *__restrict__ Ac, const uint8_t *__restrict__ dead_all) {
    const int b = blockIdx.y, nc = cb[b + 1] - cb[b];
    double *A = Ac + off[b];
    const uint8_t *dead = dead_all + cb[b];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nc * nc) return;
    const int i = (int)(idx / nc), j = (int)(idx % nc);
    if (dead[i] || dead[j]) A[idx] = 0.0;
}

// Restriction cz = P^T r: one workgroup of RW threads per aggregate g sums (P_i^T r_i)[m] over
// its members into czf[g M + m]. Each thread takes every RW-th member, RESTRICT_PER at a time with
// all their loads issued first; the offsets from the centre are read in member order. RW is the
// smallest of 64, 256 and 1024 that covers the largest aggregate in one round (else 1024), so
// that small aggregates do not launch idle wavefronts. (A split of an aggregate over several
// workgroups needs a cross-workgroup sum; on this GPU the device-scope fence that publishes a
// partial writes back the XCD's L2, which cost more than it saved.)
constexpr int RESTRICT_PER = 4;

template <typename T, int D, int VP, int RW>
__global__ __launch_bounds__(RW) void k_coarse_restrict(const int32_t *__restrict__ agg_ptr,
                                                                 const int32_t *__restrict__ agg_idx,
                                                                 const double *__restrict__ agg_dm,
                                                                 const uint8_t *__restrict__ mask, const T *__restrict__ r,
                                                                 double *__restrict__ czf, const double *__restrict__ scal) {
    constexpr int M = D == 3 ? 6 : 3;
    constexpr int NW = RW / 64;
    __shared__ double red[M][NW];
    const int g = blockIdx.x;
    const int beg = agg_ptr[g], end = agg_ptr[g + 1];
    double acc[M];
    for (int a = 0; a < M; ++a) acc[a] = 0.0;
    // the first members' ids and offsets are requested together with the stop flag
    int64_t ii0[RESTRICT_PER];
    double dm0[RESTRICT_PER][3];
#pragma unroll
    for (int u = 0; u < RESTRICT_PER; ++u) {
        const int m = beg + (int)threadIdx.x + u * RW;
        ii0[u] = m < end ? agg_idx[m] : -1;
        for (int k = 0; k < D; ++k) dm0[u][k] = m < end ? agg_dm[D * (int64_t)m + k] : 0.0;
    }
    if (scal[SC_DONE] != 0.0) return;
    for (int m0 = beg + (int)threadIdx.x; m0 < end; m0 += RW * RESTRICT_PER) {
        const bool first = m0 == beg + (int)threadIdx.x;
        int64_t ii[RESTRICT_PER];
#pragma unroll
        for (int u = 0; u < RESTRICT_PER; ++u) {
            const int m = m0 + u * RW;
            ii[u] = first ? ii0[u] : (m < end ? agg_idx[m] : -1);
        }
#pragma unroll
        for (int u = 0; u < RESTRICT_PER; ++u) {
            const int m = m0 + u * RW;
            if (ii[u] >= 0) {
                double P[3][6], d[3];
                for (int k = 0; k < D; ++k) d[k] = first ? dm0[u][k] : agg_dm[D * (int64_t)m + k];
                rigid_modes<D>(d, mask[ii[u]], P);
                double ri[3] = {0, 0, 0};
                for (int k = 0; k < D; ++k) ri[k] = (double)r[VP * ii[u] + k];
                for (int a = 0; a < M; ++a)
                    for (int k = 0; k < D; ++k) acc[a] += P[k][a] * ri[k];
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int a = 0; a < M; ++a) {
        double v = acc[a];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[a][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < M) {
        double v = 0.0;
        for (int w = 0; w < NW; ++w) v += red[threadIdx.x][w];
        czf[(int64_t)g * M + threadIdx.x] = v;
    }
}

// cy = Ac^-1 cz for group blockIdx.y: each wavefront forms one row of the group's block, lanes
// striding the row (Ac^-1 is symmetric, row = column), against cz staged in LDS.
__global__ __launch_bounds__(256) void k_coarse_gemv(const int32_t *__restrict__ cb, const int64_t *__restrict__ off,
                                                      const double *__restrict__ Ainv_all, const double *__restrict__ czf,
                                                      double *__restrict__ cy, const double *__restrict__ scal) {
    const int b = blockIdx.y, base = cb[b], nc = cb[b + 1] - base;
    if ((int)(blockIdx.x * (blockDim.x >> 6)) >= nc) return;  // uniform over the workgroup
    __shared__ double cz[1024];  // a group's block has at most 1024 DoFs
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    // this wavefront's row of Ac^-1, requested together with the stop flag
    const double *Arow = Ainv_all + off[b] + (int64_t)r * nc;
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 64 * q;
        a[q] = (r < nc && c < nc) ? Arow[c] : 0.0;
    }
    // cz is requested with the stop flag too (a group has at most 4 x 256 DoFs)
    double zq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = (int)threadIdx.x + q * (int)blockDim.x;
        zq[q] = c < nc ? czf[base + c] : 0.0;
    }
    if (scal[SC_DONE] != 0.0) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = (int)threadIdx.x + q * (int)blockDim.x;
        if (c < nc) cz[c] = zq[q];
    }
    __syncthreads();
    if (r >= nc) return;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = lane + 64 * q;
        if (c < nc) s += a[q] * cz[c];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) cy[base + r] = s;
}

// flag = 1 where the DoF masks differ
__global__ __launch_bounds__(256) void k_mask_diff(int64_t n, const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                                                    int32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) *flag = 1;
}

}  // namespace

static int ck(lpm_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LPM_OK;
}

void coarse_free(lpm_ctx *c) {
    if (c->gj_graph) hipGraphExecDestroy(c->gj_graph);
    void *ptrs[] = {c->agg_of, c->agg_ptr, c->agg_idx, c->agg_d, c->ra_ptr, c->ra_id,  c->cq_ptr,
                    c->cq_f,   c->cq_gh,   c->cvbuf,   c->Ac,    c->cz,     c->cy,     c->dofmask,
                    c->cgrp_cb, c->cgrp_off, c->agg_grp, c->dofmask_prev, c->agg_dm};
    for (void *p : ptrs)
        if (p) hipFree(p);
}

// Aggregates and index lists (host, once per context).
static int coarse_setup(lpm_ctx *c) {
    if (c->agg_of) return LPM_OK;
    const int64_t n = c->n;
    const int D = c->dim, M = D == 3 ? 6 : 3;
    std::vector<double> x(3 * n);
    std::vector<int32_t> cp(n + 1), ci(c->nconn);
    // bonds, to split cells into the pieces that intact bonds connect
    std::vector<int32_t> hbp(n + 1), hp2(std::max<int64_t>(c->nbond, 1));
    std::vector<double> hbd(std::max<int64_t>(c->nbond, 1));
    LPM_HIP(c, hipMemcpyAsync(hbp.data(), c->bptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    if (c->nbond > 0) {
        LPM_HIP(c, hipMemcpyAsync(hp2.data(), c->p2, sizeof(int32_t) * c->nbond, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipMemcpyAsync(hbd.data(), c->bdamage, sizeof(double) * c->nbond, hipMemcpyDeviceToHost, c->stream));
    }
    LPM_HIP(c, hipMemcpyAsync(x.data(), c->xyz0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(cp.data(), c->cptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipMemcpyAsync(ci.data(), c->cidx, sizeof(int32_t) * c->nconn, hipMemcpyDeviceToHost, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    const double spacing = 2.0 * c->radius;
    // Aggregates are the cells of a grid laid over each specimen's own bounding box (the whole
    // system when there are no specimens), so they never span two specimens, and each specimen is
    // one coarse group (one diagonal block of Ac). Each box edge is cut into round(extent /
    // target) equal cells: a remainder cell one or two lattice rows thick would make an aggregate
    // whose modes are dependent once its prescribed DoFs are masked. The target edge is the set
    // value, or, per group, the smallest of 8, 9, 10, ... lattice spacings that keeps the group's
    // block at most 768 DoFs (the inversion's cost grows with its cube). A cell that a crack or
    // hole cuts is split into the pieces its intact bonds connect: the two faces of a crack move
    // apart, which one set of rigid-body modes per cell cannot follow.
    std::vector<int32_t> uf(n);
    auto find = [&](int32_t a) {
        while (uf[a] != a) {
            uf[a] = uf[uf[a]];
            a = uf[a];
        }
        return a;
    };
    const int64_t nsp = std::max(c->nspec, 1);
    std::vector<int64_t> first(nsp + 1, 0);
    if (c->nspec > 0)
        for (int64_t s = 0; s <= nsp; ++s) first[s] = c->spec_first[s];
    else
        first[1] = n;
    std::vector<int32_t> agg(n), agg_grp;
    std::vector<int32_t> cb(nsp + 1, 0);
    std::vector<int64_t> off(nsp + 1, 0);
    int nagg = 0, nbmax = 0;
    for (int64_t s = 0; s < nsp; ++s) {
        double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int64_t i = first[s]; i < first[s + 1]; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], x[3 * i + k]);
                hi[k] = std::max(hi[k], x[3 * i + k]);
            }
        int ns = 0;  // aggregates of this group, numbered from nagg in order of first member
        for (int cells = c->coarse_cells > 0 ? c->coarse_cells : 8; first[s + 1] > first[s]; cells += (cells < 32 ? 1 : 8)) {
            const double target = cells * spacing;
            int64_t dims[3] = {1, 1, 1};
            double edge[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < D; ++k) {
                dims[k] = std::max<int64_t>(1, (int64_t)std::llround((hi[k] - lo[k]) / target));
                edge[k] = (hi[k] - lo[k]) / (double)dims[k];
            }
            // cell of each particle, then the connected pieces of each cell (union-find over the
            // intact bonds whose two ends share a cell), numbered in order of first member
            const int64_t ncells = dims[0] * dims[1] * dims[2];
            std::vector<int64_t> key(first[s + 1] - first[s]);
            std::vector<char> used(ncells, 0);
            int64_t ncell = 0;
            for (int64_t i = first[s]; i < first[s + 1]; ++i) {
                int64_t kk = 0;
                for (int k = D - 1; k >= 0; --k) {
                    const int64_t ck =
                        edge[k] > 0.0 ? std::min<int64_t>(dims[k] - 1, (int64_t)((x[3 * i + k] - lo[k]) / edge[k])) : 0;
                    kk = kk * dims[k] + ck;
                }
                key[i - first[s]] = kk;
                if (!used[kk]) used[kk] = 1, ++ncell;
            }
            // splitting only adds aggregates: a size whose plain cells exceed the budget is skipped
            if (c->coarse_cells <= 0 && ncell * M > c->coarse_budget && ncells > 1) continue;
            for (int64_t i = first[s]; i < first[s + 1]; ++i) uf[i] = (int32_t)i;
            for (int64_t i = first[s]; i < first[s + 1]; ++i)
                for (int b = hbp[i]; b < hbp[i + 1]; ++b) {
                    const int32_t j = hp2[b];
                    if (j < first[s] || j >= first[s + 1] || key[j - first[s]] != key[i - first[s]] ||
                        std::fabs(hbd[b] - 1.0) < LPM_EPS)
                        continue;
                    const int32_t ra = find((int32_t)i), rb = find(j);
                    if (ra != rb) uf[std::max(ra, rb)] = std::min(ra, rb);
                }
            std::vector<int32_t> id(first[s + 1] - first[s], -1);
            ns = 0;
            for (int64_t i = first[s]; i < first[s + 1]; ++i) {
                const int64_t r = find((int32_t)i) - first[s];
                if (id[r] < 0) id[r] = ns++;
                agg[i] = nagg + id[r];
            }
            if ((int64_t)ns * M > 1024) {
                // more pieces than one block holds (a shattered specimen): plain cells at this size
                std::vector<int32_t> cid(ncells, -1);
                ns = 0;
                for (int64_t i = first[s]; i < first[s + 1]; ++i) {
                    const int64_t kk = key[i - first[s]];
                    if (cid[kk] < 0) cid[kk] = ns++;
                    agg[i] = nagg + cid[kk];
                }
            }
            if (ncells == 1) break;
            if (c->coarse_cells > 0 || (int64_t)ns * M <= c->coarse_budget) break;
        }
        if ((int64_t)ns * M > 1024)
            return lpm_fail(c, LPM_E_LIMIT, "coarse block above 1024 DoFs: raise coarse_cells");
        nagg += ns;
        agg_grp.insert(agg_grp.end(), ns, (int32_t)s);
        cb[s + 1] = cb[s] + ns * M;
        off[s + 1] = off[s] + (int64_t)ns * M * ns * M;
        nbmax = std::max(nbmax, ns * M);
    }
    const int nc = nagg * M;
    // members, centres, offsets
    std::vector<int32_t> aptr(nagg + 1, 0), aidx(n);
    for (int64_t i = 0; i < n; ++i) aptr[agg[i] + 1]++;
    for (int g = 0; g < nagg; ++g) aptr[g + 1] += aptr[g];
    {
        std::vector<int32_t> fill(aptr.begin(), aptr.end() - 1);
        for (int64_t i = 0; i < n; ++i) aidx[fill[agg[i]]++] = (int32_t)i;
    }
    std::vector<double> cen(3 * nagg, 0.0), d(3 * n);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) cen[3 * agg[i] + k] += x[3 * i + k];
    for (int g = 0; g < nagg; ++g)
        for (int k = 0; k < 3; ++k) cen[3 * g + k] /= (double)(aptr[g + 1] - aptr[g]);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) d[3 * i + k] = x[3 * i + k] - cen[3 * agg[i] + k];
    // distinct aggregates of each row's columns (ascending)
    std::vector<int32_t> rptr(n + 1, 0), rid;
    rid.reserve(2 * n);
    for (int64_t i = 0; i < n; ++i) {
        std::vector<int32_t> hs;
        for (int q = cp[i]; q < cp[i + 1]; ++q) hs.push_back(agg[ci[q]]);
        std::sort(hs.begin(), hs.end());
        hs.erase(std::unique(hs.begin(), hs.end()), hs.end());
        rid.insert(rid.end(), hs.begin(), hs.end());
        rptr[i + 1] = (int32_t)rid.size();
    }
    // coarse blocks (g, h) and their contributing row entries f, in row order
    std::vector<int32_t> qgh;
    {
        std::vector<std::pair<int64_t, int32_t>> kf;  // (g*nagg + h, f)
        kf.reserve(rid.size());
        for (int64_t i = 0; i < n; ++i)
            for (int f = rptr[i]; f < rptr[i + 1]; ++f) kf.push_back({(int64_t)agg[i] * nagg + rid[f], f});
        std::stable_sort(kf.begin(), kf.end(), [](const std::pair<int64_t, int32_t> &a, const std::pair<int64_t, int32_t> &b) {
            return a.first < b.first;
        });
        std::vector<int32_t> qptr(1, 0), qf;
        for (size_t t = 0; t < kf.size(); ++t) {
            if (t > 0 && kf[t].first != kf[t - 1].first) qptr.push_back((int32_t)qf.size());
            if (t == 0 || kf[t].first != kf[t - 1].first) {
                qgh.push_back((int32_t)(kf[t].first / nagg));
                qgh.push_back((int32_t)(kf[t].first % nagg));
            }
            qf.push_back(kf[t].second);
        }
        qptr.push_back((int32_t)qf.size());
        // every coarse block must lie inside one group's diagonal block (the kernels index it so)
        for (size_t t = 0; t < qgh.size(); t += 2)
            if (agg_grp[qgh[t]] != agg_grp[qgh[t + 1]])
                return lpm_fail(c, LPM_E_TOPO, "the matrix pattern joins two specimens");
        c->ncq = (int64_t)qgh.size() / 2;
        LPM_HIP(c, hipMalloc(&c->cq_ptr, sizeof(int32_t) * qptr.size()));
        LPM_HIP(c, hipMalloc(&c->cq_f, sizeof(int32_t) * std::max<size_t>(qf.size(), 1)));
        LPM_HIP(c, hipMemcpyAsync(c->cq_ptr, qptr.data(), sizeof(int32_t) * qptr.size(), hipMemcpyHostToDevice, c->stream));
        LPM_HIP(c, hipMemcpyAsync(c->cq_f, qf.data(), sizeof(int32_t) * qf.size(), hipMemcpyHostToDevice, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
    }
    c->nagg = nagg;
    c->agg_max = 0;
    for (int g = 0; g < nagg; ++g) c->agg_max = std::max(c->agg_max, aptr[g + 1] - aptr[g]);
    c->nmode = M;
    c->ncoarse = nc;
    c->nra = (int64_t)rid.size();
    LPM_HIP(c, hipMalloc(&c->cq_gh, sizeof(int32_t) * std::max<size_t>(qgh.size(), 1)));
    LPM_HIP(c, hipMalloc(&c->agg_of, sizeof(int32_t) * n));
    LPM_HIP(c, hipMalloc(&c->agg_ptr, sizeof(int32_t) * (nagg + 1)));
    LPM_HIP(c, hipMalloc(&c->agg_idx, sizeof(int32_t) * n));
    LPM_HIP(c, hipMalloc(&c->agg_d, sizeof(double) * D * std::max<int64_t>(n, 1)));
    LPM_HIP(c, hipMalloc(&c->ra_ptr, sizeof(int32_t) * (n + 1)));
    LPM_HIP(c, hipMalloc(&c->ra_id, sizeof(int32_t) * std::max<size_t>(rid.size(), 1)));
    LPM_HIP(c, hipMalloc(&c->cvbuf, sizeof(double) * M * M * std::max<size_t>(rid.size(), 1)));
    // two ping-pong buffers of the packed blocks, the groups' NB x NB block inverses, their row
    // panels, the original diagonal and the dead-row flags
    c->ncgrp = (int)nsp;
    c->cgrp_nbmax = nbmax;
    c->ac_elems = off[nsp];
    LPM_HIP(c, hipMalloc(&c->Ac, sizeof(double) * (2 * (size_t)c->ac_elems + (size_t)nsp * GJ_NB * GJ_NB +
                                                   (size_t)nc * GJ_NB + nc) + nc));
    LPM_HIP(c, hipMalloc(&c->cgrp_cb, sizeof(int32_t) * (nsp + 1)));
    LPM_HIP(c, hipMalloc(&c->cgrp_off, sizeof(int64_t) * (nsp + 1)));
    LPM_HIP(c, hipMalloc(&c->agg_grp, sizeof(int32_t) * std::max(nagg, 1)));
    LPM_HIP(c, hipMemcpyAsync(c->cgrp_cb, cb.data(), sizeof(int32_t) * (nsp + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->cgrp_off, off.data(), sizeof(int64_t) * (nsp + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->agg_grp, agg_grp.data(), sizeof(int32_t) * nagg, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMalloc(&c->cz, sizeof(double) * std::max(nc, 1)));
    LPM_HIP(c, hipMalloc(&c->cy, sizeof(double) * nc));
    LPM_HIP(c, hipMalloc(&c->dofmask, n));
    LPM_HIP(c, hipMalloc(&c->agg_dm, sizeof(double) * D * std::max<int64_t>(n, 1)));
    LPM_HIP(c, hipMemsetAsync(c->cz, 0, sizeof(double) * std::max(nc, 1), c->stream));
    // the offsets on the device keep D entries per particle (the modes read no z offset in 2D), in
    // particle order (agg_d) and in member order (agg_dm)
    std::vector<double> dD(D * n), dm(D * n);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < D; ++k) dD[D * i + k] = d[3 * i + k];
    for (int64_t m = 0; m < n; ++m)
        for (int k = 0; k < D; ++k) dm[D * m + k] = d[3 * (int64_t)aidx[m] + k];
    LPM_HIP(c, hipMemcpyAsync(c->agg_dm, dm.data(), sizeof(double) * D * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemsetAsync(c->cy, 0, sizeof(double) * nc, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->cq_gh, qgh.data(), sizeof(int32_t) * qgh.size(), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->agg_of, agg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->agg_ptr, aptr.data(), sizeof(int32_t) * (nagg + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->agg_idx, aidx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->agg_d, dD.data(), sizeof(double) * D * n, hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->ra_ptr, rptr.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipMemcpyAsync(c->ra_id, rid.data(), sizeof(int32_t) * rid.size(), hipMemcpyHostToDevice, c->stream));
    LPM_HIP(c, hipStreamSynchronize(c->stream));
    return LPM_OK;
}

template <typename T, int D>
static int coarse_rows_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_coarse_rows<T, D>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr,
                       c->scol, c->sp, (const T *)c->aval, c->agg_of, c->agg_d, c->dofmask, c->ra_ptr, c->ra_id, c->cvbuf);
    return ck(c, "k_coarse_rows");
}

// The eliminated-DoF mask of the current K, and under coarse reuse its comparison with the mask
// of the last build; the comparison's flag goes to h_iflag[1] without a wait. A Newton loop
// queues this behind its first force pass (which leaves damage_visual current), so the flag
// arrives with the first residual and coarse_prepare need not wait for it on its own.
int coarse_check(lpm_ctx *c) {
    if (!two_level(c) || !c->k_ready) return LPM_OK;
    int rc = coarse_setup(c);
    if (rc) return rc;
    if (c->coarse_version == c->k_version || c->coarse_check_version == c->k_version) return LPM_OK;
    const int64_t n = c->n;
    hipLaunchKernelGGL(k_dofmask, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->dim, c->cons, c->damage_visual,
                       c->dofmask);
    if ((rc = ck(c, "k_dofmask"))) return rc;
    c->coarse_check_version = c->k_version;
    c->coarse_check_diff = c->coarse_reuse && c->coarse_built && !c->coarse_stale;
    if (c->coarse_check_diff) {
        LPM_HIP(c, hipMemsetAsync(c->iflag, 0, sizeof(int32_t), c->stream));
        hipLaunchKernelGGL(k_mask_diff, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, n, (const uint8_t *)c->dofmask,
                           (const uint8_t *)c->dofmask_prev, c->iflag);
        if ((rc = ck(c, "k_mask_diff"))) return rc;
        LPM_HIP(c, hipMemcpyAsync(c->h_iflag + 1, c->iflag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    }
    return LPM_OK;
}

// Rebuild and invert Ac after the matrix changed.
int coarse_prepare(lpm_ctx *c) {
    if (!two_level(c)) return LPM_OK;
    int rc = coarse_setup(c);
    if (rc) return rc;
    if (c->coarse_version == c->k_version) return LPM_OK;
    const int64_t n = c->n;
    // the mask of this K (coarse_check, queued earlier or now); the stale flag may have been set
    // by a solve since, and a stale space is rebuilt whatever the masks say
    if ((rc = coarse_check(c))) return rc;
    if (c->coarse_check_diff && c->coarse_reuse && c->coarse_built && !c->coarse_stale) {
        // keep the coarse space of the last build if the eliminated DoFs are the same
        LPM_HIP(c, hipStreamSynchronize(c->stream));  // returns at once when the flag has arrived
        if (c->h_iflag[1] == 0) {
            c->coarse_version = c->k_version;
            return LPM_OK;
        }
    }
    if (c->precision == 0) rc = c->dim == 3 ? coarse_rows_t<float, 3>(c) : coarse_rows_t<float, 2>(c);
    else rc = c->dim == 3 ? coarse_rows_t<double, 3>(c) : coarse_rows_t<double, 2>(c);
    if (rc) return rc;
    const int nc = c->ncoarse, M = c->nmode, G = c->ncgrp, nbmax = c->cgrp_nbmax;
    const int64_t ne = c->ac_elems;
    LPM_HIP(c, hipMemsetAsync(c->Ac, 0, sizeof(double) * (size_t)ne, c->stream));
    if (c->ncq > 0) {
        const int gq = (int)((c->ncq + 3) / 4);
        if (M == 6)
            hipLaunchKernelGGL((k_coarse_blocks<6>), dim3(gq), dim3(256), 0, c->stream, c->ncq, c->cq_ptr, c->cq_f, c->cq_gh,
                               c->cvbuf, c->agg_grp, c->cgrp_cb, c->cgrp_off, c->Ac);
        else
            hipLaunchKernelGGL((k_coarse_blocks<3>), dim3(gq), dim3(256), 0, c->stream, c->ncq, c->cq_ptr, c->cq_f, c->cq_gh,
                               c->cvbuf, c->agg_grp, c->cgrp_cb, c->cgrp_off, c->Ac);
        if ((rc = ck(c, "k_coarse_blocks"))) return rc;
    }
    double *A0 = c->Ac, *A1 = c->Ac + ne, *Dinv = c->Ac + 2 * ne, *Rt = Dinv + (int64_t)G * GJ_NB * GJ_NB;
    double *odiag = Rt + (int64_t)nc * GJ_NB;
    uint8_t *dead = (uint8_t *)(odiag + nc);
    const int nsteps = (nbmax + GJ_NB - 1) / GJ_NB, T = nsteps;  // T x T tiles of 32 cover the largest block
    // every group's block in the same launches (grid y = group, x over the largest block)
    const dim3 gsq(blocks_for((int64_t)nbmax * nbmax, 256), G);
    auto issue = [&]() {
        hipLaunchKernelGGL(k_coarse_sym, gsq, dim3(256), 0, c->stream, (const int32_t *)c->cgrp_cb,
                           (const int64_t *)c->cgrp_off, A0);
        hipLaunchKernelGGL(k_coarse_dead, dim3(blocks_for(nc, 4)), dim3(256), 0, c->stream, nc, M,
                           (const int32_t *)c->agg_grp, (const int32_t *)c->cgrp_cb, (const int64_t *)c->cgrp_off, A0, odiag,
                           dead);
        for (int st = 0; st < nsteps; ++st) {
            const double *Ain = (st & 1) ? A1 : A0;
            double *Aout = (st & 1) ? A0 : A1;
            hipLaunchKernelGGL(k_gj_diag, dim3(1, G), dim3(GJ_NB * GJ_NB), 0, c->stream, st * GJ_NB,
                               (const int32_t *)c->cgrp_cb, (const int64_t *)c->cgrp_off, Ain, (const double *)odiag, dead,
                               Dinv);
            hipLaunchKernelGGL(k_gj_panel, dim3(blocks_for((int64_t)nbmax * 4, 256), G), dim3(256), 0, c->stream,
                               st * GJ_NB, (const int32_t *)c->cgrp_cb, (const int64_t *)c->cgrp_off, Ain,
                               (const double *)Dinv, Rt);
            hipLaunchKernelGGL(k_gj_update, dim3(T * T, G), dim3(256), 0, c->stream, st * GJ_NB, T,
                               (const int32_t *)c->cgrp_cb, (const int64_t *)c->cgrp_off, Ain, Aout, (const double *)Dinv,
                               (const double *)Rt);
        }
        hipLaunchKernelGGL(k_coarse_undead, gsq, dim3(256), 0, c->stream, (const int32_t *)c->cgrp_cb,
                           (const int64_t *)c->cgrp_off, (nsteps & 1) ? A1 : A0, (const uint8_t *)dead);
    };
    if (nbmax > 0 && !c->coarse_graph) {
        // issued one by one: each step's kernels run for longer than the host takes to launch them
        issue();
        if ((rc = ck(c, "coarse inversion"))) return rc;
    } else if (nbmax > 0) {
        // captured once into a graph (for block steps too short to keep up with host launches)
        if (!c->gj_graph) {
            hipGraph_t graph = nullptr;
            LPM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            issue();
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) hipGraphDestroy(graph);
                return lpm_fail(c, LPM_E_HIP,
                                std::string("inversion graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
            }
            const hipError_t ie = hipGraphInstantiate(&c->gj_graph, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                c->gj_graph = nullptr;
                return lpm_fail(c, LPM_E_HIP, std::string("inversion graph: ") + hipGetErrorString(ie));
            }
        }
        LPM_HIP(c, hipGraphLaunch(c->gj_graph, c->stream));
    }
    c->Ainv = (nsteps & 1) ? A1 : A0;
    c->coarse_version = c->k_version;
    c->st_coarse_builds++;
    c->coarse_built = true;
    c->coarse_stale = false;
    c->coarse_ref_iters = -1;
    if (c->coarse_reuse) {
        if (!c->dofmask_prev) LPM_HIP(c, hipMalloc(&c->dofmask_prev, std::max<int64_t>(n, 1)));
        LPM_HIP(c, hipMemcpyAsync(c->dofmask_prev, c->dofmask, n, hipMemcpyDeviceToDevice, c->stream));
    }
    return LPM_OK;
}

template <typename T, int D, int VP>
static int coarse_apply_t(lpm_ctx *c) {
#define RESTRICT_LAUNCH(RW)                                                                                            \
    hipLaunchKernelGGL((k_coarse_restrict<T, D, VP, RW>), dim3(c->nagg), dim3(RW), 0, c->stream, c->agg_ptr, c->agg_idx, \
                       c->agg_dm, c->dofmask, (const T *)c->v_r, c->cz, (const double *)c->scal)
    if (c->agg_max <= 64 * RESTRICT_PER) RESTRICT_LAUNCH(64);
    else if (c->agg_max <= 256 * RESTRICT_PER) RESTRICT_LAUNCH(256);
    else RESTRICT_LAUNCH(1024);
#undef RESTRICT_LAUNCH
    int rc = ck(c, "k_coarse_restrict");
    if (rc) return rc;
    // 4 wavefronts per workgroup, one row of a group's block per wavefront
    hipLaunchKernelGGL(k_coarse_gemv, dim3(blocks_for(std::max(c->cgrp_nbmax, 1), 4), c->ncgrp), dim3(256), 0, c->stream,
                       (const int32_t *)c->cgrp_cb, (const int64_t *)c->cgrp_off, (const double *)c->Ainv,
                       (const double *)c->cz, c->cy, (const double *)c->scal);
    return ck(c, "k_coarse_gemv");
}

int launch_coarse_apply(lpm_ctx *c) {
    if (c->precision == 0) return c->dim == 3 ? coarse_apply_t<float, 3, 3>(c) : coarse_apply_t<float, 2, 2>(c);
    return c->dim == 3 ? coarse_apply_t<double, 3, 3>(c) : coarse_apply_t<double, 2, 2>(c);
}
