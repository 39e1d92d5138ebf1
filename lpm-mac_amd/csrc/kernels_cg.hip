// Block-sparse SpMV and the preconditioned conjugate gradient that replaces
// Eigen::ConjugateGradient<SparseMatrix<double>, Lower|Upper> (solver.h:192-217).
//
// Matrix: SELL-64 (see sell_off in reduce.h). The SpMV runs one wavefront per slice of 64
// rows with lane l owning row 64k+l, so every value load and every gather of the input
// vector touches 64 consecutive entries (neighbouring rows of a lattice have neighbouring
// columns), and no cross-lane reduction is needed.
//
// Solver: Jacobi (or 3x3 block-Jacobi) preconditioned CG in the single-reduction form of
// Chronopoulos and Gear (J. Comput. Appl. Math. 25, 1989): per iteration one SpMV (which
// also forms delta = (w, u)) and one fused vector kernel whose workgroups first reduce the
// partial sums to alpha/beta and the stopping decision, then do the p, s, x, r, u updates
// and the partials of ||r||^2 and gamma = (r, u). The stopping rule and defaults are Eigen
// 3.4's: x0 = 0, stop when ||r||^2 < max(tol^2 ||b||^2, DBL_MIN) or after maxIters iterations.
// It reaches the same solution as Eigen's recurrence (the tests check the converged
// vectors, not the iterates). Scalars are FP64 in device memory; vectors and matrix are
// FP32 (precision 0) or FP64 (precision 1). No host synchronisation happens inside an
// iteration: the "done" flag makes the remaining kernels of a batch return at once.
//
// Vectors are stored padded to VP entries per particle (VP = 4 in 3D, 2 in 2D).
#include <algorithm>
#include <cfloat>
#include <vector>

#include "lpm_internal.h"
#include "reduce.h"

namespace {

// One particle's DoFs in a CG vector: 3 entries in 3D (12 or 24 bytes, element-aligned), 2 in 2D
template <typename T, int VP>
struct alignas((VP & (VP - 1)) == 0 ? sizeof(T) * VP : sizeof(T)) Vec {
    T v[VP];
};

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD works on
// a contiguous range of rows (its L2 then holds the matching window of the input vector).
__device__ __forceinline__ int64_t xcd_block(int remap) {
    const int64_t b = blockIdx.x, g = gridDim.x;
    if (!remap || (g & 7)) return b;
    return (b & 7) * (g >> 3) + (b >> 3);
}

template <typename T, int D, int VP>
__device__ __forceinline__ void precond_apply(const T *__restrict__ invd, int block_pre, int64_t i, const Vec<T, VP> &r,
                                              Vec<T, VP> &z) {
    for (int k = 0; k < VP; ++k) z.v[k] = (T)0;
    if (block_pre) {
        for (int a = 0; a < D; ++a) {
            T s = 0;
            for (int q = 0; q < D; ++q) s += invd[(int64_t)D * D * i + a * D + q] * r.v[q];
            z.v[a] = s;
        }
    } else {
        for (int k = 0; k < D; ++k) z.v[k] = invd[D * i + k] * r.v[k];
    }
}

// ---------------------------------------------------------------- CG prologue
// x = 0, r = b, u = M^-1 r, p = s = 0; partials ||b||^2 and gamma = (r, u)
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_init(int64_t n, const double *__restrict__ b64, const T *__restrict__ invd,
                                                  int block_pre, Vec<T, VP> *__restrict__ x, Vec<T, VP> *__restrict__ r,
                                                  Vec<T, VP> *__restrict__ u, Vec<T, VP> *__restrict__ p,
                                                  Vec<T, VP> *__restrict__ s, double *partial, int cap, double *scal,
                                                  double tol, double maxit) {
    // the prologue SpMV that follows checks the flag; clear what the previous solve left and
    // set the solve's constants, which the iteration kernels read from here (so that a batch
    // of them can be captured once and replayed for every solve)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[SC_DONE] = 0.0;
        scal[SC_IT0] = 0.0;
        scal[SC_TOL] = tol;
        scal[SC_MAXIT] = maxit;
    }
    double v[2] = {0.0, 0.0};
    for (int64_t i = xcd_major_block() * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Vec<T, VP> rv, zv, zero;
        for (int k = 0; k < VP; ++k) { rv.v[k] = (T)0; zero.v[k] = (T)0; }
        for (int k = 0; k < D; ++k) rv.v[k] = (T)b64[D * i + k];
        precond_apply<T, D, VP>(invd, block_pre, i, rv, zv);
        x[i] = zero;
        p[i] = zero;
        s[i] = zero;
        r[i] = rv;
        u[i] = zv;
        for (int k = 0; k < D; ++k) {
            v[0] += (double)rv.v[k] * (double)rv.v[k];
            v[1] += (double)rv.v[k] * (double)zv.v[k];
        }
    }
    block_partial<2>(v, partial, cap);
}

// ---------------------------------------------------------------- SpMV  w = A u  (+ delta = (w, u))
// Column of a SELL slot: int32 column ids, or int16 offsets j - i (2 bytes less per block, used
// when every offset of the pattern fits).
// Loads of the matrix stream: non-temporal when NT (the values and columns are read once per
// SpMV; the input vector should keep the caches)
template <bool NT, typename X>
__device__ __forceinline__ X ld_stream(const X *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename IX>
__device__ __forceinline__ int col_of(IX v, int64_t row) {
    if constexpr (sizeof(IX) == 2) return (int)row + (int)v;
    else return (int)v;
}

// One wavefront per 64-row slice; lane l owns row 64k+l and walks its slots. UNR slots are
// loaded before they are consumed so that each lane has several independent loads in flight.
template <typename T, int D, int VP, bool FUSED, int UNR, typename IX, bool NT = false>
__global__ __launch_bounds__(512) void k_spmv(int64_t n, int64_t nslice, const int64_t *__restrict__ sp,
                                               const int32_t *__restrict__ cptr, const IX *__restrict__ scol,
                                               const T *__restrict__ aval, const Vec<T, VP> *__restrict__ u,
                                               Vec<T, VP> *__restrict__ w, double *partial, int cap, const double *scal,
                                               int remap) {
    constexpr int EB = D * D;
    const int lane = threadIdx.x & 63;
    const int64_t k = xcd_block(remap) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    // the slice's row length and slot range are requested together with the stop flag
    const int64_t row = k * 64 + lane;
    int len = 0;
    int64_t base = 0, send = 0;
    if (k < nslice) {
        len = row < n ? cptr[row + 1] - cptr[row] : 0;
        base = sp[k];
        send = sp[k + 1];
    }
    if (FUSED && scal[SC_DONE] != 0.0) return;
    double delta = 0.0;
    if (k < nslice) {
        const int W = (int)((send - base) >> 6);
        const IX *ci = scol + base + lane;
        const T *av = aval + EB * base + lane;
        T acc[D];
        for (int r = 0; r < D; ++r) acc[r] = (T)0;
        int s = 0;
        for (; s + UNR <= W; s += UNR) {
            int j[UNR];
            Vec<T, VP> pj[UNR];
            T e[UNR][EB];
#pragma unroll
            for (int q = 0; q < UNR; ++q) j[q] = (s + q < len) ? col_of(ld_stream<NT>(&ci[(int64_t)(s + q) * 64]), row) : -1;
#pragma unroll
            for (int q = 0; q < UNR; ++q) {
                if (j[q] >= 0) {
                    pj[q] = u[j[q]];
#pragma unroll
                    for (int e2 = 0; e2 < EB; ++e2) e[q][e2] = ld_stream<NT>(&av[((int64_t)(s + q) * EB + e2) * 64]);
                } else {
                    for (int t = 0; t < VP; ++t) pj[q].v[t] = (T)0;
                    for (int e2 = 0; e2 < EB; ++e2) e[q][e2] = (T)0;
                }
            }
#pragma unroll
            for (int q = 0; q < UNR; ++q)
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) acc[r] += e[q][r * D + c] * pj[q].v[c];
        }
        for (; s < W; ++s) {
            if (s < len) {
                const int j = col_of(ld_stream<NT>(&ci[(int64_t)s * 64]), row);
                const Vec<T, VP> pj = u[j];
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) acc[r] += ld_stream<NT>(&av[((int64_t)s * EB + r * D + c) * 64]) * pj.v[c];
            }
        }
        if (row < n) {
            Vec<T, VP> wv;
            for (int t = 0; t < VP; ++t) wv.v[t] = (T)0;
            for (int r = 0; r < D; ++r) wv.v[r] = acc[r];
            w[row] = wv;
            if (FUSED) {
                const Vec<T, VP> ui = u[row];
                for (int r = 0; r < D; ++r) delta += (double)ui.v[r] * (double)acc[r];
            }
        }
    }
    if (!FUSED) return;
    double v[1] = {delta};
    block_partial<1>(v, partial, cap);
}

// Cooperative variant: the NWS wavefronts of a workgroup work through SPW consecutive slices
// together, wavefront v taking slots v, v + NWS, ... of each slice, so a workgroup streams one
// contiguous range of the matrix (a few thousand concurrent streams of one slice per wavefront
// leave the DRAM pages less well used). The partial row sums meet in LDS and are added in
// wavefront order, so the result does not depend on scheduling.
template <typename T, int D, int VP, bool FUSED, int UNR, int NWS, typename IX, bool NT = false>
__global__ __launch_bounds__(64 * NWS) void k_spmv_coop(int64_t n, int64_t nslice, int spw,
                                                         const int64_t *__restrict__ sp, const int32_t *__restrict__ cptr,
                                                         const IX *__restrict__ scol, const T *__restrict__ aval,
                                                         const Vec<T, VP> *__restrict__ u, Vec<T, VP> *__restrict__ w,
                                                         double *partial, int cap, const double *scal, int remap) {
    constexpr int EB = D * D;
    __shared__ T red[NWS][D][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t grp = xcd_block(remap);
    // the first slice's row length and slot range are requested together with the stop flag
    const int64_t kf = grp * spw, rowf = kf * 64 + lane;
    int lenf = 0;
    int64_t basef = 0, endf = 0;
    if (kf < nslice) {
        lenf = rowf < n ? cptr[rowf + 1] - cptr[rowf] : 0;
        basef = sp[kf];
        endf = sp[kf + 1];
    }
    if (FUSED && scal[SC_DONE] != 0.0) return;
    double delta = 0.0;
    for (int t = 0; t < spw; ++t) {
        const int64_t k = grp * spw + t;
        if (k >= nslice) break;  // uniform over the workgroup
        const int64_t row = k * 64 + lane;
        const int len = t == 0 ? lenf : (row < n ? cptr[row + 1] - cptr[row] : 0);
        const int64_t base = t == 0 ? basef : sp[k];
        const int W = (int)(((t == 0 ? endf : sp[k + 1]) - base) >> 6);
        const IX *ci = scol + base + lane;
        const T *av = aval + EB * base + lane;
        T acc[D];
        for (int r = 0; r < D; ++r) acc[r] = (T)0;
        for (int s0 = wv; s0 < W; s0 += NWS * UNR) {
            int j[UNR];
            Vec<T, VP> pj[UNR];
            T e[UNR][EB];
#pragma unroll
            for (int q = 0; q < UNR; ++q) {
                const int s = s0 + q * NWS;
                j[q] = (s < len) ? col_of(ld_stream<NT>(&ci[(int64_t)s * 64]), row) : -1;
            }
#pragma unroll
            for (int q = 0; q < UNR; ++q) {
                const int s = s0 + q * NWS;
                if (j[q] >= 0) {
                    pj[q] = u[j[q]];
#pragma unroll
                    for (int e2 = 0; e2 < EB; ++e2) e[q][e2] = ld_stream<NT>(&av[((int64_t)s * EB + e2) * 64]);
                } else {
                    for (int t2 = 0; t2 < VP; ++t2) pj[q].v[t2] = (T)0;
                    for (int e2 = 0; e2 < EB; ++e2) e[q][e2] = (T)0;
                }
            }
#pragma unroll
            for (int q = 0; q < UNR; ++q)
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) acc[r] += e[q][r * D + c] * pj[q].v[c];
        }
        for (int r = 0; r < D; ++r) red[wv][r][lane] = acc[r];
        __syncthreads();
        if (wv == 0 && row < n) {
            Vec<T, VP> yv;
            for (int t2 = 0; t2 < VP; ++t2) yv.v[t2] = (T)0;
            for (int r = 0; r < D; ++r) {
                T y = (T)0;
                for (int q = 0; q < NWS; ++q) y += red[q][r][lane];
                yv.v[r] = y;
            }
            w[row] = yv;
            if (FUSED) {
                const Vec<T, VP> ui = u[row];
                for (int r = 0; r < D; ++r) delta += (double)ui.v[r] * (double)yv.v[r];
            }
        }
        __syncthreads();  // red is reused by the next slice
    }
    if (!FUSED) return;
    double v[1] = {delta};
    block_partial<1>(v, partial, cap);
}

// ---------------------------------------------------------------- fused scalar step + vector update
// Iteration `it` (the host passes it; every workgroup of the launch sees the same value):
//  1. every workgroup sums the partials of ||r_it||^2, gamma_it = (r_it, u_it) (written by the
//     previous vector kernel) and delta_it = (w_it, u_it) (written by the SpMV) in one fixed
//     order, so all workgroups reach the same stopping decision and the same alpha/beta without
//     a separate finalize launch;
//  2. stop if ||r_it||^2 < thr (it = 0: thr = max(tol^2 ||b||^2, DBL_MIN), zero rhs) or it >= maxit;
//  3. else p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s; u = M^-1 r, and write the
//     partials of ||r||^2 and gamma for iteration it+1 into the other parity slot.
// Workgroup 0 records the scalars; gamma and alpha live in per-parity slots so that the
// workgroups of one launch read the previous iteration's values while workgroup 0 writes its own.
// Returns alpha, beta and the stop code in coef (shared).
__device__ void cg_scalar_step(const double *prr, const double *pg, const double *pd, int np_vec, int np_spmv,
                               double *scal, int it, double tol, double maxit, double *coef) {
    __shared__ double red[3][16];
    double t[3] = {0.0, 0.0, 0.0};
    // the scalars this step reads are requested with the partials: the stop flag and the previous
    // iteration's gamma and alpha (written by an earlier launch; this launch's workgroup 0 writes
    // only the current parity slots, and a stop it records now is the decision every workgroup
    // reaches from the same sums) and the threshold (written at iteration 0, which computes its own)
    const int cur = it & 1, prev = cur ^ 1;
    // both parity slots are read, so that the loads do not wait on it
    double done0 = 0.0, thr0 = 0.0, gam0 = 1.0, gam1 = 1.0, alp0 = 1.0, alp1 = 1.0;
    if (threadIdx.x == 0) {
        done0 = scal[SC_DONE];
        thr0 = scal[SC_THR];
        gam0 = scal[SC_GAMMA_P];
        gam1 = scal[SC_GAMMA_P + 1];
        alp0 = scal[SC_ALPHA_P];
        alp1 = scal[SC_ALPHA_P + 1];
    }
    // Each thread's partials are requested SCALAR_PER at a time before they are added (a loop
    // that adds as it loads waits out one memory latency per partial); the additions keep the
    // ascending order of b, so every workgroup forms the same sums.
    constexpr int SCALAR_PER = 8;
    const int bd = blockDim.x, npmax = np_vec > np_spmv ? np_vec : np_spmv;
    for (int b0 = threadIdx.x; b0 < npmax; b0 += SCALAR_PER * bd) {
        double q[3][SCALAR_PER];
#pragma unroll
        for (int u = 0; u < SCALAR_PER; ++u) {
            const int b = b0 + u * bd;
            q[0][u] = b < np_vec ? prr[b] : 0.0;
            q[1][u] = b < np_vec ? pg[b] : 0.0;
            q[2][u] = b < np_spmv ? pd[b] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < SCALAR_PER; ++u)
            for (int k = 0; k < 3; ++k) t[k] += q[k][u];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        double v = t[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0 && done0 != 0.0) {  // the solve has stopped: change nothing
        coef[0] = coef[1] = 0.0;
        coef[2] = done0;
    } else if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        double tot[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < nw; ++q) tot[k] += red[k][q];
        const double rr = tot[0], gamma = tot[1], delta = tot[2];
        double alpha = 0.0, beta = 0.0, done = 0.0, thr;
        if (it == 0) {
            thr = fmax(tol * tol * rr, DBL_MIN);
            done = (rr == 0.0) ? 3.0 : ((rr < thr) ? 1.0 : (maxit <= 0 ? 2.0 : 0.0));
            alpha = gamma / delta;
        } else {
            thr = thr0;
            if (rr < thr) done = 1.0;
            else if ((double)it >= maxit) done = 2.0;
            else {
                beta = gamma / (prev ? gam1 : gam0);
                alpha = gamma / (delta - beta * gamma / (prev ? alp1 : alp0));
            }
        }
        coef[0] = alpha;
        coef[1] = beta;
        coef[2] = done;
        if (blockIdx.x == 0) {
            if (it == 0) {
                scal[SC_RHS2] = rr;
                scal[SC_THR] = thr;
            }
            scal[SC_RR] = rr;
            scal[SC_ITER] = it;
            scal[SC_GAMMA_P + cur] = gamma;
            scal[SC_ALPHA_P + cur] = alpha;
            scal[SC_DONE] = done;
        }
    }
    __syncthreads();
}

template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_vec(int64_t n, const T *__restrict__ invd, int block_pre,
                                                 Vec<T, VP> *__restrict__ x, Vec<T, VP> *__restrict__ r,
                                                 Vec<T, VP> *__restrict__ u, Vec<T, VP> *__restrict__ p,
                                                 Vec<T, VP> *__restrict__ s, const Vec<T, VP> *__restrict__ w,
                                                 double *partial, int cap, int np_vec, int np_spmv, double *scal, int k) {
    // iteration k of the batch that starts at scal[SC_IT0] (an even iteration: batches have an
    // even number of iterations, so the parity of the partial slots is k's); the scalar step
    // reads the stop flag
    const int it = (int)scal[SC_IT0] + k;
    const double tol = scal[SC_TOL], maxit = scal[SC_MAXIT];
    // the first particle's vectors are requested before the scalar step waits on the partials
    const int64_t i0 = xcd_major_block() * blockDim.x + threadIdx.x;  // XCD x: a contiguous particle range
    Vec<T, VP> p0, s0, x0, r0, u0, w0;
    if (i0 < n) {
        p0 = p[i0];
        s0 = s[i0];
        x0 = x[i0];
        r0 = r[i0];
        u0 = u[i0];
        w0 = w[i0];
    }
    __shared__ double coef[3];  // alpha, beta, done
    const double *pv = partial + (int64_t)(k & 1) * 2 * cap;
    cg_scalar_step(pv, pv + cap, partial + (int64_t)4 * cap, np_vec, np_spmv, scal, it, tol, maxit, coef);
    if (coef[2] != 0.0) return;
    double v[2] = {0.0, 0.0};
    const T alpha = (T)coef[0], beta = (T)coef[1];
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Vec<T, VP> pv = i == i0 ? p0 : p[i], sv = i == i0 ? s0 : s[i], xv = i == i0 ? x0 : x[i], rv = i == i0 ? r0 : r[i], zv;
        const Vec<T, VP> uv = i == i0 ? u0 : u[i], wv = i == i0 ? w0 : w[i];
        for (int k = 0; k < D; ++k) {
            pv.v[k] = uv.v[k] + beta * pv.v[k];
            sv.v[k] = wv.v[k] + beta * sv.v[k];
            xv.v[k] += alpha * pv.v[k];
            rv.v[k] -= alpha * sv.v[k];
        }
        precond_apply<T, D, VP>(invd, block_pre, i, rv, zv);
        p[i] = pv;
        s[i] = sv;
        x[i] = xv;
        r[i] = rv;
        u[i] = zv;
        for (int k = 0; k < D; ++k) {
            v[0] += (double)rv.v[k] * (double)rv.v[k];
            v[1] += (double)rv.v[k] * (double)zv.v[k];
        }
    }
    block_partial<2>(v, partial + (int64_t)((k & 1) ^ 1) * 2 * cap, cap);
}

// Two-level variant, first half: the same scalar step (gamma partials at the fixed offset cap),
// then p, s, x, r updates and the ||r||^2 partials; u = M^-1 r waits for the coarse correction.
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_vec_a(int64_t n, Vec<T, VP> *__restrict__ x, Vec<T, VP> *__restrict__ r,
                                                   const Vec<T, VP> *__restrict__ u, Vec<T, VP> *__restrict__ p,
                                                   Vec<T, VP> *__restrict__ s, const Vec<T, VP> *__restrict__ w,
                                                   double *partial, int cap, int np_vec, int np_spmv, double *scal, int k) {
    const int it = (int)scal[SC_IT0] + k;  // SC_IT0 is even: the parity of the partial slots is k's
    const double tol = scal[SC_TOL], maxit = scal[SC_MAXIT];
    // the first particle's vectors are requested before the scalar step waits on the partials
    const int64_t i0 = xcd_major_block() * blockDim.x + threadIdx.x;  // XCD x: a contiguous particle range
    Vec<T, VP> p0, s0, x0, r0, u0, w0;
    if (i0 < n) {
        p0 = p[i0];
        s0 = s[i0];
        x0 = x[i0];
        r0 = r[i0];
        u0 = u[i0];
        w0 = w[i0];
    }
    __shared__ double coef[3];
    cg_scalar_step(partial + (int64_t)(k & 1) * 2 * cap, partial + cap, partial + (int64_t)4 * cap, np_vec, np_spmv, scal,
                   it, tol, maxit, coef);
    if (coef[2] != 0.0) return;
    double v[1] = {0.0};
    const T alpha = (T)coef[0], beta = (T)coef[1];
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Vec<T, VP> pv = i == i0 ? p0 : p[i], sv = i == i0 ? s0 : s[i], xv = i == i0 ? x0 : x[i], rv = i == i0 ? r0 : r[i];
        const Vec<T, VP> uv = i == i0 ? u0 : u[i], wv = i == i0 ? w0 : w[i];
        for (int k = 0; k < D; ++k) {
            pv.v[k] = uv.v[k] + beta * pv.v[k];
            sv.v[k] = wv.v[k] + beta * sv.v[k];
            xv.v[k] += alpha * pv.v[k];
            rv.v[k] -= alpha * sv.v[k];
        }
        p[i] = pv;
        s[i] = sv;
        x[i] = xv;
        r[i] = rv;
        for (int k = 0; k < D; ++k) v[0] += (double)rv.v[k] * (double)rv.v[k];
    }
    block_partial<1>(v, partial + (int64_t)((k & 1) ^ 1) * 2 * cap, cap);
}

// Two-level variant, second half: u = D^-1 r + P Ac^-1 P^T r (the coarse correction cy is
// already formed) and the gamma = (r, u) partials at offset cap.
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_vec_b(int64_t n, const T *__restrict__ invd, int block_pre,
                                                   const Vec<T, VP> *__restrict__ r, Vec<T, VP> *__restrict__ u,
                                                   const int32_t *__restrict__ agg_of, const double *__restrict__ agg_d,
                                                   const uint8_t *__restrict__ mask, const double *__restrict__ cy,
                                                   double *partial, int cap, const double *scal) {
    constexpr int M = D == 3 ? 6 : 3;  // compile-time, so that P stays in registers
    const int64_t i0 = xcd_major_block() * blockDim.x + threadIdx.x;  // XCD x: a contiguous particle range
    // the first particle's inputs are requested together with the stop flag
    Vec<T, VP> r0;
    double d0[3] = {0.0, 0.0, 0.0};
    int a0 = 0;
    unsigned m0 = 0;
    if (i0 < n) {
        r0 = r[i0];
        a0 = agg_of[i0];
        m0 = mask[i0];
        for (int k = 0; k < D; ++k) d0[k] = agg_d[D * i0 + k];
    }
    if (scal[SC_DONE] != 0.0) return;
    double v[1] = {0.0};
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const Vec<T, VP> rv = i == i0 ? r0 : r[i];
        Vec<T, VP> zv;
        precond_apply<T, D, VP>(invd, block_pre, i, rv, zv);
        double P[3][6];
        if (i == i0) rigid_modes<D>(d0, m0, P);
        else rigid_modes<D>(&agg_d[D * i], mask[i], P);
        const double *y = cy + (int64_t)(i == i0 ? a0 : agg_of[i]) * M;
        for (int k = 0; k < D; ++k) {
            double t = 0.0;
            for (int m = 0; m < M; ++m) t += P[k][m] * y[m];
            zv.v[k] += (T)t;
        }
        u[i] = zv;
        for (int k = 0; k < D; ++k) v[0] += (double)rv.v[k] * (double)zv.v[k];
    }
    block_partial<1>(v, partial + cap, cap);
}

// end of a batch of CG iterations: the next batch starts `batch` iterations later
__global__ __launch_bounds__(64) void k_cg_advance(double *scal, int batch) {
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[SC_IT0] += (double)batch;
}

// ---------------------------------------------------------------- one-block finalize
// Sums per-block partials in a fixed order: FIN_NORMS (||r||^2, ||R||^2) for the Newton loop,
// FIN_DIAG ||diag K||^2 for the Dirichlet elimination.
__global__ __launch_bounds__(1024) void k_finalize(int kind, int na, const double *__restrict__ partial, int cap,
                                                   double *__restrict__ scal) {
    __shared__ double red[2][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nv = (kind == FIN_DIAG) ? 1 : 2;
    for (int k = 0; k < nv; ++k) {
        double t = 0;
        for (int b = threadIdx.x; b < na; b += blockDim.x) t += partial[k * cap + b];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (lane == 0) red[k][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[2] = {0.0, 0.0};
    for (int k = 0; k < nv; ++k)
        for (int w = 0; w < nw; ++w) tot[k] += red[k][w];
    if (kind == FIN_NORMS) {
        scal[SC_NORM_R] = tot[0];
        scal[SC_NORM_REAC] = tot[1];
    } else {
        scal[SC_DIAGNORM] = tot[0];
    }
}

// ---------------------------------------------------------------- preconditioner
// DiagonalPreconditioner: 1/A_dd, or 1 when the diagonal entry is zero or absent; block
// variant: inverse of the (dim x dim) diagonal block (scalar Jacobi where it is singular).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_precond(int64_t n, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                                  const int64_t *__restrict__ sp, const T *__restrict__ aval, int block_pre,
                                                  T *__restrict__ invd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c0 = cptr[i], nc = cptr[i + 1] - c0;
    int lo = 0, hi = nc - 1, ds = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = cidx[c0 + mid];
        if (v == i) { ds = mid; break; }
        if (v < i) lo = mid + 1;
        else hi = mid - 1;
    }
    double B[D][D];
    for (int r = 0; r < D; ++r)
        for (int q = 0; q < D; ++q) B[r][q] = ds >= 0 ? (double)aval[sell_off(sp, i, ds, D * D, r * D + q)] : 0.0;
    if (!block_pre) {
        for (int r = 0; r < D; ++r) invd[D * i + r] = (B[r][r] != 0.0) ? (T)(1.0 / B[r][r]) : (T)1;
        return;
    }
    double inv[D][D];
    double det;
    if (D == 3) {
        inv[0][0] = B[1][1] * B[2][2] - B[1][2] * B[2][1];
        inv[0][1] = B[0][2] * B[2][1] - B[0][1] * B[2][2];
        inv[0][2] = B[0][1] * B[1][2] - B[0][2] * B[1][1];
        inv[1][0] = B[1][2] * B[2][0] - B[1][0] * B[2][2];
        inv[1][1] = B[0][0] * B[2][2] - B[0][2] * B[2][0];
        inv[1][2] = B[0][2] * B[1][0] - B[0][0] * B[1][2];
        inv[2][0] = B[1][0] * B[2][1] - B[1][1] * B[2][0];
        inv[2][1] = B[0][1] * B[2][0] - B[0][0] * B[2][1];
        inv[2][2] = B[0][0] * B[1][1] - B[0][1] * B[1][0];
        det = B[0][0] * inv[0][0] + B[0][1] * inv[1][0] + B[0][2] * inv[2][0];
    } else {
        inv[0][0] = B[1][1];
        inv[0][1] = -B[0][1];
        inv[1][0] = -B[1][0];
        inv[1][1] = B[0][0];
        det = B[0][0] * B[1][1] - B[0][1] * B[1][0];
    }
    const bool ok = det != 0.0 && isfinite(det);
    for (int r = 0; r < D; ++r)
        for (int q = 0; q < D; ++q) {
            double v;
            if (ok) v = inv[r][q] / det;
            else v = (r == q) ? ((B[r][r] != 0.0) ? 1.0 / B[r][r] : 1.0) : 0.0;
            invd[(int64_t)D * D * i + r * D + q] = (T)v;
        }
}

// ---------------------------------------------------------------- position update
// Solver::solveLinearSystem (solver.h:97-105): every particle moves by its DoFs of disp.
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_move(int64_t n, const Vec<T, VP> *__restrict__ x, double *__restrict__ disp,
                                               double *__restrict__ xyz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Vec<T, VP> xv = x[i];
    double d[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        d[k] = (double)xv.v[k];
        disp[D * i + k] = d[k];
    }
    xyz[3 * i] = xyz[3 * i] + d[0];
    xyz[3 * i + 1] = xyz[3 * i + 1] + d[1];
    xyz[3 * i + 2] = xyz[3 * i + 2] + d[2];
}

// ---------------------------------------------------------------- copy of known size
// Streams `n16` 16-byte words from a to b: a kernel of known bytes (n16*16 read, n16*16 written)
// that calibrates hardware traffic counters and gives the practical HBM ceiling.
__global__ __launch_bounds__(256) void k_copy_calib(int64_t n16, const float4 *__restrict__ a, float4 *__restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace

// =================================================================== host side
// Events used only to time kernels: no system-scope fence, so recording one does not write
// back and invalidate the caches (which would both stretch the measured interval and slow
// the kernel that follows).
static hipError_t timing_event(hipEvent_t *e) { return hipEventCreateWithFlags(e, hipEventDisableSystemFence); }

static int check_launch(lpm_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LPM_OK;
}

hipError_t launch_finalize_raw(hipStream_t stream, int kind, int na, const double *partial, int cap, double *scal) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, stream, kind, na, partial, cap, scal);
    return hipGetLastError();
}

int launch_finalize(lpm_ctx *c, int kind, int na) {
    if (na > c->partial_cap) return lpm_fail(c, LPM_E_LIMIT, "reduction exceeds partial buffer");
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, kind, na, (const double *)c->partial,
                       c->partial_cap, c->scal);
    return check_launch(c, "k_finalize");
}

int cg_alloc(lpm_ctx *c) {
    const size_t es = c->precision == 0 ? sizeof(float) : sizeof(double);
    const int64_t ebl = (int64_t)c->dim * c->dim;
    if (c->aval && c->aval_prec != c->precision) {
        hipFree(c->aval);
        hipFree(c->invdiag);
        void *vs[] = {c->v_x, c->v_r, c->v_u, c->v_p, c->v_s, c->v_w};
        for (void *v : vs) hipFree(v);
        c->aval = c->invdiag = c->v_x = c->v_r = c->v_u = c->v_p = c->v_s = c->v_w = nullptr;
    }
    if (!c->aval) {
        const int64_t slots = c->nslots > 0 ? c->nslots : 1;
        LPM_HIP(c, hipMalloc(&c->aval, es * ebl * slots));
        // padding slots are never read (lanes past their row length skip them), but start defined
        LPM_HIP(c, hipMemsetAsync(c->aval, 0, es * ebl * slots, c->stream));
        LPM_HIP(c, hipMalloc(&c->invdiag, es * ebl * c->n));
        const size_t vb = es * c->vpad * c->n;
        void **vs[] = {&c->v_x, &c->v_r, &c->v_u, &c->v_p, &c->v_s, &c->v_w};
        for (void **v : vs) {
            LPM_HIP(c, hipMalloc(v, vb));
            LPM_HIP(c, hipMemsetAsync(*v, 0, vb, c->stream));
        }
        c->aval_prec = c->precision;
    }
    return LPM_OK;
}

template <typename T, int D, int VP>
static int precond_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_precond<T, D>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, c->sp,
                       (const T *)c->aval, block_jacobi(c), (T *)c->invdiag);
    return check_launch(c, "k_precond");
}

// The Jacobi part of the preconditioner depends on K alone: it is formed once per K (assembly
// generation, precision and preconditioner kind), which lets a Newton loop queue it before its
// first wait for the residual.
int launch_precond_setup(lpm_ctx *c) {
    const int key = 8 * c->precision + c->precond;
    if (c->precond_version == c->k_version && c->precond_key == key) return LPM_OK;
    int rc;
    if (c->precision == 0) rc = c->dim == 3 ? precond_t<float, 3, 3>(c) : precond_t<float, 2, 2>(c);
    else rc = c->dim == 3 ? precond_t<double, 3, 3>(c) : precond_t<double, 2, 2>(c);
    if (rc) return rc;
    c->precond_version = c->k_version;
    c->precond_key = key;
    return LPM_OK;
}

// workgroups of the SpMV launch (= partials the fused SpMV leaves)
// int16 column offsets: built (every offset fits) and not switched off
static bool spmv_idx16(const lpm_ctx *c) { return c->scol16 && c->spmv_idx16; }

static int spmv_grid(const lpm_ctx *c) {
    if (c->spmv_coop) return blocks_for(c->nslice, c->spmv_spw);
    return blocks_for(c->nslice, c->spmv_wpb);
}

template <typename T, int D, int VP, bool FUSED, int UNR, int NWS>
static void spmv_coop_go(lpm_ctx *c, const void *in, void *out, double *part) {
    typedef Vec<T, VP> V;
    if (spmv_idx16(c) && c->spmv_nt)
        hipLaunchKernelGGL((k_spmv_coop<T, D, VP, FUSED, UNR, NWS, int16_t, true>), dim3(spmv_grid(c)), dim3(64 * NWS), 0,
                           c->stream, c->n, c->nslice, c->spmv_spw, c->sp, c->cptr, (const int16_t *)c->scol16,
                           (const T *)c->aval, (const V *)in, (V *)out, part, c->partial_cap, (const double *)c->scal,
                           c->spmv_remap);
    else if (spmv_idx16(c))
        hipLaunchKernelGGL((k_spmv_coop<T, D, VP, FUSED, UNR, NWS, int16_t>), dim3(spmv_grid(c)), dim3(64 * NWS), 0, c->stream,
                           c->n, c->nslice, c->spmv_spw, c->sp, c->cptr, (const int16_t *)c->scol16, (const T *)c->aval,
                           (const V *)in, (V *)out, part, c->partial_cap, (const double *)c->scal, c->spmv_remap);
    else
        hipLaunchKernelGGL((k_spmv_coop<T, D, VP, FUSED, UNR, NWS, int32_t>), dim3(spmv_grid(c)), dim3(64 * NWS), 0, c->stream,
                           c->n, c->nslice, c->spmv_spw, c->sp, c->cptr, c->scol, (const T *)c->aval, (const V *)in, (V *)out,
                           part, c->partial_cap, (const double *)c->scal, c->spmv_remap);
}

template <typename T, int D, int VP, bool FUSED, int UNR>
static void spmv_go(lpm_ctx *c, const void *in, void *out, double *part) {
    typedef Vec<T, VP> V;
    const int g = spmv_grid(c), tb = 64 * c->spmv_wpb;
    if (c->spmv_coop) {
        if (c->spmv_coop == 16) spmv_coop_go<T, D, VP, FUSED, (UNR > 1 ? 2 : 1), 16>(c, in, out, part);
        else spmv_coop_go<T, D, VP, FUSED, (UNR > 1 ? 2 : 1), 8>(c, in, out, part);
        return;
    }
    if (spmv_idx16(c) && c->spmv_nt)
        hipLaunchKernelGGL((k_spmv<T, D, VP, FUSED, UNR, int16_t, true>), dim3(g), dim3(tb), 0, c->stream, c->n, c->nslice,
                           c->sp, c->cptr, (const int16_t *)c->scol16, (const T *)c->aval, (const V *)in, (V *)out, part,
                           c->partial_cap, (const double *)c->scal, c->spmv_remap);
    else if (spmv_idx16(c))
        hipLaunchKernelGGL((k_spmv<T, D, VP, FUSED, UNR, int16_t>), dim3(g), dim3(tb), 0, c->stream, c->n, c->nslice, c->sp,
                           c->cptr, (const int16_t *)c->scol16, (const T *)c->aval, (const V *)in, (V *)out, part,
                           c->partial_cap, (const double *)c->scal, c->spmv_remap);
    else
        hipLaunchKernelGGL((k_spmv<T, D, VP, FUSED, UNR, int32_t>), dim3(g), dim3(tb), 0, c->stream, c->n, c->nslice, c->sp,
                           c->cptr, c->scol, (const T *)c->aval, (const V *)in, (V *)out, part, c->partial_cap,
                           (const double *)c->scal, c->spmv_remap);
}

// w = A u (fused: + delta partials, skipped once the solve is done)
template <typename T, int D, int VP>
static int spmv_launch(lpm_ctx *c, bool fused, void *in, void *out) {
    double *part = fused ? c->partial + 4 * (int64_t)c->partial_cap : c->partial;
    switch (c->spmv_unr) {
    case 1:
        fused ? spmv_go<T, D, VP, true, 1>(c, in, out, part) : spmv_go<T, D, VP, false, 1>(c, in, out, part);
        break;
    case 4:
        fused ? spmv_go<T, D, VP, true, 4>(c, in, out, part) : spmv_go<T, D, VP, false, 4>(c, in, out, part);
        break;
    default:
        fused ? spmv_go<T, D, VP, true, 2>(c, in, out, part) : spmv_go<T, D, VP, false, 2>(c, in, out, part);
    }
    return check_launch(c, c->spmv_coop ? "k_spmv_coop" : "k_spmv");
}

// Launch-configuration tuner: times the fused SpMV the CG runs (delta partials into the CG's
// partial buffer; v_p, v_w and the partials are not live outside a solve) in every configuration
// and keeps the fastest.
template <typename T, int D, int VP>
static int spmv_tune_t(lpm_ctx *c) {
    // 4 or 8 wavefronts per workgroup: the fused SpMV leaves one partial per workgroup, which
    // every workgroup of the next vector kernel sums, so fewer workgroups keep that cheap
    // per-wavefront slices: unroll 1, 2, 4 x 4 or 8 wavefronts per workgroup; cooperative
    // slices: 8 or 16 wavefronts per slice x unroll 1, 2 x 2 or 4 slices per workgroup; each with
    // and without the XCD remap
    struct Cfg {
        int coop, unr, wpb, spw;
    };
    std::vector<Cfg> cfgs;
    for (int un : {1, 2, 4})
        for (int wp : {4, 8}) cfgs.push_back({0, un, wp, 1});
    for (int nw : {8, 16})
        for (int un : {1, 2})
            for (int sw : {2, 4}) cfgs.push_back({nw, un, c->spmv_wpb, sw});
    hipEvent_t a, b;
    LPM_HIP(c, timing_event(&a));
    LPM_HIP(c, timing_event(&b));
    const int reps = 6;
    // the fused kernel returns at once while the stop flag of the last solve is still set
    LPM_HIP(c, hipMemsetAsync(c->scal + SC_DONE, 0, sizeof(double), c->stream));
    // Two passes over all configurations; each keeps the faster of its two timings, so that one
    // noisy timing does not decide between configurations a few percent apart. The launches
    // alternate between v_p -> v_w and v_w -> v_p, so that each reads a vector the launch before
    // it has just written, as the CG's SpMV reads the vector k_cg_vec_b has just written.
    const size_t ncand = 2 * cfgs.size();
    std::vector<double> t(ncand, 1e300);
    int rc = LPM_OK;
    for (int pass = 0; pass < 2 && !rc; ++pass)
        for (size_t q = 0; q < ncand && !rc; ++q) {
            const Cfg &cf = cfgs[q / 2];
            c->spmv_coop = cf.coop;
            c->spmv_unr = cf.unr;
            c->spmv_wpb = cf.wpb;
            c->spmv_spw = cf.spw;
            c->spmv_remap = (int)(q % 2);
            if (spmv_grid(c) > c->partial_cap) continue;
            if ((rc = spmv_launch<T, D, VP>(c, true, c->v_p, c->v_w))) break;
            if (hipEventRecord(a, c->stream) != hipSuccess) rc = lpm_fail(c, LPM_E_HIP, "SpMV tuner event");
            for (int k = 0; k < reps && !rc; ++k)
                rc = spmv_launch<T, D, VP>(c, true, k % 2 ? c->v_p : c->v_w, k % 2 ? c->v_w : c->v_p);
            float ms = 0;
            if (!rc && (hipEventRecord(b, c->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                        hipEventElapsedTime(&ms, a, b) != hipSuccess))
                rc = lpm_fail(c, LPM_E_HIP, "SpMV tuner event");
            if (!rc) t[q] = std::min(t[q], (double)ms);
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc) return rc;
    // The fastest one-wavefront-per-slice configuration is kept unless a cooperative one beats it
    // by more than 3%: timed alone the two kinds come within a few percent of each other at SC
    // 64^3 and the pick between them follows the noise, while inside the CG the one-wavefront
    // kind is consistently faster (95.8 against 94.9 us per SpMV, 129.7 against 131.4 load
    // steps/s over five runs each). In 2D the cooperative kind wins by 15% and is kept.
    size_t bp = ncand, bc = ncand;  // the fastest of each kind (ncand: none timed)
    for (size_t q = 0; q < ncand; ++q) {
        size_t &best = cfgs[q / 2].coop ? bc : bp;
        if (best == ncand || t[q] < t[best]) best = q;
    }
    const size_t bq = (bc < ncand && t[bc] < 1e300 && (bp == ncand || t[bc] < 0.97 * t[bp])) ? bc : bp;
    const Cfg &cf = cfgs[bq / 2];
    c->spmv_coop = cf.coop;
    c->spmv_unr = cf.unr;
    c->spmv_wpb = cf.wpb;
    c->spmv_spw = cf.spw;
    c->spmv_remap = (int)(bq % 2);
    c->tune_best_ms = t[bq] / reps;
    return LPM_OK;
}

static int spmv_tune(lpm_ctx *c) {
    const int key = c->precision;
    if (!c->spmv_tune || c->tune_key == key) return LPM_OK;
    int rc;
    if (c->precision == 0) rc = c->dim == 3 ? spmv_tune_t<float, 3, 3>(c) : spmv_tune_t<float, 2, 2>(c);
    else rc = c->dim == 3 ? spmv_tune_t<double, 3, 3>(c) : spmv_tune_t<double, 2, 2>(c);
    if (rc) return rc;
    c->tune_key = key;
    return LPM_OK;
}

// Captures one batch of CG iterations into c->cg_exec unless the graph captured before is
// still valid: the key lists every value the batch's launches take (sizes, grids, the tuned
// SpMV configuration and every buffer), so any change re-captures.
template <typename F>
static int cg_graph_prepare(lpm_ctx *c, int gv, int gs, int batch, F issue) {
    const std::vector<uintptr_t> key = {
        (uintptr_t)c->n, (uintptr_t)c->dim, (uintptr_t)c->precision, (uintptr_t)c->precond, (uintptr_t)gv, (uintptr_t)gs,
        (uintptr_t)batch, (uintptr_t)c->nslice, (uintptr_t)c->spmv_unr, (uintptr_t)c->spmv_wpb, (uintptr_t)c->spmv_coop,
        (uintptr_t)c->spmv_spw, (uintptr_t)c->spmv_remap, (uintptr_t)c->spmv_nt, (uintptr_t)c->partial_cap,
        (uintptr_t)c->sp, (uintptr_t)c->cptr, (uintptr_t)c->scol, (uintptr_t)c->scol16, (uintptr_t)c->spmv_idx16, (uintptr_t)c->aval,
        (uintptr_t)c->v_x, (uintptr_t)c->v_r, (uintptr_t)c->v_u, (uintptr_t)c->v_p, (uintptr_t)c->v_s, (uintptr_t)c->v_w,
        (uintptr_t)c->invdiag, (uintptr_t)c->partial, (uintptr_t)c->scal, (uintptr_t)c->agg_of, (uintptr_t)c->agg_d,
        (uintptr_t)c->agg_ptr, (uintptr_t)c->agg_idx, (uintptr_t)c->dofmask, (uintptr_t)c->cz, (uintptr_t)c->cy, (uintptr_t)c->agg_dm, (uintptr_t)c->agg_max,
        (uintptr_t)c->Ainv, (uintptr_t)c->nagg, (uintptr_t)c->nmode, (uintptr_t)c->cgrp_cb,
        (uintptr_t)c->cgrp_off, (uintptr_t)c->ncgrp, (uintptr_t)c->cgrp_nbmax};
    if (c->cg_exec && key == c->cg_key) return LPM_OK;
    if (c->cg_exec) {
        hipGraphExecDestroy(c->cg_exec);
        c->cg_exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    LPM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = issue();
    const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    if (rc || ce != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        return rc ? rc : lpm_fail(c, LPM_E_HIP, std::string("CG batch capture: ") + hipGetErrorString(ce));
    }
    const hipError_t ie = hipGraphInstantiate(&c->cg_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        c->cg_exec = nullptr;
        return lpm_fail(c, LPM_E_HIP, std::string("CG batch graph: ") + hipGetErrorString(ie));
    }
    c->cg_key = key;
    return LPM_OK;
}

template <typename T, int D, int VP>
static int cg_t(lpm_ctx *c, int *iters, double *error, int *success) {
    typedef Vec<T, VP> V;
    const int64_t n = c->n;
    // vector kernels: grid-stride over at most CG_VEC_WG workgroups (4 per CU: few partials to sum)
    const int gv = std::max(1, std::min(blocks_for(n, 256), CG_VEC_WG)), gs = spmv_grid(c);
    if (gv > c->partial_cap || gs > c->partial_cap) return lpm_fail(c, LPM_E_LIMIT, "CG grid exceeds partial buffer");
    const int64_t maxit = c->cg_maxit < 0 ? 2 * n * c->dim : c->cg_maxit;
    hipLaunchKernelGGL((k_cg_init<T, D, VP>), dim3(gv), dim3(256), 0, c->stream, n, c->residual, (const T *)c->invdiag,
                       block_jacobi(c), (V *)c->v_x, (V *)c->v_r, (V *)c->v_u, (V *)c->v_p, (V *)c->v_s, c->partial,
                       c->partial_cap, c->scal, c->cg_tol, (double)maxit);
    int rc = check_launch(c, "k_cg_init");
    if (rc) return rc;
    const bool coarse = two_level(c);
    // two-level: u = D^-1 r + P Ac^-1 P^T r (restriction + coarse solve, then the fused update)
    auto coarse_u = [&]() -> int {
        int r2 = launch_coarse_apply(c);
        if (r2) return r2;
        hipLaunchKernelGGL((k_cg_vec_b<T, D, VP>), dim3(gv), dim3(256), 0, c->stream, n, (const T *)c->invdiag, block_jacobi(c),
                           (const V *)c->v_r, (V *)c->v_u, c->agg_of, c->agg_d, c->dofmask, (const double *)c->cy,
                           c->partial, c->partial_cap, (const double *)c->scal);
        return check_launch(c, "k_cg_vec_b");
    };
    if (coarse && (rc = coarse_u())) return rc;
    rc = spmv_launch<T, D, VP>(c, true, c->v_u, c->v_w);
    if (rc) return rc;
    // Batches of iterations between reads of the device's stop flag (even sizes, so that a batch
    // starts on parity 0); the flag turns the rest of a batch into no-ops once the solve has
    // stopped, so the sizes change launches only, never results. Batches of cg_batch; under
    // cg_batch_auto a Newton solve whose index has a recorded count starts with a batch covering
    // that count + 2, followed by two batches of cg_batch_next (the few iterations it may lack)
    // and batches of cg_batch after those. Graph mode replays one captured batch size throughout.
    const bool graph = c->cg_graph != 0;
    const int batch_set = ((c->cg_batch > 0 ? c->cg_batch : 16) + 1) & ~1;
    const int slot = c->cg_slot;
    const bool predicted = !graph && c->cg_batch_auto && slot >= 0 && slot < 4 && c->cg_hist[slot] >= 0;
    const int batch = predicted ? (std::min(std::max(c->cg_hist[slot] + 2, 4), 512) + 1) & ~1 : batch_set;
    const int batch_next = predicted && c->cg_batch_next > 0 ? (c->cg_batch_next + 1) & ~1 : batch_set;
    const int batch_max = std::max({batch, batch_next, batch_set});
    // timing: events around the SpMV of every `timing`-th iteration only (an event record costs
    // a few microseconds of idle GPU time, so timing every launch would slow the solve); not in
    // graph mode, whose batches are replayed whole. One pair per timed iteration of a batch.
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events() {
            for (auto &x : e) hipEventDestroy(x);
        }
    } evs;
    if (c->timing && !graph) {
        evs.e.resize(2 * (batch_max / c->timing + 2));
        for (auto &e : evs.e) LPM_HIP(c, timing_event(&e));
    }
    auto timed = [&](int64_t it) { return !graph && c->timing > 0 && it % c->timing == 0; };
    int64_t it0 = 0;
    // one batch of nb iterations: iteration k of the batch starting at scal[SC_IT0], then the
    // batch base moves on
    auto issue = [&](int64_t base, int nb) -> int {
        int r = LPM_OK, te = 0;
        for (int k = 0; k < nb && !r; ++k) {
            if (coarse) {
                hipLaunchKernelGGL((k_cg_vec_a<T, D, VP>), dim3(gv), dim3(256), 0, c->stream, n, (V *)c->v_x, (V *)c->v_r,
                                   (const V *)c->v_u, (V *)c->v_p, (V *)c->v_s, (const V *)c->v_w, c->partial, c->partial_cap,
                                   gv, gs, c->scal, k);
                r = check_launch(c, "k_cg_vec_a");
                if (!r) r = coarse_u();
            } else {
                hipLaunchKernelGGL((k_cg_vec<T, D, VP>), dim3(gv), dim3(256), 0, c->stream, n, (const T *)c->invdiag,
                                   block_jacobi(c), (V *)c->v_x, (V *)c->v_r, (V *)c->v_u, (V *)c->v_p, (V *)c->v_s,
                                   (const V *)c->v_w, c->partial, c->partial_cap, gv, gs, c->scal, k);
                r = check_launch(c, "k_cg_vec");
            }
            if (r) break;
            const bool tk = timed(base + k);
            if (tk && hipEventRecord(evs.e[2 * te], c->stream) != hipSuccess) r = lpm_fail(c, LPM_E_HIP, "event");
            if (!r) r = spmv_launch<T, D, VP>(c, true, c->v_u, c->v_w);
            if (!r && tk && hipEventRecord(evs.e[2 * te + 1], c->stream) != hipSuccess) r = lpm_fail(c, LPM_E_HIP, "event");
            if (tk) ++te;
        }
        if (r) return r;
        hipLaunchKernelGGL(k_cg_advance, dim3(1), dim3(64), 0, c->stream, c->scal, nb);
        return check_launch(c, "k_cg_advance");
    };
    if (graph && (rc = cg_graph_prepare(c, gv, gs, batch, [&]() { return issue(0, batch); }))) return rc;
    for (int nb = batch, nbatch = 1;; nb = nbatch++ < 3 ? batch_next : batch_set) {
        if (graph) LPM_HIP(c, hipGraphLaunch(c->cg_exec, c->stream));
        else if ((rc = issue(it0, nb))) return rc;
        LPM_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        // SpMVs of the batch that did work: one per update (iterations it0 .. ITER-1 updated)
        const bool done = c->h_scal[SC_DONE] != 0.0;
        int real = done ? (int)(c->h_scal[SC_ITER] - (double)it0) : nb;
        real = real < 0 ? 0 : (real > nb ? nb : real);
        c->st_spmv += real;
        for (int k = 0, te = 0; k < nb; ++k) {
            if (!timed(it0 + k)) continue;
            if (k < real) {
                float ms = 0;
                LPM_HIP(c, hipEventElapsedTime(&ms, evs.e[2 * te], evs.e[2 * te + 1]));
                c->st_spmv_ms += ms;
                c->st_spmv_timed++;
            }
            ++te;
        }
        if (done) break;
        it0 += nb;
        if (it0 > maxit + 2 * batch_max) break;  // the device cap stops earlier; this only guards the loop
    }
    c->st_spmv += 1;  // the prologue SpMV
    const double rhs2 = c->h_scal[SC_RHS2];
    const double done = c->h_scal[SC_DONE];
    // Eigen reports the updates done before the converging one (its loop breaks before i++)
    const int64_t upd = (int64_t)c->h_scal[SC_ITER];
    if (slot >= 0 && slot < 4) c->cg_hist[slot] = (int)std::min<int64_t>(upd, 1 << 20);
    *iters = (int)((done == 1.0 && upd > 0) ? upd - 1 : upd);
    *error = (done == 3.0) ? 0.0 : sqrt(c->h_scal[SC_RR] / rhs2);
    *success = (done == 3.0) || (*error <= c->cg_tol);
    return LPM_OK;
}

int cg_solve(lpm_ctx *c, int *iters, double *error, int *success) {
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "solve before the stiffness matrix was assembled");
    int rc = launch_precond_setup(c);
    if (rc) return rc;
    if ((rc = coarse_prepare(c))) return rc;
    if ((rc = spmv_tune(c))) return rc;
    if (c->precision == 0)
        rc = c->dim == 3 ? cg_t<float, 3, 3>(c, iters, error, success) : cg_t<float, 2, 2>(c, iters, error, success);
    else
        rc = c->dim == 3 ? cg_t<double, 3, 3>(c, iters, error, success) : cg_t<double, 2, 2>(c, iters, error, success);
    if (!rc && two_level(c) && c->coarse_reuse) {
        // a kept coarse space is rebuilt before the next solve once the CG needs clearly more
        // iterations than the first solve after the last build
        if (c->coarse_ref_iters < 0) c->coarse_ref_iters = *iters;
        else if (*iters > 1.25 * c->coarse_ref_iters + 4) c->coarse_stale = true;
    }
    return rc;
}

// x_i *= w[specimen of i] (undoes the per-specimen right-hand-side scaling of an ensemble solve)
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_scale_x(int64_t n, const int32_t *__restrict__ spec_of, const double *__restrict__ w,
                                                  Vec<T, VP> *__restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = w[spec_of[i]];
    Vec<T, VP> v = x[i];
    for (int k = 0; k < D; ++k) v.v[k] = (T)((double)v.v[k] * f);
    x[i] = v;
}

template <typename T, int D, int VP>
static int scale_x_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_scale_x<T, D, VP>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       (const int32_t *)c->spec_of, (const double *)c->spec_w, (Vec<T, VP> *)c->v_x);
    return check_launch(c, "k_scale_x");
}

int launch_scale_x(lpm_ctx *c) {
    if (c->precision == 0) return c->dim == 3 ? scale_x_t<float, 3, 3>(c) : scale_x_t<float, 2, 2>(c);
    return c->dim == 3 ? scale_x_t<double, 3, 3>(c) : scale_x_t<double, 2, 2>(c);
}

template <typename T, int D, int VP>
static int move_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_move<T, D, VP>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       (const Vec<T, VP> *)c->v_x, c->disp, c->xyz);
    return check_launch(c, "k_move");
}

int launch_move(lpm_ctx *c) {
    if (c->precision == 0) return c->dim == 3 ? move_t<float, 3, 3>(c) : move_t<float, 2, 2>(c);
    return c->dim == 3 ? move_t<double, 3, 3>(c) : move_t<double, 2, 2>(c);
}

template <typename T, int D, int VP>
static int bench_t(lpm_ctx *c, int reps, double *ms) {
    hipEvent_t a, b;
    LPM_HIP(c, timing_event(&a));
    LPM_HIP(c, timing_event(&b));
    int rc = spmv_launch<T, D, VP>(c, false, c->v_p, c->v_w);  // warm
    if (rc) return rc;
    LPM_HIP(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < reps; ++k) {
        rc = spmv_launch<T, D, VP>(c, false, c->v_p, c->v_w);
        if (rc) return rc;
    }
    LPM_HIP(c, hipEventRecord(b, c->stream));
    LPM_HIP(c, hipEventSynchronize(b));
    float t = 0;
    LPM_HIP(c, hipEventElapsedTime(&t, a, b));
    *ms = t / reps;
    hipEventDestroy(a);
    hipEventDestroy(b);
    return LPM_OK;
}

int spmv_bench(lpm_ctx *c, int reps, double *ms) {
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "no assembled matrix");
    if (reps <= 0) return lpm_fail(c, LPM_E_ARG, "reps must be positive");
    const int rc = spmv_tune(c);
    if (rc) return rc;
    if (c->precision == 0) return c->dim == 3 ? bench_t<float, 3, 3>(c, reps, ms) : bench_t<float, 2, 2>(c, reps, ms);
    return c->dim == 3 ? bench_t<double, 3, 3>(c, reps, ms) : bench_t<double, 2, 2>(c, reps, ms);
}

int copy_bench(lpm_ctx *c, int64_t bytes, int reps, double *ms) {
    if (bytes < 16 || reps <= 0) return lpm_fail(c, LPM_E_ARG, "copy_bench needs bytes >= 16 and reps > 0");
    const int64_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    LPM_HIP(c, hipMalloc(&a, n16 * 16));
    if (hipMalloc(&b, n16 * 16) != hipSuccess) {
        hipFree(a);
        return lpm_fail(c, LPM_E_NOMEM, "copy_bench allocation");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LPM_OK;
    const int g = (int)std::min<int64_t>(blocks_for(n16, 256), 8192);
    if (hipMemsetAsync(a, 0, n16 * 16, c->stream) != hipSuccess || timing_event(&e0) != hipSuccess ||
        timing_event(&e1) != hipSuccess) {
        rc = lpm_fail(c, LPM_E_HIP, "copy_bench setup");
    } else {
        hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, c->stream, n16, (const float4 *)a, (float4 *)b);
        hipEventRecord(e0, c->stream);
        for (int k = 0; k < reps; ++k)
            hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, c->stream, n16, (const float4 *)a, (float4 *)b);
        hipEventRecord(e1, c->stream);
        rc = check_launch(c, "k_copy_calib");
        float t = 0;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = lpm_fail(c, LPM_E_HIP, "copy_bench timing");
        *ms = t / reps;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = lpm_fail(c, LPM_E_HIP, "stream synchronisation");
    hipFree(a);
    hipFree(b);
    return rc;
}
