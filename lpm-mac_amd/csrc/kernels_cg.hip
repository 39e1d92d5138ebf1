// Block-CSR SpMV and the Jacobi-preconditioned conjugate gradient that replaces
// Eigen::ConjugateGradient<SparseMatrix<double>, Lower|Upper> (solver.h:192-217).
//
// The CG follows Eigen 3.4's conjugate_gradient(): x0 = 0, r = b, p = M^-1 r,
// threshold = max(tol^2 ||b||^2, DBL_MIN), stop when ||r||^2 < threshold (the
// breaking iteration is not counted) or after maxIters iterations. Scalars
// (dot products, alpha, beta) are FP64 in device memory; vectors and the matrix
// are FP32 (precision 0) or FP64 (precision 1). No host synchronisation happens
// inside an iteration: each kernel reads the scalars it needs from device memory
// and the "done" flag makes the remaining kernels of a batch exit at once.
//
// Vectors are stored padded to VP entries per particle (VP = 4 in 3D, 2 in 2D) so
// that the SpMV gathers one aligned 16-byte (FP32, 3D) load per block.
#include <cfloat>

#include "lpm_internal.h"

namespace {

template <typename T, int VP>
struct alignas(sizeof(T) * VP) Vec {
    T v[VP];
};

__device__ __forceinline__ double agent_load(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block sum of NV doubles; the last block to finish reduces the per-block partials in a
// fixed order and returns true with the totals (valid in every thread of that block).
template <int NV>
__device__ bool grid_reduce(double (&v)[NV], double *partial, int cap, unsigned int *counter, double (&out)[NV]) {
    __shared__ double red[NV][8];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; ++k) {
            double s = 0;
            for (int w = 0; w < nw; ++w) s += red[k][w];
            partial[k * cap + blockIdx.x] = s;
        }
        __threadfence();
        const unsigned int t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int k = 0; k < NV; ++k) {
        double s = 0;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) s += agent_load(&partial[k * cap + b]);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        __syncthreads();
        if (lane == 0) red[k][wid] = s;
        __syncthreads();
        double t = 0;
        for (int w = 0; w < nw; ++w) t += red[k][w];
        out[k] = t;
    }
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD works on
// a contiguous range of rows (its L2 then holds the matching window of p).
__device__ __forceinline__ int64_t xcd_block(int remap) {
    const int64_t b = blockIdx.x, g = gridDim.x;
    if (!remap || (g & 7)) return b;
    return (b & 7) * (g >> 3) + (b >> 3);
}

// ---------------------------------------------------------------- CG init
// r = b (x0 = 0), p = M^-1 r, absNew = r.p, threshold (Eigen conjugate_gradient prologue)
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_init(int64_t n, const double *__restrict__ b64, const T *__restrict__ invd,
                                                  int block_pre, Vec<T, VP> *__restrict__ x, Vec<T, VP> *__restrict__ r,
                                                  Vec<T, VP> *__restrict__ p, double tol, double maxit, double *partial,
                                                  int cap, unsigned int *counter, double *scal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        Vec<T, VP> rv, zv, xv;
        for (int k = 0; k < VP; ++k) { rv.v[k] = (T)0; zv.v[k] = (T)0; xv.v[k] = (T)0; }
        for (int k = 0; k < D; ++k) rv.v[k] = (T)b64[D * i + k];
        if (block_pre) {
            for (int a = 0; a < D; ++a) {
                T s = 0;
                for (int q = 0; q < D; ++q) s += invd[(int64_t)D * D * i + a * D + q] * rv.v[q];
                zv.v[a] = s;
            }
        } else {
            for (int k = 0; k < D; ++k) zv.v[k] = invd[D * i + k] * rv.v[k];
        }
        x[i] = xv;
        r[i] = rv;
        p[i] = zv;
        for (int k = 0; k < D; ++k) {
            v[0] += (double)rv.v[k] * (double)rv.v[k];
            v[1] += (double)rv.v[k] * (double)zv.v[k];
        }
    }
    double out[2];
    if (grid_reduce<2>(v, partial, cap, counter, out) && threadIdx.x == 0) {
        const double rhs2 = out[0];
        scal[SC_RHS2] = rhs2;
        scal[SC_RR] = rhs2;
        scal[SC_ITER] = 0;
        scal[SC_MAXIT] = maxit;
        const double thr = fmax(tol * tol * rhs2, DBL_MIN);
        scal[SC_THR] = thr;
        scal[SC_ABSNEW] = out[1];
        scal[SC_DONE] = (rhs2 == 0.0) ? 3.0 : ((rhs2 < thr) ? 1.0 : (maxit <= 0 ? 2.0 : 0.0));
    }
}

// ---------------------------------------------------------------- SpMV  t = A p,  pt = p.t
// One wavefront per block row, one lane per block (slot); a row with more than 64 blocks is
// walked in chunks of 64. The matrix stream uses non-temporal loads so that it does not evict
// the gathered vector from L2.
template <typename T, int D, int VP, bool FUSED>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                               const T *__restrict__ aval, const Vec<T, VP> *__restrict__ p,
                                               Vec<T, VP> *__restrict__ t, double *partial, int cap, unsigned int *counter,
                                               double *scal, int remap) {
    if (FUSED && scal[SC_DONE] != 0.0) return;
    const int lane = threadIdx.x & 63;
    const int64_t i = xcd_block(remap) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    double pt = 0.0;
    if (i < n) {
        const int c0 = cptr[i], nc = cptr[i + 1] - c0;
        const T *a = aval + (int64_t)D * D * c0;
        T acc[D];
        for (int r = 0; r < D; ++r) acc[r] = (T)0;
        for (int slot = lane; slot < nc; slot += 64) {
            const int64_t j = cidx[c0 + slot];
            const Vec<T, VP> pj = p[j];
            T e[D * D];
            for (int q = 0; q < D * D; ++q) e[q] = __builtin_nontemporal_load(a + (int64_t)q * nc + slot);
            for (int r = 0; r < D; ++r)
                for (int q = 0; q < D; ++q) acc[r] += e[r * D + q] * pj.v[q];
        }
        for (int r = 0; r < D; ++r)
            for (int o = 32; o > 0; o >>= 1) acc[r] += __shfl_xor(acc[r], o, 64);
        if (lane == 0) {
            Vec<T, VP> tv;
            for (int k = 0; k < VP; ++k) tv.v[k] = (T)0;
            for (int r = 0; r < D; ++r) tv.v[r] = acc[r];
            t[i] = tv;
            if (FUSED) {
                const Vec<T, VP> pi = p[i];
                for (int r = 0; r < D; ++r) pt += (double)pi.v[r] * (double)acc[r];
            }
        }
    }
    if (!FUSED) return;
    double v[1] = {pt}, out[1];
    if (grid_reduce<1>(v, partial, cap, counter, out) && threadIdx.x == 0) {
        scal[SC_PT] = out[0];
        scal[SC_ALPHA] = scal[SC_ABSNEW] / out[0];
    }
}

// ---------------------------------------------------------------- x += a p, r -= a t, ||r||^2, r.M^-1 r
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, const T *__restrict__ invd, int block_pre,
                                                    Vec<T, VP> *__restrict__ x, Vec<T, VP> *__restrict__ r,
                                                    const Vec<T, VP> *__restrict__ p, const Vec<T, VP> *__restrict__ t,
                                                    double *partial, int cap, unsigned int *counter, double *scal) {
    if (scal[SC_DONE] != 0.0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double alpha = scal[SC_ALPHA];
    double v[2] = {0.0, 0.0};
    if (i < n) {
        Vec<T, VP> xv = x[i], rv = r[i];
        const Vec<T, VP> pv = p[i], tv = t[i];
        const T a = (T)alpha;
        for (int k = 0; k < D; ++k) {
            xv.v[k] += a * pv.v[k];
            rv.v[k] -= a * tv.v[k];
        }
        x[i] = xv;
        r[i] = rv;
        T z[D];
        if (block_pre) {
            for (int q = 0; q < D; ++q) {
                T s = 0;
                for (int w = 0; w < D; ++w) s += invd[(int64_t)D * D * i + q * D + w] * rv.v[w];
                z[q] = s;
            }
        } else {
            for (int k = 0; k < D; ++k) z[k] = invd[D * i + k] * rv.v[k];
        }
        for (int k = 0; k < D; ++k) {
            v[0] += (double)rv.v[k] * (double)rv.v[k];
            v[1] += (double)rv.v[k] * (double)z[k];
        }
    }
    double out[2];
    if (grid_reduce<2>(v, partial, cap, counter, out) && threadIdx.x == 0) {
        scal[SC_RR] = out[0];
        if (out[0] < scal[SC_THR]) {
            scal[SC_DONE] = 1.0;
        } else {
            const double absOld = scal[SC_ABSNEW];
            scal[SC_ABSNEW] = out[1];
            scal[SC_BETA] = out[1] / absOld;
            const double it = scal[SC_ITER] + 1.0;
            scal[SC_ITER] = it;
            if (it >= scal[SC_MAXIT]) scal[SC_DONE] = 2.0;
        }
    }
}

// ---------------------------------------------------------------- p = M^-1 r + beta p
template <typename T, int D, int VP>
__global__ __launch_bounds__(256) void k_cg_direction(int64_t n, const T *__restrict__ invd, int block_pre,
                                                       const Vec<T, VP> *__restrict__ r, Vec<T, VP> *__restrict__ p,
                                                       const double *scal) {
    if (scal[SC_DONE] != 0.0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T beta = (T)scal[SC_BETA];
    const Vec<T, VP> rv = r[i];
    Vec<T, VP> pv = p[i];
    if (block_pre) {
        for (int q = 0; q < D; ++q) {
            T s = 0;
            for (int w = 0; w < D; ++w) s += invd[(int64_t)D * D * i + q * D + w] * rv.v[w];
            pv.v[q] = s + beta * pv.v[q];
        }
    } else {
        for (int k = 0; k < D; ++k) pv.v[k] = invd[D * i + k] * rv.v[k] + beta * pv.v[k];
    }
    p[i] = pv;
}

// ---------------------------------------------------------------- preconditioner
// DiagonalPreconditioner: 1/A_dd, or 1 when the diagonal entry is zero or absent; block
// variant: inverse of the (dim x dim) diagonal block (scalar Jacobi where it is singular).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_precond(int64_t n, const int32_t *__restrict__ cptr, const int32_t *__restrict__ cidx,
                                                  const T *__restrict__ aval, int block_pre, T *__restrict__ invd) {
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
        for (int q = 0; q < D; ++q)
            B[r][q] = ds >= 0 ? (double)aval[(int64_t)D * D * c0 + (int64_t)(r * D + q) * nc + ds] : 0.0;
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

}  // namespace

// =================================================================== host side
static int check_launch(lpm_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpm_fail(c, LPM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LPM_OK;
}

int cg_alloc(lpm_ctx *c) {
    const size_t es = c->precision == 0 ? sizeof(float) : sizeof(double);
    const int64_t ebl = (int64_t)c->dim * c->dim;
    if (c->aval && c->aval_prec != c->precision) {
        hipFree(c->aval); hipFree(c->invdiag); hipFree(c->v_x); hipFree(c->v_r); hipFree(c->v_p); hipFree(c->v_t);
        c->aval = c->invdiag = c->v_x = c->v_r = c->v_p = c->v_t = nullptr;
    }
    if (!c->aval) {
        LPM_HIP(c, hipMalloc(&c->aval, es * ebl * (c->nconn > 0 ? c->nconn : 1)));
        LPM_HIP(c, hipMalloc(&c->invdiag, es * ebl * c->n));
        const size_t vb = es * c->vpad * c->n;
        LPM_HIP(c, hipMalloc(&c->v_x, vb));
        LPM_HIP(c, hipMalloc(&c->v_r, vb));
        LPM_HIP(c, hipMalloc(&c->v_p, vb));
        LPM_HIP(c, hipMalloc(&c->v_t, vb));
        LPM_HIP(c, hipMemsetAsync(c->v_p, 0, vb, c->stream));
        c->aval_prec = c->precision;
    }
    return LPM_OK;
}

template <typename T, int D, int VP>
static int precond_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_precond<T, D>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx,
                       (const T *)c->aval, c->precond, (T *)c->invdiag);
    return check_launch(c, "k_precond");
}

int launch_precond_setup(lpm_ctx *c) {
    if (c->precision == 0) return c->dim == 3 ? precond_t<float, 3, 4>(c) : precond_t<float, 2, 2>(c);
    return c->dim == 3 ? precond_t<double, 3, 4>(c) : precond_t<double, 2, 2>(c);
}

template <typename T, int D, int VP>
static int spmv_launch(lpm_ctx *c, bool fused, int remap) {
    typedef Vec<T, VP> V;
    const int g = blocks_for(c->n, 4);
    if (fused)
        hipLaunchKernelGGL((k_spmv<T, D, VP, true>), dim3(g), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, (const T *)c->aval,
                           (const V *)c->v_p, (V *)c->v_t, c->partial, c->partial_cap, c->counter + 1, c->scal, remap);
    else
        hipLaunchKernelGGL((k_spmv<T, D, VP, false>), dim3(g), dim3(256), 0, c->stream, c->n, c->cptr, c->cidx, (const T *)c->aval,
                           (const V *)c->v_p, (V *)c->v_t, c->partial, c->partial_cap, c->counter + 1, c->scal, remap);
    return check_launch(c, "k_spmv");
}

template <typename T, int D, int VP>
static int cg_t(lpm_ctx *c, int *iters, double *error, int *success) {
    typedef Vec<T, VP> V;
    const int64_t n = c->n;
    const int gp = blocks_for(n, 256);
    if (gp > c->partial_cap || blocks_for(n, 4) > c->partial_cap) return lpm_fail(c, LPM_E_LIMIT, "CG grid exceeds partial buffer");
    const int64_t maxit = c->cg_maxit < 0 ? 2 * n * c->dim : c->cg_maxit;
    const int remap = 1;
    hipLaunchKernelGGL((k_cg_init<T, D, VP>), dim3(gp), dim3(256), 0, c->stream, n, c->residual, (const T *)c->invdiag, c->precond,
                       (V *)c->v_x, (V *)c->v_r, (V *)c->v_p, c->cg_tol, (double)maxit, c->partial, c->partial_cap,
                       c->counter, c->scal);
    int rc = check_launch(c, "k_cg_init");
    if (rc) return rc;
    int64_t launched = 0;
    const int batch = c->cg_batch > 0 ? c->cg_batch : 16;
    std::vector<hipEvent_t> evs;
    if (c->timing) {
        evs.resize(2 * batch);
        for (auto &e : evs) LPM_HIP(c, hipEventCreate(&e));
    }
    for (;;) {
        LPM_HIP(c, hipMemcpyAsync(c->h_scal, c->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
        LPM_HIP(c, hipStreamSynchronize(c->stream));
        if (c->timing && launched > 0) {
            for (int k = 0; k < batch; ++k) {
                float ms = 0;
                if (hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]) == hipSuccess) c->st_spmv_ms += ms;
            }
        }
        if (c->h_scal[SC_DONE] != 0.0) break;
        if (launched > maxit + 2 * batch) break;  // the device flag caps the count; this only guards the loop
        for (int k = 0; k < batch; ++k) {
            if (c->timing) LPM_HIP(c, hipEventRecord(evs[2 * k], c->stream));
            rc = spmv_launch<T, D, VP>(c, true, remap);
            if (rc) return rc;
            if (c->timing) LPM_HIP(c, hipEventRecord(evs[2 * k + 1], c->stream));
            hipLaunchKernelGGL((k_cg_update<T, D, VP>), dim3(gp), dim3(256), 0, c->stream, n, (const T *)c->invdiag, c->precond,
                               (V *)c->v_x, (V *)c->v_r, (const V *)c->v_p, (const V *)c->v_t, c->partial, c->partial_cap,
                               c->counter + 2, c->scal);
            rc = check_launch(c, "k_cg_update");
            if (rc) return rc;
            hipLaunchKernelGGL((k_cg_direction<T, D, VP>), dim3(gp), dim3(256), 0, c->stream, n, (const T *)c->invdiag,
                               c->precond, (const V *)c->v_r, (V *)c->v_p, c->scal);
            rc = check_launch(c, "k_cg_direction");
            if (rc) return rc;
        }
        launched += batch;
        c->st_spmv += batch;
    }
    for (auto &e : evs) hipEventDestroy(e);
    const double rhs2 = c->h_scal[SC_RHS2];
    const double done = c->h_scal[SC_DONE];
    *iters = (int)c->h_scal[SC_ITER];
    *error = (done == 3.0) ? 0.0 : sqrt(c->h_scal[SC_RR] / rhs2);
    *success = (done == 3.0) || (*error <= c->cg_tol);
    return LPM_OK;
}

int cg_solve(lpm_ctx *c, int *iters, double *error, int *success) {
    if (!c->k_ready) return lpm_fail(c, LPM_E_STATE, "solve before the stiffness matrix was assembled");
    int rc = launch_precond_setup(c);
    if (rc) return rc;
    if (c->precision == 0) return c->dim == 3 ? cg_t<float, 3, 4>(c, iters, error, success) : cg_t<float, 2, 2>(c, iters, error, success);
    return c->dim == 3 ? cg_t<double, 3, 4>(c, iters, error, success) : cg_t<double, 2, 2>(c, iters, error, success);
}

template <typename T, int D, int VP>
static int move_t(lpm_ctx *c) {
    hipLaunchKernelGGL((k_move<T, D, VP>), dim3(blocks_for(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       (const Vec<T, VP> *)c->v_x, c->disp, c->xyz);
    return check_launch(c, "k_move");
}

int launch_move(lpm_ctx *c) {
    if (c->precision == 0) return c->dim == 3 ? move_t<float, 3, 4>(c) : move_t<float, 2, 2>(c);
    return c->dim == 3 ? move_t<double, 3, 4>(c) : move_t<double, 2, 2>(c);
}

template <typename T, int D, int VP>
static int bench_t(lpm_ctx *c, int reps, double *ms) {
    hipEvent_t a, b;
    LPM_HIP(c, hipEventCreate(&a));
    LPM_HIP(c, hipEventCreate(&b));
    int rc = spmv_launch<T, D, VP>(c, false, 1);  // warm
    if (rc) return rc;
    LPM_HIP(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < reps; ++k) {
        rc = spmv_launch<T, D, VP>(c, false, 1);
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
    if (c->precision == 0) return c->dim == 3 ? bench_t<float, 3, 4>(c, reps, ms) : bench_t<float, 2, 2>(c, reps, ms);
    return c->dim == 3 ? bench_t<double, 3, 4>(c, reps, ms) : bench_t<double, 2, 2>(c, reps, ms);
}
