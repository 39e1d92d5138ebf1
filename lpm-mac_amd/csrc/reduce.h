// Deterministic block reductions for the multi-block dot products and norms.
//
// Each block reduces its values and writes one partial per quantity (no atomics:
// a counter shared by tens of thousands of workgroups serialises at the memory side).
// A one-block finalize kernel then sums the partials in a fixed order and updates the
// device scalars, so results do not depend on workgroup scheduling.
#pragma once

#include "lpm_internal.h"

template <int NV>
static __device__ __forceinline__ void block_partial(double (&v)[NV], double *partial, int cap) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[k][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0;
        for (int w = 0; w < nw; ++w) s += red[threadIdx.x][w];
        partial[threadIdx.x * cap + blockIdx.x] = s;
    }
}

// finalize kinds
enum { FIN_NORMS = 3, FIN_DIAG = 4 };

// SELL-64 layout of the CG matrix: block rows are grouped in slices of 64 consecutive rows;
// slice k starts at block slot sp[k] and is W_k slots wide (W_k = longest row of the slice).
// Entry e (= r*dim + s) of block (row i, slot c) lives at
//   aval[EB*sp[i>>6] + ((int64)c*EB + e)*64 + (i & 63)],  EB = dim*dim,
// so a wavefront whose lane l owns row 64k+l reads 64 consecutive values per load.
__device__ __forceinline__ int64_t sell_off(const int64_t *sp, int64_t i, int c, int EB, int e) {
    return (int64_t)EB * sp[i >> 6] + ((int64_t)c * EB + e) * 64 + (i & 63);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD b % 8). This maps
// them so that each XCD takes a contiguous range of logical workgroups, for any grid size:
// XCD x gets floor(g/8) workgroups, plus one when x < g % 8.
__device__ __forceinline__ int64_t xcd_major_block() {
    const int64_t b = blockIdx.x, g = gridDim.x, q = g >> 3, rem = g & 7, x = b & 7;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}

// Rigid-body modes of one particle at offset d from its aggregate centre: P = [I_D | R(d)] with
// u = t + w x d (3D, 6 modes) or u = t + theta e_z x d (2D, 3 modes); row k is zero when bit k
// of `mask` marks DoF k as eliminated.
template <int D>
__device__ __forceinline__ void rigid_modes(const double *d, unsigned mask, double (&P)[3][6]) {
    for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 6; ++m) P[k][m] = 0.0;
    if (D == 3) {
        P[0][0] = 1.0;
        P[0][4] = d[2];
        P[0][5] = -d[1];
        P[1][1] = 1.0;
        P[1][3] = -d[2];
        P[1][5] = d[0];
        P[2][2] = 1.0;
        P[2][3] = d[1];
        P[2][4] = -d[0];
    } else {
        P[0][0] = 1.0;
        P[0][2] = -d[1];
        P[1][1] = 1.0;
        P[1][2] = d[0];
    }
    for (int k = 0; k < D; ++k)
        if (mask & (1u << k))
            for (int m = 0; m < 6; ++m) P[k][m] = 0.0;
}
