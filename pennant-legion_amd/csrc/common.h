// Shared device helpers for the gfx950 PENNANT kernels.
//
// Every kernel in this library is compiled with -ffp-contract=off, so each
// multiply and add below rounds separately, in the order written; together
// with IEEE division and sqrt this reproduces the reference's serial CPU
// arithmetic bit for bit (src/Vec2.hh:24-203 semantics).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pennant_hip.h"

#define PN_BLOCK 256
#define PN_RED_BLOCKS 1024   // maximum blocks of a reduction launch

namespace pn {

struct d2 {
    double x, y;
};

__device__ __forceinline__ d2 mk(double x, double y) { d2 v; v.x = x; v.y = y; return v; }
__device__ __forceinline__ d2 add(d2 a, d2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ d2 sub(d2 a, d2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ d2 neg(d2 a) { return mk(-a.x, -a.y); }
__device__ __forceinline__ d2 scl(double r, d2 v) { return mk(v.x * r, v.y * r); }
// vector / scalar multiplies by the reciprocal (src/Vec2.hh:145-149)
__device__ __forceinline__ d2 dvs(d2 v, double r) { const double ri = 1. / r; return mk(v.x * ri, v.y * ri); }
__device__ __forceinline__ double dot(d2 a, d2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double cross(d2 a, d2 b) { return a.x * b.y - a.y * b.x; }
__device__ __forceinline__ double len2(d2 v) { return v.x * v.x + v.y * v.y; }
__device__ __forceinline__ double len(d2 v) { return __builtin_sqrt(v.x * v.x + v.y * v.y); }
__device__ __forceinline__ d2 rot_ccw(d2 v) { return mk(-v.y, v.x); }
__device__ __forceinline__ d2 project(d2 v, d2 u) { return sub(v, scl(dot(v, u), u)); }

// std::min / std::max as the reference's host code evaluates them
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// fields of the driver (two arrays per vector field, stride 1)
__device__ __forceinline__ d2 ld(pennant_cvec v, int64_t i) { return mk(v.x[i], v.y[i]); }
__device__ __forceinline__ d2 ld(pennant_vec v, int64_t i) { return mk(v.x[i], v.y[i]); }
__device__ __forceinline__ void st(pennant_vec v, int64_t i, d2 a) { v.x[i] = a.x; v.y[i] = a.y; }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide reductions for PN_BLOCK threads; the result is valid in thread 0.
// Each starts with a barrier, so that calling one twice in a row cannot
// overwrite a partial that thread 0 is still reading.
__device__ __forceinline__ double block_min(double v) {
    __shared__ double part[PN_BLOCK / 64];
    __syncthreads();
    v = wave_min(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < PN_BLOCK / 64; ++i) v = (part[i] < v) ? part[i] : v;
    }
    return v;
}
__device__ __forceinline__ double block_max(double v) {
    __shared__ double part[PN_BLOCK / 64];
    __syncthreads();
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < PN_BLOCK / 64; ++i) v = (part[i] > v) ? part[i] : v;
    }
    return v;
}
__device__ __forceinline__ int block_sum(int v) {
    __shared__ int part[PN_BLOCK / 64];
    __syncthreads();
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < PN_BLOCK / 64; ++i) v += part[i];
    }
    return v;
}

// The block's index in a launch whose work is laid out for the 8 XCDs:
// MI355X dispatches blocks to its XCDs round-robin by block id and each XCD
// has its own L2, so XCD x (block ids b = x mod 8) takes the x-th contiguous
// share of the logical blocks.  Neighbouring tiles -- which share points and
// the side values of zones across a row -- then meet in the same L2.  A
// bijection of [0, gridDim.x) for any grid size.
__device__ __forceinline__ int xcd_block() {
    const int n = gridDim.x, b = blockIdx.x;
    const int q = n >> 3, r = n & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

inline int nblocks(int64_t n) { return (int)((n + PN_BLOCK - 1) / PN_BLOCK); }
inline int nred_blocks(int64_t n) {
    const int b = nblocks(n);
    return b < 1 ? 1 : (b > PN_RED_BLOCKS ? PN_RED_BLOCKS : b);
}

}  // namespace pn
