// pm_reduce.h -- deterministic compensated FP64 block reductions.
#pragma once

#include "pm_kernels.h"

namespace pm {

struct NSum {
    double s, c;
};

__device__ __forceinline__ void nadd(NSum &a, double x) {
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ void nmerge(NSum &a, const NSum &b) {
    nadd(a, b.s);
    a.c += b.c;
}

struct RedAcc {
    NSum s[8];         // u, v, w (fluid); u, v, w (box fluid); p(z=zmin), p(z=zmax)
    long long n[4];    // fluid, box fluid, zmin fluid, zmax fluid
    double wmax;
};

__device__ __forceinline__ void acc_zero(RedAcc &a) {
    for (int q = 0; q < 8; q++) { a.s[q].s = 0.0; a.s[q].c = 0.0; }
    for (int q = 0; q < 4; q++) a.n[q] = 0;
    a.wmax = 0.0;
}
__device__ __forceinline__ void acc_merge(RedAcc &a, const RedAcc &b) {
    for (int q = 0; q < 8; q++) nmerge(a.s[q], b.s[q]);
    for (int q = 0; q < 4; q++) a.n[q] += b.n[q];
    a.wmax = fmax(a.wmax, b.wmax);
}


// Block of kRedThreads threads: a butterfly of lane shuffles inside each
// 64-lane wave (no LDS, no barrier), then the waves' partials through LDS,
// merged in wave order by thread 0.  The merge order is fixed, so the result
// is the same on every run; only lane 0's butterfly value is used.
__device__ __forceinline__ RedAcc acc_shfl_xor(const RedAcc &a, int m) {
    RedAcc b;
    for (int q = 0; q < 8; q++) {
        b.s[q].s = __shfl_xor(a.s[q].s, m, 64);
        b.s[q].c = __shfl_xor(a.s[q].c, m, 64);
    }
    for (int q = 0; q < 4; q++) b.n[q] = __shfl_xor(a.n[q], m, 64);
    b.wmax = __shfl_xor(a.wmax, m, 64);
    return b;
}

__device__ void block_reduce_store(RedAcc &a, RedAcc *out) {
    static_assert(kRedThreads % 64 == 0, "whole waves");
    constexpr int kWaves = kRedThreads / 64;
    __shared__ RedAcc sh[kWaves];
    for (int m = 32; m > 0; m >>= 1) {
        const RedAcc b = acc_shfl_xor(a, m);
        acc_merge(a, b);
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        RedAcc r = sh[0];
        for (int w = 1; w < kWaves; w++) acc_merge(r, sh[w]);
        *out = r;
    }
}


}  // namespace pm
