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


__device__ void block_reduce_store(RedAcc &a, RedAcc *out) {
    __shared__ RedAcc sh[kRedThreads];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int off = kRedThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) acc_merge(sh[threadIdx.x], sh[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}


}  // namespace pm
