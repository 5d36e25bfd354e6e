// pm_stream_probe: the HBM bandwidth a streaming kernel reaches on this GPU at
// a given read : write mix, the ceiling against which the sweeps' measured
// traffic rates are read (the published 8 TB/s is a peak no kernel reaches).
//
//     poremaps_amd/bin/pm_stream_probe [MiB of written data, default 1024]
//
// For R = 1 (copy), 2, 3, 4, 6, 8 float4 read per float4 written, a wave
// reads R consecutive 1 KiB blocks of one input stream and writes one 1 KiB
// block of the output stream; a pure read of one stream and R independent
// read streams (offsets i, i+n, ..., i+(R-1)n: the HBM rate drops as the
// streams multiply) are timed beside them.  Footprints are GB-sized, far
// beyond the 256 MB Infinity Cache; a pure read and a copy of 64 and 256 MiB
// show what the caches add.  Times are HIP events over 20 launches after 3
// warm-ups.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

__global__ void k_read(const float4 *__restrict__ a, long n, float *sink) {
    float s = 0.f;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    // never true for the zero-filled input; keeps the loads live
    if (s == 1.0e30f) sink[0] = s;
}

// one read stream: output element i (lane l of its wave) sums
// a[(i - l) R + r 64 + l], r < R
template <int R>
__global__ void k_mix1(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
    const long lane = threadIdx.x & 63;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float4 *src = a + (i - lane) * R + lane;
        float4 s = src[0];
#pragma unroll
        for (int r = 1; r < R; r++) {
            const float4 v = src[r * 64];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
        b[i] = s;
    }
}

// R read streams
template <int R>
__global__ void k_mix(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float4 s = a[i];
#pragma unroll
        for (int r = 1; r < R; r++) {
            const float4 v = a[i + r * n];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
        b[i] = s;
    }
}

struct Timer {
    hipEvent_t e0, e1;
    Timer() { CHECK(hipEventCreate(&e0)); CHECK(hipEventCreate(&e1)); }
};

template <class F>
float time_ms(Timer &t, F launch) {
    for (int w = 0; w < 3; w++) launch();
    CHECK(hipEventRecord(t.e0));
    const int reps = 20;
    for (int r = 0; r < reps; r++) launch();
    CHECK(hipEventRecord(t.e1));
    CHECK(hipEventSynchronize(t.e1));
    CHECK(hipGetLastError());
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, t.e0, t.e1));
    return ms / reps;
}

template <int R>
void run_mix(Timer &t, const float4 *a, float4 *b, long n, int grid, int block) {
    const double rd = (double)n * 16.0 * R, wr = (double)n * 16.0;
    const float ms1 = time_ms(t, [&] { k_mix1<R><<<grid, block>>>(a, b, n); });
    const float msk = time_ms(t, [&] { k_mix<R><<<grid, block>>>(a, b, n); });
    std::printf("%-11d %9.1f %9.1f %9.1f %8.3f %6.2f %9.1f %8.3f\n", R, rd / 1048576.0, wr / 1048576.0,
                ms1 * 1e3, (rd + wr) / (ms1 * 1e-3) / 1e12, (rd + wr) / (ms1 * 1e-3) / 8e12, msk * 1e3,
                (rd + wr) / (msk * 1e-3) / 1e12);
}

}  // namespace

int main(int argc, char **argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    if (mib < 256 || mib > 4096) {
        std::fprintf(stderr, "MiB of written data must lie in [256, 4096]\n");
        return 2;
    }
    const int kMaxR = 8;
    const long n = (mib << 20) / 16;  // float4 elements per stream
    float4 *a = nullptr, *b = nullptr;
    float *sink = nullptr;
    CHECK(hipMalloc(&a, (size_t)n * 16 * kMaxR));
    CHECK(hipMalloc(&b, (size_t)n * 16));
    CHECK(hipMalloc(&sink, 64));
    CHECK(hipMemset(a, 0, (size_t)n * 16 * kMaxR));
    CHECK(hipMemset(b, 0, (size_t)n * 16));
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int block = 256, grid = prop.multiProcessorCount * 8;
    Timer t;
    std::printf("device %s, %d CUs, grid %d x %d, %ld MiB per stream\n", prop.gcnArchName,
                prop.multiProcessorCount, grid, block, mib);
    std::printf("%-11s %9s %9s %9s %8s %6s %9s %8s\n", "reads/write", "read MiB", "write MiB", "us", "TB/s",
                "of 8", "R streams", "TB/s");
    {
        const long nr = n * 4;  // pure read over a footprint of four streams
        const float ms = time_ms(t, [&] { k_read<<<grid, block>>>(a, nr, sink); });
        const double rd = (double)nr * 16.0;
        std::printf("%-11s %9.1f %9.1f %9.1f %8.3f %6.2f\n", "read-only", rd / 1048576.0, 0.0, ms * 1e3,
                    rd / (ms * 1e-3) / 1e12, rd / (ms * 1e-3) / 8e12);
    }
    run_mix<1>(t, a, b, n, grid, block);
    run_mix<2>(t, a, b, n, grid, block);
    run_mix<3>(t, a, b, n, grid, block);
    run_mix<4>(t, a, b, n, grid, block);
    run_mix<6>(t, a, b, n, grid, block);
    run_mix<8>(t, a, b, n, grid, block);
    std::printf("cache-resident footprints:\n");
    for (long small : {64L, 256L}) {
        const long ns = (small << 20) / 16;
        const float msr = time_ms(t, [&] { k_read<<<grid, block>>>(a, ns, sink); });
        const long nc = ns / 2;
        const float msc = time_ms(t, [&] { k_mix1<1><<<grid, block>>>(a, b, nc); });
        std::printf("  %4ld MiB footprint: read-only %.3f TB/s, copy %.3f TB/s\n", small,
                    (double)ns * 16.0 / (msr * 1e-3) / 1e12, (double)nc * 32.0 / (msc * 1e-3) / 1e12);
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(sink));
    return 0;
}
