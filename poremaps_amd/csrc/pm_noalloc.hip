// pm_noalloc.hip -- pm_permeability_noalloc: the permeability reductions of
// pm_permeability (k_reduce_state + k_reduce_final, src/evaluation.cc:54-211)
// over fields the caller keeps in device memory, for callers that cannot
// allocate device memory.  The kernels' partial sums and the final sums live
// in one file-scope device array (`scratch`, 1.38 MB per GPU, held until the
// process ends) instead of a buffer the library allocates; the call itself
// allocates nothing.
#include "../../include/poremaps_amd.h"
#include "pm_ctx.h"
#include "pm_reduce.h"

#include <cstdint>
#include <mutex>

namespace {

// One block partial per reduction workgroup, then the final sums.
struct NoAllocScratch {
    pm::RedAcc part[pm::kRedBlocksMax];
    pm::RedSums sums;
};

static __device__ __align__(256) NoAllocScratch scratch;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
NoAllocScratch *g_ptr[kMaxDevices] = {};  // scratch's address on each device

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
int scratch_on(int dev, NoAllocScratch **out) {
    if (dev < 0 || dev >= kMaxDevices) return pm_fail(PM_E_ARG, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return pm_fail(PM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_ptr[dev] = (NoAllocScratch *)p;
    }
    *out = g_ptr[dev];
    return PM_OK;
}

}  // namespace

extern "C" int pm_permeability_noalloc(const uint8_t *geom, const double *p, const double *u, const double *v,
                                       const double *w, int nx, int ny, int nz, const int dom_interest[6],
                                       double voxel_size, double porosity, double out[8], void *stream) {
    if (!geom || !p || !u || !v || !w || !out) return pm_fail(PM_E_ARG, "null argument");
    if (nx < 1 || ny < 1 || nz < 1) return pm_fail(PM_E_ARG, "size %d x %d x %d", nx, ny, nz);
    const int64_t n = (int64_t)nx * ny * nz;
    if (n > PM_NOALLOC_MAX_VOXELS)
        return pm_fail(PM_E_ARG, "%lld voxels exceed the %lld of pm_permeability_noalloc", (long long)n,
                       (long long)PM_NOALLOC_MAX_VOXELS);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return pm_fail(PM_E_HIP, "hipGetDevice failed");
    NoAllocScratch *sp = nullptr;
    int rc = scratch_on(dev, &sp);
    if (rc != PM_OK) return rc;
    // Interior arrays (x fastest) seen through the halo-2 indexing of the
    // kernels: cell (i, j, k) of the halo-including grid is element
    // (i-2)*nx*ny + (j-2)*nx + (k-2).
    pm::Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.NX = nx + 4; g.NY = ny + 4; g.NZ = nz + 4;
    g.sx = nx;
    g.sy = (int64_t)nx * ny;
    g.base = -(2 * g.sy + 2 * g.sx + 2);
    g.total = n;
    g.px = g.py = g.pz = 0;
    g.cb = 0;
    pm::DomBox box;
    const int *d = dom_interest;
    box.on = d && !(d[0] == 0 && d[1] == 0 && d[2] == 0 && d[3] == 0 && d[4] == 0 && d[5] == 0);
    for (int q = 0; q < 6; q++) box.d[q] = d ? d[q] : 0;
    hipStream_t s = (hipStream_t)stream;
    pm::launch_reduce(g, geom, p, u, v, w, box, true, sp->part, &sp->sums, s);
    hipError_t e = hipGetLastError();
    pm::RedSums r;
    if (e == hipSuccess) e = hipMemcpyAsync(&r, &sp->sums, sizeof r, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return pm_fail(PM_E_HIP, "pm_permeability_noalloc: %s", hipGetErrorString(e));
    // src/calc_flow.cc:92-97: porosity -1.0 is computed from the geometry
    const double phi = porosity == -1.0 ? (double)r.count[0] / (double)(nx * ny * nz) : porosity;
    pm_permeability_from(r, box.on ? d : box.d, voxel_size, phi, out);
    return PM_OK;
}
