// noalloc.hip -- the allocation-free variant of the Toeplitz K_ff DFT
// builder, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops).  Its one scratch buffer, the real Toeplitz
// kernel st between the Okada launch and the DFT launch, is a static device
// array of fixed size instead of a hipMalloc.
#include <hip/hip_runtime.h>
#include <mutex>

#include "oq_internal.h"

namespace oq {

// nx * nxi * nxi doubles of the real kernel: 2^20 (8 MiB) holds the 256 x 64
// fault of the N = 65536 workload.  Each GPU the library runs on keeps its
// own copy until the process ends.
static constexpr int64_t kNoAllocMaxSt = (int64_t)1 << 20;
static __device__ __align__(256) double scratch[kNoAllocMaxSt];

// The array's device address per device (the host-side name `scratch` is a
// stand-in, not a device address): looked up once per device, with that
// device current, on the entry point's first call there.
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int scratch_of_current_device(double **out)
{
    int dev = -1;
    OQ_HIP(hipGetDevice(&dev));
    OQ_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        OQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        OQ_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return OQ_OK;
}

}  // namespace oq

using namespace oq;

extern "C" {

int oq_gf_okada_ff_dft_noalloc(const oq_fault_mesh *mesh, double lambda, double mu, int ftype,
                               int64_t nrept, double buffer_ratio, double *st_dft, void *stream)
{
    int rc;
    OQ_REQUIRE(st_dft != nullptr, "output is NULL");
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    OQ_REQUIRE(nrept >= 0, "nrept must be >= 0");
    OQ_REQUIRE(buffer_ratio >= 0.0, "Argument `buffer_ratio` must be ≥ 0.");
    if ((rc = check_fault(mesh))) return rc;
    OQ_REQUIRE(mesh->nx <= kMaxDftNx, "the device DFT of the Toeplitz kernel takes nx <= 5120");
    OQ_REQUIRE(mesh->nx <= kNoAllocMaxSt / mesh->nxi / mesh->nxi,
               "nx*nxi*nxi exceeds the allocation-free scratch of 1048576 doubles "
               "(use oq_gf_okada_ff_dft)");
    double *st = nullptr;
    if ((rc = scratch_of_current_device(&st))) return rc;
    DevFault f{};
    f.x = mesh->x;
    f.ax = mesh->ax;
    f.y = mesh->y;
    f.z = mesh->z;
    f.axi = mesh->axi;
    f.nx = mesh->nx;
    f.nxi = mesh->nxi;
    f.dep = mesh->dep;
    f.sd = sind(mesh->dip);
    f.cd = cosd(mesh->dip);
    f.s2d = sind(2.0 * mesh->dip);
    f.c2d = cosd(2.0 * mesh->dip);
    f.lrept = (buffer_ratio + 1.0) * (mesh->dx * (double)mesh->nx);  // GF.jl:38
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((rc = launch_gf_ff(f, lambda, mu, ftype, (int)nrept, st, s))) return rc;
    return launch_toeplitz_dft(st, mesh->nx, mesh->nxi, st_dft, s);
}

}  // extern "C"
