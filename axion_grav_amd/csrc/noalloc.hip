// axg_energy_fields: the lattice energy of axg_energy on fields the caller holds in device
// memory, for callers that cannot allocate (code under HIP graph capture, tight serving loops).
// The per-block partial sums and the three column sums live in `scratch`, a fixed device array
// of this file; the entry point itself allocates nothing and does not synchronize.
#include <array>
#include <mutex>

#include "axg_internal.h"

namespace axg {

constexpr int NOALLOC_NWG = 1024;                // energy blocks: the grid axg_energy launches
constexpr int NOALLOC_SCRATCH_DOUBLES = 8192;    // 64 KiB per device
static_assert(4 * NOALLOC_NWG + 3 <= NOALLOC_SCRATCH_DOUBLES, "energy partials exceed the scratch array");
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace axg

// The scratch array: [0, 3 nwg) per-block partials (kinetic, gradient, potential), then the
// three sums at 4 nwg.  Each device holds its own copy; host code reaches it only through the
// address hipGetSymbolAddress gives for the current device.
static __device__ __align__(256) double scratch[axg::NOALLOC_SCRATCH_DOUBLES];

namespace axg {

static std::array<double*, NOALLOC_MAX_DEVICES> g_scratch_ptr{};
static std::array<hipError_t, NOALLOC_MAX_DEVICES> g_scratch_err{};
static std::array<std::once_flag, NOALLOC_MAX_DEVICES> g_scratch_once;

// The device address of `scratch` on `dev` (current): looked up once per device.
static double* scratch_on(int dev)
{
    std::call_once(g_scratch_once[dev], [dev] {
        void* ptr = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (g_scratch_err[dev] == hipSuccess) g_scratch_ptr[dev] = (double*)ptr;
    });
    if (g_scratch_err[dev] != hipSuccess)
        throw Error(AXG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[dev]));
    return g_scratch_ptr[dev];
}

// out = the column sums scaled as axg_energy scales them on the host, in the same order
// (kinetic dv, gradient / dspa^2 dv, potential tau^2 sigma_cr^2 dv, total).
__global__ void energy_finish_kernel(const double* __restrict__ sums, double dv, double dd, double ts,
                                     double* __restrict__ out)
{
#pragma clang fp contract(off)
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double a = sums[0] * dv;
        const double b = sums[1] / dd * dv;
        const double c = ts * sums[2] * dv;
        out[0] = a;
        out[1] = b;
        out[2] = c;
        out[3] = a + b + c;
    }
}

}  // namespace axg

using namespace axg;

extern "C" int axg_energy_fields(const axg_params* p, const double* phi, const double* pi, double tau, double* out,
                                 void* stream)
{
    try {
        if (!p || !phi || !pi || !out) throw Error(AXG_EINVAL, "axg_energy_fields: null argument");
        if (!(tau > 0)) throw Error(AXG_EINVAL, "axg_energy_fields: tau must be > 0");
        if (p->ngrid < 2 || p->ngrid > 1024)
            throw Error(AXG_EINVAL, "axg_energy_fields: ngrid must be 2..1024 (got " + std::to_string(p->ngrid) + ")");
        if (!(p->dspa > 0) || !(p->sigma_cr > 0)) throw Error(AXG_EINVAL, "axg_energy_fields: dspa and sigma_cr needed");
        if (p->device < 0 || p->device >= NOALLOC_MAX_DEVICES)
            throw Error(AXG_EINVAL, "axg_energy_fields: device ordinal out of range");
        AXG_HIP(hipSetDevice(p->device));
        double* sc = scratch_on(p->device);
        hipStream_t s = (hipStream_t)stream;
        const double thrm = p->initial_loop ? 0.0 : AXG_THRM_SIGMA / tau;
        const double m2 = msquared(thrm), ma2 = maxion2(thrm, p->kappa), kt = tilt_coef(*p, tau);
        double* sums = sc + (size_t)4 * NOALLOC_NWG;
        launch_energy<double, false>(s, (int)p->ngrid, phi, nullptr, pi, nullptr, m2, ma2, kt, sc, NOALLOC_NWG, sums);
        const double dv = p->dspa * p->dspa * p->dspa;
        const double ts = tau * tau * p->sigma_cr * p->sigma_cr;
        hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(64), 0, s, (const double*)sums, dv, p->dspa * p->dspa,
                           ts, out);
        AXG_HIP(hipGetLastError());
        return AXG_OK;
    } catch (const Error& e) {
        set_handleless_error(e.msg);
        return e.code;
    } catch (...) {
        set_handleless_error("axg_energy_fields: unexpected exception");
        return AXG_EDEVICE;
    }
}
