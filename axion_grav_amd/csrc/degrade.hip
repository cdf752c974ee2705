// The reduced grid of degrade_spec > 0: the GW source and the axion spectrum are computed on
// Nc = N / deg sites per axis, deg = 2^degrade_spec, each coarse site standing for a block of
// deg^3 lattice sites.  One thread per coarse site walks its block in the reference's loop
// order (z outermost, x innermost) and sums in FP64 term by term as the reference does, so the
// coarse values are the reference's to the last bit for FP64 fields.
#include "axg_internal.h"

namespace axg {

constexpr int DEG_NT = 256;

// grav_waves.f90:384-407 data1_to_arr_dphi_degrade: per axis d the block sum of the central
// differences phi(s + e_d) - phi(s - e_d), accumulated as (sum + a) - b from zero and divided
// by d = deg^3 * dspa * 2; then the six products S_ij = 2 Re(g_i conj(g_j)) of
// grav_wave_integration (grav_waves.f90:213-221) in the packed planes the x pass reads:
// C0 = S11 + i S21, C1 = S22 + i S31, C2 = S32 + i S33.
template <typename T, bool COMP>
__global__ void __launch_bounds__(DEG_NT) degrade_source_kernel(int N, int deg, const typename V2<T>::type* __restrict__ phi,
                                                               const short2* __restrict__ cphi, double d,
                                                               double2* __restrict__ src)
{
#pragma clang fp contract(off)
    const int Nc = N / deg;
    const int64_t nc3 = (int64_t)Nc * Nc * Nc;
    const int64_t c = blockIdx.x * (int64_t)DEG_NT + threadIdx.x;
    if (c >= nc3) return;
    const int cx = (int)(c % Nc), cy = (int)((c / Nc) % Nc), cz = (int)(c / ((int64_t)Nc * Nc));
    const int64_t NN = (int64_t)N * N;
    double2 g[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int z = cz * deg; z < cz * deg + deg; z++) {
        const int64_t zm = (z == 0 ? N - 1 : z - 1) * NN, zp = (z == N - 1 ? 0 : z + 1) * NN, z0 = z * NN;
        for (int y = cy * deg; y < cy * deg + deg; y++) {
            const int64_t ym = (int64_t)(y == 0 ? N - 1 : y - 1) * N, yp = (int64_t)(y == N - 1 ? 0 : y + 1) * N,
                          y0 = (int64_t)y * N;
            for (int x = cx * deg; x < cx * deg + deg; x++) {
                const int xm = x == 0 ? N - 1 : x - 1, xp = x == N - 1 ? 0 : x + 1;
                const double2 ax = ldf<T, COMP>(phi, cphi, z0 + y0 + xp), bx = ldf<T, COMP>(phi, cphi, z0 + y0 + xm);
                const double2 ay = ldf<T, COMP>(phi, cphi, z0 + yp + x), by = ldf<T, COMP>(phi, cphi, z0 + ym + x);
                const double2 az = ldf<T, COMP>(phi, cphi, zp + y0 + x), bz = ldf<T, COMP>(phi, cphi, zm + y0 + x);
                g[0].x = g[0].x + ax.x - bx.x;
                g[0].y = g[0].y + ax.y - bx.y;
                g[1].x = g[1].x + ay.x - by.x;
                g[1].y = g[1].y + ay.y - by.y;
                g[2].x = g[2].x + az.x - bz.x;
                g[2].y = g[2].y + az.y - bz.y;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        g[k].x = g[k].x / d;
        g[k].y = g[k].y / d;
    }
    src[c] = double2{2 * (g[0].x * g[0].x + g[0].y * g[0].y), 2 * (g[1].x * g[0].x + g[1].y * g[0].y)};
    src[nc3 + c] = double2{2 * (g[1].x * g[1].x + g[1].y * g[1].y), 2 * (g[2].x * g[0].x + g[2].y * g[0].y)};
    src[2 * nc3 + c] = double2{2 * (g[2].x * g[1].x + g[2].y * g[1].y), 2 * (g[2].x * g[2].x + g[2].y * g[2].y)};
}

// spectrum.f90:761-786 data2_to_data_spec_degrade: the block mean of a field (pi), summed in
// loop order from zero and divided by deg^3.
template <typename T, bool COMP>
__global__ void __launch_bounds__(DEG_NT) degrade_field_kernel(int N, int deg, const typename V2<T>::type* __restrict__ f,
                                                              const short2* __restrict__ cf, double2* __restrict__ out)
{
#pragma clang fp contract(off)
    const int Nc = N / deg;
    const int64_t nc3 = (int64_t)Nc * Nc * Nc;
    const int64_t c = blockIdx.x * (int64_t)DEG_NT + threadIdx.x;
    if (c >= nc3) return;
    const int cx = (int)(c % Nc), cy = (int)((c / Nc) % Nc), cz = (int)(c / ((int64_t)Nc * Nc));
    const int64_t NN = (int64_t)N * N;
    double2 s = {0.0, 0.0};
    for (int z = cz * deg; z < cz * deg + deg; z++)
        for (int y = cy * deg; y < cy * deg + deg; y++)
            for (int x = cx * deg; x < cx * deg + deg; x++) {
                const double2 v = ldf<T, COMP>(f, cf, z * NN + (int64_t)y * N + x);
                s.x = s.x + v.x;
                s.y = s.y + v.y;
            }
    const double dd = (double)(deg * deg * deg);
    out[c] = double2{s.x / dd, s.y / dd};
}

// masking (spectrum.f90:900-954) with deg > 1: a coarse site is masked (flag bit 4) when any
// lattice site of its block is.
__global__ void __launch_bounds__(DEG_NT) degrade_mask_kernel(int N, int deg, const int8_t* __restrict__ flags,
                                                             int8_t* __restrict__ out)
{
    const int Nc = N / deg;
    const int64_t nc3 = (int64_t)Nc * Nc * Nc;
    const int64_t c = blockIdx.x * (int64_t)DEG_NT + threadIdx.x;
    if (c >= nc3) return;
    const int cx = (int)(c % Nc), cy = (int)((c / Nc) % Nc), cz = (int)(c / ((int64_t)Nc * Nc));
    const int64_t NN = (int64_t)N * N;
    int m = 0;
    for (int z = cz * deg; z < cz * deg + deg; z++)
        for (int y = cy * deg; y < cy * deg + deg; y++)
            for (int x = cx * deg; x < cx * deg + deg; x++) m |= flags[z * NN + (int64_t)y * N + x] & 4;
    out[c] = (int8_t)m;
}

static unsigned deg_blocks(int N, int deg)
{
    const int64_t Nc = N / deg;
    return (unsigned)((Nc * Nc * Nc + DEG_NT - 1) / DEG_NT);
}

template <typename T, bool COMP>
void launch_degrade_source(hipStream_t s, int N, int deg, const void* phi, const void* cphi, double dspa, double2* src)
{
    using T2 = typename V2<T>::type;
    const double d = (double)(deg * deg * deg) * dspa * 2;  // grav_waves.f90:384
    hipLaunchKernelGGL((degrade_source_kernel<T, COMP>), dim3(deg_blocks(N, deg)), dim3(DEG_NT), 0, s, N, deg,
                       (const T2*)phi, (const short2*)cphi, d, src);
    AXG_HIP(hipGetLastError());
}

template <typename T, bool COMP>
void launch_degrade_field(hipStream_t s, int N, int deg, const void* f, const void* cf, double2* out)
{
    using T2 = typename V2<T>::type;
    hipLaunchKernelGGL((degrade_field_kernel<T, COMP>), dim3(deg_blocks(N, deg)), dim3(DEG_NT), 0, s, N, deg,
                       (const T2*)f, (const short2*)cf, out);
    AXG_HIP(hipGetLastError());
}

void launch_degrade_mask(hipStream_t s, int N, int deg, const int8_t* flags, int8_t* out)
{
    hipLaunchKernelGGL(degrade_mask_kernel, dim3(deg_blocks(N, deg)), dim3(DEG_NT), 0, s, N, deg, flags, out);
    AXG_HIP(hipGetLastError());
}

#define AXG_DEGRADE_INST(T, C)                                                                                       \
    template void launch_degrade_source<T, C>(hipStream_t, int, int, const void*, const void*, double, double2*);      \
    template void launch_degrade_field<T, C>(hipStream_t, int, int, const void*, const void*, double2*);
AXG_DEGRADE_INST(double, false)
AXG_DEGRADE_INST(float, false)
AXG_DEGRADE_INST(float, true)
#undef AXG_DEGRADE_INST

}  // namespace axg
