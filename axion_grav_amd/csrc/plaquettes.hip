// Plaquette windings: for every elementary face of the lattice, the winding number of the phase
// of phi around its four corners, and the reference's quadrant test on the same face (the
// string-penetration test of velocity_string_YY, defects_3d.f90:850-879).  The winding is integer
// work with no transcendental function: an edge a -> b of the loop adds +1 when the shorter arc
// from arg a to arg b crosses the negative real axis counterclockwise, -1 when clockwise, decided
// by sign tests on Im a, Im b and the cross product Re a Im b - Im a Re b.  That is the lattice
// winding sum_edges wrap(arg b - arg a) / 2 pi, except where two corners are antipodal (cross
// product 0, the shorter arc undefined), which counts 0.
//
// Face o of the site s (o = 0, 1, 2 as the reference's ipla = 1, 2, 3) spans the axes (u, v) =
// (x, y), (y, z), (z, x); its loop is s -> s+u -> s+u+v -> s+v -> s, counterclockwise about
// w = u x v.  Per site: code = sum_o (n_o + 2) << 3o | quad_o << (9 + o) as a uint16.
#include "axg_internal.h"

namespace axg {

constexpr int PLAQ_NT = PLAQ_BLOCK;

// Crossing of the negative real axis by the shorter arc from arg a to arg b: +1 counterclockwise,
// -1 clockwise, 0 otherwise.  The cut belongs to the upper side (Im z >= 0, as atan2 = pi).
__device__ __forceinline__ int arc_cut(double2 a, double2 b)
{
#pragma clang fp contract(off)
    const bool ua = !(a.y < 0), ub = !(b.y < 0);
    const double c = a.x * b.y - a.y * b.x;
    return (ua && !ub && c > 0) ? 1 : ((!ua && ub && c < 0) ? -1 : 0);
}

// Quadrant index of defects_3d.f90:868: ceiling(2 ph / pi) + 2 with ph = aimag(log(z)).
__device__ __forceinline__ int quadrant(double2 z)
{
#pragma clang fp contract(off)
    const double ph = atan2(z.y, z.x);
    return (int)ceil(2 * ph / AXG_PI_T) + 2;
}

// defects_3d.f90:856-877: the four corners' phases fill at least three of the four quadrants.
__device__ __forceinline__ int quad_hit(double2 a, double2 b, double2 c, double2 d)
{
    const int m = (1 << quadrant(a)) | (1 << quadrant(b)) | (1 << quadrant(c)) | (1 << quadrant(d));
    return __popc(m) >= 3 ? 1 : 0;
}

template <typename T, bool COMP>
__global__ void __launch_bounds__(PLAQ_NT) plaquette_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                            const short2* __restrict__ cphi,
                                                            uint16_t* __restrict__ code, int64_t* __restrict__ counts)
{
    __shared__ int sc[PLAQ_COLS];
    if (threadIdx.x < PLAQ_COLS) sc[threadIdx.x] = 0;
    __syncthreads();
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    const int64_t i = blockIdx.x * (int64_t)PLAQ_NT + threadIdx.x;
    if (i < n3) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
        const int64_t dx = (x + 1 == N ? 0 : x + 1) - x, dy = ((y + 1 == N ? 0 : y + 1) - y) * (int64_t)N,
                      dz = ((z + 1 == N ? 0 : z + 1) - z) * NN;
        const double2 s = ldf<T, COMP>(phi, cphi, i);
        const double2 sx = ldf<T, COMP>(phi, cphi, i + dx), sy = ldf<T, COMP>(phi, cphi, i + dy);
        const double2 sz = ldf<T, COMP>(phi, cphi, i + dz);
        const double2 sxy = ldf<T, COMP>(phi, cphi, i + dx + dy), syz = ldf<T, COMP>(phi, cphi, i + dy + dz);
        const double2 szx = ldf<T, COMP>(phi, cphi, i + dz + dx);
        // (u, v) = (x, y), (y, z), (z, x): corners s, s+u, s+u+v, s+v
        const double2 c1[3] = {sx, sy, sz}, c2[3] = {sxy, syz, szx}, c3[3] = {sy, sz, sx};
        int cd = 0, pierced = 0;
#pragma unroll
        for (int o = 0; o < 3; o++) {
            const int n = arc_cut(s, c1[o]) + arc_cut(c1[o], c2[o]) + arc_cut(c2[o], c3[o]) + arc_cut(c3[o], s);
            const int q = quad_hit(s, c1[o], c2[o], c3[o]);
            cd |= (n + 2) << (3 * o);
            cd |= q << (9 + o);
            if (n != 0) {
                pierced++;
                atomicAdd(&sc[5 * o + 0], 1);
                atomicAdd(&sc[5 * o + 4], n);
            }
            if (n == 1) atomicAdd(&sc[5 * o + 1], 1);
            if (n == -1) atomicAdd(&sc[5 * o + 2], 1);
            if (q) atomicAdd(&sc[5 * o + 3], 1);
        }
        code[i] = (uint16_t)cd;
        if (pierced) atomicAdd(&sc[15], pierced);
    }
    __syncthreads();
    if (threadIdx.x < PLAQ_COLS) counts[blockIdx.x * (int64_t)PLAQ_COLS + threadIdx.x] = sc[threadIdx.x];
}

// Pierced faces in site order, then face order: rows (ix, iy, iz, o, n), 1-based site and o.
__global__ void __launch_bounds__(PLAQ_NT) plaquette_compact_kernel(int N, const uint16_t* __restrict__ code,
                                                                    const int64_t* __restrict__ offs,
                                                                    int32_t* __restrict__ rec)
{
    __shared__ int sh[PLAQ_NT];
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    const int64_t i = blockIdx.x * (int64_t)PLAQ_NT + threadIdx.x;
    int n[3] = {0, 0, 0}, m = 0;
    if (i < n3) {
        const int cd = code[i];
#pragma unroll
        for (int o = 0; o < 3; o++) {
            n[o] = ((cd >> (3 * o)) & 7) - 2;
            m += n[o] != 0;
        }
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int d = 1; d < PLAQ_NT; d <<= 1) {
        const int v = (int)threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    if (m) {
        int64_t pos = offs[blockIdx.x] + sh[threadIdx.x] - m;
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
#pragma unroll
        for (int o = 0; o < 3; o++) {
            if (n[o] != 0) {
                int32_t* r = rec + 5 * pos;
                r[0] = x + 1;
                r[1] = y + 1;
                r[2] = z + 1;
                r[3] = o + 1;
                r[4] = n[o];
                pos++;
            }
        }
    }
}

template <typename T, bool COMP>
void launch_plaquettes(hipStream_t s, int N, const void* phi, const void* cphi, uint16_t* code, int64_t* counts,
                       int64_t* offs, int64_t* total, int64_t* part)
{
    using T2 = typename V2<T>::type;
    const int64_t n3 = (int64_t)N * N * N, nblk = (n3 + PLAQ_NT - 1) / PLAQ_NT;
    hipLaunchKernelGGL((plaquette_kernel<T, COMP>), dim3((unsigned)nblk), dim3(PLAQ_NT), 0, s, N, (const T2*)phi,
                       (const short2*)cphi, code, counts);
    AXG_HIP(hipGetLastError());
    // offsets of the pierced totals (column 15) and the totals of all 16 columns
    launch_scan_counts(s, counts, offs, nblk, total, 15, PLAQ_COLS, part);
}

void launch_plaquette_compact(hipStream_t s, int N, const uint16_t* code, const int64_t* offs, int32_t* rec)
{
    const int64_t n3 = (int64_t)N * N * N, nblk = (n3 + PLAQ_NT - 1) / PLAQ_NT;
    hipLaunchKernelGGL(plaquette_compact_kernel, dim3((unsigned)nblk), dim3(PLAQ_NT), 0, s, N, code, offs, rec);
    AXG_HIP(hipGetLastError());
}

template void launch_plaquettes<double, false>(hipStream_t, int, const void*, const void*, uint16_t*, int64_t*,
                                               int64_t*, int64_t*, int64_t*);
template void launch_plaquettes<float, false>(hipStream_t, int, const void*, const void*, uint16_t*, int64_t*,
                                              int64_t*, int64_t*, int64_t*);
template void launch_plaquettes<float, true>(hipStream_t, int, const void*, const void*, uint16_t*, int64_t*,
                                             int64_t*, int64_t*, int64_t*);

}  // namespace axg
