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

// A block takes a PQ_TX x PQ_TY tile of (x, y) over zc consecutive planes, marching in z: the
// values and quadrants of one site plane of the tile, with the periodic upper neighbours, sit in
// LDS, so every site is loaded and its atan2 taken about (1 + 1/PQ_TX)(1 + 1/PQ_TY)(1 + 1/zc)
// times instead of 7 and 12.  The three faces of a site need its plane and the next.
constexpr int PQ_TX = 32, PQ_TY = 8, PQ_NT = PQ_TX * PQ_TY, PQ_ZC = 32;

template <typename T, bool COMP>
__global__ void __launch_bounds__(PQ_NT) plaquette_kernel(int N, int zc, const typename V2<T>::type* __restrict__ phi,
                                                          const short2* __restrict__ cphi, uint16_t* __restrict__ code)
{
    __shared__ double2 val[2][PQ_TY + 1][PQ_TX + 1];
    __shared__ int8_t qd[2][PQ_TY + 1][PQ_TX + 1];
    const int x0 = blockIdx.x * PQ_TX, y0 = blockIdx.y * PQ_TY, z0 = blockIdx.z * zc;
    const int nx = min(PQ_TX, N - x0), ny = min(PQ_TY, N - y0), nz = min(zc, N - z0);
    const int64_t NN = (int64_t)N * N;
    const int t = threadIdx.x, lx = t % PQ_TX, ly = t / PQ_TX;
    auto load_plane = [&](int buf, int z) {
        for (int k = t; k < (PQ_TY + 1) * (PQ_TX + 1); k += PQ_NT) {
            const int py = k / (PQ_TX + 1), px = k - py * (PQ_TX + 1);
            if (px <= nx && py <= ny) {
                const int gx = x0 + px == N ? 0 : x0 + px, gy = y0 + py == N ? 0 : y0 + py;
                const double2 v = ldf<T, COMP>(phi, cphi, z * NN + (int64_t)gy * N + gx);
                val[buf][py][px] = v;
                qd[buf][py][px] = (int8_t)quadrant(v);
            }
        }
    };
    load_plane(0, z0);
    for (int k = 0; k < nz; k++) {
        const int z = z0 + k, a = k & 1, b = a ^ 1;
        load_plane(b, z + 1 == N ? 0 : z + 1);
        __syncthreads();
        if (lx < nx && ly < ny) {
            const double2 s = val[a][ly][lx], sx = val[a][ly][lx + 1], sy = val[a][ly + 1][lx];
            const double2 sxy = val[a][ly + 1][lx + 1], sz = val[b][ly][lx];
            const double2 syz = val[b][ly + 1][lx], szx = val[b][ly][lx + 1];
            const int q = qd[a][ly][lx], qx = qd[a][ly][lx + 1], qy = qd[a][ly + 1][lx];
            const int qxy = qd[a][ly + 1][lx + 1], qz = qd[b][ly][lx], qyz = qd[b][ly + 1][lx];
            const int qzx = qd[b][ly][lx + 1];
            // (u, v) = (x, y), (y, z), (z, x): corners s, s+u, s+u+v, s+v
            const double2 c1[3] = {sx, sy, sz}, c2[3] = {sxy, syz, szx}, c3[3] = {sy, sz, sx};
            const int q1[3] = {qx, qy, qz}, q2[3] = {qxy, qyz, qzx}, q3[3] = {qy, qz, qx};
            int cd = 0;
#pragma unroll
            for (int o = 0; o < 3; o++) {
                const int n = arc_cut(s, c1[o]) + arc_cut(c1[o], c2[o]) + arc_cut(c2[o], c3[o]) + arc_cut(c3[o], s);
                // defects_3d.f90:856-877: the four corners fill at least three of the quadrants
                const int m = (1 << q) | (1 << q1[o]) | (1 << q2[o]) | (1 << q3[o]);
                cd |= (n + 2) << (3 * o);
                cd |= (__popc(m) >= 3 ? 1 : 0) << (9 + o);
            }
            code[z * NN + (int64_t)(y0 + ly) * N + x0 + lx] = (uint16_t)cd;
        }
        __syncthreads();
    }
}

// Per 256-site block of the site order (the granularity of the compaction): the 16 count
// columns from the face codes, one wave per block.
__global__ void __launch_bounds__(256) plaquette_count_kernel(int64_t n3, int64_t nblk, const uint16_t* __restrict__ code,
                                                              int64_t* __restrict__ counts)
{
    const int64_t b = blockIdx.x * (int64_t)4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (b >= nblk) return;  // uniform per wave
    int c[PLAQ_COLS];
#pragma unroll
    for (int j = 0; j < PLAQ_COLS; j++) c[j] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int64_t i = b * PLAQ_NT + k * 64 + lane;
        if (i < n3) {
            const int cd = code[i];
#pragma unroll
            for (int o = 0; o < 3; o++) {
                const int n = ((cd >> (3 * o)) & 7) - 2, q = (cd >> (9 + o)) & 1;
                c[5 * o + 0] += n != 0;
                c[5 * o + 1] += n == 1;
                c[5 * o + 2] += n == -1;
                c[5 * o + 3] += q;
                c[5 * o + 4] += n;
                c[15] += n != 0;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < PLAQ_COLS; j++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c[j] += __shfl_xor(c[j], o);
    }
    if (lane < PLAQ_COLS) {
        int v = 0;
#pragma unroll
        for (int j = 0; j < PLAQ_COLS; j++) v = lane == j ? c[j] : v;
        counts[b * PLAQ_COLS + lane] = v;
    }
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
    const int zc = min(N, PQ_ZC);
    const dim3 grid((unsigned)((N + PQ_TX - 1) / PQ_TX), (unsigned)((N + PQ_TY - 1) / PQ_TY), (unsigned)((N + zc - 1) / zc));
    hipLaunchKernelGGL((plaquette_kernel<T, COMP>), grid, dim3(PQ_NT), 0, s, N, zc, (const T2*)phi, (const short2*)cphi,
                       code);
    AXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(plaquette_count_kernel, dim3((unsigned)((nblk + 3) / 4)), dim3(256), 0, s, n3, nblk,
                       (const uint16_t*)code, counts);
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
