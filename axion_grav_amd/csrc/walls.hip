// Wall cells: plane fit and area, one thread per wall cell (id_defects' wall loop,
// defects_3d.f90:183-188, with position_wall 1058-1202), plus the per-block sums of
// add_wall_area (2145-2204, group_walls = F).
//
// position_wall takes the points where sin(ph) changes sign on the cell's 12 edges.  Here each
// edge owns a fixed slot in the reference's edge order (4 x-edges, 4 z-edges, 4 y-edges) and
// a flag, so every array stays in registers: the sums over the valid points run in slot order
// (the order of the reference's compacted list), and the angular bubble sort runs over all 12
// slots with +inf in the empty ones (a stable sort, so the valid points come out in the
// reference's order).  Only the interpolated coordinate of each point is stored; the other
// two are the edge's fixed corners.
#include "axg_internal.h"

namespace axg {

namespace {

constexpr double PI_T = 3.14159265359;
constexpr double TWOPI_T = 2 * PI_T;

// Point of edge slot e (0..11) with interpolated coordinate t, relative to the cell centre.
__device__ __forceinline__ void edge_point(int e, double t, double& px, double& py, double& pz)
{
    const int g = e >> 2, a = (e >> 1) & 1, b = e & 1;
    if (g == 0) { px = t; py = b - 0.5; pz = a - 0.5; }        // x-edge: slot iz*2+iy
    else if (g == 1) { px = b - 0.5; py = a - 0.5; pz = t; }   // z-edge: slot 4+iy*2+ix
    else { px = b - 0.5; py = t; pz = a - 0.5; }               // y-edge: slot 8+iz*2+ix
}

// The search grid of position_wall: 20 x 20 directions (mu, phi) with mu = (2 imu - 21) / 20 and
// phi = pi_t (2 iphi - 21) / 20, the same for every cell.  A block evaluates sqrt(1 - mu^2) and
// cos / sin of phi once into LDS (the same library calls as the per-cell loop made, so the same
// bits); the loop then costs 12 dot products per direction instead of three transcendentals.
constexpr int WALL_GRID = 20;
struct WallGrid {
    double smu[WALL_GRID];         // sqrt(1 - mu^2) of mu index imu - 1
    double cs[WALL_GRID][2];       // cos, sin of phi index iphi - 1
};

__device__ __forceinline__ double grid_mu(int imu) { return (-20 + 2 * imu - 1.0) / 20; }
__device__ __forceinline__ double grid_phi(int iphi) { return PI_T * (2 * iphi - 1.0 - 20) / 20; }

// position_wall; returns the number m of edge points (m < 3: no fit, outputs untouched).
__device__ int position_wall(const double ph[8], const WallGrid& G, double pos[3], double dir[3], double& gof,
                             double& wa)
{
#pragma clang fp contract(off)
    double sn[8];
#pragma unroll
    for (int c = 0; c < 8; c++) sn[c] = sin(ph[c]);
    double t[12];
    bool ok[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
        const int g = e >> 2, a = (e >> 1) & 1, b = e & 1;
        int c0, c1;
        if (g == 0) { c0 = (a * 2 + b) * 2; c1 = c0 + 1; }          // (0,iy,iz)-(1,iy,iz), iz = a, iy = b
        else if (g == 1) { c0 = a * 2 + b; c1 = c0 + 4; }           // (ix,iy,0)-(ix,iy,1), iy = a, ix = b
        else { c0 = a * 4 + b; c1 = c0 + 2; }                        // (ix,0,iz)-(ix,1,iz), iz = a, ix = b
        const double u = sn[c0], v = sn[c1];
        ok[e] = u * v < 0;
        t[e] = ok[e] ? (u + v) / (u - v) / 2 : 0.0;
    }
    int m = 0;
#pragma unroll
    for (int e = 0; e < 12; e++) m += ok[e] ? 1 : 0;
    if (m < 3) return m;

    double chi2min = 6.0, mut = 0, wt = 0;
    int bphi = 0;  // phi index of the best direction (0: none below chi2 = 6)
    for (int imu = 1; imu <= WALL_GRID; imu++) {
        const double mu = grid_mu(imu), smu = G.smu[imu - 1];
        for (int iphi = 1; iphi <= WALL_GRID; iphi++) {
            const double d0 = smu * G.cs[iphi - 1][0], d1 = smu * G.cs[iphi - 1][1], d2 = mu;
            double w = 0, chi2 = 0;
#pragma unroll
            for (int e = 0; e < 12; e++) {
                if (!ok[e]) continue;
                double px, py, pz;
                edge_point(e, t[e], px, py, pz);
                const double d = px * d0 + py * d1 + pz * d2;
                w = w + d;
                chi2 = chi2 + d * d;
            }
            if (w < 0) w = 0;
            else w = w / m;
            chi2 = chi2 - m * w * w;
            if (chi2 < chi2min) {
                chi2min = chi2;
                mut = mu;
                bphi = iphi;
                wt = w;
            }
        }
    }
    const double phit = bphi > 0 ? grid_phi(bphi) : 0.0;
    double dr[3] = {sqrt(1 - mut * mut) * cos(phit), sqrt(1 - mut * mut) * sin(phit), mut};
    int num = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const double d = ((c & 1) - 0.5) * dr[0] + (((c >> 1) & 1) - 0.5) * dr[1] + ((c >> 2) - 0.5) * dr[2] - wt;
        if (d * cos(ph[c]) > 0) num++;
    }
    if (num <= 4) {
        dr[0] = -dr[0];
        dr[1] = -dr[1];
        dr[2] = -dr[2];
        wt = -wt;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        dir[k] = dr[k];
        pos[k] = wt * dr[k];
    }
    gof = exp(-chi2min / 2);

    const double e1[3] = {mut * cos(phit), mut * sin(phit), -sqrt(1 - mut * mut)};
    const double e2[3] = {-sin(phit), cos(phit), 0.0};
    double th[12], u2[12], v2[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
        double px, py, pz;
        edge_point(e, t[e], px, py, pz);
        px = px - pos[0];
        py = py - pos[1];
        pz = pz - pos[2];
        const double u = e1[0] * px + e1[1] * py + e1[2] * pz;
        const double v = e2[0] * px + e2[1] * py + e2[2] * pz;
        th[e] = ok[e] ? atan2(v, u) : __builtin_inf();
        u2[e] = u;
        v2[e] = v;
    }
#pragma unroll
    for (int i = 10; i >= 0; i--)
#pragma unroll
        for (int j = 0; j <= i; j++)
            if (th[j] > th[j + 1]) {
                double a = th[j];
                th[j] = th[j + 1];
                th[j + 1] = a;
                a = u2[j];
                u2[j] = u2[j + 1];
                u2[j + 1] = a;
                a = v2[j];
                v2[j] = v2[j + 1];
                v2[j + 1] = a;
            }
    double thl = 0, ul = 0, vl = 0;  // the last valid point (slot m-1 after the sort)
#pragma unroll
    for (int e = 0; e < 12; e++)
        if (e == m - 1) {
            thl = th[e];
            ul = u2[e];
            vl = v2[e];
        }
    double area = 0, w = thl - TWOPI_T, u = sqrt(ul * ul + vl * vl);
#pragma unroll
    for (int i = 0; i < 12; i++) {
        if (i < m) {
            const double v = sqrt(u2[i] * u2[i] + v2[i] * v2[i]);
            w = th[i] - w;
            area = area + 0.5 * u * v * sin(w);
            u = v;
            w = th[i];
        }
    }
    wa = area;
    return m;
}

template <typename T, bool COMP>
__global__ void __launch_bounds__(WALL_NT) wall_fit_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                           const short2* __restrict__ cphi,
                                                           const int32_t* __restrict__ cells, int64_t nw,
                                                           double* __restrict__ rec, double* __restrict__ partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[2][WALL_NT];
    __shared__ WallGrid grid;
    static_assert(WALL_NT >= WALL_GRID, "one thread per grid value");
    if (threadIdx.x < WALL_GRID) {
        const double mu = grid_mu(threadIdx.x + 1), phi = grid_phi(threadIdx.x + 1);
        grid.smu[threadIdx.x] = sqrt(1 - mu * mu);
        grid.cs[threadIdx.x][0] = cos(phi);
        grid.cs[threadIdx.x][1] = sin(phi);
    }
    __syncthreads();
    const int64_t s = blockIdx.x * (int64_t)WALL_NT + threadIdx.x;
    double acc[2] = {0, 0};
    if (s < nw) {
        const int64_t NN = (int64_t)N * N;
        const int64_t i = cells[s];
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
        double ph[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            int cx = x + (c & 1), cy = y + ((c >> 1) & 1), cz = z + (c >> 2);
            cx = cx == N ? 0 : cx;
            cy = cy == N ? 0 : cy;
            cz = cz == N ? 0 : cz;
            const double2 f = ldf<T, COMP>(phi, cphi, (int64_t)cz * NN + (int64_t)cy * N + cx);
            ph[c] = atan2(f.y, f.x);
        }
        double pos[3] = {0, 0, 0}, dir[3] = {0, 0, 0}, gof = 0, wa = 0;
        double* o = rec + 8 * s;
        if (position_wall(ph, grid, pos, dir, gof, wa) < 3) {
            acc[1] = 1;
#pragma unroll
            for (int k = 0; k < 8; k++) o[k] = 0;
        } else {
            o[0] = pos[0] + ((double)(x + 1) - 0.5);
            o[1] = pos[1] + ((double)(y + 1) - 0.5);
            o[2] = pos[2] + ((double)(z + 1) - 0.5);
            o[3] = dir[0];
            o[4] = dir[1];
            o[5] = dir[2];
            o[6] = gof;
            o[7] = wa;
            acc[0] = wa;
        }
    }
    sh[0][threadIdx.x] = acc[0];
    sh[1][threadIdx.x] = acc[1];
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0;
        for (int j = 0; j < WALL_NT; j++) t += sh[threadIdx.x][j];
        partial[(int64_t)blockIdx.x * 2 + threadIdx.x] = t;
    }
}

}  // namespace

template <typename T, bool COMP>
void launch_wall_fit(hipStream_t s, int N, const void* phi, const void* cphi, const int32_t* cells, int64_t nw,
                     double* rec, double* partial)
{
    using T2 = typename V2<T>::type;
    if (nw <= 0) return;
    const int64_t nblk = (nw + WALL_NT - 1) / WALL_NT;
    hipLaunchKernelGGL((wall_fit_kernel<T, COMP>), dim3((unsigned)nblk), dim3(WALL_NT), 0, s, N, (const T2*)phi,
                       (const short2*)cphi, cells, nw, rec, partial);
    AXG_HIP(hipGetLastError());
}

template void launch_wall_fit<double, false>(hipStream_t, int, const void*, const void*, const int32_t*, int64_t,
                                             double*, double*);
template void launch_wall_fit<float, false>(hipStream_t, int, const void*, const void*, const int32_t*, int64_t,
                                            double*, double*);
template void launch_wall_fit<float, true>(hipStream_t, int, const void*, const void*, const int32_t*, int64_t,
                                           double*, double*);

}  // namespace axg
