// Grouping of string and wall points into networks (grouping_strings / grouping_walls,
// defects_3d.f90:1235-1260, 1441-1652, 1735-1760): two points belong to the same group when a
// chain of points joins them with every link at most th_dist apart in the periodic distance of
// lat_dist_periodic (1645-1652).  The groups are the connected components of that graph.
//
// Every point sits in its own lattice cell (one fit per defect cell), in practice within 0.87
// cells per axis of the cell centre.  With `maxoff` the largest such offset, two points within
// th_dist lie in cells at most R = ceil(th_dist + 2 maxoff) apart per axis.  A map cell -> point
// index over the N^3 cells lets each point look only at the cells of that box (the half with a
// positive offset, so each pair is seen once).  A point farther than one cell from its centre
// (a fit that wandered off, or a non-finite one) stays out of the map and is compared with
// every point instead, so the box stays small whatever the fits do.  Thresholds too large for
// the box take one of two other routes: at th_dist >= sqrt(3) N/2 every pair of mapped points
// is linked (their periodic distance cannot exceed that), so they form one group; in between,
// every pair is tested.  Linked points are merged by a lock-free union-find that always hooks the
// larger root under the smaller, so each root ends as the smallest index of its component and
// the labels do not depend on the order in which the threads run.
#include <cstring>

#include "axg_internal.h"

namespace axg {

namespace {

__device__ __forceinline__ int32_t load_parent(const int32_t* parent, int32_t i)
{
    return __hip_atomic_load(parent + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of i, halving the path on the way (a node's parent is replaced by its grandparent).
// Only non-roots are written, and a non-root never becomes a root again, so these stores do
// not race with the root hooks.  Parents only ever decrease (parent[i] <= i), so the walk ends
// within n steps even on a stale read; `budget` bounds it regardless.
__device__ __forceinline__ int32_t find_root(int32_t* parent, int32_t i, int& budget)
{
    while (budget > 0) {
        budget--;
        const int32_t p = load_parent(parent, i);
        if (p == i) return i;
        const int32_t g = load_parent(parent, p);
        if (g == p) return p;
        __hip_atomic_store(parent + i, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i = g;
    }
    return i;
}

// lat_dist_periodic, in the reference's operation order.
__device__ __forceinline__ double lat_dist(const double* a, const double* b, double n)
{
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double dp = fabs(a[k] - b[k]);
        const double m = fmin(dp, n - dp);
        s = s + m * m;
    }
    return sqrt(s);
}

__global__ void map_clear_kernel(int64_t n3, int32_t* __restrict__ map)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        map[i] = -1;
}

constexpr double MAP_OFFSET = 1.0;  // points farther from their cell centre are compared with all

// map[cell of point i] = i (or the point joins the outlier list), parent[i] = i, val[i] = the
// record's value column, and the largest per-axis offset of a mapped point from its cell
// centre (non-negative doubles order as their bit patterns, so an integer atomicMax on the
// bits is exact and order-independent).
__global__ void map_scatter_kernel(int N, const double* __restrict__ rec, int stride, int col,
                                   const int32_t* __restrict__ cells, int64_t n, int32_t* __restrict__ map,
                                   int32_t* __restrict__ parent, double* __restrict__ val,
                                   unsigned long long* __restrict__ maxoff_bits, int32_t* __restrict__ outliers,
                                   int* __restrict__ n_out, int* __restrict__ first_mapped)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double off = 0;
    if (i < n) {
        const int64_t c = cells[i];
        const int64_t NN = (int64_t)N * N;
        const int x = (int)(c % N), y = (int)((c / N) % N), z = (int)(c / NN);
        const double* p = rec + (int64_t)stride * i;
        const double o = fmax(fabs(p[0] - (x + 0.5)), fmax(fabs(p[1] - (y + 0.5)), fabs(p[2] - (z + 0.5))));
        const bool mapped = o <= MAP_OFFSET;  // false for NaN as well
        if (mapped) {
            off = o;
            atomicMin(first_mapped, (int)i);
        } else {
            outliers[atomicAdd(n_out, 1)] = (int32_t)i;
        }
        map[c] = mapped ? (int32_t)i : -1;
        parent[i] = (int32_t)i;
        val[i] = p[col];
    }
    // wave maximum, then one atomic per wave
    for (int s = 32; s > 0; s >>= 1) off = fmax(off, __shfl_xor(off, s));
    if ((threadIdx.x & 63) == 0) atomicMax(maxoff_bits, (unsigned long long)__double_as_longlong(off));
}

// union(u, v): hook the larger root under the smaller; false when the budget ran out.
__device__ __forceinline__ void unite(int32_t* parent, int32_t u, int32_t v, int& budget)
{
    while (budget > 0) {
        int32_t ru = find_root(parent, u, budget), rv = find_root(parent, v, budget);
        if (ru == rv) return;
        if (ru > rv) {
            const int32_t t = ru;
            ru = rv;
            rv = t;
        }
        const int32_t old = atomicCAS(parent + rv, rv, ru);
        if (old == rv) return;
        u = ru;
        v = old;  // rv was hooked meanwhile: continue from its new parent
        budget--;
    }
}

// One thread per point: every point in the positive half of the (2R+1)^3 box of cells around
// its own, within th of it, is merged with it.  Cells whose nearest possible point is farther
// than th are skipped (per axis the points can be no closer than |d| - 2 maxoff).
__global__ void group_link_kernel(int N, const double* __restrict__ rec, int stride, const int32_t* __restrict__ cells,
                                  int64_t n, const int32_t* __restrict__ map, int32_t* __restrict__ parent, double th,
                                  double maxoff, int R, int* __restrict__ fail)
{
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = cells[i];
    const int64_t NN = (int64_t)N * N;
    const int x = (int)(c % N), y = (int)((c / N) % N), z = (int)(c / NN);
    const double* pi = rec + (int64_t)stride * i;
    const double a[3] = {pi[0], pi[1], pi[2]};
    const double th2 = th * th;
    const double two = 2 * maxoff;
    int budget = 1 << 22;  // bounds every wait below; exhausting it reports a failure
    for (int dz = 0; dz <= R; dz++) {
        const double gz = fmax(0.0, dz - two);
        const int cz = (z + dz) % N;
        for (int dy = (dz == 0 ? 0 : -R); dy <= R; dy++) {
            const double gy = fmax(0.0, (dy < 0 ? -dy : dy) - two);
            if (gz * gz + gy * gy > th2) continue;
            const int cy = ((y + dy) % N + N) % N;
            for (int dx = (dz == 0 && dy == 0 ? 1 : -R); dx <= R; dx++) {
                const double gx = fmax(0.0, (dx < 0 ? -dx : dx) - two);
                if (gz * gz + gy * gy + gx * gx > th2) continue;
                const int cx = ((x + dx) % N + N) % N;
                const int32_t j = map[(int64_t)cz * NN + (int64_t)cy * N + cx];
                if (j < 0 || j == (int32_t)i) continue;
                const double* pj = rec + (int64_t)stride * j;
                const double b[3] = {pj[0], pj[1], pj[2]};
                if (!(lat_dist(a, b, (double)N) <= th)) continue;
                unite(parent, (int32_t)i, j, budget);
            }
        }
    }
    if (budget <= 0) atomicAdd(fail, 1);
}

// One thread per point: merge it with every outlier within th (outlier pairs once, from the
// lower index).
__global__ void group_outlier_kernel(int N, const double* __restrict__ rec, int stride, int64_t n,
                                     const int32_t* __restrict__ outliers, int n_out, int32_t* __restrict__ parent,
                                     double th, int* __restrict__ fail)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* pi = rec + (int64_t)stride * i;
    const double a[3] = {pi[0], pi[1], pi[2]};
    int budget = 1 << 22;
    for (int k = 0; k < n_out; k++) {
        const int32_t j = outliers[k];
        if (j == (int32_t)i) continue;
        const double* pj = rec + (int64_t)stride * j;
        const double b[3] = {pj[0], pj[1], pj[2]};
        if (!(lat_dist(a, b, (double)N) <= th)) continue;
        unite(parent, (int32_t)i, j, budget);
    }
    if (budget <= 0) atomicAdd(fail, 1);
}

// Every mapped point joins the first mapped point (all pairs are within th).
__global__ void group_star_kernel(int64_t n, const int32_t* __restrict__ cells, const int32_t* __restrict__ map,
                                  int32_t* __restrict__ parent, int centre, int* __restrict__ fail)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n || map[cells[i]] != (int32_t)i) return;  // outliers are not in the map
    int budget = 1 << 22;
    if ((int32_t)i != centre) unite(parent, (int32_t)i, centre, budget);
    if (budget <= 0) atomicAdd(fail, 1);
}

// Every pair (i, j > i) tested.
__global__ void group_pairs_kernel(int N, const double* __restrict__ rec, int stride, int64_t n,
                                   int32_t* __restrict__ parent, double th, int* __restrict__ fail)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* pi = rec + (int64_t)stride * i;
    const double a[3] = {pi[0], pi[1], pi[2]};
    int budget = 1 << 22;
    for (int64_t j = i + 1; j < n; j++) {
        const double* pj = rec + (int64_t)stride * j;
        const double b[3] = {pj[0], pj[1], pj[2]};
        if (!(lat_dist(a, b, (double)N) <= th)) continue;
        unite(parent, (int32_t)i, (int32_t)j, budget);
    }
    if (budget <= 0) atomicAdd(fail, 1);
}

__global__ void group_label_kernel(int64_t n, int32_t* __restrict__ parent, int32_t* __restrict__ label,
                                   int* __restrict__ fail)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int budget = 1 << 22;
    const int32_t r = find_root(parent, (int32_t)i, budget);
    label[i] = r;
    if (budget <= 0) atomicAdd(fail, 1);
}

}  // namespace

double group_prepare(hipStream_t s, int N, const double* rec, int stride, int col, const int32_t* cells, int64_t n,
                     int32_t* map, int32_t* parent, double* val, unsigned long long* scalars, int32_t* outliers,
                     int* n_out, int* first_mapped)
{
    const int64_t n3 = (int64_t)N * N * N;
    hipLaunchKernelGGL(map_clear_kernel, dim3(4096), dim3(256), 0, s, n3, map);
    AXG_HIP(hipGetLastError());
    // scalars: maxoff bits; outlier count and the first mapped point (two ints)
    unsigned long long init[2] = {0, 0x7fffffff00000000ull};
    AXG_HIP(hipMemcpyAsync(scalars, init, sizeof init, hipMemcpyHostToDevice, s));
    const unsigned nb = (unsigned)((n + 255) / 256);
    int* ints = (int*)(scalars + 1);
    hipLaunchKernelGGL(map_scatter_kernel, dim3(nb), dim3(256), 0, s, N, rec, stride, col, cells, n, map, parent, val,
                       scalars, outliers, ints, ints + 1);
    AXG_HIP(hipGetLastError());
    unsigned long long out[2] = {0, 0};
    AXG_HIP(hipMemcpyAsync(out, scalars, sizeof out, hipMemcpyDeviceToHost, s));
    AXG_HIP(hipStreamSynchronize(s));
    double m;
    std::memcpy(&m, &out[0], sizeof m);
    int k[2];
    std::memcpy(k, &out[1], sizeof k);
    *n_out = k[0];
    *first_mapped = k[1];
    return m;
}

void group_link(hipStream_t s, int N, const double* rec, int stride, const int32_t* cells, int64_t n,
                const int32_t* map, int32_t* parent, int32_t* label, double th, double maxoff, int R, int mode,
                int centre, const int32_t* outliers, int n_out, int* fail)
{
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (mode == 0) {
        hipLaunchKernelGGL(group_link_kernel, dim3(nb), dim3(256), 0, s, N, rec, stride, cells, n, map, parent, th,
                           maxoff, R, fail);
    } else if (mode == 1) {
        hipLaunchKernelGGL(group_star_kernel, dim3(nb), dim3(256), 0, s, n, cells, map, parent, centre, fail);
    } else {
        hipLaunchKernelGGL(group_pairs_kernel, dim3(nb), dim3(256), 0, s, N, rec, stride, n, parent, th, fail);
    }
    AXG_HIP(hipGetLastError());
    if (n_out > 0 && mode != 2) {
        hipLaunchKernelGGL(group_outlier_kernel, dim3(nb), dim3(256), 0, s, N, rec, stride, n, outliers, n_out, parent,
                           th, fail);
        AXG_HIP(hipGetLastError());
    }
    // the outlier list lives in `label`: labels are written last
    hipLaunchKernelGGL(group_label_kernel, dim3(nb), dim3(256), 0, s, n, parent, label, fail);
    AXG_HIP(hipGetLastError());
}

}  // namespace axg
