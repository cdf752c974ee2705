// Checkpoint diagnostics: defect cell classification and mask compaction
// (defects_3d.f90:67-103, 152-219, 309-349), the masked field spectrum and its
// deconvolution reductions (spectrum.f90:344-567, 900-954), the lattice energy, and the
// field conversions / initial-field helpers.  Every floating-point reduction is
// deterministic: fixed per-block partial order, then a fixed-order sum over blocks.
#include "axg_internal.h"

namespace axg {

constexpr int RED_NT = 256;

// Fixed-order tree reduction of one double per thread; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = RED_NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// out[b] = sum over the nwg rows of column b of partial[nwg][width]: one block per column,
// thread t sums rows t, t + 256, ... in order, then a fixed-shape tree (deterministic).
__global__ void __launch_bounds__(256) sum_columns_kernel(const double* __restrict__ partial, int nwg, int width,
                                                          double* __restrict__ out)
{
    __shared__ double sh[256];
    const int b = blockIdx.x;
    double s = 0;
    for (int w = threadIdx.x; w < nwg; w += 256) s += partial[(int64_t)w * width + b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = sh[0];
}

void sum_columns(hipStream_t s, const double* partial, int nwg, int width, double* out)
{
    hipLaunchKernelGGL(sum_columns_kernel, dim3(width), dim3(256), 0, s, partial, nwg, width, out);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ energy

template <typename T, bool COMP>
__global__ void __launch_bounds__(RED_NT) energy_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                        const short2* __restrict__ cphi,
                                                        const typename V2<T>::type* __restrict__ pi,
                                                        const short2* __restrict__ cpi, double m2, double ma2,
                                                        double kt, double* __restrict__ partial)
{
    __shared__ double sh[RED_NT];
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    double ek = 0, eg = 0, ev = 0;
    for (int64_t i = blockIdx.x * (int64_t)RED_NT + threadIdx.x; i < n3; i += (int64_t)gridDim.x * RED_NT) {
        int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
        double2 c = ldf<T, COMP>(phi, cphi, i);
        double2 p = ldf<T, COMP>(pi, cpi, i);
        ek += p.x * p.x + p.y * p.y;
        double2 nb[3] = {ldf<T, COMP>(phi, cphi, i - x + (x + 1 == N ? 0 : x + 1)),
                         ldf<T, COMP>(phi, cphi, i + ((y + 1 == N ? 0 : y + 1) - y) * (int64_t)N),
                         ldf<T, COMP>(phi, cphi, i + ((z + 1 == N ? 0 : z + 1) - z) * NN)};
        for (int d = 0; d < 3; d++) {
            double dr = nb[d].x - c.x, di = nb[d].y - c.y;
            eg += dr * dr + di * di;
        }
        double s2 = c.x * c.x + c.y * c.y;
        double h = s2 + m2 / 2;
        double v = h * h - ma2 * c.x;
        if (kt != 0.0) v = v - kt * (c.x * c.x - c.y * c.y);  // N_DW=2 tilt
        ev += v;
    }
    double a = block_sum(ek, sh), b = block_sum(eg, sh), c = block_sum(ev, sh);
    if (threadIdx.x == 0) {
        partial[(int64_t)blockIdx.x * 3 + 0] = a;
        partial[(int64_t)blockIdx.x * 3 + 1] = b;
        partial[(int64_t)blockIdx.x * 3 + 2] = c;
    }
}

template <typename T, bool COMP>
void launch_energy(hipStream_t s, int N, const void* phi, const void* cphi, const void* pi, const void* cpi, double m2,
                   double ma2, double kt, double* partial, int nwg, double* out)
{
    using T2 = typename V2<T>::type;
    hipLaunchKernelGGL((energy_kernel<T, COMP>), dim3(nwg), dim3(RED_NT), 0, s, N, (const T2*)phi, (const short2*)cphi,
                       (const T2*)pi, (const short2*)cpi, m2, ma2, kt, partial);
    AXG_HIP(hipGetLastError());
    sum_columns(s, partial, nwg, 3, out);
}

// ------------------------------------------------------------------ defect classification
// Corner i-1 of the cell at (x,y,z) has offsets (bit0, bit1, bit2) in (x, y, z)
// (defects_3d.f90:38-40); the sort below yields the same ordered list whatever the corner
// order, so only the phases matter.

__device__ __forceinline__ int id_with_phase(double* p, int id_strings, int id_walls, int spec_est)
{
    // defects_3d.f90:309-349 sorts the eight phases by bubble sort; Batcher's odd-even merge
    // network (19 min/max pairs) gives the same ordered list (the phases are never NaN, and a
    // swapped +0 / -0 pair changes no comparison below).
    constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
    for (int k = 0; k < 19; k++) {
        const double a = p[net[k][0]], b = p[net[k][1]];
        p[net[k][0]] = fmin(a, b);
        p[net[k][1]] = fmax(a, b);
    }
    int flag = 0;
    double t = p[0] - p[7] + AXG_TWOPI_T;
    double tmax = p[1] - p[0];
#pragma unroll
    for (int i = 2; i < 8; i++) { double d = p[i] - p[i - 1]; tmax = d > tmax ? d : tmax; }
    if (id_walls && tmax > t) flag += 2;
    else tmax = tmax > t ? tmax : t;
    if (id_strings && tmax < AXG_PI_T) flag += 1;
    if (spec_est && tmax < 7 * AXG_PI_T / 4) flag += 4;
    return flag;
}

template <typename T, bool COMP>
__device__ __forceinline__ double site_phase(const typename V2<T>::type* __restrict__ phi,
                                             const short2* __restrict__ cphi, int64_t i)
{
    const double2 v = ldf<T, COMP>(phi, cphi, i);
    return atan2(v.y, v.x);  // aimag(log(phi)), defects_3d.f90:92
}

// A block classifies the cells of a CL_TX x CL_TY tile of (x, y) over zc consecutive planes,
// marching in z.  The phases of one site plane of the tile, (CL_TX+1) x (CL_TY+1) with the
// periodic upper neighbours, sit in LDS; each step computes the next plane's phases once and
// classifies the cells between the two planes, so a site's atan2 is evaluated about
// (1 + 1/CL_TX)(1 + 1/CL_TY)(1 + 1/zc) times instead of 8.
constexpr int CL_TX = 32, CL_TY = 8, CL_NT = CL_TX * CL_TY, CL_ZC = 32;

template <typename T, bool COMP>
__global__ void __launch_bounds__(CL_NT) classify_kernel(int N, int zc, const typename V2<T>::type* __restrict__ phi,
                                                         const short2* __restrict__ cphi, int id_strings,
                                                         int id_walls, int spec_est, int8_t* __restrict__ flags)
{
#pragma clang fp contract(off)
    __shared__ double ph[2][CL_TY + 1][CL_TX + 1];
    const int x0 = blockIdx.x * CL_TX, y0 = blockIdx.y * CL_TY, z0 = blockIdx.z * zc;
    const int nx = min(CL_TX, N - x0), ny = min(CL_TY, N - y0), nz = min(zc, N - z0);
    const int64_t NN = (int64_t)N * N;
    const int t = threadIdx.x, lx = t % CL_TX, ly = t / CL_TX;
    auto load_plane = [&](int buf, int z) {
        for (int k = t; k < (CL_TY + 1) * (CL_TX + 1); k += CL_NT) {
            const int py = k / (CL_TX + 1), px = k - py * (CL_TX + 1);
            if (px <= nx && py <= ny) {
                const int gx = x0 + px == N ? 0 : x0 + px, gy = y0 + py == N ? 0 : y0 + py;
                ph[buf][py][px] = site_phase<T, COMP>(phi, cphi, z * NN + (int64_t)gy * N + gx);
            }
        }
    };
    load_plane(0, z0);
    for (int k = 0; k < nz; k++) {
        const int z = z0 + k, a = k & 1, b = a ^ 1;
        load_plane(b, z + 1 == N ? 0 : z + 1);
        __syncthreads();
        if (lx < nx && ly < ny) {
            // corner c has offsets (bit0, bit1, bit2) in (x, y, z), defects_3d.f90:38-40
            double p[8];
#pragma unroll
            for (int c = 0; c < 8; c++) p[c] = ph[(c & 4) ? b : a][ly + ((c >> 1) & 1)][lx + (c & 1)];
            flags[z * NN + (int64_t)(y0 + ly) * N + x0 + lx] = (int8_t)id_with_phase(p, id_strings, id_walls, spec_est);
        }
        __syncthreads();
    }
}

// Per 256-cell block of the site order (the granularity of cell_compact_kernel), the number of
// cells with flag bit 1, 2 and 4; one wave per block, the three counts packed 10 bits apart.
__global__ void __launch_bounds__(256) flag_count_kernel(int64_t n3, int64_t nblk, const int8_t* __restrict__ flags,
                                                         int64_t* __restrict__ counts)
{
    const int64_t b = blockIdx.x * (int64_t)4 + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    if (b >= nblk) return;  // uniform per wave
    int packed = 0;
    if (b * 256 + 256 <= n3) {  // a whole block: four flags per lane in one load
        const uint32_t f4 = *(const uint32_t*)(flags + b * 256 + 4 * lane);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t f = f4 >> (8 * k);
            packed += (f & 1) | (((f >> 1) & 1) << 10) | (((f >> 2) & 1) << 20);
        }
    } else {
        for (int k = 0; k < 4; k++) {
            const int64_t i = b * 256 + k * 64 + lane;
            if (i < n3) {
                const int f = flags[i];
                packed += (f & 1) | (((f >> 1) & 1) << 10) | (((f >> 2) & 1) << 20);
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) packed += __shfl_xor(packed, o);
    if (lane == 0) {
        counts[b * 3 + 0] = packed & 1023;
        counts[b * 3 + 1] = (packed >> 10) & 1023;
        counts[b * 3 + 2] = packed >> 20;
    }
}

template <typename T, bool COMP>
void launch_classify(hipStream_t s, int N, const void* phi, const void* cphi, int id_strings, int id_walls, int spec_est,
                     int8_t* flags, int64_t* block_counts, int64_t nblk)
{
    using T2 = typename V2<T>::type;
    const int zc = min(N, CL_ZC);
    const dim3 grid((unsigned)((N + CL_TX - 1) / CL_TX), (unsigned)((N + CL_TY - 1) / CL_TY), (unsigned)((N + zc - 1) / zc));
    hipLaunchKernelGGL((classify_kernel<T, COMP>), grid, dim3(CL_NT), 0, s, N, zc, (const T2*)phi, (const short2*)cphi,
                       id_strings, id_walls, spec_est, flags);
    AXG_HIP(hipGetLastError());
    const int64_t n3 = (int64_t)N * N * N;
    hipLaunchKernelGGL(flag_count_kernel, dim3((unsigned)((nblk + 3) / 4)), dim3(256), 0, s, n3, nblk,
                       (const int8_t*)flags, block_counts);
    AXG_HIP(hipGetLastError());
}

// Exclusive scan of column col of the per-block counts in[n][C] and the totals of all C columns
// (the classifier's C = 3: strings, walls, mask; the plaquettes' C = 16), over the whole chip:
// one block per SCAN_ROWS rows sums its rows, one block scans those partial sums, and every
// block then scans its own rows from its partial's offset.  Integer sums, so the result does
// not depend on the split.  part holds (C + 1) * ceil(n / SCAN_ROWS) values.
constexpr int SCAN_ROWS = 1024;

template <int C>
__global__ void __launch_bounds__(256) count_partials_kernel(const int64_t* __restrict__ in, int64_t n,
                                                             int64_t* __restrict__ part)
{
    __shared__ int64_t sh[4][C];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int64_t s[C];
#pragma unroll
    for (int c = 0; c < C; c++) s[c] = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ROWS / 256; k++) {
        const int64_t r = blockIdx.x * (int64_t)SCAN_ROWS + k * 256 + t;
        if (r < n)
#pragma unroll
            for (int c = 0; c < C; c++) s[c] += in[r * C + c];
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[c] += __shfl_xor(s[c], o);
        if (lane == 0) sh[w][c] = s[c];
    }
    __syncthreads();
    if (t < C) part[blockIdx.x * (int64_t)C + t] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}

template <int C>
__global__ void __launch_bounds__(1024) partials_scan_kernel(const int64_t* __restrict__ part, int64_t G, int col,
                                                             int64_t* __restrict__ poff, int64_t* __restrict__ total)
{
    __shared__ int64_t sh[1024];
    const int t = threadIdx.x;
    const int64_t chunk = (G + 1023) / 1024, g0 = t * chunk, g1 = min(G, g0 + chunk);
    for (int c = 0; c < C; c++) {
        int64_t s = 0;
        for (int64_t g = g0; g < g1; g++) s += part[g * C + c];
        sh[t] = s;
        __syncthreads();
        for (int h = 512; h > 0; h >>= 1) {
            if (t < h) sh[t] += sh[t + h];
            __syncthreads();
        }
        if (t == 0) total[c] = sh[0];
        __syncthreads();
    }
    int64_t s = 0;
    for (int64_t g = g0; g < g1; g++) s += part[g * C + col];
    sh[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
        const int64_t v = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int64_t run = sh[t] - s;
    for (int64_t g = g0; g < g1; g++) {
        poff[g] = run;
        run += part[g * C + col];
    }
}

template <int C>
__global__ void __launch_bounds__(256) scan_rows_kernel(const int64_t* __restrict__ in, int64_t n, int col,
                                                        const int64_t* __restrict__ poff, int64_t* __restrict__ out)
{
    __shared__ int64_t ws[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t r0 = blockIdx.x * (int64_t)SCAN_ROWS + 4 * t;  // four consecutive rows per thread
    int64_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = r0 + k < n ? in[(r0 + k) * C + col] : 0;
        s += v[k];
    }
    int64_t inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t u = __shfl_up(inc, d);
        if (lane >= d) inc += u;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int64_t run = poff[blockIdx.x] + inc - s;
    for (int k = 0; k < w; k++) run += ws[k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (r0 + k < n) out[r0 + k] = run;
        run += v[k];
    }
}

int64_t scan_scratch(int64_t n, int ncol) { return (ncol + 1) * ((n + SCAN_ROWS - 1) / SCAN_ROWS); }

template <int C>
static void scan_counts_c(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total, int col,
                          int64_t* part)
{
    const int64_t G = (n + SCAN_ROWS - 1) / SCAN_ROWS;
    int64_t* poff = part + C * G;
    hipLaunchKernelGGL(count_partials_kernel<C>, dim3((unsigned)G), dim3(256), 0, s, in, n, part);
    AXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(partials_scan_kernel<C>, dim3(1), dim3(1024), 0, s, (const int64_t*)part, G, col, poff, total);
    AXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_rows_kernel<C>, dim3((unsigned)G), dim3(256), 0, s, in, n, col, (const int64_t*)poff, out);
    AXG_HIP(hipGetLastError());
}

void launch_scan_counts(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total, int col, int ncol,
                        int64_t* part)
{
    if (ncol == 3) scan_counts_c<3>(s, in, out, n, total, col, part);
    else if (ncol == PLAQ_COLS) scan_counts_c<PLAQ_COLS>(s, in, out, n, total, col, part);
    else throw std::runtime_error("scan over an unsupported column count");
}

// Cells of block b with flag bit `bit` go to offs[b] + (rank among the block's such cells), so
// the list is in site order: z outermost, x fastest -- the loop order of defects_3d.f90:167-217.
// triplets != 0: 1-based (ix, iy, iz) per cell (the mask list); else the site index.
__global__ void __launch_bounds__(256) cell_compact_kernel(int N, const int8_t* __restrict__ flags, int bit,
                                                           const int64_t* __restrict__ offs, int32_t* __restrict__ out,
                                                           int triplets)
{
    __shared__ int sh[256];
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    int m = (i < n3 && (flags[i] & bit)) ? 1 : 0;
    sh[threadIdx.x] = m;
    __syncthreads();
    // inclusive scan (Hillis-Steele) over 256 entries
    for (int d = 1; d < 256; d <<= 1) {
        int v = (int)threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    if (m) {
        int64_t pos = offs[blockIdx.x] + sh[threadIdx.x] - 1;
        if (triplets) {
            int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
            out[3 * pos + 0] = x + 1;
            out[3 * pos + 1] = y + 1;
            out[3 * pos + 2] = z + 1;
        } else {
            out[pos] = (int32_t)i;  // N <= 1024: site indices fit in 31 bits
        }
    }
}

void launch_cell_compact(hipStream_t s, int N, const int8_t* flags, int bit, const int64_t* block_offsets,
                         int32_t* out, int triplets, int64_t nblk)
{
    hipLaunchKernelGGL(cell_compact_kernel, dim3((unsigned)nblk), dim3(256), 0, s, N, flags, bit, block_offsets, out,
                       triplets);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ field spectrum
// out = (mask ? 0 : pi) or (mask ? 0 : fill) as FP64 complex (spectrum.f90:358-374, 488-515).

template <typename T, bool COMP>
__global__ void masked_field_kernel(int64_t n3, const typename V2<T>::type* __restrict__ pi,
                                    const short2* __restrict__ cpi, const int8_t* __restrict__ flags,
                                    double2* __restrict__ out, double fill, int use_pi)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = use_pi ? ldf<T, COMP>(pi, cpi, i) : double2{fill, 0.0};
        if (flags[i] & 4) v = double2{0.0, 0.0};
        out[i] = v;
    }
}

template <typename T, bool COMP>
void launch_masked_field(hipStream_t s, int N, const void* pi, const void* cpi, const int8_t* flags, double2* out,
                         double fill_unmasked, bool use_pi)
{
    using T2 = typename V2<T>::type;
    int64_t n3 = (int64_t)N * N * N;
    hipLaunchKernelGGL((masked_field_kernel<T, COMP>), dim3(65536), dim3(256), 0, s, n3, (const T2*)pi,
                       (const short2*)cpi, flags, out, fill_unmasked, use_pi ? 1 : 0);
    AXG_HIP(hipGetLastError());
}

// Shell binning over the full spectrum (spectrum.f90:530-556).  Row (ky,kz) is folded onto
// j = |nx| = 0..N/2 (kx = j and kx = N-j), along which the bin is non-decreasing, then every
// run of equal bins is summed by the thread at its start (same deterministic scheme as the
// GW binning).  pow[b] += |d*amp|^2 (times dv_wav on the host), num[b] += 1 per mode.
// As in the GW binning, each thread loads its values of the next row while the current row is
// binned (only modes up to the last bin's |n|^2, without waiting for the bin lookup), and each run
// is summed over the length its last element recorded; the sums and their order are unchanged.
constexpr int SPEC_BIN_NT = 256, SPEC_BIN_KPT = 3;  // folded j values per thread: N / 2 + 1 <= 768

__global__ void __launch_bounds__(SPEC_BIN_NT) spec_bin_kernel(int N, const double2* __restrict__ d,
                                                               const uint16_t* __restrict__ bin_of_n2, int64_t n2max,
                                                               int nb, double amp, double* __restrict__ partial)
{
    extern __shared__ double sh[];
    double* bins = sh;
    double* v = sh + 2 * nb;
    double* cnt = v + (N / 2 + 1);
    int* bb = (int*)(cnt + (N / 2 + 1));
    int* rend = bb + (N / 2 + 1);  // nb: one past the last j of bin b's run in the row
    const int nh = N / 2 + 1;
    const int64_t NN = (int64_t)N * N;
    for (int b = threadIdx.x; b < 2 * nb; b += blockDim.x) bins[b] = 0;
    double2 a[SPEC_BIN_KPT], e[SPEC_BIN_KPT];
    int bq[SPEC_BIN_KPT];
    auto fetch = [&](int64_t row) {
        const int kz = (int)(row / N), ky = (int)(row - (int64_t)kz * N);
        const int ny = wrapn(ky, N), nz = wrapn(kz, N);
        const double2* r = d + row * N;
#pragma unroll
        for (int i = 0; i < SPEC_BIN_KPT; i++) {
            const int j = threadIdx.x + i * SPEC_BIN_NT;
            if (j < nh) {
                const int64_t n2 = (int64_t)j * j + (int64_t)ny * ny + (int64_t)nz * nz;
                bq[i] = bin_of_n2[n2];
                if (n2 <= n2max) {
                    a[i] = r[j];
                    if (j > 0 && j < N / 2) e[i] = r[N - j];
                }
            }
        }
    };
    int64_t row = blockIdx.x;
    if (row < NN) fetch(row);
    __syncthreads();
    for (; row < NN; row += gridDim.x) {
#pragma unroll
        for (int i = 0; i < SPEC_BIN_KPT; i++) {
            const int j = threadIdx.x + i * SPEC_BIN_NT;
            if (j < nh) {
                double s = 0, c = 0;
                if (bq[i] > 0) {
                    double ax = a[i].x * amp, ay = a[i].y * amp;
                    s = ax * ax + ay * ay;
                    c = 1;
                    if (j > 0 && j < N / 2) {
                        double ex = e[i].x * amp, ey = e[i].y * amp;
                        s += ex * ex + ey * ey;
                        c = 2;
                    }
                }
                v[j] = s;
                cnt[j] = c;
                bb[j] = bq[i];
            }
        }
        if (row + gridDim.x < NN) fetch(row + gridDim.x);
        __syncthreads();
        for (int j = threadIdx.x; j < nh; j += blockDim.x) {
            const int b = bb[j];
            if (b > 0 && (j == nh - 1 || bb[j + 1] != b)) rend[b - 1] = j + 1;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < nh; j += blockDim.x) {
            int b = bb[j];
            if (b > 0 && (j == 0 || bb[j - 1] != b)) {
                const int end = rend[b - 1];
                double sn = 0, sd = 0;
                for (int k = j; k < end; k++) { sn += v[k]; sd += cnt[k]; }
                bins[b - 1] += sn;
                bins[nb + b - 1] += sd;
            }
        }
        __syncthreads();
    }
    for (int b = threadIdx.x; b < 2 * nb; b += blockDim.x) partial[(int64_t)blockIdx.x * 2 * nb + b] = bins[b];
}

void launch_spec_bin(hipStream_t s, int N, const double2* data, const uint16_t* bin_of_n2, int64_t n2max, int nb,
                     double scale, double* partial, int nwg, double* out)
{
    int nh = N / 2 + 1;
    if (nh > SPEC_BIN_KPT * SPEC_BIN_NT) throw Error(AXG_EINVAL, "field spectrum: N too large for the binning kernel");
    size_t lds = (size_t)(2 * nb + 2 * nh) * sizeof(double) + (size_t)(nh + nb) * sizeof(int);
    hipLaunchKernelGGL(spec_bin_kernel, dim3(nwg), dim3(SPEC_BIN_NT), lds, s, N, data, bin_of_n2, n2max, nb, scale,
                       partial);
    AXG_HIP(hipGetLastError());
    sum_columns(s, partial, nwg, 2 * nb, out);
}

// Histogram of |W|^2 by n^2 = |n|^2, deterministic: block kz walks its plane by a = |ny|
// (a = 0 .. N/2, in order).  The rows ky = a and ky = N - a share a, so their values at each
// folded j = |nx| land on the same m = j^2 + a^2 and are added together first; within one a the
// j hit distinct m, so the per-plane partial plane[kz][m] needs no atomics.  Each thread loads
// its values of the next a while the current one is added (the input is read-only).
// hist[n2] = sum over kz (in order) of plane[kz][n2 - nz^2].
constexpr int N2P_NT = 256, N2P_KPT = 3;  // folded j values per thread: N / 2 + 1 <= 768

__global__ void __launch_bounds__(N2P_NT) n2_plane_kernel(int N, const double2* __restrict__ d,
                                                          double* __restrict__ plane)
{
    const int nh = N / 2 + 1;
    const int64_t mm = (int64_t)2 * (N / 2) * (N / 2) + 1;
    const int kz = blockIdx.x;
    double* pp = plane + (int64_t)kz * mm;
    for (int64_t m = threadIdx.x; m < mm; m += blockDim.x) pp[m] = 0;
    double2 u[N2P_KPT][4];  // rows a and N - a, each at j and N - j
    auto fetch = [&](int a) {
        const double2* r1 = d + ((int64_t)kz * N + a) * N;
        const double2* r2 = d + ((int64_t)kz * N + (N - a)) * N;
        const bool two = a > 0 && a < N / 2;
#pragma unroll
        for (int i = 0; i < N2P_KPT; i++) {
            const int j = threadIdx.x + i * N2P_NT;
            if (j < nh) {
                const bool mir = j > 0 && j < N / 2;
                u[i][0] = r1[j];
                if (mir) u[i][1] = r1[N - j];
                if (two) {
                    u[i][2] = r2[j];
                    if (mir) u[i][3] = r2[N - j];
                }
            }
        }
    };
    fetch(0);
    __syncthreads();
    for (int a = 0; a <= N / 2; a++) {
        const bool two = a > 0 && a < N / 2;
        double s[N2P_KPT];
#pragma unroll
        for (int i = 0; i < N2P_KPT; i++) {
            const int j = threadIdx.x + i * N2P_NT;
            const bool mir = j > 0 && j < N / 2;
            double t = u[i][0].x * u[i][0].x + u[i][0].y * u[i][0].y;
            if (mir) t += u[i][1].x * u[i][1].x + u[i][1].y * u[i][1].y;
            if (two) {
                double t2 = u[i][2].x * u[i][2].x + u[i][2].y * u[i][2].y;
                if (mir) t2 += u[i][3].x * u[i][3].x + u[i][3].y * u[i][3].y;
                t += t2;
            }
            s[i] = t;
        }
        if (a < N / 2) fetch(a + 1);
#pragma unroll
        for (int i = 0; i < N2P_KPT; i++) {
            const int j = threadIdx.x + i * N2P_NT;
            if (j < nh) pp[(int64_t)j * j + (int64_t)a * a] += s[i];
        }
        __syncthreads();
    }
}

__global__ void n2_merge_kernel(int N, const double* __restrict__ plane, double* __restrict__ hist)
{
    const int64_t mm = (int64_t)2 * (N / 2) * (N / 2) + 1;
    const int64_t nn2 = (int64_t)3 * (N / 2) * (N / 2) + 1;
    int64_t n2 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n2 >= nn2) return;
    double s = 0;
    // unrolled, every load made (a plane without this n2 reads its entry 0 and adds +0, which
    // leaves the non-negative sum unchanged), so the planes' loads are in flight together; the
    // sum keeps its order
#pragma unroll 8
    for (int kz = 0; kz < N; kz++) {
        int nz = wrapn(kz, N);
        int64_t m = n2 - (int64_t)nz * nz;
        const bool in = m >= 0 && m < mm;
        const double v = plane[(int64_t)kz * mm + (in ? m : 0)];
        s += in ? v : 0.0;
    }
    hist[n2] = s;
}

void launch_n2_hist(hipStream_t s, int N, const double2* data, double* plane_part, double* hist)
{
    if (N / 2 + 1 > N2P_KPT * N2P_NT) throw Error(AXG_EINVAL, "field spectrum: N too large for the window histogram");
    hipLaunchKernelGGL(n2_plane_kernel, dim3(N), dim3(N2P_NT), 0, s, N, data, plane_part);
    AXG_HIP(hipGetLastError());
    int64_t nn2 = (int64_t)3 * (N / 2) * (N / 2) + 1;
    hipLaunchKernelGGL(n2_merge_kernel, dim3((unsigned)((nn2 + 255) / 256)), dim3(256), 0, s, N, plane_part, hist);
    AXG_HIP(hipGetLastError());
}

// mat0[ij] = sum_{n2 in pair ij's shell} hist[n2] u_ij(q), tat0[ij] = sum cnt[n2] u_ij(q),
// q = sqrt(n2)*dwav, u_ij = get_mat_spec (spectrum.f90:385-428, 446-464).
__global__ void __launch_bounds__(RED_NT) deconv_mat_kernel(int nlq, const double* __restrict__ largeq,
                                                            const double* __restrict__ mat,
                                                            const double* __restrict__ dqs,
                                                            const double* __restrict__ hist,
                                                            const double* __restrict__ cnt, int64_t nn2, double dwav,
                                                            double* __restrict__ mat0, double* __restrict__ tat0)
{
    __shared__ double sh[RED_NT];
    const int ij = blockIdx.x;
    const double* lq = largeq + (int64_t)ij * (nlq + 1);
    const double* ms = mat + (int64_t)ij * (nlq + 1);
    double t1 = lq[nlq] / dwav, t0 = lq[0] / dwav;
    int64_t imax2 = (int64_t)ceil(t1 * t1) - 1;
    int64_t imin2 = (int64_t)ceil(t0 * t0);
    if (imax2 > nn2 - 1) imax2 = nn2 - 1;
    if (imin2 < 0) imin2 = 0;
    const double dq = dqs[ij], q0 = lq[0];
    double m = 0, t = 0;
    for (int64_t n2 = imin2 + threadIdx.x; n2 <= imax2; n2 += RED_NT) {
        double c = cnt[n2];
        if (c == 0) continue;
        double q = sqrt((double)n2) * dwav;
        double d = (q - q0) / dq;
        double fl = floor(d);
        double u = 0;
        if (fl < nlq && fl >= 0) {
            int iq = (int)fl;
            d = d - iq;
            u = (1 - d) * ms[iq] + d * ms[iq + 1];
        }
        m += hist[n2] * u;
        t += c * u;
    }
    double a = block_sum(m, sh), b = block_sum(t, sh);
    if (threadIdx.x == 0) {
        mat0[ij] = a;
        tat0[ij] = b;
    }
}

void launch_deconv_mat(hipStream_t s, int nb, int num_cov, int nlq, const double* largeq, const double* mat,
                       const double* dqs, const double* hist, const double* cnt, int64_t nn2, double dwav, double* mat0,
                       double* tat0)
{
    (void)nb;
    hipLaunchKernelGGL(deconv_mat_kernel, dim3(num_cov), dim3(RED_NT), 0, s, nlq, largeq, mat, dqs, hist, cnt, nn2, dwav,
                       mat0, tat0);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ conversions / IC helpers

template <typename T, bool COMP>
__global__ void store_field_kernel(int64_t n3, const double2* __restrict__ in, typename V2<T>::type* __restrict__ f,
                                   short2* __restrict__ cf)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        stf<T, COMP>(f, cf, i, in[i]);
}

template <typename T, bool COMP>
__global__ void load_field_kernel(int64_t n3, const typename V2<T>::type* __restrict__ f, const short2* __restrict__ cf,
                                  double2* __restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = ldf<T, COMP>(f, cf, i);
}

template <typename T, bool COMP>
void launch_store_field(hipStream_t s, int N, const double2* in, void* f, void* cf)
{
    using T2 = typename V2<T>::type;
    int64_t n3 = (int64_t)N * N * N;
    hipLaunchKernelGGL((store_field_kernel<T, COMP>), dim3(65536), dim3(256), 0, s, n3, in, (T2*)f, (short2*)cf);
    AXG_HIP(hipGetLastError());
}

template <typename T, bool COMP>
void launch_load_field(hipStream_t s, int N, const void* f, const void* cf, double2* out)
{
    using T2 = typename V2<T>::type;
    int64_t n3 = (int64_t)N * N * N;
    hipLaunchKernelGGL((load_field_kernel<T, COMP>), dim3(65536), dim3(256), 0, s, n3, (const T2*)f,
                       (const short2*)cf, out);
    AXG_HIP(hipGetLastError());
}

__global__ void scatter_modes_kernel(const int64_t* __restrict__ idx, const double2* __restrict__ val, int64_t nm,
                                     double2* __restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < nm) out[idx[i]] = val[i];
}

void launch_scatter_modes(hipStream_t s, int N, const int64_t* idx, const double2* val, int64_t nmodes, double2* out)
{
    (void)N;
    hipLaunchKernelGGL(scatter_modes_kernel, dim3((unsigned)((nmodes + 255) / 256)), dim3(256), 0, s, idx, val, nmodes,
                       out);
    AXG_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(RED_NT) sumsq_kernel(int64_t n, const double2* __restrict__ d,
                                                       double* __restrict__ partial)
{
    __shared__ double sh[RED_NT];
    double s = 0;
    for (int64_t i = blockIdx.x * (int64_t)RED_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_NT)
        s += d[i].x * d[i].x + d[i].y * d[i].y;
    double r = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

void launch_sumsq(hipStream_t s, int64_t n, const double2* data, double* partial, int nwg, double* out)
{
    hipLaunchKernelGGL(sumsq_kernel, dim3(nwg), dim3(RED_NT), 0, s, n, data, partial);
    AXG_HIP(hipGetLastError());
    sum_columns(s, partial, nwg, 1, out);
}

__global__ void scale_kernel(int64_t n, double2* __restrict__ d, double sc)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        d[i].x = d[i].x * sc;
        d[i].y = d[i].y * sc;
    }
}

void launch_scale(hipStream_t s, int64_t n, double2* data, double scale)
{
    hipLaunchKernelGGL(scale_kernel, dim3(65536), dim3(256), 0, s, n, data, scale);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ thermal initial state
// One mode per thread (initial.f90:105-135 with initial_phi_phidot 372-390): dv_wav times a
// unit complex Gaussian times the Bose-Einstein amplitude of the mode, in the reference's
// operation order.  The Gaussian of mode lin of field ip: from the reference's stream, draws
// pos + 4 N^3 (ip - 1) + 4 lin .. +3 (two Box-Muller pairs, the reference's order); otherwise
// Philox4x32-10 keyed by the seed, counter = (mode index, 'thml', ip), Box-Muller on two 53-bit
// uniforms in (0,1].  out is then transformed with the FORWARD sign.
__global__ void thermal_modes_kernel(int N, Draws dr, int ip, double tau_ini, double sigma_cr, double dwav,
                                     double dv_wav, double thrm_sigma, double2* __restrict__ out)
{
#pragma clang fp contract(off)
    const uint64_t seed = (uint64_t)dr.k0 | ((uint64_t)dr.k1 << 32);
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    for (int64_t lin = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; lin < n3;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(lin % N), y = (int)((lin / N) % N), z = (int)(lin / NN);
        const int64_t nx = x < N / 2 ? x : x - N, ny = y < N / 2 ? y : y - N, nz = z < N / 2 ? z : z - N;
        double g1, g2;
        if (dr.mcg) {
            uint64_t xs = mcg59_jump(dr.x0, dr.pos + 4 * (uint64_t)n3 * (ip - 1) + 4 * (uint64_t)lin);
            g1 = mcg59_gauss(xs);
            g2 = mcg59_gauss(xs);
        } else {
            unit_gauss(seed, 0x74686d6cu, (uint32_t)ip, (uint64_t)lin, g1, g2);
        }
        // initial_phi_phidot
        double q = (double)(nx * nx + ny * ny + nz * nz);
        q = sqrt(q) * dwav;
        double wav2 = q / tau_ini / sigma_cr;
        wav2 = wav2 * wav2;
        double thrm = thrm_sigma / tau_ini;
        double osc = sqrt((thrm * thrm - thrm_sigma * thrm_sigma) / 3 + wav2);
        double fbe = exp(-osc / thrm);
        fbe = fbe / (1 - fbe);
        double var = fbe * osc / dv_wav / sigma_cr;
        double f = sqrt(var / 2);
        if (ip == 1) f = f / (osc * sigma_cr);
        out[lin] = double2{(g1 * dv_wav) * f, (g2 * dv_wav) * f};
    }
}

void launch_thermal_modes(hipStream_t s, int N, const Draws& dr, int ip, double tau_ini, double sigma_cr, double dwav,
                          double dv_wav, double thrm_sigma, double2* out)
{
    hipLaunchKernelGGL(thermal_modes_kernel, dim3(4096), dim3(256), 0, s, N, dr, ip, tau_ini, sigma_cr, dwav, dv_wav,
                       thrm_sigma, out);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ spectrum covariance (MC)
// One Monte Carlo realisation of estimate_covmat (spectrum.f90:592-641): mode k gets a unit
// complex Gaussian times sigma[b] of its spectrum bin b, or 0 outside the bins.  The Gaussian of
// mode k in realisation iter (1-based): from the reference's stream, draws pos + 4 N^3 (iter - 1)
// + 4 k .. +3 (every mode draws, binned or not, as there); otherwise Philox4x32-10 keyed by the
// seed, counter = (mode index, 'cov ', iter).
__global__ void cov_modes_kernel(int N, Draws dr, int iter, const uint16_t* __restrict__ bin_of_n2,
                                 const double* __restrict__ sigma, double2* __restrict__ out)
{
#pragma clang fp contract(off)
    const uint64_t seed = (uint64_t)dr.k0 | ((uint64_t)dr.k1 << 32);
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    for (int64_t lin = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; lin < n3;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(lin % N), y = (int)((lin / N) % N), z = (int)(lin / NN);
        const int64_t nx = x < N / 2 ? x : x - N, ny = y < N / 2 ? y : y - N, nz = z < N / 2 ? z : z - N;
        const int b = bin_of_n2[nx * nx + ny * ny + nz * nz];
        if (b == 0) {
            out[lin] = double2{0.0, 0.0};
            continue;
        }
        double g1, g2;
        if (dr.mcg) {
            uint64_t xs = mcg59_jump(dr.x0, dr.pos + 4 * (uint64_t)n3 * (iter - 1) + 4 * (uint64_t)lin);
            g1 = mcg59_gauss(xs);
            g2 = mcg59_gauss(xs);
        } else {
            unit_gauss(seed, 0x636f7620u, (uint32_t)iter, (uint64_t)lin, g1, g2);
        }
        const double sg = sigma[b - 1];
        out[lin] = double2{g1 * sg, g2 * sg};
    }
}

// masking (spectrum.f90:900-954): the mask cells of the real-space field set to zero.
__global__ void mask_zero_kernel(int64_t n3, const int8_t* __restrict__ flags, double2* __restrict__ data)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x)
        if (flags[i] & 4) data[i] = double2{0.0, 0.0};
}

void launch_cov_modes(hipStream_t s, int N, const Draws& dr, int iter, const uint16_t* bin_of_n2, const double* sigma,
                      double2* out)
{
    hipLaunchKernelGGL(cov_modes_kernel, dim3(4096), dim3(256), 0, s, N, dr, iter, bin_of_n2, sigma, out);
    AXG_HIP(hipGetLastError());
}

void launch_mask_zero(hipStream_t s, int64_t n3, const int8_t* flags, double2* data)
{
    hipLaunchKernelGGL(mask_zero_kernel, dim3(4096), dim3(256), 0, s, n3, flags, data);
    AXG_HIP(hipGetLastError());
}

#define AXG_INST_DIAG(T, C)                                                                                      \
    template void launch_energy<T, C>(hipStream_t, int, const void*, const void*, const void*, const void*, double, \
                                      double, double, double*, int, double*);                                    \
    template void launch_classify<T, C>(hipStream_t, int, const void*, const void*, int, int, int, int8_t*,       \
                                        int64_t*, int64_t);                                                      \
    template void launch_masked_field<T, C>(hipStream_t, int, const void*, const void*, const int8_t*, double2*,  \
                                            double, bool);                                                       \
    template void launch_store_field<T, C>(hipStream_t, int, const double2*, void*, void*);                       \
    template void launch_load_field<T, C>(hipStream_t, int, const void*, const void*, double2*);
AXG_INST_DIAG(double, false)
AXG_INST_DIAG(float, false)
AXG_INST_DIAG(float, true)

}  // namespace axg
