// Lattice integrator kernels: the drift (evolve.f90:259-281) and the fused
// kick + GW source + drift pass (evolve.f90:283-335 with potential.f90:43-60, the
// gradient of grav_waves.f90:323-356 and the source products of grav_waves.f90:213-223).
//
// FP64 kernels keep the reference's operation order with contraction off, so one FP64 step
// is bit-identical to the reference arithmetic.  The FP32 kernels store phi and pi as FP32
// and (with COMP) an int16 residual per real that carries the bits below the FP32 value;
// all arithmetic runs in FP64 registers.
#include "axg_internal.h"
#include "fft_p2.h"

namespace axg {

// ------------------------------------------------------------------ drift
// phi_out may equal phi_in (every thread reads and then writes only its own site).

template <typename T, bool COMP>
__global__ void drift_kernel(int64_t n3, const typename V2<T>::type* phi_in, typename V2<T>::type* phi_out,
                             const typename V2<T>::type* __restrict__ pi, const short2* cphi_in, short2* cphi_out,
                             const short2* __restrict__ cpi, double dt)
{
#pragma clang fp contract(off)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        double2 f = ldf<T, COMP>(phi_in, cphi_in, i);
        double2 p = ldf<T, COMP>(pi, cpi, i);
        double2 r = {f.x + dt * p.x, f.y + dt * p.y};
        stf<T, COMP>(phi_out, cphi_out, i, r);
    }
}

template <typename T, bool COMP>
void launch_drift(hipStream_t s, int N, const void* phi_in, void* phi_out, const void* pi, const void* cphi_in,
                  void* cphi_out, const void* cpi, double dt)
{
    using T2 = typename V2<T>::type;
    int64_t n3 = (int64_t)N * N * N;
    int64_t blocks = (n3 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((drift_kernel<T, COMP>), dim3((unsigned)blocks), dim3(256), 0, s, n3, (const T2*)phi_in,
                       (T2*)phi_out, (const T2*)pi, (const short2*)cphi_in, (short2*)cphi_out, (const short2*)cpi,
                       dt);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ kick + source + drift
//
// One thread per site.  Reads phi at the site and its 6 periodic neighbours and pi at the
// site; writes pi (kicked), phi_out = phi + dt_drift * pi (the next drift, into the other
// buffer so neighbours still see the old phi) and, when gw, the 6 real source products
// S_ij = 2 Re(d_i phi conj(d_j phi)) packed as 3 complex planes
//   C0 = S11 + i S21,  C1 = S22 + i S31,  C2 = S32 + i S33.

template <typename T, bool COMP>
__global__ void kick_drift_kernel(int N, const typename V2<T>::type* __restrict__ phi_in,
                                  typename V2<T>::type* __restrict__ phi_out, typename V2<T>::type* __restrict__ pi,
                                  const short2* __restrict__ cphi_in, short2* __restrict__ cphi_out,
                                  short2* __restrict__ cpi, typename V2<T>::type* __restrict__ src, KickScalars k,
                                  double d, double dt_drift, int gw, StepRef sr)
{
#pragma clang fp contract(off)
    if (const StepEntry* se = step_entry(sr)) {
        k = se->k;
        dt_drift = se->dtd;
    }
    using T2 = typename V2<T>::type;
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    int y = blockIdx.y, z = blockIdx.z;
    int xm = x == 0 ? N - 1 : x - 1, xp = x == N - 1 ? 0 : x + 1;
    int ym = y == 0 ? N - 1 : y - 1, yp = y == N - 1 ? 0 : y + 1;
    int zm = z == 0 ? N - 1 : z - 1, zp = z == N - 1 ? 0 : z + 1;
    const int64_t NN = (int64_t)N * N;
    const int64_t i = (int64_t)z * NN + (int64_t)y * N + x;
    double2 c = ldf<T, COMP>(phi_in, cphi_in, i);
    double2 a1 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)y * N + xm);
    double2 a2 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)y * N + xp);
    double2 b1 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)ym * N + x);
    double2 b2 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)yp * N + x);
    double2 c1 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)zm * NN + (int64_t)y * N + x);
    double2 c2 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)zp * NN + (int64_t)y * N + x);
    double2 p = ldf<T, COMP>(pi, cpi, i);

    // 7-point Laplacian in the reference's summation order
    double lr = a1.x + a2.x + b1.x + b2.x + c1.x + c2.x - 6 * c.x;
    double li = a1.y + a2.y + b1.y + b2.y + c1.y + c2.y - 6 * c.y;
    // dVdphi: (2|phi|^2 + m2) phi - ma2/2  [- kt conj(phi) for the N_DW=2 tilt]
    double s2 = c.x * c.x + c.y * c.y;
    double g = 2 * s2 + k.m2;
    double dr = g * c.x - k.ma2 / 2;
    double di = g * c.y;
    if (k.kt != 0.0) {
        dr = dr - k.kt * c.x;
        di = di + k.kt * c.y;
    }
    p.x = k.f1 * p.x + k.f2 * lr + k.f3 * dr;
    p.y = k.f1 * p.y + k.f2 * li + k.f3 * di;
    stf<T, COMP>(pi, cpi, i, p);

    if (gw) {
        double2 gx = {(a2.x - a1.x) / d, (a2.y - a1.y) / d};
        double2 gy = {(b2.x - b1.x) / d, (b2.y - b1.y) / d};
        double2 gz = {(c2.x - c1.x) / d, (c2.y - c1.y) / d};
        double s11 = 2 * (gx.x * gx.x + gx.y * gx.y);
        double s21 = 2 * (gy.x * gx.x + gy.y * gx.y);
        double s22 = 2 * (gy.x * gy.x + gy.y * gy.y);
        double s31 = 2 * (gz.x * gx.x + gz.y * gx.y);
        double s32 = 2 * (gz.x * gy.x + gz.y * gy.y);
        double s33 = 2 * (gz.x * gz.x + gz.y * gz.y);
        const int64_t n3 = NN * N;
        src[i] = T2{(T)s11, (T)s21};
        src[n3 + i] = T2{(T)s22, (T)s31};
        src[2 * n3 + i] = T2{(T)s32, (T)s33};
    }

    double2 f = {c.x + dt_drift * p.x, c.y + dt_drift * p.y};
    stf<T, COMP>(phi_out, cphi_out, i, f);
}

// *counter += delta (one thread, a vector store): advances a captured graph's step counter once
// per replay.
__global__ void add_int_kernel(int* __restrict__ counter, int delta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = *counter + delta;
}

void launch_add_int(hipStream_t s, int* counter, int delta)
{
    hipLaunchKernelGGL(add_int_kernel, dim3(1), dim3(64), 0, s, counter, delta);
    AXG_HIP(hipGetLastError());
}

template <typename T, bool COMP>
void launch_kick_drift(hipStream_t s, int N, const void* phi_in, void* phi_out, void* pi, const void* cphi_in,
                       void* cphi_out, void* cpi, void* src, const KickScalars& k, double dspa, double dt_drift,
                       bool gw, const StepRef& sr)
{
    using T2 = typename V2<T>::type;
    int bx = N < 256 ? N : 256;
    dim3 grid((N + bx - 1) / bx, N, N);
    hipLaunchKernelGGL((kick_drift_kernel<T, COMP>), grid, dim3(bx), 0, s, N, (const T2*)phi_in, (T2*)phi_out,
                       (T2*)pi, (const short2*)cphi_in, (short2*)cphi_out, (short2*)cpi, (T2*)src, k, dspa * 2,
                       dt_drift, gw ? 1 : 0, sr);
    AXG_HIP(hipGetLastError());
}


// ------------------------------------------------------------------ fused pass for N = 2^L, 3 * 2^L
//
// The same kick + source + drift, with the x transform of the three packed source planes
// done in the same block (fft_p2.h): block = XR rows y at one z with one thread per site
// (XR*N = 512 threads, or N when N > 512), so the stencil phase has the memory parallelism of
// the plain kernel.  Each thread leaves its packed sources C_m = S_a + i S_b in LDS (3 x XR x N
// complex); then 3*XR*N/8 of the threads transform the 3*XR rows in place and every thread
// takes part in splitting them into the two real sources' half spectra
//   S_a^(k) = (X(k) + conj(X(N-k)))/2,  S_b^(k) = (X(k) - conj(X(N-k)))/(2i),  k = 0..N/2,
// written to W[2m] and W[2m+1] (the y and z passes follow).  The six source reals per site
// never reach HBM.

// rows per block: a power of two (it divides N) with blocks of up to 512 threads (measured at
// 512^3: one row per block beats two by 5-7 % in both precisions), one row when N >= 512
template <int N> constexpr int fused_rows() { return N >= 512 ? 1 : (512 / N < N ? p2_rows_div(N, 512 / N) : N); }
template <int N> constexpr int fused_rows_wide() { return N >= 1024 ? 1 : (1024 / N < N ? p2_rows_div(N, 1024 / N) : N); }

// One site of the fused pass: kick of p (evolve.f90:325-330 with potential.f90:43-60, in the
// reference's order), the next drift f = c + dt_drift p, and the gradient (grav_waves.f90:
// 331-351) with the packed source products of grav_waves.f90:213-223, C0 = S11 + i S21,
// C1 = S22 + i S31, C2 = S32 + i S33.  The reference divides the differences by d2 = 2 dspa.
// Multiplying by inv_d = 1/d2 gives the same bits when d2 is a power of two; for any other
// dspa the host sets div and the FP64 path divides.  FP32 storage always multiplies: its
// gradient is rounded to FP32, far coarser than the FP64 quotient's last bit.
template <typename T>
__device__ __forceinline__ void kick_site(double2 c, double2 a1, double2 a2, double2 b1, double2 b2, double2 c1v,
                                          double2 c2v, double2& p, double2& f, typename V2<T>::type (&src)[3],
                                          const KickScalars& k, double inv_d, double d2, bool div,
                                          double dt_drift)
{
#pragma clang fp contract(off)
    using T2 = typename V2<T>::type;
    double lr = a1.x + a2.x + b1.x + b2.x + c1v.x + c2v.x - 6 * c.x;
    double li = a1.y + a2.y + b1.y + b2.y + c1v.y + c2v.y - 6 * c.y;
    double s2 = c.x * c.x + c.y * c.y;
    double g = 2 * s2 + k.m2;
    double dr = g * c.x - k.ma2 / 2;
    double di = g * c.y;
    if (k.kt != 0.0) {
        dr = dr - k.kt * c.x;
        di = di + k.kt * c.y;
    }
    p.x = k.f1 * p.x + k.f2 * lr + k.f3 * dr;
    p.y = k.f1 * p.y + k.f2 * li + k.f3 * di;
    f = double2{c.x + dt_drift * p.x, c.y + dt_drift * p.y};
    if constexpr (sizeof(T) == 4) {
        // FP32 storage: the differences in FP64 (they cancel), the gradient and the products in
        // FP32, the precision the sources are stored in
        const float sx = (float)((a2.x - a1.x) * inv_d), sy = (float)((a2.y - a1.y) * inv_d);
        const float tx = (float)((b2.x - b1.x) * inv_d), ty = (float)((b2.y - b1.y) * inv_d);
        const float ux = (float)((c2v.x - c1v.x) * inv_d), uy = (float)((c2v.y - c1v.y) * inv_d);
        src[0] = T2{2 * (sx * sx + sy * sy), 2 * (tx * sx + ty * sy)};
        src[1] = T2{2 * (tx * tx + ty * ty), 2 * (ux * sx + uy * sy)};
        src[2] = T2{2 * (ux * tx + uy * ty), 2 * (ux * ux + uy * uy)};
    } else {
        double2 gx, gy, gz;
        if (div) {
            gx = double2{(a2.x - a1.x) / d2, (a2.y - a1.y) / d2};
            gy = double2{(b2.x - b1.x) / d2, (b2.y - b1.y) / d2};
            gz = double2{(c2v.x - c1v.x) / d2, (c2v.y - c1v.y) / d2};
        } else {
            gx = double2{(a2.x - a1.x) * inv_d, (a2.y - a1.y) * inv_d};
            gy = double2{(b2.x - b1.x) * inv_d, (b2.y - b1.y) * inv_d};
            gz = double2{(c2v.x - c1v.x) * inv_d, (c2v.y - c1v.y) * inv_d};
        }
        src[0] = T2{(T)(2 * (gx.x * gx.x + gx.y * gx.y)), (T)(2 * (gy.x * gx.x + gy.y * gx.y))};
        src[1] = T2{(T)(2 * (gy.x * gy.x + gy.y * gy.y)), (T)(2 * (gz.x * gx.x + gz.y * gx.y))};
        src[2] = T2{(T)(2 * (gz.x * gy.x + gz.y * gy.y)), (T)(2 * (gz.x * gz.x + gz.y * gz.y))};
    }
}

// Sub-strips per XCD strip of SG row groups of XR rows: the largest power of two dividing SG that
// keeps at least 64 rows x 512 sites per sub-strip, so the three planes of a sub-strip an XCD
// works on (about 1.2 MB in FP32C) fit its 4 MB L2 also at N = 1024.
constexpr int strip_subs(int SG, int XR, int N)
{
    int sub = 1;
    while (SG % (2 * sub) == 0 && (SG / (2 * sub)) * XR * N >= 64 * 512) sub *= 2;
    return sub;
}

// Block order of the XCD strips: block lin goes to XCD lin % 8 as its (lin / 8)-th block.  XCD k
// takes the strip of row groups [k SG, (k+1) SG), split into SUB sub-strips, and walks each
// sub-strip plane by plane before the next.
template <int N, int XR>
__device__ __forceinline__ void strip_order(int lin, int& yg, int& z)
{
    constexpr int NG = N / XR, SG = NG >= 8 ? NG / 8 : 1;  // launched only when NG % 8 == 0
    constexpr int SUB = strip_subs(SG, XR, N), SS = SG / SUB;
    const int l = lin >> 3;
    const int sub = l / (SS * N), m = l - sub * (SS * N);
    yg = (lin & 7) * SG + sub * SS + m % SS;
    z = m / SS;
}

// XCD-aware block orders: the dispatcher deals consecutive blocks to the 8 XCDs in turn.
// REMAP 1 (planes): XCD k walks a contiguous run of planes, so the blocks it holds at once are
// neighbouring rows of one plane and share their y halo rows in its L2.  REMAP 2 (strips,
// strip_order): XCD k takes a strip of NG/8 row groups and walks it plane by plane, so its
// blocks share both the y and the z halo rows.
template <typename T, bool COMP, int N, int XR, int REMAP, bool STK>
__global__ void __launch_bounds__(XR * N) kick_source_xfft_kernel(
    const typename V2<T>::type* __restrict__ phi_in, typename V2<T>::type* __restrict__ phi_out,
    typename V2<T>::type* __restrict__ pi, const short2* __restrict__ cphi_in, short2* __restrict__ cphi_out,
    short2* __restrict__ cpi, typename V2<T>::type* __restrict__ W, int P, const typename V2<T>::type* __restrict__ tw,
    KickScalars k, double inv_d, double d2, int div, double dt_drift, StepRef sr, const LatRef* __restrict__ lat)
{
#pragma clang fp contract(off)
    using T2 = typename V2<T>::type;
    if constexpr (STK) {
        // stacked launch: this block's lattice, its buffers for this step and its scalars
        const LatRef& L = lat[blockIdx.z];
        const int st = step_index(sr) & 1;
        phi_in = (const T2*)L.phi[st];
        phi_out = (T2*)L.phi[st ^ 1];
        cphi_in = (const short2*)L.cphi[st];
        cphi_out = (short2*)L.cphi[st ^ 1];
        pi = (T2*)L.pi;
        cpi = (short2*)L.cpi;
        W = (T2*)L.W;
        sr.tab = L.tab;
    }
    if (const StepEntry* se = step_entry(sr)) {
        k = se->k;
        dt_drift = se->dtd;
    }
    constexpr int TPT = N / 8, NR = 3 * XR;  // NR rows to transform
    constexpr int BS = XR * N;                 // threads per block
    constexpr int NG = N / XR;                 // row groups per plane
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);  // [m][rr][x]
    const int64_t NN = (int64_t)N * N;
    int yg = blockIdx.x, z = blockIdx.y;
    if constexpr (REMAP == 1) {
        const int lin = blockIdx.y * NG + blockIdx.x, nb = NG * N;
        const int logical = (lin & 7) * (nb >> 3) + (lin >> 3);
        yg = logical % NG;
        z = logical / NG;
    } else if constexpr (REMAP == 2) {
        strip_order<N, XR>(blockIdx.y * NG + blockIdx.x, yg, z);
    }
    {
        const int rr = threadIdx.x / N, x = threadIdx.x % N;
        const int y = yg * XR + rr;
        const int ym = ring_dn<N>(y), yp = ring_up<N>(y);
        const int zm = ring_dn<N>(z), zp = ring_up<N>(z);
        const int xm = ring_dn<N>(x), xp = ring_up<N>(x);
        const int64_t row = (int64_t)z * NN + (int64_t)y * N;
        const int64_t i = row + x;
        double2 c = ldf<T, COMP>(phi_in, cphi_in, i);
        double2 a1 = ldf<T, COMP>(phi_in, cphi_in, row + xm);
        double2 a2 = ldf<T, COMP>(phi_in, cphi_in, row + xp);
        double2 b1 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)ym * N + x);
        double2 b2 = ldf<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)yp * N + x);
        double2 c1v = ldf<T, COMP>(phi_in, cphi_in, (int64_t)zm * NN + (int64_t)y * N + x);
        double2 c2v = ldf<T, COMP>(phi_in, cphi_in, (int64_t)zp * NN + (int64_t)y * N + x);
        double2 p = ldf<T, COMP>(pi, cpi, i);
        double2 f;
        T2 src[3];
        kick_site<T>(c, a1, a2, b1, b2, c1v, c2v, p, f, src, k, inv_d, d2, div != 0, dt_drift);
        stf<T, COMP>(pi, cpi, i, p);
        stf<T, COMP>(phi_out, cphi_out, i, f);
        s[rr * N + x] = src[0];
        s[(XR + rr) * N + x] = src[1];
        s[(2 * XR + rr) * N + x] = src[2];
    }
    // FP32: twiddles of the row transforms requested before the barrier, so their latency
    // overlaps it; FP64 reads them inside the stages (fewer registers held)
    const int q = threadIdx.x / TPT, j0 = threadIdx.x % TPT;
    const bool active = q < NR;
    P2Tw<T, N> twr;
    if constexpr (sizeof(T) == 4) {
        if (active) twr = p2_load_twiddles<T, N, +1>(j0, tw);
    }
    const P2TwTable<T, N, +1> twt{tw, j0};
    __syncthreads();
    // transform the NR rows in place: row q = threadIdx / TPT
    {
        T2* rowp = s + (active ? q : 0) * N;
        T2 v[8];
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = rowp[j0 + e * TPT];
        }
        __syncthreads();
        if constexpr (sizeof(T) == 4) fft_p2_if<T, N, +1>(active, v, j0, rowp, 1, 0, twr);
        else fft_p2_if<T, N, +1>(active, v, j0, rowp, 1, 0, twt);
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) rowp[j0 + e * TPT] = v[e];
        }
        __syncthreads();
    }
    // split every transformed row into two half spectra: NR rows x (N/2+1) outputs
    constexpr int NH = N / 2 + 1, PC = half_pitch_c<N>();
    constexpr int64_t cs = (int64_t)N * N * PC;  // component stride of W
    // whole rows of pitch PC: the padding columns NH..PC-1 get zeros, so the last 128-byte line
    // of a row is written whole instead of in part
    for (int t = threadIdx.x; t < NR * PC; t += BS) {
        const int q = t / PC, kk = t - q * PC;
        const int m = q / XR, rr = q - m * XR;
        const int y = yg * XR + rr;
        T2* wa = W + (int64_t)(2 * m) * cs + ((int64_t)z * N + y) * PC;
        if (kk < NH) {
            const T2* rowp = s + q * N;
            T2 xk = rowp[kk];
            T2 xm = rowp[ring_neg<N>(kk)];
            st_stream(wa + kk, T2{(T)0.5 * (xk.x + xm.x), (T)0.5 * (xk.y - xm.y)});
            st_stream(wa + cs + kk, T2{(T)0.5 * (xk.y + xm.y), (T)0.5 * (xm.x - xk.x)});
        } else {
            st_stream(wa + kk, T2{0, 0});
            st_stream(wa + cs + kk, T2{0, 0});
        }
    }
}

// Two neighbouring sites x0 = 2t, 2t+1 per thread: phi and pi pairs move as one 16-byte
// (FP32) or 32-byte (FP64) access and one 8-byte residual access, the outer x neighbours as
// single values, and the block of XR*N/2 threads keeps 3/4 of its threads busy in the row
// transforms.  Same arithmetic per site (kick_site) as the one-site kernel.
template <typename T> struct alignas(2 * sizeof(typename V2<T>::type)) CPair {
    typename V2<T>::type a, b;
};
struct alignas(8) RPair {
    short2 a, b;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

// NT: non-temporal (streamed) accesses, for the pairs a pass touches once
template <typename T, bool COMP, bool NT = false>
__device__ __forceinline__ void ld_pair(const typename V2<T>::type* __restrict__ f, const short2* __restrict__ cf,
                                        int64_t i, double2& u, double2& w)
{
    CPair<T> v;
    if constexpr (NT && sizeof(T) == 4) {
        const f32x4 q = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(f + i));
        v.a = {q.x, q.y};
        v.b = {q.z, q.w};
    } else if constexpr (NT) {
        v.a = ld_stream(f + i);
        v.b = ld_stream(f + i + 1);
    } else {
        v = reinterpret_cast<const CPair<T>*>(f + i)[0];
    }
    if constexpr (sizeof(T) == 8) {
        u = v.a;
        w = v.b;
    } else if constexpr (COMP) {
        RPair r;
        if constexpr (NT) {
            const s16x4 q = __builtin_nontemporal_load(reinterpret_cast<const s16x4*>(cf + i));
            r.a = {q.x, q.y};
            r.b = {q.z, q.w};
        } else {
            r = reinterpret_cast<const RPair*>(cf + i)[0];
        }
        u = double2{comp_get(v.a.x, r.a.x), comp_get(v.a.y, r.a.y)};
        w = double2{comp_get(v.b.x, r.b.x), comp_get(v.b.y, r.b.y)};
    } else {
        u = double2{(double)v.a.x, (double)v.a.y};
        w = double2{(double)v.b.x, (double)v.b.y};
    }
}

template <typename T, bool COMP, bool NT = false>
__device__ __forceinline__ void st_pair(typename V2<T>::type* __restrict__ f, short2* __restrict__ cf, int64_t i,
                                        double2 u, double2 w)
{
    using T2 = typename V2<T>::type;
    CPair<T> v;
    if constexpr (sizeof(T) == 8) {
        v.a = u;
        v.b = w;
    } else if constexpr (COMP) {
        RPair r;
        v.a = T2{comp_put(u.x, r.a.x), comp_put(u.y, r.a.y)};
        v.b = T2{comp_put(w.x, r.b.x), comp_put(w.y, r.b.y)};
        if constexpr (NT)
            __builtin_nontemporal_store(s16x4{r.a.x, r.a.y, r.b.x, r.b.y}, reinterpret_cast<s16x4*>(cf + i));
        else
            reinterpret_cast<RPair*>(cf + i)[0] = r;
    } else {
        v.a = T2{(float)u.x, (float)u.y};
        v.b = T2{(float)w.x, (float)w.y};
    }
    if constexpr (NT && sizeof(T) == 4) {
        __builtin_nontemporal_store(f32x4{v.a.x, v.a.y, v.b.x, v.b.y}, reinterpret_cast<f32x4*>(f + i));
    } else if constexpr (NT) {
        st_stream(f + i, v.a);
        st_stream(f + i + 1, v.b);
    } else {
        reinterpret_cast<CPair<T>*>(f + i)[0] = v;
    }
}

template <typename T, bool COMP, int N, int XR, bool STRIP, bool STK>
__global__ void __launch_bounds__(XR * N / 2) kick_source_xfft2_kernel(
    const typename V2<T>::type* __restrict__ phi_in, typename V2<T>::type* __restrict__ phi_out,
    typename V2<T>::type* __restrict__ pi, const short2* __restrict__ cphi_in, short2* __restrict__ cphi_out,
    short2* __restrict__ cpi, typename V2<T>::type* __restrict__ W, const typename V2<T>::type* __restrict__ tw,
    KickScalars k, double inv_d, double d2, int div, double dt_drift, StepRef sr, const LatRef* __restrict__ lat)
{
#pragma clang fp contract(off)
    using T2 = typename V2<T>::type;
    if constexpr (STK) {
        // stacked launch: this block's lattice, its buffers for this step and its scalars
        const LatRef& L = lat[blockIdx.z];
        const int st = step_index(sr) & 1;
        phi_in = (const T2*)L.phi[st];
        phi_out = (T2*)L.phi[st ^ 1];
        cphi_in = (const short2*)L.cphi[st];
        cphi_out = (short2*)L.cphi[st ^ 1];
        pi = (T2*)L.pi;
        cpi = (short2*)L.cpi;
        W = (T2*)L.W;
        sr.tab = L.tab;
    }
    if (const StepEntry* se = step_entry(sr)) {
        k = se->k;
        dt_drift = se->dtd;
    }
    constexpr int TPT = N / 8, NR = 3 * XR, BS = XR * N / 2, H = N / 2;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);  // [m][rr][x]
    const int64_t NN = (int64_t)N * N;
    int yg = blockIdx.x, z = blockIdx.y;
    if constexpr (STRIP) {
        // XCD-aware strips (strip_order): the y and z neighbour rows a block reads were fetched
        // by its XCD's own recent blocks (its L2)
        strip_order<N, XR>(blockIdx.y * (N / XR) + blockIdx.x, yg, z);
    }
    {
        const int rr = threadIdx.x / H, x0 = 2 * (threadIdx.x % H);
        const int y = yg * XR + rr;
        const int ym = ring_dn<N>(y), yp = ring_up<N>(y);
        const int zm = ring_dn<N>(z), zp = ring_up<N>(z);
        const int64_t row = (int64_t)z * NN + (int64_t)y * N;
        double2 c0, c1, l, r, b10, b11, b20, b21, z10, z11, z20, z21, p0, p1;
        ld_pair<T, COMP>(phi_in, cphi_in, row + x0, c0, c1);
        l = ldf<T, COMP>(phi_in, cphi_in, row + ring_dn<N>(x0));
        r = ldf<T, COMP>(phi_in, cphi_in, row + ring_up<N, 2>(x0));
        ld_pair<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)ym * N + x0, b10, b11);
        ld_pair<T, COMP>(phi_in, cphi_in, (int64_t)z * NN + (int64_t)yp * N + x0, b20, b21);
        ld_pair<T, COMP>(phi_in, cphi_in, (int64_t)zm * NN + (int64_t)y * N + x0, z10, z11);
        ld_pair<T, COMP>(phi_in, cphi_in, (int64_t)zp * NN + (int64_t)y * N + x0, z20, z21);
        ld_pair<T, COMP, true>(pi, cpi, row + x0, p0, p1);
        double2 f0, f1;
        T2 sa[3], sb[3];
        kick_site<T>(c0, l, c1, b10, b20, z10, z20, p0, f0, sa, k, inv_d, d2, div != 0, dt_drift);
        kick_site<T>(c1, c0, r, b11, b21, z11, z21, p1, f1, sb, k, inv_d, d2, div != 0, dt_drift);
        st_pair<T, COMP, true>(pi, cpi, row + x0, p0, p1);
        st_pair<T, COMP, true>(phi_out, cphi_out, row + x0, f0, f1);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            CPair<T> v;
            v.a = sa[m];
            v.b = sb[m];
            reinterpret_cast<CPair<T>*>(s + (m * XR + rr) * N + x0)[0] = v;
        }
    }
    const int q = threadIdx.x / TPT, j0 = threadIdx.x % TPT;
    const bool active = q < NR;
    P2Tw<T, N> twr;
    if constexpr (sizeof(T) == 4) {
        if (active) twr = p2_load_twiddles<T, N, +1>(j0, tw);
    }
    const P2TwTable<T, N, +1> twt{tw, j0};
    __syncthreads();
    {
        T2* rowp = s + (active ? q : 0) * N;
        T2 v[8];
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = rowp[j0 + e * TPT];
        }
        __syncthreads();
        if constexpr (sizeof(T) == 4) fft_p2_if<T, N, +1>(active, v, j0, rowp, 1, 0, twr);
        else fft_p2_if<T, N, +1>(active, v, j0, rowp, 1, 0, twt);
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) rowp[j0 + e * TPT] = v[e];
        }
        __syncthreads();
    }
    constexpr int NH = N / 2 + 1, PC = half_pitch_c<N>();
    constexpr int64_t cs = (int64_t)N * N * PC;
    // whole rows of pitch PC: the padding columns NH..PC-1 get zeros, so the last 128-byte line
    // of a row is written whole instead of in part
    for (int t = threadIdx.x; t < NR * PC; t += BS) {
        const int qq = t / PC, kk = t - qq * PC;
        const int m = qq / XR, rr = qq - m * XR;
        const int y = yg * XR + rr;
        T2* wa = W + (int64_t)(2 * m) * cs + ((int64_t)z * N + y) * PC;
        if (kk < NH) {
            const T2* rowp = s + qq * N;
            T2 xk = rowp[kk];
            T2 xm = rowp[ring_neg<N>(kk)];
            st_stream(wa + kk, T2{(T)0.5 * (xk.x + xm.x), (T)0.5 * (xk.y - xm.y)});
            st_stream(wa + cs + kk, T2{(T)0.5 * (xk.y + xm.y), (T)0.5 * (xm.x - xk.x)});
        } else {
            st_stream(wa + kk, T2{0, 0});
            st_stream(wa + cs + kk, T2{0, 0});
        }
    }
}

template <typename T, bool COMP>
bool launch_kick_source_xfft(hipStream_t s, int N, const void* phi_in, void* phi_out, void* pi, const void* cphi_in,
                             void* cphi_out, void* cpi, void* W, int pitch, const void* tw, const KickScalars& k,
                             double dspa, double dt_drift, int variant, const StepRef& sr, const LatRef* lat,
                             int nlat)
{
    using T2 = typename V2<T>::type;
    if (pitch != half_pitch(N)) throw Error(AXG_EINVAL, "fused pass: unexpected half-spectrum pitch");
    // the gradient's divisor 2 dspa: a power of two (mantissa 1/2) lets the kernel multiply
    const double d2 = dspa * 2;
    int d2e;
    const int div = frexp(d2, &d2e) != 0.5 ? 1 : 0;
    // configuration (same results; the block shape and order only): bit 0 = wide blocks (twice
    // the rows: 1024 threads), bit 1 = XCD remap by planes (one site per thread only), bit 2 =
    // swap the sites-per-thread choice, bit 3 = XCD strips
    // a kernel with the buffers as arguments (lat == nullptr) or from the stack's LatRef table: two
    // instantiations, so that the single-lattice form addresses HBM with global instructions
    // (reached through the table, the buffers take flat ones)
    auto launch = [&]<int NN, int XR, int RM>() {
        auto go1 = [&]<bool STK>() {
            hipLaunchKernelGGL((kick_source_xfft_kernel<T, COMP, NN, XR, RM, STK>), dim3(NN / XR, NN, nlat),
                               dim3(XR * NN), (size_t)3 * XR * NN * sizeof(T2), s, (const T2*)phi_in, (T2*)phi_out,
                               (T2*)pi, (const short2*)cphi_in, (short2*)cphi_out, (short2*)cpi, (T2*)W, pitch,
                               (const T2*)tw, k, 1.0 / d2, d2, div, dt_drift, sr, lat);
        };
        if (lat) go1.template operator()<true>();
        else go1.template operator()<false>();
    };
    auto go = [&]<int NN>() {
        constexpr int XR = fused_rows<NN>(), XW = fused_rows_wide<NN>();
        if constexpr (NN >= 64) {
            // two sites per thread (blocks of XR*N/2 threads) for FP32 storage; one for FP64
            // (measured at 512^3: FP32C 2.95 vs 3.39 ms, FP64 4.97 vs 3.64 ms)
            const bool two = (sizeof(T) == 4) != ((variant & 4) != 0);
            if (two) {
                auto launch2 = [&]<int X2, bool ST>() {
                    auto go2 = [&]<bool STK>() {
                        hipLaunchKernelGGL((kick_source_xfft2_kernel<T, COMP, NN, X2, ST, STK>), dim3(NN / X2, NN, nlat),
                                           dim3(X2 * NN / 2), (size_t)3 * X2 * NN * sizeof(T2), s, (const T2*)phi_in,
                                           (T2*)phi_out, (T2*)pi, (const short2*)cphi_in, (short2*)cphi_out,
                                           (short2*)cpi, (T2*)W, (const T2*)tw, k, 1.0 / d2, d2, div, dt_drift, sr,
                                           lat);
                    };
                    if (lat) go2.template operator()<true>();
                    else go2.template operator()<false>();
                };
                // wide: twice the rows when they divide N and the tile stays within 64 KiB of LDS
                constexpr int X2W =
                    2 * XR <= NN && NN % (2 * XR) == 0 && 6 * XR * NN * (int)sizeof(T2) <= 65536 ? 2 * XR : XR;
                const bool strip = (variant & 8) != 0 && (NN / X2W) % 8 == 0;
                if (variant & 1) {
                    if (strip) launch2.template operator()<X2W, true>();
                    else launch2.template operator()<X2W, false>();
                } else {
                    if (strip && (NN / XR) % 8 == 0) launch2.template operator()<XR, true>();
                    else launch2.template operator()<XR, false>();
                }
                return;
            }
            const bool wide = (variant & 1) != 0;
            if ((variant & 8) && (NN / (wide ? XW : XR)) % 8 == 0) {
                if (wide) launch.template operator()<NN, XW, 2>();
                else launch.template operator()<NN, XR, 2>();
                return;
            }
            switch (variant & 3) {
            case 1: launch.template operator()<NN, XW, 0>(); return;
            case 2: launch.template operator()<NN, XR, 1>(); return;
            case 3: launch.template operator()<NN, XW, 1>(); return;
            default: break;
            }
        }
        launch.template operator()<NN, XR, 0>();
    };
    switch (N) {
    case 16: go.template operator()<16>(); break;
    case 32: go.template operator()<32>(); break;
    case 64: go.template operator()<64>(); break;
    case 128: go.template operator()<128>(); break;
    case 256: go.template operator()<256>(); break;
    case 512: go.template operator()<512>(); break;
    case 1024: go.template operator()<1024>(); break;
    case 24: go.template operator()<24>(); break;
    case 48: go.template operator()<48>(); break;
    case 96: go.template operator()<96>(); break;
    case 192: go.template operator()<192>(); break;
    case 384: go.template operator()<384>(); break;
    case 768: go.template operator()<768>(); break;
    default: return false;
    }
    AXG_HIP(hipGetLastError());
    return true;
}

#define AXG_INST_STEP(T, C)                                                                                  \
    template void launch_drift<T, C>(hipStream_t, int, const void*, void*, const void*, const void*, void*,   \
                                     const void*, double);                                                   \
    template void launch_kick_drift<T, C>(hipStream_t, int, const void*, void*, void*, const void*, void*,    \
                                          void*, void*, const KickScalars&, double, double, bool,        \
                                          const StepRef&);                                                \
    template bool launch_kick_source_xfft<T, C>(hipStream_t, int, const void*, void*, void*, const void*, void*, \
                                                void*, void*, int, const void*, const KickScalars&, double, double, \
                                                int, const StepRef&, const LatRef*, int);
AXG_INST_STEP(double, false)
AXG_INST_STEP(float, false)
AXG_INST_STEP(float, true)

}  // namespace axg
