// Adjoint of the leapfrog + GW step for the gradient of a band of the GW spectrum
// J = sum over bins b in the band of fact q_b^3 / denom_b * numer_b, numer_b the sum of |A_m(k)|^2
// over the bin's modes (grav_spec_estimation, grav_waves.f90:257-305).
//
// Conventions: a complex field z = a + ib has the adjoint zbar = dJ/da + i dJ/db, so the adjoint
// of multiplication by a complex c is multiplication by conj(c), and of a real-linear map its
// transpose.  The GW accumulators are linear in the six real sources S_c(x) = 2 Re(g_i conj g_j)
// (g = central-difference gradient of phi) through D_c = sum_x S_c(x) e^{+ikx} on the half
// spectrum kx <= N/2 and the per-component update of grav_waves.f90:228-242.  The adjoint of
// one kick's update, for the accumulator adjoints abar of the final state:
//   pdbar = c abar11 + s (abar21 [+ abar12 with the index quirk]),  ndbar = c abar12 + s abar22,
//   Dbar  = conj(P) pdbar + P ndbar,  P = pol dw,  c, s = cos, sin(q tau);
// under the quirk (A(2,1) <- A(1,2) + ..., A(1,2) <- new A(2,1) + ...) the adjoint state after
// the last update going backward becomes abar21 = 0, abar12 = abar21 + abar12, constant
// thereafter.  Sbar_c(x) = Re sum_half Dbar_c(k) e^{-ikx} = sum_full H_c(k) e^{-ikx} with the
// Hermitian H_c(k) = (Dbar_c(k)[k in half] + conj Dbar_c(-k)[-k in half]) / 2, so the pair
// H_a + i H_b transforms to Sbar_a + i Sbar_b in one complex FFT.
#include "axg_internal.h"
#include "fft_p2.h"

namespace axg {

namespace {

__device__ __forceinline__ int wrapi(int i, int n) { return i < n / 2 ? i : i - n; }

// pol_c = p_i p_j (halved when i == j) of grav_waves.f90:157-191 for the six (i, j) in the
// order (1,1) (2,1) (2,2) (3,1) (3,2) (3,3); q = |n| dwav.  pol = 0 at k = 0.
__device__ void pol_all(int nx, int ny, int nz, double dwav, double2 pol[6], double& q)
{
#pragma clang fp contract(off)
    double qh[3] = {(double)nx, (double)ny, (double)nz};
    const double qn = sqrt(qh[0] * qh[0] + qh[1] * qh[1] + qh[2] * qh[2]);
    q = qn * dwav;
    if (!(qn > 0)) {
#pragma unroll
        for (int c = 0; c < 6; c++) pol[c] = double2{0, 0};
        return;
    }
    qh[0] = qh[0] / qn;
    qh[1] = qh[1] / qn;
    qh[2] = qh[2] / qn;
    const double r = sqrt(qh[0] * qh[0] + qh[1] * qh[1]);
    double cc = 1, ss = 0;
    if (r > 0) {
        cc = qh[0] / r;
        ss = qh[1] / r;
    }
    const double2 p[3] = {{qh[2] * cc, -ss}, {qh[2] * ss, cc}, {-r, 0}};
    const int ii[6] = {0, 1, 1, 2, 2, 2}, jj[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const double2 a = p[ii[c]], b = p[jj[c]];
        double2 v = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
        if (ii[c] == jj[c]) v = double2{v.x / 2, v.y / 2};
        pol[c] = v;
    }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ double2 cconj(double2 a) { return {a.x, -a.y}; }

// pol_all in FP32 (for FP32 adjoint sources): the geometry to FP32 precision; q in FP64.
__device__ void pol_all_f(int nx, int ny, int nz, double dwav, double2 pol[6], double& q)
{
#pragma clang fp contract(off)
    const int n2 = nx * nx + ny * ny + nz * nz;
    q = sqrt((double)n2) * dwav;
    if (n2 == 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) pol[c] = double2{0, 0};
        return;
    }
    const float inv = rsqrtf((float)n2);
    const float hx = nx * inv, hy = ny * inv, hz = nz * inv;
    const int rr2 = nx * nx + ny * ny;
    float cc = 1, ss = 0, r = 0;
    if (rr2 > 0) {
        const float ir = rsqrtf((float)rr2);
        cc = nx * ir;
        ss = ny * ir;
        r = sqrtf((float)rr2) * inv;
    }
    const float2 p[3] = {{hz * cc, -ss}, {hz * ss, cc}, {-r, 0}};
    const int ii[6] = {0, 1, 1, 2, 2, 2}, jj[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const float2 a = p[ii[c]], b = p[jj[c]];
        float2 v = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
        if (ii[c] == jj[c]) v = float2{v.x * 0.5f, v.y * 0.5f};
        pol[c] = double2{v.x, v.y};
    }
}

// Dbar_c at one stored mode (half index hk) for all six components; FAST: the mode geometry
// and the phase's sine and cosine in FP32 (the phase itself reduced in FP64).
template <bool FAST>
__device__ void dbar_all(int N, int kx, int ky, int kz, const double2* __restrict__ abar, int64_t ms, int64_t hk,
                         const AdjStepScalars& g, double2 out[6])
{
#pragma clang fp contract(off)
    double2 pol[6];
    double q;
    double s, c;
    if constexpr (FAST) {
        pol_all_f(wrapi(kx, N), wrapi(ky, N), wrapi(kz, N), g.dwav, pol, q);
        double ph = q * g.tau;
        ph = ph - 6.283185307179586 * rint(ph * 0.15915494309189535);
        float sf, cf;
        sincosf((float)ph, &sf, &cf);
        s = sf;
        c = cf;
    } else {
        pol_all(wrapi(kx, N), wrapi(ky, N), wrapi(kz, N), g.dwav, pol, q);
        const double ph = q * g.tau;
        sincos(ph, &s, &c);
    }
    const double2 a11 = abar[hk], a21 = abar[ms + hk], a12 = abar[2 * ms + hk], a22 = abar[3 * ms + hk];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double2 pdb, ndb;
        if (g.fix_swap) {
            pdb = {c * a11.x + s * a21.x, c * a11.y + s * a21.y};
            ndb = {c * a12.x + s * a22.x, c * a12.y + s * a22.y};
        } else if (g.last && k == 5) {
            // the last update of the run: the final adjoints themselves
            pdb = {c * a11.x + s * (a21.x + a12.x), c * a11.y + s * (a21.y + a12.y)};
            ndb = {c * a12.x + s * a22.x, c * a12.y + s * a22.y};
        } else {
            // every earlier update: abar21 = 0, abar12 = abar21 + abar12 of the final state
            const double2 m12 = {a21.x + a12.x, a21.y + a12.y};
            pdb = {c * a11.x + s * m12.x, c * a11.y + s * m12.y};
            ndb = {c * m12.x + s * a22.x, c * m12.y + s * a22.y};
        }
        const double2 P = {pol[k].x * g.dw, pol[k].y * g.dw};
        const double2 u = cmul(cconj(P), pdb), v = cmul(P, ndb);
        out[k] = double2{u.x + v.x, u.y + v.y};
    }
}

template <typename TW>
__device__ __forceinline__ void store_pairs(typename V2<TW>::type* __restrict__ W0, typename V2<TW>::type* __restrict__ W1,
                                            typename V2<TW>::type* __restrict__ W2, int64_t lin, const double2 H[6])
{
    using W2T = typename V2<TW>::type;
    W0[lin] = W2T{(TW)(H[0].x - H[1].y), (TW)(H[0].y + H[1].x)};
    W1[lin] = W2T{(TW)(H[2].x - H[3].y), (TW)(H[2].y + H[3].x)};
    W2[lin] = W2T{(TW)(H[4].x - H[5].y), (TW)(H[4].y + H[5].x)};
}

// W_p(k) = H_{2p}(k) + i H_{2p+1}(k) on the full grid (x fastest), p = 0, 1, 2, one thread per
// stored mode (kz, ky, kx <= N/2): for 0 < kx < N/2 the mode fills H(k) = Dbar(k)/2 and its
// mirror H(-k) = conj Dbar(k)/2; on the planes kx = 0, N/2 (closed under k -> -k) each mode fills
// its own H(k) = (Dbar(k) + conj Dbar(-k))/2.
template <typename TW>
__global__ void adj_modes_kernel(int N, const double2* __restrict__ abar, AdjStepScalars g,
                                 typename V2<TW>::type* __restrict__ W0, typename V2<TW>::type* __restrict__ W1,
                                 typename V2<TW>::type* __restrict__ W2)
{
#pragma clang fp contract(off)
    const int nh = N / 2 + 1;
    const int64_t NN = (int64_t)N * N, ms = NN * nh;
    for (int64_t hk = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; hk < ms;
         hk += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(hk % nh);
        const int64_t r = hk / nh;
        const int y = (int)(r % N), z = (int)(r / N);
        const int ym = (N - y) % N, zm = (N - z) % N;
        double2 d[6], H[6];
        dbar_all<sizeof(TW) == 4>(N, x, y, z, abar, ms, hk, g, d);
        if (x > 0 && x < N / 2) {
#pragma unroll
            for (int c = 0; c < 6; c++) H[c] = double2{d[c].x / 2, d[c].y / 2};
            store_pairs<TW>(W0, W1, W2, ((int64_t)z * N + y) * N + x, H);
#pragma unroll
            for (int c = 0; c < 6; c++) H[c] = double2{d[c].x / 2, -d[c].y / 2};
            store_pairs<TW>(W0, W1, W2, ((int64_t)zm * N + ym) * N + (N - x), H);
        } else {
            double2 e[6];
            dbar_all<sizeof(TW) == 4>(N, x, ym, zm, abar, ms, ((int64_t)zm * N + ym) * nh + x, g, e);
#pragma unroll
            for (int c = 0; c < 6; c++) H[c] = double2{(d[c].x + e[c].x) / 2, (d[c].y - e[c].y) / 2};
            store_pairs<TW>(W0, W1, W2, ((int64_t)z * N + y) * N + x, H);
        }
    }
}

// abar_m(k) = 2 w_b mult(kx) A_m(k) in place on the packed FP64 copies [m][kz][ky][kx <= N/2]:
// d|A|^2 = 2 Re(conj A dA), summed over the full spectrum, whose modes -k mirror the stored
// ones with 0 < kx < N/2 (mult 2) and are stored themselves on the planes kx = 0, N/2 (mult 1).
__global__ void adj_seed_kernel(int N, double2* __restrict__ abar, const uint16_t* __restrict__ bin_of_n2,
                                const double* __restrict__ w)
{
    const int nh = N / 2 + 1;
    const int64_t tot = (int64_t)N * N * nh;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % nh);
        const int64_t r = i / nh;
        const int ky = (int)(r % N), kz = (int)(r / N);
        const int64_t nx = wrapi(kx, N), ny = wrapi(ky, N), nz = wrapi(kz, N);
        const int64_t n2 = nx * nx + ny * ny + nz * nz;
        const int b = n2 > 0 ? bin_of_n2[n2] : 0;
        const double f = b > 0 ? 2 * w[b - 1] * ((kx == 0 || kx == N / 2) ? 1.0 : 2.0) : 0.0;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const double2 a = abar[m * tot + i];
            abar[m * tot + i] = double2{a.x * f, a.y * f};
        }
    }
}

template <typename T, bool COMP>
__device__ __forceinline__ double2 ldphi(const typename V2<T>::type* phi, const short2* cphi, int64_t i)
{
    return ldf<T, COMP>(phi, cphi, i);
}

// gbar_i(x) = sum_j dS_ij/dg_i Sbar_ij: 4 Sbar_ii g_i + 2 sum_{j != i} Sbar_ij g_j, written over
// W (W0 <- gbar_x, W1 <- gbar_y, W2 <- gbar_z); W holds Sbar as (S11 + i S21, S22 + i S31,
// S32 + i S33).
template <typename T, bool COMP, typename TW>
__global__ void adj_gbar_kernel(int N, const typename V2<T>::type* __restrict__ phi, const short2* __restrict__ cphi,
                                double inv_d, typename V2<TW>::type* __restrict__ W0, typename V2<TW>::type* __restrict__ W1,
                                typename V2<TW>::type* __restrict__ W2)
{
#pragma clang fp contract(off)
    using W2T = typename V2<TW>::type;
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
        const int64_t base = i - x;
        const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
        const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
        const int zp = z + 1 == N ? 0 : z + 1, zm = z == 0 ? N - 1 : z - 1;
        const double2 a1 = ldphi<T, COMP>(phi, cphi, base + xm), a2 = ldphi<T, COMP>(phi, cphi, base + xp);
        const double2 b1 = ldphi<T, COMP>(phi, cphi, (int64_t)z * NN + (int64_t)ym * N + x);
        const double2 b2 = ldphi<T, COMP>(phi, cphi, (int64_t)z * NN + (int64_t)yp * N + x);
        const double2 c1 = ldphi<T, COMP>(phi, cphi, (int64_t)zm * NN + (int64_t)y * N + x);
        const double2 c2 = ldphi<T, COMP>(phi, cphi, (int64_t)zp * NN + (int64_t)y * N + x);
        const double2 g1 = {(a2.x - a1.x) * inv_d, (a2.y - a1.y) * inv_d};
        const double2 g2 = {(b2.x - b1.x) * inv_d, (b2.y - b1.y) * inv_d};
        const double2 g3 = {(c2.x - c1.x) * inv_d, (c2.y - c1.y) * inv_d};
        const W2T w0 = W0[i], w1 = W1[i], w2 = W2[i];
        const double s11 = w0.x, s21 = w0.y, s22 = w1.x, s31 = w1.y, s32 = w2.x, s33 = w2.y;
        W0[i] = W2T{(TW)(4 * s11 * g1.x + 2 * s21 * g2.x + 2 * s31 * g3.x),
                    (TW)(4 * s11 * g1.y + 2 * s21 * g2.y + 2 * s31 * g3.y)};
        W1[i] = W2T{(TW)(2 * s21 * g1.x + 4 * s22 * g2.x + 2 * s32 * g3.x),
                    (TW)(2 * s21 * g1.y + 4 * s22 * g2.y + 2 * s32 * g3.y)};
        W2[i] = W2T{(TW)(2 * s31 * g1.x + 2 * s32 * g2.x + 4 * s33 * g3.x),
                    (TW)(2 * s31 * g1.y + 2 * s32 * g2.y + 4 * s33 * g3.y)};
    }
}

// The same gbar with the last (x) transform of Sbar done in the block: block = XR rows y at one
// z, one site per thread.  3*XR*N/8 of the threads transform the rows of W0..W2 (already
// transformed along y and z) in place in LDS, then every thread forms gbar at its site from
// the three LDS rows and writes it over W.  Saves the separate row pass (a read and a write of
// the three arrays).
template <int N> constexpr int adj_rows() { return N >= 512 ? 1 : p2_rows_div(N, 512 / N); }

template <typename T, bool COMP, typename TW, int N>
__global__ void __launch_bounds__(adj_rows<N>() * N) adj_gbar_xfft_kernel(
    const typename V2<T>::type* __restrict__ phi, const short2* __restrict__ cphi, double inv_d,
    typename V2<TW>::type* __restrict__ W0, typename V2<TW>::type* __restrict__ W1,
    typename V2<TW>::type* __restrict__ W2, const typename V2<TW>::type* __restrict__ tw)
{
#pragma clang fp contract(off)
    using W2T = typename V2<TW>::type;
    constexpr int XR = adj_rows<N>(), TPT = N / 8, NR = 3 * XR;
    extern __shared__ unsigned char smem_raw[];
    W2T* s = reinterpret_cast<W2T*>(smem_raw);  // [m][rr][x]
    const int64_t NN = (int64_t)N * N;
    const int y0 = blockIdx.x * XR, z = blockIdx.y;
    {
        const int q = threadIdx.x / TPT, j0 = threadIdx.x % TPT;
        const bool active = q < NR;
        const int m = active ? q / XR : 0, rr = active ? q - m * XR : 0;
        const W2T* src = (m == 0 ? W0 : m == 1 ? W1 : W2) + (int64_t)z * NN + (int64_t)(y0 + rr) * N;
        W2T v[8];
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = src[j0 + e * TPT];
        }
        W2T* rowp = s + (active ? q : 0) * N;
        fft_p2_if<TW, N, -1>(active, v, j0, rowp, 1, 0, P2TwTable<TW, N, -1>{tw, j0});
        if (active) {
#pragma unroll
            for (int e = 0; e < 8; e++) rowp[j0 + e * TPT] = v[e];
        }
        __syncthreads();
    }
    const int rr = threadIdx.x / N, x = threadIdx.x - rr * N;
    const int y = y0 + rr;
    const int64_t i = (int64_t)z * NN + (int64_t)y * N + x;
    const int64_t base = i - x;
    const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
    const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
    const int zp = z + 1 == N ? 0 : z + 1, zm = z == 0 ? N - 1 : z - 1;
    const double2 a1 = ldphi<T, COMP>(phi, cphi, base + xm), a2 = ldphi<T, COMP>(phi, cphi, base + xp);
    const double2 b1 = ldphi<T, COMP>(phi, cphi, (int64_t)z * NN + (int64_t)ym * N + x);
    const double2 b2 = ldphi<T, COMP>(phi, cphi, (int64_t)z * NN + (int64_t)yp * N + x);
    const double2 c1 = ldphi<T, COMP>(phi, cphi, (int64_t)zm * NN + (int64_t)y * N + x);
    const double2 c2 = ldphi<T, COMP>(phi, cphi, (int64_t)zp * NN + (int64_t)y * N + x);
    const double2 g1 = {(a2.x - a1.x) * inv_d, (a2.y - a1.y) * inv_d};
    const double2 g2 = {(b2.x - b1.x) * inv_d, (b2.y - b1.y) * inv_d};
    const double2 g3 = {(c2.x - c1.x) * inv_d, (c2.y - c1.y) * inv_d};
    const W2T w0 = s[rr * N + x], w1 = s[(XR + rr) * N + x], w2 = s[(2 * XR + rr) * N + x];
    const double s11 = w0.x, s21 = w0.y, s22 = w1.x, s31 = w1.y, s32 = w2.x, s33 = w2.y;
    W0[i] = W2T{(TW)(4 * s11 * g1.x + 2 * s21 * g2.x + 2 * s31 * g3.x),
                (TW)(4 * s11 * g1.y + 2 * s21 * g2.y + 2 * s31 * g3.y)};
    W1[i] = W2T{(TW)(2 * s21 * g1.x + 4 * s22 * g2.x + 2 * s32 * g3.x),
                (TW)(2 * s21 * g1.y + 4 * s22 * g2.y + 2 * s32 * g3.y)};
    W2[i] = W2T{(TW)(2 * s31 * g1.x + 2 * s32 * g2.x + 4 * s33 * g3.x),
                (TW)(2 * s31 * g1.y + 2 * s32 * g2.y + 4 * s33 * g3.y)};
}

// Adjoint of one kick (and of the source's central differences):
//   phibar(y) += inv_d sum_i (gbar_i(y - e_i) - gbar_i(y + e_i)) + f2 Lap(pibar)(y) + f3 H(phi(y)) pibar(y),
// H the symmetric Jacobian of dVdphi as a map of (Re, Im).  Block partial sums of the parameter
// derivatives: [0] = sum <pibar, V'_0(phi)> (V' without the tilt), [1] = sum <pibar, (-a, b)>.
template <typename T, bool COMP, typename TW>
__global__ void __launch_bounds__(256) adj_kick_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                       const short2* __restrict__ cphi,
                                                       const double2* __restrict__ pibar, double2* __restrict__ phibar,
                                                       double2* __restrict__ pibar_out,
                                                       const typename V2<TW>::type* __restrict__ G0,
                                                       const typename V2<TW>::type* __restrict__ G1,
                                                       const typename V2<TW>::type* __restrict__ G2, KickScalars k,
                                                       double inv_d, double hd, double s0, double s1, int use_g,
                                                       double* __restrict__ partial)
{
#pragma clang fp contract(off)
    using W2T = typename V2<TW>::type;
    __shared__ double sh[2][256];
    const int64_t NN = (int64_t)N * N, n3 = NN * N;
    double acc0 = 0, acc1 = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / NN);
        const int64_t base = i - x;
        const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
        const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
        const int zp = z + 1 == N ? 0 : z + 1, zm = z == 0 ? N - 1 : z - 1;
        const int64_t ixm = base + xm, ixp = base + xp;
        const int64_t iym = (int64_t)z * NN + (int64_t)ym * N + x, iyp = (int64_t)z * NN + (int64_t)yp * N + x;
        const int64_t izm = (int64_t)zm * NN + (int64_t)y * N + x, izp = (int64_t)zp * NN + (int64_t)y * N + x;
        const double2 pb = pibar[i];
        const double2 pa1 = pibar[ixm], pa2 = pibar[ixp], pb1 = pibar[iym], pb2 = pibar[iyp], pc1 = pibar[izm],
                      pc2 = pibar[izp];
        const double lr = pa1.x + pa2.x + pb1.x + pb2.x + pc1.x + pc2.x - 6 * pb.x;
        const double li = pa1.y + pa2.y + pb1.y + pb2.y + pc1.y + pc2.y - 6 * pb.y;
        const double2 c = ldphi<T, COMP>(phi, cphi, i);
        const double a = c.x, b = c.y;
        const double s2 = a * a + b * b;
        const double g = 2 * s2 + k.m2;
        const double hrr = g + 4 * (a * a) - k.kt, hii = g + 4 * (b * b) + k.kt, hri = 4 * (a * b);
        double fr = phibar[i].x + k.f2 * lr + k.f3 * (hrr * pb.x + hri * pb.y);
        double fi = phibar[i].y + k.f2 * li + k.f3 * (hri * pb.x + hii * pb.y);
        if (use_g) {
            const W2T gx1 = G0[ixm], gx2 = G0[ixp], gy1 = G1[iym], gy2 = G1[iyp], gz1 = G2[izm], gz2 = G2[izp];
            fr = fr + inv_d * (((double)gx1.x - gx2.x) + ((double)gy1.x - gy2.x) + ((double)gz1.x - gz2.x));
            fi = fi + inv_d * (((double)gx1.y - gx2.y) + ((double)gy1.y - gy2.y) + ((double)gz1.y - gz2.y));
        }
        phibar[i] = double2{fr, fi};
        // the kick scales pibar by f1; the drift before it adds hd phibar
        pibar_out[i] = double2{k.f1 * pb.x + hd * fr, k.f1 * pb.y + hd * fi};
        const double vr = g * a - k.ma2 / 2, vi = g * b;
        acc0 = acc0 + (pb.x * vr + pb.y * vi);
        acc1 = acc1 + (pb.y * b - pb.x * a);
    }
    sh[0][threadIdx.x] = acc0;
    sh[1][threadIdx.x] = acc1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    // each block adds its scaled sums to its own slot: a fixed order, step after step
    if (threadIdx.x < 2) {
        const int64_t j = (int64_t)blockIdx.x * 2 + threadIdx.x;
        partial[j] = partial[j] + (threadIdx.x == 0 ? s0 : s1) * sh[threadIdx.x][0];
    }
}

// Block partial sums of <phibar, phi> + <pibar, pi>.
template <typename T, bool COMP>
__global__ void __launch_bounds__(256) adj_dot_kernel(int64_t n3, const typename V2<T>::type* __restrict__ phi,
                                                      const short2* __restrict__ cphi,
                                                      const typename V2<T>::type* __restrict__ pi,
                                                      const short2* __restrict__ cpi, const double2* __restrict__ phibar,
                                                      const double2* __restrict__ pibar, double* __restrict__ partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[256];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 f = ldphi<T, COMP>(phi, cphi, i), p = ldphi<T, COMP>(pi, cpi, i);
        const double2 fb = phibar[i], pb = pibar[i];
        acc = acc + (fb.x * f.x + fb.y * f.y) + (pb.x * p.x + pb.y * p.y);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

}  // namespace

constexpr int ADJ_GRID = ADJ_GRID_H;

void launch_adj_seed(hipStream_t s, int N, double2* abar, const uint16_t* bin_of_n2, const double* w)
{
    hipLaunchKernelGGL(adj_seed_kernel, dim3(ADJ_GRID), dim3(256), 0, s, N, abar, bin_of_n2, w);
    AXG_HIP(hipGetLastError());
}

template <typename TW>
void launch_adj_modes(hipStream_t s, int N, const double2* abar, const AdjStepScalars& g, void* W0, void* W1, void* W2)
{
    using W2T = typename V2<TW>::type;
    hipLaunchKernelGGL((adj_modes_kernel<TW>), dim3(ADJ_GRID), dim3(256), 0, s, N, abar, g, (W2T*)W0, (W2T*)W1,
                       (W2T*)W2);
    AXG_HIP(hipGetLastError());
}
template void launch_adj_modes<float>(hipStream_t, int, const double2*, const AdjStepScalars&, void*, void*, void*);
template void launch_adj_modes<double>(hipStream_t, int, const double2*, const AdjStepScalars&, void*, void*, void*);

template <typename T, bool COMP, typename TW>
void launch_adj_gbar(hipStream_t s, int N, const void* phi, const void* cphi, double inv_d, void* W0, void* W1,
                     void* W2)
{
    using T2 = typename V2<T>::type;
    using W2T = typename V2<TW>::type;
    hipLaunchKernelGGL((adj_gbar_kernel<T, COMP, TW>), dim3(ADJ_GRID), dim3(256), 0, s, N, (const T2*)phi,
                       (const short2*)cphi, inv_d, (W2T*)W0, (W2T*)W1, (W2T*)W2);
    AXG_HIP(hipGetLastError());
}

template <typename T, bool COMP, typename TW>
bool launch_adj_gbar_xfft(hipStream_t s, int N, const void* phi, const void* cphi, double inv_d, void* W0, void* W1,
                          void* W2, const void* tw)
{
    using T2 = typename V2<T>::type;
    using W2T = typename V2<TW>::type;
    const bool done = p2_dispatch(N, [&]<int NN>() {
        constexpr int XR = adj_rows<NN>();
        hipLaunchKernelGGL((adj_gbar_xfft_kernel<T, COMP, TW, NN>), dim3(NN / XR, NN), dim3(XR * NN),
                           (size_t)3 * XR * NN * sizeof(W2T), s, (const T2*)phi, (const short2*)cphi, inv_d,
                           (W2T*)W0, (W2T*)W1, (W2T*)W2, (const W2T*)tw);
    });
    if (done) AXG_HIP(hipGetLastError());
    return done;
}

template <typename T, bool COMP, typename TW>
void launch_adj_kick(hipStream_t s, int N, const void* phi, const void* cphi, const double2* pibar, double2* phibar,
                     double2* pibar_out, const void* G0, const void* G1, const void* G2, const KickScalars& k,
                     double inv_d, double hd, double s0, double s1, bool use_g, double* partial)
{
    using T2 = typename V2<T>::type;
    using W2T = typename V2<TW>::type;
    hipLaunchKernelGGL((adj_kick_kernel<T, COMP, TW>), dim3(ADJ_GRID), dim3(256), 0, s, N, (const T2*)phi,
                       (const short2*)cphi, pibar, phibar, pibar_out, (const W2T*)G0, (const W2T*)G1, (const W2T*)G2, k,
                       inv_d, hd, s0, s1, (int)use_g, partial);
    AXG_HIP(hipGetLastError());
}

template <typename T, bool COMP>
void launch_adj_dot(hipStream_t s, int64_t n3, const void* phi, const void* cphi, const void* pi, const void* cpi,
                    const double2* phibar, const double2* pibar, double* partial, double* out)
{
    using T2 = typename V2<T>::type;
    hipLaunchKernelGGL((adj_dot_kernel<T, COMP>), dim3(ADJ_GRID), dim3(256), 0, s, n3, (const T2*)phi,
                       (const short2*)cphi, (const T2*)pi, (const short2*)cpi, phibar, pibar, partial);
    AXG_HIP(hipGetLastError());
    sum_columns(s, partial, ADJ_GRID, 1, out);
}

#define AXG_INST_ADJ(T, C, TW)                                                                                       \
    template void launch_adj_gbar<T, C, TW>(hipStream_t, int, const void*, const void*, double, void*, void*, void*); \
    template bool launch_adj_gbar_xfft<T, C, TW>(hipStream_t, int, const void*, const void*, double, void*, void*,   \
                                                 void*, const void*);                                             \
    template void launch_adj_kick<T, C, TW>(hipStream_t, int, const void*, const void*, const double2*, double2*,     \
                                            double2*, const void*, const void*, const void*, const KickScalars&,      \
                                            double, double, double, double, bool, double*);                           \
    template void launch_adj_dot<T, C>(hipStream_t, int64_t, const void*, const void*, const void*, const void*,     \
                                       const double2*, const double2*, double*, double*);
AXG_INST_ADJ(double, false, double)
AXG_INST_ADJ(float, false, float)
AXG_INST_ADJ(float, true, float)

}  // namespace axg
