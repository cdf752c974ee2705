// Device-side building blocks shared by the kernels: complex arithmetic on HIP vector
// types, the small DFT butterflies, and a mixed-radix Stockham FFT that runs on a batch of
// transforms resident in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };

// The lattice sizes with compile-time kernels (fft_p2.h): N = 2^L in 16..1024 and
// N = 3 * 2^L in 24..768.  Other sizes run the runtime-plan kernels.
constexpr bool p2_size(int n)
{
    const int pw = n % 3 == 0 ? n / 3 : n;
    return pw >= 8 && (pw & (pw - 1)) == 0 && n >= 16 && n <= 1024;
}

template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return {a.x + b.x, a.y + b.y}; }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return {a.x - b.x, a.y - b.y}; }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (SIGN * i)
template <int SIGN, typename T2> __device__ __forceinline__ T2 cmuli(T2 a)
{
    if constexpr (SIGN > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}
template <typename T2, typename T> __device__ __forceinline__ T2 cscale(T2 a, T s) { return {a.x * s, a.y * s}; }

// ---------------------------------------------------------------- small DFTs
// X[k] = sum_r v[r] exp(SIGN * 2 pi i r k / R), in place.

template <int SIGN, typename T2> __device__ __forceinline__ void dft2(T2* v)
{
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int SIGN, typename T2> __device__ __forceinline__ void dft4(T2* v)
{
    T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    T2 a2 = cadd(v[1], v[3]), a3 = cmuli<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <int SIGN, typename T2> __device__ __forceinline__ void dft8(T2* v)
{
    using T = decltype(v[0].x);
    T2 e[4] = {v[0], v[2], v[4], v[6]};
    T2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    const T h = (T)0.70710678118654752440084436210484904;
    // o[k] *= exp(SIGN 2 pi i k / 8)
    T2 o1 = {h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x)};
    T2 o2 = cmuli<SIGN>(o[2]);
    T2 o3 = {h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

template <int SIGN, typename T2> __device__ __forceinline__ void dft3(T2* v)
{
    using T = decltype(v[0].x);
    const T s3 = (T)0.86602540378443864676372317075293618;
    T2 t1 = cadd(v[1], v[2]);
    T2 t2 = csub(v[1], v[2]);
    T2 m = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    T2 r = cmuli<SIGN>(cscale(t2, s3));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <int R, int SIGN, typename T2> __device__ __forceinline__ void dft_small(T2* v)
{
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 3) dft3<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else dft8<SIGN>(v);
}

// ---------------------------------------------------------------- FFT plan

constexpr int AXG_MAX_STAGES = 8;
struct FftPlan {
    int n;                        // transform length
    int nst;                      // number of stages
    int radix[AXG_MAX_STAGES];    // radix of each stage (8, 4, 3 or 2)
};

// ---------------------------------------------------------------- LDS Stockham FFT
//
// M transforms of length N live in LDS; element n of transform b sits at s[n*es + b*bs].
// One stage of radix R with span Ns: butterfly j (0 <= j < N/R) reads elements
// j + r*N/R, twiddles by w^(r*(j mod Ns)) with w = exp(SIGN 2 pi i / (Ns R)), runs an
// R-point DFT and writes elements (j/Ns)*Ns*R + (j mod Ns) + r*Ns (the Stockham autosort
// ordering, so the result comes out in natural order).  tw[m] = exp(+2 pi i m / N) for
// m < N (conjugated on the fly for SIGN < 0).  All NT threads of the block take part;
// ELEMS bounds N*M, so each thread holds at most ceil(ELEMS/R/NT) butterflies.
// BATCH_FAST selects which index runs fastest across lanes (the contiguous one in LDS).

template <typename T, int R, int SIGN, int NT, int ELEMS, bool BATCH_FAST>
__device__ __forceinline__ void stockham_stage(typename V2<T>::type* s, int N, int M, int Ns, int es, int bs,
                                               const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    constexpr int MAXI = (ELEMS / R + NT - 1) / NT;
    const int nj = N / R;
    const int total = nj * M;
    const int tstride = N / (Ns * R);
    T2 v[MAXI][R];
    int dst[MAXI];
#pragma unroll
    for (int i = 0; i < MAXI; i++) {
        int it = threadIdx.x + i * NT;
        dst[i] = -1;
        if (it < total) {
            int j, b;
            if constexpr (BATCH_FAST) { j = it / M; b = it - j * M; }
            else { b = it / nj; j = it - b * nj; }
            int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = s[(j + r * nj) * es + b * bs];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    T2 w = tw[r * k * tstride];
                    if constexpr (SIGN < 0) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            dft_small<R, SIGN>(v[i]);
            dst[i] = ((j / Ns) * Ns * R + k) * es + b * bs;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXI; i++) {
        if (dst[i] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) s[dst[i] + r * Ns * es] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int SIGN, int NT, int ELEMS, bool BATCH_FAST>
__device__ __forceinline__ void fft_lds(typename V2<T>::type* s, const FftPlan& plan, int M, int es, int bs,
                                        const typename V2<T>::type* __restrict__ tw)
{
    int Ns = 1;
    for (int st = 0; st < plan.nst; st++) {
        int R = plan.radix[st];
        if (R == 8) stockham_stage<T, 8, SIGN, NT, ELEMS, BATCH_FAST>(s, plan.n, M, Ns, es, bs, tw);
        else if (R == 4) stockham_stage<T, 4, SIGN, NT, ELEMS, BATCH_FAST>(s, plan.n, M, Ns, es, bs, tw);
        else if (R == 3) stockham_stage<T, 3, SIGN, NT, ELEMS, BATCH_FAST>(s, plan.n, M, Ns, es, bs, tw);
        else stockham_stage<T, 2, SIGN, NT, ELEMS, BATCH_FAST>(s, plan.n, M, Ns, es, bs, tw);
        Ns *= R;
    }
}

// ---------------------------------------------------------------- lattice helpers

__device__ __forceinline__ int wrapn(int i, int n) { return (i < n / 2) ? i : i - n; }

// ------------------------------------------------------------------ compensated storage
// x ~= hi + lo * 2^(e-39), where hi = (float)x and hi = m 2^e with m in [0.5, 1).  The
// residual x - hi is at most half an FP32 ulp (2^(e-25)), so lo fits in 15 bits and the
// pair carries ~39 significant bits.

// Both directions are branch-free: a per-element branch here costs the callers (the kick
// passes) more in exec-mask handling than the arithmetic it would skip.  lo = 0 adds +0, which
// leaves every value but -0 unchanged (that one reads back as +0).
__device__ __forceinline__ double comp_get(float hi, short lo)
{
    int e;
    (void)frexpf(hi, &e);
    return (double)hi + ldexp((double)lo, e - 39);
}

__device__ __forceinline__ float comp_put(double x, short& lo)
{
    const float hi = (float)x;
    int e;
    (void)frexpf(hi, &e);
    double r = rint(ldexp(x - (double)hi, 39 - e));
    r = fmin(fmax(r, -32767.0), 32767.0);
    lo = (hi == 0.0f || !isfinite(hi)) ? (short)0 : (short)r;
    return hi;
}

// Streaming accesses: data a pass touches once (its outputs, pi, the transform inputs and the
// accumulators) moves with non-temporal loads and stores, which leave L2 to the re-read halo
// rows of phi.  Measured at 512^3 FP32C: fused pass 2.49 -> 2.35 ms, y pass 1.40 -> 1.34 ms,
// z pass 1.65 -> 1.62 ms, step 5.62 -> 5.38 ms (same GPU, results bit-identical).
template <typename T2>
__device__ __forceinline__ T2 ld_stream(const T2* p)
{
    if constexpr (sizeof(T2) == 8) {
        typedef float v2 __attribute__((ext_vector_type(2)));
        const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return T2{w.x, w.y};
    } else {
        typedef double v2 __attribute__((ext_vector_type(2)));
        const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return T2{w.x, w.y};
    }
}

template <typename T2>
__device__ __forceinline__ void st_stream(T2* p, T2 v)
{
    if constexpr (sizeof(T2) == 8) {
        typedef float v2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2*>(p));
    } else {
        typedef double v2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(v2{v.x, v.y}, reinterpret_cast<v2*>(p));
    }
}

template <typename T, bool COMP>
__device__ __forceinline__ double2 ldf(const typename V2<T>::type* __restrict__ f, const short2* __restrict__ cf,
                                       int64_t i)
{
    if constexpr (sizeof(T) == 8) {
        return f[i];
    } else {
        float2 v = f[i];
        if constexpr (COMP) {
            short2 c = cf[i];
            return {comp_get(v.x, c.x), comp_get(v.y, c.y)};
        } else {
            return {(double)v.x, (double)v.y};
        }
    }
}

template <typename T, bool COMP>
__device__ __forceinline__ void stf(typename V2<T>::type* __restrict__ f, short2* __restrict__ cf, int64_t i,
                                    double2 v)
{
    if constexpr (sizeof(T) == 8) {
        f[i] = v;
    } else if constexpr (COMP) {
        short2 c;
        float hx = comp_put(v.x, c.x);
        float hy = comp_put(v.y, c.y);
        f[i] = float2{hx, hy};
        cf[i] = c;
    } else {
        f[i] = float2{(float)v.x, (float)v.y};
    }
}

// Philox4x32-10 (Salmon et al., SC'11): the counter-based generator of the seeded initial
// states and of the string fits' starting points.
__device__ __forceinline__ void philox4x32(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                           uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t h0 = __umulhi(M0, c0), l0 = M0 * c0;
        uint32_t h1 = __umulhi(M1, c2), l1 = M1 * c2;
        uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0;
        c1 = l1;
        c2 = n2;
        c3 = l0;
        k0 += W0;
        k1 += W1;
    }
}

// Unit complex Gaussian (each part unit variance) number `lin` of the stream (seed, tag, word):
// Philox4x32-10 on the counter (lin, tag, word) keyed by the seed, then Box-Muller on two
// 53-bit uniforms in (0,1].
__device__ __forceinline__ void unit_gauss(uint64_t seed, uint32_t tag, uint32_t word, uint64_t lin, double& re,
                                           double& im)
{
#pragma clang fp contract(off)
    uint32_t c0 = (uint32_t)lin, c1 = (uint32_t)(lin >> 32), c2 = tag, c3 = word;
    philox4x32(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = (((uint64_t)c0 << 21) ^ (uint64_t)(c1 >> 11)) & ((1ull << 53) - 1);
    const uint64_t b = (((uint64_t)c2 << 21) ^ (uint64_t)(c3 >> 11)) & ((1ull << 53) - 1);
    const double u1 = ((double)a + 1.0) / 9007199254740992.0;
    const double u2 = ((double)b + 1.0) / 9007199254740992.0;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    re = r * cos(th);
    im = r * sin(th);
}

// ------------------------------------------------------------------ the reference's random stream
// MKL VSL's MCG59, the generator of the reference (subroutines.f90:16, 49-159): x_{n+1} = 13^13 x_n
// mod 2^59, uniform u_n = x_n / 2^59; VSL's BOXMULLER Gaussian n is sqrt(-2 ln u_{2n})
// sin(2 pi u_{2n+1}).  Position p of the stream started at x_0 is x_0 13^(13 p) mod 2^59: 2^59
// divides 2^64, so wrapping 64-bit products reduced to 59 bits at the end give it exactly.
constexpr uint64_t MCG59_A = 302875106592253ull;  // 13^13
constexpr uint64_t MCG59_MASK = (1ull << 59) - 1;

__host__ __device__ __forceinline__ uint64_t mcg59_jump(uint64_t x0, uint64_t p)
{
    uint64_t a = MCG59_A, r = 1;
    while (p) {
        if (p & 1) r *= a;
        a *= a;
        p >>= 1;
    }
    return (r * x0) & MCG59_MASK;
}

__host__ __device__ __forceinline__ uint64_t mcg59_next(uint64_t x) { return (x * MCG59_A) & MCG59_MASK; }
__host__ __device__ __forceinline__ double mcg59_u(uint64_t x) { return (double)x * 1.7347234759768071e-18; }

// Box-Muller on the next two draws of the stream at x (advanced by two).
__device__ __forceinline__ double mcg59_gauss(uint64_t& x)
{
#pragma clang fp contract(off)
    const double u1 = mcg59_u(x);
    x = mcg59_next(x);
    const double u2 = mcg59_u(x);
    x = mcg59_next(x);
    return sqrt(-2.0 * log(u1)) * sin(6.283185307179586 * u2);
}

// Where a routine's random numbers come from: the reference's MCG59 stream (mcg != 0: its start
// x0 and the position of the routine's first draw) or the build's counter-based Philox streams
// keyed by (k0, k1).
struct Draws {
    uint64_t x0, pos;
    uint32_t k0, k1;
    int mcg;
};
