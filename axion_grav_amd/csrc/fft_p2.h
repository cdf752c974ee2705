// Compile-time radix-8 Stockham FFT for N = 2^L (16 <= N <= 1024) or N = 3 * 2^L
// (24 <= N <= 768), register resident.
//
// A transform is carried by TPT = N/8 threads; thread j0 holds the 8 elements
// x[j0 + e*TPT], e = 0..7 (for a batch in LDS, element n of the thread's column sits at
// s[n*es + off]).  The register stages cover the power-of-two factor PW (= N, or N/3): radix 8,
// the last one radix 2 or 4 when log2(PW) is not a multiple of 3.  A stage of radix R and span
// Ns runs 8/R butterflies per thread: butterfly j = j0 + m*TPT reads x[j + r*N/R] =
// v[m + r*(8/R)], twiddles by w^(r*(j mod Ns)) with w = exp(SIGN 2 pi i/(Ns R)), and writes
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  Between stages the outputs go through LDS and come back
// in the x[j0 + e*TPT] layout.  For N = 3 * 2^L a radix-3 step of span PW follows (p3_finish):
// each thread forms its own 8 outputs from the three LDS values they depend on.  Either way
// v[e] ends holding output element j0 + e*TPT (natural order), so the first load and the last
// store go straight to and from HBM, coalesced along j0.
#pragma once
#include <utility>

#include "axg_device.h"

constexpr int p2_ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

// largest power of two p <= lim that divides n (rows per block of the compile-time kernels)
constexpr int p2_rows_div(int n, int lim)
{
    int p = 1;
    while (2 * p <= lim && n % (2 * p) == 0) p *= 2;
    return p;
}

template <int N> struct P2 {
    static_assert(p2_size(N), "P2 plans cover N = 2^L in 16..1024 and N = 3 * 2^L in 24..768");
    static constexpr bool R3 = N % 3 == 0;  // a radix-3 step after the register stages
    static constexpr int PW = R3 ? N / 3 : N;
    static constexpr int L = p2_ilog2(PW);
    static constexpr int N8 = L / 3;
    static constexpr int REM = L % 3;
    static constexpr int NST = N8 + (REM ? 1 : 0);
    static constexpr int TPT = N / 8;
    static constexpr int radix(int s) { return s < N8 ? 8 : (REM == 1 ? 2 : 4); }
    static constexpr int span(int s)
    {
        int ns = 1;
        for (int i = 0; i < s; i++) ns *= radix(i);
        return ns;
    }
};

// Periodic neighbours and mirror on a ring of N sites (masks when N is a power of two).
template <int N> __device__ __forceinline__ int ring_dn(int i)
{
    if constexpr ((N & (N - 1)) == 0) return (i - 1) & (N - 1);
    else return i == 0 ? N - 1 : i - 1;
}
template <int N, int D = 1> __device__ __forceinline__ int ring_up(int i)
{
    if constexpr ((N & (N - 1)) == 0) return (i + D) & (N - 1);
    else return i + D >= N ? i + D - N : i + D;
}
template <int N> __device__ __forceinline__ int ring_neg(int i)
{
    if constexpr ((N & (N - 1)) == 0) return (N - i) & (N - 1);
    else return i == 0 ? 0 : N - i;
}

// The twiddles a thread needs depend only on j0: stage S >= 1 uses M*(R-1) <= 7 of them.
// Loading them all before the first stage (next to the data loads, or before a barrier)
// takes their latency and address arithmetic out of the stages.
template <typename T, int N> struct P2Tw {
    static constexpr int NS = P2<N>::NST > 1 ? P2<N>::NST - 1 : 1;
    typename V2<T>::type w[NS][7];
    typename V2<T>::type w3[2];  // w^j0 and w^(2 j0) of the radix-3 step (N = 3 * 2^L)
};

template <typename T, int N, int SIGN, int S = 1>
__device__ __forceinline__ void p2_twiddles(P2Tw<T, N>& t, int j0, const typename V2<T>::type* __restrict__ tw)
{
    if constexpr (S < P2<N>::NST) {
        using T2 = typename V2<T>::type;
        constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
        constexpr int TS = N / (Ns * R);
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int k = (j0 + m * TPT) & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; r++) {
                T2 w = tw[r * k * TS];
                if constexpr (SIGN < 0) w.y = -w.y;
                t.w[S - 1][m * (R - 1) + r - 1] = w;
            }
        }
        p2_twiddles<T, N, SIGN, S + 1>(t, j0, tw);
    }
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ P2Tw<T, N> p2_load_twiddles(int j0, const typename V2<T>::type* __restrict__ tw)
{
    P2Tw<T, N> t;
    p2_twiddles<T, N, SIGN>(t, j0, tw);
    if constexpr (P2<N>::R3) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            typename V2<T>::type w = tw[(i + 1) * j0];
            if constexpr (SIGN < 0) w.y = -w.y;
            t.w3[i] = w;
        }
    }
    return t;
}

// Twiddles read from the table inside each stage instead (fewer registers held: for FP64
// kernels whose register budget is already tight).
template <typename T, int N, int SIGN> struct P2TwTable {
    const typename V2<T>::type* __restrict__ tw;
    int j0;
};

template <int S, typename T, int N>
__device__ __forceinline__ typename V2<T>::type p2_tw(const P2Tw<T, N>& t, int m, int r)
{
    constexpr int R = P2<N>::radix(S);
    return t.w[S - 1][m * (R - 1) + r - 1];
}

template <int S, typename T, int N, int SIGN>
__device__ __forceinline__ typename V2<T>::type p2_tw(const P2TwTable<T, N, SIGN>& t, int m, int r)
{
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT;
    constexpr int TS = N / (Ns * R);
    const int k = (t.j0 + m * TPT) & (Ns - 1);
    typename V2<T>::type w = t.tw[r * k * TS];
    if constexpr (SIGN < 0) w.y = -w.y;
    return w;
}

template <typename T, int N>
__device__ __forceinline__ typename V2<T>::type p3_tw(const P2Tw<T, N>& t, int i)
{
    return t.w3[i];
}

template <typename T, int N, int SIGN>
__device__ __forceinline__ typename V2<T>::type p3_tw(const P2TwTable<T, N, SIGN>& t, int i)
{
    typename V2<T>::type w = t.tw[(i + 1) * t.j0];
    if constexpr (SIGN < 0) w.y = -w.y;
    return w;
}

// a * exp(SIGN 2 pi i E/8): exact swaps and negations for even E, one (1 +- i)/sqrt(2)
// product for odd E
template <int SIGN, int E, typename T2> __device__ __forceinline__ T2 rot8(T2 a)
{
    constexpr int e = E & 7;
    if constexpr (e & 1) {
        using T = decltype(a.x);
        const T h = (T)0.70710678118654752440084436210485;
        const T2 b = SIGN > 0 ? T2{(a.x - a.y) * h, (a.y + a.x) * h} : T2{(a.x + a.y) * h, (a.y - a.x) * h};
        return rot8<SIGN, e - 1>(b);
    } else if constexpr (e == 0) {
        return a;
    } else if constexpr (e == 4) {
        return T2{-a.x, -a.y};
    } else if constexpr ((e == 2) == (SIGN > 0)) {
        return T2{-a.y, a.x};  // times i
    } else {
        return T2{a.y, -a.x};  // times -i
    }
}

template <typename T, int N, int S, int SIGN, typename TW>
__device__ __forceinline__ void p2_stage(typename V2<T>::type (&v)[8], const TW& tws)
{
    using T2 = typename V2<T>::type;
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), M = 8 / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        T2 u[R];
#pragma unroll
        for (int r = 0; r < R; r++) u[r] = v[m + r * M];
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) u[r] = cmul(u[r], p2_tw<S>(tws, m, r));
        }
        dft_small<R, SIGN>(u);
#pragma unroll
        for (int r = 0; r < R; r++) v[m + r * M] = u[r];
    }
}

// Outputs of stage S to LDS in Stockham order, back as the next stage's inputs.
template <typename T, int N, int S>
__device__ __forceinline__ void p2_exchange(typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s, int es, int off)
{
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int j = j0 + m * TPT;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;  // (j/Ns)*Ns*R + (j mod Ns)
#pragma unroll
        for (int r = 0; r < R; r++) s[(base + r * Ns) * es + off] = v[m + r * M];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = s[(j0 + e * TPT) * es + off];
    __syncthreads();
}

// Variant for blocks where only some threads carry a transform: inactive threads skip the
// arithmetic and LDS traffic but take part in every barrier.
template <typename T, int N, int S>
__device__ __forceinline__ void p2_exchange_if(bool active, typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s,
                                               int es, int off)
{
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
    if (active) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int j = j0 + m * TPT;
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) s[(base + r * Ns) * es + off] = v[m + r * M];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = s[(j0 + e * TPT) * es + off];
    }
    __syncthreads();
}

// The radix-3 step of N = 3 * 2^L after the last register stage (span PW = N/3).  Its
// butterfly j < PW maps y[j + r PW] to X[j + r' PW] = sum_r y[j + r PW] w^(r j) w3^(r r')
// with w = exp(SIGN 2 pi i/N), w3 = w^PW, i.e. X[n] = sum_r y[(n mod PW) + r PW] w^(r n).  The
// thread's outputs n = j0 + e TPT (TPT = N/8) need w^n = w^j0 exp(SIGN 2 pi i e/8) and
// w^(2n) = w^(2 j0) exp(SIGN 2 pi i 2e/8), so two table twiddles serve all eight.
template <typename T, int N, int SIGN, typename TW>
__device__ __forceinline__ void p3_finish(bool active, typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s,
                                          int es, int off, const TW& tws)
{
    using T2 = typename V2<T>::type;
    constexpr int S = P2<N>::NST - 1, R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
    constexpr int PW = P2<N>::PW;
    if (active) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int j = j0 + m * TPT;
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) s[(base + r * Ns) * es + off] = v[m + r * M];
        }
    }
    __syncthreads();
    if (active) {
        const T2 a = p3_tw(tws, 0), b = p3_tw(tws, 1);
        [&]<int... E>(std::integer_sequence<int, E...>) {
            (
                [&] {
                    const int jj = (j0 + E * TPT) & (PW - 1);
                    const T2 y0 = s[jj * es + off], y1 = s[(jj + PW) * es + off], y2 = s[(jj + 2 * PW) * es + off];
                    const T2 u = cmul(y1, rot8<SIGN, E>(a)), w = cmul(y2, rot8<SIGN, 2 * E>(b));
                    v[E] = T2{y0.x + u.x + w.x, y0.y + u.y + w.y};
                }(),
                ...);
        }(std::make_integer_sequence<int, 8>{});
    }
    __syncthreads();
}

// tws: p2_load_twiddles<T, N, SIGN>(j0, tw) of the same thread (registers), or
// P2TwTable<T, N, SIGN>{tw, j0} (read in the stages).
template <typename T, int N, int SIGN, int S = 0, typename TW>
__device__ __forceinline__ void fft_p2_if(bool active, typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s,
                                          int es, int off, const TW& tws)
{
    if constexpr (S < P2<N>::NST) {
        if (active) p2_stage<T, N, S, SIGN>(v, tws);
        if constexpr (S + 1 < P2<N>::NST) p2_exchange_if<T, N, S>(active, v, j0, s, es, off);
        else if constexpr (P2<N>::R3) p3_finish<T, N, SIGN>(active, v, j0, s, es, off, tws);
        fft_p2_if<T, N, SIGN, S + 1>(active, v, j0, s, es, off, tws);
    }
}

template <typename T, int N, int SIGN, int S = 0, typename TW>
__device__ __forceinline__ void fft_p2(typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s, int es, int off,
                                       const TW& tws)
{
    if constexpr (S < P2<N>::NST) {
        p2_stage<T, N, S, SIGN>(v, tws);
        if constexpr (S + 1 < P2<N>::NST) p2_exchange<T, N, S>(v, j0, s, es, off);
        else if constexpr (P2<N>::R3) p3_finish<T, N, SIGN>(true, v, j0, s, es, off, tws);
        fft_p2<T, N, SIGN, S + 1>(v, j0, s, es, off, tws);
    }
}

// A workgroup barrier for LDS traffic only: this wave's LDS operations complete, then s_barrier.
// __syncthreads() also waits (vmcnt 0) for every direct global-to-LDS copy in flight, since such a
// copy writes LDS; a transform that runs while its block copies its next input into another LDS
// tile synchronises with this barrier instead (fft_p2_rb), which leaves the copies in flight.
__device__ __forceinline__ void lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <typename T, int N, int S>
__device__ __forceinline__ void p2_exchange_rb(typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s, int es, int off)
{
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int j = j0 + m * TPT;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) s[(base + r * Ns) * es + off] = v[m + r * M];
    }
    lds_sync();
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = s[(j0 + e * TPT) * es + off];
    lds_sync();
}

// fft_p2 with lds_sync barriers (2^k sizes).
template <typename T, int N, int SIGN, int S = 0, typename TW>
__device__ __forceinline__ void fft_p2_rb(typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s, int es, int off,
                                          const TW& tws)
{
    if constexpr (S < P2<N>::NST) {
        p2_stage<T, N, S, SIGN>(v, tws);
        if constexpr (S + 1 < P2<N>::NST) p2_exchange_rb<T, N, S>(v, j0, s, es, off);
        fft_p2_rb<T, N, SIGN, S + 1>(v, j0, s, es, off, tws);
    }
}

// Two transforms per thread, the LDS tile's columns off0 and off1, through the same stages: the
// exchanges of both share one pair of barriers.
template <typename T, int N, int S>
__device__ __forceinline__ void p2_exchange2(typename V2<T>::type (&v0)[8], typename V2<T>::type (&v1)[8], int j0,
                                             typename V2<T>::type* s, int es, int off0, int off1)
{
    constexpr int R = P2<N>::radix(S), Ns = P2<N>::span(S), TPT = P2<N>::TPT, M = 8 / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int j = j0 + m * TPT;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            s[(base + r * Ns) * es + off0] = v0[m + r * M];
            s[(base + r * Ns) * es + off1] = v1[m + r * M];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; e++) {
        v0[e] = s[(j0 + e * TPT) * es + off0];
        v1[e] = s[(j0 + e * TPT) * es + off1];
    }
    __syncthreads();
}

template <typename T, int N, int SIGN, int S = 0, typename TW>
__device__ __forceinline__ void fft_p2_2(typename V2<T>::type (&v0)[8], typename V2<T>::type (&v1)[8], int j0,
                                         typename V2<T>::type* s, int es, int off0, int off1, const TW& tws)
{
    if constexpr (S < P2<N>::NST) {
        p2_stage<T, N, S, SIGN>(v0, tws);
        p2_stage<T, N, S, SIGN>(v1, tws);
        if constexpr (S + 1 < P2<N>::NST) {
            p2_exchange2<T, N, S>(v0, v1, j0, s, es, off0, off1);
        } else if constexpr (P2<N>::R3) {
            p3_finish<T, N, SIGN>(true, v0, j0, s, es, off0, tws);
            p3_finish<T, N, SIGN>(true, v1, j0, s, es, off1, tws);
        }
        fft_p2_2<T, N, SIGN, S + 1>(v0, v1, j0, s, es, off0, off1, tws);
    }
}

// Column-tile width of the specialized passes: 128 bytes of complex per row, bounded so the
// block (TX * N/8 threads) stays within 1024 threads and the tile within 64 KiB of LDS.
template <typename T, int N> constexpr int p2_tx()
{
    constexpr int bytes = (int)sizeof(typename V2<T>::type);
    int tx = 128 / bytes;
    while (tx > 1 && (tx * (N / 8) > 1024 || tx * N * bytes > 65536)) tx /= 2;
    return tx;
}

// Runs F.template operator()<N>() for N = 2^L in 16..1024 or 3 * 2^L in 24..768 (the sizes of
// fft_p2.h); returns false otherwise.
template <typename F> bool p2_dispatch(int N, F&& f)
{
    switch (N) {
    case 24: f.template operator()<24>(); return true;
    case 48: f.template operator()<48>(); return true;
    case 96: f.template operator()<96>(); return true;
    case 192: f.template operator()<192>(); return true;
    case 384: f.template operator()<384>(); return true;
    case 768: f.template operator()<768>(); return true;
    case 16: f.template operator()<16>(); return true;
    case 32: f.template operator()<32>(); return true;
    case 64: f.template operator()<64>(); return true;
    case 128: f.template operator()<128>(); return true;
    case 256: f.template operator()<256>(); return true;
    case 512: f.template operator()<512>(); return true;
    case 1024: f.template operator()<1024>(); return true;
    default: return false;
    }
}
