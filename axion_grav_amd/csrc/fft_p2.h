// Compile-time radix-8 Stockham FFT for N = 2^L (16 <= N <= 1024), register resident.
//
// A transform is carried by TPT = N/8 threads; thread j0 holds the 8 elements
// x[j0 + e*TPT], e = 0..7 (for a batch in LDS, element n of the thread's column sits at
// s[n*es + off]).  Stages are radix 8, the last one radix 2 or 4 when L is not a multiple
// of 3.  A stage of radix R and span Ns runs 8/R butterflies per thread: butterfly
// j = j0 + m*TPT reads x[j + r*N/R] = v[m + r*(8/R)], twiddles by w^(r*(j mod Ns)) with
// w = exp(SIGN 2 pi i/(Ns R)), and writes y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  Between
// stages the outputs go through LDS and come back in the x[j0 + e*TPT] layout; after the
// last stage v[e] already holds output element j0 + e*TPT (natural order), so the first
// load and the last store go straight to and from HBM, coalesced along j0.
#pragma once
#include "axg_device.h"

template <int N> struct P2 {
    static_assert(N >= 16 && N <= 1024 && (N & (N - 1)) == 0, "P2 plans cover 16..1024, powers of two");
    static constexpr int L = N == 16 ? 4 : N == 32 ? 5 : N == 64 ? 6 : N == 128 ? 7 : N == 256 ? 8 : N == 512 ? 9 : 10;
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

// The twiddles a thread needs depend only on j0: stage S >= 1 uses M*(R-1) <= 7 of them.
// Loading them all before the first stage (next to the data loads, or before a barrier)
// takes their latency and address arithmetic out of the stages.
template <typename T, int N> struct P2Tw {
    static constexpr int NS = P2<N>::NST > 1 ? P2<N>::NST - 1 : 1;
    typename V2<T>::type w[NS][7];
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

// tws: p2_load_twiddles<T, N, SIGN>(j0, tw) of the same thread (registers), or
// P2TwTable<T, N, SIGN>{tw, j0} (read in the stages).
template <typename T, int N, int SIGN, int S = 0, typename TW>
__device__ __forceinline__ void fft_p2_if(bool active, typename V2<T>::type (&v)[8], int j0, typename V2<T>::type* s,
                                          int es, int off, const TW& tws)
{
    if constexpr (S < P2<N>::NST) {
        if (active) p2_stage<T, N, S, SIGN>(v, tws);
        if constexpr (S + 1 < P2<N>::NST) p2_exchange_if<T, N, S>(active, v, j0, s, es, off);
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
        fft_p2<T, N, SIGN, S + 1>(v, j0, s, es, off, tws);
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
