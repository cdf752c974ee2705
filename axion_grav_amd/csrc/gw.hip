// GW sourcing kernels: the 3D FFT of the six source products (grav_waves.f90:225, an
// unnormalized BACKWARD transform) as three passes over HBM, and the helicity projection +
// Green's-function accumulation of grav_waves.f90:227-247 fused into the last pass, plus
// the shell binning of grav_spec_estimation (grav_waves.f90:271-295).
//
// Layout.  The sources are real, so the transform is kept on the half spectrum
// kx = 0..N/2.  W[c][z][y][kx] (c = 0..5 in the reference's (i,j) order (1,1),(2,1),(2,2),
// (3,1),(3,2),(3,3); row pitch P = half_pitch(N)) holds the x-transformed sources.  The y
// transform writes them out of place into W2 in the tiled GwLayout [c][ky][kx/8][kz][kx%8]
// (the sizes of fft_p2.h; in place in W with rows [c][kz][ky][kx] otherwise), and the accumulation pass
// z-transforms them on the fly.  The accumulators A[m][...] (m = arr_grav (1,1),(2,1),(1,2),
// (2,2)) use the same GwLayout, so that pass reads and writes contiguous runs.
#include "axg_internal.h"
#include "fft_p2.h"

namespace axg {

constexpr int FFT_NT = 256;
// complex elements per block held in LDS (64 KiB): 4096 FP64, 8192 FP32
template <typename T> constexpr int fft_elems() { return 65536 / (int)sizeof(typename V2<T>::type); }
// Column-tile width: 128 bytes of complex per row, fewer when N columns would not fit.
template <typename T> int col_width(int N)
{
    int tx = col_tile<T>();
    while (tx > 1 && N * tx > fft_elems<T>()) tx /= 2;
    return N * tx <= fft_elems<T>() ? tx : 0;
}


// ================================================================== N = 2^L, 3 * 2^L specializations
// The same three passes with the compile-time radix-8 engine of fft_p2.h: every thread
// carries 8 elements of one transform in registers, the first load and last store go
// directly between registers and HBM, and LDS is touched only between stages.

// The accumulation pass keeps 8 k-points' projected sums per thread.  With at most 768 threads
// per block the compiler has 170 VGPRs for the pipelined FP32 form (it needs about 155).  FP32
// from N = 512 instead takes tiles of up to 1024 threads (16 columns at 512, 8 at 1024; 128
// VGPRs, no pipelining), so the y pass writes 128-byte (512) or 64-byte (1024) tile rows
// instead of half of that.  Measured: y pass 1.80 -> 1.55 ms at 512^3 and 35 -> 15 ms at
// 1024^3, z pass within 3 %.
template <typename T, int N> constexpr bool p2_zwide() { return sizeof(T) == 4 && N >= 512; }
template <typename T, int N> constexpr int p2_ztx()
{
    int tx = p2_tx<T, N>();
    while (tx > 1 && tx * (N / 8) > (p2_zwide<T, N>() ? 1024 : 768)) tx /= 2;
    return tx;
}

template <int N> constexpr int p2_xrows() { return (2048 / N) < N ? p2_rows_div(N, 2048 / N < 1 ? 1 : 2048 / N) : N; }

template <typename T, int N>
__global__ void __launch_bounds__(p2_xrows<N>() * (N / 8)) gw_xpass_p2_kernel(const typename V2<T>::type* __restrict__ src,
                                                                           typename V2<T>::type* __restrict__ W, int P,
                                                                           const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    constexpr int TPT = N / 8, XR = p2_xrows<N>();
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int rr = threadIdx.x / TPT, j0 = threadIdx.x % TPT;
    const int y = blockIdx.x * XR + rr, z = blockIdx.y, m = blockIdx.z;
    const int64_t NN = (int64_t)N * N;
    const T2* in = src + (int64_t)m * NN * N + (int64_t)z * NN + (int64_t)y * N;
    T2 v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = in[j0 + e * TPT];
    const P2Tw<T, N> tws = p2_load_twiddles<T, N, +1>(j0, tw);
    fft_p2<T, N, +1>(v, j0, s, 1, rr * N, tws);
    T2* row = s + rr * N;
#pragma unroll
    for (int e = 0; e < 8; e++) row[j0 + e * TPT] = v[e];
    __syncthreads();
    T2* wa = W + ((int64_t)(2 * m) * NN + (int64_t)z * N + y) * P;
    T2* wb = W + ((int64_t)(2 * m + 1) * NN + (int64_t)z * N + y) * P;
#pragma unroll
    for (int e = 0; e < 5; e++) {
        const int k = j0 + e * TPT;
        if (k <= N / 2) {
            T2 xk = row[k];
            T2 xm = row[ring_neg<N>(k)];
            wa[k] = T2{(T)0.5 * (xk.x + xm.x), (T)0.5 * (xk.y - xm.y)};
            wb[k] = T2{(T)0.5 * (xk.y + xm.y), (T)0.5 * (xm.x - xk.x)};
        }
    }
}

template <typename T, int N, int SIGN>
__global__ void __launch_bounds__((p2_tx<T, N>() * (N / 8))) fft_cols_p2_kernel(typename V2<T>::type* __restrict__ data,
                                                                           int64_t es, int64_t ps, int ncol,
                                                                           const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    constexpr int TX = p2_tx<T, N>(), TPT = N / 8;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int b = threadIdx.x % TX, j0 = threadIdx.x / TX;
    const int col = blockIdx.x * TX + b;
    const bool ok = col < ncol;
    T2* base = data + (int64_t)blockIdx.y * ps + col;
    T2 v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = ok ? base[(int64_t)(j0 + e * TPT) * es] : T2{0, 0};
    const P2Tw<T, N> tws = p2_load_twiddles<T, N, SIGN>(j0, tw);
    fft_p2<T, N, SIGN>(v, j0, s, TX, b, tws);
    if (ok) {
#pragma unroll
        for (int e = 0; e < 8; e++) base[(int64_t)(j0 + e * TPT) * es] = v[e];
    }
}

// y transform of the x-transformed sources W[c][z][y][kx] (pitch P) into W2 in the tiled
// layout of the accumulation pass.  Block = TX kx columns of KZ consecutive (c, z) planes: the
// planes are transformed one after the other through the same LDS tile, and their outputs for
// one (ky, kx) are adjacent tile rows, stored back to back (KZ = 2 turns the 64-byte rows of
// 8-column FP32 or 4-column FP64 tiles into whole 128-byte lines).
template <typename T, int N, int TX, int KZ, bool STK>
__global__ void __launch_bounds__((TX * (N / 8))) gw_ypass_p2_kernel(const typename V2<T>::type* __restrict__ W,
                                                                    typename V2<T>::type* __restrict__ W2,
                                                                    const typename V2<T>::type* __restrict__ tw,
                                                                    const LatRef* __restrict__ lat)
{
    using T2 = typename V2<T>::type;
    if constexpr (STK) {
        // stacked launch: this block's lattice in blockIdx.z
        W = (const T2*)lat[blockIdx.z].W;
        W2 = (T2*)lat[blockIdx.z].W2;
    }
    constexpr int TPT = N / 8, NH = N / 2 + 1, P = half_pitch_c<N>();
    using L = TiledLayout<N, p2_ztx<T, N>()>;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int b = threadIdx.x % TX, j0 = threadIdx.x / TX;
    const int kx = blockIdx.x * TX + b;
    const int c = blockIdx.y / (N / KZ), z0 = (blockIdx.y % (N / KZ)) * KZ;
    const bool ok = kx < NH;
    T2 v[KZ][8];
#pragma unroll
    for (int q = 0; q < KZ; q++) {
        const T2* base = W + ((int64_t)c * N + z0 + q) * (int64_t)N * P + kx;
#pragma unroll
        for (int e = 0; e < 8; e++) v[q][e] = ok ? ld_stream(base + (int64_t)(j0 + e * TPT) * P) : T2{0, 0};
    }
    // twiddles read inside the stages (measured: no faster with them held in registers)
#pragma unroll
    for (int q = 0; q < KZ; q++) {
        if (q > 0) __syncthreads();
        fft_p2<T, N, +1>(v[q], j0, s, TX, b, P2TwTable<T, N, +1>{tw, j0});
    }
    // the padding columns of the last tile (NH <= kx < NT * ZT) get zeros, so its rows are
    // written whole; a block wider than the tile (TX > ZT) also covers columns past the last
    // tile, which do not exist
    if (kx < L::NT * L::ZTW) {
#pragma unroll
        for (int e = 0; e < 8; e++)
#pragma unroll
            for (int q = 0; q < KZ; q++) W2[L::idx(c, z0 + q, j0 + e * TPT, kx)] = ok ? v[q][e] : T2{0, 0};
    }
}


// Column-tile width of the y pass with two adjacent kx columns per thread: a whole 128-byte line
// of complex per row, for the sizes whose one-column tile is narrower (the 1024-thread block
// limit at N = 768 and 1024 in FP32, the 64 KiB tile of the one-column kernels in FP64); 0 where
// it does not apply.  Its LDS tile, TX * N complex (up to 128 KiB), fits gfx950's 160 KiB per
// workgroup.
template <typename T, int N> constexpr int p2_tx2()
{
    constexpr int bytes = (int)sizeof(typename V2<T>::type), tx = 128 / bytes;
    return (p2_tx<T, N>() < tx && (tx / 2) * (N / 8) <= 1024 && tx * N * bytes <= 160 * 1024) ? tx : 0;
}

// The y pass of gw_ypass_p2_kernel with two adjacent columns kx, kx + 1 per thread: W is read in
// whole 128-byte lines (16-byte FP32 / 32-byte FP64 accesses per thread), and each thread writes
// its column pair as one access into the tiled W2 (pairs never straddle a tile: kx is even).
// The two columns' transforms share the stages and barriers (fft_p2_2).
template <typename T, int N, int TX, int KZ, bool STK>
__global__ void __launch_bounds__((TX / 2) * (N / 8)) gw_ypass_p2w_kernel(const typename V2<T>::type* __restrict__ W,
                                                                         typename V2<T>::type* __restrict__ W2,
                                                                         const typename V2<T>::type* __restrict__ tw,
                                                                         const LatRef* __restrict__ lat)
{
    using T2 = typename V2<T>::type;
    if constexpr (STK) {
        W = (const T2*)lat[blockIdx.z].W;
        W2 = (T2*)lat[blockIdx.z].W2;
    }
    constexpr int TPT = N / 8, NH = N / 2 + 1, P = half_pitch_c<N>(), H = TX / 2;
    using L = TiledLayout<N, p2_ztx<T, N>()>;
    static_assert(L::ZTW % 2 == 0 && P % 2 == 0, "column pairs stay within a tile and a 16-byte line");
    __shared__ __align__(16) unsigned char smem[TX * N * sizeof(T2)];
    T2* s = reinterpret_cast<T2*>(smem);
    const int b = threadIdx.x % H, j0 = threadIdx.x / H;
    const int kx = blockIdx.x * TX + 2 * b;
    const int c = blockIdx.y / (N / KZ), z0 = (blockIdx.y % (N / KZ)) * KZ;
    T2 v0[KZ][8], v1[KZ][8];
#pragma unroll
    for (int q = 0; q < KZ; q++) {
        const T2* base = W + ((int64_t)c * N + z0 + q) * (int64_t)N * P + kx;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const T2* r = base + (int64_t)(j0 + e * TPT) * P;
            if (kx + 1 < NH) {
                if constexpr (sizeof(T) == 4) {
                    typedef float f32x4 __attribute__((ext_vector_type(4)));
                    const f32x4 u = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(r));
                    v0[q][e] = T2{u.x, u.y};
                    v1[q][e] = T2{u.z, u.w};
                } else {
                    v0[q][e] = ld_stream(r);
                    v1[q][e] = ld_stream(r + 1);
                }
            } else {
                v0[q][e] = kx < NH ? ld_stream(r) : T2{0, 0};
                v1[q][e] = T2{0, 0};
            }
        }
    }
#pragma unroll
    for (int q = 0; q < KZ; q++) {
        if (q > 0) __syncthreads();
        fft_p2_2<T, N, +1>(v0[q], v1[q], j0, s, TX, 2 * b, 2 * b + 1, P2TwTable<T, N, +1>{tw, j0});
    }
    // the padding columns of the last tile get zeros, as in gw_ypass_p2_kernel; columns past the
    // last tile do not exist
    if (kx < L::NT * L::ZTW) {
#pragma unroll
        for (int e = 0; e < 8; e++)
#pragma unroll
            for (int q = 0; q < KZ; q++) {
                T2* o = W2 + L::idx(c, z0 + q, j0 + e * TPT, kx);
                const T2 a = kx < NH ? v0[q][e] : T2{0, 0}, bb = kx + 1 < NH ? v1[q][e] : T2{0, 0};
                if constexpr (sizeof(T) == 4) {
                    typedef float f32x4 __attribute__((ext_vector_type(4)));
                    *reinterpret_cast<f32x4*>(o) = f32x4{a.x, a.y, bb.x, bb.y};
                } else {
                    o[0] = a;
                    o[1] = bb;
                }
            }
    }
}

template <typename T, int N>
__global__ void __launch_bounds__((p2_ztx<T, N>() * (N / 8))) gw_zpass_p2_kernel(const typename V2<T>::type* __restrict__ W,
                                                                           typename V2<T>::type* __restrict__ A,
                                                                           const typename V2<T>::type* __restrict__ tw,
                                                                           GwStepScalars g, StepRef sr,
                                                                           const LatRef* __restrict__ lat)
{
#pragma clang fp contract(off)
    if (lat) {
        // stacked launch: this block's lattice in blockIdx.z
        W = (const typename V2<T>::type*)lat[blockIdx.z].W2;
        A = (typename V2<T>::type*)lat[blockIdx.z].A;
        sr.tab = lat[blockIdx.z].tab;
    }
    if (const StepEntry* se = step_entry(sr)) {
        g.tau = se->tau;
        g.dw = se->dw;
    }
    using L = TiledLayout<N, p2_ztx<T, N>()>;
    using T2 = typename V2<T>::type;
    constexpr int TX = p2_ztx<T, N>(), TPT = N / 8, NH = N / 2 + 1;
    __shared__ __attribute__((aligned(16))) T2 s[N * TX];
    const int b = threadIdx.x % TX, j0 = threadIdx.x / TX;
    // the block is one tile (L.zt == TX columns) of one ky: contiguous N*TX runs
    const int kx = blockIdx.x * TX + b, ky = blockIdx.y;
    const bool ok = kx < NH;
    const int nx = wrapn(kx, N), ny = wrapn(ky, N);
    const int64_t tile0 = L::idx(0, 0, ky, blockIdx.x * TX) + b;
    constexpr int64_t cst = L::COMP;
    constexpr int zs = TX;

    // polarization geometry of grav_waves.f90:165-183, p1 = (qz c, -s), p2 = (qz s, c),
    // p3 = (-r, 0) with qhat = n/|n|, r = |qhat_xy|, c = qhat_x/r = nx/rho, s = ny/rho
    // (rho = |n_xy|; c = 1, s = 0 on the kz axis): c and s are the same for the thread's 8
    // wavevectors, only qz and r vary with kz.
    T gc, gs;
    {
        const T rho = sqrt((T)nx * (T)nx + (T)ny * (T)ny);
        if (rho > 0) { gc = (T)nx / rho; gs = (T)ny / rho; }
        else { gc = 1; gs = 0; }
    }
    T2 sp[8], sn[8];
#pragma unroll
    for (int e = 0; e < 8; e++) sp[e] = sn[e] = T2{0, 0};
    // per wavevector: qz = nz/|n| and r = |n_xy|/|n| (one reciprocal; k = 0 is zeroed after the
    // component loop); held for the 8 k-points, or recomputed per component in the 1024-thread
    // form, whose 128 VGPRs cannot hold them
    const T rho2g = (T)nx * (T)nx + (T)ny * (T)ny, rhog = sqrt(rho2g);
    auto geom = [&](int e, T& z, T& r) {
        const int nz = wrapn(j0 + e * TPT, N);
        const T q = sqrt(rho2g + (T)nz * (T)nz);
        const T iq = q > 0 ? (T)1 / q : (T)0;
        z = (T)nz * iq;
        r = rhog * iq;
    };
    constexpr int NG8 = p2_zwide<T, N>() ? 1 : 8;
    T gz[NG8], gr[NG8];
    if constexpr (!p2_zwide<T, N>()) {
#pragma unroll
        for (int e = 0; e < 8; e++) geom(e, gz[e], gr[e]);
    }

    // FP32: twiddles in registers and component c+1 loaded while component c is transformed
    // (software pipelining); FP64 has no registers to spare for either.
    constexpr bool PIPE = sizeof(T) == 4 && !p2_zwide<T, N>(), PTW = PIPE;
    T2 v[8], nxt[8];
    // FP32 wide tiles (1024 threads, no registers left to load the next component ahead): each
    // component's N x TX tile (64 KiB, contiguous in the tiled layout) is copied from HBM into a
    // second LDS tile by the waves' direct global-to-LDS loads (16 B per lane, 1 KiB per wave and
    // instruction) while the previous component is transformed.  The transform then synchronises
    // with lds_sync and reads its twiddles from LDS, so that nothing waits for the copies before
    // the next component's turn.  512^3 FP32C: 1.62-1.63 -> 1.53-1.54 ms, results bit-identical.
    constexpr bool DMA = p2_zwide<T, N>();
    __shared__ __attribute__((aligned(16))) T2 pf[DMA ? N * TX : 1];
    const int64_t tilebase = L::idx(0, 0, ky, blockIdx.x * TX);
    auto dma = [&](int c) {
        constexpr int NW = DMA ? TX * (N / 8) / 64 : 1;
        constexpr int NI = DMA ? (int)(N * TX * sizeof(T2)) / (NW * 1024) : 0;
        static_assert(!DMA || NI * NW * 1024 == (int)(N * TX * sizeof(T2)), "whole wave chunks");
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
        const char* g = reinterpret_cast<const char*>(W + c * cst + tilebase);
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int chunk = i * NW + w;
            __builtin_amdgcn_global_load_lds(
                (const void __attribute__((address_space(1)))*)(g + (int64_t)chunk * 1024 + lane * 16),
                (void __attribute__((address_space(3)))*)((__attribute__((address_space(3))) char*)
                    (__attribute__((address_space(3))) T2*)pf + chunk * 1024), 16, 0, 0);
        }
    };
    // the transform's twiddles in LDS: a twiddle load from global memory would wait (vmcnt, in
    // issue order) for the copies in flight
    __shared__ T2 stw[DMA ? N : 1];
    if constexpr (DMA) {
        for (int i = threadIdx.x; i < N; i += TX * (N / 8)) stw[i] = tw[i];
        dma(0);
    }
    const P2TwTable<T, N, +1> twl{stw, j0};
    if constexpr (PIPE) {
#pragma unroll
        for (int e = 0; e < 8; e++) nxt[e] = ok ? W[tile0 + (int64_t)(j0 + e * TPT) * zs] : T2{0, 0};
    }
    P2Tw<T, N> twr;
    if constexpr (PTW) twr = p2_load_twiddles<T, N, +1>(j0, tw);
    const P2TwTable<T, N, +1> twt{tw, j0};
#pragma unroll 1
    for (int c = 0; c < 6; c++) {
        if constexpr (PIPE) {
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = nxt[e];
            if (c + 1 < 6) {
                const T2* base = W + (c + 1) * cst + tile0;
#pragma unroll
                for (int e = 0; e < 8; e++) nxt[e] = ok ? ld_stream(base + (int64_t)(j0 + e * TPT) * zs) : T2{0, 0};
            }
        } else if constexpr (DMA) {
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies have landed
            __syncthreads();                      // and every wave's; the last transform is done with s
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = ok ? pf[(j0 + e * TPT) * TX + b] : T2{0, 0};
            __syncthreads();                      // pf read by all: the next copy may start
            if (c + 1 < 6) dma(c + 1);
        } else {
            const T2* base = W + c * cst + tile0;
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = ok ? ld_stream(base + (int64_t)(j0 + e * TPT) * zs) : T2{0, 0};
        }
        if (c > 0 && !DMA) __syncthreads();
        if constexpr (DMA && !P2<N>::R3) fft_p2_rb<T, N, +1>(v, j0, s, TX, b, twl);
        else if constexpr (PTW) fft_p2<T, N, +1>(v, j0, s, TX, b, twr);
        else fft_p2<T, N, +1>(v, j0, s, TX, b, twt);
        {
            // pol_ij d into sp (pol d) and sn (conj(pol) d), (i,j) in the reference's order
            // (1,1),(2,1),(2,2),(3,1),(3,2),(3,3) (grav_waves.f90:213-247)
            const int ii = c == 0 ? 0 : (c <= 2 ? 1 : 2);
            const int jj = (c == 0 || c == 1 || c == 3) ? 0 : ((c == 2 || c == 4) ? 1 : 2);
#pragma unroll
            for (int e = 0; e < 8; e++) {
                T ge, re;
                if constexpr (p2_zwide<T, N>()) geom(e, ge, re);
                else { ge = gz[e]; re = gr[e]; }
                const T2 p1 = {ge * gc, -gs}, p2 = {ge * gs, gc}, p3 = {-re, 0};
                const T2 a = ii == 0 ? p1 : (ii == 1 ? p2 : p3);
                const T2 bb = jj == 0 ? p1 : (jj == 1 ? p2 : p3);
                T2 pol = {a.x * bb.x - a.y * bb.y, a.x * bb.y + a.y * bb.x};
                if (ii == jj) {
                    pol.x = pol.x / 2;
                    pol.y = pol.y / 2;
                }
                const T2 d = v[e];
                sp[e].x += pol.x * d.x - pol.y * d.y;
                sp[e].y += pol.x * d.y + pol.y * d.x;
                sn[e].x += pol.x * d.x + pol.y * d.y;
                sn[e].y += pol.x * d.y - pol.y * d.x;
            }
        }
    }
    const int64_t ms = cst;
    if (!ok) {
        // padding columns of the last tile: FP64 writes them zeros (what they hold from creation),
        // so the tile's rows are written whole (measured at 512^3: 3.02 -> 2.92 ms); in FP32 the
        // extra bytes cost more than the partial rows (1.64 -> 1.68 ms), so it skips them
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int64_t ai = tile0 + (int64_t)(j0 + e * TPT) * zs;
#pragma unroll
                for (int m = 0; m < 4; m++) A[m * ms + ai] = T2{0, 0};
            }
        }
        return;
    }
    if (kx == 0 && ky == 0 && j0 == 0) sp[0] = sn[0] = T2{0, 0};  // k = 0: pol = 0
    const T dw = (T)g.dw;
    const int64_t rho2 = (int64_t)nx * nx + (int64_t)ny * ny;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int kz = j0 + e * TPT;
        const int nz = wrapn(kz, N);
        const int64_t n2 = rho2 + (int64_t)nz * nz;
        // the phase q tau in FP64 (large arguments); FP32 storage evaluates the reduced phase in FP32
        double q = sqrt((double)n2) * g.dwav * g.tau;
        T cq, sq;
        if constexpr (sizeof(T) == 8) {
            cq = cos(q);
            sq = sin(q);
        } else {
            const double twopi = 6.283185307179586476925286766559;
            float r = (float)(q - twopi * rint(q * (1.0 / twopi)));
            __sincosf(r, &sq, &cq);
        }
        const T2 SP = {sp[e].x * dw, sp[e].y * dw};
        const T2 SN = {sn[e].x * dw, sn[e].y * dw};
        const int64_t ai = tile0 + (int64_t)kz * zs;
        // the accumulators are streamed: each is read and written once per step
        T2 a11 = ld_stream(A + ai), a22 = ld_stream(A + 3 * ms + ai), a12 = ld_stream(A + 2 * ms + ai);
        st_stream(A + ai, T2{a11.x + cq * SP.x, a11.y + cq * SP.y});
        st_stream(A + 3 * ms + ai, T2{a22.x + sq * SN.x, a22.y + sq * SN.y});
        if (!g.fix_swap) {
            T2 r12 = {a12.x + (sq * SP.x + cq * SN.x), a12.y + (sq * SP.y + cq * SN.y)};
            const T n6 = n2 > 0 ? ((T)rho2 / (T)n2) / 2 * dw : (T)0;  // pol_33 dw = r^2/2 dw
            st_stream(A + 2 * ms + ai, r12);
            st_stream(A + ms + ai, T2{r12.x - cq * (n6 * v[e].x), r12.y - cq * (n6 * v[e].y)});
        } else {
            T2 a21 = ld_stream(A + ms + ai);
            st_stream(A + ms + ai, T2{a21.x + sq * SP.x, a21.y + sq * SP.y});
            st_stream(A + 2 * ms + ai, T2{a12.x + cq * SN.x, a12.y + cq * SN.y});
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the k-points' loads from piling up in registers
    }
}

// ------------------------------------------------------------------ x pass (R2C via pairing)
// Block = TB rows of one packed component m at one z.  Row FFT of C_m = S_a + i S_b, then
//   S_a^(k) = (X(k) + conj(X(N-k)))/2,  S_b^(k) = (X(k) - conj(X(N-k)))/(2i),  k = 0..N/2,
// written to W[2m] and W[2m+1].

template <typename T>
__global__ void __launch_bounds__(FFT_NT) gw_xpass_kernel(int N, int P, int TB, const typename V2<T>::type* __restrict__ src,
                                                         typename V2<T>::type* __restrict__ W, FftPlan plan,
                                                         const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int m = blockIdx.z, z = blockIdx.y, y0 = blockIdx.x * TB;
    const int nrow = min(TB, N - y0);
    const int64_t NN = (int64_t)N * N;
    const T2* in = src + (int64_t)m * NN * N + (int64_t)z * NN + (int64_t)y0 * N;
    for (int e = threadIdx.x; e < nrow * N; e += FFT_NT) s[e] = in[e];
    for (int e = nrow * N + threadIdx.x; e < TB * N; e += FFT_NT) s[e] = T2{0, 0};
    __syncthreads();
    fft_lds<T, +1, FFT_NT, fft_elems<T>(), false>(s, plan, TB, 1, N, tw);
    const int nh = N / 2 + 1;
    T2* wa = W + ((int64_t)(2 * m) * NN + (int64_t)z * N + y0) * P;
    T2* wb = W + ((int64_t)(2 * m + 1) * NN + (int64_t)z * N + y0) * P;
    for (int e = threadIdx.x; e < nrow * nh; e += FFT_NT) {
        int r = e / nh, k = e - r * nh;
        T2 xk = s[r * N + k];
        T2 xm = s[r * N + (k == 0 ? 0 : N - k)];
        T2 a = {(T)0.5 * (xk.x + xm.x), (T)0.5 * (xk.y - xm.y)};
        // (xk - conj(xm)) / (2i) = (dr + i di) * (-i/2) = (di/2, -dr/2)
        T2 b = {(T)0.5 * (xk.y + xm.y), (T)0.5 * (xm.x - xk.x)};
        wa[(int64_t)r * P + k] = a;
        wb[(int64_t)r * P + k] = b;
    }
}

template <typename T>
void launch_gw_xpass(hipStream_t s, int N, int pitch, const void* src, void* W, const FftPlan& plan, const void* tw)
{
    using T2 = typename V2<T>::type;
    bool done = p2_dispatch(N, [&]<int NN>() {
        constexpr int XR = p2_xrows<NN>();
        hipLaunchKernelGGL((gw_xpass_p2_kernel<T, NN>), dim3(NN / XR, NN, 3), dim3(XR * (NN / 8)),
                           (size_t)XR * NN * sizeof(T2), s, (const T2*)src, (T2*)W, pitch, (const T2*)tw);
    });
    if (done) { AXG_HIP(hipGetLastError()); return; }
    int TB = fft_elems<T>() / N;
    if (TB > N) TB = N;
    dim3 grid((N + TB - 1) / TB, N, 3);
    size_t lds = (size_t)TB * N * sizeof(T2);
    hipLaunchKernelGGL((gw_xpass_kernel<T>), grid, dim3(FFT_NT), lds, s, N, pitch, TB, (const T2*)src, (T2*)W, plan,
                       (const T2*)tw);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ strided (column) FFT
// In place along an axis with element stride es; plane p starts at p*ps; columns are the
// contiguous index 0..ncol-1, handled TX at a time (TX * sizeof(complex) = 128 bytes).

template <typename T, int SIGN>
__global__ void __launch_bounds__(FFT_NT) fft_cols_kernel(int N, int TX, typename V2<T>::type* __restrict__ data, int64_t es,
                                                         int64_t ps, int ncol, FftPlan plan,
                                                         const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int c0 = blockIdx.x * TX;
    T2* base = data + (int64_t)blockIdx.y * ps + c0;
    for (int e = threadIdx.x; e < N * TX; e += FFT_NT) {
        int n = e / TX, b = e - n * TX;
        s[e] = (c0 + b < ncol) ? base[(int64_t)n * es + b] : T2{0, 0};
    }
    __syncthreads();
    fft_lds<T, SIGN, FFT_NT, fft_elems<T>(), true>(s, plan, TX, TX, 1, tw);
    for (int e = threadIdx.x; e < N * TX; e += FFT_NT) {
        int n = e / TX, b = e - n * TX;
        if (c0 + b < ncol) base[(int64_t)n * es + b] = s[e];
    }
}

template <typename T, int SIGN>
void launch_fft_cols(hipStream_t s, int N, void* data, int64_t es, int64_t ps, int nplanes, int ncol,
                     const FftPlan& plan, const void* tw)
{
    using T2 = typename V2<T>::type;
    bool done = p2_dispatch(N, [&]<int NN>() {
        constexpr int TXP = p2_tx<T, NN>();
        hipLaunchKernelGGL((fft_cols_p2_kernel<T, NN, SIGN>), dim3((ncol + TXP - 1) / TXP, nplanes),
                           dim3(TXP * (NN / 8)), (size_t)NN * TXP * sizeof(T2), s, (T2*)data, es, ps, ncol,
                           (const T2*)tw);
    });
    if (done) { AXG_HIP(hipGetLastError()); return; }
    int TX = col_width<T>(N);
    if (TX < 1) throw Error(AXG_EINVAL, "column FFT: N too large for the LDS tile");
    dim3 grid((ncol + TX - 1) / TX, nplanes);
    size_t lds = (size_t)N * TX * sizeof(T2);
    hipLaunchKernelGGL((fft_cols_kernel<T, SIGN>), grid, dim3(FFT_NT), lds, s, N, TX, (T2*)data, es, ps, ncol, plan,
                       (const T2*)tw);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ contiguous (row) FFT

template <typename T, int SIGN>
__global__ void __launch_bounds__(FFT_NT) fft_rows_kernel(int N, int TB, int64_t nrows, typename V2<T>::type* __restrict__ data,
                                                         FftPlan plan, const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int64_t r0 = (int64_t)blockIdx.x * TB;
    const int nrow = (int)min((int64_t)TB, nrows - r0);
    T2* base = data + r0 * N;
    for (int e = threadIdx.x; e < nrow * N; e += FFT_NT) s[e] = base[e];
    for (int e = nrow * N + threadIdx.x; e < TB * N; e += FFT_NT) s[e] = T2{0, 0};
    __syncthreads();
    fft_lds<T, SIGN, FFT_NT, fft_elems<T>(), false>(s, plan, TB, 1, N, tw);
    for (int e = threadIdx.x; e < nrow * N; e += FFT_NT) base[e] = s[e];
}

// Rows of the compile-time sizes: RB rows per block of RB * N/8 threads, each thread holding 8 elements of its
// row in registers (fft_p2.h), loads and stores coalesced along the row.
template <int N> constexpr int p2_rows_per_block() { return (N / 8) >= 256 ? 1 : 256 / (N / 8); }

template <typename T, int N, int SIGN>
__global__ void __launch_bounds__(p2_rows_per_block<N>() * (N / 8))
    fft_rows_p2_kernel(typename V2<T>::type* __restrict__ data, int64_t nrows, const typename V2<T>::type* __restrict__ tw)
{
    using T2 = typename V2<T>::type;
    constexpr int TPT = N / 8, RB = p2_rows_per_block<N>();
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int rr = threadIdx.x / TPT, j0 = threadIdx.x % TPT;
    const int64_t row = (int64_t)blockIdx.x * RB + rr;
    const bool active = row < nrows;
    T2* base = data + (active ? row : 0) * N;
    T2 v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = active ? base[j0 + e * TPT] : T2{0, 0};
    const P2Tw<T, N> tws = p2_load_twiddles<T, N, SIGN>(j0, tw);
    fft_p2_if<T, N, SIGN>(active, v, j0, s + rr * N, 1, 0, tws);
    if (active) {
#pragma unroll
        for (int e = 0; e < 8; e++) base[j0 + e * TPT] = v[e];
    }
}

template <typename T, int SIGN>
void launch_fft_rows(hipStream_t s, int N, void* data, int64_t nrows, const FftPlan& plan, const void* tw)
{
    using T2 = typename V2<T>::type;
    bool done = p2_dispatch(N, [&]<int NN>() {
        constexpr int RB = p2_rows_per_block<NN>();
        hipLaunchKernelGGL((fft_rows_p2_kernel<T, NN, SIGN>), dim3((unsigned)((nrows + RB - 1) / RB)),
                           dim3(RB * (NN / 8)), (size_t)RB * NN * sizeof(T2), s, (T2*)data, nrows, (const T2*)tw);
    });
    if (done) { AXG_HIP(hipGetLastError()); return; }
    int TB = fft_elems<T>() / N;
    if (TB < 1) throw Error(AXG_EINVAL, "row FFT: N too large for the LDS tile");
    int64_t nb = (nrows + TB - 1) / TB;
    size_t lds = (size_t)TB * N * sizeof(T2);
    hipLaunchKernelGGL((fft_rows_kernel<T, SIGN>), dim3((unsigned)nb), dim3(FFT_NT), lds, s, N, TB, nrows, (T2*)data,
                       plan, (const T2*)tw);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ z pass + accumulation
//
// Block = one ky and TX consecutive kx.  For each source component c the z column tile is
// transformed in LDS; every thread then owns k-points (kz, kx) and accumulates
// SP = sum_c pol_c(k) D_c(k) and SN = sum_c conj(pol_c(k)) D_c(k).  With c = cos(q tau),
// s = sin(q tau) and the step weight dw, the six sequential updates of
// grav_waves.f90:239-242 (P = pol*dw, N = conj(P)) collapse to
//   A11 += c dw SP;   A22 += s dw SN;   A12 += s dw SP + c dw SN;   A21 = A12 - c N6 D6,
// where A12 is the updated value, A21's old value is discarded as in the reference, and
// N6 = pol_33 dw = (r^2/2) dw is real; D6 (the (3,3) component, transformed last) is still in
// LDS when the update runs.  With fix_swap: A21 += s dw SP, A12 += c dw SN.

constexpr int ZNT = 512;  // threads of the accumulation pass

template <typename T>
__device__ __forceinline__ void pol_vec(int nx, int ny, int nz, typename V2<T>::type* p)
{
    // grav_waves.f90:165-183 (n != 0)
    T qx = (T)nx, qy = (T)ny, qz = (T)nz;
    T q = sqrt(qx * qx + qy * qy + qz * qz);
    qx = qx / q; qy = qy / q; qz = qz / q;
    T r = sqrt(qx * qx + qy * qy);
    T c, s;
    if (r > 0) { c = qx / r; s = qy / r; }
    else { c = 1; s = 0; }
    p[0] = {qz * c, -s};
    p[1] = {qz * s, c};
    p[2] = {-r, 0};
}

template <typename T>
__global__ void __launch_bounds__(ZNT) gw_zpass_acc_kernel(int N, int P, const typename V2<T>::type* __restrict__ W,
                                                          typename V2<T>::type* __restrict__ A, FftPlan plan,
                                                          const typename V2<T>::type* __restrict__ tw,
                                                          GwStepScalars g, int TX, StepRef sr)
{
#pragma clang fp contract(off)
    if (const StepEntry* se = step_entry(sr)) {
        g.tau = se->tau;
        g.dw = se->dw;
    }
    using T2 = typename V2<T>::type;
    constexpr int KPT = fft_elems<T>() / ZNT;  // k-points per thread (upper bound)
    extern __shared__ unsigned char smem_raw[];
    T2* s = reinterpret_cast<T2*>(smem_raw);
    const int nh = N / 2 + 1;
    const int c0 = blockIdx.x * TX, ky = blockIdx.y;
    const int64_t NN = (int64_t)N * N;
    const int64_t es = (int64_t)N * P;  // z stride in W and A
    const int nk = N * TX;
    const int ny = wrapn(ky, N);

    T2 sp[KPT], sn[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) sp[i] = sn[i] = T2{0, 0};

    for (int c = 0; c < 6; c++) {
        // (i,j) of component c in the order (1,1),(2,1),(2,2),(3,1),(3,2),(3,3), 0-based
        const int ii = c == 0 ? 0 : (c <= 2 ? 1 : 2);
        const int jj = (c == 0 || c == 1 || c == 3) ? 0 : ((c == 2 || c == 4) ? 1 : 2);
        if (c > 0) __syncthreads();
        const T2* base = W + (int64_t)c * NN * P + (int64_t)ky * P + c0;
        for (int e = threadIdx.x; e < nk; e += ZNT) {
            int n = e / TX, b = e - n * TX;
            s[e] = (c0 + b < nh) ? base[(int64_t)n * es + b] : T2{0, 0};
        }
        __syncthreads();
        fft_lds<T, +1, ZNT, fft_elems<T>(), true>(s, plan, TX, TX, 1, tw);
#pragma unroll
        for (int i = 0; i < KPT; i++) {
            int e = threadIdx.x + i * ZNT;
            if (e < nk && c0 + (e % TX) < nh) {
                int kz = e / TX, kx = c0 + (e - kz * TX);
                int nx = wrapn(kx, N), nz = wrapn(kz, N);
                T2 pol = {0, 0};
                if (nx != 0 || ny != 0 || nz != 0) {
                    T2 pv[3];
                    pol_vec<T>(nx, ny, nz, pv);
                    T2 a = ii == 0 ? pv[0] : (ii == 1 ? pv[1] : pv[2]);
                    T2 b = jj == 0 ? pv[0] : (jj == 1 ? pv[1] : pv[2]);
                    pol = {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
                    if (ii == jj) { pol.x = pol.x / 2; pol.y = pol.y / 2; }
                }
                T2 d = s[e];
                sp[i].x += pol.x * d.x - pol.y * d.y;
                sp[i].y += pol.x * d.y + pol.y * d.x;
                sn[i].x += pol.x * d.x + pol.y * d.y;
                sn[i].y += pol.x * d.y - pol.y * d.x;
            }
        }
    }
    // s[] still holds D6 = the transformed (3,3) source

#pragma unroll
    for (int i = 0; i < KPT; i++) {
        int e = threadIdx.x + i * ZNT;
        if (e < nk && c0 + (e % TX) < nh) {
            int kz = e / TX, kx = c0 + (e - kz * TX);
            int nx = wrapn(kx, N), nz = wrapn(kz, N);
            int64_t rho2 = (int64_t)nx * nx + (int64_t)ny * ny;
            int64_t n2 = rho2 + (int64_t)nz * nz;
            double qn = sqrt((double)n2);
            double q = qn * g.dwav * g.tau;
            double cq, sq;
            if constexpr (sizeof(T) == 8) {
                cq = cos(q);
                sq = sin(q);
            } else {
                // FP32 storage: reduce the phase in FP64, evaluate in FP32
                const double twopi = 6.283185307179586476925286766559;
                float r = (float)(q - twopi * rint(q * (1.0 / twopi)));
                float cf, sf;
                __sincosf(r, &sf, &cf);
                cq = cf;
                sq = sf;
            }
            const double dw = g.dw;
            double2 SP = {(double)sp[i].x * dw, (double)sp[i].y * dw};
            double2 SN = {(double)sn[i].x * dw, (double)sn[i].y * dw};
            int64_t ai = (int64_t)kz * es + (int64_t)ky * P + kx;
            const int64_t ms = NN * P;
            T2 a11 = A[ai], a21 = A[ms + ai], a12 = A[2 * ms + ai], a22 = A[3 * ms + ai];
            double2 r11 = {(double)a11.x + cq * SP.x, (double)a11.y + cq * SP.y};
            double2 r22 = {(double)a22.x + sq * SN.x, (double)a22.y + sq * SN.y};
            double2 r21, r12;
            if (!g.fix_swap) {
                r12 = {(double)a12.x + (sq * SP.x + cq * SN.x), (double)a12.y + (sq * SP.y + cq * SN.y)};
                double n6 = n2 > 0 ? ((double)rho2 / (double)n2) / 2 * dw : 0.0;
                T2 d6 = s[e];
                r21 = {r12.x - cq * (n6 * (double)d6.x), r12.y - cq * (n6 * (double)d6.y)};
            } else {
                r21 = {(double)a21.x + sq * SP.x, (double)a21.y + sq * SP.y};
                r12 = {(double)a12.x + cq * SN.x, (double)a12.y + cq * SN.y};
            }
            A[ai] = T2{(T)r11.x, (T)r11.y};
            A[ms + ai] = T2{(T)r21.x, (T)r21.y};
            A[2 * ms + ai] = T2{(T)r12.x, (T)r12.y};
            A[3 * ms + ai] = T2{(T)r22.x, (T)r22.y};
        }
    }
}

GwLayout gw_layout(int N, bool fp32)
{
    GwLayout L{N, half_pitch(N), 0, 0};
    int zt = 0;
    p2_dispatch(N, [&]<int NN>() { zt = fp32 ? p2_ztx<float, NN>() : p2_ztx<double, NN>(); });
    if (zt) {
        L.zt = zt;
        L.ntile = (N / 2 + 1 + zt - 1) / zt;
    }
    return L;
}

template <typename T>
void launch_gw_ypass(hipStream_t s, int N, int pitch, void* W, void* W2, const GwLayout& L, const FftPlan& plan,
                     const void* tw, int variant, const LatRef* lat, int nlat)
{
    using T2 = typename V2<T>::type;
    if (lat && !L.zt) throw Error(AXG_EINVAL, "stacked y pass needs the tiled layout");
    if (L.zt) {
        bool done = p2_dispatch(N, [&]<int NN>() {
            if (!TiledLayout<NN, p2_ztx<T, NN>()>::matches(L) || pitch != half_pitch_c<NN>())
                throw Error(AXG_EINVAL, "y pass: layout does not match the compiled tile");
            // variant 1: half-width column tiles (more, smaller blocks per CU); variant 2: two
            // z planes per block.  Where the one-column tile is narrower than a 128-byte line
            // (p2_tx2), variants 0 and 2 run the two-column kernel instead and variant 1 the
            // one-column tile.
            // buffers as arguments or from the stack's LatRef table (global or flat addressing)
            auto go = [&]<int TXP, int KZ>() {
                auto g1 = [&]<bool STK>() {
                    hipLaunchKernelGGL((gw_ypass_p2_kernel<T, NN, TXP, KZ, STK>),
                                       dim3((NN / 2 + 1 + TXP - 1) / TXP, 6 * NN / KZ, nlat), dim3(TXP * (NN / 8)),
                                       (size_t)NN * TXP * sizeof(T2), s, (const T2*)W, (T2*)W2, (const T2*)tw, lat);
                };
                if (lat) g1.template operator()<true>();
                else g1.template operator()<false>();
            };
            constexpr int TXP = p2_tx<T, NN>(), TXH = TXP > 1 ? TXP / 2 : 1, TX2 = p2_tx2<T, NN>();
            if constexpr (TX2 > 0) {
                auto go2 = [&]<int KZ>() {
                    using L = TiledLayout<NN, p2_ztx<T, NN>()>;
                    constexpr int cols = L::NT * L::ZTW > NN / 2 + 1 ? L::NT * L::ZTW : NN / 2 + 1;
                    auto g2 = [&]<bool STK>() {
                        hipLaunchKernelGGL((gw_ypass_p2w_kernel<T, NN, TX2, KZ, STK>),
                                           dim3((cols + TX2 - 1) / TX2, 6 * NN / KZ, nlat), dim3((TX2 / 2) * (NN / 8)), 0,
                                           s, (const T2*)W, (T2*)W2, (const T2*)tw, lat);
                    };
                    if (lat) g2.template operator()<true>();
                    else g2.template operator()<false>();
                };
                if (variant == 1) go.template operator()<TXP, 1>();
                else if (variant == 2) go2.template operator()<2>();
                else go2.template operator()<1>();
            } else {
                if (variant == 1) go.template operator()<TXH, 1>();
                else if (variant == 2) go.template operator()<TXP, 2>();
                else go.template operator()<TXP, 1>();
            }
        });
        if (!done) throw Error(AXG_EINVAL, "tiled y pass needs N = 2^k or 3 * 2^k");
        AXG_HIP(hipGetLastError());
        return;
    }
    launch_fft_cols<T, +1>(s, N, W, pitch, (int64_t)N * pitch, 6 * N, N / 2 + 1, plan, tw);
}

template <typename T>
void launch_gw_zpass_acc(hipStream_t s, int N, int pitch, const void* W, void* A, const GwLayout& L,
                         const FftPlan& plan, const void* tw, const GwStepScalars& g, const StepRef& sr,
                         const LatRef* lat, int nlat)
{
    using T2 = typename V2<T>::type;
    if (lat && !L.zt) throw Error(AXG_EINVAL, "stacked z pass needs the tiled layout");
    bool done = L.zt && p2_dispatch(N, [&]<int NN>() {
        constexpr int TXP = p2_ztx<T, NN>();
        if (!TiledLayout<NN, TXP>::matches(L))
            throw Error(AXG_EINVAL, "accumulation layout does not match the z pass tile");
        hipLaunchKernelGGL((gw_zpass_p2_kernel<T, NN>), dim3(L.ntile, NN, nlat), dim3(TXP * (NN / 8)), 0, s,
                           (const T2*)W, (T2*)A, (const T2*)tw, g, sr, lat);
    });
    if (done) { AXG_HIP(hipGetLastError()); return; }
    int TX = col_width<T>(N);
    if (TX < 1) throw Error(AXG_EINVAL, "z pass: N too large for the LDS tile");
    int nh = N / 2 + 1;
    dim3 grid((nh + TX - 1) / TX, N);
    size_t lds = (size_t)N * TX * sizeof(T2);
    hipLaunchKernelGGL((gw_zpass_acc_kernel<T>), grid, dim3(ZNT), lds, s, N, pitch, (const T2*)W, (T2*)A, plan,
                       (const T2*)tw, g, TX, sr);
    AXG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ GW spectrum binning
// bin_of_n2[n2] = reference bin (1..nb) of a mode with |n|^2 = n2, 0 = outside; bins are
// non-decreasing in n2, so along a half-spectrum row (kx = 0..N/2, |nx| increasing) every
// bin is one contiguous run.  Each block walks its rows in a fixed order; the thread at the
// start of a run sums the run sequentially and adds it to the block's LDS bin (runs of one
// row have distinct bins, so no two threads touch the same bin).  Block partials are then
// summed in block order: the reduction is deterministic.  Weight 2 for 0 < kx < N/2 (the
// mirrored half), 1 on the kx = 0 and kx = N/2 planes; k = 0 is excluded (q > 0).

// Each thread holds its kx values' four accumulators of the next row in registers, loaded while
// the current row is binned (the loads do not wait for the bin lookup either): a row's loads are
// in flight during the previous row's run sums, and each run's sum walks a length recorded by the
// run's last element.  The sums and their order are those of a row at a time, so the spectrum is
// the same bit for bit.  512^3, 128 bins: kernel 1.13 -> 0.77 ms in FP32C (its 1.18 GB now move
// with the lookups and run sums overlapped), whole spectrum call 1.43 -> 0.83 ms (FP64 1.36 ->
// 0.84-0.87 ms).
// block size: 512 threads for FP32 accumulators, 256 for FP64 (512^3: 0.83 -> 0.72 ms per FP32C
// spectrum with 512, FP64 0.84-0.87 -> 1.04 ms, so FP64 keeps 256); kx values per thread up to
// N / 2 + 1 <= 768
template <typename T> constexpr int gw_bin_nt() { return sizeof(T) == 4 ? 512 : 256; }
template <typename T> constexpr int gw_bin_kpt() { return (768 + gw_bin_nt<T>() - 1) / gw_bin_nt<T>(); }

template <typename T>
__global__ void __launch_bounds__(gw_bin_nt<T>()) gw_bin_kernel(int N, GwLayout L, const typename V2<T>::type* __restrict__ A,
                                                           const uint16_t* __restrict__ bin_of_n2, int64_t n2max,
                                                           int nb, double* __restrict__ partial)
{
    using T2 = typename V2<T>::type;
    constexpr int GW_BIN_NT = gw_bin_nt<T>(), GW_BIN_KPT = gw_bin_kpt<T>();
    extern __shared__ double sh[];
    double* bins = sh;              // 2*nb: numer, denom
    double* v = sh + 2 * nb;        // N/2+1 values
    int* bb = (int*)(v + (N / 2 + 1));
    int* rend = bb + (N / 2 + 1);   // nb: one past the last kx of bin b's run in the row
    const int nh = N / 2 + 1;
    const int64_t NN = (int64_t)N * N, ms = L.comp();
    for (int b = threadIdx.x; b < 2 * nb; b += blockDim.x) bins[b] = 0;
    T2 a[GW_BIN_KPT][4];
    int bq[GW_BIN_KPT];
    auto fetch = [&](int64_t row) {
        const int kz = (int)(row / N), ky = (int)(row - (int64_t)kz * N);
        const int ny = wrapn(ky, N), nz = wrapn(kz, N);
#pragma unroll
        for (int i = 0; i < GW_BIN_KPT; i++) {
            const int kx = threadIdx.x + i * GW_BIN_NT;
            if (kx < nh) {
                const int nx = wrapn(kx, N);
                const int64_t n2 = (int64_t)nx * nx + (int64_t)ny * ny + (int64_t)nz * nz;
                bq[i] = (n2 > 0) ? bin_of_n2[n2] : 0;
                // modes past the last bin (about half of the half spectrum) are not read; the test
                // is on n2 itself, so the loads do not wait for the bin lookup
                if (n2 > 0 && n2 <= n2max) {
                    const int64_t ai = L.idx(0, kz, ky, kx);
#pragma unroll
                    for (int m = 0; m < 4; m++) a[i][m] = A[m * ms + ai];
                }
            }
        }
    };
    int64_t row = blockIdx.x;
    if (row < NN) fetch(row);
    __syncthreads();
    for (; row < NN; row += gridDim.x) {
#pragma unroll
        for (int i = 0; i < GW_BIN_KPT; i++) {
            const int kx = threadIdx.x + i * GW_BIN_NT;
            if (kx < nh) {
                double s = 0;
                if (bq[i] > 0) {
#pragma unroll
                    for (int m = 0; m < 4; m++) s += (double)a[i][m].x * a[i][m].x + (double)a[i][m].y * a[i][m].y;
                }
                v[kx] = s;
                bb[kx] = bq[i];
            }
        }
        if (row + gridDim.x < NN) fetch(row + gridDim.x);
        __syncthreads();
        // the last element of each run records where the run ends, so that the run's sum walks a
        // known length instead of testing the next element's bin at every step
        for (int kx = threadIdx.x; kx < nh; kx += blockDim.x) {
            const int b = bb[kx];
            if (b > 0 && (kx == nh - 1 || bb[kx + 1] != b)) rend[b - 1] = kx + 1;
        }
        __syncthreads();
        for (int kx = threadIdx.x; kx < nh; kx += blockDim.x) {
            int b = bb[kx];
            if (b > 0 && (kx == 0 || bb[kx - 1] != b)) {
                const int e = rend[b - 1];
                double sn = 0, sd = 0;
                for (int j = kx; j < e; j++) {
                    double w = (j == 0 || j == N / 2) ? 1.0 : 2.0;
                    sn += w * v[j];
                    sd += w;
                }
                bins[b - 1] += sn;
                bins[nb + b - 1] += sd;
            }
        }
        __syncthreads();
    }
    for (int b = threadIdx.x; b < 2 * nb; b += blockDim.x) partial[(int64_t)blockIdx.x * 2 * nb + b] = bins[b];
}

__global__ void sum_partials_kernel(const double* __restrict__ partial, int nwg, int width, double* __restrict__ out)
{
    // one thread per output column; fixed block order (the unrolled loads are in flight together,
    // the additions stay in order: 283 -> 37 us for 1024 partials of 256 columns)
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= width) return;
    double s = 0;
#pragma unroll 16
    for (int w = 0; w < nwg; w++) s += partial[(int64_t)w * width + b];
    out[b] = s;
}

template <typename T>
void launch_gw_bin(hipStream_t s, int N, const GwLayout& L, const void* A, const uint16_t* bin_of_n2, int64_t n2max, int nb,
                   double* partial, int nwg, double* out)
{
    using T2 = typename V2<T>::type;
    int nh = N / 2 + 1;
    size_t lds = (size_t)(2 * nb + nh) * sizeof(double) + (size_t)(nh + nb) * sizeof(int);
    if (nh > gw_bin_kpt<T>() * gw_bin_nt<T>()) throw Error(AXG_EINVAL, "GW spectrum: N too large for the binning kernel");
    hipLaunchKernelGGL((gw_bin_kernel<T>), dim3(nwg), dim3(gw_bin_nt<T>()), lds, s, N, L, (const T2*)A, bin_of_n2, n2max,
                       nb, partial);
    AXG_HIP(hipGetLastError());
    hipLaunchKernelGGL(sum_partials_kernel, dim3((2 * nb + 255) / 256), dim3(256), 0, s, partial, nwg, 2 * nb, out);
    AXG_HIP(hipGetLastError());
}

// accumulator component m <-> compact FP64 [kz][ky][kx], kx = 0..N/2
template <typename T>
__global__ void gw_pack_kernel(int N, GwLayout L, typename V2<T>::type* A, int m, double2* compact, int dir)
{
    using T2 = typename V2<T>::type;
    const int nh = N / 2 + 1;
    const int64_t tot = (int64_t)N * N * nh;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x) {
        int kx = (int)(i % nh);
        int64_t r = i / nh;
        int ky = (int)(r % N), kz = (int)(r / N);
        int64_t a = L.idx(m, kz, ky, kx);
        if (dir > 0) compact[i] = double2{(double)A[a].x, (double)A[a].y};
        else A[a] = T2{(T)compact[i].x, (T)compact[i].y};
    }
}

template <typename T>
void launch_gw_pack(hipStream_t s, int N, const GwLayout& L, void* A, int m, double2* compact, int dir)
{
    using T2 = typename V2<T>::type;
    hipLaunchKernelGGL((gw_pack_kernel<T>), dim3(16384), dim3(256), 0, s, N, L, (T2*)A, m, compact, dir);
    AXG_HIP(hipGetLastError());
}

#define AXG_INST_GW(T)                                                                                          \
    template void launch_gw_xpass<T>(hipStream_t, int, int, const void*, void*, const FftPlan&, const void*);   \
    template void launch_fft_cols<T, +1>(hipStream_t, int, void*, int64_t, int64_t, int, int, const FftPlan&,   \
                                         const void*);                                                          \
    template void launch_fft_rows<T, +1>(hipStream_t, int, void*, int64_t, const FftPlan&, const void*);        \
    template void launch_gw_ypass<T>(hipStream_t, int, int, void*, void*, const GwLayout&, const FftPlan&,       \
                                     const void*, int, const LatRef*, int);                                     \
    template void launch_gw_zpass_acc<T>(hipStream_t, int, int, const void*, void*, const GwLayout&,             \
                                         const FftPlan&, const void*, const GwStepScalars&, const StepRef&,     \
                                         const LatRef*, int);                                                   \
    template void launch_gw_bin<T>(hipStream_t, int, const GwLayout&, const void*, const uint16_t*, int64_t, int, double*, \
                                   int, double*);                                                               \
    template void launch_gw_pack<T>(hipStream_t, int, const GwLayout&, void*, int, double2*, int);
AXG_INST_GW(double)
AXG_INST_GW(float)
template void launch_fft_cols<double, -1>(hipStream_t, int, void*, int64_t, int64_t, int, int, const FftPlan&,
                                          const void*);
template void launch_fft_rows<double, -1>(hipStream_t, int, void*, int64_t, const FftPlan&, const void*);
template void launch_fft_cols<float, -1>(hipStream_t, int, void*, int64_t, int64_t, int, int, const FftPlan&,
                                         const void*);
template void launch_fft_rows<float, -1>(hipStream_t, int, void*, int64_t, const FftPlan&, const void*);

}  // namespace axg
