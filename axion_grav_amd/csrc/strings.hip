// String cells: straight-line fit, chord length and velocity, one thread per string cell
// (id_defects' string loop, defects_3d.f90:167-205, with calc_chi2_string 384-433,
// position_string 435-603, string_length 1027-1056 and velocity_string 605-690 of the VEL = 0
// build), plus the per-block sums of add_string_length (1654-1733, group_strings = F).
//
// The fit is a small sequential search per cell (1001 random starts, then up to 1000 simplex
// iterations, 8 atan2 per chi2), so the kernel is bound by FP64 issue, not by HBM: it reads 8
// corner values of phi and pi per string and writes 11 doubles.  Every array of the search
// lives in registers: the insertion of a start point and the bubble sort are unrolled over the
// five simplex slots, so no index into them depends on data.
//
// Two stages, because the simplex phase diverges: most fits converge within a few iterations,
// but about one string in seven never reaches chi2 <= 0.05 and runs all 1000, so nearly every
// wave of 64 strings would wait for 1000 iterations while most of its lanes idle.  Stage one
// runs the random starts and the first `split` simplex iterations of every string; a string still
// searching then parks its simplex (5 points, 5 chi2, iteration count) in a list.  Stage two
// continues the parked strings, one per thread, so the long searches share waves.  A string's
// search is the same sequence of operations whichever stage ends it, so the records do not
// depend on `split`.  The per-block sums are taken afterwards from the records.
//
// Within a wave the simplex iterations run in lockstep, and a shrink (four chi2) on any lane
// used to cost the whole wave four chi2 rounds; the searches that never converge shrink at almost
// every iteration late in their run.  The four chi2 of every shrinking search are now shared out
// over the wave's 64 lanes (up to 16 shrinks per chi2 round), so an iteration costs three rounds
// (reflection, expansion or contraction, shrink) where it cost up to six.
#include "axg_internal.h"

namespace axg {

namespace {

constexpr int NSIMP = 4;       // num_simp_string_pos (defects_3d.f90:26)
constexpr int MAXIT = 1000;    // max_iter_string_pos
constexpr double PI_T = 3.14159265359;
constexpr double TWOPI_T = 2 * PI_T;

// The four uniforms of start point i of string s.  From the reference's stream (dr.mcg): its
// next four draws at xs, which fit_starts positions at pos + 4004 s (defects_3d.f90:137-162,
// 465: each string cell takes 4 x 1001 draws, in site order).  Otherwise one Philox block keyed
// by (k0, k1), counter (i, s, 'strg'): one stream per string, independent of how the strings are
// scheduled.
__device__ __forceinline__ void uniforms(const Draws& dr, uint64_t& xs, int64_t s, int i, double u[4])
{
    if (dr.mcg) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            u[k] = mcg59_u(xs);
            xs = mcg59_next(xs);
        }
        return;
    }
    const uint32_t k0 = dr.k0, k1 = dr.k1;
    uint32_t c0 = (uint32_t)i, c1 = (uint32_t)s, c2 = (uint32_t)((uint64_t)s >> 32), c3 = 0x73747267u;
    philox4x32(c0, c1, c2, c3, k0, k1);
    const double sc = 2.3283064365386963e-10;  // 2^-32
    u[0] = ((double)c0 + 0.5) * sc;
    u[1] = ((double)c1 + 0.5) * sc;
    u[2] = ((double)c2 + 0.5) * sc;
    u[3] = ((double)c3 + 0.5) * sc;
}

// calc_chi2_string: variance over the corners of the phase minus the phase of the straight
// string p = (mu, phi, rho, theta); corner i = (iz*2+iy)*2+ix sits at (ix,iy,iz) - 1/2.
struct Phases {
    double v[8];
};

// Out of line (one copy of its 8 atan2 instead of one per call site), with every argument
// passed by value so that nothing of the caller's search state has to live in memory.
__device__ __noinline__ double chi2_string_v(Phases ph, double p0, double p1, double p2, double p3)
{
#pragma clang fp contract(off)
    double mu = p0, phi = p1, u0 = p2 * cos(p3), v0 = p2 * sin(p3);
    double cm = mu, sm = sqrt(1.0 - mu * mu), cp = cos(phi), sp = sin(phi);
    double a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        double x = (i & 1) - 0.5, y = ((i >> 1) & 1) - 0.5, z = (i >> 2) - 0.5;
        double du = -u0 + cm * (x * cp + y * sp) - z * sm;
        double dv = -v0 - x * sp + y * cp;
        double d = ph.v[i] - atan2(dv, du);
        if (d >= PI_T) d = d - TWOPI_T;
        else if (d < -PI_T) d = d + TWOPI_T;
        a = a + d * d;
        b = b + d;
    }
    a = a / 8;
    b = b / 8;
    return a - b * b;
}

__device__ __forceinline__ double chi2_string(const Phases& ph, const double p[4])
{
    return chi2_string_v(ph, p[0], p[1], p[2], p[3]);
}

// string_length: chord of pos + t dir through [-1/2, 1/2]^3, zero when an exit lies beyond
// the reference's cut (just below sqrt(3)/2).
__device__ double string_length(const double pos[3], const double dir[3])
{
#pragma clang fp contract(off)
    const double atmax = 0.867, atmax0 = 0.86602540378378444;
    double tp = atmax, tn = -atmax;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            double w = j - 0.5;
            w = w - pos[i];
            if (fabs(w) > atmax * fabs(dir[i])) continue;
            double t = w / dir[i];
            if (t > 0) { if (t < tp) tp = t; }
            else { if (t > tn) tn = t; }
        }
    return (tp < atmax0 && tn > -atmax0) ? tp - tn : 0.0;
}

__device__ __forceinline__ bool out_of_box(const double t[4])
{
    const double mx[4] = {1.0, PI_T, 0.86602540378443865, PI_T};
    const double mn[4] = {-1.0, -PI_T, 0.0, -PI_T};
    bool o = false;
#pragma unroll
    for (int k = 0; k < 4; k++) o = o || t[k] > mx[k] || t[k] < mn[k];
    return o;
}

// packed-upper Cholesky inverse of a 4x4 SPD matrix, A(r,c) at r + c(c-1)/2 - 1 (1-based),
// in the operation order of the dpptrf('U') + dpptri('U') restatement of the oracle.
__device__ void pp_inverse4(double ap[10])
{
#pragma clang fp contract(off)
#define AP(r, c) ap[(r) + (c) * ((c) - 1) / 2 - 1]
#pragma unroll
    for (int j = 1; j <= 4; j++) {
#pragma unroll
        for (int i = 1; i < j; i++) {
            double s = AP(i, j);
#pragma unroll
            for (int k = 1; k < i; k++) s -= AP(k, i) * AP(k, j);
            AP(i, j) = s / AP(i, i);
        }
        double s = AP(j, j);
#pragma unroll
        for (int k = 1; k < j; k++) s -= AP(k, j) * AP(k, j);
        AP(j, j) = sqrt(s);
    }
#pragma unroll
    for (int j = 1; j <= 4; j++) {
        AP(j, j) = 1 / AP(j, j);
        double ajj = -AP(j, j);
        double col[4];
#pragma unroll
        for (int i = 1; i < j; i++) {
            double s = 0;
#pragma unroll
            for (int k = i; k < j; k++) s += AP(i, k) * AP(k, j);
            col[i - 1] = s * ajj;
        }
#pragma unroll
        for (int i = 1; i < j; i++) AP(i, j) = col[i - 1];
    }
#pragma unroll
    for (int i = 1; i <= 4; i++)
#pragma unroll
        for (int j = i; j <= 4; j++) {
            double s = 0;
#pragma unroll
            for (int k = j; k <= 4; k++) s += AP(i, k) * AP(j, k);
            AP(i, j) = s;
        }
#undef AP
}

__device__ __forceinline__ void copy4(double* d, const double* s)
{
#pragma unroll
    for (int k = 0; k < 4; k++) d[k] = s[k];
}

// position_string (defects_3d.f90:435-603) in three parts.  The search state of one string:
// the five simplex points p and their chi2 c2.
constexpr double STR_TOL = 0.05;  // chi2 at which the simplex stops

// 1001 random starts keep the best five (the first two unsorted, as in the reference).
// The search arrays are locals here and copied out at the end: with the caller's arrays
// (references) the compiler moved c2 to scratch memory inside the 1001-start loop.
__device__ __forceinline__ void fit_starts(const Phases& ph, const Draws& dr, int64_t s,
                                           double (&p_out)[NSIMP + 1][4], double (&c2_out)[NSIMP + 1])
{
#pragma clang fp contract(off)
    uint64_t xs = dr.mcg ? mcg59_jump(dr.x0, dr.pos + 4004 * (uint64_t)s) : 0;
    const double mx[4] = {1.0, PI_T, 0.86602540378443865, PI_T};
    const double mn[4] = {-1.0, -PI_T, 0.0, -PI_T};
    double p[NSIMP + 1][4], c2[NSIMP + 1], t[4];
#pragma unroll
    for (int j = 0; j <= NSIMP; j++) {
        c2[j] = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) p[j][k] = 0;
    }
    int n0 = -1;
    for (int i = 0; i <= MAXIT; i++) {
        double x[4];
        uniforms(dr, xs, s, i, x);
#pragma unroll
        for (int k = 0; k < 4; k++) t[k] = mx[k] * x[k] + mn[k] * (1 - x[k]);
        double chi2 = chi2_string(ph, t);
        if (n0 <= 0) {
            n0 = n0 + 1;
            if (n0 == 0) { c2[0] = chi2; copy4(p[0], t); }
            else { c2[1] = chi2; copy4(p[1], t); }
            continue;
        }
        // c2[n0] by selects (n0 is 1..NSIMP here), so that c2 stays in registers
        double cn = n0 == 1 ? c2[1] : n0 == 2 ? c2[2] : n0 == 3 ? c2[3] : c2[4];
        if (n0 < NSIMP || chi2 <= cn) {
            n0 = n0 + 1 < NSIMP ? n0 + 1 : NSIMP;
            // insertion from slot n0 downwards: shift while chi2 <= the slot above
            bool placed = false;
#pragma unroll
            for (int j = NSIMP; j >= 1; j--) {
                if (j <= n0 && !placed) {
                    if (chi2 <= c2[j - 1]) {
                        c2[j] = c2[j - 1];
                        copy4(p[j], p[j - 1]);
                    } else {
                        c2[j] = chi2;
                        copy4(p[j], t);
                        placed = true;
                    }
                }
            }
            if (!placed) { c2[0] = chi2; copy4(p[0], t); }
        }
    }
#pragma unroll
    for (int j = 0; j <= NSIMP; j++) {
        c2_out[j] = c2[j];
        copy4(p_out[j], p[j]);
    }
}

// Centroid of the simplex without its worst point.
__device__ __forceinline__ void centroid(const double (&p)[NSIMP + 1][4], double c[4])
{
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = 0;
#pragma unroll
    for (int i0 = 0; i0 < NSIMP; i0++)
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] = c[k] + p[i0][k];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = c[k] / NSIMP;
}

// The reference's bubble sort of the simplex by chi2 (stable: equal chi2 keep their order).
__device__ __forceinline__ void sort_simplex(double (&p)[NSIMP + 1][4], double (&c2)[NSIMP + 1])
{
#pragma unroll
    for (int i0 = NSIMP - 1; i0 >= 0; i0--)
#pragma unroll
        for (int j = 0; j <= i0; j++)
            if (c2[j] > c2[j + 1]) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    double q = p[j][k];
                    p[j][k] = p[j + 1][k];
                    p[j + 1][k] = q;
                }
                double q = c2[j];
                c2[j] = c2[j + 1];
                c2[j + 1] = q;
            }
}

// LDS of the shared shrink evaluation: SHR_SLOTS searches per chi2 round, each with its four
// shrunk points and its corner phases, and the chi2 results.  The search loop's trip count is
// decided per wave (__any), and its rounds synchronise the block, so a fit block must be exactly
// one wave.
constexpr int SHR_SLOTS = 16;
static_assert(STRING_NT == 64, "string fit blocks are one wave");
struct ShrinkLds {
    double pt[SHR_SLOTS][4][4];
    double ph[SHR_SLOTS][8];
    double res[SHR_SLOTS][4];
};

// The four chi2 of every shrinking search of the wave, shared out over all 64 lanes: up to 16
// searches (64 evaluations) per chi2 round instead of four rounds per search.  Called by every
// lane of the wave (the block is one wave) with shr set on the lanes whose search shrinks.
__device__ __forceinline__ void shrink_chi2(ShrinkLds& L, bool shr, const Phases& ph,
                                            const double (&p)[NSIMP + 1][4], double (&c2)[NSIMP + 1])
{
    const uint64_t m = __ballot(shr);
    if (m == 0) return;
    const int k = __popcll(m);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const int lane = (int)(threadIdx.x & 63);
    for (int c0 = 0; c0 < k; c0 += SHR_SLOTS) {
        const int sl = rank - c0;
        const bool mine = shr && sl >= 0 && sl < SHR_SLOTS;
        if (mine) {
#pragma unroll
            for (int a = 0; a < NSIMP; a++)
#pragma unroll
                for (int kk = 0; kk < 4; kk++) L.pt[sl][a][kk] = p[a + 1][kk];
#pragma unroll
            for (int c = 0; c < 8; c++) L.ph[sl][c] = ph.v[c];
        }
        __syncthreads();
        const int ts = lane >> 2, tp = lane & 3;
        if (c0 + ts < k) {
            Phases q;
#pragma unroll
            for (int c = 0; c < 8; c++) q.v[c] = L.ph[ts][c];
            L.res[ts][tp] = chi2_string_v(q, L.pt[ts][tp][0], L.pt[ts][tp][1], L.pt[ts][tp][2], L.pt[ts][tp][3]);
        }
        __syncthreads();
        if (mine) {
#pragma unroll
            for (int a = 0; a < NSIMP; a++) c2[a + 1] = L.res[sl][a];
        }
        __syncthreads();
    }
}

// The reference's simplex iteration, continued from iteration `it` until chi2 <= 0.05 or
// iteration itmax (the reference's limit is MAXIT).  Every lane of the wave calls it and stays
// in the loop until the wave's last search ends (`valid` false: no search on this lane), because
// the four chi2 of a shrink are evaluated by the whole wave.  A search's arithmetic is the
// reference's operation for operation (the shrunk points do not depend on each other's chi2),
// so where its chi2 are evaluated does not change its bits.
__device__ __forceinline__ void fit_simplex(const Phases& ph, double (&p)[NSIMP + 1][4], double (&c2)[NSIMP + 1],
                                            int& it, int itmax, bool valid, ShrinkLds& L)
{
#pragma clang fp contract(off)
    const double chi2max = 8 * PI_T * PI_T;
    double t[4], c[4], sv[4];
    for (;;) {
        const bool act = valid && c2[0] > STR_TOL && it < itmax;
        if (!__any(act)) break;
        bool shr = false;
        if (act) {
            it = it + 1;
            centroid(p, c);
#pragma unroll
            for (int k = 0; k < 4; k++) t[k] = 2 * c[k] - p[NSIMP][k];
            double chi2 = out_of_box(t) ? chi2max : chi2_string(ph, t);
            // expansion or contraction point, evaluated at one call site (one chi2 round for
            // the wave whichever its lanes take)
            const bool expand = chi2 < c2[NSIMP];
#pragma unroll
            for (int k = 0; k < 4; k++) sv[k] = expand ? 3 * c[k] - 4 * p[NSIMP][k] : 1.5 * c[k] - 0.5 * p[NSIMP][k];
            const double chi2s = out_of_box(sv) ? chi2max : chi2_string(ph, sv);
            if (expand) {
                if (chi2 <= chi2s) { c2[NSIMP] = chi2; copy4(p[NSIMP], t); }
                else { c2[NSIMP] = chi2s; copy4(p[NSIMP], sv); }
            } else {
                if (chi2s <= c2[NSIMP]) {
                    c2[NSIMP] = chi2s;
                    copy4(p[NSIMP], sv);
                } else {
                    // shrink towards the best point
#pragma unroll
                    for (int i0 = 1; i0 <= NSIMP; i0++)
#pragma unroll
                        for (int k = 0; k < 4; k++) p[i0][k] = 0.5 * (p[i0][k] + p[0][k]);
                    shr = true;
                }
            }
        }
        shrink_chi2(L, shr, ph, p, c2);
        if (act) sort_simplex(p, c2);
    }
}

// Position, direction, goodness of fit and chord length of the best point p0 (chi2 c20).
__device__ __forceinline__ void fit_result(const double p0[4], double c20, double pos[3], double dir[3], double& gof,
                                           double& sl)
{
#pragma clang fp contract(off)
    double mu = p0[0], phi = p0[1];
    double u0 = p0[2] * cos(p0[3]), v0 = p0[2] * sin(p0[3]);
    double sm = sqrt(1.0 - mu * mu);
    dir[0] = sm * cos(phi);
    dir[1] = sm * sin(phi);
    dir[2] = mu;
    pos[0] = u0 * mu * cos(phi) - v0 * sin(phi);
    pos[1] = u0 * mu * sin(phi) + v0 * cos(phi);
    pos[2] = -u0 * sm;
    gof = exp(-c20 / 2);
    sl = string_length(pos, dir);
}

// Lattice site of string cell s; its corner sites (periodic wrap) are formed where they are
// read, so no corner index stays live through the fit.
struct StringCell {
    int N, x, y, z;
    __device__ __forceinline__ int64_t site(int c) const
    {
        int cx = x + (c & 1), cy = y + ((c >> 1) & 1), cz = z + (c >> 2);
        cx = cx == N ? 0 : cx;
        cy = cy == N ? 0 : cy;
        cz = cz == N ? 0 : cz;
        return ((int64_t)cz * N + cy) * N + cx;
    }
};

__device__ __forceinline__ StringCell string_cell(int N, const int32_t* __restrict__ cells, int64_t s)
{
    const int64_t i = cells[s];
    const int64_t NN = (int64_t)N * N;
    return StringCell{N, (int)(i % N), (int)((i / N) % N), (int)(i / NN)};
}

template <typename T, bool COMP>
__device__ __forceinline__ Phases corner_phases(const typename V2<T>::type* __restrict__ phi,
                                                const short2* __restrict__ cphi, const StringCell& sc)
{
    Phases ph;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        double2 f = ldf<T, COMP>(phi, cphi, sc.site(c));
        ph.v[c] = atan2(f.y, f.x);
    }
    return ph;
}

// velocity_string and the record of string s (11 doubles: pos in site coordinates counted
// from 0, dir, gof, sl, vel).
template <typename T, bool COMP>
__device__ __forceinline__ void finish_string(const typename V2<T>::type* __restrict__ phi,
                                              const short2* __restrict__ cphi,
                                              const typename V2<T>::type* __restrict__ pi,
                                              const short2* __restrict__ cpi, const StringCell& sc,
                                              const double pos[3], const double dir[3], double gof, double sl,
                                              double dspa, double* __restrict__ rec, int64_t s)
{
#pragma clang fp contract(off)
    const int x = sc.x, y = sc.y, z = sc.z;
    auto corner = [&](int c) { return sc.site(c); };
    // the corner values are read again here rather than kept live through the fit
    double2 f[8], q[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        f[c] = ldf<T, COMP>(phi, cphi, corner(c));
        q[c] = ldf<T, COMP>(pi, cpi, corner(c));
    }

    // velocity_string: least-squares fit of phi and pi to a + b.d over the corners
    double wr[4] = {0, 0, 0, 0}, wi[4] = {0, 0, 0, 0}, dr[4] = {0, 0, 0, 0}, di[4] = {0, 0, 0, 0};
    double mat[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 8; c++) {
        double d[4] = {1.0, (c & 1) - pos[0] - 0.5, ((c >> 1) & 1) - pos[1] - 0.5, (c >> 2) - pos[2] - 0.5};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            wr[m] = wr[m] + d[m] * f[c].x;
            wi[m] = wi[m] + d[m] * f[c].y;
            dr[m] = dr[m] + d[m] * q[c].x;
            di[m] = di[m] + d[m] * q[c].y;
        }
#pragma unroll
        for (int cc = 1; cc <= 4; cc++)
#pragma unroll
            for (int r = 1; r <= cc; r++) mat[r + cc * (cc - 1) / 2 - 1] += d[r - 1] * d[cc - 1];
    }
#pragma unroll
    for (int m = 0; m < 4; m++) {
        wr[m] /= 8;
        wi[m] /= 8;
        dr[m] /= 8;
        di[m] /= 8;
    }
#pragma unroll
    for (int m = 0; m < 10; m++) mat[m] /= 8;
    pp_inverse4(mat);
    double w0r[4] = {0, 0, 0, 0}, w0i[4] = {0, 0, 0, 0}, d0r = 0, d0i = 0;
#pragma unroll
    for (int ii = 1; ii <= 4; ii++)
#pragma unroll
        for (int jj = 1; jj <= 4; jj++) {
            int kk = ii > jj ? ii : jj;
            kk = kk * (kk - 1) / 2 + (ii < jj ? ii : jj) - 1;
            w0r[ii - 1] += mat[kk] * wr[jj - 1];
            w0i[ii - 1] += mat[kk] * wi[jj - 1];
            if (ii == 1) {
                d0r += mat[kk] * dr[jj - 1];
                d0i += mat[kk] * di[jj - 1];
            }
        }
    double a = 0, b = 0, cc = 0;
#pragma unroll
    for (int m = 1; m < 4; m++) {
        a += w0r[m] * w0r[m];
        b += w0r[m] * w0i[m];
        cc += w0i[m] * w0i[m];
    }
    double det = a * cc - b * b;
    double vel[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        double v = (d0r * cc - d0i * b) * w0r[m + 1] + (-d0r * b + d0i * a) * w0i[m + 1];
        vel[m] = -v / det * dspa;
    }

    double* o = rec + 11 * s;
    o[0] = pos[0] + ((double)(x + 1) - 0.5);
    o[1] = pos[1] + ((double)(y + 1) - 0.5);
    o[2] = pos[2] + ((double)(z + 1) - 0.5);
    o[3] = dir[0];
    o[4] = dir[1];
    o[5] = dir[2];
    o[6] = gof;
    o[7] = sl;
    o[8] = vel[0];
    o[9] = vel[1];
    o[10] = vel[2];
}

// Searches parked by stage one: slot j holds string pend[j] and, at state + 26 j, its simplex
// points (20), their chi2 (5) and the iterations done.

// Stage one: random starts and the first `split` simplex iterations of string s = one thread.
template <typename T, bool COMP>
__global__ void __launch_bounds__(STRING_NT) string_fit_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                               const short2* __restrict__ cphi,
                                                               const typename V2<T>::type* __restrict__ pi,
                                                               const short2* __restrict__ cpi,
                                                               const int32_t* __restrict__ cells, int64_t ns,
                                                               Draws dr, double dspa, int split,
                                                               double* __restrict__ rec, int* __restrict__ npend,
                                                               int64_t* __restrict__ pend, double* __restrict__ state)
{
#pragma clang fp contract(off)
    __shared__ ShrinkLds lds;
    const int64_t s = blockIdx.x * (int64_t)STRING_NT + threadIdx.x;
    // lanes past the last string stay for the wave's shared shrink evaluations
    const bool valid = s < ns;
    const StringCell sc = string_cell(N, cells, valid ? s : 0);
    Phases ph;
    double p[NSIMP + 1][4], c2[NSIMP + 1];
    int it = 0;
    if (valid) {
        ph = corner_phases<T, COMP>(phi, cphi, sc);
        fit_starts(ph, dr, s, p, c2);
    }
    fit_simplex(ph, p, c2, it, split < MAXIT ? split : MAXIT, valid, lds);
    if (!valid) return;
    if (c2[0] > STR_TOL && it < MAXIT) {
        // still searching: park it for stage two
        const int j = atomicAdd(npend, 1);
        pend[j] = s;
        double* st = state + (int64_t)j * STRING_STATE;
#pragma unroll
        for (int a = 0; a <= NSIMP; a++) {
#pragma unroll
            for (int k = 0; k < 4; k++) st[4 * a + k] = p[a][k];
            st[20 + a] = c2[a];
        }
        st[25] = (double)it;
        return;
    }
    double pos[3], dir[3], gof, sl;
    fit_result(p[0], c2[0], pos, dir, gof, sl);
    finish_string<T, COMP>(phi, cphi, pi, cpi, sc, pos, dir, gof, sl, dspa, rec, s);
}

// Stage two: the parked searches, one per thread, continued to the end.
constexpr int STRING_NT2 = 64;
static_assert(STRING_NT2 == 64, "string fit blocks are one wave (see ShrinkLds)");
template <typename T, bool COMP>
__global__ void __launch_bounds__(STRING_NT2, 3) string_fit_rest_kernel(
    int N, const typename V2<T>::type* __restrict__ phi, const short2* __restrict__ cphi,
    const typename V2<T>::type* __restrict__ pi, const short2* __restrict__ cpi, const int32_t* __restrict__ cells,
    double dspa, double* __restrict__ rec, const int* __restrict__ npend, const int64_t* __restrict__ pend,
    const double* __restrict__ state)
{
#pragma clang fp contract(off)
    __shared__ ShrinkLds lds;
    const int np = *npend;
    if (blockIdx.x * STRING_NT2 >= (unsigned)np) return;  // the whole wave has no search
    const int j = blockIdx.x * STRING_NT2 + threadIdx.x;
    const bool valid = j < np;
    double p[NSIMP + 1][4], c2[NSIMP + 1];
    int it = 0;
    Phases ph;
    StringCell sc{N, 0, 0, 0};
    int64_t s = 0;
    if (valid) {
        s = pend[j];
        const double* st = state + (int64_t)j * STRING_STATE;
#pragma unroll
        for (int a = 0; a <= NSIMP; a++) {
#pragma unroll
            for (int k = 0; k < 4; k++) p[a][k] = st[4 * a + k];
            c2[a] = st[20 + a];
        }
        it = (int)st[25];
        sc = string_cell(N, cells, s);
        ph = corner_phases<T, COMP>(phi, cphi, sc);
    }
    fit_simplex(ph, p, c2, it, MAXIT, valid, lds);
    if (!valid) return;
    double pos[3], dir[3], gof, sl;
    fit_result(p[0], c2[0], pos, dir, gof, sl);
    finish_string<T, COMP>(phi, cphi, pi, cpi, sc, pos, dir, gof, sl, dspa, rec, s);
}

// velocity_string_YY (defects_3d.f90:824-982, the reference's VEL = 1 build) for string cell s,
// written over the record's velocity.  For each face orientation (u, v, w) = cshift(x, y, z) and
// each of the cell's two faces normal to w: the quadrant test of its four corners (three of the
// four quadrants of arg phi filled); on a pierced face, the least-squares plane of phi over the
// corners (whose normal matrix is diag(4, 1, 1), so the packed inverse is diag(1/4, 1, 1)
// exactly) gives the penetration point (du, dv), clamped to the face; the corners' pi and central
// differences of phi, weighted (0.5 + sign(du, i - 1/2))(0.5 + sign(dv, j - 1/2)) as the
// reference writes them, give the face's velocity.  The cell's velocity is the reference's
// average over the pierced faces, which adds the last face's velocity npen times (its index
// npen in place of the loop's), or 1e-10 per component when no face is pierced.  In z the
// corners of the upper face of the top cell plane lie in the halo plane: the same values as
// plane 0, but a one-sided difference.
template <typename T, bool COMP>
__global__ void __launch_bounds__(STRING_NT) string_vel_yy_kernel(int N, const typename V2<T>::type* __restrict__ phi,
                                                                  const short2* __restrict__ cphi,
                                                                  const typename V2<T>::type* __restrict__ pi,
                                                                  const short2* __restrict__ cpi,
                                                                  const int32_t* __restrict__ cells, int64_t ns,
                                                                  double dspa, double* __restrict__ rec)
{
#pragma clang fp contract(off)
    const int64_t s = blockIdx.x * (int64_t)STRING_NT + threadIdx.x;
    if (s >= ns) return;
    const int64_t cell = cells[s], NN = (int64_t)N * N;
    const int c0[3] = {(int)(cell % N), (int)((cell / N) % N), (int)(cell / NN)};
    auto site = [&](int x, int y, int z) { return (int64_t)z * NN + (int64_t)y * N + x; };
    int npen = 0;
    double last[3] = {0, 0, 0};
    for (int pl = 0; pl < 3; pl++) {
        const int au = pl, av = (pl + 1) % 3, aw = (pl + 2) % 3;
        for (int k = 0; k < 2; k++) {
            // corner (i, j): x and y wrapped, z left at N for the halo plane
            auto corner = [&](int i, int j, int c[3]) {
#pragma unroll
                for (int a = 0; a < 3; a++) c[a] = c0[a] + (a == au ? i : 0) + (a == av ? j : 0) + (a == aw ? k : 0);
                if (c[0] >= N) c[0] -= N;
                if (c[1] >= N) c[1] -= N;
            };
            int mask = 0, n = 0;
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    int c[3];
                    corner(i, j, c);
                    const double2 f = ldf<T, COMP>(phi, cphi, site(c[0], c[1], c[2] == N ? 0 : c[2]));
                    const int m = (int)ceil(2 * atan2(f.y, f.x) / AXG_PI_T) + 2;
                    if (!((mask >> m) & 1)) {
                        mask |= 1 << m;
                        n++;
                    }
                }
            if (n < 3) continue;
            npen++;
            double2 vp[3] = {{0, 0}, {0, 0}, {0, 0}};
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    int c[3];
                    corner(i, j, c);
                    const double2 f = ldf<T, COMP>(phi, cphi, site(c[0], c[1], c[2] == N ? 0 : c[2]));
                    const double w[3] = {1.0, i - 0.5, j - 0.5};
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        vp[m].x = vp[m].x + w[m] * f.x;
                        vp[m].y = vp[m].y + w[m] * f.y;
                    }
                }
            const double d01 = 0.25 * vp[0].x, d02 = 0.25 * vp[0].y;
            const double d11 = vp[1].x, d12 = vp[1].y, d21 = vp[2].x, d22 = vp[2].y;
            const double det = d11 * d22 - d21 * d12;
            double du = -((d22 * d01 - d21 * d02) / det);
            double dv = -((-(d12 * d01) + d11 * d02) / det);
            du = fmax(fmin(du, 0.5), -0.5);
            dv = fmax(fmin(dv, 0.5), -0.5);
            double2 dp[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const double wt = (0.5 + (i ? fabs(du) : -fabs(du))) * (0.5 + (j ? fabs(dv) : -fabs(dv)));
                    int c[3];
                    corner(i, j, c);
                    const int cz = c[2] == N ? 0 : c[2];
                    const double2 q = ldf<T, COMP>(pi, cpi, site(c[0], c[1], cz));
                    dp[0].x = dp[0].x + wt * q.x;
                    dp[0].y = dp[0].y + wt * q.y;
                    const int xp = c[0] + 1 == N ? 0 : c[0] + 1, xm = c[0] == 0 ? N - 1 : c[0] - 1;
                    const int yp = c[1] + 1 == N ? 0 : c[1] + 1, ym = c[1] == 0 ? N - 1 : c[1] - 1;
                    double2 a = ldf<T, COMP>(phi, cphi, site(xp, c[1], cz)), b = ldf<T, COMP>(phi, cphi, site(xm, c[1], cz));
                    dp[1].x = dp[1].x + wt * (a.x - b.x) / 2;
                    dp[1].y = dp[1].y + wt * (a.y - b.y) / 2;
                    a = ldf<T, COMP>(phi, cphi, site(c[0], yp, cz));
                    b = ldf<T, COMP>(phi, cphi, site(c[0], ym, cz));
                    dp[2].x = dp[2].x + wt * (a.x - b.x) / 2;
                    dp[2].y = dp[2].y + wt * (a.y - b.y) / 2;
                    // z: central but one-sided from the halo plane (c[2] == N)
                    int zp, zm;
                    double dz;
                    if (c[2] == N) {
                        zp = 0;
                        zm = N - 1;
                        dz = 1;
                    } else {
                        zp = c[2] + 1 == N ? 0 : c[2] + 1;
                        zm = c[2] == 0 ? N - 1 : c[2] - 1;
                        dz = 2;
                    }
                    a = ldf<T, COMP>(phi, cphi, site(c[0], c[1], zp));
                    b = ldf<T, COMP>(phi, cphi, site(c[0], c[1], zm));
                    dp[3].x = dp[3].x + wt * (a.x - b.x) / dz;
                    dp[3].y = dp[3].y + wt * (a.y - b.y) / dz;
                }
            double a11 = 0, a12 = 0, a22 = 0;
#pragma unroll
            for (int m = 1; m <= 3; m++) {
                a11 = a11 + dp[m].x * dp[m].x;
                a12 = a12 + dp[m].x * dp[m].y;
                a22 = a22 + dp[m].y * dp[m].y;
            }
            const double dt = a11 * a22 - a12 * a12;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const double v = dp[0].x * (a22 * dp[m + 1].x - a12 * dp[m + 1].y) +
                                 dp[0].y * (a11 * dp[m + 1].y - a12 * dp[m + 1].x);
                last[m] = -(v / dt * dspa);
            }
        }
    }
    double vel[3];
    if (npen != 0) {
#pragma unroll
        for (int m = 0; m < 3; m++) {
            double t = 0;
            for (int i = 0; i < npen; i++) t = t + last[m];
            vel[m] = t / npen;
        }
    } else {
        vel[0] = vel[1] = vel[2] = 1.e-10;
    }
    double* o = rec + 11 * s;
    o[8] = vel[0];
    o[9] = vel[1];
    o[10] = vel[2];
}

// add_string_length's sums from the records, per block of STRING_NT strings in string order
// (the order of the sums does not depend on which stage ended a search).
__global__ void __launch_bounds__(STRING_NT) string_sum_kernel(const double* __restrict__ rec, int64_t ns,
                                                               double* __restrict__ partial)
{
#pragma clang fp contract(off)
    __shared__ double sh[5][STRING_NT];
    const int64_t s = blockIdx.x * (int64_t)STRING_NT + threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    if (s < ns) {
        const double* o = rec + 11 * s;
        const double vel[3] = {o[8], o[9], o[10]};
        acc[0] = o[7];
        double v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
        if (v2 < 1.0) {
            acc[1] = 1;
            acc[2] = sqrt(v2);
            acc[3] = v2;
            acc[4] = 1 / sqrt(1 - v2);
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < 5) {
        double t = 0;
        for (int j = 0; j < STRING_NT; j++) t += sh[threadIdx.x][j];
        partial[(int64_t)blockIdx.x * 5 + threadIdx.x] = t;
    }
}

}  // namespace

template <typename T, bool COMP>
void launch_string_fit(hipStream_t s, int N, const void* phi, const void* cphi, const void* pi, const void* cpi,
                       const int32_t* cells, int64_t ns, const Draws& dr, double dspa, int split, int vel_yy,
                       double* rec, double* partial, int* npend, int64_t* pend, double* state)
{
    using T2 = typename V2<T>::type;
    if (ns <= 0) return;
    if (ns > INT32_MAX) throw Error(AXG_EINVAL, "string fits: more than 2^31 - 1 string cells");
    const int64_t nblk = (ns + STRING_NT - 1) / STRING_NT;
    AXG_HIP(hipMemsetAsync(npend, 0, sizeof(int), s));
    hipLaunchKernelGGL((string_fit_kernel<T, COMP>), dim3((unsigned)nblk), dim3(STRING_NT), 0, s, N, (const T2*)phi,
                       (const short2*)cphi, (const T2*)pi, (const short2*)cpi, cells, ns, dr, dspa, split, rec,
                       npend, pend, state);
    AXG_HIP(hipGetLastError());
    // as many threads as strings; waves past the parked count return at once
    hipLaunchKernelGGL((string_fit_rest_kernel<T, COMP>), dim3((unsigned)((ns + STRING_NT2 - 1) / STRING_NT2)),
                       dim3(STRING_NT2), 0, s, N, (const T2*)phi, (const short2*)cphi, (const T2*)pi,
                       (const short2*)cpi, cells, dspa, rec, (const int*)npend, (const int64_t*)pend,
                       (const double*)state);
    AXG_HIP(hipGetLastError());
    if (vel_yy) {
        hipLaunchKernelGGL((string_vel_yy_kernel<T, COMP>), dim3((unsigned)nblk), dim3(STRING_NT), 0, s, N,
                           (const T2*)phi, (const short2*)cphi, (const T2*)pi, (const short2*)cpi, cells, ns, dspa, rec);
        AXG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(string_sum_kernel, dim3((unsigned)nblk), dim3(STRING_NT), 0, s, (const double*)rec, ns, partial);
    AXG_HIP(hipGetLastError());
}

#define AXG_STRING_FIT(T, C)                                                                                        \
    template void launch_string_fit<T, C>(hipStream_t, int, const void*, const void*, const void*, const void*,    \
                                          const int32_t*, int64_t, const Draws&, double, int, int, double*,         \
                                          double*, int*, int64_t*, double*);
AXG_STRING_FIT(double, false)
AXG_STRING_FIT(float, false)
AXG_STRING_FIT(float, true)
#undef AXG_STRING_FIT

}  // namespace axg
