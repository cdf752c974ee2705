/*
 * oracle.c -- CPU FP64 restatement of axion_grav's lattice time step and its
 * checkpoint diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it.  The
 * product (axion_grav_amd/) never links or calls it.
 *
 * It restates the reference's algorithm in the reference's own dataflow (unfused:
 * drift, kick, gradient, then 6 x [source -> 3D FFT -> accumulate]) and its FP64
 * operation order, compiled with -ffp-contract=off (the shipped reference build is
 * -O0, so it has no fused multiply-adds).  Each function cites the reference
 * file:line it follows.  The FFT is this file's own mixed-radix DFT (the reference
 * calls FFTW3-MPI through MKL, which is not available); it is pinned against a
 * brute-force DFT and numpy.fft in tests/test_oracle.py.
 *
 * Layout: a complex field is interleaved (re, im) doubles, N^3 sites, x fastest:
 * site (x, y, z) -> x + N*(y + N*z), exactly the Fortran complex(C_DOUBLE_COMPLEX)
 * array data1(1:N,1:N,1:N) without halo planes.
 *
 * Parity unpinned (no reference run, golden vector or closed form fixes them):
 *  - FFT rounding: orc_fft3d is pinned to the DFT (brute force, numpy), not to the
 *    unversioned FFTW/MKL cdft the reference links;
 *  - LAPACK rounding: orc_pp_inverse is pinned to numpy's inverse, not to MKL's
 *    dpptrf/dpptri;
 *  - the positions in the MCG59 stream at which the *_mcg functions place their
 *    draws (4 N^3 per thermal field, 4004 per string cell, 4 N^3 per covariance
 *    realisation) follow the reference's skip-ahead calls as read; the generator
 *    itself is pinned to MKL's own draws (tests/golden/mcg59_mkl.npz).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef struct { double re, im; } oc;

/* potential.f90:8 -- the truncated pi the reference uses everywhere */
#define PI_T 3.14159265359
#define TWOPI_T (2 * PI_T)
/* potential.f90:13-30 */
#define TCMB_IN_K 2.726
#define MSTAR_IN_GEV 2.435e18
#define GEV_IN_K 1.16045250061657e13
#define GEV_IN_HZ 1.5192488833521e24
#define BIGH 2.1332e-42
#define G_REL_NOWS 3.9091
#define N_AXION 6.68
#define CT_AXION 6.26
#define C0_AXION 1.0
#define THRM_SIGMA 2.449489742783178

static inline int64_t wrapn(int64_t i, int64_t n) { return (i < n / 2) ? i : i - n; }

/* ------------------------------------------------------------------ potential */

/* potential.f90:34-41 msquared */
double orc_msquared(double thrm) { return (thrm * thrm - THRM_SIGMA * THRM_SIGMA) / 3; }

/* potential.f90:62-79 maxion2 (kappa is a module variable there, an argument here) */
double orc_maxion2(double thrm, double kappa)
{
    double thrm12 = kappa * pow(CT_AXION / C0_AXION, 1.0 / N_AXION);
    double m;
    if (thrm > thrm12) m = CT_AXION * pow(kappa / thrm, N_AXION);
    else m = C0_AXION;
    double k2 = kappa * kappa;
    return m * (k2 * k2);
}

/* ------------------------------------------------------------------- settings */

/* settings.f90:274-297: derived units.  out = {a_now, tau_cr, sigma_cr, zeta, width_cr,
 * thickness_cr, dspa, dwav, dv_spa, dv_wav} */
void orc_derive(double sigma, double g_rel, double kappa, double tau_fin, double boxsize,
                int64_t ngrid, double* out)
{
    const double tcmb_in_mstar = TCMB_IN_K / GEV_IN_K / MSTAR_IN_GEV;
    double a_now = sigma * THRM_SIGMA / tcmb_in_mstar * pow(g_rel / G_REL_NOWS, 1.0 / 3);
    double tau_cr = 1 / PI_T / sqrt(g_rel / 90.0) / (sigma * sigma) / (THRM_SIGMA * THRM_SIGMA);
    double sigma_cr = tau_cr * sigma;
    double zeta = sqrt(22.5 / g_rel) / sigma / PI_T;
    double width_cr = 1 / sqrt(2.0) / sigma_cr;
    double thickness_cr = 1 / sqrt(2.0) / (sqrt(orc_maxion2(1.e-3 * kappa, kappa)) * sigma_cr);
    double dspa = boxsize * tau_fin / (double)ngrid;
    double dwav = TWOPI_T / dspa / (double)ngrid;
    double dv_spa = dspa * dspa * dspa;
    double r = dwav / TWOPI_T;
    double dv_wav = r * r * r;
    out[0] = a_now; out[1] = tau_cr; out[2] = sigma_cr; out[3] = zeta; out[4] = width_cr;
    out[5] = thickness_cr; out[6] = dspa; out[7] = dwav; out[8] = dv_spa; out[9] = dv_wav;
}

/* evolve.f90:10-24 get_dtau */
void orc_get_dtau(double tau1, double tau2, double sigma_cr, double dwav, int64_t ngrid,
                  double eps_dtau, double* dtau, int32_t* n)
{
    double omega = THRM_SIGMA * tau2 * sigma_cr;
    double kmax = dwav * (double)ngrid;
    omega = sqrt(kmax * kmax + omega * omega);
    double dt = eps_dtau / omega;
    int32_t nn = (int32_t)ceil((tau2 - tau1) / dt);
    *n = nn;
    *dtau = (tau2 - tau1) / nn;
}

/* ------------------------------------------------------------------ integrator */

/* evolve.f90:259-281 evolve_phi: phi <- phi + dt*pi over every site */
void orc_drift(double* phi, const double* pi, int64_t nsite, double dt)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nsite; i++) {
        phi[2 * i] = phi[2 * i] + dt * pi[2 * i];
        phi[2 * i + 1] = phi[2 * i + 1] + dt * pi[2 * i + 1];
    }
}

/* Coefficient of the N_DW=2 tilt V_t = -kappa0 t^4 (phi^2 + phi*^2)/2, which has no
 * counterpart in the reference.  The reference scales its potential so that sigma_cr^2 dV/dphi*
 * is the force (potential.f90:43-60 with evolve.f90:303: lambda = 4 sigma_cr^2), so the tilt
 * adds -kt conj(phi) to dVdphi and -kt (re^2 - im^2) to the energy's potential term, with
 * kt = kappa0 t^4 / sigma_cr^2.  kappa0 = 0 leaves the reference's arithmetic untouched. */
static double tilt_coef(double kappa0, double t, double sigma_cr)
{
    if (kappa0 == 0.0) return 0.0;
    double t2 = t * t;
    return kappa0 * (t2 * t2) / (sigma_cr * sigma_cr);
}

/* evolve.f90:283-335 evolve_phidot (without the GW call at 337), with potential.f90:43-60
 * dVdphi inlined.  Periodic in all three directions (update_margin's halo planes are the
 * periodic wrap on one rank). */
void orc_kick(const double* phi, double* pi, int64_t N, double t, double dt, double dspa,
              double sigma_cr, double kappa, int initial_loop, double kappa0)
{
    const double kt = tilt_coef(kappa0, t, sigma_cr);
    double hubble = 1 / t;
    double thrm = initial_loop ? 0.0 : THRM_SIGMA / t;
    double m2 = orc_msquared(thrm);
    double ma2 = orc_maxion2(thrm, kappa);
    double fact0 = 1 + dt * hubble;
    double fact1 = 1 - dt * hubble;
    double fact2 = dt / (dspa * dspa);
    double fact3 = -(dt * (t * t) * (sigma_cr * sigma_cr));
    fact1 = fact1 / fact0;
    fact2 = fact2 / fact0;
    fact3 = fact3 / fact0;
    const oc* f = (const oc*)phi;
    oc* p = (oc*)pi;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++) {
            int64_t z1 = (z + N - 1) % N, z2 = (z + 1) % N;
            int64_t y1 = (y + N - 1) % N, y2 = (y + 1) % N;
            for (int64_t x = 0; x < N; x++) {
                int64_t x1 = (x + N - 1) % N, x2 = (x + 1) % N;
                oc c = f[x + N * (y + N * z)];
                oc a1 = f[x1 + N * (y + N * z)], a2 = f[x2 + N * (y + N * z)];
                oc b1 = f[x + N * (y1 + N * z)], b2 = f[x + N * (y2 + N * z)];
                oc c1 = f[x + N * (y + N * z1)], c2 = f[x + N * (y + N * z2)];
                double lr = a1.re + a2.re + b1.re + b2.re + c1.re + c2.re - 6 * c.re;
                double li = a1.im + a2.im + b1.im + b2.im + c1.im + c2.im - 6 * c.im;
                double s2 = c.re * c.re + c.im * c.im;
                double g = 2 * s2 + m2;
                double dr = g * c.re - ma2 / 2;
                double di = g * c.im;
                if (kt != 0.0) {
                    dr = dr - kt * c.re;
                    di = di + kt * c.im;
                }
                oc* q = &p[x + N * (y + N * z)];
                q->re = fact1 * q->re + fact2 * lr + fact3 * dr;
                q->im = fact1 * q->im + fact2 * li + fact3 * di;
            }
        }
}

/* ------------------------------------------------------------------------ FFT */

/* e^{sign*2*pi*i*m/n}, m = 0..n-1, from the exact pi (the DFT itself does not use the
 * truncated constant: FFTW's twiddles are exact). */
static oc* make_twiddles(int64_t n, int sign)
{
    oc* tw = (oc*)malloc(sizeof(oc) * n);
    const long double pl = 3.141592653589793238462643383279502884L;
    for (int64_t m = 0; m < n; m++) {
        long double a = 2.0L * pl * (long double)m / (long double)n;
        tw[m].re = (double)cosl(a);
        tw[m].im = (double)(sign * sinl(a));
    }
    return tw;
}

static int64_t small_factor(int64_t n)
{
    if (n % 4 == 0 && n != 8) return 4;
    if (n % 2 == 0) return 2;
    for (int64_t p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* y[k] = sum_j x[j*xs] * w^{jk}, mixed radix, decimation in time.  tw is the table of the
 * top-level length ntop; tws = ntop/n. */
static void fft_rec(const oc* x, oc* y, int64_t n, int64_t xs, const oc* tw, int64_t ntop,
                    int64_t tws, oc* tmp)
{
    if (n == 1) { y[0] = x[0]; return; }
    int64_t p = small_factor(n), m = n / p;
    for (int64_t r = 0; r < p; r++) fft_rec(x + r * xs, y + r * m, m, xs * p, tw, ntop, tws * p, tmp);
    for (int64_t k = 0; k < m; k++) {
        for (int64_t r = 0; r < p; r++) tmp[r] = y[r * m + k];
        for (int64_t q = 0; q < p; q++) {
            double sr = 0, si = 0;
            int64_t kk = k + q * m;
            for (int64_t r = 0; r < p; r++) {
                oc w = tw[((r * kk) % n) * tws];
                sr += tmp[r].re * w.re - tmp[r].im * w.im;
                si += tmp[r].re * w.im + tmp[r].im * w.re;
            }
            y[kk].re = sr;
            y[kk].im = si;
        }
    }
}

/* Unnormalized 3D DFT out[k] = sum_x in[x] e^{sign*2*pi*i k.x/N}; sign=+1 is FFTW_BACKWARD
 * (grav_waves.f90:117-118, spectrum.f90:210-211).  in and out may alias. */
void orc_fft3d(const double* in, double* out, int64_t N, int sign)
{
    int64_t n3 = N * N * N;
    oc* o = (oc*)out;
    if (out != in) memcpy(out, in, sizeof(oc) * n3);
    oc* tw = make_twiddles(N, sign);
    for (int axis = 0; axis < 3; axis++) {
        int64_t es = axis == 0 ? 1 : (axis == 1 ? N : N * N);
#pragma omp parallel
        {
            oc* a = (oc*)malloc(sizeof(oc) * N);
            oc* b = (oc*)malloc(sizeof(oc) * N);
            oc* t = (oc*)malloc(sizeof(oc) * N);
#pragma omp for schedule(static)
            for (int64_t line = 0; line < N * N; line++) {
                int64_t base;
                if (axis == 0) base = line * N;
                else if (axis == 1) base = (line % N) + (line / N) * N * N;
                else base = line;
                for (int64_t i = 0; i < N; i++) a[i] = o[base + i * es];
                fft_rec(a, b, N, 1, tw, N, 1, t);
                for (int64_t i = 0; i < N; i++) o[base + i * es] = b[i];
            }
            free(a); free(b); free(t);
        }
    }
    free(tw);
}

/* ----------------------------------------------------------------- GW sourcing */

/* grav_waves.f90:157-191 get_comwav_grav for one wavevector index (0-based ix,iy,iz).
 * Returns q (= |n|*dwav) and, when i>0, pol = p_i*p_j (halved if i==j). */
static double comwav_grav(int64_t ix, int64_t iy, int64_t iz, int64_t N, double dwav, int i, int j,
                          oc* pol)
{
    double qh[3] = {(double)wrapn(ix, N), (double)wrapn(iy, N), (double)wrapn(iz, N)};
    double q = sqrt(qh[0] * qh[0] + qh[1] * qh[1] + qh[2] * qh[2]);
    if (i > 0) {
        if (q > 0) {
            qh[0] = qh[0] / q; qh[1] = qh[1] / q; qh[2] = qh[2] / q;
            double r = sqrt(qh[0] * qh[0] + qh[1] * qh[1]);
            double c, s;
            if (r > 0) { c = qh[0] / r; s = qh[1] / r; }
            else { c = 1; s = 0; }
            oc p[3] = {{qh[2] * c, -s}, {qh[2] * s, c}, {-r, 0}};
            oc a = p[i - 1], b = p[j - 1];
            pol->re = a.re * b.re - a.im * b.im;
            pol->im = a.re * b.im + a.im * b.re;
            if (i == j) { pol->re = pol->re / 2; pol->im = pol->im / 2; }
        } else {
            pol->re = 0; pol->im = 0;
        }
    }
    return q * dwav;
}

/* grav_waves.f90:323-356 data1_to_arr_dphi: central differences, dphi[d*N^3 + site] */
void orc_gradient(const double* phi, double* dphi, int64_t N, double dspa)
{
    const oc* f = (const oc*)phi;
    oc* g = (oc*)dphi;
    double d = dspa * 2;
    int64_t n3 = N * N * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++) {
            int64_t z1 = (z + N - 1) % N, z2 = (z + 1) % N;
            int64_t y1 = (y + N - 1) % N, y2 = (y + 1) % N;
            for (int64_t x = 0; x < N; x++) {
                int64_t x1 = (x + N - 1) % N, x2 = (x + 1) % N;
                int64_t s = x + N * (y + N * z);
                oc a = f[x2 + N * (y + N * z)], b = f[x1 + N * (y + N * z)];
                g[s].re = (a.re - b.re) / d; g[s].im = (a.im - b.im) / d;
                a = f[x + N * (y2 + N * z)]; b = f[x + N * (y1 + N * z)];
                g[n3 + s].re = (a.re - b.re) / d; g[n3 + s].im = (a.im - b.im) / d;
                a = f[x + N * (y + N * z2)]; b = f[x + N * (y + N * z1)];
                g[2 * n3 + s].re = (a.re - b.re) / d; g[2 * n3 + s].im = (a.im - b.im) / d;
            }
        }
}

/* grav_waves.f90:193-255 grav_wave_integration at the kick's time tau with step dtau.
 * A holds the four accumulators arr_grav(a,b,k) as planes m = (a-1) + 2*(b-1):
 * m=0 (1,1) cos,+   m=1 (2,1) sin,+   m=2 (1,2) cos,-   m=3 (2,2) sin,-   (full N^3 each).
 * fix_swap=0 reproduces lines 239-242 exactly (A(2,1) <- A(1,2)+..., A(1,2) <- new A(2,1)+...);
 * fix_swap=1 applies the evidently intended A(2,1) += s*P*D, A(1,2) += c*N*D. */
void orc_gw_accumulate(const double* dphi, double* A, int64_t N, double tau, double dtau, double dwav,
                       double dv_spa, double sigma_cr, int fix_swap);

void orc_gw_integrate(const double* phi, double* A, int64_t N, double tau, double dtau,
                      double dspa, double dwav, double dv_spa, double sigma_cr, int fix_swap)
{
    double* dphi = (double*)malloc(sizeof(double) * 6 * N * N * N);
    orc_gradient(phi, dphi, N, dspa);
    orc_gw_accumulate(dphi, A, N, tau, dtau, dwav, dv_spa, sigma_cr, fix_swap);
    free(dphi);
}

/* grav_waves.f90:209-247: the accumulation of grav_wave_integration from arr_dphi (dphi[d*N^3 +
 * site] on the GW grid of N sites per axis, wavenumber unit dwav and cell volume dv_spa). */
void orc_gw_accumulate(const double* dphi, double* A, int64_t N, double tau, double dtau, double dwav,
                       double dv_spa, double sigma_cr, int fix_swap)
{
    int64_t n3 = N * N * N;
    double* src = (double*)malloc(sizeof(double) * 2 * n3);
    const oc* g = (const oc*)dphi;
    oc* a = (oc*)A;
    double dw = tau * dtau * dv_spa * (sigma_cr * sigma_cr);
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= i; j++) {
            oc* s = (oc*)src;
            const oc* gi = g + (i - 1) * n3;
            const oc* gj = g + (j - 1) * n3;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < n3; k++) {
                /* dphi_i*conj(dphi_j) + conj(dphi_i)*dphi_j, as the complex products */
                double zr = gi[k].re * gj[k].re - gi[k].im * (-gj[k].im);
                double zi = gi[k].im * gj[k].re + gi[k].re * (-gj[k].im);
                double wr = gi[k].re * gj[k].re - (-gi[k].im) * gj[k].im;
                double wi = gi[k].re * gj[k].im + (-gi[k].im) * gj[k].re;
                s[k].re = zr + wr;
                s[k].im = zi + wi;
            }
            orc_fft3d(src, src, N, +1);
#pragma omp parallel for collapse(2) schedule(static)
            for (int64_t z = 0; z < N; z++)
                for (int64_t y = 0; y < N; y++)
                    for (int64_t x = 0; x < N; x++) {
                        int64_t k = x + N * (y + N * z);
                        oc polp = {0, 0};
                        double q = comwav_grav(x, y, z, N, dwav, i, j, &polp);
                        q = q * tau;
                        double c = cos(q), sn = sin(q);
                        polp.re = polp.re * dw; polp.im = polp.im * dw;
                        oc poln = {polp.re, -polp.im};
                        oc d = s[k];
                        oc pd = {polp.re * d.re - polp.im * d.im, polp.re * d.im + polp.im * d.re};
                        oc nd = {poln.re * d.re - poln.im * d.im, poln.re * d.im + poln.im * d.re};
                        oc* a11 = &a[k]; oc* a21 = &a[n3 + k]; oc* a12 = &a[2 * n3 + k]; oc* a22 = &a[3 * n3 + k];
                        a11->re = a11->re + c * pd.re; a11->im = a11->im + c * pd.im;
                        if (!fix_swap) {
                            a21->re = a12->re + sn * pd.re; a21->im = a12->im + sn * pd.im;
                            a12->re = a21->re + c * nd.re; a12->im = a21->im + c * nd.im;
                        } else {
                            a21->re = a21->re + sn * pd.re; a21->im = a21->im + sn * pd.im;
                            a12->re = a12->re + c * nd.re; a12->im = a12->im + c * nd.im;
                        }
                        a22->re = a22->re + sn * nd.re; a22->im = a22->im + sn * nd.im;
                    }
        }
    free(src);
}

/* ------------------------------------------------------ reduced grid (degrade_spec > 0) */

/* grav_waves.f90:384-407 data1_to_arr_dphi_degrade (one rank): on the grid Nc = N/deg, per axis
 * the sum over the block of deg^3 sites (z outermost, x innermost) of phi(s+e) - phi(s-e),
 * accumulated as (sum + a) - b from zero, divided by d = deg^3 * dspa * 2.  dphi: 3 x Nc^3. */
void orc_degrade_gradient(const double* phi, int64_t N, int64_t deg, double dspa, double* dphi)
{
    const oc* f = (const oc*)phi;
    oc* g = (oc*)dphi;
    const int64_t Nc = N / deg, nc3 = Nc * Nc * Nc;
    const double d = (double)(deg * deg * deg) * dspa * 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t cz = 0; cz < Nc; cz++)
        for (int64_t cy = 0; cy < Nc; cy++)
            for (int64_t cx = 0; cx < Nc; cx++) {
                oc s[3] = {{0, 0}, {0, 0}, {0, 0}};
                for (int64_t z = cz * deg; z < (cz + 1) * deg; z++) {
                    const int64_t z1 = (z + N - 1) % N, z2 = (z + 1) % N;
                    for (int64_t y = cy * deg; y < (cy + 1) * deg; y++) {
                        const int64_t y1 = (y + N - 1) % N, y2 = (y + 1) % N;
                        for (int64_t x = cx * deg; x < (cx + 1) * deg; x++) {
                            const int64_t x1 = (x + N - 1) % N, x2 = (x + 1) % N;
                            const oc ax = f[x2 + N * (y + N * z)], bx = f[x1 + N * (y + N * z)];
                            const oc ay = f[x + N * (y2 + N * z)], by = f[x + N * (y1 + N * z)];
                            const oc az = f[x + N * (y + N * z2)], bz = f[x + N * (y + N * z1)];
                            s[0].re = s[0].re + ax.re - bx.re; s[0].im = s[0].im + ax.im - bx.im;
                            s[1].re = s[1].re + ay.re - by.re; s[1].im = s[1].im + ay.im - by.im;
                            s[2].re = s[2].re + az.re - bz.re; s[2].im = s[2].im + az.im - bz.im;
                        }
                    }
                }
                const int64_t c = cx + Nc * (cy + Nc * cz);
                for (int k = 0; k < 3; k++) {
                    g[k * nc3 + c].re = s[k].re / d;
                    g[k * nc3 + c].im = s[k].im / d;
                }
            }
}

/* spectrum.f90:761-786 data2_to_data_spec_degrade (one rank): the block mean of f on the grid
 * N/deg, summed in loop order from zero and divided by deg^3. */
void orc_degrade_field(const double* f, int64_t N, int64_t deg, double* out)
{
    const oc* a = (const oc*)f;
    oc* o = (oc*)out;
    const int64_t Nc = N / deg;
    const double dd = (double)(deg * deg * deg);
    for (int64_t cz = 0; cz < Nc; cz++)
        for (int64_t cy = 0; cy < Nc; cy++)
            for (int64_t cx = 0; cx < Nc; cx++) {
                oc s = {0, 0};
                for (int64_t z = cz * deg; z < (cz + 1) * deg; z++)
                    for (int64_t y = cy * deg; y < (cy + 1) * deg; y++)
                        for (int64_t x = cx * deg; x < (cx + 1) * deg; x++) {
                            s.re = s.re + a[x + N * (y + N * z)].re;
                            s.im = s.im + a[x + N * (y + N * z)].im;
                        }
                o[cx + Nc * (cy + Nc * cz)].re = s.re / dd;
                o[cx + Nc * (cy + Nc * cz)].im = s.im / dd;
            }
}

/* spectrum.f90:900-954 masking with deg > 1: the mask file's cells (ix0,iy0,iz0) zero the coarse
 * cell ((ix0-1)/deg+1, ...); here the coarse flag gets bit 4 from any fine cell's bit 4. */
void orc_degrade_mask(const int8_t* flags, int64_t N, int64_t deg, int8_t* out)
{
    const int64_t Nc = N / deg;
    memset(out, 0, (size_t)(Nc * Nc * Nc));
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++)
                if (flags[x + N * (y + N * z)] & 4) out[x / deg + Nc * (y / deg + Nc * (z / deg))] = 4;
}

/* evolve.f90:238-257 symplectic_integration over [t1, t2] with nt steps of dt (from get_dtau
 * or a fixed step): drift dt/2, then nt x (kick at the midpoint [+GW], drift dt), the last
 * drift dt/2.  A may be NULL when grav_spec is off. */
void orc_symplectic(double* phi, double* pi, double* A, int64_t N, double t1, double dt, int32_t nt,
                    double dspa, double dwav, double dv_spa, double sigma_cr, double kappa,
                    int initial_loop, int fix_swap, double kappa0, int degrade_spec)
{
    int64_t n3 = N * N * N;
    double t = t1;
    /* init_grav (grav_waves.f90:37-45): the GW grid of degrade_spec > 0 */
    const int64_t deg = (int64_t)1 << degrade_spec, Ng = N / deg;
    const double dspa_g = dspa * (double)deg, dwav_g = TWOPI_T / dspa_g / (double)Ng;
    const double dv_spa_g = dspa_g * dspa_g * dspa_g;
    double* dphi = (A && deg > 1) ? (double*)malloc(sizeof(double) * 6 * Ng * Ng * Ng) : NULL;
    for (int32_t it = 1; it <= nt; it++) {
        if (it == 1) { orc_drift(phi, pi, n3, dt / 2); t = t + dt / 2; }
        orc_kick(phi, pi, N, t, dt, dspa, sigma_cr, kappa, initial_loop, kappa0);
        if (A && deg > 1) {
            orc_degrade_gradient(phi, N, deg, dspa, dphi);
            orc_gw_accumulate(dphi, A, Ng, t, dt, dwav_g, dv_spa_g, sigma_cr, fix_swap);
        } else if (A) {
            orc_gw_integrate(phi, A, N, t, dt, dspa, dwav, dv_spa, sigma_cr, fix_swap);
        }
        if (it == nt) { orc_drift(phi, pi, n3, dt / 2); t = t + dt / 2; }
        else { orc_drift(phi, pi, n3, dt); t = t + dt; }
    }
    free(dphi);
}

/* ---------------------------------------------------------------- GW spectrum */

/* grav_waves.f90:62-71 (bin width), 257-305 grav_spec_estimation: numer[b], denom[b] over the
 * full k-space for q>0, b = nint(q/dq + 0.5) in 1..nb (returned 0-based), and the power
 * P[b] = numer/denom * qmid^3 * fact (fact from line 267, qmid from the cumulative edges). */
void orc_gw_spectrum(const double* A, int64_t N, int nb, double dspa, double dwav, double tau_cr,
                     double a_now, double dv_wav, double* numer, double* denom, double* qmid,
                     double* power)
{
    int64_t n3 = N * N * N;
    const oc* a = (const oc*)A;
    double qmax = PI_T / dspa;
    double dq0 = qmax / nb;
    double* wave = (double*)calloc(nb + 2, sizeof(double));
    wave[0] = 0;
    for (int b = 1; b <= nb; b++) wave[b] = wave[b - 1] + dq0;
    double dq = wave[1] - wave[0];
    for (int b = 0; b < nb; b++) { numer[b] = 0; denom[b] = 0; }
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                double q = comwav_grav(x, y, z, N, dwav, 0, 0, NULL);
                long iq = lround(q / dq + 0.5);
                if (q > 0 && iq >= 1 && iq <= nb) {
                    int64_t k = x + N * (y + N * z);
                    denom[iq - 1] = denom[iq - 1] + 1;
                    double s = 0;
                    for (int m = 0; m < 4; m++) {
                        oc v = a[m * n3 + k];
                        /* abs(z)**2 in the reference: hypot squared; |z|^2 = re^2+im^2 */
                        s += v.re * v.re + v.im * v.im;
                    }
                    numer[iq - 1] += s;
                }
            }
    const double bigh_in_mstar = BIGH / MSTAR_IN_GEV;
    double t6 = tau_cr * tau_cr * tau_cr; t6 = t6 * t6;
    double a4 = a_now * a_now; a4 = a4 * a4;
    double fact = 1 / t6 / (bigh_in_mstar * bigh_in_mstar) / a4 / 3 / (TWOPI_T * TWOPI_T) * dv_wav;
    for (int b = 1; b <= nb; b++) {
        double qm = (wave[b] + wave[b - 1]) / 2;
        qmid[b - 1] = qm;
        power[b - 1] = denom[b - 1] > 0 ? numer[b - 1] / denom[b - 1] * (qm * qm * qm) * fact : 0.0;
    }
    free(wave);
}

/* --------------------------------------------------------------------- defects */

/* defects_3d.f90:309-349 id_with_phase for one cell's 8 phases.  Returns the flag bits
 * (string 1, wall 2, mask 4). */
static int id_with_phase(const double* ph, int id_strings, int id_walls, int spec_est)
{
    double p[8];
    for (int i = 0; i < 8; i++) p[i] = ph[i];
    for (int i = 7; i >= 1; i--)
        for (int j = 0; j < i; j++)
            if (p[j] > p[j + 1]) { double t = p[j]; p[j] = p[j + 1]; p[j + 1] = t; }
    int flag = 0;
    double t = p[0] - p[7] + TWOPI_T;
    double tmax = p[1] - p[0];
    for (int i = 2; i < 8; i++) { double d = p[i] - p[i - 1]; if (d > tmax) tmax = d; }
    if (id_walls && tmax > t) flag += 2;
    else tmax = tmax > t ? tmax : t;
    if (id_strings && tmax < PI_T) flag += 1;
    if (spec_est && tmax < 7 * PI_T / 4) flag += 4;
    return flag;
}

/* defects_3d.f90:67-103 classification pass: for every cell (x,y,z) with corners
 * (x+{0,1}, y+{0,1}, z+{0,1}) (periodic), ph = aimag(log(phi)) = atan2(im, re).
 * flags[cell] gets the bits; counts = {strings, walls, masks}. */
void orc_classify(const double* phi, int64_t N, int id_strings, int id_walls, int spec_est,
                  int8_t* flags, int64_t* counts)
{
    const oc* f = (const oc*)phi;
    int64_t ns = 0, nw = 0, nm = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : ns, nw, nm)
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++) {
            int64_t z2 = (z + 1) % N, y2 = (y + 1) % N;
            for (int64_t x = 0; x < N; x++) {
                int64_t x2 = (x + 1) % N;
                int64_t xs[2] = {x, x2}, ys[2] = {y, y2}, zs[2] = {z, z2};
                double ph[8];
                for (int i = 0; i < 8; i++) {
                    oc v = f[xs[i & 1] + N * (ys[(i >> 1) & 1] + N * zs[i >> 2])];
                    ph[i] = atan2(v.im, v.re);
                }
                int fl = id_with_phase(ph, id_strings, id_walls, spec_est);
                flags[x + N * (y + N * z)] = (int8_t)fl;
                ns += fl & 1;
                nw += (fl >> 1) & 1;
                nm += (fl >> 2) & 1;
            }
        }
    counts[0] = ns; counts[1] = nw; counts[2] = nm;
}

/* ---------------------------------------------------------- plaquette windings */

/* Plaquette windings (a new diagnostic; the reference's plaquette test is the quadrant count of
 * velocity_string_YY, defects_3d.f90:850-879).  Face o of site s spans (u, v) = (x, y), (y, z),
 * (z, x) (ipla = o + 1, the cshift order of defects_3d.f90:854); loop s, s+u, s+u+v, s+v.  Edge
 * a -> b adds +1 when the shorter arc from arg a to arg b crosses arg = pi counterclockwise (a on
 * the closed upper side Im >= 0, b below, cross product Re a Im b - Im a Re b > 0), -1 when
 * clockwise, 0 when the cross product vanishes (antipodal or zero corners). */
static int arc_cut(oc a, oc b)
{
    int ua = !(a.im < 0), ub = !(b.im < 0);
    double c = a.re * b.im - a.im * b.re;
    if (ua && !ub && c > 0) return 1;
    if (!ua && ub && c < 0) return -1;
    return 0;
}

/* defects_3d.f90:866-873: quadrant m = ceiling(2 ph / pi) + 2 of ph = aimag(log(z)) */
static int quadrant_of(oc z)
{
    double ph = atan2(z.im, z.re);
    return (int)ceil(2 * ph / PI_T) + 2;
}

/* code[site] = sum_o (n_o + 2) << 3o | quad_o << (9 + o); counts[5 o + k] = pierced, +1, -1,
 * quadrant-test hits, net winding of face o; counts[15] = all pierced faces. */
void orc_plaquettes(const double* phi, int64_t N, uint16_t* code, int64_t* counts)
{
    const oc* f = (const oc*)phi;
    int64_t c[16] = {0};
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : c[:16])
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                int64_t p[3] = {x, y, z};
                int cd = 0;
                for (int o = 0; o < 3; o++) {
                    int u = o, v = (o + 1) % 3;
                    int64_t q[4][3];
                    for (int k = 0; k < 4; k++) {
                        q[k][0] = p[0]; q[k][1] = p[1]; q[k][2] = p[2];
                    }
                    q[1][u] = (p[u] + 1) % N;
                    q[2][u] = (p[u] + 1) % N;
                    q[2][v] = (p[v] + 1) % N;
                    q[3][v] = (p[v] + 1) % N;
                    oc w[4];
                    for (int k = 0; k < 4; k++) w[k] = f[q[k][0] + N * (q[k][1] + N * q[k][2])];
                    int n = 0, mask = 0;
                    for (int k = 0; k < 4; k++) {
                        n += arc_cut(w[k], w[(k + 1) % 4]);
                        mask |= 1 << quadrant_of(w[k]);
                    }
                    int nq = 0;
                    for (int m = 1; m <= 4; m++) nq += (mask >> m) & 1;
                    int quad = nq >= 3;
                    cd |= (n + 2) << (3 * o);
                    cd |= quad << (9 + o);
                    if (n != 0) { c[5 * o] += 1; c[5 * o + 4] += n; c[15] += 1; }
                    if (n == 1) c[5 * o + 1] += 1;
                    if (n == -1) c[5 * o + 2] += 1;
                    c[5 * o + 3] += quad;
                }
                code[x + N * (y + N * z)] = (uint16_t)cd;
            }
    for (int k = 0; k < 16; k++) counts[k] = c[k];
}

/* ---------------------------------------------------------- field (axion) spectrum */

/* spectrum.f90:466-567 compute_conv_spec (+ masking 900-954, degrade_spec=0 copy 736-792,
 * get_comwav_spec 233-240): pi with the mask cells (flag bit 4) zeroed -> BACKWARD FFT ->
 * *dv_spa -> |.|^2*dv_wav binned with i = nint(q/dq+0.5), q=0 included (bin 1). */
void orc_conv_spectrum(const double* pi, const int8_t* flags, int64_t N, int nb, double dspa,
                       double dwav, double dv_spa, double dv_wav, double* pow_out, double* num)
{
    int64_t n3 = N * N * N;
    double* w = (double*)malloc(sizeof(double) * 2 * n3);
    memcpy(w, pi, sizeof(double) * 2 * n3);
    if (flags)
        for (int64_t k = 0; k < n3; k++)
            if (flags[k] & 4) { w[2 * k] = 0; w[2 * k + 1] = 0; }
    orc_fft3d(w, w, N, +1);
    for (int64_t k = 0; k < 2 * n3; k++) w[k] = w[k] * dv_spa;
    double qmax = PI_T / dspa, dq0 = qmax / nb;
    double* wave = (double*)calloc(nb + 2, sizeof(double));
    wave[0] = 0;
    for (int b = 1; b <= nb; b++) wave[b] = wave[b - 1] + dq0;
    double dq = wave[1] - wave[0];
    for (int b = 0; b < nb; b++) { pow_out[b] = 0; num[b] = 0; }
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                int64_t nx = wrapn(x, N), ny = wrapn(y, N), nz = wrapn(z, N);
                double q = (double)(nx * nx) + (double)(ny * ny) + (double)(nz * nz);
                q = sqrt(q) * dwav;
                long i = lround(q / dq + 0.5);
                if (i >= 1 && i <= nb) {
                    int64_t k = x + N * (y + N * z);
                    double re = w[2 * k], im = w[2 * k + 1];
                    pow_out[i - 1] = pow_out[i - 1] + (re * re + im * im) * dv_wav;
                    num[i - 1] = num[i - 1] + 1.0;
                }
            }
    free(w);
    free(wave);
}

/* spectrum.f90:117-171 integrate_p_check (3D branch) */
double orc_integrate_p_check(int nint_, double li, double si, double lj, double sj, double q)
{
    double maxp = lj < q + li ? lj : q + li;
    double minp = (q - li) > (si - q) ? (q - li) : (si - q);
    minp = minp > sj ? minp : sj;
    double dp = (maxp - minp) / nint_;
    double x = 0;
    if (dp > 0) {
        for (int ip = 1; ip <= nint_; ip++) {
            double p = dp * (ip - 0.5) + minp;
            double maxmu, minmu;
            if (q == 0.0) { maxmu = 1; minmu = -1; }
            else {
                maxmu = (li * li - p * p - q * q) / 2 / p / q;
                minmu = (si * si - p * p - q * q) / 2 / p / q;
                if (maxmu > 1) maxmu = 1;
                if (minmu < -1) minmu = -1;
            }
            x = x + p * p * dp * (maxmu - minmu);
        }
    }
    return x * TWOPI_T;
}

/* spectrum.f90:23-103 init_spectrum: bin edges, and for every packed pair ij (1-based,
 * i = nint(sqrt(2 ij)), j = ij - i(i-1)/2) the 501-point table of shell-overlap integrals.
 * largeq/mat: (num_largeq+1) x num_cov, column-major per pair; dqs: num_cov. */
void orc_init_spectrum(int nb, double dspa, int num_largeq, double* wav, double* largeq,
                       double* mat, double* dqs)
{
    double qmax = PI_T / dspa, dq = qmax / nb;
    wav[0] = 0;
    for (int b = 1; b <= nb; b++) wav[b] = wav[b - 1] + dq;
    int num_cov = nb * (nb + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ij = 1; ij <= num_cov; ij++) {
        int i = (int)lround(sqrt(2.0 * ij));
        int j = ij - i * (i - 1) / 2;
        double lj = wav[j], sj = wav[j - 1], li = wav[i], si = wav[i - 1];
        double maxq = li + lj;
        double mn = li < lj ? li : lj;
        double mx = si > sj ? si : sj;
        double minq = mx - mn;
        if (minq < 0) minq = 0;
        double d = maxq - minq;
        d = d / num_largeq;
        dqs[ij - 1] = d;
        for (int iq = 0; iq <= num_largeq; iq++) {
            double q = minq + d * iq;
            largeq[(int64_t)(ij - 1) * (num_largeq + 1) + iq] = q;
            mat[(int64_t)(ij - 1) * (num_largeq + 1) + iq] = orc_integrate_p_check(200, li, si, lj, sj, q);
        }
    }
}

/* spectrum.f90:446-464 get_mat_spec */
static double get_mat_spec(const double* largeq, const double* mat, const double* dqs, int nlq, int ij0,
                           double q)
{
    const double* lq = largeq + (int64_t)ij0 * (nlq + 1);
    const double* ms = mat + (int64_t)ij0 * (nlq + 1);
    double d = (q - lq[0]) / dqs[ij0];
    double fl = floor(d);
    if (fl < nlq && fl >= 0) {
        int iq = (int)fl;
        d = d - iq;
        return (1 - d) * ms[iq] + d * ms[iq + 1];
    }
    return 0.0;
}

/* Packed-upper Cholesky inverse (the LAPACK dpptrf('U') + dpptri('U') pair used at
 * spectrum.f90:437-438): ap holds A(r,c), r<=c (1-based) at r + c(c-1)/2 - 1; returns 0 on
 * success, the failing column otherwise. */
int orc_pp_inverse(int n, double* ap)
{
#define AP(r, c) ap[(r) + (int64_t)(c) * ((c) - 1) / 2 - 1]
    /* A = U^T U */
    for (int j = 1; j <= n; j++) {
        for (int i = 1; i < j; i++) {
            double s = AP(i, j);
            for (int k = 1; k < i; k++) s -= AP(k, i) * AP(k, j);
            AP(i, j) = s / AP(i, i);
        }
        double s = AP(j, j);
        for (int k = 1; k < j; k++) s -= AP(k, j) * AP(k, j);
        if (!(s > 0)) return j;
        AP(j, j) = sqrt(s);
    }
    /* U <- inv(U) (upper triangular), column by column: the leading (j-1) block already
     * holds inv(U); column j = -inv(U)(1:j-1,1:j-1) * U(1:j-1,j) / U(j,j). */
    double* col = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int j = 1; j <= n; j++) {
        AP(j, j) = 1 / AP(j, j);
        double ajj = -AP(j, j);
        for (int i = 1; i < j; i++) {
            double s = 0;
            for (int k = i; k < j; k++) s += AP(i, k) * AP(k, j);
            col[i - 1] = s * ajj;
        }
        for (int i = 1; i < j; i++) AP(i, j) = col[i - 1];
    }
    free(col);
    /* inv(A) = inv(U) inv(U)^T, upper part */
    for (int i = 1; i <= n; i++)
        for (int j = i; j <= n; j++) {
            double s = 0;
            for (int k = j; k <= n; k++) s += AP(i, k) * AP(j, k);
            AP(i, j) = s;
        }
#undef AP
    return 0;
}

/* spectrum.f90:344-444 compute_matrix_deconv: window W = BACKWARD-FFT(dv_spa with mask cells
 * zeroed); mat(ij) = sum_k |W|^2 u_ij(|k|) / sum_k u_ij(|k|) * dv_wav^2 over each pair's
 * shell; returns mat (before inversion) and matinv (packed-upper inverse). */
int orc_matrix_deconv(const int8_t* flags, int64_t N, int nb, double dwav, double dv_spa,
                      double dv_wav, const double* largeq, const double* matspec, const double* dqs,
                      int nlq, double* mat, double* matinv)
{
    int64_t n3 = N * N * N;
    int num_cov = nb * (nb + 1) / 2;
    double* w = (double*)malloc(sizeof(double) * 2 * n3);
    for (int64_t k = 0; k < n3; k++) {
        w[2 * k] = dv_spa; w[2 * k + 1] = 0;
        if (flags && (flags[k] & 4)) w[2 * k] = 0;
    }
    orc_fft3d(w, w, N, +1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ij = 0; ij < num_cov; ij++) {
        double ql = largeq[(int64_t)ij * (nlq + 1) + nlq], q0 = largeq[(int64_t)ij * (nlq + 1)];
        double t1 = ql / dwav, t0 = q0 / dwav;
        int64_t imax2 = (int64_t)ceil(t1 * t1) - 1;
        int64_t imin2 = (int64_t)ceil(t0 * t0);
        double m0 = 0, t = 0;
        for (int64_t z = 0; z < N; z++) {
            int64_t iqz2 = wrapn(z, N) * wrapn(z, N);
            if (iqz2 > imax2) continue;
            for (int64_t y = 0; y < N; y++) {
                int64_t iqy2 = wrapn(y, N) * wrapn(y, N);
                if (iqz2 + iqy2 > imax2) continue;
                for (int64_t x = 0; x < N; x++) {
                    int64_t iq2 = wrapn(x, N) * wrapn(x, N) + iqy2 + iqz2;
                    if (!(iq2 <= imax2 && iq2 >= imin2)) continue;
                    double q = sqrt((double)iq2) * dwav;
                    double u = get_mat_spec(largeq, matspec, dqs, nlq, ij, q);
                    int64_t k = x + N * (y + N * z);
                    double v = w[2 * k] * w[2 * k] + w[2 * k + 1] * w[2 * k + 1];
                    m0 = m0 + v * u;
                    t = t + u;
                }
            }
        }
        mat[ij] = m0 / t * dv_wav * dv_wav;
    }
    free(w);
    memcpy(matinv, mat, sizeof(double) * num_cov);
    return orc_pp_inverse(nb, matinv);
}

/* ------------------------------------------------------------------ energy */

/* Comoving energy of the lattice (a diagnostic the reference does not have): kinetic
 * sum|pi|^2, gradient sum_d |phi(x+e_d)-phi(x)|^2/dspa^2 (the form whose variation is the
 * 7-point Laplacian), potential tau^2 sigma_cr^2 sum V with V = (|phi|^2+m2/2)^2 - ma2*Re(phi)
 * (dV/dphi* is potential.f90:43-60's dVdphi), each times dv_spa; out[3] = total. */
void orc_energy(const double* phi, const double* pi, int64_t N, double t, double dspa,
                double sigma_cr, double kappa, int initial_loop, double kappa0, double* out)
{
    const double kt = tilt_coef(kappa0, t, sigma_cr);
    const oc* f = (const oc*)phi;
    const oc* p = (const oc*)pi;
    double thrm = initial_loop ? 0.0 : THRM_SIGMA / t;
    double m2 = orc_msquared(thrm), ma2 = orc_maxion2(thrm, kappa);
    double ek = 0, eg = 0, ev = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : ek, eg, ev)
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                int64_t k = x + N * (y + N * z);
                oc c = f[k];
                ek += p[k].re * p[k].re + p[k].im * p[k].im;
                oc nb3[3] = {f[(x + 1) % N + N * (y + N * z)], f[x + N * ((y + 1) % N + N * z)],
                             f[x + N * (y + N * ((z + 1) % N))]};
                for (int d = 0; d < 3; d++) {
                    double dr = nb3[d].re - c.re, di = nb3[d].im - c.im;
                    eg += dr * dr + di * di;
                }
                double s2 = c.re * c.re + c.im * c.im;
                double h = s2 + m2 / 2;
                double v = h * h - ma2 * c.re;
                if (kt != 0.0) v = v - kt * (c.re * c.re - c.im * c.im);
                ev += v;
            }
    double dv = dspa * dspa * dspa;
    out[0] = ek * dv;
    out[1] = eg / (dspa * dspa) * dv;
    out[2] = t * t * sigma_cr * sigma_cr * ev * dv;
    out[3] = out[0] + out[1] + out[2];
}

/* ------------------------------------------------------------ initial condition */

/* Philox4x32-10 (Salmon et al., SC'11), the counter-based generator used for the seeded
 * initial field so that every backend draws the same numbers for the same mode. */
static void philox4x32_10(uint32_t ctr[4], uint32_t key0, uint32_t key1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key0, k1 = key1;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
        uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        c0 = h1 ^ c1 ^ k0; c1 = l1; c2 = h0 ^ c3 ^ k1; c3 = l0;
        k0 += W0; k1 += W1;
    }
    ctr[0] = c0; ctr[1] = c1; ctr[2] = c2; ctr[3] = c3;
}

void orc_philox(uint32_t* ctr, uint32_t k0, uint32_t k1) { philox4x32_10(ctr, k0, k1); }

/* Mode amplitude of the seeded flat-spectrum initial field for the wavevector index
 * (0-based ix,iy,iz): complex Gaussian (Box-Muller on two 53-bit uniforms in (0,1]). */
void orc_ic_mode(uint64_t seed, int64_t N, int64_t ix, int64_t iy, int64_t iz, double* re, double* im)
{
    uint64_t lin = (uint64_t)ix + (uint64_t)N * ((uint64_t)iy + (uint64_t)N * (uint64_t)iz);
    uint32_t c[4] = {(uint32_t)lin, (uint32_t)(lin >> 32), 0x6178696fu, 0x6e677276u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
    uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
    a &= (1ull << 53) - 1; b &= (1ull << 53) - 1;
    double u1 = ((double)a + 1.0) / 9007199254740992.0;
    double u2 = ((double)b + 1.0) / 9007199254740992.0;
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    *re = r * cos(th);
    *im = r * sin(th);
}

/* Seeded initial field: phi = BACKWARD-FFT of complex Gaussian modes for |k| < kmax
 * (|k| = sqrt(n^2)*dwav), zero above, rescaled so that <|phi|^2> = 1; pi = 0. */
void orc_ic_flat(double* phi, double* pi, int64_t N, uint64_t seed, double kmax, double dwav)
{
    int64_t n3 = N * N * N;
    memset(phi, 0, sizeof(double) * 2 * n3);
    memset(pi, 0, sizeof(double) * 2 * n3);
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                int64_t nx = wrapn(x, N), ny = wrapn(y, N), nz = wrapn(z, N);
                double q = sqrt((double)(nx * nx + ny * ny + nz * nz)) * dwav;
                if (q < kmax) {
                    int64_t k = x + N * (y + N * z);
                    orc_ic_mode(seed, N, x, y, z, &phi[2 * k], &phi[2 * k + 1]);
                }
            }
    orc_fft3d(phi, phi, N, +1);
    double s = 0;
    for (int64_t k = 0; k < 2 * n3; k++) s += phi[k] * phi[k];
    double scale = 1.0 / sqrt(s / (double)n3);
    for (int64_t k = 0; k < 2 * n3; k++) phi[k] = phi[k] * scale;
}

/* ------------------------------------------------------------ string fit and length */

#define STR_NSIMP 4      /* num_simp_string_pos (defects_3d.f90:26) */
#define STR_MAXIT 1000   /* max_iter_string_pos */

/* defects_3d.f90:384-433 calc_chi2_string: variance over the 8 corners of the phase minus the
 * phase of a straight string with direction (sqrt(1-mu^2) cos phi, sqrt(1-mu^2) sin phi, mu)
 * at distance rho, angle theta from the cell centre; p = (mu, phi, rho, theta), corner
 * i = (iz*2+iy)*2+ix at (ix,iy,iz) - 1/2. */
double orc_chi2_string(const double* ph, const double* p)
{
    double mu = p[0], phi = p[1], u0 = p[2] * cos(p[3]), v0 = p[2] * sin(p[3]);
    double cm = mu, sm = sqrt(1.0 - mu * mu), cp = cos(phi), sp = sin(phi);
    double a = 0, b = 0;
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                double x = ix - 0.5, y = iy - 0.5, z = iz - 0.5;
                int i = (iz * 2 + iy) * 2 + ix;
                double du = -u0 + cm * (x * cp + y * sp) - z * sm;
                double dv = -v0 - x * sp + y * cp;
                double d = ph[i] - atan2(dv, du);
                if (d >= PI_T) d = d - TWOPI_T;
                else if (d < -PI_T) d = d + TWOPI_T;
                a = a + d * d;
                b = b + d;
            }
    a = a / 8;
    b = b / 8;
    return a - b * b;
}

/* defects_3d.f90:1027-1056 string_length: the chord of the line pos + t dir through the cell
 * [-1/2, 1/2]^3 (zero when an exit lies beyond sqrt(3)/2). */
double orc_string_length(const double* pos, const double* dir)
{
    const double atmax = 0.867, atmax0 = 0.86602540378378444;
    double tp = atmax, tn = -atmax;
    for (int i = 0; i < 3; i++)
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

/* ------------------------------------------------ the reference's random stream (MCG59) */

/* The reference draws every random number from one MKL VSL stream of the MCG59 generator
 * (subroutines.f90:16, 49-159; Intel MKL VS notes): x_{n+1} = a x_n mod 2^59 with a = 13^13,
 * uniform u_n = x_n / 2^59 on [0, 1); vslNewStreamEx(MCG59, 2, {w0, w1}) starts at
 * x_0 = (w0 + 2^32 w1) mod 2^59 (1 if that is 0), and u_0 is the first draw.  VSL's
 * BOXMULLER Gaussian n uses the pair (u_{2n}, u_{2n+1}): sqrt(-2 ln u_{2n}) sin(2 pi u_{2n+1}).
 * Position p of the stream is x_0 a^p mod 2^59; as 2^59 divides 2^64, wrapping 64-bit products
 * reduced to 59 bits at the end give it exactly. */
#define MCG59_A 302875106592253ull /* 13^13 */
#define MCG59_MASK ((1ull << 59) - 1)

uint64_t orc_mcg59_seed(uint32_t w0, uint32_t w1)
{
    uint64_t x = ((uint64_t)w0 + ((uint64_t)w1 << 32)) & MCG59_MASK;
    return x ? x : 1;
}

uint64_t orc_mcg59_jump(uint64_t x0, uint64_t p)
{
    uint64_t a = MCG59_A, r = 1;
    while (p) {
        if (p & 1) r *= a;
        a *= a;
        p >>= 1;
    }
    return (r * x0) & MCG59_MASK;
}

static double mcg59_u(uint64_t x) { return (double)x * 1.7347234759768071e-18; /* 2^-59 */ }
static uint64_t mcg59_next(uint64_t x) { return (x * MCG59_A) & MCG59_MASK; }

/* Box-Muller on the next two draws of the stream at x (advanced by two). */
static double mcg59_gauss(uint64_t* x)
{
    double u1 = mcg59_u(*x);
    *x = mcg59_next(*x);
    double u2 = mcg59_u(*x);
    *x = mcg59_next(*x);
    return sqrt(-2.0 * log(u1)) * sin(6.283185307179586 * u2);
}

/* n uniforms / n Gaussians of the stream (w0, w1) from position pos (test vectors). */
void orc_mcg59_uniforms(uint32_t w0, uint32_t w1, uint64_t pos, int64_t n, double* out)
{
    uint64_t x = orc_mcg59_jump(orc_mcg59_seed(w0, w1), pos);
    for (int64_t i = 0; i < n; i++) {
        out[i] = mcg59_u(x);
        x = mcg59_next(x);
    }
}

void orc_mcg59_gauss(uint32_t w0, uint32_t w1, uint64_t pos, int64_t n, double* out)
{
    uint64_t x = orc_mcg59_jump(orc_mcg59_seed(w0, w1), pos);
    for (int64_t i = 0; i < n; i++) out[i] = mcg59_gauss(&x);
}

/* Where a routine's random numbers come from: the reference's stream (on != 0: x0 and the
 * position of the routine's first draw) or, when on = 0 or the pointer is NULL, the build's
 * own counter-based Philox streams keyed by the routine's seed. */
typedef struct {
    int on;
    uint64_t x0, pos;
} orc_stream;

/* The uniforms of start point i of string s.  The reference's stream: draws pos + 4004 s + 4 i
 * .. +3 (defects_3d.f90:137-162, 465: each string cell takes mult_string_pos = 4 x 1001 draws,
 * in site order).  Otherwise one Philox4x32-10 block keyed by the seed, counter (i, s, 'strg'):
 * the same "stream per string" structure, independent of the thread count. */
static void string_uniforms(uint32_t key0, uint32_t key1, const orc_stream* st, int64_t s, int i, double* u)
{
    if (st && st->on) {
        uint64_t x = orc_mcg59_jump(st->x0, st->pos + 4004 * (uint64_t)s + 4 * (uint64_t)i);
        for (int k = 0; k < 4; k++) {
            u[k] = mcg59_u(x);
            x = mcg59_next(x);
        }
        return;
    }
    uint32_t c[4] = {(uint32_t)i, (uint32_t)s, (uint32_t)((uint64_t)s >> 32), 0x73747267u};
    philox4x32_10(c, key0, key1);
    for (int k = 0; k < 4; k++) u[k] = ((double)c[k] + 0.5) * 2.3283064365386963e-10;
}

static int str_out_of_box(const double* t, const double* mx, const double* mn)
{
    for (int k = 0; k < STR_NSIMP; k++)
        if (t[k] > mx[k] || t[k] < mn[k]) return 1;
    return 0;
}

/* defects_3d.f90:435-603 position_string: 1001 random starting points keep the best five
 * (the first two unsorted, as there), then the simplex iteration of the reference (reflection
 * 2c - p, expansion 3c - 4p, contraction 1.5c - 0.5p, else shrink towards the best point,
 * bubble sort) until chi2 <= 0.05 or 1000 iterations; then position, direction, goodness of
 * fit and the chord length.  s = the string's index in site order (its RNG stream). */
static void position_string_st(const double* ph, uint32_t key0, uint32_t key1, const orc_stream* st, int64_t s,
                               double* pos, double* dir, double* gof, double* sl);

void orc_position_string(const double* ph, uint32_t key0, uint32_t key1, int64_t s, double* pos,
                         double* dir, double* gof, double* sl)
{
    position_string_st(ph, key0, key1, NULL, s, pos, dir, gof, sl);
}

static void position_string_st(const double* ph, uint32_t key0, uint32_t key1, const orc_stream* st, int64_t s,
                               double* pos, double* dir, double* gof, double* sl)
{
    const double chi2max = 8 * PI_T * PI_T, tol = 0.05;
    const double mx[STR_NSIMP] = {1.0, PI_T, 0.86602540378443865, PI_T};
    const double mn[STR_NSIMP] = {-1.0, -PI_T, 0.0, -PI_T};
    double p[STR_NSIMP + 1][STR_NSIMP], c2[STR_NSIMP + 1], t[STR_NSIMP], c[STR_NSIMP], sv[STR_NSIMP];
    memset(p, 0, sizeof p);
    memset(c2, 0, sizeof c2);
    int n0 = -1;
    for (int i = 0; i <= STR_MAXIT; i++) {
        double x[4];
        string_uniforms(key0, key1, st, s, i, x);
        for (int k = 0; k < STR_NSIMP; k++) t[k] = mx[k] * x[k] + mn[k] * (1 - x[k]);
        double chi2 = orc_chi2_string(ph, t);
        if (n0 <= 0) {
            n0 = n0 + 1;
            c2[n0] = chi2;
            memcpy(p[n0], t, sizeof t);
            continue;
        }
        if (n0 < STR_NSIMP || chi2 <= c2[n0]) {
            n0 = n0 + 1 < STR_NSIMP ? n0 + 1 : STR_NSIMP;
            int i0 = n0;
            while (chi2 <= c2[i0 - 1]) {
                if (i0 <= STR_NSIMP) {
                    c2[i0] = c2[i0 - 1];
                    memcpy(p[i0], p[i0 - 1], sizeof t);
                }
                i0 = i0 - 1;
                if (i0 == 0) break;
            }
            c2[i0] = chi2;
            memcpy(p[i0], t, sizeof t);
        }
    }
    int it = 0;
    while (c2[0] > tol && it < STR_MAXIT) {
        it = it + 1;
        for (int k = 0; k < STR_NSIMP; k++) c[k] = 0;
        for (int i0 = 0; i0 < STR_NSIMP; i0++)
            for (int k = 0; k < STR_NSIMP; k++) c[k] = c[k] + p[i0][k];
        for (int k = 0; k < STR_NSIMP; k++) c[k] = c[k] / STR_NSIMP;
        for (int k = 0; k < STR_NSIMP; k++) t[k] = 2 * c[k] - p[STR_NSIMP][k];
        double chi2 = str_out_of_box(t, mx, mn) ? chi2max : orc_chi2_string(ph, t);
        if (chi2 < c2[STR_NSIMP]) {
            for (int k = 0; k < STR_NSIMP; k++) sv[k] = 3 * c[k] - 4 * p[STR_NSIMP][k];
            double chi2s = str_out_of_box(sv, mx, mn) ? chi2max : orc_chi2_string(ph, sv);
            if (chi2 <= chi2s) { c2[STR_NSIMP] = chi2; memcpy(p[STR_NSIMP], t, sizeof t); }
            else { c2[STR_NSIMP] = chi2s; memcpy(p[STR_NSIMP], sv, sizeof t); }
        } else {
            for (int k = 0; k < STR_NSIMP; k++) sv[k] = 1.5 * c[k] - 0.5 * p[STR_NSIMP][k];
            double chi2s = str_out_of_box(sv, mx, mn) ? chi2max : orc_chi2_string(ph, sv);
            if (chi2s <= c2[STR_NSIMP]) {
                c2[STR_NSIMP] = chi2s;
                memcpy(p[STR_NSIMP], sv, sizeof t);
            } else {
                for (int i0 = 1; i0 <= STR_NSIMP; i0++) {
                    for (int k = 0; k < STR_NSIMP; k++) p[i0][k] = 0.5 * (p[i0][k] + p[0][k]);
                    c2[i0] = orc_chi2_string(ph, p[i0]);
                }
            }
        }
        for (int i0 = STR_NSIMP - 1; i0 >= 0; i0--)
            for (int j = 0; j <= i0; j++)
                if (c2[j] > c2[j + 1]) {
                    double tmp[STR_NSIMP];
                    memcpy(tmp, p[j], sizeof tmp);
                    memcpy(p[j], p[j + 1], sizeof tmp);
                    memcpy(p[j + 1], tmp, sizeof tmp);
                    double q = c2[j]; c2[j] = c2[j + 1]; c2[j + 1] = q;
                }
    }
    double mu = p[0][0], phi = p[0][1];
    double u0 = p[0][2] * cos(p[0][3]), v0 = p[0][2] * sin(p[0][3]);
    double sm = sqrt(1.0 - mu * mu);
    dir[0] = sm * cos(phi);
    dir[1] = sm * sin(phi);
    dir[2] = mu;
    pos[0] = u0 * mu * cos(phi) - v0 * sin(phi);
    pos[1] = u0 * mu * sin(phi) + v0 * cos(phi);
    pos[2] = -u0 * sm;
    *gof = exp(-c2[0] / 2);
    *sl = orc_string_length(pos, dir);
}

/* defects_3d.f90:605-690 velocity_string (the VEL = 0 build): least-squares fit of phi and pi
 * to a + b.d over the 8 corners, d = corner - (pos + 1/2), through the packed-upper inverse of
 * the 4x4 normal matrix; vel = -[(x c - y b) u + (-x b + y a) v] / det * dspa with x + i y the
 * fitted pi, u + i v the fitted gradient, a = u.u, b = u.v, c = v.v, det = a c - b^2.
 * (x,y,z) is the cell's corner (0-based); pos is relative to the cell centre. */
void orc_velocity_string(const double* phi, const double* pi, int64_t N, int64_t x, int64_t y, int64_t z,
                         const double* pos, double dspa, double* vel)
{
    const oc* f = (const oc*)phi;
    const oc* q = (const oc*)pi;
    double wr[4] = {0}, wi[4] = {0}, dr[4] = {0}, di[4] = {0}, mat[10] = {0}, d[4];
    d[0] = 1.0;
    for (int k = 0; k < 2; k++) {
        int64_t iz = (z + k) % N;
        d[3] = k - pos[2] - 0.5;
        for (int j = 0; j < 2; j++) {
            int64_t iy = (y + j) % N;
            d[2] = j - pos[1] - 0.5;
            for (int i = 0; i < 2; i++) {
                int64_t ix = (x + i) % N;
                d[1] = i - pos[0] - 0.5;
                oc a = f[ix + N * (iy + N * iz)], b = q[ix + N * (iy + N * iz)];
                for (int m = 0; m < 4; m++) {
                    wr[m] = wr[m] + d[m] * a.re;
                    wi[m] = wi[m] + d[m] * a.im;
                    dr[m] = dr[m] + d[m] * b.re;
                    di[m] = di[m] + d[m] * b.im;
                }
                /* dspr('U'): A(r,c) += d_r d_c for r <= c, packed at r + c(c-1)/2 (1-based) */
                for (int cc = 1; cc <= 4; cc++)
                    for (int r = 1; r <= cc; r++) mat[r + cc * (cc - 1) / 2 - 1] += d[r - 1] * d[cc - 1];
            }
        }
    }
    for (int m = 0; m < 4; m++) { wr[m] /= 8; wi[m] /= 8; dr[m] /= 8; di[m] /= 8; }
    for (int m = 0; m < 10; m++) mat[m] /= 8;
    orc_pp_inverse(4, mat);
    double w0r[4] = {0}, w0i[4] = {0}, d0r[4] = {0}, d0i[4] = {0};
    for (int ii = 1; ii <= 4; ii++)
        for (int jj = 1; jj <= 4; jj++) {
            int kk = ii > jj ? ii : jj;
            kk = kk * (kk - 1) / 2 + (ii < jj ? ii : jj) - 1;
            w0r[ii - 1] += mat[kk] * wr[jj - 1];
            w0i[ii - 1] += mat[kk] * wi[jj - 1];
            d0r[ii - 1] += mat[kk] * dr[jj - 1];
            d0i[ii - 1] += mat[kk] * di[jj - 1];
        }
    double xx = d0r[0], yy = d0i[0];
    double a = 0, b = 0, c = 0;
    for (int m = 1; m < 4; m++) {
        a += w0r[m] * w0r[m];
        b += w0r[m] * w0i[m];
        c += w0i[m] * w0i[m];
    }
    double det = a * c - b * b;
    for (int m = 0; m < 3; m++) {
        double v = (xx * c - yy * b) * w0r[m + 1] + (-xx * b + yy * a) * w0i[m + 1];
        vel[m] = -v / det * dspa;
    }
}

/* defects_3d.f90:824-982 velocity_string_YY (the VEL = 1 build) for the cell with corner (x,y,z)
 * (0-based), one rank: for ipla = 1..3, (u,v,w) = cshift((1,2,3), ipla-1), and k = 0, 1, the
 * quadrant test of the face's corners (i, j) in i-outer order; on a pierced face vphi = sum of
 * (1, i-1/2, j-1/2) phi, d = invmat vphi with invmat = diag(1/4, 1, 1) (the packed inverse of
 * diag(4, 1, 1), exact), det, du, dv (clamped to [-1/2, 1/2]), the weights (1/2 + sign(du, i-1/2))
 * (1/2 + sign(dv, j-1/2)) applied to pi and the central differences of phi (one-sided in z from
 * the halo plane above the top plane), a11, a12, a22, vpen = -(...)/det*dspa; then vel = the
 * sum over the pierced faces of vpen(:, npen) -- the last face's, as the reference indexes it --
 * divided by npen, or 1e-10 without a pierced face. */
void orc_velocity_string_yy(const double* phi, const double* pi, int64_t N, int64_t x, int64_t y, int64_t z,
                            double dspa, double* vel)
{
    const oc* f = (const oc*)phi;
    const oc* q = (const oc*)pi;
    const int64_t c0[3] = {x, y, z};
#define AT(a, cx, cy, cz) (a)[(cx) + N * ((cy) + N * (cz))]
    int npen = 0;
    double last[3] = {0, 0, 0};
    for (int pl = 0; pl < 3; pl++) {
        const int au = pl, av = (pl + 1) % 3, aw = (pl + 2) % 3;
        for (int k = 0; k < 2; k++) {
            int64_t cc[2][2][3];
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    for (int a = 0; a < 3; a++)
                        cc[i][j][a] = c0[a] + (a == au ? i : 0) + (a == av ? j : 0) + (a == aw ? k : 0);
                    if (cc[i][j][0] >= N) cc[i][j][0] -= N;
                    if (cc[i][j][1] >= N) cc[i][j][1] -= N;
                }
            int quad[5] = {0, 0, 0, 0, 0}, n = 0;
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const int64_t *c = cc[i][j];
                    const oc v = AT(f, c[0], c[1], c[2] % N);
                    const double ph = atan2(v.im, v.re);
                    const int m = (int)ceil(2 * ph / PI_T) + 2;
                    if (!quad[m]) { quad[m] = 1; n++; }
                }
            if (n < 3) continue;
            npen++;
            oc vp[3] = {{0, 0}, {0, 0}, {0, 0}};
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const int64_t *c = cc[i][j];
                    const oc v = AT(f, c[0], c[1], c[2] % N);
                    const double w[3] = {1.0, i - 0.5, j - 0.5};
                    for (int m = 0; m < 3; m++) { vp[m].re = vp[m].re + w[m] * v.re; vp[m].im = vp[m].im + w[m] * v.im; }
                }
            const double d01 = 0.25 * vp[0].re, d02 = 0.25 * vp[0].im;
            const double d11 = vp[1].re, d12 = vp[1].im, d21 = vp[2].re, d22 = vp[2].im;
            const double det = d11 * d22 - d21 * d12;
            double du = -((d22 * d01 - d21 * d02) / det);
            double dv = -((-(d12 * d01) + d11 * d02) / det);
            du = fmax(fmin(du, 0.5), -0.5);
            dv = fmax(fmin(dv, 0.5), -0.5);
            oc dp[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const double wt = (0.5 + (i ? fabs(du) : -fabs(du))) * (0.5 + (j ? fabs(dv) : -fabs(dv)));
                    const int64_t *c = cc[i][j];
                    const int64_t cz = c[2] % N;
                    const oc p = AT(q, c[0], c[1], cz);
                    dp[0].re = dp[0].re + wt * p.re;
                    dp[0].im = dp[0].im + wt * p.im;
                    oc a = AT(f, (c[0] + 1) % N, c[1], cz), b = AT(f, (c[0] + N - 1) % N, c[1], cz);
                    dp[1].re = dp[1].re + wt * (a.re - b.re) / 2;
                    dp[1].im = dp[1].im + wt * (a.im - b.im) / 2;
                    a = AT(f, c[0], (c[1] + 1) % N, cz);
                    b = AT(f, c[0], (c[1] + N - 1) % N, cz);
                    dp[2].re = dp[2].re + wt * (a.re - b.re) / 2;
                    dp[2].im = dp[2].im + wt * (a.im - b.im) / 2;
                    /* evolve/defects z indexing on one rank: l1 = iz+1 unless it passes the upper
                     * halo plane N+1 (1-based), then l1 = iz; divisor l1 - l2 */
                    int64_t zp, zm;
                    double dz;
                    if (c[2] == N) { zp = 0; zm = N - 1; dz = 1; }
                    else { zp = (c[2] + 1) % N; zm = (c[2] + N - 1) % N; dz = 2; }
                    a = AT(f, c[0], c[1], zp);
                    b = AT(f, c[0], c[1], zm);
                    dp[3].re = dp[3].re + wt * (a.re - b.re) / dz;
                    dp[3].im = dp[3].im + wt * (a.im - b.im) / dz;
                }
            double a11 = 0, a12 = 0, a22 = 0;
            for (int m = 1; m <= 3; m++) {
                a11 = a11 + dp[m].re * dp[m].re;
                a12 = a12 + dp[m].re * dp[m].im;
                a22 = a22 + dp[m].im * dp[m].im;
            }
            const double dt = a11 * a22 - a12 * a12;
            for (int m = 0; m < 3; m++) {
                const double v = dp[0].re * (a22 * dp[m + 1].re - a12 * dp[m + 1].im) +
                                 dp[0].im * (a11 * dp[m + 1].im - a12 * dp[m + 1].re);
                last[m] = -(v / dt * dspa);
            }
        }
    }
#undef AT
    if (npen != 0) {
        for (int m = 0; m < 3; m++) {
            double t = 0;
            for (int i = 0; i < npen; i++) t = t + last[m];
            vel[m] = t / npen;
        }
    } else {
        vel[0] = vel[1] = vel[2] = 1.e-10;
    }
}

/* id_defects' string loop (defects_3d.f90:167-205: cells with the string flag in site order,
 * 8 corner phases atan2(Im, Re), position_string, velocity_string, pos + ixyz - 1/2 with ixyz
 * the 1-based corner, i.e. site coordinates counted from 0) and add_string_length (1654-1733, group_strings = F).  rec (or NULL)
 * receives 11 doubles per string: pos(3), dir(3), gof, sl, vel(3).  sums = {sum sl, n with
 * |v| < 1, sum |v|, sum v^2, sum gamma} over those.  Returns the string count. */
int64_t orc_strings_st(const double* phi, const double* pi, int64_t N, const int8_t* flags, uint32_t key0,
                       uint32_t key1, const orc_stream* st, double dspa, int vel_yy, double* rec, double* sums);

/* vel_yy != 0: the velocities of velocity_string_YY (the VEL = 1 build) instead */
int64_t orc_strings(const double* phi, const double* pi, int64_t N, const int8_t* flags, uint32_t key0,
                    uint32_t key1, double dspa, int vel_yy, double* rec, double* sums)
{
    return orc_strings_st(phi, pi, N, flags, key0, key1, NULL, dspa, vel_yy, rec, sums);
}

/* orc_strings with the reference's stream: x0, pos = the stream's start and the position of the
 * first string's first draw. */
int64_t orc_strings_mcg(const double* phi, const double* pi, int64_t N, const int8_t* flags, uint64_t x0,
                        uint64_t pos, double dspa, int vel_yy, double* rec, double* sums)
{
    orc_stream st = {1, x0, pos};
    return orc_strings_st(phi, pi, N, flags, 0, 0, &st, dspa, vel_yy, rec, sums);
}

int64_t orc_strings_st(const double* phi, const double* pi, int64_t N, const int8_t* flags, uint32_t key0,
                       uint32_t key1, const orc_stream* st, double dspa, int vel_yy, double* rec, double* sums)
{
    const oc* f = (const oc*)phi;
    int64_t n3 = N * N * N, ns = 0;
    for (int64_t i = 0; i < n3; i++) ns += (flags[i] & 1) != 0;
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (ns > 0 ? ns : 1));
    double* r = (double*)malloc(sizeof(double) * 11 * (ns > 0 ? ns : 1));
    ns = 0;
    for (int64_t i = 0; i < n3; i++)
        if (flags[i] & 1) cell[ns++] = i;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < ns; s++) {
        int64_t i = cell[s], x = i % N, y = (i / N) % N, z = i / (N * N);
        double ph[8];
        for (int c = 0; c < 8; c++) {
            int64_t cx = (x + (c & 1)) % N, cy = (y + ((c >> 1) & 1)) % N, cz = (z + ((c >> 2) & 1)) % N;
            oc v = f[cx + N * (cy + N * cz)];
            ph[c] = atan2(v.im, v.re);
        }
        double* o = r + 11 * s;
        position_string_st(ph, key0, key1, st, s, o, o + 3, o + 6, o + 7);
        if (vel_yy) orc_velocity_string_yy(phi, pi, N, x, y, z, dspa, o + 8);
        else orc_velocity_string(phi, pi, N, x, y, z, o, dspa, o + 8);
        o[0] += (double)(x + 1) - 0.5;
        o[1] += (double)(y + 1) - 0.5;
        o[2] += (double)(z + 1) - 0.5;
    }
    double sl = 0, nv = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int64_t s = 0; s < ns; s++) {
        const double* o = r + 11 * s;
        sl += o[7];
        double v2 = o[8] * o[8] + o[9] * o[9] + o[10] * o[10];
        if (v2 >= 1.0) continue;
        nv += 1;
        s1 += sqrt(v2);
        s2 += v2;
        s3 += 1 / sqrt(1 - v2);
    }
    sums[0] = sl; sums[1] = nv; sums[2] = s1; sums[3] = s2; sums[4] = s3;
    if (rec) memcpy(rec, r, sizeof(double) * 11 * ns);
    free(cell);
    free(r);
    return ns;
}

/* ------------------------------------------------------------ thermal initial state */

/* initial.f90:372-390 initial_phi_phidot with get_comwav_initial (363-370): the factor that
 * turns a unit complex Gaussian times dv_wav into the amplitude of the mode |n|^2 = n2 of phi
 * (ip = 1) or phidot (ip = 2) at tau_ini: Bose-Einstein occupation of a mode of frequency
 * osc = sqrt(m^2(T) + k_phys^2), T = thrm_sigma / tau_ini, in the reference's order. */
double orc_thermal_factor(int ip, int64_t n2, double tau_ini, double sigma_cr, double dwav, double dv_wav)
{
    double q = (double)n2;
    q = sqrt(q) * dwav;
    double wav2 = q / tau_ini / sigma_cr;
    wav2 = wav2 * wav2;
    double thrm = THRM_SIGMA / tau_ini;
    double osc = sqrt(orc_msquared(thrm) + wav2);
    double fbe = exp(-osc / thrm);
    fbe = fbe / (1 - fbe);
    double var = fbe * osc / dv_wav / sigma_cr;
    double f = sqrt(var / 2);
    if (ip == 1) f = f / (osc * sigma_cr);
    return f;
}

/* Unit complex Gaussian number lin of the stream (seed, tag, word): Philox4x32-10 on the
 * counter (lin, tag, word) keyed by the seed, Box-Muller on two 53-bit uniforms in (0,1]. */
static void unit_gauss(uint64_t seed, uint32_t tag, uint32_t word, uint64_t lin, double* re, double* im)
{
    uint32_t c[4] = {(uint32_t)lin, (uint32_t)(lin >> 32), tag, word};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11);
    uint64_t b = ((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11);
    a &= (1ull << 53) - 1; b &= (1ull << 53) - 1;
    double u1 = ((double)a + 1.0) / 9007199254740992.0;
    double u2 = ((double)b + 1.0) / 9007199254740992.0;
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    *re = r * cos(th);
    *im = r * sin(th);
}

/* Unit complex Gaussian of mode lin (site-order index of the wavevector) for field ip.  The
 * reference draws these from an MKL VSL MCG59 stream (initial.f90:105-135); this keeps one
 * draw per mode, independent of thread and rank counts. */
void orc_thermal_gauss(uint64_t seed, int ip, uint64_t lin, double* re, double* im)
{
    unit_gauss(seed, 0x74686d6cu, (uint32_t)ip, lin, re, im);
}

/* Unit complex Gaussian of mode lin in Monte Carlo realisation iter of the covariance
 * estimate (the reference draws from its MCG59 stream, spectrum.f90:592-630). */
void orc_cov_gauss(uint64_t seed, int iter, uint64_t lin, double* re, double* im)
{
    unit_gauss(seed, 0x636f7620u, (uint32_t)iter, lin, re, im);
}

/* initial.f90:80-238 get_initial_thermal (one rank): for phi then phidot, every mode of the
 * wavevector grid (x fastest) gets dv_wav times a unit complex Gaussian times
 * orc_thermal_factor, then the unnormalised FORWARD transform (FFTW_FORWARD, 332-333) gives
 * the field in real space. */
static void ic_thermal_st(double* phi, double* pi, int64_t N, uint64_t seed, const orc_stream* st, double tau_ini,
                          double sigma_cr, double dwav, double dv_wav);

void orc_ic_thermal(double* phi, double* pi, int64_t N, uint64_t seed, double tau_ini, double sigma_cr,
                    double dwav, double dv_wav)
{
    ic_thermal_st(phi, pi, N, seed, NULL, tau_ini, sigma_cr, dwav, dv_wav);
}

/* The same from the reference's stream at position pos: mode lin of field ip takes draws
 * pos + 4 N^3 (ip - 1) + 4 lin .. +3, i.e. Gaussians (re, im) of those two pairs
 * (initial.f90:102-126: mult = 2 Gaussians x 2 draws per site, phi's N^3 sites then phidot's). */
void orc_ic_thermal_mcg(double* phi, double* pi, int64_t N, uint64_t x0, uint64_t pos, double tau_ini,
                        double sigma_cr, double dwav, double dv_wav)
{
    orc_stream st = {1, x0, pos};
    ic_thermal_st(phi, pi, N, 0, &st, tau_ini, sigma_cr, dwav, dv_wav);
}

static void ic_thermal_st(double* phi, double* pi, int64_t N, uint64_t seed, const orc_stream* st, double tau_ini,
                          double sigma_cr, double dwav, double dv_wav)
{
    int64_t n3 = N * N * N;
    oc* buf = (oc*)malloc(sizeof(oc) * n3);
    for (int ip = 1; ip <= 2; ip++) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int64_t z = 0; z < N; z++)
            for (int64_t y = 0; y < N; y++)
                for (int64_t x = 0; x < N; x++) {
                    int64_t lin = x + N * (y + N * z);
                    int64_t nx = wrapn(x, N), ny = wrapn(y, N), nz = wrapn(z, N);
                    double g1, g2;
                    if (st && st->on) {
                        uint64_t x = orc_mcg59_jump(st->x0, st->pos + 4 * (uint64_t)n3 * (ip - 1) + 4 * (uint64_t)lin);
                        g1 = mcg59_gauss(&x);
                        g2 = mcg59_gauss(&x);
                    } else {
                        orc_thermal_gauss(seed, ip, (uint64_t)lin, &g1, &g2);
                    }
                    double f = orc_thermal_factor(ip, nx * nx + ny * ny + nz * nz, tau_ini, sigma_cr, dwav, dv_wav);
                    buf[lin].re = (g1 * dv_wav) * f;
                    buf[lin].im = (g2 * dv_wav) * f;
                }
        orc_fft3d((const double*)buf, ip == 1 ? phi : pi, N, -1);
    }
    free(buf);
}

/* ------------------------------------------------------ spectrum covariance (MC) */

/* spectrum.f90:569-734 estimate_covmat (degrade_spec = 0, one rank).  power: the deconvolved
 * spectrum; matn: the packed-upper deconvolution matrix after division by num_i num_j
 * (mat_deconv as spectrum_estimation leaves it, 316-323).  Realisation iter: mode k (bins by
 * i = nint(q/dq+0.5) <= nb) = orc_cov_gauss * sqrt(power/2)/N^3, else 0; FORWARD FFT; mask
 * cells (flag bit 4) zeroed; BACKWARD FFT; pow0 = binned |.|^2; pow = matn pow0 - power;
 * cov += pow pow^T, with the leave-one-out sums tmp(:,k) of the jackknife; from iter = min_mc
 * on, stop when max over the diagonal of sd(tmp/(iter-1))/(cov/iter) < tol.  cov is divided
 * by the final loop counter, max_mc + 1 without convergence.  Returns the counter, negated
 * without convergence.  pows (optional, max_mc x nb) receives each realisation's pow. */
static int field_covariance_st(const int8_t* flags, int64_t N, int nb, double dspa, double dwav,
                               const double* power, const double* matn, uint64_t seed, const orc_stream* st,
                               int min_mc, int max_mc, double tol, double* cov, double* pows);

int orc_field_covariance(const int8_t* flags, int64_t N, int nb, double dspa, double dwav, const double* power,
                         const double* matn, uint64_t seed, int min_mc, int max_mc, double tol, double* cov,
                         double* pows)
{
    return field_covariance_st(flags, N, nb, dspa, dwav, power, matn, seed, NULL, min_mc, max_mc, tol, cov, pows);
}

/* The same from the reference's stream at position pos: realisation iter (1-based) draws
 * pos + 4 N^3 (iter - 1) + 4 k .. +3 for mode k (spectrum.f90:592-630, all N^3 modes, binned or
 * not). */
int orc_field_covariance_mcg(const int8_t* flags, int64_t N, int nb, double dspa, double dwav, const double* power,
                             const double* matn, uint64_t x0, uint64_t pos, int min_mc, int max_mc, double tol,
                             double* cov, double* pows)
{
    orc_stream st = {1, x0, pos};
    return field_covariance_st(flags, N, nb, dspa, dwav, power, matn, 0, &st, min_mc, max_mc, tol, cov, pows);
}

static int field_covariance_st(const int8_t* flags, int64_t N, int nb, double dspa, double dwav,
                               const double* power, const double* matn, uint64_t seed, const orc_stream* st,
                               int min_mc, int max_mc, double tol, double* cov, double* pows)
{
    const int64_t n3 = N * N * N;
    const int ncov = nb * (nb + 1) / 2;
    double qmax = PI_T / dspa, dq0 = qmax / nb;
    double* wave = (double*)calloc(nb + 2, sizeof(double));
    for (int b = 1; b <= nb; b++) wave[b] = wave[b - 1] + dq0;
    const double dq = wave[1] - wave[0];
    free(wave);
    int* bin = (int*)malloc(sizeof(int) * n3);
    for (int64_t z = 0; z < N; z++)
        for (int64_t y = 0; y < N; y++)
            for (int64_t x = 0; x < N; x++) {
                int64_t nx = wrapn(x, N), ny = wrapn(y, N), nz = wrapn(z, N);
                double q = (double)(nx * nx) + (double)(ny * ny) + (double)(nz * nz);
                q = sqrt(q) * dwav;
                long i = lround(q / dq + 0.5);
                bin[x + N * (y + N * z)] = (i >= 1 && i <= nb) ? (int)i : 0;
            }
    double* sigma = (double*)malloc(sizeof(double) * nb);
    for (int b = 0; b < nb; b++) sigma[b] = sqrt(power[b] / 2) / (double)n3;
    double* w = (double*)malloc(sizeof(double) * 2 * n3);
    double* pow0 = (double*)malloc(sizeof(double) * nb);
    double* pw = (double*)malloc(sizeof(double) * nb);
    double* tmp = (double*)calloc((size_t)ncov * max_mc, sizeof(double));
    for (int ij = 0; ij < ncov; ij++) cov[ij] = 0;
    int iter, converged = 0;
    for (iter = 1; iter <= max_mc; iter++) {
        uint64_t xs = st && st->on ? orc_mcg59_jump(st->x0, st->pos + 4 * (uint64_t)n3 * (iter - 1)) : 0;
        for (int64_t k = 0; k < n3; k++) {
            double g1, g2;
            if (st && st->on) {
                g1 = mcg59_gauss(&xs);
                g2 = mcg59_gauss(&xs);
            } else {
                orc_cov_gauss(seed, iter, (uint64_t)k, &g1, &g2);
            }
            if (bin[k]) {
                w[2 * k] = g1 * sigma[bin[k] - 1];
                w[2 * k + 1] = g2 * sigma[bin[k] - 1];
            } else {
                w[2 * k] = 0;
                w[2 * k + 1] = 0;
            }
        }
        orc_fft3d(w, w, N, -1);
        if (flags)
            for (int64_t k = 0; k < n3; k++)
                if (flags[k] & 4) { w[2 * k] = 0; w[2 * k + 1] = 0; }
        orc_fft3d(w, w, N, +1);
        for (int b = 0; b < nb; b++) pow0[b] = 0;
        for (int64_t k = 0; k < n3; k++)
            if (bin[k]) pow0[bin[k] - 1] += w[2 * k] * w[2 * k] + w[2 * k + 1] * w[2 * k + 1];
        /* dspmv('U', nb, 1, matn, pow0, 1, 1, pow, 1) with pow = -power */
        for (int i = 1; i <= nb; i++) {
            double acc = -power[i - 1];
            for (int j = 1; j <= nb; j++) {
                int a = i < j ? i : j, b = i < j ? j : i;
                acc += matn[a + b * (b - 1) / 2 - 1] * pow0[j - 1];
            }
            pw[i - 1] = acc;
        }
        if (pows)
            for (int b = 0; b < nb; b++) pows[(size_t)(iter - 1) * nb + b] = pw[b];
        for (int i = 1; i <= nb; i++)
            for (int j = 1; j <= i; j++) {
                int ij = i * (i - 1) / 2 + j - 1;
                double u = pw[i - 1] * pw[j - 1];
                tmp[(size_t)ij * max_mc + iter - 1] = cov[ij];
                for (int k = 0; k < iter - 1; k++) tmp[(size_t)ij * max_mc + k] += u;
                cov[ij] += u;
            }
        if (iter < min_mc) continue;
        double diff = 0;
        for (int i = 1; i <= nb; i++) {
            int ij = i * (i + 1) / 2 - 1;
            double u = 0, v = 0;
            for (int k = 0; k < iter; k++) {
                double x = tmp[(size_t)ij * max_mc + k] / (iter - 1);
                u = u + x;
                v = v + x * x;
            }
            u = u / iter;
            v = v / iter;
            v = sqrt(v - u * u);
            double x = cov[ij] / iter;
            if (v / x > diff) diff = v / x;
        }
        if (diff < tol) { converged = 1; break; }
    }
    for (int ij = 0; ij < ncov; ij++) cov[ij] = cov[ij] / iter;
    free(bin); free(sigma); free(w); free(pow0); free(pw); free(tmp);
    return converged ? iter : -iter;
}

/* ------------------------------------------------------------ wall fit and area */

/* defects_3d.f90:1058-1202 position_wall: the points where sin(ph) changes sign on the cell's
 * edges (linear interpolation), in the reference's edge order; the plane normal by a 20 x 20
 * grid search in (mu, phi) minimising the spread of the points' distance along it; the
 * orientation by the corners' cos(ph); and the area of the polygon the points span in that
 * plane (angular sort, triangle fan).  Returns the number of edge points m; m < 3 is the
 * reference's stop ('something wrong in position_wall') and leaves the outputs unset. */
int orc_position_wall(const double* ph, double* pos, double* dir, double* gof, double* wa)
{
    double pp[12][3], th[12], uv[12][2];
    int m = 0;
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++) {
            double u = sin(ph[(iz * 2 + iy) * 2 + 0]), v = sin(ph[(iz * 2 + iy) * 2 + 1]);
            if (u * v < 0) { pp[m][0] = (u + v) / (u - v) / 2; pp[m][1] = iy - 0.5; pp[m][2] = iz - 0.5; m++; }
        }
    for (int iy = 0; iy < 2; iy++)
        for (int ix = 0; ix < 2; ix++) {
            double u = sin(ph[(0 * 2 + iy) * 2 + ix]), v = sin(ph[(1 * 2 + iy) * 2 + ix]);
            if (u * v < 0) { pp[m][0] = ix - 0.5; pp[m][1] = iy - 0.5; pp[m][2] = (u + v) / (u - v) / 2; m++; }
        }
    for (int iz = 0; iz < 2; iz++)
        for (int ix = 0; ix < 2; ix++) {
            double u = sin(ph[(iz * 2 + 0) * 2 + ix]), v = sin(ph[(iz * 2 + 1) * 2 + ix]);
            if (u * v < 0) { pp[m][0] = ix - 0.5; pp[m][1] = (u + v) / (u - v) / 2; pp[m][2] = iz - 0.5; m++; }
        }
    if (m < 3) return m;
    double chi2min = 6.0, mut = 0, phit = 0, wt = 0;
    for (int imu = 1; imu <= 20; imu++) {
        double mu = (-20 + 2 * imu - 1.0) / 20;
        for (int iphi = 1; iphi <= 20; iphi++) {
            double phi = PI_T * (2 * iphi - 1.0 - 20) / 20;
            double d0 = sqrt(1 - mu * mu) * cos(phi), d1 = sqrt(1 - mu * mu) * sin(phi), d2 = mu;
            double w = 0, chi2 = 0;
            for (int i = 0; i < m; i++) {
                double d = pp[i][0] * d0 + pp[i][1] * d1 + pp[i][2] * d2;
                w = w + d;
                chi2 = chi2 + d * d;
            }
            if (w < 0) w = 0;
            else w = w / m;
            chi2 = chi2 - m * w * w;
            if (chi2 < chi2min) { chi2min = chi2; mut = mu; phit = phi; wt = w; }
        }
    }
    double dr[3] = {sqrt(1 - mut * mut) * cos(phit), sqrt(1 - mut * mut) * sin(phit), mut};
    int num = 0;
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                double d = (ix - 0.5) * dr[0] + (iy - 0.5) * dr[1] + (iz - 0.5) * dr[2] - wt;
                if (d * cos(ph[(iz * 2 + iy) * 2 + ix]) > 0) num++;
            }
    if (num <= 4) { dr[0] = -dr[0]; dr[1] = -dr[1]; dr[2] = -dr[2]; wt = -wt; }
    for (int k = 0; k < 3; k++) { dir[k] = dr[k]; pos[k] = wt * dr[k]; }
    *gof = exp(-chi2min / 2);
    double e1[3] = {mut * cos(phit), mut * sin(phit), -sqrt(1 - mut * mut)};
    double e2[3] = {-sin(phit), cos(phit), 0.0};
    for (int i = 0; i < m; i++) {
        for (int k = 0; k < 3; k++) pp[i][k] = pp[i][k] - pos[k];
        double u = e1[0] * pp[i][0] + e1[1] * pp[i][1] + e1[2] * pp[i][2];
        double v = e2[0] * pp[i][0] + e2[1] * pp[i][1] + e2[2] * pp[i][2];
        th[i] = atan2(v, u);
        uv[i][0] = u;
        uv[i][1] = v;
    }
    for (int i = m - 2; i >= 0; i--)
        for (int j = 0; j <= i; j++)
            if (th[j] > th[j + 1]) {
                double t = th[j]; th[j] = th[j + 1]; th[j + 1] = t;
                double a = uv[j][0], b = uv[j][1];
                uv[j][0] = uv[j + 1][0]; uv[j][1] = uv[j + 1][1];
                uv[j + 1][0] = a; uv[j + 1][1] = b;
            }
    double area = 0, w = th[m - 1] - TWOPI_T, u = sqrt(uv[m - 1][0] * uv[m - 1][0] + uv[m - 1][1] * uv[m - 1][1]);
    for (int i = 0; i < m; i++) {
        double v = sqrt(uv[i][0] * uv[i][0] + uv[i][1] * uv[i][1]);
        w = th[i] - w;
        area = area + 0.5 * u * v * sin(w);
        u = v;
        w = th[i];
    }
    *wa = area;
    return m;
}

/* id_defects' wall loop (defects_3d.f90:183-188: cells with the wall flag in site order,
 * position_wall, pos + ixyz - 1/2) and add_wall_area (2145-2204, group_walls = F).  rec (or
 * NULL) receives 8 doubles per wall: pos(3), dir(3), gof, wa.  sums = {sum wa, cells with
 * fewer than 3 edge points}.  Returns the wall count. */
int64_t orc_walls(const double* phi, int64_t N, const int8_t* flags, double* rec, double* sums)
{
    const oc* f = (const oc*)phi;
    int64_t n3 = N * N * N, nw = 0;
    for (int64_t i = 0; i < n3; i++) nw += (flags[i] & 2) != 0;
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (nw > 0 ? nw : 1));
    double* r = (double*)malloc(sizeof(double) * 8 * (nw > 0 ? nw : 1));
    int* bad = (int*)calloc(nw > 0 ? nw : 1, sizeof(int));
    nw = 0;
    for (int64_t i = 0; i < n3; i++)
        if (flags[i] & 2) cell[nw++] = i;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nw; s++) {
        int64_t i = cell[s], x = i % N, y = (i / N) % N, z = i / (N * N);
        double ph[8];
        for (int c = 0; c < 8; c++) {
            int64_t cx = (x + (c & 1)) % N, cy = (y + ((c >> 1) & 1)) % N, cz = (z + ((c >> 2) & 1)) % N;
            oc v = f[cx + N * (cy + N * cz)];
            ph[c] = atan2(v.im, v.re);
        }
        double* o = r + 8 * s;
        if (orc_position_wall(ph, o, o + 3, o + 6, o + 7) < 3) {
            bad[s] = 1;
            for (int k = 0; k < 8; k++) o[k] = 0;
            continue;
        }
        o[0] += (double)(x + 1) - 0.5;
        o[1] += (double)(y + 1) - 0.5;
        o[2] += (double)(z + 1) - 0.5;
    }
    double wa = 0, nb = 0;
    for (int64_t s = 0; s < nw; s++) { wa += r[8 * s + 7]; nb += bad[s]; }
    sums[0] = wa;
    sums[1] = nb;
    if (rec) memcpy(rec, r, sizeof(double) * 8 * nw);
    free(cell); free(r); free(bad);
    return nw;
}

/* ------------------------------------------------------------ string / wall networks */

/* defects_3d.f90:1645-1652 lat_dist_periodic */
double orc_lat_dist(const double* p1, const double* p2, int64_t N)
{
    double s = 0;
    for (int k = 0; k < 3; k++) {
        double dp = fabs(p1[k] - p2[k]);
        double m = dp < (double)N - dp ? dp : (double)N - dp;
        s = s + m * m;
    }
    return sqrt(s);
}

static int64_t uf_find(int64_t* par, int64_t i)
{
    while (par[i] != i) {
        par[i] = par[par[i]];
        i = par[i];
    }
    return i;
}

/* grouping_strings / grouping_walls (defects_3d.f90:1235-1260, 1441-1643, 1735-1760): the
 * connected components of the graph linking records (pos = the first 3 of `stride` doubles)
 * at most th apart in lat_dist_periodic, by brute force over all pairs.  label[i] = the
 * smallest record index of i's group. */
void orc_group_points(const double* rec, int stride, int64_t n, int64_t N, double th, int32_t* label)
{
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) par[i] = i;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            if (!(orc_lat_dist(rec + stride * i, rec + stride * j, N) <= th)) continue;
            int64_t a = uf_find(par, i), b = uf_find(par, j);
            if (a == b) continue;
            if (a < b) par[b] = a;
            else par[a] = b;
        }
    for (int64_t i = 0; i < n; i++) label[i] = (int32_t)uf_find(par, i);
    free(par);
}

/* add_string_length / add_wall_area with grouping (defects_3d.f90:1676-1702, 2172-2200):
 * groups numbered by their first record, each group's values summed in record order, then
 * out[0] = sum of all groups, out[1] = of groups above t2, out[2] = of groups above t3.
 * Returns the number of groups. */
int64_t orc_group_sums(const int32_t* label, const double* val, int64_t n, double t2, double t3, int32_t* ids,
                       double* out)
{
    double* g = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    int64_t ng = 0;
    for (int64_t i = 0; i < n; i++) {
        if (label[i] == i) {
            g[ng] = 0;
            ids[i] = (int32_t)(++ng);
        } else {
            ids[i] = ids[label[i]];
        }
        g[ids[i] - 1] += val[i];
    }
    out[0] = out[1] = out[2] = 0;
    for (int64_t k = 0; k < ng; k++) {
        out[0] = out[0] + g[k];
        if (g[k] > t2) out[1] = out[1] + g[k];
        if (g[k] > t3) out[2] = out[2] + g[k];
    }
    free(g);
    return ng;
}
