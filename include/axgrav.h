/*
 * axgrav.h -- C-ABI of the MI355X lattice step and diagnostics for axion_grav.
 *
 * A drop-in for the reference's hot path: a Fortran host keeps main.f90's driver, the
 * inifile/settings keys and the output formats, and calls these entry points through
 * ISO_C_BINDING (see axion_grav_amd/fortran/axgrav_capi.f90).  All entry points are
 * extern "C", take plain pointers and sizes, and return an int status (0 = ok, < 0 =
 * error); the message of the last failure is axg_last_error().
 *
 * Field layout at the boundary: exactly the reference's complex(C_DOUBLE_COMPLEX)
 * data1/data2 interior, N^3 sites, x fastest, interleaved (re, im) doubles, no halo
 * planes.  The handle owns every device buffer; fields stay resident in HBM between
 * calls; host pointers are borrowed only for the duration of a call; every call
 * synchronizes before it returns.  One host thread per handle; handles are independent.
 */
#ifndef AXGRAV_H
#define AXGRAV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AXG_OK = 0,
    AXG_EINVAL = -1,  /* bad argument (size, null pointer, unsupported N) */
    AXG_ENOMEM = -2,  /* device allocation failed */
    AXG_EDEVICE = -3, /* HIP runtime error */
    AXG_ESTATE = -4,  /* call out of order (e.g. spectrum before defects) */
    AXG_ENUMERIC = -5 /* numerical failure (e.g. deconvolution matrix not positive definite) */
};

/* Field storage: FP64; FP32C = FP32 value + int16 residual (about 39 significant bits) with FP64
 * arithmetic; FP32 = plain FP32 storage, FP64 arithmetic (a reference point for FP32C's
 * precision, not a drop-in mode).  GW transforms and accumulators are FP32 in both FP32 modes. */
enum { AXG_FP64 = 0, AXG_FP32C = 1, AXG_FP32 = 2 };

/* Parameters: the settings.f90 globals the hot path reads (settings.f90:17-66,
 * potential.f90:24) plus the keys this build adds.  Derive the physical ones with
 * axg_derive() from the ini keys, or fill them in from an existing settings module. */
typedef struct axg_params {
    int64_t ngrid;          /* N (settings key ngrid); even, 2^a * 3^b, 8 <= N <= 1024 */
    int32_t precision;      /* AXG_FP64, AXG_FP32C or AXG_FP32 (new key precision) */
    int32_t grav_spec;      /* accumulate GW each kick (key grav_spec) */
    int32_t initial_loop;   /* thermal mass off (key initial_loop; evolve.f90:292-296) */
    int32_t gw_fix_swap;    /* 0: reproduce grav_waves.f90:239-242 exactly; 1: intended form */
    int32_t num_bin_grav;   /* GW spectrum bins (grav_waves.f90:6 hard-codes 25) */
    int32_t num_bin;        /* axion spectrum bins (key num_bin, default 50) */
    int32_t id_strings;     /* keys id_strings / id_walls / spec_est (defect flag bits) */
    int32_t id_walls;
    int32_t spec_est;
    int32_t device;         /* HIP device ordinal */
    double dspa, dwav, dv_spa, dv_wav;   /* settings.f90:294-297 */
    double sigma_cr, tau_cr, a_now;      /* settings.f90:274-276 */
    double kappa;                        /* potential.f90:24 (key kappa) */
    double eps_dtau;                     /* key eps_dtau (get_dtau, evolve.f90:10-24) */
    double dtau_fixed;                   /* > 0: fixed step instead of get_dtau (new key) */
    double potential_kappa0;             /* N_DW=2 tilt -kappa0 tau^4 (phi^2+phi*^2)/2 beside
                                            lambda/4 (|phi|^2-1)^2, lambda = 4 sigma_cr^2; 0 = off */
    int32_t degrade_spec;   /* key degrade_spec (settings.f90:45): the GW source and the axion
                               spectrum live on the reduced grid N / 2^degrade_spec (>= 8), as
                               init_grav / init_spectrum set it up (grav_waves.f90:37-45,
                               spectrum.f90:42-49); 0 = the lattice grid */
    int32_t string_vel;     /* the string velocity estimator of the reference's build option VEL
                               (Makefile:11): 0 = velocity_string (defects_3d.f90:605-690, the
                               default build), 1 = velocity_string_YY (824-982) */
    int32_t reserved[6];
} axg_params;

/* The ini keys axg_derive() needs (settings.f90:269-297). */
typedef struct axg_keys {
    double sigma, g_rel, kappa, tau_fin, boxsize;
    int64_t ngrid;
} axg_keys;

/* Derived units of settings.f90:274-297 in the reference's FP64 operation order (with the
 * truncated pi of potential.f90:8).  Fills the physical fields of *p; leaves the rest. */
int axg_derive(const axg_keys* k, axg_params* p, double* zeta, double* width_cr,
               double* thickness_cr);

/* get_dtau (evolve.f90:10-24): the step and step count for the interval [t1, t2]. */
int axg_get_dtau(const axg_params* p, double t1, double t2, double* dtau, int32_t* nsteps);

typedef struct axg_handle axg_handle;

/* Allocate the lattice on the device (replaces init_grav/init_spectrum's allocations,
 * grav_waves.f90:27-82, spectrum.f90:23-103, initial.f90:240-285).  GW accumulators
 * start at zero (grav_waves.f90:77). */
int axg_create(const axg_params* p, axg_handle** out);
int axg_destroy(axg_handle* h);
/* Message of the last failed call on h; h == NULL gives the last failure of a call without a
 * handle (axg_create, axg_energy_fields, axg_hbm_probe) on the calling thread. */
const char* axg_last_error(const axg_handle* h);

/* Copy phi (data1) and pi = phi' (data2) in / out; nsite must be N^3.  In FP32C mode the
 * FP64 values are split into an FP32 value and its compensation. */
int axg_set_fields(axg_handle* h, const double* phi, const double* pi, int64_t nsite);
int axg_get_fields(axg_handle* h, double* phi, double* pi, int64_t nsite);

/* The reference's random stream (subroutines.f90:49-159): from this call on, the handle's random
 * draws come from Intel MKL VSL's MCG59 generator (x_{n+1} = 13^13 x_n mod 2^59, uniform
 * x/2^59, Gaussians by Box-Muller on consecutive pairs, as VSL_RNG_METHOD_GAUSSIAN_BOXMULLER)
 * started as vslNewStreamEx(VSL_BRNG_MCG59, 2, {seed1, seed2}) does -- the reference's key ss --
 * and positioned at `position` draws, in the reference's order: axg_init_thermal takes
 * 8 N^3 draws (initial.f90:102-126: per site 2 Gaussians of phi, then those of phidot),
 * axg_strings 4004 per string cell (defects_3d.f90:137-162, 465) and axg_field_covariance 4 N^3
 * per realisation run (spectrum.f90:592-630); each advances the position by what it took, as the
 * reference's stream0 advances.  The seed arguments of those calls are then unused.  Draws equal
 * MKL's: uniforms bit for bit, Gaussians to MKL's own ln/sin accuracy (~1e-8 relative).  Without
 * this call the handle uses its counter-based Philox streams keyed by the calls' seeds. */
int axg_set_stream(axg_handle* h, uint32_t seed1, uint32_t seed2, uint64_t position);
/* The position of the stream's next draw (AXG_ESTATE-like AXG_EINVAL when no stream is set). */
int axg_get_stream(axg_handle* h, uint64_t* position);

/* Seeded flat-spectrum initial field: complex Gaussian modes (Philox4x32-10 + Box-Muller)
 * for |k| < kmax, zero above, BACKWARD FFT, rescaled to <|phi|^2> = 1; pi = 0. */
int axg_init_flat(axg_handle* h, uint64_t seed, double kmax);

/* Thermal initial state before the transition (get_initial_thermal, initial.f90:80-238, with
 * initial_phi_phidot 372-390): every mode of phi and phidot gets dv_wav times a unit complex
 * Gaussian times its Bose-Einstein amplitude at temperature thrm_sigma/tau_ini, then the
 * unnormalised FORWARD transform.  The Gaussians come from the reference's stream when
 * axg_set_stream was called (draw for draw as the reference's), otherwise from Philox4x32-10
 * keyed by seed, one draw per mode.  Needs 0 < tau_ini < 1 (AXG_EINVAL otherwise: at
 * tau_ini >= 1 the reference's zero mode is not finite). */
int axg_init_thermal(axg_handle* h, uint64_t seed, double tau_ini);

/* GW accumulators arr_grav (grav_waves.f90:12,74) on the half spectrum kx = 0..Ng/2 of the GW
 * grid, Ng = N / 2^degrade_spec: count = 4 * Ng * Ng * (Ng/2+1) complex, ordered [m][kz][ky][kx]
 * with m = (1,1),(2,1),(1,2),(2,2) of arr_grav(a,b,...), interleaved (re, im).  The other half
 * is the complex conjugate (exact for every mode that reaches a spectrum bin).  With
 * degrade_spec > 0 they are FP64 whatever the field precision. */
int axg_get_gw_state(axg_handle* h, double* A, int64_t count);
int axg_set_gw_state(axg_handle* h, const double* A, int64_t count);

/* symplectic_integration(t1, t2) (evolve.f90:238-257): get_dtau (or dtau_fixed), a drift
 * of dt/2, then n x (kick at the midpoint with the GW accumulation of grav_waves.f90:193-255
 * when grav_spec, drift dt), the last drift being dt/2.  Returns the step count.  With
 * degrade_spec > 0 the GW source is data1_to_arr_dphi_degrade's block-summed gradient
 * (grav_waves.f90:358-507) and the accumulation runs on the GW grid. */
int axg_advance(axg_handle* h, double t1, double t2, int32_t* nsteps_out);
/* The same DKD schedule with an explicit step: nsteps steps of dt from t1.  dt < 0 runs it
 * backwards (t1 + nsteps*dt > 0): the leapfrog with the Crank-Nicolson friction of
 * evolve.f90:299-303 is time-reversible, so forward then backward returns the fields to
 * round-off (the reversibility residual of the energy diagnostics). */
int axg_advance_steps(axg_handle* h, double t1, double dt, int32_t nsteps);

/* grav_spec_estimation (grav_waves.f90:257-321) without the file write: qmid, power,
 * numer, denom per bin (nbins must equal num_bin_grav) on the GW grid (bin edges from
 * pi / dspa_grav, dspa_grav = dspa 2^degrade_spec).  Any pointer may be NULL. */
int axg_gw_spectrum(axg_handle* h, double* qmid, double* power, double* numer, double* denom,
                    int32_t nbins);

typedef struct axg_defect_out {
    int64_t n_string;   /* cells flagged string (defects_3d.f90:340-343) */
    int64_t n_wall;     /* cells flagged wall (defects_3d.f90:334-336) */
    int64_t n_mask;     /* cells flagged mask (defects_3d.f90:344-347) */
} axg_defect_out;

/* id_defects classification pass (defects_3d.f90:67-103, 309-349): flags every cell,
 * counts them and compacts the mask list on the device.  Flags stay resident for the
 * field spectrum. */
int axg_defects(axg_handle* h, axg_defect_out* out);
/* Plaquette windings (no counterpart in the reference as a diagnostic; its plaquette test is the
 * quadrant count of velocity_string_YY, defects_3d.f90:850-879).  Face o = 0, 1, 2 of the site
 * (x, y, z) is the elementary square spanned by (u, v) = (x, y), (y, z), (z, x) from that site
 * (the reference's ipla = 1, 2, 3), looped s -> s+u -> s+u+v -> s+v -> s, counterclockwise about
 * its normal u x v.  Its winding n is the number of times the shorter arcs between consecutive
 * corners' phases cross arg = pi counterclockwise minus clockwise (= sum of the wrapped phase
 * differences / 2 pi; an antipodal pair of corners counts 0), decided by sign tests only, so the
 * counts are exact integers.  n_quadrant counts the faces whose four corners' phases occupy at
 * least three quadrants ceiling(2 ph / pi) + 2 (the reference's penetration test).  On a periodic
 * lattice net[o] = 0, and the windings out of the six faces of every cell sum to 0. */
typedef struct axg_plaquette_out {
    int64_t n_pierced[3];   /* faces with n != 0, per orientation */
    int64_t n_plus[3];      /* faces with n = +1 */
    int64_t n_minus[3];     /* faces with n = -1 (|n| = 2 is counted in n_pierced only) */
    int64_t n_quadrant[3];  /* faces passing the quadrant test */
    int64_t net[3];         /* sum of n */
} axg_plaquette_out;

int axg_plaquettes(axg_handle* h, axg_plaquette_out* out);
/* The pierced faces of the last axg_plaquettes, compacted on the device in site order (z, then y,
 * then x fastest) and face order: 5 int32 per face, (ix, iy, iz) 1-based, o + 1, n.  cap = room in
 * faces; *count = faces written; rec == NULL or cap == 0: size query. */
int axg_get_plaquettes(axg_handle* h, int32_t* rec, int64_t cap, int64_t* count);

/* Compacted mask cells in the reference's loop order (z, then y, then x fastest), as
 * 1-based (ix, iy, iz) triplets, the '(3I6)' rows of _mask_<idx>-<rank>.dat
 * (defects_3d.f90:265-273).  cap = room in ixyz (triplets); *count = cells written.  With
 * ixyz == NULL or cap == 0 only *count is set (size query). */
int axg_get_mask(axg_handle* h, int32_t* ixyz, int64_t cap, int64_t* count);
/* The per-cell flag bits (1 string, 2 wall, 4 mask), N^3 int8 in site order. */
int axg_get_flags(axg_handle* h, int8_t* flags, int64_t nsite);

typedef struct axg_string_out {
    int64_t n_string;   /* string cells fitted */
    int64_t n_vel;      /* strings with |v| < 1 (the ones add_string_length averages over) */
    double length;      /* sum of the chord lengths sl over the string cells, in lattice units:
                           id_defects' sl(1) = sl(2) = sl(3) with group_strings = F */
    double vel[3];      /* mean |v|, mean v^2, mean 1/sqrt(1-v^2) over those (svel) */
} axg_string_out;

/* String length and velocity (defects_3d.f90:167-205 with position_string 435-603,
 * calc_chi2_string 384-433, string_length 1027-1056, velocity_string 605-690 of the VEL = 0
 * build -- or velocity_string_YY 824-982 of the VEL = 1 build when params.string_vel = 1 -- and
 * add_string_length 1654-1733 with group_strings = F), one GPU thread per string
 * cell of the last axg_defects.  Each fit draws its 1001 starting points from the reference's
 * stream when axg_set_stream was called (string s in site order at position + 4004 s, as the
 * reference positions it), otherwise from Philox4x32-10 keyed by (seed0, seed1) with the
 * string's index in site order as counter.  The driver's string parameter is
 * (tau/dspa)^2 * length / N^3 / 4 (evolve.f90:191). */
int axg_strings(axg_handle* h, uint32_t seed0, uint32_t seed1, axg_string_out* out);
/* The per-string records of the last axg_strings: 11 doubles per string, pos(3), dir(3), gof,
 * sl, vel(3), in site order (the arr_string list of defects_3d.f90:8-20).  pos is the point of
 * the fitted line nearest the cell centre plus ixyz - 1/2 with ixyz the cell's 1-based corner
 * (defects_3d.f90:202), i.e. in site coordinates counted from 0.  cap = room in strings;
 * *count = strings written; rec == NULL or cap == 0: size query. */
int axg_get_strings(axg_handle* h, double* rec, int64_t cap, int64_t* count);

typedef struct axg_wall_out {
    int64_t n_wall;     /* wall cells fitted */
    double area;        /* sum of the cells' wall areas wa, in lattice units: id_defects' wa(1) =
                           wa(2) = wa(3) with group_walls = F */
} axg_wall_out;

/* Wall area (defects_3d.f90:183-188 with position_wall 1058-1202 and add_wall_area 2145-2204,
 * group_walls = F), one GPU thread per wall cell of the last axg_defects: the points where
 * sin(ph) changes sign on the cell's edges, the plane normal from the reference's 20 x 20 grid
 * search, the polygon area in that plane.  No random numbers: the GPU and the reference
 * algorithm agree to round-off.  A wall cell with fewer than 3 such edges is the reference's
 * stop: AXG_ENUMERIC (after *out is filled).  The driver's wall parameter is
 * (tau/dspa) * area / N^3 / 2 (evolve.f90:192). */
int axg_walls(axg_handle* h, axg_wall_out* out);
/* The per-wall records of the last axg_walls: 8 doubles per wall, pos(3) (the fitted plane's
 * point nearest the cell centre plus ixyz - 1/2, site coordinates counted from 0), dir(3) (the
 * normal), gof, wa, in site order (the arr_wall list of defects_3d.f90:8-20).  Size query as
 * axg_get_strings. */
int axg_get_walls(axg_handle* h, double* rec, int64_t cap, int64_t* count);

/* Networks of strings or walls (grouping_strings / grouping_walls, defects_3d.f90:1235-1260,
 * 1441-1652, 1735-1760, with the grouped sums of add_string_length 1654-1733 and add_wall_area
 * 2145-2204).  The points (the records of the last axg_strings / axg_walls) fall into the
 * connected components of the graph linking two points at most th_dist apart in the periodic
 * distance lat_dist_periodic, th_dist = max(2, (width^2 hor_size)^(1/3)).  Arguments as the
 * reference passes them: width = width_cr/dspa/tau (strings) or thickness_cr/dspa/tau (walls),
 * hor_size = tau/dspa.  sum[0] = the total length (area); sum[1], sum[2] = the total over groups
 * longer than 2 hor_size and 2 pi hor_size (walls: larger than pi hor_size^2 and 4 pi
 * hor_size^2).  Groups are numbered 1.. in the order of their first record (site order).  The
 * reference's single-rank run also merges groups through its cross-rank step (its pair index
 * assumes the two ranks differ); this build gives the components the multi-rank run finds. */
typedef struct {
    int64_t n_group;
    double th_dist;
    double sum[3];
} axg_group_out;

int axg_group_strings(axg_handle* h, double width, double hor_size, axg_group_out* out);
int axg_group_walls(axg_handle* h, double thickness, double hor_size, axg_group_out* out);
/* Group number of each record of the last axg_group_strings (kind 0) or axg_group_walls
 * (kind 1), in record order.  Size query as axg_get_strings. */
int axg_get_groups(axg_handle* h, int32_t kind, int32_t* ids, int64_t cap, int64_t* count);

/* spectrum_estimation's deterministic part (spectrum.f90:242-286 with compute_matrix_deconv
 * 344-444 and compute_conv_spec 466-567): the masked pi spectrum deconvolved with the mask
 * window.  tau is tau_id(idx).  power = the deconvolved spectrum, conv = conv_spec (already
 * times tau^2), num = num_conv_spec; rho = the compute_crho energy (spectrum.f90:316-342, its
 * first output).  Needs axg_defects first when spec_est (the mask); nbins must equal num_bin.
 * With degrade_spec > 0 it runs on the grid N / 2^degrade_spec: pi's block means
 * (data2_to_data_spec_degrade, spectrum.f90:736-898) and a coarse cell masked when any of its
 * lattice cells is (masking, spectrum.f90:900-954); axg_field_covariance follows it there
 * (4 (N / 2^degrade_spec)^3 stream draws per realisation). */
int axg_field_spectrum(axg_handle* h, double tau, double* qmid, double* power, double* conv,
                       double* num, double* rho, int32_t nbins);

/* estimate_covmat (spectrum.f90:569-734; spectrum_estimation calls it with min_mc = 10,
 * max_mc = 100, tol = 0.1): Monte Carlo covariance of the power of the last
 * axg_field_spectrum (an ESTATE error if the fields changed since).  Each realisation draws
 * Gaussian modes of that power (from the reference's stream when axg_set_stream was called,
 * realisation after realisation as the reference draws them; otherwise Philox4x32-10 keyed by
 * seed), masks them in real space and deconvolves.
 * cov: the packed-upper num_bin x num_bin matrix (ncov = num_bin (num_bin+1)/2 entries, pair
 * (i, j <= i) at i(i-1)/2 + j - 1), the reference's power_cov(:, idx).  iters: the number of
 * realisations when the jackknife test met tol, else -(max_mc + 1), the count the reference
 * divides by in that case.  Clobbers nothing but the work buffer. */
int axg_field_covariance(axg_handle* h, uint64_t seed, int32_t min_mc, int32_t max_mc, double tol,
                         double* cov, int64_t ncov, int32_t* iters);

/* Gradient of a band of the GW spectrum by the adjoint of the leapfrog + GW step (no
 * counterpart in the reference; the step it differentiates is evolve.f90:238-339 with
 * grav_waves.f90:193-255).  Runs nsteps DKD steps of dt from the current fields at t1 exactly as
 * axg_advance_steps does, with the GW accumulators starting from zero, keeping (phi, pi) on the
 * device before kicks 0, K, 2K, ...; then J = the sum of the axg_gw_spectrum power over the bins
 * with qlo <= qmid <= qhi, and a backward sweep that undoes the steps by time reversal (each
 * segment restarts from its checkpoint, bounding the reversal drift) while it propagates the
 * adjoints.  dJ_dsigma2: d/d(sigma_cr^2) with dspa and the spectrum's prefactor fixed (through
 * the kick's force, the tilt and the GW weight; for V = lambda/4 (|phi|^2 - 1)^2 in these
 * units sigma_cr^2 = lambda/4); dJ_dkappa0: d/d(potential_kappa0); dJ_dlnamp: d/d(ln A) for
 * the initial (phi, pi) scaled by A.  On return the fields and accumulators are those at the
 * end of the run; the adjoints of the initial fields are available from axg_get_adjoint.
 * forward_ms / adjoint_ms: wall time of the two sweeps.  Needs degrade_spec = 0. */
typedef struct {
    double J;
    double dJ_dsigma2;
    double dJ_dkappa0;
    double dJ_dlnamp;
    double forward_ms;
    double adjoint_ms;
} axg_grad_out;

int axg_gradient(axg_handle* h, double t1, double dt, int32_t nsteps, int32_t K, double qlo, double qhi,
                 axg_grad_out* out);
/* dJ/dRe + i dJ/dIm of the initial phi and pi of the last axg_gradient, as N^3 complex
 * (re, im) pairs, x fastest; either pointer may be NULL. */
int axg_get_adjoint(axg_handle* h, double* phibar, double* pibar, int64_t nsite);

/* Batches of lattices (no counterpart in the reference; the reference runs one lattice per
 * job).  Advances n independent lattices -- any sizes and precisions, one device -- by nsteps
 * DKD steps of dt from t1, each exactly as axg_advance_steps would (bit-identical fields and
 * accumulators).  Every step's scalars go to a device table per lattice first.  Lattices that
 * share N, precision, dspa, dwav and the GW index order, with the fused pass, are stacked: each
 * pass of a step is one launch over all of them (the lattice index in the grid, each reading its
 * own buffers and table entry).  The others step one by one.  One HIP graph of `chunk` steps of
 * all the stacks and lattices (dealt to two streams forked from and joined into the first
 * handle's; environment variable AXG_BATCH_LANES sets another count) is captured once and
 * replayed, a device step counter advancing once per replay; the steps left over, and chunk = 0,
 * launch the same kernels directly.  AXG_BATCH_STACK=0 turns stacking off (same results).
 * chunk: 0 or an even number from 2 to 64.  *ms (may be NULL): wall time of the steps.  Errors
 * are reported on hs[0].  Handles with degrade_spec > 0 are refused (AXG_EINVAL): advance them
 * with axg_advance_steps. */
int axg_batch_advance(axg_handle* const* hs, int32_t n, double t1, double dt, int32_t nsteps, int32_t chunk,
                      double* ms);

/* Comoving lattice energy at time tau (FP64 reduction): out = {kinetic sum|pi|^2,
 * gradient sum|grad phi|^2, potential tau^2 sigma_cr^2 sum V, total}, each times dspa^3. */
int axg_energy(axg_handle* h, double tau, double out[4]);

/* axg_energy on fields the caller holds in device memory, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  phi and pi: device pointers to
 * N^3 FP64 complex (re, im) values, x fastest, N = p->ngrid (2..1024; the physics is read from
 * p: dspa, sigma_cr, kappa, initial_loop, potential_kappa0, device).  out: device pointer to 4
 * doubles, written as axg_energy writes them and bit-identical to it on the same FP64 fields.
 * Launches on `stream` (a hipStream_t; NULL = the default stream) and returns without
 * synchronizing.  The reduction's partial sums go to a fixed 64 KiB device array of the library
 * (1024 blocks x 3 partials + 3 sums, whatever N), held by every device it runs on until the
 * process ends; calls that overlap on one device race on it.  Errors: as the other entry
 * points, with the message in axg_last_error(NULL). */
int axg_energy_fields(const axg_params* p, const double* phi, const double* pi, double tau, double* out,
                      void* stream);

/* Autotuner (no counterpart in the reference).  Times each configuration of the fused
 * kick / GW source / x-transform pass (block rows, sites per thread) and of the y transform
 * (column-tile width) on this handle's lattice, reps launches each, and keeps the fastest for
 * the handle's later steps.  Every configuration gives bit-identical results; the fields are
 * left as they were (pi is saved and restored; the other buffers written are rewritten by
 * every step).  *variant = fused-pass code | (y-pass code << 8); ms[i] (i < nms, may be NULL)
 * = average ms per launch of: fused pass rows in dispatch order, wide blocks (twice the
 * rows), other sites per thread, both, and the same four with XCD strips (each XCD walking its
 * own strip of rows plane by plane); y pass 128-byte tiles, 64-byte tiles, 128-byte tiles with two z planes per
 * block (all 0 when the lattice has a single configuration: grav_spec off, or N below 64 or
 * without the compile-time kernels). */
int axg_tune(axg_handle* h, int32_t reps, int32_t* variant, double* ms, int32_t nms);

/* Kernel timing (no counterpart in the reference; for measurement).  enable = 1 resets the
 * counters and starts timing every step kernel with HIP events on the handle's stream;
 * enable = 0 stops.  ms[slot] / launches[slot] receive the totals so far for the slots
 * 0 first drift, 1 kick + GW source + drift, 2 x-transform, 3 y-transform,
 * 4 z-transform + accumulation (nslots <= 5; either pointer may be NULL). */
int axg_profile(axg_handle* h, int32_t enable, double* ms, int64_t* launches, int32_t nslots);

/* Streaming-bandwidth probe of one device's HBM (no counterpart in the reference; for
 * measurement).  Allocates two buffers of `bytes` (>= 1 MiB) and times `reps` launches each of
 * a read-only, a write-only, a copy and a copy with nontemporal hints kernel, 16 bytes per lane;
 * gbs[kind] = bytes moved / time in GB/s (a copy moves 2 x bytes).  Frees everything before
 * returning.  Errors: message in axg_last_error(NULL). */
int axg_hbm_probe(int32_t device, int64_t bytes, int32_t reps, double gbs[4]);

#ifdef __cplusplus
}
#endif
#endif
