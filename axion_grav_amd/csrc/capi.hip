// C-ABI entry points of libaxgrav (declared in include/axgrav.h) and the host-side
// scalars they need.  Every device buffer belongs to the handle; every entry point runs on
// the handle's stream and synchronizes before returning.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include <chrono>

#include "axg_internal.h"

namespace axg {

// ------------------------------------------------------------------ potential scalars

double msquared(double thrm) { return (thrm * thrm - AXG_THRM_SIGMA * AXG_THRM_SIGMA) / 3; }  // potential.f90:34-41

double maxion2(double thrm, double kappa)  // potential.f90:62-79
{
    const double n = 6.68, cT = 6.26, c0 = 1.0;
    double thrm12 = kappa * pow(cT / c0, 1.0 / n);
    double m = thrm > thrm12 ? cT * pow(kappa / thrm, n) : c0;
    double k2 = kappa * kappa;
    return m * (k2 * k2);
}

// evolve.f90:291-305 in the same FP64 operation order
KickScalars kick_scalars(const axg_params& p, double t, double dt)
{
    KickScalars k;
    double hubble = 1 / t;
    double thrm = p.initial_loop ? 0.0 : AXG_THRM_SIGMA / t;
    k.m2 = msquared(thrm);
    k.ma2 = maxion2(thrm, p.kappa);
    double fact0 = 1 + dt * hubble;
    double fact1 = 1 - dt * hubble;
    double fact2 = dt / (p.dspa * p.dspa);
    double fact3 = -(dt * (t * t) * (p.sigma_cr * p.sigma_cr));
    k.f1 = fact1 / fact0;
    k.f2 = fact2 / fact0;
    k.f3 = fact3 / fact0;
    k.kt = tilt_coef(p, t);
    return k;
}

// N_DW=2 tilt V_t = -kappa0 t^4 (phi^2 + phi*^2)/2 (no counterpart in the reference).  The
// reference's dVdphi is scaled so that sigma_cr^2 dVdphi is the force (potential.f90:43-60,
// evolve.f90:303), so the tilt adds -kt conj(phi) with kt = kappa0 t^4 / sigma_cr^2; 0 = off.
double tilt_coef(const axg_params& p, double t)
{
    if (p.potential_kappa0 == 0.0) return 0.0;
    double t2 = t * t;
    return p.potential_kappa0 * (t2 * t2) / (p.sigma_cr * p.sigma_cr);
}

bool supported_n(int64_t n)
{
    if (n < 8 || n > 1024 || (n & 1)) return false;
    int64_t m = n;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    return m == 1;
}

FftPlan make_plan(int n)
{
    FftPlan p{};
    p.n = n;
    int m = n, k = 0;
    while (m % 3 == 0) { p.radix[k++] = 3; m /= 3; }
    while (m % 8 == 0) { p.radix[k++] = 8; m /= 8; }
    if (m % 4 == 0) { p.radix[k++] = 4; m /= 4; }
    if (m % 2 == 0) { p.radix[k++] = 2; m /= 2; }
    p.nst = k;
    return p;
}

}  // namespace axg

using namespace axg;

// ------------------------------------------------------------------ handle

constexpr int AXG_NSLOT = 5;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct axg_handle {
    axg_params p{};
    int N = 0;
    int64_t n3 = 0;
    int P = 0;        // half-spectrum row pitch
    int nh = 0;       // N/2 + 1
    int prec = AXG_FP64;
    bool fp32 = false, comp = false;
    size_t t2 = 16;   // bytes of one stored complex
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false for the reduced-grid handle, which runs on its lattice's stream
    // degrade_spec > 0: deg = 2^degrade_spec and the handle of the reduced grid N/deg that holds
    // the GW state and the axion spectrum machinery (FP64, sharing this handle's stream)
    int deg = 1;
    axg_handle* gs = nullptr;
    FftPlan plan{};
    std::vector<DevBuf> bufs;
    void* phi[2] = {nullptr, nullptr};
    void* cphi[2] = {nullptr, nullptr};
    void* pi = nullptr;
    void* cpi = nullptr;
    void* src = nullptr;
    void* W = nullptr;
    void* W2 = nullptr;      // y-transformed sources in the tiled layout (p2_size(N))
    void* A = nullptr;
    GwLayout L{};            // layout of W2 (or W) and A as the accumulation pass reads them
    void* tw = nullptr;      // twiddles in the storage precision
    void* tw64 = nullptr;    // FP64 twiddles for the diagnostic transforms
    int cur = 0;
    bool fuse_x = true;      // x transform fused into the stencil pass when p2_size(N)
    int kick_variant = 0;    // configuration of the fused pass (axg_tune codes; set per size at creation)
    int ypass_variant = 0;   // configuration of the y pass (axg_tune; 0 = 128-byte column tiles)
    // diagnostics
    uint16_t* gw_bin = nullptr;
    int64_t gw_n2max = 0;     // largest |n|^2 with a GW bin
    uint16_t* spec_bin = nullptr;
    int64_t spec_n2max = 0;   // largest |n|^2 with a field-spectrum bin
    int64_t nn2 = 0;          // entries of the n^2 tables (3 (N/2)^2 + 1)
    double* red = nullptr;    // reduction scratch
    size_t red_bytes = 0;
    double2* work = nullptr;  // N^3 FP64 complex work array (spectra, IC, field copies)
    int8_t* flags = nullptr;
    int64_t* blk = nullptr;   // per-block counts / offsets of the classifier
    int32_t* mask = nullptr;  // compacted mask list (3 ints per cell)
    int64_t nmask = -1;       // -1: no classification yet
    int64_t nstring = 0;      // string cells of the last classification
    bool have_flags = false;
    // the reference's random stream (axg_set_stream): MCG59 start and the position of the next draw
    bool stream_on = false;
    uint64_t stream_x0 = 0, stream_pos = 0;
    // the random source of the next call: the stream at its position, or Philox keyed by (k0, k1)
    Draws draws(uint32_t k0, uint32_t k1) const { return Draws{stream_x0, stream_pos, k0, k1, stream_on ? 1 : 0}; }
    // plaquette windings (grown on demand): face codes, per-block counts / offsets / totals, rows
    uint16_t* plaq_code = nullptr;
    int64_t* plaq_cnt = nullptr;
    int32_t* plaq_rec = nullptr;
    int64_t plaq_rec_cap = 0;
    int64_t nplaq = 0;        // pierced faces of the last axg_plaquettes
    bool have_plaq = false;
    // string fits (grown on demand): cell list, 11 doubles per string, per-block sums
    int32_t* str_cells = nullptr;
    double* str_rec = nullptr;
    double* str_part = nullptr;
    int64_t* str_pend = nullptr;   // string fits' second-stage list (launch_string_fit)
    double* str_state = nullptr;   // its counter (first 8 bytes) and parked searches
    int64_t str_cap = 0;      // strings the five buffers hold
    int string_split = STRING_SPLIT;  // simplex iterations before a search moves to the second kernel
    int64_t nstr_rec = -1;    // strings in str_rec (-1: no fit since the fields changed)
    int64_t nwall = 0;        // wall cells of the last classification
    int32_t* wall_cells = nullptr;
    double* wall_rec = nullptr;
    double* wall_part = nullptr;
    int64_t wall_cap = 0;
    int64_t nwall_rec = -1;
    // grouping scratch (parent, label, value per point; grown on demand) and the last groups
    void* grp = nullptr;
    int64_t grp_cap = 0;
    std::vector<int32_t> grp_ids[2];  // group number (1-based) per string / wall record
    bool have_grp[2] = {false, false};
    bool wall_bad = false;  // the last axg_walls met cells with fewer than 3 edge points
    // adjoint (axg_gradient): phibar, pibar, two more N^3 FP64 complex arrays (the third is work),
    // the accumulator adjoints, and the checkpoints (grown on demand)
    double2* adj = nullptr;
    double2* abar = nullptr;
    void* ckpt = nullptr;
    int64_t ckpt_cap = 0;  // checkpoints the buffer holds
    bool have_adj = false;
    double2* adj_pibar = nullptr;  // which of the two pibar buffers holds the result
    // table-driven steps (axg_batch_advance): per-step scalars, the counter and the current entry
    StepEntry* steptab = nullptr;
    int64_t steptab_cap = 0;
    int* stepbase = nullptr;       // the batch's step counter (on the first handle of a batch)
    LatRef* latref = nullptr;      // the batch's stacked-launch tables (on the first handle)
    int64_t latref_cap = 0;
    double adj_scale = 1;          // the adjoint fields are held divided by this
    // spectrum tables (init_spectrum)
    std::vector<double> wav;  // num_bin+1 edges
    // the last field spectrum: deconvolved power and the count-normalised deconvolution matrix
    // (estimate_covmat's inputs); valid while the fields and flags are unchanged
    std::vector<double> spec_power, spec_mat;
    bool have_spec = false;
    double* d_largeq = nullptr;
    double* d_matspec = nullptr;
    double* d_dqs = nullptr;
    double* d_cnt = nullptr;  // lattice points per n^2
    double* d_hist = nullptr;
    double* d_plane = nullptr;
    // kernel timing
    bool prof = false;
    std::vector<hipEvent_t> ev;      // event pool
    std::vector<int> ev_slot;        // slot of each recorded (start, stop) pair
    size_t ev_used = 0;
    double prof_ms[AXG_NSLOT] = {0};
    int64_t prof_n[AXG_NSLOT] = {0};
    std::string err;

    void* alloc(size_t bytes)
    {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(AXG_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        bufs.push_back({q, bytes});
        return q;
    }
    // Free one buffer of bufs (stream already idle).
    void release_one(void*& q)
    {
        if (!q) return;
        for (size_t i = 0; i < bufs.size(); i++)
            if (bufs[i].p == q) {
                (void)hipFree(q);
                bufs.erase(bufs.begin() + i);
                break;
            }
        q = nullptr;
    }
    // Record a start/stop event pair around launch() on the handle's stream when timing.
    template <typename F>
    void timed(int slot, F&& launch)
    {
        if (!prof) { launch(); return; }
        while (ev.size() < 2 * (ev_used + 1)) {
            hipEvent_t e;
            AXG_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
        AXG_HIP(hipEventRecord(ev[2 * ev_used], stream));
        launch();
        AXG_HIP(hipEventRecord(ev[2 * ev_used + 1], stream));
        if (ev_slot.size() <= ev_used) ev_slot.resize(ev_used + 1);
        ev_slot[ev_used++] = slot;
    }
    // Fold the recorded pairs into the totals (stream already synchronized).
    void collect()
    {
        for (size_t i = 0; i < ev_used; i++) {
            float ms = 0;
            AXG_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            prof_ms[ev_slot[i]] += ms;
            prof_n[ev_slot[i]] += 1;
        }
        ev_used = 0;
    }
    void release()
    {
        if (gs) {
            gs->release();
            delete gs;
            gs = nullptr;
        }
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
        for (auto& b : bufs) (void)hipFree(b.p);
        bufs.clear();
        if (stream && own_stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

static thread_local std::string g_create_err;

namespace axg {
// The message axg_last_error(NULL) returns: the last failure of a call without a handle
// (axg_create, axg_energy_fields) on the calling thread.
void set_handleless_error(const std::string& msg) { g_create_err = msg; }
}  // namespace axg

template <typename F>
static int guarded(axg_handle* h, F&& f)
{
    if (!h) return AXG_EINVAL;
    try {
        AXG_HIP(hipSetDevice(h->p.device));
        f();
        return AXG_OK;
    } catch (const Error& e) {
        h->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        h->err = "host allocation failed";
        return AXG_ENOMEM;
    } catch (...) {
        h->err = "unexpected exception";
        return AXG_EDEVICE;
    }
}

static void require(bool ok, const std::string& msg)
{
    if (!ok) throw Error(AXG_EINVAL, msg);
}

// ------------------------------------------------------------------ settings

extern "C" int axg_derive(const axg_keys* k, axg_params* p, double* zeta, double* width_cr, double* thickness_cr)
{
    if (!k || !p) return AXG_EINVAL;
    // settings.f90:274-297, potential.f90:20-21
    const double tcmb_in_mstar = 2.726 / 1.16045250061657e13 / 2.435e18;
    double a_now = k->sigma * AXG_THRM_SIGMA / tcmb_in_mstar * pow(k->g_rel / 3.9091, 1.0 / 3);
    double tau_cr = 1 / AXG_PI_T / sqrt(k->g_rel / 90.0) / (k->sigma * k->sigma) / (AXG_THRM_SIGMA * AXG_THRM_SIGMA);
    double sigma_cr = tau_cr * k->sigma;
    double dspa = k->boxsize * k->tau_fin / (double)k->ngrid;
    double dwav = AXG_TWOPI_T / dspa / (double)k->ngrid;
    double r = dwav / AXG_TWOPI_T;
    p->ngrid = k->ngrid;
    p->a_now = a_now;
    p->tau_cr = tau_cr;
    p->sigma_cr = sigma_cr;
    p->kappa = k->kappa;
    p->dspa = dspa;
    p->dwav = dwav;
    p->dv_spa = dspa * dspa * dspa;
    p->dv_wav = r * r * r;
    if (zeta) *zeta = sqrt(22.5 / k->g_rel) / k->sigma / AXG_PI_T;
    if (width_cr) *width_cr = 1 / sqrt(2.0) / sigma_cr;
    if (thickness_cr) *thickness_cr = 1 / sqrt(2.0) / (sqrt(maxion2(1.e-3 * k->kappa, k->kappa)) * sigma_cr);
    return AXG_OK;
}

static void get_dtau(const axg_params& p, double t1, double t2, double* dt, int32_t* n)
{
    if (p.dtau_fixed > 0) {
        // fixed step: the largest step <= dtau_fixed that divides the interval evenly
        int32_t nn = (int32_t)ceil((t2 - t1) / p.dtau_fixed - 1e-9);
        if (nn < 1) nn = 1;
        *n = nn;
        *dt = (t2 - t1) / nn;
        return;
    }
    // evolve.f90:10-24
    double omega = AXG_THRM_SIGMA * t2 * p.sigma_cr;
    double kmax = p.dwav * (double)p.ngrid;
    omega = sqrt(kmax * kmax + omega * omega);
    double d = p.eps_dtau / omega;
    int32_t nn = (int32_t)ceil((t2 - t1) / d);
    *n = nn;
    *dt = (t2 - t1) / nn;
}

extern "C" int axg_get_dtau(const axg_params* p, double t1, double t2, double* dtau, int32_t* nsteps)
{
    if (!p || !dtau || !nsteps || !(t2 > t1)) return AXG_EINVAL;
    if (p->dtau_fixed <= 0 && !(p->eps_dtau > 0)) return AXG_EINVAL;
    get_dtau(*p, t1, t2, dtau, nsteps);
    return AXG_OK;
}

// Bin of a mode with |n|^2 = n2, in the reference's FP64 order: q = sqrt(n2)*dwav,
// i = nint(q/dq + 0.5) with dq = wave(1) - wave(0) (grav_waves.f90:62-71,265,279;
// spectrum.f90:63-72,476,538).  0 = outside 1..nb.
static void bin_table(int64_t nn2, int nb, double dspa, double dwav, std::vector<uint16_t>& out,
                      std::vector<double>* edges)
{
    double qmax = AXG_PI_T / dspa;
    double dq0 = qmax / nb;
    std::vector<double> wave(nb + 1);
    wave[0] = 0;
    for (int b = 1; b <= nb; b++) wave[b] = wave[b - 1] + dq0;
    double dq = wave[1] - wave[0];
    out.assign(nn2, 0);
    for (int64_t n2 = 0; n2 < nn2; n2++) {
        double q = sqrt((double)n2) * dwav;
        long i = lround(q / dq + 0.5);
        out[n2] = (i >= 1 && i <= nb) ? (uint16_t)i : 0;
    }
    if (edges) *edges = wave;
}

// spectrum.f90:117-171 integrate_p_check (3D branch)
static double integrate_p(int nint_, double li, double si, double lj, double sj, double q)
{
    double maxp = std::min(lj, q + li);
    double minp = std::max(q - li, si - q);
    minp = std::max(minp, sj);
    double dp = (maxp - minp) / nint_;
    double x = 0;
    if (dp > 0) {
        for (int ip = 1; ip <= nint_; ip++) {
            double pp = dp * (ip - 0.5) + minp;
            double maxmu, minmu;
            if (q == 0.0) { maxmu = 1; minmu = -1; }
            else {
                maxmu = (li * li - pp * pp - q * q) / 2 / pp / q;
                minmu = (si * si - pp * pp - q * q) / 2 / pp / q;
                maxmu = std::min(1.0, maxmu);
                minmu = std::max(-1.0, minmu);
            }
            x = x + pp * pp * dp * (maxmu - minmu);
        }
    }
    return x * AXG_TWOPI_T;
}

// ------------------------------------------------------------------ create / destroy

// The reduced grid of degrade_spec > 0 as init_grav and init_spectrum set it up
// (grav_waves.f90:37-45, spectrum.f90:42-49): ngrid / deg sites per axis of spacing dspa * deg,
// the wavenumber unit and cell volumes from those in the reference's expressions.
static axg_params reduced_params(const axg_params& p)
{
    axg_params q = p;
    const int deg = 1 << p.degrade_spec;
    q.ngrid = p.ngrid / deg;
    q.precision = AXG_FP64;
    q.degrade_spec = 0;
    q.dspa = p.dspa * deg;
    q.dwav = AXG_TWOPI_T / q.dspa / (double)q.ngrid;
    q.dv_spa = q.dspa * q.dspa * q.dspa;
    const double r = q.dwav / AXG_TWOPI_T;
    q.dv_wav = r * r * r;
    return q;
}

static void make_reduced_grid(axg_handle* h)
{
    const axg_params q = reduced_params(h->p);
    axg_handle* g = nullptr;
    const int rc = axg_create(&q, &g);
    if (rc != AXG_OK) throw Error(rc, "reduced grid: " + g_create_err);
    h->gs = g;
    // the packed source planes the degraded gradient fills (the x transform is its own pass)
    g->fuse_x = false;
    if (q.grav_spec && !g->src) g->src = g->alloc(3 * g->n3 * g->t2);
    // one stream for both grids, so the coarse passes follow the lattice's kernels in order
    AXG_HIP(hipStreamSynchronize(g->stream));
    AXG_HIP(hipStreamDestroy(g->stream));
    g->stream = h->stream;
    g->own_stream = false;
}

extern "C" int axg_create(const axg_params* pin, axg_handle** out)
{
    g_create_err.clear();
    if (!pin || !out) { g_create_err = "null argument"; return AXG_EINVAL; }
    *out = nullptr;
    std::unique_ptr<axg_handle> h(new (std::nothrow) axg_handle());
    if (!h) { g_create_err = "host allocation failed"; return AXG_ENOMEM; }
    h->p = *pin;
    const axg_params& p = h->p;
    try {
        require(supported_n(p.ngrid), "ngrid must be even, 2^a 3^b, 8..1024");
        require(p.precision == AXG_FP64 || p.precision == AXG_FP32C || p.precision == AXG_FP32,
                "precision must be AXG_FP64, AXG_FP32C or AXG_FP32");
        require(p.dspa > 0 && p.dwav > 0 && p.dv_spa > 0 && p.dv_wav > 0 && p.sigma_cr > 0, "derived units missing");
        require(p.num_bin_grav >= 1 && p.num_bin_grav <= 4096, "num_bin_grav out of range");
        require(p.string_vel == 0 || p.string_vel == 1, "string_vel must be 0 (VEL = 0 build) or 1 (VEL = 1)");
        require(p.degrade_spec >= 0 && p.degrade_spec <= 7 && supported_n(p.ngrid >> p.degrade_spec) &&
                    ((p.ngrid >> p.degrade_spec) << p.degrade_spec) == p.ngrid,
                "degrade_spec must leave an even 2^a 3^b grid of at least 8 (ngrid / 2^degrade_spec)");
        require(!p.spec_est || (p.num_bin >= 1 && p.num_bin <= 1024), "num_bin out of range");
        AXG_HIP(hipSetDevice(p.device));
        h->N = (int)p.ngrid;
        const int N = h->N;
        h->n3 = (int64_t)N * N * N;
        h->nh = N / 2 + 1;
        h->P = half_pitch(N);
        h->prec = p.precision;
        h->fp32 = p.precision != AXG_FP64;
        h->comp = p.precision == AXG_FP32C;
        h->t2 = h->fp32 ? 8 : 16;
        h->plan = make_plan(N);
        // tuning knob: AXG_XFUSE=0 runs the x transform as its own pass (same results)
        if (const char* e = getenv("AXG_XFUSE")) h->fuse_x = strcmp(e, "0") != 0;
        // the fused pass's configuration before any axg_tune, from the tuner's picks on MI355X:
        // XCD strips at N = 128, 256, 384, 768 and 1024; at 512 wide blocks with strips for FP32
        // storage and the dispatcher's order for FP64; code 0 elsewhere (within 10 % of the best)
        if (N == 512) h->kick_variant = h->fp32 ? 9 : 0;
        else if (N == 128 || N == 256 || N == 384 || N >= 768) h->kick_variant = 8;
        // the y pass at 768 and 1024 in FP32: two columns per thread over 128-byte lines, two z
        // planes per block (1024^3: 12.7 ms against 16.5 with one plane, 19.9 for the one-column tile)
        if (h->fp32 && (N == 768 || N == 1024)) h->ypass_variant = 2;
        // tuning knobs AXG_KICK_VARIANT / AXG_YPASS_VARIANT: initial configurations (axg_tune's codes)
        if (const char* e = getenv("AXG_KICK_VARIANT")) h->kick_variant = atoi(e);
        if (const char* e = getenv("AXG_YPASS_VARIANT")) h->ypass_variant = atoi(e);
        // AXG_STRING_SPLIT: simplex iterations of the string fits' first kernel (same records)
        if (const char* e = getenv("AXG_STRING_SPLIT")) h->string_split = std::max(0, atoi(e));
        // degrade_spec > 0: the GW passes run on the reduced grid's handle (made below); the
        // lattice's own step is the plain kick + drift
        const bool reduced = p.degrade_spec > 0 && (p.grav_spec || p.spec_est);
        if (reduced) {
            h->deg = 1 << p.degrade_spec;
            h->fuse_x = false;
        }
        AXG_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        const int64_t n3 = h->n3;
        for (int i = 0; i < 2; i++) h->phi[i] = h->alloc(n3 * h->t2);
        h->pi = h->alloc(n3 * h->t2);
        if (h->comp) {
            for (int i = 0; i < 2; i++) h->cphi[i] = h->alloc(n3 * 4);
            h->cpi = h->alloc(n3 * 4);
        }
        if (p.grav_spec && !reduced) {
            // the packed source planes only exist when the x transform is not fused
            if (!p2_size(N) || !h->fuse_x) h->src = h->alloc(3 * n3 * h->t2);
            h->L = gw_layout(N, h->fp32);
            h->W = h->alloc((size_t)6 * N * N * h->P * h->t2);
            if (h->L.zt) h->W2 = h->alloc((size_t)6 * h->L.comp() * h->t2);
            h->A = h->alloc((size_t)4 * h->L.comp() * h->t2);
            AXG_HIP(hipMemsetAsync(h->A, 0, (size_t)4 * h->L.comp() * h->t2, h->stream));
        }
        // twiddles e^{+2 pi i m / N}
        {
            std::vector<double> t64(2 * N);
            std::vector<float> t32(2 * N);
            const long double pl = 3.141592653589793238462643383279502884L;
            for (int m = 0; m < N; m++) {
                long double a = 2.0L * pl * (long double)m / (long double)N;
                t64[2 * m] = (double)cosl(a);
                t64[2 * m + 1] = (double)sinl(a);
                t32[2 * m] = (float)cosl(a);
                t32[2 * m + 1] = (float)sinl(a);
            }
            h->tw64 = h->alloc(N * 16);
            AXG_HIP(hipMemcpy(h->tw64, t64.data(), N * 16, hipMemcpyHostToDevice));
            if (h->fp32) {
                h->tw = h->alloc(N * 8);
                AXG_HIP(hipMemcpy(h->tw, t32.data(), N * 8, hipMemcpyHostToDevice));
            } else {
                h->tw = h->tw64;
            }
        }
        h->nn2 = 3 * (int64_t)(N / 2) * (N / 2) + 1;
        {
            std::vector<uint16_t> tab;
            bin_table(h->nn2, p.num_bin_grav, p.dspa, p.dwav, tab, nullptr);
            for (int64_t n2 = 0; n2 < h->nn2; n2++)
                if (tab[n2]) h->gw_n2max = n2;
            h->gw_bin = (uint16_t*)h->alloc(h->nn2 * 2);
            AXG_HIP(hipMemcpy(h->gw_bin, tab.data(), h->nn2 * 2, hipMemcpyHostToDevice));
            if (p.spec_est) {
                bin_table(h->nn2, p.num_bin, p.dspa, p.dwav, tab, &h->wav);
                for (int64_t n2 = 0; n2 < h->nn2; n2++)
                    if (tab[n2]) h->spec_n2max = n2;
                h->spec_bin = (uint16_t*)h->alloc(h->nn2 * 2);
                AXG_HIP(hipMemcpy(h->spec_bin, tab.data(), h->nn2 * 2, hipMemcpyHostToDevice));
            }
        }
        h->red_bytes = (size_t)2048 * 2 * (std::max(p.num_bin_grav, std::max(p.num_bin, 4)) + 8) * sizeof(double);
        h->red = (double*)h->alloc(h->red_bytes);
        h->work = (double2*)h->alloc(n3 * 16);
        h->flags = (int8_t*)h->alloc(n3);
        int64_t nblk = (n3 + 255) / 256;
        // counts [nblk][3], offsets [nblk] + totals, string / wall offsets [nblk] + totals, scan scratch
        h->blk = (int64_t*)h->alloc((6 * nblk + 16 + scan_scratch(nblk, 3)) * sizeof(int64_t));
        if (reduced) make_reduced_grid(h.get());
        AXG_HIP(hipStreamSynchronize(h->stream));
    } catch (const Error& e) {
        g_create_err = e.msg;
        h->release();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_create_err = "host allocation failed";
        h->release();
        return AXG_ENOMEM;
    }
    *out = h.release();
    return AXG_OK;
}

extern "C" int axg_destroy(axg_handle* h)
{
    if (!h) return AXG_EINVAL;
    (void)hipSetDevice(h->p.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->release();
    delete h;
    return AXG_OK;
}

extern "C" const char* axg_last_error(const axg_handle* h)
{
    if (!h) return g_create_err.c_str();
    return h->err.c_str();
}

// ------------------------------------------------------------------ fields

static void store_from_work(axg_handle* h, void* f, void* cf)
{
    if (h->comp) launch_store_field<float, true>(h->stream, h->N, h->work, f, cf);
    else if (h->fp32) launch_store_field<float, false>(h->stream, h->N, h->work, f, cf);
    else launch_store_field<double, false>(h->stream, h->N, h->work, f, cf);
}

static void load_to_work(axg_handle* h, const void* f, const void* cf)
{
    if (h->comp) launch_load_field<float, true>(h->stream, h->N, f, cf, h->work);
    else if (h->fp32) launch_load_field<float, false>(h->stream, h->N, f, cf, h->work);
    else launch_load_field<double, false>(h->stream, h->N, f, cf, h->work);
}

extern "C" int axg_set_fields(axg_handle* h, const double* phi, const double* pi, int64_t nsite)
{
    return guarded(h, [&] {
        require(nsite == h->n3, "nsite must be ngrid^3");
        size_t bytes = (size_t)h->n3 * 16;
        if (phi) {
            AXG_HIP(hipMemcpyAsync(h->work, phi, bytes, hipMemcpyHostToDevice, h->stream));
            store_from_work(h, h->phi[h->cur], h->cphi[h->cur]);
        }
        if (pi) {
            AXG_HIP(hipMemcpyAsync(h->work, pi, bytes, hipMemcpyHostToDevice, h->stream));
            store_from_work(h, h->pi, h->cpi);
        }
        h->have_flags = false;
        h->have_plaq = false;
        h->have_spec = false;
        h->nstr_rec = -1;
        h->nwall_rec = -1;
        h->have_grp[0] = h->have_grp[1] = false;
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

extern "C" int axg_get_fields(axg_handle* h, double* phi, double* pi, int64_t nsite)
{
    return guarded(h, [&] {
        require(nsite == h->n3, "nsite must be ngrid^3");
        size_t bytes = (size_t)h->n3 * 16;
        if (phi) {
            load_to_work(h, h->phi[h->cur], h->cphi[h->cur]);
            AXG_HIP(hipMemcpyAsync(phi, h->work, bytes, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
        if (pi) {
            load_to_work(h, h->pi, h->cpi);
            AXG_HIP(hipMemcpyAsync(pi, h->work, bytes, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

// ------------------------------------------------------------------ GW state

// The result of an entry point run on the reduced grid's handle, reported as h's.
static void forward(axg_handle* h, int rc)
{
    if (rc != AXG_OK) throw Error(rc, h->gs->err);
}

extern "C" int axg_get_gw_state(axg_handle* h, double* A, int64_t count)
{
    return guarded(h, [&] {
        if (h->gs) return forward(h, axg_get_gw_state(h->gs, A, count));
        const int N = h->N;
        require(h->A != nullptr, "grav_spec is off: no GW state");
        require(A && count == (int64_t)4 * N * N * h->nh, "count must be 4*N*N*(N/2+1)");
        const size_t per = (size_t)N * N * h->nh;  // complex per component
        for (int m = 0; m < 4; m++) {
            if (h->fp32) launch_gw_pack<float>(h->stream, N, h->L, h->A, m, h->work, +1);
            else launch_gw_pack<double>(h->stream, N, h->L, h->A, m, h->work, +1);
            AXG_HIP(hipMemcpyAsync(A + 2 * per * m, h->work, per * 16, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

extern "C" int axg_set_gw_state(axg_handle* h, const double* A, int64_t count)
{
    return guarded(h, [&] {
        if (h->gs) return forward(h, axg_set_gw_state(h->gs, A, count));
        const int N = h->N;
        require(h->A != nullptr, "grav_spec is off: no GW state");
        require(A && count == (int64_t)4 * N * N * h->nh, "count must be 4*N*N*(N/2+1)");
        const size_t per = (size_t)N * N * h->nh;
        for (int m = 0; m < 4; m++) {
            AXG_HIP(hipMemcpyAsync(h->work, A + 2 * per * m, per * 16, hipMemcpyHostToDevice, h->stream));
            if (h->fp32) launch_gw_pack<float>(h->stream, N, h->L, h->A, m, h->work, -1);
            else launch_gw_pack<double>(h->stream, N, h->L, h->A, m, h->work, -1);
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
    });
}

// ------------------------------------------------------------------ time stepping

// GW sourcing on the reduced grid (degrade_spec > 0; grav_waves.f90:202-206, 358-507) for the
// kick at t with step dt: the block sums of the gradient of phi at t (phi[cur], which the
// out-of-place kick + drift has not overwritten) into the coarse handle's packed source planes,
// then the coarse grid's x, y and z passes with dw = tau dtau dv_spa_grav sigma_cr^2.
template <typename T, bool COMP>
static void reduced_gw_step(axg_handle* h, double t, double dt)
{
    axg_handle* g = h->gs;
    h->timed(2, [&] {
        launch_degrade_source<T, COMP>(h->stream, h->N, h->deg, h->phi[h->cur], h->cphi[h->cur], h->p.dspa,
                                       (double2*)g->src);
        launch_gw_xpass<double>(g->stream, g->N, g->P, g->src, g->W, g->plan, g->tw);
    });
    h->timed(3, [&] {
        launch_gw_ypass<double>(g->stream, g->N, g->P, g->W, g->W2, g->L, g->plan, g->tw, g->ypass_variant);
    });
    GwStepScalars s;
    s.tau = t;
    s.dw = t * dt * g->p.dv_spa * (g->p.sigma_cr * g->p.sigma_cr);
    s.dwav = g->p.dwav;
    s.fix_swap = g->p.gw_fix_swap;
    const void* wz = g->L.zt ? g->W2 : g->W;
    h->timed(4, [&] { launch_gw_zpass_acc<double>(g->stream, g->N, g->P, wz, g->A, g->L, g->plan, g->tw, s); });
}

// The DKD loop; before_kick(s, t) runs before kick s (0-based) at its time t, when phi has
// reached t and pi is still the previous kick's.
template <typename T, bool COMP, typename Hook>
static void run_steps_hooked(axg_handle* h, double t1, double dt, int32_t nt, Hook&& before_kick)
{
    const int N = h->N;
    const axg_params& p = h->p;
    double t = t1;
    for (int32_t it = 1; it <= nt; it++) {
        if (it == 1) {
            // first half drift, in place (each site reads and writes only itself)
            h->timed(0, [&] {
                launch_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[h->cur], h->pi, h->cphi[h->cur],
                                      h->cphi[h->cur], h->cpi, dt / 2);
            });
            t = t + dt / 2;
        }
        before_kick(it - 1, t);
        KickScalars k = kick_scalars(p, t, dt);
        double dtd = (it == nt) ? dt / 2 : dt;
        int nxt = 1 - h->cur;
        bool fused = false;
        const bool lattice_gw = p.grav_spec && !h->gs;  // GW sourced on the lattice's own grid
        if (lattice_gw && h->fuse_x) {
            h->timed(1, [&] {
                fused = launch_kick_source_xfft<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi,
                                                         h->cphi[h->cur], h->cphi[nxt], h->cpi, h->W, h->P, h->tw, k,
                                                         p.dspa, dtd, h->kick_variant);
            });
        }
        if (!fused) {
            h->timed(1, [&] {
                launch_kick_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, h->cphi[h->cur],
                                           h->cphi[nxt], h->cpi, h->src, k, p.dspa, dtd, lattice_gw);
            });
        }
        if (p.grav_spec && h->gs) reduced_gw_step<T, COMP>(h, t, dt);
        if (lattice_gw) {
            if (!fused) h->timed(2, [&] { launch_gw_xpass<T>(h->stream, N, h->P, h->src, h->W, h->plan, h->tw); });
            h->timed(3, [&] {
                launch_gw_ypass<T>(h->stream, N, h->P, h->W, h->W2, h->L, h->plan, h->tw, h->ypass_variant);
            });
            GwStepScalars g;
            g.tau = t;
            g.dw = t * dt * p.dv_spa * (p.sigma_cr * p.sigma_cr);  // grav_waves.f90:209
            g.dwav = p.dwav;
            g.fix_swap = p.gw_fix_swap;
            const void* wz = h->L.zt ? h->W2 : h->W;
            h->timed(4, [&] { launch_gw_zpass_acc<T>(h->stream, N, h->P, wz, h->A, h->L, h->plan, h->tw, g); });
        }
        h->cur = nxt;
        t = t + dtd;
    }
}

template <typename T, bool COMP>
static void run_steps(axg_handle* h, double t1, double dt, int32_t nt)
{
    run_steps_hooked<T, COMP>(h, t1, dt, nt, [](int32_t, double) {});
}

// One DKD step whose scalars come from the handle's step table, entry *base + sidx: the same
// kernels as run_steps, so the fields match it bit for bit.  Launches only, so it can be
// captured into a graph; flips h->cur like a step does.
template <typename T, bool COMP>
static void table_step(axg_handle* h, int sidx, const int* base)
{
    const int N = h->N;
    const axg_params& p = h->p;
    const StepRef sr{h->steptab, base, sidx};
    const KickScalars k{};  // replaced by the table's entry in the kernels
    const int nxt = 1 - h->cur;
    bool fused = false;
    if (p.grav_spec && h->fuse_x)
        fused = launch_kick_source_xfft<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, h->cphi[h->cur],
                                                 h->cphi[nxt], h->cpi, h->W, h->P, h->tw, k, p.dspa, 0.0,
                                                 h->kick_variant, sr);
    if (!fused)
        launch_kick_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, h->cphi[h->cur], h->cphi[nxt],
                                   h->cpi, h->src, k, p.dspa, 0.0, p.grav_spec != 0, sr);
    if (p.grav_spec) {
        if (!fused) launch_gw_xpass<T>(h->stream, N, h->P, h->src, h->W, h->plan, h->tw);
        launch_gw_ypass<T>(h->stream, N, h->P, h->W, h->W2, h->L, h->plan, h->tw, h->ypass_variant);
        GwStepScalars g;
        g.tau = 0;
        g.dw = 0;
        g.dwav = p.dwav;
        g.fix_swap = p.gw_fix_swap;
        const void* wz = h->L.zt ? h->W2 : h->W;
        launch_gw_zpass_acc<T>(h->stream, N, h->P, wz, h->A, h->L, h->plan, h->tw, g, sr);
    }
    h->cur = nxt;
}

static void table_step_any(axg_handle* h, int sidx, const int* base)
{
    if (h->comp) table_step<float, true>(h, sidx, base);
    else if (h->fp32) table_step<float, false>(h, sidx, base);
    else table_step<double, false>(h, sidx, base);
}

// One DKD step of a stack of lattices that share N, precision, dspa, dwav and the index order of
// the GW update, in the pass configurations of h (the first of them): each pass is one launch
// over all of them, the lattice in blockIdx.z, each reading its buffers and its step entry *base + sidx from
// its LatRef.  Launches on h->stream; the handles' cur is left alone (the caller flips it).
template <typename T, bool COMP>
static void stacked_step(axg_handle* h, const LatRef* lat, int nlat, int sidx, const int* base)
{
    const axg_params& p = h->p;
    const StepRef sr{nullptr, base, sidx};
    const KickScalars k{};
    if (!launch_kick_source_xfft<T, COMP>(h->stream, h->N, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                          nullptr, h->P, h->tw, k, p.dspa, 0.0, h->kick_variant, sr, lat, nlat))
        throw Error(AXG_EINVAL, "stacked step: no fused pass for this size");
    launch_gw_ypass<T>(h->stream, h->N, h->P, nullptr, nullptr, h->L, h->plan, h->tw, h->ypass_variant, lat, nlat);
    GwStepScalars g;
    g.tau = 0;
    g.dw = 0;
    g.dwav = p.dwav;
    g.fix_swap = p.gw_fix_swap;
    launch_gw_zpass_acc<T>(h->stream, h->N, h->P, nullptr, nullptr, h->L, h->plan, h->tw, g, sr, lat, nlat);
}

static void stacked_step_any(axg_handle* h, const LatRef* lat, int nlat, int sidx, const int* base)
{
    if (h->comp) stacked_step<float, true>(h, lat, nlat, sidx, base);
    else if (h->fp32) stacked_step<float, false>(h, lat, nlat, sidx, base);
    else stacked_step<double, false>(h, lat, nlat, sidx, base);
}

static void advance(axg_handle* h, double t1, double dt, int32_t nt)
{
    if (h->comp) run_steps<float, true>(h, t1, dt, nt);
    else if (h->fp32) run_steps<float, false>(h, t1, dt, nt);
    else run_steps<double, false>(h, t1, dt, nt);
    h->have_flags = false;
    h->have_plaq = false;
    h->have_spec = false;
    h->nstr_rec = -1;
    h->nwall_rec = -1;
    h->have_grp[0] = h->have_grp[1] = false;
    AXG_HIP(hipStreamSynchronize(h->stream));
    if (h->prof) h->collect();
}

extern "C" int axg_profile(axg_handle* h, int32_t enable, double* ms, int64_t* launches, int32_t nslots)
{
    return guarded(h, [&] {
        require(nslots >= 0 && nslots <= AXG_NSLOT, "nslots must be 0..5");
        if (enable && !h->prof) {
            for (int i = 0; i < AXG_NSLOT; i++) { h->prof_ms[i] = 0; h->prof_n[i] = 0; }
            h->ev_used = 0;
        }
        h->prof = enable != 0;
        for (int i = 0; i < nslots; i++) {
            if (ms) ms[i] = h->prof_ms[i];
            if (launches) launches[i] = h->prof_n[i];
        }
    });
}

// Configurations of the fused kick / source / x-transform pass the tuner tries: the built-in
// choice, wide blocks (twice the rows), the other sites-per-thread choice, and both; then the
// same four with XCD strips.
static const int kTuneVariants[] = {0, 1, 4, 5, 8, 9, 12, 13};

// A start/stop event pair, destroyed with the scope.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair()
    {
        AXG_HIP(hipEventCreate(&a));
        AXG_HIP(hipEventCreate(&b));
    }
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
};

extern "C" int axg_tune(axg_handle* h, int32_t reps, int32_t* variant, double* ms, int32_t nms)
{
    return guarded(h, [&] {
        require(reps >= 1 && reps <= 1000, "reps must be 1..1000");
        const int N = h->N;
        const int nv = (int)(sizeof kTuneVariants / sizeof kTuneVariants[0]);
        for (int i = 0; ms && i < nms; i++) ms[i] = 0;
        const bool tunable = h->p.grav_spec && h->fuse_x && p2_size(N) && N >= 64;
        if (!tunable) {  // one configuration only
            if (variant) *variant = h->kick_variant | (h->ypass_variant << 8);
            return;
        }
        EventPair ev;
        // mean ms per launch of once() over reps launches after one warm-up launch
        auto time_avg = [&](auto&& once) {
            once();
            AXG_HIP(hipEventRecord(ev.a, h->stream));
            for (int r = 0; r < reps; r++) once();
            AXG_HIP(hipEventRecord(ev.b, h->stream));
            AXG_HIP(hipEventSynchronize(ev.b));
            float t = 0;
            AXG_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
            return (double)t / reps;
        };
        // fused pass: pi is kicked in place by every trial launch, so it waits in the work array
        const size_t pib = (size_t)h->n3 * h->t2;
        AXG_HIP(hipMemcpyAsync(h->work, h->pi, pib, hipMemcpyDeviceToDevice, h->stream));
        if (h->comp)
            AXG_HIP(hipMemcpyAsync((char*)h->work + pib, h->cpi, (size_t)h->n3 * 4, hipMemcpyDeviceToDevice, h->stream));
        const KickScalars k = kick_scalars(h->p, 1.0, 1e-3);
        const int nxt = 1 - h->cur;
        double best = 1e300;
        int bestv = h->kick_variant;
        for (int iv = 0; iv < nv; iv++) {
            const int v = kTuneVariants[iv];
            const double avg = time_avg([&] {
                if (h->comp)
                    launch_kick_source_xfft<float, true>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi,
                                                         h->cphi[h->cur], h->cphi[nxt], h->cpi, h->W, h->P, h->tw, k,
                                                         h->p.dspa, 1e-3, v);
                else if (h->fp32)
                    launch_kick_source_xfft<float, false>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, nullptr,
                                                          nullptr, nullptr, h->W, h->P, h->tw, k, h->p.dspa, 1e-3, v);
                else
                    launch_kick_source_xfft<double, false>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, nullptr,
                                                           nullptr, nullptr, h->W, h->P, h->tw, k, h->p.dspa, 1e-3, v);
            });
            if (ms && iv < nms) ms[iv] = avg;
            if (avg < best) {
                best = avg;
                bestv = v;
            }
        }
        AXG_HIP(hipMemcpyAsync(h->pi, h->work, pib, hipMemcpyDeviceToDevice, h->stream));
        if (h->comp)
            AXG_HIP(hipMemcpyAsync(h->cpi, (char*)h->work + pib, (size_t)h->n3 * 4, hipMemcpyDeviceToDevice, h->stream));
        // y pass: reads W and writes W2, both rewritten by every step
        double ybest = 1e300;
        int ybestv = h->ypass_variant;
        if (h->L.zt) {
            for (int yv = 0; yv < 3; yv++) {
                const double avg = time_avg([&] {
                    if (h->fp32) launch_gw_ypass<float>(h->stream, N, h->P, h->W, h->W2, h->L, h->plan, h->tw, yv);
                    else launch_gw_ypass<double>(h->stream, N, h->P, h->W, h->W2, h->L, h->plan, h->tw, yv);
                });
                if (ms && nv + yv < nms) ms[nv + yv] = avg;
                if (avg < ybest) {
                    ybest = avg;
                    ybestv = yv;
                }
            }
        }
        AXG_HIP(hipStreamSynchronize(h->stream));
        h->kick_variant = bestv;
        h->ypass_variant = ybestv;
        if (variant) *variant = bestv | (ybestv << 8);
    });
}

extern "C" int axg_advance_steps(axg_handle* h, double t1, double dt, int32_t nsteps)
{
    return guarded(h, [&] {
        // dt < 0 runs the schedule backwards in time (the time-reversal check of the integrator)
        require(nsteps >= 1 && dt != 0 && t1 > 0 && t1 + nsteps * dt > 0,
                "need nsteps >= 1, dt != 0, t1 > 0 and t1 + nsteps*dt > 0");
        advance(h, t1, dt, nsteps);
    });
}

extern "C" int axg_advance(axg_handle* h, double t1, double t2, int32_t* nsteps_out)
{
    return guarded(h, [&] {
        require(t2 > t1 && t1 > 0, "need t2 > t1 > 0");
        require(h->p.dtau_fixed > 0 || h->p.eps_dtau > 0, "eps_dtau or dtau_fixed must be positive");
        double dt;
        int32_t n;
        get_dtau(h->p, t1, t2, &dt, &n);
        advance(h, t1, dt, n);
        if (nsteps_out) *nsteps_out = n;
    });
}

// ------------------------------------------------------------------ GW spectrum

extern "C" int axg_gw_spectrum(axg_handle* h, double* qmid, double* power, double* numer, double* denom, int32_t nbins)
{
    return guarded(h, [&] {
        if (h->gs) return forward(h, axg_gw_spectrum(h->gs, qmid, power, numer, denom, nbins));
        const axg_params& p = h->p;
        const int nb = p.num_bin_grav;
        require(h->A != nullptr, "grav_spec is off: no GW state");
        require(nbins == nb, "nbins must equal num_bin_grav");
        const int nwg = 1024;
        require((size_t)(nwg + 1) * 2 * nb * sizeof(double) <= h->red_bytes, "reduction scratch too small");
        double* out = h->red + (size_t)nwg * 2 * nb;
        if (h->fp32) launch_gw_bin<float>(h->stream, h->N, h->L, h->A, h->gw_bin, h->gw_n2max, nb, h->red, nwg, out);
        else launch_gw_bin<double>(h->stream, h->N, h->L, h->A, h->gw_bin, h->gw_n2max, nb, h->red, nwg, out);
        std::vector<double> r(2 * nb);
        AXG_HIP(hipMemcpyAsync(r.data(), out, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        // grav_waves.f90:62-71, 265-267, 297-305
        double qmax = AXG_PI_T / p.dspa, dq0 = qmax / nb;
        std::vector<double> wave(nb + 1);
        wave[0] = 0;
        for (int b = 1; b <= nb; b++) wave[b] = wave[b - 1] + dq0;
        const double bigh_in_mstar = 2.1332e-42 / 2.435e18;
        double t6 = p.tau_cr * p.tau_cr * p.tau_cr;
        t6 = t6 * t6;
        double a4 = p.a_now * p.a_now;
        a4 = a4 * a4;
        double fact = 1 / t6 / (bigh_in_mstar * bigh_in_mstar) / a4 / 3 / (AXG_TWOPI_T * AXG_TWOPI_T) * p.dv_wav;
        for (int b = 0; b < nb; b++) {
            double qm = (wave[b + 1] + wave[b]) / 2;
            double nu = r[b], de = r[nb + b];
            if (qmid) qmid[b] = qm;
            if (numer) numer[b] = nu;
            if (denom) denom[b] = de;
            if (power) power[b] = de > 0 ? nu / de * (qm * qm * qm) * fact : 0.0;
        }
    });
}

// ------------------------------------------------------------------ energy

extern "C" int axg_energy(axg_handle* h, double tau, double out[4])
{
    return guarded(h, [&] {
        require(out && tau > 0, "need out and tau > 0");
        const axg_params& p = h->p;
        double thrm = p.initial_loop ? 0.0 : AXG_THRM_SIGMA / tau;
        double m2 = msquared(thrm), ma2 = maxion2(thrm, p.kappa), kt = tilt_coef(p, tau);
        const int nwg = 1024;
        double* dout = h->red + (size_t)nwg * 4;
        if (h->comp)
            launch_energy<float, true>(h->stream, h->N, h->phi[h->cur], h->cphi[h->cur], h->pi, h->cpi, m2, ma2, kt,
                                       h->red, nwg, dout);
        else if (h->fp32)
            launch_energy<float, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->pi, nullptr, m2, ma2, kt, h->red,
                                        nwg, dout);
        else
            launch_energy<double, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->pi, nullptr, m2, ma2, kt, h->red,
                                         nwg, dout);
        double r[3];
        AXG_HIP(hipMemcpyAsync(r, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        double dv = p.dspa * p.dspa * p.dspa;
        out[0] = r[0] * dv;
        out[1] = r[1] / (p.dspa * p.dspa) * dv;
        out[2] = tau * tau * p.sigma_cr * p.sigma_cr * r[2] * dv;
        out[3] = out[0] + out[1] + out[2];
    });
}

// ------------------------------------------------------------------ defects

extern "C" int axg_defects(axg_handle* h, axg_defect_out* out)
{
    return guarded(h, [&] {
        require(out != nullptr, "null output");
        const axg_params& p = h->p;
        const int64_t nblk = (h->n3 + 255) / 256;
        int64_t* counts = h->blk;             // [nblk][3]
        int64_t* offs = h->blk + 3 * nblk;    // [nblk] exclusive mask offsets
        int64_t* tot = offs + nblk;           // [4]
        if (h->comp)
            launch_classify<float, true>(h->stream, h->N, h->phi[h->cur], h->cphi[h->cur], p.id_strings, p.id_walls,
                                         p.spec_est, h->flags, counts, nblk);
        else if (h->fp32)
            launch_classify<float, false>(h->stream, h->N, h->phi[h->cur], nullptr, p.id_strings, p.id_walls,
                                          p.spec_est, h->flags, counts, nblk);
        else
            launch_classify<double, false>(h->stream, h->N, h->phi[h->cur], nullptr, p.id_strings, p.id_walls,
                                           p.spec_est, h->flags, counts, nblk);
        launch_scan_counts(h->stream, counts, offs, nblk, tot, 2, 3, h->blk + 6 * nblk + 16);
        int64_t t[4];
        AXG_HIP(hipMemcpyAsync(t, tot, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        out->n_string = t[0];
        out->n_wall = t[1];
        out->n_mask = t[2];
        h->nmask = t[2];
        h->nstring = t[0];
        h->nwall = t[1];
        h->have_flags = true;
        h->nstr_rec = -1;
        h->nwall_rec = -1;
        h->have_grp[0] = h->have_grp[1] = false;
    });
}

extern "C" int axg_plaquettes(axg_handle* h, axg_plaquette_out* out)
{
    return guarded(h, [&] {
        require(out != nullptr, "null output");
        const int64_t nblk = (h->n3 + PLAQ_BLOCK - 1) / PLAQ_BLOCK;
        if (!h->plaq_code) {
            h->plaq_code = (uint16_t*)h->alloc((size_t)h->n3 * sizeof(uint16_t));
            h->plaq_cnt = (int64_t*)h->alloc((size_t)(17 * nblk + 16 + scan_scratch(nblk, PLAQ_COLS)) * sizeof(int64_t));
        }
        int64_t* counts = h->plaq_cnt;
        int64_t* offs = counts + 16 * nblk;
        int64_t* tot = offs + nblk;
        if (h->comp)
            launch_plaquettes<float, true>(h->stream, h->N, h->phi[h->cur], h->cphi[h->cur], h->plaq_code, counts, offs,
                                           tot, tot + 16);
        else if (h->fp32)
            launch_plaquettes<float, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->plaq_code, counts, offs, tot,
                                            tot + 16);
        else
            launch_plaquettes<double, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->plaq_code, counts, offs, tot,
                                            tot + 16);
        int64_t t[16];
        AXG_HIP(hipMemcpyAsync(t, tot, sizeof(t), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        for (int o = 0; o < 3; o++) {
            out->n_pierced[o] = t[5 * o + 0];
            out->n_plus[o] = t[5 * o + 1];
            out->n_minus[o] = t[5 * o + 2];
            out->n_quadrant[o] = t[5 * o + 3];
            out->net[o] = t[5 * o + 4];
        }
        h->nplaq = t[15];
        h->have_plaq = true;
    });
}

extern "C" int axg_get_plaquettes(axg_handle* h, int32_t* rec, int64_t cap, int64_t* count)
{
    return guarded(h, [&] {
        require(h->have_plaq, "call axg_plaquettes first");
        require(count != nullptr, "null count");
        *count = h->nplaq;
        if (h->nplaq == 0 || rec == nullptr || cap == 0) return;  // size query
        require(cap >= h->nplaq, "plaquette buffer too small");
        if (h->nplaq > h->plaq_rec_cap) {
            void* a = h->plaq_rec;
            AXG_HIP(hipStreamSynchronize(h->stream));
            h->release_one(a);
            h->plaq_rec_cap = 0;
            h->plaq_rec = (int32_t*)h->alloc((size_t)5 * h->nplaq * sizeof(int32_t));
            h->plaq_rec_cap = h->nplaq;
        }
        const int64_t nblk = (h->n3 + PLAQ_BLOCK - 1) / PLAQ_BLOCK;
        launch_plaquette_compact(h->stream, h->N, h->plaq_code, h->plaq_cnt + 16 * nblk, h->plaq_rec);
        AXG_HIP(hipMemcpyAsync(rec, h->plaq_rec, (size_t)5 * h->nplaq * sizeof(int32_t), hipMemcpyDeviceToHost,
                               h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

extern "C" int axg_get_mask(axg_handle* h, int32_t* ixyz, int64_t cap, int64_t* count)
{
    return guarded(h, [&] {
        require(h->have_flags, "call axg_defects first");
        require(count != nullptr, "null count");
        *count = h->nmask;
        if (h->nmask == 0 || ixyz == nullptr || cap == 0) return;  // size query
        require(cap >= h->nmask, "mask buffer too small");
        const int64_t nblk = (h->n3 + 255) / 256;
        int64_t* offs = h->blk + 3 * nblk;
        int32_t* dm = (int32_t*)h->work;  // 3*nmask ints fit in the N^3 x 16-byte work array
        launch_cell_compact(h->stream, h->N, h->flags, 4, offs, dm, 1, nblk);
        AXG_HIP(hipMemcpyAsync(ixyz, dm, (size_t)3 * h->nmask * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

extern "C" int axg_strings(axg_handle* h, uint32_t seed0, uint32_t seed1, axg_string_out* out)
{
    return guarded(h, [&] {
        require(out != nullptr, "null output");
        require(h->have_flags, "call axg_defects first");
        require(h->p.id_strings != 0, "id_strings is off: no string cells flagged");
        const int64_t nblk = (h->n3 + 255) / 256;
        const int64_t ns = h->nstring;
        const int64_t nsb = (ns + STRING_NT - 1) / STRING_NT;
        if (ns > h->str_cap) {
            void* a = h->str_cells;
            void* b = h->str_rec;
            void* c = h->str_part;
            void* d = h->str_pend;
            void* e = h->str_state;
            AXG_HIP(hipStreamSynchronize(h->stream));
            h->release_one(a);
            h->release_one(b);
            h->release_one(c);
            h->release_one(d);
            h->release_one(e);
            h->str_cap = 0;
            h->str_cells = (int32_t*)h->alloc((size_t)ns * sizeof(int32_t));
            h->str_rec = (double*)h->alloc((size_t)11 * ns * sizeof(double));
            h->str_part = (double*)h->alloc((size_t)(5 * nsb + 8) * sizeof(double));
            h->str_pend = (int64_t*)h->alloc((size_t)ns * sizeof(int64_t));
            h->str_state = (double*)h->alloc((size_t)(STRING_STATE * ns + 1) * sizeof(double));
            h->str_cap = ns;
        }
        double sums[5] = {0, 0, 0, 0, 0};
        if (ns > 0) {
            int64_t* counts = h->blk;               // [nblk][3] from the classification
            int64_t* soffs = h->blk + 4 * nblk + 8;  // [nblk] exclusive string offsets
            int64_t* stot = soffs + nblk;            // [4]
            launch_scan_counts(h->stream, counts, soffs, nblk, stot, 0, 3, h->blk + 6 * nblk + 16);
            launch_cell_compact(h->stream, h->N, h->flags, 1, soffs, h->str_cells, 0, nblk);
            double* dsum = h->str_part + 5 * nsb;
            const Draws dr = h->draws(seed0, seed1);
            int* npend = (int*)h->str_state;
            double* state = h->str_state + 1;
            if (h->comp)
                launch_string_fit<float, true>(h->stream, h->N, h->phi[h->cur], h->cphi[h->cur], h->pi, h->cpi,
                                               h->str_cells, ns, dr, h->p.dspa, h->string_split, h->p.string_vel, h->str_rec,
                                               h->str_part, npend, h->str_pend, state);
            else if (h->fp32)
                launch_string_fit<float, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->pi, nullptr, h->str_cells,
                                                ns, dr, h->p.dspa, h->string_split, h->p.string_vel, h->str_rec, h->str_part,
                                                npend, h->str_pend, state);
            else
                launch_string_fit<double, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->pi, nullptr,
                                                 h->str_cells, ns, dr, h->p.dspa, h->string_split, h->p.string_vel,
                                                 h->str_rec, h->str_part, npend, h->str_pend, state);
            sum_columns(h->stream, h->str_part, (int)nsb, 5, dsum);
            AXG_HIP(hipMemcpyAsync(sums, dsum, sizeof sums, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
        if (h->stream_on) h->stream_pos += 4004 * (uint64_t)ns;  // 4 draws x 1001 starts per string cell
        out->n_string = ns;
        out->n_vel = (int64_t)sums[1];
        out->length = sums[0];
        for (int k = 0; k < 3; k++) out->vel[k] = out->n_vel > 0 ? sums[2 + k] / (double)out->n_vel : 0.0;
        h->nstr_rec = ns;
        h->have_grp[0] = false;
    });
}

extern "C" int axg_walls(axg_handle* h, axg_wall_out* out)
{
    return guarded(h, [&] {
        require(out != nullptr, "null output");
        require(h->have_flags, "call axg_defects first");
        require(h->p.id_walls != 0, "id_walls is off: no wall cells flagged");
        const int64_t nblk = (h->n3 + 255) / 256;
        const int64_t nw = h->nwall;
        const int64_t nwb = (nw + WALL_NT - 1) / WALL_NT;
        if (nw > h->wall_cap) {
            void* a = h->wall_cells;
            void* b = h->wall_rec;
            void* c = h->wall_part;
            AXG_HIP(hipStreamSynchronize(h->stream));
            h->release_one(a);
            h->release_one(b);
            h->release_one(c);
            h->wall_cap = 0;
            h->wall_cells = (int32_t*)h->alloc((size_t)nw * sizeof(int32_t));
            h->wall_rec = (double*)h->alloc((size_t)8 * nw * sizeof(double));
            h->wall_part = (double*)h->alloc((size_t)(2 * nwb + 8) * sizeof(double));
            h->wall_cap = nw;
        }
        double sums[2] = {0, 0};
        if (nw > 0) {
            int64_t* counts = h->blk;
            int64_t* woffs = h->blk + 4 * nblk + 8;  // shares the string offsets' room
            int64_t* wtot = woffs + nblk;
            launch_scan_counts(h->stream, counts, woffs, nblk, wtot, 1, 3, h->blk + 6 * nblk + 16);
            launch_cell_compact(h->stream, h->N, h->flags, 2, woffs, h->wall_cells, 0, nblk);
            double* dsum = h->wall_part + 2 * nwb;
            if (h->comp)
                launch_wall_fit<float, true>(h->stream, h->N, h->phi[h->cur], h->cphi[h->cur], h->wall_cells, nw,
                                             h->wall_rec, h->wall_part);
            else if (h->fp32)
                launch_wall_fit<float, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->wall_cells, nw, h->wall_rec,
                                              h->wall_part);
            else
                launch_wall_fit<double, false>(h->stream, h->N, h->phi[h->cur], nullptr, h->wall_cells, nw,
                                               h->wall_rec, h->wall_part);
            sum_columns(h->stream, h->wall_part, (int)nwb, 2, dsum);
            AXG_HIP(hipMemcpyAsync(sums, dsum, sizeof sums, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
        out->n_wall = nw;
        out->area = sums[0];
        h->nwall_rec = nw;
        h->have_grp[1] = false;
        h->wall_bad = sums[1] > 0;
        if (sums[1] > 0)
            throw Error(AXG_ENUMERIC, "position_wall: " + std::to_string((int64_t)sums[1]) +
                                          " wall cells have fewer than 3 edges where Im(phi) changes sign");
    });
}

// grouping_strings / grouping_walls (defects_3d.f90:1235-1260, 1735-1760) on the records of the
// last fit, and the grouped sums of add_string_length / add_wall_area (1654-1733, 2145-2204).
// Groups are numbered 1, 2, ... in the order of their first record (site order); the sums run
// over each group's records in site order, then over the groups in that order.
static void group_records(axg_handle* h, int kind, double width, double hor_size, axg_group_out* out)
{
    const bool st = kind == 0;
    const int64_t n = st ? h->nstr_rec : h->nwall_rec;
    require(n >= 0, st ? "call axg_strings first (after the last change of the fields)"
                       : "call axg_walls first (after the last change of the fields)");
    require(out != nullptr, "null output");
    require(width >= 0 && hor_size > 0 && std::isfinite(width) && std::isfinite(hor_size),
            "need width >= 0 and hor_size > 0");
    // threshold: the weighted geometric mean of width and separation, at least 2 (1245-1246)
    double th = pow(pow(width, 2.0) * hor_size, 1 / (2.0 + 1));
    th = std::max(2.0, th);
    require(st || !h->wall_bad, "the last axg_walls failed (cells with fewer than 3 edge points)");
    const double* rec = st ? h->str_rec : h->wall_rec;
    const int32_t* cells = st ? h->str_cells : h->wall_cells;
    const int stride = st ? 11 : 8, col = 7;  // sl and wa are both column 7
    std::vector<int32_t> lab(n);
    std::vector<double> val(n);
    if (n > 0) {
        if (n > h->grp_cap) {
            void* q = h->grp;
            AXG_HIP(hipStreamSynchronize(h->stream));
            h->release_one(q);
            h->grp = nullptr;
            h->grp_cap = 0;
            h->grp = h->alloc((size_t)n * 16 + 256);  // val, parent, label; scalars
            h->grp_cap = n;
        }
        double* d_val = (double*)h->grp;
        int32_t* d_parent = (int32_t*)(d_val + n);
        int32_t* d_label = d_parent + n;
        unsigned long long* d_scal = (unsigned long long*)((char*)h->grp + (((size_t)n * 16 + 63) & ~(size_t)63));
        int* d_fail = (int*)(d_scal + 2);
        int32_t* map = (int32_t*)h->work;  // one int per cell
        int n_out = 0, centre = 0;
        const double maxoff = group_prepare(h->stream, h->N, rec, stride, col, cells, n, map, d_parent, d_val, d_scal,
                                            d_label, &n_out, &centre);
        require(n_out <= 65536, "more than 65536 defect positions lie over a cell from their cell");
        // route: the cell box (th_dist + 2 offsets, each at most one cell, up to 16 cells), one
        // group when no two mapped points can be farther apart than th_dist, else every pair
        int mode = 0, R = 0;
        if (th >= sqrt(3.0) * h->N / 2) {
            mode = n_out < n ? 1 : 2;
        } else if (th + 2 * maxoff <= 16) {
            R = std::min((int)ceil(th + 2 * maxoff), h->N);
        } else {
            require(n <= 262144, "grouping threshold " + std::to_string(th) +
                                     " cells needs the all-pairs test, limited to 262144 points");
            mode = 2;
        }
        const int zero = 0;
        AXG_HIP(hipMemcpyAsync(d_fail, &zero, sizeof zero, hipMemcpyHostToDevice, h->stream));
        group_link(h->stream, h->N, rec, stride, cells, n, map, d_parent, d_label, th, maxoff, R, mode, centre,
                   d_label, n_out, d_fail);
        int fail = 0;
        AXG_HIP(hipMemcpyAsync(&fail, d_fail, sizeof fail, hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipMemcpyAsync(lab.data(), d_label, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipMemcpyAsync(val.data(), d_val, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        require(fail == 0, "grouping did not converge");
    }
    std::vector<int32_t>& ids = h->grp_ids[kind];
    ids.assign(n, 0);
    std::vector<double> gsum;
    for (int64_t i = 0; i < n; i++) {
        const int32_t r = lab[i];
        require(r >= 0 && r <= i, "grouping produced an invalid label");
        if (r == i) {
            gsum.push_back(0.0);
            ids[i] = (int32_t)gsum.size();
        } else {
            ids[i] = ids[r];
        }
        gsum[ids[i] - 1] += val[i];
    }
    // thresholds of add_string_length (2 hor, 2 pi hor) and add_wall_area (pi hor^2, 4 pi hor^2)
    const double t2 = st ? 2 * hor_size : AXG_PI_T * (hor_size * hor_size);
    const double t3 = st ? AXG_TWOPI_T * hor_size : 4 * AXG_PI_T * (hor_size * hor_size);
    double s[3] = {0, 0, 0};
    for (double g : gsum) {
        s[0] = s[0] + g;
        if (g > t2) s[1] = s[1] + g;
        if (g > t3) s[2] = s[2] + g;
    }
    out->n_group = (int64_t)gsum.size();
    out->th_dist = th;
    for (int k = 0; k < 3; k++) out->sum[k] = s[k];
    h->have_grp[kind] = true;
}

extern "C" int axg_group_strings(axg_handle* h, double width, double hor_size, axg_group_out* out)
{
    return guarded(h, [&] { group_records(h, 0, width, hor_size, out); });
}

extern "C" int axg_group_walls(axg_handle* h, double thickness, double hor_size, axg_group_out* out)
{
    return guarded(h, [&] { group_records(h, 1, thickness, hor_size, out); });
}

extern "C" int axg_get_groups(axg_handle* h, int32_t kind, int32_t* ids, int64_t cap, int64_t* count)
{
    return guarded(h, [&] {
        require(count != nullptr, "null count");
        require(kind == 0 || kind == 1, "kind must be 0 (strings) or 1 (walls)");
        require(h->have_grp[kind], kind == 0 ? "call axg_group_strings first" : "call axg_group_walls first");
        const std::vector<int32_t>& g = h->grp_ids[kind];
        *count = (int64_t)g.size();
        if (g.empty() || ids == nullptr || cap == 0) return;  // size query
        require(cap >= (int64_t)g.size(), "id buffer too small");
        memcpy(ids, g.data(), g.size() * sizeof(int32_t));
    });
}

extern "C" int axg_get_walls(axg_handle* h, double* rec, int64_t cap, int64_t* count)
{
    return guarded(h, [&] {
        require(count != nullptr, "null count");
        require(h->nwall_rec >= 0, "call axg_walls first (after the last change of the fields)");
        *count = h->nwall_rec;
        if (h->nwall_rec == 0 || rec == nullptr || cap == 0) return;  // size query
        require(cap >= h->nwall_rec, "wall buffer too small");
        AXG_HIP(hipMemcpyAsync(rec, h->wall_rec, (size_t)8 * h->nwall_rec * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

extern "C" int axg_get_strings(axg_handle* h, double* rec, int64_t cap, int64_t* count)
{
    return guarded(h, [&] {
        require(count != nullptr, "null count");
        require(h->nstr_rec >= 0, "call axg_strings first (after the last change of the fields)");
        *count = h->nstr_rec;
        if (h->nstr_rec == 0 || rec == nullptr || cap == 0) return;  // size query
        require(cap >= h->nstr_rec, "string buffer too small");
        AXG_HIP(hipMemcpyAsync(rec, h->str_rec, (size_t)11 * h->nstr_rec * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

extern "C" int axg_get_flags(axg_handle* h, int8_t* flags, int64_t nsite)
{
    return guarded(h, [&] {
        require(h->have_flags, "call axg_defects first");
        require(flags && nsite == h->n3, "nsite must be ngrid^3");
        AXG_HIP(hipMemcpyAsync(flags, h->flags, h->n3, hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
    });
}

// ------------------------------------------------------------------ field spectrum

// In-place unnormalised 3D transform of an N^3 FP64 array: sign +1 is FFTW_BACKWARD (the
// sign of every diagnostic transform), -1 FFTW_FORWARD (the thermal initial state).
static void fft3d_work(axg_handle* h, double2* d, int sign = +1)
{
    const int N = h->N;
    if (sign > 0) {
        launch_fft_rows<double, +1>(h->stream, N, d, (int64_t)N * N, h->plan, h->tw64);
        launch_fft_cols<double, +1>(h->stream, N, d, N, (int64_t)N * N, N, N, h->plan, h->tw64);
        launch_fft_cols<double, +1>(h->stream, N, d, (int64_t)N * N, N, N, N, h->plan, h->tw64);
    } else {
        launch_fft_rows<double, -1>(h->stream, N, d, (int64_t)N * N, h->plan, h->tw64);
        launch_fft_cols<double, -1>(h->stream, N, d, N, (int64_t)N * N, N, N, h->plan, h->tw64);
        launch_fft_cols<double, -1>(h->stream, N, d, (int64_t)N * N, N, N, N, h->plan, h->tw64);
    }
}

// Runs body(lo, hi) over [0, n) split into contiguous ranges on up to 16 host threads (the
// one-time spectrum tables: each index writes only its own outputs, so the split changes nothing).
template <typename F> static void host_parallel(int64_t n, F&& body)
{
    const unsigned hc = std::thread::hardware_concurrency();
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({(int64_t)(hc ? hc : 1), (int64_t)16, n}));
    std::vector<std::thread> pool;
    for (int64_t k = 1; k < nt; k++) pool.emplace_back([&, k] { body(n * k / nt, n * (k + 1) / nt); });
    body(0, n / nt);
    for (auto& t : pool) t.join();
}

static void ensure_spec_tables(axg_handle* h)
{
    if (h->d_largeq) return;
    const int nb = h->p.num_bin, nlq = 500;
    const int ncov = nb * (nb + 1) / 2;
    // spectrum.f90:74-98
    std::vector<double> largeq((size_t)ncov * (nlq + 1)), mat((size_t)ncov * (nlq + 1)), dqs(ncov);
    const std::vector<double>& wav = h->wav;
    host_parallel(ncov, [&](int64_t lo, int64_t hi) {
        for (int ij = (int)lo + 1; ij <= (int)hi; ij++) {
            int i = (int)lround(sqrt(2.0 * ij));
            int j = ij - i * (i - 1) / 2;
            double lj = wav[j], sj = wav[j - 1], li = wav[i], si = wav[i - 1];
            double maxq = li + lj;
            double minq = std::max(si, sj) - std::min(li, lj);
            if (minq < 0) minq = 0;
            double d = maxq - minq;
            d = d / nlq;
            dqs[ij - 1] = d;
            for (int iq = 0; iq <= nlq; iq++) {
                double q = minq + d * iq;
                largeq[(size_t)(ij - 1) * (nlq + 1) + iq] = q;
                mat[(size_t)(ij - 1) * (nlq + 1) + iq] = integrate_p(200, li, si, lj, sj, q);
            }
        }
    });
    h->d_largeq = (double*)h->alloc(largeq.size() * 8);
    h->d_matspec = (double*)h->alloc(mat.size() * 8);
    h->d_dqs = (double*)h->alloc(dqs.size() * 8);
    AXG_HIP(hipMemcpy(h->d_largeq, largeq.data(), largeq.size() * 8, hipMemcpyHostToDevice));
    AXG_HIP(hipMemcpy(h->d_matspec, mat.data(), mat.size() * 8, hipMemcpyHostToDevice));
    AXG_HIP(hipMemcpy(h->d_dqs, dqs.data(), dqs.size() * 8, hipMemcpyHostToDevice));
    // lattice points per n^2 (host, once): r2[m] = #(ny,nz) with ny^2+nz^2 = m
    const int N = h->N;
    std::vector<double> r2((size_t)2 * (N / 2) * (N / 2) + 1, 0.0), cnt(h->nn2, 0.0);
    for (int y = 0; y < N; y++)
        for (int z = 0; z < N; z++) {
            int ny = (y < N / 2) ? y : y - N, nz = (z < N / 2) ? z : z - N;
            r2[(size_t)ny * ny + (size_t)nz * nz] += 1;
        }
    // cnt[k] = sum over x of r2[k - nx^2]: each thread owns a range of k (integer counts, exact
    // in any order)
    host_parallel(h->nn2, [&](int64_t lo, int64_t hi) {
        for (int x = 0; x < N; x++) {
            const int64_t nx = (x < N / 2) ? x : x - N, s2 = nx * nx;
            const int64_t k0 = std::max<int64_t>(lo, s2), k1 = std::min<int64_t>(hi, s2 + (int64_t)r2.size());
            for (int64_t k = k0; k < k1; k++) {
                const double v = r2[(size_t)(k - s2)];
                if (v != 0) cnt[(size_t)k] += v;
            }
        }
    });
    h->d_cnt = (double*)h->alloc(h->nn2 * 8);
    AXG_HIP(hipMemcpy(h->d_cnt, cnt.data(), h->nn2 * 8, hipMemcpyHostToDevice));
    h->d_hist = (double*)h->alloc(h->nn2 * 8);
    h->d_plane = (double*)h->alloc((size_t)N * ((size_t)(N / 2) * (N / 2) * 2 + 1) * 8);
}

extern "C" int axg_field_spectrum(axg_handle* h, double tau, double* qmid, double* power, double* conv, double* num,
                                  double* rho, int32_t nbins)
{
    return guarded(h, [&] {
        const axg_params& p = h->p;
        const int nb = p.num_bin;
        require(p.spec_est, "spec_est is off");
        require(nbins == nb, "nbins must equal num_bin");
        require(h->have_flags, "call axg_defects first (the mask)");
        if (h->gs) {
            // compute_conv_spec on the reduced grid: the block means of pi
            // (data2_to_data_spec_degrade) and the blocks' mask, then the same spectrum there
            axg_handle* g = h->gs;
            if (h->comp) launch_degrade_field<float, true>(h->stream, h->N, h->deg, h->pi, h->cpi, (double2*)g->pi);
            else if (h->fp32) launch_degrade_field<float, false>(h->stream, h->N, h->deg, h->pi, nullptr, (double2*)g->pi);
            else launch_degrade_field<double, false>(h->stream, h->N, h->deg, h->pi, nullptr, (double2*)g->pi);
            launch_degrade_mask(h->stream, h->N, h->deg, h->flags, g->flags);
            g->have_flags = true;
            g->have_spec = false;
            forward(h, axg_field_spectrum(g, tau, qmid, power, conv, num, rho, nbins));
            h->have_spec = true;
            return;
        }
        ensure_spec_tables(h);
        const int ncov = nb * (nb + 1) / 2;
        const int nwg = 1024;
        // compute_matrix_deconv (spectrum.f90:358-433): window = dv_spa with mask cells zeroed
        if (h->comp) launch_masked_field<float, true>(h->stream, h->N, nullptr, nullptr, h->flags, h->work, p.dv_spa, false);
        else launch_masked_field<double, false>(h->stream, h->N, nullptr, nullptr, h->flags, h->work, p.dv_spa, false);
        fft3d_work(h, h->work);
        launch_n2_hist(h->stream, h->N, h->work, h->d_plane, h->d_hist);
        double* mat0 = h->red;
        double* tat0 = h->red + ncov;
        require((size_t)2 * ncov * 8 <= h->red_bytes, "reduction scratch too small");
        launch_deconv_mat(h->stream, nb, ncov, 500, h->d_largeq, h->d_matspec, h->d_dqs, h->d_hist, h->d_cnt, h->nn2,
                          p.dwav, mat0, tat0);
        std::vector<double> m0(ncov), t0(ncov);
        AXG_HIP(hipMemcpyAsync(m0.data(), mat0, ncov * 8, hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipMemcpyAsync(t0.data(), tat0, ncov * 8, hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        // compute_conv_spec (spectrum.f90:474-563): masked pi -> FFT -> *dv_spa -> binned
        if (h->comp) launch_masked_field<float, true>(h->stream, h->N, h->pi, h->cpi, h->flags, h->work, 0.0, true);
        else if (h->fp32) launch_masked_field<float, false>(h->stream, h->N, h->pi, nullptr, h->flags, h->work, 0.0, true);
        else launch_masked_field<double, false>(h->stream, h->N, h->pi, nullptr, h->flags, h->work, 0.0, true);
        fft3d_work(h, h->work);
        double* out = h->red + (size_t)nwg * 2 * nb;
        require((size_t)(nwg + 1) * 2 * nb * 8 <= h->red_bytes, "reduction scratch too small");
        launch_spec_bin(h->stream, h->N, h->work, h->spec_bin, h->spec_n2max, nb, p.dv_spa, h->red, nwg, out);
        std::vector<double> r(2 * nb);
        AXG_HIP(hipMemcpyAsync(r.data(), out, 2 * nb * 8, hipMemcpyDeviceToHost, h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        // host algebra: spectrum.f90:433-440, 267-275, 316-342
        std::vector<double> mat(ncov), cs(nb), nm(nb);
        for (int ij = 0; ij < ncov; ij++) mat[ij] = m0[ij] / t0[ij] * p.dv_wav * p.dv_wav;
        // packed-upper Cholesky inverse (dpptrf + dpptri)
        auto AP = [&](int r, int c) -> double& { return mat[(size_t)r + (size_t)c * (c - 1) / 2 - 1]; };
        for (int j = 1; j <= nb; j++) {
            for (int i = 1; i < j; i++) {
                double s = AP(i, j);
                for (int k = 1; k < i; k++) s -= AP(k, i) * AP(k, j);
                AP(i, j) = s / AP(i, i);
            }
            double s = AP(j, j);
            for (int k = 1; k < j; k++) s -= AP(k, j) * AP(k, j);
            if (!(s > 0)) throw Error(AXG_ENUMERIC, "deconvolution matrix is not positive definite");
            AP(j, j) = sqrt(s);
        }
        std::vector<double> col(nb);
        for (int j = 1; j <= nb; j++) {
            AP(j, j) = 1 / AP(j, j);
            double ajj = -AP(j, j);
            for (int i = 1; i < j; i++) {
                double s = 0;
                for (int k = i; k < j; k++) s += AP(i, k) * AP(k, j);
                col[i - 1] = s * ajj;
            }
            for (int i = 1; i < j; i++) AP(i, j) = col[i - 1];
        }
        for (int i = 1; i <= nb; i++)
            for (int j = i; j <= nb; j++) {
                double s = 0;
                for (int k = j; k <= nb; k++) s += AP(i, k) * AP(j, k);
                AP(i, j) = s;
            }
        double t2 = tau * tau;
        for (int i = 0; i < nb; i++) {
            cs[i] = r[i] * p.dv_wav * t2;
            nm[i] = r[nb + i];
        }
        for (int i = 1; i <= nb; i++)
            for (int j = 1; j <= i; j++) {
                int ij = i * (i - 1) / 2 + j;
                mat[ij - 1] = mat[ij - 1] / nm[i - 1] / nm[j - 1];
            }
        std::vector<double> pw(nb, 0.0);
        for (int i = 1; i <= nb; i++)
            for (int j = 1; j <= nb; j++) {
                int a = std::min(i, j), b = std::max(i, j);
                pw[i - 1] += mat[(size_t)a + (size_t)b * (b - 1) / 2 - 1] * cs[j - 1];
            }
        double rh = 0;
        for (int i = 0; i < nb; i++) {
            double q = (h->wav[i + 1] + h->wav[i]) / 2;
            double dq = h->wav[i + 1] - h->wav[i];
            double q2dq = q * q * dq / AXG_TWOPI_T / AXG_PI_T;
            rh = rh + pw[i] * q2dq;
            if (qmid) qmid[i] = q;
            if (power) power[i] = pw[i];
            if (conv) conv[i] = cs[i];
            if (num) num[i] = nm[i];
        }
        if (rho) *rho = rh;
        h->spec_power = pw;
        h->spec_mat = mat;
        h->have_spec = true;
    });
}

// estimate_covmat (spectrum.f90:569-734): Monte Carlo covariance of the deconvolved spectrum of
// the last axg_field_spectrum.  Each realisation: Gaussian modes with amplitude
// sqrt(power/2)/N^3 per bin, FORWARD transform, mask cells zeroed, BACKWARD transform, binned
// |.|^2, deconvolved, minus the power; outer products summed, with leave-one-out sums for the
// jackknife test on the diagonal after min_mc realisations (stop when below tol).
extern "C" int axg_field_covariance(axg_handle* h, uint64_t seed, int32_t min_mc, int32_t max_mc, double tol,
                                    double* cov, int64_t ncov_out, int32_t* iters)
{
    return guarded(h, [&] {
        const int nb = h->p.num_bin;
        const int ncov = nb * (nb + 1) / 2;
        require(h->have_spec && h->have_flags, "call axg_field_spectrum first (after the last change of the fields)");
        require(cov != nullptr && ncov_out == ncov, "cov needs num_bin*(num_bin+1)/2 entries");
        if (h->gs) {
            // estimate_covmat on the reduced grid, drawing from this handle's stream
            axg_handle* g = h->gs;
            g->stream_on = h->stream_on;
            g->stream_x0 = h->stream_x0;
            g->stream_pos = h->stream_pos;
            forward(h, axg_field_covariance(g, seed, min_mc, max_mc, tol, cov, ncov_out, iters));
            h->stream_pos = g->stream_pos;
            return;
        }
        require(min_mc >= 2 && max_mc >= min_mc && max_mc <= 100000, "need 2 <= min_mc <= max_mc");
        const int N = h->N;
        const double ntot = (double)h->n3;
        std::vector<double> sig(nb);
        for (int i = 0; i < nb; i++) sig[i] = sqrt(h->spec_power[i] / 2) / ntot;
        double* d_sig = h->red;
        const int nwg = 1024;
        double* out = h->red + nb + (size_t)nwg * 2 * nb;
        double* part = h->red + nb;
        require((size_t)(nb + (size_t)(nwg + 1) * 2 * nb) * 8 <= h->red_bytes, "reduction scratch too small");
        AXG_HIP(hipMemcpyAsync(d_sig, sig.data(), nb * 8, hipMemcpyHostToDevice, h->stream));
        std::vector<double> pc(ncov, 0.0), tmp((size_t)ncov * max_mc, 0.0), r(2 * nb), pw(nb);
        const std::vector<double>& mat = h->spec_mat;
        int iter = 1;
        bool converged = false;
        const Draws dr = h->draws((uint32_t)seed, (uint32_t)(seed >> 32));
        for (; iter <= max_mc; iter++) {
            launch_cov_modes(h->stream, N, dr, iter, h->spec_bin, d_sig, h->work);
            fft3d_work(h, h->work, -1);
            launch_mask_zero(h->stream, h->n3, h->flags, h->work);
            fft3d_work(h, h->work, +1);
            launch_spec_bin(h->stream, N, h->work, h->spec_bin, h->spec_n2max, nb, 1.0, part, nwg, out);
            AXG_HIP(hipMemcpyAsync(r.data(), out, 2 * nb * 8, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
            // pow = mat_deconv pow0 - power (dspmv with beta = 1, spectrum.f90:675-677)
            for (int i = 1; i <= nb; i++) {
                double acc = -h->spec_power[i - 1];
                for (int j = 1; j <= nb; j++) {
                    int a = std::min(i, j), b = std::max(i, j);
                    acc += mat[(size_t)a + (size_t)b * (b - 1) / 2 - 1] * r[j - 1];
                }
                pw[i - 1] = acc;
            }
            for (int i = 1; i <= nb; i++)
                for (int j = 1; j <= i; j++) {
                    const int ij = i * (i - 1) / 2 + j - 1;
                    const double u = pw[i - 1] * pw[j - 1];
                    tmp[(size_t)ij * max_mc + (iter - 1)] = pc[ij];
                    for (int k = 0; k < iter - 1; k++) tmp[(size_t)ij * max_mc + k] += u;
                    pc[ij] += u;
                }
            if (iter < min_mc) continue;
            double diff = 0;
            for (int i = 1; i <= nb; i++) {
                const int ij = i * (i + 1) / 2 - 1;
                double u = 0, v = 0;
                for (int k = 0; k < iter; k++) {
                    double x = tmp[(size_t)ij * max_mc + k] / (iter - 1);
                    u = u + x;
                    v = v + x * x;
                }
                u = u / iter;
                v = v / iter;
                v = sqrt(v - u * u);
                double x = pc[ij] / iter;
                diff = std::max(diff, v / x);
            }
            if (diff < tol) {
                converged = true;
                break;
            }
        }
        // without convergence the loop counter ends at max_mc + 1, which the reference divides by;
        // the stream has given 4 N^3 draws to each realisation run
        if (h->stream_on) h->stream_pos += 4 * (uint64_t)h->n3 * (uint64_t)(converged ? iter : max_mc);
        for (int ij = 0; ij < ncov; ij++) cov[ij] = pc[ij] / iter;
        if (iters) *iters = converged ? iter : -iter;
    });
}

// ------------------------------------------------------------------ batches of lattices

// A captured graph and its instance, destroyed with the scope.
struct GraphPair {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~GraphPair()
    {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
    }
};

struct EventSet {
    std::vector<hipEvent_t> e;
    explicit EventSet(size_t n)
    {
        for (size_t i = 0; i < n; i++) {
            hipEvent_t v;
            AXG_HIP(hipEventCreateWithFlags(&v, hipEventDisableTiming));
            e.push_back(v);
        }
    }
    ~EventSet()
    {
        for (hipEvent_t v : e) (void)hipEventDestroy(v);
    }
};

extern "C" int axg_batch_advance(axg_handle* const* hs, int32_t n, double t1, double dt, int32_t nsteps,
                                 int32_t chunk, double* ms)
{
    axg_handle* h0 = (hs && n > 0) ? hs[0] : nullptr;
    return guarded(h0, [&] {
        require(hs != nullptr && n >= 1, "need at least one handle");
        require(nsteps >= 1 && dt > 0 && t1 > 0 && std::isfinite(dt) && std::isfinite(t1),
                "need nsteps >= 1, dt > 0 and t1 > 0");
        require(chunk == 0 || (chunk >= 2 && chunk % 2 == 0 && chunk <= 64),
                "chunk: 0 (no graph) or an even number of steps from 2 to 64");
        for (int i = 0; i < n; i++) {
            require(hs[i] != nullptr, "null handle");
            require(!hs[i]->gs, "degrade_spec > 0: advance such a lattice with axg_advance_steps");
            require(hs[i]->p.device == h0->p.device, "all handles must be on one device");
            for (int j = 0; j < i; j++) require(hs[j] != hs[i], "a handle appears twice");
        }
        // each lattice's table: the scalars run_steps would pass, step by step
        for (int i = 0; i < n; i++) {
            axg_handle* h = hs[i];
            const axg_params& p = h->p;
            std::vector<StepEntry> tab(nsteps);
            double t = t1 + dt / 2;
            for (int32_t s = 0; s < nsteps; s++) {
                const double dtd = s == nsteps - 1 ? dt / 2 : dt;
                tab[s].k = kick_scalars(p, t, dt);
                tab[s].dtd = dtd;
                tab[s].tau = t;
                tab[s].dw = t * dt * p.dv_spa * (p.sigma_cr * p.sigma_cr);
                t = t + dtd;
            }
            if (nsteps > h->steptab_cap) {
                void* q = h->steptab;
                AXG_HIP(hipStreamSynchronize(h->stream));
                h->release_one(q);
                h->steptab = nullptr;
                h->steptab_cap = 0;
                h->steptab = (StepEntry*)h->alloc((size_t)nsteps * sizeof(StepEntry));
                h->steptab_cap = nsteps;
            }
            AXG_HIP(hipMemcpyAsync(h->steptab, tab.data(), (size_t)nsteps * sizeof(StepEntry), hipMemcpyHostToDevice,
                                   h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
        }
        // units of work: a stack of lattices that share N, precision, dspa, dwav and the GW index
        // order (one launch per pass over all of them, in the pass configurations of the stack's
        // first lattice: every configuration gives the same bits), or one lattice the stacked
        // kernels do not cover (no GW, no fused pass, runtime-plan sizes).  AXG_BATCH_STACK=0 makes
        // every lattice its own unit (same results).
        bool stack = true;
        if (const char* e = getenv("AXG_BATCH_STACK")) stack = strcmp(e, "0") != 0;
        // a stack is closed before its step working set (fields, source transforms, accumulators)
        // passes 210 MiB, so a stack's passes hand their arrays to each other through the 256 MB
        // last-level cache (measured on 16 x 64^3 + 16 x 96^3 + 16 x 128^3, 600 steps, graph of
        // 20 steps: 1.319-1.321 s at 210-240 MiB, where the 102 MiB 96^3 lattices go two to a
        // stack and the 33 MiB 64^3 ones six or seven, against 1.332-1.334 s at 160 MiB on the same
        // GPU (1.344 s at 160-200 MiB on another, 1.359-1.360 s at 128, 1.365-1.366 s at 96); a
        // 128^3 lattice takes 232 MiB and stays alone; with 10-step graphs, earlier: 1.44-1.45 s
        // at 32-64 MiB, 1.46 s unlimited); tuning knob AXG_BATCH_STACK_MB (0 = no limit)
        double stack_mb = 210;
        if (const char* e = getenv("AXG_BATCH_STACK_MB")) stack_mb = atof(e);
        auto step_mb = [](const axg_handle* h) {
            const double fields = (double)h->n3 * (3 * (double)h->t2 + (h->comp ? 12.0 : 0.0));
            const double gw = h->A ? ((double)6 * h->N * h->N * h->P + 10.0 * h->L.comp()) * (double)h->t2 : 0.0;
            return (fields + gw) / (1 << 20);
        };
        std::vector<double> unit_mb;
        struct Unit {
            std::vector<int> idx;
            bool stacked = false;
            int off = 0;  // first LatRef of a stacked unit
        };
        std::vector<Unit> units;
        for (int i = 0; i < n; i++) {
            axg_handle* h = hs[i];
            const bool can = stack && h->p.grav_spec && h->fuse_x && p2_size(h->N) && h->L.zt > 0;
            Unit* u = nullptr;
            const double mb = step_mb(h);
            if (can) {
                for (size_t v = 0; v < units.size(); v++) {
                    if (!units[v].stacked) continue;
                    const axg_handle* g = hs[units[v].idx[0]];
                    if (g->N == h->N && g->prec == h->prec && g->p.dspa == h->p.dspa && g->p.dwav == h->p.dwav &&
                        g->p.gw_fix_swap == h->p.gw_fix_swap && (stack_mb <= 0 || unit_mb[v] + mb <= stack_mb)) {
                        u = &units[v];
                        unit_mb[v] += mb;
                        break;
                    }
                }
            }
            if (!u) {
                units.emplace_back();
                unit_mb.push_back(mb);
                u = &units.back();
                u->stacked = can;
            }
            u->idx.push_back(i);
        }
        std::vector<LatRef> refs;
        for (auto& u : units) {
            if (!u.stacked) continue;
            u.off = (int)refs.size();
            for (int i : u.idx) {
                const axg_handle* h = hs[i];
                LatRef r{};
                r.phi[0] = (gptr)h->phi[h->cur];
                r.phi[1] = (gptr)h->phi[1 - h->cur];
                r.cphi[0] = (gptr)h->cphi[h->cur];
                r.cphi[1] = (gptr)h->cphi[1 - h->cur];
                r.pi = (gptr)h->pi;
                r.cpi = (gptr)h->cpi;
                r.W = (gptr)h->W;
                r.W2 = (gptr)h->W2;
                r.A = (gptr)h->A;
                r.tab = h->steptab;
                refs.push_back(r);
            }
        }
        if ((int64_t)refs.size() > h0->latref_cap) {
            void* q = h0->latref;
            h0->release_one(q);
            h0->latref = nullptr;
            h0->latref_cap = 0;
            h0->latref = (LatRef*)h0->alloc(refs.size() * sizeof(LatRef));
            h0->latref_cap = (int64_t)refs.size();
        }
        if (!h0->stepbase) h0->stepbase = (int*)h0->alloc(sizeof(int));
        if (!refs.empty())
            AXG_HIP(hipMemcpyAsync(h0->latref, refs.data(), refs.size() * sizeof(LatRef), hipMemcpyHostToDevice,
                                   h0->stream));
        const int zero = 0;
        AXG_HIP(hipMemcpyAsync(h0->stepbase, &zero, sizeof zero, hipMemcpyHostToDevice, h0->stream));
        AXG_HIP(hipStreamSynchronize(h0->stream));
        int* base = h0->stepbase;
        auto step_unit = [&](const Unit& u, int sidx) {
            if (u.stacked) stacked_step_any(hs[u.idx[0]], h0->latref + u.off, (int)u.idx.size(), sidx, base);
            else table_step_any(hs[u.idx[0]], sidx, base);
        };
        const int nu = (int)units.size();
        const auto c0 = std::chrono::steady_clock::now();
        // first half drifts
        for (int i = 0; i < n; i++) {
            axg_handle* h = hs[i];
            if (h->comp) launch_drift<float, true>(h->stream, h->N, h->phi[h->cur], h->phi[h->cur], h->pi,
                                                   h->cphi[h->cur], h->cphi[h->cur], h->cpi, dt / 2);
            else if (h->fp32) launch_drift<float, false>(h->stream, h->N, h->phi[h->cur], h->phi[h->cur], h->pi,
                                                         nullptr, nullptr, nullptr, dt / 2);
            else launch_drift<double, false>(h->stream, h->N, h->phi[h->cur], h->phi[h->cur], h->pi, nullptr,
                                             nullptr, nullptr, dt / 2);
        }
        for (int i = 0; i < n; i++) AXG_HIP(hipStreamSynchronize(hs[i]->stream));
        int32_t done = 0;
        if (chunk > 0 && nsteps >= chunk) {
            // one graph of `chunk` steps of every unit, replayed; the step counter advances by chunk
            // at its end, and an even chunk returns the ping-pong buffers to where the capture found
            // them.  The units are dealt to S lanes (the streams of the first S handles, forked from
            // and joined into the first): unit u runs on lane u % S after the lane's previous unit.
            // Two lanes by default; tuning knob AXG_BATCH_LANES.
            int S = std::min(nu, 2);
            if (const char* e = getenv("AXG_BATCH_LANES")) S = std::max(1, std::min(nu, atoi(e)));
            hipStream_t s0 = h0->stream;
            EventSet ev((size_t)n + 1);
            GraphPair gp;
            std::vector<hipStream_t> own(n);
            for (int i = 0; i < n; i++) own[i] = hs[i]->stream;
            AXG_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
            AXG_HIP(hipEventRecord(ev.e[n], s0));
            for (int l = 1; l < S; l++) AXG_HIP(hipStreamWaitEvent(own[l], ev.e[n], 0));
            try {
                for (int u = 0; u < nu; u++) {
                    axg_handle* h = hs[units[u].idx[0]];
                    h->stream = own[u % S];
                    for (int32_t c = 0; c < chunk; c++) step_unit(units[u], c);
                    h->stream = own[units[u].idx[0]];
                }
            } catch (...) {
                for (int i = 0; i < n; i++) hs[i]->stream = own[i];
                throw;
            }
            for (int l = 1; l < S; l++) {
                AXG_HIP(hipEventRecord(ev.e[l], own[l]));
                AXG_HIP(hipStreamWaitEvent(s0, ev.e[l], 0));
            }
            launch_add_int(s0, base, chunk);
            AXG_HIP(hipStreamEndCapture(s0, &gp.g));
            AXG_HIP(hipGraphInstantiate(&gp.x, gp.g, nullptr, nullptr, 0));
            for (; done + chunk <= nsteps; done += chunk) AXG_HIP(hipGraphLaunch(gp.x, s0));
            AXG_HIP(hipEventRecord(ev.e[0], s0));
            for (int i = 1; i < n; i++) AXG_HIP(hipStreamWaitEvent(hs[i]->stream, ev.e[0], 0));
            // the steps left over: entries *base + s = done + s
            for (int32_t s = 0; s < nsteps - done; s++)
                for (const Unit& u : units) step_unit(u, s);
        } else {
            for (int32_t s = 0; s < nsteps; s++)
                for (const Unit& u : units) step_unit(u, s);
        }
        for (int i = 0; i < n; i++) AXG_HIP(hipStreamSynchronize(hs[i]->stream));
        for (const Unit& u : units)
            if (u.stacked)
                for (int i : u.idx) hs[i]->cur ^= (nsteps & 1);
        const auto c1 = std::chrono::steady_clock::now();
        for (int i = 0; i < n; i++) {
            axg_handle* h = hs[i];
            h->have_flags = false;
            h->have_plaq = false;
            h->have_spec = false;
            h->nstr_rec = -1;
            h->nwall_rec = -1;
            h->have_grp[0] = h->have_grp[1] = false;
        }
        if (ms) *ms = std::chrono::duration<double, std::milli>(c1 - c0).count();
    });
}

// ------------------------------------------------------------------ gradient (adjoint)

template <typename T, bool COMP>
static void gradient_impl(axg_handle* h, double t1, double dt, int32_t nsteps, int32_t K, double qlo, double qhi,
                          axg_grad_out* out)
{
    const axg_params& p = h->p;
    const int N = h->N;
    const int64_t n3 = h->n3;
    const int nh = N / 2 + 1;
    const int64_t hs = (int64_t)N * N * nh;
    const double s2c = p.sigma_cr * p.sigma_cr;
    const size_t fsz = h->fp32 ? 8 : 16, csz = h->comp ? 4 : 0;
    const size_t state = (size_t)n3 * (2 * fsz + 2 * csz);
    const int64_t nck = (nsteps - 1) / K + 1 + 2;  // kicks 0, K, 2K, ...; the initial and final states
    if (!h->adj) {
        h->adj = (double2*)h->alloc((size_t)5 * n3 * sizeof(double2));
        h->abar = (double2*)h->alloc((size_t)4 * hs * sizeof(double2));
    }
    if (nck > h->ckpt_cap) {
        void* q = h->ckpt;
        AXG_HIP(hipStreamSynchronize(h->stream));
        h->release_one(q);
        h->ckpt = nullptr;
        h->ckpt_cap = 0;
        h->ckpt = h->alloc(state * (size_t)nck);
        h->ckpt_cap = nck;
    }
    double2* phibar = h->adj;
    double2* pibar = h->adj + n3;       // pibar ping-pongs with pibar2 across the adjoint kicks
    double2* pibar2 = h->adj + 4 * n3;
    // the adjoint sources in the storage precision (FP32 for FP32 and FP32C lattices)
    void* W0 = h->work;
    void* W1 = h->adj + 2 * n3;
    void* W2 = h->adj + 3 * n3;
    auto fft_back = [&](void* d) {
        launch_fft_rows<T, -1>(h->stream, N, d, (int64_t)N * N, h->plan, h->tw);
        launch_fft_cols<T, -1>(h->stream, N, d, N, (int64_t)N * N, N, N, h->plan, h->tw);
        launch_fft_cols<T, -1>(h->stream, N, d, (int64_t)N * N, N, N, N, h->plan, h->tw);
    };
    // y and z only: the compile-time sizes do the x transform inside the gbar kernel
    auto fft_back_yz = [&](void* d) {
        launch_fft_cols<T, -1>(h->stream, N, d, N, (int64_t)N * N, N, N, h->plan, h->tw);
        launch_fft_cols<T, -1>(h->stream, N, d, (int64_t)N * N, N, N, N, h->plan, h->tw);
    };
    const bool fused_x = p2_size(N);
    auto slot = [&](int64_t j) { return (char*)h->ckpt + state * (size_t)j; };
    auto save = [&](int64_t j) {
        char* b = slot(j);
        AXG_HIP(hipMemcpyAsync(b, h->phi[h->cur], n3 * fsz, hipMemcpyDeviceToDevice, h->stream));
        AXG_HIP(hipMemcpyAsync(b + n3 * fsz, h->pi, n3 * fsz, hipMemcpyDeviceToDevice, h->stream));
        if (csz) {
            AXG_HIP(hipMemcpyAsync(b + 2 * n3 * fsz, h->cphi[h->cur], n3 * csz, hipMemcpyDeviceToDevice, h->stream));
            AXG_HIP(hipMemcpyAsync(b + 2 * n3 * fsz + n3 * csz, h->cpi, n3 * csz, hipMemcpyDeviceToDevice, h->stream));
        }
    };
    auto load = [&](int64_t j) {
        const char* b = slot(j);
        AXG_HIP(hipMemcpyAsync(h->phi[h->cur], b, n3 * fsz, hipMemcpyDeviceToDevice, h->stream));
        AXG_HIP(hipMemcpyAsync(h->pi, b + n3 * fsz, n3 * fsz, hipMemcpyDeviceToDevice, h->stream));
        if (csz) {
            AXG_HIP(hipMemcpyAsync(h->cphi[h->cur], b + 2 * n3 * fsz, n3 * csz, hipMemcpyDeviceToDevice, h->stream));
            AXG_HIP(hipMemcpyAsync(h->cpi, b + 2 * n3 * fsz + n3 * csz, n3 * csz, hipMemcpyDeviceToDevice, h->stream));
        }
    };
    const int64_t J_INIT = nck - 2, J_FINAL = nck - 1;

    // forward: the accumulators from zero, checkpoints before kicks 0, K, 2K, ...
    AXG_HIP(hipStreamSynchronize(h->stream));
    const auto c0 = std::chrono::steady_clock::now();
    AXG_HIP(hipMemsetAsync(h->A, 0, (size_t)4 * h->L.comp() * fsz, h->stream));
    save(J_INIT);
    std::vector<double> ts(nsteps);
    run_steps_hooked<T, COMP>(h, t1, dt, nsteps, [&](int32_t s, double t) {
        ts[s] = t;
        if (s % K == 0) save(s / K);
    });
    save(J_FINAL);
    AXG_HIP(hipStreamSynchronize(h->stream));
    const auto c1 = std::chrono::steady_clock::now();

    // J and its seed: w_b = fact q_b^3 / denom_b over the band, abar = 2 w mult A
    const int nb = p.num_bin_grav;
    std::vector<double> qm(nb), pw(nb), nu(nb), de(nb), w(nb, 0.0);
    if (axg_gw_spectrum(h, qm.data(), pw.data(), nu.data(), de.data(), nb) != 0) throw Error(AXG_EDEVICE, h->err);
    double J = 0;
    {
        const double bigh_in_mstar = 2.1332e-42 / 2.435e18;
        double t6 = p.tau_cr * p.tau_cr * p.tau_cr;
        t6 = t6 * t6;
        double a4 = p.a_now * p.a_now;
        a4 = a4 * a4;
        const double fact = 1 / t6 / (bigh_in_mstar * bigh_in_mstar) / a4 / 3 / (AXG_TWOPI_T * AXG_TWOPI_T) * p.dv_wav;
        for (int b = 0; b < nb; b++)
            if (qm[b] >= qlo && qm[b] <= qhi && de[b] > 0) {
                w[b] = fact * (qm[b] * qm[b] * qm[b]) / de[b];
                J = J + pw[b];
            }
    }
    // the adjoint is linear in its seed: run it on w / wmax (the spectrum's prefactor can be
    // ~1e120, beyond FP32) and scale the results back
    double wmax = 0;
    for (int b = 0; b < nb; b++) wmax = std::max(wmax, w[b]);
    const double scale = wmax > 0 ? wmax : 1.0;
    for (int b = 0; b < nb; b++) w[b] = w[b] / scale;
    double* d_w = h->red;
    double* d_tot = h->red + nb;      // 2 totals
    double* d_col = h->red + nb + 2;  // the dot product
    double* d_part = h->red + nb + 4;  // [block][2] sums over the steps
    require((size_t)(nb + 4 + 2 * ADJ_GRID_H) * 8 <= h->red_bytes, "reduction scratch too small");
    AXG_HIP(hipMemcpyAsync(d_w, w.data(), nb * 8, hipMemcpyHostToDevice, h->stream));
    for (int m = 0; m < 4; m++) launch_gw_pack<T>(h->stream, N, h->L, h->A, m, h->abar + m * hs, +1);
    launch_adj_seed(h->stream, N, h->abar, h->gw_bin, d_w);
    AXG_HIP(hipMemsetAsync(phibar, 0, (size_t)2 * n3 * sizeof(double2), h->stream));
    AXG_HIP(hipMemsetAsync(d_part, 0, (size_t)2 * ADJ_GRID_H * sizeof(double), h->stream));

    // backward: from the final state, undo the last half drift, then per kick s (last first):
    // the adjoint of its GW update and of the kick, the kick undone (-dt), the exact state from
    // the checkpoint when there is one, and the preceding drift undone with its adjoint
    const double inv_d = 1 / (2 * p.dspa);
    launch_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[h->cur], h->pi, h->cphi[h->cur], h->cphi[h->cur],
                          h->cpi, -dt / 2);
    for (int32_t s = nsteps - 1; s >= 0; s--) {
        const double t = ts[s];
        AdjStepScalars g;
        g.tau = t;
        g.dw = t * dt * p.dv_spa * s2c;
        g.dwav = p.dwav;
        g.fix_swap = p.gw_fix_swap;
        g.last = s == nsteps - 1;
        launch_adj_modes<T>(h->stream, N, h->abar, g, W0, W1, W2);
        if (fused_x) {
            fft_back_yz(W0);
            fft_back_yz(W1);
            fft_back_yz(W2);
            if (!launch_adj_gbar_xfft<T, COMP, T>(h->stream, N, h->phi[h->cur], h->cphi[h->cur], inv_d, W0, W1, W2,
                                                  h->tw))
                throw Error(AXG_EDEVICE, "gradient: no compile-time transform for this size");
        } else {
            fft_back(W0);
            fft_back(W1);
            fft_back(W2);
            launch_adj_gbar<T, COMP, T>(h->stream, N, h->phi[h->cur], h->cphi[h->cur], inv_d, W0, W1, W2);
        }
        const KickScalars k = kick_scalars(p, t, dt);
        const double hd = s == 0 ? dt / 2 : dt;  // the drift before kick s
        const double t2 = t * t;
        launch_adj_kick<T, COMP, T>(h->stream, N, h->phi[h->cur], h->cphi[h->cur], pibar, phibar, pibar2, W0, W1, W2, k,
                                 inv_d, hd, k.f3 / s2c, k.f3 * (t2 * t2) / s2c, true, d_part);
        std::swap(pibar, pibar2);
        // undo the kick (the same kick with -dt at the same time) and the drift before it; at a
        // checkpoint the exact state comes back between the two
        const bool ck = s % K == 0;
        const int nxt = 1 - h->cur;
        launch_kick_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[nxt], h->pi, h->cphi[h->cur], h->cphi[nxt],
                                   h->cpi, nullptr, kick_scalars(p, t, -dt), p.dspa, ck ? 0.0 : -hd, false);
        h->cur = nxt;
        if (ck) {
            load(s / K);
            launch_drift<T, COMP>(h->stream, N, h->phi[h->cur], h->phi[h->cur], h->pi, h->cphi[h->cur],
                                  h->cphi[h->cur], h->cpi, -hd);
        }
    }
    sum_columns(h->stream, d_part, ADJ_GRID_H, 2, d_tot);
    // gradients: the dynamics' parts plus the explicit sigma_cr^2 in dw (J is quadratic in A)
    load(J_INIT);
    double* d_dot = d_col;
    launch_adj_dot<T, COMP>(h->stream, n3, h->phi[h->cur], h->cphi[h->cur], h->pi, h->cpi, phibar, pibar, d_part, d_dot);
    double tot[2], dot = 0;
    AXG_HIP(hipMemcpyAsync(tot, d_tot, sizeof tot, hipMemcpyDeviceToHost, h->stream));
    AXG_HIP(hipMemcpyAsync(&dot, d_dot, sizeof dot, hipMemcpyDeviceToHost, h->stream));
    AXG_HIP(hipStreamSynchronize(h->stream));
    const auto c2 = std::chrono::steady_clock::now();
    load(J_FINAL);
    AXG_HIP(hipStreamSynchronize(h->stream));
    h->adj_pibar = pibar;
    h->adj_scale = scale;
    h->have_adj = true;
    out->J = J;
    out->dJ_dsigma2 = tot[0] * scale + 2 * J / s2c;
    out->dJ_dkappa0 = tot[1] * scale;
    out->dJ_dlnamp = dot * scale;
    out->forward_ms = std::chrono::duration<double, std::milli>(c1 - c0).count();
    out->adjoint_ms = std::chrono::duration<double, std::milli>(c2 - c1).count();
}

extern "C" int axg_gradient(axg_handle* h, double t1, double dt, int32_t nsteps, int32_t K, double qlo, double qhi,
                            axg_grad_out* out)
{
    return guarded(h, [&] {
        require(out != nullptr, "null output");
        require(!h->gs, "axg_gradient differentiates the lattice-grid GW spectrum (degrade_spec = 0)");
        require(h->p.grav_spec && h->A != nullptr, "grav_spec is off: no GW spectrum to differentiate");
        require(nsteps >= 1 && K >= 1, "need nsteps >= 1 and K >= 1");
        require(dt > 0 && t1 > 0 && std::isfinite(dt) && std::isfinite(t1), "need t1 > 0 and dt > 0");
        require(qlo <= qhi, "need qlo <= qhi");
        h->have_adj = false;
        if (h->comp) gradient_impl<float, true>(h, t1, dt, nsteps, K, qlo, qhi, out);
        else if (h->fp32) gradient_impl<float, false>(h, t1, dt, nsteps, K, qlo, qhi, out);
        else gradient_impl<double, false>(h, t1, dt, nsteps, K, qlo, qhi, out);
        h->have_flags = false;
        h->have_plaq = false;
        h->have_spec = false;
        h->nstr_rec = -1;
        h->nwall_rec = -1;
        h->have_grp[0] = h->have_grp[1] = false;
    });
}

extern "C" int axg_get_adjoint(axg_handle* h, double* phibar, double* pibar, int64_t nsite)
{
    return guarded(h, [&] {
        require(h->have_adj, "call axg_gradient first");
        require(nsite == h->n3, "nsite must be N^3");
        if (phibar)
            AXG_HIP(hipMemcpyAsync(phibar, h->adj, (size_t)h->n3 * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        if (pibar)
            AXG_HIP(hipMemcpyAsync(pibar, h->adj_pibar, (size_t)h->n3 * sizeof(double2), hipMemcpyDeviceToHost,
                                   h->stream));
        AXG_HIP(hipStreamSynchronize(h->stream));
        for (int64_t i = 0; i < 2 * h->n3; i++) {
            if (phibar) phibar[i] *= h->adj_scale;
            if (pibar) pibar[i] *= h->adj_scale;
        }
    });
}

// ------------------------------------------------------------------ initial field

static void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    // Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        uint32_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = l1; c[2] = n2; c[3] = l0;
        k0 += W0; k1 += W1;
    }
}

extern "C" int axg_init_thermal(axg_handle* h, uint64_t seed, double tau_ini)
{
    return guarded(h, [&] {
        // at tau_ini >= 1 the temperature is at or below the transition, m^2 <= 0, and the
        // reference's zero mode has no finite amplitude (osc = 0 or imaginary)
        require(tau_ini > 0 && tau_ini < 1, "the thermal initial state needs 0 < tau_ini < 1 (T above T_cr)");
        const axg_params& p = h->p;
        const Draws dr = h->draws((uint32_t)seed, (uint32_t)(seed >> 32));
        for (int ip = 1; ip <= 2; ip++) {
            launch_thermal_modes(h->stream, h->N, dr, ip, tau_ini, p.sigma_cr, p.dwav, p.dv_wav, AXG_THRM_SIGMA,
                                 h->work);
            fft3d_work(h, h->work, -1);
            if (ip == 1) store_from_work(h, h->phi[h->cur], h->cphi[h->cur]);
            else store_from_work(h, h->pi, h->cpi);
        }
        AXG_HIP(hipStreamSynchronize(h->stream));
        if (h->stream_on) h->stream_pos += 8 * (uint64_t)h->n3;  // 2 fields x N^3 modes x 4 draws
        h->have_flags = false;
        h->have_plaq = false;
        h->have_spec = false;
        h->nstr_rec = -1;
        h->nwall_rec = -1;
        h->have_grp[0] = h->have_grp[1] = false;
    });
}

extern "C" int axg_set_stream(axg_handle* h, uint32_t seed1, uint32_t seed2, uint64_t position)
{
    return guarded(h, [&] {
        uint64_t x = ((uint64_t)seed1 + ((uint64_t)seed2 << 32)) & MCG59_MASK;
        h->stream_x0 = x ? x : 1;  // vslNewStreamEx's start for two seed words
        h->stream_pos = position;
        h->stream_on = true;
    });
}

extern "C" int axg_get_stream(axg_handle* h, uint64_t* position)
{
    return guarded(h, [&] {
        require(position != nullptr, "null position");
        require(h->stream_on, "no stream set (axg_set_stream)");
        *position = h->stream_pos;
    });
}

extern "C" int axg_init_flat(axg_handle* h, uint64_t seed, double kmax)
{
    return guarded(h, [&] {
        require(kmax > 0, "kmax must be positive");
        const int N = h->N;
        const double dwav = h->p.dwav;
        std::vector<int64_t> idx;
        std::vector<double> val;
        int rmax = (int)std::min<double>(N / 2, ceil(kmax / dwav) + 1);
        for (int nz = -rmax; nz <= rmax; nz++)
            for (int ny = -rmax; ny <= rmax; ny++)
                for (int nx = -rmax; nx <= rmax; nx++) {
                    if (nx < -N / 2 || nx >= N / 2 || ny < -N / 2 || ny >= N / 2 || nz < -N / 2 || nz >= N / 2) continue;
                    double q = sqrt((double)((int64_t)nx * nx + (int64_t)ny * ny + (int64_t)nz * nz)) * dwav;
                    if (!(q < kmax)) continue;
                    int64_t ix = nx < 0 ? nx + N : nx, iy = ny < 0 ? ny + N : ny, iz = nz < 0 ? nz + N : nz;
                    uint64_t lin = (uint64_t)ix + (uint64_t)N * ((uint64_t)iy + (uint64_t)N * (uint64_t)iz);
                    uint32_t c[4] = {(uint32_t)lin, (uint32_t)(lin >> 32), 0x6178696fu, 0x6e677276u};
                    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
                    uint64_t a = (((uint64_t)c[0] << 21) ^ (uint64_t)(c[1] >> 11)) & ((1ull << 53) - 1);
                    uint64_t b = (((uint64_t)c[2] << 21) ^ (uint64_t)(c[3] >> 11)) & ((1ull << 53) - 1);
                    double u1 = ((double)a + 1.0) / 9007199254740992.0;
                    double u2 = ((double)b + 1.0) / 9007199254740992.0;
                    double r = sqrt(-2.0 * log(u1));
                    double th = 6.283185307179586 * u2;
                    idx.push_back((int64_t)lin);
                    val.push_back(r * cos(th));
                    val.push_back(r * sin(th));
                }
        require(!idx.empty(), "no mode below kmax");
        int64_t nm = (int64_t)idx.size();
        AXG_HIP(hipMemsetAsync(h->work, 0, (size_t)h->n3 * 16, h->stream));
        int64_t* d_idx = (int64_t*)h->red;
        double2* d_val = nullptr;
        std::vector<void*> tmp;
        void* pidx = nullptr;
        void* pval = nullptr;
        AXG_HIP(hipMalloc(&pidx, nm * 8));
        AXG_HIP(hipMalloc(&pval, nm * 16));
        try {
            AXG_HIP(hipMemcpyAsync(pidx, idx.data(), nm * 8, hipMemcpyHostToDevice, h->stream));
            AXG_HIP(hipMemcpyAsync(pval, val.data(), nm * 16, hipMemcpyHostToDevice, h->stream));
            d_idx = (int64_t*)pidx;
            d_val = (double2*)pval;
            launch_scatter_modes(h->stream, N, d_idx, d_val, nm, h->work);
            fft3d_work(h, h->work);
            const int nwg = 1024;
            double* dsum = h->red + nwg;
            launch_sumsq(h->stream, h->n3, h->work, h->red, nwg, dsum);
            double s;
            AXG_HIP(hipMemcpyAsync(&s, dsum, 8, hipMemcpyDeviceToHost, h->stream));
            AXG_HIP(hipStreamSynchronize(h->stream));
            double scale = 1.0 / sqrt(s / (double)h->n3);
            launch_scale(h->stream, h->n3, h->work, scale);
            store_from_work(h, h->phi[h->cur], h->cphi[h->cur]);
            AXG_HIP(hipMemsetAsync(h->work, 0, (size_t)h->n3 * 16, h->stream));
            store_from_work(h, h->pi, h->cpi);
            AXG_HIP(hipStreamSynchronize(h->stream));
        } catch (...) {
            (void)hipFree(pidx);
            (void)hipFree(pval);
            throw;
        }
        (void)hipFree(pidx);
        (void)hipFree(pval);
        h->have_flags = false;
        h->have_plaq = false;
        h->have_spec = false;
        h->nstr_rec = -1;
        h->nwall_rec = -1;
        h->have_grp[0] = h->have_grp[1] = false;
    });
}
