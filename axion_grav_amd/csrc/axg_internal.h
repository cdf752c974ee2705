// Host-side internals shared by the translation units of libaxgrav: the handle, the
// kernel launchers and the small host math (get_dtau, potential scalars, bin tables).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/axgrav.h"
#include "axg_device.h"

// potential.f90:8 -- the truncated pi used by the reference for every physical constant,
// bin edge and defect threshold.
constexpr double AXG_PI_T = 3.14159265359;
constexpr double AXG_TWOPI_T = 2 * AXG_PI_T;
constexpr double AXG_THRM_SIGMA = 2.449489742783178;  // potential.f90:30

#define AXG_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) throw axg::Error(AXG_EDEVICE, std::string(#call) + ": " + \
                                               hipGetErrorString(e_));                  \
    } while (0)

namespace axg {

struct Error {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// Scalars of one kick (evolve.f90:291-305).
struct KickScalars {
    double f1, f2, f3, m2, ma2, kt;  // kt: N_DW=2 tilt coefficient kappa0*t^4 (0 = off)
};
KickScalars kick_scalars(const axg_params& p, double t, double dt);
void set_handleless_error(const std::string& msg);
// One step's scalars in device memory, for steps replayed from a captured graph: the step
// kernels take them from here instead of their by-value arguments when the pointer is set.
struct StepEntry {
    KickScalars k;
    double dtd;       // the drift after the kick
    double tau, dw;   // the GW update's time and weight
};
// Where a table-driven step finds its scalars: tab[*base + sidx] (base may be null: 0).  The
// step kernels of a captured graph take a fixed sidx and a device counter `base` that the graph
// advances once per replay, so no per-step launch selects the entry.
struct StepRef {
    const StepEntry* tab = nullptr;
    const int* base = nullptr;
    int sidx = 0;
};
__device__ __forceinline__ int step_index(const StepRef& r) { return (r.base ? *r.base : 0) + r.sidx; }
__device__ __forceinline__ const StepEntry* step_entry(const StepRef& r)
{
    return r.tab ? r.tab + step_index(r) : nullptr;
}
// One lattice of a stacked launch: lattices of one size and precision advanced by one launch per
// pass, the lattice index in blockIdx.z.  phi[0] / cphi[0] hold the fields at step 0 of the
// table, so step s reads phi[s & 1] and writes phi[(s & 1) ^ 1].
// The buffers are global-memory pointers by type, so that the stacked kernels, which read them
// from this table, address HBM with global instructions rather than flat ones (a flat access also
// counts against the LDS wait counter, so each barrier of a transform would wait for it).
using gptr = __attribute__((address_space(1))) void*;
struct LatRef {
    gptr phi[2];
    gptr cphi[2];
    gptr pi;
    gptr cpi;
    gptr W;
    gptr W2;
    gptr A;
    const StepEntry* tab;
};
void launch_add_int(hipStream_t s, int* counter, int delta);
double tilt_coef(const axg_params& p, double t);
double maxion2(double thrm, double kappa);
double msquared(double thrm);

FftPlan make_plan(int n);
bool supported_n(int64_t n);

// Column-tile width of the strided FFT passes: 128 bytes of complex per row.
template <typename T> constexpr int col_tile() { return 128 / (int)sizeof(typename V2<T>::type); }

// Row pitch (in complex elements) of the half-spectrum arrays: N/2+1 rounded up to 16.
inline int half_pitch(int n) { return ((n / 2 + 1) + 15) / 16 * 16; }

// ------------------------------------------------------------------ kernel launchers
// T = double (FP64) or float (FP32 storage); COMP = FP32 with int16 compensation.

template <typename T, bool COMP>
void launch_drift(hipStream_t s, int N, const void* phi_in, void* phi_out, const void* pi,
                  const void* cphi_in, void* cphi_out, const void* cpi, double dt);

template <typename T, bool COMP>
void launch_kick_drift(hipStream_t s, int N, const void* phi_in, void* phi_out, void* pi,
                       const void* cphi_in, void* cphi_out, void* cpi, void* src,
                       const KickScalars& k, double dspa, double dt_drift, bool gw,
                       const StepRef& sr = StepRef{});

// Fused kick + source + x transform + drift for p2_size(N); false (nothing launched) otherwise.
template <typename T, bool COMP>
bool launch_kick_source_xfft(hipStream_t s, int N, const void* phi_in, void* phi_out, void* pi, const void* cphi_in,
                             void* cphi_out, void* cpi, void* W, int pitch, const void* tw, const KickScalars& k,
                             double dspa, double dt_drift, int variant, const StepRef& sr = StepRef{},
                             const LatRef* lat = nullptr, int nlat = 1);

template <typename T>
void launch_gw_xpass(hipStream_t s, int N, int pitch, const void* src, void* W, const FftPlan& plan,
                     const void* tw);

template <typename T, int SIGN>
void launch_fft_cols(hipStream_t s, int N, void* data, int64_t es, int64_t ps, int nplanes, int ncol,
                     const FftPlan& plan, const void* tw);

template <typename T, int SIGN>
void launch_fft_rows(hipStream_t s, int N, void* data, int64_t nrows, const FftPlan& plan, const void* tw);

// Layout of the half-spectrum arrays read by the accumulation pass (the y-transformed sources
// and the accumulators).  zt == 0: rows [c][kz][ky][kx] with pitch P.  zt > 0 (p2_size(N)):
// tiles [c][ky][t][kz][zt] with t = kx / zt, so one accumulation block (one ky, one tile)
// reads and writes contiguous N*zt runs whatever the tile width.
struct GwLayout {
    int N, P, zt, ntile;
    __host__ __device__ int64_t comp() const
    {
        return zt ? (int64_t)N * ntile * N * zt : (int64_t)N * N * P;
    }
    __host__ __device__ int64_t idx(int c, int kz, int ky, int kx) const
    {
        if (zt == 0) return (int64_t)c * N * N * P + ((int64_t)kz * N + ky) * P + kx;
        return (((int64_t)c * N + ky) * ntile + kx / zt) * N * zt + (int64_t)kz * zt + (kx % zt);
    }
};
GwLayout gw_layout(int N, bool fp32);

// half_pitch(N) as a compile-time constant for the compile-time (p2_size) kernels.
template <int N> constexpr int half_pitch_c() { return ((N / 2 + 1) + 15) / 16 * 16; }

// The tiled GwLayout with compile-time N and tile width ZT (a power of two), so index
// arithmetic in the compile-time kernels reduces to shifts, masks and constant divisions.
// Launchers check that the handle's runtime layout is this one.
template <int N, int ZT> struct TiledLayout {
    static constexpr int ZTW = ZT;
    static constexpr int NT = (N / 2 + 1 + ZT - 1) / ZT;
    static constexpr int64_t COMP = (int64_t)N * NT * N * ZT;
    __host__ __device__ static int64_t idx(int c, int kz, int ky, int kx)
    {
        return (((int64_t)c * N + ky) * NT + kx / ZT) * (int64_t)(N * ZT) + (int64_t)kz * ZT + (kx % ZT);
    }
    static bool matches(const GwLayout& L) { return L.N == N && L.zt == ZT && L.ntile == NT; }
};

struct GwStepScalars {
    double tau, dw, dwav;
    int fix_swap;
};
// One kick's scalars for the adjoint of its GW update (adjoint.hip); last = the run's final kick.
struct AdjStepScalars {
    double tau, dw, dwav;
    int fix_swap, last;
};
constexpr int ADJ_GRID_H = 1024;  // blocks of the adjoint's reductions
void launch_adj_seed(hipStream_t s, int N, double2* abar, const uint16_t* bin_of_n2, const double* w);
// W0..W2: the three packed adjoint-source arrays, N^3 complex in the precision TW
template <typename TW>
void launch_adj_modes(hipStream_t s, int N, const double2* abar, const AdjStepScalars& g, void* W0, void* W1, void* W2);
template <typename T, bool COMP, typename TW>
void launch_adj_gbar(hipStream_t s, int N, const void* phi, const void* cphi, double inv_d, void* W0, void* W1,
                     void* W2);
// The same with the x transform of W0..W2 (SIGN -1, twiddle table tw of N entries in TW) done
// first in the kernel, for the sizes of the compile-time FFT engine; false (nothing launched)
// for the others.
template <typename T, bool COMP, typename TW>
bool launch_adj_gbar_xfft(hipStream_t s, int N, const void* phi, const void* cphi, double inv_d, void* W0, void* W1,
                          void* W2, const void* tw);
// phibar += adjoint of the kick and of the source's differences; pibar_out = f1 pibar + hd phibar;
// partial[block][2] += (s0, s1) times the block's sums of <pibar, V'_0(phi)> and <pibar, (-Re, Im)>.
template <typename T, bool COMP, typename TW>
void launch_adj_kick(hipStream_t s, int N, const void* phi, const void* cphi, const double2* pibar, double2* phibar,
                     double2* pibar_out, const void* G0, const void* G1, const void* G2, const KickScalars& k,
                     double inv_d, double hd, double s0, double s1, bool use_g, double* partial);
template <typename T, bool COMP>
void launch_adj_dot(hipStream_t s, int64_t n3, const void* phi, const void* cphi, const void* pi, const void* cpi,
                    const double2* phibar, const double2* pibar, double* partial, double* out);
// y transform of W ([c][z][y][kx], pitch P): in place for the row layout, or out of place
// into W2 in the tiled layout.
template <typename T>
void launch_gw_ypass(hipStream_t s, int N, int pitch, void* W, void* W2, const GwLayout& L, const FftPlan& plan,
                     const void* tw, int variant = 0, const LatRef* lat = nullptr, int nlat = 1);

template <typename T>
void launch_gw_zpass_acc(hipStream_t s, int N, int pitch, const void* W, void* A, const GwLayout& L,
                         const FftPlan& plan, const void* tw, const GwStepScalars& g,
                         const StepRef& sr = StepRef{}, const LatRef* lat = nullptr, int nlat = 1);

template <typename T>
void launch_gw_bin(hipStream_t s, int N, const GwLayout& L, const void* A, const uint16_t* bin_of_n2, int64_t n2max, int nb,
                   double* partial, int nwg, double* out);

// one accumulator component between the device layout and the compact [kz][ky][kx<=N/2]
// FP64 form (dir > 0: device -> compact)
template <typename T>
void launch_gw_pack(hipStream_t s, int N, const GwLayout& L, void* A, int m, double2* compact, int dir);

// diagnostics
template <typename T, bool COMP>
void launch_energy(hipStream_t s, int N, const void* phi, const void* cphi, const void* pi,
                   const void* cpi, double m2, double ma2, double kt, double* partial, int nwg, double* out);

template <typename T, bool COMP>
void launch_classify(hipStream_t s, int N, const void* phi, const void* cphi, int id_strings, int id_walls,
                     int spec_est, int8_t* flags, int64_t* block_counts, int64_t nblk);

void launch_cell_compact(hipStream_t s, int N, const int8_t* flags, int bit, const int64_t* block_offsets,
                         int32_t* out, int triplets, int64_t nblk);
void sum_columns(hipStream_t s, const double* partial, int nwg, int width, double* out);
// Wall cells (flag bit 2) in site order: plane fit and area per cell.  rec: 8 doubles per
// wall (pos(3) in site coordinates counted from 0, dir(3), gof, wa); partial: 2 doubles per
// block of WALL_NT walls (sum wa, cells with fewer than 3 edge points).
constexpr int WALL_NT = 64;
template <typename T, bool COMP>
void launch_wall_fit(hipStream_t s, int N, const void* phi, const void* cphi, const int32_t* cells, int64_t nw,
                     double* rec, double* partial);
// Grouping of defect points (groups.hip): clears the cell map, scatters the points into it
// (points more than one cell from their cell centre go to the outlier list instead) and
// returns the largest per-axis offset of a mapped point; then links the points within th and
// writes each point's label (the smallest index of its group).  scalars: 2 device words
// (offset bits; outlier count and first mapped point); the failure flag is separate.  mode 0:
// cell box of half-width R; 1: all mapped points in one group (centre = the first mapped
// point); 2: every pair tested.
double group_prepare(hipStream_t s, int N, const double* rec, int stride, int col, const int32_t* cells, int64_t n,
                     int32_t* map, int32_t* parent, double* val, unsigned long long* scalars, int32_t* outliers,
                     int* n_out, int* first_mapped);
void group_link(hipStream_t s, int N, const double* rec, int stride, const int32_t* cells, int64_t n,
                const int32_t* map, int32_t* parent, int32_t* label, double th, double maxoff, int R, int mode,
                int centre, const int32_t* outliers, int n_out, int* fail);
void launch_cov_modes(hipStream_t s, int N, const Draws& dr, int iter, const uint16_t* bin_of_n2, const double* sigma,
                      double2* out);
void launch_mask_zero(hipStream_t s, int64_t n3, const int8_t* flags, double2* data);
void launch_thermal_modes(hipStream_t s, int N, const Draws& dr, int ip, double tau_ini, double sigma_cr, double dwav,
                          double dv_wav, double thrm_sigma, double2* out);

// String cells (flag bit 1) in site order `cells`: fit, chord length and velocity per cell.
// rec: 11 doubles per string (pos(3) in site coordinates counted from 0, dir(3), gof, sl, vel(3));
// partial: 5 doubles per block of STRING_NT strings (sum sl, n with |v| < 1, sum |v|, sum v^2,
// sum gamma), to be summed with sum_columns.  Searches still running after `split` simplex
// iterations move to a second kernel (the records do not depend on split); its work space:
// npend (one int), pend (ns int64) and state (STRING_STATE doubles per string).
constexpr int STRING_NT = 64;
constexpr int STRING_STATE = 26;
constexpr int STRING_SPLIT = 32;  // default split (fastest of 0..128 at 512^3 on MI355X)
template <typename T, bool COMP>
void launch_string_fit(hipStream_t s, int N, const void* phi, const void* cphi, const void* pi, const void* cpi,
                       const int32_t* cells, int64_t ns, const Draws& dr, double dspa, int split, int vel_yy,
                       double* rec, double* partial, int* npend, int64_t* pend, double* state);

// Exclusive offsets of column col of per-block counts in[n][ncol] (ncol 3 or PLAQ_COLS) and the
// column totals; part: scan_scratch(n, ncol) values of device scratch.
void launch_scan_counts(hipStream_t s, const int64_t* in, int64_t* out, int64_t n, int64_t* total, int col, int ncol,
                        int64_t* part);
int64_t scan_scratch(int64_t n, int ncol);

// Plaquette windings (plaquettes.hip): per-site face codes (uint16: 3 bits of winding + 2 per face
// o, then the quadrant-test bit of face o at 9 + o), per-block counts [nblk][16] (pierced, +1, -1,
// quadrant, net per face, then the block's pierced total), their exclusive offsets of the
// pierced totals and the 16 column totals.  The compaction writes rows (ix, iy, iz, o, n).
constexpr int PLAQ_BLOCK = 256;
constexpr int PLAQ_COLS = 16;  // per block: [pierced, plus, minus, quadrant, net] x 3, then pierced total
template <typename T, bool COMP>
void launch_plaquettes(hipStream_t s, int N, const void* phi, const void* cphi, uint16_t* code, int64_t* counts,
                       int64_t* offs, int64_t* total, int64_t* part);
void launch_plaquette_compact(hipStream_t s, int N, const uint16_t* code, const int64_t* offs, int32_t* rec);

// The reduced grid of degrade_spec > 0 (degrade.hip), deg = 2^degrade_spec: the block sums of
// the lattice gradient into the packed source planes of the N/deg grid, the block mean of a
// field, and the block's mask bit.
template <typename T, bool COMP>
void launch_degrade_source(hipStream_t s, int N, int deg, const void* phi, const void* cphi, double dspa, double2* src);
template <typename T, bool COMP>
void launch_degrade_field(hipStream_t s, int N, int deg, const void* f, const void* cf, double2* out);
void launch_degrade_mask(hipStream_t s, int N, int deg, const int8_t* flags, int8_t* out);

template <typename T, bool COMP>
void launch_masked_field(hipStream_t s, int N, const void* pi, const void* cpi, const int8_t* flags,
                         double2* out, double fill_unmasked, bool use_pi);

void launch_spec_bin(hipStream_t s, int N, const double2* data, const uint16_t* bin_of_n2, int64_t n2max, int nb,
                     double scale, double* partial, int nwg, double* out);

void launch_n2_hist(hipStream_t s, int N, const double2* data, double* plane_part, double* hist);

void launch_deconv_mat(hipStream_t s, int nb, int num_cov, int nlq, const double* largeq, const double* mat,
                       const double* dqs, const double* hist, const double* cnt, int64_t nn2, double dwav,
                       double* mat0, double* tat0);

template <typename T, bool COMP>
void launch_store_field(hipStream_t s, int N, const double2* in, void* f, void* cf);

template <typename T, bool COMP>
void launch_load_field(hipStream_t s, int N, const void* f, const void* cf, double2* out);

void launch_scatter_modes(hipStream_t s, int N, const int64_t* idx, const double2* val, int64_t nmodes,
                          double2* out);

void launch_sumsq(hipStream_t s, int64_t n, const double2* data, double* partial, int nwg, double* out);
void launch_scale(hipStream_t s, int64_t n, double2* data, double scale);

}  // namespace axg
