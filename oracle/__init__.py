"""CPU FP64 oracle for the axion_grav lattice step and diagnostics (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package.  It is the checker, never the thing measured or shipped: the product in
``axion_grav_amd/`` does not import it.

The arithmetic lives in ``oracle.c`` (a restatement of the reference Fortran, each function
citing the reference file:line it follows); this module is a numpy/ctypes veneer over it.

Complex fields are numpy ``complex128`` arrays of shape ``(N, N, N)`` indexed ``[z, y, x]``
(C order, x fastest), which is byte-identical to the reference's Fortran ``data1(x, y, z)``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PI_T = 3.14159265359  # potential.f90:8


_LIB_PATH = None  # another build of oracle.c to load instead of liboracle.so (use_library)


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def use_library(path):
    """Load the oracle from `path` (a build of the same oracle.c, e.g. for the host's own CPU)
    instead of liboracle.so; before the first call only."""
    global _LIB_PATH
    if _LIB is not None:
        raise RuntimeError("the oracle library is already loaded")
    _LIB_PATH = path


def lib():
    global _LIB
    if _LIB is None:
        path = _LIB_PATH or os.path.join(_HERE, "liboracle.so")
        if _LIB_PATH is None and not os.path.exists(path):
            build()
        L = C.CDLL(path)
        d, i64, i32, p = C.c_double, C.c_int64, C.c_int32, C.c_void_p
        L.orc_msquared.restype = d
        L.orc_msquared.argtypes = [d]
        L.orc_maxion2.restype = d
        L.orc_maxion2.argtypes = [d, d]
        L.orc_derive.argtypes = [d, d, d, d, d, i64, p]
        L.orc_get_dtau.argtypes = [d, d, d, d, i64, d, p, p]
        L.orc_drift.argtypes = [p, p, i64, d]
        L.orc_kick.argtypes = [p, p, i64, d, d, d, d, d, C.c_int, d]
        L.orc_fft3d.argtypes = [p, p, i64, C.c_int]
        L.orc_gradient.argtypes = [p, p, i64, d]
        L.orc_gw_integrate.argtypes = [p, p, i64, d, d, d, d, d, d, C.c_int]
        L.orc_symplectic.argtypes = [p, p, p, i64, d, d, i32, d, d, d, d, d, C.c_int, C.c_int, d, C.c_int]
        L.orc_gw_accumulate.argtypes = [p, p, i64, d, d, d, d, d, C.c_int]
        L.orc_degrade_gradient.argtypes = [p, i64, i64, d, p]
        L.orc_degrade_field.argtypes = [p, i64, i64, p]
        L.orc_degrade_mask.argtypes = [p, i64, i64, p]
        L.orc_gw_spectrum.argtypes = [p, i64, C.c_int, d, d, d, d, d, p, p, p, p]
        L.orc_classify.argtypes = [p, i64, C.c_int, C.c_int, C.c_int, p, p]
        L.orc_plaquettes.argtypes = [p, i64, p, p]
        u32, u64 = C.c_uint32, C.c_uint64
        L.orc_mcg59_seed.restype = u64
        L.orc_mcg59_seed.argtypes = [u32, u32]
        L.orc_mcg59_jump.restype = u64
        L.orc_mcg59_jump.argtypes = [u64, u64]
        L.orc_mcg59_uniforms.argtypes = [u32, u32, u64, i64, p]
        L.orc_mcg59_gauss.argtypes = [u32, u32, u64, i64, p]
        L.orc_strings_mcg.restype = i64
        L.orc_strings_mcg.argtypes = [p, p, i64, p, u64, u64, d, C.c_int, p, p]
        L.orc_velocity_string_yy.argtypes = [p, p, i64, i64, i64, i64, d, p]
        L.orc_ic_thermal_mcg.argtypes = [p, p, i64, u64, u64, d, d, d, d]
        L.orc_field_covariance_mcg.restype = C.c_int
        L.orc_field_covariance_mcg.argtypes = [p, i64, C.c_int, d, d, p, p, u64, u64, C.c_int, C.c_int, d, p, p]
        L.orc_conv_spectrum.argtypes = [p, p, i64, C.c_int, d, d, d, d, p, p]
        L.orc_integrate_p_check.restype = d
        L.orc_integrate_p_check.argtypes = [C.c_int, d, d, d, d, d]
        L.orc_init_spectrum.argtypes = [C.c_int, d, C.c_int, p, p, p, p]
        L.orc_pp_inverse.restype = C.c_int
        L.orc_pp_inverse.argtypes = [C.c_int, p]
        L.orc_matrix_deconv.restype = C.c_int
        L.orc_matrix_deconv.argtypes = [p, i64, C.c_int, d, d, d, p, p, p, C.c_int, p, p]
        L.orc_energy.argtypes = [p, p, i64, d, d, d, d, C.c_int, d, p]
        L.orc_philox.argtypes = [p, C.c_uint32, C.c_uint32]
        L.orc_ic_mode.argtypes = [C.c_uint64, i64, i64, i64, i64, p, p]
        L.orc_ic_flat.argtypes = [p, p, i64, C.c_uint64, d, d]
        L.orc_chi2_string.restype = d
        L.orc_chi2_string.argtypes = [p, p]
        L.orc_string_length.restype = d
        L.orc_string_length.argtypes = [p, p]
        L.orc_position_string.argtypes = [p, C.c_uint32, C.c_uint32, i64, p, p, p, p]
        L.orc_velocity_string.argtypes = [p, p, i64, i64, i64, i64, p, d, p]
        L.orc_strings.restype = i64
        L.orc_strings.argtypes = [p, p, i64, p, C.c_uint32, C.c_uint32, d, C.c_int, p, p]
        L.orc_thermal_factor.restype = d
        L.orc_thermal_factor.argtypes = [C.c_int, i64, d, d, d, d]
        L.orc_thermal_gauss.argtypes = [C.c_uint64, C.c_int, C.c_uint64, p, p]
        L.orc_ic_thermal.argtypes = [p, p, i64, C.c_uint64, d, d, d, d]
        L.orc_lat_dist.restype = d
        L.orc_lat_dist.argtypes = [p, p, i64]
        L.orc_group_points.argtypes = [p, C.c_int, i64, i64, d, p]
        L.orc_group_sums.restype = i64
        L.orc_group_sums.argtypes = [p, p, i64, d, d, p, p]
        L.orc_cov_gauss.argtypes = [C.c_uint64, C.c_int, C.c_uint64, p, p]
        L.orc_field_covariance.restype = C.c_int
        L.orc_field_covariance.argtypes = [p, i64, C.c_int, d, d, p, p, C.c_uint64, C.c_int, C.c_int, d, p, p]
        L.orc_position_wall.restype = C.c_int
        L.orc_position_wall.argtypes = [p, p, p, p, p]
        L.orc_walls.restype = i64
        L.orc_walls.argtypes = [p, i64, p, p, p]
        _LIB = L
    return _LIB


def _p(a):
    if not (a.flags["C_CONTIGUOUS"] and a.dtype in (np.complex128, np.float64, np.int8, np.int64, np.uint32)):
        raise ValueError("oracle arrays must be C-contiguous complex128/float64/int8/int64/uint32")
    if a.dtype == np.complex128 and a.ndim >= 3 and len(set(a.shape[-3:])) != 1:
        raise ValueError(f"lattice arrays must be (..., N, N, N), got {a.shape}")
    return a.ctypes.data_as(C.c_void_p)


def _field(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


# ------------------------------------------------------------------ settings / potential

def msquared(thrm):
    return lib().orc_msquared(thrm)


def maxion2(thrm, kappa):
    return lib().orc_maxion2(thrm, kappa)


DERIVED_KEYS = ("a_now", "tau_cr", "sigma_cr", "zeta", "width_cr", "thickness_cr",
                "dspa", "dwav", "dv_spa", "dv_wav")


def derive(sigma, g_rel, kappa, tau_fin, boxsize, ngrid):
    out = np.zeros(10)
    lib().orc_derive(sigma, g_rel, kappa, tau_fin, boxsize, int(ngrid), _p(out))
    return dict(zip(DERIVED_KEYS, out.tolist()))


def get_dtau(tau1, tau2, sigma_cr, dwav, ngrid, eps_dtau):
    dt = C.c_double()
    n = C.c_int32()
    lib().orc_get_dtau(tau1, tau2, sigma_cr, dwav, int(ngrid), eps_dtau, C.byref(dt), C.byref(n))
    return dt.value, n.value


# ------------------------------------------------------------------ integrator

def drift(phi, pi, dt):
    lib().orc_drift(_p(phi), _p(pi), phi.size, dt)


def kick(phi, pi, t, dt, dspa, sigma_cr, kappa, initial_loop, kappa0=0.0):
    lib().orc_kick(_p(phi), _p(pi), phi.shape[0], t, dt, dspa, sigma_cr, kappa, int(initial_loop), kappa0)


def fft3d(a, sign=+1):
    a = _field(a)
    out = np.empty_like(a)
    lib().orc_fft3d(_p(a), _p(out), a.shape[0], int(sign))
    return out


def gradient(phi, dspa):
    N = phi.shape[0]
    g = np.empty((3, N, N, N), np.complex128)
    lib().orc_gradient(_p(phi), _p(g), N, dspa)
    return g


def gw_integrate(phi, A, tau, dtau, dspa, dwav, dv_spa, sigma_cr, fix_swap=False):
    lib().orc_gw_integrate(_p(phi), _p(A), phi.shape[0], tau, dtau, dspa, dwav, dv_spa, sigma_cr,
                           int(fix_swap))


def symplectic(phi, pi, A, t1, dt, nsteps, P, fix_swap=False):
    """Advance (phi, pi[, A]) in place by nsteps DKD steps of dt from t1.  P is a dict with
    dspa, dwav, dv_spa, sigma_cr, kappa, initial_loop and optionally potential_kappa0 and
    degrade_spec (then A lives on the grid N / 2^degrade_spec)."""
    lib().orc_symplectic(_p(phi), _p(pi), _p(A) if A is not None else None, phi.shape[0], t1, dt,
                         int(nsteps), P["dspa"], P["dwav"], P["dv_spa"], P["sigma_cr"], P["kappa"],
                         int(P["initial_loop"]), int(fix_swap), float(P.get("potential_kappa0", 0.0)),
                         int(P.get("degrade_spec", 0)))


def new_gw_state(N):
    return np.zeros((4, N, N, N), np.complex128)


def gw_accumulate(dphi, A, tau, dtau, dwav, dv_spa, sigma_cr, fix_swap=False):
    """grav_wave_integration's accumulation from arr_dphi (3, N, N, N) into A (4, N, N, N)."""
    dphi = np.ascontiguousarray(dphi, np.complex128)
    lib().orc_gw_accumulate(_p(dphi), _p(A), dphi.shape[1], tau, dtau, dwav, dv_spa, sigma_cr, int(fix_swap))


def reduced_units(P, degrade_spec):
    """The GW / spectrum grid of degrade_spec > 0 as init_grav and init_spectrum derive it
    (grav_waves.f90:37-45, spectrum.f90:42-49): ngrid, dspa, dwav, dv_spa, dv_wav."""
    deg = 2 ** degrade_spec
    twopi_t = 2 * 3.14159265359
    n = int(P["ngrid"] if "ngrid" in P else P["N"]) // deg
    dspa = P["dspa"] * deg
    dwav = twopi_t / dspa / n
    r = dwav / twopi_t
    return dict(ngrid=n, dspa=dspa, dwav=dwav, dv_spa=dspa * dspa * dspa, dv_wav=r * r * r)


def degrade_gradient(phi, degrade_spec, dspa):
    """data1_to_arr_dphi_degrade: (3, Nc, Nc, Nc) block-summed central differences."""
    phi = np.ascontiguousarray(phi, np.complex128)
    N, deg = phi.shape[0], 2 ** degrade_spec
    g = np.empty((3, N // deg, N // deg, N // deg), np.complex128)
    lib().orc_degrade_gradient(_p(phi), N, deg, dspa, _p(g))
    return g


def degrade_field(f, degrade_spec):
    """data2_to_data_spec_degrade: the block means (Nc, Nc, Nc)."""
    f = np.ascontiguousarray(f, np.complex128)
    N, deg = f.shape[0], 2 ** degrade_spec
    out = np.empty((N // deg,) * 3, np.complex128)
    lib().orc_degrade_field(_p(f), N, deg, _p(out))
    return out


def degrade_mask(flags, degrade_spec):
    """masking with degrade: bit 4 on the coarse cells that hold a masked lattice cell."""
    flags = np.ascontiguousarray(flags, np.int8)
    N, deg = flags.shape[0], 2 ** degrade_spec
    out = np.empty((N // deg,) * 3, np.int8)
    lib().orc_degrade_mask(_p(flags), N, deg, _p(out))
    return out


# ------------------------------------------------------------------ diagnostics

def gw_spectrum(A, nb, dspa, dwav, tau_cr, a_now, dv_wav):
    N = A.shape[1]
    numer, denom, qmid, power = (np.zeros(nb) for _ in range(4))
    lib().orc_gw_spectrum(_p(A), N, nb, dspa, dwav, tau_cr, a_now, dv_wav, _p(numer), _p(denom),
                          _p(qmid), _p(power))
    return dict(numer=numer, denom=denom, qmid=qmid, power=power)


def classify(phi, id_strings=True, id_walls=True, spec_est=True):
    N = phi.shape[0]
    flags = np.zeros((N, N, N), np.int8)
    counts = np.zeros(3, np.int64)
    lib().orc_classify(_p(phi), N, int(id_strings), int(id_walls), int(spec_est), _p(flags), _p(counts))
    return flags, counts


def plaquettes(phi):
    """Plaquette windings: (code (N, N, N) uint16, counts dict of (3,) arrays n_pierced, n_plus,
    n_minus, n_quadrant, net, records (n, 5) int32 (ix, iy, iz, o, n) in site then face order)."""
    N = phi.shape[0]
    code = np.zeros((N, N, N), np.uint16)
    c = np.zeros(16, np.int64)
    lib().orc_plaquettes(_p(np.ascontiguousarray(phi, np.complex128)), N, code.ctypes.data_as(C.c_void_p), _p(c))
    counts = {k: c[j::5][:3].copy() for j, k in enumerate(("n_pierced", "n_plus", "n_minus", "n_quadrant", "net"))}
    n = np.stack([((code.astype(np.int32) >> (3 * o)) & 7) - 2 for o in range(3)], axis=-1)  # (z, y, x, o)
    zz, yy, xx, oo = np.nonzero(n)
    rec = np.stack([xx + 1, yy + 1, zz + 1, oo + 1, n[zz, yy, xx, oo]], axis=1).astype(np.int32)
    return code, counts, rec


def conv_spectrum(pi, flags, nb, dspa, dwav, dv_spa, dv_wav):
    N = pi.shape[0]
    pw, num = np.zeros(nb), np.zeros(nb)
    lib().orc_conv_spectrum(_p(pi), _p(flags) if flags is not None else None, N, nb, dspa, dwav,
                            dv_spa, dv_wav, _p(pw), _p(num))
    return pw, num


def integrate_p_check(nint, li, si, lj, sj, q):
    return lib().orc_integrate_p_check(nint, li, si, lj, sj, q)


def init_spectrum(nb, dspa, num_largeq=500):
    ncov = nb * (nb + 1) // 2
    wav = np.zeros(nb + 1)
    largeq = np.zeros((ncov, num_largeq + 1))
    mat = np.zeros((ncov, num_largeq + 1))
    dqs = np.zeros(ncov)
    lib().orc_init_spectrum(nb, dspa, num_largeq, _p(wav), _p(largeq), _p(mat), _p(dqs))
    return dict(wav=wav, largeq=largeq, mat=mat, dqs=dqs, num_largeq=num_largeq)


def pp_inverse(nb, ap):
    ap = np.ascontiguousarray(ap, dtype=np.float64).copy()
    info = lib().orc_pp_inverse(nb, _p(ap))
    return ap, info


def matrix_deconv(flags, nb, dwav, dv_spa, dv_wav, spec_tab):
    N = flags.shape[0]
    ncov = nb * (nb + 1) // 2
    mat, inv = np.zeros(ncov), np.zeros(ncov)
    info = lib().orc_matrix_deconv(_p(flags), N, nb, dwav, dv_spa, dv_wav, _p(spec_tab["largeq"]),
                                   _p(spec_tab["mat"]), _p(spec_tab["dqs"]), spec_tab["num_largeq"],
                                   _p(mat), _p(inv))
    return mat, inv, info


def energy(phi, pi, t, dspa, sigma_cr, kappa, initial_loop, kappa0=0.0):
    out = np.zeros(4)
    lib().orc_energy(_p(phi), _p(pi), phi.shape[0], t, dspa, sigma_cr, kappa, int(initial_loop), kappa0, _p(out))
    return out


def philox(ctr, k0, k1):
    c = np.array(ctr, np.uint32)
    lib().orc_philox(_p(c), k0, k1)
    return c


def ic_mode(seed, N, ix, iy, iz):
    re, im = C.c_double(), C.c_double()
    lib().orc_ic_mode(seed, N, ix, iy, iz, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def ic_flat(N, seed, kmax, dwav):
    phi = np.zeros((N, N, N), np.complex128)
    pi = np.zeros((N, N, N), np.complex128)
    lib().orc_ic_flat(_p(phi), _p(pi), N, seed, kmax, dwav)
    return phi, pi


def sym_from_packed(ap, nb):
    """Full symmetric matrix from LAPACK packed-upper storage (ij = j + i(i-1)/2, j <= i)."""
    M = np.zeros((nb, nb))
    for i in range(1, nb + 1):
        for j in range(1, i + 1):
            v = ap[i * (i - 1) // 2 + j - 1]
            M[j - 1, i - 1] = v
            M[i - 1, j - 1] = v
    return M


# ------------------------------------------------------------------ strings

def chi2_string(ph, prm):
    return lib().orc_chi2_string(_p(np.ascontiguousarray(ph, np.float64)), _p(np.ascontiguousarray(prm, np.float64)))


def string_length(pos, dirv):
    return lib().orc_string_length(_p(np.ascontiguousarray(pos, np.float64)), _p(np.ascontiguousarray(dirv, np.float64)))


def position_string(ph, key0, key1, s):
    """Fit of one string cell: (pos relative to the cell centre, dir, gof, sl)."""
    pos, dirv, g, sl = np.zeros(3), np.zeros(3), np.zeros(1), np.zeros(1)
    lib().orc_position_string(_p(np.ascontiguousarray(ph, np.float64)), key0, key1, s, _p(pos), _p(dirv), _p(g), _p(sl))
    return pos, dirv, float(g[0]), float(sl[0])


def velocity_string(phi, pi, x, y, z, pos, dspa):
    vel = np.zeros(3)
    lib().orc_velocity_string(_p(phi), _p(pi), phi.shape[0], x, y, z, _p(np.ascontiguousarray(pos, np.float64)),
                              dspa, _p(vel))
    return vel


def velocity_string_yy(phi, pi, x, y, z, dspa):
    """velocity_string_YY (the VEL = 1 build) of the cell with corner (x, y, z), 0-based."""
    vel = np.zeros(3)
    lib().orc_velocity_string_yy(_p(phi), _p(pi), phi.shape[0], x, y, z, dspa, _p(vel))
    return vel


def strings(phi, pi, flags, key0, key1, dspa, stream=None, vel_yy=False):
    """All string cells (flag bit 1) in site order: records (ns, 11) = pos(3, site coordinates
    counted from 0), dir(3), gof, sl, vel(3); sums = (sum sl, n |v|<1, sum |v|, sum v^2, sum gamma).
    stream = (w0, w1, pos): starting points from the reference's MCG59 stream (keys unused).
    vel_yy: velocities from velocity_string_YY (the VEL = 1 build)."""
    N = phi.shape[0]
    ns = int(np.count_nonzero(flags & 1))
    rec = np.zeros((max(ns, 1), 11))
    sums = np.zeros(5)
    fl = _p(np.ascontiguousarray(flags, np.int8))
    if stream is None:
        n = lib().orc_strings(_p(phi), _p(pi), N, fl, key0, key1, dspa, int(vel_yy), _p(rec), _p(sums))
    else:
        n = lib().orc_strings_mcg(_p(phi), _p(pi), N, fl, mcg59_seed(stream[0], stream[1]), stream[2], dspa,
                                  int(vel_yy), _p(rec), _p(sums))
    return rec[:n], sums


# ------------------------------------------------------------------ thermal initial state

def thermal_factor(ip, n2, tau_ini, sigma_cr, dwav, dv_wav):
    return lib().orc_thermal_factor(ip, n2, tau_ini, sigma_cr, dwav, dv_wav)


def thermal_gauss(seed, ip, lin):
    re, im = C.c_double(), C.c_double()
    lib().orc_thermal_gauss(seed, ip, lin, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def ic_thermal(N, seed, tau_ini, sigma_cr, dwav, dv_wav, stream=None):
    """Thermal state; stream = (w0, w1, pos): the reference's MCG59 stream seeded by the two words
    and positioned at pos (seed is then unused)."""
    phi = np.zeros((N, N, N), np.complex128)
    pi = np.zeros((N, N, N), np.complex128)
    if stream is None:
        lib().orc_ic_thermal(_p(phi), _p(pi), N, seed, tau_ini, sigma_cr, dwav, dv_wav)
    else:
        x0 = mcg59_seed(stream[0], stream[1])
        lib().orc_ic_thermal_mcg(_p(phi), _p(pi), N, x0, stream[2], tau_ini, sigma_cr, dwav, dv_wav)
    return phi, pi


# ------------------------------------------------------------------ the reference's random stream

def mcg59_seed(w0, w1):
    return lib().orc_mcg59_seed(w0, w1)


def mcg59_jump(x0, p):
    return lib().orc_mcg59_jump(x0, p)


def mcg59_uniforms(w0, w1, pos, n):
    out = np.zeros(n)
    lib().orc_mcg59_uniforms(w0, w1, pos, n, _p(out))
    return out


def mcg59_gauss(w0, w1, pos, n):
    out = np.zeros(n)
    lib().orc_mcg59_gauss(w0, w1, pos, n, _p(out))
    return out


# ------------------------------------------------------------------ spectrum covariance

def cov_gauss(seed, it, lin):
    re, im = C.c_double(), C.c_double()
    lib().orc_cov_gauss(seed, it, lin, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def field_covariance(flags, nb, dspa, dwav, power, matn, seed, min_mc, max_mc, tol, want_pows=False, stream=None):
    """Monte Carlo covariance of the deconvolved spectrum: (packed cov, iteration counter
    (negative without convergence), per-realisation pow (max_mc, nb) or None).  stream =
    (w0, w1, pos): realisations from the reference's MCG59 stream (seed unused)."""
    N = flags.shape[0]
    ncov = nb * (nb + 1) // 2
    cov = np.zeros(ncov)
    pows = np.zeros((max_mc, nb)) if want_pows else None
    args = (_p(np.ascontiguousarray(flags, np.int8)), N, nb, dspa, dwav, _p(np.ascontiguousarray(power, np.float64)),
            _p(np.ascontiguousarray(matn, np.float64)))
    tail = (min_mc, max_mc, tol, _p(cov), _p(pows) if want_pows else None)
    if stream is None:
        it = lib().orc_field_covariance(*args, seed, *tail)
    else:
        it = lib().orc_field_covariance_mcg(*args, mcg59_seed(stream[0], stream[1]), stream[2], *tail)
    return cov, it, pows


# ------------------------------------------------------------------ string / wall networks

def lat_dist(p1, p2, N):
    return lib().orc_lat_dist(_p(np.ascontiguousarray(p1, np.float64)), _p(np.ascontiguousarray(p2, np.float64)), N)


def group_points(rec, N, th):
    """Labels (smallest record index of each record's group) of records (n, >= 3) by position."""
    rec = np.ascontiguousarray(rec, np.float64)
    n = rec.shape[0]
    lab = np.zeros(max(n, 1), np.int32)
    lib().orc_group_points(_p(rec), rec.shape[1] if n else 3, n, N, th, lab.ctypes.data_as(C.c_void_p))
    return lab[:n]


def group_sums(label, val, t2, t3):
    """(number of groups, group number per record (1-based), [total, above t2, above t3])."""
    label = np.ascontiguousarray(label, np.int32)
    n = label.shape[0]
    ids = np.zeros(max(n, 1), np.int32)
    out = np.zeros(3)
    ng = lib().orc_group_sums(label.ctypes.data_as(C.c_void_p), _p(np.ascontiguousarray(val, np.float64)), n, t2, t3,
                              ids.ctypes.data_as(C.c_void_p), _p(out))
    return ng, ids[:n], out


def group_threshold(width, hor_size):
    """grouping_strings / grouping_walls threshold (defects_3d.f90:1245-1246, 1745-1746)."""
    return max(2.0, (width ** 2.0 * hor_size) ** (1 / (2.0 + 1)))


# ------------------------------------------------------------------ walls

def position_wall(ph):
    """Fit of one wall cell: (m edge points, pos relative to the cell centre, dir, gof, wa)."""
    pos, dirv, g, wa = np.zeros(3), np.zeros(3), np.zeros(1), np.zeros(1)
    m = lib().orc_position_wall(_p(np.ascontiguousarray(ph, np.float64)), _p(pos), _p(dirv), _p(g), _p(wa))
    return m, pos, dirv, float(g[0]), float(wa[0])


def walls(phi, flags):
    """All wall cells (flag bit 2) in site order: records (nw, 8) = pos(3, site coordinates
    counted from 0), dir(3), gof, wa; sums = (sum wa, cells with < 3 edge points)."""
    nw = int(np.count_nonzero(flags & 2))
    rec = np.zeros((max(nw, 1), 8))
    sums = np.zeros(2)
    n = lib().orc_walls(_p(phi), phi.shape[0], _p(np.ascontiguousarray(flags, np.int8)), _p(rec), _p(sums))
    return rec[:n], sums
