"""ctypes binding of libaxgrav (include/axgrav.h) with numpy array conversion.

This is the Python face of the same C-ABI a Fortran host binds through ISO_C_BINDING.  The
library is loaded from ``axion_grav_amd/lib/libaxgrav.so`` (built in-tree by
``python -m axion_grav_amd.build``); there is no fallback: if the library or a GPU is missing,
the calls raise.

Complex fields are numpy ``complex128`` arrays of shape ``(N, N, N)`` indexed ``[z, y, x]``
(x fastest), byte-identical to the reference's Fortran ``data1(x, y, z)``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libaxgrav.so")

FP64, FP32C, FP32 = 0, 1, 2
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ESTATE", -5: "ENUMERIC"}


class Params(C.Structure):
    _fields_ = [
        ("ngrid", C.c_int64),
        ("precision", C.c_int32),
        ("grav_spec", C.c_int32),
        ("initial_loop", C.c_int32),
        ("gw_fix_swap", C.c_int32),
        ("num_bin_grav", C.c_int32),
        ("num_bin", C.c_int32),
        ("id_strings", C.c_int32),
        ("id_walls", C.c_int32),
        ("spec_est", C.c_int32),
        ("device", C.c_int32),
        ("dspa", C.c_double),
        ("dwav", C.c_double),
        ("dv_spa", C.c_double),
        ("dv_wav", C.c_double),
        ("sigma_cr", C.c_double),
        ("tau_cr", C.c_double),
        ("a_now", C.c_double),
        ("kappa", C.c_double),
        ("eps_dtau", C.c_double),
        ("dtau_fixed", C.c_double),
        ("potential_kappa0", C.c_double),
        ("degrade_spec", C.c_int32),
        ("string_vel", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class Keys(C.Structure):
    _fields_ = [("sigma", C.c_double), ("g_rel", C.c_double), ("kappa", C.c_double),
                ("tau_fin", C.c_double), ("boxsize", C.c_double), ("ngrid", C.c_int64)]


class DefectOut(C.Structure):
    _fields_ = [("n_string", C.c_int64), ("n_wall", C.c_int64), ("n_mask", C.c_int64)]


class PlaquetteOut(C.Structure):
    _fields_ = [("n_pierced", C.c_int64 * 3), ("n_plus", C.c_int64 * 3), ("n_minus", C.c_int64 * 3),
                ("n_quadrant", C.c_int64 * 3), ("net", C.c_int64 * 3)]


class WallOut(C.Structure):
    _fields_ = [("n_wall", C.c_int64), ("area", C.c_double)]


class GroupOut(C.Structure):
    _fields_ = [("n_group", C.c_int64), ("th_dist", C.c_double), ("sum", C.c_double * 3)]


class GradOut(C.Structure):
    _fields_ = [("J", C.c_double), ("dJ_dsigma2", C.c_double), ("dJ_dkappa0", C.c_double),
                ("dJ_dlnamp", C.c_double), ("forward_ms", C.c_double), ("adjoint_ms", C.c_double)]


class StringOut(C.Structure):
    _fields_ = [("n_string", C.c_int64), ("n_vel", C.c_int64), ("length", C.c_double), ("vel", C.c_double * 3)]


EXPORTS = ("axg_derive", "axg_get_dtau", "axg_create", "axg_destroy", "axg_last_error", "axg_set_fields",
           "axg_get_fields", "axg_init_flat", "axg_init_thermal", "axg_get_gw_state", "axg_set_gw_state", "axg_advance",
           "axg_advance_steps", "axg_gw_spectrum", "axg_defects", "axg_get_mask", "axg_get_flags",
           "axg_strings", "axg_get_strings", "axg_walls", "axg_get_walls",
           "axg_group_strings", "axg_group_walls", "axg_get_groups", "axg_field_spectrum",
           "axg_field_covariance", "axg_gradient", "axg_get_adjoint", "axg_batch_advance", "axg_energy",
           "axg_energy_fields", "axg_tune", "axg_profile", "axg_hbm_probe", "axg_plaquettes",
           "axg_get_plaquettes", "axg_set_stream", "axg_get_stream")

_LIB = None


def lib():
    """Load libaxgrav.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m axion_grav_amd.build`")
        L = C.CDLL(LIB_PATH)
        p, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "axg_derive": [C.POINTER(Keys), C.POINTER(Params), p, p, p],
            "axg_get_dtau": [C.POINTER(Params), d, d, p, p],
            "axg_create": [C.POINTER(Params), C.POINTER(p)],
            "axg_destroy": [p],
            "axg_set_fields": [p, p, p, i64],
            "axg_get_fields": [p, p, p, i64],
            "axg_init_flat": [p, C.c_uint64, d],
            "axg_init_thermal": [p, C.c_uint64, d],
            "axg_get_gw_state": [p, p, i64],
            "axg_set_gw_state": [p, p, i64],
            "axg_advance": [p, d, d, p],
            "axg_advance_steps": [p, d, d, i32],
            "axg_gw_spectrum": [p, p, p, p, p, i32],
            "axg_defects": [p, C.POINTER(DefectOut)],
            "axg_get_mask": [p, p, i64, p],
            "axg_get_flags": [p, p, i64],
            "axg_plaquettes": [p, C.POINTER(PlaquetteOut)],
            "axg_set_stream": [p, C.c_uint32, C.c_uint32, C.c_uint64],
            "axg_get_stream": [p, p],
            "axg_get_plaquettes": [p, p, i64, p],
            "axg_strings": [p, C.c_uint32, C.c_uint32, C.POINTER(StringOut)],
            "axg_get_strings": [p, p, i64, p],
            "axg_walls": [p, C.POINTER(WallOut)],
            "axg_get_walls": [p, p, i64, p],
            "axg_group_strings": [p, d, d, C.POINTER(GroupOut)],
            "axg_group_walls": [p, d, d, C.POINTER(GroupOut)],
            "axg_get_groups": [p, i32, p, i64, p],
            "axg_field_spectrum": [p, d, p, p, p, p, p, i32],
            "axg_field_covariance": [p, C.c_uint64, i32, i32, d, p, i64, p],
            "axg_gradient": [p, d, d, i32, i32, d, d, C.POINTER(GradOut)],
            "axg_batch_advance": [p, i32, d, d, i32, i32, p],
            "axg_get_adjoint": [p, p, p, i64],
            "axg_energy": [p, d, p],
            "axg_energy_fields": [C.POINTER(Params), p, p, d, p, p],
            "axg_hbm_probe": [i32, i64, i32, p],
            "axg_profile": [p, i32, p, p, i32],
            "axg_tune": [p, i32, p, p, i32],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.axg_last_error.argtypes = [p]
        L.axg_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


class AxgError(RuntimeError):
    pass


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def derive(sigma, g_rel, kappa, tau_fin, boxsize, ngrid, params=None):
    """settings.f90:274-297 derived units; returns (Params, extras dict)."""
    k = Keys(sigma, g_rel, kappa, tau_fin, boxsize, int(ngrid))
    P = params if params is not None else Params()
    z, w, t = C.c_double(), C.c_double(), C.c_double()
    rc = lib().axg_derive(C.byref(k), C.byref(P), C.byref(z), C.byref(w), C.byref(t))
    if rc != 0:
        raise AxgError(f"axg_derive failed ({rc})")
    return P, dict(zeta=z.value, width_cr=w.value, thickness_cr=t.value)


def batch_advance(lattices, t1, dt, nsteps, chunk=10):
    """Advance every Lattice of the list by nsteps steps in one captured graph of `chunk` steps
    of all of them (0: direct launches); returns the wall time in ms."""
    hs = (C.c_void_p * len(lattices))(*[L._h for L in lattices])
    ms = C.c_double()
    rc = lib().axg_batch_advance(hs, len(lattices), t1, dt, nsteps, chunk, C.byref(ms))
    if rc != 0:
        raise AxgError(f"axg_batch_advance failed ({ERRORS.get(rc, rc)}): {lib().axg_last_error(lattices[0]._h).decode()}")
    return ms.value


def energy_fields(params, phi_dev, pi_dev, tau, out_dev, stream=None):
    """axg_energy_fields on device pointers (ints: FP64 complex N^3 fields, 4-double output);
    launches on `stream` (a hipStream_t as an int, None = default) without synchronizing."""
    rc = lib().axg_energy_fields(C.byref(params), C.c_void_p(phi_dev), C.c_void_p(pi_dev), tau, C.c_void_p(out_dev),
                                 C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise AxgError(f"axg_energy_fields failed ({ERRORS.get(rc, rc)}): {lib().axg_last_error(None).decode()}")


def hbm_probe(nbytes=4 << 30, reps=10, device=0):
    """Streaming bandwidth of the device's HBM with the library's own 16-byte-per-lane kernels:
    dict(read, write, copy, copy_nt) in GB/s (a copy counts its reads and its writes)."""
    g = np.zeros(4)
    rc = lib().axg_hbm_probe(device, int(nbytes), int(reps), _ptr(g))
    if rc != 0:
        raise AxgError(f"axg_hbm_probe failed ({ERRORS.get(rc, rc)}): {lib().axg_last_error(None).decode()}")
    return dict(zip(("read", "write", "copy", "copy_nt"), (float(x) for x in g)))


def make_params(ngrid, dspa, sigma_cr, *, precision=FP64, grav_spec=True, initial_loop=True, kappa=1e-30,
                num_bin_grav=25, num_bin=50, id_strings=True, id_walls=True, spec_est=True, tau_cr=1.0,
                a_now=1.0, eps_dtau=0.0, dtau_fixed=0.0, gw_fix_swap=False, potential_kappa0=0.0, degrade_spec=0,
                string_vel=0, device=0):
    """Params from a lattice spacing and sigma_cr (dwav, dv_spa, dv_wav as settings.f90:295-297).
    degrade_spec > 0 puts the GW source and the axion spectrum on the grid ngrid / 2^degrade_spec;
    string_vel = 1 takes string velocities from velocity_string_YY (the reference's VEL = 1 build)."""
    twopi_t = 2 * 3.14159265359
    P = Params()
    P.ngrid = int(ngrid)
    P.precision = precision
    P.grav_spec = int(grav_spec)
    P.initial_loop = int(initial_loop)
    P.gw_fix_swap = int(gw_fix_swap)
    P.num_bin_grav = num_bin_grav
    P.num_bin = num_bin
    P.id_strings, P.id_walls, P.spec_est = int(id_strings), int(id_walls), int(spec_est)
    P.device = device
    P.dspa = dspa
    P.dwav = twopi_t / dspa / ngrid
    P.dv_spa = dspa * dspa * dspa
    r = P.dwav / twopi_t
    P.dv_wav = r * r * r
    P.sigma_cr = sigma_cr
    P.tau_cr = tau_cr
    P.a_now = a_now
    P.kappa = kappa
    P.eps_dtau = eps_dtau
    P.dtau_fixed = dtau_fixed
    P.potential_kappa0 = potential_kappa0
    P.degrade_spec = int(degrade_spec)
    P.string_vel = int(string_vel)
    return P


def get_dtau(P, t1, t2):
    dt, n = C.c_double(), C.c_int32()
    rc = lib().axg_get_dtau(C.byref(P), t1, t2, C.byref(dt), C.byref(n))
    if rc != 0:
        raise AxgError(f"axg_get_dtau failed ({rc})")
    return dt.value, n.value


class Lattice:
    """One device-resident lattice (an axg_handle)."""

    def __init__(self, params):
        self.params = params
        self.N = int(params.ngrid)
        h = C.c_void_p()
        rc = lib().axg_create(C.byref(params), C.byref(h))
        if rc != 0:
            msg = lib().axg_last_error(None).decode()
            raise AxgError(f"axg_create failed ({ERRORS.get(rc, rc)}): {msg}")
        self._h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().axg_last_error(self._h).decode()
            raise AxgError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}")

    def close(self):
        if self._h:
            lib().axg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n3(self):
        return self.N ** 3

    def set_fields(self, phi=None, pi=None):
        a = np.ascontiguousarray(phi, np.complex128) if phi is not None else None
        b = np.ascontiguousarray(pi, np.complex128) if pi is not None else None
        for x in (a, b):
            if x is not None and x.size != self.n3:
                raise ValueError("field size must be N^3")
        self._check(lib().axg_set_fields(self._h, _ptr(a), _ptr(b), self.n3), "axg_set_fields")

    def get_fields(self):
        phi = np.empty((self.N,) * 3, np.complex128)
        pi = np.empty((self.N,) * 3, np.complex128)
        self._check(lib().axg_get_fields(self._h, _ptr(phi), _ptr(pi), self.n3), "axg_get_fields")
        return phi, pi

    def set_stream(self, seed1, seed2, position=0):
        """Take the handle's random draws from the reference's MCG59 stream seeded by (seed1, seed2)
        (its key ss), positioned at `position` draws."""
        self._check(lib().axg_set_stream(self._h, seed1, seed2, int(position)), "axg_set_stream")

    def stream_position(self):
        pos = C.c_uint64()
        self._check(lib().axg_get_stream(self._h, C.byref(pos)), "axg_get_stream")
        return pos.value

    def init_flat(self, seed, kmax):
        self._check(lib().axg_init_flat(self._h, int(seed), float(kmax)), "axg_init_flat")

    def init_thermal(self, seed, tau_ini):
        """Thermal state at tau_ini < 1 (get_initial_thermal, initial.f90:80-238)."""
        self._check(lib().axg_init_thermal(self._h, seed, tau_ini), "axg_init_thermal")

    def gw_state(self):
        """The four GW accumulators on the half spectrum of the GW grid (the lattice's, or
        ngrid / 2^degrade_spec): (4, Ng, Ng, Ng/2 + 1) complex."""
        N = self.N >> int(self.params.degrade_spec)
        A = np.empty((4, N, N, N // 2 + 1), np.complex128)
        self._check(lib().axg_get_gw_state(self._h, _ptr(A), A.size), "axg_get_gw_state")
        return A

    def set_gw_state(self, A):
        A = np.ascontiguousarray(A, np.complex128)
        self._check(lib().axg_set_gw_state(self._h, _ptr(A), A.size), "axg_set_gw_state")

    def advance(self, t1, t2):
        n = C.c_int32()
        self._check(lib().axg_advance(self._h, t1, t2, C.byref(n)), "axg_advance")
        return n.value

    def advance_steps(self, t1, dt, nsteps):
        self._check(lib().axg_advance_steps(self._h, t1, dt, int(nsteps)), "axg_advance_steps")

    def gw_spectrum(self):
        nb = self.params.num_bin_grav
        q, pw, nu, de = (np.zeros(nb) for _ in range(4))
        self._check(lib().axg_gw_spectrum(self._h, _ptr(q), _ptr(pw), _ptr(nu), _ptr(de), nb), "axg_gw_spectrum")
        return dict(qmid=q, power=pw, numer=nu, denom=de)

    def defects(self):
        out = DefectOut()
        self._check(lib().axg_defects(self._h, C.byref(out)), "axg_defects")
        return np.array([out.n_string, out.n_wall, out.n_mask], np.int64)

    def mask_list(self):
        cnt = C.c_int64()
        self._check(lib().axg_get_mask(self._h, None, 0, C.byref(cnt)), "axg_get_mask")
        n = cnt.value
        buf = np.zeros((max(n, 1), 3), np.int32)
        self._check(lib().axg_get_mask(self._h, _ptr(buf), n, C.byref(cnt)), "axg_get_mask")
        return buf[:n]

    def flags(self):
        f = np.empty((self.N,) * 3, np.int8)
        self._check(lib().axg_get_flags(self._h, _ptr(f), self.n3), "axg_get_flags")
        return f

    def plaquettes(self):
        """Plaquette windings of the current phi: dict of (3,) int64 arrays n_pierced, n_plus,
        n_minus, n_quadrant, net (faces (x,y), (y,z), (z,x))."""
        out = PlaquetteOut()
        self._check(lib().axg_plaquettes(self._h, C.byref(out)), "axg_plaquettes")
        return {k: np.array(getattr(out, k)[:], np.int64) for k, _ in PlaquetteOut._fields_}

    def plaquette_records(self):
        """(n, 5) int32 rows (ix, iy, iz, o, n) of the pierced faces of the last plaquettes()."""
        cnt = C.c_int64()
        self._check(lib().axg_get_plaquettes(self._h, None, 0, C.byref(cnt)), "axg_get_plaquettes")
        n = cnt.value
        buf = np.zeros((max(n, 1), 5), np.int32)
        self._check(lib().axg_get_plaquettes(self._h, _ptr(buf), n, C.byref(cnt)), "axg_get_plaquettes")
        return buf[:n]

    def strings(self, seed0, seed1):
        """Fit every string cell of the last defects() (needs id_strings): dict(n_string, n_vel,
        length = sum of chord lengths in lattice units, vel = (mean |v|, mean v^2, mean gamma))."""
        out = StringOut()
        self._check(lib().axg_strings(self._h, seed0, seed1, C.byref(out)), "axg_strings")
        return dict(n_string=out.n_string, n_vel=out.n_vel, length=out.length, vel=np.array(out.vel[:]))

    def string_records(self):
        """(n_string, 11): pos(3, site coordinates counted from 0), dir(3), gof, sl, vel(3)."""
        cnt = C.c_int64()
        self._check(lib().axg_get_strings(self._h, None, 0, C.byref(cnt)), "axg_get_strings")
        n = cnt.value
        buf = np.zeros((max(n, 1), 11))
        self._check(lib().axg_get_strings(self._h, _ptr(buf), n, C.byref(cnt)), "axg_get_strings")
        return buf[:n]

    def walls(self):
        """Fit every wall cell of the last defects() (needs id_walls): dict(n_wall, area = sum of
        the cells' wall areas in lattice units)."""
        out = WallOut()
        self._check(lib().axg_walls(self._h, C.byref(out)), "axg_walls")
        return dict(n_wall=out.n_wall, area=out.area)

    def wall_records(self):
        """(n_wall, 8): pos(3, site coordinates counted from 0), dir(3), gof, wa."""
        cnt = C.c_int64()
        self._check(lib().axg_get_walls(self._h, None, 0, C.byref(cnt)), "axg_get_walls")
        n = cnt.value
        buf = np.zeros((max(n, 1), 8))
        self._check(lib().axg_get_walls(self._h, _ptr(buf), n, C.byref(cnt)), "axg_get_walls")
        return buf[:n]

    def group(self, kind, width, hor_size):
        """Networks of the last strings (kind 0) or walls (kind 1): dict(n_group, th_dist,
        sum (3,), ids (group number per record))."""
        o = GroupOut()
        fn = lib().axg_group_strings if kind == 0 else lib().axg_group_walls
        self._check(fn(self._h, width, hor_size, C.byref(o)), "axg_group_strings" if kind == 0 else "axg_group_walls")
        cnt = C.c_int64()
        self._check(lib().axg_get_groups(self._h, kind, None, 0, C.byref(cnt)), "axg_get_groups")
        ids = np.zeros(max(cnt.value, 1), np.int32)
        self._check(lib().axg_get_groups(self._h, kind, ids.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt)),
                    "axg_get_groups")
        return dict(n_group=o.n_group, th_dist=o.th_dist, sum=np.array(o.sum[:]), ids=ids[: cnt.value])

    def field_spectrum(self, tau):
        nb = self.params.num_bin
        q, pw, cv, nm = (np.zeros(nb) for _ in range(4))
        rho = C.c_double()
        self._check(lib().axg_field_spectrum(self._h, tau, _ptr(q), _ptr(pw), _ptr(cv), _ptr(nm), C.byref(rho), nb),
                    "axg_field_spectrum")
        return dict(qmid=q, power=pw, conv=cv, num=nm, rho=rho.value)

    def field_covariance(self, seed, min_mc=10, max_mc=100, tol=0.1):
        """Monte Carlo covariance of the last field_spectrum: (packed-upper matrix, realisations;
        negative: -(max_mc + 1) without convergence)."""
        nb = self.params.num_bin
        cov = np.zeros(nb * (nb + 1) // 2)
        it = C.c_int32()
        self._check(lib().axg_field_covariance(self._h, seed, min_mc, max_mc, tol, _ptr(cov), len(cov), C.byref(it)),
                    "axg_field_covariance")
        return cov, it.value

    def gradient(self, t1, dt, nsteps, K, qlo, qhi):
        """Run nsteps from t1 with the GW accumulators from zero and differentiate the GW band
        energy J (power summed over qlo <= qmid <= qhi): dict(J, dJ_dsigma2, dJ_dkappa0,
        dJ_dlnamp, forward_ms, adjoint_ms)."""
        o = GradOut()
        self._check(lib().axg_gradient(self._h, t1, dt, nsteps, K, qlo, qhi, C.byref(o)), "axg_gradient")
        return {k: getattr(o, k) for k, _ in GradOut._fields_}

    def adjoint(self):
        """(phibar, pibar) of the last gradient: dJ/dRe + i dJ/dIm of the initial fields."""
        N = self.params.ngrid
        a = np.zeros((N, N, N), np.complex128)
        b = np.zeros((N, N, N), np.complex128)
        self._check(lib().axg_get_adjoint(self._h, _ptr(a), _ptr(b), N ** 3), "axg_get_adjoint")
        return a, b

    TUNE_CONFIGS = ("fused: rows in dispatch order", "fused: wide blocks", "fused: other sites per thread",
                    "fused: other sites, wide blocks", "fused: XCD strips", "fused: XCD strips, wide blocks",
                    "fused: other sites, XCD strips", "fused: other sites, XCD strips, wide blocks",
                    "y: 128-byte tiles", "y: 64-byte tiles", "y: two planes per block")

    def tune(self, reps=5):
        """Autotune the fused pass and the y transform on this lattice:
        (fused-pass code, y-pass code, {config: ms per launch})."""
        v = C.c_int32()
        ms = np.zeros(len(self.TUNE_CONFIGS))
        self._check(lib().axg_tune(self._h, reps, C.byref(v), _ptr(ms), len(ms)), "axg_tune")
        return v.value & 0xFF, v.value >> 8, dict(zip(self.TUNE_CONFIGS, (float(x) for x in ms)))

    KERNEL_SLOTS = ("drift", "kick_source_drift", "gw_xfft", "gw_yfft", "gw_zfft_accumulate")

    def profile(self, enable):
        """Start (True: reset) or stop kernel timing; returns {slot: (total ms, launches)}."""
        ms = np.zeros(5)
        n = np.zeros(5, np.int64)
        self._check(lib().axg_profile(self._h, int(enable), _ptr(ms), _ptr(n), 5), "axg_profile")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNEL_SLOTS)}

    def energy(self, tau):
        out = np.zeros(4)
        self._check(lib().axg_energy(self._h, tau, _ptr(out)), "axg_energy")
        return out
