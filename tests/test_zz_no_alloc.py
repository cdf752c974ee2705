"""axg_energy_fields, the entry point that allocates nothing: its partial sums live in a fixed
device array of the library.  Two calls in a row on caller-held FP64 device fields, the second
on a smaller lattice (the array keeps the first call's contents), each bit-identical to
axg_energy on a handle holding the same fields; and the size and argument errors.

The caller's device buffers come from the HIP runtime the library itself is linked against
(libamdhip64.so.7, already loaded with it), so the test shares the library's runtime instance."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from axion_grav_amd import capi
from tests.conftest import SQ_HALF, phys

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """hipMalloc / hipMemcpy / hipFree / a stream of the runtime libaxgrav uses."""

    def __init__(self):
        capi.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.bufs = []

    def ok(self, rc, what):
        assert rc == 0, f"{what} failed: {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.bufs.append(p)
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy")
        return p.value

    def get(self, ptr, n):
        out = np.empty(n, np.float64)
        self.ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, D2H), "hipMemcpy")
        return out

    def stream(self):
        s = C.c_void_p()
        self.ok(self.rt.hipStreamCreate(C.byref(s)), "hipStreamCreate")
        return s

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _fields(N, seed):
    P = phys(N)
    phi, _ = O.ic_flat(N, seed, 4.0, P["dwav"])
    rng = np.random.default_rng(seed)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    return P, phi, pi


def test_energy_fields_matches_handle_energy_on_two_calls():
    hip = Hip()
    s = hip.stream()
    try:
        results = []
        for N, seed, kw in ((64, 5, dict(potential_kappa0=1e-3)), (32, 6, dict(initial_loop=False, kappa=1.2))):
            print(f"[case] energy_fields N={N}", flush=True)
            P, phi, pi = _fields(N, seed)
            prm = capi.make_params(N, P["dspa"], SQ_HALF, grav_spec=False, **kw)
            d_phi, d_pi = hip.put(phi.view(np.float64)), hip.put(pi.view(np.float64))
            d_out = hip.put(np.full(4, np.nan))
            capi.energy_fields(prm, d_phi, d_pi, 1.7, d_out, s.value)
            hip.ok(hip.rt.hipStreamSynchronize(s), "hipStreamSynchronize")
            got = hip.get(d_out, 4)
            with capi.Lattice(prm) as L:
                L.set_fields(phi, pi)
                ref = L.energy(1.7)
            results.append((got, ref))
        for got, ref in results:
            assert np.array_equal(got, ref), (got, ref)
        assert not np.array_equal(results[0][0], results[1][0])
    finally:
        hip.rt.hipStreamDestroy(s)
        hip.close()


def test_energy_fields_rejects_bad_arguments():
    hip = Hip()
    try:
        prm = capi.make_params(16, 0.25, SQ_HALF, grav_spec=False)
        d = hip.put(np.zeros(2 * 16 ** 3))
        out = hip.put(np.zeros(4))
        with pytest.raises(capi.AxgError, match="tau"):
            capi.energy_fields(prm, d, d, 0.0, out)
        big = capi.make_params(2048, 0.25, SQ_HALF, grav_spec=False)
        with pytest.raises(capi.AxgError, match="ngrid"):
            capi.energy_fields(big, d, d, 1.0, out)
        with pytest.raises(capi.AxgError, match="null"):
            capi.energy_fields(prm, 0, d, 1.0, out)
    finally:
        hip.close()
