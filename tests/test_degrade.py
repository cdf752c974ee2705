"""degrade_spec > 0 (grav_waves.f90:202-206, 358-507; spectrum.f90:736-898, 900-954): the GW
source and the axion spectrum on the reduced grid N / 2^degrade_spec, through the C-ABI, against
the oracle's restatement on the same inputs.  Bars: lattice fields bit-exact with and without
the reduced grid (it only reads them); FP64 GW accumulators and spectra <= 1e-10 relative
(FFT order differs); bin counts and mask bit-exact; the deconvolved spectrum <= 1e-8 and the
covariance <= 1e-6 (the bars of the lattice-grid tests); FP32C against FP64 within 1e-5."""
import numpy as np
import pytest

import oracle as O
from axion_grav_amd import capi
from tests.conftest import SQ_HALF, phys
from tests.test_gpu_parity import ic, packed, relerr, say

pytestmark = pytest.mark.gpu


def params(N, ds, **kw):
    P = phys(N)
    P["degrade_spec"] = ds
    return capi.make_params(N, P["dspa"], SQ_HALF, degrade_spec=ds, **kw), P


@pytest.mark.parametrize("N,ds", [(32, 1), (48, 1), (64, 2), (64, 1)])
def test_gw_on_reduced_grid_fp64(N, ds):
    """Accumulators after 4 steps on the N/2^ds grid (16 and 32: compile-time kernels; 24: the
    runtime plan), and the GW spectrum there (bin edges from pi / dspa_grav)."""
    say("degrade gw", N, ds)
    nb = 16
    prm, P = params(N, ds, num_bin_grav=nb, tau_cr=9.42, a_now=5.4e30)
    P.update(tau_cr=9.42, a_now=5.4e30)
    phi, pi = ic(N, 12, 3.0)
    dt = 0.05
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(1.0, dt, 4)
        f, p = L.get_fields()
        A = L.gw_state()
        s = L.gw_spectrum()
    Ng = N >> ds
    assert A.shape == (4, Ng, Ng, Ng // 2 + 1)
    OA = O.new_gw_state(Ng)
    of, op = phi.copy(), pi.copy()
    O.symplectic(of, op, OA, 1.0, dt, 4, P)
    assert np.array_equal(f, of) and np.array_equal(p, op)
    for m in range(4):
        assert relerr(A[m], OA[m][:, :, : Ng // 2 + 1]) < 1e-10, m
    u = O.reduced_units(P, ds)
    o = O.gw_spectrum(OA, nb, u["dspa"], u["dwav"], 9.42, 5.4e30, u["dv_wav"])
    assert np.array_equal(s["denom"], o["denom"]) and np.array_equal(s["qmid"], o["qmid"])
    assert relerr(s["numer"], o["numer"]) < 1e-10 and relerr(s["power"], o["power"]) < 1e-10


def test_gw_state_roundtrip_and_fp32c():
    """set_gw_state / gw_state on the reduced grid round-trip; FP32C lattices source the same
    FP64 reduced-grid accumulators from their FP32-compensated fields (within 1e-5 of FP64)."""
    say("degrade state / fp32c")
    N, ds = 32, 1
    phi, pi = ic(N, 5, 3.0)
    out = {}
    for prec in (capi.FP64, capi.FP32C):
        prm, _ = params(N, ds, precision=prec)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.advance_steps(1.0, 0.05, 3)
            out[prec] = L.gw_state()
            L.set_gw_state(out[prec] * 2)
            assert np.array_equal(L.gw_state(), out[prec] * 2)
    assert relerr(out[capi.FP32C], out[capi.FP64]) < 1e-5


@pytest.mark.parametrize("N,ds", [(32, 1), (48, 1)])
def test_field_spectrum_and_covariance_on_reduced_grid(N, ds):
    """compute_conv_spectrum with degrade: block means of pi, the coarse mask, the spectrum and its
    deconvolution on the N/2^ds grid; then estimate_covmat there from the reference's stream
    (4 (N/2^ds)^3 draws per realisation)."""
    say("degrade spectrum", N, ds)
    nb = 8
    prm, P = params(N, ds, grav_spec=False, num_bin=nb)
    phi, _ = ic(N, 31, 2.0)
    rng = np.random.default_rng(1)
    pi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    tau = 3.0
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.set_stream(3, 4, 100)
        L.defects()
        s = L.field_spectrum(tau)
        cov, it = L.field_covariance(5, 4, 6, 0.0)
        pos = L.stream_position()
    flags, _ = O.classify(phi)
    cf = O.degrade_mask(flags, ds)
    cpi = O.degrade_field(pi, ds)
    assert 0 < np.count_nonzero(cf) < cf.size
    u = O.reduced_units(P, ds)
    pw, num = O.conv_spectrum(cpi, cf, nb, u["dspa"], u["dwav"], u["dv_spa"], u["dv_wav"])
    assert np.array_equal(s["num"], num)
    assert relerr(s["conv"], pw * tau * tau) < 1e-10
    tab = O.init_spectrum(nb, u["dspa"])
    _, inv, info = O.matrix_deconv(cf, nb, u["dwav"], u["dv_spa"], u["dv_wav"], tab)
    assert info == 0
    M = O.sym_from_packed(inv, nb) / np.outer(num, num)
    assert relerr(s["power"], M @ (pw * tau * tau)) < 1e-8
    assert np.all(s["power"] > 0)
    ocov, oit, _ = O.field_covariance(cf, nb, u["dspa"], u["dwav"], s["power"], packed(M), 0, 4, 6, 0.0,
                                      stream=(3, 4, 100))
    assert it == oit == -7
    assert np.all(np.isfinite(cov)) and relerr(cov, ocov) < 1e-6
    assert pos == 100 + 4 * u["ngrid"] ** 3 * 6


def test_reduced_grid_refusals():
    """A grid below 8 sites and the entry points that need the lattice-grid GW state refuse."""
    say("degrade refusals")
    prm, _ = params(16, 2)
    with pytest.raises(RuntimeError):
        capi.Lattice(prm)
    prm, _ = params(32, 1)
    with capi.Lattice(prm) as L:
        L.init_flat(1, 2.0)
        with pytest.raises(RuntimeError):
            capi.batch_advance([L], 1.0, 0.05, 2, 0)
