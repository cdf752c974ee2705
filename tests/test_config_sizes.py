"""GPU parity at the workloads' own sizes and with the reference's default physics.

Each test compares the HIP path (through the C-ABI) with the CPU FP64 oracle on the same inputs:

* 512^3 (L = 128): one FP64 DKD step with GW from a GPU-evolved field -- fields bit-exact,
  accumulators within 1e-10, GW spectrum bin counts bit-exact and bin sums within 1e-10 -- and
  the cell classification of that field: flags and counts
  bit-exact (integer work; the thresholds at pi and 7 pi/4 depend on device atan2 agreeing with
  the host's over 134 M cells); the plaquette windings of that field, bit-exact, with the
  winding invariants (zero net winding per orientation, none out of any cell).
* 128^3: FP64 GW accumulators over two steps (the XCD-strip configuration that 128 and 256 run
  by default) and the strip / dispatcher-order configurations bit-identical at 128 and 256.
* 128^3 adjoint (L = 32, kappa0 = 1e-6, K = 4) against the oracle's adjoint, and its FP32C
  gradient against the FP64 one.
* The thermal mass and a live axion mass (initial_loop = F, kappa = 1.37): the run crosses the
  maxion2 branch point thrm12 = kappa (cT/c0)^(1/n) (potential.f90:62-79) with m2(T) of
  evolve.f90:292-297 switched on; fields bit-exact with and without GW.
* dspa = 0.3 (not a power of two): the fused pass divides by 2 dspa as the reference does
  (grav_waves.f90:331-351); fields bit-exact and equal to the unfused kernel's.
"""
import numpy as np
import pytest

import oracle as O
from oracle import adjoint as OA
from axion_grav_amd import capi
from tests.conftest import SQ_HALF, phys

pytestmark = pytest.mark.gpu

PI_T = 3.14159265359
THRM_SIGMA = 2.449489742783178


def say(*a):
    print("[case]", *a, flush=True)


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


# ---------------------------------------------------------------- 512^3, L = 128

@pytest.fixture(scope="module")
def field512():
    """A 512^3 FP64 field (L = 128, seeded flat spectrum |k| < 1, seed 7) advanced 25 steps of
    0.05 from tau = 1 on the GPU, as the host gets it back."""
    N = 512
    say("512^3 field: init + 25 steps")
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP64, grav_spec=False)
    with capi.Lattice(prm) as L:
        L.init_flat(7, 1.0)
        L.advance_steps(1.0, 0.05, 25)
        phi, pi = L.get_fields()
    return phi, pi


def test_fp64_512_step_matches_oracle(field512):
    """One DKD step with GW at 512^3 in FP64 (the precision form's reference path): phi and pi
    bit-exact, the four accumulators within 1e-10 of the oracle's full-spectrum ones, and the
    128-bin GW spectrum of the device accumulators (257 folded kx per row, so a block thread
    bins more than one; rows whose modes lie past the last bin are skipped) against the oracle's
    binning of its own: mode counts bit-exact, power sums within 1e-10."""
    N = 512
    phi0, pi0 = field512
    say("512^3 FP64 step vs oracle")
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP64, num_bin_grav=128)
    t1, dt = 2.25, 0.05
    with capi.Lattice(prm) as L:
        L.set_fields(phi0, pi0)
        L.advance_steps(t1, dt, 1)
        g_phi, g_pi = L.get_fields()
        A = L.gw_state()
        spec = L.gw_spectrum()
    phi, pi = phi0.copy(), pi0.copy()
    OA_ = O.new_gw_state(N)
    O.symplectic(phi, pi, OA_, t1, dt, 1, P)
    assert np.array_equal(g_phi, phi), relerr(g_phi, phi)
    assert np.array_equal(g_pi, pi), relerr(g_pi, pi)
    del g_phi, g_pi, phi, pi
    h = N // 2 + 1
    for m in range(4):
        e = relerr(A[m], OA_[m][:, :, :h])
        say(f"  accumulator {m}: rel err {e:.2e}")
        assert e < 1e-10, (m, e)
    del A
    o = O.gw_spectrum(OA_, 128, P["dspa"], P["dwav"], P["tau_cr"], P["a_now"], P["dv_wav"])
    assert np.array_equal(spec["denom"], o["denom"])
    e = relerr(spec["numer"], o["numer"])
    say(f"  GW spectrum: rel err {e:.2e}")
    assert e < 1e-10, e


def test_classify_512_matches_oracle(field512):
    """Cell classification of the evolved 512^3 field: every flag and the three counts equal
    the oracle's (glibc atan2 on the host, the device's on the GPU)."""
    N = 512
    phi, pi = field512
    say("512^3 classification vs oracle")
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP64, grav_spec=False)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        cnt = L.defects()
        fl = L.flags()
    ofl, ocnt = O.classify(phi)
    say(f"  counts gpu {cnt.tolist()} oracle {ocnt.tolist()}")
    assert cnt[0] > 10000 and cnt[1] > 10000 and cnt[2] > 10000
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(fl, ofl)


def test_plaquettes_512_match_oracle(field512):
    """Plaquette windings of the evolved 512^3 field: counts and the list of pierced faces equal
    the oracle's, and the size-independent invariants hold on the GPU's own list: the windings of
    each orientation sum to 0, and the six faces of every cell carry no net winding."""
    N = 512
    phi, pi = field512
    say("512^3 plaquettes vs oracle")
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP64, grav_spec=False)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        cnt = L.plaquettes()
        rec = L.plaquette_records()
    code, ocnt, orec = O.plaquettes(phi)
    say(f"  pierced {cnt['n_pierced'].tolist()} quadrant {cnt['n_quadrant'].tolist()}")
    for k in ocnt:
        assert np.array_equal(cnt[k], ocnt[k]), k
    assert np.array_equal(rec, orec)
    assert cnt["n_pierced"].min() > 10000 and np.all(cnt["net"] == 0)
    n = np.zeros((N, N, N, 3), np.int8)
    n[rec[:, 2] - 1, rec[:, 1] - 1, rec[:, 0] - 1, rec[:, 3] - 1] = rec[:, 4]
    res = (np.roll(n[..., 0], -1, axis=0) - n[..., 0] + np.roll(n[..., 1], -1, axis=2) - n[..., 1]
           + np.roll(n[..., 2], -1, axis=1) - n[..., 2])
    assert not np.any(res)


# ---------------------------------------------------------------- 128^3 and 256^3

def test_gw_128_fp64_matches_oracle():
    """Two FP64 steps with GW at 128^3 (the batch's and the adjoint's size; the fused pass runs
    in XCD strips there by default): fields bit-exact, accumulators within 1e-10."""
    N = 128
    say("128^3 FP64 GW vs oracle")
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP64, num_bin_grav=64)
    phi, pi = O.ic_flat(N, 3, 2.0, P["dwav"])
    rng = np.random.default_rng(128)
    pi = 0.1 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(1.0, 0.05, 2)
        g_phi, g_pi = L.get_fields()
        A = L.gw_state()
        s = L.gw_spectrum()
    OA_ = O.new_gw_state(N)
    O.symplectic(phi, pi, OA_, 1.0, 0.05, 2, P)
    assert np.array_equal(g_phi, phi) and np.array_equal(g_pi, pi)
    for m in range(4):
        assert relerr(A[m], OA_[m][:, :, : N // 2 + 1]) < 1e-10
    o = O.gw_spectrum(OA_, 64, P["dspa"], P["dwav"], 1.0, 1.0, P["dv_wav"])
    assert np.array_equal(s["denom"], o["denom"])
    assert relerr(s["numer"], o["numer"]) < 1e-10


@pytest.mark.parametrize("N,prec", [(128, capi.FP32C), (256, capi.FP32C), (128, capi.FP64), (256, capi.FP64)])
def test_strip_configurations_bit_identical(N, prec, monkeypatch):
    """The fused pass's configurations at 128 and 256 (dispatcher order 0, wide blocks 1, XCD
    strips 8 -- the default at these sizes -- and strips with wide blocks 9): every site is
    updated once whatever the block order, so fields, accumulators and spectra are the same bit
    for bit."""
    say("strip configurations", N, prec)
    P = phys(N, 0.25)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=prec, num_bin_grav=32)
    out = []
    for v in ("0", "1", "8", "9"):
        monkeypatch.setenv("AXG_KICK_VARIANT", v)
        with capi.Lattice(prm) as L:
            L.init_flat(11, 2.0)
            L.advance_steps(1.0, 0.05, 3)
            phi, pi = L.get_fields()
            out.append((phi, pi, L.gw_state()))
    for phi, pi, A in out[1:]:
        assert np.array_equal(phi, out[0][0]) and np.array_equal(pi, out[0][1]) and np.array_equal(A, out[0][2])


# ---------------------------------------------------------------- 128^3 adjoint (L = 32)

def _c4_fields(N):
    P = phys(N, 0.25)
    phi, _ = O.ic_flat(N, 99, 6.0, P["dwav"])
    rng = np.random.default_rng(99)
    pi = 0.2 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    return phi, pi


def test_gradient_128_matches_oracle():
    """axg_gradient at 128^3, L = 32, kappa0 = 1e-6 (the differentiation workload's lattice and
    tilt), 8 steps of 0.01875, checkpoints every 4, band 0.5 <= q <= 2, 64 bins: J, the three
    derivatives and both adjoint fields within 1e-7 of the oracle's stored-trajectory adjoint
    (FP64); the FP32C gradient within 2e-4 of the FP64 one."""
    N, nb, qlo, qhi, t1, dt, n, K = 128, 64, 0.5, 2.0, 1.0, 0.01875, 8, 4
    say("gradient 128^3")
    P = dict(phys(N, 0.25), potential_kappa0=1e-6, tau_cr=9.42, a_now=5.4e30)
    phi0, pi0 = _c4_fields(N)
    res = {}
    for prec in (capi.FP64, capi.FP32C):
        prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=prec, num_bin_grav=nb, tau_cr=9.42, a_now=5.4e30,
                               potential_kappa0=1e-6)
        with capi.Lattice(prm) as L:
            L.set_fields(phi0, pi0)
            g = L.gradient(t1, dt, n, K, qlo, qhi)
            res[prec] = (g, L.adjoint())
    J, ds2, dk0, dlna, ophib, opib = OA.gradient(phi0, pi0, P, t1, dt, n, nb, qlo, qhi, False)
    g, (phib, pib) = res[capi.FP64]
    say(f"  J {g['J']:.9e}/{J:.9e} ds2 {g['dJ_dsigma2']:.6e}/{ds2:.6e} dk0 {g['dJ_dkappa0']:.6e}/{dk0:.6e} "
        f"dlnA {g['dJ_dlnamp']:.6e}/{dlna:.6e}")
    for k, o in (("J", J), ("dJ_dsigma2", ds2), ("dJ_dkappa0", dk0), ("dJ_dlnamp", dlna)):
        assert abs(g[k] / o - 1) < 1e-7, k
    assert np.abs(phib - ophib).max() < 1e-7 * np.abs(ophib).max()
    assert np.abs(pib - opib).max() < 1e-7 * np.abs(opib).max()
    c = res[capi.FP32C][0]
    for k in ("J", "dJ_dsigma2", "dJ_dkappa0", "dJ_dlnamp"):
        assert abs(c[k] / g[k] - 1) < 2e-4, (k, c[k], g[k])


# ---------------------------------------------------------------- default physics

@pytest.mark.parametrize("gw", [False, True])
def test_thermal_and_axion_mass_bitexact(gw):
    """initial_loop = F (the reference's default) switches on the thermal mass m2(T) =
    (T^2 - T_cr^2)/3 with T = T_cr/tau (evolve.f90:292-297), and kappa = 1.37 makes the axion
    mass live: maxion2 (potential.f90:62-79) takes its power-law branch while T > thrm12 =
    kappa (6.26)^(1/6.68) ~ 1.80 and its constant branch after, and the run (kicks at tau =
    1.025 .. 1.975) crosses that point.  Fields bit-exact, accumulators within 1e-10."""
    N = 32
    say("thermal + axion mass", gw)
    kappa = 1.37
    P = dict(phys(N, 0.25), initial_loop=False, kappa=kappa)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, grav_spec=gw, initial_loop=False, kappa=kappa)
    t1, dt, n = 1.0, 0.05, 20
    thrm12 = kappa * (6.26 / 1.0) ** (1 / 6.68)
    kicks = t1 + dt * (np.arange(n) + 0.5)
    assert (THRM_SIGMA / kicks).max() > thrm12 > (THRM_SIGMA / kicks).min()
    ma2 = [O.maxion2(THRM_SIGMA / t, kappa) for t in kicks]
    assert min(ma2) > 0.1  # the axion-mass shift -ma2/2 moves every site
    phi, pi = O.ic_flat(N, 41, 4.0, P["dwav"])
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(t1, dt, n)
        g_phi, g_pi = L.get_fields()
        A = L.gw_state() if gw else None
    o_phi, o_pi = phi.copy(), pi.copy()
    OA_ = O.new_gw_state(N) if gw else None
    O.symplectic(o_phi, o_pi, OA_, t1, dt, n, P)
    assert np.array_equal(g_phi, o_phi), relerr(g_phi, o_phi)
    assert np.array_equal(g_pi, o_pi), relerr(g_pi, o_pi)
    if gw:
        for m in range(4):
            assert relerr(A[m], OA_[m][:, :, : N // 2 + 1]) < 1e-10
    # both terms changed the trajectory: against initial_loop = T with kappa ~ 0
    n_phi, n_pi = phi.copy(), pi.copy()
    O.symplectic(n_phi, n_pi, None, t1, dt, n, phys(N, 0.25))
    assert relerr(n_pi, o_pi) > 1e-3


@pytest.mark.parametrize("N", [32, 48])
def test_non_power_of_two_dspa(N, monkeypatch):
    """dspa = 0.3: the gradient's divisor 2 dspa is not a power of two, so multiplying by its
    reciprocal would differ from the reference's division in the last bit.  The fused pass and
    the unfused kick kernel both divide: fields bit-exact against the oracle and each other,
    accumulators within 1e-10."""
    say("dspa 0.3", N)
    P = phys(N, 0.3)
    prm = capi.make_params(N, 0.3, SQ_HALF)
    phi, pi = O.ic_flat(N, 23, 3.0, P["dwav"])
    rng = np.random.default_rng(N)
    pi = 0.1 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    out = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("AXG_XFUSE", fuse)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.advance_steps(1.0, 0.06, 4)
            out.append((*L.get_fields(), L.gw_state()))
    OA_ = O.new_gw_state(N)
    o_phi, o_pi = phi.copy(), pi.copy()
    O.symplectic(o_phi, o_pi, OA_, 1.0, 0.06, 4, P)
    for g_phi, g_pi, A in out:
        assert np.array_equal(g_phi, o_phi) and np.array_equal(g_pi, o_pi)
        for m in range(4):
            assert relerr(A[m], OA_[m][:, :, : N // 2 + 1]) < 1e-10
