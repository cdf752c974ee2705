"""Parity of the HIP path (called through the C-ABI) with the CPU FP64 oracle.

Every comparison runs the GPU and the oracle on the same seeded inputs.  Bars:
  * integer work (defect flags, counts, mask list, bin counts): bit-exact;
  * FP64 kick/drift without GW: bit-exact (same operation order, no contraction);
  * FP64 GW accumulators / spectra / energy: relative error <= 1e-10 (stated per test);
  * FP32C: per-output errors against FP64 with the tolerance written in the test.
"""
import numpy as np
import pytest

import oracle as O
from axion_grav_amd import capi
from tests.conftest import SQ_HALF, phys

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def params(N, **kw):
    P = phys(N)
    return capi.make_params(N, P["dspa"], SQ_HALF, **kw), P


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def ic(N, seed, kmax):
    P = phys(N)
    return O.ic_flat(N, seed, kmax, P["dwav"])


# ---------------------------------------------------------------- fields / integrator

@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_field_roundtrip(prec):
    say("roundtrip", prec)
    N = 16
    prm, _ = params(N, precision=prec, grav_spec=False)
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    pi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        a, b = L.get_fields()
    if prec == capi.FP64:
        assert np.array_equal(a, phi) and np.array_equal(b, pi)
    else:
        # FP32 value + int16 residual: ~39 significant bits
        assert relerr(a, phi) < 4e-11 and relerr(b, pi) < 4e-11


@pytest.mark.parametrize("N,nsteps", [(16, 1), (24, 3), (32, 10)])
def test_kick_drift_bitexact(N, nsteps):
    say("kick_drift", N, nsteps)
    prm, P = params(N, grav_spec=False)
    phi, pi = ic(N, 11, 4.0)
    rng = np.random.default_rng(N)
    pi = 0.1 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    t1, dt = 1.0, 0.05
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(t1, dt, nsteps)
        g_phi, g_pi = L.get_fields()
    o_phi, o_pi = phi.copy(), pi.copy()
    O.symplectic(o_phi, o_pi, None, t1, dt, nsteps, P)
    assert np.array_equal(g_phi, o_phi), relerr(g_phi, o_phi)
    assert np.array_equal(g_pi, o_pi), relerr(g_pi, o_pi)


@pytest.mark.parametrize("N,gw", [(16, False), (32, True)])
def test_kick_drift_tilt_bitexact(N, gw):
    """N_DW=2 tilt on (kappa0 = 1e-3, large enough to move every site): the plain kick and the
    fused kick+source kernel both match the oracle bit for bit."""
    say("kick_drift tilt", N, gw)
    prm, P = params(N, grav_spec=gw, potential_kappa0=1e-3)
    P = dict(P, potential_kappa0=1e-3)
    phi, pi = ic(N, 12, 4.0)
    t1, dt, nsteps = 1.0, 0.05, 4
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(t1, dt, nsteps)
        g_phi, g_pi = L.get_fields()
    o_phi, o_pi = phi.copy(), pi.copy()
    O.symplectic(o_phi, o_pi, O.new_gw_state(N) if gw else None, t1, dt, nsteps, P)
    assert np.array_equal(g_phi, o_phi), relerr(g_phi, o_phi)
    assert np.array_equal(g_pi, o_pi), relerr(g_pi, o_pi)
    # the tilt changed the trajectory
    n_phi, n_pi = phi.copy(), pi.copy()
    O.symplectic(n_phi, n_pi, None, t1, dt, nsteps, dict(P, potential_kappa0=0.0))
    assert relerr(n_pi, o_pi) > 1e-6


@pytest.mark.parametrize("prec,tol", [(capi.FP64, 1e-12), (capi.FP32C, 1e-9)])
def test_time_reversal_residual(prec, tol):
    """Forward 30 steps, then 30 steps of -dt from the end time: the fields come back to
    round-off (the integrator is time-reversible; the fused GW path runs in both legs)."""
    say("time reversal", prec)
    N = 32
    prm, P = params(N, precision=prec)
    phi, pi = ic(N, 21, 4.0)
    t1, dt, n = 1.0, 0.05, 30
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(t1, dt, n)
        mid, _ = L.get_fields()
        L.advance_steps(t1 + n * dt, -dt, n)
        b_phi, b_pi = L.get_fields()
    assert relerr(mid, phi) > 1e-2  # the fields did move
    assert relerr(b_phi, phi) < tol, relerr(b_phi, phi)
    assert np.abs(b_pi - pi).max() < tol * np.abs(phi).max(), np.abs(b_pi - pi).max()


@pytest.mark.parametrize("prec,N", [(capi.FP64, 64), (capi.FP32C, 64), (capi.FP32C, 96)])
def test_tune_keeps_fields_and_results(prec, N):
    """axg_tune times the fused pass's configurations on the live lattice: the fields are the
    same before and after, and the steps after tuning match untuned steps bit for bit."""
    say("tune", prec, N)
    prm, P = params(N, precision=prec)
    phi, pi = ic(N, 9, 4.0)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        a_phi, a_pi = L.get_fields()
        v, yv, ms = L.tune(3)
        b_phi, b_pi = L.get_fields()
        assert np.array_equal(a_phi, b_phi) and np.array_equal(a_pi, b_pi)
        assert v in (0, 1, 4, 5, 8, 9, 12, 13) and yv in (0, 1, 2) and all(t > 0 for t in ms.values())
        L.advance_steps(1.0, 0.05, 3)
        t_phi, t_pi = L.get_fields()
        t_A = L.gw_state()
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(1.0, 0.05, 3)
        u_phi, u_pi = L.get_fields()
        u_A = L.gw_state()
    assert np.array_equal(t_phi, u_phi) and np.array_equal(t_pi, u_pi) and np.array_equal(t_A, u_A)


# ---------------------------------------------------------------- GW accumulation

@pytest.mark.parametrize("N,nsteps,fix", [(16, 1, False), (16, 3, True), (24, 2, False), (32, 4, False),
                                           (64, 2, False), (48, 2, True), (36, 2, False), (36, 1, True)])
def test_gw_accumulators_fp64(N, nsteps, fix):
    say("gw", N, nsteps, fix)
    prm, P = params(N, gw_fix_swap=fix)
    phi, pi = ic(N, 5, 4.0)
    t1, dt = 1.0, 0.05
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(t1, dt, nsteps)
        A = L.gw_state()
        g_phi, _ = L.get_fields()
    o_phi, o_pi = phi.copy(), pi.copy()
    OA = O.new_gw_state(N)
    O.symplectic(o_phi, o_pi, OA, t1, dt, nsteps, P, fix_swap=fix)
    assert np.array_equal(g_phi, o_phi)
    h = N // 2 + 1
    for m in range(4):
        e = relerr(A[m], OA[m][:, :, :h])
        assert e < 1e-10, (m, e)


@pytest.mark.parametrize("N", [16, 32])
def test_gw_spectrum_fp64(N):
    say("gw_spectrum", N)
    prm, P = params(N, num_bin_grav=25, tau_cr=9.42, a_now=5.4e30)
    P.update(tau_cr=9.42, a_now=5.4e30)
    phi, pi = ic(N, 21, 4.0)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(1.0, 0.05, 3)
        s = L.gw_spectrum()
    OA = O.new_gw_state(N)
    O.symplectic(phi, pi, OA, 1.0, 0.05, 3, P)
    o = O.gw_spectrum(OA, 25, P["dspa"], P["dwav"], 9.42, 5.4e30, P["dv_wav"])
    assert np.array_equal(s["denom"], o["denom"])
    assert relerr(s["numer"], o["numer"]) < 1e-10
    assert relerr(s["power"], o["power"]) < 1e-10
    assert np.array_equal(s["qmid"], o["qmid"])


@pytest.mark.parametrize("N", [96, 192, 384])
def test_gw_radix3_short(N):
    """N = 3 * 2^k runs the compile-time kernels with their radix-3 step after the register
    stages: 96 = (8 * 4) * 3, 192 = (8 * 8) * 3, 384 = (8 * 8 * 2) * 3.  N = 36 in the
    accumulator test above runs the runtime-plan kernels."""
    say("gw radix-3", N)
    prm, P = params(N)
    phi, pi = ic(N, 4, 2.0)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.advance_steps(1.0, 0.05, 1)
        A = L.gw_state()
    OA = O.new_gw_state(N)
    O.symplectic(phi, pi, OA, 1.0, 0.05, 1, P)
    for m in range(4):
        assert relerr(A[m], OA[m][:, :, : N // 2 + 1]) < 1e-10


@pytest.mark.parametrize("N,prec", [(64, capi.FP32C), (96, capi.FP64)])
def test_ypass_variants_bit_identical(N, prec, monkeypatch):
    """The y pass's three configurations (128-byte tiles, 64-byte tiles, two z planes per block)
    give the same accumulators bit for bit."""
    say("y variants", N, prec)
    prm, _ = params(N, precision=prec)
    phi, pi = ic(N, 12, 4.0)
    out = []
    for v in ("0", "1", "2"):
        monkeypatch.setenv("AXG_YPASS_VARIANT", v)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.advance_steps(1.0, 0.05, 2)
            out.append(L.gw_state())
    for A in out[1:]:
        assert np.array_equal(A, out[0])


@pytest.mark.parametrize("N,prec", [(768, capi.FP32C), (768, capi.FP64), (1024, capi.FP32C)])
def test_ypass_line_tiles_large(N, prec, monkeypatch):
    """At N = 768 and 1024 the one-column y-pass tile is narrower than a 128-byte line (8 FP32 or
    4 FP64 columns: the 1024-thread block, the 64 KiB tile); variants 0 and 2 then run two
    adjacent columns per thread over a 128-byte-wide tile (96-128 KiB of LDS), variant 1 the
    one-column tile.  One step from the same field gives the same GW spectrum sums bit for bit."""
    say("y line tiles", N, prec)
    prm, _ = params(N, precision=prec, num_bin_grav=64)
    out = []
    for v in ("0", "1", "2"):
        monkeypatch.setenv("AXG_YPASS_VARIANT", v)
        with capi.Lattice(prm) as L:
            L.init_flat(3, 1.0)
            L.advance_steps(1.0, 0.05, 1)
            s = L.gw_spectrum()
        out.append((s["numer"], s["denom"]))
    assert np.any(out[0][0] > 0) and np.all(np.isfinite(out[0][0]))
    for nu, de in out[1:]:
        assert np.array_equal(nu, out[0][0]) and np.array_equal(de, out[0][1])


@pytest.mark.parametrize("N", [32, 48])
def test_gw_unfused_x_pass_matches_fused(N, monkeypatch):
    """AXG_XFUSE=0 runs the x transform as its own pass (kick_drift_kernel + gw_xpass_p2_kernel):
    with dspa a power of two the FP64 fields equal the fused pass's bit for bit; the accumulators
    agree to FFT rounding (the two kernels' transforms may fuse multiply-adds differently), and
    both match the oracle."""
    say("unfused x", N)
    prm, P = params(N)
    phi, pi = ic(N, 8, 4.0)
    out = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("AXG_XFUSE", fuse)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.advance_steps(1.0, 0.05, 2)
            out.append((L.get_fields()[0], L.gw_state()))
    assert np.array_equal(out[0][0], out[1][0])
    OA = O.new_gw_state(N)
    O.symplectic(phi, pi, OA, 1.0, 0.05, 2, P)
    for m in range(4):
        assert relerr(out[1][1][m], out[0][1][m]) < 1e-13
        for f in out:
            assert relerr(f[1][m], OA[m][:, :, : N // 2 + 1]) < 1e-10


# ---------------------------------------------------------------- defects / spectra / energy

@pytest.mark.parametrize("N,kmax", [(16, 6.0), (32, 6.0)])
def test_defects_bitexact(N, kmax):
    say("defects", N)
    prm, P = params(N, grav_spec=False)
    phi, pi = ic(N, 77, kmax)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        cnt = L.defects()
        fl = L.flags()
        ml = L.mask_list()
    ofl, ocnt = O.classify(phi)
    assert np.array_equal(cnt, ocnt), (cnt, ocnt)
    assert np.array_equal(fl, ofl)
    zz, yy, xx = np.nonzero(ofl & 4)
    exp = np.stack([xx + 1, yy + 1, zz + 1], axis=1).astype(np.int32)
    assert np.array_equal(ml, exp)


@pytest.mark.parametrize("N", [16, 32, 48])
def test_field_spectrum_fp64(N):
    say("field_spectrum", N)
    nb = 8 if N == 16 else 12
    prm, P = params(N, grav_spec=False, num_bin=nb)
    phi, _ = ic(N, 31, 6.0)
    rng = np.random.default_rng(1)
    pi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    tau = 3.0
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.defects()
        s = L.field_spectrum(tau)
    flags, _ = O.classify(phi)
    pw, num = O.conv_spectrum(pi, flags, nb, P["dspa"], P["dwav"], P["dv_spa"], P["dv_wav"])
    assert np.array_equal(s["num"], num)
    assert relerr(s["conv"], pw * tau * tau) < 1e-10
    tab = O.init_spectrum(nb, P["dspa"])
    mat, inv, info = O.matrix_deconv(flags, nb, P["dwav"], P["dv_spa"], P["dv_wav"], tab)
    assert info == 0
    M = O.sym_from_packed(inv, nb) / np.outer(num, num)
    power = M @ (pw * tau * tau)
    assert relerr(s["power"], power) < 1e-8


@pytest.mark.parametrize("N,kmax,min_mc,max_mc,tol", [(16, 2.0, 10, 100, 0.1), (32, 2.0, 4, 12, 0.0),
                                                       (16, 6.0, 10, 20, 0.1), (48, 2.0, 4, 6, 0.0)])
def test_field_covariance_fp64(N, kmax, min_mc, max_mc, tol):
    """estimate_covmat on the GPU against the oracle's restatement on the same mask, power (the
    GPU's) and deconvolution matrix (the oracle's, which agrees with the GPU's to 1e-8): the
    same realisation count, and the covariance to 1e-6.  The second case never converges
    (tol = 0): counter -(max_mc + 1).  In the third, most cells are masked and a bin's
    deconvolved power is negative: sqrt(power/2) is NaN there, and so is every covariance
    entry, on both sides."""
    say("field_covariance", N, kmax, min_mc, max_mc, tol)
    nb = 8 if N == 16 else 12
    prm, P = params(N, grav_spec=False, num_bin=nb)
    phi, _ = ic(N, 31, kmax)
    rng = np.random.default_rng(1)
    pi = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    tau = 3.0
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.defects()
        with pytest.raises(RuntimeError):
            L.field_covariance(5, min_mc, max_mc, tol)  # no spectrum yet
        s = L.field_spectrum(tau)
        cov, it = L.field_covariance(5, min_mc, max_mc, tol)
        cov2, it2 = L.field_covariance(5, min_mc, max_mc, tol)
    assert it == it2 and np.array_equal(cov, cov2, equal_nan=True)  # deterministic
    flags, _ = O.classify(phi)
    assert np.count_nonzero(flags & 4) > 0
    assert (s["power"].min() < 0) == (kmax == 6.0)
    pw, num = O.conv_spectrum(pi, flags, nb, P["dspa"], P["dwav"], P["dv_spa"], P["dv_wav"])
    tab = O.init_spectrum(nb, P["dspa"])
    _, inv, info = O.matrix_deconv(flags, nb, P["dwav"], P["dv_spa"], P["dv_wav"], tab)
    M = O.sym_from_packed(inv, nb) / np.outer(num, num)
    ocov, oit, _ = O.field_covariance(flags, nb, P["dspa"], P["dwav"], s["power"], packed(M), 5, min_mc,
                                      max_mc, tol)
    assert it == oit
    if tol == 0.0:
        assert it == -(max_mc + 1)
    if kmax == 6.0:
        assert np.all(np.isnan(cov)) and np.all(np.isnan(ocov))
    else:
        assert np.all(np.isfinite(cov)) and relerr(cov, ocov) < 1e-6


def packed(M):
    nb = M.shape[0]
    return np.array([M[j - 1, i - 1] for i in range(1, nb + 1) for j in range(1, i + 1)])


@pytest.mark.parametrize("prec,tol,kappa0", [(capi.FP64, 1e-12, 0.0), (capi.FP32C, 1e-9, 0.0),
                                              (capi.FP64, 1e-12, 1e-3)])
def test_energy(prec, tol, kappa0):
    say("energy", prec, kappa0)
    N = 24
    prm, P = params(N, grav_spec=False, precision=prec, potential_kappa0=kappa0)
    phi, _ = ic(N, 8, 4.0)
    rng = np.random.default_rng(2)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        e = L.energy(2.5)
    o = O.energy(phi, pi, 2.5, P["dspa"], P["sigma_cr"], P["kappa"], True, kappa0)
    assert relerr(e, o) < tol


@pytest.mark.parametrize("N,seed,kmax", [(16, 1234, 4.0), (64, 1234, 1.0)])
def test_initial_field(N, seed, kmax):
    say("ic", N)
    prm, P = params(N, grav_spec=False)
    with capi.Lattice(prm) as L:
        L.init_flat(seed, kmax)
        phi, pi = L.get_fields()
    ophi, opi = O.ic_flat(N, seed, kmax, P["dwav"])
    assert relerr(phi, ophi) < 1e-12
    assert not np.any(pi)


# ---------------------------------------------------------------- string fits

@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_strings_match_oracle(prec):
    """String fits on the GPU against the oracle on the same field, cell list and seed.  The
    cell list is bit-exact; each fit runs the same FP64 search from the same Philox starting
    points, so the searches take the same path and the records differ only by last-bit
    differences of atan2/sin/cos (GPU vs libm) carried through sqrt(1 - mu^2) near |mu| = 1 and
    the velocity fit's 1/det (measured: <= 1.2e-6 relative, ~10 % of records above 1e-9).
    Bars: every record within 1e-5 (a search that took another path differs by O(0.1)), the
    summed length within 1e-9, the mean velocities within 1e-8, n_vel equal."""
    say("strings", prec)
    N = 32
    prm, P = params(N, grav_spec=False, precision=prec, id_strings=True, id_walls=True, spec_est=True)
    phi, _ = ic(N, 31, 6.0)
    rng = np.random.default_rng(4)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        f_phi, f_pi = L.get_fields()       # the values the GPU holds (FP32C rounds them)
        cnt = L.defects()
        out = L.strings(1234, 99)
        rec = L.string_records()
        flags = L.flags()
    oflags, ocnt = O.classify(f_phi)
    assert np.array_equal(flags, oflags) and cnt[0] == ocnt[0] == out["n_string"] > 50
    orec, osums = O.strings(f_phi, f_pi, oflags, 1234, 99, P["dspa"])
    assert rec.shape == orec.shape
    rel = np.abs(rec - orec) / np.maximum(1, np.abs(orec))
    print(f"  strings {len(rec)}: max rel diff {rel.max():.2e}; length {out['length']:.10f} vs {osums[0]:.10f}")
    assert rel.max() < 1e-5
    assert abs(out["length"] - osums[0]) <= 1e-9 * osums[0]
    assert out["n_vel"] == osums[1]
    ovel = osums[2:5] / osums[1]
    assert np.all(np.abs(out["vel"] - ovel) <= 1e-8 * np.abs(ovel))


def test_strings_match_oracle_128():
    """The string fits at 128^3 on an evolved field (the oracle's own FP64 evolution, 12 DKD
    steps of dt = 0.05 from a flat-spectrum start: 11,622 string cells, 47 % of them still
    searching after 64 simplex iterations, so most long searches run in the second kernel), FP64
    storage so that the GPU holds the oracle's values: classification bit-exact, then the same
    bars as at 32^3 -- every
    record within 1e-5 (a search that took another path differs by O(0.1)), the summed length
    within 1e-9, the mean velocities within 1e-8, n_vel equal."""
    N = 128
    say("strings 128", N)
    prm, P = params(N, grav_spec=False, precision=capi.FP64, id_strings=True, id_walls=True, spec_est=True)
    phi, pi = O.ic_flat(N, 3, 1.0, P["dwav"])
    O.symplectic(phi, pi, None, 1.0, 0.05, 12, P)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        cnt = L.defects()
        out = L.strings(1234, 99)
        rec = L.string_records()
        flags = L.flags()
    oflags, ocnt = O.classify(phi)
    assert np.array_equal(flags, oflags) and cnt[0] == ocnt[0] == out["n_string"] > 10000
    orec, osums = O.strings(phi, pi, oflags, 1234, 99, P["dspa"])
    rel = np.abs(rec - orec) / np.maximum(1, np.abs(orec))
    print(f"  strings {len(rec)}: max rel diff {rel.max():.2e}, {np.count_nonzero(rel.max(1) > 1e-9)} records "
          f"above 1e-9; length {out['length']:.10f} vs {osums[0]:.10f}")
    assert rel.max() < 1e-5
    assert abs(out["length"] - osums[0]) <= 1e-9 * osums[0]
    assert out["n_vel"] == osums[1]
    ovel = osums[2:5] / osums[1]
    assert np.all(np.abs(out["vel"] - ovel) <= 1e-8 * np.abs(ovel))


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_strings_vel_yy_matches_oracle(prec):
    """string_vel = 1, the reference's VEL = 1 build: the velocities come from velocity_string_YY
    (defects_3d.f90:824-982) -- quadrant tests, a plane fit per pierced face, weighted central
    differences, the last pierced face's velocity -- against the oracle's restatement on the
    values the GPU holds: velocities within 1e-12 relative (no transcendental but the quadrants'
    atan2), the fits as in the VEL = 0 test, and the velocity sums from the same records."""
    say("strings vel_yy", prec)
    N = 32
    prm, P = params(N, grav_spec=False, precision=prec, id_strings=True, id_walls=True, spec_est=True,
                    string_vel=1)
    phi, _ = ic(N, 31, 6.0)
    rng = np.random.default_rng(4)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        f_phi, f_pi = L.get_fields()
        L.defects()
        out = L.strings(1234, 99)
        rec = L.string_records()
    oflags, _ = O.classify(f_phi)
    orec, osums = O.strings(f_phi, f_pi, oflags, 1234, 99, P["dspa"], vel_yy=True)
    assert rec.shape == orec.shape and len(rec) > 50
    vrel = np.abs(rec[:, 8:] - orec[:, 8:]) / np.maximum(1e-3, np.abs(orec[:, 8:]))
    assert vrel.max() < 1e-12, vrel.max()
    assert np.max(np.abs(rec[:, :8] - orec[:, :8]) / np.maximum(1, np.abs(orec[:, :8]))) < 1e-5
    assert out["n_vel"] == osums[1]
    ovel = osums[2:5] / osums[1]
    assert np.all(np.abs(out["vel"] - ovel) <= 1e-10 * np.abs(ovel))


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_string_fit_split_bit_identical(prec, monkeypatch):
    """The string fits run in two kernels: a search still running after `split` simplex
    iterations is parked and continued by the second.  The records and the sums are the same bit
    for bit whatever the split (0: every unconverged search is parked after the random starts;
    1000: none is).  Searches that end unconverged (chi2 > 0.05, so gof < exp(-0.025)) ran all
    1000 iterations, so at the default split they went through the second kernel."""
    say("string fit split", prec)
    N = 32
    prm, P = params(N, grav_spec=False, precision=prec, id_strings=True, id_walls=True, spec_est=True)
    phi, _ = ic(N, 31, 6.0)
    rng = np.random.default_rng(4)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    out = []
    for split in ("0", "1", "32", "1000"):
        monkeypatch.setenv("AXG_STRING_SPLIT", split)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.defects()
            o = L.strings(1234, 99)
            out.append((o, L.string_records()))
    o0, r0 = out[0]
    assert len(r0) > 50 and np.count_nonzero(r0[:, 6] < np.exp(-0.025)) > 0
    for o, r in out[1:]:
        assert np.array_equal(r, r0)
        assert o["length"] == o0["length"] and o["n_vel"] == o0["n_vel"] and np.array_equal(o["vel"], o0["vel"])


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_walls_match_oracle(prec):
    """Wall fits on the GPU against the oracle on the same field and cell list: no random
    numbers, the same grid search and polygon; records differ by last-bit libm differences
    only.  Bars: every record within 1e-9, the summed area within 1e-12."""
    say("walls", prec)
    N = 32
    prm, P = params(N, grav_spec=False, precision=prec, id_strings=True, id_walls=True, spec_est=True)
    phi, pi = ic(N, 31, 6.0)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        f_phi, _ = L.get_fields()
        cnt = L.defects()
        out = L.walls()
        rec = L.wall_records()
    flags, ocnt = O.classify(f_phi)
    orec, osums = O.walls(f_phi, flags)
    assert cnt[1] == ocnt[1] == out["n_wall"] > 50
    rel = np.abs(rec - orec) / np.maximum(1, np.abs(orec))
    print(f"  walls {len(rec)}: max rel diff {rel.max():.2e}; area {out['area']:.10f} vs {osums[0]:.10f}")
    assert rel.max() < 1e-9
    assert abs(out["area"] - osums[0]) <= 1e-12 * osums[0]


@pytest.mark.parametrize("N,kmax,width,hor,route", [(32, 1.0, 1.0, 2.0, "box"), (64, 1.5, 1.0, 3.0, "box"),
                                                      (64, 2.0, 3.0, 2.0, "box"), (64, None, 40.0, 3.0, "pairs"),
                                                      (32, 1.0, 1e6, 2.0, "one group")])
def test_grouping_matches_oracle(N, kmax, width, hor, route):
    """String and wall networks on the GPU against the oracle's brute-force grouping of the
    same records: identical group numbers per record (integer work, bit-exact) and identical
    grouped sums (the same values summed in the same order).  The thresholds exercise the
    three routes: the cell box (th_dist of 2-3.2 cells), every pair (16.9 cells, on a field of
    four straight strings 32 cells apart and two wall sheets) and the single group (th_dist
    beyond the largest periodic distance, as for walls at kappa ~ 0)."""
    say("grouping", N, kmax, width, hor, route)
    prm, P = params(N, grav_spec=False, id_strings=True, id_walls=True, spec_est=True)
    rng = np.random.default_rng(6)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    if kmax is None:
        x = np.arange(N) + 0.3
        phi = np.sin(2 * np.pi * x[None, None, :] / N) + 1j * np.sin(2 * np.pi * x[None, :, None] / N)
        phi = np.ascontiguousarray(np.broadcast_to(phi, (N, N, N)))
    else:
        phi, _ = ic(N, 17, kmax)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.defects()
        with pytest.raises(RuntimeError):
            L.group(0, width, hor)  # no fits yet
        L.strings(5, 6)
        srec = L.string_records()
        gs = L.group(0, width, hor)
        L.walls()
        wrec = L.wall_records()
        gw = L.group(1, width, hor)
        gs2 = L.group(0, width, hor)
    assert np.array_equal(gs["ids"], gs2["ids"]) and np.array_equal(gs["sum"], gs2["sum"])
    th = O.group_threshold(width, hor)
    for g, rec, t2, t3 in ((gs, srec, 2 * hor, 2 * PI_T * hor), (gw, wrec, PI_T * hor ** 2, 4 * PI_T * hor ** 2)):
        assert g["th_dist"] == th and len(g["ids"]) == len(rec) > 50
        lab = O.group_points(rec, N, th)
        ng, ids, s = O.group_sums(lab, rec[:, 7], t2, t3)
        print(f"  {len(rec)} records, {ng} groups, sums {s}")
        assert g["n_group"] == ng and np.array_equal(g["ids"], ids)
        assert (ng == 1) == (route == "one group")
        assert np.array_equal(g["sum"], s)


PI_T = 3.14159265359


def test_string_velocity_of_moving_string():
    """A straight string along z moving in x at 0.3: phi = (X - X0 - u tau) + i (Y - Y0) in
    physical units (pi = -u), so the fitted velocity of each of its cells is (u, 0, 0) to
    round-off and the chord about one cell."""
    say("moving string")
    N, u = 16, 0.3
    prm, P = params(N, grav_spec=False, id_strings=True)
    d = P["dspa"]
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    x0, y0 = 7.3, 8.6
    phi = np.ascontiguousarray((x - x0) * d + 1j * (y - y0) * d)
    pi = np.full((N, N, N), -u + 0j)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        L.defects()
        L.strings(5, 6)
        rec = L.string_records()
    on_line = (np.abs(rec[:, 0] - x0) < 1) & (np.abs(rec[:, 1] - y0) < 1)   # site coordinates from 0
    assert on_line.sum() == N   # one cell per z-plane (other cells sit on the periodic seams)
    r = rec[on_line]
    assert np.abs(r[:, 8] - u).max() < 1e-12 and np.abs(r[:, 9:11]).max() < 1e-12
    assert np.abs(r[:, 7] - 1).max() < 0.15


# ---------------------------------------------------------------- thermal initial state

@pytest.mark.parametrize("N,prec", [(32, capi.FP64), (48, capi.FP64), (32, capi.FP32C), (192, capi.FP64)])
def test_thermal_state_matches_oracle(N, prec):
    """The GPU thermal state (same Philox draws, device exp/log/sqrt, FP64 transform) against
    the oracle: <= 1e-12 relative in FP64, the FP32C storage rounding otherwise."""
    say("thermal", N, prec)
    prm, P = params(N, grav_spec=False, initial_loop=False, precision=prec)
    with capi.Lattice(prm) as L:
        L.init_thermal(77, 0.2)
        phi, pi = L.get_fields()
    ophi, opi = O.ic_thermal(N, 77, 0.2, P["sigma_cr"], P["dwav"], P["dv_wav"])
    tol = 1e-12 if prec == capi.FP64 else 1e-10
    assert relerr(phi, ophi) < tol and relerr(pi, opi) < tol


def test_thermal_state_rejects_tau_ini_at_or_above_one():
    prm, _ = params(16, grav_spec=False, initial_loop=False)
    with capi.Lattice(prm) as L:
        with pytest.raises(capi.AxgError, match="tau_ini"):
            L.init_thermal(1, 1.0)


# ---------------------------------------------------------------- FP32C vs FP64

@pytest.mark.parametrize("N,fix", [(32, False), (96, False), (32, True), (48, True)])
def test_fp32c_tracks_fp64_gw(N, fix):
    say("fp32c vs fp64", N, fix)
    phi, pi = ic(N, 6, 3.0)
    out = {}
    for prec in (capi.FP64, capi.FP32C, capi.FP32):
        prm, P = params(N, precision=prec, gw_fix_swap=fix)
        with capi.Lattice(prm) as L:
            L.set_fields(phi, pi)
            L.advance_steps(1.0, 0.05, 20)
            out[prec] = (L.get_fields()[0], L.gw_spectrum())
    ref_phi, ref_s = out[capi.FP64]
    c_phi, c_s = out[capi.FP32C]
    # compensated storage keeps phi within ~1e-9 of FP64 over 20 steps
    assert relerr(c_phi, ref_phi) < 1e-8
    # GW spectrum with FP32 transforms/accumulators: per-bin relative error
    nz = ref_s["numer"] > 0
    err = np.abs(c_s["numer"][nz] - ref_s["numer"][nz]) / ref_s["numer"][nz]
    assert err.max() < 1e-4, err.max()
    assert np.array_equal(c_s["denom"], ref_s["denom"])


@pytest.mark.parametrize("N", [384, 512])
def test_fp32_wide_tiles_track_fp64_gw(N):
    """The FP32 z tiles of the larger lattices: 16 columns with 768-thread accumulation blocks at
    N = 384, 16 columns with 1024-thread blocks (no register pipelining) from N = 512.  The FP32C
    and FP32 GW spectra after 6 steps from the same device-made field stay within FP32 transform
    rounding of FP64's.  Bins holding more than 1e-6 of the peak agree to 1e-4 each; the rest (a
    few steps from a smooth field leave the highest bins ~1e-20 of the peak) sit at the FP32
    noise floor, below 1e-6 of the peak."""
    say("fp32 wide tiles", N)
    out = {}
    for prec in (capi.FP64, capi.FP32C, capi.FP32):
        prm, _ = params(N, precision=prec, num_bin_grav=64)
        with capi.Lattice(prm) as L:
            L.init_flat(5, 4.0)
            L.advance_steps(1.0, 0.05, 6)
            out[prec] = L.gw_spectrum()
    ref = out[capi.FP64]
    top = ref["numer"].max()
    big = ref["numer"] > 1e-6 * top
    assert big.sum() > 30
    for prec in (capi.FP32C, capi.FP32):
        got = out[prec]["numer"]
        err = np.abs(got[big] - ref["numer"][big]) / ref["numer"][big]
        assert err.max() < 1e-4, (prec, err.max())
        assert np.abs(got - ref["numer"]).max() < 1e-6 * top
        assert np.array_equal(out[prec]["denom"], ref["denom"])


@pytest.mark.parametrize("N", [768, 1024])
def test_largest_lattices_fp32c_tracks_fp64(N):
    """The largest compile-time sizes: 768^3 (3 * 2^8: radix-3 finish, sub-strips, 8-column FP32
    and 4-column FP64 tiles) and 1024^3 (the largest single-GPU lattice; its FP64 handle holds
    about 213 GB), checked through properties that do not depend on the size.  From the same
    device-made field, the FP32C GW spectrum after 4 steps stays within FP32 transform rounding of
    FP64's (bins above 1e-6 of the peak to 1e-4 each, the rest below 1e-6 of the peak, bin counts
    identical).  In each precision, 4 steps forward and 4 steps back return the energy terms to
    round-off (the integrator is time-reversible; tolerances as the 32^3 field-reversal test)."""
    out = {}
    for prec, tol in ((capi.FP64, 1e-11), (capi.FP32C, 1e-8)):
        say("largest lattices", N, prec)
        prm, _ = params(N, precision=prec, num_bin_grav=64)
        with capi.Lattice(prm) as L:
            L.init_flat(5, 4.0)
            e0 = L.energy(1.0)
            L.advance_steps(1.0, 0.05, 4)
            out[prec] = L.gw_spectrum()
            e_mid = L.energy(1.2)
            L.advance_steps(1.2, -0.05, 4)
            e1 = L.energy(1.0)
        scale = np.abs(e0).max()
        assert np.abs(e_mid - e0).max() > 1e-6 * scale  # the fields did move
        assert np.abs(e1 - e0).max() <= tol * scale, (prec, e0, e1)
    ref, got = out[capi.FP64], out[capi.FP32C]
    top = ref["numer"].max()
    big = ref["numer"] > 1e-6 * top
    assert big.sum() > 30
    err = np.abs(got["numer"][big] - ref["numer"][big]) / ref["numer"][big]
    assert err.max() < 1e-4, err.max()
    assert np.abs(got["numer"] - ref["numer"]).max() < 1e-6 * top
    assert np.array_equal(got["denom"], ref["denom"])


def test_strip_orders_bit_identical_at_768(monkeypatch):
    """At N = 768 the XCD strips split into sub-strips (strip_order): the fused pass in the
    dispatcher's order (0), in strips (8) and in strips with wide blocks (9) gives the same
    energy terms and GW spectrum bit for bit (every site is updated exactly once)."""
    say("strip orders 768")
    N = 768
    prm, _ = params(N, precision=capi.FP32C, num_bin_grav=32)
    out = []
    for v in ("0", "8", "9"):
        monkeypatch.setenv("AXG_KICK_VARIANT", v)
        with capi.Lattice(prm) as L:
            L.init_flat(3, 2.0)
            L.advance_steps(1.0, 0.05, 2)
            out.append((L.energy(1.1), L.gw_spectrum()["numer"]))
    for e, g in out[1:]:
        assert np.array_equal(e, out[0][0]) and np.array_equal(g, out[0][1])


def test_advance_get_dtau_interval():
    say("advance interval")
    N = 16
    prm, P = params(N, eps_dtau=1.0)
    dt, n = capi.get_dtau(prm, 1.0, 2.0)
    odt, on = O.get_dtau(1.0, 2.0, P["sigma_cr"], P["dwav"], N, 1.0)
    assert (dt, n) == (odt, on)
    phi, pi = ic(N, 2, 4.0)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        assert L.advance(1.0, 2.0) == n
        g_phi, g_pi = L.get_fields()
        A = L.gw_state()
    OA = O.new_gw_state(N)
    O.symplectic(phi, pi, OA, 1.0, dt, n, P)
    assert np.array_equal(g_phi, phi) and np.array_equal(g_pi, pi)
    assert relerr(A[0], OA[0][:, :, : N // 2 + 1]) < 1e-10


# ---------------------------------------------------------------- plaquette windings

@pytest.mark.parametrize("N,prec,kmax", [(16, capi.FP64, 6.0), (32, capi.FP64, 6.0), (24, capi.FP32C, 4.0),
                                         (48, capi.FP64, 2.0)])
def test_plaquettes_bitexact(N, prec, kmax):
    """Face codes, counts and the compacted list of pierced faces equal the oracle's on the field
    the GPU holds (integer work: bit-exact)."""
    say("plaquettes", N, prec)
    prm, P = params(N, grav_spec=False, precision=prec)
    phi, pi = ic(N, 19, kmax)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        f_phi, _ = L.get_fields()
        with pytest.raises(capi.AxgError):
            L.plaquette_records()  # no plaquettes yet
        cnt = L.plaquettes()
        rec = L.plaquette_records()
    code, ocnt, orec = O.plaquettes(f_phi)
    for k in ocnt:
        assert np.array_equal(cnt[k], ocnt[k]), k
    assert cnt["n_pierced"].sum() > 0 and np.all(cnt["net"] == 0)
    assert np.array_equal(rec, orec)


# ---------------------------------------------------------------- the reference's random stream

@pytest.mark.parametrize("N,prec", [(32, capi.FP64), (48, capi.FP64), (32, capi.FP32C)])
def test_thermal_state_from_reference_stream(N, prec):
    """axg_set_stream: the thermal state's Gaussians come from the reference's MCG59 stream (draws
    pos + 4 N^3 (ip - 1) + 4 lin, initial.f90:102-126) and the stream advances by 8 N^3.  Against
    the oracle on the same stream: <= 1e-12 in FP64 (device vs host ln/sin/exp), FP32C storage
    rounding otherwise."""
    say("thermal stream", N, prec)
    prm, P = params(N, grav_spec=False, initial_loop=False, precision=prec)
    with capi.Lattice(prm) as L:
        L.set_stream(12345, 678, 1000)
        L.init_thermal(0, 0.2)
        phi, pi = L.get_fields()
        pos = L.stream_position()
    assert pos == 1000 + 8 * N ** 3
    ophi, opi = O.ic_thermal(N, 0, 0.2, P["sigma_cr"], P["dwav"], P["dv_wav"], stream=(12345, 678, 1000))
    tol = 1e-12 if prec == capi.FP64 else 1e-10
    assert relerr(phi, ophi) < tol and relerr(pi, opi) < tol
    # not the Philox state
    xphi, _ = O.ic_thermal(N, 0, 0.2, P["sigma_cr"], P["dwav"], P["dv_wav"])
    assert relerr(phi, xphi) > 0.1


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32C])
def test_strings_and_covariance_from_reference_stream(prec):
    """A snapshot's draws in the reference's order from one stream: the string fits take 4004
    draws per string cell (string s at position + 4004 s, defects_3d.f90:137-162, 465), then the
    covariance 4 N^3 per realisation (spectrum.f90:592-630).  Records and sums as the oracle's on
    the same stream positions (bars of the Philox tests), the covariance with the same number of
    realisations within 1e-6, and the position advanced by exactly what was drawn."""
    say("strings + covariance stream", prec)
    N, nb = 32, 12
    prm, P = params(N, grav_spec=False, precision=prec, id_strings=True, id_walls=True, spec_est=True, num_bin=nb)
    phi, _ = ic(N, 31, 2.0)  # a mask of ~25 % of the cells keeps every deconvolved bin positive
    rng = np.random.default_rng(4)
    pi = 0.3 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    p0 = 8 * N ** 3
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        f_phi, f_pi = L.get_fields()
        L.set_stream(7, 9, p0)
        cnt = L.defects()
        out = L.strings(0, 0)
        rec = L.string_records()
        p1 = L.stream_position()
        s = L.field_spectrum(3.0)
        cov, it = L.field_covariance(0, 4, 12, 0.0)
        p2 = L.stream_position()
    ns = int(cnt[0])
    assert ns > 50 and p1 == p0 + 4004 * ns
    assert it == -13 and p2 == p1 + 4 * N ** 3 * 12
    oflags, _ = O.classify(f_phi)
    orec, osums = O.strings(f_phi, f_pi, oflags, 0, 0, P["dspa"], stream=(7, 9, p0))
    rel = np.abs(rec - orec) / np.maximum(1, np.abs(orec))
    assert rel.max() < 1e-5
    assert abs(out["length"] - osums[0]) <= 1e-9 * osums[0] and out["n_vel"] == osums[1]
    prec_rec, _ = O.strings(f_phi, f_pi, oflags, 0, 0, P["dspa"])
    assert np.abs(prec_rec - orec).max() > 1e-3  # the Philox starts give other fits
    tab = O.init_spectrum(nb, P["dspa"])
    pw, num = O.conv_spectrum(f_pi, oflags, nb, P["dspa"], P["dwav"], P["dv_spa"], P["dv_wav"])
    _, inv, info = O.matrix_deconv(oflags, nb, P["dwav"], P["dv_spa"], P["dv_wav"], tab)
    M = O.sym_from_packed(inv, nb) / np.outer(num, num)
    ocov, oit, _ = O.field_covariance(oflags, nb, P["dspa"], P["dwav"], s["power"], packed(M), 0, 4, 12, 0.0,
                                      stream=(7, 9, p1))
    assert oit == it and relerr(cov, ocov) < 1e-6
