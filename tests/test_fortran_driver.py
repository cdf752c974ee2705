"""The Fortran host (axion_grav_amd/bin/axion_3d) keeps the reference's command line, ini keys,
checkpoint schedule and output files, and drives the GPU through the ISO_C_BINDING module.
CPU tests use dry_run (parameters only); the GPU test compares every output file of a short
run with the CPU oracle and checks resume from a checkpoint (evolve.f90:439-546)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests.conftest import ROOT

EXE = os.path.join(ROOT, "axion_grav_amd", "bin", "axion_3d")
PI_T = 3.14159265359
SIGMA = float(1 / (6 * PI_T * np.sqrt(0.5)))  # sigma_cr^2 = 1/2 at g_rel = 90


def ini_text(rootdir, **over):
    keys = dict(dry_run="F", rootdir=rootdir, rootname="run", ss="1 2", ngrid=32, n_chk=2, num_bin=8,
                initial_loop="T", loop_radius=0.5, loop_velocity=0.0, loop_velocity_in=0.0,
                id_strings="T", group_strings="F", out_strings="F", id_walls="T", group_walls="F",
                out_walls="F", spec_est="T", grav_spec="T", n_id=2, n_mod_eps=1, degrade_spec=0,
                sigma=repr(SIGMA), g_rel=90.0, kappa="1d-30", tau_ini=1.0, tau_fin=2.0, eps_dtau=1.0,
                boxsize=4.0, ic_seed=77, ic_kmax=6.0)
    keys.update(over)
    lines = ["# test run", "COMMENT this line is ignored = 1"] + [f"{k} = {v}" for k, v in keys.items()]
    lines += ["ngrid = 64", "END", "ngrid = 16"]  # first occurrence wins; END stops reading
    return "\n".join(lines) + "\n"


def run(tmp_path, text, *args, check=True):
    p = tmp_path / "p.ini"
    p.write_text(text)
    r = subprocess.run([EXE, str(p), *args], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    if check and r.returncode != 0:
        raise AssertionError(r.stdout + r.stderr)
    return r


needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="axion_3d not built (no amdflang)")


@needs_exe
def test_dry_run_parses_keys_and_derives(tmp_path):
    r = run(tmp_path, ini_text(str(tmp_path), dry_run="T"), "-b", "100")
    assert "dry run" in r.stdout
    assert " ngrid = 32" in r.stdout  # the first 'ngrid' wins, 'END' stops reading
    o = O.derive(SIGMA, 90.0, 1e-30, 2.0, 4.0, 32)
    line = [l for l in r.stdout.splitlines() if "delta_lattice" in l][0]
    assert float(line.split("=")[1]) == o["dspa"]
    line = [l for l in r.stdout.splitlines() if "sigma_cr" in l][0]
    assert abs(float(line.split("=")[1]) - o["sigma_cr"]) < 1e-13


@needs_exe
@pytest.mark.parametrize("key,val,msg", [("dry_run", "TRUE", "error reading logical"),
                                         ("ngrid", "3.2e1", "error reading integer")])
def test_bad_values_stop(tmp_path, key, val, msg):
    over = dict(dry_run="T")
    over[key] = val
    r = run(tmp_path, ini_text(str(tmp_path), **over), "-b", "100", check=False)
    assert r.returncode != 0 and msg in r.stdout + r.stderr


@needs_exe
def test_missing_key_and_memory_check(tmp_path):
    text = ini_text(str(tmp_path), dry_run="T").replace("sigma = ", "sigmaX = ")
    r = run(tmp_path, text, "-b", "100", check=False)
    assert r.returncode != 0 and "key not found : sigma" in r.stdout
    r = run(tmp_path, ini_text(str(tmp_path)), check=False)  # no -b: settings.f90:260
    assert r.returncode != 0 and "not enough memory" in r.stdout + r.stderr


def read_rows(path):
    return np.array([[float(x) for x in l.split()] for l in open(path) if l.strip()])


def oracle_run(N, P, phi0, pi0, taus, eps):
    """The same schedule with the oracle: from (phi0, pi0), get_dtau + DKD per interval."""
    phi, pi = phi0.copy(), pi0.copy()
    A = O.new_gw_state(N)
    states = {}
    for idx in range(1, len(taus)):
        dt, n = O.get_dtau(taus[idx - 1], taus[idx], P["sigma_cr"], P["dwav"], N, eps)
        O.symplectic(phi, pi, A, taus[idx - 1], dt, n, P)
        states[idx] = (phi.copy(), pi.copy(), A.copy())
    return states


@needs_exe
@pytest.mark.gpu
def test_run_outputs_match_oracle_and_resume(tmp_path):
    print("[case] fortran driver N=32", flush=True)
    N, nb, nbg = 32, 8, 25
    d = O.derive(SIGMA, 90.0, 1e-30, 2.0, 4.0, N)
    # the oracle's seeded field enters through ic_file, so the fields can be compared bitwise.
    # A smooth phi (few defects, a mask of 10-20% of the cells) and a broadband pi keep the
    # deconvolved power positive in every bin at both snapshots, so the covariance is finite
    # (a negative bin makes estimate_covmat's sqrt(power/2), and so the covariance, NaN).
    phi0, _ = O.ic_flat(N, 77, 1.0, d["dwav"])
    rng = np.random.default_rng(5)
    pi0 = 2.0 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    np.concatenate([phi0.ravel(), pi0.ravel()]).tofile(tmp_path / "ic.raw")
    run(tmp_path, ini_text(str(tmp_path), ic_file=str(tmp_path / "ic.raw")), "-b", "100")
    P = dict(d, kappa=1e-30, initial_loop=True)
    taus = [1.0, 1.5, 2.0]
    states = oracle_run(N, P, phi0, pi0, taus, 1.0)
    # the reference's stream seeded by ss = "1 2" (an ic_file start draws nothing): at each snapshot
    # the string fits take 4004 draws per string cell, then the covariance 4 N^3 per realisation
    pos = 0
    str_pos = {}
    for idx in (1, 2):
        phi, pi, A = states[idx]
        # gravspec: tau_id, q, q f_cr, power, denom (grav_waves.f90:313-314)
        g = read_rows(tmp_path / f"run_gravspec_{idx}.dat")
        o = O.gw_spectrum(A, nbg, d["dspa"], d["dwav"], d["tau_cr"], d["a_now"], d["dv_wav"])
        assert g.shape == (nbg, 5)
        assert np.allclose(g[:, 0], taus[idx], rtol=1e-10)
        assert np.allclose(g[:, 1], o["qmid"], rtol=1e-9)
        assert np.array_equal(g[:, 4], o["denom"])
        assert np.allclose(g[:, 3], o["power"], rtol=1e-8, atol=0)
        # mask list (defects_3d.f90:265-273), bit-exact
        flags, counts = O.classify(phi)
        ml = read_rows(tmp_path / f"run_mask_{idx}-0.dat").astype(int) if counts[2] else np.zeros((0, 3))
        zz, yy, xx = np.nonzero(flags & 4)
        assert np.array_equal(ml, np.stack([xx + 1, yy + 1, zz + 1], axis=1))
        # spectrum: qmid, power, conv/num, num (spectrum.f90:282-283)
        s = read_rows(tmp_path / f"run_spectrum_{idx}.dat")
        pw, num = O.conv_spectrum(pi, flags, nb, d["dspa"], d["dwav"], d["dv_spa"], d["dv_wav"])
        conv = pw * taus[idx] ** 2
        tab = O.init_spectrum(nb, d["dspa"])
        _, inv, info = O.matrix_deconv(flags, nb, d["dwav"], d["dv_spa"], d["dv_wav"], tab)
        power = (O.sym_from_packed(inv, nb) / np.outer(num, num)) @ conv
        assert np.array_equal(s[:, 3], num)
        assert np.allclose(s[:, 2], conv / num, rtol=1e-8)
        assert np.allclose(s[:, 1], power, rtol=1e-6, atol=1e-12 * np.abs(power).max())
        # covariance (spectrum.f90:295-304, estimate_covmat 569-734): pairs (i, j <= i); the
        # MC generator is keyed by ss = "1 2" and idx
        c = read_rows(tmp_path / f"run_covariance_{idx}.dat")
        ii, jj = np.tril_indices(nb)
        order = np.lexsort((jj, ii))
        assert c.shape == (nb * (nb + 1) // 2, 3)
        assert np.array_equal(c[:, 0], ii[order] + 1) and np.array_equal(c[:, 1], jj[order] + 1)
        matn = [(O.sym_from_packed(inv, nb) / np.outer(num, num))[j, i] for i, j in zip(ii[order], jj[order])]
        str_pos[idx] = pos
        pos += 4004 * int(counts[0])
        ocov, oit, _ = O.field_covariance(flags, nb, d["dspa"], d["dwav"], power, np.array(matn), 0, 10, 100, 0.1,
                                          stream=(1, 2, pos))
        pos += 4 * N ** 3 * (oit if oit > 0 else 100)
        assert power.min() > 0 and np.all(np.isfinite(ocov))
        assert np.allclose(c[:, 2], ocov, rtol=1e-5, atol=1e-6 * np.abs(ocov).max())
    # get_epsilon (spectrum.f90:956-1040) from the files: eps = f sum d q^2 / sum d q with
    # d the differential spectrum, and its delta-method error sqrt(w^T (C1 + C2) w) f/|den|,
    # w = q^2 - eps q / f; delta crho = sqrt(a^T C a) / (2 pi^2), a = q^2 dq, in quadrature
    e = read_rows(tmp_path / "run_eps.txt")
    assert e.shape == (1, 9) and list(e[0, :2]) == [1, 2]
    edges = np.zeros(nb + 1)
    for b in range(1, nb + 1):
        edges[b] = edges[b - 1] + PI_T / d["dspa"] / nb
    q, dq = (edges[1:] + edges[:-1]) / 2, np.diff(edges)
    spec = {i: read_rows(tmp_path / f"run_spectrum_{i}.dat")[:, 1] for i in (1, 2)}
    C = {}
    for i in (1, 2):
        rows = read_rows(tmp_path / f"run_covariance_{i}.dat")
        M = np.zeros((nb, nb))
        M[rows[:, 0].astype(int) - 1, rows[:, 1].astype(int) - 1] = rows[:, 2]
        C[i] = M + np.tril(M, -1).T
    tau = (taus[1] + taus[2]) / 2
    f = tau / 2 / (2 * PI_T)
    dd = spec[2] - spec[1]
    den = (dd * q).sum()
    eps = (dd * q * q).sum() / den * f
    w = q * q - eps * q / f
    deps = np.sqrt(w @ (C[1] + C[2]) @ w / den ** 2) * f
    a = q * q * dq
    drho = {i: np.sqrt(a @ C[i] @ a) / (2 * PI_T * PI_T) for i in (1, 2)}
    cr = read_rows(tmp_path / "run_crho.txt")
    assert np.allclose(e[0, 2:5], [taus[1], taus[2], tau], rtol=1e-9)
    assert np.allclose(e[0, 5:7], [eps, deps], rtol=1e-7)
    assert np.allclose(e[0, 7], cr[1, 3] - cr[0, 3], rtol=1e-6, atol=1e-9 * abs(cr[1, 3]))
    assert np.allclose(e[0, 8], np.hypot(drho[1], drho[2]), rtol=1e-7)
    sw = read_rows(tmp_path / "run_string_wall_parameters.txt")
    assert sw.shape == (2, 12) and np.allclose(sw[:, 1], taus[1:])
    # string parameter (evolve.f90:191) and velocities from the fits' draws of the stream
    for idx in (1, 2):
        phi, pi, _ = states[idx]
        flags, _ = O.classify(phi)
        _, sums = O.strings(phi, pi, flags, 0, 0, d["dspa"], stream=(1, 2, str_pos[idx]))
        xi = (taus[idx] / d["dspa"]) ** 2 * sums[0] / N ** 3 / 4
        assert np.allclose(sw[idx - 1, 3:6], xi, rtol=0.02), (sw[idx - 1, 3:6], xi)
        assert np.allclose(sw[idx - 1, 9:12], sums[2:5] / sums[1], rtol=0.02)
        # wall parameter (evolve.f90:192), no random numbers
        _, wsums = O.walls(phi, flags)
        largea = (taus[idx] / d["dspa"]) * wsums[0] / N ** 3 / 2
        assert np.allclose(sw[idx - 1, 6:9], largea, rtol=1e-9), (sw[idx - 1, 6:9], largea)
    # the final field checkpoint and the resume file (evolve.f90:401-437)
    idx, tau, flag = open(tmp_path / "run_param.chk").read().split()
    assert int(idx) == 2 and abs(float(tau) - 2.0) < 1e-12 and flag == "F"
    raw = np.fromfile(tmp_path / "run_phi_2.chk", dtype=np.complex128)
    phi2, pi2, _ = states[2]
    assert np.array_equal(raw[: N ** 3].reshape(N, N, N), phi2)
    assert np.array_equal(raw[N ** 3:].reshape(N, N, N), pi2)
    assert not (tmp_path / "run_phi_0.chk").exists()

    # resume from idx 1: the field comes back, the GW accumulators only with chk_gw
    for chk_gw in ("F", "T"):
        print(f"[case] fortran resume chk_gw={chk_gw}", flush=True)
        rd = tmp_path / f"resume_{chk_gw}"
        rd.mkdir()
        phi1, pi1, A1 = states[1]
        np.concatenate([phi1.ravel(), pi1.ravel()]).tofile(rd / "run_phi_1.chk")
        if chk_gw == "T":
            np.ascontiguousarray(A1[:, :, :, : N // 2 + 1]).tofile(rd / "run_grav_1.chk")
        (rd / "run_param.chk").write_text("1 1.5 T\n")
        r = run(rd, ini_text(str(rd), chk_gw=chk_gw), "-b", "100")
        assert "chk files EXIST" in r.stdout
        g = read_rows(rd / "run_gravspec_2.dat")
        if chk_gw == "T":
            A2 = states[2][2]
        else:  # the reference restarts arr_grav from zero after a resume
            phi, pi, A2 = phi1.copy(), pi1.copy(), O.new_gw_state(N)
            dt, n = O.get_dtau(1.5, 2.0, P["sigma_cr"], P["dwav"], N, 1.0)
            O.symplectic(phi, pi, A2, 1.5, dt, n, P)
        o = O.gw_spectrum(A2, nbg, d["dspa"], d["dwav"], d["tau_cr"], d["a_now"], d["dv_wav"])
        assert np.allclose(g[:, 3], o["power"], rtol=1e-8, atol=0)
        raw = np.fromfile(rd / "run_phi_2.chk", dtype=np.complex128)
        assert np.array_equal(raw[: N ** 3].reshape(N, N, N), phi2)


@needs_exe
@pytest.mark.gpu
def test_tuned_run_matches_oracle(tmp_path):
    """At N = 64 the driver tunes the step kernels (axg_tune) before its first interval; the GW
    spectrum it writes still matches the oracle's (the tuner's trial launches leave no trace)."""
    print("[case] fortran driver tuned N=64", flush=True)
    N, nbg = 64, 25
    d = O.derive(SIGMA, 90.0, 1e-30, 1.5, 8.0, N)
    phi0, _ = O.ic_flat(N, 78, 1.0, d["dwav"])
    rng = np.random.default_rng(6)
    pi0 = 0.5 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    np.concatenate([phi0.ravel(), pi0.ravel()]).tofile(tmp_path / "ic.raw")
    run(tmp_path, ini_text(str(tmp_path), ic_file=str(tmp_path / "ic.raw"), ngrid=N, n_chk=1, tau_fin=1.5,
                           boxsize=8.0, spec_est="F", id_strings="F", id_walls="F", n_id=1), "-b", "100")
    P = dict(d, kappa=1e-30, initial_loop=True)
    _, _, A = oracle_run(N, P, phi0, pi0, [1.0, 1.5], 1.0)[1]
    g = read_rows(tmp_path / "run_gravspec_1.dat")
    o = O.gw_spectrum(A, nbg, d["dspa"], d["dwav"], d["tau_cr"], d["a_now"], d["dv_wav"])
    assert np.array_equal(g[:, 4], o["denom"])
    assert np.allclose(g[:, 3], o["power"], rtol=1e-8, atol=0)


@needs_exe
@pytest.mark.gpu
def test_grouped_networks_and_defect_files(tmp_path):
    """group_strings = group_walls = T with out_strings = out_walls = T: the string and wall
    parameter columns hold the grouped sums (all, long strings / large walls, longer still) of
    the oracle's grouping of the oracle's fits with the driver's thresholds, and
    _string_<idx>.dat / _wall_<idx>.dat list every fit with its group number."""
    print("[case] fortran driver grouping N=32", flush=True)
    N = 32
    d = O.derive(SIGMA, 90.0, 1e-30, 2.0, 4.0, N)
    phi0, _ = O.ic_flat(N, 77, 1.0, d["dwav"])
    rng = np.random.default_rng(5)
    pi0 = 2.0 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    np.concatenate([phi0.ravel(), pi0.ravel()]).tofile(tmp_path / "ic.raw")
    r = run(tmp_path, ini_text(str(tmp_path), ic_file=str(tmp_path / "ic.raw"), group_strings="T", group_walls="T",
                               out_strings="T", out_walls="T", spec_est="F", grav_spec="F"), "-b", "100")
    assert "threshold for string grouping" in r.stdout and "threshold for wall grouping" in r.stdout
    P = dict(d, kappa=1e-30, initial_loop=True)
    taus = [1.0, 1.5, 2.0]
    states = oracle_run(N, P, phi0, pi0, taus, 1.0)
    sw = read_rows(tmp_path / "run_string_wall_parameters.txt")
    pos = 0  # the reference's stream seeded by ss: 4004 draws per string cell, no covariance
    for idx in (1, 2):
        phi, pi, _ = states[idx]
        tau = taus[idx]
        flags, counts = O.classify(phi)
        srec, _ = O.strings(phi, pi, flags, 0, 0, d["dspa"], stream=(1, 2, pos))
        pos += 4004 * int(counts[0])
        wrec, _ = O.walls(phi, flags)
        hor = tau / d["dspa"]
        th_s = O.group_threshold(d["width_cr"] / d["dspa"] / tau, hor)
        th_w = O.group_threshold(d["thickness_cr"] / d["dspa"] / tau, hor)
        ng_s, ids_s, s_s = O.group_sums(O.group_points(srec, N, th_s), srec[:, 7], 2 * hor, 2 * PI_T * hor)
        ng_w, ids_w, s_w = O.group_sums(O.group_points(wrec, N, th_w), wrec[:, 7], PI_T * hor ** 2,
                                        4 * PI_T * hor ** 2)
        # kappa ~ 0 makes the wall thickness, and so the wall threshold, astronomically large:
        # one network of walls; the strings split into networks at idx 2
        assert th_w > 1e30 and ng_w == 1 and ng_s == idx
        xi = (tau / d["dspa"]) ** 2 * s_s / N ** 3 / 4
        largea = (tau / d["dspa"]) * s_w / N ** 3 / 2
        assert np.allclose(sw[idx - 1, 3:6], xi, rtol=1e-6), (sw[idx - 1, 3:6], xi)
        assert np.allclose(sw[idx - 1, 6:9], largea, rtol=1e-9), (sw[idx - 1, 6:9], largea)
        # the per-defect files: the same fits to the printed 3 decimals, the same group numbers
        fs = read_rows(tmp_path / f"run_string_{idx}.dat")
        fw = read_rows(tmp_path / f"run_wall_{idx}.dat")
        assert fs.shape == (counts[0], 13) and fw.shape == (counts[1], 10)
        assert np.allclose(fs[:, 0:6], srec[:, 0:6], atol=1.5e-3) and np.allclose(fs[:, 9:11], srec[:, 6:8], atol=1.5e-3)
        assert np.allclose(fw[:, 0:8], wrec[:, 0:8], atol=1.5e-3)
        assert np.array_equal(fs[:, 11], ids_s) and np.array_equal(fs[:, 12], ids_s)
        assert np.array_equal(fw[:, 8], ids_w) and np.array_equal(fw[:, 9], ids_w)


HALO_EXE = os.path.join(ROOT, "axion_grav_amd", "bin", "halo_binding_check")


@pytest.mark.skipif(not os.path.exists(HALO_EXE), reason="halo_binding_check not built (no amdflang)")
@pytest.mark.gpu
def test_halo_padded_binding_roundtrip():
    """The reference's data1/data2 layout (ngrid x ngrid x (ngrid+2), halo planes 1 and ngrid+2,
    initial.f90:240-285) through the binding INTEGRATION.md shows: the interior sections go in and
    come back bit for bit, the halo planes are never touched, and a domain wall placed between
    interior planes ngrid/2 and ngrid/2+1 is flagged on exactly those cells (a one-plane shift of
    the fields would move it)."""
    print("[case] halo-padded Fortran binding", flush=True)
    r = subprocess.run([HALO_EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "halo binding ok" in r.stdout, r.stdout + r.stderr


@needs_exe
@pytest.mark.gpu
def test_degraded_run_matches_oracle(tmp_path):
    """degrade_spec = 1: the gravspec, spectrum and covariance files come from the 16^3 grid of
    the 32^3 lattice (grav_waves.f90:37-45, 358-507; spectrum.f90:42-49, 736-954), the GW
    checkpoint holds that grid's accumulators, and the field checkpoint the lattice's.  A coarse
    cell is masked when any of its 8 lattice cells is, so the reduced grid's mask is about twice
    the lattice's: at the first snapshot one deconvolved bin is negative and the covariance NaN on
    both sides (as in the reference: sqrt of a negative power), at the second all are finite."""
    print("[case] fortran driver N=32 degrade_spec=1", flush=True)
    N, nb, nbg, ds = 32, 8, 25, 1
    Ng = N >> ds
    d = O.derive(SIGMA, 90.0, 1e-30, 2.0, 4.0, N)
    phi0, _ = O.ic_flat(N, 5, 1.0, d["dwav"])
    rng = np.random.default_rng(5)
    pi0 = 2.0 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    np.concatenate([phi0.ravel(), pi0.ravel()]).tofile(tmp_path / "ic.raw")
    run(tmp_path, ini_text(str(tmp_path), ic_file=str(tmp_path / "ic.raw"), degrade_spec=ds, chk_gw="T"),
        "-b", "100")
    P = dict(d, kappa=1e-30, initial_loop=True, degrade_spec=ds, N=N)
    u = O.reduced_units(P, ds)
    phi, pi = phi0.copy(), pi0.copy()
    A = O.new_gw_state(Ng)
    taus = [1.0, 1.5, 2.0]
    pos = 0
    for idx in (1, 2):
        dt, n = O.get_dtau(taus[idx - 1], taus[idx], P["sigma_cr"], P["dwav"], N, 1.0)
        O.symplectic(phi, pi, A, taus[idx - 1], dt, n, P)
        g = read_rows(tmp_path / f"run_gravspec_{idx}.dat")
        o = O.gw_spectrum(A, nbg, u["dspa"], u["dwav"], d["tau_cr"], d["a_now"], u["dv_wav"])
        assert np.allclose(g[:, 1], o["qmid"], rtol=1e-9) and np.array_equal(g[:, 4], o["denom"])
        assert np.allclose(g[:, 3], o["power"], rtol=1e-8, atol=0)
        flags, counts = O.classify(phi)
        cf, cpi = O.degrade_mask(flags, ds), O.degrade_field(pi, ds)
        s = read_rows(tmp_path / f"run_spectrum_{idx}.dat")
        pw, num = O.conv_spectrum(cpi, cf, nb, u["dspa"], u["dwav"], u["dv_spa"], u["dv_wav"])
        conv = pw * taus[idx] ** 2
        tab = O.init_spectrum(nb, u["dspa"])
        _, inv, info = O.matrix_deconv(cf, nb, u["dwav"], u["dv_spa"], u["dv_wav"], tab)
        M = O.sym_from_packed(inv, nb) / np.outer(num, num)
        power = M @ conv
        assert np.array_equal(s[:, 3], num)
        assert np.allclose(s[:, 0], (np.arange(nb) + 0.5) * PI_T / u["dspa"] / nb, rtol=1e-12)
        assert np.allclose(s[:, 1], power, rtol=1e-6, atol=1e-12 * np.abs(power).max())
        c = read_rows(tmp_path / f"run_covariance_{idx}.dat")
        ii, jj = np.tril_indices(nb)
        order = np.lexsort((jj, ii))
        pos += 4004 * int(counts[0])
        ocov, oit, _ = O.field_covariance(cf, nb, u["dspa"], u["dwav"], power,
                                          np.array([M[j, i] for i, j in zip(ii[order], jj[order])]), 0, 10, 100,
                                          0.1, stream=(1, 2, pos))
        pos += 4 * Ng ** 3 * (oit if oit > 0 else 100)
        fin = np.isfinite(ocov)
        assert np.array_equal(np.isfinite(c[:, 2]), fin) and fin.all() == (idx == 2) == (power.min() > 0)
        if idx == 2:
            assert np.allclose(c[:, 2], ocov, rtol=1e-5, atol=1e-6 * np.abs(ocov).max())
    raw = np.fromfile(tmp_path / "run_phi_2.chk", dtype=np.complex128)
    assert np.array_equal(raw[: N ** 3].reshape(N, N, N), phi) and np.array_equal(raw[N ** 3:].reshape(N, N, N), pi)
    ga = np.fromfile(tmp_path / "run_grav_2.chk", dtype=np.complex128).reshape(4, Ng, Ng, Ng // 2 + 1)
    assert np.max(np.abs(ga - A[:, :, :, : Ng // 2 + 1])) <= 1e-10 * np.abs(A).max()
    assert read_rows(tmp_path / "run_eps.txt").shape == (1, 9)
