"""Batches of lattices (axg_batch_advance): lattices of different sizes and precisions advanced
together from one captured HIP graph end bit-identical to the same lattices advanced one by
one with axg_advance_steps (the same kernels; the step scalars come from device tables)."""
import numpy as np
import pytest

import oracle as O
from axion_grav_amd import capi
from tests.conftest import SQ_HALF, phys

pytestmark = pytest.mark.gpu

CASES = [(16, capi.FP64), (24, capi.FP32C), (32, capi.FP32C), (64, capi.FP32), (32, capi.FP64), (36, capi.FP32C),
         (48, capi.FP64), (32, capi.FP32C), (64, capi.FP32)]


def make(N, prec, seed):
    P = phys(N)
    prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=prec, num_bin_grav=16,
                           potential_kappa0=1e-3 if seed % 2 else 0.0)
    phi, _ = O.ic_flat(N, seed, 4.0, P["dwav"])
    rng = np.random.default_rng(seed)
    pi = 0.2 * (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N)))
    return prm, phi, pi


@pytest.mark.parametrize("nsteps,chunk,lanes,stack", [(7, 2, None, "1"), (8, 4, None, "1"), (5, 0, None, "1"),
                                                      (3, 4, None, "1"), (6, 2, "1", "1"), (6, 2, "3", "1"),
                                                      (4, 4, "64", "1"), (7, 2, None, "0"), (5, 0, None, "0")])
def test_batch_matches_single_lattices(nsteps, chunk, lanes, stack, monkeypatch):
    """Odd step counts leave a tail of direct steps after the graph replays; chunk > nsteps runs
    no graph at all; chunk = 0 launches every step directly.  AXG_BATCH_LANES deals the units
    to 1, 3 or (clamped) all streams in the graph instead of the default two.  The lattices of
    the same size and precision (two 32^3 FP32C, with and without the tilt) form a stack, the
    others step alone; AXG_BATCH_STACK=0 steps every lattice alone."""
    print(f"[case] batch nsteps={nsteps} chunk={chunk} lanes={lanes} stack={stack}", flush=True)
    monkeypatch.setenv("AXG_BATCH_STACK", stack)
    if lanes is not None:
        monkeypatch.setenv("AXG_BATCH_LANES", lanes)
    t1, dt = 1.0, 0.05
    lats, ref = [], []
    for i, (N, prec) in enumerate(CASES):
        prm, phi, pi = make(N, prec, 10 + i)
        L = capi.Lattice(prm)
        L.set_fields(phi, pi)
        lats.append(L)
        with capi.Lattice(prm) as R:
            R.set_fields(phi, pi)
            R.advance_steps(t1, dt, nsteps)
            ref.append((R.get_fields(), R.gw_state()))
    try:
        ms = capi.batch_advance(lats, t1, dt, nsteps, chunk)
        assert ms > 0
        for L, ((rphi, rpi), rA) in zip(lats, ref):
            phi, pi = L.get_fields()
            assert np.array_equal(phi, rphi) and np.array_equal(pi, rpi)
            assert np.array_equal(L.gw_state(), rA)
        # a second call continues from the first (its tables are rewritten, the counter restarts)
        capi.batch_advance(lats, t1 + nsteps * dt, dt, 2, 2)
        assert all(np.all(np.isfinite(L.get_fields()[0])) for L in lats)
    finally:
        for L in lats:
            L.close()


def test_batch_rejects_bad_arguments():
    prm, phi, pi = make(16, capi.FP64, 1)
    with capi.Lattice(prm) as L:
        L.set_fields(phi, pi)
        with pytest.raises(capi.AxgError):
            capi.batch_advance([L], 1.0, 0.05, 4, 3)  # odd chunk
        with pytest.raises(capi.AxgError):
            capi.batch_advance([L, L], 1.0, 0.05, 4, 2)  # the same lattice twice


def test_batch_mixed_workload_sizes():
    """The mixed-size workload's lattice sizes in one captured graph: two each of 64^3, 96^3 and
    128^3, FP32C with GW (the kernels the batch benchmark times), bit-identical to the same
    lattices advanced one by one, and their size-grouped mean GW spectra equal."""
    print("[case] batch 64/96/128", flush=True)
    t1, dt, nsteps = 1.0, 0.025, 7
    lats, ref = [], []
    try:
        for i, N in enumerate((64, 96, 128, 64, 96, 128)):
            P = phys(N)
            prm = capi.make_params(N, P["dspa"], SQ_HALF, precision=capi.FP32C, num_bin_grav=64)
            L = capi.Lattice(prm)
            L.init_flat(i, 4.0)
            lats.append(L)
            phi, pi = L.get_fields()
            with capi.Lattice(prm) as R:
                R.set_fields(phi, pi)
                R.advance_steps(t1, dt, nsteps)
                ref.append((R.get_fields(), R.gw_state(), R.gw_spectrum()["power"]))
        capi.batch_advance(lats, t1, dt, nsteps, 2)
        spec = {}
        for L, ((rphi, rpi), rA, rpow) in zip(lats, ref):
            phi, pi = L.get_fields()
            assert np.array_equal(phi, rphi) and np.array_equal(pi, rpi)
            assert np.array_equal(L.gw_state(), rA)
            spec.setdefault(L.N, []).append(L.gw_spectrum()["power"])
        for N, s in spec.items():
            r = [p for (_, _, p), L in zip(ref, lats) if L.N == N]
            assert np.array_equal(np.mean(s, axis=0), np.mean(r, axis=0))
    finally:
        for L in lats:
            L.close()
