#!/usr/bin/env python3
"""Throughput of the axion_grav lattice step with GW sourcing on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-fp64]

Workload: a 512^3 periodic lattice, comoving box L = 128 (dspa = 0.25), complex scalar with
lambda = 2 (sigma_cr^2 = 1/2, zero-temperature potential), radiation era from tau = 1 with
dtau = 0.2 dspa, the seeded flat-spectrum initial field (seed 7, |k| < 1), GW accumulation on.
A step is one drift-kick-drift with the six source transforms and the projected accumulation.
The headline value is the FP32-compensated path (FP32 storage + int16 residual for phi and
pi, FP32 transforms and accumulators, FP64 stencil arithmetic); the FP64 path on the same
GPU is reported beside it.  Inputs are resident in HBM before the timed region; the timed
region is exactly K steps between device synchronizations.

Prints ONE JSON line (rank 0).  `roofline` covers the dominant kernel, its achieved
algorithmic bytes per launch over its mean duration measured with HIP events on the stream
the kernels run on; `traffic` is its counter-measured HBM bytes per launch from the committed
512^3 profile.  `cpu_baseline` is the CPU FP64 oracle (the reference dataflow restated in C,
OpenMP) timed on this host on a bounded sample: three steps at 512^3 after one untimed.

Besides the headline line's FP32C and FP64 step rates, the default run also reports (each can be
skipped): the HBM streaming probe (read / write / copy with the library's own kernels), the
diagnostics' costs, the adjoint's cost at 128^3 with its finite-difference check, the
48-realization mixed-size batch, the 1024^3 largest share, and the FP32C-vs-FP64 GW spectrum error
over 2000 steps at 512^3.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the oracle loads)
os.environ.setdefault("OMP_PROC_BIND", "close")
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
TWOPI_T = 2 * 3.14159265359
SIGMA_CR = float(np.sqrt(0.5))

# Algorithmic reals moved per site by one launch of each step kernel.  The model of 40 reals per
# site-update splits as 14 + 12 + 14 over the three passes of a step: [kick + drift + sources +
# x transform], [y transform], [z transform + projected accumulation].  gw_xfft only runs when
# the x transform is not fused (lattice sizes without the compile-time transform).
KERNEL_REALS = {
    "drift": 6,                  # phi r/w, pi read
    "kick_source_drift": 14,     # phi r/w, pi r/w, 6 half-spectrum reals of the x-transformed sources
    "gw_xfft": 12,               # 6 reals in, 6 half-spectrum reals out
    "gw_yfft": 12,
    "gw_zfft_accumulate": 14,    # 6 half-spectrum reals in, 4 accumulators read + written
}
MODEL_REALS = 40

# HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE (calibrated against a copy of known
# size) for the 512^3 workload; the profile summary holds the commands and the calibration.
TRAFFIC_FILE = os.path.join(HERE, "profiles", "lattice_step_512_hbm_traffic.md")
# the fused pass's kernel for each tuner choice (0 rows in dispatch order, 1 wide blocks, 4 / 5 the other
# sites-per-thread choice), FP32C and FP64 at 512^3
def _fused_name(prec, v):
    """Kernel name of the fused pass for tuner code v at 512^3 (see launch_kick_source_xfft)."""
    T, C = ("float", "true") if prec == "fp32c" else ("double", "false")
    two = (prec == "fp32c") != bool(v & 4)
    xr = 2 if v & 1 else 1
    if two:
        return f"kick_source_xfft2_kernel<{T}, {C}, 512, {xr}, {'true' if v & 8 else 'false'}, false>"
    return f"kick_source_xfft_kernel<{T}, {C}, 512, {xr}, {2 if v & 8 else 1 if v & 2 else 0}, false>"


YPASS_KERNEL = {("fp32c", 0): "gw_ypass_p2_kernel<float, 512, 16, 1, false>",
                ("fp32c", 1): "gw_ypass_p2_kernel<float, 512, 8, 1, false>",
                ("fp32c", 2): "gw_ypass_p2_kernel<float, 512, 16, 2, false>",
                ("fp64", 0): "gw_ypass_p2_kernel<double, 512, 8, 1, false>",
                ("fp64", 1): "gw_ypass_p2_kernel<double, 512, 4, 1, false>",
                ("fp64", 2): "gw_ypass_p2_kernel<double, 512, 8, 2, false>"}
ZPASS_KERNEL = {"fp32c": "gw_zpass_p2_kernel<float, 512>", "fp64": "gw_zpass_p2_kernel<double, 512>"}


def measured_traffic(slot, prec, N, choice=0):
    """Counter-measured HBM bytes per launch of a step kernel at 512^3 (the fused pass in the
    tuner's configuration `choice`), or None when the profile does not hold that kernel."""
    if slot == "kick_source_drift":
        name = _fused_name(prec, choice)
    elif slot == "gw_yfft":
        name = YPASS_KERNEL.get((prec, choice))
    else:
        name = ZPASS_KERNEL.get(prec)
    if N != 512 or name is None or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        cols = [c.strip().strip("`") for c in line.split("|")]
        if len(cols) > 8 and cols[1] == name:
            return float(cols[8])
    return None


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def params(capi, N, L, precision, nbins):
    dspa = L / N
    return capi.make_params(N, dspa, SIGMA_CR, precision=precision, grav_spec=True, initial_loop=True,
                            kappa=1e-30, num_bin_grav=nbins, id_strings=True, id_walls=True, spec_est=True)


def time_ms(fn, reps=3):
    """Mean wall time of fn() (each call synchronizes before returning), after one untimed call."""
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    return (time.perf_counter() - t0) / reps * 1e3, out


def run_diagnostics(lat, t):
    """Wall time of each diagnostic on the evolved field (outside the timed step region)."""
    d = {}
    d["gw_spectrum_ms"], _ = time_ms(lat.gw_spectrum)
    d["defects_ms"], cells = time_ms(lat.defects)
    d["strings_ms"], st = time_ms(lambda: lat.strings(1, 2), reps=1)
    d["walls_ms"], wl = time_ms(lat.walls, reps=1)
    d["field_spectrum_ms"], _ = time_ms(lambda: lat.field_spectrum(t))
    d["energy_ms"], _ = time_ms(lambda: lat.energy(t))
    # estimate_covmat: 10 realisations (tol 0 runs all of them), reported per realisation
    ms, _ = time_ms(lambda: lat.field_covariance(3, 10, 10, 0.0), reps=1)
    d["covariance_ms_per_realisation"] = ms / 10
    d["n_string_cells"] = int(cells[0])
    d["n_wall_cells"] = int(cells[1])
    d["gw_spectra_per_s"] = 1e3 / d["gw_spectrum_ms"]
    return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}


def run_gpu(capi, N, L, precision, steps, warmup, seed, diag=False, tune=True):
    P = params(capi, N, L, precision, 128)
    real = 8 if precision == capi.FP64 else 4
    dt = 0.2 * (L / N)
    with capi.Lattice(P) as lat:
        lat.init_flat(seed, 1.0)
        tuned = lat.tune(5) if tune else None  # picks the fused pass's configuration; results unchanged
        t = 1.0
        if warmup > 0:
            lat.advance_steps(t, dt, warmup)
            t += warmup * dt
        lat.profile(True)
        t0 = time.perf_counter()
        lat.advance_steps(t, dt, steps)  # synchronizes before returning
        el = time.perf_counter() - t0
        prof = lat.profile(False)
        spec = lat.gw_spectrum()
        t = t + steps * dt
        dg = run_diagnostics(lat, t) if diag else None
    n3 = N ** 3
    kern = {}
    for k, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        avg = ms / cnt
        alg = KERNEL_REALS[k] * real * n3
        kern[k] = dict(ms_avg=avg, launches=cnt, alg_bytes=alg, gbs=alg / (avg * 1e-3) / 1e9)
    return dict(su_s=steps * n3 / el, ms_per_step=el / steps * 1e3, kernels=kern, real=real,
                finite=bool(np.all(np.isfinite(spec["numer"]))), diagnostics=dg,
                tuned=None if tuned is None else {"choice": tuned[0], "ypass": tuned[1],
                                                  "ms": {k: round(v, 4) for k, v in tuned[2].items()}})


def run_precision(capi, N, L, seed, steps=2000, every=250, dt=0.01, bins=128):
    """The precision workload: FP64 and FP32C lattices from the same seeded field advanced in
    lockstep with GW on; every `every` steps the FP32C GW spectrum's per-bin relative error against
    FP64, the cell counts and the string length (axion_grav_amd/precision_study.py).  dt = 0.01, not
    0.2 dspa: at lambda = 2 the leapfrog passes its stability bound near tau = 23, which 2000 steps of
    0.05 would cross (the FP64 oracle diverges between steps 400 and 500 there)."""
    from axion_grav_amd import precision_study as PS

    L64 = PS.lattice(N, L, capi.FP64, bins)
    L32 = PS.lattice(N, L, capi.FP32C, bins)
    try:
        L64.init_flat(seed, 1.0)
        phi, pi = L64.get_fields()
        L32.set_fields(phi, pi)
        del phi, pi
        t, rows = 1.0, []
        t0 = time.perf_counter()
        for blk in range(steps // every):
            L64.advance_steps(t, dt, every)
            L32.advance_steps(t, dt, every)
            t = 1.0 + (blk + 1) * every * dt
            step = (blk + 1) * every
            r = PS.compare(t, step, PS.diag(L64, t, step), PS.diag(L32, t, step))
            rows.append(r)
            log(f"precision: step {step} tau {t:.2f} GW max rel err {r['gw_relerr_max']:.3g}, "
                f"cells {r['cells_fp32c']} vs {r['cells_fp64']}")
        el = time.perf_counter() - t0
    finally:
        L32.close()
        L64.close()
    return dict(workload=f"{N}^3, L={L:g}, seed {seed}, {steps} steps of {dt:g} from tau 1, {bins} GW bins, "
                         f"FP32C vs FP64 every {every}",
                gw_relerr_max=max(r["gw_relerr_max"] for r in rows),
                gw_relerr_median_last=rows[-1]["gw_relerr_median"],
                gw_denom_equal=all(r["gw_denom_equal"] for r in rows),
                cells_equal=all(r["cells_fp64"] == r["cells_fp32c"] for r in rows),
                string_length_relerr_max=max(r["string_length_relerr"] for r in rows),
                energy_relerr_max=max(max(r["energy_relerr"]) for r in rows),
                per_checkpoint=[{k: r[k] for k in ("step", "tau", "gw_relerr_max", "cells_fp64")} for r in rows],
                wall_s=round(el, 1))


def run_adjoint(capi, N=128, L=32.0, steps=800, K=50, kappa0=1e-6, seed=99, t_end=16.0, check_steps=400):
    """Cost of the gradient of a GW band energy (k in [0.5, 2]) by the adjoint: a plain FP32C
    forward of `steps` steps, then the same steps with checkpoints every K and the backward sweep.
    The steps span tau = 1 .. t_end: a fixed step of 0.2 dspa would reach tau = 41, past the
    leapfrog's stability limit (near tau = 23 for lambda = 2 at dt = 0.05).  The gradients are
    checked against FP64 central differences (relative step 1e-4) over the first `check_steps`
    steps: over the whole run J varies too roughly with the parameters for differences to
    converge (at 800 steps steps of 1e-4 and 1e-6 disagree with each other)."""
    P = params(capi, N, L, capi.FP32C, 32)
    P.potential_kappa0 = kappa0
    dt = (t_end - 1.0) / steps
    with capi.Lattice(P) as lat:
        lat.init_flat(seed, 1.0)
        phi0, pi0 = lat.get_fields()
        lat.advance_steps(1.0, dt, 3)  # warm the kernels up
        lat.set_fields(phi0, pi0)
        t0 = time.perf_counter()
        lat.advance_steps(1.0, dt, steps)
        fwd = (time.perf_counter() - t0) * 1e3
        lat.set_fields(phi0, pi0)
        g = lat.gradient(1.0, dt, steps, K, 0.5, 2.0)
        lat.set_fields(phi0, pi0)
        gc = lat.gradient(1.0, dt, check_steps, K, 0.5, 2.0)

    def J64(s2f=1.0, k0f=1.0, a=1.0):
        P64 = params(capi, N, L, capi.FP64, 32)
        P64.sigma_cr = SIGMA_CR * s2f ** 0.5
        P64.potential_kappa0 = kappa0 * k0f
        with capi.Lattice(P64) as lat:
            lat.set_fields(a * phi0, a * pi0)
            lat.advance_steps(1.0, dt, check_steps)
            s = lat.gw_spectrum()
        band = (s["qmid"] >= 0.5) & (s["qmid"] <= 2.0) & (s["denom"] > 0)
        return float(s["power"][band].sum())

    h = 1e-4
    lam = 4 * SIGMA_CR ** 2
    fd = {"dJ_dlambda": (J64(s2f=1 + h) - J64(s2f=1 - h)) / (2 * h * lam),
          "dJ_dkappa0": (J64(k0f=1 + h) - J64(k0f=1 - h)) / (2 * h * kappa0),
          "dJ_dA": (J64(a=1 + h) - J64(a=1 - h)) / (2 * h)}
    ad = {"dJ_dlambda": gc["dJ_dsigma2"] / 4, "dJ_dkappa0": gc["dJ_dkappa0"], "dJ_dA": gc["dJ_dlnamp"]}
    rel = {k: float(f"{abs(ad[k] / fd[k] - 1):.3g}") for k in fd}
    return dict(workload=f"{N}^3, L={L:g}, FP32C, kappa0={kappa0:g}, seed {seed}, {steps} steps of {dt:g} from tau 1, "
                         f"checkpoints every {K}",
                J=g["J"], dJ_dlambda=g["dJ_dsigma2"] / 4, dJ_dkappa0=g["dJ_dkappa0"], dJ_dA=g["dJ_dlnamp"],
                forward_ms=round(fwd, 2), forward_with_checkpoints_ms=round(g["forward_ms"], 2),
                adjoint_ms=round(g["adjoint_ms"], 2), adjoint_over_forward=round(g["adjoint_ms"] / fwd, 3),
                gradient_over_forward=round((g["forward_ms"] + g["adjoint_ms"]) / fwd, 3),
                check=dict(steps=check_steps, adjoint_fp32c=ad, fd_fp64=fd, rel_err=rel,
                           horizon=f"finite differences checked over the first {check_steps} of the {steps} steps: "
                                   f"over all {steps} J varies too roughly with the parameters for central "
                                   f"differences to converge (relative steps 1e-4 and 1e-6 disagree by 4x)"))


def run_batch(capi, steps=600, chunk=40, t_end=16.0):
    """Independent realizations of uneven sizes advanced together: 16 x 64^3 + 16 x 96^3 +
    16 x 128^3, dspa = 0.25, seeds 0..47, FP32C, GW on (64 bins), `steps` steps spanning tau
    1 .. t_end (a fixed 0.2 dspa would pass the leapfrog's stability limit near tau = 23).
    Times one captured graph of `chunk` steps of all 48 replayed, the same kernels launched
    directly, and the lattices advanced one after another (axg_advance_steps); then the
    per-realization GW spectra, averaged per size."""
    sizes = [64] * 16 + [96] * 16 + [128] * 16
    dt = (t_end - 1.0) / steps
    lats = []
    try:
        for i, n in enumerate(sizes):
            P = params(capi, n, 0.25 * n, capi.FP32C, 64)
            lat = capi.Lattice(P)
            lat.init_flat(i, 1.0)
            lats.append(lat)
        for lat in lats:
            lat.tune(3)
        init = [lat.get_fields() for lat in lats]
        su = sum(n ** 3 for n in sizes) * steps

        def reset():
            for lat, (a, b) in zip(lats, init):
                lat.set_fields(a, b)

        capi.batch_advance(lats, 1.0, dt, 4, chunk)  # warm-up (and graph code paths)
        reset()
        t_graph = capi.batch_advance(lats, 1.0, dt, steps, chunk)
        # per realization: the 64-bin GW spectrum and the string / wall / mask cell counts; then
        # their means per size
        spec, cells = {}, {}
        per = []
        for lat, n in zip(lats, sizes):
            pw = lat.gw_spectrum()["power"]
            c = lat.defects()
            per.append((n, pw, c))
        for n in (64, 96, 128):
            spec[n] = np.mean([pw for m, pw, _ in per if m == n], axis=0)
            cells[n] = np.mean([c for m, _, c in per if m == n], axis=0)
        reset()
        t_direct = capi.batch_advance(lats, 1.0, dt, steps, 0)
        reset()
        t0 = time.perf_counter()
        for lat in lats:
            lat.advance_steps(1.0, dt, steps)
        t_seq = (time.perf_counter() - t0) * 1e3
    finally:
        for lat in lats:
            lat.close()
    return dict(workload=f"16 x 64^3 + 16 x 96^3 + 16 x 128^3, dspa 0.25, FP32C, GW on, {steps} steps of {dt:g}",
                graph_chunk=chunk, graph_ms=round(t_graph, 1), direct_ms=round(t_direct, 1),
                one_by_one_ms=round(t_seq, 1), su_s_graph=su / (t_graph * 1e-3),
                su_s_one_by_one=su / (t_seq * 1e-3), speedup_graph_vs_one_by_one=round(t_seq / t_graph, 3),
                size_mean_gw_power_finite=bool(all(np.all(np.isfinite(v)) for v in spec.values())),
                size_mean_cells={str(n): [round(float(x), 1) for x in cells[n]] for n in cells},
                size_mean_gw_power={str(n): [float(f"{x:.4g}") for x in spec[n]] for n in spec},
                realization_string_cells=[int(c[0]) for _, _, c in per])


def run_largest(capi, N=1024, L=256.0, steps=10, warmup=2, kappa0=1e-6, seed=2024):
    """The largest single-GPU share: 1024^3, L = 256, FP32C with GW (256 bins), the N_DW = 2 tilt
    kappa0 tau^4, seed 2024.  Times `steps` steps after `warmup`, then one round of the
    diagnostics the driver runs at a checkpoint (classification, string and wall fits, GW and
    field spectra with 256 bins; the first field spectrum also builds the deconvolution tables)."""
    P = params(capi, N, L, capi.FP32C, 256)
    P.potential_kappa0 = kappa0
    P.num_bin = 256
    dt = 0.2 * (L / N)
    d = {}
    with capi.Lattice(P) as lat:
        lat.init_flat(seed, 1.0)
        tuned = lat.tune(3)
        lat.advance_steps(1.0, dt, warmup)
        t = 1.0 + warmup * dt
        lat.profile(True)
        t0 = time.perf_counter()
        lat.advance_steps(t, dt, steps)
        el = time.perf_counter() - t0
        prof = lat.profile(False)
        t += steps * dt
        for name, fn in (("gw_spectrum_ms", lat.gw_spectrum), ("defects_ms", lat.defects),
                         ("strings_ms", lambda: lat.strings(1, 2)), ("walls_ms", lat.walls),
                         ("field_spectrum_first_ms", lambda: lat.field_spectrum(t)),
                         ("field_spectrum_ms", lambda: lat.field_spectrum(t))):
            t1 = time.perf_counter()
            r = fn()
            d[name] = round((time.perf_counter() - t1) * 1e3, 2)
            if name == "defects_ms":
                d["n_string_cells"], d["n_wall_cells"], d["n_mask_cells"] = (int(v) for v in r)
            if name == "gw_spectrum_ms":
                d["gw_power_finite"] = bool(np.all(np.isfinite(r["power"])))
    kern = {k: round(ms / cnt, 3) for k, (ms, cnt) in prof.items() if cnt}
    return dict(workload=f"{N}^3, L={L:g}, FP32C, GW on (256 bins), kappa0={kappa0:g}, seed {seed}, "
                         f"{steps} steps of {dt:g} after {warmup}",
                su_s=steps * N ** 3 / el, ms_per_step=round(el / steps * 1e3, 3), kernel_ms=kern,
                tuning={"choice": tuned[0], "ypass": tuned[1], "ms": {k: round(v, 3) for k, v in tuned[2].items()}},
                diagnostics=d)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_for_this_host():
    """oracle/oracle.c compiled for this host's CPU (-march=native, the same -ffp-contract=off
    FP64 order) into a temporary directory; None when the host has no compiler for it."""
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="axg_oracle_"), "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-std=c11", "-shared", "-o", out, os.path.join(HERE, "oracle", "oracle.c"), "-lm"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    return out if r.returncode == 0 else None


def run_cpu(N, L, seed, steps=3, warmup=1):
    import oracle as O

    dspa = L / N
    dwav = TWOPI_T / dspa / N
    P = dict(dspa=dspa, dwav=dwav, dv_spa=dspa ** 3, sigma_cr=SIGMA_CR, kappa=1e-30, initial_loop=True)
    phi, pi = O.ic_flat(N, seed, 1.0, dwav)
    A = O.new_gw_state(N)
    dt = 0.2 * dspa
    O.symplectic(phi, pi, A, 1.0, dt, warmup, P)
    t0 = time.perf_counter()
    O.symplectic(phi, pi, A, 1.0 + warmup * dt, dt, steps, P)
    el = time.perf_counter() - t0
    return steps * N ** 3 / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--box", type=float, default=128.0)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-diag", action="store_true", help="skip timing the diagnostics after the FP32C run")
    ap.add_argument("--no-tune", action="store_true", help="keep the fused pass's configuration chosen at creation (by lattice size)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip timing the gradient (128^3, 800 steps)")
    ap.add_argument("--no-batch", action="store_true", help="skip timing the 48-realization batch")
    ap.add_argument("--no-largest", action="store_true",
                    help="skip timing the largest single-GPU share (1024^3 FP32C with GW and walls, ~130 GB)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32C-vs-FP64 spectrum error run (512^3, 2000 steps of 0.01)")
    ap.add_argument("--no-probe", action="store_true",
                    help="skip the HBM streaming probe (read, write, copy of 4 GiB with the library's own kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the lattice lives in one MI355X's HBM and this build has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    from axion_grav_amd import capi

    probe = None
    if not args.no_probe:
        probe = {k: round(v, 1) for k, v in capi.hbm_probe(4 << 30, 10).items()}
        log(f"HBM probe (4 GiB, GB/s): {probe}")
    N, L = args.n, args.box
    log(f"FP32C {N}^3 L={L}: warmup {args.warmup}, timed {args.steps} steps")
    c = run_gpu(capi, N, L, capi.FP32C, args.steps, args.warmup, args.seed, diag=not args.no_diag,
                tune=not args.no_tune)
    log(f"FP32C tuning: {c['tuned']}")
    if c["diagnostics"]:
        log("FP32C diagnostics: " + ", ".join(f"{k} {v}" for k, v in c["diagnostics"].items()))
    log(f"FP32C {c['su_s']:.4g} site-updates/s, {c['ms_per_step']:.3f} ms/step; kernels "
        + ", ".join(f"{k} {v['ms_avg']:.3f} ms {v['gbs']:.0f} GB/s" for k, v in c["kernels"].items()))
    d = None
    if not args.no_fp64:
        log(f"FP64 {N}^3: warmup {args.warmup}, timed {args.steps} steps")
        d = run_gpu(capi, N, L, capi.FP64, args.steps, args.warmup, args.seed, tune=not args.no_tune)
        log(f"FP64 {d['su_s']:.4g} site-updates/s, {d['ms_per_step']:.3f} ms/step; kernels "
            + ", ".join(f"{k} {v['ms_avg']:.3f} ms {v['gbs']:.0f} GB/s" for k, v in d["kernels"].items()))

    adj = None
    if not args.no_adjoint:
        log("adjoint: 128^3 FP32C, 800 steps, checkpoints every 50")
        adj = run_adjoint(capi)
        log(f"adjoint: {adj}")

    bat = None
    if not args.no_batch:
        log("batch: 16 x 64^3 + 16 x 96^3 + 16 x 128^3 FP32C, 600 steps")
        bat = run_batch(capi)
        log(f"batch: {bat}")

    prec = None
    if not args.no_precision:
        log(f"precision: {N}^3 FP32C vs FP64, 2000 steps of 0.01, every 250")
        prec = run_precision(capi, N, L, args.seed)
        log(f"precision: max GW rel err {prec['gw_relerr_max']:.3g}, cells equal {prec['cells_equal']}")

    big = None
    if not args.no_largest:
        log("largest: 1024^3 FP32C, GW on, N_DW = 2 tilt")
        big = run_largest(capi)
        log(f"largest: {big}")

    dom_name, dom = max(c["kernels"].items(), key=lambda kv: kv[1]["ms_avg"] * kv[1]["launches"])
    tc = c["tuned"]  # without tuning the configuration is the one chosen at creation: no counted traffic
    traffic = None if tc is None else measured_traffic(
        dom_name, "fp32c", N, tc["choice"] if dom_name == "kick_source_drift" else tc["ypass"])
    roof = dict(bound="hbm", kernel=dom_name, achieved=round(dom["gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(dom["gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                alg_bytes_per_launch=dom["alg_bytes"], ms_per_launch=round(dom["ms_avg"], 4),
                traffic_gbs=None if traffic is None else round(traffic / (dom["ms_avg"] * 1e-3) / 1e9, 1))
    step_frac = MODEL_REALS * c["real"] * c["su_s"] / (HBM_PEAK_GBS * 1e9)

    cpu = None
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        native = oracle_for_this_host()
        march = "-march=native" if native else "-march=x86-64-v3"
        if native:
            import oracle as O
            O.use_library(native)
        log(f"CPU oracle ({march}): 1 + 3 steps at {N}^3 on {cores} threads")
        su, el = run_cpu(N, L, args.seed)
        cpu = dict(value=su, unit="site-updates/s", cores=cores, kind="port",
                   sample=f"3 DKD steps with GW at {N}^3 after 1 untimed (FP64 oracle, unfused reference dataflow, "
                          f"-O3 {march} -ffp-contract=off), {el:.1f} s; {cpu_model()}, OMP_PROC_BIND="
                          f"{os.environ.get('OMP_PROC_BIND', 'unset')}")
        log(f"CPU {su:.4g} site-updates/s ({el:.1f} s)")

    out = {
        "metric": "site-updates/s at 512^3 GW on (FP32c)",
        "value": c["su_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded flat-spectrum initial field, seed %d)" % args.seed,
        "config": {"workload": f"{N}^3 lattice, L={L:g}, GW on, FP32 compensated storage", "N": N, "L": L,
                   "dtau": 0.2 * L / N, "gw_bins": 128, "precision": "fp32c"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_model_frac": round(step_frac, 4),
        "fp64": None if d is None else {"value": d["su_s"], "ms_per_step": d["ms_per_step"],
                                        "step_model_frac": round(MODEL_REALS * 8 * d["su_s"] / 8e12, 4)},
        "fp32c_over_fp64": None if d is None else round(c["su_s"] / d["su_s"], 3),
        "kernels": {k: {"ms": round(v["ms_avg"], 4), "GB/s": round(v["gbs"], 1)} for k, v in c["kernels"].items()},
        "hbm_probe_gbs": probe,
        "diagnostics": c["diagnostics"],
        "tuning": {"fp32c": c["tuned"], "fp64": None if d is None else d["tuned"]},
        "adjoint": adj,
        "batch": bat,
        "largest": big,
        "precision": prec,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
