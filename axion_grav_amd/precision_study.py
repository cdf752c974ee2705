#!/usr/bin/env python3
"""FP32-compensated vs FP64 on the same GPU: error of every reported diagnostic.

    python -m axion_grav_amd.precision_study [--n 512] [--box 128] [--steps 2000]
                                             [--every 250] [--dt 0.01] [--seed 7] [--bins 128]
                                             [--storage fp32c|fp32]

Both precisions start from the same seeded flat-spectrum field (|k| < 1, <|phi|^2> = 1,
pi = 0, lambda = 2 at zero temperature, radiation era from tau = 1) and advance in lockstep
with GW accumulation on.  Every `--every` steps each lattice reports its 128-bin GW spectrum,
its energy (FP64 reductions in both precisions), its string / wall / mask cell counts, its
string length and velocities and its wall area, and
one JSON line gives the FP32C error against FP64 (`--storage fp32`: plain FP32 storage in
place of FP32C, the keys keep their names): the per-bin relative error of the GW power
(max and median over the bins that hold modes), the relative error of the energy terms, the
cell counts, and the relative error of the summed string length and mean string velocities
(fits seeded alike in both precisions) and of the summed wall area.

The step is dt = 0.01 by default: with lambda = 2 the radiation-era potential term grows as
tau^2, and the leapfrog passes its stability bound near tau = 23 (at dt = 0.05 the FP64 oracle
diverges between steps 400 and 500), so 2000 steps of 0.01 cover tau = 1 .. 21.
"""
import argparse
import json
import sys
import time

import numpy as np

from axion_grav_amd import capi

SIGMA_CR = float(np.sqrt(0.5))


def lattice(N, box, precision, bins):
    P = capi.make_params(N, box / N, SIGMA_CR, precision=precision, grav_spec=True, initial_loop=True,
                         kappa=1e-30, num_bin_grav=bins, id_strings=True, id_walls=True, spec_est=True)
    return capi.Lattice(P)


def relerr(a, b):
    return float(abs(a - b) / abs(b)) if b != 0 else float(abs(a))


def compare(t, step, d64, d32):
    s64, s32 = d64["spec"], d32["spec"]
    live = s64["denom"] > 0
    ok = live & (s64["power"] != 0)
    rel = np.abs(s32["power"][ok] - s64["power"][ok]) / np.abs(s64["power"][ok])
    e64, e32 = d64["energy"], d32["energy"]
    return dict(step=step, tau=round(t, 6), gw_bins=int(ok.sum()),
                gw_relerr_max=float(rel.max()) if rel.size else None,
                gw_relerr_median=float(np.median(rel)) if rel.size else None,
                gw_denom_equal=bool(np.array_equal(s32["denom"], s64["denom"])),
                energy_relerr=[relerr(a, b) for a, b in zip(e32, e64)],
                cells_fp64=[int(x) for x in d64["cells"]], cells_fp32c=[int(x) for x in d32["cells"]],
                string_length_fp64=d64["str"]["length"],
                string_length_relerr=relerr(d32["str"]["length"], d64["str"]["length"]),
                string_vel_relerr=[relerr(a, b) for a, b in zip(d32["str"]["vel"], d64["str"]["vel"])],
                wall_area_fp64=d64["wall"]["area"],
                wall_area_relerr=relerr(d32["wall"]["area"], d64["wall"]["area"]))


def diag(L, t, step):
    cells = L.defects()
    return dict(spec=L.gw_spectrum(), energy=L.energy(t), cells=cells, str=L.strings(7, step), wall=L.walls())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--box", type=float, default=128.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--bins", type=int, default=128)
    ap.add_argument("--storage", choices=("fp32c", "fp32"), default="fp32c")
    a = ap.parse_args()
    if a.steps % a.every:
        sys.exit("--steps must be a multiple of --every")
    L64 = lattice(a.n, a.box, capi.FP64, a.bins)
    L32 = lattice(a.n, a.box, capi.FP32C if a.storage == "fp32c" else capi.FP32, a.bins)
    L64.init_flat(a.seed, 1.0)
    phi, pi = L64.get_fields()
    L32.set_fields(phi, pi)
    del phi, pi
    t, worst = 1.0, 0.0
    t0 = time.perf_counter()
    for blk in range(a.steps // a.every):
        L64.advance_steps(t, a.dt, a.every)
        L32.advance_steps(t, a.dt, a.every)
        t = 1.0 + (blk + 1) * a.every * a.dt
        step = (blk + 1) * a.every
        r = compare(t, step, diag(L64, t, step), diag(L32, t, step))
        r["wall_s"] = round(time.perf_counter() - t0, 1)
        if r["gw_relerr_max"] is not None:
            worst = max(worst, r["gw_relerr_max"])
        print(json.dumps(r), flush=True)
    print(json.dumps(dict(summary=True, n=a.n, box=a.box, steps=a.steps, dt=a.dt, seed=a.seed, bins=a.bins,
                          storage=a.storage, gw_relerr_max_over_run=worst)), flush=True)
    L32.close()
    L64.close()


if __name__ == "__main__":
    main()
