import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


SQ_HALF = float(np.sqrt(0.5))  # sigma_cr for lambda = 2 (V = lambda/4 (|phi|^2-1)^2 <-> sigma_cr^2 = lambda/4)


def phys(N, dspa=0.25):
    """Oracle-side parameter dict of the lambda=2 radiation-era runs (initial_loop, kappa ~ 0)."""
    twopi_t = 2 * 3.14159265359
    dwav = twopi_t / dspa / N
    r = dwav / twopi_t
    return dict(N=N, dspa=dspa, dwav=dwav, dv_spa=dspa ** 3, dv_wav=r * r * r, sigma_cr=SQ_HALF,
                kappa=1e-30, initial_loop=True, tau_cr=1.0, a_now=1.0)
