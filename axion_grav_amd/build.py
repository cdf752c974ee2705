"""Build libaxgrav.so (the HIP kernels + C-ABI) for gfx950 with hipcc, in-tree.

    python -m axion_grav_amd.build            # objects in axion_grav_amd/_build, library in axion_grav_amd/lib
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libaxgrav.so")
SOURCES = ["capi.hip", "step.hip", "gw.hip", "diag.hip", "strings.hip", "walls.hip", "groups.hip", "adjoint.hip", "noalloc.hip", "probe.hip", "plaquettes.hip", "degrade.hip"]
FORTRAN = ["axgrav_capi.f90", "axg_ini.f90", "axg_driver.f90"]
BINDIR = os.path.join(HERE, "bin")
FC = os.environ.get("AXG_FC", "/opt/rocm/bin/amdflang")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AXG_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++20", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-result",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    for _, w in results:
        if w and verbose:
            print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


def build_fortran(verbose=True):
    """The Fortran host driver (axion_grav_amd/bin/axion_3d) over the ISO_C_BINDING module."""
    if not os.path.exists(FC):
        if verbose:
            print(f"no Fortran compiler at {FC}: skipping the Fortran host driver", file=sys.stderr)
        return None
    os.makedirs(BINDIR, exist_ok=True)
    fdir = os.path.join(HERE, "fortran")
    objs = []
    for src in FORTRAN:
        obj = os.path.join(BUILD, src.replace(".f90", ".o"))
        cmd = [FC, "-O2", "-c", os.path.join(fdir, src), "-o", obj, "-module-dir", BUILD, "-I", BUILD]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed on {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    exe = os.path.join(BINDIR, "axion_3d")
    cmd = [FC, "-o", exe, *objs, "-L", LIBDIR, "-laxgrav", f"-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of axion_3d failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    # the reference-side binding with halo-padded arrays, as a runnable check
    obj = os.path.join(BUILD, "halo_binding_check.o")
    chk = os.path.join(BINDIR, "halo_binding_check")
    for cmd in ([FC, "-O2", "-c", os.path.join(fdir, "halo_binding_check.f90"), "-o", obj, "-module-dir", BUILD,
                 "-I", BUILD],
                [FC, "-o", chk, objs[0], obj, "-L", LIBDIR, "-laxgrav", f"-Wl,-rpath,$ORIGIN/../lib"]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {chk}")
    return exe


if __name__ == "__main__":
    build()
