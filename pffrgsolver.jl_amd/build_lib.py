"""Compile the gfx950 HIP library libpffrg.so in-tree (pffrg/_lib/).

The vertex-pass kernel is instantiated once per pass type in csrc/pass_*.hip; the
translation units compile in parallel and are linked into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
SOURCES = ["pffrg.hip", "pass_kat.hip", "pass_left.hip", "pass_central.hip", "pass_bse.hip", "pass_reduced.hip",
           "chi.hip", "mlp.hip", "rescan.hip", "gather_probe.hip"]
# per-file flags: the correlation cubature and the mesh rescanning restate the oracle's arithmetic
# without FMA contraction, so that adaptive region choices and resampled values are the oracle's
FILE_FLAGS = {"chi.hip": ["-ffp-contract=off"], "rescan.hip": ["-ffp-contract=off"]}
OUT = PKG / "pffrg" / "_lib" / "libpffrg.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int | None = None, out: Path | None = None, defines=()) -> Path:
    out = Path(out) if out else OUT
    out.parent.mkdir(parents=True, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    with tempfile.TemporaryDirectory(prefix="pffrg_build_") as tmpd:
        def compile_one(src: str) -> str:
            obj = str(Path(tmpd) / (Path(src).stem + ".o"))
            r = subprocess.run([hipcc, *HIPCC_FLAGS, *FILE_FLAGS.get(src, []), *[f"-D{d}" for d in defines], *inc,
                                "-c", "-o", obj,
                                str(CSRC / src)],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}{r.stderr}")
            if verbose:
                sys.stderr.write(r.stderr)
            return obj

        n = jobs or min(len(SOURCES), os.cpu_count() or 1, 8)
        with ThreadPoolExecutor(max_workers=n) as ex:
            objs = list(ex.map(compile_one, SOURCES))
        tmp = out.with_suffix(".so.tmp")
        r = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, out)
    return out


# HDF5 observable/checkpoint files (host code, no GPU): built against the HDF5 C library of the
# image when its headers are present; pffrg.h5io fails loudly when the library is missing
H5_PREFIX = Path(os.environ.get("PFFRG_HDF5_PREFIX", "/opt/conda"))
H5_OUT = PKG / "pffrg" / "_lib" / "libpffrg_h5.so"


def build_h5(out: Path | None = None) -> Path | None:
    out = Path(out) if out else H5_OUT
    if not (H5_PREFIX / "include" / "hdf5.h").exists():
        sys.stderr.write(f"libpffrg_h5.so not built: no HDF5 headers under {H5_PREFIX}\n")
        return None
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           f"-I{ROOT / 'include'}", f"-I{H5_PREFIX / 'include'}", str(CSRC / "h5io.cpp"),
           f"-L{H5_PREFIX / 'lib'}", "-lhdf5", f"-Wl,-rpath,{H5_PREFIX / 'lib'}", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{r.stdout}{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_h5())
