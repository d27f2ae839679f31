"""Build steps for the native parts of the package.

``build_hip()`` compiles the gfx950 HIP library ``libnbody2d.so`` (kernels +
C ABI) into this directory; ``build_host()`` compiles the C++ host adapter
(``libnbody2d_host.so`` and the headless driver ``nbody2d-run``) against it.
Both always recompile. ``build_oracle()`` builds the CPU oracle under
``oracle/`` (test infrastructure, not linked by the product).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent               # n-body-2d_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
HOST = ROOT / "host"
INCLUDE = REPO / "include"
ORACLE = REPO / "oracle"
LIB = PKG / "libnbody2d.so"
HOST_LIB = PKG / "libnbody2d_host.so"
HOST_BIN = PKG / "nbody2d-run"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_hip(jobs: int = 8) -> Path:
    objdir = PKG / "_obj"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))

    def one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    return LIB


def build_host() -> list[Path]:
    srcs = sorted(HOST.glob("*.cpp"))
    lib_srcs = [s for s in srcs if s.name != "nbody2d_run.cpp"]
    common = ["-std=c++20", "-O2", "-fPIC", f"-I{INCLUDE}", f"-I{HOST}", "-pthread"]
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run(["g++", *common, "-shared", "-o", str(tmp), *map(str, lib_srcs),
          f"-L{PKG}", "-lnbody2d", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    main = HOST / "nbody2d_run.cpp"
    if main.exists():
        tmpb = HOST_BIN.with_suffix(".tmp")
        _run(["g++", *common, "-o", str(tmpb), str(main), f"-L{PKG}", "-lnbody2d_host",
              "-lnbody2d", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmpb, HOST_BIN)
    return [HOST_LIB, HOST_BIN]


def build_oracle() -> None:
    """CPU oracle (tests/ and bench.py's cpu_baseline leg only)."""
    _run(["make", "-B", "-C", str(ORACLE)])


def build_all() -> None:
    build_hip()
    if HOST.exists() and any(HOST.glob("*.cpp")):
        build_host()
    build_oracle()
