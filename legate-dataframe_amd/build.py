"""Builds the native parts of the package in-tree.

  libldf_hip.so          HIP kernels + the C ABI declared in include/ldf_hip.h (hipcc, gfx950)
  libldf.so              the C++ host API (legate::dataframe::join, groupby_aggregation, ...)
  legate_dataframe/_ldf  pybind11 bindings of the C++ host API

Run as ``python legate-dataframe_amd/build.py`` (or through ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
CPP = PKG / "cpp"
BUILD = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("LDF_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions off (they have produced silent wrong results on this hardware)
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    "-munsafe-fp-atomics",
]

HIP_SOURCES = ["common.hip", "partition.hip", "join.hip", "groupby.hip", "median.hip", "batched.hip", "distinct.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _newer(target: Path, deps: list[Path]) -> bool:
    return target.exists() and all(target.stat().st_mtime >= d.stat().st_mtime for d in deps)


def build_hip(force: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = str(ROCM / "bin" / "hipcc")
    headers = list(CSRC.glob("*.hpp")) + [ROOT / "include" / "ldf_hip.h"]
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        s = CSRC / src
        if not s.exists():
            continue
        o = BUILD / (s.stem + ".o")
        objs.append(o)
        if force or not _newer(o, [s] + headers):
            jobs.append([hipcc, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, max(1, len(jobs)))) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    lib = PKG / "libldf_hip.so"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(lib), *map(str, objs)])
    return lib


def build_cpp(force: bool = True) -> Path:
    """C++ host API on top of the C ABI."""
    srcs = sorted((CPP / "src").glob("*.cpp"))
    lib = PKG / "libldf.so"
    if not srcs:
        return lib
    hdrs = list((CPP / "include").rglob("*.hpp"))
    if not force and _newer(lib, srcs + hdrs):
        return lib
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared",
        f"-I{CPP / 'include'}", f"-I{ROOT / 'include'}", f"-I{ROCM / 'include'}",
        "-D__HIP_PLATFORM_AMD__",
        *map(str, srcs),
        "-o", str(lib),
        f"-L{PKG}", "-lldf_hip", f"-L{ROCM / 'lib'}", "-lamdhip64",
        "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}",
    ]
    _run(cmd)
    return lib


def build_python(force: bool = True) -> Path:
    import pybind11

    src = PKG / "legate_dataframe" / "_ldf.cpp"
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = PKG / "legate_dataframe" / f"_ldf{ext}"
    if not src.exists():
        return out
    hdrs = list((CPP / "include").rglob("*.hpp"))
    if not force and _newer(out, [src] + hdrs + [PKG / "libldf.so"]):
        return out
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared",
        f"-I{pybind11.get_include()}", f"-I{py_inc}",
        f"-I{CPP / 'include'}", f"-I{ROOT / 'include'}", f"-I{ROCM / 'include'}",
        "-D__HIP_PLATFORM_AMD__",
        str(src), "-o", str(out),
        f"-L{PKG}", "-lldf", "-lldf_hip", f"-L{ROCM / 'lib'}", "-lamdhip64",
        "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM / 'lib'}",
    ]
    _run(cmd)
    return out


def build_all(force: bool = True) -> None:
    build_hip(force)
    build_cpp(force)
    build_python(force)


if __name__ == "__main__":
    build_all(force="--incremental" not in sys.argv)
    print("built", PKG / "libldf_hip.so")
