"""Build the HIP C-ABI library (gfx950) in-tree.

    python -m entropical_amd.build          # library only
    python -m entropical_amd.build --oracle # also the CPU oracle (test infrastructure)

Output: entropical_amd/libentropical_kde.so.  Objects go to _build/ (ignored).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libentropical_kde.so"
ARCH = os.environ.get("ENT_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kde_density_q4.hip", "kde_density_q4w.hip", "kde_density_q2.hip", "kde_density_q1.hip",
               "kde_density_q1s3.hip", "kde_kernels.hip",
               "engine.hip", "autotune.hip", "capi.hip", "reductions.hip", "drivers.hip",
               "mlp.hip", "noalloc.hip", "colstats.hip"]
HOST_SOURCES = ["densities_shim.cpp", "xtc.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay off (see the design notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found (ROCm required to build the HIP library)")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        out = "\n".join(
            ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln
        )
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _stale(src: Path, obj: Path) -> bool:
    """An object older than its source or any header of the library."""
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    deps = [src, *CSRC.glob("*.hpp"), *(ROOT / "include").rglob("*.h*")]
    return any(d.stat().st_mtime > t for d in deps)


def build_library(jobs: int = 8, verbose: bool = True, incremental: bool = False) -> Path:
    """Compile every source (incremental: only those whose object is stale)
    and link the library."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        if not incremental or _stale(CSRC / src, obj):
            cmds.append([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = BUILD / (src + ".o")
        objs.append(obj)
        if not incremental or _stale(CSRC / src, obj):
            cmds.append([hipcc, "-x", "c++", "-O3", "-std=c++17", "-fPIC", "-Wall",
                         "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", f"-I{CSRC}",
                         "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_dropin_check(verbose: bool = True) -> Path:
    """C++ program linking the drop-in header against the library (used by
    tests/test_gpu_dropin.py)."""
    src = ROOT / "tests" / "cpp" / "dropin_check.cpp"
    exe = src.with_suffix("")
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{PKG}", "-lentropical_kde",
          f"-Wl,-rpath,$ORIGIN/../../entropical_amd", "-o", str(exe)])
    if verbose:
        print(f"built {exe}")
    return exe


def build_cli(verbose: bool = True) -> Path:
    """The entropical command-line program over the library
    (entropical_amd/bin/entropical)."""
    src = PKG / "cli" / "entropical_main.cpp"
    exe = PKG / "bin" / "entropical"
    exe.parent.mkdir(exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(src), f"-L{PKG}", "-lentropical_kde",
          "-Wl,-rpath,$ORIGIN/..", "-o", str(exe)])
    if verbose:
        print(f"built {exe}")
    return exe


def build_oracle(verbose: bool = True, with_ref: bool = True) -> None:
    od = ROOT / "oracle"
    _run(["make", "-C", str(od), "-j4", "all"])
    if with_ref and Path("/root/reference/boxedsearch.cpp").exists():
        _run(["make", "-C", str(od), "ref"])
    if verbose:
        print(f"built {od / 'libkde_oracle.so'}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--incremental", action="store_true", help="recompile only stale objects")
    a = ap.parse_args(argv)
    build_library(a.j, incremental=a.incremental)
    build_dropin_check()
    build_cli()
    if a.oracle:
        build_oracle()
    return 0


if __name__ == "__main__":
    sys.exit(main())
