"""In-tree build of the native libraries.

* lib/libsparc_hip.so   -- HIP kernels + C ABI (include/sparc_hip.h), gfx950.
* lib/libsparc_torch.so -- thin torch binding (torch.classes.sparc.Solver),
                           linked against libsparc_hip.so.

Run ``python -m sparc_flows_amd.build``.  hipcc cross-compiles for gfx950 on
machines without a GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
HIP_SO = LIB / "libsparc_hip.so"
TORCH_SO = LIB / "libsparc_torch.so"
ARCH = os.environ.get("SPARC_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["sparc_kernels.hip", "sparc_colseg.hip", "sparc_capi.hip", "sparc_3d.hip", "sparc_discrete.hip"]
# plain C++ compiled for the host (also built into the CPU check of the column pass)
HOST_SOURCES = ["sparc_segtab.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of generated code on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-8000:])
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build_hip() -> Path:
    """Compile each translation unit in parallel, then link the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    LIB.mkdir(exist_ok=True)
    objdir = LIB / "obj"
    objdir.mkdir(exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [objdir / (Path(s).stem + ".o") for s in HIP_SOURCES]
    hobjs = [objdir / (Path(s).stem + ".host.o") for s in HOST_SOURCES]
    jobs = [["hipcc", *cflags, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(HIP_SOURCES, objs)]
    jobs += [["g++", "-O2", "-std=c++17", "-fPIC", "-c", str(CSRC / s), "-o", str(o)]
             for s, o in zip(HOST_SOURCES, hobjs)]
    with ThreadPoolExecutor(len(jobs)) as ex:
        list(ex.map(_run, jobs))
    objs += hobjs
    tmp = HIP_SO.with_suffix(".so.tmp")
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(tmp)])
    os.replace(tmp, HIP_SO)
    return HIP_SO


def build_torch() -> Path:
    import torch
    from torch.utils import cpp_extension

    LIB.mkdir(exist_ok=True)
    tlib = Path(torch.__file__).resolve().parent / "lib"
    tmp = TORCH_SO.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", str(CSRC / "sparc_torch.cpp"), "-o", str(tmp),
           f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}"]
    cmd += [f"-I{p}" for p in cpp_extension.include_paths()]
    cmd += [f"-I{sysconfig.get_paths()['include']}"]
    cmd += [f"-L{LIB}", "-lsparc_hip", f"-L{tlib}", "-ltorch", "-ltorch_cpu", "-lc10",
            "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{tlib}"]
    _run(cmd)
    os.replace(tmp, TORCH_SO)
    return TORCH_SO


def build_all() -> None:
    build_hip()
    build_torch()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "hip" in what:
        build_hip()
    elif "torch" in what:
        build_torch()
    else:
        build_all()
    print("built", HIP_SO, TORCH_SO)
