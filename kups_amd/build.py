"""Build libkupsmi.so (all HIP sources, gfx950) in-tree with hipcc.

Run as ``python -m kups_amd.build`` or through ``__graft_entry__.build()``.
Every source is compiled to its own object (in parallel, objects under
``kups_amd/_build``) and the objects are linked into ``kups_amd/libkupsmi.so``,
next to this file, so the library travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libkupsmi.so"
OBJ = PKG / "_build"
INCLUDE = PKG.parent / "include"

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # packed FP32 VALU ops are disabled for every HIP source of the library
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}: {' '.join(cmd[-3:])}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link kups_amd/libkupsmi.so (always rebuilds)."""
    OBJ.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = sources()
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    cmds = [[cc, *HIPCC_FLAGS, f"-I{INCLUDE}", "-c", "-o", str(o), str(s)]
            for s, o in zip(srcs, objs)]
    n = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=n) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)],
         verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
