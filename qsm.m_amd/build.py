"""Builds libqsm_hip.so (every HIP source under csrc/, gfx950 only) and the
MEX wrappers' CI form (mex/ against the shim in mex/shim/).

    python qsm.m_amd/build.py            # library + MEX test driver
    python qsm.m_amd/build.py --lib      # library only

Outputs go to qsm.m_amd/lib/ (git-ignored, shipped with the tree).
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
MEX = PKG / "mex"
LIB = PKG / "lib"
OBJ = LIB / "obj"
INCLUDE = ROOT / "include"

ARCH = "gfx950"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: no FMA contraction, so the FP64 path reproduces the
# reference's CPU arithmetic bit for bit.  Packed-FP32 instructions are
# disabled (they have given wrong results on this hardware pool).
HIPFLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if err:
        print(err, file=sys.stderr)


def build_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    objs = []

    def one(src: Path) -> Path:
        obj = OBJ / (src.name + ".o")
        lang = [] if src.suffix == ".hip" else ["-x", "hip"]
        _run([HIPCC, *HIPFLAGS, *lang, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    out = LIB / "libqsm_hip.so"
    tmp = LIB / "libqsm_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_mex() -> Path | None:
    """Compiles the MEX wrappers against the shim into one shared object
    (mexFunction of each wrapper renamed) for tests driven from Python."""
    if not (MEX / "shim" / "mex_shim.c").exists():
        return None
    out = LIB / "libqsm_mex_ci.so"
    objs = []
    for src in sorted(MEX.glob("*_mex.c")):
        name = src.stem
        obj = OBJ / (name + ".o")
        _run(["gcc", "-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra",
              f"-I{MEX / 'shim'}", f"-I{INCLUDE}",
              f"-DmexFunction={name}_mexFunction", "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    shim_obj = OBJ / "mex_shim.o"
    _run(["gcc", "-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra",
          f"-I{MEX / 'shim'}", "-c", str(MEX / "shim" / "mex_shim.c"), "-o", str(shim_obj)])
    _run(["gcc", "-shared", "-o", str(out), *map(str, objs), str(shim_obj),
          f"-L{LIB}", "-lqsm_hip", f"-Wl,-rpath,$ORIGIN"])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    print("built", build_lib(a.j))
    if not a.lib:
        m = build_mex()
        if m:
            print("built", m)
    return 0


if __name__ == "__main__":
    sys.exit(main())
