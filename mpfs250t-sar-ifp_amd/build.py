"""Build libmi355x_pfa.so for gfx950 in-tree with hipcc.

    python mpfs250t-sar-ifp_amd/build.py [--jobs N] [--verbose]

Every source under csrc/ is compiled for gfx950 only.  Floating-point
contraction is off (the FP64 position arithmetic of the resample passes must
round each operation as numpy does), float32 division/sqrt are correctly
rounded, and packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / "csrc"
OUT = HERE / "mi355x_pfa" / "libmi355x_pfa.so"
BUILD = HERE / "_build"


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-I", str(REPO / "include"), "-I", str(CSRC)]
DEVICE = ["--offload-arch=gfx950", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    if not srcs:
        raise RuntimeError(f"no sources in {CSRC}")

    def compile_one(src: Path):
        obj = BUILD / (src.name + ".o")
        cmd = [cc, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
        if src.suffix == ".hip":
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (r.stderr or "").splitlines()
                        if "packed-fp32-ops" not in l and "warning: feature" not in l)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src.name}\n{r.stderr}")
        if verbose and msg.strip():
            print(msg)
        return obj

    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    out = build(a.jobs, a.verbose)
    print(f"built {out}")


if __name__ == "__main__":
    sys.exit(main())
