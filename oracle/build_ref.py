"""Build the reference's own C++ FFT model into oracle/_ref/ (test side only).

mpfs/fpga/fft_tb.cpp is the reference's bit-exact C++ restatement of
fixedpoint.fft1d_bfp (stdlib only).  It is compiled straight from the
reference checkout (PFA_REFERENCE, default /root/reference) when that exists;
on machines without the checkout this is a no-op and the tests that use the
binary skip.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_ref"


def build() -> Path | None:
    ref = Path(os.environ.get("PFA_REFERENCE", "/root/reference"))
    src = ref / "mpfs" / "fpga" / "fft_tb.cpp"
    if not src.exists():
        return None
    OUT.mkdir(exist_ok=True)
    exe = OUT / "fft_tb"
    subprocess.run(["g++", "-O2", "-std=c++14", str(src), "-o", str(exe)], check=True)
    return exe


if __name__ == "__main__":
    print(build())
    sys.exit(0)
