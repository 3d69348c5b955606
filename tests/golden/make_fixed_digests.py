"""Write tests/golden/fixed_digests.json: SHA-256 digests of the oracle's
fixed-point focus (oracle/ref_fixed.form_image_fixed, the restatement of
silicon_emulator.form_image pinned by the reference fixtures) at the
production shapes, on the deterministic inputs of tests/fixed_inputs.py.

    python tests/golden/make_fixed_digests.py C3 11 [C2 11 ...]

Each case is a config (C3: 8167 x 8192, C2: 5634 x 4319; grid 8192) and the
interpolators of the two passes (range, azimuth; 1 = 32-tap sinc, 0 = lerp).
The digests of the input codes and of every scene table are stored too, so a
test that regenerates them elsewhere first proves it has the same inputs.
Runs on the CPU in minutes per case; results are merged into the JSON file.
"""
from __future__ import annotations

import fcntl
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
for p in (str(REPO), str(REPO / "mpfs250t-sar-ifp_amd"), str(REPO / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import fixed_inputs as fi  # noqa: E402
from oracle import ref_fixed as rf  # noqa: E402

OUT = HERE / "fixed_digests.json"
TABLE_KEYS = ("f0", "df", "pr", "tan_s", "inv_order", "krp", "kcp", "hamr", "hamc")


def run_case(config, tag):
    m, n, tabs = fi.scene_tables(config)
    seed = {"C3": 3, "C2": 2}[config]
    t0 = time.time()
    iq = fi.phase_history_iq(m, n, seed)
    print(f"{config}: input {time.time() - t0:.1f} s", flush=True)
    rs, az = tag[0] == "1", tag[1] == "1"
    t0 = time.time()
    img, ea, eb = rf.form_image_fixed(iq[..., 0].astype(np.int64), iq[..., 1].astype(np.int64),
                                      tabs, 8192, rs, az)
    print(f"{config} {tag}: oracle {time.time() - t0:.1f} s", flush=True)
    return {"config": config, "m": m, "n": n, "grid": 8192, "seed": seed,
            "range_sinc": rs, "az_sinc": az,
            "input_sha256": fi.sha(iq),
            "tables_sha256": {k: fi.sha(tabs[k]) for k in TABLE_KEYS},
            "image_sha256": fi.sha(img.astype(np.uint16)),
            "exp_a_sha256": fi.sha(np.asarray(ea, np.int32)),
            "exp_b_sha256": fi.sha(np.asarray(eb, np.int32)),
            "image_max": int(img.max()), "emax_a": int(np.max(ea)), "emax_b": int(np.max(eb))}


def main(argv):
    cases = list(zip(argv[0::2], argv[1::2]))
    for config, tag in cases:
        res = run_case(config, tag)
        with open(HERE / ".fixed_digests.lock", "w") as lk:     # several cases may run at once
            fcntl.flock(lk, fcntl.LOCK_EX)
            data = json.loads(OUT.read_text()) if OUT.exists() else {}
            data[f"{config}_{tag}"] = res
            OUT.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
        print("wrote", f"{config}_{tag}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["C3", "11"])
