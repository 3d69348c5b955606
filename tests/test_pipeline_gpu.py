"""The accelerator seam (accel.py:23-78, ``backend.focus(signal, tables)``) fed
by the O(m + n) table preparation: the GPU backend's image from
``pipeline.prepare_tables`` equals, bit for bit, its image from the oracle's
reference-shaped tables (full (m, n) freq array), and both are within the
float contract (peak-normalised RMS <= 1e-5) of the numpy backend's image
(form_image_pfa.pfa, the oracle).  The crop case reads the signal from the
returned raw sample offset, as the reference's caller does."""
import dataclasses

import numpy as np
import pytest

from oracle import ref_float as rfl

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def reader():
    from mi355x_pfa import synth
    spec = dataclasses.replace(synth.C1, pulses=301, samples=256, targets=4, seed=7)
    return synth.SceneReader(synth.make_scene(spec))


@pytest.mark.parametrize("crop", [0, 200])
@pytest.mark.parametrize("backend", ["gpu", "gpu-sinc32"])
def test_backend_on_compact_tables(reader, crop, backend):
    from mi355x_pfa.pfa import make_backend
    from mi355x_pfa.pipeline import prepare_tables
    meta = reader.cphd_meta
    t = prepare_tables(reader, meta, 1, 1, crop)
    tr = rfl.prepare_tables(reader, meta, 1, 1, crop)
    m, n = t["dims"]
    s0 = t["samp0"]
    sig = reader.read_chip((0, t["n_vec"], 1), (s0, s0 + n, 1))
    print(f"backend {backend} crop {crop}: {m}x{n} from sample {s0}", flush=True)
    be = make_backend(backend)
    a = be.focus(sig, t)
    b = be.focus(sig, tr)
    assert a.dtype == np.float32 and a.shape == (512, 256)
    np.testing.assert_array_equal(a, b)
    if backend == "gpu":
        img, _ = rfl.pfa(sig, tr["freq"], tr["ax"], tr["ay"], tr["sgn"])
        ref = rfl.detect(img).astype(np.float64)
        err = float(np.sqrt(np.mean((a - ref) ** 2)) / ref.max())
        print(f"  vs numpy backend: pn-RMS {err:.2e}", flush=True)
        assert err <= TOL
        assert np.unravel_index(np.argmax(a), a.shape) == np.unravel_index(np.argmax(ref), ref.shape)


def test_library_before_torch_shares_one_runtime():
    """A process that calls the library before importing torch still gets a
    working torch and holds one HIP runtime (regression: the library once bound
    the system runtime, and torch's own copy then found no device)."""
    import os
    import subprocess
    import sys
    from conftest import REPO
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {str(REPO / 'mpfs250t-sar-ifp_amd')!r})\n"
        "from mi355x_pfa.pfa import quicklook\n"
        "x = (np.arange(256 * 256) % 7).astype(np.complex64).reshape(256, 256)\n"
        "img = quicklook(x, -1)\n"
        "import torch\n"
        "t = torch.ones(4, device='cuda') * 2\n"
        "torch.cuda.synchronize()\n"
        "libs = sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l})\n"
        "print(len(libs), float(t.sum()), float(abs(img).max()))\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=180,
                       env=dict(os.environ))
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    n_rt, s, peak = r.stdout.split()[-3:]
    assert n_rt == "1" and float(s) == 8.0 and float(peak) > 0
