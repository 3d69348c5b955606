"""CPU oracle for the Polar Format focus path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy, the reference's algorithms for the PFA focus
path so that the GPU library can be checked against them:

* ``ref_fixed`` -- the bit-exact fixed-point datapath of
  ``mpfs/host/silicon_emulator.py`` (with ``src/fixedpoint.py``,
  ``mpfs/host/serialize_inputs.py:67-94`` and
  ``mpfs/fpga/tb/gen_resample_vectors.py:198-296``).
* ``ref_float`` -- the float focus of ``src/form_image_pfa.py:144-246`` plus the
  SINC32 float variant and float64 forms used as error references.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results.  The product
library never calls it and has no CPU fallback through it.
"""
