"""MI355X-native Polar Format focus library (Python side of libmi355x_pfa.so).

Modules
  emulator  GPU drop-ins for mpfs/host/silicon_emulator.py (fixed point, bit-exact)
  tables    host-side table preparation (sinc taps, tapers, per-scene tables)
  synth     deterministic synthetic phase histories for tests and benchmarks
"""
from ._lib import PfaError, Context, context  # noqa: F401

__all__ = ["PfaError", "Context", "context"]
