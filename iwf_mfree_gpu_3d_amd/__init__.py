"""MI355X (gfx950) implementation of the iwf_mfree_gpu_3d SPH cutting time step.

The product is the HIP library ``libmfree_hip.so`` (C ABI in ``include/mfree_gpu.h``); this
package holds its sources (``csrc/``), the in-tree build (``build.py``), ctypes bindings
(``_lib.py``), the reference-named host API (``api.py``) and workload set-up (``scenes.py``).
"""
from . import _lib  # noqa: F401

__all__ = ["api", "scenes", "build"]
