"""MI355X-native reconstruction hot path of adaptive tensorial density fields.

Layers:
  * ``lib/libatdf.so``   -- hand-written gfx950 HIP kernels behind the C ABI
    declared in ``include/atdf.h`` (``capi`` binds it with ctypes).
  * ``atdf_torch.so``    -- libtorch host layer mirroring the reference's C++
    operator API (HyperTreeBase, IndirectGridSample3D, IntegrateSamplesXRay,
    AccumulateSampleLossPerNode) plus the fused projection with its
    hand-written adjoint.  ``native()`` returns the loaded module.

There is no CPU fallback: if the native libraries are missing, ``native()``
raises.
"""
from __future__ import annotations

import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
EXT_PATH = os.path.join(HERE, "atdf_torch.so")
LIB_PATH = os.path.join(HERE, "lib", "libatdf.so")

_native = None


def native():
    """Load and return the libtorch host layer (``atdf_torch``)."""
    global _native
    if _native is None:
        import torch  # noqa: F401  (loads libtorch before the extension)

        if not os.path.exists(EXT_PATH) or not os.path.exists(LIB_PATH):
            raise ImportError(
                "adaptivetensordf_amd native libraries are not built "
                f"({EXT_PATH}, {LIB_PATH}); run `python -m adaptivetensordf_amd.build`")
        spec = importlib.util.spec_from_file_location("atdf_torch", EXT_PATH)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _native = mod
    return _native


__all__ = ["native", "EXT_PATH", "LIB_PATH"]
