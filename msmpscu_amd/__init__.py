"""MI355X-native tabulated Finnis-Sinclair / EAM molecular-dynamics force path.

The compute lives in libmsm_hip.so (hand-written gfx950 HIP kernels behind the C-ABI of
include/msm_hip.h).  ``mdlib`` mirrors the reference MDLIB call surface on top of it and
``configs`` builds the benchmark workloads.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib", "mdlib", "configs"]
