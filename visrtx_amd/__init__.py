"""visrtx_amd -- an MI355X-native ANARI ray-tracing device.

The rendering path is native: gfx950 HIP kernels behind the C ABI of
``include/vrx_hip.h`` (``libvisrtx_amd_hip.so``) and a C++ ANARI device on top
(``libanari_library_visrtx_amd.so``).  This Python package only binds them
(``anari``, ``_native``), provides the seeded synthetic scenes (``scenes``) and
the differentiable-volume autograd wrapper (``diffvol``).
"""
from ._native import VrxError, anari_lib, check, hip  # noqa: F401

__version__ = "0.1.0"
