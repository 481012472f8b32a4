"""MI355X (gfx950) implementation of mbem's dense mollified-BEM hot path.

Host-side mirror of the reference's Python API (``kernels.tri_kernels``,
``model``, ``backends.dense``, ``la``, ``evaluate``) over hand-written HIP
kernels in ``libmbem_hip.so`` (C ABI: ``include/mbem_hip.h``).
"""
__version__ = "0.1.0"
