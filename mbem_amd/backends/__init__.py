"""Backends (mbem's src/mbem/backends): the GPU dense backend."""
from .dense import AssembledDense, DenseBackend  # noqa: F401
