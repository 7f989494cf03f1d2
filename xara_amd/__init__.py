"""MI355X-native state determination for xara models.

The package exposes the reference's model/analysis command interface for the
nonlinear state-determination path (``Model``) over the gfx950 library built
from ``xara_amd/csrc`` (C-ABI in ``include/xara_gpu.h``).
"""
from .model import Model, XaraError  # noqa: F401
from . import meshes  # noqa: F401

__all__ = ["Model", "XaraError", "meshes"]
