"""wavefield_imaging_amd: MI355X-native acoustic finite-difference time loop.

A from-scratch HIP (gfx950) implementation of the time loop behind the
Madagascar programs sfAWEFDcart2D/3D and sfAWEFDspher2D/3D of
Nacnud04/wavefield_imaging, exposed through the C ABI in include/awefd.h
(libawefd.so) and the C drop-in programs built from cli/.
"""
from . import _lib  # noqa: F401
from .awefd import Model, make_params, run  # noqa: F401

__all__ = ["Model", "make_params", "run"]
