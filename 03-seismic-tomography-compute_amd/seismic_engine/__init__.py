"""seismic_engine — MI355X (gfx950) drop-in for the reference's pybind11 module
of the same name (src/cpp/pybind/bindings.cpp:13).

All compute runs in libseismic_hip.so through its C ABI; torch provides
device memory and streams.  Importing this module does not touch the GPU;
the first compute call loads the library and fails loudly if it is missing.
"""
from ._lib import SeismicError, lib  # noqa: F401
from .fd import FDPropagator  # noqa: F401
from . import synthetic  # noqa: F401

__all__ = ["SeismicError", "lib", "FDPropagator", "synthetic"]
