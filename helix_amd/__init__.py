"""helix_amd -- an MI355X (gfx950) HEOM propagator with the public API of yjmaxpayne/HELIX.

Native parts (built in-tree by ``python -m helix_amd.build``):
  libhelix_hip.so   hand-written HIP kernels behind include/helix_hip.h
  libhelix.so       C++17 host set-up and the helix:: C++ API

This Python package is a thin ctypes front (``helix_amd.capi``) over the C entry points.
"""
from .capi import FP32, FP64, SPLIT_FP32, Batch, HelixError, Model, Plan, measure_peaks  # noqa: F401

__all__ = ["Model", "Plan", "Batch", "HelixError", "FP64", "FP32", "SPLIT_FP32", "measure_peaks"]
