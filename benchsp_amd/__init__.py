"""benchsp_amd: MI355X-native SpMV / block-SpMM library (HIP kernels for
gfx950 behind a C-ABI, libbsp.so).

Modules:
  host    deterministic synthetic matrices (numpy CSR)
  device  device-resident CSR and SpMV plans (torch memory, HIP kernels)
  build   compiles libbsp.so in-tree with hipcc
"""
from ._lib import BspError, LIB_PATH, lib  # noqa: F401

__all__ = ["BspError", "LIB_PATH", "lib"]
