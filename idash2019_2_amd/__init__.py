"""MI355X-native secure genotype imputation (the cloud step of
ssmiler/idash2019_2 and its encrypt/decrypt companions).

The compute path is libidash_hip.so (hand-written gfx950 kernels behind the C
ABI in include/idash_hip.h); this package is its Python front end. The
idash-compatible C++ host (keygen/encrypt/cloud/decrypt) lives in host/.
"""
from . import _abi
from .engine import (Band, Batch, DeviceArray, Engine, Key, PinnedArray, Server, band_plan,
                     layout_for, noise_table)

__all__ = ["Engine", "Band", "Batch", "Server", "PinnedArray", "DeviceArray", "Key", "band_plan", "layout_for", "noise_table", "_abi"]
