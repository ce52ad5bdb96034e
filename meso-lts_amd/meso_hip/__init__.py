"""meso_hip: host-side Python access to the MI355X DUGKS/CDUGKS step.

The compute path is the HIP library ``lib/libmeso_hip.so`` (gfx950 kernels
behind the C ABI of include/meso_hip.h); this package only binds it and
writes meso input files.
"""
from .capi import (Case, Device, MesoError, MODEL_CDUGKS_BGK, MODEL_CDUGKS_SHAKHOV, MODEL_DUGKS_BGK,
                   STORAGE_FP32, STORAGE_FP64, LAYOUT_CELL_MAJOR, LAYOUT_VELOCITY_MAJOR, LIB_PATH)

__all__ = ["Case", "Device", "MesoError", "MODEL_CDUGKS_BGK", "MODEL_CDUGKS_SHAKHOV", "MODEL_DUGKS_BGK",
           "STORAGE_FP32", "STORAGE_FP64", "LAYOUT_CELL_MAJOR", "LAYOUT_VELOCITY_MAJOR", "LIB_PATH"]
