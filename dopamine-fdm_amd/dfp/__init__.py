"""dfp: MI355X-native pressure projection for dopamine-fdm (host side).

The compute path is libdfp.so (HIP kernels for gfx950 behind the C ABI of
include/dfp.h); this package binds it with ctypes and mirrors the
reference's ``projection`` module interface.  There is no CPU fallback: if
the library is missing, every call raises.
"""
from ._lib import (DFP_ASYNC, DFP_CHECK_FINITE, DFP_FIELDS_ON_DEVICE, DFP_FP32_REFINE, DFP_FP64, DFP_SAVE_P,
                   DFP_SOLVE_ONLY, DfpError, LIB_PATH, HEADER_PATH, lib)
from .grid import ChannelGrid, channel_grid, duct_grid, namelist_grid, periodic_y_grid, zwall_periodic_y_grid
from .projection import RK_COEF_DIAG, ChannelState, Plan, Projection
from .synth import refresh_velocity_ghosts, synth_fields
from .hipmem import DeviceArray, Event, device_count, synchronize

__all__ = [
    "DFP_ASYNC", "DFP_CHECK_FINITE", "DFP_FIELDS_ON_DEVICE", "DFP_FP32_REFINE", "DFP_FP64", "DFP_SAVE_P",
    "DFP_SOLVE_ONLY", "DfpError", "LIB_PATH", "HEADER_PATH", "lib", "ChannelGrid", "channel_grid", "duct_grid", "namelist_grid", "periodic_y_grid", "zwall_periodic_y_grid", "RK_COEF_DIAG", "ChannelState",
    "Plan", "Projection", "refresh_velocity_ghosts", "synth_fields", "DeviceArray", "Event", "device_count",
    "synchronize",
]
