"""Coherence-stage pixel selection on the GPU (pycwb/modules/coherence_native)."""

from .selection import (_build_selection_cache, _shift_bins_from_lag_shifts,
                        select_network_pixels, select_network_pixels_lags)

__all__ = ["_build_selection_cache", "_shift_bins_from_lag_shifts", "select_network_pixels",
           "select_network_pixels_lags"]
