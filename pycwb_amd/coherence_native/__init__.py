"""Coherence-stage pixel selection and clustering on the GPU
(pycwb/modules/coherence_native)."""

from .clustering import cluster_pixels, cluster_pixels_device, cluster_pixels_lags
from .selection import (_build_selection_cache, _shift_bins_from_lag_shifts,
                        select_network_pixels, select_network_pixels_device,
                        select_network_pixels_lags)

__all__ = ["_build_selection_cache", "_shift_bins_from_lag_shifts", "select_network_pixels",
           "select_network_pixels_device", "select_network_pixels_lags", "cluster_pixels",
           "cluster_pixels_device", "cluster_pixels_lags"]
