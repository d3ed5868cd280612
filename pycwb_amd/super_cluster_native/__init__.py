"""Sub-network cut of superclusters on the GPU (pycwb/modules/super_cluster_native)."""

from .sub_net_cut import (SubnetBatch, calculate_e2or_from_acore, evaluate_subnet_cut,
                          mra_statistics_from_td, optimze_sky_loc_from_td, sub_net_cut,
                          sub_net_cut_from_pixel_arrays)
from .utils import apply_subnet_cut

__all__ = ["SubnetBatch", "apply_subnet_cut", "calculate_e2or_from_acore",
           "evaluate_subnet_cut", "mra_statistics_from_td", "optimze_sky_loc_from_td",
           "sub_net_cut", "sub_net_cut_from_pixel_arrays"]
