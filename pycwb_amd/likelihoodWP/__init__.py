"""Drop-in GPU likelihood stage (pycwb/modules/likelihoodWP).

Same names and signatures as the reference module:
``setup_likelihood``, ``likelihood``, ``likelihood_wrapper`` (likelihood.py),
``calculate_dpf`` (dpf.py) and ``scan_sky_for_best_fit`` (sky_scan.py).
"""

from .dpf import calculate_dpf
from .likelihood import (evaluate_cluster_likelihood, evaluate_fragment_clusters, likelihood,
                         likelihood_wrapper, prepare_likelihood_inputs, setup_likelihood)
from .sky_scan import find_optimal_sky_localization, scan_sky_for_best_fit

__all__ = ["calculate_dpf", "evaluate_cluster_likelihood", "evaluate_fragment_clusters",
           "find_optimal_sky_localization", "likelihood", "likelihood_wrapper",
           "prepare_likelihood_inputs", "scan_sky_for_best_fit", "setup_likelihood"]
