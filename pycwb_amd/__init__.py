"""MI355X-native coherent sky-scan likelihood and WDM transform for PycWB.

Subpackages and modules:
  likelihoodWP  drop-in for pycwb.modules.likelihoodWP (setup_likelihood,
                likelihood, likelihood_wrapper, calculate_dpf,
                scan_sky_for_best_fit)
  wdm           WDM(M, K, beta_order, precision).t2w on the GPU
  skygrad       differentiable continuous-direction sky statistic and
                off-grid refinement
  batch         batched regulator + scan + argmax (one hipGraph launch)
  skystats      best-sky statistics of many clusters in one launch
The HIP kernels live in libpycwb_amd.so (C ABI: include/pycwb_amd.h); there
is no CPU fallback.
"""

__version__ = "0.1.0"
