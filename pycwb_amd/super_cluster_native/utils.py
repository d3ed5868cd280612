"""Sub-network cut over the superclusters of a lag, as one GPU batch.

Drop-in for ``apply_subnet_cut`` and ``_top_loudest_indices`` of
pycwb/modules/super_cluster_native/utils.py:640-790.
"""

from __future__ import annotations

import logging
import time

import numpy as np

from .sub_net_cut import (_prepared_tables, _xtalk_csr, calculate_e2or_from_acore,
                          cluster_inputs, evaluate_subnet_cut)

logger = logging.getLogger(__name__)

__all__ = ["apply_subnet_cut", "calculate_e2or_from_acore", "_top_loudest_indices"]


def _top_loudest_indices(likelihood: np.ndarray, n_loudest: int) -> np.ndarray:
    """The first n_loudest of ``np.argsort(-likelihood)`` (utils.py:758-790 returns
    that ordering, using a partition only when it provably agrees)."""
    n_loudest = int(n_loudest)
    order = np.argsort(-np.asarray(likelihood))
    if n_loudest <= 0:
        return order[:0]
    return order[:n_loudest]


def apply_subnet_cut(superclusters, n_loudest_local, ml_local, FP_local, FX_local, acor_local,
                     e2or_local, n_ifo_local, n_sky_local, subnet_local, subcut_local,
                     subnorm_local, subrho_local, xtalk_local, arrays_prepared: bool = False):
    """utils.py:640-755: marks cluster_status -1 (passed) or 1 (failed) and returns
    the clusters with cluster_status <= 0.  Every cluster's sky loop runs in one
    batch; the cross-talk and MRA statistics run, as one more batch, only for
    clusters whose suball exceeds subnet."""
    timing = {"topk": 0.0, "data_prep": 0.0, "sky_scan": 0.0, "xtalk": 0.0, "mra": 0.0,
              "total": 0.0}
    t_total = time.perf_counter()
    ml_p, FP_p, FX_p = _prepared_tables(ml_local, FP_local, FX_local, arrays_prepared)
    rms_list, td_list, rows_list, live = [], [], [], []
    for i, c in enumerate(superclusters):
        t = time.perf_counter()
        rows = _top_loudest_indices(c.pixel_arrays.likelihood, n_loudest_local)
        timing["topk"] += time.perf_counter() - t
        if rows.size == 0:
            c.cluster_status = 1
            continue
        t = time.perf_counter()
        rms, td = cluster_inputs(c.pixel_arrays, rows)
        timing["data_prep"] += time.perf_counter() - t
        rms_list.append(rms)
        td_list.append(td)
        rows_list.append(rows)
        live.append(i)

    def csr_of(k):
        pa, rows = superclusters[live[k]].pixel_arrays, rows_list[k]
        return _xtalk_csr(xtalk_local, None, np.asarray(pa.layers)[rows],
                          np.asarray(pa.time)[rows])

    results = []
    if live:
        results = evaluate_subnet_cut(rms_list, td_list, ml_p, FP_p, FX_p, acor_local,
                                      e2or_local, n_ifo_local, subnet_local, subcut_local,
                                      subnorm_local, subrho_local, csr_of, timing=timing)
    for k, res in zip(live, results):
        c = superclusters[k]
        if res["subnet_passed"] and res["subrho_passed"] and res["subthr_passed"]:
            c.cluster_status = -1
        else:
            failed = [res[key] for key in ("subnet_condition", "subrho_condition",
                                           "subthr_condition")
                      if not res[key.replace("_condition", "_passed")]]
            logger.debug("Cluster %d failed %s", k, ", ".join(failed))
            c.cluster_status = 1
    timing["total"] = time.perf_counter() - t_total
    logger.info("   subnet timing detail      : clusters=%d topk=%.3fs data=%.3fs sky=%.3fs "
                "xtalk=%.3fs mra=%.3fs total=%.3fs", len(superclusters), timing["topk"],
                timing["data_prep"], timing["sky_scan"], timing["xtalk"], timing["mra"],
                timing["total"])
    return [c for c in superclusters if c.cluster_status <= 0]
