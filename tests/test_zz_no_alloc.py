"""The allocation-free batched entry point against the allocating one (GPU).

pycwb_batch_run_noalloc runs the kernels of pycwb_batch_create/launch with the
work queues, counters and argmax keys in a library-held device array; its
regulator values, picks and statistics must equal the graph launch's bit for
bit, on two calls in a row (the second smaller, so stale scratch from the
first would show), and sizes beyond its maximum must be refused.
"""

import numpy as np
import pytest
import torch

from pycwb_amd import _lib, synth
from pycwb_amd.batch import BatchedSkyScan, net_tables, synth_batch
from pycwb_amd.device import lib

pytestmark = pytest.mark.gpu


def _run(nets, sizes, net_ids, seed, dev):
    tables = net_tables(nets, device=dev)
    batch = synth_batch(nets, sizes, net_ids, seed=seed, snr=3.0, device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    bs = BatchedSkyScan(batch, tables, valid, 0.5 * 0.5 * 2 / 3, 0.5, 0.5, use_graph=True)
    bs.launch()
    ref = bs.results()
    for t in (bs.REG1, bs.l_max, bs.stat_max, bs.reg_f64):
        t.fill_(-7)
    bs.launch_noalloc()
    got = bs.results()
    bs.close()
    return ref, got


def _same(ref, got):
    for k in ref:
        a, b = np.asarray(ref[k]), np.asarray(got[k])
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), k


def test_noalloc_matches_the_graph_launch_twice():
    dev = torch.device("cuda:0")
    nets = [synth.make_network(synth.NETWORK_3, order=5, K=24),
            synth.make_network(synth.NETWORK_4, order=5, K=24)]
    rng = np.random.default_rng(5)
    sizes = np.round(np.exp(rng.uniform(np.log(16), np.log(1500), 40))).astype(np.int64)
    ids = np.arange(40) % 2
    ref, got = _run(nets, sizes, ids, seed=11, dev=dev)
    _same(ref, got)
    assert np.all(ref["sky_stat_max"] > 0)
    # a second, smaller call right after: nothing of the first may leak through
    ref2, got2 = _run(nets, sizes[:9], ids[:9], seed=12, dev=dev)
    _same(ref2, got2)


def test_noalloc_refuses_more_than_its_maximum():
    dev = torch.device("cuda:0")
    cap = int(lib().pycwb_batch_noalloc_max_clusters())
    assert cap == 16384
    nets = [synth.make_network(synth.NETWORK_3, order=1, K=4)]
    tables = net_tables(nets, device=dev)
    batch = synth_batch(nets, np.full(cap + 1, 4), np.zeros(cap + 1, np.int64), seed=3,
                        device=dev)
    valid = torch.arange(nets[0].n_sky, dtype=torch.int64, device=dev)
    bs = BatchedSkyScan(batch, tables, valid, 0.1, 0.5, 0.5, use_graph=False)
    with pytest.raises(ValueError, match="exceed the maximum of 16384"):
        bs.launch_noalloc()
    bs.close()
