"""The allocation-free Monte Carlo entry point (phx_pair_noalloc, buffers from the library's static scratch
array) through the C ABI against the allocating one it mirrors (phx_pair_ex with device draws): identical
results on two calls in a row, the second smaller, and for both clone forms; a batch above the scratch limit
is refused with a message, before any device work, and the next call still runs."""
import numpy as np
import pytest

from phygas_amd import hip

pytestmark = pytest.mark.gpu


def _batch(workload, npairs, seed):
    cl = workload.make_clusters(500, 0)
    st = workload.make_stars(50, 4)
    pul = workload.make_pulsars(1, 1)
    tg = np.concatenate([st, pul])
    rng = np.random.default_rng(seed)
    ti = rng.integers(0, len(tg), npairs).astype(np.int32)
    ci = rng.integers(0, len(cl), npairs).astype(np.int32)
    return tg, cl, ti, ci


def test_noalloc_matches_the_allocating_entry_point(workload):
    calls = [(96, 400, hip.PLAIN, 1), (17, 150, hip.FAST, 2)]   # the second call smaller
    with hip.Engine(0) as e:
        for npairs, n_clones, form, seed in calls:
            tg, cl, ti, ci = _batch(workload, npairs, seed)
            got = hip.pair_noalloc(tg, cl, ti, ci, n_clones=n_clones, form=form)
            want = e.pair(tg, cl, ti, ci, n_clones=n_clones, form=form, device_draws=True)
            assert got.tobytes() == want.tobytes(), (npairs, n_clones, form)
            assert np.all(got["n"] > 0)


def test_noalloc_refuses_a_batch_above_the_limit(workload):
    limit = hip.lib().phx_pair_noalloc_limit()
    assert limit == 64 << 20
    tg, cl, ti, ci = _batch(workload, 600, 3)
    with pytest.raises(hip.PhygasError, match="scratch"):
        hip.pair_noalloc(tg, cl, ti, ci, n_clones=1000)        # 600,000 clone slots: ~79 MB
    small = hip.pair_noalloc(tg, cl, ti[:5], ci[:5], n_clones=64)
    with hip.Engine(0) as e:
        want = e.pair(tg, cl, ti[:5], ci[:5], n_clones=64, device_draws=True)
    assert small.tobytes() == want.tobytes()
