"""hipGraph replay of the fast pass: every replay, with eager work and synchronisation between replays, gives
the eager pass's rows byte for byte."""
import numpy as np
import pytest

from phygas_amd import hip

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(workload):
    cl = workload.make_clusters(900, 21)
    st = workload.make_stars(240, 22)
    with hip.Engine(0) as e:
        e.load_clusters(cl)
        e.load_stars(st)
        eager = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
        e.set_graphs(True)
        for i in range(4):
            got = e.screen(hip.FAST, chi2_max=25.0, p_min=0.0)
            assert got.tobytes() == eager.tobytes(), i
            # eager work between replays: a plain pass on the same context
            if i == 1:
                plain = e.screen(hip.PLAIN, chi2_max=25.0, p_min=0.0)
                assert len(plain) > 0
        caps, reps = e.graph_stats()
        assert caps == 1 and reps == 4
        # a parameter change re-captures
        got = e.screen(hip.FAST, chi2_max=9.0, p_min=0.0)
        want = eager[eager["chi2"] < 9.0]
        assert np.array_equal(got["cluster"], want["cluster"]) and np.array_equal(got["p_r50"], want["p_r50"])
        assert e.graph_stats()[0] == 2
