"""Serving loop on the GPU: every request's result is the same alone and in a concurrent mix."""
import numpy as np
import pytest

from phygas_amd import hip
from phygas_amd import serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [hip.FAST, hip.PLAIN])
def test_results_same_alone_and_in_mix(workload, form):
    cl = workload.make_clusters(600, 11)
    chunks = [workload.make_stars(n, 40 + i, prefix=f"M{i}S") for i, n in enumerate((80, 300, 17, 150))]
    jobs = [(workload.make_pulsars(1, 50 + i), np.arange(i * 7, i * 7 + 25, dtype=np.int32)) for i in range(3)]
    # alone: one engine call per request
    alone_rows, alone_pairs = [], []
    with hip.Engine(0) as e:
        e.load_clusters(cl, 64, 2.0)
        for st in chunks:
            e.clear_stars()
            e.load_stars(st)
            alone_rows.append(e.screen(form, chi2_max=25.0, p_min=0.0))
        for pul, idx in jobs:
            alone_pairs.append(e.pair(pul, cl, np.zeros(len(idx), np.int32), idx, n_clones=120, dt_myr=0.1,
                                      form=hip.PLAIN if form == hip.PLAIN else hip.FAST, device_draws=True))
    # in the mix: everything submitted at once to a server with several slots
    with serve.Server(cl, form=form, screen_slots=2, pair_slots=1, max_batch_stars=400, p_min=0.0) as srv:
        ids = []
        for i in range(max(len(chunks), len(jobs))):
            if i < len(chunks):
                ids.append(("s", i, srv.submit_screen(chunks[i])))
            if i < len(jobs):
                ids.append(("p", i, srv.submit_pair(jobs[i][0], jobs[i][1], n_clones=120, dt_myr=0.1)))
        batched = 0
        for kind, i, rid in ids:
            assert srv.wait(rid, timeout=120.0)
            inf = srv.info(rid)
            assert inf["status"] == 2 and inf["arrive"] <= inf["start"] <= inf["done"]
            batched += inf["batch"] > 1
            if kind == "s":
                got = srv.rows(rid)
                assert got.tobytes() == alone_rows[i].tobytes(), i
            else:
                got = srv.pair_results(rid)
                assert got.tobytes() == alone_pairs[i].tobytes(), i
            srv.release(rid)
    assert batched > 0     # some requests shared a batch


def test_oversized_request_runs_alone(workload):
    cl = workload.make_clusters(200, 12)
    st = workload.make_stars(120, 13)
    with serve.Server(cl, form=hip.FAST, screen_slots=1, pair_slots=0, hbm_budget=1 << 20, p_min=0.0) as srv:
        a = srv.submit_screen(st)
        b = srv.submit_screen(st[:10])
        assert srv.wait(a, 120.0) and srv.wait(b, 120.0)
        assert srv.info(a)["batch"] == 1 and srv.info(b)["batch"] == 1
        assert np.array_equal(srv.rows(b), srv.rows(a)[srv.rows(a)["star"] < 10])


@pytest.mark.parametrize("form", [hip.FAST, hip.PLAIN])
def test_serving_shape_alone_and_in_mix(workload, form):
    """The serving work unit's own shapes: the 4,087-cluster tables, a 20,000-star and a 256-star screening chunk,
    pulsar jobs of 600 and 100 clusters x 1,000 clones, all submitted at once; each result equals the same
    request run alone."""
    cl = workload.make_clusters(4087, 0)
    chunks = [workload.make_stars(20000, 71, prefix="BIGS"), workload.make_stars(256, 72, prefix="SMLS")]
    rng = np.random.default_rng(5)
    jobs = [(workload.make_pulsars(1, 73 + i), np.sort(rng.choice(len(cl), k, replace=False)).astype(np.int32))
            for i, k in enumerate((600, 100))]
    pform = hip.PLAIN if form == hip.PLAIN else hip.FAST
    alone_rows, alone_pairs = [], []
    with hip.Engine(0) as e:
        e.load_clusters(cl, 64, 2.0)
        for st in chunks:
            e.clear_stars()
            e.load_stars(st)
            alone_rows.append(e.screen(form, chi2_max=25.0, p_min=0.001))
        for pul, idx in jobs:
            alone_pairs.append(e.pair(pul, cl, np.zeros(len(idx), np.int32), idx, n_clones=1000, dt_myr=0.05,
                                      form=pform, device_draws=True))
    assert len(alone_rows[0]) > 100
    with serve.Server(cl, form=form, screen_slots=2, pair_slots=1, p_min=0.001) as srv:
        ids = [("s", 0, srv.submit_screen(chunks[0])), ("p", 0, srv.submit_pair(jobs[0][0], jobs[0][1], 1000, 0.05)),
               ("s", 1, srv.submit_screen(chunks[1])), ("p", 1, srv.submit_pair(jobs[1][0], jobs[1][1], 1000, 0.05))]
        for kind, i, rid in ids:
            assert srv.wait(rid, timeout=300.0)
            if kind == "s":
                assert srv.rows(rid).tobytes() == alone_rows[i].tobytes(), i
            else:
                assert srv.pair_results(rid).tobytes() == alone_pairs[i].tobytes(), i
            srv.release(rid)
