"""Served solve stream: results are the same alone and in the mix, and the
HBM budget only delays requests."""
import numpy as np
import pytest

from tsem_amd import serve

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _reqs():
    return serve.make_requests(6, seed=3, n_range=(3, 5), p_range=(3, 5))


def test_results_same_alone_and_in_the_mix(ctx):
    _say("serve mix")
    mixed = serve.Server(slots=3, keep_solutions=True)
    reqs, summary = mixed.run(_reqs())
    mixed.close()
    assert summary["max_relres"] <= 1e-10
    assert len({r.worker for r in reqs}) > 1  # the work was spread over several streams
    alone = serve.Server(slots=1, keep_solutions=True)
    for r in _reqs():
        got, _ = alone.run([r])
        ref = reqs[r.id].solution
        # same space, geometry and deterministic kernels: bit-identical
        assert np.array_equal(got[0].solution, ref), r
    alone.close()


def test_budget_serializes_and_evicts(ctx):
    _say("serve budget")
    reqs = _reqs()
    small = max(serve.footprint(r.n, r.p) for r in reqs)
    s = serve.Server(slots=3, hbm_budget=small)
    reqs, summary = s.run(reqs)
    assert summary["max_relres"] <= 1e-10
    # with room for one space at a time, no two requests overlap
    spans = sorted((r.start, r.done) for r in reqs)
    assert all(spans[i + 1][0] >= spans[i][1] - 1e-6 for i in range(len(spans) - 1))
    assert s.in_use <= small
    s.close()


def test_foreground_and_background_classes(ctx):
    """With small topologies foreground and large ones background (large
    solves pause between chunks while a small one runs), every request still
    gets the bit-identical result it gets alone."""
    _say("serve classes")
    reqs = serve.make_requests(8, seed=5, n_range=(3, 6), p_range=(3, 6))
    costs = sorted(serve.cost(r.n, r.p) for r in reqs)
    split = costs[len(costs) // 2]
    assert costs[0] <= split < costs[-1]
    mixed = serve.Server(slots=4, small_slots=2, small_cost=split, keep_solutions=True)
    got, summary = mixed.run(reqs)
    mixed.close()
    assert summary["max_relres"] <= 1e-10
    alone = serve.Server(slots=1, keep_solutions=True, classes=False)
    for r in serve.make_requests(8, seed=5, n_range=(3, 6), p_range=(3, 6)):
        (one,), _ = alone.run([r])
        assert np.array_equal(one.solution, got[r.id].solution), r
    alone.close()
