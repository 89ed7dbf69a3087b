"""The serving loop: requests of mixed 2D / 3D sizes arriving over time are admitted against an
HBM budget into at most two concurrent runs (each on its own stream and host thread); every
request gives bit-identical fields alone and in the mix; the budget and the slot count are
never exceeded and a request that does not fit waits."""
from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests():
    from phipict_amd.serving import DuctRequest

    shapes = [(48, 24), (20, 12, 12), (64, 32), (24, 16, 12)]
    return [DuctRequest(rid=i, cells=c, re=100.0 + 150.0 * i, ha=10.0 + 7.0 * i, steps=6, dt=5e-4, tol=1e-8,
                        dtype=torch.float32, seed=20 + i, arrival=0.05 * i) for i, c in enumerate(shapes)]


def test_requests_equal_alone_and_in_the_mix(monkeypatch):
    from phipict_amd import serving
    from phipict_amd.solvers import amg

    monkeypatch.setattr(amg, "AUTOTUNE", False)  # lanes by rule: the same kernels in every run
    reqs = _requests()
    fp = sorted(serving.estimate_footprint(r) for r in reqs)
    budget = fp[-1] + fp[-2]  # the two largest fit together, not three requests
    print(f"[serving] 4 requests, two slots, budget {budget / 2**20:.0f} MiB", flush=True)
    loop = serving.ServingLoop(budget, max_slots=2, keep_fields=True)
    mixed = loop.run(reqs)
    assert all(r.error is None for r in mixed), [r.error for r in mixed]
    assert loop.peak_in_use <= budget and loop.peak_running <= 2
    assert any(r.wait > 0.0 for r in mixed)
    lat = serving.latency_percentiles(mixed)
    print(f"  mixed: p50 {lat['p50']:.3f} s, p99 {lat['p99']:.3f} s", flush=True)
    for req, rec in zip(reqs, mixed):
        alone = serving.ServingLoop(budget, max_slots=1, keep_fields=True).run(
            [serving.DuctRequest(**{**req.__dict__, "arrival": 0.0})])[0]
        assert alone.error is None
        for k in ("velocity", "pressure", "potential"):
            assert torch.equal(alone.fields[k], rec.fields[k]), (req.rid, k)
        v = rec.fields["velocity"]
        assert torch.isfinite(v).all() and float(v[0, 0].abs().max()) > 0


def test_request_beyond_budget_is_refused():
    from phipict_amd import serving

    req = _requests()[0]
    rec = serving.ServingLoop(1 << 20, max_slots=1).run([req])[0]
    assert rec.error == "footprint exceeds the HBM budget"
