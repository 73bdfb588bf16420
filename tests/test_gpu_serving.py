"""GPU test of the serving loop: continuous batching over a Poisson stream of requests of
different kinds and sizes, then every request re-run alone -- the results must be the same
bit for bit (a request's result does not depend on the rest of the batch)."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_serving_results_equal_alone():
    from godunov_mhd_ct_amd.batch import BatchGodunovMHD
    from godunov_mhd_ct_amd.serving import Server, prepare, request_stream

    reqs = request_stream(8, rate=400.0, seed=11, nsteps=12, sizes2d=(32, 64), sizes3d=(32,))
    print("[serve] requests:", [(r.kind, r.n, r.b) for r in reqs], flush=True)
    # a budget that holds only a few requests at once, so admissions wait for retirements
    budget = 3 * max(BatchGodunovMHD.footprint_bytes(prepare(r)[0]) for r in reqs)
    srv = Server(hbm_budget=budget, max_tick=5)
    out = srv.run(reqs, keep_results=True)
    assert out["requests"] == len(reqs)
    assert out["latency_ms_p50"] is not None and out["latency_ms_p99"] >= out["latency_ms_p50"]
    for r in out["done"]:
        mesh, U, B, st = prepare(r)
        bt = BatchGodunovMHD(mesh.dim)
        sid = bt.add(mesh, U, B, st)
        bt.step(r.nsteps)
        Ua, Ba = bt.state(sid)
        Us, Bs = r.result
        assert np.array_equal(Ua, Us) and np.array_equal(Ba, Bs), f"request {r.rid} differs alone vs served"
