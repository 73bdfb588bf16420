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


def _oracle_final(mesh, U, B, nsteps):
    from oracle_ref import OracleSolver, settings_array

    ora = OracleSolver(mesh, settings_array(gamma0=5.0 / 3.0, cfl=0.8, slope_type=1.0), nthreads=8)
    Uo = U.reshape(mesh.num_total, 5, -1).copy()
    Bo = B.reshape(mesh.num_total, -1).copy()
    for _ in range(nsteps):
        Uo, Bo, _ = ora.step(Uo, Bo)
    return Uo, Bo


@pytest.mark.gpu
def test_served_requests_match_oracle():
    """One request of each kind (the field loop on its 2 x 1 brick, a 2D blast, a 3D blast),
    of the serving workload's smallest sizes, arriving while the others run: each served result
    after its full 200 steps (each with its own CFL dt, reduced on the device) against the
    oracle's 200 steps of the same request, relative L-inf <= 1e-12 per field."""
    from godunov_mhd_ct_amd.serving import Request, Server, prepare
    from parity import assert_parity

    reqs = [Request(0, 0.000, "loop2d", 64, 8, 1.0, 200, seed=1),
            Request(1, 0.002, "blast2d", 64, 16, 0.3, 200, seed=2),
            Request(2, 0.004, "blast3d", 32, 8, 4.0, 200, seed=3)]
    print("[serve] oracle parity:", [(r.kind, r.n, r.b, r.beta) for r in reqs], flush=True)
    srv = Server(hbm_budget=1 << 34, max_tick=20)
    out = srv.run(reqs, keep_results=True)
    assert out["requests"] == 3
    for r in sorted(out["done"], key=lambda q: q.rid):
        mesh, U, B, _ = prepare(r)
        Uo, Bo = _oracle_final(mesh, U, B, r.nsteps)
        Us, Bs = r.result
        errs, _, _ = assert_parity(Us, Bs, Uo, Bo, mesh.num_leaves, f"request {r.rid} ({r.kind})")
        print(f"  request {r.rid} {r.kind}: max rel L-inf {max(errs.values()):.3e}", flush=True)
