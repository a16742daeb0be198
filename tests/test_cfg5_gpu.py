"""Serving at the workload's sizes: Poisson arrivals (seeded) of Laplace
(interior, linear basis, CG) and Helmholtz (exterior, Brakhage-Werner form,
k in [1, 6], GMRES(30)) Dirichlet solves with one to four point sources, on
meshes of 2048 (sphere), 6912 (cube) and 131072 (sphere) triangles, H2 by
interpolation of order 4, leaf size 64; operators cached on the device.

Checks: every request completes, requests overlap on the device (admitted
before others finished), and each result is the same alone and in the mix:
solved again alone with the one-system solver on the same operators it
gives the same steps and the same bits (reproducible mode), and its true
residual meets the stopping rule.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_serving_large_meshes_alone_equals_mix():
    import h2lib_amd as h
    from h2lib_amd import geometry as G
    from h2lib_amd.serve import Server, draw_requests, mesh_bounds
    from h2lib_amd.solve import solve_cg, solve_gmres
    meshes = {"sphere2048": G.sphere(16), "cube6912": G.cube(24),
              "sphere131072": G.sphere(128)}
    bounds = {k: mesh_bounds(m) for k, m in meshes.items()}
    reqs = draw_requests(12, 40.0, sorted(meshes), bounds, seed=21,
                         ks=(1.0, 3.5, 6.0))
    print("requests", [(r.mesh_id, r.kind, r.k) for r in reqs], flush=True)
    srv = Server(meshes, 200e9, eps=1e-10, maxiter=3000)
    srv.run(reqs, timeout=600)
    assert all(r.done is not None and r.x is not None for r in reqs)
    assert {r.mesh_id for r in reqs} == set(meshes)
    assert any(a.admitted < b.done and a.rid != b.rid
               for a in reqs for b in reqs)
    assert srv.evictions == 0
    lat = np.array([r.done - r.arrival for r in reqs])
    print("latency p50 %.0f ms p99 %.0f ms" % (
        1e3 * np.percentile(lat, 50), 1e3 * np.percentile(lat, 99)),
        flush=True)
    for r in reqs:
        e = srv.cache[r.key]
        f, b = e.prob.rhs_sources(r.sources, e.stream)
        x = h.DeviceArray(e.prob.n, b.dtype)
        x.zero(e.stream)
        if e.batch.method == "cg":
            it, _ = solve_cg(e.op, b, x, 1e-10, 3000, stream=e.stream)
        else:
            it, _ = solve_gmres(e.op, b, x, 1e-10, 3000, 30, e.stream)
        e.stream.sync()
        xa = x.download()
        rr = h.DeviceArray.from_host(b.download())
        e.op.addeval(-1.0, x, rr, e.stream)
        e.stream.sync()
        true = np.linalg.norm(rr.download()) / np.linalg.norm(b.download())
        print("request", r.rid, r.key, "steps", r.steps, it,
              "true residual %.1e" % true, flush=True)
        assert int(r.steps) == int(it), r.rid
        assert np.array_equal(r.x, xa), r.rid
        assert true <= 1.5e-10, r.rid
