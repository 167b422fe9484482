"""The request loop on the device: requests run concurrently on per-slot streams give the
same results as the same requests run alone (every kernel reduces in a fixed order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_requests_in_the_mix_match_alone():
    import torch
    torch.cuda.set_device(0)
    from sucemfem.Serving import Request, Server, device_executor, summarize
    print("serving: mix vs alone", flush=True)
    dims = [(0.3, 0.2, 0.25), (0.5, 0.4, 0.3), (0.25, 0.25, 0.4)]
    trace = []
    for i in range(8):
        d = dims[i % 3]
        kind = "eigen" if i % 2 == 0 else "driven"
        order = 1 + (i // 2) % 2
        n = 3 + i % 3
        if kind == "eigen":
            lo = (np.pi / max(d)) ** 2 + (np.pi / sorted(d)[1]) ** 2
            trace.append(Request(i, kind, d, n, order, 0.0, nev=3, sigma=0.5 * lo))
        else:
            trace.append(Request(i, kind, d, n, order, 0.0, freq=3e8 / (10 * max(d) / n)))
    # one large request: the 40^3-box order-2 cube (2,480,240 DOFs), driven
    trace.append(Request(len(trace), "driven", (1.0, 1.0, 1.0), 40, 2, 0.0, freq=1e9))
    for precision in (64, 32):
        # FP32: driven solves and the eigen requests' inner solves in FP32 with FP64
        # refinement (the large request stays in the FP64 mix only)
        reqs = trace if precision == 64 else trace[:-1]
        ex = device_executor(tol=1e-8, max_iter=20000, precision=precision)
        results, makespan = Server(ex, slots=3).run(reqs, realtime=False)
        s = summarize(results, makespan)
        assert s["done"] == len(reqs), [r.error for r in results if r.status != "done"]
        alone = Server(ex, slots=1)
        for r, req in zip(results, reqs):
            (ra,), _ = alone.run([req], realtime=False)
            assert ra.status == "done", ra.error
            assert np.array_equal(np.asarray(r.value), np.asarray(ra.value)), (precision, req)


def test_eigen_requests_in_the_mix_match_oracle():
    """Eigen requests at the configured box counts (n = 6..8 of [6, 48]; orders 1 and 2, 4
    modes each) served concurrently (2 slots): each k^2 equals the oracle pencil's on the
    same box mesh (the device's UnitCube, checked equal to the oracle's) to 1e-8."""
    import torch
    torch.cuda.set_device(0)
    import oracle as O
    from sucemfem.Mesh import Mesh
    from sucemfem.Serving import Request, Server, device_executor, summarize
    print("serving: eigen requests vs oracle", flush=True)
    cases = [((0.3, 0.2, 0.25), 6, 1), ((0.5, 0.4, 0.3), 7, 2), ((0.25, 0.25, 0.4), 8, 1),
             ((0.4, 0.3, 0.35), 6, 2)]
    trace = []
    for i, (d, n, order) in enumerate(cases):
        lo = (np.pi / max(d)) ** 2 + (np.pi / sorted(d)[1]) ** 2
        trace.append(Request(i, "eigen", d, n, order, 0.0, nev=4, sigma=0.5 * lo))
    ex = device_executor(tol=1e-8, max_iter=20000, precision=64)
    results, makespan = Server(ex, slots=2).run(trace, realtime=False)
    s = summarize(results, makespan)
    assert s["done"] == len(trace), [r.error for r in results if r.status != "done"]
    for r, req in zip(results, trace):
        co, ce = O.box_mesh(req.dims, req.n)
        m = Mesh.box_on_device(req.n, scale=req.dims)
        assert np.array_equal(np.sort(np.asarray(m.cells()), axis=1), np.sort(ce, axis=1))
        assert np.allclose(np.asarray(m.coordinates()), co, rtol=0, atol=1e-15)
        ref, _, _ = O.cavity_eigs(O.Space(co, ce, req.order), req.nev, req.sigma)
        k2 = np.asarray(r.value)
        assert r.stats["block"] == 2, r.stats
        assert np.all(np.abs(k2 - ref) / ref <= 1e-8), (req, k2, ref)
