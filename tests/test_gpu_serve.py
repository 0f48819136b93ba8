"""Continuous-batching server on the GPU: each request's result is the same
when it runs alone and when it runs in a mix with other requests in flight
(own context and stream per request; the kernels' reductions have a fixed
order, so the agreement is exact, and the test allows 1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def _requests():
    from doublyexcited_amd.serve import Request

    specs = [(2, 0.5, 0.30, 102), (3, 0.5, 0.25, 7), (2, 0.45, 0.35, 11), (3, 0.55, 0.30, 5)]
    return [Request(arrival=0.01 * i, rid=i, n=n, delta=dl, theta=th, sigma=-0.78 - 0.002j, nev=4,
                    tol=1e-8, maxit=400, seed=sd, inner_iter=4, guard=2)
            for i, (n, dl, th, sd) in enumerate(specs)]


def test_results_alone_equal_results_in_the_mix(dex):
    from doublyexcited_amd.serve import Server

    print("case: serve alone vs mix", flush=True)
    reqs = _requests()
    alone = [Server(slots=1).run([r], realtime=False)[0] for r in reqs]
    mixed = Server(slots=4).run(reqs, realtime=False)
    for a, b in zip(alone, mixed):
        assert a.ok and b.ok, (a.error, b.error)
        assert a.iterations == b.iterations
        assert a.converged_pairs == b.converged_pairs
        np.testing.assert_allclose(b.e, a.e, rtol=0, atol=1e-12)
        np.testing.assert_allclose(b.residuals, a.residuals, rtol=0, atol=1e-12)
    # every converged pair meets the tolerance
    for r in mixed:
        assert np.sum(r.residuals <= 1e-8) == r.converged_pairs


@pytest.mark.parametrize("n", [2, 5, 7])
def test_admission_footprint_is_the_library_count(dex, n):
    """The server's admission estimate equals the library's own device
    footprint (dex_gplmr_footprint), batched matvec intermediates included
    (NN = 5 and 11 batch 16 columns, NN = 15 batches 2)."""
    from doublyexcited_amd.serve import footprint_bytes

    print(f"case: footprint N={n}", flush=True)
    for prec in ("fp64", "mixed"):
        with dex.HeliumGrid(n, 0.5, 0.3, precision=prec) as g:
            for nev, ns, inner, ib, guard in ((4, 1, 0, 0, 0), (8, 4, 8, 0, 2), (16, 1, 4, 4, 0)):
                assert g.footprint(nev, ns, inner, ib, guard) == \
                    footprint_bytes(n, nev, prec, ns, inner, ib, guard)


def test_stop_ends_a_solve_at_an_iteration_boundary(dex):
    """dex_ctx_stop: set before a solve, the solve returns at its first
    boundary with the start block's Rayleigh quotients; set from another
    thread during a solve, it returns early; cleared, solves run to MAXIT."""
    import threading
    import time

    print("case: stop flag", flush=True)
    sigma = -0.78 - 0.002j
    for prec in ("fp64", "mixed"):
        with dex.HeliumGrid(3, 0.5, 0.3, precision=prec) as g:
            x = torch.empty(g.d * 4, dtype=torch.complex128, device="cuda")
            g.fill_gaussian_device(x, 4, seed=21, symmetry=1)
            g.stop()
            r = g.gplmr_eig_device(x, sigma, 4, ns=1, maxit=50, tol=1e-14, symmetry=1)
            assert r.iterations == 0 and not r.converged
            assert np.all(np.isfinite(r.e))
            g.stop(False)
            r = g.gplmr_eig_device(x, sigma, 4, ns=1, maxit=3, tol=1e-14, symmetry=1)
            assert r.iterations == 3
            g.stop(False)
    # from another thread, mid-solve (NN=11): it returns, converged or stopped, long before MAXIT
    with dex.HeliumGrid(5, 0.5, 0.3) as g:
        x = torch.empty(g.d * 4, dtype=torch.complex128, device="cuda")
        g.fill_gaussian_device(x, 4, seed=22, symmetry=1)
        t = threading.Timer(0.3, g.stop)
        t.start()
        t0 = time.time()
        r = g.gplmr_eig_device(x, sigma, 4, ns=1, maxit=100000, tol=1e-12, symmetry=1)
        t.join()
        assert 0 < r.iterations < 100000 and time.time() - t0 < 60.0


def test_results_alone_equal_results_in_a_mix_of_grids_11_to_21(dex):
    """The same at the served stream's grid sizes: NN=11, 15 and 21 (block 4,
    a few iterations each, so the test stays short) alone and together; a
    request's result does not depend on what else is on the device."""
    from doublyexcited_amd.serve import Request, Server

    print("case: serve alone vs mix, NN=11/15/21", flush=True)
    specs = [(5, 0.5, 0.30, 6), (10, 0.42, 0.39, 3), (7, 0.45, 0.35, 4)]
    reqs = [Request(arrival=0.0, rid=i, n=n, delta=dl, theta=th, sigma=-0.69 - 0.002j, nev=4, tol=1e-7,
                    maxit=its, seed=31 + i, inner_iter=4, inner_block=4, ns=1)
            for i, (n, dl, th, its) in enumerate(specs)]
    alone = [Server(slots=1).run([r], realtime=False)[0] for r in reqs]
    mixed = Server(slots=3).run(reqs, realtime=False)
    for a, b in zip(alone, mixed):
        assert a.ok and b.ok, (a.error, b.error)
        assert a.iterations == b.iterations
        np.testing.assert_allclose(b.e, a.e, rtol=0, atol=1e-12)
        np.testing.assert_allclose(b.residuals, a.residuals, rtol=0, atol=1e-12)
    assert [r.nn for r in mixed] == [11, 21, 15]
