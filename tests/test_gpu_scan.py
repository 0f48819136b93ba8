"""Theta / grid continuation on the GPU: a warm start from the neighbouring
theta's eigenvectors reaches the same pairs as a cold start in fewer
iterations, and a grid change through SIXD_FIX gives a usable start."""
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


def match(a, b):
    return max(np.min(np.abs(np.asarray(b) - z)) for z in a)


def test_warm_theta_step_same_pairs_fewer_iterations(dex, tmp_path):
    from doublyexcited_amd.eigvecs import read_eigenvectors
    from doublyexcited_amd.scan import theta_scan

    print("case: theta scan N=2", flush=True)
    pts = [(2, 0.5, 0.30), (2, 0.5, 0.31)]
    kw = dict(sigma=-0.78 - 0.002j, nev=4, tol=1e-9, inner_iter=4, guard=2, maxit=400)
    warm = theta_scan(pts, warm=True, checkpoint_dir=tmp_path, **kw)
    cold = theta_scan(pts[1:], warm=False, seed=77, **kw)
    assert warm[1].warm and not cold[0].warm
    assert warm[1].converged_pairs == cold[0].converged_pairs == 4
    assert match(warm[1].e, cold[0].e) < 1e-8 and match(cold[0].e, warm[1].e) < 1e-8
    assert warm[1].iterations < cold[0].iterations
    # the checkpoint holds the returned vectors in the reference's format
    v = read_eigenvectors(tmp_path / "eigvecs.S.N2.T0.3100")
    assert v.shape == (4, 5 ** 6)


def test_grid_change_through_sixd_fix(dex):
    from doublyexcited_amd.scan import theta_scan

    print("case: grid change N=2 -> N=3", flush=True)
    pts = [(2, 0.5, 0.30), (3, 0.5, 0.30)]
    out = theta_scan(pts, sigma=-0.78 - 0.002j, nev=4, tol=1e-8, inner_iter=4, guard=2, maxit=400)
    assert out[1].warm
    assert out[1].converged_pairs == 4


def test_scan_at_the_scan_configuration_grids(dex):
    """The theta scan's own settings (block 8 + 2 guard columns, NS 3, inner
    order 8, tol 1e-8, delta 0.5, sigma -0.78-0.002i) on its first grids:
    NN=11 at theta 0.30 and 0.31, then NN=13 at 0.31 through SIXD_FIX.  Each
    warm point converges all 8 pairs to the eigenvalues a cold seeded start
    reaches at the same point (within 1e-8 au), and the warm theta step takes
    fewer iterations than the cold start."""
    from doublyexcited_amd.scan import theta_scan

    kw = dict(sigma=-0.78 - 0.002j, nev=8, tol=1e-8, ns=3, inner_iter=8, guard=2, maxit=400)
    pts = [(5, 0.5, 0.30), (5, 0.5, 0.31), (6, 0.5, 0.31)]
    print("case: warm scan NN=11, 11, 13", flush=True)
    warm = theta_scan(pts, warm=True, **kw)
    print("case: cold starts NN=11, 13 at theta 0.31", flush=True)
    cold = theta_scan(pts[1:], warm=False, seed=91, **kw)
    for p in warm + cold:
        print(f"  NN={2 * p.n + 1} theta {p.theta} warm {p.warm}: {p.converged_pairs}/8 pairs, "
              f"{p.iterations} iterations, {p.seconds:.1f} s", flush=True)
    assert [p.warm for p in warm] == [False, True, True]
    assert all(p.converged_pairs == 8 for p in warm + cold)
    for w, c in ((warm[1], cold[0]), (warm[2], cold[1])):
        assert match(w.e, c.e) < 1e-8 and match(c.e, w.e) < 1e-8
    assert warm[1].iterations < cold[0].iterations
