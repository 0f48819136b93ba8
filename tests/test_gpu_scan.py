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
