"""The one-electron (He+) block Lanczos on the device, BLOCKLANCZOS
(lanmod.F90:18-219).

Oracles: the eigenvalues of the dense one-electron H1 (Kronecker-built on the
host, independent of the device operator) and the numpy restatement
oracle/lanczos.py (pinned in tests/test_oracle.py), plus the reference's own
logged He+ eigenvalues at NN=11, delta=2, theta=0.2 (CSCALE-WORK/Outs/
Out.S.N5:51-64; its LANTOL was 1e-2, so the levels it had converged agree to
rounding and the rest to the log's spread).  Residuals are recomputed with
the oracle's H1.  At NN=19 (M = 6859, past the dense route's limit) the run
is checked by its residuals against the oracle's H1.
"""
import json

import numpy as np
import pytest

from oracle.hamiltonian import Hamiltonian

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def _oracle_residuals(h, e, v):
    from oracle.lanczos import h1_apply

    r = h1_apply(h, v) - v * e[None, :]
    return np.linalg.norm(r, axis=0) / np.linalg.norm(v, axis=0)


def test_lanczos_matches_dense_h1_and_logged_eigenvalues(dex, golden_dir):
    print("case: He+ Lanczos NN=11 delta 2 theta 0.2 block 3", flush=True)
    g = json.loads((golden_dir / "heplus_nn11_d2_t0.2.json").read_text())
    p = g["params"]
    h = Hamiltonian(p["N"], p["delta"], p["theta"], p["charge"])
    ev = np.linalg.eigvals(h.dense_one_electron())
    want = ev[np.argsort(ev.real, kind="stable")[:14]]
    with dex.HeliumGrid(p["N"], p["delta"], p["theta"]) as grid:
        e, orb, res, it = grid.heplus_lanczos(14, blocksize=3, kdim=501, check=10, tol=1e-9, seed=3)
    assert it % 10 == 0 and np.all(res < 1e-9)
    assert max(np.min(np.abs(want - z)) for z in e) < 1e-10
    assert max(np.min(np.abs(e - z)) for z in want) < 1e-10
    assert np.all(_oracle_residuals(h, e, orb) < 1.01e-9)
    np.testing.assert_allclose(np.linalg.norm(orb, axis=0), 1.0, rtol=1e-12)
    logged = np.array([complex(a, b) for a, b in g["heplus_eigenvalues"]])
    err = np.array([np.min(np.abs(e - z)) for z in logged])
    assert err[:3].max() < 1e-10 and err.max() < 5e-6


def test_lanczos_matches_oracle_restatement(dex):
    """The device run and the numpy restatement (column Gram-Schmidt as the
    reference) reach the same eigenvalues, each to its tolerance."""
    from oracle.lanczos import block_lanczos

    print("case: He+ Lanczos NN=7 vs oracle restatement", flush=True)
    h = Hamiltonian(3, 0.8, 0.25)
    rng = np.random.default_rng(5)
    ref = block_lanczos(h, 3, kdim=100, check=5, tol=1e-10, numeigs=6,
                        v0=rng.random((h.m, 3)) + 1e-4j * rng.random((h.m, 3)))
    with dex.HeliumGrid(3, 0.8, 0.25) as grid:
        e, orb, res, it = grid.heplus_lanczos(6, blocksize=3, kdim=100, check=5, tol=1e-10, seed=5)
    assert ref.converged
    assert max(np.min(np.abs(ref.eigenvalues - z)) for z in e) < 1e-9
    assert np.all(_oracle_residuals(h, e, orb) < 1.01e-10)


def test_lanczos_beyond_the_dense_route(dex):
    """NN=19 (M = 6859): the device Lanczos returns 14 orbitals whose
    residuals, recomputed with the oracle's H1, meet the tolerance, and they
    feed DIAGSUBSPACE on the two-electron grid."""
    print("case: He+ Lanczos NN=19 delta 2 theta 0.2", flush=True)
    h = Hamiltonian(9, 2.0, 0.2)
    with dex.HeliumGrid(9, 2.0, 0.2) as grid:
        e, orb, res, it = grid.heplus_lanczos(14, blocksize=3, kdim=501, check=10, tol=1e-8, seed=7)
        assert np.all(_oracle_residuals(h, e, orb) < 1.01e-8)
        assert np.all(np.diff(e.real) >= 0)  # EIGSORT order
        psi, e0, ex = grid.diagsubspace_device(orb, 1, -0.76)
        assert psi.shape[0] == grid.d and np.isfinite(e0).all()
