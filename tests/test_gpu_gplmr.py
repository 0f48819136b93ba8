"""GPU parity of the GPLMR / harmonic-Ritz block eigensolver (GPLMR_EIG).

Oracles: a dense eigendecomposition at NN=3, the numpy GPLMR restatement
(oracle/gplmr.py) from the same start block at NN=5, and the reference's
logged orbital-subspace value and singlet resonance at NN=11, delta=2,
theta=0.2 (CSCALE-WORK/Outs/Out.S.N5:97-147).  Tolerances: |dE| <= 1e-9 au for FP64 solves
converged to residual 1e-10, <= 1e-7 au for the mixed form (complex64 search
blocks with FP32 matvecs, FP64 restarts).
"""
import json

import numpy as np
import pytest

from oracle.gplmr import basic_block_lan_solve, gplmr_eig
from oracle.hamiltonian import Hamiltonian, exchange, singlet_project

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def start_block(h, nev, seed):
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal((h.d, nev)) + 1j * rng.standard_normal((h.d, nev))
    for j in range(nev):
        x0[:, j] = singlet_project(x0[:, j], h.m)
    return np.asfortranarray(x0)


def match(a, b):
    """max over a of the distance to the nearest element of b."""
    return max(np.min(np.abs(np.asarray(b) - z)) for z in a)


def test_gplmr_matches_dense_nn3(dex):
    print("case: gplmr dense NN=3", flush=True)
    h = Hamiltonian(1, 0.6, 0.3)
    hd = h.dense()
    w, v = np.linalg.eig(hd)
    sym = np.array([np.linalg.norm(v[:, j] - exchange(v[:, j], h.m)) for j in range(v.shape[1])])
    ws = w[sym < 1e-6]
    sigma, nev = 0.5 - 0.1j, 3
    x = start_block(h, nev, 11)
    with dex.HeliumGrid(1, 0.6, 0.3) as g:
        r = g.gplmr_eig(x, sigma, nev, ns=1, maxit=300, tol=1e-10)
    assert r.converged
    near = ws[np.argsort(np.abs(ws - sigma))[:nev]]
    np.testing.assert_allclose(np.sort_complex(r.e), np.sort_complex(near), atol=1e-9)
    # returned vectors are eigenvectors
    for j in range(nev):
        xj = x[:, j]
        assert np.linalg.norm(hd @ xj - r.e[j] * xj) / np.linalg.norm(xj) < 1e-9


# Blocks cover complete O_h clusters (at NN=5, sigma=-0.78-0.002i the
# nearest singlet levels are one non-degenerate level and 3-fold levels):
# without an inner solve, a block that splits a degenerate level stalls in
# both implementations.  Two 3-fold levels lie at almost the same distance
# from sigma (1.870 and 2.09 au, next to the single level at 1.867 au), and
# which of them the iteration settles on is decided by rounding: the oracle
# from these start blocks takes the farther one in ~60 iterations, and so
# may the GPU, or it may take the nearer one.  Either is a fixed point of the
# reference's method; the test therefore requires (1) every returned pair to
# be an eigenpair of the oracle's operator to the tolerance, and (2) the
# returned set to be the oracle's set or one at least as near to sigma.
@pytest.mark.parametrize("n,nev,sigma,tol,seed", [(2, 4, -0.78 - 0.002j, 1e-10, 102),
                                                  (2, 4, -0.78 - 0.002j, 1e-10, 17)])
def test_gplmr_matches_oracle_solver(dex, n, nev, sigma, tol, seed):
    print(f"case: gplmr vs oracle N={n} nev={nev} seed={seed}", flush=True)
    h = Hamiltonian(n, 0.5, 0.3)
    v = h.potential2()
    x0 = start_block(h, nev, seed)
    ref = gplmr_eig(lambda b: np.stack([h.hammult(b[:, j], v) for j in range(b.shape[1])], axis=1),
                    x0, sigma, nev, ns=1, maxit=600, tol=tol)
    assert ref.converged
    x = x0.copy(order="F")
    with dex.HeliumGrid(n, 0.5, 0.3) as g:
        r = g.gplmr_eig(x, sigma, nev, ns=1, maxit=600, tol=tol)
    assert r.converged
    for j in range(nev):
        xj = x[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - r.e[j] * xj) <= 10 * tol * np.linalg.norm(xj)
    same = match(r.e, ref.e) < 1e-9 and match(ref.e, r.e) < 1e-9
    assert same or np.max(np.abs(r.e - sigma)) <= np.max(np.abs(ref.e - sigma)) + 1e-9


@pytest.mark.parametrize("nev,inner", [(4, 1), (1, 0), (3, 0), (4, 3), (2, 6)])
def test_iteration_trajectory_matches_oracle(dex, nev, inner):
    """Step-by-step parity: from the same start block, 15 (4 with an inner solve) GPLMR iterations on the
    GPU and in the oracle give the same Rayleigh quotients (the iterations differ
    only by rounding), including PRECBLOCK's W <- H W form (inner_iter = 1)
    and the in-house inner solve (inner_iter >= 2)."""
    print(f"case: trajectory nev={nev} inner={inner}", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    sigma = -0.78 - 0.002j
    x0 = start_block(h, nev, 7)
    # the inner solve converges within a few iterations: stop while the
    # residual is still far above rounding level
    its = 15 if inner < 2 else 4
    ref = gplmr_eig(lambda b: np.stack([h.hammult(b[:, j], v) for j in range(b.shape[1])], axis=1),
                    x0, sigma, nev, ns=1, maxit=its, tol=1e-14, inner_iter=inner)
    x = x0.copy(order="F")
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        r = g.gplmr_eig(x, sigma, nev, ns=1, maxit=its, tol=1e-14, inner_iter=inner)
    assert r.iterations == ref.iters == its
    assert match(r.e, ref.e) < 1e-8 and match(ref.e, r.e) < 1e-8
    assert r.max_residual == pytest.approx(float(ref.resnrm.max()), rel=1e-4)


@pytest.mark.parametrize("k,order,sym", [(1, 1, 0), (3, 4, 0), (4, 5, 1), (2, 8, 1)])
def test_precblock_matches_oracle(dex, k, order, sym):
    """PRECBLOCK on the GPU (one matvec per Krylov column, block Gram-Schmidt)
    against the oracle's step-by-step basicblocklansolve on the same block:
    the same least-squares solution up to rounding."""
    print(f"case: precblock k={k} order={order} sym={sym}", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    sigma = -0.78 - 0.002j
    b = start_block(h, k, 31 + order)
    amul = lambda y: np.stack([h.hammult(y[:, j], v) for j in range(y.shape[1])], axis=1)
    ref = amul(b) if order == 1 else basic_block_lan_solve(amul, b, sigma, order)
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        xd = torch.from_numpy(np.ascontiguousarray(b.T)).to("cuda")
        pd = torch.empty_like(xd)
        g.precblock_device(xd, pd, k, sigma, order, symmetry=sym)
        got = pd.cpu().numpy().T
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)


def test_guard_columns_return_nearest_pairs(dex):
    """With guard columns the solver iterates NEV + guard columns and returns
    the NEV nearest of the block: every returned pair meets the tolerance
    (residual recomputed from the returned vectors with the oracle's
    operator), the values come nearest-first, and the set is no farther from
    sigma than the plain block's (the two 3-fold levels near this sigma make
    the exact set depend on rounding, as in the solver test above)."""
    print("case: guard N=2", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    nev, sigma = 4, -0.78 - 0.002j
    x0 = start_block(h, nev, 102)
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        x = x0.copy(order="F")
        r0 = g.gplmr_eig(x, sigma, nev, ns=1, maxit=600, tol=1e-10)
        xg = x0.copy(order="F")
        rg = g.gplmr_eig(xg, sigma, nev, ns=1, maxit=600, tol=1e-10, guard=2)
    assert r0.converged and rg.converged
    dist = np.abs(rg.e - sigma)
    assert np.all(np.diff(dist) >= -1e-9)
    assert dist.max() <= np.abs(r0.e - sigma).max() + 1e-9
    for j in range(nev):
        xj = xg[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - rg.e[j] * xj) <= 10 * 1e-10 * np.linalg.norm(xj)


@pytest.mark.parametrize("nev,ns", [(12, 3), (16, 2)])
def test_wide_projection_trajectory(dex, nev, ns):
    """Projected bases wider than one Gram / combine launch takes (72 and 80
    columns > 64): the harmonic Gram and the update go in chunks.  Step-by-step
    parity with the oracle over 8 iterations, as in the trajectory test."""
    print(f"case: wide projection nev={nev} ns={ns}", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    sigma = -0.78 - 0.002j
    x0 = start_block(h, nev, 11)
    its = 8
    ref = gplmr_eig(lambda b: np.stack([h.hammult(b[:, j], v) for j in range(b.shape[1])], axis=1),
                    x0, sigma, nev, ns=ns, maxit=its, tol=1e-14)
    x = x0.copy(order="F")
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        r = g.gplmr_eig(x, sigma, nev, ns=ns, maxit=its, tol=1e-14)
    assert r.iterations == ref.iters == its
    assert match(r.e, ref.e) < 1e-8 and match(ref.e, r.e) < 1e-8
    assert r.max_residual == pytest.approx(float(ref.resnrm.max()), rel=1e-4)


def test_wide_projection_guard_converges(dex):
    """NEV=16 with 2 guard columns and NS=2 (90 projected columns): converges,
    every returned pair meets the tolerance against the oracle's operator and
    the values come nearest-first."""
    print("case: wide projection guard nev=16", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    nev, sigma, tol = 16, -0.78 - 0.002j, 1e-9
    x = start_block(h, nev, 5).copy(order="F")
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        r = g.gplmr_eig(x, sigma, nev, ns=2, maxit=400, tol=tol, guard=2)
    assert r.converged
    dist = np.abs(r.e - sigma)
    assert np.all(np.diff(dist) >= -1e-9)
    for j in range(nev):
        xj = x[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - r.e[j] * xj) <= 10 * tol * np.linalg.norm(xj)


def test_fp32_inner_solve_same_pairs(dex):
    """The inner solve with a complex64 Arnoldi basis and FP32 matvecs is only a
    different preconditioner: the outer FP64 iteration returns the same pairs
    (|dE| <= 1e-9, residuals recomputed in FP64 <= tol)."""
    print("case: inner fp32 N=2", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    nev, sigma, tol = 4, -0.78 - 0.002j, 1e-10
    x0 = start_block(h, nev, 102)
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        x = x0.copy(order="F")
        r64 = g.gplmr_eig(x, sigma, nev, ns=1, maxit=400, tol=tol, inner_iter=4, symmetry=1)
        x32 = x0.copy(order="F")
        r32 = g.gplmr_eig(x32, sigma, nev, ns=1, maxit=400, tol=tol, inner_iter=4, symmetry=1, inner_fp32=True)
    assert r64.converged and r32.converged
    assert match(r32.e, r64.e) < 1e-9 and match(r64.e, r32.e) < 1e-9
    for j in range(nev):
        xj = x32[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - r32.e[j] * xj) <= 10 * tol * np.linalg.norm(xj)


def test_mixed_precision_within_1e7(dex):
    """Mixed form (complex64 search blocks with FP32 matvecs around complex128
    X and AX, then FP64 restarts) against the FP64 solve from the same start
    block: the same pairs (|dE| <= 1e-7 au both ways) and every returned
    residual, recomputed with the oracle's HAMMULT, <= tol.  Two guard columns
    keep the block from ending inside a degenerate level (without them two
    converged solves from one start can hold different 3-fold levels)."""
    print("case: mixed N=2", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    nev, sigma, tol = 4, -0.78 - 0.002j, 1e-9
    x0 = start_block(h, nev, 5)
    kw = dict(ns=1, maxit=800, tol=tol, guard=2, symmetry=1)
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        x = x0.copy(order="F")
        r64 = g.gplmr_eig(x, sigma, nev, **kw)
    with dex.HeliumGrid(2, 0.5, 0.3, precision="mixed") as g:
        x = x0.copy(order="F")
        r32 = g.gplmr_eig(x, sigma, nev, fp32_tol=1e-5, **kw)
    assert r64.converged and r32.converged
    assert r32.fp64_restarts_used >= 1
    assert match(r32.e, r64.e) <= 1e-7 and match(r64.e, r32.e) <= 1e-7
    for j in range(nev):
        xj = x[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - r32.e[j] * xj) <= tol * np.linalg.norm(xj) * 1.01


def test_mixed_precision_singlet_inner_solve(dex):
    """Mixed form in the singlet mode with the inner solve (batched FP32
    matvecs on complex64 columns, complex64 Arnoldi basis) and guard columns:
    the same pairs as the FP64 solve (|dE| <= 1e-7 au), residuals recomputed
    with the oracle's HAMMULT <= tol."""
    print("case: mixed singlet inner N=3", flush=True)
    h = Hamiltonian(3, 0.5, 0.3)
    v = h.potential2()
    nev, sigma, tol = 4, -0.78 - 0.002j, 1e-8
    x0 = start_block(h, nev, 17)
    kw = dict(ns=2, maxit=400, tol=tol, inner_iter=4, symmetry=1, guard=2)
    with dex.HeliumGrid(3, 0.5, 0.3) as g:
        x = x0.copy(order="F")
        r64 = g.gplmr_eig(x, sigma, nev, **kw)
    with dex.HeliumGrid(3, 0.5, 0.3, precision="mixed") as g:
        x = x0.copy(order="F")
        r32 = g.gplmr_eig(x, sigma, nev, **kw)
    assert r64.converged and r32.converged
    assert match(r32.e, r64.e) <= 1e-7 and match(r64.e, r32.e) <= 1e-7
    for j in range(nev):
        xj = x[:, j]
        assert np.linalg.norm(h.hammult(xj, v) - r32.e[j] * xj) <= tol * np.linalg.norm(xj) * 1.01


def test_diagsubspace_matches_oracle(dex):
    """DIAGSUBSPACE on the GPU (orbital products built on the device, H applied
    by the library's matvec, A = psi^T H psi) against the oracle's restatement
    from the same orbitals: eigenvalues to 1e-10, vectors to 1e-9 up to a phase,
    expectation values to 1e-10."""
    from oracle.startvec import diagsubspace, heplus_orbitals

    print("case: diagsubspace N=2", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    v = h.potential2()
    _, orb = heplus_orbitals(h, 6)
    amul = lambda b: np.stack([h.hammult(b[:, j], v) for j in range(b.shape[1])], axis=1)
    for singlet in (True, False):
        e_ref, psi_ref, ex_ref = diagsubspace(amul, orb, singlet, False, -1.0 + 0.0j, 3)
        with dex.HeliumGrid(2, 0.5, 0.3) as g:
            psi, e0, ex = g.diagsubspace_device(orb, 3, -1.0 + 0.0j, singlet=singlet)
            got = psi.cpu().numpy().reshape(3, -1).T
        np.testing.assert_allclose(e0, e_ref, atol=1e-10)
        np.testing.assert_allclose(ex, ex_ref, atol=1e-10)
        for j in range(3):
            ph = np.vdot(psi_ref[:, j], got[:, j])
            assert abs(abs(ph) - 1.0) < 1e-9
            assert np.linalg.norm(got[:, j] - ph * psi_ref[:, j]) < 1e-9


def test_reference_logged_singlet_resonance(dex, golden_dir):
    """End to end against the reference's log at NN=11, delta=2, theta=0.2 (He,
    singlet): the 14 lowest He+ orbitals -> the orbital-product subspace vector
    nearest -0.76 (the log printed -0.70668286560422 + 0.03676471730071 i for
    it, from Lanczos orbitals converged to LANTOL 1e-2) -> GPLMR with NEV = 1,
    NS = 5 (the reference's default) and the inner solve of order 5, shifted by
    the vector's expectation value (the log's chosen shift, to the same
    orbital accuracy).  The reference reached E = -0.713154142874 +
    0.035227459486 i at residual 8e-8 (Out.S.N5:147); the oracle's run of the
    same recipe reaches it to 3e-10."""
    from oracle.startvec import heplus_orbitals

    print("case: logged resonance NN=11", flush=True)
    g = json.loads((golden_dir / "heplus_nn11_d2_t0.2.json").read_text())
    ref = complex(*g["singlet_resonance"])
    sub = g["subspace"]
    p = g["params"]
    h = Hamiltonian(p["N"], p["delta"], p["theta"])
    _, orb = heplus_orbitals(h, sub["numorbs"])
    with dex.HeliumGrid(p["N"], p["delta"], p["theta"]) as grid:
        psi, e0, ex = grid.diagsubspace_device(orb, 1, complex(*sub["etarget"]), singlet=True)
        assert abs(e0[0] - complex(*sub["eigenvalues"][0])) < 5e-6
        assert abs(ex[0] - complex(*sub["chosen_shift"])) < 1e-5
        r = grid.gplmr_eig_device(psi, ex[0], 1, ns=5, maxit=200, tol=1e-9, inner_iter=5, symmetry=1)
        res = grid.residuals_device(psi, r.e)
    assert r.converged, (r.iterations, r.max_residual)
    assert res[0] < 2e-9
    # the log's own value carries its residual 8e-8
    assert abs(r.e[0] - ref) < 5e-8


def test_singlet_mode_matches_general_solve(dex):
    """The exchange-shortcut solve (symmetry=+1) reaches the same eigenvalues as
    the general-HAMMULT solve from the same singlet start block."""
    print("case: singlet-mode solve N=2", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    nev, sigma = 4, -0.78 - 0.002j
    x0 = start_block(h, nev, 101)
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        xa = x0.copy(order="F")
        ra = g.gplmr_eig(xa, sigma, nev, ns=1, maxit=600, tol=1e-10)
        xb = x0.copy(order="F")
        rb = g.gplmr_eig(xb, sigma, nev, ns=1, maxit=600, tol=1e-10, symmetry=1)
    assert ra.converged and rb.converged
    assert match(rb.e, ra.e) < 1e-9 and match(ra.e, rb.e) < 1e-9
    for j in range(nev):
        xm = xb[:, j].reshape(h.m, h.m)
        assert np.array_equal(xm, xm.T)


def test_kernel_statistics_every_and_sampled(dex):
    """Kernel statistics: with timing 1 every launch is counted, with timing 4
    a random subset (about a quarter) whose algorithmic bytes per launch and
    device time per launch stay those of the full count; the solve itself is
    unchanged by timing (same eigenvalues, same iteration count)."""
    print("case: kernel statistics", flush=True)
    h = Hamiltonian(2, 0.5, 0.3)
    x0 = start_block(h, 4, 102)
    out = {}
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        for period in (0, 1, 4):
            g.reset_kernel_stats()
            g.set_kernel_timing(period)
            x = x0.copy(order="F")
            r = g.gplmr_eig(x, -0.78 - 0.002j, 4, ns=1, maxit=40, tol=1e-14, symmetry=1)
            g.set_kernel_timing(False)
            out[period] = (r, g.kernel_stats())
    r0, s0 = out[0]
    r1, s1 = out[1]
    r4, s4 = out[4]
    assert all(v[1] == 0 for v in s0.values())
    assert r1.iterations == r4.iterations == r0.iterations
    assert np.allclose(r1.e, r0.e, atol=1e-12) and np.allclose(r4.e, r0.e, atol=1e-12)
    for cls in ("pass1", "symadd", "gram", "combine"):
        n1, n4 = s1[cls][1], s4[cls][1]
        assert n1 > 30, (cls, n1)
        assert 0.05 * n1 < n4 < 0.6 * n1, (cls, n1, n4)
        # bytes per timed launch: the same population, sampled
        b1, b4 = s1[cls][2] / n1, s4[cls][2] / n4
        assert 0.5 * b1 < b4 < 2.0 * b1, cls
        assert s1[cls][0] > 0 and s4[cls][0] > 0



def test_mixed_solve_where_fp64_has_only_the_generic_kernels(dex):
    """NN=25 (no FP64 LDS cube fits, so the mixed form's FP64 AX comes from
    the generic kernels while its complex64 blocks use the exchange
    transpose-add ones): after 80 iterations the returned pairs still sit at
    the levels near sigma with FP64 residuals below 0.5.  Without the
    exchange projection of the generic output, the rounding-level odd part
    grew through the iteration and the returned pairs had |E| near 50-100 au
    and residuals near 70."""
    print("case: mixed NN=25 block 2, 80 iterations", flush=True)
    sigma = -0.78 - 0.002j
    with dex.HeliumGrid(12, 0.35, 0.3, precision="mixed") as g:
        x = torch.empty(g.d * 2, dtype=torch.complex128, device="cuda")
        g.fill_gaussian_device(x, 2, seed=1003, symmetry=1)
        r = g.gplmr_eig_device(x, sigma, 2, ns=1, maxit=80, tol=1e-7, inner_iter=4, inner_block=4, symmetry=1,
                               fp32_tol=1e-5)
        res = g.residuals_device(x, r.e)
    assert np.all(np.abs(r.e - sigma) < 0.5), r.e
    assert np.all(res < 0.5), res
