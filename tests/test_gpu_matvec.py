"""GPU parity of the matvec (HAMMULT / MYMATMUL) and the dots (MYDOTSUB).

Oracle: ``oracle.hamiltonian`` (numpy restatement of hammod.F90:19-67,138-187,
502-582 and tooth.f90:14-101), itself pinned by tests/test_oracle.py.
Tolerances: FP64 results within 1e-13 of max|y| (a different summation order
of the same products); complex64 arithmetic within 2e-6 of max|y|.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle.hamiltonian import Hamiltonian, exchange, getinverse

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_matvec_matches_dense_nn3(dex):
    print("case: dense NN=3", flush=True)
    h = Hamiltonian(1, 0.6, 0.3)
    x = _rand(h.d, 1)
    ref = h.dense() @ x
    with dex.HeliumGrid(1, 0.6, 0.3) as g:
        y = g.hammult(x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,delta,theta", [(2, 0.45, 0.35), (3, 0.5, 0.3), (5, 2.0, 0.2), (6, 0.5, 0.25)])
def test_matvec_matches_oracle(dex, n, delta, theta):
    print(f"case: oracle N={n}", flush=True)
    h = Hamiltonian(n, delta, theta)
    x = _rand(h.d, n)
    ref = h.hammult(x)
    with dex.HeliumGrid(n, delta, theta) as g:
        assert g.fast_path
        y = g.hammult(x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_block_matvec_howmany(dex):
    print("case: howmany=3 NN=7", flush=True)
    h = Hamiltonian(3, 0.5, 0.3)
    x = np.stack([_rand(h.d, 10 + j) for j in range(3)], axis=1)
    with dex.HeliumGrid(3, 0.5, 0.3) as g:
        y = np.empty_like(x, order="F")
        xf = np.asfortranarray(x)
        g.mymatmul(xf, y, 3, g.d)
    for j in range(3):
        ref = h.hammult(x[:, j])
        assert np.max(np.abs(y[:, j] - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_checksize_mismatch_is_size_error(dex):
    with dex.HeliumGrid(2, 0.5, 0.3) as g:
        x = np.zeros(g.d, np.complex128)
        y = np.zeros(g.d, np.complex128)
        with pytest.raises(dex.DexError) as ei:
            g.mymatmul(x, y, 1, g.d - 1)
        assert ei.value.status == 1


def test_tables_match_oracle(dex):
    with dex.HeliumGrid(10, 0.4, 0.35) as g:
        np.testing.assert_allclose(g.tinv(), getinverse(21, 0.4), rtol=1e-12, atol=1e-14)


def test_fp32_matvec_matches_oracle(dex):
    print("case: fp32 NN=11", flush=True)
    h = Hamiltonian(5, 0.5, 0.3)
    x = _rand(h.d, 3)
    ref = h.hammult(x)
    with dex.HeliumGrid(5, 0.5, 0.3, precision="fp32") as g:
        y = g.hammult(x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < 2e-6
    assert err > 1e-12  # the FP32 kernels really ran


def _fast_vs_generic(n, delta, theta, seed):
    """Run the generic six-pass kernels in a child process (DEX_MATVEC=generic)."""
    code = (
        "import numpy as np, sys, doublyexcited_amd as d\n"
        f"g = d.HeliumGrid({n}, {delta}, {theta})\n"
        "assert not g.fast_path\n"
        f"rng = np.random.default_rng({seed})\n"
        "x = rng.standard_normal(g.d) + 1j * rng.standard_normal(g.d)\n"
        "np.save(sys.argv[1], g.hammult(x))\n"
    )
    return code


@pytest.mark.parametrize("n,delta,theta", [(8, 0.5, 0.3), (10, 0.4, 0.35)])
def test_fast_kernels_match_generic_kernels(dex, tmp_path, n, delta, theta):
    print(f"case: fast vs generic N={n}", flush=True)
    out = tmp_path / "gen.npy"
    env = dict(os.environ, DEX_MATVEC="generic")
    subprocess.run([sys.executable, "-c", _fast_vs_generic(n, delta, theta, 5), str(out)], env=env,
                   check=True, timeout=300)
    ref = np.load(out)
    with dex.HeliumGrid(n, delta, theta) as g:
        assert g.fast_path
        rng = np.random.default_rng(5)
        x = rng.standard_normal(g.d) + 1j * rng.standard_normal(g.d)
        y = g.hammult(x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_exchange_symmetry_nn21(dex):
    print("case: exchange NN=21", flush=True)
    with dex.HeliumGrid(10, 0.4, 0.35) as g:
        x = torch.randn(g.d, dtype=torch.complex128, device="cuda")
        y = torch.empty_like(x)
        g.matvec_device(x, y)
        xp = x.view(g.m, g.m).t().contiguous().view(-1)
        yp = torch.empty_like(x)
        g.matvec_device(xp, yp)
        g.sync()
        d = (yp.view(g.m, g.m).t().contiguous().view(-1) - y).abs().max().item()
        assert d <= 1e-13 * y.abs().max().item()


def test_dots_match_numpy(dex):
    with dex.HeliumGrid(3, 0.5, 0.3) as g:
        b = np.asfortranarray(np.stack([_rand(g.d, 20 + j) for j in range(3)], axis=1))
        k = np.asfortranarray(np.stack([_rand(g.d, 30 + j) for j in range(5)], axis=1))
        out = g.mydotsub(b, k, g.d, 3, 5)
    np.testing.assert_allclose(out, b.conj().T @ k, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_dots_device_over_the_gram_shapes(dex, n):
    """MYDOTSUB on device columns over the shapes of both Gram kernels: the
    packed real/imaginary one (kl <= 32, kr <= 8) and the 16 x 16 MFMA tiles,
    one launch (<= 64 x 128) and chunked (beyond), on D = 729, 15625 and
    117649 rows (none a multiple of the row tiles).  Each entry within
    1e-13 ||b_i|| ||k_j|| of numpy's FP64 product (a bound that a dropped
    row tile, a swapped real/imaginary part or a missing conjugate exceeds)."""
    shapes = [(1, 1), (3, 5), (8, 8), (13, 7), (32, 8), (33, 8), (5, 9), (16, 16), (40, 70), (64, 128),
              (70, 130)]
    with dex.HeliumGrid(n, 0.5, 0.3) as g:
        kl_max = max(s[0] for s in shapes)
        kr_max = max(s[1] for s in shapes)
        rng = np.random.default_rng(40 + n)
        b = rng.standard_normal((kl_max, g.d)) + 1j * rng.standard_normal((kl_max, g.d))
        k = rng.standard_normal((kr_max, g.d)) + 1j * rng.standard_normal((kr_max, g.d))
        bt = torch.from_numpy(b).cuda()
        kt = torch.from_numpy(k).cuda()
        nb = np.linalg.norm(b, axis=1)
        nk = np.linalg.norm(k, axis=1)
        for kl, kr in shapes:
            out = g.dots_device(bt[:kl].contiguous().view(-1), kt[:kr].contiguous().view(-1), kl, kr)
            ref = b[:kl].conj() @ k[:kr].T
            err = np.abs(out - ref) / np.outer(nb[:kl], nk[:kr])
            assert err.max() <= 1e-13, (kl, kr, g.d, err.max())


@pytest.mark.parametrize("n,sym", [(3, 1), (5, -1), (10, 1), (10, -1)])
def test_exchange_symmetric_matvec_matches_general(dex, n, sym):
    """dex_matvec_sym_device on X = sym X^T agrees with the general HAMMULT and
    its output is exactly (anti)symmetric."""
    print(f"case: sym matvec N={n} sym={sym}", flush=True)
    with dex.HeliumGrid(n, 0.4, 0.35) as g:
        x = torch.randn(g.d, dtype=torch.complex128, device="cuda")
        xm = x.view(g.m, g.m)
        x = (0.5 * (xm + sym * xm.t())).contiguous().view(-1)
        y = torch.empty_like(x)
        ys = torch.empty_like(x)
        g.matvec_device(x, y)
        g.matvec_device(x, ys, symmetry=sym)
        g.sync()
        scale = y.abs().max().item()
        assert (ys - y).abs().max().item() <= 1e-13 * scale
        ysm = ys.view(g.m, g.m)
        assert torch.equal(ysm, sym * ysm.t())


@pytest.mark.parametrize("n,howmany", [(2, 5), (5, 20)])
def test_batched_symmetric_matvec_matches_single_column(dex, n, howmany):
    """Several exchange-symmetric columns in one call go through the batched
    launches (up to 16 columns per pass; 20 columns = 16 + 4): bitwise the
    per-column results."""
    print(f"case: batched sym matvec N={n} howmany={howmany}", flush=True)
    with dex.HeliumGrid(n, 0.5, 0.3) as g:
        x = torch.randn(howmany, g.m, g.m, dtype=torch.complex128, device="cuda")
        x = (0.5 * (x + x.transpose(1, 2))).contiguous().view(-1)
        yb = torch.empty_like(x)
        g.matvec_device(x, yb, howmany, symmetry=1)
        y1 = torch.empty_like(x)
        for j in range(howmany):
            g.matvec_device(x[j * g.d:(j + 1) * g.d], y1[j * g.d:(j + 1) * g.d], 1, symmetry=1)
        g.sync()
        assert torch.equal(yb, y1)


def test_batched_symmetric_matvec_fp32_context(dex):
    """The batched launches in an FP32 context (FP32 arithmetic, complex128
    columns): bitwise the per-column results."""
    print("case: batched sym matvec fp32 context N=5", flush=True)
    with dex.HeliumGrid(5, 0.5, 0.3, precision="fp32") as g:
        howmany = 6
        x = torch.randn(howmany, g.m, g.m, dtype=torch.complex128, device="cuda")
        x = (0.5 * (x + x.transpose(1, 2))).contiguous().view(-1)
        yb = torch.empty_like(x)
        g.matvec_device(x, yb, howmany, symmetry=1)
        y1 = torch.empty_like(x)
        for j in range(howmany):
            g.matvec_device(x[j * g.d:(j + 1) * g.d], y1[j * g.d:(j + 1) * g.d], 1, symmetry=1)
        g.sync()
        assert torch.equal(yb, y1)


def test_symmetric_matvec_fp32(dex):
    print("case: sym matvec fp32 N=10", flush=True)
    with dex.HeliumGrid(10, 0.4, 0.35, precision="fp32") as g:
        x = torch.randn(g.d, dtype=torch.complex128, device="cuda")
        xm = x.view(g.m, g.m)
        x = (0.5 * (xm + xm.t())).contiguous().view(-1)
        x32 = x.to(torch.complex64)
        y32 = torch.empty_like(x32)
        g.matvec_device(x32, y32, single=True, symmetry=1)
        g.sync()
    with dex.HeliumGrid(10, 0.4, 0.35) as g:
        y = torch.empty_like(x)
        g.matvec_device(x, y)
        g.sync()
    err = (y32.to(torch.complex128) - y).abs().max().item() / y.abs().max().item()
    assert 1e-12 < err < 2e-6
