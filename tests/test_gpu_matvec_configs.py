"""GPU parity of HAMMULT at the configurations' own grid sizes against the C oracle.

NN=21 (delta 0.4, theta 0.35: the drop-in FP64 configuration) and NN=25
(delta 0.35, theta 0.3: the mixed-precision configuration).  The oracle is
``oracle/hammult.c`` (KEMULT2's column / transpose algorithm,
hammod.F90:19-67,138-187,264-286, with the C restatement of POTENTIAL2,
hammod.F90:502-582), itself pinned against the numpy restatement and a dense
Kronecker H in tests/test_oracle.py.

Paths covered: the FP64 LDS two-pass kernels (general and exchange-symmetric)
at NN=21; at NN=25 the FP64 six-pass generic kernels (no FP64 LDS cube fits),
the FP32 LDS kernels on complex128 vectors (general and exchange-symmetric)
and on complex64 vectors.  Bounds: FP64 within 1e-13 of max|y| (another
summation order of the same products); FP32 arithmetic within 2e-6 of max|y|.
"""
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


def _oracle(n, delta, theta):
    from oracle.c_oracle import CHammult
    from oracle.hamiltonian import Hamiltonian

    return CHammult(Hamiltonian(n, delta, theta), c_potential=True)


def _start(d, m, seed, sym):
    """Seeded complex Gaussian on the host, exchange-projected when sym != 0."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    if sym:
        xm = x.reshape(m, m)
        x = (0.5 * (xm + sym * xm.T)).reshape(-1)
    return np.ascontiguousarray(x)


def _gpu_matvec(g, x, sym, single=False):
    xd = torch.from_numpy(x).to("cuda")
    if single:
        xd = xd.to(torch.complex64)
    yd = torch.empty_like(xd)
    g.matvec_device(xd, yd, 1, single=single, symmetry=sym)
    y = yd.to(torch.complex128).cpu().numpy()
    del xd, yd
    return y


@pytest.mark.parametrize("sym", [0, 1])
def test_nn21_fp64_matches_c_oracle(dex, sym):
    print(f"case: NN=21 fp64 sym={sym} vs C oracle", flush=True)
    h = _oracle(10, 0.4, 0.35)
    x = _start(h.h.d, h.h.m, 2101 + sym, sym)
    with dex.HeliumGrid(10, 0.4, 0.35) as g:
        assert g.fast_path
        y = _gpu_matvec(g, x, sym)
    ref = h(x)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(y - ref)) <= 1e-13 * scale


def test_nn25_fp64_generic_matches_c_oracle(dex):
    print("case: NN=25 fp64 generic vs C oracle", flush=True)
    h = _oracle(12, 0.35, 0.3)
    x = _start(h.h.d, h.h.m, 2501, 0)
    with dex.HeliumGrid(12, 0.35, 0.3) as g:
        assert not g.fast_path  # the FP64 cube does not fit the LDS: six axis passes
        y = _gpu_matvec(g, x, 0)
    ref = h(x)
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("sym,single", [(0, False), (1, False), (1, True)])
def test_nn25_fp32_matches_c_oracle(dex, sym, single):
    print(f"case: NN=25 fp32 sym={sym} complex64={single} vs C oracle", flush=True)
    h = _oracle(12, 0.35, 0.3)
    x = _start(h.h.d, h.h.m, 2502 + sym, sym)
    if single:
        x = x.astype(np.complex64).astype(np.complex128)  # the input the complex64 path sees
    with dex.HeliumGrid(12, 0.35, 0.3, precision="fp32") as g:
        y = _gpu_matvec(g, x, sym, single=single)
    ref = h(x)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < 2e-6
    assert err > 1e-12  # FP32 arithmetic really ran
