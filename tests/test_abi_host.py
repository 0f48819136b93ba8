"""CPU checks of the native library: it loads, exports every symbol the header
declares, and its host-side arithmetic (KEMATRIX, GETINVERSE, the projected
pencil) agrees with the oracle and with LAPACK.  No device calls."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

from doublyexcited_amd import _lib
from oracle.hamiltonian import getinverse, kematrix

ROOT = Path(__file__).resolve().parent.parent
_dp = C.POINTER(C.c_double)


def header_symbols():
    text = (ROOT / "include" / "dex.h").read_text()
    return sorted(set(re.findall(r"\b(dex_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_status_strings_and_defaults():
    lib = _lib.load()
    assert lib.dex_status_string(_lib.DEX_ESIZE).decode() == "size mismatch"
    p = _lib.DexParams()
    lib.dex_params_default(C.byref(p))
    assert (p.n, p.delta, p.charge, p.nsmall, p.tfactor) == (2, 0.5, 2, 22, 4)
    o = _lib.DexSolveOpts()
    lib.dex_solve_opts_default(C.byref(o))
    assert o.fp64_restarts == 0 and o.strict_cholesky == 0 and o.fp32_tol == 1e-5 and o.refresh_every == 0


def test_host_kematrix_matches_oracle():
    lib = _lib.load()
    t = np.zeros(21 * 21)
    assert lib.dex_host_kematrix(21, 0.4, t.ctypes.data_as(_dp)) == 0
    np.testing.assert_allclose(t.reshape(21, 21, order="F"), kematrix(21, 0.4), rtol=1e-15)


@pytest.mark.parametrize("nout,delta", [(11, 2.0), (21, 0.4), (25, 0.35)])
def test_host_getinverse_matches_oracle(nout, delta):
    lib = _lib.load()
    L = 2 * nout + 1
    t = np.zeros(L ** 3)
    assert lib.dex_host_getinverse(nout, delta, 22, 4, t.ctypes.data_as(_dp)) == 0
    ref = getinverse(nout, delta)
    np.testing.assert_allclose(t.reshape(L, L, L, order="F"), ref, rtol=1e-12, atol=1e-14)


def test_host_getinverse_rejects_small_big_grid():
    lib = _lib.load()
    t = np.zeros(1)
    assert lib.dex_host_getinverse(100, 0.5, 22, 4, t.ctypes.data_as(_dp)) == _lib.DEX_EINVAL


@pytest.mark.parametrize("k,nev,seed", [(8, 2, 0), (32, 8, 1), (64, 16, 2), (24, 8, 3)])
def test_host_pencil_matches_lapack_zggev(k, nev, seed):
    lib = _lib.load()
    rng = np.random.default_rng(seed)
    n = 400
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, _ = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    sigma = 0.3 - 0.1j
    b = a @ q - sigma * q
    kp = np.asfortranarray(b.conj().T @ b)
    mp = np.asfortranarray(b.conj().T @ q)
    coords = np.zeros((k, nev), dtype=np.complex128, order="F")
    mu = np.zeros(nev, dtype=np.complex128)
    st = lib.dex_host_solve_gevp(k, kp.ctypes.data_as(_dp), mp.ctypes.data_as(_dp), k, nev,
                                 coords.ctypes.data_as(_dp), mu.ctypes.data_as(_dp))
    assert st == 0
    w = sla.eig(kp, mp, right=False)
    ref = w[np.argsort(np.abs(w))[:nev]]
    np.testing.assert_allclose(np.sort(np.abs(mu)), np.sort(np.abs(ref)), rtol=1e-9)
    for j in range(nev):
        y = coords[:, j]
        assert np.linalg.norm(y) == pytest.approx(1.0, rel=1e-12)
        r = kp @ y - mu[j] * (mp @ y)
        assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(kp, 2)


@pytest.mark.parametrize("seed", range(6))
def test_host_pencil_degenerate_clusters(seed):
    """Exactly repeated eigenvalues (3-fold clusters, as O_h symmetry produces):
    regression for a QR shift that lost the split of a 2 x 2 block to cancellation."""
    lib = _lib.load()
    rng = np.random.default_rng(100 + seed)
    n, k, nev = 300, 32, 8
    lam = rng.standard_normal(n // 3) + 1j * rng.standard_normal(n // 3)
    d = np.repeat(lam, 3)
    v = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = v @ np.diag(d) @ np.linalg.inv(v)
    # a basis containing near-converged eigenvectors of the clusters
    q = np.concatenate([v[:, :9] + 1e-9 * rng.standard_normal((n, 9)),
                        rng.standard_normal((n, k - 9)) + 0j], axis=1)
    q, _ = np.linalg.qr(q)
    sigma = lam[0] + 0.01
    b = a @ q - sigma * q
    kp = np.asfortranarray(b.conj().T @ b)
    mp = np.asfortranarray(b.conj().T @ q)
    coords = np.zeros((k, nev), dtype=np.complex128, order="F")
    mu = np.zeros(nev, dtype=np.complex128)
    st = lib.dex_host_solve_gevp(k, kp.ctypes.data_as(_dp), mp.ctypes.data_as(_dp), k, nev,
                                 coords.ctypes.data_as(_dp), mu.ctypes.data_as(_dp))
    assert st == 0
    w = sla.eig(kp, mp, right=False)
    ref = np.sort(np.abs(w[np.argsort(np.abs(w))[:nev]]))
    np.testing.assert_allclose(np.sort(np.abs(mu)), ref, rtol=1e-7)


def _c_struct_fields(name):
    import re
    from pathlib import Path

    hdr = (Path(__file__).resolve().parent.parent / "include" / "dex.h").read_text()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), hdr, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1]
        fields += [n.strip() for n in names.split(",")]
    return fields


def _fortran_type_fields(text, name):
    import re

    body = re.search(r"type, bind\(C\) :: %s\n(.*?)end type" % name, text, re.S).group(1)
    fields = []
    for ln in body.splitlines():
        if "::" in ln:
            fields += [n.strip() for n in ln.split("::")[1].split(",")]
    return fields


def test_struct_mirrors_match_the_header():
    """ctypes mirrors and the Fortran ISO_C_BINDING types (the binding notes at
    the repository root and the compiled Fortran check) list the header's fields in the header's order:
    a missing trailing field makes dex_solve_opts_default write past the
    caller's object."""
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    for cname, pyname in (("dex_solve_opts", "DexSolveOpts"), ("dex_result", "DexResult"),
                          ("dex_params", "DexParams")):
        fields = _c_struct_fields(cname)
        assert [f[0] for f in getattr(_lib, pyname)._fields_] == fields, cname
        for text in ((root / "INTEGRATION.md").read_text(),
                     (root / "tests" / "fortran" / "dex_binding_check.f90").read_text()):
            assert _fortran_type_fields(text, cname) == fields, cname
