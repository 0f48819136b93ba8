"""The allocation-free matvec entry point, called through the C-ABI.

dex_matvec_noalloc keeps the matvec intermediate in a static device array
instead of a library allocation; its output must equal that of the allocating
entry point dex_matvec_sym_device on the same columns (FP64 within 1e-13 of
max|y|: the same kernels, so in practice bitwise).  Two calls in a row share
the array, the second on a smaller input, and a grid whose intermediate does not
fit the array is refused with DEX_ESIZE.
"""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def _symmetric(howmany, m, sym, dtype=torch.complex128, seed=0):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(howmany, m, m, dtype=dtype, device="cuda", generator=gen)
    if sym:
        x = 0.5 * (x + sym * x.transpose(1, 2))
    return x.contiguous().view(-1)


def _call(lib, fn, g, x, y, howmany, single, sym):
    torch.cuda.synchronize()
    st = getattr(lib, fn)(g._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), howmany, g.d,
                          int(single), sym)
    msg = lib.dex_last_error().decode()
    assert st == 0, (fn, st, msg)
    g.sync()


@pytest.mark.parametrize("first,second", [
    # (n, precision, howmany, single, symmetry): the second call's input is smaller
    ((5, "fp64", 3, False, 1), (4, "fp64", 1, False, 0)),
    ((6, "fp32", 2, True, 1), (5, "fp32", 1, False, -1)),
])
def test_noalloc_matvec_matches_allocating_entry_point(dex, first, second):
    from doublyexcited_amd import _lib

    lib = _lib.load()
    for k, (n, prec, howmany, single, sym) in enumerate((first, second)):
        print(f"case: noalloc call {k + 1}: N={n} {prec} howmany={howmany} complex64={single} sym={sym}",
              flush=True)
        with dex.HeliumGrid(n, 0.5, 0.3, precision=prec) as g:
            x = _symmetric(howmany, g.m, sym, torch.complex64 if single else torch.complex128, seed=10 * n + k)
            y_ref = torch.full_like(x, float("nan"))
            y = torch.full_like(x, float("nan"))
            _call(lib, "dex_matvec_sym_device", g, x, y_ref, howmany, single, sym)
            _call(lib, "dex_matvec_noalloc", g, x, y, howmany, single, sym)
            assert not torch.isnan(torch.view_as_real(y)).any()
            scale = y_ref.abs().max().item()
            assert (y - y_ref).abs().max().item() <= 1e-13 * scale


def test_noalloc_refuses_an_intermediate_beyond_the_scratch(dex):
    from doublyexcited_amd import _lib

    lib = _lib.load()
    print("case: noalloc NN=13 fp64 refused", flush=True)
    with dex.HeliumGrid(6, 0.5, 0.3) as g:  # 13^6 complex128 = 77 MB > 64 MiB
        x = torch.zeros(g.d, dtype=torch.complex128, device="cuda")
        y = torch.zeros_like(x)
        torch.cuda.synchronize()
        st = lib.dex_matvec_noalloc(g._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), 1, g.d, 0, 1)
        assert st == _lib.DEX_ESIZE
        assert "64 MiB" in lib.dex_last_error().decode()
