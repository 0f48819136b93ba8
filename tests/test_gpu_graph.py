"""The HIP graph of the allocation-free block matvec, called through the C-ABI.

dex_matvec_graph_create captures dex_matvec_noalloc's launches on the
context's stream; dex_graph_launch replays them.  Every replay must equal the
eager dex_matvec_noalloc on the input the tensor holds at replay time (the
same kernels on the same data, so bitwise), with torch work and
synchronisations between replays.  A call the allocation-free entry point
refuses is refused at creation, and the context's stream stays usable.
"""
import ctypes as C
import time

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dex():
    import doublyexcited_amd as d

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return d


def _symmetric(howmany, m, sym, dtype, seed):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(howmany, m, m, dtype=dtype, device="cuda", generator=gen)
    if sym:
        x = 0.5 * (x + sym * x.transpose(1, 2))
    return x.contiguous().view(-1)


def _eager(lib, g, x, y, howmany, single, sym):
    torch.cuda.synchronize()
    st = lib.dex_matvec_noalloc(g._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), howmany, g.d,
                                int(single), sym)
    assert st == 0, (st, lib.dex_last_error().decode())
    g.sync()


@pytest.mark.parametrize("n,prec,howmany,single,sym", [
    (5, "fp64", 4, False, 1), (4, "fp64", 3, False, 0), (6, "fp32", 2, True, 1), (5, "fp32", 2, False, -1)])
def test_graph_replays_equal_eager_calls(dex, n, prec, howmany, single, sym):
    from doublyexcited_amd import _lib

    lib = _lib.load()
    dt = torch.complex64 if single else torch.complex128
    print(f"case: graph N={n} {prec} howmany={howmany} complex64={single} sym={sym}", flush=True)
    with dex.HeliumGrid(n, 0.5, 0.3, precision=prec) as g:
        x = _symmetric(howmany, g.m, sym, dt, seed=n)
        y = torch.full_like(x, float("nan"))
        y_ref = torch.full_like(x, float("nan"))
        with g.matvec_graph(x, y, howmany, single, sym) as gr:
            for rep in range(3):
                x.copy_(_symmetric(howmany, g.m, sym, dt, seed=100 * n + rep))  # new input each replay
                y.fill_(float("nan"))
                gr.launch()
                g.sync()
                _eager(lib, g, x, y_ref, howmany, single, sym)
                assert torch.equal(y, y_ref), rep
                assert y.abs().max().item() > 0
        assert not g._graphs


def test_graph_creation_refuses_what_the_noalloc_call_refuses(dex):
    from doublyexcited_amd import _lib

    print("case: graph NN=13 fp64 refused, then an eager matvec on the same context", flush=True)
    with dex.HeliumGrid(6, 0.5, 0.3) as g:  # FP64 intermediate 77 MB > the 64 MiB array
        x = _symmetric(1, g.m, 1, torch.complex128, seed=5)
        y = torch.empty_like(x)
        with pytest.raises(_lib.DexError) as ei:
            g.matvec_graph(x, y, 1, False, 1)
        assert ei.value.status == _lib.DEX_ESIZE
        assert not g._graphs
        g.matvec_device(x, y, 1, symmetry=1)  # the stream is not left capturing
        assert torch.isfinite(y.abs()).all() and y.abs().max().item() > 0


def test_graph_launch_time_against_eager_calls(dex):
    """Prints the host-side cost of a 16-column block at NN=9 both ways (no
    bound asserted: it depends on the host)."""
    from doublyexcited_amd import _lib

    lib = _lib.load()
    with dex.HeliumGrid(4, 0.5, 0.3) as g:
        x = _symmetric(16, g.m, 1, torch.complex128, seed=9)
        y = torch.empty_like(x)
        y_ref = torch.empty_like(x)
        with g.matvec_graph(x, y, 16, False, 1) as gr:
            gr.launch()
            g.sync()
            _eager(lib, g, x, y_ref, 16, False, 1)
            assert torch.equal(y, y_ref)
            reps = 50
            t0 = time.perf_counter()
            for _ in range(reps):
                gr.launch()
            g.sync()
            t_graph = (time.perf_counter() - t0) / reps
            ctx = g._ctx
            t0 = time.perf_counter()
            for _ in range(reps):
                assert lib.dex_matvec_noalloc(ctx, C.c_void_p(x.data_ptr()), C.c_void_p(y_ref.data_ptr()), 16,
                                              g.d, 0, 1) == 0
            g.sync()
            t_eager = (time.perf_counter() - t0) / reps
        print(f"NN=9, 16 columns (32 kernel launches): graph {t_graph * 1e6:.0f} us, eager "
              f"{t_eager * 1e6:.0f} us per block", flush=True)
