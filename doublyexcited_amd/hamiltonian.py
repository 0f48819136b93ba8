"""Host-side mirror of the reference's operator / eigensolver interface.

``HeliumGrid`` owns one native context (``dex_ctx``): the reference's
``PARAMETERS`` + ``HAMILTONIAN`` built by ``INIT_STRUCTS(NELEC=2)``
(input.F90:365-417).  Its methods keep the reference's argument meaning:

* ``hammult(x)``                 -> HAMMULT(HAM,PARAMS,MPIDATA,X,Y,MM), hammod.F90:19
* ``mymatmul(inp, out, howmany, checksize)`` -> MYMATMUL, doublyexcited.F90:24-44
* ``mydotsub(bras, kets, size, nn, mm)``     -> MYDOTSUB, doublyexcited.F90:47-59
* ``gplmr_eig(x, sigma, nev, ns, maxit, tol, inner_iter, outmodulus)``
                                 -> GPLMR_EIG, gplmr.F90:9

Host arrays are numpy complex128, vector-major (vector j is ``x[j*D:(j+1)*D]``
or column j of a (D, k) Fortran-order array).  Device variants take torch
tensors on ``cuda`` (torch provides the memory; all arithmetic is in the
native library).  Size mismatches raise ``DexError`` with status ``DEX_ESIZE``
where the reference prints "SIZE ERRORRRR 66" and STOPs.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from ._lib import DexError, check

_dp = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _dev(t):
    """Device pointer of a torch tensor, after the work torch has queued on its
    current stream (which may still be writing the tensor) has finished."""
    import torch

    torch.cuda.current_stream(t.device).synchronize()
    return C.c_void_p(t.data_ptr())


def _check_dev(t, ncols: int, d: int, device: int, single: bool = False, name: str = "x"):
    """Raises before any device call unless ``t`` is a contiguous tensor on this
    context's GPU that holds ``ncols`` vectors of ``d`` entries in the format the
    call reads or writes (complex128, or complex64 with ``single``): the library
    takes a bare pointer and trusts its extent."""
    import torch

    want = torch.complex64 if single else torch.complex128
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA tensor")
    if t.device.index != device:
        raise ValueError(f"{name} is on cuda:{t.device.index}, the context on cuda:{device}")
    if t.dtype != want:
        raise TypeError(f"{name} must be {want}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if ncols < 0 or t.numel() < ncols * d:
        raise DexError(_lib.DEX_ESIZE, f"{name} holds {t.numel()} entries, the call needs {ncols} x {d}")


@dataclasses.dataclass
class SolveResult:
    e: np.ndarray
    residuals: np.ndarray | None
    iterations: int
    converged: bool
    matvecs: int
    max_residual: float
    t_total: float
    t_matvec: float
    t_prec: float
    t_gram: float
    t_update: float
    t_smalleig: float
    fallbacks: int
    fp64_restarts_used: int
    status: int


class MatvecGraph:
    """A captured block matvec of one HeliumGrid (``HeliumGrid.matvec_graph``);
    holds its input and output tensors so their memory outlives the graph."""

    def __init__(self, grid: "HeliumGrid", handle, tensors):
        self._grid = grid
        self._h = handle
        self._tensors = tensors

    def launch(self):
        """Replays the matvec on the grid's stream, after the work torch has
        queued that may still be writing the input."""
        if not self._h.value:
            raise ValueError("the graph is closed")
        _dev(self._tensors[0])
        check(self._grid._lib.dex_graph_launch(self._h))

    def close(self):
        if self._h.value:
            h, self._h = self._h, C.c_void_p()
            self._grid._graphs.remove(self)
            check(self._grid._lib.dex_graph_destroy(h))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class HeliumGrid:
    """One complex-scaled two-electron sinc-DVR Hamiltonian on one GPU."""

    def __init__(self, n: int, delta: float, theta: float, charge: int = 2, nsmall: int = 22,
                 tfactor: int = 4, precision: str = "fp64", device: int = 0):
        lib = _lib.load()
        p = _lib.DexParams()
        lib.dex_params_default(C.byref(p))
        p.n, p.delta, p.theta, p.charge = n, float(delta), float(theta), charge
        p.nsmall, p.tfactor, p.device = nsmall, tfactor, device
        p.precision = {"fp64": _lib.DEX_FP64, "fp32": _lib.DEX_FP32, "mixed": _lib.DEX_FP32}[precision]
        self._ctx = C.c_void_p()
        check(lib.dex_ctx_create(C.byref(p), C.byref(self._ctx)))
        info = _lib.DexInfo()
        check(lib.dex_ctx_info(self._ctx, C.byref(info)))
        self.n, self.nn, self.m, self.d = info.n, info.nn, info.m, info.d
        self.delta, self.theta, self.charge = float(delta), float(theta), charge
        self.precision = precision
        self.fast_path = bool(info.fast_path)
        self.device = device
        self._lib = lib
        self._graphs: list = []  # MatvecGraph objects still open (freed before the context)

    # -- lifetime -----------------------------------------------------------------
    def close(self):
        for gr in list(getattr(self, "_graphs", [])):
            gr.close()
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._lib.dex_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- tables ---------------------------------------------------------------------
    def kematrix(self) -> np.ndarray:
        t = np.zeros((self.nn, self.nn), order="F")
        check(self._lib.dex_get_kematrix(self._ctx, _ptr(t)))
        return t

    def tinv(self) -> np.ndarray:
        L = 2 * self.nn + 1
        t = np.zeros(L ** 3)
        check(self._lib.dex_get_tinv(self._ctx, _ptr(t)))
        return t.reshape(L, L, L, order="F")  # [i+NN, j+NN, k+NN]

    # -- operator -----------------------------------------------------------------------
    def mymatmul(self, inp: np.ndarray, out: np.ndarray, howmany: int, checksize: int) -> None:
        if inp.dtype != np.complex128 or out.dtype != np.complex128:
            raise TypeError("complex128 arrays required")
        if inp.size < howmany * checksize or out.size < howmany * checksize:
            raise DexError(_lib.DEX_ESIZE, "array smaller than howmany*checksize")
        if not (inp.flags.f_contiguous or inp.flags.c_contiguous) or not (
                out.flags.f_contiguous or out.flags.c_contiguous):
            raise ValueError("contiguous arrays required")
        check(self._lib.dex_matvec(self._ctx, _ptr(inp), _ptr(out), howmany, checksize))

    def hammult(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
        if x.size % self.d:
            raise DexError(_lib.DEX_ESIZE, f"vector length {x.size} is not a multiple of {self.d}")
        y = np.empty_like(x)
        self.mymatmul(x, y, x.size // self.d, self.d)
        return y

    def mydotsub(self, bras: np.ndarray, kets: np.ndarray, size: int, nn: int, mm: int) -> np.ndarray:
        out = np.zeros((nn, mm), dtype=np.complex128, order="F")
        b = np.asfortranarray(bras, dtype=np.complex128)
        k = np.asfortranarray(kets, dtype=np.complex128)
        check(self._lib.dex_dots(self._ctx, _ptr(b), _ptr(k), size, nn, mm,
                                 out.ctypes.data_as(_dp)))
        return out

    # device variants (torch tensors as memory)
    def matvec_device(self, x, y, howmany: int | None = None, single: bool = False,
                      symmetry: int = 0) -> None:
        """HAMMULT on device tensors; symmetry = +1/-1 uses the exchange shortcut."""
        howmany = x.numel() // self.d if howmany is None else howmany
        _check_dev(x, howmany, self.d, self.device, single, "x")
        _check_dev(y, howmany, self.d, self.device, single, "y")
        check(self._lib.dex_matvec_sym_device(self._ctx, _dev(x), _dev(y), howmany, self.d,
                                              int(single), int(symmetry)))
        self.sync()

    def matvec_graph(self, x, y, howmany: int | None = None, single: bool = False,
                     symmetry: int = 0) -> "MatvecGraph":
        """A HIP graph of the allocation-free matvec Y = H X on these device
        tensors (dex_matvec_graph_create): ``launch()`` replays it on the grid's
        stream, reading what ``x`` holds then (``sync()`` orders the host
        after it); ``close()`` frees it (closing the grid frees it too)."""
        howmany = x.numel() // self.d if howmany is None else howmany
        _check_dev(x, howmany, self.d, self.device, single, "x")
        _check_dev(y, howmany, self.d, self.device, single, "y")
        h = C.c_void_p()
        check(self._lib.dex_matvec_graph_create(self._ctx, _dev(x), _dev(y), howmany, self.d, int(single),
                                                int(symmetry), C.byref(h)))
        gr = MatvecGraph(self, h, (x, y))
        self._graphs.append(gr)
        return gr

    def dots_device(self, bras, kets, nn: int, mm: int) -> np.ndarray:
        out = np.zeros((nn, mm), dtype=np.complex128, order="F")
        _check_dev(bras, nn, self.d, self.device, name="bras")
        _check_dev(kets, mm, self.d, self.device, name="kets")
        check(self._lib.dex_dots_device(self._ctx, _dev(bras), _dev(kets), self.d, nn, mm,
                                        out.ctypes.data_as(_dp)))
        return out

    def fill_gaussian_device(self, x, ncols: int, seed: int, stream_id: int = 0, symmetry: int = 1):
        _check_dev(x, ncols, self.d, self.device)
        check(self._lib.dex_fill_gaussian_device(self._ctx, _dev(x), ncols, seed, stream_id,
                                                 int(symmetry)))

    def sync(self):
        check(self._lib.dex_sync(self._ctx))

    def stop(self, stop: bool = True):
        """Asks a solve running on this grid's context (from any thread) to
        return at its next iteration boundary; later solves return at their
        first until ``stop(False)``."""
        check(self._lib.dex_ctx_stop(self._ctx, int(bool(stop))))

    def set_kernel_timing(self, enable: bool | int = True):
        """False/0 off, True/1 every launch, n > 1 a random ~1 in n launches."""
        check(self._lib.dex_set_kernel_timing(self._ctx, int(enable)))

    def kernel_timing(self):
        a, b = C.c_double(), C.c_double()
        check(self._lib.dex_get_kernel_timing(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_stats(self) -> dict:
        """{class: (device ms, launches, algorithmic bytes, algorithmic flops)}
        accumulated while timing is on."""
        out = {}
        for i, name in enumerate(_lib.KERNEL_CLASSES):
            ms, n, by, fl = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
            check(self._lib.dex_get_kernel_stats(self._ctx, i, C.byref(ms), C.byref(n), C.byref(by),
                                                 C.byref(fl)))
            out[name] = (ms.value, n.value, by.value, fl.value)
        return out

    def kernel_totals(self) -> dict:
        """{class: (launches, algorithmic bytes, algorithmic flops)} over every
        launch since the context was created or the stats were reset."""
        out = {}
        for i, name in enumerate(_lib.KERNEL_CLASSES):
            n, by, fl = C.c_int64(), C.c_double(), C.c_double()
            check(self._lib.dex_get_kernel_totals(self._ctx, i, C.byref(n), C.byref(by), C.byref(fl)))
            out[name] = (n.value, by.value, fl.value)
        return out

    def reset_kernel_stats(self) -> None:
        check(self._lib.dex_reset_kernel_stats(self._ctx))

    def residuals_device(self, x, e: np.ndarray) -> np.ndarray:
        nev = e.size
        ee = np.ascontiguousarray(e, dtype=np.complex128)
        _check_dev(x, nev, self.d, self.device)
        out = np.zeros(nev)
        check(self._lib.dex_residuals_device(self._ctx, _dev(x), ee.view(np.float64).ctypes.data_as(_dp),
                                             nev, _ptr(out)))
        return out

    def precblock_device(self, x, px, nev: int, sigma: complex, inner_iter: int, symmetry: int = 0) -> None:
        """PRECBLOCK (gplmr.F90:466-502) on device blocks: px = H x for
        inner_iter = 1, else the order-inner_iter block minimal-residual
        approximation to (H - sigma)^{-1} x (PRECBLOCK_BASIC)."""
        _check_dev(x, nev, self.d, self.device, name="x")
        _check_dev(px, nev, self.d, self.device, name="px")
        if x.data_ptr() == px.data_ptr():
            raise ValueError("x and px must not alias")
        check(self._lib.dex_precblock_device(self._ctx, _dev(x), _dev(px), self.m, self.m, nev,
                                             float(sigma.real), float(sigma.imag), inner_iter, symmetry))

    def diagsubspace_device(self, orbitals: np.ndarray, numout: int, etarget: complex, singlet: bool = True,
                            hermopt: bool = False):
        """DIAGSUBSPACE (hammod.F90:1125-1256) + GET_EXPECTS: start vectors from
        the (anti)symmetrised products of one-electron orbitals (M x numorbs).
        Returns (psi: torch complex128 of numout * D on the device, e0, expects)."""
        import torch

        orb = np.asarray(orbitals, dtype=np.complex128)
        if orb.ndim != 2 or orb.shape[0] != self.m:
            raise DexError(_lib.DEX_ESIZE, "orbitals must be an M x numorbs array")
        k = orb.shape[1]
        orb = np.ascontiguousarray(orb.reshape(-1, order="F"))  # column-major, flat
        psi = torch.empty(self.d * numout, dtype=torch.complex128, device=f"cuda:{self.device}")
        e0 = np.zeros(numout, dtype=np.complex128)
        ex = np.zeros(numout, dtype=np.complex128)
        check(self._lib.dex_diagsubspace_device(self._ctx, orb.view(np.float64).ctypes.data_as(_dp), k,
                                                int(singlet), int(hermopt), float(np.real(etarget)),
                                                float(np.imag(etarget)), numout, _dev(psi),
                                                e0.view(np.float64).ctypes.data_as(_dp),
                                                ex.view(np.float64).ctypes.data_as(_dp)))
        return psi, e0, ex

    def heplus_lanczos(self, numeigs: int, blocksize: int = 3, kdim: int = 501, check: int = 10,
                       tol: float = 1e-8, seed: int = 1, allow_unconverged: bool = False):
        """BLOCKLANCZOS (lanmod.F90:18-219) of the one-electron operator on the
        device.  Returns (eigenvalues (numeigs), orbitals (M x numeigs),
        residuals (numeigs), iterations)."""
        kdim = min(kdim, self.m // blocksize)
        e = np.zeros(numeigs, dtype=np.complex128)
        flat = np.zeros(self.m * numeigs, dtype=np.complex128)  # column-major M x numeigs
        res = np.zeros(numeigs)
        it = C.c_int32()
        st = self._lib.dex_heplus_lanczos_device(self._ctx, blocksize, kdim, check, float(tol), numeigs, seed,
                                                 e.view(np.float64).ctypes.data_as(_dp),
                                                 flat.view(np.float64).ctypes.data_as(_dp), _ptr(res),
                                                 C.byref(it))
        _lib.check(st, allow=(_lib.DEX_OK, _lib.DEX_ENOTCONVERGED) if allow_unconverged else (_lib.DEX_OK,))
        return e, flat.reshape(self.m, numeigs, order="F"), res, it.value

    def footprint(self, nev: int, ns: int = 1, inner_iter: int = 0, inner_block: int = 0,
                  guard: int = 0, inner_fp32: bool = False) -> int:
        o = self._opts(False, False, 0, 2, 1e-5, 0, False, inner_block, guard, inner_fp32)
        return int(self._lib.dex_gplmr_footprint(self._ctx, nev, ns, inner_iter, C.cast(C.byref(o), C.c_void_p)))

    # -- eigensolver --------------------------------------------------------------------
    @staticmethod
    def _opts(strict: bool, verbose: bool, refresh_every: int, fp64_restarts: int, fp32_tol: float,
              symmetry: int = 0, lock: bool = False, inner_block: int = 0, guard: int = 0,
              inner_fp32: bool = False):
        o = _lib.DexSolveOpts()
        _lib.load().dex_solve_opts_default(C.byref(o))
        o.symmetry = int(symmetry)
        o.lock = int(lock)
        o.inner_block = int(inner_block)
        o.guard = int(guard)
        o.inner_fp32 = int(inner_fp32)
        o.strict_cholesky = int(strict)
        o.verbose = int(verbose)
        o.refresh_every = refresh_every
        o.fp64_restarts = fp64_restarts
        o.fp32_tol = fp32_tol
        return o

    def _result(self, st, r, e):
        return SolveResult(e=e, residuals=None, iterations=r.iterations, converged=bool(r.converged),
                           matvecs=r.matvecs, max_residual=r.max_residual, t_total=r.t_total,
                           t_matvec=r.t_matvec, t_prec=r.t_prec, t_gram=r.t_gram, t_update=r.t_update,
                           t_smalleig=r.t_smalleig, fallbacks=r.fallbacks,
                           fp64_restarts_used=r.fp64_restarts_used, status=st)

    def gplmr_eig(self, x: np.ndarray, sigma: complex, nev: int, ns: int = 1, maxit: int = 100,
                  tol: float = 1e-12, inner_iter: int = 0, outmodulus: int = 5, strict: bool = False,
                  verbose: bool = False, refresh_every: int = 0, fp64_restarts: int = 0,
                  fp32_tol: float = 1e-5, m: int | None = None, mb: int | None = None,
                  symmetry: int = 0, lock: bool = False, inner_block: int = 0,
                  guard: int = 0, inner_fp32: bool = False) -> SolveResult:
        """GPLMR_EIG on a host block ``x`` (D*NEV complex128, vector-major), updated in place."""
        if x.dtype != np.complex128 or not (x.flags.f_contiguous or x.flags.c_contiguous):
            raise TypeError("x must be a contiguous complex128 array")
        m = self.m if m is None else m
        mb = self.m if mb is None else mb
        if x.size != m * mb * nev:
            raise DexError(_lib.DEX_ESIZE, "x.size != M*MB*NEV")
        e = np.zeros(nev, dtype=np.complex128)
        r = _lib.DexResult()
        o = self._opts(strict, verbose, refresh_every, fp64_restarts, fp32_tol, symmetry, lock, inner_block,
                       guard, inner_fp32)
        st = self._lib.dex_gplmr_eig(self._ctx, _ptr(x), e.view(np.float64).ctypes.data_as(_dp), m, mb,
                                     float(np.real(sigma)), float(np.imag(sigma)), nev, ns, maxit, tol,
                                     inner_iter, outmodulus, C.byref(o), C.byref(r))
        check(st, allow=(_lib.DEX_OK, _lib.DEX_ENOTCONVERGED))
        return self._result(st, r, e)

    def gplmr_eig_device(self, x, sigma: complex, nev: int, ns: int = 1, maxit: int = 100,
                         tol: float = 1e-12, inner_iter: int = 0, outmodulus: int = 5,
                         strict: bool = False, verbose: bool = False, refresh_every: int = 0,
                         fp64_restarts: int = 0, fp32_tol: float = 1e-5,
                         symmetry: int = 0, lock: bool = False, inner_block: int = 0,
                         guard: int = 0, inner_fp32: bool = False) -> SolveResult:
        """GPLMR_EIG with the block resident on the device (torch complex128 tensor)."""
        _check_dev(x, nev, self.d, self.device)
        e = np.zeros(nev, dtype=np.complex128)
        r = _lib.DexResult()
        o = self._opts(strict, verbose, refresh_every, fp64_restarts, fp32_tol, symmetry, lock, inner_block,
                       guard, inner_fp32)
        st = self._lib.dex_gplmr_eig_device(self._ctx, _dev(x),
                                            e.view(np.float64).ctypes.data_as(_dp), self.m, self.m,
                                            float(np.real(sigma)), float(np.imag(sigma)), nev, ns,
                                            maxit, tol, inner_iter, outmodulus, C.byref(o), C.byref(r))
        check(st, allow=(_lib.DEX_OK, _lib.DEX_ENOTCONVERGED))
        return self._result(st, r, e)
