"""The oracle's whole solve against the reference's logged grid-change run
(host only; minutes of CPU time, so it runs on request: DEX_SLOW_ORACLE=1).

The reference solved the helium singlet resonance at NN=11 (delta 2, theta 0.2)
and then at NN=13 from the NN=11 eigenvector mapped by SIXD_FIX
(CSCALE-WORK/Outs/Out.S.N5, Out.S.D2d0.T0.2d0.N6.nopetsc).  The oracle
follows the same recipe: the 14 lowest He+ orbitals -> the orbital-product
subspace vector nearest -0.76 (DIAGSUBSPACE) -> GPLMR with NEV 1, NS 5 and the
inner solve of order 5 at NN=11 -> SIXD_FIX to NN=13 -> its expectation value
-> GPLMR at NN=13 shifted by the real part the log chose.  Each value is
compared with the log within the accuracy its residual allows.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.skipif(os.environ.get("DEX_SLOW_ORACLE") != "1",
                                reason="minutes of host time: run with DEX_SLOW_ORACLE=1")


def test_grid_change_run_nn11_to_nn13(golden_dir):
    from doublyexcited_amd.eigvecs import sixd_fix
    from oracle.c_oracle import CHammult
    from oracle.gplmr import gplmr_eig
    from oracle.hamiltonian import Hamiltonian
    from oracle.startvec import diagsubspace, heplus_orbitals

    g = json.loads((golden_dir / "heplus_nn11_d2_t0.2.json").read_text())
    sub = g["subspace"]
    h11, h13 = Hamiltonian(5, 2.0, 0.2), Hamiltonian(6, 2.0, 0.2)
    a11, a13 = CHammult(h11, c_potential=True), CHammult(h13, c_potential=True)
    _, orb = heplus_orbitals(h11, sub["numorbs"])
    _, psi, ex = diagsubspace(a11, orb, singlet=True, hermopt=False, etarget=complex(*sub["etarget"]), numout=1)
    r11 = gplmr_eig(a11, psi, ex[0], 1, ns=5, maxit=80, tol=1e-9, inner_iter=5)
    print(f"NN=11: {r11.e[0]:.12f} after {r11.iters} iterations, residual {r11.resnrm[0]:.1e}", flush=True)
    assert r11.converged
    # Out.S.N5:147, at residual 8e-8
    assert abs(r11.e[0] - complex(*g["singlet_resonance"])) < 5e-8
    # SIXD_FIX to NN=13 and the expectation value the NN=13 run printed
    # (Out.S.D2d0.T0.2d0.N6.nopetsc:64-68): the NN=11 eigenvalue again
    x13 = sixd_fix(np.ascontiguousarray(r11.x.T), 5, 6).T
    hx = a13(np.asfortranarray(x13))
    e13_expect = np.vdot(x13[:, 0], hx[:, 0]) / np.vdot(x13[:, 0], x13[:, 0])
    print(f"NN=13 expectation of the padded vector: {e13_expect:.14f}", flush=True)
    assert abs(e13_expect - (-0.71315414287425 + 0.03522745948576j)) < 1e-7
    # the NN=13 solve, shifted by the log's chosen shift (:74), to the log's final value (:95-96,
    # at residual 8.3e-8)
    r13 = gplmr_eig(a13, x13, complex(-0.713154142874252, 0.0), 1, ns=5, maxit=80, tol=1e-9, inner_iter=5)
    print(f"NN=13: {r13.e[0]:.12f} after {r13.iters} iterations, residual {r13.resnrm[0]:.1e}", flush=True)
    assert r13.converged
    assert abs(r13.e[0] - (-0.713140523217 + 0.035275093419j)) < 1e-7
