"""He+ orbitals from the library's host tables: the dense one-electron
Hamiltonian equals the oracle's, and its lowest eigenvalues reproduce the
reference's logged Lanczos values (Out.S.N5:51-64) at NN=11, delta=2,
theta=0.2."""
import json

import numpy as np

from doublyexcited_amd.orbitals import heplus_orbitals, one_electron_hamiltonian
from oracle.hamiltonian import Hamiltonian


def test_one_electron_hamiltonian_equals_oracle():
    h = Hamiltonian(2, 0.45, 0.3)
    np.testing.assert_allclose(one_electron_hamiltonian(2, 0.45, 0.3), h.dense_one_electron(),
                               rtol=0, atol=1e-12)


def test_heplus_orbitals_reproduce_reference_log(golden_dir):
    g = json.loads((golden_dir / "heplus_nn11_d2_t0.2.json").read_text())
    logv = np.array([complex(a, b) for a, b in g["heplus_eigenvalues"]])
    w, v = heplus_orbitals(5, 2.0, 0.2, 14, method="dense")
    # converged levels to 1e-12; the log's LANTOL 1e-2 members agree to its own spread
    assert abs(w[0] - logv[0]) < 1e-12 and abs(w[1] - logv[1]) < 1e-12
    assert np.max(np.abs(w - logv)) < 1e-5
    h1 = one_electron_hamiltonian(5, 2.0, 0.2)
    assert np.max(np.abs(h1 @ v - v * w[None, :])) < 1e-10
