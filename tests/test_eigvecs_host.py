"""Eigenvector files in the reference's format (READWRITE_EIGENVECTORS and
SIXD_FIX, utils.F90:82-198): round trips, an independent reader (scipy's
Fortran unformatted-sequential reader) for the record framing, the grid
change against explicit 6-D slicing, and the 12-byte header of readers that
keep the size field in 4 bytes."""
import numpy as np
import pytest
from scipy.io import FortranFile

from doublyexcited_amd import _lib
from doublyexcited_amd.eigvecs import eigenvector_file_info, read_eigenvectors, write_eigenvectors


def _vectors(n, num, seed):
    m = (2 * n + 1) ** 3
    rng = np.random.default_rng(seed)
    return rng.standard_normal((num, m * m)) + 1j * rng.standard_normal((num, m * m))


def test_round_trip_and_independent_reader(tmp_path):
    psi = _vectors(1, 3, 1)
    f = tmp_path / "eigvecs.S.N1"
    write_eigenvectors(f, psi)
    assert eigenvector_file_info(f) == (27, 3)
    np.testing.assert_array_equal(read_eigenvectors(f), psi)
    with FortranFile(f, "r") as ff:
        head = ff.read_record(np.dtype([("m", "<i4"), ("num", "<i4"), ("size", "<i8")]))
        assert head["m"][0] == 27 and head["num"][0] == 3 and head["size"][0] == 27 * 27 * 16
        for v in range(3):
            np.testing.assert_array_equal(ff.read_record("<c16"), psi[v])


def test_sixd_fix_matches_six_dimensional_slicing(tmp_path):
    psi = _vectors(1, 2, 2)  # NN = 3
    f = tmp_path / "small"
    write_eigenvectors(f, psi)
    big = read_eigenvectors(f, m=125)  # NN = 5
    for v in range(2):
        a = psi[v].reshape((3,) * 6, order="F")
        b = big[v].reshape((5,) * 6, order="F")
        centre = (slice(1, 4),) * 6
        np.testing.assert_array_equal(b[centre], a)
        b2 = b.copy()
        b2[centre] = 0
        assert not np.any(b2)
    # and back: truncation to the central block
    f2 = tmp_path / "big"
    write_eigenvectors(f2, big)
    np.testing.assert_array_equal(read_eigenvectors(f2, m=27), psi)


def test_reads_a_four_byte_size_field(tmp_path):
    psi = _vectors(1, 1, 3)
    f = tmp_path / "i4header"
    with FortranFile(f, "w") as ff:
        ff.write_record(np.array([27, 1, 27 * 27 * 16], dtype="<i4"))
        ff.write_record(psi[0].astype("<c16"))
    np.testing.assert_array_equal(read_eigenvectors(f), psi)


def test_inconsistent_header_is_a_size_error(tmp_path):
    f = tmp_path / "bad"
    with FortranFile(f, "w") as ff:
        ff.write_record(np.array([27, 1, 1234], dtype="<i4"))
        ff.write_record(np.zeros(27 * 27, dtype="<c16"))
    with pytest.raises(_lib.DexError) as exc:
        read_eigenvectors(f)
    assert exc.value.status == _lib.DEX_ESIZE
    # an M that is not an odd cube cannot be fixed
    g = tmp_path / "ok"
    write_eigenvectors(g, _vectors(1, 1, 4))
    with pytest.raises(_lib.DexError):
        read_eigenvectors(g, m=64)


def test_sixd_fix_alone_equals_the_file_route(tmp_path):
    from doublyexcited_amd.eigvecs import sixd_fix

    psi = _vectors(1, 2, 5)
    f = tmp_path / "v"
    write_eigenvectors(f, psi)
    np.testing.assert_array_equal(sixd_fix(psi, 1, 2), read_eigenvectors(f, m=125))
