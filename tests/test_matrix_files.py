"""ProteinDF matrix files through the engine library's readers and writers.

tests/golden/{M,V,MV,VMV}.mat are the reference's own unit-test data
(data/unit_test/, used by tl_dense_symmetric_matrix_lapack_test.cc:255-298):
M is RLHD and V is RSFD, both with 4-byte type and count fields, and the
reference checks M V = MV and V^T M V = VMV to 1e-5.  Reading them through
the engine and repeating those products pins the header detection and both
element orders (V is not symmetric, so a transposed read fails).  Writers are
pinned by the reference's save layout (1-byte type, 32-bit counts,
tl_dense_symmetric_matrix_object.cc:292-312, tl_dense_general_matrix_object.cc
:239-266) and by reading back.  No GPU is used.
"""
from pathlib import Path

import numpy as np
import pytest

from proteindf_amd import engine, guess

GOLDEN = Path(__file__).parent / "golden"


def test_reference_fixtures_multiply_as_the_reference_checks():
    assert engine.matrix_file_info(GOLDEN / "M.mat") == (engine.PDFXC_MAT_RLHD, 190, 190)
    assert engine.matrix_file_info(GOLDEN / "V.mat") == (engine.PDFXC_MAT_RSFD, 190, 190)
    M = guess.unpack_lower(engine.load_symmetric_matrix(GOLDEN / "M.mat"), 190)
    V = engine.load_general_matrix(GOLDEN / "V.mat")
    MV = engine.load_general_matrix(GOLDEN / "MV.mat")
    VMV = guess.unpack_lower(engine.load_symmetric_matrix(GOLDEN / "VMV.mat"), 190)
    assert np.abs(V - V.T).max() > 1.0
    np.testing.assert_allclose(M @ V, MV, rtol=0, atol=1e-10)
    np.testing.assert_allclose(V.T @ M @ V, VMV, rtol=0, atol=1e-9)
    # a transposed V would not reproduce MV
    assert np.abs(M @ V.T - MV).max() > 1.0


def test_symmetric_save_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n = 37
    packed = rng.normal(size=n * (n + 1) // 2)
    path = tmp_path / "F.rks1.mat"
    engine.save_symmetric_matrix(path, packed)
    raw = path.read_bytes()
    assert raw[:9] == bytes([2]) + np.int32(n).tobytes() + np.int32(n).tobytes()
    np.testing.assert_array_equal(np.frombuffer(raw[9:], dtype="<f8"), packed)
    np.testing.assert_array_equal(engine.load_symmetric_matrix(path), packed)


def test_general_save_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    m = rng.normal(size=(11, 7))
    path = tmp_path / "grid.mat"
    engine.save_general_matrix(path, m)
    raw = path.read_bytes()
    assert raw[:9] == bytes([1]) + np.int32(11).tobytes() + np.int32(7).tobytes()
    np.testing.assert_array_equal(np.frombuffer(raw[9:], dtype="<f8"), m.T.ravel())  # column-major
    np.testing.assert_array_equal(engine.load_general_matrix(path), m)


@pytest.mark.parametrize("tw,iw", [(1, 4), (1, 8), (4, 4), (4, 8), (8, 4), (8, 8)])
def test_every_header_width_the_reference_accepts(tmp_path, tw, iw):
    rng = np.random.default_rng(tw * 10 + iw)
    m = rng.normal(size=(5, 3))
    tdt, idt = {1: np.int8, 4: np.int32, 8: np.int64}[tw], {4: np.int32, 8: np.int64}[iw]
    path = tmp_path / "m.mat"
    # RSFD body: row-major
    path.write_bytes(tdt(0).tobytes() + idt(5).tobytes() + idt(3).tobytes() + m.astype("<f8").tobytes())
    assert engine.matrix_file_info(path) == (engine.PDFXC_MAT_RSFD, 5, 3)
    np.testing.assert_array_equal(engine.load_general_matrix(path), m)
    # RLHD
    n = 6
    packed = rng.normal(size=n * (n + 1) // 2)
    path.write_bytes(tdt(2).tobytes() + idt(n).tobytes() + idt(n).tobytes() + packed.tobytes())
    np.testing.assert_array_equal(engine.load_symmetric_matrix(path), packed)


def test_bad_files_are_refused(tmp_path):
    path = tmp_path / "bad.mat"
    path.write_bytes(bytes([1]) + np.int32(4).tobytes() + np.int32(4).tobytes() + np.zeros(15).tobytes())
    with pytest.raises(engine.PdfxcError) as ei:
        engine.matrix_file_info(path)  # body one double short
    assert ei.value.code == engine.ERROR_CODES["UNSUPPORTED"]
    path.write_bytes(bytes([1]) + np.int32(4).tobytes() + np.int32(4).tobytes() + np.zeros(16).tobytes())
    with pytest.raises(engine.PdfxcError) as ei:
        engine.load_symmetric_matrix(path)  # CSFD is not a symmetric-matrix format
    assert ei.value.code == engine.ERROR_CODES["UNSUPPORTED"]
    engine.save_symmetric_matrix(path, np.zeros(10))
    with pytest.raises(engine.PdfxcError) as ei:
        engine.load_general_matrix(path)  # RLHD is not a general-matrix format
    assert ei.value.code == engine.ERROR_CODES["UNSUPPORTED"]
    with pytest.raises(engine.PdfxcError) as ei:
        engine.matrix_file_info(tmp_path / "missing.mat")
    assert ei.value.code == engine.ERROR_CODES["ARGUMENT"]


def test_dimension_mismatch_is_an_argument_error(tmp_path):
    path = tmp_path / "P.mat"
    engine.save_symmetric_matrix(path, np.zeros(10))  # n = 4
    out = np.zeros(15)
    rc = engine.lib().pdfxc_load_symmetric_matrix(str(path).encode(), 5, engine._p(out))
    assert rc == engine.ERROR_CODES["ARGUMENT"]
    assert "dimension 4" in engine.lib().pdfxc_last_error().decode()
