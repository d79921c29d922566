"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/pdfxc.h declares, and its host-only routines behave (no GPU
work is launched here)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from proteindf_amd import engine

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "pdfxc.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdfxc_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(engine.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(engine.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing


def test_default_opts():
    o = engine.default_opts()
    assert o.density_cutoff == 1e-16 and o.ao_screen == 0.0
    assert o.precision == engine.PDFXC_FP64 and o.block_points == 128


def test_normalisation_matches_oracle():
    from oracle import oracle
    from proteindf_amd.basis import SHELL_L, basis_631gd

    for sym in ("H", "C", "N", "O"):
        for cg in basis_631gd(sym):
            e1, c1 = engine.normalize_cgto(SHELL_L[cg.shell], cg.exps, cg.coefs)
            e2, c2 = oracle.normalize_cgto(SHELL_L[cg.shell], cg.exps, cg.coefs)
            np.testing.assert_array_equal(e1, e2)
            np.testing.assert_allclose(c1, c2, rtol=1e-15)


def test_error_codes_are_reported():
    with pytest.raises(engine.PdfxcError) as ei:
        engine.normalize_cgto(3, [1.0], [1.0])
    assert ei.value.code == 4 and "shell" in str(ei.value)
    with pytest.raises(engine.PdfxcError) as ei:
        engine.normalize_cgto(0, [], [])
    assert ei.value.code == 1


def test_precision_modes_are_validated_before_any_device_work():
    from proteindf_amd import calc_grid, molecules

    assert (engine.PDFXC_FP64, engine.PDFXC_SPLIT_BF16, engine.PDFXC_SPLIT_BF16X3,
            engine.PDFXC_SPLIT_K_BF16X3, engine.PDFXC_SPLIT_BFP, engine.PDFXC_SPLIT_K_F16) == (0, 1, 2, 3, 4, 5)
    info = calc_grid.orbital_info(molecules.water_cluster(1, 0))
    for bad in (-1, 6, 7):
        with pytest.raises(engine.PdfxcError) as ei:
            engine.XcContext(info, engine.default_opts(precision=bad))
        assert ei.value.code == 1 and "precision" in str(ei.value)


def test_functional_ids_follow_the_reference_xc_type():
    # XC_TYPE of DfXCFunctional.h:44: UNDEFINED_XC, HF, HFS, SVWN, HFB, BLYP, B3LYP
    assert (engine.PDFXC_HF, engine.PDFXC_HFS, engine.PDFXC_SVWN, engine.PDFXC_HFB,
            engine.PDFXC_BLYP, engine.PDFXC_B3LYP) == (1, 2, 3, 4, 5, 6)
    assert engine.XC_TYPES["B88LYP"] == engine.XC_TYPES["BLYP"]
    # an unknown functional is refused before any device work
    for bad in (0, 7, -3):
        with pytest.raises(engine.PdfxcError) as ei:
            engine.eval_functional([0.1], [0.01], bad)
        assert ei.value.code == 4 and "unknown functional" in str(ei.value)


def test_xc_functional_names_follow_the_reference():
    from proteindf_amd import calc_grid

    assert calc_grid.DfXCFunctional(None, 0, "B3LYP").getFockExchangeCoefficient() == 0.2
    assert calc_grid.DfXCFunctional(None, 0, "HF").isHybridFunctional()
    assert not calc_grid.DfXCFunctional(None, 0, "SVWN~").isHybridFunctional()
    with pytest.raises(ValueError):
        calc_grid.DfXCFunctional(None, 0, "PBE")
