"""Pin the oracle's functionals against the reference's own known-answer tests.

tests/golden/functional_kats.json holds every active gtest case of
src/unit_test/DfFunctional_{B3LYP,Becke88,LYP,Slater,VWN}Test.cpp (inputs,
expected values, which outputs are checked, the file's EPS).  Closed-shell
("rks") cases are evaluated as the grid path evaluates them: the
spin-resolved functional with rhoB = rhoA and gAB = gBB = gAA
(DfFunctional.h:190-192, :221-232).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle

KATS = json.loads((Path(__file__).parent / "golden" / "functional_kats.json").read_text())["cases"]
OUT_INDEX = {"zk": 0, "vrhoA": 1, "vrhoB": 2, "vgAA": 3, "vgAB": 4, "vgBB": 5}


def _inputs(case):
    i = case["inputs"]
    return np.array([[i["rhoA"], i["rhoB"], i["gAA"], i["gAB"], i["gBB"]]])


@pytest.mark.parametrize("case", KATS, ids=[f"{c['functional']}.{c['test']}" for c in KATS])
def test_oracle_matches_reference_kat(case):
    out = oracle.functional(oracle.FUNCTIONAL_IDS[case["functional"]], _inputs(case))[0]
    for name in case["checked"]:
        exp = case["expected"][name]
        got = out[OUT_INDEX[name]]
        assert abs(got - exp) <= case["eps"], (name, got, exp)


def test_kat_inventory():
    # every reference functional test file contributes; B3LYP has UKS and RKS cases
    funcs = {c["functional"] for c in KATS}
    assert funcs == {"DfFunctional_B3LYP", "DfFunctional_Becke88", "DfFunctional_LYP",
                     "DfFunctional_Slater", "DfFunctional_VWN"}
    assert len(KATS) == 74
    assert {c["mode"] for c in KATS if c["functional"] == "DfFunctional_B3LYP"} == {"uks", "rks"}


def _fd_check(fid, x, h_rel=1e-5, tol=2e-6):
    """Central finite differences of zk against the analytic first derivatives."""
    base = oracle.functional(fid, x)[0]
    for k, oi in ((0, 1), (1, 2), (2, 3), (3, 4), (4, 5)):
        if x[0, k] == 0.0:
            continue
        h = abs(x[0, k]) * h_rel
        xp, xm = x.copy(), x.copy()
        xp[0, k] += h
        xm[0, k] -= h
        fd = (oracle.functional(fid, xp)[0, 0] - oracle.functional(fid, xm)[0, 0]) / (2 * h)
        ana = base[oi]
        assert abs(fd - ana) <= tol * max(1.0, abs(ana)), (fid, k, fd, ana)


# spin-resolved points with gAB != gAA so that each partial derivative is separate
FD_POINTS = [
    [0.35, 0.21, 0.08, 0.03, 0.05],
    [1.7, 1.1, 2.3, 0.9, 1.4],
    [0.02, 0.015, 1e-3, 4e-4, 8e-4],
    [4.5e-3, 3.0e-3, 2e-6, 1e-6, 1.5e-6],
]


@pytest.mark.parametrize("fid", [oracle.SLATER, oracle.BECKE88, oracle.BECKE88_EXCEPT_LDA,
                                 oracle.VWN5, oracle.VWN3, oracle.LYP, oracle.B3LYP,
                                 oracle.HFS, oracle.SVWN, oracle.BLYP])
@pytest.mark.parametrize("pt", FD_POINTS)
def test_oracle_derivatives_match_finite_differences(fid, pt):
    x = np.array([pt])
    if fid in (oracle.BECKE88, oracle.BECKE88_EXCEPT_LDA):
        # B88 has no gamma_AB dependence: its derivative is defined as 0
        x[0, 3] = 0.0
    _fd_check(fid, x)


def test_b3lyp_is_the_stated_linear_combination():
    x = np.array(FD_POINTS)
    b3 = oracle.functional(oracle.B3LYP, x)
    s = oracle.functional(oracle.SLATER, x)
    b = oracle.functional(oracle.BECKE88_EXCEPT_LDA, x)
    v = oracle.functional(oracle.VWN3, x)
    l = oracle.functional(oracle.LYP, x)
    comb = 0.80 * s + 0.72 * b + 0.19 * v + 0.81 * l
    np.testing.assert_allclose(b3, comb, rtol=1e-14, atol=1e-18)


def test_slater_scaling_closed_form():
    # e_x[lambda rho] = lambda^(4/3) e_x[rho] (homogeneity of Slater exchange)
    x = np.array([[0.3, 0.2, 0, 0, 0]])
    e1 = oracle.functional(oracle.SLATER, x)[0, 0]
    e2 = oracle.functional(oracle.SLATER, 2.0 * x)[0, 0]
    assert abs(e2 - 2.0 ** (4.0 / 3.0) * e1) < 1e-14


def test_b88_except_lda_difference_is_slater_like():
    # B88 - B88(except LDA) = G_PARAM (rhoA^4/3 + rhoB^4/3), G = -(3/2)(3/4pi)^(1/3)
    x = np.array([[0.4, 0.25, 0.1, 0.02, 0.07]])
    d = (oracle.functional(oracle.BECKE88, x) - oracle.functional(oracle.BECKE88_EXCEPT_LDA, x))[0, 0]
    G = -(1.5) * (3.0 / (4.0 * np.pi)) ** (1.0 / 3.0)
    assert abs(d - G * (0.4 ** (4 / 3) + 0.25 ** (4 / 3))) < 1e-14


def test_closed_shell_vwn3_has_no_spin_polarisation_terms():
    # at zeta = 0 the spin-interpolation terms vanish exactly: vrhoA == vrhoB
    x = np.array([[0.12, 0.12, 0, 0, 0]])
    o = oracle.functional(oracle.VWN3, x)[0]
    assert o[1] == o[2]


def test_composite_functionals_are_their_components():
    # DfFunctional_HFS = Slater, DfFunctional_SVWN = Slater + VWN (VWN5 sets),
    # DfFunctional_B88LYP = Becke88 (with its LDA part) + LYP, term by term;
    # each component is pinned by the reference's known answers above
    x = np.array(FD_POINTS)
    s = oracle.functional(oracle.SLATER, x)
    np.testing.assert_array_equal(oracle.functional(oracle.HFS, x), s)
    np.testing.assert_allclose(oracle.functional(oracle.SVWN, x), s + oracle.functional(oracle.VWN5, x),
                               rtol=1e-15, atol=0)
    np.testing.assert_allclose(oracle.functional(oracle.BLYP, x),
                               oracle.functional(oracle.BECKE88, x) + oracle.functional(oracle.LYP, x),
                               rtol=1e-15, atol=0)


def test_svwn_and_hfs_differ_by_vwn5_not_vwn3():
    # SVWN's correlation is DfFunctional_VWN (VWN5): at rs ~ 1 the paramagnetic
    # VWN5 and VWN3 energies per electron differ by more than 1 mEh
    rho = 3.0 / (4.0 * np.pi)  # rs = 1
    x = np.array([[rho / 2, rho / 2, 0, 0, 0]])
    ec = (oracle.functional(oracle.SVWN, x) - oracle.functional(oracle.HFS, x))[0, 0] / rho
    e5 = oracle.functional(oracle.VWN5, x)[0, 0] / rho
    e3 = oracle.functional(oracle.VWN3, x)[0, 0] / rho
    assert abs(ec - e5) < 1e-15 and abs(e5 - e3) > 1e-3
    # VWN5 fits Ceperley-Alder's unpolarised correlation, about -0.120 Ry = -0.060 Eh at rs = 1
    assert abs(e5 - (-0.0600)) < 5e-4
