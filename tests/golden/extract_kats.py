"""Extract the functional known-answer tests (KATs) of the reference into JSON.

The reference pins its XC functionals with gtest cases of the form

    TEST(DfFunctional_LYP, pointwise1) {
        const double dRhoA = 0.17E+01;
        ...
        const double zk = -0.179175399535E+00;
        ...
        DfFunctional_LYP f;
        double dFunctionalValue = f.getFunctional(dRhoA, dRhoB, ...);
        EXPECT_NEAR(zk, dFunctionalValue, EPS);
        ...
    }

(src/unit_test/DfFunctional_{B3LYP,Becke88,LYP,Slater,VWN}Test.cpp).  This
script reads those files as text and writes only the numbers: per active TEST,
the functional class, whether the spin-resolved (5-argument, "uks") or the
closed-shell (2-argument, "rks") entry point is called, the inputs, the
expected outputs, which outputs are actually checked by an active
EXPECT_NEAR, and the file's tolerance.  Commented-out lines are ignored.

Run (in a container that has the reference mounted):

    python tests/golden/extract_kats.py /root/reference > tests/golden/functional_kats.json
"""
import json
import re
import sys
from pathlib import Path

FILES = [
    "DfFunctional_B3LYPTest.cpp",
    "DfFunctional_Becke88Test.cpp",
    "DfFunctional_LYPTest.cpp",
    "DfFunctional_SlaterTest.cpp",
    "DfFunctional_VWNTest.cpp",
]
INPUTS = {"dRhoA": "rhoA", "dRhoB": "rhoB", "dGammaAA": "gAA",
          "dGammaAB": "gAB", "dGammaBB": "gBB"}
OUTPUTS = {"zk": "zk", "vRhoA": "vrhoA", "vRhoB": "vrhoB",
           "vGammaAA": "vgAA", "vGammaAB": "vgAB", "vGammaBB": "vgBB"}
# EXPECT_NEAR(expected, computed, EPS): map the computed variable to the output
COMPUTED = {"dFunctionalValue": "zk", "dRoundF_roundRhoA": "vrhoA",
            "dRoundF_roundRhoB": "vrhoB", "dRoundF_roundGammaAA": "vgAA",
            "dRoundF_roundGammaAB": "vgAB", "dRoundF_roundGammaBB": "vgBB"}

CONST_RE = re.compile(r"const\s+double\s+(\w+)\s*=\s*([-+0-9.eE]+)\s*;")
NEAR_RE = re.compile(r"EXPECT_NEAR\(\s*(\w+)\s*,\s*(\w+)\s*,\s*(\w+)\s*\)")
TEST_RE = re.compile(r"^TEST\(\s*(\w+)\s*,\s*(\w+)\s*\)")
EPS_RE = re.compile(r"^(?:static\s+)?const\s+double\s+EPS\s*=\s*([-+0-9.eE]+)\s*;")


def strip_comment(line: str) -> str:
    i = line.find("//")
    return line if i < 0 else line[:i]


def parse_file(path: Path):
    eps = None
    cases = []
    cur = None
    depth = 0
    in_block_comment = False
    for raw in path.read_text(encoding="utf-8", errors="replace").splitlines():
        line = raw
        if in_block_comment:
            if "*/" in line:
                line = line[line.index("*/") + 2:]
                in_block_comment = False
            else:
                continue
        if "/*" in line and "*/" not in line:
            in_block_comment = True
            line = line[:line.index("/*")]
        line = strip_comment(line).strip()
        if not line:
            continue
        if cur is None:
            m = EPS_RE.match(line)
            if m:
                eps = float(m.group(1))
                continue
            m = TEST_RE.match(line)
            if m:
                cur = {"functional": m.group(1), "test": m.group(2),
                       "inputs": {}, "expected": {}, "checked": [],
                       "mode": None, "file": path.name}
                depth = line.count("{") - line.count("}")
                continue
            continue
        depth += line.count("{") - line.count("}")
        m = CONST_RE.search(line)
        if m:
            name, val = m.group(1), float(m.group(2))
            if name in INPUTS:
                cur["inputs"][INPUTS[name]] = val
            elif name in OUTPUTS:
                cur["expected"][OUTPUTS[name]] = val
        if "f.getFunctional(" in line.replace(" ", ""):
            args = line.split("getFunctional(", 1)[1]
            cur["mode"] = "uks" if ("dRhoB" in args or "dGammaAB" in args) else "rks"
        for m in NEAR_RE.finditer(line):
            exp_name, got_name = m.group(1), m.group(2)
            if got_name in COMPUTED and exp_name in OUTPUTS:
                cur["checked"].append(COMPUTED[got_name])
        if depth <= 0:
            if cur["mode"] is not None and cur["checked"]:
                cur["eps"] = eps
                cases.append(cur)
            cur = None
    return cases


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = []
    for name in FILES:
        out.extend(parse_file(ref / "src" / "unit_test" / name))
    for c in out:
        if c["mode"] == "rks":
            # closed-shell entry point: rhoB = rhoA, gAB = gBB = gAA
            c["inputs"].setdefault("rhoB", c["inputs"]["rhoA"])
            for k in ("gAB", "gBB"):
                if "gAA" in c["inputs"]:
                    c["inputs"].setdefault(k, c["inputs"]["gAA"])
        for k in ("gAA", "gAB", "gBB"):
            c["inputs"].setdefault(k, 0.0)
    json.dump({"source": "src/unit_test/DfFunctional_*Test.cpp (active TESTs)",
               "cases": out}, sys.stdout, indent=1)
    sys.stdout.write("\n")


if __name__ == "__main__":
    main()
