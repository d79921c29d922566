"""Copy the 6-31G entries for H, C, N, O out of the reference's basis library.

ProteinDF reads orbital basis sets from data/basis2 (Fl_Db_Basis.cpp:36-112):
a name line such as "O-6-31G.C", a line of per-shell CGTO counts (s p d f g),
then for each CGTO the number of primitives followed by "exponent coefficient"
lines.  The library has no 6-31G(d); the workloads add one uncontracted d
shell (exponent 0.8, Hariharan-Pople) on C, N and O in the host code.

This script copies the four entries verbatim (they are published basis-set
data) into proteindf_amd/data/basis2_6-31G_HCNO.txt, in the same format, so that the
GPU runs do not need the reference tree:

    python tests/golden/extract_basis.py /root/reference > proteindf_amd/data/basis2_6-31G_HCNO.txt
"""
import sys
from pathlib import Path

NAMES = ["O-6-31G.H", "O-6-31G.C", "O-6-31G.N", "O-6-31G.O"]


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    lines = (ref / "data" / "basis2").read_text().splitlines()
    out = ["# 6-31G for H, C, N, O copied from ProteinDF data/basis2 (basis2 format)"]
    for name in NAMES:
        i = lines.index(name)
        counts = [int(t) for t in lines[i + 1].split()]
        block = [lines[i], lines[i + 1]]
        j = i + 2
        for _ in range(sum(counts)):
            nprim = int(lines[j].split()[0])
            block.extend(lines[j:j + 1 + nprim])
            j += 1 + nprim
        out.extend(block)
        out.append("")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
