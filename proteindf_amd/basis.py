"""Orbital basis handling: the basis2 library format and the AO table.

Reader for ProteinDF's basis2 format (Fl_Db_Basis.cpp:36-112) and the AO
table the grid engine consumes, laid out as TlOrbitalInfoObject builds it
(TlOrbitalInfoObject.cpp:181-259): atom order, then the CGTOs of the atom's
basis in library order (all s, then p, then d), then the 2l+1 components with
basis type shellType^2 + k (0 s; 1-3 px py pz; 4 dxy 5 dxz 6 dyz 7 dx2-y2
8 dz2).  Contraction coefficients are normalised and primitives sorted by
ascending exponent by a caller-supplied normaliser (the engine library's
pdfxc_normalize_cgto in the product, the oracle in tests).
"""
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

SHELL_L = {"s": 0, "p": 1, "d": 2}
DEFAULT_BASIS2 = Path(__file__).resolve().parent / "data" / "basis2_6-31G_HCNO.txt"


@dataclass
class CGTO:
    shell: str
    exps: np.ndarray
    coefs: np.ndarray


def read_basis2(path, name):
    """Read one entry (e.g. 'O-6-31G.C') of a basis2-format file."""
    lines = []
    for raw in Path(path).read_text().splitlines():
        s = raw.strip()
        if not s or s.startswith("#"):
            continue
        lines.append(s)
    try:
        i = lines.index(name)
    except ValueError:
        raise KeyError(f"the orbital basis set '{name}' is not found in {path}") from None
    counts = [int(t) for t in lines[i + 1].split()][:5]
    counts += [0] * (5 - len(counts))
    out = []
    j = i + 2
    for shell, cnt in zip("spdfg", counts):
        for _ in range(cnt):
            nprim = int(lines[j].split()[0])
            rows = [lines[j + 1 + k].split() for k in range(nprim)]
            out.append(CGTO(shell, np.array([float(r[0]) for r in rows]),
                            np.array([float(r[1]) for r in rows])))
            j += 1 + nprim
    return out


def basis_631gd(symbol, path=DEFAULT_BASIS2, d_exponent=0.8):
    """6-31G from the basis2 library plus one uncontracted d (exponent 0.8) on heavy atoms."""
    cg = read_basis2(path, f"O-6-31G.{symbol}")
    if symbol not in ("H", "He"):
        cg.append(CGTO("d", np.array([d_exponent]), np.array([1.0])))
    return cg


@dataclass
class OrbitalInfo:
    """Per-AO table (mirrors TlOrbitalInfo's getters used on the grid path)."""
    ao_atom: np.ndarray
    ao_center: np.ndarray
    ao_type: np.ndarray
    ao_nprim: np.ndarray
    ao_exps: list = field(default_factory=list)
    ao_coefs: list = field(default_factory=list)

    @property
    def n_aos(self):
        return len(self.ao_type)

    @classmethod
    def build(cls, mol, normalizer, basis_for=basis_631gd, labels=None):
        """basis_for(symbol) gives an atom's CGTOs; with labels (one string per
        atom) it is called as basis_for(symbol, label), the reference's
        "symbol@label" basis keys (TlOrbitalInfoObject.cpp:199-212)."""
        cache = {}
        atom, center, typ, nprim, exps, coefs = [], [], [], [], [], []
        for a, (s, p) in enumerate(zip(mol.symbols, mol.xyz)):
            if s == "X":
                continue
            key = (s,) if labels is None else (s, labels[a])
            if key not in cache:
                shells = []
                for cg in basis_for(*key):
                    if cg.shell not in SHELL_L:
                        raise ValueError(f"shell '{cg.shell}' of atom {a} ({s}): only s, p and d shells are supported")
                    e, c = normalizer(SHELL_L[cg.shell], cg.exps, cg.coefs)
                    shells.append((SHELL_L[cg.shell], np.asarray(e), np.asarray(c)))
                cache[key] = shells
            for l, e, c in cache[key]:
                for k in range(2 * l + 1):
                    atom.append(a)
                    center.append(p)
                    typ.append(l * l + k)
                    nprim.append(len(e))
                    exps.append(e)
                    coefs.append(c)
        return cls(np.array(atom, dtype=np.int32), np.array(center, dtype=np.float64),
                   np.array(typ, dtype=np.int32), np.array(nprim, dtype=np.int32), exps, coefs)

    def padded(self, max_prim):
        n = self.n_aos
        E = np.zeros((n, max_prim))
        C = np.zeros((n, max_prim))
        for i in range(n):
            k = self.ao_nprim[i]
            E[i, :k] = self.ao_exps[i]
            C[i, :k] = self.ao_coefs[i]
        return E, C
