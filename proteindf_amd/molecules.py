"""Synthetic molecules of the benchmark workloads.

* Water clusters (H2O)_n: an ideal monomer (O-H 0.957 A, H-O-H 104.5 deg)
  placed at centres drawn uniformly in a cube, a draw being rejected when any
  O-O distance is below 2.7 A, each monomer turned by a uniform random
  rotation (scipy Rotation.random on the same generator).  For n > 8 the cube
  side grows as 9 A * (n/8)^(1/3) so that the packing density of the
  (H2O)_8 / 9 A case is kept.
* alpha-helical polyglycine H-(Gly)_n-OH built atom by atom with the natural
  extension reference frame (NeRF) from standard backbone geometry
  (N-CA 1.458, CA-C 1.525, C-N 1.329, C=O 1.231, N-H 1.01, CA-H 1.09 A),
  phi = -57, psi = -47, omega = 180 deg, then jittered by N(0, 0.02 A) per
  coordinate.  7n + 3 atoms.

Coordinates are returned in bohr, converted with 0.52917721067 A/bohr as the
reference's input parser does (PdfUserInput.cpp:30, :339-342).
"""
from dataclasses import dataclass

import numpy as np

BOHR_ANGSTROM = 0.52917721067
ELEMENTS = ("X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
            "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar")
# atomic numbers up to Ar, the elements the grid's Bragg-Slater table covers; "X" is a dummy atom
Z_OF = {s: z for z, s in enumerate(ELEMENTS)}


@dataclass
class Molecule:
    symbols: list
    xyz: np.ndarray  # (n_atoms, 3) bohr
    name: str = ""

    @property
    def Z(self):
        return np.array([Z_OF[s] for s in self.symbols], dtype=np.int32)

    @property
    def n_atoms(self):
        return len(self.symbols)

    @property
    def n_electrons(self):
        return int(self.Z.sum())


def _water_monomer():
    r, ang = 0.957, np.deg2rad(104.5)
    o = np.zeros(3)
    h1 = np.array([r * np.sin(ang / 2), 0.0, r * np.cos(ang / 2)])
    h2 = np.array([-r * np.sin(ang / 2), 0.0, r * np.cos(ang / 2)])
    return np.stack([o, h1, h2])


def water_cluster_angstrom(n, rng, cube=None, min_oo=2.7):
    from scipy.spatial.transform import Rotation

    if cube is None:
        cube = 9.0 * (max(n, 8) / 8.0) ** (1.0 / 3.0)
    centres = []
    tries = 0
    while len(centres) < n:
        c = rng.uniform(0.0, cube, size=3)
        tries += 1
        if tries > 1_000_000:
            raise RuntimeError("water cluster placement did not converge")
        if all(np.linalg.norm(c - q) >= min_oo for q in centres):
            centres.append(c)
    mono = _water_monomer()
    xyz, sym = [], []
    for c in centres:
        R = Rotation.random(random_state=rng).as_matrix()
        for s, p in zip("OHH", mono):
            sym.append(s)
            xyz.append(R @ p + c)
    return sym, np.array(xyz)


def water_cluster(n, seed=0):
    rng = np.random.default_rng(seed)
    sym, xyz = water_cluster_angstrom(n, rng)
    return Molecule(sym, xyz / BOHR_ANGSTROM, f"(H2O)_{n} seed {seed}")


def water_cluster_batch(count=64, nmin=4, nmax=64, seed=3):
    """Many clusters from one generator stream: n ~ U{nmin..nmax} then the cluster."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        sym, xyz = water_cluster_angstrom(n, rng)
        out.append(Molecule(sym, xyz / BOHR_ANGSTROM, f"(H2O)_{n}"))
    return out


def _place(a, b, c, bond, angle_deg, torsion_deg):
    """NeRF: position d with |cd| = bond, angle(b,c,d) = angle, torsion(a,b,c,d) = torsion."""
    ang, tor = np.deg2rad(angle_deg), np.deg2rad(torsion_deg)
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * np.cos(ang), bond * np.sin(ang) * np.cos(tor),
                   bond * np.sin(ang) * np.sin(tor)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def polyglycine_helix_angstrom(n, phi=-57.0, psi=-47.0, omega=180.0):
    b_nca, b_cac, b_cn, b_co, b_nh, b_cah, b_coh, b_oh = 1.458, 1.525, 1.329, 1.231, 1.01, 1.09, 1.34, 0.97
    a_ncac, a_cacn, a_cnca = 111.2, 116.2, 121.7
    sym, xyz = [], []
    # first three backbone atoms in a fixed frame
    N = np.array([0.0, 0.0, 0.0])
    CA = np.array([b_nca, 0.0, 0.0])
    t = np.deg2rad(180.0 - a_ncac)
    C = CA + b_cac * np.array([np.cos(t), np.sin(t), 0.0])
    backbone = []
    for i in range(n):
        if i > 0:
            Cp, CAp, Np = backbone[-1][2], backbone[-1][1], backbone[-1][0]
            N = _place(Np, CAp, Cp, b_cn, a_cacn, psi)
            CA = _place(CAp, Cp, N, b_nca, a_cnca, omega)
            C = _place(Cp, N, CA, b_cac, a_ncac, phi)
        backbone.append((N, CA, C))
    for i, (N, CA, C) in enumerate(backbone):
        sym.append("N")
        xyz.append(N)
        if i == 0:
            # N-terminal NH2
            sym += ["H", "H"]
            xyz.append(_place(C, CA, N, b_nh, 109.5, 180.0))
            xyz.append(_place(C, CA, N, b_nh, 109.5, 60.0))
        else:
            Cp = backbone[i - 1][2]
            u = (N - Cp) / np.linalg.norm(N - Cp) + (N - CA) / np.linalg.norm(N - CA)
            sym.append("H")
            xyz.append(N + b_nh * u / np.linalg.norm(u))
        sym += ["C", "H", "H"]
        xyz.append(CA)
        xyz.append(_place(C, N, CA, b_cah, 109.5, 120.0))
        xyz.append(_place(C, N, CA, b_cah, 109.5, -120.0))
        sym += ["C", "O"]
        xyz.append(C)
        xyz.append(_place(N, CA, C, b_co, 120.8, psi + 180.0))
    # C-terminal OH
    N, CA, C = backbone[-1]
    OXT = _place(N, CA, C, b_coh, 111.0, psi)
    sym += ["O", "H"]
    xyz.append(OXT)
    xyz.append(_place(CA, C, OXT, b_oh, 109.0, 180.0))
    return sym, np.array(xyz)


def polyglycine_helix(n, seed):
    sym, xyz = polyglycine_helix_angstrom(n)
    rng = np.random.default_rng(seed)
    xyz = xyz + rng.normal(0.0, 0.02, size=xyz.shape)
    return Molecule(sym, xyz / BOHR_ANGSTROM, f"(Gly)_{n} alpha-helix seed {seed}")
