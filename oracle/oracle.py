"""ctypes binding of the C oracle (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of bench.py use this
module; the product never imports it.  The shared library is built by
``make -C oracle`` (``__graft_entry__.build()`` runs that step).
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
MAX_PRIM = 16

SLATER, BECKE88, BECKE88_EXCEPT_LDA, VWN5, VWN3, LYP, B3LYP, HFS, SVWN, BLYP = range(10)
FUNCTIONAL_IDS = {
    "DfFunctional_HFS": HFS,
    "DfFunctional_SVWN": SVWN,
    "DfFunctional_B88LYP": BLYP,
    "DfFunctional_Slater": SLATER,
    "DfFunctional_Becke88": BECKE88,
    "DfFunctional_Becke88_ExceptLDA": BECKE88_EXCEPT_LDA,
    "DfFunctional_VWN": VWN5,
    "DfFunctional_VWN3": VWN3,
    "DfFunctional_LYP": LYP,
    "DfFunctional_B3LYP": B3LYP,
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        d, i64, i32 = ctypes.c_double, ctypes.c_int64, ctypes.c_int
        P = ctypes.c_void_p
        L.oracle_functional.argtypes = [i32, i64, P, P]
        L.oracle_normalize_cgto.argtypes = [i32, i32, P, P, P, P]
        L.oracle_eval_aos.argtypes = [P, i64, P, P, P, P, P]
        L.oracle_calc_xc_rks.argtypes = [P, i32, i64, P, P, P, P, d, i32]
        L.oracle_calc_xc_rks.restype = d
        L.oracle_calc_xc_uks.argtypes = [P, i32, i64, P, P, P, P, P, P, d, i32, i32]
        L.oracle_calc_xc_uks.restype = d
        L.oracle_lebedev.argtypes = [i32, P, P]
        L.oracle_lebedev.restype = i32
        L.oracle_radial_em.argtypes = [d, i32, i32, P, P]
        L.oracle_bragg_slater.argtypes = [i32]
        L.oracle_bragg_slater.restype = d
        L.oracle_o_matrix.argtypes = [i32, P, P, P]
        L.oracle_becke_weight.argtypes = [i32, P, P, i32, P, i32]
        L.oracle_becke_weight.restype = d
        L.oracle_generate_grid.argtypes = [i32, P, P, i32, i32, P, d, P, P, i32]
        L.oracle_generate_grid.restype = i64
        L.oracle_radial_gc.argtypes = [d, i32, i32, i32, P, P]
        L.oracle_sg1_radius.argtypes = [i32]
        L.oracle_sg1_radius.restype = d
        L.oracle_lebedev_floor.argtypes = [i32]
        L.oracle_lebedev_floor.restype = i32
        L.oracle_pruned_order.argtypes = [i32, i32, d, d, i32]
        L.oracle_pruned_order.restype = i32
        L.oracle_ss_weight.argtypes = [i32, P, i32, P]
        L.oracle_ss_weight.restype = d
        L.oracle_generate_grid_ex.argtypes = [i32, P, P, P, P, P, P, i32]
        L.oracle_generate_grid_ex.restype = i64
        L.oracle_eval_ao_hessians.argtypes = [P, i64, P, P]
        L.oracle_becke_weight_gradient.argtypes = [i32, P, P, i32, P, P]
        L.oracle_becke_weight_gradient.restype = d
        L.oracle_xc_gradient_rks.argtypes = [P, P, i32, i64, P, P, i32, P, P, P, d, P, P, i32]
        L.oracle_xc_gradient_rks.restype = d
        L.oracle_xc_gradient_rks_ex.argtypes = [P, P, i32, i64, P, P, i32, P, P, P, d, i32, P, P, i32]
        L.oracle_xc_gradient_rks_ex.restype = d
        L.oracle_partition_weight_gradient.argtypes = [i32, P, P, i32, P, i32, P]
        L.oracle_partition_weight_gradient.restype = d
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def functional(fid, inputs):
    """inputs: (n, 5) rhoA, rhoB, gAA, gAB, gBB -> (n, 6) zk, vrhoA, vrhoB, vgAA, vgAB, vgBB."""
    x = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, 5)
    out = np.zeros((x.shape[0], 6))
    lib().oracle_functional(fid, x.shape[0], _p(x), _p(out))
    return out


def normalize_cgto(shell, exps, coefs):
    e = np.ascontiguousarray(exps, dtype=np.float64)
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    oe, oc = np.zeros_like(e), np.zeros_like(c)
    lib().oracle_normalize_cgto(shell, len(e), _p(e), _p(c), _p(oe), _p(oc))
    return oe, oc


class _Aos(ctypes.Structure):
    _fields_ = [("n_aos", ctypes.c_int), ("center", ctypes.c_void_p),
                ("type", ctypes.c_void_p), ("nprim", ctypes.c_void_p),
                ("exps", ctypes.c_void_p), ("coefs", ctypes.c_void_p)]


class AoTable:
    """Per-AO table (centre, basis type 0..8, normalised primitives)."""

    def __init__(self, center, types, nprim, exps, coefs):
        self.center = np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 3)
        self.type = np.ascontiguousarray(types, dtype=np.int32)
        self.nprim = np.ascontiguousarray(nprim, dtype=np.int32)
        n = len(self.type)
        self.exps = np.zeros((n, MAX_PRIM))
        self.coefs = np.zeros((n, MAX_PRIM))
        for i in range(n):
            k = self.nprim[i]
            self.exps[i, :k] = exps[i][:k]
            self.coefs[i, :k] = coefs[i][:k]
        self.n = n
        self._s = _Aos(n, _p(self.center).value, _p(self.type).value,
                       _p(self.nprim).value, _p(self.exps).value, _p(self.coefs).value)

    @classmethod
    def from_basis(cls, basis):
        """basis: object with per-AO arrays center/type/nprim/exps/coefs."""
        return cls(basis.ao_center, basis.ao_type, basis.ao_nprim, basis.ao_exps,
                   basis.ao_coefs)

    def ref(self):
        return ctypes.byref(self._s)


def eval_aos(aos, xyz, grad=True):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    phi = np.zeros((n, aos.n))
    if grad:
        gx, gy, gz = (np.zeros((n, aos.n)) for _ in range(3))
        lib().oracle_eval_aos(aos.ref(), n, _p(xyz), _p(phi), _p(gx), _p(gy), _p(gz))
        return phi, gx, gy, gz
    lib().oracle_eval_aos(aos.ref(), n, _p(xyz), _p(phi), None, None, None)
    return phi


def calc_xc_rks(aos, grid, PA, fid=B3LYP, density_cutoff=1e-16, n_threads=0,
                want_rho=False):
    """Returns (E_xc, F_alpha dense symmetric, rho (n,4) or None)."""
    grid = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1, 4)
    PA = np.ascontiguousarray(PA, dtype=np.float64)
    F = np.zeros_like(PA)
    rho = np.zeros((grid.shape[0], 4)) if want_rho else None
    e = lib().oracle_calc_xc_rks(aos.ref(), fid, grid.shape[0], _p(grid), _p(PA),
                                 _p(F), _p(rho) if want_rho else None,
                                 density_cutoff, n_threads)
    F = np.tril(F)
    F = F + np.tril(F, -1).T
    return e, F, rho


def calc_xc_uks(aos, grid, PA, PB, fid=B3LYP, density_cutoff=1e-16, n_threads=0, want_rho=False,
                exact_gab=False):
    """Returns (E_xc, F_alpha, F_beta dense symmetric, rho (n,8) or None)."""
    grid = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1, 4)
    PA = np.ascontiguousarray(PA, dtype=np.float64)
    PB = np.ascontiguousarray(PB, dtype=np.float64)
    FA, FB = np.zeros_like(PA), np.zeros_like(PB)
    rho = np.zeros((grid.shape[0], 8)) if want_rho else None
    e = lib().oracle_calc_xc_uks(aos.ref(), fid, grid.shape[0], _p(grid), _p(PA), _p(PB), _p(FA), _p(FB),
                                 _p(rho) if want_rho else None, density_cutoff, int(exact_gab), n_threads)
    sym = lambda F: np.tril(F) + np.tril(F, -1).T
    return e, sym(FA), sym(FB), rho


def eval_ao_hessians(aos, xyz):
    """(n, N, 6) second derivatives xx, xy, xz, yy, yz, zz."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    h = np.zeros((xyz.shape[0], aos.n, 6))
    lib().oracle_eval_ao_hessians(aos.ref(), xyz.shape[0], _p(xyz), _p(h))
    return h


def becke_weight_gradient(Z, xyz, iatom, r, partition=0):
    """(omega_A, (n_atoms, 3) grad_B omega_A with the point fixed); partition 0
    Becke with size adjustments, 1 Becke without, 2 Stratmann-Scuseria."""
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros((len(Z), 3))
    om = lib().oracle_partition_weight_gradient(len(Z), _p(Z), _p(xyz), iatom, _p(r), partition, _p(out))
    return om, out


def xc_gradient_rks(aos, ao_atom, grid, pt_atom, Z, xyz, PA, fid=B3LYP, density_cutoff=1e-16,
                    n_threads=0, partition=0):
    """Returns (E_xc, grad_functional (n_atoms, 3), grad_weight (n_atoms, 3)).
    partition: the grid's partition (0 Becke with size adjustments, 1 Becke
    without, 2 Stratmann-Scuseria) for the weight term."""
    grid = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1, 4)
    pt_atom = np.ascontiguousarray(pt_atom, dtype=np.int32)
    ao_atom = np.ascontiguousarray(ao_atom, dtype=np.int32)
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    PA = np.ascontiguousarray(PA, dtype=np.float64)
    gf = np.zeros((len(Z), 3))
    gw = np.zeros((len(Z), 3))
    e = lib().oracle_xc_gradient_rks_ex(aos.ref(), _p(ao_atom), fid, grid.shape[0], _p(grid), _p(pt_atom),
                                        len(Z), _p(Z), _p(xyz), _p(PA), density_cutoff, partition, _p(gf),
                                        _p(gw), n_threads)
    return e, gf, gw


def lebedev(order):
    xyz = np.zeros((order, 3))
    w = np.zeros(order)
    n = lib().oracle_lebedev(order, _p(xyz), _p(w))
    if n != order:
        raise ValueError(f"unsupported Lebedev order {order}")
    return xyz, w


def radial_em(R, Nr, i):
    r, w = ctypes.c_double(), ctypes.c_double()
    lib().oracle_radial_em(R, Nr, i, ctypes.byref(r), ctypes.byref(w))
    return r.value, w.value


def bragg_slater(Z):
    return lib().oracle_bragg_slater(Z)


def o_matrix(Z, xyz):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    O = np.zeros((3, 3))
    lib().oracle_o_matrix(len(Z), _p(Z), _p(xyz), _p(O))
    return O


def becke_weight(Z, xyz, iatom, r, size_adjust=True):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().oracle_becke_weight(len(Z), _p(Z), _p(xyz), iatom, _p(r), int(size_adjust))


class GridOpts(ctypes.Structure):
    """oracle_grid_opts: radial 0 EM / 1 GC; gc_mapping 0 Becke / 1 TA / 2 KK;
    pruning 0 none / 1 MHL / 2 SG-1; partitioning 0 Becke / 1 SS; radii 0
    Bragg-Slater / 1 SG-1."""
    _fields_ = [("n_rad", ctypes.c_int), ("n_ang", ctypes.c_int), ("radial", ctypes.c_int),
                ("gc_mapping", ctypes.c_int), ("pruning", ctypes.c_int), ("partitioning", ctypes.c_int),
                ("size_adjust", ctypes.c_int), ("radii", ctypes.c_int), ("weight_cutoff", ctypes.c_double)]


def grid_opts(n_rad=75, n_ang=302, radial=0, gc_mapping=1, pruning=0, partitioning=0, size_adjust=1,
              radii=0, weight_cutoff=1e-16):
    return GridOpts(n_rad, n_ang, radial, gc_mapping, pruning, partitioning, size_adjust, radii,
                    weight_cutoff)


def radial_gc(R, Nr, i, mapping):
    r, w = ctypes.c_double(), ctypes.c_double()
    lib().oracle_radial_gc(R, Nr, i, mapping, ctypes.byref(r), ctypes.byref(w))
    return r.value, w.value


def sg1_radius(Z):
    return lib().oracle_sg1_radius(Z)


def lebedev_floor(n):
    return lib().oracle_lebedev_floor(n)


def pruned_order(pruning, Z, r, R, n_ang):
    return lib().oracle_pruned_order(pruning, Z, r, R, n_ang)


def ss_weight(xyz, iatom, r):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().oracle_ss_weight(len(xyz), _p(xyz), iatom, _p(r))


def generate_grid_ex(Z, xyz, opts, O=None):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    if O is None:
        O = o_matrix(Z, xyz)
    O = np.ascontiguousarray(O, dtype=np.float64)
    L = lib()
    n = L.oracle_generate_grid_ex(len(Z), _p(Z), _p(xyz), ctypes.byref(opts), _p(O), None, None, 0)
    grid = np.zeros((n, 4))
    atom = np.zeros(n, dtype=np.int32)
    L.oracle_generate_grid_ex(len(Z), _p(Z), _p(xyz), ctypes.byref(opts), _p(O), _p(grid), _p(atom), 0)
    return grid, atom


def generate_grid(Z, xyz, n_rad, n_ang, O=None, weight_cutoff=1e-16):
    Z = np.ascontiguousarray(Z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    if O is None:
        O = o_matrix(Z, xyz)
    O = np.ascontiguousarray(O, dtype=np.float64)
    L = lib()
    n = L.oracle_generate_grid(len(Z), _p(Z), _p(xyz), n_rad, n_ang, _p(O),
                               weight_cutoff, None, None, 0)
    if n < 0:
        raise ValueError("unsupported grid")
    grid = np.zeros((n, 4))
    atom = np.zeros(n, dtype=np.int32)
    L.oracle_generate_grid(len(Z), _p(Z), _p(xyz), n_rad, n_ang, _p(O),
                           weight_cutoff, _p(grid), _p(atom), 0)
    return grid, atom


os.environ.setdefault("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1)))
