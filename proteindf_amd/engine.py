"""ctypes binding of the engine library (include/pdfxc.h).

The library is built in-tree by ``python -m proteindf_amd.build`` (or
``__graft_entry__.build()``) into ``proteindf_amd/_lib/libpdfxc.so``.  There
is no fallback: if the library is missing or fails to load, every entry point
raises.
"""
import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libpdfxc.so"

PDFXC_OK = 0
# functionals: the XC_TYPE values of DfXCFunctional.h:44
PDFXC_HF = 1
PDFXC_HFS = 2
PDFXC_SVWN = 3
PDFXC_HFB = 4
PDFXC_BLYP = 5
PDFXC_B3LYP = 6
XC_TYPES = {"HF": PDFXC_HF, "HFS": PDFXC_HFS, "SVWN": PDFXC_SVWN, "HFB": PDFXC_HFB,
            "BLYP": PDFXC_BLYP, "B88LYP": PDFXC_BLYP, "B3LYP": PDFXC_B3LYP}
PDFXC_FP64 = 0
PDFXC_SPLIT_BF16 = 1
PDFXC_SPLIT_BF16X3 = 2
PDFXC_SPLIT_K_BF16X3 = 3   # FP64 rho pass, hi + mid + lo K pass
PDFXC_SPLIT_BFP = 4        # block floating point rho pass (4 slices, exact sums), hi + mid + lo K pass
PDFXC_SPLIT_K_F16 = 5      # FP64 rho pass, fp16 hi + lo K pass (per-slab column scales)
PRECISIONS = {"fp64": PDFXC_FP64, "bf16x2": PDFXC_SPLIT_BF16, "bf16x3": PDFXC_SPLIT_BF16X3,
              "bf16x3k": PDFXC_SPLIT_K_BF16X3, "bfp": PDFXC_SPLIT_BFP, "f16k": PDFXC_SPLIT_K_F16}
ERROR_NAMES = {1: "ARGUMENT", 2: "HIP", 3: "NO_GRID", 4: "UNSUPPORTED", 5: "MEMORY"}
ERROR_CODES = {v: k for k, v in ERROR_NAMES.items()}

# every symbol declared in include/pdfxc.h
EXPORTED = [
    "pdfxc_default_opts", "pdfxc_normalize_cgto", "pdfxc_ctx_create", "pdfxc_ctx_destroy",
    "pdfxc_set_grid", "pdfxc_generate_grid", "pdfxc_get_grid", "pdfxc_build_rks",
    "pdfxc_build_rks_device", "pdfxc_eval_functional", "pdfxc_eval_aos", "pdfxc_get_stats",
    "pdfxc_stream", "pdfxc_last_error", "pdfxc_gradient_rks", "pdfxc_build_rks_update",
    "pdfxc_build_rks_update_device", "pdfxc_selftest_mfma", "pdfxc_selftest_mfma_split",
    "pdfxc_selftest_copy", "pdfxc_matrix_file_info", "pdfxc_load_symmetric_matrix",
    "pdfxc_save_symmetric_matrix", "pdfxc_load_general_matrix", "pdfxc_save_general_matrix",
    "pdfxc_load_grid_matrix", "pdfxc_save_grid_matrix", "pdfxc_build_uks", "pdfxc_build_uks_update",
    "pdfxc_build_uks_device", "pdfxc_build_uks_update_device", "pdfxc_eval_functional_uks",
    "pdfxc_density_at_points", "pdfxc_batch_create", "pdfxc_batch_destroy", "pdfxc_batch_offsets",
    "pdfxc_batch_build_rks_device", "pdfxc_batch_build_rks", "pdfxc_batch_stats", "pdfxc_set_tuning",
    "pdfxc_autotune", "pdfxc_grid_opts_preset", "pdfxc_generate_grid_ex", "pdfxc_generate_grid_oriented",
    "pdfxc_eval_functional_noalloc",
]
# matrix file types (TlMatrixObject::MatrixType)
PDFXC_MAT_RSFD, PDFXC_MAT_CSFD, PDFXC_MAT_RLHD = 0, 1, 2


class PdfxcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdfxc error {code} ({ERROR_NAMES.get(code, '?')}): {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("density_cutoff", ctypes.c_double), ("ao_screen", ctypes.c_double),
                ("precision", ctypes.c_int), ("block_points", ctypes.c_int),
                ("scratch_bytes", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("rho_shift", ctypes.c_int), ("k_tiles", ctypes.c_int), ("split_rho_shift", ctypes.c_int),
                ("split_k_patches", ctypes.c_int), ("split_k_group", ctypes.c_int), ("persistent", ctypes.c_int),
                ("ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# molecular grid options (pdfxc_grid_opts)
PDFXC_RADIAL_EM, PDFXC_RADIAL_GC = 0, 1
PDFXC_GC_BECKE, PDFXC_GC_TA, PDFXC_GC_KK = 0, 1, 2
PDFXC_PRUNE_NONE, PDFXC_PRUNE_MHL, PDFXC_PRUNE_SG1 = 0, 1, 2
PDFXC_PARTITION_BECKE, PDFXC_PARTITION_SS = 0, 1
PDFXC_RADII_BRAGG, PDFXC_RADII_SG1 = 0, 1


class GridOpts(ctypes.Structure):
    _fields_ = [("n_rad", ctypes.c_int), ("n_ang", ctypes.c_int), ("radial", ctypes.c_int),
                ("gc_mapping", ctypes.c_int), ("pruning", ctypes.c_int), ("partitioning", ctypes.c_int),
                ("size_adjust", ctypes.c_int), ("radii", ctypes.c_int), ("weight_cutoff", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def grid_opts_preset(grid_type="sg-1", **overrides):
    """pdfxc_grid_opts_preset for a grid type, then field overrides."""
    o = GridOpts()
    check(lib().pdfxc_grid_opts_preset(str(grid_type).encode(), ctypes.byref(o)))
    for k, v in overrides.items():
        if k not in o.as_dict():
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Stats(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("n_blocks", ctypes.c_int64),
                ("n_chunks", ctypes.c_int64), ("flops", ctypes.c_double),
                ("mean_nao", ctypes.c_double), ("max_nao", ctypes.c_int),
                ("last_build_ms", ctypes.c_double), ("last_kernel_ms", ctypes.c_double * 4),
                ("n_k_fp64_blocks", ctypes.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"engine library not built: {LIB_PATH} (run python -m proteindf_amd.build)")
        L = ctypes.CDLL(str(LIB_PATH))
        P, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.pdfxc_default_opts.argtypes = [P]
        L.pdfxc_default_opts.restype = None
        L.pdfxc_normalize_cgto.argtypes = [i32, i32, P, P, P, P]
        L.pdfxc_ctx_create.argtypes = [P, i32, P, P, P, P, P, i32, P]
        L.pdfxc_ctx_destroy.argtypes = [P]
        L.pdfxc_set_grid.argtypes = [P, i64, P, P, P, P]
        L.pdfxc_generate_grid.argtypes = [P, i32, P, P, i32, i32, d, P]
        L.pdfxc_grid_opts_preset.argtypes = [ctypes.c_char_p, P]
        L.pdfxc_generate_grid_ex.argtypes = [P, i32, P, P, P, P]
        L.pdfxc_generate_grid_oriented.argtypes = [P, i32, P, P, P, P, P]
        L.pdfxc_get_grid.argtypes = [P, P, P, P, P]
        L.pdfxc_build_rks.argtypes = [P, i32, P, P, P, P]
        L.pdfxc_build_rks_device.argtypes = [P, i32, P, P, P]
        L.pdfxc_eval_functional.argtypes = [i32, i64, P, P]
        L.pdfxc_eval_functional_noalloc.argtypes = [i32, i64, P, P]
        L.pdfxc_eval_aos.argtypes = [P, i64, P, P, P, P, P]
        L.pdfxc_get_stats.argtypes = [P, P]
        L.pdfxc_stream.argtypes = [P]
        L.pdfxc_stream.restype = P
        L.pdfxc_last_error.argtypes = []
        L.pdfxc_last_error.restype = ctypes.c_char_p
        L.pdfxc_selftest_mfma.argtypes = [P]
        L.pdfxc_gradient_rks.argtypes = [P, i32, i32, P, P, P, P, P, P, P]
        L.pdfxc_build_rks_update.argtypes = [P, i32, P, P, P, P]
        L.pdfxc_build_rks_update_device.argtypes = [P, i32, P, P, P]
        L.pdfxc_selftest_mfma_split.argtypes = [P]
        L.pdfxc_selftest_copy.argtypes = [P, P, i64]
        c = ctypes.c_char_p
        L.pdfxc_matrix_file_info.argtypes = [c, P, P, P]
        L.pdfxc_load_symmetric_matrix.argtypes = [c, i64, P]
        L.pdfxc_save_symmetric_matrix.argtypes = [c, i64, P]
        L.pdfxc_load_general_matrix.argtypes = [c, i64, i64, P]
        L.pdfxc_save_general_matrix.argtypes = [c, i64, i64, P]
        L.pdfxc_load_grid_matrix.argtypes = [P, c, P, P]
        L.pdfxc_save_grid_matrix.argtypes = [P, c, i32]
        L.pdfxc_build_uks.argtypes = [P, i32, P, P, P, P, P, P]
        L.pdfxc_build_uks_update.argtypes = [P, i32, P, P, P, P, P, P]
        L.pdfxc_build_uks_device.argtypes = [P, i32, P, P, P, P, P]
        L.pdfxc_build_uks_update_device.argtypes = [P, i32, P, P, P, P, P]
        L.pdfxc_eval_functional_uks.argtypes = [i32, i64, P, P]
        L.pdfxc_density_at_points.argtypes = [P, P, i64, P, P, P]
        L.pdfxc_batch_create.argtypes = [P, i32, P]
        L.pdfxc_batch_destroy.argtypes = [P]
        L.pdfxc_batch_offsets.argtypes = [P, P, P]
        L.pdfxc_batch_build_rks_device.argtypes = [P, i32, P, P, P]
        L.pdfxc_batch_build_rks.argtypes = [P, i32, P, P, P]
        L.pdfxc_batch_stats.argtypes = [P, P]
        L.pdfxc_set_tuning.argtypes = [P, P]
        L.pdfxc_autotune.argtypes = [P, i32, P, P, P, i32, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(code):
    if code != PDFXC_OK:
        raise PdfxcError(code, lib().pdfxc_last_error().decode())


def default_opts(**kw):
    o = Opts()
    lib().pdfxc_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def normalize_cgto(shell, exps, coefs):
    e = np.ascontiguousarray(exps, dtype=np.float64)
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    oe, oc = np.zeros_like(e), np.zeros_like(c)
    check(lib().pdfxc_normalize_cgto(shell, len(e), _p(e), _p(c), _p(oe), _p(oc)))
    return oe, oc


def eval_functional(rho_a, gamma_aa, functional=PDFXC_B3LYP):
    """Closed-shell functional kernel: returns (n, 4) f, df/drhoA, df/dgAA, df/dgAB."""
    x = np.ascontiguousarray(np.stack([np.ravel(rho_a), np.ravel(gamma_aa)], 1), dtype=np.float64)
    out = np.zeros((x.shape[0], 4))
    check(lib().pdfxc_eval_functional(functional, x.shape[0], _p(x), _p(out)))
    return out


NOALLOC_MAX_POINTS = 65536  # PDFXC_NOALLOC_MAX_POINTS


def eval_functional_noalloc(rho_a, gamma_aa, functional=PDFXC_B3LYP):
    """eval_functional through the library's fixed device staging (no device
    allocation; at most NOALLOC_MAX_POINTS points per call)."""
    x = np.ascontiguousarray(np.stack([np.ravel(rho_a), np.ravel(gamma_aa)], 1), dtype=np.float64)
    out = np.zeros((x.shape[0], 4))
    check(lib().pdfxc_eval_functional_noalloc(functional, x.shape[0], _p(x), _p(out)))
    return out


def eval_functional_uks(inputs, functional=PDFXC_B3LYP):
    """Spin-resolved functional kernel: inputs (n, 5) rhoA, rhoB, gAA, gAB, gBB ->
    (n, 6) f, df/drhoA, df/drhoB, df/dgAA, df/dgAB, df/dgBB."""
    x = np.ascontiguousarray(np.asarray(inputs, dtype=np.float64).reshape(-1, 5))
    out = np.zeros((x.shape[0], 6))
    check(lib().pdfxc_eval_functional_uks(functional, x.shape[0], _p(x), _p(out)))
    return out


def _path(path):
    return str(path).encode()


def matrix_file_info(path):
    """(type, rows, cols) of a ProteinDF dense matrix file."""
    t, r, c = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().pdfxc_matrix_file_info(_path(path), ctypes.byref(t), ctypes.byref(r), ctypes.byref(c)))
    return t.value, r.value, c.value


def load_symmetric_matrix(path):
    """RLHD file -> packed lower triangle by rows (N(N+1)/2 doubles)."""
    _, n, _ = matrix_file_info(path)
    out = np.zeros(n * (n + 1) // 2)
    check(lib().pdfxc_load_symmetric_matrix(_path(path), n, _p(out)))
    return out


def save_symmetric_matrix(path, packed):
    a = np.ascontiguousarray(packed, dtype=np.float64)
    n = int(round((np.sqrt(8 * a.size + 1) - 1) / 2))
    if n * (n + 1) // 2 != a.size:
        raise ValueError("packed symmetric matrix must hold N(N+1)/2 values")
    check(lib().pdfxc_save_symmetric_matrix(_path(path), n, _p(a)))


def load_general_matrix(path):
    """RSFD or CSFD file -> (rows, cols) array."""
    _, r, c = matrix_file_info(path)
    out = np.zeros((c, r))  # column-major buffer
    check(lib().pdfxc_load_general_matrix(_path(path), r, c, _p(out)))
    return out.T


def save_general_matrix(path, m):
    m = np.asarray(m, dtype=np.float64)
    cm = np.ascontiguousarray(m.T)
    check(lib().pdfxc_save_general_matrix(_path(path), m.shape[0], m.shape[1], _p(cm)))


def selftest_mfma():
    out = np.zeros((16, 16))
    check(lib().pdfxc_selftest_mfma(_p(out)))
    return out


def selftest_mfma_split():
    """(4, 16, 16): the 16x16x32 and 16x16x16 bf16 products, then the same two fp16 products."""
    out = np.zeros((4, 16, 16), dtype=np.float32)
    check(lib().pdfxc_selftest_mfma_split(_p(out)))
    return out


class XcContext:
    """One engine context: AO table + grid + device buffers (pdfxc_ctx)."""

    def __init__(self, info, opts=None, max_prim=None):
        n = info.n_aos
        if max_prim is None:
            max_prim = int(max(info.ao_nprim))
        self.n_aos = n
        E, C = info.padded(max_prim)
        self._center = np.ascontiguousarray(info.ao_center, dtype=np.float64)
        self._type = np.ascontiguousarray(info.ao_type, dtype=np.int32)
        self._nprim = np.ascontiguousarray(info.ao_nprim, dtype=np.int32)
        self._exps = np.ascontiguousarray(E)
        self._coefs = np.ascontiguousarray(C)
        self.opts = opts if opts is not None else default_opts()
        h = ctypes.c_void_p()
        check(lib().pdfxc_ctx_create(ctypes.byref(h), n, _p(self._center), _p(self._type),
                                     _p(self._nprim), _p(self._exps), _p(self._coefs), max_prim,
                                     ctypes.byref(self.opts)))
        self.h = h
        self.n_pts = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pdfxc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, grid):
        g = np.asarray(grid, dtype=np.float64).reshape(-1, 4)
        cols = [np.ascontiguousarray(g[:, k]) for k in range(4)]
        check(lib().pdfxc_set_grid(self.h, g.shape[0], *[_p(c) for c in cols]))
        self.n_pts = g.shape[0]

    def generate_grid(self, Z, xyz, n_rad, n_ang, weight_cutoff=1e-16):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = ctypes.c_int64()
        check(lib().pdfxc_generate_grid(self.h, len(Z), _p(Z), _p(xyz), n_rad, n_ang,
                                        weight_cutoff, ctypes.byref(n)))
        self.n_pts = n.value
        return n.value

    def generate_grid_ex(self, Z, xyz, opts):
        """The molecular grid for a GridOpts (see grid_opts_preset)."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = ctypes.c_int64()
        check(lib().pdfxc_generate_grid_ex(self.h, len(Z), _p(Z), _p(xyz), ctypes.byref(opts),
                                           ctypes.byref(n)))
        self.n_pts = n.value
        return n.value

    def generate_grid_oriented(self, Z, xyz, opts, O):
        """generate_grid_ex with the angular grids in the fixed orientation O (3 x 3)."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        O = np.ascontiguousarray(O, dtype=np.float64).reshape(3, 3)
        n = ctypes.c_int64()
        check(lib().pdfxc_generate_grid_oriented(self.h, len(Z), _p(Z), _p(xyz), ctypes.byref(opts), _p(O),
                                                 ctypes.byref(n)))
        self.n_pts = n.value
        return n.value

    def load_grid_matrix(self, path):
        """grid.mat -> the context's grid (and point atoms, and stored rho); returns its column count."""
        n, nc = ctypes.c_int64(), ctypes.c_int()
        check(lib().pdfxc_load_grid_matrix(self.h, _path(path), ctypes.byref(n), ctypes.byref(nc)))
        self.n_pts = n.value
        return nc.value

    def save_grid_matrix(self, path, n_rho_cols=4):
        check(lib().pdfxc_save_grid_matrix(self.h, _path(path), n_rho_cols))

    def set_tuning(self, rho_shift=0, k_tiles=96, split_rho_shift=0, split_k_patches=24, split_k_group=4,
                   persistent=2):
        t = Tuning(rho_shift, k_tiles, split_rho_shift, split_k_patches, split_k_group, persistent, 0.0)
        check(lib().pdfxc_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, dP_ptr, dK_ptr, dE_ptr, reps=3, functional=PDFXC_B3LYP):
        """Time the candidate launch shapes on device buffers, keep the fastest; returns it."""
        t = Tuning()
        check(lib().pdfxc_autotune(self.h, functional, ctypes.c_void_p(dP_ptr), ctypes.c_void_p(dK_ptr),
                                   ctypes.c_void_p(dE_ptr), reps, ctypes.byref(t)))
        return t

    def density_at_points(self, P_packed, xyz, want_grad=False):
        """rho = phi^T P phi at arbitrary points (gridDensity); returns rho (n,) or (rho, grad (n, 3))."""
        P = np.ascontiguousarray(P_packed, dtype=np.float64)
        if P.shape != (self.n_aos * (self.n_aos + 1) // 2,):
            raise ValueError("P_packed must hold N(N+1)/2 values")
        x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        rho = np.zeros(x.shape[0])
        grad = np.zeros((x.shape[0], 3)) if want_grad else None
        check(lib().pdfxc_density_at_points(self.h, _p(P), x.shape[0], _p(x), _p(rho), _p(grad)))
        return (rho, grad) if want_grad else rho

    def get_grid(self):
        cols = [np.zeros(self.n_pts) for _ in range(4)]
        check(lib().pdfxc_get_grid(self.h, *[_p(c) for c in cols]))
        return np.stack(cols, 1)

    def build_rks(self, P_alpha_packed, K_alpha_packed=None, want_rho=False,
                  functional=PDFXC_B3LYP):
        P = np.ascontiguousarray(P_alpha_packed, dtype=np.float64)
        n = self.n_aos
        if P.shape != (n * (n + 1) // 2,):
            raise ValueError("P_alpha_packed must hold N(N+1)/2 values")
        K = np.zeros_like(P) if K_alpha_packed is None else K_alpha_packed
        if K.dtype != np.float64 or not K.flags.c_contiguous or K.shape != P.shape:
            raise ValueError("K_alpha_packed must be a contiguous float64 array of N(N+1)/2 values")
        E = ctypes.c_double()
        rho = np.zeros((self.n_pts, 4)) if want_rho else None
        check(lib().pdfxc_build_rks(self.h, functional, _p(P), _p(K), ctypes.byref(E), _p(rho)))
        return E.value, K, rho

    def gradient_rks(self, Z, xyz, P_alpha_packed, point_atom=None, functional=PDFXC_B3LYP):
        """(E_xc, grad_functional (n_atoms, 3), grad_weight (n_atoms, 3)); dE/dR."""
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        P = np.ascontiguousarray(P_alpha_packed, dtype=np.float64)
        pa = None if point_atom is None else np.ascontiguousarray(point_atom, dtype=np.int32)
        gf = np.zeros((len(Z), 3))
        gw = np.zeros((len(Z), 3))
        E = ctypes.c_double()
        check(lib().pdfxc_gradient_rks(self.h, functional, len(Z), _p(Z), _p(xyz), _p(pa), _p(P), _p(gf),
                                       _p(gw), ctypes.byref(E)))
        return E.value, gf, gw

    def build_rks_update(self, dP_alpha_packed, K_alpha_packed=None, want_rho=False, functional=PDFXC_B3LYP):
        """Update mode: adds rho[dP_alpha] to the stored rho, then K and E from it."""
        P = np.ascontiguousarray(dP_alpha_packed, dtype=np.float64)
        n = self.n_aos
        if P.shape != (n * (n + 1) // 2,):
            raise ValueError("dP_alpha_packed must hold N(N+1)/2 values")
        K = np.zeros_like(P) if K_alpha_packed is None else K_alpha_packed
        E = ctypes.c_double()
        rho = np.zeros((self.n_pts, 4)) if want_rho else None
        check(lib().pdfxc_build_rks_update(self.h, functional, _p(P), _p(K), ctypes.byref(E), _p(rho)))
        return E.value, K, rho

    def _uks(self, fn, PA, PB, KA, KB, want_rho, functional):
        n = self.n_aos
        PA = np.ascontiguousarray(PA, dtype=np.float64)
        PB = np.ascontiguousarray(PB, dtype=np.float64)
        if PA.shape != (n * (n + 1) // 2,) or PB.shape != PA.shape:
            raise ValueError("PA and PB must hold N(N+1)/2 values")
        KA = np.zeros_like(PA) if KA is None else KA
        KB = np.zeros_like(PB) if KB is None else KB
        for K in (KA, KB):
            if K.dtype != np.float64 or not K.flags.c_contiguous or K.shape != PA.shape:
                raise ValueError("K must be a contiguous float64 array of N(N+1)/2 values")
        E = ctypes.c_double()
        rho = np.zeros((self.n_pts, 8)) if want_rho else None
        check(fn(self.h, functional, _p(PA), _p(PB), _p(KA), _p(KB), ctypes.byref(E), _p(rho)))
        return E.value, KA, KB, rho

    def build_uks(self, PA_packed, PB_packed, KA_packed=None, KB_packed=None, want_rho=False,
                  functional=PDFXC_B3LYP):
        """UKS build: returns (E_xc, K_alpha, K_beta, rho (n, 8) or None)."""
        return self._uks(lib().pdfxc_build_uks, PA_packed, PB_packed, KA_packed, KB_packed, want_rho, functional)

    def build_uks_update(self, dPA_packed, dPB_packed, KA_packed=None, KB_packed=None, want_rho=False,
                         functional=PDFXC_B3LYP):
        return self._uks(lib().pdfxc_build_uks_update, dPA_packed, dPB_packed, KA_packed, KB_packed, want_rho,
                         functional)

    def build_rks_device(self, dP_ptr, dK_ptr, dE_ptr, functional=PDFXC_B3LYP):
        check(lib().pdfxc_build_rks_device(self.h, functional, ctypes.c_void_p(dP_ptr),
                                           ctypes.c_void_p(dK_ptr), ctypes.c_void_p(dE_ptr)))

    def eval_aos(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        outs = [np.zeros((n, self.n_aos)) for _ in range(4)]
        check(lib().pdfxc_eval_aos(self.h, n, _p(xyz), *[_p(o) for o in outs]))
        return tuple(outs)

    def stats(self):
        s = Stats()
        check(lib().pdfxc_get_stats(self.h, ctypes.byref(s)))
        return s

    def stream(self):
        return lib().pdfxc_stream(self.h)


class XcBatch:
    """Several contexts' grids built together (pdfxc_batch_*): one set of
    launches, captured in a HIP graph and replayed."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = ctypes.c_void_p()
        check(lib().pdfxc_batch_create(ctypes.byref(h), len(self.contexts), arr))
        self.h = h
        n = len(self.contexts)
        self.dense_off = np.zeros(n + 1, dtype=np.int64)
        self.packed_off = np.zeros(n + 1, dtype=np.int64)
        check(lib().pdfxc_batch_offsets(self.h, _p(self.dense_off), _p(self.packed_off)))

    def close(self):
        if self.h:
            lib().pdfxc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_rks(self, P_alpha_packed_list, functional=PDFXC_B3LYP):
        """Returns (E per molecule, list of packed K)."""
        P = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64) for p in P_alpha_packed_list]))
        if P.shape != (self.packed_off[-1],):
            raise ValueError("P list does not match the molecules' sizes")
        K = np.zeros_like(P)
        E = np.zeros(len(self.contexts))
        check(lib().pdfxc_batch_build_rks(self.h, functional, _p(P), _p(K), _p(E)))
        return E, [K[self.packed_off[i]:self.packed_off[i + 1]] for i in range(len(self.contexts))]

    def build_rks_device(self, dP_ptr, dK_ptr, dE_ptr, functional=PDFXC_B3LYP):
        check(lib().pdfxc_batch_build_rks_device(self.h, functional, ctypes.c_void_p(dP_ptr),
                                                 ctypes.c_void_p(dK_ptr), ctypes.c_void_p(dE_ptr)))

    def stats(self):
        s = Stats()
        check(lib().pdfxc_batch_stats(self.h, ctypes.byref(s)))
        return s
