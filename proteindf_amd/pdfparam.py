"""Reader for pdfparam.mpac, the parameter file of a ProteinDF run (a msgpack map).

The grid XC path takes from it:

* ``coordinates`` -- {"atoms": [...], "groups": {name: {...}}} as
  Fl_Geometry::setup reads it (Fl_Geometry.cpp:212-240): the groups first,
  each read the same way (recursively), then the atoms of this level; every
  atom has "symbol", "xyz" (bohr), "charge" and "label" (nil reads as "").
* ``basis_set`` -- {"<symbol>[@<label>]": {"name": ..., "cGTOs": [{"shell_type",
  "scale_factor", "pGTOs": [{"exp", "coef"}]}]}} as Fl_Gto::setup reads it
  (Fl_Gto.cpp:474-530).  An atom takes the CGTOs of the entry whose symbol and
  label equal its own, in the file's order (TlOrbitalInfoObject.cpp:181-259);
  dummy atoms "X" have none.  scale_factor is kept but, as in the reference
  (TlOrbitalInfoObject::setCGTO, :38-106), not applied.  An atom with no entry
  raises here, where the reference logs it and goes on without its AOs.
* the method ("RKS"/"NSP", "UKS"/"SP", "ROKS"; DfObject.cpp:103-113), the
  electron counts (method/<m>/...), ``xc_functional`` (upper-cased, a trailing
  "~" marks XC fitting; DfObject.cpp:220-224), ``xc-update`` (anything but "no"
  is on; :231), ``xc-density-threshold`` (DfCalcGridX.cpp:56-59, default 1e-16)
  and the grid settings exactly as DfGenerateGrid's constructor resolves them
  (DfGenerateGrid.cpp:39-170).

Strings are decoded as UTF-8; numbers written as strings are parsed as the
reference's getDouble / getInt / getBoolean parse them (TlSerializeData.cpp:504-534).
The file is read with msgpack, which executes nothing from it.
"""
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .basis import CGTO, OrbitalInfo
from .molecules import Z_OF, Molecule


def _decode(x):
    if isinstance(x, bytes):
        return x.decode("utf-8")
    if isinstance(x, dict):
        return {_decode(k): _decode(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_decode(v) for v in x]
    return x


def _str(v):
    """TlSerializeData::getStr: nil is "", booleans are "true"/"false"."""
    if v is None:
        return ""
    if isinstance(v, bool):
        return "true" if v else "false"
    return str(v)


def _bool(v):
    if isinstance(v, str):
        return v.upper() in ("TRUE", "YES", "ON", "1")
    return bool(v) if v is not None else False


@dataclass
class GridSettings:
    """The molecular grid as DfGenerateGrid's constructor resolves the keys
    (DfGenerateGrid.cpp:39-170).  radial: "EM" (Euler-Maclaurin) or "GC"
    (Gauss-Chebyshev, gc_mapping "TA", "BECKE" or "KK"); partitioning
    "SSWEIGHT" or "BECKE"."""
    grid_type: str = "FINE"
    n_rad: int = 75
    n_ang: int = 302
    radial: str = "EM"
    gc_mapping: str = "TA"
    pruning: bool = True
    partitioning: str = "SSWEIGHT"
    atomic_size_adjustments: bool = True
    weight_cutoff: float = 1e-16

    def to_opts(self):
        """The engine's pdfxc_grid_opts for these settings: SG-1 prunes by its
        regions with its own radii (DfGenerateGrid.cpp:953-957, :1014-1019),
        other types by Murray-Handy-Laming when pruning is on."""
        from . import engine

        sg1 = self.grid_type == "SG-1"
        return engine.GridOpts(
            self.n_rad, self.n_ang,
            engine.PDFXC_RADIAL_GC if self.radial == "GC" else engine.PDFXC_RADIAL_EM,
            {"BECKE": engine.PDFXC_GC_BECKE, "TA": engine.PDFXC_GC_TA, "KK": engine.PDFXC_GC_KK}[self.gc_mapping],
            engine.PDFXC_PRUNE_SG1 if sg1 else (engine.PDFXC_PRUNE_MHL if self.pruning else engine.PDFXC_PRUNE_NONE),
            engine.PDFXC_PARTITION_SS if self.partitioning == "SSWEIGHT" else engine.PDFXC_PARTITION_BECKE,
            1 if self.atomic_size_adjustments else 0,
            engine.PDFXC_RADII_SG1 if sg1 else engine.PDFXC_RADII_BRAGG,
            self.weight_cutoff)


_PRESETS = {"COARSE": (35, 110), "MEDIUM": (32, 146), "MEDIUM-FINE": (64, 146), "MEDIUM_FINE": (64, 146),
            "FINE": (75, 302), "ULTRA-FINE": (99, 590), "ULTRA_FINE": (99, 590)}


def grid_settings(param):
    g = GridSettings()
    get = lambda k: _str(param.get(k))
    if get("grid/weight_cutoff") != "":
        g.weight_cutoff = float(get("grid/weight_cutoff"))
    t = get("xc-potential/grid-type").upper()
    if t in _PRESETS:
        g.grid_type = t.replace("_", "-")
        g.n_rad, g.n_ang = _PRESETS[t]
    elif t == "USER":
        g.grid_type = "USER"
        if get("grid/num_of_radial_shells") != "":
            g.n_rad = int(get("grid/num_of_radial_shells"))
        if get("grid/num_of_angular_points") != "":
            g.n_ang = int(get("grid/num_of_angular_points"))
        m = get("grid/radial_quadorature_method").upper()
        if m in ("GC", "GAUSS-CHEBYSHEV"):
            g.radial = "GC"
        if g.radial == "GC" and get("grid/GC_mapping_type").upper() in ("BECKE", "TA", "KK"):
            g.gc_mapping = get("grid/GC_mapping_type").upper()
        if get("grid/pruning") != "":
            g.pruning = _bool(param.get("grid/pruning"))
    elif t in ("SG-1", "SG1"):
        g.grid_type = "SG-1"
        g.n_rad, g.n_ang = 50, 194
        g.atomic_size_adjustments = False
    else:
        raise ValueError(f"unsupported grid type: {get('xc-potential/grid-type')}")
    pm = get("grid/partitioning_method").upper()
    if pm in ("BECKE", "SSWEIGHT"):
        g.partitioning = pm
    if get("grid/atomic_size_adjustments") != "":
        g.atomic_size_adjustments = _bool(param.get("grid/atomic_size_adjustments"))
    return g


@dataclass
class BasisEntry:
    name: str
    cgtos: list                      # basis.CGTO, in the file's order
    scale_factors: list = field(default_factory=list)


class PdfParam:
    """A loaded pdfparam.mpac: the raw map (``raw``, strings decoded) and the
    pieces the grid XC step uses."""

    def __init__(self, raw):
        self.raw = raw
        self.symbols, self.xyz, self.labels, self.charges = [], [], [], []
        self._read_geometry(raw.get("coordinates") or {})
        self.basis = {}
        for key, entry in (raw.get("basis_set") or {}).items():
            cg, sf = [], []
            for c in entry.get("cGTOs") or []:
                shell = _str(c.get("shell_type"))[:1]
                pg = c.get("pGTOs") or []
                cg.append(CGTO(shell, np.array([float(p["exp"]) for p in pg]),
                               np.array([float(p["coef"]) for p in pg])))
                sf.append(float(c.get("scale_factor") or 0.0))
            self.basis[key] = BasisEntry(_str(entry.get("name")), cg, sf)

    def _read_geometry(self, geom):
        for sub in (geom.get("groups") or {}).values():
            self._read_geometry(sub)
        for a in geom.get("atoms") or []:
            s = _str(a.get("symbol"))
            if s not in Z_OF:
                raise ValueError(f"unsupported element '{s}'")
            self.symbols.append(s)
            self.xyz.append([float(v) for v in a["xyz"][:3]])
            self.labels.append(_str(a.get("label")))
            self.charges.append(float(a.get("charge") or 0.0))

    def get(self, key, default=None):
        v = self.raw.get(key)
        return default if v is None else v

    @property
    def molecule(self):
        return Molecule(list(self.symbols), np.array(self.xyz, dtype=np.float64).reshape(-1, 3),
                        _str(self.raw.get("comment")))

    @property
    def method(self):
        m = _str(self.raw.get("method")).upper()
        return {"NSP": "RKS", "SP": "UKS"}.get(m, m if m in ("RKS", "UKS", "ROKS") else "RKS")

    @property
    def xc_functional(self):
        return _str(self.raw.get("xc_functional")).upper()

    @property
    def is_xc_fitting(self):
        return self.xc_functional.endswith("~")

    @property
    def xc_update(self):
        return _str(self.raw.get("xc-update")).upper() != "NO"

    @property
    def density_cutoff(self):
        v = _str(self.raw.get("xc-density-threshold"))
        return float(v) if v != "" else 1e-16

    def electrons(self):
        """RKS: the electron count; UKS: (alpha, beta); ROKS: (closed, open)."""
        num = lambda k: int(_str(self.raw.get(k)) or 0)
        if self.method == "UKS":
            return num("method/uks/alpha_electrons"), num("method/uks/beta_electrons")
        if self.method == "ROKS":
            return num("method/roks/closed_electrons"), num("method/roks/open_electrons")
        return num("method/rks/electrons")

    @property
    def grid(self):
        return grid_settings(self.raw)

    def basis_for(self, symbol, label):
        """The CGTOs of every basis_set entry whose key splits at its first '@'
        into (symbol, label) (Fl_Gto.cpp:486-492), in the file's order."""
        out = []
        for key, entry in self.basis.items():
            atom, _, lab = key.partition("@")
            if (atom, lab) == (symbol, label):
                out.extend(entry.cgtos)
        if not out:
            raise KeyError(f'the basis is not found: symbol="{symbol}", label="{label}"')
        return out

    def orbital_info(self, normalizer=None):
        """TlOrbitalInfo(coordinates, basis_set), normalised by the engine library
        unless a normaliser is given."""
        if normalizer is None:
            from . import engine
            normalizer = engine.normalize_cgto
        return OrbitalInfo.build(self.molecule, normalizer, self.basis_for, labels=self.labels)


def load(path):
    """Read pdfparam.mpac (msgpack; nothing in the file is executed)."""
    import msgpack

    raw = msgpack.unpackb(Path(path).read_bytes(), raw=True, strict_map_key=False)
    return PdfParam(_decode(raw))
