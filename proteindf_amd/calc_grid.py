"""Host-side mirror of ProteinDF's grid XC interface over the engine library.

Names, argument meaning and error behaviour follow the reference classes the
engine replaces, so that callers (and tests) read like the reference's own:

* ``DfFunctional_HFS``, ``DfFunctional_SVWN``, ``DfFunctional_Becke88``,
  ``DfFunctional_B88LYP``, ``DfFunctional_B3LYP`` -- getFunctional /
  getDerivativeFunctional for closed-shell points as the grid path calls
  them (DfFunctional.h:104-106, :126-130, :190-192, :221-232), evaluated by
  the engine's functional kernel.
* ``DfGenerateGrid`` -- the molecular grid (DfGenerateGrid.cpp:39-250,
  :730-768), generated on the device.
* ``DfCalcGridX`` -- calcXCIntegForFockAndEnergy(P_A, functional, F_A)
  (DfCalcGridX.cpp:2305-2315): P_A and F_A in packed lower-triangle order,
  F_A accumulated, E_xc returned; getWholeDensity (:2502-2537).
* ``DfXCFunctional`` -- the functional chosen by name as its constructor
  does (DfXCFunctional.cpp:62-101) and buildXcMatrix for RKS
  (DfXCFunctional.cpp:128-152, getFxc DfXCFunctional.h:132-190): P_A = P/2,
  a zeroed F_xc, the functional's build (none for HF), the electron-count
  check (checkGridAccuracy, :227-248).

* ``WorkFiles`` -- the fl_Work file names of the path (DfObject.cpp:357-389,
  :540-562, :657-683): ``~Ppq.rks<n>.mat`` (P in), ``dP.rks<n>.mat`` (update
  mode), ``Fxc.rks<n>.mat`` (out), ``grid.mat``; read and written by the
  library's RLHD / CSFD routines.

Every call goes through the HIP library; nothing here computes on the CPU.
"""
from pathlib import Path

import numpy as np

from . import engine
from .basis import OrbitalInfo, basis_631gd


class DfFunctional_GGA:
    """Closed-shell evaluation of one engine functional (functional_id)."""
    functional_id = None

    def getFunctional(self, rhoA, gammaAA=None):
        return engine.eval_functional(rhoA, _gamma(rhoA, gammaAA), self.functional_id)[:, 0]

    def getDerivativeFunctional(self, rhoA, gammaAA=None):
        """(df/drhoA, df/dgammaAA, df/dgammaAB); the gamma terms are zero for LDA."""
        o = engine.eval_functional(rhoA, _gamma(rhoA, gammaAA), self.functional_id)
        return o[:, 1], o[:, 2], o[:, 3]


def _gamma(rhoA, gammaAA):
    return np.zeros_like(np.asarray(rhoA, dtype=np.float64)) if gammaAA is None else gammaAA


class DfFunctional_HFS(DfFunctional_GGA):
    functional_id = engine.PDFXC_HFS


class DfFunctional_SVWN(DfFunctional_GGA):
    functional_id = engine.PDFXC_SVWN


class DfFunctional_Becke88(DfFunctional_GGA):
    functional_id = engine.PDFXC_HFB


class DfFunctional_B88LYP(DfFunctional_GGA):
    functional_id = engine.PDFXC_BLYP


class DfFunctional_B3LYP(DfFunctional_GGA):
    functional_id = engine.PDFXC_B3LYP


# getFxc's dispatch (DfXCFunctional.h:140-180); HF has no grid functional
FUNCTIONAL_FOR = {engine.PDFXC_HFS: DfFunctional_HFS, engine.PDFXC_SVWN: DfFunctional_SVWN,
                  engine.PDFXC_HFB: DfFunctional_Becke88, engine.PDFXC_BLYP: DfFunctional_B88LYP,
                  engine.PDFXC_B3LYP: DfFunctional_B3LYP}


def orbital_info(mol, basis_for=basis_631gd):
    """TlOrbitalInfo for a molecule, normalised by the engine's own routine."""
    return OrbitalInfo.build(mol, engine.normalize_cgto, basis_for)


class DfGenerateGrid:
    """The molecular grid.  Without opts: grid-type USER, EM radial, no
    pruning, Becke partitioning with size adjustments.  opts: an
    engine.GridOpts (engine.grid_opts_preset, or PdfParam.grid.to_opts()) for
    any grid type the reference's constructor resolves (DfGenerateGrid.cpp:39-170)."""

    def __init__(self, mol, n_rad=75, n_ang=302, weight_cutoff=1e-16, opts=None):
        self.mol, self.n_rad, self.n_ang, self.weight_cutoff = mol, n_rad, n_ang, weight_cutoff
        self.opts = opts

    def generate(self, ctx):
        if self.opts is not None:
            return ctx.generate_grid_ex(self.mol.Z, self.mol.xyz, self.opts)
        return ctx.generate_grid(self.mol.Z, self.mol.xyz, self.n_rad, self.n_ang,
                                 self.weight_cutoff)


class WorkFiles:
    """The files the XC step reads and writes in the work directory (fl_Work).
    run: the run-type suffix (DfObject.cpp:45-46): "rks", "uks_alpha", "uks_beta",
    "roks_closed", "roks_open", "roks_alpha", "roks_beta"."""

    def __init__(self, work_dir="fl_Work"):
        self.dir = Path(work_dir)

    def p_in(self, iteration, run="rks"):          # getPInMatrixPath (DfObject.cpp:677-679)
        return self.dir / f"~Ppq.{run}{iteration}.mat"

    def diff_density(self, iteration, run="rks"):  # getDiffDensityMatrixPath (:665-667)
        return self.dir / f"dP.{run}{iteration}.mat"

    def fxc(self, iteration, run="rks"):           # getFxcMatrixPath (:712-714)
        return self.dir / f"Fxc.{run}{iteration}.mat"

    def grid_matrix(self):            # getGridMatrixPath: "grid.mat", no iteration in the name
        return self.dir / "grid.mat"


class DfCalcGridX:
    """grid: an (n, 4) array, a grid.mat path (its points, atoms and stored rho),
    an engine.GridOpts to generate that grid, or None to generate the USER / EM /
    unpruned / Becke grid of n_rad x n_ang."""

    def __init__(self, mol, grid=None, n_rad=75, n_ang=302, basis_for=basis_631gd,
                 density_cutoff=1e-16, ao_screen=0.0, info=None, precision=engine.PDFXC_FP64):
        self.mol = mol
        self.info = info if info is not None else orbital_info(mol, basis_for)
        self.m_nNumOfAOs = self.info.n_aos
        opts = engine.default_opts(density_cutoff=density_cutoff, ao_screen=ao_screen, precision=precision)
        self.ctx = engine.XcContext(self.info, opts)
        if grid is None:
            DfGenerateGrid(mol, n_rad, n_ang).generate(self.ctx)
        elif isinstance(grid, engine.GridOpts):
            DfGenerateGrid(mol, opts=grid).generate(self.ctx)
        elif isinstance(grid, (str, Path)):
            self.ctx.load_grid_matrix(grid)
        else:
            self.ctx.set_grid(grid)
        self._rho = None

    @classmethod
    def from_pdfparam(cls, param, grid=None, ao_screen=0.0, precision=engine.PDFXC_FP64):
        """DfCalcGridX(pdfParam): the molecule, orbital basis, density cut-off
        (xc-density-threshold) and molecular grid of a pdfparam.PdfParam;
        grid overrides the generated grid as in the constructor."""
        mol = param.molecule
        return cls(mol, grid=param.grid.to_opts() if grid is None else grid, info=param.orbital_info(),
                   density_cutoff=param.density_cutoff, ao_screen=ao_screen, precision=precision)

    @property
    def grid(self):
        return self.ctx.get_grid()

    def calcXCIntegForFockAndEnergy(self, *args, keep_rho=False, update=False):
        """(P_A, functional, F_A) for RKS or (P_A, P_B, functional, F_A, F_B) for
        UKS, as the reference's overloads (DfCalcGridX.h:195-232).  update: the
        densities are changes and rho accumulates on the stored one
        (calcRho_*_part with m_bIsUpdateXC, DfCalcGridX.cpp:1866-1869)."""
        if len(args) == 3:
            P_A, functional, F_A = args
        elif len(args) == 5:
            P_A, P_B, functional, F_A, F_B = args
        else:
            raise TypeError("calcXCIntegForFockAndEnergy(P_A, functional, F_A) or (P_A, P_B, functional, F_A, F_B)")
        if not isinstance(functional, DfFunctional_GGA) or functional.functional_id is None:
            raise TypeError("functional must be one of the DfFunctional_* classes")
        fid = functional.functional_id
        if len(args) == 3:
            build = self.ctx.build_rks_update if update else self.ctx.build_rks
            E, _, rho = build(P_A, F_A, want_rho=keep_rho, functional=fid)
        else:
            build = self.ctx.build_uks_update if update else self.ctx.build_uks
            E, _, _, rho = build(P_A, P_B, F_A, F_B, want_rho=keep_rho, functional=fid)
        if keep_rho:
            self._rho = rho
        return E

    def saveGridMatrix(self, path, functional, uks=False):
        """grid.mat with the rho columns the functional's type uses (DfGenerateGrid.cpp:229-239)."""
        lda = functional.functional_id in (engine.PDFXC_HFS, engine.PDFXC_SVWN)
        self.ctx.save_grid_matrix(path, (1 if lda else 4) * (2 if uks else 1))

    def getWholeDensity(self):
        """Integrated rho_alpha (RKS), or (rho_alpha, rho_beta) after a UKS build
        (DfCalcGridX.cpp:2502-2537)."""
        if self._rho is None:
            raise RuntimeError("no density kept: call calcXCIntegForFockAndEnergy(..., keep_rho=True)")
        w = self.grid[:, 3]
        if self._rho.shape[1] == 8:
            return float((w * self._rho[:, 0]).sum()), float((w * self._rho[:, 4]).sum())
        return float((w * self._rho[:, 0]).sum())


class TlDensField:
    """makeDensFld (src/tools/TlDensField.cpp:35-53): the density of P at the
    given points, through the engine's gridDensity."""

    def __init__(self, calc_grid):
        self.calc_grid = calc_grid

    def makeDensFld(self, P_packed, grids):
        return self.calc_grid.ctx.density_at_points(P_packed, grids)


class DfXCFunctional:
    """xc_functional: the reference's names ("HF", "HFS", "SVWN", "HFB",
    "BLYP" or "B88LYP", "B3LYP"); a trailing "~" (XC fitting) is removed as
    the reference removes it.  An unknown name raises, as getFxc aborts."""

    def __init__(self, calc_grid, n_electrons, xc_functional="B3LYP", method="RKS"):
        name = xc_functional[:-1] if xc_functional.endswith("~") else xc_functional
        if name not in engine.XC_TYPES:
            raise ValueError(f"unknown XC functional: {xc_functional}")
        if method not in ("RKS", "UKS", "ROKS"):
            raise ValueError(f"unknown method: {method}")
        self.m_nXCFunctional = engine.XC_TYPES[name]
        self.method = method
        self.calc_grid = calc_grid
        self.n_electrons = n_electrons
        self.XC_energy_ = 0.0
        self.electrons_on_grid = None

    def isHybridFunctional(self):
        return self.m_nXCFunctional in (engine.PDFXC_HF, engine.PDFXC_B3LYP)

    def getFockExchangeCoefficient(self):
        return {engine.PDFXC_HF: 1.0, engine.PDFXC_B3LYP: 0.2}.get(self.m_nXCFunctional, 0.0)

    def buildXcMatrixFromFiles(self, work, iteration, update_xc=False):
        """DfXCFunctional::buildXcMatrix with its files (DfXCFunctional.cpp:128-222).
        RKS: P_A = P/2 from ~Ppq.rks<n>.mat; UKS: P_A, P_B from ~Ppq.uks_alpha<n>.mat,
        ~Ppq.uks_beta<n>.mat; ROKS: P_A = P_closed/2 + P_open, P_B = P_closed/2.  In
        update mode the dP.* files are read and, after the first iteration, rho
        accumulates.  Fxc.<run><n>.mat are written (rks; uks_alpha / uks_beta;
        roks_alpha / roks_beta) and grid.mat is rewritten with rho.  Returns F_xc
        (RKS) or (F_alpha, F_beta)."""
        n = self.calc_grid.m_nNumOfAOs
        load = lambda run: engine.load_symmetric_matrix(
            work.diff_density(iteration, run) if update_xc else work.p_in(iteration, run))
        accumulate = update_xc and iteration > 1
        if self.method == "RKS":
            Fxc = np.zeros(n * (n + 1) // 2)
            if self.m_nXCFunctional == engine.PDFXC_HF:
                self.XC_energy_ = 0.0
            else:
                functional = FUNCTIONAL_FOR[self.m_nXCFunctional]()
                self.XC_energy_ = self.calc_grid.calcXCIntegForFockAndEnergy(
                    0.5 * load("rks"), functional, Fxc, keep_rho=True, update=accumulate)
                self.electrons_on_grid = 2.0 * self.calc_grid.getWholeDensity()
                self.calc_grid.saveGridMatrix(work.grid_matrix(), functional)
            engine.save_symmetric_matrix(work.fxc(iteration), Fxc)
            return Fxc
        if self.method == "UKS":
            PA, PB = load("uks_alpha"), load("uks_beta")
            runs = ("uks_alpha", "uks_beta")
        else:
            Pc, Po = load("roks_closed"), load("roks_open")
            PA, PB = 0.5 * Pc + Po, 0.5 * Pc
            runs = ("roks_alpha", "roks_beta")
        FA, FB = self.buildXcMatrixUKS(PA, PB, update=accumulate)
        if self.m_nXCFunctional != engine.PDFXC_HF:
            self.calc_grid.saveGridMatrix(work.grid_matrix(), FUNCTIONAL_FOR[self.m_nXCFunctional](), uks=True)
        engine.save_symmetric_matrix(work.fxc(iteration, runs[0]), FA)
        engine.save_symmetric_matrix(work.fxc(iteration, runs[1]), FB)
        return FA, FB

    def buildXcMatrixUKS(self, PA_packed, PB_packed, update=False):
        """getFxc(P_A, P_B, ...) (DfXCFunctional.h:194-240): (F_alpha, F_beta)."""
        n = self.calc_grid.m_nNumOfAOs
        FA, FB = np.zeros(n * (n + 1) // 2), np.zeros(n * (n + 1) // 2)
        if self.m_nXCFunctional == engine.PDFXC_HF:
            self.XC_energy_ = 0.0
            return FA, FB
        self.XC_energy_ = self.calc_grid.calcXCIntegForFockAndEnergy(
            np.asarray(PA_packed, dtype=np.float64), np.asarray(PB_packed, dtype=np.float64),
            FUNCTIONAL_FOR[self.m_nXCFunctional](), FA, FB, keep_rho=True, update=update)
        self.electrons_on_grid = sum(self.calc_grid.getWholeDensity())
        return FA, FB

    def buildXcMatrix(self, P_packed):
        n = self.calc_grid.m_nNumOfAOs
        Fxc = np.zeros(n * (n + 1) // 2)
        if self.m_nXCFunctional == engine.PDFXC_HF:
            self.XC_energy_ = 0.0
            return Fxc
        P_A = 0.5 * np.asarray(P_packed, dtype=np.float64)
        self.XC_energy_ = self.calc_grid.calcXCIntegForFockAndEnergy(
            P_A, FUNCTIONAL_FOR[self.m_nXCFunctional](), Fxc, keep_rho=True)
        self.electrons_on_grid = 2.0 * self.calc_grid.getWholeDensity()
        return Fxc
