/*
 * CPU oracle for the grid exchange-correlation build (test infrastructure).
 *
 * This is a plain-C restatement of ProteinDF's CPU path, written to check the
 * HIP engine against.  It is NOT part of the product: only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it.  Every
 * function cites the reference lines it follows (paths are relative to the
 * ProteinDF source tree).
 *
 * Parity status of each part (what pins it):
 *   functionals      reference KATs src/unit_test/DfFunctional_*Test.cpp
 *                    (74 active cases) + central finite differences
 *   AO normalisation <phi|phi> = 1 by quadrature, brute-force primitive norms
 *   AO evaluation    closed forms at chosen points, central FD of phi
 *   rho / K / E      brute force on tiny inputs, <D,K> = dE/de central FD,
 *                    integral of rho = sum_i P_ii for diagonal P
 *   grid             Lebedev exactness on spherical harmonics, EM radial
 *                    closed form, Becke partition of unity
 *   O matrix signs   parity unpinned (LAPACK dsyev sign convention is
 *                    implementation-defined; see oracle_o_matrix)
 */
#ifndef PDFXC_ORACLE_H
#define PDFXC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_PRIM 16

/* functional ids */
enum {
    ORACLE_SLATER = 0,
    ORACLE_BECKE88 = 1,
    ORACLE_BECKE88_EXCEPT_LDA = 2,
    ORACLE_VWN5 = 3,
    ORACLE_VWN3 = 4,
    ORACLE_LYP = 5,
    ORACLE_B3LYP = 6,
    ORACLE_HFS = 7,  /* Slater (DfFunctional_HFS.cpp) */
    ORACLE_SVWN = 8, /* Slater + VWN5 (DfFunctional_SVWN.cpp) */
    ORACLE_BLYP = 9  /* Becke88 + LYP (DfFunctional_B88LYP.cpp) */
};

/* in: n x 5 (rhoA, rhoB, gAA, gAB, gBB); out: n x 6
 * (zk, vrhoA, vrhoB, vgAA, vgAB, vgBB), spin-resolved entry points. */
void oracle_functional(int id, int64_t n, const double* in, double* out);

/* Contracted-GTO normalisation (Fl_Gto.cpp:558-708, TlOrbitalInfoObject.cpp:38-106).
 * shell: 0 s, 1 p, 2 d.  Outputs are sorted by ascending exponent. */
void oracle_normalize_cgto(int shell, int nprim, const double* exps,
                           const double* coefs, double* out_exps,
                           double* out_coefs);

/* AO table: per AO centre (3), basis type 0..8, nprim, exps/coefs
 * (n_aos x ORACLE_MAX_PRIM, normalised and sorted). */
typedef struct {
    int n_aos;
    const double* center;  /* n_aos x 3 */
    const int* type;       /* n_aos */
    const int* nprim;      /* n_aos */
    const double* exps;    /* n_aos x ORACLE_MAX_PRIM */
    const double* coefs;   /* n_aos x ORACLE_MAX_PRIM */
} oracle_aos;

/* phi and grad phi of all AOs at n_pts points (DfCalcGridX.cpp:667-769).
 * Outputs are n_pts x n_aos, row-major. grad may be NULL. */
void oracle_eval_aos(const oracle_aos* aos, int64_t n_pts, const double* xyz,
                     double* phi, double* gx, double* gy, double* gz);

/* RKS GGA build (DfCalcGridX.cpp:2305-2315 with :1923-1962 and :2137-2193,
 * :2455-2500).  grid: n_pts x 4 (x, y, z, w).  PA: dense N x N alpha density
 * (symmetric).  FA: dense N x N, the lower triangle is ACCUMULATED (upper
 * left untouched).  rho_out (optional): n_pts x 4 (rho, drho/dx, dy, dz).
 * Returns E_xc.  functional: any id (LDA ids give zero gamma derivatives, so
 * the GGA loop reduces to the LDA one). n_threads<=0: all. */
double oracle_calc_xc_rks(const oracle_aos* aos, int functional, int64_t n_pts,
                          const double* grid, const double* PA, double* FA,
                          double* rho_out, double density_cutoff,
                          int n_threads);

/* UKS build (DfCalcGridX.cpp:2291-2330 with :1892-1921, :1964-2016,
 * :2093-2135, :2195-2278).  PA, PB: dense alpha / beta densities; FA, FB:
 * lower triangles ACCUMULATED.  rho_out (optional): n_pts x 8 (rhoA, grad
 * rhoA, rhoB, grad rhoB).  exact_gab: 0 = the reference (vgAB times the
 * same-spin gradient), 1 = the exact derivative (test use). */
double oracle_calc_xc_uks(const oracle_aos* aos, int functional, int64_t n_pts, const double* grid,
                          const double* PA, const double* PB, double* FA, double* FB, double* rho_out,
                          double density_cutoff, int exact_gab, int n_threads);

/* Lebedev rule of the given order (the 22 orders 6 ... 1730 of TlLebedevGrid);
 * returns number of points written (0 if unsupported). xyz: n x 3, w: n. */
int oracle_lebedev(int order, double* xyz, double* w);

/* Euler-Maclaurin radial abscissa/weight (DfGenerateGrid.cpp:1597-1610),
 * including the reference's 32-bit evaluation of i^5. */
void oracle_radial_em(double R, int Nr, int i, double* r, double* w);

/* Bragg-Slater radius table value as used by the grid (TlPrdctbl.cpp:150-160). */
double oracle_bragg_slater(int Z);

/* Standard-orientation matrix O (DfGenerateGrid.cpp:1739-1809): eigenvectors
 * (columns, ascending eigenvalues) of M = sum_A Z_A (|R_A-T|^2 I - (R_A-T)(R_A-T)^T).
 * Sign convention: the largest-|.| component of each column is positive
 * (parity unpinned against LAPACK). O is row-major 3x3. */
void oracle_o_matrix(int n_atoms, const int* Z, const double* xyz, double* O);

/* Becke partition weight P_A / sum_B P_B at point r (DfGenerateGrid.cpp:1811-1876,
 * :2043-2086) with atomic size adjustments. */
double oracle_becke_weight(int n_atoms, const int* Z, const double* xyz,
                           int iatom, const double* r, int size_adjust);

/* Molecular grid (DfGenerateGrid.cpp:730-768, :934-1091) for grid-type USER,
 * EM radial, no pruning, Becke partitioning with size adjustments, weight
 * cutoff.  Atom blocks are emitted in atom order.  If out_grid is NULL only
 * the count is returned.  out_grid: n x 4 (x, y, z, w); out_atom: n. */
int64_t oracle_generate_grid(int n_atoms, const int* Z, const double* xyz,
                             int n_rad, int n_ang, const double* O,
                             double weight_cutoff, double* out_grid,
                             int* out_atom, int n_threads);

/* Grid options, as DfGenerateGrid's constructor resolves them (DfGenerateGrid.cpp:39-170).
 * radial 0 EM, 1 GC; gc_mapping 0 Becke, 1 TA, 2 KK; pruning 0 none, 1 MHL,
 * 2 SG-1; partitioning 0 Becke, 1 SS; size_adjust: Becke atomic size
 * adjustments; radii 0 Bragg-Slater, 1 SG-1. */
typedef struct {
    int n_rad, n_ang, radial, gc_mapping, pruning, partitioning, size_adjust, radii;
    double weight_cutoff;
} oracle_grid_opts;

/* Gauss-Chebyshev radial abscissa/weight (DfGenerateGrid.cpp:1612-1655). */
void oracle_radial_gc(double R, int Nr, int i, int mapping, double* r, double* w);
/* SG-1 radius list (DfGenerateGrid.cpp:409-429). */
double oracle_sg1_radius(int Z);
/* the largest supported Lebedev order <= n (6 when n < 6). */
int oracle_lebedev_floor(int n);
/* angular order of a radial shell at r (DfGenerateGrid.cpp:1658-1713). */
int oracle_pruned_order(int pruning, int Z, double r, double R, int n_ang);
/* Stratmann-Scuseria partition weight (DfGenerateGrid.cpp:1878-1956). */
double oracle_ss_weight(int n_atoms, const double* xyz, int iatom, const double* r);
/* The molecular grid for any option set (DfGenerateGrid.cpp:934-1091). */
int64_t oracle_generate_grid_ex(int n_atoms, const int* Z, const double* xyz, const oracle_grid_opts* opt,
                                const double* O, double* out_grid, int* out_atom, int n_threads);

/* AO second derivatives (getD2AOs_core, DfCalcGridX.cpp:771-863): h6 is
 * n_pts x n_aos x 6 (xx, xy, xz, yy, yz, zz). */
void oracle_eval_ao_hessians(const oracle_aos* aos, int64_t n_pts, const double* xyz, double* h6);

/* grad_B omega_A (size-adjusted Becke partition of atom A) at point r for
 * every atom B, point held fixed: out n_atoms x 3 (row A is zero).
 * Returns omega_A. */
double oracle_becke_weight_gradient(int n_atoms, const int* Z, const double* xyz, int A, const double* r,
                                    double* out);

/* The same for partition 0 (Becke with size adjustments), 1 (Becke without)
 * or 2 (Stratmann-Scuseria). */
double oracle_partition_weight_gradient(int n_atoms, const int* Z, const double* xyz, int A, const double* r,
                                        int partition, double* out);

/* XC nuclear gradient, RKS GGA (energyGradient, DfCalcGridX.cpp:1215-1415)
 * on the given grid: grid n_pts x 4 (x, y, z, w), pt_atom[p] = the atom the
 * point belongs to (the point moves with it), ao_atom[i] = atom of AO i,
 * Z / xyz the atoms (Becke partition with size adjustments).  grad_f: the
 * functional term (doubled for RKS as the reference does); grad_w: the
 * partition-weight term with the exact Becke derivative (the reference's
 * term is identically zero, see the .c file).  Both n_atoms x 3, dE/dR.
 * Returns E_xc. */
double oracle_xc_gradient_rks(const oracle_aos* aos, const int* ao_atom, int functional, int64_t n_pts,
                              const double* grid, const int* pt_atom, int n_atoms, const int* Z,
                              const double* xyz, const double* PA, double density_cutoff,
                              double* grad_f, double* grad_w, int n_threads);
/* ... with the weight term of partition 0, 1 or 2 (as above). */
double oracle_xc_gradient_rks_ex(const oracle_aos* aos, const int* ao_atom, int functional, int64_t n_pts,
                                 const double* grid, const int* pt_atom, int n_atoms, const int* Z,
                                 const double* xyz, const double* PA, double density_cutoff, int partition,
                                 double* grad_f, double* grad_w, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
