/*
 * pdfxc -- MI355X (gfx950) exchange-correlation quadrature engine for
 * ProteinDF's grid XC path.  C ABI: plain pointers and sizes, no C++ types.
 *
 * The operator this library replaces is
 *
 *   double DfCalcGridX::calcXCIntegForFockAndEnergy(
 *       const TlDenseSymmetricMatrix_Lapack& P_A, DfFunctional_GGA* pFunctional,
 *       TlDenseSymmetricMatrix_Lapack* pF_A);                // DfCalcGridX.h:203-205
 *
 * called from DfXCFunctional::getFxc<> (DfXCFunctional.h:133-190) with
 * P_A = P/2 and a zeroed F_A.  The engine reads the same inputs (AO table of
 * TlOrbitalInfo, the grid that DfGenerateGrid writes to grid.mat, the alpha
 * density in packed lower-triangle order) and returns the same outputs
 * (E_xc, F_A accumulated in packed lower-triangle order).
 *
 * Every entry point returns PDFXC_OK (0) or an error code; pdfxc_last_error()
 * gives a message for the calling thread.  A context owns one HIP stream and
 * its device buffers; it is not thread-safe.  Host buffers stay owned by the
 * caller.
 */
#ifndef PDFXC_H
#define PDFXC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PDFXC_OK = 0,
    PDFXC_ERR_ARGUMENT = 1,    /* bad size, null pointer, unsupported value */
    PDFXC_ERR_HIP = 2,         /* HIP runtime error (message has details) */
    PDFXC_ERR_NO_GRID = 3,     /* build called before a grid was set */
    PDFXC_ERR_UNSUPPORTED = 4, /* functional / basis type not implemented */
    PDFXC_ERR_MEMORY = 5       /* device or host allocation failed */
};

/* functionals: the XC_TYPE values of DfXCFunctional.h:44, dispatched as
 * getFxc does (DfXCFunctional.h:140-180).  HF has no grid term: a build
 * returns E_xc = 0 and leaves K (and rho_out) untouched.  HFS and SVWN are
 * LDA (the reference's calcRho_LDA / buildVxc(LDA) path, DfCalcGridX.cpp
 * :1816-1890, :2056-2091, :2350-2368: K = sum w v_rho phi phi); HFB, BLYP
 * and B3LYP are GGA.  rho_out carries the gradient columns for all of them. */
enum {
    PDFXC_HF = 1,    /* no XC term */
    PDFXC_HFS = 2,   /* Slater (DfFunctional_HFS.cpp) */
    PDFXC_SVWN = 3,  /* Slater + VWN5 (DfFunctional_SVWN.cpp) */
    PDFXC_HFB = 4,   /* Becke 88 (DfFunctional_Becke88.cpp) */
    PDFXC_BLYP = 5,  /* Becke 88 + LYP (DfFunctional_B88LYP.cpp) */
    PDFXC_B3LYP = 6  /* DfFunctional_B3LYP.cpp */
};

/* arithmetic of the two contractions X = Phi.P and K = Phi^T.V:
 * FP64 MFMA, or bf16 / fp16 MFMA with FP32 accumulation on operands split
 * into slices (hi + lo: 3 slice products; hi + mid + lo: 6).
 *   PDFXC_SPLIT_BF16, PDFXC_SPLIT_BF16X3: both passes in bf16 hi + lo, or
 *     hi + mid + lo;
 *   PDFXC_SPLIT_K_BF16X3: the rho pass in FP64 (so E_xc and rho are the FP64
 *     build's), the K pass in bf16 hi + mid + lo;
 *   PDFXC_SPLIT_BFP: the rho pass in block floating point (4 bf16 slices of
 *     8 bits sharing one exponent per point and per P row; the FP32 sums of
 *     integer slice products are exact), the K pass in bf16 hi + mid + lo;
 *   PDFXC_SPLIT_K_F16: the rho pass in FP64, the K pass in fp16 hi + lo with
 *     one power-of-two scale per AO column and 16-point slab, so that the
 *     values stay inside fp16's exponent range.
 * AO values, the functional, the rho / grad rho dots and the K_xc sums stay
 * FP64 in every mode.  No reference counterpart (the reference is FP64 only). */
enum {
    PDFXC_FP64 = 0,
    PDFXC_SPLIT_BF16 = 1,
    PDFXC_SPLIT_BF16X3 = 2,
    PDFXC_SPLIT_K_BF16X3 = 3,
    PDFXC_SPLIT_BFP = 4,
    PDFXC_SPLIT_K_F16 = 5
};

typedef struct pdfxc_opts {
    double density_cutoff;   /* rho_alpha cutoff, xc-density-threshold (DfCalcGridX.cpp:55-59), 1e-16 */
    double ao_screen;        /* per-block AO screening bound; 0 keeps every AO the reference
                                evaluates as non-zero (only the alpha r^2 > 30 primitive cut,
                                DfCalcGridX.cpp:42, :690-700) */
    int precision;           /* PDFXC_FP64 or one of the PDFXC_SPLIT_* modes above */
    int block_points;        /* points per grid block: 128 */
    int64_t scratch_bytes;   /* reserved, must be 0 (AO values are never stored in HBM) */
} pdfxc_opts;

typedef struct pdfxc_stats {
    int64_t n_points;        /* grid points */
    int64_t n_blocks;        /* grid blocks */
    int64_t n_chunks;        /* kernel launches per build (one per kernel instantiation class) */
    double flops;            /* algorithmic sum_b 4 n_b N_b^2 (N_b = screened AOs of block b) */
    double mean_nao;         /* mean N_b */
    int max_nao;             /* max N_b */
    double last_build_ms;    /* device time of the last build (HIP events) */
    double last_kernel_ms[4];/* per kernel: k_rho (rho + functional), k_kmat (K_xc), energy sum, unused */
    int64_t n_k_fp64_blocks; /* split precision: blocks whose K pass runs in FP64 (too wide for the
                                bf16 K kernel's LDS) */
} pdfxc_stats;

/* Launch shapes of the build kernels.  rho_shift (0..3) / split_rho_shift
 * (0..2): move every block that many classes toward fewer points per
 * workgroup (rho pass; at rho_shift 0 a block of 129..256 padded AO columns
 * takes 64 points per workgroup when they fit in LDS, else 32; in the block
 * floating point rho pass a block takes the first of the 64-, 32- and
 * 16-point classes from split_rho_shift on that holds its columns and fits
 * in LDS); k_tiles (8..96): FP64 K pass work item size in lower
 * 16 x 16 tiles for blocks of more than 32 tiles; split_k_patches (4..32):
 * split K pass work item size in 2 x 2 tile patches (up to 16: workgroups of
 * 16 waves, one patch each; up to 24: 12 waves of two; up to 32: 8 waves of
 * four); split_k_group (1, 2, 4 or 8): grid blocks (128 points each,
 * neighbours in the point bisection) whose K contributions the split K pass
 * sums on the union of their AO columns before one scatter into K, i.e.
 * 128 x split_k_group points per K work unit (a group too wide for the
 * kernel's LDS is halved); persistent (bit mask, 0..3): bit 0 the FP64
 * rho-pass kernels of the widest classes (those that run one workgroup per CU:
 * their persistent form loads the next item's set-up data while the current
 * one runs) and the split rho-pass kernels, bit 1 the K-pass kernels run as
 * persistent launches, as many workgroups as stay resident on the device,
 * each taking work items (blocks, K items; of every molecule in a batch) from
 * one device counter until they run out, instead of one workgroup per item.
 * Defaults 0, 96, 0, 24, 4, 2.
 * ms: the build time measured for it (pdfxc_autotune), else 0. */
typedef struct pdfxc_tuning {
    int rho_shift;
    int k_tiles;
    int split_rho_shift;
    int split_k_patches;
    int split_k_group;
    int persistent;
    double ms;
} pdfxc_tuning;

typedef struct pdfxc_ctx pdfxc_ctx;
typedef struct pdfxc_batch pdfxc_batch;

/* defaults: density_cutoff 1e-16, ao_screen 0, FP64, 128 points per block */
void pdfxc_default_opts(pdfxc_opts* opts);

/* Contracted-GTO normalisation as TlOrbitalInfoObject::setCGTO does it
 * (TlOrbitalInfoObject.cpp:38-106 with Fl_Gto::getNormalizedfactor /
 * getNormalized, Fl_Gto.cpp:558-708).  shell: 0 s, 1 p, 2 d.  Outputs are
 * sorted by ascending exponent.  Host only. */
int pdfxc_normalize_cgto(int shell, int nprim, const double* exps, const double* coefs,
                         double* out_exps, double* out_coefs);

/* Create a context from the per-AO table TlOrbitalInfo provides
 * (getPosition, getBasisType, getCgtoContraction, getExponent,
 * getCoefficient; TlOrbitalInfoObject.cpp:181-259).  ao_exps / ao_coefs are
 * n_aos x max_prim row-major, already normalised.  The AOs of one shell
 * (types 0 | 1,2,3 | 4..8 on one centre) must be consecutive. */
int pdfxc_ctx_create(pdfxc_ctx** ctx, int n_aos, const double* ao_center,
                     const int* ao_type, const int* ao_nprim, const double* ao_exps,
                     const double* ao_coefs, int max_prim, const pdfxc_opts* opts);
int pdfxc_ctx_destroy(pdfxc_ctx* ctx);

/* Grid as grid.mat holds it (CSFD columns x, y, z, weight; DfCalcGridX.h:36-55).
 * Partitions the points into blocks and screens AOs per block. */
int pdfxc_set_grid(pdfxc_ctx* ctx, int64_t n_pts, const double* x, const double* y,
                   const double* z, const double* w);

/* Generate the molecular grid on the device: grid-type USER, Euler-Maclaurin
 * radial rule, no pruning, Becke partitioning with atomic size adjustments,
 * |w| > weight_cutoff (DfGenerateGrid.cpp:39-250, :730-768, :934-1091,
 * :1597-1610, :1715-1876): pdfxc_generate_grid_ex with those options, except
 * that n_ang must be a supported Lebedev order.  Atoms with Z = 0 are
 * dummies.  Sets the grid of the context and returns the point count. */
int pdfxc_generate_grid(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                        int n_rad, int n_ang, double weight_cutoff, int64_t* n_pts);
/* Molecular grid options, as DfGenerateGrid's constructor resolves the
 * pdfparam keys (DfGenerateGrid.cpp:39-170, setCellPara :282-430). */
enum { PDFXC_RADIAL_EM = 0, PDFXC_RADIAL_GC = 1 };                       /* Euler-Maclaurin, Gauss-Chebyshev */
enum { PDFXC_GC_BECKE = 0, PDFXC_GC_TA = 1, PDFXC_GC_KK = 2 };           /* GC radial mappings */
enum { PDFXC_PRUNE_NONE = 0, PDFXC_PRUNE_MHL = 1, PDFXC_PRUNE_SG1 = 2 }; /* angular pruning */
enum { PDFXC_PARTITION_BECKE = 0, PDFXC_PARTITION_SS = 1 };              /* Becke, Stratmann-Scuseria */
enum { PDFXC_RADII_BRAGG = 0, PDFXC_RADII_SG1 = 1 };                     /* radial scale R per element */
typedef struct pdfxc_grid_opts {
    int n_rad;          /* radial shells (nrgrid) */
    int n_ang;          /* Lebedev order before pruning (nOgrid), floored to a supported one */
    int radial;         /* PDFXC_RADIAL_* */
    int gc_mapping;     /* PDFXC_GC_*, for PDFXC_RADIAL_GC */
    int pruning;        /* PDFXC_PRUNE_*: MHL n = min(int(5 r / R_Bragg * n_ang), n_ang); SG-1 regions */
    int partitioning;   /* PDFXC_PARTITION_* */
    int size_adjust;    /* Becke atomic size adjustments (isAtomicSizeAdjustments_) */
    int radii;          /* PDFXC_RADII_*: Bragg-Slater table, or the SG-1 list */
    double weight_cutoff; /* points with |w| <= weight_cutoff are dropped (1e-16) */
} pdfxc_grid_opts;

/* Options of a grid type ("sg-1", "coarse", "medium", "medium-fine", "fine",
 * "ultra-fine", "user"; case-insensitive, '_' for '-'), with the defaults the
 * constructor gives before the grid/ keys override them:
 * presets: EM, MHL pruning, SS partitioning, size adjustments, Bragg-Slater R;
 * sg-1: 50 x 194, EM, SG-1 pruning and radii, SS, no size adjustments;
 * user: 75 x 302, EM, MHL pruning, SS, size adjustments.  Unknown type:
 * PDFXC_ERR_UNSUPPORTED (the reference aborts). */
int pdfxc_grid_opts_preset(const char* grid_type, pdfxc_grid_opts* out);

/* Generate the molecular grid for any option set (generateGrid /
 * generateGrid_atom, DfGenerateGrid.cpp:730-768, :934-1091): per atom and
 * radial shell (r_i > 30 dropped) the pruned Lebedev rule in the standard
 * orientation, times the Becke or SS partition weight of the point's atom;
 * every atom, a dummy (Z = 0) included, has a partition cell as in the
 * reference (Becke skips dummy neighbours, SS does not).  Sets the grid of
 * the context and returns the point count. */
int pdfxc_generate_grid_ex(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                           const pdfxc_grid_opts* opts, int64_t* n_pts);
/* pdfxc_generate_grid_ex with the angular grids in a given orientation O
 * (3 x 3 row-major, orthogonal: point = O . (Lebedev point) r_i + atom)
 * instead of the one the reference derives from the geometry
 * (DfGenerateGrid.cpp:1739-1809), so that a caller moving atoms (finite
 * differences, geometry scans) keeps the grid frame fixed.  A matrix that is
 * not orthogonal to 1e-10 is refused with PDFXC_ERR_ARGUMENT. */
int pdfxc_generate_grid_oriented(pdfxc_ctx* ctx, int n_atoms, const int* Z, const double* xyz,
                                 const pdfxc_grid_opts* opts, const double* O, int64_t* n_pts);

/* Copy the context's grid back in the caller's point order (x, y, z, w
 * columns, n_pts each). */
int pdfxc_get_grid(pdfxc_ctx* ctx, double* x, double* y, double* z, double* w);

/* calcXCIntegForFockAndEnergy, RKS (DfCalcGridX.cpp:2280-2289 LDA, :2305-2315 GGA):
 * P_alpha_packed: N(N+1)/2 doubles, row r, column c <= r (RLHD order).
 * K_alpha_packed: same layout, ACCUMULATED into (the caller zeroes it).
 * E_xc: total XC energy.  rho_out (may be NULL): n_pts x 4 (rho_alpha,
 * d/dx, d/dy, d/dz) in the caller's point order, the values the reference
 * writes to grid.mat columns 5-8. */
int pdfxc_build_rks(pdfxc_ctx* ctx, int functional, const double* P_alpha_packed,
                    double* K_alpha_packed, double* E_xc, double* rho_out);

/* Same build on device-resident buffers, enqueued on the context's stream
 * and synchronised before return: dP dense N x N symmetric, dK dense N x N
 * whose lower triangle is accumulated, dE one double (overwritten). */
int pdfxc_build_rks_device(pdfxc_ctx* ctx, int functional, const double* dP, double* dK,
                           double* dE);

/* Update mode (xc-update = yes; DfXCFunctional.cpp:134-136 passes P_A =
 * dP/2 and calcRho_GGA_part adds rho[dP_A] to the grid's stored rho,
 * DfCalcGridX.cpp:1925-1961): rho and grad rho of dP_alpha are ADDED to the
 * context's device-resident values (the previous build's, or zero after
 * set_grid / generate_grid), and K_xc, E_xc follow from the accumulated rho
 * exactly as in pdfxc_build_rks.  Same buffers and layouts as
 * pdfxc_build_rks / pdfxc_build_rks_device. */
int pdfxc_build_rks_update(pdfxc_ctx* ctx, int functional, const double* dP_alpha_packed,
                           double* K_alpha_packed, double* E_xc, double* rho_out);
int pdfxc_build_rks_update_device(pdfxc_ctx* ctx, int functional, const double* d_dP, double* dK,
                                  double* dE);

/* UKS (calcXCIntegForFockAndEnergy(P_A, P_B, ...), DfCalcGridX.cpp:2291-2330;
 * buildVxc UKS :2093-2135, :2195-2278): PA, PB the alpha and beta densities
 * (not halved), KA, KB ACCUMULATED, layouts as pdfxc_build_rks.  A point
 * counts when rhoA or rhoB exceeds the density cut-off.  As the reference's
 * build_XC_Matrix does, K_A takes df/dgammaAB times grad rhoA (and K_B
 * times grad rhoB).  rho_out (may be NULL): n_pts x 8 (rhoA, grad rhoA,
 * rhoB, grad rhoB).  The update forms add rho[dPA], rho[dPB] to the stored
 * two-spin rho.  ROKS is this build with PA = P_closed/2 + P_open,
 * PB = P_closed/2 (DfXCFunctional.cpp:184-205). */
int pdfxc_build_uks(pdfxc_ctx* ctx, int functional, const double* PA_packed, const double* PB_packed,
                    double* KA_packed, double* KB_packed, double* E_xc, double* rho_out);
int pdfxc_build_uks_update(pdfxc_ctx* ctx, int functional, const double* dPA_packed, const double* dPB_packed,
                           double* KA_packed, double* KB_packed, double* E_xc, double* rho_out);
int pdfxc_build_uks_device(pdfxc_ctx* ctx, int functional, const double* dPA, const double* dPB, double* dKA,
                           double* dKB, double* dE);
int pdfxc_build_uks_update_device(pdfxc_ctx* ctx, int functional, const double* d_dPA, const double* d_dPB,
                                  double* dKA, double* dKB, double* dE);
/* Spin-resolved functional kernel: in n x 5 (rhoA, rhoB, gAA, gAB, gBB),
 * out n x 6 (f, df/drhoA, df/drhoB, df/dgAA, df/dgAB, df/dgBB). */
int pdfxc_eval_functional_uks(int functional, int64_t n, const double* in, double* out);

/* XC nuclear gradient, RKS (DfCalcGridX::energyGradient,
 * DfCalcGridX.cpp:1215-1415), on the context's grid with its points moving
 * rigidly with their atoms: point_atom[p] (caller point order) gives each
 * point's atom, or NULL when pdfxc_generate_grid made the grid (its own
 * assignment is used); AO centres must coincide with atoms.  Outputs dE/dR,
 * n_atoms x 3 each: grad_functional = the reference's functional term
 * (AO-centre and point-translation parts, doubled for RKS); grad_weight = the
 * derivative of the partition weights the grid was made with (Becke with or
 * without size adjustments, or Stratmann-Scuseria, as pdfxc_generate_grid_ex
 * made it; size-adjusted Becke for a caller's grid).  Dummy atoms are refused
 * with PDFXC_ERR_UNSUPPORTED.  (The reference's
 * own weight term evaluates to zero, DfGenerateGrid.cpp:2355-2372.)  Their
 * sum is the gradient of E_xc.  E_xc (may be NULL) is returned.  The
 * context's precision applies: FP64, or, in a split mode, the functional
 * term's contractions X = P phi and Z = P W on bf16 MFMA with the mode's
 * slices (hi+lo, or hi+mid+lo for PDFXC_SPLIT_BF16X3 and PDFXC_SPLIT_K_BF16X3)
 * and rho from the build's own rho pass (FP64 in PDFXC_SPLIT_K_BF16X3); the
 * rest stays FP64.  The context's stored rho becomes rho[P_alpha] (an
 * update-mode SCF restarts from a full build). */
int pdfxc_gradient_rks(pdfxc_ctx* ctx, int functional, int n_atoms, const int* Z, const double* xyz,
                       const int* point_atom, const double* P_alpha_packed, double* grad_functional,
                       double* grad_weight, double* E_xc);

/* Functional kernel over arbitrary closed-shell points, as the grid path
 * calls it (DfFunctional.h:190-192, :221-232: rhoB = rhoA, gAB = gBB = gAA).
 * in: n x 2 (rho_alpha, gamma_alpha_alpha); out: n x 4 (f, df/drhoA,
 * df/dgammaAA, df/dgammaAB; both zero for HFS and SVWN, all four for HF).
 * Host buffers; runs on the current device. */
int pdfxc_eval_functional(int functional, int64_t n, const double* in, double* out);

/* pdfxc_eval_functional without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Inputs
 * and outputs are staged through a fixed device array of the library, one per
 * GPU, held until the process ends: n is at most PDFXC_NOALLOC_MAX_POINTS
 * (larger n: PDFXC_ERR_ARGUMENT, nothing runs), i.e. 6 x 65536 doubles = 3 MiB
 * per GPU.  Runs on the current device and the null stream; calls that overlap
 * on one GPU share that array and race.  Same results as pdfxc_eval_functional. */
#define PDFXC_NOALLOC_MAX_POINTS 65536
int pdfxc_eval_functional_noalloc(int functional, int64_t n, const double* in, double* out);

/* ProteinDF matrix files (TlMatrixObject::MatrixType, tl_matrix_object.h:50-61):
 * a header (type, rows, cols) and the doubles.  Readers accept every header
 * width the reference accepts (type 1/4/8 bytes, counts 4/8 bytes, the first
 * whose implied size is the file's, tl_matrix_utils.cc:114-200, :334-358);
 * writers use the reference's own (1-byte type, 32-bit counts). */
enum { PDFXC_MAT_RSFD = 0, PDFXC_MAT_CSFD = 1, PDFXC_MAT_RLHD = 2 };
/* Type and shape of a dense matrix file; PDFXC_ERR_UNSUPPORTED if no header fits. */
int pdfxc_matrix_file_info(const char* path, int* type, int64_t* rows, int64_t* cols);
/* Symmetric matrix <-> RLHD file, packed lower triangle by rows
 * (TlDenseSymmetricMatrixObject::load / save, tl_dense_symmetric_matrix_object.cc:245-321).
 * The load checks the file's dimension against n. */
int pdfxc_load_symmetric_matrix(const char* path, int64_t n, double* packed);
int pdfxc_save_symmetric_matrix(const char* path, int64_t n, const double* packed);
/* General matrix, column-major in memory: loads RSFD or CSFD, saves CSFD
 * (TlDenseGeneralMatrixObject::load / save, tl_dense_general_matrix_object.cc:178-266). */
int pdfxc_load_general_matrix(const char* path, int64_t rows, int64_t cols, double* colmajor);
int pdfxc_save_general_matrix(const char* path, int64_t rows, int64_t cols, const double* colmajor);
/* grid.mat (DfObject::getGridMatrix / saveGridMatrix; columns DfCalcGridX.h:37-54):
 * one row per point: x, y, z, w, atom, then per spin rho (LDA: 6 columns
 * RKS, 7 UKS) or rho and grad rho (GGA: 9 RKS, 13 UKS, alpha first).  load:
 * sets the context's grid from columns 0-3, the point atoms (for
 * pdfxc_gradient_rks) from column 4, and, for 6, 7, 9 or 13 columns, the
 * stored rho an update-mode build adds to.  save: n_rho_cols 0, 1, 2, 4 or 8
 * after x, y, z, w, atom (-1 where the context has no point atoms); rho is the
 * context's current (last built) rho. */
int pdfxc_load_grid_matrix(pdfxc_ctx* ctx, const char* path, int64_t* n_pts, int* n_cols);
int pdfxc_save_grid_matrix(pdfxc_ctx* ctx, const char* path, int n_rho_cols);

/* Set the launch shapes and rebuild the context's work lists (the grid and its
 * blocks are kept; the stored rho of update mode is cleared). */
int pdfxc_set_tuning(pdfxc_ctx* ctx, const pdfxc_tuning* tuning);
/* Autotuner: time `reps` builds of functional on device buffers dP / dK / dE
 * (as pdfxc_build_rks_device) for each launch shape of a fixed candidate list
 * for the context's precision, keep the fastest and return it (with its time)
 * in chosen.  dK is accumulated into by every trial build. */
int pdfxc_autotune(pdfxc_ctx* ctx, int functional, const double* dP, double* dK, double* dE, int reps,
                   pdfxc_tuning* chosen);

/* Density of P at arbitrary points: DfCalcGridX::gridDensity
 * (DfCalcGridX.cpp:1038-1044) as TlDensField::makeDensFld uses it
 * (src/tools/TlDensField.cpp:35-53): rho = phi^T P phi with P as given (pass
 * the total density for the total rho).  xyz: n x 3; rho: n; grad_rho
 * (may be NULL): n x 3.  The context's grid and stored rho are not changed. */
int pdfxc_density_at_points(pdfxc_ctx* ctx, const double* P_packed, int64_t n, const double* xyz, double* rho,
                            double* grad_rho);

/* Batches: the grid blocks of several molecules (contexts with their grids
 * set) in one set of launches, captured in a HIP graph on the first build and
 * replayed while the buffers stay the same.  No counterpart in the reference,
 * which builds one molecule per call; each molecule's result is the one its
 * own context gives (E_xc bit for bit, K up to the order of its atomic sums).
 * The contexts must share precision and density cut-off and must outlive
 * nothing: the batch copies what it needs.
 * Buffers: the molecules' dense N_i x N_i P (and K, accumulated) one after
 * another, molecule i at dense_off[i] (pdfxc_batch_offsets); E: one E_xc per
 * molecule.  The host form takes the molecules' packed lower triangles one
 * after another (packed_off[i]). */
int pdfxc_batch_create(pdfxc_batch** out, int n_mol, pdfxc_ctx* const* ctxs);
int pdfxc_batch_destroy(pdfxc_batch* batch);
int pdfxc_batch_offsets(pdfxc_batch* batch, int64_t* dense_off, int64_t* packed_off); /* n_mol + 1 each */
int pdfxc_batch_build_rks_device(pdfxc_batch* batch, int functional, const double* dP, double* dK, double* dE);
int pdfxc_batch_build_rks(pdfxc_batch* batch, int functional, const double* P_packed, double* K_packed, double* E);
int pdfxc_batch_stats(pdfxc_batch* batch, pdfxc_stats* stats);

/* AO values and gradients of all AOs at arbitrary points (getAOs / getDAOs,
 * DfCalcGridX.cpp:655-769).  Outputs n_pts x n_aos row-major. */
int pdfxc_eval_aos(pdfxc_ctx* ctx, int64_t n_pts, const double* xyz, double* phi,
                   double* gx, double* gy, double* gz);

int pdfxc_get_stats(pdfxc_ctx* ctx, pdfxc_stats* stats);

/* The context's HIP stream (hipStream_t) for callers that enqueue around
 * pdfxc_build_rks_device. */
void* pdfxc_stream(pdfxc_ctx* ctx);

const char* pdfxc_last_error(void);

/* Diagnostics: one v_mfma_f64_16x16x4 product of fixed operands
 * A[i][k] = i + 100 k, B[k][j] = j - 7 k (i, j < 16, k < 4), written as the
 * 16 x 16 row-major D the engine's accumulator layout assumes.  Lets a test
 * confirm the lane layout the kernels rely on. */
int pdfxc_selftest_mfma(double* out16x16);

/* Diagnostics: the integer operands A[i][k] = ((3 i + 5 k) mod 11) - 5,
 * B[k][j] = ((7 k + 2 j) mod 13) - 6 (exact in bf16, fp16 and FP32)
 * multiplied by v_mfma_f32_16x16x32_bf16 (k < 32; out[0..255]), by
 * v_mfma_f32_16x16x16_bf16 (k < 16; out[256..511]), by
 * v_mfma_f32_16x16x32_f16 (k < 32; out[512..767]) and by
 * v_mfma_f32_16x16x16_f16 (k < 16; out[768..1023]), each written as the
 * 16 x 16 row-major D the split-precision kernels' operand and accumulator
 * layouts assume. */
int pdfxc_selftest_mfma_split(float* out4x16x16);

/* Diagnostics: a plain streaming copy of n doubles between device buffers
 * (reads 8n bytes, writes 8n bytes), the reference point for HBM traffic
 * counters of the other kernels. */
int pdfxc_selftest_copy(const double* d_src, double* d_dst, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
