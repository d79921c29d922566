// Internal declarations shared by the host engine (pdfxc_host.cpp) and the
// kernels (pdfxc_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pdfxc.h"

namespace pdfxc {

// precision modes: every non-FP64 mode keeps the split structures and runs the
// K pass on bf16 slices; the rho pass is split only in SPLIT_BF16 / SPLIT_BF16X3
inline bool prec_split_rho(int p) { return p == PDFXC_SPLIT_BF16 || p == PDFXC_SPLIT_BF16X3 || p == PDFXC_SPLIT_BFP; }
// the K pass in FP16 hi + lo with per-slab column exponents (PDFXC_SPLIT_K_F16)
inline bool prec_k_f16(int p) { return p == PDFXC_SPLIT_K_F16; }
// the K pass's slices: bf16 hi + lo, FP16 hi + lo, else bf16 hi + mid + lo
inline int prec_slices(int p) { return (p == PDFXC_SPLIT_BF16 || p == PDFXC_SPLIT_K_F16) ? 2 : 3; }
// the rho pass in block floating point: 4 slices with shared exponents (PDFXC_SPLIT_BFP)
inline bool prec_bfp(int p) { return p == PDFXC_SPLIT_BFP; }
constexpr int kBfpPlanes = 4;
constexpr int kRhoFp64Class = 16;  // split rho launch classes from here on: FP64 k_rho class - kRhoFp64Class
// bf16 elements of one molecule's P slices (and, in block FP, its row exponents after the planes)
inline int64_t psplit_elems(int p, int n) {
    const int64_t nc = (n + 7) / 8 * 8;
    if (prec_bfp(p)) return (kBfpPlanes * (int64_t)n * nc + 2 * (int64_t)n + 7) / 8 * 8;
    return (int64_t)prec_slices(p) * n * nc;
}


constexpr int kBlockPts = 128;  // grid points per block (rows of the MFMA tiles)
constexpr int kColPad = 16;     // AO columns of a block are padded to a multiple of this
constexpr int kMaxPrim = 16;
constexpr double kPi = 3.14159265358979323846;

// One shell = the 2l+1 consecutive AOs that share a centre and primitives.
struct Shell {
    int first_ao;  // global index of the first AO
    int l;         // 0 s, 1 p, 2 d
    int nprim;
    double center[3];
    double exps[kMaxPrim];   // ascending
    double coefs[kMaxPrim];  // normalised
};

constexpr int kMaxDynLds = 160 * 1024 - 16 * 1024;  // dynamic LDS ceiling per workgroup (gfx950: 160 KiB total)

// Device-side view of one grid block.
struct BlockDesc {
    int pt_begin;     // first slot (block-sorted order; every block owns 128 slots)
    int npts;         // real points (<= 128); slots npts..127 have weight 0
    int shell_begin;  // offset into the block-shell list
    int nshell;
    int nao;          // screened AO columns N_b
    int ldn;          // N_b padded to a multiple of kColPad
    int col_begin;    // offset into the block-AO-index list
    int ldp;          // AO count of the block's molecule: leading dimension of its P and K
    int64_t poff;     // offset of the molecule's dense P and K in the P / K buffers
    int64_t psoff;    // offset of the molecule's P slices (split precision)
};

// Split-precision column layout of a block: nchunk 4-aligned runs of AO
// positions in the split order (chunk ids in the chunk list from
// chunk_begin), ldnp = 4 nchunk rounded up to 32.
struct SplitBlock {
    int chunk_begin, nchunk, ldnp, pad;
};

// One k_kmat work item: lower 16 x 16 tiles [t0, t1) of a block (k_kmat_split:
// lower 2 x 2 tile patches [t0, t1), ps = slab buffers).
struct KWork {
    int block, t0, t1, ncols, nsh, ps;
};

// A launch of one kernel instantiation over a contiguous range of a work list.
struct LaunchRange {
    int cls;     // template instantiation
    int offset;  // first work item
    int count;   // work items (blocks for k_rho, KWork for k_kmat)
    size_t lds;  // dynamic LDS bytes
};

// A split K pass work unit: 1 to kMaxGroup spatially adjacent blocks
// (siblings of the point bisection) whose K contributions are summed in
// registers on the union of their AO columns before one scatter into K.
constexpr int kMaxGroup = 8;
constexpr int kMaxMaskWords = 32;  // union shells per group <= 32 x 32
struct KGroup {
    int sub_begin;    // the group's blocks: kg_blocks[sub_begin .. + nsub)
    int nsub;
    int shell_begin;  // union shells: kg_shells (global ids, ascending), kg_shell_col, kg_sorder
    int nshell;
    int col_begin;    // union columns: kg_cols (global AO index, ascending)
    int nao;
    int mask_begin;   // kg_masks: nsub x nword words; bit k of block j = union shell k kept by block j
    int nword;
    int ldp;          // P / K leading dimension (the molecule's AO count)
    int pad;
    int64_t poff;     // offset of the molecule's dense P and K
};

// A k_kmat_split item: lower 2 x 2 tile patches [q0, q1) of a group's union
// in column-major order (patch column Q, then rows P >= Q); its slabs hold
// the union columns [c0, c1) (c0 = 32 Q(q0), c1 = 32 Tp), which the union
// shells k0 .. nshell - 1 cover; nbuf slab buffers (1 or 2).
struct KGWork {
    int group, q0, q1, c0, c1, k0, nbuf, pad;
};

struct DeviceShell {
    double cx, cy, cz;
    int l, nprim;
    int first_ao;
    int pad;
    double exps[kMaxPrim];
    double coefs[kMaxPrim];
};

struct Context {
    pdfxc_opts opts;
    pdfxc_tuning tune{0, 96, 0, 24, 4, 2, 0.0};  // launch shapes (pdfxc_set_tuning / pdfxc_autotune)
    int n_aos = 0;
    std::vector<Shell> shells;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};

    // grid (host copies in caller order)
    int64_t n_pts = 0;
    std::vector<double> gx, gy, gz, gw;
    std::vector<int> gen_atom;  // atom of each point when pdfxc_generate_grid made the grid
    int grid_partition = 0;     // partition weights of a generated grid (PDFXC_PARTITION_*)
    int grid_size_adjust = 1;   // ... and its Becke size adjustments

    // blocks
    std::vector<BlockDesc> blocks;     // launch order (descending cost)
    std::vector<int> block_shells;     // global shell ids per block
    std::vector<int> block_cols;       // global AO index per block column
    std::vector<int> block_shell_col;  // local first column of each block shell
    std::vector<int> block_sorder;     // local shells of each block ordered by (l, nprim)
    std::vector<int64_t> slot_of_point;  // caller point -> slot (block*128 + k)
    std::vector<LaunchRange> rho_launches, kmat_launches;
    double flops = 0.0;
    double mean_nao = 0.0;
    int max_nao = 0;

    // device buffers
    DeviceShell* d_shells = nullptr;
    double* d_shrec = nullptr;  // per shell: cx, cy, cz, l + 4 nprim, (exp, coef) x maxprim
    int rec_stride = 0;
    double* d_pts = nullptr;  // 4 x n_slots SoA: x, y, z, w
    int64_t n_slots = 0;
    BlockDesc* d_blocks = nullptr;
    int* d_block_shells = nullptr;
    int* d_block_cols = nullptr;
    int* d_block_shell_col = nullptr;
    int* d_block_sorder = nullptr;
    int* d_rho_work = nullptr;    // block indices grouped by k_rho class
    void* d_k_work = nullptr;     // KWork items grouped by k_kmat class
    double* d_coef = nullptr;     // per slot: (w vrho / 2, cx, cy, cz)
    double* d_rho = nullptr;      // per slot: rho, grad rho (4)
    double* d_rho_b = nullptr;    // UKS: beta rho, grad rho per slot (allocated by the first UKS build)
    double* d_coef_b = nullptr;   // UKS: beta point terms
    double* d_eblock = nullptr;   // per 256 slots: energy (k_xc_points partial sums)
    double* d_P = nullptr;        // dense N x N
    double* d_K = nullptr;        // dense N x N
    double* d_P2 = nullptr;       // UKS host entry points: beta P and K (allocated on first use)
    double* d_K2 = nullptr;
    double* d_E = nullptr;        // 1 double
    double* h_pinned = nullptr;   // packed staging (pinned)

    // split precision (opts.precision != PDFXC_FP64)
    int split_nc = 0;                   // N rounded up to 8: row length of the P slices
    std::vector<int> split_perm;        // split AO order: position -> AO (each atom's shells diffuse first)
    int* d_perm = nullptr;
    std::vector<SplitBlock> sblocks;
    std::vector<int> chunks;            // per block: chunk ids (split-order AO position / 4), ascending
    std::vector<int> sshell_col;        // per block shell: local column in the chunk layout
    std::vector<LaunchRange> srho_launches, sk_launches;
    unsigned short* d_Ps = nullptr;     // S x N x split_nc bf16 slices of P in the split AO order
    SplitBlock* d_sblocks = nullptr;
    int* d_chunks = nullptr;
    int* d_sshell_col = nullptr;
    int* d_srho_work = nullptr;
    // split K pass groups (pdfxc_tuning.split_k_group blocks each) and their items
    std::vector<KGroup> kgroups;
    std::vector<int> kg_blocks, kg_shells, kg_shell_col, kg_sorder, kg_cols;
    std::vector<unsigned> kg_masks;
    KGroup* d_kgroups = nullptr;
    int* d_kg_blocks = nullptr;
    int* d_kg_shells = nullptr;
    int* d_kg_shell_col = nullptr;
    int* d_kg_sorder = nullptr;
    int* d_kg_cols = nullptr;
    unsigned* d_kg_masks = nullptr;
    KGWork* d_kg_work = nullptr;        // items of k_kmat_split
    std::vector<LaunchRange> k_fp64_launches;  // k_kmat items of blocks too wide for k_kmat_split
    int64_t n_k_fp64_blocks = 0;

    int rho_accumulate = 0;             // update mode: the rho kernels add rho[dP] to d_rho

    // persistent launches: work-queue counters of one build sequence (zeroed
    // by a fill kernel when the sequence starts), the next free one, and the
    // device's compute units
    int* d_queue = nullptr;             // kMaxQueues counters
    int n_queue_used = 0;
    bool queue_open = false;            // begin_queues ran for the current sequence
    int n_cu = 0;

    // host copies of the device work lists (d_rho_work, d_k_work, d_srho_work, d_sk_work)
    std::vector<int> h_rho_work, h_srho_work;
    std::vector<KWork> h_k_work;
    std::vector<KGWork> h_kg_work;

    // batches: P slices made per molecule (empty for one molecule)
    struct PsplitJob {
        int64_t poff, psoff;
        int perm_off, n;
    };
    std::vector<PsplitJob> psplit_jobs;

    double last_build_ms = 0.0;
    double last_kernel_ms[4] = {0, 0, 0, 0};
};

constexpr int kMaxQueues = 64;  // persistent launches per build sequence

// error helpers
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define PDFXC_HIP(call)                                          \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::pdfxc::hip_fail(e_, #call); \
    } while (0)

// kernel launchers (pdfxc_kernels.hip)
int launch_rho(Context& c, const LaunchRange& r);
int rho_points_per_workgroup(int cls);
// dynamic LDS the 64-point k_rho class of ldn <= 256 blocks may take: 160 KiB
// less the kernel's static LDS, read from the code object
int rho64_max_dyn_lds(size_t* out);
int launch_xc_points(Context& c, int xc_type, double* d_ewf = nullptr);
int launch_kmat(Context& c, const LaunchRange& r);
int configure_kernels();
int launch_energy_sum(Context& c, double* dE);
int launch_zero(double* d, int64_t n, hipStream_t s);
// start a build sequence's persistent launches: zero the counters on the stream
int begin_queues(Context& c);
// kernel families of pdfxc_tuning.persistent (bit mask)
constexpr int kQueueRho = 1, kQueueK = 2;
// grid and queue of a launch of n_items work items of a kernel family whose
// persistent instantiation keeps `per_cu` workgroups resident per compute
// unit: a persistent launch of the resident workgroups over the next counter
// when the items exceed them (and the sequence's counters are open and the
// family's bit of tune.persistent is set), else one workgroup per item of the
// plain instantiation (queue = nullptr)
int* launch_queue(Context& c, int kind, int n_items, int per_cu, unsigned* grid);
// workgroups of `kernel` resident per compute unit at `threads` and `lds` bytes
// of dynamic LDS (at least 1)
template <typename Kernel>
inline int resident_per_cu(Kernel kernel, int threads, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(kernel), threads, lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb > 0 ? nb : 1;
}
int launch_energy_sums(Context& c, const int* d_bounds, int n_mol, double* dE);  // one E per molecule
int launch_xc_points_uks(Context& c, int xc_type);
int ensure_beta_buffers(Context* c);  // UKS beta rho / point-term buffers, zeroed when first made
int launch_eval_functional_uks(int xc_type, int64_t n, const double* d_in, double* d_out, hipStream_t s);
int launch_eval_functional(int xc_type, int64_t n, const double* d_in, double* d_out, hipStream_t s);
// pdfxc_eval_functional through the library's own device staging (no allocation)
int eval_functional_noalloc(int xc_type, int64_t n, const double* in, double* out);
int launch_eval_aos_points(const Context& c, int64_t n, const double* d_xyz, double* d_out,
                           hipStream_t s);
int launch_becke(const double* d_atoms, int n_atoms, int64_t n, const double* d_pts,
                 const int* d_atom_of_pt, double* d_w, int partition, hipStream_t s);
int launch_mfma_selftest(double* d_out, hipStream_t s);
// split precision (pdfxc_split.hip)
int split_points_per_workgroup(int cls);
// block-FP rho classes 0, 1, 2: 64, 32, 16 points per workgroup, blocks of at
// most 256, 384, 512 chunked columns; the dynamic LDS a class may take
int bfp_points_per_workgroup(int cls);
int bfp_max_columns(int cls);
int bfp_max_dyn_lds(int cls, size_t* out);
int launch_split_rho(Context& c);   // P slices + k_rho_split
int launch_split_kmat(Context& c);  // k_kmat_split
int split_kmat_max_dyn_lds(int S, size_t* out);
int launch_mfma_split_selftest(float* d_out, hipStream_t s);
// XC nuclear gradient (pdfxc_grad.hip); d_gacc: functional term [3 n_atoms] then weight term [3 n_atoms]
int launch_grad_fn(Context& c, const LaunchRange& r, const int* d_shell_atom, const int* d_slot_atom,
                   int n_atoms, double* d_gacc);
int launch_grad_w(Context& c, const double* d_atoms, int n_atoms, const int* d_slot_atom, const double* d_ewf,
                  double* d_gacc, int partition);
int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s);

}  // namespace pdfxc
