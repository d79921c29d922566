// Split-precision kernels of the grid XC build for gfx950 (CDNA4).
//
// The two contractions run on v_mfma_f32_16x16x32_bf16 with every FP64
// operand x written as S bf16 slices (x = h0 + h1 [+ h2] up to the last
// slice's rounding, each slice the nearest bf16 to what the previous ones
// leave) and the slice products h_a h'_b with a + b < S accumulated in FP32:
// S = 2 is hi + lo (3 products), S = 3 is hi + mid + lo (6 products).  AO
// values, the functional, rho / grad rho dots and the K_xc scatter stay
// FP64.  AOs are renumbered once per context (each atom's shells most
// diffuse first, so the shells a block keeps of a distant atom are one run)
// and block columns are 4-aligned runs of that order ("chunks"), so that the
// P operand is two 8-byte loads per lane; AOs of a chunk that the block's
// screening dropped are zero columns.
//   k_psplit      P (dense FP64 N x N) -> S bf16 slice planes [S][N][Nc], split order
//   k_rho_split   X = Phi_b P_bb for PB points, then rho, grad rho
//                 (calcRho_GGA_part, DfCalcGridX.cpp:1923-1962)
//   k_kmat_split  K_uu += Phi^T V + V^T Phi over a group of adjacent blocks on
//                 the union u of their AO columns, lower 2 x 2 tile patches,
//                 FP64 atomics into the lower triangle of K (build_XC_Matrix
//                 GGA, :2455-2500)
#include <hip/hip_runtime.h>

#include "pdfxc_device.hpp"
#include "pdfxc_internal.hpp"
#include "pdfxc_split.hpp"

namespace pdfxc {

// ---------------------------------------------------------------------------
// P slices in the split order: Ps[s][r][c] = slice s of P[perm r][perm c]
// for r < N, c < Nc (Nc = N rounded up to 8; columns c >= N are zero)
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_psplit(const double* __restrict__ P, const int* __restrict__ perm, int N,
                                                int Nc, unsigned short* __restrict__ Ps) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t plane = (int64_t)N * Nc;
    if (i >= plane) return;
    const int r = (int)(i / Nc), c = (int)(i % Nc);
    const double x = (c < N) ? P[(int64_t)perm[r] * N + perm[c]] : 0.0;
    unsigned short h[S];
    split_bf16<S>(x, h);
#pragma unroll
    for (int s = 0; s < S; ++s) Ps[s * plane + i] = h[s];
}

// ---------------------------------------------------------------------------
// rho pass, split precision.  PB points of a block per workgroup (512
// threads).  LDS plane A[s][p][c] (bf16, row stride ldnp + 8) holds the Phi
// slices; X = Phi P_bb accumulates in FP32 with wave w owning the AO column
// tiles J = w, w + 8, ... and all PB / 16 point tiles; the B operand (P_bb
// rows = P rows by symmetry) is loaded per lane as the 8 AOs of two chunks.
// X is then stored over the plane in FP64 for half of the points at a time
// (row stride ldnp + 1) and dotted with phi, grad phi in FP64.
// ---------------------------------------------------------------------------
template <int S, int PB, int NJ, int NJMAX>
__device__ inline void rho_split_tiles(int wave, int lane, int ldnp, int nch, const int* s_chunk,
                                       const unsigned short* s_a, const unsigned short* __restrict__ Ps, int N,
                                       int Nc, v4d (&acc)[PB / 16][NJMAX]) {
    constexpr int NP = PB / 16;
    const int li = lane & 15, lq = lane >> 4;
    const int RA = ldnp + 8;
    const int64_t plane = (int64_t)N * Nc;
    // each K step's slice products accumulate in FP32 from zero and are then
    // added to the FP64 accumulators, so FP32 rounding acts on one step's sum only
    int64_t prow[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const int c = (wave + 8 * jj) * 16 + li;
        const int ch = c >> 2;
        const int g = s_chunk[ch] * 4 + (c & 3);
        prow[jj] = (int64_t)((ch < nch && g < N) ? g : 0) * Nc;  // padding columns: any row (Phi is 0 there)
    }
    const int nks = ldnp / 32;
    for (int ks = 0; ks < nks; ++ks) {
        const int koff0 = s_chunk[ks * 8 + 2 * lq] * 4;
        const int koff1 = s_chunk[ks * 8 + 2 * lq + 1] * 4;
        bf16x8 bfr[NJ][S];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const unsigned short* row = Ps + s * plane + prow[jj];
                const bf16x4 lo = *reinterpret_cast<const bf16x4*>(row + koff0);
                const bf16x4 hi = *reinterpret_cast<const bf16x4*>(row + koff1);
                bfr[jj][s] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
            }
        bf16x8 afr[NP][S];
#pragma unroll
        for (int ip = 0; ip < NP; ++ip)
#pragma unroll
            for (int s = 0; s < S; ++s)
                afr[ip][s] = *reinterpret_cast<const bf16x8*>(s_a + (size_t)(s * PB + ip * 16 + li) * RA +
                                                              ks * 32 + lq * 8);
#pragma unroll
        for (int ip = 0; ip < NP; ++ip)
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj) {
                // hi x hi in an FP32 sum of its own: the corrections (at most
                // 2^-8 of it) do not add their roundings to the leading term
                const f32x4 z = f32x4{0.f, 0.f, 0.f, 0.f};
                const f32x4 a32 = mfma_bf16(afr[ip][0], bfr[jj][0], z);
                f32x4 c32 = z;
#pragma unroll
                for (int q = 1; q < SlicePairs<S>::n; ++q)
                    c32 = mfma_bf16(afr[ip][SlicePairs<S>::a[q]], bfr[jj][SlicePairs<S>::b[q]], c32);
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[ip][jj][r] += (double)a32[r] + (double)c32[r];
            }
    }
}

template <int S, int PB, int NJMAX>
__device__ __forceinline__ void k_rho_split_item(int item, const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                   const SplitBlock* __restrict__ sblocks,
                                                   const double* __restrict__ pts, int64_t n_slots,
                                                   const double* __restrict__ shrec, int RS,
                                                   const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                                   const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                                   const unsigned short* __restrict__ Ps, int N, int Nc,
                                                   double* __restrict__ rho_out, int accumulate,
        double (*s_pt)[3], double* s_dyn) {
    constexpr int SUBS = kBlockPts / PB;
    constexpr int NSL = 512 / PB;  // shell lanes per point
    const int bidx = work[item / SUBS];
    const int sub = item % SUBS;
    const BlockDesc B = blocks[bidx];
    const SplitBlock SB = sblocks[bidx];
    const int p0 = sub * PB;
    if (p0 >= B.npts) return;  // uniform: no real point in this part of the block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldnp = SB.ldnp, nch = SB.nchunk, nsh = B.nshell;
    const int RA = ldnp + 8;
    const unsigned short* Pbs = Ps + B.psoff;  // the block's molecule: N = ldp, Nc = ldp rounded up to 8
    const int ldp = B.ldp, ldpc = (B.ldp + 7) / 8 * 8;
    // LDS: plane (S x PB x RA bf16, reused as PB x (ldnp + 4) FP32 and as partial sums) | records | chunk [ldnp/4] | shcol | sorder
    unsigned short* s_a = reinterpret_cast<unsigned short*>(s_dyn);
    const int plane_dbl = (S * PB * RA * 2 + 15) / 16 * 2;  // doubles, 16-byte multiple
    const int region_dbl = max(plane_dbl, 8 * 512);  // also holds the 4 x 1024 partial sums
    double* s_rec = s_dyn + region_dbl;
    int* s_chunk = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_chunk + ldnp / 4;
    int* s_sord = s_shcol + nsh;
    for (int i = tid; i < ldnp / 4; i += 512) s_chunk[i] = (i < nch) ? chunks[SB.chunk_begin + i] : 0;
    for (int k = tid; k < nsh; k += 512) {
        s_shcol[k] = sshell_col[B.shell_begin + k];
        s_sord[k] = bsorder[B.shell_begin + k];
    }
    for (int idx = tid; idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    if (tid < PB) {
        const int64_t slot = B.pt_begin + p0 + tid;
        s_pt[tid][0] = pts[slot];
        s_pt[tid][1] = pts[n_slots + slot];
        s_pt[tid][2] = pts[2 * n_slots + slot];
    }
    {
        uint4* z = reinterpret_cast<uint4*>(s_dyn);
        for (int i = tid; i < plane_dbl / 2; i += 512) z[i] = uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();

    // lanes take consecutive points of one shell: the slice writes of a
    // column land on distinct banks, and a lane group runs one code path
    const int p = tid & (PB - 1);
    const int kk = tid / PB;
    const double px = s_pt[p][0], py = s_pt[p][1], pz = s_pt[p][2];
    // 1. Phi slices of the kept AOs (columns of dropped AOs stay zero)
    for (int j = kk; j < nsh; j += NSL) {
        const int k = s_sord[j];
        double v[5];
        rec_values0(s_rec + k * RS, px, py, pz, v);
        const int c0 = s_shcol[k];
        for_components(((int)s_rec[k * RS + 3]) & 3, [&](int m) {
            unsigned short h[S];
            split_bf16<S>(v[m], h);
#pragma unroll
            for (int s = 0; s < S; ++s) s_a[(size_t)(s * PB + p) * RA + c0 + m] = h[s];
        });
    }
    __syncthreads();

    // 2. X = Phi P_bb; 3. rho and grad rho (FP64): X (FP64, stored over the
    // plane for half of the points at a time) dotted with phi, grad phi.
    // Per round, lanes take point pl of the half and shell lanes kk2.
    double* s_xd = s_dyn;
    constexpr int PH = PB / 2;
    constexpr int NSL2 = 512 / PH;
    const int pl = tid & (PH - 1), kk2 = tid / PH;
    double part[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    auto dots = [&](int round) {
        const int RXD = ldnp + 1;
        const int pt = pl + round * PH;
        const double qx = s_pt[pt][0], qy = s_pt[pt][1], qz = s_pt[pt][2];
        double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
        for (int j = kk2; j < nsh; j += NSL2) {
            const int k = s_sord[j];
            const double* rec = s_rec + k * RS;
            const double x = qx - rec[0], y = qy - rec[1], z = qz - rec[2];
            double R0, R1;
            rec_radial01(rec, x, y, z, R0, R1);
            const int l = ((int)rec[3]) & 3;
            const double* xr = s_xd + pl * RXD + s_shcol[k];
            double X[5] = {xr[0], 0.0, 0.0, 0.0, 0.0};
            if (l >= 1) {
                X[1] = xr[1];
                X[2] = xr[2];
            }
            if (l == 2) {
                X[3] = xr[3];
                X[4] = xr[4];
            }
            shell_dots(l, x, y, z, R0, R1, X, r0, r1, r2, r3);
        }
        part[round][0] = r0;
        part[round][1] = r1;
        part[round][2] = r2;
        part[round][3] = r3;
    };
    constexpr int NP = PB / 16;
    v4d acc[NP][NJMAX];
#pragma unroll
    for (int ip = 0; ip < NP; ++ip)
#pragma unroll
        for (int jj = 0; jj < NJMAX; ++jj) acc[ip][jj] = v4d{0.0, 0.0, 0.0, 0.0};
    const int nJ = ldnp / 16;
    const int nJw = (nJ > wave) ? (nJ - wave + 7) / 8 : 0;  // the host keeps ldnp <= 128 NJMAX
    switch (nJw) {
        case 1: rho_split_tiles<S, PB, 1, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc); break;
        case 2:
            if constexpr (NJMAX >= 2) rho_split_tiles<S, PB, 2, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 3:
            if constexpr (NJMAX >= 3) rho_split_tiles<S, PB, 3, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 4:
            if constexpr (NJMAX >= 4) rho_split_tiles<S, PB, 4, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 5:
            if constexpr (NJMAX >= 5) rho_split_tiles<S, PB, 5, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 6:
            if constexpr (NJMAX >= 6) rho_split_tiles<S, PB, 6, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 7:
            if constexpr (NJMAX >= 7) rho_split_tiles<S, PB, 7, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        case 8:
            if constexpr (NJMAX >= 8) rho_split_tiles<S, PB, 8, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, acc);
            break;
        default: break;  // no column tiles
    }
    __syncthreads();  // every wave has stopped reading the Phi plane
    {
        const int li = lane & 15, lq = lane >> 4;
        const int RXD = ldnp + 1;
#pragma unroll
        for (int round = 0; round < 2; ++round) {
#pragma unroll
            for (int ip = 0; ip < NP; ++ip)
#pragma unroll
                for (int jj = 0; jj < NJMAX; ++jj)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = ip * 16 + 4 * lq + r;
                        if (jj < nJw && (row >= PH) == (round == 1))
                            s_xd[(row - round * PH) * RXD + (wave + 8 * jj) * 16 + li] = acc[ip][jj][r];
                    }
            __syncthreads();  // this round's X is stored
            dots(round);
            __syncthreads();  // X no longer read
        }
    }
    // partial sums of the NSL2 shell lanes of each point, added in a fixed
    // order (the region is free: the last round's barrier has passed)
    double* s_part = s_dyn;  // [4][NSL2][PB]
#pragma unroll
    for (int round = 0; round < 2; ++round)
#pragma unroll
        for (int q = 0; q < 4; ++q) s_part[(q * NSL2 + kk2) * PB + pl + round * PH] = part[round][q];
    __syncthreads();
    if (tid < 4 * PB) {
        const int q = tid / PB;
        const int pp = tid & (PB - 1);
        double sum = 0.0;
        for (int k2 = 0; k2 < NSL2; ++k2) sum += s_part[(q * NSL2 + k2) * PB + pp];
        const int64_t slot = B.pt_begin + p0 + pp;
        const double v = (q == 0) ? sum : 2.0 * sum;
        rho_out[4 * slot + q] = accumulate ? rho_out[4 * slot + q] + v : v;
    }
}

template <int S, int PB, int NJMAX, bool PERSIST>
__global__ __launch_bounds__(512) void k_rho_split(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                   const SplitBlock* __restrict__ sblocks,
                                                   const double* __restrict__ pts, int64_t n_slots,
                                                   const double* __restrict__ shrec, int RS,
                                                   const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                                   const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                                   const unsigned short* __restrict__ Ps, int N, int Nc,
                                                   double* __restrict__ rho_out, int accumulate,
                                                   int* __restrict__ queue, int n_items) {
    __shared__ double s_pt[PB][3];
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];
    if constexpr (PERSIST) {
        __shared__ int s_item;
        for (int iter = 0;; ++iter) {
            const int item = next_item(queue, n_items, iter, &s_item);
            if (item >= n_items) break;
            k_rho_split_item<S, PB, NJMAX>(item, work, blocks, sblocks, pts, n_slots, shrec, RS, bshells, bsorder, sshell_col, chunks, Ps, N, Nc, rho_out, accumulate, s_pt, s_dyn);
        }
    } else {
        k_rho_split_item<S, PB, NJMAX>((int)blockIdx.x, work, blocks, sblocks, pts, n_slots, shrec, RS, bshells, bsorder, sshell_col, chunks, Ps, N, Nc, rho_out, accumulate, s_pt, s_dyn);
    }
}

// ---------------------------------------------------------------------------
// rho pass in block floating point (PDFXC_SPLIT_BFP): every operand of X =
// Phi P_bb is written as 4 bf16 slices of 8 bits that share one exponent per
// point (a row of Phi: E_p) or per P row (F_j; a P row is a P column by
// symmetry): Phi_pi = 2^(E_p - 7) sum_s a_s 2^(-8 s), P_ij = 2^(F_j - 7)
// sum_t b_t 2^(-8 t), every slice an integer in [-128, 128].  A product of
// two slices is then an integer of magnitude at most 2^14; the products of
// one level l = s + t (0..3) share an FP32 accumulator L_l, and over at most
// 512 columns L_0 (<= 2^23) and L_1 (<= 2^24) are exact whatever order the
// hardware adds in; L_2 and L_3 are too unless more than 1024 of their
// products come near 2^14 with one sign, which the third and fourth digits
// of the slices do not.  X = 2^(E_p + F_j - 14) sum_l 2^(-8 l) L_l is formed
// in FP64 once, after the last K step.  The error left is the slicing's,
// 2^-32 of a row's largest value, so phi is staged in FP64 (an FP32 phi
// alone moves E_xc by ~5e-7 Eh on a water octamer).
// ---------------------------------------------------------------------------
constexpr int kBfpSlices = kBfpPlanes;

// chunked columns of the widest block a block-FP class takes: PB = 64 points
// per workgroup up to 256, 32 up to 384, 16 up to 512 (the plane, 8 bytes
// per value, and the accumulators L fit one workgroup)
template <int PB>
constexpr int bfp_max_cols() {
    return PB == 64 ? 256 : PB == 32 ? 384 : 512;
}

// slices of t, |t| < 128: a_0 = rne(t), then the rounded digits of what is
// left, 8 bits each; an integer of at most 8 bits is exact in FP32 with its
// low 16 bits zero, so its bf16 pattern is the FP32 pattern's top half
__device__ inline void split_bfp(double t, unsigned short* out) {
#pragma unroll
    for (int s = 0; s < kBfpSlices; ++s) {
        const double a = __builtin_rint(t);
        out[s] = (unsigned short)(__float_as_uint((float)a) >> 16);
        t = (t - a) * 256.0;  // exact
    }
}

// the exponent E with |x| < 2^E for every |x| <= m (m a float's bits; 0 for m = 0)
__device__ inline int bfp_exponent(unsigned mbits) {
    const float m = __uint_as_float(mbits);
    return (m > 0.0f) ? ilogbf(m) + 1 : 0;
}

// P in the split order as block-FP slices interleaved by 4-column chunk:
// Pb[r][ch][s][e] = slice s of P[perm r][perm (4 ch + e)] 2^(7 - F_r) (a row
// is 4 Nc values, a chunk's 4 slices 32 contiguous bytes), F_r the exponent
// of row r's largest |P|, the N exponents stored after the N rows (one
// workgroup per row; columns c >= N are zero)
__global__ __launch_bounds__(256) void k_psplit_bfp(const double* __restrict__ P, const int* __restrict__ perm, int N,
                                                    int Nc, unsigned short* __restrict__ Pb) {
    int* F = reinterpret_cast<int*>(Pb + (int64_t)kBfpSlices * N * Nc);  // the row exponents follow the rows
    __shared__ float s_max[256];
    const int r = blockIdx.x;
    const double* row = P + (int64_t)perm[r] * N;
    float m = 0.0f;
    for (int c = threadIdx.x; c < N; c += 256) m = fmaxf(m, (float)fabs(row[perm[c]]));
    s_max[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + k]);
        __syncthreads();
    }
    const int E = bfp_exponent(__float_as_uint(s_max[0]));
    if (threadIdx.x == 0) F[r] = E;
    unsigned short* out = Pb + (int64_t)r * kBfpSlices * Nc;
    for (int c = threadIdx.x; c < Nc; c += 256) {
        unsigned short h[kBfpSlices];
        split_bfp((c < N) ? ldexp(row[perm[c]], 7 - E) : 0.0, h);
#pragma unroll
        for (int s = 0; s < kBfpSlices; ++s) out[(c >> 2) * 16 + s * 4 + (c & 3)] = h[s];
    }
}

// level accumulators of one tile: L[l] += sum_{s + t = l} A_s B_t
__device__ inline void bfp_mfma(const bf16x8 (&a)[kBfpSlices], const bf16x8 (&b)[kBfpSlices], f32x4 (&L)[kBfpSlices]) {
    L[0] = mfma_bf16(a[0], b[0], L[0]);
    L[1] = mfma_bf16(a[0], b[1], L[1]);
    L[1] = mfma_bf16(a[1], b[0], L[1]);
    L[2] = mfma_bf16(a[0], b[2], L[2]);
    L[2] = mfma_bf16(a[2], b[0], L[2]);
    L[2] = mfma_bf16(a[1], b[1], L[2]);
    L[3] = mfma_bf16(a[0], b[3], L[3]);
    L[3] = mfma_bf16(a[3], b[0], L[3]);
    L[3] = mfma_bf16(a[1], b[2], L[3]);
    L[3] = mfma_bf16(a[2], b[1], L[3]);
}

// the wave's X tiles: column tiles J = w, w + 8, ... (NJ of them), all PB / 16
// point tiles; A = Phi slices from LDS, B = P slices from global memory (a
// lane's 8 K values are two 4-AO chunks of a P row: per chunk one 32-byte
// read holds the 4 slices)
template <int PB, int NJ, int NJMAX>
__device__ inline void rho_bfp_tiles(int wave, int lane, int ldnp, int nch, const int* s_chunk,
                                     const unsigned short* s_a, const unsigned short* __restrict__ Pb, int N, int Nc,
                                     f32x4 (&L)[PB / 16][NJMAX][kBfpSlices]) {
    constexpr int NP = PB / 16;
    const int li = lane & 15, lq = lane >> 4;
    const int RA = ldnp + 8;
    const unsigned short* prow[NJ];
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        const int c = (wave + 8 * jj) * 16 + li;
        const int ch = c >> 2;
        const int g = s_chunk[ch] * 4 + (c & 3);
        prow[jj] = Pb + (int64_t)((ch < nch && g < N) ? g : 0) * (kBfpSlices * Nc);  // padding columns: any row
    }
    const int nks = ldnp / 32;
    for (int ks = 0; ks < nks; ++ks) {
        const int k0 = s_chunk[ks * 8 + 2 * lq] * 16;
        const int k1 = s_chunk[ks * 8 + 2 * lq + 1] * 16;
        bf16x8 b[NJ][kBfpSlices];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            const bf16x8 lo01 = *reinterpret_cast<const bf16x8*>(prow[jj] + k0);
            const bf16x8 lo23 = *reinterpret_cast<const bf16x8*>(prow[jj] + k0 + 8);
            const bf16x8 hi01 = *reinterpret_cast<const bf16x8*>(prow[jj] + k1);
            const bf16x8 hi23 = *reinterpret_cast<const bf16x8*>(prow[jj] + k1 + 8);
            b[jj][0] = __builtin_shufflevector(lo01, hi01, 0, 1, 2, 3, 8, 9, 10, 11);
            b[jj][1] = __builtin_shufflevector(lo01, hi01, 4, 5, 6, 7, 12, 13, 14, 15);
            b[jj][2] = __builtin_shufflevector(lo23, hi23, 0, 1, 2, 3, 8, 9, 10, 11);
            b[jj][3] = __builtin_shufflevector(lo23, hi23, 4, 5, 6, 7, 12, 13, 14, 15);
        }
#pragma unroll
        for (int ip = 0; ip < NP; ++ip) {
            bf16x8 a[kBfpSlices];
#pragma unroll
            for (int s = 0; s < kBfpSlices; ++s)
                a[s] = *reinterpret_cast<const bf16x8*>(s_a + (size_t)(s * PB + ip * 16 + li) * RA + ks * 32 + lq * 8);
#pragma unroll
            for (int jj = 0; jj < NJ; ++jj) bfp_mfma(a, b[jj], L[ip][jj]);
        }
    }
}

typedef double f64x2 __attribute__((ext_vector_type(2)));

// PB points of a block per workgroup (512 threads; at most bfp_max_cols<PB>
// chunked columns):
//   1. phi of the kept AOs in FP64, staged as FP64 [point][column] and each
//      point's largest |phi| (LDS atomics on the float bits)
//   2. the 4 block-FP slices of every value into the plane [s][point][column],
//      which takes the staging's place (8 bytes per value either way): wave w
//      takes rows w, w + 8, ..., a lane the column pairs 2 lane + 128 q, and
//      every value is read to registers before any slice is written
//   3. the level sums L on MFMA, then X (FP64) over the plane
//   4. rho and grad rho: X dotted with phi, grad phi (shell_dots, radial sums
//      recomputed), the shell lanes' partial sums added in a fixed order
template <int PB, int NJMAX>
__device__ __forceinline__ void k_rho_bfp_item(int item, const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                               const SplitBlock* __restrict__ sblocks,
                                               const double* __restrict__ pts, int64_t n_slots,
                                               const double* __restrict__ shrec, int RS,
                                               const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                               const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                               const unsigned short* __restrict__ Pb,
                                               double* __restrict__ rho_out, int accumulate,
                                               double (*s_pt)[3], unsigned* s_emax, double* s_dyn) {
    constexpr int SUBS = kBlockPts / PB;
    constexpr int NSL = 512 / PB;  // shell lanes per point
    constexpr int NP = PB / 16;
    static_assert(bfp_max_cols<PB>() <= 128 * NJMAX, "a wave's column tiles");
    const int bidx = work[item / SUBS];
    const int sub = item % SUBS;
    const BlockDesc B = blocks[bidx];
    const SplitBlock SB = sblocks[bidx];
    const int p0 = sub * PB;
    if (p0 >= B.npts) return;  // uniform: no real point in this part of the block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldnp = SB.ldnp, nch = SB.nchunk, nsh = B.nshell;
    const int RA = ldnp + 8;
    const unsigned short* Pbs = Pb + B.psoff;  // the block's molecule: N = ldp, Nc = ldp rounded up to 8
    const int ldp = B.ldp, ldpc = (B.ldp + 7) / 8 * 8;
    const int* Fm = reinterpret_cast<const int*>(Pbs + (int64_t)kBfpSlices * ldp * ldpc);
    // LDS: plane (FP64 staging PB x RA; then 4 x PB x RA bf16 slices; then X,
    // PB x (ldnp + 1) FP64; then partial sums) | records | chunks | shcol | sorder
    static_assert(kBfpSlices * 2 == sizeof(double), "the slice planes take the FP64 staging's place");
    double* s_stage = s_dyn;
    unsigned short* s_a = reinterpret_cast<unsigned short*>(s_dyn);
    const int plane_dbl = max(PB * RA, 4 * 512);
    double* s_rec = s_dyn + plane_dbl;
    int* s_chunk = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_chunk + ldnp / 4;
    int* s_sord = s_shcol + nsh;
    for (int i = tid; i < ldnp / 4; i += 512) s_chunk[i] = (i < nch) ? chunks[SB.chunk_begin + i] : 0;
    for (int k = tid; k < nsh; k += 512) {
        s_shcol[k] = sshell_col[B.shell_begin + k];
        s_sord[k] = bsorder[B.shell_begin + k];
    }
    for (int idx = tid; idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    if (tid < PB) {
        const int64_t slot = B.pt_begin + p0 + tid;
        s_pt[tid][0] = pts[slot];
        s_pt[tid][1] = pts[n_slots + slot];
        s_pt[tid][2] = pts[2 * n_slots + slot];
        s_emax[tid] = 0u;
    }
    for (int i = tid; i < PB * RA; i += 512) s_stage[i] = 0.0;  // dropped and padding columns stay zero
    __syncthreads();

    const int p = tid & (PB - 1);
    const int kk = tid / PB;
    const double px = s_pt[p][0], py = s_pt[p][1], pz = s_pt[p][2];
    // 1. phi (FP64) and the point's largest |phi|.  The radial sums of a
    // thread's first NC shells stay in registers for step 4 (same thread-to-
    // shell mapping); the rest are recomputed there.  These classes run 2
    // waves per SIMD, so NC is what fits in 256 VGPRs beside the level
    // accumulators without spilling in the persistent forms (the 64-point
    // class has no room to spare).
    constexpr int NC = (PB == 64) ? 0 : (PB == 32) ? 6 : 8;
    double cR0[NC > 0 ? NC : 1], cR1[NC > 0 ? NC : 1];
    {
        float m = 0.0f;
        auto put = [&](int k, const double* v) {
            const int c0 = s_shcol[k];
            for_components(((int)s_rec[k * RS + 3]) & 3, [&](int q) {
                s_stage[p * RA + c0 + q] = v[q];
                m = fmaxf(m, (float)fabs(v[q]));  // rounding is monotone: 2^E above the float bounds every |v|
            });
        };
#pragma unroll
        for (int jj = 0; jj < NC; ++jj) {
            const int j = kk + jj * NSL;
            cR0[jj] = cR1[jj] = 0.0;
            if (j < nsh) {
                const int k = s_sord[j];
                const double* rec = s_rec + k * RS;
                const double x = px - rec[0], y = py - rec[1], z = pz - rec[2];
                rec_radial01(rec, x, y, z, cR0[jj], cR1[jj]);
                double v[5];
                rec_values0_r(((int)rec[3]) & 3, x, y, z, cR0[jj], v);
                put(k, v);
            }
        }
        for (int j = kk + NC * NSL; j < nsh; j += NSL) {
            const int k = s_sord[j];
            double v[5];
            rec_values0(s_rec + k * RS, px, py, pz, v);
            put(k, v);
        }
        atomicMax(&s_emax[p], __float_as_uint(m));  // non-negative floats order as their bits
    }
    __syncthreads();
    // 2. block-FP slices, each row's exponent E_p
    {
        constexpr int NR = PB / 8;
        constexpr int NQ = bfp_max_cols<PB>() / 128;
        f64x2 v[NR][NQ];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int c = 2 * lane + 128 * q;
                v[r][q] = (c < ldnp) ? *reinterpret_cast<const f64x2*>(s_stage + (wave + 8 * r) * RA + c)
                                     : f64x2{0.0, 0.0};
            }
        __syncthreads();  // the staging is read: the slices overwrite it
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int pp = wave + 8 * r;
            const int sh = 7 - bfp_exponent(s_emax[pp]);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int c = 2 * lane + 128 * q;
                if (c < ldnp) {
                    unsigned short h0[kBfpSlices], h1[kBfpSlices];
                    split_bfp(ldexp(v[r][q].x, sh), h0);
                    split_bfp(ldexp(v[r][q].y, sh), h1);
#pragma unroll
                    for (int s = 0; s < kBfpSlices; ++s)
                        *reinterpret_cast<unsigned*>(s_a + (size_t)(s * PB + pp) * RA + c) =
                            (unsigned)h0[s] | ((unsigned)h1[s] << 16);
                }
            }
        }
    }
    __syncthreads();

    // 3. level sums on MFMA
    f32x4 L[NP][NJMAX][kBfpSlices];
#pragma unroll
    for (int ip = 0; ip < NP; ++ip)
#pragma unroll
        for (int jj = 0; jj < NJMAX; ++jj)
#pragma unroll
            for (int l = 0; l < kBfpSlices; ++l) L[ip][jj][l] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nJ = ldnp / 16;
    const int nJw = (nJ > wave) ? (nJ - wave + 7) / 8 : 0;  // <= NJMAX: the host keeps ldnp <= bfp_max_cols<PB>
    switch (nJw) {
        case 1: rho_bfp_tiles<PB, 1, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, L); break;
        case 2:
            if constexpr (NJMAX >= 2) rho_bfp_tiles<PB, 2, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, L);
            break;
        case 3:
            if constexpr (NJMAX >= 3) rho_bfp_tiles<PB, 3, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, L);
            break;
        case 4:
            if constexpr (NJMAX >= 4) rho_bfp_tiles<PB, 4, NJMAX>(wave, lane, ldnp, nch, s_chunk, s_a, Pbs, ldp, ldpc, L);
            break;
        default: break;  // no column tiles
    }
    __syncthreads();  // every wave has stopped reading the plane
    // X = 2^(E_p + F_j - 14) sum_l 2^(-8 l) L_l, FP64, over the plane [point][ldnp + 1]
    double* s_xd = s_dyn;
    const int RXD = ldnp + 1;
    {
        const int li = lane & 15, lq = lane >> 4;
#pragma unroll
        for (int jj = 0; jj < NJMAX; ++jj) {
            if (jj < nJw) {
                const int c = (wave + 8 * jj) * 16 + li;
                const int ch = c >> 2;
                const int g = (ch < nch) ? s_chunk[ch] * 4 + (c & 3) : 0;
                const int Fj = (g < ldp) ? Fm[g] : 0;
#pragma unroll
                for (int ip = 0; ip < NP; ++ip)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const double x = (double)L[ip][jj][0][r] + ldexp((double)L[ip][jj][1][r], -8) +
                                         ldexp((double)L[ip][jj][2][r], -16) + ldexp((double)L[ip][jj][3][r], -24);
                        const int row = ip * 16 + 4 * lq + r;
                        s_xd[row * RXD + c] = ldexp(x, bfp_exponent(s_emax[row]) + Fj - 14);
                    }
            }
        }
    }
    __syncthreads();  // X is stored
    // 4. rho and grad rho (radial sums kept from step 1 for the first NC shells)
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    auto dot = [&](int k, double R0, double R1) {
        const double* rec = s_rec + k * RS;
        const int l = ((int)rec[3]) & 3;
        const double* xr = s_xd + p * RXD + s_shcol[k];
        double X[5] = {xr[0], 0.0, 0.0, 0.0, 0.0};
        if (l >= 1) {
            X[1] = xr[1];
            X[2] = xr[2];
        }
        if (l == 2) {
            X[3] = xr[3];
            X[4] = xr[4];
        }
        shell_dots(l, px - rec[0], py - rec[1], pz - rec[2], R0, R1, X, r0, r1, r2, r3);
    };
#pragma unroll
    for (int jj = 0; jj < NC; ++jj) {
        const int j = kk + jj * NSL;
        if (j < nsh) dot(s_sord[j], cR0[jj], cR1[jj]);
    }
    for (int j = kk + NC * NSL; j < nsh; j += NSL) {
        const int k = s_sord[j];
        const double* rec = s_rec + k * RS;
        double R0, R1;
        rec_radial01(rec, px - rec[0], py - rec[1], pz - rec[2], R0, R1);
        dot(k, R0, R1);
    }
    __syncthreads();  // X no longer read: the plane holds the partial sums
    double* s_part = s_dyn;  // [4][NSL][PB]
    s_part[(0 * NSL + kk) * PB + p] = r0;
    s_part[(1 * NSL + kk) * PB + p] = r1;
    s_part[(2 * NSL + kk) * PB + p] = r2;
    s_part[(3 * NSL + kk) * PB + p] = r3;
    __syncthreads();
    if (tid < 4 * PB) {
        const int q = tid / PB;
        const int pp = tid & (PB - 1);
        double sum = 0.0;
        for (int k2 = 0; k2 < NSL; ++k2) sum += s_part[(q * NSL + k2) * PB + pp];
        const int64_t slot = B.pt_begin + p0 + pp;
        const double v = (q == 0) ? sum : 2.0 * sum;
        rho_out[4 * slot + q] = accumulate ? rho_out[4 * slot + q] + v : v;
    }
}

template <int PB, int NJMAX, bool PERSIST>
__global__ __launch_bounds__(512) void k_rho_bfp(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                 const SplitBlock* __restrict__ sblocks,
                                                 const double* __restrict__ pts, int64_t n_slots,
                                                 const double* __restrict__ shrec, int RS,
                                                 const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                                 const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                                 const unsigned short* __restrict__ Pb,
                                                 double* __restrict__ rho_out, int accumulate,
                                                 int* __restrict__ queue, int n_items) {
    __shared__ double s_pt[PB][3];
    __shared__ unsigned s_emax[PB];
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];
    if constexpr (PERSIST) {
        __shared__ int s_item;
        for (int iter = 0;; ++iter) {
            const int item = next_item(queue, n_items, iter, &s_item);
            if (item >= n_items) break;
            k_rho_bfp_item<PB, NJMAX>(item, work, blocks, sblocks, pts, n_slots, shrec, RS, bshells, bsorder,
                                      sshell_col, chunks, Pb, rho_out, accumulate, s_pt, s_emax, s_dyn);
        }
    } else {
        k_rho_bfp_item<PB, NJMAX>((int)blockIdx.x, work, blocks, sblocks, pts, n_slots, shrec, RS, bshells, bsorder,
                                  sshell_col, chunks, Pb, rho_out, accumulate, s_pt, s_emax, s_dyn);
    }
}

// ---------------------------------------------------------------------------
// K pass, split precision, over groups of spatially adjacent blocks (siblings
// of the point bisection, 1 to kMaxGroup of them): K_uu += Phi^T V + V^T Phi
// summed over all the group's points on the union u of their AO columns, so
// that the scatter adds one union-sized lower triangle per group instead of
// one triangle per block.  An item takes the lower 2 x 2 tile patches
// [q0, q1) of the union in column-major order (patch column Q, then rows
// P >= Q), so its slabs hold only the union columns [c0, c1) = [32 Q(q0),
// 32 Tp) that those patches touch.  NW waves (16, 12 or 8), each producer
// and consumer: all evaluate phi and V = (w vrho / 2) phi + c . grad phi in
// FP64 for the next 16-point slab of the current block (the shells the
// block's screening kept; the union's other shells are written as zeros) and
// store their bf16 slices as [phi, V][s][column][16 points]; then each runs
// v_mfma_f32_16x16x32_bf16 on the current slab for its NPW patches, the
// 32-long K index covering the slab's 16 points twice, (phi | V) against
// (V | phi), so that one product gives phi_i V_j + V_i phi_j.  Every slab
// sums from zero in FP32, the slice corrections (a + b >= 1, at most 2^-8 of
// the leading term) in one sum and hi x hi in another, and their FP32 sum is
// added into FP64 accumulators per slab: the MFMA's FP32 sums do not round to
// nearest (their errors share a sign), so a sum held in FP32 over a block's
// 8 slabs lost up to ~1.6e-7 of K on coarse grids.  FP64 atomics into the
// dense lower triangle of K at the end of the item.
// ---------------------------------------------------------------------------

// slab address (bf16 units) of the 8-point half h of column c, halves of
// every eighth column swapped so that 16 lanes reading 16 consecutive
// columns hit distinct banks
__device__ inline int kslab_at(int c, int h) { return c * 16 + ((h ^ (c >> 3)) & 1) * 8; }

template <int S, int NPW>
__device__ inline void kmat_split_consume(int pb, int pe, int Tp, int c0, int wcols, int lane,
                                          const unsigned short* s_buf, v4d (&acc)[NPW][4]) {
    const int li = lane & 15, lq = lane >> 4;
    // operand of the 16x16x32 form: lane (li, lq) holds k = 8 lq .. 8 lq + 7;
    // k < 16 are the slab's points of phi (A) or V (B), k >= 16 the other
    const unsigned short* fa = s_buf + (size_t)(lq >> 1) * S * wcols * 16;         // A: phi, then V
    const unsigned short* fb = s_buf + (size_t)(1 - (lq >> 1)) * S * wcols * 16;   // B: V, then phi
    int P = 0, Q = 0;
    {
        int q = pb, rem = Tp;
        while (q >= rem) {  // column-major lower patches: column Q has Tp - Q of them
            q -= rem;
            ++Q;
            --rem;
        }
        P = Q + q;
    }
#pragma unroll
    for (int m = 0; m < NPW; ++m) {
        if (pb + m < pe) {
            bf16x8 aI[2][S], bJ[2][S];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const size_t so = (size_t)s * wcols * 16;
                    aI[h][s] = *reinterpret_cast<const bf16x8*>(fa + so + kslab_at(32 * P + 16 * h + li - c0, lq & 1));
                    bJ[h][s] = *reinterpret_cast<const bf16x8*>(fb + so + kslab_at(32 * Q + 16 * h + li - c0, lq & 1));
                }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int hi = t >> 1, hj = t & 1;
                f32x4 cs = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int q = 1; q < SlicePairs<S>::n; ++q)
                    cs = mfma_bf16(aI[hi][SlicePairs<S>::a[q]], bJ[hj][SlicePairs<S>::b[q]], cs);
                const f32x4 ls = mfma_bf16(aI[hi][0], bJ[hj][0], f32x4{0.f, 0.f, 0.f, 0.f});
                // one FP32 rounding of the slab's sum (2^-24 of it, the size of
                // those inside the MFMA's own sum) for one conversion less
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[m][t][r] += (double)(ls[r] + cs[r]);
            }
        }
        if (++P == Tp) {
            ++Q;
            P = Q;
        }
    }
}

// the FP16 form: phi and V carry their own exponents per slab column (phi's
// [0, wcols), V's [wcols, 2 wcols) of s_ex), so the two terms of
// K_PQ = Phi_P^T V_Q + V_P^T Phi_Q run as separate v_mfma_f32_16x16x16_f16
// products over the slab's 16 points (lane (li, lq) holds points 4 lq ..
// 4 lq + 3) and are scaled back in FP64 apart: element (i, j) of the first
// carries 2^(30 - eP_i - eV_j), of the second 2^(30 - eV_i - eP_j)
template <int NPW>
__device__ inline void kmat_f16_consume(int pb, int pe, int Tp, int c0, int wcols, int lane,
                                        const unsigned short* s_buf, const int* s_ex, v4d (&acc)[NPW][4]) {
    constexpr int S = 2;
    const int li = lane & 15, lq = lane >> 4;
    const unsigned short* fphi = s_buf;
    const unsigned short* fv = s_buf + (size_t)S * wcols * 16;
    const int* exP = s_ex;
    const int* exV = s_ex + wcols;
    const int hsel = lq >> 1, hoff = (lq & 1) * 4;
    int P = 0, Q = 0;
    {
        int q = pb, rem = Tp;
        while (q >= rem) {
            q -= rem;
            ++Q;
            --rem;
        }
        P = Q + q;
    }
#pragma unroll
    for (int m = 0; m < NPW; ++m) {
        if (pb + m < pe) {
            f16x4 pI[2][S], vI[2][S], pJ[2][S], vJ[2][S];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ai = kslab_at(32 * P + 16 * h + li - c0, hsel) + hoff;
                const int aj = kslab_at(32 * Q + 16 * h + li - c0, hsel) + hoff;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const size_t so = (size_t)s * wcols * 16;
                    pI[h][s] = *reinterpret_cast<const f16x4*>(fphi + so + ai);
                    vI[h][s] = *reinterpret_cast<const f16x4*>(fv + so + ai);
                    pJ[h][s] = *reinterpret_cast<const f16x4*>(fphi + so + aj);
                    vJ[h][s] = *reinterpret_cast<const f16x4*>(fv + so + aj);
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int hi = t >> 1, hj = t & 1;
                const f32x4 z = f32x4{0.f, 0.f, 0.f, 0.f};
                const f32x4 la = mfma_f16_k16(pI[hi][0], vJ[hj][0], z);
                const f32x4 ca = mfma_f16_k16(pI[hi][1], vJ[hj][0], mfma_f16_k16(pI[hi][0], vJ[hj][1], z));
                const f32x4 lb = mfma_f16_k16(vI[hi][0], pJ[hj][0], z);
                const f32x4 cb = mfma_f16_k16(vI[hi][1], pJ[hj][0], mfma_f16_k16(vI[hi][0], pJ[hj][1], z));
                const int cj = 32 * Q + 16 * hj + li - c0;
                const int ePj = exP[cj] - 30, eVj = exV[cj] - 30;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ci = 32 * P + 16 * hi + 4 * lq + r - c0;
                    acc[m][t][r] += ldexp((double)la[r] + (double)ca[r], exP[ci] + eVj) +
                                    ldexp((double)lb[r] + (double)cb[r], exV[ci] + ePj);
                }
            }
        }
        if (++P == Tp) {
            ++Q;
            P = Q;
        }
    }
}

template <int S, bool F16, int NW, int NPW>
__device__ __forceinline__ void k_kmat_split_item(int item, const KGWork* __restrict__ work, const KGroup* __restrict__ groups,
                                                    const BlockDesc* __restrict__ blocks, const int* __restrict__ gblocks,
                                                    const double* __restrict__ pts, int64_t n_slots,
                                                    const double* __restrict__ shrec, int RS,
                                                    const int* __restrict__ gshells, const int* __restrict__ gshell_col,
                                                    const int* __restrict__ gcols, const int* __restrict__ gsorder,
                                                    const unsigned* __restrict__ gmasks,
                                                    const double* __restrict__ coef, double* __restrict__ K,
        double (*s_pt)[kBlockPts][3], double (*s_cf)[kBlockPts][4], unsigned (*s_msk)[kMaxMaskWords],
        int* s_first, int* s_blk, int* s_pcount, double* s_dyn) {
    const KGWork W = work[item];
    const KGroup G = groups[W.group];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int NT = NW * 64;
    const int c0 = W.c0;
    const int wcols = W.c1 - W.c0;  // multiple of 32
    const int Tp = W.c1 / 32;
    const int nsh = G.nshell - W.k0;
    const int nbuf = W.nbuf;
    const int BUF = 2 * S * wcols * 16;
    // LDS: slab buffers (bf16 / f16) | records | cols [wcols] | shcol, shid,
    // shk [nsh] | FP16: the slab buffers' column exponents [nbuf][phi, V][wcols]
    unsigned short* s_slab = reinterpret_cast<unsigned short*>(s_dyn);
    const int slab_dbl = (nbuf * BUF * 2 + 15) / 16 * 2;
    double* s_rec = s_dyn + slab_dbl;
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + wcols;
    int* s_shid = s_shcol + nsh;
    int* s_shk = s_shid + nsh;
    int* s_ex = s_shk + nsh;
    if constexpr (F16)
        for (int i = tid; i < nbuf * 2 * wcols; i += NT) s_ex[i] = 0;  // columns past the union: zero slices
    for (int c = tid; c < wcols; c += NT) s_cols[c] = (c0 + c < G.nao) ? gcols[G.col_begin + c0 + c] : -1;
    if (tid == 0) s_pcount[0] = s_pcount[1] = 0;
    if (tid < G.nsub) {  // the group's blocks and their slab counts, one lane each (a prefix after the barrier)
        const int b = gblocks[G.sub_begin + tid];
        s_blk[tid] = b;
        s_first[tid + 1] = (blocks[b].npts + 15) / 16;
    }
    if (wave == NW / 2) {  // the item's shells (union shells k0 ..), costliest (l, nprim) last
        int base = 0;
        for (int i0 = 0; i0 < G.nshell; i0 += 64) {
            const int i = i0 + lane;
            const int sh = (i < G.nshell) ? gsorder[G.shell_begin + i] : -1;
            const bool keep = sh >= W.k0;
            const unsigned long long msk = __ballot(keep);
            const int pos = base + __popcll(msk & ((1ull << lane) - 1ull));
            if (keep) {
                s_shid[pos] = gshells[G.shell_begin + sh];
                s_shcol[pos] = gshell_col[G.shell_begin + sh] - c0;  // may be < 0 for a shell across c0
                s_shk[pos] = sh;
            }
            base += __popcll(msk);
        }
    }
    {
        uint4* z = reinterpret_cast<uint4*>(s_dyn);  // columns past the union stay zero
        for (int i = tid; i < slab_dbl / 2; i += NT) z[i] = uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();
    if (tid == 0) {  // first slab of each block of the group, then the slab count
        s_first[0] = 0;
        for (int j = 0; j < G.nsub; ++j) s_first[j + 1] += s_first[j];
    }
    for (int idx = tid; idx < nsh * RS; idx += NT) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)s_shid[k] * RS + (idx - k * RS)];
    }
    // points, point terms and kept-shell masks of block j of the group, into
    // the buffers of j's parity
    auto load_block = [&](int j) {
        const int par = j & 1;
        const BlockDesc& B = blocks[s_blk[j]];
        for (int q = tid; q < kBlockPts; q += NT) {
            const int64_t slot = B.pt_begin + q;
            s_pt[par][q][0] = pts[slot];
            s_pt[par][q][1] = pts[n_slots + slot];
            s_pt[par][q][2] = pts[2 * n_slots + slot];
            s_cf[par][q][0] = coef[4 * slot + 0];
            s_cf[par][q][1] = coef[4 * slot + 1];
            s_cf[par][q][2] = coef[4 * slot + 2];
            s_cf[par][q][3] = coef[4 * slot + 3];
        }
        for (int w = tid; w < G.nword; w += NT) s_msk[par][w] = gmasks[G.mask_begin + j * G.nword + w];
    };
    load_block(0);
    __syncthreads();
    const int nslab = s_first[G.nsub];

    // phi and V of global slab gs, FP64, sliced into the slab buffer
    auto produce = [&](int gs) {
        int j = 0;
        while (gs >= s_first[j + 1]) ++j;
        const int par = j & 1;
        const int p0 = (gs - s_first[j]) * 16;
        unsigned short* phi = s_slab + (nbuf == 2 ? (gs & 1) * BUF : 0);
        unsigned short* vv = phi + S * wcols * 16;
        int* ex = s_ex + (nbuf == 2 ? (gs & 1) * 2 * wcols : 0);
        // (shell, point) items in chunks of 64 (four shells x 16 points) taken
        // from an LDS counter, shells from the costliest (l, nprim) down: a
        // wave that consumes fewer patches produces more
        int* pc = s_pcount + (gs & 1);
        for (;;) {
            int chunk = 0;
            if (lane == 0) chunk = atomicAdd(pc, 1);
            const int r0 = __shfl(chunk, 0) * 64;
            if (r0 >= nsh * 16) break;  // wave-uniform
            const int it = r0 + lane;
            if (it >= nsh * 16) continue;  // whole 16-lane groups
            const int p = it & 15;
            const int k = nsh - 1 - (it >> 4);
            const int pp = p0 + p;
            const int kk = s_shk[k];
            const bool kept = (s_msk[par][kk >> 5] >> (kk & 31)) & 1u;
            const int cs = s_shcol[k];
            if (!kept) {  // a shell this block's screening dropped: zero columns
                const int l = ((int)s_rec[k * RS + 3]) & 3;
                for (int m = 0; m < 2 * l + 1; ++m) {
                    const int c = cs + m;
                    if (c < 0) continue;
                    if (c >= wcols) break;
                    if constexpr (F16)
                        if (p == 0) ex[c] = ex[wcols + c] = 0;
                    const int off = kslab_at(c, p >> 3) + (p & 7);
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        phi[(size_t)s * wcols * 16 + off] = 0;
                        vv[(size_t)s * wcols * 16 + off] = 0;
                    }
                }
                continue;
            }
            // radial sums once, then each component's phi and V in closed form
            const double* rec = s_rec + k * RS;
            const double x = s_pt[par][pp][0] - rec[0], y = s_pt[par][pp][1] - rec[1], z = s_pt[par][pp][2] - rec[2];
            double R0, R1;
            rec_radial01(rec, x, y, z, R0, R1);
            shell_phi_v(((int)rec[3]) & 3, x, y, z, R0, R1, s_cf[par][pp][0], s_cf[par][pp][1], s_cf[par][pp][2],
                        s_cf[par][pp][3], [&](int m, double v, double V) {
                            const int c = cs + m;
                            if (c < 0 || c >= wcols) return;  // a shell across the item's first column, or past its last
                            const int off = kslab_at(c, p >> 3) + (p & 7);
                            unsigned short h[S];
                            if constexpr (F16) {
                                // an exponent per slab column for phi and one for
                                // V: the largest of the 16 points (the 16 lanes of
                                // this shell), so every value scales into |t| < 2^15
                                float mp = (float)fabs(v), mv = (float)fabs(V);
#pragma unroll
                                for (int x = 8; x > 0; x >>= 1) {
                                    mp = fmaxf(mp, __shfl_xor(mp, x, 16));
                                    mv = fmaxf(mv, __shfl_xor(mv, x, 16));
                                }
                                const int ep = bfp_exponent(__float_as_uint(mp));
                                const int ev = bfp_exponent(__float_as_uint(mv));
                                if (p == 0) {
                                    ex[c] = ep;
                                    ex[wcols + c] = ev;
                                }
                                split_f16(ldexp(v, 15 - ep), h);
#pragma unroll
                                for (int s = 0; s < S; ++s) phi[(size_t)s * wcols * 16 + off] = h[s];
                                split_f16(ldexp(V, 15 - ev), h);
#pragma unroll
                                for (int s = 0; s < S; ++s) vv[(size_t)s * wcols * 16 + off] = h[s];
                            } else {
                                split_bf16<S>(v, h);
#pragma unroll
                                for (int s = 0; s < S; ++s) phi[(size_t)s * wcols * 16 + off] = h[s];
                                split_bf16<S>(V, h);
#pragma unroll
                                for (int s = 0; s < S; ++s) vv[(size_t)s * wcols * 16 + off] = h[s];
                            }
                        });
        }
    };

    // the item's patches over the waves, NPW each; every wave runs the same
    // barrier sequence
    const bool has = W.q0 + wave * NPW < W.q1;
    const int pb = has ? W.q0 + wave * NPW : W.q1 - 1;
    const int pe = has ? min(pb + NPW, W.q1) : pb;
    v4d acc[NPW][4];
#pragma unroll
    for (int m = 0; m < NPW; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[m][t] = v4d{0.0, 0.0, 0.0, 0.0};
    produce(0);
    __syncthreads();  // the first slab is in
    int jn = 1;       // next block of the group to load
    for (int gs = 0; gs < nslab; ++gs) {
        // slab gs's production counter is done with (an earlier barrier): zero it for slab gs + 2
        if (tid == 0) s_pcount[gs & 1] = 0;
        if (gs + 1 < nslab && jn < G.nsub && gs + 1 == s_first[jn]) {
            // the next slab starts block jn: its points into the other parity's
            // buffers (their last reader, block jn - 2, finished production
            // before an earlier barrier)
            load_block(jn);
            ++jn;
            __syncthreads();
        }
        // a wave consumes its patches of slab gs, then takes production of
        // slab gs + 1 (the other buffer) while any is left
        if constexpr (F16)
            kmat_f16_consume<NPW>(pb, pe, Tp, c0, wcols, lane, s_slab + (nbuf == 2 ? (gs & 1) * BUF : 0),
                                  s_ex + (nbuf == 2 ? (gs & 1) * 2 * wcols : 0), acc);
        else
            kmat_split_consume<S, NPW>(pb, pe, Tp, c0, wcols, lane, s_slab + (nbuf == 2 ? (gs & 1) * BUF : 0), acc);
        if (nbuf == 2 && gs + 1 < nslab) produce(gs + 1);
        __syncthreads();  // slab gs consumed (and, two buffers, slab gs + 1 in)
        if (nbuf == 1) {
            if (gs + 1 < nslab) produce(gs + 1);
            __syncthreads();  // single buffer: the next slab is in
        }
    }
    // scatter: register r of lane l is tile element (4 (l / 16) + r, l % 16)
    {
        const int li = lane & 15, lq = lane >> 4;
        double* Km = K + G.poff;
        const int N = G.ldp;
        int P = 0, Q = 0;
        {
            int q = pb, rem = Tp;
            while (q >= rem) {
                q -= rem;
                ++Q;
                --rem;
            }
            P = Q + q;
        }
#pragma unroll
        for (int m = 0; m < NPW; ++m) {
            if (pb + m < pe) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int lc = 32 * Q + 16 * (t & 1) + li;
                    const int gj = s_cols[lc - c0];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int lr = 32 * P + 16 * (t >> 1) + 4 * lq + r;
                        const int gi = s_cols[lr - c0];
                        if (gi >= 0 && gj >= 0 && lr >= lc) atomicAdd(Km + (int64_t)gi * N + gj, acc[m][t][r]);
                    }
                }
            }
            if (++P == Tp) {
                ++Q;
                P = Q;
            }
        }
    }
}

template <int S, bool F16, int NW, int NPW, bool PERSIST>
__global__ __launch_bounds__(NW * 64) void k_kmat_split(const KGWork* __restrict__ work, const KGroup* __restrict__ groups,
                                                    const BlockDesc* __restrict__ blocks, const int* __restrict__ gblocks,
                                                    const double* __restrict__ pts, int64_t n_slots,
                                                    const double* __restrict__ shrec, int RS,
                                                    const int* __restrict__ gshells, const int* __restrict__ gshell_col,
                                                    const int* __restrict__ gcols, const int* __restrict__ gsorder,
                                                    const unsigned* __restrict__ gmasks,
                                                    const double* __restrict__ coef, double* __restrict__ K,
                                                    int* __restrict__ queue, int n_items) {
    __shared__ double s_pt[2][kBlockPts][3];
    __shared__ double s_cf[2][kBlockPts][4];
    __shared__ unsigned s_msk[2][kMaxMaskWords];
    __shared__ int s_first[kMaxGroup + 1];  // first slab of each block of the group, then the slab count
    __shared__ int s_blk[kMaxGroup];
    __shared__ int s_pcount[2];  // production chunks taken, per slab parity
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];
    if constexpr (PERSIST) {
        __shared__ int s_item;
        for (int iter = 0;; ++iter) {
            const int item = next_item(queue, n_items, iter, &s_item);
            if (item >= n_items) break;
            k_kmat_split_item<S, F16, NW, NPW>(item, work, groups, blocks, gblocks, pts, n_slots, shrec, RS, gshells, gshell_col, gcols, gsorder, gmasks, coef, K, s_pt, s_cf, s_msk, s_first, s_blk, s_pcount, s_dyn);
        }
    } else {
        k_kmat_split_item<S, F16, NW, NPW>((int)blockIdx.x, work, groups, blocks, gblocks, pts, n_slots, shrec, RS, gshells, gshell_col, gcols, gsorder, gmasks, coef, K, s_pt, s_cf, s_msk, s_first, s_blk, s_pcount, s_dyn);
    }
}

typedef short s16x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x16_bf16: lane l holds A[l % 16][4 (l / 16) + e],
// B[4 (l / 16) + e][l % 16], e = 0..3; the result as for the 16x16x32 form
// (the self test below checks it beside the 16x16x32 form the kernels use)
__device__ inline f32x4 mfma_bf16_k16(s16x4 a, s16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// layout self test: A(i, k) = ((3 i + 5 k) % 11) - 5, B(k, j) = ((7 k + 2 j) % 13) - 6;
// out[0..255] from v_mfma_f32_16x16x32_bf16 (k < 32), out[256..511] from
// v_mfma_f32_16x16x16_bf16 (k < 16), out[512..767] from
// v_mfma_f32_16x16x32_f16 (k < 32), out[768..1023] from
// v_mfma_f32_16x16x16_f16 (k < 16)
__global__ void k_mfma_split_selftest(float* out) {
    const int l = threadIdx.x;
    const int li = l & 15, lq = l >> 4;
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * lq + e;
        a[e] = (__bf16)(float)(((3 * li + 5 * k) % 11) - 5);
        b[e] = (__bf16)(float)(((7 * k + 2 * li) % 13) - 6);
    }
    const f32x4 d = mfma_bf16(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(4 * lq + r) * 16 + li] = d[r];
    s16x4 a4, b4;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = 4 * lq + e;
        a4[e] = (short)bf16_bits((__bf16)(float)(((3 * li + 5 * k) % 11) - 5));
        b4[e] = (short)bf16_bits((__bf16)(float)(((7 * k + 2 * li) % 13) - 6));
    }
    const f32x4 d4 = mfma_bf16_k16(a4, b4, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; ++r) out[256 + (4 * lq + r) * 16 + li] = d4[r];
    f16x8 ah, bh;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 8 * lq + e;
        ah[e] = (_Float16)(float)(((3 * li + 5 * k) % 11) - 5);
        bh[e] = (_Float16)(float)(((7 * k + 2 * li) % 13) - 6);
    }
    const f32x4 dh = mfma_f16(ah, bh, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; ++r) out[512 + (4 * lq + r) * 16 + li] = dh[r];
    f16x4 ah4, bh4;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = 4 * lq + e;
        ah4[e] = (_Float16)(float)(((3 * li + 5 * k) % 11) - 5);
        bh4[e] = (_Float16)(float)(((7 * k + 2 * li) % 13) - 6);
    }
    const f32x4 dh4 = mfma_f16_k16(ah4, bh4, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; ++r) out[768 + (4 * lq + r) * 16 + li] = dh4[r];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return PDFXC_OK;
}

int split_points_per_workgroup(int cls) { return cls == 0 ? 64 : cls == 1 ? 32 : 16; }

template <int S>
static int psplit_one(Context& c) {
    if (c.psplit_jobs.empty()) {
        const int64_t n = (int64_t)c.n_aos * c.split_nc;
        hipLaunchKernelGGL((k_psplit<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.d_P,
                           c.d_perm, c.n_aos, c.split_nc, c.d_Ps);
        return check("k_psplit");
    }
    for (const Context::PsplitJob& j : c.psplit_jobs) {  // a batch: each molecule's P into its slices
        const int nc = (j.n + 7) / 8 * 8;
        const int64_t n = (int64_t)j.n * nc;
        hipLaunchKernelGGL((k_psplit<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream,
                           c.d_P + j.poff, c.d_perm + j.perm_off, j.n, nc, c.d_Ps + j.psoff);
        if (int rc = check("k_psplit")) return rc;
    }
    return PDFXC_OK;
}

template <int S, int PB, int NJMAX>
static int rho_split_one(Context& c, const LaunchRange& r) {
    const int items = r.count * (kBlockPts / PB);
    unsigned grid = 0;
    int* q = launch_queue(c, kQueueRho, items, resident_per_cu(&k_rho_split<S, PB, NJMAX, true>, 512, r.lds), &grid);
    if (q)
        hipLaunchKernelGGL((k_rho_split<S, PB, NJMAX, true>), dim3(grid), dim3(512), r.lds, c.stream,
                           c.d_srho_work + r.offset, c.d_blocks, c.d_sblocks, c.d_pts, c.n_slots, c.d_shrec,
                           c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col, c.d_chunks, c.d_Ps,
                           c.n_aos, c.split_nc, c.d_rho, c.rho_accumulate, q, items);
    else
        hipLaunchKernelGGL((k_rho_split<S, PB, NJMAX, false>), dim3(grid), dim3(512), r.lds, c.stream,
                           c.d_srho_work + r.offset, c.d_blocks, c.d_sblocks, c.d_pts, c.n_slots, c.d_shrec,
                           c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col, c.d_chunks, c.d_Ps,
                           c.n_aos, c.split_nc, c.d_rho, c.rho_accumulate, q, items);
    return check("k_rho_split");
}

int bfp_points_per_workgroup(int cls) { return cls == 0 ? 64 : cls == 1 ? 32 : 16; }

int bfp_max_columns(int cls) { return cls == 0 ? bfp_max_cols<64>() : cls == 1 ? bfp_max_cols<32>() : bfp_max_cols<16>(); }

// dynamic LDS a block-FP rho workgroup of class cls may take: 160 KiB less the
// kernel's static LDS, read from the code object
int bfp_max_dyn_lds(int cls, size_t* out) {
    hipFuncAttributes a;
    const void* f = cls == 0   ? reinterpret_cast<const void*>(&k_rho_bfp<64, 2, false>)
                    : cls == 1 ? reinterpret_cast<const void*>(&k_rho_bfp<32, 3, true>)
                               : reinterpret_cast<const void*>(&k_rho_bfp<16, 4, true>);
    const hipError_t e = hipFuncGetAttributes(&a, f);
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes(k_rho_bfp)");
    *out = (size_t)160 * 1024 - a.sharedSizeBytes;
    return PDFXC_OK;
}

// (the 64-point class runs one workgroup per item in every tuning: its
// persistent loop would spill registers)
template <int PB, int NJMAX>
static int rho_bfp_one(Context& c, const LaunchRange& r) {
    const int items = r.count * (kBlockPts / PB);
    unsigned grid = (unsigned)items;
    if constexpr (PB < 64) {
        int* q = launch_queue(c, kQueueRho, items, resident_per_cu(&k_rho_bfp<PB, NJMAX, true>, 512, r.lds), &grid);
        if (q) {
            hipLaunchKernelGGL((k_rho_bfp<PB, NJMAX, true>), dim3(grid), dim3(512), r.lds, c.stream,
                               c.d_srho_work + r.offset, c.d_blocks, c.d_sblocks, c.d_pts, c.n_slots, c.d_shrec,
                               c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col, c.d_chunks, c.d_Ps,
                               c.d_rho, c.rho_accumulate, q, items);
            return check("k_rho_bfp");
        }
    }
    hipLaunchKernelGGL((k_rho_bfp<PB, NJMAX, false>), dim3(grid), dim3(512), r.lds, c.stream,
                       c.d_srho_work + r.offset, c.d_blocks, c.d_sblocks, c.d_pts, c.n_slots, c.d_shrec,
                       c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col, c.d_chunks, c.d_Ps,
                       c.d_rho, c.rho_accumulate, nullptr, items);
    return check("k_rho_bfp");
}

// block-FP rho pass: P's slices and row exponents (per molecule in a batch),
// the block-FP classes, and the FP64 k_rho for the blocks no block-FP class
// takes (wider than 512 chunked columns, or too many shells for the LDS)
static int bfp_build(Context& c) {
    int rc = PDFXC_OK;
    if (c.psplit_jobs.empty()) {
        hipLaunchKernelGGL(k_psplit_bfp, dim3((unsigned)c.n_aos), dim3(256), 0, c.stream, c.d_P, c.d_perm, c.n_aos,
                           c.split_nc, c.d_Ps);
        rc = check("k_psplit_bfp");
    }
    for (const Context::PsplitJob& j : c.psplit_jobs) {
        if (rc != PDFXC_OK) break;
        hipLaunchKernelGGL(k_psplit_bfp, dim3((unsigned)j.n), dim3(256), 0, c.stream, c.d_P + j.poff,
                           c.d_perm + j.perm_off, j.n, (j.n + 7) / 8 * 8, c.d_Ps + j.psoff);
        rc = check("k_psplit_bfp");
    }
    for (const LaunchRange& r : c.srho_launches) {
        if (rc != PDFXC_OK) break;
        if (r.cls >= kRhoFp64Class) {
            int* const save = c.d_rho_work;
            c.d_rho_work = c.d_srho_work;  // the FP64 kernel over these blocks
            LaunchRange f = r;
            f.cls = r.cls - kRhoFp64Class;
            rc = launch_rho(c, f);
            c.d_rho_work = save;
            continue;
        }
        switch (r.cls) {
            case 0: rc = rho_bfp_one<64, 2>(c, r); break;
            case 1: rc = rho_bfp_one<32, 3>(c, r); break;
            case 2: rc = rho_bfp_one<16, 4>(c, r); break;
            default: set_error("k_rho_bfp: bad class"); return PDFXC_ERR_ARGUMENT;
        }
    }
    return rc;
}

template <int S, bool F16, int NW, int NPW>
static int kmat_split_launch(Context& c, const LaunchRange& r) {
    unsigned grid = 0;
    int* q = launch_queue(c, kQueueK, r.count, resident_per_cu(&k_kmat_split<S, F16, NW, NPW, true>, NW * 64, r.lds),
                          &grid);
    if (q)
        hipLaunchKernelGGL((k_kmat_split<S, F16, NW, NPW, true>), dim3(grid), dim3(NW * 64), r.lds, c.stream,
                           c.d_kg_work + r.offset, c.d_kgroups, c.d_blocks, c.d_kg_blocks, c.d_pts, c.n_slots,
                           c.d_shrec, c.rec_stride, c.d_kg_shells, c.d_kg_shell_col, c.d_kg_cols, c.d_kg_sorder,
                           c.d_kg_masks, c.d_coef, c.d_K, q, r.count);
    else
        hipLaunchKernelGGL((k_kmat_split<S, F16, NW, NPW, false>), dim3(grid), dim3(NW * 64), r.lds, c.stream,
                           c.d_kg_work + r.offset, c.d_kgroups, c.d_blocks, c.d_kg_blocks, c.d_pts, c.n_slots,
                           c.d_shrec, c.rec_stride, c.d_kg_shells, c.d_kg_shell_col, c.d_kg_cols, c.d_kg_sorder,
                           c.d_kg_masks, c.d_coef, c.d_K, q, r.count);
    return check("k_kmat_split");
}

// the workgroup shape of a launch class: items of up to 16 patches, 16 waves
// of one patch (127 VGPRs); up to 24, 12 waves of two (167); up to 32, 8
// waves of four (247)
template <int S, bool F16>
static int kmat_split_one(Context& c, const LaunchRange& r) {
    switch (r.cls) {
        case 0: return kmat_split_launch<S, F16, 16, 1>(c, r);
        case 1: return kmat_split_launch<S, F16, 12, 2>(c, r);
        case 2: return kmat_split_launch<S, F16, 8, 4>(c, r);
        default: set_error("k_kmat_split: bad class"); return PDFXC_ERR_ARGUMENT;
    }
}

template <int S>
static int split_build(Context& c) {
    int rc = psplit_one<S>(c);
    for (const LaunchRange& r : c.srho_launches) {
        if (rc != PDFXC_OK) break;
        switch (r.cls) {
            case 0: rc = rho_split_one<S, 64, 2>(c, r); break;
            case 1: rc = rho_split_one<S, 32, 4>(c, r); break;
            case 2: rc = rho_split_one<S, 16, 8>(c, r); break;
            default: set_error("k_rho_split: bad class"); return PDFXC_ERR_ARGUMENT;
        }
    }
    return rc;
}

int launch_split_rho(Context& c) {
    switch (c.opts.precision) {
        case PDFXC_SPLIT_BF16: return split_build<2>(c);
        case PDFXC_SPLIT_BF16X3: return split_build<3>(c);
        case PDFXC_SPLIT_BFP: return bfp_build(c);
        default: set_error("split build: bad precision"); return PDFXC_ERR_ARGUMENT;
    }
}

int launch_split_kmat(Context& c) {
    int rc = PDFXC_OK;
    for (const LaunchRange& r : c.sk_launches) {
        if (rc != PDFXC_OK) break;
        if (r.count <= 0) continue;
        if (prec_k_f16(c.opts.precision))
            rc = kmat_split_one<2, true>(c, r);
        else
            rc = (prec_slices(c.opts.precision) == 3) ? kmat_split_one<3, false>(c, r) : kmat_split_one<2, false>(c, r);
    }
    return rc;
}

// dynamic LDS a k_kmat_split<S> workgroup may take: the 160 KiB of a gfx950
// workgroup less the kernel's static LDS, read from the code object (the FP16
// form has the static LDS of the bf16 one)
int split_kmat_max_dyn_lds(int S, size_t* out) {
    hipFuncAttributes a;
    const hipError_t e = (S == 3) ? hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_kmat_split<3, false, 16, 1, false>))
                                  : hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_kmat_split<2, false, 16, 1, false>));
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes(k_kmat_split)");
    *out = (size_t)160 * 1024 - a.sharedSizeBytes;
    return PDFXC_OK;
}

int launch_mfma_split_selftest(float* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_split_selftest, dim3(1), dim3(64), 0, s, d_out);
    return check("k_mfma_split_selftest");
}

}  // namespace pdfxc
