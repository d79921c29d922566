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
        const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
        for (int m = 0; m < nc; ++m) {
            unsigned short h[S];
            split_bf16<S>(v[m], h);
#pragma unroll
            for (int s = 0; s < S; ++s) s_a[(size_t)(s * PB + p) * RA + c0 + m] = h[s];
        }
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
// the leading term) in one sum and hi x hi in another, and both are added
// into FP64 accumulators per slab: the MFMA's FP32 sums do not round to
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
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[m][t][r] += (double)ls[r] + (double)cs[r];
            }
        }
        if (++P == Tp) {
            ++Q;
            P = Q;
        }
    }
}

template <int S, int NW, int NPW>
__device__ __forceinline__ void k_kmat_split_item(int item, const KGWork* __restrict__ work, const KGroup* __restrict__ groups,
                                                    const BlockDesc* __restrict__ blocks, const int* __restrict__ gblocks,
                                                    const double* __restrict__ pts, int64_t n_slots,
                                                    const double* __restrict__ shrec, int RS,
                                                    const int* __restrict__ gshells, const int* __restrict__ gshell_col,
                                                    const int* __restrict__ gcols, const int* __restrict__ gsorder,
                                                    const unsigned* __restrict__ gmasks,
                                                    const double* __restrict__ coef, double* __restrict__ K,
        double (*s_pt)[kBlockPts][3], double (*s_cf)[kBlockPts][4], unsigned (*s_msk)[kMaxMaskWords],
        int* s_first, int* s_blk, double* s_dyn) {
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
    // LDS: slab buffers (bf16) | records | cols [wcols] | shcol, shid, shk [nsh]
    unsigned short* s_slab = reinterpret_cast<unsigned short*>(s_dyn);
    const int slab_dbl = (nbuf * BUF * 2 + 15) / 16 * 2;
    double* s_rec = s_dyn + slab_dbl;
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + wcols;
    int* s_shid = s_shcol + nsh;
    int* s_shk = s_shid + nsh;
    for (int c = tid; c < wcols; c += NT) s_cols[c] = (c0 + c < G.nao) ? gcols[G.col_begin + c0 + c] : -1;
    if (tid == 0) {
        int f = 0;
        for (int j = 0; j < G.nsub; ++j) {
            const int b = gblocks[G.sub_begin + j];
            s_blk[j] = b;
            s_first[j] = f;
            f += (blocks[b].npts + 15) / 16;
        }
        s_first[G.nsub] = f;
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
        // (shell, point) items in rounds of NT: shells from the costliest
        // (l, nprim) down, and every other round's waves in reverse order, so
        // a wave's rounds mix dear and cheap shells
        for (int r0 = 0; r0 < nsh * 16; r0 += NT) {
            const int ww = ((r0 / NT) & 1) ? (NW - 1 - wave) : wave;
            const int it = r0 + ww * 64 + lane;
            if (it >= nsh * 16) continue;
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
                            split_bf16<S>(v, h);
#pragma unroll
                            for (int s = 0; s < S; ++s) phi[(size_t)s * wcols * 16 + off] = h[s];
                            split_bf16<S>(V, h);
#pragma unroll
                            for (int s = 0; s < S; ++s) vv[(size_t)s * wcols * 16 + off] = h[s];
                        });
        }
    };

    // the item's patches over the waves, NPW each; every wave runs the same
    // production and barrier sequence
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
        if (gs + 1 < nslab && jn < G.nsub && gs + 1 == s_first[jn]) {
            // the next slab starts block jn: its points into the other parity's
            // buffers (their last reader, block jn - 2, finished production
            // before an earlier barrier)
            load_block(jn);
            ++jn;
            __syncthreads();
        }
        if (nbuf == 2 && gs + 1 < nslab) produce(gs + 1);  // into the other buffer
        kmat_split_consume<S, NPW>(pb, pe, Tp, c0, wcols, lane, s_slab + (nbuf == 2 ? (gs & 1) * BUF : 0), acc);
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

template <int S, int NW, int NPW, bool PERSIST>
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
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];
    if constexpr (PERSIST) {
        __shared__ int s_item;
        for (int iter = 0;; ++iter) {
            const int item = next_item(queue, n_items, iter, &s_item);
            if (item >= n_items) break;
            k_kmat_split_item<S, NW, NPW>(item, work, groups, blocks, gblocks, pts, n_slots, shrec, RS, gshells, gshell_col, gcols, gsorder, gmasks, coef, K, s_pt, s_cf, s_msk, s_first, s_blk, s_dyn);
        }
    } else {
        k_kmat_split_item<S, NW, NPW>((int)blockIdx.x, work, groups, blocks, gblocks, pts, n_slots, shrec, RS, gshells, gshell_col, gcols, gsorder, gmasks, coef, K, s_pt, s_cf, s_msk, s_first, s_blk, s_dyn);
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
// v_mfma_f32_16x16x16_bf16 (k < 16)
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

template <int S, int NW, int NPW>
static int kmat_split_launch(Context& c, const LaunchRange& r) {
    unsigned grid = 0;
    int* q = launch_queue(c, kQueueK, r.count, resident_per_cu(&k_kmat_split<S, NW, NPW, true>, NW * 64, r.lds),
                          &grid);
    if (q)
        hipLaunchKernelGGL((k_kmat_split<S, NW, NPW, true>), dim3(grid), dim3(NW * 64), r.lds, c.stream,
                           c.d_kg_work + r.offset, c.d_kgroups, c.d_blocks, c.d_kg_blocks, c.d_pts, c.n_slots,
                           c.d_shrec, c.rec_stride, c.d_kg_shells, c.d_kg_shell_col, c.d_kg_cols, c.d_kg_sorder,
                           c.d_kg_masks, c.d_coef, c.d_K, q, r.count);
    else
        hipLaunchKernelGGL((k_kmat_split<S, NW, NPW, false>), dim3(grid), dim3(NW * 64), r.lds, c.stream,
                           c.d_kg_work + r.offset, c.d_kgroups, c.d_blocks, c.d_kg_blocks, c.d_pts, c.n_slots,
                           c.d_shrec, c.rec_stride, c.d_kg_shells, c.d_kg_shell_col, c.d_kg_cols, c.d_kg_sorder,
                           c.d_kg_masks, c.d_coef, c.d_K, q, r.count);
    return check("k_kmat_split");
}

// the workgroup shape of a launch class: items of up to 16 patches, 16 waves
// of one patch (127 VGPRs); up to 24, 12 waves of two (167); up to 32, 8
// waves of four (247)
template <int S>
static int kmat_split_one(Context& c, const LaunchRange& r) {
    switch (r.cls) {
        case 0: return kmat_split_launch<S, 16, 1>(c, r);
        case 1: return kmat_split_launch<S, 12, 2>(c, r);
        case 2: return kmat_split_launch<S, 8, 4>(c, r);
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
        default: set_error("split build: bad precision"); return PDFXC_ERR_ARGUMENT;
    }
}

int launch_split_kmat(Context& c) {
    int rc = PDFXC_OK;
    for (const LaunchRange& r : c.sk_launches) {
        if (rc != PDFXC_OK) break;
        if (r.count <= 0) continue;
        rc = (prec_slices(c.opts.precision) == 3) ? kmat_split_one<3>(c, r) : kmat_split_one<2>(c, r);
    }
    return rc;
}

// dynamic LDS a k_kmat_split<S> workgroup may take: the 160 KiB of a gfx950
// workgroup less the kernel's static LDS, read from the code object
int split_kmat_max_dyn_lds(int S, size_t* out) {
    hipFuncAttributes a;
    const hipError_t e = (S == 3) ? hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_kmat_split<3, 16, 1, false>))
                                  : hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_kmat_split<2, 16, 1, false>));
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes(k_kmat_split)");
    *out = (size_t)160 * 1024 - a.sharedSizeBytes;
    return PDFXC_OK;
}

int launch_mfma_split_selftest(float* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_split_selftest, dim3(1), dim3(64), 0, s, d_out);
    return check("k_mfma_split_selftest");
}

}  // namespace pdfxc
