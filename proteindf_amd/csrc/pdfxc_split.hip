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
//   k_mfull_split M_bb = Phi_b^T V_b over a block's points, tiles of a 16 x 16
//                 tile window, FP64 atomics into a dense N x N M
//   k_ksym        K_lower += M + M^T (build_XC_Matrix GGA, :2455-2500)
#include <hip/hip_runtime.h>

#include "pdfxc_device.hpp"
#include "pdfxc_internal.hpp"

namespace pdfxc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16 (checked on gfx950 by pdfxc_selftest_mfma_split):
// lane l holds A[l % 16][8 (l / 16) + e] and B[8 (l / 16) + e][l % 16],
// e = 0..7; register r of the result holds D[4 (l / 16) + r][l % 16].
__device__ inline f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline unsigned short bf16_bits(__bf16 h) { return __builtin_bit_cast(unsigned short, h); }

// slices of x: s[0] = bf16(x), s[i] = bf16(x - s[0] - ... - s[i-1]) with the
// residual kept in FP64
template <int S>
__device__ inline void split_bf16(double x, unsigned short* out) {
    double r = x;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const __bf16 h = (__bf16)(float)r;
        out[i] = bf16_bits(h);
        r -= (double)(float)h;
    }
}

// slice products kept: (a, b) with a + b < S
template <int S>
struct SlicePairs;
template <>
struct SlicePairs<2> {
    static constexpr int n = 3;
    static constexpr int a[3] = {0, 0, 1};
    static constexpr int b[3] = {0, 1, 0};
};
template <>
struct SlicePairs<3> {
    static constexpr int n = 6;
    static constexpr int a[6] = {0, 0, 1, 0, 2, 1};
    static constexpr int b[6] = {0, 1, 0, 2, 0, 1};
};

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
// X is then stored over the plane (FP32, row stride ldnp + 4) and dotted
// with phi, grad phi in FP64.
// ---------------------------------------------------------------------------
template <int S, int PB, int NJ>
__device__ inline void rho_split_tiles(int wave, int lane, int ldnp, int nch, const int* s_chunk,
                                       const unsigned short* s_a, float* s_x,
                                       const unsigned short* __restrict__ Ps, int N, int Nc) {
    constexpr int NP = PB / 16;
    const int li = lane & 15, lq = lane >> 4;
    const int RA = ldnp + 8;
    const int64_t plane = (int64_t)N * Nc;
    // each K step's slice products accumulate in FP32 from zero and are then
    // added to FP64 accumulators, so FP32 rounding acts on one step's sum only
    v4d acc[NP][NJ];
#pragma unroll
    for (int ip = 0; ip < NP; ++ip)
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) acc[ip][jj] = v4d{0.0, 0.0, 0.0, 0.0};
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
                f32x4 a32 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int q = 0; q < SlicePairs<S>::n; ++q)
                    a32 = mfma_bf16(afr[ip][SlicePairs<S>::a[q]], bfr[jj][SlicePairs<S>::b[q]], a32);
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[ip][jj][r] += (double)a32[r];
            }
    }
    __syncthreads();  // every wave has stopped reading the Phi plane
    const int RX = ldnp + 4;
#pragma unroll
    for (int ip = 0; ip < NP; ++ip)
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_x[(ip * 16 + 4 * lq + r) * RX + (wave + 8 * jj) * 16 + li] = (float)acc[ip][jj][r];
}

template <int S, int PB, int NJMAX>
__global__ __launch_bounds__(512) void k_rho_split(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                   const SplitBlock* __restrict__ sblocks,
                                                   const double* __restrict__ pts, int64_t n_slots,
                                                   const double* __restrict__ shrec, int RS,
                                                   const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                                   const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                                   const unsigned short* __restrict__ Ps, int N, int Nc,
                                                   double* __restrict__ rho_out, int accumulate) {
    constexpr int SUBS = kBlockPts / PB;
    constexpr int NSL = 512 / PB;  // shell lanes per point
    __shared__ double s_pt[PB][3];
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];

    const int bidx = work[blockIdx.x / SUBS];
    const int sub = blockIdx.x % SUBS;
    const BlockDesc B = blocks[bidx];
    const SplitBlock SB = sblocks[bidx];
    const int p0 = sub * PB;
    if (p0 >= B.npts) return;  // uniform exit before any barrier
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldnp = SB.ldnp, nch = SB.nchunk, nsh = B.nshell;
    const int RA = ldnp + 8;
    // LDS: plane (S x PB x RA bf16, reused as PB x (ldnp + 4) FP32 and as partial sums) | records | chunk [ldnp/4] | shcol | sorder
    unsigned short* s_a = reinterpret_cast<unsigned short*>(s_dyn);
    const int plane_dbl = (S * PB * RA * 2 + 15) / 16 * 2;  // doubles, 16-byte multiple
    const int region_dbl = max(plane_dbl, 4 * 512);  // also holds the 4 x 512 partial sums
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

    // 2. X = Phi P_bb, stored over the plane (FP32)
    float* s_x = reinterpret_cast<float*>(s_dyn);
    {
        const int nJ = ldnp / 16;
        const int nJw = (nJ > wave) ? (nJ - wave + 7) / 8 : 0;
        switch (nJw) {
            case 1: rho_split_tiles<S, PB, 1>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc); break;
            case 2:
                if constexpr (NJMAX >= 2) rho_split_tiles<S, PB, 2>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();  // unreachable: the host keeps ldnp <= 128 NJMAX
                break;
            case 3:
                if constexpr (NJMAX >= 3) rho_split_tiles<S, PB, 3>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            case 4:
                if constexpr (NJMAX >= 4) rho_split_tiles<S, PB, 4>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            case 5:
                if constexpr (NJMAX >= 5) rho_split_tiles<S, PB, 5>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            case 6:
                if constexpr (NJMAX >= 6) rho_split_tiles<S, PB, 6>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            case 7:
                if constexpr (NJMAX >= 7) rho_split_tiles<S, PB, 7>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            case 8:
                if constexpr (NJMAX >= 8) rho_split_tiles<S, PB, 8>(wave, lane, ldnp, nch, s_chunk, s_a, s_x, Ps, N, Nc);
                else __syncthreads();
                break;
            default: __syncthreads(); break;  // no columns: only the barrier before the store
        }
    }
    __syncthreads();

    // 3. rho and grad rho (FP64): phi, grad phi dotted with X
    const int RX = ldnp + 4;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    for (int j = kk; j < nsh; j += NSL) {
        const int k = s_sord[j];
        double v[5], vx[5], vy[5], vz[5];
        rec_values(s_rec + k * RS, px, py, pz, v, vx, vy, vz);
        const int c0 = s_shcol[k];
        const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
        for (int m = 0; m < nc; ++m) {
            const double x = (double)s_x[p * RX + c0 + m];
            r0 += x * v[m];
            r1 += x * vx[m];
            r2 += x * vy[m];
            r3 += x * vz[m];
        }
    }
    // partial sums of the NSL shell lanes of each point, added in a fixed order
    __syncthreads();  // X no longer read: the region holds the partial sums
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

// ---------------------------------------------------------------------------
// K pass, split precision: M_bb tiles [I0, I0 + nI) x [J0, J0 + nJ) (nI, nJ
// <= 16) of M = Phi^T V over the block's points in slabs of 32 (one MFMA K
// step).  All 512 threads evaluate a slab -- phi of the window's row columns,
// V = (w vrho / 2) phi + c . grad phi of its column columns -- as bf16
// slices [s][col][32 points] (the 8-point group g of column c stored at
// group g ^ ((c >> 2) & 3)), then wave w multiplies its 4 x 8 tiles
// (rows 4 (w >> 1) .., columns 8 (w & 1) ..).  FP64 atomics put the tiles
// into the dense M; k_ksym folds M + M^T into the lower triangle of K.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(512) void k_mfull_split(const SKWork* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                     const SplitBlock* __restrict__ sblocks,
                                                     const double* __restrict__ pts, int64_t n_slots,
                                                     const double* __restrict__ shrec, int RS,
                                                     const int* __restrict__ bshells, const int* __restrict__ bsorder,
                                                     const int* __restrict__ sshell_col, const int* __restrict__ chunks,
                                                     const int* __restrict__ perm, const double* __restrict__ coef,
                                                     double* __restrict__ M, int N) {
    __shared__ double s_pt[kBlockPts][3];
    __shared__ double s_cf[kBlockPts][4];
    extern __shared__ __attribute__((aligned(16))) double s_dyn[];
    const SKWork W = work[blockIdx.x];
    const BlockDesc B = blocks[W.block];
    const SplitBlock SB = sblocks[W.block];
    const int tid = threadIdx.x;
    const int lane = tid & 63, li = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nsh = B.nshell;
    const int ncI = 16 * W.nI, ncJ = 16 * W.nJ;
    const int cI0 = 16 * W.I0, cJ0 = 16 * W.J0;
    // LDS: phi slices [S][ncI][32] | V slices [S][ncJ][32] (bf16) | records | shcol | sorder | gI [ncI] | gJ [ncJ]
    unsigned short* s_phi = reinterpret_cast<unsigned short*>(s_dyn);
    unsigned short* s_v = s_phi + (size_t)S * ncI * 32;
    const int slab_dbl = (S * (ncI + ncJ) * 32 * 2 + 15) / 16 * 2;
    double* s_rec = s_dyn + slab_dbl;
    int* s_shcol = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_sord = s_shcol + nsh;
    int* s_gI = s_sord + nsh;
    int* s_gJ = s_gI + ncI;
    for (int k = tid; k < nsh; k += 512) {
        s_shcol[k] = sshell_col[B.shell_begin + k];
        s_sord[k] = bsorder[B.shell_begin + k];
    }
    for (int c = tid; c < ncI; c += 512) s_gI[c] = -1;
    for (int c = tid; c < ncJ; c += 512) s_gJ[c] = -1;
    for (int idx = tid; idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    for (int q = tid; q < kBlockPts; q += 512) {
        const int64_t slot = B.pt_begin + q;
        s_pt[q][0] = pts[slot];
        s_pt[q][1] = pts[n_slots + slot];
        s_pt[q][2] = pts[2 * n_slots + slot];
        s_cf[q][0] = coef[4 * slot + 0];
        s_cf[q][1] = coef[4 * slot + 1];
        s_cf[q][2] = coef[4 * slot + 2];
        s_cf[q][3] = coef[4 * slot + 3];
    }
    {
        uint4* z = reinterpret_cast<uint4*>(s_dyn);
        for (int i = tid; i < slab_dbl / 2; i += 512) z[i] = uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();
    // global AO of every kept column of the window (dropped AOs stay -1)
    for (int k = tid; k < nsh; k += 512) {
        const int c0 = s_shcol[k];
        const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
        for (int m = 0; m < nc; ++m) {
            const int c = c0 + m;
            const int g = perm[chunks[SB.chunk_begin + (c >> 2)] * 4 + (c & 3)];  // caller AO index
            if (c >= cI0 && c < cI0 + ncI) s_gI[c - cI0] = g;
            if (c >= cJ0 && c < cJ0 + ncJ) s_gJ[c - cJ0] = g;
        }
    }

    const int rg = wave >> 1, cg = wave & 1;
    f32x4 acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nslab = (B.npts + 31) / 32;
    for (int sl = 0; sl < nslab; ++sl) {
        // evaluate: 32 lanes = the 32 points of one shell
        for (int it = tid; it < nsh * 32; it += 512) {
            const int q = it & 31;
            const int k = s_sord[it >> 5];
            const int c0 = s_shcol[k];
            const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
            const bool inI = c0 < cI0 + ncI && c0 + nc > cI0;
            const bool inJ = c0 < cJ0 + ncJ && c0 + nc > cJ0;
            if (!inI && !inJ) continue;
            const int pp = sl * 32 + q;
            double v[5], vx[5], vy[5], vz[5];
            rec_values(s_rec + k * RS, s_pt[pp][0], s_pt[pp][1], s_pt[pp][2], v, vx, vy, vz);
            const double a = s_cf[pp][0], cx = s_cf[pp][1], cy = s_cf[pp][2], cz = s_cf[pp][3];
            for (int m = 0; m < nc; ++m) {
                const int c = c0 + m;
                if (c >= cI0 && c < cI0 + ncI) {
                    const int lc = c - cI0;
                    unsigned short h[S];
                    split_bf16<S>(v[m], h);
                    const int off = lc * 32 + ((((q >> 3) ^ (lc >> 2)) & 3) << 3) + (q & 7);
#pragma unroll
                    for (int s = 0; s < S; ++s) s_phi[(size_t)s * ncI * 32 + off] = h[s];
                }
                if (c >= cJ0 && c < cJ0 + ncJ) {
                    const int lc = c - cJ0;
                    unsigned short h[S];
                    split_bf16<S>(a * v[m] + cx * vx[m] + cy * vy[m] + cz * vz[m], h);
                    const int off = lc * 32 + ((((q >> 3) ^ (lc >> 2)) & 3) << 3) + (q & 7);
#pragma unroll
                    for (int s = 0; s < S; ++s) s_v[(size_t)s * ncJ * 32 + off] = h[s];
                }
            }
        }
        __syncthreads();
        // multiply: A[i][k] = phi at point k of row column i, B[k][j] = V at point k of column j;
        // row-outer order keeps one row's A slices and one column's B slices live
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int I = 4 * rg + i;
            if (I < W.nI) {
                const int lci = I * 16 + li;
                const int offa = lci * 32 + (((lq ^ (lci >> 2)) & 3) << 3);
                bf16x8 afr[S];
#pragma unroll
                for (int s = 0; s < S; ++s) afr[s] = *reinterpret_cast<const bf16x8*>(s_phi + (size_t)s * ncI * 32 + offa);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int J = 8 * cg + j;
                    if (J < W.nJ) {
                        const int lc = J * 16 + li;
                        const int off = lc * 32 + (((lq ^ (lc >> 2)) & 3) << 3);
                        bf16x8 bfr[S];
#pragma unroll
                        for (int s = 0; s < S; ++s)
                            bfr[s] = *reinterpret_cast<const bf16x8*>(s_v + (size_t)s * ncJ * 32 + off);
#pragma unroll
                        for (int qq = 0; qq < SlicePairs<S>::n; ++qq)
                            acc[i][j] = mfma_bf16(afr[SlicePairs<S>::a[qq]], bfr[SlicePairs<S>::b[qq]], acc[i][j]);
                    }
                }
            }
        }
        __syncthreads();
    }
    // scatter: register r of lane l is M[row 4 (l / 16) + r][column l % 16] of the tile
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int I = 4 * rg + i;
        if (I >= W.nI) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int J = 8 * cg + j;
            if (J >= W.nJ) continue;
            const int gj = s_gJ[J * 16 + li];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = s_gI[I * 16 + 4 * lq + r];
                if (gi >= 0 && gj >= 0) atomicAdd(M + (int64_t)gi * N + gj, (double)acc[i][j][r]);
            }
        }
    }
}

// K[i][j] += M[i][j] + M[j][i] for j <= i
__global__ __launch_bounds__(256) void k_ksym(const double* __restrict__ M, int N, double* __restrict__ K) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    const int i = (int)(t / N), j = (int)(t % N);
    if (j <= i) K[t] += M[t] + M[(int64_t)j * N + i];
}

// layout self test: A(i, k) = ((3 i + 5 k) % 11) - 5, B(k, j) = ((7 k + 2 j) % 13) - 6
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
    const int64_t n = (int64_t)c.n_aos * c.split_nc;
    hipLaunchKernelGGL((k_psplit<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.d_P, c.d_perm,
                       c.n_aos, c.split_nc, c.d_Ps);
    return check("k_psplit");
}

template <int S, int PB, int NJMAX>
static int rho_split_one(Context& c, const LaunchRange& r) {
    const int subs = kBlockPts / PB;
    hipLaunchKernelGGL((k_rho_split<S, PB, NJMAX>), dim3(r.count * subs), dim3(512), r.lds, c.stream,
                       c.d_srho_work + r.offset, c.d_blocks, c.d_sblocks, c.d_pts, c.n_slots, c.d_shrec,
                       c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col, c.d_chunks, c.d_Ps,
                       c.n_aos, c.split_nc, c.d_rho, c.rho_accumulate);
    return check("k_rho_split");
}

template <int S>
static int mfull_one(Context& c, const LaunchRange& r) {
    hipLaunchKernelGGL((k_mfull_split<S>), dim3(r.count), dim3(512), r.lds, c.stream,
                       reinterpret_cast<const SKWork*>(c.d_sk_work) + r.offset, c.d_blocks, c.d_sblocks, c.d_pts,
                       c.n_slots, c.d_shrec, c.rec_stride, c.d_block_shells, c.d_block_sorder, c.d_sshell_col,
                       c.d_chunks, c.d_perm, c.d_coef, c.d_M, c.n_aos);
    return check("k_mfull_split");
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
    PDFXC_HIP(hipMemsetAsync(c.d_M, 0, sizeof(double) * (size_t)c.n_aos * c.n_aos, c.stream));
    int rc = PDFXC_OK;
    for (const LaunchRange& r : c.sk_launches) {
        if (rc != PDFXC_OK) break;
        if (r.count <= 0) continue;
        rc = (c.opts.precision == PDFXC_SPLIT_BF16X3) ? mfull_one<3>(c, r) : mfull_one<2>(c, r);
    }
    if (rc != PDFXC_OK) return rc;
    const int64_t n = (int64_t)c.n_aos * c.n_aos;
    hipLaunchKernelGGL(k_ksym, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, c.d_M, c.n_aos, c.d_K);
    return check("k_ksym");
}

int launch_mfma_split_selftest(float* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_split_selftest, dim3(1), dim3(64), 0, s, d_out);
    return check("k_mfma_split_selftest");
}

}  // namespace pdfxc
