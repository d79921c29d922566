// HIP kernels of the grid XC build for gfx950 (CDNA4), FP64 path.
//
// Per grid block b (128 points, N_b screened AOs, ldn = N_b padded to 16),
// AO values are evaluated where they are consumed and never stored in HBM:
//   k_rho   for 32 (or 16) points of a block: X = Phi_b P_bb on
//           v_mfma_f64_16x16x4 with X held in accumulators, Phi streamed in
//           16-column slabs through LDS; then rho = sum_j X_pj phi_pj and
//           grad rho = 2 sum_j X_pj grad phi_pj with phi, grad phi
//           re-evaluated per 64-column chunk, then the B3LYP point terms.
//   k_kmat  for a range of lower 16 x 16 tiles of K_bb (held in
//           accumulators): Phi_b^T V_b + V_b^T Phi_b over the block's points
//           in 16-point slabs, phi / grad phi / V evaluated into LDS, then
//           scattered into the dense lower triangle of K.
// The reference runs the same quantities point by point
// (DfCalcGridX.cpp:1923-1962, :2137-2193, :2455-2500).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pdfxc_internal.hpp"
#include "xc_functional.hpp"

namespace pdfxc {

typedef double v4d __attribute__((ext_vector_type(4)));

// MFMA f64 16x16x4 accumulator layout (checked on gfx950 by
// pdfxc_selftest_mfma): register r of lane l holds D[4 r + l / 16][l % 16];
// operand A[i][k] comes from lane i + 16 k, operand B[k][j] from lane j + 16 k.
__device__ inline v4d mfma16x16x4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr double kTooBig = 30.0;  // DfCalcGridX::TOOBIG, primitives with alpha r^2 > 30 are skipped
constexpr double kInvSqrt3 = 0.57735026918962576451;

// ---------------------------------------------------------------------------
// AO values and gradients for one block (getAOs_core / getDAOs_core,
// DfCalcGridX.cpp:667-769).  Radial sums R0 = sum c e^{-a r^2} and
// R1 = sum -2a c e^{-a r^2} are shared by the 2l+1 components; the angular
// prefactors and their gradients are those of getPrefactor (:103-141).
// ---------------------------------------------------------------------------
__device__ inline void shell_values(const DeviceShell& S, double x, double y, double z, double* v,
                                    double* vx, double* vy, double* vz) {
    const double r2 = x * x + y * y + z * z;
    double R0 = 0.0, R1 = 0.0;
    for (int k = 0; k < S.nprim; ++k) {
        const double a = S.exps[k];
        const double sh = a * r2;
        if (sh <= kTooBig) {
            const double g = S.coefs[k] * exp(-sh);
            R0 += g;
            R1 += -2.0 * a * g;
        }
    }
    if (S.l == 0) {
        v[0] = R0;
        vx[0] = x * R1;
        vy[0] = y * R1;
        vz[0] = z * R1;
    } else if (S.l == 1) {
        const double c[3] = {x, y, z};
        for (int m = 0; m < 3; ++m) {
            v[m] = c[m] * R0;
            vx[m] = (m == 0 ? R0 : 0.0) + c[m] * x * R1;
            vy[m] = (m == 1 ? R0 : 0.0) + c[m] * y * R1;
            vz[m] = (m == 2 ? R0 : 0.0) + c[m] * z * R1;
        }
    } else {
        // 4 xy, 5 xz, 6 yz, 7 (x^2-y^2)/2, 8 (z^2-(x^2+y^2)/2)/sqrt3
        const double f[5] = {x * y, z * x, y * z, 0.5 * (x * x - y * y),
                             kInvSqrt3 * (z * z - 0.5 * (x * x + y * y))};
        const double fx[5] = {y, z, 0.0, x, -kInvSqrt3 * x};
        const double fy[5] = {x, 0.0, z, -y, -kInvSqrt3 * y};
        const double fz[5] = {0.0, x, y, 0.0, 2.0 * kInvSqrt3 * z};
        for (int m = 0; m < 5; ++m) {
            v[m] = f[m] * R0;
            vx[m] = fx[m] * R0 + f[m] * x * R1;
            vy[m] = fy[m] * R0 + f[m] * y * R1;
            vz[m] = fz[m] * R0 + f[m] * z * R1;
        }
    }
}

// ---------------------------------------------------------------------------
// Shell records.  Every shell has a record of RS doubles in global memory:
// [cx, cy, cz, l + 4 nprim, e0, c0, e1, c1, ...]; a block copies the records
// it needs into LDS once, so that AO evaluation reads no global memory.
// ---------------------------------------------------------------------------
__device__ inline void rec_values(const double* rec, double px, double py, double pz, double* v,
                                  double* vx, double* vy, double* vz) {
    const double x = px - rec[0], y = py - rec[1], z = pz - rec[2];
    const int meta = (int)rec[3];
    const int l = meta & 3, nprim = meta >> 2;
    const double r2 = x * x + y * y + z * z;
    double R0 = 0.0, R1 = 0.0;
    for (int k = 0; k < nprim; ++k) {
        const double a = rec[4 + 2 * k];
        const double sh = a * r2;
        if (sh <= kTooBig) {
            const double g = rec[5 + 2 * k] * exp(-sh);
            R0 += g;
            R1 += -2.0 * a * g;
        }
    }
    if (l == 0) {
        v[0] = R0;
        vx[0] = x * R1;
        vy[0] = y * R1;
        vz[0] = z * R1;
    } else if (l == 1) {
        v[0] = x * R0;
        v[1] = y * R0;
        v[2] = z * R0;
        vx[0] = R0 + x * x * R1;
        vy[0] = x * y * R1;
        vz[0] = x * z * R1;
        vx[1] = y * x * R1;
        vy[1] = R0 + y * y * R1;
        vz[1] = y * z * R1;
        vx[2] = z * x * R1;
        vy[2] = z * y * R1;
        vz[2] = R0 + z * z * R1;
    } else {
        // 4 xy, 5 xz, 6 yz, 7 (x^2-y^2)/2, 8 (z^2-(x^2+y^2)/2)/sqrt3
        const double f[5] = {x * y, z * x, y * z, 0.5 * (x * x - y * y),
                             kInvSqrt3 * (z * z - 0.5 * (x * x + y * y))};
        const double fx[5] = {y, z, 0.0, x, -kInvSqrt3 * x};
        const double fy[5] = {x, 0.0, z, -y, -kInvSqrt3 * y};
        const double fz[5] = {0.0, x, y, 0.0, 2.0 * kInvSqrt3 * z};
#pragma unroll
        for (int m = 0; m < 5; ++m) {
            v[m] = f[m] * R0;
            vx[m] = fx[m] * R0 + f[m] * x * R1;
            vy[m] = fy[m] * R0 + f[m] * y * R1;
            vz[m] = fz[m] * R0 + f[m] * z * R1;
        }
    }
}

// ---------------------------------------------------------------------------
// rho pass for PB = 16 RT points of a block: X^T = P_bb Phi^T on
// v_mfma_f64_16x16x4 (rows = AOs, columns = points) with X^T held in
// accumulators; wave w owns the AO tiles w, w + 4, ....  Phi of all ldn
// columns is evaluated into LDS once; the A operand comes from P in global
// memory (L2), prefetched one group of G k-steps ahead.  Epilogue: phi and
// grad phi per 64-column chunk (one AO tile per wave) evaluated into the same
// LDS and dotted with X^T: rho = sum_j X_pj phi_pj, grad rho = 2 sum_j X_pj
// grad phi_pj (calcRho_GGA_part, DfCalcGridX.cpp:1923-1962); then the B3LYP
// point terms (buildVxc, :2137-2193): coef = (w vrho / 2, w (2 vgAA + vgAB)
// grad rho) and the weighted energy.
// ---------------------------------------------------------------------------
template <int RT, int MAXI, int G>
__global__ __launch_bounds__(256) void k_rho(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                             const double* __restrict__ pts, int64_t n_slots,
                                             const double* __restrict__ shrec, int RS,
                                             const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                             const int* __restrict__ bcols, const int* __restrict__ bslab,
                                             const int* __restrict__ bsorder, const double* __restrict__ P, int N,
                                             double* __restrict__ coef_out, double* __restrict__ rho_out,
                                             double* __restrict__ ewg, double dens_cut) {
    constexpr int PB = 16 * RT;
    constexpr int SUBS = kBlockPts / PB;
    __shared__ double s_pt[PB][4];
    __shared__ double s_red[4][PB][4];
    __shared__ double s_e[PB];
    extern __shared__ double s_dyn[];

    const int bidx = work[blockIdx.x / SUBS];
    const int sub = blockIdx.x % SUBS;
    const BlockDesc B = blocks[bidx];
    const int tid = threadIdx.x;
    const int p0 = sub * PB;
    const int64_t wg_slot = (int64_t)bidx * 8 + sub;
    if (p0 >= B.npts) {  // no real point in this part of the block (uniform exit, before any barrier)
        if (tid == 0) ewg[wg_slot] = 0.0;
        return;
    }
    const int lane = tid & 63, li = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldn = B.ldn, nao = B.nao, nsh = B.nshell;
    const int T = ldn >> 4;
    // LDS: phi [max(ldn, 256)][PB] | records [nsh][RS] | cols [ldn] | shcol [nsh] | sorder [nsh] | slab [T+1]
    double* s_phi = s_dyn;
    double* s_rec = s_phi + (size_t)max(ldn, 4 * 64) * PB;
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + ldn;
    int* s_sord = s_shcol + nsh;
    int* s_slab = s_sord + nsh;
    for (int c = tid; c < ldn; c += 256) s_cols[c] = (c < nao) ? bcols[B.col_begin + c] : -1;
    for (int c = tid; c < nsh; c += 256) {
        s_shcol[c] = bshell_col[B.shell_begin + c];
        s_sord[c] = bsorder[B.shell_begin + c];
    }
    for (int c = tid; c <= T; c += 256) s_slab[c] = bslab[B.slab_begin + c];
    for (int idx = tid; idx < nsh * RS; idx += 256) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    if (tid < PB) {
        const int64_t slot = B.pt_begin + p0 + tid;
        s_pt[tid][0] = pts[slot];
        s_pt[tid][1] = pts[n_slots + slot];
        s_pt[tid][2] = pts[2 * n_slots + slot];
        s_pt[tid][3] = pts[3 * n_slots + slot];
    }
    __syncthreads();

    // phase 1: phi of all columns at the PB points, shells in (l, nprim) order
    for (int it = tid; it < nsh * PB; it += 256) {
        const int p = it % PB;
        const int k = s_sord[it / PB];
        double v[5], vx[5], vy[5], vz[5];
        rec_values(s_rec + k * RS, s_pt[p][0], s_pt[p][1], s_pt[p][2], v, vx, vy, vz);
        const int c0 = s_shcol[k];
        const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
        for (int m = 0; m < nc; ++m) s_phi[(c0 + m) * PB + p] = v[m];
    }
    for (int it = tid; it < (ldn - nao) * PB; it += 256) s_phi[nao * PB + it] = 0.0;
    __syncthreads();

    // phase 2: X^T = P_bb Phi^T, A fragments from P, B fragments from LDS
    v4d acc[MAXI][RT];
#pragma unroll
    for (int i = 0; i < MAXI; ++i)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[i][rt] = v4d{0.0, 0.0, 0.0, 0.0};
    int gi[MAXI];
#pragma unroll
    for (int i = 0; i < MAXI; ++i) {
        const int I = wave + 4 * i;
        gi[i] = (I < T) ? s_cols[I * 16 + li] : -1;
    }
    const int nsteps = T * 4;
    const int ngroups = (nsteps + G - 1) / G;
    double aC[G][MAXI], aN[G][MAXI];
    auto load_a = [&](int g, double (&a)[G][MAXI]) {
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const int step = g * G + d;
            const int gk = (step < nsteps) ? s_cols[4 * step + lq] : -1;
            const double* prow = P + (int64_t)(gk < 0 ? 0 : gk) * N;
#pragma unroll
            for (int i = 0; i < MAXI; ++i) a[d][i] = (gk >= 0 && gi[i] >= 0) ? prow[gi[i]] : 0.0;
        }
    };
    load_a(0, aC);
    for (int g = 0; g < ngroups; ++g) {
        if (g + 1 < ngroups) load_a(g + 1, aN);
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const int step = g * G + d;
            if (step < nsteps) {
                double b[RT];
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) b[rt] = s_phi[(4 * step + lq) * PB + rt * 16 + li];
#pragma unroll
                for (int i = 0; i < MAXI; ++i)
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt) acc[i][rt] = mfma16x16x4(aC[d][i], b[rt], acc[i][rt]);
            }
        }
#pragma unroll
        for (int d = 0; d < G; ++d)
#pragma unroll
            for (int i = 0; i < MAXI; ++i) aC[d][i] = aN[d][i];
    }

    // phase 3: epilogue per 64-column chunk; E[q][col][p] reuses the phi region
    double part[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[rt][q] = 0.0;
    double* E = s_phi;
    constexpr int EQ = 64 * PB;  // doubles per E plane
#pragma unroll
    for (int c = 0; c < MAXI; ++c) {
        const int lo = c * 64;
        if (lo < ldn) {
            const int hi = min(lo + 64, ldn);
            const int first = s_slab[lo >> 4];
            int last = first;
            while (last < nsh && s_shcol[last] < hi) ++last;
            __syncthreads();
            for (int it = tid; it < (last - first) * PB; it += 256) {
                const int p = it % PB;
                const int k = first + it / PB;
                double v[5], vx[5], vy[5], vz[5];
                rec_values(s_rec + k * RS, s_pt[p][0], s_pt[p][1], s_pt[p][2], v, vx, vy, vz);
                const int c0 = s_shcol[k];
                const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
                for (int m = 0; m < nc; ++m) {
                    const int col = c0 + m;
                    if (col >= lo && col < hi) {
                        double* e = E + (col - lo) * PB + p;
                        e[0] = v[m];
                        e[EQ] = vx[m];
                        e[2 * EQ] = vy[m];
                        e[3 * EQ] = vz[m];
                    }
                }
            }
            __syncthreads();
            const int I = wave + 4 * c;
            if (I < T) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int cl = wave * 16 + 4 * r + lq;  // column within the chunk
                    if (s_cols[lo + cl] >= 0) {
#pragma unroll
                        for (int rt = 0; rt < RT; ++rt) {
                            const double x = acc[c][rt][r];
                            const double* e = E + cl * PB + rt * 16 + li;
                            part[rt][0] += x * e[0];
                            part[rt][1] += x * e[EQ];
                            part[rt][2] += x * e[2 * EQ];
                            part[rt][3] += x * e[3 * EQ];
                        }
                    }
                }
            }
        }
    }
    // reduce over the four lanes that share a point (lq), then over waves
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v = part[rt][q];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            part[rt][q] = v;
        }
    if (lq == 0) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int q = 0; q < 4; ++q) s_red[wave][rt * 16 + li][q] = part[rt][q];
    }
    __syncthreads();
    if (tid < PB) {
        const int p = tid;
        double rq[4];
        for (int q = 0; q < 4; ++q) rq[q] = s_red[0][p][q] + s_red[1][p][q] + s_red[2][p][q] + s_red[3][p][q];
        const double rho = rq[0];
        const double gx = 2.0 * rq[1], gy = 2.0 * rq[2], gz = 2.0 * rq[3];
        const int64_t slot = B.pt_begin + p0 + p;
        const double w = s_pt[p][3];
        if (rho_out) {
            rho_out[4 * slot + 0] = rho;
            rho_out[4 * slot + 1] = gx;
            rho_out[4 * slot + 2] = gy;
            rho_out[4 * slot + 3] = gz;
        }
        double a = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, e = 0.0;
        if (p0 + p < B.npts && rho > dens_cut) {
            const double gamma = gx * gx + gy * gy + gz * gz;
            const XcPoint f = b3lyp::eval_rks(rho, gamma);
            // build_XC_Matrix GGA (DfCalcGridX.cpp:2455-2500): coef1 = w vrho,
            // coef2 = w (2 vgAA + vgAB) grad rho; V carries half of coef1
            a = 0.5 * (w * f.vr);
            const double c2 = 2.0 * f.vgaa;
            cx = w * (c2 * gx + f.vgab * gx);
            cy = w * (c2 * gy + f.vgab * gy);
            cz = w * (c2 * gz + f.vgab * gz);
            e = w * f.e;
        }
        coef_out[4 * slot + 0] = a;
        coef_out[4 * slot + 1] = cx;
        coef_out[4 * slot + 2] = cy;
        coef_out[4 * slot + 3] = cz;
        s_e[p] = e;
    }
    __syncthreads();
    if (tid == 0) {
        double e = 0.0;
        for (int p = 0; p < PB; ++p) e += s_e[p];
        ewg[wg_slot] = e;
    }
}

// ---------------------------------------------------------------------------
// K pass: K_bb += Phi^T V + V^T Phi for the lower 16 x 16 tiles t0 <= t < t1
// of one block (t = I (I + 1) / 2 + J).  Eight waves: waves 0-3 (consumers)
// hold the tiles in accumulators -- each a contiguous quarter, padded to MAXT
// with a repeated tile that is not scattered, so the accumulate loop has no
// branches -- and run the MFMAs on point slab s while waves 4-7 (producers)
// evaluate slab s + 1 into the other half of a double-buffered LDS area:
// phi and grad phi of the needed columns from the LDS shell records, and
// V = (w vrho / 2) phi + c . grad phi with the per-point coefficients
// (build_XC_Matrix GGA, DfCalcGridX.cpp:2455-2500).  Slab rows are padded to
// a stride of 16 (mod 32) doubles so that the 16-lane groups of an MFMA
// operand read fall on disjoint LDS banks.  Tiles go into the dense lower
// triangle of K with FP64 atomics.
// ---------------------------------------------------------------------------
template <int MAXT>
__global__ __launch_bounds__(512) void k_kmat(const KWork* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                              const double* __restrict__ pts, int64_t n_slots,
                                              const double* __restrict__ shrec, int RS,
                                              const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                              const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                              const double* __restrict__ coef, double* __restrict__ K, int N) {
    __shared__ double s_pt[kBlockPts][3];
    __shared__ double s_cf[kBlockPts][4];
    extern __shared__ double s_dynd[];
    const KWork W = work[blockIdx.x];
    const BlockDesc B = blocks[W.block];
    const int tid = threadIdx.x;
    const int lane = tid & 63, li = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool producer = wave >= 4;
    const int ncols = W.ncols;  // multiple of 16, <= ldn
    const int LD = (ncols & 31) ? ncols : ncols + 16;
    const int PS = W.ps;
    const int nsh = W.nsh;
    const int SLAB = PS * LD;  // doubles per array per buffer
    // LDS: [2 buffers][phi, V][PS][LD] | records [nsh][RS] | cols [ncols] | shcol [nsh] | shid [nsh]
    double* s_buf = s_dynd;
    double* s_rec = s_buf + 4 * SLAB;
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + ncols;
    int* s_shid = s_shcol + nsh;
    for (int c = tid; c < ncols; c += 512) s_cols[c] = (c < B.nao) ? bcols[B.col_begin + c] : -1;
    // shells this item needs (local index < nsh), in (l, nprim) order, compacted by wave 4
    if (wave == 4) {
        int base = 0;
        for (int i0 = 0; i0 < B.nshell; i0 += 64) {
            const int i = i0 + lane;
            const int sh = (i < B.nshell) ? bsorder[B.shell_begin + i] : B.nshell;
            const bool keep = sh < nsh;
            const unsigned long long m = __ballot(keep);
            const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
            if (keep) {
                s_shid[pos] = bshells[B.shell_begin + sh];
                s_shcol[pos] = bshell_col[B.shell_begin + sh];
            }
            base += __popcll(m);
        }
    }
    for (int p = tid; p < kBlockPts; p += 512) {
        const int64_t slot = B.pt_begin + p;
        s_pt[p][0] = pts[slot];
        s_pt[p][1] = pts[n_slots + slot];
        s_pt[p][2] = pts[2 * n_slots + slot];
        s_cf[p][0] = coef[4 * slot + 0];
        s_cf[p][1] = coef[4 * slot + 1];
        s_cf[p][2] = coef[4 * slot + 2];
        s_cf[p][3] = coef[4 * slot + 3];
    }
    __syncthreads();
    for (int idx = tid; idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)s_shid[k] * RS + (idx - k * RS)];
    }
    __syncthreads();

    const int ptid = tid - 256;  // producer thread index
    auto produce = [&](int pb, double* phi, double* vv) {
        for (int it = ptid; it < nsh * PS; it += 256) {
            const int p = it % PS;
            const int k = it / PS;
            const int pp = pb + p;
            double v[5], vx[5], vy[5], vz[5];
            rec_values(s_rec + k * RS, s_pt[pp][0], s_pt[pp][1], s_pt[pp][2], v, vx, vy, vz);
            const double a = s_cf[pp][0], cx = s_cf[pp][1], cy = s_cf[pp][2], cz = s_cf[pp][3];
            const int c0 = s_shcol[k];
            const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
            for (int m = 0; m < nc; ++m) {
                const int c = c0 + m;
                if (c < ncols) {
                    phi[p * LD + c] = v[m];
                    vv[p * LD + c] = a * v[m] + cx * vx[m] + cy * vy[m] + cz * vz[m];
                }
            }
        }
        for (int it = ptid; it < (ncols - min(ncols, B.nao)) * PS; it += 256) {
            const int c = B.nao + it / PS;
            const int p = it % PS;
            phi[p * LD + c] = 0.0;
            vv[p * LD + c] = 0.0;
        }
    };

    const int nslab = (B.npts + PS - 1) / PS;
    if (producer) {
        // producers: slab s + 1 while the consumers multiply slab s; the
        // number of barriers matches the consumer branch
        produce(0, s_buf, s_buf + SLAB);
        __syncthreads();
        for (int s = 0; s < nslab; ++s) {
            if (s + 1 < nslab) {
                double* nxt = s_buf + ((s + 1) & 1) * 2 * SLAB;
                produce((s + 1) * PS, nxt, nxt + SLAB);
            }
            __syncthreads();
        }
        return;
    }
    // consumers: a contiguous run [tb, te) of tiles, row (I) and column (J) offsets
    const int q4 = (W.t1 - W.t0 + 3) >> 2;
    const int tb = W.t0 + wave * q4;
    const int te = min(W.t1, tb + q4);
    int offI[MAXT], offJ[MAXT];
    {
        const int t = min(tb, W.t1 - 1);
        int I = 0;
        while ((I + 1) * (I + 2) / 2 <= t) ++I;
        int J = t - I * (I + 1) / 2;
#pragma unroll
        for (int m = 0; m < MAXT; ++m) {
            offI[m] = I * 16;
            offJ[m] = J * 16;
            if (tb + m + 1 < te) {  // advance along the run; padding repeats the last tile
                if (J == I) {
                    ++I;
                    J = 0;
                } else {
                    ++J;
                }
            }
        }
    }
    v4d acc[MAXT];
#pragma unroll
    for (int m = 0; m < MAXT; ++m) acc[m] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
        const double* cur = s_buf + (s & 1) * 2 * SLAB;
        for (int ks = 0; ks < PS / 4; ++ks) {
            const double* fr = cur + (4 * ks + lq) * LD + li;
            const double* vr = cur + SLAB + (4 * ks + lq) * LD + li;
            // groups of 4 tiles: bounded operand lookahead keeps the
            // accumulators and operands within the register budget
#pragma unroll
            for (int m0 = 0; m0 < MAXT; m0 += 4) {
#pragma unroll
                for (int m = m0; m < m0 + 4 && m < MAXT; ++m) {
                    const double a1 = fr[offI[m]], b1 = vr[offJ[m]];
                    const double a2 = vr[offI[m]], b2 = fr[offJ[m]];
                    acc[m] = mfma16x16x4(a1, b1, acc[m]);
                    acc[m] = mfma16x16x4(a2, b2, acc[m]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();
    }
    // scatter: tile element (4 r + lq, li) -> K[cols[I 16 + 4 r + lq]][cols[J 16 + li]]
#pragma unroll
    for (int m = 0; m < MAXT; ++m) {
        if (tb + m < te) {
            const int lc = offJ[m] + li;
            const int gj = s_cols[lc];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int lr = offI[m] + 4 * r + lq;
                const int gi = s_cols[lr];
                if (gi >= 0 && gj >= 0 && lr >= lc) atomicAdd(K + (int64_t)gi * N + gj, acc[m][r]);
            }
        }
    }
}

// deterministic sum of the per-block energies (one workgroup)
__global__ __launch_bounds__(256) void k_energy_sum(const double* __restrict__ eb, int n,
                                                    double* __restrict__ out) {
    __shared__ double s[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += eb[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

__global__ __launch_bounds__(256) void k_eval_functional(int64_t n, const double* __restrict__ in,
                                                         double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const XcPoint f = b3lyp::eval_rks(in[2 * i], in[2 * i + 1]);
    out[4 * i + 0] = f.e;
    out[4 * i + 1] = f.vr;
    out[4 * i + 2] = f.vgaa;
    out[4 * i + 3] = f.vgab;
}

// all AOs at arbitrary points, one thread per (point, shell): out = 4 planes n x N
__global__ __launch_bounds__(256) void k_eval_aos_points(int64_t n, const double* __restrict__ xyz,
                                                         const DeviceShell* __restrict__ shells,
                                                         int nshell, int N, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n * nshell) return;
    const int64_t p = t / nshell;
    const int s = (int)(t % nshell);
    const DeviceShell& S = shells[s];
    double v[5], vx[5], vy[5], vz[5];
    shell_values(S, xyz[3 * p] - S.cx, xyz[3 * p + 1] - S.cy, xyz[3 * p + 2] - S.cz, v, vx, vy, vz);
    const int64_t plane = n * (int64_t)N;
    for (int m = 0; m < 2 * S.l + 1; ++m) {
        const int64_t o = p * N + S.first_ao + m;
        out[o] = v[m];
        out[plane + o] = vx[m];
        out[2 * plane + o] = vy[m];
        out[3 * plane + o] = vz[m];
    }
}

// ---------------------------------------------------------------------------
// Becke partition weights (calcMultiCenterWeight_Becke + Ps_uij,
// DfGenerateGrid.cpp:1811-1876, :2043-2086): w *= P_A / sum_m P_m with
// P_m = prod_{n != m} (1 - f3(nu_mn)) / 2, nu = mu + a_mn (1 - mu^2),
// mu = (r_m - r_n) / R_mn.  atoms: n_atoms x 4 (x, y, z, Z) then the
// n_atoms^2 inverse distances and size adjustments a_mn.  One thread per
// point; the point's atom distances are kept in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_becke(const double* __restrict__ atoms, int na,
                                              int64_t n, const double* __restrict__ pts,
                                              const int* __restrict__ atom_of_pt,
                                              double* __restrict__ w) {
    extern __shared__ double s_r[];  // blockDim.x x na
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* inv_d = atoms + 4 * na;
    const double* amat = inv_d + (int64_t)na * na;
    double* r = s_r + threadIdx.x * na;
    const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    for (int m = 0; m < na; ++m) {
        const double dx = x - atoms[4 * m], dy = y - atoms[4 * m + 1], dz = z - atoms[4 * m + 2];
        r[m] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    const int A = atom_of_pt[i];
    double total = 0.0, mine = 0.0;
    for (int m = 0; m < na; ++m) {
        if (atoms[4 * m + 3] <= 0.0) continue;  // dummy atom: no cell of its own
        double ps = 1.0;
        for (int k = 0; k < na && ps != 0.0; ++k) {
            if (k == m || atoms[4 * k + 3] <= 0.0) continue;
            double u = (r[m] - r[k]) * inv_d[(int64_t)m * na + k];
            const double a = amat[(int64_t)m * na + k];
            u = u + a * (1.0 - u * u);
            double f = 0.5 * u * (3.0 - u * u);
            f = 0.5 * f * (3.0 - f * f);
            f = 0.5 * f * (3.0 - f * f);
            ps *= 0.5 * (1.0 - f);
        }
        if (m == A) mine = ps;
        total += ps;
    }
    w[i] *= (mine / total);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return PDFXC_OK;
}

template <int RT, int MAXI, int G>
static int rho_one(Context& c, const LaunchRange& r) {
    const int subs = kBlockPts / (16 * RT);
    hipLaunchKernelGGL((k_rho<RT, MAXI, G>), dim3(r.count * subs), dim3(256), r.lds, c.stream,
                       c.d_rho_work + r.offset, c.d_blocks, c.d_pts, c.n_slots, c.d_shrec, c.rec_stride,
                       c.d_block_shells, c.d_block_shell_col, c.d_block_cols, c.d_block_slab,
                       c.d_block_sorder, c.d_P, c.n_aos, c.d_coef, c.d_rho, c.d_eblock,
                       c.opts.density_cutoff);
    return check_launch("k_rho");
}

int launch_rho(Context& c, const LaunchRange& r) {
    if (r.count <= 0) return PDFXC_OK;
    switch (r.cls) {
        case 0: return rho_one<2, 2, 4>(c, r);
        case 1: return rho_one<2, 4, 4>(c, r);
        case 2: return rho_one<1, 8, 2>(c, r);
        case 3: return rho_one<1, 16, 1>(c, r);
        default: set_error("k_rho: bad class"); return PDFXC_ERR_ARGUMENT;
    }
}

template <int MAXT>
static int kmat_one(Context& c, const LaunchRange& r) {
    hipLaunchKernelGGL((k_kmat<MAXT>), dim3(r.count), dim3(512), r.lds, c.stream,
                       reinterpret_cast<const KWork*>(c.d_k_work) + r.offset, c.d_blocks, c.d_pts,
                       c.n_slots, c.d_shrec, c.rec_stride, c.d_block_shells, c.d_block_shell_col,
                       c.d_block_cols, c.d_block_sorder, c.d_coef, c.d_K, c.n_aos);
    return check_launch("k_kmat");
}

int launch_kmat(Context& c, const LaunchRange& r) {
    if (r.count <= 0) return PDFXC_OK;
    switch (r.cls) {
        case 0: return kmat_one<8>(c, r);
        case 1: return kmat_one<24>(c, r);
        default: set_error("k_kmat: bad class"); return PDFXC_ERR_ARGUMENT;
    }
}

int configure_kernels() {
    // the block kernels use up to 160 KiB of LDS per workgroup (static + dynamic);
    // HIP on gfx950 needs no opt-in, but the device must offer it
    int dev = 0, lds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute(MaxSharedMemoryPerBlock)");
    if (lds < 160 * 1024) {
        set_error("device offers " + std::to_string(lds) + " bytes of LDS per workgroup; the kernels need 160 KiB (gfx950)");
        return PDFXC_ERR_UNSUPPORTED;
    }
    return PDFXC_OK;
}

int launch_energy_sum(Context& c, double* dE) {
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(256), 0, c.stream, c.d_eblock,
                       (int)c.blocks.size() * 8, dE);
    return check_launch("k_energy_sum");
}

int launch_eval_functional(int64_t n, const double* d_in, double* d_out, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    const int64_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_eval_functional, dim3((unsigned)g), dim3(256), 0, s, n, d_in, d_out);
    return check_launch("k_eval_functional");
}

int launch_eval_aos_points(const Context& c, int64_t n, const double* d_xyz, double* d_out,
                           hipStream_t s) {
    const int64_t work = n * (int64_t)c.shells.size();
    if (work <= 0) return PDFXC_OK;
    const int64_t g = (work + 255) / 256;
    hipLaunchKernelGGL(k_eval_aos_points, dim3((unsigned)g), dim3(256), 0, s, n, d_xyz, c.d_shells,
                       (int)c.shells.size(), c.n_aos, d_out);
    return check_launch("k_eval_aos_points");
}

int launch_becke(const double* d_atoms, int n_atoms, int64_t n, const double* d_pts,
                 const int* d_atom_of_pt, double* d_w, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    // point-to-atom distances live in LDS: keep them within 64 KiB per workgroup
    int threads = 64;
    while (threads > 1 && (size_t)threads * (size_t)n_atoms * sizeof(double) > 65536) threads >>= 1;
    if ((size_t)n_atoms * sizeof(double) > 65536) {
        set_error("Becke weights: too many atoms for the LDS distance table");
        return PDFXC_ERR_UNSUPPORTED;
    }
    const int64_t g = (n + threads - 1) / threads;
    const size_t lds = sizeof(double) * (size_t)threads * (size_t)n_atoms;
    hipLaunchKernelGGL(k_becke, dim3((unsigned)g), dim3(threads), lds, s, d_atoms, n_atoms, n,
                       d_pts, d_atom_of_pt, d_w);
    return check_launch("k_becke");
}

// MFMA layout self test: A[i][k] = i + 100 k, B[k][j] = j - 7 k; D = A B
__global__ void k_mfma_selftest(double* out) {
    const int l = threadIdx.x;
    const int i = l & 15, k = l >> 4;
    const double a = i + 100.0 * k;
    const double b = (double)(l & 15) - 7.0 * k;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = mfma16x16x4(a, b, acc);
    for (int r = 0; r < 4; ++r) out[(4 * r + (l >> 4)) * 16 + (l & 15)] = acc[r];
}

// plain streaming copy: the reference point for HBM traffic counters
__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ src, double* __restrict__ dst,
                                              int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}

int launch_copy(const double* src, double* dst, int64_t n, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    const int64_t g = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(256), 0, s, src, dst, n);
    return check_launch("k_copy");
}

int launch_mfma_selftest(double* d_out, hipStream_t s) {
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, s, d_out);
    return check_launch("k_mfma_selftest");
}

}  // namespace pdfxc
