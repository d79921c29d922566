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
#include <mutex>
#include <string>

#include "pdfxc_device.hpp"
#include "pdfxc_internal.hpp"
#include "xc_functional.hpp"

namespace pdfxc {

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
    for (int k = 0; k < S.nprim; ++k) {  // exponents ascend: the first cut primitive ends the sum
        const double a = S.exps[k];
        const double sh = a * r2;
        if (sh > kTooBig) break;
        const double g = S.coefs[k] * exp_neg(sh);
        R0 += g;
        R1 += -2.0 * a * g;
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
// LDS layout of the per-point AO planes of k_rho: element (column c, point p)
// of a [ldn][PB] plane sits at c PB + (p ^ 16 (c & 1)).  The swizzle puts the
// four 16-lane groups of an MFMA B-operand read (columns 4 s + 0..3) on
// alternating bank halves.
// ---------------------------------------------------------------------------
template <int PB>
__device__ inline int plane_at(int c, int p) {
    if constexpr (PB >= 32) return c * PB + (p ^ ((c & 1) << 4));
    return c * PB + p;
}

// X^T tiles of one wave: rows I = wI + WI i (i < NI), point tiles
// wP RT .. wP RT + RT - 1; A = P_bb from global memory (any valid AO index
// stands in for a padding column: its Phi row is zero), B = Phi from LDS.
// After a barrier (every wave has stopped reading Phi) the tiles are stored
// over the Phi plane as X^T[col][p].
template <int PB, int RT, int NI, int G>
__device__ inline void rho_tiles(int wI, int WI, int wP, int T, int lane, const int* s_cols,
                                 double* s_phi, const double* __restrict__ P, int N) {
    const int li = lane & 15, lq = lane >> 4;
    v4d acc[NI][RT];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[i][rt] = v4d{0.0, 0.0, 0.0, 0.0};
    int gi[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) gi[i] = max(s_cols[(wI + WI * i) * 16 + li], 0);
    const int ngroups = T * 4 / G;  // G (2 or 4) divides the 4T steps: every group is whole
    double aC[G][NI], aN[G][NI];
    auto load_a = [&](int g, double (&a)[G][NI]) {
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const double* prow = P + (int64_t)max(s_cols[4 * (g * G + d) + lq], 0) * N;
#pragma unroll
            for (int i = 0; i < NI; ++i) a[d][i] = prow[gi[i]];
        }
    };
    // the first group through aN as well, so that aC never holds a pending
    // load inside the loop (else its first MFMA waits for every load)
    load_a(0, aN);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < G; ++d)
#pragma unroll
        for (int i = 0; i < NI; ++i) aC[d][i] = aN[d][i];
    for (int g = 0; g < ngroups; ++g) {
        if (g + 1 < ngroups) load_a(g + 1, aN);
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const int c = 4 * (g * G + d) + lq;
            double b[RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) b[rt] = s_phi[plane_at<PB>(c, (wP * RT + rt) * 16 + li)];
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) acc[i][rt] = mfma16x16x4(aC[d][i], b[rt], acc[i][rt]);
        }
        // the next group's operands are taken over here, after this group's
        // MFMAs: without the scheduling barrier the compiler moves each copy
        // up to the last use of its aC element, and the copy then waits for
        // the loads just issued for the next group, before this group's
        // first MFMA
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int d = 0; d < G; ++d)
#pragma unroll
            for (int i = 0; i < NI; ++i) aC[d][i] = aN[d][i];
    }
    __syncthreads();
    // tile element (4 r + lq, li) = X^T[I 16 + 4 r + lq][point tile 16 + li]
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_phi[plane_at<PB>((wI + WI * i) * 16 + 4 * r + lq, (wP * RT + rt) * 16 + li)] = acc[i][rt][r];
}

// X^T tiles of one wave of the 64-point, 8-wave form, balanced over the four
// SIMDs (waves w and w + 4 share SIMD w mod 4).  Of the T row tiles the first
// 4q (q = T / 4) go whole to SIMD I mod 4 (its two waves alternate rows: row
// s + 4 (2 m + h) to wave s + 4 h), and each of the last r = T mod 4 rows is
// split by point tile, SIMD s taking point tile s in the wave of its two that
// holds fewer whole rows; so every SIMD runs T row x point tiles, where whole
// rows alone give one SIMD ceil(T / 4) x 4.  NF whole rows, NR split rows.
template <int NF, int NR, int G>
__device__ inline void rho_tiles_bal(int s, int h, int q, int T, int lane, const int* s_cols, double* s_phi,
                                     const double* __restrict__ P, int N) {
    constexpr int PB = 64, RT = 4, NA = NF + NR;
    const int li = lane & 15, lq = lane >> 4;
    v4d accF[NF > 0 ? NF : 1][RT], accR[NR > 0 ? NR : 1];
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) accF[i][rt] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < NR; ++j) accR[j] = v4d{0.0, 0.0, 0.0, 0.0};
    auto row = [&](int i) { return i < NF ? s + 4 * (2 * i + h) : 4 * q + (i - NF); };
    int gi[NA > 0 ? NA : 1];
#pragma unroll
    for (int i = 0; i < NA; ++i) gi[i] = max(s_cols[row(i) * 16 + li], 0);
    const int ngroups = T * 4 / G;  // G (2 or 4) divides the 4T steps: every group is whole
    double aC[G][NA > 0 ? NA : 1], aN[G][NA > 0 ? NA : 1];
    auto load_a = [&](int g, double (&a)[G][NA > 0 ? NA : 1]) {
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const double* prow = P + (int64_t)max(s_cols[4 * (g * G + d) + lq], 0) * N;
#pragma unroll
            for (int i = 0; i < NA; ++i) a[d][i] = prow[gi[i]];
        }
    };
    if constexpr (NA > 0) {
        load_a(0, aN);  // as in rho_tiles
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int d = 0; d < G; ++d)
#pragma unroll
            for (int i = 0; i < NA; ++i) aC[d][i] = aN[d][i];
        for (int g = 0; g < ngroups; ++g) {
            if (g + 1 < ngroups) load_a(g + 1, aN);
#pragma unroll
            for (int d = 0; d < G; ++d) {
                {
                    const int c = 4 * (g * G + d) + lq;
                    if constexpr (NF > 0) {
                        double b[RT];
#pragma unroll
                        for (int rt = 0; rt < RT; ++rt) b[rt] = s_phi[plane_at<PB>(c, rt * 16 + li)];
#pragma unroll
                        for (int i = 0; i < NF; ++i)
#pragma unroll
                            for (int rt = 0; rt < RT; ++rt) accF[i][rt] = mfma16x16x4(aC[d][i], b[rt], accF[i][rt]);
                    }
                    if constexpr (NR > 0) {
                        const double bs = s_phi[plane_at<PB>(c, s * 16 + li)];
#pragma unroll
                        for (int j = 0; j < NR; ++j) accR[j] = mfma16x16x4(aC[d][NF + j], bs, accR[j]);
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);  // as in rho_tiles
#pragma unroll
            for (int d = 0; d < G; ++d)
#pragma unroll
                for (int i = 0; i < NA; ++i) aC[d][i] = aN[d][i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_phi[plane_at<PB>(row(i) * 16 + 4 * r + lq, rt * 16 + li)] = accF[i][rt][r];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) s_phi[plane_at<PB>((4 * q + j) * 16 + 4 * r + lq, s * 16 + li)] = accR[j][r];
}

// ---------------------------------------------------------------------------
// rho pass for PB points of a block (512 threads, 8 waves):
//   1. Phi of all ldn columns at the PB points into an LDS plane; thread
//      (point p, shell lane kk) evaluates shells kk, kk + 512/PB, ... in
//      (l, nprim) order.
//   2. X^T = P_bb Phi^T on v_mfma_f64_16x16x4, X^T in accumulators: the
//      waves form a WI x (8/WI) grid over (AO tile rows, point tiles); wave
//      (wI, wP) owns the exact number of rows T - wI over WI rounded up, so no
//      MFMA runs on padding tiles.  X^T is then stored over the Phi plane.
//   3. phi and grad phi re-evaluated with the same thread mapping and dotted
//      with X: rho = sum_j X_pj phi_pj, grad rho = 2 sum_j X_pj grad phi_pj
//      (calcRho_GGA_part, DfCalcGridX.cpp:1923-1962); the partial sums of
//      the shell lanes are added in a fixed order.
// rho_out receives (rho, grad rho) per slot -- or adds them (update mode,
// calcRho_GGA_part's accumulation, DfCalcGridX.cpp:1925-1961); the
// functional runs over all points in k_xc_points.
// The persistent kernel (AHEAD) loads the next item's set-up data while the
// current item runs: its index during step 1, its descriptor during step 2,
// its list entries at the start of step 3 and its first shell records at the
// end, one load level per phase, each thread into registers (RhoAhead; the
// LDS is the current item's).  A workgroup runs alone on its CU, so without
// this the four dependent load levels of the set-up stall the whole CU.
// ---------------------------------------------------------------------------
struct RhoAhead {
    static constexpr int kRec = 4;  // record doubles per thread held (the rest load at the item's start)
    int col, shcol, sord;           // entry tid of the column and shell lists
    int shid[kRec];                 // shell of record double tid + 512 j
    double rec[kRec];
    double pt[3];                   // point tid (tid < PB)
};

template <int PB, int WI, int NIMAX, int G, bool AHEAD>
__device__ __forceinline__ int k_rho_item(int item, const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                             const double* __restrict__ pts, int64_t n_slots,
                                             const double* __restrict__ shrec, int RS,
                                             const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                             const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                             const double* __restrict__ P, int N, double* __restrict__ rho_out,
                                             int accumulate,
        double (*s_pt)[3], double* s_dyn, RhoAhead& ah, bool have, int* queue, int n_items, int* s_next) {
    constexpr int SUBS = kBlockPts / PB;
    constexpr int NSL = 512 / PB;  // shell lanes per point
    constexpr int WP = 8 / WI;
    constexpr int RT = PB / 16 / WP;  // point tiles per wave
    static_assert(RT * 16 * WP == PB, "point tiles must cover the points");
    const int bidx = work[item / SUBS];
    const int sub = item % SUBS;
    const BlockDesc B = blocks[bidx];
    const int p0 = sub * PB;
    if (p0 >= B.npts) return -1;  // uniform: no real point in this part of the block
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldn = B.ldn, nao = B.nao, nsh = B.nshell;
    const int T = ldn >> 4;
    const double* Pb = P + B.poff;  // the block's molecule
    const int ldp = B.ldp;
    // LDS: plane [max(ldn PB, 4 x 512)] | records [nsh][RS] | cols [ldn] | shcol [nsh] | sorder [nsh]
    double* s_phi = s_dyn;
    double* s_rec = s_phi + max(ldn * PB, 4 * 512);
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + ldn;
    int* s_sord = s_shcol + nsh;
    const bool pre = AHEAD && have;  // the first entries of every list are in ah
    if (pre) {
        if (tid < ldn) s_cols[tid] = ah.col;
        if (tid < nsh) {
            s_shcol[tid] = ah.shcol;
            s_sord[tid] = ah.sord;
        }
#pragma unroll
        for (int j = 0; j < RhoAhead::kRec; ++j)
            if (tid + 512 * j < nsh * RS) s_rec[tid + 512 * j] = ah.rec[j];
        if (tid < PB) {
            s_pt[tid][0] = ah.pt[0];
            s_pt[tid][1] = ah.pt[1];
            s_pt[tid][2] = ah.pt[2];
        }
    }
    for (int c = tid + (pre ? 512 : 0); c < ldn; c += 512) s_cols[c] = (c < nao) ? bcols[B.col_begin + c] : -1;
    for (int c = tid + (pre ? 512 : 0); c < nsh; c += 512) {
        s_shcol[c] = bshell_col[B.shell_begin + c];
        s_sord[c] = bsorder[B.shell_begin + c];
    }
    for (int idx = tid + (pre ? 512 * RhoAhead::kRec : 0); idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    if (!pre && tid < PB) {
        const int64_t slot = B.pt_begin + p0 + tid;
        s_pt[tid][0] = pts[slot];
        s_pt[tid][1] = pts[n_slots + slot];
        s_pt[tid][2] = pts[2 * n_slots + slot];
    }
    __syncthreads();
    // the next item's index, taken now and published at the end of step 1
    int nxt = n_items;
    if (AHEAD && tid == 512 - 64) nxt = atomicAdd(queue, 1);

    const int p = tid & (PB - 1);
    const int kk = tid / PB;
    const double px = s_pt[p][0], py = s_pt[p][1], pz = s_pt[p][2];
    // 1. Phi plane.  The radial sums of a thread's first NC shells are kept
    // in registers for step 3 (same thread-to-shell mapping), which then
    // forms phi and grad phi for them without the exponentials (the rest are
    // recomputed).  NC covers most of a thread's share of its class's blocks
    // within the registers the class's occupancy leaves: the 64-point class of
    // 256 columns and the 16-point class of 1024 run one workgroup per CU
    // (2 waves per SIMD); the 16-point class of 512 columns runs two, and
    // keeping even 4 shells there takes it past 128 VGPRs.
    // (12 in the persistent form, whose registers also hold the next item's set-up data)
    constexpr int kColsMax = NIMAX * WI * 16;
    constexpr int NC = (PB == 64 && kColsMax == 256) ? (AHEAD ? 12 : 16) : (PB == 16 && kColsMax == 512) ? 0 : 8;
    double cR0[NC > 0 ? NC : 1], cR1[NC > 0 ? NC : 1];
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
            const int l = ((int)rec[3]) & 3;
            rec_values0_r(l, x, y, z, cR0[jj], v);
            const int c0 = s_shcol[k];
            for_components(l, [&](int m) { s_phi[plane_at<PB>(c0 + m, p)] = v[m]; });
        }
    }
    for (int j = kk + NC * NSL; j < nsh; j += NSL) {
        const int k = s_sord[j];
        double v[5];
        rec_values0(s_rec + k * RS, px, py, pz, v);
        const int c0 = s_shcol[k];
        for_components(((int)s_rec[k * RS + 3]) & 3, [&](int m) { s_phi[plane_at<PB>(c0 + m, p)] = v[m]; });
    }
    for (int it = tid; it < (ldn - nao) * PB; it += 512) s_phi[nao * PB + it] = 0.0;
    if (AHEAD && tid == 512 - 64) *s_next = nxt;
    __syncthreads();
    BlockDesc NB{};
    int np0 = 0;
    if constexpr (AHEAD) {
        nxt = __builtin_amdgcn_readfirstlane(*s_next);
        if (nxt < n_items) {
            NB = blocks[work[nxt / SUBS]];
            np0 = (nxt % SUBS) * PB;
        }
    }

    // 2. X^T = P_bb Phi^T, stored over the plane
    if constexpr (PB == 64 && WI == 8 && NIMAX == 2) {  // T <= 16: balanced over the SIMDs
        const int sd = wave & 3, h = wave >> 2, q = T >> 2;
        const int nf = (q > h) ? (q - h + 1) / 2 : 0;
        const int nr = (h == (q & 1)) ? (T & 3) : 0;
        switch (nf * 4 + nr) {
            case 0: __syncthreads(); break;  // no tiles: only the barrier before the store
            case 1: rho_tiles_bal<0, 1, 4>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 2: rho_tiles_bal<0, 2, 4>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 3: rho_tiles_bal<0, 3, 2>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 4: rho_tiles_bal<1, 0, 4>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 5: rho_tiles_bal<1, 1, 4>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 6: rho_tiles_bal<1, 2, 2>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 7: rho_tiles_bal<1, 3, 2>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 8: rho_tiles_bal<2, 0, 4>(sd, h, q, T, lane, s_cols, s_phi, Pb, ldp); break;
            default: __syncthreads(); break;  // unreachable: T <= 16 gives nf = 2 only with nr = 0
        }
    } else {
        const int wI = wave % WI, wP = wave / WI;
        const int nI = (T > wI) ? (T - wI + WI - 1) / WI : 0;
        switch (nI) {
            case 1: rho_tiles<PB, RT, 1, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp); break;
            case 2:
                if constexpr (NIMAX >= 2) rho_tiles<PB, RT, 2, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 3:
                if constexpr (NIMAX >= 3) rho_tiles<PB, RT, 3, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 4:
                if constexpr (NIMAX >= 4) rho_tiles<PB, RT, 4, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 5:
                if constexpr (NIMAX >= 5) rho_tiles<PB, RT, 5, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 6:
                if constexpr (NIMAX >= 6) rho_tiles<PB, RT, 6, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 7:
                if constexpr (NIMAX >= 7) rho_tiles<PB, RT, 7, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            case 8:
                if constexpr (NIMAX >= 8) rho_tiles<PB, RT, 8, G>(wI, WI, wP, T, lane, s_cols, s_phi, Pb, ldp);
                else __syncthreads();  // unreachable: the host keeps T <= NIMAX WI
                break;
            default: __syncthreads(); break;  // no rows: only the barrier before the store
        }
    }
    __syncthreads();

    // the next item's list entries (its descriptor has arrived during step 2)
    const bool ahead = AHEAD && nxt < n_items && np0 < NB.npts;  // uniform
    if (ahead) {
        ah.col = (tid < NB.nao) ? bcols[NB.col_begin + tid] : -1;
        ah.shcol = (tid < NB.nshell) ? bshell_col[NB.shell_begin + tid] : 0;
        ah.sord = (tid < NB.nshell) ? bsorder[NB.shell_begin + tid] : 0;
#pragma unroll
        for (int j = 0; j < RhoAhead::kRec; ++j) {
            const int idx = tid + 512 * j;
            ah.shid[j] = (idx < NB.nshell * RS) ? bshells[NB.shell_begin + idx / RS] : 0;
        }
        if (tid < PB) {
            const int64_t slot = NB.pt_begin + np0 + tid;
            ah.pt[0] = pts[slot];
            ah.pt[1] = pts[n_slots + slot];
            ah.pt[2] = pts[2 * n_slots + slot];
        }
    }

    // 3. rho and grad rho: X dotted with phi and grad phi, per shell from its
    // radial sums (kept from step 1 for the first NC shells of a thread)
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    auto shell_x = [&](int k, double* X) {  // the shell's X values at point p (components past l: 0)
        const int c0 = s_shcol[k];
        const int l = ((int)s_rec[k * RS + 3]) & 3;
        X[0] = s_phi[plane_at<PB>(c0, p)];
        X[1] = X[2] = X[3] = X[4] = 0.0;
        if (l >= 1) {
            X[1] = s_phi[plane_at<PB>(c0 + 1, p)];
            X[2] = s_phi[plane_at<PB>(c0 + 2, p)];
        }
        if (l == 2) {
            X[3] = s_phi[plane_at<PB>(c0 + 3, p)];
            X[4] = s_phi[plane_at<PB>(c0 + 4, p)];
        }
        return l;
    };
#pragma unroll
    for (int jj = 0; jj < NC; ++jj) {
        const int j = kk + jj * NSL;
        if (j < nsh) {
            const int k = s_sord[j];
            const double* rec = s_rec + k * RS;
            double X[5];
            const int l = shell_x(k, X);
            shell_dots(l, px - rec[0], py - rec[1], pz - rec[2], cR0[jj], cR1[jj], X, r0, r1, r2, r3);
        }
    }
    for (int j = kk + NC * NSL; j < nsh; j += NSL) {
        const int k = s_sord[j];
        const double* rec = s_rec + k * RS;
        const double x = px - rec[0], y = py - rec[1], z = pz - rec[2];
        double R0, R1;
        rec_radial01(rec, x, y, z, R0, R1);
        double X[5];
        const int l = shell_x(k, X);
        shell_dots(l, x, y, z, R0, R1, X, r0, r1, r2, r3);
    }
    if (ahead) {  // the next item's first shell records (their shells have arrived during step 3)
#pragma unroll
        for (int j = 0; j < RhoAhead::kRec; ++j) {
            const int idx = tid + 512 * j;
            if (idx < NB.nshell * RS) ah.rec[j] = shrec[(int64_t)ah.shid[j] * RS + (idx - (idx / RS) * RS)];
        }
    }
    __syncthreads();  // X no longer read: the plane holds the partial sums
    double* s_part = s_phi;  // [4][NSL][PB]
    s_part[(0 * NSL + kk) * PB + p] = r0;
    s_part[(1 * NSL + kk) * PB + p] = r1;
    s_part[(2 * NSL + kk) * PB + p] = r2;
    s_part[(3 * NSL + kk) * PB + p] = r3;
    __syncthreads();
    if (tid < 4 * PB) {
        const int q = tid / PB;
        const int pp = tid & (PB - 1);
        double s = 0.0;
        for (int k2 = 0; k2 < NSL; ++k2) s += s_part[(q * NSL + k2) * PB + pp];
        const int64_t slot = B.pt_begin + p0 + pp;
        const double v = (q == 0) ? s : 2.0 * s;
        rho_out[4 * slot + q] = accumulate ? rho_out[4 * slot + q] + v : v;  // update mode adds rho[dP]
    }
    return ahead ? nxt : (AHEAD && nxt >= n_items ? n_items : -1);
}

template <int PB, int WI, int NIMAX, int G, bool PERSIST>
__global__ __launch_bounds__(512) void k_rho(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                             const double* __restrict__ pts, int64_t n_slots,
                                             const double* __restrict__ shrec, int RS,
                                             const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                             const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                             const double* __restrict__ P, int N, double* __restrict__ rho_out,
                                             int accumulate, int* __restrict__ queue, int n_items) {
    __shared__ double s_pt[PB][3];
    extern __shared__ double s_dyn[];
    RhoAhead ah;
    if constexpr (PERSIST) {
        // an item returns the next one with its set-up data in ah, n_items when
        // the queue is empty, or -1 when it took no next item (an empty part of
        // a block, or the next one empty): then the queue is asked here
        __shared__ int s_item, s_next;
        int item = next_item(queue, n_items, 0, &s_item);
        bool have = false;
        while (item < n_items) {
            const int nxt = k_rho_item<PB, WI, NIMAX, G, true>(item, work, blocks, pts, n_slots, shrec, RS, bshells,
                                                               bshell_col, bcols, bsorder, P, N, rho_out, accumulate,
                                                               s_pt, s_dyn, ah, have, queue, n_items, &s_next);
            have = nxt >= 0 && nxt < n_items;
            item = nxt >= 0 ? nxt : next_item(queue, n_items, 1, &s_item);
        }
    } else {
        k_rho_item<PB, WI, NIMAX, G, false>((int)blockIdx.x, work, blocks, pts, n_slots, shrec, RS, bshells, bshell_col,
                                            bcols, bsorder, P, N, rho_out, accumulate, s_pt, s_dyn, ah, false, nullptr,
                                            n_items, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Functional point terms for every slot (buildVxc, DfCalcGridX.cpp:2137-2193
// for GGA and :2056-2091 for LDA, and the coefficients of build_XC_Matrix,
// :2353-2368 and :2455-2500): coef = (w vrho / 2, w (2 vgAA + vgAB) grad rho),
// zero where rho <= the density cut-off or w = 0 (padding slots).  For the
// LDA functionals vgAA = vgAB = 0, so the gradient part of coef is exactly
// zero and K = sum_g w vrho phi phi, the LDA matrix.  Per-workgroup sums of
// w e_xc in a fixed order, and w e_xc per slot when ewf is given.
// ---------------------------------------------------------------------------
template <int XC>
__global__ __launch_bounds__(256) void k_xc_points(int64_t n_slots, const double* __restrict__ pts,
                                                   const double* __restrict__ rho, double* __restrict__ coef,
                                                   double* __restrict__ epart, double* __restrict__ ewf,
                                                   double dens_cut) {
    __shared__ double s_e[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < n_slots) {
        const double w = pts[3 * n_slots + i];
        const double r = rho[4 * i], gx = rho[4 * i + 1], gy = rho[4 * i + 2], gz = rho[4 * i + 3];
        double a = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        if (w != 0.0 && r > dens_cut) {
            const double gamma = gx * gx + gy * gy + gz * gz;
            const XcPoint f = eval_rks(XC, r, gamma);
            a = 0.5 * (w * f.vr);
            const double c2 = 2.0 * f.vgaa;
            cx = w * (c2 * gx + f.vgab * gx);
            cy = w * (c2 * gy + f.vgab * gy);
            cz = w * (c2 * gz + f.vgab * gz);
            e = w * f.e;
        }
        coef[4 * i + 0] = a;
        coef[4 * i + 1] = cx;
        coef[4 * i + 2] = cy;
        coef[4 * i + 3] = cz;
        if (ewf) ewf[i] = e;
    }
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s_e[threadIdx.x] += s_e[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) epart[blockIdx.x] = s_e[0];
}

// UKS point terms (buildVxc UKS, DfCalcGridX.cpp:2093-2135 LDA, :2195-2278
// GGA): a slot counts when rhoA or rhoB exceeds the cut-off; coef_a =
// (w vrhoA / 2, w (2 vgAA + vgAB) grad rhoA), coef_b = (w vrhoB / 2,
// w (2 vgBB + vgAB) grad rhoB) -- build_XC_Matrix multiplies vgAB by the
// same-spin gradient (:2455-2470).
template <int XC>
__global__ __launch_bounds__(256) void k_xc_points_uks(int64_t n_slots, const double* __restrict__ pts,
                                                       const double* __restrict__ rho_a,
                                                       const double* __restrict__ rho_b,
                                                       double* __restrict__ coef_a, double* __restrict__ coef_b,
                                                       double* __restrict__ epart, double dens_cut) {
    __shared__ double s_e[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < n_slots) {
        const double w = pts[3 * n_slots + i];
        const double ra = rho_a[4 * i], ax = rho_a[4 * i + 1], ay = rho_a[4 * i + 2], az = rho_a[4 * i + 3];
        const double rb = rho_b[4 * i], bx = rho_b[4 * i + 1], by = rho_b[4 * i + 2], bz = rho_b[4 * i + 3];
        double ca[4] = {0.0, 0.0, 0.0, 0.0}, cb[4] = {0.0, 0.0, 0.0, 0.0};
        if (w != 0.0 && (ra > dens_cut || rb > dens_cut)) {
            const double gaa = ax * ax + ay * ay + az * az;
            const double gab = ax * bx + ay * by + az * bz;
            const double gbb = bx * bx + by * by + bz * bz;
            const XcPointU f = eval_uks(XC, ra, rb, gaa, gab, gbb);
            const double fa = 2.0 * f.vgaa + f.vgab, fb = 2.0 * f.vgbb + f.vgab;
            ca[0] = 0.5 * (w * f.va);
            ca[1] = w * fa * ax;
            ca[2] = w * fa * ay;
            ca[3] = w * fa * az;
            cb[0] = 0.5 * (w * f.vb);
            cb[1] = w * fb * bx;
            cb[2] = w * fb * by;
            cb[3] = w * fb * bz;
            e = w * f.e;
        }
        for (int q = 0; q < 4; ++q) {
            coef_a[4 * i + q] = ca[q];
            coef_b[4 * i + q] = cb[q];
        }
    }
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s_e[threadIdx.x] += s_e[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) epart[blockIdx.x] = s_e[0];
}

// ---------------------------------------------------------------------------
// K pass consumer: NT accumulator tiles of K_bb (a contiguous run [tb, te)
// of lower tiles, t = I (I + 1) / 2 + J, rounded up to a multiple of 4 with
// repeats of the last tile that are not scattered) over the point slabs the
// producer waves fill.  Slab element (point p, column c) sits at
// p LD + (c ^ (p & 15)): the XOR keeps the 16 lanes of an operand read on
// one 128-byte run while spreading a producer's column writes over the banks.
// ---------------------------------------------------------------------------
template <int NT>
__device__ inline void kmat_consume(int tb, int te, int nslab, int PS, int LD, int SLAB, int lane,
                                    const double* s_buf, const int* s_cols, double* __restrict__ K, int N) {
    const int li = lane & 15, lq = lane >> 4;
    int offI[NT], offJ[NT];
    {
        int I = 0;
        while ((I + 1) * (I + 2) / 2 <= tb) ++I;
        int J = tb - I * (I + 1) / 2;
#pragma unroll
        for (int m = 0; m < NT; ++m) {
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
    v4d acc[NT];
#pragma unroll
    for (int m = 0; m < NT; ++m) acc[m] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
        const double* cur = s_buf + (s & 1) * 2 * SLAB;
        for (int ks = 0; ks < PS / 4; ++ks) {
            const int p = 4 * ks + lq;
            const double* fr = cur + p * LD + (li ^ (p & 15));
            const double* vr = fr + SLAB;
            // groups of 4 tiles: bounded operand lookahead keeps the
            // accumulators and operands within the register budget
#pragma unroll
            for (int m0 = 0; m0 < NT; m0 += 4) {
#pragma unroll
                for (int m = m0; m < m0 + 4; ++m) {
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
    for (int m = 0; m < NT; ++m) {
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

// ---------------------------------------------------------------------------
// K pass: K_bb += Phi^T V + V^T Phi for the lower 16 x 16 tiles t0 <= t < t1
// (at most 96) of one block.  Twelve waves: waves 0-7 (consumers, two per
// SIMD) each hold a contiguous eighth of the tiles in accumulators and run
// the MFMAs on point slab s while waves 8-11 (producers) evaluate slab s + 1
// into the other
// half of a double-buffered LDS area: phi and grad phi of the needed columns
// from the LDS shell records, and V = (w vrho / 2) phi + c . grad phi with
// the per-point coefficients (build_XC_Matrix GGA, DfCalcGridX.cpp:2455-2500).
// Tiles go into the dense lower triangle of K with FP64 atomics.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_kmat_item(int item, const KWork* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                              const double* __restrict__ pts, int64_t n_slots,
                                              const double* __restrict__ shrec, int RS,
                                              const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                              const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                              const double* __restrict__ coef, double* __restrict__ K, int N,
        double (*s_pt)[3], double (*s_cf)[4], double* s_dynd) {
    const KWork W = work[item];
    const BlockDesc B = blocks[W.block];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool producer = wave >= 8;
    const int ncols = W.ncols;  // multiple of 16, <= ldn
    const int LD = (ncols & 31) ? ncols : ncols + 16;
    const int PS = W.ps;
    const int ps_log = 31 - __builtin_clz(PS);
    const int nsh = W.nsh;
    const int SLAB = PS * LD;  // doubles per array per buffer
    // LDS: [2 buffers][phi, V][PS][LD] | records [nsh][RS] | cols [ncols] | shcol [nsh] | shid [nsh]
    double* s_buf = s_dynd;
    double* s_rec = s_buf + 4 * SLAB;
    int* s_cols = reinterpret_cast<int*>(s_rec + (size_t)nsh * RS);
    int* s_shcol = s_cols + ncols;
    int* s_shid = s_shcol + nsh;
    for (int c = tid; c < ncols; c += 768) s_cols[c] = (c < B.nao) ? bcols[B.col_begin + c] : -1;
    // shells this item needs (local index < nsh), in (l, nprim) order, compacted by wave 8
    if (wave == 8) {
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
    for (int p = tid; p < kBlockPts; p += 768) {
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
    for (int idx = tid; idx < nsh * RS; idx += 768) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)s_shid[k] * RS + (idx - k * RS)];
    }
    __syncthreads();

    const int ptid = tid - 512;  // producer thread index
    auto produce = [&](int pb, double* phi, double* vv) {
        for (int it = ptid; it < nsh * PS; it += 256) {
            const int p = it & (PS - 1);  // PS is a power of two (16, 8 or 4)
            const int k = it >> ps_log;
            const int pp = pb + p;
            const double* rec = s_rec + k * RS;
            const double x = s_pt[pp][0] - rec[0], y = s_pt[pp][1] - rec[1], z = s_pt[pp][2] - rec[2];
            double R0, R1;
            rec_radial01(rec, x, y, z, R0, R1);
            const int c0 = s_shcol[k];
            const int sw = p & 15;
            shell_phi_v(((int)rec[3]) & 3, x, y, z, R0, R1, s_cf[pp][0], s_cf[pp][1], s_cf[pp][2], s_cf[pp][3],
                        [&](int m, double v, double V) {
                            const int c = c0 + m;
                            if (c < ncols) {
                                phi[p * LD + (c ^ sw)] = v;
                                vv[p * LD + (c ^ sw)] = V;
                            }
                        });
        }
        for (int it = ptid; it < (ncols - min(ncols, B.nao)) * PS; it += 256) {
            const int c = B.nao + (it >> ps_log);
            const int p = it & (PS - 1);
            phi[p * LD + (c ^ (p & 15))] = 0.0;
            vv[p * LD + (c ^ (p & 15))] = 0.0;
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
        return;  // to the next item's barrier, which the consumers reach after their scatter
    }
    // consumers: a contiguous run [tb, te) of the item's tiles, dispatched on
    // the run length rounded up to 4 (one barrier before the slab loop, one per slab)
    const int q8 = (W.t1 - W.t0 + 7) >> 3;
    const int tstart = W.t0 + wave * q8;
    const int tend = min(W.t1, tstart + q8);
    const bool has = tstart < tend;
    const int tb = has ? tstart : W.t1 - 1;  // a wave without tiles keeps valid offsets
    const int te = has ? tend : tb;          // and scatters nothing
    const int ngrp = has ? (tend - tstart + 3) >> 2 : 1;
    switch (ngrp) {
        case 1: kmat_consume<4>(tb, te, nslab, PS, LD, SLAB, lane, s_buf, s_cols, K + B.poff, B.ldp); break;
        case 2: kmat_consume<8>(tb, te, nslab, PS, LD, SLAB, lane, s_buf, s_cols, K + B.poff, B.ldp); break;
        default: kmat_consume<12>(tb, te, nslab, PS, LD, SLAB, lane, s_buf, s_cols, K + B.poff, B.ldp); break;
    }
}

template <bool PERSIST>
__global__ __launch_bounds__(768) void k_kmat(const KWork* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                              const double* __restrict__ pts, int64_t n_slots,
                                              const double* __restrict__ shrec, int RS,
                                              const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                              const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                              const double* __restrict__ coef, double* __restrict__ K, int N,
                                              int* __restrict__ queue, int n_items) {
    __shared__ double s_pt[kBlockPts][3];
    __shared__ double s_cf[kBlockPts][4];
    extern __shared__ double s_dynd[];
    if constexpr (PERSIST) {
        __shared__ int s_item;
        for (int iter = 0;; ++iter) {
            const int item = next_item(queue, n_items, iter, &s_item);
            if (item >= n_items) break;
            k_kmat_item(item, work, blocks, pts, n_slots, shrec, RS, bshells, bshell_col, bcols, bsorder, coef, K, N, s_pt, s_cf, s_dynd);
        }
    } else {
        k_kmat_item((int)blockIdx.x, work, blocks, pts, n_slots, shrec, RS, bshells, bshell_col, bcols, bsorder, coef, K, N, s_pt, s_cf, s_dynd);
    }
}

// deterministic sum of the per-block energies (one workgroup)
// E_xc of molecule blockIdx.x: the partial sums [bounds[m], bounds[m + 1]) in a
// fixed order (one molecule: bounds = {0, n})
__global__ __launch_bounds__(256) void k_energy_sum(const double* __restrict__ eb, const int* __restrict__ bounds,
                                                    int n, double* __restrict__ out) {
    __shared__ double s[256];
    const int lo = bounds ? bounds[blockIdx.x] : 0, hi = bounds ? bounds[blockIdx.x + 1] : n;
    double v = 0.0;
    for (int i = lo + threadIdx.x; i < hi; i += 256) v += eb[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(64) void k_zero(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <int XC>
__global__ __launch_bounds__(256) void k_eval_functional(int64_t n, const double* __restrict__ in,
                                                         double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const XcPoint f = eval_rks(XC, in[2 * i], in[2 * i + 1]);
    out[4 * i + 0] = f.e;
    out[4 * i + 1] = f.vr;
    out[4 * i + 2] = f.vgaa;
    out[4 * i + 3] = f.vgab;
}

template <int XC>
__global__ __launch_bounds__(256) void k_eval_functional_uks(int64_t n, const double* __restrict__ in,
                                                             double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* x = in + 5 * i;
    const XcPointU f = eval_uks(XC, x[0], x[1], x[2], x[3], x[4]);
    double* o = out + 6 * i;
    o[0] = f.e;
    o[1] = f.va;
    o[2] = f.vb;
    o[3] = f.vgaa;
    o[4] = f.vgab;
    o[5] = f.vgbb;
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
// Partition weight of each point's atom (calcMultiCenterWeight_Becke /
// Ps_uij, DfGenerateGrid.cpp:1811-1876, :2043-2086; calcMultiCenterWeight_SS,
// :1878-1956): w *= P_A / sum_m P_m with P_m = prod_{k != m} s(u_mk),
// u = (r_m - r_k) / R_mk (Becke: nu = u + a_mk (1 - u^2) through f3).
// atoms: (x, y, z, Z) per atom | 1/R_mk (symmetric) | a_mk stored at [k][m].
// Every atom m, a dummy (Z = 0) included, has a cell; the Becke product skips
// dummy neighbours k, the SS product does not.  One wave per point: the
// point's atom distances in LDS, lanes over the cells m (so the [k][m] reads
// of a neighbour k are contiguous), the sums over cells in a fixed butterfly.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_becke(const double* __restrict__ atoms, int na,
                                              int64_t n, const double* __restrict__ pts,
                                              const int* __restrict__ atom_of_pt,
                                              double* __restrict__ w, int ss) {
    extern __shared__ double s_r[];  // [waves][na]
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
    const bool valid = i0 < n;
    const int64_t i = valid ? i0 : n - 1;  // a spare wave repeats the last point and writes nothing
    const double* inv_d = atoms + 4 * na;
    const double* amat_t = inv_d + (int64_t)na * na;
    double* r = s_r + (int64_t)wv * na;
    const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    for (int m = lane; m < na; m += 64) {
        const double dx = x - atoms[4 * m], dy = y - atoms[4 * m + 1], dz = z - atoms[4 * m + 2];
        r[m] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    __syncthreads();
    const int A = atom_of_pt[i];
    // the point's own atom is the first neighbour of every cell: it is (near)
    // the nearest, so a far cell's product reaches zero (Stratmann-Scuseria:
    // exactly, u > 0.64) or underflows after few factors and the loop ends
    const int A0 = (A >= 0 && A < na) ? A : 0;
    double total = 0.0, mine = 0.0;
    for (int m = lane; m < na; m += 64) {
        const double rm = r[m];
        double ps = 1.0;
        for (int j = 0; j < na && ps != 0.0; ++j) {
            const int k = (j == 0) ? A0 : (j <= A0 ? j - 1 : j);  // A0, then the others in index order
            if (k == m) continue;
            double u = (rm - r[k]) * inv_d[(int64_t)k * na + m];
            double g;
            if (ss) {
                // Stratmann-Scuseria g(u), a = 0.64
                if (u < -0.64) {
                    g = -1.0;
                } else if (u > 0.64) {
                    g = 1.0;
                } else {
                    const double q = u / 0.64, q2 = q * q, q3 = q2 * q, q5 = q2 * q3;
                    g = (1.0 / 16.0) * (35.0 * q - 35.0 * q3 + 21.0 * q5 - 5.0 * q5 * q2);
                }
            } else {
                if (atoms[4 * k + 3] <= 0.0) continue;  // Becke: dummy neighbours do not cut
                const double a = amat_t[(int64_t)k * na + m];
                u = u + a * (1.0 - u * u);
                g = 0.5 * u * (3.0 - u * u);
                g = 0.5 * g * (3.0 - g * g);
                g = 0.5 * g * (3.0 - g * g);
            }
            ps *= 0.5 * (1.0 - g);
        }
        if (m == A) mine = ps;
        total += ps;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        total += __shfl_xor(total, off);
        mine += __shfl_xor(mine, off);
    }
    if (valid && lane == 0) w[i] *= (mine / total);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return PDFXC_OK;
}

template <int PB, int WI, int NIMAX, int G>
static int rho_one(Context& c, const LaunchRange& r) {
    const int items = r.count * (kBlockPts / PB);
    unsigned grid = (unsigned)items;
    // the persistent form (the next item's set-up loaded during the current
    // one) only for the classes whose widest plane takes more than half of the
    // LDS: they run one workgroup per CU either way, while the others run two
    // as plain launches and would lose one to the persistent form's registers
    constexpr bool kAlone = (size_t)PB * NIMAX * WI * 16 * sizeof(double) > 80 * 1024;
    if constexpr (kAlone) {
        int* q = launch_queue(c, kQueueRho, items, resident_per_cu(&k_rho<PB, WI, NIMAX, G, true>, 512, r.lds),
                              &grid);
        if (q) {
            hipLaunchKernelGGL((k_rho<PB, WI, NIMAX, G, true>), dim3(grid), dim3(512), r.lds, c.stream,
                               c.d_rho_work + r.offset, c.d_blocks, c.d_pts, c.n_slots, c.d_shrec, c.rec_stride,
                               c.d_block_shells, c.d_block_shell_col, c.d_block_cols, c.d_block_sorder, c.d_P,
                               c.n_aos, c.d_rho, c.rho_accumulate, q, items);
            return check_launch("k_rho");
        }
    }
    hipLaunchKernelGGL((k_rho<PB, WI, NIMAX, G, false>), dim3(grid), dim3(512), r.lds, c.stream,
                       c.d_rho_work + r.offset, c.d_blocks, c.d_pts, c.n_slots, c.d_shrec, c.rec_stride,
                       c.d_block_shells, c.d_block_shell_col, c.d_block_cols, c.d_block_sorder, c.d_P, c.n_aos,
                       c.d_rho, c.rho_accumulate, nullptr, items);
    return check_launch("k_rho");
}

// k_rho classes by padded block width ldn: <= 128, 256, 512, 1024, and class
// 4, ldn <= 256 at 64 points per workgroup (half the workgroups per block, so
// each block's P_bb and shell records are read half as often)
int rho_points_per_workgroup(int cls) { return (cls == 0 || cls == 4) ? 64 : cls == 1 ? 32 : 16; }

int rho64_max_dyn_lds(size_t* out) {
    hipFuncAttributes a, b;  // the plain and the persistent form (a few bytes more of static LDS)
    hipError_t e = hipFuncGetAttributes(&a, reinterpret_cast<const void*>(&k_rho<64, 8, 2, 4, false>));
    if (e == hipSuccess) e = hipFuncGetAttributes(&b, reinterpret_cast<const void*>(&k_rho<64, 8, 2, 4, true>));
    if (e != hipSuccess) return hip_fail(e, "hipFuncGetAttributes(k_rho)");
    *out = (size_t)160 * 1024 - std::max(a.sharedSizeBytes, b.sharedSizeBytes);
    return PDFXC_OK;
}

int launch_rho(Context& c, const LaunchRange& r) {
    if (r.count <= 0) return PDFXC_OK;
    switch (r.cls) {
        case 0: return rho_one<64, 4, 2, 4>(c, r);
        case 1: return rho_one<32, 8, 2, 4>(c, r);
        case 2: return rho_one<16, 8, 4, 4>(c, r);
        case 3: return rho_one<16, 8, 8, 2>(c, r);
        case 4: return rho_one<64, 8, 2, 4>(c, r);
        default: set_error("k_rho: bad class"); return PDFXC_ERR_ARGUMENT;
    }
}

template <int XC>
static void xc_points_one(Context& c, double* d_ewf) {
    const int64_t g = (c.n_slots + 255) / 256;
    hipLaunchKernelGGL(k_xc_points<XC>, dim3((unsigned)g), dim3(256), 0, c.stream, c.n_slots, c.d_pts,
                       c.d_rho, c.d_coef, c.d_eblock, d_ewf, c.opts.density_cutoff);
}

int launch_xc_points(Context& c, int xc_type, double* d_ewf) {
    switch (xc_type) {
        case kHFS: xc_points_one<kHFS>(c, d_ewf); break;
        case kSVWN: xc_points_one<kSVWN>(c, d_ewf); break;
        case kHFB: xc_points_one<kHFB>(c, d_ewf); break;
        case kBLYP: xc_points_one<kBLYP>(c, d_ewf); break;
        case kB3LYP: xc_points_one<kB3LYP>(c, d_ewf); break;
        default: set_error("k_xc_points: no grid functional " + std::to_string(xc_type)); return PDFXC_ERR_UNSUPPORTED;
    }
    return check_launch("k_xc_points");
}

template <int XC>
static void xc_points_uks_one(Context& c) {
    const int64_t g = (c.n_slots + 255) / 256;
    hipLaunchKernelGGL(k_xc_points_uks<XC>, dim3((unsigned)g), dim3(256), 0, c.stream, c.n_slots, c.d_pts,
                       c.d_rho, c.d_rho_b, c.d_coef, c.d_coef_b, c.d_eblock, c.opts.density_cutoff);
}

int launch_xc_points_uks(Context& c, int xc_type) {
    switch (xc_type) {
        case kHFS: xc_points_uks_one<kHFS>(c); break;
        case kSVWN: xc_points_uks_one<kSVWN>(c); break;
        case kHFB: xc_points_uks_one<kHFB>(c); break;
        case kBLYP: xc_points_uks_one<kBLYP>(c); break;
        case kB3LYP: xc_points_uks_one<kB3LYP>(c); break;
        default: set_error("k_xc_points_uks: no grid functional " + std::to_string(xc_type)); return PDFXC_ERR_UNSUPPORTED;
    }
    return check_launch("k_xc_points_uks");
}

static int kmat_one(Context& c, const LaunchRange& r) {
    unsigned grid = 0;
    int* q = launch_queue(c, kQueueK, r.count, resident_per_cu(&k_kmat<true>, 768, r.lds), &grid);
    if (q)
        hipLaunchKernelGGL(k_kmat<true>, dim3(grid), dim3(768), r.lds, c.stream,
                           reinterpret_cast<const KWork*>(c.d_k_work) + r.offset, c.d_blocks, c.d_pts,
                           c.n_slots, c.d_shrec, c.rec_stride, c.d_block_shells, c.d_block_shell_col,
                           c.d_block_cols, c.d_block_sorder, c.d_coef, c.d_K, c.n_aos, q, r.count);
    else
        hipLaunchKernelGGL(k_kmat<false>, dim3(grid), dim3(768), r.lds, c.stream,
                           reinterpret_cast<const KWork*>(c.d_k_work) + r.offset, c.d_blocks, c.d_pts,
                           c.n_slots, c.d_shrec, c.rec_stride, c.d_block_shells, c.d_block_shell_col,
                           c.d_block_cols, c.d_block_sorder, c.d_coef, c.d_K, c.n_aos, q, r.count);
    return check_launch("k_kmat");
}

int launch_kmat(Context& c, const LaunchRange& r) {
    if (r.count <= 0) return PDFXC_OK;
    switch (r.cls) {
        case 0:
        case 1: return kmat_one(c, r);
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
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(256), 0, c.stream, c.d_eblock, (const int*)nullptr,
                       (int)((c.n_slots + 255) / 256), dE);
    return check_launch("k_energy_sum");
}

int launch_zero(double* d, int64_t n, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, d, n);
    return check_launch("k_zero");
}

__global__ __launch_bounds__(64) void k_zero_int(int* __restrict__ p, int n) {
    const int i = (int)blockIdx.x * 64 + (int)threadIdx.x;
    if (i < n) p[i] = 0;
}

int begin_queues(Context& c) {
    c.n_queue_used = 0;
    c.queue_open = false;
    if (!c.d_queue || !c.tune.persistent) return PDFXC_OK;  // no persistent kernel in this sequence
    // a fill kernel, not a memset: the sequence may be captured in a graph
    hipLaunchKernelGGL(k_zero_int, dim3((kMaxQueues + 63) / 64), dim3(64), 0, c.stream, c.d_queue, kMaxQueues);
    if (int rc = check_launch("k_zero_int")) return rc;
    c.queue_open = true;
    return PDFXC_OK;
}

int* launch_queue(Context& c, int kind, int n_items, int per_cu, unsigned* grid) {
    const int64_t resident = (int64_t)std::max(per_cu, 1) * std::max(c.n_cu, 1);
    if (!c.queue_open || !(c.tune.persistent & kind) || c.n_queue_used >= kMaxQueues || n_items <= resident) {
        *grid = (unsigned)n_items;
        return nullptr;
    }
    *grid = (unsigned)resident;
    return c.d_queue + c.n_queue_used++;
}

int launch_energy_sums(Context& c, const int* d_bounds, int n_mol, double* dE) {
    hipLaunchKernelGGL(k_energy_sum, dim3(n_mol), dim3(256), 0, c.stream, c.d_eblock, d_bounds, 0, dE);
    return check_launch("k_energy_sum");
}

int launch_eval_functional(int xc_type, int64_t n, const double* d_in, double* d_out, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    const dim3 g((unsigned)((n + 255) / 256));
    switch (xc_type) {
        case kHFS: hipLaunchKernelGGL(k_eval_functional<kHFS>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kSVWN: hipLaunchKernelGGL(k_eval_functional<kSVWN>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kHFB: hipLaunchKernelGGL(k_eval_functional<kHFB>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kBLYP: hipLaunchKernelGGL(k_eval_functional<kBLYP>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kB3LYP: hipLaunchKernelGGL(k_eval_functional<kB3LYP>, g, dim3(256), 0, s, n, d_in, d_out); break;
        default: set_error("k_eval_functional: no grid functional " + std::to_string(xc_type)); return PDFXC_ERR_UNSUPPORTED;
    }
    return check_launch("k_eval_functional");
}

// ---------------------------------------------------------------------------
// Staging that the library does not allocate: pdfxc_eval_functional_noalloc
// copies its inputs (n x 2) and outputs (n x 4) through this array, one copy
// per GPU held for the life of the process.  Host code never passes the
// array's name to a kernel or a copy (on the host it is only a stand-in for
// the symbol): the device address comes from hipGetSymbolAddress, looked up
// once per device on that device's first call and kept.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMax = PDFXC_NOALLOC_MAX_POINTS;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[6 * kNoAllocMax];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

static int scratch_on_current_device(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "pdfxc_eval_functional_noalloc: hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("pdfxc_eval_functional_noalloc: device index " + std::to_string(dev) + " beyond " +
                  std::to_string(kNoAllocMaxDevices));
        return PDFXC_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "pdfxc_eval_functional_noalloc: hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return PDFXC_OK;
}

int eval_functional_noalloc(int xc_type, int64_t n, const double* in, double* out) {
    if (n < 0 || n > kNoAllocMax) {
        set_error("pdfxc_eval_functional_noalloc: " + std::to_string(n) + " points; this entry point takes 0 to " +
                  std::to_string(kNoAllocMax) + " (PDFXC_NOALLOC_MAX_POINTS)");
        return PDFXC_ERR_ARGUMENT;
    }
    if (n > 0 && (!in || !out)) {
        set_error("pdfxc_eval_functional_noalloc: null input or output");
        return PDFXC_ERR_ARGUMENT;
    }
    if (n == 0) return PDFXC_OK;
    if (xc_type == kHF) {
        std::fill(out, out + 4 * n, 0.0);
        return PDFXC_OK;
    }
    double* s = nullptr;
    if (int rc = scratch_on_current_device(&s)) return rc;
    double* s_in = s;                  // [n][2]
    double* s_out = s + 2 * kNoAllocMax;  // [n][4]
    PDFXC_HIP(hipMemcpy(s_in, in, 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    if (int rc = launch_eval_functional(xc_type, n, s_in, s_out, nullptr)) return rc;
    PDFXC_HIP(hipMemcpy(out, s_out, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return PDFXC_OK;
}

int launch_eval_functional_uks(int xc_type, int64_t n, const double* d_in, double* d_out, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    const dim3 g((unsigned)((n + 255) / 256));
    switch (xc_type) {
        case kHFS: hipLaunchKernelGGL(k_eval_functional_uks<kHFS>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kSVWN: hipLaunchKernelGGL(k_eval_functional_uks<kSVWN>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kHFB: hipLaunchKernelGGL(k_eval_functional_uks<kHFB>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kBLYP: hipLaunchKernelGGL(k_eval_functional_uks<kBLYP>, g, dim3(256), 0, s, n, d_in, d_out); break;
        case kB3LYP: hipLaunchKernelGGL(k_eval_functional_uks<kB3LYP>, g, dim3(256), 0, s, n, d_in, d_out); break;
        default: set_error("k_eval_functional_uks: no grid functional " + std::to_string(xc_type)); return PDFXC_ERR_UNSUPPORTED;
    }
    return check_launch("k_eval_functional_uks");
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
                 const int* d_atom_of_pt, double* d_w, int partition, hipStream_t s) {
    if (n <= 0) return PDFXC_OK;
    // one wave per point, each wave's point-to-atom distances in LDS: four
    // waves per workgroup while they fit in 64 KiB, else one
    if ((size_t)n_atoms * sizeof(double) > 65536) {
        set_error("Becke weights: too many atoms for the LDS distance table");
        return PDFXC_ERR_UNSUPPORTED;
    }
    const int waves = ((size_t)4 * n_atoms * sizeof(double) <= 65536) ? 4 : 1;
    const int64_t g = (n + waves - 1) / waves;
    const size_t lds = sizeof(double) * (size_t)waves * (size_t)n_atoms;
    hipLaunchKernelGGL(k_becke, dim3((unsigned)g), dim3(64 * waves), lds, s, d_atoms, n_atoms, n,
                       d_pts, d_atom_of_pt, d_w, partition == 1 ? 1 : 0);
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
