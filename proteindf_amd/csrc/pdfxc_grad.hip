// XC nuclear gradient kernels, RKS GGA, FP64 (energyGradient /
// energyGradient_part, DfCalcGridX.cpp:1215-1415).
//
// With X = P phi, Y_d = P d(phi)/dr_d and the per-point coefficients of the
// K pass (a = w vrho / 2, c_d = w (2 vgAA + vgAB) drho_d), the derivative of
// E_xc with respect to the centre B of AO i, grid point held fixed, is
//   dE/dR_B = -4 sum_points sum_{i on B} [ grad phi_i (2 a X_i + Z_i)
//                                          + X_i sum_d c_d d(grad phi_i)/dr_d ]
// where Z = P W, W = sum_d c_d d(phi)/dr_d -- the three Y_d products of the
// reference fold into one.  The point moves with its own atom A, which takes
// the opposite sum (:1372-1377); the factor 4 is the reference's -2 w times
// its RKS doubling (:1234).
//   k_grad_fn  per 16 points of a block: phi and W into LDS, X^T = P_bb phi^T
//              and Z^T = P_bb W^T on v_mfma_f64_16x16x4, then grad phi and the
//              AO second derivatives per (point, shell), summed per atom in LDS
//   k_grad_w   per point: grad_B omega_A of the grid's partition (Becke with
//              or without size adjustments, or Stratmann-Scuseria)
//              for every atom B (point fixed), times w_single e_xc, to B and
//              with opposite sign to A (:1386-1404).  The reference's JGP_t
//              (DfGenerateGrid.cpp:2355-2372) returns 0 for every |mu| <= 1,
//              so its weight term vanishes; this kernel computes the exact one.
#include <hip/hip_runtime.h>

#include "pdfxc_device.hpp"
#include "pdfxc_internal.hpp"
#include "pdfxc_split.hpp"

namespace pdfxc {

template <int NI>
__device__ inline void grad_tiles(int wave, int lane, int T, const int* s_cols, double* s_phi, double* s_w,
                                  const double* __restrict__ P, int N) {
    constexpr int G = 4;
    const int li = lane & 15, lq = lane >> 4;
    v4d acc0[NI], acc1[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        acc0[i] = v4d{0.0, 0.0, 0.0, 0.0};
        acc1[i] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    int gi[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) gi[i] = max(s_cols[(wave + 8 * i) * 16 + li], 0);
    const int nsteps = T * 4;
    const int ngroups = (nsteps + G - 1) / G;
    double aC[G][NI], aN[G][NI];
    auto load_a = [&](int g, double (&a)[G][NI]) {
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const int step = min(g * G + d, nsteps - 1);
            const double* prow = P + (int64_t)max(s_cols[4 * step + lq], 0) * N;
#pragma unroll
            for (int i = 0; i < NI; ++i) a[d][i] = prow[gi[i]];
        }
    };
    load_a(0, aC);
    for (int g = 0; g < ngroups; ++g) {
        if (g + 1 < ngroups) load_a(g + 1, aN);
#pragma unroll
        for (int d = 0; d < G; ++d) {
            const int step = g * G + d;
            if (step < nsteps) {
                const int o = (4 * step + lq) * 16 + li;
                const double b0 = s_phi[o], b1 = s_w[o];
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    acc0[i] = mfma16x16x4(aC[d][i], b0, acc0[i]);
                    acc1[i] = mfma16x16x4(aC[d][i], b1, acc1[i]);
                }
            }
        }
#pragma unroll
        for (int d = 0; d < G; ++d)
#pragma unroll
            for (int i = 0; i < NI; ++i) aC[d][i] = aN[d][i];
    }
    __syncthreads();  // every wave has stopped reading the planes
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = ((wave + 8 * i) * 16 + 4 * r + lq) * 16 + li;
            s_phi[o] = acc0[i][r];
            s_w[o] = acc1[i][r];
        }
}

// The same X^T and Z^T tiles on v_mfma_f32_16x16x32_bf16 with S bf16 slices
// per operand (P sliced as it is loaded, phi and W as they are read from the
// planes; padding columns are zero): each 32-deep K step sums its hi x hi
// products and its slice corrections in two FP32 accumulators from zero and
// adds both into FP64, as the split rho pass does.  Register r of lane l
// holds D[4 (l / 16) + r][l % 16] (AO row, point).
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <int S, int NI>
__device__ inline void grad_tiles_split(int wave, int lane, int ldn, const int* s_cols, double* s_phi, double* s_w,
                                        const double* __restrict__ P, int N) {
    const int li = lane & 15, lq = lane >> 4;
    v4d acc0[NI], acc1[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        acc0[i] = v4d{0.0, 0.0, 0.0, 0.0};
        acc1[i] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    int gi[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) gi[i] = max(s_cols[(wave + 8 * i) * 16 + li], 0);
    const int nks = (ldn + 31) / 32;
    for (int ks = 0; ks < nks; ++ks) {
        u16x8 bp[S], bw[S];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = 32 * ks + 8 * lq + e;
            const double vp = (k < ldn) ? s_phi[k * 16 + li] : 0.0;
            const double vw = (k < ldn) ? s_w[k * 16 + li] : 0.0;
            unsigned short h[S];
            split_bf16<S>(vp, h);
#pragma unroll
            for (int q = 0; q < S; ++q) bp[q][e] = h[q];
            split_bf16<S>(vw, h);
#pragma unroll
            for (int q = 0; q < S; ++q) bw[q][e] = h[q];
        }
        bf16x8 Bp[S], Bw[S];
#pragma unroll
        for (int q = 0; q < S; ++q) {
            Bp[q] = __builtin_bit_cast(bf16x8, bp[q]);
            Bw[q] = __builtin_bit_cast(bf16x8, bw[q]);
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            u16x8 ap[S];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 32 * ks + 8 * lq + e;
                const int col = (k < ldn) ? s_cols[k] : -1;
                const double x = (col >= 0) ? P[(int64_t)col * N + gi[i]] : 0.0;
                unsigned short h[S];
                split_bf16<S>(x, h);
#pragma unroll
                for (int q = 0; q < S; ++q) ap[q][e] = h[q];
            }
            bf16x8 A[S];
#pragma unroll
            for (int q = 0; q < S; ++q) A[q] = __builtin_bit_cast(bf16x8, ap[q]);
            const f32x4 z = f32x4{0.f, 0.f, 0.f, 0.f};
            const f32x4 h0 = mfma_bf16(A[0], Bp[0], z);
            const f32x4 h1 = mfma_bf16(A[0], Bw[0], z);
            f32x4 c0 = z, c1 = z;
#pragma unroll
            for (int q = 1; q < SlicePairs<S>::n; ++q) {
                c0 = mfma_bf16(A[SlicePairs<S>::a[q]], Bp[SlicePairs<S>::b[q]], c0);
                c1 = mfma_bf16(A[SlicePairs<S>::a[q]], Bw[SlicePairs<S>::b[q]], c1);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc0[i][r] += (double)h0[r] + (double)c0[r];
                acc1[i][r] += (double)h1[r] + (double)c1[r];
            }
        }
    }
    __syncthreads();  // every wave has stopped reading the planes
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = ((wave + 8 * i) * 16 + 4 * lq + r) * 16 + li;
            s_phi[o] = acc0[i][r];
            s_w[o] = acc1[i][r];
        }
}

// X^T, Z^T tiles of one wave: FP64 MFMA (S = 0) or S bf16 slices
template <int S, int NI>
__device__ inline void grad_tiles_any(int wave, int lane, int T, int ldn, const int* s_cols, double* s_phi,
                                      double* s_w, const double* __restrict__ P, int N) {
    if constexpr (S == 0)
        grad_tiles<NI>(wave, lane, T, s_cols, s_phi, s_w, P, N);
    else
        grad_tiles_split<S, NI>(wave, lane, ldn, s_cols, s_phi, s_w, P, N);
}

template <int S, int NIMAX>
__global__ __launch_bounds__(512) void k_grad_fn(const int* __restrict__ work, const BlockDesc* __restrict__ blocks,
                                                 const double* __restrict__ pts, int64_t n_slots,
                                                 const double* __restrict__ shrec, int RS,
                                                 const int* __restrict__ bshells, const int* __restrict__ bshell_col,
                                                 const int* __restrict__ bcols, const int* __restrict__ bsorder,
                                                 const int* __restrict__ shell_atom, const int* __restrict__ slot_atom,
                                                 const double* __restrict__ coef, const double* __restrict__ P, int N,
                                                 int n_atoms, double* __restrict__ gacc) {
    constexpr int PB = 16, SUBS = kBlockPts / PB, NSL = 512 / PB;
    __shared__ double s_pt[PB][3];
    __shared__ double s_cf[PB][4];
    __shared__ int s_pa[PB];
    extern __shared__ double s_dyn[];
    const int bidx = work[blockIdx.x / SUBS];
    const int sub = blockIdx.x % SUBS;
    const BlockDesc B = blocks[bidx];
    const int p0 = sub * PB;
    if (p0 >= B.npts) return;  // uniform exit before any barrier
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldn = B.ldn, nao = B.nao, nsh = B.nshell;
    const int T = ldn >> 4;
    // LDS: phi / X [ldn][16] | W / Z [ldn][16] | records | per-atom sums [3 n_atoms] | cols | shcol | sorder | shell atom
    double* s_phi = s_dyn;
    double* s_w = s_phi + ldn * PB;
    double* s_rec = s_w + ldn * PB;
    double* s_g = s_rec + (size_t)nsh * RS;
    int* s_cols = reinterpret_cast<int*>(s_g + 3 * n_atoms);
    int* s_shcol = s_cols + ldn;
    int* s_sord = s_shcol + nsh;
    int* s_shat = s_sord + nsh;
    for (int c = tid; c < ldn; c += 512) s_cols[c] = (c < nao) ? bcols[B.col_begin + c] : -1;
    for (int k = tid; k < nsh; k += 512) {
        s_shcol[k] = bshell_col[B.shell_begin + k];
        s_sord[k] = bsorder[B.shell_begin + k];
        s_shat[k] = shell_atom[bshells[B.shell_begin + k]];
    }
    for (int idx = tid; idx < nsh * RS; idx += 512) {
        const int k = idx / RS;
        s_rec[idx] = shrec[(int64_t)bshells[B.shell_begin + k] * RS + (idx - k * RS)];
    }
    for (int i = tid; i < 3 * n_atoms; i += 512) s_g[i] = 0.0;
    if (tid < PB) {
        const int64_t slot = B.pt_begin + p0 + tid;
        s_pt[tid][0] = pts[slot];
        s_pt[tid][1] = pts[n_slots + slot];
        s_pt[tid][2] = pts[2 * n_slots + slot];
        for (int q = 0; q < 4; ++q) s_cf[tid][q] = coef[4 * slot + q];
        s_pa[tid] = slot_atom[slot];
    }
    __syncthreads();

    const int p = tid & (PB - 1);
    const int kk = tid / PB;
    const double px = s_pt[p][0], py = s_pt[p][1], pz = s_pt[p][2];
    const double a2 = 2.0 * s_cf[p][0];
    const double cd[3] = {s_cf[p][1], s_cf[p][2], s_cf[p][3]};
    // 1. phi and W = c . grad phi
    for (int j = kk; j < nsh; j += NSL) {
        const int k = s_sord[j];
        double v[5], vx[5], vy[5], vz[5];
        rec_values(s_rec + k * RS, px, py, pz, v, vx, vy, vz);
        const int c0 = s_shcol[k];
        const int nc = 2 * (((int)s_rec[k * RS + 3]) & 3) + 1;
        for (int m = 0; m < nc; ++m) {
            s_phi[(c0 + m) * PB + p] = v[m];
            s_w[(c0 + m) * PB + p] = cd[0] * vx[m] + cd[1] * vy[m] + cd[2] * vz[m];
        }
    }
    for (int it = tid; it < (ldn - nao) * PB; it += 512) {
        s_phi[nao * PB + it] = 0.0;
        s_w[nao * PB + it] = 0.0;
    }
    __syncthreads();

    // 2. X^T = P_bb phi^T, Z^T = P_bb W^T, stored over the planes
    {
        const int nI = (T > wave) ? (T - wave + 7) / 8 : 0;
        switch (nI) {
            case 1: grad_tiles_any<S, 1>(wave, lane, T, ldn, s_cols, s_phi, s_w, P + B.poff, B.ldp); break;
            case 2:
                if constexpr (NIMAX >= 2) grad_tiles_any<S, 2>(wave, lane, T, ldn, s_cols, s_phi, s_w, P + B.poff, B.ldp);
                else __syncthreads();  // unreachable: the host keeps T <= 8 NIMAX
                break;
            case 3:
                if constexpr (NIMAX >= 3) grad_tiles_any<S, 3>(wave, lane, T, ldn, s_cols, s_phi, s_w, P + B.poff, B.ldp);
                else __syncthreads();
                break;
            case 4:
                if constexpr (NIMAX >= 4) grad_tiles_any<S, 4>(wave, lane, T, ldn, s_cols, s_phi, s_w, P + B.poff, B.ldp);
                else __syncthreads();
                break;
            default: __syncthreads(); break;
        }
    }
    __syncthreads();

    // 3. grad phi and second derivatives per (point, shell), summed per atom
    double gp[3] = {0.0, 0.0, 0.0};
    for (int j = kk; j < nsh; j += NSL) {
        const int k = s_sord[j];
        double r3[3], R0, R1, R2;
        const int l = rec_radial2(s_rec + k * RS, px, py, pz, r3, R0, R1, R2);
        const int c0 = s_shcol[k];
        double g[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < 2 * l + 1; ++m) {
            double f, fg[3], fh[6];
            angular2(l, m, r3, f, fg, fh);
            const double X = s_phi[(c0 + m) * PB + p];
            const double Zv = s_w[(c0 + m) * PB + p];
            const double t = a2 * X + Zv;
            // H[a][b] = fh_ab R0 + (fg_a r_b + fg_b r_a) R1 + f (delta_ab R1 + r_a r_b R2)
            double H[3][3];
            const int hix[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    H[a][b] = fh[hix[a][b]] * R0 + (fg[a] * r3[b] + fg[b] * r3[a]) * R1 +
                              f * ((a == b ? R1 : 0.0) + r3[a] * r3[b] * R2);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double dphi = fg[a] * R0 + f * r3[a] * R1;
                g[a] += dphi * t + X * (cd[0] * H[a][0] + cd[1] * H[a][1] + cd[2] * H[a][2]);
            }
        }
        if (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0) {
            const int at = s_shat[k];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                atomicAdd(&s_g[3 * at + a], g[a]);
                gp[a] += g[a];
            }
        }
    }
    if (s_pa[p] >= 0 && (gp[0] != 0.0 || gp[1] != 0.0 || gp[2] != 0.0)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) atomicAdd(&s_g[3 * s_pa[p] + a], -gp[a]);
    }
    __syncthreads();
    for (int i = tid; i < 3 * n_atoms; i += 512)
        if (s_g[i] != 0.0) atomicAdd(gacc + i, -4.0 * s_g[i]);
}

// Becke step s(nu) = (1 - f(f(f(nu)))) / 2, f(x) = x (3 - x^2) / 2, and ds/dnu
__device__ inline void becke_step(double nu, double& s, double& ds) {
    const double f1 = 0.5 * nu * (3.0 - nu * nu);
    const double f2 = 0.5 * f1 * (3.0 - f1 * f1);
    const double f3 = 0.5 * f2 * (3.0 - f2 * f2);
    s = 0.5 * (1.0 - f3);
    ds = -0.5 * (1.5 * (1.0 - f2 * f2)) * (1.5 * (1.0 - f1 * f1)) * (1.5 * (1.0 - nu * nu));
}

// cell factor s(mu) of the grid's partition and ds/dmu: Becke with the size
// adjustment a (nu = mu + a (1 - mu^2); a = 0 without adjustments), or the
// Stratmann-Scuseria step, g(mu) of calcMultiCenterWeight_SS
// (DfGenerateGrid.cpp:1923-1938, a = 0.64)
__device__ inline void cell_step(int ss, double mu, double a, double& s, double& dsdmu) {
    if (ss) {
        if (mu < -0.64 || mu > 0.64) {
            s = (mu < 0.0) ? 1.0 : 0.0;
            dsdmu = 0.0;
        } else {
            const double q = mu / 0.64, q2 = q * q, q3 = q2 * q, q5 = q2 * q3;
            s = 0.5 * (1.0 - (1.0 / 16.0) * (35.0 * q - 35.0 * q3 + 21.0 * q5 - 5.0 * q5 * q2));
            dsdmu = -0.5 * (1.0 / 16.0) * (35.0 - 105.0 * q2 + 105.0 * q2 * q2 - 35.0 * q2 * q2 * q2) / 0.64;
        }
        return;
    }
    double ds;
    becke_step(mu + a * (1.0 - mu * mu), s, ds);
    dsdmu = ds * (1.0 - 2.0 * a * mu);
}

// atoms: n x 4 (x, y, z, Z) | 1/R_mn [n][n] | a_mn [n][n] (the partition kernel's table);
// ss: the grid's partition is Stratmann-Scuseria (else Becke)
__global__ void k_grad_w(const double* __restrict__ atoms, int na, int64_t n_slots, const double* __restrict__ pts,
                         const int* __restrict__ slot_atom, const double* __restrict__ ewf,
                         double* __restrict__ gacc, int ss) {
    extern __shared__ double s_buf[];  // r [blockDim][na] | P [blockDim][na] | sums [3 na]
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    double* r = s_buf + (size_t)tid * na;
    double* Pv = s_buf + (size_t)nt * na + (size_t)tid * na;
    double* s_g = s_buf + 2 * (size_t)nt * na;
    for (int i = tid; i < 3 * na; i += nt) s_g[i] = 0.0;
    __syncthreads();
    const double* invd = atoms + 4 * na;
    const double* amat = invd + (size_t)na * na;
    const int64_t slot = (int64_t)blockIdx.x * nt + tid;
    const int A = (slot < n_slots) ? slot_atom[slot] : -1;
    const double e = (A >= 0) ? ewf[slot] : 0.0;
    if (A >= 0 && e != 0.0) {
        const double x = pts[slot], y = pts[n_slots + slot], z = pts[2 * n_slots + slot];
        for (int i = 0; i < na; ++i) {
            const double dx = x - atoms[4 * i], dy = y - atoms[4 * i + 1], dz = z - atoms[4 * i + 2];
            r[i] = sqrt(dx * dx + dy * dy + dz * dz);
        }
        double Zs = 0.0;
        for (int i = 0; i < na; ++i) {
            double p = 0.0;
            if (atoms[4 * i + 3] > 0.0) {
                p = 1.0;
                for (int j = 0; j < na && p != 0.0; ++j) {
                    if (j == i || atoms[4 * j + 3] <= 0.0) continue;
                    const double mu = (r[i] - r[j]) * invd[(size_t)i * na + j];
                    double s, ds;
                    cell_step(ss, mu, amat[(size_t)i * na + j], s, ds);
                    p *= s;
                }
            }
            Pv[i] = p;
            Zs += p;
        }
        const double om = Pv[A] / Zs;
        if (om > 0.0) {
            const double K = e / om / Zs;  // w_single e_xc / Z
            // product of cell i's factors without the one of atom j
            auto prod_ex = [&](int i, int j, double s_ij) {
                if (s_ij > 1e-200) return Pv[i] / s_ij;
                double q = 1.0;
                for (int m = 0; m < na; ++m) {
                    if (m == i || m == j || atoms[4 * m + 3] <= 0.0) continue;
                    const double mu = (r[i] - r[m]) * invd[(size_t)i * na + m];
                    double s, ds;
                    cell_step(ss, mu, amat[(size_t)i * na + m], s, ds);
                    q *= s;
                }
                return q;
            };
            double gA[3] = {0.0, 0.0, 0.0};
            for (int Bt = 0; Bt < na; ++Bt) {
                if (Bt == A || atoms[4 * Bt + 3] <= 0.0) continue;
                const double uB[3] = {(atoms[4 * Bt] - x) / r[Bt], (atoms[4 * Bt + 1] - y) / r[Bt],
                                      (atoms[4 * Bt + 2] - z) / r[Bt]};
                double gZ[3] = {0.0, 0.0, 0.0}, gPA[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < na; ++i) {
                    if (i == Bt || atoms[4 * i + 3] <= 0.0) continue;
                    const double id = invd[(size_t)Bt * na + i];
                    const double Rv[3] = {(atoms[4 * Bt] - atoms[4 * i]) * id, (atoms[4 * Bt + 1] - atoms[4 * i + 1]) * id,
                                          (atoms[4 * Bt + 2] - atoms[4 * i + 2]) * id};
                    // cell B, factor of atom i: grad_B mu_Bi = (u_B - mu_Bi e_Bi) / R
                    {
                        const double mu = (r[Bt] - r[i]) * id;
                        double s, ds;
                        cell_step(ss, mu, amat[(size_t)Bt * na + i], s, ds);
                        const double c = prod_ex(Bt, i, s) * ds * id;
#pragma unroll
                        for (int d = 0; d < 3; ++d) gZ[d] += c * (uB[d] - mu * Rv[d]);
                    }
                    // cell i, factor of atom B: grad_B mu_iB = (-u_B - mu_iB e_Bi) / R
                    {
                        const double mu = (r[i] - r[Bt]) * id;
                        double s, ds;
                        cell_step(ss, mu, amat[(size_t)i * na + Bt], s, ds);
                        const double c = prod_ex(i, Bt, s) * ds * id;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const double gm = c * (-uB[d] - mu * Rv[d]);
                            gZ[d] += gm;
                            if (i == A) gPA[d] += gm;
                        }
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double c = K * (gPA[d] - om * gZ[d]);
                    atomicAdd(&s_g[3 * Bt + d], c);
                    gA[d] -= c;
                }
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) atomicAdd(&s_g[3 * A + d], gA[d]);
        }
    }
    __syncthreads();
    for (int i = tid; i < 3 * na; i += nt)
        if (s_g[i] != 0.0) atomicAdd(gacc + i, s_g[i]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return PDFXC_OK;
}

template <int S, int NIMAX>
static int grad_fn_one(Context& c, const LaunchRange& r, const int* d_shell_atom, const int* d_slot_atom,
                       int n_atoms, double* d_gacc) {
    hipLaunchKernelGGL((k_grad_fn<S, NIMAX>), dim3(r.count * (kBlockPts / 16)), dim3(512), r.lds, c.stream,
                       c.d_rho_work + r.offset, c.d_blocks, c.d_pts, c.n_slots, c.d_shrec, c.rec_stride,
                       c.d_block_shells, c.d_block_shell_col, c.d_block_cols, c.d_block_sorder, d_shell_atom,
                       d_slot_atom, c.d_coef, c.d_P, c.n_aos, n_atoms, d_gacc);
    return check("k_grad_fn");
}

template <int S>
static int grad_fn_cls(Context& c, const LaunchRange& r, const int* d_shell_atom, const int* d_slot_atom,
                       int n_atoms, double* d_gacc) {
    switch (r.cls) {
        case 0: return grad_fn_one<S, 1>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
        case 1: return grad_fn_one<S, 2>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
        case 2: return grad_fn_one<S, 4>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
        default: set_error("gradient: blocks wider than 512 AO columns are not supported (raise ao_screen)");
                 return PDFXC_ERR_UNSUPPORTED;
    }
}

int launch_grad_fn(Context& c, const LaunchRange& r, const int* d_shell_atom, const int* d_slot_atom,
                   int n_atoms, double* d_gacc) {
    if (r.count <= 0) return PDFXC_OK;
    // the gradient's contractions: FP64, bf16 hi + lo in PDFXC_SPLIT_BF16,
    // else bf16 hi + mid + lo (the FP16 K mode included)
    switch (c.opts.precision == PDFXC_FP64 ? 0 : c.opts.precision == PDFXC_SPLIT_BF16 ? 2 : 3) {
        case 0: return grad_fn_cls<0>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
        case 2: return grad_fn_cls<2>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
        default: return grad_fn_cls<3>(c, r, d_shell_atom, d_slot_atom, n_atoms, d_gacc);
    }
}

int launch_grad_w(Context& c, const double* d_atoms, int n_atoms, const int* d_slot_atom, const double* d_ewf,
                  double* d_gacc, int partition) {
    int threads = 64;
    while (threads > 1 && (size_t)threads * 2 * n_atoms * sizeof(double) + 3 * n_atoms * sizeof(double) > 65536)
        threads >>= 1;
    const size_t lds = ((size_t)threads * 2 * n_atoms + 3 * (size_t)n_atoms) * sizeof(double);
    if (lds > 65536) {
        set_error("gradient: too many atoms for the weight-derivative kernel's LDS");
        return PDFXC_ERR_UNSUPPORTED;
    }
    const int64_t g = (c.n_slots + threads - 1) / threads;
    hipLaunchKernelGGL(k_grad_w, dim3((unsigned)g), dim3(threads), lds, c.stream, d_atoms, n_atoms, c.n_slots,
                       c.d_pts, d_slot_atom, d_ewf, d_gacc + 3 * n_atoms, partition == PDFXC_PARTITION_SS ? 1 : 0);
    return check("k_grad_w");
}

}  // namespace pdfxc
