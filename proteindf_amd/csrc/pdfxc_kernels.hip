// HIP kernels of the grid XC build for gfx950 (CDNA4), FP64 path.
//
// Per grid block b (128 points, N_b screened AOs, ldn = N_b padded to 16):
//   k_ao_eval   phi, dphi/dx, dphi/dy, dphi/dz of the block's AOs   -> scratch
//               [4][ldn][128] (AO-major, points contiguous)
//   k_rho_vxc   X = Phi_b P_bb on v_mfma_f64_16x16x4 with the row dots
//               rho = sum_j X_pj phi_pj, grad rho = 2 sum_j X_pj grad phi_pj fused,
//               then the functional per point and V = (w vrho/2) phi + c . grad phi
//               written over the dphi/dx plane; per-block energy sum.
//   k_kmat      K_bb += Phi_b^T V_b + V_b^T Phi_b on MFMA for the lower tile
//               pairs, scattered into the dense lower triangle of K.
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

__global__ __launch_bounds__(256) void k_ao_eval(const BlockDesc* __restrict__ blocks, int b0,
                                                 const double* __restrict__ pts, int64_t n_slots,
                                                 const DeviceShell* __restrict__ shells,
                                                 const int* __restrict__ bshells,
                                                 const int* __restrict__ bshell_col,
                                                 double* __restrict__ scratch,
                                                 int64_t chunk_off) {
    const BlockDesc B = blocks[b0 + blockIdx.x];
    double* phi = scratch + (B.scratch_off - chunk_off);
    const int64_t plane = (int64_t)B.ldn * kBlockPts;
    const int p = threadIdx.x & (kBlockPts - 1);
    const int half = threadIdx.x >> 7;
    const int64_t slot = B.pt_begin + p;
    const double px = pts[slot], py = pts[n_slots + slot], pz = pts[2 * n_slots + slot];
    for (int s = half; s < B.nshell; s += 2) {
        const DeviceShell& S = shells[bshells[B.shell_begin + s]];
        const int col = bshell_col[B.shell_begin + s];
        double v[5], vx[5], vy[5], vz[5];
        shell_values(S, px - S.cx, py - S.cy, pz - S.cz, v, vx, vy, vz);
        const int nc = 2 * S.l + 1;
        for (int m = 0; m < nc; ++m) {
            const int64_t o = (int64_t)(col + m) * kBlockPts + p;
            phi[o] = v[m];
            phi[plane + o] = vx[m];
            phi[2 * plane + o] = vy[m];
            phi[3 * plane + o] = vz[m];
        }
    }
    for (int c = B.nao + half; c < B.ldn; c += 2) {
        const int64_t o = (int64_t)c * kBlockPts + p;
        phi[o] = 0.0;
        phi[plane + o] = 0.0;
        phi[2 * plane + o] = 0.0;
        phi[3 * plane + o] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// X = Phi_b P_bb and the density at the block's points, then the functional
// (buildVxc, DfCalcGridX.cpp:2137-2193) and V_b.  4 waves; wave w owns rows
// 32w..32w+31 (two 16-row MFMA tiles) of a 128 x 64 output tile, the
// workgroup sweeps the 64-column tiles of X.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rho_vxc(const BlockDesc* __restrict__ blocks, int b0,
                                                 const double* __restrict__ pts, int64_t n_slots,
                                                 const double* __restrict__ P, int N,
                                                 const int* __restrict__ bcols,
                                                 double* __restrict__ scratch, int64_t chunk_off,
                                                 double* __restrict__ coef_out,
                                                 double* __restrict__ rho_out,
                                                 double* __restrict__ eblock, double dens_cut) {
    extern __shared__ int s_cols[];  // ldn column -> global AO (or -1)
    __shared__ double s_rho[kBlockPts][4];
    __shared__ double s_coef[kBlockPts][4];
    __shared__ double s_e[kBlockPts];

    const int bidx = b0 + blockIdx.x;
    const BlockDesc B = blocks[bidx];
    const double* phi = scratch + (B.scratch_off - chunk_off);
    const int64_t plane = (int64_t)B.ldn * kBlockPts;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;
    const int lq = lane >> 4;
    const int ldn = B.ldn;

    for (int c = tid; c < ldn; c += 256) s_cols[c] = (c < B.nao) ? bcols[B.col_begin + c] : -1;
    __syncthreads();

    const int row0 = wave * 32;
    double part[2][4][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) part[rt][r][q] = 0.0;

    for (int j0 = 0; j0 < ldn; j0 += 64) {
        const int nct = min(4, (ldn - j0) >> 4);
        int gcol[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) gcol[ct] = (ct < nct) ? s_cols[j0 + ct * 16 + li] : -1;
        v4d acc[2][4];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = v4d{0.0, 0.0, 0.0, 0.0};

        for (int k0 = 0; k0 < ldn; k0 += 4) {
            const int kk = k0 + lq;
            const double a0 = phi[(int64_t)kk * kBlockPts + row0 + li];
            const double a1 = phi[(int64_t)kk * kBlockPts + row0 + 16 + li];
            const int gk = s_cols[kk];
            const double* Prow = P + (int64_t)(gk < 0 ? 0 : gk) * N;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                if (ct < nct) {
                    const double b = (gk >= 0 && gcol[ct] >= 0) ? Prow[gcol[ct]] : 0.0;
                    acc[0][ct] = mfma16x16x4(a0, b, acc[0][ct]);
                    acc[1][ct] = mfma16x16x4(a1, b, acc[1][ct]);
                }
            }
        }
        // fused row dots with phi and grad phi of the same columns
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            if (ct < nct) {
                const int64_t cbase = (int64_t)(j0 + ct * 16 + li) * kBlockPts;
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = row0 + rt * 16 + 4 * r + lq;
                        const double x = acc[rt][ct][r];
                        part[rt][r][0] += x * phi[cbase + row];
                        part[rt][r][1] += x * phi[plane + cbase + row];
                        part[rt][r][2] += x * phi[2 * plane + cbase + row];
                        part[rt][r][3] += x * phi[3 * plane + cbase + row];
                    }
            }
        }
    }
    // reduce over the 16 lanes that share rows (same lq)
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double v = part[rt][r][q];
                v += __shfl_xor(v, 1);
                v += __shfl_xor(v, 2);
                v += __shfl_xor(v, 4);
                v += __shfl_xor(v, 8);
                part[rt][r][q] = v;
            }
    if (li == 0) {
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = row0 + rt * 16 + 4 * r + lq;
                s_rho[row][0] = part[rt][r][0];
                s_rho[row][1] = 2.0 * part[rt][r][1];
                s_rho[row][2] = 2.0 * part[rt][r][2];
                s_rho[row][3] = 2.0 * part[rt][r][3];
            }
    }
    __syncthreads();

    if (tid < kBlockPts) {
        const int p = tid;
        const int64_t slot = B.pt_begin + p;
        const double w = pts[3 * n_slots + slot];
        const double rho = s_rho[p][0];
        const double gx = s_rho[p][1], gy = s_rho[p][2], gz = s_rho[p][3];
        if (rho_out) {
            rho_out[4 * slot + 0] = rho;
            rho_out[4 * slot + 1] = gx;
            rho_out[4 * slot + 2] = gy;
            rho_out[4 * slot + 3] = gz;
        }
        double a = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, e = 0.0;
        if (p < B.npts && rho > dens_cut) {
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
        s_coef[p][0] = a;
        s_coef[p][1] = cx;
        s_coef[p][2] = cy;
        s_coef[p][3] = cz;
        s_e[p] = e;
        if (coef_out) {
            coef_out[4 * slot + 0] = a;
            coef_out[4 * slot + 1] = cx;
            coef_out[4 * slot + 2] = cy;
            coef_out[4 * slot + 3] = cz;
        }
    }
    __syncthreads();
    // deterministic block energy sum
    for (int s = kBlockPts / 2; s > 0; s >>= 1) {
        if (tid < s) s_e[tid] += s_e[tid + s];
        __syncthreads();
    }
    if (tid == 0) eblock[bidx] = s_e[0];

    // V = a phi + c . grad phi, written over the d/dx plane
    double* V = const_cast<double*>(phi) + plane;
    const int p = tid & (kBlockPts - 1);
    const double a = s_coef[p][0], cx = s_coef[p][1], cy = s_coef[p][2], cz = s_coef[p][3];
    for (int c = tid >> 7; c < ldn; c += 2) {
        const int64_t o = (int64_t)c * kBlockPts + p;
        const double v = a * phi[o] + cx * V[o] + cy * phi[2 * plane + o] + cz * phi[3 * plane + o];
        V[o] = v;
    }
}

// ---------------------------------------------------------------------------
// K_bb += Phi^T V + V^T Phi over the block's points, lower tile pairs only,
// scattered into K (lower triangle, global AO indices ascending within a
// block).  Wave w of 4 owns a 32 x 32 quarter of a 64 x 64 tile pair; the
// reduction runs over the 128 points in slabs of 16, each lane loading 4
// consecutive points of one AO row (the MFMA k index is a permutation of the
// points, the same for both operands).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kmat(const BlockDesc* __restrict__ blocks, int b0,
                                              const int* __restrict__ bcols,
                                              const double* __restrict__ scratch, int64_t chunk_off,
                                              double* __restrict__ K, int N) {
    extern __shared__ int s_cols[];
    const BlockDesc B = blocks[b0 + blockIdx.x];
    const double* phi = scratch + (B.scratch_off - chunk_off);
    const int64_t plane = (int64_t)B.ldn * kBlockPts;
    const double* V = phi + plane;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int li = lane & 15;
    const int lq = lane >> 4;
    const int ldn = B.ldn;
    for (int c = tid; c < ldn; c += 256) s_cols[c] = (c < B.nao) ? bcols[B.col_begin + c] : -1;
    __syncthreads();

    const int ntile = (ldn + 63) >> 6;
    const int wi = wave >> 1, wj = wave & 1;
    for (int I = 0; I < ntile; ++I) {
        for (int J = 0; J <= I; ++J) {
            const int r0 = I * 64 + wi * 32;
            const int c0 = J * 64 + wj * 32;
            if (r0 >= ldn || c0 >= ldn) continue;  // whole quarter outside the padded width
            if (I == J && wj > wi) continue;        // strictly upper quarter of a diagonal pair
            const bool rv[2] = {r0 < ldn, r0 + 16 < ldn};
            const bool cv[2] = {c0 < ldn, c0 + 16 < ldn};
            v4d acc[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
            for (int t = 0; t < kBlockPts / 16; ++t) {
                const int pbase = 16 * t + 4 * lq;
                double4 fI[2], vI[2], fJ[2], vJ[2];
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int64_t oi = (int64_t)(r0 + a * 16 + li) * kBlockPts + pbase;
                    const int64_t oj = (int64_t)(c0 + a * 16 + li) * kBlockPts + pbase;
                    const double4 z = {0.0, 0.0, 0.0, 0.0};
                    fI[a] = rv[a] ? *reinterpret_cast<const double4*>(phi + oi) : z;
                    vI[a] = rv[a] ? *reinterpret_cast<const double4*>(V + oi) : z;
                    fJ[a] = cv[a] ? *reinterpret_cast<const double4*>(phi + oj) : z;
                    vJ[a] = cv[a] ? *reinterpret_cast<const double4*>(V + oj) : z;
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        if (!(rv[a] && cv[b])) continue;
                        acc[a][b] = mfma16x16x4(fI[a].x, vJ[b].x, acc[a][b]);
                        acc[a][b] = mfma16x16x4(vI[a].x, fJ[b].x, acc[a][b]);
                        acc[a][b] = mfma16x16x4(fI[a].y, vJ[b].y, acc[a][b]);
                        acc[a][b] = mfma16x16x4(vI[a].y, fJ[b].y, acc[a][b]);
                        acc[a][b] = mfma16x16x4(fI[a].z, vJ[b].z, acc[a][b]);
                        acc[a][b] = mfma16x16x4(vI[a].z, fJ[b].z, acc[a][b]);
                        acc[a][b] = mfma16x16x4(fI[a].w, vJ[b].w, acc[a][b]);
                        acc[a][b] = mfma16x16x4(vI[a].w, fJ[b].w, acc[a][b]);
                    }
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    if (!(rv[a] && cv[b])) continue;
                    const int lc = c0 + b * 16 + li;
                    const int gj = s_cols[lc];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int lr = r0 + a * 16 + 4 * r + lq;
                        const int gi = s_cols[lr];
                        if (gi >= 0 && gj >= 0 && lr >= lc) {
                            atomicAdd(K + (int64_t)gi * N + gj, acc[a][b][r]);
                        }
                    }
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

static int max_ldn(const Context& c, const Chunk& ch) {
    int m = kColPad;
    for (int b = ch.block_begin; b < ch.block_end; ++b) m = std::max(m, c.blocks[b].ldn);
    return m;
}

int launch_ao_eval(Context& c, const Chunk& ch) {
    const int nb = ch.block_end - ch.block_begin;
    if (nb <= 0) return PDFXC_OK;
    const int64_t off = c.blocks[ch.block_begin].scratch_off;
    hipLaunchKernelGGL(k_ao_eval, dim3(nb), dim3(256), 0, c.stream, c.d_blocks, ch.block_begin,
                       c.d_pts, c.n_slots, c.d_shells, c.d_block_shells, c.d_block_shell_col,
                       c.d_scratch, off);
    return check_launch("k_ao_eval");
}

int launch_rho_vxc(Context& c, const Chunk& ch) {
    const int nb = ch.block_end - ch.block_begin;
    if (nb <= 0) return PDFXC_OK;
    const int64_t off = c.blocks[ch.block_begin].scratch_off;
    const size_t lds = sizeof(int) * (size_t)max_ldn(c, ch);
    hipLaunchKernelGGL(k_rho_vxc, dim3(nb), dim3(256), lds, c.stream, c.d_blocks, ch.block_begin,
                       c.d_pts, c.n_slots, c.d_P, c.n_aos, c.d_block_cols, c.d_scratch, off,
                       c.d_coef, c.d_rho, c.d_eblock, c.opts.density_cutoff);
    return check_launch("k_rho_vxc");
}

int launch_kmat(Context& c, const Chunk& ch) {
    const int nb = ch.block_end - ch.block_begin;
    if (nb <= 0) return PDFXC_OK;
    const int64_t off = c.blocks[ch.block_begin].scratch_off;
    const size_t lds = sizeof(int) * (size_t)max_ldn(c, ch);
    hipLaunchKernelGGL(k_kmat, dim3(nb), dim3(256), lds, c.stream, c.d_blocks, ch.block_begin,
                       c.d_block_cols, c.d_scratch, off, c.d_K, c.n_aos);
    return check_launch("k_kmat");
}

int launch_energy_sum(Context& c, double* dE) {
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(256), 0, c.stream, c.d_eblock,
                       (int)c.blocks.size(), dE);
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
