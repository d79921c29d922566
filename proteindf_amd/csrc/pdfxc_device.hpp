// Device helpers shared by the FP64 kernels (pdfxc_kernels.hip) and the
// split-precision kernels (pdfxc_split.hip): MFMA wrappers and AO
// evaluation from LDS shell records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pdfxc {

typedef double v4d __attribute__((ext_vector_type(4)));

// MFMA f64 16x16x4 accumulator layout (checked on gfx950 by
// pdfxc_selftest_mfma): register r of lane l holds D[4 r + l / 16][l % 16];
// operand A[i][k] comes from lane i + 16 k, operand B[k][j] from lane j + 16 k.
__device__ inline v4d mfma16x16x4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Work-queue step of a persistent launch: the workgroup's next item, taken
// from *queue with one atomic (queue != nullptr; the launch has fewer
// workgroups than items and the host zeroes *queue before it), or the single
// item blockIdx.x (queue == nullptr; the second call returns n_items).  The
// barrier before the fetch keeps the previous item's LDS use apart from the
// next item's, the one after it publishes the index.  Every workgroup leaves
// its loop once the counter passes n_items.  The fetch is made by the first
// lane of the last wave: reading the atomic's result waits for every memory
// write its wave still has in flight, and the kernels that use this loop
// leave their writes (the K scatter, the rho stores) to the first waves.
__device__ inline int next_item(int* queue, int n_items, int iter, int* s_item) {
    __syncthreads();
    if (threadIdx.x == blockDim.x - 64)
        *s_item = queue ? atomicAdd(queue, 1) : (iter == 0 ? (int)blockIdx.x : n_items);
    __syncthreads();
    return *s_item;
}


constexpr double kTooBig = 30.0;  // DfCalcGridX::TOOBIG, primitives with alpha r^2 > 30 are skipped
constexpr double kInvSqrt3 = 0.57735026918962576451;

// e^{-x} for 0 <= x <= kTooBig: Cody-Waite reduction x = n ln2 + r with
// |r| <= ln2 / 2 (ln2 split in a 32-bit-exact high part and a low part),
// e^{-r} by its Taylor series to degree 13 in Horner form (truncation
// < 5e-18 relative), then scaling by 2^{-n}.  Written for the primitive
// exponentials of the AO values; agrees with the libm result to a few ulp.
__device__ inline double exp_neg(double x) {
    constexpr double kLog2e = 1.4426950408889634074;
    constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    const double n = __builtin_rint(x * kLog2e);
    const double r = -__builtin_fma(-n, kLn2Lo, __builtin_fma(-n, kLn2Hi, x));  // -(x - n ln2)
    double p = 1.0 / 6227020800.0;                        // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);           // 1/12!
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, -(int)n);
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
    // exponents ascend, so alpha r^2 > 30 holds for every primitive after the first that exceeds it
    for (int k = 0; k < nprim; ++k) {
        const double a = rec[4 + 2 * k];
        const double sh = a * r2;
        if (sh > kTooBig) break;
        const double g = rec[5 + 2 * k] * exp_neg(sh);
        R0 += g;
        R1 += -2.0 * a * g;
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

// Radial sums R0 = sum c e^{-a r^2}, R1 = sum -2a c e^{-a r^2} of a shell
// record at offset (x, y, z) (ascending exponents, TOOBIG cut), and the
// values / values and gradients formed from them: the same arithmetic as
// rec_values0 / rec_values, split so that a caller can keep R0, R1 between
// a values pass and a gradients pass.
__device__ inline void rec_radial01(const double* rec, double x, double y, double z, double& R0, double& R1) {
    const int nprim = ((int)rec[3]) >> 2;
    const double r2 = x * x + y * y + z * z;
    R0 = 0.0;
    R1 = 0.0;
    for (int k = 0; k < nprim; ++k) {
        const double a = rec[4 + 2 * k];
        const double sh = a * r2;
        if (sh > kTooBig) break;
        const double g = rec[5 + 2 * k] * exp_neg(sh);
        R0 += g;
        R1 += -2.0 * a * g;
    }
}
__device__ inline void rec_values0_r(int l, double x, double y, double z, double R0, double* v) {
    if (l == 0) {
        v[0] = R0;
    } else if (l == 1) {
        v[0] = x * R0;
        v[1] = y * R0;
        v[2] = z * R0;
    } else {
        v[0] = x * y * R0;
        v[1] = z * x * R0;
        v[2] = y * z * R0;
        v[3] = 0.5 * (x * x - y * y) * R0;
        v[4] = kInvSqrt3 * (z * z - 0.5 * (x * x + y * y)) * R0;
    }
}
// f(m) for the 2l + 1 components m of a shell, each call with a constant m:
// a loop to a run-time bound that reads a register array v[m] compiles to a
// chain of conditional moves per element
template <typename F>
__device__ inline void for_components(int l, F f) {
    if (l == 0) {
        f(0);
    } else if (l == 1) {
        f(0);
        f(1);
        f(2);
    } else {
        f(0);
        f(1);
        f(2);
        f(3);
        f(4);
    }
}

// phi_m and V_m = a phi_m + c . grad phi_m of the 2l+1 components of a shell
// at offset (x, y, z) from its centre, from its radial sums R0 and R1 =
// (dR/dr) / r: phi_m = f_m R0 and c . grad phi_m = R0 (c . grad f_m) +
// R1 f_m (c . r), f_m the angular polynomials of getPrefactor
// (DfCalcGridX.cpp:103-141: 1; x, y, z; xy, zx, yz, (x^2 - y^2)/2,
// (z^2 - (x^2 + y^2)/2)/sqrt3).  out(m, phi_m, V_m) is called per component.
template <typename Out>
__device__ inline void shell_phi_v(int l, double x, double y, double z, double R0, double R1, double a, double cx,
                                   double cy, double cz, Out out) {
    const double R1cr = R1 * (cx * x + cy * y + cz * z);
    auto put = [&](int m, double f, double cgf) {
        const double v = f * R0;
        out(m, v, a * v + R0 * cgf + f * R1cr);
    };
    if (l == 0) {
        put(0, 1.0, 0.0);
    } else if (l == 1) {
        put(0, x, cx);
        put(1, y, cy);
        put(2, z, cz);
    } else {
        put(0, x * y, cx * y + cy * x);
        put(1, z * x, cx * z + cz * x);
        put(2, y * z, cy * z + cz * y);
        put(3, 0.5 * (x * x - y * y), cx * x - cy * y);
        put(4, kInvSqrt3 * (z * z - 0.5 * (x * x + y * y)), kInvSqrt3 * (2.0 * cz * z - cx * x - cy * y));
    }
}

// Dots of one shell's values at a point with per-component weights X
// (X = P phi in the rho pass): sum_m X_m phi_m = R0 S and sum_m X_m grad
// phi_m = R0 G + R1 r S, with S = sum_m X_m f_m and G = sum_m X_m grad f_m
// over the angular polynomials f_m of getPrefactor (DfCalcGridX.cpp:103-141)
// and R1 = (dR/dr) / r; added to (d0, dx, dy, dz).  The same sums as
// forming phi and grad phi first, in far fewer operations.
__device__ inline void shell_dots(int l, double x, double y, double z, double R0, double R1, const double* X,
                                  double& d0, double& dx, double& dy, double& dz) {
    double S, Gx, Gy, Gz;
    if (l == 0) {
        S = X[0];
        Gx = Gy = Gz = 0.0;
    } else if (l == 1) {
        S = X[0] * x + X[1] * y + X[2] * z;
        Gx = X[0];
        Gy = X[1];
        Gz = X[2];
    } else {
        const double x4 = kInvSqrt3 * X[4];
        S = X[0] * (x * y) + X[1] * (z * x) + X[2] * (y * z) + X[3] * (0.5 * (x * x - y * y)) +
            x4 * (z * z - 0.5 * (x * x + y * y));
        Gx = X[0] * y + X[1] * z + (X[3] - x4) * x;
        Gy = X[0] * x + X[2] * z - (X[3] + x4) * y;
        Gz = X[1] * x + X[2] * y + 2.0 * x4 * z;
    }
    d0 += R0 * S;
    dx += R0 * Gx + R1 * x * S;
    dy += R0 * Gy + R1 * y * S;
    dz += R0 * Gz + R1 * z * S;
}

__device__ inline void rec_values_r(int l, double x, double y, double z, double R0, double R1, double* v,
                                    double* vx, double* vy, double* vz) {
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

// values only (first pass of k_rho)
__device__ inline void rec_values0(const double* rec, double px, double py, double pz, double* v) {
    const double x = px - rec[0], y = py - rec[1], z = pz - rec[2];
    const int meta = (int)rec[3];
    const int l = meta & 3, nprim = meta >> 2;
    const double r2 = x * x + y * y + z * z;
    double R0 = 0.0;
    for (int k = 0; k < nprim; ++k) {  // ascending exponents (see rec_values)
        const double a = rec[4 + 2 * k];
        const double sh = a * r2;
        if (sh > kTooBig) break;
        R0 += rec[5 + 2 * k] * exp_neg(sh);
    }
    if (l == 0) {
        v[0] = R0;
    } else if (l == 1) {
        v[0] = x * R0;
        v[1] = y * R0;
        v[2] = z * R0;
    } else {
        v[0] = x * y * R0;
        v[1] = z * x * R0;
        v[2] = y * z * R0;
        v[3] = 0.5 * (x * x - y * y) * R0;
        v[4] = kInvSqrt3 * (z * z - 0.5 * (x * x + y * y)) * R0;
    }
}

// Radial sums of a shell record at a point: R0 = sum c e^{-a r^2},
// R1 = sum -2a c e^{-a r^2}, R2 = sum 4a^2 c e^{-a r^2} (TOOBIG cut as in
// getD2AOs_core, DfCalcGridX.cpp:771-863); returns l and the offset (x, y, z).
__device__ inline int rec_radial2(const double* rec, double px, double py, double pz, double* r3, double& R0,
                                  double& R1, double& R2) {
    r3[0] = px - rec[0];
    r3[1] = py - rec[1];
    r3[2] = pz - rec[2];
    const int meta = (int)rec[3];
    const int nprim = meta >> 2;
    const double r2 = r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2];
    R0 = R1 = R2 = 0.0;
    for (int k = 0; k < nprim; ++k) {  // ascending exponents (see rec_values)
        const double a = rec[4 + 2 * k];
        const double sh = a * r2;
        if (sh > kTooBig) break;
        const double g = rec[5 + 2 * k] * exp_neg(sh);
        R0 += g;
        R1 += -2.0 * a * g;
        R2 += 4.0 * a * a * g;
    }
    return meta & 3;
}

// Angular polynomial f of component m of an l shell, its gradient and its
// second derivatives (xx, xy, xz, yy, yz, zz): s; x, y, z; xy, zx, yz,
// (x^2 - y^2)/2, (z^2 - (x^2 + y^2)/2)/sqrt3 (getPrefactor, :103-141).
__device__ inline void angular2(int l, int m, const double* r, double& f, double* fg, double* fh) {
    const double x = r[0], y = r[1], z = r[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) fg[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) fh[k] = 0.0;
    if (l == 0) {
        f = 1.0;
    } else if (l == 1) {
        f = r[m];
        fg[m] = 1.0;
    } else {
        switch (m) {
            case 0: f = x * y; fg[0] = y; fg[1] = x; fh[1] = 1.0; break;
            case 1: f = z * x; fg[0] = z; fg[2] = x; fh[2] = 1.0; break;
            case 2: f = y * z; fg[1] = z; fg[2] = y; fh[4] = 1.0; break;
            case 3: f = 0.5 * (x * x - y * y); fg[0] = x; fg[1] = -y; fh[0] = 1.0; fh[3] = -1.0; break;
            default:
                f = kInvSqrt3 * (z * z - 0.5 * (x * x + y * y));
                fg[0] = -kInvSqrt3 * x;
                fg[1] = -kInvSqrt3 * y;
                fg[2] = 2.0 * kInvSqrt3 * z;
                fh[0] = -kInvSqrt3;
                fh[3] = -kInvSqrt3;
                fh[5] = 2.0 * kInvSqrt3;
                break;
        }
    }
}

}  // namespace pdfxc
