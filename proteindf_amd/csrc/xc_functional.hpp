// The closed-shell functionals of ProteinDF's grid path, on the device.
//
// The grid path reaches a functional through the non-virtual closed-shell
// forwarders of DfFunctional_LDA / DfFunctional_GGA (DfFunctional.h:104-106,
// :126-130, :190-192, :221-232), i.e. the spin-resolved formulas at
// rhoB = rhoA and gAB = gBB = gAA.  This code evaluates those formulas
// specialised to that point: the spin-polarisation terms of VWN vanish exactly
// at zeta = 0 (they are multiplied by g(0) = 0 and g'(0) = 0), and the alpha
// and beta halves of Slater and B88 are equal.
//
// Components
//   Slater  DfFunctional_Slater.cpp:25-52, :64-76
//   B88     DfFunctional_Becke88.cpp:25-31, :40-57, :71-108, :137-163;
//           without the LDA part: DfFunctional_Becke88_ExceptLDA.cpp:26-33;
//           asinh = log(x+sqrt(x^2+1)) (TlMath.h:57-59)
//   VWN     DfFunctional_VWN.cpp:23-46, :76-124, :145-260, :262-346 (VWN5 sets);
//           DfFunctional_VWN3.cpp:23-54 (VWN3 para set)
//   LYP     DfFunctional_LYP.cpp:25-43, :64-197, :288-317, :346-386, :507-594, :652-674
// Functionals (the XC_TYPE values of DfXCFunctional.h:44; dispatch :140-180)
//   HF    no grid term (DfXCFunctional.h:141-143)
//   HFS   Slater                                   (DfFunctional_HFS.cpp)
//   SVWN  Slater + VWN5                             (DfFunctional_SVWN.cpp)
//   HFB   B88 with its LDA part                     (DfFunctional_Becke88.cpp)
//   BLYP  B88 + LYP                                 (DfFunctional_B88LYP.cpp)
//   B3LYP 0.80 Slater + 0.72 B88(non-LDA) + 0.19 VWN3 + 0.81 LYP
//         (DfFunctional_B3LYP.cpp:25-28, :59-75, :91-141)
//
// Outputs: e = f(rhoA, rhoA, g, g, g); vr = df/drhoA; vgaa = df/dgammaAA;
// vgab = df/dgammaAB (zero for the LDA functionals).  All FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace pdfxc {

struct XcPoint {
    double e, vr, vgaa, vgab;
};

// XC_TYPE of DfXCFunctional.h:44
enum XcType { kHF = 1, kHFS = 2, kSVWN = 3, kHFB = 4, kBLYP = 5, kB3LYP = 6 };

__host__ __device__ constexpr bool xc_type_valid(int t) { return t >= kHF && t <= kB3LYP; }

namespace xcf {

constexpr double kPi = 3.14159265358979323846;

// Slater at rhoA = rhoB = ra: e = C (ra^4/3 + ra^4/3), v = C' ra^1/3
__device__ __host__ inline void slater(double ra13, double ra43, double* e, double* v) {
    const double c13 = pow(3.0 / (4.0 * kPi), 1.0 / 3.0);
    const double slater_c = -(9.0 / 4.0) * c13 * (2.0 / 3.0);
    const double slater_dc = -3.0 * c13 * (2.0 / 3.0);
    *e = slater_c * (ra43 + ra43);
    *v = slater_dc * ra13;
}

// B88 at rhoA = rhoB = ra, gAA = gBB = g; x = sqrt(g)/ra^4/3.  The reference
// sets vr and vg only where ra and g both exceed 1e-16, also for the LDA part.
template <bool WITH_LDA>
__device__ __host__ inline void b88(double ra, double g, double ra13, double ra43, double* e,
                                    double* v, double* vg) {
    *e = 0.0;
    *v = 0.0;
    *vg = 0.0;
    if (ra > 1.0e-16) {
        const double sg = sqrt(g);
        const double x = sg / ra43;
        const double bx = 0.0042 * x;
        const double ash = log(x + sqrt(x * x + 1));
        const double M = 1.0 + 6.0 * bx * ash;
        double gx;
        if (WITH_LDA) {
            const double gparam = -(3.0 / 2.0) * pow(3.0 / (4.0 * kPi), 1.0 / 3.0);
            gx = gparam - ((bx * x) / M);
        } else {
            gx = -((bx * x) / M);
        }
        *e = ra43 * gx + ra43 * gx;
        if (g > 1.0e-16) {
            const double invM = 1.0 / M;
            const double gp = 6.0 * bx * bx * (ash + x / sqrt(x * x + 1.0)) * invM * invM - 2.0 * bx * invM;
            *v = (4.0 / 3.0) * ra13 * (gx - x * gp);
            *vg = 0.5 * (1.0 / sg) * gp;
        }
    }
}

// VWN eps_c(x) and its x-derivative for one parameter set (VWN eq. A13, A18)
__device__ __host__ inline void vwn_eps(double A, double b, double c, double x0, double x,
                                        double* eps, double* deps) {
    const double X = x * (x + b) + c;
    const double invX = 1.0 / X;
    const double X0 = x0 * (x0 + b) + c;
    const double invX0 = 1.0 / X0;
    const double Q = sqrt(4 * c - b * b);
    const double invQ = 1.0 / Q;
    const double at = atan(Q / (2.0 * x + b));
    const double xx0 = x - x0;
    const double t1 = log(x * x * invX);
    const double t2 = 2.0 * b * invQ * at;
    const double t3 = -b * x0 * invX0 * (log(xx0 * xx0 * invX) + 2.0 * (2.0 * x0 + b) * invQ * at);
    *eps = A * (t1 + t2 + t3);
    const double x2b = 2.0 * x + b;
    const double inv_x2bQ = 1.0 / (x2b * x2b + Q * Q);
    const double d4 = -b * x0 * invX0 * (2.0 / xx0 - x2b * invX - 4.0 * (2.0 * x0 + b) * inv_x2bQ);
    *deps = A * ((2.0 / x) + (-x2b * invX) + (-4.0 * b * inv_x2bQ) + d4);
}

// VWN at zeta = 0 (paramagnetic set only): eps_c = eps_para,
// v = eps + rho (-x / (6 rho)) eps'
template <bool VWN3>
__device__ __host__ inline void vwn(double ra, double* e, double* v) {
    const double rho = ra + ra;
    const double xv = pow((3.0 / (4.0 * kPi)) * (1.0 / rho), 1.0 / 6.0);
    double ep, epp;
    if (VWN3)
        vwn_eps(0.0310907, 13.0720, 42.7198, -0.409286, xv, &ep, &epp);
    else
        vwn_eps(0.0310907, 3.72744, 12.9352, -0.10498, xv, &ep, &epp);
    *e = rho * ep;
    *v = (ra > 1.0e-20) ? ep + rho * (-xv / (6.0 * rho) * epp) : 0.0;
}

// LYP with rhoA = rhoB = ra, gAA = gAB = gBB = g
__device__ __host__ inline void lyp(double ra, double g, double* e, double* v, double* vgaa,
                                    double* vgab) {
    const double rho = ra + ra;
    const double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349;
    const double AB = A * B;
    const double lyp_c = pow(2.0, 11.0 / 3.0) * 0.3 * pow((3.0 * kPi * kPi), 2.0 / 3.0);
    const double rm13 = pow(rho, -(1.0 / 3.0));
    const double den = 1.0 + D * rm13;
    const double omega = exp(-C * rm13) * pow(rho, -(11.0 / 3.0)) / den;
    const double delta = C * rm13 + D * rm13 / den;
    const double abw = AB * omega;
    const double rab = ra * ra;
    const double ra83 = pow(ra, 8.0 / 3.0);
    // d f / d gamma (roundLYP_roundGamma at rhoA = rhoB): AA == BB
    const double half = ra / rho;  // 1/2
    const double gaa = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * half) - rab);
    const double gab = -abw * ((1.0 / 9.0) * rab * (47.0 - 7.0 * delta) - (4.0 / 3.0) * rho * rho);
    *e = -4.0 * A / den * rab / rho - lyp_c * AB * omega * rab * (ra83 + ra83) + (gaa + gab + gaa) * g;
    *v = 0.0;
    *vgaa = 0.0;
    *vgab = 0.0;
    if (ra > 1.0e-20) {
        const double inv = 1.0 / rho;
        const double rm43 = pow(rho, -4.0 / 3.0);
        const double omegap = -(1.0 / 3.0) * ((1.0 / rho) * (1.0 / pow(rho, 1.0 / 3.0))) * omega *
                              (11.0 * pow(rho, 1.0 / 3.0) - C - (D / den));
        const double dd = 1.0 + D * rm13;
        const double deltap = (1.0 / 3.0) * ((D * D) * pow(rho, -(5.0 / 3.0)) / (dd * dd) + (-1.0 / rho * delta));
        const double wpw = -(1.0 / 3.0) * rm43 * (11.0 * pow(rho, 1.0 / 3.0) - C - D / den);
        const double s1 = 1.0 - 3.0 * delta - (delta - 11.0) * ra * inv;
        // d2f / drhoA dgammaXX (roundRoundLYP_roundRhoRoundGamma)
        const double rA_AA = wpw * gaa + (-abw * ((1.0 / 9.0) * ra * s1 +
                                                  -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap + (delta - 11.0) * ra * inv * inv)));
        const double t21 = (1.0 / 9.0) * (47.0 - 7.0 * delta);
        const double t22 = -(7.0 / 9.0) * rab * deltap - (8.0 / 3.0) * rho;
        const double rA_AB = wpw * gab + (-abw * (t21 * ra + t22));
        const double rA_BB = wpw * gaa + (-abw * ((1.0 / 9.0) * ra * s1 +
                                                  -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap - (delta - 11.0) * ra * inv * inv) +
                                                  -2.0 * ra));
        const double a1c = -4.0 * A / den * rab / rho;
        const double a1 = a1c * ((1.0 / 3.0) * D * rm43 / den + 1.0 / ra - inv);
        const double a2 = -lyp_c * AB * (omegap * rab * (ra83 + ra83) + omega * ra * ((11.0 / 3.0) * ra83 + ra83));
        *v = a1 + a2 + rA_AA * g + rA_AB * g + rA_BB * g;
        *vgaa = gaa;
        *vgab = gab;
    }
}

}  // namespace xcf

// ---------------------------------------------------------------------------
// Spin-resolved (UKS) forms: the same reference formulas at general
// (rhoA, rhoB, gAA, gAB, gBB), for buildVxc UKS (DfCalcGridX.cpp:2093-2135,
// :2195-2278).  Each returns f and its five first derivatives.
// ---------------------------------------------------------------------------
struct XcPointU {
    double e, va, vb, vgaa, vgab, vgbb;
};

namespace xcu {

using xcf::kPi;

__device__ __host__ inline void slater(double ra, double rb, XcPointU& o, double s) {
    const double c13 = pow(3.0 / (4.0 * kPi), 1.0 / 3.0);
    const double slater_c = -(9.0 / 4.0) * c13 * (2.0 / 3.0);
    const double slater_dc = -3.0 * c13 * (2.0 / 3.0);
    o.e += s * (slater_c * (pow(ra, 4.0 / 3.0) + pow(rb, 4.0 / 3.0)));
    o.va += s * (slater_dc * pow(ra, 1.0 / 3.0));
    o.vb += s * (slater_dc * pow(rb, 1.0 / 3.0));
}

// one spin of B88: e part, and (v, vg) where rho and gamma exceed 1e-16
template <bool WITH_LDA>
__device__ __host__ inline void b88_spin(double r, double g, double* e, double* v, double* vg) {
    *e = 0.0;
    *v = 0.0;
    *vg = 0.0;
    if (!(r > 1.0e-16)) return;
    const double r43 = pow(r, 4.0 / 3.0);
    const double sg = sqrt(g);
    const double x = sg / r43;
    const double bx = 0.0042 * x;
    const double ash = log(x + sqrt(x * x + 1));
    const double M = 1.0 + 6.0 * bx * ash;
    const double gx = WITH_LDA ? (-(3.0 / 2.0) * pow(3.0 / (4.0 * kPi), 1.0 / 3.0) - ((bx * x) / M)) : -((bx * x) / M);
    *e = r43 * gx;
    if (g > 1.0e-16) {
        const double invM = 1.0 / M;
        const double gp = 6.0 * bx * bx * (ash + x / sqrt(x * x + 1.0)) * invM * invM - 2.0 * bx * invM;
        *v = (4.0 / 3.0) * pow(r, 1.0 / 3.0) * (gx - x * gp);
        *vg = 0.5 * (1.0 / sg) * gp;
    }
}

template <bool WITH_LDA>
__device__ __host__ inline void b88(double ra, double rb, double gaa, double gbb, XcPointU& o, double s) {
    double ea, va, ga, eb, vb, gb;
    b88_spin<WITH_LDA>(ra, gaa, &ea, &va, &ga);
    b88_spin<WITH_LDA>(rb, gbb, &eb, &vb, &gb);
    o.e += s * (ea + eb);
    o.va += s * va;
    o.vb += s * vb;
    o.vgaa += s * ga;
    o.vgbb += s * gb;
}

// VWN with spin interpolation (eps_p + eps_a g(z) (1 + h z^4)), VWN eq. 4.4
template <bool VWN3>
__device__ __host__ inline void vwn(double ra, double rb, XcPointU& o, double s) {
    const double rho = ra + rb;
    const double inv = 1.0 / rho;
    const double x = pow((3.0 / (4.0 * kPi)) * inv, 1.0 / 6.0);
    // zeta clamped to [-1, 1]: at a one-spin point (ra - rb) * inv can round
    // past 1, and with the product contracted into 1 - z (an FMA) the
    // pow(1 - z, 4/3) below would see -1e-17 and return NaN
    const double z = fmin(fmax((ra - rb) * inv, -1.0), 1.0);
    double ep, epp, ef, efp, ea, eap;
    if (VWN3) {
        xcf::vwn_eps(0.0310907, 13.0720, 42.7198, -0.409286, x, &ep, &epp);
        xcf::vwn_eps(0.01554535, 20.1231, 101.578, -0.743294, x, &ef, &efp);
    } else {
        xcf::vwn_eps(0.0310907, 3.72744, 12.9352, -0.10498, x, &ep, &epp);
        xcf::vwn_eps(0.01554535, 7.06042, 18.0578, -0.325, x, &ef, &efp);
    }
    xcf::vwn_eps(-1.0 / (6.0 * kPi * kPi), 1.13107, 13.0045, -0.0047584, x, &ea, &eap);
    const double C = 4.0 / (9.0 * (pow(2.0, 1.0 / 3.0) - 1.0));
    const double g = (9.0 / 8.0) * (pow(1.0 + z, 4.0 / 3.0) + pow(1.0 - z, 4.0 / 3.0) - 2.0);
    const double gp = (3.0 / 2.0) * (pow(1.0 + z, 1.0 / 3.0) - pow(1.0 - z, 1.0 / 3.0));
    const double z3 = z * z * z, z4 = z3 * z;
    const double EC = ep + ea * g * (1.0 + (C * (ef - ep) / ea - 1.0) * z4);
    o.e += s * (rho * EC);
    const double h = C * ((ef - ep) / ea) - 1.0;
    const double invea = 1.0 / ea;
    const double hp = C * invea * (efp - epp - (ef - ep) * invea * eap);
    const double term1 = -x / (6.0 * rho) * (epp + eap * g * (1.0 + h * z4) + ea * g * hp * z4);
    const double t2c = ea * (gp * (1.0 + h * z4) + 4.0 * g * h * z3);
    if (ra > 1.0e-20) o.va += s * (EC + rho * (term1 + t2c * (inv * (1.0 - z))));
    if (rb > 1.0e-20) o.vb += s * (EC + rho * (term1 + t2c * (-inv * (1.0 + z))));
}

__device__ __host__ inline void lyp(double ra, double rb, double gaa, double gab, double gbb, XcPointU& o,
                                    double s) {
    const double A = 0.04918, B = 0.132, C = 0.2533, D = 0.349;
    const double AB = A * B;
    const double lyp_c = pow(2.0, 11.0 / 3.0) * 0.3 * pow((3.0 * kPi * kPi), 2.0 / 3.0);
    const double rho = ra + rb;
    const double rab = ra * rb;
    const double rm13 = pow(rho, -(1.0 / 3.0));
    const double den = 1.0 + D * rm13;
    const double omega = exp(-C * rm13) * pow(rho, -(11.0 / 3.0)) / den;
    const double delta = C * rm13 + D * rm13 / den;
    const double abw = AB * omega;
    const double ra83 = pow(ra, 8.0 / 3.0), rb83 = pow(rb, 8.0 / 3.0);
    // d f / d gamma_XY
    const double daa = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * ra / rho) - rb * rb);
    const double dbb = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * rb / rho) - ra * ra);
    const double dab = -abw * ((1.0 / 9.0) * rab * (47.0 - 7.0 * delta) - (4.0 / 3.0) * rho * rho);
    o.e += s * (-4.0 * A / den * rab / rho - lyp_c * AB * omega * rab * (ra83 + rb83) + daa * gaa + dab * gab +
                dbb * gbb);
    if (!(fabs(ra) > 1.0e-20 && fabs(rb) > 1.0e-20)) return;
    const double inv = 1.0 / rho;
    const double r13 = pow(rho, 1.0 / 3.0);
    const double rm43 = pow(rho, -4.0 / 3.0);
    const double omegap = -(1.0 / 3.0) * ((1.0 / rho) * (1.0 / r13)) * omega * (11.0 * r13 - C - (D / (1.0 + D * (1.0 / r13))));
    const double dd = 1.0 + D * rm13;
    const double deltap = (1.0 / 3.0) * ((D * D) * pow(rho, -(5.0 / 3.0)) / (dd * dd) + (-1.0 / rho * delta));
    const double wpw = -(1.0 / 3.0) * rm43 * (11.0 * pow(rho, 1.0 / 3.0) - C - D / den);
    // d2f / drhoX dgammaYZ
    const double rA_AA = wpw * daa + (-abw * ((1.0 / 9.0) * rb * (1.0 - 3.0 * delta - (delta - 11.0) * ra * inv) +
                                              -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap + (delta - 11.0) * rb * inv * inv)));
    const double rB_BB = wpw * dbb + (-abw * ((1.0 / 9.0) * ra * (1.0 - 3.0 * delta - (delta - 11.0) * rb * inv) +
                                              -(1.0 / 9.0) * rab * ((3.0 + rb * inv) * deltap + (delta - 11.0) * ra * inv * inv)));
    const double t21 = (1.0 / 9.0) * (47.0 - 7.0 * delta);
    const double t22 = -(7.0 / 9.0) * rab * deltap - (8.0 / 3.0) * rho;
    const double rA_AB = wpw * dab + (-abw * (t21 * rb + t22));
    const double rB_AB = wpw * dab + (-abw * (t21 * ra + t22));
    const double rA_BB = wpw * dbb + (-abw * ((1.0 / 9.0) * rb * (1.0 - 3.0 * delta - (delta - 11.0) * rb * inv) +
                                              -(1.0 / 9.0) * rab * ((3.0 + rb * inv) * deltap - (delta - 11.0) * rb * inv * inv) +
                                              -2.0 * ra));
    const double rB_AA = wpw * daa + (-abw * ((1.0 / 9.0) * ra * (1.0 - 3.0 * delta - (delta - 11.0) * ra * inv) +
                                              -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap - (delta - 11.0) * ra * inv * inv) +
                                              -2.0 * rb));
    const double a1c = -4.0 * A / den * rab / rho;
    const double a1A = a1c * ((1.0 / 3.0) * D * rm43 / den + 1.0 / ra - inv);
    const double a1B = a1c * ((1.0 / 3.0) * D * rm43 / den + 1.0 / rb - inv);
    const double a2A = -lyp_c * AB * (omegap * rab * (ra83 + rb83) + omega * rb * ((11.0 / 3.0) * ra83 + rb83));
    const double a2B = -lyp_c * AB * (omegap * rab * (rb83 + ra83) + omega * ra * ((11.0 / 3.0) * rb83 + ra83));
    o.va += s * (a1A + a2A + rA_AA * gaa + rA_AB * gab + rA_BB * gbb);
    o.vb += s * (a1B + a2B + rB_AA * gaa + rB_AB * gab + rB_BB * gbb);
    o.vgaa += s * daa;
    o.vgab += s * dab;
    o.vgbb += s * dbb;
}

}  // namespace xcu

// One spin-resolved point of functional `type` (an XcType other than kHF).
__device__ __host__ inline XcPointU eval_uks(int type, double ra, double rb, double gaa, double gab, double gbb) {
    XcPointU o{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    switch (type) {
        case kHFS:
            xcu::slater(ra, rb, o, 1.0);
            break;
        case kSVWN:
            xcu::slater(ra, rb, o, 1.0);
            xcu::vwn<false>(ra, rb, o, 1.0);
            break;
        case kHFB:
            xcu::b88<true>(ra, rb, gaa, gbb, o, 1.0);
            break;
        case kBLYP:
            xcu::b88<true>(ra, rb, gaa, gbb, o, 1.0);
            xcu::lyp(ra, rb, gaa, gab, gbb, o, 1.0);
            break;
        case kB3LYP:
            xcu::slater(ra, rb, o, 0.80);
            xcu::b88<false>(ra, rb, gaa, gbb, o, 0.72);
            xcu::vwn<true>(ra, rb, o, 0.19);
            xcu::lyp(ra, rb, gaa, gab, gbb, o, 0.81);
            break;
        default:
            break;
    }
    return o;
}

// One closed-shell point of functional `type` (an XcType other than kHF).
__device__ __host__ inline XcPoint eval_rks(int type, double ra, double g) {
    XcPoint o{0.0, 0.0, 0.0, 0.0};
    const double ra13 = pow(ra, 1.0 / 3.0);
    const double ra43 = pow(ra, 4.0 / 3.0);
    switch (type) {
        case kHFS:
            xcf::slater(ra13, ra43, &o.e, &o.vr);
            break;
        case kSVWN: {
            double es, vs, ev, vv;
            xcf::slater(ra13, ra43, &es, &vs);
            xcf::vwn<false>(ra, &ev, &vv);
            o.e = es + ev;
            o.vr = vs + vv;
        } break;
        case kHFB:
            xcf::b88<true>(ra, g, ra13, ra43, &o.e, &o.vr, &o.vgaa);
            break;
        case kBLYP: {
            double eb, vb, gb, el, vl, gal, gbl;
            xcf::b88<true>(ra, g, ra13, ra43, &eb, &vb, &gb);
            xcf::lyp(ra, g, &el, &vl, &gal, &gbl);
            o.e = eb + el;
            o.vr = vb + vl;
            o.vgaa = gb + gal;
            o.vgab = 0.0 + gbl;
        } break;
        case kB3LYP: {
            double es, vs, eb, vb, gb, ev, vv, el, vl, gal, gbl;
            xcf::slater(ra13, ra43, &es, &vs);
            xcf::b88<false>(ra, g, ra13, ra43, &eb, &vb, &gb);
            xcf::vwn<true>(ra, &ev, &vv);
            xcf::lyp(ra, g, &el, &vl, &gal, &gbl);
            o.e = 0.80 * es + 0.72 * eb + 0.19 * ev + 0.81 * el;
            o.vr = 0.80 * vs + 0.72 * vb + 0.19 * vv + 0.81 * vl;
            o.vgaa = 0.72 * gb + 0.81 * gal;
            o.vgab = 0.72 * 0.0 + 0.81 * gbl;
        } break;
        default:
            break;
    }
    return o;
}

}  // namespace pdfxc
