// B3LYP for closed-shell points, as ProteinDF's grid path evaluates it.
//
// The grid path reaches the functional through the non-virtual closed-shell
// forwarders of DfFunctional_GGA (DfFunctional.h:190-192, :221-232), i.e. the
// spin-resolved formulas at rhoB = rhoA, gAB = gBB = gAA.  This device code
// evaluates those formulas specialised to that point: the spin-polarisation
// terms of VWN vanish exactly at zeta = 0 (they are multiplied by g(0) = 0 and
// g'(0) = 0), and the alpha and beta halves of Slater and B88 are equal.
//
//   f = 0.80 Slater + 0.72 B88(non-LDA) + 0.19 VWN3 + 0.81 LYP
//       (DfFunctional_B3LYP.cpp:25-28, :59-75, :91-141)
//   Slater  DfFunctional_Slater.cpp:25-52, :64-76
//   B88     DfFunctional_Becke88.cpp:25-31, :40-57, :71-108, :137-163;
//           DfFunctional_Becke88_ExceptLDA.cpp:26-33; asinh = log(x+sqrt(x^2+1)) (TlMath.h:57-59)
//   VWN3    DfFunctional_VWN.cpp:76-124, :145-260, :262-346; DfFunctional_VWN3.cpp:23-54
//   LYP     DfFunctional_LYP.cpp:25-43, :64-197, :288-317, :346-386, :507-594, :652-674
//
// Outputs: e = f(rhoA, rhoA, g, g, g); vr = df/drhoA; vgaa = df/dgammaAA;
// vgab = df/dgammaAB.  All FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace pdfxc {

struct XcPoint {
    double e, vr, vgaa, vgab;
};

namespace b3lyp {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSlaterCoefs = 0.80, kB88Coef = 0.72, kVwnCoef = 0.19, kLypCoef = 0.81;

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

__device__ __host__ inline XcPoint eval_rks(double ra, double g) {
    XcPoint o;
    const double rho = ra + ra;
    const double ra13 = pow(ra, 1.0 / 3.0);
    const double ra43 = pow(ra, 4.0 / 3.0);

    // Slater: e = C (rhoA^4/3 + rhoB^4/3), v = C' rhoA^1/3
    const double c13 = pow(3.0 / (4.0 * kPi), 1.0 / 3.0);
    const double slater_c = -(9.0 / 4.0) * c13 * (2.0 / 3.0);
    const double slater_dc = -3.0 * c13 * (2.0 / 3.0);
    const double e_s = slater_c * (ra43 + ra43);
    const double v_s = slater_dc * ra13;

    // B88 without the LDA part; x = sqrt(gamma)/rho^4/3
    double e_b = 0.0, v_b = 0.0, vg_b = 0.0;
    if (ra > 1.0e-16) {
        const double sg = sqrt(g);
        const double x = sg / ra43;
        const double bx = 0.0042 * x;
        const double ash = log(x + sqrt(x * x + 1));
        const double M = 1.0 + 6.0 * bx * ash;
        const double gx = -((bx * x) / M);
        e_b = ra43 * gx + ra43 * gx;
        if (g > 1.0e-16) {
            const double invM = 1.0 / M;
            const double gp = 6.0 * bx * bx * (ash + x / sqrt(x * x + 1.0)) * invM * invM - 2.0 * bx * invM;
            v_b = (4.0 / 3.0) * ra13 * (gx - x * gp);
            vg_b = 0.5 * (1.0 / sg) * gp;
        }
    }

    // VWN3 at zeta = 0: eps_c = eps_para; v = eps + rho * (-x/(6 rho)) eps'
    const double xv = pow((3.0 / (4.0 * kPi)) * (1.0 / rho), 1.0 / 6.0);
    double ep, epp;
    vwn_eps(0.0310907, 13.0720, 42.7198, -0.409286, xv, &ep, &epp);
    const double e_v = rho * ep;
    const double v_v = (ra > 1.0e-20) ? ep + rho * (-xv / (6.0 * rho) * epp) : 0.0;

    // LYP with rhoA = rhoB = ra, gAA = gAB = gBB = g
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
    const double e_l = -4.0 * A / den * rab / rho - lyp_c * AB * omega * rab * (ra83 + ra83) +
                       (gaa + gab + gaa) * g;
    double v_l = 0.0, vgaa_l = 0.0, vgab_l = 0.0;
    if (ra > 1.0e-20) {
        const double inv = 1.0 / rho;
        const double r13 = 1.0 / rm13;
        const double rm43 = pow(rho, -4.0 / 3.0);
        const double omegap = -(1.0 / 3.0) * ((1.0 / rho) * (1.0 / pow(rho, 1.0 / 3.0))) * omega *
                              (11.0 * pow(rho, 1.0 / 3.0) - C - (D / den));
        (void)r13;
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
        v_l = a1 + a2 + rA_AA * g + rA_AB * g + rA_BB * g;
        vgaa_l = gaa;
        vgab_l = gab;
    }

    o.e = kSlaterCoefs * e_s + kB88Coef * e_b + kVwnCoef * e_v + kLypCoef * e_l;
    o.vr = kSlaterCoefs * v_s + kB88Coef * v_b + kVwnCoef * v_v + kLypCoef * v_l;
    o.vgaa = kB88Coef * vg_b + kLypCoef * vgaa_l;
    o.vgab = kB88Coef * 0.0 + kLypCoef * vgab_l;
    return o;
}

}  // namespace b3lyp
}  // namespace pdfxc
