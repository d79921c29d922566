/*
 * CPU oracle for the grid XC build -- test infrastructure only (see
 * pdfxc_oracle.h).  Plain C99 + OpenMP.  Each routine restates the reference
 * algorithm and cites the ProteinDF lines it follows.  Operation order is kept
 * close to the reference where that is cheap, so that results agree to a few
 * ulps rather than only to the physical tolerance.
 */
#include "pdfxc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Slater exchange: DfFunctional_Slater.cpp:25-34 (constants), :44-52 (e),    */
/* :64-76 (derivative).  Slater alpha = 2/3.                                  */
/* ------------------------------------------------------------------------ */
static double slater_coef(void) {
    return -(9.0 / 4.0) * pow((3.0 / (4.0 * M_PI)), (1.0 / 3.0)) * (2.0 / 3.0);
}
static double slater_dcoef(void) {
    return -3.0 * pow((3.0 / (4.0 * M_PI)), (1.0 / 3.0)) * (2.0 / 3.0);
}
static double slater_e(double ra, double rb) {
    return slater_coef() * (pow(ra, 4.0 / 3.0) + pow(rb, 4.0 / 3.0));
}
static void slater_d(double ra, double rb, double* va, double* vb) {
    *va = slater_dcoef() * pow(ra, 1.0 / 3.0);
    *vb = slater_dcoef() * pow(rb, 1.0 / 3.0);
}

/* ------------------------------------------------------------------------ */
/* Becke 88: DfFunctional_Becke88.cpp:25-31 (constants), :40-57 (e), :71-108  */
/* (derivatives), :137-163 (x, g, g'); the "except LDA" variant drops G_PARAM */
/* from g (DfFunctional_Becke88_ExceptLDA.cpp:26-33).  asinh as log+sqrt      */
/* (TlMath.h:57-59).                                                          */
/* ------------------------------------------------------------------------ */
#define B88_EPS 1.0e-16
#define B88_B 0.0042
static double b88_gparam(void) {
    return -(3.0 / 2.0) * pow(3.0 / (4.0 * M_PI), (1.0 / 3.0));
}
static double tl_arcsinh(double x) { return log(x + sqrt(x * x + 1)); }
static double b88_x(double rho, double gamma) {
    return sqrt(gamma) / pow(rho, 4.0 / 3.0);
}
static double b88_g(double x, int with_lda) {
    const double bx = B88_B * x;
    const double ash = tl_arcsinh(x);
    const double nonlocal = -((bx * x) / (1.0 + 6.0 * bx * ash));
    return with_lda ? (b88_gparam() - ((bx * x) / (1.0 + 6.0 * bx * ash)))
                    : nonlocal;
}
static double b88_gprime(double x) {
    const double bx = B88_B * x;
    const double xx = x * x;
    const double ash = tl_arcsinh(x);
    const double M = 1.0 + 6.0 * bx * ash;
    const double invM = 1.0 / M;
    const double arg1 = 6.0 * bx * bx * (ash + x / sqrt(xx + 1.0)) * invM * invM;
    const double arg2 = -2.0 * bx * invM;
    return arg1 + arg2;
}
static double b88_e(double ra, double rb, double gaa, double gbb, int with_lda) {
    double ta = 0.0, tb = 0.0;
    if (ra > B88_EPS) {
        const double xa = b88_x(ra, gaa);
        ta = pow(ra, 4.0 / 3.0) * b88_g(xa, with_lda);
    }
    if (rb > B88_EPS) {
        const double xb = b88_x(rb, gbb);
        tb = pow(rb, 4.0 / 3.0) * b88_g(xb, with_lda);
    }
    return ta + tb;
}
static void b88_d(double ra, double rb, double gaa, double gbb, int with_lda,
                  double* vra, double* vrb, double* vgaa, double* vgab,
                  double* vgbb) {
    *vgab = 0.0;
    *vra = 0.0;
    *vgaa = 0.0;
    if ((ra > B88_EPS) && (gaa > B88_EPS)) {
        const double xa = b88_x(ra, gaa);
        const double ga = b88_g(xa, with_lda);
        const double gpa = b88_gprime(xa);
        *vra = (4.0 / 3.0) * pow(ra, 1.0 / 3.0) * (ga - xa * gpa);
        *vgaa = 0.5 * (1.0 / sqrt(gaa)) * gpa;
    }
    *vrb = 0.0;
    *vgbb = 0.0;
    if ((rb > B88_EPS) && (gbb > B88_EPS)) {
        const double xb = b88_x(rb, gbb);
        const double gb = b88_g(xb, with_lda);
        const double gpb = b88_gprime(xb);
        *vrb = (4.0 / 3.0) * pow(rb, 1.0 / 3.0) * (gb - xb * gpb);
        *vgbb = 0.5 * (1.0 / sqrt(gbb)) * gpb;
    }
}

/* ------------------------------------------------------------------------ */
/* VWN correlation: DfFunctional_VWN.cpp:23-46 (parameters), :76-124 (e),     */
/* :145-260 (eps_c, its x-derivative, g, h, g', h'), :262-346 (derivative).   */
/* VWN3 replaces the para/ferro sets and keeps the VWN5 anti-ferro set        */
/* (DfFunctional_VWN3.cpp:23-54).                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    double A, b, c, x0;
} vwn_set;
static const vwn_set VWN5_PARA = {0.0310907, 3.72744, 12.9352, -0.10498};
static const vwn_set VWN5_FERR = {0.01554535, 7.06042, 18.0578, -0.325};
static const vwn_set VWN3_PARA = {0.0310907, 13.0720, 42.7198, -0.409286};
static const vwn_set VWN3_FERR = {0.01554535, 20.1231, 101.578, -0.743294};
static vwn_set vwn_anti(void) {
    vwn_set s = {-1.0 / (6.0 * M_PI * M_PI), 1.13107, 13.0045, -0.0047584};
    return s;
}
static double vwn_ec_coef(void) { return 4.0 / (9.0 * (pow(2.0, 1.0 / 3.0) - 1.0)); }

static double vwn_eps(vwn_set s, double x) {
    const double X = x * (x + s.b) + s.c;
    const double invX = 1.0 / X;
    const double X0 = s.x0 * (s.x0 + s.b) + s.c;
    const double invX0 = 1.0 / X0;
    const double Q = sqrt(4 * s.c - s.b * s.b);
    const double invQ = 1.0 / Q;
    const double at = atan(Q / (2.0 * x + s.b));
    const double arg1 = log(x * x * invX);
    const double arg2 = 2.0 * s.b * invQ * at;
    const double xx0 = x - s.x0;
    const double arg3_1 = log(xx0 * xx0 * invX);
    const double arg3_2 = 2.0 * (2.0 * s.x0 + s.b) * invQ * at;
    const double arg3 = -s.b * s.x0 * invX0 * (arg3_1 + arg3_2);
    return s.A * (arg1 + arg2 + arg3);
}
static double vwn_eps_prime(vwn_set s, double x) {
    const double X = x * (x + s.b) + s.c;
    const double invX = 1.0 / X;
    const double X0 = s.x0 * (s.x0 + s.b) + s.c;
    const double invX0 = 1.0 / X0;
    const double Q = sqrt(4 * s.c - s.b * s.b);
    const double x2b = 2.0 * x + s.b;
    const double inv_x2bQ = 1.0 / (x2b * x2b + Q * Q);
    const double d1 = 2.0 / x;
    const double d2 = -x2b * invX;
    const double d3 = -4.0 * s.b * inv_x2bQ;
    const double d4 = -s.b * s.x0 * invX0 *
                      (2.0 / (x - s.x0) - x2b * invX - 4.0 * (2.0 * s.x0 + s.b) * inv_x2bQ);
    return s.A * (d1 + d2 + d3 + d4);
}
static double vwn_gz(double z) {
    return (9.0 / 8.0) * (pow((1.0 + z), 4.0 / 3.0) + pow((1.0 - z), 4.0 / 3.0) - 2.0);
}
static double vwn_gz_prime(double z) {
    return (3.0 / 2.0) * (pow((1.0 + z), 1.0 / 3.0) - pow((1.0 - z), 1.0 / 3.0));
}
static double vwn_epsc(vwn_set para, vwn_set ferr, double x, double zeta) {
    const double ep = vwn_eps(para, x);
    const double ef = vwn_eps(ferr, x);
    const double ea = vwn_eps(vwn_anti(), x);
    const double g = vwn_gz(zeta);
    const double z4 = zeta * zeta * zeta * zeta;
    const double coef2 = 1.0 + (vwn_ec_coef() * (ef - ep) / ea - 1.0) * z4;
    return ep + ea * g * coef2;
}
static double vwn_e(double ra, double rb, int v3) {
    const vwn_set para = v3 ? VWN3_PARA : VWN5_PARA;
    const vwn_set ferr = v3 ? VWN3_FERR : VWN5_FERR;
    const double rho = ra + rb;
    const double inv = 1.0 / rho;
    const double x = pow((3.0 / (4.0 * M_PI)) * inv, 1.0 / 6.0);
    const double zeta = (ra - rb) * inv;
    return rho * vwn_epsc(para, ferr, x, zeta);
}
static void vwn_d(double ra, double rb, int v3, double* va, double* vb) {
    const vwn_set para = v3 ? VWN3_PARA : VWN5_PARA;
    const vwn_set ferr = v3 ? VWN3_FERR : VWN5_FERR;
    const vwn_set anti = vwn_anti();
    const double C = vwn_ec_coef();
    *va = 0.0;
    *vb = 0.0;
    const double rho = ra + rb;
    const double inv = 1.0 / rho;
    const double x = pow((3.0 / (4.0 * M_PI)) * inv, 1.0 / 6.0);
    const double zeta = (ra - rb) * inv;
    const double EC = vwn_epsc(para, ferr, x, zeta);
    const double ep = vwn_eps(para, x), ef = vwn_eps(ferr, x), ea = vwn_eps(anti, x);
    const double epp = vwn_eps_prime(para, x), efp = vwn_eps_prime(ferr, x),
                 eap = vwn_eps_prime(anti, x);
    const double g = vwn_gz(zeta);
    const double gp = vwn_gz_prime(zeta);
    const double h = C * ((ef - ep) / ea) - 1.0;
    const double invea = 1.0 / ea;
    const double hp = C * invea * (efp - epp - (ef - ep) * invea * eap);
    const double z3 = zeta * zeta * zeta;
    const double z4 = z3 * zeta;
    const double term1 = -x / (6.0 * rho) * (epp + eap * g * (1.0 + h * z4) + ea * g * hp * z4);
    const double t2c = ea * (gp * (1.0 + h * z4) + 4.0 * g * h * z3);
    const double dzA = inv * (1.0 - zeta);
    const double dzB = -inv * (1.0 + zeta);
    const double dECA = term1 + t2c * dzA;
    const double dECB = term1 + t2c * dzB;
    if (ra > 1.0e-20) *va = EC + rho * dECA;
    if (rb > 1.0e-20) *vb = EC + rho * dECB;
}

/* ------------------------------------------------------------------------ */
/* LYP correlation: DfFunctional_LYP.cpp:25-43 (constants), :64-197           */
/* (derivatives), :288-317 (energy), :346-386 (omega, delta, d/dgamma),       */
/* :507-594 (mixed second derivatives), :652-674 (omega', delta').            */
/* ------------------------------------------------------------------------ */
#define LYP_TOL 1.0e-20
#define LYP_A 0.04918
#define LYP_B 0.132
#define LYP_C 0.2533
#define LYP_D 0.349
static double lyp_coef(void) {
    return pow(2.0, 11.0 / 3.0) * 0.3 * pow((3.0 * M_PI * M_PI), 2.0 / 3.0);
}
static double lyp_omega(double rho) {
    const double m13 = pow(rho, -(1.0 / 3.0));
    const double m113 = pow(rho, -(11.0 / 3.0));
    const double t1 = exp(-LYP_C * m13);
    const double t2 = 1.0 + LYP_D * m13;
    return t1 * m113 / t2;
}
static double lyp_delta(double rho) {
    const double m13 = pow(rho, -(1.0 / 3.0));
    return LYP_C * m13 + LYP_D * m13 / (1.0 + LYP_D * m13);
}
static double lyp_omega_prime(double rho, double omega) {
    const double r13 = pow(rho, 1.0 / 3.0);
    const double rm13 = 1.0 / r13;
    const double rm43 = (1.0 / rho) * rm13;
    return -(1.0 / 3.0) * rm43 * omega * (11.0 * r13 - LYP_C - (LYP_D / (1.0 + LYP_D * rm13)));
}
static double lyp_delta_prime(double rho, double delta) {
    const double a11 = 1.0 + LYP_D * pow(rho, -(1.0 / 3.0));
    const double a1 = ((LYP_D * LYP_D) * pow(rho, -(5.0 / 3.0))) / (a11 * a11);
    const double a2 = -1.0 / rho * delta;
    return (1.0 / 3.0) * (a1 + a2);
}
static void lyp_dgamma(double rho, double ra, double rb, double omega, double delta,
                       double* daa, double* dab, double* dbb) {
    const double rab = ra * rb;
    const double abw = (LYP_A * LYP_B) * omega;
    *daa = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * ra / rho) - rb * rb);
    *dbb = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * rb / rho) - ra * ra);
    *dab = -abw * ((1.0 / 9.0) * rab * (47.0 - 7.0 * delta) - (4.0 / 3.0) * rho * rho);
}
static double lyp_e(double ra, double rb, double gaa, double gab, double gbb) {
    const double rho = ra + rb;
    const double rab = ra * rb;
    const double pm13 = pow(rho, -(1.0 / 3.0));
    const double pa83 = pow(ra, 8.0 / 3.0);
    const double pb83 = pow(rb, 8.0 / 3.0);
    const double omega = lyp_omega(rho);
    const double delta = lyp_delta(rho);
    double daa, dab, dbb;
    lyp_dgamma(rho, ra, rb, omega, delta, &daa, &dab, &dbb);
    const double term1 = -4.0 * LYP_A / (1.0 + LYP_D * pm13) * rab / rho;
    const double term2 = -lyp_coef() * (LYP_A * LYP_B) * omega * rab * (pa83 + pb83);
    return term1 + term2 + daa * gaa + dab * gab + dbb * gbb;
}
static void lyp_rr(double rho, double ra, double rb, double inv, double rab,
                   double omega, double delta, double deltap, double* rA_AA,
                   double* rB_AA, double* rA_AB, double* rB_AB, double* rA_BB,
                   double* rB_BB) {
    const double abw = (LYP_A * LYP_B) * omega;
    const double wpw = -(1.0 / 3.0) * pow(rho, -(4.0 / 3.0)) *
                       (11.0 * pow(rho, 1.0 / 3.0) - LYP_C -
                        LYP_D / (1.0 + LYP_D * pow(rho, -(1.0 / 3.0))));
    double daa, dab, dbb;
    lyp_dgamma(rho, ra, rb, omega, delta, &daa, &dab, &dbb);
    {
        const double t1A = wpw * daa;
        const double t2A_1 = (1.0 / 9.0) * rb * (1.0 - 3.0 * delta - (delta - 11.0) * ra * inv);
        const double t2A_2 = -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap + (delta - 11.0) * rb * inv * inv);
        const double t2A = -abw * (t2A_1 + t2A_2);
        const double t1B = wpw * dbb;
        const double t2B_1 = (1.0 / 9.0) * ra * (1.0 - 3.0 * delta - (delta - 11.0) * rb * inv);
        const double t2B_2 = -(1.0 / 9.0) * rab * ((3.0 + rb * inv) * deltap + (delta - 11.0) * ra * inv * inv);
        const double t2B = -abw * (t2B_1 + t2B_2);
        *rA_AA = t1A + t2A;
        *rB_BB = t1B + t2B;
    }
    {
        const double t1 = wpw * dab;
        const double t2_1 = (1.0 / 9.0) * (47.0 - 7.0 * delta);
        const double t2_2 = -(7.0 / 9.0) * rab * deltap - (8.0 / 3.0) * rho;
        *rA_AB = t1 + (-abw * (t2_1 * rb + t2_2));
        *rB_AB = t1 + (-abw * (t2_1 * ra + t2_2));
    }
    {
        const double t1A = wpw * dbb;
        const double t2A_1 = (1.0 / 9.0) * rb * (1.0 - 3.0 * delta - (delta - 11.0) * rb * inv);
        const double t2A_2 = -(1.0 / 9.0) * rab * ((3.0 + rb * inv) * deltap - (delta - 11.0) * rb * inv * inv);
        const double t2A_3 = -2.0 * ra;
        const double t2A = -abw * (t2A_1 + t2A_2 + t2A_3);
        const double t1B = wpw * daa;
        const double t2B_1 = (1.0 / 9.0) * ra * (1.0 - 3.0 * delta - (delta - 11.0) * ra * inv);
        const double t2B_2 = -(1.0 / 9.0) * rab * ((3.0 + ra * inv) * deltap - (delta - 11.0) * ra * inv * inv);
        const double t2B_3 = -2.0 * rb;
        const double t2B = -abw * (t2B_1 + t2B_2 + t2B_3);
        *rA_BB = t1A + t2A;
        *rB_AA = t1B + t2B;
    }
}
static void lyp_d(double ra, double rb, double gaa, double gab, double gbb,
                  double* vra, double* vrb, double* vgaa, double* vgab, double* vgbb) {
    *vra = *vrb = *vgaa = *vgab = *vgbb = 0.0;
    const double rho = ra + rb;
    if ((fabs(ra) > LYP_TOL) && (fabs(rb) > LYP_TOL)) {
        const double rab = ra * rb;
        const double inv = 1.0 / rho;
        const double invA = 1.0 / ra;
        const double invB = 1.0 / rb;
        const double r13 = pow(rho, 1.0 / 3.0);
        const double rm13 = 1.0 / r13;
        const double rm43 = pow(rho, -4.0 / 3.0);
        const double ra83 = pow(ra, 8.0 / 3.0);
        const double rb83 = pow(rb, 8.0 / 3.0);
        const double omega = lyp_omega(rho);
        const double abw = (LYP_A * LYP_B) * omega;
        const double omegap = lyp_omega_prime(rho, omega);
        const double delta = lyp_delta(rho);
        const double deltap = lyp_delta_prime(rho, delta);
        double rA_AA, rB_AA, rA_AB, rB_AB, rA_BB, rB_BB;
        lyp_rr(rho, ra, rb, inv, rab, omega, delta, deltap, &rA_AA, &rB_AA, &rA_AB,
               &rB_AB, &rA_BB, &rB_BB);
        const double a1c = -4.0 * LYP_A / (1.0 + LYP_D * rm13) * rab / rho;
        const double a1A = a1c * ((1.0 / 3.0) * LYP_D * rm43 / (1.0 + LYP_D * rm13) + invA - inv);
        const double a1B = a1c * ((1.0 / 3.0) * LYP_D * rm43 / (1.0 + LYP_D * rm13) + invB - inv);
        const double a2A = -lyp_coef() * (LYP_A * LYP_B) *
                           (omegap * rab * (ra83 + rb83) + omega * rb * ((11.0 / 3.0) * ra83 + rb83));
        const double a2B = -lyp_coef() * (LYP_A * LYP_B) *
                           (omegap * rab * (rb83 + ra83) + omega * ra * ((11.0 / 3.0) * rb83 + ra83));
        if (ra > LYP_TOL) {
            *vra = a1A + a2A + rA_AA * gaa + rA_AB * gab + rA_BB * gbb;
            *vgaa = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * ra / rho) - rb * rb);
        }
        if (rb > LYP_TOL) {
            *vrb = a1B + a2B + rB_AA * gaa + rB_AB * gab + rB_BB * gbb;
            *vgbb = -abw * ((1.0 / 9.0) * rab * (1.0 - 3.0 * delta - (delta - 11.0) * rb / rho) - ra * ra);
        }
        if ((ra > LYP_TOL) && (rb > LYP_TOL)) {
            *vgab = -abw * ((1.0 / 9.0) * rab * (47.0 - 7.0 * delta) - (4.0 / 3.0) * rho * rho);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* B3LYP: DfFunctional_B3LYP.cpp:25-28 (coefficients), :59-75 (e),            */
/* :91-141 (derivatives).  0.80 Slater + 0.72 B88(non-LDA) + 0.19 VWN3 +      */
/* 0.81 LYP.                                                                  */
/* ------------------------------------------------------------------------ */
static void b3lyp(double ra, double rb, double gaa, double gab, double gbb, double* o) {
    const double e = 0.80 * slater_e(ra, rb) + 0.72 * b88_e(ra, rb, gaa, gbb, 0) +
                     0.19 * vwn_e(ra, rb, 1) + 0.81 * lyp_e(ra, rb, gaa, gab, gbb);
    double sA, sB, vA, vB, bA, bB, bgaa, bgab, bgbb, lA, lB, lgaa, lgab, lgbb;
    slater_d(ra, rb, &sA, &sB);
    vwn_d(ra, rb, 1, &vA, &vB);
    b88_d(ra, rb, gaa, gbb, 0, &bA, &bB, &bgaa, &bgab, &bgbb);
    lyp_d(ra, rb, gaa, gab, gbb, &lA, &lB, &lgaa, &lgab, &lgbb);
    o[0] = e;
    o[1] = 0.80 * sA + 0.72 * bA + 0.19 * vA + 0.81 * lA;
    o[2] = 0.80 * sB + 0.72 * bB + 0.19 * vB + 0.81 * lB;
    o[3] = 0.72 * bgaa + 0.81 * lgaa;
    o[4] = 0.72 * bgab + 0.81 * lgab;
    o[5] = 0.72 * bgbb + 0.81 * lgbb;
}

static void functional_point(int id, const double* x, double* o) {
    const double ra = x[0], rb = x[1], gaa = x[2], gab = x[3], gbb = x[4];
    o[0] = o[1] = o[2] = o[3] = o[4] = o[5] = 0.0;
    switch (id) {
        case ORACLE_SLATER:
            o[0] = slater_e(ra, rb);
            slater_d(ra, rb, &o[1], &o[2]);
            break;
        case ORACLE_BECKE88:
        case ORACLE_BECKE88_EXCEPT_LDA: {
            const int lda = (id == ORACLE_BECKE88);
            o[0] = b88_e(ra, rb, gaa, gbb, lda);
            b88_d(ra, rb, gaa, gbb, lda, &o[1], &o[2], &o[3], &o[4], &o[5]);
        } break;
        case ORACLE_VWN5:
        case ORACLE_VWN3: {
            const int v3 = (id == ORACLE_VWN3);
            o[0] = vwn_e(ra, rb, v3);
            vwn_d(ra, rb, v3, &o[1], &o[2]);
        } break;
        case ORACLE_LYP:
            o[0] = lyp_e(ra, rb, gaa, gab, gbb);
            lyp_d(ra, rb, gaa, gab, gbb, &o[1], &o[2], &o[3], &o[4], &o[5]);
            break;
        case ORACLE_B3LYP:
            b3lyp(ra, rb, gaa, gab, gbb, o);
            break;
        case ORACLE_HFS: /* DfFunctional_HFS.cpp: Slater alone */
            o[0] = slater_e(ra, rb);
            slater_d(ra, rb, &o[1], &o[2]);
            break;
        case ORACLE_SVWN: { /* DfFunctional_SVWN.cpp: ex + ec */
            double xa, xb, ca, cb;
            slater_d(ra, rb, &xa, &xb);
            vwn_d(ra, rb, 0, &ca, &cb);
            o[0] = slater_e(ra, rb) + vwn_e(ra, rb, 0);
            o[1] = xa + ca;
            o[2] = xb + cb;
        } break;
        case ORACLE_BLYP: { /* DfFunctional_B88LYP.cpp: ex + ec term by term */
            double x[5], c[5];
            b88_d(ra, rb, gaa, gbb, 1, &x[0], &x[1], &x[2], &x[3], &x[4]);
            lyp_d(ra, rb, gaa, gab, gbb, &c[0], &c[1], &c[2], &c[3], &c[4]);
            o[0] = b88_e(ra, rb, gaa, gbb, 1) + lyp_e(ra, rb, gaa, gab, gbb);
            for (int k = 0; k < 5; ++k) o[1 + k] = x[k] + c[k];
        } break;
        default:
            break;
    }
}

void oracle_functional(int id, int64_t n, const double* in, double* out) {
    for (int64_t i = 0; i < n; ++i) functional_point(id, in + 5 * i, out + 6 * i);
}

/* ------------------------------------------------------------------------ */
/* Basis normalisation: Fl_Gto::getNormalized (Fl_Gto.cpp:644-708),           */
/* getNormalizedfactor (:558-641), TlOrbitalInfoObject::setCGTO (:38-106).    */
/* All five d functions use the (l,m,n) = (1,1,0) norm.                       */
/* ------------------------------------------------------------------------ */
static double dbfact(int x) {
    /* TlMath::dbfact: only odd arguments (2l-1 >= -1) reach it here */
    double a = 1.0;
    while (x > 1) {
        a *= (double)x;
        x -= 2;
    }
    return a;
}
static double prim_norm(double a, int l, int m, int n) {
    const double pwr = (double)(l + m + n);
    double ans = pow(2.0, pwr);
    ans *= pow(dbfact(2 * l - 1) * dbfact(2 * m - 1) * dbfact(2 * n - 1), -1.0 / 2.0);
    ans *= pow(2.0 / M_PI, 3.0 / 4.0);
    ans *= pow(a, (pwr + 3.0 / 2.0) / 2.0);
    return ans;
}
void oracle_normalize_cgto(int shell, int nprim, const double* exps,
                           const double* coefs, double* out_exps, double* out_coefs) {
    int l = 0, m = 0, n = 0;
    if (shell == 1) l = 1;
    if (shell == 2) l = m = 1;
    const double pwr = (double)(l + m + n) + 3.0 / 2.0;
    double ans = 0.0;
    for (int a = 0; a < nprim; ++a)
        for (int b = 0; b < nprim; ++b) {
            double trm = coefs[a] * coefs[b];
            trm *= prim_norm(exps[a], l, m, n) * prim_norm(exps[b], l, m, n);
            trm *= pow(exps[a] + exps[b], -pwr);
            ans += trm;
        }
    ans *= dbfact(2 * l - 1) * dbfact(2 * m - 1) * dbfact(2 * n - 1);
    ans *= pow(2.0, -(double)(l + m + n));
    ans *= pow(M_PI, 3.0 / 2.0);
    const double nf = sqrt(1.0 / ans);
    double e[ORACLE_MAX_PRIM], c[ORACLE_MAX_PRIM];
    for (int k = 0; k < nprim; ++k) {
        e[k] = exps[k];
        c[k] = coefs[k] * prim_norm(exps[k], l, m, n) * nf;
    }
    /* stable insertion sort by ascending exponent (cgto_sort_functor) */
    for (int i = 1; i < nprim; ++i) {
        const double ei = e[i], ci = c[i];
        int j = i - 1;
        while (j >= 0 && e[j] > ei) {
            e[j + 1] = e[j];
            c[j + 1] = c[j];
            --j;
        }
        e[j + 1] = ei;
        c[j + 1] = ci;
    }
    for (int k = 0; k < nprim; ++k) {
        out_exps[k] = e[k];
        out_coefs[k] = c[k];
    }
}

/* ------------------------------------------------------------------------ */
/* AO values and gradients: DfCalcGridX::getAOs_core / getDAOs_core           */
/* (DfCalcGridX.cpp:667-769) with getPrefactor (:103-141) and                 */
/* getPrefactorForDerivative (:143-235).  A primitive is skipped when         */
/* alpha*r^2 > 30 (TOOBIG, :42).                                              */
/* ------------------------------------------------------------------------ */
#define TOOBIG 30.0
static double prefactor(int t, double x, double y, double z) {
    const double inv_sqrt3 = 1.0 / sqrt(3.0);
    switch (t) {
        case 0: return 1.0;
        case 1: return x;
        case 2: return y;
        case 3: return z;
        case 4: return x * y;
        case 5: return z * x;
        case 6: return y * z;
        case 7: return 0.5 * (x * x - y * y);
        case 8: return inv_sqrt3 * (z * z - 0.5 * (x * x + y * y));
        default: return 0.0;
    }
}
static void prefactor_d(int t, double alpha, double x, double y, double z,
                        double* px, double* py, double* pz) {
    const double a2 = 2.0 * alpha;
    const double inv_sqrt3 = 1.0 / sqrt(3.0);
    switch (t) {
        case 0: *px = -a2 * x; *py = -a2 * y; *pz = -a2 * z; break;
        case 1: *px = -a2 * x * x + 1.0; *py = -a2 * x * y; *pz = -a2 * x * z; break;
        case 2: *px = -a2 * y * x; *py = -a2 * y * y + 1.0; *pz = -a2 * y * z; break;
        case 3: *px = -a2 * z * x; *py = -a2 * z * y; *pz = -a2 * z * z + 1.0; break;
        case 4: {
            const double xy = x * y;
            *px = -a2 * xy * x + y; *py = -a2 * xy * y + x; *pz = -a2 * xy * z;
        } break;
        case 5: {
            const double xz = x * z;
            *px = -a2 * xz * x + z; *py = -a2 * xz * y; *pz = -a2 * xz * z + x;
        } break;
        case 6: {
            const double yz = y * z;
            *px = -a2 * yz * x; *py = -a2 * yz * y + z; *pz = -a2 * yz * z + y;
        } break;
        case 7: {
            const double xx = x * x, yy = y * y;
            const double xxX = -a2 * xx * x + 2.0 * x, xxY = -a2 * xx * y, xxZ = -a2 * xx * z;
            const double yyX = -a2 * yy * x, yyY = -a2 * yy * y + 2.0 * y, yyZ = -a2 * yy * z;
            *px = 0.5 * (xxX - yyX); *py = 0.5 * (xxY - yyY); *pz = 0.5 * (xxZ - yyZ);
        } break;
        case 8: {
            const double xx = x * x, yy = y * y, zz = z * z;
            const double xxX = -a2 * xx * x + 2.0 * x, xxY = -a2 * xx * y, xxZ = -a2 * xx * z;
            const double yyX = -a2 * yy * x, yyY = -a2 * yy * y + 2.0 * y, yyZ = -a2 * yy * z;
            const double zzX = -a2 * zz * x, zzY = -a2 * zz * y, zzZ = -a2 * zz * z + 2.0 * z;
            *px = inv_sqrt3 * (zzX - 0.5 * (xxX + yyX));
            *py = inv_sqrt3 * (zzY - 0.5 * (xxY + yyY));
            *pz = inv_sqrt3 * (zzZ - 0.5 * (xxZ + yyZ));
        } break;
        default: *px = *py = *pz = 0.0; break;
    }
}
static void eval_point(const oracle_aos* a, const double* p, double* phi,
                       double* gx, double* gy, double* gz) {
    for (int i = 0; i < a->n_aos; ++i) {
        const double x = p[0] - a->center[3 * i + 0];
        const double y = p[1] - a->center[3 * i + 1];
        const double z = p[2] - a->center[3 * i + 2];
        const double d2 = x * x + y * y + z * z;
        const int t = a->type[i];
        double v = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
        for (int k = 0; k < a->nprim[i]; ++k) {
            const double al = a->exps[i * ORACLE_MAX_PRIM + k];
            const double sh = al * d2;
            if (sh <= TOOBIG) {
                const double g = a->coefs[i * ORACLE_MAX_PRIM + k] * exp(-sh);
                v += prefactor(t, x, y, z) * g;
                if (gx) {
                    double px, py, pz;
                    prefactor_d(t, al, x, y, z, &px, &py, &pz);
                    dx += px * g;
                    dy += py * g;
                    dz += pz * g;
                }
            }
        }
        phi[i] = v;
        if (gx) {
            gx[i] = dx;
            gy[i] = dy;
            gz[i] = dz;
        }
    }
}
void oracle_eval_aos(const oracle_aos* aos, int64_t n_pts, const double* xyz,
                     double* phi, double* gx, double* gy, double* gz) {
    const int64_t N = aos->n_aos;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < n_pts; ++p) {
        eval_point(aos, xyz + 3 * p, phi + p * N, gx ? gx + p * N : NULL,
                   gy ? gy + p * N : NULL, gz ? gz + p * N : NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* RKS GGA two-pass build: calcRho_GGA_part (DfCalcGridX.cpp:1923-1962) with  */
/* getRhoAtGridPoint (:865-881) and getGradRhoAtGridPoint (:902-931);         */
/* buildVxc GGA RKS (:2137-2193) through the non-virtual closed-shell         */
/* forwarders of DfFunctional_GGA (DfFunctional.h:190-192, :221-232) that call */
/* the spin-resolved functional with rhoB = rhoA, gAB = gBB = gAA;            */
/* build_XC_Matrix GGA (:2455-2500).  K is accumulated per thread and summed  */
/* at the end (the reference adds with omp atomic; the order differs).        */
/* ------------------------------------------------------------------------ */
double oracle_calc_xc_rks(const oracle_aos* aos, int functional, int64_t n_pts,
                          const double* grid, const double* PA, double* FA,
                          double* rho_out, double density_cutoff, int n_threads) {
    const int N = aos->n_aos;
    int nt = 1;
#ifdef _OPENMP
    nt = (n_threads > 0) ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
#endif
    double* Fpriv = (double*)calloc((size_t)nt * N * N, sizeof(double));
    double energy = 0.0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : energy)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* F = Fpriv + (size_t)tid * N * N;
        double* phi = (double*)malloc(sizeof(double) * 4 * N);
        double* gx = phi + N;
        double* gy = gx + N;
        double* gz = gy + N;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t g = 0; g < n_pts; ++g) {
            const double* p = grid + 4 * g;
            const double w = p[3];
            eval_point(aos, p, phi, gx, gy, gz);
            /* rho: triangle loop, getRhoAtGridPoint */
            double rho = 0.0;
            for (int i = 0; i < N; ++i) {
                const double ai = phi[i];
                const double* Pi = PA + (size_t)i * N;
                for (int j = 0; j < i; ++j) rho += 2.0 * Pi[j] * ai * phi[j];
                rho += Pi[i] * ai * ai;
            }
            /* grad rho: full square, getGradRhoAtGridPoint */
            double grx = 0.0, gry = 0.0, grz = 0.0;
            for (int i = 0; i < N; ++i) {
                const double ai = phi[i];
                const double* Pi = PA + (size_t)i * N;
                for (int j = 0; j < N; ++j) {
                    const double pij = Pi[j];
                    grx += pij * ai * gx[j];
                    gry += pij * ai * gy[j];
                    grz += pij * ai * gz[j];
                }
            }
            grx *= 2.0;
            gry *= 2.0;
            grz *= 2.0;
            if (rho_out) {
                rho_out[4 * g + 0] = rho;
                rho_out[4 * g + 1] = grx;
                rho_out[4 * g + 2] = gry;
                rho_out[4 * g + 3] = grz;
            }
            if (rho > density_cutoff) {
                const double gamma = grx * grx + gry * gry + grz * grz;
                const double in[5] = {rho, rho, gamma, gamma, gamma};
                double o[6];
                functional_point(functional, in, o);
                const double vr = o[1], vgaa = o[3], vgab = o[4];
                const double coef1 = w * vr;
                const double vgaa2 = 2.0 * vgaa;
                const double cx = w * (vgaa2 * grx + vgab * grx);
                const double cy = w * (vgaa2 * gry + vgab * gry);
                const double cz = w * (vgaa2 * grz + vgab * grz);
                for (int i = 0; i < N; ++i) {
                    const double ai = phi[i];
                    const double ci = cx * gx[i] + cy * gy[i] + cz * gz[i];
                    double* Fi = F + (size_t)i * N;
                    {
                        double v = coef1 * ai * ai;
                        v += ai * ci * 2.0;
                        Fi[i] += v;
                    }
                    for (int j = 0; j < i; ++j) {
                        const double aj = phi[j];
                        double v = coef1 * ai * aj;
                        v += ai * (cx * gx[j] + cy * gy[j] + cz * gz[j]);
                        v += aj * ci;
                        Fi[j] += v;
                    }
                }
                energy += w * o[0];
            }
        }
        free(phi);
    }
    for (int t = 0; t < nt; ++t) {
        const double* F = Fpriv + (size_t)t * N * N;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j) FA[(size_t)i * N + j] += F[(size_t)i * N + j];
    }
    free(Fpriv);
    return energy;
}

/* ------------------------------------------------------------------------ */
/* UKS build: calcXCIntegForFockAndEnergy(P_A, P_B, ...) (DfCalcGridX.cpp     */
/* :2291-2303, :2317-2330) with calcRho_*_part two-spin (:1892-1921,          */
/* :1964-2016) and buildVxc UKS (:2093-2135 LDA, :2195-2278 GGA): a point     */
/* counts when rhoA or rhoB exceeds the cut-off; F_A gets build_XC_Matrix     */
/* (:2455-2500) with (vrhoA, vgAA, vgAB, grad rhoA) and F_B with (vrhoB,      */
/* vgBB, vgAB, grad rhoB).  build_XC_Matrix multiplies vgAB by the same-spin  */
/* gradient, so K_A holds vgAB grad rhoA where dE/dP_A has vgAB grad rhoB.    */
/* exact_gab = 1 (test use only) takes the other-spin gradient instead, the   */
/* exact derivative, so that finite differences can pin the rest.             */
/* ------------------------------------------------------------------------ */
static void add_point_matrix(double* F, int N, const double* phi, const double* gx, const double* gy,
                             const double* gz, double coef1, double cx, double cy, double cz) {
    for (int i = 0; i < N; ++i) {
        const double ai = phi[i];
        const double ci = cx * gx[i] + cy * gy[i] + cz * gz[i];
        double* Fi = F + (size_t)i * N;
        Fi[i] += coef1 * ai * ai + ai * ci * 2.0;
        for (int j = 0; j < i; ++j) {
            const double aj = phi[j];
            double v = coef1 * ai * aj;
            v += ai * (cx * gx[j] + cy * gy[j] + cz * gz[j]);
            v += aj * ci;
            Fi[j] += v;
        }
    }
}

static void rho_at_point(int N, const double* P, const double* phi, const double* gx, const double* gy,
                         const double* gz, double* r) {
    double rho = 0.0, grx = 0.0, gry = 0.0, grz = 0.0;
    for (int i = 0; i < N; ++i) {
        const double ai = phi[i];
        const double* Pi = P + (size_t)i * N;
        for (int j = 0; j < i; ++j) rho += 2.0 * Pi[j] * ai * phi[j];
        rho += Pi[i] * ai * ai;
        for (int j = 0; j < N; ++j) {
            grx += Pi[j] * ai * gx[j];
            gry += Pi[j] * ai * gy[j];
            grz += Pi[j] * ai * gz[j];
        }
    }
    r[0] = rho;
    r[1] = 2.0 * grx;
    r[2] = 2.0 * gry;
    r[3] = 2.0 * grz;
}

double oracle_calc_xc_uks(const oracle_aos* aos, int functional, int64_t n_pts, const double* grid,
                          const double* PA, const double* PB, double* FA, double* FB, double* rho_out,
                          double density_cutoff, int exact_gab, int n_threads) {
    const int N = aos->n_aos;
    int nt = 1;
#ifdef _OPENMP
    nt = (n_threads > 0) ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
#endif
    double* Fpriv = (double*)calloc((size_t)2 * nt * N * N, sizeof(double));
    double energy = 0.0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : energy)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* Fa = Fpriv + (size_t)2 * tid * N * N;
        double* Fb = Fa + (size_t)N * N;
        double* phi = (double*)malloc(sizeof(double) * 4 * N);
        double *gx = phi + N, *gy = gx + N, *gz = gy + N;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t g = 0; g < n_pts; ++g) {
            const double* p = grid + 4 * g;
            const double w = p[3];
            eval_point(aos, p, phi, gx, gy, gz);
            double ra[4], rb[4];
            rho_at_point(N, PA, phi, gx, gy, gz, ra);
            rho_at_point(N, PB, phi, gx, gy, gz, rb);
            if (rho_out) {
                for (int q = 0; q < 4; ++q) {
                    rho_out[8 * g + q] = ra[q];
                    rho_out[8 * g + 4 + q] = rb[q];
                }
            }
            if (ra[0] > density_cutoff || rb[0] > density_cutoff) {
                const double gaa = ra[1] * ra[1] + ra[2] * ra[2] + ra[3] * ra[3];
                const double gab = ra[1] * rb[1] + ra[2] * rb[2] + ra[3] * rb[3];
                const double gbb = rb[1] * rb[1] + rb[2] * rb[2] + rb[3] * rb[3];
                const double in[5] = {ra[0], rb[0], gaa, gab, gbb};
                double o[6];
                functional_point(functional, in, o);
                const double* oa = exact_gab ? rb : ra; /* gradient multiplying vgAB in F_A */
                const double* ob = exact_gab ? ra : rb;
                double c[3], d[3];
                for (int k = 0; k < 3; ++k) {
                    c[k] = w * (2.0 * o[3] * ra[1 + k] + o[4] * oa[1 + k]);
                    d[k] = w * (2.0 * o[5] * rb[1 + k] + o[4] * ob[1 + k]);
                }
                add_point_matrix(Fa, N, phi, gx, gy, gz, w * o[1], c[0], c[1], c[2]);
                add_point_matrix(Fb, N, phi, gx, gy, gz, w * o[2], d[0], d[1], d[2]);
                energy += w * o[0];
            }
        }
        free(phi);
    }
    for (int t = 0; t < nt; ++t) {
        const double* Fa = Fpriv + (size_t)2 * t * N * N;
        const double* Fb = Fa + (size_t)N * N;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j) {
                FA[(size_t)i * N + j] += Fa[(size_t)i * N + j];
                FB[(size_t)i * N + j] += Fb[(size_t)i * N + j];
            }
    }
    free(Fpriv);
    return energy;
}

/* ------------------------------------------------------------------------ */
/* Lebedev rules.  Each rule is a list of octahedral orbits (TlLebedevGrid.cpp */
/* :93-269 lists the six orbit shapes; the per-order parameters :271-1577).   */
/* The parameters are the published Lebedev-Laikov values, all 22 orders of   */
/* TlLebedevGrid.cpp:4-6 (lebedev_rules.inc, shared with the library).         */
/* The orbit is generated here by applying all 48                             */
/* signed permutations to a seed point and removing duplicates, so the point   */
/* order differs from the reference (only the summation order changes).        */
/* ------------------------------------------------------------------------ */
#include "lebedev_rules.inc"

typedef struct {
    int type;
    double a, b, w;
} leb_orbit;

static int orbit_points(const leb_orbit* o, double* xyz) {
    double s[3];
    switch (o->type) {
        case 1: s[0] = 1.0; s[1] = 0.0; s[2] = 0.0; break;
        case 2: s[0] = sqrt(0.5); s[1] = sqrt(0.5); s[2] = 0.0; break;
        case 3: s[0] = s[1] = s[2] = sqrt(1.0 / 3.0); break;
        case 4: s[0] = o->a; s[1] = o->a; s[2] = sqrt(1.0 - 2.0 * o->a * o->a); break;
        case 5: s[0] = o->a; s[1] = sqrt(1.0 - o->a * o->a); s[2] = 0.0; break;
        case 6: s[0] = o->a; s[1] = o->b; s[2] = sqrt(1.0 - o->a * o->a - o->b * o->b); break;
        default: return 0;
    }
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int n = 0;
    for (int pi = 0; pi < 6; ++pi)
        for (int sg = 0; sg < 8; ++sg) {
            double q[3];
            for (int d = 0; d < 3; ++d) {
                const double v = s[perm[pi][d]];
                q[d] = ((sg >> d) & 1) ? -v : v;
                if (q[d] == 0.0) q[d] = 0.0; /* fold -0 */
            }
            int dup = 0;
            for (int k = 0; k < n && !dup; ++k)
                dup = (xyz[3 * k] == q[0] && xyz[3 * k + 1] == q[1] && xyz[3 * k + 2] == q[2]);
            if (!dup) {
                xyz[3 * n] = q[0];
                xyz[3 * n + 1] = q[1];
                xyz[3 * n + 2] = q[2];
                ++n;
            }
        }
    return n;
}

int oracle_lebedev(int order, double* xyz, double* w) {
    const double(*t)[4] = NULL;
    int no = 0;
#define LEB_CASE(N, T) case N: t = T; no = (int)(sizeof(T) / sizeof(T[0])); break;
    switch (order) {
        LEBEDEV_RULE_LIST(LEB_CASE)
        default: return 0;
    }
#undef LEB_CASE
    int n = 0;
    double buf[48 * 3];
    for (int k = 0; k < no; ++k) {
        const leb_orbit o = {(int)t[k][0], t[k][1], t[k][2], t[k][3]};
        const int m = orbit_points(&o, buf);
        for (int i = 0; i < m; ++i) {
            xyz[3 * (n + i)] = buf[3 * i];
            xyz[3 * (n + i) + 1] = buf[3 * i + 1];
            xyz[3 * (n + i) + 2] = buf[3 * i + 2];
            w[n + i] = o.w;
        }
        n += m;
    }
    return (n == order) ? n : 0;
}

/* ------------------------------------------------------------------------ */
/* Euler-Maclaurin radial rule (DfGenerateGrid.cpp:1597-1610):                */
/*   r_i = R i^2/(Nr+1-i)^2,  w_i = 2 R^3 (Nr+1) i^5/(Nr+1-i)^7 * 4 pi.       */
/* The reference forms i^5 in 32-bit int arithmetic, which wraps for i >= 74; */
/* the wrap is reproduced with unsigned arithmetic (well defined here).       */
/* ------------------------------------------------------------------------ */
void oracle_radial_em(double R, int Nr, int i, double* r, double* w) {
    const double Nr1i = (double)(Nr + 1 - i);
    const double inv = 1.0 / Nr1i;
    *r = R * (double)(i * i) * inv * inv;
    const uint32_t u = (uint32_t)i;
    const int32_t i5 = (int32_t)(u * u * u * u * u);
    *w = 2.0 * R * R * R * (double)(Nr + 1) * (double)i5 * inv * inv * inv * inv *
         inv * inv * inv * 4.0 * M_PI;
}

/* Bragg-Slater radii in the units the reference uses them (TlPrdctbl.cpp:150-180) */
double oracle_bragg_slater(int Z) {
    static const double t[] = {0.00, 0.25, 0.50, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60,
                               0.50, 0.50, 1.80, 1.50, 1.25, 1.10, 1.00, 1.00, 1.00,
                               1.00};
    if (Z >= 1 && Z < (int)(sizeof(t) / sizeof(t[0]))) return t[Z];
    return 0.0;
}
/* getCovalentRadiiForBecke (DfGenerateGrid.cpp:1958-1968): H uses 0.35 */
static double becke_radius(int Z) { return (Z == 1) ? 0.35 : oracle_bragg_slater(Z); }

/* 3x3 symmetric eigen-decomposition by cyclic Jacobi rotations. */
static void jacobi3(double A[3][3], double V[3][3], double ev[3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) ev[i] = A[i][i];
}

void oracle_o_matrix(int n_atoms, const int* Z, const double* xyz, double* O) {
    double T[3] = {0, 0, 0}, sz = 0.0;
    for (int a = 0; a < n_atoms; ++a) {
        if (Z[a] <= 0) continue;
        for (int d = 0; d < 3; ++d) T[d] += Z[a] * xyz[3 * a + d];
        sz += Z[a];
    }
    for (int d = 0; d < 3; ++d) T[d] /= sz;
    double M[3][3] = {{0}};
    for (int a = 0; a < n_atoms; ++a) {
        if (Z[a] <= 0) continue;
        double r[3];
        for (int d = 0; d < 3; ++d) r[d] = xyz[3 * a + d] - T[d];
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] += Z[a] * (((i == j) ? r2 : 0.0) - r[i] * r[j]);
    }
    double V[3][3], ev[3];
    jacobi3(M, V, ev);
    int idx[3] = {0, 1, 2};
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[idx[j]] < ev[idx[i]]) {
                const int t = idx[i];
                idx[i] = idx[j];
                idx[j] = t;
            }
    for (int c = 0; c < 3; ++c) {
        const int k = idx[c];
        int big = 0;
        for (int r = 1; r < 3; ++r)
            if (fabs(V[r][k]) > fabs(V[big][k])) big = r;
        const double sgn = (V[big][k] < 0) ? -1.0 : 1.0;
        for (int r = 0; r < 3; ++r) O[3 * r + c] = sgn * V[r][k];
    }
}

/* Becke cell function product P_m(r) (Ps_uij, DfGenerateGrid.cpp:2043-2086)
 * and normalisation (calcMultiCenterWeight_Becke, :1811-1876). */
static double becke_f1(double x) { return 0.5 * x * (3.0 - x * x); }
static double becke_ps(int n_atoms, const int* Z, const double* xyz, int m,
                       const double* r, int size_adjust) {
    const double Rm = becke_radius(Z[m]);
    const double dxm = r[0] - xyz[3 * m], dym = r[1] - xyz[3 * m + 1], dzm = r[2] - xyz[3 * m + 2];
    const double rm = sqrt(dxm * dxm + dym * dym + dzm * dzm);
    double ps = 1.0;
    for (int n = 0; n < n_atoms; ++n) {
        if (n == m || Z[n] <= 0) continue;
        const double Rn = becke_radius(Z[n]);
        const double dxn = r[0] - xyz[3 * n], dyn = r[1] - xyz[3 * n + 1], dzn = r[2] - xyz[3 * n + 2];
        const double rn = sqrt(dxn * dxn + dyn * dyn + dzn * dzn);
        const double ex = xyz[3 * m] - xyz[3 * n], ey = xyz[3 * m + 1] - xyz[3 * n + 1],
                     ez = xyz[3 * m + 2] - xyz[3 * n + 2];
        const double inv_dist = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
        double u = (rm - rn) * inv_dist;
        if (size_adjust) {
            const double au = (Rm - Rn) / (Rm + Rn);
            double a = au / (au * au - 1.0);
            if (a < -0.50) a = -0.50;
            else if (a > 0.50) a = 0.50;
            u = u + a * (1.0 - u * u);
        }
        const double f3 = becke_f1(becke_f1(becke_f1(u)));
        ps *= 0.5 * (1.0 - f3);
    }
    return ps;
}
double oracle_becke_weight(int n_atoms, const int* Z, const double* xyz, int iatom,
                           const double* r, int size_adjust) {
    double total = 0.0, mine = 0.0;
    for (int m = 0; m < n_atoms; ++m) {
        const double p = becke_ps(n_atoms, Z, xyz, m, r, size_adjust);
        if (m == iatom) mine = p;
        total += p;
    }
    return mine / total;
}

/* Gauss-Chebyshev (second kind) radial rule with the Becke, Treutler-Ahlrichs
 * or Knowles-Kirby style mapping (getRadialAbscissaAndWeight_GaussChebyshev,
 * DfGenerateGrid.cpp:1612-1655); the weight carries 4 pi r^2. */
void oracle_radial_gc(double R, int Nr, int i, int mapping, double* r, double* w) {
    const double inv_ln2 = 1.0 / log(2.0);
    const double Nr1 = Nr + 1;
    const double t = (double)i * M_PI / Nr1;
    const double sn = sin(t), ss = sn * sn, cs = cos(t);
    const double x = (Nr1 - 2.0 * (double)i) / Nr1 + (2.0 / M_PI) * (1.0 + (2.0 / 3.0) * ss) * cs * sn;
    const double inv_x1 = 1.0 / (x - 1.0);
    double ri = 0.0, jac = 0.0;
    switch (mapping) {
        case 0: /* Becke: r = R (1 + x) / (1 - x) */
            ri = R * (1.0 + x) / (1.0 - x);
            jac = R * ((x + 1.0) * inv_x1 * inv_x1 - inv_x1);
            break;
        case 1: /* TA: r = R / ln2 ln(2 / (1 - x)) */
            ri = inv_ln2 * log(2.0 * (-inv_x1)) * R;
            jac = R * inv_ln2 / (1.0 - x);
            break;
        default: /* KK: the TA form without R */
            ri = inv_ln2 * log(2.0 * (-inv_x1));
            jac = inv_ln2 / (1.0 - x);
            break;
    }
    *r = ri;
    *w = (16.0 / (3.0 * Nr1)) * ss * ss * jac * ri * ri * 4.0 * M_PI;
}

/* SG-1 atomic radii (setCellPara's SG-1 list, DfGenerateGrid.cpp:409-429) */
double oracle_sg1_radius(int Z) {
    static const double t[] = {0.0,    1.0000, 0.5882, 3.0769, 2.0513, 1.5385, 1.2308,
                               1.0256, 0.8791, 0.7692, 0.6838, 4.0909, 3.1579, 2.5714,
                               2.1687, 1.8750, 1.6514, 1.4754, 1.3333};
    if (Z >= 1 && Z < (int)(sizeof(t) / sizeof(t[0]))) return t[Z];
    return 0.0;
}

static const int LEB_ORDERS[] = {6,   14,  26,  38,  50,  74,  86,  110,  146,  170,  194,
                                 230, 266, 302, 350, 434, 590, 770, 974, 1202, 1454, 1730};

/* the largest supported order <= n, or 6 (getNumOfPrunedAnglarPoints' lookup, :1677-1684) */
int oracle_lebedev_floor(int n) {
    int best = LEB_ORDERS[0];
    for (size_t k = 0; k < sizeof(LEB_ORDERS) / sizeof(LEB_ORDERS[0]); ++k)
        if (LEB_ORDERS[k] <= n) best = LEB_ORDERS[k];
    return best;
}

/* Angular points of a radial shell at r (DfGenerateGrid.cpp:1658-1713).
 * pruning 0: the rule of n_ang (floored to a supported order); 1: Murray-Handy-
 * Laming, n = min(int(5 r / R_Bragg * n_ang), n_ang) then floored; 2: SG-1
 * regions of r / R by element row: 6, 38, 86, 194, 86. */
int oracle_pruned_order(int pruning, int Z, double r, double R, int n_ang) {
    if (pruning == 2) {
        double al[4] = {0.0, 0.0, 0.0, 10000.0};
        if (Z == 1 || Z == 2) {
            al[0] = 0.25; al[1] = 0.5; al[2] = 1.0; al[3] = 4.5;
        } else if (Z >= 3 && Z <= 10) {
            al[0] = 0.1667; al[1] = 0.5; al[2] = 0.9; al[3] = 3.5;
        } else if (Z >= 11 && Z <= 18) {
            al[0] = 0.1; al[1] = 0.4; al[2] = 0.8; al[3] = 2.5;
        }
        const double q = r * (1.0 / R);
        if (q < al[1]) return (q < al[0]) ? 6 : 38;
        if (q < al[2]) return 86;
        return (q < al[3]) ? 194 : 86;
    }
    int n = n_ang;
    if (pruning == 1) {
        const double v = (5.0 * r / oracle_bragg_slater(Z)) * (double)n_ang;
        n = (v < (double)n_ang) ? (int)v : n_ang;
    }
    return oracle_lebedev_floor(n);
}

/* Stratmann-Scuseria partition weight (calcMultiCenterWeight_SS, :1878-1956):
 * every atom, dummy atoms included, has a cell; no size adjustment. */
double oracle_ss_weight(int n_atoms, const double* xyz, int iatom, const double* r) {
    double* rr = (double*)malloc(sizeof(double) * (size_t)n_atoms);
    for (int p = 0; p < n_atoms; ++p) {
        const double dx = r[0] - xyz[3 * p], dy = r[1] - xyz[3 * p + 1], dz = r[2] - xyz[3 * p + 2];
        rr[p] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    double total = 0.0, mine = 0.0;
    for (int m = 0; m < n_atoms; ++m) {
        double ps = 1.0;
        for (int n = 0; n < n_atoms; ++n) {
            if (n == m) continue;
            const double ex = xyz[3 * m] - xyz[3 * n], ey = xyz[3 * m + 1] - xyz[3 * n + 1],
                         ez = xyz[3 * m + 2] - xyz[3 * n + 2];
            const double u = (rr[m] - rr[n]) * (1.0 / sqrt(ex * ex + ey * ey + ez * ez));
            double g;
            if (u < -0.64) g = -1.0;
            else if (u > 0.64) g = 1.0;
            else {
                const double q = u / 0.64, q2 = q * q, q3 = q2 * q, q5 = q2 * q3;
                g = (1.0 / 16.0) * (35.0 * q - 35.0 * q3 + 21.0 * q5 - 5.0 * q5 * q2);
            }
            ps *= 0.5 * (1.0 - g);
        }
        if (m == iatom) mine = ps;
        total += ps;
    }
    free(rr);
    return mine / total;
}

int64_t oracle_generate_grid_ex(int n_atoms, const int* Z, const double* xyz, const oracle_grid_opts* opt,
                                const double* O, double* out_grid, int* out_atom, int n_threads) {
    /* the rules a shell can use: every supported order up to the largest */
    const int top = (opt->pruning == 2) ? 194 : oracle_lebedev_floor(opt->n_ang);
    double* leb[32] = {0};
    double* lebw[32] = {0};
    int nord = 0, ord[32];
    for (size_t k = 0; k < sizeof(LEB_ORDERS) / sizeof(LEB_ORDERS[0]) && LEB_ORDERS[k] <= top; ++k) {
        ord[nord] = LEB_ORDERS[k];
        leb[nord] = (double*)malloc(sizeof(double) * 4 * (size_t)LEB_ORDERS[k]);
        lebw[nord] = leb[nord] + 3 * (size_t)LEB_ORDERS[k];
        oracle_lebedev(LEB_ORDERS[k], leb[nord], lebw[nord]);
        ++nord;
    }
    /* each atom fills its own buffer (generateGrid_atom, :934-1091); the
     * blocks are then concatenated in atom order */
    const size_t cap = (size_t)opt->n_rad * (size_t)top;
    double** buf = (double**)calloc((size_t)n_atoms, sizeof(double*));
    int64_t* cnt = (int64_t*)calloc((size_t)n_atoms, sizeof(int64_t));
    (void)n_threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < n_atoms; ++a) {
        int64_t c = 0;
        if (Z[a] > 0) {
            double* B = (double*)malloc(sizeof(double) * 4 * cap);
            buf[a] = B;
            const double R = (opt->radii == 1) ? oracle_sg1_radius(Z[a]) : oracle_bragg_slater(Z[a]);
            for (int i = 1; i <= opt->n_rad; ++i) {
                double ri, wr;
                if (opt->radial == 1) oracle_radial_gc(R, opt->n_rad, i, opt->gc_mapping, &ri, &wr);
                else oracle_radial_em(R, opt->n_rad, i, &ri, &wr);
                if (ri > 30.0) continue;
                const int na = oracle_pruned_order(opt->pruning, Z[a], ri, R, opt->n_ang);
                int q = 0;
                while (ord[q] != na) ++q;
                for (int k = 0; k < na; ++k) {
                    double p[3];
                    for (int d = 0; d < 3; ++d) {
                        const double v = O[3 * d + 0] * leb[q][3 * k] + O[3 * d + 1] * leb[q][3 * k + 1] +
                                         O[3 * d + 2] * leb[q][3 * k + 2];
                        p[d] = v * ri + xyz[3 * a + d];
                    }
                    double w = lebw[q][k] * wr;
                    w *= (opt->partitioning == 1) ? oracle_ss_weight(n_atoms, xyz, a, p)
                                                  : oracle_becke_weight(n_atoms, Z, xyz, a, p, opt->size_adjust);
                    if (fabs(w) > opt->weight_cutoff) {
                        B[4 * c + 0] = p[0];
                        B[4 * c + 1] = p[1];
                        B[4 * c + 2] = p[2];
                        B[4 * c + 3] = w;
                        ++c;
                    }
                }
            }
        }
        cnt[a] = c;
    }
    int64_t total = 0;
    for (int a = 0; a < n_atoms; ++a) {
        if (out_grid && cnt[a] > 0) {
            memcpy(out_grid + 4 * total, buf[a], sizeof(double) * 4 * (size_t)cnt[a]);
            if (out_atom)
                for (int64_t k = 0; k < cnt[a]; ++k) out_atom[total + k] = a;
        }
        total += cnt[a];
        free(buf[a]);
    }
    free(buf);
    free(cnt);
    for (int q = 0; q < nord; ++q) free(leb[q]);
    return total;
}

int64_t oracle_generate_grid(int n_atoms, const int* Z, const double* xyz, int n_rad,
                             int n_ang, const double* O, double weight_cutoff,
                             double* out_grid, int* out_atom, int n_threads) {
    if (oracle_lebedev_floor(n_ang) != n_ang) return -1;
    const oracle_grid_opts opt = {n_rad, n_ang, 0, 1, 0, 0, 1, 0, weight_cutoff};
    return oracle_generate_grid_ex(n_atoms, Z, xyz, &opt, O, out_grid, out_atom, n_threads);
}

/* ------------------------------------------------------------------------ */
/* XC nuclear gradient, RKS GGA (energyGradient / energyGradient_part,        */
/* DfCalcGridX.cpp:1215-1241, :1243-1415).                                    */
/*                                                                            */
/* Functional term, as the reference forms it per grid point of atom A with   */
/* weight w, for every AO i (on atom B = atom(i)):                            */
/*   g_i = -2 w [ v_rho dphi_i X_i + sum_d c_d (d dphi_i/dr_d X_i             */
/*                + dphi_i Y_di) ],  X = P phi, Y_d = P dphi/dr_d,            */
/*   c_d = 2 v_gAA drho_d + v_gAB drho_d,                                     */
/* added to atom B and subtracted from atom A (the point moves with A), then  */
/* doubled for RKS (:1234).  AO second derivatives: getD2AOs_core (:771-863)  */
/* with getPrefactorForSecondDerivative (:237-385), written here through the  */
/* product rule on (angular polynomial) x exp(-a r^2).                        */
/*                                                                            */
/* Weight term: sum over B != A of w_single e_xc grad_B omega_A, added to B   */
/* and subtracted from A (:1386-1404).  The reference's JGP_t                 */
/* (DfGenerateGrid.cpp:2355-2372) only evaluates for |mu| > 1, which never   */
/* happens, so its weight term is zero; this oracle uses the exact            */
/* derivative of the size-adjusted Becke weight instead (pinned by finite     */
/* differences of E_xc on a grid that moves rigidly with the atoms).          */
/* ------------------------------------------------------------------------ */
static void angular_poly(int t, double x, double y, double z, double* f, double g[3], double h[6]) {
    const double s3 = 1.0 / sqrt(3.0);
    for (int k = 0; k < 3; ++k) g[k] = 0.0;
    for (int k = 0; k < 6; ++k) h[k] = 0.0;  /* xx xy xz yy yz zz */
    switch (t) {
        case 0: *f = 1.0; break;
        case 1: *f = x; g[0] = 1.0; break;
        case 2: *f = y; g[1] = 1.0; break;
        case 3: *f = z; g[2] = 1.0; break;
        case 4: *f = x * y; g[0] = y; g[1] = x; h[1] = 1.0; break;
        case 5: *f = z * x; g[0] = z; g[2] = x; h[2] = 1.0; break;
        case 6: *f = y * z; g[1] = z; g[2] = y; h[4] = 1.0; break;
        case 7: *f = 0.5 * (x * x - y * y); g[0] = x; g[1] = -y; h[0] = 1.0; h[3] = -1.0; break;
        case 8:
            *f = s3 * (z * z - 0.5 * (x * x + y * y));
            g[0] = -s3 * x; g[1] = -s3 * y; g[2] = 2.0 * s3 * z;
            h[0] = -s3; h[3] = -s3; h[5] = 2.0 * s3;
            break;
        default: *f = 0.0; break;
    }
}
/* phi, grad phi and the 6 second derivatives of every AO at point p */
static void eval_point_d2(const oracle_aos* a, const double* p, double* phi, double* g3, double* h6) {
    static const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int i = 0; i < a->n_aos; ++i) {
        const double r[3] = {p[0] - a->center[3 * i + 0], p[1] - a->center[3 * i + 1],
                             p[2] - a->center[3 * i + 2]};
        const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        double f, fg[3], fh[6];
        angular_poly(a->type[i], r[0], r[1], r[2], &f, fg, fh);
        double v = 0.0, gv[3] = {0, 0, 0}, hv[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < a->nprim[i]; ++k) {
            const double al = a->exps[i * ORACLE_MAX_PRIM + k];
            const double sh = al * d2;
            if (sh > TOOBIG) continue;
            const double e = a->coefs[i * ORACLE_MAX_PRIM + k] * exp(-sh);
            v += f * e;
            for (int d = 0; d < 3; ++d) gv[d] += (fg[d] - 2.0 * al * r[d] * f) * e;
            for (int q = 0; q < 6; ++q) {
                const int x = ia[q], y = ib[q];
                const double t = fh[q] - 2.0 * al * (r[x] * fg[y] + r[y] * fg[x]) +
                                 f * (4.0 * al * al * r[x] * r[y] - (x == y ? 2.0 * al : 0.0));
                hv[q] += t * e;
            }
        }
        phi[i] = v;
        for (int d = 0; d < 3; ++d) g3[3 * i + d] = gv[d];
        for (int q = 0; q < 6; ++q) h6[6 * i + q] = hv[q];
    }
}
void oracle_eval_ao_hessians(const oracle_aos* aos, int64_t n_pts, const double* xyz, double* h6) {
    const int N = aos->n_aos;
#pragma omp parallel
    {
        double* phi = (double*)malloc(sizeof(double) * 4 * (size_t)N);
        double* g3 = phi + N;
#pragma omp for schedule(dynamic, 64)
        for (int64_t p = 0; p < n_pts; ++p) eval_point_d2(aos, xyz + 3 * p, phi, g3, h6 + (size_t)p * 6 * N);
        free(phi);
    }
}

/* Becke step s(nu) = (1 - f(f(f(nu)))) / 2 and ds/dnu */
static void becke_s(double nu, double* s, double* ds) {
    const double f1 = becke_f1(nu), f2 = becke_f1(f1), f3 = becke_f1(f2);
    *s = 0.5 * (1.0 - f3);
    *ds = -0.5 * (1.5 * (1.0 - f2 * f2)) * (1.5 * (1.0 - f1 * f1)) * (1.5 * (1.0 - nu * nu));
}
static double becke_a(int Zi, int Zj) {
    const double Ri = becke_radius(Zi), Rj = becke_radius(Zj);
    const double au = (Ri - Rj) / (Ri + Rj);
    double a = au / (au * au - 1.0);
    if (a < -0.50) a = -0.50;
    else if (a > 0.50) a = 0.50;
    return a;
}
/* product of row i of s without column j */
static double prod_except(const double* s, int n, int i, int j) {
    double q = 1.0;
    for (int m = 0; m < n; ++m)
        if (m != j) q *= s[i * n + m];
    return q;
}
/* Stratmann-Scuseria step s(mu) = (1 - g(mu)) / 2 and ds/dmu (the g of
 * calcMultiCenterWeight_SS, DfGenerateGrid.cpp:1923-1938, a = 0.64) */
static void ss_s(double mu, double* s, double* ds) {
    if (mu < -0.64 || mu > 0.64) {
        *s = (mu < 0.0) ? 1.0 : 0.0;
        *ds = 0.0;
        return;
    }
    const double q = mu / 0.64, q2 = q * q, q3 = q2 * q, q5 = q2 * q3;
    *s = 0.5 * (1.0 - (1.0 / 16.0) * (35.0 * q - 35.0 * q3 + 21.0 * q5 - 5.0 * q5 * q2));
    *ds = -0.5 * (1.0 / 16.0) * (35.0 - 105.0 * q2 + 105.0 * q2 * q2 - 35.0 * q2 * q2 * q2) / 0.64;
}
/* grad_B omega_A at point r for every B (point held fixed); out: n_atoms x 3.
 * partition: 0 Becke with size adjustments, 1 Becke without, 2 SS.
 * Returns omega_A. */
static double becke_weight_gradient(int n_atoms, const int* Z, const double* xyz, int A, const double* r,
                                    double* s, double* dsm, double* Pv, double* out, int partition) {
    const int n = n_atoms;
    double* rr = Pv + n;
    for (int i = 0; i < n; ++i) {
        const double dx = r[0] - xyz[3 * i], dy = r[1] - xyz[3 * i + 1], dz = r[2] - xyz[3 * i + 2];
        rr[i] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    /* s_ij and ds_ij/dmu_ij */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            s[i * n + j] = 1.0;
            dsm[i * n + j] = 0.0;
            if (i == j || Z[i] <= 0 || Z[j] <= 0) continue;
            const double ex = xyz[3 * i] - xyz[3 * j], ey = xyz[3 * i + 1] - xyz[3 * j + 1],
                         ez = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double inv_dist = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
            const double mu = (rr[i] - rr[j]) * inv_dist;  /* as becke_ps forms it */
            double sv, ds;
            if (partition == 2) {
                ss_s(mu, &sv, &ds);
                s[i * n + j] = sv;
                dsm[i * n + j] = ds;
            } else {
                const double a = (partition == 0) ? becke_a(Z[i], Z[j]) : 0.0;
                becke_s(mu + a * (1.0 - mu * mu), &sv, &ds);
                s[i * n + j] = sv;
                dsm[i * n + j] = ds * (1.0 - 2.0 * a * mu);
            }
        }
    double Zs = 0.0;
    for (int i = 0; i < n; ++i) {
        double p = (Z[i] > 0) ? 1.0 : 0.0;
        for (int j = 0; j < n; ++j) p *= s[i * n + j];
        Pv[i] = p;
        Zs += p;
    }
    const double omega = Pv[A] / Zs;
    for (int B = 0; B < n; ++B) {
        double* o = out + 3 * B;
        o[0] = o[1] = o[2] = 0.0;
        if (B == A || Z[B] <= 0) continue;
        const double uB[3] = {(xyz[3 * B] - r[0]) / rr[B], (xyz[3 * B + 1] - r[1]) / rr[B],
                              (xyz[3 * B + 2] - r[2]) / rr[B]};
        double gPA[3] = {0, 0, 0}, gZ[3] = {0, 0, 0};
        for (int i = 0; i < n; ++i) {
            if (Z[i] <= 0) continue;
            if (i == B) {
                /* grad_B P_B = sum_j prod_{k != j} s_Bk  ds_Bj grad_B mu_Bj */
                for (int j = 0; j < n; ++j) {
                    if (j == B || Z[j] <= 0) continue;
                    const double R[3] = {xyz[3 * B] - xyz[3 * j], xyz[3 * B + 1] - xyz[3 * j + 1],
                                         xyz[3 * B + 2] - xyz[3 * j + 2]};
                    const double Rl = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                    const double mu = (rr[B] - rr[j]) / Rl;
                    const double c = prod_except(s, n, B, j) * dsm[B * n + j];
                    for (int d = 0; d < 3; ++d) gZ[d] += c * (uB[d] - mu * R[d] / Rl) / Rl;
                }
            } else {
                /* grad_B P_i = prod_{k != B} s_ik  ds_iB grad_B mu_iB */
                const double R[3] = {xyz[3 * B] - xyz[3 * i], xyz[3 * B + 1] - xyz[3 * i + 1],
                                     xyz[3 * B + 2] - xyz[3 * i + 2]};
                const double Rl = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
                const double mu = (rr[i] - rr[B]) / Rl;
                const double c = prod_except(s, n, i, B) * dsm[i * n + B];
                for (int d = 0; d < 3; ++d) {
                    const double gm = (-uB[d] - mu * R[d] / Rl) / Rl;
                    gZ[d] += c * gm;
                    if (i == A) gPA[d] += c * gm;
                }
            }
        }
        for (int d = 0; d < 3; ++d) o[d] = (gPA[d] - omega * gZ[d]) / Zs;
    }
    return omega;
}

double oracle_partition_weight_gradient(int n_atoms, const int* Z, const double* xyz, int A, const double* r,
                                        int partition, double* out) {
    const int n = n_atoms;
    double* s = (double*)malloc(sizeof(double) * (2 * (size_t)n * n + 2 * (size_t)n));
    const double om =
        becke_weight_gradient(n, Z, xyz, A, r, s, s + (size_t)n * n, s + 2 * (size_t)n * n, out, partition);
    free(s);
    return om;
}
double oracle_becke_weight_gradient(int n_atoms, const int* Z, const double* xyz, int A, const double* r,
                                    double* out) {
    return oracle_partition_weight_gradient(n_atoms, Z, xyz, A, r, 0, out);
}

double oracle_xc_gradient_rks(const oracle_aos* aos, const int* ao_atom, int functional, int64_t n_pts,
                              const double* grid, const int* pt_atom, int n_atoms, const int* Z,
                              const double* xyz, const double* PA, double density_cutoff,
                              double* grad_f, double* grad_w, int n_threads) {
    return oracle_xc_gradient_rks_ex(aos, ao_atom, functional, n_pts, grid, pt_atom, n_atoms, Z, xyz, PA,
                                     density_cutoff, 0, grad_f, grad_w, n_threads);
}
double oracle_xc_gradient_rks_ex(const oracle_aos* aos, const int* ao_atom, int functional, int64_t n_pts,
                                 const double* grid, const int* pt_atom, int n_atoms, const int* Z,
                                 const double* xyz, const double* PA, double density_cutoff, int partition,
                                 double* grad_f, double* grad_w, int n_threads) {
    const int N = aos->n_aos;
    const int na = n_atoms;
    int nt = 1;
#ifdef _OPENMP
    nt = (n_threads > 0) ? n_threads : omp_get_max_threads();
#else
    (void)n_threads;
#endif
    double* acc = (double*)calloc((size_t)nt * 6 * na, sizeof(double));
    double energy = 0.0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : energy)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* gf = acc + (size_t)tid * 6 * na;
        double* gw = gf + 3 * na;
        double* phi = (double*)malloc(sizeof(double) * 10 * (size_t)N);
        double* g3 = phi + N;
        double* h6 = g3 + 3 * (size_t)N;
        double* s = (double*)malloc(sizeof(double) * (2 * (size_t)na * na + 2 * (size_t)na + 3 * (size_t)na));
        double* dsm = s + (size_t)na * na;
        double* Pv = dsm + (size_t)na * na;
        double* gom = Pv + 2 * (size_t)na;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t g = 0; g < n_pts; ++g) {
            const double* p = grid + 4 * g;
            const double w = p[3];
            const int A = pt_atom[g];
            eval_point_d2(aos, p, phi, g3, h6);
            /* rho (getRhoAtGridPoint) and grad rho (getGradRhoAtGridPoint) */
            double rho = 0.0;
            for (int i = 0; i < N; ++i) {
                const double* Pi = PA + (size_t)i * N;
                for (int j = 0; j < i; ++j) rho += 2.0 * Pi[j] * phi[i] * phi[j];
                rho += Pi[i] * phi[i] * phi[i];
            }
            double gr[3] = {0, 0, 0};
            for (int i = 0; i < N; ++i) {
                const double* Pi = PA + (size_t)i * N;
                for (int j = 0; j < N; ++j)
                    for (int d = 0; d < 3; ++d) gr[d] += Pi[j] * phi[i] * g3[3 * j + d];
            }
            for (int d = 0; d < 3; ++d) gr[d] *= 2.0;
            if (!(rho > density_cutoff)) continue;
            const double gamma = gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2];
            const double in[5] = {rho, rho, gamma, gamma, gamma};
            double o[6];
            functional_point(functional, in, o);
            energy += w * o[0];
            const double vr = o[1];
            const double c[3] = {2.0 * o[3] * gr[0] + o[4] * gr[0], 2.0 * o[3] * gr[1] + o[4] * gr[1],
                                 2.0 * o[3] * gr[2] + o[4] * gr[2]};
            /* functional term */
            for (int i = 0; i < N; ++i) {
                double rao = 0.0, rd[3] = {0, 0, 0};
                const double* Pi = PA + (size_t)i * N;
                for (int j = 0; j < N; ++j) {
                    rao += Pi[j] * phi[j];
                    for (int d = 0; d < 3; ++d) rd[d] += Pi[j] * g3[3 * j + d];
                }
                const double* di = g3 + 3 * i;
                const double* hi = h6 + 6 * i;
                /* H[a][b] of AO i: xx xy xz yy yz zz */
                const double H[3][3] = {{hi[0], hi[1], hi[2]}, {hi[1], hi[3], hi[4]}, {hi[2], hi[4], hi[5]}};
                const int B = ao_atom[i];
                for (int a = 0; a < 3; ++a) {
                    double v = vr * di[a] * rao;
                    for (int d = 0; d < 3; ++d) v += c[d] * (H[a][d] * rao + di[a] * rd[d]);
                    v *= -2.0 * w;
                    gf[3 * B + a] += v;
                    gf[3 * A + a] -= v;
                }
            }
            /* weight term (a cell weight that rounds to 0 has ds/dnu = 0 there too) */
            const double omega = becke_weight_gradient(na, Z, xyz, A, p, s, dsm, Pv, gom, partition);
            if (!(omega > 0.0)) continue;
            const double ws = w / omega;
            for (int B = 0; B < na; ++B) {
                if (B == A) continue;
                for (int a = 0; a < 3; ++a) {
                    const double v = ws * o[0] * gom[3 * B + a];
                    gw[3 * B + a] += v;
                    gw[3 * A + a] -= v;
                }
            }
        }
        free(phi);
        free(s);
    }
    for (int k = 0; k < 3 * na; ++k) {
        double f = 0.0, wv = 0.0;
        for (int t = 0; t < nt; ++t) {
            f += acc[(size_t)t * 6 * na + k];
            wv += acc[(size_t)t * 6 * na + 3 * na + k];
        }
        grad_f[k] = 2.0 * f; /* rks (:1234) */
        grad_w[k] = wv;
    }
    free(acc);
    return energy;
}
