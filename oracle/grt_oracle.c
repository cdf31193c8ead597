/*
 * grt_oracle.c -- CPU FP64 restatement of GR-ray-tracer's per-photon hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, by the smoke check and by the
 * CPU-baseline leg of the benchmark, always as the checker, never as the thing
 * measured on the GPU.  Built with gcc -O2 -fno-fast-math and no automatic
 * FMA contraction, so every expression is evaluated in the order written,
 * which is the reference's order; the only fused operations are the explicit
 * fma() calls of the integrators' right-hand side and stage sums (one
 * rounding per product-plus-sum, as a contracting compiler builds the
 * reference).  Citations are file:line inside the reference's src/ directory.
 *
 * Parity status of each function (how it is pinned, see tests/test_oracle_*.py):
 *   metric core / RHS ............ closed forms at 50 digits (mpmath), g*ginv = I,
 *                                  null l, RHS == ginv p and -1/2 dginv pp
 *                                  (test_geodesic_observables.cpp:47-66)
 *   fused RHS .................... within a few ulp of the RHS above, and the
 *                                  same closed forms
 *   RK4 / RK45 ................... exact linear system (test_rk_integrators.cpp),
 *                                  convergence order + tolerance monotonicity
 *                                  (test_convergence.cpp), hand accept/reject cases,
 *                                  the Fehlberg stability polynomial and error constant
 *   controller root x^(1/4) ...... round-to-nearest of the exact root (mpmath, 200 bits)
 *   cell search .................. 3x2x4 nonuniform grid -> (2,1,3)
 *                                  (test_npy_loader.cpp:74-97) + brute force
 *   tetrad ....................... g(e_a,e_b) = eta_ab, e0 || u
 *   radiative transfer ........... the five closed-form cases of
 *                                  test_radiative_transfer.cpp
 *   scattering ................... statistical (mu, phi uniform; E_loc kept;
 *                                  null norm), Stokes factors, Thomson P
 *   photon generators ............ thermal: KS against the Planck energy spectrum and
 *                                  its first two moments; power law: KS against the
 *                                  analytic CDF (slopes 2.5, 1, 1+5e-13, -0.5); custom
 *   disk emission ................ test_photon_generation.cpp:11-98; packet-to-cell
 *                                  assignment against a brute-force expansion; positions
 *                                  through the Boyer-Lindquist radius; fluid-frame isotropy
 *   spherical -> Cartesian u, b .. central differences of the position map; g(u,u) = -1
 *   global loop stop rule ........ flat-space (M = 0) rays with closed-form termination
 *                                  steps, ten (threshold, interval, max_steps) cases
 *   camera / image products ...... flat-space (M = 0) straight rays hit the
 *                                  analytic screen point; test_output_stride.cpp
 *   Philox4x32-10 ................ Random123 known-answer vectors
 *   reference adjoint ............ central finite differences of orc_trace
 *                                  (tests/test_oracle_adjoint.py); the increments'
 *                                  derivatives against central differences
 *   RNG-driven branches vs the reference's XorShift64 pool: parity unpinned
 *   (statistical only).
 */
#include "grt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846  /* utils.hpp:16 */

enum { PH_DISK = 0, PH_BACKWARD = 1, PH_FORWARD = 2, PH_ARRIVED = 3, PH_REJECTED = 4 };

static double sqr(double x) { return x * x; }
static double cube(double x) { return x * x * x; }
static double quad(double x) { return x * x * x * x; }

/* ------------------------------------------------------------------------- */
/* Kerr-Schild metric core: metrics/kerr_schild_core.hpp                      */
/* ------------------------------------------------------------------------- */

/* kerr_schild_core.hpp:28-33 */
double orc_compute_r(double x, double y, double z, double a)
{
    const double rsq = sqr(x) + sqr(y) + sqr(z);
    const double A = rsq - sqr(a);
    const double S = sqrt(sqr(A) + 4.0 * sqr(a * z));
    return sqrt(0.5 * (A + S));
}

/* kerr_schild_core.hpp:36-46 */
void orc_compute_r_and_deriv(double x, double y, double z, double a, double *r, double dr[3])
{
    const double rsq = sqr(x) + sqr(y) + sqr(z);
    const double A = rsq - sqr(a);
    const double S = sqrt(sqr(A) + 4.0 * sqr(a * z));
    const double rr = sqrt(0.5 * (A + S));
    const double f = 1.0 + A / S;
    *r = rr;
    dr[0] = x * f / (2.0 * rr);
    dr[1] = y * f / (2.0 * rr);
    dr[2] = (z * f + 2.0 * sqr(a) * z / S) / (2.0 * rr);
}

/* kerr_schild_core.hpp:50-56 */
static void ks_l(double x, double y, double z, double r, double a, double l[4])
{
    const double B = sqr(r) + sqr(a);
    l[0] = 1.0;
    l[1] = (r * x + a * y) / B;
    l[2] = (r * y - a * x) / B;
    l[3] = z / r;
}

/* kerr_schild_core.hpp:59-77; dl is [k][mu] row-major */
static void ks_l_and_deriv(double x, double y, double z, double r, const double dr[3], double a,
                           double l[4], double dl[12])
{
    const double B = sqr(r) + sqr(a);
    l[0] = 1.0;
    l[1] = (r * x + a * y) / B;
    l[2] = (r * y - a * x) / B;
    l[3] = z / r;
    for (int k = 0; k < 3; ++k) {
        const double ex = (k == 0) ? 1.0 : 0.0;
        const double ey = (k == 1) ? 1.0 : 0.0;
        const double ez = (k == 2) ? 1.0 : 0.0;
        const double dB = 2.0 * r * dr[k];
        dl[4 * k + 0] = 0.0;
        dl[4 * k + 1] = ((dr[k] * x + r * ex) + a * ey - l[1] * dB) / B;
        dl[4 * k + 2] = ((dr[k] * y + r * ey) - a * ex - l[2] * dB) / B;
        dl[4 * k + 3] = (ez - l[3] * dr[k]) / r;
    }
}

/* kerr_schild_core.hpp:80-83 */
static double ks_H(double r, double z, double a, double m)
{
    const double D = quad(r) + sqr(a * z);
    return m * cube(r) / D;
}

/* kerr_schild_core.hpp:86-95 */
static double ks_H_and_deriv(double r, double z, const double dr[3], double a, double m, double dH[3])
{
    const double D = quad(r) + sqr(a * z);
    const double H = m * cube(r) / D;
    for (int k = 0; k < 3; ++k) {
        const double ez = (k == 2) ? 1.0 : 0.0;
        const double dD = 4.0 * cube(r) * dr[k] + 2.0 * sqr(a) * z * ez;
        dH[k] = m * (3.0 * sqr(r) * dr[k] * D - cube(r) * dD) / sqr(D);
    }
    return H;
}

/* kerr_schild_core.hpp:99-110 ; g row-major [mu][nu] */
void orc_compute_metric(const double X[4], double a, double m, double g[16])
{
    const double r = orc_compute_r(X[1], X[2], X[3], a);
    double l[4];
    ks_l(X[1], X[2], X[3], r, a, l);
    const double H = ks_H(r, X[3], a, m);
    static const double eta[4] = {-1.0, 1.0, 1.0, 1.0};
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu)
            g[4 * mu + nu] = 2.0 * H * l[mu] * l[nu] + ((mu == nu) ? eta[mu] : 0.0);
}

/* kerr_schild_core.hpp:114-126 */
void orc_compute_inverse_metric(const double X[4], double a, double m, double ginv[16])
{
    const double r = orc_compute_r(X[1], X[2], X[3], a);
    double l[4];
    ks_l(X[1], X[2], X[3], r, a, l);
    const double lu[4] = {-l[0], l[1], l[2], l[3]};
    const double H = ks_H(r, X[3], a, m);
    static const double eta[4] = {-1.0, 1.0, 1.0, 1.0};
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu)
            ginv[4 * mu + nu] = -2.0 * H * lu[mu] * lu[nu] + ((mu == nu) ? eta[mu] : 0.0);
}

/* kerr_schild_core.hpp:131-155 ; dginv[alpha][mu][nu] */
void orc_compute_inverse_metric_deriv(const double X[4], double a, double m, double dginv[64])
{
    double r, dr[3], l[4], dl[12], dH[3];
    orc_compute_r_and_deriv(X[1], X[2], X[3], a, &r, dr);
    ks_l_and_deriv(X[1], X[2], X[3], r, dr, a, l, dl);
    const double H = ks_H_and_deriv(r, X[3], dr, a, m, dH);
    const double lu[4] = {-l[0], l[1], l[2], l[3]};
    for (int i = 0; i < 16; ++i) dginv[i] = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double dlu[4] = {-dl[4 * k], dl[4 * k + 1], dl[4 * k + 2], dl[4 * k + 3]};
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu)
                dginv[16 * (k + 1) + 4 * mu + nu] =
                    -2.0 * (dH[k] * lu[mu] * lu[nu] + H * (dlu[mu] * lu[nu] + lu[mu] * dlu[nu]));
    }
}

/* kerr_schild_core.hpp:161-192 */
void orc_hamiltonian_rhs(const double X[4], const double p[4], double a, double m, double rhs[8])
{
    double r, dr[3], l[4], dl[12], dH[3];
    orc_compute_r_and_deriv(X[1], X[2], X[3], a, &r, dr);
    ks_l_and_deriv(X[1], X[2], X[3], r, dr, a, l, dl);
    const double H = ks_H_and_deriv(r, X[3], dr, a, m, dH);
    const double lu[4] = {-l[0], l[1], l[2], l[3]};
    double lp = 0.0;
    for (int mu = 0; mu < 4; ++mu) lp += lu[mu] * p[mu];
    rhs[0] = -p[0] - 2.0 * H * lu[0] * lp;
    for (int mu = 1; mu < 4; ++mu) rhs[mu] = p[mu] - 2.0 * H * lu[mu] * lp;
    rhs[4] = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double dlu[4] = {-dl[4 * k], dl[4 * k + 1], dl[4 * k + 2], dl[4 * k + 3]};
        double dlp = 0.0;
        for (int mu = 0; mu < 4; ++mu) dlp += dlu[mu] * p[mu];
        rhs[5 + k] = dH[k] * sqr(lp) + 2.0 * H * lp * dlp;
    }
}

/* The same right-hand side as orc_hamiltonian_rhs (kerr_schild_core.hpp:
 * 161-192 with :36-46, :59-77 and :86-95 written out for k = x, y, z), with
 * the rounding the integrators use: every product that the reference's
 * expressions add to a sum is fused with that addition (fma, one rounding,
 * as a contracting compiler builds the reference); quotients and roots are
 * the IEEE operations.  Zero Kronecker terms are dropped (adding or
 * multiplying by them is exact).  A test holds it within a few ulp of
 * orc_hamiltonian_rhs. */
void orc_hamiltonian_rhs_fused(const double s[8], double a, double m, double out[8])
{
    const double x = s[1], y = s[2], z = s[3];
    const double p0 = s[4], p1 = s[5], p2 = s[6], p3 = s[7];
    const double a2 = a * a;
    /* compute_r_and_deriv */
    const double A = fma(z, z, fma(y, y, x * x)) - a2;
    const double az = a * z;
    const double az2 = az * az;
    const double S = sqrt(fma(A, A, 4.0 * az2));
    const double r = sqrt(0.5 * (A + S));
    const double tr = 2.0 * r;
    const double f = 1.0 + A / S;
    const double a2z2 = 2.0 * a2 * z;
    const double drx = x * f / tr;
    const double dry = y * f / tr;
    const double drz = fma(z, f, a2z2 / S) / tr;
    /* compute_l_and_deriv */
    const double B = fma(r, r, a2);
    const double l1 = fma(r, x, a * y) / B;
    const double l2 = fma(r, y, -(a * x)) / B;
    const double l3 = z / r;
    const double dBx = tr * drx, dBy = tr * dry, dBz = tr * drz;
    const double dl_x1 = fma(-l1, dBx, fma(drx, x, r)) / B;
    const double dl_x2 = fma(-l2, dBx, fma(drx, y, -a)) / B;
    const double dl_x3 = -(l3 * drx) / r;
    const double dl_y1 = fma(-l1, dBy, fma(dry, x, a)) / B;
    const double dl_y2 = fma(-l2, dBy, fma(dry, y, r)) / B;
    const double dl_y3 = -(l3 * dry) / r;
    const double dl_z1 = fma(-l1, dBz, drz * x) / B;
    const double dl_z2 = fma(-l2, dBz, drz * y) / B;
    const double dl_z3 = fma(-l3, drz, 1.0) / r;
    /* compute_H_and_deriv */
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double D = fma(r3, r, az2);
    const double D2 = D * D;
    const double H = m * r3 / D;
    const double c3 = 3.0 * r2;
    const double c4 = 4.0 * r3;
    const double dH_x = m * fma(c3 * drx, D, -(r3 * (c4 * drx))) / D2;
    const double dH_y = m * fma(c3 * dry, D, -(r3 * (c4 * dry))) / D2;
    const double dH_z = m * fma(c3 * drz, D, -(r3 * fma(c4, drz, a2z2))) / D2;
    /* compute_hamiltonian_rhs; l^mu = (-1, l1, l2, l3) */
    const double lp = fma(l3, p3, fma(l2, p2, fma(l1, p1, -p0)));
    const double twoH = 2.0 * H;
    out[0] = fma(twoH, lp, -p0);
    out[1] = fma(-(twoH * l1), lp, p1);
    out[2] = fma(-(twoH * l2), lp, p2);
    out[3] = fma(-(twoH * l3), lp, p3);
    out[4] = 0.0;
    const double dlp_x = fma(dl_x3, p3, fma(dl_x2, p2, dl_x1 * p1));
    const double dlp_y = fma(dl_y3, p3, fma(dl_y2, p2, dl_y1 * p1));
    const double dlp_z = fma(dl_z3, p3, fma(dl_z2, p2, dl_z1 * p1));
    const double lp2 = lp * lp;
    const double hl = twoH * lp;
    out[5] = fma(hl, dlp_x, dH_x * lp2);
    out[6] = fma(hl, dlp_y, dH_y * lp2);
    out[7] = fma(hl, dlp_z, dH_z * lp2);
}

/* kerr_schild_core.hpp:201-227 */
int orc_null_covariant_momentum(const double X[4], const double K[3], double a, double m, double p[4])
{
    double g[16];
    orc_compute_metric(X, a, m, g);
    const double qa = g[0];
    const double qb = 2.0 * (g[1] * K[0] + g[2] * K[1] + g[3] * K[2]);
    double qc = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) qc += g[4 * (i + 1) + (j + 1)] * K[i] * K[j];
    const double disc = sqr(qb) - 4.0 * qa * qc;
    if (disc < 0.0) return 0;
    const double sq = sqrt(disc);
    const double r1 = (-qb + sq) / (2.0 * qa);
    const double r2 = (-qb - sq) / (2.0 * qa);
    const double K0 = (r1 > 0.0) ? r1 : r2;
    if (!(K0 > 0.0)) return 0;
    const double Kc[4] = {K0, K[0], K[1], K[2]};
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0.0;
        for (int nu = 0; nu < 4; ++nu) s += g[4 * mu + nu] * Kc[nu];
        p[mu] = s;
    }
    return 1;
}

/* cartesian_kerr_schild.hpp:11-22 */
double orc_inverse_norm(const double X[4], const double V[4], double a, double m)
{
    double gi[16];
    orc_compute_inverse_metric(X, a, m, gi);
    double n = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) n += gi[4 * mu + nu] * V[mu] * V[nu];
    return n;
}

/* analysis/geodesic_observables.hpp:15-35 ; out = (norm, E, Lz, carter) */
void orc_geodesic_observables(const double s[8], double a, double m, double out[4])
{
    const double X[4] = {s[0], s[1], s[2], s[3]};
    const double P[4] = {s[4], s[5], s[6], s[7]};
    const double x = X[1], y = X[2], z = X[3];
    const double r = orc_compute_r(x, y, z, a);
    const double st = sqrt(fmax(0.0, 1.0 - sqr(z / r)));
    const double ct = z / r;
    const double sts = fmax(1e-15, st);
    const double E = -P[0];
    const double Lz = x * P[2] - y * P[1];
    const double pth = ((x * P[1] + y * P[2]) * ct / sts) - r * st * P[3];
    out[0] = orc_inverse_norm(X, P, a, m);
    out[1] = E;
    out[2] = Lz;
    out[3] = pth * pth + sqr(ct) * (sqr(a) * sqr(E) + sqr(Lz) / sqr(sts));
}

/* ------------------------------------------------------------------------- */
/* Runge-Kutta: rk_integrators.hpp and integrate_cart_ks.hpp:143-227          */
/* ------------------------------------------------------------------------- */

/* Fehlberg 4(5) tableau, integrate_cart_ks.hpp:166-180 */
#define A21 (1.0 / 4.0)
#define A31 (3.0 / 32.0)
#define A32 (9.0 / 32.0)
#define A41 (1932.0 / 2197.0)
#define A42 (-7200.0 / 2197.0)
#define A43 (7296.0 / 2197.0)
#define A51 (439.0 / 216.0)
#define A52 (-8.0)
#define A53 (3680.0 / 513.0)
#define A54 (-845.0 / 4104.0)
#define A61 (-8.0 / 27.0)
#define A62 (2.0)
#define A63 (-3544.0 / 2565.0)
#define A64 (1859.0 / 4104.0)
#define A65 (-11.0 / 40.0)
#define B41 (25.0 / 216.0)
#define B43 (1408.0 / 2565.0)
#define B44 (2197.0 / 4104.0)
#define B45 (-1.0 / 5.0)
#define B51 (16.0 / 135.0)
#define B53 (6656.0 / 12825.0)
#define B54 (28561.0 / 56430.0)
#define B55 (-9.0 / 50.0)
#define B56 (2.0 / 55.0)

#define ORC_NMAX 16

/* pow(x, 0.25) correctly rounded (integrate_cart_ks.hpp:217 calls the
 * platform pow, whose last bit varies between libraries; glibc's is not
 * correctly rounded for ~0.1% of inputs).  sqrt(sqrt(x)) is within an ulp;
 * of it and its neighbours keep the one whose fourth power, evaluated with
 * exact fma products, is closest to x. */
static double quarter_root(double x)
{
    const double y = sqrt(sqrt(x));
    double best = y, bestd = INFINITY;
    for (int d = -1; d <= 1; ++d) {
        const double c = d < 0 ? nextafter(y, 0.0) : (d > 0 ? nextafter(y, INFINITY) : y);
        const double s = c * c;
        const double e1 = fma(c, c, -s);
        const double q = s * s;
        const double e2 = fma(s, s, -q);
        const double r = fabs((q - x) + (e2 + 2.0 * s * e1));
        if (r < bestd) { bestd = r; best = c; }
    }
    return best;
}

/* The controller's root, exported for its test. */
double orc_quarter_root(double x) { return quarter_root(x); }


/* Rounding model of the integrators.  0 (default): the fused model that
 * the tracing kernel reproduces bit for bit (explicit fma in the RHS and the
 * stage sums).  1: the reference's expressions rounded once per operation,
 * unfused.  Both are builds of the reference (a contracting compiler and a
 * non-contracting one); tests compare whole renders under the two. */
static int g_unfused = 0;

void orc_set_unfused_rounding(int unfused) { g_unfused = unfused ? 1 : 0; }

/* rk_integrators.hpp:12-22 */
void orc_rk4_generic(int n, double *state, double dt, orc_rhs_fn f, void *user)
{
    double k1[ORC_NMAX], k2[ORC_NMAX], k3[ORC_NMAX], k4[ORC_NMAX], tmp[ORC_NMAX];
    if (n < 1 || n > ORC_NMAX) return;
    if (g_unfused) {
        f(state, k1, user);
        for (int i = 0; i < n; ++i) tmp[i] = state[i] + 0.5 * dt * k1[i];
        f(tmp, k2, user);
        for (int i = 0; i < n; ++i) tmp[i] = state[i] + 0.5 * dt * k2[i];
        f(tmp, k3, user);
        for (int i = 0; i < n; ++i) tmp[i] = state[i] + dt * k3[i];
        f(tmp, k4, user);
        for (int i = 0; i < n; ++i)
            state[i] += (dt / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
        return;
    }
    f(state, k1, user);
    for (int i = 0; i < n; ++i) tmp[i] = fma(0.5 * dt, k1[i], state[i]);
    f(tmp, k2, user);
    for (int i = 0; i < n; ++i) tmp[i] = fma(0.5 * dt, k2[i], state[i]);
    f(tmp, k3, user);
    for (int i = 0; i < n; ++i) tmp[i] = fma(dt, k3[i], state[i]);
    f(tmp, k4, user);
    for (int i = 0; i < n; ++i)
        state[i] = fma(dt / 6.0, k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i], state[i]);
}

/* rk_integrators.hpp:35-106 (== integrate_cart_ks.hpp:160-227); returns accepted */
int orc_rk45_generic(int n, double *state, double *dt, const double *k1, orc_rhs_fn f, void *user,
                     double atol, double rtol, double min_s, double max_s, double safety)
{
    const double h = *dt;
    double k2[ORC_NMAX], k3[ORC_NMAX], k4[ORC_NMAX], k5[ORC_NMAX], k6[ORC_NMAX], tmp[ORC_NMAX];
    double x4[ORC_NMAX], x5[ORC_NMAX];
    if (n < 1 || n > ORC_NMAX) return 0;
    if (g_unfused) {
        for (int i = 0; i < n; ++i) tmp[i] = state[i] + h * A21 * k1[i];
        f(tmp, k2, user);
        for (int i = 0; i < n; ++i) tmp[i] = state[i] + h * (A31 * k1[i] + A32 * k2[i]);
        f(tmp, k3, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = state[i] + h * (A41 * k1[i] + A42 * k2[i] + A43 * k3[i]);
        f(tmp, k4, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = state[i] + h * (A51 * k1[i] + A52 * k2[i] + A53 * k3[i] + A54 * k4[i]);
        f(tmp, k5, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = state[i] +
                     h * (A61 * k1[i] + A62 * k2[i] + A63 * k3[i] + A64 * k4[i] + A65 * k5[i]);
        f(tmp, k6, user);
        for (int i = 0; i < n; ++i) {
            x4[i] = state[i] + h * (B41 * k1[i] + B43 * k3[i] + B44 * k4[i] + B45 * k5[i]);
            x5[i] = state[i] +
                    h * (B51 * k1[i] + B53 * k3[i] + B54 * k4[i] + B55 * k5[i] + B56 * k6[i]);
        }
    }
    /* each product joins its running sum in one fma (left to right, as
     * written), and so does the final state + h * sum */
    else {
        for (int i = 0; i < n; ++i) tmp[i] = fma(h * A21, k1[i], state[i]);
        f(tmp, k2, user);
        for (int i = 0; i < n; ++i) tmp[i] = fma(h, fma(A32, k2[i], A31 * k1[i]), state[i]);
        f(tmp, k3, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = fma(h, fma(A43, k3[i], fma(A42, k2[i], A41 * k1[i])), state[i]);
        f(tmp, k4, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = fma(h, fma(A54, k4[i], fma(A53, k3[i], fma(A52, k2[i], A51 * k1[i]))),
                         state[i]);
        f(tmp, k5, user);
        for (int i = 0; i < n; ++i)
            tmp[i] = fma(h, fma(A65, k5[i], fma(A64, k4[i], fma(A63, k3[i],
                                                                 fma(A62, k2[i], A61 * k1[i])))),
                         state[i]);
        f(tmp, k6, user);
        for (int i = 0; i < n; ++i) {
            x4[i] = fma(h, fma(B45, k5[i], fma(B44, k4[i], fma(B43, k3[i], B41 * k1[i]))),
                        state[i]);
            x5[i] = fma(h, fma(B56, k6[i], fma(B55, k5[i], fma(B54, k4[i], fma(B53, k3[i],
                                                                                 B51 * k1[i])))),
                        state[i]);
        }
    }
    double err = 0.0;
    int finite = 1;
    for (int i = 0; i < n; ++i) {
        const double big = fmax(fabs(x5[i]), fabs(x4[i]));
        const double sc = g_unfused ? atol + rtol * big : fma(rtol, big, atol);
        const double ratio = fabs(x5[i] - x4[i]) / sc;
        if (!isfinite(x5[i]) || !isfinite(ratio)) finite = 0;
        err = fmax(err, ratio);
    }
    if (!finite) {
        *dt = h * min_s;
        return 0;
    }
    if (err == 0.0) {
        *dt = h * max_s;
        for (int i = 0; i < n; ++i) state[i] = x5[i];
        return 1;
    }
    double scale = safety * quarter_root(1.0 / err);
    scale = fmin(max_s, fmax(min_s, scale));
    *dt = h * scale;
    if (err <= 1.0) {
        for (int i = 0; i < n; ++i) state[i] = x5[i];
        return 1;
    }
    return 0;
}

typedef struct { double a, m; } ks_user;

static void ks_rhs_cb(const double *in, double *out, void *user)
{
    const ks_user *u = (const ks_user *)user;
    if (g_unfused)
        orc_hamiltonian_rhs(in, in + 4, u->a, u->m, out);
    else
        orc_hamiltonian_rhs_fused(in, u->a, u->m, out);
}

/* integrate_cart_ks.hpp:143-155 */
void orc_rk4_step_ks(double state[8], double dt, double a, double m)
{
    ks_user u = {a, m};
    orc_rk4_generic(8, state, dt, ks_rhs_cb, &u);
}

/* integrate_cart_ks.hpp:160-227 */
int orc_rk45_step_ks(const orc_params *p, double state[8], double *dlambda, const double k1[8])
{
    ks_user u = {p->a_bh, p->m_bh};
    return orc_rk45_generic(8, state, dlambda, k1, ks_rhs_cb, &u, p->atol, p->rtol,
                            p->min_step_scale, p->max_step_scale, p->safety);
}

/* ------------------------------------------------------------------------- */
/* Cell search: radiative_transfer/scattering.hpp:41-180                      */
/* ------------------------------------------------------------------------- */

static double coord3(const double *a, const orc_grid *g, int i, int j, int k)
{
    return a[((size_t)i * g->nth + j) * g->nph + k];
}

/* scattering.hpp:41-45 */
static double periodic_angle_distance(double lhs, double rhs)
{
    const double d = fmod(fabs(lhs - rhs), 2.0 * ORC_PI);
    return fmin(d, 2.0 * ORC_PI - d);
}

/* scattering.hpp:47-65 */
static int nearest_radial(const orc_grid *g, double radius, int j, int k)
{
    if (g->nr == 0 || !(radius > 0.0)) return -1;
    int best = 0;
    const double first = coord3(g->r, g, 0, j, k);
    if (!(first > 0.0)) return -1;
    double bd = fabs(log(radius / first));
    for (int i = 1; i < g->nr; ++i) {
        const double c = coord3(g->r, g, i, j, k);
        if (!(c > 0.0)) continue;
        const double d = fabs(log(radius / c));
        if (d < bd) { best = i; bd = d; }
    }
    return best;
}

/* scattering.hpp:67-81 */
static int nearest_polar(const orc_grid *g, double polar, int i, int k)
{
    if (g->nth == 0) return -1;
    int best = 0;
    double bd = fabs(polar - coord3(g->theta, g, i, 0, k));
    for (int j = 1; j < g->nth; ++j) {
        const double d = fabs(polar - coord3(g->theta, g, i, j, k));
        if (d < bd) { best = j; bd = d; }
    }
    return best;
}

/* scattering.hpp:83-97 */
static int nearest_azimuth(const orc_grid *g, double az, int i, int j)
{
    if (g->nph == 0) return -1;
    int best = 0;
    double bd = periodic_angle_distance(az, coord3(g->phi, g, i, j, 0));
    for (int k = 1; k < g->nph; ++k) {
        const double d = periodic_angle_distance(az, coord3(g->phi, g, i, j, k));
        if (d < bd) { best = k; bd = d; }
    }
    return best;
}

/* scattering.hpp:99-117 */
static int within_radial(const orc_grid *g, double radius, int i, int j, int k)
{
    const double c = coord3(g->r, g, i, j, k);
    if (!(c > 0.0) || !(radius > 0.0)) return 0;
    const double cd = fabs(log(radius / c));
    if (g->nr == 1) return cd <= 1.0e-10;
    double span = 0.0;
    if (i > 0 && coord3(g->r, g, i - 1, j, k) > 0.0)
        span = fmax(span, fabs(log(c / coord3(g->r, g, i - 1, j, k))));
    if (i + 1 < g->nr && coord3(g->r, g, i + 1, j, k) > 0.0)
        span = fmax(span, fabs(log(c / coord3(g->r, g, i + 1, j, k))));
    return cd <= 0.500001 * span;
}

/* scattering.hpp:119-133 */
static int within_polar(const orc_grid *g, double polar, int i, int j, int k)
{
    const double c = coord3(g->theta, g, i, j, k);
    const double cd = fabs(polar - c);
    if (g->nth == 1) return cd <= 1.0e-10;
    double span = 0.0;
    if (j > 0) span = fmax(span, fabs(c - coord3(g->theta, g, i, j - 1, k)));
    if (j + 1 < g->nth) span = fmax(span, fabs(c - coord3(g->theta, g, i, j + 1, k)));
    return cd <= 0.500001 * span;
}

/* scattering.hpp:135-149 */
static int within_azimuth(const orc_grid *g, double az, int i, int j, int k)
{
    const double c = coord3(g->phi, g, i, j, k);
    const double cd = periodic_angle_distance(az, c);
    if (g->nph == 1) return cd <= 1.0e-10;
    double span = 0.0;
    if (k > 0) span = fmax(span, periodic_angle_distance(c, coord3(g->phi, g, i, j, k - 1)));
    if (k + 1 < g->nph) span = fmax(span, periodic_angle_distance(c, coord3(g->phi, g, i, j, k + 1)));
    return cd <= 0.500001 * span;
}

/* scattering.hpp:151-180 */
int orc_locate_cell(const orc_grid *g, double radius, double polar, double azimuth, int32_t ijk[3])
{
    if (g->nr == 0 || g->nth == 0 || g->nph == 0) return 0;
    int i = g->nr / 2, j = g->nth / 2, k = g->nph / 2;
    for (int it = 0; it < 3; ++it) {
        i = nearest_radial(g, radius, j, k);
        if (i < 0) return 0;
        j = nearest_polar(g, polar, i, k);
        if (j < 0) return 0;
        k = nearest_azimuth(g, azimuth, i, j);
        if (k < 0) return 0;
    }
    ijk[0] = i;
    ijk[1] = j;
    ijk[2] = k;
    return within_radial(g, radius, i, j, k) && within_polar(g, polar, i, j, k) &&
           within_azimuth(g, azimuth, i, j, k);
}

/* ------------------------------------------------------------------------- */
/* Fluid frame: radiative_transfer/fluid_frame.hpp                            */
/* ------------------------------------------------------------------------- */

/* fluid_frame.hpp:6-14 */
static double metric_inner(const double g[16], const double a[4], const double b[4])
{
    double s = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) s += g[4 * mu + nu] * a[mu] * b[nu];
    return s;
}

/* fluid_frame.hpp:16-27 */
static void spherical_position(double r, double th, double ph, double spin, double pos[4])
{
    const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    pos[0] = 0.0;
    pos[1] = (r * cp - spin * sp) * st;
    pos[2] = (r * sp + spin * cp) * st;
    pos[3] = r * ct;
}

/* fluid_frame.hpp:29-49 */
void orc_spherical_to_cartesian_vector(double r, double th, double ph, double spin,
                                       const double s[4], double c[4])
{
    const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    const double rx = cp * st, ry = sp * st, rz = ct;
    const double tx = (r * cp - spin * sp) * ct;
    const double ty = (r * sp + spin * cp) * ct;
    const double tz = -r * st;
    const double px = (-r * sp - spin * cp) * st;
    const double py = (r * cp - spin * sp) * st;
    c[0] = s[0];
    c[1] = rx * s[1] + tx * s[2] + px * s[3];
    c[2] = ry * s[1] + ty * s[2] + py * s[3];
    c[3] = rz * s[1] + tz * s[2];
}

/* fluid_frame.hpp:52-80 ; e row-major [a][mu] */
int orc_build_fluid_tetrad(const double X[4], const double u[4], double spin, double mass, double e[16])
{
    double g[16];
    orc_compute_metric(X, spin, mass, g);
    const double vn = metric_inner(g, u, u);
    if (!(vn < 0.0) || !isfinite(vn)) return 0;
    const double inv = 1.0 / sqrt(-vn);
    const double ts = u[0] >= 0.0 ? 1.0 : -1.0;
    for (int mu = 0; mu < 4; ++mu) e[mu] = ts * u[mu] * inv;
    for (int axis = 0; axis < 3; ++axis) {
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        c[axis + 1] = 1.0;
        const double tp = metric_inner(g, e, c);
        for (int mu = 0; mu < 4; ++mu) c[mu] += tp * e[mu];
        for (int prev = 0; prev < axis; ++prev) {
            const double pr = metric_inner(g, e + 4 * (prev + 1), c);
            for (int mu = 0; mu < 4; ++mu) c[mu] -= pr * e[4 * (prev + 1) + mu];
        }
        const double nn = metric_inner(g, c, c);
        if (!(nn > 1.0e-20) || !isfinite(nn)) return 0;
        const double is = 1.0 / sqrt(nn);
        for (int mu = 0; mu < 4; ++mu) e[4 * (axis + 1) + mu] = c[mu] * is;
    }
    return 1;
}

/* fluid_frame.hpp:83-104 */
static void fluid_direction_to_momentum(const double X[4], const double e[16], double dx, double dy,
                                        double dz, double spin, double mass, double p[4])
{
    double w[4], g[16];
    for (int mu = 0; mu < 4; ++mu) w[mu] = e[mu] + dx * e[4 + mu] + dy * e[8 + mu] + dz * e[12 + mu];
    orc_compute_metric(X, spin, mass, g);
    for (int mu = 0; mu < 4; ++mu) {
        p[mu] = 0.0;
        for (int nu = 0; nu < 4; ++nu) p[mu] += g[4 * mu + nu] * w[nu];
    }
}

/* scattering.hpp:185-251 */
static int sample_fluid_cell(const double s[8], const orc_grid *g, orc_fluid *f, double e[16],
                             size_t *cell_out)
{
    if (!g->available) return 0;
    const double radius = orc_compute_r(s[1], s[2], s[3], g->spin);
    if (!(radius > 0.0) || !isfinite(radius)) return 0;
    const double ct = fmax(-1.0, fmin(1.0, s[3] / radius));
    const double polar = acos(ct);
    double az = atan2(radius * s[2] - g->spin * s[1], radius * s[1] + g->spin * s[2]);
    if (az < 0.0) az += 2.0 * ORC_PI;
    int32_t ijk[3] = {-1, -1, -1};
    if (!orc_locate_cell(g, radius, polar, az, ijk)) return 0;
    const int i = ijk[0], j = ijk[1], k = ijk[2];
    const size_t cell = ((size_t)i * g->nth + j) * g->nph + k;
    *cell_out = cell;
    f->position_cm[0] = s[1] * g->length_scale;
    f->position_cm[1] = s[2] * g->length_scale;
    f->position_cm[2] = s[3] * g->length_scale;
    f->coordinate_time_s = s[0] * g->time_scale;
    f->density_g_cm3 = g->rho[cell] * g->density_scale;
    f->temperature_k = g->tgas[cell] * g->temperature_scale;
    if (!(f->density_g_cm3 > 0.0) || !isfinite(f->density_g_cm3) || !isfinite(f->temperature_k))
        return 0;
    double sv[4], sb[4], cb[4];
    for (int c = 0; c < 4; ++c) {
        sv[c] = g->vel[4 * cell + c];
        sb[c] = g->mag[4 * cell + c];
    }
    const double rc = g->r[cell], thc = g->theta[cell], phc = g->phi[cell];
    orc_spherical_to_cartesian_vector(rc, thc, phc, g->spin, sv, f->coordinate_four_velocity_cm_s);
    for (int c = 0; c < 4; ++c) f->coordinate_four_velocity_cm_s[c] *= g->velocity_scale;
    const double X[4] = {s[0], s[1], s[2], s[3]};
    if (!orc_build_fluid_tetrad(X, f->coordinate_four_velocity_cm_s, g->spin, g->mass, e)) return 0;
    orc_spherical_to_cartesian_vector(rc, thc, phc, g->spin, sb, cb);
    for (int c = 0; c < 4; ++c) cb[c] *= g->magnetic_scale;
    double gm[16];
    orc_compute_metric(X, g->spin, g->mass, gm);
    f->magnetic_gauss[0] = -metric_inner(gm, e, cb);
    for (int ax = 1; ax < 4; ++ax) f->magnetic_gauss[ax] = metric_inner(gm, e + 4 * ax, cb);
    return 1;
}

int orc_sample_fluid(const double s[8], const orc_grid *g, orc_fluid *f, double e[16])
{
    size_t cell;
    return sample_fluid_cell(s, g, f, e, &cell);
}

/* ------------------------------------------------------------------------- */
/* Radiative transfer: radiative_transfer.hpp:54-96                           */
/* ------------------------------------------------------------------------- */

/* radiative_transfer.hpp:62-67 */
double orc_optical_depth_increment(const orc_fluid *f, double nu, double dl, double coeff)
{
    (void)nu;
    return fmax(0.0, coeff * f->density_g_cm3 * fabs(dl));
}

/* radiative_transfer.hpp:70-75 */
double orc_emissivity_increment(const orc_fluid *f, double nu, double dl, double coeff)
{
    (void)nu;
    return fmax(0.0, coeff * f->density_g_cm3 * f->temperature_k * fabs(dl));
}

/* Partial derivatives of the two default increments with respect to the
 * density, the temperature and the coefficient (zero where fmax clamps). */
double orc_optical_depth_increment_grad(const orc_fluid *f, double nu, double dl, double coeff,
                                        double d[3])
{
    const double v = orc_optical_depth_increment(f, nu, dl, coeff);
    const int live = v > 0.0;
    d[0] = live ? coeff * fabs(dl) : 0.0;
    d[1] = 0.0;
    d[2] = live ? f->density_g_cm3 * fabs(dl) : 0.0;
    return v;
}

double orc_emissivity_increment_grad(const orc_fluid *f, double nu, double dl, double coeff,
                                     double d[3])
{
    const double v = orc_emissivity_increment(f, nu, dl, coeff);
    const int live = v > 0.0;
    d[0] = live ? coeff * f->temperature_k * fabs(dl) : 0.0;
    d[1] = live ? coeff * f->density_g_cm3 * fabs(dl) : 0.0;
    d[2] = live ? f->density_g_cm3 * f->temperature_k * fabs(dl) : 0.0;
    return v;
}

/* radiative_transfer.hpp:81-96 */
void orc_apply_radiative_transfer(double stokes[4], const orc_fluid *f, double nu, double dl,
                                  int absorb, int emit, double kabs, double kemit)
{
    if (absorb) {
        const double att = exp(-orc_optical_depth_increment(f, nu, dl, kabs));
        for (int c = 0; c < 4; ++c) stokes[c] *= att;
    }
    if (emit) stokes[0] += orc_emissivity_increment(f, nu, dl, kemit);
}

/* ------------------------------------------------------------------------- */
/* Counter-based RNG: Philox4x32-10 (Salmon et al., SC'11, "Parallel random   */
/* numbers: as easy as 1, 2, 3").  Replaces Kokkos::Random_XorShift64_Pool    */
/* (scattering.hpp:16-17), whose streams are not reproducible.                */
/* ------------------------------------------------------------------------- */

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One uniform double in [0, 1) per draw: counter = (draw, stream, id_lo, id_hi). */
double orc_rng_drand(orc_rng *rng)
{
    const uint32_t ctr[4] = {rng->draw, rng->stream, (uint32_t)rng->id, (uint32_t)(rng->id >> 32)};
    const uint32_t key[2] = {(uint32_t)rng->seed, (uint32_t)(rng->seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    rng->draw += 1u;
    const uint64_t m = ((uint64_t)(o[0] >> 5) << 26) | (uint64_t)(o[1] >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------- */
/* Scattering and photon generation                                           */
/* ------------------------------------------------------------------------- */

#define PROTON_MASS_CGS 1.67262192369e-24   /* scattering.hpp:256 */
#define SIGMA_THOMSON_CGS 6.6524587158e-25  /* scattering.hpp:257 */

/* scattering.hpp:259-313 */
int orc_maybe_scatter(double s[8], double stokes[4], int enabled, double albedo, double dl_abs,
                      double length_scale, const orc_fluid *f, const double e[16], double spin,
                      double mass, orc_rng *rng)
{
    if (!enabled) return 0;
    const double path = dl_abs * length_scale;
    const double ne = f->density_g_cm3 / PROTON_MASS_CGS;
    const double tau = SIGMA_THOMSON_CGS * ne * path;
    const double prob = 1.0 - exp(-tau);
    if (orc_rng_drand(rng) >= prob) return 0;
    double eloc = 0.0;
    for (int mu = 0; mu < 4; ++mu) eloc -= s[4 + mu] * e[mu];
    if (!(eloc > 0.0) || !isfinite(eloc)) return 0;
    const double mu = 2.0 * orc_rng_drand(rng) - 1.0;
    const double dphi = 2.0 * ORC_PI * orc_rng_drand(rng);
    const double tr = sqrt(fmax(0.0, 1.0 - mu * mu));
    const double X[4] = {s[0], s[1], s[2], s[3]};
    double pm[4];
    fluid_direction_to_momentum(X, e, tr * cos(dphi), tr * sin(dphi), mu, spin, mass, pm);
    for (int c = 0; c < 4; ++c) s[4 + c] = eloc * pm[c];
    stokes[0] *= albedo;
    stokes[1] *= 0.5 * albedo;
    stokes[2] *= 0.5 * albedo;
    stokes[3] *= 0.25 * albedo;
    return 1;
}

/* initialize_photons.hpp:17-19.  static_cast<int> of an out-of-range double is
 * undefined in C++; both this restatement and the HIP path saturate. */
int orc_superphotons_per_cell(double density_code, double normalization)
{
    const double v = density_code * normalization;
    if (!(v == v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)(-2147483647 - 1);
    return (int)v;
}

/* photon_generation.hpp:99-156 ; generator selection integrate_cart_ks.hpp:332-347 */
int orc_generate_photon(const orc_params *p, const orc_fluid *cell, orc_rng *rng, int packets,
                        double *nu, double *energy)
{
    if (p->generator == 0) {
        if (!(cell->temperature_k > 0.0)) return 0;
        const double inverse_zeta4 = 0.9239384029215902;
        const double choose = orc_rng_drand(rng);
        double cum = 0.0;
        int idx = 1;
        for (; idx < 64; ++idx) {
            const double n = (double)idx;
            cum += inverse_zeta4 / (n * n * n * n);
            if (choose <= cum) break;
        }
        const double d1 = orc_rng_drand(rng), d2 = orc_rng_drand(rng);
        const double d3 = orc_rng_drand(rng), d4 = orc_rng_drand(rng);
        const double prod = fmax(1.0e-300, d1 * d2 * d3 * d4);
        const double x = -log(prod) / (double)idx;
        const double boltzmann_over_planck = 2.0836619123327576e10;
        *nu = x * boltzmann_over_planck * cell->temperature_k;
        *energy = p->energy_per_cell_erg / (double)packets;
        return isfinite(*nu) && *nu > 0.0;
    } else if (p->generator == 1) {
        const double draw = orc_rng_drand(rng);
        if (fabs(p->power_law_slope - 1.0) < 1.0e-12) {
            *nu = p->nu_min_hz * pow(p->nu_max_hz / p->nu_min_hz, draw);
        } else {
            const double ex = 1.0 - p->power_law_slope;
            const double lo = pow(p->nu_min_hz, ex);
            const double hi = pow(p->nu_max_hz, ex);
            *nu = pow(lo + draw * (hi - lo), 1.0 / ex);
        }
        *energy = p->energy_per_cell_erg / (double)packets;
        return isfinite(*nu) && *nu > 0.0;
    }
    *nu = p->custom_frequency_hz;
    *energy = p->energy_per_cell_erg / (double)packets;
    return 1;
}

void orc_sample_generator(const orc_params *p, const orc_fluid *cell, int packets, int64_t n,
                          double *nu, double *energy)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        orc_rng rng = {(uint64_t)i, 0u, 1u, p->seed};
        double v = 0.0, e = 0.0;
        if (!orc_generate_photon(p, cell, &rng, packets, &v, &e)) v = 0.0;
        nu[i] = v;
        energy[i] = e;
    }
}

/* ------------------------------------------------------------------------- */
/* Camera: integrate_cart_ks.hpp:96-130                                       */
/* ------------------------------------------------------------------------- */

int orc_camera_plane_crossing(const orc_params *p, const double s[8], double *u, double *v)
{
    const double th = p->camera_theta, ph = p->camera_phi, d = p->camera_distance;
    const double n[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
    const double t[3] = {cos(th) * cos(ph), cos(th) * sin(ph), -sin(th)};
    const double f[3] = {-sin(ph), cos(ph), 0.0};
    const double plane = s[1] * n[0] + s[2] * n[1] + s[3] * n[2];
    if (plane < d) return 0;
    const double rel[3] = {s[1] - d * n[0], s[2] - d * n[1], s[3] - d * n[2]};
    *u = rel[0] * t[0] + rel[1] * t[1] + rel[2] * t[2];
    *v = rel[0] * f[0] + rel[1] * f[1] + rel[2] * f[2];
    const int inside = fabs(*u) <= 0.5 * p->camera_width && fabs(*v) <= 0.5 * p->camera_height;
    return inside ? 1 : -1;
}

double orc_camera_observer_energy(const orc_params *p, const double s[8])
{
    const double X[4] = {s[0], s[1], s[2], s[3]};
    double g[16];
    orc_compute_metric(X, p->a_bh, p->m_bh, g);
    if (!(g[0] < 0.0)) return 0.0;
    const double u0 = 1.0 / sqrt(-g[0]);
    return -s[4] * u0;
}

/* ------------------------------------------------------------------------- */
/* The per-photon functor: integrate_cart_ks.hpp:231-419                      */
/* ------------------------------------------------------------------------- */

/* Per-photon record of the events the adjoint needs (reference adjoint). */
typedef struct {
    int kind;              /* 0: radiative-transfer update, 1: I set regardless of its past */
    size_t cell;
    double I_before, taken, nu;
    orc_fluid fluid;
} tape_entry;

typedef struct {
    tape_entry *e;
    int64_t n, cap;
} photon_tape;

static void tape_push(photon_tape *t, int kind, size_t cell, double I_before, double taken,
                      double nu, const orc_fluid *f)
{
    if (!t) return;
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 16;
        t->e = (tape_entry *)realloc(t->e, (size_t)t->cap * sizeof(tape_entry));
    }
    tape_entry *e = &t->e[t->n++];
    e->kind = kind;
    e->cell = cell;
    e->I_before = I_before;
    e->taken = taken;
    e->nu = nu;
    if (f) e->fluid = *f; else memset(&e->fluid, 0, sizeof e->fluid);
}

static void step_photon(const orc_params *p, const orc_grid *g, orc_photons *ph, int64_t idx,
                        uint32_t *rng_draw, orc_stats *st, photon_tape *tape)
{
    if (ph->terminate[idx]) return;
    st->functor_calls++;
    double s[8] = {ph->x0[idx], ph->x1[idx], ph->x2[idx], ph->x3[idx],
                   ph->k0[idx], ph->k1[idx], ph->k2[idx], ph->k3[idx]};
    const int phase = ph->phase[idx];
    const double dist = orc_compute_r(s[1], s[2], s[3], p->a_bh);
    if (dist <= p->r_term_min) {
        ph->I[idx] = ph->Q[idx] = ph->U[idx] = ph->V[idx] = 0.0;
        tape_push(tape, 1, 0, 0.0, 0.0, 0.0, NULL);
        ph->terminate[idx] = 1;
        ph->phase[idx] = PH_REJECTED;
        return;
    }
    if (dist > p->r_term_max) {
        ph->terminate[idx] = 1;
        ph->phase[idx] = PH_REJECTED;
        return;
    }
    double taken = 0.0;
    if (p->integrator == 0) {
        taken = ph->dlambda[idx];
        orc_rk4_step_ks(s, taken, p->a_bh, p->m_bh);
        st->accepted_steps++;
    } else if (p->integrator == 1) {
        double k1[8];
        ks_user u = {p->a_bh, p->m_bh};
        ks_rhs_cb(s, k1, &u);
        double dl = ph->dlambda[idx];
        int acc = 0;
        for (int attempt = 0; attempt < 12; ++attempt) {
            taken = dl;
            acc = orc_rk45_step_ks(p, s, &dl, k1);
            if (acc) break;
            st->rejected_attempts++;
        }
        ph->dlambda[idx] = dl;
        if (!acc) {
            ph->terminate[idx] = 1;
            st->failed_steps++;
            return;
        }
        st->accepted_steps++;
    } else {
        ph->terminate[idx] = 1;
        return;
    }

    double stokes[4] = {ph->I[idx], ph->Q[idx], ph->U[idx], ph->V[idx]};
    orc_fluid fluid;
    double e[16];
    memset(&fluid, 0, sizeof fluid);
    const int image = (p->mode == 0);
    const int needs = p->scattering_enabled || p->absorption_enabled || p->emission_enabled ||
                      (image && phase == PH_BACKWARD);
    int sampled = 0;
    size_t cell = 0;
    if (needs) {
        st->fluid_samples++;
        sampled = sample_fluid_cell(s, g, &fluid, e, &cell);
        st->fluid_hits += sampled ? 1u : 0u;
    }
    orc_rng rng = {ph->id[idx], rng_draw ? rng_draw[idx] : 0u, p->rng_stream << 1, p->seed};

    if (image && phase == PH_BACKWARD && sampled) {
        double nu = 0.0, energy = 0.0;
        int generated = 0;
        const double dcode = fluid.density_g_cm3 / g->density_scale;
        const int packets = orc_superphotons_per_cell(dcode, p->superphoton_count_normalization);
        if (packets > 0) generated = orc_generate_photon(p, &fluid, &rng, packets, &nu, &energy);
        if (generated) {
            ph->frequency[idx] = nu;
            double eem = 0.0;
            for (int c = 0; c < 4; ++c) eem -= s[4 + c] * e[c];
            ph->emission_frame_energy[idx] = eem;
            stokes[0] = energy;
            stokes[1] = stokes[2] = stokes[3] = 0.0;
            ph->dlambda[idx] = -fabs(ph->dlambda[idx]);
            ph->phase[idx] = PH_FORWARD;
        } else {
            stokes[0] = stokes[1] = stokes[2] = stokes[3] = 0.0;
            ph->terminate[idx] = 1;
            ph->phase[idx] = PH_REJECTED;
        }
        tape_push(tape, 1, 0, 0.0, 0.0, 0.0, NULL);
    } else if (sampled && phase != PH_BACKWARD) {
        if (p->absorption_enabled || p->emission_enabled)
            tape_push(tape, 0, cell, stokes[0], taken, ph->frequency[idx], &fluid);
        if (p->absorption_enabled || p->emission_enabled)
            orc_apply_radiative_transfer(stokes, &fluid, ph->frequency[idx], taken,
                                         p->absorption_enabled, p->emission_enabled,
                                         p->absorption_coefficient, p->emission_coefficient);
        if (p->scattering_enabled)
            orc_maybe_scatter(s, stokes, 1, p->albedo, fabs(ph->dlambda[idx]), g->length_scale,
                              &fluid, e, p->a_bh, p->m_bh, &rng);
    }
    if (rng_draw) rng_draw[idx] = rng.draw;

    if (image && phase == PH_FORWARD) {
        double au = 0.0, av = 0.0;
        const int crossing = orc_camera_plane_crossing(p, s, &au, &av);
        if (crossing > 0) {
            const double eobs = orc_camera_observer_energy(p, s);
            const double eem = ph->emission_frame_energy[idx];
            if (eobs > 0.0 && eem > 0.0 && isfinite(eobs) && isfinite(eem)) {
                ph->frequency[idx] *= eobs / eem;
            } else {
                stokes[0] = stokes[1] = stokes[2] = stokes[3] = 0.0;
                tape_push(tape, 1, 0, 0.0, 0.0, 0.0, NULL);
            }
            ph->theta_disp[idx] = au;
            ph->phi_disp[idx] = av;
            ph->phase[idx] = PH_ARRIVED;
            ph->terminate[idx] = 1;
        } else if (crossing < 0) {
            stokes[0] = stokes[1] = stokes[2] = stokes[3] = 0.0;
            tape_push(tape, 1, 0, 0.0, 0.0, 0.0, NULL);
            ph->phase[idx] = PH_REJECTED;
            ph->terminate[idx] = 1;
        }
    }
    ph->x0[idx] = s[0]; ph->x1[idx] = s[1]; ph->x2[idx] = s[2]; ph->x3[idx] = s[3];
    ph->k0[idx] = s[4]; ph->k1[idx] = s[5]; ph->k2[idx] = s[6]; ph->k3[idx] = s[7];
    ph->I[idx] = stokes[0]; ph->Q[idx] = stokes[1]; ph->U[idx] = stokes[2]; ph->V[idx] = stokes[3];
}

void orc_step_photon(const orc_params *p, const orc_grid *g, orc_photons *ph, int64_t idx,
                     uint32_t *rng_draw, orc_stats *st)
{
    step_photon(p, g, ph, idx, rng_draw, st, NULL);
}

/* Host loop, integrate_cart_ks.hpp:423-535 (single rank, no output files):
 * one functor pass over every photon per global step; the terminated
 * fraction is recounted only on output steps (step % interval == 0 or the last
 * step) and the loop stops once it reaches termination_percent.  Global steps
 * on which no photon is live are skipped without changing any state. */
static void trace_impl(const orc_params *p_in, const orc_grid *g, orc_photons *ph, int64_t n,
                       int32_t max_steps, double termination_percent, int64_t output_interval,
                       int32_t nthreads, orc_stats *st, photon_tape *tapes)
{
    orc_params p = *p_in;
    const int avail = g && g->available && g->nr > 0 && g->nth > 0 && g->nph > 0;
    p.scattering_enabled = p.scattering_enabled && avail;       /* :446-447 */
    p.absorption_enabled = p.absorption_enabled && avail;       /* :448-450 */
    p.emission_enabled = p.emission_enabled && avail;           /* :451-452 */
    orc_grid gl;
    if (g) gl = *g; else memset(&gl, 0, sizeof gl);
    gl.available = avail;
    uint32_t *draws = (uint32_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
    memset(st, 0, sizeof *st);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double fraction = 0.0;
    int64_t live = 0;
    for (int64_t i = 0; i < n; ++i) live += ph->terminate[i] ? 0 : 1;
    for (int32_t step = 0; step < max_steps; ++step) {
        st->global_steps++;
        if (live > 0) {
            uint64_t calls = 0, acc = 0, rej = 0, fail = 0, samp = 0, hits = 0;
            int64_t nlive = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : calls, acc, rej, fail, samp, hits, nlive)
            for (int64_t i = 0; i < n; ++i) {
                orc_stats ls;
                memset(&ls, 0, sizeof ls);
                step_photon(&p, &gl, ph, i, draws, &ls, tapes ? &tapes[i] : NULL);
                calls += ls.functor_calls; acc += ls.accepted_steps;
                rej += ls.rejected_attempts; fail += ls.failed_steps; samp += ls.fluid_samples;
                hits += ls.fluid_hits;
                nlive += ph->terminate[i] ? 0 : 1;
            }
            st->functor_calls += calls; st->accepted_steps += acc; st->rejected_attempts += rej;
            st->failed_steps += fail; st->fluid_samples += samp; st->fluid_hits += hits;
            live = nlive;
        }
        if (output_interval <= 0 || step % output_interval == 0 || step == max_steps - 1) {
            fraction = (double)(n - live) / (double)(n > 0 ? n : 1);
        }
        if (fraction >= termination_percent) break;
    }
    st->terminated = (uint64_t)(n - live);
    free(draws);
}

void orc_trace(const orc_params *p_in, const orc_grid *g, orc_photons *ph, int64_t n,
               int32_t max_steps, double termination_percent, int64_t output_interval,
               int32_t nthreads, orc_stats *st)
{
    trace_impl(p_in, g, ph, n, max_steps, termination_percent, output_interval, nthreads, st,
               NULL);
}

/* Reference gradient of L(I_final) for the differentiable form: traces the
 * photons like orc_trace while recording every radiative-transfer update and
 * every event that sets Stokes I regardless of its past, then walks each
 * photon's record backwards through I <- a I + b, a = exp(-dtau), b = dI:
 *   dL/dI_before = a dL/dI_after,
 *   dL/dtheta   += dL/dI_after (-a I_before ddtau/dtheta + d(dI)/dtheta)
 * for theta = the sampled cell's density and temperature (chained to code
 * units through the unit scales) and the two coefficients.  The emission
 * energy E/int(rho norm) is piecewise constant in rho and is held fixed.
 * grad_rho and grad_T (one per cell) and grad_coeff (absorption, emission)
 * are overwritten.  No checkpointing: the record holds every update. */
void orc_trace_adjoint(const orc_params *p_in, const orc_grid *g, orc_photons *ph, int64_t n,
                       int32_t max_steps, double termination_percent, int64_t output_interval,
                       int32_t nthreads, const double *dL_dI, double *grad_rho, double *grad_T,
                       double grad_coeff[2], orc_stats *st)
{
    photon_tape *tapes = (photon_tape *)calloc((size_t)(n > 0 ? n : 1), sizeof(photon_tape));
    trace_impl(p_in, g, ph, n, max_steps, termination_percent, output_interval, nthreads, st,
               tapes);
    const size_t G = (size_t)g->nr * g->nth * g->nph;
    memset(grad_rho, 0, G * sizeof(double));
    memset(grad_T, 0, G * sizeof(double));
    grad_coeff[0] = grad_coeff[1] = 0.0;
    const int absorb = p_in->absorption_enabled, emit = p_in->emission_enabled;
    for (int64_t i = 0; i < n; ++i) {
        double lam = dL_dI[i];
        for (int64_t q = tapes[i].n - 1; q >= 0 && lam != 0.0; --q) {
            const tape_entry *e = &tapes[i].e[q];
            if (e->kind == 1) break;
            double a = 1.0, dt[3] = {0.0, 0.0, 0.0}, di[3] = {0.0, 0.0, 0.0};
            if (absorb)
                a = exp(-orc_optical_depth_increment_grad(&e->fluid, e->nu, e->taken,
                                                          p_in->absorption_coefficient, dt));
            if (emit)
                orc_emissivity_increment_grad(&e->fluid, e->nu, e->taken,
                                              p_in->emission_coefficient, di);
            const double aI = a * e->I_before;
            grad_rho[e->cell] += lam * (di[0] - aI * dt[0]) * g->density_scale;
            grad_T[e->cell] += lam * (di[1] - aI * dt[1]) * g->temperature_scale;
            grad_coeff[0] += lam * (-aI * dt[2]);
            grad_coeff[1] += lam * di[2];
            lam *= a;
        }
        free(tapes[i].e);
    }
    free(tapes);
}

/* ------------------------------------------------------------------------- */
/* Initialisation: initialize_photons.hpp                                     */
/* ------------------------------------------------------------------------- */

/* initialize_photons.hpp:82-170, one photon with an injected screen (u, v). */
void orc_init_image_photon(const orc_params *p, double su, double sv, double initial_step,
                           double s[8], double *dlambda)
{
    const double d = p->camera_distance, th = p->camera_theta, ph = p->camera_phi;
    const double cx = d * sin(th) * cos(ph);
    const double cy = d * sin(th) * sin(ph);
    const double cz = d * cos(th);
    const double t[3] = {cos(th) * cos(ph), cos(th) * sin(ph), -sin(th)};
    const double f[3] = {-sin(ph), cos(ph), 0.0};
    const double px = cx + su * t[0] + sv * f[0];
    const double py = cy + su * t[1] + sv * f[1];
    const double pz = cz + su * t[2] + sv * f[2];
    double kx = -cx, ky = -cy, kz = -cz;
    const double kn = sqrt(kx * kx + ky * ky + kz * kz);
    kx /= kn; ky /= kn; kz /= kn;
    const double X[4] = {0.0, px, py, pz};
    const double K[3] = {kx, ky, kz};
    double pc[4] = {0.0, 0.0, 0.0, 0.0};
    orc_null_covariant_momentum(X, K, p->a_bh, p->m_bh, pc);
    s[0] = 0.0; s[1] = px; s[2] = py; s[3] = pz;
    s[4] = pc[0]; s[5] = pc[1]; s[6] = pc[2]; s[7] = pc[3];
    *dlambda = initial_step * d;
}

void orc_init_image(const orc_params *p, const double *su, const double *sv, int64_t n,
                    uint64_t id_offset, double initial_step, orc_photons *ph)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s[8], dl;
        orc_init_image_photon(p, su[i], sv[i], initial_step, s, &dl);
        ph->id[i] = id_offset + (uint64_t)i;
        ph->frequency[i] = 0.0;
        ph->emission_frame_energy[i] = 1.0;
        ph->x0[i] = s[0]; ph->x1[i] = s[1]; ph->x2[i] = s[2]; ph->x3[i] = s[3];
        ph->k0[i] = s[4]; ph->k1[i] = s[5]; ph->k2[i] = s[6]; ph->k3[i] = s[7];
        ph->I[i] = ph->Q[i] = ph->U[i] = ph->V[i] = 0.0;
        ph->dlambda[i] = dl;
        ph->terminate[i] = 0;
        ph->phase[i] = PH_BACKWARD;
        ph->theta_disp[i] = su[i];
        ph->phi_disp[i] = sv[i];
    }
}

/* initialize_photons.hpp:32-62 (prefix sum) and :173-327 (emission).  The
 * RNG stream for initialisation is (rng_stream << 1) | 1. */
void orc_init_disk(const orc_params *p, const orc_grid *g, int64_t n, uint64_t id_offset,
                   double initial_step, orc_photons *ph, uint64_t *total_out)
{
    const size_t ncell = (size_t)g->nr * g->nth * g->nph;
    uint64_t *off = (uint64_t *)malloc((ncell + 1) * sizeof(uint64_t));
    uint64_t run = 0;
    for (size_t c = 0; c < ncell; ++c) {
        off[c] = run;
        const int cnt = orc_superphotons_per_cell(g->rho[c], p->superphoton_count_normalization);
        run += (uint64_t)(cnt > 0 ? cnt : 0);
    }
    off[ncell] = run;
    if (total_out) *total_out = run;
    if (!ph || n <= 0) { free(off); return; }
    const double spin = g->spin, mass = g->mass;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        /* initialize_photons.hpp:38-51: largest cell with off[cell] <= i */
        size_t lo = 0, hi = ncell;
        while (lo + 1 < hi) {
            const size_t mid = lo + (hi - lo) / 2;
            if (off[mid] <= (uint64_t)i) lo = mid; else hi = mid;
        }
        const size_t cell = lo;
        const int packets = (int)(off[cell + 1] - off[cell]);
        const double rc = g->r[cell], thc = g->theta[cell], phc = g->phi[cell];
        double pos[4];
        spherical_position(rc, thc, phc, spin, pos);
        orc_fluid fl;
        memset(&fl, 0, sizeof fl);
        for (int c = 0; c < 3; ++c) fl.position_cm[c] = pos[c + 1] * g->length_scale;
        fl.coordinate_time_s = pos[0] * g->time_scale;
        fl.density_g_cm3 = g->rho[cell] * g->density_scale;
        fl.temperature_k = g->tgas[cell] * g->temperature_scale;
        double sv[4], sb[4], cb[4], e[16];
        for (int c = 0; c < 4; ++c) { sv[c] = g->vel[4 * cell + c]; sb[c] = g->mag[4 * cell + c]; }
        orc_spherical_to_cartesian_vector(rc, thc, phc, spin, sv, fl.coordinate_four_velocity_cm_s);
        for (int c = 0; c < 4; ++c) fl.coordinate_four_velocity_cm_s[c] *= g->velocity_scale;
        orc_spherical_to_cartesian_vector(rc, thc, phc, spin, sb, cb);
        for (int c = 0; c < 4; ++c) cb[c] *= g->magnetic_scale;
        const int valid = orc_build_fluid_tetrad(pos, fl.coordinate_four_velocity_cm_s, spin, mass, e);
        if (valid) {
            double gm[16];
            orc_compute_metric(pos, spin, mass, gm);
            fl.magnetic_gauss[0] = -metric_inner(gm, e, cb);
            for (int ax = 1; ax < 4; ++ax) fl.magnetic_gauss[ax] = metric_inner(gm, e + 4 * ax, cb);
        }
        orc_rng rng = {id_offset + (uint64_t)i, 0u, (p->rng_stream << 1) | 1u, p->seed};
        double nu = 0.0, energy = 0.0;
        int gen = 0;
        if (valid) gen = orc_generate_photon(p, &fl, &rng, packets, &nu, &energy);
        const double mu = 2.0 * orc_rng_drand(&rng) - 1.0;
        const double dphi = 2.0 * ORC_PI * orc_rng_drand(&rng);
        const double tr = sqrt(fmax(0.0, 1.0 - mu * mu));
        double pc[4] = {0.0, 0.0, 0.0, 0.0};
        if (gen) fluid_direction_to_momentum(pos, e, tr * cos(dphi), tr * sin(dphi), mu, spin, mass, pc);
        ph->id[i] = id_offset + (uint64_t)i;
        ph->frequency[i] = nu;
        double eem = 0.0;
        if (gen) for (int c = 0; c < 4; ++c) eem -= pc[c] * e[c];
        ph->emission_frame_energy[i] = eem;
        ph->x0[i] = pos[0]; ph->x1[i] = pos[1]; ph->x2[i] = pos[2]; ph->x3[i] = pos[3];
        ph->k0[i] = pc[0]; ph->k1[i] = pc[1]; ph->k2[i] = pc[2]; ph->k3[i] = pc[3];
        ph->I[i] = gen ? energy : 0.0;
        ph->Q[i] = ph->U[i] = ph->V[i] = 0.0;
        ph->dlambda[i] = initial_step;
        ph->terminate[i] = !gen;
        ph->phase[i] = gen ? PH_DISK : PH_REJECTED;
        ph->theta_disp[i] = 0.0;
        ph->phi_disp[i] = 0.0;
    }
    free(off);
}

/* ------------------------------------------------------------------------- */
/* Observation products: radiative_transfer/observation.hpp:22-84             */
/* ------------------------------------------------------------------------- */

int64_t orc_observation_bin(double value, double lo, double hi, int64_t bins)
{
    if (bins <= 1 || !(hi > lo)) return 0;
    const double fr = (value - lo) / (hi - lo);
    const double sc = fmax(0.0, fmin(1.0, fr));
    return (int64_t)(sc * (double)(bins - 1));
}

void orc_image_products(const orc_photons *ph, int64_t n, int64_t nx, int64_t ny,
                        int64_t nb, double smin, double smax, double w, double h,
                        double *image, double *spec_nu, double *spec_I, uint64_t *arrived)
{
    const int64_t np = nx * ny;
    for (int64_t i = 0; i < 4 * np; ++i) image[i] = 0.0;
    for (int64_t b = 0; b < nb; ++b) spec_I[b] = 0.0;
    const double lmin = log10(smin), lmax = log10(smax);
    for (int64_t b = 0; b < nb; ++b) {
        const double fr = nb > 1 ? (double)b / (double)(nb - 1) : 0.0;
        spec_nu[b] = pow(10.0, lmin + fr * (lmax - lmin));
    }
    uint64_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (ph->phase[i] != PH_ARRIVED) continue;
        ++count;
        const int64_t px = orc_observation_bin(ph->theta_disp[i], -0.5 * w, 0.5 * w, nx);
        const int64_t py = orc_observation_bin(ph->phi_disp[i], -0.5 * h, 0.5 * h, ny);
        const int64_t pix = py * nx + px;
        image[pix] += ph->I[i];
        image[np + pix] += ph->Q[i];
        image[2 * np + pix] += ph->U[i];
        image[3 * np + pix] += ph->V[i];
        const double nu = ph->frequency[i];
        if (nu >= smin && nu <= smax) {
            const int64_t b = orc_observation_bin(log10(nu), lmin, lmax, nb);
            spec_I[b] += ph->I[i];
        }
    }
    *arrived = count;
}
