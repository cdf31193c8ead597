// grt_device.hpp -- per-photon FP64 numerics of the traced path.
//
// Every function is __host__ __device__: the tracing kernel inlines them on
// gfx950, and the CPU harness under tests/ compiles the very same source for
// the host to cross-check the kernel's logic against the oracle.
//
// Arithmetic follows the reference's expressions term by term (same
// operations, same order) and the library is compiled without automatic FMA
// contraction.  The right-hand side and the RK stage sums fuse each product
// with the sum it joins through explicit fma calls (the rounding of a
// contracting build of the reference, which the oracle's integrators
// reproduce); everything else rounds once per +, -, *, / and sqrt.  The step
// controller's x^(1/4) is correctly rounded (quarter_root).  Citations are
// file:line in the reference's src/ directory.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grt_extension_points.h"

#include <type_traits>

#define GRT_HD __host__ __device__ __forceinline__

namespace grt {

constexpr double kPi = 3.14159265358979323846;   // utils.hpp:16
constexpr double kTwoPi = 2.0 * kPi;

GRT_HD bool is_finite(double x) { return __builtin_isfinite(x); }

// v, passed through an empty asm statement the compiler cannot see into
// (device code): what is computed from the result stays below the point of
// the call, and branches on it are not threaded across that point.
template <class V>
GRT_HD V opaque(V v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v));
#endif
    return v;
}

// Wavefront priority of the tracing loop.  The two wavefronts that share a
// SIMD alternate between a step's Runge-Kutta stages (long runs of
// independent FP64 arithmetic) and the rest of the iteration: the step
// controller, the post-step, the bookkeeping and the refills (short
// dependency chains, branches and memory waits).  The rest runs at the higher
// priority, so its next instruction, once ready, issues ahead of the other
// wavefront's stage arithmetic, which fills the remaining issue slots; less
// time then passes with both wavefronts outside their stages, when the
// vector ALU idles.  (C2 render: 3% faster than equal priorities; the
// opposite assignment is 5% slower.)  With RK4, which has no controller, the
// priority is raised only for a fluid sample inside the grid's radial range
// (the search and the events after a hit): a disk-mode trace, with transfer
// on every step, runs 21% faster that way, while raising it for every
// post-step slowed an image render, whose post-steps are short, by 3%.
// Results do not depend on it.
GRT_HD void priority_stages()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_setprio(0);
#endif
}

GRT_HD void priority_control()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_setprio(2);
#endif
}

enum : uint8_t { kDisk = 0, kBackward = 1, kForward = 2, kArrived = 3, kRejected = 4 };

// Constants captured by the step functor (integrate_cart_ks.hpp:35-61), with
// the camera basis evaluated once on the host.
struct StepParams {
    double a, m;
    double r_min, r_max;
    double atol, rtol, smin, smax, safety;
    int integrator, mode;
    int absorb, emit, scatter;
    double kabs, kemit, albedo;
    uint64_t seed;
    uint32_t stream;
    int generator;
    double norm, e_cell, slope, nu_min, nu_max, nu_custom;
    double cam_d, cam_half_w, cam_half_h;
    double cam_n[3], cam_t[3], cam_f[3];      // normal, theta_hat, phi_hat
    double cam_dn[3];                          // camera_distance * normal
    double length_scale, time_scale, density_scale;
};

// ---------------------------------------------------------------------------
// Kerr-Schild geometry, metrics/kerr_schild_core.hpp
// ---------------------------------------------------------------------------

GRT_HD double sqrt_ieee(double x);   // IEEE sqrt, defined with the exact sequences below
GRT_HD double div_rn(double n, double d, double y);   // correctly rounded n / d, below
GRT_HD double rcp_rn(double d);                       // correctly rounded 1 / d, below

// kerr_schild_core.hpp:28-33
GRT_HD double ks_radius(double x, double y, double z, double a)
{
    const double rsq = x * x + y * y + z * z;
    const double A = rsq - a * a;
    const double az = a * z;
    const double S = sqrt_ieee(A * A + 4.0 * (az * az));
    return sqrt_ieee(0.5 * (A + S));
}

// The radii for which ks_H and ks_l divide through correctly rounded
// reciprocals (div_rn, with |a| < 2^150 as well): the divisors r, r^2 + a^2
// and r^4 + a^2 z^2 (|z| <= r) then lie far inside the normal range, as in
// ks_rhs.  Outside it the IEEE division is used.
GRT_HD bool ks_div_range(double r)
{
    return r > 0x1p-150 && r < 0x1p150;
}

// kerr_schild_core.hpp:80-83 with l_0 = 1: the g_00 entry, 2 H l_0 l_0 - 1.
GRT_HD double ks_H(double r, double z, double a, double m)
{
    const double r2 = r * r;
    const double az = a * z;
    const double D = r2 * r * r + az * az;
    if (ks_div_range(r) && fabs(a) < 0x1p150) return div_rn(m * (r2 * r), D, rcp_rn(D));
    return m * (r2 * r) / D;
}

// l_mu of kerr_schild_core.hpp:50-56 (l_0 = 1 implied).
GRT_HD void ks_l(double x, double y, double z, double r, double a, double l[4])
{
    const double B = r * r + a * a;
    l[0] = 1.0;
    if (ks_div_range(r) && fabs(a) < 0x1p150) {
        const double iB = rcp_rn(B);
        l[1] = div_rn(r * x + a * y, B, iB);
        l[2] = div_rn(r * y - a * x, B, iB);
        l[3] = div_rn(z, r, rcp_rn(r));
        return;
    }
    l[1] = (r * x + a * y) / B;
    l[2] = (r * y - a * x) / B;
    l[3] = z / r;
}

// g_{mu nu} = eta + 2 H l l, kerr_schild_core.hpp:99-110 (row-major 4x4).
GRT_HD void ks_metric(const double X[4], double a, double m, double g[16])
{
    const double r = ks_radius(X[1], X[2], X[3], a);
    double l[4];
    ks_l(X[1], X[2], X[3], r, a, l);
    const double twoH = 2.0 * ks_H(r, X[3], a, m);
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) {
            const double eta = (mu != nu) ? 0.0 : (mu == 0 ? -1.0 : 1.0);
            g[4 * mu + nu] = twoH * l[mu] * l[nu] + eta;
        }
    }
}

// n / d correctly rounded, given y = 1/d correctly rounded: q = n y is within
// an ulp of n/d, its residual n - q d is exact with fma, and q + residual * y
// rounds to the IEEE quotient (Markstein's theorem for binary floating point;
// see Muller et al., Handbook of Floating-Point Arithmetic, ch. 5).  Valid
// away from overflow/underflow, which the geometry here never approaches.
GRT_HD double div_rn(double n, double d, double y)
{
    const double q = n * y;
    const double res = fma(-q, d, n);
    return fma(res, y, q);
}

// Correctly rounded sqrt(x) and 1/d for finite, positive, normal-range
// arguments (2^-700 < x, d < 2^700).  On the device these are the cores of
// the sequences the compiler emits for sqrt() and 1.0 / d (a v_rsq_f64 /
// v_rcp_f64 seed refined by Newton steps and a final exact-residual
// correction) without the exponent pre-scaling and special-value fix-up
// those carry for denormal, huge, zero or non-finite arguments: for the
// arguments here the full sequences reduce to exactly these operations, so
// the results are the same correctly rounded values, bit for bit.
GRT_HD double sqrt_rn(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
#else
    return sqrt(x);
#endif
}

GRT_HD double rcp_rn(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    const double r = fma(-d, y, 1.0);
    return fma(r, y, y);
#else
    return 1.0 / d;
#endif
}

// sqrt_rn, also returning in q an approximation of 1/sqrt(x) (relative
// error about 2^-47: twice the refined half-reciprocal of the sequence).
GRT_HD double sqrt_rn_q(double x, double& q)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    q = h + h;
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
#else
    q = 1.0 / sqrt(x);
    return sqrt(x);
#endif
}

// rcp_rn from a seed y0 of relative error below about 2^-40 instead of the
// hardware's 2^-24 one (no v_rcp_f64, one Newton step fewer): one Newton
// step leaves y within half an ulp and a hair of 1/d, where the final
// exact-residual step provably returns the correctly rounded reciprocal,
// as it does after rcp_rn's two steps.  Seeds used: the 1/sqrt(x) of
// sqrt_rn_q for 1/sqrt_rn(x), and (1/D)^2 for 1/(D D).
GRT_HD double rcp_rn_seeded(double d, double y0)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double e = fma(-d, y0, 1.0);
    const double y = fma(y0, e, y0);
    const double r = fma(-d, y, 1.0);
    return fma(r, y, y);
#else
    (void)y0;
    return 1.0 / d;
#endif
}

// sqrt(x) and 1.0 / sqrt(x) with the IEEE results: the short sequences of
// sqrt_rn / rcp_rn inside their exact range, the library ones outside it
// (zero, denormal, huge or non-finite arguments).  A file compiled with
// relaxed math (GRT_RELAXED_MATH, the mixed form) takes, inside the range,
// the hardware's v_rsq_f64 seed (relative error below 2^-24) refined by one
// coupled Newton step instead (about 2^-46 relative).
GRT_HD double sqrt_ieee(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (x > 0x1p-600 && x < 0x1p600) {
#if defined(GRT_RELAXED_MATH)
        const double y = __builtin_amdgcn_rsq(x);
        const double g = x * y, h = 0.5 * y;
        return fma(g, fma(-h, g, 0.5), g);
#else
        return sqrt_rn(x);
#endif
    }
#endif
    return sqrt(x);
}

GRT_HD double rsqrt_ieee(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (x > 0x1p-600 && x < 0x1p600) {
#if defined(GRT_RELAXED_MATH)
        const double y = __builtin_amdgcn_rsq(x);
        const double g = x * y, h = 0.5 * y;
        return 2.0 * fma(h, fma(-h, g, 0.5), h);
#else
        return rcp_rn(sqrt_rn(x));
#endif
    }
#endif
    return 1.0 / sqrt(x);
}

// Hamilton's equations for H = 1/2 g^{mu nu} p_mu p_nu in Cartesian
// Kerr-Schild form, kerr_schild_core.hpp:161-192 (with r, dr/dx^k of :36-46,
// l, dl/dx^k of :59-77 and H, dH/dx^k of :86-95 expanded for k = x, y, z).
// out[4] (dp_t/dlambda) is identically zero and is not written.
//
// Rounding: every product that the reference's expressions add to a sum is
// fused with that addition (an explicit fma: one rounding, as a contracting
// compiler builds the reference), every quotient is the correctly rounded
// IEEE quotient and every root the correctly rounded root.  The oracle's
// integrator evaluates the same operations in the same order, so the two
// agree bit for bit; the file is compiled without automatic contraction so
// that no other operation is fused.

GRT_HD void ks_rhs(const double s[8], double a, double m, double out[8])
{
    const double x = s[1], y = s[2], z = s[3];
    const double p0 = s[4], p1 = s[5], p2 = s[6], p3 = s[7];
    const double a2 = a * a;

    const double A = fma(z, z, fma(y, y, x * x)) - a2;
    const double az = a * z;
    const double az2 = az * az;
    // S > 0 and r > 0 away from the ring singularity, which photons never
    // reach (they terminate at r_term_min)
    double qS, qr;
    const double S = sqrt_rn_q(fma(A, A, 4.0 * az2), qS);
    const double r = sqrt_rn_q(0.5 * (A + S), qr);
    const double tr = 2.0 * r;
    const double B = fma(r, r, a2);
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double D = fma(r3, r, az2);
    // Every quotient n/d is the correctly rounded IEEE quotient, evaluated
    // from the correctly rounded reciprocal y = 1/d of its divisor with
    // Markstein's correction (div_rn), so the five divisors cost one
    // division each instead of one per quotient.  1/(2r) = (1/r)/2 exactly.
    // Three of the five reciprocals start from seeds at hand (the roots'
    // 1/sqrt and (1/D)^2) instead of a v_rcp_f64.
    const double D2 = D * D;
    const double iS = rcp_rn_seeded(S, qS), ir = rcp_rn_seeded(r, qr), iB = rcp_rn(B);
    const double iD = rcp_rn(D), iD2 = rcp_rn_seeded(D2, iD * iD);
    const double itr = 0.5 * ir;
    const double f = 1.0 + div_rn(A, S, iS);
    const double a2z2 = 2.0 * a2 * z;
    const double drx = div_rn(x * f, tr, itr);
    const double dry = div_rn(y * f, tr, itr);
    const double drz = div_rn(fma(z, f, div_rn(a2z2, S, iS)), tr, itr);

    const double l1 = div_rn(fma(r, x, a * y), B, iB);
    const double l2 = div_rn(fma(r, y, -(a * x)), B, iB);
    const double l3 = div_rn(z, r, ir);
    const double dBx = tr * drx, dBy = tr * dry, dBz = tr * drz;
    // dl[k][mu]; the Kronecker terms of :73-75 reduce to +-r, +-a or 1.
    const double dl_x1 = div_rn(fma(-l1, dBx, fma(drx, x, r)), B, iB);
    const double dl_x2 = div_rn(fma(-l2, dBx, fma(drx, y, -a)), B, iB);
    const double dl_x3 = div_rn(-(l3 * drx), r, ir);
    const double dl_y1 = div_rn(fma(-l1, dBy, fma(dry, x, a)), B, iB);
    const double dl_y2 = div_rn(fma(-l2, dBy, fma(dry, y, r)), B, iB);
    const double dl_y3 = div_rn(-(l3 * dry), r, ir);
    const double dl_z1 = div_rn(fma(-l1, dBz, drz * x), B, iB);
    const double dl_z2 = div_rn(fma(-l2, dBz, drz * y), B, iB);
    const double dl_z3 = div_rn(fma(-l3, drz, 1.0), r, ir);

    const double H = div_rn(m * r3, D, iD);
    const double c3 = 3.0 * r2;
    const double c4 = 4.0 * r3;
    const double dH_x = div_rn(m * fma(c3 * drx, D, -(r3 * (c4 * drx))), D2, iD2);
    const double dH_y = div_rn(m * fma(c3 * dry, D, -(r3 * (c4 * dry))), D2, iD2);
    const double dH_z = div_rn(m * fma(c3 * drz, D, -(r3 * fma(c4, drz, a2z2))), D2, iD2);

    // l^mu = (-1, l1, l2, l3)
    const double lp = fma(l3, p3, fma(l2, p2, fma(l1, p1, -p0)));
    const double twoH = 2.0 * H;
    out[0] = fma(twoH, lp, -p0);
    out[1] = fma(-(twoH * l1), lp, p1);
    out[2] = fma(-(twoH * l2), lp, p2);
    out[3] = fma(-(twoH * l3), lp, p3);
    const double dlp_x = fma(dl_x3, p3, fma(dl_x2, p2, dl_x1 * p1));
    const double dlp_y = fma(dl_y3, p3, fma(dl_y2, p2, dl_y1 * p1));
    const double dlp_z = fma(dl_z3, p3, fma(dl_z2, p2, dl_z1 * p1));
    const double lp2 = lp * lp;
    const double hl = twoH * lp;
    out[5] = fma(hl, dlp_x, dH_x * lp2);
    out[6] = fma(hl, dlp_y, dH_y * lp2);
    out[7] = fma(hl, dlp_z, dH_z * lp2);
}

// Reciprocal and square root for the relaxed right-hand side (positive,
// normal-range arguments): the hardware's FP64 seeds (v_rcp_f64 and
// v_rsq_f64, relative error below 2^-24 as measured on gfx950) refined by
// one Newton step each, to a relative error of about 2^-47: 3 and 5
// instructions instead of the 7 and 17 of the library sequences, and far
// below the 1e-7 tolerance of the step control.
GRT_HD double rcp_newton1(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rcp(d);
    return fma(y, fma(-d, y, 1.0), y);
#else
    return 1.0 / d;
#endif
}

GRT_HD double sqrt_newton1(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    return fma(g, fma(-h, g, 0.5), g);
#else
    return sqrt(x);
#endif
}

// sqrt_newton1, also returning 1/sqrt(x) in q from the same seed and Newton
// step (the coupled iteration refines both; about 2^-46 relative).
GRT_HD double sqrt_newton1_q(double x, double& q)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    const double hr = fma(h, r, h);
    q = hr + hr;
    return fma(g, r, g);
#else
    q = 1.0 / sqrt(x);
    return sqrt(x);
#endif
}

// Hamilton's equations of ks_rhs in fewer operations (mixed-precision form).
// Only the contractions the equations need are formed:
//   d_k(l.p) = (the explicit x^k dependence) + dr/dx^k * C,
//     C = U/B - 2 r (r U + a V)/B^2 - W/r^2,  U = x p1 + y p2,
//     V = y p1 - x p2, W = z p3, B = r^2 + a^2,
//   with explicit parts (r p1 - a p2)/B, (r p2 + a p1)/B, p3/r;
//   d_k H = H (3 dr_k / r - dD_k / D), D = r^4 + a^2 z^2,
//     dD_k = 4 r^3 dr_k + 2 a^2 z delta_kz,
// instead of the nine d_k l_j and the quotient form of d_k H
// (kerr_schild_core.hpp:59-95, 161-192): about 100 operations, four
// reciprocals and two square roots against 248 operations, 21 divisions
// and two square roots.  The result equals ks_rhs up to rounding (and the
// ~2^-47 of rcp_newton1 / sqrt_newton1 on the device).
GRT_HD void ks_rhs_fast(const double s[8], double a, double m, double out[8])
{
    const double x = s[1], y = s[2], z = s[3];
    const double p0 = s[4], p1 = s[5], p2 = s[6], p3 = s[7];
    const double a2 = a * a;
    const double A = (x * x + y * y + z * z) - a2;
    const double az = a * z;
    double iS, ir;
    const double S = sqrt_newton1_q(A * A + 4.0 * (az * az), iS);
    const double r = sqrt_newton1_q(0.5 * (A + S), ir);
    const double r2 = r * r;
    const double iB = rcp_newton1(r2 + a2);
    const double r3 = r2 * r;
    const double iD = rcp_newton1(r3 * r + az * az);
    const double f = (1.0 + A * iS) * (0.5 * ir);
    const double drx = x * f, dry = y * f;
    const double drz = z * f + a2 * z * iS * ir;
    const double U = x * p1 + y * p2, V = y * p1 - x * p2, W = z * p3;
    const double Q = (r * U + a * V) * iB;          // l1 p1 + l2 p2
    const double lp = (-p0 + Q) + W * ir;
    const double H = m * r3 * iD;
    const double twoH = 2.0 * H;
    const double hl = twoH * lp;
    out[0] = -p0 + hl;
    out[1] = p1 - hl * (r * x + a * y) * iB;
    out[2] = p2 - hl * (r * y - a * x) * iB;
    out[3] = p3 - hl * z * ir;
    const double Cb = U * iB - 2.0 * r * Q * iB - W * ir * ir;
    const double dlp_x = (r * p1 - a * p2) * iB + drx * Cb;
    const double dlp_y = (r * p2 + a * p1) * iB + dry * Cb;
    const double dlp_z = p3 * ir + drz * Cb;
    const double g = 3.0 * ir - 4.0 * r3 * iD;       // dH_k / H = g dr_k - (2 a^2 z / D) delta_kz
    const double lp2 = lp * lp;
    out[5] = H * (g * drx) * lp2 + hl * dlp_x;
    out[6] = H * (g * dry) * lp2 + hl * dlp_y;
    out[7] = H * (g * drz - 2.0 * a2 * z * iD) * lp2 + hl * dlp_z;
}

struct RhsFast {
    GRT_HD static void eval(const double s[8], const StepParams& P, double out[8])
    {
        ks_rhs_fast(s, P.a, P.m, out);
    }
};

// Right-hand-side policy of the integrators: the FP64 evaluation above.
struct RhsExact {
    GRT_HD static void eval(const double s[8], const StepParams& P, double out[8])
    {
        ks_rhs(s, P.a, P.m, out);
    }
};

// ---------------------------------------------------------------------------
// Integrators, integrate_cart_ks.hpp:143-227
// ---------------------------------------------------------------------------

// Fehlberg 4(5) tableau, integrate_cart_ks.hpp:166-180.
constexpr double A21 = 1.0 / 4.0;
constexpr double A31 = 3.0 / 32.0, A32 = 9.0 / 32.0;
constexpr double A41 = 1932.0 / 2197.0, A42 = -7200.0 / 2197.0, A43 = 7296.0 / 2197.0;
constexpr double A51 = 439.0 / 216.0, A52 = -8.0, A53 = 3680.0 / 513.0, A54 = -845.0 / 4104.0;
constexpr double A61 = -8.0 / 27.0, A62 = 2.0, A63 = -3544.0 / 2565.0, A64 = 1859.0 / 4104.0,
                 A65 = -11.0 / 40.0;
constexpr double B41 = 25.0 / 216.0, B43 = 1408.0 / 2565.0, B44 = 2197.0 / 4104.0, B45 = -1.0 / 5.0;
constexpr double B51 = 16.0 / 135.0, B53 = 6656.0 / 12825.0, B54 = 28561.0 / 56430.0,
                 B55 = -9.0 / 50.0, B56 = 2.0 / 55.0;

// x^(1/4) correctly rounded: sqrt(sqrt(x)) is within an ulp; of it and its two
// neighbours keep the one whose fourth power is closest to x, evaluated with
// exact products (fma).  The step controller's pow(1/err, 0.25)
// (integrate_cart_ks.hpp:217) is a libm call whose last bit varies between
// libraries; using the correctly rounded value in both this kernel and the
// oracle makes the step-size sequence reproducible bit for bit.
GRT_HD double quarter_root(double x)
{
    const double y = sqrt(sqrt(x));
    double best = y, bestd = 1.0 / 0.0;
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
        const double c = d < 0 ? nextafter(y, 0.0) : (d > 0 ? nextafter(y, 1.0 / 0.0) : y);
        const double s = c * c;
        const double e1 = fma(c, c, -s);
        const double q = s * s;
        const double e2 = fma(s, s, -q);
        const double r = fabs((q - x) + (e2 + 2.0 * s * e1));
        if (r < bestd) {
            bestd = r;
            best = c;
        }
    }
    return best;
}

// The residual c^4 - x of quarter_root's candidate check, with its exact
// products, signed.
GRT_HD double quarter_residual(double c, double x)
{
    const double s = c * c;
    const double e1 = fma(c, c, -s);
    const double q = s * s;
    const double e2 = fma(s, s, -q);
    return (q - x) + (e2 + 2.0 * s * e1);
}

// quarter_root(1.0 / err), bit for bit.  For err in (2^-500, 2^500) the
// reciprocal and both roots take the short correctly rounded sequences and
// the neighbours of y are its bit pattern -+ 1 (y is positive and finite
// there), which is what nextafter returns.  Of quarter_root's three
// candidates only two can win: the residual evaluation is exact to about
// 2^-101 x while neighbouring candidates' fourth powers differ by at least
// 2^-51 x, so the residual grows strictly away from y on the side where y^4
// already exceeds x (y^4 > x: y + ulp loses to y; y^4 < x: y - ulp loses).
// The remaining neighbour is compared as quarter_root's loop would: y - ulp
// is examined before y and keeps a tie, y + ulp after y and needs a strictly
// smaller residual.
GRT_HD double quarter_root_rcp(double err)
{
#if defined(__HIP_DEVICE_COMPILE__) && !defined(GRT_RELAXED_MATH)
    if (err > 0x1p-500 && err < 0x1p500) {
        const double x = rcp_rn(err);
        const double y = sqrt_rn(sqrt_rn(x));
        const double ry = quarter_residual(y, x);
        const bool above = ry > 0.0;
        const double c = __longlong_as_double(__double_as_longlong(y) + (above ? -1 : 1));
        const double rc = fabs(quarter_residual(c, x));
        const double r0 = fabs(ry);
        return (above ? rc <= r0 : rc < r0) ? c : y;
    }
#endif
    return quarter_root(1.0 / err);
}

// Component 4 (p_t) has a zero derivative, so every stage leaves it equal to
// s[4]; the loops below skip it.
#define GRT_FOR_STATE(i) _Pragma("unroll") for (int i = 0; i < 8; ++i) if (i != 4)

// Classic RK4, integrate_cart_ks.hpp:143-155.
template <class Rhs = RhsExact>
GRT_HD void rk4_step(double s[8], double h, const StepParams& P)
{
    double k1[8], k2[8], k3[8], k4[8], t[8];
    const double h2 = 0.5 * h;
    Rhs::eval(s, P, k1);
    t[4] = s[4];
    GRT_FOR_STATE(i) t[i] = fma(h2, k1[i], s[i]);
    Rhs::eval(t, P, k2);
    GRT_FOR_STATE(i) t[i] = fma(h2, k2[i], s[i]);
    Rhs::eval(t, P, k3);
    GRT_FOR_STATE(i) t[i] = fma(h, k3[i], s[i]);
    Rhs::eval(t, P, k4);
    const double h6 = h / 6.0;
    GRT_FOR_STATE(i) s[i] = fma(h6, k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i], s[i]);
}

// One RKF45 attempt with precomputed k1, integrate_cart_ks.hpp:160-227.
// Updates *h to the next proposal; returns true (and replaces s with the
// 5th-order solution) when accepted.
//
// The stage sums are accumulated left to right as each k_j becomes
// available (P4 = a41 k1 + a42 k2, later P4 + a43 k3, ...), which is exactly
// the evaluation order of the reference's expressions, so only the partial
// sums -- not k2..k6 -- stay live in registers.  Each product joins its sum
// in one fma (the rounding of ks_rhs), as does the final s + h * sum.
//
// kMixed evaluates the controller's err^(-1/4) = sqrt(1/sqrt(err)) from the
// hardware's FP64 seeds with one Newton step each (sqrt_newton1_q,
// sqrt_newton1; about 2^-45 relative) instead of the correctly rounded root.
// It stays FP64: an FP32 root moves every proposal by ~1e-7 relative, which
// moves the image-mode emission and arrival points (RK step
// endpoints) across cell and pixel edges in ~13% of the lit pixels of a
// 2048^2 torus image, against under 1% with FP64.
template <class Rhs = RhsExact, int kMixed = 0>
GRT_HD bool rk45_attempt(double s[8], double* h_io, const double k1[8], const StepParams& P)
{
    const double h = *h_io;
    const double h21 = h * A21;
    double k[8], t[8], P4[8], P5[8], P6[8], X4[8], X5[8];
    t[4] = s[4];
    GRT_FOR_STATE(i) t[i] = fma(h21, k1[i], s[i]);
    Rhs::eval(t, P, k);                                   // k2
    GRT_FOR_STATE(i) {
        t[i] = fma(h, fma(A32, k[i], A31 * k1[i]), s[i]);
        P4[i] = fma(A42, k[i], A41 * k1[i]);
        P5[i] = fma(A52, k[i], A51 * k1[i]);
        P6[i] = fma(A62, k[i], A61 * k1[i]);
    }
    Rhs::eval(t, P, k);                                   // k3
    GRT_FOR_STATE(i) {
        t[i] = fma(h, fma(A43, k[i], P4[i]), s[i]);
        P5[i] = fma(A53, k[i], P5[i]);
        P6[i] = fma(A63, k[i], P6[i]);
        X4[i] = fma(B43, k[i], B41 * k1[i]);
        X5[i] = fma(B53, k[i], B51 * k1[i]);
    }
    Rhs::eval(t, P, k);                                   // k4
    GRT_FOR_STATE(i) {
        t[i] = fma(h, fma(A54, k[i], P5[i]), s[i]);
        P6[i] = fma(A64, k[i], P6[i]);
        X4[i] = fma(B44, k[i], X4[i]);
        X5[i] = fma(B54, k[i], X5[i]);
    }
    Rhs::eval(t, P, k);                                   // k5
    GRT_FOR_STATE(i) {
        t[i] = fma(h, fma(A65, k[i], P6[i]), s[i]);
        X4[i] = fma(B45, k[i], X4[i]);
        X5[i] = fma(B55, k[i], X5[i]);
    }
    Rhs::eval(t, P, k);                                   // k6
    priority_control();   // the error norm and the controller's root

    double x5[8];
    double err = 0.0;
    bool finite = is_finite(s[4]);
    x5[4] = s[4];
    GRT_FOR_STATE(i) {
        const double x4 = fma(h, X4[i], s[i]);
        x5[i] = fma(h, fma(B56, k[i], X5[i]), s[i]);
        const double sc = fma(P.rtol, fmax(fabs(x5[i]), fabs(x4)), P.atol);
        // the mixed form divides with the seeded reciprocal (a zero sc, only
        // possible with atol = 0, gives a non-finite ratio either way)
        const double ratio = kMixed ? fabs(x5[i] - x4) * rcp_newton1(sc) : fabs(x5[i] - x4) / sc;
        finite = finite && is_finite(x5[i]) && is_finite(ratio);
        err = fmax(err, ratio);
    }
    // The state update is a select, not a branch: accepted and rejected
    // lanes leave the attempt with the state in the same registers.
    double scale;
    if (!finite) {
        scale = P.smin;
    } else if (err == 0.0) {
        scale = P.smax;
    } else {
        if (kMixed) {
            double q = 0.0;   // 1/sqrt(err); err^(-1/4) = sqrt(q)
            if (err > 0x1p-900 && err < 0x1p900) sqrt_newton1_q(err, q);
            scale = P.safety * ((err > 0x1p-900 && err < 0x1p900) ? sqrt_newton1(q)
                                                                  : 1.0 / sqrt(sqrt(err)));
        } else
            scale = P.safety * quarter_root_rcp(err);
        scale = fmin(P.smax, fmax(P.smin, scale));
    }
    *h_io = h * scale;
    const bool accepted = finite && err <= 1.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] = accepted ? x5[i] : s[i];
    return accepted;
}

// ---------------------------------------------------------------------------
// Fluid grid and nearest-cell search, radiative_transfer/scattering.hpp:41-251
// ---------------------------------------------------------------------------

// Cell record, built on the host from the .npy fields: density and
// temperature times their unit scales, and the fluid four-velocity and
// magnetic four-vector already mapped to Cartesian Kerr-Schild components
// (fluid_frame.hpp:29-49) and scaled (scattering.hpp:212-241).
constexpr int kCellDoubles = 10;

struct GridView {
    int nr, nth, nph;
    int available;
    int separable;            // 1: coordinates depend only on their own index
    // separable fast path: search tables, 4 * (nr + nth + nph) doubles
    // (layout and meaning: see locate_separable)
    const double* tab;
    // the 1-D axes r_i, theta_j, phi_k of a separable grid (cell centres)
    const double* axes;
    // general path: the full (nr, nth, nph) coordinate fields
    const double* r3;
    const double* th3;
    const double* ph3;
    const double* cells;      // kCellDoubles per cell, C order
    // bucket index of the separable search (axis 0 radial, 1 polar,
    // 2 azimuth; key_nb 0: none), see axis_bucket
    int key_shift[3], key_base[3], key_nb[3];
    // separable path: the radial windows' outer bounds, r_0 e^-thr and
    // r_(nr-1) e^thr (table entries 2 nr and 4 nr - 1), kept here so the test
    // that rejects most samples reads launch constants instead of LDS
    double win_lo, win_hi;
};

// Bucket of a positive double w for the separable search: the top bits of
// its IEEE pattern (monotone in w), offset and clamped to [0, nb).  With
// table entries e_j keyed the same way, every entry in a lower bucket is
// below w and every entry in a higher bucket above it, so a count of
// entries <= w needs comparisons only inside w's bucket; the host stores,
// per bucket b, the number of entries keyed below b (uint16, after the
// search tables).  NaN goes to bucket 0, where no entry compares true.
GRT_HD int axis_bucket(double w, int shift, int base, int nb)
{
    if (!(w == w)) return 0;
    uint64_t u;
    __builtin_memcpy(&u, &w, sizeof u);
    const int64_t k = (int64_t)(u >> shift) - (int64_t)base;
    return (int)(k < 0 ? 0 : (k >= nb ? nb - 1 : k));
}

// scattering.hpp:41-45.  fmod(d, 2 pi) is d itself for 0 <= d < 2 pi, the
// only range the arguments here reach; the library call is kept for the rest.
GRT_HD double periodic_distance(double lhs, double rhs)
{
    double d = fabs(lhs - rhs);
    if (!(d < kTwoPi)) d = fmod(d, kTwoPi);
    return fmin(d, kTwoPi - d);
}

// Number of entries <= v in the ascending table t[0..n).
GRT_HD int count_le(const double* t, int n, double v)
{
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (t[lo + half] <= v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// Number of leading entries >= v in the descending table t[0..n).
GRT_HD int count_ge_desc(const double* t, int n, double v)
{
    int lo = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (t[lo + half] >= v) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// The count of bucket [l, h) of an ascending (Desc: descending) table t that
// is <= v (Desc: >= v), decided on the first two entries: t[l] and t[l + 1]
// are read whatever h - l is (every table is followed by more tables in
// LDS), and the rest of the bucket matters only when it has more than two
// entries and the second one already counts, which the few crowded buckets
// near the polar axis's ends can need.  Equal to count_le / count_ge_desc
// of the bucket.
template <bool Desc>
GRT_HD int bucket_count_two(const double* t, int l, int h, double t0, double t1, double v)
{
    const bool in0 = Desc ? t0 >= v : t0 <= v;
    const bool in1 = Desc ? t1 >= v : t1 <= v;
    int c = l + ((h > l && in0) ? 1 : 0) + ((h > l + 1 && in1) ? 1 : 0);
    if (h - l > 2 && in1)
        c += Desc ? count_ge_desc(t + l + 2, h - l - 2, v) : count_le(t + l + 2, h - l - 2, v);
    return c;
}

// count_le / count_ge_desc narrowed by the bucket counts B of the table
// (axis_bucket of the keyed query w); equal to the plain counts.
GRT_HD int bucket_count_le(const double* t, const uint16_t* B, int shift, int base, int nb,
                           double w, double v)
{
    const int b = axis_bucket(w, shift, base, nb);
    const int l = B[b], h = B[b + 1];
    return bucket_count_two<false>(t, l, h, t[l], t[l + 1], v);
}

GRT_HD int bucket_count_ge_desc(const double* t, const uint16_t* B, int shift, int base, int nb,
                                double w, double v)
{
    const int b = axis_bucket(w, shift, base, nb);
    const int l = B[b], h = B[b + 1];
    return bucket_count_two<true>(t, l, h, t[l], t[l + 1], v);
}

// A monotone stand-in for atan2(Y, X) mapped to [0, 2 pi): the "diamond
// angle" in [0, 4), one division, no transcendental.  Quadrant by quadrant
// (Y >= 0, X >= 0: Y / (X + Y), or 0 at the origin; Y >= 0, X < 0:
// 1 + (-X) / (Y - X); both negative: 2 + (-Y) / (-X - Y); else
// 3 + X / (X - Y)), the same operations as the host's table values; the
// quadrant only selects the operands, so a wavefront whose lanes lie in
// several quadrants divides once.
GRT_HD double pseudo_angle(double X, double Y)
{
    const bool yp = Y >= 0.0, xp = X >= 0.0;
    const double num = yp ? (xp ? Y : -X) : (xp ? X : -Y);
    const double den = yp ? (xp ? X + Y : Y - X) : (xp ? X - Y : -X - Y);
    const double base = yp ? (xp ? 0.0 : 1.0) : (xp ? 3.0 : 2.0);
    const double q = num / den;
    if (yp && xp) return den > 0.0 ? q : 0.0;
    return base + q;
}

// Separable fast path of locate_structured_grid_cell (scattering.hpp:151-180);
// (i, j, k) are written only when the point lies in a cell.
// With r(i,j,k) = r(i), theta(i,j,k) = theta(j), phi(i,j,k) = phi(k) the three
// coordinate-descent passes converge in the first, and each nearest_* scan
// (first strict minimum) reduces to a bracket search plus the candidates
// around it.  Every comparison of the reference is replaced by an equivalent
// one on quantities the host tabulated, so no log, acos, atan2 or fmod runs
// here (decisions can differ from the reference's only when a point lies
// within rounding of a cell boundary or midpoint):
//   radial   |log(r/c_lo)| vs |log(r/c_hi)|  <=>  r vs sqrt(c_lo c_hi)
//            |log(r/c_i)| <= thr_i          <=>  c_i e^-thr_i <= r <= c_i e^thr_i
//   polar    theta = acos(ct); theta_j <= theta <=> cos theta_j >= ct, the
//            midpoint and the +-thr_j window mapped through cos
//   azimuth  periodic distance to phi_k is decreasing in the dot product of
//            (X, Y) = (r x + a y, r y - a x) with (cos phi_k, sin phi_k); the
//            bracket comes from a monotone pseudo-angle; the window test is
//            dot >= cos(thr_k) |(X, Y)|
// Table layout (doubles), n = nr, m = nth, q = nph:
//   [0,n) r_j   [n,2n) sqrt(r_j r_j+1)   [2n,3n) r_j e^-thr   [3n,4n) r_j e^thr
//   then cos th_j | cos of midpoints | cos(min(pi, th_j + thr)) | cos(max(0, th_j - thr))
//   then pseudo(phi_k) | cos phi_k | sin phi_k | cos(thr_k) (-2: single entry)
//   then the uint16 bucket counts of the three axes (see axis_bucket): keyed
//   on r, on 3 - cos theta and on pseudo + 4, so that every key is positive.
// Beyond the outer window of the last radial cell or below the inner window
// of the first, the nearest cell is that end cell and its window test fails:
// no cell can be found, and no search is needed.
GRT_HD bool separable_radial_windows(const GridView& G, double radius)
{
    return radius <= G.win_hi && radius >= G.win_lo;
}

GRT_HD bool locate_separable(const GridView& G, double radius, double ct, double X, double Y,
                             int& i_out, int& j_out, int& k_out)
{
    const int nr = G.nr, nth = G.nth, nph = G.nph;
    const double* R = G.tab;
    const double* T = R + 4 * nr;
    const double* F = T + 4 * nth;
    const uint16_t* BR = reinterpret_cast<const uint16_t*>(F + 4 * nph);
    const uint16_t* BT = BR + G.key_nb[0] + 1;
    const uint16_t* BF = BT + G.key_nb[1] + 1;
    if (!separable_radial_windows(G, radius)) return false;
    // The three axes are independent: each step below reads all three
    // axes' entries before the next step needs any of them, so a lookup
    // costs about four dependent LDS round trips instead of twelve.
    const double pa = pseudo_angle(X, Y);
    int lo_r, lo_t, lo_f;   // brackets: entries <= r, >= cos theta, <= pseudo-angle, minus one
    if (G.key_nb[0] > 0) {
        const int br = axis_bucket(radius, G.key_shift[0], G.key_base[0], G.key_nb[0]);
        const int bt = axis_bucket(3.0 - ct, G.key_shift[1], G.key_base[1], G.key_nb[1]);
        const int bf = axis_bucket(pa + 4.0, G.key_shift[2], G.key_base[2], G.key_nb[2]);
        const int lr = BR[br], hr = BR[br + 1];
        const int lt = BT[bt], ht = BT[bt + 1];
        const int lf = BF[bf], hf = BF[bf + 1];
        const double r0 = R[lr], r1 = R[lr + 1];
        const double t0 = T[lt], t1 = T[lt + 1];
        const double f0 = F[lf], f1 = F[lf + 1];
        lo_r = bucket_count_two<false>(R, lr, hr, r0, r1, radius) - 1;
        lo_t = bucket_count_two<true>(T, lt, ht, t0, t1, ct) - 1;
        lo_f = bucket_count_two<false>(F, lf, hf, f0, f1, pa) - 1;
    } else {
        lo_r = count_le(R, nr, radius) - 1;
        lo_t = count_ge_desc(T, nth, ct) - 1;
        lo_f = count_le(F, nph, pa) - 1;
    }
    // nearest entries (scattering.hpp:47-97, first strict minimum): the
    // midpoint between the bracketing entries decides (radial and polar);
    // the azimuth takes the largest dot product among candidates 0, lo,
    // lo + 1 and nph - 1.  Entries are read at clamped indices and used only
    // where the reference's scan would look.
    const int mr = lo_r < 0 ? 0 : (lo_r > nr - 1 ? nr - 1 : lo_r);
    const int mt = lo_t < 0 ? 0 : (lo_t > nth - 1 ? nth - 1 : lo_t);
    const int c1 = lo_f, c2 = lo_f + 1, c3 = nph - 1;
    const int q1 = c1 < 0 ? 0 : (c1 > nph - 1 ? nph - 1 : c1);
    const int q2 = c2 < 0 ? 0 : (c2 > nph - 1 ? nph - 1 : c2);
    const double mid_r = R[nr + mr], mid_t = T[nth + mt];
    const double d0 = X * F[nph] + Y * F[2 * nph];
    const double d1 = X * F[nph + q1] + Y * F[2 * nph + q1];
    const double d2 = X * F[nph + q2] + Y * F[2 * nph + q2];
    const double d3 = X * F[nph + c3] + Y * F[2 * nph + c3];
    const int i = lo_r < 0 ? 0 : (lo_r >= nr - 1 ? nr - 1 : (radius > mid_r ? lo_r + 1 : lo_r));
    const int j = lo_t < 0 ? 0 : (lo_t >= nth - 1 ? nth - 1 : (ct < mid_t ? lo_t + 1 : lo_t));
    int k = 0;
    double best = d0;
    if (c1 > 0 && c1 < nph && d1 > best) { best = d1; k = c1; }
    if (c2 > 0 && c2 > c1 && c2 < nph && d2 > best) { best = d2; k = c2; }
    if (c3 > 0 && c3 > c2 && d3 > best) { best = d3; k = c3; }
    i_out = i;
    j_out = j;
    k_out = k;
    // within_* tests (scattering.hpp:99-149): the cell is used only when all
    // three hold
    const bool in_r = R[2 * nr + i] <= radius && radius <= R[3 * nr + i];
    const bool in_t = T[2 * nth + j] <= ct && ct <= T[3 * nth + j];
    const double norm = sqrt(X * X + Y * Y);
    const double cthr = F[3 * nph + k];
    bool in_f;
    if (cthr >= -1.0) {
        in_f = best >= cthr * norm;
    } else {
        // single azimuth: |sin d| <= 1e-10 on the near side
        const double cross = X * F[2 * nph + k] - Y * F[nph + k];
        in_f = best >= 0.0 && fabs(cross) <= 1.0e-10 * norm;
    }
    return in_r && in_t && in_f;
}

GRT_HD double c3at(const double* a, const GridView& G, int i, int j, int k)
{
    return a[((size_t)i * G.nth + j) * G.nph + k];
}

// General path: the reference's three coordinate-descent passes of linear
// scans over arbitrary 3-D coordinate fields (scattering.hpp:47-180).
GRT_HD bool locate_general(const GridView& G, double radius, double polar, double az,
                           int& i_out, int& j_out, int& k_out)
{
    int i = G.nr / 2, j = G.nth / 2, k = G.nph / 2;
    for (int it = 0; it < 3; ++it) {
        {
            const double first = c3at(G.r3, G, 0, j, k);
            if (!(radius > 0.0) || !(first > 0.0)) return false;
            int best = 0;
            double bd = fabs(log(radius / first));
            for (int q = 1; q < G.nr; ++q) {
                const double c = c3at(G.r3, G, q, j, k);
                if (!(c > 0.0)) continue;
                const double d = fabs(log(radius / c));
                if (d < bd) { best = q; bd = d; }
            }
            i = best;
        }
        {
            int best = 0;
            double bd = fabs(polar - c3at(G.th3, G, i, 0, k));
            for (int q = 1; q < G.nth; ++q) {
                const double d = fabs(polar - c3at(G.th3, G, i, q, k));
                if (d < bd) { best = q; bd = d; }
            }
            j = best;
        }
        {
            int best = 0;
            double bd = periodic_distance(az, c3at(G.ph3, G, i, j, 0));
            for (int q = 1; q < G.nph; ++q) {
                const double d = periodic_distance(az, c3at(G.ph3, G, i, j, q));
                if (d < bd) { best = q; bd = d; }
            }
            k = best;
        }
    }
    i_out = i;
    j_out = j;
    k_out = k;
    // within_radial_cell, scattering.hpp:99-117
    {
        const double c = c3at(G.r3, G, i, j, k);
        if (!(c > 0.0) || !(radius > 0.0)) return false;
        const double cd = fabs(log(radius / c));
        if (G.nr == 1) {
            if (!(cd <= 1.0e-10)) return false;
        } else {
            double span = 0.0;
            if (i > 0 && c3at(G.r3, G, i - 1, j, k) > 0.0)
                span = fmax(span, fabs(log(c / c3at(G.r3, G, i - 1, j, k))));
            if (i + 1 < G.nr && c3at(G.r3, G, i + 1, j, k) > 0.0)
                span = fmax(span, fabs(log(c / c3at(G.r3, G, i + 1, j, k))));
            if (!(cd <= 0.500001 * span)) return false;
        }
    }
    // within_polar_cell, scattering.hpp:119-133
    {
        const double c = c3at(G.th3, G, i, j, k);
        const double cd = fabs(polar - c);
        if (G.nth == 1) {
            if (!(cd <= 1.0e-10)) return false;
        } else {
            double span = 0.0;
            if (j > 0) span = fmax(span, fabs(c - c3at(G.th3, G, i, j - 1, k)));
            if (j + 1 < G.nth) span = fmax(span, fabs(c - c3at(G.th3, G, i, j + 1, k)));
            if (!(cd <= 0.500001 * span)) return false;
        }
    }
    // within_azimuth_cell, scattering.hpp:135-149
    {
        const double c = c3at(G.ph3, G, i, j, k);
        const double cd = periodic_distance(az, c);
        if (G.nph == 1) return cd <= 1.0e-10;
        double span = 0.0;
        if (k > 0) span = fmax(span, periodic_distance(c, c3at(G.ph3, G, i, j, k - 1)));
        if (k + 1 < G.nph) span = fmax(span, periodic_distance(c, c3at(G.ph3, G, i, j, k + 1)));
        return cd <= 0.500001 * span;
    }
}

// The Kerr-Schild metric at one point in factored form,
// g_{mu nu} = eta_{mu nu} + twoH l_mu l_nu with l_0 = 1
// (kerr_schild_core.hpp:99-110).
struct KSMetric {
    double twoH, l1, l2, l3;
};

GRT_HD KSMetric ks_metric_at(const double s[8], double r, double a, double m)
{
    double l[4];
    ks_l(s[1], s[2], s[3], r, a, l);
    return KSMetric{2.0 * ks_H(r, s[3], a, m), l[1], l[2], l[3]};
}

GRT_HD KSMetric ks_metric_at(const double s[8], double a, double m)
{
    return ks_metric_at(s, ks_radius(s[1], s[2], s[3], a), a, m);
}

// metric_inner_product (fluid_frame.hpp:6-14) through the factored metric:
// eta(u, v) + twoH (l.u)(l.v).  Equal to the 16-term sum up to rounding.
// T is the arithmetic type (double; float in the mixed-precision form).
template <class T, class U, class V>
GRT_HD T inner(const KSMetric& g, const U u[4], const V v[4])
{
    const T l1 = (T)g.l1, l2 = (T)g.l2, l3 = (T)g.l3;
    const T u0 = (T)u[0], u1 = (T)u[1], u2 = (T)u[2], u3 = (T)u[3];
    const T v0 = (T)v[0], v1 = (T)v[1], v2 = (T)v[2], v3 = (T)v[3];
    if constexpr (std::is_same<T, double>::value) {
        // fused: each product joins its sum in one fma
        const double lu = fma(l3, u3, fma(l2, u2, fma(l1, u1, u0)));
        const double lv = fma(l3, v3, fma(l2, v2, fma(l1, v1, v0)));
        const double eta = fma(u3, v3, fma(u2, v2, fma(u1, v1, -u0 * v0)));
        return fma(g.twoH * lu, lv, eta);
    } else {
        const T lu = u0 + l1 * u1 + l2 * u2 + l3 * u3;
        const T lv = v0 + l1 * v1 + l2 * v2 + l3 * v3;
        return (-u0 * v0 + u1 * v1 + u2 * v2 + u3 * v3) + (T)g.twoH * lu * lv;
    }
}

GRT_HD double inner(const KSMetric& g, const double u[4], const double v[4])
{
    return inner<double>(g, u, v);
}

// The spatial legs e[1..3] of build_fluid_tetrad (fluid_frame.hpp:62-79):
// Gram-Schmidt of the coordinate x, y, z axes under g against e[0] (set),
// in arithmetic T.  With u timelike (the check that gave e[0]) and g = eta +
// 2H l l with H >= 0, the coordinate axes span a positive-definite subspace
// whose metric is at least the identity, so every Gram-Schmidt norm is >= 1:
// the construction's checks (norm > 1e-20, finite) cannot fail here, and
// none is made.
template <class T>
GRT_HD void fluid_tetrad_axes(const KSMetric& g, T e[16])
{
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        T c[4] = {(T)0, (T)0, (T)0, (T)0};
        c[ax + 1] = (T)1;
        const T tp = inner<T>(g, e, c);
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) c[mu] += tp * e[mu];
#pragma unroll
        for (int pv = 0; pv < ax; ++pv) {
            const T pr = inner<T>(g, e + 4 * (pv + 1), c);
#pragma unroll
            for (int mu = 0; mu < 4; ++mu) c[mu] -= pr * e[4 * (pv + 1) + mu];
        }
        const T nn = inner<T>(g, c, c);
        T is;
        if constexpr (std::is_same<T, double>::value) is = rsqrt_ieee(nn);
        else is = (T)1 / sqrt(nn);
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) e[4 * (ax + 1) + mu] = c[mu] * is;
    }
}

// build_fluid_tetrad, fluid_frame.hpp:52-80: e[0] = u / |u|, then (full) the
// spatial legs, in arithmetic T.  Returns false where the reference's
// construction fails; that can only be the check of e[0] (fluid_tetrad_axes).
template <class T>
GRT_HD bool fluid_tetrad(const KSMetric& g, const double u_in[4], T e[16], bool full = true)
{
    const T u[4] = {(T)u_in[0], (T)u_in[1], (T)u_in[2], (T)u_in[3]};
    const T vn = inner<T>(g, u, u);
    if (!(vn < (T)0) || !is_finite((double)vn)) return false;
    T inv;
    if constexpr (std::is_same<T, double>::value) inv = rsqrt_ieee(-vn);
    else inv = (T)1 / sqrt(-vn);
    const T sgn = u[0] >= (T)0 ? (T)1 : (T)-1;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) e[mu] = sgn * u[mu] * inv;
    if (full) fluid_tetrad_axes<T>(g, e);
    return true;
}

// Whether sample_fluid builds the tetrad's spatial legs for a scattering
// event itself (FP64) or leaves them to maybe_scatter (the mixed form's FP32
// tetrad).  Measured on the disk-mode scattering workload: eager is 1%
// faster in FP64 (a wavefront has a scattering lane on most iterations, and
// the FP64 legs cost more in a divergent pass), lazy 3% faster in FP32.
template <class T>
GRT_HD constexpr bool scatter_legs_eager()
{
    return std::is_same<T, double>::value;
}

// sample_scattering_fluid, scattering.hpp:185-251.  On success it calls
// on_sample(fluid, tetrad e, metric g at the photon position (reused by the
// scattering update), cell index) and returns true; the caller's events run
// inside that call, where the tetrad is defined on every path.
// kSep selects the separable fast search at compile time.
// `radius` is ks_radius of s's position; `enabled` false skips the sample
// (returns false).
//
// Every way of finding no cell (disabled, no grid, a bad radius, outside the
// separable grid's radial windows, a failed search) ends in the one flag
// `found`, passed through opaque() before the branch into the hit path, so
// the compiler does not thread each exit to the end of the post-step (that
// repeats the copies of the photon state the hit path may modify on every
// exit).  Cos theta and the azimuth plane are formed only for a radius that
// passes the radial windows (nearly every sample of an image ray misses
// there).
template <int kSep, class T, class OnSample>
GRT_HD bool sample_fluid(const double s[8], double radius, bool enabled, const GridView& G,
                         const StepParams& P, bool full, OnSample&& on_sample)
{
    FluidCellCGS fl;
    T e[16];
    KSMetric g;
    uint32_t cell;
    int i = 0, j = 0, k = 0;
    bool found = false;
    if (enabled && G.available && radius > 0.0 && is_finite(radius) &&
        (!kSep || separable_radial_windows(G, radius))) {
        priority_control();   // the search and what follows a hit
        const double rad = opaque(radius);
        const double ct = fmax(-1.0, fmin(1.0, s[3] / rad));
        const double X = rad * s[1] + P.a * s[2];
        const double Y = rad * s[2] - P.a * s[1];
        if (kSep) {
            found = locate_separable(G, rad, ct, X, Y, i, j, k);
        } else {
            const double polar = acos(ct);
            double az = atan2(Y, X);
            if (az < 0.0) az += kTwoPi;
            found = locate_general(G, rad, polar, az, i, j, k);
        }
    }
    found = opaque(found ? 1 : 0) != 0;
    if (!found) return false;
    cell = (uint32_t)(((size_t)i * G.nth + j) * G.nph + k);
    const double* c = G.cells + (size_t)cell * kCellDoubles;
    fl.position_cm[0] = s[1] * P.length_scale;
    fl.position_cm[1] = s[2] * P.length_scale;
    fl.position_cm[2] = s[3] * P.length_scale;
    fl.coordinate_time_s = s[0] * P.time_scale;
    fl.density_g_cm3 = c[0];
    fl.temperature_k = c[1];
    // the four-velocity is read with the density and temperature, before
    // their check, so the record costs one memory round trip, not two
#pragma unroll
    for (int q = 0; q < 4; ++q) fl.coordinate_four_velocity_cm_s[q] = c[2 + q];
    if (!(fl.density_g_cm3 > 0.0) || !is_finite(fl.density_g_cm3) || !is_finite(fl.temperature_k))
        return false;
    g = ks_metric_at(s, radius, P.a, P.m);
    if (!fluid_tetrad<T>(g, fl.coordinate_four_velocity_cm_s, e, false)) return false;
    // The spatial legs of the tetrad serve the magnetic field and a
    // scattering event.  With scattering on, the FP64 form builds them here,
    // with every sampling lane; otherwise (and in the mixed form, where they
    // are FP32) maybe_scatter builds them for the lanes that scatter, and
    // where nothing reads fl.magnetic_gauss (the default transfer plug-ins)
    // the compiler drops the ones built for the field.
    const bool legs = full && scatter_legs_eager<T>() && P.scatter;
    if (legs) fluid_tetrad_axes<T>(g, e);
    if (full) {
        T em[16];
        if (legs) {
#pragma unroll
            for (int q = 0; q < 16; ++q) em[q] = e[q];
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) em[q] = e[q];
            fluid_tetrad_axes<T>(g, em);
        }
        const double b[4] = {c[6], c[7], c[8], c[9]};
        fl.magnetic_gauss[0] = -(double)inner<T>(g, em, b);
        fl.magnetic_gauss[1] = (double)inner<T>(g, em + 4, b);
        fl.magnetic_gauss[2] = (double)inner<T>(g, em + 8, b);
        fl.magnetic_gauss[3] = (double)inner<T>(g, em + 12, b);
    }
    on_sample(fl, e, g, cell);
    return true;
}

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).
// Replaces the Kokkos XorShift64 pool (scattering.hpp:16-17).
// ---------------------------------------------------------------------------

GRT_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Per-photon stream: counter (draw, stream, id_lo, id_hi), key = seed.
struct PhotonRng {
    uint64_t id;
    uint64_t seed;
    uint32_t stream;
    uint32_t draw;
    GRT_HD double drand()
    {
        uint32_t c[4] = {draw, stream, (uint32_t)id, (uint32_t)(id >> 32)};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        draw += 1u;
        const uint64_t mant = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
        return (double)mant * (1.0 / 9007199254740992.0);
    }
};

// ---------------------------------------------------------------------------
// Emission, scattering and camera
// ---------------------------------------------------------------------------

// Frequency of a thermal packet at temperature T (GenerateBlackbody,
// photon_generation.hpp:99-127): five draws, always.
GRT_HD double thermal_frequency(double T, PhotonRng& rng)
{
    const double inverse_zeta4 = 0.9239384029215902;
    const double choose = rng.drand();
    double cum = 0.0;
    int idx = 1;
    for (; idx < 64; ++idx) {
        const double n = (double)idx;
        cum += inverse_zeta4 / (n * n * n * n);
        if (choose <= cum) break;
    }
    const double d1 = rng.drand();
    const double d2 = rng.drand();
    const double d3 = rng.drand();
    const double d4 = rng.drand();
    const double prod = fmax(1.0e-300, d1 * d2 * d3 * d4);
    const double xx = -log(prod) / (double)idx;
    return xx * 2.0836619123327576e10 * T;
}
constexpr uint32_t kThermalDraws = 5;

// GenerateBlackbody / GeneratePowerLaw (photon_generation.hpp:99-143), the
// GeneratePhotonsCustom plug-in, selected as integrate_cart_ks.hpp:332-347.
GRT_HD bool generate_photon(const StepParams& P, const FluidCellCGS& cell, PhotonRng& rng,
                            int packets, double& nu, double& energy)
{
    if (P.generator == 0) {
        if (!(cell.temperature_k > 0.0)) return false;
        nu = thermal_frequency(cell.temperature_k, rng);
        energy = P.e_cell / (double)packets;
        return is_finite(nu) && nu > 0.0;
    }
    if (P.generator == 1) {
        const double draw = rng.drand();
        if (fabs(P.slope - 1.0) < 1.0e-12) {
            nu = P.nu_min * pow(P.nu_max / P.nu_min, draw);
        } else {
            const double ex = 1.0 - P.slope;
            const double lo = pow(P.nu_min, ex);
            const double hi = pow(P.nu_max, ex);
            nu = pow(lo + draw * (hi - lo), 1.0 / ex);
        }
        energy = P.e_cell / (double)packets;
        return is_finite(nu) && nu > 0.0;
    }
    return GeneratePhotonsCustom(cell, rng, packets, P.nu_custom, P.e_cell, nu, energy);
}

constexpr double kProtonMassCgs = 1.67262192369e-24;    // scattering.hpp:256
constexpr double kSigmaThomsonCgs = 6.6524587158e-25;   // scattering.hpp:257

// maybe_scatter_photon, scattering.hpp:259-313 (g = metric at the photon).
template <class T>
GRT_HD bool maybe_scatter(double s[8], double stokes[4], const StepParams& P, double dl_abs,
                          const FluidCellCGS& fl, const T e0[16], const KSMetric& g,
                          PhotonRng& rng)
{
    // e0 holds the fluid tetrad: all of it when sample_fluid built the
    // spatial legs (scatter_legs_eager), else e[0] only, and the legs are
    // built here after the scattering draw, from the same e[0] and metric,
    // so they are those of the sample's full tetrad.
    T e[16];
    constexpr int kGiven = scatter_legs_eager<T>() ? 16 : 4;
#pragma unroll
    for (int q = 0; q < kGiven; ++q) e[q] = e0[q];
    const double path = dl_abs * P.length_scale;
    const double ne = fl.density_g_cm3 / kProtonMassCgs;
    const double tau = kSigmaThomsonCgs * ne * path;
    const double prob = 1.0 - exp(-tau);
    if (rng.drand() >= prob) return false;
    double eloc = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) eloc -= s[4 + mu] * (double)e[mu];
    if (!(eloc > 0.0) || !is_finite(eloc)) return false;
    if (!scatter_legs_eager<T>()) fluid_tetrad_axes<T>(g, e);
    const double mu = 2.0 * rng.drand() - 1.0;
    const double dphi = kTwoPi * rng.drand();
    const double tr = sqrt(fmax(0.0, 1.0 - mu * mu));
    const double dx = tr * cos(dphi), dy = tr * sin(dphi), dz = mu;
    // fluid_frame_direction_to_covariant_momentum, fluid_frame.hpp:83-104:
    // p_mu = g_{mu nu} w^nu = eta_{mu nu} w^nu + twoH l_mu (l.w)
    double w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        w[q] = (double)e[q] + dx * (double)e[4 + q] + dy * (double)e[8 + q] + dz * (double)e[12 + q];
    const double lw = w[0] + g.l1 * w[1] + g.l2 * w[2] + g.l3 * w[3];
    s[4] = eloc * (-w[0] + g.twoH * lw);
    s[5] = eloc * (w[1] + g.twoH * g.l1 * lw);
    s[6] = eloc * (w[2] + g.twoH * g.l2 * lw);
    s[7] = eloc * (w[3] + g.twoH * g.l3 * lw);
    stokes[0] *= P.albedo;
    stokes[1] *= 0.5 * P.albedo;
    stokes[2] *= 0.5 * P.albedo;
    stokes[3] *= 0.25 * P.albedo;
    return true;
}

// camera_plane_crossing, integrate_cart_ks.hpp:96-120
GRT_HD int camera_crossing(const StepParams& P, const double s[8], double& u, double& v)
{
    const double plane = s[1] * P.cam_n[0] + s[2] * P.cam_n[1] + s[3] * P.cam_n[2];
    if (plane < P.cam_d) return 0;
    const double r0 = s[1] - P.cam_dn[0], r1 = s[2] - P.cam_dn[1], r2 = s[3] - P.cam_dn[2];
    u = r0 * P.cam_t[0] + r1 * P.cam_t[1] + r2 * P.cam_t[2];
    v = r0 * P.cam_f[0] + r1 * P.cam_f[1] + r2 * P.cam_f[2];
    return (fabs(u) <= P.cam_half_w && fabs(v) <= P.cam_half_h) ? 1 : -1;
}

// camera_observer_energy, integrate_cart_ks.hpp:122-130 (g_00 = 2H - 1);
// r is ks_radius of s's position.
GRT_HD double observer_energy(double a, double m, const double s[8], double r)
{
    const double g00 = 2.0 * ks_H(r, s[3], a, m) + (-1.0);
    if (!(g00 < 0.0)) return 0.0;
    return -s[4] * rsqrt_ieee(-g00);
}

// ---------------------------------------------------------------------------
// Per-photon state and the post-step part of the functor
// ---------------------------------------------------------------------------

// Observer of the per-call events the differentiable form needs: every
// radiative-transfer update (cell index, fluid, frequency, affine step and
// Stokes I before the update) and every event that sets Stokes I regardless
// of its previous value (emission event, zeroing).  NoHook compiles away.
struct NoHook {
    GRT_HD void rt(const FluidCellCGS&, uint32_t, double, double, double) {}
    GRT_HD void reset() {}
};

struct Photon {
    double s[8];
    double stokes[4];
    double freq, eem, dl;
    uint8_t phase, term;
    uint8_t pend;   // thermal frequency deferred: freq holds -T (see post_step)
};

// Deferred thermal frequency: the frequency of the packet emitted at
// temperature T = -ph.freq, drawn from the five draws reserved before
// rng.draw (nothing draws in between: deferral is off with scattering).
GRT_HD void resolve_frequency(Photon& ph, const PhotonRng& rng)
{
    PhotonRng r = rng;
    r.draw -= kThermalDraws;
    ph.freq = thermal_frequency(-ph.freq, r);
    ph.pend = 0;
}

// The camera arrival of a photon whose frequency is deferred, completed after
// the trace from its stored state (integrate_cart_ks.hpp:122-130, 392-400):
// the redshift factor eobs / eem that the immediate form multiplies into the
// frequency at arrival (returns true), or, where eobs or eem is not a
// positive finite number, false with fr = 1: the caller zeroes the Stokes
// vector.  It reads s[1..4] (position and p_t) and eem, which are those the
// arrival test saw, so the values are the immediate form's bit for bit.
GRT_HD bool deferred_arrival_factor(double a, double m, const Photon& ph, double& fr)
{
    const double r = ks_radius(ph.s[1], ph.s[2], ph.s[3], a);
    const double eobs = observer_energy(a, m, ph.s, r);
    const double eem = ph.eem;
    if (eobs > 0.0 && eem > 0.0 && is_finite(eobs) && is_finite(eem)) {
        fr = eobs / eem;
        return true;
    }
    fr = 1.0;
    return false;
}

// After the trace: the deferred frequency, times the arrival redshift fr
// (deferred_arrival_factor) for a photon that arrived.
GRT_HD void finish_deferred(Photon& ph, const PhotonRng& rng, double fr)
{
    resolve_frequency(ph, rng);
    if (ph.phase == kArrived) ph.freq *= fr;
}

// Between these temperatures a thermal frequency is always finite and
// positive (the sampled x lies in [7e-18, 691]), so a deferred emission
// succeeds exactly when the immediate one would.
constexpr double kDeferTmin = 1.0e-280;
constexpr double kDeferTmax = 1.0e290;

// Restores g^{mu nu} p_mu p_nu = 0 (mixed-precision form, after a scattering
// event) by scaling the spatial momentum (p_t, the conserved energy, is kept): one
// Newton step on f(c) = -p_t^2 + c^2 |p|^2 - 2H (-p_t + c l.p)^2 from c = 1.
GRT_HD void null_reproject(double s[8], const StepParams& P)
{
    const double r = ks_radius(s[1], s[2], s[3], P.a);
    double l[4];
    ks_l(s[1], s[2], s[3], r, P.a, l);
    const double H = ks_H(r, s[3], P.a, P.m);
    const double p0 = s[4];
    const double pp = s[5] * s[5] + s[6] * s[6] + s[7] * s[7];
    const double lq = l[1] * s[5] + l[2] * s[6] + l[3] * s[7];
    const double u = lq - p0;
    const double f = (pp - p0 * p0) - 2.0 * H * u * u;
    const double df = 2.0 * pp - 4.0 * H * u * lq;
    if (!(df != 0.0) || !is_finite(f) || !is_finite(df)) return;
    const double c = 1.0 - f / df;
    s[5] *= c;
    s[6] *= c;
    s[7] *= c;
}

// Everything after the RK step in Geodesic_cartesian_kerr_schild::operator()
// (integrate_cart_ks.hpp:308-404): fluid sampling, the image-mode emission
// event, radiative transfer, scattering and the camera test.  `phase0` is the
// phase the call started with; `taken` the affine step just integrated.
// On arrival at the camera (phase becomes kArrived) the screen coordinates
// are returned in (au, av).  Returns kSampleTried when the call looked for
// fluid (sample_scattering_fluid was called) plus kSampleHit when it found a
// cell with valid fluid and tetrad.
//
// kMixed (the mixed-precision form) builds the fluid tetrad and evaluates
// the attenuation factor exp(-dtau) in FP32; neither feeds back into the
// geodesic or its step-size sequence.  A photon scattered with the FP32
// tetrad is put back on the null cone in FP64 (null_reproject).
//
// kDefer: the image-mode emission of a thermal packet (generator 0, no
// scattering) only reserves its five draws and keeps -T in freq (pend set);
// the frequency is drawn when a transfer update first needs it, or after the
// launch (resolve_deferred), where every photon draws it at once instead of
// one or two lanes at a time.  The arrival of a deferred photon only sets
// its phase here: its redshift factor (and the zeroing of an invalid one)
// follows after the launch from the stored state (deferred_arrival_factor).
// Results are bit for bit those of the immediate emission.  (kDefer is used
// only with NoHook: the differentiable form does not defer.)
//
// The Kerr-Schild radius of the new position is computed once here, for the
// fluid sample and the observer energy, and returned in r_out: no event
// moves the photon, so it is also the radius the next call's opening tests
// (radius_exit) read.
//
// active false (a rejected RK45 attempt) makes the call change nothing but
// r_out (the radius of the unchanged position).  Rejected lanes run the
// post-step with the accepted ones instead of branching around it, so both
// leave it with the photon in the same registers (no copies at a merge).
constexpr uint32_t kSampleTried = 1u, kSampleHit = 2u;

template <int kSep, int kMixed = 0, class Hook = NoHook, int kDefer = 0>
GRT_HD uint32_t post_step(Photon& ph, uint8_t phase0, double taken, const GridView& G,
                      const StepParams& P, PhotonRng& rng, double& au, double& av,
                      double& r_out, Hook& hook, bool active)
{
    typedef typename std::conditional<kMixed != 0, float, double>::type T;
    const double radius = ks_radius(ph.s[1], ph.s[2], ph.s[3], P.a);
    r_out = radius;
    const bool image = (P.mode == 0);
    const bool needs =
        active && (P.scatter || P.absorb || P.emit || (image && phase0 == kBackward));
    // The events run inside the fluid sample's success block (a
    // continuation), so the tetrad never flows through a merge of the
    // sample's exit paths.
    // The image-mode emission event reads only e[0] (the emission-frame
    // energy) and, for the thermal and power-law generators, nothing of the
    // magnetic field: the rest of the tetrad is skipped there (every lane
    // emits once, so the wavefront would otherwise pay for it on most
    // iterations with one or two lanes active).
    const bool full = !(image && phase0 == kBackward) || P.generator == 2;
    const bool sampled = sample_fluid<kSep, T>(
        ph.s, radius, needs, G, P, full,
        [&](const FluidCellCGS& fl, const T* e, const KSMetric& g, uint32_t cell) {
            if (image && phase0 == kBackward) {
                double nu = 0.0, energy = 0.0;
                bool generated = false;
                const double dcode = fl.density_g_cm3 / P.density_scale;
                const int packets = ComputeSuperphotonsPerCell(dcode, P.norm);
                bool defer = false;
                if (packets > 0) {
                    if (kDefer && P.generator == 0 && !P.scatter && fl.temperature_k > kDeferTmin &&
                        fl.temperature_k < kDeferTmax) {
                        nu = -fl.temperature_k;
                        energy = P.e_cell / (double)packets;
                        rng.draw += kThermalDraws;
                        generated = defer = true;
                    } else {
                        generated = generate_photon(P, fl, rng, packets, nu, energy);
                    }
                }
                if (generated) {
                    ph.freq = nu;
                    ph.pend = defer ? 1 : 0;
                    double eem = 0.0;
        #pragma unroll
                    for (int q = 0; q < 4; ++q) eem -= ph.s[4 + q] * (double)e[q];
                    ph.eem = eem;
                    ph.stokes[0] = energy;
                    ph.stokes[1] = ph.stokes[2] = ph.stokes[3] = 0.0;
                    ph.dl = -fabs(ph.dl);
                    ph.phase = kForward;
                } else {
                    ph.stokes[0] = ph.stokes[1] = ph.stokes[2] = ph.stokes[3] = 0.0;
                    ph.term = 1;
                    ph.phase = kRejected;
                }
                hook.reset();
            } else if (phase0 != kBackward) {
                if (kDefer && ph.pend) resolve_frequency(ph, rng);
                if (P.absorb || P.emit) hook.rt(fl, cell, ph.freq, taken, ph.stokes[0]);
                // apply_radiative_transfer_step, radiative_transfer.hpp:81-96
                if (P.absorb) {
                    const double dtau = ComputeOpticalDepthIncrement(fl, ph.freq, taken, P.kabs);
                    const double att = kMixed ? (double)expf((float)-dtau) : exp(-dtau);
        #pragma unroll
                    for (int q = 0; q < 4; ++q) ph.stokes[q] *= att;
                }
                if (P.emit) ph.stokes[0] += ComputeEmissivityIncrement(fl, ph.freq, taken, P.kemit);
                if (P.scatter && maybe_scatter<T>(ph.s, ph.stokes, P, fabs(ph.dl), fl, e, g, rng) &&
                    kMixed)
                    null_reproject(ph.s, P);
            }
        });

    if (active && image && phase0 == kForward) {
        const int crossing = camera_crossing(P, ph.s, au, av);
        if (crossing > 0) {
            if (!(kDefer && ph.pend)) {
                const double eobs = observer_energy(P.a, P.m, ph.s, radius);
                const double eem = ph.eem;
                if (eobs > 0.0 && eem > 0.0 && is_finite(eobs) && is_finite(eem)) {
                    ph.freq *= eobs / eem;
                } else {
                    ph.stokes[0] = ph.stokes[1] = ph.stokes[2] = ph.stokes[3] = 0.0;
                    hook.reset();
                }
            }
            ph.phase = kArrived;
            ph.term = 1;
        } else if (crossing < 0) {
            ph.stokes[0] = ph.stokes[1] = ph.stokes[2] = ph.stokes[3] = 0.0;
            hook.reset();
            ph.phase = kRejected;
            ph.term = 1;
        }
    }
    return (needs ? kSampleTried : 0u) | (sampled ? kSampleHit : 0u);
}

// The radius tests that open every functor call (integrate_cart_ks.hpp:248-263);
// dist is ks_radius of the photon's position.  Returns true when the photon
// terminated.
template <class Hook>
GRT_HD bool radius_exit(Photon& ph, double dist, const StepParams& P, Hook& hook)
{
    if (dist <= P.r_min) {
        ph.stokes[0] = ph.stokes[1] = ph.stokes[2] = ph.stokes[3] = 0.0;
        hook.reset();
        ph.term = 1;
        ph.phase = kRejected;
        return true;
    }
    if (dist > P.r_max) {
        ph.term = 1;
        ph.phase = kRejected;
        return true;
    }
    return false;
}

// null_covariant_momentum_from_spatial_direction, kerr_schild_core.hpp:201-227:
// the future-directed K^0 making (K^0, K) null, lowered to p_mu.
GRT_HD bool null_momentum(const double X[4], const double K[3], double a, double m, double p[4])
{
    double g[16];
    ks_metric(X, a, m, g);
    const double qa = g[0];
    const double qb = 2.0 * (g[1] * K[0] + g[2] * K[1] + g[3] * K[2]);
    double qc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) qc += g[4 * (i + 1) + (j + 1)] * K[i] * K[j];
    const double disc = qb * qb - 4.0 * qa * qc;
    if (disc < 0.0) return false;
    const double sq = sqrt(disc);
    const double r1 = (-qb + sq) / (2.0 * qa);
    const double r2 = (-qb - sq) / (2.0 * qa);
    const double K0 = (r1 > 0.0) ? r1 : r2;
    if (!(K0 > 0.0)) return false;
    const double Kc[4] = {K0, K[0], K[1], K[2]};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0.0;
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) s += g[4 * mu + nu] * Kc[nu];
        p[mu] = s;
    }
    return true;
}

// Camera frame of initialize_photons_image (initialize_photons.hpp:94-108),
// evaluated on the host.
struct CameraInit {
    double c[3];          // plane centre d (sin th cos ph, sin th sin ph, cos th)
    double t[3], f[3];    // theta_hat, phi_hat
    double k[3];          // -c / |c|
    double dlambda0;      // initial_step * camera_distance
};

// One image-mode ray at screen (u, v), initialize_photons.hpp:123-166.
GRT_HD void init_image_ray(const CameraInit& C, double u, double v, double a, double m,
                           double s[8])
{
    const double px = C.c[0] + u * C.t[0] + v * C.f[0];
    const double py = C.c[1] + u * C.t[1] + v * C.f[1];
    const double pz = C.c[2] + u * C.t[2] + v * C.f[2];
    const double X[4] = {0.0, px, py, pz};
    double p[4] = {0.0, 0.0, 0.0, 0.0};
    null_momentum(X, C.k, a, m, p);
    s[0] = 0.0;
    s[1] = px;
    s[2] = py;
    s[3] = pz;
    s[4] = p[0];
    s[5] = p[1];
    s[6] = p[2];
    s[7] = p[3];
}

// Work counters of one lane.
struct LaneCounters {
    uint32_t calls, accepted, rejected, failed, samples, hits;
    GRT_HD void count_sample(uint32_t code)
    {
        samples += code & kSampleTried;
        hits += code >> 1;
    }
};

// One photon resident in one lane between refills.  r_open is ks_radius of
// the photon's position whenever no call is in progress: whoever places a
// photon in the lane sets it (lane_reset, or the photon stage, which
// computes it for 64 photons at once), and every call's post-step leaves the
// radius of the position it ends at.
struct Lane {
    Photon ph;
    double r_open;
    int attempt;
    bool in_call;
    uint8_t phase0;
};

GRT_HD void lane_reset_calls(Lane& L)
{
    L.in_call = false;
    L.attempt = 0;
    L.phase0 = 0;
}

GRT_HD void lane_reset(Lane& L, const StepParams& P)
{
    lane_reset_calls(L);
    L.r_open = ks_radius(L.ph.s[1], L.ph.s[2], L.ph.s[3], P.a);
}

// One iteration of a lane: at most one RK45 attempt (or one RK4 step) of the
// functor call in progress, opening a new call first if none is in progress.
// Geodesic_cartesian_kerr_schild::operator() (integrate_cart_ks.hpp:231-419)
// is thereby split at its retry loop, so lanes whose step is rejected retry
// while their neighbours go on.  Returns true when the functor call
// completed in this iteration; (au, av) are set when the photon arrived.
//
// k1 = f(s) is evaluated on every attempt, also on a retry, where s and so
// k1 are unchanged: in a wavefront some lane opens a call on nearly every
// iteration, so the evaluation costs the same issue slots either way, and
// no k1 is carried from one iteration to the next.  A rejected attempt runs
// the post-step too, as a no-op (post_step's `active`).
template <int kIntegrator, int kSep, class Rhs = RhsExact, int kMixed = 0, class Hook = NoHook,
          int kDefer = 0>
GRT_HD bool lane_iterate(Lane& L, const GridView& G, const StepParams& P, PhotonRng& rng,
                         LaneCounters& n, double& au, double& av, Hook& hook)
{
    Photon& ph = L.ph;
    if (!L.in_call) {
        L.phase0 = ph.phase;
        if (radius_exit(ph, L.r_open, P, hook)) return true;
        if (kIntegrator != 1) {
            // RK4, integrate_cart_ks.hpp:269-272
            const double taken = ph.dl;
            priority_stages();
            rk4_step<Rhs>(ph.s, taken, P);
            ++n.accepted;
            n.count_sample(post_step<kSep, kMixed, Hook, kDefer>(ph, L.phase0, taken, G, P, rng, au,
                                                                  av, L.r_open, hook, true));
            return true;
        }
        L.attempt = 0;
        L.in_call = true;
    }
    // One attempt of the retry loop, integrate_cart_ks.hpp:273-299.
    double k1[8];
    priority_stages();
    Rhs::eval(ph.s, P, k1);
    const double taken = ph.dl;
    const bool accepted = rk45_attempt<Rhs, kMixed>(ph.s, &ph.dl, k1, P);
    ++L.attempt;
    const uint32_t code = post_step<kSep, kMixed, Hook, kDefer>(ph, L.phase0, taken, G, P, rng, au,
                                                                av, L.r_open, hook, accepted);
    if (accepted) {
        ++n.accepted;
        L.in_call = false;
        n.count_sample(code);
        return true;
    }
    ++n.rejected;
    if (L.attempt == 12) {
        ++n.failed;
        ph.term = 1;
        L.in_call = false;
        return true;
    }
    return false;
}

template <int kIntegrator, int kSep, class Rhs = RhsExact, int kMixed = 0, int kDefer = 0>
GRT_HD bool lane_iterate(Lane& L, const GridView& G, const StepParams& P, PhotonRng& rng,
                         LaneCounters& n, double& au, double& av)
{
    NoHook h;
    return lane_iterate<kIntegrator, kSep, Rhs, kMixed, NoHook, kDefer>(L, G, P, rng, n, au, av,
                                                                        h);
}

}  // namespace grt
