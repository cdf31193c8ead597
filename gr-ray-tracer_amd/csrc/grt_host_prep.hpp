// grt_host_prep.hpp -- host-side preparation of what the tracing kernel
// consumes: the functor constants and the fluid-grid tables / cell records.
// Pure C++ (no HIP runtime calls); shared by the library and the CPU test
// harness.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "grt.h"
#include "grt_device.hpp"

namespace grt {

// Functor constants (integrate_cart_ks.hpp:471-476) with the camera basis of
// integrate_cart_ks.hpp:98-107 evaluated once here.
inline StepParams make_step_params(const grt_params& p)
{
    StepParams s{};
    s.a = p.a_bh;
    s.m = p.m_bh;
    s.r_min = p.r_term_min;
    s.r_max = p.r_term_max;
    s.atol = p.atol;
    s.rtol = p.rtol;
    s.smin = p.min_step_scale;
    s.smax = p.max_step_scale;
    s.safety = p.safety;
    s.integrator = p.integrator;
    s.mode = p.mode;
    s.absorb = p.absorption_enabled ? 1 : 0;
    s.emit = p.emission_enabled ? 1 : 0;
    s.scatter = p.scattering_enabled ? 1 : 0;
    s.kabs = p.absorption_coefficient;
    s.kemit = p.emission_coefficient;
    s.albedo = p.albedo;
    s.seed = p.seed;
    s.stream = p.rng_stream;
    s.generator = p.generator;
    s.norm = p.superphoton_count_normalization;
    s.e_cell = p.energy_per_cell_erg;
    s.slope = p.power_law_slope;
    s.nu_min = p.nu_min_hz;
    s.nu_max = p.nu_max_hz;
    s.nu_custom = p.custom_frequency_hz;
    const double th = p.camera_theta, ph = p.camera_phi, d = p.camera_distance;
    s.cam_d = d;
    s.cam_half_w = 0.5 * p.camera_width;
    s.cam_half_h = 0.5 * p.camera_height;
    s.cam_n[0] = std::sin(th) * std::cos(ph);
    s.cam_n[1] = std::sin(th) * std::sin(ph);
    s.cam_n[2] = std::cos(th);
    s.cam_t[0] = std::cos(th) * std::cos(ph);
    s.cam_t[1] = std::cos(th) * std::sin(ph);
    s.cam_t[2] = -std::sin(th);
    s.cam_f[0] = -std::sin(ph);
    s.cam_f[1] = std::cos(ph);
    s.cam_f[2] = 0.0;
    for (int q = 0; q < 3; ++q) s.cam_dn[q] = d * s.cam_n[q];
    s.length_scale = p.length_cm_per_code;
    s.time_scale = p.time_s_per_code;
    s.density_scale = p.density_g_cm3_per_code;
    return s;
}

template <typename F>
void host_parallel_for(size_t n, F&& f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536) nt = 1;
    std::vector<std::thread> th;
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const size_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([&f, lo, hi] {
            for (size_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto& x : th) x.join();
}

// Everything the kernel reads about the grid, still in host memory.
struct HostGrid {
    int nr = 0, nth = 0, nph = 0;
    bool separable = false;
    std::vector<double> tables;   // separable search tables, see locate_separable
    std::vector<double> axes;     // r_i | theta_j | phi_k of a separable grid
    std::vector<double> cells;    // kCellDoubles per cell
    int key[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // key_shift[3], key_base[3], key_nb[3]
};

// Buckets per search axis at most (2-byte counts, appended to the tables).
constexpr int kMaxBuckets = 1024;

// Bucket counts of one search axis (axis_bucket on the device): `w` are the
// keyed entries, non-decreasing; picks the finest key (most mantissa bits)
// with at most kMaxBuckets buckets and appends, per bucket b = 0..nb, the
// number of entries keyed below b.  Returns false (no buckets) when the
// entries are not positive, finite and sorted.
inline bool build_axis_buckets(const std::vector<double>& w, int& shift, int& base, int& nb,
                               std::vector<uint16_t>& out)
{
    nb = 0;
    const size_t n = w.size();
    if (n == 0 || n > 65535) return false;
    for (size_t j = 0; j < n; ++j)
        if (!(w[j] > 0.0) || !std::isfinite(w[j]) || (j > 0 && !(w[j] >= w[j - 1]))) return false;
    auto bits = [](double v) {
        uint64_t u;
        std::memcpy(&u, &v, sizeof u);
        return u;
    };
    const uint64_t lo_bits = bits(w.front()), hi_bits = bits(w.back());
    for (int mant = 20; mant >= 0; --mant) {
        const int sh = 52 - mant;
        const uint64_t span = (hi_bits >> sh) - (lo_bits >> sh) + 1;
        if (span <= (uint64_t)kMaxBuckets) {
            shift = sh;
            base = (int)(lo_bits >> sh);
            nb = (int)span;
            break;
        }
    }
    if (nb == 0) return false;
    std::vector<int> key(n);
    for (size_t j = 0; j < n; ++j) {
        key[j] = axis_bucket(w[j], shift, base, nb);
        if (j > 0 && key[j] < key[j - 1]) {
            nb = 0;
            return false;
        }
    }
    size_t j = 0;
    for (int b = 0; b <= nb; ++b) {
        while (j < n && key[j] < b) ++j;
        out.push_back((uint16_t)j);
    }
    return true;
}

// Largest separable table set staged in LDS (bytes).
constexpr size_t kMaxTableBytes = 48 * 1024;

// The diamond pseudo-angle of (cos phi, sin phi), as pseudo_angle() on device.
inline double host_pseudo_angle(double X, double Y)
{
    if (Y >= 0.0) {
        if (X >= 0.0) return (X + Y > 0.0) ? Y / (X + Y) : 0.0;
        return 1.0 + (-X) / (Y - X);
    }
    if (X < 0.0) return 2.0 + (-Y) / (-X - Y);
    return 3.0 + X / (X - Y);
}

// Builds the separable search tables of locate_separable from the 1-D axes;
// returns false if the axes do not admit the fast path.
inline bool build_search_tables(const std::vector<double>& ra, const std::vector<double>& ta,
                                const std::vector<double>& pa, std::vector<double>& tab,
                                int key[9])
{
    const int nr = (int)ra.size(), nth = (int)ta.size(), nph = (int)pa.size();
    const double pi = 3.14159265358979323846, two_pi = 2.0 * pi;
    for (int i = 0; i < nr; ++i)
        if (!(ra[i] > 0.0) || !std::isfinite(ra[i]) || (i > 0 && !(ra[i] > ra[i - 1]))) return false;
    for (int j = 0; j < nth; ++j)
        if (!(ta[j] >= 0.0) || !(ta[j] <= pi) || (j > 0 && !(ta[j] > ta[j - 1]))) return false;
    for (int k = 0; k < nph; ++k)
        if (!(pa[k] >= 0.0) || !(pa[k] < two_pi) || (k > 0 && !(pa[k] > pa[k - 1]))) return false;
    if ((size_t)4 * (nr + nth + nph) * sizeof(double) > kMaxTableBytes) return false;
    tab.assign(4 * (size_t)(nr + nth + nph), 0.0);
    const double inf = 1.0 / 0.0;
    // radial: axis, geometric midpoints, e^-+thr window (scattering.hpp:99-117)
    double* R = tab.data();
    for (int i = 0; i < nr; ++i) {
        double span = 0.0;
        if (i > 0) span = std::fmax(span, std::fabs(std::log(ra[i] / ra[i - 1])));
        if (i + 1 < nr) span = std::fmax(span, std::fabs(std::log(ra[i] / ra[i + 1])));
        const double thr = nr == 1 ? 1.0e-10 : 0.500001 * span;
        R[i] = ra[i];
        R[nr + i] = i + 1 < nr ? std::sqrt(ra[i] * ra[i + 1]) : inf;
        R[2 * nr + i] = ra[i] * std::exp(-thr);
        R[3 * nr + i] = ra[i] * std::exp(thr);
    }
    // polar in cos space (scattering.hpp:119-133)
    double* T = R + 4 * nr;
    for (int j = 0; j < nth; ++j) {
        double span = 0.0;
        if (j > 0) span = std::fmax(span, std::fabs(ta[j] - ta[j - 1]));
        if (j + 1 < nth) span = std::fmax(span, std::fabs(ta[j] - ta[j + 1]));
        const double thr = nth == 1 ? 1.0e-10 : 0.500001 * span;
        T[j] = std::cos(ta[j]);
        T[nth + j] = j + 1 < nth ? std::cos(0.5 * (ta[j] + ta[j + 1])) : -inf;
        T[2 * nth + j] = std::cos(std::fmin(pi, ta[j] + thr));
        T[3 * nth + j] = std::cos(std::fmax(0.0, ta[j] - thr));
    }
    for (int j = 1; j < nth; ++j)
        if (!(T[j] < T[j - 1])) return false;
    // azimuth: pseudo-angle, cos, sin, cos(window) (scattering.hpp:135-149)
    double* F = T + 4 * nth;
    auto per = [&](double a, double b) {
        const double d = std::fmod(std::fabs(a - b), two_pi);
        return std::fmin(d, two_pi - d);
    };
    for (int k = 0; k < nph; ++k) {
        double span = 0.0;
        if (k > 0) span = std::fmax(span, per(pa[k], pa[k - 1]));
        if (k + 1 < nph) span = std::fmax(span, per(pa[k], pa[k + 1]));
        const double c = std::cos(pa[k]), sn = std::sin(pa[k]);
        F[k] = host_pseudo_angle(c, sn);
        F[nph + k] = c;
        F[2 * nph + k] = sn;
        F[3 * nph + k] = nph == 1 ? -2.0 : std::cos(std::fmin(pi, 0.500001 * span));
    }
    for (int k = 1; k < nph; ++k)
        if (!(F[k] > F[k - 1])) return false;
    // bucket counts: radial on r, polar on 3 - cos theta (ascending), azimuth
    // on pseudo-angle + 4; the device keys its query the same way
    std::vector<double> wr(ra), wt(nth), wp(nph);
    for (int j = 0; j < nth; ++j) wt[j] = 3.0 - T[j];
    for (int k = 0; k < nph; ++k) wp[k] = F[k] + 4.0;
    std::vector<uint16_t> counts;
    const std::vector<double>* w[3] = {&wr, &wt, &wp};
    for (int a = 0; a < 3; ++a) {
        if (!build_axis_buckets(*w[a], key[a], key[3 + a], key[6 + a], counts)) {
            for (int q = 0; q < 9; ++q) key[q] = 0;
            counts.clear();
            break;
        }
    }
    const size_t head = tab.size();
    tab.resize(head + (counts.size() * sizeof(uint16_t) + sizeof(double) - 1) / sizeof(double), 0.0);
    if (!counts.empty())
        std::memcpy(tab.data() + head, counts.data(), counts.size() * sizeof(uint16_t));
    return true;
}

inline void prepare_grid(const grt_params& p, int nr, int nth, int nph, const double* r,
                         const double* theta, const double* phi, const double* rho,
                         const double* tgas, const double* vel, const double* mag, HostGrid& out)
{
    const size_t G = (size_t)nr * nth * nph;
    out.nr = nr;
    out.nth = nth;
    out.nph = nph;
    auto at = [&](const double* a, int i, int j, int k) {
        return a[((size_t)i * nth + j) * nph + k];
    };
    // Separable axes: r(i,j,k) = r(i), theta = theta(j), phi = phi(k).
    bool sep = true;
    for (size_t c = 0; c < G && sep; ++c) {
        const int i = (int)(c / ((size_t)nth * nph));
        const int j = (int)((c / nph) % nth);
        const int k = (int)(c % nph);
        if (r[c] != at(r, i, 0, 0) || theta[c] != at(theta, 0, j, 0) || phi[c] != at(phi, 0, 0, k))
            sep = false;
    }
    std::vector<double> ra(nr), ta(nth), pa(nph);
    for (int i = 0; i < nr; ++i) ra[i] = at(r, i, 0, 0);
    for (int j = 0; j < nth; ++j) ta[j] = at(theta, 0, j, 0);
    for (int k = 0; k < nph; ++k) pa[k] = at(phi, 0, 0, k);
    out.tables.clear();
    out.axes.clear();
    for (int q = 0; q < 9; ++q) out.key[q] = 0;
    if (sep) sep = build_search_tables(ra, ta, pa, out.tables, out.key);
    out.separable = sep;
    if (sep) {
        out.axes = ra;
        out.axes.insert(out.axes.end(), ta.begin(), ta.end());
        out.axes.insert(out.axes.end(), pa.begin(), pa.end());
    }
    // Cell records: scaled density and temperature, and the Cartesian
    // Kerr-Schild four-velocity / magnetic vector of each cell
    // (transform_spherical_four_vector_to_cartesian, fluid_frame.hpp:29-49,
    // evaluated at the cell's own r, theta, phi as scattering.hpp:223-241 does).
    out.cells.assign(G * kCellDoubles, 0.0);
    const double spin = p.a_bh;
    const double ds = p.density_g_cm3_per_code, ts = p.temperature_k_per_code;
    const double vs = p.four_velocity_cm_s_per_code, ms = p.magnetic_gauss_per_code;
    double* cells = out.cells.data();
    host_parallel_for(G, [&](size_t c) {
        const double rc = r[c], thc = theta[c], phc = phi[c];
        const double st = std::sin(thc), ct = std::cos(thc), sp = std::sin(phc), cp = std::cos(phc);
        const double rx = cp * st, ry = sp * st, rz = ct;
        const double tx = (rc * cp - spin * sp) * ct;
        const double ty = (rc * sp + spin * cp) * ct;
        const double tz = -rc * st;
        const double px = (-rc * sp - spin * cp) * st;
        const double py = (rc * cp - spin * sp) * st;
        double* o = cells + c * kCellDoubles;
        const double* v = vel + 4 * c;
        const double* b = mag + 4 * c;
        o[0] = rho[c] * ds;
        o[1] = tgas[c] * ts;
        o[2] = v[0] * vs;
        o[3] = (rx * v[1] + tx * v[2] + px * v[3]) * vs;
        o[4] = (ry * v[1] + ty * v[2] + py * v[3]) * vs;
        o[5] = (rz * v[1] + tz * v[2]) * vs;
        o[6] = b[0] * ms;
        o[7] = (rx * b[1] + tx * b[2] + px * b[3]) * ms;
        o[8] = (ry * b[1] + ty * b[2] + py * b[3]) * ms;
        o[9] = (rz * b[1] + tz * b[2]) * ms;
    });
}

// GridView over host (or device) copies of HostGrid's arrays.
// win: host copies of the radial window bounds (GridView::win_lo, win_hi).
inline GridView grid_view(int nr, int nth, int nph, bool available, bool separable,
                          const double* tables, const double* axes, const double* coords3,
                          const double* cells, const int* key, const double* win)
{
    GridView g{};
    g.win_lo = win[0];
    g.win_hi = win[1];
    for (int a = 0; a < 3; ++a) {
        g.key_shift[a] = key[a];
        g.key_base[a] = key[3 + a];
        g.key_nb[a] = (available && separable) ? key[6 + a] : 0;
    }
    g.nr = nr;
    g.nth = nth;
    g.nph = nph;
    g.available = available ? 1 : 0;
    g.separable = separable ? 1 : 0;
    if (available && separable) {
        g.tab = tables;
        g.axes = axes;
    }
    if (available && !separable) {
        const size_t G = (size_t)nr * nth * nph;
        g.r3 = coords3;
        g.th3 = coords3 + G;
        g.ph3 = coords3 + 2 * G;
    }
    g.cells = cells;
    return g;
}

}  // namespace grt
