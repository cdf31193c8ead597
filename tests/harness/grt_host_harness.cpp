// grt_host_harness.cpp -- TEST INFRASTRUCTURE: the tracing kernel's per-lane
// logic (grt_device.hpp: lane_iterate, the separable cell search, ...) and
// the library's host-side grid preparation (grt_host_prep.hpp), compiled for
// the CPU so their results can be compared with the oracle without a GPU.
// Photons run one after another to termination or max_steps calls, which is
// the reference loop's outcome for termination_percent = 1.
#include <cmath>
#include <cstring>
#include <vector>

#include "grt.h"
#include "grt_device.hpp"
#include "grt_host_prep.hpp"

using namespace grt;

namespace {

struct Prepared {
    HostGrid hg;
    std::vector<double> coords;
    GridView view;
};

void prepare(const grt_params* p, int nr, int nth, int nph, const double* r, const double* th,
             const double* ph, const double* rho, const double* tg, const double* vel,
             const double* mag, int force_general, Prepared& out)
{
    const bool avail = nr > 0 && r != nullptr;
    if (avail) {
        prepare_grid(*p, nr, nth, nph, r, th, ph, rho, tg, vel, mag, out.hg);
        if (force_general) out.hg.separable = false;
        if (!out.hg.separable) {
            const size_t G = (size_t)nr * nth * nph;
            out.coords.resize(3 * G);
            std::memcpy(out.coords.data(), r, G * 8);
            std::memcpy(out.coords.data() + G, th, G * 8);
            std::memcpy(out.coords.data() + 2 * G, ph, G * 8);
        }
    }
    const double win[2] = {avail && out.hg.separable ? out.hg.tables[2 * (size_t)nr] : 0.0,
                           avail && out.hg.separable ? out.hg.tables[4 * (size_t)nr - 1] : 0.0};
    out.view = grid_view(nr, nth, nph, avail, avail && out.hg.separable, out.hg.tables.data(),
                         out.hg.axes.data(), out.coords.data(), out.hg.cells.data(), out.hg.key,
                         win);
}

}  // namespace

extern "C" {

int harness_locate(const grt_params* p, int nr, int nth, int nph, const double* r,
                   const double* th, const double* ph, const double* rho, const double* tg,
                   const double* vel, const double* mag, int force_general, int64_t nq,
                   const double* radius, const double* polar, const double* az, int32_t* out)
{
    Prepared pr;
    prepare(p, nr, nth, nph, r, th, ph, rho, tg, vel, mag, force_general, pr);
    for (int64_t q = 0; q < nq; ++q) {
        int i = -1, j = -1, k = -1;
        // the kernel derives (cos theta, X, Y) from the position; here they
        // come from the query's (r, theta, phi) with a = 0: X, Y = cos, sin phi
        const bool ok = pr.view.separable
                            ? locate_separable(pr.view, radius[q], std::cos(polar[q]),
                                               std::cos(az[q]), std::sin(az[q]), i, j, k)
                            : locate_general(pr.view, radius[q], polar[q], az[q], i, j, k);
        out[4 * q + 0] = ok ? 1 : 0;
        out[4 * q + 1] = i;
        out[4 * q + 2] = j;
        out[4 * q + 3] = k;
    }
    return pr.view.separable ? 1 : 0;
}

// The separable search on raw kernel inputs (radius, cos theta, X, Y), with
// the bucket index (no_buckets = 0) or with plain binary searches over the
// whole axes (no_buckets = 1).  Returns the number of buckets per axis in
// nb_out[3] (0: the grid took no buckets).
int harness_locate_raw(const grt_params* p, int nr, int nth, int nph, const double* r,
                       const double* th, const double* ph, const double* rho, const double* tg,
                       const double* vel, const double* mag, int no_buckets, int64_t nq,
                       const double* radius, const double* ct, const double* X, const double* Y,
                       int32_t* out, int32_t* nb_out)
{
    Prepared pr;
    prepare(p, nr, nth, nph, r, th, ph, rho, tg, vel, mag, 0, pr);
    for (int a = 0; a < 3; ++a) nb_out[a] = pr.view.key_nb[a];
    if (!pr.view.separable) return 0;
    if (no_buckets)
        for (int a = 0; a < 3; ++a) pr.view.key_nb[a] = 0;
    for (int64_t q = 0; q < nq; ++q) {
        int i = -1, j = -1, k = -1;
        const bool ok = locate_separable(pr.view, radius[q], ct[q], X[q], Y[q], i, j, k);
        out[4 * q + 0] = ok ? 1 : 0;
        out[4 * q + 1] = i;
        out[4 * q + 2] = j;
        out[4 * q + 3] = k;
    }
    return 1;
}

// Per axis (0 radial, 1 polar, 2 azimuth) and query, the plain count over
// the axis table and the bucketed count: out[2 q] and out[2 q + 1].  The
// queries are the kernel's compared values (r, cos theta, pseudo-angle).
int harness_axis_counts(const grt_params* p, int nr, int nth, int nph, const double* r,
                        const double* th, const double* ph, const double* rho, const double* tg,
                        const double* vel, const double* mag, int axis, int64_t nq,
                        const double* v, int32_t* out)
{
    Prepared pr;
    prepare(p, nr, nth, nph, r, th, ph, rho, tg, vel, mag, 0, pr);
    const GridView& G = pr.view;
    if (!G.separable || G.key_nb[axis] == 0) return 0;
    const double* R = G.tab;
    const double* T = R + 4 * nr;
    const double* F = T + 4 * nth;
    const uint16_t* B = reinterpret_cast<const uint16_t*>(F + 4 * nph);
    for (int a = 0; a < axis; ++a) B += G.key_nb[a] + 1;
    for (int64_t q = 0; q < nq; ++q) {
        const int sh = G.key_shift[axis], bs = G.key_base[axis], nb = G.key_nb[axis];
        if (axis == 0) {
            out[2 * q] = count_le(R, nr, v[q]);
            out[2 * q + 1] = bucket_count_le(R, B, sh, bs, nb, v[q], v[q]);
        } else if (axis == 1) {
            out[2 * q] = count_ge_desc(T, nth, v[q]);
            out[2 * q + 1] = bucket_count_ge_desc(T, B, sh, bs, nb, 3.0 - v[q], v[q]);
        } else {
            out[2 * q] = count_le(F, nph, v[q]);
            out[2 * q + 1] = bucket_count_le(F, B, sh, bs, nb, v[q] + 4.0, v[q]);
        }
    }
    return 1;
}

int harness_trace(const grt_params* p, int nr, int nth, int nph, const double* r,
                  const double* th, const double* ph, const double* rho, const double* tg,
                  const double* vel, const double* mag, int force_general, grt_photons* P,
                  uint64_t n, int max_steps, uint64_t* stats)
{
    Prepared pr;
    prepare(p, nr, nth, nph, r, th, ph, rho, tg, vel, mag, force_general, pr);
    StepParams sp = make_step_params(*p);
    const bool fluid = pr.view.available;
    sp.absorb = sp.absorb && fluid;
    sp.emit = sp.emit && fluid;
    sp.scatter = sp.scatter && fluid;
    LaneCounters cnt{0, 0, 0, 0, 0, 0};
    for (int s = 0; s < 5; ++s) stats[s] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (P->terminate[i]) continue;
        Lane L;
        Photon& f = L.ph;
        f.s[0] = P->x0[i]; f.s[1] = P->x1[i]; f.s[2] = P->x2[i]; f.s[3] = P->x3[i];
        f.s[4] = P->k0[i]; f.s[5] = P->k1[i]; f.s[6] = P->k2[i]; f.s[7] = P->k3[i];
        f.stokes[0] = P->I[i]; f.stokes[1] = P->Q[i]; f.stokes[2] = P->U[i]; f.stokes[3] = P->V[i];
        f.freq = P->frequency[i];
        f.eem = P->emission_frame_energy[i];
        f.dl = P->dlambda[i];
        f.phase = P->phase[i];
        f.term = 0;
        f.pend = 0;
        lane_reset(L, sp);
        PhotonRng rng{P->id[i], sp.seed, sp.stream << 1, 0u};
        int calls = 0;
        double au = 0.0, av = 0.0;
        // with the deferred thermal emission of the tracing kernel
        while (calls < max_steps && !f.term) {
            au = av = 0.0;
            bool done;
            const bool sep = pr.view.separable || !pr.view.available;
            if (sp.integrator == GRT_RK4)
                done = sep ? lane_iterate<GRT_RK4, 1, RhsExact, 0, 1>(L, pr.view, sp, rng, cnt, au, av)
                           : lane_iterate<GRT_RK4, 0, RhsExact, 0, 1>(L, pr.view, sp, rng, cnt, au, av);
            else
                done = sep ? lane_iterate<GRT_RK45, 1, RhsExact, 0, 1>(L, pr.view, sp, rng, cnt, au, av)
                           : lane_iterate<GRT_RK45, 0, RhsExact, 0, 1>(L, pr.view, sp, rng, cnt, au, av);
            if (done) { ++calls; ++cnt.calls; }
        }
        if (f.pend) {
            // as resolve_deferred completes it after a tracing launch
            double fr = 1.0;
            if (f.phase == kArrived && !deferred_arrival_factor(sp.a, sp.m, f, fr))
                f.stokes[0] = f.stokes[1] = f.stokes[2] = f.stokes[3] = 0.0;
            finish_deferred(f, rng, fr);
        }
        P->x0[i] = f.s[0]; P->x1[i] = f.s[1]; P->x2[i] = f.s[2]; P->x3[i] = f.s[3];
        P->k0[i] = f.s[4]; P->k1[i] = f.s[5]; P->k2[i] = f.s[6]; P->k3[i] = f.s[7];
        P->I[i] = f.stokes[0]; P->Q[i] = f.stokes[1]; P->U[i] = f.stokes[2]; P->V[i] = f.stokes[3];
        P->frequency[i] = f.freq;
        P->emission_frame_energy[i] = f.eem;
        P->dlambda[i] = f.dl;
        P->phase[i] = f.phase;
        P->terminate[i] = f.term;
        if (f.phase == kArrived) {
            P->theta_disp[i] = au;
            P->phi_disp[i] = av;
        }
    }
    stats[0] = cnt.calls;
    stats[1] = cnt.accepted;
    stats[2] = cnt.rejected;
    stats[3] = cnt.failed;
    stats[4] = cnt.samples;
    stats[5] = cnt.hits;
    return 0;
}

// Both right-hand-side evaluations of the device code on n states (8 each).
void harness_rhs(const double* states, int n, double a, double m, double* exact, double* fast)
{
    for (int q = 0; q < n; ++q) {
        grt::ks_rhs(states + 8 * q, a, m, exact + 8 * q);
        grt::ks_rhs_fast(states + 8 * q, a, m, fast + 8 * q);
        exact[8 * q + 4] = fast[8 * q + 4] = 0.0;
    }
}

}  // extern "C"
