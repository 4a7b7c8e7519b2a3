// completion.cpp -- host side of the spectral completion criterion
// (CompletionSensor in FT mode, sensors/s_completion.cpp): where the watched
// cells go and at which wavelengths, and the completion test on the ring of
// running-DFT values the device keeps (k_ft_feed).  Both run once per run or
// once per check (every cc_step steps), so they stay on the host.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>

namespace aefdtd {

namespace {

// randp(A, B) of common/mathUT.cpp:65-68 on a std::mt19937.
double randp(std::mt19937 &mte, double A, double B) {
    return A + (B - A) * mte() / (mte.max() - mte.min());
}

std::string known_layout(const char layout[4]) {
    std::string l(layout, strnlen(layout, 4));
    if (l == "v") return l;
    const std::string ok = "ndub";
    bool known = l.size() == 3;
    for (char c : l) known = known && ok.find(c) != std::string::npos;
    if (!known || l == "nnn") return "nnb";
    return l;
}

}  // namespace

// CompletionSensor::initialize (s_completion.cpp:232-398): cc_quant cells on
// the structure region's faces (or in its volume for "v"), shared among the
// faces in proportion to their areas, positions and wavelengths drawn from the
// generator in the reference's order; navg = (longest wavelength / c) / Dt.
int ft_place(const aefdtd_grid *g, double Dt, double lmin, double lmax, int Np, const char layout_in[4],
             unsigned seed, std::vector<int> &pts, std::vector<double> &w, int &navg) {
    if (Np < 1 || !(lmin > 0) || !(lmax > 0) || !(Dt > 0)) return AEFDTD_EINVAL;
    std::mt19937 mte(seed);
    const std::string layout = known_layout(layout_in);
    std::vector<int> il(Np), jl(Np), kl(Np);
    const int Nx = g->Nx, Ny = g->Ny, Nz = g->Nz;
    if (layout == "v") {
        for (int p = 0; p < Np; p++) {
            il[p] = static_cast<int>(randp(mte, g->xs_s, g->xs_e - 1) + 0.5);
            jl[p] = static_cast<int>(randp(mte, g->ys_s, g->ys_e - 1) + 0.5);
            kl[p] = static_cast<int>(randp(mte, g->zs_s, g->zs_e - 1) + 0.5);
        }
    } else {
        int n[6] = {0, 0, 0, 0, 0, 0};                      // xd xu yd yu zd zu
        const int area[3] = {Ny * Nz, Nx * Nz, Nx * Ny};
        for (int a = 0; a < 3; a++) {
            const char c = layout[a];
            if (c == 'd' || c == 'b') n[2 * a] = area[a];
            if (c == 'u' || c == 'b') n[2 * a + 1] = area[a];
        }
        double tot = 0;
        for (int f = 0; f < 6; f++) tot += n[f];
        int left = Np;
        for (int f = 0; f < 6; f++) {
            int alloc = static_cast<int>(Np * n[f] / tot + 0.5);
            alloc = alloc < left ? alloc : left;
            left -= alloc;
            n[f] = alloc;
        }
        int off = 0;
        for (int f = 0; f < 6; f++) {
            const int a = f / 2;
            const bool hi = f & 1;
            for (int p = off; p < off + n[f]; p++) {
                // the face's own coordinate is fixed; the other two are drawn in x, y, z order
                il[p] = a == 0 ? (hi ? g->xs_e - 1 : g->xs_s) : static_cast<int>(randp(mte, g->xs_s, g->xs_e - 1) + 0.5);
                jl[p] = a == 1 ? (hi ? g->ys_e - 1 : g->ys_s) : static_cast<int>(randp(mte, g->ys_s, g->ys_e - 1) + 0.5);
                kl[p] = a == 2 ? (hi ? g->zs_e - 1 : g->zs_s) : static_cast<int>(randp(mte, g->zs_s, g->zs_e - 1) + 0.5);
            }
            off += n[f];
        }
        // cells a rounding shortfall leaves unplaced stay at (0, 0, 0), as the
        // reference's zero-initialised location grids do
    }
    pts.resize(3 * (size_t)Np);
    w.resize(Np);
    for (int p = 0; p < Np; p++) {
        pts[3 * p] = il[p];
        pts[3 * p + 1] = jl[p];
        pts[3 * p + 2] = kl[p];
    }
    for (int p = 0; p < Np; p++) {
        const double lam = randp(mte, lmin, lmax);
        w[p] = 2.0 * kPi * kC / lam;                       // m_to_rad_Hz (common/phys_tools.cpp:159)
    }
    const double T = std::max(lmin, lmax) / kC;
    navg = static_cast<int>(T / Dt);
    return navg >= 1 ? AEFDTD_OK : AEFDTD_EINVAL;
}

// CompletionSensor::completion_check in FT mode (s_completion.cpp:39-121) on
// the ring [6][np][navg] (Ex Ey Ez real parts, then imaginary parts), at
// sensor step count `steps`.  As in the reference, the six averages carry over
// from one cell to the next (they are set to |sum / navg| and not cleared), a
// cell is settled when the span of its largest-average series over the ring
// is at most coeff times that average, and the run is complete when 90% of the
// cells are settled.
bool ft_complete(const std::vector<double> &ring, int np, int navg, double coeff, int steps, int nmax) {
    if (steps < 2 * nmax) return false;
    double avg[6] = {0, 0, 0, 0, 0, 0};
    int nok = 0;
    auto R = [&](int c, int p, int n) { return ring[((size_t)c * np + p) * navg + n]; };
    for (int p = 0; p < np; p++) {
        for (int n = 0; n < navg; n++)
            for (int c = 0; c < 6; c++) avg[c] += R(c, p, n);
        for (int c = 0; c < 6; c++) avg[c] = std::abs(avg[c] / (navg + 0.0));
        double amax = avg[0];
        for (int c = 1; c < 6; c++) amax = std::max(amax, avg[c]);
        int tg = 0;
        while (tg < 5 && !(amax == avg[tg])) tg++;
        double tmin = R(tg, p, 0), tmax = R(tg, p, 0);
        for (int n = 0; n < navg; n++) {
            tmin = std::min(tmin, R(tg, p, n));
            tmax = std::max(tmax, R(tg, p, n));
        }
        const double span = std::abs(tmax - tmin);
        if (span <= amax * coeff) nok++;
    }
    return nok / (np + 0.0) >= 0.9;
}

}  // namespace aefdtd
