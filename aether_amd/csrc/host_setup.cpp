// host_setup.cpp -- setup-time host arithmetic of the engine: material pole
// conversion, CPML grading profiles, the analytic pulse and the tapering curve.
// All of it is FP64 and runs once per run (or once per step for the pulse
// tables), so it stays on the host; the results are uploaded as tables.
#include "engine.hpp"

#include <cmath>

namespace aefdtd {

namespace {
const cplx Im(0.0, 1.0);
}

// Material -> update coefficients.
//  Pole time exponentials (a, b) per model: Debye a=ds/t0, b=-1/t0; Drude
//  a=-wd^2/g, b=-g, sigma=e0 wd^2/g; Lorentz a=-i A O^2/beta, b=-G/2+i beta with
//  beta=sqrt(O^2-G^2/4) taken as a real root; critical point a=-2i A O e^{-iP},
//  b=-G+iO  (materials/dielec_models.cpp:222-226,281-290,350-358,419-423), in
//  the order Debye, Drude, Lorentz, CP (dielec_models.cpp:71-107).
//  Recursive convolution: Crec=e^{b dt}, chi=a/b (e^{b dt}-1),
//  Dchi=-a/b (1-e^{b dt})^2 (core/mats_RC.cpp:143-154);
//  e_eff = eps_inf + sum Re chi + sigma dt/e0, C1=eps_inf/e_eff, C2x=dt/(Dx e0 e_eff),
//  C3=1/e_eff, C4=dt/(e0 e_eff) (core/mats_RC.cpp:159-184).  Without poles:
//  C1=1, C2x=dt/(Dx e0 eps), C3=0, C4=dt/(e0 eps) (core/mats_const.cpp:71-80).
void material_coefficients(const aefdtd_material &m, double Dx, double Dy, double Dz, double Dt, HostMat &M) {
    std::vector<cplx> va, vb;
    double sig = 0;
    for (int i = 0; i < m.n_debye; i++) {
        double ds = m.debye[2 * i], t0 = m.debye[2 * i + 1];
        va.push_back(cplx(ds / t0));
        vb.push_back(cplx(-1.0 / t0));
    }
    for (int i = 0; i < m.n_drude; i++) {
        double wd = m.drude[2 * i], g = m.drude[2 * i + 1];
        double wd2 = wd * wd;
        va.push_back(cplx(-wd2 / g));
        vb.push_back(cplx(-g));
        sig += kE0 * wd2 / g;
    }
    for (int i = 0; i < m.n_lorentz; i++) {
        double A = m.lorentz[3 * i], O = m.lorentz[3 * i + 1], G = m.lorentz[3 * i + 2];
        double alp = G / 2.0;
        cplx beta = std::sqrt(O * O - alp * alp);
        cplx eta = A * O * O / beta;
        va.push_back(-eta * Im);
        vb.push_back(-alp + beta * Im);
    }
    for (int i = 0; i < m.n_critpoint; i++) {
        double A = m.critpoint[4 * i], O = m.critpoint[4 * i + 1];
        double P = m.critpoint[4 * i + 2], G = m.critpoint[4 * i + 3];
        va.push_back(-2.0 * A * O * std::exp(-P * Im) * Im);
        vb.push_back(-G + O * Im);
    }
    M.set = true;
    M.ei = m.eps_inf;
    if (va.empty()) {
        M.rc = false;
        M.sig = 0;
        M.C1 = 1;
        M.C2x = Dt / (Dx * kE0 * M.ei);
        M.C2y = Dt / (Dy * kE0 * M.ei);
        M.C2z = Dt / (Dz * kE0 * M.ei);
        M.C3 = 0;
        M.C4 = Dt / (kE0 * M.ei);
        M.Crec.clear(); M.chi.clear(); M.Dchi.clear();
        return;
    }
    M.rc = true;
    M.sig = sig;
    size_t np = va.size();
    M.Crec.resize(np); M.chi.resize(np); M.Dchi.resize(np);
    double chisum = 0;
    for (size_t p = 0; p < np; p++) {
        cplx e = std::exp(vb[p] * Dt);
        M.Crec[p] = e;
        M.chi[p] = va[p] / vb[p] * (std::exp(vb[p] * Dt) - 1.0);
        M.Dchi[p] = -va[p] / vb[p] * std::pow(1.0 - std::exp(vb[p] * Dt), 2.0);
    }
    for (size_t p = 0; p < np; p++) chisum += std::real(M.chi[p]);
    double e_eff = M.ei + chisum + M.sig * Dt / kE0;
    M.C1 = M.ei / e_eff;
    M.C2x = Dt / (Dx * kE0 * e_eff);
    M.C2y = Dt / (Dy * kE0 * e_eff);
    M.C2z = Dt / (Dz * kE0 * e_eff);
    M.C3 = 1.0 / e_eff;
    M.C4 = Dt / (kE0 * e_eff);
}

// Graded CPML profile at fractional index ind (core/fdtd_pml.cpp:345-433):
// polynomial grading of order 4 for kappa and sigma, (1-x)^2 for alpha,
// b = exp(-(sigma/kappa+alpha) dt/e0), c = sigma (b-1)/(sigma kappa + kappa^2 alpha).
void pml_profile(double ind, int N, int pm, int pp, const double kap[2], const double sig[2],
                 const double alp[2], double Dt, double &kappa, double &b, double &c) {
    const double order = 4, order_a = 2;
    kappa = 1.0;
    auto side = [&](double x, int s) {
        kappa = 1.0 + (kap[s] - 1.0) * std::pow(x, order);
        double sg = sig[s] * std::pow(x, order);
        double al = alp[s] * std::pow(1.0 - x, order_a);
        b = std::exp(-(sg / kappa + al) * Dt / kE0);
        c = sg / (sg * kappa + kappa * kappa * al) * (b - 1.0);
    };
    if (ind < pm) side((pm - ind) / static_cast<double>(pm), 0);
    if (ind > N - pp) side((ind - (N - pp)) / static_cast<double>(pp), 1);
}

// sigma_max scale of set_pml (core/fdtd_pml.cpp:445-447, grading order 4).
double pml_sigma_scale(double D) { return 0.8 * (4.0 + 1.0) / (std::sqrt(kMu0 / kE0) * D); }

// Gaussian-envelope plane wave of ChpIn (core/chpin.cpp:479-545).
void pulse_init(aefdtd_pulse &p, double li, double lf, double inc_d, double pol_d, double dir_d,
                double sx, double sy, double sz, double n_ind) {
    double wi = 2.0 * kPi * kC / li, wf = 2.0 * kPi * kC / lf;
    double w0 = (wi + wf) / 2.0;
    double l0 = 2.0 * kPi * kC / w0;
    double dw = std::abs(0.5 * (wi - w0));
    p.lambda = l0;
    p.n_ind = n_ind;
    double inc = inc_d * kPi / 180.0, pol = pol_d * kPi / 180.0, dir = dir_d * kPi / 180.0;
    p.dz = 2.0 * kC / dw / p.n_ind;
    p.k = 2.0 * kPi / p.lambda * p.n_ind;
    p.omega = 2.0 * kPi * kC / p.lambda;
    double incr = kPi / 2.0 - inc;
    double Ca = std::cos(incr), Sa = std::sin(incr);
    double Cb = std::cos(dir), Sb = std::sin(dir);
    double Cc = std::cos(pol), Sc = std::sin(pol);
    p.kx = p.k * (Ca * Cb);
    p.ky = p.k * (Ca * Sb);
    p.kz = p.k * (-Sa);
    p.Ex0 = (Sa * Cb * Sc - Sb * Cc) * 1.0;
    p.Ey0 = (Sa * Sb * Sc + Cb * Cc) * 1.0;
    p.Ez0 = Ca * Sc * 1.0;
    double h = 1.0 / (kC * kMu0) * 1.0 * p.n_ind;
    p.Hx0 = h * (Sa * Cb * Cc + Sb * Sc);
    p.Hy0 = h * (Sa * Sb * Cc - Cb * Sc);
    p.Hz0 = h * (Ca * Cc);
    double bet = 1e-4;
    p.C1 = -1.0 / p.k * (p.kx * sx + p.ky * sy + p.kz * sz) + p.dz * std::sqrt(-std::log(bet));
}

double pulse_eval(const aefdtd_pulse &p, int comp, double x, double y, double z, double t) {
    double ph = p.kx * x + p.ky * y + p.kz * z;
    double a = ph - p.omega * t;
    double b = (1.0 / p.k * ph - kC * t / p.n_ind + p.C1) / p.dz;
    const double amp[6] = {p.Ex0, p.Ey0, p.Ez0, p.Hx0, p.Hy0, p.Hz0};
    return amp[comp] * std::exp(-b * b) * std::cos(a);
}

// Smoothstep tapering (common/mathUT.cpp:1186-1193).
double s_curve(double x, double xmin, double xmax) {
    double u = (x - xmin) / (xmax - xmin);
    if (u <= 0) return 0;
    if (u >= 1.0) return 1.0;
    return u * u * (3.0 - 2.0 * u);
}

// Oscillator::initialize (sources/sc_oscillator.cpp:70-84): centre w0 of the
// [lambda_min, lambda_max] band in rad/s, width dw from the sqrt(0.1) edge
// level, delay tshift (truncated to whole steps) from the 1e-7 onset level.
void oscillator_init(double lambda_min, double lambda_max, double Dt, OscParams &o) {
    const double alp = std::sqrt(0.1), bet = 1e-7;
    const double w1 = 2.0 * kPi * kC / lambda_max, w2 = 2.0 * kPi * kC / lambda_min;
    o.w0 = (w1 + w2) / 2.0;
    o.dw = std::abs(w2 - w1) / (2.0 * std::sqrt(-std::log(alp)));
    o.tshift = (int)(std::sqrt(-4.0 / (o.dw * o.dw) * std::log(bet)) / Dt);
    o.Dt = Dt;
}

// Oscillator::deep_inject_E's value (sc_oscillator.cpp:37-38): the complex
// carrier e^{-i w0 step Dt} times the real Gaussian envelope, real part, x50.
double oscillator_value(const OscParams &o, int step) {
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> v = std::exp(-o.w0 * step * o.Dt * I) *
                                   std::exp(-std::pow(o.dw / 2.0 * ((step - o.tshift) * o.Dt), 2.0));
    return std::real(v) * 50;
}

}  // namespace aefdtd

extern "C" {

int aefdtd_pulse_init(aefdtd_pulse *p, double li, double lf, double inc, double pol, double dir,
                      double sx, double sy, double sz, double n_ind) {
    if (!p || !(li > 0) || !(lf > 0) || !(n_ind > 0)) return AEFDTD_EINVAL;
    aefdtd::pulse_init(*p, li, lf, inc, pol, dir, sx, sy, sz, n_ind);
    return AEFDTD_OK;
}

double aefdtd_pulse_eval(const aefdtd_pulse *p, int comp, double x, double y, double z, double t) {
    if (!p || comp < 0 || comp > 5) return 0.0;
    return aefdtd::pulse_eval(*p, comp, x, y, z, t);
}

}
