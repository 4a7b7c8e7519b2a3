// fdtd_oracle.cpp -- CPU FP64 restatement of Aether's FDTD hot path.
//
// TEST INFRASTRUCTURE.  This is the checker the HIP engine is compared with
// and the CPU baseline timed beside it; the product never links it.
//
// It restates, in this project's own code, the arithmetic and the phase order
// of Aether's CPU solver (all paths relative to the reference's src/):
//   grid / padding / extension  finite_differences/FDTD/core/fdtd_core.cpp:158-285
//   E, H stencils               core/fdtd_core.cpp:363-969
//   step phases + threads       core/fdtd_core.cpp:1163-1335, core/fdtd_threads.cpp:32-235
//   CPML                        core/fdtd_pml.cpp:26-85,305-497,504-929
//   materials (const, RC)       core/mats.cpp:27-52,110-193,394-496,564-567,
//                               core/mats_RC.cpp:21-50,126-184, core/mats_const.cpp:18-80,
//                               materials/dielec_models.cpp:52-153,217-423
//   set_matsgrid / set_material core/fdtd_utils.cpp:572-796,810-820,855-881
//   local_E/H interpolation     core/fdtd_utils.cpp:457-507
//   analytic pulse              core/chpin.cpp:44-90,479-545
//   sensors                     sensors/sensors.cpp:196-299,456-644, sensors/s_spec_power.cpp:44-232
//   drivers                     modes/normal_incidence.cpp:161-607, modes/single_particle.cpp:29-415
//   helpers                     common/mathUT.cpp:1169-1193, common/phys_constants.h:19-21
//
// Threading follows the reference: a persistent pool of Nthreads workers and a
// barrier between phases, with the z->y->x slab rule of fdtd_threads.cpp:56-82.
//
// Pins (tests/test_oracle_pins.py, tests/test_field_sensors.py,
// tests/test_completion.py): pole and RC constants against Aether's own
// compiled material layer and quadrature, coefficients and CPML grading in
// closed form, the stencils by the discrete-energy invariant and the
// dispersion recurrence, CPML by absorption and the transparent-CPML wall
// test, TFSF by the field outside the box in a uniform medium, the extension
// by edge replication, the normal-mode drivers against Airy.
// Parity unpinned (only the reference's text defines them, or no check above
// reaches them): the kappa_y gating (fdtd_core.cpp:399), the sensor tapering
// step (sensors.cpp:236-240), the CPML psi slab bounds beyond absorption and
// the wall test (fdtd_pml.cpp:504-929), and the custom driver's phase sequence
// (default_fdtd.cpp:148-165) beyond the oscillator and sensor pins.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <numbers>
#include <random>
#include <string>
#include <thread>
#include <vector>

// The caller's end_computation flag, polled once per step by the drivers.
static bool cancelled(const int *flag) { return flag && __atomic_load_n(flag, __ATOMIC_RELAXED); }

// Planted mistakes (aeoracle_set_slip), so that the tests can show that each
// pin of the oracle fails when the function it pins is wrong: 0 = none,
// 1 = TFSF-E z faces read the auxiliary H one plane off, 2 = TFSF-H lower y
// face with the sign of the Hz correction flipped, 3 = no PEC wall for Ex on y,
// 4 = the far-side x extension copies the second-last structure plane,
// 5 = normal incidence without the -Dz/2 phase calibration, 6 = local_Ex
// averaging toward j-1.
static int g_slip = 0;

namespace {

using cplx = std::complex<double>;
const cplx Im(0.0, 1.0);
constexpr double e0 = 8.854187817e-12;       // phys_constants.h:19
constexpr double mu0 = 12.566370614e-7;      // phys_constants.h:20
constexpr double c_light = 299792458.0;      // phys_constants.h:21
constexpr double Pi = std::numbers::pi_v<double>;  // mathUT.h:34

std::string g_create_error;
int g_threads = 0;

int nearest_integer(double x) {               // mathUT.cpp:1169-1173
    if (x >= 0) return static_cast<int>(x + 0.5);
    return static_cast<int>(x - 0.5);
}

double s_curve(double x, double xmin, double xmax) {   // mathUT.cpp:1186-1193
    double u = (x - xmin) / (xmax - xmin);
    if (u <= 0) return 0;
    if (u >= 1.0) return 1.0;
    return u * u * (3.0 - 2.0 * u);
}

// ---------------------------------------------------------------- thread pool
// Fork/join pool: run(f) calls f(ID) on every worker and returns when all
// are done -- the signal/wait pair of ThreadsAlternator (thread_utils.cpp:82-267).
class Pool {
public:
    explicit Pool(int n) : n_(n) {
        for (int i = 1; i < n_; i++) workers_.emplace_back([this, i] { loop(i); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)> &f) {
        if (n_ == 1) { f(0); return; }
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &f;
            pending_ = n_ - 1;
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }
private:
    void loop(int id) {
        unsigned long seen = 0;
        for (;;) {
            const std::function<void(int)> *job;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (quit_) return;
                job = job_;
            }
            (*job)(id);
            {
                std::lock_guard<std::mutex> lk(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    int n_;
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int pending_ = 0;
    unsigned long gen_ = 0;
    bool quit_ = false;
};

// ---------------------------------------------------------------- materials
struct Poles {
    std::vector<cplx> a, b;
    double ei = 1.0, sig = 0.0;
    bool is_const = true;
};

// Dielec_model::get_time_exp + get_const + get_sigma (dielec_models.cpp:66-153,
// per-model constants :222-226 Debye, :281-290 Drude, :350-358 Lorentz,
// :419-423 critical point); order of poles: Debye, Drude, Lorentz, CP.
Poles material_poles(const aefdtd_material &m) {
    Poles P;
    P.ei = m.eps_inf;
    for (int i = 0; i < m.n_debye; i++) {
        double ds = m.debye[2 * i], t0 = m.debye[2 * i + 1];
        P.a.push_back(cplx(ds / t0));
        P.b.push_back(cplx(-1.0 / t0));
    }
    for (int i = 0; i < m.n_drude; i++) {
        double wd = m.drude[2 * i], g = m.drude[2 * i + 1];
        double wd2 = wd * wd;
        P.a.push_back(cplx(-wd2 / g));
        P.b.push_back(cplx(-g));
        P.sig += e0 * wd2 / g;
    }
    for (int i = 0; i < m.n_lorentz; i++) {
        double A = m.lorentz[3 * i], O = m.lorentz[3 * i + 1], G = m.lorentz[3 * i + 2];
        double alp = G / 2.0;
        cplx bet = std::sqrt(O * O - alp * alp);   // real sqrt, as the reference
        cplx eta = A * O * O / bet;
        P.a.push_back(-eta * Im);
        P.b.push_back(-alp + bet * Im);
    }
    for (int i = 0; i < m.n_critpoint; i++) {
        double A = m.critpoint[4 * i], O = m.critpoint[4 * i + 1];
        double Ph = m.critpoint[4 * i + 2], G = m.critpoint[4 * i + 3];
        P.a.push_back(-2.0 * A * O * std::exp(-Ph * Im) * Im);
        P.b.push_back(-G + O * Im);
    }
    P.is_const = P.a.empty();
    return P;
}

struct Mat {                       // FDTD_Material, the members the RC/const paths use
    bool rc = false;               // m_type==MAT_RC (comp_ante=1, comp_simp=0)
    int Np = 0;
    double C1 = 1, C2 = 0, C2x = 0, C2y = 0, C2z = 0, C3 = 0, C4 = 0, ei = 1, sig = 0;
    std::vector<cplx> Crec, chi, Dchi;
    int x1 = 0, y1 = 0, z1 = 0, xs = 0, ys = 0, zs = 0;   // link_grid bbox
    std::vector<cplx> psi;                                // Grid4 (x,y,z,3Np), 4th slowest
    cplx &P(int i, int j, int k, int c) {
        return psi[(size_t)(i - x1) + (size_t)xs * ((size_t)(j - y1) + (size_t)ys * ((size_t)(k - z1) + (size_t)zs * c))];
    }
};

// FDTD_Material::set_base_mat -> set_const / RC_dielec_treat + RC_recalc
// (mats.cpp:445-464, mats_const.cpp:18-36,71-80, mats_RC.cpp:126-184).
void material_setup(Mat &M, const aefdtd_material &m, double Dx, double Dy, double Dz, double Dt) {
    Poles P = material_poles(m);
    if (P.is_const) {
        M.rc = false;
        M.Np = 0;
        M.ei = P.ei;
        M.sig = 0;
        M.C1 = 1;
        M.C2 = Dt / (e0 * M.ei);
        M.C2x = Dt / (Dx * e0 * M.ei);
        M.C2y = Dt / (Dy * e0 * M.ei);
        M.C2z = Dt / (Dz * e0 * M.ei);
        M.C3 = 0;
        M.C4 = Dt / (e0 * M.ei);
        return;
    }
    M.rc = true;
    M.Np = (int)P.a.size();
    M.ei = P.ei;
    M.sig = P.sig;
    M.Crec.assign(M.Np, 0);
    M.chi.assign(M.Np, 0);
    M.Dchi.assign(M.Np, 0);
    for (int i = 0; i < M.Np; i++) {
        cplx pa = P.a[i], pb = P.b[i];
        M.Crec[i] = std::exp(pb * Dt);
        M.chi[i] = pa / pb * (std::exp(pb * Dt) - 1.0);
        M.Dchi[i] = -pa / pb * std::pow(1.0 - std::exp(pb * Dt), 2.0);
    }
    double chitmp = 0;
    for (int i = 0; i < M.Np; i++) chitmp += std::real(M.chi[i]);
    double e_eff = M.ei + chitmp + M.sig * Dt / e0;
    M.C1 = M.ei / e_eff;
    M.C2 = Dt / (e0 * e_eff);
    M.C2x = Dt / (Dx * e0 * e_eff);
    M.C2y = Dt / (Dy * e0 * e_eff);
    M.C2z = Dt / (Dz * e0 * e_eff);
    M.C3 = 1.0 / e_eff;
    M.C4 = Dt / (e0 * e_eff);
}

// CPML profile at fractional index ind (fdtd_pml.cpp:345-433); pm/pp PML
// widths of the axis, kap/sig/alp per side (sig already scaled by set_pml).
void pml_coeffs(double ind, int N, int pm, int pp, const double kap[2], const double sig[2],
                const double alp[2], double Dt, double &kappa, double &b, double &c) {
    const double pml_m = 4, pml_ma = 2;
    kappa = 1.0;
    if (ind < pm) {
        double x = (pm - ind) / static_cast<double>(pm);
        kappa = 1.0 + (kap[0] - 1.0) * std::pow(x, pml_m);
        double s = sig[0] * std::pow(x, pml_m);
        double a = alp[0] * std::pow(1.0 - x, pml_ma);
        b = std::exp(-(s / kappa + a) * Dt / e0);
        c = s / (s * kappa + kappa * kappa * a) * (b - 1.0);
    }
    if (ind > N - pp) {
        double x = (ind - (N - pp)) / static_cast<double>(pp);
        kappa = 1.0 + (kap[1] - 1.0) * std::pow(x, pml_m);
        double s = sig[1] * std::pow(x, pml_m);
        double a = alp[1] * std::pow(1.0 - x, pml_ma);
        b = std::exp(-(s / kappa + a) * Dt / e0);
        c = s / (s * kappa + kappa * kappa * a) * (b - 1.0);
    }
}

double pml_sig_scale(double D) {               // fdtd_pml.cpp:445-447 with pml_m=4
    return 0.8 * (4.0 + 1.0) / (std::sqrt(mu0 / e0) * D);
}

// ---------------------------------------------------------------- pulse
void pulse_setfields(aefdtd_pulse &p, double inc, double pol, double dir) {   // chpin.cpp:516-545
    p.k = 2.0 * Pi / p.lambda * p.n_ind;
    p.omega = 2.0 * Pi * c_light / p.lambda;
    double incr = Pi / 2.0 - inc;
    double Ca = std::cos(incr), Sa = std::sin(incr);
    double Cb = std::cos(dir), Sb = std::sin(dir);
    double Cc = std::cos(pol), Sc = std::sin(pol);
    p.kx = p.k * (Ca * Cb);
    p.ky = p.k * (Ca * Sb);
    p.kz = p.k * (-Sa);
    double amp = 1;
    p.Ex0 = (Sa * Cb * Sc - Sb * Cc) * amp;
    p.Ey0 = (Sa * Sb * Sc + Cb * Cc) * amp;
    p.Ez0 = Ca * Sc * amp;
    double tmp = 1.0 / (c_light * mu0) * amp * p.n_ind;
    p.Hx0 = tmp * (Sa * Cb * Cc + Sb * Sc);
    p.Hy0 = tmp * (Sa * Sb * Cc - Cb * Sc);
    p.Hz0 = tmp * (Ca * Cc);
}

void pulse_init(aefdtd_pulse &p, double li, double lf, double inci, double poli, double diri,
                double sx, double sy, double sz, double n_ind) {     // chpin.cpp:479-514
    double wi = 2.0 * Pi * c_light / li;
    double wf = 2.0 * Pi * c_light / lf;
    double w0 = (wi + wf) / 2.0;
    double l0 = 2.0 * Pi * c_light / w0;
    double dw = std::abs(0.5 * (wi - w0));
    p.lambda = l0;
    p.n_ind = n_ind;
    double inc = inci * Pi / 180.0, pol = poli * Pi / 180.0, dir = diri * Pi / 180.0;
    p.dz = 2.0 * c_light / dw / p.n_ind;
    pulse_setfields(p, inc, pol, dir);
    double bet = 1e-4;
    p.C1 = -1.0 / p.k * (p.kx * sx + p.ky * sy + p.kz * sz) + p.dz * std::sqrt(-std::log(bet));
}

double pulse_eval(const aefdtd_pulse &p, int comp, double x, double y, double z, double t) {  // chpin.cpp:44-90
    double pp = p.kx * x + p.ky * y + p.kz * z;
    double a = pp - p.omega * t;
    double b = (1.0 / p.k * pp - c_light * t / p.n_ind + p.C1) / p.dz;
    const double amp[6] = {p.Ex0, p.Ey0, p.Ez0, p.Hx0, p.Hy0, p.Hz0};
    return amp[comp] * std::exp(-b * b) * std::cos(a);
}

}  // namespace

// ================================================================= grid
struct SpectSensor {               // Spect_Poynting over SensorFieldHolder (interpolate=true)
    int type, x1, x2, y1, y2, z1, z2, span1, span2, Nl;
    int step = 0, Nt, Ntap;
    double tapE = 1.0, tapH = 1.0, Dx, Dy, Dz, Dt;
    std::vector<double> lambda;
    std::vector<double> t[6];
    std::vector<cplx> sp[6];       // (span1, span2, Nl) x-fastest
};

// Oscillator (sources/sc_oscillator.cpp:20-84) with the Source base members
// it uses (core/sources.h:219-233: tshift is an int).
struct Oscillator {
    int type;                      // field index 0..2 after initialize's XM->X folding
    int x1, x2, y1, y2, z1, z2;
    int step = 0, tshift = 0;
    double w0 = 0, dw = 0, Dt = 0;
};

// FieldBlock (sensors/s_fieldmap.cpp:28-230) and FieldMap (:290-470): a
// single wavelength, one accumulation per feed at the sensor's own step.
struct FieldSens {
    bool block = true;
    int type = 0;                  // FieldMap: NORMAL_X/Y/Z after initialize's folding (:433-440)
    bool cumulative = false, mag_map = false;
    int x1, x2, y1, y2, z1, z2;    // linked (full-grid) location
    int span1 = 0, span2 = 0, span3 = 0;
    int step = 0;
    double lambda = 0, Dt = 0;
    std::vector<cplx> acc[6];      // block: Ex..Hz (span1, span2, span3); map: acc_Ex..Ez (span1, span2)
};

struct PlaneSum {
    int k;
    bool has_inc, subtract;
    aefdtd_pulse inc;
    double zpos;
    std::vector<double> tap, out;  // out: per step 6 values
};


struct aeoracle_grid {
    int mode = AEFDTD_MODE_CUSTOM;
    int tstep = 0;
    int Nx, Ny, Nz, Nt, Ntap = 0;
    int Nx_s, Ny_s, Nz_s;
    int xs_s, xs_e, ys_s, ys_e, zs_s, zs_e;
    int pad[6];
    int pml[6];
    double Dx, Dy, Dz, Dt;
    double dte, dtm, dtdex, dtdey, dtdez, dtdmx, dtdmy, dtdmz;
    double pml_kappa[6], pml_sigma[6], pml_alpha[6];
    bool enable[6] = {true, true, true, true, true, true};
    std::vector<double> F[6];                 // Ex,Ey,Ez,Hx,Hy,Hz
    std::vector<uint32_t> matsgrid;
    std::vector<Mat> mats;
    unsigned Nmat = 1;
    // CPML state: PsiExy,PsiExz,PsiEyx,PsiEyz,PsiEzx,PsiEzy and H analogues
    std::vector<double> PsiEx[2], PsiEy[2], PsiEz[2], PsiHx[2], PsiHy[2], PsiHz[2];
    int psx[3], psy[3], psz[3];               // dims of the x-, y-, z-slab arrays
    std::vector<double> kE[3], kH[3], bE[3], bH[3], cE[3], cH[3];
    // driver actions moved into the grid object
    int inj_k = -1;
    std::vector<double> inj[4];
    const aeoracle_grid *tfsf_aux = nullptr;
    int tb[6] = {0, 0, 0, 0, 0, 0};
    bool tf_per_x = false, tf_per_y = false;
    std::vector<SpectSensor *> sensors;
    std::vector<FieldSens *> fsensors;
    std::vector<PlaneSum *> sums;
    std::vector<Oscillator> oscs;
    Pool *pool = nullptr;
    std::string err;

    size_t id(int i, int j, int k) const { return (size_t)i + (size_t)Nx * ((size_t)j + (size_t)Ny * k); }
    double &Ex(int i, int j, int k) { return F[0][id(i, j, k)]; }
    double &Ey(int i, int j, int k) { return F[1][id(i, j, k)]; }
    double &Ez(int i, int j, int k) { return F[2][id(i, j, k)]; }
    double &Hx(int i, int j, int k) { return F[3][id(i, j, k)]; }
    double &Hy(int i, int j, int k) { return F[4][id(i, j, k)]; }
    double &Hz(int i, int j, int k) { return F[5][id(i, j, k)]; }
    double cf(int c, int i, int j, int k) const { return F[c][id(i, j, k)]; }
    uint32_t mg(int i, int j, int k) const { return matsgrid[id(i, j, k)]; }

    void allocate_pml();
    void pml_coeff_calc();
    // phases
    void advMats_ante(int i1, int i2);
    void advE(int c, int i1, int i2, int j1, int j2, int k1, int k2);
    void advMats_simp(int i1, int i2);
    void app_pml_Ex(int i1, int i2);
    void app_pml_Ey(int j1, int j2);
    void app_pml_Ez(int k1, int k2);
    void advH(int c, int i1, int i2, int j1, int j2, int k1, int k2);
    void app_pml_Hx(int i1, int i2);
    void app_pml_Hy(int j1, int j2);
    void app_pml_Hz(int k1, int k2);
    void tfsf_E();
    void tfsf_H();
    void update_E();
    void update_H();
    double local(int c, int i, int j, int k) const;
};


void aeoracle_grid::allocate_pml() {             // fdtd_pml.cpp:26-85
    auto dims = [&](int ax, int *d) {
        d[0] = d[1] = d[2] = 1;
        if (pml[2 * ax] != 0 || pml[2 * ax + 1] != 0) {
            d[0] = Nx; d[1] = Ny; d[2] = Nz;
            d[ax] = pml[2 * ax] + pml[2 * ax + 1];
        }
    };
    dims(0, psx); dims(1, psy); dims(2, psz);
    size_t nx = (size_t)psx[0] * psx[1] * psx[2];
    size_t ny = (size_t)psy[0] * psy[1] * psy[2];
    size_t nz = (size_t)psz[0] * psz[1] * psz[2];
    // PsiEyx,PsiEzx,PsiHyx,PsiHzx on the x slabs; PsiExy,PsiEzy,PsiHxy,PsiHzy on y; PsiExz,PsiEyz,PsiHxz,PsiHyz on z
    PsiEy[0].assign(nx, 0); PsiEz[0].assign(nx, 0); PsiHy[0].assign(nx, 0); PsiHz[0].assign(nx, 0);
    PsiEx[0].assign(ny, 0); PsiEz[1].assign(ny, 0); PsiHx[0].assign(ny, 0); PsiHz[1].assign(ny, 0);
    PsiEx[1].assign(nz, 0); PsiEy[1].assign(nz, 0); PsiHx[1].assign(nz, 0); PsiHy[1].assign(nz, 0);
    int N[3] = {Nx, Ny, Nz};
    for (int a = 0; a < 3; a++) {
        kE[a].assign(N[a], 1.0); kH[a].assign(N[a], 1.0);
        bE[a].assign(N[a], 0.0); bH[a].assign(N[a], 0.0);
        cE[a].assign(N[a], 0.0); cH[a].assign(N[a], 0.0);
    }
}

void aeoracle_grid::pml_coeff_calc() {           // fdtd_pml.cpp:305-343
    int N[3] = {Nx, Ny, Nz};
    for (int a = 0; a < 3; a++) {
        std::fill(kE[a].begin(), kE[a].end(), 1.0); std::fill(kH[a].begin(), kH[a].end(), 1.0);
        std::fill(bE[a].begin(), bE[a].end(), 0.0); std::fill(bH[a].begin(), bH[a].end(), 0.0);
        std::fill(cE[a].begin(), cE[a].end(), 0.0); std::fill(cH[a].begin(), cH[a].end(), 0.0);
        if (pml[2 * a] || pml[2 * a + 1]) {
            const double kap[2] = {pml_kappa[2 * a], pml_kappa[2 * a + 1]};
            const double sig[2] = {pml_sigma[2 * a], pml_sigma[2 * a + 1]};
            const double alp[2] = {pml_alpha[2 * a], pml_alpha[2 * a + 1]};
            for (int i = 0; i < N[a]; i++) {
                pml_coeffs(i, N[a], pml[2 * a], pml[2 * a + 1], kap, sig, alp, Dt, kE[a][i], bE[a][i], cE[a][i]);
                pml_coeffs(i + 0.5, N[a], pml[2 * a], pml[2 * a + 1], kap, sig, alp, Dt, kH[a][i], bH[a][i], cH[a][i]);
            }
        }
    }
}

void aeoracle_grid::advMats_ante(int i1, int i2) {      // fdtd_core.cpp:971-988 + mats_RC.cpp:21-36
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = i1; i < i2; i++) {
                Mat &M = mats[mg(i, j, k)];
                if (!M.rc) continue;
                double ex = Ex(i, j, k), ey = Ey(i, j, k), ez = Ez(i, j, k);
                for (int p = 0; p < M.Np; p++) {
                    cplx &px = M.P(i, j, k, 3 * p + 0);
                    cplx &py = M.P(i, j, k, 3 * p + 1);
                    cplx &pz = M.P(i, j, k, 3 * p + 2);
                    px = M.Crec[p] * px + M.Dchi[p] * ex;
                    py = M.Crec[p] * py + M.Dchi[p] * ey;
                    pz = M.Crec[p] * pz + M.Dchi[p] * ez;
                }
            }
}

// advEx/advEy/advEz (fdtd_core.cpp:363-419,478-533,600-663); c = 0,1,2
void aeoracle_grid::advE(int c, int i1_, int i2_, int j1_, int j2_, int k1_, int k2_) {
    const bool px = pml[0] || pml[1], py_q = pml[2] || pml[1], pz = pml[4] || pml[5];
    for (int k = k1_; k < k2_; k++) {
        int km = (k == 0) ? Nz - 1 : k - 1;
        double ikz = 1.0 / (pz ? kE[2][k] : 1.0);
        for (int j = j1_; j < j2_; j++) {
            int jm = (j == 0) ? Ny - 1 : j - 1;
            double iky = 1.0 / (py_q ? kE[1][j] : 1.0);     // kappa_y quirk: pml_ym||pml_xp
            for (int i = i1_; i < i2_; i++) {
                int im = (i == 0) ? Nx - 1 : i - 1;
                double ikx = 1.0 / (px ? kE[0][i] : 1.0);
                const Mat &M = mats[mg(i, j, k)];
                if (c == 0)
                    Ex(i, j, k) = M.C1 * Ex(i, j, k) + M.C2y * iky * (Hz(i, j, k) - Hz(i, jm, k))
                                                     - M.C2z * ikz * (Hy(i, j, k) - Hy(i, j, km));
                else if (c == 1)
                    Ey(i, j, k) = M.C1 * Ey(i, j, k) + M.C2z * ikz * (Hx(i, j, k) - Hx(i, j, km))
                                                     - M.C2x * ikx * (Hz(i, j, k) - Hz(im, j, k));
                else
                    Ez(i, j, k) = M.C1 * Ez(i, j, k) + M.C2x * ikx * (Hy(i, j, k) - Hy(im, j, k))
                                                     - M.C2y * iky * (Hx(i, j, k) - Hx(i, jm, k));
            }
        }
    }
}

void aeoracle_grid::advMats_simp(int i1, int i2) {     // fdtd_core.cpp:1014-1033 + mats_RC.cpp:38-50
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = i1; i < i2; i++) {
                Mat &M = mats[mg(i, j, k)];
                if (!M.rc) continue;
                for (int d = 0; d < 3; d++) {
                    double s = 0;
                    for (int p = 0; p < M.Np; p++) s += std::real(M.P(i, j, k, 3 * p + d));
                    F[d][id(i, j, k)] += M.C3 * s;
                }
            }
}

// CPML-E (fdtd_pml.cpp:504-735).  Slab arrays: x slabs (pxm+pxp,Ny,Nz), etc.
void aeoracle_grid::app_pml_Ex(int i1, int i2) {
    double inv_Dy = 1.0 / Dy, inv_Dz = 1.0 / Dz;
    if (pml[2] || pml[3]) {
        auto P = [&](int i, int n, int k) -> double & { return PsiEx[0][(size_t)i + (size_t)psy[0] * ((size_t)n + (size_t)psy[1] * k)]; };
        for (int i = i1; i < i2; i++)
            for (int k = 0; k < Nz; k++) {
                for (int j = 1; j < pml[2]; j++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[1][j] * P(i, j, k) + cE[1][j] * inv_Dy * (Hz(i, j, k) - Hz(i, j - 1, k));
                    Ex(i, j, k) += C4 * P(i, j, k);
                }
                for (int j = Ny - pml[3] + 1; j < Ny; j++) {
                    int n = j - (Ny - pml[3]) + pml[2];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, n, k) = bE[1][j] * P(i, n, k) + cE[1][j] * inv_Dy * (Hz(i, j, k) - Hz(i, j - 1, k));
                    Ex(i, j, k) += C4 * P(i, n, k);
                }
            }
        for (int i = i1; i < i2; i++)
            for (int k = 0; k < Nz; k++)
                if (g_slip != 3) Ex(i, 0, k) = 0;
    }
    if (pml[4] || pml[5]) {
        auto P = [&](int i, int j, int n) -> double & { return PsiEx[1][(size_t)i + (size_t)psz[0] * ((size_t)j + (size_t)psz[1] * n)]; };
        for (int k = 1; k < pml[4]; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = i1; i < i2; i++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[2][k] * P(i, j, k) + cE[2][k] * inv_Dz * (Hy(i, j, k) - Hy(i, j, k - 1));
                    Ex(i, j, k) -= C4 * P(i, j, k);
                }
        for (int k = Nz - pml[5] + 1; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = i1; i < i2; i++) {
                    int n = k - (Nz - pml[5]) + pml[4];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, n) = bE[2][k] * P(i, j, n) + cE[2][k] * inv_Dz * (Hy(i, j, k) - Hy(i, j, k - 1));
                    Ex(i, j, k) -= C4 * P(i, j, n);
                }
        for (int i = i1; i < i2; i++)
            for (int j = 0; j < Ny; j++) Ex(i, j, 0) = 0;
    }
}

void aeoracle_grid::app_pml_Ey(int j1, int j2) {
    double inv_Dx = 1.0 / Dx, inv_Dz = 1.0 / Dz;
    if (pml[0] || pml[1]) {
        auto P = [&](int n, int j, int k) -> double & { return PsiEy[0][(size_t)n + (size_t)psx[0] * ((size_t)j + (size_t)psx[1] * k)]; };
        for (int j = j1; j < j2; j++)
            for (int k = 0; k < Nz; k++) {
                for (int i = 1; i < pml[0]; i++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[0][i] * P(i, j, k) + cE[0][i] * inv_Dx * (Hz(i, j, k) - Hz(i - 1, j, k));
                    Ey(i, j, k) -= C4 * P(i, j, k);
                }
                for (int i = Nx - pml[1] + 1; i < Nx; i++) {
                    int n = i - (Nx - pml[1]) + pml[0];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(n, j, k) = bE[0][i] * P(n, j, k) + cE[0][i] * inv_Dx * (Hz(i, j, k) - Hz(i - 1, j, k));
                    Ey(i, j, k) -= C4 * P(n, j, k);
                }
            }
        for (int j = j1; j < j2; j++)
            for (int k = 0; k < Nz; k++) Ey(0, j, k) = 0;
    }
    if (pml[4] || pml[5]) {
        auto P = [&](int i, int j, int n) -> double & { return PsiEy[1][(size_t)i + (size_t)psz[0] * ((size_t)j + (size_t)psz[1] * n)]; };
        for (int k = 1; k < pml[4]; k++)
            for (int j = j1; j < j2; j++)
                for (int i = 0; i < Nx; i++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[2][k] * P(i, j, k) + cE[2][k] * inv_Dz * (Hx(i, j, k) - Hx(i, j, k - 1));
                    Ey(i, j, k) += C4 * P(i, j, k);
                }
        for (int k = Nz - pml[5] + 1; k < Nz; k++)
            for (int j = j1; j < j2; j++)
                for (int i = 0; i < Nx; i++) {
                    int n = k - (Nz - pml[5]) + pml[4];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, n) = bE[2][k] * P(i, j, n) + cE[2][k] * inv_Dz * (Hx(i, j, k) - Hx(i, j, k - 1));
                    Ey(i, j, k) += C4 * P(i, j, n);
                }
        for (int i = 0; i < Nx; i++)
            for (int j = j1; j < j2; j++) Ey(i, j, 0) = 0;
    }
}

void aeoracle_grid::app_pml_Ez(int k1, int k2) {
    double inv_Dx = 1.0 / Dx, inv_Dy = 1.0 / Dy;
    if (pml[0] || pml[1]) {
        auto P = [&](int n, int j, int k) -> double & { return PsiEz[0][(size_t)n + (size_t)psx[0] * ((size_t)j + (size_t)psx[1] * k)]; };
        for (int j = 0; j < Ny; j++)
            for (int k = k1; k < k2; k++) {
                for (int i = 1; i < pml[0]; i++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[0][i] * P(i, j, k) + cE[0][i] * inv_Dx * (Hy(i, j, k) - Hy(i - 1, j, k));
                    Ez(i, j, k) += C4 * P(i, j, k);
                }
                for (int i = Nx - pml[1] + 1; i < Nx; i++) {
                    int n = i - (Nx - pml[1]) + pml[0];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(n, j, k) = bE[0][i] * P(n, j, k) + cE[0][i] * inv_Dx * (Hy(i, j, k) - Hy(i - 1, j, k));
                    Ez(i, j, k) += C4 * P(n, j, k);
                }
            }
        for (int j = 0; j < Ny; j++)
            for (int k = k1; k < k2; k++) Ez(0, j, k) = 0;
    }
    if (pml[2] || pml[3]) {
        auto P = [&](int i, int n, int k) -> double & { return PsiEz[1][(size_t)i + (size_t)psy[0] * ((size_t)n + (size_t)psy[1] * k)]; };
        for (int i = 0; i < Nx; i++)
            for (int k = k1; k < k2; k++) {
                for (int j = 1; j < pml[2]; j++) {
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, j, k) = bE[1][j] * P(i, j, k) + cE[1][j] * inv_Dy * (Hx(i, j, k) - Hx(i, j - 1, k));
                    Ez(i, j, k) -= C4 * P(i, j, k);
                }
                for (int j = Ny - pml[3] + 1; j < Ny; j++) {
                    int n = j - (Ny - pml[3]) + pml[2];
                    double C4 = mats[mg(i, j, k)].C4;
                    P(i, n, k) = bE[1][j] * P(i, n, k) + cE[1][j] * inv_Dy * (Hx(i, j, k) - Hx(i, j - 1, k));
                    Ez(i, j, k) -= C4 * P(i, n, k);
                }
            }
        for (int i = 0; i < Nx; i++)
            for (int k = k1; k < k2; k++) Ez(i, 0, k) = 0;
    }
}

// advHx/advHy/advHz (fdtd_core.cpp:715-761,811-857,915-969); c = 0,1,2
void aeoracle_grid::advH(int c, int i1_, int i2_, int j1_, int j2_, int k1_, int k2_) {
    const bool px = pml[0] || pml[1], py_q = pml[2] || pml[1], pz = pml[4] || pml[5];
    for (int k = k1_; k < k2_; k++) {
        int kp = (k == Nz - 1) ? 0 : k + 1;
        double ikz = 1.0 / (pz ? kH[2][k] : 1.0);
        for (int j = j1_; j < j2_; j++) {
            int jp = (j == Ny - 1) ? 0 : j + 1;
            double iky = 1.0 / (py_q ? kH[1][j] : 1.0);
            for (int i = i1_; i < i2_; i++) {
                int ip = (i == Nx - 1) ? 0 : i + 1;
                double ikx = 1.0 / (px ? kH[0][i] : 1.0);
                if (c == 0)
                    Hx(i, j, k) += dtdmz * ikz * (Ey(i, j, kp) - Ey(i, j, k)) - dtdmy * iky * (Ez(i, jp, k) - Ez(i, j, k));
                else if (c == 1)
                    Hy(i, j, k) += dtdmx * ikx * (Ez(ip, j, k) - Ez(i, j, k)) - dtdmz * ikz * (Ex(i, j, kp) - Ex(i, j, k));
                else
                    Hz(i, j, k) += dtdmy * iky * (Ex(i, jp, k) - Ex(i, j, k)) - dtdmx * ikx * (Ey(ip, j, k) - Ey(i, j, k));
            }
        }
    }
}

// CPML-H (fdtd_pml.cpp:737-929)
void aeoracle_grid::app_pml_Hx(int i1, int i2) {
    double inv_Dy = 1.0 / Dy, inv_Dz = 1.0 / Dz;
    if (pml[2] || pml[3]) {
        auto P = [&](int i, int n, int k) -> double & { return PsiHx[0][(size_t)i + (size_t)psy[0] * ((size_t)n + (size_t)psy[1] * k)]; };
        for (int i = i1; i < i2; i++)
            for (int k = 0; k < Nz; k++) {
                for (int j = 0; j < pml[2]; j++) {
                    P(i, j, k) = bH[1][j] * P(i, j, k) + cH[1][j] * inv_Dy * (Ez(i, j + 1, k) - Ez(i, j, k));
                    Hx(i, j, k) -= dtm * P(i, j, k);
                }
                for (int j = Ny - pml[3]; j < Ny - 1; j++) {
                    int n = j - (Ny - pml[3]) + pml[2];
                    P(i, n, k) = bH[1][j] * P(i, n, k) + cH[1][j] * inv_Dy * (Ez(i, j + 1, k) - Ez(i, j, k));
                    Hx(i, j, k) -= dtm * P(i, n, k);
                }
            }
        for (int i = i1; i < i2; i++)
            for (int k = 0; k < Nz; k++) Hx(i, Ny - 1, k) = 0;
    }
    if (pml[4] || pml[5]) {
        auto P = [&](int i, int j, int n) -> double & { return PsiHx[1][(size_t)i + (size_t)psz[0] * ((size_t)j + (size_t)psz[1] * n)]; };
        for (int k = 0; k < pml[4]; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = i1; i < i2; i++) {
                    P(i, j, k) = bH[2][k] * P(i, j, k) + cH[2][k] * inv_Dz * (Ey(i, j, k + 1) - Ey(i, j, k));
                    Hx(i, j, k) += dtm * P(i, j, k);
                }
        for (int k = Nz - pml[5]; k < Nz - 1; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = i1; i < i2; i++) {
                    int n = k - (Nz - pml[5]) + pml[4];
                    P(i, j, n) = bH[2][k] * P(i, j, n) + cH[2][k] * inv_Dz * (Ey(i, j, k + 1) - Ey(i, j, k));
                    Hx(i, j, k) += dtm * P(i, j, n);
                }
        // Top PEC (also zeroes Hy, as the reference; app_pml_Hy re-zeroes it when Hy is enabled)
        for (int i = i1; i < i2; i++)
            for (int j = 0; j < Ny; j++) Hx(i, j, Nz - 1) = Hy(i, j, Nz - 1) = 0;
    }
}

void aeoracle_grid::app_pml_Hy(int j1, int j2) {
    double inv_Dx = 1.0 / Dx, inv_Dz = 1.0 / Dz;
    if (pml[0] || pml[1]) {
        auto P = [&](int n, int j, int k) -> double & { return PsiHy[0][(size_t)n + (size_t)psx[0] * ((size_t)j + (size_t)psx[1] * k)]; };
        for (int j = j1; j < j2; j++)
            for (int k = 0; k < Nz; k++) {
                for (int i = 0; i < pml[0]; i++) {
                    P(i, j, k) = bH[0][i] * P(i, j, k) + cH[0][i] * inv_Dx * (Ez(i + 1, j, k) - Ez(i, j, k));
                    Hy(i, j, k) += dtm * P(i, j, k);
                }
                for (int i = Nx - pml[1]; i < Nx - 1; i++) {
                    int n = i - (Nx - pml[1]) + pml[0];
                    P(n, j, k) = bH[0][i] * P(n, j, k) + cH[0][i] * inv_Dx * (Ez(i + 1, j, k) - Ez(i, j, k));
                    Hy(i, j, k) += dtm * P(n, j, k);
                }
            }
        for (int j = j1; j < j2; j++)
            for (int k = 0; k < Nz; k++) Hy(Nx - 1, j, k) = 0;
    }
    if (pml[4] || pml[5]) {
        auto P = [&](int i, int j, int n) -> double & { return PsiHy[1][(size_t)i + (size_t)psz[0] * ((size_t)j + (size_t)psz[1] * n)]; };
        for (int k = 0; k < pml[4]; k++)
            for (int j = j1; j < j2; j++)
                for (int i = 0; i < Nx; i++) {
                    P(i, j, k) = bH[2][k] * P(i, j, k) + cH[2][k] * inv_Dz * (Ex(i, j, k + 1) - Ex(i, j, k));
                    Hy(i, j, k) -= dtm * P(i, j, k);
                }
        for (int k = Nz - pml[5]; k < Nz - 1; k++)
            for (int j = j1; j < j2; j++)
                for (int i = 0; i < Nx; i++) {
                    int n = k - (Nz - pml[5]) + pml[4];
                    P(i, j, n) = bH[2][k] * P(i, j, n) + cH[2][k] * inv_Dz * (Ex(i, j, k + 1) - Ex(i, j, k));
                    Hy(i, j, k) -= dtm * P(i, j, n);
                }
        for (int i = 0; i < Nx; i++)
            for (int j = j1; j < j2; j++) Hy(i, j, Nz - 1) = 0;
    }
}

void aeoracle_grid::app_pml_Hz(int k1, int k2) {
    double inv_Dx = 1.0 / Dx, inv_Dy = 1.0 / Dy;
    if (pml[0] || pml[1]) {
        auto P = [&](int n, int j, int k) -> double & { return PsiHz[0][(size_t)n + (size_t)psx[0] * ((size_t)j + (size_t)psx[1] * k)]; };
        for (int j = 0; j < Ny; j++)
            for (int k = k1; k < k2; k++) {
                for (int i = 0; i < pml[0]; i++) {
                    P(i, j, k) = bH[0][i] * P(i, j, k) + cH[0][i] * inv_Dx * (Ey(i + 1, j, k) - Ey(i, j, k));
                    Hz(i, j, k) -= dtm * P(i, j, k);
                }
                for (int i = Nx - pml[1]; i < Nx - 1; i++) {
                    int n = i - (Nx - pml[1]) + pml[0];
                    P(n, j, k) = bH[0][i] * P(n, j, k) + cH[0][i] * inv_Dx * (Ey(i + 1, j, k) - Ey(i, j, k));
                    Hz(i, j, k) -= dtm * P(n, j, k);
                }
            }
        for (int j = 0; j < Ny; j++)
            for (int k = k1; k < k2; k++) Hz(Nx - 1, j, k) = 0;
    }
    if (pml[2] || pml[3]) {
        auto P = [&](int i, int n, int k) -> double & { return PsiHz[1][(size_t)i + (size_t)psy[0] * ((size_t)n + (size_t)psy[1] * k)]; };
        for (int i = 0; i < Nx; i++)
            for (int k = k1; k < k2; k++) {
                for (int j = 0; j < pml[2]; j++) {
                    P(i, j, k) = bH[1][j] * P(i, j, k) + cH[1][j] * inv_Dy * (Ex(i, j + 1, k) - Ex(i, j, k));
                    Hz(i, j, k) += dtm * P(i, j, k);
                }
                for (int j = Ny - pml[3]; j < Ny - 1; j++) {
                    int n = j - (Ny - pml[3]) + pml[2];
                    P(i, n, k) = bH[1][j] * P(i, n, k) + cH[1][j] * inv_Dy * (Ex(i, j + 1, k) - Ex(i, j, k));
                    Hz(i, j, k) += dtm * P(i, n, k);
                }
            }
        for (int i = 0; i < Nx; i++)
            for (int k = k1; k < k2; k++) Hz(i, Ny - 1, k) = 0;
    }
}

// TFSF E corrections of single_particle.cpp:277-332 (main grid, aux H).
void aeoracle_grid::tfsf_E() {
    const aeoracle_grid &A = *tfsf_aux;
    auto aH = [&](int c, int k) { return A.cf(3 + c, 0, 0, k); };
    int i1 = tb[0], i2 = tb[1], j1 = tb[2], j2 = tb[3], k1 = tb[4], k2 = tb[5];
    for (int i = i1; i < i2; i++)
        for (int j = j1; j < j2; j++) {
            Ex(i, j, k1) += mats[mg(i, j, k1)].C2z * aH(1, g_slip == 1 ? k1 : k1 - 1);
            Ex(i, j, k2) -= mats[mg(i, j, k2)].C2z * aH(1, k2 - 1);
            Ey(i, j, k1) -= mats[mg(i, j, k1)].C2z * aH(0, k1 - 1);
            Ey(i, j, k2) += mats[mg(i, j, k2)].C2z * aH(0, k2 - 1);
        }
    if (!tf_per_y)
        for (int i = i1; i < i2; i++)
            for (int k = k1; k < k2; k++) {
                Ex(i, j1, k) -= mats[mg(i, j1, k)].C2y * aH(2, k);
                Ex(i, j2, k) += mats[mg(i, j2, k)].C2y * aH(2, k);
                Ez(i, j1, k) += mats[mg(i, j1, k)].C2y * aH(0, k);
                Ez(i, j2, k) -= mats[mg(i, j2, k)].C2y * aH(0, k);
            }
    if (!tf_per_x)
        for (int k = k1; k < k2; k++)
            for (int j = j1; j < j2; j++) {
                Ey(i1, j, k) += mats[mg(i1, j, k)].C2x * aH(2, k);
                Ey(i2, j, k) -= mats[mg(i2, j, k)].C2x * aH(2, k);
                Ez(i1, j, k) -= mats[mg(i1, j, k)].C2x * aH(1, k);
                Ez(i2, j, k) += mats[mg(i2, j, k)].C2x * aH(1, k);
            }
}

// TFSF H corrections of single_particle.cpp:350-378 (main grid, aux E).
void aeoracle_grid::tfsf_H() {
    const aeoracle_grid &A = *tfsf_aux;
    auto aE = [&](int c, int k) { return A.cf(c, 0, 0, k); };
    int i1 = tb[0], i2 = tb[1], j1 = tb[2], j2 = tb[3], k1 = tb[4], k2 = tb[5];
    for (int i = i1; i < i2; i++)
        for (int j = j1; j < j2; j++) {
            Hx(i, j, k1 - 1) -= dtdmz * aE(1, k1);
            Hx(i, j, k2 - 1) += dtdmz * aE(1, k2);
            Hy(i, j, k1 - 1) += dtdmz * aE(0, k1);
            Hy(i, j, k2 - 1) -= dtdmz * aE(0, k2);
        }
    if (!tf_per_y)
        for (int i = i1; i < i2; i++)
            for (int k = k1; k < k2; k++) {
                Hx(i, j1 - 1, k) += dtdmy * aE(2, k);
                Hx(i, j2 - 1, k) -= dtdmy * aE(2, k);
                Hz(i, j1 - 1, k) -= (g_slip == 2 ? -dtdmy : dtdmy) * aE(0, k);
                Hz(i, j2 - 1, k) += dtdmy * aE(0, k);
            }
    if (!tf_per_x)
        for (int k = k1; k < k2; k++)
            for (int j = j1; j < j2; j++) {
                Hy(i1 - 1, j, k) -= dtdmx * aE(2, k);
                Hy(i2 - 1, j, k) += dtdmx * aE(2, k);
                Hz(i1 - 1, j, k) += dtdmx * aE(1, k);
                Hz(i2 - 1, j, k) -= dtdmx * aE(1, k);
            }
}

// update_E: phases of threaded_process_E (fdtd_threads.cpp:32-150), with the
// driver's TFSF-E between the stencil and RC-apply (single_particle.cpp:262-335)
// and the driver's plane injection after the whole update (normal_incidence.cpp:334-342).
void aeoracle_grid::update_E() {
    if (tstep == 0) pml_coeff_calc();
    const int NT = pool->size();
    auto xs = [&](int ID, auto fn) {
        if (Nx > NT) fn((ID * Nx) / NT, ((ID + 1) * Nx) / NT);
        else if (ID == 0) fn(0, Nx);
    };
    pool->run([&](int ID) { xs(ID, [&](int a, int b) { advMats_ante(a, b); }); });
    pool->run([&](int ID) {
        int x1 = (ID * Nx) / NT, x2 = ((ID + 1) * Nx) / NT;
        int y1 = (ID * Ny) / NT, y2 = ((ID + 1) * Ny) / NT;
        int z1 = (ID * Nz) / NT, z2 = ((ID + 1) * Nz) / NT;
        for (int c = 0; c < 3; c++) {
            if (!enable[c]) continue;
            if (Nz > NT) advE(c, 0, Nx, 0, Ny, z1, z2);
            else if (Ny > NT) advE(c, 0, Nx, y1, y2, 0, Nz);
            else if (Nx > NT) advE(c, x1, x2, 0, Ny, 0, Nz);
            else if (ID == 0) advE(c, 0, Nx, 0, Ny, 0, Nz);
        }
    });
    if (tfsf_aux) tfsf_E();
    pool->run([&](int ID) { xs(ID, [&](int a, int b) { advMats_simp(a, b); }); });
    // advMats_post / advMats_self do nothing for const and RC materials (mats.cpp:167-184)
    pool->run([&](int ID) {
        if (enable[0]) {
            if (Nx > NT) app_pml_Ex((ID * Nx) / NT, ((ID + 1) * Nx) / NT);
            else if (ID == 0) app_pml_Ex(0, Nx);
        }
        if (enable[1]) {
            if (Ny > NT) app_pml_Ey((ID * Ny) / NT, ((ID + 1) * Ny) / NT);
            else if (ID == 0) app_pml_Ey(0, Ny);
        }
        if (enable[2]) {
            if (Nz > NT) app_pml_Ez((ID * Nz) / NT, ((ID + 1) * Nz) / NT);
            else if (ID == 0) app_pml_Ez(0, Nz);
        }
    });
    if (inj_k >= 0 && tstep < (int)inj[0].size()) {
        double dx = inj[0][tstep], dy = inj[1][tstep];
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                Ex(i, j, inj_k) += dx;
                Ey(i, j, inj_k) += dy;
            }
    }
    // sources' inject_E (default_fdtd.cpp:158-159): Oscillator::deep_inject_E (sc_oscillator.cpp:34-67)
    for (Oscillator &o : oscs) {
        cplx im_val = std::exp(-o.w0 * o.step * o.Dt * Im) * std::exp(-std::pow(o.dw / 2.0 * ((o.step - o.tshift) * o.Dt), 2.0));
        double val = std::real(im_val) * 50;
        for (int i = o.x1; i < o.x2; i++)
            for (int j = o.y1; j < o.y2; j++)
                for (int k = o.z1; k < o.z2; k++) F[o.type][id(i, j, k)] += val;
    }
}

// update_H: threaded_process_H (fdtd_threads.cpp:165-235), then the driver's
// plane injection (normal_incidence.cpp:349-357) and TFSF-H (single_particle.cpp:350-378).
void aeoracle_grid::update_H() {
    const int NT = pool->size();
    pool->run([&](int ID) {
        int x1 = (ID * Nx) / NT, x2 = ((ID + 1) * Nx) / NT;
        int y1 = (ID * Ny) / NT, y2 = ((ID + 1) * Ny) / NT;
        int z1 = (ID * Nz) / NT, z2 = ((ID + 1) * Nz) / NT;
        for (int c = 0; c < 3; c++) {
            if (!enable[3 + c]) continue;
            if (Nz > NT) advH(c, 0, Nx, 0, Ny, z1, z2);
            else if (Ny > NT) advH(c, 0, Nx, y1, y2, 0, Nz);
            else if (Nx > NT) advH(c, x1, x2, 0, Ny, 0, Nz);
            else if (ID == 0) advH(c, 0, Nx, 0, Ny, 0, Nz);
        }
    });
    // The x-slab Hx pass and y-slab Hy pass race on Hy(:,:,Nz-1) in the reference
    // (both write 0 last), so running the three passes one after another is exact.
    pool->run([&](int ID) {
        if (enable[3]) {
            if (Nx > NT) app_pml_Hx((ID * Nx) / NT, ((ID + 1) * Nx) / NT);
            else if (ID == 0) app_pml_Hx(0, Nx);
        }
    });
    pool->run([&](int ID) {
        if (enable[4]) {
            if (Ny > NT) app_pml_Hy((ID * Ny) / NT, ((ID + 1) * Ny) / NT);
            else if (ID == 0) app_pml_Hy(0, Ny);
        }
        if (enable[5]) {
            if (Nz > NT) app_pml_Hz((ID * Nz) / NT, ((ID + 1) * Nz) / NT);
            else if (ID == 0) app_pml_Hz(0, Nz);
        }
    });
    if (inj_k >= 0 && tstep < (int)inj[2].size()) {
        double dx = inj[2][tstep], dy = inj[3][tstep];
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                Hx(i, j, inj_k) += dx;
                Hy(i, j, inj_k) += dy;
            }
    }
    if (tfsf_aux) tfsf_H();
    tstep += 1;
    // sources' inject_H: an oscillator only counts its step (sources.cpp:240-249);
    // the sensors' feed in between does not read it
    for (Oscillator &o : oscs) o.step += 1;
}

// local_Ex..local_Hz (fdtd_utils.cpp:457-507); c = 0..5
double aeoracle_grid::local(int c, int i, int j, int k) const {
    int ip = (i + 1 == Nx) ? 0 : i + 1, jp = (j + 1 == Ny) ? 0 : j + 1, kp = (k + 1 == Nz) ? 0 : k + 1;
    switch (c) {
        case 0:
            if (g_slip == 6) jp = j == 0 ? Ny - 1 : j - 1;
            return (cf(0, i, j, k) + cf(0, i, jp, k) + cf(0, i, j, kp) + cf(0, i, jp, kp)) / 4.0;
        case 1: return (cf(1, i, j, k) + cf(1, ip, j, k) + cf(1, i, j, kp) + cf(1, ip, j, kp)) / 4.0;
        case 2: return (cf(2, i, j, k) + cf(2, ip, j, k) + cf(2, i, jp, k) + cf(2, ip, jp, k)) / 4.0;
        case 3: return (cf(3, i, j, k) + cf(3, ip, j, k)) / 2.0;
        case 4: return (cf(4, i, j, k) + cf(4, i, jp, k)) / 2.0;
        default: return (cf(5, i, j, k) + cf(5, i, j, kp)) / 2.0;
    }
}

namespace {

int set_err(aeoracle_grid *g, int code, const std::string &msg) {
    if (g) g->err = msg; else g_create_error = msg;
    return code;
}

// Sensor::feed + SensorFieldHolder::deep_feed/update_t_interp/FT_comp
// (sensors.cpp:232-243,456-494,601-644): the interpolated plane is gathered on
// the calling thread, the running DFT is split over wavelengths.
void sensor_feed(aeoracle_grid &g, SpectSensor &s) {
    for (int b = 0; b < s.span2; b++)
        for (int a = 0; a < s.span1; a++) {
            int i, j, k;
            if (s.type == AEFDTD_NORMAL_X || s.type == AEFDTD_NORMAL_XM) { i = s.x1; j = s.y1 + a; k = s.z1 + b; }
            else if (s.type == AEFDTD_NORMAL_Y || s.type == AEFDTD_NORMAL_YM) { i = s.x1 + a; j = s.y1; k = s.z1 + b; }
            else { i = s.x1 + a; j = s.y1 + b; k = s.z1; }
            for (int c = 0; c < 6; c++) s.t[c][a + (size_t)s.span1 * b] = g.local(c, i, j, k);
        }
    // FT_comp over a slice of wavelengths per worker (sensors.cpp:456-479,534-550)
    auto ft = [&](int l1, int l2) {
        size_t P = (size_t)s.span1 * s.span2;
        for (int l = l1; l < l2; l++) {
            double w = 2.0 * Pi * c_light / s.lambda[l];
            cplx cE = s.tapE * std::exp(w * s.step * s.Dt * Im);
            cplx cH = s.tapH * std::exp(w * (s.step + 0.5) * s.Dt * Im);
            for (size_t p = 0; p < P; p++) {
                for (int c = 0; c < 3; c++) s.sp[c][p + P * l] += cE * s.t[c][p];
                for (int c = 3; c < 6; c++) s.sp[c][p + P * l] += cH * s.t[c][p];
            }
        }
    };
    const int NT = g.pool->size();
    g.pool->run([&](int ID) {
        if (s.Nl > NT) ft((ID * s.Nl) / NT, ((ID + 1) * s.Nl) / NT);
        else if (ID < s.Nl) ft(ID, ID + 1);
    });
    if (s.step >= s.Nt - s.Ntap) {
        s.tapE = s_curve(s.step, s.Nt - 1.0, s.Nt - s.Ntap);
        s.tapH = s_curve(s.step + 0.5, s.Nt - 1.0, s.Nt - s.Ntap);
    }
    s.step += 1;
}

// Spect_Poynting::treat (s_spec_power.cpp:44-119)
void sensor_treat(const SpectSensor &s, double *rsp) {
    size_t P = (size_t)s.span1 * s.span2;
    for (int l = 0; l < s.Nl; l++) {
        double Sx = 0, Sy = 0, Sz = 0;
        for (int a = 0; a < s.span1; a++)
            for (int b = 0; b < s.span2; b++) {
                size_t p = a + (size_t)s.span1 * b + P * l;
                const cplx Ex = s.sp[0][p], Ey = s.sp[1][p], Ez = s.sp[2][p];
                const cplx Hx = s.sp[3][p], Hy = s.sp[4][p], Hz = s.sp[5][p];
                Sx += std::real(Ey * std::conj(Hz) - Ez * std::conj(Hy));
                Sy += std::real(Ez * std::conj(Hx) - Ex * std::conj(Hz));
                Sz += std::real(Ex * std::conj(Hy) - Ey * std::conj(Hx));
            }
        double f;
        if (s.type == AEFDTD_NORMAL_X || s.type == AEFDTD_NORMAL_XM) f = s.Dy * s.Dz;
        else if (s.type == AEFDTD_NORMAL_Y || s.type == AEFDTD_NORMAL_YM) f = s.Dx * s.Dz;
        else f = s.Dx * s.Dy;
        Sx *= f; Sy *= f; Sz *= f;
        double r = 0;
        switch (s.type) {
            case AEFDTD_NORMAL_X: r = Sx; break;
            case AEFDTD_NORMAL_Y: r = Sy; break;
            case AEFDTD_NORMAL_Z: r = Sz; break;
            case AEFDTD_NORMAL_XM: r = -Sx; break;
            case AEFDTD_NORMAL_YM: r = -Sy; break;
            default: r = -Sz; break;
        }
        rsp[l] = r;
    }
}

void plane_sum_feed(aeoracle_grid &g, PlaneSum &s) {     // normal_incidence.cpp:378-415
    int t = g.tstep - 1;
    double tb = (t + 0.5) * g.Dt;
    double tap = (t >= 0 && t < (int)s.tap.size()) ? s.tap[t] : 1.0;
    double Dx2 = g.Dx / 2.0, Dy2 = g.Dy / 2.0, Dz2 = g.Dz / 2.0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double z = s.zpos;
    for (int j = 0; j < g.Ny; j++) {
        double y = j * g.Dy;
        for (int i = 0; i < g.Nx; i++) {
            double x = i * g.Dx;
            double inc[3] = {0, 0, 0};
            if (s.has_inc) {
                inc[0] = pulse_eval(s.inc, 0, x + Dx2, y, z, tb);
                inc[1] = pulse_eval(s.inc, 1, x, y + Dy2, z, tb);
                inc[2] = pulse_eval(s.inc, 2, x, y, z + Dz2, tb);
            }
            for (int c = 0; c < 3; c++) {
                acc[3 + c] += inc[c] * tap;
                if (s.subtract) acc[c] += (g.cf(c, i, j, s.k) - inc[c]) * tap;
                else acc[c] += g.cf(c, i, j, s.k) * tap;
            }
        }
    }
    for (int c = 0; c < 6; c++) s.out.push_back(acc[c]);
}

int threads_now() {
    if (g_threads > 0) return g_threads;
    int n = (int)std::thread::hardware_concurrency();
    return n > 0 ? n : 1;
}

}  // namespace

// ================================================================= C ABI
extern "C" {

void aeoracle_set_threads(int n) { g_threads = n; }
void aeoracle_set_slip(int code) { g_slip = code; }

// The full-grid material indices (x fastest), after the extension of set_matsgrid.
int aeoracle_read_matsgrid(const aeoracle_grid *g, uint32_t *out) {
    if (!g || !out) return AEFDTD_EINVAL;
    std::copy(g->matsgrid.begin(), g->matsgrid.end(), out);
    return AEFDTD_OK;
}
int aeoracle_get_threads(void) { return threads_now(); }

int aeoracle_create(aeoracle_grid **out, int Nx, int Ny, int Nz, int Nt, double Dx, double Dy,
                    double Dz, double Dt, int mode, const int pml[6], const int pad[6], int precision) {
    if (!out) return set_err(nullptr, AEFDTD_EINVAL, "null output handle");
    *out = nullptr;
    if (Nx < 1 || Ny < 1 || Nz < 1 || Nt < 0) return set_err(nullptr, AEFDTD_EINVAL, "bad grid dimensions");
    if (precision != AEFDTD_FP64) return set_err(nullptr, AEFDTD_EINVAL, "the oracle is FP64 only");
    for (int f = 0; f < 6; f++)
        if (pml[f] < 0 || pad[f] < 0) return set_err(nullptr, AEFDTD_EINVAL, "negative pml/pad");
    auto *g = new aeoracle_grid();
    g->mode = mode;
    g->Nt = Nt;
    g->Dx = Dx; g->Dy = Dy; g->Dz = Dz; g->Dt = Dt;
    for (int f = 0; f < 6; f++) {
        g->pml[f] = pml[f]; g->pad[f] = pad[f];
        g->pml_kappa[f] = 1.0; g->pml_sigma[f] = 0; g->pml_alpha[f] = 0;
    }
    // padding rules, fdtd_core.cpp:212-243 (NORM and CUSTOM modes)
    if (mode == AEFDTD_MODE_NORM || mode == AEFDTD_MODE_CUSTOM) {
        for (int a = 0; a < 3; a++)
            if (pml[2 * a] || pml[2 * a + 1]) {
                g->pad[2 * a] = std::max(5, g->pad[2 * a]);
                g->pad[2 * a + 1] = std::max(5, g->pad[2 * a + 1]);
            }
    }
    g->Nx_s = Nx; g->Ny_s = Ny; g->Nz_s = Nz;
    g->Nx = Nx + pml[0] + g->pad[0] + g->pad[1] + pml[1];
    g->Ny = Ny + pml[2] + g->pad[2] + g->pad[3] + pml[3];
    g->Nz = Nz + pml[4] + g->pad[4] + g->pad[5] + pml[5];
    g->xs_s = pml[0] + g->pad[0]; g->xs_e = g->xs_s + Nx;
    g->ys_s = pml[2] + g->pad[2]; g->ys_e = g->ys_s + Ny;
    g->zs_s = pml[4] + g->pad[4]; g->zs_e = g->zs_s + Nz;
    size_t N = (size_t)g->Nx * g->Ny * g->Nz;
    for (int c = 0; c < 6; c++) g->F[c].assign(N, 0.0);
    g->matsgrid.assign(N, 0);
    g->mats.assign(1, Mat());
    // basic_differentials_compute, fdtd_core.cpp:1137-1148
    g->dte = Dt / e0; g->dtm = Dt / mu0;
    g->dtdex = Dt / (Dx * e0); g->dtdey = Dt / (Dy * e0); g->dtdez = Dt / (Dz * e0);
    g->dtdmx = Dt / (Dx * mu0); g->dtdmy = Dt / (Dy * mu0); g->dtdmz = Dt / (Dz * mu0);
    g->allocate_pml();
    g->pml_coeff_calc();
    g->pool = new Pool(threads_now());
    *out = g;
    return AEFDTD_OK;
}

int aeoracle_destroy(aeoracle_grid *g) {
    if (!g) return AEFDTD_OK;
    for (auto *s : g->sensors) delete s;
    for (auto *s : g->fsensors) delete s;
    for (auto *s : g->sums) delete s;
    delete g->pool;
    delete g;
    return AEFDTD_OK;
}

const char *aeoracle_last_error(const aeoracle_grid *g) { return g ? g->err.c_str() : g_create_error.c_str(); }

int aeoracle_layout(const aeoracle_grid *g, int out[16]) {
    int v[16] = {g->Nx, g->Ny, g->Nz, g->xs_s, g->xs_e, g->ys_s, g->ys_e, g->zs_s, g->zs_e,
                 g->pad[0], g->pad[1], g->pad[2], g->pad[3], g->pad[4], g->pad[5], g->tstep};
    std::memcpy(out, v, sizeof(v));
    return AEFDTD_OK;
}

int aeoracle_set_pml(aeoracle_grid *g, int face, double kap, double sig, double alp) {
    if (face < 0 || face > 5) return set_err(g, AEFDTD_EINVAL, "bad face");
    double D = face < 2 ? g->Dx : (face < 4 ? g->Dy : g->Dz);
    g->pml_kappa[face] = kap;
    g->pml_sigma[face] = pml_sig_scale(D) * sig;
    g->pml_alpha[face] = alp;
    return AEFDTD_OK;
}

int aeoracle_set_tapering(aeoracle_grid *g, int Ntap) { g->Ntap = Ntap; return AEFDTD_OK; }

int aeoracle_set_matsgrid(aeoracle_grid *g, const uint32_t *ids, int nx, int ny, int nz) {
    if (nx != g->Nx_s || ny != g->Ny_s || nz != g->Nz_s) return set_err(g, AEFDTD_EINVAL, "matsgrid size mismatch");
    aeoracle_grid &G = *g;
    auto M = [&](int i, int j, int k) -> uint32_t & { return G.matsgrid[G.id(i, j, k)]; };
    for (int i = G.xs_s; i < G.xs_e; i++)
        for (int j = G.ys_s; j < G.ys_e; j++)
            for (int k = G.zs_s; k < G.zs_e; k++)
                M(i, j, k) = ids[(size_t)(i - G.xs_s) + (size_t)nx * ((size_t)(j - G.ys_s) + (size_t)ny * (k - G.zs_s))];
    for (int i = 0; i < G.xs_s; i++)
        for (int j = G.ys_s; j < G.ys_e; j++)
            for (int k = G.zs_s; k < G.zs_e; k++) M(i, j, k) = M(G.xs_s, j, k);
    for (int i = G.xs_e; i < G.Nx; i++)
        for (int j = G.ys_s; j < G.ys_e; j++)
            for (int k = G.zs_s; k < G.zs_e; k++) M(i, j, k) = M(g_slip == 4 ? G.xs_e - 2 : G.xs_e - 1, j, k);
    for (int i = 0; i < G.Nx; i++)
        for (int j = 0; j < G.ys_s; j++)
            for (int k = G.zs_s; k < G.zs_e; k++) M(i, j, k) = M(i, G.ys_s, k);
    for (int i = 0; i < G.Nx; i++)
        for (int j = G.ys_e; j < G.Ny; j++)
            for (int k = G.zs_s; k < G.zs_e; k++) M(i, j, k) = M(i, G.ys_e - 1, k);
    for (int i = 0; i < G.Nx; i++)
        for (int j = 0; j < G.Ny; j++)
            for (int k = 0; k < G.zs_s; k++) M(i, j, k) = M(i, j, G.zs_s);
    for (int i = 0; i < G.Nx; i++)
        for (int j = 0; j < G.Ny; j++)
            for (int k = G.zs_e; k < G.Nz; k++) M(i, j, k) = M(i, j, G.zs_e - 1);
    unsigned nmat = G.Nmat;
    for (uint32_t v : G.matsgrid) nmat = std::max(nmat, v + 1);
    G.Nmat = nmat;
    G.mats.assign(nmat, Mat());
    return AEFDTD_OK;
}

int aeoracle_set_material(aeoracle_grid *g, unsigned int index, const aefdtd_material *m) {
    aeoracle_grid &G = *g;
    if (index >= G.Nmat) return AEFDTD_OK;                          // not in grid
    int x1 = G.Nx, x2 = -1, y1 = G.Ny, y2 = -1, z1 = G.Nz, z2 = -1;   // link_grid bbox (mats.cpp:394-443)
    for (int k = 0; k < G.Nz; k++)
        for (int j = 0; j < G.Ny; j++)
            for (int i = 0; i < G.Nx; i++)
                if (G.mg(i, j, k) == index) {
                    x1 = std::min(x1, i); x2 = std::max(x2, i);
                    y1 = std::min(y1, j); y2 = std::max(y2, j);
                    z1 = std::min(z1, k); z2 = std::max(z2, k);
                }
    if (x2 < 0) return AEFDTD_OK;                                     // mats_in_grid false
    Mat &M = G.mats[index];
    material_setup(M, *m, G.Dx, G.Dy, G.Dz, G.Dt);
    M.x1 = x1; M.y1 = y1; M.z1 = z1;
    M.xs = x2 + 1 - x1; M.ys = y2 + 1 - y1; M.zs = z2 + 1 - z1;
    if (M.rc) M.psi.assign((size_t)M.xs * M.ys * M.zs * 3 * M.Np, cplx(0, 0));
    return AEFDTD_OK;
}

int aeoracle_disable_fields(aeoracle_grid *g, const int *fields, int n) {
    for (int i = 0; i < n; i++)
        if (fields[i] >= 0 && fields[i] < 6) g->enable[fields[i]] = false;
    return AEFDTD_OK;
}

int aeoracle_bootstrap(aeoracle_grid *g) {
    g->allocate_pml();
    g->pml_coeff_calc();
    return AEFDTD_OK;
}

int aeoracle_update_E(aeoracle_grid *g) { g->update_E(); return AEFDTD_OK; }
int aeoracle_update_H(aeoracle_grid *g) { g->update_H(); return AEFDTD_OK; }

int aeoracle_step(aeoracle_grid *g, int n) {
    for (int t = 0; t < n; t++) { g->update_E(); g->update_H(); }
    return AEFDTD_OK;
}

int aeoracle_reset_fields(aeoracle_grid *g) {
    for (int c = 0; c < 6; c++) std::fill(g->F[c].begin(), g->F[c].end(), 0.0);
    for (auto *v : {g->PsiEx, g->PsiEy, g->PsiEz, g->PsiHx, g->PsiHy, g->PsiHz})
        for (int a = 0; a < 2; a++) std::fill(v[a].begin(), v[a].end(), 0.0);
    for (auto &M : g->mats) std::fill(M.psi.begin(), M.psi.end(), cplx(0, 0));
    return AEFDTD_OK;
}

int aeoracle_read_field(aeoracle_grid *g, int f, double *out) {
    if (f < 0 || f > 5) return set_err(g, AEFDTD_EINVAL, "bad field id");
    std::memcpy(out, g->F[f].data(), g->F[f].size() * sizeof(double));
    return AEFDTD_OK;
}

int aeoracle_write_field(aeoracle_grid *g, int f, const double *in) {
    if (f < 0 || f > 5) return set_err(g, AEFDTD_EINVAL, "bad field id");
    std::memcpy(g->F[f].data(), in, g->F[f].size() * sizeof(double));
    return AEFDTD_OK;
}

int aeoracle_set_plane_injection(aeoracle_grid *g, int k, int n, const double *dEx, const double *dEy,
                                 const double *dHx, const double *dHy) {
    if (k < 0 || k >= g->Nz || n < 0) return set_err(g, AEFDTD_EINVAL, "bad injection plane");
    g->inj_k = k;
    g->inj[0].assign(dEx, dEx + n); g->inj[1].assign(dEy, dEy + n);
    g->inj[2].assign(dHx, dHx + n); g->inj[3].assign(dHy, dHy + n);
    return AEFDTD_OK;
}

int aeoracle_set_tfsf(aeoracle_grid *g, const aeoracle_grid *aux, int i1, int i2, int j1, int j2,
                      int k1, int k2, int px, int py) {
    if ((!px && (i1 < 1 || i2 >= g->Nx)) || (!py && (j1 < 1 || j2 >= g->Ny)) || i1 < 0 || j1 < 0 ||
        i2 > g->Nx || j2 > g->Ny || k1 < 1 || k2 >= g->Nz || aux->Nz != g->Nz)
        return set_err(g, AEFDTD_EINVAL, "bad TFSF box");
    g->tfsf_aux = aux;
    int b[6] = {i1, i2, j1, j2, k1, k2};
    std::memcpy(g->tb, b, sizeof(b));
    g->tf_per_x = px != 0; g->tf_per_y = py != 0;
    return AEFDTD_OK;
}

static int add_field(aeoracle_grid *g, FieldSens *s, int x1, int x2, int y1, int y2, int z1, int z2, double lambda,
                     int *id) {
    // Sensor::link offsets and clamps (sensors.cpp:214-224)
    s->x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); s->x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);
    s->y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); s->y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    s->z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); s->z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    if (s->x2 <= s->x1 || s->y2 <= s->y1 || s->z2 <= s->z1 || !(lambda > 0)) {
        delete s;
        return set_err(g, AEFDTD_EINVAL, "empty sensor");
    }
    s->lambda = lambda;                   // set_spectrum(lambda) (sensors.cpp:254-263)
    s->Dt = g->Dt;
    if (s->block) {                       // FieldBlock::initialize (s_fieldmap.cpp:177-191)
        s->span1 = s->x2 - s->x1; s->span2 = s->y2 - s->y1; s->span3 = s->z2 - s->z1;
        for (int c = 0; c < 6; c++) s->acc[c].assign((size_t)s->span1 * s->span2 * s->span3, cplx(0, 0));
    } else {                              // FieldMap::initialize (:433-446)
        if (s->type == AEFDTD_NORMAL_XM) s->type = AEFDTD_NORMAL_X;
        else if (s->type == AEFDTD_NORMAL_YM) s->type = AEFDTD_NORMAL_Y;
        else if (s->type == AEFDTD_NORMAL_ZM) s->type = AEFDTD_NORMAL_Z;
        if (s->type == AEFDTD_NORMAL_X) { s->span1 = s->y2 - s->y1; s->span2 = s->z2 - s->z1; s->span3 = s->x2 - s->x1; }
        if (s->type == AEFDTD_NORMAL_Y) { s->span1 = s->x2 - s->x1; s->span2 = s->z2 - s->z1; s->span3 = s->y2 - s->y1; }
        if (s->type == AEFDTD_NORMAL_Z) { s->span1 = s->x2 - s->x1; s->span2 = s->y2 - s->y1; s->span3 = s->z2 - s->z1; }
        for (int c = 0; c < 3; c++) s->acc[c].assign((size_t)s->span1 * s->span2, cplx(0, 0));
    }
    g->fsensors.push_back(s);
    *id = (int)g->fsensors.size() - 1;
    return AEFDTD_OK;
}

int aeoracle_add_field_block(aeoracle_grid *g, int x1, int x2, int y1, int y2, int z1, int z2, double lambda,
                             int *id) {
    auto *s = new FieldSens();
    s->block = true;
    return add_field(g, s, x1, x2, y1, y2, z1, z2, lambda, id);
}

int aeoracle_add_field_map(aeoracle_grid *g, int orientation, int x1, int x2, int y1, int y2, int z1, int z2,
                           double lambda, int cumulative, int magnetic, int *id) {
    if (orientation < 0 || orientation > 5) return set_err(g, AEFDTD_EINVAL, "bad orientation");
    auto *s = new FieldSens();
    s->block = false;
    s->type = orientation;
    s->cumulative = cumulative != 0;
    s->mag_map = magnetic != 0;
    return add_field(g, s, x1, x2, y1, y2, z1, z2, lambda, id);
}

int aeoracle_field_sensor_read(aeoracle_grid *g, int id, double *out, int dims[4]) {
    if (!g || !dims || id < 0 || id >= (int)g->fsensors.size()) return AEFDTD_EINVAL;
    const FieldSens &s = *g->fsensors[id];
    const int nc = s.block ? 6 : 3;
    dims[0] = nc; dims[1] = s.span1; dims[2] = s.span2; dims[3] = s.block ? s.span3 : 1;
    if (!out) return AEFDTD_OK;
    size_t n = s.acc[0].size(), o = 0;
    for (int c = 0; c < nc; c++)
        for (size_t q = 0; q < n; q++) { out[o++] = s.acc[c][q].real(); out[o++] = s.acc[c][q].imag(); }
    return AEFDTD_OK;
}

// FieldBlock::deep_feed / FieldMap::deep_feed at step 0 (s_fieldmap.cpp:100-110, 331-345)
int aeoracle_field_sensor_mats(aeoracle_grid *g, int id, uint32_t *out) {
    if (!g || !out || id < 0 || id >= (int)g->fsensors.size()) return AEFDTD_EINVAL;
    const FieldSens &s = *g->fsensors[id];
    if (s.block) {
        for (int k = 0; k < s.span3; k++)
            for (int j = 0; j < s.span2; j++)
                for (int i = 0; i < s.span1; i++)
                    out[(size_t)i + (size_t)s.span1 * ((size_t)j + (size_t)s.span2 * k)] = g->mg(s.x1 + i, s.y1 + j, s.z1 + k);
        return AEFDTD_OK;
    }
    for (int j = 0; j < s.span2; j++)
        for (int i = 0; i < s.span1; i++) {
            int a = 0, b = 0, c = 0;
            if (s.type == AEFDTD_NORMAL_X) { b = i; c = j; }
            else if (s.type == AEFDTD_NORMAL_Y) { a = i; c = j; }
            else { a = i; b = j; }
            out[(size_t)i + (size_t)s.span1 * j] = g->mg(s.x1 + a, s.y1 + b, s.z1 + c);
        }
    return AEFDTD_OK;
}

int aeoracle_add_spect_poynting(aeoracle_grid *g, int type, int x1, int x2, int y1, int y2, int z1,
                                int z2, int Nl, const double *lambda, int *id) {
    auto *s = new SpectSensor();
    s->type = type;
    // Sensor::link offsets and clamps (sensors.cpp:214-224)
    s->x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); s->x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);
    s->y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); s->y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    s->z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); s->z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    if (type == AEFDTD_NORMAL_X || type == AEFDTD_NORMAL_XM) { s->span1 = s->y2 - s->y1; s->span2 = s->z2 - s->z1; }
    else if (type == AEFDTD_NORMAL_Y || type == AEFDTD_NORMAL_YM) { s->span1 = s->x2 - s->x1; s->span2 = s->z2 - s->z1; }
    else { s->span1 = s->x2 - s->x1; s->span2 = s->y2 - s->y1; }
    if (s->span1 <= 0 || s->span2 <= 0 || Nl < 1) { delete s; return set_err(g, AEFDTD_EINVAL, "empty sensor"); }
    s->Nl = Nl;
    s->lambda.assign(lambda, lambda + Nl);
    s->Nt = g->Nt; s->Ntap = g->Ntap;
    s->Dx = g->Dx; s->Dy = g->Dy; s->Dz = g->Dz; s->Dt = g->Dt;
    size_t P = (size_t)s->span1 * s->span2;
    for (int c = 0; c < 6; c++) { s->t[c].assign(P, 0.0); s->sp[c].assign(P * Nl, cplx(0, 0)); }
    g->sensors.push_back(s);
    *id = (int)g->sensors.size() - 1;
    return AEFDTD_OK;
}

// generate_fdtd_source for SOURCE_GEN_OSCILLATOR (sources/sources.cpp:323-348):
// Oscillator ctor, set_spectrum, Source::link (offset by xs_s.., clamp, :251-283)
// and Oscillator::initialize (sc_oscillator.cpp:70-84).
int aeoracle_add_oscillator(aeoracle_grid *g, int type, int x1, int x2, int y1, int y2, int z1, int z2,
                            double lambda_min, double lambda_max, int *id) {
    *id = -1;
    if (type < 0 || type > 6) return set_err(g, AEFDTD_EINVAL, "bad oscillator orientation");
    if (type == 6) return AEFDTD_OK;
    Oscillator o;
    o.type = type / 2;
    o.x1 = std::clamp(x1 + g->xs_s, 0, g->Nx); o.x2 = std::clamp(x2 + g->xs_s, 0, g->Nx);
    o.y1 = std::clamp(y1 + g->ys_s, 0, g->Ny); o.y2 = std::clamp(y2 + g->ys_s, 0, g->Ny);
    o.z1 = std::clamp(z1 + g->zs_s, 0, g->Nz); o.z2 = std::clamp(z2 + g->zs_s, 0, g->Nz);
    o.Dt = g->Dt;
    double alp = std::sqrt(0.1);
    double bet = 1e-7;
    double w1 = 2.0 * Pi * c_light / lambda_max;    // m_to_rad_Hz (common/phys_tools.cpp:159)
    double w2 = 2.0 * Pi * c_light / lambda_min;
    o.w0 = (w1 + w2) / 2.0;
    o.dw = std::abs(w2 - w1) / (2.0 * std::sqrt(-std::log(alp)));
    o.tshift = std::sqrt(-4.0 / (o.dw * o.dw) * std::log(bet)) / o.Dt;
    g->oscs.push_back(o);
    *id = (int)g->oscs.size() - 1;
    return AEFDTD_OK;
}

// FieldBlock::threaded_computation (s_fieldmap.cpp:196-229) and
// FieldMap::FT_compute (:352-431), then Sensor::feed's step += 1 (sensors.cpp:232-243).
static void field_feed(const aeoracle_grid &g, FieldSens &s) {
    const double w = 2.0 * Pi * c_light / s.lambda;
    const cplx I(0.0, 1.0);
    if (s.block) {
        cplx tcoeff = std::exp(w * s.step * s.Dt * I);
        cplx tcoeff2 = std::exp(w * (s.step + 0.5) * s.Dt * I);
        for (int i = 0; i < s.span1; i++)
            for (int j = 0; j < s.span2; j++)
                for (int k = 0; k < s.span3; k++) {
                    size_t q = (size_t)i + (size_t)s.span1 * ((size_t)j + (size_t)s.span2 * k);
                    for (int c = 0; c < 3; c++) s.acc[c][q] += g.local(c, i + s.x1, j + s.y1, k + s.z1) * tcoeff;
                    for (int c = 3; c < 6; c++) s.acc[c][q] += g.local(c, i + s.x1, j + s.y1, k + s.z1) * tcoeff2;
                }
    } else {
        cplx tf_coeff = std::exp(w * s.step * s.Dt * I);
        const int o = s.mag_map ? 3 : 0;
        for (int j = 0; j < s.span2; j++)
            for (int i = 0; i < s.span1; i++) {
                int a = 0, b = 0, c = 0;
                if (s.type == AEFDTD_NORMAL_X) { b = i; c = j; }
                else if (s.type == AEFDTD_NORMAL_Y) { a = i; c = j; }
                else { a = i; b = j; }
                double S[3] = {0, 0, 0};
                if (s.cumulative) {
                    for (int k = 0; k < s.span3; k++) {
                        int aa = a, bb = b, cc = c;
                        if (s.type == AEFDTD_NORMAL_X) aa = k;
                        else if (s.type == AEFDTD_NORMAL_Y) bb = k;
                        else cc = k;
                        for (int q = 0; q < 3; q++) S[q] += g.local(o + q, s.x1 + aa, s.y1 + bb, s.z1 + cc);
                    }
                } else {
                    for (int q = 0; q < 3; q++) S[q] = g.local(o + q, s.x1 + a, s.y1 + b, s.z1 + c);
                }
                size_t p = (size_t)i + (size_t)s.span1 * j;
                for (int q = 0; q < 3; q++) s.acc[q][p] += tf_coeff * S[q];
            }
    }
    s.step += 1;
}

int aeoracle_feed_sensors(aeoracle_grid *g) {
    for (auto *s : g->sensors) sensor_feed(*g, *s);
    for (auto *s : g->fsensors) field_feed(*g, *s);
    for (auto *s : g->sums) plane_sum_feed(*g, *s);
    return AEFDTD_OK;
}

int aeoracle_sensor_poynting(aeoracle_grid *g, int id, double *rsp) {
    if (id < 0 || id >= (int)g->sensors.size()) return set_err(g, AEFDTD_EINVAL, "bad sensor id");
    sensor_treat(*g->sensors[id], rsp);
    return AEFDTD_OK;
}

int aeoracle_sensor_spectra(aeoracle_grid *g, int id, double *out, int dims[3]) {
    if (id < 0 || id >= (int)g->sensors.size()) return set_err(g, AEFDTD_EINVAL, "bad sensor id");
    const SpectSensor &s = *g->sensors[id];
    dims[0] = s.span1; dims[1] = s.span2; dims[2] = s.Nl;
    if (out) {
        size_t n = (size_t)s.span1 * s.span2 * s.Nl;
        for (int c = 0; c < 6; c++)
            for (size_t p = 0; p < n; p++) {
                out[2 * (c * n + p)] = s.sp[c][p].real();
                out[2 * (c * n + p) + 1] = s.sp[c][p].imag();
            }
    }
    return AEFDTD_OK;
}

int aeoracle_add_plane_sum(aeoracle_grid *g, int k, const aefdtd_pulse *inc, double zpos, int subtract,
                           const double *tap, int *id) {
    if (k < 0 || k >= g->Nz) return set_err(g, AEFDTD_EINVAL, "bad plane");
    auto *s = new PlaneSum();
    s->k = k;
    s->has_inc = inc != nullptr;
    s->subtract = subtract != 0 && inc != nullptr;
    if (inc) s->inc = *inc;
    s->zpos = zpos;
    s->tap.assign(tap, tap + g->Nt);
    g->sums.push_back(s);
    *id = (int)g->sums.size() - 1;
    return AEFDTD_OK;
}

int aeoracle_read_plane_sum(aeoracle_grid *g, int id, int n, double *out) {
    if (id < 0 || id >= (int)g->sums.size()) return set_err(g, AEFDTD_EINVAL, "bad sum id");
    const auto &v = g->sums[id]->out;
    size_t m = std::min((size_t)n * 6, v.size());
    std::memcpy(out, v.data(), m * sizeof(double));
    return AEFDTD_OK;
}

int aeoracle_pulse_init(aefdtd_pulse *p, double li, double lf, double inc, double pol, double dir,
                        double sx, double sy, double sz, double n_ind) {
    pulse_init(*p, li, lf, inc, pol, dir, sx, sy, sz, n_ind);
    return AEFDTD_OK;
}

double aeoracle_pulse_eval(const aefdtd_pulse *p, int comp, double x, double y, double z, double t) {
    return pulse_eval(*p, comp, x, y, z, t);
}

int aeoracle_material_coeffs(const aefdtd_material *m, double Dx, double Dy, double Dz, double Dt,
                             double *out, int *n_poles) {
    Poles P = material_poles(*m);
    Mat M;
    material_setup(M, *m, Dx, Dy, Dz, Dt);
    int np = (int)P.a.size();
    *n_poles = np;
    if (!out) return AEFDTD_OK;
    for (int p = 0; p < np; p++) {
        double *o = out + 10 * p;
        o[0] = P.a[p].real(); o[1] = P.a[p].imag(); o[2] = P.b[p].real(); o[3] = P.b[p].imag();
        o[4] = M.Crec[p].real(); o[5] = M.Crec[p].imag(); o[6] = M.chi[p].real(); o[7] = M.chi[p].imag();
        o[8] = M.Dchi[p].real(); o[9] = M.Dchi[p].imag();
    }
    double *o = out + 10 * np;
    o[0] = M.C1; o[1] = M.C2x; o[2] = M.C2y; o[3] = M.C2z; o[4] = M.C3; o[5] = M.C4; o[6] = M.ei; o[7] = M.sig;
    return AEFDTD_OK;
}

int aeoracle_material_eval(const aefdtd_material *m, double w, double *re, double *im) {   // dielec_models.cpp:52-64
    cplx r = 0;
    r += m->eps_inf;
    for (int i = 0; i < m->n_debye; i++) r += m->debye[2 * i] / (1.0 - w * m->debye[2 * i + 1] * Im);
    for (int i = 0; i < m->n_drude; i++) {
        double wd = m->drude[2 * i], g = m->drude[2 * i + 1];
        r += -(wd * wd) / (w * w + w * g * Im);
    }
    for (int i = 0; i < m->n_lorentz; i++) {
        double A = m->lorentz[3 * i], O = m->lorentz[3 * i + 1], G = m->lorentz[3 * i + 2];
        r += A * (O * O) / ((O * O) - w * w - w * G * Im);
    }
    for (int i = 0; i < m->n_critpoint; i++) {
        double A = m->critpoint[4 * i], O = m->critpoint[4 * i + 1], P = m->critpoint[4 * i + 2], G = m->critpoint[4 * i + 3];
        r += A * O * (std::exp(P * Im) / (O - w - G * Im) + std::exp(-P * Im) / (O + w + G * Im));
    }
    *re = r.real(); *im = r.imag();
    return AEFDTD_OK;
}

int aeoracle_pml_profile(int N, int pm, int pp, double D, double Dt, const double kap[2], const double sig[2],
                         const double alp[2], double ind, double out[3]) {
    double s[2] = {pml_sig_scale(D) * sig[0], pml_sig_scale(D) * sig[1]};
    double b = 0, c = 0, k = 1;
    pml_coeffs(ind, N, pm, pp, kap, s, alp, Dt, k, b, c);
    out[0] = k; out[1] = b; out[2] = c;
    return AEFDTD_OK;
}

double aeoracle_s_curve(double x, double xmin, double xmax) { return s_curve(x, xmin, xmax); }
int aeoracle_nearest_integer(double x) { return nearest_integer(x); }

// ---------------------------------------------------------------- drivers
namespace {

struct GridGuard {
    aeoracle_grid *g = nullptr;
    ~GridGuard() { aeoracle_destroy(g); }
};

void lambda_grid(int Nl, double lmin, double lmax, std::vector<double> &L) {   // normal_incidence.cpp:235
    L.resize(Nl);
    for (int l = 0; l < Nl; l++) L[l] = lmin + (lmax - lmin) * l / (Nl - 1.0);
}

// Sensors of a driver: a box becomes 6 planar sensors (Box_Spect_Poynting,
// s_spec_power.cpp:125-232); `first` records the planar ids per request.
struct SensorSet {
    std::vector<std::vector<int>> ids;   // per request: 1 (planar) or 6 (box, -1 when disabled)
};

int add_requests(aeoracle_grid *g, int n, const aefdtd_sensor_req *req, SensorSet &S) {
    for (int r = 0; r < n; r++) {
        const aefdtd_sensor_req &q = req[r];
        std::vector<int> ids;
        if (q.kind < AEFDTD_SENSOR_POYNTING || q.kind > AEFDTD_SENSOR_FIELDMAP2) return AEFDTD_EINVAL;
        if (q.kind == AEFDTD_SENSOR_FIELDBLOCK || q.kind == AEFDTD_SENSOR_FIELDMAP) {
            // FieldBlock / FieldMap at lambda_min (sensors.cpp:362-378)
            int id = -1;
            int rc = q.kind == AEFDTD_SENSOR_FIELDBLOCK
                         ? aeoracle_add_field_block(g, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, &id)
                         : aeoracle_add_field_map(g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, 0, 0, &id);
            if (rc) return rc;
            ids.push_back(id);
            S.ids.push_back(ids);
            continue;
        }
        if (q.Nl < 1) return AEFDTD_EINVAL;
        std::vector<double> L;
        lambda_grid(q.Nl, q.lambda_min, q.lambda_max, L);
        if (q.box && q.kind == AEFDTD_SENSOR_POYNTING) {
            const int t[6] = {AEFDTD_NORMAL_XM, AEFDTD_NORMAL_X, AEFDTD_NORMAL_YM, AEFDTD_NORMAL_Y, AEFDTD_NORMAL_ZM, AEFDTD_NORMAL_Z};
            const int b[6][6] = {{q.x1, q.x1 + 1, q.y1, q.y2, q.z1, q.z2}, {q.x2 - 1, q.x2, q.y1, q.y2, q.z1, q.z2},
                                 {q.x1, q.x2, q.y1, q.y1 + 1, q.z1, q.z2}, {q.x1, q.x2, q.y2 - 1, q.y2, q.z1, q.z2},
                                 {q.x1, q.x2, q.y1, q.y2, q.z1, q.z1 + 1}, {q.x1, q.x2, q.y1, q.y2, q.z2 - 1, q.z2}};
            for (int f = 0; f < 6; f++) {
                int id = -1;
                if (!q.disable[f]) {
                    int rc = aeoracle_add_spect_poynting(g, t[f], b[f][0], b[f][1], b[f][2], b[f][3], b[f][4], b[f][5], q.Nl, L.data(), &id);
                    if (rc) return rc;
                }
                ids.push_back(id);
            }
        } else {
            int id = -1;
            int rc = aeoracle_add_spect_poynting(g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.Nl, L.data(), &id);
            if (rc) return rc;
            ids.push_back(id);
        }
        S.ids.push_back(ids);
    }
    return AEFDTD_OK;
}

// Box_Spect_Poynting::treat (s_spec_power.cpp:176-232): total then per face.
void collect_requests(aeoracle_grid *g, int n, const aefdtd_sensor_req *req, SensorSet &S, double *out) {
    size_t off = 0;
    for (int r = 0; r < n; r++) {
        int Nl = req[r].Nl;
        if (req[r].kind == AEFDTD_SENSOR_FIELDBLOCK || req[r].kind == AEFDTD_SENSOR_FIELDMAP) {
            int dims[4];
            aeoracle_field_sensor_read(g, S.ids[r][0], out ? out + off : nullptr, dims);
            off += (size_t)dims[0] * dims[1] * dims[2] * dims[3] * 2;
            continue;
        }
        if (req[r].kind == AEFDTD_SENSOR_FIELDMAP2) {
            // FieldMap2: the raw spectral accumulators (s_fieldmap.cpp:624-672)
            int dims[3];
            aeoracle_sensor_spectra(g, S.ids[r][0], nullptr, dims);
            if (out) aeoracle_sensor_spectra(g, S.ids[r][0], out + off, dims);
            off += 6 * (size_t)Nl * dims[0] * dims[1] * 2;
            continue;
        }
        std::vector<double> tmp(Nl);
        if (req[r].box) {
            std::vector<double> tot(Nl, 0.0);
            std::vector<double> faces(6 * (size_t)Nl, 0.0);
            for (int f = 0; f < 6; f++) {
                if (S.ids[r][f] < 0) continue;
                aeoracle_sensor_poynting(g, S.ids[r][f], tmp.data());
                for (int l = 0; l < Nl; l++) faces[f * (size_t)Nl + l] = tmp[l];
            }
            for (int l = 0; l < Nl; l++)
                for (int f = 0; f < 6; f++)
                    if (S.ids[r][f] >= 0) tot[l] += faces[f * (size_t)Nl + l];
            if (out) {
                std::memcpy(out + off, tot.data(), Nl * sizeof(double));
                std::memcpy(out + off + Nl, faces.data(), 6 * Nl * sizeof(double));
            }
            off += 7 * (size_t)Nl;
        } else {
            aeoracle_sensor_poynting(g, S.ids[r][0], tmp.data());
            if (out) std::memcpy(out + off, tmp.data(), Nl * sizeof(double));
            off += Nl;
        }
    }
}

int setup_grid(aeoracle_grid *g, const aefdtd_mode *m) {
    for (int f = 0; f < 6; f++) aeoracle_set_pml(g, f, m->kappa[f], m->sigma[f], m->alpha[f]);
    aeoracle_set_tapering(g, m->tapering);
    int rc = aeoracle_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz);
    if (rc) return rc;
    for (int i = 0; i < m->n_materials; i++) aeoracle_set_material(g, i, &m->materials[i]);
    aeoracle_disable_fields(g, m->disable_fields, m->n_disable);
    return AEFDTD_OK;
}

void copy_fields(aeoracle_grid *g, double *out) {
    if (!out) return;
    size_t N = g->F[0].size();
    for (int c = 0; c < 6; c++) std::memcpy(out + c * N, g->F[c].data(), N * sizeof(double));
}

// CompletionSensor (sensors/s_completion.cpp), fed after the other sensors
// every step; the drivers check it when t % cc_step == 0 and it never
// completes before its step count (t + 1 after the feed of step t) reaches
// twice the longest axis (:39-41).
//  Energy mode (:122-128,160-180): every feed sums Ex^2+Ey^2+Ez^2 over the
//  whole grid in k, j, i order and keeps the running maximum; complete when
//  last/max < cc_coeff.
//  FT mode (:26-36,43-120,132-158,232-398): cc_quant cells placed on the
//  structure region's faces by cc_layout (or in its volume), at wavelengths
//  drawn in [cc_lmin, cc_lmax] from a std::mt19937 (randp, common/mathUT.cpp:
//  65-68); each feed adds local_E * exp(i w step Dt) to a running DFT per cell
//  and keeps its last navg values; complete when 90% of the cells' largest
//  averaged series vary by at most cc_coeff of their average over the ring
//  (the six averages carry over from one cell to the next, as written).
std::vector<double> g_energy_trace;   // every energy feed's sum, for the tests
std::vector<int> g_ft_pts;            // the last FT run's cells (i, j, k), wavelengths, navg, ring, checks
std::vector<double> g_ft_lambda, g_ft_ring;
std::vector<int> g_ft_checks;         // per check: sensor step count, settled cells
int g_ft_navg = 0;

struct Completion {
    const aefdtd_mode *m;
    const aeoracle_grid *g;
    double last = 0, max = 0;
    // FT mode
    int np = 0, navg = 0;
    std::vector<int> I, J, K;
    std::vector<double> w;
    std::vector<std::complex<double>> acc;   // [3 * np]
    std::vector<double> ring;                // [6][np][navg]
    double Dt = 0;

    Completion(const aefdtd_mode *m_, const aeoracle_grid *g_, double Dt_) : m(m_), g(g_), Dt(Dt_) {
        g_energy_trace.clear();
        if (m->time_type == 2) place();
    }
    bool on() const { return m->time_type != 0; }

    static double randp(std::mt19937 &mte, double A, double B) {
        return A + (B - A) * mte() / (mte.max() - mte.min());
    }

    void place() {   // CompletionSensor::initialize (s_completion.cpp:232-398)
        std::mt19937 mte(m->cc_seed);
        std::string layout(m->cc_layout, strnlen(m->cc_layout, 4));
        bool known = layout == "v";
        if (layout.size() == 3) {
            known = true;
            for (char c : layout) known = known && (c == 'n' || c == 'd' || c == 'u' || c == 'b');
        }
        if (layout == "nnn" || !known) layout = "nnb";
        np = m->cc_quant;
        I.assign(np, 0); J.assign(np, 0); K.assign(np, 0);
        auto rx = [&]() { return static_cast<int>(randp(mte, g->xs_s, g->xs_e - 1) + 0.5); };
        auto ry = [&]() { return static_cast<int>(randp(mte, g->ys_s, g->ys_e - 1) + 0.5); };
        auto rz = [&]() { return static_cast<int>(randp(mte, g->zs_s, g->zs_e - 1) + 0.5); };
        if (layout == "v") {
            for (int i = 0; i < np; i++) { I[i] = rx(); J[i] = ry(); K[i] = rz(); }
        } else {
            int Np_xd = 0, Np_xu = 0, Np_yd = 0, Np_yu = 0, Np_zd = 0, Np_zu = 0;
            const int Nx = g->Nx, Ny = g->Ny, Nz = g->Nz;
            if (layout[0] == 'd') Np_xd = Ny * Nz;
            if (layout[0] == 'u') Np_xu = Ny * Nz;
            if (layout[0] == 'b') Np_xd = Np_xu = Ny * Nz;
            if (layout[1] == 'd') Np_yd = Nx * Nz;
            if (layout[1] == 'u') Np_yu = Nx * Nz;
            if (layout[1] == 'b') Np_yd = Np_yu = Nx * Nz;
            if (layout[2] == 'd') Np_zd = Nx * Ny;
            if (layout[2] == 'u') Np_zu = Nx * Ny;
            if (layout[2] == 'b') Np_zd = Np_zu = Nx * Ny;
            const double tot = Np_xd + Np_xu + Np_yd + Np_yu + Np_zd + Np_zu;
            int left = np;
            for (int *q : {&Np_xd, &Np_xu, &Np_yd, &Np_yu, &Np_zd, &Np_zu}) {
                int a = static_cast<int>(np * *q / tot + 0.5);
                a = a < left ? a : left;
                left -= a;
                *q = a;
            }
            int o = 0;
            for (int i = o; i < o + Np_xd; i++) { I[i] = g->xs_s; J[i] = ry(); K[i] = rz(); }
            o += Np_xd;
            for (int i = o; i < o + Np_xu; i++) { I[i] = g->xs_e - 1; J[i] = ry(); K[i] = rz(); }
            o += Np_xu;
            for (int i = o; i < o + Np_yd; i++) { I[i] = rx(); J[i] = g->ys_s; K[i] = rz(); }
            o += Np_yd;
            for (int i = o; i < o + Np_yu; i++) { I[i] = rx(); J[i] = g->ys_e - 1; K[i] = rz(); }
            o += Np_yu;
            for (int i = o; i < o + Np_zd; i++) { I[i] = rx(); J[i] = ry(); K[i] = g->zs_s; }
            o += Np_zd;
            for (int i = o; i < o + Np_zu; i++) { I[i] = rx(); J[i] = ry(); K[i] = g->zs_e - 1; }
        }
        std::vector<double> lam(np);
        w.assign(np, 0);
        for (int i = 0; i < np; i++) {
            lam[i] = randp(mte, m->cc_lmin, m->cc_lmax);
            w[i] = 2.0 * Pi * c_light / lam[i];             // m_to_rad_Hz (common/phys_tools.cpp:159)
        }
        const double T = std::max(m->cc_lmin, m->cc_lmax) / c_light;
        navg = static_cast<int>(T / Dt);
        acc.assign(3 * (size_t)np, 0.0);
        ring.assign(6 * (size_t)np * navg, 0.0);
        g_ft_pts.clear();
        for (int i = 0; i < np; i++) { g_ft_pts.push_back(I[i]); g_ft_pts.push_back(J[i]); g_ft_pts.push_back(K[i]); }
        g_ft_lambda = lam;
        g_ft_navg = navg;
        g_ft_checks.clear();
    }

    double local(int c, int i, int j, int k) const {   // FDTD::local_Ex..Ez (core/fdtd_utils.cpp:457-479)
        const int ip = i + 1 == g->Nx ? 0 : i + 1, jp = j + 1 == g->Ny ? 0 : j + 1, kp = k + 1 == g->Nz ? 0 : k + 1;
        auto F = [&](int x, int y, int z) { return g->F[c][(size_t)x + (size_t)g->Nx * ((size_t)y + (size_t)g->Ny * z)]; };
        if (c == 0) return (F(i, j, k) + F(i, jp, k) + F(i, j, kp) + F(i, jp, kp)) / 4.0;
        if (c == 1) return (F(i, j, k) + F(ip, j, k) + F(i, j, kp) + F(ip, j, kp)) / 4.0;
        return (F(i, j, k) + F(ip, j, k) + F(i, jp, k) + F(ip, jp, k)) / 4.0;
    }

    bool feed_and_check(const aeoracle_grid *, int t) {
        const int steps = t + 1, nmax = std::max(g->Nx, std::max(g->Ny, g->Nz));
        if (m->time_type == 1) {
            last = 0;
            const size_t N = g->F[0].size();
            for (size_t c = 0; c < N; c++) {
                const double ex = g->F[0][c], ey = g->F[1][c], ez = g->F[2][c];
                last += ex * ex + ey * ey + ez * ez;
            }
            max = std::max(max, last);
            g_energy_trace.push_back(last);
            if (t % m->cc_step) return false;
            if (steps < 2 * nmax) return false;
            return last / max < m->cc_coeff;
        }
        const std::complex<double> Im(0.0, 1.0);
        const int off = t % navg;
        for (int i = 0; i < np; i++) {
            const std::complex<double> coeff = std::exp(w[i] * t * Dt * Im);
            for (int c = 0; c < 3; c++) {
                acc[3 * i + c] += local(c, I[i], J[i], K[i]) * coeff;
                ring[((size_t)c * np + i) * navg + off] = acc[3 * i + c].real();
                ring[((size_t)(3 + c) * np + i) * navg + off] = acc[3 * i + c].imag();
            }
        }
        g_ft_ring = ring;
        if (t % m->cc_step) return false;
        if (steps < 2 * nmax) return false;
        double ar[6] = {0, 0, 0, 0, 0, 0};
        int nok = 0;
        for (int i = 0; i < np; i++) {
            for (int j = 0; j < navg; j++)
                for (int c = 0; c < 6; c++) ar[c] += ring[((size_t)c * np + i) * navg + j];
            for (int c = 0; c < 6; c++) ar[c] = std::abs(ar[c] / (navg + 0.0));
            const double amax = std::max({ar[0], ar[1], ar[2], ar[3], ar[4], ar[5]});
            int tg = 0;
            for (int c = 5; c >= 0; c--)
                if (amax == ar[c]) tg = c;                    // the first series with the largest average
            const double *row = &ring[((size_t)tg * np + i) * navg];
            double tmin = row[0], tmax = row[0];
            for (int k = 0; k < navg; k++) { tmin = std::min(tmin, row[k]); tmax = std::max(tmax, row[k]); }
            if (std::abs(tmax - tmin) <= amax * m->cc_coeff) nok++;
        }
        g_ft_checks.push_back(steps);
        g_ft_checks.push_back(nok);
        return nok / (np + 0.0) >= 0.9;
    }
};

int check_time_type(const aefdtd_mode *m) {
    if (m->time_type == 0) return AEFDTD_OK;
    if (m->time_type < 0 || m->time_type > 2 || m->cc_step < 1) return AEFDTD_EINVAL;
    if (m->time_type == 2 && (m->cc_quant < 1 || !(m->cc_lmin > 0) || !(m->cc_lmax > 0))) return AEFDTD_EINVAL;
    return AEFDTD_OK;
}

int g_last_steps = 0;

}  // namespace

int aeoracle_last_steps(void) { return g_last_steps; }

int aeoracle_ft_info(int *navg, int *np, int *nchecks) {
    *navg = g_ft_navg;
    *np = (int)g_ft_lambda.size();
    *nchecks = (int)g_ft_checks.size() / 2;
    return AEFDTD_OK;
}

int aeoracle_ft_read(int *pts, double *lambda, double *ring, int *checks) {
    if (pts) std::copy(g_ft_pts.begin(), g_ft_pts.end(), pts);
    if (lambda) std::copy(g_ft_lambda.begin(), g_ft_lambda.end(), lambda);
    if (ring) std::copy(g_ft_ring.begin(), g_ft_ring.end(), ring);
    if (checks) std::copy(g_ft_checks.begin(), g_ft_checks.end(), checks);
    return AEFDTD_OK;
}

int aeoracle_energy_trace(double *out, int n) {
    const int m = (int)g_energy_trace.size();
    for (int i = 0; out && i < std::min(n, m); i++) out[i] = g_energy_trace[i];
    return m;
}

// FDTD_normal_incidence, modes/normal_incidence.cpp:161-607
int aeoracle_run_normal_incidence(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                                  aefdtd_normal_result *res, double *sensor_out, double *fields_out,
                                  const int *end_computation) {
    if (check_time_type(m)) return AEFDTD_EINVAL;
    int Nt = m->Nt;
    double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    int pad_zm = std::max(5, m->pad[4]), pad_zp = std::max(5, m->pad[5]);
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * c_light) * 0.99 * m->time_mod;
    const int pml[6] = {0, 0, 0, 0, m->pml[4], m->pml[5]};
    const int pad[6] = {0, 0, 0, 0, pad_zm, pad_zp};
    GridGuard G;
    int rc = aeoracle_create(&G.g, m->nx, m->ny, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_NORM, pml, pad, AEFDTD_FP64);
    if (rc) return rc;
    aeoracle_grid *g = G.g;
    aeoracle_set_pml(g, 4, m->kappa[4], m->sigma[4], m->alpha[4]);
    aeoracle_set_pml(g, 5, m->kappa[5], m->sigma[5], m->alpha[5]);
    aeoracle_set_tapering(g, m->tapering);
    if ((rc = aeoracle_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz))) return rc;
    for (int i = 0; i < m->n_materials; i++) aeoracle_set_material(g, i, &m->materials[i]);
    aeoracle_disable_fields(g, m->disable_fields, m->n_disable);
    int Nx = g->Nx, Ny = g->Ny, zs_s = g->zs_s, zs_e = g->zs_e;
    int Nl = m->Nl;
    std::vector<double> Lambda, w(Nl);
    lambda_grid(Nl, m->lambda_min, m->lambda_max, Lambda);
    for (int l = 0; l < Nl; l++) w[l] = 2.0 * Pi * c_light / Lambda[l];
    double pol = (m->polarization == 1) ? 90 : 0;
    double eps_sub = g->mats[g->mg(0, 0, zs_s)].ei;
    double eps_sup = g->mats[g->mg(0, 0, zs_e)].ei;
    double index_sub = std::sqrt(eps_sub), index_sup = std::sqrt(eps_sup);
    aefdtd_pulse chp;
    pulse_init(chp, m->lambda_min, m->lambda_max, 0, pol, 0, 0, 0, (zs_e - 1) * Dz, index_sup);
    aeoracle_bootstrap(g);
    // injection tables (normal_incidence.cpp:334-357)
    std::vector<double> dEx(Nt), dEy(Nt), dHx(Nt), dHy(Nt), tap(Nt);
    for (int t = 0; t < Nt; t++) {
        double tb = t * Dt;
        double iHx = pulse_eval(chp, 3, 0, 0, (zs_e - 1 + 0.5) * Dz, tb);
        double iHy = pulse_eval(chp, 4, 0, 0, (zs_e - 1 + 0.5) * Dz, tb);
        dEx[t] = -(g->dtdez * iHy / eps_sup);
        dEy[t] = g->dtdez * iHx / eps_sup;
        tb = (t + 0.5) * Dt;
        double iEx = pulse_eval(chp, 0, 0, 0, (zs_e - 1) * Dz, tb);
        double iEy = pulse_eval(chp, 1, 0, 0, (zs_e - 1) * Dz, tb);
        dHx[t] = g->dtdmz * iEy;
        dHy[t] = -(g->dtdmz * iEx);
        tap[t] = (t >= Nt - m->tapering) ? s_curve(t, Nt - 1.0, Nt - m->tapering) : 1.0;
    }
    aeoracle_set_plane_injection(g, zs_e - 1, Nt, dEx.data(), dEy.data(), dHx.data(), dHy.data());
    SensorSet S;
    if ((rc = add_requests(g, n_sensors, sensors, S))) return rc;
    int idR, idT;
    aeoracle_add_plane_sum(g, zs_e - 2, &chp, (zs_e - 2) * Dz, 1, tap.data(), &idR);
    aeoracle_add_plane_sum(g, zs_s, &chp, zs_s * Dz, 0, tap.data(), &idT);
    Completion E(m, g, Dt);
    int t;
    bool stopped = false;
    for (t = 0; t < Nt; t++) {
        g->update_E();
        g->update_H();
        aeoracle_feed_sensors(g);
        if (E.on() && E.feed_and_check(g, t)) break;     // Nt = t (normal_incidence.cpp:417-423)
        if (cancelled(end_computation)) { stopped = true; break; }    // Nt = t (:431-435)
    }
    int Nrun = t;
    g_last_steps = (t < Nt && E.on()) ? t + 1 : t;
    std::vector<double> R(6 * (size_t)Nrun), T(6 * (size_t)Nrun);
    aeoracle_read_plane_sum(g, idR, Nrun, R.data());
    aeoracle_read_plane_sum(g, idT, Nrun, T.data());
    // find_slab (fdtd_utils.cpp:347-392)
    int Nsub = 0, Nsup = 0;
    long IDsup = 0, IDsub = 0, IDmoy;
    for (int i = g->xs_s; i < g->xs_e; i++)
        for (int j = g->ys_s; j < g->ys_e; j++) { IDsup += g->mg(i, j, zs_e); IDsub += g->mg(i, j, zs_s); }
    IDmoy = IDsub;
    int k = zs_s;
    while (IDmoy == IDsub && k < zs_e - 1) {
        IDmoy = 0;
        for (int i = g->xs_s; i < g->xs_e; i++)
            for (int j = g->ys_s; j < g->ys_e; j++) IDmoy += g->mg(i, j, k + 1);
        Nsub++; k++;
    }
    IDmoy = IDsup;
    k = zs_e - 1;
    while (IDmoy == IDsup && k >= zs_s) {
        IDmoy = 0;
        for (int i = g->xs_s; i < g->xs_e; i++)
            for (int j = g->ys_s; j < g->ys_e; j++) IDmoy += g->mg(i, j, k - 1);
        Nsup++; k--;
    }
    double hsup = (Nsup - 1) * Dz, hsub = Nsub * Dz, hstruc = (g->Nz_s - Nsup - Nsub) * Dz;
    if (g_slip != 5) {
        hsup -= Dz / 2.0;
        hsub -= Dz / 2.0;
    }
    // post-run DFT (normal_incidence.cpp:455-488); Nt is the number of steps run
    int NtE = Nrun;
    for (int l = 0; l < Nl; l++) {
        double k0 = -w[l] / c_light, k1 = k0 * index_sup, k3 = k0 * index_sub;
        cplx tmp_sup = std::exp(2.0 * k1 * hsup * Im);
        cplx tmp_sub = std::exp((k3 * hsub - k0 * (hstruc + hsub)) * Im);
        cplx r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
        for (int tt = 0; tt < NtE; tt++) {
            cplx tmp1 = std::exp(tt * Dt * w[l] * Im) / (NtE - 1.0);
            cplx tmp2 = tmp1 * tmp_sup, tmp3 = tmp1 * tmp_sub;
            for (int c = 0; c < 3; c++) {
                r1[c] += R[6 * (size_t)tt + 3 + c] * tmp1;
                r2[c] += R[6 * (size_t)tt + c] * tmp2;
                t1[c] += T[6 * (size_t)tt + 3 + c] * tmp1;
                t2[c] += T[6 * (size_t)tt + c] * tmp3;
            }
        }
        double a = std::sqrt(std::norm(r2[0]) + std::norm(r2[1]) + std::norm(r2[2]));
        double b = std::sqrt(std::norm(r1[0]) + std::norm(r1[1]) + std::norm(r1[2]));
        res->lambda[l] = Lambda[l];
        res->R[l] = a / b;
        a = std::sqrt(std::norm(t2[0]) + std::norm(t2[1]) + std::norm(t2[2]));
        b = std::sqrt(std::norm(t1[0]) + std::norm(t1[1]) + std::norm(t1[2]));
        res->T[l] = a / b;
        cplx ref0 = r1[1], trans0 = t1[1];
        if (m->polarization == 1) { ref0 = r1[0]; trans0 = t1[0]; }
        cplx v[4] = {r2[0] / ref0, r2[1] / ref0, t2[0] / trans0, t2[1] / trans0};
        double *dst[4] = {res->refx, res->refy, res->transx, res->transy};
        for (int q = 0; q < 4; q++) { dst[q][2 * l] = v[q].real(); dst[q][2 * l + 1] = v[q].imag(); }
    }
    res->index_sup = index_sup;
    res->index_sub = index_sub;
    res->Nt_run = Nrun;
    collect_requests(g, n_sensors, sensors, S, sensor_out);
    copy_fields(g, fields_out);
    (void)Nx; (void)Ny;
    return stopped ? AEFDTD_ECANCELLED : AEFDTD_OK;
}

// FDTD_single_particle, modes/single_particle.cpp:29-415
// FDTD_Mode::finalize (modes/lua_fdtd.cpp:74-203), the part that sizes the
// grid, written out axis by axis as the reference does.
int aeoracle_mode_finalize(aefdtd_mode *m, int kind, int n_sensors, const aefdtd_sensor_req *sensors) {
    if (!m) return AEFDTD_EINVAL;
    const int Nx = m->nx, Ny = m->ny, Nz = m->nz;
    if (kind == 0) {
        m->pml[0] = m->pml[1] = 0;
        m->pml[2] = m->pml[3] = 0;
        if (m->pml[4] <= 0) m->pml[4] = 25;
        if (m->pml[5] <= 0) m->pml[5] = 25;
    }
    auto near = [&](int &pad, int pml, auto lo_of, auto src_lo_of) {
        if (pml) {
            pad = std::max(pad, 5);
            for (int i = 0; i < n_sensors; i++) pad = std::max(pad, 5 - lo_of(sensors[i]));
            for (int i = 0; i < m->n_sources; i++) {
                if (m->sources[i].type == 0) pad = std::max(pad, 50 - src_lo_of(m->sources[i]));
                else pad = std::max(pad, 5 - src_lo_of(m->sources[i]));
            }
        } else pad = 0;
    };
    auto far = [&](int &pad, int pml, int N, auto hi_of, auto src_hi_of) {
        if (pml) {
            pad = std::max(pad, 5);
            for (int i = 0; i < n_sensors; i++) pad = std::max(pad, 5 + hi_of(sensors[i]) - N);
            for (int i = 0; i < m->n_sources; i++) {
                if (m->sources[i].type == 0) pad = std::max(pad, 50 + src_hi_of(m->sources[i]) - N);
                else pad = std::max(pad, 5 + src_hi_of(m->sources[i]) - N);
            }
        } else pad = 0;
    };
    if (Nx > 1) {
        near(m->pad[0], m->pml[0], [](const aefdtd_sensor_req &q) { return q.x1; }, [](const aefdtd_source_req &q) { return q.x1; });
        far(m->pad[1], m->pml[1], Nx, [](const aefdtd_sensor_req &q) { return q.x2; }, [](const aefdtd_source_req &q) { return q.x2; });
    }
    if (Ny > 1) {
        near(m->pad[2], m->pml[2], [](const aefdtd_sensor_req &q) { return q.y1; }, [](const aefdtd_source_req &q) { return q.y1; });
        far(m->pad[3], m->pml[3], Ny, [](const aefdtd_sensor_req &q) { return q.y2; }, [](const aefdtd_source_req &q) { return q.y2; });
    }
    if (Nz > 1) {
        near(m->pad[4], m->pml[4], [](const aefdtd_sensor_req &q) { return q.z1; }, [](const aefdtd_source_req &q) { return q.z1; });
        far(m->pad[5], m->pml[5], Nz, [](const aefdtd_sensor_req &q) { return q.z2; }, [](const aefdtd_source_req &q) { return q.z2; });
    }
    return AEFDTD_OK;
}

// mode_default_fdtd, modes/default_fdtd.cpp:24-210 (fixed step count).
int aeoracle_run_custom(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors, double *sensor_out,
                        double *fields_out, const int *end_computation) {
    if (check_time_type(m)) return AEFDTD_EINVAL;
    int Nt = m->Nt;
    double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * c_light) * 0.99 * m->time_mod;
    GridGuard G;
    int rc = aeoracle_create(&G.g, m->nx, m->ny, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, m->pml, m->pad, AEFDTD_FP64);
    if (rc) return rc;
    aeoracle_grid *g = G.g;
    if ((rc = setup_grid(g, m))) return rc;
    aeoracle_bootstrap(g);
    SensorSet S;
    if ((rc = add_requests(g, n_sensors, sensors, S))) return rc;
    for (int i = 0; i < m->n_sources; i++) {
        const aefdtd_source_req &q = m->sources[i];
        if (q.type != 0) return set_err(g, AEFDTD_EINVAL, "only oscillator sources are supported");
        int id;
        if ((rc = aeoracle_add_oscillator(g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, q.lambda_max, &id)))
            return rc;
    }
    Completion E(m, g, Dt);
    int t;
    bool stopped = false;
    for (t = 0; t < Nt; t++) {
        g->update_E();       // + sources' inject_E
        g->update_H();       // + sources' inject_H (step count)
        aeoracle_feed_sensors(g);
        if (E.on() && E.feed_and_check(g, t)) { t++; break; }     // default_fdtd.cpp:193-195
        if (cancelled(end_computation)) { t++; stopped = true; break; }   // :201
    }
    g_last_steps = t;
    collect_requests(g, n_sensors, sensors, S, sensor_out);
    copy_fields(g, fields_out);
    return stopped ? AEFDTD_ECANCELLED : AEFDTD_OK;
}

int aeoracle_run_single_particle(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                                 double *sensor_out, double *fields_out, const int *end_computation) {
    if (check_time_type(m)) return AEFDTD_EINVAL;
    int Nt = m->Nt;
    double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    int pml_x = std::max(m->pml[0], m->pml[1]), pml_y = std::max(m->pml[2], m->pml[3]);
    bool periodic_x = pml_x <= 0, periodic_y = pml_y <= 0;
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * c_light) * 0.99 * m->time_mod;
    int pad[6];
    for (int f = 0; f < 6; f++) pad[f] = std::max(10, m->pad[f]);
    if (pml_x <= 0) pad[0] = pad[1] = 0;
    if (pml_y <= 0) pad[2] = pad[3] = 0;
    GridGuard G, A;
    int rc = aeoracle_create(&G.g, m->nx, m->ny, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, m->pml, pad, AEFDTD_FP64);
    if (rc) return rc;
    const int apml[6] = {0, 0, 0, 0, m->pml[4], m->pml[5]};
    const int apad[6] = {0, 0, 0, 0, pad[4], pad[5]};
    if ((rc = aeoracle_create(&A.g, 1, 1, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, apml, apad, AEFDTD_FP64))) return rc;
    aeoracle_grid *g = G.g, *a = A.g;
    for (int f = 0; f < 6; f++) aeoracle_set_pml(g, f, m->kappa[f], m->sigma[f], m->alpha[f]);
    aeoracle_set_pml(a, 4, m->kappa[4], m->sigma[4], m->alpha[4]);
    aeoracle_set_pml(a, 5, m->kappa[5], m->sigma[5], m->alpha[5]);
    // the driver never calls set_tapering: Ntap stays 0 (single_particle.cpp:101-165)
    if ((rc = aeoracle_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz))) return rc;
    std::vector<uint32_t> col(m->nz);
    for (int k = 0; k < m->nz; k++) col[k] = m->matsgrid[(size_t)m->nx * m->ny * k];
    if ((rc = aeoracle_set_matsgrid(a, col.data(), 1, 1, m->nz))) return rc;
    for (int i = 0; i < m->n_materials; i++) {
        aeoracle_set_material(g, i, &m->materials[i]);
        aeoracle_set_material(a, i, &m->materials[i]);
    }
    aeoracle_disable_fields(g, m->disable_fields, m->n_disable);
    int zs_e = g->zs_e;
    double pol = 0;
    if (m->polarization == 1) pol = 90;
    else if (m->polarization == 2) pol = m->polar_angle;
    double eps_sup = g->mats[g->mg(0, 0, zs_e)].ei;
    aefdtd_pulse chp;
    pulse_init(chp, m->lambda_min, m->lambda_max, 0, pol, 0, 0, 0, (zs_e - 1 + 5) * Dz, 1.0);
    aeoracle_bootstrap(g);
    std::vector<double> dEx(Nt), dEy(Nt), dHx(Nt), dHy(Nt);
    for (int t = 0; t < Nt; t++) {
        double tb = t * Dt;
        double iHx = pulse_eval(chp, 3, 0, 0, (zs_e - 1 + 5.5) * Dz, tb);
        double iHy = pulse_eval(chp, 4, 0, 0, (zs_e - 1 + 5.5) * Dz, tb);
        dEx[t] = -(g->dtdez * iHy / eps_sup);
        dEy[t] = g->dtdez * iHx / eps_sup;
        tb = (t + 0.5) * Dt;
        double iEx = pulse_eval(chp, 0, 0, 0, (zs_e - 1 + 5) * Dz, tb);
        double iEy = pulse_eval(chp, 1, 0, 0, (zs_e - 1 + 5) * Dz, tb);
        dHx[t] = g->dtdmz * iEy;
        dHy[t] = -(g->dtdmz * iEx);
    }
    aeoracle_set_plane_injection(a, zs_e - 1 + 5, Nt, dEx.data(), dEy.data(), dHx.data(), dHy.data());
    if ((rc = aeoracle_set_tfsf(g, a, g->xs_s, g->xs_e, g->ys_s, g->ys_e, g->zs_s, g->zs_e, periodic_x, periodic_y))) return rc;
    SensorSet S;
    if ((rc = add_requests(g, n_sensors, sensors, S))) return rc;
    Completion E(m, g, Dt);
    int t;
    bool stopped = false;
    for (t = 0; t < Nt; t++) {
        a->update_E();
        g->update_E();
        a->update_H();
        g->update_H();
        aeoracle_feed_sensors(g);
        if (E.on() && E.feed_and_check(g, t)) { t++; break; }     // single_particle.cpp:399-401
        if (cancelled(end_computation)) { t++; stopped = true; break; }   // :403
    }
    g_last_steps = t;
    collect_requests(g, n_sensors, sensors, S, sensor_out);
    copy_fields(g, fields_out);
    return stopped ? AEFDTD_ECANCELLED : AEFDTD_OK;
}

}  // extern "C"
