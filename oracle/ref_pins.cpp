// ref_pins.cpp -- TEST INFRASTRUCTURE.  C entry points over the parts of
// Aether that compile from their own sources in this image (no stand-in
// headers): the dispersive-model layer (src/materials/dielec_models.cpp) and
// the math helpers (src/common/mathUT.cpp).  Built by oracle/Makefile into
// oracle/_ref/ from the sources where they lie; used only by tests to pin the
// oracle's pole constants, permittivity evaluation, tapering curve and index
// rounding against the reference's own compiled code.
#include <dielec_models.h>
#include <mathUT.h>

extern "C" {

// kind: 0 Debye(ds,t0) 1 Drude(wd,g) 2 Lorentz(A,O,G) 3 CritPoint(A,O,P,G).
// out: a.re, a.im, b.re, b.im, sigma (Drude only, else 0)
int aeref_time_exp(int kind, const double *p, double *out) {
    Imdouble a, b;
    double sig = 0;
    if (kind == 0) { DebyeModel m; m.set(p[0], p[1]); m.get_time_exp(a, b); }
    else if (kind == 1) { DrudeModel m; m.set(p[0], p[1]); m.get_time_exp(a, b); sig = m.get_sigma(); }
    else if (kind == 2) { LorentzModel m; m.set(p[0], p[1], p[2]); m.get_time_exp(a, b); }
    else if (kind == 3) { CritpointModel m; m.set(p[0], p[1], p[2], p[3]); m.get_time_exp(a, b); }
    else return 1;
    out[0] = a.real(); out[1] = a.imag(); out[2] = b.real(); out[3] = b.imag(); out[4] = sig;
    return 0;
}

// Dielec_model::eval(w) for one model of `kind` on top of eps_inf.
int aeref_eval(int kind, const double *p, double eps_inf, double w, double *out) {
    Dielec_model d;
    d.eps_inf = eps_inf;
    if (kind == 0) { DebyeModel m; m.set(p[0], p[1]); d.debye_arr.push_back(m); }
    else if (kind == 1) { DrudeModel m; m.set(p[0], p[1]); d.drude_arr.push_back(m); }
    else if (kind == 2) { LorentzModel m; m.set(p[0], p[1], p[2]); d.lorentz_arr.push_back(m); }
    else if (kind == 3) { CritpointModel m; m.set(p[0], p[1], p[2], p[3]); d.cp_arr.push_back(m); }
    else if (kind != -1) return 1;
    Imdouble r = d.eval(w);
    out[0] = r.real(); out[1] = r.imag();
    return 0;
}

double aeref_s_curve(double x, double xmin, double xmax) { return s_curve(x, xmin, xmax); }
// The global generator of mathUT.cpp: seedp (rng_seed) and randp(A, B).
void aeref_seedp(int s) { seedp(s); }
double aeref_randp(double A, double B) { return randp(A, B); }
int aeref_nearest_integer(double x) { return nearest_integer(x); }

}
