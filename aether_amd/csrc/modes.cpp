// modes.cpp -- host mode drivers over the grid C ABI.  They replace Aether's
// FDTD_normal_incidence (modes/normal_incidence.cpp:161-607),
// FDTD_single_particle (modes/single_particle.cpp:29-415) and the custom mode
// mode_default_fdtd (modes/default_fdtd.cpp:24-210): same setup, same
// per-step sequence, same post-processing, with the fields resident on the
// device and every per-step field access done by device passes.
//
// A driver is split into a run object (create / step n / finish) so that a
// caller can interleave many runs (the serving loop) or time the stepping
// alone (the benchmark); aefdtd_run_normal_incidence/_single_particle are the
// one-call forms.
#include "runs.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

using namespace aefdtd;

namespace {

void lambda_grid(int Nl, double lmin, double lmax, std::vector<double> &L) {   // normal_incidence.cpp:235
    L.resize(Nl);
    for (int l = 0; l < Nl; l++) L[l] = lmin + (lmax - lmin) * l / (Nl - 1.0);
}

}  // namespace

namespace {

#define TRY(x)                   \
    do {                         \
        int rc_ = (x);           \
        if (rc_) return rc_;     \
    } while (0)

int rfail(aefdtd_run *r, int rc);

// A driver in one call: create, step to the end (or the caller's cancel
// flag), post-process.  A cancelled run is post-processed with the steps it
// made and returns AEFDTD_ECANCELLED.
int run_one_call(int kind, const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                 aefdtd_normal_result *res, double *sensor_out, double *fields_out, const int *end_computation) {
    aefdtd_run *r = nullptr;
    TRY(aefdtd_run_create(&r, kind, m, n_sensors, sensors));
    r->end_flag = end_computation;
    int rc = AEFDTD_OK;
    while (r->t < r->Nt && !r->done && rc == AEFDTD_OK) rc = aefdtd_run_step(r, 1);
    if (rc == AEFDTD_ECANCELLED) rc = AEFDTD_OK;
    if (rc == AEFDTD_OK) rc = aefdtd_run_finish(r, res, sensor_out, fields_out);
    const bool cancelled = r->cancelled;
    aefdtd_run_destroy(r);
    if (rc) return rc;
    return cancelled ? AEFDTD_ECANCELLED : AEFDTD_OK;
}

int rfail(aefdtd_run *r, int rc) {
    if (r && r->g && !r->g->err.empty()) r->err = r->g->err;
    else if (r && r->a && !r->a->err.empty()) r->err = r->a->err;
    else if (r) r->err = aefdtd_last_error(nullptr);
    return rc;
}

// Box_Spect_Poynting (sensors/s_spec_power.cpp:125-232) becomes six planar
// sensors; a planar request one sensor.
int add_requests(aefdtd_run *r) {
    for (const aefdtd_sensor_req &q : r->reqs) {
        if (q.Nl < 1) return AEFDTD_EINVAL;
        std::vector<double> L;
        lambda_grid(q.Nl, q.lambda_min, q.lambda_max, L);
        std::vector<int> v;
        if (q.box) {
            const int t[6] = {AEFDTD_NORMAL_XM, AEFDTD_NORMAL_X, AEFDTD_NORMAL_YM, AEFDTD_NORMAL_Y, AEFDTD_NORMAL_ZM, AEFDTD_NORMAL_Z};
            const int b[6][6] = {{q.x1, q.x1 + 1, q.y1, q.y2, q.z1, q.z2}, {q.x2 - 1, q.x2, q.y1, q.y2, q.z1, q.z2},
                                 {q.x1, q.x2, q.y1, q.y1 + 1, q.z1, q.z2}, {q.x1, q.x2, q.y2 - 1, q.y2, q.z1, q.z2},
                                 {q.x1, q.x2, q.y1, q.y2, q.z1, q.z1 + 1}, {q.x1, q.x2, q.y1, q.y2, q.z2 - 1, q.z2}};
            for (int f = 0; f < 6; f++) {
                int id = -1;
                if (!q.disable[f])
                    TRY(aefdtd_add_spect_poynting(r->g, t[f], b[f][0], b[f][1], b[f][2], b[f][3], b[f][4], b[f][5], q.Nl, L.data(), &id));
                v.push_back(id);
            }
        } else {
            int id = -1;
            TRY(aefdtd_add_spect_poynting(r->g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.Nl, L.data(), &id));
            v.push_back(id);
        }
        r->ids.push_back(v);
    }
    return AEFDTD_OK;
}

// Spect_Poynting::treat per plane; Box_Spect_Poynting::treat sums the faces
// in the order xm, xp, ym, yp, zm, zp (s_spec_power.cpp:199-207).
int collect_requests(aefdtd_run *r, double *out) {
    size_t off = 0;
    for (size_t q = 0; q < r->reqs.size(); q++) {
        int Nl = r->reqs[q].Nl;
        std::vector<double> tmp(Nl);
        if (r->reqs[q].box) {
            std::vector<double> tot(Nl, 0.0), faces(6 * (size_t)Nl, 0.0);
            for (int f = 0; f < 6; f++) {
                if (r->ids[q][f] < 0) continue;
                TRY(aefdtd_sensor_poynting(r->g, r->ids[q][f], tmp.data()));
                std::memcpy(&faces[f * (size_t)Nl], tmp.data(), Nl * sizeof(double));
            }
            for (int l = 0; l < Nl; l++)
                for (int f = 0; f < 6; f++)
                    if (r->ids[q][f] >= 0) tot[l] += faces[f * (size_t)Nl + l];
            if (out) {
                std::memcpy(out + off, tot.data(), Nl * sizeof(double));
                std::memcpy(out + off + Nl, faces.data(), 6 * (size_t)Nl * sizeof(double));
            }
            off += 7 * (size_t)Nl;
        } else {
            TRY(aefdtd_sensor_poynting(r->g, r->ids[q][0], tmp.data()));
            if (out) std::memcpy(out + off, tmp.data(), Nl * sizeof(double));
            off += Nl;
        }
    }
    return AEFDTD_OK;
}

int copy_fields(aefdtd_grid *g, double *out) {
    if (!out) return AEFDTD_OK;
    size_t N = g->ncell();
    for (int c = 0; c < 6; c++) TRY(aefdtd_read_field(g, c, out + c * N));
    return AEFDTD_OK;
}

uint32_t mg_at(const aefdtd_grid *g, int i, int j, int k) {
    return g->matsgrid[(size_t)i + (size_t)g->Nx * ((size_t)j + (size_t)g->Ny * k)];
}

// Setup of FDTD_normal_incidence up to the time loop (normal_incidence.cpp:161-324).
int setup_normal(aefdtd_run *r, const aefdtd_mode *m) {
    const int Nt = m->Nt;
    const double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    int pad_zm = std::max(5, m->pad[4]), pad_zp = std::max(5, m->pad[5]);
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * kC) * 0.99 * m->time_mod;
    r->Dt = Dt;
    const int pml[6] = {0, 0, 0, 0, m->pml[4], m->pml[5]};
    const int pad[6] = {0, 0, 0, 0, pad_zm, pad_zp};
    TRY(aefdtd_create(&r->g, m->nx, m->ny, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_NORM, pml, pad, m->precision));
    aefdtd_grid *g = r->g;
    TRY(aefdtd_set_pml(g, 4, m->kappa[4], m->sigma[4], m->alpha[4]));
    TRY(aefdtd_set_pml(g, 5, m->kappa[5], m->sigma[5], m->alpha[5]));
    TRY(aefdtd_set_tapering(g, m->tapering));
    TRY(aefdtd_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz));
    for (int i = 0; i < m->n_materials; i++) TRY(aefdtd_set_material(g, i, &m->materials[i]));
    TRY(aefdtd_disable_fields(g, m->disable_fields, m->n_disable));
    const int zs_s = g->zs_s, zs_e = g->zs_e;
    double pol = (m->polarization == 1) ? 90 : 0;
    double eps_sub = g->mats[mg_at(g, 0, 0, zs_s)].ei, eps_sup = g->mats[mg_at(g, 0, 0, zs_e)].ei;
    r->index_sub = std::sqrt(eps_sub);
    r->index_sup = std::sqrt(eps_sup);
    aefdtd_pulse chp;
    pulse_init(chp, m->lambda_min, m->lambda_max, 0, pol, 0, 0, 0, (zs_e - 1) * Dz, r->index_sup);
    TRY(aefdtd_bootstrap(g));
    // per-step injection values (normal_incidence.cpp:334-357) and tapering (:384-385)
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
    TRY(aefdtd_set_plane_injection(g, zs_e - 1, Nt, dEx.data(), dEy.data(), dHx.data(), dHy.data()));
    TRY(add_requests(r));
    TRY(aefdtd_add_plane_sum(g, zs_e - 2, &chp, (zs_e - 2) * Dz, 1, tap.data(), &r->idR));
    TRY(aefdtd_add_plane_sum(g, zs_s, &chp, zs_s * Dz, 0, tap.data(), &r->idT));
    return AEFDTD_OK;
}

// Post-processing of FDTD_normal_incidence (normal_incidence.cpp:442-570).
int finish_normal(aefdtd_run *r, aefdtd_normal_result *res) {
    aefdtd_grid *g = r->g;
    const int Nt = r->Nt, Nrun = r->nrun >= 0 ? r->nrun : r->t, Nl = r->Nl;
    const double Dz = g->Dz, Dt = r->Dt;
    const int zs_s = g->zs_s, zs_e = g->zs_e;
    std::vector<double> Lambda, w(Nl);
    lambda_grid(Nl, r->lambda_min, r->lambda_max, Lambda);
    for (int l = 0; l < Nl; l++) w[l] = 2.0 * kPi * kC / Lambda[l];
    std::vector<double> R(6 * (size_t)Nt), T(6 * (size_t)Nt);
    TRY(aefdtd_read_plane_sum(g, r->idR, Nt, R.data()));
    TRY(aefdtd_read_plane_sum(g, r->idT, Nt, T.data()));
    // find_slab (core/fdtd_utils.cpp:347-392)
    int Nsub = 0, Nsup = 0;
    long IDsup = 0, IDsub = 0, IDmoy;
    for (int i = g->xs_s; i < g->xs_e; i++)
        for (int j = g->ys_s; j < g->ys_e; j++) { IDsup += mg_at(g, i, j, zs_e); IDsub += mg_at(g, i, j, zs_s); }
    IDmoy = IDsub;
    int k = zs_s;
    while (IDmoy == IDsub && k < zs_e - 1) {
        IDmoy = 0;
        for (int i = g->xs_s; i < g->xs_e; i++)
            for (int j = g->ys_s; j < g->ys_e; j++) IDmoy += mg_at(g, i, j, k + 1);
        Nsub++; k++;
    }
    IDmoy = IDsup;
    k = zs_e - 1;
    while (IDmoy == IDsup && k >= zs_s) {
        IDmoy = 0;
        for (int i = g->xs_s; i < g->xs_e; i++)
            for (int j = g->ys_s; j < g->ys_e; j++) IDmoy += mg_at(g, i, j, k - 1);
        Nsup++; k--;
    }
    double hsup = (Nsup - 1) * Dz, hsub = Nsub * Dz, hstruc = (g->Nz_s - Nsup - Nsub) * Dz;
    hsup -= Dz / 2.0;   // "phase calibration" of normal_incidence.cpp:448-449
    hsub -= Dz / 2.0;
    // post-run DFT of the plane sums (normal_incidence.cpp:455-488)
    const std::complex<double> I(0.0, 1.0);
    for (int l = 0; l < Nl; l++) {
        double k0 = -w[l] / kC, k1 = k0 * r->index_sup, k3 = k0 * r->index_sub;
        std::complex<double> tmp_sup = std::exp(2.0 * k1 * hsup * I);
        std::complex<double> tmp_sub = std::exp((k3 * hsub - k0 * (hstruc + hsub)) * I);
        std::complex<double> r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0}, t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
        for (int tt = 0; tt < Nrun; tt++) {
            std::complex<double> tmp1 = std::exp(tt * Dt * w[l] * I) / (Nrun - 1.0);
            std::complex<double> tmp2 = tmp1 * tmp_sup, tmp3 = tmp1 * tmp_sub;
            for (int c = 0; c < 3; c++) {
                r1[c] += R[6 * (size_t)tt + 3 + c] * tmp1;
                r2[c] += R[6 * (size_t)tt + c] * tmp2;
                t1[c] += T[6 * (size_t)tt + 3 + c] * tmp1;
                t2[c] += T[6 * (size_t)tt + c] * tmp3;
            }
        }
        double a = std::sqrt(std::norm(r2[0]) + std::norm(r2[1]) + std::norm(r2[2]));
        double b = std::sqrt(std::norm(r1[0]) + std::norm(r1[1]) + std::norm(r1[2]));
        if (res->lambda) res->lambda[l] = Lambda[l];
        if (res->R) res->R[l] = a / b;
        a = std::sqrt(std::norm(t2[0]) + std::norm(t2[1]) + std::norm(t2[2]));
        b = std::sqrt(std::norm(t1[0]) + std::norm(t1[1]) + std::norm(t1[2]));
        if (res->T) res->T[l] = a / b;
        std::complex<double> ref0 = r1[1], trans0 = t1[1];
        if (r->polarization == 1) { ref0 = r1[0]; trans0 = t1[0]; }
        std::complex<double> v[4] = {r2[0] / ref0, r2[1] / ref0, t2[0] / trans0, t2[1] / trans0};
        double *dst[4] = {res->refx, res->refy, res->transx, res->transy};
        for (int q = 0; q < 4; q++)
            if (dst[q]) { dst[q][2 * l] = v[q].real(); dst[q][2 * l + 1] = v[q].imag(); }
    }
    res->index_sup = r->index_sup;
    res->index_sub = r->index_sub;
    res->Nt_run = Nrun;
    return AEFDTD_OK;
}

// Setup of FDTD_single_particle up to the time loop (single_particle.cpp:29-255).
int setup_single(aefdtd_run *r, const aefdtd_mode *m) {
    const int Nt = m->Nt;
    const double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    int pml_x = std::max(m->pml[0], m->pml[1]), pml_y = std::max(m->pml[2], m->pml[3]);
    bool periodic_x = pml_x <= 0, periodic_y = pml_y <= 0;
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * kC) * 0.99 * m->time_mod;
    r->Dt = Dt;
    int pad[6];
    for (int f = 0; f < 6; f++) pad[f] = std::max(10, m->pad[f]);
    if (pml_x <= 0) pad[0] = pad[1] = 0;
    if (pml_y <= 0) pad[2] = pad[3] = 0;
    TRY(aefdtd_create(&r->g, m->nx, m->ny, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, m->pml, pad, m->precision));
    const int apml[6] = {0, 0, 0, 0, m->pml[4], m->pml[5]};
    const int apad[6] = {0, 0, 0, 0, pad[4], pad[5]};
    TRY(aefdtd_create(&r->a, 1, 1, m->nz, Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, apml, apad, m->precision));
    aefdtd_grid *g = r->g, *a = r->a;
    for (int f = 0; f < 6; f++) TRY(aefdtd_set_pml(g, f, m->kappa[f], m->sigma[f], m->alpha[f]));
    TRY(aefdtd_set_pml(a, 4, m->kappa[4], m->sigma[4], m->alpha[4]));
    TRY(aefdtd_set_pml(a, 5, m->kappa[5], m->sigma[5], m->alpha[5]));
    TRY(aefdtd_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz));
    std::vector<uint32_t> col(m->nz);
    for (int k = 0; k < m->nz; k++) col[k] = m->matsgrid[(size_t)m->nx * m->ny * k];   // aux_grid(0,0,k)=matsgrid(0,0,k)
    TRY(aefdtd_set_matsgrid(a, col.data(), 1, 1, m->nz));
    for (int i = 0; i < m->n_materials; i++) {
        TRY(aefdtd_set_material(g, i, &m->materials[i]));
        TRY(aefdtd_set_material(a, i, &m->materials[i]));
    }
    TRY(aefdtd_disable_fields(g, m->disable_fields, m->n_disable));
    const int zs_e = g->zs_e;
    double pol = 0;
    if (m->polarization == 1) pol = 90;
    else if (m->polarization == 2) pol = m->polar_angle;
    double eps_sup = g->mats[mg_at(g, 0, 0, zs_e)].ei;
    aefdtd_pulse chp;
    pulse_init(chp, m->lambda_min, m->lambda_max, 0, pol, 0, 0, 0, (zs_e - 1 + 5) * Dz, 1.0);
    TRY(aefdtd_bootstrap(g));
    TRY(aefdtd_bootstrap(a));
    // auxiliary-grid injection values (single_particle.cpp:270-275,343-348)
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
    TRY(aefdtd_set_plane_injection(a, zs_e - 1 + 5, Nt, dEx.data(), dEy.data(), dHx.data(), dHy.data()));
    TRY(aefdtd_set_tfsf(g, a, g->xs_s, g->xs_e, g->ys_s, g->ys_e, g->zs_s, g->zs_e, periodic_x, periodic_y));
    TRY(add_requests(r));
    return AEFDTD_OK;
}

// Setup of mode_default_fdtd up to the time loop (default_fdtd.cpp:24-98):
// the grid as given by the finalized mode, then sensors and sources.
int setup_custom(aefdtd_run *r, const aefdtd_mode *m) {
    const double Dx = m->Dx, Dy = m->Dy, Dz = m->Dz;
    double Dt = std::min(std::min(Dx, Dy), Dz) / (std::sqrt(3.0) * kC) * 0.99 * m->time_mod;
    r->Dt = Dt;
    TRY(aefdtd_create(&r->g, m->nx, m->ny, m->nz, m->Nt, Dx, Dy, Dz, Dt, AEFDTD_MODE_CUSTOM, m->pml, m->pad,
                      m->precision));
    aefdtd_grid *g = r->g;
    if (m->precision == AEFDTD_FP16 && m->n_sources > 0) {
        // oscillators inject amplitude 50 (sc_oscillator.cpp:37-38) and the fields near them
        // grow past the unit-wave scale: store E unscaled and H at 2^8 (about 377 x E)
        const int ex[6] = {0, 0, 0, 8, 8, 8};
        TRY(aefdtd_set_storage_scale(g, ex));
    }
    for (int f = 0; f < 6; f++) TRY(aefdtd_set_pml(g, f, m->kappa[f], m->sigma[f], m->alpha[f]));
    TRY(aefdtd_set_tapering(g, m->tapering));
    TRY(aefdtd_set_matsgrid(g, m->matsgrid, m->nx, m->ny, m->nz));
    for (int i = 0; i < m->n_materials; i++) TRY(aefdtd_set_material(g, i, &m->materials[i]));
    TRY(aefdtd_disable_fields(g, m->disable_fields, m->n_disable));
    TRY(aefdtd_bootstrap(g));
    TRY(add_requests(r));
    for (const aefdtd_source_req &q : r->srcs) {
        if (q.type != 0) {
            g->err = "only oscillator sources are supported";
            return AEFDTD_EINVAL;
        }
        int id;
        TRY(aefdtd_add_oscillator(g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, q.lambda_max,
                                  &id));
    }
    return AEFDTD_OK;
}

}  // namespace

// Grid sizes of the drivers' setups (setup_normal/_single/_custom) without
// building them, for admission control.
void aefdtd::run_dims(int kind, const aefdtd_mode *m, int dims[3], int pads[6]) {
    int pml[6], pad[6];
    for (int f = 0; f < 6; f++) { pml[f] = m->pml[f]; pad[f] = m->pad[f]; }
    if (kind == 0) {
        for (int f = 0; f < 4; f++) pml[f] = pad[f] = 0;
        pad[4] = std::max(5, m->pad[4]);
        pad[5] = std::max(5, m->pad[5]);
    } else if (kind == 1) {
        for (int f = 0; f < 6; f++) pad[f] = std::max(10, m->pad[f]);
        if (std::max(pml[0], pml[1]) <= 0) pad[0] = pad[1] = 0;
        if (std::max(pml[2], pml[3]) <= 0) pad[2] = pad[3] = 0;
    }
    for (int a = 0; a < 3; a++)                 // FDTD::FDTD (core/fdtd_core.cpp:212-243)
        if (pml[2 * a] || pml[2 * a + 1]) {
            pad[2 * a] = std::max(5, pad[2 * a]);
            pad[2 * a + 1] = std::max(5, pad[2 * a + 1]);
        }
    const int n[3] = {m->nx, m->ny, m->nz};
    for (int a = 0; a < 3; a++) dims[a] = n[a] + pml[2 * a] + pml[2 * a + 1] + pad[2 * a] + pad[2 * a + 1];
    if (pads)
        for (int f = 0; f < 6; f++) pads[f] = pad[f];
}

extern "C" {

int aefdtd_mode_finalize(aefdtd_mode *m, int kind, int n_sensors, const aefdtd_sensor_req *sensors) {
    if (!m || kind < 0 || kind > 2 || (n_sensors > 0 && !sensors) || (m->n_sources > 0 && !m->sources))
        return AEFDTD_EINVAL;
    if (kind == 0) {                       // lua_fdtd.cpp:90-97
        m->pml[0] = m->pml[1] = m->pml[2] = m->pml[3] = 0;
        if (m->pml[4] <= 0) m->pml[4] = 25;
        if (m->pml[5] <= 0) m->pml[5] = 25;
    }
    const int N[3] = {m->nx, m->ny, m->nz};
    for (int a = 0; a < 3; a++) {          // lua_fdtd.cpp:99-202
        if (N[a] <= 1) continue;
        for (int side = 0; side < 2; side++) {
            const int f = 2 * a + side;
            if (!m->pml[f]) {
                m->pad[f] = 0;
                continue;
            }
            int p = std::max(m->pad[f], 5);
            for (int i = 0; i < n_sensors; i++) {
                const int lo[3] = {sensors[i].x1, sensors[i].y1, sensors[i].z1};
                const int hi[3] = {sensors[i].x2, sensors[i].y2, sensors[i].z2};
                p = std::max(p, side == 0 ? 5 - lo[a] : 5 + hi[a] - N[a]);
            }
            for (int i = 0; i < m->n_sources; i++) {
                const aefdtd_source_req &q = m->sources[i];
                const int lo[3] = {q.x1, q.y1, q.z1}, hi[3] = {q.x2, q.y2, q.z2};
                const int margin = q.type == 0 ? 50 : 5;
                p = std::max(p, side == 0 ? margin - lo[a] : margin + hi[a] - N[a]);
            }
            m->pad[f] = p;
        }
    }
    return AEFDTD_OK;
}

int aefdtd_synchronize(aefdtd_grid *g) {
    if (!g) return AEFDTD_EINVAL;
    if (hipStreamSynchronize(g->stream) != hipSuccess) {
        g->err = "hipStreamSynchronize failed";
        return AEFDTD_EDEVICE;
    }
    return AEFDTD_OK;
}

int aefdtd_run_create(aefdtd_run **out, int kind, const aefdtd_mode *m, int n_sensors,
                      const aefdtd_sensor_req *sensors) {
    if (!out) return AEFDTD_EINVAL;
    *out = nullptr;
    if (!m || !m->matsgrid || kind < 0 || kind > 2 || m->Nt < 1 || (n_sensors > 0 && !sensors) ||
        (m->n_sources > 0 && !m->sources))
        return AEFDTD_EINVAL;
    if (kind == 0 && (m->Nt < 2 || m->Nl < 2)) return AEFDTD_EINVAL;
    auto *r = new aefdtd_run();
    r->kind = kind;
    r->Nt = m->Nt;
    r->polarization = m->polarization;
    r->Nl = m->Nl;
    r->lambda_min = m->lambda_min;
    r->lambda_max = m->lambda_max;
    r->reqs.assign(sensors, sensors + std::max(0, n_sensors));
    if (kind == 2 && m->n_sources > 0) r->srcs.assign(m->sources, m->sources + m->n_sources);
    if (m->time_type < 0 || m->time_type > 2 || (m->time_type != 0 && m->cc_step < 1) ||
        (m->time_type == 2 && (m->cc_quant < 1 || !(m->cc_lmin > 0) || !(m->cc_lmax > 0)))) {
        delete r;
        return AEFDTD_EINVAL;
    }
    r->time_type = m->time_type;
    r->cc_step = m->cc_step;
    r->cc_coeff = m->cc_coeff;
    int rc = kind == 0 ? setup_normal(r, m) : (kind == 1 ? setup_single(r, m) : setup_custom(r, m));
    if (rc) {
        // keep the failure text reachable through aefdtd_last_error(NULL)
        std::string e = r->g && !r->g->err.empty() ? r->g->err : (r->a && !r->a->err.empty() ? r->a->err : "");
        if (!e.empty()) set_global_error(e);
        delete r;
        return rc;
    }
    r->g->en_fuse = r->time_type == 1;      // the E pass then also sums the energy (3D, FP32/FP64)
    if (r->time_type == 2) {
        // the sensor is linked to the main grid after its setup (single_particle.cpp:216-229)
        std::vector<int> pts;
        std::vector<double> w;
        int navg = 0;
        rc = aefdtd::ft_place(r->g, r->Dt, m->cc_lmin, m->cc_lmax, m->cc_quant, m->cc_layout, m->cc_seed, pts, w, navg);
        if (!rc) rc = aefdtd::ft_setup(r->g, r->ft, pts, w, navg);
        if (rc) {
            set_global_error(r->g->err.empty() ? "bad spectral completion parameters" : r->g->err);
            delete r;
            return rc;
        }
    }
    *out = r;
    return AEFDTD_OK;
}

int aefdtd_run_step(aefdtd_run *r, int n) {
    if (!r) return AEFDTD_EINVAL;
    if (r->batch) {
        r->err = "run is stepped by a batch";
        return AEFDTD_ESTATE;
    }
    for (int s = 0; s < n && r->t < r->Nt && !r->done; s++) {
        int rc;
        if (r->kind == 1) {
            // single_particle.cpp:256-381: aux E, main E (+TFSF-E), aux H, main H (+TFSF-H), sensors;
            // one step of the grid and its column (the one-pass kernel on 3D grids)
            if ((rc = aefdtd::step_pair(r->g, r->a)) || (rc = aefdtd_feed_sensors(r->g))) return rfail(r, rc);
        } else {
            // normal_incidence.cpp:325-415: E (+injection), H (+injection), sensors + R/T sums;
            // default_fdtd.cpp:148-165: E (+sources' inject_E), H, sensors (inject_H of an
            // oscillator only counts its step, which update_H does)
            if ((rc = aefdtd::step_pair(r->g, nullptr)) || (rc = aefdtd_feed_sensors(r->g))) return rfail(r, rc);
        }
        if (r->time_type != 0) {
            // the CompletionSensor is fed with the other sensors; the drivers check it
            // when t % cc_step == 0 (single_particle.cpp:396-401, normal_incidence.cpp:417-423,
            // default_fdtd.cpp:193-195): complete once its step count reaches twice the
            // longest grid axis and, for the energy criterion, the last energy is below
            // cc_coeff of the largest (s_completion.cpp:39-41,122-128), for the spectral
            // one, 90% of the watched cells have settled (:43-120)
            if (r->time_type == 1) rc = aefdtd::energy_feed(r->g);
            else rc = aefdtd::ft_feed(r->g, r->ft, r->t, r->Dt);
            if (rc) return rfail(r, rc);
            if (r->t % r->cc_step == 0) {
                const int steps = r->t + 1, nmax = std::max(r->g->Nx, std::max(r->g->Ny, r->g->Nz));
                bool complete = false;
                if (r->time_type == 1) {
                    double last = 0, mx = 0;
                    if ((rc = aefdtd::energy_read(r->g, &last, &mx))) return rfail(r, rc);
                    complete = steps >= 2 * nmax && last / mx < r->cc_coeff;
                } else if (steps >= 2 * nmax) {
                    if ((rc = aefdtd::ft_read(r->g, r->ft, r->ft_ring))) return rfail(r, rc);
                    complete = aefdtd::ft_complete(r->ft_ring, r->ft.np, r->ft.navg, r->cc_coeff, steps, nmax);
                }
                if (complete) {
                    r->done = true;
                    if (r->kind == 0) r->nrun = r->t;      // normal incidence: Nt = t, then break
                    r->t++;
                    break;
                }
            }
        }
        r->t++;
        if (r->cancel.load(std::memory_order_relaxed) ||
            (r->end_flag && __atomic_load_n(r->end_flag, __ATOMIC_RELAXED))) {
            // the run ends after this step; normal incidence keeps Nt = t steps for its spectra
            r->done = true;
            r->cancelled = true;
            if (r->kind == 0) r->nrun = r->t - 1;
            return AEFDTD_ECANCELLED;
        }
    }
    return r->cancelled ? AEFDTD_ECANCELLED : AEFDTD_OK;
}

int aefdtd_run_cancel(aefdtd_run *r) {
    if (!r) return AEFDTD_EINVAL;
    r->cancel.store(true, std::memory_order_relaxed);
    return AEFDTD_OK;
}

int aefdtd_run_progress(const aefdtd_run *r, int *t, int *Nt) {
    if (!r) return AEFDTD_EINVAL;
    if (t) *t = r->t;
    if (Nt) *Nt = r->done ? r->t : r->Nt;     // a run stopped by its completion check has ended
    return AEFDTD_OK;
}

aefdtd_grid *aefdtd_run_grid(aefdtd_run *r) { return r ? r->g : nullptr; }

const char *aefdtd_run_error(const aefdtd_run *r) { return r ? r->err.c_str() : ""; }

int aefdtd_run_finish(aefdtd_run *r, aefdtd_normal_result *res, double *sensor_out, double *fields_out) {
    if (!r) return AEFDTD_EINVAL;
    if (r->batch) {
        r->err = "run is in a batch: remove it first";
        return AEFDTD_ESTATE;
    }
    int rc;
    if (r->kind == 0 && res && (rc = finish_normal(r, res))) return rfail(r, rc);
    if ((rc = collect_requests(r, sensor_out))) return rfail(r, rc);
    if ((rc = copy_fields(r->g, fields_out))) return rfail(r, rc);
    return AEFDTD_OK;
}

int aefdtd_run_destroy(aefdtd_run *r) {
    if (r && r->batch) aefdtd_batch_remove(r->batch, r);
    delete r;
    return AEFDTD_OK;
}

// Result files as the reference writes them; the same stream formatting
// (std::ostream, setprecision(15) or the default 6 digits) on the same values.
int aefdtd_run_write(aefdtd_run *r, const char *directory, const char *prefix, const char *const *names,
                     const double *sensor_out, const aefdtd_normal_result *res) {
    if (!r || !directory || !r->g) return AEFDTD_EINVAL;
    namespace fs = std::filesystem;
    const fs::path dir(directory);
    auto open = [&](const fs::path &p, std::ofstream &f) {
        f.open(p, std::ios::out | std::ios::trunc);
        if (!f) r->err = "cannot write " + p.string();
        return (bool)f;
    };
    const aefdtd_grid *g = r->g;
    size_t off = 0;
    for (size_t q = 0; q < r->reqs.size() && sensor_out; q++) {
        const aefdtd_sensor_req &s = r->reqs[q];
        const int Nl = s.Nl;
        std::vector<double> L;
        lambda_grid(Nl, s.lambda_min, s.lambda_max, L);
        const std::string name = names && names[q] ? names[q] : "sensor" + std::to_string(q);
        const double *v = sensor_out + off;
        if (s.box) {
            // Box_Spect_Poynting::treat (sensors/s_spec_power.cpp:174-232)
            std::ofstream f, fm;
            if (!open(dir / (name + "_bspdft"), f) || !open(dir / (name + ".m"), fm)) return AEFDTD_EINVAL;
            for (int l = 0; l < Nl; l++) {
                f << std::setprecision(15) << L[l] << " " << v[l];
                for (int face = 0; face < 6; face++)
                    if (!s.disable[face]) f << std::setprecision(15) << " " << v[(size_t)(1 + face) * Nl + l];
                f << std::endl;
            }
            const fs::path core(name);
            fm << "function [L,tot]=" << core.stem() << "\n";
            fm << "   data=dlmread('" << core.stem() << "_bspdft');\n";
            fm << "   L=data(:,1);\n";
            fm << "   tot=data(:,2);\n";
            fm << "end";
            off += 7 * (size_t)Nl;
        } else {
            // Spect_Poynting::treat (sensors/s_spec_power.cpp:44-118)
            int span1 = 0, span2 = 0;
            TRY(aefdtd::sensor_spans(g, r->ids[q][0], &span1, &span2));
            double area = span1 * span2;
            const int t = s.orientation;
            if (t == AEFDTD_NORMAL_X || t == AEFDTD_NORMAL_XM) area *= g->Dy * g->Dz;
            else if (t == AEFDTD_NORMAL_Y || t == AEFDTD_NORMAL_YM) area *= g->Dx * g->Dz;
            else area *= g->Dx * g->Dy;
            std::ofstream f;
            if (!open(dir / (name + "_pspdft"), f)) return AEFDTD_EINVAL;
            for (int l = 0; l < Nl; l++) f << std::setprecision(15) << L[l] << " " << v[l] << " " << v[l] / area << std::endl;
            off += Nl;
        }
    }
    if (r->kind == 0 && res && res->lambda) {
        // FDTD_normal_incidence's spectra (modes/normal_incidence.cpp:490-557)
        const std::string pre = prefix ? prefix : "";
        std::ofstream f1, f2;
        if (!open(dir / (pre + "spectdata_norm"), f1) || !open(dir / (pre + "spectdata_norm2"), f2))
            return AEFDTD_EINVAL;
        using C = std::complex<double>;
        for (int l = 0; l < r->Nl; l++) {
            f1 << res->lambda[l] << " " << res->R[l] << " " << res->T[l] << std::endl;
            const C rx(res->refx[2 * l], res->refx[2 * l + 1]), ry(res->refy[2 * l], res->refy[2 * l + 1]);
            const C tx(res->transx[2 * l], res->transx[2 * l + 1]), ty(res->transy[2 * l], res->transy[2 * l + 1]);
            const double prx = std::norm(rx), pry = std::norm(ry);
            const double ptx = std::norm(tx) * res->index_sub / res->index_sup;
            const double pty = std::norm(ty) * res->index_sub / res->index_sup;
            f2 << res->lambda[l] << " " << std::abs(rx) << " " << std::arg(rx) << " " << prx << " " << std::abs(ry)
               << " " << std::arg(ry) << " " << pry << " " << std::abs(tx) << " " << std::arg(tx) << " " << ptx << " "
               << std::abs(ty) << " " << std::arg(ty) << " " << pty << std::endl;
        }
    }
    return AEFDTD_OK;
}

// FDTD_normal_incidence (modes/normal_incidence.cpp:161-607) in one call.
int aefdtd_run_normal_incidence(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                                aefdtd_normal_result *res, double *sensor_out, double *fields_out,
                                const int *end_computation) {
    if (!res) return AEFDTD_EINVAL;
    return run_one_call(0, m, n_sensors, sensors, res, sensor_out, fields_out, end_computation);
}

// ---- batches of custom-mode runs
int aefdtd_batch_create(aefdtd_batch **out, int precision) {
    if (!out) return AEFDTD_EINVAL;
    *out = nullptr;
    Batch *b = nullptr;
    TRY(batch_create(&b, precision));
    *out = new aefdtd_batch();
    (*out)->b = b;
    return AEFDTD_OK;
}

int aefdtd_batch_destroy(aefdtd_batch *b) {
    if (!b) return AEFDTD_OK;
    for (aefdtd_run *r : b->runs) r->batch = nullptr;
    batch_destroy(b->b);
    delete b;
    return AEFDTD_OK;
}

const char *aefdtd_batch_error(const aefdtd_batch *b) {
    if (!b) return "";
    return b->err.empty() ? batch_error(b->b) : b->err.c_str();
}

int aefdtd_batch_add(aefdtd_batch *b, aefdtd_run *r) {
    if (!b || !r) return AEFDTD_EINVAL;
    b->err.clear();
    if (r->time_type != 0) {
        b->err = "runs that stop on their completion check are stepped on their own";
        return AEFDTD_ESTATE;
    }
    if (r->kind != 2) {
        b->err = "only custom-mode runs are batched";
        return AEFDTD_ESTATE;
    }
    if (r->batch) {
        b->err = "run is already in a batch";
        return AEFDTD_ESTATE;
    }
    TRY(batch_add(b->b, r->g));
    b->runs.push_back(r);
    r->batch = b;
    return AEFDTD_OK;
}

int aefdtd_batch_remove(aefdtd_batch *b, aefdtd_run *r) {
    if (!b || !r) return AEFDTD_EINVAL;
    b->err.clear();
    auto it = std::find(b->runs.begin(), b->runs.end(), r);
    if (it == b->runs.end()) {
        b->err = "run is not in the batch";
        return AEFDTD_EINVAL;
    }
    TRY(batch_remove(b->b, r->g));
    b->runs.erase(it);
    r->batch = nullptr;
    return AEFDTD_OK;
}

int aefdtd_batch_step(aefdtd_batch *b, int chunks) {
    if (!b) return AEFDTD_EINVAL;
    b->err.clear();
    TRY(batch_step(b->b, chunks));
    for (aefdtd_run *r : b->runs) r->t = r->g->tstep;
    return AEFDTD_OK;
}

int aefdtd_batch_running(const aefdtd_batch *b) {
    if (!b) return 0;
    int n = 0;
    for (const aefdtd_run *r : b->runs) n += r->t < r->Nt;
    return n;
}

int aefdtd_batch_synchronize(aefdtd_batch *b) {
    if (!b) return AEFDTD_EINVAL;
    return batch_synchronize(b->b);
}

int aefdtd_batch_set_tiles(aefdtd_batch *b, int tile_e, int tile_h) {
    if (!b) return AEFDTD_EINVAL;
    return batch_set_tiles(b->b, tile_e, tile_h);
}

int aefdtd_batch_set_skew(aefdtd_batch *b, int code) {
    if (!b) return AEFDTD_EINVAL;
    return batch_set_skew(b->b, code);
}

int aefdtd_batch_info(const aefdtd_batch *b, int out[4]) {
    if (!b || !out) return AEFDTD_EINVAL;
    out[0] = batch_members(b->b);
    out[1] = aefdtd_batch_running(b);
    out[2] = batch_skewed_members(b->b);
    out[3] = batch_skew(b->b);
    return AEFDTD_OK;
}

namespace {

// Times one chunk of the batch with its current configuration (the first
// replay after a capture also uploads the graph, so the second one is timed);
// returns false when the members ran out of steps first.
bool time_chunk(aefdtd_batch *b, double &ms, int &rc) {
    ms = 1e30;
    if (aefdtd_batch_running(b) == 0) return false;
    if ((rc = aefdtd_batch_step(b, 1))) return false;
    if (aefdtd_batch_running(b) == 0) return false;
    long n = 0;
    batch_profile(b->b, 1, nullptr, nullptr);
    if ((rc = aefdtd_batch_step(b, 1))) return false;
    batch_profile(b->b, 0, &ms, &n);
    return n > 0;
}

}  // namespace

// Online tuner over the real run.  Candidates: the per-step passes and the
// large time-skewed tiles at depths 4 and 8 when 2D TE/TM members are
// present; then, for the members left to the per-step passes, the tile codes
// with the XCD fold and BY in {2,4,8}, KZ in {1,2,4,8} (E and H share a
// code).  Every candidate computes the same steps, so tuning advances the
// members like any other chunk.
int aefdtd_batch_autotune(aefdtd_batch *b, int out[3]) {
    if (!b) return AEFDTD_EINVAL;
    int rc = AEFDTD_OK;
    if (batch_skewed_members(b->b) > 0 || batch_skew(b->b) < 0) {
        int best_s = batch_skew(b->b);
        double best = 1e30;
        for (int code : {-1, 1 | (1 << 2), 1 | (2 << 2)}) {
            if (batch_set_skew(b->b, code) != AEFDTD_OK) continue;
            if (code >= 0 && batch_skewed_members(b->b) == 0) continue;
            double ms;
            if (!time_chunk(b, ms, rc)) break;
            if (ms < best) { best = ms; best_s = code; }
        }
        if (rc) return rc;
        TRY(batch_set_skew(b->b, best_s));
    }
    // members left to the per-step passes
    const bool stepped = batch_skew(b->b) < 0 || batch_skewed_members(b->b) < (int)b->runs.size();
    int best_c = 2 | (2 << 2) | (1 << 4);
    if (stepped) {
        double best = 1e30;
        for (int by = 1; by < 4; by++)
            for (int kz = 0; kz < 4; kz++) {
                const int c = by | (kz << 2) | (1 << 4);
                TRY(batch_set_tiles(b->b, c, c));
                double ms;
                if (!time_chunk(b, ms, rc)) break;
                if (ms < best) { best = ms; best_c = c; }
            }
        if (rc) return rc;
    }
    TRY(batch_set_tiles(b->b, best_c, best_c));
    if (out) { out[0] = best_c; out[1] = best_c; out[2] = batch_skew(b->b); }
    return AEFDTD_OK;
}

int aefdtd_batch_profile(aefdtd_batch *b, int on, double *ms, long *chunks) {
    if (!b) return AEFDTD_EINVAL;
    batch_profile(b->b, on, ms, chunks);
    return AEFDTD_OK;
}

// mode_default_fdtd (modes/default_fdtd.cpp:24-210) in one call.
int aefdtd_run_custom(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors, double *sensor_out,
                      double *fields_out, const int *end_computation) {
    return run_one_call(2, m, n_sensors, sensors, nullptr, sensor_out, fields_out, end_computation);
}

// FDTD_single_particle (modes/single_particle.cpp:29-415) in one call.
int aefdtd_run_single_particle(const aefdtd_mode *m, int n_sensors, const aefdtd_sensor_req *sensors,
                               double *sensor_out, double *fields_out, const int *end_computation) {
    return run_one_call(1, m, n_sensors, sensors, nullptr, sensor_out, fields_out, end_computation);
}

}  // extern "C"
