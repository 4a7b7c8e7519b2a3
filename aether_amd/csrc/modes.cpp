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
#include <sstream>
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

// In-plane extents (span1, span2) of a planar sensor of orientation t over
// the box x1..z2 (Sensor::link's spans, sensors.cpp:196-230).
void plane_spans(int t, int dx, int dy, int dz, int &s1, int &s2) {
    if (t == AEFDTD_NORMAL_X || t == AEFDTD_NORMAL_XM) { s1 = dy; s2 = dz; }
    else if (t == AEFDTD_NORMAL_Y || t == AEFDTD_NORMAL_YM) { s1 = dx; s2 = dz; }
    else { s1 = dx; s2 = dy; }
}

}  // namespace

// Doubles a request contributes to sensor_out (the AEFDTD_SENSOR_* table of
// aether_fdtd.h).
size_t aefdtd::request_out_len(const aefdtd_sensor_req &q) {
    const size_t dx = std::max(0, q.x2 - q.x1), dy = std::max(0, q.y2 - q.y1), dz = std::max(0, q.z2 - q.z1);
    int s1 = 0, s2 = 0;
    switch (q.kind) {
        case AEFDTD_SENSOR_FIELDBLOCK: return 6 * dx * dy * dz * 2;
        case AEFDTD_SENSOR_FIELDMAP:
            plane_spans(2 * (q.orientation / 2), (int)dx, (int)dy, (int)dz, s1, s2);
            return 3 * (size_t)s1 * s2 * 2;
        case AEFDTD_SENSOR_FIELDMAP2:
            plane_spans(q.orientation, (int)dx, (int)dy, (int)dz, s1, s2);
            return 6 * (size_t)std::max(q.Nl, 0) * s1 * s2 * 2;
        default: return q.box ? 7 * (size_t)std::max(q.Nl, 0) : (size_t)std::max(q.Nl, 0);
    }
}

extern "C" size_t aefdtd_sensor_out_len(const aefdtd_sensor_req *q) { return q ? aefdtd::request_out_len(*q) : 0; }

namespace {

// Box_Spect_Poynting (sensors/s_spec_power.cpp:125-232) becomes six planar
// sensors; a planar request or a FieldMap2 one sensor; FieldBlock and
// FieldMap their field sensors (ids kept per request).
int add_requests(aefdtd_run *r) {
    for (const aefdtd_sensor_req &q : r->reqs) {
        if (q.kind < AEFDTD_SENSOR_POYNTING || q.kind > AEFDTD_SENSOR_FIELDMAP2) return AEFDTD_EINVAL;
        if ((q.kind == AEFDTD_SENSOR_POYNTING || q.kind == AEFDTD_SENSOR_FIELDMAP2) && q.Nl < 1) return AEFDTD_EINVAL;
        std::vector<int> v;
        if (q.kind == AEFDTD_SENSOR_FIELDBLOCK || q.kind == AEFDTD_SENSOR_FIELDMAP) {
            int id = -1;
            if (q.kind == AEFDTD_SENSOR_FIELDBLOCK)
                TRY(aefdtd_add_field_block(r->g, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, &id));
            else
                TRY(aefdtd_add_field_map(r->g, q.orientation, q.x1, q.x2, q.y1, q.y2, q.z1, q.z2, q.lambda_min, 0, 0, &id));
            int dims[4];
            TRY(aefdtd_field_sensor_read(r->g, id, nullptr, dims));
            if ((size_t)dims[0] * dims[1] * dims[2] * dims[3] * 2 != aefdtd::request_out_len(q)) {
                r->g->err = "field sensor reaches outside the grid";
                return AEFDTD_EINVAL;
            }
            v.push_back(id);
            r->ids.push_back(v);
            continue;
        }
        std::vector<double> L;
        lambda_grid(q.Nl, q.lambda_min, q.lambda_max, L);
        if (q.box && q.kind == AEFDTD_SENSOR_POYNTING) {
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
            if (q.kind == AEFDTD_SENSOR_FIELDMAP2) {
                int s1 = 0, s2 = 0;
                TRY(aefdtd::sensor_spans(r->g, id, &s1, &s2));
                if (6 * (size_t)q.Nl * s1 * s2 * 2 != aefdtd::request_out_len(q)) {
                    r->g->err = "field sensor reaches outside the grid";
                    return AEFDTD_EINVAL;
                }
            }
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
        const int kind = r->reqs[q].kind;
        if (kind != AEFDTD_SENSOR_POYNTING) {
            // field sensors: their accumulators as they are
            const size_t n = aefdtd::request_out_len(r->reqs[q]);
            if (out) {
                if (kind == AEFDTD_SENSOR_FIELDMAP2) {
                    int dims[3];
                    TRY(aefdtd_sensor_spectra(r->g, r->ids[q][0], out + off, dims));
                } else {
                    int dims[4];
                    TRY(aefdtd_field_sensor_read(r->g, r->ids[q][0], out + off, dims));
                }
            }
            off += n;
            continue;
        }
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
        if (r->g->precision == AEFDTD_FP16 && r->t % 64 == 0) {
            // FP16 storage range control every 64 steps (a batch checks after each
            // 64-step chunk, at the same steps)
            bool ch = false;
            if ((rc = aefdtd::fp16_range(r->g, &ch)) || (r->a && (rc = aefdtd::fp16_range(r->a, &ch))))
                return rfail(r, rc);
        }
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

}  // extern "C"

namespace {

template <typename V>
void put(std::ofstream &f, V v) { f.write(reinterpret_cast<const char *>(&v), sizeof(V)); }

// FieldBlock::treat (sensors/s_fieldmap.cpp:232-287): <name>.afblock, binary.
bool write_fieldblock(aefdtd_run *r, const std::filesystem::path &dir, const std::string &name, int id,
                      const double *v) {
    aefdtd_grid *g = r->g;
    int geo[8];
    double lambda = 0;
    if (aefdtd::field_sensor_geom(g, id, geo, &lambda)) return false;
    const int x1 = geo[2], y1 = geo[3], z1 = geo[4], s1 = geo[5], s2 = geo[6], s3 = geo[7];
    const size_t n = (size_t)s1 * s2 * s3;
    std::vector<uint32_t> mats(n);
    if (aefdtd_field_sensor_mats(g, id, mats.data())) return false;
    std::ofstream f(dir / (name + ".afblock"), std::ios::out | std::ios::trunc | std::ios::binary);
    if (!f) return false;
    put(f, lambda); put(f, x1); put(f, s1); put(f, y1); put(f, s2); put(f, z1); put(f, s3);
    put(f, g->Dx); put(f, g->Dy); put(f, g->Dz);
    for (int i = 0; i < s1; i++)
        for (int j = 0; j < s2; j++)
            for (int k = 0; k < s3; k++) {
                const size_t q = (size_t)i + (size_t)s1 * ((size_t)j + (size_t)s2 * k);
                for (int c = 0; c < 6; c++) { put(f, v[2 * (c * n + q)]); put(f, v[2 * (c * n + q) + 1]); }
                put(f, mats[q]);
            }
    return (bool)f;
}

// FieldMap::treat (s_fieldmap.cpp:470-506, PNG excepted): <name>_fieldmap
// (binary) and, through fmap_script / fmap_raw / fmap_mats_raw
// (sensors/fieldblock_holder.cpp:83-361), <name>.m, <name>_E{x,y,z}_raw and
// <name>_mats_raw.
bool write_fieldmap(aefdtd_run *r, const std::filesystem::path &dir, const std::string &name, int id,
                    const double *v) {
    namespace fs = std::filesystem;
    aefdtd_grid *g = r->g;
    int geo[8];
    double lambda = 0;
    if (aefdtd::field_sensor_geom(g, id, geo, &lambda)) return false;
    const int normal = geo[1], x1 = geo[2], y1 = geo[3], z1 = geo[4], s1 = geo[5], s2 = geo[6];
    const size_t n = (size_t)s1 * s2;
    std::vector<uint32_t> mats(n);
    if (aefdtd_field_sensor_mats(g, id, mats.data())) return false;
    auto re = [&](int c, int i, int j) { return v[2 * (c * n + (size_t)i + (size_t)s1 * j)]; };
    auto im = [&](int c, int i, int j) { return v[2 * (c * n + (size_t)i + (size_t)s1 * j) + 1]; };
    {
        std::ofstream f(dir / (name + "_fieldmap"), std::ios::out | std::ios::trunc | std::ios::binary);
        if (!f) return false;
        const int type = 2 * normal;
        put(f, lambda); put(f, type);
        put(f, normal == 0 ? y1 : x1);
        put(f, s1);
        put(f, normal == 2 ? y1 : z1);
        put(f, s2);
        put(f, g->Dx); put(f, g->Dy); put(f, g->Dz);
        for (int i = 0; i < s1; i++)
            for (int j = 0; j < s2; j++)
                for (int c = 0; c < 3; c++) { put(f, re(c, i, j)); put(f, im(c, i, j)); }
        if (!f) return false;
    }
    const fs::path base = dir / name;
    const std::string stem = base.stem().generic_string();
    const char *raw[3] = {"_Ex_raw", "_Ey_raw", "_Ez_raw"};
    {
        fs::path mfile = base;
        if (mfile.extension() != ".m") mfile.replace_extension(".m");
        std::ofstream f(mfile, std::ios::out | std::ios::trunc | std::ios::binary);
        if (!f) return false;
        const double D1 = 1.0, D2 = 1.0;
        f << "function out=" << stem << "(varargin)" << '\n' << "" << '\n' << "baseline=1;" << '\n' << "" << '\n'
          << "x=linspace(0,6,1000);" << '\n' << "" << '\n'
          << "r=(x<2)+(3.0-x).*(x>=2).*(x<3)+(x-5).*(x>=5).*(x<6)+(x>=6);" << '\n'
          << "g=(x-1.0).*(x>=1).*(x<2)+(x>=2).*(x<4)+(5.0-x).*(x>=4).*(x<5);" << '\n'
          << "b=(1.0-x).*(x<1)+(x-3).*(x>=3).*(x<4)+(x>=4);" << '\n' << "" << '\n'
          << "cmap=[transpose(r) transpose(g) transpose(b)];" << '\n' << "" << '\n'
          << "Fx=dlmread('" << stem << raw[0] << "')/baseline;" << '\n'
          << "Fy=dlmread('" << stem << raw[1] << "')/baseline;" << '\n'
          << "Fz=dlmread('" << stem << raw[2] << "')/baseline;" << '\n'
          << "mats=dlmread('" << stem << "_mats_raw" << "');" << '\n' << "" << '\n'
          << "Nx=size(Fx,2);" << '\n' << "Ny=size(Fy,2);" << '\n' << "Nz=size(Fz,2);" << '\n' << "" << '\n'
          << "Fx_r=Fx(:,1:2:Nx);" << '\n' << "Fx_i=Fx(:,2:2:Nx);" << '\n' << "Fx=Fx_r+i*Fx_i;" << '\n' << "" << '\n'
          << "Fy_r=Fy(:,1:2:Ny);" << '\n' << "Fy_i=Fy(:,2:2:Ny);" << '\n' << "Fy=Fy_r+i*Fy_i;" << '\n' << "" << '\n'
          << "Fz_r=Fz(:,1:2:Nz);" << '\n' << "Fz_i=Fz(:,2:2:Nz);" << '\n' << "Fz=Fz_r+i*Fz_i;" << '\n'
          << "" << '\n' << "N1=size(Fx,2);" << '\n' << "N2=size(Fx,1);" << '\n'
          << "N1t=ones(N2,1)*(1:N1)*" << D1 << ";" << '\n' << "N2t=transpose((1:N2)*" << D2 << ")*ones(1,N1);" << '\n'
          << "N1=N1t;" << '\n' << "N2=N2t;" << '\n' << '\n' << "silent=0;" << '\n' << '\n'
          << "if nargin>0" << '\n' << "    j=1;" << '\n' << "    while j<=nargin" << '\n'
          << "        if strcmp(varargin{j},'saturate')" << '\n' << "            val=varargin{j+1}/sqrt(3.0);" << '\n'
          << "            Fx=Fx.*(abs(Fx)<val)+val*exp(i*angle(Fx)).*(abs(Fx)>=val);" << '\n'
          << "            Fy=Fy.*(abs(Fy)<val)+val*exp(i*angle(Fy)).*(abs(Fy)>=val);" << '\n'
          << "            Fz=Fz.*(abs(Fz)<val)+val*exp(i*angle(Fz)).*(abs(Fz)>=val);" << '\n'
          << "            j=j+2;" << '\n' << "            continue;" << '\n' << "        end" << '\n'
          << "        if strcmp(varargin{j},'silent')" << '\n' << "            silent=1;" << '\n' << "            j=j+1;" << '\n'
          << "            continue;" << '\n' << "        end" << '\n' << "    end" << '\n' << "end" << '\n' << '\n'
          << "F_tot=sqrt(abs(Fx).^2+abs(Fy).^2+abs(Fz).^2);" << '\n' << '\n' << "if silent==0" << '\n';
        const char *lab[3] = {"x", "y", "z"};
        for (int c = 0; c < 3; c++)
            f << "    figure(" << c + 1 << ")" << '\n' << "    pcolor(N1,N2,abs(F" << lab[c] << "))" << '\n'
              << "    colormap(jet(500))" << '\n' << "    colorbar" << '\n' << "    title('|E_" << lab[c] << "|')" << '\n'
              << "    shading flat" << '\n';
        f << "" << '\n';
        for (int c = 0; c < 3; c++)
            f << "    figure(" << c + 4 << ")" << '\n' << "    pcolor(N1,N2,angle(F" << lab[c] << "))" << '\n'
              << "    colormap(cmap)" << '\n' << "    shading flat" << '\n' << "    title('arg(E_" << lab[c] << ")')" << '\n';
        f << "" << '\n' << "    figure(7)" << '\n' << "    pcolor(N1,N2,F_tot)" << '\n' << "    title('|E|')" << '\n'
          << "    colormap(jet(500))" << '\n' << "    colorbar" << '\n' << "    shading flat" << '\n' << "end" << '\n'
          << "" << '\n' << "out.Ex=Fx;" << '\n' << "out.Ey=Fy;" << '\n' << "out.Ez=Fz;" << '\n' << "out.mats=mats;" << '\n'
          << "" << '\n' << "end" << '\n';
        if (!f) return false;
    }
    for (int c = 0; c < 3; c++) {
        std::ofstream f(base.parent_path() / (stem + raw[c]), std::ios::out | std::ios::trunc);
        if (!f) return false;
        std::stringstream strm;
        for (int j = 0; j < s2; j++) {
            for (int i = 0; i < s1; i++) strm << re(c, i, j) << " " << im(c, i, j) << " ";
            strm << std::endl;
        }
        f << strm.str();
    }
    std::ofstream f(base.parent_path() / (stem + "_mats_raw"), std::ios::out | std::ios::trunc);
    if (!f) return false;
    std::stringstream strm;
    for (int j = 0; j < s2; j++) {
        for (int i = 0; i < s1; i++) strm << mats[(size_t)i + (size_t)s1 * j] << " ";
        strm << std::endl;
    }
    f << strm.str();
    return (bool)f;
}

// FieldMap2::treat (s_fieldmap.cpp:594-672): <name>_fieldmap, text.  The
// coordinate lines add the structure offset to the linked (already offset)
// indices, and the z line is always written (its test is `type != NORMAL_Z
// || type != NORMAL_ZM`), as the reference does.
bool write_fieldmap2(aefdtd_run *r, const std::filesystem::path &dir, const std::string &name, int id,
                     const aefdtd_sensor_req &q, const double *v) {
    aefdtd_grid *g = r->g;
    int b[6], s1 = 0, s2 = 0;
    if (aefdtd::sensor_box(g, id, b) || aefdtd::sensor_spans(g, id, &s1, &s2)) return false;
    std::vector<double> L;
    lambda_grid(q.Nl, q.lambda_min, q.lambda_max, L);
    const int type = q.orientation;
    std::ofstream f(dir / (name + "_fieldmap"), std::ios::out | std::ios::trunc | std::ios::binary);
    if (!f) return false;
    f << 0 << " " << type << " 1 1 1 1 1 1 1\n";
    for (int l = 0; l < q.Nl; l++) {
        f << L[l];
        if (l + 1 != q.Nl) f << " ";
    }
    f << "\n";
    if (type != AEFDTD_NORMAL_X && type != AEFDTD_NORMAL_XM) {
        for (int i = b[0]; i < b[1]; i++) {
            f << (g->xs_s + i) * g->Dx;
            if (i + 1 != b[1]) f << " ";
        }
        f << "\n";
    }
    if (type != AEFDTD_NORMAL_Y && type != AEFDTD_NORMAL_YM) {
        for (int j = b[2]; j < b[3]; j++) {
            f << (g->ys_s + j) * g->Dy;
            if (j + 1 != b[3]) f << " ";
        }
        f << "\n";
    }
    for (int k = b[4]; k < b[5]; k++) {
        f << (g->zs_s + k) * g->Dz;
        if (k + 1 != b[5]) f << " ";
    }
    f << "\n";
    for (int j = 0; j < s2; j++) {
        for (int i = 0; i < s1; i++) {
            f << 0;
            if (i + 1 != s1) f << " ";
        }
        f << "\n";
    }
    const size_t P = (size_t)s1 * s2;
    for (int l = 0; l < q.Nl; l++)
        for (int c = 0; c < 6; c++)
            for (int j = 0; j < s2; j++) {
                for (int i = 0; i < s1; i++) {
                    const size_t o = 2 * (((size_t)c * q.Nl + l) * P + (size_t)i + (size_t)s1 * j);
                    f << v[o] << " " << v[o + 1];
                    if (i + 1 != s1) f << " ";
                }
                f << "\n";
            }
    return (bool)f;
}

}  // namespace

extern "C" {

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
        if (s.kind != AEFDTD_SENSOR_POYNTING) {
            const int id = r->ids[q][0];
            bool ok = s.kind == AEFDTD_SENSOR_FIELDBLOCK ? write_fieldblock(r, dir, name, id, v)
                      : s.kind == AEFDTD_SENSOR_FIELDMAP ? write_fieldmap(r, dir, name, id, v)
                                                         : write_fieldmap2(r, dir, name, id, s, v);
            if (!ok) {
                r->err = "cannot write the files of sensor " + name;
                return AEFDTD_EINVAL;
            }
            off += aefdtd::request_out_len(s);
            continue;
        }
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
        {
            // norm_octave_script (normal_incidence.cpp:24-59, called at :601)
            std::ofstream f;
            if (!open(dir / (pre + "show_norm.m"), f)) return AEFDTD_EINVAL;
            const std::string stem = fs::path((dir / pre).generic_string()).stem().generic_string();
            using std::endl;
            f << "function out=" << stem << "show_norm(varargin)" << endl << endl;
            f << "D=load('" << stem + "spectdata_norm" << "');" << endl;
            f << "L=D(:,1)*1e6;" << endl << "r=D(:,2);" << endl << "t=D(:,3);" << endl << endl;
            f << "R=r.^2;" << endl << "T=" << res->index_sub << "/" << res->index_sup << "*t.^2;" << endl
              << "A=1-R-T;" << endl << endl;
            f << "figure(1)" << endl << "plot(L,r,L,t)" << endl << "axis([min(L) max(L) 0 1])" << endl << endl;
            f << "figure(2)" << endl << "plot(L,R,L,T,L,A)" << endl << "legend('R','T','A')" << endl
              << "xlabel('Lambda (microns)')" << endl << "axis([min(L) max(L) 0 1])" << endl << endl;
            f << "if nargin>0" << endl << "    out=[L,R,T,A];" << endl << "end" << endl << endl << "end" << endl;
        }
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
    if (r->kind == 0) {
        b->err = "normal-incidence runs (plane sums) are stepped on their own";
        return AEFDTD_ESTATE;
    }
    if (r->batch) {
        b->err = "run is already in a batch";
        return AEFDTD_ESTATE;
    }
    // a single-particle run brings its TFSF source column, stepped in the same launches
    if (r->a) TRY(batch_add(b->b, r->a));
    if (int rc = batch_add(b->b, r->g)) {
        if (r->a) (void)batch_remove(b->b, r->a);
        return rc;
    }
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
    if (r->a) TRY(batch_remove(b->b, r->a));
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
