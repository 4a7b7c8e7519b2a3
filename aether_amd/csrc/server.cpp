// server.cpp -- the request server of the serving form.  Builder threads
// admit requests in FIFO order while their estimated bytes fit a device-memory
// budget and build their runs (host set-up and uploads on each run's own
// stream); a worker thread keeps every admitted run resident and steps all of
// them together in rounds of 64 steps until their spectra are out (continuous
// batching).  Rounds are pipelined one deep: round R is queued on every stream
// before the worker waits for round R-1, so the device always has a round of
// work while the host admits and retires runs.  The reference has no such
// layer: it runs one mode driver per call (modes/lua_fdtd.cpp:296-320).
#include "runs.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

using namespace aefdtd;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kRound = 64;     // steps per round (one batch chunk)
constexpr int kBuilders = 3;   // threads admitting and building runs
// Runs stepped on their own streams in one round, in admission order, until
// their cells reach this many (at least one run): enough concurrent work to
// fill the device with small grids, without large grids contending for it.
constexpr size_t kActiveCells = 48u << 20;

double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// A request with its own copy of every input the driver reads.
struct Request {
    long id = 0;
    int kind = 0;
    int state = AEFDTD_REQ_QUEUED;
    int rc = AEFDTD_OK;
    std::string err;
    aefdtd_mode m{};
    std::vector<uint32_t> mg;
    std::vector<aefdtd_material> mats;
    std::vector<std::vector<double>> poles;
    std::vector<int> disable;
    std::vector<aefdtd_source_req> srcs;
    std::vector<aefdtd_sensor_req> sens;
    size_t est = 0, bytes = 0;
    aefdtd_run *run = nullptr;
    aefdtd_batch *batch = nullptr;   // set when stepped by a batch
    hipEvent_t ev[2] = {nullptr, nullptr};   // end of the run's rounds (even, odd) on its stream
    bool ev_set[2] = {false, false};
    long fin_round = -1;             // round that queued the run's last steps
    std::vector<double> out;
    Clock::time_point t_submit, t_admit, t_done;
    bool cancel = false;
};

void copy_inputs(Request &q, const aefdtd_mode *m, int ns, const aefdtd_sensor_req *sens) {
    q.m = *m;
    q.mg.assign(m->matsgrid, m->matsgrid + (size_t)m->nx * m->ny * m->nz);
    q.m.matsgrid = q.mg.data();
    q.mats.assign(m->materials, m->materials + std::max(0, m->n_materials));
    q.poles.reserve(4 * q.mats.size());
    for (aefdtd_material &a : q.mats) {
        auto keep = [&](const double *p, int n, int w) -> const double * {
            if (n <= 0 || !p) return nullptr;
            q.poles.emplace_back(p, p + (size_t)n * w);
            return q.poles.back().data();
        };
        a.debye = keep(a.debye, a.n_debye, 2);
        a.drude = keep(a.drude, a.n_drude, 2);
        a.lorentz = keep(a.lorentz, a.n_lorentz, 3);
        a.critpoint = keep(a.critpoint, a.n_critpoint, 4);
    }
    q.m.materials = q.mats.data();
    if (m->n_disable > 0) q.disable.assign(m->disable_fields, m->disable_fields + m->n_disable);
    q.m.disable_fields = q.disable.data();
    if (m->n_sources > 0) q.srcs.assign(m->sources, m->sources + m->n_sources);
    q.m.sources = q.srcs.data();
    if (ns > 0) q.sens.assign(sens, sens + ns);
}

size_t out_len(const Request &q) {
    size_t n = 0;
    for (const aefdtd_sensor_req &s : q.sens) n += aefdtd::request_out_len(s);
    if (q.kind == 0) n += 2 * (size_t)q.m.Nl;
    return n;
}

bool has_poles(const aefdtd_material &a) { return a.n_debye + a.n_drude + a.n_lorentz + a.n_critpoint > 0; }

}  // namespace

struct aefdtd_server {
    size_t budget = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> queue;
    std::vector<Request *> built;         // admitted and built, not yet stepped
    std::vector<Request *> resident;
    std::map<long, std::unique_ptr<Request>> all;
    long next_id = 1;
    bool stop = false;
    size_t used = 0, peak = 0;
    long rounds = 0, n_done = 0, n_failed = 0;
    aefdtd_batch *batch[3] = {nullptr, nullptr, nullptr};   // per precision
    hipEvent_t bev[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    bool bev_set[3][2] = {{false, false}, {false, false}, {false, false}};
    std::thread worker;
    std::vector<std::thread> builders;
    std::string err;
};

namespace {

// Resident bytes of a run: its grid and the auxiliary grid of a single-particle run.
size_t run_bytes(aefdtd_run *r) {
    size_t a = 0, b = 0;
    if (r->g) aefdtd_device_bytes(r->g, &a);
    if (r->a) aefdtd_device_bytes(r->a, &b);
    return a + b;
}

void finish_request(aefdtd_server *S, Request *q, int state, int rc, const std::string &err) {
    // called with S->mu held
    q->state = state;
    q->rc = rc;
    if (!err.empty()) q->err = err;
    q->t_done = Clock::now();
    if (state == AEFDTD_REQ_DONE) S->n_done++;
    else if (state == AEFDTD_REQ_FAILED) S->n_failed++;
    S->cv.notify_all();
}

void free_events(Request *q) {
    for (hipEvent_t &e : q->ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

// Builds the run of a request (builder thread, lock not held).
int build(Request *q) {
    int rc = aefdtd_run_create(&q->run, q->kind, &q->m, (int)q->sens.size(), q->sens.data());
    if (rc) {
        q->err = aefdtd_last_error(nullptr);
        return rc;
    }
    q->bytes = run_bytes(q->run);
    return AEFDTD_OK;
}

// Puts a built run into the batch of its precision when the batched passes
// can step it (worker thread: the batches are the worker's).  A run the batch
// refuses still runs on its own.
int join_batch(aefdtd_server *S, Request *q) {
    if (q->kind != 2 || q->run->g->n_disp || q->m.time_type) return AEFDTD_OK;
    const int p = q->m.precision;
    int rc;
    if (!S->batch[p] && (rc = aefdtd_batch_create(&S->batch[p], p))) {
        q->err = "batch creation failed";
        return rc;
    }
    if (aefdtd_batch_add(S->batch[p], q->run) == AEFDTD_OK) q->batch = S->batch[p];
    else q->err = aefdtd_batch_error(S->batch[p]);
    return AEFDTD_OK;
}

// Spectra of a finished run, then its device memory is released (worker thread).
int retire(Request *q) {
    int rc = AEFDTD_OK;
    q->bytes = std::max(q->bytes, run_bytes(q->run));   // stage and DFT tables grow after admission
    if (q->batch) {
        rc = aefdtd_batch_remove(q->batch, q->run);
        q->batch = nullptr;
    }
    if (!rc) {
        q->out.assign(out_len(*q), 0.0);
        size_t nsens = q->out.size() - (q->kind == 0 ? 2 * (size_t)q->m.Nl : 0);
        aefdtd_normal_result res{};
        std::vector<double> lam(q->m.Nl), cx(4 * 2 * (size_t)std::max(1, q->m.Nl));
        if (q->kind == 0) {
            res.lambda = lam.data();
            res.R = q->out.data() + nsens;
            res.T = q->out.data() + nsens + q->m.Nl;
            res.refx = cx.data(); res.refy = cx.data() + 2 * q->m.Nl;
            res.transx = cx.data() + 4 * q->m.Nl; res.transy = cx.data() + 6 * q->m.Nl;
        }
        rc = aefdtd_run_finish(q->run, q->kind == 0 ? &res : nullptr, q->out.data(), nullptr);
        if (rc) q->err = aefdtd_run_error(q->run);
    }
    aefdtd_run_destroy(q->run);
    q->run = nullptr;
    free_events(q);
    return rc;
}

void drop(Request *q) {
    if (q->batch) aefdtd_batch_remove(q->batch, q->run);
    q->batch = nullptr;
    aefdtd_run_destroy(q->run);
    q->run = nullptr;
    free_events(q);
}

// Marks the end of round R on a stream (event slot R mod 2).
void mark(hipEvent_t &e, bool &set, hipStream_t st) {
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        e = nullptr;
        return;
    }
    set = hipEventRecord(e, st) == hipSuccess;
}

void wait_mark(hipEvent_t e, bool &set) {
    if (set) (void)hipEventSynchronize(e);
    set = false;
}

// Admission: FIFO under the budget; builds each admitted run outside the lock.
void builder_loop(aefdtd_server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // cancelled requests leave the queue wherever they are
        for (auto it = S->queue.begin(); it != S->queue.end();) {
            if ((*it)->cancel) {
                finish_request(S, *it, AEFDTD_REQ_CANCELLED, AEFDTD_ECANCELLED, "cancelled");
                it = S->queue.erase(it);
            } else {
                ++it;
            }
        }
        if (S->stop) break;
        if (S->queue.empty()) {
            S->cv.wait(lk);
            continue;
        }
        Request *q = S->queue.front();
        if (q->est > S->budget) {
            S->queue.pop_front();
            finish_request(S, q, AEFDTD_REQ_FAILED, AEFDTD_ENOMEM, "request exceeds the device-memory budget");
            continue;
        }
        if (S->used + q->est > S->budget) {      // wait for a retirement
            S->cv.wait(lk);
            continue;
        }
        S->queue.pop_front();
        S->used += q->est;                        // held while the run is built
        lk.unlock();
        int rc = build(q);
        lk.lock();
        S->used -= q->est;
        if (rc) {
            if (q->run) drop(q);
            finish_request(S, q, AEFDTD_REQ_FAILED, rc, q->err.empty() ? "admission failed" : q->err);
            continue;
        }
        S->used += std::max(q->bytes, q->est);
        S->peak = std::max(S->peak, S->used);
        q->state = AEFDTD_REQ_RUNNING;
        q->t_admit = Clock::now();
        S->built.push_back(q);
        S->cv.notify_all();
    }
}

void worker_loop(aefdtd_server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // ---- built runs join the resident set (cancelled ones are released)
        std::vector<Request *> fresh;
        fresh.swap(S->built);
        for (Request *q : fresh) {
            if (q->cancel) {
                lk.unlock();
                drop(q);
                lk.lock();
                S->used -= std::max(q->bytes, q->est);
                finish_request(S, q, AEFDTD_REQ_CANCELLED, AEFDTD_ECANCELLED, "cancelled");
                continue;
            }
            lk.unlock();
            int rc = join_batch(S, q);
            lk.lock();
            if (rc) {
                lk.unlock();
                drop(q);
                lk.lock();
                S->used -= std::max(q->bytes, q->est);
                finish_request(S, q, AEFDTD_REQ_FAILED, rc, q->err.empty() ? "admission failed" : q->err);
                continue;
            }
            S->resident.push_back(q);
        }
        if (S->resident.empty()) {
            if (S->stop && S->built.empty()) break;
            S->cv.wait(lk, [&] { return S->stop || !S->built.empty(); });
            continue;
        }
        std::vector<Request *> res = S->resident;
        const long R = S->rounds;
        lk.unlock();
        // ---- queue round R: every resident run with steps left advances up to 64
        std::vector<std::pair<Request *, int>> failed;
        const int cur = (int)(R & 1), prev = cur ^ 1;
        for (int p = 0; p < 3; p++) {
            aefdtd_batch *b = S->batch[p];
            if (!b || aefdtd_batch_running(b) == 0) continue;
            int rc = aefdtd_batch_step(b, 1);
            if (rc) {
                for (Request *q : res)
                    if (q->batch == b) failed.push_back({q, rc});
            } else {
                mark(S->bev[p][cur], S->bev_set[p][cur], batch_stream(b->b));
            }
        }
        size_t active = 0;
        for (Request *q : res) {
            if (q->batch || q->fin_round >= 0) continue;
            if (active >= kActiveCells) break;
            active += aefdtd_run_grid(q->run)->ncell();
            int rc = aefdtd_run_step(q->run, kRound);
            if (rc) failed.push_back({q, rc});
            else mark(q->ev[cur], q->ev_set[cur], aefdtd_run_grid(q->run)->stream);
        }
        for (Request *q : res) {
            int t = 0, nt = 0;
            aefdtd_run_progress(q->run, &t, &nt);
            if (q->fin_round < 0 && t >= nt) q->fin_round = R;
        }
        // ---- wait for round R-1 (round R stays queued behind it)
        for (int p = 0; p < 3; p++) wait_mark(S->bev[p][prev], S->bev_set[p][prev]);
        for (Request *q : res) wait_mark(q->ev[prev], q->ev_set[prev]);
        // ---- retire runs whose last steps were in an earlier round, failed and cancelled runs
        std::vector<std::pair<Request *, int>> ended;   // request, final state
        for (Request *q : res) {
            bool f = std::any_of(failed.begin(), failed.end(), [&](auto &p) { return p.first == q; });
            const bool done = q->fin_round >= 0 && q->fin_round < R;
            if (f || q->cancel || done)
                ended.push_back({q, f ? AEFDTD_REQ_FAILED : (q->cancel ? AEFDTD_REQ_CANCELLED : AEFDTD_REQ_DONE)});
        }
        std::vector<int> rcs;
        for (auto &e : ended) {
            Request *q = e.first;
            int rc = AEFDTD_OK;
            if (e.second == AEFDTD_REQ_DONE) {
                rc = retire(q);
            } else {
                if (e.second == AEFDTD_REQ_FAILED) {
                    for (auto &p : failed)
                        if (p.first == q) rc = p.second;
                    if (q->err.empty()) q->err = aefdtd_run_error(q->run);
                }
                drop(q);
            }
            rcs.push_back(rc);
        }
        lk.lock();
        S->rounds++;
        for (size_t i = 0; i < ended.size(); i++) {
            Request *q = ended[i].first;
            S->resident.erase(std::find(S->resident.begin(), S->resident.end(), q));
            S->used -= std::max(q->bytes, q->est);
            int st = ended[i].second;
            if (st == AEFDTD_REQ_DONE && rcs[i]) st = AEFDTD_REQ_FAILED;
            int rc = st == AEFDTD_REQ_CANCELLED ? AEFDTD_ECANCELLED : (st == AEFDTD_REQ_FAILED ? (rcs[i] ? rcs[i] : AEFDTD_EDEVICE) : AEFDTD_OK);
            finish_request(S, q, st, rc, "");
        }
    }
}

}  // namespace

extern "C" {

int aefdtd_estimate_bytes(int kind, const aefdtd_mode *m, int ns, const aefdtd_sensor_req *sens, size_t *bytes) {
    if (!m || !bytes || kind < 0 || kind > 2 || (ns > 0 && !sens) || !m->matsgrid) return AEFDTD_EINVAL;
    int d[3], pads[6];
    run_dims(kind, m, d, pads);
    const size_t N = (size_t)d[0] * d[1] * d[2];
    // cell arrays of 3D grids have rows padded to 64 cells
    const size_t px = (d[2] > 1 && d[0] > 1) ? (size_t)(d[0] + 63) / 64 * 64 : (size_t)d[0];
    const size_t NA = px * d[1] * d[2];
    const int p = m->precision;
    const size_t ss = p == AEFDTD_FP64 ? 8 : (p == AEFDTD_FP32 ? 4 : 2), cs = p == AEFDTD_FP64 ? 8 : 4;
    size_t b = 6 * NA * ss + NA;                            // fields, material ids
    // dispersive cells of the extended grid: edge planes repeat into pads and PML
    int npmax = 0;
    bool any = false;
    std::vector<char> rc(std::max(1, m->n_materials), 0);
    for (int i = 0; i < m->n_materials; i++)
        if (has_poles(m->materials[i])) {
            rc[i] = 1;
            any = true;
            const aefdtd_material &a = m->materials[i];
            npmax = std::max(npmax, a.n_debye + a.n_drude + a.n_lorentz + a.n_critpoint);
        }
    if (any) {
        int pml[6];
        for (int f = 0; f < 6; f++) pml[f] = kind == 0 && f < 4 ? 0 : m->pml[f];
        const int lo[3] = {pml[0] + pads[0], pml[2] + pads[2], pml[4] + pads[4]};
        const int nn[3] = {m->nx, m->ny, m->nz};
        auto mult = [&](int a, int v) {
            size_t w = 1;
            if (v == 0) w += lo[a];
            if (v == nn[a] - 1) w += d[a] - lo[a] - nn[a];
            return w;
        };
        size_t nd = 0;
        for (int k = 0; k < m->nz; k++)
            for (int j = 0; j < m->ny; j++) {
                const size_t wjk = mult(1, j) * mult(2, k);
                const uint32_t *row = m->matsgrid + ((size_t)k * m->ny + j) * m->nx;
                for (int i = 0; i < m->nx; i++)
                    if (row[i] < (uint32_t)m->n_materials && rc[row[i]]) nd += wjk * mult(0, i);
            }
        b += 4 * NA + nd * (8 + 3 * cs + 3 * (size_t)npmax * 16);   // slot map, (cell, material), increments, poles
    }
    const size_t cross[3] = {(size_t)d[1] * d[2], (size_t)d[0] * d[2], (size_t)d[0] * d[1]};
    size_t cpml = 0;
    for (int a = 0; a < 3; a++) {
        int pm = kind == 0 && a < 2 ? 0 : m->pml[2 * a], pp = kind == 0 && a < 2 ? 0 : m->pml[2 * a + 1];
        if (pm || pp) cpml += 4 * (size_t)(pm + pp) * cross[a] * cs;
    }
    b += cpml;
    // a 2D custom run stepped by the time-skewed batch passes keeps a second
    // copy of its three live components and their CPML psi
    if (kind == 2 && d[2] == 1 && !any && p != AEFDTD_FP16) b += 3 * N * ss + cpml / 2;
    b += 6 * (size_t)(d[0] + d[1] + d[2]) * cs;              // CPML profiles
    size_t Ptot = 0;
    for (int i = 0; i < ns; i++) {
        const aefdtd_sensor_req &q = sens[i];
        const size_t sx = std::max(0, q.x2 - q.x1), sy = std::max(0, q.y2 - q.y1), sz = std::max(0, q.z2 - q.z1);
        size_t P[6] = {0, 0, 0, 0, 0, 0};
        int nf = 0;
        if (q.kind == AEFDTD_SENSOR_FIELDBLOCK || q.kind == AEFDTD_SENSOR_FIELDMAP) {
            // FP64 complex accumulators (6 or 3 components) and a material index per cell
            const size_t cells = q.kind == AEFDTD_SENSOR_FIELDBLOCK ? sx * sy * sz
                                 : q.orientation / 2 == 0 ? sy * sz
                                 : q.orientation / 2 == 1 ? sx * sz : sx * sy;
            b += cells * ((q.kind == AEFDTD_SENSOR_FIELDBLOCK ? 6 : 3) * 16 + 4);
            continue;
        }
        if (q.box && q.kind == AEFDTD_SENSOR_POYNTING) {
            P[0] = P[1] = sy * sz; P[2] = P[3] = sx * sz; P[4] = P[5] = sx * sy;
            nf = 6;
        } else {
            P[0] = (q.orientation <= 1) ? sy * sz : (q.orientation <= 3 ? sx * sz : sx * sy);
            nf = 1;
        }
        for (int f = 0; f < nf; f++) {
            if (nf == 6 && q.disable[f]) continue;
            // accumulators, the FP64 coefficient table and its copy for the
            // matrix-core flush (in the arithmetic type)
            b += 6 * (size_t)q.Nl * P[f] * 16 + (size_t)std::max(m->Nt, 1) * 2 * q.Nl * 16;
            b += (size_t)std::max(m->Nt, 1) * 2 * ((2 * q.Nl + 31) / 32 * 32) * (p == AEFDTD_FP64 ? 8 : 4);
            Ptot += P[f];
        }
    }
    b += (size_t)std::max(kRound, stage_depth(p)) * 6 * Ptot * cs;   // DFT stage
    // sensor strips of planes normal to x on 3D grids (two columns x 6 values per (y, z))
    if (d[2] > 1 && d[0] > 1) {
        int nx_planes = 0;
        for (int i = 0; i < ns; i++) {
            const aefdtd_sensor_req &q = sens[i];
            if (q.kind == AEFDTD_SENSOR_FIELDBLOCK || q.kind == AEFDTD_SENSOR_FIELDMAP) continue;
            nx_planes += q.box && q.kind == AEFDTD_SENSOR_POYNTING ? 2 : (q.orientation <= 1 ? 1 : 0);
        }
        b += (size_t)std::min(nx_planes, 4) * 12 * d[1] * d[2] * cs;
    }
    if (kind == 2) b += (size_t)std::max(0, m->n_sources) * std::max(m->Nt, 1) * 8;
    if (kind == 0) b += 2 * (size_t)std::max(m->Nt, 1) * 64 * 6 * 8 + 2 * (size_t)m->Nt * 8;   // R/T plane sums
    if (kind == 1) b += (size_t)d[2] * (6 * ss + 1 + 8 * cs) + (size_t)d[2] * 4 * 4;           // auxiliary column
    *bytes = b + b / 50 + (8u << 20);
    return AEFDTD_OK;
}

int aefdtd_server_create(aefdtd_server **out, size_t budget) {
    if (!out || budget == 0) return AEFDTD_EINVAL;
    auto *S = new aefdtd_server();
    S->budget = budget;
    S->worker = std::thread(worker_loop, S);
    for (int t = 0; t < kBuilders; t++) S->builders.emplace_back(builder_loop, S);
    *out = S;
    return AEFDTD_OK;
}

int aefdtd_server_destroy(aefdtd_server *S) {
    if (!S) return AEFDTD_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto &kv : S->all) kv.second->cancel = true;
        S->stop = true;
    }
    S->cv.notify_all();
    for (std::thread &t : S->builders) t.join();
    S->worker.join();
    // a run built after the worker left is released here
    for (Request *q : S->built) drop(q);
    S->built.clear();
    for (aefdtd_batch *b : S->batch) aefdtd_batch_destroy(b);
    for (auto &pair : S->bev)
        for (hipEvent_t e : pair)
            if (e) (void)hipEventDestroy(e);
    delete S;
    return AEFDTD_OK;
}

int aefdtd_server_submit(aefdtd_server *S, int kind, const aefdtd_mode *m, int ns, const aefdtd_sensor_req *sens,
                         long *id) {
    if (!S || !m || !id || kind < 0 || kind > 2 || !m->matsgrid || (ns > 0 && !sens)) return AEFDTD_EINVAL;
    if (m->precision < AEFDTD_FP64 || m->precision > AEFDTD_FP16) return AEFDTD_EINVAL;
    auto q = std::make_unique<Request>();
    q->kind = kind;
    copy_inputs(*q, m, ns, sens);
    int rc = aefdtd_estimate_bytes(kind, &q->m, ns, q->sens.data(), &q->est);
    if (rc) return rc;
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return AEFDTD_ESTATE;
        q->id = S->next_id++;
        *id = q->id;
        S->queue.push_back(q.get());
        S->all[q->id] = std::move(q);
    }
    S->cv.notify_all();
    return AEFDTD_OK;
}

int aefdtd_server_cancel(aefdtd_server *S, long id) {
    if (!S) return AEFDTD_EINVAL;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return AEFDTD_EINVAL;
        it->second->cancel = true;
    }
    S->cv.notify_all();
    return AEFDTD_OK;
}

int aefdtd_server_wait(aefdtd_server *S, long id, double timeout_s, int *state) {
    if (!S) return AEFDTD_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return AEFDTD_EINVAL;
    Request *q = it->second.get();
    auto ended = [&] { return q->state >= AEFDTD_REQ_DONE; };
    if (timeout_s < 0) S->cv.wait(lk, ended);
    else S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
    if (state) *state = q->state;
    return AEFDTD_OK;
}

int aefdtd_server_result(aefdtd_server *S, long id, double *out, size_t n, double times[4]) {
    if (!S) return AEFDTD_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return AEFDTD_EINVAL;
    Request *q = it->second.get();
    if (q->state < AEFDTD_REQ_DONE) return AEFDTD_ESTATE;
    if (q->state != AEFDTD_REQ_DONE) return q->rc ? q->rc : AEFDTD_ESTATE;
    if (out) std::memcpy(out, q->out.data(), std::min(n, q->out.size()) * sizeof(double));
    if (times) {
        times[0] = secs(q->t_submit, q->t_admit);
        times[1] = secs(q->t_admit, q->t_done);
        times[2] = secs(q->t_submit, q->t_done);
        times[3] = (double)q->bytes;
    }
    return AEFDTD_OK;
}

const char *aefdtd_server_error(aefdtd_server *S, long id) {
    if (!S) return "";
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return "unknown request";
    return it->second->err.c_str();
}

int aefdtd_server_release(aefdtd_server *S, long id) {
    if (!S) return AEFDTD_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return AEFDTD_EINVAL;
    if (it->second->state < AEFDTD_REQ_DONE) return AEFDTD_ESTATE;
    S->all.erase(it);
    return AEFDTD_OK;
}

int aefdtd_server_stats(aefdtd_server *S, double out[8]) {
    if (!S || !out) return AEFDTD_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = (double)S->queue.size();
    out[1] = (double)S->resident.size();
    out[2] = (double)S->n_done;
    out[3] = (double)S->n_failed;
    out[4] = (double)S->used;
    out[5] = (double)S->peak;
    out[6] = (double)S->budget;
    out[7] = (double)S->rounds;
    return AEFDTD_OK;
}

}  // extern "C"
