// runs.hpp -- host objects of the mode drivers (modes.cpp) shared with the
// request server (server.cpp).
#pragma once

#include "engine.hpp"

#include <atomic>
#include <string>
#include <vector>

// One mode driver between create and finish (modes.cpp).
struct aefdtd_run {
    int kind = 0;                     // 0 normal incidence, 1 single particle, 2 custom
    int Nt = 0, t = 0;
    int polarization = 0, Nl = 0;
    double lambda_min = 0, lambda_max = 0, Dt = 0;
    aefdtd_grid *g = nullptr, *a = nullptr;
    std::vector<aefdtd_sensor_req> reqs;
    std::vector<aefdtd_source_req> srcs;
    std::vector<std::vector<int>> ids;
    int idR = -1, idT = -1;
    double index_sup = 1, index_sub = 1;
    std::string err;
    struct aefdtd_batch *batch = nullptr;   // set while a member of a batch
    // auto_tsteps: energy (time_type 1) or spectral (2) criterion; done once it stopped
    int time_type = 0, cc_step = 1;
    double cc_coeff = 0;
    bool done = false;
    int nrun = -1;                    // normal incidence: the step count its spectra use
    // cancellation, polled after every step as the drivers poll end_computation
    // (normal_incidence.cpp:431-435, single_particle.cpp:403, default_fdtd.cpp:201):
    // this run's own flag (aefdtd_run_cancel) and the caller's flag of a one-call driver
    std::atomic<bool> cancel{false};
    const int *end_flag = nullptr;
    bool cancelled = false;
    aefdtd::FTState ft;               // the spectral criterion's watched cells
    std::vector<double> ft_ring;
    ~aefdtd_run() {
        aefdtd::ft_free(ft);
        aefdtd_destroy(g);
        aefdtd_destroy(a);
    }
};

struct aefdtd_batch {
    aefdtd::Batch *b = nullptr;
    std::vector<aefdtd_run *> runs;
    std::string err;
};

namespace aefdtd {
// Full-grid dimensions a mode driver of `kind` builds for mode m (the padding
// rules of the drivers and of FDTD::FDTD), and the auxiliary grid's cells.
void run_dims(int kind, const aefdtd_mode *m, int dims[3], int pads[6]);
// Doubles a sensor request contributes to the drivers' sensor_out.
size_t request_out_len(const aefdtd_sensor_req &q);
}  // namespace aefdtd
