// C ABI: solver entry points (declared in include/giftbte.h).
#include <cstdio>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "solver.hpp"

struct gbte_handle {
    std::unique_ptr<gbte::Solver> s;
    std::string err;
};

namespace {

int finish(gbte_handle *h, int rc) {
    if (rc != GBTE_OK && h->s) h->err = h->s->error();
    return rc;
}

}  // namespace

extern "C" {

int gbte_create(const gbte_problem *p, const gbte_options *opts, gbte_handle **out) {
    if (!p || !out) return GBTE_BAD_INPUT;
    *out = nullptr;
    gbte_options o{};
    if (opts) o = *opts;
    if (o.precision != GBTE_FP64 && o.precision != GBTE_FP32_SWEEP) {
        auto *h = new gbte_handle();
        h->err = "unknown precision option";
        *out = h;
        return GBTE_BAD_INPUT;
    }
    if (o.precision == GBTE_FP32_SWEEP && p->p.deck.order == "2limit") {
        // the limiter factor jumps (from 1 to 3/4 at its threshold), so the
        // float rounding of e moves it: the fields differ from FP64 by ~1e-4
        auto *h = new gbte_handle();
        h->err = "FP32 storage is not supported with the 2limit limiter, whose switch float rounding moves";
        *out = h;
        return GBTE_UNSUPPORTED;
    }
    if (o.precision == GBTE_FP32_SWEEP && o.fallback_kernels) {
        auto *h = new gbte_handle();
        h->err = "FP32 storage needs the tiled face kernel and the wave-specialised sweep, which fallback_kernels "
                 "turns off";
        *out = h;
        return GBTE_BAD_INPUT;
    }
    auto *h = new gbte_handle();
    h->s.reset(new gbte::Solver(p->p, o));
    *out = h;
    if (!h->s->ok()) {
        h->err = h->s->error();
        return h->err.find("no HIP device") != std::string::npos || h->err.find("HIP error") != std::string::npos
                   ? GBTE_HIP_ERROR
                   : GBTE_UNSUPPORTED;
    }
    return GBTE_OK;
}

int gbte_fourier_solve(gbte_handle *h, gbte_stats *stats) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    return finish(h, h->s->fourier(stats));
}

int gbte_set_initial(gbte_handle *h, const double *T, const double *qx, const double *qy, const double *qz) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    return finish(h, h->s->set_initial(T, qx, qy, qz));
}

int gbte_solve(gbte_handle *h, int max_iter, double tol_temp, double tol_flux, gbte_stats *stats) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    if (max_iter < 0) {
        h->err = "max_iter must be non-negative";
        return GBTE_BAD_INPUT;
    }
    return finish(h, h->s->solve(max_iter, tol_temp, tol_flux, stats));
}

int gbte_solve_batch(gbte_handle *const *hs, int n, const int *max_iter, const double *tol_temp,
                     const double *tol_flux, gbte_stats *stats, int *status) {
    if (!hs || n <= 0 || !max_iter || !tol_temp || !tol_flux) return GBTE_BAD_INPUT;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (hs[i] == hs[j]) return GBTE_BAD_INPUT;  // one solve per handle at a time
    std::vector<int> rc(n, GBTE_OK);
    // every handle is checked before any is changed; status[] is always filled
    int bad = GBTE_OK;
    for (int i = 0; i < n; ++i) {
        rc[i] = (!hs[i] || !hs[i]->s || !hs[i]->s->ok()) ? GBTE_HIP_ERROR : GBTE_OK;
        if (rc[i] != GBTE_OK && hs[i]) hs[i]->err = "solve_batch: handle was not created successfully";
        if (bad == GBTE_OK) bad = rc[i];
    }
    if (bad != GBTE_OK) {
        if (status)
            for (int i = 0; i < n; ++i) status[i] = rc[i];
        return bad;
    }
    // graphs are recorded here, one handle after another: a capture must not
    // overlap another thread's allocations and synchronous copies
    for (int i = 0; i < n; ++i) {
        hs[i]->s->set_single_stream(true);
        int r = hs[i]->s->prepare();
        if (r != GBTE_OK) {
            for (int j = 0; j <= i; ++j) hs[j]->s->set_single_stream(false);
            hs[i]->err = hs[i]->s->error();
            if (status)
                for (int j = 0; j < n; ++j) status[j] = j == i ? r : GBTE_OK;
            return r;
        }
    }
    std::vector<std::thread> th;
    th.reserve(n);
    for (int i = 0; i < n; ++i)
        th.emplace_back([&, i]() {
            try {
                rc[i] = gbte_solve(hs[i], max_iter[i], tol_temp[i], tol_flux[i], stats ? &stats[i] : nullptr);
            } catch (const std::exception &e) {  // a thread must not end the process
                hs[i]->err = std::string("solve failed: ") + e.what();
                rc[i] = GBTE_HIP_ERROR;
            }
        });
    for (auto &t : th) t.join();
    for (int i = 0; i < n; ++i) hs[i]->s->set_single_stream(false);
    int first = GBTE_OK;
    for (int i = 0; i < n; ++i) {
        if (status) status[i] = rc[i];
        if (first == GBTE_OK && rc[i] != GBTE_OK) first = rc[i];
    }
    return first;
}

int gbte_get_fields(gbte_handle *h, double *T_lattice, double *T_energy, double *qx, double *qy, double *qz,
                    double *T_band) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    return finish(h, h->s->get_fields(T_lattice, T_energy, qx, qy, qz, T_band));
}

int gbte_write_outputs(gbte_handle *h, const gbte_problem *p, const char *dir) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    if (!p) return GBTE_BAD_INPUT;
    if (p->p.n_cells != h->s->n_cells() || p->p.n_bands != h->s->n_bands()) {
        // the field buffers below are sized from p and filled from the handle
        h->err = "write_outputs: the problem is not the one the handle was created from (" +
                 std::to_string(p->p.n_cells) + " cells x " + std::to_string(p->p.n_bands) + " bands, handle " +
                 std::to_string(h->s->n_cells()) + " x " + std::to_string(h->s->n_bands()) + ")";
        return GBTE_BAD_INPUT;
    }
    size_t nc = p->p.n_cells, nB = p->p.n_bands;
    std::vector<double> T(nc), E(nc), qx(nc), qy(nc), qz(nc), Tb(nB * nc);
    int rc = finish(h, h->s->get_fields(T.data(), E.data(), qx.data(), qy.data(), qz.data(), Tb.data()));
    if (rc != GBTE_OK) return rc;
    rc = gbte_write_fields(p, T.data(), E.data(), qx.data(), qy.data(), qz.data(), Tb.data(), dir);
    if (rc != GBTE_OK) h->err = "cannot write the output files";
    return rc;
}

int gbte_keff(gbte_handle *h, int axis, double delta_T, double *keff) {
    if (!h || !h->s || !keff) return GBTE_BAD_INPUT;
    return finish(h, h->s->keff(axis, delta_T, keff));
}

int gbte_adjoint(gbte_handle *h, int axis, double delta_T, int max_iter, double tol, double *dk_dtau_scale,
                 double *dk_dbound_temp, gbte_adjoint_stats *stats) {
    if (!h || !h->s || !h->s->ok()) return GBTE_HIP_ERROR;
    if (!dk_dtau_scale) {
        h->err = "gbte_adjoint: dk_dtau_scale is required";
        return GBTE_BAD_INPUT;
    }
    return finish(h, h->s->adjoint(axis, delta_T, max_iter, tol, dk_dtau_scale, dk_dbound_temp, stats));
}

const char *gbte_last_error(const gbte_handle *h) { return h ? h->err.c_str() : "null handle"; }
const char *gbte_tuning_log(const gbte_handle *h) { return h && h->s ? h->s->tuning_log().c_str() : ""; }

void gbte_destroy(gbte_handle *h) { delete h; }

}  // extern "C"
