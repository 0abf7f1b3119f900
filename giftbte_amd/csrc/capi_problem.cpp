// C ABI: problem construction (declared in include/giftbte.h).
#include <cstring>
#include <string>

#include "problem.hpp"

namespace {
thread_local std::string g_problem_error;
}

namespace gbte {
// The calling thread's last message of a call without a handle (problem
// construction, gbte_flux_keff*), read by gbte_problem_error.
std::string &thread_error() { return g_problem_error; }
}  // namespace gbte

extern "C" {

int gbte_problem_load(const char *deck_dir, gbte_problem **out) {
    if (!deck_dir || !out) {
        g_problem_error = "gbte_problem_load: null argument";
        return GBTE_BAD_INPUT;
    }
    *out = nullptr;
    auto *h = new gbte_problem();
    std::string err;
    try {
        err = gbte::load_problem(deck_dir, h->p);
    } catch (const std::exception &e) {
        err = std::string("deck parse failure: ") + e.what();
    }
    if (!err.empty()) {
        g_problem_error = err;
        delete h;
        return (err.find("not supported") != std::string::npos) ? GBTE_UNSUPPORTED : GBTE_BAD_INPUT;
    }
    g_problem_error.clear();
    *out = h;
    return GBTE_OK;
}

int gbte_problem_tables(const gbte_problem *p, gbte_tables *out) {
    if (!p || !out) return GBTE_BAD_INPUT;
    *out = p->p.tables();
    return GBTE_OK;
}

void gbte_problem_free(gbte_problem *p) { delete p; }

const char *gbte_problem_error(void) { return g_problem_error.c_str(); }

}  // extern "C"
