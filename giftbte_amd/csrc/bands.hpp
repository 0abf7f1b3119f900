// Band-data front ends: the spectral binning of ShengBTE and ALAMODE mode
// data into gray bands (src/BTEBand/BTEBand.cpp:81-740).
#pragma once

#include <string>
#include <vector>

namespace gbte {

// One material's bands: group velocity, relaxation time, heat capacity,
// lattice ratio and heat ratio, nband entries each.
struct BandColumns {
    std::vector<double> v, tau, C, lr, hr;
};

// ShengBTE output directory `dir` (CONTROL, BTE.omega, BTE.qpoints_full,
// BTE.w_final, BTE.v), BTEBand.cpp:81-628.
std::string bin_shengbte(const std::string &dir, int nband, BandColumns &out);

// ALAMODE lifetime file `file`; `key_len` is the length of the PHONON line
// that names it, which bounds where the reference accepts its "Phonon
// lifetime" header (BTEBand.cpp:629-740).
std::string bin_alamode(const std::string &file, size_t key_len, int nband, BandColumns &out);

}  // namespace gbte
