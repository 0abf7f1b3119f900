// Band-data front ends (src/BTEBand/BTEBand.cpp:81-740): mode data from
// ShengBTE or ALAMODE binned into gray bands.
//
// The binning is the reference's, quirks included, because the band tables
// must come out bit-identical:
//   - modes go to bands in file order (ShengBTE: branch-major over the full
//     q grid), not by frequency; band k takes modes (w k - 1, w (k + 1) - 1]
//     with w = ceil(nmode / nband) of an integer quotient, the last band up to
//     the last mode (:564-577, :690-701);
//   - the ShengBTE cell volume repeats lattvec[0][1] where the determinant has
//     lattvec[1][0] (:294), and T is read from every CONTROL line with a
//     capital T past its first column whose value is not .TRUE. (:276-292);
//   - numbers are the tokens between single spaces, each read with a prefix
//     parse (std::stod), and a data file ends at its first line of at most one
//     character (:310-350).  A line with trailing spaces makes the reference's
//     scan repeat its last token forever; here the scan stops there.
#include "bands.hpp"

#include <cmath>
#include <fstream>
#include <string>

namespace gbte {
namespace {

bool prefix_double(const std::string &s, double &v) {
    const char *p = s.c_str();
    char *end = nullptr;
    v = std::strtod(p, &end);
    return end != p;
}

// The reference's token scan of one data line (BTEBand.cpp:310-350).
bool ref_tokens(const std::string &l, std::vector<double> &out) {
    out.clear();
    size_t n = l.size(), pos1 = 0, pos2 = 0;
    bool first = true;
    while (pos2 >= pos1) {
        size_t prev1 = pos1;
        for (size_t si = pos2; si < n; ++si)
            if (l[si] != ' ') {
                pos1 = si;
                break;
            }
        if (!first && pos1 == prev1) break;  // only spaces left
        first = false;
        bool found = false;
        for (size_t si = pos1; si < n; ++si)
            if (l[si] == ' ') {
                pos2 = si;
                found = true;
                break;
            }
        double v;
        if (found || pos2 >= pos1) {
            if (!prefix_double(l.substr(pos1, pos2 - pos1), v)) return false;
        } else {
            if (!prefix_double(l.substr(pos1), v)) return false;
        }
        out.push_back(v);
        if (!found) break;
    }
    return true;
}

// Rows of a ShengBTE data file up to its first line of length <= 1.
std::string read_rows(const std::string &path, std::vector<std::vector<double>> &rows) {
    std::ifstream f(path);
    if (!f.is_open()) return "Error: " + path + " not open";
    std::string s;
    std::vector<double> r;
    while (std::getline(f, s)) {
        if (s.length() <= 1) break;
        if (!ref_tokens(s, r)) return "Error: unreadable number in " + path;
        rows.push_back(r);
    }
    return "";
}

// Binning of the modes into bands (BTEBand.cpp:561-628, :684-730).
std::string bin_modes(const std::vector<double> &C, const std::vector<double> &vel, const std::vector<double> &kap,
                      int nband, BandColumns &out) {
    int total = (int)C.size();
    if (nband <= 0 || total <= 0) return "Error: no phonon modes to bin";
    std::vector<double> bin(nband + 1);
    bin[0] = -1;
    for (int j = 1; j <= nband; ++j) {
        bin[j] = bin[0] + std::ceil(total / nband) * j;
        if (j == nband) bin[j] = total;
    }
    std::vector<double> sC(nband, 0.0), sCv(nband, 0.0), sk(nband, 0.0);
    for (int j = 0; j < total; ++j)
        for (int k = 0; k < nband; ++k)
            if (j > bin[k] && j <= bin[k + 1]) {
                sC[k] += C[j];
                sCv[k] += C[j] * vel[j];
                sk[k] += kap[j];
            }
    out.v.assign(nband, 0.0);
    out.tau.assign(nband, 0.0);
    out.C.assign(nband, 0.0);
    out.lr.assign(nband, 0.0);
    out.hr.assign(nband, 0.0);
    double sumL = 0, sumheat = 0;
    for (int j = 0; j < nband; ++j) {
        out.v[j] = sCv[j] / sC[j];
        out.C[j] = sC[j];
        out.tau[j] = sk[j] / out.C[j] / out.v[j] / out.v[j];
        sumL += out.C[j] / out.tau[j];
        sumheat += out.C[j];
    }
    for (int j = 0; j < nband; ++j) {
        out.lr[j] = out.C[j] / out.tau[j] / sumL;
        out.hr[j] = out.C[j] / sumheat;
    }
    return "";
}

// The value after '=' of a CONTROL line, from its first non-space character.
std::string after_eq(const std::string &line) {
    size_t e = line.find('=');
    std::string v = line.substr(e == std::string::npos ? 0 : e + 1);
    size_t p = v.find_first_not_of(' ');
    return p == std::string::npos ? std::string() : v.substr(p);
}

bool has_at(const std::string &s, const char *key) {
    size_t p = s.find(key);
    return p != std::string::npos && p > 0;
}

}  // namespace

std::string bin_shengbte(const std::string &dir, int nband, BandColumns &out) {
    std::ifstream control(dir + "/CONTROL");
    if (!control.is_open()) return "Error: ShengBTE CONTROL file not open";
    double lfactor = 1.0, T = 0.0, lat[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    std::string s;
    std::vector<double> tok;
    const char *lv[3] = {"lattvec(:,1)", "lattvec(:,2)", "lattvec(:,3)"};
    while (std::getline(control, s)) {
        if (has_at(s, "lfactor")) {
            double v;
            if (!prefix_double(after_eq(s), v)) return "Error: ShengBTE lfactor unreadable";
            lfactor = v;
        }
        for (int k = 0; k < 3; ++k)
            if (has_at(s, lv[k])) {
                if (!ref_tokens(after_eq(s), tok) || tok.size() < 3) return "Error: ShengBTE lattvec unreadable";
                for (int c = 0; c < 3; ++c) lat[k][c] = tok[c];
            }
        if (has_at(s, "T")) {
            std::string v = after_eq(s);
            double x;
            if (v.find("TRUE") == std::string::npos && v.find("True") == std::string::npos) {
                if (!prefix_double(v, x)) return "Error: ShengBTE temperature unreadable";
                T = x;
            }
        }
    }
    double volume = (lat[0][0] * lat[1][1] * lat[2][2] + lat[1][0] * lat[2][1] * lat[0][2] +
                     lat[0][1] * lat[1][2] * lat[2][0] - lat[0][2] * lat[1][1] * lat[2][0] -
                     lat[0][1] * lat[0][1] * lat[2][2] - lat[0][0] * lat[1][2] * lat[2][1]) *
                    lfactor * lfactor * lfactor;
    std::vector<std::vector<double>> omega, qpt, wfin, vel;
    std::string err = read_rows(dir + "/BTE.omega", omega);
    if (err.empty()) err = read_rows(dir + "/BTE.qpoints_full", qpt);
    if (err.empty()) err = read_rows(dir + "/BTE.w_final", wfin);
    if (err.empty()) err = read_rows(dir + "/BTE.v", vel);
    if (!err.empty()) return err;
    int nirr = (int)omega.size(), nq = (int)qpt.size();
    if (nirr == 0 || nq == 0) return "Error: empty ShengBTE data";
    int nbr = (int)omega.back().size();  // branches: the token count of the last line read
    size_t nw = (size_t)nbr * nirr;
    if (wfin.size() < nw || vel.size() < nw) return "Error: ShengBTE BTE.w_final / BTE.v shorter than BTE.omega";
    std::vector<std::vector<double>> om(nq), ta(nq), ve(nq), ca(nq);
    for (int i1 = 0; i1 < nq; ++i1) {
        if (qpt[i1].size() < 2) return "Error: ShengBTE BTE.qpoints_full row too short";
        int q = (int)qpt[i1][1] - 1;
        if (q < 0 || q >= nirr) return "Error: ShengBTE q-point index out of range";
        for (int i2 = 0; i2 < nbr; ++i2) {
            size_t r = (size_t)i2 * nirr + q;
            if ((int)omega[q].size() <= i2 || wfin[r].size() < 2 || vel[r].size() < 3)
                return "Error: ShengBTE data row too short";
            om[i1].push_back(omega[q][i2] * 1E12);
            ta[i1].push_back(1 / wfin[r][1] * 1E-12);
            ve[i1].push_back(std::pow(std::pow(vel[r][0] * 1000, 2) + std::pow(vel[r][1] * 1000, 2) +
                                          std::pow(vel[r][2] * 1000, 2),
                                      0.5));
        }
    }
    const double kb = 1.38062E-23, hbar = 1.05459E-34;
    volume = std::fabs(volume) * 1E-27;
    for (int i1 = 0; i1 < nq; ++i1)
        for (int i2 = 0; i2 < nbr; ++i2) {
            double con = hbar * om[i1][i2] / kb / T;
            ca[i1].push_back(1 / volume * kb * con * con * std::exp(con) / std::pow(std::exp(con) - 1, 2));
        }
    std::vector<double> C, V, K;
    for (int i2 = 0; i2 < nbr; ++i2)
        for (int i1 = 0; i1 < nq; ++i1) {
            double value = ca[i1][i2] / nq * ve[i1][i2] * ve[i1][i2] * ta[i1][i2];
            if (std::isnan(value)) continue;
            if (value >= 0) {
                C.push_back(ca[i1][i2] / nq);
                V.push_back(ve[i1][i2]);
                K.push_back(value);
            }
        }
    return bin_modes(C, V, K, nband, out);
}

std::string bin_alamode(const std::string &file, size_t key_len, int nband, BandColumns &out) {
    std::ifstream f(file);
    if (!f.is_open()) return "Error: ALAMODE band file not open";
    std::string s;
    bool got = false;
    while (std::getline(f, s)) {
        size_t p = s.find("Phonon lifetime");
        if (p == std::string::npos || p == 0 || p >= key_len) continue;
        std::string a1, a2;
        int q0 = 0, q1 = -1, b0 = 0, b1 = -1;
        f >> a1 >> a2 >> a1 >> q0 >> q1;
        f >> a1 >> a2 >> a1 >> b0 >> b1;
        std::getline(f, s);
        std::getline(f, s);
        long total = (long)(b1 - b0 + 1) * (q1 - q0 + 1);
        if (!f || total <= 0) return "Error: ALAMODE lifetime header unreadable";
        std::vector<double> freq(total), tau(total), v(total), mfp(total), mult(total), kx(total), ky(total),
            kz(total), kap(total), C(total);
        for (long j = 0; j < total; ++j) {
            int iq, ib;
            f >> iq >> ib >> freq[j] >> tau[j] >> v[j] >> mfp[j] >> mult[j] >> kx[j] >> ky[j] >> ky[j] >> ky[j] >>
                ky[j] >> kz[j] >> kz[j] >> kz[j] >> kz[j];
        }
        if (!f) return "Error: ALAMODE lifetime records unreadable";
        for (long j = 0; j < total; ++j) {
            tau[j] = tau[j] * 1e-12;
            kap[j] = (kx[j] + ky[j] + kz[j]) * mult[j];
            C[j] = kap[j] / v[j] / v[j] / tau[j];
            if (tau[j] == 0 || v[j] == 0) C[j] = 0;
        }
        std::string err = bin_modes(C, v, kap, nband, out);
        if (!err.empty()) return err;
        got = true;
    }
    return got ? "" : "Error: no \"Phonon lifetime\" block in the ALAMODE file";
}

}  // namespace gbte
