// params.cpp -- defaults, the mpDPFT.input reader and the mpDPFT_V.dat reader (host only).
//
// The reader keeps the reference's positional contract (GetInputParameters,
// mpDPFT.cpp:621-721): line n holds parameter n, each value followed by a
// "//" comment; S-vectors read exactly S values; open-ended vectors read
// numbers until the first token that is not a number; Mixer reads 3 values.
// Lines after the 47th are extension keys written as "value //Key ...".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mpdpft.h"

extern "C" int mpdpft_abi_version(void) { return MPDPFT_ABI_VERSION; }
extern "C" size_t mpdpft_params_sizeof(void) { return sizeof(mpdpft_params); }
extern "C" size_t mpdpft_result_sizeof(void) { return sizeof(mpdpft_result); }

// Structural checks every consumer of a params record runs first: the
// layout version and every list count against its fixed array (a larger count
// would index past the struct).  Returns nullptr or the reason.
const char *params_structure_error(const mpdpft_params *p) {
    if (p->abi_version != MPDPFT_ABI_VERSION) return "abi_version does not match this library (mpdpft_params_default)";
    if (p->S < 1 || p->S > MPDPFT_MAX_SPECIES) return "S out of range";
    if (p->n_interactions < 0 || p->n_interactions > MPDPFT_MAX_LIST) return "n_interactions out of range";
    if (p->n_schedule < 0 || p->n_schedule > MPDPFT_MAX_LIST) return "n_schedule out of range";
    if (p->n_regularize < 0 || p->n_regularize > MPDPFT_MAX_LIST) return "n_regularize out of range";
    if (p->n_ekin < 0 || p->n_ekin > MPDPFT_MAX_LIST) return "n_ekin out of range";
    if (p->n_Mpp < 0 || p->n_Mpp > MPDPFT_MAX_LIST) return "n_Mpp out of range";
    if (p->n_mpp < 0 || p->n_mpp > MPDPFT_MAX_MPP) return "n_mpp out of range";
    if (p->n_centres < 0 || p->n_centres > MPDPFT_MAX_CENTRES) return "n_centres out of range";
    if (p->n_hyperbox < 0 || p->n_hyperbox > 2 * MPDPFT_MAX_SPECIES) return "n_hyperbox out of range";
    return nullptr;
}

extern "C" void mpdpft_params_default(mpdpft_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->abi_version = MPDPFT_ABI_VERSION;
    p->DIM = 2;
    p->Units = 1;
    p->edge = 16.;
    p->steps = 64;
    p->S = 1;
    p->Abundances[0] = 1.;
    p->RelAcc = 1e-10;
    p->InternalAcc = 1e-6;
    p->DensityExpression = 7;
    for (int s = 0; s < MPDPFT_MAX_SPECIES; s++) {
        p->tauVec[s] = 1.;
        p->tVec[s] = 1.;
        p->TVec[s] = 1.;
        p->Environments[s] = 1;
        p->muVec[s] = 1.;
        p->thetaVec[s] = 0.1;
    }
    p->n_mpp = 1;
    p->mpp[0] = 0.5;
    p->stretchfactor = 1.;
    p->n_interactions = 1;
    p->Interactions[0] = 0;
    p->incrementalV = 1.;
    p->DeltamuModifier = -3.;
    p->Mixer[0] = 0.;
    p->Mixer[1] = 8.;
    p->Mixer[2] = 0.;
    p->SCcriterion = 1e-10;
    p->maxSCcount = 100;
    p->method = 1;
    p->DelFieldMethod = 5;
    p->n_schedule = 1;
    p->n_ekin = 1;
    p->degeneracy = 1.;
    p->n_regularize = 1;
    p->RegularizationThreshold = 0.375;
    p->Print = 0;
    p->TaperOff = 1;
    p->RegularizeCoulombSingularity = 1;
    p->SmearingExtentDx = 6.;
    p->OutPrecision = 16;
    p->MaxIteration = 4;
    p->HartreeMode = 0;
    p->SmoothingKWidth = 0.;
    p->ConvergenceMode = 0;
    p->Precision = 0;
    p->WriteFiles = 1;
    p->n_hyperbox = 2;  // the shipped input's "0. 0."
}

namespace {

// Numbers at the start of the line until the first non-number token.
template <class T> std::vector<T> leading_numbers(const std::string &line) {
    std::istringstream iss(line);
    std::vector<T> out;
    T v;
    while (iss >> v) out.push_back(v);
    return out;
}

template <class T> int fill_fixed(const std::string &line, int n, T *dst) {
    std::vector<T> v = leading_numbers<T>(line);
    if ((int)v.size() < n) return -1;
    for (int i = 0; i < n; i++) dst[i] = v[i];
    return 0;
}

template <class T> int fill_open(const std::string &line, int maxn, T *dst, int *count) {
    std::vector<T> v = leading_numbers<T>(line);
    if ((int)v.size() > maxn) return -1;
    for (size_t i = 0; i < v.size(); i++) dst[i] = v[i];
    *count = (int)v.size();
    return 0;
}

template <class T> int first(const std::string &line, T *dst) {
    std::istringstream iss(line);
    T v;
    if (!(iss >> v)) return -1;
    *dst = v;
    return 0;
}

// "value //Key ..." -> key, value text
bool extension_line(const std::string &line, std::string *key, std::string *val) {
    size_t c = line.find("//");
    if (c == std::string::npos) return false;
    std::istringstream ks(line.substr(c + 2));
    if (!(ks >> *key)) return false;
    *val = line.substr(0, c);
    return true;
}

int parse_extension(const std::string &key, const std::string &val, mpdpft_params *p) {
    if (key == "HartreeMode") return first(val, &p->HartreeMode);
    if (key == "SmoothingKWidth") return first(val, &p->SmoothingKWidth);
    if (key == "ConvergenceMode") return first(val, &p->ConvergenceMode);
    if (key == "Precision") return first(val, &p->Precision);
    if (key == "TaperOff") return first(val, &p->TaperOff);
    if (key == "RegularizeCoulombSingularity") return first(val, &p->RegularizeCoulombSingularity);
    if (key == "SmearingExtent") return first(val, &p->SmearingExtentDx);
    if (key == "OutPrecision") return first(val, &p->OutPrecision);
    if (key == "MaxIteration") return first(val, &p->MaxIteration);
    if (key == "WriteFiles") return first(val, &p->WriteFiles);
    if (key == "CubeFile") return first(val, &p->CubeFile);
    if (key == "NoiseSeed") return first(val, &p->NoiseSeed);
    if (key == "Centre") {  // x y z Z a
        if (p->n_centres >= MPDPFT_MAX_CENTRES) return -1;
        if (fill_fixed(val, 5, p->centres[p->n_centres])) return -1;
        p->n_centres++;
        return 0;
    }
    return -1;
}

}  // namespace

extern "C" int mpdpft_parse_input(const char *path, mpdpft_params *p) {
    std::ifstream in(path);
    if (!in) return MPDPFT_EIO;
    mpdpft_params q = *p;
    std::string line;
    int n = 0, bad = 0;
    while (std::getline(in, line)) {
        int rc = 0;
        switch (n) {
        case 0: rc = first(line, &q.System); break;
        case 1: rc = first(line, &q.Symmetry); break;
        case 2: rc = first(line, &q.TaskType); break;
        case 3: rc = first(line, &q.TaskParameterCount); break;
        case 4: rc = fill_open(line, 2 * MPDPFT_MAX_SPECIES, q.TaskHyperBox, &q.n_hyperbox); break;
        case 5: rc = first(line, &q.DIM); break;
        case 6: rc = first(line, &q.Units); break;
        case 7: rc = first(line, &q.edge); break;
        case 8: rc = first(line, &q.steps); break;
        case 9:
            rc = first(line, &q.S);
            if (!rc && (q.S < 1 || q.S > MPDPFT_MAX_SPECIES)) rc = -1;
            break;
        case 10: rc = fill_fixed(line, q.S, q.Abundances); break;
        case 11: rc = first(line, &q.RelAcc); break;
        case 12: rc = first(line, &q.InternalAcc); break;
        case 13: rc = first(line, &q.K); break;
        case 14: break;  // Resources: ecology, not on this path
        case 15: rc = first(line, &q.DensityExpression); break;
        case 16: rc = fill_open(line, MPDPFT_MAX_MPP, q.mpp, &q.n_mpp); break;
        case 17: rc = fill_open(line, MPDPFT_MAX_LIST, q.Mpp, &q.n_Mpp); break;
        case 18: rc = fill_fixed(line, q.S, q.tauVec); break;
        case 19: rc = fill_fixed(line, q.S, q.tVec); break;
        case 20: rc = fill_fixed(line, q.S, q.TVec); break;
        case 21: rc = fill_fixed(line, q.S, q.Environments); break;
        case 22: rc = first(line, &q.stretchfactor); break;
        case 23: rc = first(line, &q.Wall); break;
        case 24: rc = first(line, &q.Noise); break;
        case 25: rc = first(line, &q.InterpolVQ); break;
        case 26: rc = fill_open(line, MPDPFT_MAX_LIST, q.Interactions, &q.n_interactions); break;
        case 27: rc = first(line, &q.incrementalV); break;
        case 28: rc = fill_fixed(line, q.S, q.muVec); break;
        case 29: rc = first(line, &q.DeltamuModifier); break;
        case 30: rc = fill_fixed(line, 3, q.Mixer); break;
        case 31: rc = fill_fixed(line, q.S, q.thetaVec); break;
        case 32: rc = first(line, &q.SCcriterion); break;
        case 33: rc = first(line, &q.maxSCcount); break;
        case 34: rc = first(line, &q.method); break;
        case 35: rc = first(line, &q.DelFieldMethod); break;
        case 36: rc = first(line, &q.MovieQ); break;
        case 37: rc = fill_open(line, MPDPFT_MAX_LIST, q.Schedule, &q.n_schedule); break;
        case 38: rc = fill_open(line, MPDPFT_MAX_LIST, q.EkinTypes, &q.n_ekin); break;
        case 39: rc = first(line, &q.alpha); break;
        case 40: rc = first(line, &q.beta); break;
        case 41: rc = first(line, &q.gamma); break;
        case 42: rc = first(line, &q.gammaH); break;
        case 43: rc = first(line, &q.degeneracy); break;
        case 44: rc = fill_open(line, MPDPFT_MAX_LIST, q.regularize, &q.n_regularize); break;
        case 45: rc = first(line, &q.RegularizationThreshold); break;
        case 46: rc = first(line, &q.Print); break;
        default: {
            std::string key, val;
            if (line.find_first_not_of(" \t\r") == std::string::npos) break;
            if (!extension_line(line, &key, &val) || parse_extension(key, val, &q)) rc = -1;
        }
        }
        if (rc) bad++;
        n++;
    }
    if (n < 47 || bad) return MPDPFT_EINVAL;
    *p = q;
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// LoadV/ImportV (InterpolVQ = 1, mpDPFT.cpp:6430-6502): mpDPFT_V.dat read and
// interpolated multilinearly onto this run's lattice (host only).
// ---------------------------------------------------------------------------
// Cell index and fraction of t on the sorted axis X (n >= 2): the cell
// [X[l], X[l+1]] that holds t, the first or last cell outside the axis range
// (so the multilinear form extrapolates linearly there).
static void cell(const std::vector<double> &X, double t, int &l, double &f) {
    int n = (int)X.size(), lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int m = (lo + hi) / 2;
        if (X[m] >= t) hi = m;
        else lo = m;
    }
    l = lo;
    f = (t - X[l]) / (X[l + 1] - X[l]);
}

extern "C" int mpdpft_import_v(const char *path, const mpdpft_params *p, double *out, int *warnings) {
    if (!path || !p || !out) return MPDPFT_EINVAL;
    if (params_structure_error(p)) return MPDPFT_EINVAL;
    if (p->DIM < 1 || p->DIM > 3 || p->S < 1 || p->S > MPDPFT_MAX_SPECIES || p->steps < 1) return MPDPFT_EINVAL;
    const int D = p->DIM, S = p->S, EL = p->steps + 1;
    FILE *f = std::fopen(path, "r");
    if (!f) return MPDPFT_EIO;
    std::vector<std::vector<double>> rows;
    std::vector<char> buf(1 << 16);
    std::string line;
    while (std::fgets(buf.data(), (int)buf.size(), f)) {
        line += buf.data();
        if (line.empty() || line.back() != '\n') {
            if (!std::feof(f)) continue;
        }
        std::istringstream is(line);
        std::vector<double> r;
        double v;
        while (is >> v) r.push_back(v);
        if (!r.empty()) rows.push_back(std::move(r));
        line.clear();
    }
    std::fclose(f);
    int warn = 0;
    const int N = (int)rows.size();
    if (N < 2) return MPDPFT_EIO;
    const int n = (int)(std::pow((double)N, 1. / (double)D) + 0.5);
    size_t nD = 1;
    for (int d = 0; d < D; d++) nD *= (size_t)n;
    if (n < 2 || nD != (size_t)N) return MPDPFT_EIO;
    if ((int)rows[0].size() != D + S) warn++;  // "LoadV: Error!!! dimensional mismatch"
    for (const auto &r : rows)
        if ((int)r.size() < D + S) return MPDPFT_EIO;
    // axes from the first columns: x slowest (row i n^(D-1)), ..., last axis fastest
    std::vector<std::vector<double>> X(D, std::vector<double>(n));
    for (int d = 0; d < D; d++) {
        size_t stride = 1;
        for (int e = d + 1; e < D; e++) stride *= (size_t)n;
        for (int i = 0; i < n; i++) X[d][i] = rows[(size_t)i * stride][d];
    }
    // target lattice (as mpdpft_create builds it)
    std::vector<double> lat(EL);
    const double x1 = -p->edge / 2., x2 = p->edge / 2., del = (x2 - x1) / (double)p->steps;
    for (int i = 0; i < EL; i++) lat[i] = x1 + del * (double)i;
    std::vector<std::vector<int>> cl(D, std::vector<int>(EL));
    std::vector<std::vector<double>> cf(D, std::vector<double>(EL));
    for (int d = 0; d < D; d++)
        for (int i = 0; i < EL; i++) cell(X[d], lat[i], cl[d][i], cf[d][i]);
    size_t G = 1;
    for (int d = 0; d < D; d++) G *= (size_t)EL;
    for (int s = 0; s < S; s++) {
        const int col = D + s;
        bool nan = false;
        for (const auto &r : rows) nan |= r[col] != r[col];
        if (nan) warn++;  // "LoadV: Error!!! ImportField ... is NAN"
        for (size_t g = 0; g < G; g++) {
            int idx[3] = {0, 0, 0};
            size_t rem = g;
            for (int d = D - 1; d >= 0; d--) {
                idx[d] = (int)(rem % EL);
                rem /= EL;
            }
            // multilinear: sum over the 2^D cell corners
            double acc = 0.;
            for (int corner = 0; corner < (1 << D); corner++) {
                double w = 1.;
                size_t r = 0;
                for (int d = 0; d < D; d++) {
                    int b = (corner >> (D - 1 - d)) & 1;
                    double t = cf[d][idx[d]];
                    w *= b ? t : 1. - t;
                    r = r * (size_t)n + (size_t)(cl[d][idx[d]] + b);
                }
                acc += w * rows[r][col];
            }
            out[(size_t)s * G + g] = acc;
        }
    }
    // integral cross-check (the reference warns above 5% asymmetry)
    double imp = 0., cur = 0.;
    for (int s = 0; s < S; s++) {
        for (const auto &r : rows) imp += r[D + s];
        for (size_t g = 0; g < G; g++) cur += out[(size_t)s * G + g];
    }
    imp *= std::pow(2. * std::fabs(rows[0][0]), (double)D) / (double)N;
    cur *= std::pow(del, (double)D);
    if (std::fabs((imp - cur) / (imp + cur)) > 0.05) warn++;
    if (warnings) *warnings = warn;
    return MPDPFT_OK;
}
