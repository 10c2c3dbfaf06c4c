// sanitize_driver.cpp -- the host-side readers and the oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer (built and run by
// tests/test_sanitizers.py; no GPU).
//
//   sanitize_driver <dir>
//
// <dir> holds the inputs the test wrote: c0.input (2D, two species),
// c1.input (3D nAiT), v2.dat and v3.dat (mpDPFT_V.dat files for LoadV).
// The driver parses both inputs with mpdpft_parse_input, imports both V files
// with mpdpft_import_v (onto the parsed lattices and onto a coarser and a finer
// one), and runs the oracle on them: a short 2D solve, the 3D environment,
// DelField, the nAiT density and mu search, mixing and the Pulay small solve.
// Any sanitizer report aborts with a non-zero status; a failed call returns 2.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../include/mpdpft.h"
#include "../oracle/dpft_oracle.h"

static int fails = 0;
#define CHECK(cond, what)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAILED: %s\n", what);    \
            fails++;                                       \
        }                                                  \
    } while (0)

static size_t grid(const mpdpft_params &p) {
    size_t G = 1;
    for (int d = 0; d < p.DIM; d++) G *= (size_t)(p.steps + 1);
    return G;
}

static bool finite(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    mpdpft_params p2, p3;
    mpdpft_params_default(&p2);
    mpdpft_params_default(&p3);
    CHECK(mpdpft_parse_input((dir + "/c0.input").c_str(), &p2) == MPDPFT_OK, "parse c0.input");
    CHECK(mpdpft_parse_input((dir + "/c1.input").c_str(), &p3) == MPDPFT_OK, "parse c1.input");
    CHECK(mpdpft_parse_input((dir + "/missing.input").c_str(), &p2) == MPDPFT_EIO, "missing input is EIO");
    CHECK(p2.DIM == 2 && p2.S == 2 && p3.DIM == 3 && p3.S == 1, "parsed dimensions");

    // LoadV onto the parsed lattice and onto coarser / finer ones
    for (int k = 0; k < 3; k++) {
        mpdpft_params q = k == 0 ? p2 : p3;
        const char *vf = k == 0 ? "/v2.dat" : "/v3.dat";
        for (int steps : {q.steps / 2, q.steps, 2 * q.steps}) {
            mpdpft_params r = q;
            r.steps = steps;
            std::vector<double> v((size_t)r.S * grid(r));
            int warn = -1;
            CHECK(mpdpft_import_v((dir + vf).c_str(), &r, v.data(), &warn) == MPDPFT_OK, "import_v");
            CHECK(finite(v) && warn >= 0, "imported V finite");
        }
    }
    {
        std::vector<double> v((size_t)p2.S * grid(p2));
        CHECK(mpdpft_import_v((dir + "/c0.input").c_str(), &p2, v.data(), nullptr) != MPDPFT_OK,
              "a non-V file is rejected");
    }

    // oracle: a short 2D two-species solve
    {
        mpdpft_params q = p2;
        q.maxSCcount = 12;
        mpdpft_result r;
        std::vector<double> den((size_t)q.S * grid(q)), V(den.size());
        int rc = oracle_solve(&q, &r, den.data(), V.data());
        CHECK(rc == MPDPFT_OK || rc == MPDPFT_ENOTCONV, "oracle_solve 2D");
        CHECK(finite(den) && finite(V) && std::isfinite(r.TotalEnergy), "2D solve finite");
    }
    // oracle: 3D environment, DelField, nAiT density and mu search, mixing
    {
        const size_t G = grid(p3);
        std::vector<double> env(G), dx(G), dy(G), dz(G), g2(G), lap(G), den(G);
        CHECK(oracle_env(&p3, env.data()) == MPDPFT_OK, "oracle_env");
        CHECK(oracle_delfield(&p3, env.data(), 5, dx.data(), dy.data(), dz.data(), g2.data(), lap.data()) ==
                  MPDPFT_OK,
              "oracle_delfield");
        double N = 0.;
        CHECK(oracle_density(&p3, 0, -0.3, env.data(), den.data(), &N) == MPDPFT_OK && std::isfinite(N),
              "oracle_density nAiT");
        double mu = -0.3, Nm = 0.;
        int nob = 0;
        CHECK(oracle_enforce(&p3, env.data(), &mu, &Nm, den.data(), &nob) == MPDPFT_OK && std::isfinite(mu),
              "oracle_enforce");
        std::vector<double> outs(3 * G), mixed(G);
        for (int k = 0; k < 3; k++)
            for (size_t i = 0; i < G; i++) outs[k * G + i] = den[i] * (1. + 0.01 * k);
        int rej = 0;
        CHECK(oracle_mix_sequence(&p3, 3, den.data(), outs.data(), mixed.data(), &rej) == MPDPFT_OK &&
                  finite(mixed),
              "oracle_mix_sequence");
        double A[9] = {2., 1., 0.5, 1., 2., 1., 0.5, 1., 2.}, c[3], tv = 0.;
        CHECK(oracle_pulay_coefficients(3, A, c, &tv) == 1, "oracle_pulay_coefficients");
        CHECK(std::isfinite(oracle_nait_point(-0.5, 0.3, 0.1, 0.05, 1e-6, 10, 20, 1e-6)), "nait point");
        CHECK(std::isfinite(oracle_ekin_point(-0.5, 0.3, 0.1, 1e-6, 1e-6, 8)), "ekin point");
    }
    if (fails) {
        std::fprintf(stderr, "%d check(s) failed\n", fails);
        return 2;
    }
    std::printf("sanitize_driver: all checks passed\n");
    return 0;
}
