// output.cpp -- the drop-in directory contract: read mpDPFT.input, solve,
// write mpDPFT_Den.dat, mpDPFT_V.dat and mpDPFT_Control.dat in the reference
// formats (ExportData, mpDPFT.cpp:12399-12520; final block :12716-12800;
// WriteControlFile :11989-11995).
#include <cstdio>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "solver.h"

static std::string join(const std::string &dir, const char *name) {
    if (dir.empty()) return name;
    return dir + (dir.back() == '/' ? "" : "/") + name;
}

// std::setprecision(p) in the default float format is printf("%.{p}g").
static void put(FILE *f, double v, int prec) { std::fprintf(f, "%.*g", prec, v); }

static std::string vec_to_str(const double *v, int n) {
    std::ostringstream o;
    for (int i = 0; i < n; i++) {
        std::ostringstream e;
        e << std::setprecision(16) << v[i];
        o << e.str() << " ";
    }
    return o.str();
}

static int write_den(mpdpft_ctx *c, const std::string &path, const std::vector<double> &den) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return MPDPFT_EIO;
    int prec = c->p.OutPrecision;
    for (size_t i = 0; i < c->G; i++)
        for (int s = 0; s < c->S; s++) {
            put(f, den[s * c->G + i], prec);
            std::fputc(s == c->S - 1 ? '\n' : ' ', f);
        }
    std::fclose(f);
    return MPDPFT_OK;
}

static int write_v(mpdpft_ctx *c, const std::string &path, const std::vector<double> &V) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return MPDPFT_EIO;
    int prec = c->p.OutPrecision, EL = c->EL;
    for (size_t i = 0; i < c->G; i++) {
        size_t rem = i;
        int idx[3] = {0, 0, 0};
        for (int d = c->D - 1; d >= 0; d--) {
            idx[d] = (int)(rem % EL);
            rem /= EL;
        }
        for (int d = 0; d < c->D; d++) {
            put(f, c->lat_h[idx[d]], prec);
            std::fputc(' ', f);
        }
        for (int s = 0; s < c->S; s++) {
            put(f, V[s * c->G + i], prec);
            std::fputc(' ', f);
        }
        std::fputc('\n', f);
    }
    std::fclose(f);
    return MPDPFT_OK;
}

static std::string final_block(mpdpft_ctx *c, const mpdpft_result &r) {
    const mpdpft_params &p = c->p;
    std::ostringstream t;
    t << std::setprecision(16);
    double tot = 0.;
    for (int s = 0; s < c->S; s++) tot += r.FinalAbundances[s];
    t << "**** Final Output Variables ****\\\\";
    t << "FinalAbundances              = " << vec_to_str(r.FinalAbundances, c->S) << "\\\\";
    t << "TargetAbundances             = " << vec_to_str(p.Abundances, c->S) << "\\\\";
    t << "muVec                        = " << vec_to_str(r.muVec, c->S) << "\\\\";
    t << "DeltamuModifier              = " << std::setprecision(6) << r.DeltamuModifier << std::setprecision(16) << "\\\\";
    t << "Linear/Pulay-Mixings         = " << r.LinMix << "/" << r.PulMix << "\\\\";
    for (int s = 0; s < c->S; s++)
        t << "AuxilliaryEkin[s=" << s << "]          = " << vec_to_str(r.AuxEkin[s], 4) << "\\\\";
    double ed[MPDPFT_MAX_SPECIES], ee[MPDPFT_MAX_SPECIES];
    for (int s = 0; s < c->S; s++) {
        ed[s] = r.DispersalEnergies[s];
        ee[s] = r.EnvironmentalEnergies[s];
    }
    t << "DispersalEnergies            = " << vec_to_str(ed, c->S) << "\\\\";
    t << "EnvironmentalEnergies        = " << vec_to_str(ee, c->S) << "\\\\";
    t << "InteractionEnergies          = " << vec_to_str(r.InteractionEnergies, p.n_interactions) << "\\\\";
    t << "TotalEnergy                  = " << r.TotalEnergy << "\\\\";
    t << "SCiterations                 = " << r.SCcount - 1 << "/" << p.maxSCcount << "\\\\";
    t << "ConvergenceCheck             = " << std::setprecision(6) << r.CC << "\\\\";
    t << "ComputationalTime            = " << std::setprecision(6) << r.solve_seconds << "\\\\";
    t << "\\newpage \n";
    return t.str();
}

extern "C" int mpdpft_run_dir(const char *dir, int device, mpdpft_result *res) {
    std::string d = dir ? dir : "";
    mpdpft_params p;
    mpdpft_params_default(&p);
    int rc = mpdpft_parse_input(join(d, "mpDPFT.input").c_str(), &p);
    if (rc) return rc;
    mpdpft_ctx *c = nullptr;
    rc = mpdpft_create(&p, device, &c);
    if (rc) {
        if (c) {
            FILE *f = std::fopen(join(d, "mpDPFT_Control.dat").c_str(), "w");
            if (f) {
                std::fprintf(f, "mpdpft_create failed (%d): %s\n", rc, mpdpft_last_error(c));
                std::fclose(f);
            }
        }
        mpdpft_destroy(c);
        return rc;
    }
    mpdpft_result r;
    int src = mpdpft_solve(c, &r);
    if (res) *res = r;
    if (src != MPDPFT_OK && src != MPDPFT_ENOTCONV) {
        mpdpft_destroy(c);
        return src;
    }
    if (c->p.WriteFiles) {
        std::vector<double> den(c->SG), V(c->SG);
        rc = mpdpft_get_density(c, den.data());
        if (!rc) rc = mpdpft_get_potential(c, V.data());
        c->control << "\n O U T P U T ...\n\n ***** export Den ****\n";
        if (!rc) rc = write_den(c, join(d, "mpDPFT_Den.dat"), den);
        if (!(c->D == 3 && c->s > 400)) {
            c->control << " ***** export V ****\n";
            if (!rc) rc = write_v(c, join(d, "mpDPFT_V.dat"), V);
        } else
            c->control << " ***** skip V ****\n";
        c->control << final_block(c, r);
        c->control << " " << r.errorCount << " Error(s);  " << r.warningCount << " Warning(s)\n EndOfControlFile\n";
        FILE *f = std::fopen(join(d, "mpDPFT_Control.dat").c_str(), "w");
        if (!f) rc = MPDPFT_EIO;
        else {
            std::string sc = c->control.str();
            std::fwrite(sc.data(), 1, sc.size(), f);
            std::fclose(f);
        }
    }
    mpdpft_destroy(c);
    return rc ? rc : src;
}
