// output.cpp -- the drop-in directory contract: read mpDPFT.input, solve,
// write mpDPFT_Den.dat, mpDPFT_V.dat, mpDPFT_ContourData.dat, mpDPFT_CutData.dat
// and mpDPFT_Control.dat in the reference formats (ExportData,
// mpDPFT.cpp:12399-12660; final block :12716-12800; WriteControlFile
// :11989-11995).
#include <cmath>
#include <fstream>
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

// Euclidean norm as the reference's Norm (mpDPFT.h:259): squares summed in order
static double norm(const double *v, int n) {
    double s = 0.;
    for (int i = 0; i < n; i++) s += v[i] * v[i];
    return std::sqrt(s);
}

// Grid index of (i, j) in the x-y plane: 2D row-major, 3D at z index steps/2
// (GetFieldInterpolations, mpDPFT.cpp:4545-4546; ContourData :12537-12538).
static size_t plane_index(const mpdpft_ctx *c, int i, int j) {
    const size_t EL = (size_t)c->EL;
    return c->D == 2 ? i * EL + j : (i * EL + j) * EL + (size_t)(c->s / 2);
}

// mpDPFT_ContourData.dat (ExportData, mpDPFT.cpp:12527-12550), DIM 2 and 3: per
// point of the x-y plane its position, |position| and, per species, density,
// environment and potential; a blank line before each new x row and one at
// the end (gnuplot pm3d blocks).
static int write_contour(mpdpft_ctx *c, const std::string &path, const std::vector<double> &den,
                         const std::vector<double> &env, const std::vector<double> &V) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return MPDPFT_EIO;
    const int prec = c->p.OutPrecision, EL = c->EL;
    double posxinit = c->lat_h[0];
    for (int i = 0; i < EL; i++)
        for (int j = 0; j < EL; j++) {
            const size_t idx = plane_index(c, i, j);
            double pos[3] = {c->lat_h[i], c->lat_h[j], c->D == 3 ? c->lat_h[c->s / 2] : 0.};
            if (pos[0] > posxinit) {
                std::fputc('\n', f);
                posxinit = pos[0];
            }
            put(f, pos[0], prec);
            std::fputc(' ', f);
            put(f, pos[1], prec);
            std::fputc(' ', f);
            put(f, norm(pos, c->D), prec);
            for (int s = 0; s < c->S; s++) {
                const size_t k = (size_t)s * c->G + idx;
                std::fputc(' ', f);
                put(f, den[k], prec);
                std::fputc(' ', f);
                put(f, env[k], prec);
                std::fputc(' ', f);
                put(f, V[k], prec);
            }
            std::fputc('\n', f);
        }
    std::fputc('\n', f);
    std::fclose(f);
    return MPDPFT_OK;
}

// Bilinear interpolation on the square grid X x X of F[i n + j] = field(X[i], X[j])
// at (px, py): the cell of each coordinate by bisection (x[l] < t <= x[l+1], the
// first cell at the left end), then the four-corner weights.  The reference
// builds ALGLIB's bilinear spline over the same plane (spline2dbuildbilinearv,
// mpDPFT.cpp:4559) and evaluates it with the coordinates swapped to match its
// row order (:12241); the value is this one.
static void find_cell(const std::vector<double> &X, double t, int &l, double &w) {
    int lo = 0, hi = (int)X.size() - 1;
    while (lo != hi - 1) {
        int h = (lo + hi) / 2;
        if (X[h] >= t) hi = h;
        else lo = h;
    }
    l = lo;
    w = (t - X[l]) / (X[l + 1] - X[l]);
}
static double bilinear(const std::vector<double> &X, const std::vector<double> &F, double px, double py) {
    const int n = (int)X.size();
    int ix, iy;
    double tx, ty;
    find_cell(X, px, ix, tx);
    find_cell(X, py, iy, ty);
    const double f00 = F[(size_t)ix * n + iy], f10 = F[(size_t)(ix + 1) * n + iy];
    const double f11 = F[(size_t)(ix + 1) * n + iy + 1], f01 = F[(size_t)ix * n + iy + 1];
    return (1. - ty) * (1. - tx) * f00 + ty * (1. - tx) * f01 + ty * tx * f11 + (1. - ty) * tx * f10;
}

// mpDPFT_CutData.dat along the default cut OutCut = (-edge/2, 0) -> (edge/2, 0)
// (mpDPFT.cpp:1367): GetCutPositions (:12151-12173), GetFieldInterpolationsAlongCut
// (:12217-12253) and GetCutDataStream (:12175-12213).  Row i: the cut coordinate,
// per species the density, environment and potential at the cut point (DIM 1:
// the grid values; DIM 2/3: interpolated in the x-y plane), and in 3D the lattice
// coordinate and per species the column density sum_k n(i, steps/2, k) dx.
// The cut (GetCutPositions :12151-12173, GetFieldInterpolationsAlongCut
// :12217-12253): its length and fi[(s 3 + p) EL + i], p = density, environment,
// potential of species s at cut point i.
double cut_fields(const mpdpft_ctx *c, const std::vector<double> &den, const std::vector<double> &env,
                  const std::vector<double> &V, std::vector<double> &fi) {
    const int EL = c->EL, steps = c->s;
    std::vector<double> cx(EL), cy(EL, 0.);
    if (c->D == 1) {
        for (int i = 0; i < EL; i++) cx[i] = c->lat_h[i];
    } else {
        const double x0 = -c->p.edge / 2., x1 = c->p.edge / 2.;
        const double DeltaX = (x1 - x0) / (double)steps, DeltaY = (0. - 0.) / (double)steps;
        for (int i = 0; i < EL; i++) {
            cx[i] = x0 + (double)i * DeltaX;
            cy[i] = 0. + (double)i * DeltaY;
        }
    }
    double dv[2] = {cx[0] - cx[EL - 1], cy[0] - cy[EL - 1]};
    const double CutLine = norm(dv, 2);
    fi.assign((size_t)c->S * 3 * EL, 0.);
    const std::vector<double> *src[3] = {&den, &env, &V};
    if (c->D == 1) {
        for (int s = 0; s < c->S; s++)
            for (int pp = 0; pp < 3; pp++)
                for (int i = 0; i < EL; i++) fi[((size_t)s * 3 + pp) * EL + i] = (*src[pp])[(size_t)s * c->G + i];
    } else {
        std::vector<double> X(EL), F((size_t)EL * EL);
        for (int i = 0; i < EL; i++) X[i] = -c->p.edge / 2. + (double)i * c->p.edge / ((double)steps);
        for (int s = 0; s < c->S; s++)
            for (int pp = 0; pp < 3; pp++) {
                for (int i = 0; i < EL; i++)
                    for (int j = 0; j < EL; j++)
                        F[(size_t)i * EL + j] = (*src[pp])[(size_t)s * c->G + plane_index(c, i, j)];
                for (int i = 0; i < EL; i++) fi[((size_t)s * 3 + pp) * EL + i] = bilinear(X, F, cx[i], cy[i]);
            }
    }
    return CutLine;
}

static int write_cut(mpdpft_ctx *c, const std::string &path, const std::vector<double> &den,
                     const std::vector<double> &env, const std::vector<double> &V) {
    const int EL = c->EL, steps = c->s, prec = c->p.OutPrecision;
    std::vector<double> fi;
    const double CutLine = cut_fields(c, den, env, V, fi);
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return MPDPFT_EIO;
    for (int i = 0; i < EL; i++) {
        put(f, -CutLine / 2. + (double)i * CutLine / ((double)steps), prec);
        for (int s = 0; s < c->S; s++)
            for (int pp = 0; pp < 3; pp++) {
                std::fputc(' ', f);
                put(f, fi[((size_t)s * 3 + pp) * EL + i], prec);
            }
        if (c->D == 3) {
            std::fputc(' ', f);
            put(f, c->lat_h[i], prec);
            const int j = steps / 2;
            for (int s = 0; s < c->S; s++) {
                double col = 0.;
                for (int k = 0; k < EL; k++) col += den[(size_t)s * c->G + ((size_t)i * EL + j) * EL + k];
                col *= c->dx;
                std::fputc(' ', f);
                put(f, col, prec);
            }
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
    for (int s = 0; s < c->S; s++)
        t << "AuxilliaryEint[s=" << s << "]          = " << vec_to_str(r.AuxEint[s], 4) << "\\\\";
    double ed[MPDPFT_MAX_SPECIES], ee[MPDPFT_MAX_SPECIES];
    for (int s = 0; s < c->S; s++) {
        ed[s] = r.DispersalEnergies[s];
        ee[s] = r.EnvironmentalEnergies[s];
    }
    t << "DispersalEnergies            = " << vec_to_str(ed, c->S) << "\\\\";
    const double zeros[MPDPFT_MAX_SPECIES] = {};  // the reference leaves these at 0 (:5416, :5644)
    t << "AlternativeDispersalEnergies = " << vec_to_str(zeros, c->S) << "\\\\";
    t << "EnvironmentalEnergies        = " << vec_to_str(ee, c->S) << "\\\\";
    t << "InteractionEnergies          = " << vec_to_str(r.InteractionEnergies, p.n_interactions) << "\\\\";
    t << "TotalEnergy                  = " << r.TotalEnergy << "\\\\";
    t << "FigureOfMerit                = " << c->fom << "\\\\";  // GetFigureOfMerit (:13743-13793); 0 outside task types 5-7
    t << "SCiterations                 = " << r.SCcount - 1 << "/" << p.maxSCcount << "\\\\";
    t << "ConvergenceCheck             = " << std::setprecision(6) << r.CC << "\\\\";
    t << "ComputationalTime            = " << std::setprecision(6) << r.solve_seconds << "\\\\";
    t << "\\newpage \n";
    return t.str();
}

// mpDPFT_Den_Cube.dat (ExportData, mpDPFT.cpp:12458-12490), written in place of
// mpDPFT_Den.dat when the extension key CubeFile is 1 (the reference writes it
// for its chemistry environment 10; here the nuclei are the environment-41
// centres): the nucleus count, three axis lines of steps and spacing, one line
// per nucleus (charge and position), then species 0's particles per voxel,
// max(1e-14, n dx^3), over [0, steps)^3 (the periodic images left out), x
// slowest.  With Units 2 lengths are converted from Angstrom to bohr.
static int write_cube(mpdpft_ctx *c, const std::string &path, const std::vector<double> &den) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return MPDPFT_EIO;
    const mpdpft_params &p = c->p;
    const int prec = p.OutPrecision, EL = c->EL, steps = c->s;
    double cv = 1., c3 = 1.;
    if (p.Units == 2) {
        cv = 0.52917721067;
        c3 = cv * cv * cv;
    }
    const double h = c->dx / cv, vol = h * h * h;
    const int nuc = p.Environments[0] == MPDPFT_ENV_SOFT_COULOMB_CENTRES ? p.n_centres : 0;
    std::fprintf(f, "%d 0 0 0\n", nuc);
    for (int a = 0; a < 3; a++) {
        std::fprintf(f, "%d ", steps);
        for (int b = 0; b < 3; b++) {
            if (a == b) put(f, h, prec);
            else std::fputc('0', f);
            if (b < 2) std::fputc(' ', f);
        }
        std::fputs(a == 2 ? " \n" : "\n", f);
    }
    if (nuc > 0) {
        for (int k = 0; k < nuc; k++) {
            std::fprintf(f, "%.16g ", p.centres[k][3]);
            for (int d = 0; d < 3; d++) std::fprintf(f, "%.16g ", p.centres[k][d] / cv);
            std::fputc('\n', f);
        }
    } else
        std::fputs("0 0 0 0\n", f);
    for (int i = 0; i < steps; i++)
        for (int j = 0; j < steps; j++)
            for (int k = 0; k < steps; k++) {
                const size_t idx = ((size_t)i * EL + j) * EL + k;
                put(f, std::max(1.0e-14, c3 * den[idx] * vol), prec);
                std::fputc('\n', f);
            }
    std::fclose(f);
    return MPDPFT_OK;
}

// mpDPFT_MovieData.tmp and mpDPFT_Movie.sh (ExportData, mpDPFT.cpp:12660-12691),
// MovieQ 1 (linear) or 2 (log scale): the stored cut frames, one row per cut
// point (the coordinate, then per frame and species the density, environment
// and potential; 4 significant digits, a space before each value, no newline
// after the last row), and a script that draws every frame's densities with
// gnuplot (image.<frame>.png) and joins them with ffmpeg.  The script reads the
// .tmp file (the reference's script names a .dat file it never writes).
static int write_movie(mpdpft_ctx *c, const std::string &dir) {
    if (c->movie.empty() || c->frames < 1) return MPDPFT_OK;
    FILE *f = std::fopen(join(dir, "mpDPFT_MovieData.tmp").c_str(), "w");
    if (!f) return MPDPFT_EIO;
    const int EL = c->EL, S = c->S;
    double lo = 0., hi = 0.;
    bool first = true;
    for (int i = 0; i < EL; i++) {
        for (size_t j = 0; j < c->movie[i].size(); j++) {
            std::fprintf(f, " %.4g", c->movie[i][j]);
            if (j > 0 && (j - 1) % 3 == 0) {  // a density column
                const double v = c->movie[i][j];
                lo = first ? v : std::min(lo, v);
                hi = first ? v : std::max(hi, v);
                first = false;
            }
        }
        if (i < EL - 1) std::fputc('\n', f);
    }
    std::fclose(f);
    const double d = hi - lo;
    std::ostringstream sh;
    sh << std::setprecision(10);
    sh << "#!/bin/bash\n# frames of the density cuts in mpDPFT_MovieData.tmp (gnuplot, png), joined by ffmpeg\n"
       << "GNUPLOT=${GNUPLOT:-gnuplot}\ncd \"$(dirname \"$0\")\" || exit 1\n$GNUPLOT <<'EOF'\n"
       << "set terminal pngcairo size 800,600\nset grid\n";
    if (c->p.MovieQ == 2) sh << "set logscale y\nset yrange [" << std::max(1e-16, lo - 0.1 * d) << ":" << hi + 0.1 * d << "]\n";
    else sh << "set yrange [" << lo - 0.1 * d << ":" << hi + 0.1 * d << "]\n";
    sh << "do for [frame=0:" << c->frames - 1 << "] {\n  set output sprintf('image.%d.png', frame)\n  plot";
    for (int s = 0; s < S; s++)
        sh << (s ? "," : "") << " 'mpDPFT_MovieData.tmp' using 1:(column(" << 2 + 3 * s << " + frame*" << 3 * S
           << ")) with lines lw 3 title sprintf('Den" << s << " frame %d', frame)";
    sh << "\n}\nEOF\n"
       << "ffmpeg -y -framerate " << ((double)c->frames + 1.) / 20. << " -i image.%d.png -pix_fmt yuv420p mpDPFT_Movie.mp4\n";
    f = std::fopen(join(dir, "mpDPFT_Movie.sh").c_str(), "w");
    if (!f) return MPDPFT_EIO;
    const std::string t = sh.str();
    std::fwrite(t.data(), 1, t.size(), f);
    std::fclose(f);
    return MPDPFT_OK;
}

// mpDPFT_Plots.sh and mpDPFT_CombinedPlots.tex (the plot shells of ExportData,
// mpDPFT.cpp:12802-13063): per species the x-y maps of density, environment
// and potential from mpDPFT_ContourData.dat (DIM 2, 3), the density and the
// environment + potential (with mu) along the cut from mpDPFT_CutData.dat, the
// density cut on a log scale, and for several species an overview of all
// density cuts.  The script drives gnuplot's pdfcairo terminal, one PDF per
// figure named after the figure (mpDPFT_Plots-<figure>.pdf); the LaTeX file
// holds the run's result block and includes every figure.
static int write_plots(mpdpft_ctx *c, const std::string &dir, const mpdpft_result &r, const std::vector<double> &den,
                       const std::vector<double> &env, const std::vector<double> &V) {
    const int S = c->S;
    const double L2 = c->p.edge / 2.;
    auto range = [&](const std::vector<double> &f, int s, double &lo, double &hi) {
        lo = hi = f[(size_t)s * c->G];
        for (size_t i = 0; i < c->G; i++) {
            const double v = f[(size_t)s * c->G + i];
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
        lo -= 0.1 * std::fabs(lo);
        hi += 0.1 * std::fabs(hi);
        if (std::fabs(hi - lo) < 1e-12) {
            lo = 0.;
            hi = 1.;
        }
    };
    std::ostringstream sh, inc;
    sh << std::setprecision(10);
    sh << "#!/bin/bash\n"
       << "# gnuplot figures of mpDPFT_ContourData.dat and mpDPFT_CutData.dat (pdfcairo), one PDF per figure,\n"
       << "# named as mpDPFT_CombinedPlots.tex includes them\n"
       << "GNUPLOT=${GNUPLOT:-gnuplot}\n"
       << "cd \"$(dirname \"$0\")\" || exit 1\n"
       << "figure() {  # name width height; gnuplot commands on stdin\n"
       << "  { printf 'set terminal pdfcairo enhanced size %s,%s\\nset output \"%s.pdf\"\\n' \"$2\" \"$3\" \"$1\"; cat; } "
          "| $GNUPLOT && echo \"Generated $1.pdf\"\n"
       << "}\n";
    const char *fields[3] = {"Den", "Env", "V"};
    const std::vector<double> *src[3] = {&den, &env, &V};
    auto add = [&](const std::string &name) { inc << "\\includegraphics[width=0.32\\textwidth]{mpDPFT_Plots-" << name << "}\n"; };
    std::ostringstream cut;
    cut << " along (" << -L2 << ",0) to (" << L2 << ",0)";
    for (int s = 0; s < S; s++) {
        double lo, hi;
        if (c->D > 1)
            for (int k = 0; k < 3; k++) {
                range(*src[k], s, lo, hi);
                const std::string name = std::string("2Dcontour") + fields[k] + std::to_string(s);
                sh << "figure mpDPFT_Plots-" << name << " 10cm 10cm <<'EOF'\n"
                   << "set title '" << fields[k] << s << "'\nset view map\nunset surface\nset style data pm3d\n"
                   << "set xrange [" << -L2 << ":" << L2 << "]\nset yrange [" << -L2 << ":" << L2 << "]\n"
                   << "set cbrange [" << lo << ":" << hi << "]\n"
                   << "splot 'mpDPFT_ContourData.dat' using 1:2:" << 3 * s + 4 + k << " with pm3d notitle\nEOF\n";
                add(name);
            }
        range(den, s, lo, hi);
        std::string name = "CutDen" + std::to_string(s);
        sh << "figure mpDPFT_Plots-" << name << " 10cm 6.875cm <<'EOF'\n"
           << "set title 'Den" << s << cut.str() << "'\nset grid\nset yrange [" << lo << ":" << hi << "]\n"
           << "plot 'mpDPFT_CutData.dat' using 1:" << 3 * s + 2 << " with lines lw 3 title 'Den" << s << "'\nEOF\n";
        add(name);
        name = "LogPlot-CutDen" + std::to_string(s);
        sh << "figure mpDPFT_Plots-" << name << " 10cm 6.875cm <<'EOF'\n"
           << "set title 'LogPlotDen" << s << cut.str() << "'\nset logscale y\nset format y '%.0e'\n"
           << "set yrange [" << std::max(lo, 1e-16) << ":" << std::max(hi, 1e-15) << "]\n"
           << "plot 'mpDPFT_CutData.dat' using 1:" << 3 * s + 2 << " with lines lw 3 title 'Den" << s << "'\nEOF\n";
        add(name);
        double elo, ehi, vlo, vhi;
        range(env, s, elo, ehi);
        range(V, s, vlo, vhi);
        lo = std::min(std::min(elo, vlo), r.muVec[s] - 0.1 * std::fabs(r.muVec[s]));
        hi = std::max(std::max(ehi, vhi), r.muVec[s] + 0.1 * std::fabs(r.muVec[s]));
        name = "CutEnv+V" + std::to_string(s);
        sh << "figure mpDPFT_Plots-" << name << " 10cm 6.875cm <<'EOF'\n"
           << "set title 'Env+V" << s << cut.str() << "'\nset grid\nset yrange [" << lo << ":" << hi << "]\n"
           << "mu(x) = " << std::setprecision(17) << r.muVec[s] << std::setprecision(10) << "\n"
           << "plot 'mpDPFT_CutData.dat' using 1:" << 3 * s + 3 << " with lines lw 3 title 'Env" << s
           << "', mu(x) dt 2 lw 3 title 'mu', 'mpDPFT_CutData.dat' using 1:" << 3 * s + 4
           << " with lines lw 2 title 'V" << s << "'\nEOF\n";
        add(name);
    }
    if (S > 1) {
        double lo = 0., hi = 0.;
        for (int s = 0; s < S; s++) {
            double a, b;
            range(den, s, a, b);
            lo = s ? std::min(lo, a) : a;
            hi = s ? std::max(hi, b) : b;
        }
        sh << "figure mpDPFT_Plots-Overview-CutDen 10cm 6.875cm <<'EOF'\n"
           << "set title 'Overview Den" << cut.str() << "'\nset grid\nset yrange [" << lo << ":" << hi << "]\nplot";
        for (int s = 0; s < S; s++)
            sh << (s ? ", " : " ") << "'mpDPFT_CutData.dat' using 1:" << 3 * s + 2 << " with lines lw 3 title 'Den"
               << s << "'";
        sh << "\nEOF\n";
        add("Overview-CutDen");
    }
    FILE *f = std::fopen(join(dir, "mpDPFT_Plots.sh").c_str(), "w");
    if (!f) return MPDPFT_EIO;
    std::string t = sh.str();
    std::fwrite(t.data(), 1, t.size(), f);
    std::fclose(f);
    // the result block is LaTeX already ("\\" line ends); escape the underscores of its names
    std::string block;
    for (char ch : final_block(c, r)) {
        if (ch == '_') block += "\\_";
        else block += ch;
    }
    std::ostringstream tex;
    tex << "\\documentclass[10pt]{article}\n\\usepackage{graphicx}\n\\usepackage[margin=1.5cm]{geometry}\n"
        << "\\begin{document}\n\\noindent " << block << "\n\\noindent\n" << inc.str() << "\\end{document}\n";
    f = std::fopen(join(dir, "mpDPFT_CombinedPlots.tex").c_str(), "w");
    if (!f) return MPDPFT_EIO;
    t = tex.str();
    std::fwrite(t.data(), 1, t.size(), f);
    std::fclose(f);
    return MPDPFT_OK;
}

// One GetData call (mpDPFT.cpp:135-244) in directory d with parameters p:
// LoadV when InterpolVQ is 1, the SC loop, and ExportData when write_files.
// LoadReferenceDensity (mpDPFT.cpp:13247-13280), 2D: the n x n table (x index
// fastest, on the lattice -edge/2 + edge i/(n-1)) interpolated bilinearly onto
// the run's lattice, value at (x_i, y_j) stored at j EL + i as the reference
// does; then GetFigureOfMerit for task types 4-7 (:13771-13775):
// 1 - 2 <n_0, ref> / (|n_0|^2 + |ref|^2), plain sums over the grid points.
static double figure_of_merit(const mpdpft_ctx *c, const std::vector<double> &tab, const std::vector<double> &den) {
    const int n = (int)(std::sqrt((double)tab.size()) + 0.5), EL = c->EL;
    std::vector<double> X(n), ref(c->G);
    for (int i = 0; i < n; i++) X[i] = -0.5 * c->p.edge + c->p.edge * (double)i / ((double)(n - 1));
    auto cellof = [&](double t, int &l, double &f) {
        l = (int)std::floor((t - X[0]) / (X[1] - X[0]));
        l = std::min(std::max(l, 0), n - 2);
        f = (t - X[l]) / (X[l + 1] - X[l]);
    };
    for (int i = 0; i < EL; i++)
        for (int j = 0; j < EL; j++) {
            int a, b;
            double t, u;
            cellof(c->lat_h[i], a, t);
            cellof(c->lat_h[j], b, u);
            const double f00 = tab[(size_t)b * n + a], f10 = tab[(size_t)b * n + a + 1],
                         f01 = tab[(size_t)(b + 1) * n + a], f11 = tab[(size_t)(b + 1) * n + a + 1];
            ref[(size_t)j * EL + i] = (1. - t) * (1. - u) * f00 + t * (1. - u) * f10 + (1. - t) * u * f01 + t * u * f11;
        }
    double dot = 0., nn = 0., rr = 0.;
    for (size_t k = 0; k < c->G; k++) {
        dot += den[k] * ref[k];
        nn += den[k] * den[k];
        rr += ref[k] * ref[k];
    }
    return 1. - 2. * dot / (nn + rr);
}

int run_one(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res, bool write_files,
            const std::string &notes, double *mpp_out, double *task_grad, int *grad_rc,
            const std::vector<double> *ref, double *fom, const std::vector<double> *v_start,
            std::vector<double> *v_out) {
    mpdpft_ctx *c = nullptr;
    int rc = mpdpft_create(&p, device, &c);
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
    c->notes << notes;
    mpdpft_set_break_file(c, join(d, "mpDPFTmanualSCloopBreakQ.dat").c_str());
    if (p.InterpolVQ == 1) {  // LoadV (:184)
        std::vector<double> v0(c->SG);
        int warn = 0;
        rc = mpdpft_import_v(join(d, "mpDPFT_V.dat").c_str(), &p, v0.data(), &warn);
        if (!rc) rc = mpdpft_set_initial_potential(c, v0.data());
        if (rc) {
            mpdpft_destroy(c);
            return rc;
        }
        c->notes << " ***** load existing V ***** \nLoadV: Interpolate V from mpDPFT_V.dat\n";
        if (warn) c->notes << "LoadV: " << warn << " warning(s) in the imported field\n";
        c->importWarnings = warn;
    } else if (v_start) {  // InterpolVQ = 2: data.V = task.V (LoadV :6500)
        rc = mpdpft_set_initial_potential(c, v_start->data());
        if (rc) {
            mpdpft_destroy(c);
            return rc;
        }
    }
    mpdpft_result r;
    int src = mpdpft_solve(c, &r);
    if (res) *res = r;
    if (mpp_out && p.DensityExpression == 5)  // the adapted (BoolePairs, maxIter) per species
        for (int s = 0; s < c->S; s++) nait_state(c, s, &mpp_out[2 * s], &mpp_out[2 * s + 1]);
    if (src != MPDPFT_OK && src != MPDPFT_ENOTCONV) {
        mpdpft_destroy(c);
        return src;
    }
    if (task_grad) *grad_rc = mpdpft_task_gradient(c, task_grad);
    if (v_out) {
        v_out->resize(c->SG);
        rc = mpdpft_get_potential(c, v_out->data());
        if (rc) {
            mpdpft_destroy(c);
            return rc;
        }
    }
    if (ref) {
        std::vector<double> den0(c->SG);
        rc = mpdpft_get_density(c, den0.data());  // [S][G]: species 0 first
        if (rc) {
            mpdpft_destroy(c);
            return rc;
        }
        c->fom = figure_of_merit(c, *ref, den0);
        if (fom) *fom = c->fom;
    }
    if (write_files && c->p.WriteFiles) {
        std::vector<double> den(c->SG), V(c->SG);
        rc = mpdpft_get_density(c, den.data());
        if (!rc) rc = mpdpft_get_potential(c, V.data());
        c->control << "\n O U T P U T ...\n\n";
        if (c->p.CubeFile && c->D == 3) {
            c->control << " ***** export Cube file ****\n";
            if (!rc) rc = write_cube(c, join(d, "mpDPFT_Den_Cube.dat"), den);
        } else {
            c->control << " ***** export Den ****\n";
            if (!rc) rc = write_den(c, join(d, "mpDPFT_Den.dat"), den);
        }
        if (!(c->D == 3 && c->s > 400)) {
            c->control << " ***** export V ****\n";
            if (!rc) rc = write_v(c, join(d, "mpDPFT_V.dat"), V);
        } else
            c->control << " ***** skip V ****\n";
        std::vector<double> env;
        if (!rc) {
            env.resize(c->SG);
            rc = mpdpft_get_environment(c, env.data());
        }
        if (c->D > 1) {
            c->control << " ***** export ContourData ****\n";
            if (!rc) rc = write_contour(c, join(d, "mpDPFT_ContourData.dat"), den, env, V);
        }
        c->control << " ***** export CutData ****\n";
        if (!rc) rc = write_cut(c, join(d, "mpDPFT_CutData.dat"), den, env, V);
        c->control << " ***** produce plot shells ****\n";
        if (!rc) rc = write_plots(c, d, r, den, env, V);
        if (c->p.MovieQ > 0 && r.SCcount > 0) {
            c->control << " ***** export MovieData ****\n";
            if (!rc) rc = write_movie(c, d);
        }
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

// Task types 5, 6, 7 (mpDPFTmain.cpp:110-166, RetrieveTask mpDPFT.cpp:6086-6100,
// ProcessTASK :170-185): a rectangular scan of TaskParameterCount values per
// variable over the TaskHyperBox, v = lo + (hi - lo) t / (count - 1); type 7
// varies mpp[0], type 5 mpp[0] and gammaH, type 6 tauVec[0], mpp[0] and
// gammaH (the last variable fastest).  Each run's figure of merit against
// TabFunc_ReferenceDensity.dat (2D) goes with its variables into
// mpDPFT_TASKParameterExploration.dat; the last run writes its files.
static int run_exploration(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res) {
    const int type = p.TaskType, nv = type == 5 ? 2 : (type == 6 ? 3 : 1), count = p.TaskParameterCount;
    std::ostringstream task;
    task << std::setprecision(16);
    task << " ***** mpdpft_hip task loop: TaskType " << type << " *****\n";
    auto fail = [&](int code, const std::string &msg) {
        task << msg << "\n";
        FILE *f = std::fopen(join(d, "mpDPFT_ControlTask.dat").c_str(), "w");
        if (f) {
            std::string t = task.str();
            std::fwrite(t.data(), 1, t.size(), f);
            std::fclose(f);
        }
        return code;
    };
    if (p.DIM != 2) return fail(MPDPFT_EINVAL, "task types 5-7 compare with a 2D reference density (DIM 2)");
    if (count < 2) return fail(MPDPFT_EINVAL, "task types 5-7 need TaskParameterCount >= 2");
    if (p.n_hyperbox < 2 * nv) return fail(MPDPFT_EINVAL, "task types 5-7 need one TaskHyperBox pair per variable");
    std::vector<double> tab;
    {
        std::ifstream in(join(d, "TabFunc_ReferenceDensity.dat"));
        if (!in) return fail(MPDPFT_EIO, "TabFunc_ReferenceDensity.dat not found");
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream iss(line);
            double v = 0.;
            if (!(iss >> v)) v = 0.;
            tab.push_back(v);
        }
        const int n = (int)(std::sqrt((double)tab.size()) + 0.5);
        if (n < 2 || (size_t)n * n != tab.size())
            return fail(MPDPFT_EIO, "TabFunc_ReferenceDensity.dat does not hold n x n values");
    }
    auto value = [&](int k, int t) {
        return p.TaskHyperBox[2 * k] + (p.TaskHyperBox[2 * k + 1] - p.TaskHyperBox[2 * k]) * (double)t /
                                           ((double)count - 1.);
    };
    const int total = nv == 1 ? count : (nv == 2 ? count * count : count * count * count);
    std::ostringstream rows;
    rows << std::setprecision(16);
    mpdpft_result r{};
    int last_rc = MPDPFT_OK;
    for (int k = 0; k < total; k++) {
        int t[3] = {0, 0, 0}, rem = k;
        for (int v = nv - 1; v >= 0; v--) {
            t[v] = rem % count;
            rem /= count;
        }
        double vec[3];
        for (int v = 0; v < nv; v++) vec[v] = value(v, t[v]);
        mpdpft_params q = p;
        if (type == 7) q.mpp[0] = vec[0];
        else if (type == 5) {
            q.mpp[0] = vec[0];
            q.gammaH = vec[1];
        } else {
            q.tauVec[0] = vec[0];
            q.mpp[0] = vec[1];
            q.gammaH = vec[2];
        }
        double fom = 0.;
        last_rc = run_one(d, q, device, &r, k == total - 1, "", nullptr, nullptr, nullptr, &tab, &fom);
        if (last_rc != MPDPFT_OK && last_rc != MPDPFT_ENOTCONV) return fail(last_rc, "run failed");
        rows << fom << " ";
        for (int v = 0; v < nv; v++) rows << vec[v] << " ";
        rows << "\n";
        task << "task " << k << ": VEC " << vec_to_str(vec, nv) << " FigureOfMerit " << fom << "  TotalEnergy "
             << r.TotalEnergy << "  converged " << r.converged << "\n";
    }
    if (res) *res = r;
    FILE *f = std::fopen(join(d, "mpDPFT_TASKParameterExploration.dat").c_str(), "w");
    if (!f) return MPDPFT_EIO;
    std::string t = rows.str();
    std::fwrite(t.data(), 1, t.size(), f);
    std::fclose(f);
    f = std::fopen(join(d, "mpDPFT_ControlTask.dat").c_str(), "w");
    if (f) {
        t = task.str();
        std::fwrite(t.data(), 1, t.size(), f);
        std::fclose(f);
    }
    return last_rc;
}

// TaskType 10 (mpDPFTmain.cpp:179-198, RetrieveTask mpDPFT.cpp:6124-6131):
// max(1, TaskParameterCount) repetitions of the solve, which differ by their
// Noise (repetition k seeds the generator with NoiseSeed + k; the reference's
// seed comes from the clock), then one more run started from the potential
// and mu of the repetition with the lowest total energy (InterpolVQ 2) with
// Schedule 0, Noise 0 and maxSCcount 0: no SC loop, the densities of that
// potential and the output files.  mpDPFT_ControlTask.dat gets the
// reference's lines on each new minimum (or no improvement) with the sorted
// energy history.
static int run_repetitions(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res) {
    const int count = std::max(1, p.TaskParameterCount);
    std::ostringstream task;
    task << std::setprecision(16);
    task << " ***** mpdpft_hip task loop: TaskType 10, " << count << " repetition(s) *****\n";
    std::vector<double> hist, bestV, V;
    double minE = 0., bestMu[MPDPFT_MAX_SPECIES] = {};
    mpdpft_result r{};
    int last_rc = MPDPFT_OK;
    for (int k = 0; k <= count; k++) {
        const bool last = k == count;
        mpdpft_params q = p;
        std::ostringstream notes;
        notes << std::setprecision(16);
        if (!last) {
            if (p.NoiseSeed) q.NoiseSeed = p.NoiseSeed + (unsigned long long)k;
        } else {
            q.InterpolVQ = 0;  // the potential comes from bestV (InterpolVQ 2)
            for (int s = 0; s < q.S; s++) q.muVec[s] = bestMu[s];
            q.n_schedule = 1;
            q.Schedule[0] = 0;
            q.Noise = 0.;
            q.maxSCcount = 0;
            notes << "InterpolVQ -> 2\nmuVec -> " << vec_to_str(q.muVec, q.S)
                  << "\nSchedule -> 0\nNoise -> 0\nmaxSCcount -> 0\n";
        }
        last_rc = run_one(d, q, device, &r, last, notes.str(), nullptr, nullptr, nullptr, nullptr, nullptr,
                          last ? &bestV : nullptr, last ? nullptr : &V);
        if (last_rc != MPDPFT_OK && last_rc != MPDPFT_ENOTCONV) break;
        if (last) break;
        hist.push_back(r.TotalEnergy);
        std::sort(hist.begin(), hist.end());
        std::ostringstream h;
        h << std::setprecision(16);
        for (size_t i = 0; i < hist.size(); i++) h << (i ? " " : "") << hist[i];
        if (k == 0 || r.TotalEnergy < minE) {
            minE = r.TotalEnergy;
            bestV.swap(V);
            for (int s = 0; s < q.S; s++) bestMu[s] = r.muVec[s];
            task << " ***** task.Type==10: new minEtot = " << minE << " @ task.count = " << k << "\n";
            task << "\n       Etot history (sorted) = " << h.str() << "\n\n";
        } else {
            task << " ***** task.Type==10: no improvement in minEtot @ task.count = " << k
                 << "\n       Etot history (sorted) = " << h.str() << "\n\n";
        }
    }
    task << "final: muVec " << vec_to_str(r.muVec, p.S) << "  NVec " << vec_to_str(r.FinalAbundances, p.S)
         << "  TotalEnergy " << r.TotalEnergy << "\n";
    if (res) *res = r;
    FILE *f = std::fopen(join(d, "mpDPFT_ControlTask.dat").c_str(), "w");
    if (f) {
        std::string sc = task.str();
        std::fwrite(sc.data(), 1, sc.size(), f);
        std::fclose(f);
    }
    return last_rc;
}

// The outer task loop of main (mpDPFTmain.cpp:77-214) for task types 0, 1 and 8
// (2 and 3, the conjugate-gradient minimisations, are in tasks.cpp),
// with RetrieveTask's per-task parameter changes (mpDPFT.cpp:6070-6121):
// 1: rectangular grid of starting chemical potentials mu_s (1 + delta)^p_s,
//    delta = 1.1^(1/AxesPoints) - 1 (two species); files of the last task only;
// 8: successive refinement: task k runs on steps 2^k with InterpolVQ = 1 (LoadV of
//    the previous task's mpDPFT_V.dat), the previous mu and adapted nAiT Mpp, and
//    incrementalV = 1; every task writes its files.
// mpDPFT_ControlTask.dat gets one line per task.
extern "C" int mpdpft_run_dir(const char *dir, int device, mpdpft_result *res) {
    std::string d = dir ? dir : "";
    mpdpft_set_data_dir(d.c_str());  // the TabFunc files sit beside mpDPFT.input, as for the reference
    mpdpft_params p;
    mpdpft_params_default(&p);
    int rc = mpdpft_parse_input(join(d, "mpDPFT.input").c_str(), &p);
    if (rc) return rc;
    const int type = p.TaskType;
    bool sched11 = false;
    for (int i = 0; i < p.n_schedule; i++) sched11 = sched11 || p.Schedule[i] == 11;
    const char *refuse = nullptr;
    if (type < 0 || type > 10 || type == 4) refuse = "is not built (0-3, 5-10 are)";
    else if (type == 8 && sched11 && p.Noise > 0.)  // its last task zeroes NoiseOriginal alone (:903)
        refuse = "with Schedule 11 and Noise > 0 is not built";
    if (refuse) {
        FILE *f = std::fopen(join(d, "mpDPFT_Control.dat").c_str(), "w");
        if (f) {
            std::fprintf(f, "TaskType %d %s\n", type, refuse);
            std::fclose(f);
        }
        return MPDPFT_EINVAL;
    }
    if (type >= 5 && type <= 7) return run_exploration(d, p, device, res);
    if (type == 10) return run_repetitions(d, p, device, res);
    if (type == 2 || type == 3 || type == 9) {
        std::ostringstream task;
        task << " ***** mpdpft_hip task loop: TaskType " << type << " *****\n";
        mpdpft_result r{};
        int rc2 = type == 9 ? run_gd_task(d, p, device, &r, task) : run_cgd_task(d, p, device, &r, task);
        if (res) *res = r;
        FILE *f = std::fopen(join(d, "mpDPFT_ControlTask.dat").c_str(), "w");
        if (f) {
            std::string sc = task.str();
            std::fwrite(sc.data(), 1, sc.size(), f);
            std::fclose(f);
        }
        return rc2;
    }
    int count = (type == 0 || p.TaskParameterCount < 1) ? 1 : p.TaskParameterCount;
    if (type == 1 && p.TaskParameterCount <= 1) count = 1;
    std::ostringstream task;
    task << std::setprecision(16);
    task << " ***** mpdpft_hip task loop: TaskType " << type << ", " << count << " task(s) *****\n";
    mpdpft_result r{};
    double mpp[2 * MPDPFT_MAX_SPECIES] = {};
    int last_rc = MPDPFT_OK;
    for (int k = 0; k < count; k++) {
        mpdpft_params q = p;
        std::ostringstream notes;
        bool files = type != 1 || k == count - 1;
        if (type == 1 && count > 1) {
            if (q.S < 2) return MPDPFT_EINVAL;
            int axes = (int)(std::pow((double)count, 1. / (double)q.S) + 0.5);
            int p0 = (k - k % axes) / axes, p1 = k % axes;
            double delta = std::pow(1.1, 1. / (double)axes) - 1.;
            q.muVec[0] *= std::pow(1. + delta, (double)p0);
            q.muVec[1] *= std::pow(1. + delta, (double)p1);
            notes << "-------> TASK = " << k + 1 << "/" << count << "\n";
        } else if (type == 8 && k > 0) {
            q.steps = p.steps * (int)(std::pow(2., (double)k) + 0.5);
            q.InterpolVQ = 1;
            if (q.incrementalV < 1.) q.incrementalV = 1.;
            for (int s = 0; s < q.S; s++) q.muVec[s] = r.muVec[s];
            if (q.DensityExpression == 5) {  // the adapted Mpp (one list for all species here)
                q.Mpp[0] = mpp[0];
                q.Mpp[1] = mpp[1];
                if (k == count - 1) q.DeltamuModifier = r.DeltamuModifier;
            }
            notes << "successive InterpolV; task.count = " << k << "\nsteps -> " << q.steps << "\n";
        }
        last_rc = run_one(d, q, device, &r, files, notes.str(), mpp);
        if (last_rc != MPDPFT_OK && last_rc != MPDPFT_ENOTCONV) return last_rc;
        task << "task " << k << ": steps " << q.steps << "  muVec " << vec_to_str(r.muVec, q.S) << "  NVec "
             << vec_to_str(r.FinalAbundances, q.S) << "  TotalEnergy " << r.TotalEnergy << "  SCiterations "
             << r.SCcount - 1 << "  converged " << r.converged << "\n";
    }
    if (res) *res = r;
    FILE *f = std::fopen(join(d, "mpDPFT_ControlTask.dat").c_str(), "w");
    if (f) {
        std::string sc = task.str();
        std::fwrite(sc.data(), 1, sc.size(), f);
        std::fclose(f);
    }
    return last_rc;
}
