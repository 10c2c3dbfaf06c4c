// tasks.cpp -- TaskType 2 and 3: the minimum of the total energy over the
// chemical potentials (2, GetGlobalMinimum_muVec, mpDPFT.cpp:246-253) or over
// the abundances (3, GetGlobalMinimum_Abundances, :256-285).  The reference
// runs ALGLIB's nonlinear conjugate gradients (ConjugateGradientDescent
// :8365-8522) on finite-difference gradients of whole solves
// (functionToBeMinimized :8524-8584, ALGLIB's 4-point formula: four solves per
// coordinate).  Here a one-species solve's gradient is the fixed point's
// adjoint (mpdpft_task_gradient: one linear solve after the SC loop); other
// solves fall back to the same 4-point differences of whole solves.  The
// optimiser is this file's own: Polak-Ribiere+ conjugate gradients
// preconditioned by the squared HyperBox widths (the reference scales its
// variables by them: mincgsetscale + mincgsetprecscale) with a bracketing line
// search on the directional derivative (sufficient decrease and curvature
// conditions).  Stopping: the reference's EpsF rule |f_k+1 - f_k| <=
// EpsF max(|f_k|, |f_k+1|, 1) (EpsF 1e-4 for System 0 and 2, 0 for System 1),
// a scaled-gradient rule max_i |w_i g_i| <= 1e-9 max(|f|, 1), a line search
// that cannot lower f (ALGLIB's termination type 7) or 200 iterations.
// The start is the HyperBox centre (the reference draws a uniform point of the
// box from its own random stream).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "solver.h"

namespace {

constexpr double kMPt = 1.0e-12;  // MP (mpDPFT.cpp:111)

std::string vstr(const std::vector<double> &v, int prec = 16) {
    std::ostringstream o;
    o << std::setprecision(prec);
    for (size_t i = 0; i < v.size(); i++) o << (i ? " " : "") << v[i];
    return o.str();
}

struct Objective {
    std::string dir;
    mpdpft_params p;
    int device = 0;
    int type = 3;
    bool adjoint = true;  // try mpdpft_task_gradient first; differences after it refuses once
    double diffstep = 1e-6;
    int solves = 0, grads = 0;
    std::ostringstream *log = nullptr;

    mpdpft_params at(const std::vector<double> &x) const {
        mpdpft_params q = p;
        for (int s = 0; s < q.S; s++) {
            if (type == 2) q.muVec[s] = x[s];
            else q.Abundances[s] = std::max(0., x[s]);  // POS (functionToBeMinimized :8557)
        }
        return q;
    }

    int value(const std::vector<double> &x, double *f, double *grad, mpdpft_result *r) {
        int grc = MPDPFT_EINVAL;
        mpdpft_result rr{};
        int rc = run_one(dir, at(x), device, &rr, false, "", nullptr, grad, grad ? &grc : nullptr);
        solves++;
        if (rc != MPDPFT_OK && rc != MPDPFT_ENOTCONV) return rc;
        *f = rr.TotalEnergy;
        if (r) *r = rr;
        if (grad && grc != MPDPFT_OK) return -1000 + grc;  // the adjoint refused: caller switches to differences
        return MPDPFT_OK;
    }

    // f and its gradient at x
    int eval(const std::vector<double> &x, double *f, std::vector<double> &g, mpdpft_result *r) {
        const int n = (int)x.size();
        g.assign(n, 0.);
        if (adjoint && p.S == 1) {
            int rc = value(x, f, g.data(), r);
            if (rc == MPDPFT_OK) {
                if (type == 3 && x[0] < 0.) g[0] = 0.;  // d POS(x)/dx
                grads++;
                return MPDPFT_OK;
            }
            if (rc > -900) return rc;
            adjoint = false;
            *log << "ConjugateGradientDescent: the adjoint gradient is not available for this functional (code "
                 << rc + 1000 << "); 4-point differences of whole solves from here on\n";
        }
        int rc = value(x, f, nullptr, r);
        if (rc) return rc;
        for (int i = 0; i < n; i++) {
            double fs[4];
            const double off[4] = {-2., -1., 1., 2.};
            for (int k = 0; k < 4; k++) {
                std::vector<double> y = x;
                y[i] += off[k] * diffstep;
                if (int e = value(y, &fs[k], nullptr, nullptr)) return e;
            }
            g[i] = (fs[0] - 8. * fs[1] + 8. * fs[2] - fs[3]) / (12. * diffstep);
        }
        grads++;
        return MPDPFT_OK;
    }
};

double dotv(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

}  // namespace

int run_cgd_task(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res,
                 std::ostringstream &log) {
    const int type = p.TaskType, n = p.S;
    if (p.n_hyperbox != 2 * n) {
        log << "TaskType " << type << " needs one TaskHyperBox pair per species (" << 2 * n << " values, got "
            << p.n_hyperbox << ")\n";
        return MPDPFT_EINVAL;
    }
    if (type == 2 && p.RelAcc > kMPt) {
        log << "TaskType 2 varies muVec: RelAcc must be 0 (no mu search)\n";
        return MPDPFT_EINVAL;
    }
    if (type == 3 && !(p.RelAcc > kMPt)) {
        log << "TaskType 3 varies the abundances: RelAcc must be positive (mu search on)\n";
        return MPDPFT_EINVAL;
    }
    std::vector<double> W(n), x(n);
    double scale = 0.;
    for (int h = 0; h < n; h++) {
        W[h] = p.TaskHyperBox[2 * h + 1] - p.TaskHyperBox[2 * h];
        if (!(W[h] > 0.)) {
            log << "TaskHyperBox pair " << h << " is empty\n";
            return MPDPFT_EINVAL;
        }
        x[h] = p.TaskHyperBox[2 * h] + 0.5 * W[h];
        if (type == 2 && std::fabs(x[h]) < 1e-10) x[h] += x[h] > -1e-10 ? 1e-10 : -1e-10;  // muThreshold (:8423)
        scale += W[h] / n;
    }
    double epsf = 1e-4, diffstep = 1e-6 * scale;  // System 0 (:8444-8451)
    if (p.System == 1) {
        epsf = 0.;
        diffstep = 0.1 * scale;
    } else if (p.System == 2) {
        diffstep = scale;
    }
    Objective ob;
    ob.dir = d;
    ob.p = p;
    ob.p.TaskType = 0;
    ob.device = device;
    ob.type = type;
    ob.diffstep = diffstep;
    ob.log = &log;
    log << std::setprecision(16);
    log << (type == 2 ? "GetGlobalMinimum muVec" : "GetGlobalMinimum Abundances") << "\n";
    log << "ConjugateGradientDescent 1: Start from " << vstr(x) << "\n";

    double f;
    std::vector<double> g, dvec(n), xn(n), gn;
    mpdpft_result r{};
    int rc = ob.eval(x, &f, g, &r);
    if (rc) return rc;
    const double c1 = 1e-4, c2 = 0.1;
    double t = 0., gd_prev = 0., t_prev = 0.;
    int it = 0, term = 5;
    std::string why = "200 steps were taken";
    for (int h = 0; h < n; h++) dvec[h] = -W[h] * W[h] * g[h];
    for (it = 0; it < 200; it++) {
        double gmax = 0.;
        for (int h = 0; h < n; h++) gmax = std::max(gmax, std::fabs(W[h] * g[h]));
        if (gmax <= 1e-9 * std::max(std::fabs(f), 1.)) {
            term = 4;
            why = "scaled gradient max_i |w_i g_i| <= 1e-9 max(|f|, 1)";
            break;
        }
        double gd = dotv(g, dvec);
        if (!(gd < 0.)) {  // not a descent direction: restart along the preconditioned gradient
            for (int h = 0; h < n; h++) dvec[h] = -W[h] * W[h] * g[h];
            gd = dotv(g, dvec);
        }
        double dmax = 0.;
        for (int h = 0; h < n; h++) dmax = std::max(dmax, std::fabs(dvec[h] / W[h]));
        if (it == 0) t = 0.1 / dmax;  // a tenth of the box along the largest coordinate
        else t = std::min(t_prev * gd_prev / gd, 0.5 / dmax);
        // line search on phi(t) = f(x + t d): bracket, then shrink by safeguarded secant steps
        double lo = 0., flo = f, dlo = gd, hi = -1., dhi = 0., fn = f;
        bool ok = false;
        for (int ls = 0; ls < 30; ls++) {
            for (int h = 0; h < n; h++) xn[h] = x[h] + t * dvec[h];
            rc = ob.eval(xn, &fn, gn, &r);
            if (rc) return rc;
            const double dn = dotv(gn, dvec);
            if (fn > f + c1 * t * gd || (ls > 0 && fn >= flo && hi < 0.)) {
                hi = t;
                dhi = dn;
            } else if (std::fabs(dn) <= c2 * std::fabs(gd)) {
                ok = true;
                break;
            } else if (dn > 0.) {
                hi = t;
                dhi = dn;
            } else {
                lo = t;
                flo = fn;
                dlo = dn;
            }
            if (hi < 0.) {
                t *= 2.;
                continue;
            }
            double ts = (dhi > 0. && dlo < 0.) ? lo + (hi - lo) * dlo / (dlo - dhi) : 0.5 * (lo + hi);
            const double m = 0.1 * (hi - lo);
            t = std::min(std::max(ts, lo + m), hi - m);
        }
        if (!ok) {
            if (!(fn < f)) {
                term = 7;
                why = "stopping conditions are too stringent, further improvement is impossible";
                break;
            }
        }
        log << "ConjugateGradientDescent: iteration " << it + 1 << ": f = " << fn << " @ {" << vstr(xn)
            << "}  (step " << t << ", solves so far " << ob.solves << ")\n";
        // Polak-Ribiere+ in the preconditioned metric
        double num = 0., den = 0.;
        for (int h = 0; h < n; h++) {
            num += gn[h] * W[h] * W[h] * (gn[h] - g[h]);
            den += g[h] * W[h] * W[h] * g[h];
        }
        const double beta = den > 0. ? std::max(0., num / den) : 0.;
        for (int h = 0; h < n; h++) dvec[h] = -W[h] * W[h] * gn[h] + beta * dvec[h];
        const double fold = f;
        x = xn;
        f = fn;
        g = gn;
        gd_prev = gd;
        t_prev = t;
        if (std::fabs(f - fold) <= epsf * std::max(std::max(std::fabs(fold), std::fabs(f)), 1.)) {
            it++;
            term = 1;
            std::ostringstream w;
            w << "relative function improvement is no more than EpsF = " << epsf;
            why = w.str();
            break;
        }
    }
    if (type == 3)
        for (int h = 0; h < n; h++)
            if (x[h] < 0.) {
                log << "ConjugateGradientDescent: x[" << h << "] = " << x[h] << "->0\n";
                x[h] = 0.;
            }
    // recompute the result at the minimum, with the files (lastQ)
    log << "ConjugateGradientDescent: recompute result from conjugate gradient descent\n";
    mpdpft_result fr{};
    rc = run_one(d, ob.at(x), device, &fr, true, "", nullptr);
    if (rc != MPDPFT_OK && rc != MPDPFT_ENOTCONV) return rc;
    log << "ConjugateGradientDescent: 1. Candidate for global minimum " << std::setprecision(10) << fr.TotalEnergy
        << " @ {" << vstr(x, 8) << "}\n"
        << "                         [ #Iterations = " << it << " #GradEvals = " << ob.grads
        << " #Solves = " << ob.solves + 1 << " #Gradient = " << (ob.adjoint && n == 1 ? "adjoint" : "4-point differences")
        << " #TerminationType = " << term << " (" << why << ") ]\n";
    log << std::setprecision(16) << "ConjugateGradientDescent: Emin = " << fr.TotalEnergy << " @ "
        << (type == 2 ? "muVec" : "Abundances") << " = " << vstr(x) << "\n";
    log << "ConjugateGradientDescent: gradient at the minimum = " << vstr(g) << "\n";
    if (res) *res = fr;
    return rc;
}

namespace {

// libstdc++'s std::uniform_real_distribution<double>(a, b) on std::mt19937_64:
// one 64-bit output x, u = double(x) / 2^64 (1 becomes the double below 1),
// a + (b - a) u -- spelled out so that the draws do not depend on the
// library's distribution object.
double draw(std::mt19937_64 &g, double a, double b) {
    double u = (double)g() / 18446744073709551616.0;
    if (u >= 1.) u = std::nextafter(1., 0.);
    return u * (b - a) + a;
}

double norm(const std::vector<double> &v) {
    double s = 0.;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}

}  // namespace

// TaskType 9 (GetGlobalMinimum_Abundances_AnnealedGradientDescent
// mpDPFT.cpp:604-608, GradientDescent :8586-8633, gradient_descent
// :8635-8710): the total energy's minimum over the abundances by
// Barzilai-Borwein gradient descent on forward differences (step gammaStart,
// one solve per abundance plus the base point) with random annealing kicks,
// restarted from uniform points of the TaskHyperBox until TaskParameterCount
// descents end at the same minimum (within 10 AbsoluteTargetAcc in f and in
// |x - xmin|^2).  The random stream is the reference's (std::mt19937_64, the
// start point's uniform draws and the kicks' RNpos, RN, RNpos in that order),
// seeded with NoiseSeed (the reference seeds it from the clock).  Where the
// reference reads ErrorHistory[-2] at the second iteration (the anneal step's
// comparison with a history of one entry) this treats the error as not grown.
// A cap of 10 TaskParameterCount + 10 descents ends a search whose count of
// matching descents steps past TaskParameterCount (the reference's test is an
// equality).  The files are those of a last solve at the final point.
int run_gd_task(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res,
                std::ostringstream &log) {
    log << std::setprecision(16);
    const int n = p.n_hyperbox / 2;
    if (n != p.S || n < 1) return MPDPFT_EINVAL;
    const double *box = p.TaskHyperBox;
    double scale = 0., avgN = 0.;
    for (int i = 0; i < n; i++) scale += (box[2 * i + 1] - box[2 * i]) / (double)n;
    for (int s = 0; s < p.S; s++) avgN += p.Abundances[s] / (double)p.S;
    scale = std::min(scale, avgN);
    const double acc = p.RelAcc > kMPt ? p.RelAcc * scale : 1.0e-4;
    double gammaStart = 1.0e-2 * scale, AnnealMagnitude = 3., fmin = 0.;
    const double AnnealThreshold = 0.7;
    const int maxIter = 200, maxDescents = 10 * std::max(1, p.TaskParameterCount) + 10;
    std::mt19937_64 rng;
    unsigned long long seed = p.NoiseSeed;
    if (seed == 0)
        seed = (unsigned long long)std::chrono::high_resolution_clock::now().time_since_epoch().count() % 2147483647ULL +
               1000000ULL;
    rng.seed(seed);
    std::vector<std::vector<double>> reports, starts;
    std::vector<double> x(n), xmin(n);
    int solves = 0;
    bool success = false, aborted = false;
    mpdpft_result last{};
    auto energy = [&](const std::vector<double> &at, double *f, mpdpft_result *r) -> int {
        mpdpft_params q = p;
        for (int s = 0; s < n; s++) q.Abundances[s] = at[s];  // data.Abundances = GradDesc.x (:6122)
        mpdpft_result rr{};
        int rc = run_one(d, q, device, &rr, false, "", nullptr);
        solves++;
        if (rc != MPDPFT_OK && rc != MPDPFT_ENOTCONV) return rc;
        *f = rr.TotalEnergy;
        if (r) *r = rr;
        return rr.SCcount >= p.maxSCcount ? 1 : 0;  // the solve hit maxSCcount: abort the descent
    };
    while (!success && (int)reports.size() < maxDescents) {
        for (int i = 0; i < n; i++) x[i] = box[2 * i] + draw(rng, 0., 1.) * (box[2 * i + 1] - box[2 * i]);
        starts.push_back(x);
        // gradient_descent
        double f = 0., gamma = gammaStart, Error = acc + 1., anneal = 0.;
        int iter = 0, successCount = 0, fokCount = 0;
        bool ok = false;
        std::vector<double> xOld(x), df(n, 0.), dfOld(n, 0.), fdiff(n), dfHist, fHist, errHist;
        std::vector<std::vector<double>> xHist;
        while (iter < maxIter && (Error > acc || !ok) && fokCount < 10) {
            xHist.push_back(x);
            if (iter > 0) xOld = xHist[xHist.size() - 2];
            dfOld = df;
            bool stop = false;
            for (int i = 0; i <= n && !stop; i++) {
                std::vector<double> at(x);
                if (i < n) at[i] += gammaStart;
                double fi = 0.;
                int rc = energy(at, &fi, i == n ? &last : nullptr);
                if (rc < 0) return rc;
                if (rc == 1) stop = true;
                else if (i == n) f = fi;
                else fdiff[i] = fi;
            }
            if (stop) {
                aborted = true;
                break;
            }
            for (int i = 0; i < n; i++) df[i] = (fdiff[i] - f) / gammaStart;
            dfHist.push_back(norm(df));
            fHist.push_back(f);
            const double I = std::min(10., (double)dfHist.size());
            double dfAv = 0., fAv = 0.;
            for (int k = 1; k <= (int)(I + 0.5); k++) {
                dfAv += dfHist[dfHist.size() - k] / I;
                fAv += fHist[fHist.size() - k] / I;
            }
            if (iter > 0) {
                if (std::fabs(f - fAv) < p.RelAcc * std::fabs(f + fAv)) fokCount++;
                else fokCount = 0;
                if (norm(df) < AnnealThreshold * dfAv) {
                    const double fp = fHist[fHist.size() - 2];
                    const double asym = (f == 0. && fp == 0.) ? 0. : (f - fp) / (f + fp);
                    const double r1 = draw(rng, 0., 1.), r2 = draw(rng, -1., 1.), r3 = draw(rng, 0., 1.);
                    const double heav = (0.5 + 0.5 * std::tanh(10. * asym)) / std::sqrt((double)iter) - r1 > 0. ? 1. : 0.;
                    anneal = heav * r2 * (AnnealMagnitude + std::pow(AnnealMagnitude * AnnealMagnitude, r3));
                    const bool grew = errHist.size() >= 2 && errHist[errHist.size() - 1] > errHist[errHist.size() - 2];
                    gammaStart *= grew ? 0.5 : 1.2;
                } else
                    anneal = 0.;
                double dxdf = 0., df2 = 0.;
                for (int i = 0; i < n; i++) {
                    dxdf += (x[i] - xOld[i]) * (df[i] - dfOld[i]);
                    df2 += (df[i] - dfOld[i]) * (df[i] - dfOld[i]);
                }
                gamma = df2 > 0. ? std::fabs(dxdf) / df2 : gammaStart;
                if (anneal != 0.) gamma *= anneal;
            }
            std::vector<double> xDiff(n), xNew(n);
            for (int i = 0; i < n; i++) xDiff[i] = df[i] * gamma;
            if (norm(xDiff) == 0.) break;
            for (int i = 0; i < n; i++) xNew[i] = std::min(std::max(x[i] - xDiff[i], box[2 * i]), box[2 * i + 1]);
            Error = norm(xDiff);
            errHist.push_back(Error);
            successCount = Error <= acc ? successCount + 1 : 0;
            if (successCount >= 3) ok = true;
            log << "gradient_descent: iter = " << iter + 1 << "/" << maxIter << ": f = " << f << "\n gamma = " << gamma
                << "(anneal=" << anneal << ") -> Error = " << Error << "(" << acc << ") ---> xNew = " << vstr(xNew)
                << " successCount = " << successCount << "/3 step = " << gammaStart << "\n";
            x = xNew;
            iter++;
        }
        std::vector<double> rep(x);
        rep.insert(rep.end(), df.begin(), df.end());
        rep.push_back(f);
        rep.push_back((double)iter);
        rep.push_back(Error);
        rep.push_back(gamma);
        reports.push_back(rep);
        if (iter == maxIter) AnnealMagnitude *= 0.7;
        else AnnealMagnitude /= 0.8;
        if (reports.size() == 1 || f < fmin) {
            fmin = f;
            xmin = x;
        }
        int enc = 0;
        for (size_t r = 0; r < reports.size(); r++) {
            double dot = 0., a2 = 0., b2 = 0.;
            for (int i = 0; i < n; i++) {
                dot += reports[r][i] * xmin[i];
                a2 += reports[r][i] * reports[r][i];
                b2 += xmin[i] * xmin[i];
            }
            if (std::fabs(reports[r][2 * n] - fmin) < 10. * acc && std::fabs(2. * dot - (a2 + b2)) < 10. * acc) enc++;
            log << "f = " << reports[r][2 * n] << " @ x = " << vstr(std::vector<double>(reports[r].begin(),
                                                                                      reports[r].begin() + n))
                << " [xStart = " << vstr(starts[r]) << "] <--report #" << r + 1 << "\n";
        }
        if (enc == p.TaskParameterCount) success = true;
        log << "GradientDescent: report " << reports.size() << " (x df f iter Error gamma)\n"
            << "                 fminEncounter   = " << enc << " fmin = " << fmin << " xmin = " << vstr(xmin) << "\n"
            << "                 AnnealMagnitude = " << AnnealMagnitude << "\n";
        if (aborted) break;
    }
    log << "                 Local minimum f = " << reports.back()[2 * n] << " @ x = " << vstr(x) << "\n";
    log << "GradientDescent: " << reports.size() << " descent(s), " << solves << " solves"
        << (success ? "" : (aborted ? " (a solve hit maxSCcount)" : " (descent cap reached)")) << "\n";
    // the final point's files
    mpdpft_params q = p;
    for (int s = 0; s < n; s++) q.Abundances[s] = x[s];
    mpdpft_result fr{};
    int rc = run_one(d, q, device, &fr, true, "", nullptr);
    if (res) *res = fr;
    return rc;
}
