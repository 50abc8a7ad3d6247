// gh_tmcmc.cpp -- the TMCMC sampler of the reference's inference
// (tools/pi4u_lite/Inference/engine_tmcmc.c, tmcmc_stats.c, auxil.c,
// posdef.c), restated as a host state machine that hands out whole
// generations of candidate points, so that a caller can evaluate them in one
// batched device solve (gh_batch) instead of one process per sample.
//
// Stage logic, as in the reference:
//   generation 0    PopSize points uniform in the bounds (engine_tmcmc.c:1133-1140)
//   statistics      next annealing exponent p_{j+1}: the x > p_j where the
//                   coefficient of variation of w_i = exp((F_i - F_max)(x - p_j))
//                   equals TolCOV (Objlogp, tmcmc_stats.c:18-50), by a 1-D
//                   Nelder-Mead from p_j with step 1e-4 (fminsearch, :315-408)
//                   and the grid scan fzerofind (:96-214) as fallback; accepted
//                   only if above p_j, else p_j + 1e-5; capped at 1 (:598-612).
//                   Weights exp(F_i (p_{j+1}-p_j) - max), log-selection
//                   log(sum w) + max - log n (:616-648), multinomial selection
//                   of PopSize leaders, weighted mean and covariance SS made
//                   positive definite (posdef.c force_pos_def2) (:655-760).
//   new chains      leaders sorted by selection count; counts above
//                   MaxChainLength split into chains of MaxChainLength, counts
//                   below MinChainLength raised (engine_tmcmc.c:754-1001).
//   chain step      candidate ~ N(leader, bbeta * SS) redrawn until inside the
//                   bounds; accepted with probability
//                   min(1, exp((prior_c - prior_l) + (F_c - F_l) p_j)); the
//                   leader (new or kept) enters the next generation (:661-738).
//   end             after the generation whose exponent is 1, or MaxStages;
//                   log evidence = sum of the log-selections (:1270-1280).
// F is the log-likelihood; the uniform prior makes prior_c - prior_l = 0.
//
// Random numbers come from this file's own generator (xoshiro256**, seeded
// through splitmix64): the reference draws from GSL, which is not part of
// this build, so sample paths differ from the reference's for the same seed.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "gh_internal.h"

namespace gh {
namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double normal() {  // Marsaglia's polar method
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        double u, v, q;
        do {
            u = 2 * uniform() - 1;
            v = 2 * uniform() - 1;
            q = u * u + v * v;
        } while (q >= 1 || q == 0);
        const double f = std::sqrt(-2 * std::log(q) / q);
        spare = v * f;
        has_spare = true;
        return u * f;
    }
    bool has_spare = false;
    double spare = 0;
};

double mean_of(const std::vector<double>& v) {
    double s = 0;
    for (double x : v) s += x;
    return s / (double)v.size();
}
// sample standard deviation, n - 1 (auxil.c compute_std)
double std_of(const std::vector<double>& v, double m) {
    double s = 0;
    for (double x : v) s += (x - m) * (x - m);
    return std::sqrt(s / (double)(v.size() - 1));
}

// (CoV(w) - tol)^2 with w_i = exp((f_i - max f)(x - p)) (Objlogp)
double objlogp(const std::vector<double>& f, double x, double p, double tol) {
    const double fmax = *std::max_element(f.begin(), f.end());
    std::vector<double> w(f.size());
    double sum = 0;
    for (size_t i = 0; i < f.size(); ++i) sum += (w[i] = std::exp((f[i] - fmax) * (x - p)));
    for (double& v : w) v /= sum;
    const double m = mean_of(w);
    const double c = std_of(w, m) / m - tol;
    return c * c;
}

// 1-D Nelder-Mead from x0 with initial step: reflection 1, expansion 2,
// contraction 1/2, shrink 1/2; converged when the simplex is smaller than tol
// or the value is within tol of 0 (fminsearch's tests), at most max_iter steps.
bool nelder_mead_1d(const std::vector<double>& f, double p, double tolcov, double x0, double step, double tol,
                    int max_iter, double* xmin, double* fmin) {
    double a = x0, b = x0 + step;
    double fa = objlogp(f, a, p, tolcov), fb = objlogp(f, b, p, tolcov);
    bool conv = false;
    for (int it = 0; it < max_iter; ++it) {
        if (fb < fa) {
            std::swap(a, b);
            std::swap(fa, fb);
        }  // a best
        const double xr = a + (a - b);
        const double fr = objlogp(f, xr, p, tolcov);
        if (fr < fa) {
            const double xe = a + 2 * (a - b);
            const double fe = objlogp(f, xe, p, tolcov);
            if (fe < fr) { b = xe; fb = fe; } else { b = xr; fb = fr; }
        } else if (fr < fb) {
            b = xr;
            fb = fr;
        } else {
            const double xc = a + 0.5 * (b - a);
            const double fc = objlogp(f, xc, p, tolcov);
            if (fc < fb) { b = xc; fb = fc; } else { b = a + 0.5 * (b - a); fb = objlogp(f, b, p, tolcov); }
        }
        const double best = std::min(fa, fb);
        if (std::fabs(a - b) < tol || std::fabs(best) <= tol) {
            conv = true;
            break;
        }
    }
    if (fb < fa) {
        std::swap(a, b);
        std::swap(fa, fb);
    }
    if (conv && std::fabs(fa) > tol) conv = false;  // "converged but not found minimum"
    *xmin = conv ? a : 0.0;
    *fmin = conv ? fa : 0.0;
    return conv;
}

// fzerofind (tmcmc_stats.c:96-214): scan [0, 4] with step 1e-4 for the first
// |f| <= tol, else refine around the best point with a 10x finer step.
bool fzerofind(const std::vector<double>& f, double p, double tolcov, double tol, double* xmin, double* fmin) {
    double lo = 0.0, hi = 4.0, step = 1e-4;
    while (true) {
        const size_t niters = (size_t)((hi - lo) / step);
        double m = 0, fm = DBL_MAX;
        for (size_t it = 0; it < niters; ++it) {
            const double x = lo + (double)it * step;
            const double fx = objlogp(f, x, p, tolcov);
            if (fx < fm) {
                fm = fx;
                m = x;
            }
            if (std::fabs(fx) <= tol) {
                *xmin = m;
                *fmin = fm;
                return true;
            }
        }
        lo = std::max(0.0, m - 10 * step);
        hi = std::min(4.0, m + 10 * step);
        step *= 0.1;
        if (step < 1e-16) return false;
    }
}

// eigenvalues of a symmetric matrix (cyclic Jacobi)
std::vector<double> sym_eigenvalues(std::vector<double> a, int n) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    std::vector<double> ev(n);
    for (int i = 0; i < n; ++i) ev[i] = a[i * n + i];
    return ev;
}

bool is_pos_def(const std::vector<double>& a, int n) {
    for (double e : sym_eigenvalues(a, n))
        if (!(e > 0)) return false;
    return true;
}

// force_pos_def2 (posdef.c:273-342): grow the diagonal until positive definite
void make_posdef(std::vector<double>& a, int n) {
    const double fixrate = 0.01, min_limit = 1e-15, convfix = 1 + fixrate;
    for (int loops = 0; loops < 100000 && !is_pos_def(a, n); ++loops) {
        bool below = false;
        double maxabs = std::fabs(a[0]), mn = a[0];
        for (int i = 0; i < n; ++i) {
            const double d = a[i * n + i];
            if (d <= min_limit) below = true;
            maxabs = std::max(maxabs, std::fabs(d));
            mn = std::min(mn, d);
        }
        double m = maxabs * fixrate;
        if (below) {
            double add;
            if (mn < 0) {
                add = m - mn;
            } else {
                if (m < min_limit) m = min_limit;
                add = m;
            }
            for (int i = 0; i < n; ++i) a[i * n + i] += add;
        } else {
            for (int i = 0; i < n; ++i) a[i * n + i] *= convfix;
        }
    }
}

// lower Cholesky factor of a (positive definite)
std::vector<double> cholesky(const std::vector<double>& a, int n) {
    std::vector<double> l(n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double d = a[j * n + j];
        for (int k = 0; k < j; ++k) d -= l[j * n + k] * l[j * n + k];
        d = std::sqrt(std::max(d, 0.0));
        l[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = a[i * n + j];
            for (int k = 0; k < j; ++k) s -= l[i * n + k] * l[j * n + k];
            l[i * n + j] = d > 0 ? s / d : 0.0;
        }
    }
    return l;
}

struct Entry {
    std::vector<double> x;
    double F;
};
struct Chain {
    std::vector<double> x;
    double F;
    int steps_left;
};

}  // namespace

struct Tmcmc {
    int nth, pop, max_stages, min_chain, max_chain;
    double tol_cov, bbeta, opt_tol = 1e-12;
    int opt_max_iter = 1000;
    std::vector<double> lo, hi;
    Rng rng;
    int gen = 0;
    bool done = false;
    std::vector<double> p{0.0}, cov{0.0}, logsel;
    std::vector<int> uniques;
    std::vector<double> SS, chol;
    std::vector<Entry> curgen, lastgen;
    std::vector<Chain> chains;
    std::vector<std::vector<double>> pending;  // candidates handed out
    std::vector<int> pending_chain;            // chain of each candidate (-1: prior sample)
    Tmcmc(int n_, uint64_t seed) : nth(n_), rng(seed) {}

    void start() {
        pending.clear();
        pending_chain.clear();
        for (int i = 0; i < pop; ++i) {
            std::vector<double> x(nth);
            for (int d = 0; d < nth; ++d) x[d] = lo[d] + (hi[d] - lo[d]) * rng.uniform();
            pending.push_back(x);
            pending_chain.push_back(-1);
        }
    }

    std::vector<double> candidate(const std::vector<double>& leader) {
        std::vector<double> z(nth), c(nth);
        while (true) {
            for (int d = 0; d < nth; ++d) z[d] = rng.normal();
            bool inside = true;
            for (int i = 0; i < nth; ++i) {
                double s = leader[i];
                for (int k = 0; k <= i; ++k) s += chol[i * nth + k] * z[k];
                c[i] = s;
                if (c[i] < lo[i] || c[i] > hi[i]) inside = false;
            }
            if (inside) return c;
        }
    }

    void next_round() {  // candidates of the chains that still have steps
        pending.clear();
        pending_chain.clear();
        for (size_t k = 0; k < chains.size(); ++k)
            if (chains[k].steps_left > 0) {
                pending.push_back(candidate(chains[k].x));
                pending_chain.push_back((int)k);
            }
    }

    // p_{j+1}, weights, selection, SS and the new chains (prepare_newgen)
    void statistics() {
        const int n = (int)curgen.size();
        // unique points (|dx| <= 1e-6 in every coordinate)
        int un = 0;
        std::vector<int> uidx;
        for (int i = 0; i < n; ++i) {
            bool uniq = true;
            for (int j : uidx) {
                bool same = true;
                for (int d = 0; d < nth && same; ++d)
                    if (std::fabs(curgen[i].x[d] - curgen[j].x[d]) > 1e-6) same = false;
                if (same) {
                    uniq = false;
                    break;
                }
            }
            if (uniq) {
                uidx.push_back(i);
                ++un;
            }
        }
        uniques.push_back(un);
        std::vector<double> f(n);
        for (int i = 0; i < n; ++i) f[i] = curgen[i].F;
        const double pj = p[gen];
        double xmin = 0, fmin = 0;
        bool conv = nelder_mead_1d(f, pj, tol_cov, pj, 1e-4, opt_tol, opt_max_iter, &xmin, &fmin);
        if (!conv) conv = fzerofind(f, pj, tol_cov, opt_tol, &xmin, &fmin);
        double pn, cn;
        if (conv && xmin > pj) {
            pn = xmin;
            cn = fmin;
        } else {
            pn = pj + 0.1 * 1e-4;
            cn = cov[gen];
        }
        if (pn > 1) pn = 1;
        p.push_back(pn);
        cov.push_back(cn);
        // weights and log-selection
        std::vector<double> fl(n), w(n), q(n);
        for (int i = 0; i < n; ++i) fl[i] = f[i] * (pn - pj);
        const double fmax = *std::max_element(fl.begin(), fl.end());
        double sw = 0;
        for (int i = 0; i < n; ++i) sw += (w[i] = std::exp(fl[i] - fmax));
        for (int i = 0; i < n; ++i) q[i] = w[i] / sw;
        logsel.push_back(std::log(sw) + fmax - std::log((double)n));
        const double mq = mean_of(q);
        cov[gen] = std_of(q, mq) / mq;
        // multinomial selection of pop leaders (inverse-CDF draws)
        std::vector<double> cdf(n);
        double acc = 0;
        for (int i = 0; i < n; ++i) cdf[i] = (acc += q[i]);
        std::vector<int> sel(n, 0);
        for (int k = 0; k < pop; ++k) {
            const double u = rng.uniform() * acc;
            int i = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            if (i >= n) i = n - 1;
            ++sel[i];
        }
        // weighted mean and covariance
        std::vector<double> mean(nth, 0.0);
        for (int d = 0; d < nth; ++d)
            for (int i = 0; i < n; ++i) mean[d] += curgen[i].x[d] * q[i];
        SS.assign(nth * nth, 0.0);
        for (int a = 0; a < nth; ++a)
            for (int b = a; b < nth; ++b) {
                double s = 0;
                for (int i = 0; i < n; ++i) s += q[i] * (curgen[i].x[a] - mean[a]) * (curgen[i].x[b] - mean[b]);
                SS[a * nth + b] = SS[b * nth + a] = s;
            }
        make_posdef(SS, nth);
        std::vector<double> bss(SS);
        for (double& v : bss) v *= bbeta;
        chol = cholesky(bss, nth);
        // chains: sorted by selection count, long ones split, short ones raised
        struct Sel {
            int idx, nsel;
        };
        std::vector<Sel> list;
        for (int i = 0; i < n; ++i)
            if (sel[i] > 0) list.push_back({i, sel[i]});
        std::stable_sort(list.begin(), list.end(), [](const Sel& a, const Sel& b) { return a.nsel > b.nsel; });
        if (max_chain > 0) {
            const size_t n0 = list.size();
            for (size_t i = 0; i < n0; ++i)
                while (list[i].nsel > max_chain) {
                    list.push_back({list[i].idx, max_chain});
                    list[i].nsel -= max_chain;
                }
            std::stable_sort(list.begin(), list.end(), [](const Sel& a, const Sel& b) { return a.nsel > b.nsel; });
        }
        for (Sel& s : list)
            if (s.nsel > 0 && s.nsel < min_chain) s.nsel = min_chain;
        chains.clear();
        for (const Sel& s : list) chains.push_back({curgen[s.idx].x, curgen[s.idx].F, s.nsel});
        lastgen = curgen;
        curgen.clear();
    }

    void submit(const double* F) {
        if (pending_chain.empty() || pending_chain[0] < 0) {  // generation 0
            for (size_t i = 0; i < pending.size(); ++i) curgen.push_back({pending[i], F[i]});
        } else {
            const double pj = p[gen];
            for (size_t i = 0; i < pending.size(); ++i) {
                Chain& ch = chains[pending_chain[i]];
                const double L = std::exp((F[i] - ch.F) * pj);  // uniform prior: equal prior terms
                const double u = rng.uniform();
                if (u < std::min(L, 1.0)) {
                    ch.x = pending[i];
                    ch.F = F[i];
                }
                curgen.push_back({ch.x, ch.F});
                --ch.steps_left;
            }
        }
        bool more = false;
        for (const Chain& c : chains)
            if (c.steps_left > 0) more = true;
        if (pending_chain.empty() || pending_chain[0] < 0 || !more) {
            // generation complete
            statistics();
            if (p[gen] == 1 || gen + 1 >= max_stages) {
                done = true;
                pending.clear();
                pending_chain.clear();
                return;
            }
            ++gen;
        }
        next_round();
    }
};

}  // namespace gh

using namespace gh;

namespace {
Tmcmc* as_tmcmc(gh_tmcmc* t) { return reinterpret_cast<Tmcmc*>(t); }
}

extern "C" {

int gh_tmcmc_create(int nth, const double* lo, const double* hi, int pop_size, int max_stages, int min_chain,
                    int max_chain, double tol_cov, double bbeta, uint64_t seed, gh_tmcmc** out) {
    if (!out) return set_error(GH_EINVAL, "gh_tmcmc_create: out is NULL");
    *out = nullptr;
    if (nth < 1 || nth > 64 || !lo || !hi || pop_size < 2 || max_stages < 1 || min_chain < 0 || max_chain < 0 ||
        !(tol_cov > 0) || !(bbeta > 0))
        return set_error(GH_EINVAL, "gh_tmcmc_create: bad configuration");
    for (int d = 0; d < nth; ++d)
        if (!(hi[d] > lo[d])) return set_error(GH_EINVAL, "gh_tmcmc_create: every bound needs hi > lo");
    std::unique_ptr<Tmcmc> t(new Tmcmc(nth, seed));
    t->pop = pop_size;
    t->max_stages = max_stages;
    t->min_chain = min_chain;
    t->max_chain = max_chain;
    t->tol_cov = tol_cov;
    t->bbeta = bbeta;
    t->lo.assign(lo, lo + nth);
    t->hi.assign(hi, hi + nth);
    t->start();
    *out = reinterpret_cast<gh_tmcmc*>(t.release());
    return GH_OK;
}

int gh_tmcmc_destroy(gh_tmcmc* th) {
    delete as_tmcmc(th);
    return GH_OK;
}

int gh_tmcmc_candidates(gh_tmcmc* th, int* n, double* points) {
    Tmcmc* t = as_tmcmc(th);
    if (!t || !n) return set_error(GH_EINVAL, "gh_tmcmc_candidates: NULL argument");
    *n = (int)t->pending.size();
    if (points)
        for (size_t i = 0; i < t->pending.size(); ++i)
            std::memcpy(points + i * t->nth, t->pending[i].data(), t->nth * sizeof(double));
    return GH_OK;
}

int gh_tmcmc_submit(gh_tmcmc* th, int n, const double* loglik) {
    Tmcmc* t = as_tmcmc(th);
    if (!t || !loglik) return set_error(GH_EINVAL, "gh_tmcmc_submit: NULL argument");
    if (t->done) return set_error(GH_ESTATE, "gh_tmcmc_submit: the run is finished");
    if (n != (int)t->pending.size()) return set_error(GH_EINVAL, "gh_tmcmc_submit: one value per candidate");
    std::vector<double> F(loglik, loglik + n);
    for (double& v : F)  // fitfun_glioma.c:198-207: a failed evaluation is a penalty, -1e12
        if (!std::isfinite(v)) v = -1.0e12;
    t->submit(F.data());
    return GH_OK;
}

int gh_tmcmc_state(gh_tmcmc* th, int* gen, double* p, double* cov, double* log_evidence, int* done) {
    Tmcmc* t = as_tmcmc(th);
    if (!t) return set_error(GH_EINVAL, "null sampler");
    if (gen) *gen = t->gen;
    if (p) *p = t->p[t->gen];
    if (cov) *cov = t->cov[t->gen];
    if (log_evidence) {
        double s = 0;
        for (double v : t->logsel) s += v;
        *log_evidence = s;
    }
    if (done) *done = t->done;
    return GH_OK;
}

int gh_tmcmc_population(gh_tmcmc* th, int* n, double* points, double* loglik) {
    Tmcmc* t = as_tmcmc(th);
    if (!t || !n) return set_error(GH_EINVAL, "gh_tmcmc_population: NULL argument");
    *n = (int)t->lastgen.size();
    for (size_t i = 0; i < t->lastgen.size(); ++i) {
        if (points) std::memcpy(points + i * t->nth, t->lastgen[i].x.data(), t->nth * sizeof(double));
        if (loglik) loglik[i] = t->lastgen[i].F;
    }
    return GH_OK;
}

int gh_tmcmc_next_exponent(int n, const double* loglik, double p, double tol_cov, double* p_next, double* cov) {
    if (n < 2 || !loglik || !p_next) return set_error(GH_EINVAL, "gh_tmcmc_next_exponent: bad argument");
    std::vector<double> f(loglik, loglik + n);
    double xmin = 0, fmin = 0;
    bool conv = nelder_mead_1d(f, p, tol_cov, p, 1e-4, 1e-12, 1000, &xmin, &fmin);
    if (!conv) conv = fzerofind(f, p, tol_cov, 1e-12, &xmin, &fmin);
    double pn = (conv && xmin > p) ? xmin : p + 1e-5;
    if (pn > 1) pn = 1;
    *p_next = pn;
    if (cov) {
        std::vector<double> w(n);
        const double fmax = *std::max_element(f.begin(), f.end());
        double s = 0;
        for (int i = 0; i < n; ++i) s += (w[i] = std::exp((f[i] - fmax) * (pn - p)));
        for (double& v : w) v /= s;
        const double m = mean_of(w);
        *cov = std_of(w, m) / m;
    }
    return GH_OK;
}

}  // extern "C"

// ---- the glioma fitness function's parameter map and file round trips ----
// fitfun_glioma.c:118-152 writes InputParameters.txt and LikelihoodInput.txt
// with "%.16lf" and TumorIC.txt with "%.4f"; brain reads the first two files
// into Real (Glioma_ReactionDiffusion.cpp:33-47, 279-289), likelihood reads
// doubles (HGG_Likelihood.cpp:13-22); Likelihood.txt holds "%Lf" of -logL
// and is read back with "%lf" (fitfun_glioma.c:190-207).
namespace {
double through_text(double v, const char* fmt, bool as_float) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), fmt, v);
    return as_float ? (double)std::strtof(buf, nullptr) : std::strtod(buf, nullptr);
}
}  // namespace

extern "C" {

int gh_glioma_params(int n, const double* x, int real_kind, double* sample, double* lik) {
    if (n < 0 || (n > 0 && (!x || !sample || !lik))) return set_error(GH_EINVAL, "gh_glioma_params: bad argument");
    const bool f32 = real_kind == GH_REAL_F32;
    for (int i = 0; i < n; ++i) {
        const double* xi = x + 11 * i;
        double prm[11];
        for (int k = 0; k < 11; ++k) prm[k] = std::exp(xi[k]);
        prm[6] = prm[6] * prm[6];          // PETsigma2
        prm[2] = std::sqrt(prm[2] / prm[1]);  // Tend = sqrt(c / rho)
        double* s = sample + 6 * i;
        for (int k = 0; k < 3; ++k) s[k] = through_text(prm[k], "%.16lf", f32);
        for (int k = 0; k < 3; ++k) s[3 + k] = through_text(prm[3 + k], "%.4f", f32);
        double* l = lik + 5 * i;
        for (int k = 0; k < 5; ++k) l[k] = through_text(prm[6 + k], "%.16lf", false);
    }
    return GH_OK;
}

int gh_glioma_fitness(int n, const double* neg_loglik, double* F) {
    if (n < 0 || (n > 0 && (!neg_loglik || !F))) return set_error(GH_EINVAL, "gh_glioma_fitness: bad argument");
    for (int i = 0; i < n; ++i) {
        double v = neg_loglik[i];
        if (std::isfinite(v)) {
            char buf[512];
            std::snprintf(buf, sizeof(buf), "%Lf", (long double)v);
            v = std::strtod(buf, nullptr);
        }
        if (!std::isfinite(v)) v = 1.0e12;
        F[i] = -v;
    }
    return GH_OK;
}

}  // extern "C"
