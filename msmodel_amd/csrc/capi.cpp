// capi.cpp -- the exported C ABI (include/msgpu.h): drop-in replacements with
// the reference signatures and the batched msgpu_* entry points.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "aberth.h"
#include "muller.h"
#include "capi_internal.h"

using namespace msg;
using cd = std::complex<double>;

namespace {

inline cx<double> to_cx(const msgpu_cplx& z) { return cx<double>{z.real(), z.imag()}; }
inline msgpu_cplx to_c(const cx<double>& z) { return msgpu_cplx(z.x, z.y); }
inline cx<double>* as_cx(msgpu_cplx* p) { return reinterpret_cast<cx<double>*>(p); }
inline const cx<double>* as_cx(const msgpu_cplx* p) { return reinterpret_cast<const cx<double>*>(p); }

// Drop-in entry points have no error return: a device failure is fatal, loudly.
void die_on(int rc, const char* where) {
    if (rc < 0) {
        fprintf(stderr, "[ERROR] %s: %s\n", where, last_error());
        fflush(stderr);
        exit(EXIT_FAILURE);
    }
}

}  // namespace

extern "C" {

// complex64 arithmetic for the modes that use it (MSGPU_FP64_LDLT computes in complex128)
static inline int reduced(int mode) { return mode == MSGPU_FP64 || mode == MSGPU_FP64_LDLT ? 0 : 1; }

// ------------------------------------------------------------------ batched API
MSGPU_API int msgpu_init(void) { return engine_init(); }
MSGPU_API int msgpu_autotune(int n, int mode, int nitems, double* report) {
    if (engine_init() != 0) return -1;
    return autotune(padded(n), reduced(mode), nitems, report);
}
MSGPU_API int msgpu_tune_save(const char* path) { return save_tuning(path); }
MSGPU_API int msgpu_tune_load(const char* path) { return load_tuning(path); }
MSGPU_API const char* msgpu_last_error(void) { return last_error(); }
MSGPU_API int msgpu_profile_begin(void) { return profile_begin(); }
MSGPU_API int msgpu_profile_end(double* ms, double* flops, double* bytes, int* count, int ncat) {
    return profile_end(ms, flops, bytes, count, ncat);
}

}  // extern "C"

namespace {

std::vector<MediumView> views_of(int nmed, const msgpu_medium* const* meds) {
    std::vector<MediumView> v;
    for (int m = 0; m < nmed; m++) v.push_back(view_of(meds[m]));
    return v;
}

// Trace-log of items (meds[midx[i]], ks[i]) in FP64, FP32 or MIXED.  MIXED: complex64 for
// every item, then complex128 for the items whose N*Re(tracelog) lies below their own
// medium's maximum + ln(reeval_rel) (|det| below reeval_rel of that medium's largest
// |det| in the batch: near zeros of det M, where the complex64 factor loses the most
// relative accuracy) and for non-finite complex64 results.  Returns #re-evaluated.
// MSGPU_FP64_LDLT: the pivot-free LDL^T, with the pivoted LU wherever its error estimate
// is large (or its result is not finite).  Two parts:
//  - Order.  The LDL^T factors M with its rows and columns in order of increasing
//    distance of the scatterers from their centroid (a symmetric permutation: det M is
//    unchanged).  Below the real axis the Green functions grow with distance,
//    e^{|Im k| r} / (4 pi r); eliminating the inner scatterers first keeps the factors
//    small.  On the d = 3, N = 1000 map over Re k in [5, 30] x Im k in [-2, 0] the largest
//    |L| fell from up to 1.6e8 in the medium's own order to at most 1.2e5.
//  - Gate.  The factorisation also returns rho = sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r|
//    (the diagonal of |L| |D| |L|^T over D: eps rho_r bounds the relative rounding error of
//    D_r).  A cell is kept when rho <= kLdltAccept (default 1e7, eps rho = 2.2e-9);
//    measured in numpy over 278 cells of that map, the kept cells agreed with LAPACK's
//    pivoted LU to 2.8e-10 in log det, and every error above 1e-10 stayed within
//    2 eps rho.  The largest |L| alone does not bound the error (|L| up to 1e10 with an
//    error of 6e-11 in a Gaussian medium; 6e3 with 1.1e-9 near Re k = 5).
// Cells whose growth in the medium's own order, max(e^{|Im k| D} / (4 pi D), 1) / |F^-1(k)|
// at the medium's diameter D, exceeds MSGPU_LDLT_PREDICT go to the LU without an attempt
// (default: no such bound); MSGPU_LDLT_ACCEPT overrides the bound on rho.
static double ldlt_bound(const char* var, double dflt) {
    const char* e = getenv(var);
    const double v = e ? atof(e) : 0.0;
    return v > 0.0 ? v : dflt;
}
static const double kLdltPredict = ldlt_bound("MSGPU_LDLT_PREDICT", HUGE_VAL);
static const double kLdltAccept = ldlt_bound("MSGPU_LDLT_ACCEPT", 1e7);

// The positions of v in order of increasing distance from their centroid (ties in input
// order), and the diameter bound 2 max |x_i - c|.
static std::vector<double> centre_out(const MediumView& v, double* diam) {
    double c[4] = {0, 0, 0, 0};
    for (int i = 0; i < v.n; i++)
        for (int q = 0; q < v.d; q++) c[q] += v.pos[(size_t)i * v.d + q] / v.n;
    std::vector<double> r2(v.n);
    double rmax = 0.0;
    for (int i = 0; i < v.n; i++) {
        double s = 0.0;
        for (int q = 0; q < v.d; q++) {
            const double e = v.pos[(size_t)i * v.d + q] - c[q];
            s += e * e;
        }
        r2[i] = s;
        rmax = std::max(rmax, s);
    }
    *diam = std::max(2.0 * std::sqrt(rmax), 1e-300);
    std::vector<int> idx(v.n);
    for (int i = 0; i < v.n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return r2[a] < r2[b]; });
    std::vector<double> out((size_t)v.n * v.d);
    for (int i = 0; i < v.n; i++)
        for (int q = 0; q < v.d; q++) out[(size_t)i * v.d + q] = v.pos[(size_t)idx[i] * v.d + q];
    return out;
}

// The media in centre-out order (views into opos), for the LDL^T.
static std::vector<MediumView> centre_out_views(const std::vector<MediumView>& mv,
                                                std::vector<std::vector<double>>& opos, std::vector<double>& diam) {
    std::vector<MediumView> ov(mv);
    opos.resize(mv.size());
    diam.resize(mv.size());
    for (size_t m = 0; m < mv.size(); m++) {
        opos[m] = centre_out(mv[m], &diam[m]);
        ov[m].pos = opos[m].data();
        ov[m].dpos = nullptr;
    }
    return ov;
}

static int tracelog_ldlt(const std::vector<MediumView>& mv, int nk, const int* midx, const msgpu_cplx* ks,
                         msgpu_cplx* out, int* flags) {
    if (flags && nk > 0) memset(flags, 0, sizeof(int) * nk);
    if (nk <= 0) return 0;
    std::vector<double> diam;
    std::vector<std::vector<double>> opos;
    const std::vector<MediumView> ov = centre_out_views(mv, opos, diam);
    std::vector<int> sym, lu;
    for (int l = 0; l < nk; l++) {
        const int m = midx ? midx[l] : 0;
        const cx<double> k = to_cx(ks[l]);
        bool attempt = true;
        if (kLdltPredict < HUGE_VAL) {
            const double a = std::max(0.0, -k.y), D = diam[m];
            const cx<double> f = invf(mv[m].m, mv[m].d, k);
            attempt = std::max(std::exp(a * D) / (4.0 * kPi * D), 1.0) / std::hypot(f.x, f.y) <= kLdltPredict;
        }
        (attempt ? sym : lu).push_back(l);
    }
    auto run = [&](const std::vector<int>& cells, int precision, std::vector<msgpu_cplx>& res,
                   std::vector<double>* st) -> int {
        std::vector<msgpu_cplx> sub(cells.size());
        std::vector<int> sm(cells.size());
        for (size_t i = 0; i < cells.size(); i++) {
            sub[i] = ks[cells[i]];
            sm[i] = midx ? midx[cells[i]] : 0;
        }
        res.resize(cells.size());
        if (st) st->resize(2 * cells.size());
        return tracelog_multi(precision == 2 ? ov : mv, (int)cells.size(), sm.data(), as_cx(sub.data()),
                              as_cx(res.data()), st ? st->data() : nullptr, precision);
    };
    if (!sym.empty()) {
        std::vector<msgpu_cplx> res;
        std::vector<double> st;
        if (int rc = run(sym, 2, res, &st); rc < 0) return rc;
        for (size_t i = 0; i < sym.size(); i++) {
            const cx<double> t = to_cx(res[i]);
            if (st[2 * i + 1] <= kLdltAccept && std::isfinite(t.x) && std::isfinite(t.y)) out[sym[i]] = res[i];
            else lu.push_back(sym[i]);
        }
    }
    if (!lu.empty()) {
        std::sort(lu.begin(), lu.end());
        std::vector<msgpu_cplx> res;
        if (int rc = run(lu, 0, res, nullptr); rc < 0) return rc;
        for (size_t i = 0; i < lu.size(); i++) {
            out[lu[i]] = res[i];
            if (flags) flags[lu[i]] = 1;
        }
    }
    return (int)lu.size();
}

int tracelog_any(const std::vector<MediumView>& mv, int nk, const int* midx, const msgpu_cplx* ks, msgpu_cplx* out,
                 int mode, double reeval_rel, int* flags) {
    if (mode == MSGPU_FP64_LDLT) return tracelog_ldlt(mv, nk, midx, ks, out, flags);
    if (mode == MSGPU_FP64 || mode == MSGPU_FP32) {
        if (flags && nk > 0) memset(flags, 0, sizeof(int) * nk);
        return tracelog_multi(mv, nk, midx, as_cx(ks), as_cx(out), nullptr, mode == MSGPU_FP32 ? 1 : 0);
    }
    if (mode != MSGPU_MIXED && mode != MSGPU_MIXED_GATED) { set_error("msgpu: unknown mode"); return -4; }
    if (nk <= 0) return 0;
    const bool gated = mode == MSGPU_MIXED_GATED;
    std::vector<double> st(gated ? 2 * (size_t)nk : 0);
    int rc = tracelog_multi(mv, nk, midx, as_cx(ks), as_cx(out), gated ? st.data() : nullptr, 1);
    if (rc < 0) return rc;
    std::vector<int> idx, sub_m;
    if (flags) memset(flags, 0, sizeof(int) * nk);
    if (gated) {        // the pivot-growth gate, cell by cell
        for (int l : gated_cells(nk, st.data(), out, reeval_rel)) {
            if (flags) flags[l] = 1;
            idx.push_back(l);
            sub_m.push_back(midx ? midx[l] : 0);
        }
    } else {            // per medium, the cells below reeval_rel of its largest |det|
        std::vector<std::vector<int>> per(mv.size());
        for (int l = 0; l < nk; l++) per[midx ? midx[l] : 0].push_back(l);
        for (size_t m = 0; m < mv.size(); m++) {
            if (per[m].empty()) continue;
            std::vector<msgpu_cplx> t(per[m].size());
            for (size_t q = 0; q < per[m].size(); q++) t[q] = out[per[m][q]];
            for (int q : reeval_cells(mv[m].n, (int)t.size(), t.data(), reeval_rel)) {
                const int l = per[m][q];
                if (flags) flags[l] = 1;
                idx.push_back(l);
                sub_m.push_back((int)m);
            }
        }
    }
    if (!idx.empty()) {
        std::vector<msgpu_cplx> sub(idx.size()), res(idx.size());
        for (size_t i = 0; i < idx.size(); i++) sub[i] = ks[idx[i]];
        rc = tracelog_multi(mv, (int)idx.size(), sub_m.data(), as_cx(sub.data()), as_cx(res.data()), nullptr, 0);
        if (rc < 0) return rc;
        for (size_t i = 0; i < idx.size(); i++) out[idx[i]] = res[i];
    }
    return (int)idx.size();
}

// Maehly-Aberth-Ehrlich on det M(k) for several media at once (root_finder.c:522-577,
// stepped by AberthRun): the background fits of all media run as one batched FP64
// call, then every sweep evaluates f'/f for the live roots of all media in one batched
// call per precision.
int resonances_any(const std::vector<MediumView>& mv, const double* doms, int* nz, msgpu_cplx* z, int nexpsup,
                   int maxit, double toler, int mode) {
    const int nmed = (int)mv.size();
    std::vector<int> offs(nmed, 0);
    for (int m = 0, o = 0; m < nmed; m++) { offs[m] = o; o += nz[m]; }
    std::vector<AberthRun> runs;
    runs.reserve(nmed);
    for (int m = 0; m < nmed; m++)
        runs.emplace_back(z + offs[m], nz[m], doms[4 * m], doms[4 * m + 1], doms[4 * m + 2], doms[4 * m + 3],
                          nexpsup, maxit, toler, reduced(mode) != 0);
    // background fits
    {
        std::vector<msgpu_cplx> zs, fs;
        std::vector<int> ms;
        for (int m = 0; m < nmed; m++)
            for (const cd& c : runs[m].fit_points()) { zs.push_back(c); ms.push_back(m); }
        if (!zs.empty()) {
            fs.resize(zs.size());
            const int rc = dk_log_det_multi(mv, (int)zs.size(), ms.data(), as_cx(zs.data()), as_cx(fs.data()), 0);
            if (rc < 0) return rc;
            for (size_t q = 0; q < zs.size();) {
                const int m = ms[q];
                std::vector<cd> f;
                while (q < zs.size() && ms[q] == m) f.push_back(fs[q++]);
                runs[m].take_fit(f);
            }
        }
    }
    int sweeps = 0;
    for (;;) {
        std::vector<msgpu_cplx> zq[2], fq[2];
        std::vector<int> mq[2];
        for (int m = 0; m < nmed; m++) {
            int p = 0;
            for (const cd& c : runs[m].sweep_points(&p)) { zq[p].push_back(c); mq[p].push_back(m); }
        }
        if (zq[0].empty() && zq[1].empty()) break;
        for (int p = 0; p < 2; p++) {
            if (zq[p].empty()) continue;
            fq[p].resize(zq[p].size());
            const int rc =
                dk_log_det_multi(mv, (int)zq[p].size(), mq[p].data(), as_cx(zq[p].data()), as_cx(fq[p].data()), p);
            if (rc < 0) return rc;
        }
        for (int p = 0; p < 2; p++)
            for (size_t q = 0; q < zq[p].size();) {
                const int m = mq[p][q];
                std::vector<cd> f;
                while (q < zq[p].size() && mq[p][q] == m) f.push_back(fq[p][q++]);
                runs[m].take_sweep(f);
            }
        sweeps++;
    }
    for (int m = 0; m < nmed; m++) {
        const std::vector<cd> r = runs[m].roots();
        nz[m] = (int)r.size();
        for (size_t i = 0; i < r.size(); i++) z[offs[m] + i] = r[i];
    }
    return sweeps;
}

}  // namespace

extern "C" {

MSGPU_API int msgpu_tracelog_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, int mode,
                                   double reeval_rel, int* flags) {
    return tracelog_any(views_of(1, &med), nk, nullptr, ks, out, mode, reeval_rel, flags);
}

MSGPU_API int msgpu_tracelog_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of,
                                   const msgpu_cplx* ks, msgpu_cplx* out, int mode, double reeval_rel, int* flags) {
    if (nmed <= 0 || !meds) { set_error("msgpu: no medium given"); return -2; }
    return tracelog_any(views_of(nmed, meds), nk, med_of, ks, out, mode, reeval_rel, flags);
}

MSGPU_API int msgpu_dk_log_det_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of,
                                     const msgpu_cplx* ks, msgpu_cplx* out, int mode) {
    if (nmed <= 0 || !meds) { set_error("msgpu: no medium given"); return -2; }
    return dk_log_det_multi(views_of(nmed, meds), nk, med_of, as_cx(ks), as_cx(out), reduced(mode));
}

MSGPU_API int msgpu_resonances_multi(int nmed, const msgpu_medium* const* meds, const double* doms, int* nz,
                                     msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode) {
    if (nmed <= 0 || !meds) { set_error("msgpu: no medium given"); return -2; }
    return resonances_any(views_of(nmed, meds), doms, nz, z, nexpsup, maxit, toler, mode);
}

MSGPU_API int msgpu_tracelog_stats(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out,
                                   double* stats, int mode) {
    if (mode == MSGPU_FP64_LDLT) {      // every cell by the pivot-free LDL^T, no fallback
        std::vector<double> diam;
        std::vector<std::vector<double>> opos;
        const std::vector<MediumView> ov = centre_out_views({view_of(med)}, opos, diam);
        return tracelog_multi(ov, nk, nullptr, as_cx(ks), as_cx(out), stats, 2);
    }
    return tracelog_points(view_of(med), nk, as_cx(ks), as_cx(out), stats, reduced(mode));
}

MSGPU_API int msgpu_dk_log_det_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out,
                                     int mode) {
    return dk_log_det_points(view_of(med), nk, as_cx(ks), as_cx(out), reduced(mode));
}

MSGPU_API int msgpu_solve(const msgpu_medium* med, const msgpu_iwave* iwave, msgpu_cplx* a, int mode, double tol) {
    return solve_amplitudes(view_of(med), to_cx(iwave->k), iwave->source, as_cx(a), reduced(mode),
                            tol);
}

MSGPU_API int msgpu_wavefunction_grid(const msgpu_medium* med, const msgpu_iwave* iwave, const msgpu_cplx* a, int nx,
                                      int ny, double xmin, double xmax, double ymin, double ymax, msgpu_cplx* psi) {
    return wave_grid(view_of(med), to_cx(iwave->k), iwave->source, as_cx(a), nx, ny, xmin, xmax, ymin, ymax,
                     as_cx(psi));
}

// Random functions and cross sections of many (medium, real k) items in one batch:
// run_random_function_omp (ms_core.c:476-519: total cross section, or the density of
// states -Im tr(M^-1 M')/(2 pi k)) and run_differential_cross_section_omp
// (ms_core.c:526-587: prefac |b_theta^H a|^2 per angle).
MSGPU_API int msgpu_scattering_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of,
                                     const double* ks, int type, int nang, const double* theta_deg, double* out) {
    if (nk < 0 || (nk > 0 && (!ks || !out)) || (type == MSGPU_DCS && (nang < 0 || (nang > 0 && !theta_deg)))) {
        set_error("msgpu_scattering_multi: bad arguments");
        return -2;
    }
    if (type != MSGPU_CRSEC && type != MSGPU_DOS && type != MSGPU_DCS) {
        set_error("msgpu_scattering_multi: unknown type");
        return -2;
    }
    if (nk == 0) return 0;
    const std::vector<MediumView> mv = views_of(nmed, meds);
    std::vector<msgpu_cplx> kc(nk);
    for (int i = 0; i < nk; i++) kc[i] = msgpu_cplx(ks[i], 0.0);
    if (type == MSGPU_DOS) {
        std::vector<msgpu_cplx> dk(nk);
        if (int rc = dk_log_det_multi(mv, nk, med_of, as_cx(kc.data()), as_cx(dk.data()), 0)) return rc;
        for (int i = 0; i < nk; i++) out[i] = -dk[i].imag() / (2 * ks[i] * kPi);
        return 0;
    }
    size_t tot = 0;
    for (int i = 0; i < nk; i++) {
        const int m = med_of ? med_of[i] : 0;
        if (m < 0 || m >= nmed) { set_error("msgpu: medium index out of range"); return -2; }
        tot += mv[m].n;
    }
    std::vector<msgpu_cplx> a(tot);
    if (int rc = amplitudes_multi(mv, nk, med_of, as_cx(kc.data()), as_cx(a.data()))) return rc;
    size_t off = 0;
    for (int i = 0; i < nk; i++) {
        const MediumView& m = mv[med_of ? med_of[i] : 0];
        const double k = ks[i];
        const msgpu_cplx* ai = a.data() + off;
        off += m.n;
        if (type == MSGPU_CRSEC) {      // -Im(phi^H a)/k, phi_j = e^{i k x_j0} (ms_matrix_util.c:344-366)
            msgpu_cplx dot = 0.0;
            for (int j = 0; j < m.n; j++) dot += std::conj(std::exp(msgpu_cplx(0.0, k * m.pos[(size_t)j * m.d]))) * ai[j];
            out[i] = -dot.imag() / k;
            continue;
        }
        // prefac |b^H a|^2, b the plane wave of angle theta (build_plane_wave, ms_matrix_util.c:88-106)
        const double prefac = std::pow(k / (2 * kPi), m.d - 3) / (16 * kPi * kPi);
        for (int t = 0; t < nang; t++) {
            const double th = theta_deg[t] * kPi / 180;
            double ct = std::cos(th), st = std::sin(th);
            if (m.d == 1) ct = ct > 0.0 ? 1.0 : -1.0;
            msgpu_cplx dot = 0.0;
            for (int j = 0; j < m.n; j++) {
                const double* x = m.pos + (size_t)j * m.d;
                const double ph = m.d == 1 ? ct * x[0] : ct * x[0] + st * x[1];
                dot += std::conj(std::exp(msgpu_cplx(0.0, 1.0) * msgpu_cplx(k, 0.0) * ph)) * ai[j];
            }
            out[(size_t)i * nang + t] = prefac * (dot.real() * dot.real() + dot.imag() * dot.imag());
        }
    }
    return 0;
}

// Roots of det M(k) by the damped Muller method on another function of M(k), all starts
// together (find_eigenstates_serial_2/_3, ms_core.c:335-401): every round evaluates f
// at every pending point of every live start in one batch.
//   MSGPU_ROOT_INVTRACEINV: f(k) = 1/tr(M(k)^-1)
//   MSGPU_ROOT_MINEIGVAL:   f(k) = smallest eigenvalue of M(k) by inverse iteration (50
//     steps, 1e-10) from the start's previous eigenvector (the constant unit vector at
//     first); a converged root must also have |f| < sqrt(toler) there.
// Then non-converged starts are dropped (filter_cvector) and duplicates within 1e-14
// relative (uniq_cvector, complex_vector_util.c:130-177).
MSGPU_API int msgpu_roots_muller(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                                 msgpu_cplx* z, int maxit, double toler, int method) {
    if (!med || !nz || *nz < 0 || (*nz > 0 && !z)) { set_error("msgpu_roots_muller: bad arguments"); return -2; }
    if (method != MSGPU_ROOT_MINEIGVAL && method != MSGPU_ROOT_INVTRACEINV) {
        set_error("msgpu_roots_muller: unknown method");
        return -2;
    }
    const MediumView mv = view_of(med);
    const int nst = *nz, n = mv.n;
    const bool mineig = method == MSGPU_ROOT_MINEIGVAL;
    std::vector<MullerRun> runs;
    for (int i = 0; i < nst; i++) runs.emplace_back(z[i], xmin, xmax, ymin, ymax, maxit, toler);
    // per start, its current eigenvector (mineigval); the constant unit vector first
    std::vector<msgpu_cplx> vec(mineig ? (size_t)nst * n : 0, msgpu_cplx(1.0 / std::sqrt((double)n), 0.0));
    // f at points pts (start owner[j]); eigenvectors of the last point of each start kept
    auto evaluate = [&](const std::vector<msgpu_cplx>& pts, const std::vector<int>& owner,
                        std::vector<msgpu_cplx>& f) -> int {
        const int np_ = (int)pts.size();
        f.assign(np_, 0.0);
        if (!mineig) {
            std::vector<msgpu_cplx> tr(np_);
            if (int rc = trace_inv_multi({mv}, np_, nullptr, as_cx(pts.data()), as_cx(tr.data()), 0)) return rc;
            for (int j = 0; j < np_; j++) f[j] = 1.0 / tr[j];
            return 0;
        }
        std::vector<msgpu_cplx> vin((size_t)np_ * n), vout((size_t)np_ * n);
        std::vector<int> conv(np_);
        for (int j = 0; j < np_; j++)
            std::copy(vec.begin() + (size_t)owner[j] * n, vec.begin() + (size_t)(owner[j] + 1) * n,
                      vin.begin() + (size_t)j * n);
        if (int rc = eigenstates(mv, np_, as_cx(pts.data()), 50, 1e-10, as_cx(vout.data()), as_cx(f.data()),
                                 conv.data(), as_cx(vin.data())))
            return rc;
        for (int j = 0; j < np_; j++)
            std::copy(vout.begin() + (size_t)j * n, vout.begin() + (size_t)(j + 1) * n,
                      vec.begin() + (size_t)owner[j] * n);
        return 0;
    };
    int rounds = 0;
    for (;;) {
        std::vector<msgpu_cplx> pts;
        std::vector<int> owner, first(nst, -1), count(nst, 0);
        for (int i = 0; i < nst; i++) {
            if (runs[i].finished()) continue;
            const std::vector<std::complex<double>> p = runs[i].points();
            first[i] = (int)pts.size();
            count[i] = (int)p.size();
            for (const auto& q : p) { pts.push_back(q); owner.push_back(i); }
        }
        if (pts.empty()) break;
        std::vector<msgpu_cplx> f;
        if (int rc = evaluate(pts, owner, f)) return rc;
        for (int i = 0; i < nst; i++)
            if (count[i]) runs[i].take(std::vector<std::complex<double>>(f.begin() + first[i], f.begin() + first[i] + count[i]));
        rounds++;
    }
    std::vector<int> keep(nst);
    for (int i = 0; i < nst; i++) keep[i] = runs[i].converged() ? 1 : 0;
    if (mineig) {       // |mineigval(root)| < sqrt(toler) (ms_core.c:355)
        std::vector<msgpu_cplx> pts;
        std::vector<int> owner;
        for (int i = 0; i < nst; i++)
            if (keep[i]) { pts.push_back(runs[i].root()); owner.push_back(i); }
        std::vector<msgpu_cplx> f;
        if (!pts.empty()) {
            if (int rc = evaluate(pts, owner, f)) return rc;
            for (size_t j = 0; j < pts.size(); j++)
                if (!(std::abs(f[j]) < std::sqrt(toler))) keep[owner[j]] = 0;
        }
    }
    std::vector<msgpu_cplx> roots;
    for (int i = 0; i < nst; i++) {
        if (!keep[i]) continue;
        const msgpu_cplx r = runs[i].root();
        bool uniq = true;
        for (const msgpu_cplx& o : roots)
            if (std::abs(r - o) < 1e-14 * std::abs(r) || std::abs(r - o) < 1e-14) { uniq = false; break; }
        if (uniq) roots.push_back(r);
    }
    std::copy(roots.begin(), roots.end(), z);
    *nz = (int)roots.size();
    return rounds;
}

// Eigenstates at resonances (compute_eigenstates, ms_core.c:296-313) and the checks of
// check_eigenstate_list (eigenstate_list_util.c:130-163): a state is suspect if
// |mu| > 1e-6 or an earlier root lies within 1e-6 |k_r| of it.
MSGPU_API int msgpu_eigenstates(const msgpu_medium* med, int nroot, const msgpu_cplx* kroot, int maxit, double toler,
                                msgpu_cplx* vecs, msgpu_cplx* mu, int* conv, int* suspect) {
    if (nroot < 0 || (nroot > 0 && (!kroot || !vecs || !mu || !conv))) {
        set_error("msgpu_eigenstates: bad arguments");
        return -2;
    }
    if (int rc = eigenstates(view_of(med), nroot, as_cx(kroot), maxit, toler, as_cx(vecs), as_cx(mu), conv)) return rc;
    constexpr double tolmu = 1e-6;
    int nsuspect = 0;
    for (int i = 0; i < nroot; i++) {
        bool s = std::abs(mu[i]) > tolmu;
        for (int j = 0; j < i; j++)
            if (std::abs(kroot[i] - kroot[j]) < tolmu * std::abs(kroot[i])) s = true;
        if (suspect) suspect[i] = s ? 1 : 0;
        nsuspect += s ? 1 : 0;
    }
    return nsuspect;
}

// Aberth-Ehrlich on det M(k) of one medium (root_finder.c:522-577); see resonances_any.
MSGPU_API int msgpu_resonances(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                               msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode) {
    const double dom[4] = {xmin, xmax, ymin, ymax};
    return resonances_any(views_of(1, &med), dom, nz, z, nexpsup, maxit, toler, mode);
}

// ------------------------------------------------------------------ k-point generators
// The sample points of the reference's k-plane tasks, in their storage order, so that a
// caller can hand a whole map, cut or log cut to the batched entry points.  The
// arithmetic is that of the reference's generators, operation for operation (the
// results are equal bit for bit).
// get_arg_cmap (complex_map_util.c:27-31, 90-91): row-major from the top-left corner.
MSGPU_API int msgpu_kgrid_map(int nx, int ny, double xmin, double xmax, double ymin, double ymax, msgpu_cplx* k) {
    if (nx < 2 || ny < 2 || !k) { set_error("msgpu_kgrid_map: nx, ny >= 2 and an output array needed"); return -2; }
    const double dx = (xmax - xmin) / (nx - 1), dy = (ymax - ymin) / (ny - 1);
    for (int l = 0; l < nx * ny; l++) {
        const int ix = l % nx, iy = l / nx;
        k[l] = msgpu_cplx(xmin + ix * dx, ymax - iy * dy);
    }
    return nx * ny;
}

// get_arg_ccut (complex_cut_util.c:32-36, 68): 3 (ns + 2) points, triplet-major, each
// triplet across the cut (offsets -i dz, 0, +i dz with the complex step dz).
MSGPU_API int msgpu_kgrid_cut(int ns, msgpu_cplx zmin, msgpu_cplx zmax, msgpu_cplx* k) {
    if (ns < 2 || !k) { set_error("msgpu_kgrid_cut: ns >= 2 and an output array needed"); return -2; }
    const double dzr = (zmax.real() - zmin.real()) / (ns - 1), dzi = (zmax.imag() - zmin.imag()) / (ns - 1);
    for (int l = 0; l < 3 * (ns + 2); l++) {
        const int il = l / 3, it = l % 3;
        const double re = ((ns - il) * zmin.real() + (il - 1) * zmax.real()) / (ns - 1);
        const double im = ((ns - il) * zmin.imag() + (il - 1) * zmax.imag()) / (ns - 1);
        k[l] = msgpu_cplx(re + (-dzi) * (it - 1), im + dzr * (it - 1));      // + i dz (it - 1)
    }
    return 3 * (ns + 2);
}

// init_logcut (log_cut_util.c:72-101): rek + i imkmin (imkmax/imkmin)^(i/(ns-1)), computed as
// imkmin exp(i log(imkmax/imkmin) / (ns - 1)).
MSGPU_API int msgpu_kgrid_logcut(int ns, double rek, double imkmin, double imkmax, msgpu_cplx* k) {
    if (ns < 2 || !k || imkmin >= 0 || imkmax >= 0) {
        set_error("msgpu_kgrid_logcut: ns >= 2, negative end points and an output array needed");
        return -2;
    }
    const double lfac = std::log(imkmax / imkmin);
    for (int i = 0; i < ns; i++) k[i] = msgpu_cplx(rek, imkmin * std::exp((lfac * i) / (ns - 1)));
    return ns;
}

// ------------------------------------------------------------------ drop-in: ms_matrix_util.h
MSGPU_API void build_ms_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix) {
    die_on(build_matrix_host(view_of(med), to_cx(k), 0, as_cx(matrix)), "build_ms_matrix");
}

MSGPU_API void build_normalized_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix) {
    die_on(build_matrix_host(view_of(med), to_cx(k), 2, as_cx(matrix)), "build_normalized_matrix");
}

MSGPU_API void build_dk_ms_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix) {
    die_on(build_matrix_host(view_of(med), to_cx(k), 1, as_cx(matrix)), "build_dk_ms_matrix");
}

MSGPU_API void build_plane_wave(msgpu_medium* med, msgpu_cplx k, double theta, msgpu_cplx* vec) {
    die_on(plane_wave_host(view_of(med), to_cx(k), theta, as_cx(vec)), "build_plane_wave");
}

MSGPU_API msgpu_cplx tracelog_lu(int n, msgpu_cplx* matrix) {
    std::vector<int> ipiv(n);
    cx<double> tl{0, 0};
    int info = 0;
    // concurrent calls of the caller's threads share one batched factorization
    die_on(tracelog_host_coalesced(n, as_cx(matrix), n, ipiv.data(), &tl, &info), "tracelog_lu");
    if (info > 0) printf("[LAPACK] Element U(%d,%d) is exactly zero, the determinant will be singular.\n", info, info);
    return to_c(tl);
}

MSGPU_API msgpu_cplx dk_log_det_ms(msgpu_medium* med, msgpu_cplx k) {
    msgpu_cplx out;
    die_on(dk_log_det_points(view_of(med), 1, as_cx(&k), as_cx(&out), 0), "dk_log_det_ms");
    return out;
}

MSGPU_API msgpu_cplx traceinv(int n, msgpu_cplx* a) {
    std::vector<int> ipiv(n);
    std::vector<msgpu_cplx> e((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) e[i + (size_t)i * n] = 1.0;
    int info = 0;
    die_on(lu_host(n, as_cx(a), n, ipiv.data(), n, as_cx(e.data()), n, nullptr, &info), "traceinv");
    if (info > 0) printf("[LAPACK] Element U(%d,%d) is exactly zero, the inverse will be singular.\n", info, info);
    msgpu_cplx tr = 0.0;
    for (int i = 0; i < n; i++) tr += e[i + (size_t)i * n];
    return tr;
}

static msgpu_cplx normalize_vec(int n, msgpu_cplx* x) {   // cnormalize (complex_vector_util.c:47-60)
    int i = 0;
    while (i < n && x[i] == 0.0) i++;
    if (i == n) return 1.0;
    double s = 0;
    for (int j = 0; j < n; j++) s += std::norm(x[j]);
    const msgpu_cplx fac = std::sqrt(s) * x[i] / std::abs(x[i]);
    for (int j = 0; j < n; j++) x[j] /= fac;
    return fac;
}

MSGPU_API int inverse_iteration(int n, msgpu_cplx* a, msgpu_cplx* vec, msgpu_cplx* lambda, int maxit, double toler,
                                int verb) {
    std::vector<int> ipiv(n);
    int info = 0;
    die_on(lu_host(n, as_cx(a), n, ipiv.data(), 0, nullptr, 0, nullptr, &info), "inverse_iteration");
    if (info > 0)
        printf("[LAPACK] Element U(%d,%d) is exactly zero, the inverse power iteration will fail.\n", info, info);
    std::vector<msgpu_cplx> old(n);
    msgpu_cplx fac = normalize_vec(n, vec);
    int conv = 0;
    for (int s = 0; s < maxit; s++) {
        std::copy(vec, vec + n, old.begin());
        die_on(lu_solve_host(n, as_cx(a), n, ipiv.data(), 1, as_cx(vec), n), "inverse_iteration");
        fac = normalize_vec(n, vec);
        double d2 = 0;
        for (int i = 0; i < n; i++) d2 += std::norm(vec[i] - old[i]);
        if (verb >= 1) printf("[STEP] #%03d | fac=%18.12g%+18.12gi, delta=%g \n", s, fac.real(), fac.imag(), std::sqrt(d2));
        if (std::sqrt(d2) < toler) { conv = 1; break; }
    }
    *lambda = 1.0 / fac;
    return conv;
}

MSGPU_API double total_cross_section(msgpu_medium* med, double k) {
    const MediumView mv = view_of(med);
    std::vector<msgpu_cplx> a(mv.n), phi(mv.n);
    die_on(solve_amplitudes(mv, cx<double>{k, 0.0}, MSGPU_PLANE, as_cx(a.data()), 0, 0.0), "total_cross_section");
    die_on(incident_host(mv, cx<double>{k, 0.0}, MSGPU_PLANE, as_cx(phi.data())), "total_cross_section");
    msgpu_cplx dot = 0.0;
    for (int i = 0; i < mv.n; i++) dot += std::conj(phi[i]) * a[i];
    return -dot.imag() / k;
}

// ------------------------------------------------------------------ drop-in: incident_wave_util.h
MSGPU_API msgpu_cplx eval_incident_wave(msgpu_iwave* iw, int d, double* pos) {
    if (iw->source == MSGPU_PLANE) return std::exp(msgpu_cplx(0, 1) * iw->k * pos[0]);
    if (iw->source == MSGPU_SPHERICAL) {
        double s = 0;
        for (int c = 0; c < d; c++) s += pos[c] * pos[c];
        const cx<double> g = green<double>(d, to_cx(iw->k), std::sqrt(s));
        return to_c(g);
    }
    return 0.0;
}

// Called once per pixel by every thread of run_wave_function_omp (ms_core.c:629-640):
// concurrent calls on the same content run as one launch.
MSGPU_API msgpu_cplx eval_total_wave(msgpu_iwave* iw, msgpu_medium* med, msgpu_cplx* a, double* pos) {
    msgpu_cplx psi;
    die_on(wave_point_coalesced(view_of(med), to_cx(iw->k), iw->source, as_cx(a), pos, as_cx(&psi)),
           "eval_total_wave");
    return psi;
}

MSGPU_API msgpu_cplx invf(msgpu_medium* med, msgpu_cplx k) {
    return to_c(msg::invf(scmodel_of(med), med->d, to_cx(k)));
}

// ------------------------------------------------------------------ drop-in: LAPACK (LP64)
MSGPU_API void zgetrf_(int* m, int* n, msgpu_cplx* a, int* lda, int* ipiv, int* info) {
    if (*m != *n || *lda < std::max(1, *m)) { *info = (*m != *n) ? -1 : -4; return; }
    die_on(lu_host(*n, as_cx(a), *lda, ipiv, 0, nullptr, 0, nullptr, info), "zgetrf_");
}

MSGPU_API void zgetrs_(char* trans, int* n, int* nrhs, msgpu_cplx* a, int* lda, int* ipiv, msgpu_cplx* b, int* ldb,
                       int* info) {
    if (trans[0] != 'N' && trans[0] != 'n') { *info = -1; return; }
    *info = 0;
    if (*n == 0 || *nrhs == 0) return;
    die_on(lu_solve_host(*n, as_cx(a), *lda, ipiv, *nrhs, as_cx(b), *ldb), "zgetrs_");
}

MSGPU_API void zgesv_(int* n, int* nrhs, msgpu_cplx* a, int* lda, int* ipiv, msgpu_cplx* b, int* ldb, int* info) {
    die_on(lu_host(*n, as_cx(a), *lda, ipiv, *nrhs, as_cx(b), *ldb, nullptr, info), "zgesv_");
}

}  // extern "C"
