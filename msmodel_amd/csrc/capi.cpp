// capi.cpp -- the exported C ABI (include/msgpu.h): drop-in replacements with
// the reference signatures and the batched msgpu_* entry points.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

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

// Small dense complex least squares min |A x - b| (Householder QR), A m x nc column-major.
void lstsq(int m, int nc, std::vector<cd>& a, std::vector<cd>& b) {
    for (int j = 0; j < nc; j++) {
        double nrm = 0;
        for (int i = j; i < m; i++) nrm += std::norm(a[i + j * m]);
        nrm = std::sqrt(nrm);
        if (nrm == 0) continue;
        cd ajj = a[j + j * m];
        cd alpha = (std::abs(ajj) > 0 ? -ajj / std::abs(ajj) : cd(-1)) * nrm;
        std::vector<cd> v(m, 0.0);
        for (int i = j; i < m; i++) v[i] = a[i + j * m];
        v[j] -= alpha;
        double vn = 0;
        for (int i = j; i < m; i++) vn += std::norm(v[i]);
        if (vn == 0) continue;
        for (int c = j; c < nc; c++) {
            cd s = 0;
            for (int i = j; i < m; i++) s += std::conj(v[i]) * a[i + c * m];
            s *= 2.0 / vn;
            for (int i = j; i < m; i++) a[i + c * m] -= s * v[i];
        }
        cd s = 0;
        for (int i = j; i < m; i++) s += std::conj(v[i]) * b[i];
        s *= 2.0 / vn;
        for (int i = j; i < m; i++) b[i] -= s * v[i];
    }
    for (int j = nc - 1; j >= 0; j--) {
        cd s = b[j];
        for (int c = j + 1; c < nc; c++) s -= a[j + c * m] * b[c];
        b[j] = s / a[j + j * m];
    }
}

double halton(int i, int base) {
    double h = 0, f = 1;
    while (i > 0) { f /= base; h += f * (i % base); i /= base; }
    return h;
}

bool in_open(cd z, double x0, double x1, double y0, double y1) {
    return x0 < z.real() && z.real() < x1 && y0 < z.imag() && z.imag() < y1;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ batched API
MSGPU_API int msgpu_init(void) { return engine_init(); }
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
int tracelog_any(const std::vector<MediumView>& mv, int nk, const int* midx, const msgpu_cplx* ks, msgpu_cplx* out,
                 int mode, double reeval_rel, int* flags) {
    if (mode == MSGPU_FP64 || mode == MSGPU_FP32) {
        if (flags && nk > 0) memset(flags, 0, sizeof(int) * nk);
        return tracelog_multi(mv, nk, midx, as_cx(ks), as_cx(out), nullptr, mode == MSGPU_FP32 ? 1 : 0);
    }
    if (mode != MSGPU_MIXED) { set_error("msgpu: unknown mode"); return -4; }
    if (nk <= 0) return 0;
    int rc = tracelog_multi(mv, nk, midx, as_cx(ks), as_cx(out), nullptr, 1);
    if (rc < 0) return rc;
    std::vector<double> best(mv.size(), -INFINITY);
    for (int l = 0; l < nk; l++) {
        const int m = midx ? midx[l] : 0;
        if (std::isfinite(out[l].real())) best[m] = std::max(best[m], mv[m].n * out[l].real());
    }
    std::vector<int> idx, sub_m;
    for (int l = 0; l < nk; l++) {
        const int m = midx ? midx[l] : 0;
        const double cut = reeval_rel > 0 ? best[m] + std::log(reeval_rel) : -INFINITY;
        const bool f = (mv[m].n * out[l].real() < cut) || !std::isfinite(out[l].real());
        if (flags) flags[l] = f ? 1 : 0;
        if (f) { idx.push_back(l); sub_m.push_back(m); }
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

// Maehly-Aberth-Ehrlich on det M(k) for several media at once (root_finder.c:522-577).
// Per medium: duplicate removal and the open-domain filter of the starts, the
// exponential background fitted from 6*nexpsup Halton samples of f'/f
// (root_finder.c:161-172), then simultaneous sweeps; every sweep evaluates f'/f for
// the live roots of all media in one batched call per precision.  In MSGPU_MIXED a
// medium's sweeps run on complex64 factors until all its steps are below
// max(toler, 1e-5)|z|, then complex128 sweeps polish to toler.
struct AberthState {
    double x0, x1, y0, y1;
    std::vector<cd> live, poly;
    std::vector<int> flag;      // 0 live, 1 converged, -1 left the domain
    int prec = 0;
    bool done = false;
};

int resonances_any(const std::vector<MediumView>& mv, const double* doms, int* nz, msgpu_cplx* z, int nexpsup,
                   int maxit, double toler, int mode) {
    const int nmed = (int)mv.size();
    std::vector<AberthState> S(nmed);
    std::vector<int> offs(nmed, 0);
    for (int m = 0, o = 0; m < nmed; m++) { offs[m] = o; o += nz[m]; }
    for (int m = 0; m < nmed; m++) {
        AberthState& A = S[m];
        A.x0 = doms[4 * m]; A.x1 = doms[4 * m + 1]; A.y0 = doms[4 * m + 2]; A.y1 = doms[4 * m + 3];
        std::vector<cd> zs;
        for (int i = 0; i < nz[m]; i++) {
            const cd c = z[offs[m] + i];
            bool uniq = true;
            for (const cd& o : zs)
                if (std::abs(c - o) < toler * std::abs(c) || std::abs(c - o) < toler) { uniq = false; break; }
            if (uniq) zs.push_back(c);
        }
        for (const cd& c : zs)
            if (in_open(c, A.x0, A.x1, A.y0, A.y1)) A.live.push_back(c);
        A.flag.assign(A.live.size(), 0);
        A.poly.assign(std::max(nexpsup, 1), 0.0);
        A.prec = (mode == MSGPU_FP64) ? 0 : 1;
        A.done = A.live.size() <= 1;    // zero or one start: nothing to iterate (as before)
        if (A.done) std::fill(A.flag.begin(), A.flag.end(), 1);
    }
    int rc;
    if (nexpsup > 0) {      // background fit, all media in one FP64 call
        const int ns = 6 * nexpsup;
        std::vector<msgpu_cplx> zs, fs;
        std::vector<int> ms;
        for (int m = 0; m < nmed; m++) {
            if (S[m].done) continue;
            for (int i = 0; i < ns; i++) {
                zs.push_back(cd(S[m].x0 + (S[m].x1 - S[m].x0) * halton(i + 1, 2),
                                S[m].y0 + (S[m].y1 - S[m].y0) * halton(i + 1, 3)));
                ms.push_back(m);
            }
        }
        fs.resize(zs.size());
        rc = dk_log_det_multi(mv, (int)zs.size(), ms.data(), as_cx(zs.data()), as_cx(fs.data()), 0);
        if (rc < 0) return rc;
        for (size_t base = 0; base < zs.size(); base += ns) {
            const int m = ms[base];
            std::vector<cd> a((size_t)ns * nexpsup), b(ns);
            for (int i = 0; i < ns; i++) {
                cd p = 1.0;
                for (int j = 0; j < nexpsup; j++) { a[i + (size_t)j * ns] = p; p *= cd(zs[base + i]); }
                b[i] = fs[base + i];
            }
            lstsq(ns, nexpsup, a, b);
            for (int j = 0; j < nexpsup; j++) S[m].poly[j] = b[j];
        }
    }
    const double coarse = std::max(toler, 1e-5);
    int sweeps = 0;
    for (int s = 0; s < maxit + 40; s++) {
        // live roots of all media, split by the precision of their medium
        std::vector<msgpu_cplx> zq[2], fq[2];
        std::vector<int> mq[2], iq[2];
        for (int m = 0; m < nmed; m++) {
            if (S[m].done) continue;
            for (size_t i = 0; i < S[m].live.size(); i++)
                if (S[m].flag[i] == 0) {
                    zq[S[m].prec].push_back(S[m].live[i]);
                    mq[S[m].prec].push_back(m);
                    iq[S[m].prec].push_back((int)i);
                }
        }
        if (zq[0].empty() && zq[1].empty()) break;
        sweeps++;
        for (int p = 0; p < 2; p++) {
            if (zq[p].empty()) continue;
            fq[p].resize(zq[p].size());
            rc = dk_log_det_multi(mv, (int)zq[p].size(), mq[p].data(), as_cx(zq[p].data()), as_cx(fq[p].data()), p);
            if (rc < 0) return rc;
        }
        std::vector<std::vector<cd>> next(nmed);
        std::vector<char> all_coarse(nmed, 1);
        for (int m = 0; m < nmed; m++) next[m] = S[m].live;
        for (int p = 0; p < 2; p++)
            for (size_t q = 0; q < zq[p].size(); q++) {
                const int m = mq[p][q], i = iq[p][q];
                AberthState& A = S[m];
                cd pv = 0.0;
                for (int j = nexpsup - 1; j >= 0; j--) pv = pv * A.live[i] + A.poly[j];
                cd invdz = cd(fq[p][q]) - pv;
                for (size_t j = 0; j < A.live.size(); j++)
                    if ((int)j != i) invdz -= 1.0 / (A.live[i] - A.live[j]);
                const cd dz = 1.0 / invdz, zn = A.live[i] - dz;
                const double tol_now = A.prec ? coarse : toler;
                if (!in_open(zn, A.x0, A.x1, A.y0, A.y1)) A.flag[i] = -1;
                else if (std::abs(dz) < tol_now * std::abs(A.live[i])) { if (!A.prec) A.flag[i] = 1; next[m][i] = zn; }
                else { next[m][i] = zn; all_coarse[m] = 0; }
            }
        for (int m = 0; m < nmed; m++) {
            if (S[m].done) continue;
            S[m].live = next[m];
            if (S[m].prec && all_coarse[m]) S[m].prec = 0;     // switch to the complex128 polish
        }
    }
    for (int m = 0; m < nmed; m++) {
        std::vector<cd> roots;
        for (size_t i = 0; i < S[m].live.size(); i++)
            if (S[m].flag[i] == 1) roots.push_back(S[m].live[i]);
        std::sort(roots.begin(), roots.end(), [](const cd& a, const cd& b) { return a.real() < b.real(); });
        nz[m] = (int)roots.size();
        for (size_t i = 0; i < roots.size(); i++) z[offs[m] + i] = roots[i];
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
    return dk_log_det_multi(views_of(nmed, meds), nk, med_of, as_cx(ks), as_cx(out), mode == MSGPU_FP64 ? 0 : 1);
}

MSGPU_API int msgpu_resonances_multi(int nmed, const msgpu_medium* const* meds, const double* doms, int* nz,
                                     msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode) {
    if (nmed <= 0 || !meds) { set_error("msgpu: no medium given"); return -2; }
    return resonances_any(views_of(nmed, meds), doms, nz, z, nexpsup, maxit, toler, mode);
}

MSGPU_API int msgpu_tracelog_stats(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out,
                                   double* stats, int mode) {
    return tracelog_points(view_of(med), nk, as_cx(ks), as_cx(out), stats, mode == MSGPU_FP64 ? 0 : 1);
}

MSGPU_API int msgpu_dk_log_det_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out,
                                     int mode) {
    return dk_log_det_points(view_of(med), nk, as_cx(ks), as_cx(out), mode == MSGPU_FP64 ? 0 : 1);
}

MSGPU_API int msgpu_solve(const msgpu_medium* med, const msgpu_iwave* iwave, msgpu_cplx* a, int mode, double tol) {
    return solve_amplitudes(view_of(med), to_cx(iwave->k), iwave->source, as_cx(a), mode == MSGPU_FP64 ? 0 : 1,
                            tol);
}

MSGPU_API int msgpu_wavefunction_grid(const msgpu_medium* med, const msgpu_iwave* iwave, const msgpu_cplx* a, int nx,
                                      int ny, double xmin, double xmax, double ymin, double ymax, msgpu_cplx* psi) {
    return wave_grid(view_of(med), to_cx(iwave->k), iwave->source, as_cx(a), nx, ny, xmin, xmax, ymin, ymax,
                     as_cx(psi));
}

// Aberth-Ehrlich on det M(k) of one medium (root_finder.c:522-577); see resonances_any.
MSGPU_API int msgpu_resonances(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                               msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode) {
    const double dom[4] = {xmin, xmax, ymin, ymax};
    return resonances_any(views_of(1, &med), dom, nz, z, nexpsup, maxit, toler, mode);
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
    die_on(lu_host(n, as_cx(matrix), n, ipiv.data(), 0, nullptr, 0, &tl, &info), "tracelog_lu");
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

MSGPU_API msgpu_cplx eval_total_wave(msgpu_iwave* iw, msgpu_medium* med, msgpu_cplx* a, double* pos) {
    msgpu_cplx psi;
    die_on(wave_points(view_of(med), to_cx(iw->k), iw->source, as_cx(a), 1, pos, as_cx(&psi)), "eval_total_wave");
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
