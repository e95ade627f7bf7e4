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

MSGPU_API int msgpu_tracelog_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, int mode,
                                   double reeval_rel, int* flags) {
    const MediumView mv = view_of(med);
    if (mode == MSGPU_FP64 || mode == MSGPU_FP32) {
        if (flags) memset(flags, 0, sizeof(int) * (nk > 0 ? nk : 0));
        return tracelog_points(mv, nk, as_cx(ks), as_cx(out), nullptr, mode == MSGPU_FP32 ? 1 : 0);
    }
    if (mode != MSGPU_MIXED) { set_error("msgpu: unknown mode"); return -4; }
    std::vector<double> stats(2 * (size_t)nk);
    int rc = tracelog_points(mv, nk, as_cx(ks), as_cx(out), stats.data(), 1);
    if (rc < 0) return rc;
    // Cells to recompute in FP64: |det| below reeval_rel of the map maximum (near zeros of
    // det M, where the complex64 factor loses the most relative accuracy) and cells whose
    // complex64 result is not finite.
    double best = -INFINITY;
    for (int l = 0; l < nk; l++)
        if (std::isfinite(out[l].real())) best = std::max(best, mv.n * out[l].real());
    const double cut = reeval_rel > 0 ? best + std::log(reeval_rel) : -INFINITY;
    std::vector<int> idx;
    for (int l = 0; l < nk; l++) {
        const bool f = (mv.n * out[l].real() < cut) || !std::isfinite(out[l].real());
        if (flags) flags[l] = f ? 1 : 0;
        if (f) idx.push_back(l);
    }
    if (!idx.empty()) {
        std::vector<msgpu_cplx> sub(idx.size()), res(idx.size());
        for (size_t i = 0; i < idx.size(); i++) sub[i] = ks[idx[i]];
        rc = tracelog_points(mv, (int)idx.size(), as_cx(sub.data()), as_cx(res.data()), nullptr, 0);
        if (rc < 0) return rc;
        for (size_t i = 0; i < idx.size(); i++) out[idx[i]] = res[i];
    }
    return (int)idx.size();
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

// Aberth-Ehrlich on det M(k) (root_finder.c:522-577): the exponential background is
// fitted from 6*nexpsup Halton samples of f'/f (root_finder.c:161-172); then every
// sweep evaluates f'/f for all live roots in one batched call and moves them
// simultaneously.  In MSGPU_MIXED, sweeps run on complex64 factors until the step
// falls below max(toler, 1e-5)|z|, then complex128 sweeps polish to toler.
MSGPU_API int msgpu_resonances(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                               msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode) {
    const MediumView mv = view_of(med);
    int n0 = *nz;
    std::vector<cd> zs;
    for (int i = 0; i < n0; i++) {   // duplicate removal, then the open-domain filter
        cd c = z[i];
        bool uniq = true;
        for (const cd& o : zs)
            if (std::abs(c - o) < toler * std::abs(c) || std::abs(c - o) < toler) { uniq = false; break; }
        if (uniq) zs.push_back(c);
    }
    std::vector<cd> live;
    for (const cd& c : zs)
        if (in_open(c, xmin, xmax, ymin, ymax)) live.push_back(c);
    const int nr = (int)live.size();
    if (nr <= 1) { *nz = nr; for (int i = 0; i < nr; i++) z[i] = live[i]; return 0; }
    std::vector<cd> poly(std::max(nexpsup, 1), 0.0);
    int rc;
    if (nexpsup > 0) {
        const int ns = 6 * nexpsup;
        std::vector<msgpu_cplx> zsmp(ns), fs(ns);
        for (int i = 0; i < ns; i++)
            zsmp[i] = cd(xmin + (xmax - xmin) * halton(i + 1, 2), ymin + (ymax - ymin) * halton(i + 1, 3));
        rc = dk_log_det_points(mv, ns, as_cx(zsmp.data()), as_cx(fs.data()), 0);
        if (rc < 0) return rc;
        std::vector<cd> a((size_t)ns * nexpsup), b(ns);
        for (int i = 0; i < ns; i++) {
            cd p = 1.0;
            for (int j = 0; j < nexpsup; j++) { a[i + (size_t)j * ns] = p; p *= cd(zsmp[i]); }
            b[i] = fs[i];
        }
        lstsq(ns, nexpsup, a, b);
        for (int j = 0; j < nexpsup; j++) poly[j] = b[j];
    }
    std::vector<int> flag(nr, 0);
    int prec = (mode == MSGPU_FP64) ? 0 : 1;
    const double coarse = std::max(toler, 1e-5);
    int sweeps = 0;
    for (int s = 0; s < maxit + 40 && sweeps < maxit + 40; s++, sweeps++) {
        std::vector<int> act;
        for (int i = 0; i < nr; i++) if (flag[i] == 0) act.push_back(i);
        if (act.empty()) break;
        std::vector<msgpu_cplx> za(act.size()), fa(act.size());
        for (size_t q = 0; q < act.size(); q++) za[q] = live[act[q]];
        rc = dk_log_det_points(mv, (int)act.size(), as_cx(za.data()), as_cx(fa.data()), prec);
        if (rc < 0) return rc;
        const double tol_now = prec ? coarse : toler;
        bool all_coarse = true;
        std::vector<cd> next = live;
        for (size_t q = 0; q < act.size(); q++) {
            const int i = act[q];
            cd pv = 0.0;
            for (int j = nexpsup - 1; j >= 0; j--) pv = pv * live[i] + poly[j];
            cd invdz = cd(fa[q]) - pv;
            for (int j = 0; j < nr; j++) if (j != i) invdz -= 1.0 / (live[i] - live[j]);
            const cd dz = 1.0 / invdz, zn = live[i] - dz;
            if (!in_open(zn, xmin, xmax, ymin, ymax)) flag[i] = -1;
            else if (std::abs(dz) < tol_now * std::abs(live[i])) { if (!prec) flag[i] = 1; next[i] = zn; }
            else { next[i] = zn; all_coarse = false; }
        }
        live = next;
        if (prec && all_coarse) prec = 0;       // switch to the complex128 polish
        if (s >= maxit && prec == 0) {
            bool busy = false;
            for (int i = 0; i < nr; i++) if (flag[i] == 0) busy = true;
            if (!busy) break;
        }
        if (s >= maxit + 39) break;
    }
    std::vector<cd> roots;
    for (int i = 0; i < nr; i++) if (flag[i] == 1) roots.push_back(live[i]);
    std::sort(roots.begin(), roots.end(), [](const cd& a, const cd& b) { return a.real() < b.real(); });
    *nz = (int)roots.size();
    for (size_t i = 0; i < roots.size(); i++) z[i] = roots[i];
    return sweeps;
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
