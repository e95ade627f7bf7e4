// aberth.cpp -- step-wise Maehly-Aberth-Ehrlich iteration (see aberth.h).
#include "aberth.h"

#include <algorithm>
#include <cmath>

namespace msg {

namespace {

using cd = std::complex<double>;

// Dense complex least squares min |A x - b| by Householder QR; A is m x nc, column-major.
void lstsq(int m, int nc, std::vector<cd>& a, std::vector<cd>& b) {
    for (int j = 0; j < nc; j++) {
        double nrm = 0;
        for (int i = j; i < m; i++) nrm += std::norm(a[i + j * m]);
        nrm = std::sqrt(nrm);
        if (nrm == 0) continue;
        const cd ajj = a[j + j * m];
        const cd alpha = (std::abs(ajj) > 0 ? -ajj / std::abs(ajj) : cd(-1)) * nrm;
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

// Radical inverse of i in the given base (Halton sequence).
double halton(int i, int base) {
    double h = 0, f = 1;
    while (i > 0) { f /= base; h += f * (i % base); i /= base; }
    return h;
}

}  // namespace

AberthRun::AberthRun(const cd* starts, int nstart, double xmin, double xmax, double ymin, double ymax, int nexpsup,
                     int maxit, double toler, bool mixed)
    : x0_(xmin), x1_(xmax), y0_(ymin), y1_(ymax), toler_(toler), nexpsup_(std::max(nexpsup, 0)), maxit_(maxit),
      prec_(mixed ? 1 : 0), mixed_(mixed), fitted_(nexpsup <= 0) {
    std::vector<cd> zs;
    for (int i = 0; i < nstart; i++) {
        const cd c = starts[i];
        bool uniq = true;
        for (const cd& o : zs)
            if (std::abs(c - o) < toler * std::abs(c) || std::abs(c - o) < toler) { uniq = false; break; }
        if (uniq) zs.push_back(c);
    }
    for (const cd& c : zs)
        if (in_open(c)) live_.push_back(c);
    poly_.assign(std::max(nexpsup_, 1), 0.0);
    flag_.assign(live_.size(), 0);
    if (live_.size() <= 1) {            // nothing to iterate: the start is returned as is
        std::fill(flag_.begin(), flag_.end(), 1);
        fitted_ = true;
    }
}

std::vector<cd> AberthRun::fit_points() const {
    std::vector<cd> z;
    if (fitted_) return z;
    const int ns = 6 * nexpsup_;
    for (int i = 0; i < ns; i++)
        z.push_back(cd(x0_ + (x1_ - x0_) * halton(i + 1, 2), y0_ + (y1_ - y0_) * halton(i + 1, 3)));
    return z;
}

void AberthRun::take_fit(const std::vector<cd>& f) {
    const std::vector<cd> z = fit_points();
    const int ns = (int)z.size();
    std::vector<cd> a((size_t)ns * nexpsup_), b(ns);
    for (int i = 0; i < ns; i++) {
        cd p = 1.0;
        for (int j = 0; j < nexpsup_; j++) { a[i + (size_t)j * ns] = p; p *= z[i]; }
        b[i] = f[i];
    }
    lstsq(ns, nexpsup_, a, b);
    for (int j = 0; j < nexpsup_; j++) poly_[j] = b[j];
    fitted_ = true;
}

// The reference stops after maxit sweeps (root_finder.c:546); a mixed-precision run
// may add up to kPolish complex128 sweeps after its complex64 phase.
static constexpr int kPolish = 10;

bool AberthRun::finished() const {
    if (!fitted_) return false;
    if (sweeps_ >= maxit_ + (mixed_ ? kPolish : 0)) return true;
    for (int f : flag_)
        if (f == 0) return false;
    return true;
}

std::vector<cd> AberthRun::sweep_points(int* precision) const {
    std::vector<cd> z;
    *precision = prec_;
    if (!fitted_ || finished()) return z;
    for (size_t i = 0; i < live_.size(); i++)
        if (flag_[i] == 0) z.push_back(live_[i]);
    return z;
}

// One sweep as root_finder.c:546-565: f'/f of every live root was evaluated at its
// position before the sweep (a root's own evaluation never sees its update); the
// repulsion terms then use the positions already updated in this sweep for the roots
// before it (Gauss-Seidel order).  A converged root keeps its previous position.
void AberthRun::take_sweep(const std::vector<cd>& f) {
    act_.clear();
    for (size_t i = 0; i < live_.size(); i++)
        if (flag_[i] == 0) act_.push_back((int)i);
    const double tol_now = prec_ ? std::max(toler_, 1e-5) : toler_;
    bool all_coarse = true;
    std::vector<char> coarse(live_.size(), 0);
    for (size_t q = 0; q < act_.size(); q++) {
        const int i = act_[q];
        cd pv = 0.0;
        for (int j = nexpsup_ - 1; j >= 0; j--) pv = pv * live_[i] + poly_[j];
        cd invdz = f[q] - pv;
        for (size_t j = 0; j < live_.size(); j++)
            if ((int)j != i) invdz -= 1.0 / (live_[i] - live_[j]);
        const cd dz = 1.0 / invdz, zn = live_[i] - dz;
        if (!in_open(zn)) {
            flag_[i] = -1;
        } else if (std::abs(dz) < tol_now * std::abs(live_[i])) {
            if (!prec_) flag_[i] = 1;
            else { live_[i] = zn; coarse[i] = 1; }   // complex64 phase: keep moving until the polish
        } else {
            live_[i] = zn;
            all_coarse = false;
        }
    }
    sweeps_++;
    // switch to the complex128 polish once every live root is coarse-converged, or when
    // the complex64 phase has used its maxit sweeps (roots still far off are dropped)
    if (prec_ && (all_coarse || sweeps_ >= maxit_)) {
        for (int i : act_)
            if (flag_[i] == 0 && !coarse[i]) flag_[i] = -1;
        prec_ = 0;
    }
}

std::vector<cd> AberthRun::roots() const {
    std::vector<cd> r;
    for (size_t i = 0; i < live_.size(); i++)
        if (flag_[i] == 1) r.push_back(live_[i]);
    std::sort(r.begin(), r.end(), [](const cd& a, const cd& b) { return a.real() < b.real(); });
    return r;
}

}  // namespace msg
