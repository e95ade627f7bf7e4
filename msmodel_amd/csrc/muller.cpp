// muller.cpp -- step-wise damped Muller iteration (see muller.h).
#include "muller.h"

#include <cmath>

namespace msg {

MullerRun::MullerRun(cd z0, double xmin, double xmax, double ymin, double ymax, int maxit, double toler)
    : x0_(xmin), x1_(xmax), y0_(ymin), y1_(ymax), toler_(toler), maxit_(maxit), z_(z0) {
    if (maxit <= 1 || toler <= 0.0) {       // invalid parameters: no iteration (root_finder.c:355-358)
        state_ = kDone;
        return;
    }
    const double delta = 1e-4;             // the finite step of root_finder.c:19
    za_ = cd(1.0, -delta) * z0;
    zb_ = z0;
    zc_ = cd(1.0, delta) * z0;
}

std::vector<MullerRun::cd> MullerRun::points() const {
    if (state_ == kInit) return {za_, zb_, zc_};
    if (state_ == kDamp) return {zn_};
    return {};
}

void MullerRun::take(const std::vector<cd>& f) {
    if (state_ == kInit) {
        fa_ = f[0];
        fb_ = f[1];
        fc_ = f[2];
        absfc_ = std::abs(fc_);
        begin_iteration();
        return;
    }
    if (state_ != kDamp) return;
    const cd fn = f[0];
    const double absfn = std::abs(fn);
    if (absfn > absfc_ && mu_ * absdz_ > tol_) {     // damp further
        zn_ = zc_ - mu_ * dz_;
        mu_ /= pow2_;
        pow2_ *= 2.0;
        return;
    }
    za_ = zb_; fa_ = fb_;
    zb_ = zc_; fb_ = fc_;
    zc_ = zn_; fc_ = fn; absfc_ = absfn;
    s_++;
    begin_iteration();
}

void MullerRun::begin_iteration() {
    if (s_ >= maxit_) { state_ = kDone; return; }
    const cd dfab = (fa_ - fb_) / (za_ - zb_), dfac = (fa_ - fc_) / (za_ - zc_), dfbc = (fb_ - fc_) / (zb_ - zc_);
    const cd d1f = dfbc + dfac - dfab, d2f = (dfac - dfbc) / (za_ - zb_);
    const cd delta = std::sqrt(d1f * d1f - 4.0 * fc_ * d2f);
    cd denom = d1f + delta;
    const cd denom1 = d1f - delta;
    if (std::abs(denom1) > std::abs(denom)) denom = denom1;
    if (denom == 0.0 || denom != denom) { state_ = kDone; return; }      // zero or NaN: give up
    dz_ = 2.0 * fc_ / denom;
    absdz_ = std::abs(dz_);
    tol_ = toler_ * std::abs(zc_);
    if (!in_open(zc_)) { state_ = kDone; return; }
    if (absdz_ < tol_ || absfc_ == 0.0) {
        z_ = zc_;
        conv_ = true;
        state_ = kDone;
        return;
    }
    mu_ = 1.0;
    pow2_ = 2.0;
    zn_ = zc_ - mu_ * dz_;
    mu_ /= pow2_;
    pow2_ *= 2.0;
    state_ = kDamp;
}

}  // namespace msg
