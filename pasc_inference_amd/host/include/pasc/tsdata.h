// TSData<HipVector> / Signal1DData<HipVector>: owners of the data, gamma and theta of one
// time series (include/general/data/tsdata.h, signal1Ddata.h).  Data, gamma and theta live in a
// pasc_problem on the device; this class keeps the host-side descriptions and file I/O.
#pragma once

#include "common.h"

namespace pascinference {
namespace data {

template <class VectorBase> class TSData;

template <> class TSData<HipVector> {
  public:
    // T time steps of R nodes with xdim values each ([T][R][xdim]), K clusters, precision
    // PASC_FP64 / PASC_FP32_BULK.  R > 1 with grid_width > 0: the nodes form the 4-neighbour
    // grid of that width (the BGMGraphGrid2D of test_image.cpp:156-168).  fem_reduce < 1 solves
    // the gamma QP on a coarse grid (the reference passes a Fem / FemHat object to
    // GraphH1FEMModel, test_signal1D.cpp:218-228; here the problem owns it).
    TSData(pasc_ctx *ctx, const std::vector<double> &x, int K, int xdim = 1,
           int precision = PASC_FP64, const std::vector<uint8_t> *heldout = nullptr,
           double fem_reduce = 1.0, int fem_type = PASC_FEM_HAT, int64_t R = 1,
           int32_t grid_width = 0)
        : ctx_(ctx), x_(x), R_(R), K_(K), xdim_(xdim) {
        T_ = (int64_t)x.size() / ((int64_t)xdim * R);
        pasc_problem_desc d;
        std::memset(&d, 0, sizeof(d));
        d.T = T_;
        d.R = R;
        d.K = K;
        d.xdim = xdim;
        d.scalef = 1;
        d.precision = precision;
        d.x = x_.data();
        d.grid_width = grid_width;
        d.fem_reduce = fem_reduce;
        d.fem_type = fem_type;
        if (heldout) {
            heldout_ = *heldout;
            d.heldout = heldout_.data();
        }
        ok_ = PASC_TRY(pasc_problem_create(ctx, &d, &problem_));
        theta_.assign((size_t)K * xdim, 0.0);
    }
    ~TSData() {
        if (problem_) pasc_problem_destroy(problem_);
    }
    TSData(const TSData &) = delete;
    TSData &operator=(const TSData &) = delete;

    bool ok() const { return ok_; }
    pasc_problem *get_problem() const { return problem_; }
    int64_t get_T() const { return T_; }
    int64_t get_R() const { return R_; }
    int get_K() const { return K_; }
    int get_xdim() const { return xdim_; }
    const std::vector<double> &get_datavector() const { return x_; }

    // gamma in the reference's t-major layout
    std::vector<double> get_gammavector() const {
        std::vector<double> g((size_t)T_ * R_ * K_);
        PASC_TRY(pasc_get_gamma(problem_, g.data(), PASC_LAYOUT_TMAJOR));
        return g;
    }
    void set_gammavector(const std::vector<double> &g) {
        PASC_TRY(pasc_set_gamma(problem_, g.data(), PASC_LAYOUT_TMAJOR));
    }
    std::vector<double> get_thetavector() const {
        std::vector<double> th((size_t)K_ * xdim_);
        PASC_TRY(pasc_get_theta(problem_, th.data()));
        return th;
    }
    void set_thetavector(const std::vector<double> &th) {
        PASC_TRY(pasc_set_theta(problem_, th.data()));
    }
    // load_gammavector (tsdata.cpp:548-627): k-major file permuted to t-major
    void load_gammavector(const std::string &filename) {
        const std::vector<double> g = petscio::read_vec(filename);
        if ((int64_t)g.size() != T_ * R_ * K_) throw std::runtime_error("gamma0 has the wrong length");
        PASC_TRY(pasc_set_gamma(problem_, g.data(), PASC_LAYOUT_KMAJOR));
    }

    // printstats (tsdata.cpp:310-383): length, blocks and the sum, max, min and mean of the data
    void printstats(std::ostream &o) const {
        double sum = 0.0, mx = x_.empty() ? 0.0 : x_[0], mn = mx;
        for (double v : x_) {
            sum += v;
            mx = v > mx ? v : mx;
            mn = v < mn ? v : mn;
        }
        const std::streamsize prec = o.precision();
        o << std::setprecision(17) << "STATS: TSData" << std::endl;
        o << " - total length:    " << std::setw(25) << x_.size() << std::endl;
        o << " - nmb of blocks:   " << std::setw(25) << R_ * K_ << std::endl;
        o << " - length of block: " << std::setw(25) << T_ << std::endl;
        o << " - sum:             " << std::setw(25) << sum << std::endl;
        o << " - max:             " << std::setw(25) << mx << std::endl;
        o << " - min:             " << std::setw(25) << mn << std::endl;
        o << " - avg:             " << std::setw(25) << (x_.empty() ? 0.0 : sum / x_.size())
          << std::endl;
        o << std::setprecision(prec);
    }

    // cutdata (tsdata.cpp:461-478): clamp the data to [a, b]
    void cutdata(double a, double b) {
        for (double &v : x_) v = v > b ? b : (v < a ? a : v);
        PASC_TRY(pasc_set_data(problem_, x_.data()));
    }
    // scaledata (tsdata.cpp:384-422): map [min x, max x] linearly onto [a, b], theta with it
    void scaledata(double a, double b) {
        double lo = x_.empty() ? 0.0 : x_[0], hi = lo;
        for (double v : x_) {
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        scaledata(a, b, lo, hi);
    }
    // scaledata (tsdata.cpp:511-545): map [scale_min, scale_max] linearly onto [a, b]
    void scaledata(double a, double b, double scale_min, double scale_max) {
        scale_min_ = scale_min;
        scale_max_ = scale_max;
        const double k = (b - a) / (scale_max - scale_min), q = a - k * scale_min;
        for (double &v : x_) v = k * v + q;
        PASC_TRY(pasc_set_data(problem_, x_.data()));
        std::vector<double> th = get_thetavector();
        for (double &v : th) v = k * v + q;
        set_thetavector(th);
    }
    // unscaledata (tsdata.cpp:424-459): the inverse of the last scaledata(a, b)
    void unscaledata(double a, double b) {
        const double k = (b - a) / (scale_max_ - scale_min_), q = a - k * scale_min_;
        for (double &v : x_) v = (v - q) / k;
        PASC_TRY(pasc_set_data(problem_, x_.data()));
        std::vector<double> th = get_thetavector();
        for (double &v : th) v = (v - q) / k;
        set_thetavector(th);
    }
    // cutgamma (tsdata.cpp:196-232): every node's gamma set to the indicator of its largest
    // component (the first one on a tie).  The reference starts each node's search from
    // gamma_arr[t] (tsdata.cpp:212), an element of another node for t > 0; here the search
    // starts from the node's own component 0.
    void cutgamma() {
        std::vector<double> g = get_gammavector();
        const size_t n = (size_t)(T_ * R_);
        for (size_t i = 0; i < n; i++) {
            int best = 0;
            for (int k = 1; k < K_; k++)
                if (g[i * K_ + k] > g[i * K_ + best]) best = k;
            for (int k = 0; k < K_; k++) g[i * K_ + k] = k == best ? 1.0 : 0.0;
        }
        set_gammavector(g);
    }
    double get_aic() const { return aic_; }
    void set_aic(double a) { aic_ = a; }
    std::string print_thetavector() const {
        std::ostringstream o;
        for (double v : get_thetavector()) o << v << ",";
        return o.str();
    }

  protected:
    pasc_ctx *ctx_;
    pasc_problem *problem_ = nullptr;
    std::vector<double> x_;
    std::vector<uint8_t> heldout_;
    std::vector<double> theta_;
    int64_t T_ = 0, R_ = 1;
    int K_ = 0, xdim_ = 1;
    double scale_min_ = 0.0, scale_max_ = 1.0;
    double aic_ = std::numeric_limits<double>::max();
    bool ok_ = false;
};

template <class VectorBase> class Signal1DData;

// 1-D signal read from a PETSc binary Vec (signal1Ddata.cpp:7-30)
template <> class Signal1DData<HipVector> : public TSData<HipVector> {
  public:
    Signal1DData(pasc_ctx *ctx, const std::string &filename, int K, int precision = PASC_FP64,
                 double fem_reduce = 1.0, int fem_type = PASC_FEM_HAT)
        : TSData<HipVector>(ctx, petscio::read_vec(filename), K, 1, precision, nullptr,
                            fem_reduce, fem_type) {}
    Signal1DData(pasc_ctx *ctx, const std::vector<double> &x, int K, int precision = PASC_FP64,
                 double fem_reduce = 1.0, int fem_type = PASC_FEM_HAT)
        : TSData<HipVector>(ctx, x, K, 1, precision, nullptr, fem_reduce, fem_type) {}

    // compute_abserr_reconstructed (signal1Ddata.cpp:133-191): || sum_k theta_k gamma_k - sol ||_1
    double compute_abserr_reconstructed(const std::vector<double> &solution) const {
        double a = 0.0;
        PASC_TRY(pasc_abserr_reconstructed(problem_, solution.data(), &a));
        return a;
    }
    // saveSignal1D (signal1Ddata.cpp:54-131): recovered signal and k-major gamma files
    void saveSignal1D(const std::string &prefix) const {
        const std::vector<double> th = get_thetavector();
        std::vector<double> gk((size_t)T_ * K_);
        PASC_TRY(pasc_get_gamma(problem_, gk.data(), PASC_LAYOUT_KMAJOR));
        std::vector<double> rec((size_t)T_, 0.0);
        for (int k = 0; k < K_; k++)
            for (int64_t t = 0; t < T_; t++) rec[(size_t)t] += th[(size_t)k] * gk[(size_t)(k * T_ + t)];
        petscio::write_vec(prefix + "_recovered.bin", rec);
        petscio::write_vec(prefix + "_gamma.bin", gk);
    }
};

}  // namespace data
}  // namespace pascinference
