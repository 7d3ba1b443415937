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
    // T time steps of xdim values (R = 1), K clusters, precision PASC_FP64 / PASC_FP32_BULK.
    // fem_reduce < 1 solves the gamma QP on a coarse grid (the reference passes a Fem / FemHat
    // object to GraphH1FEMModel, test_signal1D.cpp:218-228; here the problem owns it).
    TSData(pasc_ctx *ctx, const std::vector<double> &x, int K, int xdim = 1,
           int precision = PASC_FP64, const std::vector<uint8_t> *heldout = nullptr,
           double fem_reduce = 1.0, int fem_type = PASC_FEM_HAT)
        : ctx_(ctx), x_(x), K_(K), xdim_(xdim) {
        T_ = (int64_t)x.size() / xdim;
        pasc_problem_desc d;
        std::memset(&d, 0, sizeof(d));
        d.T = T_;
        d.R = 1;
        d.K = K;
        d.xdim = xdim;
        d.scalef = 1;
        d.precision = precision;
        d.x = x_.data();
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
    int get_R() const { return 1; }
    int get_K() const { return K_; }
    int get_xdim() const { return xdim_; }
    const std::vector<double> &get_datavector() const { return x_; }

    // gamma in the reference's t-major layout
    std::vector<double> get_gammavector() const {
        std::vector<double> g((size_t)T_ * K_);
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
        if ((int64_t)g.size() != T_ * K_) throw std::runtime_error("gamma0 has the wrong length");
        PASC_TRY(pasc_set_gamma(problem_, g.data(), PASC_LAYOUT_KMAJOR));
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
    int64_t T_ = 0;
    int K_ = 0, xdim_ = 1;
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
