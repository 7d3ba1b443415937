// TSModel / GraphH1FEMModel over the HipVector backend (include/general/model/tsmodel.h:29-133,
// graphh1fem.h; the PETSc specialisation src/petscvector/model/graphh1fem.cpp).
#pragma once

#include "common.h"
#include "solvers.h"
#include "tsdata.h"

namespace pascinference {
namespace model {

template <class VectorBase> class TSModel {
  public:
    virtual ~TSModel() {}
    virtual std::string get_name() const { return "TSModel"; }
    virtual void initialize_gammasolver(solver::GeneralSolver **gammasolver) { *gammasolver = nullptr; }
    virtual void initialize_thetasolver(solver::GeneralSolver **thetasolver) { *thetasolver = nullptr; }
    virtual void finalize_gammasolver(solver::GeneralSolver **gammasolver) { (void)gammasolver; }
    virtual void finalize_thetasolver(solver::GeneralSolver **thetasolver) { (void)thetasolver; }
    virtual void updatebeforesolve_gammasolver(solver::GeneralSolver *) {}
    virtual void updateaftersolve_gammasolver(solver::GeneralSolver *) {}
    virtual void updatebeforesolve_thetasolver(solver::GeneralSolver *) {}
    virtual void updateaftersolve_thetasolver(solver::GeneralSolver *) {}
    virtual double get_L(solver::GeneralSolver *gammasolver, solver::GeneralSolver *thetasolver) {
        (void)gammasolver;
        (void)thetasolver;
        return std::numeric_limits<double>::max();
    }
    virtual double get_aic(double L) const { return L; }
    virtual int get_thetavectorlength_local() const { return 0; }
};

template <class VectorBase> class GraphH1FEMModel;

template <> class GraphH1FEMModel<HipVector> : public TSModel<HipVector> {
  public:
    // graphh1fem.cpp:7-65; the epsilon coefficient is the reference's epssqr
    GraphH1FEMModel(data::TSData<HipVector> &tsdata, double epssqr, const Options &opt = Options())
        : tsdata_(&tsdata), opt_(opt) {
        set_epssqr(epssqr);
    }
    std::string get_name() const override { return "Graph-H1-FEM Time-Series Model"; }

    // set_epssqr (graphh1fem.h:286-306): A = epssqr/T * A0 with scalef
    void set_epssqr(double epssqr) {
        epssqr_ = epssqr;
        PASC_TRY(pasc_set_epssqr(tsdata_->get_problem(), epssqr));
    }
    double get_epssqr() const { return epssqr_; }

    // initialize_gammasolver (graphh1fem.cpp:115-214): the automatic choice is SPG-QP over the
    // per-time-step simplices
    void initialize_gammasolver(solver::GeneralSolver **gammasolver) override {
        *gammasolver = new solver::SPGQPSolver<HipVector>(*tsdata_, opt_);
    }
    void initialize_thetasolver(solver::GeneralSolver **thetasolver) override {
        *thetasolver = new solver::SimpleSolver<HipVector>();
    }
    void finalize_gammasolver(solver::GeneralSolver **gammasolver) override {
        delete *gammasolver;
        *gammasolver = nullptr;
    }
    void finalize_thetasolver(solver::GeneralSolver **thetasolver) override {
        delete *thetasolver;
        *thetasolver = nullptr;
    }
    // updatebeforesolve_gammasolver (graphh1fem.cpp:236-289) assembles b = -(x - theta)^2 / T;
    // the device iteration forms b on the fly from x and theta, so nothing is assembled here
    void updatebeforesolve_gammasolver(solver::GeneralSolver *) override {}
    // updatebeforesolve_thetasolver (graphh1fem.cpp:305-396): k-means weighted means (FP64)
    void updatebeforesolve_thetasolver(solver::GeneralSolver *) override {
        PASC_TRY(pasc_theta_update(tsdata_->get_problem()));
    }
    // get_L (graphh1fem.h:367-371): the gamma solver's objective value
    double get_L(solver::GeneralSolver *gammasolver, solver::GeneralSolver *) override {
        return static_cast<solver::QPSolver<HipVector> *>(gammasolver)->get_fx();
    }
    // get_aic (graphh1fem.h:443-447)
    double get_aic(double L) const override { return 2.0 * std::log(L) + tsdata_->get_K(); }
    int get_thetavectorlength_local() const override {
        return tsdata_->get_K() * tsdata_->get_xdim();
    }

  private:
    data::TSData<HipVector> *tsdata_;
    Options opt_;
    double epssqr_ = 0.0;
};

}  // namespace model
}  // namespace pascinference
