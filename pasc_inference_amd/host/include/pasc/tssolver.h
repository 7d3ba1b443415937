// TSSolver<HipVector>: the outer theta/gamma alternation (include/general/solver/tssolver.h,
// solve() :491-715, set_solution_theta :777-790, defaults :13-15).
#pragma once

#include "common.h"
#include "model.h"

namespace pascinference {
namespace solver {

template <class VectorBase> class TSSolver;

template <> class TSSolver<HipVector> : public GeneralSolver {
  public:
    TSSolver(data::TSData<HipVector> &tsdata, model::TSModel<HipVector> &model,
             const Options &opt = Options())
        : tsdata_(&tsdata), model_(&model) {
        opt.set_option_value("tssolver_maxit", &maxit, 300);
        opt.set_option_value("tssolver_eps", &eps, 1e-6);
        opt.set_option_value("tssolver_debugmode", &debugmode, 0);
        model_->initialize_gammasolver(&gammasolver_);
        model_->initialize_thetasolver(&thetasolver_);
    }
    ~TSSolver() override {
        model_->finalize_gammasolver(&gammasolver_);
        model_->finalize_thetasolver(&thetasolver_);
    }
    TSSolver(const TSSolver &) = delete;
    TSSolver &operator=(const TSSolver &) = delete;

    std::string get_name() const override { return "TSSolver<HipVector>"; }

    // set_solution_theta: theta is given and never updated
    void set_solution_theta(const std::vector<double> &theta) {
        tsdata_->set_thetavector(theta);
        thetasolved_ = true;
    }
    // start with a gamma step from the theta already set (the first theta update of the
    // reference would collapse a uniform gamma0 to equal thetas)
    void set_skip_first_theta(bool s) { skip_first_theta_ = s; }

    // TSSolver::set_annealing (tssolver.h:486-488): runs of the alternation per solve
    void set_annealing(int annealing) { annealing_ = annealing; }
    void set_restart_seed(uint64_t seed) { restart_seed_ = seed; }

    // tssolver.h:491-715.  With annealing > 1 every run after the first restarts from the
    // random gamma of GeneralVector::set_random, which re-creates its generator with seed 13 at
    // every call (generalvector.cpp:7-29) -- here pasc_generate_uniform(restart_seed) -- and the
    // gamma and theta of the run with the smallest AIC are kept (tssolver.h:669-692).  The best
    // run is the best of this solve.
    void solve() override {
        timer_solve.start();
        const int runs = annealing_ > 1 ? annealing_ : 1;
        std::vector<double> gamma_best, theta_best, gamma_rand;
        double aic_best = std::numeric_limits<double>::max();
        if (runs > 1) {
            gamma_rand.resize((size_t)(tsdata_->get_T() * tsdata_->get_R() * tsdata_->get_K()));
            PASC_TRY(pasc_generate_uniform(restart_seed_, (int64_t)gamma_rand.size(),
                                           gamma_rand.data()));
        }
        for (int a = 0; a < runs; a++) {
            double L = std::numeric_limits<double>::max(), L_old, deltaL = L;
            int it;
            for (it = 0; it < maxit; it++) {
                if (!thetasolved_ && !(a == 0 && it == 0 && skip_first_theta_)) {
                    timer_theta_update.start();
                    model_->updatebeforesolve_thetasolver(thetasolver_);
                    timer_theta_update.stop();
                    thetasolver_->solve();
                    model_->updateaftersolve_thetasolver(thetasolver_);
                }
                timer_gamma_update.start();
                model_->updatebeforesolve_gammasolver(gammasolver_);
                timer_gamma_update.stop();
                timer_gamma_solve.start();
                gammasolver_->solve();
                timer_gamma_solve.stop();
                model_->updateaftersolve_gammasolver(gammasolver_);

                L_old = L;
                L = model_->get_L(gammasolver_, thetasolver_);
                deltaL = std::fabs(L - L_old);
                if (debugmode > 0)
                    std::cout << " outer loop: it = " << it << ", L = " << L
                              << ", |L - L_old| = " << deltaL << std::endl;
                if (L - L_old > 0) {
                    std::cout << "ERROR: objective function increased" << std::endl;  // tssolver.h:637-639
                    increased_++;
                }
                if (deltaL < eps) break;  // before it++: it_last is the breaking iteration
            }
            it_sum_ += it;
            it_last_ = it;
            const double aic = model_->get_aic(L);
            if (debugmode > 0)
                std::cout << "  AIC=" << aic << ", L=" << L << ", it=" << it << std::endl;
            if (runs <= 1 || aic < aic_best) {
                if (runs > 1) {
                    gamma_best = tsdata_->get_gammavector();
                    theta_best = tsdata_->get_thetavector();
                }
                aic_best = aic;
                tsdata_->set_aic(aic);
                L_ = L;
                deltaL_ = deltaL;
            }
            if (a < runs - 1) tsdata_->set_gammavector(gamma_rand);
        }
        if (runs > 1) {
            tsdata_->set_gammavector(gamma_best);
            tsdata_->set_thetavector(theta_best);
        }
        timer_solve.stop();
    }

    double get_L() const { return L_; }
    double get_deltaL() const { return deltaL_; }
    int get_it() const override { return it_sum_; }
    int get_increased() const { return increased_; }
    GeneralSolver *get_gammasolver() const { return gammasolver_; }
    GeneralSolver *get_thetasolver() const { return thetasolver_; }

    void printstatus(std::ostream &o) const override {
        o << get_name() << std::endl;
        o << " - it:      " << it_last_ << std::endl;
        o << " - L:       " << L_ << std::endl;
        o << " - deltaL:  " << deltaL_ << std::endl;
        gammasolver_->printstatus(o);
    }
    void printtimer(std::ostream &o) const override {
        o << get_name() << std::endl;
        o << " - t_solve =        " << timer_solve.get_value_sum() << std::endl;
        o << " - t_gamma_update = " << timer_gamma_update.get_value_sum() << std::endl;
        o << " - t_gamma_solve =  " << timer_gamma_solve.get_value_sum() << std::endl;
        o << " - t_theta_update = " << timer_theta_update.get_value_sum() << std::endl;
        gammasolver_->printtimer(o);
    }
    void printshort(std::ostringstream &header, std::ostringstream &values) const override {
        header << "TSSolver it, TSSolver L, TSSolver t all, t gamma solve, ";
        values << it_last_ << ", " << L_ << ", " << timer_solve.get_value_last() << ", "
               << timer_gamma_solve.get_value_sum() << ", ";
        gammasolver_->printshort(header, values);
    }

  private:
    data::TSData<HipVector> *tsdata_;
    model::TSModel<HipVector> *model_;
    GeneralSolver *gammasolver_ = nullptr;
    GeneralSolver *thetasolver_ = nullptr;
    bool thetasolved_ = false;
    bool skip_first_theta_ = false;
    int annealing_ = 1;
    uint64_t restart_seed_ = 13;
    double L_ = std::numeric_limits<double>::max();
    double deltaL_ = std::numeric_limits<double>::max();
    int it_last_ = 0, it_sum_ = 0, increased_ = 0;
    Timer timer_solve, timer_gamma_solve, timer_gamma_update, timer_theta_update;
};

}  // namespace solver
}  // namespace pascinference
