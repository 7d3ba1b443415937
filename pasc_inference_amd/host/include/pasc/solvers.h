// GeneralSolver / QPSolver / SimpleSolver / SPGQPSolver over the HipVector backend
// (include/general/solver/generalsolver.h:25-141, qpsolver.h:24-92, simplesolver.h,
// spgqpsolver.h:47-202; the PETSc specialisation src/petscvector/solver/spgqpsolver.cpp).
#pragma once

#include "common.h"
#include "tsdata.h"

namespace pascinference {
namespace solver {

class GeneralSolver {
  public:
    virtual ~GeneralSolver() {}
    virtual void solve() {}
    virtual std::string get_name() const { return "GeneralSolver"; }
    virtual int get_it() const { return 0; }
    virtual int get_maxit() const { return maxit; }
    virtual void set_maxit(int m) { maxit = m; }
    virtual double get_eps() const { return eps; }
    virtual void set_eps(double e) { eps = e; }
    virtual int get_debugmode() const { return debugmode; }
    virtual void set_debugmode(int d) { debugmode = d; }
    virtual void printstatus(std::ostream &o) const { o << get_name() << std::endl; }
    virtual void printtimer(std::ostream &o) const { o << get_name() << std::endl; }
    virtual void printshort(std::ostringstream &header, std::ostringstream &values) const {
        (void)header;
        (void)values;
    }

  protected:
    int maxit = 0;
    double eps = 0.0;
    int debugmode = 0;
};

template <class VectorBase> class QPSolver : public GeneralSolver {
  public:
    virtual double get_fx() const { return fx; }
    // the reference returns the cumulative count (qpsolver.h:260-263)
    int get_it() const override { return it_sum; }
    virtual int get_it_last() const { return it_last; }
    virtual int get_hessmult() const { return hessmult_sum; }

  protected:
    int it_sum = 0, hessmult_sum = 0, it_last = 0, hessmult_last = 0;
    double fx = std::numeric_limits<double>::max();
};

// theta "solver" of the k-means model: solve() does nothing (simplesolver.h:208-216)
template <class VectorBase> class SimpleSolver : public GeneralSolver {
  public:
    std::string get_name() const override { return "SimpleSolver"; }
};

template <class VectorBase> class SPGQPSolver;

template <> class SPGQPSolver<HipVector> : public QPSolver<HipVector> {
  public:
    explicit SPGQPSolver(data::TSData<HipVector> &tsdata, const Options &opt = Options())
        : tsdata_(&tsdata) {
        set_settings_from_console(opt);
        timer_solve.restart();
    }

    // spgqpsolver.h:161-202 (flag names and defaults of the reference)
    void set_settings_from_console(const Options &o) {
        pasc_spg_default_params(&p_);
        o.set_option_value("spgqpsolver_maxit", &maxit, (int)p_.maxit);
        o.set_option_value("spgqpsolver_eps", &eps, p_.eps);
        int m;
        o.set_option_value("spgqpsolver_m", &m, (int)p_.m);
        o.set_option_value("spgqpsolver_gamma", &p_.gamma, p_.gamma);
        o.set_option_value("spgqpsolver_sigma1", &p_.sigma1, p_.sigma1);
        o.set_option_value("spgqpsolver_sigma2", &p_.sigma2, p_.sigma2);
        o.set_option_value("spgqpsolver_alphainit", &p_.alphainit, p_.alphainit);
        bool normgp, anormgp, normgp_normb, anormgp_normb, difff, kkt;
        o.set_option_value("spgqpsolver_stop_normgp", &normgp, false);
        o.set_option_value("spgqpsolver_stop_Anormgp", &anormgp, false);
        o.set_option_value("spgqpsolver_stop_normgp_normb", &normgp_normb, false);
        o.set_option_value("spgqpsolver_stop_Anormgp_normb", &anormgp_normb, false);
        o.set_option_value("spgqpsolver_stop_difff", &difff, true);
        o.set_option_value("spgqpsolver_stop_kkt", &kkt, false);
        o.set_option_value("spgqpsolver_kkt_tol", &p_.kkt_tol, p_.kkt_tol);
        o.set_option_value("spgqpsolver_debugmode", &debugmode, 0);
        int graph;
        o.set_option_value("spgqpsolver_use_graph", &graph, 0);
        o.set_option_value("spgqpsolver_check_every", &p_.check_every, p_.check_every);
        p_.m = m;
        p_.use_graph = graph;
        auto bit = [](bool on, unsigned b) { return on ? b : 0u; };
        p_.stop_mask = bit(difff, PASC_STOP_DIFFF) | bit(normgp, PASC_STOP_NORMGP) |
                       bit(normgp_normb, PASC_STOP_NORMGP_NORMB) | bit(anormgp, PASC_STOP_ANORMGP) |
                       bit(anormgp_normb, PASC_STOP_ANORMGP_NORMB) | bit(kkt, PASC_STOP_KKT);
    }
    pasc_spg_params &params() { return p_; }

    // SPGQPSolver<PetscVector>::solve (spgqpsolver.cpp:55-312): the whole iteration runs on
    // the device; x is the model's gamma (warm start, x0 aliases x)
    void solve() override {
        p_.maxit = maxit;
        p_.eps = eps;
        timer_solve.start();
        pasc_spg_result r;
        std::memset(&r, 0, sizeof(r));
        ok_ = PASC_TRY(pasc_gamma_solve(tsdata_->get_problem(), &p_, &r));
        timer_solve.stop();
        last_ = r;
        it_last = r.it;
        hessmult_last = r.hessmult;
        it_sum += r.it;
        hessmult_sum += r.hessmult;
        fx = r.fx;
        gP = r.dd;
        if (debugmode > 0)
            std::cout << "   it = " << r.it << ", fx = " << r.fx << ", kkt = " << r.kkt << std::endl;
    }
    std::string get_name() const override { return "SPGQPSolver for HIP"; }
    double get_gP() const { return gP; }
    const pasc_spg_result &get_result() const { return last_; }
    bool ok() const { return ok_; }

    void printstatus(std::ostream &o) const override {
        o << " - it:          " << it_last << std::endl;
        o << " - hess_mult:   " << hessmult_last << std::endl;
        o << " - fx:          " << fx << std::endl;
        o << " - used norm:   " << gP << std::endl;
        o << " - kkt:         " << last_.kkt << std::endl;
    }
    void printtimer(std::ostream &o) const override {
        o << get_name() << std::endl;
        o << " - it all =        " << it_sum << std::endl;
        o << " - hessmult all =  " << hessmult_sum << std::endl;
        o << " - timers" << std::endl;
        o << "  - t_solve =      " << timer_solve.get_value_sum() << std::endl;
    }
    void printshort(std::ostringstream &header, std::ostringstream &values) const override {
        header << "SPGQP it, SPGQP hessmult, SPGQP t all, SPGQP fx, SPGQP kkt, ";
        values << it_last << ", " << hessmult_last << ", " << timer_solve.get_value_last()
               << ", " << fx << ", " << last_.kkt << ", ";
    }

  private:
    data::TSData<HipVector> *tsdata_;
    pasc_spg_params p_;
    pasc_spg_result last_{};
    double gP = std::numeric_limits<double>::max();
    Timer timer_solve;
    bool ok_ = true;
};

}  // namespace solver
}  // namespace pascinference
