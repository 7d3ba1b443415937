// test_signal1d: the 1-D FEM-H1 K-means driver over libpasc_hip.so, with the flags and the
// epsilon loop of examples/test_signal1D/test_signal1D.cpp:17-425 -- sorted epsilon list,
// warm start from the previous gamma, absolute error against the true signal, best epsilon by
// abserr, shortinfo CSV and the best solution saved as PETSc binary files.
//
//   test_signal1d --test_filename=data.bin --test_filename_solution=truth.bin --test_K=2
//                 --test_epssqr=1 --test_epssqr=10 [--test_Theta=1 --test_Theta=2]
//                 [--test_precision=fp32] [--spgqpsolver_* ...] [--tssolver_* ...]
#include <pasc/pasc.h>

#include <algorithm>
#include <iomanip>

using namespace pascinference;

int main(int argc, char **argv) {
    Options opt(argc, argv);
    std::vector<double> epssqr_list;
    if (!opt.set_option_list("test_epssqr", &epssqr_list)) {
        std::cout << "test_epssqr has to be set!" << std::endl;
        return 1;
    }
    std::sort(epssqr_list.begin(), epssqr_list.end());
    int K, device, annealing;
    std::string filename, filename_solution, filename_out, shortinfo_filename, precision;
    bool saveresult, shortinfo, skip_first_theta;
    opt.set_option_value("test_K", &K, 2);
    opt.set_option_value("test_annealing", &annealing, 1);
    opt.set_option_value("test_filename", &filename, std::string("data/samplesignal.bin"));
    opt.set_option_value("test_filename_solution", &filename_solution,
                         std::string("data/samplesignal_solution.bin"));
    opt.set_option_value("test_filename_out", &filename_out, std::string("samplesignal"));
    opt.set_option_value("test_shortinfo_filename", &shortinfo_filename,
                         std::string("shortinfo_samplesignal.txt"));
    opt.set_option_value("test_saveresult", &saveresult, true);
    opt.set_option_value("test_shortinfo", &shortinfo, true);
    opt.set_option_value("test_precision", &precision, std::string("fp64"));
    opt.set_option_value("test_skip_first_theta", &skip_first_theta, true);
    opt.set_option_value("test_device", &device, 0);
    // test_signal1D.cpp:81-82: 0 = Fem (cell sums), 1 = FemHat; fem_reduce 1 = no reduction
    int fem_type_ref;
    double fem_reduce;
    opt.set_option_value("test_fem_type", &fem_type_ref, 1);
    opt.set_option_value("test_fem_reduce", &fem_reduce, 1.0);
    if (annealing < 1) {
        std::cout << "test_annealing has to be >= 1" << std::endl;
        return 1;
    }
    std::vector<double> theta_list;
    const bool given_theta = opt.set_option_list("test_Theta", &theta_list);
    if (given_theta && (int)theta_list.size() != K) {
        std::cout << "number of provided Theta solutions is different then number of clusters!"
                  << std::endl;
        return 1;
    }

    pasc_ctx *ctx = nullptr;
    if (!PASC_TRY(pasc_ctx_create(device, &ctx))) return 2;
    int rc = 0;
    {
        const int prec = precision == "fp32" ? PASC_FP32_BULK : PASC_FP64;
        data::Signal1DData<HipVector> mydata(ctx, filename, K, prec, fem_reduce,
                                             fem_type_ref == 0 ? PASC_FEM_CONSTANT : PASC_FEM_HAT);
        if (!mydata.ok()) return 2;
        std::vector<double> solution = petscio::read_vec(filename_solution);
        std::cout << "- PROBLEM INFO ----------------------------" << std::endl;
        std::cout << " T = " << mydata.get_T() << ", K = " << K << ", precision = " << precision
                  << ", fem_reduce = " << fem_reduce << " (type " << fem_type_ref << ")" << std::endl;

        model::GraphH1FEMModel<HipVector> mymodel(mydata, epssqr_list[0], opt);
        solver::TSSolver<HipVector> mysolver(mydata, mymodel, opt);
        mysolver.set_annealing(annealing);  // TSSolver<PetscVector> mysolver(mydata, annealing), test_signal1D.cpp:237
        // gamma0: uniform (the reference draws PETSc random numbers with seed 13)
        PASC_TRY(pasc_set_gamma_uniform(mydata.get_problem()));
        if (given_theta) {
            mysolver.set_solution_theta(theta_list);
        } else {
            // initial theta: K levels spread over the data range, then a gamma step first
            std::vector<double> th(K);
            const auto &x = mydata.get_datavector();
            const double lo = *std::min_element(x.begin(), x.end());
            const double hi = *std::max_element(x.begin(), x.end());
            for (int k = 0; k < K; k++) th[k] = lo + (hi - lo) * (k + 0.5) / K;
            mydata.set_thetavector(th);
            mysolver.set_skip_first_theta(skip_first_theta);
        }

        double abserr_best = std::numeric_limits<double>::max(), epssqr_best = -1;
        std::vector<double> gamma_best, theta_best;
        std::ostringstream head, vals;
        std::ofstream sfile;
        if (shortinfo) sfile.open(shortinfo_filename);
        for (size_t depth = 0; depth < epssqr_list.size(); depth++) {
            const double epssqr = epssqr_list[depth];
            std::cout << "--- SOLVING THE PROBLEM with epssqr = " << epssqr << " ---" << std::endl;
            mymodel.set_epssqr(epssqr);
            mysolver.solve();
            const double abserr = mydata.compute_abserr_reconstructed(solution);
            std::cout << " - abserr = " << abserr << std::endl;
            if (shortinfo) {
                head.str("");
                vals.str("");
                head << "K,epssqr,abserr,";
                vals << K << "," << epssqr << "," << abserr << ",";
                for (int k = 0; k < K; k++) head << "Theta" << k << ",";
                vals << mydata.print_thetavector();
                mysolver.printshort(head, vals);
                if (depth == 0) sfile << head.str() << "\n";
                sfile << vals.str() << "\n";
            }
            if (abserr < abserr_best) {
                abserr_best = abserr;
                epssqr_best = epssqr;
                gamma_best = mydata.get_gammavector();
                theta_best = mydata.get_thetavector();
            }
        }
        mydata.set_gammavector(gamma_best);
        mydata.set_thetavector(theta_best);
        if (saveresult) {
            std::cout << "--- SAVING OUTPUT ---" << std::endl;
            std::cout << " - with best epssqr = " << epssqr_best << std::endl;
            mydata.saveSignal1D(filename_out);
        }
        std::cout << "--- THETA SOLUTION ---" << std::endl << mydata.print_thetavector() << std::endl;
        std::cout << "--- TIMERS INFO ---" << std::endl;
        mysolver.printtimer(std::cout);
        std::cout << "--- FINAL SOLVER INFO ---" << std::endl;
        mysolver.printstatus(std::cout);
        std::cout << std::setprecision(17) << "best_epssqr=" << epssqr_best
                  << " abserr=" << abserr_best << std::endl;
    }
    pasc_ctx_destroy(ctx);
    return rc;
}
