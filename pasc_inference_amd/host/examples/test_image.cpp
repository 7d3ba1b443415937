// test_image: the image FEM-H1 K-means driver over libpasc_hip.so, with the flags and the
// epsilon loop of examples/test_image/test_image.cpp:17-340 -- a grayscale image on the
// 4-neighbour grid (T = 1 step of width x height nodes), data cut to [0, 1] (or scaled to
// [-1, 1]) before the solves, a sorted epsilon list solved in turn from the previous gamma and
// theta, and per epsilon the image, gamma and recovered image saved as PETSc binary files plus
// one shortinfo CSV line.
//
//   test_image --test_image_filename=usi_250_150_02.bin --test_width=250 --test_height=150
//              --test_K=2 --test_epssqr=1e-5 --test_epssqr=1e-4 [--test_Theta=0 --test_Theta=1]
//              [--test_results_dir=results] [--test_precision=fp32] [--spgqpsolver_* ...]
//
// Not supported: test_fem_reduce < 1 (the FEM coarsening of this library is along the time
// axis of 1-D signals) and test_graph_save (VTK output of the graph).
#include <pasc/pasc.h>

#include <algorithm>
#include <filesystem>
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
    int K, annealing, width, height, fem_type, device;
    double fem_reduce;
    bool cutgamma, scaledata, cutdata, printstats, shortinfo, graph_save;
    std::string image_filename, image_out, shortinfo_filename, shortinfo_header, shortinfo_values,
        results_dir, precision;
    opt.set_option_value("test_K", &K, 2);
    opt.set_option_value("test_fem_type", &fem_type, 3);
    opt.set_option_value("test_fem_reduce", &fem_reduce, 1.0);
    opt.set_option_value("test_width", &width, 250);
    opt.set_option_value("test_height", &height, 150);
    opt.set_option_value("test_graph_save", &graph_save, false);
    opt.set_option_value("test_cutgamma", &cutgamma, false);
    opt.set_option_value("test_scaledata", &scaledata, false);
    opt.set_option_value("test_cutdata", &cutdata, true);
    opt.set_option_value("test_printstats", &printstats, false);
    opt.set_option_value("test_annealing", &annealing, 1);
    opt.set_option_value("test_image_filename", &image_filename,
                         std::string("data/usi_text/usi_250_150_02.bin"));
    opt.set_option_value("test_image_out", &image_out, std::string("usi_test_250_150_02"));
    opt.set_option_value("test_shortinfo", &shortinfo, true);
    opt.set_option_value("test_shortinfo_header", &shortinfo_header, std::string(""));
    opt.set_option_value("test_shortinfo_values", &shortinfo_values, std::string(""));
    opt.set_option_value("test_shortinfo_filename", &shortinfo_filename,
                         std::string("shortinfo/usi_250_150_02.txt"));
    opt.set_option_value("test_results_dir", &results_dir, std::string("results"));
    opt.set_option_value("test_precision", &precision, std::string("fp64"));
    opt.set_option_value("test_device", &device, 0);
    if (annealing < 1) {
        std::cout << "test_annealing has to be >= 1" << std::endl;
        return 1;
    }
    if (fem_reduce > 0.0 && fem_reduce < 1.0) {
        std::cout << "test_fem_reduce < 1 is not supported for images" << std::endl;
        return 1;
    }
    if (graph_save) std::cout << "test_graph_save: VTK output is not supported, ignored" << std::endl;
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
    try {
        const int prec = precision == "fp32" ? PASC_FP32_BULK : PASC_FP64;
        // test_image.cpp:156-179: the 2-D grid and the decomposition of T = 1 step of its nodes
        Decomposition<HipVector> decomposition(1, (int64_t)width * height, K, 1);
        data::ImageData<HipVector> mydata(ctx, decomposition, image_filename, width, height, prec);
        if (!mydata.ok()) {
            pasc_ctx_destroy(ctx);
            return 2;
        }
        std::cout << "- PROBLEM INFO ----------------------------" << std::endl;
        std::cout << " image = " << image_filename << ", " << width << " x " << height
                  << ", K = " << K << ", precision = " << precision << std::endl;
        if (printstats) mydata.printstats(std::cout);
        if (cutdata) mydata.cutdata(0, 1);
        if (scaledata) mydata.scaledata(-1, 1, 0, 1);

        model::GraphH1FEMModel<HipVector> mymodel(mydata, epssqr_list[0], opt);
        solver::TSSolver<HipVector> mysolver(mydata, mymodel, opt);
        mysolver.set_annealing(annealing);
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
            mysolver.set_skip_first_theta(true);
        }

        std::filesystem::create_directories(results_dir);
        std::ofstream sfile;
        if (shortinfo) {
            const std::filesystem::path sp(shortinfo_filename);
            if (sp.has_parent_path()) std::filesystem::create_directories(sp.parent_path());
            sfile.open(shortinfo_filename);
        }
        for (size_t depth = 0; depth < epssqr_list.size(); depth++) {
            const double epssqr = epssqr_list[depth];
            if (depth > 0) {
                mymodel.set_epssqr(epssqr);
                if (scaledata) mydata.scaledata(-1, 1, 0, 1);
            }
            std::cout << "--- SOLVING THE PROBLEM with epssqr = " << epssqr << " ---" << std::endl;
            mysolver.solve();
            if (cutgamma) mydata.cutgamma();
            if (scaledata) mydata.scaledata(0, 1, -1, 1);
            std::cout << "--- SAVING OUTPUT ---" << std::endl;
            std::ostringstream name;
            name << results_dir << "/" << image_out << "_epssqr" << epssqr;
            mydata.saveImage(name.str(), depth == 0);
            if (shortinfo) {
                std::ostringstream head, vals;
                head << shortinfo_header << "width,height,K,depth,epssqr,";
                vals << shortinfo_values << width << "," << height << "," << K << "," << depth
                     << "," << epssqr << ",";
                for (int k = 0; k < K; k++) head << "Theta" << k << ",";
                vals << mydata.print_thetavector();
                mysolver.printshort(head, vals);
                if (depth == 0) sfile << head.str() << "\n";
                sfile << vals.str() << "\n";
            }
        }
        std::cout << "--- THETA SOLUTION ---" << std::endl << mydata.print_thetavector() << std::endl;
        std::cout << "--- TIMERS INFO ---" << std::endl;
        mysolver.printtimer(std::cout);
        std::cout << "--- FINAL SOLVER INFO ---" << std::endl;
        mysolver.printstatus(std::cout);
    } catch (const std::exception &ex) {
        std::cerr << "test_image: " << ex.what() << std::endl;
        rc = 2;
    }
    pasc_ctx_destroy(ctx);
    return rc;
}
