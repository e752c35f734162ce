// spvi — drop-in for the reference solver executable (src/spvi_2.cu:730-796).
//
//   spvi [--thresh X] [--max-sweeps N] [--graph]
//
// Reads temp_modelOp_dirName.txt (5 lines, :741-747) and the model files it names
// (master_cooS1/S2 int32 or int64, master_cooVal, master_R, DP_relv_params; :400-410),
// solves on the GPU, and writes data_solverOutput/<prob_type>/<prob_name>/<prob_specs>/
// [<alpha>/ for custom1] policy.npy (<u4 [Ns,1]) and value_function.npy (<f4 [Ns,1])
// (:749-768, :790-791). The default threshold is the reference's 0.1 (:757).
#include <chrono>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/e2edp.h"
#include "npy.h"

namespace fs = std::filesystem;

static std::vector<int32_t> as_i32(const npy::Array& a, const std::string& name) {
    std::vector<int32_t> o(a.size());
    if (a.kind() == 'i' && a.itemsize() == 4) {
        std::copy(a.as<int32_t>(), a.as<int32_t>() + a.size(), o.begin());
    } else if (a.kind() == 'i' && a.itemsize() == 8) {
        const int64_t* p = a.as<int64_t>();
        for (int64_t i = 0; i < a.size(); i++) {
            if (p[i] < INT32_MIN || p[i] > INT32_MAX) throw std::runtime_error(name + ": index exceeds int32");
            o[i] = (int32_t)p[i];
        }
    } else {
        throw std::runtime_error(name + ": expected int32 or int64, got " + a.descr);
    }
    return o;
}

int main(int argc, char** argv) {
    try {
        e2e_solve_opts o{};
        o.thresh = 1e-1f;
        o.discount = 1.0f;
        o.sweeps_per_check = 16;
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "--thresh" && i + 1 < argc) o.thresh = std::stof(argv[++i]);
            else if (a == "--max-sweeps" && i + 1 < argc) o.max_sweeps = std::stoi(argv[++i]);
            else if (a == "--graph") o.use_graph = 1;
            else if (a == "-h" || a == "--help") {
                std::cout << "usage: spvi [--thresh X] [--max-sweeps N] [--graph]\n";
                return 0;
            } else throw std::runtime_error("unknown argument " + a);
        }
        std::ifstream tf("temp_modelOp_dirName.txt");
        if (!tf) throw std::runtime_error("temp_modelOp_dirName.txt not found (run build_model first)");
        std::string prob_type, prob_name, prob_specs, alpha_str, model_path;
        std::getline(tf, prob_type);
        std::getline(tf, prob_name);
        std::getline(tf, prob_specs);
        std::getline(tf, alpha_str);
        std::getline(tf, model_path);
        std::string results = "data_solverOutput/" + prob_type + "/" + prob_name + "/" + prob_specs + "/";
        if (prob_type == "custom1") results += alpha_str + "/";
        fs::create_directories(results);

        auto t0 = std::chrono::steady_clock::now();
        npy::Array aS1 = npy::load(model_path + "master_cooS1.npy");
        npy::Array aS2 = npy::load(model_path + "master_cooS2.npy");
        npy::Array aV = npy::load(model_path + "master_cooVal.npy");
        npy::Array aR = npy::load(model_path + "master_R.npy");
        npy::Array aD = npy::load(model_path + "DP_relv_params.npy");
        std::vector<int32_t> S1 = as_i32(aS1, "master_cooS1"), S2 = as_i32(aS2, "master_cooS2");
        int64_t ns, na;
        if (aD.kind() == 'i' && aD.itemsize() == 8) {
            ns = aD.as<int64_t>()[0];
            na = aD.as<int64_t>()[1];
        } else {
            ns = aD.as<int32_t>()[0];
            na = aD.as<int32_t>()[1];
        }
        if (aV.descr != "<f4" || aR.descr != "<f4") throw std::runtime_error("cooVal / R must be float32");
        if (aV.size() != (int64_t)S1.size() || aS2.size() != aS1.size())
            throw std::runtime_error("S1, S2, val lengths differ");
        if (aR.size() != ns * na) throw std::runtime_error("master_R has the wrong length");
        std::cout << "s_Ns = " << ns << "\ns_Na = " << na << "\ns_nnz = " << S1.size() << std::endl;

        void* model = nullptr;
        if (e2e_model_import_coo(ns, na, (int64_t)S1.size(), S1.data(), S2.data(), aV.as<float>(),
                                 aR.as<float>(), nullptr, &model) != E2E_OK)
            throw std::runtime_error(e2e_last_error());
        std::vector<uint32_t> policy(ns);
        std::vector<float> value(ns);
        e2e_solve_stats st{};
        e2e_status rc = e2e_vi_solve(model, &o, policy.data(), value.data(), &st);
        e2e_model_destroy(model);
        if (rc != E2E_OK && rc != E2E_ERR_NOT_CONVERGED) throw std::runtime_error(e2e_last_error());
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "num_iterations" << st.sweeps << "\nfinal residual " << st.final_residual
                  << (rc == E2E_OK ? "" : " (not converged)") << "\n---- total solve time = " << secs << std::endl;
        npy::save(results + "policy.npy", policy.data(), "<u4", {ns, 1});
        npy::save(results + "value_function.npy", value.data(), "<f4", {ns, 1});
        std::cout << "saved " << results << "{policy,value_function}.npy" << std::endl;
        return rc == E2E_OK ? 0 : 3;
    } catch (const std::exception& e) {
        std::cerr << "spvi: " << e.what() << std::endl;
        return 1;
    }
}
