// build_model — drop-in for the reference builder executable (src/build_model.cu:926-1238).
//
//   build_model [ALPHA_x100] [--params FILE] [--int32] [--dry-run] [--slot-k K]
//
// Same inputs and outputs as the reference:
//   reads  data_input/<prob_name>/{all_u_mat,all_v_mat,all_ui_mat,all_vi_mat,all_Yi,
//          all_s_mat,obstacle_mask}.npy                                   (:978-985)
//   writes data_modelOutput/<prob_type>/<prob_name>/<prob_specs>/[<alpha>/]
//          master_cooS1.npy master_cooS2.npy master_cooVal.npy master_R.npy
//          DP_relv_params.npy prob_params.npy                             (:1305-1310)
//   and temp_modelOp_dirName.txt (5 lines, :969-975) for the solver.
// The parameters come from a runtime file in the syntax of src/input_to_build_model.h
// (default: ./input_to_build_model.h, else ./input_to_build_model.txt). --int32 writes
// the S1/S2 arrays as int32 directly (src/convert_model_datatype.py:19-23).
#include <chrono>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/e2edp.h"
#include "npy.h"
#include "params.h"

namespace fs = std::filesystem;

// get_reward_type, src/build_model.cu:898-918
static int reward_type_of(const std::string& t) {
    if (t == "time") return 0;
    if (t == "energy1") return 1;
    if (t == "energy2") return 2;
    if (t == "energy3") return 3;
    if (t == "custom1") return 4;
    return -1;
}

static npy::Array load_checked(const std::string& path, char kind, int64_t expect) {
    npy::Array a = npy::load(path);
    if (a.kind() != kind || a.itemsize() != 4)
        throw std::runtime_error(path + ": expected " + (kind == 'f' ? "float32" : "int32") + ", got " + a.descr);
    if (a.size() != expect)
        throw std::runtime_error(path + ": " + std::to_string(a.size()) + " elements, expected " +
                                 std::to_string(expect));
    return a;
}

int main(int argc, char** argv) {
    try {
        std::string pfile;
        bool as_i32 = false, dry = false;
        int slot_k = 0;
        std::string alpha_arg;
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "--params" && i + 1 < argc) pfile = argv[++i];
            else if (a == "--int32") as_i32 = true;
            else if (a == "--dry-run") dry = true;
            else if (a == "--slot-k" && i + 1 < argc) slot_k = std::stoi(argv[++i]);
            else if (a == "-h" || a == "--help") {
                std::cout << "usage: build_model [ALPHA_x100] [--params FILE] [--int32] [--dry-run]\n";
                return 0;
            } else alpha_arg = a;
        }
        if (pfile.empty()) pfile = fs::exists("input_to_build_model.h") ? "input_to_build_model.h" : "input_to_build_model.txt";
        params::Parser P;
        P.parse_file(pfile);

        const std::string prob_type = P.str("prob_type"), prob_name = P.str("prob_name");
        float alpha = (float)P.num("alpha_header", 0.0);
        if (!alpha_arg.empty()) alpha = std::stof(alpha_arg) / 100;  // :936-943
        const int reward_type = reward_type_of(prob_type);
        if (reward_type < 0) throw std::runtime_error("unknown prob_type " + prob_type);

        e2e_params p{};
        p.gsize = (int32_t)P.num("gsize");
        p.nt = (int32_t)P.num("nt");
        p.nrzns = (int32_t)P.num("nrzns");
        p.nmodes = (int32_t)P.num("nmodes");
        p.num_ac_speeds = (int32_t)P.num("num_ac_speeds");
        p.num_ac_angles = (int32_t)P.num("num_ac_angles");
        p.i_term = (int32_t)P.num("i_term");
        p.j_term = (int32_t)P.num("j_term");
        p.term_subgrid_size = (int32_t)P.num("term_subgrid_size");
        p.reward_type = reward_type;
        p.neighb_gsize = (int32_t)P.num("neighb_gsize", 21);
        p.is_stationary = (int32_t)P.num("is_stationary", 0);
        p.F = (float)P.num("F");
        p.dt = (float)P.num("dt");
        p.dx = (float)P.num("dx");
        p.dy = (float)P.num("dy", p.dx);
        p.x0 = (float)P.num("x0", p.dx / 2);
        p.y0 = (float)P.num("y0", p.dy / 2);
        p.r_outbound = (float)P.num("r_outbound");
        p.r_terminal = (float)P.num("r_terminal");
        p.alpha = alpha;
        p.z = (float)P.num("z", -9999);
        if (p.neighb_gsize > p.gsize) throw std::runtime_error("neighb_gsize > gsize (build_model.cu:1027)");

        // output locations, :953-975
        const std::string prob_specs = "a" + std::to_string(p.num_ac_speeds) + "x" + std::to_string(p.num_ac_angles) +
                                       "_i" + std::to_string(p.i_term) + "_j" + std::to_string(p.j_term) + "_ref" +
                                       std::to_string(p.term_subgrid_size);
        std::string out = "data_modelOutput/" + prob_type + "/" + prob_name + "/" + prob_specs + "/";
        if (reward_type > 3) out += std::to_string(alpha) + "/";
        fs::create_directories(out);
        {
            std::ofstream tf("temp_modelOp_dirName.txt");
            tf << prob_type << "\n" << prob_name << "\n" << prob_specs << "\n" << std::to_string(alpha) << "\n" << out;
        }

        const int64_t g = p.gsize, nt = p.nt, nc = g * g, nm = p.nmodes, nr = p.nrzns;
        const int64_t na = (int64_t)p.num_ac_speeds * p.num_ac_angles, ns = nc * nt;
        std::vector<float> pp(32);
        e2e_params_pack(&p, pp.data());
        int64_t dp[2] = {ns, na};

        const std::string in = "data_input/" + prob_name + "/";
        npy::Array u = load_checked(in + "all_u_mat.npy", 'f', nt * nc);
        npy::Array v = load_checked(in + "all_v_mat.npy", 'f', nt * nc);
        npy::Array ui = load_checked(in + "all_ui_mat.npy", 'f', nt * nm * nc);
        npy::Array vi = load_checked(in + "all_vi_mat.npy", 'f', nt * nm * nc);
        npy::Array yi = load_checked(in + "all_Yi.npy", 'f', nt * nr * nm);
        npy::Array sf = load_checked(in + "all_s_mat.npy", 'f', nt * nc);
        npy::Array mk = load_checked(in + "obstacle_mask.npy", 'i', nt * nc);
        std::cout << "Finished reading Velocity Field Data !" << std::endl;

        npy::save(out + "DP_relv_params.npy", dp, "<i8", {2, 1});
        npy::save(out + "prob_params.npy", pp.data(), "<f4", {32, 1});
        if (dry) {
            std::cout << "dry run: wrote " << out << "{DP_relv_params,prob_params}.npy\n";
            return 0;
        }

        e2e_fields f{u.as<float>(), v.as<float>(), ui.as<float>(), vi.as<float>(),
                     yi.as<float>(), sf.as<float>(), mk.as<int32_t>()};
        e2e_build_opts o{};
        o.slot_k = slot_k;
        void* model = nullptr;
        auto t0 = std::chrono::steady_clock::now();
        if (e2e_model_build(&p, &f, &o, &model) != E2E_OK) throw std::runtime_error(e2e_last_error());
        e2e_model_info info{};
        e2e_model_info_get(model, &info);
        const int64_t nnz = info.nnz;
        std::vector<float> val(nnz), R(na * ns);
        if (as_i32) {
            std::vector<int32_t> S1(nnz), S2(nnz);
            if (e2e_model_export_coo_i32(model, S1.data(), S2.data(), val.data(), R.data()) != E2E_OK)
                throw std::runtime_error(e2e_last_error());
            npy::save(out + "master_cooS1.npy", S1.data(), "<i4", {nnz, 1});
            npy::save(out + "master_cooS2.npy", S2.data(), "<i4", {nnz, 1});
        } else {
            std::vector<int64_t> S1(nnz), S2(nnz);
            if (e2e_model_export_coo(model, S1.data(), S2.data(), val.data(), R.data()) != E2E_OK)
                throw std::runtime_error(e2e_last_error());
            npy::save(out + "master_cooS1.npy", S1.data(), "<i8", {nnz, 1});
            npy::save(out + "master_cooS2.npy", S2.data(), "<i8", {nnz, 1});
        }
        npy::save(out + "master_cooVal.npy", val.data(), "<f4", {nnz, 1});
        npy::save(out + "master_R.npy", R.data(), "<f4", {na * ns, 1});
        e2e_model_destroy(model);
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "nnz = " << nnz << "  lanes = " << info.lanes << "  out-of-window lanes = "
                  << info.n_out_of_window << "\nbuild " << info.build_ms << " ms (move kernel "
                  << info.move_kernel_ms << " ms), total with export " << secs << " s\nsaved files \n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "build_model: " << e.what() << std::endl;
        return 1;
    }
}
