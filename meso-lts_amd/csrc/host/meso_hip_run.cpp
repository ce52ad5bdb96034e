// meso_hip_run: meso's command-line run of a case (src/meso/main.cpp:47-57,
// src/meso/exec.cpp:141-165) with the step on the GPU through the C ABI of
// libmeso_hip.so. The solver class keeps meso's BasicSolver contract
// (include/solver/solver.h:22-46): initial(), do_step(), output(),
// get_run_state(), max_step(), write_interval(); main() is handle_solver's loop.
//
// usage: meso_hip_run --case <case-file> [--storage fp64|fp32|fp32x|fp32e] [--out <dir>] [--not-run]
//   --out: directory under which <case-name>/ is written (meso: the working directory)
//   --storage: FP64 (meso's precision), FP32 (4 bytes per value, stochastically
//     rounded: its residual settles near 1e-6, so a residual-limit below that
//     is not reached), FP32X (5 bytes per value, the FP64 kernels; converges
//     with FP64) or FP32E (the FP32X format with the FP32 kernels)
#include "meso_hip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

void check(int rc, const char *what) {
    if (rc != MESO_OK) throw std::runtime_error(std::string(what));
}

std::string setting_str(const meso_case *cs, const char *key, const char *dflt) {
    char buf[1024];
    if (meso_case_setting_str(cs, key, dflt, buf, sizeof(buf)) != MESO_OK) return dflt;
    return buf;
}

int setting_int(const meso_case *cs, const char *key, int dflt) {
    int32_t v = dflt;
    meso_case_setting_int(cs, key, dflt, &v);
    return v;
}

double setting_double(const meso_case *cs, const char *key, double dflt) {
    double v = dflt;
    meso_case_setting_double(cs, key, dflt, &v);
    return v;
}

// read_np_data (src/solver/field.cpp:252-273): one value per line, or "x y z".
void read_np(const std::string &file, double *dst, int n, int width) {
    std::ifstream fp(file);
    if (!fp) throw std::runtime_error("cannot open " + file);
    std::string line;
    int i = 0;
    while (i < n && std::getline(fp, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ss(line);
        for (int k = 0; k < width; ++k)
            if (!(ss >> dst[width * i + k])) throw std::runtime_error("bad line in " + file);
        ++i;
    }
    if (i < n) throw std::runtime_error(file + ": fewer lines than cells");
}

class HipSolver {
  public:
    HipSolver(const std::string &case_path, int storage, const std::string &out_root) : root_(out_root) {
        if (meso_case_open(case_path.c_str(), 0.0, &cs_) != MESO_OK)
            throw std::runtime_error(meso_case_last_error());
        meso_mesh_desc m;
        meso_dvs_desc d;
        meso_step_params p;
        meso_case_mesh(cs_, &m);
        meso_case_dvs(cs_, &d);
        meso_case_params(cs_, &p);
        if (meso_hip_create(&m, &d, &p, storage, &ctx_) != MESO_OK) {
            const std::string e = meso_hip_last_error(nullptr);
            meso_case_close(cs_);
            throw std::runtime_error(e);
        }
        nc_ = m.n_cells;
        dim_ = m.dim;
        model_ = p.model;
        dt_ = p.dt;
        case_name_ = setting_str(cs_, "case-name", "unnamed");
        residual_limit_ = setting_double(cs_, "residual-limit", 1e-6);
        rho_.assign(nc_, 0.0);
        T_.assign(nc_, 0.0);
        u_.assign(3 * size_t(nc_), 0.0);
        q_.assign(3 * size_t(nc_), 0.0);
        m0_.assign(nc_, 0.0);
        m1_.assign(3 * size_t(nc_), 0.0);
        m2_.assign(nc_, 0.0);
        // cell groups with `mask switch True` (src/solver/boundary/patch.cpp:48-51)
        std::vector<int32_t> mask(nc_, 0);
        check(meso_case_cell_mask(cs_, mask.data()), meso_case_last_error());
        check(meso_hip_set_mask(ctx_, mask.data()), meso_hip_last_error(ctx_));
    }
    ~HipSolver() {
        meso_hip_destroy(ctx_);
        meso_case_close(cs_);
    }

    int step = 0;
    double solution_time = 0.0;

    // initial(): the cell-group patches, or read-np-data (cdugks_shakhov.cpp:62-168,
    // cdugks.cpp:41-114, dugks.cpp:41-105)
    void initial() {
        step = 0;
        solution_time = 0.0;
        const std::string dir = root_ + "/" + case_name_;
        if (setting_str(cs_, "read-np-data", "False") == "True") {  // meso's booleans are the literal True
            if (model_ == MESO_MODEL_CDUGKS_SHAKHOV) {
                read_np(dir + "/latest/Rho", rho_.data(), nc_, 1);
                read_np(dir + "/latest/T", T_.data(), nc_, 1);
                read_np(dir + "/latest/U", u_.data(), nc_, 3);
                read_np(dir + "/latest/q", q_.data(), nc_, 3);
            } else if (model_ == MESO_MODEL_CDUGKS_BGK) {
                read_np(dir + "/latest/Rho", rho_.data(), nc_, 1);
                read_np(dir + "/latest/U", u_.data(), nc_, 3);
            } else {
                read_np(dir + "/np-data/Rho.np.dat", rho_.data(), nc_, 1);
                read_np(dir + "/np-data/vel.np.dat", u_.data(), nc_, 3);
            }
            check(meso_hip_init_from_macros(ctx_, rho_.data(), u_.data(), T_.data(), q_.data()),
                  meso_hip_last_error(ctx_));
        } else {
            meso_case_initial_patch(cs_, rho_.data(), u_.data(), T_.data());
            check(meso_hip_init_equilibrium(ctx_, rho_.data(), u_.data(), T_.data()), meso_hip_last_error(ctx_));
        }
        fetch();
    }

    // do_step(): one step, the residual check every 100 steps and the
    // re-read of [settings] (cdugks_shakhov.cpp:724-764)
    void do_step() {
        meso_step_report rep{};
        const int rc = meso_hip_step(ctx_, 1, &rep);
        // cells of masked groups were repaired by the step (inverse-distance
        // average of the neighbours, cdugks_shakhov.cpp:698-719, 766-789)
        if (rep.masked_cells > 0)
            std::cerr << "[WARN] step " << step + 1 << ": " << rep.masked_cells
                      << " masked cell(s) caught rho=nan or T < 0 and were repaired" << std::endl;
        if (rc == MESO_E_NUMERIC) {
            // an unmasked cell: meso logs it and ends the run (run_state = false)
            // after completing the step, whose state is written as the final output
            std::cerr << "[WARN] cell<" << rep.first_bad_cell + 1 << "> caught rho=nan or T < 0 at step "
                      << step + 1 << std::endl;
            failed_ = true;
            run_state_ = false;
            ++step;
            solution_time += dt_;
            return;
        }
        check(rc, meso_hip_last_error(ctx_));
        stepped_ = true;
        ++step;
        solution_time += dt_;
        if (step % residual_interval == 0) {
            double r[8];
            check(meso_hip_residual(ctx_, r), meso_hip_last_error(ctx_));
            // r = rho, u, v, w, T, qx, qy, qz; meso's list order (:733-744)
            std::vector<double> res;
            if (model_ == MESO_MODEL_CDUGKS_SHAKHOV)
                res = dim_ == 2 ? std::vector<double>{r[0], r[4], r[1], r[2], r[5], r[6]}
                                : std::vector<double>{r[0], r[4], r[1], r[2], r[3], r[5], r[6], r[7]};
            else
                res = dim_ == 2 ? std::vector<double>{r[0], r[1], r[2]} : std::vector<double>{r[0], r[1], r[2], r[3]};
            std::cout << "step: " << step;
            for (double x : res) std::cout << " " << x;
            std::cout << std::endl;
            bool conv = true;  // Utils::is_converged, src/core/core.cpp:45-51
            for (double x : res) conv = conv && (x <= residual_limit_);
            converge_state_ = conv ? converge_state_ + 1 : 0;
            if (converge_state_ >= 10) {
                std::cout << "Step " << step << " - Convergence achieved: residual all below " << residual_limit_
                          << std::endl;
                run_state_ = false;
            }
        }
        check(meso_case_reload_settings(cs_), meso_case_last_error());
        meso_step_params p;
        meso_case_params(cs_, &p);
        check(meso_hip_set_params(ctx_, &p), meso_hip_last_error(ctx_));
        dt_ = p.dt;
        residual_limit_ = setting_double(cs_, "residual-limit", 1e-6);
    }

    // output(): the macros (and Shakhov flux moments) in meso's files
    void output() {
        fetch();
        const bool shak = model_ == MESO_MODEL_CDUGKS_SHAKHOV;
        const bool flux = shak && stepped_;  // meso's flux fields are empty until a step ran
        if (meso_case_write_output(cs_, root_.c_str(), step, solution_time, rho_.data(), u_.data(),
                                   shak ? T_.data() : nullptr, shak ? q_.data() : nullptr,
                                   flux ? m0_.data() : nullptr, flux ? m1_.data() : nullptr,
                                   flux ? m2_.data() : nullptr) != MESO_OK)
            std::cerr << "[Warn] " << meso_case_output_error() << std::endl;
    }

    bool get_run_state() const { return run_state_; }
    bool failed() const { return failed_; }
    int max_step() const { return setting_int(cs_, "max-step", 100000); }
    int write_interval() const { return setting_int(cs_, "write-interval", 1000); }

  private:
    void fetch() {
        check(meso_hip_get_macros(ctx_, rho_.data(), u_.data(), T_.data(), q_.data()), meso_hip_last_error(ctx_));
        if (model_ == MESO_MODEL_CDUGKS_SHAKHOV && stepped_)
            check(meso_hip_get_flux_moments(ctx_, m0_.data(), m1_.data(), m2_.data()), meso_hip_last_error(ctx_));
    }

    static constexpr int residual_interval = 100;  // include/solver/solver.h:49
    meso_case *cs_ = nullptr;
    meso_hip_ctx *ctx_ = nullptr;
    std::string root_, case_name_;
    int nc_ = 0, dim_ = 2, model_ = 0, converge_state_ = 0;
    double dt_ = 0.0, residual_limit_ = 1e-6;
    bool run_state_ = true, failed_ = false, stepped_ = false;
    std::vector<double> rho_, T_, u_, q_, m0_, m1_, m2_;
};

}  // namespace

int main(int argc, char **argv) {
    std::string case_path, out = ".", storage_s = "fp64";
    bool not_run = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--case" && i + 1 < argc) case_path = argv[++i];
        else if (a == "--storage" && i + 1 < argc) storage_s = argv[++i];
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--not-run") not_run = true;
        else {
            std::cerr << "usage: meso_hip_run --case <case-file> [--storage fp64|fp32|fp32x|fp32e] [--out <dir>] [--not-run]"
                      << std::endl;
            return 2;
        }
    }
    if (case_path.empty() || (storage_s != "fp64" && storage_s != "fp32" && storage_s != "fp32x" &&
                               storage_s != "fp32e")) {
        std::cerr << "usage: meso_hip_run --case <case-file> [--storage fp64|fp32|fp32x|fp32e] [--out <dir>] [--not-run]"
                  << std::endl;
        return 2;
    }
    try {
        const int storage = storage_s == "fp32"    ? MESO_STORAGE_FP32
                            : storage_s == "fp32x" ? MESO_STORAGE_FP32X
                            : storage_s == "fp32e" ? MESO_STORAGE_FP32E
                                                   : MESO_STORAGE_FP64;
        HipSolver solver(case_path, storage, out);
        solver.initial();
        solver.output();
        if (not_run) return 0;
        for (int i = 0; (i < solver.max_step()) || (solver.max_step() < 0); ++i) {
            solver.do_step();
            if (!solver.get_run_state()) break;
            if (solver.step % solver.write_interval() == 0) solver.output();
        }
        if (solver.failed()) {
            solver.output();
            return 3;
        }
        std::cout << "Reached stop step: " << solver.max_step() << std::endl;
        solver.output();
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "meso_hip_run: " << e.what() << std::endl;
        return 1;
    }
}
