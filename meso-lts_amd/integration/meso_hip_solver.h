// meso_hip_solver.h — the reference-side binding of the gfx950 step: a meso
// solver class that keeps meso's case file, time loop and output files and
// runs the step on the GPU through libmeso_hip.so (include/meso_hip.h).
//
// It fills meso's solver contract:
//   BasicSolver base and members        include/solver/solver.h:22-46
//   public step / solution_time         include/solver/boltzmann/cdugks_shakhov.h:17-18
//   initial() / do_step() / output()    called by handle_solver<T>, src/meso/exec.cpp:141-165
//   residual print + convergence rule   src/solver/boltzmann/cdugks_shakhov.cpp:724-764,
//                                       cdugks.cpp:337-371, dugks.cpp:346-383
//   NaN / T<0 monitor and mask repair   cdugks_shakhov.cpp:698-719, 766-789
// and specialises handle_solver<T> with the loop of exec.cpp:141-165.
//
// One class serves the three solvers; the case key `solver` picks the model
// (src/meso/main.cpp:47-57). The case key `storage` (fp64 | fp32 | fp32e | fp32x) picks
// the device storage of the distributions; fp64 is meso's own arithmetic.
//
// A maintainer adds to meso:
//   include/solver/solver.h, after its namespace:   #include "meso_hip_solver.h"
//   src/meso/main.cpp:47-57, in the solver dispatch:
//       else if (solver == "cdugks@shakhov-hip" || solver == "cdugks@incompressible-hip" ||
//                solver == "dugks@incompressible-hip")
//           solver_status = MESO::Solver::handle_solver<MESO::Solver::MESO_HIP_SOLVER>(parser, config);
//   the link line:   -L<repo>/meso-lts_amd/lib -lmeso_hip
// The device step holds the whole velocity set, so the class runs on one MPI rank.
#ifndef MESO_HIP_SOLVER_H
#define MESO_HIP_SOLVER_H

#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "meso_hip.h"
#include "solver/solver.h"

namespace MESO::Solver {

class MESO_HIP_SOLVER : public BasicSolver {
public:
    int step{};
    double solution_time{};
    double dt{};

    MESO_HIP_SOLVER(ArgParser &parser, Config &config) : BasicSolver(parser, config) {
        if (MPI::processor_num != 1)
            fail("MESO_HIP_SOLVER holds the whole velocity set on one GPU: run it on one MPI rank");
        const auto case_path = parser.parse_param<std::string>("case", "<case-file>", false);
        if (meso_case_open(case_path.c_str(), 0.0, &cs_) != MESO_OK) fail(meso_case_last_error());
        meso_mesh_desc m{};
        meso_dvs_desc d{};
        meso_case_mesh(cs_, &m);
        meso_case_dvs(cs_, &d);
        meso_case_params(cs_, &params_);
        const auto storage = config.get<std::string>("storage", "fp64", false);
        int st = MESO_STORAGE_FP64;
        if (storage == "fp32") st = MESO_STORAGE_FP32;
        else if (storage == "fp32x") st = MESO_STORAGE_FP32X;
        else if (storage == "fp32e") st = MESO_STORAGE_FP32E;
        else if (storage != "fp64") fail("storage must be fp64, fp32, fp32e or fp32x, not " + storage);
        if (meso_hip_create(&m, &d, &params_, st, &ctx_) != MESO_OK) {
            const std::string e = meso_hip_last_error(nullptr);
            meso_case_close(cs_);
            cs_ = nullptr;
            fail(e);
        }
        nc_ = m.n_cells;
        dim_ = m.dim;
        model_ = params_.model;
        dt = params_.dt;
        // cell groups with `mask switch True` (src/solver/boundary/patch.cpp:48-51)
        std::vector<int32_t> mask(nc_, 0);
        check_case(meso_case_cell_mask(cs_, mask.data()));
        check(meso_hip_set_mask(ctx_, mask.data()));
        rho_.assign(nc_, 0.0);
        T_.assign(nc_, 0.0);
        u_.assign(3 * size_t(nc_), 0.0);
        q_.assign(3 * size_t(nc_), 0.0);
        m0_.assign(nc_, 0.0);
        m1_.assign(3 * size_t(nc_), 0.0);
        m2_.assign(nc_, 0.0);
    }

    ~MESO_HIP_SOLVER() {
        meso_hip_destroy(ctx_);
        meso_case_close(cs_);
    }

    MESO_HIP_SOLVER(const MESO_HIP_SOLVER &) = delete;
    MESO_HIP_SOLVER &operator=(const MESO_HIP_SOLVER &) = delete;

    // Shakhov: re-read [settings] and hand dt, Pr, mu0, ... to the device
    // (CDUGKS_SHAKHOV::update_config, cdugks_shakhov.cpp:7-33).
    void update_config() override {
        BasicSolver::update_config();
        check_case(meso_case_reload_settings(cs_));
        meso_case_params(cs_, &params_);
        check(meso_hip_set_params(ctx_, &params_));
        dt = params_.dt;
    }

    // initial(): the cell-group patches, or the `read-np-data` restart
    // (cdugks_shakhov.cpp:62-168, cdugks.cpp:41-114, dugks.cpp:41-105).
    void initial() {
        step = 0;
        solution_time = 0.0;
        run_state = true;
        if (config.get<bool>("read-np-data", false, false)) {
            if (model_ == MESO_MODEL_CDUGKS_SHAKHOV) {
                read_np("latest/Rho", rho_.data(), 1);
                read_np("latest/T", T_.data(), 1);
                read_np("latest/U", u_.data(), 3);
                read_np("latest/q", q_.data(), 3);
            } else if (model_ == MESO_MODEL_CDUGKS_BGK) {
                read_np("latest/Rho", rho_.data(), 1);
                read_np("latest/U", u_.data(), 3);
            } else {
                read_np("np-data/Rho.np.dat", rho_.data(), 1);
                read_np("np-data/vel.np.dat", u_.data(), 3);
            }
            check(meso_hip_init_from_macros(ctx_, rho_.data(), u_.data(), T_.data(), q_.data()));
        } else {
            check_case(meso_case_initial_patch(cs_, rho_.data(), u_.data(), T_.data()));
            check(meso_hip_init_equilibrium(ctx_, rho_.data(), u_.data(), T_.data()));
        }
    }

    // do_step(): reconstruct() + fvm_update() on the device, then meso's
    // bookkeeping (cdugks_shakhov.cpp:724-764).
    void do_step() {
        meso_step_report rep{};
        const int rc = meso_hip_step(ctx_, 1, &rep);
        if (rep.masked_cells > 0)
            logger.warn << "[Warn] Mask working on " << rep.masked_cells << " cell(s) (NaN density or T < 0)"
                        << std::endl;
        if (rc == MESO_E_NUMERIC) {
            // an unmasked cell went bad: meso logs it and stops the run
            logger.warn << "[WARN] cell<" << rep.first_bad_cell + 1 << "> caught rho=nan or T < 0 ("
                        << rep.bad_cells << " cell(s))." << std::endl;
            run_state = false;
        } else {
            check(rc);
        }
        stepped_ = true;
        step++;
        solution_time += dt;
        if (step % residual_interval == 0) {
            double r[8];  // rho, u, v, w, T, qx, qy, qz
            check(meso_hip_residual(ctx_, r));
            logger.note << "step: " << step << std::endl;
            List<Scalar> res;
            StringList names;
            if (model_ == MESO_MODEL_CDUGKS_SHAKHOV) {
                if (dim_ == 2) {
                    res = {r[0], r[4], r[1], r[2], r[5], r[6]};
                    names = {"Res[Rho]", "Res[T]", "Res[u]", "Res[v]", "Res[qx]", "Res[qy]"};
                } else {
                    res = {r[0], r[4], r[1], r[2], r[3], r[5], r[6], r[7]};
                    names = {"Res[Rho]", "Res[T]", "Res[u]", "Res[v]", "Res[w]", "Res[qx]", "Res[qy]", "Res[qz]"};
                }
            } else if (dim_ == 2) {
                res = {r[0], r[1], r[2]};
                names = {"Res[Rho]", "Res[U]", "Res[V]"};
            } else {
                res = {r[0], r[1], r[2], r[3]};
                names = {"Res[Rho]", "Res[U]", "Res[V]", "Res[W]"};
            }
            Utils::print_names_and_values(names, res);
            converge_state = Utils::is_converged(res, residual_limit) ? converge_state + 1 : 0;
            if (converge_state >= 10) {
                logger.note << "Step " << step << " - Convergence achieved: residual all below " << residual_limit
                            << std::endl;
                run_state = false;
            }
        }
        // CDUGKS_SHAKHOV re-reads its settings every step; DUGKS re-reads the
        // file without touching its members (dugks.cpp:381); CDUGKS does neither.
        if (model_ == MESO_MODEL_CDUGKS_SHAKHOV) update_config();
        else if (model_ == MESO_MODEL_DUGKS_BGK) BasicSolver::update_config();
    }

    // output(): macros (and the Shakhov flux moments) copied back and written
    // in meso's files (cdugks_shakhov.cpp:791-855, cdugks.cpp:373-420, dugks.cpp:385-411).
    void output() {
        check(meso_hip_get_macros(ctx_, rho_.data(), u_.data(), T_.data(), q_.data()));
        const bool shak = model_ == MESO_MODEL_CDUGKS_SHAKHOV;
        const bool flux = shak && stepped_;  // meso's flux fields are empty before the first step
        if (flux) check(meso_hip_get_flux_moments(ctx_, m0_.data(), m1_.data(), m2_.data()));
        if (meso_case_write_output(cs_, "", step, solution_time, rho_.data(), u_.data(), shak ? T_.data() : nullptr,
                                   shak ? q_.data() : nullptr, flux ? m0_.data() : nullptr,
                                   flux ? m1_.data() : nullptr, flux ? m2_.data() : nullptr) != MESO_OK)
            logger.warn << "Solver::output() " << meso_case_output_error() << std::endl;
    }

private:
    [[noreturn]] static void fail(const std::string &msg) {
        logger.warn << "[Error] " << msg << std::endl;
        throw std::runtime_error(msg);
    }
    void check(int rc) const {
        if (rc != MESO_OK) fail(meso_hip_last_error(ctx_));
    }
    static void check_case(int rc) {
        if (rc != MESO_OK) fail(meso_case_last_error());
    }
    // Field::read_np_data (src/solver/field.cpp:252-273), under <case-name>/
    void read_np(const std::string &name, double *dst, int width) const {
        const std::string file = case_name + "/" + name;
        std::ifstream fp(file);
        if (!fp) fail("cannot open " + file);
        std::string line;
        int i = 0;
        while (i < nc_ && std::getline(fp, line)) {
            if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
            std::istringstream ss(line);
            for (int k = 0; k < width; ++k)
                if (!(ss >> dst[size_t(width) * i + k])) fail("bad line in " + file);
            ++i;
        }
        if (i < nc_) fail(file + ": fewer lines than cells");
    }

    meso_case *cs_ = nullptr;
    meso_hip_ctx *ctx_ = nullptr;
    meso_step_params params_{};
    int nc_ = 0, dim_ = 2, model_ = 0;
    bool stepped_ = false;
    std::vector<double> rho_, T_, u_, q_, m0_, m1_, m2_;
};

// The time loop of src/meso/exec.cpp:141-165 for this class.
template <>
inline int handle_solver<MESO_HIP_SOLVER>(MESO::ArgParser &parser, MESO::Solver::Config &config) {
    MESO_HIP_SOLVER solver(parser, config);
    solver.initial();
    solver.output();
    if (parser.parse_switch("not-run")) {
        MPI_Barrier(MPI_COMM_WORLD);
        return 0;
    }
    for (int i = 0; (i < solver.max_step()) || (solver.max_step() < 0); ++i) {
        solver.do_step();
        if (not solver.get_run_state()) break;
        if (solver.step % solver.write_interval() == 0) solver.output();
        if (solver_state != 0) {
            solver.output();
            MPI_Barrier(MPI_COMM_WORLD);
            return solver_state;
        }
    }
    logger.note << "Reached stop step: " << solver.max_step() << std::endl;
    solver.output();
    MPI_Barrier(MPI_COMM_WORLD);
    return 0;
}

}  // namespace MESO::Solver

#endif  // MESO_HIP_SOLVER_H
