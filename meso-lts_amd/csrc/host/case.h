// A loaded case: Config + gambit mesh + DVS + step parameters, and the flat
// arrays behind the C-ABI descriptors (include/meso_hip.h).
#pragma once

#include "meso_hip.h"
#include "meso_host.h"

namespace meso_host {

int model_from_solver(const std::string &solver);

// Step parameters from [settings] (keys and defaults of cdugks_shakhov.cpp:7-33,
// cdugks.cpp:9-18, dugks.cpp:9-15; DUGKS dt from CFL as dugks.cpp:47).
// first: the solver's construction (meso reads venkata-limiter-k only then);
// later calls re-read [settings] and keep venkata_k.
void load_params(const MESO::Solver::Config &cfg, int model, int dim, double min_cell_size, double dvs_max_mag,
                 meso_step_params &p, bool first);

struct Case {
    MESO::Solver::Config config;
    std::string solver;
    int model = 0;
    MESO::fvmMesh::Mesh mesh;
    MESO::DVS::VelocitySet dvs;
    meso_step_params params{};

    std::vector<double> cell_pos, cell_vol, face_pos, face_area, face_normal, lsq_dr, lsq_w, lsq_C, xi, w;
    std::vector<int32_t> face_cells, cell_face_ptr, cell_face_ids, nbr_ptr, nbr_ids, face_mark, mirror;
    std::vector<meso_mark_desc> marks;

    Case(const std::string &path, double scale_override);
    void build_arrays();
    void mesh_desc(meso_mesh_desc &d) const;
    void dvs_desc(meso_dvs_desc &d) const;
    void initial_patch(double *rho, double *u3, double *T) const;
    // 1 for the cells whose group sets `mask switch True` (src/solver/boundary/patch.cpp:48-51)
    void cell_mask(int32_t *mask) const;
};

}  // namespace meso_host
