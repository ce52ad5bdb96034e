// meso_case_*: case file -> plain descriptors of the C ABI.
//
// Parameter keys and defaults follow meso's solver constructors:
//   CDUGKS_SHAKHOV::update_config  src/solver/boltzmann/cdugks_shakhov.cpp:7-33
//   CDUGKS ctor / initial          src/solver/boltzmann/cdugks.cpp:7-47
//   DUGKS ctor / initial           src/solver/boltzmann/dugks.cpp:7-48
// DVS selection follows cdugks_shakhov.cpp:43-55 and cdugks.cpp:25-34.
#include "case.h"

#include <cstring>
#include <mutex>

using namespace MESO;

namespace {
thread_local std::string g_case_error;

int fail(int code, const std::string &msg) {
    g_case_error = msg;
    return code;
}

}  // namespace

namespace meso_host {

// meso's solver names (src/meso/main.cpp:47-57), also with the suffix "-hip"
// under which a meso build registers the GPU solver class next to its own.
int model_from_solver(const std::string &name) {
    const std::string sfx = "-hip";
    const std::string s = name.size() > sfx.size() && name.compare(name.size() - sfx.size(), sfx.size(), sfx) == 0
                              ? name.substr(0, name.size() - sfx.size())
                              : name;
    if (s == "cdugks@shakhov") return MESO_MODEL_CDUGKS_SHAKHOV;
    if (s == "cdugks@incompressible") return MESO_MODEL_CDUGKS_BGK;
    if (s == "dugks@incompressible") return MESO_MODEL_DUGKS_BGK;
    throw std::invalid_argument("unsupported solver: " + name);
}

void load_params(const Solver::Config &cfg, int model, int dim, double min_cell_size, double dvs_max_mag,
                 meso_step_params &p, bool first) {
    const double venkata_k = first ? 0.0 : p.venkata_k;  // kept across reloads, as meso's member
    std::memset(&p, 0, sizeof(p));
    p.model = model;
    p.dim = dim;
    if (model == MESO_MODEL_CDUGKS_SHAKHOV) {
        p.Pr = cfg.get("Pr", 0.67);
        p.dt = cfg.get("timeStep", 1.0e-5);
        p.R = cfg.get("gas-constant", 0.5);
        p.K = cfg.get("gas-k", 0);
        p.T0 = cfg.get("ref-temperature", 1.0);
        p.mu0 = cfg.get("ref-viscosity", 1.60e-05);
        p.omega = cfg.get("vhs-index", 0.81);
        p.gradient_switch = cfg.get("gradient-switch", true) ? 1 : 0;
        p.limiter_switch = cfg.get("limiter-switch", false) ? 1 : 0;
        // venkata_k is read only at construction with the gradient off and the
        // limiter on (cdugks_shakhov.cpp:21-28); the limiter is applied only
        // with the gradient on, so it normally sees 0
        p.venkata_k = venkata_k;
        if (first && !p.gradient_switch && p.limiter_switch) p.venkata_k = cfg.get("venkata-limiter-k", 1.0);
    } else {
        double RT = cfg.get("gas-constant", 0.5) * cfg.get("ref-temperature", 1.0);
        double L0 = cfg.get("ref-length", 1.0);
        double Ma = cfg.get("Ma", 0.1);
        double Re = cfg.get("Re", 400.0);
        p.RT = RT;
        p.tau = (L0 * Ma) / (Re * std::sqrt(RT));
        p.gradient_switch = cfg.get("gradient-switch", true) ? 1 : 0;
        if (model == MESO_MODEL_CDUGKS_BGK) {
            p.dt = cfg.get("timeStep", 2.5e-3);
        } else {
            double CFL = cfg.get("CFL", 0.8);
            p.dt = CFL * (min_cell_size / dvs_max_mag);
        }
    }
}

Case::Case(const std::string &path, double scale_override) : config(path) {
    solver = config.get("solver", String("<solver>"));
    model = model_from_solver(solver);
    double scale = scale_override > 0 ? scale_override : config.get("mesh-scale", 1.0);
    mesh = fvmMesh::load_gambit(config.resolve(config.get("mesh-file", String("<mesh-file>"))), scale);
    const int D = mesh.dimension();
    if (D != 2 && D != 3) throw std::invalid_argument("mesh dimension must be 2 or 3");
    String dvs_file = config.resolve(config.get("dvs-file", String("<dvs-file>")));
    String dvs_type = config.get("dvs-type", String("<dvs-type>"));
    if (model == MESO_MODEL_CDUGKS_SHAKHOV) {
        double RT0 = config.get("gas-constant", 0.5) * config.get("ref-temperature", 1.0);
        if (dvs_type == "Newton-Cotes") dvs = DVS::newton_cotes(dvs_file, D);
        else if (dvs_type == "half-range-Gauss-Hermite") dvs = DVS::half_range_gauss_hermite(dvs_file, D, RT0);
        else if (dvs_type == "Sparse") dvs = DVS::sparse(dvs_file);
        else dvs = DVS::from_gambit(dvs_file);
    } else {
        if (dvs_type != "Gauss-Hermite") throw std::invalid_argument("dvs-mesh type error.");
        double RT = config.get("gas-constant", 0.5) * config.get("ref-temperature", 1.0);
        dvs = DVS::gauss_hermite(dvs_file, D, RT);
    }
    load_params(config, model, D, mesh.min_cell_size, dvs.max_magnitude, params, true);
    build_arrays();
}

void Case::build_arrays() {
    const int NC = mesh.NCELL, NF = mesh.NFACE;
    cell_pos.resize(3 * NC);
    cell_vol.resize(NC);
    for (auto &c : mesh.cells) {
        for (int k = 0; k < 3; ++k) cell_pos[3 * c.id + k] = c.position[k];
        cell_vol[c.id] = c.volume;
    }
    face_cells.resize(2 * NF);
    face_pos.resize(3 * NF);
    face_area.resize(NF);
    face_normal.resize(3 * NF);
    face_mark.resize(NF);
    // marks: [0] = fluid-interior, then one entry per mesh face group name in use
    marks.clear();
    meso_mark_desc interior;
    std::memset(&interior, 0, sizeof(interior));
    marks.push_back(interior);
    std::vector<int> group_to_mark(mesh.face_names.size(), -1);
    for (size_t g = 0; g < mesh.face_names.size(); ++g) {
        if (g >= mesh.face_groups.size() || mesh.face_groups[g].empty()) continue;
        const Solver::Mark *mk = nullptr;
        auto it = config.marks.find(mesh.face_names[g]);
        if (it == config.marks.end()) throw std::invalid_argument("get_face_group() missing key=" + mesh.face_names[g]);
        mk = &it->second;
        if (mk->type == Solver::fluid_interior) {
            group_to_mark[g] = 0;
            continue;
        }
        meso_mark_desc d;
        std::memset(&d, 0, sizeof(d));
        d.type = mk->type;
        d.T_kind = mk->patch.get_type("temperature");
        d.direction = mk->patch.get_int("direction");
        Vector u = mk->patch.get_vector("velocity");
        d.velocity[0] = u.x;
        d.velocity[1] = u.y;
        d.velocity[2] = u.z;
        d.temperature = mk->patch.get_scalar("temperature");
        d.density = mk->patch.get_scalar("density");
        d.pressure = mk->patch.get_scalar("pressure");
        d.alpha_u = mk->patch.get_scalar("slip-wall-alpha-u");
        d.alpha_T = mk->patch.get_scalar("slip-wall-alpha-T");
        if ((d.type == MESO_BC_SYMMETRY || d.type == MESO_BC_SLIP_WALL) && dvs.mirror.empty())
            throw std::invalid_argument("symmetry / slip-wall need a DVS with mirror ids (Newton-Cotes or half-range GH)");
        if (d.direction < 0 || d.direction > 2) throw std::invalid_argument("mirror direction must be 0, 1 or 2");
        if (d.type == MESO_BC_FREESTREAM_INLET)
            throw std::invalid_argument("freestream-inlet is not supported: its density uses the loop index "
                                        "(cdugks_shakhov.cpp:419)");
        group_to_mark[g] = int(marks.size());
        marks.push_back(d);
    }
    for (auto &f : mesh.faces) {
        face_cells[2 * f.id] = f.cell_id[0];
        face_cells[2 * f.id + 1] = f.cell_id[1];
        for (int k = 0; k < 3; ++k) {
            face_pos[3 * f.id + k] = f.position[k];
            face_normal[3 * f.id + k] = f.normal_vector[0][k];
        }
        face_area[f.id] = f.area;
        int m = group_to_mark.at(f.group_id);
        if (m < 0) throw std::invalid_argument("face group without mark");
        face_mark[f.id] = m;
    }
    cell_face_ptr.assign(1, 0);
    cell_face_ids.clear();
    nbr_ptr.assign(1, 0);
    nbr_ids.clear();
    lsq_dr.clear();
    lsq_w.clear();
    lsq_C.resize(9 * NC);
    for (auto &c : mesh.cells) {
        for (int f : c.face_id) cell_face_ids.push_back(f);
        cell_face_ptr.push_back(int(cell_face_ids.size()));
        const auto &ls = c.least_square;
        for (int j = 0; j < ls.neighbor_num; ++j) {
            nbr_ids.push_back(c.neighbors[j]);
            for (int k = 0; k < 3; ++k) lsq_dr.push_back(ls.dr[j][k]);
            lsq_w.push_back(ls.weight[j]);
        }
        nbr_ptr.push_back(int(nbr_ids.size()));
        for (int k = 0; k < 3; ++k) {
            lsq_C[9 * c.id + k] = ls.Cx[k];
            lsq_C[9 * c.id + 3 + k] = ls.Cy[k];
            lsq_C[9 * c.id + 6 + k] = ls.Cz[k];
        }
    }
    const int NV = dvs.size();
    xi.resize(3 * NV);
    w.resize(NV);
    mirror.clear();
    for (int p = 0; p < NV; ++p) {
        for (int k = 0; k < 3; ++k) xi[3 * p + k] = dvs.xi[p][k];
        w[p] = dvs.weight[p];
    }
    if (!dvs.mirror.empty())
        for (auto &m : dvs.mirror) mirror.insert(mirror.end(), m.begin(), m.end());
}

void Case::mesh_desc(meso_mesh_desc &d) const {
    std::memset(&d, 0, sizeof(d));
    d.dim = mesh.dimension();
    d.n_cells = mesh.NCELL;
    d.n_faces = mesh.NFACE;
    d.n_marks = int(marks.size());
    d.n_nbr = int(nbr_ids.size());
    d.n_cell_faces = int(cell_face_ids.size());
    d.total_volume = mesh.total_volume;
    d.min_cell_size = mesh.min_cell_size;
    d.cell_pos = cell_pos.data();
    d.cell_vol = cell_vol.data();
    d.face_cells = face_cells.data();
    d.face_pos = face_pos.data();
    d.face_area = face_area.data();
    d.face_normal = face_normal.data();
    d.cell_face_ptr = cell_face_ptr.data();
    d.cell_face_ids = cell_face_ids.data();
    d.nbr_ptr = nbr_ptr.data();
    d.nbr_ids = nbr_ids.data();
    d.lsq_dr = lsq_dr.data();
    d.lsq_w = lsq_w.data();
    d.lsq_C = lsq_C.data();
    d.face_mark = face_mark.data();
    d.marks = marks.data();
}

void Case::dvs_desc(meso_dvs_desc &d) const {
    std::memset(&d, 0, sizeof(d));
    d.n_vel = dvs.size();
    d.has_mirror = mirror.empty() ? 0 : 1;
    d.max_magnitude = dvs.max_magnitude;
    d.xi = xi.data();
    d.w = w.data();
    d.mirror = mirror.empty() ? nullptr : mirror.data();
}

void Case::initial_patch(double *rho, double *u3, double *T) const {
    for (auto &cell : mesh.cells) {
        const auto &g = config.get_cell_group(cell, mesh);
        using Solver::PatchType::fromFile;
        double r = g.patch.get_type("density") == fromFile ? g.patch.get_file_scalar("density", cell.id)
                                                            : g.patch.get_scalar("density");
        Vector u = g.patch.get_type("velocity") == fromFile ? g.patch.get_file_vector("velocity", cell.id)
                                                             : g.patch.get_vector("velocity");
        double t = g.patch.get_type("temperature") == fromFile ? g.patch.get_file_scalar("temperature", cell.id)
                                                                : g.patch.get_scalar("temperature");
        if (rho) rho[cell.id] = r;
        if (u3) {
            u3[3 * cell.id] = u.x;
            u3[3 * cell.id + 1] = u.y;
            u3[3 * cell.id + 2] = u.z;
        }
        if (T) T[cell.id] = t;
    }
}

void Case::cell_mask(int32_t *mask) const {
    for (auto &cell : mesh.cells) mask[cell.id] = config.get_cell_group(cell, mesh).patch.get_switch("mask") ? 1 : 0;
}

}  // namespace meso_host

// ------------------------------------------------------------------ C ABI
extern "C" {

const char *meso_case_last_error(void) { return g_case_error.c_str(); }

int meso_case_open(const char *case_path, double scale, meso_case **out) {
    if (!case_path || !out) return fail(MESO_E_INVALID, "null argument");
    try {
        *out = reinterpret_cast<meso_case *>(new meso_host::Case(case_path, scale));
        return MESO_OK;
    } catch (const std::invalid_argument &e) {
        return fail(std::string(e.what()).find("open") != std::string::npos ? MESO_E_IO : MESO_E_INVALID, e.what());
    } catch (const std::exception &e) {
        return fail(MESO_E_IO, e.what());
    }
}

void meso_case_close(meso_case *c) { delete reinterpret_cast<meso_host::Case *>(c); }

int meso_case_mesh(const meso_case *c, meso_mesh_desc *out) {
    if (!c || !out) return fail(MESO_E_INVALID, "null argument");
    reinterpret_cast<const meso_host::Case *>(c)->mesh_desc(*out);
    return MESO_OK;
}

int meso_case_dvs(const meso_case *c, meso_dvs_desc *out) {
    if (!c || !out) return fail(MESO_E_INVALID, "null argument");
    reinterpret_cast<const meso_host::Case *>(c)->dvs_desc(*out);
    return MESO_OK;
}

int meso_case_params(const meso_case *c, meso_step_params *out) {
    if (!c || !out) return fail(MESO_E_INVALID, "null argument");
    *out = reinterpret_cast<const meso_host::Case *>(c)->params;
    return MESO_OK;
}

int meso_case_reload_settings(meso_case *c) {
    if (!c) return fail(MESO_E_INVALID, "null argument");
    auto *cs = reinterpret_cast<meso_host::Case *>(c);
    try {
        cs->config.update_config(false);
        meso_host::load_params(cs->config, cs->model, cs->mesh.dimension(), cs->mesh.min_cell_size,
                               cs->dvs.max_magnitude, cs->params, false);
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_IO, e.what());
    }
}

int meso_case_initial_patch(const meso_case *c, double *rho, double *u3, double *T) {
    if (!c) return fail(MESO_E_INVALID, "null argument");
    try {
        reinterpret_cast<const meso_host::Case *>(c)->initial_patch(rho, u3, T);
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_INVALID, e.what());
    }
}

int meso_case_cell_mask(const meso_case *c, int32_t *mask) {
    if (!c || !mask) return fail(MESO_E_INVALID, "null argument");
    try {
        reinterpret_cast<const meso_host::Case *>(c)->cell_mask(mask);
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_INVALID, e.what());
    }
}

int meso_case_setting_int(const meso_case *c, const char *key, int32_t dflt, int32_t *out) {
    if (!c || !key || !out) return fail(MESO_E_INVALID, "null argument");
    try {
        *out = reinterpret_cast<const meso_host::Case *>(c)->config.get(key, int(dflt));
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_INVALID, e.what());
    }
}

int meso_case_setting_double(const meso_case *c, const char *key, double dflt, double *out) {
    if (!c || !key || !out) return fail(MESO_E_INVALID, "null argument");
    try {
        *out = reinterpret_cast<const meso_host::Case *>(c)->config.get(key, dflt);
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_INVALID, e.what());
    }
}

int meso_case_setting_str(const meso_case *c, const char *key, const char *dflt, char *out, int32_t cap) {
    if (!c || !key || !out || cap <= 0) return fail(MESO_E_INVALID, "null argument");
    std::string v = reinterpret_cast<const meso_host::Case *>(c)->config.get(key, std::string(dflt ? dflt : ""));
    std::strncpy(out, v.c_str(), size_t(cap) - 1);
    out[cap - 1] = 0;
    return MESO_OK;
}

}  // extern "C"
