// Driver for meso itself, compiled from the reference sources where they lie
// (see oracle/Makefile.ref). It constructs meso's own solver class from a case
// file, runs initial() and do_step() exactly as handle_solver<T>() does
// (src/meso/exec.cpp:141-160; output() only with --write), and dumps geometry
// and state as raw little-endian float64/int32 files for golden vectors.
//
// TEST INFRASTRUCTURE ONLY: used to write tests/golden/ fixtures and to pin
// the restatement in oracle/meso_oracle.cpp. Nothing in the product uses it.
//
// usage: meso_ref --case <case.txt> --out <dir> --dump 0,1,2,10 [-state] [-write]
#include "solver/solver.h"

#include <cstdio>
#include <set>
#include <sys/stat.h>

using namespace MESO;

namespace {

std::string g_out;

void put(const std::string &name, const void *data, size_t bytes) {
    std::string path = g_out + "/" + name;
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot write " + path);
    std::fwrite(data, 1, bytes, fp);
    std::fclose(fp);
}

void put_scalar(const std::string &name, const Field<Scalar> &f) { put(name, f.values.data(), f.values.size() * 8); }

void put_vector(const std::string &name, const Field<Vector> &f) {
    std::vector<double> v;
    for (auto &x : f.values) {
        v.push_back(x.x);
        v.push_back(x.y);
        v.push_back(x.z);
    }
    put(name, v.data(), v.size() * 8);
}

void put_dist(const std::string &name, const Solver::DistributionFunction &f) {
    std::vector<double> v;
    for (auto &fp : f) v.insert(v.end(), fp.values.begin(), fp.values.end());
    put(name, v.data(), v.size() * 8);
}

void put_geometry(fvmMesh::Mesh &mesh, fvmMesh::Mesh &dvs) {
    std::vector<double> cp, cv, fp, fa, fn, C, xi, w;
    std::vector<int> fc, nb;
    for (auto &c : mesh.cells) {
        cp.insert(cp.end(), {c.position.x, c.position.y, c.position.z});
        cv.push_back(c.volume);
        auto &ls = c.least_square;
        C.insert(C.end(), {ls.Cx.x, ls.Cx.y, ls.Cx.z, ls.Cy.x, ls.Cy.y, ls.Cy.z, ls.Cz.x, ls.Cz.y, ls.Cz.z});
        nb.push_back(ls.neighbor_num);
    }
    for (auto &f : mesh.faces) {
        fp.insert(fp.end(), {f.position.x, f.position.y, f.position.z});
        fa.push_back(f.area);
        fn.insert(fn.end(), {f.normal_vector[0].x, f.normal_vector[0].y, f.normal_vector[0].z});
        fc.push_back(f.cell_id[0]);
        fc.push_back(f.cell_id[1]);
    }
    for (auto &p : dvs.cells) {
        xi.insert(xi.end(), {p.position.x, p.position.y, p.position.z});
        w.push_back(p.volume);
    }
    put("cell_pos.f64", cp.data(), cp.size() * 8);
    put("cell_vol.f64", cv.data(), cv.size() * 8);
    put("lsq_C.f64", C.data(), C.size() * 8);
    put("nbr_count.i32", nb.data(), nb.size() * 4);
    put("face_pos.f64", fp.data(), fp.size() * 8);
    put("face_area.f64", fa.data(), fa.size() * 8);
    put("face_normal.f64", fn.data(), fn.size() * 8);
    put("face_cells.i32", fc.data(), fc.size() * 4);
    put("dvs_xi.f64", xi.data(), xi.size() * 8);
    put("dvs_w.f64", w.data(), w.size() * 8);
    double extra[2] = {mesh.total_volume, mesh.min_cell_size};
    put("mesh_scalars.f64", extra, sizeof(extra));
    FILE *m = std::fopen((g_out + "/meta.txt").c_str(), "w");
    std::fprintf(m, "%d %d %d %d\n", mesh.NCELL, mesh.NFACE, int(dvs.cells.size()), mesh.dimension());
    std::fclose(m);
}

std::set<int> parse_steps(const std::string &s) {
    std::set<int> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) out.insert(std::stoi(tok));
    return out;
}

template <class S>
void dump_macros(S &s, int step, bool state);

template <>
void dump_macros(Solver::CDUGKS_SHAKHOV &s, int step, bool state) {
    std::string p = "step" + std::to_string(step) + "_";
    put_scalar(p + "rho.f64", s.rho_cell);
    put_vector(p + "u.f64", s.vel_cell);
    put_scalar(p + "T.f64", s.T_cell);
    put_vector(p + "q.f64", s.q_cell);
    if (step > 0) {
        put_scalar(p + "fluxM0.f64", s.flux_m0);
        put_vector(p + "fluxM1.f64", s.flux_m1);
        put_scalar(p + "fluxM2.f64", s.flux_m2);
    }
    if (state) {
        put_dist(p + "g.f64", s.g_cell);
        put_dist(p + "h.f64", s.h_cell);
    }
}

template <class S>
void dump_bgk(S &s, int step, bool state) {
    std::string p = "step" + std::to_string(step) + "_";
    put_scalar(p + "rho.f64", s.rho_cell);
    put_vector(p + "u.f64", s.vel_cell);
    if (state) put_dist(p + "f.f64", s.f_cell);
}
template <>
void dump_macros(Solver::CDUGKS &s, int step, bool state) { dump_bgk(s, step, state); }
template <>
void dump_macros(Solver::DUGKS &s, int step, bool state) { dump_bgk(s, step, state); }

// The residuals meso's do_step() forms at a check step (every
// Solver::residual_interval steps, cdugks_shakhov.cpp:730-744, cdugks.cpp:343-345,
// dugks.cpp:352-354), recomputed with meso's own Solver::residual() on copies of
// the fields it compares: rho, u (3), T, q (3) (BGK: T and q are 0).
struct ResBase {
    std::vector<Field<Scalar>> s;
    std::vector<Field<Vector>> v;
};
ResBase res_base(Solver::CDUGKS_SHAKHOV &x) { return {{x.rho_cell_res, x.T_cell_res}, {x.vel_cell_res, x.q_cell_res}}; }
template <class S>
ResBase res_base(S &x) { return {{x.rho_cell_res}, {x.vel_cell_res}}; }
void dump_residual(Solver::CDUGKS_SHAKHOV &x, ResBase &b, int step) {
    double r[8];
    r[0] = Solver::residual(b.s[0], x.rho_cell);
    Vector u = Solver::residual(b.v[0], x.vel_cell);
    r[4] = Solver::residual(b.s[1], x.T_cell);
    Vector q = Solver::residual(b.v[1], x.q_cell);
    r[1] = u.x, r[2] = u.y, r[3] = u.z, r[5] = q.x, r[6] = q.y, r[7] = q.z;
    put("step" + std::to_string(step) + "_residual.f64", r, sizeof(r));
}
template <class S>
void dump_residual(S &x, ResBase &b, int step) {
    double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    r[0] = Solver::residual(b.s[0], x.rho_cell);
    Vector u = Solver::residual(b.v[0], x.vel_cell);
    r[1] = u.x, r[2] = u.y, r[3] = u.z;
    put("step" + std::to_string(step) + "_residual.f64", r, sizeof(r));
}

// write: also call the solver's own output() at the dumped steps (files under
// the working directory, as meso writes them)
template <class S>
int run(ArgParser &parser, Solver::Config &config, const std::set<int> &steps, bool state, bool write) {
    S solver(parser, config);
    solver.initial();
    put_geometry(solver.mesh, solver.dvs_mesh);
    {
        double dt = solver.dt;
        put("dt.f64", &dt, 8);
    }
    if (steps.count(0)) {
        dump_macros(solver, 0, state);
        if (write) solver.output();
    }
    int last = steps.empty() ? 0 : *steps.rbegin();
    for (int i = 1; i <= last; ++i) {
        const bool check = i % Solver::residual_interval == 0;
        ResBase base;
        if (check) base = res_base(solver);
        solver.do_step();
        if (check) dump_residual(solver, base, i);
        if (steps.count(i)) {
            dump_macros(solver, i, state);
            if (write) solver.output();
        }
        if (!solver.get_run_state()) break;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    MESO::MPI::Initialize(&argc, &argv);
    ArgParser parser(argc, argv);
    std::string case_file = parser.parse_param<std::string>("case", "<case>", true);
    g_out = parser.parse_param<std::string>("out", "ref_out", true);
    std::string dump = parser.parse_param<std::string>("dump", "0,1", false);
    bool state = parser.parse_switch("state");
    bool write = parser.parse_switch("write");
    ::mkdir(g_out.c_str(), 0755);
    Solver::Config config(case_file);
    std::string solver = config.get<std::string>("solver", "<solver>", true);
    auto steps = parse_steps(dump);
    int rc = 0;
    if (solver == "cdugks@shakhov") rc = run<Solver::CDUGKS_SHAKHOV>(parser, config, steps, state, write);
    else if (solver == "cdugks@incompressible") rc = run<Solver::CDUGKS>(parser, config, steps, state, write);
    else if (solver == "dugks@incompressible") rc = run<Solver::DUGKS>(parser, config, steps, state, write);
    else rc = 2;
    MESO::MPI::Finalize();
    return rc;
}
