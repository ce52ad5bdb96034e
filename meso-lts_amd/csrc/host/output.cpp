// meso_case_write_output: the files meso's solvers write in output().
//
// Formats follow meso's writers, so that meso's post-processing reads them:
//   Field<Scalar>::output, Field<Vector>::output   src/solver/field.cpp:139-173
//   Mesh::output_grid, Mesh::output_data           src/mesh/mesh_output.cpp:182-247
//   Mesh::output (DUGKS result files)              src/mesh/mesh_output.cpp:111-172
//   write_mesh_node, write_mesh_geom               src/mesh/mesh_output.cpp:249-330
// and the directory layout of the solvers' output():
//   CDUGKS_SHAKHOV::output                         src/solver/boltzmann/cdugks_shakhov.cpp:791-855
//   CDUGKS::output                                 src/solver/boltzmann/cdugks.cpp:373-420
//   DUGKS::output                                  src/solver/boltzmann/dugks.cpp:385-411
#include "case.h"

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>

using namespace MESO;

namespace {

thread_local std::string g_out_error;

int fail(int code, const std::string &msg) {
    g_out_error = msg;
    return code;
}

bool make_dir(const std::string &p) {
    std::error_code ec;
    std::filesystem::create_directories(p, ec);
    return !ec;
}

// Field<Scalar>::output: one value per line, 18 significant digits.
bool write_scalar(const std::string &file, const double *v, int n) {
    std::ofstream fp(file, std::ios::out | std::ios::trunc);
    if (!fp) return false;
    for (int i = 0; i < n; ++i) fp << std::setprecision(18) << v[i] << std::endl;
    return bool(fp);
}

// Field<Vector>::output: "x y z" per line.
bool write_vector(const std::string &file, const double *v3, int n) {
    std::ofstream fp(file, std::ios::out | std::ios::trunc);
    if (!fp) return false;
    for (int i = 0; i < n; ++i)
        fp << std::setprecision(18) << v3[3 * i] << " " << std::setprecision(18) << v3[3 * i + 1] << " "
           << std::setprecision(18) << v3[3 * i + 2] << std::endl;
    return bool(fp);
}

void write_nodes(std::ofstream &fp, const fvmMesh::Mesh &m) {
    const int dim = m.dimension();
    const int LINE = 30;
    for (int a = 0; a < dim; ++a) {
        int count = 0;
        fp << std::endl << "## Node-" << "xyz"[a] << std::endl;
        for (const auto &node : m.nodes) {
            const double x = a == 0 ? node.position.x : (a == 1 ? node.position.y : node.position.z);
            fp << "\t" << std::setprecision(15) << x;
            if (count++ >= LINE) {
                fp << std::endl;
                count = 0;
            }
        }
        fp << std::endl;
    }
}

// Tecplot FE connectivity: 4 nodes per 2D cell, 8 per 3D cell, degenerate
// elements repeating nodes; node ids 1-based.
void write_geom(std::ofstream &fp, const fvmMesh::Mesh &m) {
    fp << std::endl << "## geom" << std::endl;
    for (const auto &cell : m.cells) {
        std::vector<ObjectId> n(cell.node_id.size());
        for (size_t i = 0; i < n.size(); ++i) n[i] = cell.node_id[i] + 1;
        switch (cell.geom_type) {
            case Geom::Quad: fp << " " << n[0] << " " << n[1] << " " << n[2] << " " << n[3]; break;
            case Geom::Tria: fp << " " << n[0] << " " << n[1] << " " << n[2] << " " << n[0]; break;
            case Geom::Brick:
                fp << " " << n[0] << " " << n[1] << " " << n[3] << " " << n[2] << " " << n[4] << " " << n[5] << " "
                   << n[7] << " " << n[6];
                break;
            case Geom::Wedge:
                fp << " " << n[0] << " " << n[1] << " " << n[2] << " " << n[2] << " " << n[3] << " " << n[4] << " "
                   << n[5] << " " << n[5];
                break;
            case Geom::Tetra:
                fp << " " << n[0] << " " << n[1] << " " << n[2] << " " << n[2] << " " << n[3] << " " << n[3] << " "
                   << n[3] << " " << n[3];
                break;
            case Geom::Pyram:
                fp << " " << n[0] << " " << n[1] << " " << n[3] << " " << n[2] << " " << n[4] << " " << n[4] << " "
                   << n[4] << " " << n[4];
                break;
            default: break;
        }
        fp << std::endl;
    }
}

// Mesh::output_grid
bool write_grid(const fvmMesh::Mesh &m, const std::string &file) {
    std::ofstream fp(file, std::ios::out | std::ios::trunc);
    if (!fp) return false;
    const int dim = m.dimension();
    fp << "TITLE = \"Grid\"\n"
          "FileType = \"GRID\"\n"
          "VARIABLES = \"X\",\"Y\"\n";
    if (dim == 3) fp << ",\"Z\"";  // meso writes it after the line break
    fp << std::endl;
    fp << "ZONE T=\"Grid\", N=" << m.NNODE << ", E=" << m.NCELL
       << ", DATAPACKING=BLOCK, ZONETYPE=" << (dim == 2 ? "FEQUADRILATERAL" : "FEBRICK") << std::endl;
    write_nodes(fp, m);
    write_geom(fp, m);
    return bool(fp);
}

// Mesh::output_data: cell-centred solution; the line counter runs on across
// the variables, as in meso.
bool write_data(const fvmMesh::Mesh &m, const std::string &file, const std::vector<std::vector<double>> &vals,
                const std::vector<std::string> &names, double solution_time) {
    std::ofstream fp(file, std::ios::out | std::ios::trunc);
    if (!fp) return false;
    const int dim = m.dimension();
    const int LINE = 30;
    fp << "TITLE = \"SolutionData\"\n"
          "FileType = \"SOLUTION\"\n"
          "VARIABLES = ";
    for (size_t i = 0; i < vals.size(); ++i) {
        fp << "\"" << names[i] << "\"";
        if (i + 1 == vals.size()) break;
        fp << ", ";
    }
    fp << "\nZONE T=\"Data\", N=" << m.NNODE << ", E=" << m.NCELL << ", VARLOCATION=(["
       << (vals.size() == 1 ? std::string("1") : "1-" + std::to_string(vals.size()))
       << "]=CELLCENTERED), DATAPACKING=BLOCK, ZONETYPE=" << (dim == 2 ? "FEQUADRILATERAL" : "FEBRICK")
       << ", SOLUTIONTIME=" << solution_time << std::endl;
    int count = 0;
    for (size_t i = 0; i < vals.size(); ++i) {
        fp << std::endl << "## " << names[i] << std::endl;
        for (double v : vals[i]) {
            fp << "\t" << std::setprecision(18) << v;
            if (count++ >= LINE) {
                fp << std::endl;
                count = 0;
            }
        }
        fp << std::endl;
    }
    return bool(fp);
}

// Mesh::output (the DUGKS result file): nodes, then cell variables.
bool write_result(const fvmMesh::Mesh &m, const std::string &file, const std::vector<std::vector<double>> &vals,
                  const std::vector<std::string> &names, int step, double solution_time) {
    std::ofstream fp(file + ".dat.plt", std::ios::out | std::ios::trunc);
    if (!fp) return false;
    const int vc = int(names.size());
    const int dim = m.dimension();
    const int LINE = 30;
    fp << R"(VARIABLES = "X", "Y")";
    if (dim == 3) fp << R"( ,"Z")";
    for (int i = 0; i < vc; ++i) fp << " ,\"" << names[i] << "\"";
    fp << std::endl;
    fp << "ZONE T=\"" << (step >= 0 ? "SolutionData(step=" + std::to_string(step) + ")" : "SolutionData")
       << "\", N=" << m.NNODE << ", E=" << m.NCELL << ", VARLOCATION=([1-" << dim << "]=NODAL, [";
    if (vc == 1) fp << dim + 1;
    else fp << dim + 1 << "-" << dim + vc;
    fp << "]=CELLCENTERED), DATAPACKING=BLOCK, ZONETYPE=" << (dim == 2 ? "FEQUADRILATERAL" : "FEBRICK");
    if (solution_time >= 0.0) fp << ", SOLUTIONTIME=" << solution_time;
    fp << std::endl;
    write_nodes(fp, m);
    for (int i = 0; i < vc; ++i) {
        int count = 0;
        fp << std::endl << "## CellVar-" << names[i] << std::endl;
        for (double v : vals[i]) {
            fp << "\t" << std::setprecision(15) << v;
            if (count++ >= LINE) {
                fp << std::endl;
                count = 0;
            }
        }
        fp << std::endl;
    }
    write_geom(fp, m);
    return bool(fp);
}

std::vector<double> component(const double *v3, int n, int k) {
    std::vector<double> r(n);
    for (int i = 0; i < n; ++i) r[i] = v3[3 * i + k];
    return r;
}

}  // namespace

extern "C" {

const char *meso_case_output_error(void) { return g_out_error.c_str(); }

int meso_case_write_output(const meso_case *c, const char *root, int32_t step, double solution_time,
                           const double *rho, const double *u3, const double *T, const double *q3,
                           const double *flux_m0, const double *flux_m1_3, const double *flux_m2) {
    if (!c || !rho || !u3) return fail(MESO_E_INVALID, "null argument");
    const auto *cs = reinterpret_cast<const meso_host::Case *>(c);
    const bool shak = cs->model == MESO_MODEL_CDUGKS_SHAKHOV;
    if (shak && (!T || !q3))
        return fail(MESO_E_INVALID, "the Shakhov output needs T and q");
    // meso's flux fields are empty until the first step has run: then all three are null
    const bool have_flux = flux_m0 && flux_m1_3 && flux_m2;
    if (shak && !have_flux && (flux_m0 || flux_m1_3 || flux_m2))
        return fail(MESO_E_INVALID, "the flux moments are given all three or none");
    try {
        const auto &m = cs->mesh;
        const int nc = m.NCELL, dim = m.dimension();
        const std::string base = std::string(root && *root ? root : ".") + "/";
        const std::string name = cs->config.get("case-name", "unnamed");
        const bool latest = cs->config.get("output-latest", true);
        const bool np = cs->config.get("output-np", false);
        const std::string dir = base + name;
        if (!make_dir(dir)) return fail(MESO_E_IO, "cannot mkdir " + dir);
        bool ok = true;
        if (cs->model == MESO_MODEL_DUGKS_BGK) {  // dugks.cpp:385-411
            std::vector<std::vector<double>> vals{std::vector<double>(rho, rho + nc)};
            std::vector<std::string> names{"Rho", "U", "V"};
            for (int k = 0; k < dim; ++k) vals.push_back(component(u3, nc, k));
            if (dim == 3) names.push_back("W");
            std::ostringstream fn;
            fn << "./" << name << "/result-" << step;
            ok = ok && write_result(m, base + fn.str(), vals, names, step, solution_time);
            if (np) {
                ok = ok && make_dir(dir + "/np-data");
                ok = ok && write_scalar(dir + "/np-data/Rho", rho, nc) && write_vector(dir + "/np-data/vel", u3, nc);
            }
        } else {  // cdugks_shakhov.cpp:791-855, cdugks.cpp:373-420
            if (step == 0) ok = ok && write_grid(m, dir + "/grid.plt");
            std::ostringstream p;
            p << dir << "/step-" << step << "/";
            const std::string path = p.str(), path_latest = dir + "/latest/";
            ok = ok && make_dir(path) && make_dir(path_latest);
            std::vector<std::vector<double>> vals{std::vector<double>(rho, rho + nc)};
            std::vector<std::string> names{"Rho"};
            const char *un[3] = {"U", "V", "W"}, *qn[3] = {"qx", "qy", "qz"};
            for (int k = 0; k < dim; ++k) {
                vals.push_back(component(u3, nc, k));
                names.push_back(un[k]);
            }
            if (shak) {
                vals.push_back(std::vector<double>(T, T + nc));
                names.push_back("T");
                for (int k = 0; k < dim; ++k) {
                    vals.push_back(component(q3, nc, k));
                    names.push_back(qn[k]);
                }
            }
            ok = ok && write_data(m, path + "/data.plt", vals, names, solution_time);
            for (int pass = 0; pass < 2; ++pass) {
                if (pass == 0 && !latest) continue;
                if (pass == 1 && !np) continue;
                const std::string d = pass == 0 ? path_latest : path;
                ok = ok && write_scalar(d + "Rho", rho, nc) && write_vector(d + "U", u3, nc);
                if (shak)
                    ok = ok && write_scalar(d + "T", T, nc) && write_vector(d + "q", q3, nc) &&
                         write_scalar(d + "fluxM0", flux_m0, have_flux ? nc : 0) &&
                         write_vector(d + "fluxM1", flux_m1_3, have_flux ? nc : 0) &&
                         write_scalar(d + "fluxM2", flux_m2, have_flux ? nc : 0);
            }
        }
        return ok ? MESO_OK : fail(MESO_E_IO, "cannot write the output files under " + dir);
    } catch (const std::exception &e) {
        return fail(MESO_E_IO, e.what());
    }
}

}  // extern "C"
