// Case-file parser, boundary/group patches, Field<T> text I/O, residual, and
// the discrete-velocity-set generators.
//
// Behaviour follows meso:
//   Config::update_config   src/solver/config.cpp:30-108 (keys, "#" lines, sections)
//   Config::get<T>          src/solver/config.cpp:110-152 ("True" booleans)
//   Patch::set_value        src/solver/boundary/patch.cpp:5-52
//   Field::output / np I/O  src/solver/field.cpp:139-173, 252-273; src/core/core.cpp:100-133
//   residual                src/solver/field.cpp:222-249
//   DVS generators          src/solver/dvs/dvs_generator.cpp:15-276
#include "meso_host.h"

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <sys/stat.h>

namespace MESO {

namespace {
bool file_exists(const String &p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}
String dir_of(const String &p) {
    auto pos = p.find_last_of('/');
    return pos == String::npos ? String(".") : p.substr(0, pos);
}
template <class T>
List<T> read_np_list(const String &file);
template <>
List<ObjectType> read_np_list<ObjectType>(const String &file) {
    List<ObjectType> r;
    for (auto &l : Utils::read_lines(file)) r.push_back(std::stoi(Utils::split(l).at(0)));
    return r;
}
template <>
List<Scalar> read_np_list<Scalar>(const String &file) {
    List<Scalar> r;
    for (auto &l : Utils::read_lines(file)) r.push_back(std::stod(Utils::split(l).at(0)));
    return r;
}
template <>
List<Vector> read_np_list<Vector>(const String &file) {
    List<Vector> r;
    for (auto &l : Utils::read_lines(file)) {
        auto d = Utils::split(l);
        r.emplace_back(std::stod(d.at(0)), std::stod(d.at(1)), std::stod(d.at(2)));
    }
    return r;
}
}  // namespace

// ---------------------------------------------------------------- Field<T>
template <>
Field<Scalar> Field<Vector>::heft(int id) const {
    Field<Scalar> r(*mesh_ptr_, flag);
    for (int i = 0; i < len; ++i) r.values[i] = values[i][id];
    return r;
}

template <>
void Field<Scalar>::output(const String &file_name) const {
    std::ofstream fp(file_name, std::ios::out | std::ios::trunc);
    if (!fp.is_open()) throw std::runtime_error("Cannot open file: " + file_name);
    fp << std::setprecision(18);
    for (auto &v : values) fp << v << "\n";
}

template <>
void Field<Vector>::output(const String &file_name) const {
    std::ofstream fp(file_name, std::ios::out | std::ios::trunc);
    if (!fp.is_open()) throw std::runtime_error("Cannot open file: " + file_name);
    fp << std::setprecision(18);
    for (auto &v : values) fp << v.x << " " << v.y << " " << v.z << "\n";
}

template <>
void read_np_data<Scalar>(const String &file, Field<Scalar> &field) {
    auto lines = Utils::read_lines(file);
    for (int i = 0; i < int(lines.size()) && i < field.len; ++i) field[i] = std::stod(lines[i]);
}

template <>
void read_np_data<Vector>(const String &file, Field<Vector> &field) {
    auto lines = Utils::read_lines(file);
    for (int i = 0; i < int(lines.size()) && i < field.len; ++i) {
        auto d = Utils::split(lines[i]);
        field[i] = {std::stod(d.at(0)), std::stod(d.at(1)), std::stod(d.at(2))};
    }
}

namespace Solver {

const Dict<int> mark_type_map = {
    {"fluid-interior", fluid_interior}, {"farfield-inlet", farfield_inlet},
    {"pressure-inlet", pressure_inlet}, {"freestream-inlet", freestream_inlet},
    {"farfield-outlet", farfield_outlet}, {"pressure-outlet", pressure_outlet},
    {"wall", wall}, {"slip-wall", slip_wall}, {"symmetry", symmetry},
};

// ---------------------------------------------------------------- Patch
void Patch::set_value(const StringList &t, const String &base_dir) {
    if (t.size() < 2) return;
    const String &name = t[0];
    const String &kind = t[1];
    auto need = [&](size_t n) {
        if (t.size() < n) throw std::invalid_argument("Patch::set_value: too few tokens for " + name);
    };
    auto resolve = [&](const String &p) {
        if (p.empty() || p[0] == '/' || file_exists(p)) return p;
        return base_dir + "/" + p;
    };
    if (kind == "calculated") {
        type_[name] = PatchType::calculated;
    } else if (kind == "fixedValue") {
        need(4);
        type_[name] = PatchType::fixedValue;
        const String &vt = t[2];
        if (vt == "objectType") ints_[name] = std::stoi(t[3]);
        else if (vt == "scalar") scalars_[name] = std::stod(t[3]);
        else if (vt == "vector") {
            need(6);
            vectors_[name] = {std::stod(t[3]), std::stod(t[4]), std::stod(t[5])};
        } else
            throw std::invalid_argument("Patch::set_value caught unsupported type: " + name + " - " + kind);
    } else if (kind == "zeroGradient") {
        type_[name] = PatchType::zeroGradient;
    } else if (kind == "fromFile") {
        need(4);
        type_[name] = PatchType::fromFile;
        const String &vt = t[2];
        String file = resolve(t[3]);
        if (vt == "objectType") ints_list_[name] = read_np_list<ObjectType>(file);
        else if (vt == "scalar") scalars_list_[name] = read_np_list<Scalar>(file);
        else if (vt == "vector") vectors_list_[name] = read_np_list<Vector>(file);
        else
            throw std::invalid_argument("Patch::set_value caught unsupported type: " + name + " - " + kind);
    } else if (kind == "switch") {
        need(3);
        type_[name] = PatchType::switchType;
        switches_[name] = (t[2] == "True");
    }
}

template <class M, class K>
static typename M::mapped_type lookup(const M &m, const K &k) {
    auto it = m.find(k);
    return it == m.end() ? typename M::mapped_type{} : it->second;
}

ObjectType Patch::get_type(const String &k) const { return lookup(type_, k); }
bool Patch::get_switch(const String &k) const { return lookup(switches_, k); }
ObjectType Patch::get_int(const String &k) const { return lookup(ints_, k); }
Scalar Patch::get_scalar(const String &k) const { return lookup(scalars_, k); }
Vector Patch::get_vector(const String &k) const { return lookup(vectors_, k); }
ObjectType Patch::get_file_int(const String &k, ObjectId id) const {
    auto it = ints_list_.find(k);
    if (it == ints_list_.end() || id >= int(it->second.size())) throw std::invalid_argument("fromFile patch too short: " + k);
    return it->second[id];
}
Scalar Patch::get_file_scalar(const String &k, ObjectId id) const {
    auto it = scalars_list_.find(k);
    if (it == scalars_list_.end() || id >= int(it->second.size())) throw std::invalid_argument("fromFile patch too short: " + k);
    return it->second[id];
}
Vector Patch::get_file_vector(const String &k, ObjectId id) const {
    auto it = vectors_list_.find(k);
    if (it == vectors_list_.end() || id >= int(it->second.size())) throw std::invalid_argument("fromFile patch too short: " + k);
    return it->second[id];
}

// ---------------------------------------------------------------- Config
Config::Config(const String &file_path) : file_path_(file_path), base_dir_(dir_of(file_path)) {
    Mark interior;
    interior.id = 1;
    interior.name = interior.type_name = "fluid-interior";
    interior.type = fluid_interior;
    marks[interior.name] = interior;
    update_config(true);
}

String Config::resolve(const String &p) const {
    if (p.empty() || p[0] == '/' || file_exists(p)) return p;
    return base_dir_ + "/" + p;
}

void Config::update_config(bool update_patch) {
    StringList lines = Utils::read_lines(file_path_);
    const int n = int(lines.size());
    auto is_section = [](const String &s) { return s == "[settings]" || s == "[mark]" || s == "[group]"; };
    for (int i = 0; i < n; ++i) {
        StringList d = Utils::split(lines[i]);
        if (d.empty() || d[0] == "#") continue;
        if (d[0] == "[settings]") {
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d.empty() || d[0] == "#") continue;
                if (is_section(d[0])) { --i; break; }
                if (d.size() >= 2) settings[d[0]] = d[1];
            }
        } else if (d[0] == "[mark]") {
            Mark mark;
            mark.id = int(marks.size()) + 1;
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d.empty() || d[0] == "#") continue;
                if (is_section(d[0])) { --i; break; }
                if (!update_patch) continue;
                if (d[0] == "name" && d.size() >= 2) mark.name = d[1];
                else if (d[0] == "type" && d.size() >= 2) {
                    mark.type_name = d[1];
                    auto it = mark_type_map.find(d[1]);
                    if (it == mark_type_map.end()) throw std::invalid_argument("unknown mark type: " + d[1]);
                    mark.type = it->second;
                } else
                    mark.patch.set_value(d, base_dir_);
            }
            if (update_patch) marks[mark.name] = mark;
        } else if (d[0] == "[group]") {
            Group group;
            group.id = int(groups.size()) + 1;
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d.empty() || d[0] == "#") continue;
                if (is_section(d[0])) { --i; break; }
                if (!update_patch) continue;
                if (d[0] == "name" && d.size() >= 2) group.name = d[1];
                else group.patch.set_value(d, base_dir_);
            }
            if (update_patch) groups[group.name] = group;
        }
    }
}

String Config::get(const String &key, const String &d) const {
    auto it = settings.find(key);
    return it == settings.end() ? d : it->second;
}
int Config::get(const String &key, int d) const {
    auto it = settings.find(key);
    return it == settings.end() ? d : std::stoi(it->second);
}
double Config::get(const String &key, double d) const {
    auto it = settings.find(key);
    return it == settings.end() ? d : std::stod(it->second);
}
bool Config::get(const String &key, bool d) const {
    auto it = settings.find(key);
    return it == settings.end() ? d : (it->second == "True");
}

const Group &Config::get_cell_group(const fvmMesh::Cell &cell, const fvmMesh::Mesh &mesh) const {
    const String &key = mesh.cell_names.at(cell.group_id);
    auto it = groups.find(key);
    if (it == groups.end()) throw std::invalid_argument("get_cell_group() missing key=" + key);
    return it->second;
}

const Mark &Config::get_face_group(const fvmMesh::Face &face, const fvmMesh::Mesh &mesh) const {
    const String &key = mesh.face_names.at(face.group_id);
    auto it = marks.find(key);
    if (it == marks.end()) throw std::invalid_argument("get_face_group() missing key=" + key);
    return it->second;
}

// ---------------------------------------------------------------- residual
Scalar residual(Field<Scalar> &o, const Field<Scalar> &n) {
    auto &mesh = *o.get_mesh();
    Scalar r = 0.0;
    for (auto &c : mesh.cells) r += std::abs((n[c.id] - o[c.id]) / o[c.id]) * (c.volume / mesh.total_volume);
    o.values = n.values;
    return r;
}

Vector residual(Field<Vector> &o, const Field<Vector> &n) {
    auto &mesh = *o.get_mesh();
    Vector r;
    for (auto &c : mesh.cells) {
        const Vector &a = n[c.id], &b = o[c.id];
        Vector v{std::fabs((a.x - b.x) / b.x), std::fabs((a.y - b.y) / b.y), std::fabs((a.z - b.z) / b.z)};
        r += v * (c.volume / mesh.total_volume);
    }
    o.values = n.values;
    return r;
}

bool is_converged(const List<Scalar> &res, Scalar limit) {
    bool ok = true;
    for (auto v : res) ok = ok && (v <= limit);
    return ok;
}

}  // namespace Solver

// ---------------------------------------------------------------- DVS
namespace DVS {
namespace {
StringList dvs_lines(const String &file, const char *want) {
    auto lines = Utils::read_lines(file);
    if (lines.empty() || Utils::split(lines[0]).at(0) != want)
        throw std::invalid_argument(String("generate_dvs type error: expected ") + want);
    return lines;
}
void finish(VelocitySet &s) {
    s.max_magnitude = 0.0;
    for (auto &x : s.xi) s.max_magnitude = std::max(s.max_magnitude, x.magnitude());
}
}  // namespace

VelocitySet gauss_hermite(const String &file, int dim, double RT) {
    auto lines = dvs_lines(file, "Gauss-Hermite");
    const int n = int(lines.size()) - 1;
    const double xs = std::sqrt(2.0 * RT), ws = 1.0 / std::sqrt(M_PI);
    List<double> x(n), w(n);
    for (int i = 0; i < n; ++i) {
        auto d = Utils::split(lines[i + 1]);
        x[i] = xs * std::stod(d.at(0));
        w[i] = ws * std::stod(d.at(1));
    }
    VelocitySet s;
    // storage order: i outer, j inner (dvs_generator.cpp:43-52)
    if (dim == 2) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                s.xi.emplace_back(x[i], x[j], 0.0);
                s.weight.push_back(w[i] * w[j]);
            }
    } else {
        // meso multiplies the abscissae, not the weights, in 3D (dvs_generator.cpp:59);
        // the same product is kept so that a 3D Gauss-Hermite case behaves identically.
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) {
                    s.xi.emplace_back(x[i], x[j], x[k]);
                    s.weight.push_back(x[i] * x[j] * x[k]);
                }
    }
    finish(s);
    return s;
}

VelocitySet half_range_gauss_hermite(const String &file, int dim, double RT) {
    auto lines = dvs_lines(file, "half-range-Gauss-Hermite");
    const int n = int(lines.size()) - 1;
    const double xs = std::sqrt(2.0 * RT);
    List<double> x(n), w(n);
    for (int i = 0; i < n; ++i) {
        auto d = Utils::split(lines[i + 1]);
        x[i] = xs * std::stod(d.at(0));
        w[i] = std::stod(d.at(1));
    }
    VelocitySet s;
    auto mir = [n](int i) { return n - i - 1; };
    if (dim == 2) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                s.xi.emplace_back(x[i], x[j], 0.0);
                s.weight.push_back(w[i] * w[j]);
                s.mirror.push_back({mir(i) + j * n, i + mir(j) * n, i + j * n});
            }
    } else {
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    s.xi.emplace_back(x[i], x[j], x[k]);
                    s.weight.push_back(x[i] * x[j] * x[k]);  // as dvs_generator.cpp:133
                    s.mirror.push_back({mir(i) + (j + k * n) * n, i + (mir(j) + k * n) * n, i + (j + mir(k) * n) * n});
                }
    }
    finish(s);
    return s;
}

VelocitySet newton_cotes(const String &file, int dim) {
    auto lines = dvs_lines(file, "Newton-Cotes");
    auto d = Utils::split(lines.at(1));
    const int deg = std::stoi(d.at(0));
    const int mount = std::stoi(d.at(1));
    const double scale = std::stod(d.at(2));
    // closed Newton-Cotes panel weights of degree 1..6 (dvs_generator.cpp:152-159)
    static const std::vector<std::vector<double>> table = {
        {1. / 2., 1. / 2.},
        {1. / 6., 4. / 6., 1. / 6.},
        {1. / 8., 3. / 8., 3. / 8., 1. / 8.},
        {7. / 90., 16. / 45., 2. / 15., 16. / 45., 7. / 90.},
        {19. / 288., 25. / 96., 25. / 144., 25. / 144., 25. / 96., 19. / 288.},
        {41. / 840., 9. / 35., 9. / 280., 34. / 105., 9. / 280., 9. / 35., 41. / 840.}};
    if (deg < 1 || deg > 6 || mount < 1) throw std::invalid_argument("Newton-Cotes: degree must be 1..6");
    const auto &pw = table[deg - 1];
    const int N = deg * mount + 1;
    const double dh = 2.0 * scale / mount;
    List<double> c(N), w(N, 0.0);
    for (int i = 0; i < N; ++i) c[i] = scale * double(2 * i + 1 - N) / double(N - 1);
    for (int m = 0; m < mount; ++m)
        for (int j = 0; j <= deg; ++j) w[m * deg + j] += pw[j] * dh;
    VelocitySet s;
    auto mir = [N](int i) { return N - i - 1; };
    if (dim == 2) {
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                s.xi.emplace_back(c[i], c[j], 0.0);
                s.weight.push_back(w[i] * w[j]);
                s.mirror.push_back({mir(i) + j * N, i + mir(j) * N, i + j * N});
            }
    } else {
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    s.xi.emplace_back(c[i], c[j], c[k]);
                    s.weight.push_back(w[i] * w[j] * w[k]);
                    s.mirror.push_back({mir(i) + (j + k * N) * N, i + (mir(j) + k * N) * N, i + (j + mir(k) * N) * N});
                }
    }
    finish(s);
    return s;
}

VelocitySet sparse(const String &file) {
    auto lines = dvs_lines(file, "Sparse");
    VelocitySet s;
    for (size_t i = 1; i < lines.size(); ++i) {
        auto d = Utils::split(lines[i]);
        s.xi.emplace_back(std::stod(d.at(0)), std::stod(d.at(1)), std::stod(d.at(2)));
        s.weight.push_back(std::stod(d.at(3)));
    }
    finish(s);
    return s;
}

VelocitySet from_gambit(const String &file) {
    auto m = fvmMesh::load_gambit(file, 1.0);
    VelocitySet s;
    for (auto &c : m.cells) {
        s.xi.push_back(c.position);
        s.weight.push_back(c.volume);
    }
    finish(s);
    return s;
}

}  // namespace DVS
}  // namespace MESO
