// Host front-end of the MI355X DUGKS/CDUGKS step: the case-file, gambit mesh,
// discrete-velocity-set and Field<T> API of meso, rewritten for this library.
//
// Interface parity (names, argument meaning, error behaviour) follows meso:
//   Vector                 include/core/math.h:10-38, src/core/math/vector.cpp
//   fvmMesh::Mesh/Cell/... include/mesh/object.h:83-179
//   Field<T>               include/mesh/object.h:42-75, src/solver/field.cpp
//   Solver::Config/Patch   include/solver/config.h:38-114, src/solver/config.cpp
//   DVS::generate_dvs      include/solver/dvs/dvs_reader.h:26-42
// Errors are reported by throwing std::invalid_argument / std::runtime_error
// as meso does (e.g. src/solver/field.cpp:20-21); the C ABI converts them to
// status codes.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace MESO {

using Scalar = double;
using ObjectId = int;
using ObjectType = int;
using String = std::string;
using StringList = std::vector<std::string>;
template <class T> using List = std::vector<T>;
template <class T> using Dict = std::unordered_map<std::string, T>;

// Upwind dead-band of the reconstruction (include/core/math.h:4).
constexpr double VSMALL = 1e-30;

struct Vector {
    double x = 0.0, y = 0.0, z = 0.0;
    Vector() = default;
    Vector(double a, double b, double c) : x(a), y(b), z(c) {}
    double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Vector operator-() const { return {-x, -y, -z}; }
    Vector operator+(const Vector &o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector operator-(const Vector &o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector operator*(double s) const { return {x * s, y * s, z * s}; }
    // meso divides by multiplying with the reciprocal (vector.cpp:79-82).
    Vector operator/(double s) const {
        double r = 1.0 / s;
        return {x * r, y * r, z * r};
    }
    Vector &operator+=(const Vector &o) { x += o.x; y += o.y; z += o.z; return *this; }
    Vector &operator-=(const Vector &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    Vector &operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
    // dot product, evaluated left to right as x*x' + y*y' + z*z'
    double operator*(const Vector &o) const { return x * o.x + y * o.y + z * o.z; }
    // cross product
    Vector operator^(const Vector &o) const {
        return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
    }
    double magnitude() const { return std::sqrt(x * x + y * y + z * z); }
    Vector normalize() const {
        double m = magnitude();
        if (m == 0.0) return {0.0, 0.0, 0.0};
        return *this / m;
    }
};
inline Vector operator*(double k, const Vector &v) { return {v.x * k, v.y * k, v.z * k}; }

namespace Geom {
constexpr ObjectType Edge = 1, Quad = 2, Tria = 3, Brick = 4, Wedge = 5, Tetra = 6, Pyram = 7;
int face_num(ObjectType t);
int node_num(ObjectType t);
}  // namespace Geom

namespace fvmMesh {
class Mesh;

struct LeastSquare {
    int neighbor_num = 0;
    List<Scalar> weight;
    List<Vector> dr;
    Vector Cx, Cy, Cz;
};

struct Symmetry {
    int id[3] = {0, 0, 0};
};

struct Node {
    ObjectId id = 0;
    ObjectId group_id = 0;
    Vector position;
};

struct Face {
    ObjectType geom_type = 0;
    ObjectId id = 0;
    ObjectId group_id = 0;  // 0: fluid-interior, k: k-th boundary set of the .neu file
    Vector position;
    Scalar area = 0.0;
    List<ObjectId> node_id;
    std::array<ObjectId, 2> cell_id{{-1, -1}};       // owner, neighbour (== owner on a boundary)
    std::array<ObjectId, 2> cell_face_id{{-1, -1}};
    std::array<Vector, 2> normal_vector;              // [0] out of owner, [1] = -[0]
};

struct Cell {
    ObjectType geom_type = 0;
    ObjectId id = 0;
    ObjectId group_id = 0;
    Vector position;
    Scalar volume = 0.0;
    List<ObjectId> node_id;
    List<ObjectId> face_id;
    List<ObjectId> neighbors;
    LeastSquare least_square;
    Symmetry symmetry;
};

class Mesh {
public:
    int NNODE = 0, NFACE = 0, NCELL = 0, NZONE = 0, NMARK = 0, NDFCD = 0, NDFVL = 0;
    double total_volume = 0.0;
    double min_cell_size = -1.0;
    double max_cell_magnitude = -1.0;
    List<Node> nodes;
    List<Face> faces;
    List<Cell> cells;
    List<String> cell_names;
    List<List<ObjectId>> cell_groups;
    List<String> face_names{"fluid-interior"};
    List<List<ObjectId>> face_groups;

    int dimension() const { return NDFVL; }
    void generate_face();
    void update_mesh_params();
    void build_geom(double scale_ratio = 1.0);
};

// Parse a gambit neutral file and build topology + geometry (src/mesh/mesh.cpp:4-11).
Mesh load_gambit(const String &file_path, double mesh_scale = 1.0);
}  // namespace fvmMesh

enum { cell_field_flag, face_field_flag, node_field_flag };

template <class T>
class Field {
    fvmMesh::Mesh *mesh_ptr_ = nullptr;

public:
    int len = 0;
    int flag = cell_field_flag;
    List<T> values;

    Field() = default;
    Field(fvmMesh::Mesh &mesh, int flag_) : mesh_ptr_(&mesh), flag(flag_) {
        switch (flag_) {
            case cell_field_flag: len = mesh.NCELL; break;
            case face_field_flag: len = mesh.NFACE; break;
            case node_field_flag: len = mesh.NNODE; break;
            default: throw std::invalid_argument("Field caught wrong flag");
        }
        values.assign(len, T{});
    }
    fvmMesh::Mesh *get_mesh() const { return mesh_ptr_; }
    T &operator[](int i) { return values[i]; }
    const T &operator[](int i) const { return values[i]; }
    void set_zero() { values.assign(values.size(), T{}); }
    // component extraction of a vector field (src/solver/field.cpp:97-105)
    Field<Scalar> heft(int id) const;
    // 18-significant-digit text, one value (or "x y z") per line (field.cpp:139-173)
    void output(const String &file_name) const;
};

template <class T>
void read_np_data(const String &file, Field<T> &field);

namespace Solver {

enum BoundaryType {
    fluid_interior,
    farfield_inlet,
    pressure_inlet,
    freestream_inlet,
    farfield_outlet,
    pressure_outlet,
    wall,
    slip_wall,
    symmetry,
};
extern const Dict<int> mark_type_map;

namespace PatchType {
constexpr ObjectType zeroGradient = 0;
constexpr ObjectType fixedValue = 1;
constexpr ObjectType calculated = 2;
constexpr ObjectType fromFile = 3;
constexpr ObjectType switchType = 4;
}  // namespace PatchType

class Patch {
    Dict<ObjectType> type_;
    Dict<ObjectType> ints_;
    Dict<bool> switches_{{"mask", false}};
    Dict<Scalar> scalars_;
    Dict<Vector> vectors_;
    Dict<List<ObjectType>> ints_list_;
    Dict<List<Scalar>> scalars_list_;
    Dict<List<Vector>> vectors_list_;

public:
    // "<key> <kind> [<valtype> <values...> | <valtype> <file>]" (src/solver/boundary/patch.cpp:5-52)
    void set_value(const StringList &tokens, const String &base_dir);
    // Missing keys read as 0 / false / zero vector, as std::unordered_map::operator[] does in meso.
    ObjectType get_type(const String &key) const;
    bool has(const String &key) const { return type_.count(key) != 0; }
    bool get_switch(const String &key) const;
    ObjectType get_int(const String &key) const;
    Scalar get_scalar(const String &key) const;
    Vector get_vector(const String &key) const;
    ObjectType get_file_int(const String &key, ObjectId id) const;
    Scalar get_file_scalar(const String &key, ObjectId id) const;
    Vector get_file_vector(const String &key, ObjectId id) const;
};

struct Mark {
    String name;
    ObjectId id = 0;
    ObjectType type = fluid_interior;
    String type_name;
    Patch patch;
};

struct Group {
    String name;
    ObjectId id = 0;
    Patch patch;
};

class Config {
    String file_path_;
    String base_dir_;

public:
    Dict<String> settings;
    Dict<Group> groups;
    Dict<Mark> marks;

    Config() = default;
    explicit Config(const String &file_path);
    // Re-read the case file; with update_patch=false only [settings] change
    // (src/solver/config.cpp:30-108).
    void update_config(bool update_patch = true);
    const String &file_path() const { return file_path_; }
    // Directory of the case file: relative mesh/dvs/fromFile paths resolve against it.
    const String &base_dir() const { return base_dir_; }
    String resolve(const String &path) const;

    bool has(const String &key) const { return settings.count(key) != 0; }
    String get(const String &key, const String &d) const;
    String get(const String &key, const char *d) const { return get(key, String(d)); }
    int get(const String &key, int d) const;
    double get(const String &key, double d) const;
    bool get(const String &key, bool d) const;

    const Group &get_cell_group(const fvmMesh::Cell &cell, const fvmMesh::Mesh &mesh) const;
    const Mark &get_face_group(const fvmMesh::Face &face, const fvmMesh::Mesh &mesh) const;
};

// Residual of src/solver/field.cpp:222-249: sum |(new-old)/old| * V_c / V_total,
// then old := new.
Scalar residual(Field<Scalar> &old_, const Field<Scalar> &new_);
Vector residual(Field<Vector> &old_, const Field<Vector> &new_);
constexpr int residual_interval = 100;  // include/solver/solver.h:49
bool is_converged(const List<Scalar> &residuals, Scalar limit);  // src/core/core.cpp:45-51

}  // namespace Solver

namespace DVS {
// A discrete velocity set is a list of "cells": position = xi, volume = weight,
// symmetry.id = mirror velocity ids (src/solver/dvs/dvs_generator.cpp).
struct VelocitySet {
    List<Vector> xi;
    List<Scalar> weight;
    List<std::array<int, 3>> mirror;  // empty when the generator defines none
    double max_magnitude = 0.0;
    int size() const { return int(xi.size()); }
};
VelocitySet gauss_hermite(const String &file, int dimension, double RT);        // :15-70
VelocitySet half_range_gauss_hermite(const String &file, int dimension, double RT);  // :72-148
VelocitySet newton_cotes(const String &file, int dimension);                   // :150-244
VelocitySet sparse(const String &file);                                       // :246-276
VelocitySet from_gambit(const String &file);  // cell centroid = xi, cell volume = w
}  // namespace DVS

namespace Utils {
StringList split(const String &s);
StringList read_lines(const String &file);  // non-empty lines, '\r' stripped
}  // namespace Utils

}  // namespace MESO
