// Gambit neutral-file reader, face generation and cell/face geometry.
//
// Behaviour follows meso:
//   reader            src/mesh/mesh_reader.cpp:8-139
//   face generation   src/mesh/object.cpp:20-84 (local face tables :20-36)
//   geometry          src/mesh/object.cpp:94-172, src/mesh/geom.cpp:68-131,213-258
//   least squares     src/mesh/geom.cpp:165-211
// The 2D face normal is built from arc-cosines of the edge direction exactly as
// geom.cpp:113-127 does: the result differs from the rotated tangent in the last
// bits (cos(pi/2) != 0), and those bits decide the upwind side of velocities
// that are parallel to a face, so the same evaluation is kept here.
#include "meso_host.h"

#include <algorithm>
#include <fstream>
#include <map>
#include <cstdlib>
#include <sstream>

namespace MESO {

namespace Geom {
int face_num(ObjectType t) {
    switch (t) {
        case Tria: return 3;
        case Quad: case Tetra: return 4;
        case Wedge: case Pyram: return 5;
        case Brick: return 6;
        default: throw std::invalid_argument("<Geom::face_num> unsupported geom_type");
    }
}
int node_num(ObjectType t) {
    switch (t) {
        case Edge: return 2;
        case Tria: return 3;
        case Quad: case Tetra: return 4;
        case Pyram: return 5;
        case Wedge: return 6;
        case Brick: return 8;
        default: throw std::invalid_argument("<Geom::node_num> unsupported geom_type");
    }
}
}  // namespace Geom

namespace Utils {
namespace {
inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }
}  // namespace
// Whitespace-separated tokens (the tokens of an istringstream >> loop).
StringList split(const String &s) {
    StringList out;
    const char *p = s.data(), *e = p + s.size();
    while (p < e) {
        while (p < e && is_space(*p)) ++p;
        const char *b = p;
        while (p < e && !is_space(*p)) ++p;
        if (p > b) out.emplace_back(b, size_t(p - b));
    }
    return out;
}
// Non-empty lines of a text file with '\r' removed (read in one piece: a
// 640,000-cell gambit file is 76 MB).
StringList read_lines(const String &file) {
    std::ifstream fp(file, std::ios::binary | std::ios::ate);
    if (!fp.is_open()) throw std::invalid_argument("file open error: " + file);
    const std::streamsize n = fp.tellg();
    std::string buf(size_t(std::max<std::streamsize>(n, 0)), '\0');
    fp.seekg(0);
    if (n > 0 && !fp.read(&buf[0], n)) throw std::invalid_argument("file read error: " + file);
    StringList lines;
    const char *p = buf.data(), *e = p + buf.size();
    while (p < e) {
        const char *b = p;
        while (p < e && *p != '\n') ++p;
        const char *q = p;
        if (p < e) ++p;
        if (std::find(b, q, '\r') == q) {
            if (q > b) lines.emplace_back(b, size_t(q - b));
        } else {
            String line(b, size_t(q - b));
            line.erase(std::remove(line.begin(), line.end(), '\r'), line.end());
            if (!line.empty()) lines.push_back(std::move(line));
        }
    }
    return lines;
}
}  // namespace Utils

namespace fvmMesh {
namespace {

using namespace Geom;

// Face kinds and local node tables of each cell kind (object.cpp:20-36).
const std::vector<int> &face_kinds(ObjectType t) {
    static const std::map<int, std::vector<int>> m = {
        {Quad, {Edge, Edge, Edge, Edge}},
        {Tria, {Edge, Edge, Edge}},
        {Brick, {Quad, Quad, Quad, Quad, Quad, Quad}},
        {Wedge, {Quad, Quad, Quad, Tria, Tria}},
        {Tetra, {Tria, Tria, Tria, Tria}},
        {Pyram, {Quad, Tria, Tria, Tria, Tria}}};
    auto it = m.find(t);
    if (it == m.end()) throw std::invalid_argument("<generate_face> unsupported cell kind");
    return it->second;
}
const std::vector<std::vector<int>> &face_nodes(ObjectType t) {
    static const std::map<int, std::vector<std::vector<int>>> m = {
        {Quad, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}},
        {Tria, {{0, 1}, {1, 2}, {2, 0}}},
        {Brick, {{0, 1, 5, 4}, {1, 3, 7, 5}, {3, 2, 6, 7}, {2, 0, 4, 6}, {1, 0, 2, 3}, {4, 5, 7, 6}}},
        {Wedge, {{0, 1, 4, 3}, {1, 2, 5, 4}, {2, 0, 3, 5}, {0, 2, 1}, {3, 4, 5}}},
        {Tetra, {{1, 0, 2}, {0, 1, 3}, {1, 2, 3}, {2, 0, 3}}},
        {Pyram, {{0, 2, 3, 1}, {0, 1, 4}, {1, 3, 4}, {3, 2, 4}, {2, 0, 4}}}};
    auto it = m.find(t);
    if (it == m.end()) throw std::invalid_argument("<generate_face> unsupported cell kind");
    return it->second;
}

Vector centroid(const std::vector<const Node *> &nl) {
    Vector p;
    for (auto *n : nl) p += n->position;
    return p * (1.0 / double(nl.size()));
}

double tria_area(const Vector &a, const Vector &b, const Vector &c) {
    return 0.5 * ((b - a) ^ (c - a)).magnitude();
}
double tetra_volume(const Vector &a, const Vector &b, const Vector &c, const Vector &d) {
    Vector p1 = b - a, p2 = c - a, p3 = d - a;
    return std::abs(p1 * (p2 ^ p3)) / 6.0;
}
// pyramid of base (0,1,3,2) split into two tetrahedra (geom.cpp:248-251)
double pyram_volume(const Vector *n) {
    return tetra_volume(n[0], n[1], n[2], n[4]) + tetra_volume(n[1], n[2], n[3], n[4]);
}
// wedge = pyramid + tetrahedron (geom.cpp:243-246)
double wedge_volume(const Vector *n) {
    Vector p[5] = {n[0], n[2], n[5], n[3], n[1]};
    return pyram_volume(p) + tetra_volume(n[1], n[3], n[4], n[5]);
}
// brick = two wedges (geom.cpp:238-241)
double brick_volume(const Vector *n) {
    Vector a[6] = {n[0], n[1], n[5], n[2], n[3], n[7]};
    Vector b[6] = {n[0], n[4], n[5], n[2], n[6], n[7]};
    return wedge_volume(a) + wedge_volume(b);
}

double face_area(ObjectType t, const std::vector<const Node *> &nl) {
    switch (t) {
        case Edge: return (nl[1]->position - nl[0]->position).magnitude();
        case Tria: return tria_area(nl[0]->position, nl[1]->position, nl[2]->position);
        case Quad:
            return tria_area(nl[0]->position, nl[1]->position, nl[2]->position) +
                   tria_area(nl[0]->position, nl[3]->position, nl[2]->position);
        default: throw std::invalid_argument("<Geom::calculate_area> unsupported geom_type");
    }
}

double cell_volume(ObjectType t, const std::vector<const Node *> &nl) {
    Vector p[8];
    for (size_t i = 0; i < nl.size() && i < 8; ++i) p[i] = nl[i]->position;
    switch (t) {
        case Tria: return tria_area(p[0], p[1], p[2]);
        case Quad: return tria_area(p[0], p[1], p[2]) + tria_area(p[0], p[3], p[2]);
        case Brick: return brick_volume(p);
        case Wedge: return wedge_volume(p);
        case Tetra: return tetra_volume(p[0], p[1], p[2], p[3]);
        case Pyram: return pyram_volume(p);
        default: throw std::invalid_argument("<Geom::calculate_volume> unsupported geom_type");
    }
}

double angle_2d(const Vector &a, const Vector &b) {
    return std::acos((a * b) / (a.magnitude() * b.magnitude()));
}

// Unit normal of a face, oriented away from the given cell centre (geom.cpp:109-131).
Vector face_normal(ObjectType t, const Vector &cell_pos, const Vector &face_pos,
                   const std::vector<const Node *> &nl) {
    Vector vcf = face_pos - cell_pos;
    Vector t01 = (nl[1]->position - nl[0]->position).normalize();
    Vector nv;
    if (t == Edge) {
        double ax = angle_2d(t01, Vector(1.0, 0.0, 0.0));
        double ay = angle_2d(t01, Vector(0.0, 1.0, 0.0));
        ax = (ay <= M_PI_2) ? (M_PI_2 + ax) : (M_PI_2 - ax);
        nv = Vector(std::cos(ax), std::sin(ax), 0.0).normalize();
    } else {
        Vector e10 = nl[0]->position - nl[1]->position;
        Vector e12 = nl[2]->position - nl[1]->position;
        nv = (e10 ^ e12).normalize();
    }
    return (nv * vcf >= 0.0) ? nv : -nv;
}

}  // namespace

void Mesh::update_mesh_params() {
    NNODE = int(nodes.size());
    NFACE = int(faces.size());
    NCELL = int(cells.size());
    NZONE = int(cell_groups.size());
    NMARK = int(face_groups.size());
}

namespace {
// A face's sorted node ids (at most 4) and kind: the key under which the
// cells that share the face meet.
struct FaceKey {
    std::array<ObjectId, 5> v{{-1, -1, -1, -1, 0}};
    bool operator==(const FaceKey &o) const { return v == o.v; }
};
}  // namespace

void Mesh::generate_face() {
    // Faces are numbered in order of first appearance; the first cell that
    // meets a face owns it. Two cells share a face when its node sets match.
    // The faces met so far are listed per smallest node id (a linked list in
    // flat arrays: head per node, next per face), so a lookup walks the few
    // faces around one node.
    ObjectId nmax = 0;
    for (const auto &cell : cells)
        for (ObjectId n : cell.node_id) nmax = std::max(nmax, n + 1);
    std::vector<ObjectId> head(size_t(nmax), -1), next;
    std::vector<FaceKey> keys;
    faces.clear();
    faces.reserve(cells.size() * 3);
    next.reserve(cells.size() * 3);
    keys.reserve(cells.size() * 3);
    for (auto &cell : cells) {
        const auto &kinds = face_kinds(cell.geom_type);
        const auto &tbl = face_nodes(cell.geom_type);
        cell.face_id.assign(kinds.size(), -1);
        for (size_t fi = 0; fi < kinds.size(); ++fi) {
            std::vector<ObjectId> nl;
            for (int k : tbl[fi]) nl.push_back(cell.node_id.at(k));
            if (nl.size() > 4) throw std::invalid_argument("<generate_face> a face with more than 4 nodes");
            FaceKey key;
            std::copy(nl.begin(), nl.end(), key.v.begin());
            std::sort(key.v.begin(), key.v.begin() + nl.size());
            key.v[4] = -kinds[fi];  // keep edge/tria/quad namespaces apart
            const ObjectId m = key.v[0];
            if (m < 0) throw std::invalid_argument("<generate_face> negative node id");
            ObjectId j = head[m];
            while (j >= 0 && !(keys[j] == key)) j = next[j];
            if (j < 0) {
                Face f;
                f.id = int(faces.size());
                f.geom_type = kinds[fi];
                f.node_id = nl;
                f.cell_id = {cell.id, cell.id};
                f.cell_face_id = {int(fi), int(fi)};
                keys.push_back(key);
                next.push_back(head[m]);
                head[m] = f.id;
                cell.face_id[fi] = f.id;
                faces.push_back(std::move(f));
            } else {
                Face &f = faces[j];
                f.cell_id[1] = cell.id;
                f.cell_face_id[1] = int(fi);
                cell.face_id[fi] = f.id;
            }
        }
    }
    NFACE = int(faces.size());
}

void Mesh::build_geom(double scale) {
    for (auto &n : nodes) n.position *= scale;
    update_mesh_params();
    const int D = dimension();
    total_volume = 0.0;
    min_cell_size = -1.0;
    max_cell_magnitude = -1.0;
    std::vector<const Node *> nl;
    for (auto &cell : cells) {
        nl.clear();
        for (int id : cell.node_id) nl.push_back(&nodes.at(id));
        cell.position = centroid(nl);
        cell.volume = cell_volume(cell.geom_type, nl);
        total_volume += cell.volume;
        double mcs = (D == 2) ? 2.0 * std::sqrt(cell.volume / M_PI)
                              : 2.0 * std::pow(3.0 * cell.volume / (4.0 * M_PI), 1.0 / 3.0);
        double mcm = cell.position.magnitude();
        min_cell_size = (min_cell_size < 0.0) ? mcs : std::min(min_cell_size, mcs);
        max_cell_magnitude = (max_cell_magnitude < 0.0) ? mcm : std::max(max_cell_magnitude, mcm);
    }
    for (auto &face : faces) {
        nl.clear();
        for (int id : face.node_id) {
            nl.push_back(&nodes.at(id));
            nodes[id].group_id = face.group_id;
        }
        face.position = centroid(nl);
        face.area = face_area(face.geom_type, nl);
        face.normal_vector[0] = face_normal(face.geom_type, cells[face.cell_id[0]].position, face.position, nl);
        face.normal_vector[1] = -face.normal_vector[0];
    }
    // face-sharing neighbours in cell.face_id order, then least-squares rows
    for (auto &cell : cells) {
        cell.neighbors.clear();
        for (int fid : cell.face_id) {
            for (int cid : faces[fid].cell_id) {
                if (cid == cell.id) continue;
                if (std::find(cell.neighbors.begin(), cell.neighbors.end(), cid) != cell.neighbors.end()) continue;
                cell.neighbors.push_back(cid);
            }
        }
        auto &ls = cell.least_square;
        ls.neighbor_num = int(cell.neighbors.size());
        ls.dr.assign(ls.neighbor_num, Vector());
        ls.weight.assign(ls.neighbor_num, 0.0);
        double Sxx = 0, Sxy = 0, Sxz = 0, Syy = 0, Syz = 0, Szz = 0;
        for (int i = 0; i < ls.neighbor_num; ++i) {
            Vector dr = cells[cell.neighbors[i]].position - cell.position;
            double wi = 1.0 / (dr * dr);
            ls.dr[i] = dr;
            ls.weight[i] = wi;
            Sxx += wi * (dr.x * dr.x);
            Sxy += wi * (dr.x * dr.y);
            Sxz += wi * (dr.x * dr.z);
            Syy += wi * (dr.y * dr.y);
            Syz += wi * (dr.y * dr.z);
            Szz += wi * (dr.z * dr.z);
        }
        if (D == 2) {
            double FM = Sxx * Syy - Sxy * Sxy;
            ls.Cx = {Syy / FM, -Sxy / FM, 0.0};
            ls.Cy = {-Sxy / FM, Sxx / FM, 0.0};
            ls.Cz = {0.0, 0.0, 0.0};
        } else {
            // inverse of the symmetric 3x3 normal matrix by cofactors
            double FM = -Sxz * Sxz * Syy + 2.0 * Sxy * Sxz * Syz - Sxx * Syz * Syz - Sxy * Sxy * Szz +
                        Sxx * Syy * Szz;
            double a = Sxy * Syz - Sxz * Syy;
            double b = Sxz * Syz - Sxy * Szz;
            double c = Syy * Szz - Syz * Syz;
            double d = Sxx * Szz - Sxz * Sxz;
            double e = Sxy * Sxz - Sxx * Syz;
            double f = Sxx * Syy - Sxy * Sxy;
            ls.Cx = {c / FM, b / FM, a / FM};
            ls.Cy = {b / FM, d / FM, e / FM};
            ls.Cz = {a / FM, e / FM, f / FM};
        }
    }
}

namespace {
int to_int(const String &s) {
    char *end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) throw std::invalid_argument("not an integer: " + s);
    return int(v);
}
}

Mesh load_gambit(const String &file_path, double mesh_scale) {
    StringList lines = Utils::read_lines(file_path);
    const int n = int(lines.size());
    Mesh mesh;
    for (int i = 0; i < n; ++i) {
        StringList d = Utils::split(lines[i]);
        if (d.empty()) continue;
        if (d[0] == "NUMNP") {
            d = Utils::split(lines.at(++i));
            if (d.size() < 6) throw std::invalid_argument("gambit header: expected 6 counts");
            mesh.NNODE = to_int(d[0]);
            mesh.NCELL = to_int(d[1]);
            mesh.NZONE = to_int(d[2]);
            mesh.NMARK = to_int(d[3]);
            mesh.NDFCD = to_int(d[4]);
            mesh.NDFVL = to_int(d[5]);
            mesh.nodes.reserve(mesh.NNODE);
            mesh.cells.reserve(mesh.NCELL);
            mesh.cell_groups.resize(mesh.NZONE);
            mesh.face_groups.resize(mesh.NMARK + 1);
            continue;
        }
        if (d.size() >= 2 && d[0] == "NODAL" && d[1] == "COORDINATES") {
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                Node node;
                node.id = to_int(d[0]) - 1;
                if (d.size() >= 3) {
                    node.position.x = std::stod(d[1]);
                    node.position.y = std::stod(d[2]);
                }
                if (d.size() >= 4) node.position.z = std::stod(d[3]);
                mesh.nodes.push_back(node);
            }
            continue;
        }
        if (d[0] == "ELEMENTS/CELLS") {
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                Cell cell;
                cell.id = to_int(d[0]) - 1;
                cell.geom_type = to_int(d[1]);
                size_t count = size_t(to_int(d[2]));
                for (size_t j = 3; j < d.size(); ++j) cell.node_id.push_back(to_int(d[j]) - 1);
                if (cell.node_id.size() < count) {  // list continues on one more line
                    d = Utils::split(lines.at(++i));
                    for (auto &t : d) cell.node_id.push_back(to_int(t) - 1);
                }
                if (cell.id != int(mesh.cells.size()))
                    throw std::invalid_argument("gambit cells must be numbered 1..NELEM in order");
                mesh.cells.push_back(std::move(cell));
            }
            mesh.generate_face();
            continue;
        }
        if (d.size() >= 2 && d[0] == "ELEMENT" && d[1] == "GROUP") {
            d = Utils::split(lines.at(++i));
            int gid = to_int(d.at(1)) - 1;
            d = Utils::split(lines.at(++i));
            mesh.cell_names.push_back(d.at(0));
            if (gid >= int(mesh.cell_groups.size())) mesh.cell_groups.resize(gid + 1);
            for (i += 2; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                for (auto &t : d) {
                    int cid = to_int(t) - 1;
                    mesh.cells.at(cid).group_id = gid;
                    mesh.cell_groups[gid].push_back(cid);
                }
            }
            continue;
        }
        if (d.size() >= 2 && d[0] == "BOUNDARY" && d[1] == "CONDITIONS") {
            d = Utils::split(lines.at(++i));
            int gid = int(mesh.face_names.size());
            mesh.face_names.push_back(d.at(0));
            if (gid >= int(mesh.face_groups.size())) mesh.face_groups.resize(gid + 1);
            for (++i; i < n; ++i) {
                d = Utils::split(lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                int cid = to_int(d.at(0)) - 1;
                int lf = to_int(d.at(2)) - 1;
                mesh.faces.at(mesh.cells.at(cid).face_id.at(lf)).group_id = gid;
            }
            continue;
        }
    }
    for (auto &f : mesh.faces) {
        if (f.group_id >= int(mesh.face_groups.size())) mesh.face_groups.resize(f.group_id + 1);
        mesh.face_groups[f.group_id].push_back(f.id);
    }
    mesh.build_geom(mesh_scale);
    return mesh;
}

}  // namespace fvmMesh
}  // namespace MESO
