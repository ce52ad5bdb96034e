// Cell renumbering of a gambit neutral file (meso_mesh_renumber).
//
// meso steps its cells in the order the mesh file numbers them
// (src/mesh/mesh_reader.cpp:52-80) and generates faces in that order
// (src/mesh/object.cpp:38-84), so a file written in a poor order scatters the
// rows of neighbouring cells over memory. This utility writes the same mesh
// with its cells in a locality order, for meso and for this library alike: the
// result is an ordinary gambit file (nodes, their ids and every section but
// the cell numbering unchanged), so the solver run on it is meso's run on a
// renumbered mesh.
//
// Orders: recursive coordinate bisection of the cell centres (split at the
// median of the longest extent, down to single cells), or a seeded random
// permutation (to measure what a poor order costs).
#include "meso_host.h"
#include "meso_hip.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <numeric>
#include <random>
#include <sstream>

namespace {

thread_local std::string g_renumber_error;

struct Neu {
    std::vector<std::string> lines;       // non-empty lines of the file
    int cells_begin = -1, cells_end = -1;  // lines of the ELEMENTS/CELLS section (exclusive)
    std::vector<std::array<double, 3>> node_pos;
    std::vector<int> node_index;          // node id (1-based) -> position in node_pos, -1: none
    struct Cell {
        int type = 0;
        std::vector<int> nodes;  // 1-based node ids
    };
    std::vector<Cell> cells;
};

int to_int(const std::string &s) {
    char *end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str()) throw std::invalid_argument("not an integer: " + s);
    return int(v);
}

Neu parse(const std::string &path) {
    Neu m;
    m.lines = MESO::Utils::read_lines(path);
    const int n = int(m.lines.size());
    for (int i = 0; i < n; ++i) {
        auto d = MESO::Utils::split(m.lines[i]);
        if (d.empty()) continue;
        if (d.size() >= 2 && d[0] == "NODAL" && d[1] == "COORDINATES") {
            for (++i; i < n; ++i) {
                d = MESO::Utils::split(m.lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                const int id = to_int(d[0]);
                if (id < 1) throw std::invalid_argument("gambit node id < 1");
                if (id >= int(m.node_index.size())) m.node_index.resize(size_t(id) + 1, -1);
                m.node_index[id] = int(m.node_pos.size());
                std::array<double, 3> p{{0.0, 0.0, 0.0}};
                for (size_t k = 1; k < d.size() && k < 4; ++k) p[k - 1] = std::stod(d[k]);
                m.node_pos.push_back(p);
            }
            continue;
        }
        if (d[0] == "ELEMENTS/CELLS") {
            m.cells_begin = i + 1;
            for (++i; i < n; ++i) {
                d = MESO::Utils::split(m.lines[i]);
                if (d[0] == "ENDOFSECTION") break;
                Neu::Cell c;
                if (to_int(d[0]) != int(m.cells.size()) + 1)
                    throw std::invalid_argument("gambit cells must be numbered 1..NELEM in order");
                c.type = to_int(d.at(1));
                const size_t count = size_t(to_int(d.at(2)));
                for (size_t j = 3; j < d.size(); ++j) c.nodes.push_back(to_int(d[j]));
                if (c.nodes.size() < count) {  // the list continues on one more line
                    d = MESO::Utils::split(m.lines.at(++i));
                    for (auto &t : d) c.nodes.push_back(to_int(t));
                }
                if (c.nodes.size() != count) throw std::invalid_argument("gambit cell with a wrong node count");
                m.cells.push_back(std::move(c));
            }
            m.cells_end = i;
            continue;
        }
    }
    if (m.cells_begin < 0 || m.cells.empty()) throw std::invalid_argument("no ELEMENTS/CELLS section");
    if (m.cells_end >= n) throw std::invalid_argument("ELEMENTS/CELLS section without ENDOFSECTION");
    return m;
}

// Cell centres (node averages, as meso's geometry takes them).
std::vector<std::array<double, 3>> centres(const Neu &m) {
    std::vector<std::array<double, 3>> c(m.cells.size());
    for (size_t k = 0; k < m.cells.size(); ++k) {
        std::array<double, 3> s{{0.0, 0.0, 0.0}};
        for (int id : m.cells[k].nodes) {
            if (id < 1 || id >= int(m.node_index.size()) || m.node_index[id] < 0)
                throw std::invalid_argument("gambit cell refers to a missing node");
            const auto &p = m.node_pos[size_t(m.node_index[id])];
            for (int a = 0; a < 3; ++a) s[a] += p[a];
        }
        for (int a = 0; a < 3; ++a) s[a] /= double(m.cells[k].nodes.size());
        c[k] = s;
    }
    return c;
}

// Recursive coordinate bisection: [b, e) of idx split at the median along the
// longest extent of its centres, each half ordered the same way.
void rcb(std::vector<int> &idx, int b, int e, const std::vector<std::array<double, 3>> &x) {
    while (e - b > 1) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], x[size_t(idx[i])][a]);
                hi[a] = std::max(hi[a], x[size_t(idx[i])][a]);
            }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        const int mid = b + (e - b) / 2;
        std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](int p, int q) {
            const double xp = x[size_t(p)][ax], xq = x[size_t(q)][ax];
            return xp < xq || (xp == xq && p < q);
        });
        rcb(idx, b, mid, x);
        b = mid;  // the second half, iteratively
    }
}

std::string ids_line(const std::vector<int> &v, size_t b, size_t e) {
    std::string s;
    char buf[16];
    for (size_t i = b; i < e; ++i) {
        std::snprintf(buf, sizeof(buf), "%8d", v[i]);
        s += buf;
    }
    return s;
}

void write(const Neu &m, const std::vector<int> &order, const std::vector<int> &new_of_old, const std::string &path) {
    std::ofstream out(path);
    if (!out.is_open()) throw std::invalid_argument("cannot write " + path);
    const int n = int(m.lines.size());
    char buf[64];
    for (int i = 0; i < n; ++i) {
        if (i == m.cells_begin) {
            // the cells in their new order, numbered 1..NELEM
            for (size_t k = 0; k < order.size(); ++k) {
                const Neu::Cell &c = m.cells[size_t(order[k])];
                std::snprintf(buf, sizeof(buf), "%8d %2d %2d ", int(k) + 1, c.type, int(c.nodes.size()));
                std::string line = buf;
                line += ids_line(c.nodes, 0, c.nodes.size());
                out << line << '\n';
            }
            i = m.cells_end;  // the ENDOFSECTION line is written below
            out << m.lines[size_t(i)] << '\n';
            continue;
        }
        auto d = MESO::Utils::split(m.lines[size_t(i)]);
        if (d.size() >= 2 && d[0] == "ELEMENT" && d[1] == "GROUP") {
            // header, GROUP line, name line and flags line as they are, then
            // the member cells under their new ids
            out << m.lines[size_t(i)] << '\n';
            for (int k = 1; k <= 3 && i + k < n; ++k) out << m.lines[size_t(i + k)] << '\n';
            i += 4;
            std::vector<int> ids;
            for (; i < n; ++i) {
                d = MESO::Utils::split(m.lines[size_t(i)]);
                if (d[0] == "ENDOFSECTION") break;
                for (auto &t : d) {
                    const int old = to_int(t) - 1;
                    if (old < 0 || old >= int(new_of_old.size())) throw std::invalid_argument("group refers to a missing cell");
                    ids.push_back(new_of_old[size_t(old)] + 1);
                }
            }
            std::sort(ids.begin(), ids.end());
            for (size_t b = 0; b < ids.size(); b += 10) out << ids_line(ids, b, std::min(ids.size(), b + 10)) << '\n';
            if (i < n) out << m.lines[size_t(i)] << '\n';
            continue;
        }
        if (d.size() >= 2 && d[0] == "BOUNDARY" && d[1] == "CONDITIONS") {
            out << m.lines[size_t(i)] << '\n';
            if (++i < n) out << m.lines[size_t(i)] << '\n';  // name, type, count
            for (++i; i < n; ++i) {
                d = MESO::Utils::split(m.lines[size_t(i)]);
                if (d[0] == "ENDOFSECTION") {
                    out << m.lines[size_t(i)] << '\n';
                    break;
                }
                const int old = to_int(d.at(0)) - 1;
                if (old < 0 || old >= int(new_of_old.size())) throw std::invalid_argument("boundary refers to a missing cell");
                std::snprintf(buf, sizeof(buf), "%10d %5d %5d", new_of_old[size_t(old)] + 1, to_int(d.at(1)),
                              to_int(d.at(2)));
                out << buf << '\n';
            }
            continue;
        }
        out << m.lines[size_t(i)] << '\n';
    }
    if (!out.good()) throw std::runtime_error("write error: " + path);
}

}  // namespace

extern "C" {

const char *meso_mesh_renumber_error(void) { return g_renumber_error.c_str(); }

int meso_mesh_renumber(const char *in_path, const char *out_path, int32_t method, uint64_t seed,
                       int32_t *new_of_old, int32_t capacity, int32_t *n_cells) {
    if (!in_path || !out_path) {
        g_renumber_error = "null path";
        return MESO_E_INVALID;
    }
    if (method != MESO_RENUMBER_RCB && method != MESO_RENUMBER_RANDOM && method != MESO_RENUMBER_KEEP) {
        g_renumber_error = "unknown renumbering method";
        return MESO_E_INVALID;
    }
    try {
        Neu m = parse(in_path);
        const int nc = int(m.cells.size());
        if (n_cells) *n_cells = nc;
        if (new_of_old && capacity < nc) {
            g_renumber_error = "new_of_old holds fewer entries than the mesh has cells";
            return MESO_E_INVALID;
        }
        std::vector<int> order(static_cast<size_t>(nc));
        std::iota(order.begin(), order.end(), 0);
        if (method == MESO_RENUMBER_RCB) {
            rcb(order, 0, nc, centres(m));
        } else if (method == MESO_RENUMBER_RANDOM) {
            std::mt19937_64 rng(seed);
            std::shuffle(order.begin(), order.end(), rng);
        }
        std::vector<int> inv(static_cast<size_t>(nc));
        for (int k = 0; k < nc; ++k) inv[size_t(order[size_t(k)])] = k;
        write(m, order, inv, out_path);
        if (new_of_old) std::copy(inv.begin(), inv.end(), new_of_old);
    } catch (const std::invalid_argument &e) {
        g_renumber_error = e.what();
        return MESO_E_INVALID;
    } catch (const std::exception &e) {
        g_renumber_error = e.what();
        return MESO_E_IO;
    }
    g_renumber_error.clear();
    return MESO_OK;
}

}  // extern "C"
