// Spatial patches of a 2D or 3D mesh for the LDS-staged step kernels (see patches.h).
#include "patches.h"

#include <algorithm>
#include <cmath>
#include <unordered_map>

namespace meso_patch {

namespace {

void rcb_rec(std::vector<int> &ids, size_t lo, size_t hi, const double *pos, int stride, int cap, int dims,
             std::vector<std::vector<int>> &out) {
    const size_t n = hi - lo;
    if (n <= size_t(cap)) {
        out.emplace_back(ids.begin() + lo, ids.begin() + hi);
        return;
    }
    const size_t k = (n + cap - 1) / cap;  // groups this range becomes
    const size_t k1 = k / 2;
    const size_t n1 = (n * k1 + k / 2) / k;  // cells of the first k1 groups, proportional
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (size_t i = lo; i < hi; ++i)
        for (int a = 0; a < dims; ++a) {
            const double x = pos[size_t(ids[i]) * stride + a];
            mn[a] = std::min(mn[a], x);
            mx[a] = std::max(mx[a], x);
        }
    int ax = 0;
    for (int a = 1; a < dims; ++a)
        if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
    auto key = [&](int id) { return pos[size_t(id) * stride + ax]; };
    std::nth_element(ids.begin() + lo, ids.begin() + lo + n1, ids.begin() + hi, [&](int a, int b) {
        const double ka = key(a), kb = key(b);
        return ka < kb || (ka == kb && a < b);
    });
    rcb_rec(ids, lo, lo + n1, pos, stride, cap, dims, out);
    rcb_rec(ids, lo + n1, hi, pos, stride, cap, dims, out);
}

template <int D>
struct Builder {
    using Item = ItemT<D>;
    using Face = FaceT<D>;
    static constexpr int MAXN = Lim<D>::MAXN;
    const MeshView &m;
    PatchSet<D> &ps;
    std::unordered_map<int, int> loc;  // global cell -> item index in the current patch
    std::vector<Item> items;
    std::vector<Face> faces;
    std::unordered_map<int, int> floc;  // global face -> local face index

    Builder(const MeshView &m_, PatchSet<D> &p) : m(m_), ps(p) {}

    int add_item(int c) {
        auto it = loc.find(c);
        if (it != loc.end()) return it->second;
        Item r{};
        r.cell = c;
        const int idx = int(items.size());
        items.push_back(r);
        loc.emplace(c, idx);
        return idx;
    }

    int add_face(int f) {
        auto it = floc.find(f);
        if (it != floc.end()) return it->second;
        Face r{};
        r.face = f;
        const int idx = int(faces.size());
        faces.push_back(r);
        floc.emplace(f, idx);
        return idx;
    }

    void add_entry(Item &it, int f) {
        const int o = m.face_cells[2 * size_t(f)], n = m.face_cells[2 * size_t(f) + 1];
        const int mode = (o == n) ? SIDE_BOUNDARY : (o == it.cell ? SIDE_OWNER : SIDE_NEIGHBOUR);
        it.face[it.nface++] = floc.at(f) | (mode << 16);
    }

    // Stencils, face records; appends the patch to ps.
    void finish(int n_owned) {
        std::vector<int> rows;
        for (auto &it : items) rows.push_back(it.cell);
        std::unordered_map<int, int> rloc(loc);
        for (auto &it : items) {
            const int c = it.cell;
            const int j0 = m.nb_ptr[c], nj = m.nb_ptr[c + 1] - j0;
            it.nnb = nj;
            const int self = loc.at(c);
            for (int j = 0; j < MAXN; ++j) {
                if (j < nj) {
                    const int g = m.nb_ids[j0 + j];
                    auto f = rloc.find(g);
                    if (f == rloc.end()) {
                        f = rloc.emplace(g, int(rows.size())).first;
                        rows.push_back(g);
                    }
                    it.nb[j] = f->second;
                    for (int a = 0; a < D; ++a) it.a[j][a] = m.nb_a[3 * size_t(j0 + j) + a];
                } else {
                    it.nb[j] = self;
                    for (int a = 0; a < D; ++a) it.a[j][a] = 0.0;
                }
            }
        }
        for (auto &fr : faces) {
            const int f = fr.face;
            const int o = m.face_cells[2 * size_t(f)], n = m.face_cells[2 * size_t(f) + 1];
            fr.own = loc.at(o);
            fr.nei = loc.at(n);
            fr.area = m.face_area[f];
            for (int a = 0; a < D; ++a) fr.n[a] = m.face_normal[3 * size_t(f) + a];
            for (int a = 0; a < D; ++a) {
                fr.d[0][a] = m.face_pos[3 * size_t(f) + a] - m.cell_pos[3 * size_t(o) + a];
                fr.d[1][a] = m.face_pos[3 * size_t(f) + a] - m.cell_pos[3 * size_t(n) + a];
            }
        }
        ps.items.insert(ps.items.end(), items.begin(), items.end());
        ps.faces.insert(ps.faces.end(), faces.begin(), faces.end());
        ps.rows.insert(ps.rows.end(), rows.begin(), rows.end());
        ps.item_ptr.push_back(int(ps.items.size()));
        ps.face_ptr.push_back(int(ps.faces.size()));
        ps.row_ptr.push_back(int(ps.rows.size()));
        ps.max_rows = std::max(ps.max_rows, int(rows.size()));
        ps.n_owned.push_back(n_owned);
        ps.max_items = std::max(ps.max_items, int(items.size()));
        ps.max_faces = std::max(ps.max_faces, int(faces.size()));
        ps.max_owned = std::max(ps.max_owned, n_owned);
        ++ps.n_patch;
        items.clear();
        faces.clear();
        loc.clear();
        floc.clear();
    }
};

template <int D>
bool check_limits(const MeshView &m, std::string &err) {
    for (int c = 0; c < m.nc; ++c) {
        if (m.cf_ptr[c + 1] - m.cf_ptr[c] > Lim<D>::MAXF || m.nb_ptr[c + 1] - m.nb_ptr[c] > Lim<D>::MAXN) {
            err = "patch kernels: a cell has more faces or LSQ neighbours than the patch records hold";
            return false;
        }
    }
    return true;
}

template <int D>
void reset(PatchSet<D> &ps) {
    ps = PatchSet<D>();
    ps.item_ptr.push_back(0);
    ps.face_ptr.push_back(0);
    ps.row_ptr.push_back(0);
}

}  // namespace

std::vector<std::vector<int>> rcb(const std::vector<int> &ids, const double *pos, int stride, int cap, int dims) {
    std::vector<std::vector<int>> out;
    if (ids.empty()) return out;
    std::vector<int> work(ids);
    rcb_rec(work, 0, work.size(), pos, stride, std::max(cap, 1), dims, out);
    return out;
}

template <int D>
bool build_update_patches(const MeshView &m, const std::vector<char> &own, int max_owned, PatchSet<D> &ps,
                          std::string &err) {
    reset(ps);
    if (!check_limits<D>(m, err)) return false;
    std::vector<int> owned;
    for (int c = 0; c < m.nc; ++c)
        if (own[c]) owned.push_back(c);
    Builder<D> b(m, ps);
    for (const auto &grp : rcb(owned, m.cell_pos, 3, max_owned, D)) {
        for (int c : grp) b.add_item(c);
        // faces of the owned cells in meso's cell-face order, and the cells across them
        for (int c : grp)
            for (int k = m.cf_ptr[c]; k < m.cf_ptr[c + 1]; ++k) {
                const int f = m.cf_ids[k];
                b.add_face(f);
                const int o = m.face_cells[2 * size_t(f)], n = m.face_cells[2 * size_t(f) + 1];
                b.add_item(o == c ? n : o);
            }
        for (size_t i = 0; i < b.items.size(); ++i) {
            auto &it = b.items[i];
            const int c = it.cell;
            for (int k = m.cf_ptr[c]; k < m.cf_ptr[c + 1]; ++k) {
                const int f = m.cf_ids[k];
                if (b.floc.count(f)) b.add_entry(it, f);
            }
        }
        b.finish(int(grp.size()));
    }
    return true;
}

template <int D>
bool build_moment_patches(const MeshView &m, int max_faces, PatchSet<D> &ps, std::string &err) {
    reset(ps);
    if (!check_limits<D>(m, err)) return false;
    std::vector<int> all(m.nf);
    for (int f = 0; f < m.nf; ++f) all[f] = f;
    Builder<D> b(m, ps);
    for (const auto &grp : rcb(all, m.face_pos, 3, max_faces, D)) {
        for (int f : grp) {
            b.add_face(f);
            b.add_item(m.face_cells[2 * size_t(f)]);
            b.add_item(m.face_cells[2 * size_t(f) + 1]);
        }
        for (size_t i = 0; i < b.items.size(); ++i) {
            auto &it = b.items[i];
            const int c = it.cell;
            for (int k = m.cf_ptr[c]; k < m.cf_ptr[c + 1]; ++k) {
                const int f = m.cf_ids[k];
                if (b.floc.count(f)) b.add_entry(it, f);
            }
        }
        b.finish(0);
    }
    return true;
}

template bool build_update_patches<2>(const MeshView &, const std::vector<char> &, int, PatchSet<2> &, std::string &);
template bool build_update_patches<3>(const MeshView &, const std::vector<char> &, int, PatchSet<3> &, std::string &);
template bool build_moment_patches<2>(const MeshView &, int, PatchSet<2> &, std::string &);
template bool build_moment_patches<3>(const MeshView &, int, PatchSet<3> &, std::string &);

}  // namespace meso_patch
