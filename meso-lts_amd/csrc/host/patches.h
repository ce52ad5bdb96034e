// Spatial patches of a 2D or 3D mesh for the LDS-staged step kernels.
//
// The cells (or faces) of the mesh are split by recursive coordinate
// bisection into compact patches. A patch kernel stages the distribution
// values of a patch's "items" (the cells whose value and least-squares
// gradient it needs) in LDS once per velocity tile, forms every gradient
// and every face value once, and reuses them from LDS. Two patch sets exist:
//   update patches: owned cells = the cells the patch updates (all but those
//     whose boundary needs other velocities, i.e. mirror boundaries); items =
//     owned cells, then the cells across their faces; faces = all faces of
//     the owned cells.
//   moment patches: faces = a compact group of faces (boundary ones included)
//     whose f-bar+ moments the patch sums; items = owners and neighbours of
//     those faces.
// The stage list of a patch is its items followed by the other LSQ
// neighbours of its items; all of it is staged in LDS, so the gradient and
// face work reads LDS only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace meso_patch {

// Face and LSQ-neighbour slots of a cell: 2D triangles and quadrilaterals
// have at most 4, 3D hexahedra 6 (tetrahedra, prisms and pyramids fewer).
template <int D>
struct Lim {
    static constexpr int MAXF = D == 2 ? 4 : 6;  // faces of a cell
    static constexpr int MAXN = MAXF;            // LSQ neighbours (face neighbours)
};

// Face modes of an item's face entry.
enum FaceMode : int { SIDE_OWNER = 0, SIDE_NEIGHBOUR = 1, SIDE_BOUNDARY = 2 };

// Device record of one item (2D: 128 bytes, 3D: 208 bytes). Items are the
// first rows of the patch's stage list; the remaining rows are cells read
// only as neighbours.
template <int D>
struct ItemT;
template <>
struct alignas(16) ItemT<2> {
    int cell;         // global cell id
    int nnb;          // LSQ neighbours (<= MAXN); padding entries repeat the item with zero weights
    int nface;        // face entries (<= MAXF)
    int pad0;
    int nb[4];        // stage row of each LSQ neighbour
    int face[4];      // local face index | (mode << 16)
    int pad1[4];
    double a[4][2];   // gradient weights: grad f = sum_j a_j (f_j - f_item)
};
template <>
struct alignas(16) ItemT<3> {
    int cell, nnb, nface, pad0;
    int nb[6];
    int face[6];
    double a[6][3];
};
static_assert(sizeof(ItemT<2>) == 128 && sizeof(ItemT<3>) == 208, "Item layout");

// Device record of one patch face (2D: 80 bytes, 3D: 96 bytes).
template <int D>
struct FaceT;
template <>
struct alignas(16) FaceT<2> {
    int face;         // global face id
    int own, nei;     // item indices of owner and neighbour (equal on a boundary face)
    int pad0;
    double n[2];      // unit normal out of the owner (x, y)
    double area;
    double pad1;
    double d[2][2];   // x_f - x_owner, x_f - x_neighbour
};
template <>
struct alignas(16) FaceT<3> {
    int face, own, nei, pad0;
    double n[3];
    double area;
    double d[2][3];
};
static_assert(sizeof(FaceT<2>) == 80 && sizeof(FaceT<3>) == 96, "Face layout");

template <int D>
struct PatchSet {
    int n_patch = 0;
    int max_items = 0, max_faces = 0, max_owned = 0, max_rows = 0;
    std::vector<int> item_ptr;   // [n_patch + 1]
    std::vector<int> face_ptr;   // [n_patch + 1]
    std::vector<int> row_ptr;    // [n_patch + 1] stage rows: the items, then neighbour-only cells
    std::vector<int> n_owned;    // [n_patch] (update patches: owned items come first)
    std::vector<ItemT<D>> items;
    std::vector<FaceT<D>> faces;
    std::vector<int> rows;       // global cell id of each stage row
};

struct MeshView {
    int nc = 0, nf = 0;
    const int *face_cells = nullptr;  // [nf * 2] owner, neighbour (equal on a boundary face)
    const double *cell_pos = nullptr; // [nc * 3]
    const double *face_pos = nullptr; // [nf * 3]
    const double *face_area = nullptr;
    const double *face_normal = nullptr;  // [nf * 3]
    const int *cf_ptr = nullptr, *cf_ids = nullptr;  // cell -> faces, in meso's order
    const int *nb_ptr = nullptr, *nb_ids = nullptr;  // LSQ neighbours
    const double *nb_a = nullptr;                    // [nnz * 3] gradient weights
};

// The cells flagged in `own` grouped into update patches of at most
// `max_owned` cells (their boundary faces become SIDE_BOUNDARY entries).
// Returns false with a message if a cell has more than MAXF faces or MAXN
// LSQ neighbours.
template <int D>
bool build_update_patches(const MeshView &m, const std::vector<char> &own, int max_owned, PatchSet<D> &out,
                          std::string &err);

// All faces grouped into moment patches of at most `max_faces` faces.
template <int D>
bool build_moment_patches(const MeshView &m, int max_faces, PatchSet<D> &out, std::string &err);

// Recursive coordinate bisection of points (the first `dims` coordinates of
// pos[id * stride]) into groups of at most `cap` points; returns the groups
// in depth-first (spatially coherent) order.
std::vector<std::vector<int>> rcb(const std::vector<int> &ids, const double *pos, int stride, int cap, int dims = 2);

}  // namespace meso_patch
