// Spatial patches of a 2D mesh for the LDS-staged step kernels.
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

constexpr int MAXN = 4;  // LSQ neighbours of a 2D cell (face neighbours)
constexpr int MAXF = 4;  // faces of a 2D cell (triangle, quadrilateral)

// Face modes of an item's face entry.
enum FaceMode : int { SIDE_OWNER = 0, SIDE_NEIGHBOUR = 1, SIDE_BOUNDARY = 2 };

// Device record of one item (128 bytes). Items are the first rows of the
// patch's stage list; the remaining rows are cells read only as neighbours.
struct alignas(16) Item {
    int cell;         // global cell id
    int nnb;          // LSQ neighbours (<= MAXN); padding entries repeat the item with zero weights
    int nface;        // face entries (<= MAXF)
    int pad0;
    int nb[MAXN];     // stage row of each LSQ neighbour
    int face[MAXF];   // local face index | (mode << 16)
    int pad1[4];
    double a[MAXN][2];   // gradient weights: grad f = sum_j a_j (f_j - f_item)
};
static_assert(sizeof(Item) == 128, "Item layout");

// Device record of one patch face (80 bytes).
struct alignas(16) Face {
    int face;         // global face id
    int own, nei;     // item indices of owner and neighbour (equal on a boundary face)
    int pad0;
    double n[2];      // unit normal out of the owner (x, y)
    double area;
    double pad1;
    double d[2][2];   // x_f - x_owner, x_f - x_neighbour (x, y)
};
static_assert(sizeof(Face) == 80, "Face layout");

struct PatchSet {
    int n_patch = 0;
    int max_items = 0, max_faces = 0, max_owned = 0, max_rows = 0;
    std::vector<int> item_ptr;   // [n_patch + 1]
    std::vector<int> face_ptr;   // [n_patch + 1]
    std::vector<int> row_ptr;    // [n_patch + 1] stage rows: the items, then neighbour-only cells
    std::vector<int> n_owned;    // [n_patch] (update patches: owned items come first)
    std::vector<Item> items;
    std::vector<Face> faces;
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
// Returns false with a message if the mesh does not fit the 2D limits (more
// than MAXF faces or MAXN neighbours per cell).
bool build_update_patches(const MeshView &m, const std::vector<char> &own, int max_owned, PatchSet &out,
                          std::string &err);

// All faces grouped into moment patches of at most `max_faces` faces.
bool build_moment_patches(const MeshView &m, int max_faces, PatchSet &out, std::string &err);

// Recursive coordinate bisection of points (x, y) into groups of at most
// `cap` points; returns the groups in depth-first (spatially coherent) order.
std::vector<std::vector<int>> rcb(const std::vector<int> &ids, const double *pos, int stride, int cap);

}  // namespace meso_patch
