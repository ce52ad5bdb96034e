// gfx950 kernels of one DUGKS/CDUGKS time step.
//
// One step is five launches (six for classic DUGKS), all on the context's
// stream:
//   k_face_moments  (face, v-tile) segments: half-step reconstruction f-bar+ of
//                   every face from its upwind cell and that cell's LSQ
//                   gradient, and the raw velocity moments of f-bar+ summed in
//                   FP64 per tile  (cdugks_shakhov.cpp:213-316)
//   k_face_eq       per face: moments -> rho, u, T, q, tau, C_s; equilibrium
//                   constants; FP64 flux moments of the mixed face value along
//                   n0 (cdugks_shakhov.cpp:317-362 and the face part of :616-641)
//   k_boundary      per boundary face (one block, all v): wall density sums,
//                   inlet/outlet/wall/slip/symmetry values, exact flux moments
//                   (cdugks_shakhov.cpp:364-607)
//   k_cell_update   (cell, v-tile) segments: conservative macro update from
//                   the face flux moments, face values recomputed from the
//                   2-ring of f^n, cell flux, f-tilde-plus update, q partials
//                   (cdugks_shakhov.cpp:610-722)
//   k_cell_q        per cell: q^{n+1} from the partials (cdugks_shakhov.cpp:721)
//
// Nothing of size NF x NV is ever stored: face values are recomputed from cell
// values wherever they are needed, so a step streams f^n twice and writes
// f^{n+1} once.
#pragma once

#include "patches.h"
#include "step_common.h"

namespace meso_dev {

// A pointer into device memory as a kernel argument member: it holds the
// global-address-space pointer and converts to an ordinary pointer where it
// is used, so accesses through it compile to global instructions whether the
// arguments come by value or from a table in memory (where a plain pointer
// member would give flat instructions, which also wait on the LDS counter).
template <class T>
struct gptr {
    __attribute__((address_space(1))) T *p;
    __host__ __device__ operator T *() const { return (T *)p; }
    template <class U>
    __host__ __device__ explicit operator U *() const {
        return (U *)(T *)p;
    }
    template <class U>
    __host__ __device__ U *as() const {
        return (U *)(T *)p;
    }
    __host__ __device__ gptr &operator=(T *q) {
        p = (__attribute__((address_space(1))) T *)q;
        return *this;
    }
};

// Everything a kernel needs; passed by value, or read from a device table by
// the fused batch kernels.
struct KArgs {
    StepConst s;
    // mesh
    gptr<const int2> face_cells; // [nf] owner, neighbour
    gptr<const double> face_n;  // [nf*3] unit normal out of the owner
    gptr<const double> face_d;  // [nf*6] x_f - x_owner, x_f - x_neighbour
    gptr<const float> face_df;  // float copy of face_d
    gptr<const int> cf_ptr;     // [nc+1] cell -> faces
    gptr<const int> cf_ids;     // faces
    gptr<const double> cf_sA;   // +-area (+ when the cell owns the face)
    gptr<const int> nb_ptr;     // [nc+1] LSQ neighbours
    gptr<const int> nb_ids;
    gptr<const double> nb_a;    // [nnz*3] w_j (Cx.dr_j, Cy.dr_j, Cz.dr_j)
    gptr<const double> nb_w;    // [nnz] w_j = 1/|dr_j|^2 (inverse-distance weights of the mask repair)
    gptr<const float> nb_af;
    gptr<const double> cell_vol;
    gptr<const double> cell_S;  // [nc*3] sum over the cell's faces of +-A n (0 for a closed, planar-faced cell)
    gptr<const int> face_bnd;   // [nf] boundary index or -1
    // per-cell stencil tables, padded to MAXF faces / MAXN neighbours
    // (2D: 4, 3D: 6), in the cell's local face order
    gptr<const int> cs_nf;      // [nc] faces of the cell
    gptr<const int> cs_face;    // [nc*MAXF] face id
    gptr<const int> cs_other;   // [nc*MAXF] cell across the face, -1 on a boundary
    gptr<const int> cs_own;     // [nc*MAXF] 1 when the cell owns the face
    gptr<const int> cs_bnd;     // [nc*MAXF] boundary index or -1
    gptr<const double> cs_sA;   // [nc*MAXF] +-area
    gptr<const double> cs_d;    // [nc*MAXF*6] x_f - x_c, x_f - x_other
    gptr<const float> cs_df;
    gptr<const int> lsq_cnt;    // [nc] LSQ neighbours
    gptr<const int> lsq_nb;     // [nc*MAXN]
    gptr<const double> lsq_a;   // [nc*MAXN*3] w_j (Cx.dr_j, Cy.dr_j, Cz.dr_j)
    gptr<const float> lsq_af;
    gptr<const float> face_nf;  // [nf*3] float copy of the normals
    gptr<EqF> face_eqf;         // [nf] FP32 copy of face_eq
    // velocity slots
    gptr<const double4> vel_d;  // [nv] x, y, z, w
    gptr<const float4> vel_f;   // [nv] x, y, z, |xi|^2
    gptr<const double2> fref_d2; // [nv] g_ref, h_ref
    gptr<const float2> fref_f2;
    // axis tables of tensor-product velocity sets (nt == 0: none)
    int nt;
    gptr<const int> vtab;       // [nv] packed offsets of a velocity's axis entries
    gptr<const double> tab_x;   // [nt] abscissa of each entry
    gptr<const int> tab_ax;     // [nt] axis of each entry
    gptr<const double> tab_w;   // [nt] 1D weight of each entry (w(xi) = prod of its axis weights), or null
    gptr<void> face_tab;        // [nf*nt] per-face axis factors (storage type)
    // gather stencils: group = centre cell + its LSQ neighbours (padded with
    // the centre and zero weights); cells: groups of the cell and of the cell
    // across each face; faces: groups of owner and neighbour
    gptr<const int> cell_list;  // interior cells, then cells with a boundary face
    gptr<const int> st_id;      // [nc][(1+MAXF)*(1+MAXN)]
    gptr<const double> st_a;    // [nc][(1+MAXF)*MAXN*3]
    gptr<const float> st_af;
    gptr<const int> fs_id;      // [nf][2*(1+MAXN)]
    gptr<const double> fs_a;    // [nf][2*MAXN*3]
    gptr<const float> fs_af;
    gptr<const int> bnd_face;   // [nb]
    gptr<const int> bnd_mark;   // [nb] index into marks
    gptr<const void> marks;     // meso_mark_desc[] (host layout)
    // dvs
    gptr<const double> xi;      // [nv*3]
    gptr<const float> xif;
    gptr<const double> w;       // [nv]
    gptr<const int> mirror;     // [nv*3] or null
    // state
    gptr<const void> fin_g, fin_h; // cell-major, f^n
    gptr<void> fout_g, fout_h;  // cell-major, f^{n+1}
    // FP32X storage: f^n / f^{n+1} hold FP32 words of the deviation from f_ref
    // and these arrays the extension byte of every word (same indexing); null
    // for FP32 and FP64 storage
    gptr<const unsigned char> xin_g, xin_h;
    gptr<unsigned char> xout_g, xout_h;
    gptr<const double> rho_in, u_in, T_in, q_in;
    gptr<double> rho_out, u_out, T_out, q_out;
    // scratch
    gptr<double> part_f;        // [nf][ntile_f][ncol]
    gptr<EqParams> face_eq;     // [nf]
    gptr<double> face_flux;     // [nf][D+2] flux moments along n0 of the final face value
    gptr<BndParams> bnd;        // [nb]
    gptr<double> part_c;        // [nc][ntile_c][4]
    gptr<double> fm;            // [nc][5] cell flux moments: m0, m1 (3), m2
    gptr<double> sflux;         // small velocity sets: [nd][nf][LP] (xi . n0) f_face per face and velocity
    // control
    gptr<int> bad_count;
    gptr<int> first_bad;
    gptr<int> bad_step;
    gptr<int> halt;
    gptr<int> step_ctr;         // steps since the initial state (device counter)
    gptr<const int> mask;       // [nc] 1: a bad cell is repaired (Shakhov); null: no masked group
    gptr<int> masked;           // bad cells repaired so far (device counter)
    // LDS-staged patch kernels (2D): update patches (interior cells) and
    // moment patches (all faces); the velocity range is split into p_chunks
    // chunks of p_vchunk velocities
    int n_up, up_maxi, up_maxf, up_maxo, up_maxr;
    gptr<const void> up_items;    // meso_patch::ItemT<D>
    gptr<const void> up_faces;    // meso_patch::FaceT<D>
    gptr<const int> up_item_ptr, up_face_ptr, up_nown, up_row_ptr, up_rows;
    int n_mp, mp_maxi, mp_maxf, mp_maxr;
    gptr<const void> mp_items;
    gptr<const void> mp_faces;
    gptr<const int> mp_item_ptr, mp_face_ptr, mp_row_ptr, mp_rows;
    gptr<const int> mp_nown;    // cell-based moment patches: owned cells per patch
    int p_chunks, p_vchunk;     // velocity chunks of the patch kernels and their size (a multiple of 64)
    int n_int_cells;            // interior cells: cell_list[0, n_int_cells)
    gptr<void> crec;            // [nc] CellRec of the storage type
    // patch kernels: f-bar+ of every face, written by the moment kernel and
    // read by the update kernel (face-major [nf][nvp], storage type)
    gptr<void> fbar_g, fbar_h;
    gptr<void> cell_tab;        // [nc][2][nt] axis tables of the cell equilibria
    // FP32 storage keeps f - f_ref; f_ref is the equilibrium `ref` at rest.
    EqParams ref;
    gptr<const double> fref_d; // [2][nv] g_ref then h_ref (FP64)
    gptr<const float> fref_f; // same, FP32
    double mref[17];       // raw face moments of f_ref (face-moment column layout)
    double rflux[5][3];    // per flux column k: sum_v w xi psi_k f_ref
};

// A context's arguments from a device table (the fused batch kernels).
__device__ __forceinline__ KArgs batch_args(const KArgs *__restrict__ tab, int k) { return tab[k]; }

// Equilibrium in the stored form: the full value for FP64 storage, the
// deviation from f_ref for FP32 storage.
template <int MODEL, int D, class T>
__device__ __forceinline__ void eq_store(const KArgs &a, const EqParams &e, const T x[3], int v, T &g, T &h) {
    if constexpr (sizeof(T) == 8) {
        if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<T>(e, x, D, a.s.K, g, h);
        else {
            g = bgk_eq<T>(e, x);
            h = T(0);
        }
    } else {
        if (MODEL == CDUGKS_SHAKHOV) shakhov_dev(e, x, D, a.s.K, a.fref_f[v], g, h);
        else {
            g = bgk_dev(e, x);
            h = 0.0f;
        }
    }
}

// Full FP64 value of a stored value of distribution k (0: g, 1: h) at velocity v.
template <class T>
__device__ __forceinline__ double full_value(const KArgs &a, T stored, int k, int v) {
    if constexpr (sizeof(T) == 8) return stored;
    else return double(stored) + a.fref_d[k * a.s.nv + v];
}

// (1 - m) x + m y, evaluated as x + m (y - x) for FP32 storage so that a
// weight near 0 or 1 does not bias the result by its own rounding.
template <class T>
__device__ __forceinline__ T lerp_store(T x, T y, T m) {
    if constexpr (sizeof(T) == 8) return (T(1) - m) * x + m * y;
    else return x + m * (y - x);
}

template <int MODEL, int D>
__host__ __device__ constexpr int ncol_face() {
    // g: M0, M1 (D), M2 (D(D+1)/2), [M3 (D), H0, H1 (D) for Shakhov]
    return MODEL == CDUGKS_SHAKHOV ? 1 + D + D * (D + 1) / 2 + D + 1 + D : 1 + D + D * (D + 1) / 2;
}

// LSQ gradient of distribution f at cell c, velocity v (mesh_numerical.cpp:24-49),
// as sum_j a_j (f_j - f_c).
template <int D, class T>
__device__ __forceinline__ void lsq_grad(const KArgs &a, const T *__restrict__ f, int c, int v, T fc, T gr[3]) {
    gr[0] = gr[1] = gr[2] = T(0);
    const int j0 = a.nb_ptr[c], j1 = a.nb_ptr[c + 1];
    const int nvp = a.s.nvp;
    for (int j = j0; j < j1; ++j) {
        T df = f[size_t(a.nb_ids[j]) * nvp + v] - fc;
        if constexpr (sizeof(T) == 8) {
            gr[0] += T(a.nb_a[3 * j]) * df;
            gr[1] += T(a.nb_a[3 * j + 1]) * df;
            if (D == 3) gr[2] += T(a.nb_a[3 * j + 2]) * df;
        } else {
            gr[0] += a.nb_af[3 * j] * df;
            gr[1] += a.nb_af[3 * j + 1] * df;
            if (D == 3) gr[2] += a.nb_af[3 * j + 2] * df;
        }
    }
}

template <class T>
__device__ __forceinline__ T face_d(const KArgs &a, int f, int side, int k) {
    if constexpr (sizeof(T) == 8) return a.face_d[6 * f + 3 * side + k];
    else return a.face_df[6 * f + 3 * side + k];
}

template <class T>
__device__ __forceinline__ void load_xi(const KArgs &a, int v, T x[3]) {
    if constexpr (sizeof(T) == 8) {
        x[0] = a.xi[3 * v];
        x[1] = a.xi[3 * v + 1];
        x[2] = a.xi[3 * v + 2];
    } else {
        x[0] = a.xif[3 * v];
        x[1] = a.xif[3 * v + 1];
        x[2] = a.xif[3 * v + 2];
    }
}

// n0 . xi in FP64 with meso's evaluation order (x*x' + y*y' + z*z'); its sign
// against +-VSMALL picks the upwind side.
__device__ __forceinline__ double ndot(const KArgs &a, int f, int v) {
    return a.face_n[3 * f] * a.xi[3 * v] + a.face_n[3 * f + 1] * a.xi[3 * v + 1] +
           a.face_n[3 * f + 2] * a.xi[3 * v + 2];
}

// (d - xi h) . grad
template <int D, class T>
__device__ __forceinline__ T extrap(const KArgs &a, int f, int side, const T xh[3], const T gr[3]) {
    T r = (face_d<T>(a, f, side, 0) - xh[0]) * gr[0] + (face_d<T>(a, f, side, 1) - xh[1]) * gr[1];
    if (D == 3) r += (face_d<T>(a, f, side, 2) - xh[2]) * gr[2];
    return r;
}

// Limiter factor of cell c (Shakhov limiter-switch) for the increment dw at
// velocity v: the range of the values fd over c and its face neighbours.
template <class T>
__device__ __forceinline__ double nb_phi(const KArgs &a, const T *__restrict__ fd, int c, int v, double fc, double dw,
                                         double eps) {
    double mn = fc, mx = fc;
    const int nvp = a.s.nvp;
    for (int j = a.nb_ptr[c]; j < a.nb_ptr[c + 1]; ++j) {
        const double w = double(fd[size_t(a.nb_ids[j]) * nvp + v]);
        mn = fmin(mn, w);
        mx = fmax(mx, w);
    }
    return venkata_phi(dw, fc, mn, mx, venkata_omega(a.s.venkata_k, a.cell_vol[c], a.s.D), eps);
}

}  // namespace meso_dev
