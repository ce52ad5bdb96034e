// LDS-staged patch kernels of the step (included by step.hip inside its
// anonymous namespace; they use its helpers).
//
// A block owns one spatial patch (csrc/host/patches.h) and one chunk of the
// velocity range, and walks the chunk in tiles of VT velocities. Per tile the
// lanes hold VT consecutive velocities of 64/VT items at once:
//   A  stage the FP32/FP64 values of the patch's items in LDS
//   B  per item: its LSQ gradient (neighbours from LDS, or from HBM when they
//      are not items), then the half-step value of each of its faces on the
//      velocities for which the item is upwind; every gradient and every face
//      value is formed exactly once
//   C  moments:  raw velocity moments of f-bar+ per face (lane = face)
//      update:   cell flux from the face rows, f-tilde-plus update, q partials
// All arithmetic of the transport part (gradients, reconstruction, fluxes,
// the update itself) is FP64; FP32 storage only rounds the stored deviation
// f - f_ref and the handed-over face values, both stochastically (unbiased:
// near a steady state the per-step increments fall below half a unit in the
// last place, and round-to-nearest drops them and stalls the state). The
// equilibria of FP32 storage are evaluated in FP32 in deviation form (their
// weight in one step is dt/tau or C_s).
//
// LIM: the Venkatakrishnan limiter of the reconstruction (Shakhov
// limiter-switch); a template argument, so that the kernels without it keep
// their registers.
//
// Upwind dead-band (|xi.n| < 1e-30): the moment kernel forms meso's average of
// both sides; the update kernel gives the face no flux, since the flux term is
// (xi.n) A f with |xi.n| < 1e-30.
#pragma once

constexpr int BLOCKP = 512;  // threads of a patch block
// Update patches own at most UPD_CELLS cells; a block has (BLOCKP / 64) *
// (64 / VT) thread slots, each owning upd_qj(VT) of them (and keeping their
// heat-flux partials in registers over all tiles).
constexpr int UPD_CELLS = 32;
// Velocities per tile of the update kernel: as wide as the moment tiles (64
// FP32, 32 FP64) in 2D; in 3D the face-flux rows (6 faces per cell) would
// shrink the patches to ~18 cells at 64, so FP32 takes 32 there (32^3 hexes
// x 16^3 velocities: update 6.7 -> 4.0 ms).
template <class S, int D>
__host__ __device__ constexpr int upd_vt() {
    return sizeof(S) == 4 ? (D == 2 ? 64 : 32) : 32;
}
__host__ __device__ constexpr int upd_qj(int vt) {
    return (UPD_CELLS + (BLOCKP / 64) * (64 / vt) - 1) / ((BLOCKP / 64) * (64 / vt));
}

// Per-cell constants of the update (written by k_cell_prep for interior cells).
template <class S, int D>
struct CellRec;
template <int D>
struct CellRec<float, D> {
    EqF en, e;                        // equilibria at t_n and t_{n+1}, deviation form
    double cm, C, ai, an, Cs, dtv;    // see k_patch_update
    double u[D];                      // u^{n+1}
};
template <int D>
struct CellRec<double, D> {
    EqParams en, e;
    double cm, C, itau, itaun, Cs, dtv;
    double u[D];
};

// ------------------------------------------------------------------- prep
// Per cell of the update patches: conservative macro update (as k_cell_update
// does for its cells), the constants of the f update, the cell's axis tables.
// bid: the block's index in its context's own grid.
template <int MODEL, class S, bool TAB, int D>
__device__ __forceinline__ void d_cell_prep(const KArgs &a, int bid) {
    if (*a.halt) return;
    constexpr int WPB = BLOCK / 64;
    const int lane = threadIdx.x % 64;
    const int li = __builtin_amdgcn_readfirstlane(bid * WPB + int(threadIdx.x / 64));
    if (li >= a.n_int_cells) return;
    const int c = a.cell_list[li];
    CellStep<MODEL, D> cs;
    const bool ok = cell_macros<MODEL, D, S>(a, c, cs);
    if (lane == 0) {
        a.rho_out[c] = cs.rho;
        for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = cs.u[i];
        if (MODEL == CDUGKS_SHAKHOV) a.T_out[c] = cs.T;
        for (int k = 0; k < 5; ++k) a.fm[5 * c + k] = 0.0;
        a.fm[5 * c] = cs.phi[0];
        for (int i = 0; i < D; ++i) a.fm[5 * c + 1 + i] = cs.phi[1 + i];
        if (MODEL == CDUGKS_SHAKHOV) a.fm[5 * c + 4] = cs.phi[D + 1];
        if (!ok) flag_bad<MODEL>(a, c);
        CellRec<S, D> r;
        if constexpr (sizeof(S) == 4) {
            r.en = to_eqf(cs.en);
            r.e = to_eqf(cs.e);
            if (MODEL != CDUGKS_SHAKHOV) {
                r.en.hk = float(cs.en.A);
                r.e.hk = float(cs.e.A);
            }
            r.ai = a.s.half_dt / cs.tau;
            r.an = a.s.half_dt / cs.taun;
        } else {
            r.en = cs.en;
            r.e = cs.e;
            r.itau = 1.0 / cs.tau;
            r.itaun = 1.0 / cs.taun;
        }
        r.cm = cs.cm;
        r.C = cs.C;
        r.Cs = cs.Cs;
        r.dtv = cs.dt_v;
        for (int i = 0; i < D; ++i) r.u[i] = cs.u[i];
        static_cast<CellRec<S, D> *>(a.crec)[c] = r;
    }
    if (TAB) {
        S *dst = static_cast<S *>(a.cell_tab) + size_t(c) * 2 * a.nt;
        for (int t = lane; t < a.nt; t += 64) {
            dst[t] = tab_entry<S>(a, cs.en, t);
            dst[a.nt + t] = tab_entry<S>(a, cs.e, t);
        }
    }
}
template <int MODEL, class S, bool TAB, int D>
__global__ __launch_bounds__(BLOCK) void k_cell_prep(KArgs a) {
    d_cell_prep<MODEL, S, TAB, D>(a, blockIdx.x);
}
template <int MODEL, class S, bool TAB, int D>
__global__ __launch_bounds__(BLOCK) void kb_cell_prep(const KArgs *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_cell_prep<MODEL, S, TAB, D>(batch_args(tab, m.x), m.y);
}

// LSQ gradient g[0..D-1] of one item from the staged rows of distribution d:
// sum_j a_j (f_j - f_c) (mesh_numerical.cpp:24-49).
template <class S, int VT, int D>
__device__ __forceinline__ void item_grad(const meso_patch::ItemT<D> *it, const S *val_d, int vl, double fc,
                                          double *g) {
#pragma unroll
    for (int k = 0; k < D; ++k) g[k] = 0.0;
#pragma unroll
    for (int j = 0; j < meso_patch::Lim<D>::MAXN; ++j) {
        const double df = double(val_d[it->nb[j] * VT + vl]) - fc;
#pragma unroll
        for (int k = 0; k < D; ++k) g[k] += it->a[j][k] * df;
    }
}

// Value range [mn, mx] of one item and its LSQ (face) neighbours from the
// staged rows (padding slots repeat the item); for the limiter.
template <class S, int VT, int D>
__device__ __forceinline__ void item_range(const meso_patch::ItemT<D> *it, const S *val_d, int vl, double fc,
                                           double &mn, double &mx) {
    mn = mx = fc;
#pragma unroll
    for (int j = 0; j < meso_patch::Lim<D>::MAXN; ++j) {
        const double w = double(val_d[it->nb[j] * VT + vl]);
        mn = fmin(mn, w);
        mx = fmax(mx, w);
    }
}

// Block prologue: copy n records of R bytes (a multiple of 16) to LDS.
template <class R>
__device__ __forceinline__ void stage_copy(const R *src, int n, R *dst) {
    const int4 *si = reinterpret_cast<const int4 *>(src);
    int4 *di = reinterpret_cast<int4 *>(dst);
    for (int r = threadIdx.x; r < n * int(sizeof(R) / 16); r += BLOCKP) di[r] = si[r];
}

// FP32X storage (XS kernels, FP64 arithmetic): the rows hold the FP32 word and
// the extension byte of the deviation from f_ref; they are staged as FP64
// full values (decoded, f_ref of the tile velocity added back).
template <int ND, int VT>
__device__ __forceinline__ void stage_values_x(const KArgs &a, const int *rows, int nr, int nvp, int v, bool ok,
                                               const double ref[2], int slot, int vl, double *val, int maxr) {
    constexpr int NSLOT = (BLOCKP / 64) * (64 / VT);
    constexpr int B = 4;  // rows per batch
    const float *wg = (const float *)a.fin_g, *wh = (const float *)a.fin_h;
    const unsigned *xg = a.xin_g.as<const unsigned>(), *xh = a.xin_h.as<const unsigned>();
    for (int base = slot; base < nr; base += NSLOT * B) {
        float w[B][ND];
        unsigned e[B][ND];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int r = base + k * NSLOT;
            const bool on = ok && r < nr;
            const size_t o = size_t(on ? rows[r] : 0) * nvp + (on ? v : 0);
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                w[k][d] = on ? (d == 0 ? wg : wh)[o] : 0.0f;
                // the extension byte from the dword holding it (rows start
                // on 64-byte boundaries): dword loads shared by four lanes
                // took the FP32X moment kernel from 10.7 to 9.8 ms against
                // byte loads
                e[k][d] = on ? (d == 0 ? xg : xh)[o >> 2] >> (8 * (o & 3)) : 0u;
            }
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int r = base + k * NSLOT;
            if (r < nr)
#pragma unroll
                for (int d = 0; d < ND; ++d) val[(d * maxr + r) * VT + vl] = ok ? x32_decode(w[k][d], e[k][d]) + ref[d] : 0.0;
        }
    }
}

// Patch records: with one item per wave (IPW == 1) every record a wave reads
// is wave-uniform and is read from HBM through the scalar cache; with several
// items per wave the records are staged in LDS.
template <int IPW, int D>
struct Records {
    const meso_patch::ItemT<D> *items;
    const meso_patch::FaceT<D> *faces;
    __device__ __forceinline__ const meso_patch::ItemT<D> *item(int s) const { return items + s; }
    __device__ __forceinline__ const meso_patch::FaceT<D> *face(int lf) const { return faces + lf; }
};

// Bytes of the moment kernel's tile-velocity table: per velocity (x, y, z, w)
// in 3D; in 2D w and its products with the moment monomials
// (w, w x, w y, w x x, w x y, w y y, w |xi|^2 x, w |xi|^2 y).
__host__ __device__ constexpr size_t mp_vtile_bytes(int D, int VT) { return size_t(VT) * (D == 2 ? 64 : 32); }

// Bytes of LDS the records take (0 when they are read from HBM).
template <int IPW, int D>
__host__ __device__ constexpr size_t record_bytes(int maxi, int maxf) {
    return IPW > 1 ? size_t(maxi) * sizeof(meso_patch::ItemT<D>) + size_t(maxf) * sizeof(meso_patch::FaceT<D>) : 0;
}

// Stage the rows of one velocity tile: all loads of a batch are issued before
// their LDS stores.
template <int ND, class S, int VT>
__device__ __forceinline__ void stage_values(const int *rows, int nr, const S *fin_g, const S *fin_h, int nvp, int v,
                                             bool ok, int slot, int vl, S *val, int maxr) {
    constexpr int NSLOT = (BLOCKP / 64) * (64 / VT);
    constexpr int B = 4;  // rows per batch (FP32 moment kernel, 8 rows: within 1%, 5.02-5.06 against 5.06-5.13 ms)
    for (int base = slot; base < nr; base += NSLOT * B) {
        S x[B][ND];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int r = base + k * NSLOT;
            const bool on = ok && r < nr;
            const size_t o = size_t(on ? rows[r] : 0) * nvp + (on ? v : 0);
#pragma unroll
            for (int d = 0; d < ND; ++d) x[k][d] = on ? (d == 0 ? fin_g : fin_h)[o] : S(0);
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int r = base + k * NSLOT;
            if (r < nr)
#pragma unroll
                for (int d = 0; d < ND; ++d) val[(d * maxr + r) * VT + vl] = x[k][d];
        }
    }
}

// Lane velocity data of one tile.
struct LaneVel {
    double x, y, z, w;  // abscissae and weight (FP64); z = 0 in 2D
    float xf[3];        // FP32 abscissae
    int vt;           // packed axis-table offsets
    float gref;       // FP32 g_ref
    double gref_d, href_d;
    bool ok;          // v < nv
};

template <int D>
__device__ __forceinline__ LaneVel lane_vel(const KArgs &a, int v) {
    LaneVel L;
    L.ok = v < a.s.nv;
    const int vv = L.ok ? v : 0;
    const double4 X = a.vel_d[vv];
    L.x = X.x;
    L.y = X.y;
    L.z = D == 3 ? X.z : 0.0;
    L.w = L.ok ? X.w : 0.0;
    const float4 xf = a.vel_f[vv];
    L.xf[0] = xf.x;
    L.xf[1] = xf.y;
    L.xf[2] = D == 3 ? xf.z : 0.0f;
    L.vt = a.nt > 0 ? a.vtab[vv] : 0;
    L.gref = a.fref_f2[vv].x;
    const double2 fr = a.fref_d2[vv];
    L.gref_d = fr.x;
    L.href_d = fr.y;
    return L;
}

// Face equilibrium record of the storage kind (FP32 copy or FP64).
template <class S>
struct UpdateRecs;
template <>
struct UpdateRecs<float> {
    using Eq = EqF;
    __device__ static EqF face_eq(const KArgs &a, int f) { return a.face_eqf[f]; }
};
template <>
struct UpdateRecs<double> {
    using Eq = EqParams;
    __device__ static EqParams face_eq(const KArgs &a, int f) { return a.face_eq[f]; }
};

// f_ref flux factor sum(+-A n) of a cell as the update kernel keeps it
// (8-byte aligned: it follows records of 8-byte multiples in LDS).
struct CellS2 {
    double x, y;
};
struct CellS3 {
    double x, y, z, pad;
};
template <int D>
using CellS = typename std::conditional<D == 2, CellS2, CellS3>::type;

template <int D>
__device__ __forceinline__ CellS<D> cell_s(const double *p) {
    CellS<D> r;
    r.x = p[0];
    r.y = p[1];
    if constexpr (D == 3) {
        r.z = p[2];
        r.pad = 0.0;
    }
    return r;
}

// xi . sum(+-A n) of a cell
template <int D>
__device__ __forceinline__ double cell_xs(const CellS<D> &s, const double *xv) {
    double r = xv[0] * s.x + xv[1] * s.y;
    if constexpr (D == 3) r += xv[2] * s.z;
    return r;
}

// LDS layout of a k_patch_update block (byte offsets), shared by the kernel
// and the host that sizes the launch: item records of the owned cells | face
// records | stage rows | values | face fluxes | face equilibria | mix weights |
// cell constants | f_ref flux factors (FP32 storage, which stores deviations
// from f_ref).
__host__ __device__ constexpr size_t lds_r16(size_t x) { return (x + 15) & ~size_t(15); }
template <class S, int D>
struct UpdateLds {
    size_t faces, rows, val, flx, feq, fmix, crl, csl, bytes;
    __host__ __device__ UpdateLds(int maxo, int maxf, int vt, int nd) {
        faces = size_t(maxo) * sizeof(meso_patch::ItemT<D>);
        rows = faces + size_t(maxf) * sizeof(meso_patch::FaceT<D>);
        val = lds_r16(rows + 4 * size_t(maxo));
        flx = val + size_t(nd) * maxo * vt * sizeof(S);
        feq = lds_r16(flx + size_t(nd) * maxf * vt * sizeof(S));
        fmix = feq + size_t(maxf) * sizeof(typename UpdateRecs<S>::Eq);
        crl = fmix + size_t(maxf) * 8;
        csl = crl + size_t(maxo) * sizeof(CellRec<S, D>);
        bytes = csl + (sizeof(S) == 4 ? size_t(maxo) * sizeof(CellS<D>) : 0);
    }
};

// Equilibrium pair of face f at the lane velocity from its record e, in the
// storage form (deviation for FP32, full value for FP64), as FP64.
// With PF the axis-table entries of the face come in tav (read earlier).
template <int MODEL, class S, bool TAB, int D, bool PF = false>
__device__ __forceinline__ void face_eq_rec(const KArgs &a, const typename UpdateRecs<S>::Eq &e, int f,
                                            const LaneVel &L, double &gs, double &hs, const S *tav = nullptr) {
    if constexpr (sizeof(S) == 4) {
        float g, h = 0.0f;
        if (MODEL == CDUGKS_SHAKHOV) {
            if constexpr (TAB && PF)
                shakhov_tab_vals_f(e, tav[0], tav[1], tav[2], L.xf, D, a.s.K, L.gref, g, h);
            else if (TAB)
                shakhov_tab_f(e, static_cast<const float *>(a.face_tab) + size_t(f) * a.nt, L.vt, L.xf, D, a.s.K,
                              L.gref, g, h);
            else
                shakhov_devf(e, L.xf, D, a.s.K, L.gref, g, h);
        } else {
            g = bgk_devf(e, L.xf);
        }
        gs = g;
        hs = h;
    } else {
        const double x[3] = {L.x, L.y, L.z};
        if (MODEL == CDUGKS_SHAKHOV) {
            if constexpr (TAB && PF)
                shakhov_tab_vals_d(e, tav[0], tav[1], tav[2], x, D, a.s.K, gs, hs);
            else if (TAB)
                shakhov_tab_d(e, static_cast<const double *>(a.face_tab) + size_t(f) * a.nt, L.vt, x, D, a.s.K,
                              gs, hs);
            else
                shakhov_eq<double>(e, x, D, a.s.K, gs, hs);
        } else {
            gs = bgk_eq<double>(e, x);
            hs = 0.0;
        }
    }
}

// Value of boundary face f at the lane velocity, in the storage form, from the
// mixed value fm (gm, hm) there: the incoming half takes the boundary
// equilibrium (FP64), the outgoing half keeps fm; on the dead-band a wall takes
// the average (cdugks_shakhov.cpp:364-607). Mirror boundaries (symmetry,
// slip-wall) never reach this kernel.
template <int MODEL, class S, int D>
__device__ __forceinline__ void bnd_value_lane(const KArgs &a, int f, double sn, const LaneVel &L, double gm,
                                               double hm, double &g, double &h) {
    g = gm;
    h = hm;
    const int bi = a.face_bnd[f];
    if (bi < 0) return;
    const BndParams &P = a.bnd[bi];
    const double kn = -sn;  // xi . n1, n1 = -n0 points into the domain
    if (P.kind == BC_INTERIOR || kn < -VSMALL) return;
    const double x[3] = {L.x, L.y, L.z};
    double gi, hi = 0.0;
    if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(P.in, x, D, a.s.K, gi, hi);
    else gi = bgk_eq<double>(P.in, x);
    if constexpr (sizeof(S) == 4) {
        gi -= L.gref_d;
        hi -= L.href_d;
    }
    if (kn >= VSMALL) {
        g = gi;
        h = hi;
    } else if (P.kind == BC_WALL) {
        g = 0.5 * (gi + gm);
        h = 0.5 * (hi + hm);
    }
}

// ------------------------------------------------------------- moments
// Raw moments of f-bar+ per face of the moment patches, summed per velocity
// chunk into part_f[face][tile][col] (the layout k_face_eq reads).
//   CELLS false: face patches; every face of the patch gets all of its terms
//     (tile = chunk), the items are the cells on either side.
//   CELLS true: cell patches; the patch forms the gradient of its owned
//     cells only and adds, per face, the terms of the velocities for which an
//     owned cell is upwind (half of each dead-band term), into the slot of
//     that side (tile = side * p_chunks + chunk; the slot of a side the patch
//     does not own is written by the patch that owns it, a face with both
//     sides owned puts its sum in slot 0 and zeros in slot 1).
// XS: FP32X storage (S = double): the cell rows in HBM as FP32 word and
// extension byte of the deviation from f_ref (f-bar+ stays FP64).
// b: the block's logical index in its context's grid (XCD-mapped, xcd_block).
template <int MODEL, class S, int VT, int D, bool LIM, bool CELLS, bool XS = false>
__device__ __forceinline__ void d_patch_moments(const KArgs &a, const int b) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (*a.bad_count > 0) {
        if (b == 0 && threadIdx.x == 0) *a.halt = 1;
        return;
    }
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    constexpr int ND = SH ? 2 : 1;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr int IPW = 64 / VT;
    constexpr int NW = BLOCKP / 64;
    constexpr int NSLOT = NW * IPW;
    constexpr int VTP = VT + 1;  // odd row stride: conflict-free reads with lane = face
    using Item = meso_patch::ItemT<D>;
    using Face = meso_patch::FaceT<D>;
    const Item *mp_items = static_cast<const Item *>(a.mp_items);
    const Face *mp_faces = static_cast<const Face *>(a.mp_faces);
    const int np = a.n_mp;
    const int chunk = b / np, p = b - chunk * np;
    const int i0 = a.mp_item_ptr[p], ni = a.mp_item_ptr[p + 1] - i0;
    const int f0 = a.mp_face_ptr[p], nfp = a.mp_face_ptr[p + 1] - f0;
    const int nown = CELLS ? a.mp_nown[p] : ni;
    // cell patches stage the items only (the owned cells' LSQ neighbours are
    // the cells across their faces)
    const int r0 = a.mp_row_ptr[p], nr = CELLS ? ni : a.mp_row_ptr[p + 1] - r0;
    const int maxi = a.mp_maxi, maxf = a.mp_maxf, maxr = CELLS ? a.mp_maxi : a.mp_maxr;
    // LDS: [items | faces] | rows | tile velocities | values | f-bar+ rows
    unsigned char *q = lds_raw;
    Records<IPW, D> rec;
    if constexpr (IPW > 1) {
        auto *li = reinterpret_cast<Item *>(q);
        auto *lf = reinterpret_cast<Face *>(li + maxi);
        stage_copy(mp_items + i0, ni, li);
        stage_copy(mp_faces + f0, nfp, lf);
        rec.items = li;
        rec.faces = lf;
        q += record_bytes<IPW, D>(maxi, maxf);
    } else {
        rec.items = mp_items + i0;
        rec.faces = mp_faces + f0;
    }
    int *rows = reinterpret_cast<int *>(q);
    q += (size_t(maxr) * 4 + 31) & ~size_t(31);
    constexpr int VW = D == 2 ? 8 : 4;  // doubles per tile velocity (mp_vtile_bytes)
    double *vtd = reinterpret_cast<double *>(q);
    S *val = reinterpret_cast<S *>(vtd + VT * VW);
    S *fb = val + size_t(ND) * maxr * VT;
    for (int r = threadIdx.x; r < nr; r += BLOCKP) rows[r] = a.mp_rows[r0 + r];
    const int lane = threadIdx.x % 64;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x / 64));
    const int sub = IPW > 1 ? lane / VT : 0, vl = IPW > 1 ? lane % VT : lane;
    const int slot = wave * IPW + sub;
    const int nvp = a.s.nvp;
    const double hdt = a.s.half_dt;
    const S *fin_g = (const S *)a.fin_g, *fin_h = (const S *)a.fin_h;
    // phase-C layout: thread = (face, velocity subset); FG faces per subset
    // (FG = the patch's faces, nvs = BLOCKP / FG subsets; the last
    // BLOCKP mod FG threads idle)
    const int FG = max(nfp, 1);
    const int my_face = threadIdx.x % FG, vs = threadIdx.x / FG, nvs = BLOCKP / FG;
    const bool face_act = vs < nvs;
    double acc[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = 0.0;
    const int ntiles = (a.s.nv + VT - 1) / VT;
    const int t0 = chunk * (a.p_vchunk / VT), t1 = min(t0 + a.p_vchunk / VT, ntiles);
    // FP32 storage: seed of the stochastic rounding of the handed-over face
    // values (the device step counter; its own stream, apart from the state's)
    const unsigned fseed = sizeof(S) == 4 ? mix32(unsigned(*a.step_ctr) * 0x85ebca6bu + 0x27d4eb2fu) : 0u;
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int v = t * VT + vl;
        const bool ok = v < a.s.nv;
        const double4 X = a.vel_d[ok ? v : 0];
        // A: stage values and tile velocities, clear the face rows
        double ref[2] = {0.0, 0.0};  // XS: f_ref of the tile velocity (g, h)
        if constexpr (XS && sizeof(S) == 8) {
            ref[0] = ok ? a.fref_d[v] : 0.0;
            ref[1] = ok && SH ? a.fref_d[a.s.nv + v] : 0.0;
            stage_values_x<ND, VT>(a, rows, nr, nvp, v, ok, ref, slot, vl, (double *)val, maxr);
        } else {
            // FP32 and FP32E storage: the rows' FP32 words (with FP32E, the
            // words are the 31-bit deviations rounded to nearest FP32)
            stage_values<ND, S, VT>(rows, nr, fin_g, fin_h, nvp, v, ok, slot, vl, val, maxr);
        }
        if (threadIdx.x < VT) {
            double *y = vtd + threadIdx.x * VW;
            if constexpr (D == 2) {
                // the weights of the moment sums: one FMA per moment and value
                const double w = ok ? X.w : 0.0, x = X.x, z = X.y, kk = x * x + z * z;
                y[0] = w;
                y[1] = w * x;
                y[2] = w * z;
                y[3] = w * x * x;
                y[4] = w * x * z;
                y[5] = w * z * z;
                y[6] = w * kk * x;
                y[7] = w * kk * z;
            } else {
                *reinterpret_cast<double4 *>(y) = ok ? X : make_double4(0.0, 0.0, 0.0, 0.0);
            }
        }
        for (int r = threadIdx.x; r < nfp * VTP; r += BLOCKP) {
            fb[r] = S(0);
            if (SH) fb[maxf * VTP + r] = S(0);
        }
        __syncthreads();
        // B: gradients and f-bar+ of the patch faces
        // FP32: random bits of the face values' rounding, one hashed word per
        // lane and tile, stepped by a Weyl increment per (item, face)
        unsigned rb = sizeof(S) == 4 ? mix32(fseed ^ (unsigned(t) << 20) ^ unsigned(b * BLOCKP + int(threadIdx.x)))
                                     : 0u;
        for (int s = slot; s < nown; s += NSLOT, rb += unsigned(meso_patch::Lim<D>::MAXF) * 0x9e3779b9u) {
            const Item *it = rec.item(s);
            const int nface = it->nface;
            double fc[2], g[2][D], mn[2], mx[2];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                fc[d] = double(val[(d * maxr + s) * VT + vl]);
#pragma unroll
                for (int k = 0; k < D; ++k) g[d][k] = 0.0;
                if (a.s.grad) item_grad<S, VT, D>(it, val + d * maxr * VT, vl, fc[d], g[d]);
                if constexpr (LIM) item_range<S, VT, D>(it, val + d * maxr * VT, vl, fc[d], mn[d], mx[d]);
            }
            const double omega = LIM ? venkata_omega(a.s.venkata_k, a.cell_vol[it->cell], D) : 0.0;
            const double xv[3] = {X.x, X.y, X.z};
            // the item's face records (entry, face id, normal, offset from
            // this item) are read for all faces before any is used: one round
            // trip through the scalar cache per item instead of three per
            // face (not for the limited and FP64 3D instances, whose
            // registers are full)
            constexpr int MF = meso_patch::Lim<D>::MAXF;
            constexpr bool PRE = !LIM && !(sizeof(S) == 8 && D == 3);
            int fe[MF], fgl[MF];
            double fnv[MF][D], fdv[MF][D];
            if constexpr (PRE) {
#pragma unroll
                for (int k = 0; k < MF; ++k) {
                    fe[k] = it->face[k];
                    // entries past nface are padding: read face 0 instead
                    const Face *F = rec.face(k < nface ? fe[k] & 0xffff : 0);
                    const int sd = (fe[k] >> 16) == meso_patch::SIDE_NEIGHBOUR ? 1 : 0;
                    fgl[k] = F->face;
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        fnv[k][i] = F->n[i];
                        fdv[k][i] = F->d[sd][i];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < MF; ++k) {
                if (k >= nface) continue;
                const int e = PRE ? fe[k] : it->face[k];
                const int lf = e & 0xffff, mode = e >> 16;
                const int side = mode == meso_patch::SIDE_NEIGHBOUR ? 1 : 0;
                double fn[D], fd[D];
                int fg;
                if constexpr (PRE) {
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        fn[i] = fnv[k][i];
                        fd[i] = fdv[k][i];
                    }
                    fg = fgl[k];
                } else {
                    const Face *F = rec.face(lf);
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        fn[i] = F->n[i];
                        fd[i] = F->d[side][i];
                    }
                    fg = F->face;
                }
                double sn = 0.0;  // in 2D: + n_z xi_z, both 0
#pragma unroll
                for (int i = 0; i < D; ++i) sn += fn[i] * xv[i];
                const bool up = mode == meso_patch::SIDE_BOUNDARY ||
                                (mode == meso_patch::SIDE_OWNER ? sn >= VSMALL : sn < -VSMALL);
                const bool dead = !(sn >= VSMALL) && !(sn < -VSMALL);
                double ev[D];
#pragma unroll
                for (int i = 0; i < D; ++i) ev[i] = fd[i] - xv[i] * hdt;
                if (!ok) continue;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    double Xd = 0.0;
#pragma unroll
                    for (int i = 0; i < D; ++i) Xd += ev[i] * g[d][i];
                    if constexpr (LIM) {
                        double dw = 0.0;
#pragma unroll
                        for (int i = 0; i < D; ++i) dw += fd[i] * g[d][i];
                        Xd *= venkata_phi(dw, fc[d], mn[d], mx[d], omega, sn >= VSMALL ? 1.0e-10 : VSMALL);
                    }
                    S *dst = &fb[(d * maxf + lf) * VTP + vl];
                    // the face values also go to HBM for the update kernel
                    // (an item's lanes are consecutive velocities of one
                    // face); interior dead-band values are not needed there
                    // (no flux)
                    const size_t fi = size_t(fg) * nvp + v;
                    if (mode == meso_patch::SIDE_BOUNDARY || up) {
                        // (FP32 storage: the boundary kernel reconstructs
                        // its faces' values again in FP64 rather than reading
                        // these rounded ones; keeping unrounded copies here
                        // took the 2D kernel from 69 to 96 VGPRs, 3 to 2
                        // blocks per CU, 4.9 to 6.9 ms)
                        const double fvd = mode == meso_patch::SIDE_BOUNDARY && dead ? 0.5 * (fc[d] + fc[d] + Xd + Xd)
                                                                                     : fc[d] + Xd;
                        // FP32: stochastically rounded (unbiased): round to
                        // nearest hands the update the same rounding error
                        // every step, a bias that shifts the steady state
                        S fv;
                        if constexpr (sizeof(S) == 4) {
                            const unsigned rf = rb + unsigned(k) * 0x9e3779b9u;
                            fv = f32_sr(fvd, d == 0 ? rf : rot16(rf));
                        } else {
                            fv = S(fvd);
                        }
                        *dst = fv;
                        (d == 0 ? (S *)a.fbar_g : (S *)a.fbar_h)[fi] = fv;
                    } else if (dead) {
                        atomicAdd(dst, S(0.5 * (fc[d] + Xd)));
                    }
                }
            }
        }
        __syncthreads();
        // C: moments, thread = (face, velocity subset)
        if (face_act) {
            for (int vv = vs; vv < VT; vv += nvs) {
                const double *y = vtd + vv * VW;
                if constexpr (D == 2) {
                    const double g = double(fb[my_face * VTP + vv]);
                    acc[Col<D>::M0] += y[0] * g;
                    acc[Col<D>::M1] += y[1] * g;
                    acc[Col<D>::M1 + 1] += y[2] * g;
                    acc[Col<D>::M2] += y[3] * g;
                    acc[Col<D>::M2 + 1] += y[4] * g;
                    acc[Col<D>::M2 + 2] += y[5] * g;
                    if (SH) {
                        const double h = double(fb[(maxf + my_face) * VTP + vv]);
                        acc[Col<D>::M3] += y[6] * g;
                        acc[Col<D>::M3 + 1] += y[7] * g;
                        acc[Col<D>::H0] += y[0] * h;
                        acc[Col<D>::H1] += y[1] * h;
                        acc[Col<D>::H1 + 1] += y[2] * h;
                    }
                    continue;
                }
                const double4 Y = *reinterpret_cast<const double4 *>(y);
                const double wg = Y.w * double(fb[my_face * VTP + vv]);
                const double wgx = wg * Y.x, wgy = wg * Y.y, wgz = D == 3 ? wg * Y.z : 0.0;
                acc[Col<D>::M0] += wg;
                acc[Col<D>::M1] += wgx;
                acc[Col<D>::M1 + 1] += wgy;
                if (D == 3) acc[Col<D>::M1 + 2] += wgz;
                if (D == 2) {
                    acc[Col<D>::M2] += wgx * Y.x;
                    acc[Col<D>::M2 + 1] += wgx * Y.y;
                    acc[Col<D>::M2 + 2] += wgy * Y.y;
                } else {
                    acc[Col<D>::M2] += wgx * Y.x;
                    acc[Col<D>::M2 + 1] += wgx * Y.y;
                    acc[Col<D>::M2 + 2] += wgx * Y.z;
                    acc[Col<D>::M2 + 3] += wgy * Y.y;
                    acc[Col<D>::M2 + 4] += wgy * Y.z;
                    acc[Col<D>::M2 + 5] += wgz * Y.z;
                }
                if (SH) {
                    const double wh = Y.w * double(fb[(maxf + my_face) * VTP + vv]);
                    const double wgk = wg * (Y.x * Y.x + Y.y * Y.y + Y.z * Y.z);
                    acc[Col<D>::M3] += wgk * Y.x;
                    acc[Col<D>::M3 + 1] += wgk * Y.y;
                    if (D == 3) acc[Col<D>::M3 + 2] += wgk * Y.z;
                    acc[Col<D>::H0] += wh;
                    acc[Col<D>::H1] += wh * Y.x;
                    acc[Col<D>::H1 + 1] += wh * Y.y;
                    if (D == 3) acc[Col<D>::H1 + 2] += wh * Y.z;
                }
            }
        }
        __syncthreads();
    }
    // combine the velocity subsets of each face in a fixed order
    double *red = reinterpret_cast<double *>(lds_raw);  // [nvs][FG][NCOL]
    if (face_act)
#pragma unroll
        for (int k = 0; k < NCOL; ++k) red[(size_t(vs) * FG + my_face) * NCOL + k] = acc[k];
    __syncthreads();
    for (int r = threadIdx.x; r < nfp * NCOL; r += BLOCKP) {
        const int lf = r / NCOL, k = r - lf * NCOL;
        double sum = 0.0;
        for (int qq = 0; qq < nvs; ++qq) sum += red[(size_t(qq) * FG + lf) * NCOL + k];
        const Face &F = mp_faces[f0 + lf];
        const size_t base = size_t(F.face) * a.s.ntile_f;
        if constexpr (CELLS) {
            const bool oo = F.own < nown, on = F.nei < nown;
            if (oo) a.part_f[(base + chunk) * NCOL + k] = sum;
            if (on && !oo) a.part_f[(base + a.p_chunks + chunk) * NCOL + k] = sum;
            else if (oo && (on || F.own == F.nei)) a.part_f[(base + a.p_chunks + chunk) * NCOL + k] = 0.0;
        } else {
            a.part_f[(base + chunk) * NCOL + k] = sum;
        }
    }
}
template <int MODEL, class S, int VT, int D, bool LIM, bool CELLS, bool XS = false>
__global__ __launch_bounds__(BLOCKP, 4) void k_patch_moments(KArgs a) {
    d_patch_moments<MODEL, S, VT, D, LIM, CELLS, XS>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <int MODEL, class S, int VT, int D, bool LIM, bool CELLS, bool XS = false>
__global__ __launch_bounds__(BLOCKP, 4) void kb_patch_moments(const KArgs *__restrict__ tab,
                                                               const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_patch_moments<MODEL, S, VT, D, LIM, CELLS, XS>(batch_args(tab, m.x), m.y);
}

// ------------------------------------------------------------- update
// Cells of the update patches: face fluxes from the f-bar+ rows the moment
// kernel wrote (limited there when LIM), cell flux, f-tilde-plus update and
// the Shakhov heat-flux partials part_c[cell][chunk][0..D-1]. The equilibrium
// records of the patch faces and the update constants of the owned cells are
// copied to LDS once per block (read per lane from HBM in every tile they
// would cost address arithmetic and L2 latency). LIM only selects the
// instance that pairs with the limited moment kernel.
template <int MODEL, class S, int VT, bool TAB, int D, bool LIM, bool XS = false>
__device__ __forceinline__ void d_patch_update(const KArgs &a, const int b) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (*a.halt) return;
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    constexpr int ND = SH ? 2 : 1;
    constexpr int IPW = 64 / VT;
    constexpr int NW = BLOCKP / 64;
    constexpr int NSLOT = NW * IPW;
    using Item = meso_patch::ItemT<D>;
    using Face = meso_patch::FaceT<D>;
    const Item *up_items = static_cast<const Item *>(a.up_items);
    const Face *up_faces = static_cast<const Face *>(a.up_faces);
    const int np = a.n_up;
    const int chunk = b / np, p = b - chunk * np;
    const int i0 = a.up_item_ptr[p];
    const int f0 = a.up_face_ptr[p], nfp = a.up_face_ptr[p + 1] - f0;
    const int r0 = a.up_row_ptr[p];
    const int nown = a.up_nown[p];
    // the face values f-bar+ come from the moment kernel (HBM), so only the
    // owned cells' items and rows are staged (they come first in the patch)
    const int maxi = a.up_maxo, maxf = a.up_maxf, maxr = a.up_maxo;
    // LDS (UpdateLds): items | faces | rows | values | face fluxes | face
    // equilibria and mix weights of the patch faces and update constants of
    // the owned cells (copied once per block)
    using EqS = typename UpdateRecs<S>::Eq;
    const UpdateLds<S, D> lay(maxi, maxf, VT, ND);
    auto *items = reinterpret_cast<Item *>(lds_raw);
    auto *faces = reinterpret_cast<Face *>(lds_raw + lay.faces);
    int *rows = reinterpret_cast<int *>(lds_raw + lay.rows);
    S *val = reinterpret_cast<S *>(lds_raw + lay.val);
    S *flx = reinterpret_cast<S *>(lds_raw + lay.flx);
    EqS *feq = reinterpret_cast<EqS *>(lds_raw + lay.feq);
    double *fmix = reinterpret_cast<double *>(lds_raw + lay.fmix);
    CellRec<S, D> *crl = reinterpret_cast<CellRec<S, D> *>(lds_raw + lay.crl);
    CellS<D> *csl = reinterpret_cast<CellS<D> *>(lds_raw + lay.csl);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int sub = lane / VT, vl = lane % VT;
    const int slot = wave * IPW + sub;
    const int nvp = a.s.nvp;
    const double hdt = a.s.half_dt;
    const S *fin_g = (const S *)a.fin_g, *fin_h = (const S *)a.fin_h;
    S *fout_g = (S *)a.fout_g, *fout_h = (S *)a.fout_h;
    stage_copy(up_items + i0, nown, items);
    stage_copy(up_faces + f0, nfp, faces);
    for (int r = threadIdx.x; r < nown; r += BLOCKP) rows[r] = a.up_rows[r0 + r];
    for (int lf = threadIdx.x; lf < nfp; lf += BLOCKP) {
        const int f = up_faces[f0 + lf].face;
        feq[lf] = UpdateRecs<S>::face_eq(a, f);
        fmix[lf] = a.face_eq[f].mix;
    }
    for (int r = threadIdx.x; r < nown; r += BLOCKP) {
        const int c = up_items[i0 + r].cell;
        crl[r] = static_cast<const CellRec<S, D> *>(a.crec)[c];
        if constexpr (sizeof(S) == 4) csl[r] = cell_s<D>(a.cell_S + 3 * c);
    }
    const int ntiles = (a.s.nv + VT - 1) / VT;
    const int t0 = chunk * (a.p_vchunk / VT), t1 = min(t0 + a.p_vchunk / VT, ntiles);
    // FP32 and FP32X: seed of the stochastic rounding of the stored values (the
    // device step counter)
    const unsigned seed = XS || sizeof(S) == 4 ? mix32(unsigned(*a.step_ctr) * 0x9e3779b9u + 0x632be5abu) : 0u;
    const unsigned fxseed = mix32(seed ^ 0x5bd1e995u);  // FP32: the face flux rows' own stream
    // Shakhov heat-flux partials of this thread's owned cells over the chunk
    constexpr int UPD_QJ = upd_qj(VT);
    double qreg[UPD_QJ][D];
#pragma unroll
    for (int j = 0; j < UPD_QJ; ++j)
#pragma unroll
        for (int i = 0; i < D; ++i) qreg[j][i] = 0.0;
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int v = t * VT + vl;
        const LaneVel L = lane_vel<D>(a, v);
        const double xv[3] = {L.x, L.y, L.z};
        // A: stage the owned cells' values (read in C only: no barrier before B)
        double ref[2] = {0.0, 0.0};  // XS: f_ref of the lane velocity (g, h)
        if constexpr (XS && sizeof(S) == 8) {
            ref[0] = L.ok ? a.fref_d[v] : 0.0;
            ref[1] = L.ok && SH ? a.fref_d[a.s.nv + v] : 0.0;
            stage_values_x<ND, VT>(a, rows, nown, nvp, v, L.ok, ref, slot, vl, (double *)val, maxr);
        } else {
            stage_values<ND, S, VT>(rows, nown, fin_g, fin_h, nvp, v, L.ok, slot, vl, val, maxr);
        }
        // B: per face and velocity (all lanes busy): f-bar+ (written by the
        // moment kernel), its equilibrium mix,
        // the boundary value, flux (xi.n) A f_face into the same slot
        if (L.ok) {
            constexpr int FS = BLOCKP / VT;  // faces per pass of the block
            const S *fbg = (const S *)a.fbar_g, *fbh = (const S *)a.fbar_h;
            // Shakhov with axis tables: each face's table entries are read with
            // its f-bar+ values (PF)
            constexpr bool PF = TAB && SH;
            // FP32: random bits of the flux rows' rounding, one hashed word
            // per lane and tile, stepped by a Weyl increment per face
            unsigned rq = sizeof(S) == 4 ? mix32(fxseed ^ (unsigned(t) << 20) ^ unsigned(b * BLOCKP + int(threadIdx.x)))
                                         : 0u;
            auto face_flux = [&](const int lf, const S *fbv, const S *tav, const unsigned rk) {
                const Face *F = faces + lf;
                double sn = 0.0;
#pragma unroll
                for (int i = 0; i < D; ++i) sn += F->n[i] * xv[i];
                const bool bnd = F->own == F->nei;
                const bool dead = !(sn >= VSMALL) && !(sn < -VSMALL);
                const int f = F->face;
                double gs, hs = 0.0;
                face_eq_rec<MODEL, S, TAB, D, PF>(a, feq[lf], f, L, gs, hs, tav);
                const double mix = fmix[lf];
                double fm[2] = {0.0, 0.0};
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double fbar = double(fbv[d]);
                    const double eq = d == 0 ? gs : hs;
                    if constexpr (sizeof(S) == 8) fm[d] = (1.0 - mix) * fbar + mix * eq;
                    else fm[d] = fbar + mix * (eq - fbar);
                }
                if (bnd) bnd_value_lane<MODEL, S, D>(a, f, sn, L, fm[0], fm[1], fm[0], fm[1]);
                const double knA = sn * F->area;
                if constexpr (sizeof(S) == 4) {
                    // FP32: the face flux rows in LDS are stochastically rounded
                    // as well. Rounded to nearest, their rounding error is a
                    // fixed function of the state, a bias that moves the steady
                    // state: on the Kn = 1 cavity at quarter resolution it was
                    // 57% of the FP32 run's error at convergence (qy 1.36e-6
                    // against FP64; 5.8e-7 with this rounding)
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                        flx[(d * maxf + lf) * VT + vl] = dead && !bnd ? 0.0f : f32_sr(knA * fm[d], d == 0 ? rk : rot16(rk));
                } else {
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                        flx[(d * maxf + lf) * VT + vl] = S(dead && !bnd ? 0.0 : knA * fm[d]);
                }
            };
            // FP32: four faces per pass, all four faces' f-bar+ loads (and
            // axis-table entries) issued before any is used (more loads in
            // flight per wave; 2 and 8 were slower); FP64 one (two were slower)
            constexpr int NB = sizeof(S) == 4 ? 4 : 1;
            const S *ftab = static_cast<const S *>(a.face_tab);
            for (int lf = threadIdx.x / VT; lf < nfp; lf += NB * FS) {
                S fbv[NB][2];
                S tav[NB][3];
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    const int l = lf + k * FS < nfp ? lf + k * FS : lf;
                    const int fk = faces[l].face;
                    const size_t o = size_t(fk) * nvp + v;
                    fbv[k][0] = fbg[o];
                    fbv[k][1] = SH ? fbh[o] : S(0);
                    if constexpr (PF) {
                        const S *tb = ftab + size_t(fk) * a.nt;
                        tav[k][0] = tb[L.vt & 1023];
                        tav[k][1] = tb[(L.vt >> 10) & 1023];
                        tav[k][2] = D == 3 ? tb[(L.vt >> 20) & 1023] : S(0);
                    }
                }
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (lf + k * FS < nfp) face_flux(lf + k * FS, fbv[k], tav[k], rq + unsigned(k) * 0x9e3779b9u);
                rq += unsigned(NB) * 0x9e3779b9u;
            }
        }
        __syncthreads();
        // C: owned cells (this thread's are slot + j NSLOT, the same in every
        // tile); FP32 2D two cells per step, so that the second cell's table
        // and LDS reads overlap the first cell's arithmetic (FP64 and 3D:
        // slower; FP32E: one, as two spill registers: 7.07 against 6.68 ms)
        constexpr int CUNR = sizeof(S) == 4 && D == 2 && !XS ? 2 : 1;
        // FP32: random bits of the stored state's rounding, one hashed word per
        // lane and tile, stepped by a Weyl increment per owned cell
        const unsigned rs = mix32(seed ^ (unsigned(t) << 20) ^ unsigned(b * BLOCKP + int(threadIdx.x)) ^ 0x2545f491u);
#pragma unroll CUNR
        for (int s = slot, j = 0; s < nown; s += NSLOT, ++j) {
            const Item *it = items + s;
            const int c = it->cell;
            const int nface = it->nface;
            // Shakhov: the cell's axis-table entries (tables of g_S^n and
            // g_S^{n+1}, in the storage type) are read first, so that their
            // latency overlaps the flux sums
            S ta[2][3];
            if constexpr (TAB && SH) {
                const S *tb = static_cast<const S *>(a.cell_tab) + size_t(c) * 2 * a.nt;
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    ta[q][0] = tb[q * a.nt + (L.vt & 1023)];
                    ta[q][1] = tb[q * a.nt + ((L.vt >> 10) & 1023)];
                    ta[q][2] = D == 3 ? tb[q * a.nt + ((L.vt >> 20) & 1023)] : S(0);
                }
            }
            // FP32E: the dwords holding the cell's extension bytes (shared by
            // four lanes), read here for the same reason
            unsigned xe[2] = {0u, 0u};
            if constexpr (XS && sizeof(S) == 4) {
                const size_t o = size_t(c) * nvp + (L.ok ? v : 0);
                xe[0] = a.xin_g.template as<const unsigned>()[o >> 2] >> (8 * (o & 3));
                if (SH) xe[1] = a.xin_h.template as<const unsigned>()[o >> 2] >> (8 * (o & 3));
            }
            double F[2] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < meso_patch::Lim<D>::MAXF; ++k) {
                if (k >= nface) continue;
                const int e = it->face[k];
                const int lf = e & 0xffff;
                const bool nei = (e >> 16) == meso_patch::SIDE_NEIGHBOUR;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double fx = double(flx[(d * maxf + lf) * VT + vl]);
                    F[d] += nei ? -fx : fx;
                }
            }
            const CellRec<S, D> &R = crl[s];
            double qv[3] = {0.0, 0.0, 0.0};
            double fcg = double(val[s * VT + vl]);
            double fch = SH ? double(val[(maxr + s) * VT + vl]) : 0.0;
            if constexpr (XS && sizeof(S) == 4) {
                // FP32E: the staged word and its extension byte
                fcg = x32_decode(val[s * VT + vl], xe[0]);
                if (SH) fch = x32_decode(val[(maxr + s) * VT + vl], xe[1]);
            }
            double og, oh = 0.0;
            if (SH) {
                double gsn, hsn, gs, hs;
                if constexpr (sizeof(S) == 4) {
                    // f_ref flux through the cell: f_ref xi . sum(+-A n), non-zero only for unclosed cells
                    const double xs = cell_xs<D>(csl[s], xv);
                    F[0] += L.gref_d * xs;
                    F[1] += L.href_d * xs;
                    float g1, h1, g2, h2;
                    if (TAB) {
                        shakhov_tab_vals_f(R.en, ta[0][0], ta[0][1], ta[0][2], L.xf, D, a.s.K, L.gref, g1, h1);
                        shakhov_tab_vals_f(R.e, ta[1][0], ta[1][1], ta[1][2], L.xf, D, a.s.K, L.gref, g2, h2);
                    } else {
                        shakhov_devf(R.en, L.xf, D, a.s.K, L.gref, g1, h1);
                        shakhov_devf(R.e, L.xf, D, a.s.K, L.gref, g2, h2);
                    }
                    gsn = g1;
                    hsn = h1;
                    gs = g2;
                    hs = h2;
                    // increments of the stored deviation (cdugks_shakhov.cpp:679-695 rearranged):
                    // g = g_n + C [a (g_S - g_n) + a_n (g_S^n - g_n) - dt/V F], a = h/tau, a_n = h/tau_n
                    const double gn = fcg + R.cm * (gsn - fcg), hn = fch + R.cm * (hsn - fch);
                    const double g = gn + R.C * (R.ai * (gs - gn) + R.an * (gsn - gn) - R.dtv * F[0]);
                    const double h = hn + R.C * (R.ai * (hs - hn) + R.an * (hsn - hn) - R.dtv * F[1]);
                    og = g + R.Cs * (gs - g);
                    oh = h + R.Cs * (hs - h);
                    const double gf = g + L.gref_d, hf = h + L.href_d;
                    double cv[D], cc = 0.0;
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        cv[i] = xv[i] - R.u[i];
                        cc += cv[i] * cv[i];
                    }
                    const double tq = 0.5 * L.w * (cc * gf + hf);
#pragma unroll
                    for (int i = 0; i < D; ++i) qv[i] = tq * cv[i];
                } else {
                    if (TAB) {
                        shakhov_tab_vals_d(R.en, ta[0][0], ta[0][1], ta[0][2], xv, D, a.s.K, gsn, hsn);
                        shakhov_tab_vals_d(R.e, ta[1][0], ta[1][1], ta[1][2], xv, D, a.s.K, gs, hs);
                    } else {
                        shakhov_eq<double>(R.en, xv, D, a.s.K, gsn, hsn);
                        shakhov_eq<double>(R.e, xv, D, a.s.K, gs, hs);
                    }
                    // cdugks_shakhov.cpp:679-695 as written
                    const double gn = R.cm * gsn + (1.0 - R.cm) * fcg;
                    const double hn = R.cm * hsn + (1.0 - R.cm) * fch;
                    const double g = R.C * (gn + hdt * (gs * R.itau + (gsn - gn) * R.itaun) - R.dtv * F[0]);
                    const double h = R.C * (hn + hdt * (hs * R.itau + (hsn - hn) * R.itaun) - R.dtv * F[1]);
                    og = (1.0 - R.Cs) * g + R.Cs * gs;
                    oh = (1.0 - R.Cs) * h + R.Cs * hs;
                    double cv[D], cc = 0.0;
#pragma unroll
                    for (int i = 0; i < D; ++i) {
                        cv[i] = xv[i] - R.u[i];
                        cc += cv[i] * cv[i];
                    }
                    const double tq = 0.5 * L.w * (cc * g + h);
#pragma unroll
                    for (int i = 0; i < D; ++i) qv[i] = tq * cv[i];
                }
            } else {
                if constexpr (sizeof(S) == 4) {
                    const double xs = cell_xs<D>(csl[s], xv);
                    F[0] += L.gref_d * xs;
                    const double feqn = bgk_devf(R.en, L.xf), feq = bgk_devf(R.e, L.xf);
                    const double fn = fcg + R.cm * (feqn - fcg);
                    const double g = fn + R.C * (R.ai * (feq - fn) + R.ai * (feqn - fn) - R.dtv * F[0]);
                    og = g + R.Cs * (feq - g);
                } else {
                    const double feqn = bgk_eq<double>(R.en, xv), feq = bgk_eq<double>(R.e, xv);
                    // cdugks.cpp:325-332 as written
                    const double fn = R.cm * feqn + (1.0 - R.cm) * fcg;
                    og = (1.0 - R.Cs) * (R.C * (fn + hdt * (feq * R.itau + (feqn - fn) * R.itau) - R.dtv * F[0])) +
                         R.Cs * feq;
                }
            }
            if (L.ok) {
                const size_t o = size_t(c) * nvp + v;
                if constexpr (XS && sizeof(S) == 4) {
                    // FP32E: the deviation (this kernel's arithmetic form),
                    // stochastically rounded at bit 31
                    float wd;
                    unsigned char e;
                    const unsigned rg = rs + unsigned(j) * 0x9e3779b9u;
                    x32_encode_sr(og, rg, wd, e);
                    fout_g[o] = wd;
                    a.xout_g[o] = e;
                    if (SH) {
                        x32_encode_sr(oh, rot16(rg), wd, e);
                        fout_h[o] = wd;
                        a.xout_h[o] = e;
                    }
                } else if constexpr (XS) {
                    // deviation from f_ref, stochastically rounded at bit 31
                    // (random bits from the entry and the step)
                    float wd;
                    unsigned char e;
                    const unsigned rg = mix32(unsigned(o) ^ seed);
                    x32_encode_sr(og - ref[0], rg, wd, e);
                    ((float *)fout_g)[o] = wd;
                    a.xout_g[o] = e;
                    if (SH) {
                        x32_encode_sr(oh - ref[1], mix32(rg), wd, e);
                        ((float *)fout_h)[o] = wd;
                        a.xout_h[o] = e;
                    }
                } else if constexpr (sizeof(S) == 4) {
                    // FP32: the stored deviation stochastically rounded
                    const unsigned rg = rs + unsigned(j) * 0x9e3779b9u;
                    fout_g[o] = f32_sr(og, rg);
                    if (SH) fout_h[o] = f32_sr(oh, rot16(rg));
                } else {
                    fout_g[o] = S(og);
                    if (SH) fout_h[o] = S(oh);
                }
            } else {
#pragma unroll
                for (int i = 0; i < D; ++i) qv[i] = 0.0;
            }
            if (SH)
#pragma unroll
                for (int jj = 0; jj < UPD_QJ; ++jj)
                    if (jj == j)
#pragma unroll
                        for (int i = 0; i < D; ++i) qreg[jj][i] += qv[i];
        }
        __syncthreads();
    }
    if (SH)
#pragma unroll
        for (int j = 0; j < UPD_QJ; ++j) {
            // sum over the VT lanes of each owned cell (all lanes take part)
            const int s = slot + j * NSLOT;
            double qv[D];
#pragma unroll
            for (int i = 0; i < D; ++i) qv[i] = qreg[j][i];
#pragma unroll
            for (int o = VT / 2; o > 0; o >>= 1)
#pragma unroll
                for (int i = 0; i < D; ++i) qv[i] += __shfl_xor(qv[i], o, 64);
            if (s < nown && vl == 0) {
                double *dst = a.part_c + (size_t(items[s].cell) * a.s.ntile_c + chunk) * 4;
#pragma unroll
                for (int i = 0; i < 4; ++i) dst[i] = i < D ? qv[i] : 0.0;
            }
        }
}
template <int MODEL, class S, int VT, bool TAB, int D, bool LIM, bool XS = false>
__global__ __launch_bounds__(BLOCKP, 4) void k_patch_update(KArgs a) {
    d_patch_update<MODEL, S, VT, TAB, D, LIM, XS>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <int MODEL, class S, int VT, bool TAB, int D, bool LIM, bool XS = false>
__global__ __launch_bounds__(BLOCKP, 4) void kb_patch_update(const KArgs *__restrict__ tab,
                                                              const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_patch_update<MODEL, S, VT, TAB, D, LIM, XS>(batch_args(tab, m.x), m.y);
}
