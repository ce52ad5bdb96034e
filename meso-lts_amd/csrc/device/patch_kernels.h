// LDS-staged patch kernels of the 2D step (included by step.hip inside its
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
// f - f_ref. The equilibria of FP32 storage are evaluated in FP32 in
// deviation form (their weight in one step is dt/tau or C_s).
//
// Upwind dead-band (|xi.n| < 1e-30): the moment kernel forms meso's average of
// both sides; the update kernel gives the face no flux, since the flux term is
// (xi.n) A f with |xi.n| < 1e-30.
#pragma once

constexpr int BLOCKP = 512;  // threads of a patch block

// Per-cell constants of the update (written by k_cell_prep for interior cells).
template <class S>
struct CellRec;
template <>
struct CellRec<float> {
    EqF en, e;                        // equilibria at t_n and t_{n+1}, deviation form
    double cm, C, ai, an, Cs, dtv;    // see k_patch_update
    double u0, u1;                    // u^{n+1}
};
template <>
struct CellRec<double> {
    EqParams en, e;
    double cm, C, itau, itaun, Cs, dtv;
    double u0, u1;
};

// ------------------------------------------------------------------- prep
// Per cell of the update patches: conservative macro update (as k_cell_update
// does for its cells), the constants of the f update, the cell's axis tables.
template <int MODEL, class S, bool TAB>
__global__ __launch_bounds__(BLOCK) void k_cell_prep(KArgs a) {
    if (*a.halt) return;
    constexpr int WPB = BLOCK / 64;
    const int lane = threadIdx.x % 64;
    const int li = __builtin_amdgcn_readfirstlane(blockIdx.x * WPB + int(threadIdx.x / 64));
    if (li >= a.n_int_cells) return;
    const int c = a.cell_list[li];
    CellStep<MODEL, 2> cs;
    const bool ok = cell_macros<MODEL, 2, S>(a, c, cs);
    if (lane == 0) {
        a.rho_out[c] = cs.rho;
        for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = cs.u[i];
        if (MODEL == CDUGKS_SHAKHOV) a.T_out[c] = cs.T;
        for (int k = 0; k < 5; ++k) a.fm[5 * c + k] = 0.0;
        a.fm[5 * c] = cs.phi[0];
        for (int i = 0; i < 2; ++i) a.fm[5 * c + 1 + i] = cs.phi[1 + i];
        if (MODEL == CDUGKS_SHAKHOV) a.fm[5 * c + 4] = cs.phi[3];
        if (!ok) {
            atomicAdd(a.bad_count, 1);
            atomicMin(a.first_bad, c);
            atomicMin(a.bad_step, a.step_index);
        }
        CellRec<S> r;
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
        r.u0 = cs.u[0];
        r.u1 = cs.u[1];
        reinterpret_cast<CellRec<S> *>(a.crec)[c] = r;
    }
    if (TAB) {
        S *dst = reinterpret_cast<S *>(a.cell_tab) + size_t(c) * 2 * a.nt;
        for (int t = lane; t < a.nt; t += 64) {
            dst[t] = tab_entry<S>(a, cs.en, t);
            dst[a.nt + t] = tab_entry<S>(a, cs.e, t);
        }
    }
}

// LSQ gradient (gx, gy) of one item from the staged rows of distribution d:
// sum_j a_j (f_j - f_c) (mesh_numerical.cpp:24-49).
template <class S, int VT>
__device__ __forceinline__ void item_grad(const meso_patch::Item *it, const S *val_d, int vl, double fc, double &gx,
                                          double &gy) {
    gx = gy = 0.0;
#pragma unroll
    for (int j = 0; j < meso_patch::MAXN; ++j) {
        const double df = double(val_d[it->nb[j] * VT + vl]) - fc;
        gx += it->a[j][0] * df;
        gy += it->a[j][1] * df;
    }
}

// Block prologue: copy n records of R bytes (a multiple of 16) to LDS.
template <class R>
__device__ __forceinline__ void stage_copy(const R *src, int n, R *dst) {
    const int4 *si = reinterpret_cast<const int4 *>(src);
    int4 *di = reinterpret_cast<int4 *>(dst);
    for (int r = threadIdx.x; r < n * int(sizeof(R) / 16); r += BLOCKP) di[r] = si[r];
}

// Patch records: with one item per wave (IPW == 1) every record a wave reads
// is wave-uniform and is read from HBM through the scalar cache; with several
// items per wave the records are staged in LDS.
template <int IPW>
struct Records {
    const meso_patch::Item *items;
    const meso_patch::Face *faces;
    __device__ __forceinline__ const meso_patch::Item *item(int s) const { return items + s; }
    __device__ __forceinline__ const meso_patch::Face *face(int lf) const { return faces + lf; }
};

// Bytes of LDS the records take (0 when they are read from HBM).
template <int IPW>
__host__ __device__ constexpr size_t record_bytes(int maxi, int maxf) {
    return IPW > 1 ? size_t(maxi) * sizeof(meso_patch::Item) + size_t(maxf) * sizeof(meso_patch::Face) : 0;
}

// Stage the rows of one velocity tile: all loads of a batch are issued before
// their LDS stores.
template <int ND, class S, int VT>
__device__ __forceinline__ void stage_values(const int *rows, int nr, const S *fin_g, const S *fin_h, int nvp, int v,
                                             bool ok, int slot, int vl, S *val, int maxr) {
    constexpr int NSLOT = (BLOCKP / 64) * (64 / VT);
    constexpr int B = 8;
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
    double x, y, w;   // abscissae and weight (FP64)
    float xf[3];      // FP32 abscissae (z = 0)
    int vt;           // packed axis-table offsets
    float gref;       // FP32 g_ref
    double gref_d, href_d;
    bool ok;          // v < nv
};

__device__ __forceinline__ LaneVel lane_vel(const KArgs &a, int v) {
    LaneVel L;
    L.ok = v < a.s.nv;
    const int vv = L.ok ? v : 0;
    const double4 X = a.vel_d[vv];
    L.x = X.x;
    L.y = X.y;
    L.w = L.ok ? X.w : 0.0;
    const float4 xf = a.vel_f[vv];
    L.xf[0] = xf.x;
    L.xf[1] = xf.y;
    L.xf[2] = 0.0f;
    L.vt = a.nt > 0 ? a.vtab[vv] : 0;
    L.gref = a.fref_f2[vv].x;
    const double2 fr = a.fref_d2[vv];
    L.gref_d = fr.x;
    L.href_d = fr.y;
    return L;
}

// Equilibrium pair of face f at the lane velocity, in the storage form
// (deviation for FP32, full value for FP64), as FP64.
template <int MODEL, class S, bool TAB>
__device__ __forceinline__ void face_eq_lane(const KArgs &a, int f, const LaneVel &L, double &gs, double &hs) {
    if constexpr (sizeof(S) == 4) {
        const EqF &e = a.face_eqf[f];
        float g, h = 0.0f;
        if (MODEL == CDUGKS_SHAKHOV) {
            if (TAB)
                shakhov_tab_f(e, reinterpret_cast<const float *>(a.face_tab) + size_t(f) * a.nt, L.vt, L.xf, 2, a.s.K,
                              L.gref, g, h);
            else
                shakhov_devf(e, L.xf, 2, a.s.K, L.gref, g, h);
        } else {
            g = bgk_devf(e, L.xf);
        }
        gs = g;
        hs = h;
    } else {
        const EqParams &e = a.face_eq[f];
        const double x[3] = {L.x, L.y, 0.0};
        if (MODEL == CDUGKS_SHAKHOV) {
            if (TAB)
                shakhov_tab_d(e, reinterpret_cast<const double *>(a.face_tab) + size_t(f) * a.nt, L.vt, x, 2, a.s.K,
                              gs, hs);
            else
                shakhov_eq<double>(e, x, 2, a.s.K, gs, hs);
        } else {
            gs = bgk_eq<double>(e, x);
            hs = 0.0;
        }
    }
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

// Bytes of LDS the face and cell records of k_patch_update take with RL.
template <class S>
__host__ __device__ constexpr size_t update_rec_bytes(int maxf, int maxo) {
    return size_t(maxf) * (sizeof(typename UpdateRecs<S>::Eq) + 8) + size_t(maxo) * (sizeof(CellRec<S>) + 8);
}

// face_eq_lane with the face's equilibrium record e at hand.
template <int MODEL, class S, bool TAB>
__device__ __forceinline__ void face_eq_rec(const KArgs &a, const typename UpdateRecs<S>::Eq &e, int f,
                                            const LaneVel &L, double &gs, double &hs) {
    if constexpr (sizeof(S) == 4) {
        float g, h = 0.0f;
        if (MODEL == CDUGKS_SHAKHOV) {
            if (TAB)
                shakhov_tab_f(e, reinterpret_cast<const float *>(a.face_tab) + size_t(f) * a.nt, L.vt, L.xf, 2, a.s.K,
                              L.gref, g, h);
            else
                shakhov_devf(e, L.xf, 2, a.s.K, L.gref, g, h);
        } else {
            g = bgk_devf(e, L.xf);
        }
        gs = g;
        hs = h;
    } else {
        const double x[3] = {L.x, L.y, 0.0};
        if (MODEL == CDUGKS_SHAKHOV) {
            if (TAB)
                shakhov_tab_d(e, reinterpret_cast<const double *>(a.face_tab) + size_t(f) * a.nt, L.vt, x, 2, a.s.K,
                              gs, hs);
            else
                shakhov_eq<double>(e, x, 2, a.s.K, gs, hs);
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
template <int MODEL, class S>
__device__ __forceinline__ void bnd_value_lane(const KArgs &a, int f, double sn, const LaneVel &L, double gm,
                                               double hm, double &g, double &h) {
    g = gm;
    h = hm;
    const int bi = a.face_bnd[f];
    if (bi < 0) return;
    const BndParams &P = a.bnd[bi];
    const double kn = -sn;  // xi . n1, n1 = -n0 points into the domain
    if (P.kind == BC_INTERIOR || kn < -VSMALL) return;
    const double x[3] = {L.x, L.y, 0.0};
    double gi, hi = 0.0;
    if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(P.in, x, 2, a.s.K, gi, hi);
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
// chunk into part_f[face][chunk][col] (the layout k_face_eq reads).
template <int MODEL, class S, int VT>
__global__ __launch_bounds__(BLOCKP, 4) void k_patch_moments(KArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (*a.bad_count > 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *a.halt = 1;
        return;
    }
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    constexpr int ND = SH ? 2 : 1;
    constexpr int NCOL = ncol_face<MODEL, 2>();
    constexpr int IPW = 64 / VT;
    constexpr int NW = BLOCKP / 64;
    constexpr int NSLOT = NW * IPW;
    constexpr int VTP = VT + 1;  // odd row stride: conflict-free reads with lane = face
    const int np = a.n_mp;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int chunk = b / np, p = b - chunk * np;
    const int i0 = a.mp_item_ptr[p], ni = a.mp_item_ptr[p + 1] - i0;
    const int f0 = a.mp_face_ptr[p], nfp = a.mp_face_ptr[p + 1] - f0;
    const int r0 = a.mp_row_ptr[p], nr = a.mp_row_ptr[p + 1] - r0;
    const int maxi = a.mp_maxi, maxf = a.mp_maxf, maxr = a.mp_maxr;
    // LDS: [items | faces] | rows | tile velocities | values | f-bar+ rows
    unsigned char *q = lds_raw;
    Records<IPW> rec;
    if constexpr (IPW > 1) {
        auto *li = reinterpret_cast<meso_patch::Item *>(q);
        auto *lf = reinterpret_cast<meso_patch::Face *>(li + maxi);
        stage_copy(a.mp_items + i0, ni, li);
        stage_copy(a.mp_faces + f0, nfp, lf);
        rec.items = li;
        rec.faces = lf;
        q += record_bytes<IPW>(maxi, maxf);
    } else {
        rec.items = a.mp_items + i0;
        rec.faces = a.mp_faces + f0;
    }
    int *rows = reinterpret_cast<int *>(q);
    q += (size_t(maxr) * 4 + 31) & ~size_t(31);
    double4 *vtile = reinterpret_cast<double4 *>(q);
    S *val = reinterpret_cast<S *>(vtile + VT);
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
    const int FG = maxf > 32 ? 64 : 32;
    const int my_face = threadIdx.x % FG, vs = threadIdx.x / FG, nvs = BLOCKP / FG;
    const bool face_act = my_face < nfp;
    double acc[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = 0.0;
    const int ntiles = (a.s.nv + VT - 1) / VT;
    const int t0 = chunk * (a.p_vchunk / VT), t1 = min(t0 + a.p_vchunk / VT, ntiles);
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int v = t * VT + vl;
        const bool ok = v < a.s.nv;
        const double4 X = a.vel_d[ok ? v : 0];
        // A: stage values and tile velocities, clear the face rows
        stage_values<ND, S, VT>(rows, nr, fin_g, fin_h, nvp, v, ok, slot, vl, val, maxr);
        if (threadIdx.x < VT) vtile[threadIdx.x] = ok ? X : make_double4(0.0, 0.0, 0.0, 0.0);
        for (int r = threadIdx.x; r < nfp * VTP; r += BLOCKP) {
            fb[r] = S(0);
            if (SH) fb[maxf * VTP + r] = S(0);
        }
        __syncthreads();
        // B: gradients and f-bar+ of the patch faces
        for (int s = slot; s < ni; s += NSLOT) {
            const meso_patch::Item *it = rec.item(s);
            const int nface = it->nface;
            double fc[2], gx[2], gy[2];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                fc[d] = double(val[(d * maxr + s) * VT + vl]);
                gx[d] = gy[d] = 0.0;
                if (a.s.grad) item_grad<S, VT>(it, val + d * maxr * VT, vl, fc[d], gx[d], gy[d]);
            }
#pragma unroll
            for (int k = 0; k < meso_patch::MAXF; ++k) {
                if (k >= nface) continue;
                const int e = it->face[k];
                const int lf = e & 0xffff, mode = e >> 16;
                const meso_patch::Face *F = rec.face(lf);
                const double sn = F->n[0] * X.x + F->n[1] * X.y;  // + n_z xi_z, both 0 in 2D
                const bool up = mode == meso_patch::SIDE_BOUNDARY ||
                                (mode == meso_patch::SIDE_OWNER ? sn >= VSMALL : sn < -VSMALL);
                const bool dead = !(sn >= VSMALL) && !(sn < -VSMALL);
                const int side = mode == meso_patch::SIDE_NEIGHBOUR ? 1 : 0;
                const double ex = F->d[side][0] - X.x * hdt, ey = F->d[side][1] - X.y * hdt;
                if (!ok) continue;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double Xd = ex * gx[d] + ey * gy[d];
                    S *dst = &fb[(d * maxf + lf) * VTP + vl];
                    if (mode == meso_patch::SIDE_BOUNDARY) {
                        *dst = S(dead ? 0.5 * (fc[d] + fc[d] + Xd + Xd) : fc[d] + Xd);
                    } else if (up) {
                        *dst = S(fc[d] + Xd);
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
                const double4 Y = vtile[vv];
                const double wg = Y.w * double(fb[my_face * VTP + vv]);
                const double wgx = wg * Y.x, wgy = wg * Y.y;
                acc[Col<2>::M0] += wg;
                acc[Col<2>::M1] += wgx;
                acc[Col<2>::M1 + 1] += wgy;
                acc[Col<2>::M2] += wgx * Y.x;
                acc[Col<2>::M2 + 1] += wgx * Y.y;
                acc[Col<2>::M2 + 2] += wgy * Y.y;
                if (SH) {
                    const double wh = Y.w * double(fb[(maxf + my_face) * VTP + vv]);
                    const double wgk = wg * (Y.x * Y.x + Y.y * Y.y + Y.z * Y.z);
                    acc[Col<2>::M3] += wgk * Y.x;
                    acc[Col<2>::M3 + 1] += wgk * Y.y;
                    acc[Col<2>::H0] += wh;
                    acc[Col<2>::H1] += wh * Y.x;
                    acc[Col<2>::H1 + 1] += wh * Y.y;
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
        const int f = a.mp_faces[f0 + lf].face;
        a.part_f[(size_t(f) * a.s.ntile_f + chunk) * NCOL + k] = sum;
    }
}

// ------------------------------------------------------------- update
// Cells of the update patches: cell flux, f-tilde-plus update and the Shakhov
// heat-flux partials part_c[cell][chunk][0..1]. RL: the equilibrium records of
// the patch faces and the update constants of the owned cells are copied to
// LDS once per block; otherwise every tile re-reads them per lane from HBM
// (L2), which costs address arithmetic and L2 latency in every tile.
template <int MODEL, class S, int VT, bool TAB, bool RL>
__global__ __launch_bounds__(BLOCKP) void k_patch_update(KArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (*a.halt) return;
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    constexpr int ND = SH ? 2 : 1;
    constexpr int IPW = 64 / VT;
    constexpr int NW = BLOCKP / 64;
    constexpr int NSLOT = NW * IPW;
    const int np = a.n_up;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int chunk = b / np, p = b - chunk * np;
    const int i0 = a.up_item_ptr[p], ni = a.up_item_ptr[p + 1] - i0;
    const int f0 = a.up_face_ptr[p], nfp = a.up_face_ptr[p + 1] - f0;
    const int r0 = a.up_row_ptr[p], nr = a.up_row_ptr[p + 1] - r0;
    const int nown = a.up_nown[p];
    const int maxi = a.up_maxi, maxf = a.up_maxf, maxr = a.up_maxr;
    // LDS: items | faces | rows | values | face fluxes | q partials
    auto *items = reinterpret_cast<meso_patch::Item *>(lds_raw);
    auto *faces = reinterpret_cast<meso_patch::Face *>(items + maxi);
    int *rows = reinterpret_cast<int *>(faces + maxf);
    S *val = reinterpret_cast<S *>(lds_raw + ((reinterpret_cast<unsigned char *>(rows + maxr) - lds_raw + 15) & ~15));
    S *flx = val + size_t(ND) * maxr * VT;
    double *qacc = reinterpret_cast<double *>(
        lds_raw + ((reinterpret_cast<unsigned char *>(flx + size_t(ND) * maxf * VT) - lds_raw + 15) & ~15));
    // with RL: the face equilibria and mix weights of the patch faces and the
    // update constants of the owned cells, copied once per block
    using EqS = typename UpdateRecs<S>::Eq;
    EqS *feq = reinterpret_cast<EqS *>(qacc + 2 * a.up_maxo);
    double *fmix = reinterpret_cast<double *>(feq + maxf);
    CellRec<S> *crl = reinterpret_cast<CellRec<S> *>(fmix + maxf);
    float2 *csl = reinterpret_cast<float2 *>(crl + a.up_maxo);
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int sub = lane / VT, vl = lane % VT;
    const int slot = wave * IPW + sub;
    const int nvp = a.s.nvp;
    const double hdt = a.s.half_dt;
    const S *fin_g = (const S *)a.fin_g, *fin_h = (const S *)a.fin_h;
    S *fout_g = (S *)a.fout_g, *fout_h = (S *)a.fout_h;
    stage_copy(a.up_items + i0, ni, items);
    stage_copy(a.up_faces + f0, nfp, faces);
    for (int r = threadIdx.x; r < nr; r += BLOCKP) rows[r] = a.up_rows[r0 + r];
    for (int r = threadIdx.x; r < 2 * nown; r += BLOCKP) qacc[r] = 0.0;
    if constexpr (RL) {
        for (int lf = threadIdx.x; lf < nfp; lf += BLOCKP) {
            const int f = a.up_faces[f0 + lf].face;
            feq[lf] = UpdateRecs<S>::face_eq(a, f);
            fmix[lf] = a.face_eq[f].mix;
        }
        for (int r = threadIdx.x; r < nown; r += BLOCKP) {
            const int c = a.up_items[i0 + r].cell;
            crl[r] = reinterpret_cast<const CellRec<S> *>(a.crec)[c];
            csl[r] = make_float2(a.cell_S[3 * c], a.cell_S[3 * c + 1]);
        }
    }
    const int ntiles = (a.s.nv + VT - 1) / VT;
    const int t0 = chunk * (a.p_vchunk / VT), t1 = min(t0 + a.p_vchunk / VT, ntiles);
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int v = t * VT + vl;
        const LaneVel L = lane_vel(a, v);
        // A: stage values
        stage_values<ND, S, VT>(rows, nr, fin_g, fin_h, nvp, v, L.ok, slot, vl, val, maxr);
        __syncthreads();
        // B: gradients; flux (xi.n) A f_face of each face on its upwind side's lanes
        for (int s = slot; s < ni; s += NSLOT) {
            const meso_patch::Item *it = items + s;
            const int nface = it->nface;
            double fc[2], gx[2], gy[2];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                fc[d] = double(val[(d * maxr + s) * VT + vl]);
                gx[d] = gy[d] = 0.0;
                if (a.s.grad) item_grad<S, VT>(it, val + d * maxr * VT, vl, fc[d], gx[d], gy[d]);
            }
#pragma unroll
            for (int k = 0; k < meso_patch::MAXF; ++k) {
                if (k >= nface) continue;
                const int e = it->face[k];
                const int lf = e & 0xffff, mode = e >> 16;
                const meso_patch::Face *F = faces + lf;
                const double sn = F->n[0] * L.x + F->n[1] * L.y;
                const bool bnd = mode == meso_patch::SIDE_BOUNDARY;
                const bool own = mode == meso_patch::SIDE_OWNER;
                const bool dead = !(sn >= VSMALL) && !(sn < -VSMALL);
                const bool up = bnd || (own ? sn >= VSMALL : sn < -VSMALL);
                if (!__any(up || (own && dead))) continue;
                if (up) {
                    const int f = F->face;
                    double gs, hs = 0.0;
                    double mix;
                    if constexpr (RL) {
                        face_eq_rec<MODEL, S, TAB>(a, feq[lf], f, L, gs, hs);
                        mix = fmix[lf];
                    } else {
                        face_eq_lane<MODEL, S, TAB>(a, f, L, gs, hs);
                        mix = a.face_eq[f].mix;
                    }
                    const int side = mode == meso_patch::SIDE_NEIGHBOUR ? 1 : 0;
                    const double ex = F->d[side][0] - L.x * hdt, ey = F->d[side][1] - L.y * hdt;
                    const double knA = sn * F->area;
                    double fm[2] = {0.0, 0.0};
#pragma unroll
                    for (int d = 0; d < ND; ++d) {
                        const double Xd = ex * gx[d] + ey * gy[d];
                        // a boundary face's dead-band value 0.5 (f + f + X + X) is f + X
                        const double fbar = fc[d] + Xd;
                        const double eq = d == 0 ? gs : hs;
                        if constexpr (sizeof(S) == 8) fm[d] = (1.0 - mix) * fbar + mix * eq;
                        else fm[d] = fbar + mix * (eq - fbar);
                    }
                    if (bnd) bnd_value_lane<MODEL, S>(a, F->face, sn, L, fm[0], fm[1], fm[0], fm[1]);
#pragma unroll
                    for (int d = 0; d < ND; ++d) flx[(d * maxf + lf) * VT + vl] = S(knA * fm[d]);
                } else if (own && dead) {
#pragma unroll
                    for (int d = 0; d < ND; ++d) flx[(d * maxf + lf) * VT + vl] = S(0);
                }
            }
        }
        __syncthreads();
        // C: owned cells
        for (int s = slot; s < nown; s += NSLOT) {
            const meso_patch::Item *it = items + s;
            const int c = it->cell;
            const int nface = it->nface;
            double F[2] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < meso_patch::MAXF; ++k) {
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
            const CellRec<S> &R = RL ? crl[s] : reinterpret_cast<const CellRec<S> *>(a.crec)[c];
            const double fcg = double(val[s * VT + vl]);
            const double fch = SH ? double(val[(maxr + s) * VT + vl]) : 0.0;
            double og, oh = 0.0;
            if (SH) {
                double gsn, hsn, gs, hs;
                if constexpr (sizeof(S) == 4) {
                    // f_ref flux through the cell: f_ref xi . sum(+-A n), non-zero only for unclosed cells
                    const float2 cS = RL ? csl[s] : make_float2(a.cell_S[3 * c], a.cell_S[3 * c + 1]);
                    const double xs = L.x * double(cS.x) + L.y * double(cS.y);
                    F[0] += L.gref_d * xs;
                    F[1] += L.href_d * xs;
                    float g1, h1, g2, h2;
                    if (TAB) {
                        const float *tb = reinterpret_cast<const float *>(a.cell_tab) + size_t(c) * 2 * a.nt;
                        shakhov_tab_f(R.en, tb, L.vt, L.xf, 2, a.s.K, L.gref, g1, h1);
                        shakhov_tab_f(R.e, tb + a.nt, L.vt, L.xf, 2, a.s.K, L.gref, g2, h2);
                    } else {
                        shakhov_devf(R.en, L.xf, 2, a.s.K, L.gref, g1, h1);
                        shakhov_devf(R.e, L.xf, 2, a.s.K, L.gref, g2, h2);
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
                    const double c0 = L.x - R.u0, c1 = L.y - R.u1;
                    const double tq = 0.5 * L.w * ((c0 * c0 + c1 * c1) * gf + hf);
                    F[0] = tq * c0;  // reuse as q partials
                    F[1] = tq * c1;
                } else {
                    const double x[3] = {L.x, L.y, 0.0};
                    if (TAB) {
                        const double *tb = reinterpret_cast<const double *>(a.cell_tab) + size_t(c) * 2 * a.nt;
                        shakhov_tab_d(R.en, tb, L.vt, x, 2, a.s.K, gsn, hsn);
                        shakhov_tab_d(R.e, tb + a.nt, L.vt, x, 2, a.s.K, gs, hs);
                    } else {
                        shakhov_eq<double>(R.en, x, 2, a.s.K, gsn, hsn);
                        shakhov_eq<double>(R.e, x, 2, a.s.K, gs, hs);
                    }
                    // cdugks_shakhov.cpp:679-695 as written
                    const double gn = R.cm * gsn + (1.0 - R.cm) * fcg;
                    const double hn = R.cm * hsn + (1.0 - R.cm) * fch;
                    const double g = R.C * (gn + hdt * (gs * R.itau + (gsn - gn) * R.itaun) - R.dtv * F[0]);
                    const double h = R.C * (hn + hdt * (hs * R.itau + (hsn - hn) * R.itaun) - R.dtv * F[1]);
                    og = (1.0 - R.Cs) * g + R.Cs * gs;
                    oh = (1.0 - R.Cs) * h + R.Cs * hs;
                    const double c0 = L.x - R.u0, c1 = L.y - R.u1;
                    const double tq = 0.5 * L.w * ((c0 * c0 + c1 * c1) * g + h);
                    F[0] = tq * c0;
                    F[1] = tq * c1;
                }
            } else {
                if constexpr (sizeof(S) == 4) {
                    const float2 cS = RL ? csl[s] : make_float2(a.cell_S[3 * c], a.cell_S[3 * c + 1]);
                    const double xs = L.x * double(cS.x) + L.y * double(cS.y);
                    F[0] += L.gref_d * xs;
                    const double feqn = bgk_devf(R.en, L.xf), feq = bgk_devf(R.e, L.xf);
                    const double fn = fcg + R.cm * (feqn - fcg);
                    const double g = fn + R.C * (R.ai * (feq - fn) + R.ai * (feqn - fn) - R.dtv * F[0]);
                    og = g + R.Cs * (feq - g);
                } else {
                    const double x[3] = {L.x, L.y, 0.0};
                    const double feqn = bgk_eq<double>(R.en, x), feq = bgk_eq<double>(R.e, x);
                    // cdugks.cpp:325-332 as written
                    const double fn = R.cm * feqn + (1.0 - R.cm) * fcg;
                    og = (1.0 - R.Cs) * (R.C * (fn + hdt * (feq * R.itau + (feqn - fn) * R.itau) - R.dtv * F[0])) +
                         R.Cs * feq;
                }
                F[0] = F[1] = 0.0;
            }
            if (L.ok) {
                fout_g[size_t(c) * nvp + v] = S(og);
                if (SH) fout_h[size_t(c) * nvp + v] = S(oh);
            } else {
                F[0] = F[1] = 0.0;
            }
            if (SH) {
                // sum over the VT lanes of this item
#pragma unroll
                for (int o = VT / 2; o > 0; o >>= 1) {
                    F[0] += __shfl_xor(F[0], o, 64);
                    F[1] += __shfl_xor(F[1], o, 64);
                }
                if (vl == 0) {
                    qacc[2 * s] += F[0];
                    qacc[2 * s + 1] += F[1];
                }
            }
        }
        __syncthreads();
    }
    if (SH)
        for (int r = threadIdx.x; r < nown; r += BLOCKP) {
            const int c = items[r].cell;
            double *dst = a.part_c + (size_t(c) * a.s.ntile_c + chunk) * 4;
            dst[0] = qacc[2 * r];
            dst[1] = qacc[2 * r + 1];
            dst[2] = 0.0;
            dst[3] = 0.0;
        }
}
