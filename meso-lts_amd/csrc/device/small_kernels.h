// Step kernels for small velocity sets (NV <= 64, FP64 storage, CDUGKS models,
// no limiter, no mirror boundaries): two launches per step.
//
//   k_small_faces  one LP-lane segment per face, one lane per velocity:
//                  f-bar+ from the upwind cell's LSQ gradient, the raw
//                  moments summed across the segment, the face macros and
//                  equilibrium, the mixed face value, the boundary value
//                  (wall density sums across the segment), the flux moments
//                  of the face (cdugks_shakhov.cpp:213-607, cdugks.cpp:123-246)
//                  and the per-velocity face flux (xi . n0) f_face
//   k_small_cells  one LP-lane segment per cell: conservative macro update from
//                  the face flux moments, the cell flux from the per-velocity
//                  face fluxes, the f-tilde update and the Shakhov heat flux
//                  summed across the segment (cdugks_shakhov.cpp:610-722,
//                  cdugks.cpp:248-335)
//
// A velocity set this small leaves most lanes of a one-wave-per-face layout
// idle, and the per-stage kernels of the large sets are each a chain of
// dependent loads that takes several microseconds: here a face or a cell
// takes LP lanes (16, 32 or 64), each step is two chains, and the device step
// counter is advanced by the face kernel (block 0), which reads nothing that
// depends on it, so no third launch is needed.
#pragma once

template <int LP>
__host__ __device__ constexpr int small_spb() {
    return BLOCK / LP;  // faces or cells per block
}

// Bad-cell flag of the cell kernel: the face kernel of the same step has
// already advanced the step counter, hence the step index is one less.
template <int MODEL>
__device__ __forceinline__ void small_flag_bad(const KArgs &a, int c) {
    flag_bad<MODEL>(a, c, 1);
}

template <int MODEL, int D, int LP>
__device__ __forceinline__ void d_small_faces(const KArgs &a, const int b) {
    if (b == 0 && threadIdx.x == 0) *a.step_ctr += 1;
    if (*a.bad_count > 0) {
        if (b == 0 && threadIdx.x == 0) *a.halt = 1;
        return;
    }
    constexpr int SPB = small_spb<LP>();
    constexpr int MAXN = D == 2 ? 4 : 6;
    constexpr int GS = 1 + MAXN;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr int NFL = D + 2;
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    const int f = b * SPB + int(threadIdx.x) / LP;
    if (f >= a.s.nf) return;  // whole segments leave together
    const int v = int(threadIdx.x) % LP;
    const bool act = v < a.s.nv;
    const int vv = act ? v : 0;
    const int nvp = a.s.nvp;
    const double *fg = (const double *)a.fin_g;
    const double *fh = (const double *)a.fin_h;
    const int *ids = a.fs_id + size_t(f) * 2 * GS;
    const double *av = a.fs_a + size_t(f) * 2 * MAXN * 3;
    double vg[2 * GS], vh[2 * GS];
#pragma unroll
    for (int k = 0; k < 2 * GS; ++k) {
        const size_t r = size_t(ids[k]) * nvp + vv;
        vg[k] = fg[r];
        if (SH) vh[k] = fh[r];
    }
    double d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = a.face_d[6 * f + k];
    const int bi = a.face_bnd[f];
    const double4 X = a.vel_d[vv];
    const double x[3] = {X.x, X.y, X.z};
    const double wv = act ? X.w : 0.0;
    const double hd = a.s.half_dt;
    const double xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
    double sn;
    const int up = upwind<double>(a, f, vv, x, sn);
    // f-bar+: the upwind side's value shifted by its gradient; in the dead-band
    // the average of both sides (cdugks_shakhov.cpp:227-254)
    double go_g[3] = {0.0, 0.0, 0.0}, go_h[3] = {0.0, 0.0, 0.0};
    double gn_g[3] = {0.0, 0.0, 0.0}, gn_h[3] = {0.0, 0.0, 0.0};
    if (a.s.grad) {
        if (up >= 0) {
            group_grad<D, MAXN, double>(vg, av, go_g);
            if (SH) group_grad<D, MAXN, double>(vh, av, go_h);
        }
        if (up <= 0) {
            group_grad<D, MAXN, double>(vg + GS, av + 3 * MAXN, gn_g);
            if (SH) group_grad<D, MAXN, double>(vh + GS, av + 3 * MAXN, gn_h);
        }
    }
    const double Xog = dot_ext<D, double>(d[0], d[1], d[2], xh, go_g);
    const double Xng = dot_ext<D, double>(d[3], d[4], d[5], xh, gn_g);
    const double gb = up > 0 ? vg[0] + Xog : (up < 0 ? vg[GS] + Xng : 0.5 * (vg[0] + vg[GS] + Xog + Xng));
    double hb = 0.0;
    if (SH) {
        const double Xoh = dot_ext<D, double>(d[0], d[1], d[2], xh, go_h);
        const double Xnh = dot_ext<D, double>(d[3], d[4], d[5], xh, gn_h);
        hb = up > 0 ? vh[0] + Xoh : (up < 0 ? vh[GS] + Xnh : 0.5 * (vh[0] + vh[GS] + Xoh + Xnh));
    }
    // raw moments of f-bar+ over the segment
    double M[NCOL];
    {
        const double wg = wv * gb;
        const double wgx = wg * x[0], wgy = wg * x[1], wgz = wg * x[2];
        M[Col<D>::M0] = wg;
        M[Col<D>::M1] = wgx;
        M[Col<D>::M1 + 1] = wgy;
        if (D == 3) M[Col<D>::M1 + 2] = wgz;
        if (D == 2) {
            M[Col<D>::M2] = wgx * x[0];
            M[Col<D>::M2 + 1] = wgx * x[1];
            M[Col<D>::M2 + 2] = wgy * x[1];
        } else {
            M[Col<D>::M2] = wgx * x[0];
            M[Col<D>::M2 + 1] = wgx * x[1];
            M[Col<D>::M2 + 2] = wgx * x[2];
            M[Col<D>::M2 + 3] = wgy * x[1];
            M[Col<D>::M2 + 4] = wgy * x[2];
            M[Col<D>::M2 + 5] = wgz * x[2];
        }
        if (SH) {
            const double wh = wv * hb;
            const double wgk = wg * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            M[Col<D>::M3] = wgk * x[0];
            M[Col<D>::M3 + 1] = wgk * x[1];
            if (D == 3) M[Col<D>::M3 + 2] = wgk * x[2];
            M[Col<D>::H0] = wh;
            M[Col<D>::H1] = wh * x[0];
            M[Col<D>::H1 + 1] = wh * x[1];
            if (D == 3) M[Col<D>::H1 + 2] = wh * x[2];
        }
    }
#pragma unroll
    for (int k = 0; k < NCOL; ++k) M[k] = seg_sum<LP>(M[k]);
    const double n[3] = {a.face_n[3 * f], a.face_n[3 * f + 1], a.face_n[3 * f + 2]};
    EqParams e;
    double Fb[NFL];
#pragma unroll
    for (int k = 0; k < NFL; ++k) Fb[k] = 0.0;
    face_macros<MODEL, D>(a, M, n, e, Fb);
    double gs, hs = 0.0;
    eq_store<MODEL, D, double>(a, e, x, vv, gs, hs);
    const double mix = e.mix;
    const double gm = lerp_store<double>(gb, gs, mix);
    const double hm = SH ? lerp_store<double>(hb, hs, mix) : 0.0;
    double gv = gm, hv = hm;
    double fl[NFL];
    const double wsn = wv * sn;
    if (bi < 0) {
        // flux moments of the equilibrium part; f-bar+'s own follow from M
        const double wg = wsn * gs;
        fl[0] = wg;
        fl[1] = wg * x[0];
        fl[2] = wg * x[1];
        if (D == 3) fl[3] = wg * x[2];
        fl[D + 1] = SH ? wg * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) + wsn * hs : 0.0;
#pragma unroll
        for (int k = 0; k < NFL; ++k) fl[k] = (1.0 - mix) * Fb[k] + mix * seg_sum<LP>(fl[k]);
        if (!SH) fl[D + 1] = 0.0;
    } else {
        const int c = a.face_cells[f].x;
        const meso_mark_desc &mk = mark_of(a, a.bnd_mark[bi]);
        BndParams P;
        EqParams e1;
        if (bnd_begin<MODEL, D>(a, mk, f, c, P, e1)) {
            double s[2] = {0.0, 0.0};
            if (act) wall_terms<MODEL, D>(a, e1, vv, -sn, gm, s);
            s[0] = seg_sum<LP>(s[0]);
            s[1] = seg_sum<LP>(s[1]);
            bnd_wall_end<MODEL, D>(a, mk, P, s);
        }
        bnd_value_full<MODEL, D, double, false>(a, P, f, c, vv, sn, e, gm, hm, gv, hv);
        const double wg = wsn * gv;
        fl[0] = wg;
        fl[1] = wg * x[0];
        fl[2] = wg * x[1];
        if (D == 3) fl[3] = wg * x[2];
        fl[D + 1] = SH ? wg * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) + wsn * hv : 0.0;
#pragma unroll
        for (int k = 0; k < NFL; ++k) fl[k] = seg_sum<LP>(fl[k]);
        if (v == 0) a.bnd[bi] = P;
    }
    if (act) {
        double *ph = a.sflux + size_t(f) * LP;
        ph[v] = sn * gv;
        if (SH) ph[size_t(a.s.nf) * LP + v] = sn * hv;
    }
    if (v == 0) {
#pragma unroll
        for (int k = 0; k < NFL; ++k) a.face_flux[size_t(f) * NFL + k] = fl[k];
    }
}
template <int MODEL, int D, int LP>
__global__ __launch_bounds__(BLOCK) void k_small_faces(KArgs a) {
    d_small_faces<MODEL, D, LP>(a, xcd_block(blockIdx.x, gridDim.x));
}

template <int MODEL, int D, int LP>
__device__ __forceinline__ void d_small_cells(const KArgs &a, const int b) {
    if (*a.halt) return;
    constexpr int SPB = small_spb<LP>();
    constexpr int MAXF = D == 2 ? 4 : 6;
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    const int c = b * SPB + int(threadIdx.x) / LP;
    if (c >= a.s.nc) return;  // whole segments leave together
    const int v = int(threadIdx.x) % LP;
    const bool act = v < a.s.nv;
    const int vv = act ? v : 0;
    const int nvp = a.s.nvp;
    const size_t o = size_t(c) * nvp + vv;
    const double fcg = ((const double *)a.fin_g)[o];
    const double fch = SH ? ((const double *)a.fin_h)[o] : 0.0;
    // cell flux sum over faces of +-A (xi . n0) f_face
    const int nfc = a.cs_nf[c];
    const double *phg = a.sflux;
    const double *phh = a.sflux + size_t(a.s.nf) * LP;
    double Fg = 0.0, Fh = 0.0;
#pragma unroll
    for (int k = 0; k < MAXF; ++k) {
        if (k < nfc) {
            const int idx = c * MAXF + k;
            const size_t fo = size_t(a.cs_face[idx]) * LP + vv;
            const double sA = a.cs_sA[idx];
            Fg += sA * phg[fo];
            if (SH) Fh += sA * phh[fo];
        }
    }
    CellStep<MODEL, D> cs;
    const bool ok = cell_macros<MODEL, D, double>(a, c, cs);
    if (v == 0) {
        a.rho_out[c] = cs.rho;
        for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = cs.u[i];
        if (SH) a.T_out[c] = cs.T;
        for (int k = 0; k < 5; ++k) a.fm[5 * c + k] = 0.0;
        a.fm[5 * c] = cs.phi[0];
        for (int i = 0; i < D; ++i) a.fm[5 * c + 1 + i] = cs.phi[1 + i];
        if (SH) a.fm[5 * c + 4] = cs.phi[D + 1];
        if (!ok) small_flag_bad<MODEL>(a, c);
    }
    const double4 X = a.vel_d[vv];
    const double x[3] = {X.x, X.y, X.z};
    const double hd = a.s.half_dt;
    const double dtv = cs.dt_v;
    double *og = (double *)a.fout_g;
    double *oh = (double *)a.fout_h;
    if (SH) {
        double gsn, hsn, gs, hs;
        eq_store<MODEL, D, double>(a, cs.en, x, vv, gsn, hsn);
        eq_store<MODEL, D, double>(a, cs.e, x, vv, gs, hs);
        // cdugks_shakhov.cpp:679-695 as written
        const double cm = cs.cm, C = cs.C, Cs = cs.Cs;
        const double itau = 1.0 / cs.tau, itaun = 1.0 / cs.taun;
        const double gn = cm * gsn + (1.0 - cm) * fcg;
        const double hn = cm * hsn + (1.0 - cm) * fch;
        const double g = C * (gn + hd * (gs * itau + (gsn - gn) * itaun) - dtv * Fg);
        const double h = C * (hn + hd * (hs * itau + (hsn - hn) * itaun) - dtv * Fh);
        const double c0 = x[0] - cs.u[0], c1 = x[1] - cs.u[1], c2 = x[2] - cs.u[2];
        const double t = act ? 0.5 * X.w * ((c0 * c0 + c1 * c1 + c2 * c2) * g + h) : 0.0;
        double q[3] = {t * c0, t * c1, t * c2};
#pragma unroll
        for (int k = 0; k < 3; ++k) q[k] = seg_sum<LP>(q[k]);
        if (act) {
            og[o] = (1.0 - Cs) * g + Cs * gs;
            oh[o] = (1.0 - Cs) * h + Cs * hs;
        }
        if (v == 0)
            for (int k = 0; k < 3; ++k) a.q_out[3 * c + k] = 0.0 + q[k];
    } else {
        double feqn, feq, dummy;
        eq_store<MODEL, D, double>(a, cs.en, x, vv, feqn, dummy);
        eq_store<MODEL, D, double>(a, cs.e, x, vv, feq, dummy);
        // cdugks.cpp:325-332 as written
        const double cm = cs.cm, C = cs.C, ceq = cs.Cs;
        const double itau = 1.0 / cs.tau;
        const double fn = cm * feqn + (1.0 - cm) * fcg;
        if (act) og[o] = (1.0 - ceq) * (C * (fn + hd * (feq * itau + (feqn - fn) * itau) - dtv * Fg)) + ceq * feq;
    }
}
template <int MODEL, int D, int LP>
__global__ __launch_bounds__(BLOCK) void k_small_cells(KArgs a) {
    d_small_cells<MODEL, D, LP>(a, xcd_block(blockIdx.x, gridDim.x));
}
