// gfx950 DUGKS/CDUGKS step: kernels, device context and the C ABI of
// include/meso_hip.h (meso_hip_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "meso_hip.h"
#include "step_kernels.h"

using namespace meso_dev;

namespace {

constexpr int BLOCK = 256;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ const meso_mark_desc &mark_of(const KArgs &a, int m) {
    return static_cast<const meso_mark_desc *>(a.marks)[m];
}

// Compressible::solve_pressure (pressure_boundary.cpp:15-18)
__device__ __forceinline__ double solve_pressure(double P0, double R, double T, const double U[3], double gam) {
    double Ma = sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]) / sqrt(gam * R * T);
    return P0 / pow(1.0 + (gam - 1.0) / 2.0 * (Ma * Ma), gam / (gam - 1.0));
}

// Column layout of the raw face moments.
template <int D>
struct Col {
    static constexpr int M0 = 0;
    static constexpr int M1 = 1;                        // D entries
    static constexpr int M2 = 1 + D;                    // xx, xy, (xz), yy, (yz, zz)
    static constexpr int M3 = 1 + D + D * (D + 1) / 2;  // D entries
    static constexpr int H0 = M3 + D;
    static constexpr int H1 = H0 + 1;                   // D entries
};

template <int D>
__device__ __forceinline__ double m2_at(const double *M, int i, int j) {
    if (i > j) {
        int t = i;
        i = j;
        j = t;
    }
    if (D == 2) return M[Col<2>::M2 + (i == 0 ? j : 2)];
    // 3D order: xx xy xz yy yz zz
    int idx = (i == 0) ? j : (i == 1 ? 2 + j : 5);
    return M[Col<3>::M2 + idx];
}

// Value of face f at velocity v for its owner-side reconstruction on a
// boundary face (own == nei == c), mixed with the face equilibrium.
template <int D, class T, bool LIM>
__device__ __forceinline__ T bnd_mix(const KArgs &a, const T *__restrict__ fd, int f, int c, double sn, int v,
                                     const T xh[3], T eqv, T mix) {
    const T fc = fd[size_t(c) * a.s.nvp + v];
    T fb;
    if (!a.s.grad) {
        fb = (sn >= VSMALL || sn < -VSMALL) ? fc : T(0.5) * (fc + fc);
    } else {
        T gr[3];
        lsq_grad<D, T>(a, fd, c, v, fc, gr);
        T X = extrap<D, T>(a, f, 0, xh, gr);
        if constexpr (LIM) {
            double dw = 0.0;
            for (int i = 0; i < D; ++i) dw += double(face_d<T>(a, f, 0, i)) * double(gr[i]);
            X *= T(nb_phi<T>(a, fd, c, v, double(fc), dw, sn >= VSMALL ? 1.0e-10 : VSMALL));
        }
        fb = (sn >= VSMALL || sn < -VSMALL) ? fc + X : T(0.5) * (fc + fc + X + X);
    }
    return lerp_store<T>(fb, eqv, mix);
}

// Face-mixed pair (g, h) of boundary face f at velocity v.
template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void bnd_mix_pair(const KArgs &a, int f, int c, int v, const EqParams &ef, T &g, T &h) {
    T x[3];
    load_xi<T>(a, v, x);
    const T hd = T(a.s.half_dt);
    T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
    double sn = ndot(a, f, v);
    T mix = T(ef.mix);
    T gs, hs;
    eq_store<MODEL, D, T>(a, ef, x, v, gs, hs);
    g = bnd_mix<D, T, LIM>(a, (const T *)a.fin_g, f, c, sn, v, xh, gs, mix);
    h = MODEL == CDUGKS_SHAKHOV ? bnd_mix<D, T, LIM>(a, (const T *)a.fin_h, f, c, sn, v, xh, hs, mix) : T(0);
}

// Final boundary value at velocity v given the mixed value (gm_, hm_) there.
template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void bnd_value(const KArgs &a, const BndParams &P, int f, int c, int v, double sn,
                                          const EqParams &ef, T gmix, T hmix, T &g, T &h) {
    g = gmix;
    h = hmix;
    const double kn = -sn;  // xi . n1, n1 = -n0 (exact negation)
    if (P.kind == BC_INTERIOR || kn < -VSMALL) return;
    T x[3];
    load_xi<T>(a, v, x);
    T gi = T(0), hi = T(0);
    if (P.kind != BC_SYMMETRY) eq_store<MODEL, D, T>(a, P.in, x, v, gi, hi);
    switch (P.kind) {
        case BC_FARFIELD_INLET:
        case BC_PRESSURE_INLET:
        case BC_PRESSURE_OUTLET:
        case BC_FARFIELD_OUTLET:
            if (kn >= VSMALL) {
                g = gi;
                h = hi;
            }
            return;
        case BC_WALL:
            if (kn >= VSMALL) {
                g = gi;
                h = hi;
            } else {
                g = T(0.5) * (gi + gmix);
                h = T(0.5) * (hi + hmix);
            }
            return;
        case BC_SLIP_WALL:
        case BC_SYMMETRY: {
            int vm = a.mirror[3 * v + P.dir];
            T gr, hr;
            bnd_mix_pair<MODEL, D, T, LIM>(a, f, c, vm, ef, gr, hr);
            T gv, hv;
            if (P.kind == BC_SLIP_WALL) {
                gv = (T(1) - T(P.alpha_u)) * gi + T(P.alpha_u) * gr;
                hv = (T(1) - T(P.alpha_T)) * hi + T(P.alpha_T) * hr;
            } else {
                gv = gr;
                hv = hr;
            }
            if (kn >= VSMALL) {
                g = gv;
                h = hv;
            } else if (P.kind == BC_SLIP_WALL) {
                g = T(0.5) * (gv + gmix);
                h = T(0.5) * (hv + hmix);
            } else {
                g = T(0.5) * (gmix + gv);
                h = T(0.5) * (hmix + hv);
            }
            return;
        }
        default: return;
    }
}

// FP64 boundary-face values from any storage, as full values (the stored
// deviation of FP32 storage plus f_ref). k_boundary uses them for the wall
// density sums and the exact boundary flux moments, so the conservative update
// of boundary cells carries no FP32 rounding of weights or geometry.
template <class T, bool LIM>
__device__ __forceinline__ double bnd_mix_full(const KArgs &a, const T *__restrict__ fd, int k, int f, int c,
                                               double sn, int v, const double xh[3], double eq_full, double mix) {
    const int nvp = a.s.nvp;
    if constexpr (sizeof(T) == 8) {
        // FP64 and FP32X storage with the patch kernels: the moment kernel
        // wrote this face's f-bar+ (the same reconstruction, FP64)
        if (a.fbar_g) {
            const T *fbr = k == 0 ? (const T *)a.fbar_g : (const T *)a.fbar_h;
            return (1.0 - mix) * double(fbr[size_t(f) * nvp + v]) + mix * eq_full;
        }
    }
    if constexpr (sizeof(T) == 4) {
        // FP32 storage with the patch kernels: the moment kernel's f-bar+ (the
        // FP64 reconstruction, its deviation rounded to FP32), the values the
        // update kernel uses for the outgoing half of the same face
        if (a.fbar_g) {
            const T *fbr = k == 0 ? (const T *)a.fbar_g : (const T *)a.fbar_h;
            const double fb = double(fbr[size_t(f) * nvp + v]) + a.fref_d[k * a.s.nv + v];
            return (1.0 - mix) * fb + mix * eq_full;
        }
    }
    // (FP32X storage always runs the patch kernels, so its values come from f-bar+ above)
    auto ld = [&](size_t i) -> double { return double(fd[i]); };
    const double fs = ld(size_t(c) * nvp + v);
    double fb;
    const bool dead = !(sn >= VSMALL || sn < -VSMALL);
    if (!a.s.grad) {
        fb = dead ? 0.5 * (fs + fs) : fs;
    } else {
        double gr[3] = {0.0, 0.0, 0.0};
        for (int j = a.nb_ptr[c]; j < a.nb_ptr[c + 1]; ++j) {
            const double df = ld(size_t(a.nb_ids[j]) * nvp + v) - fs;
            gr[0] += a.nb_a[3 * j] * df;
            gr[1] += a.nb_a[3 * j + 1] * df;
            gr[2] += a.nb_a[3 * j + 2] * df;
        }
        double X = (a.face_d[6 * f] - xh[0]) * gr[0] + (a.face_d[6 * f + 1] - xh[1]) * gr[1] +
                   (a.face_d[6 * f + 2] - xh[2]) * gr[2];
        if constexpr (LIM) {
            const double dw = a.face_d[6 * f] * gr[0] + a.face_d[6 * f + 1] * gr[1] + a.face_d[6 * f + 2] * gr[2];
            X *= nb_phi<T>(a, fd, c, v, fs, dw, sn >= VSMALL ? 1.0e-10 : VSMALL);
        }
        fb = dead ? 0.5 * (fs + fs + X + X) : fs + X;
    }
    if constexpr (sizeof(T) == 4) fb += a.fref_d[k * a.s.nv + v];
    return (1.0 - mix) * fb + mix * eq_full;
}

template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void bnd_mix_pair_full(const KArgs &a, int f, int c, int v, const EqParams &ef, double &g,
                                                  double &h) {
    const double x[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
    const double hd = a.s.half_dt;
    const double xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
    const double sn = ndot(a, f, v);
    double gs, hs = 0.0;
    if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(ef, x, D, a.s.K, gs, hs);
    else gs = bgk_eq<double>(ef, x);
    g = bnd_mix_full<T, LIM>(a, (const T *)a.fin_g, 0, f, c, sn, v, xh, gs, ef.mix);
    h = MODEL == CDUGKS_SHAKHOV ? bnd_mix_full<T, LIM>(a, (const T *)a.fin_h, 1, f, c, sn, v, xh, hs, ef.mix) : 0.0;
}

// Final full value at velocity v of boundary face f from the mixed value there.
template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void bnd_value_full(const KArgs &a, const BndParams &P, int f, int c, int v, double sn,
                                               const EqParams &ef, double gmix, double hmix, double &g, double &h) {
    g = gmix;
    h = hmix;
    const double kn = -sn;
    if (P.kind == BC_INTERIOR || kn < -VSMALL) return;
    const double x[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
    double gi = 0.0, hi = 0.0;
    if (P.kind != BC_SYMMETRY) {
        if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(P.in, x, D, a.s.K, gi, hi);
        else gi = bgk_eq<double>(P.in, x);
    }
    switch (P.kind) {
        case BC_FARFIELD_INLET:
        case BC_PRESSURE_INLET:
        case BC_PRESSURE_OUTLET:
        case BC_FARFIELD_OUTLET:
            if (kn >= VSMALL) {
                g = gi;
                h = hi;
            }
            return;
        case BC_WALL:
            if (kn >= VSMALL) {
                g = gi;
                h = hi;
            } else {
                g = 0.5 * (gi + gmix);
                h = 0.5 * (hi + hmix);
            }
            return;
        case BC_SLIP_WALL:
        case BC_SYMMETRY: {
            const int vm = a.mirror[3 * v + P.dir];
            double gr, hr;
            bnd_mix_pair_full<MODEL, D, T, LIM>(a, f, c, vm, ef, gr, hr);
            double gv, hv;
            if (P.kind == BC_SLIP_WALL) {
                gv = (1.0 - P.alpha_u) * gi + P.alpha_u * gr;
                hv = (1.0 - P.alpha_T) * hi + P.alpha_T * hr;
            } else {
                gv = gr;
                hv = hr;
            }
            if (kn >= VSMALL) {
                g = gv;
                h = hv;
            } else if (P.kind == BC_SLIP_WALL) {
                g = 0.5 * (gv + gmix);
                h = 0.5 * (hv + hmix);
            } else {
                g = 0.5 * (gmix + gv);
                h = 0.5 * (hmix + hv);
            }
            return;
        }
        default: return;
    }
}

// ================================================================ kernels
// A bad cell (NaN density, T < 0 or NaN T) after the conservative update. A
// Shakhov cell of a masked group is marked (mask 2) for k_mask_repair and
// counted; any other stops the run after this step (cdugks_shakhov.cpp:698-719).
// step_off: how far the device step counter has run ahead of this step.
template <int MODEL>
__device__ __forceinline__ void flag_bad(const KArgs &a, int c, int step_off = 0) {
    if (MODEL == CDUGKS_SHAKHOV && a.mask && a.mask[c]) {
        const_cast<int *>(static_cast<const int *>(a.mask))[c] = 2;
        atomicAdd(a.masked, 1);
        return;
    }
    atomicAdd(a.bad_count, 1);
    atomicMin(a.first_bad, c);
    atomicMin(a.bad_step, *a.step_ctr - step_off);
}

// Sign of n0 . xi against the +-VSMALL dead-band (meso's upwind test,
// cdugks_shakhov.cpp:227-254). FP64 storage evaluates it in FP64 exactly as
// meso does; FP32 storage uses the FP32 product when it is clearly away from
// zero (its rounding error is below 1e-5 for |xi| < 50) and falls back to the
// FP64 product otherwise, so both paths take the same branch as meso.
// Returns +1 (owner upwind), -1 (neighbour upwind) or 0 (dead-band), and the
// product in sn.
template <class T>
__device__ __forceinline__ int upwind(const KArgs &a, int f, int v, const T x[3], T &sn) {
    if constexpr (sizeof(T) == 4) {
        const float s = a.face_nf[3 * f] * x[0] + a.face_nf[3 * f + 1] * x[1] + a.face_nf[3 * f + 2] * x[2];
        sn = s;
        if (fabsf(s) > 1e-4f) return s > 0.0f ? 1 : -1;
    }
    const double d = ndot(a, f, v);
    sn = T(d);
    return d >= VSMALL ? 1 : (d < -VSMALL ? -1 : 0);
}

// LSQ gradient of cell c at velocity v from the padded per-cell tables.
template <int D, class T>
__device__ __forceinline__ void grad_tab(const KArgs &a, const T *__restrict__ fd, int c, int v, T fc, T gr[3]) {
    constexpr int MAXN = D == 2 ? 4 : 6;
    gr[0] = gr[1] = gr[2] = T(0);
    const int n = a.lsq_cnt[c];
    const int nvp = a.s.nvp;
#pragma unroll
    for (int j = 0; j < MAXN; ++j) {
        if (j < n) {
            const int nb = a.lsq_nb[c * MAXN + j];
            const T df = fd[size_t(nb) * nvp + v] - fc;
            if constexpr (sizeof(T) == 8) {
                const double *aj = a.lsq_a + 3 * (c * MAXN + j);
                gr[0] += aj[0] * df;
                gr[1] += aj[1] * df;
                if (D == 3) gr[2] += aj[2] * df;
            } else {
                const float *aj = a.lsq_af + 3 * (c * MAXN + j);
                gr[0] += aj[0] * df;
                gr[1] += aj[1] * df;
                if (D == 3) gr[2] += aj[2] * df;
            }
        }
    }
}

template <int D, class T>
__device__ __forceinline__ T dot_ext(const T d0, const T d1, const T d2, const T xh[3], const T gr[3]) {
    T r = (d0 - xh[0]) * gr[0] + (d1 - xh[1]) * gr[1];
    if (D == 3) r += (d2 - xh[2]) * gr[2];
    return r;
}

// Gradient of one stencil group from its loaded values: sum_j a_j (f_j - f_0).
template <int D, int MAXN, class T>
__device__ __forceinline__ void group_grad(const T *vals, const T *__restrict__ av, T gr[3]) {
    gr[0] = gr[1] = gr[2] = T(0);
#pragma unroll
    for (int j = 0; j < MAXN; ++j) {
        const T df = vals[1 + j] - vals[0];
        gr[0] += av[3 * j] * df;
        gr[1] += av[3 * j + 1] * df;
        if (D == 3) gr[2] += av[3 * j + 2] * df;
    }
}

template <class T>
__device__ __forceinline__ const T *coef_ptr(const double *d, const float *f, size_t off) {
    if constexpr (sizeof(T) == 8) return reinterpret_cast<const T *>(d + off);
    else return reinterpret_cast<const T *>(f + off);
}

// Limiter factor of one stencil group (the cell first, then its face
// neighbours, padded with the cell) for the face offset d and gradient gr.
template <int D, int GS, class T>
__device__ __forceinline__ double group_phi(const T *vals, const T *d, const T *gr, double omega, double eps) {
    double mn = double(vals[0]), mx = mn;
#pragma unroll
    for (int j = 1; j < GS; ++j) {
        const double w = double(vals[j]);
        mn = fmin(mn, w);
        mx = fmax(mx, w);
    }
    double dw = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) dw += double(d[i]) * double(gr[i]);
    return venkata_phi(dw, double(vals[0]), mn, mx, omega, eps);
}

// (1) raw moments of f-bar+ per (face, velocity tile); one wave per pair.
// All stencil values of an iteration are loaded before any is used.
// b: the block's logical index in its context's grid (XCD-mapped, xcd_block).
template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void d_face_moments(const KArgs &a, const int b) {
    if (*a.bad_count > 0) {
        if (b == 0 && threadIdx.x == 0) *a.halt = 1;
        return;
    }
    constexpr int WPB = BLOCK / 64;
    constexpr int MAXN = D == 2 ? 4 : 6;
    constexpr int GS = 1 + MAXN;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    const int nf = a.s.nf;
    const int nblk = (nf + WPB - 1) / WPB;
    const int tile = b / nblk, fb = b - tile * nblk;
    const int lane = threadIdx.x % 64;
    const int f = __builtin_amdgcn_readfirstlane(fb * WPB + int(threadIdx.x / 64));
    if (f >= nf) return;
    const int nvp = a.s.nvp;
    const T hd = T(a.s.half_dt);
    const T *fg = (const T *)a.fin_g;
    const T *fh = (const T *)a.fin_h;
    const int *ids = a.fs_id + size_t(f) * 2 * GS;
    const T *av = coef_ptr<T>(a.fs_a, a.fs_af, size_t(f) * 2 * MAXN * 3);
    T d[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = face_d<T>(a, f, k / 3, k % 3);
    double acc[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = 0.0;
    const int v0 = tile * a.s.vt_f, v1 = min(v0 + a.s.vt_f, a.s.nv);
    for (int v = v0 + lane; v < v1; v += 64) {
        T vg[2 * GS], vh[2 * GS];
#pragma unroll
        for (int k = 0; k < 2 * GS; ++k) {
            const size_t r = size_t(ids[k]) * nvp + v;
            vg[k] = fg[r];
            if (SH) vh[k] = fh[r];
        }
        const double4 X = a.vel_d[v];
        T x[3];
        if constexpr (sizeof(T) == 8) {
            x[0] = X.x;
            x[1] = X.y;
            x[2] = X.z;
        } else {
            const float4 xf = a.vel_f[v];
            x[0] = xf.x;
            x[1] = xf.y;
            x[2] = xf.z;
        }
        const T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
        T sn;
        const int up = upwind<T>(a, f, v, x, sn);
        T go_g[3] = {T(0), T(0), T(0)}, go_h[3] = {T(0), T(0), T(0)};
        T gn_g[3] = {T(0), T(0), T(0)}, gn_h[3] = {T(0), T(0), T(0)};
        if (a.s.grad) {
            if (__any(up >= 0)) {
                group_grad<D, MAXN, T>(vg, av, go_g);
                if (SH) group_grad<D, MAXN, T>(vh, av, go_h);
            }
            if (__any(up <= 0)) {
                group_grad<D, MAXN, T>(vg + GS, av + 3 * MAXN, gn_g);
                if (SH) group_grad<D, MAXN, T>(vh + GS, av + 3 * MAXN, gn_h);
            }
        }
        T Xog = dot_ext<D, T>(d[0], d[1], d[2], xh, go_g), Xng = dot_ext<D, T>(d[3], d[4], d[5], xh, gn_g);
        double om_o = 0.0, om_n = 0.0;
        if constexpr (SH && LIM) {  // cdugks_shakhov.cpp:231-266
            om_o = venkata_omega(a.s.venkata_k, a.cell_vol[ids[0]], D);
            om_n = venkata_omega(a.s.venkata_k, a.cell_vol[ids[GS]], D);
            Xog *= T(group_phi<D, GS, T>(vg, d, go_g, om_o, up > 0 ? 1.0e-10 : VSMALL));
            Xng *= T(group_phi<D, GS, T>(vg + GS, d + 3, gn_g, om_n, VSMALL));
        }
        const T gb = up > 0 ? vg[0] + Xog : (up < 0 ? vg[GS] + Xng : T(0.5) * (vg[0] + vg[GS] + Xog + Xng));
        T hb = T(0);
        if (SH) {
            T Xoh = dot_ext<D, T>(d[0], d[1], d[2], xh, go_h), Xnh = dot_ext<D, T>(d[3], d[4], d[5], xh, gn_h);
            if constexpr (LIM) {
                Xoh *= T(group_phi<D, GS, T>(vh, d, go_h, om_o, up > 0 ? 1.0e-10 : VSMALL));
                Xnh *= T(group_phi<D, GS, T>(vh + GS, d + 3, gn_h, om_n, VSMALL));
            }
            hb = up > 0 ? vh[0] + Xoh : (up < 0 ? vh[GS] + Xnh : T(0.5) * (vh[0] + vh[GS] + Xoh + Xnh));
        }
        const double X0 = X.x, X1 = X.y, X2 = X.z;
        const double wv = X.w;
        const double wg = wv * double(gb);
        const double wgx = wg * X0, wgy = wg * X1, wgz = wg * X2;
        acc[Col<D>::M0] += wg;
        acc[Col<D>::M1] += wgx;
        acc[Col<D>::M1 + 1] += wgy;
        if (D == 3) acc[Col<D>::M1 + 2] += wgz;
        if (D == 2) {
            acc[Col<D>::M2] += wgx * X0;
            acc[Col<D>::M2 + 1] += wgx * X1;
            acc[Col<D>::M2 + 2] += wgy * X1;
        } else {
            acc[Col<D>::M2] += wgx * X0;
            acc[Col<D>::M2 + 1] += wgx * X1;
            acc[Col<D>::M2 + 2] += wgx * X2;
            acc[Col<D>::M2 + 3] += wgy * X1;
            acc[Col<D>::M2 + 4] += wgy * X2;
            acc[Col<D>::M2 + 5] += wgz * X2;
        }
        if (SH) {
            const double wh = wv * double(hb);
            const double wgk = wg * (X0 * X0 + X1 * X1 + X2 * X2);
            acc[Col<D>::M3] += wgk * X0;
            acc[Col<D>::M3 + 1] += wgk * X1;
            if (D == 3) acc[Col<D>::M3 + 2] += wgk * X2;
            acc[Col<D>::H0] += wh;
            acc[Col<D>::H1] += wh * X0;
            acc[Col<D>::H1 + 1] += wh * X1;
            if (D == 3) acc[Col<D>::H1 + 2] += wh * X2;
        }
    }
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = seg_sum<64>(acc[k]);
    if (lane == 0) {
        double *dst = a.part_f + (size_t(f) * a.s.ntile_f + tile) * NCOL;
#pragma unroll
        for (int k = 0; k < NCOL; ++k) dst[k] = acc[k];
    }
}
template <int MODEL, int D, class T, bool LIM>
__global__ __launch_bounds__(BLOCK) void k_face_moments(KArgs a) {
    d_face_moments<MODEL, D, T, LIM>(a, xcd_block(blockIdx.x, gridDim.x));
}

// Axis-table entry t of equilibrium e in the storage type's form.
template <class T>
__device__ __forceinline__ T tab_entry(const KArgs &a, const EqParams &e, int t) {
    const double X = a.tab_x[t];
    const int ax = a.tab_ax[t];
    const double c = X - (ax == 0 ? e.u[0] : (ax == 1 ? e.u[1] : e.u[2]));
    if constexpr (sizeof(T) == 8) return exp(-e.B * c * c);
    else return T(expm1(a.ref.B * X * X - e.B * c * c));
}

// Flux moments of the Shakhov equilibrium along n for a tensor-product
// velocity set, from 1D sums: with c = xi - u, every moment
// sum_v w (xi.n) psi(xi) g_S(xi) is A times a polynomial in c of degree <= 6
// against w_x w_y w_z E_x E_y E_z (E_a = exp(-B c_a^2)), so it is a
// combination of the products Cx[a] Cy[b] Cz[c] of the central 1D moments
// C_d[k] = sum_i w_d,i E_d,i c_d,i^k (k <= 6). Returns the D + 2 moments
// psi = 1, xi_1..xi_D, |xi|^2 g + h (the order of Fb in k_face_eq).
// g_S = g_M (1 + P1 c.q (|c|^2/RT - D - 2)),
// h_S = g_M (hk + P2 c.q ((|c|^2/RT - D)(K + 3 - D) - 2K)) (cdugks_shakhov.cpp:198-211).
template <int D>
__device__ __forceinline__ void shakhov_flux_sep(const EqParams &e, const double (&C)[3][7], const double n[3], int K,
                                                 double *out) {
    const double kD = double(K + 3 - D);
    auto M = [&](int a, int b, int c) { return C[0][a] * C[1][b] * C[2][c]; };
    // <X^a Y^b Z^c (c.q) (iRT |c|^2 - s)>
    auto Q = [&](int a, int b, int c, double r, double s) {
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int a1 = a + (d == 0), b1 = b + (d == 1), c1 = c + (d == 2);
            double sq = M(a1 + 2, b1, c1) + M(a1, b1 + 2, c1);
            if (D == 3) sq += M(a1, b1, c1 + 2);
            t += e.q[d] * (r * sq - s * M(a1, b1, c1));
        }
        return t;
    };
    auto G = [&](int a, int b, int c) { return M(a, b, c) + e.P1 * Q(a, b, c, e.iRT, double(D + 2)); };
    auto H = [&](int a, int b, int c) {
        return e.hk * M(a, b, c) + e.P2 * Q(a, b, c, kD * e.iRT, kD * double(D) + 2.0 * double(K));
    };
    double s0 = 0.0, uu = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        s0 += n[d] * e.u[d];
        uu += e.u[d] * e.u[d];
    }
    // <(xi.n) X^a Y^b Z^c g_S / g_M-factor>, xi.n = s0 + n.c
    auto SG = [&](int a, int b, int c) {
        double t = s0 * G(a, b, c) + n[0] * G(a + 1, b, c) + n[1] * G(a, b + 1, c);
        if (D == 3) t += n[2] * G(a, b, c + 1);
        return t;
    };
    const double sg0 = SG(0, 0, 0);
    double sgd[3] = {SG(1, 0, 0), SG(0, 1, 0), D == 3 ? SG(0, 0, 1) : 0.0};
    out[0] = e.A * sg0;
#pragma unroll
    for (int d = 0; d < D; ++d) out[1 + d] = e.A * (sgd[d] + e.u[d] * sg0);
    double sh0 = s0 * H(0, 0, 0) + n[0] * H(1, 0, 0) + n[1] * H(0, 1, 0);
    if (D == 3) sh0 += n[2] * H(0, 0, 1);
    double e2 = SG(2, 0, 0) + SG(0, 2, 0);
    if (D == 3) e2 += SG(0, 0, 2);
    double ud = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) ud += e.u[d] * sgd[d];
    out[D + 1] = e.A * (e2 + 2.0 * ud + uu * sg0 + sh0);
}

// Face macros from the raw moments M of f-bar+ (face-moment column layout):
// rho, u, T, the Shakhov heat flux (cdugks_shakhov.cpp:317-362; BGK
// cdugks.cpp:146-170), the equilibrium constants e with its mix weight, and Fb,
// the flux moments of f-bar+ itself along n.
template <int MODEL, int D>
__device__ __forceinline__ void face_macros(const KArgs &a, const double *M, const double n[3], EqParams &e,
                                            double *Fb) {
    const double rho = M[Col<D>::M0];
    const double inv = 1.0 / rho;
    double u[3] = {M[Col<D>::M1] * inv, M[Col<D>::M1 + 1] * inv, D == 3 ? M[Col<D>::M1 + 2] * inv : 0.0};
    double M1[3] = {M[Col<D>::M1], M[Col<D>::M1 + 1], D == 3 ? M[Col<D>::M1 + 2] : 0.0};
    double Mn[3];  // M2 . n
    for (int i = 0; i < 3; ++i) {
        Mn[i] = 0.0;
        if (i < D)
            for (int j = 0; j < D; ++j) Mn[i] += m2_at<D>(M, i, j) * n[j];
    }
    Fb[0] = n[0] * M1[0] + n[1] * M1[1] + n[2] * M1[2];
    for (int i = 0; i < D; ++i) Fb[1 + i] = Mn[i];
    const double hdt = a.s.half_dt;
    if (MODEL == CDUGKS_SHAKHOV) {
        double trM2 = 0.0;
        for (int i = 0; i < D; ++i) trM2 += m2_at<D>(M, i, i);
        const double H0 = M[Col<D>::H0];
        const double rhoE = 0.5 * (trM2 + H0);
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double T_ = (rhoE / rho - 0.5 * uu) / a.s.Cv;
        const double tau = tau_of(a.s, rho, T_);
        // central heat flux sum w c (|c|^2 g + h), c = xi - u, from raw moments
        double Mu[3];
        for (int i = 0; i < 3; ++i) {
            Mu[i] = 0.0;
            if (i < D)
                for (int j = 0; j < D; ++j) Mu[i] += m2_at<D>(M, i, j) * u[j];
        }
        const double M1u = M1[0] * u[0] + M1[1] * u[1] + M1[2] * u[2];
        double q[3];
        const double cq = tau / (2.0 * tau + hdt * a.s.Pr);
        for (int i = 0; i < 3; ++i) {
            const double M3i = i < D ? M[Col<D>::M3 + i] : 0.0;
            const double H1i = i < D ? M[Col<D>::H1 + i] : 0.0;
            const double m3 = M3i - 2.0 * Mu[i] + uu * M1[i] - u[i] * trM2 + 2.0 * u[i] * M1u -
                              u[i] * uu * rho + H1i - u[i] * H0;
            q[i] = cq * m3;
        }
        make_shakhov(e, rho, u, T_, q, a.s.R, a.s.Pr, D, a.s.K);
        e.mix = hdt / (2.0 * tau + hdt);
        double M3n = 0.0, H1n = 0.0;
        for (int i = 0; i < D; ++i) {
            M3n += n[i] * M[Col<D>::M3 + i];
            H1n += n[i] * M[Col<D>::H1 + i];
        }
        Fb[D + 1] = M3n + H1n;
    } else {
        make_bgk(e, rho, u, a.s.RT);
        e.mix = hdt / (2.0 * a.s.tau + hdt);
    }
}

// (2) face macros, equilibrium constants, axis tables and flux moments of the
// mixed value of interior faces; one wave per face
// b: the block's logical index in its context's grid (XCD-mapped, xcd_block).
template <int MODEL, int D, class T, bool TAB>
__device__ __forceinline__ void d_face_eq(const KArgs &a, const int b) {
    extern __shared__ unsigned char lds_raw[];
    if (*a.halt) return;
    constexpr int WPB = BLOCK / 64;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr int NFL = D + 2;
    const int nf = a.s.nf;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int f = __builtin_amdgcn_readfirstlane(b * WPB + wave);
    const bool act = f < nf;
    T *tb = reinterpret_cast<T *>(lds_raw) + size_t(wave) * a.nt;
    EqParams e;
    EqF ef;
    double Fb[NFL];
    double acc[NFL];
#pragma unroll
    for (int k = 0; k < NFL; ++k) acc[k] = Fb[k] = 0.0;
    bool interior = false;
    double n[3] = {0.0, 0.0, 0.0};
    if (act) {
        double M[NCOL];
#pragma unroll
        for (int k = 0; k < NCOL; ++k) M[k] = 0.0;
        const double *src = a.part_f + size_t(f) * a.s.ntile_f * NCOL;
        for (int t = 0; t < a.s.ntile_f; ++t)
#pragma unroll
            for (int k = 0; k < NCOL; ++k) M[k] += src[t * NCOL + k];
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int k = 0; k < NCOL; ++k) M[k] += a.mref[k];
        }
        for (int i = 0; i < 3; ++i) n[i] = a.face_n[3 * f + i];
        face_macros<MODEL, D>(a, M, n, e, Fb);
        if constexpr (sizeof(T) == 4) {
            if (MODEL == CDUGKS_SHAKHOV) relate_shakhov(e, a.ref);
            else relate_bgk(e, a.ref);
        }
        if constexpr (sizeof(T) == 4) {
            ef = to_eqf(e);
            if (MODEL != CDUGKS_SHAKHOV) ef.hk = float(e.A);
        }
        if (lane == 0) {
            a.face_eq[f] = e;
            if constexpr (sizeof(T) == 4) a.face_eqf[f] = ef;
        }
        if (TAB) {
            T *gt = static_cast<T *>(a.face_tab) + size_t(f) * a.nt;
            for (int t = lane; t < a.nt; t += 64) {
                const T val = tab_entry<T>(a, e, t);
                tb[t] = val;
                gt[t] = val;
            }
        }
        interior = a.face_bnd[f] < 0;
    }
    if (TAB) __syncthreads();
    if (MODEL == CDUGKS_SHAKHOV && TAB && a.tab_w != nullptr) {
        // tensor-product set with factorised weights: 1D central moments per
        // axis (lanes over the axis entries), then the closed combination
        if (act && interior) {
            double C[3][7];
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int k = 0; k < 7; ++k) C[d][k] = 0.0;
            for (int t = lane; t < a.nt; t += 64) {
                const int ax = a.tab_ax[t];
                const double c = a.tab_x[t] - (ax == 0 ? e.u[0] : (ax == 1 ? e.u[1] : e.u[2]));
                double m = a.tab_w[t] * exp(-e.B * c * c);
#pragma unroll
                for (int k = 0; k < 7; ++k) {
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        if (d == ax) C[d][k] += m;
                    m *= c;
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d)
#pragma unroll
                for (int k = 0; k < 7; ++k) C[d][k] = seg_sum<64>(C[d][k]);
            if (D == 2) {
                C[2][0] = 1.0;  // the third axis is the single abscissa 0 with weight 1
#pragma unroll
                for (int k = 1; k < 7; ++k) C[2][k] = 0.0;
            }
            if (lane == 0) {
                double full[NFL];
                shakhov_flux_sep<D>(e, C, n, a.s.K, full);
                const double mix = e.mix;
                for (int k = 0; k < NFL; ++k) a.face_flux[size_t(f) * NFL + k] = (1.0 - mix) * Fb[k] + mix * full[k];
            }
        }
        return;
    }
    if (act && interior) {
        const int nv = a.s.nv;
        for (int v = lane; v < nv; v += 64) {
            const double4 X = a.vel_d[v];
            const double wsn = X.w * (n[0] * X.x + n[1] * X.y + n[2] * X.z);
            T gs, hs = T(0);
            if (MODEL == CDUGKS_SHAKHOV && TAB) {
                if constexpr (sizeof(T) == 4) {
                    const float4 xf = a.vel_f[v];
                    const float x[3] = {xf.x, xf.y, xf.z};
                    shakhov_tab_f(ef, reinterpret_cast<const float *>(tb), a.vtab[v], x, D, a.s.K, a.fref_f2[v].x, gs,
                                  hs);
                } else {
                    const double x[3] = {X.x, X.y, X.z};
                    shakhov_tab_d(e, reinterpret_cast<const double *>(tb), a.vtab[v], x, D, a.s.K, gs, hs);
                }
            } else {
                T x[3];
                load_xi<T>(a, v, x);
                eq_store<MODEL, D, T>(a, e, x, v, gs, hs);
            }
            const double wg = wsn * double(gs);
            acc[0] += wg;
            acc[1] += wg * X.x;
            acc[2] += wg * X.y;
            if (D == 3) acc[3] += wg * X.z;
            if (MODEL == CDUGKS_SHAKHOV) acc[D + 1] += wg * (X.x * X.x + X.y * X.y + X.z * X.z) + wsn * double(hs);
        }
#pragma unroll
        for (int k = 0; k < NFL; ++k) acc[k] = seg_sum<64>(acc[k]);
        if (lane == 0) {
            if constexpr (sizeof(T) == 4) {  // flux moments of f_ref along n
                for (int k = 0; k < NFL; ++k)
                    acc[k] += n[0] * a.rflux[k][0] + n[1] * a.rflux[k][1] + n[2] * a.rflux[k][2];
            }
            const double mix = e.mix;
            const int nfl = MODEL == CDUGKS_SHAKHOV ? D + 2 : D + 1;
            for (int k = 0; k < nfl; ++k) a.face_flux[size_t(f) * NFL + k] = (1.0 - mix) * Fb[k] + mix * acc[k];
            if (MODEL != CDUGKS_SHAKHOV) a.face_flux[size_t(f) * NFL + D + 1] = 0.0;
        }
    }
}
template <int MODEL, int D, class T, bool TAB>
__global__ __launch_bounds__(BLOCK) void k_face_eq(KArgs a) {
    d_face_eq<MODEL, D, T, TAB>(a, xcd_block(blockIdx.x, gridDim.x));
}

// (2') The same stage for Shakhov with a tensor-product velocity set whose
// weights factorise (the flux moments from 1D central moments): LP lanes per
// face, BLOCK / LP faces per block. The work of a face that does not scale with
// the set (its macros, tau, the equilibrium constants, the closed combination
// of the 1D moments) is the same for 101^2 as for 16^2 velocities; with one
// face per wave it ran on one lane or on 64 lanes alike, and the axis-table
// loops left most lanes idle for short axes.
template <int D, class T, int LP>
__device__ __forceinline__ void d_face_eq_sep(const KArgs &a, const int b) {
    if (*a.halt) return;
    constexpr int MODEL = CDUGKS_SHAKHOV;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr int NFL = D + 2;
    constexpr int FPB = BLOCK / LP;
    const int f = b * FPB + int(threadIdx.x) / LP;
    if (f >= a.s.nf) return;  // whole segments leave together
    const int sl = int(threadIdx.x) % LP;
    double M[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; ++k) M[k] = 0.0;
    const double *src = a.part_f + size_t(f) * a.s.ntile_f * NCOL;
    for (int t = 0; t < a.s.ntile_f; ++t)
#pragma unroll
        for (int k = 0; k < NCOL; ++k) M[k] += src[t * NCOL + k];
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < NCOL; ++k) M[k] += a.mref[k];
    }
    const double n[3] = {a.face_n[3 * f], a.face_n[3 * f + 1], a.face_n[3 * f + 2]};
    EqParams e;
    double Fb[NFL];
#pragma unroll
    for (int k = 0; k < NFL; ++k) Fb[k] = 0.0;
    face_macros<MODEL, D>(a, M, n, e, Fb);
    if constexpr (sizeof(T) == 4) relate_shakhov(e, a.ref);
    if (sl == 0) {
        a.face_eq[f] = e;
        if constexpr (sizeof(T) == 4) a.face_eqf[f] = to_eqf(e);
    }
    T *gt = static_cast<T *>(a.face_tab) + size_t(f) * a.nt;
    for (int t = sl; t < a.nt; t += LP) gt[t] = tab_entry<T>(a, e, t);
    if (a.face_bnd[f] >= 0) return;  // k_boundary forms a boundary face's flux moments
    double C[3][7];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 7; ++k) C[d][k] = 0.0;
    for (int t = sl; t < a.nt; t += LP) {
        const int ax = a.tab_ax[t];
        const double c = a.tab_x[t] - (ax == 0 ? e.u[0] : (ax == 1 ? e.u[1] : e.u[2]));
        double m = a.tab_w[t] * exp(-e.B * c * c);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (d == ax) C[d][k] += m;
            m *= c;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int k = 0; k < 7; ++k) C[d][k] = seg_sum<LP>(C[d][k]);
    if (D == 2) {
        C[2][0] = 1.0;  // the third axis is the single abscissa 0 with weight 1
#pragma unroll
        for (int k = 1; k < 7; ++k) C[2][k] = 0.0;
    }
    if (sl == 0) {
        double full[NFL];
        shakhov_flux_sep<D>(e, C, n, a.s.K, full);
        const double mix = e.mix;
        for (int k = 0; k < NFL; ++k) a.face_flux[size_t(f) * NFL + k] = (1.0 - mix) * Fb[k] + mix * full[k];
    }
}
template <int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_face_eq_sep(KArgs a) {
    d_face_eq_sep<D, T, LP>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void kb_face_eq_sep(const KArgs *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_face_eq_sep<D, T, LP>(batch_args(tab, m.x), m.y);
}
constexpr int FEQ_LP = 16;  // lanes per face of k_face_eq_sep
template <int MODEL, int D, class T, bool TAB>
__global__ __launch_bounds__(BLOCK) void kb_face_eq(const KArgs *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_face_eq<MODEL, D, T, TAB>(batch_args(tab, m.x), m.y);
}

// Block-wide sum of N doubles (BLOCK threads); result valid in all threads.
template <int N>
__device__ __forceinline__ void block_sum(double (&x)[N], double *lds /* [N * BLOCK/64] */) {
    const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
    constexpr int NW = BLOCK / 64;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = seg_sum<64>(x[k]);
        if (lane == 0) lds[k * NW + wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double s = 0.0;
        for (int i = 0; i < NW; ++i) s += lds[k * NW + i];
        x[k] = s;
    }
    __syncthreads();
}

// Boundary face set-up (cdugks_shakhov.cpp:364-607, cdugks.cpp:194-238): the
// kind as this model handles it and, for every kind but walls, the imposed
// equilibrium P.in. Returns true for walls and slip walls, whose P.in needs
// the wall density sums; e1 is then the wall Maxwellian at unit density.
template <int MODEL, int D>
__device__ __forceinline__ bool bnd_begin(const KArgs &a, const meso_mark_desc &mk, int f, int c, BndParams &P,
                                          EqParams &e1) {
    P.kind = mk.type;
    P.dir = mk.direction;
    P.alpha_u = mk.alpha_u;
    P.alpha_T = mk.alpha_T;
    P.pad = 0.0;
    if (MODEL != CDUGKS_SHAKHOV &&
        !(P.kind == BC_FARFIELD_INLET || P.kind == BC_FARFIELD_OUTLET || P.kind == BC_WALL))
        P.kind = BC_INTERIOR;  // cdugks.cpp:194-238 handles only these three
    if (P.kind == BC_FREESTREAM_INLET) P.kind = BC_INTERIOR;
    const double n0[3] = {a.face_n[3 * f], a.face_n[3 * f + 1], a.face_n[3 * f + 2]};
    const double n1[3] = {-n0[0], -n0[1], -n0[2]};
    const double rm = a.rho_in[c], Tm = a.T_in[c];
    const double um[3] = {a.u_in[3 * c], a.u_in[3 * c + 1], a.u_in[3 * c + 2]};
    const double uw[3] = {mk.velocity[0], mk.velocity[1], mk.velocity[2]};
    const double zero[3] = {0.0, 0.0, 0.0};
    const double R = a.s.R, Pr = a.s.Pr;
    const int K = a.s.K;
    make_bgk(P.in, 1.0, zero, 1.0);
    P.in.A = 0.0;
    switch (P.kind) {
        case BC_FARFIELD_INLET:
            if (MODEL == CDUGKS_SHAKHOV) make_shakhov(P.in, mk.density, uw, mk.temperature, zero, R, Pr, D, K);
            else make_bgk(P.in, mk.density, uw, a.s.RT);
            break;
        case BC_PRESSURE_INLET: {  // cdugks_shakhov.cpp:374-394
            double un = um[0] * n1[0] + um[1] * n1[1] + um[2] * n1[2];
            double uin[3] = {un * n1[0], un * n1[1], un * n1[2]};
            double pin = solve_pressure(mk.pressure, R, Tm, uin, a.s.gamma);
            double rin = pin / (R * mk.temperature);
            make_shakhov(P.in, rin, uin, mk.temperature, zero, R, Pr, D, K);
        } break;
        case BC_PRESSURE_OUTLET: {  // :430-452
            double pe = solve_pressure(mk.pressure, R, Tm, um, a.s.gamma);
            double re = pe / (R * Tm);
            double un = um[0] * n0[0] + um[1] * n0[1] + um[2] * n0[2];
            double ue[3] = {un * n0[0], un * n0[1], un * n0[2]};
            make_shakhov(P.in, re, ue, mk.temperature, zero, R, Pr, D, K);
        } break;
        case BC_FARFIELD_OUTLET:
            if (MODEL == CDUGKS_SHAKHOV) make_shakhov(P.in, rm, um, Tm, zero, R, Pr, D, K);
            else make_bgk(P.in, rm, um, a.s.RT);
            break;
        case BC_WALL:
        case BC_SLIP_WALL: {  // :468-497 wall density sums, then :531-602
            double Tw = 0.0;
            if (mk.T_kind == MESO_PATCH_ZERO_GRADIENT) Tw = Tm;
            else if (mk.T_kind == MESO_PATCH_FIXED_VALUE) Tw = mk.temperature;
            if (MODEL == CDUGKS_SHAKHOV) make_shakhov(e1, 1.0, uw, Tw, zero, R, Pr, D, K);
            else make_bgk(e1, 1.0, uw, a.s.RT);
            return true;
        }
        default: break;
    }
    return false;
}

// Terms of the wall density sums at velocity v (kn = xi . n1, n1 into the
// domain): s1 gathers the unit-density wall Maxwellian over incoming
// velocities, s0 the mixed face value over outgoing ones.
template <int MODEL, int D>
__device__ __forceinline__ void wall_terms(const KArgs &a, const EqParams &e1, int v, double kn, double gmix,
                                           double s[2]) {
    if (kn >= VSMALL) {
        const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
        double gm, hm;
        if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(e1, X, D, a.s.K, gm, hm);
        else gm = bgk_eq<double>(e1, X);
        s[1] += kn * a.w[v] * gm;
    } else {
        s[0] -= kn * a.w[v] * gmix;
    }
}

// P.in of a wall or slip wall from the summed wall density terms.
template <int MODEL, int D>
__device__ __forceinline__ void bnd_wall_end(const KArgs &a, const meso_mark_desc &mk, BndParams &P, const double s[2]) {
    const double rho_w = s[0] / s[1];
    const double uw[3] = {mk.velocity[0], mk.velocity[1], mk.velocity[2]};
    const double zero[3] = {0.0, 0.0, 0.0};
    if (MODEL == CDUGKS_SHAKHOV) make_shakhov(P.in, rho_w, uw, mk.temperature, zero, a.s.R, a.s.Pr, D, a.s.K);
    else make_bgk(P.in, rho_w, uw, a.s.RT);
}

// (3) boundary faces: one block per face
template <int MODEL, int D, class T, bool LIM>
__device__ __forceinline__ void d_boundary(const KArgs &a, const int b) {
    if (*a.halt) return;
    constexpr int NFL = D + 2;
    __shared__ double lds[(NFL + 2) * (BLOCK / 64)];
    const int f = a.bnd_face[b];
    const int c = a.face_cells[f].x;
    const meso_mark_desc &mk = mark_of(a, a.bnd_mark[b]);
    const EqParams ef = a.face_eq[f];
    const int nv = a.s.nv;
    BndParams P;
    EqParams e1;
    if (bnd_begin<MODEL, D>(a, mk, f, c, P, e1)) {
        double s[2] = {0.0, 0.0};
        for (int v = threadIdx.x; v < nv; v += BLOCK) {
            const double kn = -ndot(a, f, v);
            double gmx = 0.0, hmx;
            if (!(kn >= VSMALL)) bnd_mix_pair_full<MODEL, D, T, LIM>(a, f, c, v, ef, gmx, hmx);
            wall_terms<MODEL, D>(a, e1, v, kn, gmx, s);
        }
        block_sum<2>(s, lds);
        bnd_wall_end<MODEL, D>(a, mk, P, s);
    }
    if constexpr (sizeof(T) == 4) {
        if (P.kind != BC_SYMMETRY && P.kind != BC_INTERIOR) {
            if (MODEL == CDUGKS_SHAKHOV) relate_shakhov(P.in, a.ref);
            else relate_bgk(P.in, a.ref);
        }
    }
    double acc[NFL];
#pragma unroll
    for (int k = 0; k < NFL; ++k) acc[k] = 0.0;
    for (int v = threadIdx.x; v < nv; v += BLOCK) {
        double gmx, hmx, g, h;
        bnd_mix_pair_full<MODEL, D, T, LIM>(a, f, c, v, ef, gmx, hmx);
        const double sn = ndot(a, f, v);
        bnd_value_full<MODEL, D, T, LIM>(a, P, f, c, v, sn, ef, gmx, hmx, g, h);
        const double X0 = a.xi[3 * v], X1 = a.xi[3 * v + 1], X2 = a.xi[3 * v + 2];
        const double wsn = a.w[v] * sn;
        const double wg = wsn * g;
        acc[0] += wg;
        acc[1] += wg * X0;
        acc[2] += wg * X1;
        if (D == 3) acc[3] += wg * X2;
        if (MODEL == CDUGKS_SHAKHOV) acc[D + 1] += wg * (X0 * X0 + X1 * X1 + X2 * X2) + wsn * h;
    }
    block_sum<NFL>(acc, lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NFL; ++k) a.face_flux[size_t(f) * NFL + k] = acc[k];
        a.bnd[b] = P;
    }
}
template <int MODEL, int D, class T, bool LIM>
__global__ __launch_bounds__(BLOCK) void k_boundary(KArgs a) {
    d_boundary<MODEL, D, T, LIM>(a, blockIdx.x);
}
template <int MODEL, int D, class T, bool LIM>
__global__ __launch_bounds__(BLOCK) void kb_boundary(const KArgs *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_boundary<MODEL, D, T, LIM>(batch_args(tab, m.x), m.y);
}

// New cell macros from the face flux moments (cdugks_shakhov.cpp:644-668,
// cdugks.cpp:306-320). Returns false when rho is NaN or T < 0 / NaN.
template <int MODEL, int D>
struct CellStep {
    EqParams en, e;  // equilibria at t_n and t_{n+1}
    double rho, u[3], T;
    double tau, taun, cm, C, Cs, dt_v;
    double phi[D + 2];
};

template <int MODEL, int D, class T>
__device__ __forceinline__ bool cell_macros(const KArgs &a, int c, CellStep<MODEL, D> &cs) {
    constexpr int NFL = D + 2;
    for (int k = 0; k < NFL; ++k) cs.phi[k] = 0.0;
    for (int j = a.cf_ptr[c]; j < a.cf_ptr[c + 1]; ++j) {
        const int f = a.cf_ids[j];
        const double sA = a.cf_sA[j];
        for (int k = 0; k < NFL; ++k) cs.phi[k] += sA * a.face_flux[size_t(f) * NFL + k];
    }
    const double hdt = a.s.half_dt;
    const double dt_v = a.s.dt / a.cell_vol[c];
    cs.dt_v = dt_v;
    const double rn = a.rho_in[c];
    const double un[3] = {a.u_in[3 * c], a.u_in[3 * c + 1], a.u_in[3 * c + 2]};
    if (MODEL == CDUGKS_SHAKHOV) {
        const double Tn = a.T_in[c];
        const double qn[3] = {a.q_in[3 * c], a.q_in[3 * c + 1], a.q_in[3 * c + 2]};
        const double Cv = a.s.Cv;
        const double rEn = rn * (0.5 * (un[0] * un[0] + un[1] * un[1] + un[2] * un[2]) + Cv * Tn);
        const double taun = tau_of(a.s, rn, Tn);
        const double r = rn - dt_v * cs.phi[0];
        double rU[3];
        for (int i = 0; i < 3; ++i) rU[i] = rn * un[i] - (i < D ? dt_v * cs.phi[1 + i] : 0.0);
        const double rE = rEn - dt_v * 0.5 * cs.phi[D + 1];
        const double ir = 1.0 / r;
        for (int i = 0; i < 3; ++i) cs.u[i] = rU[i] * ir;
        const double uu = cs.u[0] * cs.u[0] + cs.u[1] * cs.u[1] + cs.u[2] * cs.u[2];
        cs.T = (rE / r - uu * 0.5) / Cv;
        cs.rho = r;
        cs.tau = tau_of(a.s, r, cs.T);
        cs.taun = taun;
        cs.cm = hdt / (hdt - 2.0 * taun);
        cs.C = cs.tau / (cs.tau + hdt);
        cs.Cs = hdt / (2.0 * cs.tau);
        make_shakhov(cs.en, rn, un, Tn, qn, a.s.R, a.s.Pr, D, a.s.K);
        make_shakhov(cs.e, r, cs.u, cs.T, qn, a.s.R, a.s.Pr, D, a.s.K);
        if constexpr (sizeof(T) == 4) {
            relate_shakhov(cs.en, a.ref);
            relate_shakhov(cs.e, a.ref);
        }
        return !(isnan(r) || cs.T < 0.0 || isnan(cs.T));
    } else if (MODEL == CDUGKS_BGK) {
        const double tau = a.s.tau;
        const double r = rn - dt_v * cs.phi[0];
        const double irn = 1.0 / rn;
        for (int i = 0; i < 3; ++i) cs.u[i] = (rn * un[i] - (i < D ? dt_v * cs.phi[1 + i] : 0.0)) * irn;
        cs.rho = r;
        cs.T = 0.0;
        cs.tau = cs.taun = tau;
        cs.cm = hdt / (hdt - 2.0 * tau);
        cs.C = tau / (tau + hdt);
        cs.Cs = hdt / (2.0 * tau);
        make_bgk(cs.en, rn, un, a.s.RT);
        make_bgk(cs.e, r, cs.u, a.s.RT);
        if constexpr (sizeof(T) == 4) {
            relate_bgk(cs.en, a.ref);
            relate_bgk(cs.e, a.ref);
        }
        return !isnan(r);
    } else {  // classic DUGKS: macros follow from f-tilde^{n+1} (dugks.cpp:313-335)
        cs.rho = rn;
        for (int i = 0; i < 3; ++i) cs.u[i] = un[i];
        cs.T = 0.0;
        cs.tau = cs.taun = a.s.tau;
        cs.cm = cs.C = cs.Cs = 0.0;
        make_bgk(cs.en, rn, un, a.s.RT);
        if constexpr (sizeof(T) == 4) relate_bgk(cs.en, a.ref);
        cs.e = cs.en;
        return true;
    }
}

// Face equilibrium pair in the stored form at velocity x (slot v).
template <int MODEL, int D, class T>
__device__ __forceinline__ void face_eq_store(const KArgs &a, int f, const T x[3], int v, T &gs, T &hs) {
    if constexpr (sizeof(T) == 4) {
        const EqF &e = a.face_eqf[f];
        if (MODEL == CDUGKS_SHAKHOV) shakhov_devf(e, x, D, a.s.K, a.fref_f[v], gs, hs);
        else {
            gs = bgk_devf(e, x);
            hs = 0.0f;
        }
    } else {
        eq_store<MODEL, D, T>(a, a.face_eq[f], x, v, gs, hs);
    }
}

template <int MODEL, int D, class T>
__device__ __forceinline__ void cell_eq_store(const KArgs &a, const EqParams &e, const EqF &ef, const T x[3], int v,
                                              T &gs, T &hs) {
    if constexpr (sizeof(T) == 4) {
        if (MODEL == CDUGKS_SHAKHOV) shakhov_devf(ef, x, D, a.s.K, a.fref_f[v], gs, hs);
        else {
            gs = bgk_devf(ef, x);
            hs = 0.0f;
        }
    } else {
        eq_store<MODEL, D, T>(a, e, x, v, gs, hs);
    }
}

// (4) cell flux + update; one wave per (cell, velocity tile). BND selects the
// cells that have a boundary face (launched separately, so the interior
// kernel carries no boundary code). TAB: tensor-product DVS with axis tables.
template <int MODEL, int D, class T, bool TAB, bool BND, bool LIM>
__device__ __forceinline__ void d_cell_update(const KArgs &a, const int b, int list_begin, int list_count) {
    extern __shared__ unsigned char lds_raw[];
    if (*a.halt) return;
    constexpr int WPB = BLOCK / 64;
    constexpr int MAXF = D == 2 ? 4 : 6;
    constexpr int MAXN = MAXF;
    constexpr int GS = 1 + MAXN, NG = 1 + MAXF;
    constexpr bool SH = MODEL == CDUGKS_SHAKHOV;
    const int nblk = (list_count + WPB - 1) / WPB;
    const int tile = b / nblk, cb = b - tile * nblk;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int li = cb * WPB + wave;
    const bool act = li < list_count;
    const int c = __builtin_amdgcn_readfirstlane(act ? a.cell_list[list_begin + li] : 0);
    T *tabs = reinterpret_cast<T *>(lds_raw) + size_t(wave) * (MAXF + 2) * a.nt;
    CellStep<MODEL, D> cs;
    int nfc = 0;
    if (act) {
        const bool ok = cell_macros<MODEL, D, T>(a, c, cs);
        if (tile == 0 && lane == 0) {
            if (MODEL != DUGKS_BGK) {
                a.rho_out[c] = cs.rho;
                for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = cs.u[i];
                if (SH) a.T_out[c] = cs.T;
                for (int k = 0; k < 5; ++k) a.fm[5 * c + k] = 0.0;
                a.fm[5 * c] = cs.phi[0];
                for (int i = 0; i < D; ++i) a.fm[5 * c + 1 + i] = cs.phi[1 + i];
                if (SH) a.fm[5 * c + 4] = cs.phi[D + 1];
            }
            if (!ok) flag_bad<MODEL>(a, c);
        }
        nfc = a.cs_nf[c];
        if (TAB) {
            const int nt = a.nt;
            for (int k = 0; k < nfc; ++k) {
                const T *src = static_cast<const T *>(a.face_tab) + size_t(a.cs_face[c * MAXF + k]) * nt;
                for (int t = lane; t < nt; t += 64) tabs[k * nt + t] = src[t];
            }
            for (int t = lane; t < nt; t += 64) {
                tabs[MAXF * nt + t] = tab_entry<T>(a, cs.en, t);
                tabs[(MAXF + 1) * nt + t] = tab_entry<T>(a, cs.e, t);
            }
        }
    }
    if (TAB) __syncthreads();
    if (!act) return;
    EqF en_f, e_f;
    if constexpr (sizeof(T) == 4) {
        en_f = to_eqf(cs.en);
        e_f = to_eqf(cs.e);
        if (!SH) {
            en_f.hk = float(cs.en.A);
            e_f.hk = float(cs.e.A);
        }
    }
    const T *fg = (const T *)a.fin_g;
    const T *fh = (const T *)a.fin_h;
    T *og = (T *)a.fout_g;
    T *oh = (T *)a.fout_h;
    const int nvp = a.s.nvp;
    const int nt = a.nt;
    const T hd = T(a.s.half_dt);
    const T dtv = T(cs.dt_v);
    const size_t row = size_t(c) * nvp;
    const int *sid = a.st_id + size_t(c) * NG * GS;
    const T *sa = coef_ptr<T>(a.st_a, a.st_af, size_t(c) * NG * MAXN * 3);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int v0 = tile * a.s.vt_c, v1 = min(v0 + a.s.vt_c, a.s.nv);
    for (int v = v0 + lane; v < v1; v += 64) {
        // every stencil value this velocity needs, loaded before any use
        T vg[NG * GS], vh[NG * GS];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            if (g <= nfc) {
#pragma unroll
                for (int j = 0; j < GS; ++j) {
                    const size_t r = size_t(sid[g * GS + j]) * nvp + v;
                    vg[g * GS + j] = fg[r];
                    if (SH) vh[g * GS + j] = fh[r];
                }
            }
        }
        const double4 X = a.vel_d[v];
        T x[3];
        if constexpr (sizeof(T) == 8) {
            x[0] = X.x;
            x[1] = X.y;
            x[2] = X.z;
        } else {
            const float4 xf = a.vel_f[v];
            x[0] = xf.x;
            x[1] = xf.y;
            x[2] = xf.z;
        }
        const int vt = TAB ? a.vtab[v] : 0;
        T gref = T(0);
        if constexpr (sizeof(T) == 4) gref = a.fref_f2[v].x;
        const T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
        const T fcg = vg[0], fch = SH ? vh[0] : T(0);
        T gcg[3] = {T(0), T(0), T(0)}, gch[3] = {T(0), T(0), T(0)};
        if (a.s.grad) {
            group_grad<D, MAXN, T>(vg, sa, gcg);
            if (SH) group_grad<D, MAXN, T>(vh, sa, gch);
        }
        T Fg = T(0), Fh = T(0);
        const double om_c = SH && LIM ? venkata_omega(a.s.venkata_k, a.cell_vol[c], D) : 0.0;
#pragma unroll
        for (int k = 0; k < MAXF; ++k) {
            if (k < nfc) {
                const int idx = c * MAXF + k;
                const int f = a.cs_face[idx];
                T sn;
                const int up = upwind<T>(a, f, v, x, sn);
                T gs, hs = T(0);
                if (SH && TAB) {
                    if constexpr (sizeof(T) == 4)
                        shakhov_tab_f(a.face_eqf[f], reinterpret_cast<const float *>(tabs) + k * nt, vt, x, D, a.s.K,
                                      gref, gs, hs);
                    else
                        shakhov_tab_d(a.face_eq[f], reinterpret_cast<const double *>(tabs) + k * nt, vt, x, D, a.s.K,
                                      gs, hs);
                } else {
                    face_eq_store<MODEL, D, T>(a, f, x, v, gs, hs);
                }
                T mix;
                if constexpr (sizeof(T) == 4) mix = a.face_eqf[f].mix;
                else mix = a.face_eq[f].mix;
                T dc[3], dn[3];
                if constexpr (sizeof(T) == 8) {
                    for (int i = 0; i < 3; ++i) {
                        dc[i] = a.cs_d[6 * idx + i];
                        dn[i] = a.cs_d[6 * idx + 3 + i];
                    }
                } else {
                    for (int i = 0; i < 3; ++i) {
                        dc[i] = a.cs_df[6 * idx + i];
                        dn[i] = a.cs_df[6 * idx + 3 + i];
                    }
                }
                T gv, hv = T(0);
                const bool bface = BND && a.cs_other[idx] < 0;
                if (bface) {
                    // boundary face: owner == neighbour == c
                    T Xg = dot_ext<D, T>(dc[0], dc[1], dc[2], xh, gcg);
                    const double eps_c = up > 0 ? 1.0e-10 : VSMALL;
                    if constexpr (SH && LIM) Xg *= T(group_phi<D, GS, T>(vg, dc, gcg, om_c, eps_c));
                    const T bg = up != 0 ? fcg + Xg : T(0.5) * (fcg + fcg + Xg + Xg);
                    const T gmx = lerp_store<T>(bg, gs, mix);
                    T hmx = T(0);
                    if (SH) {
                        T Xh = dot_ext<D, T>(dc[0], dc[1], dc[2], xh, gch);
                        if constexpr (LIM) Xh *= T(group_phi<D, GS, T>(vh, dc, gch, om_c, eps_c));
                        const T bh = up != 0 ? fch + Xh : T(0.5) * (fch + fch + Xh + Xh);
                        hmx = lerp_store<T>(bh, hs, mix);
                    }
                    const int bi = a.cs_bnd[idx];
                    if (bi >= 0) {
                        bnd_value<MODEL, D, T, LIM>(a, a.bnd[bi], f, c, v, double(sn), a.face_eq[f], gmx, hmx, gv,
                                                    hv);
                    } else {
                        gv = gmx;
                        hv = hmx;
                    }
                } else {
                    const bool own_c = a.cs_own[idx] != 0;
                    const bool up_c = own_c ? up > 0 : up < 0;
                    const bool up_o = own_c ? up < 0 : up > 0;
                    const T *vgo = vg + (1 + k) * GS;
                    const T *vho = vh + (1 + k) * GS;
                    T gog[3] = {T(0), T(0), T(0)}, goh[3] = {T(0), T(0), T(0)};
                    if (a.s.grad && __any(!up_c)) {
                        group_grad<D, MAXN, T>(vgo, sa + (1 + k) * MAXN * 3, gog);
                        if (SH) group_grad<D, MAXN, T>(vho, sa + (1 + k) * MAXN * 3, goh);
                    }
                    const T fog = vgo[0];
                    T Xcg = dot_ext<D, T>(dc[0], dc[1], dc[2], xh, gcg);
                    T Xog = dot_ext<D, T>(dn[0], dn[1], dn[2], xh, gog);
                    // limiter: eps 1e-10 for the face's owner when it is upwind
                    const double eps_c = own_c && up > 0 ? 1.0e-10 : VSMALL;
                    const double eps_o = !own_c && up > 0 ? 1.0e-10 : VSMALL;
                    double om_o = 0.0;
                    if constexpr (SH && LIM) {
                        om_o = venkata_omega(a.s.venkata_k, a.cell_vol[sid[(1 + k) * GS]], D);
                        Xcg *= T(group_phi<D, GS, T>(vg, dc, gcg, om_c, eps_c));
                        Xog *= T(group_phi<D, GS, T>(vgo, dn, gog, om_o, eps_o));
                    }
                    // dead-band: 0.5 (f_own + f_nei + X_own + X_nei), owner first
                    const T bg = up_c ? fcg + Xcg
                                      : (up_o ? fog + Xog
                                              : (own_c ? T(0.5) * (fcg + fog + Xcg + Xog)
                                                       : T(0.5) * (fog + fcg + Xog + Xcg)));
                    gv = lerp_store<T>(bg, gs, mix);
                    if (SH) {
                        const T foh = vho[0];
                        T Xch = dot_ext<D, T>(dc[0], dc[1], dc[2], xh, gch);
                        T Xoh = dot_ext<D, T>(dn[0], dn[1], dn[2], xh, goh);
                        if constexpr (LIM) {
                            Xch *= T(group_phi<D, GS, T>(vh, dc, gch, om_c, eps_c));
                            Xoh *= T(group_phi<D, GS, T>(vho, dn, goh, om_o, eps_o));
                        }
                        const T bh = up_c ? fch + Xch
                                          : (up_o ? foh + Xoh
                                                  : (own_c ? T(0.5) * (fch + foh + Xch + Xoh)
                                                           : T(0.5) * (foh + fch + Xoh + Xch)));
                        hv = lerp_store<T>(bh, hs, mix);
                    }
                }
                const T knA = sn * T(a.cs_sA[idx]);
                Fg += knA * gv;
                if (SH) Fh += knA * hv;
            }
        }
        if constexpr (sizeof(T) == 4) {
            // flux of f_ref itself: f_ref xi . sum(+-A n); non-zero where faces are not planar
            const double xs = X.x * a.cell_S[3 * c] + X.y * a.cell_S[3 * c + 1] + X.z * a.cell_S[3 * c + 2];
            const double2 fr = a.fref_d2[v];
            Fg += T(fr.x * xs);
            if (SH) Fh += T(fr.y * xs);
        }
        const double X0 = X.x, X1 = X.y, X2 = X.z;
        const double wv = X.w;
        if (SH) {
            T gsn, hsn, gs, hs;
            if (TAB) {
                if constexpr (sizeof(T) == 4) {
                    shakhov_tab_f(en_f, reinterpret_cast<const float *>(tabs) + MAXF * nt, vt, x, D, a.s.K, gref, gsn,
                                  hsn);
                    shakhov_tab_f(e_f, reinterpret_cast<const float *>(tabs) + (MAXF + 1) * nt, vt, x, D, a.s.K, gref,
                                  gs, hs);
                } else {
                    shakhov_tab_d(cs.en, reinterpret_cast<const double *>(tabs) + MAXF * nt, vt, x, D, a.s.K, gsn,
                                  hsn);
                    shakhov_tab_d(cs.e, reinterpret_cast<const double *>(tabs) + (MAXF + 1) * nt, vt, x, D, a.s.K, gs,
                                  hs);
                }
            } else {
                cell_eq_store<MODEL, D, T>(a, cs.en, en_f, x, v, gsn, hsn);
                cell_eq_store<MODEL, D, T>(a, cs.e, e_f, x, v, gs, hs);
            }
            const T cm = T(cs.cm), C = T(cs.C), Cs = T(cs.Cs);
            T g, h, og_, oh_;
            if constexpr (sizeof(T) == 8) {  // cdugks_shakhov.cpp:679-695 as written
                const T itau = T(1.0 / cs.tau), itaun = T(1.0 / cs.taun);
                const T gn = cm * gsn + (T(1) - cm) * fcg;
                const T hn = cm * hsn + (T(1) - cm) * fch;
                g = C * (gn + hd * (gs * itau + (gsn - gn) * itaun) - dtv * Fg);
                h = C * (hn + hd * (hs * itau + (hsn - hn) * itaun) - dtv * Fh);
                og_ = (T(1) - Cs) * g + Cs * gs;
                oh_ = (T(1) - Cs) * h + Cs * hs;
            } else {
                // the same update as increments of the stored deviation:
                // g = g_n + C [a (g_S - g_n) + a_n (g_S^n - g_n) - dt/V F],
                // a = h/tau, a_n = h/tau_n, C = 1/(1 + a)
                const T ai = T(a.s.half_dt / cs.tau), an = T(a.s.half_dt / cs.taun);
                const T gn = fcg + cm * (gsn - fcg);
                const T hn = fch + cm * (hsn - fch);
                g = gn + C * (ai * (gs - gn) + an * (gsn - gn) - dtv * Fg);
                h = hn + C * (ai * (hs - hn) + an * (hsn - hn) - dtv * Fh);
                og_ = g + Cs * (gs - g);
                oh_ = h + Cs * (hs - h);
            }
            double gfull = double(g), hfull = double(h);
            if constexpr (sizeof(T) == 4) {
                const double2 fr = a.fref_d2[v];
                gfull += fr.x;
                hfull += fr.y;
            }
            const double c0 = X0 - cs.u[0], c1 = X1 - cs.u[1], c2 = X2 - cs.u[2];
            const double t = 0.5 * wv * ((c0 * c0 + c1 * c1 + c2 * c2) * gfull + hfull);
            acc[0] += t * c0;
            acc[1] += t * c1;
            acc[2] += t * c2;
            og[row + v] = og_;
            oh[row + v] = oh_;
        } else if (MODEL == CDUGKS_BGK) {
            T feqn, feq, dummy;
            cell_eq_store<MODEL, D, T>(a, cs.en, en_f, x, v, feqn, dummy);
            cell_eq_store<MODEL, D, T>(a, cs.e, e_f, x, v, feq, dummy);
            const T cm = T(cs.cm), C = T(cs.C), ceq = T(cs.Cs);
            if constexpr (sizeof(T) == 8) {  // cdugks.cpp:325-332 as written
                const T itau = T(1.0 / cs.tau);
                const T fn = cm * feqn + (T(1) - cm) * fcg;
                og[row + v] = (T(1) - ceq) * (C * (fn + hd * (feq * itau + (feqn - fn) * itau) - dtv * Fg)) + ceq * feq;
            } else {
                const T ai = T(a.s.half_dt / cs.tau);
                const T fn = fcg + cm * (feqn - fcg);
                const T g = fn + C * (ai * (feq - fn) + ai * (feqn - fn) - dtv * Fg);
                og[row + v] = g + ceq * (feq - g);
            }
        } else {
            const double dt = a.s.dt, tau = a.s.tau;
            const T Cbt = T((dt + a.s.half_dt) / (a.s.half_dt - 2.0 * tau));
            T feq, dummy;
            cell_eq_store<MODEL, D, T>(a, cs.en, en_f, x, v, feq, dummy);
            T fnew;
            if constexpr (sizeof(T) == 8) {  // dugks.cpp:306-309 as written
                const T ft = (T(1) - Cbt) * fcg + Cbt * feq;
                fnew = (T(4) * fcg - ft) / T(3) - dtv * Fg;
            } else {
                fnew = fcg - Cbt * (feq - fcg) / T(3) - dtv * Fg;
            }
            og[row + v] = fnew;
            const double wf = wv * full_value<T>(a, fnew, 0, v);
            acc[0] += wf;
            acc[1] += wf * X0;
            acc[2] += wf * X1;
            acc[3] += wf * X2;
        }
    }
    if (MODEL == CDUGKS_BGK) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = seg_sum<64>(acc[k]);
    if (lane == 0) {
        double *dst = a.part_c + (size_t(c) * a.s.ntile_c + tile) * 4;
        for (int k = 0; k < 4; ++k) dst[k] = acc[k];
    }
}
template <int MODEL, int D, class T, bool TAB, bool BND, bool LIM>
__global__ __launch_bounds__(BLOCK) void k_cell_update(KArgs a, int list_begin, int list_count) {
    d_cell_update<MODEL, D, T, TAB, BND, LIM>(a, xcd_block(blockIdx.x, gridDim.x), list_begin, list_count);
}

// (5) q^{n+1} (Shakhov) or the DUGKS macros from the per-tile partials
template <int MODEL>
__device__ __forceinline__ void d_cell_finish(const KArgs &a, const int bid) {
    if (*a.halt) return;
    const int c = bid * BLOCK + threadIdx.x;
    if (c >= a.s.nc) return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const double *src = a.part_c + size_t(c) * a.s.ntile_c * 4;
    for (int t = 0; t < a.s.ntile_c; ++t)
        for (int k = 0; k < 4; ++k) s[k] += src[4 * t + k];
    if (MODEL == CDUGKS_SHAKHOV) {
        for (int k = 0; k < 3; ++k) a.q_out[3 * c + k] = s[k];
    } else {
        const double ir = 1.0 / s[0];
        a.rho_out[c] = s[0];
        for (int k = 0; k < 3; ++k) a.u_out[3 * c + k] = s[1 + k] * ir;
        if (isnan(s[0])) {
            atomicAdd(a.bad_count, 1);
            atomicMin(a.first_bad, c);
            atomicMin(a.bad_step, *a.step_ctr);
        }
    }
}
template <int MODEL>
__global__ __launch_bounds__(BLOCK) void k_cell_finish(KArgs a) {
    d_cell_finish<MODEL>(a, blockIdx.x);
}
template <int MODEL>
__global__ __launch_bounds__(BLOCK) void kb_cell_finish(const KArgs *__restrict__ tab, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    if (m.x < 0) return;  // padding of a short XCD list
    d_cell_finish<MODEL>(batch_args(tab, m.x), m.y);
}

// (6) classic DUGKS: f-tilde -> f-tilde-plus with the new equilibrium (dugks.cpp:329-343)
template <int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_dugks_tobp(KArgs a) {
    if (*a.halt) return;
    constexpr int IPB = BLOCK / LP;
    const int c = blockIdx.x * IPB + threadIdx.x / LP, li = threadIdx.x % LP;
    if (c >= a.s.nc) return;
    EqParams e;
    const double u[3] = {a.u_out[3 * c], a.u_out[3 * c + 1], a.u_out[3 * c + 2]};
    make_bgk(e, a.rho_out[c], u, a.s.RT);
    if constexpr (sizeof(T) == 4) relate_bgk(e, a.ref);
    const T C = T((a.s.dt + a.s.half_dt) / (2.0 * a.s.tau + a.s.dt));
    T *o = (T *)a.fout_g + size_t(c) * a.s.nvp;
    for (int v = li; v < a.s.nv; v += LP) {
        T x[3];
        load_xi<T>(a, v, x);
        T feq, dummy;
        eq_store<DUGKS_BGK, D, T>(a, e, x, v, feq, dummy);
        o[v] = lerp_store<T>(o[v], feq, C);
    }
}

// initial(): f from the patch values, macros from the discrete moments
// Stores v at entry i of a distribution array: with an extension array x
// (FP32X storage) the array holds FP32 words and v goes in as word and
// byte, else as the storage type.
template <class T>
__device__ __forceinline__ void st_val(T *p, unsigned char *x, size_t i, double v) {
    if (x) {
        float w;
        unsigned char e;
        x32_encode(v, w, e);
        reinterpret_cast<float *>(p)[i] = w;
        x[i] = e;
        return;
    }
    p[i] = T(v);
}

template <int MODEL, int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_init_eq(KArgs a, const double *rp, const double *up, const double *Tp) {
    constexpr int IPB = BLOCK / LP;
    const int c = blockIdx.x * IPB + threadIdx.x / LP, li = threadIdx.x % LP;
    const bool act = c < a.s.nc;
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // m0, m1(3), m2, m3(3)
    EqParams e;
    double u[3] = {0, 0, 0};
    if (act) {
        for (int i = 0; i < 3; ++i) u[i] = up[3 * c + i];
        const double zero[3] = {0, 0, 0};
        if (MODEL == CDUGKS_SHAKHOV) make_shakhov(e, rp[c], u, Tp[c], zero, a.s.R, a.s.Pr, D, a.s.K);
        else make_bgk(e, rp[c], u, a.s.RT);
        for (int v = li; v < a.s.nv; v += LP) {
            const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
            const double wv = a.w[v];
            double g, h = 0.0;
            const size_t o = size_t(c) * a.s.nvp + v;
            if (MODEL == CDUGKS_SHAKHOV) {
                shakhov_eq<double>(e, X, D, a.s.K, g, h);
                st_val((T *)a.fout_h, a.xout_h, o, h - (sizeof(T) == 4 || a.xout_h ? a.fref_d[a.s.nv + v] : 0.0));
            } else {
                g = bgk_eq<double>(e, X);
            }
            st_val((T *)a.fout_g, a.xout_g, o, g - (sizeof(T) == 4 || a.xout_g ? a.fref_d[v] : 0.0));
            const double kk = X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
            const double c0 = X[0] - u[0], c1 = X[1] - u[1], c2 = X[2] - u[2];
            const double cc = c0 * c0 + c1 * c1 + c2 * c2;
            m[0] += wv * g;
            m[1] += wv * g * X[0];
            m[2] += wv * g * X[1];
            m[3] += wv * g * X[2];
            m[4] += wv * (kk * g + h);
            const double t = wv * (cc * g + h);
            m[5] += t * c0;
            m[6] += t * c1;
            m[7] += t * c2;
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = seg_sum<LP>(m[k]);
    if (act && li == 0) {
        const double rho = m[0];
        const double ir = 1.0 / rho;
        double uu[3] = {m[1] * ir, m[2] * ir, m[3] * ir};
        a.rho_out[c] = rho;
        for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = uu[i];
        if (MODEL == CDUGKS_SHAKHOV) {
            const double rE = 0.5 * m[4];
            a.T_out[c] = (rE / rho - 0.5 * (uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2])) / a.s.Cv;
            for (int i = 0; i < 3; ++i) a.q_out[3 * c + i] = 0.5 * m[5 + i];
        }
    }
}

// read-np-data restart: f = g_S/h_S (or f_eq) of the given macros
template <int MODEL, int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_init_macros(KArgs a) {
    constexpr int IPB = BLOCK / LP;
    const int c = blockIdx.x * IPB + threadIdx.x / LP, li = threadIdx.x % LP;
    if (c >= a.s.nc) return;
    EqParams e;
    const double u[3] = {a.u_out[3 * c], a.u_out[3 * c + 1], a.u_out[3 * c + 2]};
    if (MODEL == CDUGKS_SHAKHOV) {
        const double q[3] = {a.q_out[3 * c], a.q_out[3 * c + 1], a.q_out[3 * c + 2]};
        make_shakhov(e, a.rho_out[c], u, a.T_out[c], q, a.s.R, a.s.Pr, D, a.s.K);
    } else {
        make_bgk(e, a.rho_out[c], u, a.s.RT);
    }
    for (int v = li; v < a.s.nv; v += LP) {
        const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
        const bool dev = sizeof(T) == 4 || a.xout_g;  // deviation from f_ref stored
        const double rg = dev ? a.fref_d[v] : 0.0;
        const size_t o = size_t(c) * a.s.nvp + v;
        T *g0 = (T *)a.fout_g, *h0 = (T *)a.fout_h;
        if (MODEL == CDUGKS_SHAKHOV) {
            double g, h;
            shakhov_eq<double>(e, X, D, a.s.K, g, h);
            st_val(g0, a.xout_g, o, g - rg);
            st_val(h0, a.xout_h, o, h - (dev ? a.fref_d[a.s.nv + v] : 0.0));
        } else {
            st_val(g0, a.xout_g, o, bgk_eq<double>(e, X) - rg);
        }
    }
}

// FP64 velocity-major [nv][nc] <-> device cell-major [nc][nvp] (32x32 tiles through LDS)
template <class T, bool TO_DEVICE>
__global__ __launch_bounds__(BLOCK) void k_transpose(const double *vm_in, double *vm_out, const T *cm_in, T *cm_out,
                                                     int nc, int nv, int nvp, const double *fref, const int *slot_of,
                                                     const unsigned char *x_in, unsigned char *x_out) {
    __shared__ double tile[32][33];
    const int c0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    if (TO_DEVICE) {
        for (int k = ty; k < 32; k += 8) {
            int v = v0 + k, c = c0 + tx;
            tile[k][tx] = (v < nv && c < nc) ? vm_in[size_t(v) * nc + c] : 0.0;
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            int c = c0 + k, v = v0 + tx;
            if (c < nc && v < nv) {
                const int sl = slot_of ? slot_of[v] : v;
                st_val(cm_out, x_out, size_t(c) * nvp + sl, tile[tx][k] - (fref ? fref[sl] : 0.0));
            }
        }
    } else {
        for (int k = ty; k < 32; k += 8) {
            int c = c0 + k, v = v0 + tx;
            const int sl = (v < nv && slot_of) ? slot_of[v] : v;
            double x = 0.0;
            if (c < nc && v < nv) {
                const size_t i = size_t(c) * nvp + sl;
                x = double(cm_in[i]);
                if constexpr (sizeof(T) == 4) {
                    if (x_in) x = x32_decode(float(cm_in[i]), x_in[i]);
                }
                x += fref ? fref[sl] : 0.0;
            }
            tile[tx][k] = x;
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            int v = v0 + k, c = c0 + tx;
            if (v < nv && c < nc) vm_out[size_t(v) * nc + c] = tile[k][tx];
        }
    }
}

// residual terms (field.cpp:222-249): per block partial sums of 8 values
__global__ __launch_bounds__(BLOCK) void k_residual(const double *rho, const double *u, const double *T,
                                                    const double *q, const double *rho0, const double *u0,
                                                    const double *T0, const double *q0, const double *vol,
                                                    double total, int nc, int shakhov, double *part) {
    __shared__ double lds[8 * (BLOCK / 64)];
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c < nc) {
        const double wv = vol[c] / total;
        s[0] = fabs((rho[c] - rho0[c]) / rho0[c]) * wv;
        for (int i = 0; i < 3; ++i) s[1 + i] = fabs((u[3 * c + i] - u0[3 * c + i]) / u0[3 * c + i]) * wv;
        if (shakhov) {
            s[4] = fabs((T[c] - T0[c]) / T0[c]) * wv;
            for (int i = 0; i < 3; ++i) s[5 + i] = fabs((q[3 * c + i] - q0[3 * c + i]) / q0[3 * c + i]) * wv;
        }
    }
    block_sum<8>(s, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) part[8 * blockIdx.x + k] = s[k];
}

__global__ void k_residual_final(const double *part, int nblk, double *out) {
    if (threadIdx.x < 8) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += part[8 * b + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

__global__ void k_reset_flags(int *bad_count, int *first_bad, int *bad_step, int *halt, int *step_ctr, int steps) {
    *bad_count = 0;
    step_ctr[1] = 0;  // masked cells repaired
    *first_bad = 0x7fffffff;
    *bad_step = 0x7fffffff;
    *halt = 0;
    *step_ctr = steps;
}

// Repair of the marked cells, after the step's update (one block, cells in
// ascending order as meso's cell loop meets them; CDUGKS_SHAKHOV::
// cell_value_interp, cdugks_shakhov.cpp:766-789, with fvmMesh::interp_IDW,
// src/mesh/mesh_numerical.cpp:51-94): rho, u, T and q become the average of
// the LSQ (face) neighbours weighted by 1/|dr|^2, where a neighbour before the
// cell in that loop has its new values (and, if it was repaired, its repaired
// q) and one after it its old values; g and h become the Shakhov equilibrium
// of those. q_out keeps the cell's own heat-flux sum, as meso's q field is
// overwritten by the sums after its loop. The repaired q is kept in q_in (the
// old buffer, not read again in this step) for the later cells.
template <int D, class T>
__global__ __launch_bounds__(BLOCK) void k_mask_repair(KArgs a) {
    if (*a.halt) return;
    __shared__ int flag[BLOCK];
    __shared__ double mac[8];  // rho, u[3], T, q[3]
    for (int base = 0; base < a.s.nc; base += BLOCK) {
        const int c0 = base + int(threadIdx.x);
        flag[threadIdx.x] = (c0 < a.s.nc && a.mask[c0] == 2) ? 1 : 0;
        if (__syncthreads_count(flag[threadIdx.x]) == 0) continue;
        for (int t = 0; t < BLOCK; ++t) {
            if (!flag[t]) continue;  // uniform: every thread reads the same entry
            const int c = base + t;
            if (threadIdx.x == 0) {
                double sw = 0.0, sr = 0.0, sT = 0.0, su[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
                for (int j = a.nb_ptr[c]; j < a.nb_ptr[c + 1]; ++j) {
                    const int n = a.nb_ids[j];
                    const double w = a.nb_w[j];
                    const bool upd = n < c;  // already met by meso's loop
                    sw += w;
                    sr += w * (upd ? a.rho_out[n] : a.rho_in[n]);
                    sT += w * (upd ? a.T_out[n] : a.T_in[n]);
                    for (int i = 0; i < 3; ++i) {
                        su[i] += w * (upd ? a.u_out[3 * n + i] : a.u_in[3 * n + i]);
                        sq[i] += w * a.q_in[3 * n + i];
                    }
                }
                mac[0] = sr / sw;
                mac[4] = sT / sw;
                for (int i = 0; i < 3; ++i) {
                    mac[1 + i] = su[i] / sw;
                    mac[5 + i] = sq[i] / sw;
                }
                a.rho_out[c] = mac[0];
                a.T_out[c] = mac[4];
                for (int i = 0; i < 3; ++i) {
                    a.u_out[3 * c + i] = mac[1 + i];
                    const_cast<double *>(static_cast<const double *>(a.q_in))[3 * c + i] = mac[5 + i];
                }
                const_cast<int *>(static_cast<const int *>(a.mask))[c] = 1;
            }
            __syncthreads();
            EqParams e;
            const double u[3] = {mac[1], mac[2], mac[3]}, q[3] = {mac[5], mac[6], mac[7]};
            make_shakhov(e, mac[0], u, mac[4], q, a.s.R, a.s.Pr, D, a.s.K);
            for (int v = threadIdx.x; v < a.s.nv; v += BLOCK) {
                const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
                double g, h;
                shakhov_eq<double>(e, X, D, a.s.K, g, h);
                const size_t o = size_t(c) * a.s.nvp + v;
                const bool dev = sizeof(T) == 4 || a.xout_g;
                st_val((T *)a.fout_g, a.xout_g, o, g - (dev ? a.fref_d[v] : 0.0));
                st_val((T *)a.fout_h, a.xout_h, o, h - (dev ? a.fref_d[a.s.nv + v] : 0.0));
            }
            __syncthreads();
        }
    }
}

// Last launch of every step: the device-side step index (read by the kernels
// that record the first bad step; a captured step pair then needs no host
// value that changes between replays).
__global__ void k_step_end(int *step_ctr) { *step_ctr += 1; }
// The same for every context of a batch: thread k advances context k's counter.
__global__ void kb_step_end(int *const *__restrict__ ctrs, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) *ctrs[k] += 1;
}

#include "patch_kernels.h"
#include "small_kernels.h"

}  // namespace

// ====================================================================== host
// Host copy of the mesh arrays the patch builder reads, kept so that the
// autotuner can rebuild the patch sets with other sizes.
struct PatchSrc {
    int nc = 0, nf = 0;
    std::vector<int> face_cells, cf_ptr, cf_ids, nb_ptr, nb_ids;
    std::vector<double> cell_pos, face_pos, face_area, face_normal, nb_a;
    std::vector<char> pown;  // cells the update patches own
    meso_patch::MeshView view() const {
        meso_patch::MeshView mv;
        mv.nc = nc;
        mv.nf = nf;
        mv.face_cells = face_cells.data();
        mv.cell_pos = cell_pos.data();
        mv.face_pos = face_pos.data();
        mv.face_area = face_area.data();
        mv.face_normal = face_normal.data();
        mv.cf_ptr = cf_ptr.data();
        mv.cf_ids = cf_ids.data();
        mv.nb_ptr = nb_ptr.data();
        mv.nb_ids = nb_ids.data();
        mv.nb_a = nb_a.data();
        return mv;
    }
};

// Device flags of a context: bad_count, first_bad, bad_step, halt, step
// counter, masked count
constexpr int FLAG_INTS = 6;

struct meso_hip_ctx {
    StepConst s{};
    int storage = MESO_STORAGE_FP64;
    int nd = 1;
    int lp = 64;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t bytes = 0;
    std::vector<void *> allocs;
    std::unordered_map<void *, size_t> alloc_bytes;
    PatchSrc psrc;
    // requested patch sizes (0 / -1: the defaults): update cells, moment
    // cells or faces, moment patch kind
    int preq_up = 0, preq_mom = 0, preq_cells = -1;
    std::vector<meso_mark_desc> marks_h;
    std::vector<int> face_bnd_h, bnd_face_h;
    int n_marks = 0;
    // device pointers
    int2 *face_cells = nullptr;
    double *face_n = nullptr, *face_d = nullptr, *cf_sA = nullptr, *nb_a = nullptr, *nb_w = nullptr,
           *cell_vol = nullptr;
    float *face_df = nullptr, *nb_af = nullptr, *xif = nullptr;
    double *cell_S = nullptr;
    int *cs_nf = nullptr, *cs_face = nullptr, *cs_other = nullptr, *cs_own = nullptr, *cs_bnd = nullptr,
        *lsq_cnt = nullptr, *lsq_nb = nullptr;
    double *cs_sA = nullptr, *cs_d = nullptr, *lsq_a = nullptr;
    float *cs_df = nullptr, *lsq_af = nullptr, *face_nf = nullptr;
    EqF *face_eqf = nullptr;
    double4 *vel_d = nullptr;
    float4 *vel_f = nullptr;
    double2 *fref_d2 = nullptr;
    float2 *fref_f2 = nullptr;
    int nt = 0;
    int *vtab = nullptr, *tab_ax = nullptr;
    double *tab_x = nullptr, *tab_w = nullptr;
    void *face_tab = nullptr;
    int *cell_list = nullptr, *st_id = nullptr, *fs_id = nullptr;
    double *st_a = nullptr, *fs_a = nullptr;
    float *st_af = nullptr, *fs_af = nullptr;
    // cell_list = [first group | second group]: without patch kernels the interior
    // cells, then the cells with a boundary face; with patch kernels the cells
    // of the update patches, then the cells with a mirror boundary
    int n_int = 0, n_bcell = 0;
    // patch kernels (2D, large velocity sets)
    bool patches = false;
    int n_up = 0, up_maxi = 0, up_maxf = 0, up_maxo = 0, up_maxr = 0, n_mp = 0, mp_maxi = 0, mp_maxf = 0,
        mp_maxr = 0;
    int p_chunks = 0, p_vchunk = 0;
    // small velocity sets (k_small_faces / k_small_cells): lanes per face or
    // cell (16, 32 or 64), 0 when the context has no such step
    int small_lp = 0;
    double *sflux = nullptr;  // [nd][nf][small_lp]
    int p_chunk_cap = 0;        // chunks the part_f / part_c buffers hold (>= p_chunks)
    // captured step pairs (meso_hip_set_graphs): gexec[p] runs two steps from
    // buffer parity p; valid while gver[p] == kver, which changes with every
    // change of the kernel arguments
    bool graphs = false;
    int kver = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int gver[2] = {-1, -1};
    size_t lds_up = 0, lds_mp = 0;
    bool upd_attr[2] = {false, false};  // the dynamic-LDS limit of k_patch_update<.., LIM> has been raised
    bool mp_attr[2][2] = {{false, false}, {false, false}};  // the same for k_patch_moments [cells][LIM]
    bool mp_cells = false;              // cell-based moment patches (k_patch_moments<..., true>)
    int mp_size = 0;                    // largest moment patch: owned cells or faces
    const void *up_items = nullptr, *mp_items = nullptr;  // meso_patch::ItemT<D>
    const void *up_faces = nullptr, *mp_faces = nullptr;  // meso_patch::FaceT<D>
    int *mp_nown = nullptr;
    int *up_item_ptr = nullptr, *up_face_ptr = nullptr, *up_nown = nullptr, *mp_item_ptr = nullptr,
        *mp_face_ptr = nullptr, *up_row_ptr = nullptr, *up_rows = nullptr, *mp_row_ptr = nullptr,
        *mp_rows = nullptr;
    void *crec = nullptr, *cell_tab = nullptr;
    void *fbar[2] = {nullptr, nullptr};  // patch path: face values g, h [nf][nvp] (storage type)
    int *slot_of = nullptr;  // meso's velocity index -> storage slot (null: identity)
    int *cf_ptr = nullptr, *cf_ids = nullptr, *nb_ptr = nullptr, *nb_ids = nullptr, *face_bnd = nullptr,
        *bnd_face = nullptr, *bnd_mark = nullptr, *mirror = nullptr;
    meso_mark_desc *marks = nullptr;
    double *xi = nullptr, *w = nullptr;
    void *f[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buffer][distribution]
    // FP32X storage (storage is then MESO_STORAGE_FP64, the kernels' arithmetic,
    // with ext set): f holds FP32 words of the deviation from f_ref and fx the
    // extension bytes, [buffer][distribution]
    bool ext = false;
    unsigned char *fx[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    // masked cells (meso_hip_set_mask): 1 per cell of a group with `mask switch True`
    int *mask = nullptr;
    bool has_mask = false;
    int masked_seen = 0;  // repaired cells counted up to the last report
    double *rho[2] = {nullptr, nullptr}, *u[2] = {nullptr, nullptr}, *T[2] = {nullptr, nullptr},
           *q[2] = {nullptr, nullptr};
    double *rho_res = nullptr, *u_res = nullptr, *T_res = nullptr, *q_res = nullptr;
    double *part_f = nullptr, *face_flux = nullptr, *part_c = nullptr, *fm = nullptr, *res_part = nullptr,
           *res_out = nullptr;
    EqParams *face_eq = nullptr;
    BndParams *bnd = nullptr;
    int *flags = nullptr;  // bad_count, first_bad, bad_step, halt, step counter, masked count
    // reference equilibrium of the FP32 deviation storage
    EqParams ref{};
    double *fref_d = nullptr;
    float *fref_f = nullptr;
    double mref[17] = {};
    double rflux[5][3] = {};
    std::vector<double> xi_h, w_h;
    int *flags_host = nullptr;
    int cur = 0;           // buffer holding f^n / macros^n
    int steps_done = 0;
    // profiling
    bool profiling = false;
    struct Rec {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<std::pair<const char *, std::pair<double, int>>> totals;
};

namespace {

thread_local std::string g_hip_err;

int set_err(meso_hip_ctx *c, int code, const std::string &m) {
    if (c) c->err = m;
    g_hip_err = m;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return set_err(ctx, MESO_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <class P>
int dalloc(meso_hip_ctx *c, P **p, size_t n) {
    void *q = nullptr;
    size_t bytes = std::max<size_t>(n * sizeof(P), 16);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return set_err(c, MESO_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(q, 0, bytes, c->stream);
    if (e != hipSuccess) return set_err(c, MESO_E_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    c->alloc_bytes[q] = bytes;
    c->bytes += int64_t(bytes);
    *p = reinterpret_cast<P *>(q);
    return MESO_OK;
}

// Releases one allocation of the context (the caller has synchronised).
template <class P>
void dfree(meso_hip_ctx *c, P *&p) {
    if (!p) return;
    void *q = const_cast<void *>(reinterpret_cast<const void *>(p));
    auto it = std::find(c->allocs.begin(), c->allocs.end(), q);
    if (it != c->allocs.end()) c->allocs.erase(it);
    auto b = c->alloc_bytes.find(q);
    if (b != c->alloc_bytes.end()) {
        c->bytes -= int64_t(b->second);
        c->alloc_bytes.erase(b);
    }
    (void)hipFree(q);
    p = nullptr;
}

template <class P>
int upload(meso_hip_ctx *c, P **p, const P *src, size_t n) {
    int rc = dalloc(c, p, n);
    if (rc) return rc;
    if (n) HIP_TRY(c, hipMemcpyAsync(*p, src, n * sizeof(P), hipMemcpyHostToDevice, c->stream));
    return MESO_OK;
}

int lanes_for(int nv) {
    if (nv >= 48) return 64;
    if (nv >= 24) return 32;
    return 16;
}

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

void tiles_for(int nv, int lp, int target, int &nt, int &vt) {
    nt = std::max(1, (nv + target - 1) / target);
    int per = (nv + nt - 1) / nt;
    vt = ((per + lp - 1) / lp) * lp;
    nt = (nv + vt - 1) / vt;
}

KArgs make_args(meso_hip_ctx *c) {
    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.s = c->s;
    a.face_cells = c->face_cells;
    a.face_n = c->face_n;
    a.face_d = c->face_d;
    a.face_df = c->face_df;
    a.cf_ptr = c->cf_ptr;
    a.cf_ids = c->cf_ids;
    a.cf_sA = c->cf_sA;
    a.nb_ptr = c->nb_ptr;
    a.nb_ids = c->nb_ids;
    a.nb_a = c->nb_a;
    a.nb_w = c->nb_w;
    a.nb_af = c->nb_af;
    a.cell_vol = c->cell_vol;
    a.cell_S = c->cell_S;
    a.cs_nf = c->cs_nf;
    a.cs_face = c->cs_face;
    a.cs_other = c->cs_other;
    a.cs_own = c->cs_own;
    a.cs_bnd = c->cs_bnd;
    a.cs_sA = c->cs_sA;
    a.cs_d = c->cs_d;
    a.cs_df = c->cs_df;
    a.lsq_cnt = c->lsq_cnt;
    a.lsq_nb = c->lsq_nb;
    a.lsq_a = c->lsq_a;
    a.lsq_af = c->lsq_af;
    a.face_nf = c->face_nf;
    a.face_eqf = c->face_eqf;
    a.vel_d = c->vel_d;
    a.vel_f = c->vel_f;
    a.fref_d2 = c->fref_d2;
    a.fref_f2 = c->fref_f2;
    a.nt = c->nt;
    a.vtab = c->vtab;
    a.tab_x = c->tab_x;
    a.tab_ax = c->tab_ax;
    a.tab_w = c->tab_w;
    a.face_tab = c->face_tab;
    a.cell_list = c->cell_list;
    a.st_id = c->st_id;
    a.st_a = c->st_a;
    a.st_af = c->st_af;
    a.fs_id = c->fs_id;
    a.fs_a = c->fs_a;
    a.fs_af = c->fs_af;
    a.face_bnd = c->face_bnd;
    a.bnd_face = c->bnd_face;
    a.bnd_mark = c->bnd_mark;
    a.marks = c->marks;
    a.xi = c->xi;
    a.xif = c->xif;
    a.w = c->w;
    a.mirror = c->mirror;
    const int i = c->cur, o = 1 - c->cur;
    a.fin_g = c->f[i][0];
    a.fin_h = c->f[i][1];
    a.fout_g = c->f[o][0];
    a.fout_h = c->f[o][1];
    a.xin_g = c->fx[i][0];
    a.xin_h = c->fx[i][1];
    a.xout_g = c->fx[o][0];
    a.xout_h = c->fx[o][1];
    a.mask = c->has_mask ? c->mask : nullptr;
    a.masked = c->flags + 5;
    a.fbar_g = c->fbar[0];
    a.fbar_h = c->fbar[1];
    a.rho_in = c->rho[i];
    a.u_in = c->u[i];
    a.T_in = c->T[i];
    a.q_in = c->q[i];
    a.rho_out = c->rho[o];
    a.u_out = c->u[o];
    a.T_out = c->T[o];
    a.q_out = c->q[o];
    a.part_f = c->part_f;
    a.face_eq = c->face_eq;
    a.face_flux = c->face_flux;
    a.bnd = c->bnd;
    a.part_c = c->part_c;
    a.fm = c->fm;
    a.sflux = c->sflux;
    a.bad_count = c->flags;
    a.first_bad = c->flags + 1;
    a.bad_step = c->flags + 2;
    a.halt = c->flags + 3;
    a.step_ctr = c->flags + 4;
    a.n_up = c->n_up;
    a.up_maxi = c->up_maxi;
    a.up_maxf = c->up_maxf;
    a.up_maxo = c->up_maxo;
    a.up_items = c->up_items;
    a.up_faces = c->up_faces;
    a.up_item_ptr = c->up_item_ptr;
    a.up_face_ptr = c->up_face_ptr;
    a.up_nown = c->up_nown;
    a.up_maxr = c->up_maxr;
    a.up_row_ptr = c->up_row_ptr;
    a.up_rows = c->up_rows;
    a.mp_maxr = c->mp_maxr;
    a.mp_row_ptr = c->mp_row_ptr;
    a.mp_rows = c->mp_rows;
    a.n_mp = c->n_mp;
    a.mp_maxi = c->mp_maxi;
    a.mp_maxf = c->mp_maxf;
    a.mp_items = c->mp_items;
    a.mp_nown = c->mp_nown;
    a.mp_faces = c->mp_faces;
    a.mp_item_ptr = c->mp_item_ptr;
    a.mp_face_ptr = c->mp_face_ptr;
    a.p_chunks = c->p_chunks;
    a.p_vchunk = c->p_vchunk;
    a.n_int_cells = c->n_int;
    a.crec = c->crec;
    a.cell_tab = c->cell_tab;
    a.ref = c->ref;
    a.fref_d = c->fref_d;
    a.fref_f = c->fref_f;
    std::memcpy(a.mref, c->mref, sizeof(a.mref));
    std::memcpy(a.rflux, c->rflux, sizeof(a.rflux));
    return a;
}

// Reference equilibrium f_ref of the FP32 deviation storage: the Shakhov
// (Maxwellian) pair, or the BGK polynomial, of (rho_ref, u = 0, T_ref); its
// discrete raw moments in the face-moment column layout; and its flux vectors
// sum_v w xi psi_k f_ref. Storage then holds f - f_ref, whose rounding error is
// proportional to the deviation from rest instead of to f.
int set_reference(meso_hip_ctx *c, double rho_ref, double T_ref) {
    const StepConst &s = c->s;
    const int nv = s.nv, D = s.D;
    const bool shak = s.model == MESO_MODEL_CDUGKS_SHAKHOV;
    const double zero[3] = {0.0, 0.0, 0.0};
    if (shak) make_shakhov(c->ref, rho_ref, zero, T_ref, zero, s.R, s.Pr, D, s.K);
    else make_bgk(c->ref, rho_ref, zero, s.RT);
    std::vector<double> fr(2 * size_t(nv), 0.0);
    std::vector<float> ff(2 * size_t(nv), 0.0f);
    std::memset(c->mref, 0, sizeof(c->mref));
    std::memset(c->rflux, 0, sizeof(c->rflux));
    for (int v = 0; v < nv; ++v) {
        const double *x = &c->xi_h[3 * v];
        const double xx = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        double g, h = 0.0;
        if (shak) {
            g = c->ref.A * std::exp(-c->ref.B * xx);
            h = c->ref.hk * g;
        } else {
            g = rho_ref;
        }
        fr[v] = g;
        fr[nv + v] = h;
        ff[v] = float(g);
        ff[nv + v] = float(h);
        const double wg = c->w_h[v] * g, wh = c->w_h[v] * h;
        int k = 0;
        c->mref[k++] += wg;
        for (int i = 0; i < D; ++i) c->mref[k++] += wg * x[i];
        for (int i = 0; i < D; ++i)
            for (int j = i; j < D; ++j) c->mref[k++] += wg * x[i] * x[j];
        if (shak) {
            for (int i = 0; i < D; ++i) c->mref[k++] += wg * xx * x[i];
            c->mref[k++] += wh;
            for (int i = 0; i < D; ++i) c->mref[k++] += wh * x[i];
        }
        for (int d = 0; d < 3; ++d) {
            c->rflux[0][d] += wg * x[d];
            for (int i = 0; i < D; ++i) c->rflux[1 + i][d] += wg * x[d] * x[i];
            if (shak) c->rflux[D + 1][d] += x[d] * (wg * xx + wh);
        }
    }
    HIP_TRY(c, hipMemcpyAsync(c->fref_d, fr.data(), fr.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->fref_f, ff.data(), ff.size() * 4, hipMemcpyHostToDevice, c->stream));
    std::vector<double2> fr2(nv);
    std::vector<float2> ff2(nv);
    for (int v = 0; v < nv; ++v) {
        fr2[v] = make_double2(fr[v], fr[nv + v]);
        ff2[v] = make_float2(ff[v], ff[nv + v]);
    }
    HIP_TRY(c, hipMemcpyAsync(c->fref_d2, fr2.data(), nv * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->fref_f2, ff2.data(), nv * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int set_reference_from(meso_hip_ctx *c, const double *rho, const double *T) {
    ++c->kver;  // the reference state is a kernel argument
    if (c->storage != MESO_STORAGE_FP32 && !c->ext) return MESO_OK;
    const int nc = c->s.nc;
    double r = 0.0, t = 0.0;
    for (int i = 0; i < nc; ++i) {
        r += rho[i];
        t += T ? T[i] : 0.0;
    }
    r /= nc;
    t = T ? t / nc : 1.0;
    if (!(r > 0.0) || !(t > 0.0)) return set_err(c, MESO_E_INVALID, "reference state needs positive density and T");
    return set_reference(c, r, t);
}

struct Launch {
    meso_hip_ctx *c;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    Launch(meso_hip_ctx *c_, const char *n) : c(c_), name(n) {
        if (c->profiling) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Launch() {
        if (c->profiling) {
            (void)hipEventRecord(b, c->stream);
            c->recs.push_back({name, a, b});
        }
    }
};

#define LAUNCH(ctx, name, kern, grid, ...)                                                           \
    do {                                                                                             \
        Launch l_(ctx, name);                                                                        \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, (ctx)->stream, __VA_ARGS__);            \
    } while (0)

#define LAUNCH_SH(ctx, name, kern, grid, shmem, ...)                                                 \
    do {                                                                                             \
        Launch l_(ctx, name);                                                                        \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), shmem, (ctx)->stream, __VA_ARGS__);        \
    } while (0)

#define LAUNCH_P(ctx, name, kern, grid, shmem, ...)                                                  \
    do {                                                                                             \
        Launch l_(ctx, name);                                                                        \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCKP), shmem, (ctx)->stream, __VA_ARGS__);       \
    } while (0)

// Whether the face stage is k_face_eq_sep: Shakhov with axis tables and
// factorised weights.
bool face_sep(const meso_hip_ctx *c) { return c->s.model == CDUGKS_SHAKHOV && c->nt > 0 && c->tab_w != nullptr; }

// Step with the LDS-staged patch kernels for the face moments and the
// interior cells; boundary faces and mirror-boundary cells keep their kernels.
template <int MODEL, class T, bool TAB, int D, bool LIM, bool X>
int step_patch(meso_hip_ctx *c, KArgs a) {
    const StepConst &s = c->s;
    constexpr int WPB = BLOCK / 64;
    constexpr int VTM = sizeof(T) == 4 ? 64 : 32;  // velocities per tile: moments
    constexpr int VTU = upd_vt<T, D>();             // velocities per tile: update
    constexpr int MAXF = D == 2 ? 4 : 6;
    constexpr bool XM = X && sizeof(T) == 8;  // FP32E moments read the words only (the FP32 kernel)
    const int fwblk = (s.nf + WPB - 1) / WPB;
    if (c->mp_cells) {
        if (!c->mp_attr[1][LIM] && c->lds_mp > 65536) {
            HIP_TRY(c, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_patch_moments<MODEL, T, VTM, D, LIM, true, XM>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(c->lds_mp)));
            c->mp_attr[1][LIM] = true;
        }
        LAUNCH_P(c, "face_moments", (k_patch_moments<MODEL, T, VTM, D, LIM, true, XM>), c->n_mp * c->p_chunks, c->lds_mp,
                 a);
    } else {
        if (!c->mp_attr[0][LIM] && c->lds_mp > 65536) {
            HIP_TRY(c, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_patch_moments<MODEL, T, VTM, D, LIM, false, XM>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(c->lds_mp)));
            c->mp_attr[0][LIM] = true;
        }
        LAUNCH_P(c, "face_moments", (k_patch_moments<MODEL, T, VTM, D, LIM, false, XM>), c->n_mp * c->p_chunks,
                 c->lds_mp, a);
    }
    if (MODEL == CDUGKS_SHAKHOV && TAB && face_sep(c))
        LAUNCH(c, "face_eq", (k_face_eq_sep<D, T, FEQ_LP>), (s.nf + BLOCK / FEQ_LP - 1) / (BLOCK / FEQ_LP), a);
    else
        LAUNCH_SH(c, "face_eq", (k_face_eq<MODEL, D, T, TAB>), fwblk, TAB ? size_t(WPB) * c->nt * sizeof(T) : 0, a);
    if (s.nb > 0) LAUNCH_SH(c, "boundary", (k_boundary<MODEL, D, T, LIM>), s.nb, 0, a);
    if (c->n_int > 0) {
        LAUNCH(c, "cell_prep", (k_cell_prep<MODEL, T, TAB, D>), (c->n_int + WPB - 1) / WPB, a);
        if (!c->upd_attr[LIM] && c->lds_up > 65536) {
            HIP_TRY(c, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_patch_update<MODEL, T, VTU, TAB, D, LIM, X>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(c->lds_up)));
            c->upd_attr[LIM] = true;
        }
        LAUNCH_P(c, "cell_update", (k_patch_update<MODEL, T, VTU, TAB, D, LIM, X>), c->n_up * c->p_chunks, c->lds_up,
                 a);
    }
    const size_t lds = TAB ? size_t(WPB) * (MAXF + 2) * c->nt * sizeof(T) : 0;
    if (c->n_bcell > 0)
        LAUNCH_SH(c, "cell_update_bnd", (k_cell_update<MODEL, D, T, TAB, true, LIM>),
                  ((c->n_bcell + WPB - 1) / WPB) * s.ntile_c, lds, a, c->n_int, c->n_bcell);
    if (MODEL == CDUGKS_SHAKHOV) LAUNCH(c, "cell_finish", (k_cell_finish<MODEL>), (s.nc + BLOCK - 1) / BLOCK, a);
    if (MODEL == CDUGKS_SHAKHOV && c->has_mask) LAUNCH(c, "mask_repair", (k_mask_repair<D, T>), 1, a);
    HIP_TRY(c, hipGetLastError());
    return MESO_OK;
}

template <int MODEL, int D, class T, bool TAB, bool LIM>
int step_body(meso_hip_ctx *c, KArgs a) {
    const StepConst &s = c->s;
    constexpr int WPB = BLOCK / 64;
    constexpr int MF = MODEL == CDUGKS_SHAKHOV ? CDUGKS_SHAKHOV : CDUGKS_BGK;
    constexpr int MAXF = D == 2 ? 4 : 6;
    if constexpr (MODEL != DUGKS_BGK) {
        if (c->patches) {
            // FP32X / FP32E: the FP64 / FP32 kernels on FP32 words + extension bytes
            if (c->ext) return step_patch<MODEL, T, TAB, D, LIM, true>(c, a);
            return step_patch<MODEL, T, TAB, D, LIM, false>(c, a);
        }
    }
    const int fwblk = (s.nf + WPB - 1) / WPB;
    LAUNCH_SH(c, "face_moments", (k_face_moments<MF, D, T, LIM>), fwblk * s.ntile_f, 0, a);
    if (MF == CDUGKS_SHAKHOV && TAB && face_sep(c))
        LAUNCH(c, "face_eq", (k_face_eq_sep<D, T, FEQ_LP>), (s.nf + BLOCK / FEQ_LP - 1) / (BLOCK / FEQ_LP), a);
    else
        LAUNCH_SH(c, "face_eq", (k_face_eq<MF, D, T, TAB>), fwblk, TAB ? size_t(WPB) * c->nt * sizeof(T) : 0, a);
    if (s.nb > 0) LAUNCH_SH(c, "boundary", (k_boundary<MF, D, T, LIM>), s.nb, 0, a);
    const size_t lds = TAB ? size_t(WPB) * (MAXF + 2) * c->nt * sizeof(T) : 0;
    if (c->n_int > 0)
        LAUNCH_SH(c, "cell_update", (k_cell_update<MODEL, D, T, TAB, false, LIM>),
                  ((c->n_int + WPB - 1) / WPB) * s.ntile_c, lds, a, 0, c->n_int);
    if (c->n_bcell > 0)
        LAUNCH_SH(c, "cell_update_bnd", (k_cell_update<MODEL, D, T, TAB, true, LIM>),
                  ((c->n_bcell + WPB - 1) / WPB) * s.ntile_c, lds, a, c->n_int, c->n_bcell);
    if (MODEL == CDUGKS_SHAKHOV) LAUNCH(c, "cell_finish", (k_cell_finish<MODEL>), (s.nc + BLOCK - 1) / BLOCK, a);
    if (MODEL == CDUGKS_SHAKHOV && c->has_mask) LAUNCH(c, "mask_repair", (k_mask_repair<D, T>), 1, a);
    if (MODEL == DUGKS_BGK) {
        LAUNCH(c, "cell_finish", (k_cell_finish<MODEL>), (s.nc + BLOCK - 1) / BLOCK, a);
        LAUNCH(c, "dugks_tobp", (k_dugks_tobp<D, T, 64>), (s.nc + WPB - 1) / WPB, a);
    }
    HIP_TRY(c, hipGetLastError());
    return MESO_OK;
}

// The limiter (Shakhov only) selects kernels compiled with it.
template <int MODEL, int D, class T, bool TAB>
int step_impl(meso_hip_ctx *c, KArgs a) {
    if constexpr (MODEL == CDUGKS_SHAKHOV) {
        if (c->s.limiter) return step_body<MODEL, D, T, TAB, true>(c, a);
    }
    return step_body<MODEL, D, T, TAB, false>(c, a);
}

// Whether the context's step is the small-velocity-set step (two launches;
// the face kernel advances the step counter).
bool uses_small(const meso_hip_ctx *c) { return c->small_lp > 0 && !c->s.limiter; }

template <int MODEL, int D, int LP>
int step_small(meso_hip_ctx *c, KArgs a) {
    constexpr int SPB = small_spb<LP>();
    LAUNCH(c, "small_faces", (k_small_faces<MODEL, D, LP>), (c->s.nf + SPB - 1) / SPB, a);
    LAUNCH(c, "small_cells", (k_small_cells<MODEL, D, LP>), (c->s.nc + SPB - 1) / SPB, a);
    if (MODEL == CDUGKS_SHAKHOV && c->has_mask) LAUNCH(c, "mask_repair", (k_mask_repair<D, double>), 1, a);
    HIP_TRY(c, hipGetLastError());
    return MESO_OK;
}

template <int MODEL, int D>
int step_small_lp(meso_hip_ctx *c, KArgs a) {
    switch (c->small_lp) {
        case 16: return step_small<MODEL, D, 16>(c, a);
        case 32: return step_small<MODEL, D, 32>(c, a);
        default: return step_small<MODEL, D, 64>(c, a);
    }
}

int step_dispatch(meso_hip_ctx *c, KArgs a) {
    const int m = c->s.model, D = c->s.D;
    if (uses_small(c)) {
        if (m == CDUGKS_SHAKHOV) return D == 2 ? step_small_lp<CDUGKS_SHAKHOV, 2>(c, a) : step_small_lp<CDUGKS_SHAKHOV, 3>(c, a);
        if (m == CDUGKS_BGK) return D == 2 ? step_small_lp<CDUGKS_BGK, 2>(c, a) : step_small_lp<CDUGKS_BGK, 3>(c, a);
    }
    const bool f32 = c->storage == MESO_STORAGE_FP32;
    const bool tab = c->nt > 0;
#define MESO_S(M, DD)                                                                                \
    if (m == M && D == DD) {                                                                         \
        if (M == CDUGKS_SHAKHOV && tab)                                                              \
            return f32 ? step_impl<M, DD, float, true>(c, a) : step_impl<M, DD, double, true>(c, a);  \
        return f32 ? step_impl<M, DD, float, false>(c, a) : step_impl<M, DD, double, false>(c, a);   \
    }
    MESO_S(CDUGKS_SHAKHOV, 2)
    MESO_S(CDUGKS_SHAKHOV, 3)
    MESO_S(CDUGKS_BGK, 2)
    MESO_S(CDUGKS_BGK, 3)
    MESO_S(DUGKS_BGK, 2)
    MESO_S(DUGKS_BGK, 3)
#undef MESO_S
    return set_err(c, MESO_E_INVALID, "unsupported model/dimension combination");
}

template <int MODEL, int D, class T, int LP>
int init_impl(meso_hip_ctx *c, KArgs a, const double *rp, const double *up, const double *Tp, bool from_macros) {
    constexpr int IPB = BLOCK / LP;
    const int cblk = (c->s.nc + IPB - 1) / IPB;
    if (from_macros) LAUNCH(c, "init_macros", (k_init_macros<MODEL, D, T, LP>), cblk, a);
    else LAUNCH(c, "init_eq", (k_init_eq<MODEL, D, T, LP>), cblk, a, rp, up, Tp);
    HIP_TRY(c, hipGetLastError());
    return MESO_OK;
}

// Dispatch over (model, dimension, storage, lanes).
template <template <int, int, class, int> class F, class... Args>
int dispatch(meso_hip_ctx *c, Args... args) {
    const int m = c->s.model, D = c->s.D, lp = c->lp;
    const bool f32 = c->storage == MESO_STORAGE_FP32;
#define MESO_CASE(M, DD, TT, L) \
    if (m == M && D == DD && (std::is_same<TT, float>::value == f32) && lp == L) return F<M, DD, TT, L>::run(c, args...);
#define MESO_LP(M, DD, TT) MESO_CASE(M, DD, TT, 16) MESO_CASE(M, DD, TT, 32) MESO_CASE(M, DD, TT, 64)
#define MESO_T(M, DD) MESO_LP(M, DD, float) MESO_LP(M, DD, double)
#define MESO_D(M) MESO_T(M, 2) MESO_T(M, 3)
    MESO_D(CDUGKS_SHAKHOV)
    MESO_D(CDUGKS_BGK)
    MESO_D(DUGKS_BGK)
#undef MESO_D
#undef MESO_T
#undef MESO_LP
#undef MESO_CASE
    return set_err(c, MESO_E_INVALID, "unsupported model/dimension/storage combination");
}

template <int M, int D, class T, int L>
struct InitF {
    static int run(meso_hip_ctx *c, KArgs a, const double *rp, const double *up, const double *Tp, bool fm) {
        return init_impl<M, D, T, L>(c, a, rp, up, Tp, fm);
    }
};

void set_const(StepConst &s, const meso_step_params &p) {
    s.model = p.model;
    s.D = p.dim;
    s.K = p.K;
    s.grad = p.gradient_switch ? 1 : 0;
    s.limiter = p.model == MESO_MODEL_CDUGKS_SHAKHOV && p.limiter_switch ? 1 : 0;
    s.venkata_k = p.venkata_k;
    s.dt = p.dt;
    s.half_dt = 0.5 * p.dt;
    s.Pr = p.Pr;
    s.R = p.R;
    s.T0 = p.T0;
    s.mu0 = p.mu0;
    s.omega = p.omega;
    s.gamma = (p.K + 5.0) / (p.K + 3.0);
    s.Cv = (p.K + 3.0) * p.R * 0.5;
    s.RT = p.RT;
    s.tau = p.tau;
}

int sync_flags(meso_hip_ctx *c) {
    HIP_TRY(c, hipMemcpyAsync(c->flags_host, c->flags, FLAG_INTS * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int reset_flags(meso_hip_ctx *c) {
    hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(1), 0, c->stream, c->flags, c->flags + 1, c->flags + 2,
                       c->flags + 3, c->flags + 4, c->steps_done);
    HIP_TRY(c, hipGetLastError());
    c->masked_seen = 0;
    return MESO_OK;
}

int copy_residual_base(meso_hip_ctx *c) {
    const size_t nc = c->s.nc;
    const int i = c->cur;
    HIP_TRY(c, hipMemcpyAsync(c->rho_res, c->rho[i], nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->u_res, c->u[i], 3 * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->T_res, c->T[i], nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->q_res, c->q[i], 3 * nc * 8, hipMemcpyDeviceToDevice, c->stream));
    return MESO_OK;
}

// Velocities per block of the patch kernels (a multiple of 64). At most 1280
// (the records a block stages once are amortised over its chunk's tiles;
// 200^2 x 101^2 cavity, every storage: 1-1.7% faster per step than 2560);
// and enough chunks that the smaller patch grid has 16 blocks per CU, so that
// small cases fill the GPU (as far as 64-velocity chunks allow).
int default_vchunk(int nv, int npatch, int ncu) {
    const int most = (nv + 63) / 64;
    int chunks = std::max((nv + 1279) / 1280, (16 * ncu + npatch - 1) / std::max(npatch, 1));
    chunks = std::max(1, std::min(chunks, most));
    return ((nv + chunks - 1) / chunks + 63) / 64 * 64;
}

void set_vchunk(meso_hip_ctx *c, int vchunk) {
    ++c->kver;
    c->p_vchunk = vchunk;
    c->p_chunks = (c->s.nv + vchunk - 1) / vchunk;
    c->s.ntile_c = c->p_chunks;
    // cell-based moment patches keep a partial sum per face side and chunk
    c->s.ntile_f = (c->mp_cells ? 2 : 1) * c->p_chunks;
    c->s.vt_c = c->s.vt_f = c->p_vchunk;
}

// Spatial patches for the LDS-staged kernels: builds the update and moment
// patch sets, shrinking the patches until each block fits its LDS budget
// (half the LDS of a CU, so that two blocks share a CU). Sets the patch fields
// of c; returns false when the mesh does not fit the patch records.
template <int D>
bool setup_patches(meso_hip_ctx *c, int storage, bool keep_chunk, meso_patch::PatchSet<D> &ups,
                   meso_patch::PatchSet<D> &mps) {
    const meso_patch::MeshView mv = c->psrc.view();
    const std::vector<char> &pown = c->psrc.pown;
    const int nc = mv.nc, nv = c->s.nv;
    // velocities per tile (must match step_patch): moment kernel 64 for FP32
    // storage (one item per wave), 32 for FP64 (two items per wave); update
    // kernel upd_vt
    const int VTM = storage == MESO_STORAGE_FP32 ? 64 : 32;
    const int VTU = storage == MESO_STORAGE_FP32 ? upd_vt<float, D>() : upd_vt<double, D>();
    const int IPWM = 64 / VTM, IPWU = 64 / VTU;
    const size_t es = storage == MESO_STORAGE_FP32 ? 4 : 8;
    const int ND = c->nd;
    const int ncol = D == 2 ? (ND == 2 ? 11 : 6) : (ND == 2 ? 17 : 10);
    constexpr int NWP = BLOCKP / 64;
    std::string perr;
    const size_t si = sizeof(meso_patch::ItemT<D>), sf = sizeof(meso_patch::FaceT<D>);
    auto rec = [&](int ipw, int mi, int mf) { return ipw > 1 ? size_t(mi) * si + size_t(mf) * sf : size_t(0); };
    int lds_blk = 65536, lds_cu = 65536;
    (void)hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    (void)hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device);
    const size_t cap = std::min<size_t>(size_t(lds_blk), size_t(std::max(lds_cu, lds_blk)) / 2);
    const size_t cap_up = cap;
    bool ok = true;
    // owned cells per update patch: at most upd_qj per thread slot
    const int pmax = upd_qj(VTU) * NWP * IPWU;
    // 16 owned cells for FP64 in 3D (cube 32^3 x 16^3 update: 5.6 -> 5.2 ms),
    // else 32 (FP32 3D: 16 cells gave 4.5 against 4.0 ms)
    const int pdef = D == 3 && storage == MESO_STORAGE_FP64 ? 16 : 32;
    const int preq = c->preq_up > 0 ? c->preq_up : env_int("MESO_HIP_PATCH_CELLS", pdef);
    for (int P = std::min(preq, pmax); ok; P = P * 3 / 4) {
        ok = meso_patch::build_update_patches<D>(mv, pown, P, ups, perr);
        // the layout k_patch_update uses (owned cells only: it reads the face
        // values the moment kernel wrote)
        c->lds_up = storage == MESO_STORAGE_FP32
                        ? UpdateLds<float, D>(ups.max_owned, ups.max_faces, VTU, ND).bytes
                        : UpdateLds<double, D>(ups.max_owned, ups.max_faces, VTU, ND).bytes;
        if (c->lds_up <= cap_up || P <= 4) break;
    }
    // moment patches: cell-based (each owned cell's gradient once; the
    // default) or face-based (MESO_HIP_MOMENT_CELLS=0). Measured on 200^2 x
    // 101^2 moments: FP32 4.9 ms (16-cell patches) against 5.3 (32-face
    // patches), FP64 6.9 (32 cells) against 7.9 (64 faces); cube 32^3 x 16^3
    // FP32 3.5 ms against 4.6
    c->mp_cells = c->preq_cells >= 0 ? c->preq_cells != 0 : env_int("MESO_HIP_MOMENT_CELLS", 1) != 0;
    if (c->mp_cells) {
        const std::vector<char> all(nc, 1);
        // 16 owned cells (cube 32^3 x 16^3 moments: FP32 4.6 -> 3.8 ms, FP64
        // 5.7 -> 4.3 ms against 32), 32 for FP64 in 2D (6.9 against 7.5 ms)
        const int mdef = D == 2 && storage == MESO_STORAGE_FP64 ? 32 : 16;
        const int mreq = c->preq_mom > 0 ? c->preq_mom : env_int("MESO_HIP_MOMENT_PATCH_CELLS", mdef);
        for (int P = std::min(mreq, 64); ok; P = P * 3 / 4) {
            ok = meso_patch::build_update_patches<D>(mv, all, P, mps, perr);
            // [items | faces] | rows (the items) | tile velocities | values | f-bar+ rows
            const size_t a1 = rec(IPWM, mps.max_items, mps.max_faces) + ((4 * size_t(mps.max_items) + 31) & ~size_t(31)) +
                              mp_vtile_bytes(D, VTM) + size_t(ND) * mps.max_items * VTM * es +
                              size_t(ND) * mps.max_faces * (VTM + 1) * es;
            const size_t a2 = size_t(NWP) * 64 * ncol * 8;  // reduction buffer [threads][ncol]
            c->lds_mp = std::max(a1, a2);
            // the moment phase takes at most 128 faces per block
            if ((c->lds_mp <= cap && mps.max_faces <= 128) || P <= 4) break;
        }
        ok = ok && mps.max_faces <= 128;
    } else {
        const int fcap = c->preq_mom > 0 ? c->preq_mom : env_int("MESO_HIP_PATCH_FACES", IPWM > 1 ? 64 : 32);
        for (int F = std::min(fcap, 64); ok; F = F * 3 / 4) {
            ok = meso_patch::build_moment_patches<D>(mv, F, mps, perr);
            // [items | faces] | rows | tile velocities | values | f-bar+ rows (k_patch_moments)
            const size_t a1 = rec(IPWM, mps.max_items, mps.max_faces) + ((4 * size_t(mps.max_rows) + 31) & ~size_t(31)) +
                              mp_vtile_bytes(D, VTM) + size_t(ND) * mps.max_rows * VTM * es +
                              size_t(ND) * mps.max_faces * (VTM + 1) * es;
            const size_t a2 = size_t(NWP) * 64 * ncol * 8;  // reduction buffer [threads][ncol]
            c->lds_mp = std::max(a1, a2);
            if (c->lds_mp <= cap || F <= 4) break;
        }
    }
    if (!ok || c->lds_up > cap_up || c->lds_mp > cap || ups.n_patch == 0 || mps.n_patch == 0 ||
        ups.max_owned > pmax)
        return false;
    c->patches = true;
    // the kernels' dynamic-LDS attributes are set again for the new sizes
    c->upd_attr[0] = c->upd_attr[1] = false;
    c->mp_attr[0][0] = c->mp_attr[0][1] = c->mp_attr[1][0] = c->mp_attr[1][1] = false;
    if (keep_chunk) {
        set_vchunk(c, c->p_vchunk);  // ntile_f follows the moment patch kind
    } else {
        int ncu = 256;
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
        const int env_chunk = env_int("MESO_HIP_PCHUNK", 0);
        set_vchunk(c, env_chunk > 0 ? std::max(64, (env_chunk + 63) / 64 * 64)
                                    : default_vchunk(nv, std::min(ups.n_patch, mps.n_patch), ncu));
        // room for twice as many chunks (meso_hip_autotune tries them)
        c->p_chunk_cap = std::min((nv + 63) / 64, 2 * c->p_chunks);
    }
    c->n_up = ups.n_patch;
    c->up_maxi = ups.max_items;
    c->up_maxf = ups.max_faces;
    c->up_maxo = ups.max_owned;
    c->up_maxr = ups.max_rows;
    c->n_mp = mps.n_patch;
    c->mp_maxi = mps.max_items;
    c->mp_maxf = mps.max_faces;
    c->mp_maxr = mps.max_rows;
    c->mp_size = c->mp_cells ? mps.max_owned : mps.max_faces;
    return true;
}

template <class P>
int upload_rec(meso_hip_ctx *c, const void **dst, const std::vector<P> &v) {
    P *q = nullptr;
    const int rc = upload(c, &q, v.data(), v.size());
    *dst = q;
    return rc;
}

// Device copies of the patch sets.
template <int D>
int upload_patches(meso_hip_ctx *c, const meso_patch::PatchSet<D> &ups, const meso_patch::PatchSet<D> &mps) {
    int rc;
    // the previous patch sets (autotuner), once no launch uses them
    if (c->up_items) HIP_TRY(c, hipStreamSynchronize(c->stream));
    dfree(c, c->up_items);
    dfree(c, c->up_faces);
    dfree(c, c->up_item_ptr);
    dfree(c, c->up_face_ptr);
    dfree(c, c->up_nown);
    dfree(c, c->up_row_ptr);
    dfree(c, c->up_rows);
    dfree(c, c->mp_row_ptr);
    dfree(c, c->mp_rows);
    dfree(c, c->mp_items);
    dfree(c, c->mp_faces);
    dfree(c, c->mp_item_ptr);
    dfree(c, c->mp_face_ptr);
    dfree(c, c->mp_nown);
    ++c->kver;
    if ((rc = upload_rec(c, &c->up_items, ups.items)) || (rc = upload_rec(c, &c->up_faces, ups.faces)) ||
        (rc = upload(c, &c->up_item_ptr, ups.item_ptr.data(), ups.item_ptr.size())) ||
        (rc = upload(c, &c->up_face_ptr, ups.face_ptr.data(), ups.face_ptr.size())) ||
        (rc = upload(c, &c->up_nown, ups.n_owned.data(), ups.n_owned.size())) ||
        (rc = upload(c, &c->up_row_ptr, ups.row_ptr.data(), ups.row_ptr.size())) ||
        (rc = upload(c, &c->up_rows, ups.rows.data(), ups.rows.size())) ||
        (rc = upload(c, &c->mp_row_ptr, mps.row_ptr.data(), mps.row_ptr.size())) ||
        (rc = upload(c, &c->mp_rows, mps.rows.data(), mps.rows.size())) ||
        (rc = upload_rec(c, &c->mp_items, mps.items)) || (rc = upload_rec(c, &c->mp_faces, mps.faces)) ||
        (rc = upload(c, &c->mp_item_ptr, mps.item_ptr.data(), mps.item_ptr.size())) ||
        (rc = upload(c, &c->mp_face_ptr, mps.face_ptr.data(), mps.face_ptr.size())) ||
        (rc = upload(c, &c->mp_nown, mps.n_owned.data(), mps.n_owned.size())))
        return rc;
    return MESO_OK;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

const char *meso_hip_last_error(const meso_hip_ctx *ctx) { return ctx ? ctx->err.c_str() : g_hip_err.c_str(); }

int meso_hip_create(const meso_mesh_desc *m, const meso_dvs_desc *d, const meso_step_params *p, int32_t storage,
                    meso_hip_ctx **out) {
    if (!m || !d || !p || !out) return set_err(nullptr, MESO_E_INVALID, "null argument");
    if (p->model != MESO_MODEL_CDUGKS_BGK && p->model != MESO_MODEL_CDUGKS_SHAKHOV &&
        p->model != MESO_MODEL_DUGKS_BGK)
        return set_err(nullptr, MESO_E_INVALID, "unknown model");
    if (m->dim != 2 && m->dim != 3) return set_err(nullptr, MESO_E_INVALID, "dimension must be 2 or 3");
    if (p->dim != m->dim) return set_err(nullptr, MESO_E_INVALID, "params.dim differs from the mesh dimension");
    if (storage != MESO_STORAGE_FP64 && storage != MESO_STORAGE_FP32 && storage != MESO_STORAGE_FP32X &&
        storage != MESO_STORAGE_FP32E)
        return set_err(nullptr, MESO_E_INVALID,
                       "storage must be MESO_STORAGE_FP64, MESO_STORAGE_FP32, MESO_STORAGE_FP32X or MESO_STORAGE_FP32E");
    // FP32X runs the FP64 kernels, FP32E the FP32 kernels, on words +
    // extension bytes in HBM
    const bool ext = storage == MESO_STORAGE_FP32X || storage == MESO_STORAGE_FP32E;
    if (storage == MESO_STORAGE_FP32X) storage = MESO_STORAGE_FP64;
    if (storage == MESO_STORAGE_FP32E) storage = MESO_STORAGE_FP32;
    if (m->n_cells <= 0 || m->n_faces <= 0 || d->n_vel <= 0) return set_err(nullptr, MESO_E_INVALID, "empty mesh or DVS");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(nullptr, MESO_E_DEVICE, "no HIP device available");
    auto *c = new meso_hip_ctx();
    auto fail = [&](int rc) {
        meso_hip_destroy(c);
        return rc;
    };
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(nullptr, MESO_E_DEVICE, "hipStreamCreate failed"));
    c->storage = storage;
    c->ext = ext;
    c->nd = p->model == MESO_MODEL_CDUGKS_SHAKHOV ? 2 : 1;
    set_const(c->s, *p);
    const int nc = m->n_cells, nf = m->n_faces, nv = d->n_vel, D = m->dim;
    c->s.nc = nc;
    c->s.nf = nf;
    c->s.nv = nv;
    c->s.nvp = nv >= 64 ? ((nv + 63) / 64) * 64 : nv;
    c->lp = lanes_for(nv);
    const int target = env_int("MESO_HIP_VTILE", 1024);
    tiles_for(nv, 64, target, c->s.ntile_f, c->s.vt_f);
    tiles_for(nv, 64, target, c->s.ntile_c, c->s.vt_c);
    // ---- validate and repack mesh arrays
    for (int f = 0; f < nf; ++f) {
        int o = m->face_cells[2 * f], n = m->face_cells[2 * f + 1];
        if (o < 0 || o >= nc || n < 0 || n >= nc) return fail(set_err(nullptr, MESO_E_INVALID, "face cell index out of range"));
        if (m->face_mark[f] < 0 || m->face_mark[f] >= m->n_marks)
            return fail(set_err(nullptr, MESO_E_INVALID, "face mark out of range"));
    }
    if (m->cell_face_ptr[0] != 0 || m->cell_face_ptr[nc] != m->n_cell_faces || m->nbr_ptr[0] != 0 ||
        m->nbr_ptr[nc] != m->n_nbr)
        return fail(set_err(nullptr, MESO_E_INVALID, "CSR pointers inconsistent"));
    for (int k = 0; k < m->n_cell_faces; ++k)
        if (m->cell_face_ids[k] < 0 || m->cell_face_ids[k] >= nf)
            return fail(set_err(nullptr, MESO_E_INVALID, "cell face index out of range"));
    for (int k = 0; k < m->n_nbr; ++k)
        if (m->nbr_ids[k] < 0 || m->nbr_ids[k] >= nc)
            return fail(set_err(nullptr, MESO_E_INVALID, "neighbour index out of range"));
    std::vector<int2> fc(nf);
    std::vector<double> fd(6 * size_t(nf)), sA(m->n_cell_faces), na(3 * size_t(m->n_nbr));
    std::vector<float> fdf(6 * size_t(nf)), naf(3 * size_t(m->n_nbr)), xif(3 * size_t(nv));
    // FP64: with non-planar (warped hexahedron) faces |S| / V is O(jitter / h), and
    // the f_ref flux it carries is as large as the deviation fluxes
    std::vector<double> cS(3 * size_t(nc), 0.0);
    for (int f = 0; f < nf; ++f) {
        fc[f] = make_int2(m->face_cells[2 * f], m->face_cells[2 * f + 1]);
        for (int side = 0; side < 2; ++side) {
            int cc = m->face_cells[2 * f + side];
            for (int k = 0; k < 3; ++k) {
                double dv = m->face_pos[3 * f + k] - m->cell_pos[3 * cc + k];
                fd[6 * f + 3 * side + k] = dv;
                fdf[6 * f + 3 * side + k] = float(dv);
            }
        }
    }
    for (int cc = 0; cc < nc; ++cc) {
        for (int k = m->cell_face_ptr[cc]; k < m->cell_face_ptr[cc + 1]; ++k) {
            int f = m->cell_face_ids[k];
            sA[k] = (m->face_cells[2 * f] == cc) ? m->face_area[f] : -m->face_area[f];
        }
        double S3[3] = {0.0, 0.0, 0.0};
        for (int k = m->cell_face_ptr[cc]; k < m->cell_face_ptr[cc + 1]; ++k)
            for (int r = 0; r < 3; ++r) S3[r] += sA[k] * m->face_normal[3 * size_t(m->cell_face_ids[k]) + r];
        for (int r = 0; r < 3; ++r) cS[3 * size_t(cc) + r] = S3[r];
        const double *C = m->lsq_C + 9 * size_t(cc);
        for (int j = m->nbr_ptr[cc]; j < m->nbr_ptr[cc + 1]; ++j) {
            const double *dr = m->lsq_dr + 3 * size_t(j);
            const double wj = m->lsq_w[j];
            for (int r = 0; r < 3; ++r) {
                double v = wj * (C[3 * r] * dr[0] + C[3 * r + 1] * dr[1] + C[3 * r + 2] * dr[2]);
                na[3 * j + r] = v;
                naf[3 * j + r] = float(v);
            }
        }
    }
    for (int k = 0; k < 3 * nv; ++k) xif[k] = float(d->xi[k]);
    // boundary faces: every face whose mark is not fluid-interior, plus
    // boundary faces without a mark (own == nei) are handled in place
    c->marks_h.assign(m->marks, m->marks + m->n_marks);
    c->n_marks = m->n_marks;
    c->face_bnd_h.assign(nf, -1);
    std::vector<int> bmark;
    for (int f = 0; f < nf; ++f) {
        const meso_mark_desc &mk = m->marks[m->face_mark[f]];
        if (mk.type == MESO_BC_FLUID_INTERIOR) continue;
        if (m->face_cells[2 * f] != m->face_cells[2 * f + 1])
            return fail(set_err(nullptr, MESO_E_INVALID, "boundary mark on an interior face"));
        if ((mk.type == MESO_BC_SYMMETRY || mk.type == MESO_BC_SLIP_WALL) && p->model == MESO_MODEL_CDUGKS_SHAKHOV) {
            if (!d->mirror) return fail(set_err(nullptr, MESO_E_INVALID, "symmetry/slip-wall need DVS mirror ids"));
            if (mk.direction < 0 || mk.direction > 2)
                return fail(set_err(nullptr, MESO_E_INVALID, "mirror direction must be 0, 1 or 2"));
        }
        c->face_bnd_h[f] = int(c->bnd_face_h.size());
        c->bnd_face_h.push_back(f);
        bmark.push_back(m->face_mark[f]);
    }
    c->s.nb = int(c->bnd_face_h.size());
    if (d->mirror)
        for (int k = 0; k < 3 * nv; ++k)
            if (d->mirror[k] < 0 || d->mirror[k] >= nv)
                return fail(set_err(nullptr, MESO_E_INVALID, "mirror index out of range"));
    // per-cell stencil tables (faces, cells across them, LSQ neighbours), padded
    const int MAXF = D == 2 ? 4 : 6;
    std::vector<int> t_nf(nc), t_face(size_t(nc) * MAXF, -1), t_other(size_t(nc) * MAXF, -1),
        t_own(size_t(nc) * MAXF, 0), t_bnd(size_t(nc) * MAXF, -1), t_lcnt(nc), t_lnb(size_t(nc) * MAXF, 0);
    std::vector<double> t_sA(size_t(nc) * MAXF, 0.0), t_d(size_t(nc) * MAXF * 6, 0.0),
        t_la(size_t(nc) * MAXF * 3, 0.0);
    std::vector<float> t_df(size_t(nc) * MAXF * 6, 0.f), t_laf(size_t(nc) * MAXF * 3, 0.f), t_nf32(3 * size_t(nf));
    for (int cc = 0; cc < nc; ++cc) {
        const int k0 = m->cell_face_ptr[cc], nk = m->cell_face_ptr[cc + 1] - k0;
        const int j0 = m->nbr_ptr[cc], nj = m->nbr_ptr[cc + 1] - j0;
        if (nk > MAXF || nj > MAXF)
            return fail(set_err(nullptr, MESO_E_INVALID, "cell with more faces/neighbours than its kind allows"));
        t_nf[cc] = nk;
        t_lcnt[cc] = nj;
        for (int k = 0; k < nk; ++k) {
            const int f = m->cell_face_ids[k0 + k];
            const int o = m->face_cells[2 * f], n = m->face_cells[2 * f + 1];
            const size_t idx = size_t(cc) * MAXF + k;
            t_face[idx] = f;
            t_own[idx] = (o == cc) ? 1 : 0;
            t_other[idx] = (o == n) ? -1 : (o == cc ? n : o);
            t_bnd[idx] = c->face_bnd_h[f];
            t_sA[idx] = sA[k0 + k];
            const int oth = t_other[idx] >= 0 ? t_other[idx] : cc;
            for (int r = 0; r < 3; ++r) {
                const double ds = m->face_pos[3 * f + r] - m->cell_pos[3 * size_t(cc) + r];
                const double dn = m->face_pos[3 * f + r] - m->cell_pos[3 * size_t(oth) + r];
                t_d[6 * idx + r] = ds;
                t_d[6 * idx + 3 + r] = dn;
                t_df[6 * idx + r] = float(ds);
                t_df[6 * idx + 3 + r] = float(dn);
            }
        }
        for (int j = 0; j < nj; ++j) {
            const size_t idx = size_t(cc) * MAXF + j;
            t_lnb[idx] = m->nbr_ids[j0 + j];
            for (int r = 0; r < 3; ++r) {
                t_la[3 * idx + r] = na[3 * size_t(j0 + j) + r];
                t_laf[3 * idx + r] = float(na[3 * size_t(j0 + j) + r]);
            }
        }
    }
    for (size_t k = 0; k < 3 * size_t(nf); ++k) t_nf32[k] = float(m->face_normal[k]);
    // gather stencils: group 0 = the cell, group 1+k = the cell across face k
    // (the cell itself again on a boundary face, with zero weights)
    const int GS = 1 + MAXF, NG = 1 + MAXF;
    std::vector<int> t_sid(size_t(nc) * NG * GS), t_fsid(size_t(nf) * 2 * GS);
    std::vector<double> t_sa(size_t(nc) * NG * MAXF * 3, 0.0), t_fsa(size_t(nf) * 2 * MAXF * 3, 0.0);
    std::vector<float> t_saf(t_sa.size(), 0.f), t_fsaf(t_fsa.size(), 0.f);
    auto fill_group = [&](int cell, int *ids, double *ad) {
        ids[0] = cell;
        const int j0 = m->nbr_ptr[cell], nj = m->nbr_ptr[cell + 1] - j0;
        for (int j = 0; j < MAXF; ++j) {
            ids[1 + j] = j < nj ? m->nbr_ids[j0 + j] : cell;
            for (int r = 0; r < 3; ++r) ad[3 * j + r] = j < nj ? na[3 * size_t(j0 + j) + r] : 0.0;
        }
    };
    std::vector<int> int_cells, bnd_cells;
    for (int cc = 0; cc < nc; ++cc) {
        int *ids = &t_sid[size_t(cc) * NG * GS];
        double *ad = &t_sa[size_t(cc) * NG * MAXF * 3];
        fill_group(cc, ids, ad);
        bool has_bnd = false;
        for (int k = 0; k < MAXF; ++k) {
            const size_t idx = size_t(cc) * MAXF + k;
            const int o = (k < t_nf[cc]) ? t_other[idx] : -1;
            if (k < t_nf[cc] && o < 0) has_bnd = true;
            fill_group(o >= 0 ? o : cc, ids + (1 + k) * GS, ad + (1 + k) * MAXF * 3);
            if (o < 0)
                for (int j = 0; j < MAXF * 3; ++j) ad[(1 + k) * MAXF * 3 + j] = 0.0;
        }
        (has_bnd ? bnd_cells : int_cells).push_back(cc);
    }
    for (int f = 0; f < nf; ++f) {
        fill_group(m->face_cells[2 * f], &t_fsid[size_t(f) * 2 * GS], &t_fsa[size_t(f) * 2 * MAXF * 3]);
        fill_group(m->face_cells[2 * f + 1], &t_fsid[size_t(f) * 2 * GS + GS],
                   &t_fsa[size_t(f) * 2 * MAXF * 3 + MAXF * 3]);
    }
    for (size_t k = 0; k < t_sa.size(); ++k) t_saf[k] = float(t_sa[k]);
    for (size_t k = 0; k < t_fsa.size(); ++k) t_fsaf[k] = float(t_fsa[k]);
    c->n_int = int(int_cells.size());
    c->n_bcell = int(bnd_cells.size());
    std::vector<int> t_list(int_cells);
    t_list.insert(t_list.end(), bnd_cells.begin(), bnd_cells.end());
    // velocity slots and, for a tensor-product set, the axis tables
    std::vector<double4> t_vd(nv);
    std::vector<float4> t_vf(nv);
    for (int v = 0; v < nv; ++v) {
        const double *x = d->xi + 3 * size_t(v);
        t_vd[v] = make_double4(x[0], x[1], x[2], d->w[v]);
        t_vf[v] = make_float4(float(x[0]), float(x[1]), float(x[2]), float(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]));
    }
    std::vector<int> t_vtab, t_tax;
    std::vector<double> t_tx, t_tw;
    if (p->model == MESO_MODEL_CDUGKS_SHAKHOV && env_int("MESO_HIP_NO_TABLES", 0) == 0) {
        // unique abscissae per axis; tensor when every (i,j,k) combination occurs once
        std::vector<std::vector<double>> axes(D);
        for (int dd = 0; dd < D; ++dd) {
            for (int v = 0; v < nv; ++v) axes[dd].push_back(d->xi[3 * size_t(v) + dd]);
            std::sort(axes[dd].begin(), axes[dd].end());
            axes[dd].erase(std::unique(axes[dd].begin(), axes[dd].end()), axes[dd].end());
        }
        size_t prod = 1, tot = 0;
        for (int dd = 0; dd < D; ++dd) {
            prod *= axes[dd].size();
            tot += axes[dd].size();
        }
        bool third_zero = true;
        if (D == 2)
            for (int v = 0; v < nv; ++v) third_zero = third_zero && d->xi[3 * size_t(v) + 2] == 0.0;
        if (prod == size_t(nv) && tot < 1024 && third_zero) {
            std::vector<int> seen(prod, 0);
            bool ok = true;
            t_vtab.resize(nv);
            int off[3] = {0, 0, 0};
            for (int dd = 1; dd < D; ++dd) off[dd] = off[dd - 1] + int(axes[dd - 1].size());
            for (int v = 0; v < nv && ok; ++v) {
                int idx[3] = {0, 0, 0};
                size_t lin = 0;
                for (int dd = D - 1; dd >= 0; --dd) {
                    const double xv = d->xi[3 * size_t(v) + dd];
                    idx[dd] = int(std::lower_bound(axes[dd].begin(), axes[dd].end(), xv) - axes[dd].begin());
                    lin = lin * axes[dd].size() + idx[dd];
                }
                if (seen[lin]++) ok = false;
                int packed = 0;
                for (int dd = 0; dd < D; ++dd) packed |= (off[dd] + idx[dd]) << (10 * dd);
                t_vtab[v] = packed;
            }
            if (ok) {
                for (int dd = 0; dd < D; ++dd)
                    for (double xv : axes[dd]) {
                        t_tx.push_back(xv);
                        t_tax.push_back(dd);
                    }
                c->nt = int(t_tx.size());
                // 1D weights when w(v) is the product of per-axis weights (to
                // 1e-13): taken along the axis lines through the heaviest velocity
                int vm = 0;
                for (int v = 1; v < nv; ++v)
                    if (d->w[v] > d->w[vm]) vm = v;
                std::vector<int> im(D);
                for (int dd = 0; dd < D; ++dd) im[dd] = (t_vtab[vm] >> (10 * dd)) & 1023;
                std::vector<double> tw(c->nt, 0.0);
                bool fact = d->w[vm] > 0.0;
                for (int v = 0; v < nv && fact; ++v) {
                    int same = 0, axis = -1;
                    for (int dd = 0; dd < D; ++dd) {
                        if (((t_vtab[v] >> (10 * dd)) & 1023) == im[dd]) ++same;
                        else axis = dd;
                    }
                    if (same == D) {
                        for (int dd = 0; dd < D; ++dd) tw[im[dd]] = dd == 0 ? d->w[vm] : 1.0;
                    } else if (same == D - 1) {
                        const int tt = (t_vtab[v] >> (10 * axis)) & 1023;
                        tw[tt] = axis == 0 ? d->w[v] : d->w[v] / d->w[vm];
                    }
                }
                for (int v = 0; v < nv && fact; ++v) {
                    double prod = 1.0;
                    for (int dd = 0; dd < D; ++dd) prod *= tw[(t_vtab[v] >> (10 * dd)) & 1023];
                    fact = std::fabs(prod - d->w[v]) <= 1e-13 * std::fabs(d->w[v]) + 1e-300;
                }
                if (fact && env_int("MESO_HIP_SEP_FLUX", 1) != 0) t_tw = tw;
            }
        }
    }
    // ---- velocity order. A tile of consecutive velocity slots should be compact
    // in velocity space: its axis-table entries are then few and shared with the
    // next tile, and most faces see one upwind side across it. A tensor-product
    // set is therefore stored in Morton order of its axis indices; the state
    // transfers map slots back to meso's order.
    std::vector<int> perm(nv), slot_of(nv);
    for (int v = 0; v < nv; ++v) perm[v] = v;
    if (c->nt > 0 && env_int("MESO_HIP_VORDER", 1) != 0) {
        int axis0[3] = {0, 0, 0};
        for (int t = 1; t < c->nt; ++t)
            if (t_tax[t] != t_tax[t - 1]) axis0[t_tax[t]] = t;
        auto spread = [D](uint64_t x) {  // bit b of x to bit D*b
            uint64_t r = 0;
            for (int b = 0; b < 21; ++b) r |= ((x >> b) & 1ull) << (D * b);
            return r;
        };
        std::vector<uint64_t> key(nv);
        for (int v = 0; v < nv; ++v) {
            uint64_t k = 0;
            for (int dd = 0; dd < D; ++dd) k |= spread(uint64_t(((t_vtab[v] >> (10 * dd)) & 1023) - axis0[dd])) << dd;
            key[v] = k;
        }
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return key[x] < key[y]; });
    }
    bool identity = true;
    for (int sl = 0; sl < nv; ++sl) {
        slot_of[perm[sl]] = sl;
        identity = identity && perm[sl] == sl;
    }
    std::vector<double> xi_s(3 * size_t(nv)), w_s(nv);
    std::vector<int> mirror_s;
    {
        std::vector<double4> vd(nv);
        std::vector<float4> vf(nv);
        std::vector<float> xf(3 * size_t(nv));
        std::vector<int> vt(t_vtab.empty() ? 0 : nv);
        for (int sl = 0; sl < nv; ++sl) {
            const int v = perm[sl];
            vd[sl] = t_vd[v];
            vf[sl] = t_vf[v];
            for (int k = 0; k < 3; ++k) {
                xi_s[3 * size_t(sl) + k] = d->xi[3 * size_t(v) + k];
                xf[3 * size_t(sl) + k] = xif[3 * size_t(v) + k];
            }
            w_s[sl] = d->w[v];
            if (!vt.empty()) vt[sl] = t_vtab[v];
        }
        t_vd.swap(vd);
        t_vf.swap(vf);
        xif.swap(xf);
        t_vtab.swap(vt);
        if (d->mirror) {
            mirror_s.resize(3 * size_t(nv));
            for (int sl = 0; sl < nv; ++sl)
                for (int k = 0; k < 3; ++k) mirror_s[3 * size_t(sl) + k] = slot_of[d->mirror[3 * size_t(perm[sl]) + k]];
        }
    }
    // ---- LDS-staged patch kernels: CDUGKS models with large velocity sets
    meso_patch::PatchSet<2> ups2, mps2;
    meso_patch::PatchSet<3> ups3, mps3;
    if (p->model != MESO_MODEL_DUGKS_BGK && nv >= 64 && env_int("MESO_HIP_PATCH", 1) != 0) {
        // every cell is updated by a patch except those with a mirror boundary
        // (symmetry, slip-wall), whose face values need other velocities
        std::vector<char> pown(nc, 1);
        for (int f = 0; f < nf; ++f) {
            const int b = c->face_bnd_h[f];
            if (b < 0) continue;
            const int kind = m->marks[m->face_mark[f]].type;
            if (kind == MESO_BC_SYMMETRY || kind == MESO_BC_SLIP_WALL) pown[m->face_cells[2 * f]] = 0;
        }
        PatchSrc &ps = c->psrc;
        ps.nc = nc;
        ps.nf = nf;
        ps.face_cells.assign(m->face_cells, m->face_cells + 2 * size_t(nf));
        ps.cf_ptr.assign(m->cell_face_ptr, m->cell_face_ptr + nc + 1);
        ps.cf_ids.assign(m->cell_face_ids, m->cell_face_ids + m->n_cell_faces);
        ps.nb_ptr.assign(m->nbr_ptr, m->nbr_ptr + nc + 1);
        ps.nb_ids.assign(m->nbr_ids, m->nbr_ids + m->n_nbr);
        ps.cell_pos.assign(m->cell_pos, m->cell_pos + 3 * size_t(nc));
        ps.face_pos.assign(m->face_pos, m->face_pos + 3 * size_t(nf));
        ps.face_area.assign(m->face_area, m->face_area + nf);
        ps.face_normal.assign(m->face_normal, m->face_normal + 3 * size_t(nf));
        ps.nb_a = na;
        ps.pown = pown;
        const bool ok = D == 2 ? setup_patches<2>(c, storage, false, ups2, mps2)
                               : setup_patches<3>(c, storage, false, ups3, mps3);
        if (ok) {
            // cell lists: patch cells first (k_cell_prep), then the mirror-boundary cells (k_cell_update)
            t_list.clear();
            for (int cc = 0; cc < nc; ++cc)
                if (pown[cc]) t_list.push_back(cc);
            c->n_int = int(t_list.size());
            for (int cc = 0; cc < nc; ++cc)
                if (!pown[cc]) t_list.push_back(cc);
            c->n_bcell = nc - c->n_int;
        }
    }
    {
        // the cell kernel keeps MAXF + 2 tables per wave in LDS: stay within 64 KiB
        const size_t lds = size_t(BLOCK / 64) * (MAXF + 2) * size_t(c->nt) * (storage == MESO_STORAGE_FP32 ? 4 : 8);
        if (lds > 65536) c->nt = 0;
    }
    int rc = 0;
#define UP(dst, src, n) \
    if ((rc = upload(c, &c->dst, src, n))) return fail(rc);
    UP(face_cells, fc.data(), nf);
    UP(face_n, m->face_normal, 3 * size_t(nf));
    UP(face_d, fd.data(), fd.size());
    UP(face_df, fdf.data(), fdf.size());
    UP(cf_ptr, m->cell_face_ptr, nc + 1);
    UP(cf_ids, m->cell_face_ids, m->n_cell_faces);
    UP(cf_sA, sA.data(), sA.size());
    UP(nb_ptr, m->nbr_ptr, nc + 1);
    UP(nb_ids, m->nbr_ids, m->n_nbr);
    UP(nb_w, m->lsq_w, m->n_nbr);
    UP(nb_a, na.data(), na.size());
    UP(nb_af, naf.data(), naf.size());
    UP(cell_vol, m->cell_vol, nc);
    UP(cell_S, cS.data(), cS.size());
    UP(cs_nf, t_nf.data(), t_nf.size());
    UP(cs_face, t_face.data(), t_face.size());
    UP(cs_other, t_other.data(), t_other.size());
    UP(cs_own, t_own.data(), t_own.size());
    UP(cs_bnd, t_bnd.data(), t_bnd.size());
    UP(cs_sA, t_sA.data(), t_sA.size());
    UP(cs_d, t_d.data(), t_d.size());
    UP(cs_df, t_df.data(), t_df.size());
    UP(lsq_cnt, t_lcnt.data(), t_lcnt.size());
    UP(lsq_nb, t_lnb.data(), t_lnb.size());
    UP(lsq_a, t_la.data(), t_la.size());
    UP(lsq_af, t_laf.data(), t_laf.size());
    UP(face_nf, t_nf32.data(), t_nf32.size());
    UP(cell_list, t_list.data(), t_list.size());
    UP(st_id, t_sid.data(), t_sid.size());
    UP(st_a, t_sa.data(), t_sa.size());
    UP(st_af, t_saf.data(), t_saf.size());
    UP(fs_id, t_fsid.data(), t_fsid.size());
    UP(fs_a, t_fsa.data(), t_fsa.size());
    UP(fs_af, t_fsaf.data(), t_fsaf.size());
    if (c->patches) {
        if ((rc = D == 2 ? upload_patches<2>(c, ups2, mps2) : upload_patches<3>(c, ups3, mps3))) return fail(rc);
    }
    UP(vel_d, t_vd.data(), t_vd.size());
    UP(vel_f, t_vf.data(), t_vf.size());
    if (c->nt > 0) {
        UP(vtab, t_vtab.data(), t_vtab.size());
        UP(tab_x, t_tx.data(), t_tx.size());
        UP(tab_ax, t_tax.data(), t_tax.size());
        if (!t_tw.empty()) UP(tab_w, t_tw.data(), t_tw.size());
    }
    UP(face_bnd, c->face_bnd_h.data(), nf);
    UP(bnd_face, c->bnd_face_h.data(), c->bnd_face_h.size());
    UP(bnd_mark, bmark.data(), bmark.size());
    UP(marks, m->marks, m->n_marks);
    UP(xi, xi_s.data(), xi_s.size());
    UP(xif, xif.data(), xif.size());
    UP(w, w_s.data(), w_s.size());
    if (d->mirror) UP(mirror, mirror_s.data(), mirror_s.size());
    if (!identity) UP(slot_of, slot_of.data(), slot_of.size());
#undef UP
    const size_t es = storage == MESO_STORAGE_FP32 ? 4 : 8;
    const size_t fes = ext ? 4 : es;  // bytes per stored distribution value (FP32X: the word; bytes apart)
    const size_t fsz = size_t(nc) * c->s.nvp;
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < c->nd; ++k) {
            char *q = nullptr;
            if ((rc = dalloc(c, &q, fsz * fes))) return fail(rc);
            c->f[b][k] = q;
            if (ext && (rc = dalloc(c, &c->fx[b][k], fsz))) return fail(rc);
        }
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(c, &c->rho[b], nc)) || (rc = dalloc(c, &c->u[b], 3 * size_t(nc))) ||
            (rc = dalloc(c, &c->T[b], nc)) || (rc = dalloc(c, &c->q[b], 3 * size_t(nc))))
            return fail(rc);
    }
    const int ncol = c->nd == 2 ? (D == 2 ? 11 : 17) : (D == 2 ? 6 : 10);
    if ((rc = dalloc(c, &c->rho_res, nc)) || (rc = dalloc(c, &c->u_res, 3 * size_t(nc))) ||
        (rc = dalloc(c, &c->T_res, nc)) || (rc = dalloc(c, &c->q_res, 3 * size_t(nc))) ||
        (rc = dalloc(c, &c->part_f, size_t(nf) * std::max(c->s.ntile_f, 2 * c->p_chunk_cap) * ncol)) ||
        (rc = dalloc(c, &c->face_flux, size_t(nf) * (D + 2))) ||
        (rc = dalloc(c, &c->part_c, size_t(nc) * std::max(c->s.ntile_c, c->p_chunk_cap) * 4)) || (rc = dalloc(c, &c->fm, 5 * size_t(nc))) ||
        (rc = dalloc(c, &c->res_part, 8 * size_t((nc + BLOCK - 1) / BLOCK))) || (rc = dalloc(c, &c->res_out, 8)) ||
        (rc = dalloc(c, &c->face_eq, nf)) || (rc = dalloc(c, &c->face_eqf, nf)) || (rc = dalloc(c, &c->bnd, std::max(1, c->s.nb))) ||
        (rc = dalloc(c, &c->flags, FLAG_INTS)) || (rc = dalloc(c, &c->mask, nc)) || (rc = dalloc(c, &c->fref_d, 2 * size_t(nv))) ||
        (rc = dalloc(c, &c->fref_f, 2 * size_t(nv))) || (rc = dalloc(c, &c->fref_d2, nv)) ||
        (rc = dalloc(c, &c->fref_f2, nv)))
        return fail(rc);
    if (c->nt > 0) {
        char *ft = nullptr;
        if ((rc = dalloc(c, &ft, size_t(nf) * c->nt * es))) return fail(rc);
        c->face_tab = ft;
    }
    if (c->patches) {
        char *cr = nullptr;
        const size_t rs = storage == MESO_STORAGE_FP32
                              ? (D == 2 ? sizeof(CellRec<float, 2>) : sizeof(CellRec<float, 3>))
                              : (D == 2 ? sizeof(CellRec<double, 2>) : sizeof(CellRec<double, 3>));
        if ((rc = dalloc(c, &cr, size_t(nc) * rs))) return fail(rc);
        c->crec = cr;
        if (c->nt > 0) {
            char *ct = nullptr;
            if ((rc = dalloc(c, &ct, size_t(nc) * 2 * c->nt * es))) return fail(rc);
            c->cell_tab = ct;
        }
        for (int k = 0; k < c->nd; ++k) {
            char *fb = nullptr;
            if ((rc = dalloc(c, &fb, size_t(nf) * c->s.nvp * es))) return fail(rc);
            c->fbar[k] = fb;
        }
    }
    // small velocity sets: the two-launch step (k_small_faces, k_small_cells)
    // for the CDUGKS models with FP64 storage; a Shakhov mirror boundary
    // (symmetry, slip wall) needs the mirrored velocity's face value and keeps
    // the per-stage kernels
    if (p->model != MESO_MODEL_DUGKS_BGK && nv <= 64 && !ext && storage == MESO_STORAGE_FP64 &&
        env_int("MESO_HIP_SMALL", 1) != 0) {
        bool mirror = false;
        for (int f = 0; f < nf && p->model == MESO_MODEL_CDUGKS_SHAKHOV; ++f) {
            if (c->face_bnd_h[f] < 0) continue;
            const int kind = m->marks[m->face_mark[f]].type;
            mirror = mirror || kind == MESO_BC_SYMMETRY || kind == MESO_BC_SLIP_WALL;
        }
        if (!mirror) {
            c->small_lp = nv <= 16 ? 16 : (nv <= 32 ? 32 : 64);
            if ((rc = dalloc(c, &c->sflux, size_t(c->nd) * nf * c->small_lp))) return fail(rc);
        }
    }
    c->xi_h = xi_s;
    c->w_h = w_s;
    if (ext && (!c->patches || c->n_bcell > 0 || p->model == MESO_MODEL_DUGKS_BGK))
        return fail(set_err(nullptr, MESO_E_INVALID,
                            "MESO_STORAGE_FP32X and FP32E need the patch kernels: a CDUGKS model, at least 64 velocities "
                            "and no symmetry or slip-wall boundary"));
    if (hipHostMalloc((void **)&c->flags_host, FLAG_INTS * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(set_err(nullptr, MESO_E_DEVICE, "hipHostMalloc failed"));
    if ((rc = reset_flags(c))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_err(nullptr, MESO_E_DEVICE, "upload failed"));
    *out = c;
    return MESO_OK;
}

void meso_hip_destroy(meso_hip_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->flags_host) (void)hipHostFree(c->flags_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int meso_hip_set_mask(meso_hip_ctx *c, const int32_t *mask) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    const int nc = c->s.nc;
    std::vector<int> m(nc, 0);
    bool any = false;
    if (mask)
        for (int i = 0; i < nc; ++i) {
            m[i] = mask[i] ? 1 : 0;
            any = any || m[i];
        }
    HIP_TRY(c, hipMemcpyAsync(c->mask, m.data(), size_t(nc) * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->has_mask = any && c->s.model == MESO_MODEL_CDUGKS_SHAKHOV;
    ++c->kver;  // the launch sequence and the kernel arguments change
    return MESO_OK;
}

int meso_hip_set_params(meso_hip_ctx *c, const meso_step_params *p) {
    if (!c || !p) return set_err(c, MESO_E_INVALID, "null argument");
    if (p->model != c->s.model || p->dim != c->s.D) return set_err(c, MESO_E_INVALID, "model/dimension cannot change");
    const StepConst before = c->s;
    set_const(c->s, *p);
    // meso re-reads its settings every step (cdugks_shakhov.cpp:762): unchanged
    // values keep the captured steps and batch plans
    if (std::memcmp(&before, &c->s, sizeof(StepConst)) != 0) ++c->kver;
    return MESO_OK;
}

// scr (non-null: the allocation-free entry point) holds 5 doubles per cell for
// the staged macros of init_equilibrium, which otherwise allocates them.
static int init_common(meso_hip_ctx *c, const double *rho, const double *u3, const double *T, const double *q3,
                       bool from_macros, double *scr = nullptr) {
    const bool shak = c->s.model == MESO_MODEL_CDUGKS_SHAKHOV;
    if (!rho || !u3 || (shak && !T) || (from_macros && shak && !q3))
        return set_err(c, MESO_E_INVALID, "null argument");
    const size_t nc = c->s.nc;
    double *tr = nullptr, *tu = nullptr, *tT = nullptr;
    int rc;
    if ((rc = set_reference_from(c, rho, shak ? T : nullptr))) return rc;
    // the init kernels write buffer `1 - cur`, then it becomes current
    KArgs a = make_args(c);
    if (from_macros) {
        const int o = 1 - c->cur;
        HIP_TRY(c, hipMemcpyAsync(c->rho[o], rho, nc * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->u[o], u3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->T[o], 0, nc * 8, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->q[o], 0, 3 * nc * 8, c->stream));
        if (shak) {
            HIP_TRY(c, hipMemcpyAsync(c->T[o], T, nc * 8, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(c, hipMemcpyAsync(c->q[o], q3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
        }
        rc = dispatch<InitF>(c, a, (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, true);
    } else {
        if (scr) {
            tr = scr;
            tu = scr + nc;
            tT = scr + 4 * nc;
        } else if ((rc = dalloc(c, &tr, nc)) || (rc = dalloc(c, &tu, 3 * nc)) || (rc = dalloc(c, &tT, nc))) {
            return rc;
        }
        HIP_TRY(c, hipMemcpyAsync(tr, rho, nc * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(tu, u3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
        if (T) HIP_TRY(c, hipMemcpyAsync(tT, T, nc * 8, hipMemcpyHostToDevice, c->stream));
        rc = dispatch<InitF>(c, a, (const double *)tr, (const double *)tu, (const double *)tT, false);
    }
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // temporaries are released right away (the scratch of the allocation-free entry point stays)
    for (double *t : {tr, tu, tT}) {
        if (!t || scr) continue;
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)t);
        if (it != c->allocs.end()) c->allocs.erase(it);
        (void)hipFree(t);
        c->bytes -= int64_t(std::max<size_t>(t == tu ? 3 * nc * 8 : nc * 8, 16));
    }
    c->cur = 1 - c->cur;
    c->steps_done = 0;
    if ((rc = reset_flags(c))) return rc;
    return copy_residual_base(c);
}

int meso_hip_init_equilibrium(meso_hip_ctx *c, const double *rho, const double *u3, const double *T) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    return init_common(c, rho, u3, T, nullptr, false);
}

}  // extern "C"

// ------------------------------------------- allocation-free initial state
// Scratch of meso_hip_init_equilibrium_noalloc: the per-cell macros (rho,
// u[3], T; 40 B per cell) that the initial-state kernel reads. The module holds
// one copy per device for the life of the process. Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device,
// kept per device in g_scratch_ptr; the array's name is not a device address
// on the host.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;  // 16 MiB: up to 419,430 cells
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu[NOALLOC_MAX_DEVICES];

extern "C" {

int meso_hip_init_equilibrium_noalloc(meso_hip_ctx *c, const double *rho, const double *u3, const double *T) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    const size_t nc = size_t(c->s.nc);
    if (nc * 5 * sizeof(double) > NOALLOC_SCRATCH_BYTES)
        return set_err(c, MESO_E_INVALID,
                       "meso_hip_init_equilibrium_noalloc: " + std::to_string(nc) + " cells need " +
                           std::to_string(nc * 40) + " B of scratch; its limit is 16 MiB (" +
                           std::to_string(NOALLOC_SCRATCH_BYTES / 40) + " cells)");
    if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES)
        return set_err(c, MESO_E_DEVICE, "meso_hip_init_equilibrium_noalloc: device index out of range");
    // one call at a time per device: the call owns the device's scratch until
    // its kernel has finished (init_common synchronises the stream)
    std::lock_guard<std::mutex> lock(g_scratch_mu[c->device]);
    if (!g_scratch_ptr[c->device]) {
        int cur = -1;
        HIP_TRY(c, hipGetDevice(&cur));
        if (cur != c->device) HIP_TRY(c, hipSetDevice(c->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != c->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return set_err(c, MESO_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[c->device] = p;
    }
    return init_common(c, rho, u3, T, nullptr, false, static_cast<double *>(g_scratch_ptr[c->device]));
}

int meso_hip_init_from_macros(meso_hip_ctx *c, const double *rho, const double *u3, const double *T,
                              const double *q3) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    return init_common(c, rho, u3, T, q3, true);
}

static int transpose(meso_hip_ctx *c, bool to_device, double *host, int which, int layout) {
    const int nc = c->s.nc, nv = c->s.nv, nvp = c->s.nvp;
    const size_t n = size_t(nc) * nv;
    std::vector<double> tmp;
    const double *src = host + size_t(which) * n;
    double *stage = nullptr;
    int rc;
    if ((rc = dalloc(c, &stage, n))) return rc;
    auto release = [&]() {
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)stage);
        if (it != c->allocs.end()) c->allocs.erase(it);
        (void)hipFree(stage);
        c->bytes -= int64_t(std::max<size_t>(n * 8, 16));
    };
    if (layout == MESO_LAYOUT_CELL_MAJOR) {  // host [c][v] -> stage [v][c] via CPU transpose
        tmp.resize(n);
        if (to_device)
            for (int cc = 0; cc < nc; ++cc)
                for (int v = 0; v < nv; ++v) tmp[size_t(v) * nc + cc] = src[size_t(cc) * nv + v];
        src = tmp.data();
    }
    dim3 grid((nc + 31) / 32, (nv + 31) / 32);
    void *dev = c->f[c->cur][which];
    unsigned char *xdev = c->fx[c->cur][which];
    // FP32X: FP32 words of the deviation plus extension bytes (float instantiation)
    const bool f32 = c->storage == MESO_STORAGE_FP32 || c->ext;
    const double *fref = f32 ? c->fref_d + size_t(which) * nv : nullptr;
    if (to_device) {
        HIP_TRY(c, hipMemcpyAsync(stage, src, n * 8, hipMemcpyHostToDevice, c->stream));
        if (f32)
            hipLaunchKernelGGL((k_transpose<float, true>), grid, dim3(BLOCK), 0, c->stream, stage, nullptr, nullptr,
                               (float *)dev, nc, nv, nvp, fref, c->slot_of, nullptr, xdev);
        else
            hipLaunchKernelGGL((k_transpose<double, true>), grid, dim3(BLOCK), 0, c->stream, stage, nullptr, nullptr,
                               (double *)dev, nc, nv, nvp, fref, c->slot_of, nullptr, nullptr);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
        if (f32)
            hipLaunchKernelGGL((k_transpose<float, false>), grid, dim3(BLOCK), 0, c->stream, nullptr, stage,
                               (const float *)dev, nullptr, nc, nv, nvp, fref, c->slot_of, xdev, nullptr);
        else
            hipLaunchKernelGGL((k_transpose<double, false>), grid, dim3(BLOCK), 0, c->stream, nullptr, stage,
                               (const double *)dev, nullptr, nc, nv, nvp, fref, c->slot_of, nullptr, nullptr);
        HIP_TRY(c, hipGetLastError());
        double *dst = host + size_t(which) * n;
        if (layout == MESO_LAYOUT_CELL_MAJOR) {
            tmp.resize(n);
            HIP_TRY(c, hipMemcpyAsync(tmp.data(), stage, n * 8, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            for (int cc = 0; cc < nc; ++cc)
                for (int v = 0; v < nv; ++v) dst[size_t(cc) * nv + v] = tmp[size_t(v) * nc + cc];
        } else {
            HIP_TRY(c, hipMemcpyAsync(dst, stage, n * 8, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
    }
    release();
    return MESO_OK;
}

int meso_hip_set_state(meso_hip_ctx *c, const double *f, int32_t layout, const double *rho, const double *u3,
                       const double *T, const double *q3) {
    if (!c || !f || !rho || !u3) return set_err(c, MESO_E_INVALID, "null argument");
    if (layout != MESO_LAYOUT_VELOCITY_MAJOR && layout != MESO_LAYOUT_CELL_MAJOR)
        return set_err(c, MESO_E_INVALID, "bad layout");
    int rc;
    if ((rc = set_reference_from(c, rho, c->s.model == MESO_MODEL_CDUGKS_SHAKHOV ? T : nullptr))) return rc;
    for (int k = 0; k < c->nd; ++k)
        if ((rc = transpose(c, true, const_cast<double *>(f), k, layout))) return rc;
    const size_t nc = c->s.nc;
    const int i = c->cur;
    HIP_TRY(c, hipMemcpyAsync(c->rho[i], rho, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->u[i], u3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->T[i], 0, nc * 8, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->q[i], 0, 3 * nc * 8, c->stream));
    if (T) HIP_TRY(c, hipMemcpyAsync(c->T[i], T, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (q3) HIP_TRY(c, hipMemcpyAsync(c->q[i], q3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
    if ((rc = reset_flags(c))) return rc;
    if ((rc = copy_residual_base(c))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int meso_hip_get_state(meso_hip_ctx *c, double *f, int32_t layout) {
    if (!c || !f) return set_err(c, MESO_E_INVALID, "null argument");
    if (layout != MESO_LAYOUT_VELOCITY_MAJOR && layout != MESO_LAYOUT_CELL_MAJOR)
        return set_err(c, MESO_E_INVALID, "bad layout");
    for (int k = 0; k < c->nd; ++k) {
        int rc = transpose(c, false, f, k, layout);
        if (rc) return rc;
    }
    return MESO_OK;
}

// Launches one step (no synchronisation).

static int step_one(meso_hip_ctx *c) {
    KArgs a = make_args(c);
    const int rc = step_dispatch(c, a);
    if (rc) return rc;
    if (!uses_small(c)) hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, c->stream, c->flags + 4);
    HIP_TRY(c, hipGetLastError());
    c->cur = 1 - c->cur;
    return MESO_OK;
}

// The executable graph of two steps from the current buffer parity, captured
// (or captured again, after an argument change) on the context's stream.
static int pair_graph(meso_hip_ctx *c) {
    const int p = c->cur;
    if (c->gexec[p] && c->gver[p] == c->kver) return MESO_OK;
    if (c->gexec[p]) (void)hipGraphExecDestroy(c->gexec[p]);
    c->gexec[p] = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = step_one(c);
    if (rc == MESO_OK) rc = step_one(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = p;  // capture launches nothing
    if (rc == MESO_OK && e != hipSuccess) rc = set_err(c, MESO_E_DEVICE, std::string("step capture: ") + hipGetErrorString(e));
    if (rc == MESO_OK) {
        const hipError_t ei = hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            c->gexec[p] = nullptr;
            rc = set_err(c, MESO_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    if (rc == MESO_OK) c->gver[p] = c->kver;
    return rc;
}

// Launches up to `left` steps (two as one graph replay when the context
// replays captured pairs); returns the steps launched in *n.
static int step_issue(meso_hip_ctx *c, int left, int *n) {
    *n = 0;
    if (c->graphs && !c->profiling && left >= 2) {
        int rc = pair_graph(c);
        if (rc) return rc;
        HIP_TRY(c, hipGraphLaunch(c->gexec[c->cur], c->stream));
        *n = 2;
        return MESO_OK;
    }
    const int rc = step_one(c);
    if (rc == MESO_OK) *n = 1;
    return rc;
}

// Waits for the issued steps of a run of nsteps and reports them.
static int step_finish(meso_hip_ctx *c, int32_t nsteps, meso_step_report *rep) {
    int rc;
    if ((rc = sync_flags(c))) return rc;
    const int bad = c->flags_host[0];
    int done = nsteps;
    if (bad > 0) {
        const int bs = c->flags_host[2];
        done = std::max(0, std::min(nsteps, bs - c->steps_done + 1));
        // the buffers of the halted steps were not written: f^n of the bad
        // step's successor sits in buffer parity of the last written step
        const int skipped = nsteps - done;
        if (skipped % 2) c->cur = 1 - c->cur;
    }
    c->steps_done += done;
    if (rep) {
        rep->steps_done = done;
        rep->bad_cells = bad;
        rep->first_bad_cell = bad > 0 ? c->flags_host[1] : -1;
        rep->masked_cells = c->flags_host[5] - c->masked_seen;
    }
    c->masked_seen = c->flags_host[5];
    return bad > 0 ? set_err(c, MESO_E_NUMERIC, "NaN density or T<0 in a cell") : MESO_OK;
}

int meso_hip_step(meso_hip_ctx *c, int32_t nsteps, meso_step_report *rep) {
    if (!c || nsteps < 0) return set_err(c, MESO_E_INVALID, "bad argument");
    int rc, n = 0;
    for (int i = 0; i < nsteps; i += n)
        if ((rc = step_issue(c, nsteps - i, &n))) return rc;
    return step_finish(c, nsteps, rep);
}

int meso_hip_set_graphs(meso_hip_ctx *c, int32_t enable) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    c->graphs = enable != 0;
    return MESO_OK;
}

int meso_hip_set_vchunk(meso_hip_ctx *c, int32_t vchunk) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    if (!c->patches) return vchunk == 0 ? MESO_OK : set_err(c, MESO_E_INVALID, "the case runs the per-cell kernels");
    const int nv = c->s.nv;
    int ncu = 256;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    const int vc = vchunk == 0 ? default_vchunk(nv, std::min(c->n_up, c->n_mp), ncu) : vchunk;
    if (vc < 64 || vc % 64 != 0) return set_err(c, MESO_E_INVALID, "vchunk must be a positive multiple of 64");
    if ((nv + vc - 1) / vc > c->p_chunk_cap) return set_err(c, MESO_E_INVALID, "vchunk too small for the partial-sum buffers");
    if (vc != c->p_vchunk) set_vchunk(c, vc);
    return MESO_OK;
}

int meso_hip_step_batch(meso_hip_ctx *const *ctxs, int32_t n, int32_t nsteps, meso_step_report *reps) {
    if (!ctxs || n < 0 || nsteps < 0) return set_err(nullptr, MESO_E_INVALID, "bad argument");
    for (int k = 0; k < n; ++k)
        if (!ctxs[k]) return set_err(nullptr, MESO_E_INVALID, "null context in the batch");
    // round-robin issue: step i of every context before step i + 1 of any, each
    // on its context's stream; then one wait per context
    std::vector<int> rc(n, MESO_OK), issued(n, 0);
    for (bool more = true; more;) {
        more = false;
        for (int k = 0; k < n; ++k)
            if (rc[k] == MESO_OK && issued[k] < nsteps) {
                int got = 0;
                rc[k] = step_issue(ctxs[k], nsteps - issued[k], &got);
                issued[k] += got;
                more = more || (rc[k] == MESO_OK && issued[k] < nsteps);
            }
    }
    int first = MESO_OK;
    for (int k = 0; k < n; ++k) {
        meso_step_report r{};
        int f = step_finish(ctxs[k], issued[k], &r);
        if (rc[k] != MESO_OK) f = rc[k];
        if (reps) reps[k] = r;
        if (first == MESO_OK && f != MESO_OK) {
            first = f;
            set_err(nullptr, f, ctxs[k]->err);
        }
    }
    return first;
}

// Rebuilds and uploads the patch sets for the requested sizes (velocity chunk
// kept); MESO_E_INVALID when they do not fit.
static int apply_patches(meso_hip_ctx *c, int up, int mom, int cells) {
    c->preq_up = up;
    c->preq_mom = mom;
    c->preq_cells = cells;
    if (c->s.D == 2) {
        meso_patch::PatchSet<2> u, m;
        if (!setup_patches<2>(c, c->storage, true, u, m)) return MESO_E_INVALID;
        return upload_patches<2>(c, u, m);
    }
    meso_patch::PatchSet<3> u, m;
    if (!setup_patches<3>(c, c->storage, true, u, m)) return MESO_E_INVALID;
    return upload_patches<3>(c, u, m);
}

// Milliseconds of `steps` trial steps of the current state (computed, not
// committed), after one warm-up step.
static int time_trial(meso_hip_ctx *c, int steps, hipEvent_t e0, hipEvent_t e1, float *ms) {
    KArgs a = make_args(c);
    int rc = step_dispatch(c, a);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(e0, c->stream));
    for (int i = 0; i < steps && rc == MESO_OK; ++i) rc = step_dispatch(c, a);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(e1, c->stream));
    HIP_TRY(c, hipEventSynchronize(e1));
    HIP_TRY(c, hipEventElapsedTime(ms, e0, e1));
    return MESO_OK;
}

int meso_hip_autotune(meso_hip_ctx *c, int32_t trial_steps, int32_t *vchunk_out) {
    if (!c || trial_steps < 1) return set_err(c, MESO_E_INVALID, "bad argument");
    if (!c->patches) {  // the per-cell kernels have no tiles to choose
        if (vchunk_out) *vchunk_out = 0;
        return MESO_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    int rc = MESO_OK;
    // trial steps read the current state and write the other buffer without
    // making it current, so the state is not advanced
    struct Req {
        int up, mom, cells;
    };
    Req best{c->preq_up, c->preq_mom, c->preq_cells}, cur = best;
    float best_ms = 0.0f, ms = 0.0f;
    rc = time_trial(c, trial_steps, e0, e1, &best_ms);
    auto consider = [&](Req r) {
        if (rc) return;
        const int ar = apply_patches(c, r.up, r.mom, r.cells);
        cur = r;
        if (ar == MESO_E_INVALID) return;  // does not fit: skipped
        if ((rc = ar)) return;
        if ((rc = time_trial(c, trial_steps, e0, e1, &ms))) return;
        if (ms < best_ms) {
            best_ms = ms;
            best = r;
        }
    };
    // 1: the moment patches (kind and size), 2: the update patches, 3: the
    // velocity chunk; each stage keeps the fastest
    for (Req r : {Req{best.up, 16, 1}, Req{best.up, 32, 1}, Req{best.up, 32, 0}, Req{best.up, 64, 0}}) consider(r);
    const Req m = best;
    for (int up : {16, 24, 32}) consider(Req{up, m.mom, m.cells});
    if (rc == MESO_OK && (cur.up != best.up || cur.mom != best.mom || cur.cells != best.cells))
        rc = apply_patches(c, best.up, best.mom, best.cells);
    if (rc == MESO_OK) {
        const int nv = c->s.nv, base = c->p_chunks, keep = c->p_vchunk;
        int bvc = keep;
        for (int k : {base / 2, 2 * base}) {
            if (k < 1 || k > c->p_chunk_cap) continue;
            const int vc = ((nv + k - 1) / k + 63) / 64 * 64;
            if (vc == keep || (nv + vc - 1) / vc > c->p_chunk_cap) continue;
            set_vchunk(c, vc);
            if ((rc = time_trial(c, trial_steps, e0, e1, &ms))) break;
            if (ms < best_ms) {
                best_ms = ms;
                bvc = vc;
            }
        }
        set_vchunk(c, bvc);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = sync_flags(c))) return rc;
    if (c->flags_host[0] > 0) return set_err(c, MESO_E_NUMERIC, "NaN density or T<0 in a trial step");
    if (vchunk_out) *vchunk_out = c->p_vchunk;
    return MESO_OK;
}

int meso_hip_tiles(const meso_hip_ctx *c, meso_tile_config *out) {
    if (!c || !out) return MESO_E_INVALID;
    meso_tile_config t{};
    t.patches = c->patches ? 1 : 0;
    t.update_owned = c->n_int;
    if (c->patches) {
        t.update_cells = c->up_maxo;
        t.moment_by_cells = c->mp_cells ? 1 : 0;
        t.moment_size = c->mp_size;
        t.vchunk = c->p_vchunk;
        t.moment_lds = int(c->lds_mp);
        t.update_lds = int(c->lds_up);
        const bool f32 = c->storage == MESO_STORAGE_FP32;
        t.moment_vt = f32 ? 64 : 32;
        t.update_vt = c->s.D == 2 ? (f32 ? upd_vt<float, 2>() : upd_vt<double, 2>())
                                  : (f32 ? upd_vt<float, 3>() : upd_vt<double, 3>());
    }
    *out = t;
    return MESO_OK;
}

int meso_hip_get_macros(meso_hip_ctx *c, double *rho, double *u3, double *T, double *q3) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    const size_t nc = c->s.nc;
    const int i = c->cur;
    if (rho) HIP_TRY(c, hipMemcpyAsync(rho, c->rho[i], nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (u3) HIP_TRY(c, hipMemcpyAsync(u3, c->u[i], 3 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (T) HIP_TRY(c, hipMemcpyAsync(T, c->T[i], nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (q3) HIP_TRY(c, hipMemcpyAsync(q3, c->q[i], 3 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int meso_hip_get_flux_moments(meso_hip_ctx *c, double *m0, double *m1, double *m2) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    const size_t nc = c->s.nc;
    std::vector<double> h(5 * nc);
    HIP_TRY(c, hipMemcpyAsync(h.data(), c->fm, 5 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < nc; ++k) {
        if (m0) m0[k] = h[5 * k];
        if (m1)
            for (int i = 0; i < 3; ++i) m1[3 * k + i] = h[5 * k + 1 + i];
        if (m2) m2[k] = h[5 * k + 4];
    }
    return MESO_OK;
}

int meso_hip_residual(meso_hip_ctx *c, double out[8]) {
    if (!c || !out) return set_err(c, MESO_E_INVALID, "null argument");
    const int nc = c->s.nc, nblk = (nc + BLOCK - 1) / BLOCK;
    const int i = c->cur;
    // total volume on the host: sum of cell volumes as meso accumulates it
    static_assert(sizeof(double) == 8, "");
    std::vector<double> vol(nc);
    HIP_TRY(c, hipMemcpyAsync(vol.data(), c->cell_vol, nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double total = 0.0;
    for (double v : vol) total += v;
    hipLaunchKernelGGL(k_residual, dim3(nblk), dim3(BLOCK), 0, c->stream, c->rho[i], c->u[i], c->T[i], c->q[i],
                       c->rho_res, c->u_res, c->T_res, c->q_res, c->cell_vol, total, nc,
                       c->s.model == MESO_MODEL_CDUGKS_SHAKHOV ? 1 : 0, c->res_part);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, c->stream, c->res_part, nblk, c->res_out);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(out, c->res_out, 8 * 8, hipMemcpyDeviceToHost, c->stream));
    int rc = copy_residual_base(c);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int meso_hip_profile(meso_hip_ctx *c, int32_t enable) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->recs.clear();
    c->totals.clear();
    c->profiling = enable != 0;
    return MESO_OK;
}

int meso_hip_kernel_times(meso_hip_ctx *c, int32_t max_entries, const char **names, double *total_ms,
                          int32_t *launches, int32_t *n_entries) {
    if (!c || !n_entries) return set_err(c, MESO_E_INVALID, "null argument");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto &r : c->recs) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        bool found = false;
        for (auto &t : c->totals)
            if (std::strcmp(t.first, r.name) == 0) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
            }
        if (!found) c->totals.push_back({r.name, {double(ms), 1}});
    }
    c->recs.clear();
    int n = 0;
    for (auto &t : c->totals) {
        if (n >= max_entries) break;
        if (names) names[n] = t.first;
        if (total_ms) total_ms[n] = t.second.first;
        if (launches) launches[n] = t.second.second;
        ++n;
    }
    *n_entries = n;
    return MESO_OK;
}

int64_t meso_hip_device_bytes(const meso_hip_ctx *c) { return c ? c->bytes : 0; }

int meso_hip_sync(meso_hip_ctx *c) {
    if (!c) return set_err(c, MESO_E_INVALID, "null context");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

}  // extern "C"

// ============================================================== fused batch
// A fixed set of contexts advanced together: the contexts whose step runs the
// patch kernels are grouped by kernel instance, and each stage of a step is
// one launch over the blocks of every context of a group. A block reads its
// (context, block) pair from a per-block map and the context's kernel
// arguments from a device table, so every context computes exactly what its
// own launches compute (same blocks, same arithmetic: bit-identical).
namespace {

enum BatchStage { BS_MOM, BS_FEQ, BS_BND, BS_PREP, BS_UPD, BS_FIN, BS_N };

struct BatchGroup {
    int sub = 0;                         // the sub-batch (stream) that launches it
    std::vector<int> members;            // positions in meso_hip_batch::ctx
    int blocks[BS_N] = {};               // grid of each stage (0: no launch)
    int2 *map[BS_N] = {};                // per block: (member, block of its own grid)
    KArgs *tab[2] = {nullptr, nullptr};  // per phase: the members' arguments
    size_t lds_mp = 0, lds_up = 0, lds_feq = 0;
};

// Key of the kernel instance a context's step launches (contexts that share
// it share the launches).
long batch_key(const meso_hip_ctx *c) {
    return ((((long(c->s.model) * 4 + c->s.D) * 2 + (c->storage == MESO_STORAGE_FP32)) * 2 + c->ext) * 3 +
            (c->nt > 0 ? (face_sep(c) ? 2 : 1) : 0)) * 4 + c->s.limiter * 2 + c->mp_cells;
}

// A context can join the fused launches when its whole step is the patch
// path: no mirror-boundary cells (their own per-cell launch), no masked
// groups (the repair launch), not profiling (per-launch events).
bool batch_fusable(const meso_hip_ctx *c) {
    return c->patches && c->s.model != DUGKS_BGK && c->n_bcell == 0 && c->n_int > 0 && !c->has_mask &&
           !c->profiling;
}

// Per-block map of one stage over the members: member k has n[k] blocks of
// work w[k] each, in its own logical block order. Hardware block B runs on XCD
// B % 8, so the j-th block of XCD x is B = 8 j + x. Every XCD takes a
// contiguous eighth of every member's blocks (neighbouring patches share that
// XCD's L2, and each XCD gets the same mix of light and heavy blocks whatever
// their real cost), members with the heaviest blocks first. Short XCD lists
// are padded with (-1, -1).
std::vector<int2> batch_map(const std::vector<int> &n, const std::vector<double> &w) {
    constexpr int NX = 8;
    const int m = int(n.size());
    std::vector<int> order(m);
    for (int k = 0; k < m; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x] > w[y]; });
    std::vector<std::vector<int2>> lists(NX);
    for (int k : order)
        for (int x = 0; x < NX; ++x) {
            const int b0 = int(int64_t(n[k]) * x / NX), b1 = int(int64_t(n[k]) * (x + 1) / NX);
            for (int b = b0; b < b1; ++b) lists[x].push_back(make_int2(k, b));
        }
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    std::vector<int2> map(len * NX, make_int2(-1, -1));
    for (int x = 0; x < NX; ++x)
        for (size_t j = 0; j < lists[x].size(); ++j) map[j * NX + x] = lists[x][j];
    return map;
}

}  // namespace

struct meso_hip_batch {
    std::vector<meso_hip_ctx *> ctx;
    std::vector<int> solo;               // contexts stepped with their own launches
    std::vector<BatchGroup> groups;
    std::vector<int> kver0, cur0;        // per context, at the last build
    // sub-batches: the fused contexts are split (by DV-cells) over up to
    // nsub_req streams; each sub-batch launches its groups' stages on its own
    // stream, so the small stages of one overlap the large ones of another
    int nsub_req = 4;
    int nsub = 0;
    std::vector<int **> ctrs;            // per sub-batch: its contexts' step counters
    std::vector<int> nctr;
    int nfused = 0;
    int phase = 0;                       // fused steps since the build, mod 2
    bool built = false;
    int device = 0;
    hipStream_t stream = nullptr;        // origin stream (sub-batch 0)
    hipEvent_t ev = nullptr;
    std::vector<hipStream_t> sstream;    // sub-batch streams 1.. (0 is `stream`)
    std::vector<hipEvent_t> sev;         // join events of the sub-batch streams
    bool graphs = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<void *> dmem;            // device memory of the plan
    bool rechunk = true;                 // set the fused contexts' velocity chunks for the batch grid
    std::string err;
};

namespace {

int berr(meso_hip_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    g_hip_err = m;
    return code;
}

#define BHIP(b, expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return berr(b, MESO_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void batch_release(meso_hip_batch *b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t st : b->sstream) (void)hipStreamSynchronize(st);
    for (auto &g : b->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    for (void *p : b->dmem) (void)hipFree(p);
    b->dmem.clear();
    b->groups.clear();
    b->solo.clear();
    b->ctrs.clear();
    b->nctr.clear();
    b->nfused = 0;
    b->nsub = 0;
    b->built = false;
}

template <class P>
int batch_upload(meso_hip_batch *b, P **dst, const std::vector<P> &v) {
    void *q = nullptr;
    BHIP(b, hipMalloc(&q, std::max<size_t>(v.size() * sizeof(P), 16)));
    b->dmem.push_back(q);
    if (!v.empty()) BHIP(b, hipMemcpyAsync(q, v.data(), v.size() * sizeof(P), hipMemcpyHostToDevice, b->stream));
    *dst = reinterpret_cast<P *>(q);
    return MESO_OK;
}

// Velocity chunk of a context in the fused grid. Alone, a context takes
// enough chunks that its patch grid has 16 blocks per CU (default_vchunk);
// in a batch the contexts fill the grid together, so each takes only the
// chunks its share of the batch's DV-cells needs (at least one per 2560
// velocities): fewer, larger blocks, and fewer partial sums per face for
// k_face_eq to add (the 64-case sweep: up to 128 partial sums per face
// otherwise, and its face_eq launch 10 ms).
int batch_vchunk(const meso_hip_ctx *c, double share, int ncu) {
    const int nv = c->s.nv;
    const int npatch = std::max(1, std::min(c->n_up, c->n_mp));
    int chunks = std::max((nv + 2559) / 2560, int(std::ceil(share * 16.0 * ncu / npatch)));
    chunks = std::max(1, std::min({chunks, (nv + 63) / 64, c->p_chunk_cap}));
    return ((nv + chunks - 1) / chunks + 63) / 64 * 64;
}

// Groups, per-stage block maps and argument tables from the contexts' current
// state; phase 0 is their current buffer parity. The fused contexts' velocity
// chunks are set for the batch grid first (batch_vchunk).
int batch_build(meso_hip_batch *b) {
    batch_release(b);
    constexpr int WPB = BLOCK / 64;
    const int n = int(b->ctx.size());
    {
        int ncu = 256;
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device);
        double total = 0.0;
        for (meso_hip_ctx *c : b->ctx)
            if (batch_fusable(c)) total += double(c->s.nc) * c->s.nv;
        for (meso_hip_ctx *c : b->ctx)
            if (batch_fusable(c) && b->rechunk) {
                const int vc = batch_vchunk(c, double(c->s.nc) * c->s.nv / total, ncu);
                if (vc != c->p_vchunk) set_vchunk(c, vc);
            }
    }
    b->kver0.assign(n, 0);
    b->cur0.assign(n, 0);
    std::vector<int> fus;
    for (int k = 0; k < n; ++k) {
        meso_hip_ctx *c = b->ctx[k];
        b->kver0[k] = c->kver;
        b->cur0[k] = c->cur;
        if (batch_fusable(c)) fus.push_back(k);
        else b->solo.push_back(k);
    }
    b->nfused = int(fus.size());
    b->nsub = std::max(1, std::min(b->nsub_req, b->nfused));
    // sub-batches: largest context first to the least loaded
    std::vector<int> sub_of(n, 0);
    {
        std::vector<int> ord = fus;
        auto dv = [&](int k) { return double(b->ctx[k]->s.nc) * b->ctx[k]->s.nv; };
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return dv(x) > dv(y); });
        std::vector<double> load(b->nsub, 0.0);
        for (int k : ord) {
            const int sb = int(std::min_element(load.begin(), load.end()) - load.begin());
            sub_of[k] = sb;
            load[sb] += dv(k);
        }
    }
    std::vector<std::pair<int, long>> keys;
    std::vector<std::vector<int *>> ctrs(b->nsub);
    for (int k : fus) {
        meso_hip_ctx *c = b->ctx[k];
        const std::pair<int, long> key(sub_of[k], batch_key(c));
        size_t g = 0;
        while (g < keys.size() && keys[g] != key) ++g;
        if (g == keys.size()) {
            keys.push_back(key);
            b->groups.emplace_back();
            b->groups.back().sub = key.first;
        }
        b->groups[g].members.push_back(k);
        ctrs[sub_of[k]].push_back(c->flags + 4);
    }
    b->ctrs.assign(b->nsub, nullptr);
    b->nctr.assign(b->nsub, 0);
    for (int sb = 0; sb < b->nsub; ++sb) {
        b->nctr[sb] = int(ctrs[sb].size());
        if (b->nctr[sb]) {
            int rc = batch_upload(b, &b->ctrs[sb], ctrs[sb]);
            if (rc) return rc;
        }
    }
    while (int(b->sstream.size()) < b->nsub - 1) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        BHIP(b, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->sstream.push_back(st);
        BHIP(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->sev.push_back(ev);
    }
    for (BatchGroup &g : b->groups) {
        const int m = int(g.members.size());
        std::vector<int> nb[BS_N];
        std::vector<double> wb[BS_N];
        std::vector<KArgs> tab[2];
        for (int i = 0; i < m; ++i) {
            meso_hip_ctx *c = b->ctx[g.members[i]];
            const StepConst &s = c->s;
            const double dvc = double(s.nc) * s.nv;
            const int fpb = face_sep(c) ? BLOCK / FEQ_LP : WPB;  // faces per block of the face stage
            const int cnt[BS_N] = {c->n_mp * c->p_chunks, (s.nf + fpb - 1) / fpb, s.nb, (c->n_int + WPB - 1) / WPB,
                                   c->n_up * c->p_chunks,
                                   s.model == CDUGKS_SHAKHOV ? (s.nc + BLOCK - 1) / BLOCK : 0};
            const double work[BS_N] = {dvc, double(s.nf) * std::max(c->nt, 1), double(s.nb) * s.nv,
                                       double(c->n_int) * std::max(c->nt, 1), dvc, double(s.nc)};
            for (int st = 0; st < BS_N; ++st) {
                nb[st].push_back(cnt[st]);
                wb[st].push_back(cnt[st] > 0 ? work[st] / cnt[st] : 0.0);
            }
            g.lds_mp = std::max(g.lds_mp, c->lds_mp);
            g.lds_up = std::max(g.lds_up, c->lds_up);
            const size_t tsz = c->storage == MESO_STORAGE_FP32 ? 4 : 8;
            g.lds_feq = std::max(g.lds_feq, c->nt > 0 ? size_t(WPB) * c->nt * tsz : size_t(0));
            for (int p = 0; p < 2; ++p) {
                const int cur = c->cur;
                c->cur = cur ^ p;
                tab[p].push_back(make_args(c));
                c->cur = cur;
            }
        }
        for (int st = 0; st < BS_N; ++st) {
            const std::vector<int2> map = batch_map(nb[st], wb[st]);
            bool any = false;
            for (int x : nb[st]) any = any || x > 0;
            g.blocks[st] = any ? int(map.size()) : 0;
            if (!any) continue;
            int rc = batch_upload(b, &g.map[st], map);
            if (rc) return rc;
        }
        for (int p = 0; p < 2; ++p) {
            int rc = batch_upload(b, &g.tab[p], tab[p]);
            if (rc) return rc;
        }
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    b->phase = 0;
    b->built = true;
    return MESO_OK;
}

// The plan is stale when a context's kernel arguments changed, or it was
// stepped on its own, since the build.
bool batch_stale(const meso_hip_batch *b) {
    if (!b->built) return true;
    for (size_t k = 0; k < b->ctx.size(); ++k) {
        const meso_hip_ctx *c = b->ctx[k];
        if (c->kver != b->kver0[k] || c->cur != (b->cur0[k] ^ b->phase)) return true;
        if (batch_fusable(c) != (std::find(b->solo.begin(), b->solo.end(), int(k)) == b->solo.end())) return true;
    }
    return false;
}

#define BLAUNCH(st, kern, grid, block, shmem, ...) \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, st, __VA_ARGS__)

hipStream_t sub_stream(const meso_hip_batch *b, int sb) { return sb == 0 ? b->stream : b->sstream[sb - 1]; }

// The sub-batch streams follow the origin stream (fork) ...
int batch_fork(meso_hip_batch *b) {
    if (b->nsub < 2) return MESO_OK;
    BHIP(b, hipEventRecord(b->ev, b->stream));
    for (int sb = 1; sb < b->nsub; ++sb) BHIP(b, hipStreamWaitEvent(sub_stream(b, sb), b->ev, 0));
    return MESO_OK;
}

// ... and the origin stream follows them (join).
int batch_join(meso_hip_batch *b) {
    for (int sb = 1; sb < b->nsub; ++sb) {
        BHIP(b, hipEventRecord(b->sev[sb - 1], sub_stream(b, sb)));
        BHIP(b, hipStreamWaitEvent(b->stream, b->sev[sb - 1], 0));
    }
    return MESO_OK;
}

template <class K>
int batch_lds_attr(meso_hip_batch *b, K *kern, size_t lds) {
    if (lds > 65536)
        BHIP(b, hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)));
    return MESO_OK;
}

// One step of a group: the launches of step_patch, each over every member.
template <int MODEL, class T, bool TAB, int D, bool LIM, bool X>
int batch_group_step(meso_hip_batch *b, const BatchGroup &g, int phase, bool cells) {
    constexpr int VTM = sizeof(T) == 4 ? 64 : 32;
    constexpr int VTU = upd_vt<T, D>();
    constexpr bool XM = X && sizeof(T) == 8;  // FP32E moments read the words only (the FP32 kernel)
    const KArgs *tab = g.tab[phase];
    const hipStream_t st = sub_stream(b, g.sub);
    int rc;
    if (cells) {
        auto *km = &kb_patch_moments<MODEL, T, VTM, D, LIM, true, XM>;
        if ((rc = batch_lds_attr(b, km, g.lds_mp))) return rc;
        BLAUNCH(st, km, g.blocks[BS_MOM], BLOCKP, g.lds_mp, tab, g.map[BS_MOM]);
    } else {
        auto *km = &kb_patch_moments<MODEL, T, VTM, D, LIM, false, XM>;
        if ((rc = batch_lds_attr(b, km, g.lds_mp))) return rc;
        BLAUNCH(st, km, g.blocks[BS_MOM], BLOCKP, g.lds_mp, tab, g.map[BS_MOM]);
    }
    if (MODEL == CDUGKS_SHAKHOV && TAB && face_sep(b->ctx[g.members[0]]))
        BLAUNCH(st, (kb_face_eq_sep<D, T, FEQ_LP>), g.blocks[BS_FEQ], BLOCK, 0, tab, g.map[BS_FEQ]);
    else
        BLAUNCH(st, (kb_face_eq<MODEL, D, T, TAB>), g.blocks[BS_FEQ], BLOCK, TAB ? g.lds_feq : 0, tab, g.map[BS_FEQ]);
    if (g.blocks[BS_BND]) BLAUNCH(st, (kb_boundary<MODEL, D, T, LIM>), g.blocks[BS_BND], BLOCK, 0, tab, g.map[BS_BND]);
    BLAUNCH(st, (kb_cell_prep<MODEL, T, TAB, D>), g.blocks[BS_PREP], BLOCK, 0, tab, g.map[BS_PREP]);
    auto *ku = &kb_patch_update<MODEL, T, VTU, TAB, D, LIM, X>;
    if ((rc = batch_lds_attr(b, ku, g.lds_up))) return rc;
    BLAUNCH(st, ku, g.blocks[BS_UPD], BLOCKP, g.lds_up, tab, g.map[BS_UPD]);
    if (MODEL == CDUGKS_SHAKHOV) BLAUNCH(st, (kb_cell_finish<MODEL>), g.blocks[BS_FIN], BLOCK, 0, tab, g.map[BS_FIN]);
    BHIP(b, hipGetLastError());
    return MESO_OK;
}

int batch_group_dispatch(meso_hip_batch *b, const BatchGroup &g, int phase) {
    const meso_hip_ctx *c = b->ctx[g.members[0]];
    const int m = c->s.model, D = c->s.D;
    // as step_dispatch: axis tables and the limiter only for Shakhov
    const bool sh = m == CDUGKS_SHAKHOV;
    const bool f32 = c->storage == MESO_STORAGE_FP32, tab = sh && c->nt > 0, lim = sh && c->s.limiter != 0,
               x = c->ext, cells = c->mp_cells;
#define MESO_BL(M, DD, TB, L)                                                                                    \
    if (m == M && D == DD && tab == TB && lim == L) {                                                            \
        if (f32)                                                                                                 \
            return x ? batch_group_step<M, float, TB, DD, L, true>(b, g, phase, cells)                          \
                     : batch_group_step<M, float, TB, DD, L, false>(b, g, phase, cells);                        \
        return x ? batch_group_step<M, double, TB, DD, L, true>(b, g, phase, cells)                             \
                 : batch_group_step<M, double, TB, DD, L, false>(b, g, phase, cells);                           \
    }
#define MESO_BD(DD)                                                                                              \
    MESO_BL(CDUGKS_SHAKHOV, DD, true, false) MESO_BL(CDUGKS_SHAKHOV, DD, true, true)                             \
    MESO_BL(CDUGKS_SHAKHOV, DD, false, false) MESO_BL(CDUGKS_SHAKHOV, DD, false, true)                           \
    MESO_BL(CDUGKS_BGK, DD, false, false)
    MESO_BD(2)
    MESO_BD(3)
#undef MESO_BD
#undef MESO_BL
    return berr(b, MESO_E_INVALID, "batch: unsupported kernel instance");
}

// One fused step of every group (and the step counters), then the host-side
// buffer parity of the fused contexts.
int batch_fused_step(meso_hip_batch *b) {
    for (const BatchGroup &g : b->groups) {
        const int rc = batch_group_dispatch(b, g, b->phase);
        if (rc) return rc;
    }
    for (int sb = 0; sb < b->nsub; ++sb)
        if (b->nctr[sb]) BLAUNCH(sub_stream(b, sb), kb_step_end, (b->nctr[sb] + 63) / 64, 64, 0, b->ctrs[sb], b->nctr[sb]);
    BHIP(b, hipGetLastError());
    for (const BatchGroup &g : b->groups)
        for (int k : g.members) b->ctx[k]->cur ^= 1;
    b->phase ^= 1;
    return MESO_OK;
}

// Two fused steps from the current phase as a graph (captured on first use).
int batch_pair_graph(meso_hip_batch *b) {
    const int p = b->phase;
    if (b->gexec[p]) return MESO_OK;
    std::vector<int> cur(b->ctx.size());
    for (size_t k = 0; k < b->ctx.size(); ++k) cur[k] = b->ctx[k]->cur;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_fork(b);
    if (rc == MESO_OK) rc = batch_fused_step(b);
    if (rc == MESO_OK) rc = batch_fused_step(b);
    if (rc == MESO_OK) rc = batch_join(b);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &gr);
    // capture launched nothing: undo the host-side parities of the captured steps
    b->phase = p;
    for (size_t k = 0; k < b->ctx.size(); ++k) b->ctx[k]->cur = cur[k];
    if (rc == MESO_OK && e != hipSuccess)
        rc = berr(b, MESO_E_DEVICE, std::string("batch capture: ") + hipGetErrorString(e));
    if (rc == MESO_OK) {
        const hipError_t ei = hipGraphInstantiate(&b->gexec[p], gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            b->gexec[p] = nullptr;
            rc = berr(b, MESO_E_DEVICE, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
        }
    }
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
}

}  // namespace

extern "C" {

int meso_hip_batch_create(meso_hip_ctx *const *ctxs, int32_t n, meso_hip_batch **out) {
    if (!ctxs || n <= 0 || !out) return set_err(nullptr, MESO_E_INVALID, "bad argument");
    *out = nullptr;
    for (int k = 0; k < n; ++k) {
        if (!ctxs[k]) return set_err(nullptr, MESO_E_INVALID, "null context in the batch");
        if (ctxs[k]->device != ctxs[0]->device)
            return set_err(nullptr, MESO_E_INVALID, "batch contexts must share one device");
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == ctxs[k]) return set_err(nullptr, MESO_E_INVALID, "a context appears twice in the batch");
    }
    auto *b = new meso_hip_batch();
    b->ctx.assign(ctxs, ctxs + n);
    b->device = ctxs[0]->device;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        const int rc = set_err(nullptr, MESO_E_DEVICE, std::string("batch stream: ") + hipGetErrorString(e));
        meso_hip_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MESO_OK;
}

void meso_hip_batch_destroy(meso_hip_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    batch_release(b);
    for (hipEvent_t e : b->sev) (void)hipEventDestroy(e);
    for (hipStream_t st : b->sstream) (void)hipStreamDestroy(st);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int meso_hip_batch_set_graphs(meso_hip_batch *b, int32_t enable) {
    if (!b) return set_err(nullptr, MESO_E_INVALID, "null batch");
    b->graphs = enable != 0;
    return MESO_OK;
}

int meso_hip_batch_set_streams(meso_hip_batch *b, int32_t n) {
    if (!b || n < 1 || n > 16) return set_err(nullptr, MESO_E_INVALID, "bad argument");
    if (n != b->nsub_req) b->built = false;
    b->nsub_req = n;
    return MESO_OK;
}

int meso_hip_batch_set_rechunk(meso_hip_batch *b, int32_t enable) {
    if (!b) return set_err(nullptr, MESO_E_INVALID, "null batch");
    if ((enable != 0) != b->rechunk) b->built = false;
    b->rechunk = enable != 0;
    return MESO_OK;
}

int meso_hip_batch_info(meso_hip_batch *b, int32_t *fused, int32_t *groups, int32_t *launches_per_step) {
    if (!b) return set_err(nullptr, MESO_E_INVALID, "null batch");
    if (batch_stale(b)) {
        (void)hipSetDevice(b->device);
        const int rc = batch_build(b);
        if (rc) return rc;
    }
    int l = 0;
    for (int sb = 0; sb < b->nsub; ++sb) l += b->nctr[sb] > 0;
    for (const BatchGroup &g : b->groups)
        for (int st = 0; st < BS_N; ++st) l += g.blocks[st] > 0;
    if (fused) *fused = b->nfused;
    if (groups) *groups = int(b->groups.size());
    if (launches_per_step) *launches_per_step = l;
    return MESO_OK;
}

int meso_hip_batch_step(meso_hip_batch *b, int32_t nsteps, meso_step_report *reports) {
    if (!b || nsteps < 0) return set_err(nullptr, MESO_E_INVALID, "bad argument");
    BHIP(b, hipSetDevice(b->device));
    if (batch_stale(b)) {
        const int rc = batch_build(b);
        if (rc) return rc;
    }
    const int n = int(b->ctx.size());
    // the batch stream follows earlier work on the contexts' streams
    for (int k = 0; k < n; ++k) {
        BHIP(b, hipEventRecord(b->ev, b->ctx[k]->stream));
        BHIP(b, hipStreamWaitEvent(b->stream, b->ev, 0));
    }
    // the sub-batch streams follow the origin for eager steps (forked); a
    // graph replay runs on the origin stream and forks inside, so the origin
    // first takes in the sub-batch streams' eager work (joined)
    bool forked = false;
    std::vector<int> rc(n, MESO_OK), issued(n, 0);
    int frc = MESO_OK, fdone = 0;
    // fused steps on the batch streams, round-robin with the solo contexts' steps
    for (bool more = true; more;) {
        more = false;
        if (frc == MESO_OK && fdone < nsteps && b->nfused) {
            const bool graph = b->graphs && nsteps - fdone >= 2;
            if (graph == forked) {
                frc = graph ? batch_join(b) : batch_fork(b);
                forked = !graph;
            }
            if (frc != MESO_OK) {
            } else if (graph) {
                frc = batch_pair_graph(b);
                if (frc == MESO_OK) {
                    const hipError_t e = hipGraphLaunch(b->gexec[b->phase], b->stream);
                    if (e != hipSuccess) frc = berr(b, MESO_E_DEVICE, std::string("batch graph launch: ") + hipGetErrorString(e));
                }
                if (frc == MESO_OK) fdone += 2;  // two steps: parities and phase are back where they were
            } else {
                frc = batch_fused_step(b);
                if (frc == MESO_OK) ++fdone;
            }
            more = more || (frc == MESO_OK && fdone < nsteps);
        }
        for (int k : b->solo)
            if (rc[k] == MESO_OK && issued[k] < nsteps) {
                int got = 0;
                rc[k] = step_issue(b->ctx[k], nsteps - issued[k], &got);
                issued[k] += got;
                more = more || (rc[k] == MESO_OK && issued[k] < nsteps);
            }
    }
    std::vector<char> fused(n, 0);
    for (const BatchGroup &g : b->groups)
        for (int k : g.members) {
            issued[k] = fdone;
            rc[k] = frc;
            fused[k] = 1;
        }
    // the contexts' streams (their reports, later calls) follow the batch
    if (forked) {
        const int rj = batch_join(b);
        if (rj) return rj;
    }
    BHIP(b, hipEventRecord(b->ev, b->stream));
    for (int k = 0; k < n; ++k) BHIP(b, hipStreamWaitEvent(b->ctx[k]->stream, b->ev, 0));
    int first = MESO_OK;
    for (int k = 0; k < n; ++k) {
        meso_step_report r{};
        int f = step_finish(b->ctx[k], issued[k], &r);
        if (rc[k] != MESO_OK) f = rc[k];
        if (reports) reports[k] = r;
        if (first == MESO_OK && f != MESO_OK) {
            first = f;
            const std::string msg = rc[k] != MESO_OK && fused[k] ? b->err : b->ctx[k]->err;
            berr(b, f, msg);
        }
    }
    // a halted context's buffers: step_finish put its parity back, and the
    // plan follows it at the next call (stale)
    return first;
}

}  // extern "C"
