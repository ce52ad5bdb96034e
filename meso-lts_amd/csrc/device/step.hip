// gfx950 DUGKS/CDUGKS step: kernels, device context and the C ABI of
// include/meso_hip.h (meso_hip_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "meso_hip.h"
#include "step_kernels.h"

using namespace meso_dev;

namespace {

constexpr int BLOCK = 256;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ const meso_mark_desc &mark_of(const KArgs &a, int m) {
    return reinterpret_cast<const meso_mark_desc *>(a.marks)[m];
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
template <int D, class T>
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
        fb = (sn >= VSMALL || sn < -VSMALL) ? fc + X : T(0.5) * (fc + fc + X + X);
    }
    return lerp_store<T>(fb, eqv, mix);
}

// Face-mixed pair (g, h) of boundary face f at velocity v.
template <int MODEL, int D, class T>
__device__ __forceinline__ void bnd_mix_pair(const KArgs &a, int f, int c, int v, const EqParams &ef, T &g, T &h) {
    T x[3];
    load_xi<T>(a, v, x);
    const T hd = T(a.s.half_dt);
    T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
    double sn = ndot(a, f, v);
    T mix = T(ef.mix);
    T gs, hs;
    eq_store<MODEL, D, T>(a, ef, x, v, gs, hs);
    g = bnd_mix<D, T>(a, (const T *)a.fin_g, f, c, sn, v, xh, gs, mix);
    h = MODEL == CDUGKS_SHAKHOV ? bnd_mix<D, T>(a, (const T *)a.fin_h, f, c, sn, v, xh, hs, mix) : T(0);
}

// Final boundary value at velocity v given the mixed value (gm_, hm_) there.
template <int MODEL, int D, class T>
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
            bnd_mix_pair<MODEL, D, T>(a, f, c, vm, ef, gr, hr);
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

// ================================================================ kernels
// (1) raw moments of f-bar+ per (face, velocity tile)
template <int MODEL, int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_face_moments(KArgs a) {
    if (*a.bad_count > 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *a.halt = 1;
        return;
    }
    constexpr int IPB = BLOCK / LP;
    constexpr int NCOL = ncol_face<MODEL, D>();
    const int nf = a.s.nf;
    const int nblk = (nf + IPB - 1) / IPB;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tile = b / nblk, fb = b - tile * nblk;
    const int slot = threadIdx.x / LP, li = threadIdx.x % LP;
    const int f = fb * IPB + slot;
    double acc[NCOL];
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = 0.0;
    if (f < nf) {
        const int2 oc = a.face_cells[f];
        const T hd = T(a.s.half_dt);
        const T *fg = (const T *)a.fin_g;
        const T *fh = (const T *)a.fin_h;
        const int v0 = tile * a.s.vt_f, v1 = min(v0 + a.s.vt_f, a.s.nv);
        for (int v = v0 + li; v < v1; v += LP) {
            T x[3];
            load_xi<T>(a, v, x);
            T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
            const double sn = ndot(a, f, v);
            const double X0 = a.xi[3 * v], X1 = a.xi[3 * v + 1], X2 = a.xi[3 * v + 2];
            const double wv = a.w[v];
            const double wg = wv * double(fbar_at<D, T>(a, fg, f, oc.x, oc.y, sn, v, xh));
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
            if (MODEL == CDUGKS_SHAKHOV) {
                const double wh = wv * double(fbar_at<D, T>(a, fh, f, oc.x, oc.y, sn, v, xh));
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
    }
#pragma unroll
    for (int k = 0; k < NCOL; ++k) acc[k] = seg_sum<LP>(acc[k]);
    if (f < nf && li == 0) {
        double *dst = a.part_f + (size_t(f) * a.s.ntile_f + tile) * NCOL;
#pragma unroll
        for (int k = 0; k < NCOL; ++k) dst[k] = acc[k];
    }
}

// (2) face macros, equilibrium constants and flux moments of interior faces
template <int MODEL, int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_face_eq(KArgs a) {
    if (*a.halt) return;
    constexpr int IPB = BLOCK / LP;
    constexpr int NCOL = ncol_face<MODEL, D>();
    constexpr int NFL = D + 2;
    const int nf = a.s.nf;
    const int slot = threadIdx.x / LP, li = threadIdx.x % LP;
    const int f = xcd_block(blockIdx.x, gridDim.x) * IPB + slot;
    const bool act = f < nf;
    EqParams e;
    double Fb[NFL];
    double acc[NFL];
#pragma unroll
    for (int k = 0; k < NFL; ++k) acc[k] = Fb[k] = 0.0;
    bool interior = false;
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
        const double n[3] = {a.face_n[3 * f], a.face_n[3 * f + 1], a.face_n[3 * f + 2]};
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
                double M3i = i < D ? M[Col<D>::M3 + i] : 0.0;
                double H1i = i < D ? M[Col<D>::H1 + i] : 0.0;
                double m3 = M3i - 2.0 * Mu[i] + uu * M1[i] - u[i] * trM2 + 2.0 * u[i] * M1u -
                            u[i] * uu * rho + H1i - u[i] * H0;
                q[i] = cq * m3;
            }
            make_shakhov(e, rho, u, T_, q, a.s.R, a.s.Pr, D, a.s.K);
            if constexpr (sizeof(T) == 4) relate_shakhov(e, a.ref);
            e.mix = hdt / (2.0 * tau + hdt);
            double M3n = 0.0, H1n = 0.0;
            for (int i = 0; i < D; ++i) {
                M3n += n[i] * M[Col<D>::M3 + i];
                H1n += n[i] * M[Col<D>::H1 + i];
            }
            Fb[D + 1] = M3n + H1n;
        } else {
            make_bgk(e, rho, u, a.s.RT);
            if constexpr (sizeof(T) == 4) relate_bgk(e, a.ref);
            e.mix = hdt / (2.0 * a.s.tau + hdt);
        }
        if (li == 0) a.face_eq[f] = e;
        interior = a.face_bnd[f] < 0;
        if (interior) {
            const int nv = a.s.nv;
            for (int v = li; v < nv; v += LP) {
                T x[3];
                load_xi<T>(a, v, x);
                const double X0 = a.xi[3 * v], X1 = a.xi[3 * v + 1], X2 = a.xi[3 * v + 2];
                const double wsn = a.w[v] * (n[0] * X0 + n[1] * X1 + n[2] * X2);
                T gs, hs;
                eq_store<MODEL, D, T>(a, e, x, v, gs, hs);
                const double wg = wsn * double(gs);
                acc[0] += wg;
                acc[1] += wg * X0;
                acc[2] += wg * X1;
                if (D == 3) acc[3] += wg * X2;
                if (MODEL == CDUGKS_SHAKHOV) acc[D + 1] += wg * (X0 * X0 + X1 * X1 + X2 * X2) + wsn * double(hs);
            }
        }
    }
    // the whole LP segment shares `act` and `interior`
#pragma unroll
    for (int k = 0; k < NFL; ++k) acc[k] = seg_sum<LP>(acc[k]);
    if (act && interior && li == 0) {
        if constexpr (sizeof(T) == 4) {  // flux moments of f_ref along n
            const double n[3] = {a.face_n[3 * f], a.face_n[3 * f + 1], a.face_n[3 * f + 2]};
            for (int k = 0; k < NFL; ++k)
                acc[k] += n[0] * a.rflux[k][0] + n[1] * a.rflux[k][1] + n[2] * a.rflux[k][2];
        }
        const double mix = e.mix;
        const int nfl = MODEL == CDUGKS_SHAKHOV ? D + 2 : D + 1;
        for (int k = 0; k < nfl; ++k) a.face_flux[size_t(f) * NFL + k] = (1.0 - mix) * Fb[k] + mix * acc[k];
        if (MODEL != CDUGKS_SHAKHOV) a.face_flux[size_t(f) * NFL + D + 1] = 0.0;
    }
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

// (3) boundary faces: one block per face
template <int MODEL, int D, class T>
__global__ __launch_bounds__(BLOCK) void k_boundary(KArgs a) {
    if (*a.halt) return;
    constexpr int NFL = D + 2;
    __shared__ double lds[(NFL + 2) * (BLOCK / 64)];
    const int b = blockIdx.x;
    const int f = a.bnd_face[b];
    const int c = a.face_cells[f].x;
    const meso_mark_desc &mk = mark_of(a, a.bnd_mark[b]);
    const EqParams ef = a.face_eq[f];
    const int nv = a.s.nv;
    BndParams P;
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
            EqParams e1;
            if (MODEL == CDUGKS_SHAKHOV) make_shakhov(e1, 1.0, uw, Tw, zero, R, Pr, D, K);
            else make_bgk(e1, 1.0, uw, a.s.RT);
            double s[2] = {0.0, 0.0};
            for (int v = threadIdx.x; v < nv; v += BLOCK) {
                const double kn = -ndot(a, f, v);
                if (kn >= VSMALL) {
                    const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
                    double gm, hm;
                    if (MODEL == CDUGKS_SHAKHOV) shakhov_eq<double>(e1, X, D, K, gm, hm);
                    else gm = bgk_eq<double>(e1, X);
                    s[1] += kn * a.w[v] * gm;
                } else {
                    T gmx, hmx;
                    bnd_mix_pair<MODEL, D, T>(a, f, c, v, ef, gmx, hmx);
                    s[0] -= kn * a.w[v] * full_value<T>(a, gmx, 0, v);
                }
            }
            block_sum<2>(s, lds);
            const double rho_w = s[0] / s[1];
            if (MODEL == CDUGKS_SHAKHOV) make_shakhov(P.in, rho_w, uw, mk.temperature, zero, R, Pr, D, K);
            else make_bgk(P.in, rho_w, uw, a.s.RT);
        } break;
        default: break;
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
        T gmx, hmx, g, h;
        bnd_mix_pair<MODEL, D, T>(a, f, c, v, ef, gmx, hmx);
        const double sn = ndot(a, f, v);
        bnd_value<MODEL, D, T>(a, P, f, c, v, sn, ef, gmx, hmx, g, h);
        const double X0 = a.xi[3 * v], X1 = a.xi[3 * v + 1], X2 = a.xi[3 * v + 2];
        const double wsn = a.w[v] * sn;
        const double wg = wsn * full_value<T>(a, g, 0, v);
        acc[0] += wg;
        acc[1] += wg * X0;
        acc[2] += wg * X1;
        if (D == 3) acc[3] += wg * X2;
        if (MODEL == CDUGKS_SHAKHOV)
            acc[D + 1] += wg * (X0 * X0 + X1 * X1 + X2 * X2) + wsn * full_value<T>(a, h, 1, v);
    }
    block_sum<NFL>(acc, lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NFL; ++k) a.face_flux[size_t(f) * NFL + k] = acc[k];
        a.bnd[b] = P;
    }
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

// Mixed value of interior face f at velocity v seen from cell c, whose
// f-bar+ side is extrapolated with gc (c's own gradient) or the other cell's.
template <int D, class T>
__device__ __forceinline__ T interior_fbar(const KArgs &a, const T *__restrict__ fd, int f, int2 oc, int c, double sn,
                                           int v, const T xh[3], T fc, const T gc[3]) {
    const int nvp = a.s.nvp;
    const int side_c = (oc.x == c) ? 0 : 1;
    const int other = side_c == 0 ? oc.y : oc.x;
    if (sn >= VSMALL || sn < -VSMALL) {
        const int up_side = (sn >= VSMALL) ? 0 : 1;
        if (up_side == side_c) return a.s.grad ? fc + extrap<D, T>(a, f, side_c, xh, gc) : fc;
        const T fo = fd[size_t(other) * nvp + v];
        if (!a.s.grad) return fo;
        T go[3];
        lsq_grad<D, T>(a, fd, other, v, fo, go);
        return fo + extrap<D, T>(a, f, up_side, xh, go);
    }
    const T fo = fd[size_t(other) * nvp + v];
    const T f_own = side_c == 0 ? fc : fo, f_nei = side_c == 0 ? fo : fc;
    if (!a.s.grad) return T(0.5) * (f_own + f_nei);
    T go[3];
    lsq_grad<D, T>(a, fd, other, v, fo, go);
    const T X_own = side_c == 0 ? extrap<D, T>(a, f, 0, xh, gc) : extrap<D, T>(a, f, 0, xh, go);
    const T X_nei = side_c == 0 ? extrap<D, T>(a, f, 1, xh, go) : extrap<D, T>(a, f, 1, xh, gc);
    return T(0.5) * (f_own + f_nei + X_own + X_nei);
}

// (4) cell flux + update per (cell, velocity tile)
template <int MODEL, int D, class T, int LP>
__global__ __launch_bounds__(BLOCK) void k_cell_update(KArgs a) {
    if (*a.halt) return;
    constexpr int IPB = BLOCK / LP;
    const int nc = a.s.nc;
    const int nblk = (nc + IPB - 1) / IPB;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int tile = b / nblk, cb = b - tile * nblk;
    const int slot = threadIdx.x / LP, li = threadIdx.x % LP;
    const int c = cb * IPB + slot;
    const bool act = c < nc;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (act) {
        CellStep<MODEL, D> cs;
        const bool ok = cell_macros<MODEL, D, T>(a, c, cs);
        if (tile == 0 && li == 0) {
            if (MODEL != DUGKS_BGK) {
                a.rho_out[c] = cs.rho;
                for (int i = 0; i < 3; ++i) a.u_out[3 * c + i] = cs.u[i];
                if (MODEL == CDUGKS_SHAKHOV) a.T_out[c] = cs.T;
                for (int k = 0; k < 5; ++k) a.fm[5 * c + k] = 0.0;
                a.fm[5 * c] = cs.phi[0];
                for (int i = 0; i < D; ++i) a.fm[5 * c + 1 + i] = cs.phi[1 + i];
                if (MODEL == CDUGKS_SHAKHOV) a.fm[5 * c + 4] = cs.phi[D + 1];
            }
            if (!ok) {
                atomicAdd(a.bad_count, 1);
                atomicMin(a.first_bad, c);
                atomicMin(a.bad_step, a.step_index);
            }
        }
        const T *fg = (const T *)a.fin_g;
        const T *fh = (const T *)a.fin_h;
        T *og = (T *)a.fout_g;
        T *oh = (T *)a.fout_h;
        const T hd = T(a.s.half_dt);
        const T dtv = T(cs.dt_v);
        const size_t row = size_t(c) * a.s.nvp;
        const int j0 = a.cf_ptr[c], j1 = a.cf_ptr[c + 1];
        const int v0 = tile * a.s.vt_c, v1 = min(v0 + a.s.vt_c, a.s.nv);
        for (int v = v0 + li; v < v1; v += LP) {
            T x[3];
            load_xi<T>(a, v, x);
            T xh[3] = {x[0] * hd, x[1] * hd, x[2] * hd};
            const T fcg = fg[row + v];
            T gcg[3] = {T(0), T(0), T(0)};
            if (a.s.grad) lsq_grad<D, T>(a, fg, c, v, fcg, gcg);
            T fch = T(0), gch[3] = {T(0), T(0), T(0)};
            if (MODEL == CDUGKS_SHAKHOV) {
                fch = fh[row + v];
                if (a.s.grad) lsq_grad<D, T>(a, fh, c, v, fch, gch);
            }
            T Fg = T(0), Fh = T(0);
            for (int j = j0; j < j1; ++j) {
                const int f = a.cf_ids[j];
                const int2 oc = a.face_cells[f];
                const double sn = ndot(a, f, v);
                const EqParams &ef = a.face_eq[f];
                const T mix = T(ef.mix);
                T gv, hv = T(0);
                if (oc.x == oc.y) {
                    const int bi = a.face_bnd[f];
                    T gmx, hmx;
                    bnd_mix_pair<MODEL, D, T>(a, f, c, v, ef, gmx, hmx);
                    if (bi >= 0) {
                        const BndParams P = a.bnd[bi];
                        bnd_value<MODEL, D, T>(a, P, f, c, v, sn, ef, gmx, hmx, gv, hv);
                    } else {
                        gv = gmx;
                        hv = hmx;
                    }
                } else {
                    T gs, hs;
                    eq_store<MODEL, D, T>(a, ef, x, v, gs, hs);
                    gv = lerp_store<T>(interior_fbar<D, T>(a, fg, f, oc, c, sn, v, xh, fcg, gcg), gs, mix);
                    if (MODEL == CDUGKS_SHAKHOV)
                        hv = lerp_store<T>(interior_fbar<D, T>(a, fh, f, oc, c, sn, v, xh, fch, gch), hs, mix);
                }
                const T knA = T(sn) * T(a.cf_sA[j]);
                Fg += knA * gv;
                if (MODEL == CDUGKS_SHAKHOV) Fh += knA * hv;
            }
            if constexpr (sizeof(T) == 4) {
                // flux of f_ref itself: f_ref xi . sum(+-A n); non-zero where faces are not planar
                const float xs = x[0] * a.cell_S[3 * c] + x[1] * a.cell_S[3 * c + 1] + x[2] * a.cell_S[3 * c + 2];
                Fg += a.fref_f[v] * xs;
                if (MODEL == CDUGKS_SHAKHOV) Fh += a.fref_f[a.s.nv + v] * xs;
            }
            const double X0 = a.xi[3 * v], X1 = a.xi[3 * v + 1], X2 = a.xi[3 * v + 2];
            const double wv = a.w[v];
            if (MODEL == CDUGKS_SHAKHOV) {
                T gsn, hsn, gs, hs;
                eq_store<MODEL, D, T>(a, cs.en, x, v, gsn, hsn);
                eq_store<MODEL, D, T>(a, cs.e, x, v, gs, hs);
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
                const double c0 = X0 - cs.u[0], c1 = X1 - cs.u[1], c2 = X2 - cs.u[2];
                const double t = 0.5 * wv * ((c0 * c0 + c1 * c1 + c2 * c2) * full_value<T>(a, g, 0, v) +
                                             full_value<T>(a, h, 1, v));
                acc[0] += t * c0;
                acc[1] += t * c1;
                acc[2] += t * c2;
                og[row + v] = og_;
                oh[row + v] = oh_;
            } else if (MODEL == CDUGKS_BGK) {
                T feqn, feq, dummy;
                eq_store<MODEL, D, T>(a, cs.en, x, v, feqn, dummy);
                eq_store<MODEL, D, T>(a, cs.e, x, v, feq, dummy);
                const T cm = T(cs.cm), C = T(cs.C), ceq = T(cs.Cs);
                if constexpr (sizeof(T) == 8) {  // cdugks.cpp:325-332 as written
                    const T itau = T(1.0 / cs.tau);
                    const T fn = cm * feqn + (T(1) - cm) * fcg;
                    og[row + v] =
                        (T(1) - ceq) * (C * (fn + hd * (feq * itau + (feqn - fn) * itau) - dtv * Fg)) + ceq * feq;
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
                eq_store<MODEL, D, T>(a, cs.en, x, v, feq, dummy);
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
    }
    if (MODEL == CDUGKS_BGK) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] = seg_sum<LP>(acc[k]);
    if (act && li == 0) {
        double *dst = a.part_c + (size_t(c) * a.s.ntile_c + tile) * 4;
        for (int k = 0; k < 4; ++k) dst[k] = acc[k];
    }
}

// (5) q^{n+1} (Shakhov) or the DUGKS macros from the per-tile partials
template <int MODEL>
__global__ __launch_bounds__(BLOCK) void k_cell_finish(KArgs a) {
    if (*a.halt) return;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
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
            atomicMin(a.bad_step, a.step_index);
        }
    }
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
        T *og = (T *)a.fout_g + size_t(c) * a.s.nvp;
        T *oh = (T *)a.fout_h + size_t(c) * a.s.nvp;
        for (int v = li; v < a.s.nv; v += LP) {
            const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
            const double wv = a.w[v];
            double g, h = 0.0;
            if (MODEL == CDUGKS_SHAKHOV) {
                shakhov_eq<double>(e, X, D, a.s.K, g, h);
                oh[v] = T(h - (sizeof(T) == 4 ? a.fref_d[a.s.nv + v] : 0.0));
            } else {
                g = bgk_eq<double>(e, X);
            }
            og[v] = T(g - (sizeof(T) == 4 ? a.fref_d[v] : 0.0));
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
    T *og = (T *)a.fout_g + size_t(c) * a.s.nvp;
    T *oh = (T *)a.fout_h + size_t(c) * a.s.nvp;
    for (int v = li; v < a.s.nv; v += LP) {
        const double X[3] = {a.xi[3 * v], a.xi[3 * v + 1], a.xi[3 * v + 2]};
        const double rg = sizeof(T) == 4 ? a.fref_d[v] : 0.0;
        if (MODEL == CDUGKS_SHAKHOV) {
            double g, h;
            shakhov_eq<double>(e, X, D, a.s.K, g, h);
            og[v] = T(g - rg);
            oh[v] = T(h - (sizeof(T) == 4 ? a.fref_d[a.s.nv + v] : 0.0));
        } else {
            og[v] = T(bgk_eq<double>(e, X) - rg);
        }
    }
}

// FP64 velocity-major [nv][nc] <-> device cell-major [nc][nvp] (32x32 tiles through LDS)
template <class T, bool TO_DEVICE>
__global__ __launch_bounds__(BLOCK) void k_transpose(const double *vm_in, double *vm_out, const T *cm_in, T *cm_out,
                                                     int nc, int nv, int nvp, const double *fref) {
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
            if (c < nc && v < nv) cm_out[size_t(c) * nvp + v] = T(tile[tx][k] - (fref ? fref[v] : 0.0));
        }
    } else {
        for (int k = ty; k < 32; k += 8) {
            int c = c0 + k, v = v0 + tx;
            tile[tx][k] = (c < nc && v < nv) ? double(cm_in[size_t(c) * nvp + v]) + (fref ? fref[v] : 0.0) : 0.0;
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

__global__ void k_reset_flags(int *bad_count, int *first_bad, int *bad_step, int *halt) {
    *bad_count = 0;
    *first_bad = 0x7fffffff;
    *bad_step = 0x7fffffff;
    *halt = 0;
}

}  // namespace

// ====================================================================== host
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
    std::vector<meso_mark_desc> marks_h;
    std::vector<int> face_bnd_h, bnd_face_h;
    int n_marks = 0;
    // device pointers
    int2 *face_cells = nullptr;
    double *face_n = nullptr, *face_d = nullptr, *cf_sA = nullptr, *nb_a = nullptr, *cell_vol = nullptr;
    float *face_df = nullptr, *nb_af = nullptr, *xif = nullptr, *cell_S = nullptr;
    int *cf_ptr = nullptr, *cf_ids = nullptr, *nb_ptr = nullptr, *nb_ids = nullptr, *face_bnd = nullptr,
        *bnd_face = nullptr, *bnd_mark = nullptr, *mirror = nullptr;
    meso_mark_desc *marks = nullptr;
    double *xi = nullptr, *w = nullptr;
    void *f[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buffer][distribution]
    double *rho[2] = {nullptr, nullptr}, *u[2] = {nullptr, nullptr}, *T[2] = {nullptr, nullptr},
           *q[2] = {nullptr, nullptr};
    double *rho_res = nullptr, *u_res = nullptr, *T_res = nullptr, *q_res = nullptr;
    double *part_f = nullptr, *face_flux = nullptr, *part_c = nullptr, *fm = nullptr, *res_part = nullptr,
           *res_out = nullptr;
    EqParams *face_eq = nullptr;
    BndParams *bnd = nullptr;
    int *flags = nullptr;  // bad_count, first_bad, bad_step, halt
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
    c->bytes += int64_t(bytes);
    *p = reinterpret_cast<P *>(q);
    return MESO_OK;
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
    a.nb_af = c->nb_af;
    a.cell_vol = c->cell_vol;
    a.cell_S = c->cell_S;
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
    a.bad_count = c->flags;
    a.first_bad = c->flags + 1;
    a.bad_step = c->flags + 2;
    a.halt = c->flags + 3;
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
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int set_reference_from(meso_hip_ctx *c, const double *rho, const double *T) {
    if (c->storage != MESO_STORAGE_FP32) return MESO_OK;
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
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, c->stream);
        }
    }
    ~Launch() {
        if (c->profiling) {
            hipEventRecord(b, c->stream);
            c->recs.push_back({name, a, b});
        }
    }
};

#define LAUNCH(ctx, name, kern, grid, ...)                                                           \
    do {                                                                                             \
        Launch l_(ctx, name);                                                                        \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, (ctx)->stream, __VA_ARGS__);            \
    } while (0)

template <int MODEL, int D, class T, int LP>
int step_impl(meso_hip_ctx *c, KArgs a) {
    const StepConst &s = c->s;
    constexpr int IPB = BLOCK / LP;
    const int fblk = (s.nf + IPB - 1) / IPB, cblk = (s.nc + IPB - 1) / IPB;
    LAUNCH(c, "face_moments", (k_face_moments<MODEL == CDUGKS_SHAKHOV ? CDUGKS_SHAKHOV : CDUGKS_BGK, D, T, LP>),
           fblk * s.ntile_f, a);
    LAUNCH(c, "face_eq", (k_face_eq<MODEL == CDUGKS_SHAKHOV ? CDUGKS_SHAKHOV : CDUGKS_BGK, D, T, LP>), fblk, a);
    if (s.nb > 0)
        LAUNCH(c, "boundary", (k_boundary<MODEL == CDUGKS_SHAKHOV ? CDUGKS_SHAKHOV : CDUGKS_BGK, D, T>), s.nb, a);
    LAUNCH(c, "cell_update", (k_cell_update<MODEL, D, T, LP>), cblk * s.ntile_c, a);
    if (MODEL == CDUGKS_SHAKHOV) LAUNCH(c, "cell_finish", (k_cell_finish<MODEL>), (s.nc + BLOCK - 1) / BLOCK, a);
    if (MODEL == DUGKS_BGK) {
        LAUNCH(c, "cell_finish", (k_cell_finish<MODEL>), (s.nc + BLOCK - 1) / BLOCK, a);
        LAUNCH(c, "dugks_tobp", (k_dugks_tobp<D, T, LP>), cblk, a);
    }
    HIP_TRY(c, hipGetLastError());
    return MESO_OK;
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
struct StepF {
    static int run(meso_hip_ctx *c, KArgs a) { return step_impl<M, D, T, L>(c, a); }
};
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
    HIP_TRY(c, hipMemcpyAsync(c->flags_host, c->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return MESO_OK;
}

int reset_flags(meso_hip_ctx *c) {
    hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(1), 0, c->stream, c->flags, c->flags + 1, c->flags + 2,
                       c->flags + 3);
    HIP_TRY(c, hipGetLastError());
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
    if (storage != MESO_STORAGE_FP64 && storage != MESO_STORAGE_FP32)
        return set_err(nullptr, MESO_E_INVALID, "storage must be MESO_STORAGE_FP64 or MESO_STORAGE_FP32");
    if (m->n_cells <= 0 || m->n_faces <= 0 || d->n_vel <= 0) return set_err(nullptr, MESO_E_INVALID, "empty mesh or DVS");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(nullptr, MESO_E_DEVICE, "no HIP device available");
    auto *c = new meso_hip_ctx();
    auto fail = [&](int rc) {
        meso_hip_destroy(c);
        return rc;
    };
    hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(nullptr, MESO_E_DEVICE, "hipStreamCreate failed"));
    c->storage = storage;
    c->nd = p->model == MESO_MODEL_CDUGKS_SHAKHOV ? 2 : 1;
    set_const(c->s, *p);
    const int nc = m->n_cells, nf = m->n_faces, nv = d->n_vel, D = m->dim;
    c->s.nc = nc;
    c->s.nf = nf;
    c->s.nv = nv;
    c->s.nvp = nv >= 64 ? ((nv + 63) / 64) * 64 : nv;
    c->lp = lanes_for(nv);
    const int target = env_int("MESO_HIP_VTILE", 1024);
    tiles_for(nv, c->lp, target, c->s.ntile_f, c->s.vt_f);
    tiles_for(nv, c->lp, target, c->s.ntile_c, c->s.vt_c);
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
    std::vector<float> fdf(6 * size_t(nf)), naf(3 * size_t(m->n_nbr)), xif(3 * size_t(nv)), cS(3 * size_t(nc), 0.f);
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
        for (int r = 0; r < 3; ++r) cS[3 * size_t(cc) + r] = float(S3[r]);
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
    UP(nb_a, na.data(), na.size());
    UP(nb_af, naf.data(), naf.size());
    UP(cell_vol, m->cell_vol, nc);
    UP(cell_S, cS.data(), cS.size());
    UP(face_bnd, c->face_bnd_h.data(), nf);
    UP(bnd_face, c->bnd_face_h.data(), c->bnd_face_h.size());
    UP(bnd_mark, bmark.data(), bmark.size());
    UP(marks, m->marks, m->n_marks);
    UP(xi, d->xi, 3 * size_t(nv));
    UP(xif, xif.data(), xif.size());
    UP(w, d->w, nv);
    if (d->mirror) UP(mirror, d->mirror, 3 * size_t(nv));
#undef UP
    const size_t es = storage == MESO_STORAGE_FP32 ? 4 : 8;
    const size_t fsz = size_t(nc) * c->s.nvp;
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < c->nd; ++k) {
            char *q = nullptr;
            if ((rc = dalloc(c, &q, fsz * es))) return fail(rc);
            c->f[b][k] = q;
        }
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(c, &c->rho[b], nc)) || (rc = dalloc(c, &c->u[b], 3 * size_t(nc))) ||
            (rc = dalloc(c, &c->T[b], nc)) || (rc = dalloc(c, &c->q[b], 3 * size_t(nc))))
            return fail(rc);
    }
    const int ncol = c->nd == 2 ? (D == 2 ? 11 : 17) : (D == 2 ? 6 : 10);
    if ((rc = dalloc(c, &c->rho_res, nc)) || (rc = dalloc(c, &c->u_res, 3 * size_t(nc))) ||
        (rc = dalloc(c, &c->T_res, nc)) || (rc = dalloc(c, &c->q_res, 3 * size_t(nc))) ||
        (rc = dalloc(c, &c->part_f, size_t(nf) * c->s.ntile_f * ncol)) ||
        (rc = dalloc(c, &c->face_flux, size_t(nf) * (D + 2))) ||
        (rc = dalloc(c, &c->part_c, size_t(nc) * c->s.ntile_c * 4)) || (rc = dalloc(c, &c->fm, 5 * size_t(nc))) ||
        (rc = dalloc(c, &c->res_part, 8 * size_t((nc + BLOCK - 1) / BLOCK))) || (rc = dalloc(c, &c->res_out, 8)) ||
        (rc = dalloc(c, &c->face_eq, nf)) || (rc = dalloc(c, &c->bnd, std::max(1, c->s.nb))) ||
        (rc = dalloc(c, &c->flags, 4)) || (rc = dalloc(c, &c->fref_d, 2 * size_t(nv))) ||
        (rc = dalloc(c, &c->fref_f, 2 * size_t(nv))))
        return fail(rc);
    c->xi_h.assign(d->xi, d->xi + 3 * size_t(nv));
    c->w_h.assign(d->w, d->w + nv);
    if (hipHostMalloc((void **)&c->flags_host, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return fail(set_err(nullptr, MESO_E_DEVICE, "hipHostMalloc failed"));
    if ((rc = reset_flags(c))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_err(nullptr, MESO_E_DEVICE, "upload failed"));
    *out = c;
    return MESO_OK;
}

void meso_hip_destroy(meso_hip_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &r : c->recs) {
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    for (void *p : c->allocs) hipFree(p);
    if (c->flags_host) hipHostFree(c->flags_host);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int meso_hip_set_params(meso_hip_ctx *c, const meso_step_params *p) {
    if (!c || !p) return set_err(c, MESO_E_INVALID, "null argument");
    if (p->model != c->s.model || p->dim != c->s.D) return set_err(c, MESO_E_INVALID, "model/dimension cannot change");
    set_const(c->s, *p);
    return MESO_OK;
}

static int init_common(meso_hip_ctx *c, const double *rho, const double *u3, const double *T, const double *q3,
                       bool from_macros) {
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
        if ((rc = dalloc(c, &tr, nc)) || (rc = dalloc(c, &tu, 3 * nc)) || (rc = dalloc(c, &tT, nc))) return rc;
        HIP_TRY(c, hipMemcpyAsync(tr, rho, nc * 8, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(tu, u3, 3 * nc * 8, hipMemcpyHostToDevice, c->stream));
        if (T) HIP_TRY(c, hipMemcpyAsync(tT, T, nc * 8, hipMemcpyHostToDevice, c->stream));
        rc = dispatch<InitF>(c, a, (const double *)tr, (const double *)tu, (const double *)tT, false);
    }
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // temporaries are released right away
    for (double *t : {tr, tu, tT}) {
        if (!t) continue;
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void *)t);
        if (it != c->allocs.end()) c->allocs.erase(it);
        hipFree(t);
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
        hipFree(stage);
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
    const bool f32 = c->storage == MESO_STORAGE_FP32;
    const double *fref = f32 ? c->fref_d + size_t(which) * nv : nullptr;
    if (to_device) {
        HIP_TRY(c, hipMemcpyAsync(stage, src, n * 8, hipMemcpyHostToDevice, c->stream));
        if (f32)
            hipLaunchKernelGGL((k_transpose<float, true>), grid, dim3(BLOCK), 0, c->stream, stage, nullptr, nullptr,
                               (float *)dev, nc, nv, nvp, fref);
        else
            hipLaunchKernelGGL((k_transpose<double, true>), grid, dim3(BLOCK), 0, c->stream, stage, nullptr, nullptr,
                               (double *)dev, nc, nv, nvp, fref);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    } else {
        if (f32)
            hipLaunchKernelGGL((k_transpose<float, false>), grid, dim3(BLOCK), 0, c->stream, nullptr, stage,
                               (const float *)dev, nullptr, nc, nv, nvp, fref);
        else
            hipLaunchKernelGGL((k_transpose<double, false>), grid, dim3(BLOCK), 0, c->stream, nullptr, stage,
                               (const double *)dev, nullptr, nc, nv, nvp, fref);
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

int meso_hip_step(meso_hip_ctx *c, int32_t nsteps, meso_step_report *rep) {
    if (!c || nsteps < 0) return set_err(c, MESO_E_INVALID, "bad argument");
    int rc;
    for (int i = 0; i < nsteps; ++i) {
        KArgs a = make_args(c);
        a.step_index = c->steps_done + i;
        if ((rc = dispatch<StepF>(c, a))) return rc;
        c->cur = 1 - c->cur;
    }
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
        rep->reserved = 0;
    }
    return bad > 0 ? set_err(c, MESO_E_NUMERIC, "NaN density or T<0 in a cell") : MESO_OK;
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
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
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
        hipEventElapsedTime(&ms, r.a, r.b);
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
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
