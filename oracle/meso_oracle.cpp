// FP64 CPU restatement of meso's DUGKS/CDUGKS time step — the oracle.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker. The product never links it.
//
// Each routine restates one block of the reference, cited file:line below.
// Loop order inside every reduction over velocities is ascending p, the order
// meso uses with one MPI rank, so sums round the same way up to the
// association of the per-term products. Outer loops over velocities, faces or
// cells run under OpenMP; they carry no cross-iteration dependence.
//
// Pinning (how each part is checked, see tests/):
//   grad (A2)              closed forms: exact on linear fields; numpy lstsq brute force
//   reconstruct (A3,A4)    linear-f closed form; rest-state fixed point; golden vectors
//                          from the compiled reference (tests/golden/)
//   boundaries (A5)        wall: golden vectors + mass conservation; farfield and
//                          pressure: golden vectors; symmetry / slip-wall: parity
//                          unpinned against the reference (its GatherFieldList reads
//                          displacements[-1], field.cpp:203-205) — pinned by mirror
//                          invariance only
//   fvm_update (A6,A7)     golden vectors; mass conservation; rest state
//   residual (A8)          hand example
#include "meso_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double VSMALL = 1e-30;  // include/core/math.h:4

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit V3(const double *p) : x(p[0]), y(p[1]), z(p[2]) {}
    V3 operator+(const V3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3 operator/(double s) const {
        double r = 1.0 / s;  // vector.cpp:79-82 divides through the reciprocal
        return {x * r, y * r, z * r};
    }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    double operator*(const V3 &o) const { return x * o.x + y * o.y + z * o.z; }
    double mag() const { return std::sqrt(x * x + y * y + z * z); }
};
inline V3 operator*(double k, const V3 &v) { return {v.x * k, v.y * k, v.z * k}; }

thread_local std::string g_err;
int g_threads = 0;

struct Mark {
    meso_mark_desc d;
};

}  // namespace

struct oracle_ctx {
    // mesh
    int D = 2, NC = 0, NF = 0, NV = 0;
    double total_volume = 0;
    std::vector<V3> cpos, fpos, fn0;
    std::vector<double> cvol, farea;
    std::vector<int> own, nei;
    std::vector<int> cf_ptr, cf_ids, nb_ptr, nb_ids;
    std::vector<V3> dr;
    std::vector<double> lw;
    std::vector<V3> Cx, Cy, Cz;
    std::vector<int> fmark;
    std::vector<meso_mark_desc> marks;
    // dvs
    std::vector<V3> xi;
    std::vector<double> w;
    std::vector<int> mirror;
    // params
    meso_step_params P{};
    double Cv = 0, gamma = 0, half_dt = 0;
    // state (velocity-major)
    int nd = 1;
    std::vector<double> g, h, gface, hface, flux_g, flux_h;
    std::vector<double> rho, T;
    std::vector<V3> u, q;
    std::vector<double> fm0, fm2;
    std::vector<V3> fm1;
    std::vector<double> rho_res, T_res;
    std::vector<int> mask;  // 1: the cell's group sets `mask switch True`
    int masked_cells = 0;   // repaired cells, accumulated over oracle_step calls
    std::vector<V3> u_res, q_res;
    int bad_cells = 0, first_bad = -1;

    double &G(int p, int c) { return g[size_t(p) * NC + c]; }
    double &H(int p, int c) { return h[size_t(p) * NC + c]; }
    double &GF(int p, int f) { return gface[size_t(p) * NF + f]; }
    double &HF(int p, int f) { return hface[size_t(p) * NF + f]; }

    bool shakhov() const { return P.model == MESO_MODEL_CDUGKS_SHAKHOV; }

    void set_params(const meso_step_params &p) {
        P = p;
        D = p.dim;
        gamma = (p.K + 5.0) / (p.K + 3.0);  // cdugks_shakhov.cpp:30-32
        Cv = (p.K + 3.0) * p.R * 0.5;
        half_dt = p.dt * 0.5;
    }

    // ---------------- equilibria (cdugks_shakhov.cpp:181-211, cdugks.cpp:116-121)
    double tau_f(double r, double t) const { return P.mu0 * std::pow(t / P.T0, P.omega) / (r * P.R * t); }
    double g_maxwell(double r, double t, double cc) const {
        if (D == 2) {
            double over_RT2 = 0.5 / (P.R * t);
            return r * over_RT2 / M_PI * std::exp(-cc * over_RT2);
        }
        double RT = P.R * t;
        return r / std::pow(2.0 * M_PI * RT, 1.5) * std::exp(-cc * 0.5 / RT);
    }
    double h_maxwell(double t, double gm) const { return (P.K + 3.0 - D) * (P.R * t) * gm; }
    double g_shakhov(double r, double t, double cc, double cq, double gm) const {
        double RT = P.R * t, p = r * RT;
        return (1.0 - P.Pr) * (cq / (5.0 * p * RT)) * (cc / RT - D - 2.0) * gm + gm;
    }
    double h_shakhov(double r, double t, double cc, double cq, double gm) const {
        double RT = P.R * t, p = r * RT;
        double hm = h_maxwell(t, gm);
        return (1.0 - P.Pr) * (cq / (5.0 * p)) * ((cc / RT - D) * (P.K + 3.0 - D) - 2.0 * P.K) * gm + hm;
    }
    double f_eq(double r, const V3 &uu, const V3 &x) const {
        double over_RT = 1.0 / P.RT;
        double ku = (uu * x) * over_RT;
        double uu2 = (uu * uu) * over_RT;
        return r * (1.0 + ku + (ku * ku - uu2) * 0.5);
    }

    // ---------------- LSQ gradient (mesh_numerical.cpp:24-49)
    void grad(const double *f, std::vector<V3> &out) const {
        out.assign(NC, V3());
        for (int c = 0; c < NC; ++c) {
            V3 S;
            for (int j = nb_ptr[c]; j < nb_ptr[c + 1]; ++j) S += lw[j] * (f[nb_ids[j]] - f[c]) * dr[j];
            out[c] = V3(Cx[c] * S, Cy[c] * S, Cz[c] * S);
        }
    }

    // Venkatakrishnan limiter of cell c for the increment dw = (x_f - x_c) . grad f
    // (Solver::venkata_limiter, src/solver/limiter.cpp:7-31): min/max over the
    // cell's face neighbours, omega = (k V^(1/D))^3.
    double venkata(const double *fc, double dw, int c, double eps) const {
        if (std::fabs(dw) <= eps) return 1.0;
        const double kh = P.venkata_k * std::pow(cvol[c], 1.0 / double(D));
        const double omega = kh * kh * kh;
        const double wc = fc[c];
        double wmax = wc, wmin = wc;
        for (int j = nb_ptr[c]; j < nb_ptr[c + 1]; ++j) {
            const double wn = fc[nb_ids[j]];
            if (wn > wmax) wmax = wn;
            if (wn < wmin) wmin = wn;
        }
        const double a = dw > 0.0 ? wmax - wc : wmin - wc;
        const double aa = a * a, ab = a * dw, bb = dw * dw;
        return (aa + 2.0 * ab + omega) / (aa + 2.0 * bb + ab + omega);
    }

    // ---------------- face reconstruction (cdugks_shakhov.cpp:213-299, cdugks.cpp:124-153)
    // limit: the Shakhov solver's limiter-switch (cdugks_shakhov.cpp:231-266;
    // eps 1e-10 on the owner-upwind branch, VSMALL on the others)
    void reconstruct_faces(const std::vector<double> &cellv, std::vector<double> &facev, bool limit = false) {
#pragma omp parallel
        {
            std::vector<V3> gr;
#pragma omp for schedule(static)
            for (int p = 0; p < NV; ++p) {
                const double *fc = cellv.data() + size_t(p) * NC;
                double *ff = facev.data() + size_t(p) * NF;
                const V3 &x = xi[p];
                if (P.gradient_switch) {
                    grad(fc, gr);
                    for (int f = 0; f < NF; ++f) {
                        int o = own[f], n = nei[f];
                        double nx = fn0[f] * x;
                        if (limit) {
                            const V3 dro = fpos[f] - cpos[o], drn = fpos[f] - cpos[n];
                            if (nx >= VSMALL) {
                                const double phi = venkata(fc, dro * gr[o], o, 1.0e-10);
                                ff[f] = fc[o] + phi * ((dro - x * half_dt) * gr[o]);
                            } else if (nx < -VSMALL) {
                                const double phi = venkata(fc, drn * gr[n], n, VSMALL);
                                ff[f] = fc[n] + phi * ((drn - x * half_dt) * gr[n]);
                            } else {
                                const double po = venkata(fc, dro * gr[o], o, VSMALL);
                                const double pn = venkata(fc, drn * gr[n], n, VSMALL);
                                ff[f] = 0.5 * (fc[o] + fc[n] + po * ((dro - x * half_dt) * gr[o]) +
                                               pn * ((drn - x * half_dt) * gr[n]));
                            }
                        } else if (nx >= VSMALL) {
                            ff[f] = fc[o] + ((fpos[f] - cpos[o]) - x * half_dt) * gr[o];
                        } else if (nx < -VSMALL) {
                            ff[f] = fc[n] + ((fpos[f] - cpos[n]) - x * half_dt) * gr[n];
                        } else {
                            ff[f] = 0.5 * (fc[o] + fc[n] + ((fpos[f] - cpos[o]) - x * half_dt) * gr[o] +
                                           ((fpos[f] - cpos[n]) - x * half_dt) * gr[n]);
                        }
                    }
                } else {
                    for (int f = 0; f < NF; ++f) {
                        int o = own[f], n = nei[f];
                        double nx = fn0[f] * x;
                        if (nx >= VSMALL) ff[f] = fc[o];
                        else if (nx < -VSMALL) ff[f] = fc[n];
                        else ff[f] = 0.5 * (fc[o] + fc[n]);
                    }
                }
            }
        }
    }

    static double solve_pressure(double P0, double R, double Tm, const V3 &U, double gam) {
        // Compressible::solve_pressure, pressure_boundary.cpp:15-18
        double Ma = U.mag() / std::sqrt(gam * R * Tm);
        return P0 / std::pow(1.0 + (gam - 1.0) / 2.0 * (Ma * Ma), gam / (gam - 1.0));
    }

    // ---------------- Shakhov reconstruct (cdugks_shakhov.cpp:213-608)
    void reconstruct_shakhov() {
        reconstruct_faces(g, gface, P.limiter_switch != 0);
        reconstruct_faces(h, hface, P.limiter_switch != 0);
        // face macros + mixing (:300-363)
#pragma omp parallel for schedule(static)
        for (int f = 0; f < NF; ++f) {
            double m0 = 0, m2 = 0;
            V3 m1, m3;
            for (int p = 0; p < NV; ++p) {
                double kk = xi[p] * xi[p];
                double gg = GF(p, f), hh = HF(p, f);
                m0 += w[p] * gg;
                m1 += w[p] * gg * xi[p];
                m2 += w[p] * (kk * gg + hh);
            }
            V3 uu = m1 / m0;
            for (int p = 0; p < NV; ++p) {
                V3 c = xi[p] - uu;
                double cc = c * c;
                m3 += w[p] * c * (cc * GF(p, f) + HF(p, f));
            }
            double r = m0;
            double rhoE = 0.5 * m2;
            double t = (rhoE / r - 0.5 * (uu * uu)) / Cv;
            double tau = tau_f(r, t);
            V3 qf = (tau / (2.0 * tau + half_dt * P.Pr)) * m3;
            double Cs = half_dt / (2.0 * tau + half_dt);
            for (int p = 0; p < NV; ++p) {
                V3 c = xi[p] - uu;
                double cc = c * c, cq = c * qf;
                double gm = g_maxwell(r, t, cc);
                double gs = g_shakhov(r, t, cc, cq, gm);
                double hs = h_shakhov(r, t, cc, cq, gm);
                GF(p, f) = (1.0 - Cs) * GF(p, f) + Cs * gs;
                HF(p, f) = (1.0 - Cs) * HF(p, f) + Cs * hs;
            }
        }
        // boundaries (:364-607); every face's column is independent
#pragma omp parallel
        {
            std::vector<double> gall, hall;
#pragma omp for schedule(dynamic, 16)
            for (int f = 0; f < NF; ++f) {
                const meso_mark_desc &mk = marks[fmark[f]];
                if (mk.type == MESO_BC_FLUID_INTERIOR) continue;
                const V3 n1 = -fn0[f];
                const int cn = own[f];
                const V3 uw(mk.velocity);
                switch (mk.type) {
                    case MESO_BC_PRESSURE_INLET: {  // :374-394
                        double Tm = T[cn];
                        V3 um = u[cn];
                        V3 uin = (um * n1) * n1;
                        double Tin = mk.temperature;
                        double pin = solve_pressure(mk.pressure, P.R, Tm, uin, gamma);
                        double rin = pin / (P.R * Tin);
                        for (int p = 0; p < NV; ++p)
                            if (xi[p] * n1 >= VSMALL) {
                                V3 c = xi[p] - uin;
                                double gm = g_maxwell(rin, Tin, c * c);
                                GF(p, f) = gm;
                                HF(p, f) = h_maxwell(Tin, gm);
                            }
                    } break;
                    case MESO_BC_FARFIELD_INLET: {  // :395-411
                        for (int p = 0; p < NV; ++p)
                            if (xi[p] * n1 >= VSMALL) {
                                V3 c = xi[p] - uw;
                                double gm = g_maxwell(mk.density, mk.temperature, c * c);
                                GF(p, f) = gm;
                                HF(p, f) = h_maxwell(mk.temperature, gm);
                            }
                    } break;
                    case MESO_BC_PRESSURE_OUTLET: {  // :430-452
                        V3 um = u[cn];
                        double Tm = T[cn];
                        double pe = solve_pressure(mk.pressure, P.R, Tm, um, gamma);
                        double re = pe / (P.R * Tm);
                        V3 ne = -n1;
                        V3 ue = (um * ne) * ne;
                        double Te = mk.temperature;
                        for (int p = 0; p < NV; ++p)
                            if (xi[p] * n1 >= VSMALL) {
                                V3 c = xi[p] - ue;
                                double gm = g_maxwell(re, Te, c * c);
                                GF(p, f) = gm;
                                HF(p, f) = h_maxwell(Te, gm);
                            }
                    } break;
                    case MESO_BC_FARFIELD_OUTLET: {  // :453-467
                        for (int p = 0; p < NV; ++p)
                            if (xi[p] * n1 >= VSMALL) {
                                V3 c = xi[p] - u[cn];
                                double gm = g_maxwell(rho[cn], T[cn], c * c);
                                GF(p, f) = gm;
                                HF(p, f) = h_maxwell(T[cn], gm);
                            }
                    } break;
                    case MESO_BC_WALL:
                    case MESO_BC_SLIP_WALL: {  // :468-497 then :528-602
                        double Tw = 0.0;
                        if (mk.T_kind == MESO_PATCH_ZERO_GRADIENT) Tw = T[cn];
                        else if (mk.T_kind == MESO_PATCH_FIXED_VALUE) Tw = mk.temperature;
                        double rw = 0.0, rw0 = 0.0;
                        for (int p = 0; p < NV; ++p) {
                            double kn = xi[p] * n1;
                            if (kn >= VSMALL) {
                                V3 c = xi[p] - uw;
                                rw0 += kn * w[p] * g_maxwell(1.0, Tw, c * c);
                            } else {
                                rw -= kn * w[p] * GF(p, f);
                            }
                        }
                        double rho_w = rw / rw0;
                        double Tp = mk.temperature;
                        if (mk.type == MESO_BC_WALL) {
                            for (int p = 0; p < NV; ++p) {
                                double kn = xi[p] * n1;
                                if (kn < -VSMALL) continue;
                                V3 c = xi[p] - uw;
                                double gm = g_maxwell(rho_w, Tp, c * c);
                                double hm = h_maxwell(Tp, gm);
                                if (kn >= VSMALL) {
                                    GF(p, f) = gm;
                                    HF(p, f) = hm;
                                } else {
                                    GF(p, f) = 0.5 * (gm + GF(p, f));
                                    HF(p, f) = 0.5 * (hm + HF(p, f));
                                }
                            }
                        } else {
                            gall.resize(NV);
                            hall.resize(NV);
                            for (int p = 0; p < NV; ++p) {
                                gall[p] = GF(p, f);
                                hall[p] = HF(p, f);
                            }
                            for (int p = 0; p < NV; ++p) {
                                double kn = xi[p] * n1;
                                if (kn < -VSMALL) continue;
                                V3 c = xi[p] - uw;
                                double gm = g_maxwell(rho_w, Tp, c * c);
                                double hm = h_maxwell(Tp, gm);
                                int pm = mirror[3 * p + mk.direction];
                                double gv = (1.0 - mk.alpha_u) * gm + mk.alpha_u * gall[pm];
                                double hv = (1.0 - mk.alpha_T) * hm + mk.alpha_T * hall[pm];
                                if (kn >= VSMALL) {
                                    GF(p, f) = gv;
                                    HF(p, f) = hv;
                                } else {
                                    GF(p, f) = 0.5 * (gv + GF(p, f));
                                    HF(p, f) = 0.5 * (hv + HF(p, f));
                                }
                            }
                        }
                    } break;
                    case MESO_BC_SYMMETRY: {  // :498-517
                        gall.resize(NV);
                        hall.resize(NV);
                        for (int p = 0; p < NV; ++p) {
                            gall[p] = GF(p, f);
                            hall[p] = HF(p, f);
                        }
                        for (int p = 0; p < NV; ++p) {
                            double kn = xi[p] * n1;
                            int pm = mirror[3 * p + mk.direction];
                            if (kn >= VSMALL) {
                                GF(p, f) = gall[pm];
                                HF(p, f) = hall[pm];
                            } else if (kn >= -VSMALL) {
                                GF(p, f) = 0.5 * (GF(p, f) + gall[pm]);
                                HF(p, f) = 0.5 * (HF(p, f) + hall[pm]);
                            }
                        }
                    } break;
                    default: break;
                }
            }
        }
    }

    // ---------------- mask repair (CDUGKS_SHAKHOV::cell_value_interp,
    // cdugks_shakhov.cpp:766-789, called from the monitor at :698-719, with
    // fvmMesh::interp_IDW, src/mesh/mesh_numerical.cpp:51-94). meso meets the
    // cells in ascending order inside its update loop, so a neighbour before
    // the cell has its new (or repaired) macros and one after it still has the
    // old ones; q_cell holds q^n until the loop ends (repaired cells set theirs
    // to the average), and after the loop every q is the heat-flux sum m3, the
    // repaired cells' included. Rows are the oracle's velocity-major arrays.
    void mask_repair(int c, const std::vector<double> &rho_o, const std::vector<V3> &u_o,
                     const std::vector<double> &T_o, std::vector<V3> &q_loop) {
        double sw = 0.0, sr = 0.0, sT = 0.0;
        V3 su, sq;
        for (int k = nb_ptr[c]; k < nb_ptr[c + 1]; ++k) {
            const int n = nb_ids[k];
            const double wi = 1.0 / (dr[k] * dr[k]);
            const bool upd = n < c;
            sw += wi;
            sr += wi * (upd ? rho[n] : rho_o[n]);
            sT += wi * (upd ? T[n] : T_o[n]);
            su += wi * (upd ? u[n] : u_o[n]);
            sq += wi * q_loop[n];
        }
        const double r = sr / sw, t = sT / sw;
        const V3 uu = su / sw, qq = sq / sw;
        rho[c] = r;
        u[c] = uu;
        T[c] = t;
        q_loop[c] = qq;
        for (int p = 0; p < NV; ++p) {
            const V3 cv = xi[p] - uu;
            const double cc = cv * cv, cq = cv * qq;
            const double gm = g_maxwell(r, t, cc);
            G(p, c) = g_shakhov(r, t, cc, cq, gm);
            H(p, c) = h_shakhov(r, t, cc, cq, gm);
        }
    }

    // ---------------- Shakhov fvm_update (cdugks_shakhov.cpp:610-722)
    void fvm_update_shakhov() {
        const bool any_mask = std::find(mask.begin(), mask.end(), 1) != mask.end();
        std::vector<double> rho_o, T_o;
        std::vector<V3> u_o, q_loop;
        if (any_mask) {
            rho_o = rho;
            T_o = T;
            u_o = u;
            q_loop = q;
        }
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NC; ++c) {
            double m0 = 0, m2 = 0;
            V3 m1;
            for (int p = 0; p < NV; ++p) {
                double fg = 0.0, fh = 0.0;
                for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
                    int f = cf_ids[k];
                    V3 nv = (own[f] == c) ? fn0[f] : -fn0[f];
                    double knA = (xi[p] * nv) * farea[f];
                    fg += knA * GF(p, f);
                    fh += knA * HF(p, f);
                }
                flux_g[size_t(p) * NC + c] = fg;
                flux_h[size_t(p) * NC + c] = fh;
                m0 += w[p] * fg;
                m1 += w[p] * fg * xi[p];
                m2 += w[p] * ((xi[p] * xi[p]) * fg + fh);
            }
            fm0[c] = m0;
            fm1[c] = m1;
            fm2[c] = m2;
        }
        int bad = 0, first = -1;
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int c = 0; c < NC; ++c) {
            double dt_v = P.dt / cvol[c];
            double rn = rho[c], Tn = T[c];
            V3 un = u[c], qn = q[c];
            V3 rUn = rn * un;
            double rEn = rn * (0.5 * (un * un) + Cv * Tn);
            double taun = tau_f(rn, Tn);
            double r = rn - dt_v * fm0[c];
            V3 rU = rUn - dt_v * fm1[c];
            double rE = rEn - dt_v * 0.5 * fm2[c];
            V3 uu = rU / r;
            double t = (rE / r - (uu * uu) * 0.5) / Cv;
            double tau = tau_f(r, t);
            rho[c] = r;
            u[c] = uu;
            T[c] = t;
            double cm = half_dt / (half_dt - 2.0 * taun);
            double C = tau / (tau + half_dt);
            double Cs = half_dt / (2.0 * tau);
            V3 m3;
            for (int p = 0; p < NV; ++p) {
                V3 cn = xi[p] - un;
                double ccn = cn * cn, cqn = cn * qn;
                double gmn = g_maxwell(rn, Tn, ccn);
                double gsn = g_shakhov(rn, Tn, ccn, cqn, gmn);
                double hsn = h_shakhov(rn, Tn, ccn, cqn, gmn);
                double gn = cm * gsn + (1.0 - cm) * G(p, c);
                double hn = cm * hsn + (1.0 - cm) * H(p, c);
                V3 cc_ = xi[p] - uu;
                double cc = cc_ * cc_, cq = cc_ * qn;
                double gm = g_maxwell(r, t, cc);
                double gs = g_shakhov(r, t, cc, cq, gm);
                double hs = h_shakhov(r, t, cc, cq, gm);
                double gg = C * (gn + half_dt * (gs / tau + (gsn - gn) / taun) - dt_v * flux_g[size_t(p) * NC + c]);
                double hh = C * (hn + half_dt * (hs / tau + (hsn - hn) / taun) - dt_v * flux_h[size_t(p) * NC + c]);
                m3 += 0.5 * w[p] * cc_ * (cc * gg + hh);
                G(p, c) = (1.0 - Cs) * gg + Cs * gs;
                H(p, c) = (1.0 - Cs) * hh + Cs * hs;
            }
            q[c] = m3;
            if (std::isnan(r) || t < 0.0 || std::isnan(t)) ++bad;
        }
        if (bad) {
            bad = 0;
            for (int c = 0; c < NC; ++c) {
                if (!(std::isnan(rho[c]) || T[c] < 0.0 || std::isnan(T[c]))) continue;
                if (any_mask && mask[c]) {
                    mask_repair(c, rho_o, u_o, T_o, q_loop);
                    ++masked_cells;
                    continue;
                }
                if (first < 0) first = c;
                ++bad;
            }
        }
        bad_cells = bad;
        first_bad = first;
    }

    // ---------------- BGK reconstruct (cdugks.cpp:123-278, dugks.cpp:124-279)
    void reconstruct_bgk() {
        reconstruct_faces(g, gface);
        double c_eq = half_dt / (2.0 * P.tau + half_dt);
#pragma omp parallel for schedule(static)
        for (int f = 0; f < NF; ++f) {
            double m0 = 0;
            V3 m1;
            for (int p = 0; p < NV; ++p) {
                double ff = GF(p, f);
                m0 += w[p] * ff;
                m1 += w[p] * ff * xi[p];
            }
            V3 uu = m1 / m0;
            for (int p = 0; p < NV; ++p) GF(p, f) = (1.0 - c_eq) * GF(p, f) + c_eq * f_eq(m0, uu, xi[p]);
        }
#pragma omp parallel for schedule(dynamic, 16)
        for (int f = 0; f < NF; ++f) {
            const meso_mark_desc &mk = marks[fmark[f]];
            const V3 n1 = -fn0[f];
            const int cn = own[f];
            const V3 uw(mk.velocity);
            switch (mk.type) {
                case MESO_BC_FARFIELD_INLET:
                    for (int p = 0; p < NV; ++p)
                        if (xi[p] * n1 >= VSMALL) GF(p, f) = f_eq(mk.density, uw, xi[p]);
                    break;
                case MESO_BC_FARFIELD_OUTLET:
                    for (int p = 0; p < NV; ++p)
                        if (xi[p] * n1 >= VSMALL) GF(p, f) = f_eq(rho[cn], u[cn], xi[p]);
                    break;
                case MESO_BC_WALL: {
                    double rw = 0, rw0 = 0;
                    for (int p = 0; p < NV; ++p) {
                        double kn = xi[p] * n1;
                        if (kn >= VSMALL) rw0 += kn * w[p] * f_eq(1.0, uw, xi[p]);
                        else rw -= kn * w[p] * GF(p, f);
                    }
                    double rho_w = rw / rw0;
                    for (int p = 0; p < NV; ++p) {
                        double kn = xi[p] * n1;
                        if (kn >= VSMALL) GF(p, f) = f_eq(rho_w, uw, xi[p]);
                        else if (kn >= -VSMALL) GF(p, f) = 0.5 * (f_eq(rho_w, uw, xi[p]) + GF(p, f));
                    }
                } break;
                default: break;
            }
        }
    }

    void flux_bgk() {
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NC; ++c) {
            double m0 = 0;
            V3 m1;
            for (int p = 0; p < NV; ++p) {
                double fl = 0.0;
                for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
                    int f = cf_ids[k];
                    V3 nv = (own[f] == c) ? fn0[f] : -fn0[f];
                    fl += (xi[p] * nv) * farea[f] * GF(p, f);
                }
                flux_g[size_t(p) * NC + c] = fl;
                m0 += w[p] * fl;
                m1 += w[p] * fl * xi[p];
            }
            fm0[c] = m0;
            fm1[c] = m1;
            fm2[c] = 0.0;
        }
    }

    // ---------------- CDUGKS BGK update (cdugks.cpp:281-335)
    void fvm_update_cdugks() {
        flux_bgk();
        double tau = P.tau;
        double cm = half_dt / (half_dt - 2.0 * tau);
        double C = tau / (tau + half_dt);
        double c_eq = half_dt / (2.0 * tau);
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int c = 0; c < NC; ++c) {
            double dt_v = P.dt / cvol[c];
            double rn = rho[c];
            V3 un = u[c];
            double r = rn - dt_v * fm0[c];
            V3 uu = (rn * un - dt_v * fm1[c]) / rn;
            rho[c] = r;
            u[c] = uu;
            for (int p = 0; p < NV; ++p) {
                double feqn = f_eq(rn, un, xi[p]);
                double fn = cm * feqn + (1.0 - cm) * G(p, c);
                double feq = f_eq(r, uu, xi[p]);
                G(p, c) = (1.0 - c_eq) * (C * (fn + half_dt * (feq / tau + (feqn - fn) / tau) -
                                                dt_v * flux_g[size_t(p) * NC + c])) +
                          c_eq * feq;
            }
            if (std::isnan(r)) ++bad;
        }
        bad_cells = bad;
        first_bad = -1;
        if (bad)
            for (int c = 0; c < NC; ++c)
                if (std::isnan(rho[c])) { first_bad = c; break; }
    }

    // ---------------- classic DUGKS update (dugks.cpp:282-344)
    void fvm_update_dugks() {
        flux_bgk();
        double tau = P.tau, dt = P.dt;
        double C_bp_to_t = (dt + half_dt) / (half_dt - 2.0 * tau);
        double C_t_to_bp = (dt + half_dt) / (2.0 * tau + dt);
        int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int c = 0; c < NC; ++c) {
            double dt_v = dt / cvol[c];
            double r = rho[c];
            V3 uu = u[c];
            for (int p = 0; p < NV; ++p) {
                double feq = f_eq(r, uu, xi[p]);
                double fbp = G(p, c);
                double ft = (1.0 - C_bp_to_t) * fbp + C_bp_to_t * feq;
                G(p, c) = (4.0 * fbp - ft) / 3.0 - dt_v * flux_g[size_t(p) * NC + c];
            }
            double m0 = 0;
            V3 m1;
            for (int p = 0; p < NV; ++p) {
                m0 += w[p] * G(p, c);
                m1 += w[p] * G(p, c) * xi[p];
            }
            V3 un = m1 / m0;
            rho[c] = m0;
            u[c] = un;
            for (int p = 0; p < NV; ++p) G(p, c) = (1.0 - C_t_to_bp) * G(p, c) + C_t_to_bp * f_eq(m0, un, xi[p]);
            if (std::isnan(m0)) ++bad;
        }
        bad_cells = bad;
        first_bad = -1;
        if (bad)
            for (int c = 0; c < NC; ++c)
                if (std::isnan(rho[c])) { first_bad = c; break; }
    }

    void do_step() {
        if (shakhov()) {
            reconstruct_shakhov();
            fvm_update_shakhov();
        } else {
            reconstruct_bgk();
            if (P.model == MESO_MODEL_CDUGKS_BGK) fvm_update_cdugks();
            else fvm_update_dugks();
        }
    }

    void reset_residual() {
        rho_res = rho;
        u_res = u;
        T_res = T;
        q_res = q;
    }

    // ---------------- initial() (cdugks_shakhov.cpp:102-168, cdugks.cpp:66-104, dugks.cpp:67-105)
    void init_equilibrium(const double *r0, const double *u0, const double *T0p) {
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NC; ++c) {
            V3 up(u0 + 3 * c);
            double m0 = 0, m2 = 0;
            V3 m1, m3;
            for (int p = 0; p < NV; ++p) {
                if (shakhov()) {
                    V3 cc_ = xi[p] - up;
                    double cc = cc_ * cc_;
                    double kk = xi[p] * xi[p];
                    double gg = g_maxwell(r0[c], T0p[c], cc);
                    double hh = h_maxwell(T0p[c], gg);
                    G(p, c) = gg;
                    H(p, c) = hh;
                    m0 += w[p] * gg;
                    m1 += w[p] * gg * xi[p];
                    m2 += w[p] * (kk * gg + hh);
                    m3 += w[p] * cc_ * (cc * gg + hh);
                } else {
                    double ff = f_eq(r0[c], up, xi[p]);
                    G(p, c) = ff;
                    m0 += w[p] * ff;
                    m1 += w[p] * ff * xi[p];
                }
            }
            rho[c] = m0;
            u[c] = m1 / m0;
            if (shakhov()) {
                double rE = 0.5 * m2;
                T[c] = (rE / m0 - 0.5 * (u[c] * u[c])) / Cv;
                q[c] = 0.5 * m3;
            }
        }
        reset_residual();
    }

    void init_from_macros(const double *r0, const double *u0, const double *T0p, const double *q0) {
#pragma omp parallel for schedule(static)
        for (int c = 0; c < NC; ++c) {
            rho[c] = r0[c];
            u[c] = V3(u0 + 3 * c);
            if (shakhov()) {
                T[c] = T0p[c];
                q[c] = V3(q0 + 3 * c);
            }
            for (int p = 0; p < NV; ++p) {
                if (shakhov()) {
                    V3 cc_ = xi[p] - u[c];
                    double cc = cc_ * cc_, cq = cc_ * q[c];
                    double gm = g_maxwell(rho[c], T[c], cc);
                    G(p, c) = g_shakhov(rho[c], T[c], cc, cq, gm);
                    H(p, c) = h_shakhov(rho[c], T[c], cc, cq, gm);
                } else {
                    G(p, c) = f_eq(rho[c], u[c], xi[p]);
                }
            }
        }
        reset_residual();
    }
};

namespace {

int fail(int code, const std::string &m) {
    g_err = m;
    return code;
}

void build(oracle_ctx &o, const meso_mesh_desc &m, const meso_dvs_desc &d, const meso_step_params &p) {
    o.set_params(p);
    o.D = m.dim;
    o.NC = m.n_cells;
    o.NF = m.n_faces;
    o.NV = d.n_vel;
    o.total_volume = m.total_volume;
    o.cpos.resize(o.NC);
    o.cvol.assign(m.cell_vol, m.cell_vol + o.NC);
    for (int c = 0; c < o.NC; ++c) o.cpos[c] = V3(m.cell_pos + 3 * c);
    o.fpos.resize(o.NF);
    o.fn0.resize(o.NF);
    o.own.resize(o.NF);
    o.nei.resize(o.NF);
    o.farea.assign(m.face_area, m.face_area + o.NF);
    o.fmark.assign(m.face_mark, m.face_mark + o.NF);
    for (int f = 0; f < o.NF; ++f) {
        o.fpos[f] = V3(m.face_pos + 3 * f);
        o.fn0[f] = V3(m.face_normal + 3 * f);
        o.own[f] = m.face_cells[2 * f];
        o.nei[f] = m.face_cells[2 * f + 1];
    }
    o.cf_ptr.assign(m.cell_face_ptr, m.cell_face_ptr + o.NC + 1);
    o.cf_ids.assign(m.cell_face_ids, m.cell_face_ids + m.n_cell_faces);
    o.nb_ptr.assign(m.nbr_ptr, m.nbr_ptr + o.NC + 1);
    o.nb_ids.assign(m.nbr_ids, m.nbr_ids + m.n_nbr);
    o.dr.resize(m.n_nbr);
    for (int j = 0; j < m.n_nbr; ++j) o.dr[j] = V3(m.lsq_dr + 3 * j);
    o.lw.assign(m.lsq_w, m.lsq_w + m.n_nbr);
    o.Cx.resize(o.NC);
    o.Cy.resize(o.NC);
    o.Cz.resize(o.NC);
    for (int c = 0; c < o.NC; ++c) {
        o.Cx[c] = V3(m.lsq_C + 9 * c);
        o.Cy[c] = V3(m.lsq_C + 9 * c + 3);
        o.Cz[c] = V3(m.lsq_C + 9 * c + 6);
    }
    o.marks.assign(m.marks, m.marks + m.n_marks);
    o.xi.resize(o.NV);
    for (int v = 0; v < o.NV; ++v) o.xi[v] = V3(d.xi + 3 * v);
    o.w.assign(d.w, d.w + o.NV);
    if (d.mirror) o.mirror.assign(d.mirror, d.mirror + 3 * o.NV);
    o.nd = (p.model == MESO_MODEL_CDUGKS_SHAKHOV) ? 2 : 1;
    size_t cn = size_t(o.NV) * o.NC, fnn = size_t(o.NV) * o.NF;
    o.g.assign(cn, 0.0);
    o.gface.assign(fnn, 0.0);
    o.flux_g.assign(cn, 0.0);
    if (o.nd == 2) {
        o.h.assign(cn, 0.0);
        o.hface.assign(fnn, 0.0);
        o.flux_h.assign(cn, 0.0);
    }
    o.rho.assign(o.NC, 0.0);
    o.T.assign(o.NC, 0.0);
    o.u.assign(o.NC, V3());
    o.q.assign(o.NC, V3());
    o.fm0.assign(o.NC, 0.0);
    o.fm2.assign(o.NC, 0.0);
    o.fm1.assign(o.NC, V3());
    o.reset_residual();
}

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

void oracle_set_threads(int32_t n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int32_t oracle_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int oracle_create(const meso_mesh_desc *m, const meso_dvs_desc *d, const meso_step_params *p, oracle_ctx **out) {
    if (!m || !d || !p || !out) return fail(MESO_E_INVALID, "null argument");
    if (p->model == MESO_MODEL_CDUGKS_SHAKHOV || p->model == MESO_MODEL_CDUGKS_BGK || p->model == MESO_MODEL_DUGKS_BGK) {
    } else
        return fail(MESO_E_INVALID, "unknown model");
    try {
        auto *o = new oracle_ctx();
        build(*o, *m, *d, *p);
        *out = o;
        return MESO_OK;
    } catch (const std::exception &e) {
        return fail(MESO_E_INVALID, e.what());
    }
}

void oracle_destroy(oracle_ctx *o) { delete o; }

int oracle_set_mask(oracle_ctx *o, const int32_t *m) {
    if (!o) return fail(MESO_E_INVALID, "null argument");
    o->mask.assign(o->NC, 0);
    if (m)
        for (int c = 0; c < o->NC; ++c) o->mask[c] = m[c] ? 1 : 0;
    return MESO_OK;
}

int oracle_set_params(oracle_ctx *o, const meso_step_params *p) {
    if (!o || !p) return fail(MESO_E_INVALID, "null argument");
    o->set_params(*p);
    return MESO_OK;
}

int oracle_init_equilibrium(oracle_ctx *o, const double *r, const double *u3, const double *T) {
    if (!o || !r || !u3 || (o->shakhov() && !T)) return fail(MESO_E_INVALID, "null argument");
    o->init_equilibrium(r, u3, T);
    return MESO_OK;
}

int oracle_init_from_macros(oracle_ctx *o, const double *r, const double *u3, const double *T, const double *q3) {
    if (!o || !r || !u3 || (o->shakhov() && (!T || !q3))) return fail(MESO_E_INVALID, "null argument");
    o->init_from_macros(r, u3, T, q3);
    return MESO_OK;
}

int oracle_set_state(oracle_ctx *o, const double *f, int32_t layout, const double *r, const double *u3,
                     const double *T, const double *q3) {
    if (!o || !f || !r || !u3) return fail(MESO_E_INVALID, "null argument");
    const size_t NC = o->NC, NV = o->NV;
    for (int d = 0; d < o->nd; ++d) {
        std::vector<double> &dst = d == 0 ? o->g : o->h;
        const double *src = f + size_t(d) * NC * NV;
        for (size_t p = 0; p < NV; ++p)
            for (size_t c = 0; c < NC; ++c)
                dst[p * NC + c] = layout == MESO_LAYOUT_VELOCITY_MAJOR ? src[p * NC + c] : src[c * NV + p];
    }
    for (int c = 0; c < o->NC; ++c) {
        o->rho[c] = r[c];
        o->u[c] = V3(u3 + 3 * c);
        o->T[c] = T ? T[c] : 0.0;
        o->q[c] = q3 ? V3(q3 + 3 * c) : V3();
    }
    o->reset_residual();
    return MESO_OK;
}

int oracle_get_state(oracle_ctx *o, double *f, int32_t layout) {
    if (!o || !f) return fail(MESO_E_INVALID, "null argument");
    const size_t NC = o->NC, NV = o->NV;
    for (int d = 0; d < o->nd; ++d) {
        const std::vector<double> &src = d == 0 ? o->g : o->h;
        double *dst = f + size_t(d) * NC * NV;
        for (size_t p = 0; p < NV; ++p)
            for (size_t c = 0; c < NC; ++c)
                (layout == MESO_LAYOUT_VELOCITY_MAJOR ? dst[p * NC + c] : dst[c * NV + p]) = src[p * NC + c];
    }
    return MESO_OK;
}

int oracle_step(oracle_ctx *o, int32_t n, meso_step_report *rep) {
    if (!o || n < 0) return fail(MESO_E_INVALID, "bad argument");
    const int masked0 = o->masked_cells;
    int done = 0;
    for (; done < n; ++done) {
        o->do_step();
        if (o->bad_cells) {
            ++done;
            break;
        }
    }
    if (rep) {
        rep->steps_done = done;
        rep->bad_cells = o->bad_cells;
        rep->first_bad_cell = o->first_bad;
        rep->masked_cells = o->masked_cells - masked0;
    }
    return o->bad_cells ? fail(MESO_E_NUMERIC, "NaN density or T<0 in a cell") : MESO_OK;
}

int oracle_get_macros(oracle_ctx *o, double *r, double *u3, double *T, double *q3) {
    if (!o) return fail(MESO_E_INVALID, "null argument");
    for (int c = 0; c < o->NC; ++c) {
        if (r) r[c] = o->rho[c];
        if (T) T[c] = o->T[c];
        for (int k = 0; k < 3; ++k) {
            if (u3) u3[3 * c + k] = k == 0 ? o->u[c].x : (k == 1 ? o->u[c].y : o->u[c].z);
            if (q3) q3[3 * c + k] = k == 0 ? o->q[c].x : (k == 1 ? o->q[c].y : o->q[c].z);
        }
    }
    return MESO_OK;
}

int oracle_get_flux_moments(oracle_ctx *o, double *m0, double *m1, double *m2) {
    if (!o) return fail(MESO_E_INVALID, "null argument");
    for (int c = 0; c < o->NC; ++c) {
        if (m0) m0[c] = o->fm0[c];
        if (m2) m2[c] = o->fm2[c];
        if (m1) {
            m1[3 * c] = o->fm1[c].x;
            m1[3 * c + 1] = o->fm1[c].y;
            m1[3 * c + 2] = o->fm1[c].z;
        }
    }
    return MESO_OK;
}

int oracle_residual(oracle_ctx *o, double out[8]) {
    if (!o || !out) return fail(MESO_E_INVALID, "null argument");
    // field.cpp:222-249
    double rr = 0, rt = 0;
    V3 ru, rq;
    for (int c = 0; c < o->NC; ++c) {
        double wv = o->cvol[c] / o->total_volume;
        rr += std::abs((o->rho[c] - o->rho_res[c]) / o->rho_res[c]) * wv;
        const V3 &a = o->u[c], &b = o->u_res[c];
        ru += V3(std::fabs((a.x - b.x) / b.x), std::fabs((a.y - b.y) / b.y), std::fabs((a.z - b.z) / b.z)) * wv;
        if (o->shakhov()) {
            rt += std::abs((o->T[c] - o->T_res[c]) / o->T_res[c]) * wv;
            const V3 &qa = o->q[c], &qb = o->q_res[c];
            rq += V3(std::fabs((qa.x - qb.x) / qb.x), std::fabs((qa.y - qb.y) / qb.y),
                     std::fabs((qa.z - qb.z) / qb.z)) *
                  wv;
        }
    }
    o->reset_residual();
    out[0] = rr;
    out[1] = ru.x;
    out[2] = ru.y;
    out[3] = ru.z;
    out[4] = rt;
    out[5] = rq.x;
    out[6] = rq.y;
    out[7] = rq.z;
    return MESO_OK;
}

int oracle_grad(const oracle_ctx *o, const double *field, double *grad) {
    if (!o || !field || !grad) return fail(MESO_E_INVALID, "null argument");
    std::vector<V3> g;
    o->grad(field, g);
    for (int c = 0; c < o->NC; ++c) {
        grad[3 * c] = g[c].x;
        grad[3 * c + 1] = g[c].y;
        grad[3 * c + 2] = g[c].z;
    }
    return MESO_OK;
}

int oracle_face_values(const oracle_ctx *o, double *ff) {
    if (!o || !ff) return fail(MESO_E_INVALID, "null argument");
    std::memcpy(ff, o->gface.data(), o->gface.size() * sizeof(double));
    if (o->nd == 2) std::memcpy(ff + o->gface.size(), o->hface.data(), o->hface.size() * sizeof(double));
    return MESO_OK;
}

int oracle_fbar(oracle_ctx *o, const double *cell_f, double *face_f) {
    if (!o || !cell_f || !face_f) return fail(MESO_E_INVALID, "null argument");
    std::vector<double> cv(cell_f, cell_f + size_t(o->NV) * o->NC), fv(size_t(o->NV) * o->NF);
    o->reconstruct_faces(cv, fv);
    std::memcpy(face_f, fv.data(), fv.size() * sizeof(double));
    return MESO_OK;
}

}  // extern "C"
