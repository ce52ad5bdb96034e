// Device-side building blocks of the gfx950 DUGKS/CDUGKS step.
//
// Layout in HBM: every distribution is cell-major, f[c * nvp + v], the
// velocity index fastest so that the lanes of a wave (lanes along v) read and
// write contiguous 256 B (FP32) / 512 B (FP64) segments of one cell row.
// nvp is nv rounded up to 64 when nv >= 64 (rows start on 256 B boundaries);
// padded entries stay zero and are never read as velocities.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace meso_dev {

constexpr double VSMALL = 1e-30;  // upwind dead-band, include/core/math.h:4
constexpr double PI = 3.14159265358979323846;

enum Model : int { CDUGKS_BGK = 1, CDUGKS_SHAKHOV = 2, DUGKS_BGK = 3 };

// Boundary kinds = meso BoundaryType values.
enum Bc : int {
    BC_INTERIOR = 0,
    BC_FARFIELD_INLET = 1,
    BC_PRESSURE_INLET = 2,
    BC_FREESTREAM_INLET = 3,
    BC_FARFIELD_OUTLET = 4,
    BC_PRESSURE_OUTLET = 5,
    BC_WALL = 6,
    BC_SLIP_WALL = 7,
    BC_SYMMETRY = 8,
};

// FP32X storage: a value is an FP32 word plus one byte that extends its
// significand by 8 bits. The FP64 value is first rounded at 31 explicit
// mantissa bits; in units of 2^21 of its FP64 pattern that is an integer q.
// The word is q rounded to nearest at 23 explicit bits (a multiple of 256 in
// those units, so the word converted to FP64 has zeros below), and the byte
// is the remainder q - word, a signed value in [-128, 127]. Decoding adds the
// sign-extended byte to the word's FP64 pattern: one add. The word alone is
// the value rounded to nearest FP32, so a kernel that needs only FP32 values
// reads the words and not the bytes.
__device__ __forceinline__ void x32_pack(unsigned long long b, float &hi, unsigned char &ext) {
    const unsigned long long q = b >> 21;
    const unsigned long long w = (q + 128u) & ~0xffull;
    ext = (unsigned char)((q - w) & 0xffu);
    hi = float(__longlong_as_double((long long)(w << 21)));  // exact
}
__device__ __forceinline__ void x32_encode(double v, float &hi, unsigned char &ext) {
    x32_pack((unsigned long long)__double_as_longlong(v) + (1ull << 20), hi, ext);
}
// The same with stochastic rounding at bit 31: r holds 21 random bits, so the
// value rounds up with the probability of its position between the two
// neighbours. Unbiased, so that the small per-step increments of a run near
// its steady state are not lost the way round-to-nearest loses the ones below
// half a unit in the last place.
__device__ __forceinline__ void x32_encode_sr(double v, unsigned r, float &hi, unsigned char &ext) {
    x32_pack((unsigned long long)__double_as_longlong(v) + (r & ((1u << 21) - 1u)), hi, ext);
}
// 32-bit mixing function (xorshift-multiply; "lowbias32" constants): the
// random bits of the stochastic rounding, from (entry index, step seed).
__host__ __device__ __forceinline__ unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// FP32 with stochastic rounding: r holds 29 random bits, added to the FP64
// significand below its 23 kept bits before truncating, so that v rounds up
// with the probability of its position between its two FP32 neighbours
// (unbiased; the stored state then does not stall near a steady state the way
// round-to-nearest makes it). Non-finite values are converted as they are.
__device__ __forceinline__ float f32_sr(double v, unsigned r) {
    if (!(fabs(v) <= 1.0e300)) return float(v);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v) + (r & ((1u << 29) - 1u));
    return float(__longlong_as_double((long long)(b & ~((1ull << 29) - 1))));
}
// Second random word from a first (the other distribution of the same entry).
__device__ __forceinline__ unsigned rot16(unsigned x) { return (x << 16) | (x >> 16); }
// The low byte of ext is the signed remainder.
__device__ __forceinline__ double x32_decode(float hi, unsigned ext) {
    const long long b = __double_as_longlong(double(hi));
    return __longlong_as_double(b + (long long)(signed char)(ext & 0xffu) * (1ll << 21));
}

// Per-face equilibrium description, written by the face-macro kernel.
// Shakhov: g_S = A exp(-B|c|^2) [1 + P1 (c.q)(|c|^2 iRT - D - 2)]
//          h_S = g_M [hk + P2 (c.q)((|c|^2 iRT - D)(K+3-D) - 2K)],  c = xi - u
// BGK:     f_eq = rho (1 + xi.u iRT + ((xi.u)^2 iRT^2 - |u|^2 iRT)/2), A = rho
struct EqParams {
    double u[3];
    double q[3];
    double A, B, iRT, P1, P2, hk;
    double mix;  // weight of the equilibrium in the face value (C_s or c_eq)
    // Deviation from the reference equilibrium (FP32 storage only):
    // Shakhov: g_S = g_ref exp(lnAr + dB|xi|^2 + 2B xi.u - B|u|^2) [...],
    //          dhk = hk - hk_ref;  BGK: lnAr = rho - rho_ref.
    double lnAr, dB, dhk;
};

// FP32 copy of the deviation form of EqParams, converted once per face so the
// per-velocity loop reads 16 floats: g_S - g_ref = g_ref [expm1(dl)(1 + S) + S],
// dl = lnAr + dB |xi|^2 + B2 xi.u - Buu.
struct EqF {
    float u[3], q[3];
    float lnAr, dB, B2, Buu, iRT, P1, P2, hk, dhk, mix;
    float a0, oa0;  // expm1(lnAr), 1 + expm1(lnAr) for the axis-table evaluation
};

__host__ __device__ inline EqF to_eqf(const EqParams &e) {
    EqF r;
    for (int k = 0; k < 3; ++k) {
        r.u[k] = float(e.u[k]);
        r.q[k] = float(e.q[k]);
    }
    r.lnAr = float(e.lnAr);
    r.dB = float(e.dB);
    r.B2 = float(2.0 * e.B);
    r.Buu = float(e.B * (e.u[0] * e.u[0] + e.u[1] * e.u[1] + e.u[2] * e.u[2]));
    r.iRT = float(e.iRT);
    r.P1 = float(e.P1);
    r.P2 = float(e.P2);
    r.hk = float(e.hk);
    r.dhk = float(e.dhk);
    r.mix = float(e.mix);
    const double a0 = expm1(e.lnAr);
    r.a0 = float(a0);
    r.oa0 = float(1.0 + a0);
    return r;
}

// Axis tables for tensor-product velocity sets. With xi = (X_i, Y_j, Z_k) the
// Maxwellian factorises over axes:
//   FP64 storage: g_M = A Tx[i] Ty[j] Tz[k],  T_d[i] = exp(-B (X_i - u_d)^2)
//   FP32 storage: g_M / g_ref - 1 = (1 + a0) prod_d (1 + T_d[i]) - 1,
//                 T_d[i] = expm1(B_ref X_i^2 - B (X_i - u_d)^2), a0 = expm1(lnAr)
// so the per-velocity work is a few multiplies instead of an exponential.
// vt packs the table offsets of a velocity's axis entries (10 bits each).
// FP32 Shakhov pair (deviation form) of one velocity from its axis-table
// entries ax, ay (and az in 3D), read beforehand.
__device__ __forceinline__ void shakhov_tab_vals_f(const EqF &e, float ax, float ay, float az, const float xi[3], int D,
                                                   int K, float gref, float &gs, float &hs) {
    float t = fmaf(ax, ay, ax + ay);
    if (D == 3) t = fmaf(t, az, t + az);
    const float em1 = fmaf(t, e.oa0, e.a0);
    const float c0 = xi[0] - e.u[0], c1 = xi[1] - e.u[1], c2 = xi[2] - e.u[2];
    const float cc = c0 * c0 + c1 * c1 + c2 * c2;
    const float cq = c0 * e.q[0] + c1 * e.q[1] + c2 * e.q[2];
    const float ccr = cc * e.iRT;
    const float S = e.P1 * cq * (ccr - float(D) - 2.0f);
    const float H = e.P2 * cq * ((ccr - float(D)) * float(K + 3 - D) - float(2 * K));
    gs = gref * fmaf(S, 1.0f + em1, em1);
    hs = gref * fmaf(em1, e.hk + H, e.dhk + H);
}

// The same, reading the entries from the table tb of one face or cell.
__device__ __forceinline__ void shakhov_tab_f(const EqF &e, const float *tb, int vt, const float xi[3], int D, int K,
                                              float gref, float &gs, float &hs) {
    shakhov_tab_vals_f(e, tb[vt & 1023], tb[(vt >> 10) & 1023], D == 3 ? tb[(vt >> 20) & 1023] : 0.0f, xi, D, K, gref,
                       gs, hs);
}

// FP64 Shakhov pair of one velocity from its axis-table entries ax, ay (and
// az in 3D), read beforehand.
__device__ __forceinline__ void shakhov_tab_vals_d(const EqParams &e, double ax, double ay, double az,
                                                   const double xi[3], int D, int K, double &gs, double &hs) {
    double gm = e.A * ax * ay;
    if (D == 3) gm *= az;
    const double c0 = xi[0] - e.u[0], c1 = xi[1] - e.u[1], c2 = xi[2] - e.u[2];
    const double cc = c0 * c0 + c1 * c1 + c2 * c2;
    const double cq = c0 * e.q[0] + c1 * e.q[1] + c2 * e.q[2];
    const double ccr = cc * e.iRT;
    gs = gm * (1.0 + e.P1 * cq * (ccr - double(D) - 2.0));
    hs = gm * (e.hk + e.P2 * cq * ((ccr - double(D)) * double(K + 3 - D) - double(2 * K)));
}

// The same, reading the entries from the table tb of one face or cell.
__device__ __forceinline__ void shakhov_tab_d(const EqParams &e, const double *tb, int vt, const double xi[3], int D,
                                              int K, double &gs, double &hs) {
    shakhov_tab_vals_d(e, tb[vt & 1023], tb[(vt >> 10) & 1023], D == 3 ? tb[(vt >> 20) & 1023] : 0.0, xi, D, K, gs,
                       hs);
}

__device__ __forceinline__ void shakhov_devf(const EqF &e, const float xi[3], int D, int K, float gref, float &gs,
                                             float &hs) {
    const float c0 = xi[0] - e.u[0], c1 = xi[1] - e.u[1], c2 = xi[2] - e.u[2];
    const float cc = c0 * c0 + c1 * c1 + c2 * c2;
    const float cq = c0 * e.q[0] + c1 * e.q[1] + c2 * e.q[2];
    const float xx = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    const float xu = xi[0] * e.u[0] + xi[1] * e.u[1] + xi[2] * e.u[2];
    const float dl = e.lnAr + e.dB * xx + e.B2 * xu - e.Buu;
    const float em1 = expm1f(dl);
    const float ccr = cc * e.iRT;
    const float S = e.P1 * cq * (ccr - float(D) - 2.0f);
    const float H = e.P2 * cq * ((ccr - float(D)) * float(K + 3 - D) - float(2 * K));
    gs = gref * (em1 * (1.0f + S) + S);
    hs = gref * (em1 * (e.hk + H) + e.dhk + H);
}

__device__ __forceinline__ float bgk_devf(const EqF &e, const float xi[3]) {
    // for BGK the EqF fields hold: lnAr = rho - rho_ref, hk = rho
    const float ku = (xi[0] * e.u[0] + xi[1] * e.u[1] + xi[2] * e.u[2]) * e.iRT;
    const float uu = (e.u[0] * e.u[0] + e.u[1] * e.u[1] + e.u[2] * e.u[2]) * e.iRT;
    return e.lnAr + e.hk * (ku + (ku * ku - uu) * 0.5f);
}

// Parameters of one boundary face after the wall / inlet / outlet stage: the
// value imposed on incoming velocities is the equilibrium `in` (Maxwellian
// of rho_b, u_b, T_b; BGK polynomial), optionally mixed with the mirror
// velocity's value.
struct BndParams {
    EqParams in;
    int kind;       // Bc
    int dir;        // mirror axis
    double alpha_u; // weight of the mirror value for g (slip-wall)
    double alpha_T; // weight of the mirror value for h (slip-wall)
    double pad;
};

struct StepConst {
    int model, D, K, grad;
    int nc, nf, nv, nvp;
    int nb;            // boundary faces
    int ntile_f, vt_f; // velocity tiles of the face-moment pass
    int ntile_c, vt_c; // velocity tiles of the cell pass
    double dt, half_dt;
    double Pr, R, T0, mu0, omega, Cv, gamma;
    double RT, tau;    // BGK
    int limiter;       // Shakhov: Venkatakrishnan limiter of the reconstruction
    double venkata_k;  // its k (meso's venkata_k; 0 unless read at construction)
};

// Venkatakrishnan limiter factor of a cell for the face increment
// dw = (x_f - x_c) . grad f (Solver::venkata_limiter, src/solver/limiter.cpp:7-31):
// w the cell value, [mn, mx] the range over the cell and its face neighbours,
// omega = (k V^(1/D))^3. eps is 1e-10 where meso's owner is upwind and VSMALL
// otherwise (cdugks_shakhov.cpp:234, 247, 262-265).
__host__ __device__ inline double venkata_phi(double dw, double w, double mn, double mx, double omega, double eps) {
    if (fabs(dw) <= eps) return 1.0;
    const double a = dw > 0.0 ? mx - w : mn - w;
    const double aa = a * a, ab = a * dw, bb = dw * dw;
    return (aa + 2.0 * ab + omega) / (aa + 2.0 * bb + ab + omega);
}

// omega of cell c (0 for k = 0, the value meso's limiter normally sees).
__host__ __device__ inline double venkata_omega(double k, double vol, int D) {
    if (k == 0.0) return 0.0;
    const double kh = k * pow(vol, 1.0 / double(D));
    return kh * kh * kh;
}

template <class R>
__device__ __forceinline__ R dexp(R x);
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <>
__device__ __forceinline__ float dexp<float>(float x) { return expf(x); }

// Shakhov equilibrium pair at velocity xi for face/cell constants e.
template <class R>
__device__ __forceinline__ void shakhov_eq(const EqParams &e, const R xi[3], int D, int K, R &gs, R &hs) {
    R c0 = xi[0] - R(e.u[0]), c1 = xi[1] - R(e.u[1]), c2 = xi[2] - R(e.u[2]);
    R cc = c0 * c0 + c1 * c1 + c2 * c2;
    R cq = c0 * R(e.q[0]) + c1 * R(e.q[1]) + c2 * R(e.q[2]);
    R gm = R(e.A) * dexp<R>(-cc * R(e.B));
    R ccr = cc * R(e.iRT);
    gs = gm * (R(1) + R(e.P1) * cq * (ccr - R(D) - R(2)));
    hs = gm * (R(e.hk) + R(e.P2) * cq * ((ccr - R(D)) * R(K + 3 - D) - R(2 * K)));
}

// Incompressible BGK polynomial equilibrium (cdugks.cpp:116-121).
template <class R>
__device__ __forceinline__ R bgk_eq(const EqParams &e, const R xi[3]) {
    R ku = (xi[0] * R(e.u[0]) + xi[1] * R(e.u[1]) + xi[2] * R(e.u[2])) * R(e.iRT);
    R uu = (R(e.u[0]) * R(e.u[0]) + R(e.u[1]) * R(e.u[1]) + R(e.u[2]) * R(e.u[2])) * R(e.iRT);
    return R(e.A) * (R(1) + ku + (ku * ku - uu) * R(0.5));
}

// Host/device: constants of the Shakhov equilibrium of (rho, u, T, q).
__host__ __device__ inline void make_shakhov(EqParams &e, double rho, const double u[3], double T,
                                             const double q[3], double R, double Pr, int D, int K) {
    double RT = R * T;
    double p = rho * RT;
    for (int k = 0; k < 3; ++k) {
        e.u[k] = u[k];
        e.q[k] = q[k];
    }
    if (D == 2) {
        double over_RT2 = 0.5 / RT;
        e.A = rho * over_RT2 / PI;
        e.B = over_RT2;
    } else {
        e.A = rho / pow(2.0 * PI * RT, 1.5);
        e.B = 0.5 / RT;
    }
    e.iRT = 1.0 / RT;
    e.P1 = (1.0 - Pr) / (5.0 * p * RT);
    e.P2 = (1.0 - Pr) / (5.0 * p);
    e.hk = (K + 3.0 - D) * RT;
    e.mix = 0.0;
    e.lnAr = e.dB = e.dhk = 0.0;
}

// Fill the deviation terms of e against the reference equilibrium r.
__host__ __device__ inline void relate_shakhov(EqParams &e, const EqParams &r) {
    e.lnAr = log(e.A / r.A);
    e.dB = r.B - e.B;
    e.dhk = e.hk - r.hk;
}

__host__ __device__ inline void make_bgk(EqParams &e, double rho, const double u[3], double RT) {
    for (int k = 0; k < 3; ++k) {
        e.u[k] = u[k];
        e.q[k] = 0.0;
    }
    e.A = rho;
    e.B = 0.0;
    e.iRT = 1.0 / RT;
    e.P1 = e.P2 = e.hk = 0.0;
    e.mix = 0.0;
    e.lnAr = e.dB = e.dhk = 0.0;
}

__host__ __device__ inline void relate_bgk(EqParams &e, const EqParams &r) { e.lnAr = e.A - r.A; }

// Shakhov pair as deviations from the reference pair (gref, href = hk_ref gref):
//   g_S - gref = gref [expm1(Dl)(1 + S) + S]
//   h_S - href = gref [expm1(Dl)(hk + H) + dhk + H]
// with Dl = lnAr + dB|xi|^2 + 2B xi.u - B|u|^2, so the rounding error scales
// with the deviation, not with the distribution.
__device__ __forceinline__ void shakhov_dev(const EqParams &e, const float xi[3], int D, int K, float gref,
                                            float &gs, float &hs) {
    const float u0 = float(e.u[0]), u1 = float(e.u[1]), u2 = float(e.u[2]);
    const float c0 = xi[0] - u0, c1 = xi[1] - u1, c2 = xi[2] - u2;
    const float cc = c0 * c0 + c1 * c1 + c2 * c2;
    const float cq = c0 * float(e.q[0]) + c1 * float(e.q[1]) + c2 * float(e.q[2]);
    const float xx = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    const float xu = xi[0] * u0 + xi[1] * u1 + xi[2] * u2;
    const float dl = float(e.lnAr) + float(e.dB) * xx + float(2.0 * e.B) * xu - float(e.B * (e.u[0] * e.u[0] + e.u[1] * e.u[1] + e.u[2] * e.u[2]));
    const float em1 = expm1f(dl);
    const float ccr = cc * float(e.iRT);
    const float S = float(e.P1) * cq * (ccr - float(D) - 2.0f);
    const float H = float(e.P2) * cq * ((ccr - float(D)) * float(K + 3 - D) - float(2 * K));
    gs = gref * (em1 * (1.0f + S) + S);
    hs = gref * (em1 * (float(e.hk) + H) + float(e.dhk) + H);
}

// BGK polynomial equilibrium minus rho_ref.
__device__ __forceinline__ float bgk_dev(const EqParams &e, const float xi[3]) {
    const float u0 = float(e.u[0]), u1 = float(e.u[1]), u2 = float(e.u[2]);
    const float ku = (xi[0] * u0 + xi[1] * u1 + xi[2] * u2) * float(e.iRT);
    const float uu = (u0 * u0 + u1 * u1 + u2 * u2) * float(e.iRT);
    return float(e.lnAr) + float(e.A) * (ku + (ku * ku - uu) * 0.5f);
}

// tau(rho, T) = mu0 (T/T0)^omega / (rho R T)   (cdugks_shakhov.cpp:181-183)
__host__ __device__ inline double tau_of(const StepConst &s, double rho, double T) {
    return s.mu0 * pow(T / s.T0, s.omega) / (rho * s.R * T);
}

// Sum over the LP lanes of one segment (LP a power of two <= 64).
template <int LP>
__device__ __forceinline__ double seg_sum(double x) {
#pragma unroll
    for (int o = LP / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Logical block id that gives each XCD a contiguous range of block ids:
// hardware dispatches block b to XCD b % 8, so consecutive blocks of one
// XCD get consecutive logical ids and share cell rows in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    constexpr int NX = 8;
    int per = (nblocks + NX - 1) / NX;
    int x = b % NX, k = b / NX;
    int full = nblocks - (per - 1) * NX;  // XCDs that receive `per` blocks
    int base = (x < full) ? x * per : full * per + (x - full) * (per - 1);
    return base + k;
}

}  // namespace meso_dev
