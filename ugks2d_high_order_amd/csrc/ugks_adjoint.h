// Reverse-mode building blocks of the S2CN2 step (the discrete adjoint).
//
// Every function here is the hand-written reverse of one forward piece of the
// step: given the adjoint of the piece's outputs it accumulates the adjoint of
// its inputs.  Per-velocity pieces are reversed by hand; the short per-node
// scalar chains (primitive variables, collision time, time integrals, the
// multiscale wall's micro slopes and half-range moments) are written once as
// templates and differentiated with a tangent (value, derivative) type, one
// input direction at a time, to form the transposed Jacobian-vector product.
#pragma once
#include "ugks_node.h"

// ---------------------------------------------------------------------------
// tangent numbers for the scalar chains
// ---------------------------------------------------------------------------
struct Tan {
    double v, d;
    __device__ Tan() : v(0.0), d(0.0) {}
    __device__ Tan(double x) : v(x), d(0.0) {}
    __device__ Tan(double x, double dx) : v(x), d(dx) {}
};
__device__ inline Tan operator+(Tan a, Tan b) { return Tan(a.v + b.v, a.d + b.d); }
__device__ inline Tan operator-(Tan a, Tan b) { return Tan(a.v - b.v, a.d - b.d); }
__device__ inline Tan operator-(Tan a) { return Tan(-a.v, -a.d); }
__device__ inline Tan operator*(Tan a, Tan b) { return Tan(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline Tan operator/(Tan a, Tan b) {
    const double q = a.v / b.v;
    return Tan(q, (a.d - q * b.d) / b.v);
}
__device__ inline Tan operator+(Tan a, double b) { return Tan(a.v + b, a.d); }
__device__ inline Tan operator+(double a, Tan b) { return Tan(a + b.v, b.d); }
__device__ inline Tan operator-(Tan a, double b) { return Tan(a.v - b, a.d); }
__device__ inline Tan operator-(double a, Tan b) { return Tan(a - b.v, -b.d); }
__device__ inline Tan operator*(Tan a, double b) { return Tan(a.v * b, a.d * b); }
__device__ inline Tan operator*(double a, Tan b) { return Tan(a * b.v, a * b.d); }
__device__ inline Tan operator/(Tan a, double b) { return Tan(a.v / b, a.d / b); }
__device__ inline Tan operator/(double a, Tan b) {
    const double q = a / b.v;
    return Tan(q, -q * b.d / b.v);
}
__device__ inline double tval(double x) { return x; }
__device__ inline double tval(Tan x) { return x.v; }
__device__ inline double texp(double x) { return exp(x); }
__device__ inline Tan texp(Tan x) {
    const double e = exp(x.v);
    return Tan(e, e * x.d);
}
__device__ inline double tsqrt(double x) { return sqrt(x); }
__device__ inline Tan tsqrt(Tan x) {
    const double r = sqrt(x.v);
    return Tan(r, 0.5 * x.d / r);
}
__device__ inline double tpow(double x, double p) { return pow(x, p); }
__device__ inline Tan tpow(Tan x, double p) {
    const double r = pow(x.v, p);
    return Tan(r, p * pow(x.v, p - 1.0) * x.d);
}
__device__ inline double terfc(double x) { return erfc(x); }
__device__ inline Tan terfc(Tan x) { return Tan(erfc(x.v), -2.0 / sqrt(UGKS_PI) * exp(-x.v * x.v) * x.d); }

// collision time of conserved w (tau_of_w)
template <class T>
__device__ inline T tau_t(int ck, double omega, const T w[4], T mu_ref) {
    T p = (2.0 / (double)(ck + 2)) * (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
    T lam = 0.5 * w[0] / p;
    T mu = mu_ref * tpow(1.0 / lam, omega);
    return mu / (0.5 * w[0] / lam);
}

// Scalars of one UGKS face node as functions of its moments w0, Pr and
// mu_ref (ugks_node): out = {rho, ux, vy, lam, coeff, tna, lama, m0..m4}.
enum { FC_RHO, FC_UX, FC_VY, FC_LAM, FC_COEFF, FC_TNA, FC_LAMA, FC_M0, FC_N = FC_M0 + 5 };
template <class T>
__device__ inline void face_chain(int ck, double omega, double dt, const T w0[4], T pr, T mu_ref, T out[FC_N]) {
    T rho = w0[0], ux = w0[1] / w0[0], vy = w0[2] / w0[0];
    T eint = w0[3] - 0.5 * w0[0] * (ux * ux + vy * vy);
    T thg = ((2.0 / (double)(ck + 2)) * eint) / w0[0];
    T lam = 0.5 / thg;
    T coeff = 0.8 * (1.0 - pr) * lam * lam / rho;
    T kin = 0.5 * (ux * ux + vy * vy);
    T tna = 4.0 * (w0[3] / rho - kin) / (double)(ck + 2);
    T lama = 1.0 / tna;
    T tau = tau_t(ck, omega, w0, mu_ref);
    T ex = texp(-dt / tau);
    T m3 = tau * (1.0 - ex);
    T m4 = -tau * dt * ex + tau * m3;
    T m0 = dt - m3;
    T m1 = -tau * m0 + m4;
    T m2 = 0.5 * dt * dt - tau * m0;
    out[FC_RHO] = rho; out[FC_UX] = ux; out[FC_VY] = vy; out[FC_LAM] = lam; out[FC_COEFF] = coeff;
    out[FC_TNA] = tna; out[FC_LAMA] = lama;
    out[FC_M0] = m0; out[FC_M0 + 1] = m1; out[FC_M0 + 2] = m2; out[FC_M0 + 3] = m3; out[FC_M0 + 4] = m4;
}

// multiscale wall pieces (MicroSlope, CalcMoment, Moment_auvxi, the time
// coefficients) as templates of the scalar type
template <class T>
__device__ inline void micro_slope_t(int ck, const T prim[4], const T sw[4], T a[4]) {
    a[3] = 4.0 * prim[3] * prim[3] / ((double)(ck + 2) * prim[0]) *
           (2.0 * sw[3] - 2.0 * prim[1] * sw[1] - 2.0 * prim[2] * sw[2] +
            sw[0] * (prim[1] * prim[1] + prim[2] * prim[2] - 0.5 * (double)(ck + 2) / prim[3]));
    a[2] = 2.0 * prim[3] / prim[0] * (sw[2] - prim[2] * sw[0]) - prim[2] * a[3];
    a[1] = 2.0 * prim[3] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * a[3];
    a[0] = sw[0] / prim[0] - prim[1] * a[1] - prim[2] * a[2] -
           0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + 0.5 * (double)(ck + 2) / prim[3]) * a[3];
}

template <class T>
__device__ inline void half_moments_t(int ck, const T prim[4], T mu[7], T mv[6], T mxi[3]) {
    T mul[7], mur[7];
    T sl = tsqrt(prim[3]);
    T g = 0.5 * texp(-prim[3] * prim[1] * prim[1]) / tsqrt(UGKS_PI * prim[3]);
    mul[0] = 0.5 * terfc(-sl * prim[1]);
    mur[0] = 0.5 * terfc(sl * prim[1]);
    mul[1] = prim[1] * mul[0] + g;
    mur[1] = prim[1] * mur[0] - g;
    for (int i = 2; i <= 6; ++i) {
        mul[i] = prim[1] * mul[i - 1] + 0.5 * (double)(i - 1) * mul[i - 2] / prim[3];
        mur[i] = prim[1] * mur[i - 1] + 0.5 * (double)(i - 1) * mur[i - 2] / prim[3];
    }
    for (int i = 0; i <= 6; ++i) mu[i] = mul[i] + mur[i];
    mv[0] = T(1.0);
    mv[1] = prim[2];
    for (int i = 2; i <= 5; ++i) mv[i] = prim[2] * mv[i - 1] + 0.5 * (double)(i - 1) * mv[i - 2] / prim[3];
    mxi[0] = T(1.0);
    mxi[1] = 0.5 * (double)ck / prim[3];
    mxi[2] = (double)ck * (ck + 2.0) / (4.0 * prim[3] * prim[3]);
}

template <class T>
__device__ inline void m_uvxi_t(const T mu[7], const T mv[6], const T mxi[3], int al, int be, int de, T o[4]) {
    o[0] = mu[al] * mv[be] * mxi[de / 2];
    o[1] = mu[al + 1] * mv[be] * mxi[de / 2];
    o[2] = mu[al] * mv[be + 1] * mxi[de / 2];
    o[3] = 0.5 * (mu[al + 2] * mv[be] * mxi[de / 2] + mu[al] * mv[be + 2] * mxi[de / 2] +
                  mu[al] * mv[be] * mxi[(de + 2) / 2]);
}

template <class T>
__device__ inline void m_auvxi_t(const T a[4], const T mu[7], const T mv[6], const T mxi[3], int al, int be, T o[4]) {
    T t0[4], t1[4], t2[4], t3[4], t4[4], t5[4];
    m_uvxi_t(mu, mv, mxi, al, be, 0, t0);
    m_uvxi_t(mu, mv, mxi, al + 1, be, 0, t1);
    m_uvxi_t(mu, mv, mxi, al, be + 1, 0, t2);
    m_uvxi_t(mu, mv, mxi, al + 2, be, 0, t3);
    m_uvxi_t(mu, mv, mxi, al, be + 2, 0, t4);
    m_uvxi_t(mu, mv, mxi, al, be, 2, t5);
    for (int c = 0; c < 4; ++c)
        o[c] = a[0] * t0[c] + a[1] * t1[c] + a[2] * t2[c] + 0.5 * a[3] * t3[c] + 0.5 * a[3] * t4[c] +
               0.5 * a[3] * t5[c];
}

template <class T>
__device__ inline void wall_time_coeffs_t(double dt, T tau, T mt[5]) {
    T x = dt / tau;
    if (tval(x) < 0.1) {
        T t(1.0);
        for (int a = 0; a < 5; ++a) mt[a] = T(0.0);
        for (int n = 1; n <= 24; ++n) {
            t = t * (-x) / (double)n;
            mt[3] = mt[3] - tau * t;
            if (n >= 2) {
                mt[0] = mt[0] + tau * t;
                mt[4] = mt[4] + tau * tau * (double)(n - 1) * t;
            }
            if (n >= 3) {
                mt[1] = mt[1] + tau * tau * (double)(n - 2) * t;
                mt[2] = mt[2] - tau * tau * t;
            }
        }
    } else {
        T ex = texp(-x);
        mt[3] = tau * (1.0 - ex);
        mt[4] = -tau * dt * ex + tau * mt[3];
        mt[0] = dt - mt[3];
        mt[1] = -tau * mt[0] + mt[4];
        mt[2] = 0.5 * dt * dt - tau * mt[0];
    }
}

// Scalars of one wall face node (k_wall_flux) as functions of the wall
// cell's W, the wall speed, mu_ref, Pr and the node's six velocity sums red
// (tangential slope moments swt and the Shakhov heat moments hl).
enum { WC_MT = 0, WC_AN = 5, WC_AT = 9, WC_ATT = 13, WC_COEFF = 17, WC_MXI2 = 18, WC_IRG = 19, WC_WW = 20,
       WC_N = 24 };
enum { WI_WC = 0, WI_UX = 4, WI_MU = 5, WI_PR = 6, WI_RED = 7, WI_N = 13 };
template <class T>
__device__ inline void wall_chain(int ck, double omega, double dt, double dn, double side, double theta,
                                  const T in[WI_N], T out[WC_N]) {
    const T* Wc = in + WI_WC;
    const T ux = in[WI_UX], mu_ref = in[WI_MU], pr = in[WI_PR];
    const T* red = in + WI_RED;
    T p = (2.0 / (double)(ck + 2)) * (Wc[3] - 0.5 * (Wc[1] * Wc[1] + Wc[2] * Wc[2]) / Wc[0]);
    const double lambda = 0.5 / theta;
    T rg = p / theta;
    T prim[4] = {rg, T(0.0), -ux, T(lambda)};
    T wc[4] = {Wc[0], Wc[2], -Wc[1], Wc[3]};
    T wg[4] = {rg, T(0.0), -rg * ux, rg * ((double)(ck + 2) / (4.0 * lambda) + 0.5 * ux * ux)};
    T sw[4], an[4];
    for (int a = 0; a < 4; ++a) sw[a] = (wg[a] - wc[a]) / (side * 0.5 * dn);
    micro_slope_t(ck, prim, sw, an);
    T mu[7], mv[6], mxi[3];
    half_moments_t(ck, prim, mu, mv, mxi);
    T wtau[4] = {wg[0], -wg[2], wg[1], wg[3]};
    T tau_w = tau_t(ck, omega, wtau, mu_ref);
    T mt[5];
    wall_time_coeffs_t(dt, tau_w, mt);
    T coeff = 0.8 * (1.0 - pr) * lambda * lambda / rg;
    T rr[4] = {red[0], red[1], red[2], red[3]};
    T at[4], att[4], m1[4], m2[4], sw2[4];
    micro_slope_t(ck, prim, rr, at);
    m_auvxi_t(an, mu, mv, mxi, 1, 0, m1);
    m_auvxi_t(at, mu, mv, mxi, 0, 1, m2);
    for (int q = 0; q < 4; ++q) sw2[q] = -rg * (m1[q] + m2[q]);
    micro_slope_t(ck, prim, sw2, att);
    for (int a = 0; a < 5; ++a) out[WC_MT + a] = mt[a];
    for (int a = 0; a < 4; ++a) {
        out[WC_AN + a] = an[a];
        out[WC_AT + a] = at[a];
        out[WC_ATT + a] = att[a];
    }
    out[WC_COEFF] = coeff;
    out[WC_MXI2] = mxi[2];
    out[WC_IRG] = 1.0 / rg;
    out[WC_WW + 0] = rg;
    out[WC_WW + 1] = rg * ux;
    out[WC_WW + 2] = T(0.0);
    out[WC_WW + 3] = rg * ((double)(ck + 2) * theta / 2.0 + ux * ux / 2.0);
}

// ---------------------------------------------------------------------------
// small per-node linear algebra
// ---------------------------------------------------------------------------
// moment adjoint of one velocity: the h and b weights of g . moments
__device__ inline double mom_adj_h(const double g[4], double w, double u, double v) {
    return w * (g[0] + g[1] * u + g[2] * v + 0.5 * g[3] * (u * u + v * v));
}

// J^T x = rhs for J column-major (J[c*4 + r] = J_rc)
__device__ inline bool solve4T(const double* J, const double* rhs, double* x) {
    double t[16];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) t[c * 4 + r] = J[r * 4 + c];
    return solve4(t, rhs, x);
}

// gJ[r][a] -= g[r] * td[a] (gJ column-major like J)
__device__ inline void outer_sub(double gJ[16], const double g[4], const double td[4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) gJ[a * 4 + r] -= g[r] * td[a];
}

// Reverse of q(theta), qb(theta) and their theta-derivatives qt_a, qbt_a for
// one velocity: accumulates into gth the theta-gradient of
//   g_q q + g_qb qb + sum_a (g_qt[a] qt_a + g_qbt[a] qbt_a).
// q = rho/(pi T) exp(-c^2/T), qb = ck/2 T q, qt_a = q l_a, qbt_a = qb lb_a
// with l = (1/rho, 2du/T, 2dv/T, -1/T + c^2/T^2), lb = l + (0, 0, 0, 1/T).
__device__ inline void qtheta_adj(int ck, const double th[4], double i0, double iT, double u, double v, double q,
                                  double g_q, double g_qb, const double g_qt[4], const double g_qbt[4],
                                  double gth[4]) {
    const double du = u - th[1], dv = v - th[2], c2 = du * du + dv * dv;
    const double qb = 0.5 * (double)ck * th[3] * q;
    const double iT2 = iT * iT;
    const double l[4] = {i0, 2.0 * du * iT, 2.0 * dv * iT, -iT + c2 * iT2};
    const double lb[4] = {l[0], l[1], l[2], l[3] + iT};
    double A = 0.0, B = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) { A += g_qt[a] * l[a]; B += g_qbt[a] * lb[a]; }
    const double ch = g_q * q + q * A, cb = g_qb * qb + qb * B;
    double e[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) e[a] = q * g_qt[a] + qb * g_qbt[a];
    gth[0] += ch * l[0] + cb * lb[0] - e[0] * i0 * i0;
    gth[1] += ch * l[1] + cb * lb[1] - 2.0 * iT * e[1] - 2.0 * du * iT2 * e[3];
    gth[2] += ch * l[2] + cb * lb[2] - 2.0 * iT * e[2] - 2.0 * dv * iT2 * e[3];
    gth[3] += ch * l[3] + cb * lb[3] - 2.0 * du * iT2 * e[1] - 2.0 * dv * iT2 * e[2] +
              (iT2 - 2.0 * c2 * iT2 * iT) * e[3] - qb * g_qbt[3] * iT2;
}

// ---------------------------------------------------------------------------
// WENO5-AO and the wall quartic, reversed
// ---------------------------------------------------------------------------
// reverse of weno_weights: gq += (d fw / d q)^T gfw
__device__ inline void weno_weights_adj(const double q[5], const double gfw[4], double gq[5]) {
    const double g0 = 0.01125, g1 = 0.1275, g2 = 0.01125, g3 = 0.85, eps = 1.0e-8;
    const double t0 = q[0] - 2.0 * q[1] + q[2], s0 = q[0] - 4.0 * q[1] + 3.0 * q[2];
    const double t1 = q[1] - 2.0 * q[2] + q[3], s1 = q[1] - q[3];
    const double t2 = q[2] - 2.0 * q[3] + q[4], s2 = 3.0 * q[2] - 4.0 * q[3] + q[4];
    const double b0 = 13.0 / 12.0 * t0 * t0 + 0.25 * s0 * s0;
    const double b1 = 13.0 / 12.0 * t1 * t1 + 0.25 * s1 * s1;
    const double b2 = 13.0 / 12.0 * t2 * t2 + 0.25 * s2 * s2;
    const double b3 = (231153.0 * q[2] * q[2] + 104963.0 * q[1] * q[1] + 6908.0 * q[0] * q[0] - 38947.0 * q[0] * q[3] +
                       104963.0 * q[3] * q[3] + q[1] * (-51001.0 * q[0] + 179098.0 * q[3] - 38947.0 * q[4]) -
                       3.0 * q[2] * (99692.0 * q[1] - 22641.0 * q[0] + 99692.0 * q[3] - 22641.0 * q[4]) +
                       8209.0 * q[0] * q[4] - 51001.0 * q[3] * q[4] + 6908.0 * q[4] * q[4]) /
                      5040.0;
    const double d0 = b3 - b0, d1 = b3 - b1, d2 = b3 - b2;
    const double tau = (fabs(d0) + fabs(d1) + fabs(d2)) / 3.0;
    const double e0 = b0 + eps, e1 = b1 + eps, e2 = b2 + eps, e3 = b3 + eps;
    const double r0 = tau / e0, r1 = tau / e1, r2 = tau / e2, r3 = tau / e3;
    const double a0 = g0 * (1.0 + r0 * r0), a1 = g1 * (1.0 + r1 * r1), a2 = g2 * (1.0 + r2 * r2),
                 a3 = g3 * (1.0 + r3 * r3);
    const double S = a0 + a1 + a2 + a3, iS = 1.0 / S;
    const double o0 = a0 * iS, o1 = a1 * iS, o2 = a2 * iS, o3 = a3 * iS;
    const double go0 = gfw[0], go1 = gfw[1], go2 = gfw[2];
    const double go3 = (gfw[3] - g0 * gfw[0] - g1 * gfw[1] - g2 * gfw[2]) / g3;
    const double dot = go0 * o0 + go1 * o1 + go2 * o2 + go3 * o3;
    const double ga0 = (go0 - dot) * iS, ga1 = (go1 - dot) * iS, ga2 = (go2 - dot) * iS, ga3 = (go3 - dot) * iS;
    const double gr0 = ga0 * g0 * 2.0 * r0, gr1 = ga1 * g1 * 2.0 * r1, gr2 = ga2 * g2 * 2.0 * r2,
                 gr3 = ga3 * g3 * 2.0 * r3;
    const double gtau = gr0 / e0 + gr1 / e1 + gr2 / e2 + gr3 / e3;
    double gb0 = -gr0 * r0 / e0, gb1 = -gr1 * r1 / e1, gb2 = -gr2 * r2 / e2, gb3 = -gr3 * r3 / e3;
    const double sg0 = (d0 > 0.0) - (d0 < 0.0), sg1 = (d1 > 0.0) - (d1 < 0.0), sg2 = (d2 > 0.0) - (d2 < 0.0);
    const double gt3 = gtau / 3.0;
    gb3 += gt3 * (sg0 + sg1 + sg2);
    gb0 -= gt3 * sg0;
    gb1 -= gt3 * sg1;
    gb2 -= gt3 * sg2;
    const double gt0 = gb0 * (13.0 / 6.0) * t0, gs0 = gb0 * 0.5 * s0;
    const double gt1 = gb1 * (13.0 / 6.0) * t1, gs1 = gb1 * 0.5 * s1;
    const double gt2 = gb2 * (13.0 / 6.0) * t2, gs2 = gb2 * 0.5 * s2;
    const double c = gb3 / 5040.0;
    gq[0] += gt0 + gs0 + c * (13816.0 * q[0] - 51001.0 * q[1] + 67923.0 * q[2] - 38947.0 * q[3] + 8209.0 * q[4]);
    gq[1] += -2.0 * gt0 - 4.0 * gs0 + gt1 + gs1 +
             c * (-51001.0 * q[0] + 209926.0 * q[1] - 299076.0 * q[2] + 179098.0 * q[3] - 38947.0 * q[4]);
    gq[2] += gt0 + 3.0 * gs0 - 2.0 * gt1 + gt2 + 3.0 * gs2 +
             c * (67923.0 * q[0] - 299076.0 * q[1] + 462306.0 * q[2] - 299076.0 * q[3] + 67923.0 * q[4]);
    gq[3] += gt1 - gs1 - 2.0 * gt2 - 4.0 * gs2 +
             c * (-38947.0 * q[0] + 179098.0 * q[1] - 299076.0 * q[2] + 209926.0 * q[3] - 51001.0 * q[4]);
    gq[4] += gt2 + gs2 + c * (8209.0 * q[0] - 38947.0 * q[1] + 67923.0 * q[2] - 51001.0 * q[3] + 13816.0 * q[4]);
}

// reverse of weno_polys (linear): gq += maps^T (gcq, gc4)
__device__ inline void weno_polys_adj(const Consts& K, const double gcq[3][3], const double gc4[5], double gq[5]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int t = c; t < c + 3; ++t) {
            double s = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) s += K.cmap[c][t][d] * gcq[c][d];
            gq[t] += s;
        }
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 5; ++d) s += K.cmap[3][t][d] * gc4[d];
        gq[t] += s;
    }
}

// reverse of weno_at at xi for value adjoint gv and derivative adjoint gd
__device__ inline void weno_at_adj(const double fw[4], const double cq[3][3], const double c4[5], double xi, double gv,
                                   double gd, double gfw[4], double gcq[3][3], double gc4[5]) {
    const double x2 = xi * xi, x3 = x2 * xi, x4 = x3 * xi;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double P = cq[c][0] + xi * cq[c][1] + x2 * cq[c][2], dP = cq[c][1] + 2.0 * xi * cq[c][2];
        gfw[c] += gv * P + gd * dP;
        gcq[c][0] += fw[c] * gv;
        gcq[c][1] += fw[c] * (gv * xi + gd);
        gcq[c][2] += fw[c] * (gv * x2 + gd * 2.0 * xi);
    }
    const double P = c4[0] + xi * c4[1] + x2 * c4[2] + x3 * c4[3] + x4 * c4[4];
    const double dP = c4[1] + 2.0 * xi * c4[2] + 3.0 * x2 * c4[3] + 4.0 * x3 * c4[4];
    gfw[3] += gv * P + gd * dP;
    gc4[0] += fw[3] * gv;
    gc4[1] += fw[3] * (gv * xi + gd);
    gc4[2] += fw[3] * (gv * x2 + gd * 2.0 * xi);
    gc4[3] += fw[3] * (gv * x3 + gd * 3.0 * x2);
    gc4[4] += fw[3] * (gv * x4 + gd * 4.0 * x3);
}

// reverse of wall_at (linear in the samples)
__device__ inline void wall_at_adj(const Consts& K, int sh, double xi, double gv, double gd, double gq[5]) {
    const double x2 = xi * xi, x3 = x2 * xi, x4 = x3 * xi;
    const double e[5] = {gv, gv * xi + gd, gv * x2 + gd * 2.0 * xi, gv * x3 + gd * 3.0 * x2, gv * x4 + gd * 4.0 * x3};
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 5; ++d) s += K.wallmap[sh][t][d] * e[d];
        gq[t] += s;
    }
}

// reverse of one WENO5-AO line evaluated at NP points (value adjoint gv[p],
// derivative adjoint gd[p])
template <int NP>
__device__ inline void weno_line_adj(const Consts& K, const double q[5], const double xi[NP], const double gv[NP],
                                     const double gd[NP], double gq[5]) {
    double fw[4], cq[3][3], c4[5];
    weno_weights(q, fw);
    weno_polys(K, q, cq, c4);
    double gfw[4] = {0, 0, 0, 0}, gcq[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gc4[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int p = 0; p < NP; ++p) weno_at_adj(fw, cq, c4, xi[p], gv[p], gd[p], gfw, gcq, gc4);
    weno_polys_adj(K, gcq, gc4, gq);
    weno_weights_adj(q, gfw, gq);
}

// reverse of companion3: g[a] = {value, companion, derivative} adjoints
__device__ inline void companion3_adj(const Consts& K, const double q[5], const double comp[5], const double g[3][3],
                                      double gq[5], double gc[5]) {
    double fw[4], cq[3][3], c4[5], kq[3][3], k4[5];
    weno_weights(q, fw);
    weno_polys(K, q, cq, c4);
    weno_polys(K, comp, kq, k4);
    double gfw[4] = {0, 0, 0, 0}, gcq[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gc4[5] = {0, 0, 0, 0, 0};
    double gkq[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gk4[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        weno_at_adj(fw, cq, c4, K.gnode[a], g[a][0], g[a][2], gfw, gcq, gc4);
        weno_at_adj(fw, kq, k4, K.gnode[a], g[a][1], 0.0, gfw, gkq, gk4);
    }
    weno_polys_adj(K, gcq, gc4, gq);
    weno_weights_adj(q, gfw, gq);
    weno_polys_adj(K, gkq, gk4, gc);
}

// reverse of side_data for side s into the first-layer adjoint gL
__device__ inline void side_adj(const Consts& K, const FirstLayer& L, bool near, int sh, int s, const double g[3][3],
                                FirstLayer& gL) {
    double q[5], c[5], gq[5] = {0, 0, 0, 0, 0}, gc[5] = {0, 0, 0, 0, 0};
    if (s < 2) {
#pragma unroll
        for (int r = 0; r < 5; ++r) { q[r] = L.xv[r][s]; c[r] = L.xd[r][s]; }
        if (near) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                wall_at_adj(K, sh, K.gnode[a], g[a][0], g[a][2], gq);
                wall_at_adj(K, sh, K.gnode[a], g[a][1], 0.0, gc);
            }
        } else {
            companion3_adj(K, q, c, g, gq, gc);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) { gL.xv[r][s] += gq[r]; gL.xd[r][s] += gc[r]; }
    } else {
#pragma unroll
        for (int r = 0; r < 5; ++r) { q[r] = L.yv[r][s - 2]; c[r] = L.yd[r][s - 2]; }
        companion3_adj(K, q, c, g, gq, gc);
#pragma unroll
        for (int r = 0; r < 5; ++r) { gL.yv[r][s - 2] += gq[r]; gL.yd[r][s - 2] += gc[r]; }
    }
}

// reverse of vol_data into gL
__device__ inline void vol_adj(const Consts& K, const FirstLayer& L, bool near, int sh, const double gvol[3][3],
                               FirstLayer& gL) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double q[5], gq[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 5; ++r) q[r] = L.rows[r][a];
        if (near) {
#pragma unroll
            for (int b = 0; b < 3; ++b) wall_at_adj(K, sh, K.gnode[b], gvol[a][b], 0.0, gq);
        } else {
            const double zero[3] = {0, 0, 0};
            weno_line_adj<3>(K, q, K.gnode, gvol[a], zero, gq);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) gL.rows[r][a] += gq[r];
    }
}

// reverse of first_layer: the first-layer adjoint gL onto the 5x5 block
__device__ inline void first_layer_adj(const Consts& K, const double (&blk)[5][5], bool near, int sh,
                                       const FirstLayer& gL, double (&gblk)[5][5]) {
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const double xi[5] = {-0.5, 0.5, K.gnode[0], K.gnode[1], K.gnode[2]};
        const double gv[5] = {gL.xv[r][0], gL.xv[r][1], gL.rows[r][0], gL.rows[r][1], gL.rows[r][2]};
        const double gd[5] = {gL.xd[r][0], gL.xd[r][1], 0.0, 0.0, 0.0};
        weno_line_adj<5>(K, blk[r], xi, gv, gd, gblk[r]);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double q[5], gq[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 5; ++r) q[r] = blk[r][c];
        if (near) {
            wall_at_adj(K, sh, -0.5, gL.yv[c][0], gL.yd[c][0], gq);
            wall_at_adj(K, sh, 0.5, gL.yv[c][1], gL.yd[c][1], gq);
        } else {
            const double xi[2] = {-0.5, 0.5};
            const double gv[2] = {gL.yv[c][0], gL.yv[c][1]}, gd[2] = {gL.yd[c][0], gL.yd[c][1]};
            weno_line_adj<2>(K, q, xi, gv, gd, gq);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) gblk[r][c] += gq[r];
    }
}

__device__ inline void zero_layer(FirstLayer& L) {
#pragma unroll
    for (int r = 0; r < 5; ++r) {
#pragma unroll
        for (int s = 0; s < 2; ++s) { L.xv[r][s] = 0.0; L.xd[r][s] = 0.0; L.yv[r][s] = 0.0; L.yd[r][s] = 0.0; }
#pragma unroll
        for (int a = 0; a < 3; ++a) L.rows[r][a] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// table forms: the candidate polynomials' values and derivatives at the
// points p (0: -1/2, 1: +1/2, 2-4: Gauss nodes) as weights on the samples
// ---------------------------------------------------------------------------
// reverse of one WENO5-AO line at NP table points: gq += d/dq of
// sum_p (gv[p] value_p + gd[p] derivative_p)
template <int NP>
__device__ inline void weno_pts_adj(const Consts& K, const double q[5], const int (&pts)[NP], const double (&gv)[NP],
                                    const double (&gd)[NP], double gq[5]) {
    double fw[4], gfw[4];
    weno_weights(q, fw);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double A[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int t = 0; t < 5; ++t) A[t] += gv[p] * K.pv[pts[p]][c][t] + gd[p] * K.pd[pts[p]][c][t];
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 5; ++t) { s += A[t] * q[t]; gq[t] += fw[c] * A[t]; }
        gfw[c] = s;
    }
    weno_weights_adj(q, gfw, gq);
}

// reverse of companion3 (g[a] = {value, companion, derivative} adjoints)
__device__ inline void comp_adj(const Consts& K, const double q[5], const double cmp[5], const double g[3][3],
                                double gq[5], double gc[5]) {
    double fw[4], gfw[4];
    weno_weights(q, fw);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double A[5] = {0, 0, 0, 0, 0}, B[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                A[t] += g[a][0] * K.pv[2 + a][c][t] + g[a][2] * K.pd[2 + a][c][t];
                B[t] += g[a][1] * K.pv[2 + a][c][t];
            }
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            s += A[t] * q[t] + B[t] * cmp[t];
            gq[t] += fw[c] * A[t];
            gc[t] += fw[c] * B[t];
        }
        gfw[c] = s;
    }
    weno_weights_adj(q, gfw, gq);
}

// reverse of the near-wall x-side data (wall quartics of the samples and of
// the companion samples at the Gauss nodes)
__device__ inline void comp_wall_adj(const Consts& K, int sh, const double g[3][3], double gq[5], double gc[5]) {
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            a0 += g[a][0] * K.wv[sh][2 + a][t] + g[a][2] * K.wd[sh][2 + a][t];
            a1 += g[a][1] * K.wv[sh][2 + a][t];
        }
        gq[t] += a0;
        gc[t] += a1;
    }
}

// positivity factor of one field with the candidate that sets it: arg
// s*3+a for a face node value, 12+a*3+b for a volume node, -1 when none
__device__ inline double theta_arg(double m, const double fval[4][3], const double vol[3][3], int* arg) {
    const double keep = 1.0 - 64.0 * UGKS_DBL_EPS;
    double theta = 1.0;
    int g = -1;
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (fval[s][a] < 0.0) {
                const double t = keep * m / (m - fval[s][a]);
                if (t < theta) { theta = t; g = s * 3 + a; }
            }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            if (vol[a][b] < 0.0) {
                const double t = keep * m / (m - vol[a][b]);
                if (t < theta) { theta = t; g = 12 + a * 3 + b; }
            }
    *arg = g;
    return theta;
}
