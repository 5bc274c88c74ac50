// Low-order Explicit entry of UGKS2D on gfx950: Van Leer slopes, the
// analytic-moment UGKS face flux with the projected face equilibrium and its
// Shakhov part, the multiscale (W2B) diffuse walls and the trapezoidal
// Shakhov collision update.  Restates, in the reference:
//   Interp_VanLeer, one-sided wall slopes     src/mod_Interpolation.f90:9-58,142-171
//   CalcFlux_MultiDimension                   src/mod_Evolution.f90:186-492
//   W2BMultiscaleFluxBoundary                 src/mod_Evolution.f90:1888-2449
//   PrepareNativeFaceEquilibrium              src/mod_Evolution.f90:16-30
//   ProjectToConservedMoments                 src/mod_ImplicitCommon.f90:42-113
//   ApplyExplicitUpdate (+ Write_residuals)   src/mod_TimeIntegrator.f90:16-134, src/mod_Solve.f90:427-461
//   InitFlow_Freestream                       src/mod_InitializeFlow.f90:826-856
//
// Launch shape: one workgroup per face or cell; thread t loops over the
// velocities k = t, t + NT, ...  Nothing per velocity is held in registers
// between the velocity reductions: each pass re-reads its stencil, which the
// first pass of the workgroup has brought into cache, and the per-face
// constants live in scalar registers.  Registers stay near 128 per lane, so
// several workgroups share a CU and hide each other's serial sections.
//
// The Maxwellian of a frame factorises on the tensor velocity grid:
// exp(-lam((u-a)^2 + (v-c)^2)) = exp(-lam(u-a)^2) exp(-lam(v-c)^2), so a
// workgroup evaluates Nu + Nv exponentials into LDS tables instead of Nu*Nv.
// The product differs from the single exponential by a few ulp.
#include "ugks_state.h"
#include "ugks_prof.h"

namespace {

constexpr double LO_TINY = 1.17549435082228750797e-38;   // tiny(0.0_real32), the limiter guard
constexpr int LO_MAXAX = 64;                              // largest Nu, Nv of the tables

// Interp_VanLeer with the cell-size divisions as reciprocal products
__device__ inline double vanleer(double fl, double fc, double fr, double idl, double idr) {
    double sl = (fc - fl) * idl, sr = (fr - fc) * idr;
    return (copysign(1.0, sr) + copysign(1.0, sl)) * fabs(sr) * fabs(sl) / (fabs(sr) + fabs(sl) + LO_TINY);
}

// GetConserved (prim = rho, U, V, lambda)
__device__ inline void conserved(double gamma, const double prim[4], double w[4]) {
    w[0] = prim[0];
    w[1] = prim[0] * prim[1];
    w[2] = prim[0] * prim[2];
    w[3] = 0.5 * prim[0] / (prim[3] * (gamma - 1.0)) + 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]);
}

__device__ inline void lo_micro_slope(int ck, const double prim[4], const double sw[4], double a[4]) {
    a[3] = 4.0 * prim[3] * prim[3] / ((ck + 2) * prim[0]) *
           (2.0 * sw[3] - 2.0 * prim[1] * sw[1] - 2.0 * prim[2] * sw[2] +
            sw[0] * (prim[1] * prim[1] + prim[2] * prim[2] - 0.5 * (ck + 2) / prim[3]));
    a[2] = 2.0 * prim[3] / prim[0] * (sw[2] - prim[2] * sw[0]) - prim[2] * a[3];
    a[1] = 2.0 * prim[3] / prim[0] * (sw[1] - prim[1] * sw[0]) - prim[1] * a[3];
    a[0] = sw[0] / prim[0] - prim[1] * a[1] - prim[2] * a[2] -
           0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + 0.5 * (ck + 2) / prim[3]) * a[3];
}

// Moments of the Maxwellian over the half ranges u>0 (mur) and u<0 (mul) and
// in v and xi (CalcMoment, src/mod_Evolution.f90:1810-1842)
__device__ inline void lo_moments(int ck, const double prim[4], double mu[7], double mul[7], double mur[7],
                                  double mv[6], double mxi[3]) {
    const double sl = sqrt(prim[3]), g = 0.5 * exp(-prim[3] * prim[1] * prim[1]) / sqrt(UGKS_PI * prim[3]);
    const double il = 1.0 / prim[3];
    mul[0] = 0.5 * erfc(-sl * prim[1]);
    mul[1] = prim[1] * mul[0] + g;
    mur[0] = 0.5 * erfc(sl * prim[1]);
    mur[1] = prim[1] * mur[0] - g;
#pragma unroll
    for (int i = 2; i <= 6; ++i) {
        mul[i] = prim[1] * mul[i - 1] + 0.5 * (i - 1) * mul[i - 2] * il;
        mur[i] = prim[1] * mur[i - 1] + 0.5 * (i - 1) * mur[i - 2] * il;
    }
#pragma unroll
    for (int i = 0; i <= 6; ++i) mu[i] = mul[i] + mur[i];
    mv[0] = 1.0;
    mv[1] = prim[2];
#pragma unroll
    for (int i = 2; i <= 5; ++i) mv[i] = prim[2] * mv[i - 1] + 0.5 * (i - 1) * mv[i - 2] * il;
    mxi[0] = 1.0;
    mxi[1] = 0.5 * ck * il;
    mxi[2] = ck * (ck + 2.0) * 0.25 * il * il;
}

__device__ inline void lo_uvxi(const double mu[7], const double mv[6], const double mxi[3], int al, int be, int de,
                               double o[4]) {
    o[0] = mu[al] * mv[be] * mxi[de / 2];
    o[1] = mu[al + 1] * mv[be] * mxi[de / 2];
    o[2] = mu[al] * mv[be + 1] * mxi[de / 2];
    o[3] = 0.5 * (mu[al + 2] * mv[be] * mxi[de / 2] + mu[al] * mv[be + 2] * mxi[de / 2] +
                  mu[al] * mv[be] * mxi[(de + 2) / 2]);
}

// Moment_auvxi: <a . psi u^al v^be psi>
__device__ inline void lo_auvxi(const double a[4], const double mu[7], const double mv[6], const double mxi[3], int al,
                                int be, double o[4]) {
    double t0[4], t1[4], t2[4], t3[4], t4[4], t5[4];
    lo_uvxi(mu, mv, mxi, al, be, 0, t0);
    lo_uvxi(mu, mv, mxi, al + 1, be, 0, t1);
    lo_uvxi(mu, mv, mxi, al, be + 1, 0, t2);
    lo_uvxi(mu, mv, mxi, al + 2, be, 0, t3);
    lo_uvxi(mu, mv, mxi, al, be + 2, 0, t4);
    lo_uvxi(mu, mv, mxi, al, be, 2, t5);
#pragma unroll
    for (int c = 0; c < 4; ++c)
        o[c] = a[0] * t0[c] + a[1] * t1[c] + a[2] * t2[c] + 0.5 * a[3] * t3[c] + 0.5 * a[3] * t4[c] +
               0.5 * a[3] * t5[c];
}

// time integrals of exp(-t/tau) over the step (CalcFlux_MultiDimension, global dt)
__device__ inline void lo_time_coeffs(double dt, double tau, double Mt[5]) {
    const double ex = exp(-dt / tau);
    Mt[3] = tau * (1.0 - ex);
    Mt[4] = -tau * dt * ex + tau * Mt[3];
    Mt[0] = dt - Mt[3];
    Mt[1] = -tau * Mt[0] + Mt[4];
    Mt[2] = 0.5 * dt * dt - tau * Mt[0];
}

// Gaussian factor tables of a Maxwellian with inverse temperature lam in the
// frame whose velocities are (vn, vt) = (su*u, v) (x frame, su = +1) or
// (v, -u) (y frame): tu[iu] = exp(-lam (u - au)^2), tv[jv] = exp(-lam (v - av)^2)
// with (au, av) chosen so that (vn - U)^2 + (vt - V)^2 = (u - au)^2 + (v - av)^2.
// Also the maxima of both tables.  Ends with a barrier.
template <int NT>
__device__ inline void gauss_tables(const Dev& D, double lam, double au, double av, double* tu, double* tv,
                                    double* tmax, double* lds) {
    const int nu = D.nu, nv = D.nv, a = threadIdx.x;   // NT >= nu + nv
    double m[2] = {0.0, 0.0};
    if (a < nu) {
        const double d = D.u[a] - au;
        tu[a] = m[0] = exp(-lam * (d * d));
    } else if (a < nu + nv) {
        const double d = D.v[(a - nu) * nu] - av;
        tv[a - nu] = m[1] = exp(-lam * (d * d));
    }
    block_max<2>(m, lds);
    if (threadIdx.x == 0) { tmax[0] = m[0]; tmax[1] = m[1]; }
    __syncthreads();
}

// entries (a,c), a <= c, of the symmetric projection Gram matrix for one velocity
__device__ inline void gram_add(double g[10], double t, double tb, double vn, double vt) {
    const double e = 0.5 * (vn * vn + vt * vt);
    g[0] += t;
    g[1] += t * vn;
    g[2] += t * vt;
    g[3] += t * e;
    g[4] += t * vn * vn;
    g[5] += t * vn * vt;
    g[6] += t * vn * e;
    g[7] += t * vt * vt;
    g[8] += t * vt * e;
    g[9] += t * e * e + tb * 0.25;
}

__device__ inline void gram_full(const double g[10], double m[16]) {
    const int idx[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) m[c * 4 + a] = g[idx[a][c]];
}

// moment gate of ProjectToConservedMoments: max |m - target| <= 1e-10 max(1, |target|)
__device__ inline bool project_ok(const double mom[4], const double target[4]) {
    double rs = 1.0, e = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        rs = fmax(rs, fabs(target[a]));
        e = fmax(e, fabs(mom[a] - target[a]));
    }
    return e <= 1.0e-10 * rs;
}

__device__ inline void flag(const Dev& D, unsigned bits) { atomicOr(D.status, bits); }


}  // namespace

// Upwind reconstruction at an interior face for velocity k: the face values
// of h and b, their normal slopes and their tangential slopes in the face
// frame (out: h, b, dh/dn, db/dn, dh/dt, db/dt).  x face between (i,j) and
// (i+1,j); y face between (i,j) and (i,j+1).  Van Leer slopes, x periodic,
// one-sided y differences on the wall rows.
__device__ inline void face_data(const Dev& D, bool xf, int i, int j, int k, bool left, double idx_, double idy,
                                 double len_n, double out[6]) {
    const int nx = D.nx, ny = D.ny;
    const size_t nvel = D.nvel, row = (size_t)nx * nvel;
    double hc, bc, nh, nb, th, tb;
    if (xf) {
        const int ic = left ? i : (i + 1 == nx ? 0 : i + 1);
        const int iw = ic == 0 ? nx - 1 : ic - 1, ie = ic + 1 == nx ? 0 : ic + 1;
        const size_t o = ((size_t)j * nx + ic) * nvel + k;
        const size_t ow = ((size_t)j * nx + iw) * nvel + k, oe = ((size_t)j * nx + ie) * nvel + k;
        const size_t os = j > 0 ? o - row : o, on = j < ny - 1 ? o + row : o;
        hc = D.h[o]; bc = D.b[o];
        nh = vanleer(D.h[ow], hc, D.h[oe], idx_, idx_);
        nb = vanleer(D.b[ow], bc, D.b[oe], idx_, idx_);
        const double hs = D.h[os], hn = D.h[on], bs = D.b[os], bn = D.b[on];
        const bool edge = j == 0 || j == ny - 1;
        th = edge ? (hn - hs) * idy : vanleer(hs, hc, hn, idy, idy);
        tb = edge ? (bn - bs) * idy : vanleer(bs, bc, bn, idy, idy);
    } else {
        const int jc = left ? j : j + 1;
        const int iw = i == 0 ? nx - 1 : i - 1, ie = i + 1 == nx ? 0 : i + 1;
        const size_t o = ((size_t)jc * nx + i) * nvel + k;
        const size_t os = jc > 0 ? o - row : o, on = jc < ny - 1 ? o + row : o;
        const size_t ow = ((size_t)jc * nx + iw) * nvel + k, oe = ((size_t)jc * nx + ie) * nvel + k;
        hc = D.h[o]; bc = D.b[o];
        const double hs = D.h[os], hn = D.h[on], bs = D.b[os], bn = D.b[on];
        const bool edge = jc == 0 || jc == ny - 1;
        nh = edge ? (hn - hs) * idy : vanleer(hs, hc, hn, idy, idy);
        nb = edge ? (bn - bs) * idy : vanleer(bs, bc, bn, idy, idy);
        th = -vanleer(D.h[ow], hc, D.h[oe], idx_, idx_);
        tb = -vanleer(D.b[ow], bc, D.b[oe], idx_, idx_);
    }
    const double half = left ? 0.5 * len_n : -0.5 * len_n;
    out[0] = hc + half * nh;
    out[1] = bc + half * nb;
    out[2] = nh; out[3] = nb; out[4] = th; out[5] = tb;
}

// ===========================================================================
// interior faces: xf, block f is the east face of cell f (nx*ny blocks);
// otherwise block f is the face between cell f and the cell above it
// (nx*(ny-1) blocks)
// ===========================================================================
template <int NT, bool xf>
__global__ void __launch_bounds__(NT) k_lo_face(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[(NT / 16) * 14];
    __shared__ double tu[LO_MAXAX], tv[LO_MAXAX], tmax[2];
    const int nx = D.nx, ny = D.ny, nu = D.nu, nvel = D.nvel;
    const int f = xcd_contiguous(blockIdx.x, gridDim.x);
    const int g = f;                                // cell below/west of the face
    const int i = g % nx, j = g / nx;
    const double dt = D.scal[SC_DT];
    const double idx_ = 1.0 / D.dx, idy = 1.0 / D.dy;
    const double len_n = xf ? D.dx : D.dy, flen = xf ? D.dy : D.dx;
    const int ck = D.gas.ck;

    // ---- reconstruction of the upwind side, and the moment sums ----------
    // Nothing per velocity is held across the reductions: each pass
    // re-reads the stencil, which the first pass has brought into cache.
    double s1[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // cv(4), swt(4)
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = xf ? u : v, vt = xf ? v : -u;
        double fd[6];
        face_data(D, xf, i, j, k, !signbit(vn), idx_, idy, len_n, fd);
        const double v2 = vn * vn + vt * vt;
        s1[0] += w * fd[0]; s1[1] += w * vn * fd[0]; s1[2] += w * vt * fd[0]; s1[3] += 0.5 * w * (v2 * fd[0] + fd[1]);
        s1[4] += w * fd[4]; s1[5] += w * vn * fd[4]; s1[6] += w * vt * fd[4]; s1[7] += 0.5 * w * (v2 * fd[4] + fd[5]);
    }
    block_sum<8>(s1, lds);

    // ---- face state, slopes and time integrals (uniform) ------------------
    const double* WL = D.W + 4 * (size_t)g;
    const double* WR = D.W + 4 * (size_t)(xf ? (j * nx + (i + 1 == nx ? 0 : i + 1)) : g + nx);
    const double wL[4] = {WL[0], xf ? WL[1] : WL[2], xf ? WL[2] : -WL[1], WL[3]};
    const double wR[4] = {WR[0], xf ? WR[1] : WR[2], xf ? WR[2] : -WR[1], WR[3]};
    const double cv[4] = {s1[0], s1[1], s1[2], s1[3]};
    if (!physical_w(cv)) {
        if (threadIdx.x == 0) flag(D, ST_NONFINITE | ST_AT_FACE);
        return;
    }
    double prim[4], sw[4], aL[4], aR[4], bs[4], aT[4];
    primary(D.gas.gamma, cv, prim);
    const double i2 = 2.0 / len_n;
    for (int m = 0; m < 4; ++m) sw[m] = (cv[m] - wL[m]) * i2;
    lo_micro_slope(ck, prim, sw, aL);
    for (int m = 0; m < 4; ++m) sw[m] = (wR[m] - cv[m]) * i2;
    lo_micro_slope(ck, prim, sw, aR);
    const double swt[4] = {s1[4], s1[5], s1[6], s1[7]};
    lo_micro_slope(ck, prim, swt, bs);
    double mu[7], mul[7], mur[7], mv[6], mxi[3], mauL[4], mauR[4], mbv[4];
    lo_moments(ck, prim, mu, mul, mur, mv, mxi);
    lo_auvxi(aL, mul, mv, mxi, 1, 0, mauL);
    lo_auvxi(aR, mur, mv, mxi, 1, 0, mauR);
    lo_auvxi(bs, mu, mv, mxi, 0, 1, mbv);
    for (int m = 0; m < 4; ++m) sw[m] = -prim[0] * (mauL[m] + mauR[m] + mbv[m]);
    lo_micro_slope(ck, prim, sw, aT);
    double Mt[5];
    lo_time_coeffs(dt, tau_of_prim(D.gas, prim), Mt);
    uni<4>(prim); uni<4>(aL); uni<4>(aR); uni<4>(bs); uni<4>(aT); uni<5>(Mt);
    const double mxi2 = uni(mxi[2]);

    // ---- face equilibrium: Maxwellian of prim projected on its moments ----
    // frame velocities (vn, vt): x face (u, v), y face (v, -u)
    gauss_tables<NT>(D, prim[3], xf ? prim[1] : -prim[2], xf ? prim[2] : prim[1], tu, tv, tmax, lds);
    const double pre = uni(prim[0] * (prim[3] / UGKS_PI)), bf = uni(ck / (2.0 * prim[3]));
    const double sh = fmax(1.0, pre * (tmax[0] * tmax[1])), sb = fmax(1.0, pre * (tmax[0] * tmax[1]) * bf);
    const double eh = uni(UGKS_DBL_EPS * sh), eb = uni(UGKS_DBL_EPS * sb);
    double s3[14];
    for (int a = 0; a < 14; ++a) s3[a] = 0.0;
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = xf ? u : v, vt = xf ? v : -u;
        const double H = pre * (tu[k & (nu - 1)] * tv[k / nu]), B = H * bf;
        mom_add(s3, w, vn, vt, H, B);
        gram_add(s3 + 4, w * fmax(H, eh), w * fmax(B, eb), vn, vt);
    }
    block_sum<14>(s3, lds);
    double target[4], rhs[4], G[16], lam[4];
    conserved(D.gas.gamma, prim, target);
    for (int a = 0; a < 4; ++a) rhs[a] = target[a] - s3[a];
    gram_full(s3 + 4, G);
    if (!solve4_gj_reg(G, rhs, lam)) {
        if (threadIdx.x == 0) flag(D, ST_PROJECT | ST_AT_FACE);
        return;
    }
    uni<4>(lam);
    // projected equilibrium, its moment check, and the heat flux of the face data
    double s4[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = xf ? u : v, vt = xf ? v : -u;
        const double H = pre * (tu[k & (nu - 1)] * tv[k / nu]), B = H * bf;
        const double Hp = H + fmax(H, eh) * (lam[0] + vn * lam[1] + vt * lam[2] + 0.5 * (vn * vn + vt * vt) * lam[3]);
        const double Bp = B + 0.5 * fmax(B, eb) * lam[3];
        mom_add(s4, w, vn, vt, Hp, Bp);
        double fd[6];
        face_data(D, xf, i, j, k, !signbit(vn), idx_, idy, len_n, fd);
        const double dvn = vn - prim[1], dvt = vt - prim[2], v2 = dvn * dvn + dvt * dvt;
        s4[4] += 0.5 * w * dvn * (v2 * fd[0] + fd[1]);
        s4[5] += 0.5 * w * dvt * (v2 * fd[0] + fd[1]);
    }
    block_sum<6>(s4, lds);
    if (!project_ok(s4, target)) {
        if (threadIdx.x == 0) flag(D, ST_GATE | ST_AT_FACE);
        return;
    }
    const double coeff = uni(0.8 * (1.0 - D.gas.pr) * prim[3] * prim[3] / prim[0]);

    // ---- micro flux and its moments ----------------------------------------
    double* oh = xf ? D.xfl_h + (size_t)f * nvel : D.yfl_h + (size_t)(g + nx) * nvel;
    double* ob = xf ? D.xfl_b + (size_t)f * nvel : D.yfl_b + (size_t)(g + nx) * nvel;
    double s5[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = xf ? u : v, vt = xf ? v : -u, v2 = vn * vn + vt * vt;
        const bool left = !signbit(vn);
        const double H = pre * (tu[k & (nu - 1)] * tv[k / nu]), B = H * bf;
        const double H0 = H + fmax(H, eh) * (lam[0] + vn * lam[1] + vt * lam[2] + 0.5 * v2 * lam[3]);
        const double B0 = B + 0.5 * fmax(B, eb) * lam[3];
        const double dvn = vn - prim[1], dvt = vt - prim[2], c2 = dvn * dvn + dvt * dvt;
        const double qdot = dvn * s4[4] + dvt * s4[5], shH = 2.0 * prim[3] * c2 - 4.0, shB = shH + 2.0 / ck;
        const double Hs = coeff * qdot * shH * H0, Bs = coeff * qdot * shB * B0;
        const double an0 = left ? aL[0] : aR[0], an1 = left ? aL[1] : aR[1], an2 = left ? aL[2] : aR[2],
                     an3 = left ? aL[3] : aR[3];
        const double pa = an0 + an1 * vn + an2 * vt, pb = bs[0] + bs[1] * vn + bs[2] * vt,
                     pt = aT[0] + aT[1] * vn + aT[2] * vt;
        const double h_an = pa * H0 + 0.5 * an3 * (v2 * H0 + B0), b_an = pa * B0 + 0.5 * an3 * (v2 * B0 + mxi2 * H0);
        const double h_b = pb * H0 + 0.5 * bs[3] * (v2 * H0 + B0), b_b = pb * B0 + 0.5 * bs[3] * (v2 * B0 + mxi2 * H0);
        const double h_t = pt * H0 + 0.5 * aT[3] * (v2 * H0 + B0), b_t = pt * B0 + 0.5 * aT[3] * (v2 * B0 + mxi2 * H0);
        double fd[6];
        face_data(D, xf, i, j, k, left, idx_, idy, len_n, fd);
        const double slh = vn * fd[2] + vt * fd[4], slb = vn * fd[3] + vt * fd[5];
        const double Fh = Mt[0] * vn * (H0 + Hs) + Mt[1] * vn * vn * h_an + Mt[1] * vn * vt * h_b + Mt[2] * vn * h_t +
                          Mt[3] * vn * fd[0] - Mt[4] * vn * slh;
        const double Fb = Mt[0] * vn * (B0 + Bs) + Mt[1] * vn * vn * b_an + Mt[1] * vn * vt * b_b + Mt[2] * vn * b_t +
                          Mt[3] * vn * fd[1] - Mt[4] * vn * slb;
        const double Gh = flen * Fh, Gb = flen * Fb;
        oh[k] = Gh;
        ob[k] = Gb;
        mom_add(s5, w, u, v, Gh, Gb);
    }
    block_sum<4>(s5, lds);
    if (threadIdx.x < 4) (xf ? D.xfw + 4 * (size_t)f : D.yfw + 4 * (size_t)(g + nx))[threadIdx.x] = s5[threadIdx.x];
}

// Wall-side data of the wall face of cell c = (i, row 0 or ny-1) for
// velocity k (out: h, b at the face, dh/dn, db/dn with the one-sided y
// difference, dh/dt, db/dt = -(x Van Leer slope)).  hoff = rot * dy/2.
__device__ inline void wall_data(const Dev& D, int top, int i, int c, int k, double idx_, double idy, double hoff,
                                 double out[6]) {
    const int nx = D.nx;
    const size_t nvel = D.nvel;
    const int iw = i == 0 ? nx - 1 : i - 1, ie = i + 1 == nx ? 0 : i + 1;
    const size_t o = (size_t)c * nvel + k;
    const size_t oy = top ? o - (size_t)nx * nvel : o + (size_t)nx * nvel;
    const size_t ow = ((size_t)(c - i + iw)) * nvel + k, oe = ((size_t)(c - i + ie)) * nvel + k;
    const double hc = D.h[o], bc = D.b[o];
    const double nh = top ? (hc - D.h[oy]) * idy : (D.h[oy] - hc) * idy;
    const double nb = top ? (bc - D.b[oy]) * idy : (D.b[oy] - bc) * idy;
    out[0] = hc - hoff * nh;
    out[1] = bc - hoff * nb;
    out[2] = nh; out[3] = nb;
    out[4] = -vanleer(D.h[ow], hc, D.h[oe], idx_, idx_);
    out[5] = -vanleer(D.b[ow], bc, D.b[oe], idx_, idx_);
}

// ===========================================================================
// multiscale diffuse walls: blocks [0, nx) bottom faces, [nx, 2nx) top faces
// ===========================================================================
template <int NT>
__global__ void __launch_bounds__(NT) k_lo_wall(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[(NT / 16) * 28];
    __shared__ double tu[LO_MAXAX], tv[LO_MAXAX], tmax[2];
    const int nx = D.nx, ny = D.ny, nu = D.nu, nvel = D.nvel, ck = D.gas.ck;
    const int top = blockIdx.x >= nx ? 1 : 0, i = blockIdx.x - top * nx;
    const double side = top ? 1.0 : -1.0, rot = -side;
    const int c = top ? (ny - 1) * nx + i : i;
    const double dt = D.scal[SC_DT], flen = D.dx, len = D.dy;
    const double idx_ = 1.0 / D.dx, idy = 1.0 / D.dy;
    const double* W = D.W + 4 * (size_t)c;

    // wall-side face data (one-sided y slope of the wall row, x Van Leer),
    // re-read from cache in each pass
    const double lam_w = 1.0 / D.wall_tn[top];
    const double prim_w[4] = {1.0, 0.0, -D.wall_ux[top], lam_w};
    double pc[4];
    primary(D.gas.gamma, W, pc);
    double prim_g[4] = {pc[0] / pc[3] * prim_w[3], prim_w[1], prim_w[2], prim_w[3]};
    uni<4>(prim_g);
    double s1[6] = {0, 0, 0, 0, 0, 0};   // swt(4), heat(2)
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = v, vt = -u;
        double fd[6];
        wall_data(D, top, i, c, k, idx_, idy, rot * 0.5 * len, fd);
        const double v2 = vn * vn + vt * vt;
        s1[0] += w * fd[4]; s1[1] += w * vn * fd[4]; s1[2] += w * vt * fd[4]; s1[3] += 0.5 * w * (v2 * fd[4] + fd[5]);
        const double dvn = vn - prim_g[1], dvt = vt - prim_g[2], c2 = dvn * dvn + dvt * dvt;
        s1[4] += 0.5 * w * dvn * (c2 * fd[0] + fd[1]);
        s1[5] += 0.5 * w * dvt * (c2 * fd[0] + fd[1]);
    }
    block_sum<6>(s1, lds);

    double Mt[5];
    lo_time_coeffs(dt, tau_of_prim(D.gas, prim_g), Mt);
    double cg[4], sw[4], a_gn[4], a_gt[4], a_T[4];
    conserved(D.gas.gamma, prim_g, cg);
    const double cg1[4] = {W[0], W[2], -W[1], W[3]};
    for (int m = 0; m < 4; ++m) sw[m] = (cg[m] - cg1[m]) / (side * 0.5 * len);
    lo_micro_slope(ck, prim_g, sw, a_gn);
    const double swt[4] = {s1[0], s1[1], s1[2], s1[3]};
    lo_micro_slope(ck, prim_g, swt, a_gt);
    double mu[7], mul[7], mur[7], mv[6], mxi[3], mau[4], mbv[4];
    lo_moments(ck, prim_g, mu, mul, mur, mv, mxi);
    lo_auvxi(a_gn, mu, mv, mxi, 1, 0, mau);
    lo_auvxi(a_gt, mu, mv, mxi, 0, 1, mbv);
    for (int m = 0; m < 4; ++m) sw[m] = -prim_g[0] * (mau[m] + mbv[m]);
    lo_micro_slope(ck, prim_g, sw, a_T);
    uni<4>(a_gn); uni<4>(a_gt); uni<4>(a_T); uni<5>(Mt);
    const double mxi2 = uni(mxi[2]);

    // the two face equilibria: the gas state prim_g and the unit-density wall
    // shape; both share U, V and lambda, so one pair of Gaussian tables
    gauss_tables<NT>(D, prim_w[3], -prim_w[2], prim_w[1], tu, tv, tmax, lds);
    const double pg = prim_g[0] * (prim_w[3] / UGKS_PI), ps = prim_w[3] / UGKS_PI, bf = ck / (2.0 * prim_w[3]);
    const double shg = fmax(1.0, pg * (tmax[0] * tmax[1])), sbg = fmax(1.0, pg * (tmax[0] * tmax[1]) * bf);
    const double shs = fmax(1.0, ps * (tmax[0] * tmax[1])), sbs = fmax(1.0, ps * (tmax[0] * tmax[1]) * bf);
    const double ehg = UGKS_DBL_EPS * shg, ebg = UGKS_DBL_EPS * sbg, ehs = UGKS_DBL_EPS * shs, ebs = UGKS_DBL_EPS * sbs;
    double s2[28];
    for (int a = 0; a < 28; ++a) s2[a] = 0.0;
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = v, vt = -u, T = tu[k & (nu - 1)] * tv[k / nu];
        const double Hg = pg * T, Bg = Hg * bf, Hs = ps * T, Bs = Hs * bf;
        mom_add(s2, w, vn, vt, Hg, Bg);
        gram_add(s2 + 4, w * fmax(Hg, ehg), w * fmax(Bg, ebg), vn, vt);
        mom_add(s2 + 14, w, vn, vt, Hs, Bs);
        gram_add(s2 + 18, w * fmax(Hs, ehs), w * fmax(Bs, ebs), vn, vt);
    }
    block_sum<28>(s2, lds);
    double tg[4], ts[4], rg[4], rs[4], G[16], lg[4], ls[4];
    conserved(D.gas.gamma, prim_g, tg);
    const double prim_s[4] = {1.0, prim_w[1], prim_w[2], prim_w[3]};
    conserved(D.gas.gamma, prim_s, ts);
    for (int a = 0; a < 4; ++a) { rg[a] = tg[a] - s2[a]; rs[a] = ts[a] - s2[14 + a]; }
    gram_full(s2 + 4, G);
    bool ok = solve4_gj_reg(G, rg, lg);
    gram_full(s2 + 18, G);
    ok = ok && solve4_gj_reg(G, rs, ls);
    if (!ok) {
        if (threadIdx.x == 0) flag(D, ST_PROJECT | ST_AT_WALL);
        return;
    }
    uni<4>(lg); uni<4>(ls);
    const double coeff = 0.8 * (1.0 - D.gas.pr) * prim_g[3] * prim_g[3] / prim_g[0];
    // projection checks, and the incidence/reflection of the zero-mass-flux wall density
    double s3[10];
    for (int a = 0; a < 10; ++a) s3[a] = 0.0;
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = v, vt = -u, v2 = vn * vn + vt * vt, T = tu[k & (nu - 1)] * tv[k / nu];
        const double Hg0 = pg * T, Bg0 = Hg0 * bf, Hs0 = ps * T, Bs0 = Hs0 * bf;
        const double kl = 0.5 * v2;
        const double Hg = Hg0 + fmax(Hg0, ehg) * (lg[0] + vn * lg[1] + vt * lg[2] + kl * lg[3]);
        const double Bg = Bg0 + 0.5 * fmax(Bg0, ebg) * lg[3];
        const double Hw = Hs0 + fmax(Hs0, ehs) * (ls[0] + vn * ls[1] + vt * ls[2] + kl * ls[3]);
        const double Bw = Bs0 + 0.5 * fmax(Bs0, ebs) * ls[3];
        mom_add(s3, w, vn, vt, Hg, Bg);
        mom_add(s3 + 4, w, vn, vt, Hw, Bw);
        const double dvn = vn - prim_g[1], dvt = vt - prim_g[2], c2 = dvn * dvn + dvt * dvt;
        const double qdot = dvn * s1[4] + dvt * s1[5], shH = 2.0 * prim_g[3] * c2 - 4.0;
        const double Hp = coeff * qdot * shH * Hg;
        const double delta = (copysign(1.0, vn) * rot + 1.0) * 0.5, inflow = 1.0 - delta;
        const double Ha_gn = (a_gn[0] + a_gn[1] * vn + a_gn[2] * vt) * Hg + 0.5 * a_gn[3] * (v2 * Hg + Bg);
        const double Ha_gt = (a_gt[0] + a_gt[1] * vn + a_gt[2] * vt) * Hg + 0.5 * a_gt[3] * (v2 * Hg + Bg);
        const double Ha_T = (a_T[0] + a_T[1] * vn + a_T[2] * vt) * Hg + 0.5 * a_T[3] * (v2 * Hg + Bg);
        double fd[6];
        wall_data(D, top, i, c, k, idx_, idy, rot * 0.5 * len, fd);
        const double slh = vn * fd[2] + vt * fd[4];
        s3[8] += w * vn * inflow *
                 (Mt[0] * (Hg + Hp) + Mt[1] * (vn * Ha_gn + vt * Ha_gt) + Mt[2] * Ha_T + Mt[3] * fd[0] - Mt[4] * slh);
        s3[9] += w * vn * Hw * delta;
    }
    block_sum<10>(s3, lds);
    if (!project_ok(s3, tg) || !project_ok(s3 + 4, ts)) {
        if (threadIdx.x == 0) flag(D, ST_GATE | ST_AT_WALL);
        return;
    }
    const double rho_w = -s3[8] / (dt * s3[9]);
    if (!isfinite(rho_w)) {
        if (threadIdx.x == 0) flag(D, ST_WALL | ST_AT_WALL);
        return;
    }
    const size_t fo = (size_t)(top ? ny : 0) * nx + i;
    double* oh = D.yfl_h + fo * nvel;
    double* ob = D.yfl_b + fo * nvel;
    double s4[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k];
        const double vn = v, vt = -u, v2 = vn * vn + vt * vt, T = tu[k & (nu - 1)] * tv[k / nu];
        const double Hg0 = pg * T, Bg0 = Hg0 * bf, Hs0 = ps * T, Bs0 = Hs0 * bf;
        const double kl = 0.5 * v2;
        const double Hg = Hg0 + fmax(Hg0, ehg) * (lg[0] + vn * lg[1] + vt * lg[2] + kl * lg[3]);
        const double Bg = Bg0 + 0.5 * fmax(Bg0, ebg) * lg[3];
        const double Hw = rho_w * (Hs0 + fmax(Hs0, ehs) * (ls[0] + vn * ls[1] + vt * ls[2] + kl * ls[3]));
        const double Bw = rho_w * (Bs0 + 0.5 * fmax(Bs0, ebs) * ls[3]);
        double fd[6];
        wall_data(D, top, i, c, k, idx_, idy, rot * 0.5 * len, fd);
        const double dvn = vn - prim_g[1], dvt = vt - prim_g[2], c2 = dvn * dvn + dvt * dvt;
        const double qdot = dvn * s1[4] + dvt * s1[5], shH = 2.0 * prim_g[3] * c2 - 4.0, shB = shH + 2.0 / ck;
        const double Hp = coeff * qdot * shH * Hg, Bp = coeff * qdot * shB * Bg;
        const double delta = (copysign(1.0, vn) * rot + 1.0) * 0.5, inflow = 1.0 - delta;
        const double pn = a_gn[0] + a_gn[1] * vn + a_gn[2] * vt, pt = a_gt[0] + a_gt[1] * vn + a_gt[2] * vt,
                     pT = a_T[0] + a_T[1] * vn + a_T[2] * vt;
        const double Ha_gn = pn * Hg + 0.5 * a_gn[3] * (v2 * Hg + Bg), Ba_gn = pn * Bg + 0.5 * a_gn[3] * (v2 * Bg + mxi2 * Hg);
        const double Ha_gt = pt * Hg + 0.5 * a_gt[3] * (v2 * Hg + Bg), Ba_gt = pt * Bg + 0.5 * a_gt[3] * (v2 * Bg + mxi2 * Hg);
        const double Ha_T = pT * Hg + 0.5 * a_T[3] * (v2 * Hg + Bg), Ba_T = pT * Bg + 0.5 * a_T[3] * (v2 * Bg + mxi2 * Hg);
        const double Fh = dt * vn * Hw * delta + Mt[0] * vn * (Hg + Hp) * inflow + Mt[1] * vn * vn * Ha_gn * inflow +
                          Mt[1] * vn * vt * Ha_gt * inflow + Mt[2] * vn * Ha_T * inflow + Mt[3] * vn * fd[0] * inflow -
                          Mt[4] * vn * (vn * fd[2] + vt * fd[4]) * inflow;
        const double Fb = dt * vn * Bw * delta + Mt[0] * vn * (Bg + Bp) * inflow + Mt[1] * vn * vn * Ba_gn * inflow +
                          Mt[1] * vn * vt * Ba_gt * inflow + Mt[2] * vn * Ba_T * inflow + Mt[3] * vn * fd[1] * inflow -
                          Mt[4] * vn * (vn * fd[3] + vt * fd[5]) * inflow;
        oh[k] = flen * Fh;
        ob[k] = flen * Fb;
        mom_add(s4, w, vn, vt, Fh, Fb);
    }
    block_sum<4>(s4, lds);
    if (threadIdx.x == 0) {
        double* fw = D.yfw + 4 * fo;
        fw[0] = flen * s4[0]; fw[1] = flen * (-s4[2]); fw[2] = flen * s4[1]; fw[3] = flen * s4[3];
    }
}

// ===========================================================================
// cell update: W^{n+1} from the macro fluxes, trapezoidal Shakhov collision
// with the projected old and new equilibria (ApplyExplicitUpdate)
// ===========================================================================
template <int NT>
__global__ void __launch_bounds__(NT) k_lo_cell(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[(NT / 16) * 14];
    __shared__ double tuo[LO_MAXAX], tvo[LO_MAXAX], tmo[2], tun[LO_MAXAX], tvn[LO_MAXAX], tmn[2];
    const int nx = D.nx, nu = D.nu, nvel = D.nvel, ck = D.gas.ck;
    const int c = xcd_contiguous(blockIdx.x, gridDim.x);
    const int i = c % nx, j = c / nx, cw = j * nx + (i == 0 ? nx - 1 : i - 1);
    const size_t fs = (size_t)j * nx + i, fn = (size_t)(j + 1) * nx + i;
    const double dt = D.scal[SC_DT], area = D.dx * D.dy, ainv = 1.0 / area;
    const double* Wo = D.W + 4 * (size_t)c;
    double po[4], pn[4], wn[4];
    primary(D.gas.gamma, Wo, po);
    const double tau_o = tau_of_prim(D.gas, po);
    for (int m = 0; m < 4; ++m)
        wn[m] = Wo[m] + (D.xfw[4 * (size_t)cw + m] - D.xfw[4 * (size_t)c + m] + D.yfw[4 * fs + m] - D.yfw[4 * fn + m]) / area;
    if (!physical_w(wn)) {
        if (threadIdx.x == 0) flag(D, ST_NONFINITE | ST_AT_CELL);
        return;
    }
    primary(D.gas.gamma, wn, pn);
    const double tau = tau_of_prim(D.gas, pn);

    // heat flux of the old state (h, b are re-read from cache by the last pass)
    double q[2] = {0, 0};
    uni<4>(po); uni<4>(pn);
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const size_t o = (size_t)c * nvel + k;
        const double h = D.h[o], b = D.b[o];
        const double dvn = D.u[k] - po[1], dvt = D.v[k] - po[2], v2 = dvn * dvn + dvt * dvt;
        q[0] += 0.5 * D.w[k] * dvn * (v2 * h + b);
        q[1] += 0.5 * D.w[k] * dvt * (v2 * h + b);
    }
    block_sum<2>(q, lds);
    gauss_tables<NT>(D, po[3], po[1], po[2], tuo, tvo, tmo, lds);
    gauss_tables<NT>(D, pn[3], pn[1], pn[2], tun, tvn, tmn, lds);
    const double preo = uni(po[0] * (po[3] / UGKS_PI)), bfo = uni(ck / (2.0 * po[3]));
    const double pren = uni(pn[0] * (pn[3] / UGKS_PI)), bfn = uni(ck / (2.0 * pn[3]));
    const double co = uni(0.8 * (1.0 - D.gas.pr) * po[3] * po[3] / po[0]);
    const double cn = uni(0.8 * (1.0 - D.gas.pr) * pn[3] * pn[3] / pn[0]);
    // Maxwellian + Shakhov part of prim p with heat flux q at velocity k
    auto ms = [&](const double* p, const double* tu, const double* tv, double pre, double bf, double cf, int k,
                  double* eh, double* eb) {
        const double hk = pre * (tu[k & (nu - 1)] * tv[k / nu]), bk = hk * bf;
        const double dvn = D.u[k] - p[1], dvt = D.v[k] - p[2], v2 = dvn * dvn + dvt * dvt;
        const double qdot = dvn * q[0] + dvt * q[1], shH = 2.0 * p[3] * v2 - 4.0, shB = shH + 2.0 / ck;
        *eh = hk + cf * qdot * shH * hk;
        *eb = bk + cf * qdot * shB * bk;
    };
    // projection weights' scales: max |.| of both equilibria
    double mx[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        double a, bb;
        ms(po, tuo, tvo, preo, bfo, co, k, &a, &bb);
        mx[0] = fmax(mx[0], fabs(a)); mx[1] = fmax(mx[1], fabs(bb));
        ms(pn, tun, tvn, pren, bfn, cn, k, &a, &bb);
        mx[2] = fmax(mx[2], fabs(a)); mx[3] = fmax(mx[3], fabs(bb));
    }
    block_max<4>(mx, lds);
    const double eho = uni(UGKS_DBL_EPS * fmax(1.0, mx[0])), ebo = uni(UGKS_DBL_EPS * fmax(1.0, mx[1]));
    const double ehn = uni(UGKS_DBL_EPS * fmax(1.0, mx[2])), ebn = uni(UGKS_DBL_EPS * fmax(1.0, mx[3]));
    // moments and Gram matrices of both equilibria, one at a time
    double to[4], ro[4], rn[4], G[16], lo[4], ln[4];
    conserved(D.gas.gamma, po, to);
    bool ok = true;
#pragma unroll 1
    for (int e = 0; e < 2; ++e) {
        double s[14];
        for (int a = 0; a < 14; ++a) s[a] = 0.0;
        const double* p = e ? pn : po;
        const double *tu = e ? tun : tuo, *tv = e ? tvn : tvo;
        const double pre = e ? pren : preo, bf = e ? bfn : bfo, cf = e ? cn : co;
        const double eh = e ? ehn : eho, eb = e ? ebn : ebo;
#pragma unroll 1
        for (int k = threadIdx.x; k < nvel; k += NT) {
            const double u = D.u[k], v = D.v[k], w = D.w[k];
            double a, bb;
            ms(p, tu, tv, pre, bf, cf, k, &a, &bb);
            mom_add(s, w, u, v, a, bb);
            gram_add(s + 4, w * fmax(fabs(a), eh), w * fmax(fabs(bb), eb), u, v);
        }
        block_sum<14>(s, lds);
        double* rr = e ? rn : ro;
        const double* tt = e ? wn : to;
        for (int a = 0; a < 4; ++a) rr[a] = tt[a] - s[a];
        gram_full(s + 4, G);
        ok = ok && solve4_gj_reg(G, rr, e ? ln : lo);
    }
    if (!ok) {
        if (threadIdx.x == 0) flag(D, ST_PROJECT | ST_AT_CELL);
        return;
    }
    uni<4>(lo); uni<4>(ln);
    const double inv = uni(1.0 / (1.0 + 0.5 * dt / tau)), itau = uni(1.0 / tau), itau_o = uni(1.0 / tau_o);
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double u = D.u[k], v = D.v[k], w = D.w[k], kl = 0.5 * (u * u + v * v);
        double Ho, Bo, H, B;
        ms(po, tuo, tvo, preo, bfo, co, k, &Ho, &Bo);
        ms(pn, tun, tvn, pren, bfn, cn, k, &H, &B);
        Ho = Ho + fmax(fabs(Ho), eho) * (lo[0] + u * lo[1] + v * lo[2] + kl * lo[3]);
        Bo = Bo + 0.5 * fmax(fabs(Bo), ebo) * lo[3];
        H = H + fmax(fabs(H), ehn) * (ln[0] + u * ln[1] + v * ln[2] + kl * ln[3]);
        B = B + 0.5 * fmax(fabs(B), ebn) * ln[3];
        mom_add(m, w, u, v, Ho, Bo);
        mom_add(m + 4, w, u, v, H, B);
        const size_t o = (size_t)c * nvel + k;
        const double h = D.h[o], b = D.b[o];
        const double rh = (D.xfl_h[(size_t)cw * nvel + k] - D.xfl_h[o] + D.yfl_h[fs * nvel + k] - D.yfl_h[fn * nvel + k]) * ainv;
        const double rb = (D.xfl_b[(size_t)cw * nvel + k] - D.xfl_b[o] + D.yfl_b[fs * nvel + k] - D.yfl_b[fn * nvel + k]) * ainv;
        D.hn[o] = (h + rh + 0.5 * dt * (H * itau + (Ho - h) * itau_o)) * inv;
        D.bn[o] = (b + rb + 0.5 * dt * (B * itau + (Bo - b) * itau_o)) * inv;
    }
    block_sum<8>(m, lds);
    if (threadIdx.x == 0) {
        if (!project_ok(m, to) || !project_ok(m + 4, wn)) flag(D, ST_GATE | ST_AT_CELL);
        double* cp = D.cellpart + 16 * (size_t)c;
        for (int a = 0; a < 4; ++a) {
            D.Wn[4 * (size_t)c + a] = wn[a];
            const double d = (Wo[a] - wn[a]) * area / dt;
            cp[a] = d * d;
            cp[4 + a] = fabs(wn[a]) * area;
        }
    }
}

// residual res(1:4) of the step, in fixed order (Write_residuals semantics)
__global__ void __launch_bounds__(256) k_lo_finish(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[16 * 8];
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = threadIdx.x; c < D.ncell; c += 256)
        for (int a = 0; a < 8; ++a) s[a] += D.cellpart[16 * (size_t)c + a];
    block_sum<8>(s, lds);
    if (threadIdx.x == 0) {
        const double area = D.dx * D.dy, areaSq = (double)D.ncell * area * area;
        const double ord = fmax(fmax(s[4], s[7]), LO_TINY);
        for (int m = 0; m < 4; ++m) {
            const double num = sqrt((double)D.ncell * s[m]), den = s[4 + m] + LO_TINY;
            const double rel = num / den, ab = sqrt(s[m] / fmax(areaSq, LO_TINY));
            D.scal[SC_LO_RES + m] = s[4 + m] <= sqrt(UGKS_DBL_EPS) * fmax(ord, UGKS_DBL_MIN) ? ab : rel;
        }
    }
}

// freestream cold start: every cell holds W = GetConserved(prim) and the
// Maxwellian of prim projected onto W
template <int NT>
__global__ void __launch_bounds__(NT) k_lo_cold_start(const Dev* __restrict__ DS, double rho, double ux, double tn) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[(NT / 16) * 14];
    __shared__ double tu[LO_MAXAX], tv[LO_MAXAX], tmax[2];
    const int nu = D.nu, nvel = D.nvel, ck = D.gas.ck;
    const double prim[4] = {rho, ux, 0.0, 1.0 / tn};
    double target[4];
    conserved(D.gas.gamma, prim, target);
    gauss_tables<NT>(D, prim[3], prim[1], prim[2], tu, tv, tmax, lds);
    const double pre = prim[0] * (prim[3] / UGKS_PI), bf = ck / (2.0 * prim[3]);
    const double sh = fmax(1.0, pre * (tmax[0] * tmax[1])), sb = fmax(1.0, pre * (tmax[0] * tmax[1]) * bf);
    const double eh = UGKS_DBL_EPS * sh, eb = UGKS_DBL_EPS * sb;
    double s[14];
    for (int a = 0; a < 14; ++a) s[a] = 0.0;
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        const double H = pre * (tu[k & (nu - 1)] * tv[k / nu]), B = H * bf;
        mom_add(s, D.w[k], D.u[k], D.v[k], H, B);
        gram_add(s + 4, D.w[k] * fmax(H, eh), D.w[k] * fmax(B, eb), D.u[k], D.v[k]);
    }
    block_sum<14>(s, lds);
    double rhs[4], G[16], lam[4];
    for (int a = 0; a < 4; ++a) rhs[a] = target[a] - s[a];
    gram_full(s + 4, G);
    if (!solve4_gj_reg(G, rhs, lam)) {
        if (threadIdx.x == 0) flag(D, ST_PROJECT | ST_AT_COLD);
        return;
    }
    // the projected Maxwellian of velocity k
    auto fk = [&](int k, double* hk, double* bk) {
        const double u = D.u[k], v = D.v[k];
        const double H = pre * (tu[k & (nu - 1)] * tv[k / nu]), B = H * bf;
        *hk = H + fmax(H, eh) * (lam[0] + u * lam[1] + v * lam[2] + 0.5 * (u * u + v * v) * lam[3]);
        *bk = B + 0.5 * fmax(B, eb) * lam[3];
    };
    double m[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int k = threadIdx.x; k < nvel; k += NT) {
        double hk, bk;
        fk(k, &hk, &bk);
        mom_add(m, D.w[k], D.u[k], D.v[k], hk, bk);
    }
    block_sum<4>(m, lds);
    if (!project_ok(m, target)) {
        if (threadIdx.x == 0) flag(D, ST_GATE | ST_AT_COLD);
        return;
    }
    for (int c = blockIdx.x; c < D.ncell; c += gridDim.x) {
#pragma unroll 1
        for (int k = threadIdx.x; k < nvel; k += NT) {
            double hk, bk;
            fk(k, &hk, &bk);
            D.h[(size_t)c * nvel + k] = hk;
            D.b[(size_t)c * nvel + k] = bk;
        }
        if (threadIdx.x < 4) D.W[4 * (size_t)c + threadIdx.x] = target[threadIdx.x];
    }
}

// ===========================================================================
// launchers
// ===========================================================================
extern "C" hipError_t ugks_launch_timestep(const LaunchSet* L, hipStream_t s, Prof* P);

namespace {
template <int NT>
hipError_t launch_lo_step(const LaunchSet& L, hipStream_t s, Prof* P) {
    hipError_t e = ugks_launch_timestep(&L, s, P);
    if (e != hipSuccess) return e;
    const Dev& D = *L.size;
    const int n = L.n;
    UGKS_TIMED(P, K_LO_FACE, s, hipLaunchKernelGGL((k_lo_face<NT, true>), dim3(D.ncell, 1, n), dim3(NT), 0, s, L.dev));
    UGKS_TIMED(P, K_LO_FACE, s,
               hipLaunchKernelGGL((k_lo_face<NT, false>), dim3(D.nx * (D.ny - 1), 1, n), dim3(NT), 0, s, L.dev));
    UGKS_TIMED(P, K_LO_WALL, s, hipLaunchKernelGGL(k_lo_wall<NT>, dim3(2 * D.nx, 1, n), dim3(NT), 0, s, L.dev));
    UGKS_TIMED(P, K_LO_CELL, s, hipLaunchKernelGGL(k_lo_cell<NT>, dim3(D.ncell, 1, n), dim3(NT), 0, s, L.dev));
    UGKS_TIMED(P, K_LO_FINISH, s, hipLaunchKernelGGL(k_lo_finish, dim3(1, 1, n), dim3(256), 0, s, L.dev));
    return hipGetLastError();
}
}  // namespace

// One step of the low-order Explicit entry (one instance, L.n == 1: the face
// and cell kernels map workgroups to faces XCD-contiguously over the grid);
// the proposal goes to Wn/hn/bn.  wg: workgroup size of the face, wall and
// cell kernels (64 .. 1024, at most Nvel; each thread loops over the
// velocities k = t, t + wg, ...).
extern "C" hipError_t ugks_launch_lo_step(const LaunchSet* L, int wg, hipStream_t s, Prof* P) {
    const Dev& D = *L->size;
    if (L->n != 1 || wg > D.nvel || D.nu + D.nv > wg) return hipErrorInvalidValue;
    switch (wg) {
    case 64: return launch_lo_step<64>(*L, s, P);
    case 128: return launch_lo_step<128>(*L, s, P);
    case 256: return launch_lo_step<256>(*L, s, P);
    case 512: return launch_lo_step<512>(*L, s, P);
    case 1024: return launch_lo_step<1024>(*L, s, P);
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t ugks_launch_lo_cold_start(const LaunchSet* L, double rho, double ux, double tn, hipStream_t s) {
    const Dev& D = *L->size;
    if (D.nu + D.nv > 256) return hipErrorInvalidValue;
    const int nb = D.ncell < 1024 ? D.ncell : 1024;
    hipLaunchKernelGGL(k_lo_cold_start<256>, dim3(nb, 1, L->n), dim3(256), 0, s, L->dev, rho, ux, tn);
    return hipGetLastError();
}
