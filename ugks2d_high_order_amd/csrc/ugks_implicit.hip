// R6 pseudo-time solver of the HighOrderImplicit step on gfx950:
//   macro PR(2) symmetric Gauss-Seidel       src/mod_ImplicitMacro.f90:15-123,187-263
//   predicted equilibrium and low-order      src/mod_ImplicitMicro.f90:43-88,
//   collision of the old state               src/mod_Tools.f90:18-107,355-373,
//                                            src/mod_ImplicitCommon.f90:42-113
//   micro LU-SGS forward/backward pair with  src/mod_ImplicitMicro.f90:366-487,647-791,
//   the frozen diffuse-wall closure          src/mod_ImplicitBoundary.f90:282-326
//   admissibility, apply, macro recovery     src/mod_SteadyImplicit.f90:53-117,398-462
//   closed-Couette mass rescale              src/mod_ClosedCouetteMass.f90:243-313
//
// Both Gauss-Seidel solvers run as hyperplane wavefronts (cells with i+j = d
// on step d), which is exactly the lexicographic order's dependency order:
// the forward sweep sees new values of the west and south neighbours and old
// values of the east and north ones.  The x-periodic neighbours across the
// seam are read from the values at the last halo exchange, as the reference
// does on one rank.  In the micro sweeps every velocity is independent (the
// wall closure of a cell is frozen from the previous sweep), so the
// wavefront runs inside velocity-tile workgroups with no grid-wide barrier.
#include "ugks_state.h"
#include "ugks_prof.h"

// ===========================================================================
// PR(2) macro solver
// ===========================================================================
__device__ inline void euler_flux(double gamma, const double w[4], double nx, double ny, double f[4]) {
    double u = w[1] / w[0], v = w[2] / w[0], un = u * nx + v * ny;
    double p = (gamma - 1.0) * (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
    f[0] = w[0] * un;
    f[1] = w[1] * un + p * nx;
    f[2] = w[2] * un + p * ny;
    f[3] = (w[3] + p) * un;
}

__device__ inline void zhu_map(const double w[4], double uw, double vw, double g[4]) {
    g[0] = w[0];
    g[1] = 2.0 * uw * w[0] - w[1];
    g[2] = 2.0 * vw * w[0] - w[2];
    g[3] = w[3] + 2.0 * (uw * uw + vw * vw) * w[0] - 2.0 * uw * w[1] - 2.0 * vw * w[2];
}

__device__ inline double spectral_radius(const Gas& g, const double wi[4], const double wj[4], double nxo, double nyo,
                                         double dist) {
    double pi_[4], pj[4];
    primary(g.gamma, wi, pi_);
    primary(g.gamma, wj, pj);
    double rho = 0.5 * (pi_[0] + pj[0]), ua = 0.5 * (pi_[1] + pj[1]), va = 0.5 * (pi_[2] + pj[2]);
    double p_i = 0.5 * pi_[0] / pi_[3], p_j = 0.5 * pj[0] / pj[3], pa = 0.5 * (p_i + p_j);
    double ti = tau_of_prim(g, pi_), tj = tau_of_prim(g, pj);
    double mu = 0.5 * (ti * p_i + tj * p_j);
    double c = sqrt(g.gamma * pa / rho);
    return fabs(ua * nxo + va * nyo) + c + 2.0 * mu / (rho * fmax(dist, 1000.0 * UGKS_DBL_EPS));
}

// spectral radii of all faces (they depend on W only)
__global__ void __launch_bounds__(256) k_macro_gamma(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = D.nx, ny = D.ny;
    if (t < D.ncell) {   // east face of cell t
        int i = t % nx, j = t / nx, e = j * nx + (i == nx - 1 ? 0 : i + 1);
        double wi[4], wj[4];
        for (int m = 0; m < 4; ++m) { wi[m] = D.W[4 * t + m]; wj[m] = D.W[4 * e + m]; }
        D.gx[t] = spectral_radius(D.gas, wi, wj, 1.0, 0.0, D.dx);
    } else if (t < D.ncell + (ny + 1) * nx) {   // south face of row j
        int g = t - D.ncell, i = g % nx, j = g / nx;
        double wi[4], wj[4];
        if (j == 0 || j == ny) {
            int top = j == ny, c = top ? (ny - 1) * nx + i : i;
            for (int m = 0; m < 4; ++m) wi[m] = D.W[4 * c + m];
            zhu_map(wi, D.wall_ux[top], 0.0, wj);
            D.gy[g] = spectral_radius(D.gas, wi, wj, 0.0, top ? 1.0 : -1.0, D.dy);
        } else {
            int cs = (j - 1) * nx + i, cn = j * nx + i;
            for (int m = 0; m < 4; ++m) { wi[m] = D.W[4 * cn + m]; wj[m] = D.W[4 * cs + m]; }
            D.gy[g] = spectral_radius(D.gas, wi, wj, 0.0, -1.0, D.dy);
        }
    }
}

// RelaxMacroPoint of cell (i,j); returns false on a non-finite result
__device__ inline bool relax_macro(const Dev& D, int i, int j, double dtp) {
    const int nx = D.nx, ny = D.ny, c = j * nx + i;
    double wi[4], di[4];
    for (int m = 0; m < 4; ++m) { wi[m] = D.W[4 * c + m]; di[m] = D.dW[4 * c + m]; }
    double sdt[4] = {0, 0, 0, 0}, sgn[4] = {0, 0, 0, 0}, sg = 0.0;
    for (int f = 0; f < 4; ++f) {
        double nxo = 0.0, nyo = 0.0, len, g, wj[4], dn[4];
        bool wall = false;
        int top = 0;
        if (f == 0) { nxo = -1.0; len = D.dy; }
        else if (f == 1) { nxo = 1.0; len = D.dy; }
        else if (f == 2) { nyo = -1.0; len = D.dx; wall = j == 0; }
        else { nyo = 1.0; len = D.dx; wall = j == ny - 1; top = 1; }
        if (wall) {
            double uw = D.wall_ux[top];
            zhu_map(wi, uw, 0.0, wj);
            zhu_map(di, uw, 0.0, dn);
            g = D.gy[(size_t)(top ? ny : 0) * nx + i];
        } else {
            int ni = i, nj = j;
            bool halo = false;
            if (f == 0) { ni = i - 1; if (ni < 0) { ni = nx - 1; halo = true; } g = D.gx[j * nx + ni]; }
            else if (f == 1) { ni = i + 1; if (ni == nx) { ni = 0; halo = true; } g = D.gx[c]; }
            else if (f == 2) { nj = j - 1; g = D.gy[(size_t)j * nx + i]; }
            else { nj = j + 1; g = D.gy[(size_t)(j + 1) * nx + i]; }
            int cn = nj * nx + ni;
            const double* src = halo ? D.dWs : D.dW;
            for (int m = 0; m < 4; ++m) { wj[m] = D.W[4 * cn + m]; dn[m] = src[4 * cn + m]; }
        }
        double a[4], fnew[4], fold[4], ti[4], tj[4];
        for (int m = 0; m < 4; ++m) a[m] = wi[m] + di[m];
        euler_flux(D.gas.gamma, a, nxo, nyo, fnew);
        euler_flux(D.gas.gamma, wi, nxo, nyo, fold);
        for (int m = 0; m < 4; ++m) ti[m] = fnew[m] - fold[m];
        for (int m = 0; m < 4; ++m) a[m] = wj[m] + dn[m];
        euler_flux(D.gas.gamma, a, nxo, nyo, fnew);
        euler_flux(D.gas.gamma, wj, nxo, nyo, fold);
        for (int m = 0; m < 4; ++m) tj[m] = fnew[m] - fold[m];
        for (int m = 0; m < 4; ++m) {
            sdt[m] += len * (ti[m] + tj[m]);
            sgn[m] += len * g * dn[m];
        }
        sg += len * g;
    }
    const double area = D.dx * D.dy;
    double diag = 1.0 / dtp + 0.5 * sg / area;
    double rhs[4];
    bool ok = isfinite(diag) && diag > 0.0;
    for (int m = 0; m < 4; ++m) {
        rhs[m] = D.rate_w[4 * c + m] - 0.5 * sdt[m] / area + 0.5 * sgn[m] / area;
        ok = ok && isfinite(rhs[m]);
    }
    if (ok)
        for (int m = 0; m < 4; ++m) D.dW[4 * c + m] = rhs[m] / diag;
    return ok;
}

// one workgroup: 2 x (forward + backward) hyperplane sweeps, then the
// predicted macro state (alpha must stay 1 under the closed-Couette policy)
__global__ void __launch_bounds__(256) k_macro_pr2(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    const int nx = D.nx, ny = D.ny, nd = nx + ny - 1;
    const double dtp = D.scal[SC_DTP];
    for (int c = threadIdx.x; c < D.ncell; c += blockDim.x)
        for (int m = 0; m < 4; ++m) { D.dW[4 * c + m] = 0.0; D.dWs[4 * c + m] = 0.0; }
    __syncthreads();
    int ok = 1;
    for (int sweep = 0; sweep < D.macro_sweeps; ++sweep) {
        for (int dd = 0; dd < 2 * nd; ++dd) {
            int d = dd < nd ? dd : 2 * nd - 1 - dd;
            int ilo = d - (ny - 1) > 0 ? d - (ny - 1) : 0, ihi = d < nx - 1 ? d : nx - 1;
            for (int i = ilo + (int)threadIdx.x; i <= ihi; i += blockDim.x) ok &= relax_macro(D, i, d - i, dtp) ? 1 : 0;
            __threadfence_block();
            __syncthreads();
        }
        for (int c = threadIdx.x; c < D.ncell; c += blockDim.x)
            for (int m = 0; m < 4; ++m) D.dWs[4 * c + m] = D.dW[4 * c + m];
        __threadfence_block();
        __syncthreads();
    }
    int good = 1;
    for (int c = threadIdx.x; c < D.ncell; c += blockDim.x) {
        double t[4];
        for (int m = 0; m < 4; ++m) {
            double d = D.dW[4 * c + m];
            good &= isfinite(d) ? 1 : 0;
            t[m] = D.W[4 * c + m] + d;
            D.w_tilde[4 * c + m] = t[m];
        }
        good &= physical_w(t) ? 1 : 0;
    }
    ok = __syncthreads_and(ok);
    good = __syncthreads_and(good);
    if (threadIdx.x == 0) {
        if (!ok || !good) atomicOr(D.status, !ok ? (unsigned)ST_NONFINITE : (unsigned)ST_MACRO_ALPHA);
    }
}

// ===========================================================================
// predicted equilibrium: per-cell right-hand side of the micro sweeps
// ===========================================================================
template <int KPT, int NT>
struct CellVel {
    double u[KPT], v[KPT], w[KPT];
};

// Maxwellian plus Shakhov part at prim (rho,U,V,lambda) with heat flux q
__device__ inline void maxwell_shakhov(const Gas& g, const double prim[4], const double q[2], double u, double v,
                                       double* eh, double* eb) {
    double pre = prim[0] * (prim[3] / UGKS_PI), bf = g.ck / (2.0 * prim[3]);
    double coeff = 0.8 * (1.0 - g.pr) * prim[3] * prim[3] / prim[0];
    double dvn = u - prim[1], dvt = v - prim[2], v2 = dvn * dvn + dvt * dvt;
    double hk = pre * exp(-prim[3] * v2);
    double bk = hk * bf;
    double qdot = dvn * q[0] + dvt * q[1];
    double shH = 2.0 * prim[3] * v2 - 4.0, shB = shH + 2.0 / g.ck;
    *eh = hk + coeff * qdot * shH * hk;
    *eb = bk + coeff * qdot * shB * bk;
}

// Gram matrix of the weighted moment projection (upper triangle, 10 entries)
__device__ inline void gram_add(double G[10], double w, double u, double v, double wh, double wb) {
    double k[4] = {1.0, u, v, 0.5 * (u * u + v * v)};
    int n = 0;
    for (int a = 0; a < 4; ++a)
        for (int c = a; c < 4; ++c) {
            double kb = (a == 3 && c == 3) ? 0.25 : 0.0;
            G[n++] += w * (wh * k[a] * k[c] + wb * kb);
        }
}

__device__ inline void gram_full(const double G[10], double M[16]) {
    int n = 0;
    for (int a = 0; a < 4; ++a)
        for (int c = a; c < 4; ++c) { M[c * 4 + a] = G[n]; M[a * 4 + c] = G[n]; ++n; }
}

template <int KPT, int NT>
__global__ void __launch_bounds__(NT) k_predict(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[(NT / 16) * 20];
    const int c = blockIdx.x, nvel = D.nvel;
    if (c >= D.ncell) return;
    const Gas& g = D.gas;
    double u[KPT], v[KPT], w[KPT], h[KPT], b[KPT];
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        int k = threadIdx.x + NT * r;
        u[r] = D.u[k]; v[r] = D.v[k]; w[r] = D.w[k];
        h[r] = D.h[(size_t)c * nvel + k]; b[r] = D.b[(size_t)c * nvel + k];
    }
    double W[4], Wt[4], po[4], pt[4];
    for (int m = 0; m < 4; ++m) { W[m] = D.W[4 * c + m]; Wt[m] = D.w_tilde[4 * c + m]; }
    primary(g.gamma, W, po);
    primary(g.gamma, Wt, pt);
    // frozen heat flux of the current state (GetHeatFlux_Vec)
    double q[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        double dvn = u[r] - po[1], dvt = v[r] - po[2], v2 = dvn * dvn + dvt * dvt;
        q[0] += 0.5 * w[r] * dvn * (v2 * h[r] + b[r]);
        q[1] += 0.5 * w[r] * dvt * (v2 * h[r] + b[r]);
    }
    block_sum<2>(q, lds);
    const double tt = tau_of_prim(g, pt), to = tau_of_prim(g, po);
    // unprojected targets: moments and magnitudes
    double s8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, mx[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        double eh, eb, lh, lb;
        maxwell_shakhov(g, pt, q, u[r], v[r], &eh, &eb);
        maxwell_shakhov(g, po, q, u[r], v[r], &lh, &lb);
        mom_add(s8, w[r], u[r], v[r], eh, eb);
        mom_add(s8 + 4, w[r], u[r], v[r], lh, lb);
        mx[0] = fmax(mx[0], fabs(eh)); mx[1] = fmax(mx[1], fabs(eb));
        mx[2] = fmax(mx[2], fabs(lh)); mx[3] = fmax(mx[3], fabs(lb));
    }
    block_sum<8>(s8, lds);
    block_max<4>(mx, lds);
    double sc[4];
    for (int a = 0; a < 4; ++a) sc[a] = fmax(mx[a], 1.0) * UGKS_DBL_EPS;
    double G[20];
    for (int a = 0; a < 20; ++a) G[a] = 0.0;
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        double eh, eb, lh, lb;
        maxwell_shakhov(g, pt, q, u[r], v[r], &eh, &eb);
        maxwell_shakhov(g, po, q, u[r], v[r], &lh, &lb);
        gram_add(G, w[r], u[r], v[r], fmax(fabs(eh), sc[0]), fmax(fabs(eb), sc[1]));
        gram_add(G + 10, w[r], u[r], v[r], fmax(fabs(lh), sc[2]), fmax(fabs(lb), sc[3]));
    }
    block_sum<20>(G, lds);
    double Me[16], Ml[16], re[4], rl[4], le[4], ll[4];
    gram_full(G, Me);
    gram_full(G + 10, Ml);
    for (int m = 0; m < 4; ++m) { re[m] = Wt[m] - s8[m]; rl[m] = W[m] - s8[4 + m]; }
    unsigned st = 0;
    if (!solve4_gj(Me, re, le) || !solve4_gj(Ml, rl, ll)) st = ST_PROJECT;
    double chk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool fin = true;
    const double dtp = D.scal[SC_DTP];
    (void)dtp;
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
        int k = threadIdx.x + NT * r;
        double eh, eb, lh, lb;
        maxwell_shakhov(g, pt, q, u[r], v[r], &eh, &eb);
        maxwell_shakhov(g, po, q, u[r], v[r], &lh, &lb);
        double kh = 0.5 * (u[r] * u[r] + v[r] * v[r]);
        double weh = fmax(fabs(eh), sc[0]), web = fmax(fabs(eb), sc[1]);
        double wlh = fmax(fabs(lh), sc[2]), wlb = fmax(fabs(lb), sc[3]);
        eh = eh + weh * (le[0] + u[r] * le[1] + v[r] * le[2] + kh * le[3]);
        eb = eb + 0.5 * web * le[3];
        lh = lh + wlh * (ll[0] + u[r] * ll[1] + v[r] * ll[2] + kh * ll[3]);
        lb = lb + 0.5 * wlb * ll[3];
        mom_add(chk, w[r], u[r], v[r], eh, eb);
        mom_add(chk + 4, w[r], u[r], v[r], lh, lb);
        size_t o = (size_t)c * nvel + k;
        double rh = D.rh[o] + (eh - h[r]) / tt - (lh - h[r]) / to;
        double rb = D.rb[o] + (eb - b[r]) / tt - (lb - b[r]) / to;
        fin = fin && isfinite(rh) && isfinite(rb);
        D.rh[o] = rh;
        D.rb[o] = rb;
    }
    block_sum<8>(chk, lds);
    if (!__syncthreads_and(fin ? 1 : 0)) st |= ST_NONFINITE;
    if (threadIdx.x == 0) {
        double se = 1.0, sl = 1.0, ee = 0.0, el = 0.0;
        for (int m = 0; m < 4; ++m) {
            se = fmax(se, fabs(Wt[m])); sl = fmax(sl, fabs(W[m]));
            ee = fmax(ee, fabs(chk[m] - Wt[m])); el = fmax(el, fabs(chk[4 + m] - W[m]));
        }
        if (!(ee <= 1.0e-10 * se) || !(el <= 1.0e-10 * sl)) st |= ST_PROJECT;
        if (!(tt > 0.0)) st |= ST_NONFINITE;
        D.tau_t[c] = tt;
        if (st) atomicOr(D.status, st);
    }
}

// ===========================================================================
// micro LU-SGS sweeps
// ===========================================================================
// per-velocity diagonal and upwind coefficients of cell c
struct MicroCoef {
    double diag_base;   // 1/dtp (1/tau_tilde added per cell)
};

// forward sweep: delta = (R + upwind new neighbours)/D; neighbours across the
// x seam and on the old side are zero (reset), the wall closure of the
// forward sweep is frozen from a zero increment
__global__ void __launch_bounds__(256) k_sweep_fwd(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    UGKS_DYN_SHARED(double, sm);   // [2][2][nx][T]
    const int T = D.tile, nx = D.nx, ny = D.ny, nvel = D.nvel;
    const int lane = threadIdx.x % T, slot = threadIdx.x / T, nslot = blockDim.x / T;
    const int k = blockIdx.x * T + lane;
    const double uk = D.u[k], vk = D.v[k], wk = D.w[k];
    const double area = D.dx * D.dy, dtp = D.scal[SC_DTP];
    double* buf[2] = {sm, sm + 2 * nx * T};
    const int nd = nx + ny - 1;
    double pb0 = 0.0, pb1 = 0.0;   // incidence partials of the owned wall cells are written directly
    (void)pb0; (void)pb1;
    for (int d = 0; d < nd; ++d) {
        double* cur = buf[d & 1];
        const double* prv = buf[(d + 1) & 1];
        const int ilo = d - (ny - 1) > 0 ? d - (ny - 1) : 0, ihi = d < nx - 1 ? d : nx - 1;
        // every thread runs the same trip count so the lane-group
        // reductions below execute with converged workgroups
        for (int base = ilo; base <= ihi; base += nslot) {
            const int i = base + slot;
            const bool act = i <= ihi;
            const int j = d - i, c = j * nx + i;
            double part = 0.0;
            if (act) {
            const size_t o = (size_t)c * nvel + k;
            const double tt = D.tau_t[c];
            double rh = D.rh[o], rb = D.rb[o];
            double diag = 1.0 / dtp + 1.0 / tt;
            // west (un = -u): new value of (i-1,j), zero halo at the seam
            {
                double un = -uk, nh = 0.0, nb = 0.0;
                if (i > 0) { nh = prv[(i - 1) * T + lane]; nb = prv[nx * T + (i - 1) * T + lane]; }
                diag = diag + fmax(un, 0.0) * D.dy / area;
                rh = rh - fmin(un, 0.0) * D.dy * nh / area;
                rb = rb - fmin(un, 0.0) * D.dy * nb / area;
            }
            { double un = uk; diag = diag + fmax(un, 0.0) * D.dy / area; }   // east: old value 0
            {   // south (un = -v): new value of (i,j-1); wall closure of a zero increment
                double un = -vk, nh = 0.0, nb = 0.0;
                if (j > 0) { nh = prv[i * T + lane]; nb = prv[nx * T + i * T + lane]; }
                diag = diag + fmax(un, 0.0) * D.dx / area;
                rh = rh - fmin(un, 0.0) * D.dx * nh / area;
                rb = rb - fmin(un, 0.0) * D.dx * nb / area;
            }
            { double un = vk; diag = diag + fmax(un, 0.0) * D.dx / area; }   // north: old value 0
            double dh = rh / diag, db = rb / diag;
            cur[i * T + lane] = dh;
            cur[nx * T + i * T + lane] = db;
            D.dh[o] = dh;
            D.db[o] = db;
            // incidence partial of the wall closure used by the backward sweep
            if (j == 0 || j == ny - 1) {
                const double vn = (j == ny - 1) ? vk : -vk;
                part = (vn < 0.0) ? 0.0 : wk * vn * dh;
            }
            }
            for (int m = T >> 1; m >= 1; m >>= 1) part += __shfl_xor(part, m, 64);
            if (act && lane == 0) {
                // a one-row grid has both walls on the same cells
                if (j == 0) D.wallpart[((size_t)blockIdx.x * 2 + 0) * nx + i] = part;
                if (j == ny - 1) D.wallpart[((size_t)blockIdx.x * 2 + 1) * nx + i] = part;
            }
        }
        __syncthreads();
    }
}

// backward sweep with the apply step: f' = f + delta (alpha = 1), per-tile
// moment partials of f' for the macro recovery
__global__ void __launch_bounds__(256) k_sweep_bwd(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    UGKS_DYN_SHARED(double, sm);   // [2][2][nx][T] + [2][nx]
    const int T = D.tile, nx = D.nx, ny = D.ny, nvel = D.nvel, ck = D.gas.ck;
    const int lane = threadIdx.x % T, slot = threadIdx.x / T, nslot = blockDim.x / T;
    const int k = blockIdx.x * T + lane;
    const double uk = D.u[k], vk = D.v[k], wk = D.w[k];
    const double area = D.dx * D.dy, dtp = D.scal[SC_DTP];
    double* buf[2] = {sm, sm + 2 * nx * T};
    double* drho = sm + 4 * nx * T;   // [2][nx]
    for (int t = threadIdx.x; t < 2 * nx; t += blockDim.x) {
        int top = t / nx, i = t % nx;
        double inc = 0.0;
        for (int q = 0; q < D.ntiles; ++q) inc += D.wallpart[((size_t)q * 2 + top) * nx + i];
        drho[t] = -inc / D.scal[top ? SC_REFL_TOP : SC_REFL_BOT];
    }
    __syncthreads();
    // unit-density wall Maxwellians in each wall's local frame
    double hu[2], bu[2];
    for (int top = 0; top < 2; ++top) {
        double lam = 1.0 / D.wall_tn[top], nyo = top ? 1.0 : -1.0;
        double vn = vk * nyo, vt = -uk * nyo, pl2 = -D.wall_ux[top] * nyo;
        double dvt = vt - pl2;
        hu[top] = 1.0 * (lam / UGKS_PI) * exp(-lam * (vn * vn + dvt * dvt));
        bu[top] = hu[top] * (ck / (2.0 * lam));
    }
    const int nd = nx + ny - 1;
    bool fin = true;
    for (int dd = 0; dd < nd; ++dd) {
        const int d = nd - 1 - dd;
        double* cur = buf[dd & 1];
        const double* prv = buf[(dd + 1) & 1];
        const int ilo = d - (ny - 1) > 0 ? d - (ny - 1) : 0, ihi = d < nx - 1 ? d : nx - 1;
        for (int base = ilo; base <= ihi; base += nslot) {
            const int i = base + slot;
            const bool act = i <= ihi;
            const int j = d - i, c = j * nx + i;
            double m[4] = {0, 0, 0, 0};
            if (act) {
            const size_t o = (size_t)c * nvel + k;
            const double tt = D.tau_t[c];
            double rh = D.rh[o], rb = D.rb[o];
            double diag = 1.0 / dtp + 1.0 / tt;
            {   // west: forward value of (i-1,j); across the seam the halo holds (nx-1,j)'s forward value
                double un = -uk;
                size_t on = (size_t)(j * nx + (i > 0 ? i - 1 : nx - 1)) * nvel + k;
                double nh = D.dh[on], nb = D.db[on];
                diag = diag + fmax(un, 0.0) * D.dy / area;
                rh = rh - fmin(un, 0.0) * D.dy * nh / area;
                rb = rb - fmin(un, 0.0) * D.dy * nb / area;
            }
            {   // east: backward value of (i+1,j); across the seam the halo holds (0,j)'s forward value
                double un = uk, nh, nb;
                if (i < nx - 1) { nh = prv[(i + 1) * T + lane]; nb = prv[nx * T + (i + 1) * T + lane]; }
                else { size_t on = (size_t)(j * nx) * nvel + k; nh = D.dh[on]; nb = D.db[on]; }
                diag = diag + fmax(un, 0.0) * D.dy / area;
                rh = rh - fmin(un, 0.0) * D.dy * nh / area;
                rb = rb - fmin(un, 0.0) * D.dy * nb / area;
            }
            {   // south: forward value of (i,j-1), or the frozen bottom wall closure
                double un = -vk, nh, nb;
                if (j > 0) { size_t on = (size_t)((j - 1) * nx + i) * nvel + k; nh = D.dh[on]; nb = D.db[on]; }
                else {
                    bool sel = un < 0.0;
                    nh = sel ? drho[i] * hu[0] : D.dh[o];
                    nb = sel ? drho[i] * bu[0] : D.db[o];
                }
                diag = diag + fmax(un, 0.0) * D.dx / area;
                rh = rh - fmin(un, 0.0) * D.dx * nh / area;
                rb = rb - fmin(un, 0.0) * D.dx * nb / area;
            }
            {   // north: backward value of (i,j+1), or the frozen top wall closure
                double un = vk, nh, nb;
                if (j < ny - 1) { nh = prv[i * T + lane]; nb = prv[nx * T + i * T + lane]; }
                else {
                    bool sel = un < 0.0;
                    nh = sel ? drho[nx + i] * hu[1] : D.dh[o];
                    nb = sel ? drho[nx + i] * bu[1] : D.db[o];
                }
                diag = diag + fmax(un, 0.0) * D.dx / area;
                rh = rh - fmin(un, 0.0) * D.dx * nh / area;
                rb = rb - fmin(un, 0.0) * D.dx * nb / area;
            }
            double dh = rh / diag, db = rb / diag;
            cur[i * T + lane] = dh;
            cur[nx * T + i * T + lane] = db;
            double hn = D.h[o] + 1.0 * dh, bn = D.b[o] + 1.0 * db;
            fin = fin && isfinite(hn) && isfinite(bn);
            D.hn[o] = hn;
            D.bn[o] = bn;
            mom_add(m, wk, uk, vk, hn, bn);
            }
            for (int q = 0; q < 4; ++q)
                for (int s = T >> 1; s >= 1; s >>= 1) m[q] += __shfl_xor(m[q], s, 64);
            if (act && lane == 0)
                for (int q = 0; q < 4; ++q) D.mompart[((size_t)blockIdx.x * D.ncell + c) * 4 + q] = m[q];
        }
        __syncthreads();
    }
    if (!fin) atomicOr(D.status, (unsigned)ST_MICRO_ALPHA);
}

// reflection normalisers of the frozen wall closure (depend on the table only)
__global__ void __launch_bounds__(256) k_wall_reflection(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[16 * 4];
    double s[4] = {0, 0, 0, 0};
    for (int k = threadIdx.x; k < D.nvel; k += blockDim.x) {
        for (int top = 0; top < 2; ++top) {
            double lam = 1.0 / D.wall_tn[top], nyo = top ? 1.0 : -1.0;
            double vn = D.v[k] * nyo, vt = -D.u[k] * nyo, pl2 = -D.wall_ux[top] * nyo;
            double dvt = vt - pl2;
            double hu = 1.0 * (lam / UGKS_PI) * exp(-lam * (vn * vn + dvt * dvt));
            if (vn < 0.0) {
                s[top] += D.w[k] * vn * hu;
                s[2 + top] += fabs(D.w[k] * vn * hu);
            }
        }
    }
    block_sum<4>(s, lds);
    if (threadIdx.x == 0) {
        D.scal[SC_REFL_BOT] = s[0];
        D.scal[SC_REFL_TOP] = s[1];
        for (int top = 0; top < 2; ++top)
            if (!(fabs(s[top]) > 1000.0 * UGKS_DBL_EPS * fmax(s[2 + top], UGKS_DBL_MIN))) atomicOr(D.status, (unsigned)ST_WALL);
    }
}

// macro recovery of the proposal, admissibility, mass of the proposal
__global__ void __launch_bounds__(256) k_recover(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.ncell) return;
    double m[4] = {0, 0, 0, 0};
    for (int q = 0; q < D.ntiles; ++q)
        for (int a = 0; a < 4; ++a) m[a] += D.mompart[((size_t)q * D.ncell + c) * 4 + a];
    for (int a = 0; a < 4; ++a) D.Wn[4 * c + a] = m[a];
    if (!physical_w(m)) atomicOr(D.status, (unsigned)ST_MICRO_ALPHA);
}

__global__ void __launch_bounds__(256) k_mass(const Dev* __restrict__ DS, int rescale) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[16 * 4];
    double s[4] = {0.0, 0.0, 0.0, 0.0}, s0[1] = {0.0};
    const double area = D.dx * D.dy;
    for (int c = threadIdx.x; c < D.ncell; c += blockDim.x) {
        for (int a = 0; a < 4; ++a) s[a] += area * D.Wn[4 * c + a];
        s0[0] += area * D.W[4 * c];
    }
    block_sum<4>(s, lds);
    block_sum<1>(s0, lds);
    if (threadIdx.x == 0) {
        if (D.scal[SC_M0_INIT] == 0.0) { D.scal[SC_M0] = s0[0]; D.scal[SC_M0_INIT] = 1.0; }
        D.scal[SC_MASS_BEFORE] = s[0];
        for (int a = 0; a < 4; ++a) D.scal[SC_MB4 + a] = s[a];
        D.scal[SC_SCALE] = rescale ? D.scal[SC_M0] / s[0] : 1.0;
    }
}

__global__ void __launch_bounds__(256) k_rescale(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    const size_t n = (size_t)D.ncell * D.nvel;
    const double c = D.scal[SC_SCALE];
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        D.hn[t] = c * D.hn[t];
        D.bn[t] = c * D.bn[t];
    }
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < 4 * (size_t)D.ncell;
         t += (size_t)gridDim.x * blockDim.x)
        D.Wn[t] = c * D.Wn[t];
}

// totals after the rescale; the correction ledger the proposal would carry
// (correction_ledger + mass_after(1:4) - mass_before(1:4),
// src/mod_ClosedCouetteMass.f90:284-310), committed on acceptance
__global__ void __launch_bounds__(256) k_mass_after(const Dev* __restrict__ DS) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[16 * 4];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < D.ncell; c += blockDim.x)
        for (int a = 0; a < 4; ++a) s[a] += D.dx * D.dy * D.Wn[4 * c + a];
    block_sum<4>(s, lds);
    if (threadIdx.x == 0) {
        D.scal[SC_MASS_AFTER] = s[0];
        for (int a = 0; a < 4; ++a) D.scal[SC_LEDGER_NEXT + a] = D.scal[SC_LEDGER + a] + (s[a] - D.scal[SC_MB4 + a]);
    }
}

// ===========================================================================
// MassStatistics of the accepted state (src/mod_ClosedCouetteMass.f90:173-204)
// ===========================================================================
// Per cell (one workgroup): the discrete moments mm of (h, b)
// (MomentConserved, src/mod_ImplicitCommon.f90:32-40), the moment error
// max|W - mm| / max(|W|, 1), the negative-mass fractions of h and b, rho and
// the internal energy, and the cell's admissibility.  part[c*12 + ...]:
// A*W(4), A*mm(1), moment error, nh, nb, rho (or +inf), e (or +inf), ok.
__global__ void __launch_bounds__(256) k_mass_stats_cell(const Dev* __restrict__ DS, double* __restrict__ part) {
    const Dev& D = DS[blockIdx.z];
    const int c = blockIdx.x;
    const int nv = D.nvel;
    __shared__ double lds[16 * 9];
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double* h = D.h + (size_t)c * nv;
    const double* b = D.b + (size_t)c * nv;
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        const double w = D.w[k], u = D.u[k], v = D.v[k], hk = h[k], bk = b[k];
        s[0] += w * hk;
        s[1] += w * u * hk;
        s[2] += w * v * hk;
        s[3] += w * ((u * u + v * v) * hk + bk);
        s[4] += w * fmax(-hk, 0.0);
        s[5] += w * fabs(hk);
        s[6] += w * fmax(-bk, 0.0);
        s[7] += w * fabs(bk);
        s[8] += (isfinite(hk) && isfinite(bk)) ? 0.0 : 1.0;
    }
    block_sum<9>(s, lds);
    if (threadIdx.x != 0) return;
    const double area = D.dx * D.dy;
    double W[4], mm[4] = {s[0], s[1], s[2], 0.5 * s[3]};
    for (int a = 0; a < 4; ++a) W[a] = D.W[4 * (size_t)c + a];
    double err = 0.0;
    for (int a = 0; a < 4; ++a) err = fmax(err, fabs(W[a] - mm[a]) / fmax(fabs(W[a]), 1.0));
    double* o = part + (size_t)c * 12;
    for (int a = 0; a < 4; ++a) o[a] = area * W[a];
    o[4] = area * mm[0];
    o[5] = err;
    o[6] = s[4] / fmax(s[5], UGKS_DBL_MIN);
    o[7] = s[6] / fmax(s[7], UGKS_DBL_MIN);
    const bool valid = isfinite(W[0]) && W[0] > 0.0;
    o[8] = valid ? W[0] : __builtin_inf();
    o[9] = valid ? W[3] - 0.5 * (W[1] * W[1] + W[2] * W[2]) / W[0] : __builtin_inf();
    o[10] = (s[8] == 0.0 && physical_w(W) && physical_w(mm) && area > 0.0) ? 1.0 : 0.0;
}

// totals(5), metrics(5) and ok into scal[SC_MSTAT..SC_MSTAT+10] (one workgroup)
__global__ void __launch_bounds__(256) k_mass_stats_reduce(const Dev* __restrict__ DS, const double* __restrict__ part,
                                                             double moment_tol) {
    const Dev& D = DS[blockIdx.z];
    __shared__ double lds[16 * 5];
    double t[5] = {0, 0, 0, 0, 0}, mx[5] = {0.0, 0.0, 0.0, -__builtin_inf(), -__builtin_inf()}, bad[1] = {0.0};
    for (int c = threadIdx.x; c < D.ncell; c += blockDim.x) {
        const double* o = part + (size_t)c * 12;
        for (int a = 0; a < 5; ++a) t[a] += o[a];
        mx[0] = fmax(mx[0], o[5]);
        mx[1] = fmax(mx[1], o[6]);
        mx[2] = fmax(mx[2], o[7]);
        mx[3] = fmax(mx[3], -o[8]);
        mx[4] = fmax(mx[4], -o[9]);
        bad[0] += 1.0 - o[10];
    }
    block_sum<5>(t, lds);
    block_max<5>(mx, lds);
    block_sum<1>(bad, lds);
    if (threadIdx.x != 0) return;
    double* S = D.scal + SC_MSTAT;
    bool ok = bad[0] == 0.0;
    for (int a = 0; a < 5; ++a) { S[a] = t[a]; ok = ok && isfinite(t[a]); }
    S[5] = mx[0]; S[6] = mx[1]; S[7] = mx[2]; S[8] = -mx[3]; S[9] = -mx[4];
    for (int a = 5; a < 10; ++a) ok = ok && isfinite(S[a]);
    S[10] = (ok && S[5] <= moment_tol) ? 1.0 : 0.0;
}

extern "C" hipError_t ugks_launch_mass_stats(const LaunchSet* L, double* part, double moment_tol, hipStream_t s) {
    const Dev& D = *L->size;
    hipLaunchKernelGGL(k_mass_stats_cell, dim3(D.ncell, 1, L->n), dim3(256), 0, s, L->dev, part);
    hipLaunchKernelGGL(k_mass_stats_reduce, dim3(1, 1, L->n), dim3(256), 0, s, L->dev, part, moment_tol);
    return hipGetLastError();
}

extern "C" hipError_t ugks_launch_wall_reflection(const LaunchSet* L, hipStream_t s, Prof* P) {
    UGKS_TIMED(P, K_WALL_REFLECTION, s,
               hipLaunchKernelGGL(k_wall_reflection, dim3(1, 1, L->n), dim3(256), 0, s, L->dev));
    return hipGetLastError();
}

// The implicit update of every instance in L (the operator of the current
// state must be in place).  The proposal goes to Wn/hn/bn.
extern "C" hipError_t ugks_launch_implicit(const LaunchSet* L, int rescale, hipStream_t s, Prof* P) {
    const Dev& D = *L->size;
    const int n = L->n;
    const Dev* DS = L->dev;
    int ng = (D.ncell + (D.ny + 1) * D.nx + 255) / 256;
    UGKS_TIMED(P, K_MACRO_GAMMA, s, hipLaunchKernelGGL(k_macro_gamma, dim3(ng, 1, n), dim3(256), 0, s, DS));
    UGKS_TIMED(P, K_MACRO_PR2, s, hipLaunchKernelGGL(k_macro_pr2, dim3(1, 1, n), dim3(256), 0, s, DS));
    switch (D.nvel) {
    case 256:
        UGKS_TIMED(P, K_PREDICT, s, hipLaunchKernelGGL((k_predict<4, 64>), dim3(D.ncell, 1, n), dim3(64), 0, s, DS));
        break;
    case 1024:
        UGKS_TIMED(P, K_PREDICT, s, hipLaunchKernelGGL((k_predict<4, 256>), dim3(D.ncell, 1, n), dim3(256), 0, s, DS));
        break;
    case 4096:
        UGKS_TIMED(P, K_PREDICT, s, hipLaunchKernelGGL((k_predict<8, 512>), dim3(D.ncell, 1, n), dim3(512), 0, s, DS));
        break;
    default: return hipErrorInvalidValue;
    }
    size_t shf = sizeof(double) * 4 * D.nx * D.tile;
    size_t shb = shf + sizeof(double) * 2 * D.nx;
    UGKS_TIMED(P, K_SWEEP_FWD, s, hipLaunchKernelGGL(k_sweep_fwd, dim3(D.ntiles, 1, n), dim3(256), shf, s, DS));
    UGKS_TIMED(P, K_SWEEP_BWD, s, hipLaunchKernelGGL(k_sweep_bwd, dim3(D.ntiles, 1, n), dim3(256), shb, s, DS));
    UGKS_TIMED(P, K_RECOVER, s, hipLaunchKernelGGL(k_recover, dim3((D.ncell + 255) / 256, 1, n), dim3(256), 0, s, DS));
    UGKS_TIMED(P, K_MASS, s, hipLaunchKernelGGL(k_mass, dim3(1, 1, n), dim3(256), 0, s, DS, rescale));
    UGKS_TIMED(P, K_RESCALE, s, hipLaunchKernelGGL(k_rescale, dim3(1024, 1, n), dim3(256), 0, s, DS));
    UGKS_TIMED(P, K_MASS_AFTER, s, hipLaunchKernelGGL(k_mass_after, dim3(1, 1, n), dim3(256), 0, s, DS));
    return hipGetLastError();
}

// ===========================================================================
// batched instances: device-side acceptance of an iteration
// ===========================================================================
// An instance whose status word is clear swaps its state and proposal buffers
// (the pointer swap the single-instance host does after an accepted update)
// and counts the iteration; any status bit leaves its accepted state as it
// is, so a rejected instance keeps re-evaluating the same state.  One thread
// per instance.
__global__ void __launch_bounds__(64) k_batch_accept(Dev* DS, int n, int* count, unsigned* status_out) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= n) return;
    Dev& D = DS[z];
    const unsigned st = *D.status;
    status_out[z] = st;
    if (st == 0) {
        double* t;
        t = D.W; D.W = D.Wn; D.Wn = t;
        t = D.h; D.h = D.hn; D.hn = t;
        t = D.b; D.b = D.bn; D.bn = t;
        for (int a = 0; a < 4; ++a) D.scal[SC_LEDGER + a] = D.scal[SC_LEDGER_NEXT + a];
        count[z] += 1;
    }
}

extern "C" hipError_t ugks_launch_batch_accept(Dev* DS, int n, int* count, unsigned* status_out, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_accept, dim3((n + 63) / 64), dim3(64), 0, s, DS, n, count, status_out);
    return hipGetLastError();
}

// zero n ints (a kernel, not a memset node, so that it may sit in a graph)
__global__ void __launch_bounds__(64) k_zero_ints(int* p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

extern "C" hipError_t ugks_launch_zero_ints(int* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}
