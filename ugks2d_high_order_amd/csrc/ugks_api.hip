// C ABI of the gfx950 HighOrderImplicit step (declared in include/ugks_hip.h).
// Host code only: device memory management, the launch sequence of one
// pseudo-time iteration, and the fail-stop checks of the reference's gates.
#include "../../include/ugks_hip.h"

#ifndef UGKS_EMULATE
#include <hip/hip_runtime.h>
#endif

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ugks_state.h"
#include "ugks_prof.h"

extern "C" hipError_t ugks_set_consts(const Consts* k);
extern "C" hipError_t ugks_launch_operator(const Dev* D, double umax, double vmax, double residual_tol, hipStream_t s,
                                           Prof* P);
extern "C" hipError_t ugks_launch_implicit(const Dev* D, int rescale, hipStream_t s, Prof* P);
extern "C" hipError_t ugks_launch_wall_reflection(const Dev* D, hipStream_t s, Prof* P);
extern "C" hipError_t ugks_launch_cn(const Dev* D, int build_old, double umax, double vmax, hipStream_t s, Prof* P);
extern "C" hipError_t ugks_launch_cold_start(const Dev* D, double rho, double ux, double tn, hipStream_t s);
extern "C" hipError_t ugks_launch_observables(const Dev* D, double* out, hipStream_t s);
extern "C" hipError_t ugks_launch_lo_step(const Dev* D, int kpt, double umax, double vmax, hipStream_t s, Prof* P);
extern "C" hipError_t ugks_launch_lo_cold_start(const Dev* D, double rho, double ux, double tn, hipStream_t s);

namespace {

thread_local std::string g_err;

enum { E_OK = 0, E_NONFINITE = 1, E_GATE = 2, E_INADMISSIBLE = 3, E_POLICY = 4, E_NEWTON = 5, E_DEVICE = 6, E_ARG = 7 };

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return fail(E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Instance {
    ugks_params_t p;
    Dev D;
    hipStream_t stream = nullptr;
    double umax = 0, vmax = 0;
    bool operator_fresh = false;
    std::vector<void*> allocs;
    double* obs = nullptr;
    Prof prof;
    bool cn_alloc = false;        // S2CN2 target buffers allocated
    bool cn_targets_fresh = false; // tg holds the volume targets of the current state
    int lo_wg = 256;              // workgroup size of the low-order Explicit kernels
};

// Polynomial coefficients (degrees 0..n-1) whose averages over the unit cells
// at the given offsets are the samples: solve the cell-average Vandermonde
// system in long double for each unit sample.
void average_map(int n, const int* off, double map[5][5]) {
    for (int s = 0; s < 5; ++s)
        for (int d = 0; d < 5; ++d) map[s][d] = 0.0;
    long double A[5][5];
    for (int r = 0; r < n; ++r)
        for (int d = 0; d < n; ++d) {
            long double hi = off[r] + 0.5L, lo = off[r] - 0.5L, ph = 1.0L, pl = 1.0L;
            for (int e = 0; e <= d; ++e) { ph *= hi; pl *= lo; }
            A[r][d] = (ph - pl) / (long double)(d + 1);
        }
    for (int c = 0; c < n; ++c) {
        long double M[5][6];
        for (int r = 0; r < n; ++r) {
            for (int d = 0; d < n; ++d) M[r][d] = A[r][d];
            M[r][n] = (r == c) ? 1.0L : 0.0L;
        }
        for (int col = 0; col < n; ++col) {
            int piv = col;
            for (int r = col + 1; r < n; ++r)
                if (fabsl(M[r][col]) > fabsl(M[piv][col])) piv = r;
            for (int q = 0; q <= n; ++q) std::swap(M[col][q], M[piv][q]);
            for (int r = 0; r < n; ++r) {
                if (r == col) continue;
                long double f = M[r][col] / M[col][col];
                for (int q = col; q <= n; ++q) M[r][q] -= f * M[col][q];
            }
        }
        for (int d = 0; d < n; ++d) map[c][d] = (double)(M[d][n] / M[d][d]);
    }
}

int build_consts() {
    Consts K;
    memset(&K, 0, sizeof K);
    K.gnode[0] = -0.5 * std::sqrt(3.0 / 5.0);
    K.gnode[1] = 0.0;
    K.gnode[2] = 0.5 * std::sqrt(3.0 / 5.0);
    K.gw[0] = 5.0 / 18.0;
    K.gw[1] = 4.0 / 9.0;
    K.gw[2] = 5.0 / 18.0;
    int o[4][5] = {{-2, -1, 0}, {-1, 0, 1}, {0, 1, 2}, {-2, -1, 0, 1, 2}};
    double m[5][5];
    for (int c = 0; c < 3; ++c) {
        average_map(3, o[c], m);
        for (int s = 0; s < 3; ++s)
            for (int d = 0; d < 3; ++d) K.cmap[c][s + c][d] = m[s][d];
    }
    average_map(5, o[3], m);
    memcpy(K.cmap[3], m, sizeof m);
    for (int sh = 0; sh < 5; ++sh) {
        int off[5];
        for (int s = 0; s < 5; ++s) off[s] = s - sh;
        average_map(5, off, m);
        memcpy(K.wallmap[sh], m, sizeof m);
    }
    HIPCHK(ugks_set_consts(&K));
    return E_OK;
}

int status_to_code(unsigned st) {
    if (st & ST_NEWTON) return E_NEWTON;
    if (st & (ST_NONFINITE | ST_SOLVE)) return E_NONFINITE;
    if (st & (ST_MACRO_ALPHA | ST_MICRO_ALPHA)) return E_INADMISSIBLE;
    if (st & (ST_GATE | ST_WALL | ST_PROJECT)) return E_GATE;
    return E_OK;
}

std::string status_text(unsigned st) {
    std::string s;
    const char* names[] = {"newton", "gate", "nonfinite", "wall", "macro_alpha", "micro_alpha", "projection", "solve",
                           "in_face_flux", "in_wall_flux", "in_cell", "in_predict", "in_sweep", "in_macro", "in_cold_start"};
    for (int b = 0; b < 15; ++b)
        if (st & (1u << b)) s += std::string(s.empty() ? "" : ",") + names[b];
    return s;
}

template <typename T>
int dmalloc(Instance* I, T** ptr, size_t n) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    HIPCHK(hipMemset(p, 0, sizeof(T) * (n ? n : 1)));
    I->allocs.push_back(p);
    *ptr = (T*)p;
    return E_OK;
}

int check_status(Instance* I, unsigned* out_status) {
    unsigned st = 0;
    HIPCHK(hipMemcpyAsync(&st, I->D.status, sizeof st, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    if (out_status) *out_status = st;
    if (st) {
        unsigned z = 0;
        HIPCHK(hipMemcpyAsync(I->D.status, &z, sizeof z, hipMemcpyHostToDevice, I->stream));
        HIPCHK(hipStreamSynchronize(I->stream));
        return fail(status_to_code(st), "device gate failed: " + status_text(st));
    }
    return E_OK;
}

int run_operator(Instance* I) {
    HIPCHK(ugks_launch_operator(&I->D, I->umax, I->vmax, I->p.residual_tol, I->stream, &I->prof));
    int rc = check_status(I, nullptr);
    if (rc) return rc;
    I->operator_fresh = true;
    return E_OK;
}

}  // namespace

extern "C" const char* ugks_last_error(void) { return g_err.c_str(); }

extern "C" int ugks_create(const ugks_params_t* p, const double* u, const double* v, const double* w, void** handle) {
    if (!p || !u || !v || !w || !handle) return fail(E_ARG, "null argument");
    *handle = nullptr;
    int nvel = p->nu * p->nv;
    if (!(nvel == 256 || nvel == 1024 || nvel == 4096)) return fail(E_ARG, "Nu*Nv must be 256, 1024 or 4096");
    if (p->nx < 5 || p->ny < 5) return fail(E_ARG, "the reconstruction needs at least five cells per axis");
    if (p->ck < 1) return fail(E_ARG, "ck must be positive");
    if (p->macro_sweeps < 1 || p->micro_sweeps != 1)
        return fail(E_ARG, "one micro LU-SGS pair and at least one macro sweep pair are supported");
    for (int k = 0; k < nvel; ++k)
        if (!(std::isfinite(u[k]) && std::isfinite(v[k]) && std::isfinite(w[k]) && w[k] > 0.0))
            return fail(E_ARG, "invalid velocity table");
    int rc = build_consts();
    if (rc) return rc;
    Instance* I = new Instance();
    I->p = *p;
    Dev& D = I->D;
    memset(&D, 0, sizeof D);
    D.nx = p->nx; D.ny = p->ny; D.nu = p->nu; D.nv = p->nv; D.nvel = nvel; D.ncell = p->nx * p->ny;
    D.dx = p->dx; D.dy = p->dy; D.cfl = p->cfl; D.ncfl = p->ncfl;
    D.gas.ck = p->ck; D.gas.pr = p->pr; D.gas.omega = p->omega; D.gas.mu_ref = p->mu_ref;
    D.gas.gamma = (double)(p->ck + 4) / (double)(p->ck + 2);
    D.wall_ux[0] = p->wall_ux[0]; D.wall_ux[1] = p->wall_ux[1];
    D.wall_tn[0] = p->wall_tn[0]; D.wall_tn[1] = p->wall_tn[1];
    D.macro_sweeps = p->macro_sweeps;
    D.tile = 16;
    D.wg_vel = 128;
    D.ntiles = nvel / D.tile;
    I->umax = -1e300; I->vmax = -1e300;
    for (int k = 0; k < nvel; ++k) { I->umax = std::max(I->umax, u[k]); I->vmax = std::max(I->vmax, v[k]); }
    const size_t nc = D.ncell, cv = nc * nvel, nx = D.nx, ny = D.ny;
    auto bail = [&](int code) { ugks_destroy(I); return code; };
    if (hipStreamCreate(&I->stream) != hipSuccess) { delete I; return fail(E_DEVICE, "hipStreamCreate failed"); }
    double *du, *dv, *dw;
    if ((rc = dmalloc(I, &du, nvel)) || (rc = dmalloc(I, &dv, nvel)) || (rc = dmalloc(I, &dw, nvel))) return bail(rc);
    D.u = du; D.v = dv; D.w = dw;
    double** arrs[] = {&D.W, &D.Wn, &D.rate_w, &D.dW, &D.dWs, &D.w_tilde};
    for (auto a : arrs) if ((rc = dmalloc(I, a, 4 * nc))) return bail(rc);
    double** cvs[] = {&D.h, &D.b, &D.hn, &D.bn, &D.xfl_h, &D.xfl_b, &D.rh, &D.rb, &D.dh, &D.db};
    for (auto a : cvs) if ((rc = dmalloc(I, a, cv))) return bail(rc);
    if ((rc = dmalloc(I, &D.yfl_h, (ny + 1) * nx * nvel)) || (rc = dmalloc(I, &D.yfl_b, (ny + 1) * nx * nvel)) ||
        (rc = dmalloc(I, &D.xfw, 4 * nc)) || (rc = dmalloc(I, &D.yfw, 4 * (ny + 1) * nx)) ||
        (rc = dmalloc(I, &D.xf, nc * REC_ROWS * nvel)) || (rc = dmalloc(I, &D.yf, nx * (ny - 1) * REC_ROWS * nvel)) ||
        (rc = dmalloc(I, &D.wb, nx * REC_ROWS * nvel)) || (rc = dmalloc(I, &D.wt, nx * REC_ROWS * nvel)) ||
        (rc = dmalloc(I, &D.ws, nc * 4 * 24)) || (rc = dmalloc(I, &D.wv, nc * 4 * 9)) ||
        (rc = dmalloc(I, &D.cellpart, nc * 16)) || (rc = dmalloc(I, &D.gx, nc)) ||
        (rc = dmalloc(I, &D.gy, (ny + 1) * nx)) || (rc = dmalloc(I, &D.tau_t, nc)) ||
        (rc = dmalloc(I, &D.wallpart, (size_t)D.ntiles * 2 * nx)) ||
        (rc = dmalloc(I, &D.mompart, (size_t)D.ntiles * nc * 4)) || (rc = dmalloc(I, &D.scal, SC_COUNT)) ||
        (rc = dmalloc(I, &D.status, 1)) || (rc = dmalloc(I, &I->obs, nc * 10)))
        return bail(rc);
    if (hipMemcpy(du, u, sizeof(double) * nvel, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dv, v, sizeof(double) * nvel, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dw, w, sizeof(double) * nvel, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(E_DEVICE, "velocity table upload failed"));
    if (ugks_launch_wall_reflection(&D, I->stream, &I->prof) != hipSuccess) return bail(fail(E_DEVICE, "launch failed"));
    if ((rc = check_status(I, nullptr))) return bail(rc);
    *handle = I;
    return E_OK;
}

extern "C" void ugks_destroy(void* handle) {
    Instance* I = (Instance*)handle;
    if (!I) return;
    if (I->stream) hipStreamSynchronize(I->stream);
    I->prof.release();
    for (void* p : I->allocs) hipFree(p);
    if (I->stream) hipStreamDestroy(I->stream);
    delete I;
}

// a new state starts a new S2CN2 run: its targets and conservation ledger
static int reset_cn(Instance* I) {
    static const double zero = 0.0;
    I->cn_targets_fresh = false;
    HIPCHK(hipMemcpyAsync(I->D.scal + SC_CN_INIT, &zero, sizeof(double), hipMemcpyHostToDevice, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

static int set_state_impl(Instance* I, const double* W, const double* h, const double* b, hipMemcpyKind kind) {
    const size_t nc = I->D.ncell, cv = nc * I->D.nvel;
    HIPCHK(hipMemcpyAsync(I->D.W, W, sizeof(double) * 4 * nc, kind, I->stream));
    HIPCHK(hipMemcpyAsync(I->D.h, h, sizeof(double) * cv, kind, I->stream));
    HIPCHK(hipMemcpyAsync(I->D.b, b, sizeof(double) * cv, kind, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    I->operator_fresh = false;
    return reset_cn(I);
}

static int get_state_impl(Instance* I, double* W, double* h, double* b, hipMemcpyKind kind) {
    const size_t nc = I->D.ncell, cv = nc * I->D.nvel;
    if (W) HIPCHK(hipMemcpyAsync(W, I->D.W, sizeof(double) * 4 * nc, kind, I->stream));
    if (h) HIPCHK(hipMemcpyAsync(h, I->D.h, sizeof(double) * cv, kind, I->stream));
    if (b) HIPCHK(hipMemcpyAsync(b, I->D.b, sizeof(double) * cv, kind, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

extern "C" int ugks_set_state(void* handle, const double* W, const double* h, const double* b) {
    if (!handle || !W || !h || !b) return fail(E_ARG, "null argument");
    return set_state_impl((Instance*)handle, W, h, b, hipMemcpyHostToDevice);
}

extern "C" int ugks_get_state(void* handle, double* W, double* h, double* b) {
    if (!handle) return fail(E_ARG, "null handle");
    return get_state_impl((Instance*)handle, W, h, b, hipMemcpyDeviceToHost);
}

extern "C" int ugks_set_state_device(void* handle, const double* W, const double* h, const double* b) {
    if (!handle || !W || !h || !b) return fail(E_ARG, "null argument");
    return set_state_impl((Instance*)handle, W, h, b, hipMemcpyDeviceToDevice);
}

extern "C" int ugks_get_state_device(void* handle, double* W, double* h, double* b) {
    if (!handle) return fail(E_ARG, "null handle");
    return get_state_impl((Instance*)handle, W, h, b, hipMemcpyDeviceToDevice);
}

extern "C" int ugks_cold_start(void* handle, double rho, double ux, double t_native) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    if (!(rho > 0.0) || !(t_native > 0.0)) return fail(E_ARG, "cold start needs rho > 0 and T > 0");
    HIPCHK(ugks_launch_cold_start(&I->D, rho, ux, t_native, I->stream));
    I->operator_fresh = false;
    int rc = check_status(I, nullptr);
    return rc ? rc : reset_cn(I);
}

extern "C" int ugks_sync(void* handle) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

extern "C" int ugks_evolution(void* handle, double* dt, double* rate_w, double* rate_h, double* rate_b,
                              double* xflux_h, double* xflux_b, double* yflux_h, double* yflux_b) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    int rc = run_operator(I);
    if (rc) return rc;
    const Dev& D = I->D;
    const size_t nc = D.ncell, cv = nc * D.nvel, fy = (size_t)(D.ny + 1) * D.nx * D.nvel;
    if (dt) HIPCHK(hipMemcpyAsync(dt, D.scal + SC_DT, sizeof(double), hipMemcpyDeviceToHost, I->stream));
    if (rate_w) HIPCHK(hipMemcpyAsync(rate_w, D.rate_w, sizeof(double) * 4 * nc, hipMemcpyDeviceToHost, I->stream));
    if (rate_h) HIPCHK(hipMemcpyAsync(rate_h, D.rh, sizeof(double) * cv, hipMemcpyDeviceToHost, I->stream));
    if (rate_b) HIPCHK(hipMemcpyAsync(rate_b, D.rb, sizeof(double) * cv, hipMemcpyDeviceToHost, I->stream));
    if (xflux_h) HIPCHK(hipMemcpyAsync(xflux_h, D.xfl_h, sizeof(double) * cv, hipMemcpyDeviceToHost, I->stream));
    if (xflux_b) HIPCHK(hipMemcpyAsync(xflux_b, D.xfl_b, sizeof(double) * cv, hipMemcpyDeviceToHost, I->stream));
    if (yflux_h) HIPCHK(hipMemcpyAsync(yflux_h, D.yfl_h, sizeof(double) * fy, hipMemcpyDeviceToHost, I->stream));
    if (yflux_b) HIPCHK(hipMemcpyAsync(yflux_b, D.yfl_b, sizeof(double) * fy, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

extern "C" int ugks_norms(void* handle, double* out24, double* walls16) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    if (out24) HIPCHK(hipMemcpyAsync(out24, I->D.scal + SC_NORMS, 24 * sizeof(double), hipMemcpyDeviceToHost, I->stream));
    if (walls16) HIPCHK(hipMemcpyAsync(walls16, I->D.scal + SC_WALLS, 16 * sizeof(double), hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

extern "C" int ugks_implicit_step(void* handle, int* accepted, double* diag16) {
    Instance* I = (Instance*)handle;
    if (!I || !accepted) return fail(E_ARG, "null argument");
    *accepted = 0;
    const ugks_params_t& p = I->p;
    if (p.mass_rescale && !(p.ncfl == 1000.0))
        return fail(E_POLICY, "closed-Couette policy requires nCFL = 1000 (CheckR5Policy)");
    int rc = E_OK;
    if (!I->operator_fresh && (rc = run_operator(I))) return rc;
    HIPCHK(ugks_launch_implicit(&I->D, p.mass_rescale ? 1 : 0, I->stream, &I->prof));
    double sc[SC_COUNT];
    unsigned st = 0;
    HIPCHK(hipMemcpyAsync(sc, I->D.scal, sizeof sc, hipMemcpyDeviceToHost, I->stream));
    rc = check_status(I, &st);
    if (diag16) {
        for (int a = 0; a < 16; ++a) diag16[a] = 0.0;
        diag16[0] = (st & ST_MICRO_ALPHA) ? 0.5 : 1.0;
        diag16[1] = (st & ST_MACRO_ALPHA) ? 0.5 : 1.0;
        diag16[2] = sc[SC_SCALE] - 1.0;
        diag16[4] = sc[SC_MASS_BEFORE];
        diag16[5] = sc[SC_MASS_AFTER];
        diag16[6] = sc[SC_M0];
        diag16[7] = sc[SC_DT];
        diag16[8] = sc[SC_DTP];
        diag16[9] = (double)st;
    }
    if (rc) {
        // the proposal lives in the n-buffers; the accepted state is untouched
        if (st & (ST_MACRO_ALPHA | ST_MICRO_ALPHA)) return fail(E_POLICY, "closed-Couette policy rejected the update (alpha < 1)");
        return rc;
    }
    std::swap(I->D.W, I->D.Wn);
    std::swap(I->D.h, I->D.hn);
    std::swap(I->D.b, I->D.bn);
    I->operator_fresh = false;
    I->cn_targets_fresh = false;
    *accepted = 1;
    return E_OK;
}

// S2CN2 explicit step (no counterpart as one call in the reference: it is
// S2NativeEvolution's fluxes followed by S2NativeApplyUpdate, i.e. advance_step
// with the conservation gate, src/mod_S2NativeAdapter.f90:469-599,867-1018).
extern "C" int ugks_cn_step(void* handle, int* accepted, double* diag16) {
    Instance* I = (Instance*)handle;
    if (!I || !accepted) return fail(E_ARG, "null argument");
    *accepted = 0;
    Dev& D = I->D;
    if (!I->cn_alloc) {
        const size_t cv = (size_t)D.ncell * D.nvel;
        int rc;
        for (int q = 0; q < 6; ++q)
            if ((rc = dmalloc(I, &D.tg[q], cv)) || (rc = dmalloc(I, &D.tgn[q], cv))) return rc;
        if ((rc = dmalloc(I, &D.cnpart, (size_t)D.ncell * 8))) return rc;
        I->cn_alloc = true;
        I->cn_targets_fresh = false;
    }
    HIPCHK(ugks_launch_cn(&D, I->cn_targets_fresh ? 0 : 1, I->umax, I->vmax, I->stream, &I->prof));
    double sc[SC_COUNT];
    unsigned st = 0;
    HIPCHK(hipMemcpyAsync(sc, D.scal, sizeof sc, hipMemcpyDeviceToHost, I->stream));
    int rc = check_status(I, &st);
    if (diag16)
        for (int a = 0; a < 16; ++a) diag16[a] = sc[SC_CN_DIAG + a];
    if (rc) {
        // the state is untouched; the old targets were built only if valid
        if (!(st & ST_AT_CELL)) I->cn_targets_fresh = true;
        return rc;
    }
    I->cn_targets_fresh = true;   // tg now describes the current state
    I->operator_fresh = false;
    if (sc[SC_CN_ACCEPT] != 1.0) return E_OK;
    std::swap(D.W, D.Wn);
    std::swap(D.h, D.hn);
    std::swap(D.b, D.bn);
    for (int q = 0; q < 6; ++q) std::swap(D.tg[q], D.tgn[q]);
    *accepted = 1;
    return E_OK;
}

// S2CN2 conservation ledger: totals at the start of the run (4) and the
// cumulative wall balance (4).
extern "C" int ugks_cn_ledger(void* handle, double* out8) {
    Instance* I = (Instance*)handle;
    if (!I || !out8) return fail(E_ARG, "null argument");
    double sc[SC_COUNT];
    HIPCHK(hipMemcpyAsync(sc, I->D.scal, sizeof sc, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    for (int m = 0; m < 4; ++m) { out8[m] = sc[SC_CN_TOTAL0 + m]; out8[4 + m] = sc[SC_CN_BAL + m]; }
    return E_OK;
}

// the low-order Explicit kernels keep Gaussian tables per axis (Nu, Nv <= 64,
// powers of two)
static int lo_check(const Instance* I) {
    const int nu = I->D.nu, nv = I->D.nv;
    if (nu > 64 || nv > 64 || (nu & (nu - 1)) || (nv & (nv - 1)))
        return fail(E_ARG, "the Explicit entry needs Nu, Nv powers of two up to 64");
    if (I->lo_wg > I->D.nvel || nu + nv > I->lo_wg)
        return fail(E_ARG, "Explicit workgroup size must be at least Nu+Nv and at most Nu*Nv");
    return E_OK;
}

// Freestream cold start of the Explicit entry (InitFlow_Freestream): W of the
// primitive state (rho, ux, 0, 1/t_native) and its Maxwellian projected on W.
extern "C" int ugks_explicit_cold_start(void* handle, double rho, double ux, double t_native) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    if (!(rho > 0.0) || !(t_native > 0.0)) return fail(E_ARG, "cold start needs rho > 0 and T > 0");
    int rc = lo_check(I);
    if (rc) return rc;
    HIPCHK(ugks_launch_lo_cold_start(&I->D, rho, ux, t_native, I->stream));
    I->operator_fresh = false;
    rc = check_status(I, nullptr);
    return rc ? rc : reset_cn(I);
}

// One step of the low-order Explicit entry (TimeIntegration_method =
// 'Explicit': TimeStep, EvolutionNative, ApplyExplicitUpdate,
// src/mod_Solve.f90:84-165, src/mod_TimeIntegrator.f90:16-134).
extern "C" int ugks_explicit_step(void* handle, double* res4, double* dt) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    int rc = lo_check(I);
    if (rc) return rc;
    Dev& D = I->D;
    HIPCHK(ugks_launch_lo_step(&D, I->lo_wg, I->umax, I->vmax, I->stream, &I->prof));
    double sc[SC_COUNT];
    HIPCHK(hipMemcpyAsync(sc, D.scal, sizeof sc, hipMemcpyDeviceToHost, I->stream));
    if ((rc = check_status(I, nullptr))) return rc;
    if (res4)
        for (int m = 0; m < 4; ++m) res4[m] = sc[SC_LO_RES + m];
    if (dt) *dt = sc[SC_DT];
    std::swap(D.W, D.Wn);
    std::swap(D.h, D.hn);
    std::swap(D.b, D.bn);
    I->operator_fresh = false;
    I->cn_targets_fresh = false;
    return E_OK;
}

extern "C" int ugks_run(void* handle, int n, int* done) {
    if (!handle || !done) return fail(E_ARG, "null argument");
    *done = 0;
    for (int it = 0; it < n; ++it) {
        int acc = 0;
        int rc = ugks_implicit_step(handle, &acc, nullptr);
        if (rc) return rc;
        if (!acc) break;
        ++*done;
    }
    return E_OK;
}

// Diagnostics: copy an internal buffer to the host.  which: 0 xf (upwind x
// face data), 1 yf, 2 wb, 3 wt, 4 ws (macro face slopes), 5 wv (macro volume
// nodes), 6 scalars.  n = number of doubles copied (clamped to the buffer).
extern "C" int ugks_debug_fetch(void* handle, int which, double* out, long n) {
    Instance* I = (Instance*)handle;
    if (!I || !out) return fail(E_ARG, "null argument");
    const Dev& D = I->D;
    const size_t nv = D.nvel, nc = D.ncell;
    const double* src[] = {D.xf, D.yf, D.wb, D.wt, D.ws, D.wv, D.scal};
    const size_t len[] = {nc * REC_ROWS * nv, (size_t)D.nx * (D.ny - 1) * REC_ROWS * nv, (size_t)D.nx * REC_ROWS * nv,
                          (size_t)D.nx * REC_ROWS * nv, nc * 96, nc * 36, SC_COUNT};
    if (which < 0 || which > 6) return fail(E_ARG, "unknown buffer");
    size_t m = std::min((size_t)n, len[which]);
    HIPCHK(hipMemcpyAsync(out, src[which], sizeof(double) * m, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}

// Kernel timing (no counterpart in the reference): enable != 0 clears the
// totals and brackets every following launch with HIP events on the
// instance's stream; enable == 0 stops recording and keeps the totals.
extern "C" int ugks_profile(void* handle, int enable) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(I->stream));
    if (enable) I->prof.reset();
    else I->prof.collect();
    I->prof.on = enable != 0;
    return E_OK;
}

// Totals of kernel `idx` since profiling was enabled: name, summed event
// time in milliseconds and number of launches.  Returns 7 past the last id.
extern "C" int ugks_profile_read(void* handle, int idx, const char** name, double* total_ms, long* launches) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    if (idx < 0 || idx >= K_COUNT) return fail(E_ARG, "kernel index out of range");
    HIPCHK(hipStreamSynchronize(I->stream));
    I->prof.collect();
    if (name) *name = kernel_name(idx);
    if (total_ms) *total_ms = I->prof.total_ms[idx];
    if (launches) *launches = I->prof.launches[idx];
    return E_OK;
}

// Launch-shape tuning (no counterpart in the reference).  key 1: workgroup
// size of the velocity-parallel operator kernels (64, 128 or 256).
extern "C" int ugks_set_tuning(void* handle, int key, int value) {
    Instance* I = (Instance*)handle;
    if (!I) return fail(E_ARG, "null handle");
    switch (key) {
    case 1:
        if (value != 64 && value != 128 && value != 256) return fail(E_ARG, "velocity workgroup must be 64, 128 or 256");
        I->D.wg_vel = value;
        return E_OK;
    case 2:
        if (value != 64 && value != 128 && value != 256 && value != 512 && value != 1024)
            return fail(E_ARG, "Explicit workgroup size must be 64, 128, 256, 512 or 1024");
        if (value > I->D.nvel || I->D.nu + I->D.nv > value)
            return fail(E_ARG, "Explicit workgroup size must be at least Nu+Nv and at most Nu*Nv");
        I->lo_wg = value;
        return E_OK;
    default: return fail(E_ARG, "unknown tuning key");
    }
}

extern "C" int ugks_get_tuning(void* handle, int key, int* value) {
    Instance* I = (Instance*)handle;
    if (!I || !value) return fail(E_ARG, "null argument");
    switch (key) {
    case 1: *value = I->D.wg_vel; return E_OK;
    case 2: *value = I->lo_wg; return E_OK;
    default: return fail(E_ARG, "unknown tuning key");
    }
}

extern "C" int ugks_observables(void* handle, double* fields10) {
    Instance* I = (Instance*)handle;
    if (!I || !fields10) return fail(E_ARG, "null argument");
    HIPCHK(ugks_launch_observables(&I->D, I->obs, I->stream));
    HIPCHK(hipMemcpyAsync(fields10, I->obs, sizeof(double) * 10 * I->D.ncell, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return E_OK;
}
