// Cell reconstruction of one scalar plane for the S2 operator: WENO5-AO in
// x (periodic) and in y away from the walls, the one-sided quartic in y for
// the three rows next to each wall, tensor second stage with frozen weights
// for the normal-derivative companion (restates
// src/mod_S2D1Reconstruction.f90:379-460 and :501-578 in the reference).
#pragma once
#include "ugks_device.h"

// First stage on a 5x5 block of cell averages blk[row][col] (rows bottom to
// top, columns west to east, cell at column 2 and at the row given by the
// stencil): x lines evaluated at the faces and Gauss nodes, y lines (or the
// wall quartic) evaluated at the faces.
struct FirstLayer {
    double xv[5][2], xd[5][2], rows[5][3];
    double yv[5][2], yd[5][2];
};

__device__ inline void first_layer(const Consts& K, const double (&blk)[5][5], bool near, int sh, FirstLayer& F) {
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        double fw[4], cq[3][3], c4[5];
        weno_weights(blk[r], fw);
        weno_polys(K, blk[r], cq, c4);
        weno_at(fw, cq, c4, -0.5, &F.xv[r][0], &F.xd[r][0]);
        weno_at(fw, cq, c4, 0.5, &F.xv[r][1], &F.xd[r][1]);
#pragma unroll
        for (int a = 0; a < 3; ++a) F.rows[r][a] = weno_val(fw, cq, c4, K.gnode[a]);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        double q[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) q[r] = blk[r][c];
        if (near) {
            wall_at(K, sh, q, -0.5, &F.yv[c][0], &F.yd[c][0]);
            wall_at(K, sh, q, 0.5, &F.yv[c][1], &F.yd[c][1]);
        } else {
            double fw[4], cq[3][3], c4[5];
            weno_weights(q, fw);
            weno_polys(K, q, cq, c4);
            weno_at(fw, cq, c4, -0.5, &F.yv[c][0], &F.yd[c][0]);
            weno_at(fw, cq, c4, 0.5, &F.yv[c][1], &F.yd[c][1]);
        }
    }
}

// The first stage split by cell, for the distributions.  The x line of a
// cell is its own row's reconstruction centred on it (value and derivative
// at its two x faces, values at the three Gauss nodes); the y line of a cell
// is its own column's reconstruction evaluated at its two y faces (the wall
// quartic of its row's stencil near a wall).  A cell's FirstLayer is the x
// lines of the five cells of its column stencil and the y lines of the five
// cells of its row stencil, so each line is computed once instead of once
// per neighbour.  Line rows: xv(-1/2), xv(+1/2), xd(-1/2), xd(+1/2), three
// Gauss-node values, then yv(-1/2), yv(+1/2), yd(-1/2), yd(+1/2).
enum { LINE_ROWS = 11 };

__device__ inline void x_line(const Consts& K, const double q[5], double o[7]) {
    double fw[4], cq[3][3], c4[5];
    weno_weights(q, fw);
    weno_polys(K, q, cq, c4);
    weno_at(fw, cq, c4, -0.5, &o[0], &o[2]);
    weno_at(fw, cq, c4, 0.5, &o[1], &o[3]);
#pragma unroll
    for (int a = 0; a < 3; ++a) o[4 + a] = weno_val(fw, cq, c4, K.gnode[a]);
}

__device__ inline void y_line(const Consts& K, const double q[5], bool near, int sh, double o[4]) {
    if (near) {
        wall_at(K, sh, q, -0.5, &o[0], &o[2]);
        wall_at(K, sh, q, 0.5, &o[1], &o[3]);
    } else {
        double fw[4], cq[3][3], c4[5];
        weno_weights(q, fw);
        weno_polys(K, q, cq, c4);
        weno_at(fw, cq, c4, -0.5, &o[0], &o[2]);
        weno_at(fw, cq, c4, 0.5, &o[1], &o[3]);
    }
}

// companion evaluation at the three Gauss nodes: value and derivative of the
// reconstruction of q, and the companion reconstructed with q's weights
__device__ inline void companion3(const Consts& K, const double q[5], const double comp[5], double d[3][3]) {
    double fw[4], cq[3][3], c4[5], kq[3][3], k4[5];
    weno_weights(q, fw);
    weno_polys(K, q, cq, c4);
    weno_polys(K, comp, kq, k4);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        weno_at(fw, cq, c4, K.gnode[a], &d[a][0], &d[a][2]);
        d[a][1] = weno_val(fw, kq, k4, K.gnode[a]);
    }
}

// side s (0 west, 1 east, 2 south, 3 north): d[a] = {value, normal derivative,
// tangential derivative} per unit cell width at Gauss node a
__device__ inline void side_data(const Consts& K, const FirstLayer& F, bool near, int sh, int s, double d[3][3]) {
    if (s < 2) {
        double q[5], c[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) { q[r] = F.xv[r][s]; c[r] = F.xd[r][s]; }
        if (near) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                wall_at(K, sh, q, K.gnode[a], &d[a][0], &d[a][2]);
                wall_at(K, sh, c, K.gnode[a], &d[a][1], nullptr);
            }
        } else {
            companion3(K, q, c, d);
        }
    } else {
        double q[5], c[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) { q[r] = F.yv[r][s - 2]; c[r] = F.yd[r][s - 2]; }
        companion3(K, q, c, d);
    }
}

// 3x3 volume Gauss nodes vol[a][b] (a: x node, b: y node)
__device__ inline void vol_data(const Consts& K, const FirstLayer& F, bool near, int sh, double vol[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double q[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) q[r] = F.rows[r][a];
        if (near) {
#pragma unroll
            for (int b = 0; b < 3; ++b) wall_at(K, sh, q, K.gnode[b], &vol[a][b], nullptr);
        } else {
            double fw[4], cq[3][3], c4[5];
            weno_weights(q, fw);
            weno_polys(K, q, cq, c4);
#pragma unroll
            for (int b = 0; b < 3; ++b) vol[a][b] = weno_val(fw, cq, c4, K.gnode[b]);
        }
    }
}

// stencil rows of cell row j: rows rbase..rbase+4, near-wall flag and wall map shift
__device__ inline void stencil_rows(int j, int ny, int* rbase, bool* near, int* sh) {
    if (j <= 2) { *near = true; *rbase = 0; *sh = j; }
    else if (j >= ny - 3) { *near = true; *rbase = ny - 5; *sh = j - (ny - 5); }
    else { *near = false; *rbase = j - 2; *sh = 0; }
}
