// VAM level-1 (2-D horizontal) Chorin pressure split: the shape contract, tables and the
// three per-cell operators of the reference's generated ChorinOps.H (ChorinOps.H:1-138),
// restated term by term in factored form (k = dt / rho, h = Q[1]).
//   full state  Q  = [b, h, q_x_0, q_x_1, q_y_0, q_y_1, r_0, r_1, P_0, P_1]
//   derivative aux  = VAM_CA (P derivatives first, then frozen b, h, q derivatives)
#pragma once
#include "Numerics.H"

static constexpr int VAM_NF = 10;
static constexpr int VAM_NP = 2;
static constexpr int VAM_NS = VAM_NF - VAM_NP;
static constexpr int VAM_NCA = 20;
static constexpr int VAM_P_IDX[VAM_NP] = {8, 9};
static constexpr int VAM_CORR_IDX[6] = {2, 3, 4, 5, 6, 7};
static constexpr int VAM_N_CORR = 6;

// derivative-aux fill table: field (index into Q), first axis, second axis (-1: first
// derivative only), pressure aux (recomputed every operator application) or frozen
struct VamAuxSpec { int field; int ax1; int ax2; bool is_P; };
static constexpr VamAuxSpec VAM_CA[VAM_NCA] = {
    {8, 0, -1, true},  {8, 0, 0, true},  {8, 1, -1, true},  {8, 1, 1, true},     // P0 x, xx, y, yy
    {9, 0, -1, true},  {9, 0, 0, true},  {9, 1, -1, true},  {9, 1, 1, true},     // P1 x, xx, y, yy
    {0, 0, -1, false}, {0, 0, 0, false}, {0, 1, -1, false}, {0, 1, 1, false},    // b  x, xx, y, yy
    {1, 0, -1, false}, {1, 0, 0, false}, {1, 1, -1, false}, {1, 1, 1, false},    // h  x, xx, y, yy
    {2, 0, -1, false}, {3, 0, -1, false}, {4, 1, -1, false}, {5, 1, -1, false},  // qx0_x qx1_x qy0_y qy1_y
};

// pressure-operator blocks: A0 (P itself), Ax, Ay (first derivatives), Axx, Ayy
struct VamPBlockSpec { int ax1; int ax2; };
static constexpr int VAM_N_PBLOCK = 5;
static constexpr VamPBlockSpec VAM_PBLOCKS[VAM_N_PBLOCK] = {{-1, -1}, {0, -1}, {1, -1}, {0, 0}, {1, 1}};

// out[(blk*NP + k)*NP + l] = coefficient of block blk, row k, column l at this cell
template <typename T>
PORTABLE_FN inline void vam_pressure_blocks(const T* Q, const T* A, const T* p, T dt, T* out) {
    const T k = dt / p[1], h = Q[1];
    const T bx = A[8], bxx = A[9], by = A[10], byy = A[11];
    const T hx = A[12], hxx = A[13], hy = A[14], hyy = A[15];
    const T b2 = bx * bx + by * by, h2 = hx * hx + hy * hy, hb = hx * bx + hy * by;
    const T bl = bxx + byy, hl = hxx + hyy;
    out[0] = k * ((T)2 * bl + ((T)2 * h2 + (T)4 * hb) / h);
    out[1] = k * ((T)4 * bl + ((T)2 / 3) * hl + (((T)2 / 3) * h2 - (T)4 * b2 - (T)4) / h);
    out[2] = k * (-hl + ((T)12 * b2 + (T)12 * hb + (T)4 * h2 + (T)12) / h);
    out[3] = k * ((T)2 * bl + ((T)12 * b2 + (T)8 * hb + (T)2 * h2 + (T)12) / h);
    out[4] = (T)4 * bx * k;
    out[5] = ((T)4 / 3) * hx * k + (T)4 * bx * k;
    out[6] = -hx * k;
    out[7] = hx * k + (T)4 * bx * k;
    out[8] = (T)4 * by * k;
    out[9] = (T)4 * by * k + ((T)4 / 3) * hy * k;
    out[10] = -hy * k;
    out[11] = (T)4 * by * k + hy * k;
    out[12] = -h * k;
    out[13] = ((T)1 / 3) * h * k;
    out[14] = -h * k;
    out[15] = 0;
    out[16] = -h * k;
    out[17] = ((T)1 / 3) * h * k;
    out[18] = -h * k;
    out[19] = 0;
}

// residual(P) (affine in P): the pressure equations at this cell.  Written term by term,
// independently of vam_pressure_blocks, so that the assembled-operator consistency check
// (Pmat P == residual(P) - residual(0)) compares two separate transcriptions.
template <typename T>
PORTABLE_FN inline void vam_pressure_residual(const T* Q, const T* A, const T* p, T dt, T* out) {
    const T k = dt / p[1], h = Q[1], P0 = Q[8], P1 = Q[9];
    const T P0x = A[0], P0xx = A[1], P0y = A[2], P0yy = A[3];
    const T P1x = A[4], P1xx = A[5], P1y = A[6], P1yy = A[7];
    const T bx = A[8], bxx = A[9], by = A[10], byy = A[11];
    const T hx = A[12], hxx = A[13], hy = A[14], hyy = A[15];
    const T qx0x = A[16], qx1x = A[17], qy0y = A[18], qy1y = A[19];
    const T qx0 = Q[2], qx1 = Q[3], qy0 = Q[4], qy1 = Q[5], r0 = Q[6], r1 = Q[7];
    // second derivatives of P scaled by h
    T s0 = -h * P0xx - h * P0yy + (h * P1xx + h * P1yy) / (T)3;
    // P times the bathymetry / depth curvature
    s0 += (T)2 * P0 * (byy + bxx) + (T)4 * P1 * (byy + bxx) + ((T)2 / 3) * P1 * (hxx + hyy);
    // first derivatives of P against the bed and depth slopes
    s0 += (T)4 * (P0x * bx + by * P0y + by * P1y + P1x * bx) + ((T)4 / 3) * (hx * P1x + hy * P1y);
    // P against products of slopes, over h
    s0 += ((T)4 * P0 * by * hy + (T)2 * P0 * hx * hx + (T)4 * P0 * hx * bx + (T)2 * P0 * hy * hy -
           (T)4 * P1 * by * by + ((T)2 / 3) * P1 * hx * hx + ((T)2 / 3) * P1 * hy * hy -
           (T)4 * P1 * bx * bx - (T)4 * P1) / h;
    const T v0 = qx0x - qx1x / (T)3 + qy0y - qy1y / (T)3 -
                 (qx0 * hx + (T)2 * qx0 * bx + qx1 * hx / (T)3 + (T)2 * qy0 * by + qy0 * hy +
                  qy1 * hy / (T)3) / h + (T)2 * r0 / h;
    T s1 = -h * P0xx - h * P0yy - P0 * (hxx + hyy) + (T)2 * P1 * (byy + bxx);
    s1 += -P0x * hx + (T)4 * by * P1y + hx * P1x - hy * P0y + hy * P1y + (T)4 * P1x * bx;
    s1 += ((T)12 * P0 * by * by + (T)12 * P0 * by * hy + (T)4 * P0 * hx * hx + (T)12 * P0 * hx * bx +
           (T)4 * P0 * hy * hy + (T)12 * P0 * bx * bx + (T)12 * P0 + (T)12 * P1 * by * by +
           (T)8 * P1 * by * hy + (T)2 * P1 * hx * hx + (T)8 * P1 * hx * bx + (T)2 * P1 * hy * hy +
           (T)12 * P1 * bx * bx + (T)12 * P1) / h;
    const T v1 = qx0x + qy0y -
                 (qx0 * hx + qx1 * hx + (T)2 * qx1 * bx + qy0 * hy + (T)2 * qy1 * by + qy1 * hy) / h +
                 (T)2 * r1 / h;
    out[0] = k * s0 + v0;
    out[1] = k * s1 + v1;
}

// corrector: new values of the VAM_CORR_IDX slots from the solved pressure
template <typename T>
PORTABLE_FN inline void vam_corrector(const T* Q, const T* A, const T* p, T dt, T* out) {
    const T k = dt / p[1], h = Q[1], P0 = Q[8], P1 = Q[9];
    const T P0x = A[0], P0y = A[2], P1x = A[4], P1y = A[6];
    const T bx = A[8], by = A[10], hx = A[12], hy = A[14];
    out[0] = Q[2] - k * (h * P0x + P0 * hx - (T)2 * P1 * bx);
    out[1] = Q[3] - k * (h * P1x + (T)3 * P0 * hx + (T)6 * P0 * bx + (T)2 * P1 * hx + (T)6 * P1 * bx);
    out[2] = Q[4] - k * (h * P0y + P0 * hy - (T)2 * P1 * by);
    out[3] = Q[5] - k * (h * P1y + (T)6 * P0 * by + (T)3 * P0 * hy + (T)6 * P1 * by + (T)2 * P1 * hy);
    out[4] = Q[6] - (T)2 * P1 * k;
    out[5] = Q[7] + (T)6 * P0 * k + (T)6 * P1 * k;
}
