/*
 * sps_oracle.c - CPU restatement of the SymPhas SolverFT step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP solver, never part of the
 * product.  Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, so
 * every product and sum rounds exactly as written).
 *
 * Citations are to the reference SymPhas tree.
 */
#include "sps_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

size_t orc_desc_size(void) { return sizeof(orc_desc); }

static int64_t pdim(const orc_desc* d, int a) {
  return (int64_t)d->n[a] + 2 * ORC_DEPTH;
}

size_t orc_padded_len(const orc_desc* d) {
  size_t len = (size_t)pdim(d, 0) * (size_t)pdim(d, 1);
  if (d->dim == 3) len *= (size_t)pdim(d, 2);
  return len;
}

static int valid_points(const orc_desc* d) {
  if (d->dim == 2) {
    int L = d->lap_points, G = d->gradlap_points, B = d->bilap_points;
    return (L == 5 || L == 9) && (G == 6 || G == 8 || G == 12 || G == 16) &&
           (B == 13 || B == 17 || B == 21);
  }
  int L = d->lap_points, G = d->gradlap_points, B = d->bilap_points;
  return (L == 7 || L == 15 || L == 19 || L == 21 || L == 27) &&
         (G == 10 || G == 12 || G == 28 || G == 36 || G == 40) &&
         (B == 21 || B == 25 || B == 41 || B == 52 || B == 57);
}

int orc_check(const orc_desc* d) {
  if (!d || (d->dim != 2 && d->dim != 3)) return -1;
  for (int a = 0; a < d->dim; ++a)
    if (d->n[a] < ORC_DEPTH) return -2; /* ghost copy needs n >= 3 */
  if (!(d->h > 0)) return -3;
  if (!valid_points(d)) return -4;
  if (d->nterms < 0 || d->nterms > ORC_MAX_TERMS) return -5;
  if (d->nprov < 0 || d->nprov > ORC_MAX_PROV) return -9;
  /* group 0: the right-hand side, operands psi or any var; group k: var(k),
   * operands psi or var(j), j < k */
  for (int g = 0; g <= d->nprov; ++g) {
    const orc_term* terms = g == 0 ? d->terms : d->prov[g - 1];
    const int nt = g == 0 ? d->nterms : d->prov_nterms[g - 1];
    if (g > 0 && (nt < 1 || nt > ORC_MAX_TERMS)) return -10;
    for (int t = 0; t < nt; ++t) {
      const orc_term* tm = &terms[t];
      if (tm->op < ORC_OP_ID || tm->op > ORC_OP_GRADLAP) return -6;
      if (tm->operand < 0 || tm->operand > (g == 0 ? d->nprov : g - 1))
        return -11;
      if (tm->op == ORC_OP_GRADLAP) {
        if (tm->axis < 0 || tm->axis >= d->dim) return -7;
        if (tm->component < 0 ||
            tm->component >= ((d->dim == 2 && d->gradlap_points == 6) ? 2 : 1))
          return -8;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------- layout */

void orc_pad(const orc_desc* d, const double* in, double* out) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  const int zoff = d->dim == 3 ? ORC_DEPTH : 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < d->n[1]; ++y)
      for (int x = 0; x < d->n[0]; ++x)
        out[(x + ORC_DEPTH) + X * ((y + ORC_DEPTH) + Y * (z + zoff))] =
            in[x + (int64_t)d->n[0] * (y + (int64_t)d->n[1] * z)];
}

void orc_unpad(const orc_desc* d, const double* in, double* out) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  const int zoff = d->dim == 3 ? ORC_DEPTH : 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < d->n[1]; ++y)
      for (int x = 0; x < d->n[0]; ++x)
        out[x + (int64_t)d->n[0] * (y + (int64_t)d->n[1] * z)] =
            in[(x + ORC_DEPTH) + X * ((y + ORC_DEPTH) + Y * (z + zoff))];
}

/* ---------------------------------------------------- periodic refresh */

/* 2D: the four side updates of boundaryupdate.h:650-708 in the order
 * BoundaryGroup applies them (LEFT, RIGHT, TOP, BOTTOM; boundarygroup.h:
 * 373-381).  LEFT/RIGHT copy the side strips over the interior rows plus the
 * two corner blocks on their side; TOP/BOTTOM copy the interior columns of
 * the top/bottom strips.  Every source cell is an interior cell. */
static void refresh2d(const orc_desc* d, double* g) {
  const int64_t L = pdim(d, 0), M = pdim(d, 1);
  const int64_t ox = L - 2 * ORC_DEPTH, oy = M - 2 * ORC_DEPTH;
  /* LEFT: columns [0,3), rows [3, M-3), from column + ox */
  for (int64_t r = ORC_DEPTH; r < M - ORC_DEPTH; ++r)
    for (int64_t c = 0; c < ORC_DEPTH; ++c) g[r * L + c] = g[r * L + c + ox];
  /* LEFT, top-left corner block: rows [0,3), from row + oy, col + ox */
  for (int64_t r = 0; r < ORC_DEPTH; ++r)
    for (int64_t c = 0; c < ORC_DEPTH; ++c)
      g[r * L + c] = g[(r + oy) * L + c + ox];
  /* LEFT, bottom-left corner block: rows [M-3, M), from row - oy, col + ox */
  for (int64_t r = M - ORC_DEPTH; r < M; ++r)
    for (int64_t c = 0; c < ORC_DEPTH; ++c)
      g[r * L + c] = g[(r - oy) * L + c + ox];
  /* RIGHT: columns [L-3, L), rows [3, M-3), from column - ox */
  for (int64_t r = ORC_DEPTH; r < M - ORC_DEPTH; ++r)
    for (int64_t c = L - ORC_DEPTH; c < L; ++c)
      g[r * L + c] = g[r * L + c - ox];
  /* RIGHT, bottom-right corner block */
  for (int64_t r = M - ORC_DEPTH; r < M; ++r)
    for (int64_t c = L - ORC_DEPTH; c < L; ++c)
      g[r * L + c] = g[(r - oy) * L + c - ox];
  /* RIGHT, top-right corner block */
  for (int64_t r = 0; r < ORC_DEPTH; ++r)
    for (int64_t c = L - ORC_DEPTH; c < L; ++c)
      g[r * L + c] = g[(r + oy) * L + c - ox];
  /* TOP: rows [0,3), interior columns, from row + oy */
  for (int64_t r = 0; r < ORC_DEPTH; ++r)
    for (int64_t c = ORC_DEPTH; c < L - ORC_DEPTH; ++c)
      g[r * L + c] = g[(r + oy) * L + c];
  /* BOTTOM: rows [M-3, M), interior columns, from row - oy */
  for (int64_t r = M - ORC_DEPTH; r < M; ++r)
    for (int64_t c = ORC_DEPTH; c < L - ORC_DEPTH; ++c)
      g[r * L + c] = g[(r - oy) * L + c];
}

static int64_t wrap_padded(int64_t p, int64_t n) {
  /* padded coordinate -> padded coordinate of its interior image */
  int64_t i = (p - ORC_DEPTH) % n;
  if (i < 0) i += n;
  return i + ORC_DEPTH;
}

void orc_refresh_modular(const orc_desc* d, double* g) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int64_t Z = d->dim == 3 ? pdim(d, 2) : 1;
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        int64_t sx = wrap_padded(x, d->n[0]);
        int64_t sy = wrap_padded(y, d->n[1]);
        int64_t sz = d->dim == 3 ? wrap_padded(z, d->n[2]) : z;
        g[x + X * (y + Y * z)] = g[sx + X * (sy + Y * sz)];
      }
}

void orc_refresh(const orc_desc* d, double* g) {
  if (d->dim == 2)
    refresh2d(d, g);
  else
    orc_refresh_modular(d, g);
}

/* -------------------------------------------------------------- stencils
 *
 * Each formula keeps the reference's grouping and operand order
 * (sol/inc/stencilh2.h) so that it rounds identically; V(i,j) / W(i,j,k) is the
 * value at offset i*s0 + j*s1 (+ k*s2), where s is the stride list of
 * get_stride<axis> (datatypes/inc/grid.h:537-578).
 */
#define V(i, j) v[(i) * s[0] + (j) * s[1]]
#define W(i, j, k) v[(i) * s[0] + (j) * s[1] + (k) * s[2]]

static double lap2_5(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:87-95 */
  return (V(-1, 0) + V(1, 0) + V(0, 1) + V(0, -1) - 4. * V(0, 0)) * divh2;
}

static double lap2_9(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:97-107 */
  return (V(-1, 1) + V(1, -1) + V(1, 1) + V(-1, -1) + 4. * V(1, 0) +
          4. * V(0, 1) + 4. * V(0, -1) + 4. * V(-1, 0) - 20. * V(0, 0)) *
         divh2 * (1.0 / 6);
}

static double bilap2_13(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:112-125 */
  return (V(0, -2) + V(0, 2) + V(-2, 0) + V(2, 0) +
          2. * (V(-1, -1) + V(1, -1) + V(-1, 1) + V(1, 1)) -
          8. * (V(-1, 0) + V(1, 0) + V(0, -1) + V(0, 1)) + 20. * V(0, 0)) *
         divh4;
}

static double bilap2_17(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:127-141 */
  return ((V(2, 2) + V(-2, 2) + V(2, -2) + V(-2, -2)) +
          10. * (V(2, 0) + V(-2, 0) + V(0, 2) + V(0, -2)) +
          8. * (V(1, 1) + V(-1, 1) + V(1, -1) + V(-1, -1)) -
          64. * (V(1, 0) + V(-1, 0) + V(0, 1) + V(0, -1)) + 180. * V(0, 0)) *
         divh4 * (1.0 / 12);
}

static double bilap2_21(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:143-157; the reference divides by 12 where a consistent
   * bilaplacian with these weights needs 3, so this is bilap/4. Reproduced. */
  return ((V(2, 1) + V(-2, 1) + V(2, -1) + V(-2, -1) + V(1, 2) + V(-1, 2) +
           V(1, -2) + V(-1, -2)) +
          (V(2, 0) + V(-2, 0) + V(0, 2) + V(0, -2)) -
          2. * (V(1, 1) + V(-1, 1) + V(1, -1) + V(-1, -1)) -
          10. * (V(1, 0) + V(-1, 0) + V(0, 1) + V(0, -1)) + 36. * V(0, 0)) *
         divh4 * (1.0 / 12);
}

static double gradlap2_6(const double* v, const int64_t* s, double divh3,
                         int component) {
  /* stencilh2.h:161-175: returns {x, y}; component selects one */
  if (component == 0)
    return ((V(2, 1) - V(-2, 1) + V(2, -1) - V(-2, -1)) -
            4. * (V(1, 0) - V(-1, 0))) *
           divh3 * (1.0 / 4);
  return ((V(1, 2) - V(1, -2) + V(-1, 2) - V(-1, -2)) -
          4. * (V(0, 1) - V(0, -1))) *
         divh3 * (1.0 / 4);
}

static double gradlap2_8(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:177-189 */
  return ((V(2, 0) - V(-2, 0)) +
          (V(1, 1) - V(-1, 1) + V(1, -1) - V(-1, -1)) -
          4. * (V(1, 0) - V(-1, 0))) *
         divh3 * (1.0 / 2);
}

static double gradlap2_12(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:191-205 */
  return ((V(2, 2) - V(-2, 2) + V(2, -2) - V(-2, -2)) +
          10. * (V(2, 0) - V(-2, 0)) +
          4. * (V(1, 1) - V(-1, 1) + V(1, -1) - V(-1, -1)) -
          32. * (V(1, 0) - V(-1, 0))) *
         divh3 * (1.0 / 24);
}

static double gradlap2_16(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:207-221 */
  return (2. * (V(2, 1) - V(-2, 1) + V(2, -1) - V(-2, -1)) +
          2. * (V(2, 0) - V(-2, 0)) +
          (V(1, 2) - V(-1, 2) + V(1, -2) - V(-1, -2)) -
          2. * (V(1, 1) - V(-1, 1) + V(1, -1) - V(-1, -1)) -
          10. * (V(1, 0) - V(-1, 0))) *
         divh3 * (1.0 / 12);
}

/* 3D neighbour groups in the orders the reference sums them. */
#define AX6_A \
  (W(-1, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) + W(0, 0, 1) + W(0, 0, -1))
#define AX6_B \
  (W(1, 0, 0) + W(-1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) + W(0, 0, 1) + W(0, 0, -1))
#define AX2_6 \
  (W(2, 0, 0) + W(-2, 0, 0) + W(0, 2, 0) + W(0, -2, 0) + W(0, 0, 2) + W(0, 0, -2))
/* corners, order A: x_y_z_, xy_z_, x_yz_, xyz_, x_y_z, xy_z, x_yz, xyz */
#define CORNER8_A                                                     \
  (W(-1, -1, -1) + W(1, -1, -1) + W(-1, 1, -1) + W(1, 1, -1) +        \
   W(-1, -1, 1) + W(1, -1, 1) + W(-1, 1, 1) + W(1, 1, 1))
/* corners, order B: xyz, x_yz, xy_z, x_y_z, xyz_, x_yz_, xy_z_, x_y_z_ */
#define CORNER8_B                                                     \
  (W(1, 1, 1) + W(-1, 1, 1) + W(1, -1, 1) + W(-1, -1, 1) +            \
   W(1, 1, -1) + W(-1, 1, -1) + W(1, -1, -1) + W(-1, -1, -1))
/* edges, order A: x_y_, xy_, x_y, xy, x_z_, xz_, x_z, xz, y_z_, yz_, y_z, yz */
#define EDGE12_A                                                          \
  (W(-1, -1, 0) + W(1, -1, 0) + W(-1, 1, 0) + W(1, 1, 0) + W(-1, 0, -1) + \
   W(1, 0, -1) + W(-1, 0, 1) + W(1, 0, 1) + W(0, -1, -1) + W(0, 1, -1) +  \
   W(0, -1, 1) + W(0, 1, 1))
/* edges, order B: xy, x_y, xy_, x_y_, xz_, x_z_, xz, x_z, yz_, yz, y_z_, y_z */
#define EDGE12_B                                                         \
  (W(1, 1, 0) + W(-1, 1, 0) + W(1, -1, 0) + W(-1, -1, 0) + W(1, 0, -1) + \
   W(-1, 0, -1) + W(1, 0, 1) + W(-1, 0, 1) + W(0, 1, -1) + W(0, 1, 1) +  \
   W(0, -1, -1) + W(0, -1, 1))
/* second-neighbour corners: x2y2z2, x2_y2z2, x2y2_z2, x2_y2_z2, then z2_ */
#define CORNER8_2                                                     \
  (W(2, 2, 2) + W(-2, 2, 2) + W(2, -2, 2) + W(-2, -2, 2) +            \
   W(2, 2, -2) + W(-2, 2, -2) + W(2, -2, -2) + W(-2, -2, -2))

static double lap3_7(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:235-246 */
  return (W(-1, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) + W(0, 0, 1) +
          W(0, 0, -1) - 6. * W(0, 0, 0)) *
         divh2;
}

static double lap3_15(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:248-260 */
  return (W(-1, -1, -1) + W(1, -1, -1) + W(-1, 1, -1) + W(1, 1, -1) +
          W(-1, -1, 1) + W(1, -1, 1) + W(-1, 1, 1) + W(1, 1, 1) +
          8. * (W(-1, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) +
                W(0, 0, 1) + W(0, 0, -1) - 7. * W(0, 0, 0))) *
         divh2 * (1.0 / 12);
}

static double lap3_19(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:262-276, including the reference's "vz + vz" (:273): the
   * -z axial neighbour is replaced by a second +z neighbour. */
  return (W(-1, -1, 0) + W(1, -1, 0) + W(-1, 1, 0) + W(1, 1, 0) +
          W(-1, 0, -1) + W(1, 0, -1) + W(-1, 0, 1) + W(1, 0, 1) +
          W(0, -1, -1) + W(0, 1, -1) + W(0, -1, 1) + W(0, 1, 1) +
          2. * (W(-1, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) +
                W(0, 0, 1) + W(0, 0, 1)) -
          24. * W(0, 0, 0)) *
         divh2 * (1.0 / 6);
}

static double lap3_21(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:278-293 */
  return (-CORNER8_A + 4. * EDGE12_A + -40. * W(0, 0, 0)) * divh2 *
         (1.0 / 12);
}

static double lap3_27(const double* v, const int64_t* s, double divh2) {
  /* stencilh2.h:295-310 */
  return (CORNER8_A + 3. * EDGE12_A + 14. * AX6_A - 128. * W(0, 0, 0)) *
         divh2 * (1.0 / 30);
}

static double bilap3_21(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:312-327 */
  return ((W(2, 2, 0) + W(-2, 2, 0) + W(2, -2, 0) + W(-2, -2, 0) +
           W(2, 0, -2) + W(-2, 0, -2) + W(2, 0, 2) + W(-2, 0, 2) +
           W(0, 2, -2) + W(0, 2, 2) + W(0, -2, -2) + W(0, -2, 2)) -
          4. * CORNER8_B + 20. * W(0, 0, 0)) *
         divh4 * (1.0 / 4);
}

static double bilap3_25(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:329-344 */
  return (2. * EDGE12_B + AX2_6 - 12. * AX6_B + 42. * W(0, 0, 0)) * divh4;
}

static double bilap3_41(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:346-364 */
  return (CORNER8_2 - 40. * CORNER8_B + 96. * EDGE12_B + 20. * AX2_6 -
          320. * AX6_B + 960. * W(0, 0, 0)) *
         divh4 * (1.0 / 24);
}

static double bilap3_52(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:366-389 */
  return (-CORNER8_2 +
          10. * (W(1, 1, 2) + W(-1, 1, 2) + W(1, -1, 2) + W(-1, -1, 2) +
                 W(1, 1, -2) + W(-1, 1, -2) + W(1, -1, -2) + W(-1, -1, -2) +
                 W(1, 2, 1) + W(-1, 2, 1) + W(1, -2, 1) + W(-1, -2, 1) +
                 W(1, 2, -1) + W(-1, 2, -1) + W(1, -2, -1) + W(-1, -2, -1) +
                 W(2, 1, 1) + W(-2, 1, 1) + W(2, -1, 1) + W(-2, -1, 1) +
                 W(2, 1, -1) + W(-2, 1, -1) + W(2, -1, -1) + W(-2, -1, -1)) -
          20. * CORNER8_B - 36. * EDGE12_B + 360. * W(0, 0, 0)) *
         divh4 * (1.0 / 36);
}

static double bilap3_57(const double* v, const int64_t* s, double divh4) {
  /* stencilh2.h:391-415 */
  return ((W(0, 1, 2) + W(0, -1, 2) + W(0, 1, -2) + W(0, -1, -2) +
           W(0, 2, 1) + W(0, -2, 1) + W(0, 2, -1) + W(0, -2, -1) +
           W(2, 0, 1) + W(-2, 0, 1) + W(2, 0, -1) + W(-2, 0, -1) +
           W(1, 0, 2) + W(-1, 0, 2) + W(1, 0, -2) + W(-1, 0, -2) +
           W(1, 2, 0) + W(-1, 2, 0) + W(1, -2, 0) + W(-1, -2, 0) +
           W(2, 1, 0) + W(-2, 1, 0) + W(2, -1, 0) + W(-2, -1, 0)) +
          3. * CORNER8_B - 8. * EDGE12_B - AX2_6 + 4. * AX6_B +
          30. * W(0, 0, 0)) *
         divh4 * (1.0 / 3);
}

/* odd differences along s0 used by the gradlaplacians */
#define DX1 (W(1, 0, 0) - W(-1, 0, 0))
#define DX2 (W(2, 0, 0) - W(-2, 0, 0))
/* xy, x_y, xy_, x_y_, xz_, x_z_, xz, x_z with signs + - + - + - + - */
#define DEDGE8                                                            \
  (W(1, 1, 0) - W(-1, 1, 0) + W(1, -1, 0) - W(-1, -1, 0) + W(1, 0, -1) - \
   W(-1, 0, -1) + W(1, 0, 1) - W(-1, 0, 1))
/* xyz, x_yz, xy_z, x_y_z, xyz_, x_yz_, xy_z_, x_y_z_ with signs + - + - ... */
#define DCORNER8                                                          \
  (W(1, 1, 1) - W(-1, 1, 1) + W(1, -1, 1) - W(-1, -1, 1) + W(1, 1, -1) - \
   W(-1, 1, -1) + W(1, -1, -1) - W(-1, -1, -1))

static double gradlap3_10(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:417-429 */
  return ((W(2, 1, 1) - W(-2, 1, 1) + W(2, -1, 1) - W(-2, -1, 1) +
           W(2, 1, -1) - W(-2, 1, -1) + W(2, -1, -1) - W(-2, -1, -1)) -
          8. * DX1) *
         divh3 * (1.0 / 8);
}

static double gradlap3_12(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:431-444 */
  return (DX2 + DEDGE8 - 6. * DX1) * divh3 * (1.0 / 2);
}

static double gradlap3_28(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:446-463 */
  return ((W(2, 2, 0) - W(-2, 2, 0) + W(2, -2, 0) - W(-2, -2, 0) +
           W(2, 0, -2) - W(-2, 0, -2) + W(2, 0, 2) - W(-2, 0, 2)) +
          8. * DX2 + 6. * DCORNER8 - 8. * DEDGE8 - 16. * DX1) *
         divh3 * (1.0 / 24);
}

static double gradlap3_36(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:465-484 */
  return (2. * (W(2, 0, 1) - W(-2, 0, 1) + W(2, 0, -1) - W(-2, 0, -1) +
                W(2, 1, 0) - W(-2, 1, 0) + W(2, -1, 0) - W(-2, -1, 0)) -
          2. * DX2 +
          (W(1, 0, 2) - W(-1, 0, 2) + W(1, 0, -2) - W(-1, 0, -2) +
           W(1, 2, 0) - W(-1, 2, 0) + W(1, -2, 0) - W(-1, -2, 0)) +
          3. * DCORNER8 - 8. * DEDGE8 + 4. * DX1) *
         divh3 * (1.0 / 12);
}

static double gradlap3_40(const double* v, const int64_t* s, double divh3) {
  /* stencilh2.h:486-507 */
  return (16. * (W(2, 0, 1) - W(-2, 0, 1) + W(2, 0, -1) - W(-2, 0, -1) +
                 W(2, 1, 0) - W(-2, 1, 0) + W(2, -1, 0) - W(-2, -1, 0)) -
          16. * DX2 +
          3. * (W(1, 2, 1) - W(-1, 2, 1) + W(1, -2, 1) - W(-1, -2, 1) +
                W(1, 2, -1) - W(-1, 2, -1) + W(1, -2, -1) - W(-1, -2, -1) +
                W(1, 1, 2) - W(-1, 1, 2) + W(1, -1, 2) - W(-1, -1, 2) +
                W(1, 1, -2) - W(-1, 1, -2) + W(1, -1, -2) - W(-1, -1, -2)) +
          2. * (W(1, 0, 2) - W(-1, 0, 2) + W(1, 0, -2) - W(-1, 0, -2) +
                W(1, 2, 0) - W(-1, 2, 0) + W(1, -2, 0) - W(-1, -2, 0)) -
          22. * DEDGE8 - 40. * DX1) *
         divh3 * (1.0 / 96);
}

#undef V
#undef W

/* get_stride<axis> on the padded dims (datatypes/inc/grid.h:537-578). */
static void strides(const orc_desc* d, int axis, int64_t s[3]) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  if (d->dim == 2) {
    if (axis == 0) {
      s[0] = 1;
      s[1] = X;
    } else {
      s[0] = X;
      s[1] = -1;
    }
    s[2] = 0;
    return;
  }
  if (axis == 0) {
    s[0] = 1;
    s[1] = X;
    s[2] = X * Y;
  } else if (axis == 1) {
    s[0] = X;
    s[1] = -1;
    s[2] = X * Y;
  } else {
    s[0] = -X * Y;
    s[1] = X;
    s[2] = -1;
  }
}

/* divh, divh2, divh3, divh4 as StencilBase{2,3}d2h computes them
 * (sol/inc/stencilincludes.h:100-104,123-127; divh = 1/h, stencil.h:750). */
static void divhs(double h, double* d2, double* d3, double* d4) {
  double divh = 1.0 / h;
  *d2 = divh * divh;
  *d3 = divh * *d2;
  *d4 = *d2 * *d2;
}

double orc_stencil(const orc_desc* d, int op, int axis, int component,
                   const double* g, int64_t idx) {
  int64_t s[3];
  double d2, d3, d4;
  divhs(d->h, &d2, &d3, &d4);
  const double* v = g + idx;
  if (d->dim == 2) {
    if (op == ORC_OP_LAP) {
      strides(d, 0, s);
      return d->lap_points == 5 ? lap2_5(v, s, d2) : lap2_9(v, s, d2);
    }
    if (op == ORC_OP_BILAP) {
      strides(d, 0, s);
      switch (d->bilap_points) {
        case 13: return bilap2_13(v, s, d4);
        case 17: return bilap2_17(v, s, d4);
        default: return bilap2_21(v, s, d4);
      }
    }
    strides(d, axis, s);
    switch (d->gradlap_points) {
      case 6: return gradlap2_6(v, s, d3, component);
      case 8: return gradlap2_8(v, s, d3);
      case 12: return gradlap2_12(v, s, d3);
      default: return gradlap2_16(v, s, d3);
    }
  }
  if (op == ORC_OP_LAP) {
    strides(d, 0, s);
    switch (d->lap_points) {
      case 7: return lap3_7(v, s, d2);
      case 15: return lap3_15(v, s, d2);
      case 19: return lap3_19(v, s, d2);
      case 21: return lap3_21(v, s, d2);
      default: return lap3_27(v, s, d2);
    }
  }
  if (op == ORC_OP_BILAP) {
    strides(d, 0, s);
    switch (d->bilap_points) {
      case 21: return bilap3_21(v, s, d4);
      case 25: return bilap3_25(v, s, d4);
      case 41: return bilap3_41(v, s, d4);
      case 52: return bilap3_52(v, s, d4);
      default: return bilap3_57(v, s, d4);
    }
  }
  strides(d, axis, s);
  switch (d->gradlap_points) {
    case 10: return gradlap3_10(v, s, d3);
    case 12: return gradlap3_12(v, s, d3);
    case 28: return gradlap3_28(v, s, d3);
    case 36: return gradlap3_36(v, s, d3);
    default: return gradlap3_40(v, s, d3);
  }
}

/* ------------------------------------------------------------- RHS */

/* psi^k as a product of k copies (lib/inc/spslibmath.h:255-269); P summed in
 * increasing degree, first non-zero coefficient starting the sum. */
static double poly_eval(const double* c, double x) {
  double acc = 0.0;
  int started = 0;
  for (int k = 0; k <= ORC_MAX_DEGREE; ++k) {
    if (c[k] == 0.0) continue;
    double p = 1.0;
    if (k >= 1) {
      p = x;
      for (int j = 1; j < k; ++j) p = p * x;
    }
    double t = c[k] * p;
    acc = started ? acc + t : t;
    started = 1;
  }
  return acc;
}

static int poly_is_identity(const double* c) {
  if (c[1] != 1.0) return 0;
  for (int k = 0; k <= ORC_MAX_DEGREE; ++k)
    if (k != 1 && c[k] != 0.0) return 0;
  return 1;
}

/* iterate the interior sites in linear (x fastest) order */
#define FOR_INTERIOR(d, IDX, BODY)                                         \
  do {                                                                     \
    const int64_t X_ = pdim(d, 0), Y_ = pdim(d, 1);                        \
    const int64_t NZ_ = (d)->dim == 3 ? (d)->n[2] : 1;                     \
    const int64_t Z0_ = (d)->dim == 3 ? ORC_DEPTH : 0;                     \
    const int64_t NY_ = (d)->n[1], NX_ = (d)->n[0];                        \
    _Pragma("omp parallel for collapse(2) schedule(static) num_threads(nt_)") \
    for (int64_t z_ = 0; z_ < NZ_; ++z_)                                   \
      for (int64_t y_ = 0; y_ < NY_; ++y_) {                               \
        int64_t row_ = X_ * ((y_ + ORC_DEPTH) + Y_ * (z_ + Z0_));          \
        for (int64_t x_ = 0; x_ < NX_; ++x_) {                             \
          const int64_t IDX = row_ + x_ + ORC_DEPTH;                       \
          BODY;                                                            \
        }                                                                  \
      }                                                                    \
  } while (0)

static int nthreads_or_default(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void orc_apply(const orc_desc* d, int op, int axis, int component,
               const double* g, double* out, int nthreads) {
  int nt_ = nthreads_or_default(nthreads);
  FOR_INTERIOR(d, idx, out[idx] = orc_stencil(d, op, axis, component, g, idx));
}

/* sum of terms[0..nterms) over the interior into out; fields[0] = psi,
 * fields[k] = var(k).  A derivative of a non-trivial operand goes through a
 * temporary grid filled over the whole padded array
 * (expressionderivatives.h:1583-1588). */
static int eval_terms(const orc_desc* d, const orc_term* terms, int nterms,
                      const double* const* fields, double* out, int nthreads) {
  const size_t len = orc_padded_len(d);
  int nt_ = nthreads_or_default(nthreads);
  double* tmp[ORC_MAX_TERMS] = {0};
  for (int t = 0; t < nterms; ++t) {
    const orc_term* tm = &terms[t];
    if (tm->op == ORC_OP_ID || poly_is_identity(tm->poly)) continue;
    tmp[t] = (double*)malloc(len * sizeof(double));
    if (!tmp[t]) {
      for (int u = 0; u < t; ++u) free(tmp[u]);
      return -2;
    }
    double* tt = tmp[t];
    const double* pc = tm->poly;
    const double* src = fields[tm->operand];
#pragma omp parallel for schedule(static) num_threads(nt_)
    for (int64_t i = 0; i < (int64_t)len; ++i) tt[i] = poly_eval(pc, src[i]);
  }
  FOR_INTERIOR(d, idx, {
    double acc = 0.0;
    for (int t = 0; t < nterms; ++t) {
      const orc_term* tm = &terms[t];
      const double* u = fields[tm->operand];
      double val;
      if (tm->op == ORC_OP_ID) {
        val = tm->coeff * poly_eval(tm->poly, u[idx]);
      } else {
        const double* src = tmp[t] ? tmp[t] : u;
        val = tm->coeff *
              orc_stencil(d, tm->op, tm->axis, tm->component, src, idx);
      }
      acc = t == 0 ? val : acc + val;
    }
    out[idx] = acc;
  });
  for (int t = 0; t < nterms; ++t) free(tmp[t]);
  return 0;
}

/* var(1..upto) into vars[0..upto), each refreshed after its evaluation */
static int eval_provisionals(const orc_desc* d, const double* psi, int upto,
                             double** vars, int nthreads) {
  const double* fields[ORC_MAX_PROV + 1] = {psi, 0, 0};
  for (int k = 1; k <= upto; ++k) {
    int rc = eval_terms(d, d->prov[k - 1], d->prov_nterms[k - 1], fields,
                        vars[k - 1], nthreads);
    if (rc) return rc;
    orc_refresh(d, vars[k - 1]);
    fields[k] = vars[k - 1];
  }
  return 0;
}

int orc_provisional(const orc_desc* d, const double* psi, int k, double* out,
                    int nthreads) {
  if (orc_check(d) != 0 || k < 1 || k > d->nprov) return -1;
  const size_t len = orc_padded_len(d);
  double* vars[ORC_MAX_PROV] = {0};
  for (int j = 0; j < k - 1; ++j) {
    vars[j] = (double*)calloc(len, sizeof(double));
    if (!vars[j]) {
      for (int u = 0; u < j; ++u) free(vars[u]);
      return -2;
    }
  }
  vars[k - 1] = out;
  int rc = eval_provisionals(d, psi, k, vars, nthreads);
  for (int j = 0; j < k - 1; ++j) free(vars[j]);
  return rc;
}

int orc_equation(const orc_desc* d, const double* psi, double* dframe,
                 int nthreads) {
  if (orc_check(d) != 0) return -1;
  const size_t len = orc_padded_len(d);
  double* vars[ORC_MAX_PROV] = {0};
  const double* fields[ORC_MAX_PROV + 1] = {psi, 0, 0};
  int rc = 0;
  for (int j = 0; j < d->nprov && rc == 0; ++j) {
    vars[j] = (double*)calloc(len, sizeof(double));
    if (!vars[j]) rc = -2;
  }
  if (rc == 0 && d->nprov > 0)
    rc = eval_provisionals(d, psi, d->nprov, vars, nthreads);
  for (int j = 0; j < d->nprov; ++j) fields[j + 1] = vars[j];
  if (rc == 0) rc = eval_terms(d, d->terms, d->nterms, fields, dframe, nthreads);
  for (int j = 0; j < d->nprov; ++j) free(vars[j]);
  return rc;
}

void orc_euler(const orc_desc* d, double* psi, const double* dframe,
               int nthreads) {
  /* psi + dt*dframe over iterable_domain (solverft.h:46-50) */
  const double dt = d->dt;
  int nt_ = nthreads_or_default(nthreads);
  FOR_INTERIOR(d, idx, psi[idx] = psi[idx] + dt * dframe[idx]);
}

int orc_run(const orc_desc* d, double* psi, int nsteps, int nthreads) {
  if (orc_check(d) != 0) return -1;
  const size_t len = orc_padded_len(d);
  double* dframe = (double*)calloc(len, sizeof(double));
  if (!dframe) return -2;
  orc_refresh(d, psi); /* modules-io.h:450-452 */
  for (int i = 0; i < nsteps; ++i) {
    int rc = orc_equation(d, psi, dframe, nthreads);
    if (rc) {
      free(dframe);
      return rc;
    }
    orc_euler(d, psi, dframe, nthreads);
    orc_refresh(d, psi);
  }
  free(dframe);
  return 0;
}

/* ------------------------------------------------------------- sums */

double orc_mass(const orc_desc* d, const double* psi) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int64_t nz = d->dim == 3 ? d->n[2] : 1;
  const int64_t z0 = d->dim == 3 ? ORC_DEPTH : 0;
  double m = 0.0;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < d->n[1]; ++y)
      for (int64_t x = 0; x < d->n[0]; ++x)
        m += psi[(x + ORC_DEPTH) + X * ((y + ORC_DEPTH) + Y * (z + z0))];
  return m;
}

/* Landau free energy (sol/inc/modelspecialized.h:31-45):
 *   f = -c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2,
 * with grad psi per axis the forward difference (psi[i+1] - psi[i]) / h
 * (gradient 0) or the central one (psi[i+1] - psi[i-1]) / (2h) (gradient 1,
 * the reference's generated first-order stencil at accuracy 2,
 * sol/inc/stencil.h:45,610-660).  Sequential sum over the interior, then
 * times h^D (mean 0) or divided by the site count (mean 1, the Energy data
 * process, dataprocesses/dataenergy.h:11-40).  psi must have fresh ghosts. */
double orc_energy_ex(const orc_desc* d, const double* psi, double c1,
                     double c2, double kappa, int gradient, int mean) {
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int64_t nz = d->dim == 3 ? d->n[2] : 1;
  const int64_t z0 = d->dim == 3 ? ORC_DEPTH : 0;
  const double divh = gradient == 1 ? 1.0 / (2.0 * d->h) : 1.0 / d->h;
  const int64_t st[3] = {1, X, X * Y};
  double e = 0.0;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < d->n[1]; ++y)
      for (int64_t x = 0; x < d->n[0]; ++x) {
        int64_t i = (x + ORC_DEPTH) + X * ((y + ORC_DEPTH) + Y * (z + z0));
        double p = psi[i];
        double g2 = 0.0;
        for (int a = 0; a < d->dim; ++a) {
          double back = gradient == 1 ? psi[i - st[a]] : p;
          double gd = (psi[i + st[a]] - back) * divh;
          g2 += gd * gd;
        }
        e += -c1 * 0.5 * (p * p) + c2 * 0.25 * (p * p * p * p) +
             kappa * 0.5 * g2;
      }
  if (mean) return e / ((double)d->n[0] * d->n[1] * nz);
  double hd = d->h * d->h;
  if (d->dim == 3) hd *= d->h;
  return e * hd;
}

double orc_energy(const orc_desc* d, const double* psi) {
  return orc_energy_ex(d, psi, 1.0, 1.0, 1.0, 0, 0);
}

/* ------------------------------------------------------------- inputs */

void orc_fill_uniform(double* out, size_t count, uint64_t seed, double lo,
                      double hi) {
  /* SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then the
   * variant-13 finaliser; the top 53 bits give u in [0,1). */
  uint64_t st = seed;
  const double scale = 1.0 / 9007199254740992.0; /* 2^-53 */
  for (size_t i = 0; i < count; ++i) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double u = (double)(z >> 11) * scale;
    out[i] = lo + (hi - lo) * u;
  }
}

/* ------------------------------------------------------------- checkpoint */

int64_t orc_checkpoint_text(const orc_desc* d, const double* psi, int index,
                            const double* origin, char* out, size_t cap) {
  if (orc_check(d) != 0) return -1;
  size_t n = 0;
  char tmp[128];
#define EMIT(...)                                                     \
  do {                                                                \
    int k_ = snprintf(tmp, sizeof(tmp), __VA_ARGS__);                 \
    for (int q_ = 0; q_ < k_; ++q_, ++n)                              \
      if (n < cap) out[n] = tmp[q_];                                  \
  } while (0)
  /* print_gp_header: dimension, counts, domain bounds (x, y, z) */
  EMIT("%d ", d->dim);
  for (int a = 0; a < d->dim; ++a) EMIT("%d ", d->n[a]);
  for (int a = 0; a < d->dim; ++a) {
    const double left = origin ? origin[a] : 0.0;
    EMIT("%lf %lf ", left, left + (d->n[a] - 1) * d->h);
  }
  EMIT("%d", index);
  EMIT(" %c", '@');
  EMIT("\n");
  /* save_grid: k (z), j (y), i (x) */
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < d->n[1]; ++j) {
      const int64_t row =
          X * ((j + ORC_DEPTH) + Y * (d->dim == 3 ? k + ORC_DEPTH : 0));
      for (int i = 0; i < d->n[0]; ++i) EMIT("%.10E ", psi[row + i + ORC_DEPTH]);
      EMIT("\n");
    }
#undef EMIT
  if (n < cap) out[n] = 0;
  return (int64_t)n;
}

int orc_checkpoint_read(const orc_desc* d, const char* text, double* psi,
                        int* index) {
  if (orc_check(d) != 0) return -1;
  const char* p = text;
  int used = 0, dim = 0;
  if (sscanf(p, "%d%n", &dim, &used) != 1 || dim != d->dim) return -2;
  p += used;
  for (int a = 0; a < dim; ++a) {
    int c = 0;
    if (sscanf(p, "%d%n", &c, &used) != 1 || c != d->n[a]) return -3;
    p += used;
  }
  for (int a = 0; a < dim; ++a) {
    double l, r;
    if (sscanf(p, "%lf %lf%n", &l, &r, &used) != 2) return -4;
    p += used;
  }
  int idx = 0;
  if (sscanf(p, "%d%n", &idx, &used) != 1) return -5;
  p += used;
  char buf[4] = {0};
  if (sscanf(p, " %3s%n", buf, &used) != 1 || buf[0] != '@') return -6;
  p += used;
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < d->n[1]; ++j) {
      const int64_t row =
          X * ((j + ORC_DEPTH) + Y * (d->dim == 3 ? k + ORC_DEPTH : 0));
      for (int i = 0; i < d->n[0]; ++i) {
        double v;
        if (sscanf(p, "%lf%n", &v, &used) != 1) return -7;
        p += used;
        psi[row + i + ORC_DEPTH] = v;
      }
    }
  if (index) *index = idx;
  return 0;
}

int64_t orc_plot_text(const orc_desc* d, const double* psi, int format,
                      const double* origin, char* out, size_t cap) {
  if (orc_check(d) != 0 || (format != 1 && format != 2)) return -1;
  size_t n = 0;
  char tmp[128];
#define EMIT(...)                                                     \
  do {                                                                \
    int k_ = snprintf(tmp, sizeof(tmp), __VA_ARGS__);                 \
    for (int q_ = 0; q_ < k_; ++q_, ++n)                              \
      if (n < cap) out[n] = tmp[q_];                                  \
  } while (0)
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  const double x0 = origin ? origin[0] : 0.0, y0 = origin ? origin[1] : 0.0;
  for (int k = 0; k < nz; ++k) {
    if (format == 1) {
      EMIT("% 13d ", k);
      for (int i = 0; i < d->n[0]; ++i) EMIT("% 13.2f ", x0 + i * d->h);
      EMIT("\n");
    }
    for (int j = 0; j < d->n[1]; ++j) {
      const int64_t row =
          X * ((j + ORC_DEPTH) + Y * (d->dim == 3 ? k + ORC_DEPTH : 0));
      if (format == 1) EMIT("% 13.2f ", y0 + j * d->h);
      for (int i = 0; i < d->n[0]; ++i) {
        if (format == 1) {
          EMIT("% 10lE ", psi[row + i + ORC_DEPTH]);
        } else {
          EMIT("% 10lE", psi[row + i + ORC_DEPTH]);
          if (i < d->n[0] - 1) EMIT(",");
        }
      }
      EMIT("\n");
    }
    if (format == 1 && k < nz - 1) EMIT("\n");
  }
  if (format == 1) EMIT("\n");
#undef EMIT
  if (n < cap) out[n] = 0;
  return (int64_t)n;
}

static void xdr_put_int(unsigned char* out, size_t cap, size_t* n, int v) {
  const uint32_t u = (uint32_t)v;
  for (int b = 3; b >= 0; --b, ++*n)
    if (*n < cap) out[*n] = (unsigned char)(u >> (8 * b));
}

static void xdr_put_double(unsigned char* out, size_t cap, size_t* n,
                           double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  for (int b = 7; b >= 0; --b, ++*n)
    if (*n < cap) out[*n] = (unsigned char)(u >> (8 * b));
}

int64_t orc_xdr_write(const orc_desc* d, const double* psi, int index,
                      const double* origin, unsigned char* out, size_t cap) {
  if (orc_check(d) != 0) return -1;
  size_t n = 0;
  xdr_put_int(out, cap, &n, d->dim);
  for (int a = 0; a < d->dim; ++a) xdr_put_int(out, cap, &n, d->n[a]);
  for (int a = 0; a < d->dim; ++a) {
    const double left = origin ? origin[a] : 0.0;
    xdr_put_double(out, cap, &n, left);
    xdr_put_double(out, cap, &n, left + (d->n[a] - 1) * d->h);
  }
  xdr_put_int(out, cap, &n, index);
  xdr_put_int(out, cap, &n, d->dim);
  for (int a = 0; a < d->dim; ++a) {
    const double left = origin ? origin[a] : 0.0;
    xdr_put_double(out, cap, &n, left);
    xdr_put_double(out, cap, &n, left + (d->n[a] - 1) * d->h);
  }
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < d->n[1]; ++j) {
      const int64_t row =
          X * ((j + ORC_DEPTH) + Y * (d->dim == 3 ? k + ORC_DEPTH : 0));
      for (int i = 0; i < d->n[0]; ++i)
        xdr_put_double(out, cap, &n, psi[row + i + ORC_DEPTH]);
    }
  return (int64_t)n;
}

static int xdr_get_int(const unsigned char* in, size_t len, size_t* p,
                       int* v) {
  if (*p + 4 > len) return -1;
  uint32_t u = 0;
  for (int b = 0; b < 4; ++b) u = (u << 8) | in[(*p)++];
  *v = (int)u;
  return 0;
}

static int xdr_get_double(const unsigned char* in, size_t len, size_t* p,
                          double* v) {
  if (*p + 8 > len) return -1;
  uint64_t u = 0;
  for (int b = 0; b < 8; ++b) u = (u << 8) | in[(*p)++];
  memcpy(v, &u, 8);
  return 0;
}

int orc_xdr_read(const orc_desc* d, const unsigned char* in, size_t len,
                 double* psi, int* index) {
  if (orc_check(d) != 0) return -1;
  size_t p = 0;
  int dim = 0, c = 0, idx = 0, dim2 = 0;
  double lr;
  if (xdr_get_int(in, len, &p, &dim) || dim != d->dim) return -2;
  for (int a = 0; a < dim; ++a)
    if (xdr_get_int(in, len, &p, &c) || c != d->n[a]) return -3;
  for (int a = 0; a < 2 * dim; ++a)
    if (xdr_get_double(in, len, &p, &lr)) return -4;
  if (xdr_get_int(in, len, &p, &idx)) return -5;
  if (xdr_get_int(in, len, &p, &dim2) || dim2 != dim) return -6;
  for (int a = 0; a < 2 * dim; ++a)
    if (xdr_get_double(in, len, &p, &lr)) return -4;
  const int64_t X = pdim(d, 0), Y = pdim(d, 1);
  const int nz = d->dim == 3 ? d->n[2] : 1;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < d->n[1]; ++j) {
      const int64_t row =
          X * ((j + ORC_DEPTH) + Y * (d->dim == 3 ? k + ORC_DEPTH : 0));
      for (int i = 0; i < d->n[0]; ++i)
        if (xdr_get_double(in, len, &p, &psi[row + i + ORC_DEPTH])) return -7;
    }
  if (index) *index = idx;
  return 0;
}

/* ------------------------------------------------------------- coupled fields */

size_t orcm_desc_size(void) { return sizeof(orcm_desc); }

/* the single-field descriptor with the same grid and stencils */
static orc_desc orcm_geometry(const orcm_desc* d) {
  orc_desc g;
  memset(&g, 0, sizeof(g));
  g.dim = d->dim;
  g.dtype = d->dtype;
  for (int a = 0; a < 3; ++a) g.n[a] = d->n[a];
  g.lap_points = d->lap_points;
  g.gradlap_points = d->gradlap_points;
  g.bilap_points = d->bilap_points;
  g.path = d->path;
  g.h = d->h;
  g.dt = d->dt;
  return g;
}

int orcm_check(const orcm_desc* d) {
  if (!d) return -1;
  orc_desc g = orcm_geometry(d);
  int rc = orc_check(&g);
  if (rc) return rc;
  if (d->nfields < 1 || d->nfields > ORC_MAX_FIELDS) return -12;
  for (int f = 0; f < d->nfields; ++f) {
    if (d->nterms[f] < 0 || d->nterms[f] > ORC_MAX_MTERMS) return -13;
    for (int t = 0; t < d->nterms[f]; ++t) {
      const orcm_term* tm = &d->terms[f][t];
      if (tm->op < ORC_OP_ID || tm->op > ORC_OP_GRADLAP) return -6;
      int deg = 0;
      for (int g2 = 0; g2 < ORC_MAX_FIELDS; ++g2) {
        if (tm->pow[g2] < 0 || (g2 >= d->nfields && tm->pow[g2] != 0)) return -14;
        deg += tm->pow[g2];
      }
      if (deg > ORC_MAX_DEGREE) return -14;
      if (tm->op == ORC_OP_GRADLAP) {
        if (tm->axis < 0 || tm->axis >= d->dim) return -7;
        if (tm->component < 0 ||
            tm->component >= ((d->dim == 2 && d->gradlap_points == 6) ? 2 : 1))
          return -8;
      }
    }
  }
  return 0;
}

/* left fold of the factors psi_0^p0, psi_1^p1, ... in field order; 1 for the
 * empty product */
static double monomial(const orcm_term* tm, const double* fields, size_t len,
                       int nf, int64_t idx) {
  double acc = 1.0;
  int started = 0;
  for (int f = 0; f < nf; ++f)
    for (int k = 0; k < tm->pow[f]; ++k) {
      const double x = fields[(size_t)f * len + idx];
      acc = started ? acc * x : x;
      started = 1;
    }
  return acc;
}

/* the field read directly when the monomial is exactly one field, else -1 */
static int monomial_field(const orcm_term* tm, int nf) {
  int which = -1, deg = 0;
  for (int f = 0; f < nf; ++f) {
    deg += tm->pow[f];
    if (tm->pow[f] == 1) which = f;
  }
  return deg == 1 ? which : -1;
}

int orcm_equation(const orcm_desc* d, const double* fields, double* dframe,
                  int nthreads) {
  if (orcm_check(d) != 0) return -1;
  const orc_desc g = orcm_geometry(d);
  const orc_desc* gd = &g;
  const size_t len = orc_padded_len(gd);
  const int nf = d->nfields;
  int nt_ = nthreads_or_default(nthreads);
  for (int f = 0; f < nf; ++f) {
    double* tmp[ORC_MAX_MTERMS] = {0};
    for (int t = 0; t < d->nterms[f]; ++t) {
      const orcm_term* tm = &d->terms[f][t];
      if (tm->op == ORC_OP_ID || monomial_field(tm, nf) >= 0) continue;
      tmp[t] = (double*)malloc(len * sizeof(double));
      if (!tmp[t]) {
        for (int u = 0; u < t; ++u) free(tmp[u]);
        return -2;
      }
      double* tt = tmp[t];
#pragma omp parallel for schedule(static) num_threads(nt_)
      for (int64_t i = 0; i < (int64_t)len; ++i)
        tt[i] = monomial(tm, fields, len, nf, i);
    }
    double* out = dframe + (size_t)f * len;
    FOR_INTERIOR(gd, idx, {
      double acc = 0.0;
      for (int t = 0; t < d->nterms[f]; ++t) {
        const orcm_term* tm = &d->terms[f][t];
        double val;
        if (tm->op == ORC_OP_ID) {
          val = tm->coeff * monomial(tm, fields, len, nf, idx);
        } else {
          const int which = monomial_field(tm, nf);
          const double* src =
              tmp[t] ? tmp[t] : fields + (size_t)(which < 0 ? 0 : which) * len;
          val = tm->coeff *
                orc_stencil(gd, tm->op, tm->axis, tm->component, src, idx);
        }
        acc = t == 0 ? val : acc + val;
      }
      out[idx] = acc;
    });
    for (int t = 0; t < d->nterms[f]; ++t) free(tmp[t]);
  }
  return 0;
}

int orcm_run(const orcm_desc* d, double* fields, int nsteps, int nthreads) {
  if (orcm_check(d) != 0) return -1;
  const orc_desc g = orcm_geometry(d);
  const size_t len = orc_padded_len(&g);
  const int nf = d->nfields;
  double* dframe = (double*)calloc(len * (size_t)nf, sizeof(double));
  if (!dframe) return -2;
  for (int f = 0; f < nf; ++f) orc_refresh(&g, fields + (size_t)f * len);
  for (int i = 0; i < nsteps; ++i) {
    int rc = orcm_equation(d, fields, dframe, nthreads);
    if (rc) {
      free(dframe);
      return rc;
    }
    for (int f = 0; f < nf; ++f) {
      orc_euler(&g, fields + (size_t)f * len, dframe + (size_t)f * len, nthreads);
      orc_refresh(&g, fields + (size_t)f * len);
    }
  }
  free(dframe);
  return 0;
}

