/*
 * sps_oracle_regional.c - CPU restatement of SolverFT on a regional grid
 * (SolverSystemFDwSD), FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP regional plan, never part
 * of the product.  Built with sps_oracle.c by oracle/Makefile.
 *
 * The reference stores a field on a RegionalGrid: a box of the periodic
 * domain with a boundary layer of BOUNDARY_DEPTH cells on each side
 * (datatypes/inc/grid.h:2038-2157, select_region :1222-1300); every site
 * outside the box reads as the grid's `empty` value.  The solver evaluates
 * the right-hand side and the Euler update over the box interior only
 * (the iterable domain of a regional grid), and SolverSystemFDwSD::update
 * (sol/inc/solversystem.h:62-77) moves and resizes the box to the bounding
 * box of the sites above a cutoff (sol/inc/boundarysystem.h:437-479,
 * datatypes/inc/gridfunctions.h:2166-2238).  This file restates that as:
 *
 *   box        origin o[a] = interior coordinate of the box's first interior
 *              site, dims b[a] = interior m[a] + 2*3; local position l along
 *              axis a is the domain site (o[a] + l - 3) mod n[a];
 *   step       the descriptor's right-hand side and forward Euler over the
 *              box interior, with the box array as the padded array (the
 *              layer supplies the neighbours);
 *   layer      along an axis the box spans (m[a] == n[a]) the layer is the
 *              periodic image of the interior, refreshed before every
 *              right-hand side (regional_update_boundary,
 *              datatypes/inc/boundaryupdatedata.h:1398-1433, updates the
 *              layer where the region meets the domain edge); along other
 *              axes it keeps the values it was given when the box last
 *              moved;
 *   update     at the first update: cutoff = min + eps and empty = min of
 *              the field (boundarysystem.h:441-447), then resize by the
 *              padding factor; afterwards, once time / resize_time reaches
 *              the next resize count: with fixed_resize re-centre the box of
 *              the same dims on the active sites (adjust_region,
 *              gridfunctions.h:2166-2176), else resize with the relative
 *              dims tolerance (resize_adjust_region, :2194-2238);
 *   active box per axis, the shortest periodic interval covering the
 *              coordinates of every interior site whose value exceeds the
 *              cutoff (the complement of the largest periodic gap; ties go
 *              to the gap that starts lowest);
 *   move       a site of the new box takes the value of the old box's
 *              interior at the same domain site, or `empty` where the old
 *              interior does not cover it.
 *
 * Where this departs from the reference, and why, is stated in the product's
 * header (include/sps_fd.h, sps_fdr_*): the reference grows its box from the
 * first active site until an empty slab (get_view_resized_periodic,
 * gridfunctions.h:2011-2163), which drops clusters separated by an empty
 * slab; the covering interval here keeps every active site.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sps_oracle.h"

static int64_t box_len(const orc_desc* d, const int32_t* b) {
  int64_t len = (int64_t)b[0] * b[1];
  if (d->dim == 3) len *= b[2];
  return len;
}

static int mod_n(int64_t v, int n) {
  int64_t r = v % n;
  return (int)(r < 0 ? r + n : r);
}

int orc_cover_interval(const unsigned char* occ, int n, int* start,
                       int* length) {
  int first = -1;
  for (int i = 0; i < n; ++i)
    if (occ[i]) {
      first = i;
      break;
    }
  if (first < 0) {
    *start = 0;
    *length = 0;
    return 0;
  }
  /* walk the circle from the first occupied coordinate; every gap is a run
   * of empty coordinates ending at an occupied one */
  int best_len = 0, best_start = 0;
  int run = 0, run_start = 0;
  for (int k = 1; k <= n; ++k) {
    const int i = (first + k) % n;
    if (!occ[i]) {
      if (run == 0) run_start = i;
      ++run;
    } else {
      if (run > best_len || (run == best_len && run > 0 && run_start < best_start)) {
        best_len = run;
        best_start = run_start;
      }
      run = 0;
    }
  }
  if (best_len == 0) {
    *start = 0;
    *length = n;
  } else {
    *start = (best_start + best_len) % n;
    *length = n - best_len;
  }
  return 1;
}

/* Active box of the sites above the cutoff, per axis (start, length). */
static int active_box(const orc_desc* d, const double* box, const int32_t* o,
                      const int32_t* b, double cutoff, int* s, int* L) {
  const int D = d->dim;
  unsigned char* occ[3] = {0, 0, 0};
  int rc = 1;
  for (int a = 0; a < D; ++a) {
    occ[a] = (unsigned char*)calloc((size_t)d->n[a], 1);
    if (!occ[a]) rc = -1;
  }
  if (rc == 1) {
    const int nz = D == 3 ? b[2] - 2 * ORC_DEPTH : 1;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < b[1] - 2 * ORC_DEPTH; ++y)
        for (int x = 0; x < b[0] - 2 * ORC_DEPTH; ++x) {
          const int64_t idx =
              (x + ORC_DEPTH) +
              (int64_t)b[0] * ((y + ORC_DEPTH) +
                               (int64_t)b[1] * (D == 3 ? z + ORC_DEPTH : 0));
          if (box[idx] > cutoff) {
            occ[0][mod_n((int64_t)o[0] + x, d->n[0])] = 1;
            occ[1][mod_n((int64_t)o[1] + y, d->n[1])] = 1;
            if (D == 3) occ[2][mod_n((int64_t)o[2] + z, d->n[2])] = 1;
          }
        }
    for (int a = 0; a < D && rc == 1; ++a)
      rc = orc_cover_interval(occ[a], d->n[a], &s[a], &L[a]);
  }
  for (int a = 0; a < D; ++a) free(occ[a]);
  return rc;
}

/* Copy the old box into a new one: a new site takes the old interior value
 * at the same domain site, else `empty`. */
static void move_box(const orc_desc* d, const double* old, const int32_t* oo,
                     const int32_t* ob, double* nw, const int32_t* no,
                     const int32_t* nb, double empty) {
  const int D = d->dim;
  const int nz = D == 3 ? nb[2] : 1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < nb[1]; ++y)
      for (int x = 0; x < nb[0]; ++x) {
        const int l[3] = {x, y, z};
        int64_t src = 0, stride = 1;
        int inside = 1;
        for (int a = 0; a < D; ++a) {
          const int g = mod_n((int64_t)no[a] + l[a] - ORC_DEPTH, d->n[a]);
          const int k = mod_n((int64_t)g - oo[a], d->n[a]);
          if (k >= ob[a] - 2 * ORC_DEPTH) inside = 0;
          src += (int64_t)(k + ORC_DEPTH) * stride;
          stride *= ob[a];
        }
        const int64_t dst = x + (int64_t)nb[0] * (y + (int64_t)nb[1] * z);
        nw[dst] = inside ? old[src] : empty;
      }
}

/* Periodic layer along every axis the box spans, axes in order x, y, z. */
static void refresh_full_axes(const orc_desc* d, double* box,
                              const int32_t* b) {
  const int D = d->dim;
  for (int a = 0; a < D; ++a) {
    if (b[a] - 2 * ORC_DEPTH != d->n[a]) continue;
    const int n = d->n[a];
    const int nz = D == 3 ? b[2] : 1;
    int64_t sa = 1;
    for (int j = 0; j < a; ++j) sa *= b[j];
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < b[1]; ++y)
        for (int x = 0; x < b[0]; ++x) {
          const int l[3] = {x, y, z};
          const int64_t idx = x + (int64_t)b[0] * (y + (int64_t)b[1] * z);
          if (l[a] < ORC_DEPTH)
            box[idx] = box[idx + (int64_t)n * sa];
          else if (l[a] >= b[a] - ORC_DEPTH)
            box[idx] = box[idx - (int64_t)n * sa];
        }
  }
}

/* New box from the active box (s, L): resize_adjust_region when resize,
 * else adjust_region (same dims, re-centred).  Integer division truncates
 * toward zero, as in the reference. */
static void place_box(const orc_desc* d, const int* s, const int* L,
                      const int32_t* ob, int resize, double factor,
                      double rel_eps, int32_t* no, int32_t* nb) {
  const int D = d->dim;
  int32_t dims_set[3] = {0, 0, 0}, origin_set[3] = {0, 0, 0};
  int same = 1;
  for (int a = 0; a < D; ++a) {
    const int dims = L[a] + 2 * ORC_DEPTH;
    if (!resize) {
      origin_set[a] = mod_n((int64_t)s[a] - (ob[a] - dims) / 2, d->n[a]);
      dims_set[a] = ob[a];
      continue;
    }
    const int grid = d->n[a] + 2 * ORC_DEPTH;
    int dim0 = (int)(dims * factor);
    if (dim0 > grid) dim0 = grid;
    /* at least 3 interior sites, and never more than the domain */
    if (dim0 < 3 + 2 * ORC_DEPTH) dim0 = 3 + 2 * ORC_DEPTH < grid ? 3 + 2 * ORC_DEPTH : grid;
    origin_set[a] = mod_n((int64_t)s[a] - (dim0 - dims) / 2, d->n[a]);
    dims_set[a] = dim0;
    if (fabs((double)(dim0 - ob[a])) / (double)ob[a] > rel_eps) same = 0;
  }
  if (resize && same) {
    for (int a = 0; a < D; ++a) {
      const int dims = L[a] + 2 * ORC_DEPTH;
      origin_set[a] = mod_n((int64_t)s[a] - (ob[a] - dims) / 2, d->n[a]);
      dims_set[a] = ob[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    no[a] = a < D ? origin_set[a] : 0;
    nb[a] = a < D ? dims_set[a] : 1;
  }
}

int orc_regional_run(const orc_desc* d, const orc_regional* rp, double* field,
                     int nsteps, orc_region* st, int nthreads) {
  if (orc_check(d) != 0 || d->nprov != 0 || !rp || !st) return -1;
  if (!(rp->resize_factor >= 1.0) || !(rp->resize_time > 0)) return -1;
  const int D = d->dim;
  int32_t o[3] = {0, 0, 0}, b[3] = {1, 1, 1};
  for (int a = 0; a < D; ++a) b[a] = d->n[a] + 2 * ORC_DEPTH;
  const int64_t cap = box_len(d, b);
  double* box = (double*)calloc((size_t)cap, sizeof(double));
  double* tmp = (double*)calloc((size_t)cap, sizeof(double));
  double* dframe = (double*)calloc((size_t)cap, sizeof(double));
  if (!box || !tmp || !dframe) {
    free(box);
    free(tmp);
    free(dframe);
    return -2;
  }
  orc_pad(d, field, box); /* the whole domain: the box is the padded array */
  int next_resize = 0, has_box = 1;
  double empty = 0.0, cutoff = 0.0, time = 0.0;
  int rc = 0;
  for (int i = 0; i < nsteps && rc == 0; ++i) {
    /* update(index, time) */
    if (next_resize == 0) {
      double mn = INFINITY;
      const int64_t len = (int64_t)d->n[0] * d->n[1] * (D == 3 ? d->n[2] : 1);
      for (int64_t k = 0; k < len; ++k) mn = field[k] < mn ? field[k] : mn;
      cutoff = mn + rp->cutoff_eps;
      empty = mn;
      int s[3], L[3];
      const int act = active_box(d, box, o, b, cutoff, s, L);
      if (act < 0) rc = -2;
      if (act == 0) {
        has_box = 0;
      } else if (act > 0) {
        int32_t no[3], nb[3];
        place_box(d, s, L, b, 1, rp->resize_factor, 0.0, no, nb);
        move_box(d, box, o, b, tmp, no, nb, empty);
        double* t = box; box = tmp; tmp = t;
        memcpy(o, no, sizeof(o));
        memcpy(b, nb, sizeof(b));
      }
      next_resize = 1;
    } else if (has_box) {
      const int cur = (int)(time / rp->resize_time);
      if (cur >= next_resize) {
        int s[3], L[3];
        const int act = active_box(d, box, o, b, cutoff, s, L);
        if (act < 0) rc = -2;
        if (act > 0) {
          int32_t no[3], nb[3];
          place_box(d, s, L, b, !rp->fixed_resize, rp->resize_factor,
                    rp->dims_relative_eps, no, nb);
          move_box(d, box, o, b, tmp, no, nb, empty);
          double* t = box; box = tmp; tmp = t;
          memcpy(o, no, sizeof(o));
          memcpy(b, nb, sizeof(b));
        }
        next_resize = cur + 1;
      }
    }
    if (has_box && rc == 0) {
      refresh_full_axes(d, box, b);
      /* the box as a padded array of its own interior */
      orc_desc bd = *d;
      for (int a = 0; a < D; ++a) bd.n[a] = b[a] - 2 * ORC_DEPTH;
      rc = orc_equation(&bd, box, dframe, nthreads);
      if (rc == 0) orc_euler(&bd, box, dframe, nthreads);
    }
    time += d->dt;
  }
  /* the whole domain: box interior values inside, empty outside */
  if (rc == 0) {
    const int nz = D == 3 ? d->n[2] : 1;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < d->n[1]; ++y)
        for (int x = 0; x < d->n[0]; ++x) {
          const int g[3] = {x, y, z};
          int64_t src = 0, stride = 1;
          int inside = has_box;
          for (int a = 0; a < D && inside; ++a) {
            const int k = mod_n((int64_t)g[a] - o[a], d->n[a]);
            if (k >= b[a] - 2 * ORC_DEPTH) inside = 0;
            src += (int64_t)(k + ORC_DEPTH) * stride;
            stride *= b[a];
          }
          const int64_t dst = x + (int64_t)d->n[0] * (y + (int64_t)d->n[1] * z);
          field[dst] = inside ? box[src] : (next_resize == 0 ? field[dst] : empty);
        }
    for (int a = 0; a < 3; ++a) {
      st->origin[a] = o[a];
      st->dims[a] = has_box ? b[a] : 0;
    }
    st->empty = empty;
    st->cutoff = cutoff;
    st->time = time;
  }
  free(box);
  free(tmp);
  free(dframe);
  return rc;
}
