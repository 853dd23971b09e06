/*
 * sps_oracle.h - CPU restatement of the SymPhas SolverFT step, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  Nothing in the product links or calls it.
 *
 * It follows the reference's algorithm on the reference's padded layout
 * (BOUNDARY_DEPTH = 3 ghost cells per side, datatypes/inc/gridinfo.h:46; x
 * fastest, datatypes/inc/grid.h:537-578):
 *   - periodic ghost refresh       datatypes/inc/boundaryupdate.h:154-356,650-708
 *   - stencils, reference grouping sol/inc/stencilh2.h:87-507
 *   - strides per axis             datatypes/inc/grid.h:537-578
 *   - derivative of a non-trivial operand through a temporary grid filled
 *     over the whole padded array  sym/inc/expressionderivatives.h:1583-1588
 *   - RHS over the interior        examples/solvers/solverft.h:90-103,
 *                                  datatypes/inc/dataiterator.h:2443-2450
 *   - forward Euler, interior only examples/solvers/solverft.h:46-50
 *   - sequencing                   modules-models.h:436-492, modules-io.h:450-452
 *   - provisional variables: each evaluated over the interior into its own
 *     grid after the field's refresh, then that grid's ghosts refreshed
 *                                  sol/inc/modelspecialized.h:430-434,
 *                                  sol/inc/provisionalsystem.h:72-99,
 *                                  examples/solvers/solverft.h:163-183
 *
 * Parity pins (tests/test_oracle.py): every stencil is checked against the
 * moment conditions of a consistent 2nd-order stencil, the Fourier symbols of
 * lap5/lap9/lap7 and bilap = lap∘lap for <13>/<25>, an independent numpy
 * np.roll implementation on tiny grids, the exact single-mode recursion of
 * dpsi/dt = lap psi, constant-field recursions, mass conservation, and the
 * 2D ghost-strip copies against modular indexing.  Provisional variables are
 * pinned by MBB (var = P(psi), dpsi = -bilap psi + lap var) matching the
 * same right-hand side written without a provisional (STATE_SPEED,
 * tests/testdefs.h:23-26) bit for bit, by BB (var = lap psi + P(psi),
 * dpsi = -lap var) against the exact single-mode recursion with symbol
 * -(lambda^2 + a*lambda), and by an np.roll restatement.  The gradlaplacian Z-axis
 * sign and the lap19 "vz + vz" term are reproduced and pinned by special
 * cases, and so is the 2D 21-point bilaplacian's 1/12 factor (stencilh2.h:
 * 143-157), which makes it return a quarter of the bilaplacian.  Parity is
 * unpinned only for the reference's choice of term order
 * inside its expression optimiser (expr::transform::optimize may reorder or
 * distribute terms, e.g. -(psi^2-1)*psi versus psi - psi^3); the oracle sums
 * terms in descriptor order and powers in increasing degree.  That choice
 * changes FP64 rounding only (a relative 1e-16 per step).
 */
#ifndef SPS_ORACLE_H
#define SPS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DEPTH 3
#define ORC_MAX_DEGREE 6
#define ORC_MAX_TERMS 8
#define ORC_MAX_PROV 2

enum { ORC_OP_ID = 0, ORC_OP_LAP = 1, ORC_OP_BILAP = 2, ORC_OP_GRADLAP = 3 };

/* operand: 0 = the field psi, k >= 1 = provisional variable var(k) */
typedef struct {
  int32_t op, axis, component, operand;
  double coeff;
  double poly[ORC_MAX_DEGREE + 1];
} orc_term;

/* Same memory layout as the product's problem descriptor, so one host-side
 * structure can describe a problem to both. */
typedef struct {
  int32_t dim, dtype;
  int32_t n[3];
  int32_t lap_points, gradlap_points, bilap_points;
  int32_t nterms, path;
  double h, dt;
  orc_term terms[ORC_MAX_TERMS];
  int32_t nprov;
  int32_t prov_nterms[ORC_MAX_PROV];
  int32_t prov_reserved;
  orc_term prov[ORC_MAX_PROV][ORC_MAX_TERMS];
} orc_desc;

size_t orc_desc_size(void);
/* 0 if the descriptor is valid for the oracle, else a negative code. */
int orc_check(const orc_desc* d);
/* Number of values in the padded array. */
size_t orc_padded_len(const orc_desc* d);

/* interior <-> padded copies (ghosts of the padded array are left alone) */
void orc_pad(const orc_desc* d, const double* interior, double* padded);
void orc_unpad(const orc_desc* d, const double* padded, double* interior);

/* Periodic ghost refresh of a padded array. 2D follows the reference's strip
 * copies; 3D writes each ghost from its modular interior image (the combined
 * effect of the 6 face, 12 edge and 8 corner copies). */
void orc_refresh(const orc_desc* d, double* g);
/* Modular-index refresh, used by the tests to pin orc_refresh. */
void orc_refresh_modular(const orc_desc* d, double* g);

/* One derivative at padded linear index idx (op: ORC_OP_LAP/BILAP/GRADLAP). */
double orc_stencil(const orc_desc* d, int op, int axis, int component,
                   const double* g, int64_t idx);
/* Apply one derivative over the interior into out (padded array; ghosts of
 * out untouched). */
void orc_apply(const orc_desc* d, int op, int axis, int component,
               const double* g, double* out, int nthreads);

/* Provisional variable var(k) (1-based) over the interior into out (padded),
 * its ghosts then refreshed periodically: the order of Model::update
 * (sol/inc/modelspecialized.h:430-434).  psi must have fresh ghosts.
 * Variables below k are evaluated first, as the reference does in turn. */
int orc_provisional(const orc_desc* d, const double* psi, int k, double* out,
                    int nthreads);
/* RHS of the descriptor's terms over the interior into dframe (padded);
 * provisional variables, if any, are evaluated first. */
int orc_equation(const orc_desc* d, const double* psi, double* dframe,
                 int nthreads);
/* psi = psi + dt*dframe over the interior. */
void orc_euler(const orc_desc* d, double* psi, const double* dframe,
               int nthreads);
/* Refresh, then nsteps x (equation, euler, refresh).  psi is padded. */
int orc_run(const orc_desc* d, double* psi, int nsteps, int nthreads);

/* Sequential FP64 sum of the interior (dataprocesses/datadensity.h:11-41). */
double orc_mass(const orc_desc* d, const double* psi);
/* Landau free energy h^D * sum[-c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2]
 * (mean: the sum divided by the site count instead), grad by forward
 * (gradient 0) or central (gradient 1) differences; psi must have fresh
 * ghosts.  orc_energy is c1 = c2 = kappa = 1, forward, integral. */
double orc_energy_ex(const orc_desc* d, const double* psi, double c1,
                     double c2, double kappa, int gradient, int mean);
double orc_energy(const orc_desc* d, const double* psi);

/* The reference's GNU checkpoint file of a field (io/src/writegp.cpp:
 * print_gp_header :427-465, save_grid :780-803): a header line
 * "dim n0 n1 [n2] left0 right0 ... index @" (counts with %d, domain bounds
 * with %lf, right = left + (n-1)*h), then every interior value as "%.10E "
 * in x-fastest order with "\n" after each x row.  Writes at most cap bytes
 * into out; returns the full length, or -1 on a bad descriptor.  origin may
 * be NULL (all zeros).  psi is padded. */
int64_t orc_checkpoint_text(const orc_desc* d, const double* psi, int index,
                            const double* origin, char* out, size_t cap);
/* Read such a file as io/src/readgp.cpp does (read_header, read_block with
 * fscanf("%lf")): checks the header against d, stores the index and the
 * values into the interior of psi (padded).  Returns 0, or a negative code. */
int orc_checkpoint_read(const orc_desc* d, const char* text, double* psi,
                        int* index);

/* The reference's plot data files of a real field.
 * format 1, GNU matrix (io/src/writegp.cpp save_grid_plotting :470-507): per
 *   z slice k a line "% 13d " k, then "% 13.2f " per x position, "\n"; then
 *   per y row "% 13.2f " y position and "% 10lE " per value, "\n"; a blank
 *   line between slices and a final "\n".  Positions are origin + i*h.
 * format 2, CSV (io/src/writecsv.cpp save_grid_plotting): every x row as
 *   "% 10lE" values joined by ',' and ended by "\n", rows in y then z order.
 * Returns the full length (at most cap bytes written), or -1. */
int64_t orc_plot_text(const orc_desc* d, const double* psi, int format,
                      const double* origin, char* out, size_t cap);

/* The reference's XDR file of a field (io/src/writexdr.cpp print_xdr_header
 * :45-87, save_xdr :92-98; xdrfile's XDR encoding, RFC 4506: 4-byte
 * big-endian ints, 8-byte big-endian IEEE doubles): int dim, int counts[dim],
 * per axis double (left, right) of the domain, int index, int dim, per axis
 * double (left, right) of the interval, then every interior value as a
 * double in x-fastest order.  Returns the length (at most cap bytes
 * written), or -1. */
int64_t orc_xdr_write(const orc_desc* d, const double* psi, int index,
                      const double* origin, unsigned char* out, size_t cap);
/* Read such a file as io/src/readxdr.cpp does (read_header, read_block):
 * checks dim and counts against d, stores the index and the values into the
 * interior of psi (padded).  Returns 0, or a negative code. */
int orc_xdr_read(const orc_desc* d, const unsigned char* in, size_t len,
                 double* psi, int* index);

/* ------------------------------------------------------------ coupled fields
 * Several fields stepped together (Model C, Gray-Scott, Turing:
 * examples/models/modeldefinitions.h:63-81,123-137).  Each field's
 * right-hand side is a sum of coeff * D[prod_f psi_f^pow[f]]; a monomial is
 * the left-fold product of its factors in field order (psi*rho*rho =
 * (psi*rho)*rho, as the reference evaluates a written product), a derivative
 * of a non-trivial monomial goes through a temporary grid over the padded
 * array, and every field gets Euler and the periodic refresh.  Same layout
 * as the product's sps_fdm_desc. */
#define ORC_MAX_FIELDS 4
#define ORC_MAX_MTERMS 12

typedef struct {
  int32_t op, axis, component, reserved;
  double coeff;
  int32_t pow[ORC_MAX_FIELDS];
} orcm_term;

typedef struct {
  int32_t dim, dtype;
  int32_t n[3];
  int32_t lap_points, gradlap_points, bilap_points;
  int32_t nfields, path;
  double h, dt;
  int32_t nterms[ORC_MAX_FIELDS];
  orcm_term terms[ORC_MAX_FIELDS][ORC_MAX_MTERMS];
} orcm_desc;

size_t orcm_desc_size(void);
int orcm_check(const orcm_desc* d);
/* fields: nfields padded arrays back to back (field f at f * padded_len);
 * dframe likewise. */
int orcm_equation(const orcm_desc* d, const double* fields, double* dframe,
                  int nthreads);
/* Refresh, then nsteps x (all right-hand sides, Euler on each field,
 * refresh of each field). */
int orcm_run(const orcm_desc* d, double* fields, int nsteps, int nthreads);

/* ------------------------------------------------------------ regional grid
 * SolverFT on a regional grid (SolverSystemFDwSD, sol/inc/solversystem.h:62-77):
 * the field is stored only on a box of the periodic domain plus a boundary
 * layer of ORC_DEPTH cells, every other site reads as `empty`, and the box
 * follows the sites above a cutoff.  sps_oracle_regional.c states the rules.
 * Same layout as the product's sps_fd_regional. */
typedef struct {
  double cutoff_eps;        /* params::regional_resize_cutoff_eps (1e-6) */
  double resize_factor;     /* params::regional_resize_factor (2.0), >= 1 */
  double resize_time;       /* params::regional_resize_time (3.0), > 0 */
  int32_t fixed_resize;     /* params::regional_resize_is_fixed (1) */
  int32_t reserved;
  double dims_relative_eps; /* REGIONAL_GRID_RELATIVE_DIMS_EPS (0.05) */
} orc_regional;

/* The box after a run: origin = domain coordinate of the first interior site
 * per axis, dims = interior + 2*ORC_DEPTH (all 0 when no site was above the
 * cutoff), and the value outside, the cutoff and the time reached. */
typedef struct {
  int32_t origin[3];
  int32_t dims[3];
  double empty, cutoff, time;
} orc_region;

/* Shortest periodic interval of [0, n) covering every coordinate with
 * occ[i] != 0: the complement of the largest run of zeros on the circle (the
 * lowest-starting one among equals).  Returns 0 (start 0, length 0) when no
 * coordinate is set, else 1. */
int orc_cover_interval(const unsigned char* occ, int n, int* start,
                       int* length);

/* nsteps of (update, equation, Euler) from a whole-domain field (interior
 * layout, n values, replaced by the result: box interior values inside the
 * box, `empty` outside).  The box starts as the whole domain; the first
 * update fixes cutoff = min + eps and empty = min and resizes.  Provisional
 * variables are not supported. */
int orc_regional_run(const orc_desc* d, const orc_regional* rp, double* field,
                     int nsteps, orc_region* st, int nthreads);

/* SplitMix64 stream -> 53-bit uniform doubles in [lo, hi). */
void orc_fill_uniform(double* out, size_t count, uint64_t seed, double lo,
                      double hi);

#ifdef __cplusplus
}
#endif

#endif
