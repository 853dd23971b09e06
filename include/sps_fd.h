/*
 * sps_fd.h - C ABI of the MI355X explicit finite-difference solver step.
 *
 * This is the boundary between host code that drives a SymPhas-style model
 * (the C++ adapter SolverFT_MI355X in symphas_amd/host, ctypes from Python, or
 * any FFI) and the hand-written HIP kernels for gfx950.  Signatures carry only
 * plain C types, pointers and sizes.
 *
 * Semantics follow the reference SolverFT step (SymPhas,
 * examples/solvers/solverft.h:41-184):
 *   one step = RHS evaluation into dframe  (solverft.h:90-103)
 *            + forward Euler psi += dt*dframe over the interior (solverft.h:46-50)
 *            + periodic ghost refresh (sol/inc/boundarygroup.h:373-381,
 *              datatypes/inc/boundaryupdate.h:154-356 (3D), 650-708 (2D)).
 * Ghosts are treated as fresh before the first RHS evaluation, which is what
 * the IO driver does (modules-io.h:450-452).  On the device the field is stored
 * without ghost cells and the periodic wrap is applied inside the kernels; the
 * SymPhas padded layout (BOUNDARY_DEPTH = 3 ghost cells per side,
 * datatypes/inc/gridinfo.h:46, x fastest, datatypes/inc/grid.h:537-578) exists
 * only at upload/download.
 *
 * Errors: every entry point returns an sps_status; sps_fd_last_error() gives
 * the text of the most recent failure on the calling thread.  (The reference
 * prints and exit()s or throws, lib/inc/definitions.h:42-50,
 * sol/inc/stencil.h:671-679, sol/inc/boundarygroup.h:315-316.)
 *
 * Threading: a plan or batch may be used by one host thread at a time.  Work is
 * enqueued on the object's own HIP stream; calls that return host data
 * (download, reduce) synchronise that stream.
 */
#ifndef SPS_FD_H
#define SPS_FD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPS_FD_ABI_VERSION 3

/* Ghost width of the SymPhas padded layout (datatypes/inc/gridinfo.h:46). */
#define SPS_BOUNDARY_DEPTH 3
/* Highest power of psi inside one term; pow<N> with N <= 6 is a product of N
 * copies of psi in the reference (lib/inc/spslibmath.h:255-269). */
#define SPS_MAX_DEGREE 6
#define SPS_MAX_TERMS 8
/* Provisional variables per problem (PROVISIONAL_DEF, sol/inc/modelmacros.h). */
#define SPS_MAX_PROV 2

typedef enum {
  SPS_OK = 0,
  SPS_E_INVALID = -1,     /* bad argument or descriptor */
  SPS_E_UNSUPPORTED = -2, /* valid request this library does not implement */
  SPS_E_HIP = -3,         /* HIP runtime error */
  SPS_E_NOMEM = -4,       /* device or host allocation failed */
  SPS_E_STATE = -5        /* call made in the wrong state (e.g. no data) */
} sps_status;

typedef enum { SPS_F32 = 0, SPS_F64 = 1 } sps_dtype;

/* Derivative applied to a term's operand (sol/inc/solver.h:877-915: order 2
 * laplacian, 3 gradlaplacian, 4 bilaplacian; order 0 = the operand itself). */
typedef enum {
  SPS_OP_ID = 0,
  SPS_OP_LAP = 1,
  SPS_OP_BILAP = 2,
  SPS_OP_GRADLAP = 3
} sps_op;

typedef enum { SPS_AXIS_X = 0, SPS_AXIS_Y = 1, SPS_AXIS_Z = 2 } sps_axis;

typedef enum {
  SPS_LAYOUT_INTERIOR = 0, /* n[0]*n[1]*(n[2]) values, x fastest, no ghosts */
  SPS_LAYOUT_PADDED = 1    /* SymPhas BoundaryGrid layout: (n+6) per axis */
} sps_layout;

/* Which kernel family advances the field.
 *  REFERENCE: one site per thread, every term evaluated in descriptor order
 *             with the reference's stencil groupings and no fused
 *             multiply-add; the FP64 form reproduces the CPU SolverFT bits.
 *             Provisional variables are evaluated by a pass of their own.
 *  FUSED:     register-streaming kernels on the canonical form
 *             dpsi/dt = A(psi) + lap(B(psi)) + kc*bilap(psi), valid when the
 *             bilaplacian stencil equals laplacian∘laplacian (2D <5,*,13>,
 *             3D <7,*,25>), no gradlaplacian term is present and every
 *             provisional variable substitutes into that form; several steps
 *             may be fused per launch.
 *  AUTO:      FUSED when the descriptor allows it, else REFERENCE. */
typedef enum {
  SPS_PATH_AUTO = 0,
  SPS_PATH_REFERENCE = 1,
  SPS_PATH_FUSED = 2
} sps_path;

/* One additive term of the EVOLUTION right-hand side (or of a provisional
 * variable's definition):
 *     coeff * D[ P(u) ],   P(u) = sum_k poly[k] * u^k,
 * where the operand u is the field psi or a provisional variable var(k).
 * This is the canonical form SolverFT::form_expr_one produces
 * (solverft.h:112-119): apply_operators distributes derivatives over sums
 * (sym/inc/expressionapply.h:1476-1488) and a derivative of a non-trivial
 * operand is evaluated through a temporary grid
 * (sym/inc/expressionderivatives.h:1545-1593).  When P is exactly u the
 * stencil reads the grid directly (expressionderivatives.h:1658-1682). */
typedef struct {
  int32_t op;        /* sps_op */
  int32_t axis;      /* sps_axis, GRADLAP only: get_stride<axis> convention */
  int32_t component; /* GRADLAP with the 2D 6-point stencil only: 0 or 1 */
  int32_t operand;   /* 0: the field psi; k in 1..nprov: var(k) */
  double coeff;
  double poly[SPS_MAX_DEGREE + 1];
} sps_fd_term;

/* Problem descriptor: SolverFT<Stencil{2,3}d2h<L,G,B>> on one periodic field. */
typedef struct {
  int32_t dim;       /* 2 or 3 */
  int32_t dtype;     /* sps_dtype: storage and arithmetic of the field */
  int32_t n[3];      /* interior counts per axis (n[2] ignored in 2D) */
  int32_t lap_points, gradlap_points, bilap_points; /* <L, G, B> */
  int32_t nterms;
  int32_t path;      /* sps_path */
  double h;          /* grid spacing, the same on every axis */
  double dt;
  sps_fd_term terms[SPS_MAX_TERMS];
  /* Provisional variables (PROVISIONAL_DEF; MODELBB at
   * examples/models/modeldefinitions.h:150-155 and tests/testdefs.h:29-37).
   * var(k) = sum of prov[k-1][0 .. prov_nterms[k-1]), whose operands are psi
   * or var(j), j < k.  As in the reference, each variable is evaluated over
   * the interior into a grid of its own after the field's ghosts are
   * refreshed, and that grid's ghosts are refreshed periodically before the
   * right-hand side reads it (sol/inc/modelspecialized.h:430-434,
   * sol/inc/provisionalsystem.h:72-99, solverft.h:163-183). */
  int32_t nprov;                      /* 0 .. SPS_MAX_PROV */
  int32_t prov_nterms[SPS_MAX_PROV];  /* 1 .. SPS_MAX_TERMS each */
  int32_t prov_reserved;
  sps_fd_term prov[SPS_MAX_PROV][SPS_MAX_TERMS];
} sps_fd_desc;

typedef struct sps_fd_plan sps_fd_plan;
typedef struct sps_fd_batch sps_fd_batch;

/* ---------------------------------------------------------------- library */

/* Text of the last error on this thread ("" if none). */
const char* sps_fd_last_error(void);
/* ABI version and sizeof(sps_fd_desc) the library was built with. */
int sps_fd_abi_version(void);
size_t sps_fd_desc_size(void);
/* Number of HIP devices visible (0 when none); never fails. */
int sps_fd_device_count(void);

/* ------------------------------------------------------------ single field */

/* Validate the descriptor, lower its terms, allocate device storage for the
 * field on the current device.
 * Replaces: SolverFT::make_solver (solverft.h:121-161), the SolverSystemFD
 * constructor (sol/inc/solversystem.h:50-60, sol/inc/boundarysystem.h:
 * 155-214) and SolverFT::form_expr_one (solverft.h:112-119). */
int sps_fd_create(const sps_fd_desc* desc, sps_fd_plan** out);
int sps_fd_destroy(sps_fd_plan* plan);

/* Copy a host field to the device.  PADDED ignores the ghost cells.
 * Replaces: populating the system's interior from initial conditions
 * (sol/inc/boundarysystem.h:498-508, sol/inc/systemlib.h:1949-1959). */
int sps_fd_upload(sps_fd_plan* plan, const void* host, int layout);

/* Copy the field to the host.  PADDED fills the ghost cells with their
 * periodic values, i.e. the state the reference holds after update().
 * Replaces: reading system.values / persist (sol/inc/solversystem.h:835-838). */
int sps_fd_download(sps_fd_plan* plan, void* host, int layout);

/* Advance nsteps forward-Euler steps.
 * Replaces: model_iteration's equation → step → update sequence
 * (modules-models.h:436-467) driven by run_model (modules-models.h:481-492). */
int sps_fd_step(sps_fd_plan* plan, int nsteps);

/* Evaluate the right-hand side (dframe over the interior) of the current
 * field without advancing it, and copy it to host in the given layout
 * (PADDED leaves the ghosts of dframe at zero, as the reference does).
 * Replaces: SolverFT::equation (solverft.h:90-103). */
int sps_fd_equation(sps_fd_plan* plan, void* host_dframe, int layout);

/* Apply one derivative (op, axis, component as in sps_fd_term) to the current
 * field and copy the interior result to host.
 * Replaces: Stencil::applied_laplacian / applied_bilaplacian /
 * applied_gradlaplacian<ax> (sol/inc/stencil.h:130-159). */
int sps_fd_apply(sps_fd_plan* plan, int op, int axis, int component,
                 void* host_out, int layout);

/* Free-energy functional of the reductions: the Landau density of landau_fe
 * (sol/inc/modelspecialized.h:31-45),
 *     f(psi) = -c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2,
 * whose gradient coefficient is 1/2 in the reference (kappa = 1); kappa lets a
 * model with mu = -c1 psi + c2 psi^3 - kappa lap psi report the energy its
 * dynamics decrease.  grad psi is taken per axis as
 *   SPS_GRAD_FORWARD: (psi[i+1] - psi[i]) / h.  With it the gradient sum obeys
 *                     discrete summation by parts on the periodic grid,
 *                     sum |grad psi|^2 = -sum psi * lap psi for the 5-point
 *                     (2D) / 7-point (3D) laplacian, so Model B with those
 *                     stencils decreases exactly this energy;
 *   SPS_GRAD_CENTRAL: (psi[i+1] - psi[i-1]) / (2h), the reference's generated
 *                     first-derivative stencil at its default second-order
 *                     accuracy (make_operator_derivative<1>,
 *                     sol/inc/stencil.h:610-660, DEFAULT_STENCIL_ACCURACY 2
 *                     at :45).
 * mean = 0 reports the integral h^D * sum f; mean = 1 reports the mean of f
 * over the sites, which is what the reference's Energy data process prints
 * (energy * (1 / len), dataprocesses/dataenergy.h:11-40). */
#define SPS_GRAD_FORWARD 0
#define SPS_GRAD_CENTRAL 1
typedef struct {
  double c1, c2, kappa;
  int32_t gradient; /* SPS_GRAD_FORWARD or SPS_GRAD_CENTRAL */
  int32_t mean;     /* 0: h^D * sum; 1: sum / sites */
} sps_fd_energy;

/* FP64 global sums of the current field: mass = sum psi over the interior,
 * energy = the functional set by sps_fd_set_energy (by default c1 = c2 =
 * kappa = 1, forward differences, the integral h^D * sum f).  Either pointer
 * may be NULL.
 * Replaces: DensityData (dataprocesses/datadensity.h:11-41) and the Energy
 * data process over landau_fe (dataprocesses/dataenergy.h:11-40,
 * sol/inc/modelspecialized.h:31-45). */
int sps_fd_reduce(sps_fd_plan* plan, double* mass, double* energy);
/* Set the functional sps_fd_reduce evaluates (NULL restores the default).
 * Coefficients must be finite. */
int sps_fd_set_energy(sps_fd_plan* plan, const sps_fd_energy* energy);

/* Allocation-free variant of sps_fd_reduce for a caller-owned device field.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It enqueues the same two kernels as
 * sps_fd_reduce, in the same summation order, so its sums are bit-identical
 * to sps_fd_reduce's for the same field and functional; it calls no
 * hipMalloc, hipFree or other allocating routine.
 *   dim, dtype, n, h: the field's shape, element type and grid spacing;
 *   energy:     the functional (NULL: c1 = c2 = kappa = 1, forward
 *               differences, the integral h^D * sum f);
 *   field:      n[0]*n[1](*n[2]) values of dtype on the current device,
 *               interior layout, x fastest;
 *   sums:       2 doubles of device memory on the current device, which
 *               receive {mass, energy};
 *   hip_stream: the stream the kernels are enqueued on (NULL: the null
 *               stream).  The call returns without waiting for them.
 * Limits: 1 <= n[a] <= 2^20 and at most SPS_REDUCE_DEVICE_MAX_SITES sites;
 * anything beyond is rejected with SPS_E_INVALID before any HIP call.
 * Scratch: the pass-1 partial sums live in a fixed 16 KiB array of the
 * library's device code (2 x 1024 doubles; pass 1 never uses more than 1024
 * workgroups, whatever the field size), one copy per GPU, held until the
 * process ends and shared by every call on that GPU.  Calls that overlap on
 * one GPU (on different streams or host threads) race on it.  The first call
 * on a device looks up that device's copy (hipGetSymbolAddress); make that
 * first call outside stream capture. */
#define SPS_REDUCE_DEVICE_MAX_SITES 2147483647
int sps_fd_reduce_device(int dim, int dtype, const int32_t* n, double h,
                         const sps_fd_energy* energy, const void* field,
                         double* sums, void* hip_stream);

/* Wait for all work enqueued on the plan. */
int sps_fd_sync(sps_fd_plan* plan);

/* Enqueue the plan's later work on a caller-owned hipStream_t; NULL is the
 * HIP null stream, as in every HIP call (step sequences are then launched
 * one by one instead of through a captured graph, since the null stream
 * cannot be captured).  The caller keeps the stream alive while the plan uses
 * it.  No counterpart in the reference, whose CUDA path has no streams
 * (sym/inc/expressioncuda.cuh:1989-2031 synchronises the device after every
 * evaluation); it lets a host share one stream between the solver and its own
 * work, events and captures. */
int sps_fd_set_stream(sps_fd_plan* plan, void* hip_stream);

/* Device pointer of the current field (interior layout, dtype of the plan),
 * valid until the next sps_fd_step on the plan. */
int sps_fd_device_field(sps_fd_plan* plan, void** dptr);

/* Copy the current field to (direction 0) or from (direction 1) caller-owned
 * device memory of n sites in the interior layout, on the plan's stream.
 * No counterpart in the reference (its fields live in host memory); it lets
 * device-resident consumers such as the surrogate's training read solver
 * states without a host round trip. */
int sps_fd_copy_device(sps_fd_plan* plan, void* dev, int direction);

/* ---------------------------------------------------------- regional grid
 * SolverFT on a regional grid: the reference's SolverSystemFDwSD
 * (sol/inc/solversystem.h:62-77), the system SolverFT associates second
 * (examples/solvers/solverft.h:220), whose field lives on a RegionalGrid
 * (datatypes/inc/grid.h:2038-2157).  The field is stored on a box of the
 * periodic domain with a boundary layer of SPS_BOUNDARY_DEPTH cells per
 * side; every site outside the box reads as `empty`.  A step evaluates the
 * right-hand side and forward Euler over the box interior only, so its cost
 * follows the box, not the domain.  desc->path selects the arithmetic as for
 * sps_fd_create: SPS_PATH_REFERENCE runs the reference order (FP64 results
 * equal the CPU SolverFT's on the box); the canonical form (SPS_PATH_AUTO
 * or SPS_PATH_FUSED) runs the fused form of the streamed kernels (a box that
 * spans the domain equals sps_fd_step at one step per launch bit for bit).
 * A 2D box whose layer is frozen along both axes runs up to 4 such steps per
 * launch between updates, with the same results as one step per launch.
 *
 * Update before each right-hand side (SolverSystemFDwSD::update,
 * update_regional_system, sol/inc/boundarysystem.h:437-479):
 *   - the first update fixes cutoff = min(psi) + cutoff_eps and
 *     empty = min(psi), then places a box of resize_factor times the active
 *     box's dims (at most the domain), centred on it
 *     (resize_adjust_region, datatypes/inc/gridfunctions.h:2194-2238);
 *   - later, once time / resize_time reaches the next resize count, the box
 *     is re-centred on the active box with its dims kept (fixed_resize = 1,
 *     adjust_region, :2166-2176), or resized as above unless every axis
 *     changes by at most dims_relative_eps relatively (fixed_resize = 0);
 *   - along an axis the box spans, its layer is the periodic image of its
 *     interior (regional_update_boundary,
 *     datatypes/inc/boundaryupdatedata.h:1398-1433); along other axes the
 *     layer keeps the values it was given when the box last moved.
 * The active box along an axis is the shortest periodic interval covering
 * every coordinate of an interior site whose value exceeds the cutoff.
 * When the box moves, a site of the new box takes the old interior's value
 * at the same domain site, or empty.  Box interiors have at least 3 sites
 * per axis (or the whole axis).
 * Departures from the reference: its periodic search grows a box from the
 * first active site until an empty slab (get_view_resized_periodic,
 * gridfunctions.h:2011-2163), so a second cluster beyond an empty slab falls
 * outside the box and is set to empty; the covering interval here keeps
 * every active site.  Provisional variables are not supported. */
typedef struct {
  double cutoff_eps;        /* params::regional_resize_cutoff_eps (1e-6) */
  double resize_factor;     /* params::regional_resize_factor (2.0), >= 1 */
  double resize_time;       /* params::regional_resize_time (3.0), > 0 */
  int32_t fixed_resize;     /* params::regional_resize_is_fixed (1) */
  int32_t steps_per_launch; /* fused form, 2D box with a frozen layer: 0 or
                               4 = four steps per launch between updates,
                               1..3 = one (no counterpart in the reference;
                               the results are the same) */
  double dims_relative_eps; /* REGIONAL_GRID_RELATIVE_DIMS_EPS (0.05) */
} sps_fd_regional;         /* defaults: sol/src/boundarysystem.cpp:24-27,
                              sol/inc/boundarysystem.h:137 */

typedef struct {
  int32_t origin[3]; /* domain coordinate of the box's first interior site */
  int32_t dims[3];   /* box dims with the layer; all 0 when no site is active */
  double empty, cutoff, time;
  int32_t index, reserved;
} sps_fdr_region_info;

typedef struct sps_fdr_plan sps_fdr_plan;

/* params NULL: the reference's defaults.  Replaces constructing a
 * SolverSystemFDwSD (solversystem.h:67-71). */
int sps_fdr_create(const sps_fd_desc* desc, const sps_fd_regional* params,
                   sps_fdr_plan** out);
int sps_fdr_destroy(sps_fdr_plan* plan);
/* The whole domain's field; the box becomes the whole domain and the update
 * state (cutoff, empty, resize count, time) restarts. */
int sps_fdr_upload(sps_fdr_plan* plan, const void* host, int layout);
/* nsteps x (update, right-hand side, Euler) (modules-models.h:436-467). */
int sps_fdr_step(sps_fdr_plan* plan, int nsteps);
/* The whole domain's field: box interior values inside, empty outside. */
int sps_fdr_download(sps_fdr_plan* plan, void* host, int layout);
int sps_fdr_region(const sps_fdr_plan* plan, sps_fdr_region_info* out);
/* Wait for all work enqueued on the regional plan. */
int sps_fdr_sync(sps_fdr_plan* plan);

/* ------------------------------------------------------------- checkpoints */

/* The reference's GNU checkpoint file of the current field, formatted on the
 * device: the header line of print_gp_header (io/src/writegp.cpp:427-465),
 * "dim n0 n1 [n2] left0 right0 ... index @", with domain bounds
 * left = origin[a] (0 when origin is NULL) and right = left + (n[a]-1)*h,
 * then every value as fprintf("%.10E ") writes it, x fastest, with a newline
 * after each x row (save_grid, io/src/writegp.cpp:780-803; called from
 * checkpoint_system, modules-io.h:152-230).  The text is byte-identical to the
 * C library's.  With text NULL only *length is computed; otherwise text must
 * hold capacity >= *length bytes (a terminating NUL is added when it fits).
 * FP32 fields are written as the doubles of their values. */
int sps_fd_checkpoint_text(sps_fd_plan* plan, int index, const double* origin,
                           char* text, size_t capacity, size_t* length);

/* Read a GNU checkpoint file into the plan's field, as read_header and
 * read_block (io/src/readgp.cpp) do with fscanf("%lf"): the header's
 * dimension and counts must match the plan (a subdomain header is
 * SPS_E_INVALID); values beyond the field's sites are ignored.  *index (if
 * not NULL) receives the header's solution index.  Values are parsed on the
 * device and rounded as strtod rounds; FP32 fields round the double. */
int sps_fd_checkpoint_read(sps_fd_plan* plan, const char* text, size_t length,
                           int* index);

/* The reference's plot data file of the current field, formatted on the
 * device: SPS_PLOT_GNU is the GNU matrix of gp::save_grid_plotting
 * (io/src/writegp.cpp:470-507: per z slice a line of "% 13d " index and
 * "% 13.2f " x positions, then rows of "% 13.2f " y position and "% 10lE "
 * values, slices separated by a blank line, an empty last line);
 * SPS_PLOT_CSV is csv::save_grid_plotting (io/src/writecsv.cpp: "% 10lE"
 * values joined by ',' per x row).  Positions are origin + i*h (origin NULL:
 * zeros).  Byte-identical to the C library's output.  text NULL: only the
 * length. */
#define SPS_PLOT_GNU 1
#define SPS_PLOT_CSV 2
int sps_fd_plot_text(sps_fd_plan* plan, int format, const double* origin,
                     char* text, size_t capacity, size_t* length);

/* The reference's XDR file of the current field (io/src/writexdr.cpp
 * print_xdr_header :45-87 + save_xdr :92-98, the XDR writer's checkpoint and
 * data file): big-endian int dim, counts, domain and interval bounds, the
 * index, then every value as a big-endian IEEE double in x-fastest order,
 * converted on the device.  data NULL: only the length. */
int sps_fd_checkpoint_xdr(sps_fd_plan* plan, int index, const double* origin,
                          void* data, size_t capacity, size_t* length);
/* Read such a file into the plan's field (io/src/readxdr.cpp read_header,
 * read_block); dimension and counts must match the plan.  FP32 plans round
 * each double to nearest. */
int sps_fd_checkpoint_read_xdr(sps_fd_plan* plan, const void* data,
                               size_t length, int* index);

/* ------------------------------------------------ checkpoint directory */

/* Checkpoint the field as field `id` of a run whose output directory is dir,
 * the way the reference's IO driver does at a checkpoint index
 * (save_if_checkpoint → checkpoint_system → save_grid and record_index,
 * modules-io.h:306-327, modules-io.cpp:29-60):
 *  - single_file = 1 (the reference's default, params::single_output_file):
 *    "<dir>/checkpoint/data<id>" holds every checkpoint of the field; the
 *    first one this process writes truncates the file and carries the full
 *    header, later ones are appended under the header "<index> @"
 *    (io/src/writedefines.cpp:160-185, io/src/writegp.cpp:427-465);
 *  - single_file = 0: "<dir>/checkpoint/data<id>_<index>", full header
 *    (OUTPUT_CHECKPOINT_FMT / _INDEX_FMT, io/inc/io.h:185-188);
 *  - "#$<index>" is appended to "<dir>/checkpoint/progress.txt", followed by
 *    "#$DONE" when done is nonzero (the run reached its stop index).
 * Directories are created as needed.  The text is sps_fd_checkpoint_text's. */
int sps_fd_checkpoint_save(sps_fd_plan* plan, const char* dir, int id,
                           int index, int single_file, int done,
                           const double* origin);
/* Resume: read field `id`'s checkpoint for `index` into the plan.  With
 * single_file the blocks of "<dir>/checkpoint/data<id>" are read in order and
 * the last one read is kept, continuing while the next block's index does
 * not pass `index` (read_grid_standardized, io/inc/readdefines.h:340-420);
 * otherwise "<dir>/checkpoint/data<id>_<index>" is read.  *index_read (if not
 * NULL) receives the index of the block loaded. */
int sps_fd_checkpoint_load(sps_fd_plan* plan, const char* dir, int id,
                           int index, int single_file, int* index_read);
/* The progress record "<dir>/checkpoint/progress.txt": the last recorded
 * index (-1 when none), how many indices it lists and whether the run was
 * recorded as done; a missing record is no error (count 0).  The reference
 * collects the same list when it loads a backup configuration
 * (get_completed_indices, conf/src/conf.cpp:117-155).  Host only. */
int sps_fd_checkpoint_progress(const char* dir, int* last_index, int* count,
                               int* done);

/* Path the plan runs (sps_path, never AUTO) and the steps fused per launch. */
int sps_fd_query(const sps_fd_plan* plan, int* path, int* steps_per_launch);

/* Override the fused path's launch shape (0 keeps the current value).
 * 2D: steps_per_launch in 1..4, rows_per_tile = rows per streamed chunk,
 *     4..512; planes_per_chunk must be 0; values_per_lane =
 *     consecutive x values per lane, 16 or 8 bytes' worth (FP32: 4 or 2,
 *     FP64: 2 or 1), dividing n[0].
 * 3D: (steps_per_launch, rows_per_tile) = steps fused per launch and y rows
 *     per wave, one of (1, 4), (1, 6), (1, 8), (1, 12), (1, 16), (2, 4) and,
 *     FP32 only, (2, 6), (2, 8), (2, 12);
 *     planes_per_chunk = z planes per streamed chunk, 1..1024;
 *     values_per_lane must be 0 or 1. */
int sps_fd_tune(sps_fd_plan* plan, int steps_per_launch, int rows_per_tile,
                int planes_per_chunk, int values_per_lane);

/* Pick the fused path's launch shape by timing candidates on the device:
 * 2D steps_per_launch {1,2,3,4} x (16-byte lanes, rows_per_tile
 * {8,16,32,64} | 8-byte lanes, rows {16,32,64,128}), each lane width and
 * steps_per_launch also with the chunk heights that give every CU of the
 * device the same number (2..6) of 4-wave workgroups; 3D (steps, rows) as
 * sps_fd_tune allows x planes {8, 16, 32} and n[2] split into 1..8 equal
 * chunks.  The field is saved before and restored after the trial launches,
 * so the plan's state is unchanged.  When the environment variable
 * SPS_FD_TUNE_CACHE names a file, a choice recorded there for the same (dim,
 * shape, dtype, lowered form, GPU) is used without timing (unless use_cache
 * is 0) and new choices are appended.  The chosen shape is applied to the
 * plan and returned through the optional pointers (values_per_lane is 1 in
 * 3D).  Plans on the reference path have nothing to tune (SPS_OK).
 * No counterpart in the reference (its CPU loops have no launch shape). */
int sps_fd_autotune(sps_fd_plan* plan, int use_cache, int* steps_per_launch,
                    int* rows_per_tile, int* planes_per_chunk,
                    int* values_per_lane);

/* ---------------------------------------------------------------- batched */

/* Many independent 2D fields of uneven sizes and coefficients, advanced by one
 * persistent launch per step group (resident waves walk the strip x row-chunk
 * items of every instance), captured once in a hipGraph and replayed.
 * Every descriptor must be 2D FP32, lower to the fused form (polynomials of
 * degree <= 3) and have an even n[0]; extents may differ per instance.
 * Replaces: find_solution over an array of models, which steps each model in
 * turn (modules-io.h:368-474, loop at :462-465). */
int sps_fd_batch_create(const sps_fd_desc* descs, int count,
                        sps_fd_batch** out);
int sps_fd_batch_destroy(sps_fd_batch* batch);
int sps_fd_batch_upload(sps_fd_batch* batch, int index, const void* host,
                        int layout);
int sps_fd_batch_download(sps_fd_batch* batch, int index, void* host,
                          int layout);
int sps_fd_batch_step(sps_fd_batch* batch, int nsteps);
/* mass and energy are arrays of count doubles (either may be NULL); each
 * instance's energy is the functional set for it by sps_fd_batch_set_energy
 * (default as sps_fd_reduce). */
int sps_fd_batch_reduce(sps_fd_batch* batch, double* mass, double* energy);
/* Per-instance free-energy functionals: energy points to count entries (NULL
 * restores the default for every instance). */
int sps_fd_batch_set_energy(sps_fd_batch* batch, const sps_fd_energy* energy);
int sps_fd_batch_sync(sps_fd_batch* batch);
/* Launch shape of the batched kernel (0 keeps the current value):
 * steps_per_launch in 1..4, rows_per_item = rows of one work item, 4..512. */
int sps_fd_batch_tune(sps_fd_batch* batch, int steps_per_launch,
                      int rows_per_item);
/* Copy instance `index` to (direction 0) or from (direction 1) caller-owned
 * device memory, interior layout, on the batch's stream (no reference
 * counterpart, as sps_fd_copy_device). */
int sps_fd_batch_copy_device(sps_fd_batch* batch, int index, void* dev,
                             int direction);
/* Time steps_per_launch {1..4} x rows_per_item {16,32,64,128, and for m =
 * 2..6 the lowest height whose item count fits m items per SIMD of the
 * device} on the device (fields saved and restored) and apply the fastest;
 * returns the choice through the optional pointers. */
int sps_fd_batch_autotune(sps_fd_batch* batch, int* steps_per_launch,
                          int* rows_per_item);
/* Caller-owned hipStream_t for later work (NULL: the HIP null stream, as
 * sps_fd_set_stream). */
int sps_fd_batch_set_stream(sps_fd_batch* batch, void* hip_stream);

/* ------------------------------------------------------------ coupled fields */

/* Models of several coupled fields (MODEL(MC, (SCALARS(2)), ...), GRAYSCOTT,
 * Turing: examples/models/modeldefinitions.h:63-81,123-137): each field's
 * EVOLUTION is a sum of terms coeff * D[ prod_f psi_f^pow[f] ], a derivative
 * of a monomial of the fields, which is what apply_operators leaves of those
 * right-hand sides (products distributed, a derivative of a non-trivial
 * operand through a temporary grid, expressionderivatives.h:1583-1588).  Every
 * field is stepped with the same SolverFT semantics as a single field: all
 * right-hand sides from the current fields, forward Euler on each, periodic
 * refresh of each. */
#define SPS_MAX_FIELDS 4
#define SPS_MAX_MTERMS 12

typedef struct {
  int32_t op;        /* sps_op */
  int32_t axis;      /* GRADLAP only */
  int32_t component; /* GRADLAP with the 2D 6-point stencil only */
  int32_t reserved;
  double coeff;
  int32_t pow[SPS_MAX_FIELDS]; /* exponent of each field, 0..6, sum <= 6 */
} sps_fdm_term;

typedef struct {
  int32_t dim, dtype;
  int32_t n[3];
  int32_t lap_points, gradlap_points, bilap_points;
  int32_t nfields; /* 1 .. SPS_MAX_FIELDS */
  int32_t path;    /* sps_path */
  double h, dt;
  int32_t nterms[SPS_MAX_FIELDS];
  sps_fdm_term terms[SPS_MAX_FIELDS][SPS_MAX_MTERMS];
} sps_fdm_desc;

typedef struct sps_fdm_plan sps_fdm_plan;

/* Create / destroy a coupled-field plan.  FUSED (2D, two fields): each
 * field's right-hand side lowered to A_f(psi) + lap(B_f(psi)) + kc_f*bilap
 * (psi_f) with A_f, B_f polynomials of degree <= 3 in each field, 5-point
 * laplacian, 13-point bilaplacian and no gradlaplacian; anything else runs
 * the reference-order kernel (FP64: the oracle's bits). */
int sps_fdm_create(const sps_fdm_desc* desc, sps_fdm_plan** out);
int sps_fdm_destroy(sps_fdm_plan* plan);
/* Field f (0-based) in the interior or padded layout, as sps_fd_upload. */
int sps_fdm_upload(sps_fdm_plan* plan, int field, const void* host,
                   int layout);
int sps_fdm_download(sps_fdm_plan* plan, int field, void* host, int layout);
int sps_fdm_step(sps_fdm_plan* plan, int nsteps);
int sps_fdm_query(const sps_fdm_plan* plan, int* path, int* steps_per_launch);
/* Launch shape of the fused coupled kernel (0 keeps the current value):
 * steps_per_launch 1..4, rows_per_tile 4..512 (rows per streamed chunk),
 * values_per_lane 16/size or 8/size (16- or 8-byte lanes; must divide the x
 * extent). */
int sps_fdm_tune(sps_fdm_plan* plan, int steps_per_launch, int rows_per_tile,
                 int values_per_lane);
/* Time steps_per_launch {1..4} x rows_per_tile {16,32,64, and the chunk
 * heights that give every CU the same number (2..6) of 4-wave workgroups} x
 * both lane widths in 2D, (steps, rows) as sps_fdm_tune allows in 3D, on the
 * device (fields saved and restored) and keep the fastest (no reference
 * counterpart, as sps_fd_autotune). */
int sps_fdm_autotune(sps_fdm_plan* plan, int* steps_per_launch,
                     int* rows_per_tile, int* values_per_lane);
int sps_fdm_sync(sps_fdm_plan* plan);
/* The data files of field `field` (0-based; the reference writes one file per
 * field of a multi-field model, its id being the field's index:
 * OUTPUT_CHECKPOINT_FMT, io/inc/io.h:185-188), with the formats and behaviour
 * of sps_fd_checkpoint_text / _read, sps_fd_plot_text and
 * sps_fd_checkpoint_xdr / _read_xdr. */
int sps_fdm_checkpoint_text(sps_fdm_plan* plan, int field, int index,
                            const double* origin, char* text, size_t capacity,
                            size_t* length);
int sps_fdm_checkpoint_read(sps_fdm_plan* plan, int field, const char* text,
                            size_t length, int* index);
int sps_fdm_plot_text(sps_fdm_plan* plan, int field, int format,
                      const double* origin, char* text, size_t capacity,
                      size_t* length);
int sps_fdm_checkpoint_xdr(sps_fdm_plan* plan, int field, int index,
                           const double* origin, void* data, size_t capacity,
                           size_t* length);
int sps_fdm_checkpoint_read_xdr(sps_fdm_plan* plan, int field,
                                const void* data, size_t length, int* index);
/* Every field of a coupled plan checkpointed / resumed as
 * sps_fd_checkpoint_save / _load do for one, field f being id f; one progress
 * line per call.  Loading requires every field's block to carry the same
 * index. */
int sps_fdm_checkpoint_save(sps_fdm_plan* plan, const char* dir, int index,
                            int single_file, int done, const double* origin);
int sps_fdm_checkpoint_load(sps_fdm_plan* plan, const char* dir, int index,
                            int single_file, int* index_read);
/* Caller-owned hipStream_t for later work (NULL: the HIP null stream). */
int sps_fdm_set_stream(sps_fdm_plan* plan, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* SPS_FD_H */
