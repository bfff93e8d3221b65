/*
 * pdipc.h -- C ABI of the MI355X PD-IPC step library (libpdipc.so).
 *
 * Drop-in for the reference simulator boundary of lanlei/PD-IPC-ArmadilloDemo:
 * the entry points below replace, one for one, the calls a host makes on
 * PD_IPC::PdIpcSimulator (Simulator/PD-IPC/PdIpcSimulator.h) and its
 * configuration path (Simulator/BaseSimulator.cpp:4-65).  Plain pointers and
 * sizes only; positions and velocities cross the boundary as AoS xyz doubles
 * (the reference's devPoints layout, PdIpcSimulator.cpp:456-457) and are kept
 * SoA on the device.
 *
 * Every function returns PDIPC_OK (0) or a negative status; the message of the
 * last failure on the calling thread is available from pdipc_last_error().
 * No function aborts the process.  A handle is not thread-safe; distinct
 * handles may be used from distinct threads.
 */
#ifndef PDIPC_H
#define PDIPC_H

#ifdef __cplusplus
extern "C" {
#endif

#define PDIPC_OK 0
#define PDIPC_E_ARG -1      /* invalid argument / call order */
#define PDIPC_E_HIP -2      /* HIP runtime error (message has the HIP string) */
#define PDIPC_E_CAPACITY -3 /* a fixed-capacity buffer could not be grown */
#define PDIPC_E_BADMESH -4  /* degenerate or inconsistent mesh input */
#define PDIPC_E_NAN -5      /* the step produced a non-finite state */
#define PDIPC_E_NODEVICE -6 /* no HIP device available */

#define PDIPC_FP64 0       /* every kernel in double precision */
#define PDIPC_FP32_SPLIT 1 /* FP32 local step and Chebyshev sweeps on an FP64 residual; FP64 CCD / barrier */

typedef struct pdipc_handle pdipc_handle;

/* Step parameters.  Every constant the reference hard-codes is a field whose
 * default (pdipc_default_params) is the reference value. */
typedef struct pdipc_params {
  double dt;           /* BaseSimulator _timeStep (config <timeStep>) */
  double gravity;      /* along -y: force (0, gravity*m, 0), PdIpcSimulator.cpp:249-252 */
  double dhat;         /* <dhat>, PdIpcSimulator.h:618-623 (ctor default 1e-3) */
  double kappa;        /* <kappa>, PdIpcSimulator.h:624-628 (ctor default 1e4) */
  int lg_iters;        /* local/global passes per outer iteration, PdIpcSimulator.cpp:1592 (10) */
  int cheb_iters;      /* Chebyshev sweeps per solve, PdIpcSimulator.cpp:1597 (15) */
  int max_outer;       /* outer (CCD) iterations, PdIpcSimulator.cpp:1685 (30) */
  double residual_tol; /* |d|^2 break threshold, PdIpcSimulator.cpp:1776 (1e-5) */
  double omega;        /* Jacobi relaxation, PdIpcGpuFunc.cu:949,1051 (2/3) */
  double rho;          /* Chebyshev spectral estimate, PdIpcGpuFunc.cu:1318 (0.9992) */
  double sigma_lo;     /* strain clamp, PdIpcGpuFunc.cu:307-316 (0.999) */
  double sigma_hi;     /* (1.001) */
  double ccd_filter;   /* CCD swept-box filter, PdIpcGpuFunc.cu:2656,2676 (5e-3) */
  double toi_scale;    /* TOI safety factor, PdIpcGpuFunc.cu:2698-2699 (0.98) */
  int ccd_iters;       /* ACCD iterations, PdIpcGpuFunc.cu:2497 (10) */
  double ccd_step;     /* ACCD step factor (0.9) */
  double ccd_gap;      /* ACCD stop fraction (0.1) */
  double static_mass;  /* collision mass of static points, PdIpcSimulator.cpp:1494 (1e4) */
  double inflation;    /* broad-phase box inflation; < 0 means 0.5*dhat (PdIpcSimulator.cpp:1154) */
  int skip_static_pairs; /* 1: drop patch pairs whose patches are both static (they provably
                            contribute nothing to the step); 0: the reference's full pair set */
  int touch_policy;    /* 1: a pair already touching gives TOI 0 (conservative);
                          0: the reference rule (NaN, pair ignored, PdIpcGpuFunc.cu:2499) */
  int precision;       /* PDIPC_FP64 or PDIPC_FP32_SPLIT */
  int sell_lanes;      /* lanes per row in the Chebyshev sweep (1, 2, 4, 8); 0 = default */
  int use_graphs;      /* 1: replay the local/global loop from a HIP graph */
  int sell_stage;      /* 1: the sweeps stage their gathered operand in LDS, per workgroup, when
                          every workgroup's column union fits (default; the FP32 and FP64
                          sweeps of a lane-major layout use narrow tiles with one-byte indices
                          when they fit); 2: the same with 16-bit tiles only; 0: gather from L2 */
  int capacity_hint;   /* 0: default initial capacities of the growable broad-phase and contact
                          buffers (sized from the patch count); > 0: start each at this many
                          entries (they grow on demand either way) */
  int track_toi_argmin; /* 1: also report the choices a derivative of the step holds fixed, at
                           every outer iteration: the record attaining the minimum TOI and its
                           ACCD exit path (pdipc_counters.toi_argmins) and the set of active
                           barrier records (pdipc_counters.active_sets); diagnostics, two more
                           passes and one more host wait per outer iteration */
} pdipc_params;

typedef struct pdipc_counters {
  int outer_iters;      /* outer iterations of the last step */
  int lg_iters;         /* local/global passes of the last step */
  int patch_pairs;      /* overlapping patch pairs of the last broad phase */
  int candidates;       /* candidate primitive records (K) of the last broad phase */
  int active_pairs;     /* records within dHat at the last constraint assembly */
  double toi_min;       /* smallest TOI applied in the last step */
  double residual;      /* |d|^2 of the last outer iteration */
  long long nnz_r1;     /* off-diagonal nonzeros of R1 (plus one diagonal slot per row) */
  long long nnz_r2;     /* nonzeros of the M^2 = R2 row operator */
  long long n_2paths;   /* 2-path terms feeding the R2 operator update */
  long long sell_entries; /* stored M^2 entries including SELL padding */
  int sim_points, total_points, cloth_points, tets, patches, boundary_faces, boundary_edges;
  long long toi_argmins;  /* with track_toi_argmin: a hash of the record attaining the minimum TOI
                             at each outer iteration of the last step and of the iteration and
                             test at which its ACCD loop ended (0 where none is below 1) */
  long long active_sets;  /* with track_toi_argmin: a hash of the active barrier records of each
                             outer iteration of the last step */
} pdipc_counters;

/* --- lifecycle: PdIpcSimulator ctor + readExtraAttributeFromConfigFile (PdIpcSimulator.h:441-450,602-642) */
void pdipc_default_params(pdipc_params* out);
int pdipc_create(const pdipc_params* params, pdipc_handle** out);
int pdipc_destroy(pdipc_handle* h);
const char* pdipc_last_error(void);

/* --- models: PdIpcSimulator::addModelFromConfigFile / addStaticModelFromConfigFile
 * (PdIpcSimulator.h:462-600) with the mesh already loaded and transformed.
 * verts: 3*nverts doubles (AoS); tris: 3*ntris; tets: 4*ntets; fixed: vertex ids
 * (fixed_points.txt, PdIpcModel.h:56-77).  s_strain, s_bend_or_volume, s_pos are
 * the <type> stiffness triple; patch_dim is <patchDim>. */
int pdipc_add_cloth(pdipc_handle* h, const double* verts, int nverts, const int* tris, int ntris, double s_strain,
                    double s_bend, double s_pos, double density, int patch_dim, const int* fixed, int nfixed);
int pdipc_add_tet(pdipc_handle* h, const double* verts, int nverts, const int* tets, int ntets, double s_strain,
                  double s_volume, double s_pos, double density, int patch_dim, const int* fixed, int nfixed);
int pdipc_add_static(pdipc_handle* h, const double* verts, int nverts, const int* tris, int ntris, int patch_dim);

/* --- PdIpcSimulator::initialization (PdIpcSimulator.cpp:34-268): host precompute
 * then device upload.  pdipc_host_setup runs only the host precompute (no device
 * needed); pdipc_init runs it if not done yet and uploads. */
int pdipc_host_setup(pdipc_handle* h);
int pdipc_init(pdipc_handle* h);

/* --- PdIpcSimulator::run -> doTime -> advance (PdIpcSimulator.cpp:18,1538-1552,1683-1796) */
int pdipc_step(pdipc_handle* h, int nsteps);
int pdipc_set_precision(pdipc_handle* h, int precision);

/* --- state: postRun (PdIpcSimulator.cpp:23-32) reads positions back; set_* restart
 * from a given state.  x covers all points (simulated first, static after),
 * v covers simulated points. */
int pdipc_get_x(pdipc_handle* h, double* out, int npoints);
int pdipc_set_x(pdipc_handle* h, const double* in, int npoints);
int pdipc_get_v(pdipc_handle* h, double* out, int npoints);
int pdipc_set_v(pdipc_handle* h, const double* in, int npoints);
int pdipc_get_counters(pdipc_handle* h, pdipc_counters* out);

/* --- saveSimulator (PdIpcSimulator.h:879-895): writes the positions as
 * EigenMatrixIO::write_binary does (int64 rows = 3, int64 cols = npoints, then
 * column-major data) in float (the reference's qeal) or double. */
int pdipc_save_frame(pdipc_handle* h, const char* path, int as_double);

/* --- full-state checkpoint (no counterpart in the reference, which only dumps
 * frames): the step reads nothing from earlier steps but x and v, so a saved
 * state restarts a run bit for bit.  Binary file: "PDIPCST1", int32 version 1,
 * int32 nSim, int32 nTotal, int32 0, uint64 topology fingerprint, then x (3
 * nTotal doubles, AoS) and v (3 nSim doubles, AoS).  Loading checks the magic,
 * the point counts and the fingerprint (the scene's ring, tet and R1
 * patterns) and rejects a short or over-long file with PDIPC_E_ARG. */
int pdipc_save_state(pdipc_handle* h, const char* path);
int pdipc_load_state(pdipc_handle* h, const char* path);

/* --- candidate records of the last broad phase (int[5]: type, a, b, c, d as in
 * getCcdPotentialICPairsListKernel, PdIpcGpuFunc.cu:1915-1945); returns the
 * count in *count and copies min(count, cap) records. */
int pdipc_get_candidates(pdipc_handle* h, int* out, int cap, int* count);
int pdipc_get_patch_pairs(pdipc_handle* h, int* out, int cap, int* count);

/* --- host precompute results for fixture comparison: the int32 image of a
 * simulated model's boundary.dat (which = 0, PdIpcModel.h:219-251) or
 * patchInfo.dat (which = 1, PdIpcModel.h:291-318).  Available after
 * pdipc_host_setup; *count receives the int count. */
int pdipc_get_model_dat(pdipc_handle* h, int model, int which, int* out, int cap, int* count);

/* --- teacher-forced phase entry points (parity tests and profiling).  Arrays
 * are AoS; x-like inputs cover all points. */
/* gpuCcdCulling (PdIpcSimulator.cpp:1799-1869) on (xold, dx); returns K in *count */
int pdipc_debug_broad_phase(pdipc_handle* h, const double* xold, const double* dx, int* count);
/* getImpactOfTimeHost (PdIpcGpuFunc.cu:2706-2742) over the current candidates */
int pdipc_debug_toi(pdipc_handle* h, const double* x0, const double* x1, double* toi);
/* setCollisionConstraintsHost (PdIpcGpuFunc.cu:2894-2930): diag (sim), rhs (3*sim) */
int pdipc_debug_collision(pdipc_handle* h, const double* x, double* diag, double* rhs, int* active);
/* doLocalProjection (PdIpcSimulator.cpp:1624-1681) without the collision term */
int pdipc_debug_local(pdipc_handle* h, const double* x, const double* init_rhs, double* rhs);
/* updateSuperJacobiOperatorsHost + chebyshevJacobiR2SolverHost (PdIpcGpuFunc.cu:976-1346) */
int pdipc_debug_solve(pdipc_handle* h, const double* diag, const double* rhs, const double* x0, double* xout);
/* Validation, no counterpart in the reference: the number of FP64
 * edge-triangle crossings (triangle-triangle intersections) at the current
 * positions, over every edge/face pair whose patches' boxes overlap (and
 * within each patch), pairs sharing a vertex excluded.  Runs a broad phase
 * at the current positions (the candidate lists then describe it). */
int pdipc_count_intersections(pdipc_handle* h, long long* count);
/* LBVH of the last broad phase: sorted 64-bit keys (P), internal nodes
 * (left, right child as node ids: leaves are P-1+leaf) and parents (2P-1) */
int pdipc_debug_lbvh(pdipc_handle* h, unsigned long long* keys, int* left, int* right, int* parent, int* npatches);
/* Validation, no counterpart in the reference (the reference keeps R2 as
 * per-row lists, PdIpcGpuFunc.cu:976-1059): the SELL layout of the M^2 rows
 * that the device upload builds, for `lanes` lanes per row (1, 2, 4 or 8),
 * from the host precompute alone (after pdipc_host_setup; no device needed).
 * sizes[0..5] receive the slice count, the entry count (padding included), the
 * common slice width (0: per-slice widths), 1 if 16-bit columns are used, the
 * slot count of the row order (0: slot = row) and the rows per slice.  The
 * arrays may be null (sizes only); otherwise they hold slice count + 1 offsets,
 * slice count widths (in 64-entry columns), entry count columns and row flags
 * (row | 0x80000000 on the diagonal; padding entries repeat their row's index
 * unflagged), entry count 16-bit columns (col - slot; only if used) and the
 * slot -> row order (only if rows are reordered). */
int pdipc_debug_sell_layout(pdipc_handle* h, int lanes, long long* sizes, long long* slice_off, int* slice_w,
                            int* col, int* row_flag, short* col16, int* perm);
/* Validation, no counterpart in the reference: the LDS staging tables of the
 * sweeps for `lanes` lanes per row (after pdipc_host_setup; no device needed).
 * A tile is sizes[5] consecutive slices (one workgroup).  sizes[0..5] receive
 * 1 if staging applies (every tile's column union spans at most 65536 columns
 * and holds at most 1024), the tile count, the total union length, the
 * largest union, the SELL entry count and the slices per tile.  When staging
 * applies the arrays (may be null) receive tile count + 1 union offsets, the
 * tiles' smallest columns, the union columns minus their tile's smallest, and
 * per SELL entry its index in its tile's union. */
int pdipc_debug_stage_layout(pdipc_handle* h, int lanes, long long* sizes, int* tile_off, int* tile_base,
                             unsigned short* tile_col, unsigned short* loc);
/* The narrow staging tables of the sweeps (8 lanes per row, lane-major
 * layouts only), as pdipc_debug_stage_layout: sizes[0] is 1 when every tile of
 * sizes[5] = 2 slices has a column union of at most 256 columns; loc receives
 * one byte per SELL entry. */
int pdipc_debug_narrow_stage_layout(pdipc_handle* h, long long* sizes, int* tile_off, int* tile_base,
                                    unsigned short* tile_col, unsigned char* loc);
/* How this handle's sweeps (FP32 and FP64, the first of a solve aside) read
 * their operand: 0 gathered from L2,
 * 1 staged in LDS with 16-bit union indices, 2 narrow staging (one-byte
 * indices; taken by default wherever the narrow tiles fit). */
int pdipc_sweep_staging(pdipc_handle* h, int* path);

/* --- no counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same minimum time of impact as pdipc_debug_toi over
 * the current candidate records, on device-resident AoS positions x0, x1 (3 *
 * total points doubles each), written to the device double *toi.  The work is
 * only enqueued on `stream` (NULL: the handle's stream); nothing is allocated
 * and the simulation state is untouched.  Per-workgroup minima go to a
 * library-wide device array of 512 KiB per GPU, held until the process ends,
 * so at most 16,777,216 candidate records (more: PDIPC_E_CAPACITY).  Every
 * call on one GPU uses that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race. */
int pdipc_toi_noalloc(pdipc_handle* h, const double* x0, const double* x1, double* toi, void* stream);

/* --- measurement support (no counterpart in the reference): replays one
 * kernel of the step `reps` times on the handle's stream, bracketed by HIP
 * events, on the state left by the last step (work buffers are clobbered, the
 * simulation state is not).  which = 0: FP64 Chebyshev sweep; 1: FP32-split
 * sweep; 2: cloth local step; 3: CCD over the current candidates; 4: R2
 * operator update.  Returns the mean duration per launch in microseconds and
 * the algorithmic HBM bytes one launch moves (matrix entries once, vectors once). */
int pdipc_time_kernel(pdipc_handle* h, int which, int reps, double* avg_us, double* bytes);

/* --- measurement support (no counterpart in the reference; the reference's
 * timing members are declared but never written, PdIpcSimulator.h:54-64):
 * in-step kernel timing.  After pdipc_kernel_timing(h, 1) the following steps
 * run with direct launches (no graph replay) and every launch of the timed
 * kernels records its own start and stop with HIP events, the duration a
 * profiler's kernel trace reports; pdipc_kernel_timing(h, 0) stops and a new
 * enable starts from zero.  pdipc_get_kernel_times returns, for the first
 * nclasses classes, the summed microseconds and launch counts since then:
 * 0 FP64 sweeps, 1 FP32-split sweeps, 2 local-step rhs gather, 3 cloth
 * curvature and tet projection, 4 first FP64 sweep and FP32-split residual
 * passes, 5 M^2 operator update, 6 time of impact, 7 barrier terms, 8 broad
 * phase (one span per broad phase, its host waits included). */
int pdipc_kernel_timing(pdipc_handle* h, int enable);
int pdipc_get_kernel_times(pdipc_handle* h, double* us, long long* counts, int nclasses);

/* --- per-triangle strain stiffness of a cloth (an extension: the reference
 * takes one strain weight per cloth model, PdIpcClothModel.h:282,300).  Each
 * edge's spring weight becomes the mean stiffness of its adjacent triangles.
 * `model` indexes the simulated models in add order; one value per triangle.
 * Before pdipc_host_setup it is recorded; afterwards the matrix values are
 * recomputed and uploaded (the sparsity pattern must not change). */
int pdipc_set_cloth_stiffness(pdipc_handle* h, int model, const double* k, int ntris);

/* --- reverse-mode derivative of the FP64 step (an extension: the reference
 * has no derivative).  pdipc_rollout runs `nsteps` steps from the current
 * state keeping (x, v) every `checkpoint_every` steps; pdipc_backward takes
 * dL/dx and dL/dv of the final state (AoS over the simulated points; NULL =
 * zero) and returns dL/dx and dL/dv of the initial state and dL/dk for every
 * cloth triangle (all cloth models in add order, nk >= their total), treating
 * candidate records, barrier active sets, minimum-TOI records and iteration
 * counts as fixed.  It re-runs one checkpoint segment at a time with every
 * solver iterate recorded (the rollout records the last segment as it runs,
 * so the first backward after it re-runs all but that one); afterwards the
 * handle is left at the final state.
 * checkpoint_every = 0 records every step during the rollout instead (device
 * memory for the whole run's tape, ~0.4 GB per step of a 128^2 cloth): the
 * backward then runs no forward step and may be called again on the same
 * rollout. */
int pdipc_rollout(pdipc_handle* h, int nsteps, int checkpoint_every);
int pdipc_backward(pdipc_handle* h, const double* dL_dx, const double* dL_dv, double* dL_dx0, double* dL_dv0,
                   double* dL_dk, int nk);
/* outer iterations recorded by the last backward pass, peak tape bytes,
 * segments, steps */
int pdipc_adjoint_stats(pdipc_handle* h, double* out4);
/* phase-level adjoints for parity tests: the vector-Jacobian products of one
 * Chebyshev solve (diag, rhs, x0 -> x_out; returns dL/dx0 and dL/drhs for a
 * given dL/dx_out) and of the local step at x (dL/dx for a given dL/drhs) */
int pdipc_debug_solve_vjp(pdipc_handle* h, const double* diag, const double* rhs, const double* x0,
                          const double* xbar_out, double* xbar0, double* bbar);
int pdipc_debug_local_vjp(pdipc_handle* h, const double* x, const double* bbar, double* xbar);
/* a vector recorded by the last backward pass (its last segment): step within
 * the segment, outer iteration, which = 0 xold, 1 x at the local/global start,
 * 2 system diagonal (nsim), 3 rhs of pass l, 4 Chebyshev iterate k of pass l,
 * 5 the scalars (toi, clamped, argmin record, collision terms active) */
int pdipc_debug_tape(pdipc_handle* h, int step, int outer, int which, int l, int k, double* out);

#ifdef __cplusplus
}
#endif
#endif /* PDIPC_H */
