/*
 * stencilparareal.h -- C ABI of the MI355X (gfx950) StencilParareal hot path.
 *
 * One shared library, libsp_hip.so, built from stencilparareal_amd/csrc/.
 * Every entry point takes plain pointers and sizes, returns an int status
 * (SP_OK == 0) and never exits the process; sp_last_error() gives a
 * thread-local message for the last failure.  Device pointers are HIP device
 * memory; all work is enqueued on the context's stream (asynchronous unless the
 * function says otherwise).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   Convection ctor ................ src/Convection/Convection.h:25
 *   Convection::DoRK4 / DoEuler .... src/Convection/Convection.h:44-45, :63-64
 *   DoRK4Timestep / DoEulerTimestep  src/Convection/Convection.h:69-75
 *   ApplyImpl (periodic halo fill) . src/Communication/Periodicity.h:103
 *   UpdateStage (Parareal update) .. src/Parareal/Parareal.h:19-35
 *   Parareal::DoParallel / DoSerial  src/Parareal/Parareal.h:110-171, :250-253
 *   Parareal::DoTimedParallel ...... src/Parareal/Parareal.h:173-248
 *   computeErrorReference .......... src/Parareal/main.cpp:98-139
 *   fillQ (initial condition) ...... src/Convection/ConvectionSolution.h:23-52
 */
#ifndef STENCILPARAREAL_H_
#define STENCILPARAREAL_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define SP_OK 0
#define SP_ERR_INVALID 1     /* bad argument: null pointer, size, layout, steps < 0 */
#define SP_ERR_HIP 2         /* a HIP runtime call failed */
#define SP_ERR_UNSUPPORTED 3 /* valid request this build does not implement */
#define SP_ERR_DIVERGED 4    /* non-finite or growing Parareal increment */
#define SP_ERR_NOMEM 5       /* device allocation failed */

/* ---- element types ----------------------------------------------------------- */
#define SP_F64 0
#define SP_F32 1
#define SP_BF16 2 /* storage only; arithmetic in f32 */

/* ---- coarse propagator kinds -------------------------------------------------- */
#define SP_COARSE_EULER 0 /* DoEuler: Advection2 + Laplace2 (upwind per sp_prop_set_upwind) */
#define SP_COARSE_RK4 1   /* DoRK4 with the coarse step */
#define SP_COARSE_RK4_SPECTRAL 2 /* the same RK4 map, each coarse segment applied as its exact
                                    Fourier multiplier (FFT, product, inverse FFT; FP64 only;
                                    no counterpart in the reference) */
#define SP_COARSE_EULER_SPECTRAL 3 /* the DoEuler map (upwind per sp_prop_set_upwind) applied
                                      the same way (FP64 only; no counterpart in the reference) */

/* ---- kernel paths for the fine propagator ------------------------------------- */
#define SP_PATH_FUSED 0  /* one kernel per RK4 step, 4 stages fused (default) */
#define SP_PATH_STAGED 1 /* reference structure: 4 RHS + 3 Euler + 1 RK sweep per step */

/* ---- fused-kernel variants ------------------------------------------------------ */
#define SP_KERNEL_AUTO -1
#define SP_KERNEL_SCALAR 0 /* one column per lane, any sizes and storage types */
#define SP_KERNEL_PAIR 1   /* two x-adjacent columns per lane: FP64 (16-byte accesses) or FP32 arithmetic over FP32 / BF16 storage */
#define SP_KERNEL_QUAD 2   /* FP64, 2x2 columns per lane in stages 1-3 (32x32 tiles, 512 threads) */
#define SP_KERNEL_SPLIT 3  /* FP64, two launches of a two-stage x-pair kernel through a work field */

/*
 * Field layout.  Element (i,j,k) of the interior, 0 <= i < ni etc., lives at
 *   base[origin + i + j*sj + k*sk]
 * (i has unit stride).  halo is the number of ghost cells present on every side
 * in the storage; the propagators never read or write halos (periodicity is
 * index arithmetic), sp_periodic_fill writes them.
 *   compact layout:   halo 0, sj = ni, sk = ni*nj, origin 0
 *   STELLA CUDA ConvectionField (KJI order, i padded to 32 doubles, 2-cell halo;
 *   src/Convection/Convection.h:17; STELLA DataFieldStorageFormat.h:146-174):
 *                     halo 2, pi = pad32(ni+4), sj = pi, sk = pi*(nj+4),
 *                     origin = 2 + 2*sj + 2*sk (base = pStorageBase)
 */
typedef struct sp_layout {
    int ni, nj, nk;
    int halo;
    long long sj, sk;
    long long origin;
    int dtype; /* SP_F64 / SP_F32 / SP_BF16 */
} sp_layout;

typedef struct sp_ctx sp_ctx;
typedef struct sp_prop sp_prop;
typedef struct sp_parareal sp_parareal;

/* ---- library and context ------------------------------------------------------- */
const char* sp_version(void);
const char* sp_last_error(void);
/* Opens device `device`, creates a stream of default priority. */
int sp_ctx_create(int device, sp_ctx** out);
int sp_ctx_destroy(sp_ctx* ctx);
/* Enqueue on an external hipStream_t; NULL selects HIP's null stream.  The context's
 * own stream is what sp_ctx_get_stream returns right after sp_ctx_create. */
int sp_ctx_set_stream(sp_ctx* ctx, void* hip_stream);
int sp_ctx_get_stream(sp_ctx* ctx, void** hip_stream);
/* Blocks until all work enqueued on the context's stream has finished. */
int sp_ctx_sync(sp_ctx* ctx);
int sp_malloc(sp_ctx* ctx, long long bytes, void** out);
int sp_free(sp_ctx* ctx, void* ptr);
/* Synchronous copies between host and device memory. */
int sp_memcpy_h2d(sp_ctx* ctx, void* dst, const void* src, long long bytes);
int sp_memcpy_d2h(sp_ctx* ctx, void* dst, const void* src, long long bytes);

int sp_layout_compact(int ni, int nj, int nk, int dtype, sp_layout* out);
int sp_layout_stella_cuda(int ni, int nj, int nk, sp_layout* out);
/* Number of elements the layout's storage spans (allocation size). */
long long sp_layout_storage_elems(const sp_layout* l);

/* ---- propagator: Convection (src/Convection/Convection.h:22-108) -------------- */
/* Convection(isize, jsize, ksize, dx, nu0, nufreq, cx, cy, cz).  nu(t) =
 * nu0*(1 + sin(nufreq*t)/2) is evaluated once per step, at the step start. */
int sp_prop_create(sp_ctx* ctx, int isize, int jsize, int ksize, double dx, double nu0,
                   double nufreq, double cx, double cy, double cz, sp_prop** out);
int sp_prop_destroy(sp_prop* p);
/* Advection2 branch of DoEuler: 1 = first-order upwind (the reference build
 * default, CMakeLists.txt:101-105), 0 = second-order central. */
int sp_prop_set_upwind(sp_prop* p, int upwind);
/* SP_PATH_FUSED (default) or SP_PATH_STAGED for DoRK4. */
int sp_prop_set_path(sp_prop* p, int path);
/* Fused-kernel choice: variant SP_KERNEL_*, tile tx x ty columns, nthreads per block,
 * kchunk planes per block (0: whole k extent; the quad variant runs an odd kchunk as
 * kchunk + 1, with the same results).  variant SP_KERNEL_AUTO restores the
 * built-in choice.  Compiled shapes: scalar {32x32x1024, 32x24x1024, 24x24x1024,
 * 32x16x512, 16x16x256}, pair {32x32x512, 32x16x512, 16x16x256}, quad {32x32x512}
 * (the built-in choice for FP64 launches of 256 or more blocks; FP32 / BF16 storage takes
 * the pair kernel in its place), split {32x32x512,
 * 32x16x512, 32x16x256, 16x16x256}.  The split variant (FP64 fields, even ni) runs each
 * step as two launches through a work field the propagator allocates (8 bytes per point
 * of the batch) when the tile is chosen; its results equal the other FP64 variants' bit
 * for bit.  Grouped batches (sp_batch_*) and sp_do_rk4_noalloc run the built-in FP64
 * choice in its place. */
int sp_prop_set_tile(sp_prop* p, int variant, int tx, int ty, int nthreads, int kchunk);
/* The fused-kernel tile sp_do_rk4 (nslices = 1) or sp_do_rk4_batched (nslices > 1) will
 * use for FP64 compute (f64 = 1) or reduced-precision storage (f64 = 0): the one set with
 * sp_prop_set_tile, else the built-in choice for this size.  Output pointers may be NULL.
 * Fields the pair, quad and split kernels cannot take (odd ni, odd strides, origins not
 * aligned to two elements; quad and split: storage other than FP64) run the scalar
 * 32x16x512 kernel instead; the report covers odd ni and f64 = 0 (no counterpart in the
 * reference). */
int sp_prop_get_tile(const sp_prop* p, int nslices, int f64, int* variant, int* tx, int* ty, int* nthreads,
                     int* kchunk);
/* Autotuner (no counterpart in the reference, whose CUDA block size is fixed at 32x4):
 * times every compiled fused-kernel tile shape and k chunk for `nslices` slices of this
 * propagator's size with the given storage types, and keeps the fastest for later
 * sp_do_rk4* calls with that slice count and compute type (FP64 when both types are
 * SP_F64, else FP32), unless a tile is set with sp_prop_set_tile.  Each candidate runs for
 * about 2 ms of timed launches and the two fastest are timed again.  Results are kept for
 * the process per (device, sizes, slice count, storage types), so later propagators and
 * Parareal objects of the same sizes reuse them without timing; with the environment
 * variable SP_AUTOTUNE_CACHE set to a file path they are also read from and appended to
 * that file.  Synchronous; allocates two scratch batches while it times (none when a kept
 * result applies); refused while the stream is being captured.  *best_us
 * (nullable) receives the fastest step time in microseconds. */
int sp_prop_autotune(sp_prop* p, int nslices, int in_dtype, int out_dtype, double* best_us);

/* DoRK4: `steps` RK4 steps of size dt from tStart; step i runs at tStart + i*dt.
 * in == out is allowed.  steps == 0 leaves out untouched.  in and out may have
 * different layouts; their dtype must be SP_F64 (the fine propagator) or match
 * the coarse precision, see sp_do_rk4_ex. */
int sp_do_rk4(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
              double tStart, double dt, int steps);
/* DoEuler: forward Euler with the RHS2 stencil (SP_F64 only). */
int sp_do_euler(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                double tStart, double dt, int steps);
int sp_do_rk4_timestep(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout,
                       void* out, double t, double dt);
int sp_do_euler_timestep(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout,
                         void* out, double t, double dt);
/* RK4 with arbitrary storage types: in/out dtypes from the layouts
 * (f64/f32/bf16); arithmetic in f64 when both are f64, in f32 otherwise. */
int sp_do_rk4_ex(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                 double tStart, double dt, int steps);

/* DoRK4 that allocates nothing (no counterpart in the reference): for callers that cannot
 * call hipMalloc -- code under hipGraph capture, tight serving loops, harnesses that forbid
 * allocation.  Same arguments and result as sp_do_rk4 (bit for bit) on SP_F64 fields with
 * the fused kernels; the one intermediate field of a multi-step or in-place run lives in a
 * device array of the library (2^22 points, 32 MiB per device, held until the process
 * ends).  Fields of more than 4194304 points, the reference-order path
 * (SP_PATH_STAGED) and k extents below 8 return SP_ERR_UNSUPPORTED.  Every call on one
 * device shares that array: calls that overlap on a device (other streams or threads)
 * race. */
int sp_do_rk4_noalloc(sp_prop* p, const sp_layout* lin, const void* in, const sp_layout* lout, void* out,
                      double tStart, double dt, int steps);

/* Batched DoRK4 over `nslices` independent fields of one layout, slice b at
 * in + b*slice_stride (elements), all with the same start time; one fused kernel
 * launch per step covers every slice.  in != out required (no aliasing). */
int sp_do_rk4_batched(sp_prop* p, int nslices, const sp_layout* lin, const void* in,
                      const sp_layout* lout, void* out, long long slice_stride, double tStart,
                      double dt, int steps);

/* ---- halo fill, update, norms -------------------------------------------------- */
/* ApplyImpl: fills the layout's halo periodically, I faces (interior j,k), then J
 * faces (all i, interior k), then K faces (all i, j).  SP_F64, halo >= 1. */
int sp_periodic_fill(sp_ctx* ctx, const sp_layout* l, void* base);
/* UpdateStage: q = cnew + fine - cold on the interior (all SP_F64, same layout). */
int sp_parareal_update(sp_ctx* ctx, const sp_layout* l, void* q, const void* fine,
                       const void* cnew, const void* cold);
/* out[0] = max|a-b|, out[1] = max|a|, out[2] = sum (a-b)^2, out[3] = sum b^2 over
 * the interior (SP_F64).  Synchronous.  computeErrorReference = out[0]/out[1];
 * relative L2 = sqrt(out[2]/out[3]). */
int sp_norms(sp_ctx* ctx, const sp_layout* la, const void* a, const sp_layout* lb, const void* b,
             double out[4]);
/* fillQ with exactQ at time t on the unit cube (cell-centred points). */
int sp_fill_exact(sp_ctx* ctx, const sp_layout* l, void* base, double nu0, double nufreq,
                  double cx, double cy, double cz, double t);
/* Sum of ng Gaussians amp*exp(-d^2/(2 sigma^2)), d the periodic minimum-image
 * distance to the centre; host arrays of length ng. */
int sp_fill_gaussians(sp_ctx* ctx, const sp_layout* l, void* base, int ng, const double* xc,
                      const double* yc, const double* zc, const double* sigma, const double* amp);

/* ---- Parareal (src/Parareal/Parareal.h) --------------------------------------- */
/* The S time slices the reference gives one MPI rank each are the batch index of
 * one GPU: the fine propagation of all slices runs as one launch per step, the
 * coarse correction chain runs slice after slice, the update is fused with an
 * increment norm.  Step totals are rounded up to multiples of S as in
 * RuntimeConfiguration.cpp:133-151; dt = T / total steps. */
typedef struct sp_parareal_cfg {
    int slices;          /* S (the reference's MPI world size) */
    int kmax;            /* maximum iterations */
    double end_time;     /* T */
    int steps_fine;      /* total fine steps over [0, T] */
    int steps_coarse;    /* total coarse steps over [0, T] */
    int coarse_kind;     /* SP_COARSE_EULER (reference), SP_COARSE_RK4, SP_COARSE_RK4_SPECTRAL or
                            SP_COARSE_EULER_SPECTRAL */
    int coarse_dtype;    /* SP_F64, SP_F32 or SP_BF16 (RK4 only for the latter two) */
    int coarse_substeps; /* >= 1: each coarse step is this many sub-steps */
    double tol;          /* stop once the relative L2 increment <= tol; 0: run kmax */
    int async_schedule;  /* 1: fine-step times of the reference's async mode */
    int use_graphs;      /* 1: capture each iteration as two hipGraphs, the fine sweep and the coarse
                            chain with its updates, replayed back to back */
    int autotune;        /* 1: autotune the fused-kernel tiles of F and G at create (sp_prop_autotune) */
    int delta_correction; /* 0: the reference's update G(U_n^{k+1}) + F(U_n^k) - G(U_n^k)
                             (Parareal.h:19-35); 1: F(U_n^k) + G(U_n^{k+1} - U_n^k), equal for
                             this linear G, whose rounding then shrinks with the increment --
                             the form for SP_F32 / SP_BF16 coarse storage */
    int initial_fp64;    /* 1 (with delta_correction and SP_F32 / SP_BF16 coarse storage): the
                            initial coarse run U_n^0 = G^n(u0) in FP64, the iterations' coarse
                            steps in coarse_dtype; 0: all in coarse_dtype (no counterpart in
                            the reference, whose G is one type) */
} sp_parareal_cfg;

#define SP_MAX_ITERS 64
typedef struct sp_parareal_report {
    int iterations;                    /* iterations actually run */
    int steps_fine_per_slice, steps_coarse_per_slice;
    double dt_fine, dt_coarse;         /* effective steps after rounding */
    double increment[SP_MAX_ITERS];    /* relative L2 change of the slice states */
    double t_init;                     /* seconds: initial coarse sweep (wall clock) */
    double t_fine[SP_MAX_ITERS];       /* seconds per iteration: the batched fine sweep, between
                                          HIP events on the stream (with use_graphs the sweep and
                                          the chain are two graphs, timed the same way) */
    double t_coarse[SP_MAX_ITERS];     /* seconds: the coarse chain incl. its corrections (events) */
    double t_total;                    /* seconds: the whole run (wall clock) */
    double ms_fine_events[SP_MAX_ITERS]; /* ms: t_fine in milliseconds */
    double t_update[SP_MAX_ITERS];     /* seconds: the corrections alone, each launch between
                                          events (iterations without a graph; 0 with graphs) */
} sp_parareal_report;

int sp_parareal_create(sp_prop* p, const sp_parareal_cfg* cfg, sp_parareal** out);
int sp_parareal_destroy(sp_parareal* pr);
/* Initialisation of DoParallel (Parareal.h:113-114): slice states U_n = G^n(u0),
 * coarse values G(U_n).  u0: SP_F64 field on the device. */
int sp_parareal_begin(sp_parareal* pr, const sp_layout* l0, const void* u0);
/* One iteration (Parareal.h:116-165).  *increment (nullable) receives the
 * relative L2 change of all slice states.  Synchronous. */
int sp_parareal_iterate(sp_parareal* pr, double* increment);
/* Phase times of the last sp_parareal_iterate, in seconds, between HIP events on the
 * stream: the fine sweep, the coarse chain with its corrections, and the corrections alone
 * (0 when the iteration ran as graphs).  Outputs may be NULL.  Counterpart of the per-
 * iteration stamps of Parareal::DoTimedParallel (Parareal.h:173-248). */
int sp_parareal_iteration_times(const sp_parareal* pr, double* fine_s, double* chain_s, double* update_s);
/* Copies the current last-slice solution (the last rank's q) into q. */
int sp_parareal_solution(sp_parareal* pr, const sp_layout* lq, void* q);
/* DoParallel: begin + up to kmax iterations (stopping at tol) + solution. */
int sp_parareal_run(sp_parareal* pr, const sp_layout* l0, const void* u0, const sp_layout* lq,
                    void* q, sp_parareal_report* rep);
/* DoSerial: fine propagation over all slices, DoRK4(u0 -> q, 0, dtF, tsF*S). */
int sp_parareal_serial(sp_parareal* pr, const sp_layout* l0, const void* u0, const sp_layout* lq,
                       void* q);

/* ---- batches of independent Parareal problems ------------------------------------ */
/* No counterpart in the reference, which runs one problem per MPI job.  N independent
 * problems (own grid, slice count, coefficients, horizon) run as one Parareal whose
 * launches each cover all of them: one grouped fused-RK4 launch per fine step (per tile
 * shape in use), one grouped launch per coarse sub-step at each position of the coarse
 * chains, one grouped correction launch per position.  F is FP64; G is RK4 in SP_F64,
 * SP_F32 or SP_BF16 storage; the correction is the difference form (see
 * sp_parareal_cfg.delta_correction).  Constant diffusion (nufreq = 0). */
typedef struct sp_problem {
    int n;                        /* grid points per dimension (n^3), even, >= 8 */
    int slices;                   /* time slices S_p */
    double nu, cx, cy, cz;        /* diffusion, advection velocity */
    double end_time;              /* T_p */
    int steps_fine, steps_coarse; /* totals over [0, T_p], rounded up to multiples of S_p */
} sp_problem;

typedef struct sp_batch_cfg {
    int kmax;            /* maximum iterations */
    int coarse_dtype;    /* SP_F64, SP_F32 or SP_BF16 */
    int coarse_substeps; /* >= 1 sub-steps per coarse step */
    double tol;          /* stop once every problem's relative L2 increment <= tol; 0: run kmax */
    int use_graphs;      /* 1: capture one iteration as a hipGraph and replay it */
    int autotune;        /* 1: autotune the fused tiles once per (grid, slice count) */
} sp_batch_cfg;

typedef struct sp_batch_report {
    int iterations;
    double increment[SP_MAX_ITERS]; /* largest relative L2 increment over the problems */
    double t_init, t_total;         /* seconds: initial coarse chains; whole run */
    double lattice_updates;         /* fine + coarse point updates of the run */
} sp_batch_report;

typedef struct sp_batch_info {
    int steps_fine_per_slice, steps_coarse_per_slice;
    double dt_fine, dt_coarse;
    int fine_variant, fine_tx, fine_ty, fine_nthreads, fine_kchunk; /* fine-sweep tile */
} sp_batch_info;

typedef struct sp_batch sp_batch;
int sp_batch_create(sp_ctx* ctx, int nprob, const sp_problem* probs, const sp_batch_cfg* cfg, sp_batch** out);
int sp_batch_destroy(sp_batch* b);
/* u0[p]: problem p's initial field, q[p]: receives its last slice's state at T_p; both
 * compact SP_F64 device arrays of n_p^3 doubles.  Synchronous. */
int sp_batch_run(sp_batch* b, void* const* u0, void* const* q, sp_batch_report* rep);
/* Problem p's relative L2 increments of the last run, one per iteration. */
int sp_batch_increments(const sp_batch* b, int p, double* inc);
/* Iteration k (0-based) of the last run, timed by HIP events: the fine sweep (every
 * problem's fine steps) and the coarse chains with the corrections, in seconds. */
int sp_batch_iteration_times(const sp_batch* b, int k, double* fine_s, double* chain_s);
int sp_batch_problem_info(const sp_batch* b, int p, sp_batch_info* info);
/* Problem p's serial FP64 fine solution: DoRK4(u0 -> q, 0, dtF_p, tsF_p * S_p). */
int sp_batch_serial(sp_batch* b, int p, const void* u0, void* q);

#ifdef __cplusplus
}
#endif
#endif /* STENCILPARAREAL_H_ */
