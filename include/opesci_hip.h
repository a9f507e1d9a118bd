/*
 * opesci-fd MI355X backend -- public C ABI of libopesci_hip.so.
 *
 * The reference compiles one C++ translation unit per model and loads it with
 * ctypes (opesci/grid.py:85-156); that TU exports opesci_execute,
 * opesci_convergence and opesci_free over the structs OpesciGrid,
 * OpesciProfiling and OpesciConvergence (opesci/templates/regular3d_tmpl.py:
 * 19-23, 41-59, 99-134).  Here the compute lives in one prebuilt gfx950
 * library and the per-model part the generator used to bake into the TU
 * (constants, folded stencil coefficients, the boundary sequence, the
 * analytic initial state) arrives as data: an OpesciModelDesc produced by the
 * Python generator (opesci/plan.py).  Plain pointers and sizes only.
 *
 * Field order everywhere: U, V, W, Txx, Tyy, Tzz, Txy, Tyz, Txz
 * (StaggeredGrid.fields = vfields + sfields, opesci/staggeredgrid.py:69-71).
 * Host arrays use the reference layout real_t[tp=2][dim1][dim2][dim3],
 * dim3 contiguous (opesci/regulargrid.py:466-496).
 *
 * Every entry point returns 0 on success and a nonzero OPESCI_E* code
 * otherwise (the reference always returns 0); opesci_last_error() gives the
 * message.  Calls on one context must not overlap; different contexts are
 * independent.
 */
#ifndef OPESCI_HIP_H
#define OPESCI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPESCI_ABI_VERSION 2
#define OPESCI_NFIELDS 9
#define OPESCI_MAX_HALF_ORDER 4

enum {
    OPESCI_OK = 0,
    OPESCI_EINVAL = 1,      /* malformed model description or argument */
    OPESCI_EHIP = 2,        /* HIP runtime error */
    OPESCI_ENOMEM = 3,      /* device or host allocation failed */
    OPESCI_ESTATE = 4,      /* call out of order (e.g. convergence before execute) */
    OPESCI_EIO = 5,         /* media file missing or short */
    OPESCI_EUNSUPPORTED = 6 /* configuration without a kernel instantiation */
};

enum { OPESCI_REAL_F32 = 4, OPESCI_REAL_F64 = 8 };
enum { OPESCI_STORE_NATIVE = 0, OPESCI_STORE_F16 = 1, OPESCI_STORE_BF16 = 2 };

/* Media identifiers (constant values or arrays; opesci/staggeredgrid.py:234-282) */
enum {
    OPESCI_M_RHO = 0, OPESCI_M_VP, OPESCI_M_VS, OPESCI_M_BETA, OPESCI_M_LAMBDA, OPESCI_M_MU,
    OPESCI_M_BETA1, OPESCI_M_BETA2, OPESCI_M_BETA3, OPESCI_M_MU12, OPESCI_M_MU13, OPESCI_M_MU23,
    OPESCI_M_COUNT
};

/* Expression bytecode (analytic solutions, media-dependent boundary
 * coefficients).  Evaluated in double on the device. */
enum {
    OPESCI_OP_CONST = 0, /* push val */
    OPESCI_OP_COORD,     /* push coordinate of axis a[0] (real_t-rounded, field staggering) */
    OPESCI_OP_TIME,      /* push t */
    OPESCI_OP_MEDIA,     /* push media a[0] at anchor + (a[1],a[2],a[3]) */
    OPESCI_OP_ADD, OPESCI_OP_SUB, OPESCI_OP_MUL, OPESCI_OP_DIV, OPESCI_OP_NEG,
    OPESCI_OP_POWC,      /* top = pow(top, val) */
    OPESCI_OP_SQRT, OPESCI_OP_SIN, OPESCI_OP_COS, OPESCI_OP_EXP, OPESCI_OP_LOG,
    OPESCI_OP_ROUNDF     /* top = (double)(float)top */
};
typedef struct OpesciExprOp {
    int32_t op;
    int32_t a[4];
    int32_t pad;
    double val;
} OpesciExprOp;

/* One boundary assignment dst[plane][..][..] = sum_i coef_i * src_i over a
 * plane loop (StaggeredGrid.stress_bc_getter / velocity_bc,
 * opesci/staggeredgrid.py:754-864).  Statements carry the launch index the
 * generator's hazard analysis gave them: statements with the same sched in
 * one phase touch disjoint cells and run in one launch; ascending sched
 * reproduces the reference's sequential order exactly. */
typedef struct OpesciBcTerm {
    int32_t field;   /* source field */
    int32_t level;   /* 0: t0, 1: t1 */
    int32_t plane;   /* absolute index along the statement's axis */
    int32_t off[2];  /* offsets along the two other axes (ascending axis order) */
    int32_t prog;    /* -1: coefficient is coef; else offset of a media program in bc_code */
    int32_t prog_len;
    int32_t pad;
    double coef;
} OpesciBcTerm;

typedef struct OpesciBcStmt {
    int32_t field;   /* destination field */
    int32_t level;   /* 0: t0, 1: t1 */
    int32_t axis;    /* plane normal: 0 x, 1 y, 2 z */
    int32_t plane;   /* absolute destination index along axis */
    int32_t lo[2];   /* loop range over the other two axes */
    int32_t hi[2];
    int32_t term0;   /* first term in the term table */
    int32_t nterms;
    int32_t phase;   /* 0: init stress, 1: init velocity, 2: step stress, 3: step velocity,
                        4: step stress, rows of z-face loops away from the x/y faces: they
                           commute and run before every phase-2 launch (the generator
                           proves the result equals the reference order) */
    int32_t sched;   /* launch index within the phase */
} OpesciBcStmt;

/* Derivative of a media-dependent boundary coefficient with respect to one
 * media cell it reads (for the adjoint): term = global term index, media id
 * and offsets as in OPESCI_OP_MEDIA, prog = bytecode of the partial
 * derivative (without the term's literal factor).  Entries of one term are
 * contiguous and ordered by term. */
typedef struct OpesciBcDTerm {
    int32_t term;
    int32_t media;
    int32_t off[3];
    int32_t prog;
    int32_t prog_len;
    int32_t pad;
} OpesciBcDTerm;

/* Everything the generator bakes into the reference's emitted TU. */
/* OpesciModelDesc.flags: run both half steps of a time step in one sweep
 * (the fused step) over the rows clear of the x and y faces, where it
 * applies (FP32 arithmetic, constant media, half order 1 or 2, the deep
 * z-face statements in the stress sweep).  Results equal the separate
 * launches bit for bit; on MI355X it is slower than they are (256^3, 4th
 * order: 694 us against 386 us per step), so it is off by default. */
#define OPESCI_FLAG_FUSED_STEP 1
/* OpesciModelDesc.flags: evaluate the analytic-solution programs (initial
 * state, L2 norms) at every point as given.  By default every part of a
 * program that reads one coordinate (or none) and no media is evaluated once
 * per index along that axis into a table the point evaluation reads: the
 * same double operations in the same order, so the values are bit for bit
 * those of this flag, at a fraction of the cost. */
#define OPESCI_FLAG_POINTWISE_SOLUTION 2

typedef struct OpesciModelDesc {
    int32_t abi_version;
    int32_t dims[3];        /* dim1..dim3 incl. ghosts: grid_size + 1 + 2m */
    int32_t space_order;    /* 2M */
    int32_t ntsteps;
    int32_t real_bytes;     /* OPESCI_REAL_F32 / OPESCI_REAL_F64 */
    int32_t storage;        /* OPESCI_STORE_* (16-bit storage needs real_bytes 4) */
    int32_t media_arrays;   /* 0 constant media, 1 arrays from rho/vp/vs */
    int32_t flags;          /* OPESCI_FLAG_* bits, 0 by default */
    double dt;
    double spacing[3];
    double media_const[OPESCI_M_COUNT];  /* constant media values (float-rounded as the reference declares them) */
    /* interior update coefficients, float-rounded like the reference's F literals:
     *   normal stress i:  T_ii += sum_j sum_k cn[i][j][k] * m * (V_j[+k] - V_j[-1-k]) along j,
     *                      m = 1 (constant media) or lambda(+2mu for j == i) (arrays)
     *   shear pair s (xy,yz,xz), group g: T_s += cs[s][g][k] * m * (V[+1+k] - V[-k])
     *   velocity i:       V_i += sum_j cv[i][j][k] * m * (T[a+k] - T[a-1-k]) along j */
    double cn[3][3][OPESCI_MAX_HALF_ORDER];
    double cs[3][2][OPESCI_MAX_HALF_ORDER];
    double cv[3][3][OPESCI_MAX_HALF_ORDER];
    /* media arrays (media_arrays = 1): dim1*dim2*dim3 float32 each */
    const float* rho;
    const float* vp;
    const float* vs;
    /* boundary program */
    int32_t nstmts;
    int32_t nterms;
    int32_t ncode;
    int32_t pad0;
    const OpesciBcStmt* stmts;
    const OpesciBcTerm* terms;
    const OpesciExprOp* bc_code;
    /* analytic solution per field: code[sol_off[f] .. sol_off[f]+sol_len[f]) */
    int32_t sol_off[OPESCI_NFIELDS];
    int32_t sol_len[OPESCI_NFIELDS];
    int32_t nsol;
    int32_t pad1;
    const OpesciExprOp* sol_code;
    /* point source (Ricker etc.) into Txx/Tyy/Tzz and receivers
     * (tests/src/test_ref_iso_elastic.cpp:227-290) */
    int32_t src_on;             /* 0 none; 1 one trace for Txx, Tyy, Tzz; 3 traces x, y, z */
    int32_t src_idx[3];
    const float* src_wavelet;   /* src_on * ntsteps values (x trace, then y, then z);
                                   Txx -= wx[t]/3, Tyy -= wy[t]/3, Tzz -= wz[t]/3 after the
                                   velocity step and the receiver sampling */
    int32_t nrec;
    int32_t pad2;
    const int32_t* rec_idx;     /* 3*nrec grid indices */
    /* launch tuning (0 = library default) */
    int32_t tile_y;
    int32_t x_chunk;
    int32_t kernel_variant;     /* 0 default (LDS/register-queue sweep, two z points per lane
                                   for FP32 arithmetic), 1 plain gather, 2 sweep with one
                                   point per lane (the FP64 default) */
    int32_t use_graph;          /* capture the time loop in a hipGraph */
    /* adjoint: derivative programs of the media-dependent boundary
     * coefficients (may be 0 / null when no gradient is asked for) */
    int32_t ndterms;
    int32_t pad3;
    const OpesciBcDTerm* dterms;
} OpesciModelDesc;

/* Reference ABI structs.  Field pointers are real_t* (float or double). */
typedef struct OpesciGrid {
    void* U; void* V; void* W;
    void* Txx; void* Tyy; void* Tzz;
    void* Txy; void* Tyz; void* Txz;
} OpesciGrid;

typedef struct OpesciProfiling {
    float g_rtime;   /* seconds of the device time loop */
    float g_ptime;   /* same (there is no separate process time on the device) */
    float g_mflops;  /* algorithmic MFlop/s (reference op count) */
} OpesciProfiling;

typedef struct OpesciContext OpesciContext;

/* --- reference-named entry points ---------------------------------------
 * opesci_model_load binds a model to the process-default context, the way
 * loading a generated .so did (opesci/grid.py:35-42).  Then:
 * opesci_execute      <- generated TU, opesci/templates/regular3d_tmpl.py:44-59
 *   allocates the fields on the device, applies the initial condition and the
 *   initial boundary conditions, runs ntsteps on the GPU, and hands back host
 *   arrays real_t[2][dim1][dim2][dim3] (both time levels) through grid, which
 *   the caller releases with opesci_free.
 * opesci_convergence  <- regular3d_tmpl.py:106-112, staggeredgrid.py:892-945
 *   conv points to 9 real_t: sqrt(sum (F - F_exact)^2 dx dy dz) per field,
 *   accumulated in double on the device.
 * opesci_free         <- regular3d_tmpl.py:114-118 */
int opesci_model_load(const OpesciModelDesc* desc);
int opesci_execute(OpesciGrid* grid, OpesciProfiling* profiling);
int opesci_convergence(OpesciGrid* grid, void* conv);
int opesci_free(OpesciGrid* grid);

/* --- context API (no reference counterpart: the reference has one model per
 * generated library; these expose the same steps per handle) ------------- */
int opesci_ctx_create(const OpesciModelDesc* desc, OpesciContext** out);
int opesci_ctx_destroy(OpesciContext* ctx);
/* initial condition + initial boundary conditions (staggered3d_tmpl.py:21-22) */
int opesci_ctx_init(OpesciContext* ctx);
/* advance nsteps on the device (staggered3d_tmpl.py:40-58); seconds = device time */
int opesci_ctx_run(OpesciContext* ctx, int32_t nsteps, double* seconds);
/* per-field L2 vs the analytic solution at the current time (double) */
int opesci_ctx_l2(OpesciContext* ctx, double out[OPESCI_NFIELDS]);
/* The same L2 norms without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The per-block
 * partial sums live in a fixed library-owned device array of 9 x 512
 * doubles (36 KiB per device, held until the process ends; the partial
 * sums are grid-stride, so any grid size fits); `d_out` is device memory
 * for 9 doubles; the work is enqueued on `stream` (a hipStream_t; null =
 * the context's stream) with no host synchronisation.  The context's device
 * must be current.  Calls that overlap on one device race on the array, and
 * calls that overlap on one context also on the context's tables of the
 * analytic solution (allocated with the context, refilled by every L2).
 * Same block count and summation order as opesci_ctx_l2. */
int opesci_ctx_l2_noalloc(OpesciContext* ctx, double* d_out, void* stream);
/* copy one field, one time level (0/1, or -1 = latest) to/from a dense host array real_t[dim1][dim2][dim3] */
int opesci_ctx_download(OpesciContext* ctx, int32_t field, int32_t level, void* host);
int opesci_ctx_upload(OpesciContext* ctx, int32_t field, int32_t level, const void* host);
/* receiver traces recorded so far: float[nsteps_done][nrec][4] (U, V, W, pressure);
   replaces the receiver arrays that tests/src/test_ref_iso_elastic.cpp:227-272
   fills.  FP64 contexts record FP64 traces and round them here. */
int opesci_ctx_receivers(OpesciContext* ctx, float* host, int64_t capacity_floats);
/* the same traces as double[nsteps_done][nrec][4], unrounded for FP64 contexts
   (no counterpart in the reference, which records float) */
int opesci_ctx_receivers_f64(OpesciContext* ctx, double* host, int64_t capacity);
/* steps completed since init */
int opesci_ctx_steps(OpesciContext* ctx, int32_t* steps);
/* average device time per step of the hot kernels over nsteps steps (HIP
 * events on the context stream): [0] stress, [1] velocity, [2] all BC
 * launches of a step; with the fused step [0] is the fused kernel and [1]
 * the stress and velocity launches of the frame around its box */
int opesci_ctx_kernel_times(OpesciContext* ctx, int32_t nsteps, double out_ms[3]);
/* Execution plan of a context: out[0] = boundary launches per time step,
 * out[1] = phase-4 statements (deep z-face rows of the stress phase),
 * out[2] = launches they take (0 or 1), out[3] = 1 when the stress sweep
 * applies them in place (FP32 pair sweeps), else 0, out[4] = the fused
 * step's margin F (its box is [F, dim - F) in x and y; 0: separate stress
 * and velocity launches).  No reference counterpart. */
int opesci_ctx_stats(OpesciContext* ctx, int32_t out[5]);

/* Launch configuration of the interior kernels (kernel_variant 0 = sweep,
 * two z points per lane for FP32 arithmetic, tile_y 0 | 4 | 8 | 16: stress
 * tiles have 4 or 8 rows (default 8, 16 runs as 8), velocity tiles 16 at
 * orders 2 and 4 when asked or by default with FP32 storage, else as the
 * stress tiles; 1 = gather; 2 = sweep with one point per lane, tile_y 0
 * (default: 16 for FP32 constant media, else 8) | 4 | 8 | 16; x_chunk 0 =
 * default), used by the autotuner.  Results do not depend on it (same
 * arithmetic in every configuration).  No reference counterpart. */
int opesci_ctx_configure(OpesciContext* ctx, int32_t kernel_variant, int32_t tile_y, int32_t x_chunk);

/* Batches: independent contexts (any shapes, orders, media) stepped together
 * by one graph launch per two time steps, their kernels running side by
 * side.  Every context must be initialised and at the same step.  The
 * contexts stay owned by the caller and must outlive the batch.  No
 * reference counterpart (the reference runs one model per process). */
/* Gradient of the receiver misfit J = 1/2 sum_t sum_r sum_{c in U,V,W}
 * (rec_t[r][c] - observed[t][r][c])^2 with respect to the context's media
 * arrays (read path, native storage, receivers set), by the discrete adjoint
 * of the whole time loop (interior updates and boundary program), with the
 * forward states recomputed from checkpoints taken every `checkpoint_every`
 * steps.  observed: ntsteps * nrec * 3 floats.  Runs the forward from the
 * initial state; afterwards the context holds a recomputed early state
 * (call opesci_ctx_init before stepping again).  The gradients stay on the
 * device: opesci_ctx_download_gradient(ctx, OPESCI_M_*, host) gives one
 * media array's, dense [dim1][dim2][dim3] in the compute type; nonzero for
 * lambda, mu, mu12, mu13, mu23, beta1, beta2, beta3.  The initial state is
 * taken as independent of the media (exact for a start from rest).  No
 * reference counterpart. */
int opesci_ctx_gradient(OpesciContext* ctx, const float* observed, int64_t count, int32_t checkpoint_every,
                        double* misfit);
/* opesci_ctx_gradient in two halves, for callers that want the misfit
 * first and the gradient only on demand (an autograd backward): the forward
 * pass with checkpoints and J, then the reverse sweep.  Any init, run,
 * upload or media upload in between drops the forward pass (the reverse
 * sweep then fails with OPESCI_ESTATE).  `observed` and the host/device
 * buffers of the media and gradient copies may be host or device memory.
 * No reference counterpart. */
int opesci_ctx_misfit_forward(OpesciContext* ctx, const float* observed, int64_t count, int32_t checkpoint_every,
                              double* misfit);
int opesci_ctx_gradient_reverse(OpesciContext* ctx);
int opesci_ctx_download_gradient(OpesciContext* ctx, int32_t media, void* host);
/* Replace one media array (dense [dim1][dim2][dim3], compute type) and
 * recompute the media-dependent boundary coefficients: the hook through
 * which effective media computed elsewhere (e.g. by autograd) drive the
 * propagator.  No reference counterpart. */
int opesci_ctx_upload_media(OpesciContext* ctx, int32_t media, const void* host);
/* One effective-media array as the device holds it (dense [dim1][dim2][dim3],
 * compute type): the read path's beta1..3 / lambda / mu / mu12.. built by
 * the media kernels from rho, Vp, Vs, the arrays the reference computes in
 * StaggeredGrid.read_data (opesci/staggeredgrid.py:572-597). */
int opesci_ctx_download_media(OpesciContext* ctx, int32_t media, void* host);

typedef struct OpesciBatch OpesciBatch;
int opesci_batch_create(OpesciContext** ctxs, int32_t n, OpesciBatch** out);
int opesci_batch_run(OpesciBatch* batch, int32_t nsteps, double* seconds);
int opesci_batch_destroy(OpesciBatch* batch);

const char* opesci_last_error(void);
int opesci_abi_version(void);
/* device name and architecture of device 0 ("" if none) */
int opesci_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len);

#ifdef __cplusplus
}
#endif
#endif /* OPESCI_HIP_H */
