/*
 * pyfwi_hip.h -- C ABI of the MI355X (gfx950) elastic propagator and
 * adjoint-state gradient.
 *
 * The ABI replaces the OpenCL layer under PyFWI's WavePropagator
 * (src/PyFWI/wave_propagation.py): the kernel build with macro-pasted
 * constants (:181-192, :409-509), the device buffers (:196-277), the per-step
 * kernel enqueues of the forward loop (:625-733) and of the adjoint loop
 * (:735-856) and the gradient read-back (:279-331).  The Python mirror of
 * WavePropagator (pyfwi_amd/wave_propagation.py) binds it with ctypes.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a PYFWI_E* code; the text of
 *     the last error on the calling thread is pyfwi_last_error().
 *   - Arrays are float32, row-major, x fastest, over the grid including the
 *     absorbing layer (tnz x tnx), like the reference's buffers.
 *   - `mem` says where caller pointers live: PYFWI_MEM_HOST or
 *     PYFWI_MEM_DEVICE (a device pointer on the plan's GPU, e.g. a torch
 *     tensor's data_ptr()).
 *   - A plan is used by one thread at a time; calls on one plan are
 *     serialised on the plan's stream.  Plans are independent.
 */
#ifndef PYFWI_HIP_H
#define PYFWI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PYFWI_OK = 0,
    PYFWI_EINVAL = 1,   /* invalid argument or geometry             */
    PYFWI_ENOMEM = 2,   /* device or host allocation failed         */
    PYFWI_EHIP = 3,     /* HIP runtime error                        */
    PYFWI_ESTATE = 4,   /* call out of order (e.g. no forward yet)  */
    PYFWI_ENONFINITE = 5, /* model holds NaN or Inf                  */
    PYFWI_ECFL = 6       /* dt above the stability limit of the model
                            (Fdm.dt_computation, fwi_tools.py:634-647)  */
};

enum { PYFWI_MEM_HOST = 0, PYFWI_MEM_DEVICE = 1 };

/* Gradient modes (pyfwi_gradient `mode`). */
enum {
    /* The reference's adjoint: reconstruct the forward state by restoring the
     * snapshots taken at `chp` and reversing single steps in between, inject
     * residual row k at backward step k, adjoint stresses then adjoint
     * velocities, correlate (wave_propagation.py:751-848, elastic.cl:372-526). */
    PYFWI_MODE_PYFWI = 0,
    /* Exact discrete adjoint of the forward (FD-checkable), forward states
     * recomputed from checkpoints kept every `ckpt_every` steps. */
    PYFWI_MODE_EXACT = 1
};

/* pyfwi_forward `ckpt_every`: keep the exact adjoint's checkpoints at the
 * interval ceil(sqrt(nt)) when they fit in device memory. */
enum { PYFWI_CKPT_AUTO = -1 };

/* Flags (pyfwi_gradient `flags`). */
enum {
    PYFWI_F_PRECOND = 1 /* also accumulate the squared-energy preconditioners */
};

/* Grid and stencil (the macro block of wave_propagation.py:434-490). */
typedef struct pyfwi_grid {
    int32_t tnz, tnx;   /* rows, columns including the absorbing layer (Nz, Nx) */
    int32_t npml;       /* absorbing-layer thickness in cells (npml)            */
    int32_t sdo;        /* half stencil width: 2 for order 4, 4 for order 8      */
    double c[4];        /* stencil coefficients as the kernels see them (c1..c4) */
    double dt;          /* internal time step as the kernels see it (dt)         */
    double dx, dz;      /* grid spacing as the kernels see it (dx, dz)           */
} pyfwi_grid;

/* Receiver layouts (pyfwi_shot.rec_layout). */
enum {
    /* One horizontal line: receiver r at row rec_z, column rec_x0 + r*rec_dx
     * (the surface gather of elastic_surface.cl:7-72). */
    PYFWI_REC_LINE = 0,
    /* Any cells, from the plan's receiver table (pyfwi_set_receivers): the
     * well gathers of elastic_crosswell.cl:1-58 and elastic.cl:43-98,
     * 316-363. */
    PYFWI_REC_TABLE = 1
};

/* Receiver flags (pyfwi_shot.rec_flags). */
enum {
    /* The vz record reads vx, as the crosswell gather does
     * (elastic_crosswell.cl:27); the residual of vz still goes to the
     * adjoint vz in PYFWI_MODE_PYFWI (elastic_crosswell.cl:51). */
    PYFWI_REC_VZ_FROM_VX = 1
};

/* One shot: a point source and its receivers
 * (injSrc / Adj_injSrc of elastic_surface.cl:7-72, elastic_crosswell.cl:1-58,
 * elastic.cl:43-98, 316-363). */
typedef struct pyfwi_shot {
    int32_t src_x, src_z;   /* source cell (column, row), absorbing layer incl.   */
    int32_t rec_z;          /* PYFWI_REC_LINE: receiver row (rec_surface_const)   */
    int32_t rec_x0;         /* PYFWI_REC_LINE: first receiver column (rec_surface_var) */
    int32_t rec_dx;         /* PYFWI_REC_LINE: receiver spacing in cells (dxr), >= 1 */
    int32_t n_rec;          /* record columns filled; all receivers inside the grid */
    int32_t rec_stride;     /* floats per record row (>= n_rec; the reference's Nr) */
    int32_t nt;             /* internal time steps                                 */
    int32_t dt_scale;       /* internal steps per record sample, >= 1              */
    int32_t n_samples;      /* record rows (nt_ext)                                */
    int64_t src_offset;     /* float offset of this shot's source table
                               [ceil(nt/dt_scale)][5] (vx, vz, txx, tzz, txz)     */
    int64_t rec_offset;     /* float offset of this shot's record / residual
                               [5][n_samples][rec_stride]                         */
    int32_t rec_layout;     /* PYFWI_REC_LINE or PYFWI_REC_TABLE                  */
    int32_t rec_flags;      /* PYFWI_REC_* flags                                   */
    int64_t rec_table;      /* PYFWI_REC_TABLE: offset of this shot's entries in
                               the plan's receiver table: [2][n_rec] cells
                               (row*tnx + column) -- first the cell each record
                               column samples, then the cell each residual column
                               is injected at by PYFWI_MODE_PYFWI (the exact mode
                               injects at the sampled cells); -1 for a column
                               that no receiver fills (recorded as zero, never
                               injected) */
} pyfwi_shot;

typedef struct pyfwi_plan pyfwi_plan;

/* Library version (major*100 + minor). */
int32_t pyfwi_version(void);

/* Text of the last error on this thread ("" if none). */
const char *pyfwi_last_error(void);

/* Create a plan on the current HIP device for one grid.  Replaces the
 * program build and buffer set-up of WavePreparation.__init__
 * (wave_propagation.py:181-246).  `stream` is a hipStream_t or NULL for a
 * plan-owned stream. */
int32_t pyfwi_plan_create(const pyfwi_grid *grid, void *stream, pyfwi_plan **out);

/* Release every device buffer of the plan. */
int32_t pyfwi_plan_destroy(pyfwi_plan *plan);

/* Use another stream for later calls (e.g. torch's current stream). */
int32_t pyfwi_plan_set_stream(pyfwi_plan *plan, void *stream);

/* Upload the medium: lam, mu, buoyancy (1/rho) [tnz*tnx] and the damping
 * profiles dx_prof [tnx] (per column) and dz_prof [tnz] (per row), already
 * scaled by the model's maximum P velocity.  Replaces elastic_buffers and
 * pml_preparation (wave_propagation.py:570-591, :400-407); the reference's
 * full-grid damping arrays are constant along one axis.
 * Returns PYFWI_ENONFINITE when the model holds NaN or Inf, and PYFWI_ECFL
 * when the grid's dt exceeds the stability limit of Fdm.dt_computation
 * (fwi_tools.py:634-647): 2 / (vp_max sqrt(5 sum|c_k| (1/dx^2 + 1/dz^2))),
 * vp_max = max sqrt((lam + 2 mu) b); the reference runs such a model and
 * diverges.  The model is kept (not used) after either error until a valid
 * one is uploaded. */
int32_t pyfwi_set_model(pyfwi_plan *plan, const float *lam, const float *mu,
                        const float *buoy, const float *dx_prof,
                        const float *dz_prof, int32_t mem);

/* Receiver table of the plan, read by the shots whose rec_layout is
 * PYFWI_REC_TABLE; `table` is host memory, `count` int32 entries, each a cell
 * row*tnx + column inside the grid or -1.  Replaces the receiver macros
 * rec_top_left_const .. rec_surface_var that kernel_caller pastes into the
 * program (wave_propagation.py:475-490, set at :116-143).  The table stays
 * until the next call; a call with count 0 clears it. */
int32_t pyfwi_set_receivers(pyfwi_plan *plan, const int32_t *table, int64_t count);

/* Forward modelling of `nshots` shots in one batch (forward_propagator +
 * __kernel + make_seismogram, wave_propagation.py:625-733, 334-366).
 *   src   source tables indexed by shot.src_offset
 *   seis  records indexed by shot.rec_offset: seis[f][k][r] = field f at
 *         receiver r after k*dt_scale steps (before that step's injection);
 *         rows that are never sampled are written as zero
 *   chp, nchp  steps after which the state is kept on the device for a later
 *         PYFWI_MODE_PYFWI gradient (the reference's checkpoint list); may be
 *         NULL/0.  A list of every step 0..nt-1 (nt the largest shot's) keeps
 *         each shot's own steps 0..nt_s-1, so shots may differ in nt; any
 *         other list needs equal nt for all shots.
 *   ckpt_every  > 0: also keep the state at the start of every ckpt_every-th
 *         step for a later PYFWI_MODE_EXACT gradient (0: none);
 *         PYFWI_CKPT_AUTO: every ceil(sqrt(nt)) steps, kept only when all
 *         shots' checkpoints and one shot's adjoint working set fit in 90 %
 *         of the free device memory (else none: the gradient then recomputes
 *         them pass by pass) */
int32_t pyfwi_forward(pyfwi_plan *plan, const pyfwi_shot *shots, int32_t nshots,
                      const float *src, float *seis, const int32_t *chp,
                      int32_t nchp, int32_t ckpt_every, int32_t mem);

/* Forward modelling that allocates nothing: pyfwi_forward's records for
 * `nshots` shots in one pass, with no checkpoint list, on caller device
 * memory only (`src`, `seis` as with PYFWI_MEM_DEVICE).  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * forward states and the shot table live in one fixed device array of
 * PYFWI_SCRATCH_FLOATS floats (64 MiB), which each GPU holds from the first
 * plan created on it until the process ends; a call whose shots need more
 * (shot table + nshots * 5 padded fields) returns PYFWI_EINVAL.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (two threads,
 * two streams) race.  The plan's model (pyfwi_set_model) and receiver table
 * are used as they are; no states are kept, so no gradient may follow.
 * Records equal pyfwi_forward's bit for bit. */
enum { PYFWI_SCRATCH_FLOATS = 16777216 };
int32_t pyfwi_forward_noalloc(pyfwi_plan *plan, const pyfwi_shot *shots, int32_t nshots,
                              const float *src, float *seis);

/* Adjoint-state gradient for the shots of the last pyfwi_forward call
 * (__adjoint_modelling_per_source + __kernel_gradient,
 * wave_propagation.py:735-856).
 *   res   residual records, same layout and offsets as `seis`
 *   grad  output [6][tnz*tnx] summed over shots:
 *         g_mu, g_mu_precond, g_lam, g_lam_precond, g_rho, g_rho_precond
 *         (the buffers the reference reads back, wave_propagation.py:286-297)
 *         for PYFWI_MODE_PYFWI; for PYFWI_MODE_EXACT slots 0, 2, 4 hold
 *         dJ/dmu, dJ/dlam, dJ/drho (rho through b = 1/rho) and slots 1, 3, 5
 *         are zero.
 *   mode  PYFWI_MODE_PYFWI or PYFWI_MODE_EXACT
 *   ckpt_every  exact mode: checkpoint interval in steps; when the last
 *         forward kept no checkpoints at this interval the forward is run
 *         again here, pass by pass (0: the forward's interval when it kept
 *         checkpoints, else ceil(sqrt(nt))).  The shots run in passes sized
 *         to the device memory (unless pyfwi_plan_set_tuning fixes them):
 *         per shot, the checkpoints (ceil(nt/K) states, when recomputed),
 *         the K-state segment recomputed from each checkpoint, and the
 *         forward and adjoint states -- the two-level schedule, one extra
 *         forward sweep.  PYFWI_ENOMEM when one shot does not fit.  When the
 *         last forward kept every step of these shots (checkpoint list
 *         0..nt-1) and their sources lie on active cells, the exact adjoint
 *         reads those states instead and ckpt_every is not used: no
 *         recomputation, the same gradient bit for bit.
 *   flags PYFWI_F_PRECOND (pyfwi mode) */
int32_t pyfwi_gradient(pyfwi_plan *plan, const pyfwi_shot *shots,
                       int32_t nshots, const float *src, const float *res,
                       float *grad, int32_t mode, int32_t ckpt_every,
                       int32_t flags, int32_t mem);

/* L2 misfit and its reference-semantics gradient in one call, the sequence
 * of FWI.fprime (src/PyFWI/fwi.py:229-252): per pass of shots, forward
 * modelling, residual res = seis - obs with misfit 0.5*sum(res^2)
 * (CostFunction.l2, src/PyFWI/fwi_tools.py:1402-1422), then the adjoint of
 * __kernel_gradient (wave_propagation.py:751-856) with the checkpoint list
 * `chp` (as in pyfwi_forward: shots may differ in nt only when chp is every
 * step).  When chp is every step 0..nt-1
 * (chpr = 100) the forward states are kept in HBM for the pass and read
 * directly by the adjoint, so no time reversal is needed; a shot then keeps
 * nt_s states (5 fields each), and the shots, by decreasing nt, are packed
 * first-fit into passes whose kept states fill 90 % of the free device
 * memory (pyfwi_last_order gives the order).  With another checkpoint list
 * each shot keeps nchp states and the passes take as many shots, in the
 * order given, as that memory holds.  With the plan's shots_per_pass set,
 * passes of that many shots.  PYFWI_ENOMEM when not even one shot fits.
 *   obs        observed records, layout of `seis` (rec_offset per shot)
 *   seis       optional output of the modelled records (may be NULL)
 *   grad       output [6][tnz*tnx] as pyfwi_gradient
 *   misfit     output scalar
 *   comp_mask  bit f selects component f (vx, vz, txx, tzz, txz); 0 compares
 *              the pressure (txx + tzz)/2 with obs[txx] (components = 0) */
int32_t pyfwi_misfit_gradient(pyfwi_plan *plan, const pyfwi_shot *shots, int32_t nshots,
                              const float *src, const float *obs, float *seis, float *grad,
                              double *misfit, const int32_t *chp, int32_t nchp,
                              int32_t comp_mask, int32_t flags, int32_t mem);

/* Record filter of pyfwi_misfit_gradient: the multiscale lowpass of
 * cost_seismic (src/PyFWI/fwi_tools.py:1290-1345, lowpass / adj_lowpass
 * :1189-1255).  The reference pads each trace of n samples with `pad` zeros
 * (512), multiplies the N-point spectrum (N = n + pad) by the Butterworth
 * response H and by conj(H), and keeps n samples; that is the circular
 * convolution y[t] = sum_{|m| <= L} k[|m|] x[(t - m) mod N] with the
 * symmetric kernel k = ifft(|H|^2), whose taps k[0..half_len] the caller
 * computes (pyfwi_amd.fwi_tools.lowpass_taps).  With a filter set,
 * pyfwi_misfit_gradient takes the misfit of the filtered records,
 * K seis - K obs, computed as K (seis - obs) since K is linear, and injects
 * K applied to that filtered residual (the filter is its own adjoint); the
 * `seis` output stays unfiltered.  half_len < 0 removes the filter; half_len at most
 * PYFWI_FIR_MAX_HALF. */
enum { PYFWI_FIR_MAX_HALF = 6000 };
int32_t pyfwi_set_record_filter(pyfwi_plan *plan, const double *taps, int32_t half_len,
                                int32_t pad);

/* Execution shape of later calls (no counterpart in the reference, which runs
 * one shot at a time): shots_per_pass > 0 runs a call's shots in passes of
 * that many shots (0: all in one pass; smaller passes need less memory and
 * can keep the working set in the 256 MB on-package cache); `fused` is a
 * set of bits: PYFWI_TUNE_FUSED_STEPS does each forward and time-reversed
 * step as one launch instead of two half-step launches (the forward step
 * reads one state and writes the next: ping-pong buffers, or from one kept
 * slot into the next; records and kept states bit for bit those of two
 * launches); PYFWI_TUNE_FWD_TWO runs the kept-state forward two steps a
 * launch where the same shots are live at both (order 4; bit for bit those
 * of single steps; slower than single one-launch steps at the grids
 * measured);
 * PYFWI_TUNE_ADJ_FUSED runs the reference adjoint of each backward step as
 * one launch instead of two (stresses, then velocities; order 4 with
 * receivers on the active cells, else ignored); strip_rows > 0 fixes the
 * rows each thread walks (0: chosen per launch).  Results do not depend on
 * these choices beyond rounding. */
enum { PYFWI_TUNE_FUSED_STEPS = 1, PYFWI_TUNE_ADJ_FUSED = 2, PYFWI_TUNE_FWD_TWO = 4 };
int32_t pyfwi_plan_set_tuning(pyfwi_plan *plan, int32_t shots_per_pass, int32_t fused,
                              int32_t strip_rows);

/* hipGraph replay (no counterpart in the reference, which launches every
 * kernel from Python): enable = 1 captures each pass's forward and backward
 * time loop into a graph the first time it runs with a given configuration
 * (shot table, buffers, model, checkpoint list, flags) and replays it with
 * one hipGraphLaunch on later calls with the same configuration, instead of
 * submitting two to four launches per time step from the host.  Off by
 * default; not used while timing is enabled or with fused steps.  Results do
 * not depend on it.  enable = 0 releases the graphs. */
int32_t pyfwi_plan_set_graphs(pyfwi_plan *plan, int32_t enable);

/* Time loops captured into graphs and replayed since the plan was created. */
int32_t pyfwi_graph_stats(pyfwi_plan *plan, int64_t *captures, int64_t *replays);

/* Copy kept snapshot `slot` of shot `shot` ([5][tnz*tnx]: vx, vz, txx, tzz,
 * txz) to `out` (the reference's WavePropagator.W, wave_propagation.py:150-156,
 * 710-727). */
int32_t pyfwi_get_snapshot(pyfwi_plan *plan, int32_t shot, int32_t slot,
                           float *out, int32_t mem);

/* Shots per pass of the last pyfwi_forward, pyfwi_gradient or
 * pyfwi_misfit_gradient call, in the order they ran (no counterpart in the
 * reference, which runs one shot at a time).  The passes take consecutive
 * runs of the order pyfwi_last_order gives.  `sizes` receives up to `cap`
 * entries, `count` the number of passes, `ckpt_every` (may be NULL) the
 * checkpoint interval of an exact-mode gradient or of the forward's kept
 * checkpoints (0: none). */
int32_t pyfwi_last_passes(pyfwi_plan *plan, int32_t *sizes, int32_t cap, int32_t *count,
                          int32_t *ckpt_every);

/* The caller's shot indices (0-based positions in the `shots` argument) of
 * the last call, in the order its passes ran them (no counterpart in the
 * reference).  pyfwi_forward and pyfwi_gradient run the shots in the order
 * given.  pyfwi_misfit_gradient with every step kept orders them by
 * decreasing nt and packs them into passes first-fit (each shot joins the
 * first pass whose kept states have room for its nt states), a pass's shots
 * again by decreasing nt; with a tuned pass size (pyfwi_plan_set_tuning),
 * consecutive runs of that order.  `order` receives up to `cap` entries,
 * `count` the number of shots. */
int32_t pyfwi_last_order(pyfwi_plan *plan, int32_t *order, int32_t cap, int32_t *count);

/* Wait for all work queued on the plan's stream. */
int32_t pyfwi_synchronize(pyfwi_plan *plan);

/* Timing with HIP events on the plan's stream.  enable = 1: every stencil
 * launch is bracketed by an event pair.  A recorded event stalls the stream
 * for several microseconds, part of which falls inside the bracket, so these
 * durations serve to split a loop's time between its kernels.  enable = 2:
 * only the forward and backward time loops of each pass are bracketed
 * (PYFWI_K_FWD_LOOP, PYFWI_K_BWD_LOOP, counted in time steps), and the kernels
 * inside run as in an untimed call.  pyfwi_kernel_timing returns the summed
 * duration and count of one kind since the last enable.  Kinds: see PYFWI_K_*
 * below. */
enum {
    PYFWI_K_FWD_VEL = 0,   /* velocity half step, forward             */
    PYFWI_K_FWD_STR = 1,   /* stress half step, forward               */
    PYFWI_K_REV_STR = 2,   /* stress half step, time reversed         */
    PYFWI_K_REV_VEL = 3,   /* velocity half step, time reversed       */
    PYFWI_K_ADJ_STR = 4,   /* adjoint stresses                        */
    PYFWI_K_ADJ_VEL = 5,   /* adjoint velocities + imaging condition  */
    PYFWI_K_EXA_STR = 6,   /* exact adjoint of the stress half step   */
    PYFWI_K_EXA_VEL = 7,   /* exact adjoint of the velocity half step */
    PYFWI_K_FWD_STEP = 8,  /* fused forward step (velocities + stresses) */
    PYFWI_K_REV_STEP = 9,  /* fused time-reversed step                   */
    PYFWI_K_FWD_LOOP = 10, /* forward time loop of a pass (enable = 2)   */
    PYFWI_K_BWD_LOOP = 11, /* backward time loop of a pass (enable = 2)  */
    PYFWI_K_FWD_SPARSE = 12, /* record gather / source injection launches  */
    PYFWI_K_BWD_SPARSE = 13, /* residual injection launches                */
    PYFWI_K_ADJ_STEP = 14,   /* adjoint step in one launch (stresses, velocities,
                                imaging condition)                         */
    PYFWI_K_FWD_TWO = 15,    /* two forward steps in one launch            */
    PYFWI_K_COUNT = 16
};
int32_t pyfwi_plan_set_timing(pyfwi_plan *plan, int32_t enable);
int32_t pyfwi_kernel_timing(pyfwi_plan *plan, int32_t kind, double *total_ms,
                            int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* PYFWI_HIP_H */
