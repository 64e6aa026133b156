/*
 * swit_fd.h -- C ABI of the MI355X (gfx950) 2-D acoustic forward/adjoint solver.
 *
 * The reference has no in-process FFI: its boundary is the process call
 *     mpirun -np N fd2dmpi par=<parfile>          (toolbox/solver.py:54,79)
 * whose jobs are forward_modeling (fd2dmpi/solver.f90:36) and adjoint_modeling
 * (fd2dmpi/solver.f90:120).  The entry points below replace that process boundary
 * one-for-one; the Python drop-in (swit_amd/solver.py) binds them with ctypes.
 *
 * Conventions
 *   - Every array pointer is DEVICE memory unless stated otherwise; the caller owns
 *     inputs and outputs, the plan owns its scratch (wavefields, strips).
 *   - 2-D model/gradient arrays are [nx][nz] with iz fastest: the byte layout of the
 *     reference's vel.bin / rho.bin / src0_kernel_vp.bin (fd2dmpi/io.f90:231-264,
 *     :590-606) and of numpy (nx, nz).
 *   - Return 0 on success; a nonzero SWIT_E* code otherwise, with a message in
 *     swit_last_error() (thread-local).
 *   - One plan per stream; calls on one plan are serialised by the caller.
 */
#ifndef SWIT_FD_H
#define SWIT_FD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define SWIT_OK 0
#define SWIT_EINVAL 1       /* bad grid / geometry / argument                       */
#define SWIT_ENOMEM 2       /* device allocation failed                             */
#define SWIT_EHIP 3         /* a HIP runtime call failed                            */
#define SWIT_ESTATE 4       /* adjoint without a matching strip-saving forward      */

/* Plan flags. */
#define SWIT_PATH_PLAIN 0x0     /* per-shot, per-step kernel launches               */
#define SWIT_PATH_BATCHED 0x1   /* all shots per launch, fused steps, graph-captured */
#define SWIT_STRIPS_F16 0x10    /* boundary strips stored as FP16 with per-step scale */
#define SWIT_STRIPS_BF16 0x20   /* boundary strips stored as BF16 with per-step scale */
#define SWIT_PLAN_EXACT 0x40   /* also allocate the exact adjoint: divergence history of
                                   ns * nt * (nx+2npml) * ld floats */
#define SWIT_PLAN_EXACT_CKPT 0x80 /* exact adjoint with checkpoints instead: full states
                                   every c = round(sqrt(nt)) steps and one c-step segment
                                   of history, about 4 sqrt(nt) * ns * (nx+2npml) * ld
                                   floats; the reverse sweep recomputes each segment's
                                   forward (same gradient, bit for bit) */
#define SWIT_PLAN_NO_STRIPS 0x100 /* allocate no boundary strips: for plans that only model
                                   (observed data, line-search forwards, the exact
                                   adjoint); a strip-saving forward and swit_adjoint are
                                   refused, the autotuner tunes the forward only */

/* Grid and time axis: the reference's param type (fd2dmpi/datatype.f90:17-22) as
 * read from the parfile (io.f90:72-96).  dx and dt are the float32 values the
 * reference reads (the %f text of toolbox/solver.py:145-149). */
typedef struct swit_grid {
    int32_t nx, nz;         /* interior grid (NX, NZ)                                */
    int32_t npml;           /* absorbing-layer width (NPML, default 40)              */
    int32_t nt;             /* time samples (NT_WORK)                                */
    int32_t free_surface;   /* FREESURFACE; the reference forces the top row anyway  */
    float dx, dt;           /* DX, DT_WORK                                           */
} swit_grid;

/* Acquisition: the reference's acquisition type (datatype.f90:24-30) as read from
 * coord.txt (io.f90:126-228).  HOST pointers, copied at plan creation. */
typedef struct swit_shots {
    int32_t ns;             /* number of shots                                       */
    int32_t ngmax;          /* max receivers per shot                                */
    const int32_t *ng;      /* [ns] receivers per shot                               */
    const float *src_xz;    /* [ns][2] source x, z in metres                         */
    const float *rec_xz;    /* [ns][ngmax][2] receiver x, z in metres                */
} swit_shots;

typedef struct swit_plan swit_plan;

/* Allocates the plan: wavefields, strips, accumulators for all shots.
 * Replaces the setup half of forward_modeling/adjoint_modeling
 * (fd2dmpi/solver.f90:47-76 and :131-178): grid extents (pml.f90:24-38), coordinate
 * to cell mapping (fdcore.f90:70-72,124-126,216-218,350-352), free-surface row
 * (io.f90:777-800). */
int swit_plan_create(const swit_grid *grid, const swit_shots *shots, int flags,
                     swit_plan **out);

/* Loads vp/rho ([nx][nz] float32, device) into the plan: PML padding
 * (io.f90:231-300), vmin (math.f90:342-361, solver.f90:73), damping profile
 * (pml.f90:74-101), and the per-cell coefficients den*c^2*dt/dx and dt/(den*dx)
 * (fdcore.f90:79,99,108).  Invalidates saved strips.  Refuses (SWIT_EINVAL, no model
 * loaded) a model with a vp or rho value that is not finite and positive, or whose
 * largest vp breaks the 4th-order stability limit dt < sqrt(3/8) dx / max(vp) that
 * toolbox/base.py:57-59 enforces; the check reads two values back, so the call waits
 * for `stream`. */
int swit_plan_set_model(swit_plan *plan, const float *vp, const float *rho,
                        hipStream_t stream);

/* Forward job for every shot of the plan: replaces forward_modeling
 * (solver.f90:36-116) / staggered42_modeling_is (fdcore.f90:30-198).
 *   src:  [ns][nt] integrated source traces (the reference's src%d.bin).
 *   seis: [ns][ngmax][nt] pressure seismograms (the SU trace data of src%d_sg.su),
 *         sample 0 is zero as in the reference; traces past a shot's own receiver
 *         count are zero (dead traces).
 *   save_strips: 1 keeps boundary strips and the final state on the device for a
 *         following swit_adjoint (the reference's STORE_SNAP=1 files,
 *         fdcore.f90:130-176); 0 invalidates them (the reference's cleandata of
 *         data/syn/); 2 saves nothing and leaves earlier saved strips valid (an
 *         observed-data run, whose files the reference writes under data/obs/). */
int swit_forward(swit_plan *plan, const float *src, float *seis, int save_strips,
                 hipStream_t stream);

/* Adjoint job for every shot: replaces adjoint_modeling (solver.f90:120-289):
 * reverse-time reconstruction (fdcore.f90:202-289), adjoint propagation
 * (fdcore.f90:293-375), zero-lag imaging (solver.f90:237-239), cross-shot sum
 * (solver.f90:263-265, here FP64 on the device) and g = -2 g / c (solver.f90:270).
 *   src:    [ns][nt] the same integrated source traces as the forward.
 *   adjsrc: [ns][ngmax][nt] adjoint sources (the reference's adjoint src%d.bin).
 *   grad, illum_forw, illum_back: [nx][nz] float64 outputs (src0_kernel_vp.bin,
 *         src0_illum_forw.bin, src0_illum_back.bin). */
int swit_adjoint(swit_plan *plan, const float *src, const float *adjsrc, double *grad,
                 double *illum_forw, double *illum_back, hipStream_t stream);

/* Releases every device buffer of the plan. */
int swit_plan_destroy(swit_plan *plan);

/* Message of the last failure on this thread ("" if none). */
const char *swit_last_error(void);

/* Plan options (no counterpart in the reference: tuning and measurement of the
 * batched path). */
#define SWIT_OPT_GRAPH 1        /* 1 (default): replay the time loop as a hipGraph    */
#define SWIT_OPT_PROFILE 2      /* 1: time each step-kernel time loop with HIP events */
#define SWIT_OPT_FWD_GROUP 4    /* G: the forward step kernel's workgroups take
                                   ceil(ns/G) near-equal groups of at most G shots     */
#define SWIT_OPT_FWD_COLGROUP 8 /* tile columns per group of the forward work order   */
#define SWIT_OPT_BWD_COLGROUP 16 /* tile columns per group of the backward work order */
#define SWIT_OPT_BWD_STEPS 32   /* 2 (default): backward kernel doing two time steps per
                                   launch; 1: one step per launch (same results)      */
#define SWIT_OPT_EXACT_FUSED 64 /* 1 (default): one fused kernel per exact-adjoint step;
                                   0: the seven separate passes (same results when no
                                   two receivers share a cell)                        */
#define SWIT_OPT_BWD_TILE_X 128 /* tile width of the one-step backward kernel: 24
                                   (default) or 16 columns (same results)             */
#define SWIT_OPT_STREAMS 256    /* 1 (default); 2: the first and second half of the shots
                                   run their forward and reverse time loops on two
                                   streams (the caller's and the plan's own), forked
                                   from and joined back into the caller's stream, so
                                   that each step's two launches overlap. Results are
                                   bit-identical to one stream.                        */
#define SWIT_OPT_BWD_BALANCE 512 /* split of the one-step backward kernel's work items
                                   over the 8 XCDs: 1 (default) contiguous ranges of
                                   equal estimated bytes moved; 0 equal item counts
                                   (same results)                                      */
int swit_plan_set_option(swit_plan *plan, int option, int value);

/* Autotuner (no counterpart in the reference): times `steps` back-to-back launches of
 * each batched step kernel (best of three bursts) on this plan's grid for every candidate forward
 * shot group size (1..8, 10, 12) x column-group order (forward with strip saving) and
 * every backward kernel (two steps per launch, or one step per launch at each tile
 * width and XCD work split) x column-group order, and keeps the fastest per time step.
 * Call between jobs with the model set; it uses the wavefield and strip buffers as
 * scratch, so saved strips (and an exact-adjoint history) are invalidated.
 * choice[6] = {forward shots per workgroup, forward column group, backward column
 * group, backward steps per launch, one-step backward tile width, one-step backward
 * XCD work split}; us[2] = the chosen forward / backward
 * step-kernel time per time step in microseconds. */
int swit_plan_autotune(swit_plan *plan, int steps, int *choice, double *us, hipStream_t stream);

/* Accumulated duration and time-step count of the step kernels timed while
 * SWIT_OPT_PROFILE was on: which = 0 forward step kernel, 1 backward step kernel
 * (a two-step backward launch counts two steps).
 * Each time loop is bracketed by one pair of events (the step kernels run back to
 * back in between); this call waits for the loops recorded so far to finish. */
int swit_plan_kernel_time(swit_plan *plan, int which, double *total_ms, int64_t *launches);

/* The saved boundary strips of the last strip-saving forward, decoded to FP32:
 * out[ns][nt][S] (device), S = 9nx+6nz, row it-1 = step it, the layout of the
 * reference's boundary.bin per shot (fdcore.f90:130-159,166-176).
 * SWIT_ESTATE if no strips are saved. */
int swit_plan_strips(swit_plan *plan, float *out, hipStream_t stream);

/* ---- exact adjoint (no counterpart in the reference) ----------------------------------
 * The transpose of the forward scheme itself (the reference's adjoint, swit_adjoint,
 * uses a different discretisation and is not the exact gradient).  For a misfit J of
 * the seismograms, swit_adjoint_exact turns dJ/dseis into dJ/dvp through
 * K = den c^2 dt/dx and the PML padding; density and the damping profile are held
 * fixed.  Needs a batched plan created with SWIT_PLAN_EXACT or SWIT_PLAN_EXACT_CKPT.
 *   swit_forward_exact: as swit_forward (same seismograms, bit for bit), also storing
 *     the divergence term of every pressure update (no strips are saved);
 *   swit_adjoint_exact: gseis [ns][ngmax][nt] = dJ/dseis; grad [nx][nz] float64. */
int swit_forward_exact(swit_plan *plan, const float *src, float *seis, hipStream_t stream);
int swit_adjoint_exact(swit_plan *plan, const float *gseis, double *grad, hipStream_t stream);

/* ---- misfits and adjoint sources (toolbox/misfit.py) -------------------------------
 * obs, syn, adj: [ntraces][nt] device arrays (a plan's [ns][ngmax][nt] seismograms
 * flattened); work: [ntraces] doubles of device scratch (per-trace misfits on return);
 * misfit: one double on the device.  A trace whose observed or synthetic norm is zero
 * contributes nothing and gets a zero adjoint source (misfit.py:128-148 and peers).
 *   SWIT_MISFIT_WAVEFORM: adj = syn - obs, misfit = sum sqrt(sum r^2 dt)
 *     (misfit.py:128-148, :263-282);
 *   SWIT_MISFIT_GLOBAL_CORRELATION: o = obs/|obs|, s = syn/|syn|, c = Pearson(s, o),
 *     misfit = -sum c, adj = (s c - o) / |syn| (misfit.py:219-242, :338-355);
 *   SWIT_MISFIT_ENVELOPE: power-2 envelope difference with the Hilbert transform of
 *     scipy.signal.hilbert, misfit = dt sum r^2 (misfit.py:151-185, :286-308);
 *   SWIT_MISFIT_TRAVELTIME: cross-correlation lag from the FFT correlation maximum,
 *     misfit = 0.5 lag^2 dt, adj = scaled central difference of syn
 *     (misfit.py:188-215, :312-334, :359-373); traces whose difference is all zero
 *     are dead too.  These two take nt <= 9216 and run their transforms in FP64.
 * Sums run in FP64; the sum over traces has a fixed order (reproducible). */
#define SWIT_MISFIT_WAVEFORM 0
#define SWIT_MISFIT_GLOBAL_CORRELATION 1
#define SWIT_MISFIT_ENVELOPE 2
#define SWIT_MISFIT_TRAVELTIME 3
int swit_misfit(int kind, int ntraces, int nt, float dt, const float *obs, const float *syn, float *adj,
                double *work, double *misfit, hipStream_t stream);

/* ---- plan groups: uneven grids in one launch ------------------------------------
 * A group runs several batched plans (different grids, shot counts, record lengths,
 * strip precisions) as one job: each group forward / adjoint is a single hipGraph
 * launch holding every plan's whole time loop on its own branch, so plans with short
 * records fill the GPU while long ones finish.  No counterpart in the reference (one
 * fd2dmpi job per grid, toolbox/solver.py:24-92); results are those of
 * swit_forward / swit_adjoint on each plan alone.  The arrays below hold one entry
 * per plan, in the order of `plans`; each entry is what the single-plan call takes.
 * Plans must not be used by other calls while a group call on them is in flight, and
 * must outlive the group. */
typedef struct swit_group swit_group;
int swit_group_create(swit_plan *const *plans, int nplans, swit_group **out);
int swit_group_forward(swit_group *group, const float *const *src, float *const *seis, int save_strips,
                       hipStream_t stream);
int swit_group_adjoint(swit_group *group, const float *const *src, const float *const *adjsrc,
                       double *const *grad, double *const *illum_forw, double *const *illum_back,
                       hipStream_t stream);
int swit_group_destroy(swit_group *group);

/* A whole gradient of every plan in one graph launch: each plan's branch runs the
 * strip-saving forward (seismograms to seis), the misfit and adjoint source `kind`
 * (SWIT_MISFIT_*) against obs, exactly as swit_misfit computes them (adjoint source to
 * adj, per-trace misfits to work, their sum to misfit), and the adjoint (grad,
 * illum_forw, illum_back).  Each plan's results equal swit_forward, swit_misfit and
 * swit_adjoint run one after the other.  The branches overlap, so one plan's adjoint
 * runs beside another's forward.  Per plan: obs, seis, adj [ns][ngmax][nt]; work
 * [ns*ngmax] doubles; misfit one double; grad, illum_* [nx][nz] doubles.  No counterpart in the
 * reference (toolbox/inversion.py:48-60 runs these steps one after the other, through files). */
int swit_group_gradient(swit_group *group, int kind, const float *const *src, const float *const *obs,
                        float *const *seis, float *const *adj, double *const *work, double *const *misfit,
                        double *const *grad, double *const *illum_forw, double *const *illum_back,
                        hipStream_t stream);

/* Bytes of device scratch the plan holds (wavefields, strips, accumulators). */
int64_t swit_plan_scratch_bytes(const swit_plan *plan);

/* Generation of the divergence history an exact plan holds: a number that each
 * swit_forward_exact raises, or 0 while no valid history is held (none yet, or a new
 * model or the autotuner has invalidated it).  A caller that keeps a forward's
 * generation and compares it before swit_adjoint_exact knows that the history still
 * belongs to that forward (swit_amd.autograd does).  No counterpart in the reference. */
int64_t swit_plan_exact_generation(const swit_plan *plan);

/* ---- gradient preconditioning (toolbox/postprocess.py:20-111) --------------------- */

typedef struct swit_precond {
    int nx, nz;
    int marine;         /* 1: 'Marine'/'Offshore' (taper zeroes rows above grad_mute, smoothing
                           below it); 0: land (smoothed Hamming damping, smoothing everywhere) */
    int grad_mute;      /* optim.grad_mute: taper size in rows (0: no taper) */
    float grad_thred;   /* optim.grad_thred (land taper floor) */
    int grad_smooth;    /* optim.grad_smooth: gradient smoothing span (0: none) */
    double vpmax;       /* optim.vpmax: the scaled gradient's maximum magnitude */
} swit_precond;

/* Bytes of device workspace swit_grad_precond needs for these parameters. */
int swit_grad_precond_work(const swit_precond *p, size_t *bytes);

/* grad_precond (postprocess.py:20-86) on the device, FP64 as the reference's numpy:
 * grad, forw, back: float32 [nx][nz] (the solver's kernel and illuminations, as the
 * reference reads them from its .bin files); mask: float64 [nx][nz] or NULL;
 * out: float64 [nx][nz]; work: device memory of swit_grad_precond_work bytes.
 * Smoothing (tools.py:255-275) runs as two exact-arithmetic-equivalent 1-D passes, so
 * results equal the numpy restatement to FP64 rounding. */
int swit_grad_precond(const swit_precond *p, const float *grad, const float *forw, const float *back,
                      const double *mask, double *out, void *work, size_t work_bytes, hipStream_t stream);

/* Size of the device-resident workspace of swit_grad_precond_noalloc: 64 MiB, enough for
 * models of up to about 2.09 million cells (swit_grad_precond_work reports the need). */
#define SWIT_PRECOND_SCRATCH_BYTES ((size_t)64 << 20)

/* swit_grad_precond without a workspace argument and without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight loops, harnesses
 * that forbid hipMalloc).  It has no counterpart in the reference.  The workspace is a
 * statically allocated device array of SWIT_PRECOND_SCRATCH_BYTES that each device holds
 * for the life of the process; its address is resolved once per device, on that device's
 * first call.  Arguments and results are swit_grad_precond's.  The stream must belong to
 * the current device (NULL: its default stream).  Returns SWIT_EINVAL when
 * swit_grad_precond_work exceeds SWIT_PRECOND_SCRATCH_BYTES.  Calls on one device share
 * the workspace: calls that overlap in time on one device (two streams, two threads)
 * race, and must be serialised by the caller. */
int swit_grad_precond_noalloc(const swit_precond *p, const float *grad, const float *forw, const float *back,
                              const double *mask, double *out, hipStream_t stream);

/* ---- shot-gather processing (toolbox/preprocess.py:23-200) ----------------------------
 * process_workflow_serial's steps on device gathers, in its order: offset mutes and the
 * first-break window mute (apply_mute, :98-151), Butterworth filtfilt followed by
 * demean, linear detrend and a 5 % Hann taper (apply_filter, :64-95), trace and event
 * normalisation (apply_normalize, :154-200).  Data stay float32 between steps as the
 * reference's obspy traces do; the filter runs in FP64 with scipy's recurrence order. */
#define SWIT_PROCESS_MAX_TAPS 16
#define SWIT_NORM_NONE 0
#define SWIT_NORM_L1 1
#define SWIT_NORM_L2 2
#define SWIT_NORM_MAX 3          /* trace normalisation only */

typedef struct swit_process_params {
    int ngathers, ntr, nt;       /* data [ngathers][ntr][nt]; an event is one gather */
    double dt;                   /* sample interval, s */
    int mute_offset_short;       /* zero traces with |offset| < mute_offset_short_dis */
    int mute_offset_long;        /* zero traces with |offset| > mute_offset_long_dis */
    double mute_offset_short_dis, mute_offset_long_dis;   /* m */
    int mute_late_arrival;       /* window mute around the first break */
    double mute_late_window;     /* s; > 0 mutes arrivals later than pick + window,
                                    otherwise earlier ones (preprocess.py:121-125) */
    int filter_taps;             /* len(b) = len(a), 2..16; 0: no filter */
    double b[SWIT_PROCESS_MAX_TAPS], a[SWIT_PROCESS_MAX_TAPS];  /* a[0] = 1 */
    double zi[SWIT_PROCESS_MAX_TAPS];   /* scipy.signal.lfilter_zi(b, a) */
    int detrend_taper;           /* demean, linear detrend and taper (after the filter) */
    int norm_trace;              /* SWIT_NORM_* */
    int norm_event;              /* SWIT_NORM_NONE, _L1 or _L2 */
} swit_process_params;

/* Bytes of device workspace swit_process needs for these parameters. */
int swit_process_work(const swit_process_params *p, size_t *bytes);

/* Processes data [ngathers][ntr][nt] (device, float32) in place.
 *   offsets: device float [ngathers*ntr], group x - source x per trace (the SU
 *            header's offset; required with an offset mute);
 *   mute_win: device double [100], sin(linspace(0, pi, 200))[:100] as numpy computes
 *            it (required with the window mute);
 *   taper:   device double [nt], obspy's Trace.taper(0.05, 'hann') weights
 *            (required with detrend_taper);
 *   picks:   device int32 [ngathers*ntr] or NULL: the window mute's centre per trace in
 *            samples, int((mute_time_ms / 1000 + |mute_late_window|) / dt) from the
 *            traces' SU header first breaks (preprocess.py:134-139); NULL picks the
 *            first break from the data (the first sample above 0.1 % of the trace
 *            maximum, preprocess.py:140-142, 203-209) and shifts it by
 *            ceil(|mute_late_window| / dt);
 *   work:    device memory of swit_process_work bytes.
 * The filter needs nt > 3 * filter_taps (scipy's padding). */
int swit_process(const swit_process_params *p, const float *offsets, const double *mute_win, const double *taper,
                 const int32_t *picks, float *data, void *work, size_t work_bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SWIT_FD_H */
