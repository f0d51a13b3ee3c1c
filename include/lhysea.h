/*
 * lhysea.h -- C-ABI of the MI355X-native LandSlide-HySEA solver step.
 *
 * Plain C types only (pointers, sizes, doubles): no HIP or torch types cross
 * this boundary.  Every entry point returns an int status (LHYSEA_OK == 0)
 * and never throws; lhysea_last_error() gives a message for the calling thread.
 *
 * Reference interfaces this library replaces:
 *   lhysea_device_probe   <- extern "C" int comprobarSoporteCUDA()
 *                            (src/GPU/main.cxx:11, src/GPU/ComprobarSoporteCUDA.cu:5-25)
 *   lhysea_shallow_water  <- extern "C" int shallowWater(TDatoCluster*, ...)
 *                            (src/GPU/main.cxx:12-18, src/GPU/ShallowWater.cu:125-486),
 *                            minus the MPI fields (num_procs, id_hebra, num_voly_otros):
 *                            the whole grid lives on one GPU.
 *   lhysea_create/upload/step/download/destroy
 *                         <- the phases of shallowWater: device allocation and
 *                            upload (inicializarDatosCuda, src/GPU/ShallowWater.cu:14-110),
 *                            the step loop (:269-461), copies back (:274-275, :466)
 *                            and release (liberarSWCuda, :112-122).
 *   lhysea_initial_dt     <- the initial-dt pass (src/GPU/ShallowWater.cu:236-257)
 *   lhysea_eta_max        <- d_eta1_maxima download (src/GPU/ShallowWater.cu:466)
 *   lhysea_set_gauges / lhysea_get_gauges
 *                         <- the gauge time series (src/GPU/ShallowWater.cu:318-331)
 *
 * Threading: a handle owns its device memory and one HIP stream; handles are
 * not thread-safe but may be used from different threads one after another.
 */
#ifndef LHYSEA_H
#define LHYSEA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; 1 and 2 keep shallowWater's meaning (src/GPU/ShallowWater.cu:124) */
#define LHYSEA_OK 0
#define LHYSEA_ERR_GPU_MEMORY 1
#define LHYSEA_ERR_HOST_MEMORY 2
#define LHYSEA_ERR_ARGUMENT 3
#define LHYSEA_ERR_HIP 4
#define LHYSEA_ERR_NO_DEVICE 5
#define LHYSEA_ERR_IO 6
#define LHYSEA_ERR_NUMERIC 7   /* the state went bad: see lhysea_health */

/* health bits of a state (lhysea_health, lhysea_batch_health) */
#define LHYSEA_HEALTH_NONFINITE 1       /* a cell update produced a NaN or infinity */
#define LHYSEA_HEALTH_NEGATIVE_DEPTH 2  /* a layer depth below -eps_h after the flux
                                           update, before the clip to zero */

/* arithmetic of the step */
#define LHYSEA_FP64 0   /* plain FP64 */
#define LHYSEA_FP32C 1  /* FP32 Riemann solver and fluxes, compensated FP32 state,
                           FP64 depth, hydrostatic terms, dt and time */

/* step kernels (FP64 only offers both) */
#define LHYSEA_KERNEL_FUSED 1     /* one tiled launch per step (default) */
#define LHYSEA_KERNEL_FOURPASS 0  /* one launch per edge class, as the reference */

/* Problem parameters in the reference's internal units (after the optional
 * normalisation of src/GPU/Problema.cxx:267-287).  Angles in radians. */
typedef struct lhysea_params {
    int32_t nx, ny;           /* cells in x (columns) and y (rows); row 0 is "superior" */
    int32_t friction_law;     /* 0 Coulomb (reference default), 1 Pouliquen */
    int32_t precision;        /* LHYSEA_FP64 or LHYSEA_FP32C */
    double dx, dy;            /* ancho_vol, alto_vol */
    double borde_sup, borde_inf, borde_izq, borde_der;  /* 1 open, 0 / -1 wall */
    double cfl, r;            /* CFL number, density ratio */
    double angle1, angle2, angle3, angle4;              /* friction angles (rad) */
    double mfc, mf0, mfs;     /* interlayer, water-bottom, sediment-bottom friction */
    double vmax1, vmax2;      /* speed caps (<= 0: off) */
    double g, eps_h, L, H;    /* gravity, dry threshold, length and depth scales */
    double peso, beta;        /* 1, 1 in the reference (src/GPU/main.cxx:161) */
} lhysea_params;

typedef struct lhysea_sim lhysea_sim;

/* 0 ok, 1 no GPU, 2 no usable GPU (comprobarSoporteCUDA's codes) */
int lhysea_device_probe(void);
const char *lhysea_last_error(void);

int lhysea_create(const lhysea_params *p, lhysea_sim **out);
int lhysea_destroy(lhysea_sim *s);

/* Host FP64 planes, each nx*ny row-major: W = h1,q1x,q1y,h2,q2x,q2y (6 planes
 * back to back), H = depth (positive down, already Hmin-shifted).  Upload also
 * resets time to 0, eta max to h1+h2-H (src/GPU/Problema.cxx:541-546) and the
 * step counter. */
int lhysea_upload(lhysea_sim *s, const double *W, const double *H);
int lhysea_download(lhysea_sim *s, double *W);

/* Initial dt from the wave-speed pass on the current state; becomes the dt of
 * the next step.  *dt may be NULL. */
int lhysea_initial_dt(lhysea_sim *s, double *dt);
/* Force the dt of the next step. */
int lhysea_set_dt(lhysea_sim *s, double dt);
/* Run nsteps steps asynchronously on the handle's stream (the host reads the
 * control block before each piece of at most 1024 steps, not after the last).
 * Returns LHYSEA_ERR_NUMERIC, launching nothing more, when the health word
 * (lhysea_health) shows a bad state at one of those reads. */
int lhysea_step(lhysea_sim *s, int nsteps);
/* Step while t < t_end (the reference's loop condition, src/GPU/ShallowWater.cu:269)
 * and at most max_steps steps; *steps_done (may be NULL) gets the steps run.
 * The fused kernels stop on the device, so the host waits once per chunk of
 * at most 512 steps, not once per step; a chunk that leaves the health word
 * set ends the run with LHYSEA_ERR_NUMERIC (*steps_done is still set). */
int lhysea_run_until(lhysea_sim *s, double t_end, int32_t max_steps, int32_t *steps_done);
/* Failure detection (no counterpart in the reference, which has no NaN
 * checks): every cell update of the step kernels checks its new state, and a
 * workgroup that finds a non-finite value or a depth below -eps_h raises the
 * corresponding LHYSEA_HEALTH_* bit in the handle's control block, with the
 * first step that raised one.  *flags gets the bits raised since the last
 * upload (0: healthy), *first_step (may be NULL) that step (-1 if none).
 * Waits for the handle's stream. */
int lhysea_health(lhysea_sim *s, uint32_t *flags, int32_t *first_step);
int lhysea_sync(lhysea_sim *s);
/* time at the current state, dt of the next step, completed steps */
int lhysea_get_time(lhysea_sim *s, double *t, double *dt_next, int32_t *steps);
/* Enable (1) or disable (0) the running max of eta = h1 - H
 * (src/GPU/Volumen_kernel.cu:182-213); val/time are nx*ny each. */
int lhysea_track_eta_max(lhysea_sim *s, int enable);
int lhysea_eta_max(lhysea_sim *s, double *val, double *time);
/* Gauges: eta = h1 + h2 - H sampled at flat cell indices after each step, for
 * at most max_steps steps counted from the call. */
int lhysea_set_gauges(lhysea_sim *s, const int32_t *cells, int32_t ng, int32_t max_steps);
int lhysea_get_gauges(lhysea_sim *s, double *out, int32_t *nrec);
/* dt used by each step since the last upload (at most max_steps kept). */
int lhysea_dt_history(lhysea_sim *s, double *out, int32_t max_steps, int32_t *nrec);
/* LHYSEA_KERNEL_FUSED or LHYSEA_KERNEL_FOURPASS */
int lhysea_set_kernel(lhysea_sim *s, int32_t mode);
/* Form of the FP32C adjoint's reverse step (no counterpart in the reference):
 * LHYSEA_KERNEL_FUSED (default; one tiled kernel after the recomputation) or
 * LHYSEA_KERNEL_FOURPASS (cell adjoint, then one launch per edge class with
 * the partial-sum adjoints in HBM).  Gradients agree to rounding; the
 * four-pass form exists as a cross-check.  The FP64 reverse step is four-pass
 * either way. */
int lhysea_set_adjoint_kernel(lhysea_sim *s, int32_t mode);
/* FP32C step by tile class (no counterpart in the reference; same results
 * either way).  The grid is cut into 62 x 14 class tiles; a tile whose cells
 * and 1-cell halo hold no sediment (h2 = 0, q2 = 0) takes a reduced kernel
 * (single layer in LDS and registers, twice the occupancy), the others the
 * general two-layer tile function.  mode -1 (default): the split launch is
 * used when at least 1000 tiles are sediment-free (its general launch costs
 * a fixed latency), chosen again every 1024 steps and every run_until
 * chunk; 0: every tile general; 1: always split.  lhysea_tile_classes
 * reports the class tiles, the general ones of the next step and whether
 * the last launches were split. */
int lhysea_set_tile_split(lhysea_sim *s, int32_t mode);
int lhysea_tile_classes(lhysea_sim *s, int32_t *tiles, int32_t *general, int32_t *split);
/* Time nsteps steps with HIP events on the handle's stream: total wall time
 * of the whole sequence and the average duration of the main step kernel
 * (bracketed launch by launch in a second pass when kernel_ms is not NULL). */
int lhysea_time_steps(lhysea_sim *s, int32_t nsteps, double *total_ms, double *kernel_ms);

/* Discrete adjoint (no counterpart in the reference: an extension of the
 * step for gradient-based inversion), as a forward/backward pair.
 *
 * lhysea_adjoint_forward runs nsteps forward steps from the freshly uploaded
 * state, records the dt sequence and the gauge series eta = h1 + h2 - H
 * after each step (eta: nsteps*ng doubles, row n = after step n+1; may be
 * NULL) and keeps the checkpoints the reverse sweep needs: nsnap state
 * snapshots in HBM placed by binomial checkpointing, or, with nsnap > nsteps,
 * every state (nsteps + 1 states, no step re-advanced).  The handle then
 * holds the workspace and refuses the calls that step or reconfigure it
 * (upload, initial_dt, set_dt, step, run_until, set_gauges, track_eta_max,
 * set_kernel, time_steps) until lhysea_adjoint_backward, which consumes it;
 * lhysea_download returns the state after step nsteps meanwhile.  A second
 * forward, or lhysea_destroy, discards a pending workspace.
 *
 * lhysea_adjoint_backward takes the seeds dL/deta (nsteps*ng doubles, the
 * layout of eta) of any loss L of the gauge series and runs the reverse sweep
 * (branches and the dt sequence frozen at the forward choices).  Outputs:
 * dL/dW0 (6 planes like lhysea_upload; may be NULL), dL/dmu of the Coulomb
 * coefficient mu = |tan(angle1)| L/H and dL/dangle1 (may be NULL).
 * stats (6 doubles, may be NULL): forward ms, reverse ms, re-advanced steps,
 * reverse steps, snapshots, optimal re-advance count.  Afterwards the handle
 * must be uploaded again before further steps.  Both friction laws: under the
 * Pouliquen law (friction_law 1) the coefficient depends on the state and on
 * all four angles, grad_mu must be NULL (LHYSEA_ERR_ARGUMENT otherwise, the
 * workspace kept) and grad_angle1 is dL/dangle1.
 *
 * lhysea_adjoint_backward_angles is the same sweep with the gradient of every
 * friction angle: grad_angles[i] = dL/dangle(i+1), i = 0..3 (Coulomb: only
 * angle1 enters, the other three are 0; Pouliquen: the four angles of
 * src/GPU/Arista_kernel.cu:20-58).
 *
 * lhysea_adjoint is both for the gauge-misfit loss
 *     L = sum_n sum_g (eta_g^n - obs[n*ng + g])^2,   seeds 2 (eta - obs). */
int lhysea_adjoint_forward(lhysea_sim *s, int32_t nsteps, int32_t nsnap, const int32_t *gauges, int32_t ng,
                           double *eta);
int lhysea_adjoint_backward(lhysea_sim *s, const double *seeds, double *grad_W0, double *grad_mu,
                            double *grad_angle1, double *stats);
int lhysea_adjoint_backward_angles(lhysea_sim *s, const double *seeds, double *grad_W0, double *grad_angles,
                                   double *stats);
int lhysea_adjoint(lhysea_sim *s, int32_t nsteps, int32_t nsnap, const int32_t *gauges, int32_t ng,
                   const double *obs, double *loss, double *grad_W0, double *grad_mu, double *grad_angle1,
                   double *stats);

/* ---------------------------------------------------------------- ensembles
 * A batch steps many independent FP32C scenarios (own grid, parameters, dt,
 * time and stop condition) with one batched launch per step, replayed from a
 * captured hipGraph.  No counterpart in the reference, which runs one
 * scenario per process (src/GPU/main.cxx:157-164); each scenario's results
 * equal those of a single lhysea_sim with the same inputs.  Scenario sid of a
 * batch follows the single-simulation calls above, with the same meaning:
 *   lhysea_batch_upload        <- lhysea_upload
 *   lhysea_batch_download      <- lhysea_download
 *   lhysea_batch_initial_dt    <- lhysea_initial_dt, for every scenario
 *   lhysea_batch_get_time      <- lhysea_get_time
 *   lhysea_batch_track_eta_max / lhysea_batch_eta_max <- lhysea_track_eta_max / lhysea_eta_max
 *   lhysea_batch_set_gauges / lhysea_batch_get_gauges <- lhysea_set_gauges / lhysea_get_gauges
 *   lhysea_batch_dt_history    <- lhysea_dt_history
 * Every p[i] must have precision LHYSEA_FP32C. */
typedef struct lhysea_batch lhysea_batch;

int lhysea_batch_create(const lhysea_params *p, int32_t nscen, lhysea_batch **out);
int lhysea_batch_destroy(lhysea_batch *b);
int lhysea_batch_size(const lhysea_batch *b, int32_t *nscen);
int lhysea_batch_upload(lhysea_batch *b, int32_t sid, const double *W, const double *H);
int lhysea_batch_download(lhysea_batch *b, int32_t sid, double *W);
int lhysea_batch_initial_dt(lhysea_batch *b);
/* Scenario sid steps while t < t_end (the reference's loop condition,
 * src/GPU/ShallowWater.cu:269) and steps < max_steps; INFINITY / -1: no limit. */
int lhysea_batch_set_stop(lhysea_batch *b, int32_t sid, double t_end, int32_t max_steps);
/* nsteps batch steps (graph replays of 16 steps plus direct launches), asynchronous */
int lhysea_batch_step(lhysea_batch *b, int32_t nsteps);
/* the same steps, every launch made directly (no graph) */
int lhysea_batch_step_eager(lhysea_batch *b, int32_t nsteps);
int lhysea_batch_sync(lhysea_batch *b);
int lhysea_batch_get_time(lhysea_batch *b, int32_t sid, double *t, double *dt_next, int32_t *steps);
/* lhysea_health of scenario sid (raised by the batch's step kernels). */
int lhysea_batch_health(lhysea_batch *b, int32_t sid, uint32_t *flags, int32_t *first_step);
int lhysea_batch_track_eta_max(lhysea_batch *b, int32_t sid, int enable);
int lhysea_batch_eta_max(lhysea_batch *b, int32_t sid, double *val, double *time);
int lhysea_batch_set_gauges(lhysea_batch *b, int32_t sid, const int32_t *cells, int32_t ng, int32_t max_steps);
int lhysea_batch_get_gauges(lhysea_batch *b, int32_t sid, double *out, int32_t *nrec);
int lhysea_batch_dt_history(lhysea_batch *b, int32_t sid, double *out, int32_t max_steps, int32_t *nrec);
/* Tile shapes of the batched kernel: shapes gets 3 ints (TX, TY, threads) per
 * shape (may be NULL), count the number of shapes. */
int lhysea_batch_tile_shapes(int32_t *shapes, int32_t *count);
int lhysea_batch_set_tile(lhysea_batch *b, int32_t tile);
int lhysea_batch_get_tile(const lhysea_batch *b, int32_t *tile);
/* HIP-event time of nsteps batch steps (they advance the state). */
int lhysea_batch_time_steps(lhysea_batch *b, int32_t nsteps, double *total_ms);
/* Time nsteps steps with every tile shape and keep the fastest; the state,
 * times and outputs are left as they were.  ms_per_tile (one per shape) and
 * best may be NULL. */
int lhysea_batch_autotune(lhysea_batch *b, int32_t nsteps, double *ms_per_tile, int32_t *best);

/* Steps without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The caller owns
 * every buffer, all in device memory: H (nx*ny doubles, Hmin-shifted depth),
 * W (the state: LHYSEA_FP64 6 planes of nx*ny doubles as in lhysea_upload;
 * LHYSEA_FP32C 12 planes of nx*ny floats, the 6 FP32 values then their 6
 * compensations) and W_work (a second buffer of W's layout).  nsteps fused
 * steps run on `stream` (a hipStream_t; NULL: the default stream) from time
 * *t with first dt *dt (<= 0: the initial-dt pass), as lhysea_step does on a
 * handle (same kernels, same results); on return *t and *dt are the time and
 * the next dt, and *result_in_work says where the final state is (0: W, 1:
 * W_work; the other buffer holds the state one step earlier).  No eta max,
 * gauges or dt history.  The library's own scratch (control block and tile
 * class tables) is a fixed per-GPU array: grids of at most 65536 class tiles
 * (FP64 tiles of 30 x 14 cells, FP32C 62 x 14; every grid up to 4096 x 4096
 * fits), larger ones are refused with LHYSEA_ERR_ARGUMENT.  Calls that
 * overlap on one GPU share that array and race; run them one after another.
 * The call waits for its stream once after the tile classification and once
 * at the end (it reads the control block back); it returns
 * LHYSEA_ERR_NUMERIC (outputs set) when a step raised a health bit
 * (lhysea_health). */
int lhysea_step_device(const lhysea_params *p, const double *H, void *W, void *W_work, int32_t nsteps,
                       double *t, double *dt, int32_t *result_in_work, void *stream);

/* One-shot driver: runs while t < t_end and returns the final state in W
 * (in place), eta max value/time (nx*ny each, may be NULL) and the loop wall
 * time in *elapsed_s; writes no files (lhysea_shallow_water_output below is
 * the form that writes shallowWater's output). */
int lhysea_shallow_water(const lhysea_params *p, double *W, const double *H, double t_end,
                         double *eta_max, double *eta_max_time, double *elapsed_s, int32_t *steps_done);

/* Output files of a run as shallowWater writes them (src/GPU/ShallowWater.cu:211-334,
 * 462-474; layout of src/GPU/netcdf.cu:29-411): gauges NULL gives six NetCDF
 * files <prefix>_{eta1,q1x,q1y,eta2,q2x,q2y}.nc (axes, bathymetry, one record
 * per write, max_height at close); otherwise the text file
 * <prefix>_eta_puntos.txt, one line per write (time, then eta1 at each gauge
 * cell; -1 entries, gauges outside the grid, print -999).  Values in physical
 * units: depths (H + Hmin) and lengths times p->H and p->L, discharges times
 * Q, times times T (the reference's normalisation scales; all 1 when the run
 * is not normalised).  Host code: no GPU is needed to write. */
typedef struct lhysea_output_spec {
    const char *prefix;        /* prefijo */
    const char *grid_name;     /* nombre_bati: the grid_name attribute (may be NULL) */
    double save_interval;      /* tiempo_guardar (internal units): < 0 never, 0 every step */
    double xmin, ymin;         /* grid origin (internal units) */
    double Hmin;               /* depth shift of H (<= 0) */
    double Q, T;               /* discharge and time scales */
    double sim_time;           /* tiempo_tot, the simulation_time attribute */
    const int32_t *gauges;     /* NULL: NetCDF; else ngauges flat cell indices or -1 */
    int32_t ngauges;
} lhysea_output_spec;
typedef struct lhysea_output lhysea_output;
/* Creates the files (LHYSEA_ERR_IO if they cannot be created); H is the run's depth. */
int lhysea_output_open(const lhysea_output_spec *spec, const lhysea_params *p, const double *H,
                       lhysea_output **out);
/* One record (NetCDF) or line (gauges) of the state W (6 planes, internal units) at time t. */
int lhysea_output_write(lhysea_output *out, const double *W, double t);
/* max_height from eta_max (h1 + h2 - H of lhysea_eta_max; NULL: not written),
 * then the files are closed and out freed. */
int lhysea_output_close(lhysea_output *out, const double *eta_max);

/* shallowWater with its output (replaces src/GPU/ShallowWater.cu:125-484 as
 * main.cxx calls it): the state is written at t = 0 and every save_interval
 * while t < t_end, at most max_steps steps (INT32_MAX: no limit), then
 * max_height; verbose prints the reference's per-step line
 * ("Iteracion %3d, deltaT = %e seg, Tiempo = %g seg") and steps one step per
 * call, otherwise the run advances to each save time in one call
 * (lhysea_run_until).  W holds the final state on return. */
int lhysea_shallow_water_output(const lhysea_params *p, double *W, const double *H, double t_end,
                                int32_t max_steps, int32_t verbose, const lhysea_output_spec *spec,
                                double *elapsed_s, int32_t *steps_done);

#ifdef __cplusplus
}
#endif

#endif /* LHYSEA_H */
