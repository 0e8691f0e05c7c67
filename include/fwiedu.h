/*
 * fwiedu.h -- C ABI of the MI355X FWI hot-path library (libfwiedu.so).
 *
 * The first part declares, with the same names, argument order and meaning,
 * the entry points of FWILab2d C v1.2.6 (FWILab2d/FWILab2dCLangV1.2.6/include/)
 * that the reference's drivers call on the 2D acoustic FWI path.  A driver
 * written against the reference headers links against this library instead;
 * the work runs on the GPU (HIP, gfx950) and the results are written back into
 * the caller's host arrays exactly where the reference writes them.
 *
 * 2D arrays are Seismic Unix alloc2float(n1, n2) blocks: a row-pointer table
 * over ONE contiguous block, n1 (z) fastest.  As in the reference, the library
 * reads and writes through a[0] and assumes that contiguity.
 *
 * Errors follow the reference's err() convention: a message on stderr and
 * exit(1).  HIP failures print the HIP error string.
 *
 * The second part (fwiedu_*) has no counterpart in the reference: it selects
 * batching, checkpoint precision and timing, and exposes device-level
 * measurements.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* ---- types ------------------------------------------------------------- */

/* Record2D: same fields, order and types as observation.h:18-48. */
typedef struct
{
    int shotID;
    float sx;
    float sz;
    int isx;
    int isz;
    float fpeak;
    float *src;
    float *gx;
    float *gz;
    int *igx;
    int *igz;
    int sampleRate;
    int ntr;
    int ns;
    float dtr;
    float dt;
    int ix_min;
    int ix_max;
    float **traces_vx;
    float **traces_vz;
    float **traces_p;
    float **precond;
    float **grad;
} Record2D;

/* Acoustic2dCheckPoint: same fields as checkpoint.h:16-26 (host storage). */
typedef struct
{
    int nt;
    int dCheckPoint;
    int nCheckPoint;
    int ntPoint;
    int *status;
    float memCost;
    float ***p;
    float ***storage_p;
} Acoustic2dCheckPoint;

#define MST 5 /* propagation.h:21 */

/* ---- propagation.h ----------------------------------------------------- */

/* Replaces iso_acoustic2d_propagation, propagation.h:133-141 / propagation.c:777-1036.
 * problem 0: model records into record2d_pre[].traces_p.
 * problem 1: gradient; record2d_pre[].traces_p holds the adjoint source on return
 *            and grad[nx][nz] the (preconditioned when precondition==1) gradient.
 * problem 2: model + adjoint source (line-search trial) into record2d_pre[].traces_p.
 * Shots that share a simulation window and fpeak advance together on the GPU.
 * verbose >= 1 writes the reference's debug files of a gradient into the working
 * directory (predicted_*, adjsource_*, gradient_raw_* of every shotInterval-th
 * shot, gradient_raw_* and precond_* of the call: propagation.c:917-944,
 * 996-1005); verbose >= 2 also the wavefields of those shots every snapInterval
 * steps (wavefield_forward_*, wavefield_forward_reconstruction_*,
 * wavefield_backward_*: :597-601, 717-723), named after the shot's index in the
 * call, from the per-shot-arena stencils run kernel by kernel.  shotInterval
 * (default 20) and snapInterval (default 1000) are read from the environment. */
void iso_acoustic2d_propagation(float **vp, float **rho, Record2D *record2d_obs, int Nshot, int nz,
                                int nx, float dx, float dz, float dt, int nt, int pmlThick,
                                int freesurface, float freq, float filter, int precondition,
                                int problem, int iter, int verbose, Record2D *record2d_pre,
                                float **grad);

/* Replaces iso_acoustic2d_propagation_engine, propagation.h:95-100 / propagation.c:356-775.
 * Note the (nx, nz) order.  mode 0 forward; mode 1 forward that fills the host
 * checkpoint chkpt (pairs at multiples of dCheckPoint, status set) and, when
 * precondition==1, the pseudo-Hessian into grad[nx][nz]; mode 2 adjoint that
 * reads the checkpoint pairs from chkpt and writes the gradient to grad.
 * verbose >= 2 writes the wavefield files of iso_acoustic2d_propagation, named
 * as shot 0. */
void iso_acoustic2d_propagation_engine(Record2D *record2d, Acoustic2dCheckPoint *chkpt,
                                       float **vpm, float **rhom, int nx, int nz, float dx,
                                       float dz, float dt, int nt, int pmlThick, int freesurface,
                                       int precondition, int mode, int iter, int verbose,
                                       float **grad);

/* Replaces acoustic_iso2d_fd_it_6order, propagation.h:23-28: one sweep of the
 * stencil over ix in [ixStart, ixEnd], iz in [izStart, izEnd] of host row
 * tables, run on the GPU.  Only the cells the sweep reads (the box widened by
 * 3 in each direction for p1) and writes are touched, so any row table works. */
void acoustic_iso2d_fd_it_6order(float **p2, float **p1, float **p0, float **invrho,
                                 float **invrhox, float **invrhoz, float **K, float **coef1,
                                 float **coef2, float **coef3, float invdx, float invdz,
                                 float invdxdx, float invdzdz, int ixStart, int ixEnd, int izStart,
                                 int izEnd);

/* Replaces acoustic_iso2d_freesurface, propagation.h:58. */
void acoustic_iso2d_freesurface(float **p2, int Nx, int Nz, int pmlThick);

/* Replaces mealCoefs2d, propagation.h:72-74 (vs == NULL or vs > 0 per cell). */
void mealCoefs2d(float **D, float **Beta, float **vp, float **vs, float fpeak, float eal_alpha,
                 float dz, float dx, int NZ, int NX, int nlayer);

/* ---- checkpoint.h (host-side bookkeeping, same semantics) --------------- */
void acoustic2dCheckPoint_init(Acoustic2dCheckPoint *chkpt);                  /* checkpoint.h:29 */
void alloc_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, int NZ, int NX, int nt); /* :41 */
void free_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt);                  /* :58 */
void save_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, float **p0, float **p1, int NP,
                               int it);                                        /* :71 */
void get_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, float **p0, float **p1, int NP,
                              int it);                                         /* :88 */
void get_acoustic2dWavefield(Acoustic2dCheckPoint *chkpt, float **p, int NP, int it); /* :104 */
void save_acoustic2dWavefield(Acoustic2dCheckPoint *chkpt, float **p, int NP, int it); /* :120 */
int get_acoustic2dCheckPointStatus(Acoustic2dCheckPoint *chkpt, int it);     /* :136 */
float calculate_memory_cost_checkpoint(int nt, int NX, int NZ);              /* :152 */

/* ---- optimization.h (GPU reductions / elementwise) ---------------------- */

/* optimization.h:17-18.  filter 0 (low-pass) and 1 (band-pass) run the
 * reference's FFT chain through hipFFT; 2 is no filter. */
void compute_adjoint_source2d(Record2D *record2d_pre, Record2D *record2d_obs, float **vp,
                              float freq, int filter, Record2D *record2d_adjsource);
void apply_gradient_precondition(float *grad, const float *precond, int n, float alpha); /* :40 */
void apply_gradient_shot_precondition(float *grad, float *precond, int n);             /* :60 */
float calculate_objective_value(Record2D *record2d, int Nshot);                       /* :79 */
float compute_model_perturbation(float fraction, float *slowness, float *gradient, float *out,
                                 int len);                                             /* :93 */
float calculate_step_length(Record2D *record2d_pre0, Record2D *record2d_pre1, int Nshot); /* :112 */

/* ---- fwiedu_*: no counterpart in the reference -------------------------- */

/* Shots advanced per launch by iso_acoustic2d_propagation (0 = as many as fit
 * in the device-memory budget; 1 = one shot per launch, the plain path). */
void fwiedu_set_batch(int shots_per_launch);
/* Stencil launch shape: tile height tz (32 or 64 rows) and shots per block spb
 * (1, 2, 4, 8).  tz = 0 (default) lets the library's autotuner time every shape
 * once per geometry and keep the fastest. */
void fwiedu_set_tile(int tz, int spb);
/* Block order of a forced launch shape.  Square tiles: 1 puts vertically
 * adjacent tiles next to each other in launch order, 0 horizontally adjacent
 * ones.  Shots in lanes (set before fwiedu_set_lanes): 0 z-runs fastest, then
 * the 64-shot groups of a column group; 1 z-runs fastest, then column groups,
 * shot groups outermost; 2 column groups fastest, then z-runs, shot groups
 * outermost. */
void fwiedu_set_tile_order(int order);
/* Force the column-marching stencil: txm (16, 32, 64) columns per block, spb
 * (1, 2, 4) shots per block, pd (1, 2, 4) columns of loads kept in flight.
 * fwiedu_set_tile(0, 0) returns to the autotuner. */
void fwiedu_set_march(int txm, int spb, int pd);
/* Block order of a forced launch shape: 1 numbers stencil blocks so that each
 * XCD works on one contiguous range of tiles, 0 keeps the hardware's
 * round-robin order.  (The autotuner tries both.) */
void fwiedu_set_xcd_remap(int on);
/* Force the shots-in-lanes stencil (fields laid out [x][z][shot], the 64 lanes
 * of a wavefront are 64 shots of one cell): wpb columns per block, each wave
 * marching zrun rows with the loads of the next pd rows in flight.
 * lds = 0: x-neighbours loaded from memory (wpb 1, 2, 4, 8; zrun 8, 16, 32,
 * 64, 128 or 256; pd 1 ... 3);
 * lds = 1: x-neighbours exchanged through LDS (wpb 4, 8, 16; zrun 8 ... 256;
 * pd 1 ... 4);
 * lds = 2: the same with the rows each wave holds in a register ring (wpb 8;
 * zrun 8 ... 256, best a multiple of the ring's unroll, 10 rows for pd 2 or
 * 3 and 12 for pd 4 or 5; pd 2 ... 5).
 * Applies to iso_acoustic2d_propagation and sessions; fwiedu_set_tile(0, 0)
 * returns to the autotuner. */
void fwiedu_set_lanes(int wpb, int zrun, int pd, int lds);
/* Stencil arithmetic.  0 (default): every product and sum rounded in the
 * order of the reference's C source (bit-identical to a strict IEEE build of
 * it).  1: each product fused into the following sum (FMA), as a contracting
 * compile of the reference (it is built with -ffast-math) rounds; fewer
 * instructions, results within FP32 rounding of the strict ones. */
void fwiedu_set_fma(int on);
/* 1 (default): with shots in lanes, when two work fields of the batch fit the
 * GPU's 256 MB Infinity Cache (the 221 x 601 demo size with 120 shots), a time
 * step whose previous field is a work buffer writes the new field over it
 * (two work fields instead of three, which the cache then holds between
 * steps); the arithmetic and the results are the same.  0: every step writes
 * a third field. */
void fwiedu_set_inplace(int on);
/* 1: run each batch's forward and adjoint time loops as HIP graphs: a loop
 * runs kernel by kernel the first time, is captured when the same batch (same
 * device buffers, sizes, schedule and launch shape) comes a second time, and
 * is replayed from then on: the next gradients of a session, the next FWI
 * iterations.  0 (default): launch kernel by kernel.  Sampled kernel timing
 * skips captured launches. */
void fwiedu_set_graphs(int on);
/* Field layout of batched launches when the stencil shape is autotuned: -1
 * (default) shots in lanes from 48 shots per launch up, per-shot arenas below;
 * 0 always per-shot arenas; 1 always shots in lanes. */
void fwiedu_set_layout(int layout);
/* Fraction of free device memory the engine may use (default 0.85). */
void fwiedu_set_memory_fraction(double frac);
/* Sample every `every`-th launch of each stencil mode (forward,
 * reconstruction sweep, adjoint) with HIP events on the library's stream
 * (0 = off).  Launches replayed from a captured HIP graph are not sampled. */
void fwiedu_set_kernel_timing(int every);
/* Checkpoint storage precision of iso_acoustic2d_propagation and sessions:
 * 32 (default) or 16.  With 16, each checkpointed field u(kd-1), u(kd) of each
 * shot is stored as FP16 of u * 2^k, 2^k the largest power of two that keeps
 * max|u| * 2^k <= 65504, and restored exactly as float(h) * 2^-k; arithmetic
 * stays FP32.  Records are unchanged, the gradient sees the rounding of the
 * restarted segments.  The per-shot engine with a host Acoustic2dCheckPoint
 * (the reference's FP32 structure) always stores FP32. */
void fwiedu_set_checkpoint_bits(int bits);
/* Named counters of the last calls: "fd_launches", "fd_inplace" (of them the
 * launches that wrote the new field over p0), "fd_timed", "fd_ms_avg",
 * "fd_cells_per_launch" (sampled reconstruction launches), "fd_shots_per_launch",
 * "batch", "device_bytes" ...; per stencil mode m = 0 forward, 1 reconstruction,
 * 2 adjoint: "fd<m>_timed", "fd<m>_ms_avg", "fd<m>_cells_per_launch",
 * "fd<m>_shot_cells_per_launch", "fd<m>_acc_shot_cells_per_launch" (shot-cells
 * of the sampled launches whose pseudo-Hessian or imaging epilogue ran, per
 * sampled launch); the launch shape of mode m in the last batch, "tile<m>_kind",
 * "tile<m>_wpb", "tile<m>_zrun", "tile<m>_pd", "tile<m>_remap", "tile<m>_zfast"
 * ..., and of its launches that stage their accumulation term (v = 1) or add
 * the staged terms to the FP64 accumulator (v = 2): "tile<m><v>_zrun", ...;
 * "graphs_built", "graph_launches", "shapes_reused" (launch shapes a serving
 * worker took from a similar geometry instead of timing its own).
 * Unknown names return -1. */
double fwiedu_stat(const char *name);
void fwiedu_reset_stats(void);
/* No counterpart in the reference: compute_model_perturbation (above) on
 * arrays already on the device, for callers that cannot allocate device memory
 * (code that forbids hipMalloc, serving loops that keep allocation out of their
 * steady state).  slowness, gradient and out are device pointers of len floats
 * (out may be slowness, not gradient); returns the same scalar and writes the
 * same out as compute_model_perturbation on the same values.  Its scratch is a
 * fixed 32 KiB device array per GPU, held until the process ends, whose size
 * does not depend on len (any len from 1 to INT_MAX); every call on a GPU
 * shares it, so calls that overlap on one GPU race.  Allocates no device
 * memory. */
float fwiedu_model_perturbation_noalloc(float fraction, const float *slowness,
                                        const float *gradient, float *out, int len);
/* Shots per launch the shots-in-lanes layout can hold on an nz x nx model with
 * pmlThick absorbing cells: the largest multiple of 64 whose field (padded
 * grid x shots x 4 bytes) stays below 2 GiB, 0 if not even 64 fit (the layout
 * is then not used).  Host arithmetic only; needs no device. */
long fwiedu_lanes_max_shots(int nz, int nx, int pmlThick);
/* Device-resident gradient session: the model, the observed records and all
 * wavefield state stay in HBM between calls, so repeated gradients (FWI
 * iterations, serving) move no traces over PCIe.  Same arguments and meaning as
 * iso_acoustic2d_propagation with problem 1; the records are copied at create. */
void *fwiedu_session_create(float **vp, float **rho, Record2D *record2d_obs, int Nshot, int nz,
                            int nx, float dx, float dz, float dt, int nt, int pmlThick,
                            int freesurface, float freq, float filter, int precondition);
/* Replace the model (host arrays, copied). */
void fwiedu_session_set_model(void *session, float **vp, float **rho);
/* One gradient; returns 0.5 * sum of the squared adjoint sources (the misfit of
 * the processed residual).  grad may be NULL (the gradient stays on the device). */
float fwiedu_session_gradient(void *session, float **grad);
void fwiedu_session_destroy(void *session);
/* One iteration of the FWI driver (acoustic2d_fwi.c:409-465) with the model,
 * slowness, gradient, perturbation and both sets of adjoint sources on the
 * device: gradient at (freq, filter), smooth2D along z (rz) then x (rx),
 * objective, compute_model_perturbation(fraction) of the smoothed gradient,
 * water rows of the gradient zeroed, trial model (slow2v, water, clip, Gardner
 * density when gardner = 1), line-search trial, objective, step length, model
 * update.  The session's model becomes the updated one; fun1, fun2, steplen and
 * scalar are written back; vp, rho, grad receive the new model and the smoothed
 * gradient when not NULL. */
typedef struct FwieduIteration {
    float freq, filter;
    int rz, rx;
    int waterdepth;
    float vp_water, rho_water;
    float vpmin, vpmax, rhomin, rhomax;
    int gardner;
    float fraction;
    float fun1, fun2, steplen, scalar;
    /* iter and verbose of the driver's two iso_acoustic2d_propagation calls
     * (acoustic2d_fwi.c:411-416, 441-446): the debug files of the verbose
     * levels, as that function writes them; 0 writes none */
    int iter, verbose;
} FwieduIteration;
void fwiedu_session_iterate(void *session, FwieduIteration *it, float **vp, float **rho,
                            float **grad);

/* ---- Data formats of the drivers (host only) -------------------------------- */
/* calculate_peak_frequency (common.h:230, common.c:363-413): frequency of the
 * first maximum of |rfft(s)|. */
float calculate_peak_frequency(const float *s, int nt, float dt);
/* create_grid_observe2d (observation.h:152, observation.c:338-420): grid
 * indices, window [min igx - 30, max igx + 30] and window-local isx / igx. */
void create_grid_observe2d(Record2D *record2d, int Nshot, float dx, float dz, int nx);
/* write_acoustic2d_traces (dataio.h, dataio.c:306-372): all shots to the SU
 * file "<prefix>_<Nshot>_<ns>_<traces>_<dt ms %05.2f>ms.su". */
void write_acoustic2d_traces(char *prefix, Record2D *record2d, int Nshot);
/* get_shots_number (dataio.c:400-443): gathers in an SU file. */
int get_shots_number(const char *filename);
/* read_gather2d_from_sufile (dataio.c:445-503): geometry and traces of Nshot
 * gathers (gx, gz, traces_p allocated with the SU conventions). */
void read_gather2d_from_sufile(Record2D *record2d, const char *recordfile, int Nshot, int nt);
/* init_observe2d_from_JSON (observation.c:422-532) on the "record" array of a
 * JSON acquisition file (the reference's takes the array parsed by cJSON). */
void fwiedu_observe2d_from_json(Record2D *record2d, const char *jsonfile, int Nshot, float *x,
                                int nx, float *z, int nz, float dt, int nt, float *src,
                                int sampleRate);
/* Shots in the "record" array of a JSON acquisition file. */
int fwiedu_json_shot_count(const char *jsonfile);
/* generate_shot_positions (observation.h:84-89, observation.c:102-165): shots
 * from fsx every ds at depth sdepth, at most Nshot_in of them and none past
 * x[nx-1]; *sx, *sz allocated (alloc1float); returns the shot count. */
int generate_shot_positions(const float *x, int nx, const float *z, int nz, float fsx, float ds,
                            float sdepth, int Nshot_in, float **sx, float **sz);
/* init_observe2d (observation.h:127-129, observation.c:229-336): receivers at
 * the offsets -offsmax..-offsmin and offsmin..offsmax (step dtr) from each
 * shot, inside the model, at depth gdepth; gx, gz, traces_p and src allocated
 * with the SU conventions. */
void init_observe2d(Record2D *record2d, int Nshot, float *x, int nx, float *z, int nz, float dt,
                    int nt, float *sx, float *sz, float *src, float offsmin, float offsmax,
                    float dtr, float gdepth, int sampleRate);

/* ---- Serving loop (no counterpart in the reference) ------------------------
 * A job is one iso_acoustic2d_propagation call of problem 0 (modeling: records
 * into pre) or 1 (gradient: adjoint sources into pre, gradient into grad),
 * filter 2.  The caller keeps every array of a job alive until
 * fwiedu_server_wait returns for it. */
typedef struct FwieduJob {
    int problem;
    float **vp, **rho;   /* [nx][nz] row tables over one block, as the reference */
    Record2D *obs;       /* geometry, and the observed traces of a gradient job */
    Record2D *pre;       /* output traces */
    float **grad;        /* output gradient (problem 1) */
    int nshot, nz, nx;
    float dx, dz, dt;
    int nt, pml, freesurface;
    float freq;
    int precondition;
} FwieduJob;
/* Start a server with `slots` workers (1-16), each a host thread with its own
 * HIP stream and device buffers, sharing mem_fraction of the free device
 * memory.  Jobs are admitted in arrival order as a worker and their device
 * footprint become free (a later, smaller job may pass a waiting one during
 * its first second of waiting); the kernels of running jobs overlap on the
 * device.  graphs: replay the time loops of repeated job shapes from HIP
 * graphs.  ckpt_bits: checkpoint storage of gradient jobs (32 or 16). */
void *fwiedu_server_create(int slots, double mem_fraction, int graphs, int ckpt_bits);
/* Queue a job; returns its id.  Invalid jobs end the process like err(). */
long fwiedu_server_submit(void *server, const FwieduJob *job);
/* 1 once the job has completed. */
int fwiedu_server_done(void *server, long id);
/* Block until the job completes; latency_s = submit to completion, queue_s =
 * submit to start, misfit = 0.5 * sum of squared adjoint sources (gradient
 * jobs).  Any pointer may be NULL.  Returns 0, or -1 for an unknown id. */
int fwiedu_server_wait(void *server, long id, double *latency_s, double *queue_s,
                       double *misfit);
/* "completed", "queued", "peak_concurrency" (most jobs running at once),
 * "budget_bytes", "p50_s", "p99_s", "max_s" (latencies of completed jobs). */
double fwiedu_server_stat(void *server, const char *name);
/* Finish the queued jobs, stop the workers, free their device memory. */
void fwiedu_server_destroy(void *server);

/* Free all device memory held by the library. */
void fwiedu_release(void);
/* 1 if a HIP device is present and usable. */
int fwiedu_device_ok(void);
/* Library version string. */
const char *fwiedu_version(void);

#ifdef __cplusplus
}
#endif
