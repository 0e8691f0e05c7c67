/*
 * fwi_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the FWILab2d C v1.2.6 2D acoustic FWI hot path
 * (FWILab2d/FWILab2dCLangV1.2.6/src/{propagation,checkpoint,optimization,common}.c).
 * It exists to check the HIP library; nothing in the shipped library links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may use it.
 *
 * One source, two builds:
 *   liboracle_f32.so  REAL=float  -- the reference's own FP32/FP64 mix
 *   liboracle_f64.so  REAL=double -- the same propagator in FP64, used to state
 *                                    the rounding error of the FP32 and FP16 paths
 * Both export the same symbols; load them in separate ctypes handles.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as Record2D, observation.h:18-48. */
typedef struct
{
    int shotID;
    float sx, sz;
    int isx, isz;
    float fpeak;
    float *src;
    float *gx, *gz;
    int *igx, *igz;
    int sampleRate;
    int ntr;
    int ns;
    float dtr;
    float dt;
    int ix_min, ix_max;
    float **traces_vx, **traces_vz, **traces_p;
    float **precond, **grad;
} OracleRecord2D;

/* iso_acoustic2d_propagation, propagation.c:777-1036.  problem 0/1/2. */
void oracle_propagation(float **vp, float **rho, OracleRecord2D *obs, int Nshot, int nz, int nx,
                        float dx, float dz, float dt, int nt, int pmlThick, int freesurface,
                        float freq, float filter, int precondition, int problem,
                        OracleRecord2D *pre, float **grad);

/* oracle_propagation with iter and verbose: the reference's verbose debug
 * files (predicted / adjsource / gradient_raw / precond at level >= 1,
 * wavefield snapshots at level >= 2) in the working directory. */
void oracle_propagation_verbose(float **vp, float **rho, OracleRecord2D *obs, int Nshot, int nz,
                                int nx, float dx, float dz, float dt, int nt, int pmlThick,
                                int freesurface, float freq, float filter, int precondition,
                                int problem, OracleRecord2D *pre, float **grad, int iter,
                                int verbose);
/* Same as oracle_propagation(problem=1) but the adjoint pass reads forward
 * wavefields stored during the forward pass (every MST steps) instead of the
 * checkpoint/recompute schedule.  Used to pin the schedule logic. */
void oracle_gradient_fullstore(float **vp, float **rho, OracleRecord2D *obs, int Nshot, int nz,
                               int nx, float dx, float dz, float dt, int nt, int pmlThick,
                               int freesurface, float freq, float filter, int precondition,
                               OracleRecord2D *pre, float **grad);

/* Checkpoint plan of alloc_acoustic2dCheckPoint (checkpoint.c:29-99) and the
 * number of recompute sweeps the reverse loop performs (propagation.c:638-673). */
void oracle_checkpoint_plan(int nt, int *dCheckPoint, int *nCheckPoint, int *ntPoint,
                            int *nRecompute);
float oracle_memory_cost_checkpoint(int nt, int NX, int NZ);

/* One stencil sweep, propagation.c:13-74, on flat column-major [Nx][Nz] arrays. */
void oracle_fd_step(float *p2, const float *p1, const float *p0, const float *invrho,
                    const float *invrhox, const float *invrhoz, const float *K,
                    const float *coef1, const float *coef2, const float *coef3, int Nx, int Nz,
                    float invdx, float invdz, float invdxdx, float invdzdz, int ixStart,
                    int ixEnd, int izStart, int izEnd);
void oracle_freesurface(float *p2, int Nx, int Nz, int pmlThick);

/* Engine set-up (propagation.c:443-537): padded vp/rho -> the 7 coefficient
 * arrays used by the stencil.  vpm/rhom are the un-padded window [nx][nz]. */
void oracle_setup_coefs(const float *vpm, const float *rhom, int nx, int nz, int pmlThick,
                        float dx, float dz, float dt, float fpeak, float *D, float *Beta,
                        float *coef1, float *coef2, float *coef3, float *K, float *invrho,
                        float *invrhox, float *invrhoz);
void oracle_meal_coefs(float *D, float *Beta, const float *vp, float fpeak, float dz, float dx,
                       int NZ, int NX, int nlayer);
void oracle_pad_model(const float *mod, int nz, int nx, float *modp, int nlayer);

/* optimization.c */
void oracle_adjoint_source(OracleRecord2D *pre, OracleRecord2D *obs, float **vp, float freq,
                           int filter, OracleRecord2D *adj);
float oracle_objective(OracleRecord2D *rec, int Nshot);
float oracle_step_length(OracleRecord2D *pre0, OracleRecord2D *pre1, int Nshot);
float oracle_model_perturbation(float fraction, const float *slow, const float *grad, float *out,
                                int len);
void oracle_shot_precondition(float *grad, float *precond, int n);
void oracle_global_precondition(float *grad, const float *precond, int n, float alpha);

/* common.c subset */
void oracle_taper_traces(float **tr, int nt, int ntr, int taperLenTime, int taperLenTrace);
void oracle_smooth2d(float **in, float **out, int nx, int nz, int N, int iszdirect);
float oracle_peak_frequency(const float *s, int nt, float dt);
void oracle_vp2density(const float *vp, float *rho, int n);
void oracle_clip(float *mod, int n, float lo, float hi);
void oracle_v2slow(const float *v, float *slow, long n);
void oracle_slow2v(const float *slow, float *v, long n);
void oracle_set_water(float **model, float value, int waterdepth, int nx, int nz);
void oracle_lowpass_traces(float **tr, float **out, int nt, int ntr, float fpass, float fcut,
                           float dt);
void oracle_bandpass_traces(float **tr, float **out, int nt, int ntr, float fpass, float fcut,
                            float dt);
void oracle_interpolate_traces(float **tr, int ntr, int nt, int r, float **trs);

/* 1 if REAL is double */
int oracle_is_f64(void);

#ifdef __cplusplus
}
#endif
