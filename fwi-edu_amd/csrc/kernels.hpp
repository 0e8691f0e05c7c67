// Device kernels of the FWI hot path and their host-side launchers.
//
// Field layout in HBM: every padded field is [Nx][ld] floats, z fastest, with
// ld = Nz rounded up to a multiple of 64 so each column starts on a 256-byte
// boundary and a wave's 64 lanes read one column segment in one coalesced
// request.  A batch of S shots keeps one arena per shot; the same buffer role of
// shot s sits at  base + s * shot_stride.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fwiedu {

enum FdMode { FD_FWD = 0, FD_RECON = 1, FD_ADJ = 2 };
// accumulation steps staged (as float terms) before one pass over the FP64
// accumulators adds them in order (7: a pass every 8th accumulation step; 3,
// every 4th, made a 221 x 601 gradient 0.8 % slower)
constexpr int ACC_STAGES = 7;

// Geometry of one padded simulation window.
struct Geom {
    int nx, nz;          // physical window (nz x nx)
    int pml, fs;         // absorbing-layer width, free surface flag
    int Nx, Nz, ld;      // padded extents and leading dimension
    int ixs, ixe, izs, ize;  // stencil bounds (propagation.c:534-537)
    float dx, dz, dt;
    float invdx, invdz, invdxdx, invdzdz, invdtdt;
    long fsz;            // floats per field = Nx * ld
};

// The coefficient arrays, consecutive blocks of fsz floats.  The stencil streams
// the first six; K (bulk modulus) is read only at source and receiver cells.
// CO_A holds coef3*K (propagation.c:71 evaluates that product first).
enum { CO_C1 = 0, CO_C2, CO_A, CO_IRX, CO_IRZ, CO_IR, CO_K, CO_N };
// floats per cell of a coefficient buffer: the CO_N arrays and their
// cell-interleaved copy (8 per cell) that the shots-in-lanes stencils read
constexpr int CO_ALLOC = CO_N + 8;
constexpr int CO_STREAMED = 6;

// Stencil launch shape.  kind 0: square tiles of tz (32 or 64) rows, spb shots
// per block (1, 2, 4, 8).  kind 1: column march over txm columns per block,
// spb shots per block (1, 2, 4), pd columns of loads in flight.  kind 2: shots
// in wavefront lanes (fields laid out [x][z][shot]), wpb columns per block,
// each wave marching zrun rows.  kind 3: kind 2 with the row's x-neighbours
// exchanged through LDS (wpb 4, 8 or 16), pd rows of loads in flight (1-4).
// kind 4: kind 3 with 8 columns per block and the rows a wave holds in a
// register ring (pd 2-5 rows in flight, unrolled ring_unroll(pd) rows deep).
struct FdTile {
    int tz = 64;
    int spb = 1;
    int kind = 0;
    int txm = 0;
    int zfast = 0;  // kind 0: tile order; kinds 2-4: block order 0, 1, 2 (see lanes_block)
    int pd = 1;     // kind 1: columns prefetched ahead; kinds 2-4: rows of loads in flight
    int remap = 0;  // XCD-aware block order
    int wpb = 4;    // kinds 2-4: waves (= adjacent columns) per block
    int zrun = 32;  // kinds 2-4: rows marched by one wave
};

struct FdArgs {
    const float *coef;      // CO_ALLOC * fsz (see CO_ALLOC)
    float *out;             // shot 0 pointers
    const float *cur;
    const float *prev;
    long shot_stride;       // floats between shots (per-shot arenas: the arena size;
                            // shots in lanes: 1)
    long cell_stride;       // floats between cells (per-shot arenas: 1; lanes: Sp)
    int nshot;
    int Nx, ld, ixs, ixe, izs, ize, pml, fs, nx, nz;
    float invdx, invdz, invdxdx, invdzdz, invdtdt, dt;
    int xtiles, ztiles, ngroups;
    int remap;              // XCD-aware block order (1) or hardware order (0)
    int zfast;              // tile order: z-neighbours adjacent (1) or x-neighbours (0)
    int mx_len, mx_nseg, mz_len, mz_nseg;  // column-march plan
    int lz_run, lz_nrun, lz_ngrp, lz_nblk; // shots-in-lanes plan
    int lz_nxg, lz_order;
    int fma;                // 1: FMA-contracted update arithmetic, 0: strict source order
    // source (FD_FWD / FD_RECON)
    const int *src_cell;    // [nshot] flat padded index or -1
    const float *src_wav;   // [nshot][nt]
    int nt, it;
    // pseudo-Hessian (FD_FWD) / imaging (FD_ADJ) every MST steps
    double *acc;            // shot 0, [nx][nz]
    long acc_stride;        // doubles between shots
    long acc_cell;          // doubles between cells
    const float *u;         // FD_ADJ: forward field at `it`, shot 0
    int acc_on;
    // Deferred accumulation (shots-in-lanes LDS stencil): with acc_stage the
    // launch writes its term (float) to staging slot acc_slot of accstage
    // ([ACC_STAGES][nx*nz][lanes], slots accstage_stride floats apart);
    // otherwise it adds the acc_n staged terms, oldest first, and its own to
    // the accumulator.
    float *accstage;
    long accstage_stride;
    int acc_stage, acc_slot, acc_n;
    // Receivers inside the shots-in-lanes LDS stencil (one receiver row per
    // column): rrow[ix] is the row of column ix's receivers or -1, rcol[ix]
    // its index among those columns; stage is time-major [nt][nslot] with
    // slot = rcol[ix] * cell_stride + lane.
    const int *rrow;
    const int *rcol;
    float *stage;
    long nslot;
    int rec_on;             // FD_FWD: stage[it][slot] = p1 at the receiver row
    int inj_on;             // FD_ADJ: p2 at the receiver row -= stage[it-1][slot] * K dt^2/(dx dz)
    // Row ranges [lo, hi] of the source cells (src_cell) and of the receiver
    // rows (rrow); lo > hi: none.  The register-ring stencil tests sources and
    // receivers only in the blocks whose rows meet them; whoever sets
    // src_cell or rrow sets these.
    int src_zlo, src_zhi, rec_zlo, rec_zhi;
    // FD_RECON with the register-ring stencil: a field of the same layout
    // that receives the new field's model cells as well (nullptr: none)
    float *snap;
};

void launch_fd(int mode, const FdArgs &a, FdTile tile, hipStream_t st);
// Events the next stencil launches of this thread carry in their dispatch
// (start, stop; nullptr: none), and the error of the last such launch.
void set_stencil_events(hipEvent_t start, hipEvent_t stop);
hipError_t take_stencil_launch_error();
int fd_tile_rows_padded(int tz, int izs, int ize);
// rows per unrolled group of the register-ring stencil (launch kind 4) with
// pd rows of loads in flight
int ring_unroll(int pd);
// largest multiple of 64 shots whose lanes field of fsz cells stays below 2 GiB
long lanes_max_shots(long fsz);
int march_lds_bytes(const FdArgs &a, FdTile tile);

// ---- set-up --------------------------------------------------------------
// window model [nx][nz] (contiguous) -> padded [Nx][ld] by edge replication
void launch_pad(const float *win, int nx, int nz, int pml, float *out, const Geom &g,
                hipStream_t st);
// MEAL D / Beta with host-computed scalars (layer weights w[1..n], ln R, sqrt 3)
void launch_meal(const float *vp, const float *vs, float *D, float *Beta, const Geom &g,
                 int nlayer, const float *w_dev, float lnR, float pd, float sqrt3, float fpeak,
                 hipStream_t st);
// D/=Beta, vp/=Beta, rho*=Beta, coef1..3 (double from FP32 products), K, 1/rho
void launch_coefs(float *vp, float *rho, float *D, const float *Beta, float *coef, const Geom &g,
                  hipStream_t st);
// d(1/rho)/dx, d(1/rho)/dz: 2nd-order everywhere, 12-point optimised inside
void launch_grad_invrho(float *coef, const Geom &g, hipStream_t st);
// the cell-interleaved copy [fsz][8] of the coefficients after the CO_N arrays
// (CO_ALLOC * fsz floats in all)
void launch_coef_pack(float *coef, const Geom &g, hipStream_t st);

// ---- receivers -----------------------------------------------------------
struct RcvTable {
    int n;                  // entries
    const int *shot;        // [n]
    const int *cell;        // [n] flat padded index
    const long *trace_off;  // [n] offset of this trace in the record buffer
    const int *sr;          // [n] sample rate of the entry's shot
};
void launch_record(const RcvTable &t, const float *field0, long shot_stride, long cell_stride,
                   float *rec, int it, hipStream_t st);
// stage[t][slot] = src[slot_off[slot] + t] (0 where slot_off < 0), t < nt
void launch_stage_gather(const long *slot_off, const float *src, float *stage, long nslot, int nt,
                         hipStream_t st);
// rec[slot_off[slot] + t / sr[slot]] = stage[t][slot] for t % sr[slot] == 0
// (slots with slot_off < 0 skipped), t < nt
void launch_stage_scatter(const long *slot_off, const int *slot_sr, const float *stage, float *rec,
                          long nslot, int nt, hipStream_t st);
struct InjTable {
    int n;                  // unique (shot, cell) pairs
    const int *shot, *cell; // [n]
    const int *beg;         // [n+1] into trace_off (receiver order kept)
    const long *trace_off;  // offset of the trace's sample 0 in the adjoint-source buffer
};
void launch_inject(const InjTable &t, float *field0, long shot_stride, long cell_stride,
                   const float *K, const float *adj, int it, float dt, float invdx, float invdz,
                   hipStream_t st);

// ---- adjoint source ------------------------------------------------------
// out = taper(pre - obs) for one shot [ntr][ns]; the cosine tables are built
// on the host (coefT[0..Lt], coefX[0..Lx]), Lt/Lx < 0 means no taper.
void launch_adjsrc(const float *pre, const float *obs, float *out, int ntr, int ns,
                   const float *coefT, int Lt, const float *coefX, int Lx, hipStream_t st);

// ---- gradient finalisation / preconditioning ------------------------------
// g[i] = (float)(scale * acc[i] / (double)(rho*vp) or /(double)((rho*vp)^2))
void launch_finalize(const double *acc, long acc_stride, long acc_cell, float *out,
                     long out_stride, int nshot, const float *vpw, const float *rhow, int n,
                     double scale, int squared, hipStream_t st);
// per-shot sum of p (double) -> avg -> g /= 1 + p/(7 avg), p /= same
void launch_shot_precondition(float *g, float *p, long stride, int nshot, int n, double *scratch,
                              hipStream_t st);
// sum[ix0+i] += v[s][i] over shots in order (window offsets per shot)
void launch_accumulate(const float *v, long stride, int nshot, const int *ix0_dev,
                       const int *nxx_dev, int nz, int nx, double *sum, hipStream_t st);
void launch_to_float(const double *in, float *out, long n, hipStream_t st);
// grad = grad / (p + alpha * rms(p)), rms from a deterministic double reduction
void launch_global_precondition(float *g, const float *p, long n, float alpha, double *scratch,
                                hipStream_t st);

// ---- reductions ----------------------------------------------------------
// deterministic double sums; results land in res[0..k) on the device
void launch_sum_sq(const float *x, long n, double *scratch, double *res, hipStream_t st);
void launch_sum_step(const float *a, const float *b, long n, double *scratch, double *res,
                     hipStream_t st);
void launch_sum(const float *x, long n, double *scratch, double *res, hipStream_t st);
void launch_minmax(const float *x, long n, float *scratch, float *res, hipStream_t st);
void launch_histogram(const float *g, long n, float mn, float interval, int *hist,
                      hipStream_t st);
void launch_scale(const float *in, float *out, long n, float s, hipStream_t st);
size_t reduce_scratch_doubles();

// ---- trace filtering (FFT band filters of the adjoint source) ---------------
void launch_taper(float *x, int ntr, int ns, const float *coefT, int Lt, const float *coefX,
                  int Lx, hipStream_t st);
void launch_spec_window(float2 *spec, long total, int nb, const float *win, hipStream_t st);
void launch_div(float *x, long n, float d, hipStream_t st);
void launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st);
void launch_zero_stuff(const float *in, int ntr, int ns, int r, float2 *buf, int M, hipStream_t st);
void launch_interp_window(float2 *buf, int ntr, int M, int r, hipStream_t st);
void launch_interp_out(const float2 *buf, int ntr, int M, float *out, int nt_out, hipStream_t st);

// ---- model update of the FWI iteration -------------------------------------
enum { MOP_ADD = 0, MOP_SCALE, MOP_RECIP, MOP_CLIP, MOP_GARDNER, MOP_WATER };
void launch_smooth2d(const float *in, float *out, int nx, int nz, int N, int zdir, const float *w,
                     hipStream_t st);
void launch_model_op(int op, const float *a, const float *b, float *out, long n, int nz, float s,
                     float lo, float hi, int wd, hipStream_t st);

// ---- small helpers -------------------------------------------------------
void launch_freesurface(float *p, int Nx, int ld, int pml, hipStream_t st);
void launch_fill_u32(unsigned *p, int n, unsigned v, hipStream_t st);

// ---- FP16 checkpoints ------------------------------------------------------
// Pack the FP32 field f of nshot shots (strides SS, CS, ncell cells each) into
// FP16 h (strides SS16, CS16) with a per-shot power-of-two scale; inv[s] gets
// the inverse scale.  mx: ck_mx_words(nshot) words of scratch.  lanes selects
// the thread mapping of the layout (shot fastest or cell fastest).
// The per-shot maxima are reduced into CK_MX_COPIES copies of the shot words,
// each copy on its own cache lines, so that the blocks' atomics spread out.
constexpr int CK_MX_COPIES = 32;
inline long ck_mx_stride(int nshot) { return (long)(nshot + 31) / 32 * 32; }
inline long ck_mx_words(int nshot) { return CK_MX_COPIES * ck_mx_stride(nshot); }
void launch_ck_pack(const float *f, long SS, long CS, __half *h, long SS16, long CS16, long ncell,
                    int nshot, int lanes, unsigned *mx, float *inv, hipStream_t st);
void launch_ck_unpack(const __half *h, long SS16, long CS16, float *f, long SS, long CS,
                      long ncell, int nshot, int lanes, const float *inv, hipStream_t st);

}  // namespace fwiedu
