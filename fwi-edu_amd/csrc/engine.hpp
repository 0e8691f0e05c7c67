// Host-side runtime shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/fwiedu.h"
#include "kernels.hpp"

namespace fwiedu {

// SU err() convention: message on stderr, exit(1).
[[noreturn]] void fail(const char *fmt, ...);

#define HIPCHECK(x)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            ::fwiedu::fail("<error>: %s failed: %s (%s:%d)", #x, hipGetErrorString(e_),  \
                           __FILE__, __LINE__);                                          \
    } while (0)

// A grow-only device allocation from the device's stream-ordered memory pool
// (allocated and returned in the order of the calling thread's engine
// stream), or from hipMalloc where pools are missing.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool pooled = false;
    void *get(size_t bytes);
    template <class T> T *as(size_t n) { return static_cast<T *>(get(n * sizeof(T))); }
    void release(hipStream_t st);
};
// Bytes the memory pool holds but no buffer uses (free for the next
// allocation, counted as used by hipMemGetInfo); trim_pool returns them to
// the device.
size_t pool_idle_bytes();
void trim_pool();

struct Stats {
    long fd_launches = 0;
    long fd_inplace = 0;      // of which wrote the new field over p0
    long fd_cells = 0;        // interior cell-updates issued (all shots)
    long fd_timed = 0;            // sampled reconstruction launches (no accumulation epilogue)
    double fd_ms_timed = 0;
    double fd_cells_timed = 0;
    // sampled launches per stencil mode (0 forward, 1 reconstruction, 2 adjoint):
    // count, milliseconds, interior cells, interior cells x shots, and the
    // interior cells x shots of the launches whose accumulation epilogue ran
    long md_timed[3] = {0, 0, 0};
    double md_ms[3] = {0, 0, 0};
    double md_cells[3] = {0, 0, 0};
    double md_shot_cells[3] = {0, 0, 0};
    double md_acc_shot_cells[3] = {0, 0, 0};
    long batch = 0;
    long shots_per_launch = 0;
    double device_bytes = 0;
    double h2d_bytes = 0, d2h_bytes = 0;
    int tile_tz = 0, tile_spb = 0, tile_kind = 0, tile_txm = 0, tile_zfast = 0, tile_pd = 0, tile_remap = 0;
    int tile_wpb = 0, tile_zrun = 0, layout = 0, ckpt_bits = 32;   // the reconstruction sweep's
    FdTile tiles[3];              // launch shape per stencil mode of the last batch
    FdTile tiles_acc[3][3];       // ... and of its launches with the accumulation epilogue
    long graphs_built = 0, graph_launches = 0;
    long shapes_reused = 0;  // launch shapes taken from a similar tuned geometry
};

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    long fd_launches = 0;     // stencil launches and cells one replay runs
    long fd_cells = 0;
};

struct Ctx {
    bool ready = false;
    int device = 0;
    hipStream_t st = nullptr;
    hipStream_t cst = nullptr;    // host-copy stream (records to and from the caller's arrays)
    int batch_req = 0;
    double memfrac = 0.85;
    int ktiming = 0;
    long mode_launches[3] = {0, 0, 0};
    int ckpt_bits = 32;
    FdTile tile_req = {0, 1};     // tz == 0: autotune
    int layout_req = -1;          // -1 automatic, 0 per-shot arenas, 1 shots in lanes
    int graphs = 0;               // replay time loops from captured HIP graphs
    int fma = 0;                  // stencil arithmetic: 0 strict source order, 1 FMA-contracted
    int inplace = 1;              // lanes time steps write the new field over p0 when it is a work field
    size_t mem_cap = 0;           // serving worker: device bytes its batches may use (0: free memory)
    bool capturing = false;
    std::map<std::vector<long>, GraphEntry> graph_cache;
    std::set<std::vector<long>> graph_seen;  // loops run once; captured when they come again
    int reuse_similar = 0;        // take launch shapes tuned for similar geometries (serving)
    int quick_tune = 0;           // time a short list of launch shapes, once (serving)
    std::map<std::tuple<int, int, int, int, int, int, int, int, long>, FdTile> tuned;
    std::map<std::tuple<int, int, int, int, int, int, int, int, long>, FdTile> tuned_acc;
    Stats stats;
    std::vector<hipEvent_t> ev;   // pool of events for sampled kernel timing
    size_t ev_used = 0;
    struct TimedLaunch {
        long ev;            // index of the launch's first event
        int mode;
        long cells;         // interior cell-updates (all shots of the launch)
        long shots;
        int acc_on;
    };
    std::vector<TimedLaunch> ev_cells;
    // device buffers
    DevBuf model[2];              // full vp, rho [nx][nz]
    DevBuf setup;                 // padded vp, rho, D, Beta + weights
    DevBuf coef;                  // CO_N fields
    DevBuf arena;                 // shot arenas
    DevBuf acc;                   // double accumulators
    DevBuf rec, obs;              // records of the batch
    DevBuf src;                   // wavelets of the batch
    DevBuf tables;                // receiver tables
    DevBuf taper;                 // taper tables
    DevBuf shot_out;              // per-shot gradient and pseudo-Hessian of all shots
    DevBuf gsum;                  // double sums
    DevBuf scratch;               // reductions
    DevBuf misc[4];
    DevBuf ck16;                  // FP16 checkpoint slots
    DevBuf ckaux;                 // per-slot inverse scales and max scratch
    DevBuf fftspec, fftwin, ffttmp;   // trace filters
    DevBuf adjfull;               // adjoint sources interpolated to every time step
    DevBuf stage;                 // time-major receiver values of the lanes stencil
    DevBuf accstage;              // staged accumulation terms of the lanes stencil
    DevBuf ftables;               // receiver-row maps and slot tables of the lanes stencil
    std::map<std::pair<int, int>, std::pair<void *, void *>> fft_plans;  // (n, batch) -> R2C, C2R
};

Ctx &ctx();
void bind_ctx(Ctx *c);  // make c the calling thread's engine context (nullptr: the default)
void ensure_ready();
void sync();
void h2d(void *dst, const void *src, size_t bytes);
void d2h(void *dst, const void *src, size_t bytes);

// Geometry of a padded window (engine loop bounds, propagation.c:443-537).
Geom make_geom(int nx, int nz, int pml, int fs, float dx, float dz, float dt);

std::vector<float> meal_weights(float alpha, int nl);
float meal_lnR(int nl, float pd);

// Build the 7 coefficient arrays of a window on the device.
void build_coefs(const Geom &g, const float *vpw_dev, const float *rhow_dev, float fpeak,
                 float *coef_dev);

// compute_adjoint_source2d with filter 2 on device records [ntr][ns].
void adjoint_source_device(const Record2D *r, float **vp_full, float freq, int filter,
                           const float *pre_dev, const float *obs_dev, float *out_dev);

// Sampled timing of stencil launches.
void fd_launch(int mode, const FdArgs &a, long cells, FdTile tile);
// inplace: time the shapes as the batch runs its time steps, each writing the
// new field over p0 (out == prev)
FdTile autotune_tile(int mode, const FdArgs &base, float *out, const float *cur, const float *prev,
                     bool inplace = false);
FdTile autotune_acc_tile(int mode, int accv, const FdTile &mode_tile, const FdArgs &base,
                         float *out, const float *cur, const float *prev, bool inplace = false);
void collect_timing();
void drop_graphs();

// Debug files of the reference's verbose levels, written into the working
// directory with its names: at level >= 1 iso_acoustic2d_propagation writes,
// for every shotInterval-th shot of a gradient, the predicted records, the
// adjoint source and the raw shot gradient, and after the shot loop the raw
// gradient and the pseudo-Hessian (propagation.c:917-943, 996-1005); at level
// >= 2 the engine writes the wavefields of those shots every snapInterval steps
// (:597-601, 717-723).  The intervals come from the environment variables
// shotInterval (default 20) and snapInterval (default 1000), as there
// (:423-441, 846-852).
struct Verbose {
    int level = 0, iter = 0;
    int shot_interval = 20, snap_interval = 1000;
    bool shot(int ishot) const { return level > 0 && ishot % shot_interval == 0; }
    bool snaps() const { return level > 1; }
};
Verbose verbose_from_env(int level, int iter);
// write_bin (dataio.c:277-304): n floats to a new file, exit(1) on failure.
void write_bin_file(const char *name, const float *data, size_t n);

// The shot loop of iso_acoustic2d_propagation (propagation.c:854-1035) with
// the choice of where the observed records come from and whether the
// predicted / adjoint-source records go back to the host.
struct Call {
    float **vp, **rho;          // host models (vp also feeds the adjoint-source taper)
    Record2D *obs, *pre;
    int Nshot, nz, nx;
    float dx, dz, dt;
    int nt, pml, fs;
    float freq;
    int filter, precondition, problem;
    float **grad;               // host gradient (problem 1), may be null with grad_dev
    const float *obs_dev = nullptr;   // all observed records on the device (session)
    const long *obs_off = nullptr;    // per-shot offsets into obs_dev
    bool records_to_host = true;
    float *grad_dev = nullptr;        // device gradient output (session)
    double *misfit = nullptr;         // 0.5 * sum of squared adjoint sources (problem 1/2)
    float *pre_dev = nullptr;         // device copy of every adjoint source, at obs_off
    bool models_on_device = false;    // c.model[] already hold vp, rho
    Verbose vb;                       // debug files (level 0: none)
};

void propagate_core(const Call &P);
float objective_device(const float *d, long n);
float step_length_device(const float *a, const float *b, long n);
float model_perturbation_device(float fraction, const float *sl, const float *g, float *out,
                                int len);
float model_perturbation_with(float fraction, const float *sl, const float *g, float *out, int len,
                              float *mm, int *hist, double *red);
void smooth2d_device(const float *in, float *out, int nx, int nz, int N, int zdir);
void call_footprint(const Call &P, bool ck16, size_t *fixed, size_t *per_shot);
void check_common(int nz, int nx, float dx, float dz, float dt, int nt, int pml, int fs);
size_t held_bytes(const Ctx &c);
void release_buffers(Ctx &c);

// FFT band filters (filters.cpp): kind 0 low-pass, 1 band-pass; x is
// [batch][n] on the device, filtered in place.
std::vector<float> filter_window(int kind, int n, float fpass, float fcut, float dt);
void filter_traces_device(int kind, float *x, int n, int batch, float fpass, float fcut, float dt);
// interpolate_traces (common.c:482-581): in [ntr][ns] at r times the step ->
// out [ntr][nt_out], samples (ns-1)*r+1 ... nt_out-1 zero; r <= 1 copies.
void interpolate_traces_device(const float *in, int ntr, int ns, int r, float *out, int nt_out);
void release_fft(Ctx &c);

}  // namespace fwiedu
