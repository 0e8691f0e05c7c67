// Host engine of the FWI hot path: device buffers, the shot batcher, the
// zero-copy checkpoint schedule and the reference's propagation entry points.
//
// Reference: FWILab2d/FWILab2dCLangV1.2.6/src/propagation.c (engine and shot
// loop), checkpoint.c (schedule), optimization.c (adjoint source, per-shot and
// global preconditioning).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

#include "engine.hpp"

namespace fwiedu {

// ---------------------------------------------------------------------------
// runtime

void fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    fflush(stderr);
    exit(1);
}

// Device buffers come from the device's stream-ordered memory pool.  hipFree
// performs an implicit hipDeviceSynchronize: in a serving process a worker
// whose buffers grow for a larger job would wait there for every other
// worker's queued time loops (seconds), and so would its job.  A pooled
// buffer is returned with hipFreeAsync in the order of the owner's stream,
// and the pool keeps what is returned (release threshold: everything), so
// no allocation or return waits for the device; pool_idle_bytes counts that
// memory as free where the engine sizes its batches, fwiedu_release trims it.
static std::once_flag g_pool_once;
static bool g_pool = false;
static hipMemPool_t g_mempool = nullptr;

static void pool_setup()
{
    std::call_once(g_pool_once, [] {
        int dev = 0, ok = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        if (hipDeviceGetAttribute(&ok, hipDeviceAttributeMemoryPoolsSupported, dev) != hipSuccess ||
            !ok)
            return;
        if (hipDeviceGetDefaultMemPool(&g_mempool, dev) != hipSuccess) return;
        uint64_t keep = UINT64_MAX;
        HIPCHECK(hipMemPoolSetAttribute(g_mempool, hipMemPoolAttrReleaseThreshold, &keep));
        g_pool = true;
    });
}

size_t pool_idle_bytes()
{
    if (!g_pool) return 0;
    uint64_t reserved = 0, used = 0;
    HIPCHECK(hipMemPoolGetAttribute(g_mempool, hipMemPoolAttrReservedMemCurrent, &reserved));
    HIPCHECK(hipMemPoolGetAttribute(g_mempool, hipMemPoolAttrUsedMemCurrent, &used));
    return reserved > used ? (size_t)(reserved - used) : 0;
}

void trim_pool()
{
    if (g_pool) HIPCHECK(hipMemPoolTrimTo(g_mempool, 0));
}

void *DevBuf::get(size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (bytes > cap) {
        ensure_ready();
        Ctx &c = ctx();
        // inside a stream capture an allocation or a return would become a
        // graph node: the captured time loops allocate nothing
        if (c.capturing) fail("<error>: a device buffer grew inside a captured time loop");
        release(c.st);
        if (g_pool) {
            hipError_t e = hipMallocAsync(&p, bytes, c.st);
            if (e == hipErrorOutOfMemory) {
                // the pool's idle memory may be split in pieces too small:
                // give it back to the device and try once more
                (void)hipGetLastError();
                HIPCHECK(hipStreamSynchronize(c.st));
                trim_pool();
                e = hipMallocAsync(&p, bytes, c.st);
            }
            if (e != hipSuccess)
                fail("<error>: hipMallocAsync(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
            pooled = true;
        } else {
            HIPCHECK(hipMalloc(&p, bytes));
            pooled = false;
        }
        cap = bytes;
    }
    return p;
}

void DevBuf::release(hipStream_t st)
{
    if (p) {
        if (pooled)
            (void)hipFreeAsync(p, st);
        else
            (void)hipFree(p);
    }
    p = nullptr;
    cap = 0;
    pooled = false;
}

// The engine state of the calling thread: the process-wide default context,
// or the one a serving worker bound to its thread.
static Ctx g_default;
static thread_local Ctx *t_bound = nullptr;
Ctx &ctx() { return t_bound ? *t_bound : g_default; }
void bind_ctx(Ctx *c) { t_bound = c; }

void ensure_ready()
{
    Ctx &c = ctx();
    if (c.ready) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail("<error>: no HIP device available (libfwiedu runs the FWI path on the GPU only)");
    HIPCHECK(hipGetDevice(&c.device));
    HIPCHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&c.cst, hipStreamNonBlocking));
    pool_setup();
    c.ready = true;
}

static std::vector<DevBuf *> all_buffers(Ctx &c)
{
    return {&c.model[0], &c.model[1], &c.setup,   &c.coef,    &c.arena,   &c.acc,
            &c.rec,      &c.obs,      &c.src,     &c.tables,  &c.taper,   &c.shot_out,
            &c.gsum,     &c.scratch,  &c.misc[0], &c.misc[1], &c.misc[2], &c.misc[3],
            &c.ck16,     &c.ckaux,    &c.fftspec, &c.fftwin,  &c.ffttmp,  &c.adjfull,
            &c.stage,    &c.ftables,  &c.accstage};
}

size_t held_bytes(const Ctx &c)
{
    size_t n = 0;
    for (DevBuf *d : all_buffers(const_cast<Ctx &>(c))) n += d->cap;
    return n;
}

// Free every device buffer of c (its stream must be idle); graphs that
// captured them go too.
void release_buffers(Ctx &c)
{
    for (auto &e : c.graph_cache) (void)hipGraphExecDestroy(e.second.exec);
    c.graph_cache.clear();
    c.graph_seen.clear();
    release_fft(c);
    for (DevBuf *d : all_buffers(c)) d->release(c.st);
}

void sync() { HIPCHECK(hipStreamSynchronize(ctx().st)); }

// Host-to-device copies complete before returning: sources are often
// temporaries, and pageable memory gives no asynchrony to keep.
void h2d(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return;
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().st));
    HIPCHECK(hipStreamSynchronize(ctx().st));
    ctx().stats.h2d_bytes += (double)bytes;
}

void d2h(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return;
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().st));
    ctx().stats.d2h_bytes += (double)bytes;
}

static int env_interval(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

// propagation.c:846-852 reads shotInterval on every call; the engine reads
// both intervals at verbose > 1 (:423-441).  A zero or negative interval would
// divide by zero there; here it is an error once a level needs it.
Verbose verbose_from_env(int level, int iter)
{
    Verbose v;
    v.level = level;
    v.iter = iter;
    v.shot_interval = env_interval("shotInterval", 20);
    if (level > 1) v.snap_interval = env_interval("snapInterval", 1000);
    if (level > 0 && v.shot_interval <= 0)
        fail("<error>: shotInterval must be a positive integer (got %d)", v.shot_interval);
    if (level > 1 && v.snap_interval <= 0)
        fail("<error>: snapInterval must be a positive integer (got %d)", v.snap_interval);
    return v;
}

void write_bin_file(const char *name, const float *data, size_t n)
{
    FILE *fp = fopen(name, "wb");
    if (!fp) fail("write_bin: failed to open file: %s", name);
    const size_t k = fwrite(data, sizeof(float), n, fp);
    fclose(fp);
    if (k != n) fail("write_bin: expected to write %zu elements, but only wrote %zu", n, k);
}

// n floats of device memory into a verbose file (the stream's work first).
static void write_bin_device(const char *name, const float *dev, size_t n)
{
    std::vector<float> h(n);
    HIPCHECK(hipMemcpyAsync(h.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, ctx().st));
    HIPCHECK(hipStreamSynchronize(ctx().st));
    write_bin_file(name, h.data(), n);
}

Geom make_geom(int nx, int nz, int pml, int fs, float dx, float dz, float dt)
{
    Geom g;
    g.nx = nx;
    g.nz = nz;
    g.pml = pml;
    g.fs = fs;
    g.Nx = nx + 2 * pml;
    g.Nz = nz + 2 * pml;
    g.ld = (g.Nz + 63) / 64 * 64;
    const int N = 3;
    g.ixs = N;
    g.ixe = g.Nx - N - 1;
    g.izs = (fs == 0) ? N : pml;
    g.ize = g.Nz - N - 1;
    g.dx = dx;
    g.dz = dz;
    g.dt = dt;
    g.invdx = 1.0f / dx;
    g.invdz = 1.0f / dz;
    g.invdxdx = 1.0f / (dx * dx);
    g.invdzdz = 1.0f / (dz * dz);
    g.invdtdt = 1.0f / (dt * dt);
    g.fsz = (long)g.Nx * g.ld;
    return g;
}

// absorbWeight (propagation.c:173-190) for l = 1..nlayer, host libm.
std::vector<float> meal_weights(float alpha, int nl)
{
    std::vector<float> w(nl + 1, 0.0f);
    for (int l = 1; l <= nl; ++l) {
        const float d = (0.1f * (float)l) / (0.1f * (float)nl);
        w[l] = expf(logf(2.0f) * powf(d, alpha)) - 1.0f;
    }
    return w;
}

// ln of the reflection coefficient of propagation.c:225.
float meal_lnR(int nl, float pd)
{
    const float Rcoef =
        powf(10.0f, -1.0f * ((log10f((float)nl) - 1.0f) / (log10f(2.0f))) - 3.0f - 0.5f * pd);
    return logf(Rcoef);
}

// The MEAL scalars are evaluated here with the host libm; the per-cell
// arithmetic runs on the device.
void build_coefs(const Geom &g, const float *vpw, const float *rhow, float fpeak, float *coef)
{
    Ctx &c = ctx();
    const int nl = g.pml;
    float *base = c.setup.as<float>(4 * g.fsz + nl + 1 + 64);
    float *vp = base, *rho = base + g.fsz, *D = base + 2 * g.fsz, *Beta = base + 3 * g.fsz;
    float *wdev = base + 4 * g.fsz;
    std::vector<float> w = meal_weights(2.5f, nl);
    const float pd = 2.0f;
    const float lnR = meal_lnR(nl, pd);
    const float sqrt3 = sqrtf(3.0f);
    h2d(wdev, w.data(), w.size() * sizeof(float));
    launch_pad(vpw, g.nx, g.nz, g.pml, vp, g, c.st);
    launch_pad(rhow, g.nx, g.nz, g.pml, rho, g, c.st);
    launch_meal(vp, nullptr, D, Beta, g, nl, wdev, lnR, pd, sqrt3, fpeak, c.st);
    launch_coefs(vp, rho, D, Beta, coef, g, c.st);
    launch_grad_invrho(coef, g, c.st);
    launch_coef_pack(coef, g, c.st);
}

// ---------------------------------------------------------------------------
// sampled stencil timing

void fd_launch(int mode, const FdArgs &a, long cells, FdTile tile)
{
    Ctx &c = ctx();
    if (tile.kind >= 3 && mode != FD_RECON && a.acc_on && !a.accstage)
        fail("<error>: internal: accumulating lanes stencil launch without its staging buffer");
    // every ktiming-th launch of each stencil mode, outside graph capture
    const bool timed = c.ktiming > 0 && !c.capturing && mode >= 0 && mode < 3 &&
                       (c.mode_launches[mode]++ % c.ktiming) == 0;
    if (timed) {
        if (c.ev_used + 2 > 4096) collect_timing();
        while (c.ev.size() < c.ev_used + 2) {
            hipEvent_t e;
            // timing only: no system-scope fence (cache writeback and
            // invalidation) when the event completes, which would slow the
            // following launches and add to the measured time
            HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            c.ev.push_back(e);
        }
        // the events ride in the launch's own dispatch: kernel time only
        set_stencil_events(c.ev[c.ev_used], c.ev[c.ev_used + 1]);
    }
    launch_fd(mode, a, tile, c.st);
    if (timed) {
        set_stencil_events(nullptr, nullptr);
        HIPCHECK(take_stencil_launch_error());
        c.ev_cells.push_back({(long)c.ev_used, mode, cells, (long)a.nshot, a.acc_on});
        c.ev_used += 2;
    }
    c.stats.fd_launches++;
    if (a.out == a.prev) c.stats.fd_inplace++;
    c.stats.fd_cells += cells;
    c.stats.shots_per_launch = a.nshot;
}

void collect_timing()
{
    Ctx &c = ctx();
    if (c.ev_cells.empty()) return;
    sync();
    Stats &s = c.stats;
    for (const auto &t : c.ev_cells) {
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, c.ev[t.ev], c.ev[t.ev + 1]));
        // t.cells counts interior cell-updates of all the launch's shots
        const double sc = (double)t.cells;
        s.md_timed[t.mode]++;
        s.md_ms[t.mode] += ms;
        s.md_cells[t.mode] += t.shots > 0 ? sc / (double)t.shots : 0.0;
        s.md_shot_cells[t.mode] += sc;
        if (t.acc_on) s.md_acc_shot_cells[t.mode] += sc;
        if (t.mode == FD_RECON) {
            s.fd_ms_timed += ms;
            s.fd_cells_timed += (double)t.cells;
            s.fd_timed++;
        }
    }
    c.ev_cells.clear();
    c.ev_used = 0;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner.  For each new (stencil mode, geometry,
// shots-per-launch, layout) it times the launch shapes of the layout
// (per-shot arenas: tile height, shots per block, block order, column march;
// shots in lanes: kernel, columns per block, rows per wave, rows of loads in
// flight, block order) on that mode's stencil over three distinct zero-filled
// work fields of the batch, rotating as in a time loop, with no source and no
// accumulation epilogue (the output is zero again, so nothing the passes read
// changes), and keeps the fastest.  The modes are tuned apart because their
// epilogues change the register budget: deeper prefetch pays in the
// reconstruction sweep and not in the forward and adjoint stencils; the
// launches with the accumulation epilogue are tuned apart again
// (autotune_acc_tile).

// Average time of one launch of shape t over reps launches (after one
// untimed), the three work fields rotating as in a time loop: each launch
// reads the field the previous one wrote.
static float time_shape(int mode, FdArgs a, const FdTile &t, float *const f[3], int reps,
                        hipEvent_t e0, hipEvent_t e1, bool inplace)
{
    Ctx &c = ctx();
    // three rotating fields, or (inplace) two, each step writing over p0
    auto go = [&](int r) {
        if (inplace) {
            a.cur = f[r % 2];
            a.out = f[(r + 1) % 2];
            a.prev = a.out;
        } else {
            a.cur = f[r % 3];
            a.out = f[(r + 1) % 3];
            a.prev = f[(r + 2) % 3];
        }
        launch_fd(mode, a, t, c.st);
    };
    go(0);
    HIPCHECK(hipEventRecord(e0, c.st));
    for (int r = 1; r <= reps; ++r) go(r);
    HIPCHECK(hipEventRecord(e1, c.st));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms / (float)reps;
}

// The fastest of the candidate shapes.  Every shape is timed over 9 launches
// (3 once a launch takes more than 0.2 ms); the shapes within 3 % of the
// fastest (four at most) are then timed again, twice each in alternating
// order over three times as many launches, and the best of those averages
// wins, so that shapes within a percent of each other are not told apart by
// one short sample.  inplace: timed as the batch runs its steps in place (the
// fields f[0] and f[1] only).
static FdTile fastest_shape(int mode, const FdArgs &a, const std::vector<FdTile> &cand,
                            float *const f[3], bool once, bool inplace)
{
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
    HIPCHECK(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
    if (cand.empty()) fail("<error>: internal: no launch shape to tune");
    std::vector<std::pair<float, int>> ms;
    int reps = 9;
    for (size_t i = 0; i < cand.size(); ++i) {
        const float t = time_shape(mode, a, cand[i], f, reps, e0, e1, inplace);
        ms.push_back({t, (int)i});
        if (reps == 9 && t > 0.2f) reps = 3;
    }
    std::sort(ms.begin(), ms.end());
    int top = 1;
    while (!once && top < 4 && top < (int)ms.size() && ms[top].first <= 1.03f * ms[0].first) ++top;
    if (top == 1) {
        HIPCHECK(hipEventDestroy(e0));
        HIPCHECK(hipEventDestroy(e1));
        return cand[ms[0].second];
    }
    std::vector<float> again(top, 1e30f);
    for (int pass = 0; pass < 2; ++pass)
        for (int j = 0; j < top; ++j) {
            const int k = pass ? top - 1 - j : j;
            again[k] = std::min(again[k],
                                time_shape(mode, a, cand[ms[k].second], f, 3 * reps, e0, e1,
                                           inplace));
        }
    int best = 0;
    for (int k = 1; k < top; ++k)
        if (again[k] < again[best]) best = k;
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
    return cand[ms[best].second];
}

// Shapes tuned by any context of the process, by geometry class, for callers
// that take a shape tuned for a similar geometry instead of timing their own
// (serving workers, whose requests rarely repeat a geometry exactly: tuning a
// new geometry times a few hundred launches per stencil mode, and beside the
// other workers' kernels each timed group of launches waits behind them, so
// that one new geometry cost its request seconds).  Any shape gives the same
// bits, so reuse changes the speed only.  Similar: same stencil mode,
// accumulation variant, layout, 64-lane groups (lanes) or shot-count bucket
// (arenas: 1, 2-3, 4-7, 8 or more, the shots per block a shape may take), free
// surface and arithmetic, steps in place or not, and column count and stencil
// height within a quarter of the tuned one's.
struct TunedClass {
    int mode, accv, lanes, group, fs, fma, inplace, ncol, nrow;
    FdTile tile;
};
static std::mutex g_similar_mu;
static std::vector<TunedClass> g_similar;

// The largest power of two up to 8 that is at most n: shots per block that a
// per-shot-arena shape tuned for n shots may use.
static int spb_bucket(int n) { return n >= 8 ? 8 : n >= 4 ? 4 : n >= 2 ? 2 : 1; }

static TunedClass tuned_class(int mode, int accv, const FdArgs &a, bool inplace)
{
    TunedClass t;
    t.mode = mode;
    t.accv = accv;
    t.inplace = inplace ? 1 : 0;
    t.lanes = a.cell_stride > 1 ? 1 : 0;
    t.group = t.lanes ? (a.nshot + 63) / 64 : spb_bucket(a.nshot);
    t.fs = a.fs;
    t.fma = a.fma;
    t.ncol = a.ixe - a.ixs + 1;
    t.nrow = a.ize - a.izs + 1;
    return t;
}

static bool near(int a, int b) { return 4L * std::abs(a - b) <= (long)b; }

static bool find_similar(const TunedClass &q, const FdArgs &base, FdTile &out)
{
    std::lock_guard<std::mutex> lk(g_similar_mu);
    for (auto it = g_similar.rbegin(); it != g_similar.rend(); ++it) {
        const TunedClass &t = *it;
        if (t.mode != q.mode || t.accv != q.accv || t.lanes != q.lanes || t.group != q.group ||
            t.fs != q.fs || t.fma != q.fma || t.inplace != q.inplace || !near(q.ncol, t.ncol) ||
            !near(q.nrow, t.nrow))
            continue;
        if (t.tile.kind == 1 && march_lds_bytes(base, t.tile) > 65536) continue;
        out = t.tile;
        return true;
    }
    return false;
}

static void remember_similar(TunedClass q, const FdTile &tile)
{
    std::lock_guard<std::mutex> lk(g_similar_mu);
    q.tile = tile;
    if (g_similar.size() >= 256) g_similar.erase(g_similar.begin());
    g_similar.push_back(q);
}

FdTile autotune_tile(int mode, const FdArgs &base, float *out, const float *cur, const float *prev,
                     bool inplace)
{
    Ctx &c = ctx();
    const bool lanes = base.cell_stride > 1;
    if (c.tile_req.tz > 0) {
        if ((c.tile_req.kind >= 2) != lanes)
            fail("<error>: internal: stencil shape kind %d does not match the field layout",
                 c.tile_req.kind);
        if (c.tile_req.kind == 1 && march_lds_bytes(base, c.tile_req) > 65536)
            fail("<error>: column march with %d shots per block needs more than 64 KiB of LDS "
                 "for this grid height",
                 c.tile_req.spb);
        return c.tile_req;
    }
    const auto key = std::make_tuple(base.Nx, base.ld, base.ixs, base.ixe, base.izs, base.ize,
                                     base.fs, ((base.nshot * 2 + base.fma) * 4 + mode) * 2 + inplace,
                                     base.cell_stride);
    auto it = c.tuned.find(key);
    if (it != c.tuned.end()) return it->second;
    const TunedClass cls = tuned_class(mode, 0, base, inplace);
    FdTile sim;
    if (c.reuse_similar && find_similar(cls, base, sim)) {
        c.stats.shapes_reused++;
        c.tuned[key] = sim;
        return sim;
    }
    FdArgs a = base;
    a.out = out;
    a.cur = cur;
    a.prev = prev;
    a.src_cell = nullptr;
    a.acc = nullptr;
    a.acc_on = 0;
    a.u = nullptr;
    std::vector<FdTile> cand;
    if (c.quick_tune) {
        // serving: a short list, timed once (the shapes that won most often:
        // the register ring with 20 or 30 rows and the plain lanes kernel for
        // small grids; square tiles with one or the most shots per block)
        for (int rm : {0, 1}) {
            if (lanes) {
                for (int k : {2, 3}) {
                    FdTile t;
                    t.tz = 64;
                    t.kind = 4;
                    t.wpb = 8;
                    t.pd = 2;
                    t.zrun = k * ring_unroll(2);
                    t.remap = rm;
                    cand.push_back(t);
                }
                FdTile t;
                t.tz = 64;
                t.kind = 2;
                t.wpb = 4;
                t.zrun = 8;
                t.pd = 1;
                t.remap = rm;
                cand.push_back(t);
            } else {
                // (tile rows, shots per block)
                const int b = spb_bucket(base.nshot);
                const std::vector<std::pair<int, int>> shapes =
                    b > 1 ? std::vector<std::pair<int, int>>{{64, b}, {32, b}, {64, 1}}
                          : std::vector<std::pair<int, int>>{{64, 1}, {32, 1}};
                for (const auto &sh : shapes) {
                    FdTile t;
                    t.tz = sh.first;
                    t.spb = sh.second;
                    t.remap = rm;
                    cand.push_back(t);
                }
            }
        }
        float *const f[3] = {out, const_cast<float *>(cur), const_cast<float *>(prev)};
        const FdTile best = fastest_shape(mode, a, cand, f, true, inplace);
        c.tuned[key] = best;
        remember_similar(cls, best);
        return best;
    }
    if (lanes) {
        // the shapes that won somewhere among grids of 128^2 to 3000 x 1000
        // cells and 48 to 256 shots: the LDS kernel with 8 columns per block
        // (16 were never faster) and runs of 16 to 32 rows, and for the
        // smallest grids the plain lanes kernel with 4 columns and 8 rows per
        // block
        for (int rm : {0, 1})
            for (int sgo : {0, 2}) {
                FdTile t;
                t.tz = 64;
                t.kind = 2;
                t.wpb = 4;
                t.zrun = 8;
                t.remap = rm;
                t.zfast = sgo;
                t.pd = 1;
                cand.push_back(t);
            }
        for (int zrun : {16, 20, 24, 32})
            for (int rm : {0, 1})
                for (int order : {0, 2})
                    for (int pd : {2, 4}) {
                        FdTile t;
                        t.tz = 64;
                        t.kind = 3;
                        t.zfast = order;
                        t.wpb = 8;
                        t.zrun = zrun;
                        t.remap = rm;
                        t.pd = pd;
                        cand.push_back(t);
                    }
        // the register-ring LDS stencil: runs of two or three unrolled groups
        // of 10 rows, 2 or 3 rows of loads in flight (4 or 5 rows in flight,
        // 12-row groups, never won at 221 x 601 or 400 x 1200)
        for (int pd : {2, 3})
            for (int k : {2, 3})
                for (int rm : {0, 1})
                    for (int order : {0, 2}) {
                        FdTile t;
                        t.tz = 64;
                        t.kind = 4;
                        t.zfast = order;
                        t.wpb = 8;
                        t.zrun = k * ring_unroll(pd);
                        t.remap = rm;
                        t.pd = pd;
                        cand.push_back(t);
                    }
    }
    for (int tz : {64, 32})
        for (int spb : {1, 2, 4, 8})
            for (int zf : {0, 1})
                for (int rm : {0, 1}) {
                    if (lanes) continue;
                    if (spb > 1 && spb > base.nshot) continue;
                    FdTile t;
                    t.tz = tz;
                    t.spb = spb;
                    t.zfast = zf;
                    t.remap = rm;
                    cand.push_back(t);
                }
    for (int txm : {16, 64})
        for (int spb : {1, 2})
            for (int pd : {4}) {
                if (lanes || (spb > 1 && spb > base.nshot)) continue;
                FdTile t;
                t.kind = 1;
                t.txm = txm;
                t.spb = spb;
                t.pd = pd;
                if (march_lds_bytes(base, t) > 65536) continue;
                cand.push_back(t);
            }
    float *const f[3] = {out, const_cast<float *>(cur), const_cast<float *>(prev)};
    const FdTile best = fastest_shape(mode, a, cand, f, false, inplace);
    c.tuned[key] = best;
    remember_similar(cls, best);
    return best;
}

// The launches of the shots-in-lanes LDS stencil that carry the accumulation
// epilogue (accv 1: stage this step's term, 2: add the staged terms and this
// step's to the FP64 accumulator) hold more registers than the plain ones, so
// their shape is tuned apart: rows per wave and rows in flight, with the
// block orders of the mode's own shape.  base carries the batch's accumulator
// and staging buffers, timed right after the accumulator is cleared and with
// zero work fields: the staging slot 0 receives zeros (nothing is staged yet)
// and the accumulator adds +0 terms to +0, so nothing a later launch reads
// changes.
FdTile autotune_acc_tile(int mode, int accv, const FdTile &mode_tile, const FdArgs &base,
                         float *out, const float *cur, const float *prev, bool inplace)
{
    Ctx &c = ctx();
    if (c.tile_req.tz > 0 || mode_tile.kind < 3 || mode == FD_RECON || !base.acc ||
        !base.accstage)
        return mode_tile;
    const auto key = std::make_tuple(base.Nx, base.ld, base.ixs, base.ixe, base.izs, base.ize,
                                     base.fs,
                                     (((base.nshot * 2 + base.fma) * 4 + mode) * 4 + accv) * 2 +
                                         inplace,
                                     base.cell_stride);
    auto it = c.tuned_acc.find(key);
    if (it != c.tuned_acc.end()) return it->second;
    TunedClass cls = tuned_class(mode, accv, base, inplace);
    cls.group = cls.group * 8 + mode_tile.kind;  // its candidates derive from the mode's shape
    FdTile sim;
    if (c.reuse_similar && find_similar(cls, base, sim)) {
        c.stats.shapes_reused++;
        c.tuned_acc[key] = sim;
        return sim;
    }
    FdArgs a = base;
    a.out = out;
    a.cur = cur;
    a.prev = prev;
    a.src_cell = nullptr;
    a.u = prev;
    a.acc_on = 1;
    a.acc_stage = accv == 1 ? 1 : 0;
    a.acc_slot = 0;
    a.acc_n = 0;
    std::vector<FdTile> cand;
    if (accv == 1 && mode_tile.kind == 4)
        // the register-ring stencil stages the term itself (runs of two or
        // three unrolled groups, 2 or 3 rows of loads in flight)
        for (int pd : {2, 3})
            for (int k : {2, 3}) {
                if (c.quick_tune && (pd != 2 || k != 2)) continue;
                FdTile t = mode_tile;
                t.pd = pd;
                t.zrun = k * ring_unroll(pd);
                cand.push_back(t);
            }
    for (int zrun : {16, 20, 24, 32})
        for (int pd : {2, 4}) {
            // the LDS stencil of the ring's block shape and order: the
            // launches that add the staged terms to the accumulator, and
            // staging candidates (serving: two of these, timed once)
            if (c.quick_tune && !((zrun == 16 && pd == 4) || (zrun == 24 && pd == 2))) continue;
            FdTile t = mode_tile;
            t.kind = 3;
            t.wpb = 8;
            t.zrun = zrun;
            t.pd = pd;
            cand.push_back(t);
        }
    // (timed in place, the forward field of the imaging condition is the
    // third field, as the reconstruction's stored fields are apart from the
    // adjoint's work fields)
    float *const f[3] = {out, const_cast<float *>(cur), const_cast<float *>(prev)};
    const FdTile best = fastest_shape(mode, a, cand, f, c.quick_tune != 0, inplace);
    c.tuned_acc[key] = best;
    remember_similar(cls, best);
    return best;
}

// ---------------------------------------------------------------------------
// checkpoint schedule (checkpoint.c:29-265), zero-copy on the device.
//
// The reference copies (u(kd-1), u(kd)) into storage at every multiple kd of
// d during the forward pass, and in the reverse pass recomputes a segment from
// the pair, storing every reconstructed field.  Here the stencil writes those
// fields straight into their storage slot (no copies), and a recomputed field is
// kept only when the imaging condition will read it (it % MST == 0).  The
// fields computed, and the order of the arithmetic, are the reference's.
struct Schedule {
    int nt, d;
    std::vector<int> ckslot;  // it -> forward checkpoint slot, -1 none, -2 the zero field
    int nck = 0;              // checkpoint slots
    int nring = 0;            // ring slots (it % MST == 0 inside one segment)

    explicit Schedule(int nt_) : nt(nt_)
    {
        d = std::max((int)roundf(sqrtf(2.0f * (float)(nt - 1))), 3);
        ckslot.assign(nt, -1);
        if (nt > 0) ckslot[0] = -2;
        for (int kd = d; kd <= nt - 1; kd += d) {
            ckslot[kd - 1] = nck++;
            ckslot[kd] = nck++;
        }
        nring = (d + MST - 1) / MST + 1;
    }
};

// ---------------------------------------------------------------------------
// A batch of shots sharing one simulation window (same ix_min/ix_max) and
// fpeak, i.e. the same coefficient arrays.

struct ShotIn {
    Record2D *pre;   // predicted / adjoint-source record (output)
    Record2D *obs;   // observed record (problems 1, 2)
    int index;       // shot index in the call
};

struct BatchPlan {
    Geom g;
    int S = 0;
    int nt = 0;
    bool adjoint = false;       // problem 1
    int nbuf = 0;               // fields per shot arena
    long shot_stride = 0;       // floats
    // arena field indices: adjoint/forward work, the zero field, reconstruction
    // work, then (gradient batches only) the checkpoint slots and the imaging
    // ring; with FP16 checkpoints the slots live in a separate half-precision
    // buffer and three FP32 fields R0, R1 (restart pair) and RU (imaging
    // field at a checkpoint time) take their place
    enum { A0 = 0, A1, A2, ZERO, F0, F1, F2, CK0 };
    enum { R0 = CK0, R1, RU, RING16 };
    int ring0 = 0;
};

static long fields_per_shot(const Schedule &sc, bool adjoint, bool ck16)
{
    if (!adjoint) return 4;
    return ck16 ? (BatchPlan::RING16 + sc.nring) : (BatchPlan::CK0 + sc.nck + sc.nring);
}

struct Tables {
    RcvTable rcv;
    InjTable inj;       // adjoint sources at the recording rate ([ntr][ns] per shot)
    InjTable inj_full;  // adjoint sources interpolated to every time step ([ntr][nt] per shot)
    // Shots in lanes: the receivers the LDS stencil samples and injects itself
    // (per column at most one row, in the stencil's interior, one receiver of
    // a shot per cell), as time-major stage slots rcol[ix] * Sp + lane; the
    // rest stays in the *_left tables of the separate kernels.
    int nci = 0;                  // columns with such a receiver row
    int rec_zlo = 1 << 30;        // least and greatest receiver row of rrow
    int rec_zhi = -1;
    const int *rrow = nullptr;    // [Nx]
    const int *rcol = nullptr;    // [Nx]
    const long *rec_slot = nullptr;      // [nci*Sp] record trace offset or -1
    const int *rec_sr = nullptr;         // [nci*Sp] sample rate of the slot's shot
    const long *inj_slot = nullptr;      // [nci*Sp] adjoint-source trace offset or -1
    const long *injfull_slot = nullptr;  // [nci*Sp] interpolated adjoint-source offset or -1
    RcvTable rcv_left;
    InjTable inj_left, inj_full_left;
};

// Receiver tables of the batch.  trace offsets index the batch record buffer
// (shot-major, [ntr][ns] per shot).
static Tables build_tables(const std::vector<ShotIn> &shots, const Geom &g, int nt,
                           std::vector<long> &roff, long &rec_total, std::vector<long> &foff,
                           long &full_total, int Sp_lanes)
{
    Ctx &c = ctx();
    const int S = (int)shots.size();
    roff.assign(S, 0);
    foff.assign(S, 0);
    rec_total = 0;
    full_total = 0;
    int nent = 0;
    for (int s = 0; s < S; ++s) {
        roff[s] = rec_total;
        foff[s] = full_total;
        rec_total += (long)shots[s].pre->ntr * shots[s].pre->ns;
        full_total += (long)shots[s].pre->ntr * nt;
        nent += shots[s].pre->ntr;
    }
    std::vector<int> e_shot(nent), e_cell(nent), e_sr(nent), e_itr(nent);
    std::vector<long> e_off(nent);
    int k = 0;
    for (int s = 0; s < S; ++s) {
        const Record2D *r = shots[s].pre;
        for (int itr = 0; itr < r->ntr; ++itr, ++k) {
            const int gx = r->igx[itr] + g.pml, gz = r->igz[itr] + g.pml;
            if (gx < 0 || gx >= g.Nx || gz < 0 || gz >= g.Nz)
                fail("<error>: receiver %d of shot %d lies outside the padded grid", itr, r->shotID);
            e_shot[k] = s;
            e_cell[k] = gx * g.ld + gz;
            e_sr[k] = r->sampleRate;
            e_itr[k] = itr;
            e_off[k] = roff[s] + (long)itr * r->ns;
        }
    }
    // Entries in (cell, shot) order: with shots in lanes the threads of one
    // wave then sample or update consecutive shots of one cell (one coalesced
    // segment) instead of cells one column apart.  Different (shot, cell)
    // pairs touch different field values, so their order is free; the
    // receivers of one pair keep their order (the sort is stable).
    {
        std::vector<int> pe(nent);
        for (int i = 0; i < nent; ++i) pe[i] = i;
        std::stable_sort(pe.begin(), pe.end(), [&](int x, int y) {
            return e_cell[x] != e_cell[y] ? e_cell[x] < e_cell[y] : e_shot[x] < e_shot[y];
        });
        auto perm = [&](auto &v) {
            auto w = v;
            for (int i = 0; i < nent; ++i) w[i] = v[pe[i]];
            v.swap(w);
        };
        perm(e_shot);
        perm(e_cell);
        perm(e_sr);
        perm(e_itr);
        perm(e_off);
    }
    // unique (shot, cell) for the injection, in the same order, with the
    // receivers of the pair in receiver order
    std::vector<int> u_shot, u_cell, u_beg;
    std::vector<long> u_off, u_full;
    for (int e = 0; e < nent; ++e) {
        if (e == 0 || e_cell[e] != e_cell[e - 1] || e_shot[e] != e_shot[e - 1]) {
            u_shot.push_back(e_shot[e]);
            u_cell.push_back(e_cell[e]);
            u_beg.push_back((int)u_off.size());
        }
        u_off.push_back(e_off[e]);
        u_full.push_back(foff[e_shot[e]] + (long)e_itr[e] * nt);
    }
    u_beg.push_back((int)u_off.size());
    const int nu = (int)u_cell.size();

    // receivers the lanes stencil takes over (see Tables)
    std::vector<int> rrow, rcol;
    std::vector<char> e_fused(nent, 0), u_fused(nu, 0);
    int nci = 0;
    if (Sp_lanes > 0) {
        // per column the interior row with the most receivers (the first of
        // equal counts): entries are in cell order, so a column's rows come
        // in runs of increasing row
        rrow.assign(g.Nx, -1);
        rcol.assign(g.Nx, -1);
        std::vector<int> rcnt(g.Nx, 0);
        for (int e = 0; e < nent;) {
            int f = e + 1;
            while (f < nent && e_cell[f] == e_cell[e]) ++f;
            const int gx = e_cell[e] / g.ld, gz = e_cell[e] % g.ld;
            if (gx >= g.ixs && gx <= g.ixe && gz >= g.izs && gz <= g.ize && f - e > rcnt[gx]) {
                rcnt[gx] = f - e;
                rrow[gx] = gz;
            }
            e = f;
        }
        for (int x = 0; x < g.Nx; ++x)
            if (rrow[x] >= 0) rcol[x] = nci++;
        // one receiver per (shot, cell) in a slot: entries are in (cell, shot)
        // order, so a repeat follows its first
        for (int e = 0; e < nent; ++e) {
            const int gx = e_cell[e] / g.ld, gz = e_cell[e] % g.ld;
            const bool repeat = e > 0 && e_cell[e - 1] == e_cell[e] && e_shot[e - 1] == e_shot[e];
            e_fused[e] = rrow[gx] == gz && !repeat;
        }
        for (int u = 0; u < nu; ++u) {
            const int gx = u_cell[u] / g.ld, gz = u_cell[u] % g.ld;
            u_fused[u] = rrow[gx] == gz && u_beg[u + 1] - u_beg[u] == 1;
        }
    }
    // one device block: ints then longs
    const size_t ni = (size_t)3 * nent + 2 * nu + nu + 1;
    const size_t nl = (size_t)nent + 2 * u_off.size();
    char *base = c.tables.as<char>(ni * sizeof(int) + nl * sizeof(long) + 64);
    long *lb = reinterpret_cast<long *>(base);
    int *ib = reinterpret_cast<int *>(base + nl * sizeof(long));
    std::vector<long> lh(nl);
    std::copy(e_off.begin(), e_off.end(), lh.begin());
    std::copy(u_off.begin(), u_off.end(), lh.begin() + nent);
    std::copy(u_full.begin(), u_full.end(), lh.begin() + nent + u_off.size());
    std::vector<int> ih;
    ih.reserve(ni);
    ih.insert(ih.end(), e_shot.begin(), e_shot.end());
    ih.insert(ih.end(), e_cell.begin(), e_cell.end());
    ih.insert(ih.end(), e_sr.begin(), e_sr.end());
    ih.insert(ih.end(), u_shot.begin(), u_shot.end());
    ih.insert(ih.end(), u_cell.begin(), u_cell.end());
    ih.insert(ih.end(), u_beg.begin(), u_beg.end());
    h2d(lb, lh.data(), nl * sizeof(long));
    h2d(ib, ih.data(), ih.size() * sizeof(int));
    Tables t;
    t.rcv.n = nent;
    t.rcv.trace_off = lb;
    t.rcv.shot = ib;
    t.rcv.cell = ib + nent;
    t.rcv.sr = ib + 2 * nent;
    t.inj.n = nu;
    t.inj.trace_off = lb + nent;
    t.inj.shot = ib + 3 * nent;
    t.inj.cell = ib + 3 * nent + nu;
    t.inj.beg = ib + 3 * nent + 2 * nu;
    t.inj_full = t.inj;
    t.inj_full.trace_off = lb + nent + u_off.size();
    t.rcv_left = t.rcv;
    t.inj_left = t.inj;
    t.inj_full_left = t.inj_full;
    if (nci == 0) return t;

    // the stencil's slots, and the left-over tables of the separate kernels
    const long nslot = (long)nci * Sp_lanes;
    std::vector<long> rslot(nslot, -1), islot(nslot, -1), fslot(nslot, -1);
    std::vector<int> rsr(nslot, 1);
    std::vector<int> le_shot, le_cell, le_sr;
    std::vector<long> le_off;
    for (int e = 0; e < nent; ++e) {
        if (e_fused[e]) {
            const long sl = (long)rcol[e_cell[e] / g.ld] * Sp_lanes + e_shot[e];
            rslot[sl] = e_off[e];
            rsr[sl] = e_sr[e];
        } else {
            le_shot.push_back(e_shot[e]);
            le_cell.push_back(e_cell[e]);
            le_sr.push_back(e_sr[e]);
            le_off.push_back(e_off[e]);
        }
    }
    std::vector<int> lu_shot, lu_cell, lu_beg;
    std::vector<long> lu_off, lu_full;
    for (int u = 0; u < nu; ++u) {
        if (u_fused[u]) {
            const long sl = (long)rcol[u_cell[u] / g.ld] * Sp_lanes + u_shot[u];
            islot[sl] = u_off[u_beg[u]];
            fslot[sl] = u_full[u_beg[u]];
        } else {
            lu_shot.push_back(u_shot[u]);
            lu_cell.push_back(u_cell[u]);
            lu_beg.push_back((int)lu_off.size());
            for (int e = u_beg[u]; e < u_beg[u + 1]; ++e) {
                lu_off.push_back(u_off[e]);
                lu_full.push_back(u_full[e]);
            }
        }
    }
    lu_beg.push_back((int)lu_off.size());
    const int nle = (int)le_cell.size(), nlu = (int)lu_cell.size();
    // longs: 3 slot tables, left-over record offsets, injection offsets (both
    // rates); ints: rrow, rcol, slot sample rates, left-over tables
    const size_t fl = (size_t)3 * nslot + nle + 2 * lu_off.size();
    const size_t fi = (size_t)2 * g.Nx + nslot + 3 * nle + 2 * nlu + nlu + 1;
    char *fb = c.ftables.as<char>(fl * sizeof(long) + fi * sizeof(int) + 64);
    long *fL = reinterpret_cast<long *>(fb);
    int *fI = reinterpret_cast<int *>(fb + fl * sizeof(long));
    std::vector<long> hl;
    hl.reserve(fl);
    hl.insert(hl.end(), rslot.begin(), rslot.end());
    hl.insert(hl.end(), islot.begin(), islot.end());
    hl.insert(hl.end(), fslot.begin(), fslot.end());
    hl.insert(hl.end(), le_off.begin(), le_off.end());
    hl.insert(hl.end(), lu_off.begin(), lu_off.end());
    hl.insert(hl.end(), lu_full.begin(), lu_full.end());
    std::vector<int> hi;
    hi.reserve(fi);
    hi.insert(hi.end(), rrow.begin(), rrow.end());
    hi.insert(hi.end(), rcol.begin(), rcol.end());
    hi.insert(hi.end(), rsr.begin(), rsr.end());
    hi.insert(hi.end(), le_shot.begin(), le_shot.end());
    hi.insert(hi.end(), le_cell.begin(), le_cell.end());
    hi.insert(hi.end(), le_sr.begin(), le_sr.end());
    hi.insert(hi.end(), lu_shot.begin(), lu_shot.end());
    hi.insert(hi.end(), lu_cell.begin(), lu_cell.end());
    hi.insert(hi.end(), lu_beg.begin(), lu_beg.end());
    h2d(fL, hl.data(), hl.size() * sizeof(long));
    h2d(fI, hi.data(), hi.size() * sizeof(int));
    t.nci = nci;
    for (int x = 0; x < (int)rrow.size(); ++x)
        if (rrow[x] >= 0) {
            t.rec_zlo = std::min(t.rec_zlo, rrow[x]);
            t.rec_zhi = std::max(t.rec_zhi, rrow[x]);
        }
    t.rec_slot = fL;
    t.inj_slot = fL + nslot;
    t.injfull_slot = fL + 2 * nslot;
    t.rrow = fI;
    t.rcol = fI + g.Nx;
    t.rec_sr = fI + 2 * g.Nx;
    const int *li = fI + 2 * g.Nx + nslot;
    t.rcv_left.n = nle;
    t.rcv_left.trace_off = fL + 3 * nslot;
    t.rcv_left.shot = li;
    t.rcv_left.cell = li + nle;
    t.rcv_left.sr = li + 2 * nle;
    t.inj_left.n = nlu;
    t.inj_left.trace_off = fL + 3 * nslot + nle;
    t.inj_left.shot = li + 3 * nle;
    t.inj_left.cell = li + 3 * nle + nlu;
    t.inj_left.beg = li + 3 * nle + 2 * nlu;
    t.inj_full_left = t.inj_left;
    t.inj_full_left.trace_off = fL + 3 * nslot + nle + lu_off.size();
    return t;
}

static long interior_cells(const Geom &g) { return (long)(g.ixe - g.ixs + 1) * (g.ize - g.izs + 1); }

static FdArgs base_args(const Geom &g, const float *coef, int S, long shot_stride,
                        long cell_stride)
{
    FdArgs a;
    memset(&a, 0, sizeof(a));
    a.coef = coef;
    a.shot_stride = shot_stride;
    a.cell_stride = cell_stride;
    a.acc_cell = 1;
    a.nshot = S;
    a.src_zlo = a.rec_zlo = 1 << 30;
    a.src_zhi = a.rec_zhi = -1;
    a.Nx = g.Nx;
    a.ld = g.ld;
    a.ixs = g.ixs;
    a.ixe = g.ixe;
    a.izs = g.izs;
    a.ize = g.ize;
    a.pml = g.pml;
    a.fs = g.fs;
    a.nx = g.nx;
    a.nz = g.nz;
    a.invdx = g.invdx;
    a.invdz = g.invdz;
    a.invdxdx = g.invdxdx;
    a.invdzdz = g.invdzdz;
    a.invdtdt = g.invdtdt;
    a.dt = g.dt;
    a.remap = 0;
    a.fma = ctx().fma;
    return a;
}

static inline bool is_work(float *const w[3], const float *x)
{
    return x == w[0] || x == w[1] || x == w[2];
}

static inline float *pick_free(float *const w[3], const float *x, const float *y)
{
    for (int i = 0; i < 3; ++i)
        if (w[i] != x && w[i] != y) return w[i];
    fail("<error>: internal buffer rotation");
}

// Taper length and cosine table of taper_traces (common.c:796-895).  L < 0 on
// return means no taper; a requested length clamped to 0 keeps the reference's
// 0/0 table entry.
static std::vector<float> taper_table(int &L, int n)
{
    std::vector<float> t;
    if (L <= 0) {
        L = -1;
        return t;
    }
    if (L > n / 3) L = n / 3;
    t.resize(L + 1);
    const float PIf = (float)3.141592653589793;
    for (int i = 0; i <= L; ++i) t[i] = 0.5f * (1.0f + cosf(PIf * (float)i / (float)L));
    return t;
}

// compute_adjoint_source2d (optimization.c:300-419) for the shots of a batch:
// scalars and taper tables on the host (every shot's tables in one upload, so
// the shots need no host round trip each), residual, tapers and filters on
// the device.  pre/obs/out + off[k] are shot k's [ntr][ns] records.
static void adjoint_sources(const Record2D *const *r, int n, float **vp_full, float freq,
                            int filter, const float *pre, const float *obs, float *out,
                            const long *off)
{
    Ctx &c = ctx();
    std::vector<float> tab;
    std::vector<long> oT(n), oX(n);
    std::vector<int> Lt(n), Lx(n);
    std::vector<float> fr(n), kp(n);
    for (int k = 0; k < n; ++k) {
        const int ntr = r[k]->ntr, ns = r[k]->ns;
        double vsum = 0.0;
        for (int itr = 0; itr < ntr; ++itr) vsum += vp_full[r[k]->igx[itr]][r[k]->igz[itr]];
        const float vavg = (float)(vsum / ntr);
        fr[k] = freq == 0.0f ? r[k]->fpeak : freq;
        kp[k] = fr[k] / vavg;
        Lt[k] = (int)roundf((2.0f / fr[k]) / r[k]->dt) + 1;
        Lx[k] = (int)roundf((2.0f / kp[k]) / r[k]->dtr) + 1;
        const std::vector<float> cT = taper_table(Lt[k], ns), cX = taper_table(Lx[k], ntr);
        oT[k] = (long)tab.size();
        tab.insert(tab.end(), cT.begin(), cT.end());
        oX[k] = (long)tab.size();
        tab.insert(tab.end(), cX.begin(), cX.end());
    }
    // the table buffer is only rewritten by a later call, in stream order
    float *tb = c.taper.as<float>(tab.size() + 64);
    if (!tab.empty()) h2d(tb, tab.data(), tab.size() * sizeof(float));
    for (int k = 0; k < n; ++k) {
        const int ntr = r[k]->ntr, ns = r[k]->ns;
        const float *tT = tb + oT[k], *tX = tb + oX[k];
        const float *pk = pre + off[k], *ok = obs + off[k];
        float *dk = out + off[k];
        if (filter == 0 || filter == 1) {
            // optimization.c:368-415: taper in t, filter in t (freq, 2 freq),
            // taper in x, transpose, filter in x (kpass, 2 kpass), transpose back
            launch_adjsrc(pk, ok, dk, ntr, ns, tT, Lt[k], tX, -1, c.st);
            filter_traces_device(filter, dk, ns, ntr, fr[k], 2.0f * fr[k], r[k]->dt);
            launch_taper(dk, ntr, ns, tT, -1, tX, Lx[k], c.st);
            float *tr = c.ffttmp.as<float>((size_t)ntr * ns + 64);
            launch_transpose(dk, tr, ntr, ns, c.st);
            filter_traces_device(filter, tr, ntr, ns, kp[k], 2.0f * kp[k], r[k]->dtr);
            launch_transpose(tr, dk, ns, ntr, c.st);
            // optimization.c:417-418: the final taper in t and x
            launch_taper(dk, ntr, ns, tT, Lt[k], tX, Lx[k], c.st);
        } else {
            launch_adjsrc(pk, ok, dk, ntr, ns, tT, Lt[k], tX, Lx[k], c.st);
        }
    }
}

// compute_adjoint_source2d of one shot; the result is complete on return.
void adjoint_source_device(const Record2D *r, float **vp_full, float freq, int filter,
                           const float *pre_dev, const float *obs_dev, float *out_dev)
{
    const long off = 0;
    adjoint_sources(&r, 1, vp_full, freq, filter, pre_dev, obs_dev, out_dev, &off);
    sync();
}

// Device-to-device copies of n pieces (piece k: bytes[k] from src + soff[k] to
// dst + doff[k], offsets in floats), pieces that continue each other on both
// sides merged into one copy.
static void copy_pieces(float *dst, const long *doff, const float *src, const long *soff,
                        const long *len, int n)
{
    Ctx &c = ctx();
    for (int k = 0; k < n;) {
        long m = len[k];
        int j = k + 1;
        while (j < n && doff[j] == doff[k] + m && soff[j] == soff[k] + m) m += len[j++];
        if (m > 0)
            HIPCHECK(hipMemcpyAsync(dst + doff[k], src + soff[k], (size_t)m * sizeof(float),
                                    hipMemcpyDeviceToDevice, c.st));
        k = j;
    }
}

// Copy a host record [ntr][ns] (contiguous block traces_p[0]) to / from the device.
static void rec_h2d(float *dst, const Record2D *r)
{
    h2d(dst, r->traces_p[0], (size_t)r->ntr * r->ns * sizeof(float));
}
static void rec_d2h(Record2D *r, const float *src)
{
    d2h(r->traces_p[0], src, (size_t)r->ntr * r->ns * sizeof(float));
}

// Copies between a batch's device records and the caller's pageable record
// arrays, on the context's copy stream from a helper thread, so that they run
// while the calling thread enqueues, and the device runs, the stencil
// launches: the observed records go up during the forward pass (which does not
// read them) and the adjoint sources come down during the adjoint pass (which
// only reads them).  The copies start once the work enqueued on the compute
// stream before `after` is done; join() returns when they have completed.
// The copy stream's wait for `after` is enqueued here, on the calling thread,
// before the batch's time loop may begin a graph capture on the compute
// stream: a wait issued by the helper thread could land while that stream is
// capturing, and a stream that waits on an event of a capturing stream joins
// the capture, where the helper's copies and stream synchronisation are not
// allowed (the capture then ends with an error).
struct RecordCopies {
    struct Piece {
        void *dst;
        const void *src;
        size_t bytes;
    };
    std::thread th;
    hipEvent_t after = nullptr;

    void start(Ctx &c, std::vector<Piece> pieces, hipMemcpyKind kind)
    {
        join();
        HIPCHECK(hipEventCreateWithFlags(&after, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(after, c.st));
        for (const Piece &p : pieces)
            (kind == hipMemcpyHostToDevice ? c.stats.h2d_bytes : c.stats.d2h_bytes) +=
                (double)p.bytes;
        const int dev = c.device;
        hipStream_t cs = c.cst;
        HIPCHECK(hipStreamWaitEvent(cs, after, 0));
        th = std::thread([dev, cs, kind, pieces] {
            HIPCHECK(hipSetDevice(dev));
            for (const Piece &p : pieces)
                HIPCHECK(hipMemcpyAsync(p.dst, p.src, p.bytes, kind, cs));
            HIPCHECK(hipStreamSynchronize(cs));
        });
    }
    void join()
    {
        if (th.joinable()) th.join();
        if (after) HIPCHECK(hipEventDestroy(after));
        after = nullptr;
    }
    ~RecordCopies() { join(); }
};

// ---------------------------------------------------------------------------
// One batch of S shots on the device.

// Field layouts.  Per-shot arenas: field k of shot s is the [Nx][ld] array at
// arena + s*SS + k*KS (SS = fields per shot * fsz, KS = fsz, cell stride 1).
// Shots in lanes: the batch is padded to Sp = 64*ceil(S/64) shots and field k
// is [Nx][ld][Sp] at arena + k*KS (KS = fsz*Sp); cell i of shot s is at
// s + i*Sp (SS = 1, CS = Sp).  Accumulators follow: [S][nx*nz] or
// [nx*nz][Sp].
struct BatchState {
    Geom g;
    const float *coef = nullptr;
    int S = 0, nt = 0;
    bool adjoint = false;
    bool lanes = false;
    bool ck16 = false;         // FP16 checkpoint slots
    int Sp = 0;                // shots the arena holds (S, or S padded to 64)
    Schedule *sc = nullptr;
    long KS = 0, SS = 0, CS = 1;  // field, shot and cell strides (floats)
    long ASS = 0, ACS = 1;        // accumulator shot and cell strides (doubles)
    long arena_floats = 0;
    float *arena = nullptr;
    __half *h16 = nullptr;        // FP16 slots: slot k of shot s, cell i at k*KS16 + s*SS16 + i*CS16
    long KS16 = 0, SS16 = 0, CS16 = 1;
    float *ckinv = nullptr;       // [nck][Sp] inverse scales
    unsigned *ckmx = nullptr;     // [ck_mx_words(Sp)] scratch
    int ring0 = 0;                // first imaging-ring field
    int *cell_dev = nullptr;
    int src_zlo = 1 << 30, src_zhi = -1;  // least and greatest row of the source cells
    float *wav_dev = nullptr;
    Tables tb;
    std::vector<long> roff;
    long rec_total = 0;
    float *rec = nullptr;      // predicted records, then the adjoint source
    std::vector<long> foff;    // per-shot offsets of the interpolated adjoint sources
    long full_total = 0;
    float *adjfull = nullptr;  // adjoint sources at every time step (sampleRate > 1), or null
    double *acc = nullptr;     // [S][nx*nz]
    long nacc = 0;
    long cells_per_launch = 0;
    float *accstage = nullptr; // [ACC_STAGES][nx*nz][Sp] staged accumulation terms (lanes)
    long accstage_stride = 0;
    float *stage = nullptr;    // [nt][nslot] receiver values of the lanes stencil
    long nslot = 0;
    FdTile tile_m[3];          // launch shape per stencil mode (FD_FWD, FD_RECON, FD_ADJ)
    FdTile tile_acc[3][3];     // [mode][0 plain, 1 staging, 2 adding to the accumulator]
    const FdTile &tile_for(int mode, const FdArgs &a) const
    {
        return tile_acc[mode][!a.acc_on ? 0 : (a.acc_stage ? 1 : 2)];
    }
    FdTile tile;               // the reconstruction sweep's (reported)
    // verbose >= 2: wavefield files of the batch's shots whose index in the
    // call (snap_shot[k], -1 for none) is a multiple of shotInterval
    const Verbose *vb = nullptr;
    std::vector<int> snap_shot;
    float *F(int k) const { return arena + (long)k * KS; }
};

// The padded field f (per-shot arenas) of every selected shot of the batch, as
// the reference's engine writes p1, p1_fwd or p0..p2 at verbose > 1
// (propagation.c:597-601, 717-723): [Nx][Nz] floats.  The arena stencils
// write the full free-surface mirror of the cells they compute; what the
// reference's acoustic_iso2d_freesurface (:76-101) also writes is added here
// when the field went through it (fs_applied): -0.0 in rows 0..pml of the six
// edge columns no stencil writes (T9), and in the mirror rows of rows the
// stencil never reaches (pml + i > Nz - 4, models shallower than pml + 4).
static void snapshot(const BatchState &b, const char *what, int it, const float *f,
                     bool fs_applied)
{
    Ctx &c = ctx();
    const Geom &g = b.g;
    std::vector<float> h((size_t)g.Nx * g.Nz);
    char name[512];
    for (int k = 0; k < b.S; ++k) {
        const int ishot = b.snap_shot[k];
        if (ishot < 0) continue;
        HIPCHECK(hipMemcpy2DAsync(h.data(), g.Nz * sizeof(float), f + (long)k * b.SS,
                                  g.ld * sizeof(float), g.Nz * sizeof(float), g.Nx,
                                  hipMemcpyDeviceToHost, c.st));
        HIPCHECK(hipStreamSynchronize(c.st));
        if (fs_applied && g.fs) {
            for (int ix = 0; ix < g.Nx; ++ix) {
                float *col = h.data() + (size_t)ix * g.Nz;
                const bool edge = ix < g.ixs || ix > g.ixe;
                for (int i = 0; i <= g.pml; ++i)
                    if (edge || g.pml + i > g.ize) col[g.pml - i] = -0.0f;
            }
        }
        snprintf(name, sizeof(name), "%s_iter%03d_ishot%04d_it%05d_%d_%d_%.2fm_%.2fm.bin", what,
                 b.vb->iter, ishot, it, g.Nz, g.Nx, g.dz, g.dx);
        write_bin_file(name, h.data(), h.size());
    }
}

static bool snap_step(const BatchState &b, int it)
{
    return b.vb && b.vb->snaps() && it % b.vb->snap_interval == 0;
}

// Time steps of the shots-in-lanes stencils may write the new field over p0:
// every lanes stencil reads p0 only at the cell it updates, before it stores
// that cell.  Only a work field is overwritten (never the zero field, a
// checkpoint or a stored imaging field), and not when wavefield snapshots
// are written.  It pays where the two work fields fit the 256 MB Infinity
// Cache, which then holds them from one step to the next (the in-place
// launches load p0 and store the output at the default cache policy); larger
// batches keep three rotating fields streamed past the caches (221 x 601 with
// 120 shots: 2 x 111.6 MB in place, with the shapes timed that way, 6 % off a
// gradient; 400 x 1200 with 256 shots, 2 x 671 MB, in place and at the default
// policy: 148 against 156-162 gradient shots/s).
static constexpr size_t INFINITY_CACHE_BYTES = (size_t)256 << 20;
static bool in_place(const BatchState &b)
{
    return ctx().inplace && b.lanes && !(b.vb && b.vb->snaps()) &&
           (size_t)2 * b.KS * sizeof(float) <= INFINITY_CACHE_BYTES;
}

static void prepare_batch(BatchState &b, const Geom &g, const float *coef,
                          const std::vector<Record2D *> &recs, int nt, bool adjoint, Schedule &sc,
                          bool lanes, bool ck16)
{
    Ctx &c = ctx();
    b.g = g;
    b.coef = coef;
    b.S = (int)recs.size();
    b.nt = nt;
    b.adjoint = adjoint;
    b.sc = &sc;
    b.lanes = lanes;
    b.src_zlo = 1 << 30;
    b.src_zhi = -1;
    b.ck16 = ck16 && adjoint && sc.nck > 0;
    b.ring0 = b.ck16 ? BatchPlan::RING16 : BatchPlan::CK0 + sc.nck;
    const int S = b.S;
    const long nf = fields_per_shot(sc, adjoint, b.ck16);
    b.nacc = (long)g.nx * g.nz;
    if (lanes) {
        b.Sp = (S + 63) / 64 * 64;
        b.KS = g.fsz * b.Sp;
        b.SS = 1;
        b.CS = b.Sp;
        b.ASS = 1;
        b.ACS = b.Sp;
    } else {
        b.Sp = S;
        b.KS = g.fsz;
        b.SS = nf * g.fsz;
        b.CS = 1;
        b.ASS = b.nacc;
        b.ACS = 1;
    }
    b.arena_floats = nf * g.fsz * b.Sp;
    b.arena = c.arena.as<float>((size_t)b.arena_floats);
    HIPCHECK(hipMemsetAsync(b.arena, 0, (size_t)b.arena_floats * sizeof(float), c.st));
    if (b.ck16) {
        b.h16 = c.ck16.as<__half>((size_t)sc.nck * g.fsz * b.Sp);
        if (lanes) {
            b.KS16 = g.fsz * b.Sp;
            b.SS16 = 1;
            b.CS16 = b.Sp;
        } else {
            b.KS16 = g.fsz;
            b.SS16 = (long)sc.nck * g.fsz;
            b.CS16 = 1;
        }
        char *aux = c.ckaux.as<char>(((size_t)sc.nck * b.Sp + ck_mx_words(b.Sp)) * 4 + 256);
        b.ckinv = reinterpret_cast<float *>(aux);
        b.ckmx = reinterpret_cast<unsigned *>(aux + (size_t)sc.nck * b.Sp * 4);
    }

    std::vector<int> cells(b.Sp, -1);
    std::vector<float> wav((size_t)S * nt);
    for (int s = 0; s < S; ++s) {
        const Record2D *r = recs[s];
        const int isx = r->isx + g.pml, isz = r->isz + g.pml;
        if (isx < g.ixs || isx > g.ixe || isz < g.izs || isz > g.ize)
            fail("<error>: source of shot %d at padded (%d,%d) is outside the stencil region "
                 "[%d,%d]x[%d,%d]",
                 r->shotID, isx, isz, g.ixs, g.ixe, g.izs, g.ize);
        if (r->sampleRate < 1) fail("<error>: shot %d has sampleRate %d", r->shotID, r->sampleRate);
        if (r->ns < (nt - 1) / r->sampleRate + 1 || r->ntr < 1 || !r->src || !r->traces_p)
            fail("<error>: shot %d: record needs ntr >= 1 and ns >= (nt-1)/sampleRate+1 = %d "
                 "samples (has ntr=%d, ns=%d)",
                 r->shotID, (nt - 1) / r->sampleRate + 1, r->ntr, r->ns);
        cells[s] = isx * g.ld + isz;
        b.src_zlo = std::min(b.src_zlo, isz);
        b.src_zhi = std::max(b.src_zhi, isz);
        memcpy(&wav[(size_t)s * nt], r->src, nt * sizeof(float));
    }
    char *sb = c.src.as<char>(wav.size() * sizeof(float) + cells.size() * sizeof(int) + 256);
    b.wav_dev = reinterpret_cast<float *>(sb);
    b.cell_dev = reinterpret_cast<int *>(sb + wav.size() * sizeof(float));
    h2d(b.wav_dev, wav.data(), wav.size() * sizeof(float));
    h2d(b.cell_dev, cells.data(), cells.size() * sizeof(int));

    std::vector<ShotIn> si(S);
    for (int s = 0; s < S; ++s) si[s] = {recs[s], nullptr, s};
    b.tb = build_tables(si, g, nt, b.roff, b.rec_total, b.foff, b.full_total,
                        lanes ? b.Sp : 0);
    if (b.tb.nci > 0) {
        b.nslot = (long)b.tb.nci * b.Sp;
        b.stage = c.stage.as<float>((size_t)nt * b.nslot + 64);
    }
    b.rec = c.rec.as<float>(b.rec_total + 64);
    HIPCHECK(hipMemsetAsync(b.rec, 0, b.rec_total * sizeof(float), c.st));
    if (adjoint) {
        b.acc = c.acc.as<double>((size_t)b.Sp * b.nacc);
        HIPCHECK(hipMemsetAsync(b.acc, 0, (size_t)b.Sp * b.nacc * sizeof(double), c.st));
        if (lanes) {
            b.accstage_stride = (long)b.Sp * b.nacc;
            b.accstage = c.accstage.as<float>((size_t)ACC_STAGES * b.accstage_stride + 64);
        }
    }
    b.cells_per_launch = interior_cells(g) * S;
    // the stencil modes the batch launches are tuned (without an adjoint pass,
    // the forward alone)
    for (int m = 0; m < 3; ++m) {
        if (!adjoint && m != FD_FWD) {
            b.tile_m[m] = b.tile_m[FD_FWD];
            for (int v = 0; v < 3; ++v) b.tile_acc[m][v] = b.tile_m[m];
            continue;
        }
        b.tile_m[m] = autotune_tile(m, base_args(g, coef, S, b.SS, b.CS), b.F(BatchPlan::A2),
                                    b.F(BatchPlan::A0), b.F(BatchPlan::A1), in_place(b));
        c.stats.tiles[m] = b.tile_m[m];
        FdArgs aa = base_args(g, coef, S, b.SS, b.CS);
        aa.acc = b.acc;
        aa.acc_stride = b.ASS;
        aa.acc_cell = b.ACS;
        aa.accstage = b.accstage;
        aa.accstage_stride = b.accstage_stride;
        b.tile_acc[m][0] = b.tile_m[m];
        for (int v = 1; v <= 2; ++v)
            b.tile_acc[m][v] = autotune_acc_tile(m, v, b.tile_m[m], aa, b.F(BatchPlan::A2),
                                                 b.F(BatchPlan::A0), b.F(BatchPlan::A1),
                                                 in_place(b));
        for (int v = 0; v < 3; ++v) c.stats.tiles_acc[m][v] = b.tile_acc[m][v];
    }
    b.tile = b.tile_m[adjoint ? FD_RECON : FD_FWD];
    c.stats.tile_tz = b.tile.tz;
    c.stats.tile_spb = b.tile.spb;
    c.stats.tile_kind = b.tile.kind;
    c.stats.tile_zfast = b.tile.zfast;
    c.stats.tile_pd = b.tile.pd;
    c.stats.tile_remap = b.tile.remap;
    c.stats.tile_txm = b.tile.txm;
    c.stats.tile_wpb = b.tile.wpb;
    c.stats.tile_zrun = b.tile.zrun;
    c.stats.layout = lanes ? 1 : 0;
    c.stats.ckpt_bits = b.ck16 ? 16 : 32;
}

// FP16 checkpoint slot k <-> FP32 field f of the batch.
static void ck_store(BatchState &b, int k, const float *f)
{
    launch_ck_pack(f, b.SS, b.CS, b.h16 + (long)k * b.KS16, b.SS16, b.CS16, b.g.fsz, b.S,
                   b.lanes ? 1 : 0, b.ckmx, b.ckinv + (long)k * b.Sp, ctx().st);
}
static void ck_load(BatchState &b, int k, float *f)
{
    launch_ck_unpack(b.h16 + (long)k * b.KS16, b.SS16, b.CS16, f, b.SS, b.CS, b.g.fsz, b.S,
                     b.lanes ? 1 : 0, b.ckinv + (long)k * b.Sp, ctx().st);
}

// Deferred accumulation of an accumulating launch: stage its term while a
// staging slot is free and a later accumulation step follows, else add the
// staged terms and its own.
static void set_deferral(FdArgs &a, bool defer, int &staged, bool last)
{
    a.acc_stage = 0;
    a.acc_n = 0;
    if (!a.acc_on || !defer) return;
    if (staged < ACC_STAGES && !last) {
        a.acc_stage = 1;
        a.acc_slot = staged++;
    } else {
        a.acc_n = staged;
        staged = 0;
    }
}

// propagation.c:539-608.  With save_ckpt the fields u(kd-1), u(kd) land in
// their checkpoint slots; with precond the pseudo-Hessian accumulates in acc.
// Kernel launches only, so a HIP graph can capture it.
static void forward_loop(BatchState &b, bool save_ckpt, bool precond)
{
    Ctx &c = ctx();
    const Schedule &sc = *b.sc;
    FdArgs a = base_args(b.g, b.coef, b.S, b.SS, b.CS);
    a.src_cell = b.cell_dev;
    a.src_zlo = b.src_zlo;
    a.src_zhi = b.src_zhi;
    a.src_wav = b.wav_dev;
    a.nt = b.nt;
    a.acc = b.acc;
    a.acc_stride = b.ASS;
    a.acc_cell = b.ACS;
    // the LDS stencil samples the receivers of its stage slots itself
    const bool fused = b.tb.nci > 0 && b.tile_m[FD_FWD].kind >= 3;
    if (fused) {
        a.rrow = b.tb.rrow;
        a.rcol = b.tb.rcol;
        a.rec_zlo = b.tb.rec_zlo;
        a.rec_zhi = b.tb.rec_zhi;
        a.stage = b.stage;
        a.nslot = b.nslot;
        a.rec_on = 1;
    }
    const RcvTable &rt = fused ? b.tb.rcv_left : b.tb.rcv;
    // the LDS and register-ring stencils stage the pseudo-Hessian terms of
    // ACC_STAGES steps and the LDS stencil adds them with the next one's (one
    // pass over the FP64 accumulator instead of ACC_STAGES + 1)
    const bool defer = precond && b.accstage && b.tile_m[FD_FWD].kind >= 3;
    a.accstage = b.accstage;
    a.accstage_stride = b.accstage_stride;
    int staged = 0;
    float *W[3] = {b.F(BatchPlan::A0), b.F(BatchPlan::A1), b.F(BatchPlan::A2)};
    const float *zero = b.F(BatchPlan::ZERO);
    const float *cur = zero, *prev = zero;
    const bool inpl = in_place(b);
    for (int it = 0; it < b.nt; ++it) {
        float *out;
        const bool ck = save_ckpt && it + 1 < b.nt && sc.ckslot[it + 1] >= 0;
        if (ck && !b.ck16)
            out = b.F(BatchPlan::CK0 + sc.ckslot[it + 1]);
        else if (inpl && is_work(W, prev))
            out = const_cast<float *>(prev);
        else
            out = pick_free(W, cur, prev);
        launch_record(rt, cur, b.SS, b.CS, b.rec, it, c.st);
        a.out = out;
        a.cur = cur;
        a.prev = prev;
        a.it = it;
        a.acc_on = (precond && it % MST == 0) ? 1 : 0;
        set_deferral(a, defer, staged, it + MST > b.nt - 1);
        fd_launch(FD_FWD, a, b.cells_per_launch, b.tile_for(FD_FWD, a));
        if (ck && b.ck16) ck_store(b, sc.ckslot[it + 1], out);
        // propagation.c:597-601: p1 = u(it); u(0) is the zero start field
        if (snap_step(b, it)) snapshot(b, "wavefield_forward", it, cur, it > 0);
        prev = cur;
        cur = out;
    }
    if (fused)
        launch_stage_scatter(b.tb.rec_slot, b.tb.rec_sr, b.stage, b.rec, b.nslot, b.nt, c.st);
}

// The adjoint wavefield starts from zero (propagation.c:504-506); only the
// three work fields of every arena are cleared, the checkpoints stay.
static void adjoint_clear(BatchState &b)
{
    Ctx &c = ctx();
    if (b.lanes)
        HIPCHECK(hipMemsetAsync(b.arena, 0, 3 * b.KS * sizeof(float), c.st));
    else
        for (int s = 0; s < b.S; ++s)
            HIPCHECK(
                hipMemsetAsync(b.arena + (long)s * b.SS, 0, 3 * b.g.fsz * sizeof(float), c.st));
    HIPCHECK(hipMemsetAsync(b.acc, 0, (size_t)b.Sp * b.nacc * sizeof(double), c.st));
}

// propagation.c:633-744.  b.rec holds the adjoint source [ntr][nt] per shot.
// Kernel launches only (the clearing is adjoint_clear), so a HIP graph can
// capture it.
static void adjoint_loop(BatchState &b)
{
    Ctx &c = ctx();
    const Schedule &sc = *b.sc;
    const Geom &g = b.g;
    const int nt = b.nt;
    std::vector<char> stored(nt, 0);
    std::vector<const float *> where(nt, nullptr);
    for (int it = 0; it < nt; ++it) {
        if (sc.ckslot[it] == -2) {
            stored[it] = 1;
            where[it] = b.F(BatchPlan::ZERO);
        } else if (sc.ckslot[it] >= 0) {
            stored[it] = 1;
            // FP16 slots are restored into R0, R1 or RU when they are read
            where[it] = b.ck16 ? nullptr : b.F(BatchPlan::CK0 + sc.ckslot[it]);
        }
    }
    const int ring0 = b.ring0;
    float *FW[3] = {b.F(BatchPlan::F0), b.F(BatchPlan::F1), b.F(BatchPlan::F2)};
    float *AW[3] = {b.F(BatchPlan::A0), b.F(BatchPlan::A1), b.F(BatchPlan::A2)};
    FdArgs fr = base_args(g, b.coef, b.S, b.SS, b.CS);
    fr.src_cell = b.cell_dev;
    fr.src_zlo = b.src_zlo;
    fr.src_zhi = b.src_zhi;
    fr.src_wav = b.wav_dev;
    fr.nt = nt;
    FdArgs ad = base_args(g, b.coef, b.S, b.SS, b.CS);
    ad.acc = b.acc;
    ad.acc_stride = b.ASS;
    ad.acc_cell = b.ACS;
    const float *K = b.coef + CO_K * g.fsz;
    // the LDS stencil of step it injects the adjoint sources of its stage
    // slots for step it - 1 into its output; step nt - 1 and the left-over
    // receivers keep the separate injection
    const bool fused = b.tb.nci > 0 && b.tile_m[FD_ADJ].kind >= 3;
    if (fused) {
        launch_stage_gather(b.adjfull ? b.tb.injfull_slot : b.tb.inj_slot,
                            b.adjfull ? b.adjfull : b.rec, b.stage, b.nslot, nt, c.st);
        ad.rrow = b.tb.rrow;
        ad.rcol = b.tb.rcol;
        ad.rec_zlo = b.tb.rec_zlo;
        ad.rec_zhi = b.tb.rec_zhi;
        ad.stage = b.stage;
        ad.nslot = b.nslot;
    }
    // imaging terms staged as in the forward (the last imaging step is it = 0)
    const bool defer = b.accstage && b.tile_m[FD_ADJ].kind >= 3;
    ad.accstage = b.accstage;
    ad.accstage_stride = b.accstage_stride;
    int staged = 0;
    const float *cur = AW[0], *prev = AW[1];
    // what the reference's p1_fwd holds (verbose snapshots): zeros at first,
    // then u(t) of the latest step t that ended a recomputation or imaged
    // (get_acoustic2dWavefield copies u(t) into it, propagation.c:709)
    const float *fwd_seen = b.F(BatchPlan::ZERO);
    int fwd_t = -1;
    // The imaging condition reads the forward field at it % MST == 0 only, so
    // a segment is recomputed when the reverse loop first meets such a step
    // that is not stored, up to that step: the reference recomputes up to the
    // segment's last step (kd + d - 2), whose fields past its last imaging
    // step nothing reads (the fields computed are the same).  Verbose
    // wavefield files follow the reference's steps, whose recomputation ends
    // are what its p1_fwd shows.
    const bool trim = !(b.vb && b.vb->snaps());
    const bool inpl = in_place(b);
    const bool snap2 = inpl && b.tile_m[FD_RECON].kind == 4;
    for (int it = nt - 1; it >= 0; --it) {
        if (!stored[it] && (!trim || it % MST == 0)) {
            // get_acoustic2dCheckPointStatus: the segment restarts at the last
            // multiple of d below it, from the stored pair (kd-1, kd)
            const int s0 = (it / sc.d) * sc.d;
            const float *fprev = (s0 == 0) ? b.F(BatchPlan::ZERO) : where[s0 - 1];
            const float *fcur = where[s0];
            if (b.ck16 && s0 > 0) {
                ck_load(b, sc.ckslot[s0 - 1], b.F(BatchPlan::R0));
                ck_load(b, sc.ckslot[s0], b.F(BatchPlan::R1));
                fprev = b.F(BatchPlan::R0);
                fcur = b.F(BatchPlan::R1);
            }
            for (int itt = s0; itt < it; ++itt) {
                const int t = itt + 1;
                // the field of an imaging step goes to its ring slot: as the
                // launch's output, or (steps in place, register-ring stencil)
                // as a second store of the model cells beside the in-place
                // update of the work field, so that the next step stays in
                // place as well
                float *const ring = (t % MST == 0) ? b.F(ring0 + (t - s0) / MST) : nullptr;
                const bool inw = inpl && is_work(FW, fprev);
                float *fout;
                fr.snap = nullptr;
                if (ring && inw && snap2) {
                    fout = const_cast<float *>(fprev);
                    fr.snap = ring;
                } else if (ring) {
                    fout = ring;
                } else {
                    fout = inw ? const_cast<float *>(fprev) : pick_free(FW, fcur, fprev);
                }
                fr.out = fout;
                fr.cur = fcur;
                fr.prev = fprev;
                fr.it = itt;
                fd_launch(FD_RECON, fr, b.cells_per_launch, b.tile_m[FD_RECON]);
                fprev = fcur;
                fcur = fout;
                stored[t] = 1;
                where[t] = ring ? ring : fout;
            }
            fr.snap = nullptr;
            fwd_seen = where[it];
            fwd_t = it;
        }
        const bool left = fused && it < nt - 1;
        if (b.adjfull)
            launch_inject(left ? b.tb.inj_full_left : b.tb.inj_full, const_cast<float *>(cur),
                          b.SS, b.CS, K, b.adjfull, it, g.dt, g.invdx, g.invdz, c.st);
        else
            launch_inject(left ? b.tb.inj_left : b.tb.inj, const_cast<float *>(cur), b.SS, b.CS,
                          K, b.rec, it, g.dt, g.invdx, g.invdz, c.st);
        ad.inj_on = (fused && it >= 1) ? 1 : 0;
        float *out =
            (inpl && is_work(AW, prev)) ? const_cast<float *>(prev) : pick_free(AW, cur, prev);
        ad.out = out;
        ad.cur = cur;
        ad.prev = prev;
        ad.it = it;
        ad.acc_on = (it % MST == 0) ? 1 : 0;
        set_deferral(ad, defer, staged, it < MST);
        ad.u = (it % MST == 0) ? where[it] : nullptr;
        if (ad.acc_on && !ad.u) {
            // an FP16 checkpoint time: restore the forward field for imaging
            ck_load(b, sc.ckslot[it], b.F(BatchPlan::RU));
            ad.u = b.F(BatchPlan::RU);
        }
        fd_launch(FD_ADJ, ad, b.cells_per_launch, b.tile_for(FD_ADJ, ad));
        if (ad.acc_on) {
            fwd_seen = ad.u;
            fwd_t = it;
        }
        if (snap_step(b, it)) {
            // propagation.c:717-723: p1_fwd, and the adjoint p1 of level it
            // (injected; the start field of it = nt - 1 never went through
            // the free surface)
            snapshot(b, "wavefield_forward_reconstruction", it, fwd_seen, fwd_t > 0);
            snapshot(b, "wavefield_backward", it, cur, it < nt - 1);
        }
        prev = cur;
        cur = out;
    }
}

// ---------------------------------------------------------------------------
// HIP graphs of the time loops.
//
// A time loop of a batch (forward: nt record + stencil launches; adjoint:
// recomputation, injection and adjoint launches) depends only on the batch's
// device pointers, sizes, schedule and launch shape, so once captured it can be
// replayed with one hipGraphLaunch whenever the same batch comes again (the
// next gradient of a session, the next FWI iteration, a serving job of the
// same shape).  The key holds every value the captured kernel arguments
// depend on; buffer contents (models, coefficients, wavelets, records) are
// read at replay.

static void key_push_geom(std::vector<long> &k, const Geom &g)
{
    const float fl[] = {g.dx, g.dz, g.dt};
    uint32_t bits[3];
    memcpy(bits, fl, sizeof(bits));
    for (long v : {(long)g.nx, (long)g.nz, (long)g.pml, (long)g.fs, (long)g.Nx, (long)g.ld,
                   (long)bits[0], (long)bits[1], (long)bits[2]})
        k.push_back(v);
}

static std::vector<long> batch_key(int pass, const BatchState &b, bool f1, bool f2)
{
    std::vector<long> k = {pass, (long)f1, (long)f2, (long)ctx().fma, (long)in_place(b),
                           (long)b.nt, (long)b.S, (long)b.Sp, (long)b.lanes, (long)b.ck16, b.sc ? (long)b.sc->d : 0L,
                           b.tb.rcv.n, b.tb.inj.n, (long)b.tb.nci, b.nslot, b.tb.rcv_left.n,
                           b.tb.inj_left.n};
    key_push_geom(k, b.g);
    for (const void *p : {(const void *)b.arena, (const void *)b.coef, (const void *)b.cell_dev,
                          (const void *)b.wav_dev, (const void *)b.rec, (const void *)b.acc,
                          (const void *)b.h16, (const void *)b.ckinv, (const void *)b.ckmx,
                          (const void *)b.tb.rcv.trace_off, (const void *)b.tb.rcv.shot,
                          (const void *)b.tb.inj.trace_off, (const void *)b.tb.inj.shot,
                          (const void *)b.adjfull, (const void *)b.stage,
                          (const void *)b.accstage,
                          (const void *)b.tb.rrow, (const void *)b.tb.rec_slot,
                          (const void *)b.tb.rcv_left.trace_off, (const void *)b.tb.inj_left.beg})
        k.push_back((long)(uintptr_t)p);
    for (const auto &tm : b.tile_acc)
        for (const FdTile &t : tm)
            for (long v : {(long)t.kind, (long)t.tz, (long)t.spb, (long)t.txm, (long)t.zfast,
                           (long)t.pd, (long)t.remap, (long)t.wpb, (long)t.zrun})
                k.push_back(v);
    return k;
}

// Run body on the library stream, through a cached graph when graphs are on.
template <class F>
static void run_loop(const std::vector<long> &key, F body)
{
    Ctx &c = ctx();
    if (!c.graphs) {
        body();
        return;
    }
    auto it = c.graph_cache.find(key);
    if (it == c.graph_cache.end() && c.graph_seen.insert(key).second) {
        // first sight of this loop: run it as it is (a graph pays its capture
        // and instantiation only when the same loop comes again)
        if (c.graph_seen.size() > 256) {
            c.graph_seen.clear();
            c.graph_seen.insert(key);
        }
        body();
        return;
    }
    if (it == c.graph_cache.end()) {
        if (c.graph_cache.size() >= 16) {
            for (auto &e : c.graph_cache) (void)hipGraphExecDestroy(e.second.exec);
            c.graph_cache.clear();
        }
        const Stats before = c.stats;
        hipGraph_t gr = nullptr;
        c.capturing = true;
        HIPCHECK(hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal));
        body();
        HIPCHECK(hipStreamEndCapture(c.st, &gr));
        c.capturing = false;
        GraphEntry e;
        HIPCHECK(hipGraphInstantiate(&e.exec, gr, nullptr, nullptr, 0));
        HIPCHECK(hipGraphDestroy(gr));
        e.fd_launches = c.stats.fd_launches - before.fd_launches;
        e.fd_cells = c.stats.fd_cells - before.fd_cells;
        c.stats.graphs_built++;
        it = c.graph_cache.emplace(key, e).first;
        // the launches were counted while capturing; the replay below runs them
        c.stats.fd_launches = before.fd_launches;
        c.stats.fd_cells = before.fd_cells;
    }
    HIPCHECK(hipGraphLaunch(it->second.exec, c.st));
    c.stats.fd_launches += it->second.fd_launches;
    c.stats.fd_cells += it->second.fd_cells;
    c.stats.graph_launches++;
}

void drop_graphs()
{
    Ctx &c = ctx();
    for (auto &e : c.graph_cache) (void)hipGraphExecDestroy(e.second.exec);
    c.graph_cache.clear();
}

// Loops that write verbose wavefield files run kernel by kernel (the files
// are written between launches).
static void forward_pass(BatchState &b, bool save_ckpt, bool precond)
{
    if (b.vb && b.vb->snaps()) {
        forward_loop(b, save_ckpt, precond);
        return;
    }
    run_loop(batch_key(0, b, save_ckpt, precond), [&] { forward_loop(b, save_ckpt, precond); });
}

static void adjoint_pass(BatchState &b)
{
    adjoint_clear(b);
    if (b.vb && b.vb->snaps()) {
        adjoint_loop(b);
        return;
    }
    run_loop(batch_key(1, b, false, false), [&] { adjoint_loop(b); });
}

// ---------------------------------------------------------------------------
// batch size

static size_t per_shot_bytes(const Geom &g, const Schedule &sc, bool adjoint, bool ck16,
                             long rec_floats, int nt)
{
    size_t b = (size_t)fields_per_shot(sc, adjoint, ck16) * g.fsz * sizeof(float);
    if (adjoint && ck16) b += (size_t)sc.nck * g.fsz * sizeof(__half);
    b += (size_t)2 * rec_floats * sizeof(float) + (size_t)nt * sizeof(float);
    // the lanes stencil's time-major receiver stage: at most one receiver row
    // per column
    b += (size_t)nt * g.Nx * sizeof(float);
    if (adjoint) b += (size_t)g.nx * g.nz * (sizeof(double) + (2 + ACC_STAGES) * sizeof(float));
    return b;
}

// Shots per launch and field layout.  Shots in lanes pads the batch to a
// multiple of 64; it is used when at least 48 shots share a launch (below
// that a third of the lanes or more would idle) unless a layout or a stencil
// shape is forced.
// The lanes stencils address a field through 32-bit buffer offsets and form
// row and column byte offsets in signed 32-bit arithmetic: one field of a
// padded batch (fsz cells x Sp lanes x 4 bytes) must stay below 2 GiB.  The
// largest multiple of 64 shots that does (0 if not even 64 do).
long lanes_max_shots(long fsz)
{
    if (fsz <= 0) return 0;
    return (0x7FFFFFFFL / (fsz * 4)) / 64 * 64;
}

// arenas: the batch must run in per-shot arenas (verbose wavefield files need
// the arena stencils' full free-surface mirror)
static int choose_batch(int n, size_t per_shot, long fsz, bool &lanes, bool arenas)
{
    Ctx &c = ctx();
    long Smax;
    if (c.batch_req > 0) {
        Smax = c.batch_req;
    } else {
        double budget;
        if (c.mem_cap > 0) {
            // a serving worker: the batch buffers must fit its reservation
            budget = (double)c.mem_cap;
        } else {
            size_t fr = 0, tot = 0;
            HIPCHECK(hipMemGetInfo(&fr, &tot));
            // buffers the engine already holds are reused
            size_t held = c.arena.cap + c.acc.cap + c.rec.cap + c.obs.cap + c.src.cap +
                          c.misc[0].cap + c.misc[1].cap + c.ck16.cap + c.stage.cap +
                          c.accstage.cap;
            budget = ((double)fr + (double)held + (double)pool_idle_bytes()) * c.memfrac;
        }
        Smax = std::max<long>(1, (long)(budget / (double)per_shot));
    }
    long S = std::min<long>(n, Smax);
    if (arenas) {
        if (c.tile_req.tz > 0 && c.tile_req.kind >= 2)
            fail("<error>: verbose 2 writes wavefields from the per-shot-arena stencils; a forced "
                 "shots-in-lanes stencil shape cannot run with it");
        lanes = false;
    } else if (c.tile_req.tz > 0)
        lanes = c.tile_req.kind >= 2;
    else if (c.layout_req >= 0)
        lanes = c.layout_req == 1;
    else
        lanes = S >= 48;
    if (lanes && c.batch_req <= 0 && (S + 63) / 64 * 64 > Smax) {
        // the padded batch must fit the budget too
        if (Smax >= 64)
            S = Smax / 64 * 64;
        else if (c.tile_req.tz <= 0 && c.layout_req < 0)
            lanes = false;
    }
    if (lanes) {
        const long max_lanes = lanes_max_shots(fsz);
        if (max_lanes < 64) {
            if (c.tile_req.tz > 0 && c.tile_req.kind >= 2)
                fail("<error>: the shots-in-lanes stencil cannot address this grid (one field of 64 "
                     "shots reaches 2 GiB)");
            lanes = false;
        } else if ((S + 63) / 64 * 64 > max_lanes) {
            S = max_lanes;
        }
    }
    return (int)S;
}

// ---------------------------------------------------------------------------
// groups of shots with identical coefficient arrays

struct GroupKey {
    int ix_min, ix_max;
    uint32_t fpeak_bits;
    bool operator<(const GroupKey &o) const
    {
        return std::tie(ix_min, ix_max, fpeak_bits) < std::tie(o.ix_min, o.ix_max, o.fpeak_bits);
    }
};

static std::vector<std::vector<int>> group_shots(const Record2D *obs, int Nshot, int nx)
{
    std::map<GroupKey, int> idx;
    std::vector<std::vector<int>> groups;
    for (int s = 0; s < Nshot; ++s) {
        const Record2D &r = obs[s];
        if (r.ix_min < 0 || r.ix_max > nx - 1 || r.ix_min > r.ix_max)
            fail("<error>: shot %d has an invalid window [%d, %d] for nx=%d", s + 1, r.ix_min,
                 r.ix_max, nx);
        uint32_t fb;
        memcpy(&fb, &r.fpeak, 4);
        GroupKey k{r.ix_min, r.ix_max, fb};
        auto it = idx.find(k);
        if (it == idx.end()) {
            idx[k] = (int)groups.size();
            groups.push_back({s});
        } else {
            groups[it->second].push_back(s);
        }
    }
    return groups;
}

void check_common(int nz, int nx, float dx, float dz, float dt, int nt, int pml, int fs)
{
    if (nz < 4 || nx < 1) fail("<error>: model too small (nz=%d, nx=%d)", nz, nx);
    if (nt < 1) fail("<error>: nt must be positive (nt=%d)", nt);
    if (pml < 3) fail("<error>: pmlThick must be at least 3 (the stencil half-width), got %d", pml);
    if (fs != 0 && fs != 1) fail("<error>: freesurface must be 0 or 1, got %d", fs);
    if (!(dx > 0.0f) || !(dz > 0.0f) || !(dt > 0.0f))
        fail("<error>: dx, dz, dt must be positive");
}


// Device bytes a call of propagate_core needs: `fixed` for the whole call and
// `per_shot` for every shot advanced in the same launch (the batcher's own
// accounting plus the per-call buffers), so that a serving worker can reserve
// fixed + per_shot * S and run batches of S shots inside it.
void call_footprint(const Call &P, bool ck16, size_t *fixed, size_t *per_shot)
{
    const long nmod = (long)P.nx * P.nz;
    long win_max = 0, recmax = 0, rec_total = 0, ntr_total = 0;
    for (int s = 0; s < P.Nshot; ++s) {
        win_max = std::max<long>(win_max, (long)(P.obs[s].ix_max - P.obs[s].ix_min + 1));
        recmax = std::max<long>(recmax, (long)P.obs[s].ntr * P.obs[s].ns);
        rec_total += (long)P.obs[s].ntr * P.obs[s].ns;
        ntr_total += P.obs[s].ntr;
    }
    Geom g = make_geom((int)win_max, P.nz, P.pml, P.fs, P.dx, P.dz, P.dt);
    Schedule sc(P.nt);
    const bool grad = P.problem == 1;
    size_t f = (size_t)2 * nmod * 4;                                     // models
    f += (size_t)(CO_N + 4) * g.fsz * 4;                               // coefficients + setup
    f += (size_t)rec_total * 4 + (size_t)ntr_total * 32;               // observed records, tables
    if (grad) {
        f += (size_t)2 * P.Nshot * win_max * P.nz * 4;                 // per-shot outputs
        f += (size_t)2 * nmod * (8 + 4) + (size_t)2 * nmod * 4;        // sums, float copies
    }
    f += (size_t)64 << 20;                                             // reductions, slack
    *fixed = f;
    *per_shot = per_shot_bytes(g, sc, grad, ck16, recmax, P.nt) + (size_t)P.nt * 4 + 4096;
}

void propagate_core(const Call &P)
{
    Ctx &c = ctx();
    const int Nshot = P.Nshot, nz = P.nz, nx = P.nx, nt = P.nt, problem = P.problem;
    const long nmod = (long)nx * nz;
    float *vpd = c.model[0].as<float>(nmod);
    float *rhod = c.model[1].as<float>(nmod);
    if (!P.models_on_device) {
        h2d(vpd, P.vp[0], nmod * sizeof(float));
        h2d(rhod, P.rho[0], nmod * sizeof(float));
    }
    std::vector<std::vector<int>> groups = group_shots(P.obs, Nshot, nx);
    long win_max = 0;
    for (int s = 0; s < Nshot; ++s)
        win_max = std::max<long>(win_max, (long)(P.obs[s].ix_max - P.obs[s].ix_min + 1) * nz);
    float *gout = nullptr, *pout = nullptr;
    if (problem == 1) {
        float *so = c.shot_out.as<float>((size_t)2 * Nshot * win_max);
        gout = so;
        pout = so + (size_t)Nshot * win_max;
    }
    double *mis = nullptr;
    if (P.misfit) {
        mis = c.gsum.as<double>((size_t)2 * nmod + 2 * Nshot + 64) + 2 * nmod;
        HIPCHECK(hipMemsetAsync(mis, 0, (size_t)2 * Nshot * sizeof(double), c.st));
    }
    Schedule sc(nt);
    for (auto &grp : groups) {
        const Record2D &r0 = P.obs[grp[0]];
        const int nxx = r0.ix_max - r0.ix_min + 1;
        Geom g = make_geom(nxx, nz, P.pml, P.fs, P.dx, P.dz, P.dt);
        float *coef = c.coef.as<float>((size_t)CO_ALLOC * g.fsz);
        const float *vpw = vpd + (long)r0.ix_min * nz, *rhow = rhod + (long)r0.ix_min * nz;
        build_coefs(g, vpw, rhow, r0.fpeak, coef);
        long recmax = 0;
        for (int s : grp) recmax = std::max<long>(recmax, (long)P.pre[s].ntr * P.pre[s].ns);
        bool lanes = false;
        const bool ck16 = c.ckpt_bits == 16;
        const int S = choose_batch((int)grp.size(),
                                   per_shot_bytes(g, sc, problem == 1, ck16, recmax, nt), g.fsz,
                                   lanes, P.vb.snaps());
        c.stats.batch = S;
        c.stats.layout = lanes ? 1 : 0;
        for (size_t b0 = 0; b0 < grp.size(); b0 += S) {
            const int nb = (int)std::min<size_t>(S, grp.size() - b0);
            std::vector<Record2D *> recs(nb);
            for (int k = 0; k < nb; ++k) {
                Record2D *pr = &P.pre[grp[b0 + k]];
                if (pr->ntr != P.obs[grp[b0 + k]].ntr || pr->ns != P.obs[grp[b0 + k]].ns)
                    fail("<error>: predicted and observed records of shot %d differ in shape",
                         grp[b0 + k] + 1);
                recs[k] = pr;
            }
            BatchState b;
            prepare_batch(b, g, coef, recs, nt, problem == 1, sc, lanes, ck16);
            if (P.vb.snaps()) {
                b.vb = &P.vb;
                b.snap_shot.assign(nb, -1);
                for (int k = 0; k < nb; ++k)
                    if (P.vb.shot(grp[b0 + k])) b.snap_shot[k] = grp[b0 + k];
            }
            // verbose >= 1: the gradient's per-shot files (propagation.c:917-943)
            auto shot_files = [&](const char *what, bool records, const float *src, long stride) {
                if (problem != 1 || P.vb.level < 1) return;
                char name[512];
                for (int k = 0; k < nb; ++k) {
                    const int ishot = grp[b0 + k];
                    if (!P.vb.shot(ishot)) continue;
                    const Record2D &r = *recs[k];
                    if (records) {
                        snprintf(name, sizeof(name), "%s_iter%03d_shot%05d_%d_%d_%.2fms_%.2fm.bin",
                                 what, P.vb.iter, ishot + 1, r.ns, r.ntr, r.dt * 1e3, r.dtr);
                        write_bin_device(name, src + b.roff[k], (size_t)r.ntr * r.ns);
                    } else {
                        snprintf(name, sizeof(name), "%s_iter%03d_shot%05d_%d_%d_%.2fm_%.2fm.bin",
                                 what, P.vb.iter, ishot + 1, nz, nxx, P.dz, P.dx);
                        write_bin_device(name, src + (long)k * stride, (size_t)nxx * nz);
                    }
                }
            };
            // the batch's records in the caller's device buffers (sessions)
            std::vector<long> boff(nb), blen(nb);
            for (int k = 0; k < nb; ++k) {
                const int sidx = grp[b0 + k];
                boff[k] = P.obs_off ? P.obs_off[sidx] : 0;
                blen[k] = (long)P.obs[sidx].ntr * P.obs[sidx].ns;
            }
            float *obsd = nullptr;
            RecordCopies up, down;
            if (problem != 0) {
                obsd = c.obs.as<float>(b.rec_total + 64);
                if (P.obs_dev) {
                    copy_pieces(obsd, b.roff.data(), P.obs_dev, boff.data(), blen.data(), nb);
                } else {
                    // uploaded while the forward pass runs; joined before the
                    // adjoint sources read them
                    std::vector<RecordCopies::Piece> pc(nb);
                    for (int k = 0; k < nb; ++k) {
                        const Record2D &r = P.obs[grp[b0 + k]];
                        pc[k] = {obsd + b.roff[k], r.traces_p[0],
                                 (size_t)r.ntr * r.ns * sizeof(float)};
                    }
                    up.start(c, std::move(pc), hipMemcpyHostToDevice);
                }
            }
            const long nacc = b.nacc;
            float *gb = nullptr, *pb = nullptr;
            if (problem == 1) {
                float *m = c.misc[0].as<float>((size_t)2 * nb * nacc);
                gb = m;
                pb = m + (size_t)nb * nacc;
            }
            forward_pass(b, problem == 1, problem == 1 && P.precondition == 1);
            if (problem == 1 && P.precondition == 1) {
                // propagation.c:610-623
                launch_finalize(b.acc, b.ASS, b.ACS, pb, nacc, nb, vpw, rhow, (int)nacc, 4.0, 1,
                                c.st);
            }
            if (problem == 0) {
                if (P.records_to_host)
                    for (int k = 0; k < nb; ++k) rec_d2h(recs[k], b.rec + b.roff[k]);
                continue;
            }
            shot_files("predicted", true, b.rec, 0);
            up.join();
            adjoint_sources(recs.data(), nb, P.vp, P.freq, P.filter, b.rec, obsd, b.rec,
                            b.roff.data());
            shot_files("adjsource", true, b.rec, 0);
            if (mis) {
                double *scr = c.scratch.as<double>(reduce_scratch_doubles() + (size_t)Nshot * 512 + Nshot + 64);
                for (int k = 0; k < nb; ++k)
                    launch_sum_sq(b.rec + b.roff[k], (long)recs[k]->ntr * recs[k]->ns, scr,
                                  mis + 2 * grp[b0 + k], c.st);
            }
            if (P.records_to_host) {
                // downloaded while the adjoint pass runs (it only reads
                // b.rec); `down` is joined when this batch's scope ends
                std::vector<RecordCopies::Piece> pc(nb);
                for (int k = 0; k < nb; ++k)
                    pc[k] = {recs[k]->traces_p[0], b.rec + b.roff[k],
                             (size_t)recs[k]->ntr * recs[k]->ns * sizeof(float)};
                down.start(c, std::move(pc), hipMemcpyDeviceToHost);
            }
            if (P.pre_dev) copy_pieces(P.pre_dev, boff.data(), b.rec, b.roff.data(), blen.data(), nb);
            if (problem == 2) continue;
            bool full = false;
            for (int k = 0; k < nb; ++k) full |= recs[k]->sampleRate != 1 || recs[k]->ns != nt;
            if (full) {
                // propagation.c:635: the adjoint sources at every time step
                b.adjfull = c.adjfull.as<float>((size_t)b.full_total + 64);
                for (int k = 0; k < nb; ++k)
                    interpolate_traces_device(b.rec + b.roff[k], recs[k]->ntr, recs[k]->ns,
                                              recs[k]->sampleRate, b.adjfull + b.foff[k], nt);
            }
            adjoint_pass(b);
            launch_finalize(b.acc, b.ASS, b.ACS, gb, nacc, nb, vpw, rhow, (int)nacc, 2.0, 0, c.st);
            shot_files("gradient_raw", false, gb, nacc);
            if (P.precondition == 1) {
                double *scr = c.scratch.as<double>((size_t)nb * 512 + nb + reduce_scratch_doubles());
                launch_shot_precondition(gb, pb, nacc, nb, (int)nacc, scr, c.st);
            }
            std::vector<long> go(nb), bo(nb), gl(nb, nacc);
            for (int k = 0; k < nb; ++k) {
                go[k] = (long)grp[b0 + k] * win_max;
                bo[k] = (long)k * nacc;
            }
            copy_pieces(gout, go.data(), gb, bo.data(), gl.data(), nb);
            if (P.precondition == 1) copy_pieces(pout, go.data(), pb, bo.data(), gl.data(), nb);
        }
    }
    if (problem == 1) {
        // propagation.c:953-1011: shot-order FP64 sums, then global preconditioning
        std::vector<int> ix0(Nshot), nxx(Nshot);
        for (int s = 0; s < Nshot; ++s) {
            ix0[s] = P.obs[s].ix_min;
            nxx[s] = P.obs[s].ix_max - P.obs[s].ix_min + 1;
        }
        int *wi = c.misc[2].as<int>((size_t)2 * Nshot);
        h2d(wi, ix0.data(), Nshot * sizeof(int));
        h2d(wi + Nshot, nxx.data(), Nshot * sizeof(int));
        double *sums = c.gsum.as<double>((size_t)2 * nmod + 2 * Nshot + 64);
        HIPCHECK(hipMemsetAsync(sums, 0, (size_t)2 * nmod * sizeof(double), c.st));
        launch_accumulate(gout, win_max, Nshot, wi, wi + Nshot, nz, nx, sums, c.st);
        float *gf = P.grad_dev ? P.grad_dev : c.misc[3].as<float>((size_t)2 * nmod);
        float *pf = c.misc[3].as<float>((size_t)2 * nmod) + nmod;
        launch_to_float(sums, gf, nmod, c.st);
        char name[512];
        if (P.vb.level > 0) {
            // propagation.c:996-1005: before the global preconditioning
            snprintf(name, sizeof(name), "gradient_raw_iter%03d_%d_%d_%.2fm_%.2fm.bin", P.vb.iter,
                     nz, nx, P.dz, P.dx);
            write_bin_device(name, gf, (size_t)nmod);
        }
        if (P.precondition == 1) {
            launch_accumulate(pout, win_max, Nshot, wi, wi + Nshot, nz, nx, sums + nmod, c.st);
            launch_to_float(sums + nmod, pf, nmod, c.st);
            if (P.vb.level > 0) {
                snprintf(name, sizeof(name), "precond_iter%03d_%d_%d_%.2fm_%.2fm.bin", P.vb.iter,
                         nz, nx, P.dz, P.dx);
                write_bin_device(name, pf, (size_t)nmod);
            }
            double *scr = c.scratch.as<double>(reduce_scratch_doubles() + (size_t)Nshot * 512 + Nshot + 64);
            launch_global_precondition(gf, pf, nmod, 0.005f, scr, c.st);
        }
        if (P.grad && P.grad[0]) d2h(P.grad[0], gf, nmod * sizeof(float));
    }
    if (mis) {
        std::vector<double> h((size_t)2 * Nshot);
        d2h(h.data(), mis, h.size() * sizeof(double));
        sync();
        double f = 0.0;
        for (int s = 0; s < Nshot; ++s) f += h[2 * s];
        *P.misfit = 0.5 * f;
    }
    sync();
    collect_timing();
}

}  // namespace fwiedu

using namespace fwiedu;

extern "C" void iso_acoustic2d_propagation(float **vp, float **rho, Record2D *record2d_obs,
                                           int Nshot, int nz, int nx, float dx, float dz, float dt,
                                           int nt, int pmlThick, int freesurface, float freq,
                                           float filter, int precondition, int problem, int iter,
                                           int verbose, Record2D *record2d_pre, float **grad)
{
    check_common(nz, nx, dx, dz, dt, nt, pmlThick, freesurface);
    if (Nshot <= 0) fail("<error>: Nshot must be positive");
    if (problem < 0 || problem > 2) fail("<error>: problem must be 0, 1 or 2 (got %d)", problem);
    if (problem == 1 && (!grad || !grad[0])) fail("<error>: problem 1 needs a gradient array");
    ensure_ready();
    Call P;
    P.vp = vp;
    P.rho = rho;
    P.obs = record2d_obs;
    P.pre = record2d_pre;
    P.Nshot = Nshot;
    P.nz = nz;
    P.nx = nx;
    P.dx = dx;
    P.dz = dz;
    P.dt = dt;
    P.nt = nt;
    P.pml = pmlThick;
    P.fs = freesurface;
    P.freq = freq;
    P.filter = (int)filter;  // float here, int in compute_adjoint_source2d
    P.precondition = precondition;
    P.problem = problem;
    P.grad = grad;
    P.vb = verbose_from_env(verbose, iter);
    propagate_core(P);
}

// ---------------------------------------------------------------------------
// Device-resident session: model, observed records and all wavefield state in
// HBM; repeated gradients move no traces over PCIe.

struct Session {
    Call P;
    std::vector<float> vp, rho;            // host copies (the taper reads vp on the host)
    std::vector<float *> vrow, rrow;
    std::vector<Record2D> obs, pre;        // headers; traces stay on the device
    DevBuf obs_dev;
    std::vector<long> off;
    long rec_total = 0;
    DevBuf grad_dev;
    DevBuf pre0, pre1, slow, pert;         // FWI iteration state
};

extern "C" void *fwiedu_session_create(float **vp, float **rho, Record2D *record2d_obs, int Nshot,
                                       int nz, int nx, float dx, float dz, float dt, int nt,
                                       int pmlThick, int freesurface, float freq, float filter,
                                       int precondition)
{
    check_common(nz, nx, dx, dz, dt, nt, pmlThick, freesurface);
    if (Nshot <= 0) fail("<error>: Nshot must be positive");
    ensure_ready();
    Session *S = new Session();
    const long nmod = (long)nx * nz;
    S->vp.assign(vp[0], vp[0] + nmod);
    S->rho.assign(rho[0], rho[0] + nmod);
    S->vrow.resize(nx);
    S->rrow.resize(nx);
    for (int ix = 0; ix < nx; ++ix) {
        S->vrow[ix] = S->vp.data() + (long)ix * nz;
        S->rrow[ix] = S->rho.data() + (long)ix * nz;
    }
    S->obs.assign(record2d_obs, record2d_obs + Nshot);
    S->pre = S->obs;
    S->off.resize(Nshot);
    long tot = 0;
    for (int s = 0; s < Nshot; ++s) {
        S->off[s] = tot;
        tot += (long)record2d_obs[s].ntr * record2d_obs[s].ns;
    }
    S->rec_total = tot;
    float *d = S->obs_dev.as<float>(tot + 64);
    for (int s = 0; s < Nshot; ++s)
        h2d(d + S->off[s], record2d_obs[s].traces_p[0],
            (size_t)record2d_obs[s].ntr * record2d_obs[s].ns * sizeof(float));
    Call &P = S->P;
    P.vp = S->vrow.data();
    P.rho = S->rrow.data();
    P.obs = S->obs.data();
    P.pre = S->pre.data();
    P.Nshot = Nshot;
    P.nz = nz;
    P.nx = nx;
    P.dx = dx;
    P.dz = dz;
    P.dt = dt;
    P.nt = nt;
    P.pml = pmlThick;
    P.fs = freesurface;
    P.freq = freq;
    P.filter = (int)filter;
    P.precondition = precondition;
    P.problem = 1;
    P.obs_dev = d;
    P.obs_off = S->off.data();
    P.records_to_host = false;
    P.grad_dev = S->grad_dev.as<float>(nmod + 64);
    return S;
}

extern "C" void fwiedu_session_set_model(void *sp, float **vp, float **rho)
{
    Session *S = static_cast<Session *>(sp);
    const long nmod = (long)S->P.nx * S->P.nz;
    memcpy(S->vp.data(), vp[0], nmod * sizeof(float));
    memcpy(S->rho.data(), rho[0], nmod * sizeof(float));
}

// One gradient (problem 1) on resident data; returns the misfit of the
// processed residual.  grad (host) may be null: the gradient then stays on
// the device.
extern "C" float fwiedu_session_gradient(void *sp, float **grad)
{
    Session *S = static_cast<Session *>(sp);
    double f = 0.0;
    Call P = S->P;
    P.grad = grad;
    P.misfit = &f;
    propagate_core(P);
    return (float)f;
}

extern "C" void fwiedu_session_destroy(void *sp)
{
    Session *S = static_cast<Session *>(sp);
    if (!S) return;
    sync();
    for (DevBuf *d : {&S->obs_dev, &S->grad_dev, &S->pre0, &S->pre1, &S->slow, &S->pert})
        d->release(ctx().st);
    delete S;
}

// One iteration of the FWI driver (acoustic2d_fwi.c:409-465) with every array
// on the device: gradient (problem 1), smoothing along z then x, objective,
// perturbation from the smoothed gradient (before its water rows are zeroed,
// T15), trial model, line-search trial (problem 2), objective, step length,
// update of slowness, velocity and (Gardner) density, water rows and clipping.
extern "C" void fwiedu_session_iterate(void *sp, FwieduIteration *it, float **vp_out,
                                       float **rho_out, float **grad_out)
{
    Session *S = static_cast<Session *>(sp);
    if (!S || !it) fail("<error>: fwiedu_session_iterate: null argument");
    if (it->rz < 1 || it->rx < 1) fail("<error>: fwiedu_session_iterate: smoothing lengths >= 1");
    Ctx &c = ctx();
    const int nx = S->P.nx, nz = S->P.nz;
    const long n = (long)nx * nz;
    float *pre0 = S->pre0.as<float>(S->rec_total + 64);
    float *pre1 = S->pre1.as<float>(S->rec_total + 64);
    float *slow = S->slow.as<float>(n + 64);
    float *pert = S->pert.as<float>(n + 64);
    float *grad = S->P.grad_dev;
    // slowness of the current model (acoustic2d_fwi.c:365, 463)
    float *vpd = c.model[0].as<float>(n), *rhod = c.model[1].as<float>(n);
    h2d(vpd, S->vp.data(), n * sizeof(float));
    h2d(rhod, S->rho.data(), n * sizeof(float));
    launch_model_op(MOP_RECIP, vpd, nullptr, slow, n, nz, 0, 0, 0, 0, c.st);

    // GRADIENT (models already on the device)
    Call P = S->P;
    P.freq = it->freq;
    P.filter = (int)it->filter;
    P.pre_dev = pre0;
    P.models_on_device = true;
    P.vb = verbose_from_env(it->verbose, it->iter);
    double f = 0.0;
    P.misfit = &f;
    propagate_core(P);
    smooth2d_device(grad, pert, nx, nz, it->rz, 1);
    smooth2d_device(pert, grad, nx, nz, it->rx, 0);

    // STEP-LENGTH
    it->fun1 = objective_device(pre0, S->rec_total);
    it->scalar = model_perturbation_device(it->fraction, slow, grad, pert, (int)n);
    launch_model_op(MOP_WATER, nullptr, nullptr, grad, n, nz, 0.0f, 0, 0, it->waterdepth, c.st);
    auto set_model = [&](const float *sl) {
        // slow2v, water, clip; density from Gardner, water, clip
        launch_model_op(MOP_RECIP, sl, nullptr, vpd, n, nz, 0, 0, 0, 0, c.st);
        launch_model_op(MOP_WATER, nullptr, nullptr, vpd, n, nz, it->vp_water, 0, 0,
                        it->waterdepth, c.st);
        launch_model_op(MOP_CLIP, vpd, nullptr, vpd, n, nz, 0, it->vpmin, it->vpmax, 0, c.st);
        if (it->gardner) {
            launch_model_op(MOP_GARDNER, vpd, nullptr, rhod, n, nz, 0, 0, 0, 0, c.st);
            launch_model_op(MOP_WATER, nullptr, nullptr, rhod, n, nz, it->rho_water, 0, 0,
                            it->waterdepth, c.st);
            launch_model_op(MOP_CLIP, rhod, nullptr, rhod, n, nz, 0, it->rhomin, it->rhomax, 0,
                            c.st);
        }
    };
    float *trial = c.misc[1].as<float>(n + 64);
    launch_model_op(MOP_ADD, slow, pert, trial, n, nz, 0, 0, 0, 0, c.st);
    set_model(trial);
    // the adjoint-source taper reads vp on the host: keep the host copy current
    d2h(S->vp.data(), vpd, n * sizeof(float));
    d2h(S->rho.data(), rhod, n * sizeof(float));
    sync();

    // line-search trial (problem 2)
    Call P2 = P;
    P2.problem = 2;
    P2.pre_dev = pre1;
    P2.grad_dev = nullptr;
    double f2 = 0.0;
    P2.misfit = &f2;
    propagate_core(P2);
    it->fun2 = objective_device(pre1, S->rec_total);
    it->steplen = step_length_device(pre0, pre1, S->rec_total);

    // UPDATE Model
    launch_model_op(MOP_SCALE, pert, nullptr, pert, n, nz, it->steplen, 0, 0, 0, c.st);
    launch_model_op(MOP_ADD, slow, pert, slow, n, nz, 0, 0, 0, 0, c.st);
    set_model(slow);
    d2h(S->vp.data(), vpd, n * sizeof(float));
    d2h(S->rho.data(), rhod, n * sizeof(float));
    if (vp_out && vp_out[0]) d2h(vp_out[0], vpd, n * sizeof(float));
    if (rho_out && rho_out[0]) d2h(rho_out[0], rhod, n * sizeof(float));
    if (grad_out && grad_out[0]) d2h(grad_out[0], grad, n * sizeof(float));
    sync();
}

// ---------------------------------------------------------------------------
// Optimisation pieces on device arrays, shared by the C entry points of
// optimization.c and the device-resident FWI iteration.
namespace fwiedu {

// calculate_objective_value (optimization.c:99-131) of n samples.
float objective_device(const float *d, long n)
{
    Ctx &c = ctx();
    double *scr = c.scratch.as<double>(reduce_scratch_doubles() + 64);
    double *res = scr + reduce_scratch_doubles();
    launch_sum_sq(d, n, scr, res, c.st);
    double h[2];
    d2h(h, res, 2 * sizeof(double));
    sync();
    return (float)(h[0] * 0.5);
}

// calculate_step_length (optimization.c:134-175).
float step_length_device(const float *a, const float *b, long n)
{
    Ctx &c = ctx();
    double *scr = c.scratch.as<double>(reduce_scratch_doubles() + 64);
    double *res = scr + reduce_scratch_doubles();
    launch_sum_step(a, b, n, scr, res, c.st);
    double h[2];
    d2h(h, res, 2 * sizeof(double));
    sync();
    if (h[1] == 0.0) {
        fprintf(stderr, "<warning>: denominator is zero in step length calculation, returning 0.\n");
        return 0.0f;
    }
    return (float)(h[0] / h[1]);
}

// compute_model_perturbation (optimization.c:178-298): min/max, histogram and
// the slowness sum on the device, the 1001-bin growth on the host; out =
// -scalar * g.  g and out may not alias.
float model_perturbation_device(float fraction, const float *sl, const float *g, float *out,
                                int len)
{
    Ctx &c = ctx();
    float *fs = c.misc[2].as<float>(2 * 512 + 64 + 1001 + 64);
    int *hist = reinterpret_cast<int *>(fs + 2 * 512 + 64);
    double *scr = c.scratch.as<double>(reduce_scratch_doubles() + 64);
    return model_perturbation_with(fraction, sl, g, out, len, fs, hist, scr);
}

// The same with the caller's scratch: mm 2 * 512 + 2 floats (min/max block
// partials, then the result), hist 1001 ints (cleared here), red
// reduce_scratch_doubles() + 2 doubles (FP64 block partials, then the sums).
float model_perturbation_with(float fraction, const float *sl, const float *g, float *out, int len,
                              float *fs, int *hist, double *scr)
{
    Ctx &c = ctx();
    launch_minmax(g, len, fs, fs + 2 * 512, c.st);
    float mm[2];
    d2h(mm, fs + 2 * 512, 2 * sizeof(float));
    sync();
    const float Min = mm[0], Max = mm[1];
    const float interval = (float)((Max - Min) / 1000.0);
    launch_fill_u32(reinterpret_cast<unsigned *>(hist), 1001, 0u, c.st);
    launch_histogram(g, len, Min, interval, hist, c.st);
    int h[1001];
    d2h(h, hist, sizeof(h));
    double *res = scr + reduce_scratch_doubles();
    launch_sum(sl, len, scr, res, c.st);
    double avg[2];
    d2h(avg, res, 2 * sizeof(double));
    sync();
    int peak = 0, pos = 0;
    for (int i = 0; i < 1001; ++i)
        if (h[i] > peak) {
            peak = h[i];
            pos = i;
        }
    int hi = pos, lo = pos;
    const float cap = 0.99f;
    float acc = (float)h[hi] / (float)len;
    while (acc < cap) {
        bool moved = false;
        if (hi < 1000) {
            hi++;
            acc += (float)h[hi] / (float)len;
            moved = true;
        }
        if (lo > 0) {
            lo--;
            acc += (float)h[lo] / (float)len;
            moved = true;
        }
        if (!moved) break;
    }
    float a = Min + (float)lo * interval, b = Min + (float)hi * interval;
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    float scalar = (a > b) ? a : b;
    scalar = (float)((avg[0] / len * (double)fraction) / (double)scalar);
    launch_scale(g, out, len, scalar, c.st);
    return scalar;
}

// smooth2D (common.c:927-1040) of a device [nx][nz] array.
void smooth2d_device(const float *in, float *out, int nx, int nz, int N, int zdir)
{
    Ctx &c = ctx();
    std::vector<float> w(2 * N + 1);
    for (int i = 0; i < 2 * N + 1; ++i) {
        w[i] = (float)cos(3.141592653589793 * (i - N) / (2.0 * N));
        w[i] *= w[i];
    }
    float *wd = c.taper.as<float>(w.size() + 64);
    h2d(wd, w.data(), w.size() * sizeof(float));
    launch_smooth2d(in, out, nx, nz, N, zdir, wd, c.st);
    sync();  // the weight table is reused by the next call
}

}  // namespace fwiedu

// Per-shot engine (propagation.c:356-775) with the reference's host checkpoint.
extern "C" void iso_acoustic2d_propagation_engine(Record2D *record2d, Acoustic2dCheckPoint *chkpt,
                                                  float **vpm, float **rhom, int nx, int nz,
                                                  float dx, float dz, float dt, int nt,
                                                  int pmlThick, int freesurface, int precondition,
                                                  int mode, int iter, int verbose, float **grad)
{
    check_common(nz, nx, dx, dz, dt, nt, pmlThick, freesurface);
    // the reference's engine names its verbose wavefield files after a shot
    // index it never assigns (propagation.c:399, 597); a call of this engine
    // is one shot, named shot 0 here
    const Verbose vb = verbose_from_env(verbose > 1 ? verbose : 0, iter);
    if (mode < 0 || mode > 2) fail("<error>: engine mode must be 0, 1 or 2 (got %d)", mode);
    if (mode != 0 && !chkpt) fail("<error>: engine mode %d needs a checkpoint", mode);
    if ((mode == 2 || (mode == 1 && precondition == 1)) && (!grad || !grad[0]))
        fail("<error>: engine mode %d needs an output array", mode);
    if (mode == 2) {
        if (chkpt->nt != nt || !chkpt->p || !chkpt->status)
            fail("<error>: checkpoint not allocated for nt=%d", nt);
        // the reverse pass restarts each segment from the pair the forward
        // pass stored (get_acoustic2dCheckPointStatus, checkpoint.c:236-265,
        // finds it through status): every pair must be there.  Fields the
        // status marks beyond the pairs are recomputed, as the reference's
        // first reverse pass after mode 1 does; their values are the same.
        const Schedule sc0(nt);
        for (int it = 0; it < nt; ++it)
            if (sc0.ckslot[it] != -1 && chkpt->status[it] != 1)
                fail("<error>: iso_acoustic2d_propagation_engine mode 2: the checkpoint holds no "
                     "field at it=%d (status %d); run mode 1 with this checkpoint first",
                     it, chkpt->status[it]);
    }
    ensure_ready();
    Ctx &c = ctx();
    const long nmod = (long)nx * nz;
    float *vpd = c.model[0].as<float>(nmod);
    float *rhod = c.model[1].as<float>(nmod);
    h2d(vpd, vpm[0], nmod * sizeof(float));
    h2d(rhod, rhom[0], nmod * sizeof(float));
    Geom g = make_geom(nx, nz, pmlThick, freesurface, dx, dz, dt);
    float *coef = c.coef.as<float>((size_t)CO_ALLOC * g.fsz);
    build_coefs(g, vpd, rhod, record2d->fpeak, coef);
    Schedule sc(nt);
    BatchState b;
    std::vector<Record2D *> recs = {record2d};
    if (c.tile_req.tz > 0 && c.tile_req.kind >= 2)
        fail("<error>: iso_acoustic2d_propagation_engine runs one shot; the shots-in-lanes "
             "stencil needs a batch");
    // the host checkpoint of the reference is FP32
    prepare_batch(b, g, coef, recs, nt, mode != 0, sc, false, false);
    if (vb.snaps()) {
        b.vb = &vb;
        b.snap_shot.assign(1, 0);
    }
    const long nacc = b.nacc;
    if (mode == 0 || mode == 1) {
        forward_pass(b, mode == 1, mode == 1 && precondition == 1);
        rec_d2h(record2d, b.rec);
        if (mode == 1) {
            // hand the pairs (kd-1, kd) to the host checkpoint, status set
            if (chkpt->nt != nt || !chkpt->p || !chkpt->status)
                fail("<error>: checkpoint not allocated for nt=%d", nt);
            for (int it = 0; it < nt; ++it) {
                if (sc.ckslot[it] == -1) continue;
                if (sc.ckslot[it] == -2)
                    memset(chkpt->p[it][0], 0, (size_t)g.Nx * g.Nz * sizeof(float));
                else
                    HIPCHECK(hipMemcpy2DAsync(chkpt->p[it][0], g.Nz * sizeof(float),
                                              b.F(BatchPlan::CK0 + sc.ckslot[it]),
                                              g.ld * sizeof(float), g.Nz * sizeof(float), g.Nx,
                                              hipMemcpyDeviceToHost, c.st));
                chkpt->status[it] = 1;
            }
            if (precondition == 1) {
                float *pb = c.misc[0].as<float>(nacc);
                launch_finalize(b.acc, b.ASS, b.ACS, pb, nacc, 1, vpd, rhod, (int)nacc, 4.0, 1, c.st);
                d2h(grad[0], pb, nacc * sizeof(float));
            }
        }
    } else {
        for (int it = 0; it < nt; ++it) {
            if (sc.ckslot[it] < 0) continue;
            HIPCHECK(hipMemcpy2DAsync(b.F(BatchPlan::CK0 + sc.ckslot[it]), g.ld * sizeof(float),
                                      chkpt->p[it][0], g.Nz * sizeof(float), g.Nz * sizeof(float),
                                      g.Nx, hipMemcpyHostToDevice, c.st));
        }
        rec_h2d(b.rec, record2d);
        if (record2d->sampleRate != 1 || record2d->ns != nt) {
            // propagation.c:635
            b.adjfull = c.adjfull.as<float>((size_t)b.full_total + 64);
            interpolate_traces_device(b.rec, record2d->ntr, record2d->ns, record2d->sampleRate,
                                      b.adjfull, nt);
        }
        adjoint_pass(b);
        float *gb = c.misc[0].as<float>(nacc);
        launch_finalize(b.acc, b.ASS, b.ACS, gb, nacc, 1, vpd, rhod, (int)nacc, 2.0, 0, c.st);
        d2h(grad[0], gb, nacc * sizeof(float));
    }
    sync();
    collect_timing();
}
