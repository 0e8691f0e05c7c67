// C-ABI entry points besides the propagation engine: optimisation helpers
// (optimization.c), host checkpoint bookkeeping (checkpoint.c), single-sweep
// helpers (propagation.c:13-101, 192-354) and the fwiedu_* controls.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "engine.hpp"

using namespace fwiedu;

namespace {

// Upload the trace blocks [ntr][ns] of Nshot records back to back.
float *upload_records(DevBuf &buf, Record2D *r, int Nshot, long &n)
{
    n = 0;
    for (int s = 0; s < Nshot; ++s) n += (long)r[s].ntr * r[s].ns;
    float *d = buf.as<float>(n + 64);
    long off = 0;
    for (int s = 0; s < Nshot; ++s) {
        const long k = (long)r[s].ntr * r[s].ns;
        h2d(d + off, r[s].traces_p[0], k * sizeof(float));
        off += k;
    }
    return d;
}

double *scratch_res()
{
    Ctx &c = ctx();
    return c.scratch.as<double>(reduce_scratch_doubles() + 64);
}

}  // namespace

// ---------------------------------------------------------------------------
// optimization.c

// compute_adjoint_source2d, optimization.c:300-419
extern "C" void compute_adjoint_source2d(Record2D *record2d_pre, Record2D *record2d_obs,
                                         float **vp, float freq, int filter,
                                         Record2D *record2d_adjsource)
{
    ensure_ready();
    Ctx &c = ctx();
    Record2D *a = record2d_adjsource;
    const long n = (long)a->ntr * a->ns;
    float *pre = c.rec.as<float>(n + 64), *obs = c.obs.as<float>(n + 64);
    h2d(pre, record2d_pre->traces_p[0], n * sizeof(float));
    h2d(obs, record2d_obs->traces_p[0], n * sizeof(float));
    adjoint_source_device(a, vp, freq, filter, pre, obs, pre);
    d2h(a->traces_p[0], pre, n * sizeof(float));
    sync();
}

// calculate_objective_value, optimization.c:99-131
extern "C" float calculate_objective_value(Record2D *record2d, int Nshot)
{
    ensure_ready();
    Ctx &c = ctx();
    long n;
    float *d = upload_records(c.rec, record2d, Nshot, n);
    return objective_device(d, n);
}

// calculate_step_length, optimization.c:134-175
extern "C" float calculate_step_length(Record2D *record2d_pre0, Record2D *record2d_pre1, int Nshot)
{
    ensure_ready();
    Ctx &c = ctx();
    long n0, n1;
    float *a = upload_records(c.rec, record2d_pre0, Nshot, n0);
    float *b = upload_records(c.obs, record2d_pre1, Nshot, n1);
    if (n0 != n1) fail("<error>: calculate_step_length: record sizes differ");
    return step_length_device(a, b, n0);
}

// compute_model_perturbation, optimization.c:178-298: min/max, histogram and
// the slowness mean on the device; the 1001-bin growth loop on the host.
extern "C" float compute_model_perturbation(float fraction, float *slowness, float *gradient,
                                            float *out, int len)
{
    ensure_ready();
    Ctx &c = ctx();
    float *g = c.misc[0].as<float>((size_t)len + 64);
    float *sl = c.misc[1].as<float>((size_t)len + 64);
    h2d(g, gradient, (size_t)len * sizeof(float));
    h2d(sl, slowness, (size_t)len * sizeof(float));
    // the slowness sum is taken before the result overwrites sl (stream order)
    const float scalar = model_perturbation_device(fraction, sl, g, sl, len);
    d2h(out, sl, (size_t)len * sizeof(float));
    sync();
    return scalar;
}

// apply_gradient_shot_precondition, optimization.c:58-97
extern "C" void apply_gradient_shot_precondition(float *grad, float *precond, int n)
{
    ensure_ready();
    Ctx &c = ctx();
    float *g = c.misc[0].as<float>((size_t)n + 64), *p = c.misc[1].as<float>((size_t)n + 64);
    h2d(g, grad, (size_t)n * sizeof(float));
    h2d(p, precond, (size_t)n * sizeof(float));
    double *scr = c.scratch.as<double>(reduce_scratch_doubles() + 1024);
    launch_shot_precondition(g, p, n, 1, n, scr, c.st);
    d2h(grad, g, (size_t)n * sizeof(float));
    d2h(precond, p, (size_t)n * sizeof(float));
    sync();
}

// apply_gradient_precondition, optimization.c:13-56
extern "C" void apply_gradient_precondition(float *grad, const float *precond, int n, float alpha)
{
    ensure_ready();
    Ctx &c = ctx();
    float *g = c.misc[0].as<float>((size_t)n + 64), *p = c.misc[1].as<float>((size_t)n + 64);
    h2d(g, grad, (size_t)n * sizeof(float));
    h2d(p, precond, (size_t)n * sizeof(float));
    launch_global_precondition(g, p, n, alpha, scratch_res(), c.st);
    d2h(grad, g, (size_t)n * sizeof(float));
    sync();
}

// ---------------------------------------------------------------------------
// propagation.c single-sweep helpers on host arrays

// acoustic_iso2d_fd_it_6order, propagation.c:13-74.  The box the sweep reads
// is gathered into a local padded layout, swept on the device, and the written
// box is scattered back.
extern "C" void acoustic_iso2d_fd_it_6order(float **p2, float **p1, float **p0, float **invrho,
                                            float **invrhox, float **invrhoz, float **K,
                                            float **coef1, float **coef2, float **coef3,
                                            float invdx, float invdz, float invdxdx,
                                            float invdzdz, int ixStart, int ixEnd, int izStart,
                                            int izEnd)
{
    if (ixEnd < ixStart || izEnd < izStart) return;
    ensure_ready();
    Ctx &c = ctx();
    const int nxi = ixEnd - ixStart + 1, nzi = izEnd - izStart + 1;
    const int Nx = nxi + 6, Nz = nzi + 6, ld = (Nz + 63) / 64 * 64;
    const long fsz = (long)Nx * ld;
    std::vector<float> h((size_t)(CO_N + 2) * fsz, 0.0f);
    float **co[CO_N];
    co[CO_C1] = coef1;
    co[CO_C2] = coef2;
    co[CO_A] = coef3;  // multiplied by K below
    co[CO_K] = K;
    co[CO_IR] = invrho;
    co[CO_IRX] = invrhox;
    co[CO_IRZ] = invrhoz;
    for (int lx = 0; lx < Nx; ++lx) {
        const int ix = ixStart - 3 + lx;
        const bool inner = lx >= 3 && lx < 3 + nxi;
        for (int lz = 0; lz < Nz; ++lz) {
            const int iz = izStart - 3 + lz;
            const long i = (long)lx * ld + lz;
            h[(size_t)(CO_N)*fsz + i] = p1[ix][iz];  // p1 with its halo
            if (inner && lz >= 3 && lz < 3 + nzi) {
                h[(size_t)(CO_N + 1) * fsz + i] = p0[ix][iz];
                for (int k = 0; k < CO_N; ++k) h[(size_t)k * fsz + i] = co[k][ix][iz];
                h[(size_t)CO_A * fsz + i] = coef3[ix][iz] * K[ix][iz];  // coef3*K, as at l.71
            }
        }
    }
    float *d = c.misc[0].as<float>((size_t)(CO_N + 3) * fsz);
    h2d(d, h.data(), h.size() * sizeof(float));
    HIPCHECK(hipMemsetAsync(d + (CO_N + 2) * fsz, 0, fsz * sizeof(float), c.st));
    FdArgs a;
    memset(&a, 0, sizeof(a));
    a.coef = d;
    a.cur = d + CO_N * fsz;
    a.prev = d + (CO_N + 1) * fsz;
    a.out = d + (CO_N + 2) * fsz;
    a.nshot = 1;
    a.Nx = Nx;
    a.ld = ld;
    a.ixs = 3;
    a.ixe = 3 + nxi - 1;
    a.izs = 3;
    a.ize = 3 + nzi - 1;
    a.invdx = invdx;
    a.invdz = invdz;
    a.invdxdx = invdxdx;
    a.invdzdz = invdzdz;
    a.cell_stride = 1;
    a.fma = c.fma;
    launch_fd(FD_ADJ, a, FdTile(), c.st);
    std::vector<float> o((size_t)fsz);
    d2h(o.data(), d + (CO_N + 2) * fsz, fsz * sizeof(float));
    sync();
    for (int lx = 3; lx < 3 + nxi; ++lx)
        for (int lz = 3; lz < 3 + nzi; ++lz)
            p2[ixStart - 3 + lx][izStart - 3 + lz] = o[(size_t)lx * ld + lz];
}

// acoustic_iso2d_freesurface, propagation.c:76-101
extern "C" void acoustic_iso2d_freesurface(float **p2, int Nx, int Nz, int pmlThick)
{
    ensure_ready();
    Ctx &c = ctx();
    const int rows = 2 * pmlThick + 1;
    if (rows > Nz) fail("<error>: acoustic_iso2d_freesurface: 2*pmlThick+1 > Nz");
    const int ld = (rows + 63) / 64 * 64;
    std::vector<float> h((size_t)Nx * ld, 0.0f);
    for (int ix = 0; ix < Nx; ++ix) memcpy(&h[(size_t)ix * ld], p2[ix], rows * sizeof(float));
    float *d = c.misc[0].as<float>(h.size());
    h2d(d, h.data(), h.size() * sizeof(float));
    launch_freesurface(d, Nx, ld, pmlThick, c.st);
    d2h(h.data(), d, h.size() * sizeof(float));
    sync();
    for (int ix = 0; ix < Nx; ++ix) memcpy(p2[ix], &h[(size_t)ix * ld], rows * sizeof(float));
}

// mealCoefs2d, propagation.c:192-354 (note the dz, dx argument order)
extern "C" void mealCoefs2d(float **D, float **Beta, float **vp, float **vs, float fpeak,
                           float eal_alpha, float dz, float dx, int NZ, int NX, int nlayer)
{
    ensure_ready();
    Ctx &c = ctx();
    if (eal_alpha == -1.0f) eal_alpha = 2.5f;
    Geom g;
    memset(&g, 0, sizeof(g));
    g.Nx = NX;
    g.Nz = NZ;
    g.ld = (NZ + 63) / 64 * 64;
    g.dx = dx;
    g.dz = dz;
    g.fsz = (long)NX * g.ld;
    const int nvs = vs ? 1 : 0;
    std::vector<float> h((size_t)(1 + nvs) * g.fsz, 0.0f);
    for (int ix = 0; ix < NX; ++ix) {
        memcpy(&h[(size_t)ix * g.ld], vp[ix], NZ * sizeof(float));
        if (vs) memcpy(&h[(size_t)(g.fsz + ix * g.ld)], vs[ix], NZ * sizeof(float));
    }
    std::vector<float> w = meal_weights(eal_alpha, nlayer);
    float *d = c.misc[0].as<float>((size_t)(3 + nvs) * g.fsz + w.size() + 64);
    float *dv = d, *dvs = vs ? d + g.fsz : nullptr;
    float *dD = d + (1 + nvs) * g.fsz, *dB = dD + g.fsz, *dw = dB + g.fsz;
    h2d(d, h.data(), h.size() * sizeof(float));
    h2d(dw, w.data(), w.size() * sizeof(float));
    launch_meal(dv, dvs, dD, dB, g, nlayer, dw, meal_lnR(nlayer, 2.0f), 2.0f, sqrtf(3.0f), fpeak,
                c.st);
    std::vector<float> o((size_t)2 * g.fsz);
    d2h(o.data(), dD, o.size() * sizeof(float));
    sync();
    for (int ix = 0; ix < NX; ++ix) {
        memcpy(D[ix], &o[(size_t)ix * g.ld], NZ * sizeof(float));
        memcpy(Beta[ix], &o[(size_t)(g.fsz + ix * g.ld)], NZ * sizeof(float));
    }
}

// ---------------------------------------------------------------------------
// checkpoint.c host bookkeeping (same schedule and slot map)

extern "C" void acoustic2dCheckPoint_init(Acoustic2dCheckPoint *chkpt)
{
    chkpt->status = nullptr;
    chkpt->p = nullptr;
    chkpt->storage_p = nullptr;
}

static void free_storage(float ***s)
{
    if (!s) return;
    free(s[0][0]);
    free(s[0]);
    free(s);
}

extern "C" void alloc_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, int NZ, int NX, int nt)
{
    chkpt->nt = nt;
    chkpt->dCheckPoint = std::max((int)roundf(sqrtf(2.0f * (nt - 1))), 3);
    const int d = chkpt->dCheckPoint;
    free(chkpt->status);
    chkpt->status = (int *)calloc(nt, sizeof(int));
    free(chkpt->p);
    chkpt->p = (float ***)calloc(nt, sizeof(float **));
    int nc = 0;
    for (int it = 0; it < nt; ++it)
        if (it % d == 0 || (it + 1) % d == 0) nc++;
    chkpt->nCheckPoint = nc;
    chkpt->ntPoint = nc + d - 2;
    // checkpoint.c:296: the int product times FSIZE (a size_t), so the byte
    // count itself is formed in size_t (3.3e9 bytes at 3080 x 1080, nt 8000)
    chkpt->memCost =
        1.0f * (float)((size_t)(chkpt->ntPoint * NZ * NX) * sizeof(float)) / powf(1024.0f, 3);
    free_storage(chkpt->storage_p);
    const int np = chkpt->ntPoint;
    float ***s = (float ***)malloc((size_t)np * sizeof(float **));
    float **rows = (float **)malloc((size_t)np * NX * sizeof(float *));
    float *blk = (float *)calloc((size_t)np * NX * NZ, sizeof(float));
    if (!s || !rows || !blk || !chkpt->status || !chkpt->p)
        fail("<error>: alloc_acoustic2dCheckPoint: out of host memory");
    for (int k = 0; k < np; ++k) {
        s[k] = rows + (size_t)k * NX;
        for (int ix = 0; ix < NX; ++ix) s[k][ix] = blk + ((size_t)k * NX + ix) * NZ;
    }
    chkpt->storage_p = s;
    int ichk = 0, icur = nc;
    for (int it = 0; it < nt; ++it) {
        if (it % d == 0 || (it + 1) % d == 0) {
            chkpt->p[it] = s[ichk++];
            icur = nc;
        } else {
            chkpt->p[it] = s[icur++];
        }
    }
}

extern "C" void free_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt)
{
    free(chkpt->status);
    chkpt->status = nullptr;
    free(chkpt->p);
    chkpt->p = nullptr;
    free_storage(chkpt->storage_p);
    chkpt->storage_p = nullptr;
}

extern "C" void save_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, float **p0, float **p1,
                                          int NP, int it)
{
    if (it < 0 || it >= chkpt->nt || it % chkpt->dCheckPoint) return;
    if (it - 1 >= 0) {
        memcpy(chkpt->p[it - 1][0], p0[0], (size_t)NP * sizeof(float));
        chkpt->status[it - 1] = 1;
    }
    memcpy(chkpt->p[it][0], p1[0], (size_t)NP * sizeof(float));
    chkpt->status[it] = 1;
}

extern "C" void get_acoustic2dCheckPoint(Acoustic2dCheckPoint *chkpt, float **p0, float **p1, int NP,
                                         int it)
{
    if (it < 0 || it >= chkpt->nt || it % chkpt->dCheckPoint) return;
    if (it - 1 >= 0)
        memcpy(p0[0], chkpt->p[it - 1][0], (size_t)NP * sizeof(float));
    else
        memset(p0[0], 0, (size_t)NP * sizeof(float));
    memcpy(p1[0], chkpt->p[it][0], (size_t)NP * sizeof(float));
}

extern "C" void save_acoustic2dWavefield(Acoustic2dCheckPoint *chkpt, float **p, int NP, int it)
{
    if (it < 0 || it >= chkpt->nt) return;
    memcpy(chkpt->p[it][0], p[0], (size_t)NP * sizeof(float));
    chkpt->status[it] = 1;
}

extern "C" void get_acoustic2dWavefield(Acoustic2dCheckPoint *chkpt, float **p, int NP, int it)
{
    if (it < 0 || it >= chkpt->nt) return;
    memcpy(p[0], chkpt->p[it][0], (size_t)NP * sizeof(float));
}

extern "C" int get_acoustic2dCheckPointStatus(Acoustic2dCheckPoint *chkpt, int it)
{
    if (chkpt->status[it] == 1) return -1;
    int s = it;
    while (s >= 0 && chkpt->status[s] == 0) s--;
    return s;
}

extern "C" float calculate_memory_cost_checkpoint(int nt, int NX, int NZ)
{
    const int d = std::max((int)roundf(sqrtf(2.0f * (nt - 1))), 3);
    int nc = 0;
    for (int it = 0; it < nt; ++it)
        if (it % d == 0 || (it + 1) % d == 0) nc++;
    const int np = nc + d - 2;
    return 1.0f * (float)((size_t)(np * NZ * NX) * sizeof(float)) / powf(1024.0f, 3);
}

// ---------------------------------------------------------------------------
// fwiedu_* controls

extern "C" void fwiedu_set_batch(int s) { ctx().batch_req = s < 0 ? 0 : s; }
extern "C" void fwiedu_set_tile_order(int order)
{
    if (order < 0 || order > 2) fail("<error>: fwiedu_set_tile_order: 0, 1 or 2 (got %d)", order);
    ctx().tile_req.zfast = order;
}

extern "C" void fwiedu_set_tile(int tz, int spb)
{
    Ctx &c = ctx();
    if (tz == 0) {
        c.tile_req = FdTile();
        c.tile_req.tz = 0;
        return;
    }
    if ((tz != 32 && tz != 64) || (spb != 1 && spb != 2 && spb != 4 && spb != 8))
        fail("<error>: fwiedu_set_tile: tile height 32 or 64, shots per block 1, 2, 4 or 8");
    c.tile_req.tz = tz;
    c.tile_req.spb = spb;
    c.tile_req.kind = 0;
    c.tile_req.txm = 0;
}
extern "C" void fwiedu_set_march(int txm, int spb, int pd)
{
    Ctx &c = ctx();
    if ((txm != 16 && txm != 32 && txm != 64) || (spb != 1 && spb != 2 && spb != 4) ||
        (pd != 1 && pd != 2 && pd != 4))
        fail("<error>: fwiedu_set_march: txm 16, 32 or 64 columns, shots per block 1, 2 or 4, "
             "prefetch depth 1, 2 or 4");
    c.tile_req.tz = 64;
    c.tile_req.kind = 1;
    c.tile_req.txm = txm;
    c.tile_req.spb = spb;
    c.tile_req.pd = pd;
}
extern "C" void fwiedu_set_xcd_remap(int on) { ctx().tile_req.remap = on ? 1 : 0; }
extern "C" void fwiedu_set_lanes(int wpb, int zrun, int pd, int lds)
{
    Ctx &c = ctx();
    const bool zok = zrun == 8 || zrun == 16 || zrun == 32 || zrun == 64 || zrun == 128 ||
                     zrun == 256;
    if (lds == 2) {
        if (wpb != 8 || zrun < 8 || zrun > 256 || pd < 2 || pd > 5)
            fail("<error>: fwiedu_set_lanes (register ring): 8 columns per block, 8 to 256 rows "
                 "per block, 2 to 5 rows of loads in flight");
    } else if (lds) {
        if ((wpb != 4 && wpb != 8 && wpb != 16) || zrun < 8 || zrun > 256 || pd < 1 || pd > 4)
            fail("<error>: fwiedu_set_lanes (LDS row exchange): 4, 8 or 16 columns per block, "
                 "8 to 256 rows per block, 1 to 4 rows of loads in flight");
    } else if ((wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) || !zok || pd < 1 || pd > 3) {
        fail("<error>: fwiedu_set_lanes: 1, 2, 4 or 8 columns per block, 8 to 256 rows per wave, "
             "1 to 3 rows of loads in flight");
    }
    const int rm = c.tile_req.remap, zf = c.tile_req.zfast;
    c.tile_req = FdTile();
    c.tile_req.remap = rm;
    c.tile_req.zfast = zf;
    c.tile_req.tz = 64;
    c.tile_req.kind = lds == 2 ? 4 : (lds ? 3 : 2);
    c.tile_req.wpb = wpb;
    c.tile_req.zrun = zrun;
    c.tile_req.pd = pd;
}
extern "C" void fwiedu_set_fma(int on) { ctx().fma = on ? 1 : 0; }

extern "C" void fwiedu_set_inplace(int on)
{
    Ctx &c = ctx();
    c.inplace = on ? 1 : 0;
    if (c.ready) drop_graphs();  // captured loops hold the other buffer sequence
}

extern "C" void fwiedu_set_graphs(int on)
{
    Ctx &c = ctx();
    c.graphs = on ? 1 : 0;
    if (!on && c.ready) drop_graphs();
}

extern "C" void fwiedu_set_layout(int layout)
{
    if (layout < -1 || layout > 1)
        fail("<error>: fwiedu_set_layout: -1 (automatic), 0 (per-shot arenas) or 1 (shots in "
             "lanes)");
    ctx().layout_req = layout;
}
extern "C" void fwiedu_set_memory_fraction(double f)
{
    ctx().memfrac = (f > 0.0 && f <= 0.98) ? f : 0.85;
}
extern "C" void fwiedu_set_kernel_timing(int every) { ctx().ktiming = every < 0 ? 0 : every; }
extern "C" long fwiedu_lanes_max_shots(int nz, int nx, int pmlThick)
{
    if (nz <= 0 || nx <= 0 || pmlThick < 0)
        fail("<error>: fwiedu_lanes_max_shots: nz %d, nx %d, pmlThick %d", nz, nx, pmlThick);
    const Geom g = make_geom(nx, nz, pmlThick, 0, 1.0f, 1.0f, 1.0f);
    return lanes_max_shots(g.fsz);
}

extern "C" void fwiedu_set_checkpoint_bits(int bits)
{
    if (bits != 32 && bits != 16) fail("<error>: checkpoint bits must be 32 or 16 (got %d)", bits);
    ctx().ckpt_bits = bits;
}

extern "C" double fwiedu_stat(const char *name)
{
    const Stats &s = ctx().stats;
    const std::string n(name ? name : "");
    if (n == "fd_launches") return (double)s.fd_launches;
    if (n == "fd_cells") return (double)s.fd_cells;
    if (n == "fd_timed") return (double)s.fd_timed;
    if (n == "fd_ms_avg") return s.fd_timed ? s.fd_ms_timed / s.fd_timed : 0.0;
    if (n == "fd_ms_timed") return s.fd_ms_timed;
    if (n == "fd_cells_timed") return s.fd_cells_timed;
    if (n == "fd_cells_per_launch") return s.fd_timed ? s.fd_cells_timed / s.fd_timed : 0.0;
    if (n == "fd_shots_per_launch") return (double)s.shots_per_launch;
    // launch shape per stencil mode m: "tile<m>_kind", "tile<m>_wpb", ...; of
    // its launches that stage (v = 1) or add to the accumulator (v = 2):
    // "tile<m><v>_kind", ...
    const bool tm = n.size() > 6 && n.compare(0, 4, "tile") == 0 && n[4] >= '0' && n[4] <= '2';
    const bool tv = tm && n.size() > 7 && n[5] >= '1' && n[5] <= '2' && n[6] == '_';
    if (tm && (n[5] == '_' || tv)) {
        const FdTile &t = tv ? s.tiles_acc[n[4] - '0'][n[5] - '0'] : s.tiles[n[4] - '0'];
        const std::string k = n.substr(tv ? 7 : 6);
        if (k == "kind") return t.kind;
        if (k == "tz") return t.tz;
        if (k == "spb") return t.spb;
        if (k == "txm") return t.txm;
        if (k == "zfast") return t.zfast;
        if (k == "pd") return t.pd;
        if (k == "remap") return t.remap;
        if (k == "wpb") return t.wpb;
        if (k == "zrun") return t.zrun;
    }
    // per stencil mode m (0 forward, 1 reconstruction, 2 adjoint): "fd<m>_..."
    if (n.size() > 4 && n[0] == 'f' && n[1] == 'd' && n[2] >= '0' && n[2] <= '2' && n[3] == '_') {
        const int m = n[2] - '0';
        const std::string k = n.substr(4);
        const double t = (double)s.md_timed[m];
        if (k == "timed") return t;
        if (k == "ms_avg") return t ? s.md_ms[m] / t : 0.0;
        if (k == "cells_per_launch") return t ? s.md_cells[m] / t : 0.0;
        if (k == "shot_cells_per_launch") return t ? s.md_shot_cells[m] / t : 0.0;
        if (k == "acc_shot_cells_per_launch") return t ? s.md_acc_shot_cells[m] / t : 0.0;
    }
    if (n == "batch") return (double)s.batch;
    if (n == "tile_tz") return (double)s.tile_tz;
    if (n == "tile_spb") return (double)s.tile_spb;
    if (n == "tile_kind") return (double)s.tile_kind;
    if (n == "tile_txm") return (double)s.tile_txm;
    if (n == "tile_zfast") return (double)s.tile_zfast;
    if (n == "tile_pd") return (double)s.tile_pd;
    if (n == "tile_remap") return (double)s.tile_remap;
    if (n == "tile_wpb") return (double)s.tile_wpb;
    if (n == "tile_zrun") return (double)s.tile_zrun;
    if (n == "layout") return (double)s.layout;
    if (n == "ckpt_bits") return (double)s.ckpt_bits;
    if (n == "graphs_built") return (double)s.graphs_built;
    if (n == "graph_launches") return (double)s.graph_launches;
    if (n == "fd_inplace") return (double)s.fd_inplace;
    if (n == "shapes_reused") return (double)s.shapes_reused;
    if (n == "h2d_bytes") return s.h2d_bytes;
    if (n == "d2h_bytes") return s.d2h_bytes;
    if (n == "device_bytes") return (double)held_bytes(ctx());
    return -1.0;
}

extern "C" void fwiedu_reset_stats(void) { ctx().stats = Stats(); }

extern "C" void fwiedu_release(void)
{
    Ctx &c = ctx();
    if (!c.ready) return;
    sync();
    release_buffers(c);
    // what the buffers held goes back to the device, not only to the pool
    sync();
    trim_pool();
}

extern "C" int fwiedu_device_ok(void)
{
    int n = 0;
    return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
}

extern "C" const char *fwiedu_version(void) { return "fwiedu-mi355x 0.1 (gfx950)"; }
