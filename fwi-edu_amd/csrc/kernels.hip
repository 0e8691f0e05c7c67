// HIP kernels for gfx950 (CDNA4) of the 2D acoustic FWI hot path.
//
// Arithmetic follows the reference expressions operand for operand
// (FWILab2d/FWILab2dCLangV1.2.6/src/propagation.c, optimization.c, common.c) and
// FMA contraction is off in this file, so every FP32 result is the IEEE result
// of the same sequence of adds and multiplies as a strict C evaluation of the
// reference.  The stencil is HBM-bound (about 1.3 flop per byte), so the fused
// multiply-adds would buy nothing.
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include "kernels.hpp"

#include <float.h>
#include <math.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace fwiedu {

// Stencil launches go through fd_go: with events set (set_stencil_events) the
// launch carries them in its own dispatch (hipExtLaunchKernel), so their
// elapsed time is the kernel's execution alone, as a kernel trace shows it.
static thread_local hipEvent_t t_kev[2] = {nullptr, nullptr};
static thread_local hipError_t t_kerr = hipSuccess;

void set_stencil_events(hipEvent_t start, hipEvent_t stop)
{
    t_kev[0] = start;
    t_kev[1] = stop;
}

hipError_t take_stencil_launch_error()
{
    const hipError_t e = t_kerr;
    t_kerr = hipSuccess;
    return e;
}

template <class K>
static void fd_go(K kern, dim3 grid, dim3 block, size_t shm, hipStream_t st, const FdArgs &a)
{
    if (t_kev[0]) {
        FdArgs arg = a;
        void *args[] = {&arg};
        const hipError_t e = hipExtLaunchKernel(reinterpret_cast<const void *>(kern), grid, block,
                                                args, shm, st, t_kev[0], t_kev[1], 0);
        if (e != hipSuccess) t_kerr = e;
    } else {
        kern<<<grid, block, shm, st>>>(a);
    }
}

// ---------------------------------------------------------------------------
// The stencil step.
//
// One block = 4 waves.  A tile is TZ rows of z by TX columns of x; a wave's 64
// lanes cover 64/TZ adjacent columns of TZ rows, so every global access of a
// wave is 64/TZ column segments of TZ*4 bytes (coalesced), and each wave makes
// CXP such passes.  TZ = 64 gives 16-column tiles, TZ = 32 gives 32-column
// tiles; the engine picks the height that wastes the fewest rows.
//
// A block advances SPB shots of the same tile in turn.  The six coefficient
// streams of its cells are loaded once into registers and reused for every
// shot; per shot, p1 of the tile plus its 3-cell halo is staged in LDS, and p0
// is loaded before the barrier so all of a thread's loads are in flight
// together.
//
// Epilogue, fused per cell (propagation.c):
//   FD_FWD   : +source in double (l.558), free-surface mirror (l.76-101),
//              pseudo-Hessian every MST steps (l.583-593)
//   FD_RECON : free-surface mirror, then +source in float (l.655-662)
//   FD_ADJ   : free-surface mirror, imaging every MST steps (l.700-713)
// The mirror is written by the cell being mirrored: a cell at row pml+k,
// 1 <= k <= pml, also writes -value to row pml-k; row pml itself becomes -0.
//
// coef3*K is stored as one array (CO_A): the reference's expression
// coef3[i]*K[i]*(...) evaluates coef3*K first, so the product is the same
// FP32 value.
constexpr int NWAVE = 4;
constexpr int CXP = 4;

template <int TZ>
struct Tile {
    static constexpr int CPW = 64 / TZ;           // columns per wave pass
    static constexpr int TX = NWAVE * CPW * CXP;  // tile columns
    static constexpr int HX = TX + 6;
    static constexpr int HZ = TZ + 6;
};

// Blocks are dispatched to the 8 XCDs round-robin by linear id.  Renumber so
// that each XCD receives one contiguous range of logical blocks: neighbouring
// tiles (shared p1 halos) and the shot groups of one tile (shared
// coefficients) then meet in the same L2.
__device__ inline int xcd_remap(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// The update of propagation.c:44-71 for one cell: 6th-order first and second
// derivatives from the stencil arms (zmK / zpK: rows iz -/+ K, xmK / xpK:
// columns ix -/+ K), then c1*p1 + c2*p0 + (c3*K)*(rho^-1_x dp/dx + rho^-1_z
// dp/dz + rho^-1 (d2p/dx2 + d2p/dz2)).  FMA = false rounds every product and
// sum in the order of the C source (bit-identical to the strict-order FP32
// build); FMA = true fuses each product into the following sum, as an
// FMA-contracting compile of the reference does (it is built with -ffast-math).
template <bool FMA>
__device__ __forceinline__ float fd_update(float pc, float p0v, float zm1, float zp1, float zm2,
                                           float zp2, float zm3, float zp3, float xm1, float xp1,
                                           float xm2, float xp2, float xm3, float xp3, float k1,
                                           float k2, float kA, float kx, float kz, float kr,
                                           float invdx, float invdz, float invdxdx, float invdzdz)
{
    const float C1 = 0.77793115f, C2 = -0.17388691f, C3 = 0.02338713f;
    const float D0 = -2.8215452f, D1 = 1.57663238f, D2 = -0.18347238f, D3 = 0.01761260f;
    if (!FMA) {
        const float dpdz = (C1 * (zp1 - zm1) + C2 * (zp2 - zm2) + C3 * (zp3 - zm3)) * invdz;
        const float dpdx = (C1 * (xp1 - xm1) + C2 * (xp2 - xm2) + C3 * (xp3 - xm3)) * invdx;
        const float dzz = (D0 * pc + D1 * (zp1 + zm1) + D2 * (zp2 + zm2) + D3 * (zp3 + zm3)) * invdzdz;
        const float dxx = (D0 * pc + D1 * (xp1 + xm1) + D2 * (xp2 + xm2) + D3 * (xp3 + xm3)) * invdxdx;
        return k1 * pc + k2 * p0v + kA * (kx * dpdx + kz * dpdz + kr * (dxx + dzz));
    } else {
        const float dpdz =
            __builtin_fmaf(C3, zp3 - zm3, __builtin_fmaf(C2, zp2 - zm2, C1 * (zp1 - zm1))) * invdz;
        const float dpdx =
            __builtin_fmaf(C3, xp3 - xm3, __builtin_fmaf(C2, xp2 - xm2, C1 * (xp1 - xm1))) * invdx;
        const float d0 = D0 * pc;
        const float dzz =
            __builtin_fmaf(D3, zp3 + zm3, __builtin_fmaf(D2, zp2 + zm2, __builtin_fmaf(D1, zp1 + zm1, d0))) *
            invdzdz;
        const float dxx =
            __builtin_fmaf(D3, xp3 + xm3, __builtin_fmaf(D2, xp2 + xm2, __builtin_fmaf(D1, xp1 + xm1, d0))) *
            invdxdx;
        const float t = __builtin_fmaf(kx, dpdx, __builtin_fmaf(kz, dpdz, kr * (dxx + dzz)));
        return __builtin_fmaf(kA, t, __builtin_fmaf(k1, pc, k2 * p0v));
    }
}

template <int MODE, int TZ, int SPB>
__global__ __launch_bounds__(256) void fd_kernel(FdArgs a)
{
    using T = Tile<TZ>;
    constexpr int NB = (T::HX + NWAVE * T::CPW - 1) / (NWAVE * T::CPW);  // body loads / thread
    static_assert(T::HX * 6 <= 256, "one halo load per thread");
    __shared__ float t[T::HX][T::HZ];

    const int L = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int tile = L / a.ngroups, grp = L - tile * a.ngroups;
    int xt, zt;
    if (a.zfast) {
        zt = tile % a.ztiles;
        xt = tile / a.ztiles;
    } else {
        xt = tile % a.xtiles;
        zt = tile / a.xtiles;
    }
    // tile rows start on 128-byte lines (zbase = izs rounded down to 32 rows);
    // rows outside [izs, ize] only stage data
    const int x0 = a.ixs + xt * T::TX, z0 = ((a.izs >> 5) << 5) + zt * TZ;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int zl = lane % TZ, xl = lane / TZ;
    const int ld = a.ld;
    const long fsz = (long)a.Nx * ld;
    const float *__restrict__ cf = a.coef;

    // this thread's cells and their coefficients (reused for all SPB shots)
    int cell[CXP];
    bool ok[CXP];
    float k1[CXP], k2[CXP], kA[CXP], kx[CXP], kz[CXP], kr[CXP];
#pragma unroll
    for (int j = 0; j < CXP; ++j) {
        const int ix = x0 + w * T::CPW * CXP + j * T::CPW + xl, iz = z0 + zl;
        ok[j] = ix <= a.ixe && iz >= a.izs && iz <= a.ize;
        cell[j] = ix * ld + iz;
        if (ok[j]) {
            k1[j] = cf[CO_C1 * fsz + cell[j]];
            k2[j] = cf[CO_C2 * fsz + cell[j]];
            kA[j] = cf[CO_A * fsz + cell[j]];
            kx[j] = cf[CO_IRX * fsz + cell[j]];
            kz[j] = cf[CO_IRZ * fsz + cell[j]];
            kr[j] = cf[CO_IR * fsz + cell[j]];
        }
    }

    // the LDS slots this thread stages: NB body slots (column c, row zl+3) and
    // at most one halo slot (rows 0..2 and TZ+3..TZ+5)
    int boff[NB];
    bool bok[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int c = w * T::CPW + xl + b * NWAVE * T::CPW;
        const int ix = x0 - 3 + c, iz = z0 + zl;
        bok[b] = c < T::HX;
        boff[b] = (bok[b] && ix < a.Nx && iz < ld) ? ix * ld + iz : -1;
    }
    const bool hok = threadIdx.x < T::HX * 6;
    const int hc = threadIdx.x / 6, hh = threadIdx.x - hc * 6;
    const int hr = hh < 3 ? hh : TZ + hh;
    int hoff = -1;
    {
        const int ix = x0 - 3 + hc, iz = z0 - 3 + hr;
        if (hok && ix < a.Nx && iz >= 0 && iz < ld) hoff = ix * ld + iz;
    }

    // loads of one shot, issued together; the next shot's are issued before the
    // current one is computed so they are in flight during the arithmetic
    float stg[NB], hal = 0.f, q0[CXP];
    auto issue = [&](int s) {
        const long so = (long)s * a.shot_stride;
        const float *__restrict__ p1 = a.cur + so;
        const float *__restrict__ p0 = a.prev + so;
#pragma unroll
        for (int b = 0; b < NB; ++b) stg[b] = boff[b] >= 0 ? p1[boff[b]] : 0.f;
        hal = hoff >= 0 ? p1[hoff] : 0.f;
#pragma unroll
        for (int j = 0; j < CXP; ++j) q0[j] = ok[j] ? p0[cell[j]] : 0.f;
    };

    const int s0 = grp * SPB;
    if (s0 < a.nshot) issue(s0);
    for (int si = 0; si < SPB; ++si) {
        const int s = s0 + si;
        if (s >= a.nshot) break;  // uniform across the block
        const long so = (long)s * a.shot_stride;
        float *__restrict__ p2 = a.out + so;
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (bok[b]) t[w * T::CPW + xl + b * NWAVE * T::CPW][zl + 3] = stg[b];
        if (hok) t[hc][hr] = hal;
        float p0c[CXP];
#pragma unroll
        for (int j = 0; j < CXP; ++j) p0c[j] = q0[j];
        int scell = -1;
        if (MODE != FD_ADJ && a.src_cell) scell = a.src_cell[s];
        __syncthreads();
        if (si + 1 < SPB && s + 1 < a.nshot) issue(s + 1);

#pragma unroll
        for (int j = 0; j < CXP; ++j) {
            if (!ok[j]) continue;
            const int c = 3 + w * T::CPW * CXP + j * T::CPW + xl, r = zl + 3;
            const int i = cell[j];
            const float pc = t[c][r];
            const float p0v = p0c[j];
            const float v =
                a.fma ? fd_update<true>(pc, p0v, t[c][r - 1], t[c][r + 1], t[c][r - 2], t[c][r + 2],
                                        t[c][r - 3], t[c][r + 3], t[c - 1][r], t[c + 1][r],
                                        t[c - 2][r], t[c + 2][r], t[c - 3][r], t[c + 3][r], k1[j],
                                        k2[j], kA[j], kx[j], kz[j], kr[j], a.invdx, a.invdz,
                                        a.invdxdx, a.invdzdz)
                      : fd_update<false>(pc, p0v, t[c][r - 1], t[c][r + 1], t[c][r - 2], t[c][r + 2],
                                         t[c][r - 3], t[c][r + 3], t[c - 1][r], t[c + 1][r],
                                         t[c - 2][r], t[c + 2][r], t[c - 3][r], t[c + 3][r], k1[j],
                                         k2[j], kA[j], kx[j], kz[j], kr[j], a.invdx, a.invdz,
                                         a.invdxdx, a.invdzdz);

            float val = v;
            if (MODE == FD_FWD && i == scell) {
                const double srcv = 1.0 * (double)a.src_wav[(long)s * a.nt + a.it] *
                                    (double)cf[CO_K * fsz + i] * (double)a.dt * (double)a.dt *
                                    (double)a.invdx * (double)a.invdz;
                val = (float)((double)v + srcv);
            }
            float stored = val;
            const int ix = i / ld, iz = i - ix * ld;
            if (a.fs) {
                const int k = iz - a.pml;
                if (k == 0)
                    stored = -0.0f;
                else if (k <= a.pml)
                    p2[(long)ix * ld + (a.pml - k)] = -val;
            }
            if (MODE == FD_RECON && i == scell) {
                stored = stored + a.src_wav[(long)s * a.nt + a.it] * cf[CO_K * fsz + i] * a.dt *
                                      a.dt * a.invdx * a.invdz;
            }
            p2[i] = stored;

            if (MODE != FD_RECON && a.acc_on) {
                const int px = ix - a.pml, pz = iz - a.pml;
                if (px >= 0 && px < a.nx && pz >= 0 && pz < a.nz) {
                    double *acc = a.acc + (long)s * a.acc_stride + ((long)px * a.nz + pz) * a.acc_cell;
                    if (MODE == FD_FWD) {
                        const float q = (stored - 2.0f * pc + p0v) * a.invdtdt;
                        *acc += (double)(q * q);
                    } else {
                        const float q = (stored - 2.0f * pc + p0v) * a.invdtdt * a.u[so + i];
                        *acc += (double)q;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Column-marching stencil.
//
// A block owns a strip of rows (one thread per row, the strip aligned down to a
// 128-byte line so every column access is whole lines) and marches along x
// over mx_len columns for SPB shots.  p1 columns pass through an 8-slot LDS
// ring: each column is loaded from HBM once per block (plus 6 columns at the
// start of the march), the z-neighbours come from the ring column, the
// x-neighbours from the adjacent ring slots.  One barrier per column: the slot
// written for column ix+3 held column ix-5, which no thread still computing
// column ix-1 reads.  The loads of column ix+1 are issued before column ix is
// computed.  Arithmetic and epilogue are those of fd_kernel.
template <int MODE, int SPB, int PD>
__global__ __launch_bounds__(1024) void fd_march_kernel(FdArgs a)
{
    extern __shared__ float ring[];  // [SPB][8][ZB]
    const int ZB = blockDim.x, tid = threadIdx.x;
    // block order: shot group fastest, so the groups of one strip run together
    // and share its coefficient lines in L2
    int b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int grp = b % a.ngroups;
    b /= a.ngroups;
    const int zs = b % a.mz_nseg, xs = b / a.mz_nseg;
    const int xc0 = a.ixs + xs * a.mx_len;
    const int xc1 = min(a.ixe, xc0 + a.mx_len - 1);
    const int cz0 = a.izs + zs * a.mz_len;
    const int cz1 = min(a.ize, cz0 + a.mz_len - 1);
    const int zb = ((cz0 - 3) >> 5) << 5;
    const int rr = zb + tid;
    const bool rok = rr < a.ld;
    const bool comp = rr >= cz0 && rr <= cz1;
    const int s0 = grp * SPB;
    const int ld = a.ld;
    const long fsz = (long)a.Nx * ld;
    const float *__restrict__ cf = a.coef;
    auto R = [&](int si, int col) -> float & { return ring[(si * 8 + (col & 7)) * ZB + tid]; };
    auto Rz = [&](int si, int col, int dz) -> float { return ring[(si * 8 + (col & 7)) * ZB + tid + dz]; };

    // prologue: columns xc0-3 .. xc0+2 (xc0 >= 3)
#pragma unroll
    for (int si = 0; si < SPB; ++si) {
        const int s = s0 + si;
        if (s >= a.nshot) break;
        const float *__restrict__ p1 = a.cur + (long)s * a.shot_stride;
        for (int c = xc0 - 3; c < xc0 + 3; ++c) R(si, c) = (c < a.Nx && rok) ? p1[(long)c * ld + rr] : 0.f;
    }
    // register queue of PD columns in flight: slot q holds column xc0 + q (mod PD)
    float n1[PD][SPB], n0[PD][SPB], nk[PD][6];
    auto issue = [&](int ix, int q) {
#pragma unroll
        for (int si = 0; si < SPB; ++si) {
            const int s = s0 + si;
            n1[q][si] = 0.f;
            n0[q][si] = 0.f;
            if (s >= a.nshot) continue;
            const long so = (long)s * a.shot_stride;
            const int c = ix + 3;
            if (c < a.Nx && rok) n1[q][si] = a.cur[so + (long)c * ld + rr];
            if (comp) n0[q][si] = a.prev[so + (long)ix * ld + rr];
        }
        if (comp) {
            const long i = (long)ix * ld + rr;
            nk[q][0] = cf[CO_C1 * fsz + i];
            nk[q][1] = cf[CO_C2 * fsz + i];
            nk[q][2] = cf[CO_A * fsz + i];
            nk[q][3] = cf[CO_IRX * fsz + i];
            nk[q][4] = cf[CO_IRZ * fsz + i];
            nk[q][5] = cf[CO_IR * fsz + i];
        }
    };
#pragma unroll
    for (int q = 0; q < PD; ++q)
        if (xc0 + q <= xc1) issue(xc0 + q, q);

    for (int ix0 = xc0; ix0 <= xc1; ix0 += PD) {
#pragma unroll
        for (int q = 0; q < PD; ++q) {
            const int ix = ix0 + q;
            if (ix > xc1) break;  // uniform across the block
            float p0c[SPB], k[6];
#pragma unroll
            for (int si = 0; si < SPB; ++si) {
                R(si, ix + 3) = n1[q][si];
                p0c[si] = n0[q][si];
            }
#pragma unroll
            for (int m = 0; m < 6; ++m) k[m] = nk[q][m];
            __syncthreads();
            if (ix + PD <= xc1) issue(ix + PD, q);
            if (!comp) continue;
            const int i = ix * ld + rr;
#pragma unroll
            for (int si = 0; si < SPB; ++si) {
                const int s = s0 + si;
                if (s >= a.nshot) break;
                const long so = (long)s * a.shot_stride;
                float *__restrict__ p2 = a.out + so;
                const float pc = R(si, ix);
                const float p0v = p0c[si];
                const float zm1 = Rz(si, ix, -1), zp1 = Rz(si, ix, 1), zm2 = Rz(si, ix, -2);
                const float zp2 = Rz(si, ix, 2), zm3 = Rz(si, ix, -3), zp3 = Rz(si, ix, 3);
                const float xm1 = R(si, ix - 1), xp1 = R(si, ix + 1), xm2 = R(si, ix - 2);
                const float xp2 = R(si, ix + 2), xm3 = R(si, ix - 3), xp3 = R(si, ix + 3);
                const float v =
                    a.fma ? fd_update<true>(pc, p0v, zm1, zp1, zm2, zp2, zm3, zp3, xm1, xp1, xm2, xp2,
                                            xm3, xp3, k[0], k[1], k[2], k[3], k[4], k[5], a.invdx,
                                            a.invdz, a.invdxdx, a.invdzdz)
                          : fd_update<false>(pc, p0v, zm1, zp1, zm2, zp2, zm3, zp3, xm1, xp1, xm2,
                                             xp2, xm3, xp3, k[0], k[1], k[2], k[3], k[4], k[5],
                                             a.invdx, a.invdz, a.invdxdx, a.invdzdz);
                int scell = -1;
                if (MODE != FD_ADJ && a.src_cell) scell = a.src_cell[s];
                float val = v;
                if (MODE == FD_FWD && i == scell) {
                    const double srcv = 1.0 * (double)a.src_wav[(long)s * a.nt + a.it] *
                                        (double)cf[CO_K * fsz + i] * (double)a.dt * (double)a.dt *
                                        (double)a.invdx * (double)a.invdz;
                    val = (float)((double)v + srcv);
                }
                float stored = val;
                if (a.fs) {
                    const int kk = rr - a.pml;
                    if (kk == 0)
                        stored = -0.0f;
                    else if (kk <= a.pml)
                        p2[(long)ix * ld + (a.pml - kk)] = -val;
                }
                if (MODE == FD_RECON && i == scell) {
                    stored = stored + a.src_wav[(long)s * a.nt + a.it] * cf[CO_K * fsz + i] * a.dt *
                                          a.dt * a.invdx * a.invdz;
                }
                p2[i] = stored;
                if (MODE != FD_RECON && a.acc_on) {
                    const int px = ix - a.pml, pz = rr - a.pml;
                    if (px >= 0 && px < a.nx && pz >= 0 && pz < a.nz) {
                        double *acc = a.acc + (long)s * a.acc_stride + ((long)px * a.nz + pz) * a.acc_cell;
                        if (MODE == FD_FWD) {
                            const float qq = (stored - 2.0f * pc + p0v) * a.invdtdt;
                            *acc += (double)(qq * qq);
                        } else {
                            const float qq = (stored - 2.0f * pc + p0v) * a.invdtdt * a.u[so + i];
                            *acc += (double)qq;
                        }
                    }
                }
            }
        }
    }
}

// Strip plan of the marching kernel: z segments of at most mz_len rows whose
// line-aligned load span (rows cz0-3 .. cz1+3) fits one block of ZB threads.
static void march_plan(FdArgs &a, int mx_len, int &zb_threads)
{
    const int H = a.ize - a.izs + 1;
    int nseg = 1;
    while ((H + nseg - 1) / nseg + 6 + 31 > 1024) ++nseg;
    a.mz_nseg = nseg;
    a.mz_len = (H + nseg - 1) / nseg;
    zb_threads = (a.mz_len + 6 + 31 + 63) / 64 * 64;
    a.mx_len = mx_len;
    a.mx_nseg = (a.ixe - a.ixs + 1 + mx_len - 1) / mx_len;
}

int march_lds_bytes(const FdArgs &a0, FdTile tl)
{
    FdArgs a = a0;
    int zb = 0;
    march_plan(a, tl.txm, zb);
    return tl.spb * 8 * zb * (int)sizeof(float);
}

template <int MODE, int SPB, int PD>
static void launch_march_t(const FdArgs &a0, int mx_len, hipStream_t st)
{
    FdArgs a = a0;
    int zb = 0;
    march_plan(a, mx_len, zb);
    a.ngroups = (a.nshot + SPB - 1) / SPB;
    const long nb = (long)a.mx_nseg * a.mz_nseg * a.ngroups;
    const size_t lds = (size_t)SPB * 8 * zb * sizeof(float);
    fd_go(fd_march_kernel<MODE, SPB, PD>, dim3((unsigned)nb), dim3(zb), lds, st, a);
}

template <int MODE, int PD>
static void launch_march_spb(const FdArgs &a, FdTile tl, hipStream_t st)
{
    switch (tl.spb) {
    case 1: launch_march_t<MODE, 1, PD>(a, tl.txm, st); break;
    case 2: launch_march_t<MODE, 2, PD>(a, tl.txm, st); break;
    default: launch_march_t<MODE, 4, PD>(a, tl.txm, st); break;
    }
}

template <int MODE>
static void launch_march_mode(const FdArgs &a, FdTile tl, hipStream_t st)
{
    if (tl.pd >= 4)
        launch_march_spb<MODE, 4>(a, tl, st);
    else if (tl.pd == 2)
        launch_march_spb<MODE, 2>(a, tl, st);
    else
        launch_march_spb<MODE, 1>(a, tl, st);
}

template <int MODE, int TZ, int SPB>
static void launch_fd_t(const FdArgs &a0, hipStream_t st)
{
    FdArgs a = a0;
    a.xtiles = (a.ixe - a.ixs + 1 + Tile<TZ>::TX - 1) / Tile<TZ>::TX;
    a.ztiles = (a.ize - ((a.izs >> 5) << 5) + 1 + TZ - 1) / TZ;
    a.ngroups = (a.nshot + SPB - 1) / SPB;
    const long nb = (long)a.xtiles * a.ztiles * a.ngroups;
    fd_go(fd_kernel<MODE, TZ, SPB>, dim3((unsigned)nb), dim3(256), 0, st, a);
}

template <int MODE, int TZ>
static void launch_fd_spb(const FdArgs &a0, int spb, int zfast, hipStream_t st)
{
    FdArgs a = a0;
    a.zfast = zfast;
    switch (spb) {
    case 1: launch_fd_t<MODE, TZ, 1>(a, st); break;
    case 2: launch_fd_t<MODE, TZ, 2>(a, st); break;
    case 4: launch_fd_t<MODE, TZ, 4>(a, st); break;
    default: launch_fd_t<MODE, TZ, 8>(a, st); break;
    }
}

template <int MODE>
static void launch_fd_mode(const FdArgs &a, FdTile tl, hipStream_t st)
{
    if (tl.tz == 32)
        launch_fd_spb<MODE, 32>(a, tl.spb, tl.zfast, st);
    else
        launch_fd_spb<MODE, 64>(a, tl.spb, tl.zfast, st);
}

// ---------------------------------------------------------------------------
// Shots in wavefront lanes.
//
// Fields of a batch are laid out [x][z][shot] (shot fastest, padded to a
// multiple of 64), so the 64 lanes of a wave are 64 shots of ONE cell: every
// field access of a wave is one coalesced 256-byte segment, the coefficients
// of the cell are the same for all lanes and come through scalar loads (one
// fetch per 64 shots), and a wave owns one column and marches down it, keeping
// the 7 rows of its z-stencil in registers (one new row per step).

typedef const __attribute__((address_space(4))) float cfloat;
// read-only tables indexed by a wave-uniform value (the receiver row and
// column of a wave's column): through the scalar cache, so that waiting for
// them never waits for the vector loads in flight
typedef const __attribute__((address_space(4))) int cint;

// Field access in the lanes layout through buffer instructions: the row
// offset (cell * CS * 4 bytes) is wave-uniform and goes in the scalar offset,
// the lane (shot) offset is a 32-bit vector offset, so a load or store costs
// no vector address arithmetic.  Byte offsets are formed in signed 32-bit
// arithmetic, so a field must span less than 2 GiB (the host caps the shots
// of a lanes batch accordingly, lanes_max_shots).
typedef __amdgpu_buffer_rsrc_t frsrc;
__device__ __forceinline__ frsrc field_rsrc(const float *p, long bytes)
{
    // dword3 0x00020000: raw 32-bit buffer, the gfx9 family encoding
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0,
                                             (int)(unsigned)bytes, 0x00020000);
}
// Cache policy of the shots-in-lanes stencils' streams that no block of the
// same launch reads again: the previous field (p0, dead once read), the output
// field, the accumulators and staged terms, the forward field of the imaging
// condition.  The non-temporal bit (gfx9 family buffer/global "nt") lets the
// L2 of each XCD evict those lines first, so the lines of the current field
// that neighbouring blocks read as x-halo stay resident (at 221 x 601 with
// 120 shots the stencil launches of a gradient take 6-8 % less time).  The
// current field's loads keep the default policy.
constexpr int CP_STREAM = 2;

template <int CP = 0>
__device__ __forceinline__ float lget(frsrc f, long cell, long CS, unsigned voff)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_raw_buffer_load_b32(f, voff, (int)(cell * CS * 4), CP));
}
template <int CP = 0>
__device__ __forceinline__ void lput(frsrc f, long cell, long CS, unsigned voff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), f, voff,
                                          (int)(cell * CS * 4), CP);
}

// One cell of the stencil for the 64 shots of a wave (shots-in-lanes layout):
// q0..q6 are rows iz-3..iz+3 of the wave's column, x** the row-iz values of
// the columns ix-3..ix+3, p0v the previous field, k the six coefficients c1,
// c2, c3*K, rho^-1_x, rho^-1_z, rho^-1 of the cell; P2 is the output field, U
// the forward field of the imaging condition and s the shot of the lane.
// Accumulators are [cell][shot] (acc_stride 1).  Arithmetic order and the
// epilogue (source, free-surface mirror, accumulation) are those of fd_kernel.
// EPI selects the accumulation epilogue: 0 when a.acc_on says so, loading
// the accumulator (and the adjoint's forward field) here; 1 never (a launch
// with a.acc_on == 0); 2 always, with the cell's accumulator (accv), the
// a.acc_n terms staged by earlier launches (tv[], oldest first) and, in the
// adjoint, the forward-field value (uv) loaded by the caller rows ahead: the
// accumulator becomes ((accv + t0) + t1 ...) + this step's term, the
// reference's sequence of double additions; 3 always, writing this step's
// term (a float: the reference adds its exact double image) to staging slot
// a.acc_slot instead of the accumulator.  Returns the value of the new field
// at the cell (after source and free surface; before the adjoint injection
// that belongs to the next step).  OCP: the cache policy of the output store.
template <int MODE, bool FMA, int EPI = 0, bool K8 = false, int OCP = CP_STREAM>
__device__ __forceinline__ float lanes_update(const FdArgs &a, cfloat *cf, long fsz, long colb,
                                             int ix, int iz, unsigned s, int scell, frsrc P2,
                                             frsrc U, long CS,
                                             float q0, float q1, float q2, float q3, float q4,
                                             float q5, float q6, float xm1, float xp1, float xm2,
                                             float xp2, float xm3, float xp3, float p0v,
                                             const float *k, double accv = 0.0, float uv = 0.f,
                                             bool inj = false, float injv = 0.f,
                                             const float *tv = nullptr)
{
    const long i = colb + iz;
    const float pc = q3;
    const float v = fd_update<FMA>(pc, p0v, q2, q4, q1, q5, q0, q6, xm1, xp1, xm2, xp2, xm3, xp3,
                                   k[0], k[1], k[2], k[3], k[4], k[5], a.invdx, a.invdz, a.invdxdx,
                                   a.invdzdz);

    float val = v;
    if (MODE == FD_FWD && (int)i == scell) {
        const double srcv = 1.0 * (double)a.src_wav[(long)s * a.nt + a.it] *
                            (double)(K8 ? k[6] : cf[CO_K * fsz + i]) * (double)a.dt * (double)a.dt *
                            (double)a.invdx * (double)a.invdz;
        val = (float)((double)v + srcv);
    }
    float stored = val;
    if (a.fs) {
        // free-surface mirror (propagation.c:94-100): of the rows above the
        // surface only the three within the stencil's reach (pml-1..pml-3) are
        // ever read again, by the next step's stencil; the reference's other
        // mirror rows (pml-4..0) are dead stores in this layout, whose fields
        // never leave the device, and are not written (they stay zero)
        const int kk = iz - a.pml;
        if (kk == 0)
            stored = -0.0f;
        else if (kk <= 3 && kk <= a.pml)
            lput(P2, colb + (a.pml - kk), CS, s * 4u, -val);
    }
    if (MODE == FD_RECON && (int)i == scell) {
        stored = stored + a.src_wav[(long)s * a.nt + a.it] * (K8 ? k[6] : cf[CO_K * fsz + i]) * a.dt *
                              a.dt * a.invdx * a.invdz;
    }
    if (MODE == FD_ADJ && inj) {
        // the adjoint source of the next (earlier) step, injected where the
        // separate injection would add it before that step's stencil
        // (propagation.c:681); the imaging below sees the field without it,
        // as it does there
        lput<OCP>(
            P2, i, CS, s * 4u,
            stored - injv * (K8 ? k[6] : cf[CO_K * fsz + i]) * a.dt * a.dt * a.invdx * a.invdz);
    } else {
        lput<OCP>(P2, i, CS, s * 4u, stored);
    }
    if (MODE != FD_RECON && EPI != 1 && (EPI >= 2 || a.acc_on)) {
        const int px = ix - a.pml, pz = iz - a.pml;
        if (px >= 0 && px < a.nx && pz >= 0 && pz < a.nz) {
            const long ai = ((long)px * a.nz + pz) * a.acc_cell + s;
            float term;
            if (MODE == FD_FWD) {
                const float qq = (stored - 2.0f * pc + p0v) * a.invdtdt;
                term = qq * qq;
            } else {
                const float u = EPI >= 2 ? uv : lget<CP_STREAM>(U, i, CS, s * 4u);
                term = (stored - 2.0f * pc + p0v) * a.invdtdt * u;
            }
            if (EPI == 3) {
                __builtin_nontemporal_store(
                    term, &a.accstage[(long)a.acc_slot * a.accstage_stride + ai]);
            } else {
                double old = EPI == 2 ? accv : __builtin_nontemporal_load(&a.acc[ai]);
                if (EPI == 2) {
#pragma unroll
                    for (int j = 0; j < ACC_STAGES; ++j)
                        if (j < a.acc_n) old = old + (double)tv[j];
                }
                __builtin_nontemporal_store(old + (double)term, &a.acc[ai]);
            }
        }
    }
    return stored;
}

// Block -> (z-run, 64-shot group, column group) of the lanes stencils.
// order 0: z-runs fastest, then shot groups (the shot groups of one column
// strip together, sharing its coefficient lines); 1: z-runs, then column
// groups, shot groups outermost; 2: column groups fastest (the blocks that
// share halo columns run side by side and meet in L2), then z-runs, then
// shot groups.
__device__ __forceinline__ void lanes_block(const FdArgs &a, int b, int &zr, int &sg, int &xg)
{
    if (a.lz_order == 2) {
        xg = b % a.lz_nxg;
        b /= a.lz_nxg;
        zr = b % a.lz_nrun;
        sg = b / a.lz_nrun;
        return;
    }
    zr = b % a.lz_nrun;
    b /= a.lz_nrun;
    sg = a.lz_order ? b / a.lz_nxg : b % a.lz_ngrp;
    xg = a.lz_order ? b % a.lz_nxg : b / a.lz_ngrp;
}

// Shots in lanes, no LDS: the 6 x-neighbours of a wave's cell are loaded from
// global memory (the columns of the neighbouring waves, mostly L1/L2 hits).
// The eight loads of the next PD rows are issued before a row is computed, so
// a wave always has memory requests in flight.  (Issuing the six scalar
// coefficient loads that early as well was measured slower: they stay with
// the row's arithmetic.)
template <int MODE, int PD, bool FMA>
__global__ __launch_bounds__(512) void fd_lanes_kernel(FdArgs a)
{
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    int b = blockIdx.x;
    if (a.remap) {
        // hardware dispatch sends block b to XCD b % 8; give each XCD one
        // contiguous range of the logical order so that the x-halo columns of
        // a block are those of blocks on the same XCD (same L2)
        const int per = (int)((gridDim.x + 7) >> 3);
        b = (b & 7) * per + (b >> 3);
        if (b >= a.lz_nblk) return;
    }
    int zr, sg, xg;
    lanes_block(a, b, zr, sg, xg);
    const int ix = a.ixs + xg * wpb + w;
    if (ix > a.ixe) return;  // the whole wave
    const int iz0 = a.izs + zr * a.lz_run;
    const int iz1 = min(a.ize, iz0 + a.lz_run - 1);
    const unsigned s = (unsigned)(sg * 64 + lane);
    const long CS = a.cell_stride;
    const int ld = a.ld;
    const long fsz = (long)a.Nx * ld;
    // coefficients are the same for the 64 lanes: read through the scalar
    // data cache (constant address space, wave-uniform addresses)
    cfloat *cf = (cfloat *)a.coef;
    const long fbytes = fsz * CS * 4;
    const frsrc P1 = field_rsrc(a.cur, fbytes), P0 = field_rsrc(a.prev, fbytes);
    const frsrc P2 = field_rsrc(a.out, fbytes);
    const frsrc U = field_rsrc(MODE == FD_ADJ ? a.u : a.out, fbytes);
    const unsigned vo = s * 4u;
    const int scell = (MODE != FD_ADJ && a.src_cell) ? a.src_cell[s] : -1;
    const long colb = (long)ix * ld;

    // the loads of a row: the new z row (iz+3), the six x neighbours, p0
    auto load_row = [&](int iz, float *r) {
        const long i = colb + iz;
        r[0] = lget(P1, i + 3, CS, vo);
        r[1] = lget(P1, i - ld, CS, vo);
        r[2] = lget(P1, i + ld, CS, vo);
        r[3] = lget(P1, i - 2 * ld, CS, vo);
        r[4] = lget(P1, i + 2 * ld, CS, vo);
        r[5] = lget(P1, i - 3 * ld, CS, vo);
        r[6] = lget(P1, i + 3 * ld, CS, vo);
        r[7] = lget<CP_STREAM>(P0, i, CS, vo);
    };
    float q0 = lget(P1, colb + iz0 - 3, CS, vo), q1 = lget(P1, colb + iz0 - 2, CS, vo);
    float q2 = lget(P1, colb + iz0 - 1, CS, vo), q3 = lget(P1, colb + iz0, CS, vo);
    float q4 = lget(P1, colb + iz0 + 1, CS, vo), q5 = lget(P1, colb + iz0 + 2, CS, vo);
    // software pipeline: the loads of the next PD rows are in flight while a
    // row is computed (rows past iz1 re-load iz1, which stays in bounds)
    float nx_[PD][8];
#pragma unroll
    for (int p = 0; p < PD; ++p) load_row(min(iz0 + p, iz1), nx_[p]);
#pragma unroll 2
    for (int iz = iz0; iz <= iz1; ++iz) {
        float r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = nx_[0][k];
#pragma unroll
        for (int p = 0; p + 1 < PD; ++p)
#pragma unroll
            for (int k = 0; k < 8; ++k) nx_[p][k] = nx_[p + 1][k];
        load_row(min(iz + PD, iz1), nx_[PD - 1]);
        const float q6 = r[0];
        const float xm1 = r[1], xp1 = r[2], xm2 = r[3], xp2 = r[4], xm3 = r[5], xp3 = r[6];
        const float p0v = r[7];
        const long i = colb + iz;
        const float k[6] = {cf[CO_C1 * fsz + i],  cf[CO_C2 * fsz + i],  cf[CO_A * fsz + i],
                            cf[CO_IRX * fsz + i], cf[CO_IRZ * fsz + i], cf[CO_IR * fsz + i]};
        lanes_update<MODE, FMA>(a, cf, fsz, colb, ix, iz, s, scell, P2, U, CS, q0, q1, q2, q3, q4,
                                q5, q6, xm1, xp1, xm2, xp2, xm3, xp3, p0v, k);
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = q4;
        q4 = q5;
        q5 = q6;
    }
}

template <int MODE>
static void launch_lanes_t(const FdArgs &a0, FdTile tl, hipStream_t st)
{
    FdArgs a = a0;
    const int H = a.ize - a.izs + 1;
    a.lz_run = tl.zrun;
    a.lz_nrun = (H + tl.zrun - 1) / tl.zrun;
    a.lz_ngrp = (a.nshot + 63) / 64;
    const int nxg = (a.ixe - a.ixs + 1 + tl.wpb - 1) / tl.wpb;
    const long nb = (long)nxg * a.lz_nrun * a.lz_ngrp;
    a.lz_nblk = (int)nb;
    a.lz_nxg = nxg;
    a.lz_order = tl.zfast;
    const long nlaunch = a.remap ? (nb + 7) / 8 * 8 : nb;
    const dim3 grid((unsigned)nlaunch), block(tl.wpb * 64);
    if (a.fma) {
        if (tl.pd >= 3)
            fd_go(fd_lanes_kernel<MODE, 3, true>, grid, block, 0, st, a);
        else if (tl.pd == 2)
            fd_go(fd_lanes_kernel<MODE, 2, true>, grid, block, 0, st, a);
        else
            fd_go(fd_lanes_kernel<MODE, 1, true>, grid, block, 0, st, a);
    } else {
        if (tl.pd >= 3)
            fd_go(fd_lanes_kernel<MODE, 3, false>, grid, block, 0, st, a);
        else if (tl.pd == 2)
            fd_go(fd_lanes_kernel<MODE, 2, false>, grid, block, 0, st, a);
        else
            fd_go(fd_lanes_kernel<MODE, 1, false>, grid, block, 0, st, a);
    }
}

// Shots in lanes with an LDS row exchange.  A block of WPB waves owns WPB
// adjacent columns and marches down rows [iz0, iz1] in step: at each row every
// wave writes the centre value of its column (64 shots, 256 B) to LDS and the
// six halo columns of the block are loaded once (by waves 0..5 mod WPB), so
// one row of the block costs WPB + 6 global loads of the current field instead
// of 7 * WPB.  The LDS row is double buffered: one barrier per row.  Waves
// whose column lies past ixe still load and publish their column (an active
// neighbour may need it) but compute nothing.
typedef float float8v __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(4))) float8v cfloat8v;

// PCP: the cache policy of the p0 loads and the output stores (the
// non-temporal policy; the default policy for a launch that writes over p0,
// whose two fields the Infinity Cache may then keep)
template <int MODE, int WPB, int PD, bool FMA, int ACC, int PCP = CP_STREAM>
__global__ __launch_bounds__(WPB * 64) void fd_lanes_lds_kernel(FdArgs a)
{
    constexpr int NC = WPB + 6;                    // columns of the LDS row
    constexpr int NH = (6 + WPB - 1) / WPB;        // halo columns per wave (max)
    __shared__ float row[2][NC][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int b = blockIdx.x;
    if (a.remap) {
        const int per = (int)((gridDim.x + 7) >> 3);
        b = (b & 7) * per + (b >> 3);
        if (b >= a.lz_nblk) return;  // the whole block
    }
    int zr, sg, xg;
    lanes_block(a, b, zr, sg, xg);
    const int x0 = a.ixs + xg * WPB;
    const int ix = x0 + w;
    const bool active = ix <= a.ixe;
    const int iz0 = a.izs + zr * a.lz_run;
    const int iz1 = min(a.ize, iz0 + a.lz_run - 1);
    const unsigned s = (unsigned)(sg * 64 + lane);
    const long CS = a.cell_stride;
    const int ld = a.ld;
    const long fsz = (long)a.Nx * ld;
    cfloat *cf = (cfloat *)a.coef;
    const cfloat8v *c8 = (const cfloat8v *)(a.coef + CO_N * fsz);
    const long fbytes = fsz * CS * 4;
    const frsrc P1 = field_rsrc(a.cur, fbytes), P0 = field_rsrc(a.prev, fbytes);
    const frsrc P2 = field_rsrc(a.out, fbytes);
    const frsrc U = field_rsrc(MODE == FD_ADJ ? a.u : a.out, fbytes);
    const unsigned vo = s * 4u;
    const int scell = (MODE != FD_ADJ && a.src_cell) ? a.src_cell[s] : -1;
    const long colb = (long)min(ix, a.Nx - 1) * ld;
    // halo columns carried by this wave: h = w, w + WPB, ... < 6; LDS column
    // j = h (left halo) or WPB + h (right halo)
    int hj[NH];
    long hcol[NH];
#pragma unroll
    for (int k = 0; k < NH; ++k) {
        const int h = w + k * WPB;
        hj[k] = h < 3 ? h : WPB + h;
        const int c = h < 3 ? x0 - 3 + h : x0 + WPB + (h - 3);
        hcol[k] = (long)min(max(c, 0), a.Nx - 1) * ld;
    }
    const int nh = w < 6 ? (6 - w + WPB - 1) / WPB : 0;

    // receiver row of the column inside this run (wave-uniform), its stage
    // slot, and (adjoint) the adjoint source value to inject there, loaded
    // first so that waiting for it never waits for the row loads
    int rr = -1;
    long slot = 0;
    float injv = 0.f;
    if ((MODE == FD_FWD && a.rec_on) || (MODE == FD_ADJ && a.inj_on)) {
        if (active) {
            const int r = ((cint *)a.rrow)[ix];
            if (r >= iz0 && r <= iz1) {
                rr = r;
                slot = (long)((cint *)a.rcol)[ix] * CS + s;
                if (MODE == FD_ADJ) injv = a.stage[(long)(a.it - 1) * a.nslot + slot];
            }
        }
    }
    float q0 = lget(P1, colb + iz0 - 3, CS, vo), q1 = lget(P1, colb + iz0 - 2, CS, vo);
    float q2 = lget(P1, colb + iz0 - 1, CS, vo), q3 = lget(P1, colb + iz0, CS, vo);
    float q4 = lget(P1, colb + iz0 + 1, CS, vo), q5 = lget(P1, colb + iz0 + 2, CS, vo);
    // prefetched rows: own column z+3, p0 at z, halo columns at z; with ACC
    // (the launches whose accumulation epilogue runs: 1 staging this step's
    // term, 2 adding the staged terms and this step's to the accumulator)
    // also the (adjoint) forward-field value and, with 2, the cell's
    // accumulator and staged terms, loaded unconditionally at clamped
    // in-range indices (the epilogue skips cells outside the model), so that
    // the epilogue never waits for a load issued after the prefetched rows
    float pn[PD], pp[PD], ph[PD][NH], pu[PD] = {}, pt[PD][ACC_STAGES] = {};
    double pa[PD] = {};
    long acol = 0;
    if (ACC) acol = (long)min(max(ix - a.pml, 0), a.nx - 1) * a.nz * a.acc_cell + s;
    // byte offsets: column bases and the row stride, so that a row's loads
    // cost one multiply and an add each
    const int rs4 = (int)(CS * 4);
    const int colb4 = (int)(colb * CS * 4);
    int hcol4[NH];
#pragma unroll
    for (int k = 0; k < NH; ++k) hcol4[k] = (int)(hcol[k] * CS * 4);
    auto ldb = [&](frsrc f, int boff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(f, vo, boff, 0));
    };
    auto ldb_stream = [&](frsrc f, int boff) {
        return __builtin_bit_cast(float,
                                  __builtin_amdgcn_raw_buffer_load_b32(f, vo, boff, CP_STREAM));
    };
    auto load_row = [&](int iz, int p) {
        const int ro = iz * rs4;
        pn[p] = ldb(P1, colb4 + ro + 3 * rs4);
        pp[p] = __builtin_bit_cast(float,
                                   __builtin_amdgcn_raw_buffer_load_b32(P0, vo, colb4 + ro, PCP));
#pragma unroll
        for (int k = 0; k < NH; ++k)
            if (k < nh) ph[p][k] = ldb(P1, hcol4[k] + ro);
    };
    // issued after the row that used the slot's previous values is done, so
    // that old and new value can share registers (no copy, hence no wait, at
    // the loop's back edge)
    auto load_acc = [&](int iz, int p) {
        if (ACC) {
            const long ai = acol + (long)min(max(iz - a.pml, 0), a.nz - 1) * a.acc_cell;
            if (ACC == 2) {
                pa[p] = __builtin_nontemporal_load(&a.acc[ai]);
#pragma unroll
                for (int j = 0; j < ACC_STAGES; ++j)
                    pt[p][j] = __builtin_nontemporal_load(&a.accstage[(long)j * a.accstage_stride + ai]);
            }
            if (MODE == FD_ADJ) pu[p] = lget<CP_STREAM>(U, colb + iz, CS, vo);
        }
    };
#pragma unroll
    for (int p = 0; p < PD; ++p) {
        load_row(min(iz0 + p, iz1), p);
        load_acc(min(iz0 + p, iz1), p);
    }
    // the cell's coefficients (c1, c2, c3*K, rho^-1_x, rho^-1_z, rho^-1, K),
    // one 32-byte scalar load per row; with an even PD the next row's are
    // loaded once this row's neighbours are read from LDS, so their latency
    // passes during this row's arithmetic (a scalar load in flight would make
    // every LDS wait a full wait, so none is issued before the LDS reads)
    constexpr int CPD = (PD % 2 == 0) ? 2 : 1;
    float8v kc[CPD];
    if (CPD == 2) kc[0] = c8[colb + iz0];
    // PD rows per outer iteration, each with its own prefetch slot: slot p is
    // consumed and then reloaded with the row PD further down, so no register
    // is copied while its load is in flight and the wait before a row covers
    // only that row's loads (a rotating slot array makes the compiler wait for
    // every outstanding load at each row)
#pragma unroll 1
    for (int izb = iz0; izb <= iz1; izb += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            // rows past iz1 (the last group of a short run) go through the
            // loads and barriers but compute nothing: no early exit, so the
            // prefetched values never merge from two paths (a merge copies a
            // register whose load is in flight and has to wait for it)
            const int iz = izb + p;
            const bool live = iz <= iz1;
            const int buf = iz & 1;
            const float q6 = pn[p], p0v = pp[p];
            row[buf][3 + w][lane] = q3;
#pragma unroll
            for (int k = 0; k < NH; ++k)
                if (k < nh) row[buf][hj[k]][lane] = ph[p][k];
            load_row(min(iz + PD, iz1), p);
            __syncthreads();
            // the receiver sample of the forward: p1 at the receiver row
            // (propagation.c:574), time-major in the stage
            if (MODE == FD_FWD && iz == rr) a.stage[(long)a.it * a.nslot + slot] = q3;
            if (active && live) {
                const float xm1 = row[buf][2 + w][lane], xp1 = row[buf][4 + w][lane];
                const float xm2 = row[buf][1 + w][lane], xp2 = row[buf][5 + w][lane];
                const float xm3 = row[buf][0 + w][lane], xp3 = row[buf][6 + w][lane];
                float8v kv;
                if (CPD == 2) {
                    kv = kc[p & 1];
                    // touch the neighbours first: the next row's coefficient
                    // load goes after the LDS reads
                    __builtin_amdgcn_sched_barrier(0);
                    kc[(p + 1) & 1] = c8[colb + min(iz + 1, iz1)];
                } else {
                    kv = c8[colb + iz];
                }
                const float k[7] = {kv[0], kv[1], kv[2], kv[3], kv[4], kv[5], kv[6]};
                lanes_update<MODE, FMA, ACC == 2 ? 2 : (ACC == 1 ? 3 : 1), true, PCP>(
                    a, cf, fsz, colb, ix, iz, s, scell, P2, U, CS, q0, q1, q2, q3, q4, q5, q6,
                    xm1, xp1, xm2, xp2, xm3, xp3, p0v, k, pa[p], pu[p],
                    MODE == FD_ADJ && iz == rr, injv, pt[p]);
            }
            load_acc(min(iz + PD, iz1), p);
            q0 = q1;
            q1 = q2;
            q2 = q3;
            q3 = q4;
            q4 = q5;
            q5 = q6;
        }
    }
}

template <int MODE, int WPB, bool FMA, int ACC, int PCP = CP_STREAM>
static void launch_lanes_lds_p(const FdArgs &a, const dim3 &grid, int pd, hipStream_t st)
{
    if (pd >= 4)
        fd_go(fd_lanes_lds_kernel<MODE, WPB, 4, FMA, ACC, PCP>, grid, dim3(WPB * 64), 0, st, a);
    else if (pd == 3)
        fd_go(fd_lanes_lds_kernel<MODE, WPB, 3, FMA, ACC, PCP>, grid, dim3(WPB * 64), 0, st, a);
    else if (pd == 2)
        fd_go(fd_lanes_lds_kernel<MODE, WPB, 2, FMA, ACC, PCP>, grid, dim3(WPB * 64), 0, st, a);
    else
        fd_go(fd_lanes_lds_kernel<MODE, WPB, 1, FMA, ACC, PCP>, grid, dim3(WPB * 64), 0, st, a);
}

// The accumulation epilogue is a template choice: the launches without it
// (the reconstruction sweep, four in five forward and adjoint steps) carry
// none of its registers or loads.
template <int MODE, int WPB, bool FMA>
static void launch_lanes_lds_f(const FdArgs &a, const dim3 &grid, int pd, hipStream_t st)
{
    if (MODE == FD_RECON || !a.acc_on)
        launch_lanes_lds_p<MODE, WPB, FMA, 0>(a, grid, pd, st);
    else if (a.acc_stage)
        launch_lanes_lds_p<MODE, WPB, FMA, MODE == FD_RECON ? 0 : 1>(a, grid, pd, st);
    else if (WPB == 8 && !FMA && a.out == a.prev)
        // the launches adding the staged terms beside the register-ring
        // stencil's steps in place (its block shape, strict arithmetic)
        launch_lanes_lds_p<MODE, WPB, FMA, MODE == FD_RECON ? 0 : 2, 0>(a, grid, pd, st);
    else
        launch_lanes_lds_p<MODE, WPB, FMA, MODE == FD_RECON ? 0 : 2>(a, grid, pd, st);
}

template <int MODE, int WPB>
static void launch_lanes_lds_w(const FdArgs &a, const dim3 &grid, int pd, hipStream_t st)
{
    if (a.fma)
        launch_lanes_lds_f<MODE, WPB, true>(a, grid, pd, st);
    else
        launch_lanes_lds_f<MODE, WPB, false>(a, grid, pd, st);
}

template <int MODE>
static void launch_lanes_lds_t(const FdArgs &a0, FdTile tl, hipStream_t st)
{
    FdArgs a = a0;
    const int H = a.ize - a.izs + 1;
    a.lz_run = tl.zrun;
    a.lz_nrun = (H + tl.zrun - 1) / tl.zrun;
    a.lz_ngrp = (a.nshot + 63) / 64;
    const int nxg = (a.ixe - a.ixs + 1 + tl.wpb - 1) / tl.wpb;
    const long nb = (long)nxg * a.lz_nrun * a.lz_ngrp;
    a.lz_nblk = (int)nb;
    a.lz_nxg = nxg;
    a.lz_order = tl.zfast;
    const dim3 grid((unsigned)(a.remap ? (nb + 7) / 8 * 8 : nb));
    if (tl.wpb >= 16)
        launch_lanes_lds_w<MODE, 16>(a, grid, tl.pd, st);
    else if (tl.wpb >= 8)
        launch_lanes_lds_w<MODE, 8>(a, grid, tl.pd, st);
    else
        launch_lanes_lds_w<MODE, 4>(a, grid, tl.pd, st);
}

// Shots in lanes, LDS row exchange, register ring (launch kind 4).  The same
// block shape and arithmetic as fd_lanes_lds_kernel with WPB = 8 (8 adjacent
// columns per block, one wave per column, the row's x-neighbours through a
// double-buffered LDS row, one barrier per row), built so that a row costs
// little beside its arithmetic:
//  * the rows a wave holds (the 7-row z-stencil of its column and the D rows
//    of loads in flight) live in a ring of U registers indexed by
//    (row - first row) mod U, and the row loop is unrolled U rows deep, so
//    every value keeps its register from load to last use and no register is
//    copied at the loop's back edge (a copy of a value whose load is in flight
//    waits for it); U is even (the LDS buffer of a row is its parity) and
//    at least D + 7 (the slot a load fills is free when it is issued);
//  * every field access of a wave goes through a buffer resource based at its
//    own column (one for p1 three rows down, whose rows are the z+3 loads), so
//    all loads of a row share one scalar row offset, advanced by one add per
//    row; the store's resource is based D rows up so that it uses the same
//    offset;
//  * the coefficients of row r + 1 (one 32-byte scalar load of the
//    cell-interleaved copy) come from a group base plus an immediate offset,
//    loaded during row r;
//  * blocks whose rows hold no free-surface row and miss the row ranges of
//    the sources and receivers (kernel arguments: a scalar test, made before
//    any load) take a loop with none of those tests; the others load their
//    source cells and receiver rows and run the epilogue of lanes_update, as
//    fd_lanes_lds_kernel does.  Each branch issues its own first loads, so
//    that no value in flight is moved between registers where they join;
//  * STG (forward and adjoint launches that stage their accumulation term,
//    every 5th step seven times in eight): the cell's pseudo-Hessian or
//    imaging term goes to the launch's staging slot (non-temporal), the
//    adjoint's forward field coming through one more ring of the wave's
//    column, loaded with the rows; the launches that add the staged terms to
//    the FP64 accumulator run fd_lanes_lds_kernel.  In the reconstruction
//    STG stores the new field's model cells into a.snap as well;
//  * PCP, the cache policy of the p0 loads and the output stores:
//    non-temporal for a launch that writes a third field, the default policy
//    for one that writes over p0.
// Rows past a run's end (the last group of a run) still load, publish and
// pass the barrier, computing nothing.
template <int D>
struct RingShape {
    static constexpr int U = ((D + 8) / 2) * 2;
    static_assert(U >= D + 7 && U % 2 == 0, "ring too small");
};

__device__ __forceinline__ frsrc col_rsrc(const float *base, long off_floats, long bytes_left)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base + off_floats), (short)0,
                                             (int)(unsigned)bytes_left, 0x00020000);
}

template <int MODE, int D, bool STG, int PCP>
__global__ __launch_bounds__(512) void fd_ring_kernel(FdArgs a)
{
    constexpr int WPB = 8;
    constexpr int NC = WPB + 8;  // 6 halo columns and 2 spare (see below)
    constexpr int U = RingShape<D>::U;
    constexpr bool UF = STG && MODE == FD_ADJ;  // the forward field rides with the rows
    __shared__ float row[2][NC][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int b = blockIdx.x;
    if (a.remap) {
        const int per = (int)((gridDim.x + 7) >> 3);
        b = (b & 7) * per + (b >> 3);
        if (b >= a.lz_nblk) return;  // the whole block
    }
    int zr, sg, xg;
    lanes_block(a, b, zr, sg, xg);
    const int x0 = a.ixs + xg * WPB;
    const int ix = x0 + w;
    const bool active = ix <= a.ixe;
    const int iz0 = a.izs + zr * a.lz_run;
    const int iz1 = min(a.ize, iz0 + a.lz_run - 1);
    const int nrow = iz1 - iz0 + 1;
    const unsigned s = (unsigned)(sg * 64 + lane);
    const unsigned vo = s * 4u;
    const long CS = a.cell_stride;
    const int ld = a.ld;
    const long fsz = (long)a.Nx * ld;
    const long fbytes = fsz * CS * 4;
    const int rs4 = (int)(CS * 4);
    const long colb = (long)min(ix, a.Nx - 1) * ld;
    const long cb = colb * CS * 4;
    // column resources: row r of the column at byte offset r * rs4 (P1z: row
    // r + 3; P2: row r - D, so that the store of row r uses the load offset
    // of row r + D)
    const frsrc P1c = col_rsrc(a.cur, colb * CS, fbytes - cb);
    const frsrc P1z = col_rsrc(a.cur, (colb + 3) * CS, fbytes - cb - 3L * rs4);
    const frsrc P0c = col_rsrc(a.prev, colb * CS, fbytes - cb);
    const frsrc P2c = col_rsrc(a.out, (colb - D) * CS, fbytes - cb + (long)D * rs4);
    // halo column of this wave: waves 0..5 carry the six halo columns (LDS
    // columns 0..2 and WPB+3..WPB+5); waves 6 and 7 run the same loads
    // through a resource of zero range (no memory request, value 0) into the
    // two spare LDS columns that nothing reads, so that every wave runs the
    // same instructions and no branch splits the count of loads in flight (a
    // join of two such paths makes the compiler wait for every outstanding
    // load)
    const int hw = w < 6 ? w : w - 6;
    const int hj = w < 3 ? w : WPB + w;
    const int hc = hw < 3 ? x0 - 3 + hw : x0 + WPB + (hw - 3);
    const long hcolb = (long)min(max(hc, 0), a.Nx - 1) * ld;
    const frsrc Hc = col_rsrc(a.cur, hcolb * CS, w < 6 ? fbytes - hcolb * CS * 4 : 0);
    auto ldc = [&](frsrc f, int boff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(f, vo, boff, 0));
    };
    auto ldc_stream = [&](frsrc f, int boff) {
        return __builtin_bit_cast(float,
                                  __builtin_amdgcn_raw_buffer_load_b32(f, vo, boff, CP_STREAM));
    };
    // p0 and the output: PCP is the non-temporal policy when they are
    // separate fields (each touched once by the launch, so the caches may let
    // them go first) and the default policy when the launch updates p0 in
    // place (two fields in all, which the Infinity Cache then keeps between
    // launches)
    auto ldc_p0 = [&](frsrc f, int boff) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(f, vo, boff, PCP));
    };

    // does the block meet a free-surface row, a source row or a receiver row?
    const bool recv = (MODE == FD_FWD && a.rec_on) || (MODE == FD_ADJ && a.inj_on);
    const bool special = (a.fs && iz0 <= a.pml + 3) ||
                         (MODE != FD_ADJ && a.src_cell && iz0 <= a.src_zhi && iz1 >= a.src_zlo) ||
                         (recv && iz0 <= a.rec_zhi && iz1 >= a.rec_zlo);
    int scell = -1, rr = -1;
    long slot = 0;
    float injv = 0.f;

    // rings: V holds p1 row iz0 - 3 + j in slot j mod U, P and H (p0 and the
    // halo column) row iz0 + j in slot j mod U, and with UF the forward field
    // of the imaging condition (Fu) too
    float V[U], P[U], H[U], Fu[UF ? U : 1];
    // staging: the column's place in the model ([nx][nz] cells x lanes), or
    // none outside it.  In the reconstruction STG stores the new field's model
    // cells a second time, into a.snap (a field of the same layout): the
    // forward field that an imaging step of the adjoint reads, while the
    // launch itself updates its work field in place
    const bool mcol = STG && active && ix >= a.pml && ix < a.pml + a.nx;
    // the forward field is read where the imaging term is formed: model
    // columns, in runs that meet the model's rows (elsewhere a zero-range
    // resource: no request, value 0, never used)
    const bool urun = UF && mcol && iz0 < a.pml + a.nz && iz1 >= a.pml;
    const frsrc Uc = col_rsrc(UF ? a.u : a.prev, colb * CS, urun ? fbytes - cb : 0);
    const frsrc Snc = col_rsrc(STG && MODE == FD_RECON ? a.snap : a.out, (colb - D) * CS,
                               STG && MODE == FD_RECON ? fbytes - cb + (long)D * rs4 : 0);
    const long abase = STG ? (long)(ix - a.pml) * a.nz * a.acc_cell + s : 0;
    float *const stg = STG ? a.accstage + (long)a.acc_slot * a.accstage_stride : nullptr;
    const cfloat8v *c8 = (const cfloat8v *)(a.coef + CO_N * fsz);
    cfloat *cf = (cfloat *)a.coef;
    const frsrc P2 = field_rsrc(a.out, fbytes);
    int rl = (iz0 + D) * rs4;  // byte offset of the row whose loads are issued next
    float8v kc[2];
    auto preload = [&]() {
#pragma unroll
        for (int j = 0; j < 6; ++j) V[j] = ldc(P1c, (iz0 - 3 + j) * rs4);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            V[(6 + d) % U] = ldc(P1z, (iz0 + d) * rs4);
            P[d % U] = ldc_p0(P0c, (iz0 + d) * rs4);
            H[d % U] = ldc(Hc, (iz0 + d) * rs4);
            if (UF) Fu[d % U] = ldc_stream(Uc, (iz0 + d) * rs4);
        }
        kc[0] = c8[colb + iz0];
    };

    // one group of U rows; CHECK: rows past the run's end compute nothing;
    // ACT: the wave's column is inside the stencil (a compile-time choice on
    // the plain path, so that its loop has no branch at all and the waits for
    // loads in flight are exact)
    auto group = [&](int g, auto check_t, auto special_t, auto act_t) {
        constexpr bool CHECK = decltype(check_t)::value;
        constexpr bool SPECIAL = decltype(special_t)::value;
        constexpr bool ACT = decltype(act_t)::value;
        const cfloat8v *cg = c8 + colb + iz0 + g * U;  // the group's coefficients
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int rel = g * U + u;
            const bool live = !CHECK || rel < nrow;
            row[u & 1][3 + w][lane] = V[(u + 3) % U];
            row[u & 1][hj][lane] = H[u % U];
            V[(u + 6 + D) % U] = ldc(P1z, rl);
            P[(u + D) % U] = ldc_p0(P0c, rl);
            H[(u + D) % U] = ldc(Hc, rl);
            if (UF) Fu[(u + D) % U] = ldc_stream(Uc, rl);
            __syncthreads();
            const int iz = iz0 + rel;
            if (SPECIAL && MODE == FD_FWD && iz == rr) a.stage[(long)a.it * a.nslot + slot] = V[(u + 3) % U];
            if (ACT && (!SPECIAL || active) && live) {
                const float xm1 = row[u & 1][2 + w][lane], xp1 = row[u & 1][4 + w][lane];
                const float xm2 = row[u & 1][1 + w][lane], xp2 = row[u & 1][5 + w][lane];
                const float xm3 = row[u & 1][0 + w][lane], xp3 = row[u & 1][6 + w][lane];
                // this row's coefficients were loaded a row ago; the next
                // row's load goes after the LDS reads (a scalar load in
                // flight makes every LDS wait a full wait)
                const float8v kv = kc[u & 1];
                __builtin_amdgcn_sched_barrier(0);
                kc[(u + 1) & 1] = cg[u + 1];
                const float q0 = V[u % U], q1 = V[(u + 1) % U], q2 = V[(u + 2) % U];
                const float q3 = V[(u + 3) % U], q4 = V[(u + 4) % U], q5 = V[(u + 5) % U];
                const float q6 = V[(u + 6) % U];
                if (SPECIAL) {
                    const float k[7] = {kv[0], kv[1], kv[2], kv[3], kv[4], kv[5], kv[6]};
                    const float sv = lanes_update<MODE, false, STG ? 3 : 1, true>(
                        a, cf, fsz, colb, ix, iz, s, scell, P2, P2, CS, q0, q1, q2, q3, q4, q5,
                        q6, xm1, xp1, xm2, xp2, xm3, xp3, P[u % U], k, 0.0,
                        UF ? Fu[u % U] : 0.f, MODE == FD_ADJ && iz == rr, injv, nullptr);
                    const int pz = iz - a.pml;
                    if (STG && MODE == FD_RECON && mcol && pz >= 0 && pz < a.nz)
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, sv), Snc,
                                                              vo, rl, CP_STREAM);
                } else {
                    const float v = fd_update<false>(q3, P[u % U], q2, q4, q1, q5, q0, q6, xm1,
                                                     xp1, xm2, xp2, xm3, xp3, kv[0], kv[1], kv[2],
                                                     kv[3], kv[4], kv[5], a.invdx, a.invdz,
                                                     a.invdxdx, a.invdzdz);
                    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), P2c, vo, rl,
                                                          PCP);
                    const int pz = iz - a.pml;
                    if (STG && MODE == FD_RECON && mcol && pz >= 0 && pz < a.nz) {
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), Snc, vo,
                                                              rl, CP_STREAM);
                    } else if (STG && MODE != FD_RECON && mcol && pz >= 0 && pz < a.nz) {
                        // propagation.c:590 (forward) and :711 (adjoint), as
                        // lanes_update forms them
                        const float d2 = (v - 2.0f * q3 + P[u % U]) * a.invdtdt;
                        const float term = MODE == FD_FWD ? d2 * d2 : d2 * Fu[UF ? u % U : 0];
                        __builtin_nontemporal_store(term, stg + abase + (long)pz * a.acc_cell);
                    }
                }
            }
            rl += rs4;
        }
    };
    const int nfull = nrow / U;
    using T = std::true_type;
    using F = std::false_type;
    if (special) {
        // the column's receiver row and stage column (scalar loads), then the
        // adjoint source to inject there (adjoint) or the source cells of the
        // block's lanes, ahead of the rings; the per-lane loads go through
        // resources of zero range where there is nothing to read (value 0: no
        // cell of the stencil has index 0), so that no load sits in a branch
        // whose join would wait for it
        int r = -1, rc = 0;
        if (recv && active) {
            r = ((cint *)a.rrow)[ix];
            rc = ((cint *)a.rcol)[ix];
        }
        const bool hasr = r >= iz0 && r <= iz1;
        rr = hasr ? r : -1;
        slot = (long)rc * CS + s;
        if (MODE == FD_ADJ) {
            const frsrc St = hasr ? col_rsrc(a.stage, (long)(a.it - 1) * a.nslot, a.nslot * 4)
                                  : col_rsrc(a.out, 0, 0);
            injv = __builtin_bit_cast(float,
                                      __builtin_amdgcn_raw_buffer_load_b32(St, vo, rc * rs4, 0));
        } else {
            const frsrc Sc = a.src_cell ? __builtin_amdgcn_make_buffer_rsrc(
                                              const_cast<int *>(a.src_cell), (short)0,
                                              a.lz_ngrp * 256, 0x00020000)
                                        : col_rsrc(a.out, 0, 0);
            scell = (int)__builtin_amdgcn_raw_buffer_load_b32(Sc, vo, 0, 0);
        }
        preload();
#pragma unroll 1
        for (int g = 0; g < nfull; ++g) group(g, F(), T(), T());
        if (nfull * U < nrow) group(nfull, T(), T(), T());
    } else {
        preload();
        if (active) {
#pragma unroll 1
            for (int g = 0; g < nfull; ++g) group(g, F(), F(), T());
            if (nfull * U < nrow) group(nfull, T(), F(), T());
        } else {
#pragma unroll 1
            for (int g = 0; g < nfull; ++g) group(g, F(), F(), F());
            if (nfull * U < nrow) group(nfull, T(), F(), F());
        }
    }
}

template <int MODE, bool STG, int PCP>
static void launch_ring_t(const FdArgs &a0, FdTile tl, hipStream_t st)
{
    FdArgs a = a0;
    const int H = a.ize - a.izs + 1;
    a.lz_run = tl.zrun;
    a.lz_nrun = (H + tl.zrun - 1) / tl.zrun;
    a.lz_ngrp = (a.nshot + 63) / 64;
    const int nxg = (a.ixe - a.ixs + 1 + 7) / 8;
    const long nb = (long)nxg * a.lz_nrun * a.lz_ngrp;
    a.lz_nblk = (int)nb;
    a.lz_nxg = nxg;
    a.lz_order = tl.zfast;
    const dim3 grid((unsigned)(a.remap ? (nb + 7) / 8 * 8 : nb)), block(512);
    if (tl.pd >= 5)
        fd_go(fd_ring_kernel<MODE, 5, STG, PCP>, grid, block, 0, st, a);
    else if (tl.pd == 4)
        fd_go(fd_ring_kernel<MODE, 4, STG, PCP>, grid, block, 0, st, a);
    else if (tl.pd == 3)
        fd_go(fd_ring_kernel<MODE, 3, STG, PCP>, grid, block, 0, st, a);
    else
        fd_go(fd_ring_kernel<MODE, 2, STG, PCP>, grid, block, 0, st, a);
}

int ring_unroll(int pd)
{
    const int d = pd >= 5 ? 5 : (pd < 2 ? 2 : pd);
    return ((d + 8) / 2) * 2;
}

void launch_fd(int mode, const FdArgs &a0, FdTile tl, hipStream_t st)
{
    FdArgs a = a0;
    a.remap = tl.remap;
    if (a.snap && (mode != FD_RECON || tl.kind != 4)) {
        fprintf(stderr, "<error>: internal: a second store of the new field is a reconstruction "
                        "launch of the register-ring stencil only\n");
        exit(1);
    }
    const bool stage = mode != FD_RECON && a.acc_on && a.acc_stage && a.accstage;
    if (tl.kind == 4 && mode != FD_RECON && a.acc_on && !stage) {
        // the ring stencil stages accumulation terms but does not add them
        // to the accumulator: the LDS stencil of the same block shape runs
        // these launches
        tl.kind = 3;
        tl.wpb = 8;
        tl.pd = min(tl.pd, 4);
    }
    if (tl.kind == 4) {
        // a launch that writes over p0 (out == prev) keeps both fields at
        // the default cache policy
        if (a.out == a.prev) {
            if (mode == FD_FWD)
                stage ? launch_ring_t<FD_FWD, true, 0>(a, tl, st)
                      : launch_ring_t<FD_FWD, false, 0>(a, tl, st);
            else if (mode == FD_RECON)
                a.snap ? launch_ring_t<FD_RECON, true, 0>(a, tl, st)
                       : launch_ring_t<FD_RECON, false, 0>(a, tl, st);
            else
                stage ? launch_ring_t<FD_ADJ, true, 0>(a, tl, st)
                      : launch_ring_t<FD_ADJ, false, 0>(a, tl, st);
        } else {
            if (mode == FD_FWD)
                stage ? launch_ring_t<FD_FWD, true, CP_STREAM>(a, tl, st)
                      : launch_ring_t<FD_FWD, false, CP_STREAM>(a, tl, st);
            else if (mode == FD_RECON)
                a.snap ? launch_ring_t<FD_RECON, true, CP_STREAM>(a, tl, st)
                       : launch_ring_t<FD_RECON, false, CP_STREAM>(a, tl, st);
            else
                stage ? launch_ring_t<FD_ADJ, true, CP_STREAM>(a, tl, st)
                      : launch_ring_t<FD_ADJ, false, CP_STREAM>(a, tl, st);
        }
        return;
    }
    if (tl.kind == 3) {
        if (mode == FD_FWD)
            launch_lanes_lds_t<FD_FWD>(a, tl, st);
        else if (mode == FD_RECON)
            launch_lanes_lds_t<FD_RECON>(a, tl, st);
        else
            launch_lanes_lds_t<FD_ADJ>(a, tl, st);
        return;
    }
    if (tl.kind == 2) {
        if (mode == FD_FWD)
            launch_lanes_t<FD_FWD>(a, tl, st);
        else if (mode == FD_RECON)
            launch_lanes_t<FD_RECON>(a, tl, st);
        else
            launch_lanes_t<FD_ADJ>(a, tl, st);
        return;
    }
    if (tl.kind == 1) {
        if (mode == FD_FWD)
            launch_march_mode<FD_FWD>(a, tl, st);
        else if (mode == FD_RECON)
            launch_march_mode<FD_RECON>(a, tl, st);
        else
            launch_march_mode<FD_ADJ>(a, tl, st);
        return;
    }
    if (mode == FD_FWD)
        launch_fd_mode<FD_FWD>(a, tl, st);
    else if (mode == FD_RECON)
        launch_fd_mode<FD_RECON>(a, tl, st);
    else
        launch_fd_mode<FD_ADJ>(a, tl, st);
}

int fd_tile_rows_padded(int tz, int izs, int ize) { return (ize - izs + 1 + tz - 1) / tz * tz; }

// ---------------------------------------------------------------------------
// Set-up kernels (propagation.c:103-171, 192-354, 508-532; common.c:110-176,416-479)

static inline dim3 grid2(const Geom &g, int ldz) { return dim3((ldz + 63) / 64, g.Nx); }

__global__ void pad_kernel(const float *win, int nx, int nz, int pml, float *out, int Nz, int ld)
{
    const int iz = blockIdx.x * 64 + threadIdx.x, ix = blockIdx.y;
    if (iz >= ld) return;
    float v = 0.f;
    if (iz < Nz) {
        int sx = ix - pml, sz = iz - pml;
        sx = sx < 0 ? 0 : (sx > nx - 1 ? nx - 1 : sx);
        sz = sz < 0 ? 0 : (sz > nz - 1 ? nz - 1 : sz);
        v = win[(long)sx * nz + sz];
    }
    out[(long)ix * ld + iz] = v;
}

void launch_pad(const float *win, int nx, int nz, int pml, float *out, const Geom &g,
                hipStream_t st)
{
    pad_kernel<<<grid2(g, g.ld), 64, 0, st>>>(win, nx, nz, pml, out, g.Nz, g.ld);
}

__global__ void meal_kernel(const float *vp, const float *vs, float *D, float *Beta, int NX,
                            int NZ, int ld, int nl, const float *wt, float lnR, float pd,
                            float sqrt3, float fpeak, float dx, float dz)
{
    const int iz = blockIdx.x * 64 + threadIdx.x, ix = blockIdx.y;
    if (iz >= ld) return;
    const long i = (long)ix * ld + iz;
    if (iz >= NZ) {
        D[i] = 0.f;
        Beta[i] = 1.f;
        return;
    }
    const float v = vp[i];
    float Vs = v / sqrt3;
    if (vs && vs[i] != 0.0f) Vs = vs[i];
    const float thx = (float)nl * dx, thz = (float)nl * dz;
    const float PPWfd = 10.0f;
    float Dx = 0.f, Dz = 0.f, Bx = 1.f, Bz = 1.f;
    // top then bottom (propagation.c:285-309), left then right (:311-335)
    if (iz < nl) {
        const float w = wt[nl - iz];
        const float d0 = -1.0f * (1.0f + pd) * v * lnR / (2.0f * thz);
        const float b = 2.0f * (Vs / (dz * fpeak)) / PPWfd;
        Dz = d0 * w;
        Bz = 1.0f + (b - 1.0f) * w;
    }
    if (iz >= NZ - nl) {
        const float w = wt[iz - (NZ - nl - 1)];
        const float d0 = -1.0f * (1.0f + pd) * v * lnR / (2.0f * thz);
        const float b = 2.0f * (Vs / (dz * fpeak)) / PPWfd;
        Dz = d0 * w;
        Bz = 1.0f + (b - 1.0f) * w;
    }
    if (ix < nl) {
        const float w = wt[nl - ix];
        const float d0 = -1.0f * (1.0f + pd) * v * lnR / (2.0f * thx);
        const float b = 2.0f * (Vs / (dx * fpeak)) / PPWfd;
        Dx = d0 * w;
        Bx = 1.0f + (b - 1.0f) * w;
    }
    if (ix >= NX - nl) {
        const float w = wt[ix - (NX - nl - 1)];
        const float d0 = -1.0f * (1.0f + pd) * v * lnR / (2.0f * thx);
        const float b = 2.0f * (Vs / (dx * fpeak)) / PPWfd;
        Dx = d0 * w;
        Bx = 1.0f + (b - 1.0f) * w;
    }
    D[i] = sqrtf(Dx * Dx + Dz * Dz);
    Beta[i] = sqrtf((Bx - 1.0f) * (Bx - 1.0f) + (Bz - 1.0f) * (Bz - 1.0f)) + 1.0f;
}

void launch_meal(const float *vp, const float *vs, float *D, float *Beta, const Geom &g, int nlayer,
                 const float *w_dev, float lnR, float pd, float sqrt3, float fpeak, hipStream_t st)
{
    meal_kernel<<<grid2(g, g.ld), 64, 0, st>>>(vp, vs, D, Beta, g.Nx, g.Nz, g.ld, nlayer, w_dev,
                                               lnR, pd, sqrt3, fpeak, g.dx, g.dz);
}

__global__ void coef_kernel(float *vp, float *rho, float *D, const float *Beta, float *coef,
                            int Nz, int ld, long fsz, float dt)
{
    const int iz = blockIdx.x * 64 + threadIdx.x, ix = blockIdx.y;
    if (iz >= ld) return;
    const long i = (long)ix * ld + iz;
    if (iz >= Nz) {
        for (int k = 0; k < CO_N; ++k) coef[k * fsz + i] = 0.f;
        return;
    }
    const float b = Beta[i];
    const float d = D[i] / b;
    const float v = vp[i] / b;
    const float r = rho[i] * b;
    D[i] = d;
    vp[i] = v;
    rho[i] = r;
    // propagation.c:525-527: the products are FP32, the rest FP64
    const float dt2D2 = dt * dt * d * d;
    const float dtD = dt * d;
    const float dt2 = dt * dt;
    const float c3 = (float)((double)dt2 / (1.0 + (double)dtD));
    const float K = r * v * v;
    coef[CO_C1 * fsz + i] = (float)((2.0 - (double)dt2D2) / (1.0 + (double)dtD));
    coef[CO_C2 * fsz + i] = (float)(-1.0 * (1.0 - (double)dtD) / (1.0 + (double)dtD));
    coef[CO_A * fsz + i] = c3 * K;
    coef[CO_K * fsz + i] = K;
    coef[CO_IR * fsz + i] = 1.0f / r;
}

void launch_coefs(float *vp, float *rho, float *D, const float *Beta, float *coef, const Geom &g,
                  hipStream_t st)
{
    coef_kernel<<<grid2(g, g.ld), 64, 0, st>>>(vp, rho, D, Beta, coef, g.Nz, g.ld, g.fsz, g.dt);
}

__global__ void grad_invrho_kernel(float *coef, int Nx, int Nz, int ld, long fsz, float dx,
                                   float dz)
{
    const int iz = blockIdx.x * 64 + threadIdx.x, ix = blockIdx.y;
    if (iz >= Nz) return;
    const float *f = coef + CO_IR * fsz;
    const long i = (long)ix * ld + iz;
    float gx, gz;
    if (ix >= 6 && ix < Nx - 6 && iz >= 6 && iz < Nz - 6) {
        const float C1 = 0.91067892f, C2 = -0.34187892f, C3 = 0.13833962f;
        const float C4 = -0.04880710f, C5 = 0.01302148f, C6 = -0.00199047f;
        const float invdx = 1.0f / dx, invdz = 1.0f / dz;
        gx = (C1 * (f[i + ld] - f[i - ld]) + C2 * (f[i + 2 * ld] - f[i - 2 * ld]) +
              C3 * (f[i + 3 * ld] - f[i - 3 * ld]) + C4 * (f[i + 4 * ld] - f[i - 4 * ld]) +
              C5 * (f[i + 5 * ld] - f[i - 5 * ld]) + C6 * (f[i + 6 * ld] - f[i - 6 * ld])) *
             invdx;
        gz = (C1 * (f[i + 1] - f[i - 1]) + C2 * (f[i + 2] - f[i - 2]) + C3 * (f[i + 3] - f[i - 3]) +
              C4 * (f[i + 4] - f[i - 4]) + C5 * (f[i + 5] - f[i - 5]) + C6 * (f[i + 6] - f[i - 6])) *
             invdz;
    } else {
        if (ix == 0)
            gx = (f[i + ld] - f[i]) / dx;
        else if (ix == Nx - 1)
            gx = (f[i] - f[i - ld]) / dx;
        else
            gx = (f[i + ld] - f[i - ld]) / (2.0f * dx);
        if (iz == 0)
            gz = (f[i + 1] - f[i]) / dz;
        else if (iz == Nz - 1)
            gz = (f[i] - f[i - 1]) / dz;
        else
            gz = (f[i + 1] - f[i - 1]) / (2.0f * dz);
    }
    coef[CO_IRX * fsz + i] = gx;
    coef[CO_IRZ * fsz + i] = gz;
}

void launch_grad_invrho(float *coef, const Geom &g, hipStream_t st)
{
    grad_invrho_kernel<<<grid2(g, g.Nz), 64, 0, st>>>(coef, g.Nx, g.Nz, g.ld, g.fsz, g.dx, g.dz);
}

// The coefficients of a cell side by side, [fsz][8] after the CO_N arrays:
// c1, c2, c3*K, rho^-1_x, rho^-1_z, rho^-1, K, 0 -- one 32-byte scalar load per
// cell for the shots-in-lanes stencils instead of seven.
__global__ void coef_pack_kernel(float *coef, long fsz)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fsz) return;
    float *o = coef + CO_N * fsz + i * 8;
#pragma unroll
    for (int k = 0; k < CO_N; ++k) o[k] = coef[k * fsz + i];
    o[7] = 0.f;
}

void launch_coef_pack(float *coef, const Geom &g, hipStream_t st)
{
    coef_pack_kernel<<<(unsigned)((g.fsz + 255) / 256), 256, 0, st>>>(coef, g.fsz);
}

// ---------------------------------------------------------------------------
// Receivers (propagation.c:567-576 record p1; :676-682 adjoint injection)

__global__ void record_kernel(RcvTable t, const float *f0, long stride, long cstride, float *rec,
                              int it)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.n) return;
    const int sr = t.sr[e];
    if (it % sr) return;
    rec[t.trace_off[e] + it / sr] = f0[(long)t.shot[e] * stride + (long)t.cell[e] * cstride];
}

void launch_record(const RcvTable &t, const float *field0, long shot_stride, long cell_stride,
                   float *rec, int it, hipStream_t st)
{
    if (t.n <= 0) return;
    record_kernel<<<(t.n + 255) / 256, 256, 0, st>>>(t, field0, shot_stride, cell_stride, rec, it);
}

// Time-major staging of receiver values for the lanes stencil: 64 x 64 tiles
// (slots x time steps) transposed through LDS, so that both the trace-major
// side (records, adjoint sources: consecutive samples of one trace) and the
// time-major stage (consecutive slots of one step) are accessed in rows.
__global__ __launch_bounds__(256) void stage_gather_kernel(const long *slot_off, const float *src,
                                                           float *stage, long nslot, int nt)
{
    __shared__ float tile[64][65];
    const long s0 = (long)blockIdx.x * 64;
    const int t0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // read: one trace per row, 64 consecutive samples
    for (int j = ty; j < 64; j += 4) {
        const long sl = s0 + j;
        float v = 0.f;
        if (sl < nslot && t0 + tx < nt) {
            const long off = slot_off[sl];
            if (off >= 0) v = src[off + t0 + tx];
        }
        tile[j][tx] = v;
    }
    __syncthreads();
    // write: one step per row, 64 consecutive slots
    for (int j = ty; j < 64; j += 4) {
        const int t = t0 + j;
        if (t < nt && s0 + tx < nslot) stage[(long)t * nslot + s0 + tx] = tile[tx][j];
    }
}

void launch_stage_gather(const long *slot_off, const float *src, float *stage, long nslot, int nt,
                         hipStream_t st)
{
    if (nslot <= 0 || nt <= 0) return;
    const dim3 grid((unsigned)((nslot + 63) / 64), (unsigned)((nt + 63) / 64));
    stage_gather_kernel<<<grid, 256, 0, st>>>(slot_off, src, stage, nslot, nt);
}

__global__ __launch_bounds__(256) void stage_scatter_kernel(const long *slot_off,
                                                            const int *slot_sr, const float *stage,
                                                            float *rec, long nslot, int nt)
{
    __shared__ float tile[64][65];
    const long s0 = (long)blockIdx.x * 64;
    const int t0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int j = ty; j < 64; j += 4) {
        const int t = t0 + j;
        tile[j][tx] = (t < nt && s0 + tx < nslot) ? stage[(long)t * nslot + s0 + tx] : 0.f;
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const long sl = s0 + j;
        const int t = t0 + tx;
        if (sl >= nslot || t >= nt) continue;
        const long off = slot_off[sl];
        if (off < 0) continue;
        const int sr = slot_sr[sl];
        if (t % sr == 0) rec[off + t / sr] = tile[tx][j];
    }
}

void launch_stage_scatter(const long *slot_off, const int *slot_sr, const float *stage, float *rec,
                          long nslot, int nt, hipStream_t st)
{
    if (nslot <= 0 || nt <= 0) return;
    const dim3 grid((unsigned)((nslot + 63) / 64), (unsigned)((nt + 63) / 64));
    stage_scatter_kernel<<<grid, 256, 0, st>>>(slot_off, slot_sr, stage, rec, nslot, nt);
}

__global__ void inject_kernel(InjTable t, float *f0, long stride, long cstride, const float *K,
                              const float *adj, int it, float dt, float invdx, float invdz)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= t.n) return;
    const int cell = t.cell[u];
    float *p = f0 + (long)t.shot[u] * stride + (long)cell * cstride;
    float v = *p;
    const float k = K[cell];
    for (int e = t.beg[u]; e < t.beg[u + 1]; ++e)
        v -= adj[t.trace_off[e] + it] * k * dt * dt * invdx * invdz;
    *p = v;
}

void launch_inject(const InjTable &t, float *field0, long shot_stride, long cell_stride,
                   const float *K, const float *adj, int it, float dt, float invdx, float invdz,
                   hipStream_t st)
{
    if (t.n <= 0) return;
    inject_kernel<<<(t.n + 255) / 256, 256, 0, st>>>(t, field0, shot_stride, cell_stride, K, adj,
                                                      it, dt, invdx, invdz);
}

// ---------------------------------------------------------------------------
// Adjoint source (optimization.c:335-345, 417-418; taper_traces common.c:796-895)

__global__ void adjsrc_kernel(const float *pre, const float *obs, float *out, int ntr, int ns,
                              const float *cT, int Lt, const float *cX, int Lx)
{
    const long n = (long)ntr * ns;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int itr = (int)(e / ns), it = (int)(e - (long)itr * ns);
    float r = pre[e] - obs[e];
    if (Lt >= 0) {  // Lt < 0: no time taper
        if (it <= Lt) r *= cT[Lt - it];
        if (it >= ns - Lt - 1) r *= cT[it - (ns - Lt - 1)];
    }
    if (Lx >= 0) {
        if (itr <= Lx) r *= cX[Lx - itr];
        if (itr >= ntr - Lx - 1) r *= cX[itr - (ntr - Lx - 1)];
    }
    out[e] = r;
}

void launch_adjsrc(const float *pre, const float *obs, float *out, int ntr, int ns,
                   const float *coefT, int Lt, const float *coefX, int Lx, hipStream_t st)
{
    const long n = (long)ntr * ns;
    if (n <= 0) return;
    adjsrc_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(pre, obs, out, ntr, ns, coefT, Lt,
                                                                coefX, Lx);
}

// In-place cosine tapers of [ntr][ns] traces (taper_traces common.c:796-895):
// time taper first, then the offset taper, each skipped when its length is < 0.
__global__ void taper_kernel(float *x, int ntr, int ns, const float *cT, int Lt, const float *cX,
                             int Lx)
{
    const long n = (long)ntr * ns;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int itr = (int)(e / ns), it = (int)(e - (long)itr * ns);
    float r = x[e];
    if (Lt >= 0) {
        if (it <= Lt) r *= cT[Lt - it];
        if (it >= ns - Lt - 1) r *= cT[it - (ns - Lt - 1)];
    }
    if (Lx >= 0) {
        if (itr <= Lx) r *= cX[Lx - itr];
        if (itr >= ntr - Lx - 1) r *= cX[itr - (ntr - Lx - 1)];
    }
    x[e] = r;
}

void launch_taper(float *x, int ntr, int ns, const float *coefT, int Lt, const float *coefX,
                  int Lx, hipStream_t st)
{
    const long n = (long)ntr * ns;
    if (n <= 0 || (Lt < 0 && Lx < 0)) return;
    taper_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x, ntr, ns, coefT, Lt, coefX, Lx);
}

// Spectrum [batch][nb] (interleaved complex) times a real window of nb bins.
__global__ void spec_window_kernel(float2 *spec, long total, int nb, const float *win)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const float w = win[e % nb];
    float2 v = spec[e];
    v.x *= w;
    v.y *= w;
    spec[e] = v;
}

void launch_spec_window(float2 *spec, long total, int nb, const float *win, hipStream_t st)
{
    if (total > 0)
        spec_window_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(spec, total, nb, win);
}

// x[e] = x[e] / d (the reference divides the unnormalised inverse transform).
__global__ void div_kernel(float *x, long n, float d)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) x[e] = x[e] / d;
}

void launch_div(float *x, long n, float d, hipStream_t st)
{
    if (n > 0) div_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x, n, d);
}

// out[c][r] = in[r][c] for an rows x cols matrix, through 32 x 32 LDS tiles.
__global__ void transpose_kernel(const float *in, float *out, int rows, int cols)
{
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) t[k][threadIdx.x] = in[(long)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(long)c * rows + r] = t[threadIdx.x][k];
    }
}

void launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st)
{
    if (rows <= 0 || cols <= 0) return;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32), block(32, 8);
    transpose_kernel<<<grid, block, 0, st>>>(in, out, rows, cols);
}

// interpolate_traces (common.c:482-581) pieces: zero-stuffing into a complex
// buffer [ntr][M], the rectangular window (gain r on bins k < M/(2r) and
// k > M - M/(2r)), and the real part divided by M into [ntr][nt_out] with a
// zero tail.
__global__ void zero_stuff_kernel(const float *in, int ntr, int ns, int r, float2 *buf, int M)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)ntr * M) return;
    const int itr = (int)(e / M), j = (int)(e - (long)itr * M);
    float2 v = make_float2(0.0f, 0.0f);
    if (j % r == 0) v.x = in[(long)itr * ns + j / r];
    buf[e] = v;
}

void launch_zero_stuff(const float *in, int ntr, int ns, int r, float2 *buf, int M, hipStream_t st)
{
    const long n = (long)ntr * M;
    zero_stuff_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, ntr, ns, r, buf, M);
}

__global__ void interp_window_kernel(float2 *buf, int ntr, int M, int r)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)ntr * M) return;
    const int k = (int)(e % M);
    const int keep = M / (2 * r);
    const float w = (k < keep || k >= M - keep + 1) ? (float)r : 0.0f;
    float2 v = buf[e];
    v.x *= w;
    v.y *= w;
    buf[e] = v;
}

void launch_interp_window(float2 *buf, int ntr, int M, int r, hipStream_t st)
{
    const long n = (long)ntr * M;
    interp_window_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(buf, ntr, M, r);
}

__global__ void interp_out_kernel(const float2 *buf, int ntr, int M, float *out, int nt_out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)ntr * nt_out) return;
    const int itr = (int)(e / nt_out), j = (int)(e - (long)itr * nt_out);
    out[e] = j < M ? buf[(long)itr * M + j].x / (float)M : 0.0f;
}

void launch_interp_out(const float2 *buf, int ntr, int M, float *out, int nt_out, hipStream_t st)
{
    const long n = (long)ntr * nt_out;
    interp_out_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(buf, ntr, M, out, nt_out);
}

// ---------------------------------------------------------------------------
// Gradient finalisation and preconditioning (propagation.c:610-623, 732-743,
// 953-993; optimization.c:13-97)

__global__ void finalize_kernel(const double *acc, long acc_stride, long acc_cell, float *out,
                                long out_stride, const float *vpw, const float *rhow, int n,
                                double scale, int squared)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    if (i >= n) return;
    const float rv = rhow[i] * vpw[i];
    const double den = squared ? (double)(rv * rv) : (double)rv;
    out[(long)s * out_stride + i] =
        (float)(scale * acc[(long)s * acc_stride + (long)i * acc_cell] / den);
}

void launch_finalize(const double *acc, long acc_stride, long acc_cell, float *out,
                     long out_stride, int nshot, const float *vpw, const float *rhow, int n,
                     double scale, int squared, hipStream_t st)
{
    dim3 grid((n + 255) / 256, nshot);
    finalize_kernel<<<grid, 256, 0, st>>>(acc, acc_stride, acc_cell, out, out_stride, vpw, rhow,
                                          n, scale, squared);
}

// Deterministic reductions: a fixed grid of RB blocks, each thread strides
// through the input in a fixed order, then a fixed-shape tree in LDS, then one
// block folds the RB partials.  The result depends only on n.
constexpr int RB = 512;
constexpr int RT = 256;
size_t reduce_scratch_doubles() { return (size_t)4 * RB + 64; }

template <int OP>
__device__ inline void term(const float *a, const float *b, long i, double &s0, double &s1)
{
    if (OP == 0) {  // sum of float squares
        const float v = a[i];
        s0 += (double)(v * v);
    } else if (OP == 1) {  // sum
        s0 += (double)a[i];
    } else if (OP == 2) {  // step length: r0*(r0-r1), (r0-r1)^2
        const float r0 = a[i];
        const float d = r0 - b[i];
        s0 += (double)(r0 * d);
        s1 += (double)(d * d);
    } else {  // sum of double squares
        s0 += (double)a[i] * (double)a[i];
    }
}

__device__ inline void block_sum2(double &s0, double &s1, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = s0;
    sh[RT + t] = s1;
    __syncthreads();
    for (int k = RT / 2; k > 0; k >>= 1) {
        if (t < k) {
            sh[t] += sh[t + k];
            sh[RT + t] += sh[RT + t + k];
        }
        __syncthreads();
    }
    s0 = sh[0];
    s1 = sh[RT];
}

template <int OP>
__global__ __launch_bounds__(RT) void reduce_stage1(const float *a, const float *b, long n,
                                                    double *part)
{
    __shared__ double sh[2 * RT];
    double s0 = 0.0, s1 = 0.0;
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)RB * RT)
        term<OP>(a, b, i, s0, s1);
    block_sum2(s0, s1, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0;
        part[RB + blockIdx.x] = s1;
    }
}

__global__ __launch_bounds__(RT) void reduce_stage2(const double *part, double *res)
{
    __shared__ double sh[2 * RT];
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < RB; i += RT) {
        s0 += part[i];
        s1 += part[RB + i];
    }
    block_sum2(s0, s1, sh);
    if (threadIdx.x == 0) {
        res[0] = s0;
        res[1] = s1;
    }
}

template <int OP>
static void reduce(const float *a, const float *b, long n, double *scratch, double *res,
                   hipStream_t st)
{
    reduce_stage1<OP><<<RB, RT, 0, st>>>(a, b, n, scratch);
    reduce_stage2<<<1, RT, 0, st>>>(scratch, res);
}

void launch_sum_sq(const float *x, long n, double *scratch, double *res, hipStream_t st)
{
    reduce<0>(x, nullptr, n, scratch, res, st);
}
void launch_sum(const float *x, long n, double *scratch, double *res, hipStream_t st)
{
    reduce<1>(x, nullptr, n, scratch, res, st);
}
void launch_sum_step(const float *a, const float *b, long n, double *scratch, double *res,
                     hipStream_t st)
{
    reduce<2>(a, b, n, scratch, res, st);
}

// per-shot sums: grid (RB, nshot) -> part[s][RB] -> sums[s]
__global__ __launch_bounds__(RT) void shot_sum_stage1(const float *p, long stride, int n,
                                                      double *part)
{
    __shared__ double sh[2 * RT];
    const int s = blockIdx.y;
    double s0 = 0.0, s1 = 0.0;
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)RB * RT)
        s0 += (double)p[(long)s * stride + i];
    block_sum2(s0, s1, sh);
    if (threadIdx.x == 0) part[(long)s * RB + blockIdx.x] = s0;
}

__global__ __launch_bounds__(RT) void shot_sum_stage2(const double *part, double *sums)
{
    __shared__ double sh[2 * RT];
    const int s = blockIdx.x;
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < RB; i += RT) s0 += part[(long)s * RB + i];
    block_sum2(s0, s1, sh);
    if (threadIdx.x == 0) sums[s] = s0;
}

__global__ void shot_precond_apply(float *g, float *p, long stride, int n, const double *sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    if (i >= n) return;
    const float avg = (float)(sums[s] / n);
    const long k = (long)s * stride + i;
    const float f = 1.0f + p[k] / (7.0f * avg);
    g[k] = g[k] / f;
    p[k] = p[k] / f;
}

void launch_shot_precondition(float *g, float *p, long stride, int nshot, int n, double *scratch,
                              hipStream_t st)
{
    // scratch holds nshot*RB partials followed by nshot sums
    double *part = scratch, *sums = scratch + (long)nshot * RB;
    shot_sum_stage1<<<dim3(RB, nshot), RT, 0, st>>>(p, stride, n, part);
    shot_sum_stage2<<<nshot, RT, 0, st>>>(part, sums);
    shot_precond_apply<<<dim3((n + 255) / 256, nshot), 256, 0, st>>>(g, p, stride, n, sums);
}

__global__ void accumulate_kernel(const float *v, long stride, int nshot, const int *ix0,
                                  const int *nxx, int nz, int nx, double *sum)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nx * nz) return;
    const int ix = (int)(e / nz), iz = (int)(e - (long)ix * nz);
    double acc = sum[e];
    for (int s = 0; s < nshot; ++s) {
        const int lx = ix - ix0[s];
        if (lx >= 0 && lx < nxx[s]) acc += v[(long)s * stride + (long)lx * nz + iz];
    }
    sum[e] = acc;
}

void launch_accumulate(const float *v, long stride, int nshot, const int *ix0_dev,
                       const int *nxx_dev, int nz, int nx, double *sum, hipStream_t st)
{
    const long n = (long)nx * nz;
    accumulate_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(v, stride, nshot, ix0_dev,
                                                                   nxx_dev, nz, nx, sum);
}

__global__ void to_float_kernel(const double *in, float *out, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = (float)in[e];
}

void launch_to_float(const double *in, float *out, long n, hipStream_t st)
{
    to_float_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, out, n);
}

__global__ void global_precond_apply(float *g, const float *p, long n, float alpha,
                                     const double *res)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float rms = (float)sqrt(res[0] / (double)n);
    g[e] = g[e] / (p[e] + alpha * rms);
}

void launch_global_precondition(float *g, const float *p, long n, float alpha, double *scratch,
                                hipStream_t st)
{
    if (alpha <= 0.0f) alpha = 0.005f;
    double *res = scratch + 2 * RB;
    reduce<3>(p, nullptr, n, scratch, res, st);
    global_precond_apply<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, p, n, alpha, res);
}

// min / max (exact in any order)
__global__ __launch_bounds__(RT) void minmax_stage1(const float *x, long n, float *part)
{
    __shared__ float mn[RT], mx[RT];
    float a = 0.0f, b = 0.0f;  // optimization.c:203-204 start both at 0
    for (long i = (long)blockIdx.x * RT + threadIdx.x; i < n; i += (long)RB * RT) {
        a = fminf(a, x[i]);
        b = fmaxf(b, x[i]);
    }
    mn[threadIdx.x] = a;
    mx[threadIdx.x] = b;
    __syncthreads();
    for (int k = RT / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            mn[threadIdx.x] = fminf(mn[threadIdx.x], mn[threadIdx.x + k]);
            mx[threadIdx.x] = fmaxf(mx[threadIdx.x], mx[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = mn[0];
        part[RB + blockIdx.x] = mx[0];
    }
}

__global__ void minmax_stage2(const float *part, float *res)
{
    if (threadIdx.x) return;
    float a = 0.0f, b = 0.0f;
    for (int i = 0; i < RB; ++i) {
        a = fminf(a, part[i]);
        b = fmaxf(b, part[RB + i]);
    }
    res[0] = a;
    res[1] = b;
}

void launch_minmax(const float *x, long n, float *scratch, float *res, hipStream_t st)
{
    minmax_stage1<<<RB, RT, 0, st>>>(x, n, scratch);
    minmax_stage2<<<1, 64, 0, st>>>(scratch, res);
}

// histogram of optimization.c:229-233 (integer counts: order-free)
__global__ void histogram_kernel(const float *g, long n, float mn, float interval, int *hist)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double x = (double)((g[e] - mn) / interval) + 0.5;
    int pos = (x >= 0.0) ? (int)x : 0;  // NaN (all-zero gradient) lands in bin 0
    if (pos > 1000) pos = 1000;
    atomicAdd(&hist[pos], 1);
}

void launch_histogram(const float *g, long n, float mn, float interval, int *hist, hipStream_t st)
{
    histogram_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, n, mn, interval, hist);
}

__global__ void scale_kernel(const float *in, float *out, long n, float s)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = -s * in[e];
}

void launch_scale(const float *in, float *out, long n, float s, hipStream_t st)
{
    scale_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, out, n, s);
}

// ---------------------------------------------------------------------------
__global__ void freesurface_kernel(float *p, int Nx, int ld, int pml)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= Nx) return;
    float *c = p + (long)ix * ld;
    c[pml] = 0.0f;
    for (int iz = 0; iz <= pml; ++iz) c[pml - iz] = -c[pml + iz];
}

void launch_freesurface(float *p, int Nx, int ld, int pml, hipStream_t st)
{
    freesurface_kernel<<<(Nx + 63) / 64, 64, 0, st>>>(p, Nx, ld, pml);
}

// ---------------------------------------------------------------------------
// Model update of the FWI driver on the device (acoustic2d_fwi.c:418-465).
// Models are [nx][nz] (z fastest).

// smooth2D (common.c:927-1040): cos^2 window of half-length N along z (zdir 1)
// or x; float products accumulated in double in window order, divided by
// (sum of weights + FLT_MIN).  w: the 2N+1 float weights.
__global__ void smooth2d_kernel(const float *in, float *out, int nx, int nz, int N, int zdir,
                                const float *w)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nx * nz) return;
    const int ix = (int)(e / nz), iz = (int)(e - (long)ix * nz);
    const int j = zdir ? iz : ix, n = zdir ? nz : nx;
    const int ws = (j >= N) ? 0 : N - j;
    const int ds = (j >= N) ? j - N : 0;
    const int we = (n - 1 - j < N) ? 2 * N - (N - (n - 1 - j)) : 2 * N;
    double a1 = 0.0, a2 = 0.0;
    for (int i = ws, i2 = ds; i <= we; ++i, ++i2) {
        const float v = zdir ? in[(long)ix * nz + i2] : in[(long)i2 * nz + iz];
        a1 += (double)(w[i] * v);
        a2 += (double)w[i];
    }
    out[e] = (float)(a1 / (a2 + (double)FLT_MIN));
}

void launch_smooth2d(const float *in, float *out, int nx, int nz, int N, int zdir, const float *w,
                     hipStream_t st)
{
    const long n = (long)nx * nz;
    smooth2d_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, out, nx, nz, N, zdir, w);
}

// op: 0 out = a + b (vadd), 1 out = s * a (nvmul), 2 out = a != 0 ? 1 / a : 0
// (v2slow, slow2v), 3 clip a to [lo, hi] (clip_model), 4 out = 310 a^0.25
// (vp2density, common.c:298-330: 310.0f * powf(vp, 0.25f)), 5 rows iz < wd of
// every column set to s (set_waterlayer_value)
__global__ void model_op_kernel(int op, const float *a, const float *b, float *out, long n, int nz,
                                float s, float lo, float hi, int wd)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    switch (op) {
    case 0: out[e] = a[e] + b[e]; break;
    case 1: out[e] = s * a[e]; break;
    case 2: out[e] = (a[e] != 0.0f) ? 1.0f / a[e] : 0.0f; break;
    case 3: {
        const float v = a[e];
        out[e] = v > hi ? hi : (v < lo ? lo : v);
        break;
    }
    case 4: {
        // powf(v, 0.25f) correctly rounded: the double power rounded once to float
        const float q = (float)pow((double)a[e], 0.25);
        out[e] = 310.0f * q;
        break;
    }
    case 5:
        if ((int)(e % nz) < wd) out[e] = s;
        break;
    }
}

void launch_model_op(int op, const float *a, const float *b, float *out, long n, int nz, float s,
                     float lo, float hi, int wd, hipStream_t st)
{
    if (n > 0)
        model_op_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(op, a, b, out, n, nz, s, lo,
                                                                     hi, wd);
}

// ---------------------------------------------------------------------------
// FP16 checkpoint storage with a per-snapshot scale.
//
// A checkpoint field of shot s is stored as half(u * 2^k_s), where 2^k_s is the
// largest power of two with max|u_s| * 2^k_s <= 65504 (k_s = 0 for an all-zero
// field).  A power-of-two scale is exact in FP32, so the only rounding is the
// conversion to FP16 (round to nearest even), and restoring is exact:
// u' = float(h) * 2^-k_s.  Shot s, cell i of a field is at s*SS + i*CS
// (per-shot arenas: SS = arena, CS = 1; shots in lanes: SS = 1, CS = Sp); the
// FP16 slot uses strides SS16, CS16 of the same shape.

__global__ void fill_u32_kernel(unsigned *p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u32(unsigned *p, int n, unsigned v, hipStream_t st)
{
    if (n > 0) fill_u32_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, n, v);
}

// max |u| per shot, as the bit pattern of a non-negative float (ordered like
// the float).  Each block reduces its cells in registers and LDS and takes its
// maximum into copy blockIdx.x % CK_MX_COPIES of the shot words (mxs words
// per copy, cleared to 0 beforehand), so that the blocks' atomics meet rarely;
// ck_maxabs_fold_kernel then takes the maximum over the copies into copy 0.
// The maximum is exact whatever the order.  lanes: a wave reads one cell of 64
// shots at a time; otherwise block y is one shot and the block strides cells.
__global__ __launch_bounds__(256) void ck_maxabs_kernel(const float *f, long SS, long CS,
                                                        long ncell, int nshot, int lanes, long mxs,
                                                        unsigned *mx)
{
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned *dst = mx + (long)(blockIdx.x % CK_MX_COPIES) * mxs;
    float m0 = 0.0f, m1 = 0.0f;
    if (lanes) {
        const int s = blockIdx.y * 64 + lane;
        const long nw = (long)gridDim.x * 4;
        long i = (long)blockIdx.x * 4 + w;
        // two loads in flight per wave
        for (; i + nw < ncell; i += 2 * nw) {
            m0 = fmaxf(m0, fabsf(f[s * SS + i * CS]));
            m1 = fmaxf(m1, fabsf(f[s * SS + (i + nw) * CS]));
        }
        if (i < ncell) m0 = fmaxf(m0, fabsf(f[s * SS + i * CS]));
        red[w][lane] = fmaxf(m0, m1);
        __syncthreads();
        if (w == 0) {
            const float m = fmaxf(fmaxf(red[0][lane], red[1][lane]), fmaxf(red[2][lane], red[3][lane]));
            if (s < nshot) atomicMax(&dst[s], __float_as_uint(m));
        }
    } else {
        const int s = blockIdx.y;
        for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ncell;
             i += (long)gridDim.x * blockDim.x)
            m0 = fmaxf(m0, fabsf(f[s * SS + i * CS]));
        for (int off = 32; off > 0; off >>= 1) m0 = fmaxf(m0, __shfl_xor(m0, off));
        if (lane == 0) red[w][0] = m0;
        __syncthreads();
        if (threadIdx.x == 0) {
            const float m = fmaxf(fmaxf(red[0][0], red[1][0]), fmaxf(red[2][0], red[3][0]));
            atomicMax(&dst[s], __float_as_uint(m));
        }
    }
}

__global__ void ck_maxabs_fold_kernel(unsigned *mx, long mxs, int nshot)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshot) return;
    unsigned m = mx[s];
    for (int k = 1; k < CK_MX_COPIES; ++k) m = max(m, mx[(long)k * mxs + s]);
    mx[s] = m;
}

__device__ __forceinline__ float ck_scale_of(unsigned bits)
{
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f)) return 1.0f;
    int e;
    (void)frexpf(65504.0f / m, &e);  // 65504/m = f * 2^e, f in [0.5, 1)
    float sc = ldexpf(1.0f, e - 1);
    if (m * sc > 65504.0f) sc *= 0.5f;  // rounding of the quotient at a power of two
    return sc;
}

__global__ void ck_pack_kernel(const float *f, long SS, long CS, __half *h, long SS16, long CS16,
                               long ncell, int nshot, int lanes, const unsigned *mx, float *inv)
{
    long i;
    int s;
    if (lanes) {
        s = blockIdx.y * 64 + (threadIdx.x & 63);
        i = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    } else {
        s = blockIdx.y;
        i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    }
    if (i >= ncell || s >= nshot) return;
    const float sc = ck_scale_of(mx[s]);
    h[s * SS16 + i * CS16] = __float2half_rn(f[s * SS + i * CS] * sc);
    if (i == 0) inv[s] = 1.0f / sc;  // a power of two: exact
}

__global__ void ck_unpack_kernel(const __half *h, long SS16, long CS16, float *f, long SS, long CS,
                                 long ncell, int nshot, int lanes, const float *inv)
{
    long i;
    int s;
    if (lanes) {
        s = blockIdx.y * 64 + (threadIdx.x & 63);
        i = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    } else {
        s = blockIdx.y;
        i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    }
    if (i >= ncell || s >= nshot) return;
    f[s * SS + i * CS] = __half2float(h[s * SS16 + i * CS16]) * inv[s];
}

static dim3 ck_grid(long ncell, int nshot, int lanes, int reduce)
{
    if (lanes) {
        const long nb = reduce ? std::min<long>((ncell + 3) / 4, 1024) : (ncell + 3) / 4;
        return dim3((unsigned)nb, (unsigned)((nshot + 63) / 64));
    }
    const long nb = reduce ? std::min<long>((ncell + 255) / 256, 512) : (ncell + 255) / 256;
    return dim3((unsigned)nb, (unsigned)nshot);
}

void launch_ck_pack(const float *f, long SS, long CS, __half *h, long SS16, long CS16, long ncell,
                    int nshot, int lanes, unsigned *mx, float *inv, hipStream_t st)
{
    const long mxs = ck_mx_stride(nshot);
    launch_fill_u32(mx, (int)ck_mx_words(nshot), 0u, st);
    ck_maxabs_kernel<<<ck_grid(ncell, nshot, lanes, 1), 256, 0, st>>>(f, SS, CS, ncell, nshot,
                                                                        lanes, mxs, mx);
    ck_maxabs_fold_kernel<<<(nshot + 63) / 64, 64, 0, st>>>(mx, mxs, nshot);
    ck_pack_kernel<<<ck_grid(ncell, nshot, lanes, 0), 256, 0, st>>>(f, SS, CS, h, SS16, CS16,
                                                                      ncell, nshot, lanes, mx, inv);
}

void launch_ck_unpack(const __half *h, long SS16, long CS16, float *f, long SS, long CS,
                      long ncell, int nshot, int lanes, const float *inv, hipStream_t st)
{
    ck_unpack_kernel<<<ck_grid(ncell, nshot, lanes, 0), 256, 0, st>>>(h, SS16, CS16, f, SS, CS,
                                                                        ncell, nshot, lanes, inv);
}

}  // namespace fwiedu
