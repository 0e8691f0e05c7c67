"""Seeded synthetic FWI workloads (no datasets are available offline).

Each workload is the set of inputs the reference's FWI driver would read
(acoustic2d_fwi.c:339-393): a true model for the observed data, a starting model,
a Ricker source, a shot/receiver geometry, and the driver parameters.  Models
are float32 arrays of shape (nx, nz), z fastest, the layout of the reference's
``alloc2float(nz, nx)`` blocks and of its SU-style binary files.

Names and sizes:
  c0     101 x 201 @ 10 m, nt 1500, 4 shots, 201 receivers, pml 30, no free surface
  c1     221 x 601 @ 12.5 m, nt 4001, 120 shots, 601 receivers, pml 40, free surface
  c2     400 x 1200 @ 10 m, nt 5000, 256 shots, 1200 receivers, pml 40, free surface
  c4     1000 x 3000 @ 5 m, nt 8000, 64 shots, 3000 receivers, pml 40, free surface
  tiny   30 x 48 @ 10 m, nt 240, 3 shots -- unit-test size
Any field can be overridden (e.g. nshot=2, nt=500) to cut a case down.
"""
from dataclasses import dataclass, field

import numpy as np

from .records import ShotSet


def ricker(nt, dt, f0, t0):
    """(1 - 2a) exp(-a), a = (pi f0 (t - t0))^2, t = it*dt, evaluated in float64 and
    rounded to float32 (reproduces the reference's shipped
    ricker_6001_1ms_10Hz_delay0.15ms.bin bit for bit)."""
    t = np.arange(nt, dtype=np.float64) * float(dt)
    a = (np.pi * f0 * (t - t0)) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def peak_frequency(src, dt):
    """calculate_peak_frequency (common.c:363-413): first bin of max |rfft|, times
    df = 1/(nt*dt) formed in float32."""
    nt = src.size
    amp = np.abs(np.fft.rfft(src.astype(np.float64)))
    i = int(np.argmax(amp))
    df = np.float32(1.0) / (np.float32(nt) * np.float32(dt))
    return float(np.float32(i) * df)


def gardner(vp):
    """vp2density (common.c:298-330): 310 * vp^0.25 in float32."""
    return (np.float32(310.0) * np.power(vp.astype(np.float32), np.float32(0.25))).astype(np.float32)


def smooth_gaussian(m, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(m.astype(np.float64), sigma=sigma, mode="nearest").astype(np.float32)


def layered_model(nz, nx, seed, water_rows, nlayers=12, vmin=1600.0, vmax=4700.0,
                  nfaults=2, throw=(5, 15), noise=0.02, dip_deg=10.0):
    """Water layer, `nlayers` randomly dipping layers with sorted velocities in
    [vmin, vmax], `nfaults` vertical faults with seeded throws, and a
    (1 + noise*N(0,1)) texture below the water."""
    rng = np.random.default_rng(seed)
    vel = np.sort(rng.uniform(vmin, vmax, nlayers))
    tops = np.sort(rng.uniform(water_rows, nz * 0.95, nlayers))
    tops[0] = water_rows
    dips = np.tan(np.deg2rad(rng.uniform(-dip_deg, dip_deg, nlayers)))
    dips[0] = 0.0
    xc = nx / 2.0
    x = np.arange(nx, dtype=np.float64)[:, None]
    z = np.arange(nz, dtype=np.float64)[None, :]
    shift = np.zeros((nx, 1))
    fx = np.sort(rng.uniform(0.2 * nx, 0.8 * nx, nfaults))
    thr = rng.uniform(throw[0], throw[1], nfaults)
    for f, t in zip(fx, thr):
        shift += np.where(x > f, t, 0.0)
    zz = z - shift
    idx = np.zeros((nx, nz), dtype=np.int64)
    for k in range(nlayers):
        iface = tops[k] + dips[k] * (x - xc)
        idx = np.where(zz >= iface, k, idx)
    v = vel[idx]
    v *= 1.0 + noise * rng.standard_normal((nx, nz))
    v[:, :water_rows] = 1500.0
    return v.astype(np.float32)


def anomaly_model(nz, nx, seed, water_rows=20):
    """Water rows, then a linear 1800 -> 3500 m/s depth gradient, times three
    +/-10% Gaussian anomalies at seeded centres; at 101 x 201 the anomalies have
    sigma 8 cells and centres in rows 30-90, columns 20-180."""
    rng = np.random.default_rng(seed)
    iz = np.arange(nz, dtype=np.float64)[None, :]
    ix = np.arange(nx, dtype=np.float64)[:, None]
    v = np.where(iz < water_rows, 1500.0,
                 1800.0 + 1700.0 * (iz - water_rows) / (nz - 1 - water_rows))
    v = np.broadcast_to(v, (nx, nz)).copy()
    zc = rng.uniform(0.3 * (nz - 1), 0.9 * (nz - 1), 3)
    xc = rng.uniform(0.1 * (nx - 1), 0.9 * (nx - 1), 3)
    sig = 8.0 * nz / 101.0
    for sgn, a, b in zip((1.0, -1.0, 1.0), zc, xc):
        r2 = (iz - a) ** 2 + (ix - b) ** 2
        v *= 1.0 + sgn * 0.10 * np.exp(-r2 / (2.0 * sig ** 2))
    v[:, :water_rows] = 1500.0
    return v.astype(np.float32)


@dataclass
class Workload:
    name: str
    nz: int
    nx: int
    dx: float
    dz: float
    dt: float
    nt: int
    pml: int
    freesurface: int
    nshot: int
    shot_depth_cells: int
    rcv_depth_cells: int
    f0: float
    t0: float
    seed: int
    waterdepth: int
    vp_water: float = 1500.0
    rho_water: float = 1000.0
    vpmin: float = 0.001
    vpmax: float = 10000.0
    rhomin: float = 0.001
    rhomax: float = 10000.0
    filter: int = 2
    freq: float = 0.0
    precondition: int = 1
    sample_rate: int = 1
    smooth_sigma: float = 15.0
    kind: str = "layered"
    nrcv: int = 0
    vp_true: np.ndarray = field(default=None, repr=False)
    rho_true: np.ndarray = field(default=None, repr=False)
    vp_start: np.ndarray = field(default=None, repr=False)
    rho_start: np.ndarray = field(default=None, repr=False)
    src: np.ndarray = field(default=None, repr=False)
    fpeak: float = 0.0

    def build(self):
        nz, nx = self.nz, self.nx
        if self.kind == "anomaly":
            vp = anomaly_model(nz, nx, self.seed, self.waterdepth)
        else:
            vp = layered_model(nz, nx, self.seed, self.waterdepth)
        rho = gardner(vp)
        vp[:, :self.waterdepth] = self.vp_water
        rho[:, :self.waterdepth] = self.rho_water
        self.vp_true, self.rho_true = vp, rho
        vs = smooth_gaussian(vp, self.smooth_sigma)
        vs[:, :self.waterdepth] = self.vp_water
        rs = gardner(vs)
        rs[:, :self.waterdepth] = self.rho_water
        self.vp_start, self.rho_start = vs, rs
        self.src = ricker(self.nt, self.dt, self.f0, self.t0)
        self.fpeak = peak_frequency(self.src, self.dt)
        return self

    def shot_columns(self):
        n = self.nshot
        return [((2 * i + 1) * self.nx) // (2 * n) for i in range(n)]

    def receiver_columns(self):
        if self.nrcv and self.nrcv < self.nx:
            return np.round(np.linspace(0, self.nx - 1, self.nrcv)).astype(np.int64)
        return np.arange(self.nx)

    def shots(self, with_traces=True):
        dx, dz = np.float32(self.dx), np.float32(self.dz)
        sx = [np.float32(c) * dx for c in self.shot_columns()]
        sz = [np.float32(self.shot_depth_cells) * dz] * self.nshot
        rc = self.receiver_columns()
        gx = [(rc.astype(np.float32) * dx).astype(np.float32) for _ in range(self.nshot)]
        gz = [np.full(rc.size, np.float32(self.rcv_depth_cells) * dz, np.float32)
              for _ in range(self.nshot)]
        return ShotSet(sx, sz, gx, gz, self.src, self.nt, self.dt, self.dx, self.dz, self.nx,
                       self.fpeak, self.sample_rate, with_traces)

    def cell_updates_per_gradient_shot(self):
        """(2*nt + R) interior sweeps times interior cells (loop bounds of
        propagation.c:534-537); R from the checkpoint schedule."""
        Nz, Nx = self.nz + 2 * self.pml, self.nx + 2 * self.pml
        zs = self.pml if self.freesurface else 3
        cells = (Nx - 6) * (Nz - 3 - zs)
        return cells * (2 * self.nt + recompute_steps(self.nt))

    def cell_updates_per_modeling_shot(self):
        Nz, Nx = self.nz + 2 * self.pml, self.nx + 2 * self.pml
        zs = self.pml if self.freesurface else 3
        return (Nx - 6) * (Nz - 3 - zs) * self.nt


def checkpoint_plan(nt):
    """alloc_acoustic2dCheckPoint (checkpoint.c:29-99): d, nCheckPoint, ntPoint."""
    d = max(int(np.round(np.float32(np.sqrt(np.float32(2.0) * np.float32(nt - 1))))), 3)
    n = sum(1 for it in range(nt) if it % d == 0 or (it + 1) % d == 0)
    return d, n, n + d - 2


def recompute_steps(nt):
    """Recompute sweeps of the reverse loop: each segment [kd, kd+d) below the top
    recomputes kd..kd+d-3 (d-2 sweeps); the top segment recomputes up to nt-1."""
    d, _, _ = checkpoint_plan(nt)
    stored = np.zeros(nt, bool)
    for it in range(0, nt, d):
        stored[it] = True
        if it >= 1:
            stored[it - 1] = True
    r = 0
    for it in range(nt - 1, -1, -1):
        if stored[it]:
            continue
        s = it
        while not stored[s]:
            s -= 1
        r += it - s
        stored[s + 1:it + 1] = True
    return r


PRESETS = {
    "tiny": dict(nz=30, nx=48, dx=10.0, dz=10.0, dt=0.001, nt=240, pml=12, freesurface=0,
                 nshot=3, shot_depth_cells=2, rcv_depth_cells=2, f0=25.0, t0=0.05, seed=7,
                 waterdepth=5, kind="anomaly", smooth_sigma=4.0),
    "c0": dict(nz=101, nx=201, dx=10.0, dz=10.0, dt=0.001, nt=1500, pml=30, freesurface=0,
               nshot=4, shot_depth_cells=2, rcv_depth_cells=2, f0=10.0, t0=0.15, seed=101,
               waterdepth=20, kind="anomaly"),
    "c1": dict(nz=221, nx=601, dx=12.5, dz=12.5, dt=0.001, nt=4001, pml=40, freesurface=1,
               nshot=120, shot_depth_cells=1, rcv_depth_cells=1, f0=10.0, t0=0.15, seed=202,
               waterdepth=30, vpmin=1400.0, vpmax=5000.0, kind="layered"),
    "c2": dict(nz=400, nx=1200, dx=10.0, dz=10.0, dt=0.0008, nt=5000, pml=40, freesurface=1,
               nshot=256, shot_depth_cells=1, rcv_depth_cells=1, f0=12.0, t0=0.15, seed=303,
               waterdepth=54, vpmin=1400.0, vpmax=5000.0, kind="layered"),
    "c4": dict(nz=1000, nx=3000, dx=5.0, dz=5.0, dt=0.0004, nt=8000, pml=40, freesurface=1,
               nshot=64, shot_depth_cells=1, rcv_depth_cells=1, f0=25.0, t0=0.15, seed=505,
               waterdepth=136, vpmin=1400.0, vpmax=5000.0, kind="layered"),
}


def workload(name, build=True, **over):
    p = dict(PRESETS[name])
    p.update(over)
    w = Workload(name=name, **p)
    return w.build() if build else w
