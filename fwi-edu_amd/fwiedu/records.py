"""ctypes mirror of the reference's shot-record structure and 2D array convention.

`Record2D` has the field order and types of ``Record2D`` in
FWILab2d/FWILab2dCLangV1.2.6/include/observation.h:18-48, so an array of it can be
handed to any C entry point that takes ``Record2D *``.

2D arrays follow Seismic Unix ``alloc2float(n1, n2)``: one contiguous block of
n2*n1 floats with n1 fastest plus a table of n2 row pointers.  Here that is a
C-contiguous numpy array of shape (n2, n1) plus a pointer table built from it.
"""
import ctypes
from ctypes import POINTER, c_float, c_int

import numpy as np

c_float_p = POINTER(c_float)
c_float_pp = POINTER(c_float_p)
c_int_p = POINTER(c_int)


class Record2D(ctypes.Structure):
    _fields_ = [
        ("shotID", c_int),
        ("sx", c_float),
        ("sz", c_float),
        ("isx", c_int),
        ("isz", c_int),
        ("fpeak", c_float),
        ("src", c_float_p),
        ("gx", c_float_p),
        ("gz", c_float_p),
        ("igx", c_int_p),
        ("igz", c_int_p),
        ("sampleRate", c_int),
        ("ntr", c_int),
        ("ns", c_int),
        ("dtr", c_float),
        ("dt", c_float),
        ("ix_min", c_int),
        ("ix_max", c_int),
        ("traces_vx", c_float_pp),
        ("traces_vz", c_float_pp),
        ("traces_p", c_float_pp),
        ("precond", c_float_pp),
        ("grad", c_float_pp),
    ]


def fptr(a):
    """float* to the first element of a float32 numpy array (kept alive by the caller)."""
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_float_p)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int_p)


class Array2D:
    """A (n2, n1) float32 block with its SU-style row-pointer table."""

    def __init__(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        assert arr.ndim == 2
        self.a = arr
        n2 = arr.shape[0]
        base = arr.ctypes.data
        stride = arr.strides[0]
        self.rows = (c_float_p * n2)(*[ctypes.cast(base + i * stride, c_float_p) for i in range(n2)])

    @property
    def pp(self):
        return ctypes.cast(self.rows, c_float_pp)


class ShotSet:
    """An array of Record2D plus the numpy buffers it points into.

    Geometry follows create_grid_observe2d (observation.c:338-420): grid indices
    are floorf(coordinate / spacing); the simulation window is the receiver span
    widened by 30 cells and clipped to the model; isx and igx are window-local,
    isz and igz global.  Records hold ns = (nt-1)/sampleRate + 1 samples
    (observation.c:313) at dt*sampleRate.
    """

    def __init__(self, sx, sz, gx_list, gz_list, src, nt, dt, dx, dz, nx, fpeak,
                 sample_rate=1, with_traces=True):
        self.nshot = len(sx)
        self.recs = (Record2D * self.nshot)()
        self.keep = []
        self.traces = []
        self.trace_tables = []
        src = np.ascontiguousarray(src, dtype=np.float32)
        self.src = src
        dx32, dz32 = np.float32(dx), np.float32(dz)
        for s in range(self.nshot):
            r = self.recs[s]
            gx = np.ascontiguousarray(gx_list[s], dtype=np.float32)
            gz = np.ascontiguousarray(gz_list[s], dtype=np.float32)
            ntr = gx.size
            igx = np.floor(gx / dx32).astype(np.int32)
            igz = np.floor(gz / dz32).astype(np.int32)
            ix_min = max(int(igx.min()) - 30, 0)
            ix_max = min(int(igx.max()) + 30, nx - 1)
            igx = np.ascontiguousarray(igx - ix_min, dtype=np.int32)
            igz = np.ascontiguousarray(igz, dtype=np.int32)
            r.shotID = s + 1
            r.sx = float(np.float32(sx[s]))
            r.sz = float(np.float32(sz[s]))
            r.isx = int(np.floor(np.float32(sx[s]) / dx32)) - ix_min
            r.isz = int(np.floor(np.float32(sz[s]) / dz32))
            r.fpeak = float(fpeak)
            r.src = fptr(src)
            r.gx, r.gz, r.igx, r.igz = fptr(gx), fptr(gz), iptr(igx), iptr(igz)
            r.sampleRate = sample_rate
            r.ntr = ntr
            r.ns = (nt - 1) // sample_rate + 1
            r.dt = float(np.float32(dt) * np.float32(sample_rate))
            if ntr > 1:
                d = np.float32(0.0)
                for k in range(ntr - 1, 0, -1):
                    d = np.float32(d + (gx[k] - gx[k - 1]))
                r.dtr = float(d / np.float32(ntr - 1))
            else:
                r.dtr = 0.0
            r.ix_min, r.ix_max = ix_min, ix_max
            self.keep += [gx, gz, igx, igz]
            if with_traces:
                tr = Array2D(np.zeros((ntr, r.ns), np.float32))
                self.traces.append(tr)
                r.traces_p = tr.pp
            else:
                self.traces.append(None)

    def header_copy(self):
        """copy_record2d_header (observation.c:173-226) with fresh zeroed traces.
        Geometry and source pointers are shared with self, as in the reference."""
        other = ShotSet.__new__(ShotSet)
        other.nshot = self.nshot
        other.recs = (Record2D * self.nshot)()
        other.keep = [self]
        other.src = self.src
        other.traces = []
        for s in range(self.nshot):
            ctypes.pointer(other.recs[s])[0] = self.recs[s]
            r = other.recs[s]
            tr = Array2D(np.zeros((r.ntr, r.ns), np.float32))
            other.traces.append(tr)
            r.traces_p = tr.pp
        return other

    def trace(self, s):
        return self.traces[s].a

    def all_traces(self):
        return [t.a for t in self.traces]

    def set_traces(self, s, data):
        self.traces[s].a[...] = data
