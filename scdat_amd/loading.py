"""Synthetic inputs: Philox4x32-10 stream and the Maxwellian box loader.

Both are thin calls into the host part of the C ABI (no device needed), so
the device path and the CPU checker consume identical seeded states.
"""
import ctypes as C

import numpy as np


def philox4x32(ctr, key):
    """Philox4x32-10 of a 4-word counter and 2-word key (Salmon et al. SC'11)."""
    from . import lib
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (C.c_uint32 * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (C.c_uint32 * 4)()
    lib().scdat_philox4x32(c, k, o)
    return [int(v) for v in o]


def box_locate(spec, x):
    from . import lib, ScdatError, OK
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    cell = np.zeros(len(x), np.int32)
    rc = lib().scdat_box_locate(C.byref(spec), len(x), x.ctypes.data_as(C.POINTER(C.c_double)),
                                cell.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != OK:
        raise ScdatError(rc, "box_locate failed")
    return cell


def maxwellian_box(spec, n, seed, stream, vth, drift=None):
    """n particles uniform in the box minus the conductor, N(drift, vth^2)
    velocities; returns (x[n,3], v[n,3], cell[n])."""
    from . import lib, ScdatError, OK
    x = np.zeros((n, 3))
    v = np.zeros((n, 3))
    cell = np.zeros(n, np.int32)
    d = None if drift is None else np.ascontiguousarray(drift, np.float64)
    rc = lib().scdat_load_maxwellian(
        C.byref(spec), int(n), int(seed), int(stream), float(vth),
        d.ctypes.data_as(C.POINTER(C.c_double)) if d is not None else None,
        x.ctypes.data_as(C.POINTER(C.c_double)), v.ctypes.data_as(C.POINTER(C.c_double)),
        cell.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != OK:
        raise ScdatError(rc, "load_maxwellian failed")
    return x, v, cell
