"""Mesh import (Gmsh 2.x/4.x, VTK legacy, STL ASCII) and result export (VTK
fields, the columnar CSV/VTK/text-archive contract of the reference's
ResultExporter), thin calls into the host part of the C ABI."""
import ctypes as C

import numpy as np

MESH_AUTO, MESH_GMSH, MESH_VTK, MESH_STL = 0, 1, 2, 3


def _lib():
    from . import lib
    return lib()


def _check(rc, what):
    from . import OK, ScdatError
    if rc != OK:
        msg = _lib().scdat_io_last_error()
        raise ScdatError(rc, (msg.decode() if msg else what))


def read_mesh(path, fmt=MESH_AUTO, tag_groups=None, default_group=0):
    """Nodes, tetrahedra and triangles with their physical tags; with
    tag_groups ({physical tag: boundary group}) also the per-face group array
    for Context.set_mesh (-1 on interior faces)."""
    h = C.c_void_p()
    _check(_lib().scdat_mesh_read(str(path).encode(), fmt, C.byref(h)), "mesh_read")
    try:
        nn, nt, nf = C.c_int64(), C.c_int64(), C.c_int64()
        _check(_lib().scdat_mesh_sizes(h, C.byref(nn), C.byref(nt), C.byref(nf)), "mesh_sizes")
        out = {"xyz": np.zeros((nn.value, 3)), "tets": np.zeros((nt.value, 4), np.int32),
               "tet_tag": np.zeros(nt.value, np.int32), "tris": np.zeros((nf.value, 3), np.int32),
               "tri_tag": np.zeros(nf.value, np.int32)}
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        _check(_lib().scdat_mesh_arrays(h, p(out["xyz"], C.c_double), p(out["tets"], C.c_int32),
                                        p(out["tet_tag"], C.c_int32), p(out["tris"], C.c_int32),
                                        p(out["tri_tag"], C.c_int32)), "mesh_arrays")
        if tag_groups is not None:
            tags = np.array(list(tag_groups.keys()), np.int32)
            groups = np.array(list(tag_groups.values()), np.int32)
            fg = np.zeros((nt.value, 4), np.int32)
            m = C.c_int64()
            _check(_lib().scdat_mesh_face_groups(h, len(tags), p(tags, C.c_int32), p(groups, C.c_int32),
                                                 int(default_group), p(fg, C.c_int32), C.byref(m)),
                   "mesh_face_groups")
            out["group"] = fg
            out["faces_matched"] = m.value
        return out
    finally:
        _lib().scdat_mesh_free(h)


def write_vtk(path, xyz, tets, fields=None):
    """Tetrahedral mesh with point fields ({name: (n,) or (n, 3) array})."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    tets = np.ascontiguousarray(tets, np.int32)
    fields = fields or {}
    arrs = [np.ascontiguousarray(v, np.float64) for v in fields.values()]
    names = (C.c_char_p * max(1, len(arrs)))(*[k.encode() for k in fields.keys()])
    comps = np.array([1 if a.ndim == 1 else a.shape[1] for a in arrs] or [1], np.int32)
    data = (C.POINTER(C.c_double) * max(1, len(arrs)))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    _check(_lib().scdat_write_vtk(str(path).encode(), len(xyz), xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                  len(tets), tets.ctypes.data_as(C.POINTER(C.c_int32)), len(arrs), names,
                                  comps.ctypes.data_as(C.POINTER(C.c_int32)), data), "write_vtk")


def export_columns(csv_path, axis_name, axis, series, module="PIC-MCC"):
    """ResultExporter::exportDataSet contract: csv_path plus <stem>.vtk and
    <stem>.h5 with .metadata.json sidecars; series = {name: array}."""
    axis = np.ascontiguousarray(axis, np.float64)
    arrs = [np.ascontiguousarray(v, np.float64) for v in series.values()]
    names = (C.c_char_p * max(1, len(arrs)))(*[k.encode() for k in series.keys()])
    data = (C.POINTER(C.c_double) * max(1, len(arrs)))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    _check(_lib().scdat_export_columns(str(csv_path).encode(), module.encode(), axis_name.encode(), len(axis),
                                       axis.ctypes.data_as(C.POINTER(C.c_double)), len(arrs), names, data),
           "export_columns")
