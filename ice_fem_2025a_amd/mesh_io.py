"""IceFEM HDF5 meshes (the format of TestCases/TriAxialCompression/Meshgen/SaveToHDF.m:1-83,
read by Mesh::LoadFromFile, mesh/mesh.cpp:53-303) through the engine's C-ABI
(ife_mesh_load_hdf5 / ife_mesh_save_hdf5, host code over the HDF5 C library).

    m = load_hdf5("Mesh.h5")            # element group "internal", as the models name it
    prob.mesh = m                       # the engine and the oracle take it like a generated mesh
    save_hdf5("Mesh.h5", patch_mesh, node_groups={"bottom": ids})
"""
import ctypes as C

import numpy as np

from . import abi


class MeshFile:
    """One element group of an IceFEM HDF5 mesh, as ife_mesh_desc holds it: coords
    (n_nodes, dim), weights (n_nodes,), elems (n_elems, n) int64, dense extraction
    operators (n_elems, n, n) with C[e][j][k] = ElemData(j, k) (elementgroup.cpp:62), plus
    the file's node groups {name: node ids}."""

    extraction_1d = None  # the file holds dense operators

    def __init__(self, dim, ipcount1D, coords, weights, elems, extraction, group_name, group_type, node_groups):
        self.dim, self.ipcount1D = dim, ipcount1D
        self.coords, self.weights, self.elems, self.extraction = coords, weights, elems, extraction
        self.group_name, self.group_type, self.node_groups = group_name, group_type, node_groups

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elems(self):
        return self.elems.shape[0]

    @property
    def nodes_per_elem(self):
        return self.elems.shape[1]

    def dense_extraction(self, elements=None):
        return self.extraction if elements is None else self.extraction[elements]


def load_hdf5(path, group="internal", ipcount1D=3):
    """Read `path`; `group` is an element-group name or index."""
    L = abi.load_library()
    h = C.c_void_p()
    abi.check(L.ife_mesh_load_hdf5(str(path).encode(), ipcount1D, C.byref(h)))
    try:
        dim, nn, neg, nng = C.c_int32(), C.c_int64(), C.c_int32(), C.c_int32()
        abi.check(L.ife_mesh_file_info(h, C.byref(dim), C.byref(nn), C.byref(neg), C.byref(nng)))
        names = []
        for g in range(neg.value):
            nm, tp, ne, npe = C.c_char_p(), C.c_char_p(), C.c_int64(), C.c_int32()
            abi.check(L.ife_mesh_file_element_group(h, g, C.byref(nm), C.byref(tp), C.byref(ne), C.byref(npe)))
            names.append((nm.value.decode(), tp.value.decode()))
        gi = group if isinstance(group, int) else [n for n, _ in names].index(group)
        d = abi.MeshDesc()
        abi.check(L.ife_mesh_file_desc(h, gi, C.byref(d)))
        n = d.nodes_per_elem
        coords = np.ctypeslib.as_array(d.coords, shape=(d.n_nodes, d.dim)).copy()
        weights = np.ctypeslib.as_array(d.weights, shape=(d.n_nodes,)).copy()
        elems = np.ctypeslib.as_array(d.elems, shape=(d.n_elems, n)).copy()
        ext = np.ctypeslib.as_array(d.extraction, shape=(d.n_elems, n, n)).copy()
        ngroups = {}
        for g in range(nng.value):
            nm, cnt, ptr = C.c_char_p(), C.c_int64(), C.POINTER(C.c_int64)()
            abi.check(L.ife_mesh_file_node_group(h, g, C.byref(nm), C.byref(cnt), C.byref(ptr)))
            ngroups[nm.value.decode()] = (np.ctypeslib.as_array(ptr, shape=(cnt.value,)).copy() if cnt.value
                                          else np.zeros(0, np.int64))
        return MeshFile(d.dim, d.ipcount1D, coords, weights, elems, ext, names[gi][0], names[gi][1], ngroups)
    finally:
        L.ife_mesh_file_free(h)


def save_hdf5(path, mesh, group_name="internal", node_groups=None):
    """Write `mesh` (a generated PatchMesh or a MeshFile) in IceFEM's layout."""
    L = abi.load_library()
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    weights = np.ascontiguousarray(mesh.weights, dtype=np.float64)
    elems = np.ascontiguousarray(mesh.elems, dtype=np.int64)
    d = abi.MeshDesc()
    d.dim, d.ipcount1D = mesh.dim, mesh.ipcount1D
    d.n_nodes, d.n_elems, d.nodes_per_elem = mesh.n_nodes, mesh.n_elems, mesh.nodes_per_elem
    d.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
    d.weights = weights.ctypes.data_as(C.POINTER(C.c_double))
    d.elems = elems.ctypes.data_as(C.POINTER(C.c_int64))
    if getattr(mesh, "extraction_1d", None) is not None:
        e1 = np.ascontiguousarray(mesh.extraction_1d, dtype=np.float64)
        d.extraction_1d = e1.ctypes.data_as(C.POINTER(C.c_double))
    else:
        ex = np.ascontiguousarray(mesh.extraction, dtype=np.float64)
        d.extraction = ex.ctypes.data_as(C.POINTER(C.c_double))
    node_groups = node_groups or {}
    names = list(node_groups)
    arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    ptr = np.zeros(len(names) + 1, dtype=np.int64)
    for i, n in enumerate(names):
        ptr[i + 1] = ptr[i] + len(node_groups[n])
    nodes = np.ascontiguousarray(np.concatenate([np.asarray(node_groups[n], dtype=np.int64) for n in names])
                                 if names else np.zeros(1, np.int64))
    abi.check(L.ife_mesh_save_hdf5(str(path).encode(), C.byref(d), group_name.encode(), len(names),
                                   C.cast(arr, C.POINTER(C.c_char_p)), ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                   nodes.ctypes.data_as(C.POINTER(C.c_int64))))
