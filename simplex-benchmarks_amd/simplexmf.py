"""Python host mirror of the simplexmf C ABI (ctypes).

Mirrors the reference operator's interface -- ``reinit`` (construction),
``vmult``, ``Tvmult``, ``m``, ``initialize_dof_vector``,
``compute_inverse_diagonal`` (CG/cells_batched/application.cc:1038-1196,
application_hp.cc:1386-1409) -- over the HIP library
``lib/libsimplexmf.so``.  Device vectors are :class:`DeviceVector` objects
allocated through the library's own HIP runtime.

The library is required: if it is missing every entry point raises; there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsimplexmf.so")

MF_CG_POISSON, MF_DG_SIPG, MF_CG_VECTOR, MF_DG_VECTOR, MF_DG_HELMHOLTZ = range(5)
MF_F64, MF_F32 = 0, 1
KINDS = {"cg": MF_CG_POISSON, "dg": MF_DG_SIPG, "cg_vector": MF_CG_VECTOR,
         "dg_vector": MF_DG_VECTOR, "dg_helmholtz": MF_DG_HELMHOLTZ}
PRECISIONS = {"f64": MF_F64, "f32": MF_F32}
NP_DTYPES = {MF_F64: np.float64, MF_F32: np.float32}


class MeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("vertices", C.POINTER(C.c_double)),
                ("n_cells", C.c_int64), ("cells", C.POINTER(C.c_int32)),
                ("kuhn_n", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("reorder", C.c_int32), ("patch_cells", C.c_int32),
                ("mass_coefficient", C.c_double), ("laplace_coefficient", C.c_double),
                ("penalty_kind", C.c_int32), ("host_only", C.c_int32), ("accumulate", C.c_int32),
                ("dg_kernel", C.c_int32), ("boundary", C.c_int32), ("cells_per_wave", C.c_int32),
                ("manifold", C.c_int32), ("manifold_center", C.c_double * 3)]


class Info(C.Structure):
    _fields_ = [("n_dofs", C.c_int64), ("n_cells", C.c_int64), ("n_constrained", C.c_int64),
                ("degree", C.c_int32), ("dofs_per_cell", C.c_int32), ("n_q_points", C.c_int32),
                ("n_components", C.c_int32), ("n_patches", C.c_int64),
                ("n_patch_colors", C.c_int32), ("max_cell_colors", C.c_int32),
                ("max_local_dofs", C.c_int32), ("precision", C.c_int32),
                ("n_interior_faces", C.c_int64), ("n_boundary_faces", C.c_int64),
                ("algorithmic_bytes", C.c_double), ("flops", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CsrInfo(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("nnz", C.c_int64), ("n_row_blocks", C.c_int64),
                ("precision", C.c_int32), ("block_entries", C.c_int32), ("algorithmic_bytes", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"simplexmf HIP library not built: {LIB_PATH} missing "
                           "(run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    sig = {
        "mf_mesh_kuhn_cube": (C.c_int, [i32, C.c_double, u64, vp, vp]),
        "mf_default_options": (None, [C.POINTER(Options)]),
        "mf_create": (C.c_int, [C.POINTER(MeshDesc), i32, i32, i32, C.POINTER(Options), C.POINTER(vp)]),
        "mf_destroy": (C.c_int, [vp]),
        "mf_vmult": (C.c_int, [vp, vp, vp, vp]),
        "mf_tvmult": (C.c_int, [vp, vp, vp, vp]),
        "mf_compute_diagonal": (C.c_int, [vp, vp, vp]),
        "mf_compute_inverse_diagonal": (C.c_int, [vp, vp, vp]),
        "mf_m": (i64, [vp]),
        "mf_get_info": (C.c_int, [vp, C.POINTER(Info)]),
        "mf_get_cell_order": (C.c_int, [vp, vp]),
        "mf_get_cell_dofs": (C.c_int, [vp, vp]),
        "mf_get_cell_vertices": (C.c_int, [vp, vp]),
        "mf_get_dof_points": (C.c_int, [vp, vp]),
        "mf_get_constrained": (C.c_int, [vp, vp]),
        "mf_debug_host_apply": (C.c_int, [vp, i32, vp, vp]),
        "mf_last_error": (C.c_char_p, []),
        "mf_device_count": (C.c_int, [C.POINTER(i32)]),
        "mf_device_name": (C.c_int, [C.c_char_p, i32]),
        "mf_malloc": (C.c_int, [C.POINTER(vp), u64]),
        "mf_free": (C.c_int, [vp]),
        "mf_memcpy_htod": (C.c_int, [vp, vp, u64]),
        "mf_memcpy_dtoh": (C.c_int, [vp, vp, u64]),
        "mf_memcpy_dtod": (C.c_int, [vp, vp, u64, vp]),
        "mf_copy_kernel": (C.c_int, [vp, vp, u64, vp]),
        "mf_memset": (C.c_int, [vp, i32, u64, vp]),
        "mf_synchronize": (C.c_int, []),
        "mf_stream_create": (C.c_int, [C.POINTER(vp)]),
        "mf_stream_destroy": (C.c_int, [vp]),
        "mf_stream_synchronize": (C.c_int, [vp]),
        "mf_event_create": (C.c_int, [C.POINTER(vp)]),
        "mf_event_destroy": (C.c_int, [vp]),
        "mf_event_record": (C.c_int, [vp, vp]),
        "mf_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
        "mf_mg_default_options": (None, [C.POINTER(MGOptions)]),
        "mf_mg_create": (C.c_int, [vp, C.POINTER(MGOptions), C.POINTER(vp)]),
        "mf_mg_vmult": (C.c_int, [vp, vp, vp, vp]),
        "mf_mg_get_info": (C.c_int, [vp, C.POINTER(MGInfo)]),
        "mf_mg_destroy": (C.c_int, [vp]),
        "mf_mg_transfer": (C.c_int, [vp, i32, i32, vp, vp, vp]),
        "mf_mg_level_dofs": (C.c_int, [vp, i32, vp, vp]),
        "mf_solve_cg": (C.c_int, [vp, vp, vp, vp, C.c_double, i32, i32, C.POINTER(SolveStats), vp]),
        "mf_solve_cg_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_double, i32, i32, C.POINTER(SolveStats), vp]),
        "mf_assemble_rhs": (C.c_int, [vp, i32, u64, i32, C.c_double, vp]),
        "mf_exact_solution": (C.c_int, [vp, vp]),
        "mf_ls_default_train_options": (None, [C.POINTER(TrainOptions)]),
        "mf_ls_default_fcg_options": (None, [C.POINTER(FcgOptions)]),
        "mf_ls_create": (C.c_int, [vp, u64, C.POINTER(vp)]),
        "mf_ls_destroy": (C.c_int, [vp]),
        "mf_ls_set_params": (C.c_int, [vp, vp, i64]),
        "mf_ls_get_params": (C.c_int, [vp, vp, i64]),
        "mf_ls_save": (C.c_int, [vp, vp, u64, C.POINTER(u64)]),
        "mf_ls_load": (C.c_int, [vp, vp, u64]),
        "mf_ls_apply": (C.c_int, [vp, vp, vp, vp]),
        "mf_ls_forward": (C.c_int, [vp, vp, i64, vp, vp]),
        "mf_ls_features": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "mf_ls_gradient": (C.c_int, [vp, vp, vp, i64, vp, C.POINTER(C.c_double), vp]),
        "mf_ls_train_step": (C.c_int, [vp, vp, vp, i64, C.c_double, C.POINTER(C.c_double), vp]),
        "mf_ls_train": (C.c_int, [vp, vp, C.POINTER(TrainOptions), C.POINTER(TrainStats), vp]),
        "mf_solve_fcg_learned": (C.c_int, [vp, vp, vp, vp, vp, C.c_double, i32, C.POINTER(FcgOptions),
                                           C.POINTER(FcgStats), vp]),
        "mf_get_quadrature": (C.c_int, [vp, vp, vp]),
        "mf_get_face_quadrature": (C.c_int, [vp, C.POINTER(C.c_int32), vp, vp]),
        "mf_mesh_read_vtk": (C.c_int, [C.c_char_p, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), vp, vp]),
        "mf_get_system_matrix": (C.c_int, [vp, C.POINTER(vp)]),
        "mf_csr_vmult": (C.c_int, [vp, vp, vp, vp]),
        "mf_csr_get_info": (C.c_int, [vp, C.POINTER(CsrInfo)]),
        "mf_csr_copy_to_host": (C.c_int, [vp, vp, vp, vp]),
        "mf_csr_destroy": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class MGOptions(C.Structure):
    _fields_ = [("smoother_degree", C.c_int32), ("smoothing_range", C.c_double),
                ("eig_cg_iterations", C.c_int32), ("level_precision", C.c_int32),
                ("max_coarse_dofs", C.c_int32), ("fuse_steps", C.c_int32), ("coarse_solver", C.c_int32),
                ("amg_threshold", C.c_double)]


MF_MAX_LEVELS = 16


class MGInfo(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("degree", C.c_int32 * MF_MAX_LEVELS),
                ("kuhn_n", C.c_int32 * MF_MAX_LEVELS), ("n_dofs", C.c_int64 * MF_MAX_LEVELS),
                ("lambda_max", C.c_double * MF_MAX_LEVELS), ("vcycles", C.c_int64),
                ("level_kind", C.c_int32 * MF_MAX_LEVELS), ("nitsche", C.c_int32 * MF_MAX_LEVELS),
                ("amg_levels", C.c_int32), ("amg_rows", C.c_int64 * MF_MAX_LEVELS)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("rhs_norm", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TrainOptions(C.Structure):
    _fields_ = [("n_rhs", C.c_int32), ("pairs_per_rhs", C.c_int32), ("steps_per_pair", C.c_int32),
                ("lr", C.c_double), ("n_features", C.c_int32), ("corr_length", C.c_double),
                ("seed", C.c_uint64), ("loss", C.c_int32), ("trajectory", C.c_int32)]


class TrainStats(C.Structure):
    _fields_ = [("pairs", C.c_int32), ("steps", C.c_int32), ("first_loss", C.c_double),
                ("last_loss", C.c_double), ("zero_loss", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FcgOptions(C.Structure):
    _fields_ = [("stall_ratio", C.c_double), ("force_fallback_at", C.c_int32), ("use_learned", C.c_int32)]


class FcgStats(C.Structure):
    _fields_ = [("base", SolveStats), ("fallback", C.c_int32), ("fallback_iteration", C.c_int32),
                ("learned_steps", C.c_int32), ("mean_omega", C.c_double)]

    def as_dict(self):
        d = self.base.as_dict()
        d.update({k: getattr(self, k) for k, _ in self._fields_[1:]})
        return d


LS_N_INPUTS, LS_N_OUTPUTS, LS_N_PARAMS = 20, 10, 6154


class MFError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise MFError(lib().mf_last_error().decode())


def exported_symbols():
    """Names every caller may bind (for the ABI completeness test)."""
    lib()
    return sorted(_SIGNATURE_NAMES)


_SIGNATURE_NAMES = [
    "mf_mesh_kuhn_cube", "mf_default_options", "mf_create", "mf_destroy", "mf_vmult",
    "mf_tvmult", "mf_compute_diagonal", "mf_compute_inverse_diagonal", "mf_m", "mf_get_info",
    "mf_get_cell_order", "mf_get_cell_dofs", "mf_get_cell_vertices", "mf_get_dof_points", "mf_get_constrained",
    "mf_debug_host_apply", "mf_last_error", "mf_device_count", "mf_device_name", "mf_malloc",
    "mf_free", "mf_memcpy_htod", "mf_memcpy_dtoh", "mf_memcpy_dtod", "mf_copy_kernel", "mf_memset",
    "mf_synchronize", "mf_stream_create", "mf_stream_destroy", "mf_stream_synchronize",
    "mf_event_create", "mf_event_destroy", "mf_event_record", "mf_event_elapsed_ms",
    "mf_mg_default_options", "mf_mg_create", "mf_mg_vmult", "mf_mg_get_info", "mf_mg_destroy",
    "mf_mg_transfer", "mf_mg_level_dofs", "mf_solve_cg", "mf_solve_cg_noalloc", "mf_assemble_rhs", "mf_exact_solution",
    "mf_ls_default_train_options", "mf_ls_default_fcg_options", "mf_ls_create", "mf_ls_destroy",
    "mf_ls_set_params", "mf_ls_get_params", "mf_ls_save", "mf_ls_load", "mf_ls_apply", "mf_ls_forward",
    "mf_ls_features", "mf_ls_gradient", "mf_ls_train_step", "mf_ls_train", "mf_solve_fcg_learned",
    "mf_get_quadrature", "mf_get_face_quadrature", "mf_mesh_read_vtk", "mf_get_system_matrix", "mf_csr_vmult", "mf_csr_get_info", "mf_csr_copy_to_host", "mf_csr_destroy",
]


# --------------------------------------------------------------------------
# meshes

class Mesh:
    """Tetrahedral mesh on the host (vertices (nv,3) float64, cells (nc,4) int32)."""

    def __init__(self, vertices, cells, kuhn_n=0):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.kuhn_n = int(kuhn_n)

    @staticmethod
    def kuhn_cube(n, perturb=0.2, seed=0):
        """Unit cube, n^3 hexes x 6 Kuhn tets, interior vertices perturbed by
        perturb*h*U[-1,1) per coordinate (splitmix64 seeded)."""
        v = np.empty(((n + 1) ** 3, 3), np.float64)
        c = np.empty((6 * n ** 3, 4), np.int32)
        check(lib().mf_mesh_kuhn_cube(n, perturb, seed, v.ctypes.data, c.ctypes.data))
        return Mesh(v, c, kuhn_n=n)

    @staticmethod
    def read_vtk(path):
        """Tetrahedral mesh from a legacy ASCII VTK unstructured grid (the
        external-mesh input of CG/cells_batched/application.cc:1699-1756).
        Returns the mesh; .n_flipped counts cells whose local vertices 1 and 2
        were swapped to make det J > 0."""
        L = lib()
        nv, nc, nf = C.c_int64(), C.c_int64(), C.c_int64()
        p = os.fsencode(path)
        check(L.mf_mesh_read_vtk(p, C.byref(nv), C.byref(nc), C.byref(nf), None, None))
        v = np.empty((nv.value, 3), np.float64)
        c = np.empty((nc.value, 4), np.int32)
        check(L.mf_mesh_read_vtk(p, C.byref(nv), C.byref(nc), C.byref(nf), v.ctypes.data, c.ctypes.data))
        m = Mesh(v, c)
        m.n_flipped = nf.value
        return m

    def write_vtk(self, path, layout="classic"):
        """Writes the mesh as a legacy ASCII VTK unstructured grid of
        tetrahedra; layout "classic" (version 3.0 CELLS) or "vtk5"
        (OFFSETS/CONNECTIVITY)."""
        nv, nc = self.vertices.shape[0], self.cells.shape[0]
        with open(path, "w") as f:
            f.write("# vtk DataFile Version %s\nsimplexmf mesh\nASCII\nDATASET UNSTRUCTURED_GRID\n"
                    % ("3.0" if layout == "classic" else "5.1"))
            f.write(f"POINTS {nv} double\n")
            np.savetxt(f, self.vertices, fmt="%.17g")
            if layout == "classic":
                f.write(f"CELLS {nc} {5 * nc}\n")
                np.savetxt(f, np.hstack([np.full((nc, 1), 4, np.int64), self.cells]), fmt="%d")
            elif layout == "vtk5":
                f.write(f"CELLS {nc + 1} {4 * nc}\nOFFSETS vtktypeint64\n")
                np.savetxt(f, 4 * np.arange(nc + 1, dtype=np.int64)[None, :], fmt="%d")
                f.write("CONNECTIVITY vtktypeint64\n")
                np.savetxt(f, self.cells.reshape(1, -1), fmt="%d")
            else:
                raise ValueError("layout must be 'classic' or 'vtk5'")
            f.write(f"CELL_TYPES {nc}\n")
            np.savetxt(f, np.full((1, nc), 10, np.int64), fmt="%d")

    def scrambled(self, seed=0):
        """The same mesh with vertices and cells in a seeded random order and
        each cell's vertices rotated (an unstructured input order; no
        lattice, so only p-multigrid levels are available)."""
        rng = np.random.default_rng(seed)
        vp = rng.permutation(self.vertices.shape[0])
        inv = np.empty_like(vp)
        inv[vp] = np.arange(vp.size)
        cells = inv[self.cells[rng.permutation(self.cells.shape[0])]]
        # even permutations of the local vertices keep the orientation
        rot = np.array([[0, 1, 2, 3], [1, 2, 0, 3], [2, 0, 1, 3], [0, 2, 3, 1]])
        cells = np.take_along_axis(cells, rot[rng.integers(0, len(rot), cells.shape[0])], axis=1)
        return Mesh(self.vertices[vp], cells.astype(np.int32))

    def desc(self):
        d = MeshDesc()
        d.n_vertices = self.vertices.shape[0]
        d.vertices = self.vertices.ctypes.data_as(C.POINTER(C.c_double))
        d.n_cells = self.cells.shape[0]
        d.cells = self.cells.ctypes.data_as(C.POINTER(C.c_int32))
        d.kuhn_n = self.kuhn_n
        return d


# --------------------------------------------------------------------------
# device memory

class Stream:
    def __init__(self):
        h = C.c_void_p()
        check(lib().mf_stream_create(C.byref(h)))
        self.handle = h.value

    def synchronize(self):
        check(lib().mf_stream_synchronize(self.handle))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.mf_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(lib().mf_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        check(lib().mf_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, stop):
        ms = C.c_float()
        check(lib().mf_event_elapsed_ms(self.handle, stop.handle, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.mf_event_destroy(self.handle)
            self.handle = None


class DeviceVector:
    """A device array of n elements of dtype (float64/float32/...)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().mf_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a)
        v = DeviceVector(a.size, a.dtype)
        check(lib().mf_memcpy_htod(v.ptr, a.ctypes.data, v.nbytes))
        return v

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().mf_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        check(lib().mf_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self, stream=None):
        check(lib().mf_memset(self.ptr, 0, self.nbytes, stream.handle if stream else None))

    def free(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.mf_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()


def synchronize():
    check(lib().mf_synchronize())


def device_count():
    n = C.c_int32()
    if lib().mf_device_count(C.byref(n)) != 0:
        return 0
    return n.value


def device_name():
    buf = C.create_string_buffer(256)
    check(lib().mf_device_name(buf, 256))
    return buf.value.decode()


# --------------------------------------------------------------------------
# operator

class Operator:
    """Matrix-free operator; construction plays the role of ``reinit``."""

    def __init__(self, mesh: Mesh, degree: int, kind="cg", precision="f64", reorder=True,
                 patch_cells=0, mass_coefficient=1.0, laplace_coefficient=1.0,
                 penalty_kind=0, host_only=False, accumulate=0, dg_kernel=0, boundary="dirichlet",
                 cells_per_wave=0, manifold="none", manifold_center=None):
        L = lib()
        opt = Options()
        L.mf_default_options(C.byref(opt))
        opt.reorder = 1 if reorder else 0
        opt.patch_cells = patch_cells
        opt.mass_coefficient = mass_coefficient
        opt.laplace_coefficient = laplace_coefficient
        opt.penalty_kind = penalty_kind
        opt.host_only = 1 if host_only else 0
        opt.accumulate = accumulate
        opt.dg_kernel = dg_kernel
        opt.boundary = {"dirichlet": 0, "nitsche": 1}[boundary] if isinstance(boundary, str) else int(boundary)
        opt.cells_per_wave = cells_per_wave
        opt.manifold = {"none": 0, "spherical": 1, "straight": 2}[manifold] if isinstance(manifold, str) \
            else int(manifold)
        if manifold_center is not None:
            for d in range(3):
                opt.manifold_center[d] = float(manifold_center[d])
        self.mesh = mesh
        self.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = C.c_void_p()
        d = mesh.desc()
        check(L.mf_create(C.byref(d), degree, self.kind, self.precision, C.byref(opt), C.byref(h)))
        self.handle = h.value
        self.dtype = NP_DTYPES[self.precision]
        inf = Info()
        check(L.mf_get_info(self.handle, C.byref(inf)))
        self.info = inf.as_dict()

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.mf_destroy(self.handle)
            self.handle = None

    # -- reference surface --
    def m(self):
        return lib().mf_m(self.handle)

    def initialize_dof_vector(self):
        v = DeviceVector(self.m(), self.dtype)
        v.zero()
        return v

    def vmult(self, dst: DeviceVector, src: DeviceVector, stream: Stream | None = None):
        self._check_vec(dst)
        self._check_vec(src)
        check(lib().mf_vmult(self.handle, dst.ptr, src.ptr, stream.handle if stream else None))

    def Tvmult(self, dst, src, stream=None):
        self._check_vec(dst)
        self._check_vec(src)
        check(lib().mf_tvmult(self.handle, dst.ptr, src.ptr, stream.handle if stream else None))

    def compute_diagonal(self, diag: DeviceVector, stream=None):
        self._check_vec(diag)
        check(lib().mf_compute_diagonal(self.handle, diag.ptr, stream.handle if stream else None))

    def compute_inverse_diagonal(self, inv: DeviceVector, stream=None):
        self._check_vec(inv)
        check(lib().mf_compute_inverse_diagonal(self.handle, inv.ptr, stream.handle if stream else None))

    def _check_vec(self, v):
        if v.n != self.m() or v.dtype != self.dtype:
            raise ValueError(f"vector of {v.n} {v.dtype} does not match operator ({self.m()} {self.dtype})")

    # -- numbering queries (host) --
    def cell_order(self):
        out = np.empty(self.info["n_cells"], np.int64)
        check(lib().mf_get_cell_order(self.handle, out.ctypes.data))
        return out

    def cell_vertices(self):
        out = np.empty((self.info["n_cells"], 4), np.int32)
        check(lib().mf_get_cell_vertices(self.handle, out.ctypes.data))
        return out

    def is_dg(self):
        return self.kind in (MF_DG_SIPG, MF_DG_VECTOR, MF_DG_HELMHOLTZ)

    def cell_dofs(self):
        w = self.info["dofs_per_cell"] * (self.info["n_components"] if self.is_dg() else 1)
        out = np.empty((self.info["n_cells"], w), np.int64)
        check(lib().mf_get_cell_dofs(self.handle, out.ctypes.data))
        return out

    def n_scalar_dofs(self):
        if self.is_dg():
            return self.info["n_cells"] * self.info["dofs_per_cell"]
        return self.info["n_dofs"] // max(1, self.info["n_components"])

    def dof_points(self):
        out = np.empty((self.n_scalar_dofs(), 3), np.float64)
        check(lib().mf_get_dof_points(self.handle, out.ctypes.data))
        return out

    def constrained(self):
        out = np.empty(self.n_scalar_dofs(), np.uint8)
        check(lib().mf_get_constrained(self.handle, out.ctypes.data))
        return out

    def quadrature(self):
        """(points (n_q, 3), weights (n_q,)) of the cell rule"""
        nq = self.info["n_q_points"]
        pts, w = np.empty((nq, 3)), np.empty(nq)
        check(lib().mf_get_quadrature(self.handle, pts.ctypes.data, w.ctypes.data))
        return pts, w

    def face_quadrature(self):
        """(points (n, 2) as (s, t) on the unit triangle, weights (n,)) of the
        face rule (DG operators); on a face with local vertices a < b < c the
        point is x_a + s (x_b - x_a) + t (x_c - x_a)"""
        n = C.c_int32()
        check(lib().mf_get_face_quadrature(self.handle, C.byref(n), None, None))
        pts, w = np.empty((n.value, 2)), np.empty(n.value)
        check(lib().mf_get_face_quadrature(self.handle, C.byref(n), pts.ctypes.data, w.ctypes.data))
        return pts, w

    def host_apply(self, src=None, mode=0):
        """Host execution of the kernels' schedule (table self-check, no GPU)."""
        dst = np.zeros(self.m(), self.dtype)
        s = None if src is None else np.ascontiguousarray(src, dtype=self.dtype)
        check(lib().mf_debug_host_apply(self.handle, mode, dst.ctypes.data,
                                         None if s is None else s.ctypes.data))
        return dst


class SystemMatrix:
    """Assembled sparse matrix of a CG operator (``get_system_matrix``,
    CG/cells_batched/application_hp.cc:1411-1438) with a device CSR product
    (the matrix-based timing of CG/cells_batched/application.cc:1944-1988)."""

    def __init__(self, op: Operator):
        h = C.c_void_p()
        check(lib().mf_get_system_matrix(op.handle, C.byref(h)))
        self.handle = h.value
        self.op = op
        self.dtype = op.dtype
        inf = CsrInfo()
        check(lib().mf_csr_get_info(self.handle, C.byref(inf)))
        self.info = inf.as_dict()

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.mf_csr_destroy(self.handle)
            self.handle = None

    def m(self):
        return self.info["n_rows"]

    def vmult(self, dst: DeviceVector, src: DeviceVector, stream: Stream | None = None):
        for v in (dst, src):
            if v.n != self.m() or v.dtype != self.dtype:
                raise ValueError(f"vector of {v.n} {v.dtype} does not match the matrix ({self.m()} {self.dtype})")
        check(lib().mf_csr_vmult(self.handle, dst.ptr, src.ptr, stream.handle if stream else None))

    def to_host(self):
        """(ptr int64[n+1], col int32[nnz], val float64[nnz])"""
        n, nnz = self.info["n_rows"], self.info["nnz"]
        ptr = np.empty(n + 1, np.int64)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float64)
        check(lib().mf_csr_copy_to_host(self.handle, ptr.ctypes.data, col.ctypes.data, val.ctypes.data))
        return ptr, col[:nnz], val[:nnz]


# --------------------------------------------------------------------------
# multigrid and solvers

class Multigrid:
    """Hybrid p/h multigrid preconditioner below a CG operator
    (MultigridPreconditioner of CG/cells_batched/application_ph.cc:1754-2026)."""

    def __init__(self, fine: Operator, smoother_degree=5, smoothing_range=20.0, eig_cg_iterations=10,
                 level_precision="f32", max_coarse_dofs=4096, fuse_steps=True, coarse_solver="auto",
                 amg_threshold=None):
        L = lib()
        opt = MGOptions()
        L.mf_mg_default_options(C.byref(opt))
        opt.smoother_degree = smoother_degree
        opt.smoothing_range = smoothing_range
        opt.eig_cg_iterations = eig_cg_iterations
        opt.level_precision = PRECISIONS[level_precision] if isinstance(level_precision, str) else level_precision
        opt.max_coarse_dofs = max_coarse_dofs
        opt.fuse_steps = 1 if fuse_steps else 0
        opt.coarse_solver = {"auto": 0, "dense": 1, "amg": 2}[coarse_solver] if isinstance(coarse_solver, str) \
            else int(coarse_solver)
        if amg_threshold is not None:
            opt.amg_threshold = float(amg_threshold)
        self.level_dtype = NP_DTYPES[opt.level_precision]
        self.fine = fine
        h = C.c_void_p()
        check(L.mf_mg_create(fine.handle, C.byref(opt), C.byref(h)))
        self.handle = h.value

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.mf_mg_destroy(self.handle)
            self.handle = None

    def info(self):
        inf = MGInfo()
        check(lib().mf_mg_get_info(self.handle, C.byref(inf)))
        n = inf.n_levels
        return {"n_levels": n, "degree": list(inf.degree[:n]), "kuhn_n": list(inf.kuhn_n[:n]),
                "n_dofs": list(inf.n_dofs[:n]), "lambda_max": list(inf.lambda_max[:n]),
                "vcycles": inf.vcycles, "level_kind": list(inf.level_kind[:n]), "nitsche": list(inf.nitsche[:n]),
                "amg_rows": list(inf.amg_rows[:inf.amg_levels])}

    def vmult(self, dst, src, stream=None):
        check(lib().mf_mg_vmult(self.handle, dst.ptr, src.ptr, stream.handle if stream else None))

    def transfer(self, level, direction, dst, src, stream=None):
        check(lib().mf_mg_transfer(self.handle, level, direction, dst.ptr, src.ptr,
                                   stream.handle if stream else None))

    def level_dofs(self, level):
        n = self.info()["n_dofs"][level]
        xyz = np.empty((n, 3))
        con = np.empty(n, np.uint8)
        check(lib().mf_mg_level_dofs(self.handle, level, xyz.ctypes.data, con.ctypes.data))
        return xyz, con


def solve_cg(A: Operator, x: DeviceVector, b: DeviceVector, mg: Multigrid | None = None, rel_tol=1e-10,
             max_iter=1000, flexible=False, stream=None):
    st = SolveStats()
    check(lib().mf_solve_cg(A.handle, mg.handle if mg is not None else None, x.ptr, b.ptr, rel_tol,
                            max_iter, 1 if flexible else 0, C.byref(st), stream.handle if stream else None))
    return st.as_dict()


def solve_cg_noalloc(A: Operator, x: DeviceVector, b: DeviceVector, mg: Multigrid | None = None, rel_tol=1e-10,
                     max_iter=1000, flexible=False, stream=None):
    """solve_cg with the work memory in the library's static device array
    (no allocation; operators up to 1048576 rows; calls on one GPU must not
    overlap)"""
    st = SolveStats()
    check(lib().mf_solve_cg_noalloc(A.handle, mg.handle if mg is not None else None, x.ptr, b.ptr, rel_tol,
                                    max_iter, 1 if flexible else 0, C.byref(st), stream.handle if stream else None))
    return st.as_dict()


def assemble_rhs(op: Operator, kind="sine", seed=0, n_features=256, corr_length=0.1):
    b = np.empty(op.m())
    k = {"sine": 0, "grf": 1}[kind] if isinstance(kind, str) else int(kind)
    check(lib().mf_assemble_rhs(op.handle, k, seed, n_features, corr_length, b.ctypes.data))
    return b


def exact_solution(op: Operator):
    u = np.empty(op.m())
    check(lib().mf_exact_solution(op.handle, u.ctypes.data))
    return u


# --------------------------------------------------------------------------
# learned per-cell smoother and flexible CG

class LearnedSmoother:
    """Per-cell MLP smoother of a CG P2 FP64 Poisson operator (device weights)."""

    def __init__(self, fine: Operator, seed=0):
        self._lib = lib()
        self.op = fine
        h = C.c_void_p()
        check(self._lib.mf_ls_create(fine.handle, seed, C.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.mf_ls_destroy(self.handle)
            self.handle = None

    @property
    def n_cells(self):
        return self.op.info["n_cells"]

    def params(self):
        p = np.empty(LS_N_PARAMS, np.float32)
        check(self._lib.mf_ls_get_params(self.handle, p.ctypes.data, LS_N_PARAMS))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(self._lib.mf_ls_set_params(self.handle, p.ctypes.data, p.size))

    def save(self) -> bytes:
        n = C.c_uint64()
        check(self._lib.mf_ls_save(self.handle, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self._lib.mf_ls_save(self.handle, buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    def load(self, blob: bytes):
        check(self._lib.mf_ls_load(self.handle, blob, len(blob)))

    def apply(self, dst, src, stream=None):
        check(self._lib.mf_ls_apply(self.handle, dst.ptr, src.ptr, stream.handle if stream else None))

    def forward(self, x: DeviceVector, n, y: DeviceVector, stream=None):
        check(self._lib.mf_ls_forward(self.handle, x.ptr, n, y.ptr, stream.handle if stream else None))

    def features(self, r, e=None, stream=None):
        nc = self.n_cells
        x = DeviceVector(nc * LS_N_INPUTS, np.float32)
        t = DeviceVector(nc * LS_N_OUTPUTS, np.float32) if e is not None else None
        check(self._lib.mf_ls_features(self.handle, r.ptr, e.ptr if e is not None else None, x.ptr,
                                       t.ptr if t is not None else None, stream.handle if stream else None))
        return x, t

    def gradient(self, x: DeviceVector, t: DeviceVector, n, stream=None):
        g = np.empty(LS_N_PARAMS, np.float32)
        loss = C.c_double()
        check(self._lib.mf_ls_gradient(self.handle, x.ptr, t.ptr, n, g.ctypes.data, C.byref(loss),
                                       stream.handle if stream else None))
        return loss.value, g

    def train_step(self, x, t, n, lr=1e-3, stream=None):
        loss = C.c_double()
        check(self._lib.mf_ls_train_step(self.handle, x.ptr, t.ptr, n, lr, C.byref(loss),
                                         stream.handle if stream else None))
        return loss.value

    def train(self, mg: Multigrid, stream=None, **kw):
        o = TrainOptions()
        self._lib.mf_ls_default_train_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        st = TrainStats()
        check(self._lib.mf_ls_train(self.handle, mg.handle, C.byref(o), C.byref(st),
                                    stream.handle if stream else None))
        return st.as_dict()


def solve_fcg_learned(A: Operator, x: DeviceVector, b: DeviceVector, mg: Multigrid,
                      ls: LearnedSmoother | None = None, rel_tol=1e-10, max_iter=1000, stream=None, **kw):
    o = FcgOptions()
    lib().mf_ls_default_fcg_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    st = FcgStats()
    check(lib().mf_solve_fcg_learned(A.handle, mg.handle, ls.handle if ls is not None else None, x.ptr, b.ptr,
                                     rel_tol, max_iter, C.byref(o), C.byref(st),
                                     stream.handle if stream else None))
    return st.as_dict()
