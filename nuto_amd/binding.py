"""ctypes binding of libnuto_amd.so (the C ABI in include/nuto_amd.h).

Device arrays are torch tensors on cuda:0 (float64 / int32); host arrays are numpy. There is no
CPU fallback: if the HIP library is missing or no GPU is present the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnuto_amd.so")

QUAD4, HEX8, TET4, TET10, HEX20 = 0, 1, 2, 3, 4
NUM_NODES = {QUAD4: 4, HEX8: 8, TET4: 4, TET10: 10, HEX20: 20}
DIM = {QUAD4: 2, HEX8: 3, TET4: 3, TET10: 3, HEX20: 3}

MOMENTUM_LINEAR_ELASTIC, GRADIENT_DAMAGE, MOMENTUM_LOCAL_DAMAGE = 0, 1, 2
PLANE_STRESS, PLANE_STRAIN = 0, 1
FULL, UNILATERAL = 0, 1
CONSTANT, DECREASING = 0, 1  # NonlocalInteraction of gradient damage
# methods of Context.solve (nuto_solve): EigenConjugateGradient, EigenBiCGSTAB
SOLVER_CG, SOLVER_BICGSTAB = 0, 1
NO_CONVERGENCE = 6  # NutoError.code of nuto_quasistatic_step without convergence (NewtonRaphson::NoConvergence)
# fields of Context.eval_ip (nuto_eval_ip)
IP_STRAIN, IP_STRESS, IP_KAPPA, IP_DAMAGE, IP_EQ_STRAIN = 0, 1, 2, 3, 4

_ERRORS = {1: "argument", 2: "state", 3: "device", 4: "memory", 5: "unsupported", 6: "no convergence"}


class NutoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nuto_amd error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class SolverOptions(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("relative_residual", ctypes.c_double),
                ("independent_dofs", ctypes.c_int)]


class NewtonOptions(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("max_line_search_steps", ctypes.c_int), ("linear", SolverOptions)]


class NewtonInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("linear_iterations", ctypes.c_int),
                ("residual_evaluations", ctypes.c_int), ("residual_norm", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("E", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("plane_state", ctypes.c_int),
        ("k", ctypes.c_double),
        ("kappa0", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("c", ctypes.c_double),
        ("damage_application", ctypes.c_int),
        ("interaction", ctypes.c_int),
        ("interaction_r", ctypes.c_double),
        ("interaction_eta", ctypes.c_double),
    ]


def gradient_damage_params(E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0,
                           damage_application=FULL, interaction=CONSTANT, interaction_r=0.005, interaction_eta=5.0):
    """GradientDamage<3> parameters; interaction DECREASING selects NonlocalInteraction::Decreasing(R, eta)."""
    return Params(kind=GRADIENT_DAMAGE, dim=3, E=E, nu=nu, plane_state=0, k=k, kappa0=kappa0, beta=beta,
                  alpha=alpha, c=c, damage_application=damage_application, interaction=interaction,
                  interaction_r=interaction_r, interaction_eta=interaction_eta)


def linear_elastic_params(dim, E, nu, plane_state=PLANE_STRESS):
    return Params(kind=MOMENTUM_LINEAR_ELASTIC, dim=dim, E=E, nu=nu, plane_state=plane_state)


def local_damage_params(E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99,
                        damage_application=FULL):
    return Params(kind=MOMENTUM_LOCAL_DAMAGE, dim=3, E=E, nu=nu, plane_state=0, k=k, kappa0=kappa0, beta=beta,
                  alpha=alpha, c=0.0, damage_application=damage_application)


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch brings its own libamdhip64 (SONAME libamdhip64.so.7). Loading it first makes the
        # dynamic linker bind libnuto_amd.so to that same runtime instead of a second copy from
        # /opt/rocm, so both share one HIP runtime, device context and stream namespace.
        import torch  # noqa: F401

        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                              " (make -C nuto_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.nuto_last_error.restype = ctypes.c_char_p
        L.nuto_last_error.argtypes = [vp]
        L.nuto_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.nuto_destroy.argtypes = [vp]
        L.nuto_set_stream.argtypes = [vp, vp]
        L.nuto_synchronize.argtypes = [vp]
        L.nuto_mesh_structured.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, ip,
                                           ip, vp, vp]
        L.nuto_dof_layer.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ip]
        L.nuto_set_cells.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.nuto_set_dof_type.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.nuto_set_integrand.argtypes = [vp, ctypes.POINTER(Params)]
        L.nuto_set_rule.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.nuto_set_history_ld.argtypes = [vp, ctypes.c_int]
        L.nuto_set_material_fields.argtypes = [vp, vp, vp, vp]
        L.nuto_set_material_fields_device.argtypes = [vp, vp, vp, vp]
        L.nuto_vjp.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.nuto_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.nuto_batch_destroy.argtypes = [vp]
        L.nuto_batch_last_error.restype = ctypes.c_char_p
        L.nuto_batch_last_error.argtypes = [vp]
        L.nuto_batch_set_stream.argtypes = [vp, vp]
        L.nuto_batch_synchronize.argtypes = [vp]
        L.nuto_batch_set_integrand.argtypes = [vp, ctypes.POINTER(Params)]
        L.nuto_batch_add_mesh.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp,
                                          ctypes.c_int, vp, ip]
        L.nuto_batch_set_group_limit.argtypes = [vp, ctypes.c_longlong]
        L.nuto_batch_set_stage_bytes.argtypes = [vp, ctypes.c_longlong]
        L.nuto_batch_build.argtypes = [vp]
        L.nuto_batch_sizes.argtypes = [vp, vp]
        L.nuto_batch_mesh_info.argtypes = [vp, ctypes.c_int, vp]
        L.nuto_batch_group_info.argtypes = [vp, ctypes.c_int, vp]
        L.nuto_batch_copy_pattern.argtypes = [vp, ctypes.c_int, vp, vp]
        L.nuto_batch_set_material_fields.argtypes = [vp, vp, vp, vp]
        L.nuto_batch_assemble.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int]
        L.nuto_batch_autotune.argtypes = [vp, vp, vp, vp, vp, vp, vp, ip]
        L.nuto_batch_vjp.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.nuto_batch_residual_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.nuto_build_pattern.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ip]
        L.nuto_copy_pattern.argtypes = [vp, vp, vp]
        L.nuto_block_pattern.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), vp, vp, vp]
        L.nuto_block_values.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.nuto_pattern_device.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        for f in ("nuto_build_vector", "nuto_build_matrix"):
            getattr(L, f).argtypes = [vp, vp, vp, vp, vp]
        L.nuto_update_history.argtypes = [vp, vp, vp, vp]
        L.nuto_build_lumped_mass.argtypes = [vp, ctypes.c_double, vp]
        L.nuto_eval_ip.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
        L.nuto_set_constraints.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.nuto_solve.argtypes = [vp, vp, vp, vp, ctypes.POINTER(SolverOptions), ctypes.POINTER(SolveInfo)]
        L.nuto_quasistatic_step.argtypes = [vp, vp, vp, vp, vp, ctypes.POINTER(NewtonOptions),
                                            ctypes.POINTER(NewtonInfo)]
        L.nuto_solve_trial_state.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.POINTER(SolverOptions),
                                              ctypes.POINTER(SolveInfo)]
        L.nuto_build_vector_matrix.argtypes = [vp, vp, vp, vp, vp, vp]
        L.nuto_build_vector_matrix_noalloc.argtypes = [vp, vp, vp, vp, vp, vp]
        L.nuto_device_alloc.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_ulonglong]
        L.nuto_device_free.argtypes = [vp, vp]
        L.nuto_copy_to_device.argtypes = [vp, vp, vp, ctypes.c_ulonglong]
        L.nuto_copy_to_host.argtypes = [vp, vp, vp, ctypes.c_ulonglong]
        gp = ctypes.c_void_p
        L.nuto_gmsh_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(gp)]
        L.nuto_gmsh_free.argtypes = [gp]
        L.nuto_gmsh_free.restype = None
        L.nuto_gmsh_last_error.argtypes = [gp]
        L.nuto_gmsh_last_error.restype = ctypes.c_char_p
        L.nuto_gmsh_sizes.argtypes = [gp, ip, ip, ip, ip]
        L.nuto_gmsh_nodes.argtypes = [gp, vp]
        L.nuto_gmsh_elements.argtypes = [gp, vp, vp, vp, vp]
        L.nuto_gmsh_group.argtypes = [gp, ctypes.c_int, ip, vp]
        L.nuto_gmsh_group_id.argtypes = [gp, ctypes.c_char_p, ip]
        L.nuto_gmsh_cells.argtypes = [gp, ctypes.c_int, ctypes.c_int, ip, vp]
        u = ctypes.c_uint
        L.nuto_morton_encode3d.argtypes = [u, u, u]
        L.nuto_morton_encode3d.restype = u
        L.nuto_morton_encode2d.argtypes = [u, u]
        L.nuto_morton_encode2d.restype = u
        L.nuto_morton_decode3d.argtypes = [u, ctypes.POINTER(u), ctypes.POINTER(u), ctypes.POINTER(u)]
        L.nuto_morton_decode3d.restype = None
        L.nuto_morton_decode2d.argtypes = [u, ctypes.POINTER(u), ctypes.POINTER(u)]
        L.nuto_morton_decode2d.restype = None
        L.nuto_morton_neighbor3d.argtypes = [u, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.nuto_morton_neighbor3d.restype = u
        L.nuto_morton_order.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        _lib = L
    return _lib


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _dev_ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _f64(t, n, name, optional=True):
    """t, checked to be a float64 tensor of at least n values (None stays None when optional). The C
    ABI takes bare pointers without lengths, so a float32 or short tensor is refused here, before a
    kernel could read or write past its end. Every argument of a call is checked this way before any
    is turned into a device pointer (_ptrs)."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: a tensor is required")
    dt = str(t.dtype)
    if dt != "torch.float64":
        raise ValueError(f"{name}: expected float64, got {dt}")
    if t.numel() < n:
        raise ValueError(f"{name}: {t.numel()} values, needs {n}")
    return t


def _ptrs(*ts):
    """Device pointers of checked tensors (contiguous, on the device), in order."""
    return tuple(_dev_ptr(t) for t in ts)


# ---- set-up helpers (host) --------------------------------------------------------------------
def structured_mesh(coord_type, nx, ny, nz=0, lengths=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """Structured mesh: returns (coords n_nodes x dim, conn n_cells x nodes)."""
    L = lib()
    nn, nc = ctypes.c_int(), ctypes.c_int()
    args = (coord_type, nx, ny, nz, float(lengths[0]), float(lengths[1]), float(lengths[2]), float(jitter), seed)
    r = L.nuto_mesh_structured(*args, ctypes.byref(nn), ctypes.byref(nc), None, None)
    if r:
        raise NutoError(r, "invalid structured mesh arguments")
    dim = DIM[coord_type]
    X = np.zeros((nn.value, dim))
    C = np.zeros((nc.value, NUM_NODES[coord_type]), np.int32)
    r = L.nuto_mesh_structured(*args, None, None, _np_ptr(X), _np_ptr(C))
    if r:
        raise NutoError(r, "structured mesh failed")
    return X, C


def dof_layer(coord_type, conn, dof_type):
    """First-touch dof node ranks per cell (n_cells x nodes(dof_type)) and the dof node count."""
    L = lib()
    conn = np.ascontiguousarray(conn, np.int32)
    n = ctypes.c_int()
    r = L.nuto_dof_layer(coord_type, conn.shape[0], _np_ptr(conn), dof_type, None, ctypes.byref(n))
    if r:
        raise NutoError(r, "dof layer failed")
    ranks = np.zeros((conn.shape[0], NUM_NODES[dof_type]), np.int32)
    r = L.nuto_dof_layer(coord_type, conn.shape[0], _np_ptr(conn), dof_type, _np_ptr(ranks), ctypes.byref(n))
    if r:
        raise NutoError(r, "dof layer failed")
    return ranks, n.value


def cell_dofs(ranks, ncomp):
    """Cell dof numbering (node-major, component-minor) from per-cell node ranks."""
    ranks = np.asarray(ranks, np.int32)
    return (ranks[:, :, None] * ncomp + np.arange(ncomp, dtype=np.int32)[None, None, :]).reshape(
        ranks.shape[0], -1).astype(np.int32)


# ---- mesh input and Morton codes (host) ---------------------------------------------------------
GMSH_CELL_TYPE = {QUAD4: 3, HEX8: 5, TET4: 4, TET10: 11}  # gmsh element type of each coordinate element


class GmshMesh:
    """A gmsh MSH 2 file read by nuto_gmsh_read (MeshGmsh, MeshGmsh.cpp:440-550).

    coords: n_nodes x dim (file order); elements: gmsh types, CSR offsets and 0-based node indices
    (node orders of the interpolations); physical: first tag of each element.
    Reading errors raise NutoError with MeshGmsh's message.
    """

    def __init__(self, path):
        L = lib()
        self._h = ctypes.c_void_p()
        r = L.nuto_gmsh_read(os.fsencode(path), ctypes.byref(self._h))
        if r:
            msg = L.nuto_gmsh_last_error(self._h).decode()
            L.nuto_gmsh_free(self._h)
            self._h = None
            raise NutoError(r, msg)
        dim, nn, ne, nconn = (ctypes.c_int() for _ in range(4))
        L.nuto_gmsh_sizes(self._h, *(ctypes.byref(v) for v in (dim, nn, ne, nconn)))
        self.dim = dim.value
        self.coords = np.zeros((nn.value, self.dim))
        L.nuto_gmsh_nodes(self._h, _np_ptr(self.coords))
        self.gmsh_type = np.zeros(ne.value, np.int32)
        self.offsets = np.zeros(ne.value + 1, np.int32)
        self.conn = np.zeros(nconn.value, np.int32)
        self.physical = np.zeros(ne.value, np.int32)
        L.nuto_gmsh_elements(self._h, *(_np_ptr(a) for a in (self.gmsh_type, self.offsets, self.conn, self.physical)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().nuto_gmsh_free(self._h)
            self._h = None

    def _check(self, r):
        if r:
            raise NutoError(r, lib().nuto_gmsh_last_error(self._h).decode())

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elements(self):
        return self.gmsh_type.shape[0]

    def element_nodes(self, i):
        return self.conn[self.offsets[i]:self.offsets[i + 1]]

    def physical_group(self, group):
        """Element indices of a physical group, by id or (case-insensitive) name
        (MeshGmsh::GetPhysicalGroup, MeshGmsh.cpp:446-461)."""
        if isinstance(group, str):
            gid = ctypes.c_int()
            self._check(lib().nuto_gmsh_group_id(self._h, group.encode(), ctypes.byref(gid)))
            group = gid.value
        n = ctypes.c_int()
        self._check(lib().nuto_gmsh_group(self._h, int(group), ctypes.byref(n), None))
        out = np.zeros(n.value, np.int32)
        self._check(lib().nuto_gmsh_group(self._h, int(group), ctypes.byref(n), _np_ptr(out)))
        return out

    def cells(self, coord_type, group=None):
        """(coords, conn) of the elements of coord_type (of a group, by id or name), ready for
        Context.set_cells (nuto_set_cells)."""
        if isinstance(group, str):
            gid = ctypes.c_int()
            self._check(lib().nuto_gmsh_group_id(self._h, group.encode(), ctypes.byref(gid)))
            group = gid.value
        g = -1 if group is None else int(group)
        n = ctypes.c_int()
        self._check(lib().nuto_gmsh_cells(self._h, g, coord_type, ctypes.byref(n), None))
        conn = np.zeros((n.value, NUM_NODES[coord_type]), np.int32)
        self._check(lib().nuto_gmsh_cells(self._h, g, coord_type, ctypes.byref(n), _np_ptr(conn)))
        return self.coords, conn


def morton_encode3d(x, y, z):
    """MortonOrder::EncodeMorton3D (MortonOrder.h:14-17): 10 bits per axis, x lowest."""
    return lib().nuto_morton_encode3d(x, y, z)


def morton_encode2d(x, y):
    """The 2D code whose DecodeMorton2X/Y (MortonOrder.h:56-64) are x and y: 16 bits per axis."""
    return lib().nuto_morton_encode2d(x, y)


def morton_decode3d(code):
    """MortonOrder::DecodeMorton3X/Y/Z (MortonOrder.h:66-79)."""
    v = [ctypes.c_uint() for _ in range(3)]
    lib().nuto_morton_decode3d(code, *(ctypes.byref(a) for a in v))
    return tuple(a.value for a in v)


def morton_decode2d(code):
    """MortonOrder::DecodeMorton2X/Y (MortonOrder.h:56-64)."""
    v = [ctypes.c_uint() for _ in range(2)]
    lib().nuto_morton_decode2d(code, *(ctypes.byref(a) for a in v))
    return tuple(a.value for a in v)


def morton_neighbor3d(key, dx, dy, dz):
    """MortonOrder::NegNeighbor3D (MortonOrder.h:91-103); 0xffffffff below zero."""
    return lib().nuto_morton_neighbor3d(key, dx, dy, dz)


def morton_order(coords):
    """Permutation sorting points (n x dim) along a Morton curve over their bounding box."""
    X = np.ascontiguousarray(coords, np.float64)
    if X.ndim == 1:
        X = X[:, None]
    order = np.zeros(X.shape[0], np.int32)
    r = lib().nuto_morton_order(X.shape[0], X.shape[1], _np_ptr(X), _np_ptr(order))
    if r:
        raise NutoError(r, "invalid points")
    return order


def adaptive_solve(do_step, t_end, t_start=0.0, post_process=None, dt=0.1, dt_max=0.1, dt_min=1e-6,
                   increase_factor=1.5, decrease_factor=0.5):
    """AdaptiveSolve::Solve (AdaptiveSolve.cpp:18-67): calls do_step(t + dt), which returns the Newton
    iterations (Context.quasistatic_step's "iterations") or raises NutoError with code NO_CONVERGENCE;
    on failure dt is halved (below dt_min: an error), after a step with fewer than 3 iterations dt
    grows by increase_factor up to dt_max, and the last step ends exactly at t_end. post_process(t)
    runs at t_start and after every converged step. Returns the list of converged times."""
    t = t_start
    times = []
    if post_process:
        post_process(t)
    while True:
        try:
            iterations = do_step(t + dt)
        except NutoError as e:
            if e.code != NO_CONVERGENCE:
                raise
            dt *= decrease_factor
            if dt < dt_min:
                raise NutoError(NO_CONVERGENCE, f"Time step smaller than prescribed minimal time step {dt_min}.")
            continue
        t += dt
        times.append(t)
        if post_process:
            post_process(t)
        if t >= t_end:
            break
        if iterations < 3 and dt < dt_max:
            dt = min(dt * increase_factor, dt_max)
        if t + dt > t_end:
            dt = t_end - t
    return times


class Context:
    """One assembly context on one device (the C ABI nuto_ctx)."""

    def __init__(self, device=0, use_torch_stream=True):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("nuto_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self._torch = torch
        L = lib()
        self._L = L  # kept for close() during interpreter shutdown
        h = ctypes.c_void_p()
        r = L.nuto_create(ctypes.byref(h), device)
        if r:
            raise NutoError(r, "nuto_create failed")
        self.h = h
        self.device = device
        if use_torch_stream:
            self.set_stream(torch.cuda.current_stream(device).cuda_stream)
        self._init_sizes()

    def _init_sizes(self):
        self.n_rows = 0
        self.nnz = 0
        self.n_cells = 0
        self.n_ip = 0
        self._n_dofs = {}
        self._n_eq = {}
        self._history_ld = None

    # lengths of the device arrays the C ABI reads or writes (checked before each call)
    def _n_u(self):
        return self._n_dofs.get(0, 0)

    def _n_e(self):
        return self._n_dofs.get(1, 0)

    def _n_hist(self):
        return (self.n_cells if self._history_ld is None else self._history_ld) * self.n_ip

    def _n_rhs(self):
        return sum(self._n_eq.values())

    def _state(self, u, e, kappa):
        return (_f64(u, self._n_u(), "u", optional=False), _f64(e, self._n_e(), "e"),
                _f64(kappa, self._n_hist(), "kappa"))

    def close(self):
        h, L = getattr(self, "h", None), getattr(self, "_L", None)
        if h and L is not None:
            L.nuto_destroy(h)
        self.h = None

    __del__ = close

    def _chk(self, r):
        if r:
            raise NutoError(r, lib().nuto_last_error(self.h).decode())

    def set_stream(self, stream):
        # 0 / None selects the null stream (torch's default stream)
        self._chk(lib().nuto_set_stream(self.h, ctypes.c_void_p(stream) if stream else None))

    def synchronize(self):
        self._chk(lib().nuto_synchronize(self.h))

    def set_cells(self, coord_type, coords, conn, cell_ids=None):
        coords = np.ascontiguousarray(coords, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        cid = None if cell_ids is None else np.ascontiguousarray(cell_ids, np.int32)
        self.n_cells = conn.shape[0]
        self.dim = coords.shape[1]
        self._chk(lib().nuto_set_cells(self.h, coord_type, conn.shape[0], coords.shape[0], _np_ptr(coords),
                                       _np_ptr(conn), _np_ptr(cid)))

    def set_dof_type(self, slot, dof_type, ncomp, n_dofs, cell_dofs_):
        cd = np.ascontiguousarray(cell_dofs_, np.int32)
        self._chk(lib().nuto_set_dof_type(self.h, slot, dof_type, ncomp, n_dofs, _np_ptr(cd)))
        self._n_dofs[slot] = n_dofs
        self._n_eq[slot] = 0  # setting the dof type removes the slot's constraint equations
        if slot == 0:
            self.n_dofs0, self.ncomp0 = n_dofs, ncomp

    def set_integrand(self, params):
        self._chk(lib().nuto_set_integrand(self.h, ctypes.byref(params)))

    def set_rule(self, kind, n):
        self._chk(lib().nuto_set_rule(self.h, kind, n))
        # IntegrationTypeTensorProduct: n^dim points; IntegrationTypeTetrahedron order 1: 1, order 2: 4
        self.n_ip = n ** self.dim if kind == 0 else {1: 1, 2: 4}[n]

    def set_history_ld(self, ld):
        self._chk(lib().nuto_set_history_ld(self.h, ld))
        self._history_ld = ld if ld > 0 else None

    def set_material_fields(self, E=None, ft=None, gf=None):
        a = [None if v is None else np.ascontiguousarray(v, np.float64) for v in (E, ft, gf)]
        self._chk(lib().nuto_set_material_fields(self.h, *[_np_ptr(v) for v in a]))

    def set_material_fields_device(self, theta):
        """Per-cell (E, f_t, G_f) from a (3, n_cells) float64 device tensor (None: scalar parameters)."""
        if theta is None:
            self._chk(lib().nuto_set_material_fields_device(self.h, None, None, None))
            return
        if theta.shape != (3, self.n_cells) or theta.dtype != self._torch.float64:
            raise ValueError("theta must be a (3, n_cells) float64 tensor")
        theta = theta.contiguous()
        self._chk(lib().nuto_set_material_fields_device(self.h, _dev_ptr(theta[0]), _dev_ptr(theta[1]),
                                                        _dev_ptr(theta[2])))
        self._theta_keep = theta  # the copy is stream-ordered; keep the source alive until it ran

    def build_pattern(self):
        nnz, nr = ctypes.c_longlong(), ctypes.c_int()
        self._chk(lib().nuto_build_pattern(self.h, ctypes.byref(nnz), ctypes.byref(nr)))
        self.nnz, self.n_rows = nnz.value, nr.value
        return self.nnz, self.n_rows

    def pattern(self):
        rp = np.zeros(self.n_rows + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        self._chk(lib().nuto_copy_pattern(self.h, _np_ptr(rp), _np_ptr(ci)))
        return rp, ci

    def block_pattern(self, slot_i, slot_j):
        """Block (slot_i, slot_j) of the matrix, K(dofI, dofJ) of NuTo's DofMatrixSparse: (row_ptr, col,
        run_begin) as numpy arrays; columns numbered within slot_j; run_begin[r] is the block row's first
        entry in the monolithic values (nuto_block_pattern)."""
        nnz = ctypes.c_longlong()
        self._chk(lib().nuto_block_pattern(self.h, slot_i, slot_j, ctypes.byref(nnz), None, None, None))
        rows = self._n_dofs.get(slot_i, 0)
        rp = np.zeros(rows + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        rb = np.zeros(rows, np.int64)
        self._chk(lib().nuto_block_pattern(self.h, slot_i, slot_j, ctypes.byref(nnz), _np_ptr(rp), _np_ptr(ci),
                                           _np_ptr(rb)))
        return rp, ci, rb

    def block_values(self, slot_i, slot_j, K, out=None):
        """The values of block (slot_i, slot_j) of K (a device tensor as build_matrix returns it), in the
        block's CSR order (nuto_block_values)."""
        nnz = ctypes.c_longlong()
        self._chk(lib().nuto_block_pattern(self.h, slot_i, slot_j, ctypes.byref(nnz), None, None, None))
        kp = _f64(K, self.nnz, "K", optional=False)
        out = self._new(nnz.value) if out is None else _f64(out, nnz.value, "out")
        self._chk(lib().nuto_block_values(self.h, slot_i, slot_j, *_ptrs(kp, out)))
        return out

    def pattern_device(self):
        """(row_ptr, col) as torch int32 tensors on the device (copied through the host)."""
        torch = self._torch
        rp, ci = self.pattern()
        return torch.from_numpy(rp).cuda(), torch.from_numpy(ci).cuda()

    def _new(self, n):
        return self._torch.empty(n, dtype=self._torch.float64, device="cuda")

    def build_vector(self, u, e=None, kappa=None, out=None):
        R = self._new(self.n_rows) if out is None else out
        self._chk(lib().nuto_build_vector(self.h, *_ptrs(*self._state(u, e, kappa), _f64(R, self.n_rows, "out"))))
        return R

    def build_matrix(self, u, e=None, kappa=None, out=None):
        K = self._new(self.nnz) if out is None else out
        self._chk(lib().nuto_build_matrix(self.h, *_ptrs(*self._state(u, e, kappa), _f64(K, self.nnz, "out"))))
        return K

    def build_vector_matrix(self, u, e=None, kappa=None, R=None, K=None):
        R = self._new(self.n_rows) if R is None else R
        K = self._new(self.nnz) if K is None else K
        self._chk(lib().nuto_build_vector_matrix(self.h, *_ptrs(*self._state(u, e, kappa), _f64(R, self.n_rows, "R"),
                                                         _f64(K, self.nnz, "K"))))
        return R, K

    def build_vector_matrix_noalloc(self, u, e, kappa, R, K):
        """build_vector_matrix into the caller's R and K without any device allocation (the records go
        to the library's per-device scratch; at most 29127 hex8 gradient-damage cells)."""
        self._chk(lib().nuto_build_vector_matrix_noalloc(
            self.h, *_ptrs(*self._state(u, e, kappa), _f64(R, self.n_rows, "R", optional=False),
                           _f64(K, self.nnz, "K", optional=False))))
        return R, K

    def vjp(self, u, e, kappa, lam, want_theta=True):
        """Adjoint of build_vector: (K^T lam, dR/dtheta^T lam as (3, n_cells) or None)."""
        args = (*self._state(u, e, kappa), _f64(lam, self.n_rows, "lam", optional=False))
        d_state = self._new(self.n_rows)
        d_theta = self._new(3 * self.n_cells) if want_theta else None
        self._chk(lib().nuto_vjp(self.h, *_ptrs(*args, d_state, d_theta)))
        return d_state, (None if d_theta is None else d_theta.view(3, self.n_cells))

    def update_history(self, u, e, kappa):
        self._chk(lib().nuto_update_history(self.h, *_ptrs(*self._state(u, e, kappa))))
        return kappa

    def build_lumped_mass(self, rho, out=None):
        """BuildDiagonallyLumpedMatrix with DynamicMomentumBalance::Hessian2 (N^T N rho): one value
        per dof of slot 0."""
        M = self._new(self.n_dofs0) if out is None else out
        self._chk(lib().nuto_build_lumped_mass(self.h, float(rho), *_ptrs(_f64(M, self._n_u(), "out"))))
        return M

    def set_constraints(self, slot, equations):
        """Constraints::Add for one dof type slot: equations = [(dependent dof, [(dof, coef), ...]), ...],
        each reading 1 x[dependent] + sum coef x[dof] = rhs."""
        dep = np.array([q[0] for q in equations], np.int32)
        tp = np.zeros(len(equations) + 1, np.int32)
        td, tc = [], []
        for i, (_, terms) in enumerate(equations):
            for d, c in terms:
                td.append(d)
                tc.append(c)
            tp[i + 1] = len(td)
        td = np.array(td, np.int32)
        tc = np.array(tc, np.float64)
        self._chk(lib().nuto_set_constraints(self.h, slot, len(dep), _np_ptr(dep), _np_ptr(tp), _np_ptr(td),
                                             _np_ptr(tc)))
        self._n_eq[slot] = len(dep)

    def _solve_opts(self, method, tol, max_iter):
        return SolverOptions(method, tol, max_iter)

    def solve(self, K, f, method=SOLVER_BICGSTAB, tol=0.0, max_iter=0, out=None):
        """NuTo::Solve: x = C (C^T K C)^-1 C^T f on the device. Returns (x, info dict)."""
        x = self._new(self.n_rows) if out is None else out
        info = SolveInfo()
        opt = self._solve_opts(method, tol, max_iter)
        self._chk(lib().nuto_solve(self.h, *_ptrs(_f64(K, self.nnz, "K", optional=False),
                                           _f64(f, self.n_rows, "f", optional=False), _f64(x, self.n_rows, "out")),
                                   ctypes.byref(opt), ctypes.byref(info)))
        return x, {"iterations": info.iterations, "relative_residual": info.relative_residual,
                   "independent_dofs": info.independent_dofs}

    def solve_trial_state(self, K, f, rhs_old, rhs_new, method=SOLVER_BICGSTAB, tol=0.0, max_iter=0, out=None):
        """NuTo::SolveTrialState: the negative increment C (C^T K C)^-1 C^T (f + K delta) - delta."""
        x = self._new(self.n_rows) if out is None else out
        info = SolveInfo()
        opt = self._solve_opts(method, tol, max_iter)
        self._chk(lib().nuto_solve_trial_state(self.h, *_ptrs(_f64(K, self.nnz, "K", optional=False),
                                                       _f64(f, self.n_rows, "f", optional=False),
                                                       _f64(rhs_old, self._n_rhs(), "rhs_old"),
                                                       _f64(rhs_new, self._n_rhs(), "rhs_new"),
                                                       _f64(x, self.n_rows, "out")),
                                               ctypes.byref(opt), ctypes.byref(info)))
        return x, {"iterations": info.iterations, "relative_residual": info.relative_residual,
                   "independent_dofs": info.independent_dofs}

    def quasistatic_step(self, x, kappa=None, rhs_old=None, rhs_new=None, tol=1e-10, max_iter=6,
                         line_search_steps=6, method=SOLVER_BICGSTAB, linear_tol=0.0, linear_max_iter=0):
        """QuasistaticSolver::DoStep on the device: trial state, Newton-Raphson with line search
        (line_search_steps=0: none), history update. x ([u; e], n_rows) and kappa are updated in place
        on convergence; otherwise NutoError with code NO_CONVERGENCE is raised and both are unchanged.
        Returns the info dict (iterations is DoStep's return value)."""
        o = NewtonOptions(float(tol), int(max_iter), int(line_search_steps),
                          SolverOptions(int(method), float(linear_tol), int(linear_max_iter)))
        info = NewtonInfo()
        self._chk(lib().nuto_quasistatic_step(self.h, *_ptrs(_f64(x, self.n_rows, "x", optional=False),
                                                      _f64(kappa, self._n_hist(), "kappa"),
                                                      _f64(rhs_old, self._n_rhs(), "rhs_old"),
                                                      _f64(rhs_new, self._n_rhs(), "rhs_new")),
                                              ctypes.byref(o), ctypes.byref(info)))
        return {"iterations": info.iterations, "linear_iterations": info.linear_iterations,
                "residual_evaluations": info.residual_evaluations, "residual_norm": info.residual_norm}

    def eval_ip(self, field, u, e=None, kappa=None, out=None):
        """Cell::Eval of one field (IP_STRAIN, IP_STRESS, IP_KAPPA, IP_DAMAGE, IP_EQ_STRAIN) at every
        (cell, IP): a (n_cells, n_ip, n_comp) tensor."""
        nc = (3 if self.dim == 2 else 6) if field in (IP_STRAIN, IP_STRESS) else 1
        o = self._new(self.n_cells * self.n_ip * nc) if out is None else out
        pu, pe, pk, po = _ptrs(*self._state(u, e, kappa), _f64(o, self.n_cells * self.n_ip * nc, "out"))
        self._chk(lib().nuto_eval_ip(self.h, pu, pe, pk, field, po))
        return o.view(self.n_cells, self.n_ip, nc)


class Batch:
    """Many independent meshes assembled together (the C ABI nuto_batch).

    The batch runs on a torch stream of its own (graph replay needs a non-null stream); every call
    first orders that stream after torch's current stream and afterwards orders the current stream
    after it, so results can be used by torch right away.
    """

    MESH_INFO = ("group", "dof_nodes", "cells", "nip", "u", "e", "R_u", "R_e", "rows", "K_u", "nnz_u", "K_e", "nnz_e",
                 "history", "history_ld", "theta", "theta_stride")

    def __init__(self, device=0):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("nuto_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self._torch = torch
        L = lib()
        self._L = L
        h = ctypes.c_void_p()
        r = L.nuto_batch_create(ctypes.byref(h), device)
        if r:
            raise NutoError(r, "nuto_batch_create failed")
        self.h = h
        self.stream = torch.cuda.Stream(device=device)
        self._chk(L.nuto_batch_set_stream(self.h, ctypes.c_void_p(self.stream.cuda_stream)))
        self.n_meshes = 0
        self.sizes = None

    def close(self):
        h, L = getattr(self, "h", None), getattr(self, "_L", None)
        if h and L is not None:
            L.nuto_batch_destroy(h)
        self.h = None

    __del__ = close

    def _chk(self, r):
        if r:
            raise NutoError(r, lib().nuto_batch_last_error(self.h).decode())

    def _enter(self):
        self.stream.wait_stream(self._torch.cuda.current_stream())

    def _leave(self):
        self._torch.cuda.current_stream().wait_stream(self.stream)

    def set_integrand(self, params):
        self._chk(lib().nuto_batch_set_integrand(self.h, ctypes.byref(params)))

    def _state(self, u, e, kappa):
        if self.sizes is None:
            raise ValueError("build() the batch first")
        sz = self.sizes
        return (_f64(u, sz["u"], "u", optional=False), _f64(e, sz["e"], "e"),
                _f64(kappa, sz["history"], "kappa"))

    def add_mesh(self, coord_type, dof_type, coords, conn, ranks, n_dof_nodes):
        coords = np.ascontiguousarray(coords, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        ranks = np.ascontiguousarray(ranks, np.int32)
        m = ctypes.c_int()
        self._chk(lib().nuto_batch_add_mesh(self.h, coord_type, dof_type, conn.shape[0], coords.shape[0],
                                            _np_ptr(coords), _np_ptr(conn), n_dof_nodes, _np_ptr(ranks),
                                            ctypes.byref(m)))
        self.n_meshes = m.value + 1
        return m.value

    def set_group_limit(self, max_dof_nodes):
        """Dof nodes per group (super-mesh) of one element pair (nuto_batch_set_group_limit; 0: default)."""
        self._chk(lib().nuto_batch_set_group_limit(self.h, int(max_dof_nodes)))

    def set_stage_bytes(self, nbytes):
        """Staging budget of each group's two-phase launches (nuto_batch_set_stage_bytes; 0: whole group)."""
        self._chk(lib().nuto_batch_set_stage_bytes(self.h, int(nbytes)))

    def build(self):
        self._chk(lib().nuto_batch_build(self.h))
        sz = np.zeros(6, np.int64)
        self._chk(lib().nuto_batch_sizes(self.h, _np_ptr(sz)))
        self.sizes = dict(zip(("u", "e", "R", "K", "history", "theta"), (int(v) for v in sz)))
        return self.sizes

    def mesh_info(self, m):
        info = np.zeros(17, np.int64)
        self._chk(lib().nuto_batch_mesh_info(self.h, m, _np_ptr(info)))
        return dict(zip(self.MESH_INFO, (int(v) for v in info)))

    def group_info(self, g):
        """A group's dof element type, meshes, dof nodes, cells, cell-matrix kernel (0 by columns,
        1 / 2 Gram form with one / two row nodes per lane, 3 Gram form on the matrix cores, -1 the
        element's default), the number of groups and the cells per staging chunk (0: the whole
        group, -1: not chosen yet)."""
        info = np.zeros(7, np.int64)
        self._chk(lib().nuto_batch_group_info(self.h, g, _np_ptr(info)))
        return dict(zip(("dof_type", "meshes", "dof_nodes", "cells", "cell_kernel", "groups", "stage_cells"),
                        (int(v) for v in info)))

    def pattern(self, m):
        i = self.mesh_info(m)
        rp = np.zeros(i["rows"] + 1, np.int32)
        ci = np.zeros(i["nnz_u"] + i["nnz_e"], np.int32)
        self._chk(lib().nuto_batch_copy_pattern(self.h, m, _np_ptr(rp), _np_ptr(ci)))
        return rp, ci

    def set_material_fields(self, E=None, ft=None, gf=None):
        a = [None if v is None else np.ascontiguousarray(v, np.float64) for v in (E, ft, gf)]
        self._chk(lib().nuto_batch_set_material_fields(self.h, *[_np_ptr(v) for v in a]))

    def _new(self, n):
        return self._torch.empty(max(n, 1), dtype=self._torch.float64, device="cuda")

    def assemble(self, u, e=None, kappa=None, R=None, K=None, mode=-1):
        """R and K of every mesh (nuto_batch_assemble). mode -1: the tuned launch replayed from a
        HIP graph; 0..2 fused row-owner launch (64 / 128 / 256 nodes per workgroup); 3 row-owner per
        element type; 4 row-owner per mesh; 5 two-phase per element type; 6 two-phase per mesh."""
        state = self._state(u, e, kappa)
        R = self._new(self.sizes["R"]) if R is None else _f64(R, self.sizes["R"], "R")
        K = self._new(self.sizes["K"]) if K is None else _f64(K, self.sizes["K"], "K")
        args = _ptrs(*state, R, K)
        self._enter()
        self._chk(lib().nuto_batch_assemble(self.h, *args, mode))
        self._leave()
        return R, K

    def autotune(self, u, e=None, kappa=None, R=None, K=None):
        state = self._state(u, e, kappa)
        R = self._new(self.sizes["R"]) if R is None else _f64(R, self.sizes["R"], "R")
        K = self._new(self.sizes["K"]) if K is None else _f64(K, self.sizes["K"], "K")
        args = _ptrs(*state, R, K)
        ms = np.zeros(7)
        best = ctypes.c_int()
        self._enter()
        self._chk(lib().nuto_batch_autotune(self.h, *args, _np_ptr(ms), ctypes.byref(best)))
        self._leave()
        return ms, best.value

    def vjp(self, u, e, kappa, lam):
        args = (*self._state(u, e, kappa), _f64(lam, self.sizes["R"], "lam", optional=False))
        d_state = self._new(self.sizes["R"])
        d_theta = self._new(self.sizes["theta"])
        ptrs = _ptrs(*args, d_state, d_theta)
        self._enter()
        self._chk(lib().nuto_batch_vjp(self.h, *ptrs))
        self._leave()
        return d_state, d_theta

    def residual_adjoint(self, u, e, kappa, r_obs, R=None, K=None, d_state=None, d_theta=None):
        state = self._state(u, e, kappa)
        sz = self.sizes
        r_obs = _f64(r_obs, sz["R"], "r_obs", optional=False)
        R = self._new(sz["R"]) if R is None else _f64(R, sz["R"], "R")
        K = self._new(sz["K"]) if K is None else _f64(K, sz["K"], "K")
        d_state = self._new(sz["R"]) if d_state is None else _f64(d_state, sz["R"], "d_state")
        d_theta = self._new(sz["theta"]) if d_theta is None else _f64(d_theta, sz["theta"], "d_theta")
        ptrs = _ptrs(*state, r_obs, R, K, d_state, d_theta)
        self._enter()
        self._chk(lib().nuto_batch_residual_adjoint(self.h, *ptrs))
        self._leave()
        return R, K, d_state, d_theta

    def synchronize(self):
        self._chk(lib().nuto_batch_synchronize(self.h))
