"""ctypes wrapper of the CPU oracle (liboracle.so built from nuto_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg. The product package (nuto_amd) never imports it.

The oracle restates NuTo's assembly path (SimpleAssembler.cpp:38-156 and the files listed at the
top of nuto_oracle.cpp). Element types: 0 quad4, 1 hex8, 2 tet4, 3 tet10, 4 hex20.
Integrand kinds: 0 MomentumBalance + LinearElastic, 1 GradientDamage<3>,
2 MomentumBalance<3> + LocalIsotropicDamage<3>.

Beyond the assembly: lumped mass and IP fields (Problem.lumped_mass / eval_ip), the unit
constraint matrix and the constrained direct solve (unit_constraint_matrix, solve_constrained),
and the quasistatic load step (quasistatic_step: QuasistaticSolver::DoStep with NewtonRaphson and
LineSearch, in Python over the C restatement). Their pins: tests/test_cell_outputs_oracle.py,
tests/test_constraints_oracle.py (the reference's known-answer constraint matrix, a patch test)
and tests/test_newton_oracle.py (the homogeneous damage patch in closed form).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

QUAD4, HEX8, TET4, TET10, HEX20 = 0, 1, 2, 3, 4
NUM_NODES = {QUAD4: 4, HEX8: 8, TET4: 4, TET10: 10, HEX20: 20}
DIM = {QUAD4: 2, HEX8: 3, TET4: 3, TET10: 3, HEX20: 3}


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


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_mesh_quads.restype = P
        L.orc_mesh_bricks.restype = P
        L.orc_mesh_from_arrays.restype = P
        L.orc_mesh_from_arrays.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip]
        L.orc_mesh_free.argtypes = [P]
        L.orc_mesh_scale.argtypes = [P, ctypes.c_double]
        L.orc_mesh_info.argtypes = [P, ip]
        L.orc_mesh_coords.argtypes = [P, dp]
        L.orc_mesh_conn.argtypes = [P, ip]
        L.orc_mesh_add_dof.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_dof_info.argtypes = [P, ctypes.c_int, ip]
        L.orc_dof_conn.argtypes = [P, ctypes.c_int, ip]
        L.orc_dof_numbers.argtypes = [P, ctypes.c_int, ip]
        L.orc_problem.restype = P
        L.orc_problem.argtypes = [P, ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_problem_free.argtypes = [P]
        L.orc_problem_set_theta.argtypes = [P, dp, dp, dp]
        L.orc_problem_ndof.argtypes = [P, ctypes.c_int]
        L.orc_problem_nip.argtypes = [P]
        L.orc_problem_assemble.restype = ctypes.c_longlong
        L.orc_problem_assemble.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_problem_R.argtypes = [P, dp, P]
        L.orc_problem_K.argtypes = [P, ip, ip, P, P]
        L.orc_problem_cell.argtypes = [P, ctypes.c_int, P, P, P, P, P]
        L.orc_problem_cell_dofs.argtypes = [P, ctypes.c_int, ip]
        L.orc_set_mutation.argtypes = [ctypes.c_int]
        L.orc_problem_lumped_mass.argtypes = [P, ctypes.c_double, dp, P]
        L.orc_unit_constraint_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip, dp, P]
        L.orc_solve_constrained.argtypes = [ctypes.c_int, ip, ip, ip, ip, ip, dp, ip, ip, dp, dp, ctypes.c_int, P, P, dp]
        L.orc_problem_eval_ip.argtypes = [P, P, P, P, ctypes.c_int, P]
        L.orc_mises.restype = ctypes.c_double
        L.orc_damage_exp.restype = ctypes.c_double
        L.orc_interaction.restype = ctypes.c_double
        L.orc_interaction.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, P]
        L.orc_mock_assemble.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def legendre_roots(n):
    out = np.zeros(n)
    lib().orc_legendre_roots(n, out.ctypes.data_as(ctypes.c_void_p))
    return out


def gauss_1d(n):
    w = np.zeros(n)
    x = np.zeros(n)
    lib().orc_gauss_1d(n, _ptr(w), _ptr(x))
    return w, x


def rule(kind, dim, n):
    """kind 0: tensor Gauss with n points per direction; kind 1: tetrahedron rule of order n."""
    num = lib().orc_rule(kind, dim, n, None, None)
    xi = np.zeros((num, 3))
    w = np.zeros(num)
    lib().orc_rule(kind, dim, n, _ptr(xi), _ptr(w))
    return xi, w


def shape(etype, xi):
    n = NUM_NODES[etype]
    N = np.zeros(n)
    dN = np.zeros((n, DIM[etype]))
    x = np.zeros(3)
    x[: len(xi)] = xi
    lib().orc_shape(etype, _ptr(x), _ptr(N), _ptr(dN))
    return N, dN


def local_coords(etype):
    out = np.zeros((NUM_NODES[etype], 3))
    lib().orc_local_coords(etype, _ptr(out))
    return out


def jacobian(X, dN):
    X = np.ascontiguousarray(X, dtype=np.float64)
    dN = np.ascontiguousarray(dN, dtype=np.float64)
    n, dim = dN.shape
    J = np.zeros((dim, dim))
    inv = np.zeros((dim, dim))
    det = ctypes.c_double()
    lib().orc_jacobian(dim, n, _ptr(X), _ptr(dN), _ptr(J), _ptr(inv), ctypes.byref(det))
    return J, inv, det.value


def led(E, nu, app, strain, omega):
    s = np.zeros(6)
    D = np.zeros((6, 6))
    w = np.zeros(6)
    strain = np.ascontiguousarray(strain, dtype=np.float64)
    lib().orc_led_stress(ctypes.c_double(E), ctypes.c_double(nu), app, _ptr(strain), ctypes.c_double(omega),
                         _ptr(s), _ptr(D), _ptr(w))
    return s, D, w


def mises(nu, k, strain):
    d = np.zeros(6)
    strain = np.ascontiguousarray(strain, dtype=np.float64)
    v = lib().orc_mises(ctypes.c_double(nu), ctypes.c_double(k), _ptr(strain), _ptr(d))
    return v, d


def damage_exp(k0, beta, alpha, kappa):
    d = ctypes.c_double()
    v = lib().orc_damage_exp(ctypes.c_double(k0), ctypes.c_double(beta), ctypes.c_double(alpha),
                             ctypes.c_double(kappa), ctypes.byref(d))
    return v, d.value


def interaction(kind, r, eta, omega):
    """NonlocalInteraction factor g(omega) and derivative g'(omega) (kind 0 Constant, 1 Decreasing)."""
    d = ctypes.c_double(0.0)
    g = lib().orc_interaction(int(kind), float(r), float(eta), float(omega), ctypes.byref(d))
    return g, d.value


def elastic_c(dim, E, nu, plane_state=0):
    n = 3 if dim == 2 else 6
    C = np.zeros((n, n))
    lib().orc_elastic_c(dim, ctypes.c_double(E), ctypes.c_double(nu), plane_state, _ptr(C))
    return C


class Mesh:
    """MeshFem restatement: coordinate nodes/elements plus dof layers (AddDofInterpolation)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def quads(cls, nx, ny):
        return cls(lib().orc_mesh_quads(nx, ny))

    @classmethod
    def bricks(cls, nx, ny, nz):
        return cls(lib().orc_mesh_bricks(nx, ny, nz))

    @classmethod
    def from_arrays(cls, coord_type, X, conn):
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        return cls(lib().orc_mesh_from_arrays(coord_type, X.shape[0], X, conn.shape[0], conn))

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:  # module globals are gone at interpreter exit
            lib().orc_mesh_free(self.h)
            self.h = None

    def scale(self, s):
        lib().orc_mesh_scale(self.h, ctypes.c_double(s))

    def info(self):
        o = np.zeros(5, dtype=np.int32)
        lib().orc_mesh_info(self.h, o)
        return dict(dim=int(o[0]), coord_type=int(o[1]), n_cells=int(o[2]), n_nodes=int(o[3]), n_layers=int(o[4]))

    def coords(self):
        i = self.info()
        X = np.zeros((i["n_nodes"], i["dim"]))
        lib().orc_mesh_coords(self.h, X)
        return X

    def conn(self):
        i = self.info()
        c = np.zeros((i["n_cells"], NUM_NODES[i["coord_type"]]), dtype=np.int32)
        lib().orc_mesh_conn(self.h, c)
        return c

    def add_dof(self, etype, ncomp):
        r = lib().orc_mesh_add_dof(self.h, etype, ncomp)
        if r < 0:
            raise RuntimeError("AddDofInterpolation failed (shape mismatch)")
        return r

    def dof_layer(self, layer):
        o = np.zeros(3, dtype=np.int32)
        lib().orc_dof_info(self.h, layer, o)
        etype, ncomp, nnodes = int(o[0]), int(o[1]), int(o[2])
        conn = np.zeros((self.info()["n_cells"], NUM_NODES[etype]), dtype=np.int32)
        lib().orc_dof_conn(self.h, layer, conn)
        num = np.zeros((nnodes, ncomp), dtype=np.int32)
        lib().orc_dof_numbers(self.h, layer, num)
        return dict(type=etype, ncomp=ncomp, n_nodes=nnodes, conn=conn, numbers=num)


class Problem:
    """A cell group + integrand + integration rule, assembled by the one-thread restatement."""

    def __init__(self, mesh, params, rule_kind, rule_n, layer0, layer1=-1):
        self.mesh = mesh  # keep alive
        self.p = params
        self.h = lib().orc_problem(mesh.h, ctypes.byref(params), rule_kind, rule_n, layer0, layer1)
        self.ntypes = 2 if layer1 >= 0 else 1
        self.ndof = [lib().orc_problem_ndof(self.h, t) for t in range(self.ntypes)]
        self.nip = lib().orc_problem_nip(self.h)
        self.ncells = mesh.info()["n_cells"]

    def __del__(self):
        if getattr(self, "h", None) and lib is not None:  # module globals are gone at interpreter exit
            lib().orc_problem_free(self.h)
            self.h = None

    def set_theta(self, E, ft, gf):
        lib().orc_problem_set_theta(self.h, np.ascontiguousarray(E, np.float64), np.ascontiguousarray(ft, np.float64),
                                    np.ascontiguousarray(gf, np.float64))

    def assemble(self, u, e=None, kappa=None, want_r=True, want_k=True, nthreads=1):
        u = np.ascontiguousarray(u, np.float64)
        e = None if e is None else np.ascontiguousarray(e, np.float64)
        kappa = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
        nnz = lib().orc_problem_assemble(self.h, _ptr(u), _ptr(e), _ptr(kappa), int(want_r), int(want_k), nthreads)
        if nnz < 0:
            raise RuntimeError("oracle assembly failed")
        out = {}
        ntot = sum(self.ndof)
        if want_r:
            R = np.zeros(ntot)
            Rb = np.zeros(ntot)
            lib().orc_problem_R(self.h, R, _ptr(Rb))
            out["R"] = R
            out["R_bound"] = Rb
        if want_k:
            rp = np.zeros(ntot + 1, dtype=np.int32)
            ci = np.zeros(nnz, dtype=np.int32)
            v = np.zeros(nnz)
            b = np.zeros(nnz)
            lib().orc_problem_K(self.h, rp, ci, _ptr(v), _ptr(b))
            out.update(row_ptr=rp, col=ci, val=v, bound=b)
        return out

    def cell(self, c, u, e=None, kappa=None, want_r=True, want_k=True):
        ntot_loc = lib().orc_problem_cell(self.h, c, _ptr(np.ascontiguousarray(u, np.float64)),
                                          _ptr(None if e is None else np.ascontiguousarray(e, np.float64)),
                                          _ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float64)),
                                          None, None)
        Re = np.zeros(ntot_loc) if want_r else None
        Ke = np.zeros((ntot_loc, ntot_loc)) if want_k else None
        lib().orc_problem_cell(self.h, c, _ptr(np.ascontiguousarray(u, np.float64)),
                               _ptr(None if e is None else np.ascontiguousarray(e, np.float64)),
                               _ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float64)), _ptr(Re),
                               _ptr(Ke))
        return Re, Ke

    def lumped_mass(self, rho):
        """BuildDiagonallyLumpedMatrix with Hessian2 = N^T N rho on dof type 0: (M, magnitude scale)."""
        M = np.zeros(self.ndof[0])
        Mb = np.zeros(self.ndof[0])
        lib().orc_problem_lumped_mass(self.h, float(rho), M, _ptr(Mb))
        return M, Mb

    def eval_ip(self, field, u, e=None, kappa=None):
        """Cell::Eval of one field at every (cell, IP): (n_cells, n_ip, n_comp)."""
        nc = lib().orc_problem_eval_ip(self.h, None, None, None, field, None)
        if nc == 0:
            raise ValueError("field not defined for this integrand")
        out = np.zeros(self.ncells * self.nip * nc)
        lib().orc_problem_eval_ip(self.h, _ptr(np.ascontiguousarray(u, np.float64)),
                                  _ptr(None if e is None else np.ascontiguousarray(e, np.float64)),
                                  _ptr(None if kappa is None else np.ascontiguousarray(kappa, np.float64)), field,
                                  _ptr(out))
        return out.reshape(self.ncells, self.nip, nc)

    def cell_dofs(self, c):
        out = np.zeros(64, dtype=np.int32)
        n = lib().orc_problem_cell_dofs(self.h, c, out)
        return out[:n].copy()


def mock_assemble(numbering, Re, Ke, ndof):
    """SimpleAssembler over mock cells with fixed local vectors/matrices (no integrand)."""
    numbering = np.ascontiguousarray(numbering, np.int32)
    Re = np.ascontiguousarray(Re, np.float64)
    Ke = np.ascontiguousarray(Ke, np.float64)
    ncells, nloc = numbering.shape
    R = np.zeros(ndof)
    nnz = lib().orc_mock_assemble(ncells, nloc, _ptr(numbering), _ptr(Re), _ptr(Ke), ndof, _ptr(R), None, None,
                                  None)
    rp = np.zeros(ndof + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    lib().orc_mock_assemble(ncells, nloc, _ptr(numbering), _ptr(Re), _ptr(Ke), ndof, _ptr(R), _ptr(rp), _ptr(ci),
                            _ptr(v))
    return R, rp, ci, v


def csr_to_dense(rp, ci, v, n, m=None):
    m = n if m is None else m
    A = np.zeros((n, m))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    return A


def set_mutation(m):
    """Plant mistake m in the oracle (0: none); for tests that show the parity rule rejects it."""
    lib().orc_set_mutation(int(m))


def _eq_arrays(equations):
    """[(dep, [(dof, coef), ...]), ...] -> (dep, term_ptr, term_dof, coef) int32 / float64 arrays."""
    dep = np.array([e[0] for e in equations], np.int32)
    tp = np.zeros(len(equations) + 1, np.int32)
    td, tc = [], []
    for i, (_, terms) in enumerate(equations):
        for d, c in terms:
            td.append(d)
            tc.append(c)
        tp[i + 1] = len(td)
    return dep, tp, np.array(td, np.int32), np.array(tc, np.float64)


def unit_constraint_matrix(n_dofs, equations):
    """Constraints::BuildUnitConstraintMatrix of one dof type (dense n_dofs x n_independent)."""
    dep, tp, td, tc = _eq_arrays(equations)
    n_ind = lib().orc_unit_constraint_matrix(n_dofs, len(dep), dep, tp, td, tc, None)
    if n_ind < 0:
        raise ValueError("dependent dof out of range")
    C = np.zeros((n_dofs, n_ind))
    lib().orc_unit_constraint_matrix(n_dofs, len(dep), dep, tp, td, tc, _ptr(C))
    return C


def solve_constrained(n_dofs, equations, row_ptr, col, val, f, rhs_old=None, rhs_new=None):
    """NuTo::Solve (rhs_new None) or NuTo::SolveTrialState on the monolithic system: n_dofs and
    equations per slot ([slot0, slot1]), K in CSR, f of length sum(n_dofs)."""
    eqs = [e for s in equations for e in s]
    dep, tp, td, tc = _eq_arrays(eqs)
    n = np.array(n_dofs, np.int32)
    ne = np.array([len(s) for s in equations], np.int32)
    x = np.zeros(int(n.sum()))
    trial = rhs_new is not None
    ro = None if rhs_old is None else np.ascontiguousarray(rhs_old, np.float64)
    rn = None if rhs_new is None else np.ascontiguousarray(rhs_new, np.float64)
    m = lib().orc_solve_constrained(len(n_dofs), n, ne, dep, tp, td, tc, np.ascontiguousarray(row_ptr, np.int32),
                                    np.ascontiguousarray(col, np.int32), np.ascontiguousarray(val, np.float64),
                                    np.ascontiguousarray(f, np.float64), int(trial), _ptr(ro), _ptr(rn), x)
    if m < 0:
        raise ValueError("dependent dof out of range")
    return x


class NoConvergence(RuntimeError):
    """NewtonRaphson::NoConvergence (NewtonRaphson.h:12-20)."""


def quasistatic_step(problem, n_dofs, equations, x, kappa=None, rhs_old=None, rhs_new=None, tol=1e-10, max_iter=6,
                     line_search_steps=6):
    """QuasistaticSolver::DoStep (QuasistaticSolver.cpp:116-147) on the oracle: the trial state
    x - SolveTrialState(K(x), R(x)) (TrialState, :51-66), NewtonRaphson::Solve (NewtonRaphson.h:62-94)
    with LineSearch (LineSearch.h:39-63) or, for line_search_steps == 0, NoLineSearch (:78-88), the
    norm |C^T R| (QuasistaticSolver::Norm, :84-90), and the history update on convergence
    (GradientDamage::Update, kappa = max(kappa, eeq), read back through Cell::Eval of kappa).
    The linear systems are solved directly (solve_constrained). x is the monolithic [u; e] vector,
    kappa the history (cell + ip * n_cells). Returns (x, kappa, iterations); raises NoConvergence."""
    nd = list(n_dofs)
    two = len(nd) > 1 and nd[1] > 0
    Cs = [unit_constraint_matrix(nd[s], equations[s]) for s in range(len(nd))]

    def split(v):
        return v[:nd[0]], (v[nd[0]:nd[0] + nd[1]] if two else None)

    def norm(v):
        parts, o = [], 0
        for s, C in enumerate(Cs):
            parts.append(C.T @ v[o:o + nd[s]])
            o += nd[s]
        return float(np.sqrt(sum(float(p @ p) for p in parts)))

    def assemble(v, want_r=True, want_k=True):
        u, e = split(v)
        return problem.assemble(u, e, kappa, want_r=want_r, want_k=want_k)

    # TrialState
    a = assemble(x)
    dx = solve_constrained(nd, equations, a["row_ptr"], a["col"], a["val"], a["R"],
                           rhs_old=np.zeros(sum(len(q) for q in equations)) if rhs_old is None else rhs_old,
                           rhs_new=np.zeros(sum(len(q) for q in equations)) if rhs_new is None else rhs_new)
    xk = x - dx
    # NewtonRaphson::Solve
    r = assemble(xk, want_k=False)["R"]
    nr = norm(r)
    it = 0
    converged = nr < tol
    while not converged and it < max_iter:
        k = assemble(xk, want_r=False)
        dx = solve_constrained(nd, equations, k["row_ptr"], k["col"], k["val"], r)
        it += 1
        if line_search_steps == 0:
            xk = xk - dx
            r = assemble(xk, want_k=False)["R"]
            nr = norm(r)
            converged = nr < tol
            continue
        x0, previous, alpha = xk.copy(), nr, 1.0
        for _ in range(line_search_steps):
            xk = x0 - alpha * dx
            r = assemble(xk, want_k=False)["R"]
            nr = norm(r)
            if nr < tol:
                converged = True
                break
            alpha *= 0.5
            if nr < (1.0 - alpha) * previous:
                break
    if not converged:
        raise NoConvergence(f"No convergence after {it} iterations.")
    if norm(xk) > 1e10:
        raise NoConvergence("floating point exception")
    new_kappa = None
    if kappa is not None:
        u, e = split(xk)
        kap = problem.eval_ip(2, u, e, kappa)[:, :, 0]  # Cell::Eval of kappa = max(history, eeq)
        new_kappa = np.ascontiguousarray(kap.T).ravel()  # cell + ip * n_cells
    return xk, new_kappa, it
