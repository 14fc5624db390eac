"""ctypes wrapper of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product library never does.  The C restatement of the
reference algorithm lives in oracle/dg_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

GLL, GL = 0, 1
KG, IR, CH, RA = 0, 1, 2, 3
LXF, ROE, CENTRAL, TWO_POINT, TWO_POINT_LXF, TWO_POINT_ROE = 0, 1, 2, 3, 4, 5
EULER, NAVIER_STOKES = 0, 1
SIPG, BR2 = 0, 1
ORDER_DEALII, ORDER_LEXICOGRAPHIC = 0, 1
METHODS = {"euler": 0, "heun": 1, "ssprk3": 2, "rk4": 3,
           "rk4_3_5_3sstar": 10, "rk3_2_5f_3sstarplus": 11,
           "rk4_3_9f_3sstarplus": 12, "rk5_4_10f_3sstarplus": 13}
# rk_order of the low-storage methods (parameters_ode_solver.cpp:270-296)
RK_ORDER = {"rk4_3_5_3sstar": 4, "rk3_2_5f_3sstarplus": 3, "rk4_3_9f_3sstarplus": 4, "rk5_4_10f_3sstarplus": 5}


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("poly_degree", ctypes.c_int), ("flux_nodes", ctypes.c_int),
        ("use_split_form", ctypes.c_int), ("two_point_flux", ctypes.c_int), ("conv_flux", ctypes.c_int),
        ("pde", ctypes.c_int), ("diss_flux", ctypes.c_int), ("use_constant_viscosity", ctypes.c_int),
        ("gamma_gas", ctypes.c_double), ("mach_inf", ctypes.c_double), ("reynolds_inf", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("constant_viscosity", ctypes.c_double),
        ("temperature_inf", ctypes.c_double), ("sipg_penalty_factor", ctypes.c_double),
        ("n_cells_1d", ctypes.c_int), ("domain_left", ctypes.c_double), ("domain_right", ctypes.c_double),
        ("cell_order", ctypes.c_int),
        ("boundary_type", ctypes.c_int * 6), ("angle_of_attack", ctypes.c_double), ("grid_warp", ctypes.c_int),
        ("thermal_bc", ctypes.c_int), ("wall_temperature", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    P = ctypes.POINTER
    d = ctypes.c_double
    i = ctypes.c_int
    vp = ctypes.c_void_p
    dp = P(d)
    lib.dgo_create.restype = vp
    lib.dgo_boundary_state.argtypes = [vp, i, dp, dp, dp]
    lib.dgo_create.argtypes = [P(Params)]
    lib.dgo_destroy.argtypes = [vp]
    lib.dgo_default_params.argtypes = [P(Params)]
    lib.dgo_n_dofs.restype = ctypes.c_long
    lib.dgo_n_dofs.argtypes = [vp]
    lib.dgo_n_cells.argtypes = [vp]
    lib.dgo_n_faces.argtypes = [vp]
    lib.dgo_gauss_legendre.argtypes = [i, dp, dp]
    lib.dgo_gauss_lobatto.argtypes = [i, dp, dp]
    lib.dgo_lagrange_values.argtypes = [i, dp, d, dp, dp]
    lib.dgo_get_1d_operators.argtypes = [vp] + [dp] * 10
    lib.dgo_tensor_apply.argtypes = [i, dp, i, i, dp, i, i, dp, i, i, dp, dp, i, d]
    lib.dgo_hadamard_volume_pattern.argtypes = [i, i, P(ctypes.c_uint), P(ctypes.c_uint)]
    lib.dgo_hadamard_surface_pattern.argtypes = [i, i, i, P(ctypes.c_uint), P(ctypes.c_uint)]
    lib.dgo_cell_ijk.argtypes = [vp, P(i)]
    lib.dgo_neighbors.argtypes = [vp, P(i)]
    lib.dgo_face_list.argtypes = [vp, P(i)]
    lib.dgo_node_coords.argtypes = [vp, dp]
    lib.dgo_set_grid_nodes.argtypes = [vp, dp]
    lib.dgo_set_grid_nodes.restype = i
    lib.dgo_metric_terms.argtypes = [vp, dp, dp]
    lib.dgo_pressure.restype = d
    lib.dgo_pressure.argtypes = [vp, dp]
    lib.dgo_convective_flux.argtypes = [vp, dp, dp]
    lib.dgo_entropy_variables.argtypes = [vp, dp, dp]
    lib.dgo_conservative_from_entropy.argtypes = [vp, dp, dp]
    lib.dgo_two_point_flux.argtypes = [vp, i, dp, dp, dp]
    lib.dgo_log_mean.restype = d
    lib.dgo_log_mean.argtypes = [d, d]
    lib.dgo_numerical_flux.argtypes = [vp, i, dp, dp, dp, dp]
    lib.dgo_dissipative_flux.argtypes = [vp, dp, dp, dp]
    lib.dgo_max_convective_eigenvalue.restype = d
    lib.dgo_max_convective_eigenvalue.argtypes = [vp, dp]
    lib.dgo_assemble_residual.argtypes = [vp, dp, dp]
    lib.dgo_apply_inverse_mass.argtypes = [vp, dp, dp, i]
    lib.dgo_get_auxiliary.argtypes = [vp, i, dp]
    lib.dgo_ic_tgv.argtypes = [vp, i, dp]
    lib.dgo_ic_vortex.argtypes = [vp, d, d, d, dp]
    lib.dgo_n_stages.argtypes = [i]
    lib.dgo_step.argtypes = [vp, i, d, dp]
    lib.dgo_constant_time_step.restype = d
    lib.dgo_constant_time_step.argtypes = [vp, d]
    lib.dgo_vortex_time_step.restype = d
    lib.dgo_vortex_time_step.argtypes = [vp, d, d]
    lib.dgo_integrated_quantities.argtypes = [vp, dp, i, dp]
    lib.dgo_entropy_rate.restype = d
    lib.dgo_entropy_rate.argtypes = [vp, dp]
    lib.dgo_numerical_entropy.restype = d
    lib.dgo_numerical_entropy.argtypes = [vp, dp]
    lib.dgo_max_wave_speed.restype = d
    lib.dgo_max_wave_speed.argtypes = [vp, dp, i]
    lib.dgo_adaptive_time_step.restype = d
    lib.dgo_adaptive_time_step.argtypes = [vp, dp, d]
    lib.dgo_run.argtypes = [vp, i, d, d, dp, dp, i]
    lib.dgo_set_threads.argtypes = [i]
    lib.dgo_set_mutation.argtypes = [i]
    lib.dgo_step_error_norm.restype = d
    lib.dgo_step_error_norm.argtypes = [vp, i, d, dp, d, d]
    lib.dgo_step_rrk.restype = d
    lib.dgo_step_rrk.argtypes = [vp, i, d, dp, dp, d]
    lib.dgo_initial_step_size.restype = d
    lib.dgo_initial_step_size.argtypes = [vp, dp]
    lib.dgo_set_fsal.argtypes = [vp, i]
    lib.dgo_residual_magnitude.argtypes = [vp, dp, dp]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(**kw):
    p = Params()
    lib().dgo_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "boundary_type":
            v = (ctypes.c_int * 6)(*v)
        setattr(p, k, v)
    return p


def gauss_legendre(n):
    x = np.zeros(n); w = np.zeros(n)
    lib().dgo_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def gauss_lobatto(n):
    x = np.zeros(n); w = np.zeros(n)
    lib().dgo_gauss_lobatto(n, _dp(x), _dp(w))
    return x, w


def log_mean(a, b):
    return lib().dgo_log_mean(a, b)


def tensor_apply(dim, Ax, Ay, Az, x, adding=False, factor=1.0, out=None):
    Ax = np.ascontiguousarray(Ax, dtype=np.float64)
    Ay = np.ascontiguousarray(Ay, dtype=np.float64)
    Az = np.ascontiguousarray(Az if Az is not None else np.ones((1, 1)), dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    nout = Ax.shape[0] * Ay.shape[0] * (Az.shape[0] if dim == 3 else 1)
    y = np.zeros(nout) if out is None else out
    lib().dgo_tensor_apply(dim, _dp(Ax), Ax.shape[0], Ax.shape[1], _dp(Ay), Ay.shape[0], Ay.shape[1],
                           _dp(Az), Az.shape[0], Az.shape[1], _dp(x), _dp(y), int(adding), factor)
    return y


def hadamard_volume_pattern(dim, n):
    m = n ** dim * n * dim
    r = np.zeros(m, dtype=np.uint32); c = np.zeros(m, dtype=np.uint32)
    lib().dgo_hadamard_volume_pattern(dim, n, r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                      c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return r.reshape(-1, dim), c.reshape(-1, dim)


def hadamard_surface_pattern(dim, n, dnz):
    m = n ** (dim - 1) * n
    r = np.zeros(m, dtype=np.uint32); c = np.zeros(m, dtype=np.uint32)
    lib().dgo_hadamard_surface_pattern(dim, n, dnz, r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)))
    return r, c


class Oracle:
    """One discretisation: mesh + operators + physics."""

    def __init__(self, **kw):
        self.p = params(**kw)
        self.h = lib().dgo_create(ctypes.byref(self.p))
        if not self.h:
            raise ValueError("invalid oracle parameters")
        self.dim = self.p.dim
        self.ns = self.dim + 2
        self.n = self.p.poly_degree + 1
        self.Np = self.n ** self.dim
        self.ncells = lib().dgo_n_cells(self.h)
        self.ndofs = lib().dgo_n_dofs(self.h)

    def __del__(self):
        try:
            if self.h:
                lib().dgo_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def vec(self):
        return np.zeros(self.ndofs)

    def operators(self):
        n = self.n
        out = dict(nodes=np.zeros(n), weights=np.zeros(n), chi=np.zeros((n, n)), chi_grad=np.zeros((n, n)),
                   flux_grad=np.zeros((n, n)), skew=np.zeros((n, n)), proj=np.zeros((n, n)),
                   minv=np.zeros((n, n)), chi_surf=np.zeros((2, n)), flux_surf=np.zeros((2, n)))
        lib().dgo_get_1d_operators(self.h, *[_dp(out[k]) for k in
                                             ["nodes", "weights", "chi", "chi_grad", "flux_grad", "skew",
                                              "proj", "minv", "chi_surf", "flux_surf"]])
        return out

    def cell_ijk(self):
        a = np.zeros((self.ncells, 3), dtype=np.int32)
        lib().dgo_cell_ijk(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return a

    def neighbors(self):
        a = np.zeros((self.ncells, 2 * self.dim), dtype=np.int32)
        lib().dgo_neighbors(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return a

    def face_list(self):
        nf = lib().dgo_n_faces(self.h)
        a = np.zeros((nf, 4), dtype=np.int32)
        lib().dgo_face_list(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return a

    def node_coords(self):
        a = np.zeros((self.ncells, self.Np, 3))
        lib().dgo_node_coords(self.h, _dp(a))
        return a

    # point physics
    def pressure(self, u):
        return lib().dgo_pressure(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)))

    def convective_flux(self, u):
        f = np.zeros((self.ns, self.dim))
        lib().dgo_convective_flux(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)), _dp(f))
        return f

    def entropy_variables(self, u):
        v = np.zeros(self.ns)
        lib().dgo_entropy_variables(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)), _dp(v))
        return v

    def conservative_from_entropy(self, v):
        u = np.zeros(self.ns)
        lib().dgo_conservative_from_entropy(self.h, _dp(np.ascontiguousarray(v, dtype=np.float64)), _dp(u))
        return u

    def two_point_flux(self, which, u1, u2):
        f = np.zeros((self.ns, self.dim))
        lib().dgo_two_point_flux(self.h, which, _dp(np.ascontiguousarray(u1, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(u2, dtype=np.float64)), _dp(f))
        return f

    def numerical_flux(self, which, ui, ue, n):
        f = np.zeros(self.ns)
        lib().dgo_numerical_flux(self.h, which, _dp(np.ascontiguousarray(ui, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(ue, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(n, dtype=np.float64)), _dp(f))
        return f

    def dissipative_flux(self, u, g):
        f = np.zeros((self.ns, self.dim))
        lib().dgo_dissipative_flux(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(g, dtype=np.float64)), _dp(f))
        return f

    # discretisation
    def ic_tgv(self, isothermal=False):
        u = self.vec()
        lib().dgo_ic_tgv(self.h, int(isothermal), _dp(u))
        return u

    def ic_vortex(self, strength=5.0, u_inf=None, v_inf=None):
        M = np.sqrt(2 / 1.4)
        if u_inf is None:
            u_inf = M * np.cos(np.pi / 4)
        if v_inf is None:
            v_inf = M * np.sin(np.pi / 4)
        u = self.vec()
        lib().dgo_ic_vortex(self.h, strength, u_inf, v_inf, _dp(u))
        return u

    def set_grid_nodes(self, xyz):
        """Curvilinear grid from node coordinates [cell][node][3] (GLL, Euler)."""
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.ncells, self.Np, 3)
        if lib().dgo_set_grid_nodes(self.h, _dp(x)) != 0:
            raise ValueError("curvilinear grids need GLL flux nodes and the Euler equations")

    def metric_terms(self):
        """(C [cell][node][3][3] with C[m][n] = J dxi_n/dx_m, det J [cell][node])."""
        C = np.zeros((self.ncells, self.Np, 3, 3))
        J = np.zeros((self.ncells, self.Np))
        lib().dgo_metric_terms(self.h, _dp(C), _dp(J))
        return C, J

    def boundary_state(self, bc_type, u_int, normal):
        """Euler::boundary_face_values: exterior state of a boundary face node."""
        ui = np.ascontiguousarray(u_int, dtype=np.float64)
        nrm = np.zeros(3)
        nrm[:len(normal)] = normal
        out = np.zeros(self.ns)
        lib().dgo_boundary_state(self.h, bc_type, _dp(ui), _dp(nrm), _dp(out))
        return out

    def residual(self, u):
        r = self.vec()
        lib().dgo_assemble_residual(self.h, _dp(np.ascontiguousarray(u)), _dp(r))
        return r

    def residual_magnitude(self, u):
        """Per-DoF rounding scale of the residual: the sum of the absolute values of the terms
        summed into each DoF (GLL nodes)."""
        m = self.vec()
        lib().dgo_residual_magnitude(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)), _dp(m))
        return m

    def set_fsal(self, on):
        """3S*+ methods: evaluate the FSAL residual of the embedded error estimate (default on,
        as the reference); it does not change the solution."""
        lib().dgo_set_fsal(self.h, int(on))

    def inverse_mass(self, r):
        o = self.vec()
        lib().dgo_apply_inverse_mass(self.h, _dp(np.ascontiguousarray(r)), _dp(o), 0)
        return o

    def rhs(self, u):
        return self.inverse_mass(self.residual(u))

    def auxiliary(self, k):
        o = self.vec()
        lib().dgo_get_auxiliary(self.h, k, _dp(o))
        return o

    def step(self, method, dt, u):
        m = METHODS[method] if isinstance(method, str) else method
        lib().dgo_step(self.h, m, dt, _dp(u))
        return u

    def step_error_norm(self, method, dt, u, atol=1e-3, rtol=1e-3):
        """One low-storage RK step of u in place; returns the error norm w of its embedded pair."""
        m = METHODS[method] if isinstance(method, str) else method
        return lib().dgo_step_error_norm(self.h, m, dt, _dp(u), atol, rtol)

    def step_rrk(self, method, dt, u, gamma=1.0, tol=5e-10):
        """One relaxation-RK step of u in place; returns (time advanced gamma*dt, gamma)."""
        m = METHODS[method] if isinstance(method, str) else method
        g = ctypes.c_double(gamma)
        adv = lib().dgo_step_rrk(self.h, m, dt, _dp(u), ctypes.byref(g), tol)
        return adv, g.value

    def initial_step_size(self, u):
        return lib().dgo_initial_step_size(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)))

    def run_error_adaptive(self, method, final_time, u, atol=1e-3, rtol=1e-3, beta=(0.70, -0.23, 0.0)):
        """FlowSolver::run with error_adaptive_time_step (flow_solver.cpp:505-507, 554-607): the first
        step from get_automatic_initial_step_size, every later one from the PID update of
        get_automatic_error_adaptive_step_size (low_storage_runge_kutta_ode_solver.cpp:89-118)
        with epsilon = {1, 1, 1} initially; the last step clipped to the final time.  Returns the
        steps taken."""
        k = RK_ORDER[method]
        eps = [1.0, 1.0, 1.0]
        dt_next = self.initial_step_size(u)
        t, dts = 0.0, []
        while t < final_time:
            dt = dt_next
            if t + dt > final_time:
                dt = final_time - t
            w = self.step_error_norm(method, dt, u, atol, rtol)
            t += dt
            dts.append(dt)
            eps = [1.0 / w, eps[0], eps[1]]
            dt_next = pow(eps[0], 1.0 * beta[0] / k) * pow(eps[1], 1.0 * beta[1] / k) * pow(eps[2], 1.0 * beta[2] / k) * dt
        return dts

    def constant_time_step(self, cfl):
        return lib().dgo_constant_time_step(self.h, cfl)

    def numerical_entropy(self, u):
        return lib().dgo_numerical_entropy(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)))

    def max_wave_speed(self, u, overint=10):
        return lib().dgo_max_wave_speed(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)), overint)

    def adaptive_time_step(self, u, cfl):
        return lib().dgo_adaptive_time_step(self.h, _dp(np.ascontiguousarray(u, dtype=np.float64)), cfl)

    def vortex_time_step(self, cfl, speed):
        return lib().dgo_vortex_time_step(self.h, cfl, speed)

    def integrated(self, u, overint=10):
        o = np.zeros(4)
        lib().dgo_integrated_quantities(self.h, _dp(np.ascontiguousarray(u)), overint, _dp(o))
        return dict(kinetic_energy=o[0], enstrophy=o[1], incompressible_kinetic_energy=o[2],
                    vorticity_dissipation=o[3])

    def entropy_rate(self, u):
        return lib().dgo_entropy_rate(self.h, _dp(np.ascontiguousarray(u)))

    def run(self, method, dt, final_time, u, max_hist=0):
        m = METHODS[method] if isinstance(method, str) else method
        hist = np.zeros(max(max_hist, 1))
        nsteps = lib().dgo_run(self.h, m, dt, final_time, _dp(u), _dp(hist) if max_hist else None, max_hist)
        return nsteps, hist[:min(nsteps, max_hist)]


def set_threads(n):
    lib().dgo_set_threads(int(n))


(MUT_NONE, MUT_LXF_MIN, MUT_LXF_NO_HALF, MUT_ROE_NO_HALF, MUT_ROE_NO_HARTEN, MUT_ROE_NO_SHEAR,
 MUT_NO_METRIC_AVERAGE, MUT_CROSS_PRODUCT_METRIC, MUT_TRANSPOSED_METRIC, MUT_CELL_CONSTANT_DETJ) = range(10)


def set_mutation(m):
    """Plant one of the DGO_MUT_* mistakes in the oracle (tests of its checks only)."""
    lib().dgo_set_mutation(int(m))
