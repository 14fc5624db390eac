"""Python binding of the MI355X device DG library (libphilip_dg.so, C ABI in
include/philip_dg.h).  The library is built in-tree by `make -C philip_amd`;
importing `DG` without it raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphilip_dg.so")

GLL, GL = 0, 1
KG, IR, CH, RA = 0, 1, 2, 3
LXF, ROE, CENTRAL, TWO_POINT, TWO_POINT_LXF, TWO_POINT_ROE = 0, 1, 2, 3, 4, 5
EULER, NAVIER_STOKES = 0, 1
SIPG, BR2 = 0, 1
FP64, MIXED = 0, 1
RETURN_BIG_NUMBER, ABORT_RUN, PRINT_WARNING = 0, 1, 2   # non_physical_behavior (all_parameters.h:304)
ERR_NONPHYSICAL = -3
METHODS = {"euler": 0, "heun": 1, "ssprk3": 2, "rk4": 3,
           "rk4_3_5_3sstar": 10, "rk3_2_5f_3sstarplus": 11,
           "rk4_3_9f_3sstarplus": 12, "rk5_4_10f_3sstarplus": 13}

EXPORTED = [
    "dgx_default_params", "dgx_create", "dgx_destroy", "dgx_last_error", "dgx_n_dofs", "dgx_n_cells",
    "dgx_set_solution", "dgx_get_solution", "dgx_get_right_hand_side", "dgx_get_auxiliary_solution",
    "dgx_get_neighbors", "dgx_set_grid_nodes", "dgx_get_grid_nodes", "dgx_get_metric_terms",
    "dgx_assemble_residual", "dgx_apply_inverse_mass_to_rhs",
    "dgx_assemble_residual_device", "dgx_rhs_device", "dgx_apply_inverse_mass_device",
    "dgx_allocate_ode_system", "dgx_step_in_time", "dgx_current_time", "dgx_current_iteration", "dgx_set_time",
    "dgx_set_track_error_estimate", "dgx_constant_time_step", "dgx_integrated_quantities",
    "dgx_integrated_quantities_static",
    "dgx_numerical_entropy", "dgx_max_wave_speed", "dgx_adaptive_time_step", "dgx_set_stream",
    "dgx_enable_timing", "dgx_kernel_timing", "dgx_set_initial_condition_tgv",
    "dgx_set_initial_condition_vortex", "dgx_synchronize", "dgx_set_non_physical_behavior", "dgx_check_state",
    "dgx_nonphysical_events", "dgx_set_error_adaptive_parameters", "dgx_get_automatic_initial_step_size",
    "dgx_get_error_adaptive_step_size", "dgx_set_relaxation_runge_kutta", "dgx_relaxation_parameter",
    "dgx_batch_create", "dgx_batch_destroy", "dgx_batch_last_error", "dgx_batch_size", "dgx_batch_instance",
    "dgx_batch_allocate", "dgx_batch_step_in_time", "dgx_batch_synchronize", "dgx_batch_variant",
    "dgx_batch_tile", "dgx_batch_set_tile",
    "dgx_request_bytes", "dgx_request_n_dofs", "dgx_server_create", "dgx_server_destroy", "dgx_server_last_error",
    "dgx_server_submit", "dgx_server_step", "dgx_server_result", "dgx_server_release", "dgx_server_memory",
]


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("poly_degree", ctypes.c_int), ("flux_nodes", ctypes.c_int),
        ("use_split_form", ctypes.c_int), ("two_point_flux", ctypes.c_int), ("conv_flux", ctypes.c_int),
        ("pde", ctypes.c_int), ("diss_flux", ctypes.c_int), ("use_constant_viscosity", ctypes.c_int),
        ("gamma_gas", ctypes.c_double), ("mach_inf", ctypes.c_double), ("reynolds_inf", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("constant_viscosity", ctypes.c_double),
        ("temperature_inf", ctypes.c_double), ("sipg_penalty_factor", ctypes.c_double),
        ("n_cells_1d", ctypes.c_int), ("domain_left", ctypes.c_double), ("domain_right", ctypes.c_double),
        ("cell_order", ctypes.c_int), ("precision", ctypes.c_int),
        ("boundary_type", ctypes.c_int * 6), ("angle_of_attack", ctypes.c_double), ("grid_warp", ctypes.c_int),
        ("thermal_bc", ctypes.c_int), ("wall_temperature", ctypes.c_double),
    ]


class DeviceError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None


class Request(ctypes.Structure):
    """dgx_request: one parametric solve for the server."""
    _fields_ = [("params", Params), ("initial_condition", ctypes.c_int), ("tgv_isothermal", ctypes.c_int),
                ("vortex_strength", ctypes.c_double), ("vortex_u", ctypes.c_double), ("vortex_v", ctypes.c_double),
                ("final_time", ctypes.c_double), ("cfl", ctypes.c_double), ("wave_speed", ctypes.c_double)]


IC_TGV, IC_VORTEX = 0, 1


def request(final_time, cfl, wave_speed=1.0, ic=IC_TGV, vortex=(5.0, 0.0, 0.0), isothermal=False, **kw):
    r = Request()
    r.params = params(**kw)
    r.initial_condition = ic
    r.tgv_isothermal = int(isothermal)
    r.vortex_strength, r.vortex_u, r.vortex_v = vortex
    r.final_time, r.cfl, r.wave_speed = final_time, cfl, wave_speed
    return r


def load():
    """Load the in-tree device library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DeviceError(f"{LIB_PATH} is missing: build it with `make -C philip_amd` (no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    dp = P(d)
    i64 = ctypes.c_int64
    lib.dgx_default_params.argtypes = [P(Params)]
    lib.dgx_create.argtypes = [P(Params), P(vp)]
    lib.dgx_destroy.argtypes = [vp]
    lib.dgx_last_error.restype = ctypes.c_char_p
    lib.dgx_last_error.argtypes = [vp]
    lib.dgx_n_dofs.restype = i64
    lib.dgx_n_dofs.argtypes = [vp]
    lib.dgx_n_cells.argtypes = [vp]
    for fn in ["dgx_set_solution", "dgx_get_solution", "dgx_get_right_hand_side", "dgx_apply_inverse_mass_to_rhs",
               "dgx_set_grid_nodes", "dgx_get_grid_nodes", "dgx_get_metric_terms"]:
        getattr(lib, fn).argtypes = [vp, dp]
    lib.dgx_get_auxiliary_solution.argtypes = [vp, i, dp]
    lib.dgx_set_initial_condition_tgv.argtypes = [vp, i]
    lib.dgx_set_initial_condition_vortex.argtypes = [vp, d, d, d]
    lib.dgx_get_neighbors.argtypes = [vp, P(ctypes.c_int32)]
    lib.dgx_assemble_residual.argtypes = [vp]
    lib.dgx_assemble_residual_device.argtypes = [vp, vp, vp]
    lib.dgx_rhs_device.argtypes = [vp, vp, vp]
    lib.dgx_allocate_ode_system.argtypes = [vp, i]
    lib.dgx_step_in_time.argtypes = [vp, d, i]
    lib.dgx_current_time.restype = d
    lib.dgx_current_time.argtypes = [vp]
    lib.dgx_current_iteration.restype = i64
    lib.dgx_current_iteration.argtypes = [vp]
    lib.dgx_set_time.argtypes = [vp, d, i64]
    lib.dgx_set_track_error_estimate.argtypes = [vp, i]
    lib.dgx_constant_time_step.restype = d
    lib.dgx_constant_time_step.argtypes = [vp, d]
    lib.dgx_integrated_quantities.argtypes = [vp, i, dp]
    lib.dgx_numerical_entropy.argtypes = [vp, dp]
    lib.dgx_integrated_quantities_static.argtypes = [vp, vp, i, vp]
    lib.dgx_max_wave_speed.argtypes = [vp, i, dp]
    lib.dgx_adaptive_time_step.argtypes = [vp, d, dp]
    lib.dgx_set_stream.argtypes = [vp, vp]
    lib.dgx_enable_timing.argtypes = [vp, i]
    lib.dgx_kernel_timing.argtypes = [vp, dp, P(i64)]
    lib.dgx_synchronize.argtypes = [vp]
    lib.dgx_set_non_physical_behavior.argtypes = [vp, i]
    lib.dgx_check_state.argtypes = [vp]
    lib.dgx_nonphysical_events.restype = i64
    lib.dgx_nonphysical_events.argtypes = [vp]
    lib.dgx_set_error_adaptive_parameters.argtypes = [vp, d, d, d, d, d]
    lib.dgx_get_automatic_initial_step_size.argtypes = [vp, dp]
    lib.dgx_get_error_adaptive_step_size.argtypes = [vp, d, dp]
    lib.dgx_set_relaxation_runge_kutta.argtypes = [vp, i, d]
    lib.dgx_relaxation_parameter.restype = d
    lib.dgx_relaxation_parameter.argtypes = [vp]
    lib.dgx_apply_inverse_mass_device.argtypes = [vp, vp, vp]
    lib.dgx_batch_create.argtypes = [P(Params), i, P(vp)]
    lib.dgx_batch_destroy.argtypes = [vp]
    lib.dgx_batch_last_error.restype = ctypes.c_char_p
    lib.dgx_batch_last_error.argtypes = [vp]
    lib.dgx_batch_size.argtypes = [vp]
    lib.dgx_batch_instance.restype = vp
    lib.dgx_batch_instance.argtypes = [vp, i]
    lib.dgx_batch_allocate.argtypes = [vp, i, dp]
    lib.dgx_batch_step_in_time.argtypes = [vp, i]
    lib.dgx_batch_synchronize.argtypes = [vp]
    lib.dgx_batch_variant.argtypes = [vp, i]
    lib.dgx_batch_tile.argtypes = [vp, i, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
    lib.dgx_batch_set_tile.argtypes = [vp, i]
    lib.dgx_request_bytes.restype = i64
    lib.dgx_request_bytes.argtypes = [P(Request), i]
    lib.dgx_request_n_dofs.restype = i64
    lib.dgx_request_n_dofs.argtypes = [P(Request)]
    lib.dgx_server_create.argtypes = [i64, i, i, P(vp)]
    lib.dgx_server_destroy.argtypes = [vp]
    lib.dgx_server_last_error.restype = ctypes.c_char_p
    lib.dgx_server_last_error.argtypes = [vp]
    lib.dgx_server_submit.argtypes = [vp, P(Request), P(i64)]
    lib.dgx_server_step.argtypes = [vp, P(i), P(i), P(i)]
    lib.dgx_server_result.argtypes = [vp, i64, dp, dp, P(i64), dp]
    lib.dgx_server_release.argtypes = [vp, i64]
    lib.dgx_server_memory.argtypes = [vp, P(i64), P(i64)]
    _lib = lib
    return lib


def params(**kw):
    p = Params()
    load().dgx_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "boundary_type":
            v = (ctypes.c_int * 6)(*v)
        setattr(p, k, v)
    return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DG:
    """DGBase-shaped handle over the device context (one per HIP device/stream)."""

    def __init__(self, _handle=None, _params=None, **kw):
        lib = load()
        if _handle is not None:  # a context owned by a Batch
            self.p, self.h, self._owned = _params, ctypes.c_void_p(_handle), False
        else:
            self.p = params(**kw)
            h = ctypes.c_void_p()
            rc = lib.dgx_create(ctypes.byref(self.p), ctypes.byref(h))
            if rc != 0:
                raise DeviceError(f"dgx_create failed with code {rc}")
            self.h, self._owned = h, True
        self.dim = self.p.dim
        self.ns = self.dim + 2
        self.n = self.p.poly_degree + 1
        self.Np = self.n ** self.dim
        self.ndofs = lib.dgx_n_dofs(self.h)
        self.ncells = lib.dgx_n_cells(self.h)

    def _chk(self, rc, what):
        if rc != 0:
            raise DeviceError(f"{what}: code {rc}: {load().dgx_last_error(self.h).decode()}", rc)

    def set_non_physical_behavior(self, policy):
        """RETURN_BIG_NUMBER (reference default), ABORT_RUN (calls return ERR_NONPHYSICAL, raised
        here as DeviceError with .code == ERR_NONPHYSICAL) or PRINT_WARNING."""
        self._chk(load().dgx_set_non_physical_behavior(self.h, policy), "set_non_physical_behavior")

    def check_state(self):
        """Waits for the device, reads and clears the non-physical flag, applies the policy."""
        self._chk(load().dgx_check_state(self.h), "check_state")

    @property
    def nonphysical_events(self):
        return load().dgx_nonphysical_events(self.h)

    def right_hand_side(self):
        r = np.empty(self.ndofs)
        self._chk(load().dgx_get_right_hand_side(self.h, _dp(r)), "get_right_hand_side")
        return r

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                load().dgx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solution(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.ndofs:
            raise ValueError(f"solution: expected {self.ndofs} values, got {u.size}")
        self._chk(load().dgx_set_solution(self.h, _dp(u)), "set_solution")

    def set_initial_condition_tgv(self, isothermal=False):
        self._chk(load().dgx_set_initial_condition_tgv(self.h, int(isothermal)), "set_initial_condition_tgv")

    def set_initial_condition_vortex(self, strength=5.0, u_inf=None, v_inf=None):
        m = np.sqrt(2 / 1.4)
        u_inf = m * np.cos(np.pi / 4) if u_inf is None else u_inf
        v_inf = m * np.sin(np.pi / 4) if v_inf is None else v_inf
        self._chk(load().dgx_set_initial_condition_vortex(self.h, strength, u_inf, v_inf),
                  "set_initial_condition_vortex")

    def synchronize(self):
        """Wait for the context stream (a 1-element blocking copy)."""
        self._chk(load().dgx_synchronize(self.h), "synchronize")

    def solution(self):
        u = np.empty(self.ndofs)
        self._chk(load().dgx_get_solution(self.h, _dp(u)), "get_solution")
        return u

    def assemble_residual(self):
        self._chk(load().dgx_assemble_residual(self.h), "assemble_residual")
        r = np.empty(self.ndofs)
        self._chk(load().dgx_get_right_hand_side(self.h, _dp(r)), "get_right_hand_side")
        return r

    def apply_inverse_mass_to_rhs(self):
        o = np.empty(self.ndofs)
        self._chk(load().dgx_apply_inverse_mass_to_rhs(self.h, _dp(o)), "apply_inverse_mass")
        return o

    def auxiliary_solution(self, k):
        q = np.empty(self.ndofs)
        self._chk(load().dgx_get_auxiliary_solution(self.h, k, _dp(q)), "get_auxiliary_solution")
        return q

    def neighbors(self):
        a = np.empty((self.ncells, 2 * self.dim), dtype=np.int32)
        self._chk(load().dgx_get_neighbors(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "neighbors")
        return a

    def set_grid_nodes(self, xyz):
        """Curvilinear grid from node coordinates [cell][node][3] (grid degree = p)."""
        npn = self.ndofs // (self.ncells * (self.dim + 2))
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        if x.size != self.ncells * npn * 3:
            raise ValueError(f"node coordinates: expected {self.ncells} x {npn} x 3 values, got {x.size}")
        x = x.reshape(self.ncells, npn, 3)
        self._chk(load().dgx_set_grid_nodes(self.h, _dp(x)), "set_grid_nodes")

    def grid_nodes(self):
        """Node coordinates in use, [cell][node][3]."""
        x = np.empty((self.ncells, self.ndofs // (self.ncells * (self.dim + 2)), 3))
        self._chk(load().dgx_get_grid_nodes(self.h, _dp(x)), "grid_nodes")
        return x

    def metric_terms(self):
        """(C [cell][node][3][3] with C[m][n] = J dxi_n/dx_m, det J [cell][node]) of a curvilinear grid."""
        d = self.dim
        npn = self.ndofs // (self.ncells * (d + 2))
        raw = np.empty((self.ncells, d * d + 1, npn))
        self._chk(load().dgx_get_metric_terms(self.h, _dp(raw)), "metric_terms")
        C = np.zeros((self.ncells, npn, 3, 3))
        for m in range(d):
            for n in range(d):
                C[:, :, m, n] = raw[:, m * d + n, :]
        return C, raw[:, d * d, :].copy()

    def allocate_ode_system(self, method):
        m = METHODS[method] if isinstance(method, str) else method
        self._chk(load().dgx_allocate_ode_system(self.h, m), "allocate_ode_system")

    def step_in_time(self, dt, n_steps=1):
        self._chk(load().dgx_step_in_time(self.h, dt, n_steps), "step_in_time")

    def set_track_error_estimate(self, on=True):
        self._chk(load().dgx_set_track_error_estimate(self.h, int(on)), "set_track_error_estimate")

    @property
    def current_time(self):
        return load().dgx_current_time(self.h)

    @property
    def current_iteration(self):
        return load().dgx_current_iteration(self.h)

    def constant_time_step(self, cfl):
        return load().dgx_constant_time_step(self.h, cfl)

    def integrated(self, overint=10):
        o = np.zeros(4)
        self._chk(load().dgx_integrated_quantities(self.h, overint, _dp(o)), "integrated_quantities")
        return dict(kinetic_energy=o[0], enstrophy=o[1], incompressible_kinetic_energy=o[2],
                    vorticity_dissipation=o[3])

    def numerical_entropy(self):
        o = np.zeros(1)
        self._chk(load().dgx_numerical_entropy(self.h, _dp(o)), "numerical_entropy")
        return float(o[0])

    def max_wave_speed(self, overint=10):
        o = np.zeros(1)
        self._chk(load().dgx_max_wave_speed(self.h, overint, _dp(o)), "max_wave_speed")
        return float(o[0])

    def adaptive_time_step(self, cfl):
        o = np.zeros(1)
        self._chk(load().dgx_adaptive_time_step(self.h, cfl, _dp(o)), "adaptive_time_step")
        return float(o[0])

    def enable_timing(self, on=True):
        self._chk(load().dgx_enable_timing(self.h, int(on)), "enable_timing")

    def kernel_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(load().dgx_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)), "kernel_timing")
        return ms.value, n.value

    def set_time(self, time, iteration):
        self._chk(load().dgx_set_time(self.h, time, iteration), "set_time")

    def unsteady_row(self, numerical_entropy=False):
        """One row of the unsteady data table (periodic_turbulence.cpp:923-975): the quantities this
        device path computes, in the reference's column names."""
        q = self.integrated()
        row = {"time": self.current_time}
        if numerical_entropy:
            row["numerical_entropy"] = self.numerical_entropy()
        row["kinetic_energy"] = q["kinetic_energy"]
        row["enstrophy"] = q["enstrophy"]
        if self.p.pde == NAVIER_STOKES:
            row["eps_vorticity"] = q["vorticity_dissipation"]
        row["incompressible_kinetic_energy"] = q["incompressible_kinetic_energy"]
        return row

    def set_relaxation_runge_kutta(self, enable=True, root_tolerance=5e-10):
        self._chk(load().dgx_set_relaxation_runge_kutta(self.h, int(enable), root_tolerance),
                  "set_relaxation_runge_kutta")

    @property
    def relaxation_parameter(self):
        return load().dgx_relaxation_parameter(self.h)

    def set_error_adaptive_parameters(self, atol=1e-3, rtol=1e-3, beta=(0.70, -0.23, 0.0)):
        self._chk(load().dgx_set_error_adaptive_parameters(self.h, atol, rtol, *beta), "set_error_adaptive_parameters")

    def automatic_initial_step_size(self):
        o = ctypes.c_double()
        self._chk(load().dgx_get_automatic_initial_step_size(self.h, ctypes.byref(o)), "automatic_initial_step_size")
        return o.value

    def error_adaptive_step_size(self, dt):
        o = ctypes.c_double()
        self._chk(load().dgx_get_error_adaptive_step_size(self.h, dt, ctypes.byref(o)), "error_adaptive_step_size")
        return o.value

    def run(self, method, dt, final_time, adaptive_cfl=None, table=None, allocate=True, error_adaptive=None):
        """FlowSolver::run time loop (flow_solver.cpp:495-610): constant dt; or with adaptive_cfl the
        step from the device's maximum wave speed before every step (flow_solver.cpp:502-504,
        603-604); or with error_adaptive = dict(atol=, rtol=, beta=) (a low-storage method) the first
        step from get_automatic_initial_step_size and every later one from the embedded-error PID
        controller (flow_solver.cpp:505-507, 605-606).  The last step is clipped to the final time.
        table: list that receives one unsteady-data row per step.  allocate=False continues with the
        registers already allocated (a restarted run).  The steps taken are kept in self.last_dts."""
        if allocate:
            self.allocate_ode_system(method)
        if error_adaptive is not None:
            self.set_error_adaptive_parameters(**error_adaptive)
            next_dt = self.automatic_initial_step_size()
        t = self.current_time
        steps = 0
        self.last_dts = []
        while t < final_time:
            if error_adaptive is not None:
                step = next_dt
            else:
                step = self.adaptive_time_step(adaptive_cfl) if adaptive_cfl is not None else dt
            if t + step > final_time:
                step = final_time - t
            self.step_in_time(step, 1)
            self.last_dts.append(step)
            t = self.current_time
            steps += 1
            if table is not None:
                table.append(self.unsteady_row())
            if error_adaptive is not None:
                next_dt = self.error_adaptive_step_size(step)
        return steps

    def save_restart(self, path, table=None):
        """Restart data (flow_solver.cpp:365-426 writes the solution, the time and the data table):
        solution in the reference DoF order, time, iteration, the parameters, the table so far."""
        import json
        params = {f: getattr(self.p, f) for f, _ in Params._fields_}
        params = {k: (list(v) if isinstance(v, ctypes.Array) else v) for k, v in params.items()}
        cols = sorted({k for r in (table or []) for k in r})
        rows = np.array([[r.get(c, np.nan) for c in cols] for r in (table or [])], dtype=np.float64)
        np.savez(path, solution=self.solution(), time=np.float64(self.current_time),
                 iteration=np.int64(self.current_iteration), params=np.array(json.dumps(params)),
                 table_columns=np.array(cols, dtype=str), table_rows=rows.reshape(len(table or []), len(cols)))

    def load_restart(self, path):
        """Restart from save_restart data (flow_solver.cpp:58-100): checks the discretisation matches,
        restores the solution and the clock; returns the saved data-table rows."""
        import json
        with np.load(path, allow_pickle=False) as z:
            saved = json.loads(str(z["params"]))
            for k in ("dim", "poly_degree", "n_cells_1d", "domain_left", "domain_right", "boundary_type"):
                mine = getattr(self.p, k)
                mine = list(mine) if isinstance(mine, ctypes.Array) else mine
                if saved.get(k, mine) != mine:
                    raise DeviceError(f"restart data is for {k}={saved[k]}, this solver has {mine}")
            self.set_solution(z["solution"])
            self.set_time(float(z["time"]), int(z["iteration"]))
            cols = [str(c) for c in z["table_columns"]]
            return [dict(zip(cols, map(float, r))) for r in z["table_rows"]]


def write_unsteady_table(path, rows):
    """Unsteady data table as text (periodic_turbulence.cpp:971-975 writes a dealii::TableHandler):
    a header of column names, then one row per step, values in scientific notation with 16
    digits (flow_solver_case_base.cpp:229-237)."""
    cols = list(rows[0]) if rows else []
    with open(path, "w") as f:
        f.write(" ".join(cols) + "\n")
        for r in rows:
            f.write(" ".join(f"{r[c]:.16e}" for c in cols) + "\n")


def read_unsteady_table(path):
    with open(path) as f:
        cols = f.readline().split()
        return [dict(zip(cols, map(float, line.split()))) for line in f if line.strip()]


class Batch:
    """Independent solves advanced together: one HIP graph launch per RK step
    (dgx_batch_* in the C ABI).  ``instances[i]`` is a DG handle on instance i."""

    def __init__(self, param_dicts):
        lib = load()
        self.params = (Params * len(param_dicts))()
        for k, kw in enumerate(param_dicts):
            self.params[k] = params(**kw)
        h = ctypes.c_void_p()
        rc = lib.dgx_batch_create(self.params, len(param_dicts), ctypes.byref(h))
        if rc != 0:
            raise DeviceError(f"dgx_batch_create failed with code {rc}")
        self.h = h
        self.instances = [DG(_handle=lib.dgx_batch_instance(h, k), _params=self.params[k])
                          for k in range(len(param_dicts))]

    def _chk(self, rc, what):
        if rc != 0:
            raise DeviceError(f"{what}: code {rc}: {load().dgx_batch_last_error(self.h).decode()}")

    def allocate(self, method, dts):
        m = METHODS[method] if isinstance(method, str) else method
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        if dts.size != len(self.instances):
            raise ValueError(f"time steps: expected one per instance ({len(self.instances)}), got {dts.size}")
        self._chk(load().dgx_batch_allocate(self.h, m, _dp(dts)), "batch_allocate")

    def step_in_time(self, n_steps=1):
        self._chk(load().dgx_batch_step_in_time(self.h, n_steps), "batch_step_in_time")

    def synchronize(self):
        self._chk(load().dgx_batch_synchronize(self.h), "batch_synchronize")

    def variant(self, i):
        return load().dgx_batch_variant(self.h, i)

    def set_tile(self, tile):
        """Force candidate tile `tile` (-1: autotune) for the next allocate."""
        self._chk(load().dgx_batch_set_tile(self.h, tile), "batch_set_tile")

    def tile(self, i):
        """(elements per block, waves/SIMD bound, threads per block) of instance i's group; zeros if
        the instance is launched on its own."""
        e, w, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(load().dgx_batch_tile(self.h, i, ctypes.byref(e), ctypes.byref(w), ctypes.byref(t)), "batch_tile")
        return e.value, w.value, t.value

    def close(self):
        if getattr(self, "h", None):
            for d in self.instances:
                d.h = None
            load().dgx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with HBM-budget admission and continuous batching (dgx_server_* in the C ABI)."""

    def __init__(self, hbm_budget_bytes, max_active=64, method="rk4"):
        m = METHODS[method] if isinstance(method, str) else method
        h = ctypes.c_void_p()
        rc = load().dgx_server_create(int(hbm_budget_bytes), max_active, m, ctypes.byref(h))
        if rc != 0:
            raise DeviceError(f"dgx_server_create failed with code {rc}")
        self.h = h
        self.method = m
        self.ndofs = {}
        self.nonphysical = set()   # finished requests that met a non-physical state

    def _chk(self, rc, what):
        if rc != 0:
            raise DeviceError(f"{what}: code {rc}: {load().dgx_server_last_error(self.h).decode()}")

    def submit(self, req):
        i = ctypes.c_int64()
        self._chk(load().dgx_server_submit(self.h, ctypes.byref(req), ctypes.byref(i)), "server_submit")
        self.ndofs[i.value] = load().dgx_request_n_dofs(ctypes.byref(req))
        return i.value

    def step(self):
        a, q, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(load().dgx_server_step(self.h, ctypes.byref(a), ctypes.byref(q), ctypes.byref(d)), "server_step")
        return a.value, q.value, d.value

    def result(self, rid):
        """(solution, latency seconds, steps, DOF-updates) of a finished request, or None."""
        u = np.empty(self.ndofs[rid])
        lat, dofu = ctypes.c_double(), ctypes.c_double()
        st = ctypes.c_int64()
        rc = load().dgx_server_result(self.h, rid, _dp(u), ctypes.byref(lat), ctypes.byref(st), ctypes.byref(dofu))
        if rc == ERR_NONPHYSICAL:
            self.nonphysical.add(rid)
        elif rc != 0:
            return None
        return u, lat.value, st.value, dofu.value

    def release(self, rid):
        load().dgx_server_release(self.h, rid)

    def memory(self):
        """(bytes held by active requests, budget)."""
        u, b = ctypes.c_int64(), ctypes.c_int64()
        self._chk(load().dgx_server_memory(self.h, ctypes.byref(u), ctypes.byref(b)), "server_memory")
        return u.value, b.value

    def close(self):
        if getattr(self, "h", None):
            load().dgx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_bytes(req, method="rk4"):
    m = METHODS[method] if isinstance(method, str) else method
    return load().dgx_request_bytes(ctypes.byref(req), m)
