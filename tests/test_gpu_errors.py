"""Error contract of the boundary on the device: a non-physical stage state (negative pressure)
is detected by the stage kernels and handled by the reference's non_physical_behavior policies
(physics.cpp:298-317): return_big_number counts it and returns success, print_warning warns,
abort_run returns DGX_ERR_NONPHYSICAL and, for assemble_residual, fills the right-hand side with
ones as the reference does after an invalid assembly (dg_base.cpp:2850-2875).  The adaptive time
step refuses a non-finite maximum wave speed.  The offending state is ordinary data: the kernels
compute NaNs from it and exit normally."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SCHEMES = [
    ("ch_roe", dict(use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE)),
    ("lxf", dict(conv_flux=P.LXF)),
]


def bad_state(o, whole_cell=False):
    """TGV with the total energy of one node (or of every node of one cell) lowered so that the
    pressure there is negative."""
    u = o.ic_tgv().reshape(o.ncells, o.ns, o.Np)
    nodes = slice(None) if whole_cell else 7
    u[5, o.ns - 1, nodes] = 0.5 * (u[5, 1:4, nodes] ** 2).sum(axis=0) / u[5, 0, nodes] - 1.0
    return u.ravel()


@pytest.mark.parametrize("prec", [P.FP64, P.MIXED])
@pytest.mark.parametrize("name,scheme", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_negative_pressure_policies(prec, name, scheme):
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, mach_inf=0.1, **scheme)
    o = O.Oracle(**kw)
    good, bad = o.ic_tgv(), bad_state(o)
    d = P.DG(precision=prec, **kw)
    log(f"non-physical state: {name} precision={prec}")
    # reference default: counted, no error
    d.set_solution(bad)
    d.assemble_residual()
    assert d.nonphysical_events == 1
    d.set_solution(good)
    d.assemble_residual()
    assert d.nonphysical_events == 1  # the flag was cleared by the previous check
    # print_warning: no error either
    d.set_non_physical_behavior(P.PRINT_WARNING)
    d.set_solution(bad)
    d.assemble_residual()
    assert d.nonphysical_events == 2
    # abort_run: the call fails and the right-hand side is ones
    d.set_non_physical_behavior(P.ABORT_RUN)
    d.set_solution(bad)
    with pytest.raises(P.DeviceError) as ei:
        d.assemble_residual()
    assert ei.value.code == P.ERR_NONPHYSICAL
    assert np.array_equal(d.right_hand_side(), np.ones(d.ndofs))
    # a physical state assembles normally afterwards
    d.set_solution(good)
    r = d.assemble_residual()
    assert np.isfinite(r).all()
    # stepping: reported once per call, after its steps
    d.set_solution(bad)
    d.allocate_ode_system("rk4")
    with pytest.raises(P.DeviceError) as ei:
        d.step_in_time(1e-4, 2)
    assert ei.value.code == P.ERR_NONPHYSICAL
    d.set_solution(good)
    d.step_in_time(1e-4, 2)
    # the adaptive step refuses the NaN wave speed of a cell whose interpolated pressure is
    # negative at the quadrature points, under any policy
    d.set_non_physical_behavior(P.RETURN_BIG_NUMBER)
    d.set_solution(bad_state(o, whole_cell=True))
    with pytest.raises(P.DeviceError) as ei:
        d.adaptive_time_step(0.1)
    assert ei.value.code == P.ERR_NONPHYSICAL
    d.close()


def test_device_pointer_entry_points_report_through_check_state():
    """dgx_rhs_device does not wait for the device: its flag is read by dgx_check_state."""
    from hipmem import DeviceBuffer
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, mach_inf=0.1, use_split_form=1, two_point_flux=P.RA,
              conv_flux=P.TWO_POINT_LXF)
    o = O.Oracle(**kw)
    d = P.DG(**kw)
    d.set_non_physical_behavior(P.ABORT_RUN)
    u = DeviceBuffer.from_array(bad_state(o))
    out = DeviceBuffer(u.n)
    assert P.load().dgx_rhs_device(d.h, u.value, out.value) == 0
    with pytest.raises(P.DeviceError) as ei:
        d.check_state()
    assert ei.value.code == P.ERR_NONPHYSICAL
    u.upload(o.ic_tgv())
    assert P.load().dgx_rhs_device(d.h, u.value, out.value) == 0
    d.check_state()
    assert d.nonphysical_events == 1 and np.isfinite(out.download()).all()


def test_batch_reports_the_instance_policy():
    kw = dict(dim=3, poly_degree=3, n_cells_1d=4, mach_inf=0.1, use_split_form=1, two_point_flux=P.CH,
              conv_flux=P.TWO_POINT_ROE)
    o = O.Oracle(**kw)
    b = P.Batch([kw, kw])
    b.instances[0].set_solution(o.ic_tgv())
    b.instances[1].set_solution(bad_state(o))
    b.instances[1].set_non_physical_behavior(P.ABORT_RUN)
    b.allocate("rk4", [1e-4, 1e-4])
    b.step_in_time(1)
    with pytest.raises(P.DeviceError):
        b.synchronize()
    assert b.instances[0].nonphysical_events == 0 and b.instances[1].nonphysical_events == 1
    b.close()
