"""The five workload configurations, each checked end to end on the device through the C ABI.

* Configuration 1 (3D Euler TGV, p=2, 4^3, conservative strong DG, LxF, RK4, 100 steps): the
  device's kinetic-energy history against the FP64 oracle's at every step (bar 1e-12 relative)
  and the step-0 residual (per-DoF bar 1e-12 of the term magnitude).  The configuration's CFL 0.1
  diverges within five steps in the oracle as well (the reference's dt rule CFL*dx ignores the
  sound speed 1/M = 10), so the history is run at CFL 0.02.
* Configuration 3 (Navier-Stokes TGV, Re 1600, Pr 0.71, M 0.1, p=5, BR2, Ranocha split form,
  constant viscosity, RK4) on 4^3 cells: the FP64 device solution after three steps against the
  oracle (1e-12 relative), enstrophy and dissipation rate of the FP64 (1e-12) and mixed (1e-5)
  device paths against the oracle's.
* Configuration 4 (independent 2D vortex solves, N ~ U{8..64}, p ~ U{2..6}, strength ~ U[1,5],
  advection ~ U[0.5,1]^2, RK4, 200 steps, one graph launch per step): every sampled mixed-
  precision instance against its own FP64 solve (state within 1e-5 relative, kinetic energy within
  1e-6), and an FP64 instance at N = 32 against the oracle over 20 steps (1e-12).
* Configuration 5 (serving stream, 70 % configuration-4 vortices / 30 % 3D TGV with
  M ~ U[0.05,0.3], p ~ U{2..5}, 4^3..16^3 cells, final time ~ U[0.5,2], mixed and FP64): every
  request's result in the mix equals its solve alone bitwise, and an FP64 request equals the
  oracle's FlowSolver loop at 1e-12.  Final times are capped at 40 steps of each request's dt (a
  step cap stated here: the uncapped 3D requests take thousands of steps, and the solve alone is
  repeated for every request).
"""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

MACH_VORTEX = 1.195228609334  # sqrt(2/gamma): the vortex's p_inf = 1/(gamma M^2) = 1/2


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def state_rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_config1_lxf_rk4_kinetic_energy_history():
    kw = dict(dim=3, poly_degree=2, n_cells_1d=4, conv_flux=O.LXF, mach_inf=0.1)
    o, d = O.Oracle(**kw), P.DG(**kw)
    u = o.ic_tgv()
    d.set_initial_condition_tgv()
    assert state_rel(d.solution(), u) < 1e-15
    r_ref, mag = o.residual(u), o.residual_magnitude(u)
    r = d.assemble_residual()
    e0 = float((np.abs(r - r_ref) / np.maximum(mag, 1e-8 * mag.max())).max())
    dt = o.constant_time_step(0.02)
    d.allocate_ode_system("rk4")
    worst = 0.0
    for _ in range(100):
        o.step("rk4", dt, u)
        d.step_in_time(dt, 1)
        ke_o, ke_d = o.integrated(u)["kinetic_energy"], d.integrated()["kinetic_energy"]
        worst = max(worst, abs(ke_d - ke_o) / ke_o)
    es = state_rel(d.solution(), u)
    log(f"config 1: step-0 residual per-DoF err {e0:.2e}, KE history over 100 RK4 steps {worst:.2e}, "
        f"final state {es:.2e}, KE(100) {ke_o:.12f}")
    assert e0 <= 1e-12 and worst <= 1e-12 and es <= 1e-12


def ns_cfg3(**extra):
    return dict(dim=3, poly_degree=5, n_cells_1d=4, pde=O.NAVIER_STOKES, diss_flux=O.BR2, use_constant_viscosity=1,
                constant_viscosity=1.0, reynolds_inf=1600.0, prandtl=0.71, mach_inf=0.1, use_split_form=1,
                two_point_flux=O.RA, conv_flux=O.TWO_POINT_ROE, **extra)


def test_config3_navier_stokes_p5_fp64_and_mixed():
    o = O.Oracle(**ns_cfg3())
    d64, dmx = P.DG(**ns_cfg3(precision=P.FP64)), P.DG(**ns_cfg3(precision=P.MIXED))
    u = o.ic_tgv()
    for d in (d64, dmx):
        d.set_solution(u)
        d.allocate_ode_system("rk4")
    dt = o.constant_time_step(0.01)
    w64 = wmx = 0.0
    for step in range(3):
        o.step("rk4", dt, u)
        for d in (d64, dmx):
            d.step_in_time(dt, 1)
        qo, q64, qmx = o.integrated(u), d64.integrated(), dmx.integrated()
        for k in ("enstrophy", "vorticity_dissipation"):
            w64 = max(w64, abs(q64[k] - qo[k]) / qo[k])
            wmx = max(wmx, abs(qmx[k] - qo[k]) / qo[k])
    es = state_rel(d64.solution(), u)
    log(f"config 3 (4^3): FP64 state after 3 RK4 steps {es:.2e}, enstrophy/dissipation FP64 {w64:.2e}, "
        f"mixed {wmx:.2e}")
    assert es <= 1e-12 and w64 <= 1e-12 and wmx <= 1e-5


def vortex_instances(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kw = dict(dim=2, poly_degree=int(rng.integers(2, 7)), n_cells_1d=int(rng.integers(8, 65)),
                  domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX, use_split_form=1,
                  two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE)
        out.append((kw, (float(rng.uniform(1, 5)), *(float(x) for x in rng.uniform(0.5, 1.0, 2)))))
    return out


def test_config4_mixed_instances_against_their_fp64_solves():
    cases = vortex_instances(2024, 24)
    b = P.Batch([dict(kw, precision=P.MIXED) for kw, _ in cases])
    dts = []
    for d, (kw, ic) in zip(b.instances, cases):
        d.set_initial_condition_vortex(*ic)
        dts.append(d.constant_time_step(0.05) / 2.0)
    b.allocate("rk4", dts)
    steps = 200
    b.step_in_time(steps)
    b.synchronize()
    # sample: the largest instance, the highest degree, and every fourth one
    big = max(range(len(cases)), key=lambda k: cases[k][0]["n_cells_1d"] ** 2 * (cases[k][0]["poly_degree"] + 1) ** 2)
    top = max(range(len(cases)), key=lambda k: cases[k][0]["poly_degree"])
    sample = sorted({big, top, *range(0, len(cases), 4)})
    worst_u = worst_ke = 0.0
    for k in sample:
        kw, ic = cases[k]
        s = P.DG(**dict(kw, precision=P.FP64))
        s.set_initial_condition_vortex(*ic)
        s.allocate_ode_system("rk4")
        s.step_in_time(dts[k], steps)
        um, u64 = b.instances[k].solution(), s.solution()
        eu = state_rel(um, u64)
        ke_m, ke_64 = b.instances[k].integrated()["kinetic_energy"], s.integrated()["kinetic_energy"]
        eke = abs(ke_m - ke_64) / ke_64
        log(f"  instance {k}: N={kw['n_cells_1d']} p={kw['poly_degree']} mixed vs FP64 after {steps} steps: "
            f"state {eu:.2e}, KE {eke:.2e}")
        worst_u, worst_ke = max(worst_u, eu), max(worst_ke, eke)
        s.close()
    b.close()
    log(f"config 4: {len(sample)} sampled mixed instances vs their FP64 solves: state {worst_u:.2e}, "
        f"KE {worst_ke:.2e}")
    assert worst_u <= 1e-5 and worst_ke <= 1e-6


def test_config4_fp64_instance_at_n32_matches_oracle():
    kw, ic = vortex_instances(7, 1)[0]
    kw = dict(kw, n_cells_1d=32, poly_degree=4)
    b = P.Batch([dict(kw, precision=P.FP64), dict(kw, precision=P.MIXED, n_cells_1d=16)])
    for d in b.instances:
        d.set_initial_condition_vortex(*ic)
    o = O.Oracle(**kw)
    u = o.ic_vortex(strength=ic[0], u_inf=ic[1], v_inf=ic[2])
    dt = o.vortex_time_step(0.05, 2.0)
    b.allocate("rk4", [dt, dt])
    b.step_in_time(20)
    for _ in range(20):
        o.step("rk4", dt, u)
    e = state_rel(b.instances[0].solution(), u)
    log(f"config 4: FP64 instance N=32 p=4 in a batch vs the oracle after 20 RK4 steps: {e:.2e}")
    assert e <= 1e-12
    b.close()


STEP_CAP = 40


def serving_requests(seed, n):
    """Configuration 5's mix; final time U[0.5, 2] capped at STEP_CAP steps of the request's dt."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        prec = P.MIXED if k % 2 == 0 else P.FP64
        tf = float(rng.uniform(0.5, 2.0))
        if rng.uniform() < 0.7:
            kw = dict(dim=2, poly_degree=int(rng.integers(2, 7)), n_cells_1d=int(rng.integers(8, 65)),
                      domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX, use_split_form=1,
                      two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            spec = dict(cfl=0.05, wave_speed=2.0, ic=P.IC_VORTEX,
                        vortex=(float(rng.uniform(1, 5)), *(float(x) for x in rng.uniform(0.5, 1.0, 2))))
        else:
            m = float(rng.uniform(0.05, 0.3))
            kw = dict(dim=3, poly_degree=int(rng.integers(2, 6)), n_cells_1d=int(rng.choice([4, 8, 16])),
                      mach_inf=m, use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            spec = dict(cfl=0.1, wave_speed=1.0 + 1.0 / m, ic=P.IC_TGV)
        probe = P.DG(**kw)
        dt = probe.constant_time_step(spec["cfl"]) / spec["wave_speed"]
        probe.close()
        out.append((kw, dict(spec, final_time=min(tf, STEP_CAP * dt))))
    return out


def solve_alone(kw, spec, method="rk4"):
    d = P.DG(**kw)
    if spec["ic"] == P.IC_VORTEX:
        d.set_initial_condition_vortex(*spec["vortex"])
    else:
        d.set_initial_condition_tgv()
    dt = d.constant_time_step(spec["cfl"]) / spec["wave_speed"]
    steps = d.run(method, dt, spec["final_time"])
    u = d.solution()
    d.close()
    return u, steps


def test_config5_serving_mix_equals_alone_and_fp64_request_matches_oracle():
    reqs = serving_requests(77, 12)
    # one small FP64 request checked against the oracle's FlowSolver loop
    okw = dict(dim=2, poly_degree=3, n_cells_1d=8, domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX,
               use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE)
    o = O.Oracle(**okw)
    odt = o.vortex_time_step(0.05, 2.0)
    ospec = dict(cfl=0.05, wave_speed=2.0, ic=P.IC_VORTEX, vortex=(3.0, 0.6, 0.9), final_time=25.5 * odt)
    reqs.append((dict(okw, precision=P.FP64), ospec))
    s = P.Server(32 << 30, max_active=64)
    ids = [s.submit(P.request(**spec, **kw)) for kw, spec in reqs[:6]]
    for _ in range(3):
        s.step()
    ids += [s.submit(P.request(**spec, **kw)) for kw, spec in reqs[6:]]  # arrive while others run
    max_active = 0
    while True:
        a, q, done = s.step()
        max_active = max(max_active, a)
        if done == len(reqs):
            break
    res = [s.result(i) for i in ids]
    s.close()
    assert max_active > 6
    for k, ((kw, spec), r) in enumerate(zip(reqs, res)):
        u_alone, steps = solve_alone(kw, spec)
        log(f"  request {k}: dim={kw['dim']} p={kw['poly_degree']} N={kw['n_cells_1d']} prec={kw['precision']} "
            f"steps={r[2]} max|mix-alone|={np.abs(r[0] - u_alone).max():.1e}")
        assert r[2] == steps and np.array_equal(r[0], u_alone), k
    u = o.ic_vortex(strength=3.0, u_inf=0.6, v_inf=0.9)
    n_o, _ = o.run("rk4", odt, ospec["final_time"], u)
    e = state_rel(res[-1][0], u)
    log(f"config 5: {len(reqs)} requests (up to {max_active} active), FP64 request vs oracle "
        f"({n_o} steps, last clipped): {e:.2e}")
    assert res[-1][2] == n_o == 26 and e <= 1e-12
