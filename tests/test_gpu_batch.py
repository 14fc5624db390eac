"""Batched independent solves (dgx_batch_*): many instances of uneven element
count, polynomial degree and precision advanced by one HIP graph launch per RK
step.  Bar: every instance after the batched steps equals the same instance
solved alone with dgx_step_in_time (bitwise: the batched kernel runs the single
solve's arithmetic, and the owner-rule face flux is the same whichever block
evaluates it); FP64 instances also match the FP64 oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import philip_amd as P

pytestmark = pytest.mark.gpu

MACH_VORTEX = 1.195228609334  # sqrt(2/gamma)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return max(np.abs(a - b).max() / np.abs(b).max(), np.linalg.norm(a - b) / np.linalg.norm(b))


def vortex_cases(seed, n):
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(n):
        N = int(rng.integers(4, 11))
        p = int(rng.integers(2, 7))
        prec = P.MIXED if k % 2 == 0 else P.FP64
        beta = float(rng.uniform(1, 5))
        u, v = (float(x) for x in rng.uniform(0.5, 1.0, 2))
        kw = dict(dim=2, poly_degree=p, n_cells_1d=N, domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX,
                  use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
        cases.append((kw, (beta, u, v)))
    return cases


def setup(d, kw, ic):
    if kw["dim"] == 2:
        d.set_initial_condition_vortex(*ic)
    else:
        d.set_initial_condition_tgv()


def test_batch_matches_single_solves():
    cases = vortex_cases(7, 10)
    # one conservative instance (launched on its own inside the graph) and one 3D instance
    cases.append((dict(dim=2, poly_degree=3, n_cells_1d=6, domain_left=-5.0, domain_right=5.0,
                       mach_inf=MACH_VORTEX, conv_flux=P.LXF), (3.0, 0.7, 0.8)))
    cases.append((dict(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=P.CH,
                       conv_flux=P.TWO_POINT_ROE, precision=P.MIXED), None))
    # long 3D lines (shared by two or three items; the FP64 p=6 block holds ~135 KB of LDS), 4^3
    # cells so that the -x and +x neighbours of a cell differ
    for p_, prec in ((4, P.FP64), (5, P.FP64), (5, P.MIXED), (6, P.FP64), (6, P.MIXED)):
        cases.append((dict(dim=3, poly_degree=p_, n_cells_1d=4, use_split_form=1, two_point_flux=P.CH,
                           conv_flux=P.TWO_POINT_ROE, precision=prec), None))
    # solves outside the batched line kernel, run through their own kernels inside the graph: the
    # curved periodic grid (Euler and Navier-Stokes), boundary faces, Navier-Stokes walls
    tgv_box = dict(domain_left=0.0, domain_right=2 * np.pi)
    cases.append((dict(dim=3, poly_degree=3, n_cells_1d=4, grid_warp=1, use_split_form=1, two_point_flux=P.CH,
                       conv_flux=P.TWO_POINT_ROE, precision=P.FP64, **tgv_box), None))
    cases.append((dict(dim=3, poly_degree=2, n_cells_1d=4, grid_warp=1, pde=P.NAVIER_STOKES, reynolds_inf=1600.0,
                       use_split_form=1, two_point_flux=P.RA, conv_flux=P.TWO_POINT_ROE, precision=P.MIXED,
                       **tgv_box), None))
    cases.append((dict(dim=2, poly_degree=3, n_cells_1d=6, domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX,
                       boundary_type=[1004, 1004, 1006, 1006], use_split_form=1, two_point_flux=P.CH,
                       conv_flux=P.TWO_POINT_ROE, precision=P.MIXED), (3.0, 0.7, 0.8)))
    cases.append((dict(dim=3, poly_degree=3, n_cells_1d=4, pde=P.NAVIER_STOKES, reynolds_inf=1600.0,
                       boundary_type=[0, 0, 1001, 1001, 1006, 1006], thermal_bc=1, wall_temperature=1.1,
                       use_split_form=1, two_point_flux=P.RA, conv_flux=P.TWO_POINT_ROE, precision=P.FP64,
                       **tgv_box), None))
    log(f"batch of {len(cases)} instances")
    b = P.Batch([kw for kw, _ in cases])
    dts = []
    for d, (kw, ic) in zip(b.instances, cases):
        setup(d, kw, ic)
        h = (kw.get("domain_right", 2 * np.pi) - kw.get("domain_left", 0.0)) / kw["n_cells_1d"]
        dts.append(0.01 * h / (kw["poly_degree"] + 1))
    b.allocate("rk4", dts)
    steps = 5
    b.step_in_time(steps)
    b.synchronize()
    for k, (d, (kw, ic)) in enumerate(zip(b.instances, cases)):
        s = P.DG(**kw)
        setup(s, kw, ic)
        s.allocate_ode_system("rk4")
        s.step_in_time(dts[k], steps)
        ub, us = d.solution(), s.solution()
        log(f"  instance {k}: dim={kw['dim']} p={kw['poly_degree']} N={kw['n_cells_1d']} "
            f"prec={kw.get('precision', 0)} variant={b.variant(k)} max|diff|={np.abs(ub - us).max():.3e}")
        assert np.array_equal(ub, us), k
        assert d.current_iteration == steps and abs(d.current_time - steps * dts[k]) < 1e-15
        s.close()
    b.close()


def test_batch_fp64_instance_matches_oracle():
    kw, ic = vortex_cases(11, 3)[1]
    kw = dict(kw, precision=P.FP64)
    b = P.Batch([kw, dict(kw, n_cells_1d=5, poly_degree=4)])
    d = b.instances[0]
    d.set_initial_condition_vortex(*ic)
    b.instances[1].set_initial_condition_vortex(*ic)
    okw = {k: v for k, v in kw.items() if k != "precision"}
    o = O.Oracle(**okw)
    u = o.ic_vortex(strength=ic[0], u_inf=ic[1], v_inf=ic[2])
    assert rel(d.solution(), u) < 1e-14
    dt = 0.002
    b.allocate("rk4", [dt, dt])
    b.step_in_time(3)
    for _ in range(3):
        o.step("rk4", dt, u)
    e = np.abs(d.solution() - u).max() / np.abs(u).max()
    log(f"  batch FP64 vs oracle rel err {e:.3e} (bar 1e-12)")
    assert e <= 1e-12
    b.close()


def test_every_tile_gives_the_single_solve_bits():
    """Each candidate tile of the autotuner (elements per block x waves bound x threads per block),
    forced in turn, on 2D instances of every degree at both precisions and 3D p=3 (which has the
    first two tiles): 3 RK4 steps equal the single solves bit for bit."""
    cases = []
    for p_ in (2, 3, 4, 5, 6):
        for prec in (P.MIXED, P.FP64):
            cases.append((dict(dim=2, poly_degree=p_, n_cells_1d=7, domain_left=-5.0, domain_right=5.0,
                               mach_inf=MACH_VORTEX, use_split_form=1, two_point_flux=P.CH,
                               conv_flux=P.TWO_POINT_ROE, precision=prec), (3.0, 0.7, 0.8)))
    cases.append((dict(dim=3, poly_degree=3, n_cells_1d=4, use_split_form=1, two_point_flux=P.CH,
                       conv_flux=P.TWO_POINT_ROE, precision=P.MIXED), None))
    dts = [0.002] * len(cases)
    alone = []
    for (kw, ic), dt in zip(cases, dts):
        s = P.DG(**kw)
        setup(s, kw, ic)
        s.allocate_ode_system("rk4")
        s.step_in_time(dt, 3)
        alone.append(s.solution())
        s.close()
    for tile in range(12):
        b = P.Batch([kw for kw, _ in cases])
        b.set_tile(tile)
        for d, (kw, ic) in zip(b.instances, cases):
            setup(d, kw, ic)
        b.allocate("rk4", dts)
        b.step_in_time(3)
        b.synchronize()
        shapes = set()
        for k, d in enumerate(b.instances):
            shapes.add(b.tile(k))
            assert np.array_equal(d.solution(), alone[k]), (tile, k)
        log(f"  tile {tile}: (elements/block, waves bound, threads/block) {sorted(shapes)}: all equal")
        b.close()


def test_batch_rejects_one_stage_methods_and_locks_registers():
    kw, ic = vortex_cases(3, 1)[0]
    b = P.Batch([kw])
    b.instances[0].set_initial_condition_vortex(*ic)
    with pytest.raises(P.DeviceError):
        b.allocate("euler", [1e-3])
    b.allocate("ssprk3", [1e-3])
    with pytest.raises(P.DeviceError):
        b.instances[0].allocate_ode_system("rk4")
    b.close()
