"""Serving loop (dgx_server_*): requests admitted by HBM budget, stages of all
active requests batched together, continuous admission while others run.
Bar: every request's result in the mix equals the same request solved alone
(a single DG context running FlowSolver::run's loop: constant dt, last step
clipped to the final time) bitwise, also under a budget that holds only two
requests at a time; requests that can never fit are rejected at submission."""
import sys

import numpy as np
import pytest

import philip_amd as P

pytestmark = pytest.mark.gpu

MACH_VORTEX = 1.195228609334


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_requests(seed, n):
    """70% 2D isentropic vortex (uneven N, p, precision), 30% 3D TGV (M in [0.05, 0.3])."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        prec = P.MIXED if k % 2 == 0 else P.FP64
        if rng.uniform() < 0.7:
            N, p = int(rng.integers(4, 9)), int(rng.integers(2, 6))
            vort = (float(rng.uniform(1, 5)), *(float(x) for x in rng.uniform(0.5, 1.0, 2)))
            kw = dict(dim=2, poly_degree=p, n_cells_1d=N, domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX,
                      use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            out.append((kw, dict(final_time=float(rng.uniform(0.02, 0.05)), cfl=0.02, wave_speed=2.0,
                                 ic=P.IC_VORTEX, vortex=vort)))
        else:
            p = int(rng.integers(2, 4))
            kw = dict(dim=3, poly_degree=p, n_cells_1d=4, mach_inf=float(rng.uniform(0.05, 0.3)), use_split_form=1,
                      two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            out.append((kw, dict(final_time=float(rng.uniform(0.01, 0.03)), cfl=0.01, wave_speed=1.0, ic=P.IC_TGV)))
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


def serve(reqs, budget, max_active=64, stagger=True, trace=None):
    s = P.Server(budget, max_active=max_active)
    ids = []
    half = len(reqs) // 2 if stagger else len(reqs)
    for kw, spec in reqs[:half]:
        ids.append(s.submit(P.request(**spec, **kw)))
    for _ in range(2):
        s.step()
    for kw, spec in reqs[half:]:  # arrive while the first half is on the GPU
        ids.append(s.submit(P.request(**spec, **kw)))
    max_seen = 0
    while True:
        a, q, d = s.step()
        max_seen = max(max_seen, a)
        if trace is not None:
            trace.append((a, q, *s.memory()))
        if d == len(reqs):
            break
    res = [s.result(i) for i in ids]
    s.close()
    return res, max_seen


def test_results_equal_alone_and_in_the_mix():
    reqs = make_requests(5, 10)
    res, max_active = serve(reqs, budget=8 << 30)
    log(f"mix of {len(reqs)} requests, up to {max_active} active at once")
    assert max_active > len(reqs) // 2  # the second half joined while the first half was running
    for k, ((kw, spec), r) in enumerate(zip(reqs, res)):
        u_alone, steps = solve_alone(kw, spec)
        log(f"  request {k}: dim={kw['dim']} p={kw['poly_degree']} N={kw['n_cells_1d']} prec={kw['precision']} "
            f"steps={r[2]} latency={r[1] * 1e3:.1f} ms max|diff|={np.abs(r[0] - u_alone).max():.2e}")
        assert r[2] == steps
        assert np.array_equal(r[0], u_alone), k


def test_budget_admission_throttles():
    reqs = make_requests(9, 6)
    sizes = [P.request_bytes(P.request(**spec, **kw)) for kw, spec in reqs]
    budget = max(sizes) + min(sizes)  # never room for all six
    trace = []
    res, max_active = serve(reqs, budget=budget, stagger=False, trace=trace)
    log(f"budget {budget} B (request sizes {sorted(sizes)}): at most {max_active} active")
    assert all(used <= b for _, _, used, b in trace)
    assert any(a > 0 and q > 0 for a, q, _, _ in trace)  # requests waited for memory
    ref, _ = serve(reqs, budget=8 << 30, stagger=False)
    for a, b in zip(res, ref):
        assert np.array_equal(a[0], b[0])


def test_request_larger_than_budget_is_rejected():
    kw, spec = make_requests(1, 1)[0]
    r = P.request(**spec, **kw)
    s = P.Server(P.request_bytes(r) - 1)
    with pytest.raises(P.DeviceError):
        s.submit(r)
    s.close()
