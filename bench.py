"""Benchmark of the explicit strong-form DG RK stage on one MI355X.

Workload (the configuration the metric is quoted on): 3D Euler Taylor-Green
vortex, p=3, 32^3 periodic hex cells (2,097,152 nodes x 5 states =
10,485,760 DOF), Chandrashekar split-form volume flux with entropy-conserving
+ Roe (Harten) surface flux, collocated GLL nodes, low-storage RK3(2)5F[3S*+]
(5 stages per step).  The state and all RK registers are resident in HBM before
the timed region; a "step" is one RK step of the whole mesh.

Time step: the reference's TGV rule dt = CFL*L/(n_dofs/nstate)^(1/3) with CFL 0.01
(the rule ignores the sound speed 1/M = 10; CFL 0.01 is an acoustic CFL of ~0.14,
CFL 0.1 diverges within ten steps in the FP64 oracle as well).

Metric: DOF-updates/s per RK stage = DOF x stages x steps / wall time.  The
headline `value` is the FP64 path (the reference's precision: every flux and the
state in double; the entropy round trip u(v(u)) is the identity on collocated GLL
nodes and is not formed).  The mixed path (FP32 fluxes and
contractions, FP64 state, RK registers and reductions) is reported beside it in
`mixed_path`, with its own roofline and its error against the FP64 path.
Roofline: HBM-bound; algorithmic bytes per stage launch (FP64 registers of the
fused 3S* stage) = read S1,S2,S3 + write S1',S2 = 40 B/DOF (32 B/DOF for the
first stage, which writes S3 = S1 and starts S2 from zero).  Both paths move the
same algorithmic bytes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KERNEL = {"fp64": "line_kernel_wb<3,4,double,3>", "mixed": "line_kernel<3,4,float>"}
TRAFFIC_FILE = {"fp64": os.path.join(ROOT, "profiles", "line_kernel_fp64_tgv_p3_32cubed_traffic.txt"),
                "mixed": os.path.join(ROOT, "profiles", "line_kernel_mixed_tgv_p3_32cubed_traffic.txt")}


def measured_traffic(precision):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3 counter summary
    (read requests by size + WRITE_SIZE, calibrated on a plain copy), or None."""
    try:
        with open(TRAFFIC_FILE[precision]) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def progress(msg):
    """One line per finished part of the run, on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def workload_params(precision):
    import philip_amd as P
    return dict(dim=3, poly_degree=3, n_cells_1d=32, use_split_form=1, two_point_flux=P.CH,
                conv_flux=P.TWO_POINT_ROE, pde=P.EULER, mach_inf=0.1, precision=precision)


def host_cores():
    """CPUs this process may run on (the box's share) and the host's total."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def cpu_baseline(threads, min_seconds=10.0, max_steps=8):
    """Oracle (FP64 C restatement of the reference loop structure, OpenMP over cells) timed on
    the host: RK3(2)5F[3S*+] steps of the full workload mesh, 5 stage residuals per step (the FSAL
    residual of the embedded estimate is skipped, as the device path skips it at a fixed time
    step), repeated until at least `min_seconds` of CPU time have passed."""
    import oracle as O
    O.set_threads(threads)
    o = O.Oracle(dim=3, poly_degree=3, n_cells_1d=32, use_split_form=1, two_point_flux=O.CH,
                 conv_flux=O.TWO_POINT_ROE, mach_inf=0.1)
    o.set_fsal(False)
    u = o.ic_tgv()
    dt = o.constant_time_step(0.01)
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and (n == 0 or time.perf_counter() - t0 < min_seconds):
        o.step("rk3_2_5f_3sstarplus", dt, u)
        n += 1
    t = time.perf_counter() - t0
    usable, total = host_cores()
    return dict(value=o.ndofs * 5 * n / t, unit="DOF-updates/s", cores=threads, kind="port",
                host_cpus_usable=usable, host_cpu_count=total,
                sample=f"{n} RK3(2)5F[3S*+] step(s) (5 stage residuals each, no FSAL residual) of the 32^3 p=3 "
                       f"mesh, {t:.1f} s on {threads} OpenMP threads")


def run_gpu(precision, steps, warmup):
    import philip_amd as P

    prec = P.FP64 if precision == "fp64" else P.MIXED
    d = P.DG(**workload_params(prec))
    d.set_initial_condition_tgv()
    method = "rk3_2_5f_3sstarplus"
    stages = 5
    dt = d.constant_time_step(0.01)
    d.allocate_ode_system(method)
    d.step_in_time(dt, warmup)
    d.solution()  # blocking copy: synchronizes the context stream
    t0 = time.perf_counter()
    d.step_in_time(dt, steps)
    d.synchronize()
    t = time.perf_counter() - t0
    # per-launch kernel time with HIP events on the context stream
    d.enable_timing(True)
    d.step_in_time(dt, 2)
    avg_ms, _ = d.kernel_timing()
    d.enable_timing(False)
    # host-buffer boundary: solution copied in, one RK step, solution copied out (PCIe included)
    u = d.solution()
    reps = 3
    t1 = time.perf_counter()
    for _ in range(reps):
        d.set_solution(u)
        d.step_in_time(dt, 1)
        u = d.solution()
    pcie = d.ndofs * stages * reps / (time.perf_counter() - t1)
    ke = d.integrated()["kinetic_energy"]
    return dict(value=d.ndofs * stages * steps / t, ms_per_step=1e3 * t / steps, avg_ms=avg_ms,
                ndofs=d.ndofs, stages=stages, ke=ke, pcie=pcie)


def tgv_history(steps=1000, every=50):
    """The headline workload (Euler TGV p=3 32^3, CH split + EC/Roe, RK3(2)5F[3S*+], CFL 0.01) run
    for `steps` steps at both precisions; kinetic energy and numerical entropy every `every` steps.
    Reports the largest relative difference of the mixed KE history from the FP64 one, each run's
    entropy drift |S(t) - S(0)| / |S(0)| and the mixed run's entropy drift from the FP64 one."""
    import numpy as np
    import philip_amd as P
    hist, out = {}, {}
    for prec in ("fp64", "mixed"):
        d = P.DG(**workload_params(P.FP64 if prec == "fp64" else P.MIXED))
        d.set_initial_condition_tgv()
        dt = d.constant_time_step(0.01)
        d.allocate_ode_system("rk3_2_5f_3sstarplus")
        rows = [(0.0, d.integrated()["kinetic_energy"], d.numerical_entropy())]
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(every, steps - done)
            d.step_in_time(dt, n)
            done += n
            rows.append((d.current_time, d.integrated()["kinetic_energy"], d.numerical_entropy()))
        d.synchronize()
        wall = time.perf_counter() - t0
        h = np.array(rows)
        hist[prec] = h
        out[prec] = dict(steps=steps, final_time=d.current_time, wall_s=wall, kinetic_energy=h[-1, 1],
                         entropy_drift=float(np.max(np.abs(h[:, 2] - h[0, 2])) / abs(h[0, 2])))
        d.close()
    a, b = hist["mixed"], hist["fp64"]
    out["max_rel_diff_kinetic_energy"] = float(np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(b[:, 1])))
    out["max_entropy_diff_rel_S0"] = float(np.max(np.abs(a[:, 2] - b[:, 2])) / abs(b[0, 2]))
    out["samples"] = len(a)
    out["workload"] = f"Euler TGV p=3 32^3, CH split + EC/Roe, RK3(2)5F[3S*+], {steps} steps"
    return out


MACH_VORTEX = 1.195228609334  # sqrt(2/gamma): p_inf = 1/(gamma M^2) = 1/2 for the isentropic vortex


def batched_vortex(n_inst=256, steps=200, seed=2024, precision="mixed"):
    """Batched uneven instances: 2D isentropic vortex on [-5,5]^2, N ~ U{8..64} elements per side,
    p ~ U{2..6}, strength ~ U[1,5], advection ~ U[0.5,1]^2 (seeded), CH split form + EC/Roe surface
    flux, RK4; all instances advance with one graph launch per step.  Returns DOF-updates/s over
    the timed steps."""
    import numpy as np
    import philip_amd as P
    rng = np.random.default_rng(seed)
    prec = P.MIXED if precision == "mixed" else P.FP64
    kws, ics = [], []
    for _ in range(n_inst):
        kws.append(dict(dim=2, poly_degree=int(rng.integers(2, 7)), n_cells_1d=int(rng.integers(8, 65)),
                        domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX, use_split_form=1,
                        two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec))
        ics.append((float(rng.uniform(1, 5)), *(float(x) for x in rng.uniform(0.5, 1.0, 2))))
    b = P.Batch(kws)
    dts = []
    for d, ic in zip(b.instances, ics):
        d.set_initial_condition_vortex(*ic)
        dts.append(d.constant_time_step(0.05) / 2.0)
    b.allocate("rk4", dts)
    tiles = {}
    for k, kw in enumerate(kws):
        epb, waves, threads = b.tile(k)
        tiles[f"p{kw['poly_degree']}"] = dict(elements_per_block=epb, waves_per_simd_bound=waves,
                                              threads_per_block=threads)
    b.step_in_time(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.step_in_time(steps)
    b.synchronize()
    t = time.perf_counter() - t0
    total = sum(d.ndofs for d in b.instances)
    b.close()
    return dict(value=total * 4 * steps / t, unit="DOF-updates/s", instances=n_inst, dof_total=total,
                steps=steps, ms_per_step=1e3 * t / steps, precision=precision,
                tiles=dict(sorted(tiles.items())),
                workload="256 independent 2D vortex solves, N~U{8..64}, p~U{2..6}, RK4, one graph launch per step")


def serving(n_req=400, rate=8.0, seed=77, precision="mixed", time_scale=1.0):
    """Serving stream: Poisson arrivals (rate req/s) of parametric solves, 70% 2D vortex
    (N~U{8..64}, p~U{2..6}) and 30% 3D Euler TGV (M~U[0.05,0.3], p~U{2..5}, 4^3..16^3), final
    time ~U[0.5,2] x time_scale; HBM budget 64 GiB.  Latency = arrival to result on the host."""
    import numpy as np
    import philip_amd as P
    rng = np.random.default_rng(seed)
    prec = P.MIXED if precision == "mixed" else P.FP64
    reqs = []
    for _ in range(n_req):
        tf = float(rng.uniform(0.5, 2.0)) * time_scale
        if rng.uniform() < 0.7:
            kw = dict(dim=2, poly_degree=int(rng.integers(2, 7)), n_cells_1d=int(rng.integers(8, 65)),
                      domain_left=-5.0, domain_right=5.0, mach_inf=MACH_VORTEX, use_split_form=1,
                      two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            vort = (float(rng.uniform(1, 5)), *(float(x) for x in rng.uniform(0.5, 1.0, 2)))
            reqs.append(P.request(tf, 0.05, 2.0, P.IC_VORTEX, vort, **kw))
        else:
            m = float(rng.uniform(0.05, 0.3))
            kw = dict(dim=3, poly_degree=int(rng.integers(2, 6)), n_cells_1d=int(rng.choice([4, 8, 16])),
                      mach_inf=m, use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, precision=prec)
            reqs.append(P.request(tf, 0.1, 1.0 + 1.0 / m, P.IC_TGV, **kw))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    s = P.Server(64 << 30, max_active=256)
    ids, delay = [], []
    t0 = time.perf_counter()
    k = 0
    done = 0
    max_active = max_queued = steps = 0
    while done < n_req:
        now = time.perf_counter() - t0
        while k < n_req and arrivals[k] <= now:
            ids.append(s.submit(reqs[k]))
            delay.append(now - arrivals[k])  # host polling delay, added to the latency
            k += 1
        a, q, done_now = s.step()
        if done_now // 100 > done // 100:
            progress(f"serving {precision}: {done_now} of {n_req} requests done")
        done = done_now
        steps += 1
        max_active, max_queued = max(max_active, a), max(max_queued, q)
        if a == 0 and q == 0 and k < n_req:
            time.sleep(max(0.0, arrivals[k] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    lat, dofu = [], 0.0
    for rid, dl in zip(ids, delay):
        _, l, _, du = s.result(rid)
        lat.append(l + dl)
        dofu += du
    s.close()
    lat = np.array(lat)
    return dict(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                max_s=float(lat.max()), throughput=dofu / wall, unit="DOF-updates/s", requests=n_req,
                rate_per_s=rate, wall_s=wall, scheduler_steps=steps, max_active=max_active, max_queued=max_queued,
                precision=precision,
                workload=f"70% 2D vortex / 30% 3D TGV, final time U[0.5,2]x{time_scale}, RK4, HBM budget 64 GiB")


def ns_tgv_params():
    import philip_amd as P
    return dict(dim=3, poly_degree=5, n_cells_1d=16, pde=P.NAVIER_STOKES, diss_flux=P.BR2, use_constant_viscosity=1,
                constant_viscosity=1.0, reynolds_inf=1600.0, prandtl=0.71, mach_inf=0.1, use_split_form=1,
                two_point_flux=P.RA, conv_flux=P.TWO_POINT_ROE)


def navier_stokes_tgv(precision="mixed", steps=3, curved=False):
    """Navier-Stokes TGV, Re 1600, Pr 0.71, M 0.1, p=5 on 16^3 hex (4,423,680 DOF), BR2, Ranocha
    split form + EC/Roe surface flux, constant viscosity, classical RK4 (auxiliary-gradient kernel +
    stage kernel per stage).  curved: on the reference's curved periodic grid (beta 0.1), the
    lifted gradient and the stage kernel with the metric terms.  Returns DOF-updates/s over the
    timed steps."""
    import philip_amd as P
    d = P.DG(precision=P.MIXED if precision == "mixed" else P.FP64, grid_warp=1 if curved else 0,
             **ns_tgv_params())
    d.set_initial_condition_tgv()
    dt = d.constant_time_step(0.01)
    d.allocate_ode_system("rk4")
    d.step_in_time(dt, 1)
    d.synchronize()
    t0 = time.perf_counter()
    d.step_in_time(dt, steps)
    d.synchronize()
    t = time.perf_counter() - t0
    n = d.ndofs
    d.close()
    grid = " on the curved periodic grid" if curved else ""
    return dict(value=n * 4 * steps / t, unit="DOF-updates/s", dof=n, ms_per_stage=1e3 * t / (4 * steps),
                precision=precision, workload=f"NS TGV Re1600 p=5 16^3{grid}, BR2, Ranocha split + EC/Roe, RK4")


def curvilinear_tgv(precision="fp64", steps=3):
    """Euler TGV p=3 on the reference's curved periodic grid (NonsymmetricCurvedGridManifold, beta
    0.1) of [0, 2 pi]^3 with 32^3 cells (10,485,760 DOF), CH split form + EC/Roe surface flux,
    RK3(2)5F[3S*+]: the curvilinear stage kernel (stage2_kernel with the metric terms).  Algorithmic
    bytes per DOF-update: the RK registers' 38.4 B plus the metric (9 cofactors and det J, 80 B per
    node = 16 B per DOF)."""
    import philip_amd as P
    d = P.DG(precision=P.MIXED if precision == "mixed" else P.FP64, dim=3, poly_degree=3, n_cells_1d=32,
             grid_warp=1, use_split_form=1, two_point_flux=P.CH, conv_flux=P.TWO_POINT_ROE, mach_inf=0.1)
    d.set_initial_condition_tgv()
    dt = d.constant_time_step(0.01)
    d.allocate_ode_system("rk3_2_5f_3sstarplus")
    d.step_in_time(dt, 1)
    d.synchronize()
    t0 = time.perf_counter()
    d.step_in_time(dt, steps)
    d.synchronize()
    t = time.perf_counter() - t0
    d.enable_timing(True)
    d.step_in_time(dt, 1)
    avg_ms, _ = d.kernel_timing()
    d.enable_timing(False)
    n = d.ndofs
    d.close()
    bytes_per_launch = n * (38.4 + 16.0)
    return dict(value=n * 5 * steps / t, unit="DOF-updates/s", dof=n, ms_per_stage=1e3 * t / (5 * steps),
                kernel_ms=avg_ms, hbm_frac=bytes_per_launch / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                precision=precision,
                workload="Euler TGV p=3 on the curved periodic grid 32^3, CH split + EC/Roe, RK3(2)5F[3S*+]")


def navier_stokes_history(final_time=2.0, every=100):
    """The Navier-Stokes TGV above run to t = final_time (CFL 0.01, last step clipped) at both
    precisions, enstrophy and the vorticity-based dissipation rate (QGauss(p+11)) every `every`
    steps and at the end.  Reports each run's steps and wall time and the largest relative
    difference of the mixed history from the FP64 one (the FP64 device path restates the oracle
    to 1e-12 per stage; the oracle itself would need ~1 h of CPU time for this run)."""
    import numpy as np
    import philip_amd as P
    hist, out = {}, {}
    for prec in ("fp64", "mixed"):
        d = P.DG(precision=P.MIXED if prec == "mixed" else P.FP64, **ns_tgv_params())
        d.set_initial_condition_tgv()
        dt = d.constant_time_step(0.01)
        d.allocate_ode_system("rk4")
        n_full = int(final_time / dt)
        rows = []
        t0 = time.perf_counter()
        done = 0
        while done < n_full:
            n = min(every, n_full - done)
            d.step_in_time(dt, n)
            done += n
            q = d.integrated()
            rows.append((d.current_time, q["enstrophy"], q["vorticity_dissipation"]))
        rem = final_time - d.current_time
        if rem > 1e-12 * final_time:
            d.step_in_time(rem, 1)
            q = d.integrated()
            rows.append((d.current_time, q["enstrophy"], q["vorticity_dissipation"]))
        d.synchronize()
        wall = time.perf_counter() - t0
        hist[prec] = np.array(rows)
        out[prec] = dict(steps=d.current_iteration, final_time=d.current_time, wall_s=wall,
                         enstrophy=rows[-1][1], vorticity_dissipation=rows[-1][2])
        d.close()
    a, b = hist["mixed"], hist["fp64"]
    out["max_rel_diff_enstrophy"] = float(np.max(np.abs(a[:, 1] - b[:, 1]) / np.abs(b[:, 1])))
    out["max_rel_diff_dissipation"] = float(np.max(np.abs(a[:, 2] - b[:, 2]) / np.abs(b[:, 2])))
    out["samples"] = len(a)
    out["workload"] = f"NS TGV Re1600 p=5 16^3, BR2, Ranocha split + EC/Roe, RK4 to t={final_time}"
    return out


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size (read nbytes + write nbytes per launch): the
    reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to HBM bytes."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)          # runtime copy kernel
    for _ in range(reps):
        torch.mul(x, 2.0, out=y)  # elementwise kernel, same bytes
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": nbytes, "copy_calibration_bytes_written": nbytes,
                      "launches": reps}), flush=True)


def roofline(precision, r, bytes_per_launch):
    achieved = bytes_per_launch / (r["avg_ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(precision),
            "kernel": KERNEL[precision], "avg_launch_ms": r["avg_ms"],
            "algorithmic_bytes_per_launch": bytes_per_launch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="precision of the headline line (the other is reported in its own block)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-second-precision", "--no-fp64-line", dest="no_second", action="store_true",
                    help="skip the block of the other precision")
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline "
                    "(default: every CPU this process may run on)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for scaling rocprofv3 byte counters)")
    ap.add_argument("--navier-stokes-only", choices=["fp64", "mixed"], default=None,
                    help="only run the Navier-Stokes TGV stage rate at this precision (for profiling)")
    ap.add_argument("--curvilinear-only", choices=["fp64", "mixed"], default=None,
                    help="only run the curved-grid Euler TGV stage rate at this precision (for profiling)")
    ap.add_argument("--curvilinear-ns-only", choices=["fp64", "mixed"], default=None,
                    help="only run the curved-grid Navier-Stokes TGV stage rate at this precision (for profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-instances line")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-stream line")
    ap.add_argument("--serving-rate", type=float, default=8.0,
                    help="Poisson arrival rate of the mixed-precision serving stream (FP64: half)")
    ap.add_argument("--serving-sweep", action="store_true", help="also sweep the arrival rate")
    ap.add_argument("--no-ns-history", action="store_true",
                    help="skip the long runs at both precisions (Euler TGV 1000 steps, Navier-Stokes TGV to t=2)")
    args = ap.parse_args()
    if args.navier_stokes_only:
        print(json.dumps(navier_stokes_tgv(args.navier_stokes_only)), flush=True)
        return
    if args.curvilinear_only:
        print(json.dumps(curvilinear_tgv(args.curvilinear_only)), flush=True)
        return
    if args.curvilinear_ns_only:
        print(json.dumps(navier_stokes_tgv(args.curvilinear_ns_only, curved=True)), flush=True)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver runs on a single MI355X "
              "(a sharded multi-GPU run is not implemented)", file=sys.stderr)
        sys.exit(2)

    r = run_gpu(args.precision, args.steps, args.warmup)
    ndofs = r["ndofs"]
    # first stage of a 3S* step moves 32 B/DOF, the other four 40 B/DOF
    bytes_per_launch = ndofs * (32 + 4 * 40) / 5.0
    other = "mixed" if args.precision == "fp64" else "fp64"
    line = {
        "metric": "DOF-updates/s per RK stage (p=3 Euler TGV)",
        "value": r["value"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-flux/f64-state",
        "data": "synthetic (TGV initial condition, reference's closed form)",
        "config": {"workload": "Euler TGV p=3 32^3 hex, CH split + EC/Roe surface flux, GLL, RK3(2)5F[3S*+]",
                   "dof": ndofs, "stages_per_step": r["stages"], "precision": args.precision,
                   "kinetic_energy_after_run": r["ke"],
                   "host_buffer_rate": r["pcie"]},
        "roofline": roofline(args.precision, r, bytes_per_launch),
    }
    progress(f"{args.precision} stage: {r['avg_ms'] * 1e3:.1f} us per launch")
    if not args.no_second:
        r2 = run_gpu(other, args.steps, args.warmup)
        progress(f"{other} stage: {r2['avg_ms'] * 1e3:.1f} us per launch")
        line[other + "_path"] = {"value": r2["value"], "ms_per_step": r2["ms_per_step"],
                                 "dtype": "f64" if other == "fp64" else "f32-flux/f64-state",
                                 "roofline": roofline(other, r2, bytes_per_launch)}
    if not args.no_batch:
        line["batched"] = batched_vortex()
        progress(f"batched: {line['batched']['value']:.3e} DOF-updates/s")
        line["navier_stokes"] = [navier_stokes_tgv("fp64"), navier_stokes_tgv("mixed")]
        progress("Navier-Stokes stage rates done")
        line["curvilinear"] = [curvilinear_tgv("fp64"), curvilinear_tgv("mixed"),
                               navier_stokes_tgv("fp64", curved=True), navier_stokes_tgv("mixed", curved=True)]
        progress("curvilinear-grid stage rates done")
    if not args.no_ns_history:
        line["tgv_1000_steps"] = tgv_history()
        progress("TGV 1000-step histories done")
        line["navier_stokes_t2"] = navier_stokes_history()
        progress("Navier-Stokes TGV to t=2 done")
        if "mixed_path" in line:
            h = line["tgv_1000_steps"]
            line["mixed_path"]["error_vs_fp64"] = {
                "max_rel_diff_kinetic_energy_1000_steps": h["max_rel_diff_kinetic_energy"],
                "entropy_drift_mixed": h["mixed"]["entropy_drift"],
                "max_entropy_diff_rel_S0": h["max_entropy_diff_rel_S0"]}
    if not args.no_serving:
        line["serving"] = []
        for pr, rate in (("mixed", args.serving_rate), ("fp64", args.serving_rate / 2)):
            line["serving"].append(serving(precision=pr, rate=rate))
            progress(f"serving {pr}: p50 {line['serving'][-1]['p50_s']:.3f} s, p99 {line['serving'][-1]['p99_s']:.3f} s")
        if args.serving_sweep:
            # arrival rate swept at both precisions (the same 100-request stream per rate, the
            # configuration's final times)
            line["serving_sweep"] = []
            for pr, rates in (("mixed", (16.0, 32.0, 64.0)), ("fp64", (8.0, 16.0, 32.0))):
                for r in rates:
                    line["serving_sweep"].append(serving(n_req=100, rate=r, precision=pr))
                    progress(f"serving sweep {pr} {r:g}/s: p50 {line['serving_sweep'][-1]['p50_s']:.3f} s, "
                             f"p99 {line['serving_sweep'][-1]['p99_s']:.3f} s")
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cores()[0]
        line["cpu_baseline"] = cpu_baseline(threads)
        progress(f"CPU baseline: {line['cpu_baseline']['value']:.3e} DOF-updates/s on {threads} threads")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
