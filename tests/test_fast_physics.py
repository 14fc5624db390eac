"""CPU check of the line kernels' cached-node point physics (dg_fast.hpp), built for the host
with g++ in float (mixed path) and in double (FP64 path), against the oracle's FP64 point
physics.

The functions restate the reference formulas with fewer divisions (per-node reciprocals, log
means from (a-b)/(a+b), Roe averages from sqrt(rho)); a dropped term, a wrong sign or a swapped
side shows up as an O(1) relative error, while rounding of this restatement stays below 2e-5 of
the flux magnitude in float and below 1e-13 in double, on states near the TGV/vortex regimes
(M = 0.1 and M ~ 1.2, density and pressure varying by up to 30% between the two sides so that
both branches of the log mean are exercised)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def host_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fast") / "fast_physics_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "host", "fast_physics_host.cpp")],
                   check=True)
    return exe


def states(rng, n, mach, spread):
    gam = 1.4
    rho = 1.0 + spread * rng.uniform(-1, 1, n)
    vel = 0.5 * rng.standard_normal((n, 3))
    p = (1.0 / (gam * mach * mach)) * (1.0 + spread * rng.uniform(-1, 1, n))
    u = np.zeros((n, 5))
    u[:, 0] = rho
    u[:, 1:4] = rho[:, None] * vel
    u[:, 4] = p / (gam - 1) + 0.5 * rho * (vel ** 2).sum(1)
    return u


@pytest.mark.parametrize("tp", [O.KG, O.IR, O.CH, O.RA])
@pytest.mark.parametrize("cf", [O.LXF, O.ROE, O.CENTRAL, O.TWO_POINT, O.TWO_POINT_LXF, O.TWO_POINT_ROE])
@pytest.mark.parametrize("mach,spread", [(0.1, 0.01), (0.1, 0.3), (1.2, 0.3)])
@pytest.mark.parametrize("prec,bar", [(0, 2e-5), (1, 1e-13)], ids=["float", "double"])
def test_cached_node_fluxes_match_oracle(host_exe, tp, cf, mach, spread, prec, bar):
    rng = np.random.default_rng(17 + tp + 7 * cf)
    n = 64
    ui = states(rng, n, mach, spread)
    # the second side: small perturbation of the first for half the samples
    ue = states(rng, n, mach, spread)
    ue[: n // 2] = ui[: n // 2] * (1 + 1e-3 * rng.standard_normal((n // 2, 5)))
    inp = f"{n} {tp} {cf} 1.4 {mach} {prec}\n" + "\n".join(" ".join(f"{x:.17g}" for x in np.r_[a, b])
                                                     for a, b in zip(ui, ue))
    out = subprocess.run([host_exe], input=inp, capture_output=True, text=True, check=True).stdout
    got = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()]).reshape(n, 3, 2, 5)
    o = O.Oracle(dim=3, poly_degree=1, n_cells_1d=2, two_point_flux=tp, conv_flux=cf, mach_inf=mach,
                 use_split_form=1 if cf >= O.TWO_POINT else 0)
    worst = 0.0
    for i in range(n):
        F = o.two_point_flux(tp, ui[i], ue[i])
        for k in range(3):
            nrm = np.zeros(3)
            nrm[k] = 1.0
            G = o.numerical_flux(cf, ui[i], ue[i], nrm)
            scale_f = np.abs(F[:, k]).max()
            scale_g = np.abs(G).max()
            worst = max(worst, np.abs(got[i, k, 0] - F[:, k]).max() / scale_f,
                        np.abs(got[i, k, 1] - G).max() / scale_g)
    assert worst < bar, worst


@pytest.fixture(scope="module")
def boundary_exe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bnd") / "boundary_state_host")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "host", "boundary_state_host.cpp")],
                   check=True)
    return exe


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mach", [0.5, 1.6])
@pytest.mark.parametrize("prec,bar", [(0, 2e-5), (1, 1e-14)], ids=["float", "double"])
def test_device_boundary_states_match_oracle(boundary_exe, dim, mach, prec, bar):
    """The device's boundary states (every type, every face direction and orientation) against the
    oracle's, which tests/test_oracle_boundary.py pins by what each condition is defined to do."""
    rng = np.random.default_rng(dim + int(10 * mach))
    aoa = 0.3
    o = O.Oracle(dim=dim, poly_degree=1, n_cells_1d=2, mach_inf=mach, angle_of_attack=aoa)
    pinf = 1.0 / (1.4 * mach * mach)
    rows, expect = [], []
    for _ in range(20):
        rho = rng.uniform(0.7, 1.3)
        v = rng.uniform(-2.0, 2.0, dim) * (1.0 if mach < 1 else 0.8)
        p = pinf * rng.uniform(0.7, 1.3)
        u = np.r_[rho, rho * v, p / 0.4 + 0.5 * rho * v @ v]
        for bc in range(1001, 1008):
            for k in range(dim):
                for sg in (-1.0, 1.0):
                    nrm = np.zeros(dim)
                    nrm[k] = sg
                    rows.append(f"{bc} {k} {sg} " + " ".join(f"{x:.17g}" for x in u))
                    expect.append(o.boundary_state(bc, u, nrm))
    inp = f"{len(rows)} {dim} {mach} {aoa} {prec}\n" + "\n".join(rows)
    out = subprocess.run([boundary_exe], input=inp, capture_output=True, text=True, check=True).stdout
    got = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    expect = np.array(expect)
    err = (np.abs(got - expect) / np.abs(expect).max(axis=1, keepdims=True)).max()
    assert err < bar, err
