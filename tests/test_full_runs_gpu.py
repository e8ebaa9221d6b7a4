"""Full-duration runs against the oracle, through the C ABI.

* The Sod tube of the modal-DG configuration at full size (40 x 4 x 4
  shuffled hexahedra, p = 2, exact solver, smart ends and inviscid walls,
  demo/euler/shock_tube.cpp:32-68) to t = 0.2 with the time step evaluated
  every step at CFL 0.1 (sourceless.cpp:187-205): 359 steps.
* The four-quadrant Riemann problem on jittered tetrahedra (the C4 layout at
  8 x 8 x 2 hexahedra x 6, p = 2, HLLC, weno::Eigen detector with
  weno::Lazy and with weno::Eigen reconstruction) to t = 0.25, with smart
  boundaries holding the initial quadrant states;
* the reference's own riemann2d demo set-up with its moving boundary states
  to t = 0.25.

Both sides advance with the oracle's time step (the device's own value is
checked against it every step), so the comparison is of the states alone.
The L2 norm of a difference is exact through the orthonormal basis:
sqrt(sum over cells and coefficients of the squared coefficient
differences).  Tolerances: the exact solver's star pressure is defined to
about 1e-8 (exact.hpp:140-152), so after 359 steps the Sod states agree to
1e-7 relative (max norm); with HLLC every step agrees to 1e-12 and so does
the final state after the full run, and the troubled-cell flags of every
step are identical.
"""
import numpy as np
import pytest

import oracle_lib as ol
from pvcriemannsolvers_amd import meshgen, spatial

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def cons(rho, vel, p):
    rho = np.asarray(rho, dtype=np.float64)
    u = np.zeros(rho.shape + (5,))
    u[..., 0] = rho
    for d in range(3):
        u[..., 1 + d] = rho * vel[d]
    u[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return u


def norms(got, ref, n_per_cell):
    d = (got - ref).reshape(-1, n_per_cell, 5)
    r = ref.reshape(-1, n_per_cell, 5)
    out = {}
    for k, name in ((0, "rho"), (4, "E")):
        out["l2_" + name] = float(np.sqrt((d[..., k] ** 2).sum()))
        out["linf_" + name] = float(np.abs(d[..., k]).max() / np.abs(r[..., k]).max())
    return out


def test_sod_tube_to_final_time():
    mesh = meshgen.box_hex(40, 4, 4, hi=(1.0, 0.1, 0.1), shuffle_seed=2023)
    left, right = cons(1.0, (0, 0, 0), 1.0), cons(0.125, (0, 0, 0), 0.1)
    bcs = [("smart", left), ("smart", right)] + [("inviscid_wall", None)] * 4
    o = ol.DgOracle(mesh, degree=2, riemann=0, bcs=[(0, left), (0, right)] + [(1, [0] * 5)] * 4)
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="exact", bcs=bcs)
    x = g.sample_points()
    vals = np.where((x[..., 0] < 0.5)[..., None], left, right)
    g.approximate(vals)
    u = o.approximate(vals)
    t, t_end, steps = 0.0, 0.2, 0
    while t < t_end:
        dt = o.time_step(u, 0.1, t_end - t)
        assert abs(g.get_time_step(0.1, t_end - t) - dt) <= 1e-9 * dt, steps
        g.rk3_update(t, dt)
        u = o.rk3(u, dt, 1)
        t += dt
        steps += 1
    got = g.get_solution_column()
    n = norms(got, u, g.n_per_cell)
    print(f"Sod tube to t = 0.2: {steps} steps, {n}", flush=True)
    assert steps == 359
    assert np.isfinite(got).all()
    assert np.abs(got - u).max() <= 1e-7 * np.abs(u).max(), n
    g.close()


def quadrant_state(x, seed):
    """Four constant quadrants of [0, 1]^2 split at 0.5, rho in [0.1, 1.5],
    p in [0.03, 1.5], u, v in [-1, 1] per quadrant, seeded."""
    rng = np.random.default_rng(seed)
    rho, p = rng.uniform(0.1, 1.5, 4), rng.uniform(0.03, 1.5, 4)
    vel = rng.uniform(-1, 1, (4, 2))
    q = (x[..., 0] > 0.5).astype(int) + 2 * (x[..., 1] > 0.5).astype(int)
    return cons(rho[q], (vel[q, 0], vel[q, 1], 0.0 * q), p[q])


# The reference's riemann2d demo (demo/euler/riemann2d.cpp:5-53): quadrant
# states a..d split at 0.8, inviscid walls on top and right, and smart
# boundaries on the left and bottom whose states switch at interfaces moving
# with the Rankine-Hugoniot speeds of the c|b and c|d shocks.
PRIM = {"a": (1.5, 0.0, 0.0, 1.5), "b": (0.5323, 1.206, 0.0, 0.3), "c": (0.1380, 1.206, 1.206, 0.029),
        "d": (0.5323, 0.0, 1.206, 0.3)}
VAL = {k: cons(r, (u, v, 0.0), p) for k, (r, u, v, p) in PRIM.items()}
SPEED_CD = (VAL["d"][1] - VAL["c"][1]) / (VAL["d"][0] - VAL["c"][0])
SPEED_BC = (VAL["c"][2] - VAL["b"][2]) / (VAL["c"][0] - VAL["b"][0])


def riemann2d_ic(x):
    low_x, low_y = x[..., 0] < 0.8, x[..., 1] < 0.8
    return np.where((low_x & low_y)[..., None], VAL["c"], np.where(low_x[..., None], VAL["b"],
                    np.where(low_y[..., None], VAL["d"], VAL["a"])))


def state_left(xyz, t):
    return np.where(((xyz[:, 1] - SPEED_BC * t) < 0.8)[:, None], VAL["c"], VAL["b"])


def state_bottom(xyz, t):
    return np.where(((xyz[:, 0] - SPEED_CD * t) < 0.8)[:, None], VAL["c"], VAL["d"])


def run_tets(mode, code, ic, funcs, bcs, n=8, t_end=0.25):
    """Both sides to t_end with the oracle's time step; returns the report."""
    mesh = meshgen.box_tet(n, n, 2, hi=(1.0, 1.0, 2.0 / n), jitter=0.15, seed=5)
    kinds = {"smart": 0, "inviscid_wall": 1}
    o = ol.DgOracle(mesh, degree=2, riemann=1, bcs=[(kinds[k], s if s is not None else [0] * 5) for k, s in bcs])
    o.set_limiter(code)
    g = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc", bcs=bcs, limiter=mode)
    for tag, f in funcs.items():
        g.set_boundary_function(tag, f)
    constants = np.array([s if s is not None else [0.0] * 5 for _, s in bcs])
    vals = ic(g.sample_points())
    # start-up as the reference's driver does for tetrahedra: Approximate,
    # then two limiter passes (demo/euler/sourceless.cpp:131-147)
    g.approximate(vals)
    g.set_solution_column(g.get_solution_column())
    u, _ = o.limit(o.limit(o.approximate(vals))[0])
    t, steps, flagged, first_divergence = 0.0, 0, 0, None
    while t < t_end:
        dt = o.time_step(u, 0.1, t_end - t)
        assert abs(g.get_time_step(0.1, t_end - t) - dt) <= 1e-9 * dt, steps
        g.rk3_update(t, dt)
        u = ol.rk3_timed(o, u, t, dt, funcs, constants, lambda v: o.limit(v)[0])
        flags = o.last_flags()
        flagged += int(flags.sum())
        if first_divergence is None and not np.array_equal(g.limiter_flags(), flags):
            first_divergence = steps
        t += dt
        steps += 1
    got = g.get_solution_column()
    g.close()
    return got, u, steps, flagged, first_divergence, norms(got, u, 10)


@pytest.mark.parametrize("mode,code", [("weno", 1), ("weno_eigen", 3)])
def test_four_quadrant_tets_to_final_time(mode, code):
    """Seeded quadrant states (seed 22) on jittered tetrahedra, HLLC, WENO, to
    t = 0.25.  The x and y sides are smart boundaries that hold the initial
    state of the adjacent quadrant at each boundary point (a boundary
    function of the point), the z sides inviscid walls."""
    ic = lambda x: quadrant_state(x, 22)  # noqa: E731
    hold = lambda xyz, t: quadrant_state(xyz, 22)  # noqa: E731
    bcs = [("smart", [1, 0, 0, 0, 2.5])] * 4 + [("inviscid_wall", None)] * 2
    got, u, steps, flagged, first_div, n = run_tets(mode, code, ic, {k: hold for k in range(4)}, bcs)
    print(f"four-quadrant tets ({mode}) to t = 0.25: {steps} steps, {flagged} flagged cell-steps, "
          f"first flag divergence {first_div}, {n}", flush=True)
    assert np.isfinite(got).all()
    assert first_div is None
    assert flagged > 0
    assert np.abs(got - u).max() <= 1e-12 * np.abs(u).max(), n


def test_reference_riemann2d_to_final_time():
    """The reference's riemann2d set-up on jittered tetrahedra with its moving
    smart-boundary states (time-dependent boundary functions), weno::Lazy on
    the flagged cells, to t = 0.25."""
    bcs = [("smart", VAL["c"]), ("inviscid_wall", None), ("smart", VAL["c"]), ("inviscid_wall", None),
           ("inviscid_wall", None), ("inviscid_wall", None)]
    got, u, steps, flagged, first_div, n = run_tets("weno", 1, riemann2d_ic, {0: state_left, 2: state_bottom}, bcs)
    print(f"riemann2d tets to t = 0.25: {steps} steps, {flagged} flagged cell-steps, "
          f"first flag divergence {first_div}, {n}", flush=True)
    assert np.isfinite(got).all()
    assert first_div is None
    assert np.abs(got - u).max() <= 1e-12 * np.abs(u).max(), n
