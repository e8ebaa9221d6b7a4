"""The C++ host surface (include/mini/**): mini::temporal,
mini::riemann::rotated::Euler and mini::spatial::mi355x::FiniteElement,
exercised by compiled drivers.

The drivers are built in-tree by __graft_entry__.build() (tests/cpp/Makefile
-> tests/cpp/_build/).  The CPU test builds its driver itself when missing;
the GPU test requires the prebuilt driver, which links libminicfd.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _run(binary):
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    return r


def test_runge_kutta_bitwise_cpp():
    """RungeKutta<1|2|3> of the C++ surface against closed-form combinations (test/temporal/rk.cpp:21-72)."""
    binary = os.path.join(CPP, "_build", "test_temporal")
    if not os.path.exists(binary):
        subprocess.run(["make", "-C", CPP, os.path.join("_build", "test_temporal")], check=True)
    r = _run(binary)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 18


GAMMA, R_GAS = 1.4, 287.05


def _cons(rho, vel, p):
    return np.array([rho, *(rho * vel), p / (GAMMA - 1) + 0.5 * rho * vel @ vel])


def _riemann_points(n, seed):
    """Frames, conservative pairs, subsonic givens, and a second pair per point;
    every other second pair opens a vacuum, where the exact solver takes the
    tangential velocity from the side its kept star velocity points to
    (exact.hpp:36-43, the star velocity of its last non-vacuum call)."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    frames = q * np.sign(np.linalg.det(q))[:, None, None]
    rows = []
    for i in range(n):
        nrm = frames[i, 0]
        rho_l, rho_r, p_l, p_r = rng.uniform(0.3, 3.0, 4)
        # the left state flows in subsonically along d, as a subsonic inlet needs
        d = -nrm + 0.3 * rng.standard_normal(3)
        d /= np.linalg.norm(d)
        if d @ nrm > -0.3:
            d = -nrm
        v_l = rng.uniform(0.05, 0.8) * np.sqrt(GAMMA * p_l / rho_l) * d
        v_r = rng.normal(0, 0.8, 3)
        left, right = _cons(rho_l, v_l, p_l), _cons(rho_r, v_r, p_r)
        t_l = p_l / (rho_l * R_GAS)
        inlet = [p_l * rng.uniform(1.05, 1.6), *d, t_l * rng.uniform(1.02, 1.3)]
        outlet = [0, 0, 0, 0, rng.uniform(0.5, 2.0)]
        if i % 2:
            def shear():
                return frames[i, 1] * rng.normal() + frames[i, 2] * rng.normal()
            drift = rng.choice([-2.0, 2.0])  # puts the t-axis inside one of the fans
            l2 = _cons(1.0, (drift - 4.5) * nrm + shear(), 0.4)
            r2 = _cons(1.0, (drift + 4.5) * nrm + shear(), 0.4)
        else:
            l2, r2 = _cons(rho_r, v_r, p_r), _cons(rho_l, v_l, p_l)
        rows.append(np.concatenate([frames[i].ravel(), left, right, inlet, outlet, l2, r2]))
    return np.array(rows)


@pytest.mark.parametrize("riemann", [0, 1, 2], ids=["exact", "hllc", "roe"])
def test_riemann_host_types_match_oracle(riemann):
    """rotated::Euler<Exact|HartenLaxLeerContact|Roe> on the host (the
    Convective concept is a static_assert of the driver) against the oracle's
    solver objects, every flux method, with each point's exact-solver state
    carried from call to call as one reference solver object does."""
    binary = os.path.join(CPP, "_build", "test_riemann_host")
    if not os.path.exists(binary):
        subprocess.run(["make", "-C", CPP, os.path.join("_build", "test_riemann_host")], check=True)
    pts = _riemann_points(96, 40 + riemann)
    text = f"{riemann} {len(pts)}\n" + "\n".join(" ".join(f"{x:.17g}" for x in row) for row in pts) + "\n"
    r = subprocess.run([binary], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.splitlines()]).reshape(len(pts), 8, 5)
    fr, left, right = pts[:, :9], pts[:, 9:14], pts[:, 14:19]
    inlet, outlet, l2, r2 = pts[:, 19:24], pts[:, 24:29], pts[:, 29:34], pts[:, 34:39]
    star = np.zeros(len(pts))
    want = np.stack([
        ol.face_fluxes(riemann, 0, fr, left, right, star),
        ol.face_fluxes(riemann, 1, fr, left),
        ol.face_fluxes(riemann, 2, fr, left),
        ol.face_fluxes(riemann, 2, fr, right),
        ol.face_fluxes(riemann, 3, fr, left, inlet),
        ol.face_fluxes(riemann, 4, fr, left, outlet),
        ol.face_fluxes(riemann, 0, fr, left, right, star),
        ol.face_fluxes(riemann, 0, fr, l2, r2, star),
    ], axis=1)
    scale = np.maximum(np.abs(want).max(axis=2, keepdims=True), 1e-300)
    rel = np.abs(got - want) / scale
    # HLLC, Roe and the boundary fluxes: rounding only; the exact solver's
    # star pressure iteration stops at f >= -1e-8 (exact.hpp:140-152)
    tol = 1e-12 if riemann == 0 else 1e-13
    assert rel.max() < tol, np.unravel_index(rel.argmax(), rel.shape)


@pytest.mark.parametrize("tet", [False, True], ids=["fr_trilinear_hex", "dg_tet"])
def test_vtk_writer_cpp_equals_python_writer(tmp_path, tet):
    """mini::mesh::vtk::Writer (C++) writes the frame that pvcriemannsolvers_amd.vtk
    writes for the same host-only element and column."""
    from pvcriemannsolvers_amd import meshgen, spatial, vtk
    binary = os.path.join(CPP, "_build", "test_vtk_writer")
    if not os.path.exists(binary):
        subprocess.run(["make", "-C", CPP, os.path.join("_build", "test_vtk_writer")], check=True)
    if tet:
        mesh = meshgen.box_tet(2, 2, 1, jitter=0.1, seed=3)
        fe = spatial.FiniteElement(mesh, degree=2, scheme="dg", riemann="hllc",
                                   bcs=[("inviscid_wall", None)] * 6, host_only=True)
    else:
        mesh = meshgen.box_hex(3, 2, 2, jitter=0.2, seed=5)
        fe = spatial.FiniteElement(mesh, degree=2, riemann="hllc", bcs=[("inviscid_wall", None)] * 6,
                                   host_only=True)
    column = np.random.default_rng(4).uniform(0.5, 2.0, fe.column_size)
    data = os.path.join(str(tmp_path), "in.bin")
    head = np.array([len(mesh["coords"]), len(mesh["cells"]), len(mesh["bface_tags"]), 2, int(tet)], np.int64)
    with open(data, "wb") as f:
        for a in (head, mesh["coords"].astype(np.float64), mesh["cells"].astype(np.int64),
                  mesh["bface_nodes"].astype(np.int64), mesh["bface_tags"].astype(np.int64), column):
            f.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([binary, data, os.path.join(str(tmp_path), "cpp")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
    pressure = ("Pressure", lambda xyz, v: 0.4 * (v[:, 4] - 0.5 * (v[:, 1:4] ** 2).sum(1) / v[:, 0]))
    vtk.write_solutions(fe, os.path.join(str(tmp_path), "py"), "frame", column=column, point_data=[pressure],
                        cell_data=[("CellId", lambda ids: ids.astype(float))])
    got = vtk.read_piece(os.path.join(str(tmp_path), "cpp", "frame", "0.vtu"))
    want = vtk.read_piece(os.path.join(str(tmp_path), "py", "frame", "0.vtu"))
    assert np.array_equal(got["points"], want["points"])
    for name in vtk.FIELD_NAMES:
        assert np.array_equal(got["point_data"][name], want["point_data"][name]), name
    p_got, p_want = got["point_data"]["Pressure"], want["point_data"]["Pressure"]
    assert np.abs(p_got - p_want).max() <= 1e-15 * np.abs(p_want).max()
    for key in ("connectivity", "offsets", "types"):
        assert np.array_equal(got[key], want[key]), key
    assert np.array_equal(got["cell_data"]["CellId"], want["cell_data"]["CellId"])
    assert open(os.path.join(str(tmp_path), "cpp", "frame.pvtu")).read() == \
        open(os.path.join(str(tmp_path), "py", "frame.pvtu")).read()
    fe.close()


@pytest.mark.gpu
def test_system_contract_cpp_on_gpu():
    """FiniteElement driven by RungeKutta<3> through the System contract equals the fused device step bit for bit;
    the limiter objects, the start-up limiters, GetTimeStep(dt_guess, rk_order), Approximate of a Value(Global)
    callable and the boundary functions behave as the reference's."""
    binary = os.path.join(CPP, "_build", "test_mi355x_system")
    assert os.path.exists(binary), "the C++ drivers are built by __graft_entry__.build()"
    r = _run(binary)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 8 and "FAIL" not in r.stdout, r.stdout
