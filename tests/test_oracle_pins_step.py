"""Pins of the oracle's time integrators, collision boundary and global
quantities against closed forms (CPU only).

  symplectic Euler, leap-frog, RK4 ..... one isolated particle: a fluid
      particle feels only gravity, so v_n = v0 + n dt g and
      x_n = x0 + dt sum_k v_k for every integrator (advect_pos runs after the
      substep, solver_sph_base.py:41-51); a soil particle (mu(I), DP) also
      feels the viscous damping -c v_tmp with c = 5e-5 sqrt(E/(rho h^2))
      (solver_sph_base.py:713-715), so its velocity follows an affine
      one-step map v' = A v + B whose A and B differ between symplectic
      Euler (:79-91), leap-frog (:96-120) and the reference's RK4 (:134-180,
      every intermediate state at half a step, the fourth included); the
      closed form v_n = A^n v0 + B (1 - A^n) / (1 - A) then tells the three
      apart (and a wrong weight or step fraction from each)
  collision boundary ................... a particle driven into the floor,
      the side walls and a corner with boundary 1 (solver_sph_base.py:524-601):
      clamped to the wall minus (radius - 1e-8), velocity minus 1.3 times its
      component along the normalised sum of the wall normals
  global quantities .................... front = max x, potential energy
      sum m (-g . x), kinetic energy sum m |v|^2 / 2, mass sum m over the
      flow particles only: the isolated particle's closed form, and numpy
      on the downloaded state of a scene with walls (walls excluded)
  whole step ........................... the numpy restatement
      (oracle/sph_numpy.py) with symplectic Euler, RK4 and the collision
      boundary against the C++ FP64 oracle
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.sph_numpy import NumpySPH
from tests.common import tiny_cfg
from tisphi_amd.scene import build_scene


def _isolated(method, ti, dim=2, v0=(0.3, 0.0, 0.0), dt=None, young=None, boundary=0,
              xyz=(0.5, 0.9, 0.2)):
    """One particle (block of one lattice site) in a 1 m box."""
    cfg = tiny_cfg(method=method, dim=dim, boundary=boundary, ti=ti, r=0.005,
                   domain=(1.0, 1.0, 0.4), block=(0.01, 0.01, 0.01))
    c = cfg.config
    b = c["Blocks"][0]
    b["translation"] = list(xyz) if dim == 3 else [xyz[0], xyz[1], 0.0]
    b["velocity"] = list(v0)
    if dt is not None:
        c["Configuration"]["timeStepSizeMin"] = dt
    if young is not None:
        c["Materials"][0]["EYoungMod"] = young
    sc = build_scene(cfg)
    assert sc.n == 1
    return sc


def _affine_map(step_fn):
    """(A, B) of a one-step map v -> A v + B (per axis, B = the map at 0)."""
    b = step_fn(0.0)
    return step_fn(1.0) - b, b


def _closed_form(A, B, v0, x0, n, dt):
    k = np.arange(1, n + 1, dtype=np.float64)
    Ak = A ** k
    geo = (1.0 - Ak) / (1.0 - A) if A != 1.0 else k
    vk = Ak * v0 + B * geo  # v_1 .. v_n
    return vk[-1], x0 + dt * np.sum(vk)


@pytest.mark.parametrize("ti", [1, 2, 4])
@pytest.mark.parametrize("dim", [2, 3])
def test_isolated_fluid_particle_all_integrators(ti, dim):
    sc = _isolated(1, ti, dim)
    o = O.Oracle(sc.desc(), 64)
    O.load_scene(o, sc)
    x0 = sc.pos[0].copy()
    n, dt, g = 150, sc.dt, np.array([0.0, -9.81, 0.0])
    o.step(n)
    v0 = np.array([0.3, 0.0, 0.0])
    np.testing.assert_allclose(o.get("vel")[0], v0 + n * dt * g, rtol=1e-12, atol=1e-14)
    xn = x0 + n * dt * v0 + dt * dt * g * n * (n + 1) / 2
    np.testing.assert_allclose(o.get("pos")[0], xn, rtol=0, atol=1e-12)
    assert o.get("density")[0] == sc.material["density0"]


@pytest.mark.parametrize("method", [2, 3])
@pytest.mark.parametrize("ti", [1, 2, 4])
def test_damped_soil_particle_integrator_maps(method, ti):
    # dt = 0.01 (timeStepSizeMin above the CFL step) and a stiff material,
    # so c dt ~ 0.18: the three maps differ from the second order on
    sc = _isolated(method, ti, v0=(0.4, 0.25, 0.0), dt=0.01, young=5.84e10)
    d = sc.desc()
    dt, g = d["dt"], -9.81
    assert dt == 0.01
    c = 5e-5 * np.sqrt(d["young"] / (d["density0"] * d["smoothing_len"] ** 2))
    k = lambda v: g - c * v  # noqa: E731  d_vel of the isolated soil particle (y axis)
    kx = lambda v: -c * v  # noqa: E731  (x axis: no gravity)

    def se(v, f):
        return v + dt * f(v)

    def lf(v, f):
        return v + dt * f(v + 0.5 * dt * f(v))

    def rk(v, f):
        k1 = f(v)
        k2 = f(v + 0.5 * dt * k1)
        k3 = f(v + 0.5 * dt * k2)
        k4 = f(v + 0.5 * dt * k3)  # the reference's fourth stage: half a step too
        return v + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)

    stepper = {1: se, 2: lf, 4: rk}[ti]
    Ay, By = _affine_map(lambda v: stepper(v, k))
    Ax, Bx = _affine_map(lambda v: stepper(v, kx))
    o = O.Oracle(d, 64)
    O.load_scene(o, sc)
    n = 40
    o.step(n)
    x0 = sc.pos[0]
    vy, yn = _closed_form(Ay, By, 0.25, x0[1], n, dt)
    vx, xn = _closed_form(Ax, Bx, 0.4, x0[0], n, dt)
    v = o.get("vel")[0]
    x = o.get("pos")[0]
    np.testing.assert_allclose([v[0], v[1]], [vx, vy], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose([x[0], x[1]], [xn, yn], rtol=1e-11, atol=1e-13)
    assert v[2] == 0.0 and x[2] == 0.0
    # the maps of the other integrators are far off: the pin tells them apart
    for other in {1: se, 2: lf, 4: rk}.values():
        if other is stepper:
            continue
        A2, B2 = _affine_map(lambda v: other(v, k))
        v2, _ = _closed_form(A2, B2, 0.25, x0[1], n, dt)
        assert abs(v2 - vy) > 1e-6 * abs(vy)


def _collision_expect(x, v, dt, g, de, ds, r, dim):
    """One step of an isolated fluid particle with boundary 1: v += dt g,
    x += dt v, then enforce_boundary (solver_sph_base.py:540-601)."""
    v = v + dt * g
    x = x + dt * v
    dist = r - 1e-8
    nrm = np.zeros(3)
    pos = x.copy()
    axes = (0, 1, 2) if dim == 3 else (0, 1)
    for a in axes:
        if a != 1 and pos[a] > de[a] - dist:
            nrm[a] += 1.0
            x[a] = de[a] - dist
        if pos[a] <= ds[a] + dist:
            nrm[a] -= 1.0
            x[a] = ds[a] + dist
    ln = np.linalg.norm(nrm)
    if ln > 1e-8:
        u = nrm / ln
        v = v - 1.3 * np.dot(v, u) * u
    return x, v, ln > 1e-8


@pytest.mark.parametrize("dim,xyz,v0", [
    (2, (0.5, 0.00506, 0.0), (0.2, -0.5, 0.0)),       # floor
    (2, (0.99492, 0.5, 0.0), (0.7, 0.1, 0.0)),        # right wall
    (2, (0.00508, 0.5, 0.0), (-0.6, 0.0, 0.0)),       # left wall
    (2, (0.99496, 0.00509, 0.0), (0.8, -0.9, 0.0)),   # corner: normal (1, -1)/sqrt 2
    (3, (0.5, 0.00505, 0.39497), (0.1, -0.4, 0.6)),   # floor + back wall
    (3, (0.00504, 0.3, 0.00507), (-0.5, 0.2, -0.7)),  # left + front wall
])
def test_collision_boundary_closed_form(dim, xyz, v0):
    # xyz is the particle's position: the block's lattice site sits half a
    # diameter (0.005 m) past its translation
    sc = _isolated(1, 2, dim, v0=v0, boundary=1, xyz=tuple(c - 0.005 for c in xyz))
    d = sc.desc()
    x = sc.pos[0].copy()
    np.testing.assert_allclose(x[:dim], xyz[:dim], atol=1e-15)
    v = np.array(v0, dtype=np.float64)
    if dim == 2:
        v[2] = 0.0
    g = np.array(d["gravity"])
    de, ds = np.array(d["domain_end"]), np.array(d["domain_start"])
    o = O.Oracle(d, 64)
    O.load_scene(o, sc)
    hits = 0
    for _ in range(30):
        x, v, hit = _collision_expect(x, v, d["dt"], g, de, ds, d["particle_radius"], dim)
        hits += int(hit)
        o.step(1)
        np.testing.assert_allclose(o.get("pos")[0], x, rtol=0, atol=1e-12)
        np.testing.assert_allclose(o.get("vel")[0], v, rtol=1e-12, atol=1e-13)
    assert hits >= 1  # the particle reached the wall(s) and was clamped


def test_globals_isolated_particle_closed_form():
    sc = _isolated(1, 2, 2, v0=(0.3, -0.2, 0.0))
    o = O.Oracle(sc.desc(), 64)
    O.load_scene(o, sc)
    o.step(7)
    x = o.get("pos")[0]
    v = o.get("vel")[0]
    m = sc.mass[0]
    front, pe, ke, mass = o.globals()
    assert front == x[0]
    assert pe == pytest.approx(m * 9.81 * x[1], rel=1e-15)
    assert ke == pytest.approx(0.5 * m * (v[0] ** 2 + v[1] ** 2), rel=1e-15)
    assert mass == m


@pytest.mark.parametrize("prec", [32, 64])
def test_globals_flow_particles_only(prec):
    """Walls (zero mass, at the largest x of the scene) must not enter."""
    sc = build_scene(tiny_cfg(method=1, dim=2))
    o = O.Oracle(sc.desc(), prec)
    O.load_scene(o, sc)
    o.step(5)
    t = o.get("type")
    fl = t == 1
    assert (~fl).any()
    x = o.get("pos")
    v = o.get("vel")
    m = o.get("mass")
    assert x[~fl, 0].max() > x[fl, 0].max()  # a wall particle lies farther right
    g = np.array(sc.desc()["gravity"])
    front, pe, ke, mass = o.globals()
    assert front == x[fl, 0].max()
    np.testing.assert_allclose(pe, np.sum(m[fl] * -(x[fl] @ g)), rtol=1e-12)
    np.testing.assert_allclose(ke, np.sum(0.5 * m[fl] * np.sum(v[fl] ** 2, axis=1)), rtol=1e-12)
    np.testing.assert_allclose(mass, np.sum(m[fl]), rtol=1e-14)
    assert mass == pytest.approx(fl.sum() * sc.material["density0"] * sc.geo.particle_diameter ** 2,
                                 rel=1e-6)


@pytest.mark.parametrize("method,dim,ti,boundary,xsph", [
    (1, 2, 1, 2, False), (1, 2, 4, 2, False), (1, 3, 4, 2, False), (2, 2, 4, 2, True),
    (3, 2, 4, 2, False), (3, 2, 1, 2, True), (2, 2, 1, 2, False), (1, 2, 2, 1, False),
    (1, 3, 2, 1, False), (1, 2, 4, 1, True), (2, 2, 2, 1, False), (3, 2, 1, 1, False)])
def test_oracle_matches_numpy_integrators_and_collision(method, dim, ti, boundary, xsph):
    sc = build_scene(tiny_cfg(method=method, dim=dim, xsph=xsph, boundary=boundary, ti=ti))
    d = sc.desc()
    o = O.Oracle(d, 64)
    O.load_scene(o, sc)
    ref = NumpySPH(d, sc.pos, sc.vel, sc.density, sc.mass, sc.m_V, sc.mat_type, sc.stress)
    for _ in range(25):
        o.step(1)
        ref.step()
    idx = o.get("id0")

    def close(a, b, tol=1e-9):
        s = np.max(np.abs(b))
        assert np.max(np.abs(a - b)) <= tol * max(s, 1e-300), (np.max(np.abs(a - b)), s)

    close(o.get("pos"), ref.x[idx])
    close(o.get("vel"), ref.v[idx])
    close(o.get("density"), ref.rho[idx])
    close(o.get("d_vel"), ref.dv[idx])
    if method != 1:
        close(o.get("stress"), ref.stress6("sig")[idx])
    if boundary == 1:
        # the walls were reached: some flow particle sits exactly at a clamp
        dist = d["particle_radius"] - 1e-8
        x = o.get("pos")[o.get("type") > 0]
        assert np.any(np.abs(x[:, 1] - (d["domain_start"][1] + dist)) < 1e-12)


def test_neighbor_csr_matches_walk_and_cells():
    """The oracle's all-particle neighbour export (the GPU list check's
    reference) equals its per-particle walk, with each entry's stencil code
    equal to the neighbour's cell minus the particle's."""
    sc = build_scene(tiny_cfg(method=1, dim=3))
    o = O.Oracle(sc.desc(), 32)
    O.load_scene(o, sc)
    o.step(3)
    o.run("sort")
    cnt, ids, codes = o.neighbor_csr()
    id0 = o.get("id0")
    cell = o.get("cell")
    gn = sc.geo.grid_num
    xyz = np.stack([cell // (gn[1] * gn[2]), (cell // gn[2]) % gn[1], cell % gn[2]], axis=1)
    pos_of = np.empty(sc.n, dtype=np.int64)
    pos_of[id0] = np.arange(sc.n)
    k = 0
    for i in range(sc.n):
        nb = o.neighbors(i)
        assert cnt[i] == nb.size
        np.testing.assert_array_equal(ids[k:k + nb.size], id0[nb])
        d = xyz[nb] - xyz[i] + 1
        np.testing.assert_array_equal(codes[k:k + nb.size], (d[:, 0] << 4) | (d[:, 1] << 2) | d[:, 2])
        k += nb.size
    assert k == ids.size


class _ListsFromOracle:
    """Stands in for the engine: lists filtered from a float32 oracle's walk
    (what the engine records), to exercise the GPU tests' list comparison."""

    def __init__(self, o, drop=None):
        from tests.common import _list_keep
        cnt, ids, codes = o.neighbor_csr()
        typ = o.get("type")
        tb = np.empty_like(typ)
        tb[o.get("id0")] = typ
        owner = np.repeat(np.arange(cnt.size), cnt)
        keep = _list_keep(typ[owner], tb[ids])
        if drop is not None:  # leave out the drop-th recorded entry
            keep[np.nonzero(keep)[0][drop]] = False
        self.out = (np.bincount(owner[keep], minlength=cnt.size).astype(np.int32), ids[keep], codes[keep])

    def neighbor_lists(self, slot):
        return self.out


def test_neighbor_list_comparison_helper():
    from tests.common import compare_neighbor_lists
    sc = build_scene(tiny_cfg(method=1, dim=2))
    o32 = O.Oracle(sc.desc(), 32)
    O.load_scene(o32, sc)
    o32.step(20)
    o32.run("sort")
    o64 = O.Oracle(sc.desc(), 64)
    for f in ("cell", "offset", "vel", "density", "mass", "m_V", "type", "id0"):
        o64.set(f, o32.get(f))
    o64.run("sort")
    nent, knife, diff = compare_neighbor_lists(_ListsFromOracle(o32), 0, o32, o64)
    assert nent > 0 and knife > 0 and diff <= knife  # the lattice holds pairs at r_s
    with pytest.raises(AssertionError):  # one dropped entry is found
        compare_neighbor_lists(_ListsFromOracle(o32, drop=123), 0, o32, o64)
