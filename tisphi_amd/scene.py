"""Scene construction: JSON configuration -> engine parameters + host arrays.

This is the init-time part of the reference's ParticleSystem and solver
constructors, done on the host in float64:

* geometry  (A1)  — eng/particle_system.py:14-58: d = 2r, h = kh*d,
  r_s = kappa*h, grid = ceil(kappa*kh)*d, virtual domain = domain +- grid
  (2D: z extent d and grid_num[2] = 1), grid_num = ceil(vsize/grid).
* particles (A2)  — eng/particle_system.py:72-108, 135-200, 274-314:
  blocks in file order, then OBJ bodies (voxelised, eng/voxelize.py),
  then dummy walls, then the repulsive layer
  (boundary 3 / 4: spacing r, particle_func.py:346-374); m_V = d^dim,
  mass = m_V*rho with rho = 0 for boundary particles; id0 = generation
  index (set_id0, :208-211).
* dt        (A16) — eng/solver_sph_base.py:209-212 with the solvers' sound
  speeds (solver_sph_wc.py:17-21, solver_sph_muI.py:23-25,
  solver_sph_dp.py:20-22).
* DP initial geostatic stress — eng/solver_sph_base.py:249-260.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .eng.particle_func import (
    MAT_DUMMY, MAT_FLUID, MAT_REP, MAT_RIGID, MAT_SOIL, calc_cube_particle_num,
    calc_dummy_boundary, calc_rep_boundary, chk_block_in_domain, count_boundary, cube_positions,
    get_material, set_material)

METHOD_WC, METHOD_MUI, METHOD_DP = 1, 2, 3
BDY_NONE, BDY_COLLISION, BDY_DUMMY, BDY_REP, BDY_DUMMY_REP = 0, 1, 2, 3, 4


def calc_dt_cfl(cfl, smoothing_len, vsound, dt_min):
    """dt = max(dt_min, c_dt - c_dt % dt_min), c_dt = cfl*h/c."""
    dt = cfl * smoothing_len / vsound
    return max(dt_min, dt - dt % dt_min)


@dataclass
class SceneGeometry:
    dim: int
    domain_start: np.ndarray
    domain_end: np.ndarray
    domain_size: np.ndarray
    particle_radius: float
    kappa: float
    kh: float
    particle_diameter: float
    smoothing_len: float
    support_radius: float
    m_V0: float
    grid_size: float
    vdomain_start: np.ndarray
    vdomain_end: np.ndarray
    vdomain_size: np.ndarray
    grid_num: np.ndarray
    grid_num_total: int


def derive_geometry(cfg) -> SceneGeometry:
    dim = 3 if not cfg.get_cfg("is2D") else 2
    ds = np.array(cfg.get_cfg("domainStart"), dtype=np.float64)
    de = np.array(cfg.get_cfg("domainEnd"), dtype=np.float64)
    dsize = de - ds
    r = cfg.get_cfg("particleRadius")
    kappa = cfg.get_cfg("kappa")
    kh = cfg.get_cfg("kh")
    d = 2 * r
    h = kh * d
    rs = kappa * h
    grid = float(math.ceil(kappa * kh)) * d
    vds = ds - grid
    vde = de + grid
    vsize = vde - vds
    if dim == 2:
        de[2] = ds[2] + d
        dsize[2] = d
        vds[2] = ds[2]
        vde[2] = vds[2] + d
        vsize[2] = d
    gnum = np.ceil(vsize / grid).astype(np.int64)
    return SceneGeometry(dim, ds, de, dsize, r, kappa, kh, d, h, rs, d ** dim, grid,
                         vds, vde, vsize, gnum, int(np.prod(gnum[0:dim])))


@dataclass
class Scene:
    """Everything the engine needs for one scene."""
    cfg: object
    geo: SceneGeometry
    method: int
    boundary: int
    kernel: int
    kernel_corr: int
    time_integration: int
    xsph: int
    gravity: np.ndarray
    dt: float
    dt_min: float
    material: dict
    vsound: float
    mat_index: list
    mat_fluid: list
    mat_soil: list
    mat_rigid: list
    block_counts: list = field(default_factory=list)
    body_counts: list = field(default_factory=list)  # voxels per OBJ body
    n_dummy: int = 0
    n_rep: int = 0
    # host arrays, generation order
    pos: np.ndarray = None
    vel: np.ndarray = None
    density: np.ndarray = None
    mass: np.ndarray = None
    m_V: np.ndarray = None
    mat_type: np.ndarray = None
    obj_id: np.ndarray = None
    mat_id: np.ndarray = None
    color0: np.ndarray = None  # initial colour per particle (float32 values)
    # rigid blocks: dicts obj, dynamic, rest_cm (mass-weighted mean of the
    # initial positions, calc_rigid_rest_m, solver_sph_base.py:517-518)
    rigid: list = field(default_factory=list)
    stress: np.ndarray = None  # (n, 6) xx,yy,zz,xy,yz,zx

    @property
    def n(self):
        if self.pos is None:
            return int(sum(self.block_counts) + sum(self.body_counts) + self.n_dummy + self.n_rep)
        return int(self.pos.shape[0])

    @property
    def n_flow(self):
        return int(np.count_nonzero((self.mat_type == MAT_FLUID) | (self.mat_type == MAT_SOIL)))

    def desc(self):
        """Fields of tsph_scene_desc as a dict (see include/tisphi.h)."""
        g = self.geo
        m = self.material
        return dict(
            dim=g.dim, method=self.method, boundary=self.boundary, kernel=self.kernel,
            kernel_corr=self.kernel_corr, time_integration=self.time_integration,
            xsph=int(bool(self.xsph)), grid_num=[int(x) for x in g.grid_num],
            n_particles=self.n, vdomain_start=list(g.vdomain_start),
            domain_start=list(g.domain_start), domain_end=list(g.domain_end),
            grid_size=g.grid_size, particle_radius=g.particle_radius,
            particle_diameter=g.particle_diameter, smoothing_len=g.smoothing_len,
            support_radius=g.support_radius, gravity=list(self.gravity), dt=self.dt,
            density0=float(m["density0"]),
            viscosity=float(m.get("viscosity", 0.0)),
            stiffness=float(m.get("stiffness", 0.0)),
            exponent=float(m.get("exponent", 0.0)),
            cohesion=float(m.get("cohesion", 0.0)),
            friction=float(m.get("friction", 0.0)) / 180 * math.pi,
            young=float(m.get("EYoungMod", 0.0)),
            poisson=float(m.get("poison", 0.0)),
            dilatancy=float(m.get("dilatancy", 0.0)) / 180 * math.pi,
            vsound=self.vsound,
        )


def build_scene(cfg, generate=True) -> Scene:
    """Derive parameters and (optionally) generate particles for ``cfg``."""
    geo = derive_geometry(cfg)
    method = cfg.get_cfg("simulationMethod")
    if method not in (METHOD_WC, METHOD_MUI, METHOD_DP):
        raise NotImplementedError(f"Solver type {method} has not been implemented.")
    boundary = cfg.get_cfg("boundary")
    if boundary not in (BDY_NONE, BDY_COLLISION, BDY_DUMMY, BDY_REP, BDY_DUMMY_REP):
        raise NotImplementedError(f"boundary {boundary} is not built")
    kernel = cfg.get_cfg("kernel")
    kcorr = cfg.get_cfg("kernelCorrection")
    if kcorr not in (0, 1, 2):
        # 2 is the reference's MLS switch: kernel_deriv_corr's branch for it
        # is empty, so every corrected gradient is the zero vector
        raise NotImplementedError(f"kernelCorrection {kcorr} is not built")
    ti_flag = cfg.get_cfg("timeIntegration")
    if ti_flag not in (1, 2, 4):
        raise NotImplementedError(f"timeIntegration {ti_flag} is not built")

    mat_index, fluids, soils, rigids = set_material(cfg.get_materials())
    if method == METHOD_WC:
        material = fluids[0]
        vsound = 60.0
    elif method == METHOD_MUI:
        material = soils[0]
        vsound = 24.0
    else:
        material = soils[0]
        vsound = math.sqrt(material["EYoungMod"] / material["density0"])
    dt_min = cfg.get_cfg("timeStepSizeMin")
    dt = calc_dt_cfl(0.2, geo.smoothing_len, vsound, dt_min)
    gravity = np.array(cfg.get_cfg("gravitation"), dtype=np.float64)
    sc = Scene(cfg=cfg, geo=geo, method=method, boundary=boundary, kernel=kernel,
               kernel_corr=kcorr, time_integration=ti_flag, xsph=cfg.get_cfg("xsph"),
               gravity=gravity, dt=dt, dt_min=dt_min, material=material, vsound=vsound,
               mat_index=mat_index, mat_fluid=fluids, mat_soil=soils, mat_rigid=rigids)

    blocks = cfg.get_blocks()
    for b in blocks:
        chk_block_in_domain(geo.domain_start, geo.domain_end, b["translation"], b["size"], geo.dim)
        n, _ = calc_cube_particle_num(b["translation"], b["size"], geo.dim, offset=geo.particle_diameter)
        sc.block_counts.append(n)
    bodies = cfg.get_bodies()
    body_pts = []
    if bodies and geo.dim != 3:
        # a voxelised mesh is a 3D point set; the reference has no 2D form
        raise NotImplementedError("OBJ bodies need a 3D scene")
    if bodies:
        import os
        from .eng.voxelize import load_body
        path = getattr(cfg, "scene_file_path", None)
        base = os.path.dirname(os.path.abspath(path)) if path else None
        for b in bodies:
            # voxelised at the particle diameter (particle_system.py:82-91)
            body_pts.append(load_body(b, geo.particle_diameter, base))
            p = body_pts[-1]
            if p.size and (np.any(p < np.asarray(geo.domain_start)) or np.any(p > np.asarray(geo.domain_end))):
                # the reference would index its cell grid out of range
                raise ValueError(f"body {b['objectId']} leaves the domain")
            sc.body_counts.append(int(body_pts[-1].shape[0]))
    dummy = None
    if boundary in (BDY_DUMMY, BDY_DUMMY_REP):
        dummy = calc_dummy_boundary(geo.dim, geo.domain_start, geo.domain_end,
                                    geo.vdomain_start, geo.vdomain_end)
        sc.n_dummy = count_boundary(dummy, geo.dim, geo.particle_diameter)
    rep = None
    if boundary in (BDY_REP, BDY_DUMMY_REP):
        # one layer at spacing r on the walls (particle_system.py:101-104, 147-148)
        rep = calc_rep_boundary(geo.dim, geo.domain_start, geo.domain_end, geo.particle_radius)
        sc.n_rep = count_boundary(rep, geo.dim, geo.particle_radius)
    if not generate:
        return sc

    pos, vel, rho, mtype, oid, mid, col = [], [], [], [], [], [], []
    lists = (fluids, soils, rigids)
    for b in blocks:
        mat = get_material(mat_index, lists, b["materialId"])
        # the material colour / 255 as float32 (particle_system.py:159)
        col.append(np.array([ic / 255 for ic in mat.get("color", [0, 0, 0])], dtype=np.float32))
        p = cube_positions(np.array(b["translation"], dtype=np.float64),
                           np.array(b["size"], dtype=np.float64), geo.dim, geo.particle_diameter)
        n = p.shape[0]
        pos.append(p)
        vel.append(np.tile(np.array(b["velocity"], dtype=np.float64), (n, 1)))
        rho.append(np.full(n, float(mat["density0"])))
        mtype.append(np.full(n, mat["matType"], dtype=np.int32))
        oid.append(np.full(n, b["objectId"], dtype=np.int32))
        mid.append(np.full(n, b["materialId"], dtype=np.int32))
        col[-1] = np.tile(col[-1], (n, 1))
    for b, p in zip(bodies, body_pts):
        # init_body (particle_system.py:176-200): after every block, before
        # the boundary particles
        mat = get_material(mat_index, lists, b["materialId"])
        n = p.shape[0]
        pos.append(p)
        vel.append(np.tile(np.array(b["velocity"], dtype=np.float64), (n, 1)))
        rho.append(np.full(n, float(mat["density0"])))
        mtype.append(np.full(n, mat["matType"], dtype=np.int32))
        oid.append(np.full(n, b["objectId"], dtype=np.int32))
        mid.append(np.full(n, b["materialId"], dtype=np.int32))
        col.append(np.tile(np.array([ic / 255 for ic in mat.get("color", [0, 0, 0])],
                                    dtype=np.float32), (n, 1)))
    if dummy is not None:
        for lo, hi in dummy:
            p = cube_positions(lo, hi - lo, geo.dim, geo.particle_diameter)
            n = p.shape[0]
            pos.append(p)
            vel.append(np.zeros((n, 3)))
            rho.append(np.zeros(n))
            mtype.append(np.full(n, MAT_DUMMY, dtype=np.int32))
            oid.append(np.full(n, MAT_DUMMY, dtype=np.int32))
            mid.append(np.zeros(n, dtype=np.int32))
            # add_boundary colour [153, 153, 255] / 255 (particle_system.py:146)
            col.append(np.tile((np.array([153, 153, 255]) / 255).astype(np.float32), (n, 1)))
    if rep is not None:
        for lo, hi in rep:
            p = cube_positions(lo, hi - lo, geo.dim, geo.particle_radius)
            n = p.shape[0]
            pos.append(p)
            vel.append(np.zeros((n, 3)))
            rho.append(np.zeros(n))
            mtype.append(np.full(n, MAT_REP, dtype=np.int32))
            oid.append(np.full(n, MAT_REP, dtype=np.int32))
            mid.append(np.zeros(n, dtype=np.int32))
            # [170, 17, 255] / 255 (particle_system.py:148)
            col.append(np.tile((np.array([170, 17, 255]) / 255).astype(np.float32), (n, 1)))
    sc.pos = np.concatenate(pos) if pos else np.zeros((0, 3))
    sc.vel = np.concatenate(vel) if vel else np.zeros((0, 3))
    sc.density = np.concatenate(rho) if rho else np.zeros(0)
    sc.mat_type = np.concatenate(mtype) if mtype else np.zeros(0, np.int32)
    sc.obj_id = np.concatenate(oid) if oid else np.zeros(0, np.int32)
    sc.mat_id = np.concatenate(mid) if mid else np.zeros(0, np.int32)
    sc.color0 = np.concatenate(col) if col else np.zeros((0, 3), np.float32)
    sc.m_V = np.full(sc.n, geo.m_V0)
    sc.mass = geo.m_V0 * sc.density
    for b in list(blocks) + list(bodies):
        mat = get_material(mat_index, lists, b["materialId"])
        if mat["matType"] != MAT_RIGID:
            continue
        sel = (sc.mat_type == MAT_RIGID) & (sc.obj_id == b["objectId"])
        m = sc.mass[sel]
        cm = (m[:, None] * sc.pos[sel]).sum(axis=0) / m.sum()
        sc.rigid.append(dict(obj=int(b["objectId"]), dynamic=bool(b.get("isDynamic", False)),
                             rest_cm=cm))
    sc.stress = np.zeros((sc.n, 6))
    if method == METHOD_DP:
        sc.stress = init_stress_dp(sc)
    return sc


def init_stress_dp(sc):
    """Geostatic K0 stress of soil particles (solver_sph_base.py:249-260)."""
    st = np.zeros((sc.n, 6))
    soil = sc.mat_type == MAT_SOIL
    if not np.any(soil):
        return st
    ymax = np.max(sc.pos[soil, 1])
    fric = sc.material["friction"] / 180 * math.pi
    k0 = 1.0 - math.sin(fric)
    rho0 = sc.material["density0"]
    ver = rho0 * sc.gravity[1] * (ymax - sc.pos[soil, 1])
    st[soil, 0] = k0 * ver
    st[soil, 1] = ver
    st[soil, 2] = k0 * ver
    return st
