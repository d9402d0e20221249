"""Host-side particle generation: lattices, materials, dummy walls.

Follows the reference's generation order and counts exactly, because the
particle order fixes the summation order of every neighbour sum:

* ``calc_cube_particle_num``  — eng/particle_func.py:253-259 (``np.arange``
  per axis from ``t + off/2`` to ``t + size + 1e-5`` in steps of ``off``).
* ``cube_positions``          — eng/particle_func.py:176-208 (``meshgrid``
  with ``indexing='ij'``, x slowest; z = 0 in 2D).
* ``calc_dummy_boundary``     — eng/particle_func.py:316-343 (wall boxes; no
  top wall; 2D order x-, y-, x+; 3D order x-, z+, x+, z-, y-).
* ``set_material/get_material`` — eng/particle_func.py:218-244 (a block's
  ``materialId`` indexes the Materials list position).
* ``chk_block_in_domain``     — eng/particle_func.py:246-251.

Everything here is float64 numpy, like the reference's host code.
"""
from functools import reduce

import numpy as np

MAT_FLUID, MAT_SOIL, MAT_RIGID = 1, 2, 11
MAT_DUMMY, MAT_REP = -1, -2


def calc_cube_particle_num(translation, size, dim=3, offset=None):
    """Per-axis lattice coordinates and their product count."""
    step = offset if offset is not None else 0.1
    axes = []
    for a in range(dim):
        off = step if size[a] >= 0 else -step
        axes.append(np.arange(translation[a] + off / 2.0,
                              translation[a] + size[a] + 1e-5, off))
    return reduce(lambda x, y: x * y, [len(ax) for ax in axes]), axes


def count_cube_num(pos_bld, pos_fru, dim, offset):
    n, _ = calc_cube_particle_num(pos_bld, np.asarray(pos_fru) - np.asarray(pos_bld),
                                  dim, offset=offset)
    return n


def cube_positions(lower_corner, cube_size, dim, offset):
    """(n, 3) float64 lattice positions of one box, x slowest."""
    n, axes = calc_cube_particle_num(lower_corner, cube_size, dim, offset=offset)
    axes = list(axes)
    if dim == 2:
        axes.append(np.array([0.0]))
    grid = np.array(np.meshgrid(*axes, sparse=False, indexing="ij"), dtype=np.float64)
    pos = grid.reshape(3, -1).T.copy()
    assert pos.shape[0] == n
    return pos


def chk_block_in_domain(domain_start, domain_end, translation, size, dim):
    ok_start = all(translation[a] - domain_start[a] >= 0.0 for a in range(dim))
    ok_end = all(translation[a] + size[a] - domain_end[a] <= 0.0 for a in range(dim))
    assert ok_start and ok_end, "Block is not in domain!"


def calc_dummy_boundary(dim, ds, de, vds, vde):
    """Wall boxes [lower, upper] of the dummy-particle boundary."""
    A = np.array
    if dim == 3:
        return [
            [A([vds[0], ds[1], vds[2]]), A([ds[0], de[1], de[2]])],
            [A([vds[0], ds[1], de[2]]), A([de[0], de[1], vde[2]])],
            [A([de[0], ds[1], ds[2]]), A([vde[0], de[1], vde[2]])],
            [A([ds[0], ds[1], vds[2]]), A([vde[0], de[1], ds[2]])],
            [A(vds, dtype=np.float64), A([vde[0], ds[1], vde[2]])],
        ]
    return [
        [A([vds[0], ds[1], ds[2]]), A([ds[0], de[1], de[2]])],
        [A([vds[0], vds[1], ds[2]]), A([vde[0], ds[1], de[2]])],
        [A([de[0], ds[1], ds[2]]), A([vde[0], de[1], de[2]])],
    ]


def calc_rep_boundary(dim, ds, de, r):
    """Wall boxes [lower, upper] of the repulsive-particle layer: one layer
    of particles at spacing r on the walls, the top left open
    (particle_func.py:346-374)."""
    A = np.array
    t = r / 2
    if dim == 3:
        return [
            [A([ds[0] - t, ds[1] + t, ds[2] - t]), A([ds[0] + t, de[1] - t, de[2] - t])],
            [A([ds[0] - t, ds[1] + t, de[2] - t]), A([de[0] - t, de[1] - t, de[2] + t])],
            [A([de[0] - t, ds[1] + t, ds[2] + t]), A([de[0] + t, de[1] - t, de[2] + t])],
            [A([ds[0] + t, ds[1] + t, ds[2] - t]), A([de[0] + t, de[1] - t, ds[2] + t])],
            [A(ds, dtype=np.float64) - t, A([de[0], ds[1], de[2]]) + t],
        ]
    return [
        [A([ds[0] - t, ds[1] + t, ds[2]]), A([ds[0] + t, de[1] - t, de[2]])],
        [A([ds[0] - t, ds[1] - t, ds[2]]), A([de[0] + t, ds[1] + t, de[2]])],
        [A([de[0] - t, ds[1] + t, ds[2]]), A([de[0] + t, de[1] - t, de[2]])],
    ]


def count_boundary(boundary, dim, offset):
    return sum(count_cube_num(b[0], b[1], dim, offset) for b in boundary)


def set_material(materials):
    """Split materials by type; returns (mat_index, fluids, soils, rigids)
    where mat_index[k] = [type, position within its type list]."""
    mat_index, water, soil, rigid = [], [], [], []
    for mat in materials:
        t = mat["matType"]
        if t == MAT_FLUID:
            mat_index.append([MAT_FLUID, len(water)])
            water.append(mat)
        elif t == MAT_SOIL:
            mat_index.append([MAT_SOIL, len(soil)])
            soil.append(mat)
        elif t == MAT_RIGID:
            mat_index.append([MAT_RIGID, len(rigid)])
            rigid.append(mat)
    return mat_index, water, soil, rigid


def get_material(mat_index, lists, i):
    """Material dict of list position ``i`` (lists = (fluids, soils, rigids))."""
    t, k = mat_index[i]
    if t == MAT_FLUID:
        return lists[0][k]
    if t == MAT_SOIL:
        return lists[1][k]
    if t == MAT_RIGID:
        return lists[2][k]
    raise KeyError(i)
