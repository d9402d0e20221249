"""CPU restatement of the reference's per-frame colouring, the checker of the
device colour feed (test infrastructure: only tests import it).

    assign_value_color   eng/solver_sph_base.py:721-789
    v_maxmin             eng/particle_system.py:387-397
    set_color            eng/particle_system.py:399-407
    jet color_map        eng/colormap.py:8-48

FP64 throughout (the reference runs with default_fp=f64,
run_simulation.py:23), on fields given as numpy arrays in the current
particle order (for instance the engine's downloads): the colour of a
particle is float32 (its colour field is a vec3f32, particle_func.py:16).
"""
import numpy as np

FLUID, SOIL, DUMMY = 1, 2, -1

# (h, wl, wr, c) of jetR / jetG / jetB (colormap.py:40-42)
JET = ((1.5, 0.37, 0.37, 0.75), (1.5, 0.37, 0.37, 0.5), (1.5, 0.37, 0.37, 0.25))


def _clamp(x):
    """ColorMap.clamp (colormap.py:17-18): max(0, min(1, x)), with C's
    fmin / fmax for a NaN argument (vmax == vmin gives 0 * inf): the number
    wins.  What Taichi's min / max do there is parity unpinned."""
    return np.fmax(0.0, np.fmin(1.0, x))


def jet(x):
    """color_map (colormap.py:44-48): per channel
    clamp((w - |clamp(x) - c|) / w * h), w = wl if x < c else wr
    (ColorMap.map, colormap.py:20-27).  Returns (n, 3) float32."""
    x = np.asarray(x, dtype=np.float64)
    out = []
    for h, wl, wr, c in JET:
        w = np.where(x < c, wl, wr)
        out.append(_clamp((w - np.abs(_clamp(x) - c)) / w * h))
    return np.stack(out, axis=-1).astype(np.float32)


def _norm(v):
    v = np.asarray(v, dtype=np.float64)
    return np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])


def assign_value_color(f, title, show_bdy):
    """Colour value per particle (solver_sph_base.py:721-789) from the field
    dict ``f`` (keys as the engine's field names, plus ``x0_by_id``: the
    initial positions indexed by id0).  Particles that are neither real
    nor, with show_bdy, dummy get 0: the value the reference never writes.
    Title 101 (ti.random) is not restated."""
    t = f["type"]
    sel = (t > 0) | (bool(show_bdy) & (t == DUMMY))
    n = t.shape[0]
    st = f.get("stress", np.zeros((n, 6)))
    table = {
        1: lambda: f["id0"].astype(np.float64),
        2: lambda: f["density"],
        21: lambda: f["d_density"],
        3: lambda: _norm(f["vel"]),
        31: lambda: f["vel"][:, 0],
        32: lambda: -f["vel"][:, 1],
        33: lambda: f["vel"][:, 2],
        34: lambda: _norm(f["d_vel"]),
        35: lambda: f["d_vel"][:, 0],
        36: lambda: f["d_vel"][:, 1],
        37: lambda: f["d_vel"][:, 2],
        4: lambda: _norm(f["pos"]),
        41: lambda: f["pos"][:, 0],
        42: lambda: f["pos"][:, 1],
        43: lambda: f["pos"][:, 2],
        44: lambda: _norm(f["pos"] - f["x0_by_id"][f["id0"]]),
        51: lambda: st[:, 0],
        52: lambda: -st[:, 1],
        53: lambda: -st[:, 2],
        54: lambda: st[:, 3],
        55: lambda: st[:, 4],
        56: lambda: st[:, 5],
        57: lambda: -(st[:, 0] + st[:, 1] + st[:, 2]) / 3.0,
        61: lambda: f.get("strain_equ", np.zeros(n)),
        62: lambda: f.get("strain_equ_p", np.zeros(n)),
        7: lambda: f["pressure"],
        8: lambda: f.get("flag_retmap", np.zeros(n)).astype(np.float64),
        100: lambda: f["cell"].astype(np.float64),
        102: lambda: f["cspm_L"][:, 0] if f.get("kernel_corr", 0) == 1 else np.zeros(n),
        103: lambda: _norm(f["v_tmp"]),
    }
    v = np.asarray(table[title](), dtype=np.float64) if title in table else np.zeros(n)
    return np.where(sel, v, 0.0)


def v_maxmin(val, typ, givenmax=-1.0, givenmin=-1.0, fixmax=0, fixmin=0):
    """particle_system.py:387-397: max / min over flow particles
    (-inf / +inf when there are none), then
    vmax if givenmax == -1 or (vmax < givenmax and not fixmax) else givenmax,
    and the mirror rule for vmin."""
    flow = (typ == FLUID) | (typ == SOIL)
    vmax = float(np.max(val[flow])) if flow.any() else -np.inf
    vmin = float(np.min(val[flow])) if flow.any() else np.inf
    vmax = vmax if givenmax == -1 or (vmax < givenmax and not fixmax) else givenmax
    vmin = vmin if givenmin == -1 or (vmin > givenmin and not fixmin) else givenmin
    return vmax, vmin


def set_color(val, typ, id0, color0_by_id, vmax, vmin, group, show_bdy):
    """particle_system.py:399-407: jet((val - vmin) / (vmax - vmin)) for the
    colour group (0 flow, 1 real, 2 real and, with show_bdy, dummy); every
    other particle keeps its initial colour."""
    flow = (typ == FLUID) | (typ == SOIL)
    real = typ > 0
    member = {0: flow, 1: real, 2: real | (bool(show_bdy) & (typ == DUMMY))}[group]
    with np.errstate(divide="ignore", invalid="ignore"):
        vrange1 = np.float64(1.0) / (np.float64(vmax) - np.float64(vmin))  # inf if equal
        c = jet((val - vmin) * vrange1)
    return np.where(member[:, None], c, color0_by_id[id0].astype(np.float32))
