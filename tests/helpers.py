"""Shared test helpers: mesh coordinates, CG node positions, simple mesh texts."""
import gzip
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

DEFAULT_PARAMS = dict(
    rho_ice=900.0, rho_atm=1.3, rho_ocean=1026.0, C_atm=1.2e-3, C_ocean=5.5e-3,
    F_atm=1.2e-3 * 1.3, F_ocean=5.5e-3 * 1026.0, fc=1.45842e-4, ocean_turning_angle=0.0,
    gravity=9.81, Pstar=27500.0, DeltaMin=2e-9, alpha=1500.0, beta=1500.0, nSteps=100,
)


def box_coords(nx, ny, dx, dy=None, skew=0.0):
    """Vertex coordinates (N_v x 2, node ix + (nx+1) iy) of an affine box mesh."""
    dy = dx if dy is None else dy
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    x = i * dx + skew * j * dy
    y = j * dy
    return np.stack([x.ravel(), y.ravel()], axis=1).astype(np.float64)


def cg_node_coords(nx, ny, dx, dy=None, cg=2, skew=0.0):
    """Physical coordinates of the CG nodes (node ix + (cg nx + 1) iy) of an affine box mesh."""
    dy = dx if dy is None else dy
    j, i = np.meshgrid(np.arange(cg * ny + 1), np.arange(cg * nx + 1), indexing="ij")
    y = j * dy / cg
    x = i * dx / cg + skew * y
    return x.ravel(), y.ravel()


def smesh_text(coords, nx, ny, landmask=None, dirichlet=(), periodic=()):
    """A version-2.0 .smesh text (format of dynamics/src/ParametricMesh.cpp:8-190)."""
    lines = ["ParametricMesh 2.0", f"{nx} {ny}"]
    lines += [f"{float(x)!r} {float(y)!r}" for x, y in coords]
    if landmask is None:
        lines.append("landmask 0")
    else:
        lines.append(f"landmask {nx * ny}")
        lines += [str(int(v)) for v in landmask]
    lines.append(f"dirichlet {len(dirichlet)}")
    lines += [f"{e} {s}" for e, s in dirichlet]
    lines.append(f"periodic {len(periodic)}")
    for seg in periodic:
        lines.append(str(len(seg)))
        lines += [f"{a} {b} {c}" for a, b, c in seg]
    return "\n".join(lines) + "\n"


def read_nh_smesh():
    with gzip.open(os.path.join(GOLDEN, "25km_NH.smesh.gz"), "rt") as f:
        return f.read()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
