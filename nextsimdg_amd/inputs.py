"""Synthetic inputs of the benchmark box (the Mehlmann et al. 2021 test the reference
ships: run/make_init_benchmark.py:1-40, physics/src/BenchmarkCoordinates.cpp:20-42,
physics/src/modules/OceanBoundaryModule/BenchmarkOcean.cpp:15-37).

Host-side input generation only (numpy): coordinates, land mask, initial ice state and
the steady ocean gyre. The moving cyclone wind is evaluated on the device by
MEVPDynamicsKernel.benchmark_wind. Cells are indexed x-fastest with x = i dx; the
reference's netCDF generator writes hice[ix, iy] into the (y, x) slot, i.e. it
transposes the initial thickness pattern, which is not reproduced here.
"""
import numpy as np

L_BOX = 512e3  # metres


def box_coords(nx, ny, L=L_BOX):
    """Vertex coordinates, (nx+1)(ny+1) x 2, vertex ix + (nx+1) iy at (ix dx, iy dy)."""
    dx, dy = L / nx, L / ny
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return np.stack([(i * dx).ravel(), (j * dy).ravel()], axis=1).astype(np.float64)


def box_mask(nx, ny):
    """Ocean everywhere except the outer ring of cells (make_init_benchmark.py:20-24)."""
    m = np.ones((ny, nx))
    m[0, :] = m[-1, :] = 0.0
    m[:, 0] = m[:, -1] = 0.0
    return m.ravel()


def cell_corners(nx, ny, L=L_BOX):
    """Lower-left corner of every cell (BenchmarkCoordinates.cpp:29-38)."""
    dx, dy = L / nx, L / ny
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    return (i * dx).ravel(), (j * dy).ravel()


def ocean_gyre(nx, ny, L=L_BOX, vmax=0.01):
    """Steady rotating ocean current (BenchmarkOcean.cpp:30-35)."""
    x, y = cell_corners(nx, ny, L)
    fx, fy = x / (nx * (L / nx)), y / (ny * (L / ny))
    return vmax * (2 * fy - 1), vmax * (1 - 2 * fx)


def ice_plain(nx, ny, seed=0, L=L_BOX, noise=1e-3):
    """A = 1, h = 0.3 + 0.005 (sin(6e-5 x) + sin(3e-5 y)) + N(0, noise) (x, y in metres)."""
    x, y = cell_corners(nx, ny, L)
    rng = np.random.default_rng(seed)
    h = 0.3 + 0.005 * (np.sin(60e-3 * x / 1e3) + np.sin(30e-3 * y / 1e3)) + rng.normal(0.0, noise, x.size)
    return h, np.ones(x.size)


def ice_heterogeneous(nx, ny, seed=0):
    """h = 0.3 + LogNormal(0, 0.3), A ~ Beta(8, 1)."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    return 0.3 + rng.lognormal(0.0, 0.3, n), rng.beta(8.0, 1.0, n)


def ice_leads(nx, ny, seed=0):
    """Heterogeneous ice with open-water leads: A in [0, 1], zero along random lines."""
    rng = np.random.default_rng(seed)
    h, A = ice_heterogeneous(nx, ny, seed)
    A = A.reshape(ny, nx)
    for _ in range(max(1, nx // 64)):
        r = rng.integers(0, ny)
        A[r, :] = 0.0
        c = rng.integers(0, nx)
        A[:, c] *= 0.2
    return h, A.ravel()


def cyclone_offset(seed, L=L_BOX, frac=0.05):
    """Seeded shift of the cyclone centre, a fraction of the box."""
    rng = np.random.default_rng(seed + 12345)
    return tuple(rng.uniform(-frac * L, frac * L, 2))


def storms(seed, n=3, L=L_BOX, peak=(10.0, 25.0), frac=0.3):
    """n seeded storms, each a copy of the benchmark cyclone on its track shifted by up to
    frac L, with a peak wind speed uniform in peak (m/s): rows (xoff, yoff, vmax), where
    vmax = peak e (the cyclone's speed A vmax r e^(-k r) peaks at vmax / e, r = 1 / k)."""
    rng = np.random.default_rng(seed + 54321)
    off = rng.uniform(-frac * L, frac * L, (n, 2))
    vmax = rng.uniform(peak[0], peak[1], n) * np.e
    return np.column_stack([off, vmax])


def benchmark_setup(nx, ny, seed=0, kind="plain"):
    """Everything a run needs: coords, mask, h, A, ocean u, v (all float64)."""
    coords, mask = box_coords(nx, ny), box_mask(nx, ny)
    if kind == "plain":
        h, A = ice_plain(nx, ny, seed)
    elif kind == "heterogeneous":
        h, A = ice_heterogeneous(nx, ny, seed)
    else:
        h, A = ice_leads(nx, ny, seed)
    uo, vo = ocean_gyre(nx, ny)
    return dict(coords=coords, mask=mask, h=h, A=A, uocean=uo, vocean=vo)


def rotate_pole_from_greenland(lon, lat):
    """RotatePoleFromGreenland (ParametricMesh.hpp:159-180), the inverse of the rotation
    the kernel applies to spherical meshes: vertices given to the kernel after this
    rotation end up at (lon, lat) inside it."""
    x, y, z = np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)
    aw, bw = np.radians(-40.0), np.radians(-15.0)
    x1, y1, z1 = np.cos(bw) * x - np.sin(bw) * z, y, np.sin(bw) * x + np.cos(bw) * z
    x2, y2, z2 = np.cos(aw) * x1 - np.sin(aw) * y1, np.sin(aw) * x1 + np.cos(aw) * y1, z1
    return np.arctan2(y2, x2), np.arcsin(z2)


def sphere_patch_coords(nx, ny, lon0=-0.6, lon1=0.6, lat0=0.9, lat1=1.3):
    """Vertices (radians) of a longitude-latitude patch, as the kernel sees it after its
    pole rotation (the default: 1.2 rad x 0.4 rad around 63N, cells of ~5-7 km at 512^2)."""
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    lon = lon0 + (lon1 - lon0) * i.ravel() / nx
    lat = lat0 + (lat1 - lat0) * j.ravel() / ny
    return np.stack(rotate_pole_from_greenland(lon, lat), axis=1)
