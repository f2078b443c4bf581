"""Restart files (nsdg_dyn_write_restart / nsdg_dyn_read_restart) in the layout of the
reference's ParaGridIO::dumpModelState (core/src/ParaGridIO.cpp:272-349), netCDF classic
with 64-bit offsets.

  * an independent reader (scipy.io.netcdf_file) opens the file and finds the reference's
    dimensions (ModelArrayDetails.cpp:28-41) with the reference's lengths, the variables
    with the reference's dimension order (components fastest) and missing_value, and data
    equal to what the kernel returns through getDGData / getDG0Data / get_cg;
  * a device-state round trip: a fresh kernel loaded from the file continues bit for bit
    like the kernel that wrote it (mEVP and BBM);
  * a file written by that independent writer loads, and mismatched meshes or rheologies
    are refused with NSDG_ERR_ARG.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(kind, nx, steps, seed=4):
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    s = inputs.benchmark_setup(nx, nx, seed)
    k = nd.BBMDynamicsKernel(2, 6) if kind == "bbm" else nd.MEVPDynamicsKernel(2, 6)
    k.initialise(s["coords"], False, s["mask"], nx, nx)
    for name, key in (("hice", "h"), ("cice", "A"), ("uocean", "uocean"), ("vocean", "vocean")):
        k.setData(name, s[key], 1)
    k.setData("ssh", 0 * s["h"], 1)
    for i in range(steps):
        k.benchmark_wind(inputs.L_BOX, 120.0 * i, 0.0, 0.0, 30.0)
        k.update(120.0)
    return s, k


@pytest.mark.parametrize("kind", ["mevp", "bbm"])
def test_restart_file_layout_read_independently(kind, tmp_path):
    from scipy.io import netcdf_file

    _say(f"restart layout {kind}")
    nx = 24
    s, k = _run(kind, nx, 3)
    path = str(tmp_path / "restart.nc")
    k.write_restart(path, 3600.0)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.type == b"parametric_rectangular"
        assert f.time == 3600.0
        dims = dict(f.dimensions)
        assert dims == {"xdim": nx, "ydim": nx, "zdim": 1, "xvertex": nx + 1, "yvertex": nx + 1,
                        "x_cg": 2 * nx + 1, "y_cg": 2 * nx + 1, "dg_comp": 6, "dgstress_comp": 8, "ncoords": 2}
        v = f.variables
        assert v["hice"].dimensions == ("ydim", "xdim", "dg_comp")
        assert v["s11"].dimensions == ("ydim", "xdim", "dgstress_comp")
        assert v["u"].dimensions == ("ydim", "xdim")
        assert v["u_cg"].dimensions == ("y_cg", "x_cg")
        assert v["coords"].dimensions == ("yvertex", "xvertex", "ncoords")
        assert v["hice"].missing_value == 1.7e38
        np.testing.assert_array_equal(v["hice"][:].reshape(-1, 6), k.getDGData("hice", 6))
        np.testing.assert_array_equal(v["cice"][:].reshape(-1, 6), k.getDGData("cice", 6))
        for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
            np.testing.assert_array_equal(v[n][:].reshape(-1, 8), k.getDGData(n, 8))
        np.testing.assert_array_equal(v["u"][:].ravel(), k.getDG0Data("u"))
        np.testing.assert_array_equal(v["u_cg"][:].ravel(), k.get_cg("u"))
        np.testing.assert_array_equal(v["v_cg"][:].ravel(), k.get_cg("v"))
        np.testing.assert_array_equal(v["mask"][:].ravel(), s["mask"])
        np.testing.assert_array_equal(v["coords"][:].reshape(-1, 2), s["coords"])
        if kind == "bbm":
            np.testing.assert_array_equal(v["damage"][:].reshape(-1, 6), k.getDGData("damage", 6))
            np.testing.assert_array_equal(v["avgU_cg"][:].ravel(), k.get_cg("avgU"))
        assert np.abs(v["u_cg"][:]).max() > 1e-4  # a developed state


@pytest.mark.parametrize("kind", ["mevp", "bbm"])
def test_restart_round_trip_continues_bit_for_bit(kind, tmp_path):
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    _say(f"restart round trip {kind}")
    nx = 32
    s, a = _run(kind, nx, 3)
    path = str(tmp_path / "r.nc")
    a.write_restart(path, 360.0)
    b = nd.BBMDynamicsKernel(2, 6) if kind == "bbm" else nd.MEVPDynamicsKernel(2, 6)
    b.initialise(s["coords"], False, s["mask"], nx, nx)
    for name, key in (("uocean", "uocean"), ("vocean", "vocean")):
        b.setData(name, s[key], 1)
    b.setData("ssh", 0 * s["h"], 1)
    assert b.read_restart(path) == 360.0
    for i in range(3, 5):
        for k in (a, b):
            k.benchmark_wind(inputs.L_BOX, 120.0 * i, 0.0, 0.0, 30.0)
            k.update(120.0)
    for n in ("u", "v"):
        np.testing.assert_array_equal(a.get_cg(n), b.get_cg(n))
    for n, c in (("hice", 6), ("cice", 6), ("s11", 8), ("s22", 8), ("e12", 8)):
        np.testing.assert_array_equal(a.getDGData(n, c), b.getDGData(n, c))
    if kind == "bbm":
        np.testing.assert_array_equal(a.getDGData("damage", 6), b.getDGData("damage", 6))


def test_restart_reads_a_file_from_another_writer_and_refuses_mismatches(tmp_path):
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs
    from scipy.io import netcdf_file

    _say("restart from scipy's writer")
    nx = 16
    s = inputs.benchmark_setup(nx, nx, 1)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(s["coords"], False, s["mask"], nx, nx)
    rng = np.random.default_rng(0)
    fields = {"hice": rng.normal(1, 0.1, (nx, nx, 6)), "cice": rng.uniform(0, 1, (nx, nx, 6))}
    for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
        fields[n] = rng.normal(0, 1e3, (nx, nx, 8))
    ucg, vcg = rng.normal(0, 0.1, (2 * nx + 1, 2 * nx + 1)), rng.normal(0, 0.1, (2 * nx + 1, 2 * nx + 1))
    path = str(tmp_path / "other.nc")
    with netcdf_file(path, "w", version=2) as f:
        f.type = b"parametric_rectangular"
        f.time = 7.5
        f.rheology = b"mevp"
        for d, n in (("xdim", nx), ("ydim", nx), ("x_cg", 2 * nx + 1), ("y_cg", 2 * nx + 1), ("dg_comp", 6),
                     ("dgstress_comp", 8)):
            f.createDimension(d, n)
        for n, a in fields.items():
            f.createVariable(n, "d", ("ydim", "xdim", "dg_comp" if a.shape[2] == 6 else "dgstress_comp"))[:] = a
        f.createVariable("u_cg", "d", ("y_cg", "x_cg"))[:] = ucg
        f.createVariable("v_cg", "d", ("y_cg", "x_cg"))[:] = vcg
    assert k.read_restart(path) == 7.5
    np.testing.assert_array_equal(k.getDGData("hice", 6), fields["hice"].reshape(-1, 6))
    np.testing.assert_array_equal(k.getDGData("e22", 8), fields["e22"].reshape(-1, 8))
    np.testing.assert_array_equal(k.get_cg("u"), ucg.ravel())
    # another mesh size, and the other rheology
    other = nd.MEVPDynamicsKernel(2, 6)
    s2 = inputs.benchmark_setup(nx + 2, nx, 1)
    other.initialise(s2["coords"], False, s2["mask"], nx + 2, nx)
    with pytest.raises(nd.NsdgError) as e:
        other.read_restart(path)
    assert e.value.code == -1 and "xdim" in str(e.value)
    bbm = nd.BBMDynamicsKernel(2, 6)
    bbm.initialise(s["coords"], False, s["mask"], nx, nx)
    with pytest.raises(nd.NsdgError) as e:
        bbm.read_restart(path)
    assert e.value.code == -1 and "rheology" in str(e.value)
