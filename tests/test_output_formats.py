"""On-disk formats (CPU): the VTK ImageData writer, the .npy writer and the
radial density profile of the Sedov post-processing.

The oracle's restatements are pinned here by independent means: an XML
parser and read-back of the .vti values, numpy's own .npy loader, and for the
profile a uniform field (whose mean density is the field value in every bin)
plus brute-force binning.  The product's host-side writers (e2d_write_vti,
e2d_write_npy) need no GPU and must reproduce the oracle byte for byte.
"""
import ctypes as C
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle_lib as O
from euler2d_kokkos_amd import _capi, write_npy, write_vti

VARS = ("rho", "E", "mx", "my")


def small_run(steps=3, **over):
    p = O.default_params(nx=7, ny=5, problemType=0, nStepmax=steps, tEnd=1e9, **over)
    run = O.OracleRun(p)
    run.run()
    return p, run


def interior_vji(a):
    """[i][j][v] array with ghosts -> [v][j][i] interior."""
    return np.ascontiguousarray(np.transpose(O.interior(a), (2, 1, 0)))


def test_oracle_vti_structure(tmp_path):
    p, run = small_run(xmin=0.25, ymin=-1.0, xmax=1.0, ymax=2.0)
    path = O.save_vtk(run, 1, tmp_path, "blob", 3)
    assert path.name == "blob_0000003.vti" and path.exists()
    text = path.read_text()
    assert text.splitlines()[0] == '<?xml version="1.0"?>'
    root = ET.fromstring(text.split("\n", 1)[1])
    assert root.tag == "VTKFile"
    assert root.attrib == {"type": "ImageData", "version": "0.1", "byte_order": "LittleEndian"}
    img = root.find("ImageData")
    assert img.attrib["WholeExtent"] == "0 7 0 5 0 0"
    assert [float(x) for x in img.attrib["Origin"].split()] == [0.25, -1.0, 0.0]
    sp = [float(x) for x in img.attrib["Spacing"].split()]
    assert sp[:2] == pytest.approx([p.dx, p.dy], rel=1e-5) and sp[2] == 0.0
    assert img.find("Piece").attrib["Extent"] == "0 7 0 5 0 0 "
    arrays = img.find("Piece").find("CellData").findall("DataArray")
    assert [a.attrib["Name"] for a in arrays] == list(VARS)
    state = interior_vji(run.get(1))
    for v, a in enumerate(arrays):
        assert a.attrib["type"] == "Float64" and a.attrib["format"] == "ascii"
        vals = np.array([float(x) for x in a.text.split()])
        assert vals.size == 7 * 5
        # i fastest, then j; values at 6 significant digits
        np.testing.assert_allclose(vals, state[v].ravel(), rtol=5e-6, atol=0)


@pytest.mark.parametrize("float32", [False, True])
def test_product_vti_equals_oracle(tmp_path, hip_lib, float32):
    p, run = small_run(steps=4)
    a = O.save_vtk(run, 0, tmp_path, "ora", 4, float32=float32)
    state = interior_vji(run.get(0))
    b = write_vti(p, state, tmp_path, "prod", 4, float32=float32)
    assert a.read_bytes() == b.read_bytes()
    if float32:
        assert b'type="Float32"' in b.read_bytes()


def test_vti_special_values(tmp_path, hip_lib):
    p = O.default_params(nx=3, ny=2)
    state = np.array([0.0, -0.0, 1e-5, 1e20, -123456.789, 2.5] * 4).reshape(4, 2, 3)
    path = write_vti(p, state, tmp_path, "s", 12)
    body = path.read_text().split('format="ascii" >\n')[1].split("\n")[0]
    assert body == "0 -0 1e-05 1e+20 -123457 2.5 "


def test_vti_errors(tmp_path, hip_lib):
    p = O.default_params(nx=3, ny=2)
    state = np.zeros((4, 2, 3))
    lib = _capi.load_library()
    rc = lib.e2d_write_vti(C.byref(p), state.ctypes.data, 0, str(tmp_path / "missing").encode(),
                           b"x", 0)
    assert rc == _capi.E2D_ERR_IO
    rc = lib.e2d_write_vti(C.byref(p), None, 0, str(tmp_path).encode(), b"x", 0)
    assert rc == _capi.E2D_ERR_ARG
    with pytest.raises(ValueError):
        write_vti(p, np.zeros((4, 3, 2)), tmp_path)


@pytest.mark.parametrize("n", [0, 1, 5, 100, 117])
def test_npy_writers(tmp_path, hip_lib, n):
    data = np.random.default_rng(n).standard_normal(n)
    O.write_npy(tmp_path / "ora.npy", data)
    write_npy(tmp_path / "prod.npy", data)
    a = (tmp_path / "ora.npy").read_bytes()
    assert a == (tmp_path / "prod.npy").read_bytes()
    # numpy's own loader; the header (magic + 2-byte length + dict) is padded
    # to a multiple of 16 and ends in a newline
    back = np.load(tmp_path / "prod.npy")
    assert back.dtype == np.float64 and back.shape == (n,)
    np.testing.assert_array_equal(back, data)
    hlen = a[8] | (a[9] << 8)
    assert (10 + hlen) % 16 == 0 and a[10 + hlen - 1:10 + hlen] == b"\n"
    assert a[:8] == b"\x93NUMPY\x01\x00"


def brute_profile(p, state_ijv, nbins):
    """Independent binning: every cell of the array (ghosts included)."""
    i = np.arange(p.isize)[:, None]
    j = np.arange(p.jsize)[None, :]
    x = p.xmin + p.dx / 2 + (i - 2) * p.dx
    y = p.ymin + p.dy / 2 + (j - 2) * p.dy
    cx, cy = (p.xmin + p.xmax) / 2, (p.ymin + p.ymax) / 2
    maxr = np.hypot((p.xmax - p.xmin) / 2, (p.ymax - p.ymin) / 2)
    d = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    b = (d / maxr * nbins).astype(int)
    keep = b < nbins
    cnt = np.bincount(b[keep], minlength=nbins)
    s = np.bincount(b[keep], weights=state_ijv[..., 0][keep], minlength=nbins)
    return cnt, s, maxr


def test_oracle_radial_profile_uniform_field():
    p = O.default_params(nx=40, ny=30, xmax=1.0, ymax=0.75)
    run = O.OracleRun(p)
    a = np.zeros_like(run.get(0))
    a[..., 0] = 0.3125  # uniform density, ghosts included
    run.set(0, a)
    d, m, c = O.radial_profile(run, 0, 25)
    cnt, _, maxr = brute_profile(p, a, 25)
    np.testing.assert_array_equal(c, cnt)
    np.testing.assert_allclose(d, (np.arange(25) + 0.5) * maxr / 25, rtol=1e-15)
    full = c > 0
    np.testing.assert_allclose(m[full], 0.3125, rtol=1e-15)
    assert np.isnan(m[~full]).all()
    # every interior cell lies inside the half-diagonal
    assert c.sum() >= p.nx * p.ny


def test_oracle_radial_profile_brute_force():
    p = O.default_params(nx=33, ny=21, problemType=1, blast_radius=0.2, nStepmax=6, tEnd=1e9)
    run = O.OracleRun(p)
    run.run()
    st = run.get(0)
    d, m, c = O.radial_profile(run, 0, 17)
    cnt, s, _ = brute_profile(p, st, 17)
    np.testing.assert_array_equal(c, cnt)
    np.testing.assert_allclose(m, s / cnt, rtol=1e-13)
