"""VAC ascii configuration files: round trip and the reference fixture's header layout."""
import json
import os

import numpy as np

from smaugamd import configs, vacio

HERE = os.path.dirname(os.path.abspath(__file__))


def test_roundtrip_2d_and_3d(tmp_path):
    for s in (configs.orszag_tang(nx=12, ny=10),
              configs.stratified(nz=8, nx=6, ny=7, lz=1e6, lx=1e6, ly=1e6)):
        path = str(tmp_path / f"{s.name}.ini")
        vacio.write_ascii(path, s, it=3, t=0.25)
        hdr, pos, data = vacio.read_ascii(path)
        assert hdr["ndim"] == s.ndim and hdr["nw"] == s.nve + s.nbg
        assert hdr["dims"] == list(s.n[: s.ndim]) and hdr["it"] == 3 and hdr["t"] == 0.25
        assert hdr["eqpar"][0] == s.gamma
        np.testing.assert_allclose(data[: s.nve], s.w, rtol=1e-15, atol=1e-300)
        np.testing.assert_allclose(data[s.nve:], s.bg, rtol=1e-15, atol=1e-300)
        assert pos.shape == (s.ndim, s.ncell)


def test_header_layout_matches_reference_fixture(tmp_path):
    fx = json.load(open(os.path.join(HERE, "golden", "bw_ini_summary.json")))
    s = configs.brio_wu(ncells=fx["shape"][0] - 4, ny=fx["shape"][1] - 4)
    path = str(tmp_path / "bw.ini")
    vacio.write_ascii(path, s)
    hdr, pos, data = vacio.read_ascii(path)
    assert hdr["ndim"] == fx["header"]["ndim"]
    assert hdr["neqpar"] == fx["header"]["neqpar"]
    assert hdr["nw"] == fx["header"]["nw"]
    assert hdr["dims"] == fx["shape"]
    assert hdr["names"][: len(fx["names"])] == fx["names"]
    # eqpar holds gamma first (the reference's files carry other values in
    # the later slots, which its driver does not read)
    assert len(hdr["eqpar"]) == len(fx["eqpar"]) and hdr["eqpar"][0] == fx["eqpar"][0]


# ---------------------------------------------------------------- the C readers/writers
import gzip  # noqa: E402
import shutil  # noqa: E402
import struct  # noqa: E402
import subprocess  # noqa: E402

ROOT = os.path.dirname(HERE)
HOST = os.path.join(ROOT, "smaug-all_amd", "host")


def _harness(tmp_path, ndim):
    exe = tmp_path / f"vacio_harness{ndim}d"
    flag = "-DUSE_SAC_3D" if ndim == 3 else "-DUSE_SAC"
    subprocess.check_call(["gcc", "-O1", "-std=c11", flag, "-I", os.path.join(ROOT, "include"),
                           "-I", HOST, "-o", str(exe), os.path.join(HERE, "vacio_harness.c"),
                           os.path.join(HOST, "vacio.c")])
    return str(exe)


def _run_harness(exe, ini, base, dims, nvar, ndim):
    r = subprocess.run([exe, ini, base] + [str(d) for d in dims], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "binary roundtrip identical" in r.stdout
    nc = int(np.prod(dims))
    raw = open(base + ".raw", "rb").read()
    it, t = struct.unpack_from("<id", raw, 0)
    vals = np.frombuffer(raw, dtype="<f8", offset=12)
    assert vals.size == (nvar + ndim) * nc
    return it, t, vals[: nvar * nc].reshape(nvar, nc), vals[nvar * nc:].reshape(ndim, nc)


def _binary_layout(path, ndim, nvar, nc):
    """the reference's single-domain binary layout (readwrite.c:160-368)"""
    b = open(path, "rb").read()
    off = 79
    it, t = struct.unpack_from("<id", b, off)
    off += 12
    nd, neq, nw = struct.unpack_from("<3i", b, off)
    off += 12
    dims = struct.unpack_from(f"<{nd}i", b, off)
    off += 4 * nd
    eq = struct.unpack_from(f"<{neq}d", b, off)
    off += 8 * neq
    names = b[off:off + 79].rstrip(b"\0").decode()
    off += 79
    payload = np.frombuffer(b, dtype="<f8", count=(ndim + nvar) * nc, offset=off)
    assert len(b) == off + 8 * (ndim + nvar) * nc + 1 and b[-1:] == b"\n"
    return it, t, (nd, neq, nw), list(dims), eq, names, payload


def test_reference_bw_ini_through_c_reader_every_cell(tmp_path):
    """the reference's own Brio-Wu initial file (configs/zero1_BW_asc.ini,
    committed gzipped) through readasciivacconfig: every cell's position and
    state, the binary writer/reader round trip in the reference's layout, and
    the ascii writer read back exactly"""
    fx = json.load(open(os.path.join(HERE, "golden", "bw_ini_summary.json")))
    ini = str(tmp_path / "zero1_BW_asc.ini")
    with gzip.open(os.path.join(HERE, "golden", "zero1_BW_asc.ini.gz"), "rb") as fi, \
            open(ini, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    n0, n1 = fx["shape"]
    base = str(tmp_path / "bw")
    it, t, w, pos = _run_harness(_harness(tmp_path, 2), ini, base, (n0, n1), 10, 2)
    assert (it, t) == (fx["header"]["it"], fx["header"]["t"])
    i = np.tile(np.arange(n0), n1)
    j = np.repeat(np.arange(n1), n0)
    # positions: cell centres of a uniform grid (printed to 11 digits)
    np.testing.assert_allclose(pos[0], fx["x_first"] + i * fx["x_step"], rtol=1e-10)
    np.testing.assert_allclose(pos[1], fx["y_first"] + j * fx["y_step"], rtol=1e-9)
    # state: the left row up to the interface, the right row after it, in
    # every cell (ghost rows and columns included)
    left = i <= fx["interface_last_left_index"]
    for f in range(10):
        expect = np.where(left, fx["rows"]["left"][f], fx["rows"]["right"][f])
        assert np.array_equal(w[f], expect), f
    # the Python reader sees the same values
    hdr, ppos, data = vacio.read_ascii(ini)
    assert np.array_equal(data, w) and np.array_equal(ppos, pos)
    # binary file in the reference's layout
    bit, bt, counts, dims, eq, names, payload = _binary_layout(base + "_0.out", 2, 10, n0 * n1)
    assert (bit, bt, counts, dims) == (0, 0.0, (2, 6, 10), [n0, n1])
    assert eq == (2.0, 0.0, 1.0, 0.0, 0.0, 0.0)  # gamma eta g1 g2 and two zeros
    assert names == "x y rho mx my e bx by  gamma eta g1 g2"
    assert np.array_equal(payload[: 2 * n0 * n1].reshape(2, -1), pos)
    assert np.array_equal(payload[2 * n0 * n1:].reshape(10, -1), w)
    # ascii writer: 17 significant digits read back exactly
    hdr2, pos2, data2 = vacio.read_ascii(base + ".ini")
    assert hdr2["dims"] == [n0, n1] and hdr2["eqpar"][0] == 2.0
    assert np.array_equal(data2, w) and np.array_equal(pos2, pos)


def test_c_writers_roundtrip_3d(tmp_path):
    s = configs.stratified(nz=8, nx=6, ny=7, lz=1e6, lx=1e6, ly=1e6)
    rng = np.random.default_rng(4)
    s.w = s.w + rng.standard_normal(s.w.shape) * 1e-3
    ini = str(tmp_path / "in3.ini")
    vacio.write_ascii(ini, s, it=12, t=0.75)
    base = str(tmp_path / "s3")
    it, t, w, pos = _run_harness(_harness(tmp_path, 3), ini, base, s.n, 13, 3)
    assert (it, t) == (12, 0.75)
    assert np.array_equal(w, np.vstack([s.w, s.bg]))
    bit, bt, counts, dims, eq, names, payload = _binary_layout(base + "_12.out", 3, 13, s.ncell)
    assert (bit, bt, counts, dims) == (12, 0.75, (3, 7, 13), list(s.n))
    assert eq[0] == s.gamma and eq[2] == s.g[0]
    assert np.array_equal(payload[3 * s.ncell:].reshape(13, -1), w)
    hdr2, pos2, data2 = vacio.read_ascii(base + ".ini")
    assert np.array_equal(data2, w) and np.array_equal(pos2, pos)
