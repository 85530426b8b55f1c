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
