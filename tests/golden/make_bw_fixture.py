"""Writes bw_ini_summary.json from the reference's Brio-Wu initial
configuration (dev/smaug/configs/zero1_BW_asc.ini, a VAC ascii config file),
and the file itself, gzipped, as zero1_BW_asc.ini.gz (the I/O tests read it
through the C reader of the drop-in driver).

Run in the build container only (the reference tree is not present on the
GPU box).  The summary holds data, not code: the header values, the grid
shape, the cell-centre spacing, the index of the left/right interface and
the two distinct state rows (columns h m1 m2 e b1 b2 eb rhob bg1 bg2).
"""
import gzip
import json
import os
import sys

import numpy as np

SRC = "/root/reference/dev/smaug/configs/zero1_BW_asc.ini"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bw_ini_summary.json")


def main():
    lines = open(SRC).read().splitlines()
    it, t, ndim, neqpar, nw = lines[1].split()
    n0, n1 = (int(v) for v in lines[2].split())
    eqpar = [float(v) for v in lines[3].split()]
    names = lines[4].split()
    data = np.array([[float(v) for v in l.split()] for l in lines[5:]])
    assert data.shape == (n0 * n1, 2 + int(nw))
    x = data[:n0, 0]
    y = data[::n0, 1]
    h = data[:n0, 2]
    iface = int(np.where(np.diff(h) != 0)[0][0])
    rows = {"left": data[0, 2:].tolist(), "right": data[n0 - 1, 2:].tolist()}
    out = {
        "source": "dev/smaug/configs/zero1_BW_asc.ini",
        "header": {"it": int(it), "t": float(t), "ndim": int(ndim), "neqpar": int(neqpar),
                   "nw": int(nw)},
        "shape": [n0, n1],
        "eqpar": eqpar,
        "names": names,
        "x_first": float(x[0]), "x_step": float(x[1] - x[0]),
        "y_first": float(y[0]), "y_step": float(y[1] - y[0]),
        "interface_last_left_index": iface,
        "rows": rows,
        "distinct_rows": int(len(np.unique(data[:, 2:], axis=0))),
        "y_uniform": bool(np.all(data[:, 2:].reshape(n1, n0, -1) == data[:n0, 2:][None])),
    }
    json.dump(out, open(OUT, "w"), indent=1)
    print("wrote", OUT)
    gz = os.path.join(os.path.dirname(OUT), "zero1_BW_asc.ini.gz")
    with open(SRC, "rb") as fi, gzip.GzipFile(gz, "wb", compresslevel=9, mtime=0) as fo:
        fo.write(fi.read())
    print("wrote", gz)


if __name__ == "__main__":
    sys.exit(main())
