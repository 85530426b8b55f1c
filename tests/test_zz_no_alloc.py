"""smaug_sums_noalloc (the conservation sums without any allocation) against
smaug_sums on the same states, through the C-ABI: three calls in a row on one
device, each on a smaller grid than the one before, so the second and third
run over partials the previous call left in the shared static array.  The
device-side merge repeats the host merge's operations in the same order, so
the sums agree bit for bit."""
import numpy as np
import pytest
import torch

from helpers import perturbed

pytestmark = pytest.mark.gpu

from smaugamd import configs  # noqa: E402
from smaugamd import engine as eng  # noqa: E402
from smaugamd.engine import FP64, MIXED, Engine  # noqa: E402


def _cases():
    big = configs.stratified(nz=28, nx=16, ny=20, lz=6e6 * 28 / 256 * 4,
                             lx=4e6 * 16 / 256 * 4, ly=4e6 * 20 / 256 * 4)
    mid = configs.stratified(nz=12, nx=10, ny=9, lz=6e6 * 12 / 256 * 4,
                             lx=4e6 * 10 / 256 * 4, ly=4e6 * 9 / 256 * 4)
    small = configs.orszag_tang(nx=16, ny=12)
    return [("3-D 32x20x24 mixed", big, MIXED, 3), ("3-D 16x14x13 fp64", mid, FP64, 5),
            ("2-D 20x16 mixed", small, MIXED, 7)]


def test_noalloc_sums_match_smaug_sums_on_shrinking_grids():
    dev = torch.device("cuda:0")
    side = torch.cuda.Stream(device=dev)
    sizes = []
    for label, s, prec, seed in _cases():
        print(f"[case] sums_noalloc {label}", flush=True)
        e = Engine(s, prec)
        e.set_state(perturbed(s, amp=1e-2, seed=seed))
        ref = e.sums()
        stored = e.get_state()  # the engine's stored values, as FP64
        e.close()
        dt = torch.float64 if prec == FP64 else torch.float32
        d_w = torch.tensor(stored, dtype=dt, device=dev).contiguous()
        d_out = torch.full((s.nve,), float("nan"), dtype=torch.float64, device=dev)
        n = list(s.n) + [1] * (3 - len(s.n))
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            rc = eng.sums_noalloc(s.ndim, n, prec, d_w.data_ptr(), d_out.data_ptr(),
                                  side.cuda_stream)
        assert rc == 0, eng.last_error()
        side.synchronize()
        got = d_out.cpu().numpy()
        np.testing.assert_array_equal(got, ref)
        sizes.append(s.ncell)
    assert sizes[0] > sizes[1] > sizes[2]
