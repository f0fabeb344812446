"""Launch grids above the default workgroup cap (``ftmi_set_grid_cap``; the autotuner raises it
up to 8192): kernels that write one value per workgroup into a caller's buffer stay within the
buffer's documented size (2048 scan block sums, 2048 * 8 reduction partials) whatever the cap,
and the cap changes launch shapes only, never a value. Both buffers are given a sentinel tail
here, so a write past the documented size lands in memory the test owns and fails the check."""
import numpy as np
import pytest
import torch

from gputest_amd import _native as nat
from gputest_amd import case as C
from gputest_amd.mesh import duct_mesh
from gputest_amd.simple import SimpleSolver

pytestmark = pytest.mark.gpu
DEV = "cuda"
HIGH_CAP = 8192


class _Cap:
    def __init__(self, cap):
        self.cap = cap

    def __enter__(self):
        self.was = nat.lib().ftmi_set_grid_cap(self.cap)
        return self

    def __exit__(self, *exc):
        nat.lib().ftmi_set_grid_cap(self.was)


def test_scan_block_sums_stay_within_2048_at_a_high_cap():
    rng = np.random.default_rng(0)
    n = 3_000_000  # 11719 workgroups of 256 rows without a cap
    cnt = rng.integers(0, 40, n).astype(np.int32)
    d_in = torch.as_tensor(cnt, device=DEV)
    out = torch.empty(n + 1, dtype=torch.int32, device=DEV)
    bsum = torch.full((HIGH_CAP,), -7, dtype=torch.int64, device=DEV)
    with _Cap(HIGH_CAP):
        nat.check(nat.lib().ftmi_scan_i32(n, nat.ptr(d_in), nat.ptr(out), nat.ptr(bsum), nat.stream_ptr()), "scan")
    torch.cuda.synchronize()
    ref = np.zeros(n + 1, np.int64)
    ref[1:] = np.cumsum(cnt)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert bool((bsum[2048:] == -7).all())


def _duct(precision):
    def parabolic(x):
        return np.stack([4.0 * x[:, 1] * (1.0 - x[:, 1]), np.zeros(x.shape[0])], axis=1)

    mu = 1.0 / 500.0
    B = C.BoundarySpec
    case = C.UnstructuredCase(
        boundaries={"inlet": B("inlet", velocity=[1.0, 0.0], value=300.0, profile=parabolic),
                    "outlet": B("outlet", value=300.0),
                    "bottom": B("wall", thermal="flux", value=5000.0),
                    "top": B("wall", thermal="flux", value=5000.0)},
        viscosity=mu, energy=True, conductivity=mu / 0.7, cp=1.0, init_T=300.0, pressure_tol=0.1,
        max_iter=10 ** 9, tol=0.0, precision=precision)
    return SimpleSolver(duct_mesh(1200, 480), case, DEV)  # 576,000 cells: 2250 workgroups of 256


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solver_steps_at_a_high_cap_are_identical_and_in_bounds(precision):
    ref = _duct(precision)
    for _ in range(3):
        ref.step()
    s = _duct(precision)
    big = torch.full((HIGH_CAP * 8,), float("nan"), dtype=torch.float64, device=DEV)
    s.red_partials = big
    s.c_red.partials = nat.ptr(big)
    with _Cap(HIGH_CAP):
        for _ in range(3):
            s.step()
        torch.cuda.synchronize()
    print(precision, s.last_linear, ref.last_linear)
    assert bool(torch.isnan(big[2048 * 8:]).all()), "reduction partials written past 2048 * 8"
    for name in ("U", "p", "F", "Fb", "T"):
        a, b = getattr(s, name), getattr(ref, name)
        assert torch.equal(a, b), (name, float((a - b).abs().max()))
