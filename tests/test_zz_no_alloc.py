"""GPU test of ent_combined_densities_dev_noalloc (the density entry point
that allocates nothing; its workspace is the device's fixed scratch array):
through the C ABI, two calls in a row, the second with a smaller input,
each bit-identical to ent_combined_densities_dev on the same inputs and
within 1e-5 of FP64, plus the size limit."""
import numpy as np
import pytest

from entropical_amd import Context, DeviceArray, EntropicalError, combined_densities_dev

pytestmark = pytest.mark.gpu

MAX_POINTS = 1_000_000


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def run(ctx, dx, n, cols, h, tau, out, noalloc, method="cell"):
    combined_densities_dev(ctx, dx.ptr, n, dx.shape[0], cols, h, out, tau=tau, method=method, noalloc=noalloc)
    m = n - (max(tau) if tau else 0)
    got = np.empty(m)
    ctx.d2h(got, out)
    return got


def test_noalloc_matches_allocating_entry_point(orc):
    r = np.random.default_rng(301)
    with Context(0) as ctx:
        # first call: 2-D, 120k points with a lag; second: 3-D, 9k points;
        # third: 1-D brute force, 5k points
        cases = [(120_000, [0, 1], [0.2, 0.25], [0, 3], "cell"),
                 (9_000, [0, 1, 2], [0.4, 0.35, 0.5], None, "cell"),
                 (5_000, [2], [0.1], None, "brute")]
        for n, cols, h, tau, method in cases:
            a = r.standard_normal((3, n))
            a[1] += 0.6 * a[0]
            x = a.astype(np.float32)
            dx = DeviceArray(ctx, x)
            out = ctx.alloc(8 * n)
            try:
                ref_alloc = run(ctx, dx, n, cols, h, tau, out, False, method)
                got = run(ctx, dx, n, cols, h, tau, out, True, method)
            finally:
                ctx.free(out)
                dx.free()
            np.testing.assert_array_equal(got, ref_alloc)
            ref = orc.density_fp64(x, n, cols, h, tau) if n <= 10_000 else orc.density_fp64_window(x, n, cols, h, tau)
            assert rel_err(got, ref) < 1e-5, (n, cols)


def test_noalloc_rejects_sizes_beyond_the_maximum():
    with Context(0) as ctx:
        n = MAX_POINTS + 1
        dx = DeviceArray(ctx, np.random.default_rng(302).standard_normal((1, n)).astype(np.float32))
        out = ctx.alloc(8 * n)
        try:
            with pytest.raises(EntropicalError, match="exceed the maximum"):
                combined_densities_dev(ctx, dx.ptr, n, 1, [0], [0.1], out, noalloc=True)
            # a lag that brings the points within the limit is accepted
            combined_densities_dev(ctx, dx.ptr, n, 1, [0], [0.1], out, tau=[1], noalloc=True)
        finally:
            ctx.free(out)
            dx.free()
