"""slgpu_opacity_build_noalloc (the opacity build whose per-call staging lives in the
library's static device scratch array) against slgpu_opacity_build, through the C ABI:
identical outputs on two calls in a row, the second on a smaller input, and the size
limit rejected with SLGPU_EINVAL.  Marked gpu."""
import sys

import numpy as np
import pytest

from superlite_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(ctx, g):
    args = (g["temp"], g["natom"], g["vol"], g["mass"], g["nelec"], g["natomfr"], g["oc"], g["ginv"])
    a = ctx.opacity_build(*args)
    b = ctx.opacity_build(*args, noalloc=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    return a


def test_opacity_build_noalloc_matches(tables, golden):
    from superlite_amd import slgpu as S
    say("case: opacity_build_noalloc, c1 (50 cells, 18 ions) then the 3-cell edge case")
    w, st, wl, lines = synth.build("c1")
    ctx = S.setup_workload(w, st, wl, lines, tables, npartmax=1024)
    a = _both(ctx, st)
    assert np.count_nonzero(a["cap64"]) > 0.9 * a["cap64"].size
    ctx.close()
    # second call: a smaller input on the same GPU (the scratch array still holds c1's data)
    e = golden("ref_edge_opacity.npz")
    ctx = S.Context(3, 5, 2, 2, npartmax=16)
    ctx.set_grid(e["xarr"], e["vxarr"], e["divv"], e["vol"], 0.5 * (e["xarr"][1:] + e["xarr"][:-1]))
    ctx.set_groups(e["wl"])
    ctx.set_ions(e["ion_iz"], e["ion_ii"])
    ctx.set_tables(*tables)
    ctx.set_lines(e["wl0"], e["gxs"], e["chilw"], e["liz"], e["lii"])
    a = _both(ctx, e)
    assert np.count_nonzero(a["cap64"]) > 0
    ctx.close()


def test_opacity_build_noalloc_limit(tables):
    """2*nion*nc doubles beyond the 4 MiB scratch array: SLGPU_EINVAL, nothing launched"""
    from superlite_amd import slgpu as S
    say("case: opacity_build_noalloc size limit")
    nx, ng = 600, 2
    iz = np.array([z for z in range(1, 31) for _ in range(z)], np.int32)        # 465 ions
    ii = np.array([k for z in range(1, 31) for k in range(1, z + 1)], np.int32)
    assert len(iz) * nx > 262144
    ctx = S.Context(nx, ng, 1, 2, npartmax=16)
    x = np.linspace(0.0, 1e15, nx + 1)
    ctx.set_grid(x, x / 1.728e6, np.full(nx, 3 / 1.728e6), np.diff(x ** 3) * 4.19, 0.5 * (x[1:] + x[:-1]))
    ctx.set_groups(np.array([1000e-8, 2000e-8, 3000e-8]))
    ctx.set_ions(iz, ii)
    ctx.set_tables(*tables)
    ctx.set_lines(np.array([1500.0], np.float32), np.array([0.1], np.float32),
                  np.array([0.0], np.float32), np.array([26], np.int16), np.array([1], np.int16))
    one = np.ones(nx)
    fr = np.zeros((30, nx))
    fr[25] = 1.0
    pops = np.ones((nx, len(iz)))
    with pytest.raises(S.SlgpuError) as ei:
        ctx.opacity_build(5000.0 * one, one, one, one, one, fr, pops, pops, noalloc=True)
    assert ei.value.code == S.SLGPU_EINVAL_CODE
    ctx.close()
