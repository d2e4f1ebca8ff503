"""GPU test of the device index buffer (wt_grid_triangle_indices) against the
reference's index order (src/program.cpp:686-717)."""
import numpy as np
import pytest

import wavetool_amd as wt
from wavetool_amd import scenes
from helpers import reference_indices, reference_indices_vectorized

pytestmark = pytest.mark.gpu


def say(msg):
    print("[case] " + msg, flush=True)


def device_indices(torch, ctx, n, base=0):
    out = torch.full((6 * (n - 1) ** 2,), -1, dtype=torch.int32, device="cuda")
    ctx.grid_indices(n, out, base=base)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32).astype(np.int64)


@pytest.mark.parametrize("n,base", [(2, 0), (3, 0), (17, 0), (17, 1000), (300, 7), (1024, 0), (513, 4_000_000_000 - 513 * 513)])
def test_grid_indices_match_reference(torch_cuda, wtctx, n, base):
    say("indices n=%d base=%d" % (n, base))
    got = device_indices(torch_cuda, wtctx, n, base)
    ref = reference_indices(n) if n <= 17 else reference_indices_vectorized(n)
    assert np.array_equal(got, ref + base)


def test_grid_indices_for_batch_patches(torch_cuda, wtctx):
    """One index list per patch of a batch, each offset by the patch's first vertex, in
    one buffer: every triangle stays inside its own patch's vertex range."""
    torch = torch_cuda
    specs = scenes.batched_patches(seed=5, count=6)
    for s in specs:
        s["N"] = max(2, s["N"] // 8)
    patches, _, V = wt.batch_inputs(specs)
    counts = [6 * (p.gridLength - 1) ** 2 for p in patches]
    buf = torch.full((sum(counts),), -1, dtype=torch.int32, device="cuda")
    off = 0
    for p, cnt in zip(patches, counts):
        wtctx.grid_indices(p.gridLength, buf[off:off + cnt], base=p.vertex_offset)
        off += cnt
    torch.cuda.synchronize()
    got = buf.cpu().numpy().view(np.uint32).astype(np.int64)
    off = 0
    for p, cnt in zip(patches, counts):
        seg = got[off:off + cnt]
        assert np.array_equal(seg, reference_indices_vectorized(p.gridLength) + p.vertex_offset)
        assert seg.min() >= p.vertex_offset and seg.max() < p.vertex_offset + p.gridLength ** 2 <= V
        off += cnt


def test_grid_indices_reject_bad_arguments(torch_cuda, wtctx):
    torch = torch_cuda
    out = torch.zeros(6, dtype=torch.int32, device="cuda")
    for n, base in ((1, 0), (2, 2 ** 32 - 3)):
        with pytest.raises((wt.WtError, ValueError)):
            wtctx.grid_indices(n, out if n == 2 else torch.zeros(0, dtype=torch.int32, device="cuda"), base=base)
    with pytest.raises(wt.WtError) as e:
        wtctx.grid_indices(2, out.cpu())  # host memory
    assert e.value.status == wt.WT_ERR_INVALID_ARG
