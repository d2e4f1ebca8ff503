"""The reference's own heightmap asset (assets/textures/noise/waves/waves3/00.png, loaded
at src/program.cpp:719 through RenderEngine::load2DTexture, src/render-engine.cpp:
1308-1322) decoded by the host loader, against the committed fixture of its texels.

The fixture (tests/golden/waves3_00_r8.npz, written by make_golden.py) holds the R8
texels the reference's texture upload receives: stb_image forcing RGBA, rows flipped so
row 0 is the image's bottom, red channel. Where the reference tree is present the C++
loader decodes the PNG itself and must give the same bytes; elsewhere the fixture's own
integrity is checked."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CHECK = os.environ.get("WT_STAGE_CHECK") or os.path.join(ROOT, "wave-tool_amd", "host", "water_stage_check")
ASSET = "/root/reference/assets/textures/noise/waves/waves3/00.png"


def fixture():
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))["waves3_00_r8"]
    tex = np.load(os.path.join(GOLDEN, "waves3_00_r8.npz"))["texels"]
    return tex, meta


def test_fixture_integrity():
    tex, meta = fixture()
    assert tex.dtype == np.uint8 and tex.shape == (meta["height"], meta["width"]) == (512, 512)
    assert hashlib.sha256(tex.tobytes()).hexdigest() == meta["texels_sha256"]
    # a grey noise image centred on mid-grey (bump 2 (r - 0.5) averages about zero)
    assert tex.min() == 45 and tex.max() == 207 and abs(tex.mean() - 127.5) < 0.5 and tex.std() > 20


@pytest.mark.skipif(not os.path.exists(ASSET), reason="the reference tree is not on this machine")
def test_host_loader_decodes_the_reference_asset(tmp_path):
    if not os.path.exists(CHECK):
        pytest.fail("water_stage_check is not built (run __graft_entry__.build())")
    tex, meta = fixture()
    assert hashlib.sha256(open(ASSET, "rb").read()).hexdigest() == meta["png_sha256"]
    out = tmp_path / "waves3.bin"
    subprocess.run([CHECK, "--png", ASSET, str(out)], check=True)
    raw = out.read_bytes()
    W, H = np.frombuffer(raw[:8], dtype=np.uint32)
    got = np.frombuffer(raw[8:], dtype=np.uint8).reshape(H, W)
    assert (W, H) == (512, 512)
    assert np.array_equal(got, tex)


@pytest.mark.skipif(not os.path.exists(ASSET), reason="the reference tree is not on this machine")
def test_fixture_matches_an_independent_decode():
    """PIL's decode with the stb conventions (RGBA expansion of grey, vertical flip)
    gives the fixture's bytes; the flip is visible (row 0 is the image's last row)."""
    from PIL import Image
    tex, _ = fixture()
    img = np.asarray(Image.open(ASSET).convert("RGBA"))
    assert (img[:, :, 0] == img[:, :, 1]).all() and (img[:, :, 3] == 255).all()
    assert np.array_equal(tex, img[::-1, :, 0])
    assert not np.array_equal(tex, img[:, :, 0])
