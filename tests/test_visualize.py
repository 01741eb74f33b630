"""Demo-frame tiling and writing (examples/visualize.py's role): each env's
image lands in its own grid cell, gaps stay black, the planar layout tiles to
the same picture, and PNG output reads back bit for bit."""
import numpy as np
import pytest
import torch

from minecu import visualize as vz


def _batch(B, H, W, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.randint(1, 256, (B, H, W, 3), generator=g, dtype=torch.int64).to(torch.uint8)


@pytest.mark.parametrize("B,cols", [(5, None), (1, None), (7, 2), (9, 3)])
def test_tile_places_each_env_in_its_cell(B, cols):
    H, W, pad = 4, 6, 2
    obs = _batch(B, H, W)
    img = vz.tile(obs, cols, pad=pad).numpy()
    c = cols or int(np.ceil(np.sqrt(B)))
    r = (B + c - 1) // c
    assert img.shape == (r * (H + pad) - pad, c * (W + pad) - pad, 3)
    covered = np.zeros(img.shape[:2], bool)
    for i in range(B):
        y, x = (i // c) * (H + pad), (i % c) * (W + pad)
        assert np.array_equal(img[y:y + H, x:x + W], obs[i].numpy()), i
        covered[y:y + H, x:x + W] = True
    assert not img[~covered].any()  # gaps and empty cells are black (inputs are >= 1)


def test_planar_and_float_inputs_tile_alike():
    obs = _batch(6, 5, 3, seed=1)
    ref = vz.tile(obs)
    assert torch.equal(vz.tile(obs.permute(0, 3, 1, 2).contiguous(), layout="chw"), ref)
    f = obs.float() / 255.0
    # float images scale by 255 and truncate: k / 255 * 255 rounds to k or k - 1 ulp below
    got = vz.tile(f).int()
    assert (ref.int() - got).abs().max() <= 1
    with pytest.raises(ValueError):
        vz.tile(obs, layout="nhwc")


def test_png_frames_read_back_exactly(tmp_path):
    from PIL import Image
    frames = [vz.tile(_batch(4, 8, 8, seed=s)).numpy() for s in range(3)]
    out = vz.save(frames, str(tmp_path / "frames"), fps=10)
    for i, f in enumerate(frames):
        assert np.array_equal(np.asarray(Image.open(f"{out}/frame_{i:05d}.png").convert("RGB")), f)
    gif = vz.save(frames, str(tmp_path / "demo.gif"), fps=10)
    im = Image.open(gif)
    assert im.n_frames == 3 and im.size == (frames[0].shape[1], frames[0].shape[0])
