"""GPU parity of the stage-side export (hvd_gpu_woven_picture): the woven
active picture and its RGB24 rendering (hvd_chroma_decoder_stage.cpp:
316-413, :79-88, :500-638) against the oracle restatement in
oracle/bridge.py, which tests/test_export_oracle.py pins.  Index work and
the double-precision conversion are held bit-exact."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402
import bridge as OB  # noqa: E402

DEV = "cuda:0"


def fparams(fpd):
    fp = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    for i, e in enumerate(fpd["extra_sample_lines"]):
        fp.extra_sample_lines[i] = e
    fp.extra_sample_line_count = len(fpd["extra_sample_lines"])
    return fp


@pytest.mark.parametrize("std", ["ntsc", "pal"])
@pytest.mark.parametrize("full_raster,field", [(False, -1), (True, -1), (False, 0), (False, 1), (True, 1)])
def test_woven_picture_matches_oracle(engine, std, full_raster, field):
    fpd = synth.FRAME_PAL if std == "pal" else synth.FRAME_NTSC
    fh, fw = fpd["frame_height"], fpd["frame_width"]
    rng = np.random.default_rng(7)
    luma = rng.integers(0, 1024, (fh, fw)).astype(np.int16)
    u = rng.normal(0.0, 150.0, (fh, fw))
    v = rng.normal(0.0, 150.0, (fh, fw))
    wy, wu, wv = OB.reorder_to_woven(luma, u, v, OB.FrameParams(**fpd), full_raster, field)
    wrgb = OB.woven_rgb24(wy, wu, wv, OB.FrameParams(**fpd))
    t = [torch.from_numpy(a).to(DEV) for a in (luma, u, v)]
    y, uo, vo, rgb = engine.woven_picture(*t, fparams(fpd), full_raster, field)
    assert y.shape == wy.shape and rgb.shape == wrgb.shape
    assert np.array_equal(y.cpu().numpy(), wy)
    assert np.array_equal(uo.cpu().numpy(), wu) and np.array_equal(vo.cpu().numpy(), wv)
    assert np.array_equal(rgb.cpu().numpy(), wrgb)


def test_woven_picture_of_a_decoded_window(engine):
    # the chain the stage runs: decode a frame window, weave it for export
    fpd = synth.FRAME_NTSC
    frames = synth.frames_s16(fpd, 2, seed=9)
    fp = fparams(fpd)
    p = gpu.SolverParams.make(cg_iterations=8, irls_outer=2, fast=0)
    outs, _ = engine.decode_frame_window_s16([torch.from_numpy(f).to(DEV) for f in frames], 0, 2, fp, p,
                                             gpu.BridgeParams.make())
    luma, _, u, v = outs[0]
    y, uo, vo, rgb = engine.woven_picture(luma, u, v, fp)
    wy, wu, wv = OB.reorder_to_woven(luma.cpu().numpy(), u.cpu().numpy(), v.cpu().numpy(), OB.FrameParams(**fpd))
    assert np.array_equal(y.cpu().numpy(), wy)
    assert np.array_equal(rgb.cpu().numpy(), OB.woven_rgb24(wy, wu, wv, OB.FrameParams(**fpd)))
    assert rgb.shape == (483, 760, 3)


def test_woven_picture_rejects_bad_levels(engine):
    fpd = dict(synth.FRAME_NTSC, white_level=200.0)
    fh, fw = fpd["frame_height"], fpd["frame_width"]
    z16 = torch.zeros((fh, fw), dtype=torch.int16, device=DEV)
    z64 = torch.zeros((fh, fw), dtype=torch.float64, device=DEV)
    with pytest.raises(gpu.HvdGpuError, match="white level"):
        engine.woven_picture(z16, z64, z64, fparams(fpd))
