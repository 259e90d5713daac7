"""Write golden vectors from hvd-core's own engine code (oracle/_ref, compiled
from the reference checkout) so the oracle stays pinned where the reference
is absent.  Run in a container that has the reference:  python
tests/golden/make_golden.py.  Inputs come from the seeded generator; only
inputs and hvd-core's outputs are stored.
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "hvd-cvbs-decoding_amd"))

import oracle as O  # noqa: E402
from hvd_amd import synth  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent


def burst_field(g, burst_raw):
    """IRE field that is zero except for the stored burst window."""
    f = np.zeros((g.field_height, g.field_width), np.float32)
    f[:, g.colour_burst_start:g.colour_burst_end] = synth.ire_from_raw(burst_raw)
    return f


def main():
    O.build(with_ref=True)
    ref = O.Reference()
    orc64 = O.Oracle(False)
    out = {}
    # lock-in: full NTSC and PAL fields (seeded)
    # only the burst window is stored: samples outside it are zero, which the
    # lock-in and amplitude estimators never read
    fields = {}
    for name, g, fn, seed in (("ntsc", synth.NTSC, synth.ntsc_field, 101), ("pal", synth.PAL, synth.pal_field, 102)):
        raw = fn(seed)[0]
        burst = raw[:, g.colour_burst_start:g.colour_burst_end].copy()
        out[f"{name}_burst_raw"] = burst
        f = burst_field(g, burst)
        fields[name] = (raw, f)
        if g.uses_vswitch:
            th, vs = ref.burst_lockin_phase_pal(f, g)
            out[f"{name}_vswitch"] = vs
        else:
            th = ref.burst_lockin_phase(f, g)
        out[f"{name}_theta"] = th
        out[f"{name}_burst_amp"] = np.float32(ref.burst_amplitude_ire(f, g))
    # refine: a 40 x 96 crop of the NTSC field, AUTO aniso, 4 x 10 CG
    g = synth.NTSC
    f = synth.ire_from_raw(fields["ntsc"][0])
    th = ref.burst_lockin_phase(f, g)
    y0, x0 = 60, 300
    s = f[g.first_active_field_line + y0:g.first_active_field_line + y0 + 40,
          g.active_video_start + x0:g.active_video_start + x0 + 96].copy()
    gc = synth.Geometry(g.field_width, g.field_height, g.active_video_start + x0, g.active_video_start + x0 + 96,
                        g.colour_burst_start, g.colour_burst_end, 0, 0, g.sample_rate, 0.0, 0)
    car = ref.make_carrier(th[g.first_active_field_line + y0:g.first_active_field_line + y0 + 40], None, gc)
    _, chi0 = orc64.holographic_init(s, car, g.sample_rate)
    chi0 = chi0.astype(np.complex64)
    cfg = O.Cfg.make(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-12)
    yr, cr = ref.variational_refine(s, car, chi0, cfg)
    out.update(refine_s=s, refine_carrier=car.astype(np.complex64), refine_chi0=chi0, refine_y=yr,
               refine_chi=cr.astype(np.complex64))
    # motion: the 128 x 160 textured scene shifted by (2, -3)
    sc = synth.make_scene(synth.NTSC, np.random.default_rng(104), canvas_pad=16)
    Y = sc.y.astype(np.float32)
    A = Y[16:144, 16:176].copy()
    B = Y[14:142, 19:179].copy()
    dy, dx, cf = ref.estimate_motion(B, A, 32, 8)
    out.update(motion_ref=B, motion_cur=A, motion_dy=dy, motion_dx=dx, motion_conf=cf)
    np.savez_compressed(OUT / "hvdcore_vectors.npz", **out)
    print("wrote", OUT / "hvdcore_vectors.npz", sum(v.nbytes for v in out.values() if hasattr(v, "nbytes")), "bytes")


if __name__ == "__main__":
    main()
