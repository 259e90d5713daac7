"""GPU: each forward solve group's launches replayed from one cached HIP graph.

A graph holds the same kernels with the same arguments as the launch-by-launch
path, so every replay must give the eager result bit for bit -- checked on
several replays in a row, with other eager work and synchronisations between
them, with other batch shapes (other graphs) interleaved, and with the
profiling event nodes in the graph.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from hvd_amd import gpu, synth  # noqa: E402

DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture()
def eng():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    e = gpu.Engine(0)
    yield e
    e.close()


def decode(eng, raws, geoms, p):
    ys, chis, _ = eng.decode_fields_u16([dev(r.view(np.int16)) for r in raws], geoms, p)
    return [(y.cpu().numpy(), c.cpu().numpy()) for y, c in zip(ys, chis)]


def same(a, b):
    return all(np.array_equal(y1, y2) and np.array_equal(c1, c2) for (y1, c1), (y2, c2) in zip(a, b))


def test_graph_replays_match_eager(eng):
    p = gpu.SolverParams.make(cg_iterations=24, irls_outer=3, fast=0, cg_tol=1e-12)
    batch_a = ([synth.ntsc_field(70 + k)[0] for k in range(3)], [synth.NTSC] * 3)
    batch_b = ([synth.pal_field(80)[0], synth.ntsc_field(81)[0]], [synth.PAL, synth.NTSC])
    eng.set_graphs(False)
    want_a = decode(eng, *batch_a, p)
    want_b = decode(eng, *batch_b, p)
    eng.set_graphs(True)
    for rep in range(3):
        got_a = decode(eng, *batch_a, p)
        x = torch.randn(1 << 20, device=DEV)  # unrelated eager work between replays
        (x * 2.0 + 1.0).sum().item()
        torch.cuda.synchronize()
        got_b = decode(eng, *batch_b, p)
        print(f"replay {rep}: batch a {'ok' if same(got_a, want_a) else 'DIFF'}, "
              f"batch b {'ok' if same(got_b, want_b) else 'DIFF'}", flush=True)
        assert same(got_a, want_a) and same(got_b, want_b)


def test_graph_with_early_exits(eng):
    # a relative tolerance stops fields at different iterations inside the graph
    p = gpu.SolverParams.make(cg_iterations=160, irls_outer=4, fast=0, cg_tol=1e-3)
    raws = [synth.ntsc_field(90 + k)[0] for k in range(4)]
    eng.set_graphs(False)
    want = decode(eng, raws, [synth.NTSC] * 4, p)
    eng.set_graphs(True)
    for _ in range(2):
        assert same(decode(eng, raws, [synth.NTSC] * 4, p), want)


def test_graph_with_profiling_events(eng):
    # the profiling graph adds event-record nodes around each stage's CG loop:
    # same results, and the timing record counts every loop of every replay
    p = gpu.SolverParams.make(cg_iterations=40, irls_outer=4, fast=0, cg_tol=1e-12)
    raws = [synth.ntsc_field(95 + k)[0] for k in range(3)]
    eng.set_graphs(False)
    want = decode(eng, raws, [synth.NTSC] * 3, p)
    eng.set_graphs(True)
    eng.set_solver_tuning(8, 0)  # one solve group
    eng.set_profiling(True)
    for _ in range(3):
        assert same(decode(eng, raws, [synth.NTSC] * 3, p), want)
    ms, iters, pxit = eng.cg_timing()
    eng.set_profiling(False)
    print(f"profiled replays: {ms:.3f} ms over {iters} group-iterations", flush=True)
    assert iters == 3 * 4 * 10  # calls x stages x iterations per stage
    assert ms > 0.0 and 0 < pxit <= 3 * 3 * 4 * 10 * 242 * 760  # fields may stop early
