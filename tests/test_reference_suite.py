"""hvd-core's own engine test programs, compiled against this package's
hvd-core C++ API (libhvd_core.so) instead of hvd-core, run on the GPU.

The programs are built by oracle/Makefile's `reftests` target from the
reference checkout (hvd-core/tests/engine/*.cpp with its check.h harness);
each exits non-zero on its first failed CHECK.  Where they were not built
(no reference checkout at build time) the tests skip.
"""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
BIN = ROOT / "oracle" / "_ref" / "tests"
NAMES = ["lockin_test", "pal_lockin_test", "gain_chain_test", "chroma_reference_test", "field_order_test",
         "frame_bridge_test", "motion_test", "gradients_test", "lossless_identity_test", "temporal_test",
         "sequence_test", "colour_test", "variational_v2_test", "motion_v2_test", "pal_aniso_group_test",
         "preview_aspect_test"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hvdcore_engine_test_passes_on_gpu_library(name):
    exe = BIN / name
    if not exe.exists():
        pytest.skip(f"{exe} not built (needs the reference checkout at build time)")
    print(f"running {name}", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{name} rc={r.returncode}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
