import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT / "hvd-cvbs-decoding_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

# 1e-4 of the -40..+100 IRE code range (full scale 140 IRE).
TOL_IRE = 1e-4 * 140.0


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs the C ABI on the GPU)")


@pytest.fixture(scope="session")
def orc64():
    import oracle
    return oracle.Oracle(fp32=False)


@pytest.fixture(scope="session")
def orc32():
    import oracle
    return oracle.Oracle(fp32=True)


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.Reference.available():
        if oracle.REF_SRC.is_dir():
            oracle.build(with_ref=True)
        else:
            pytest.skip("hvd-core reference build (oracle/_ref) not available")
    return oracle.Reference()


@pytest.fixture(scope="session")
def engine():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from hvd_amd import gpu
    eng = gpu.Engine(0)
    yield eng
    eng.close()
