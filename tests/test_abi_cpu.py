"""The C ABI library without a GPU: it loads, exports every entry point that
include/hvd_gpu.h declares, and refuses to create a context when no device
is present (no compute is issued)."""
import ctypes
import pathlib
import re

import pytest

from hvd_amd import gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "hvd_gpu.h").read_text()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(hvd_gpu_\w+)\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(gpu.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not gpu.LIB_PATH.exists():
        pytest.skip("libhvd_gpu.so not built")
    lib = ctypes.CDLL(str(gpu.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.hvd_gpu_abi_version() == 103


def test_create_without_device_fails_cleanly():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lib = gpu.load()
    ctx = ctypes.c_void_p()
    assert lib.hvd_gpu_create(0, ctypes.byref(ctx)) != 0
    assert not ctx
    # entry points reject a null context
    assert lib.hvd_gpu_synchronize(None, None) != 0


def test_struct_layouts():
    # the ctypes mirrors must match the C structs' sizes
    assert ctypes.sizeof(gpu.Geometry) == 8 * 4 + 8 + 8 + 8
    assert ctypes.sizeof(gpu.SolverParams) == 6 * 8 + 2 * 4 + 8 + 2 * 4 + 8
    assert ctypes.sizeof(gpu.Plane) == 8 + 4 * 8 + 8 + 8 + 8 + 8 + 8


CPP_API = [
    "hvd::HvdEngine::HvdEngine()", "hvd::HvdEngine::DecodeFrame(", "hvd::HvdEngine::DecodeChromaOnly(",
    "hvd::HvdEngine::DecodeSequenceWindow(", "hvd::HvdEngine::SetFftThreads(int)",
    "hvd::DecodeFrameBuffer(short const*, hvd::FrameParams const&, hvd::HvdConfig const&, hvd::HvdEngine&)",
    "hvd::DecodeFrameBuffer(short const*, hvd::FrameParams const&, hvd::HvdConfig const&)",
    "hvd::DecodeFrameSequenceWindow(", "hvd::DecodeYcFrameBuffer(", "hvd::SplitFrameFields(",
    "hvd::DetectFieldParity(", "hvd::ResolveField1IsBottom(", "hvd::FieldGeometryFromParams(",
    "hvd::BurstLockinPhase(", "hvd::BurstLockinPhasePal(", "hvd::BurstAmplitudeIre(", "hvd::TridiagSmooth(",
    "hvd::MakeCarrier(", "hvd::MakeCarrierPal(", "hvd::HolographicInit(", "hvd::VariationalRefine(",
    "hvd::EstimateMotion(", "hvd::BoxBlur(", "hvd::AccGain(float, float)", "hvd::PrepareFieldObs(",
    "hvd::DecodeFieldWindow(", "hvd::GpuLastError",
]


def test_cpp_api_library_exports_hvdcore_entry_points():
    """libhvd_core.so (the hvd-core C++ API over the C ABI) defines every
    hvd-core entry point the stage plugin and the engine tests call, and
    links the HIP library by its own rpath."""
    import shutil
    import subprocess
    lib = ROOT / "hvd-cvbs-decoding_amd" / "lib" / "libhvd_core.so"
    if not lib.exists():
        pytest.skip("libhvd_core.so not built")
    if not shutil.which("nm"):
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-DC", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    for sym in CPP_API:
        assert sym in out, sym
    deps = subprocess.run(["ldd", str(lib)], capture_output=True, text=True).stdout
    assert "libhvd_gpu.so" in deps and "not found" not in deps.split("libhvd_gpu.so")[1].split("\n")[0]
