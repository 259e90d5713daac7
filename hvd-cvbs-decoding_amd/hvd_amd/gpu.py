"""ctypes binding of the C ABI in include/hvd_gpu.h (libhvd_gpu.so).

Device buffers are torch tensors on the HIP device; torch is only the
allocator and the stream here.  Every call goes through the C ABI -- there is
no CPU fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import math
import pathlib
import subprocess

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent.parent  # hvd-cvbs-decoding_amd/
LIB_PATH = PKG / "lib" / "libhvd_gpu.so"

ABI_SYMBOLS = (
    "hvd_gpu_create", "hvd_gpu_destroy", "hvd_gpu_last_error", "hvd_gpu_synchronize",
    "hvd_gpu_abi_version", "hvd_gpu_sample_to_ire_u16", "hvd_gpu_burst_lockin_phase",
    "hvd_gpu_burst_amplitude_ire", "hvd_gpu_holographic_init", "hvd_gpu_variational_refine",
    "hvd_gpu_decode_fields_u16", "hvd_gpu_decode_fields_ire", "hvd_gpu_block_match_sad_u16",
    "hvd_gpu_set_profiling", "hvd_gpu_cg_timing", "hvd_gpu_estimate_motion",
    "hvd_gpu_decode_frame_window_s16", "hvd_gpu_split_frame_fields", "hvd_gpu_field_parity_vote",
    "hvd_gpu_make_carrier", "hvd_gpu_tridiag_smooth", "hvd_gpu_box_blur", "hvd_gpu_chroma_plane",
    "hvd_gpu_weave_frame_s16", "hvd_gpu_variational_refine_stages", "hvd_gpu_refine_adjoint",
    "hvd_gpu_set_solver_tuning", "hvd_gpu_autotune_solver", "hvd_gpu_set_graphs",
    "hvd_gpu_variational_refine_temporal", "hvd_gpu_vectors_per_pixel", "hvd_gpu_warp_by_tiles",
    "hvd_gpu_warp_bilinear", "hvd_gpu_upsample_confidence", "hvd_gpu_complex_coherence",
    "hvd_gpu_motion_compensate_prev", "hvd_gpu_variational_refine_noalloc", "hvd_gpu_woven_size",
    "hvd_gpu_woven_picture", "hvd_gpu_variational_refine_temporal_batch", "hvd_gpu_variational_refine_joint",
    "hvd_gpu_verify_motion", "hvd_gpu_decode_window", "hvd_gpu_decode_frame_window_ex_s16",
    "hvd_gpu_difference", "hvd_gpu_estimate_noise_ire", "hvd_gpu_fft2d", "hvd_gpu_trajectory_fit",
    "hvd_gpu_trajectory_snap", "hvd_gpu_drizzle_frame", "hvd_gpu_ambiguous_boxes",
)


def build(arch: str = "gfx950") -> pathlib.Path:
    """Compile, from scratch, libhvd_gpu.so (hipcc, the kernels and the C ABI)
    and libhvd_core.so (g++, the hvd-core C++ API over it), in-tree."""
    subprocess.run(["make", "-s", "-B", "-j8", "-C", str(PKG), f"ARCH={arch}"], check=True)
    return LIB_PATH


class Geometry(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "field_width", "field_height", "active_video_start", "active_video_end",
        "colour_burst_start", "colour_burst_end", "first_active_field_line",
        "last_active_field_line")] + [("sample_rate", C.c_double), ("subcarrier_hz", C.c_double),
                                      ("standard", C.c_int)]

    @classmethod
    def of(cls, g) -> "Geometry":
        return cls(g.field_width, g.field_height, g.active_video_start, g.active_video_end,
                   g.colour_burst_start, g.colour_burst_end, g.first_active_field_line,
                   g.last_active_field_line, g.sample_rate, g.subcarrier_hz, g.standard)


class SolverParams(C.Structure):
    _fields_ = [("lambda_c", C.c_double), ("chroma_aniso", C.c_double),
                ("charbonnier_eps", C.c_double), ("chroma_eps", C.c_double),
                ("structure_coupling", C.c_double), ("diag_prior", C.c_double),
                ("cg_iterations", C.c_int), ("irls_outer", C.c_int), ("cg_tol", C.c_double),
                ("fast", C.c_int), ("is_pal", C.c_int), ("lambda_y", C.c_double)]

    @classmethod
    def make(cls, **kw) -> "SolverParams":
        """hvd::HvdConfig defaults (engine/hvd_config.h:26-102), overridable."""
        d = dict(lambda_c=1.0, chroma_aniso=0.0, charbonnier_eps=0.5, chroma_eps=1.0,
                 structure_coupling=0.25, diag_prior=0.0, cg_iterations=2, irls_outer=4,
                 cg_tol=0.0, fast=1, is_pal=0, lambda_y=1.0)
        d.update(kw)
        return cls(**d)


class SolverTuning(C.Structure):
    """Solver launch shape: lines per strip, most pixels per solve group (0: built-in)."""
    _fields_ = [("strip_rows", C.c_int), ("group_pixels", C.c_longlong)]


class Plane(C.Structure):
    _fields_ = [("lines", C.c_int), ("width", C.c_int), ("s", C.c_void_p), ("theta", C.c_void_p),
                ("vswitch", C.c_void_p), ("carrier", C.c_void_p), ("active_video_start", C.c_int),
                ("phase_per_sample", C.c_double), ("sample_rate", C.c_double), ("chi", C.c_void_p),
                ("y", C.c_void_p)]


class Neighbor(C.Structure):
    """hvd::NeighborTerm (variational.h:54-58) as device pointers."""
    _fields_ = [("composite", C.c_void_p), ("carrier", C.c_void_p), ("confidence", C.c_void_p)]


class WindowParams(C.Structure):
    """hvd_window_params: the window driver's HvdConfig knobs (hvd_config.h defaults)."""
    _fields_ = [(n, C.c_double) for n in ("temporal_strength", "temporal_eps", "nr_anchor", "nr_eps",
                                          "coherence_gate", "odd_gate_floor")] + \
        [(n, C.c_int) for n in ("passes", "nr_radius", "mc_tile", "mc_search", "bidirectional", "extended_temporal",
                                "trajectory_fit", "psi_init", "output_fidelity", "parallel_across_fields",
                                "selective_3d")] + [("selective_max_area", C.c_float)]

    @classmethod
    def make(cls, **kw) -> "WindowParams":
        d = dict(temporal_strength=0.0, temporal_eps=0.0, nr_anchor=1.0, nr_eps=0.0, coherence_gate=0.6,
                 odd_gate_floor=0.35, passes=2, nr_radius=2, mc_tile=32, mc_search=16, bidirectional=1,
                 extended_temporal=1, trajectory_fit=1, psi_init=0, output_fidelity=1, parallel_across_fields=1,
                 selective_3d=0, selective_max_area=0.45)
        d.update(kw)
        return cls(**d)


class WindowDiag(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("resolved_strength", "ambiguity_ire", "sigma_ire", "temporal_eps",
                                         "nr_eps")] + [("selective_boxes", C.c_int)]


class Anchor(C.Structure):
    """hvd::AnchorTerm (variational.h:60-66) as device pointers."""
    _fields_ = [("luma", C.c_void_p), ("chroma", C.c_void_p), ("weight", C.c_void_p)]


class FrameParams(C.Structure):
    """hvd::FrameParams (frame_bridge.h:40-101)."""
    _fields_ = [(n, C.c_int) for n in (
        "standard", "frame_width", "frame_height", "field1_lines", "active_video_start", "active_video_end",
        "colour_burst_start", "colour_burst_end", "first_active_frame_line", "last_active_frame_line")] + \
        [(n, C.c_float) for n in ("black_level", "white_level", "blanking_level", "chroma_dc")] + \
        [("sample_rate", C.c_double), ("subcarrier_hz", C.c_double), ("extra_sample_lines", C.c_int32 * 8),
         ("extra_sample_line_count", C.c_int)]


class BridgeParams(C.Structure):
    _fields_ = [("chroma_gain", C.c_float), ("acc", C.c_int), ("field_order", C.c_int), ("monochrome", C.c_int),
                ("chroma_phase_deg", C.c_float), ("temporal_strength", C.c_float)]

    @classmethod
    def make(cls, **kw) -> "BridgeParams":
        d = dict(chroma_gain=1.0, acc=1, field_order=0, monochrome=0, chroma_phase_deg=0.0, temporal_strength=-1.0)
        d.update(kw)
        return cls(**d)


class YcFrame(C.Structure):
    _fields_ = [("luma", C.c_void_p), ("chroma", C.c_void_p), ("u_plane", C.c_void_p), ("v_plane", C.c_void_p)]


class HvdGpuError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HvdGpuError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        lib = C.CDLL(str(LIB_PATH))
        vp, i, pp = C.c_void_p, C.c_int, C.POINTER(C.c_void_p)
        lib.hvd_gpu_create.argtypes = [i, C.POINTER(vp)]
        lib.hvd_gpu_destroy.argtypes = [vp]
        lib.hvd_gpu_destroy.restype = None
        lib.hvd_gpu_last_error.argtypes = [vp]
        lib.hvd_gpu_last_error.restype = C.c_char_p
        lib.hvd_gpu_synchronize.argtypes = [vp, vp]
        lib.hvd_gpu_sample_to_ire_u16.argtypes = [vp, vp, C.c_size_t, C.c_float, C.c_float, vp, vp]
        lib.hvd_gpu_burst_lockin_phase.argtypes = [vp, pp, i, C.POINTER(Geometry), pp, pp, vp]
        lib.hvd_gpu_burst_amplitude_ire.argtypes = [vp, pp, i, C.POINTER(Geometry), vp, vp]
        lib.hvd_gpu_holographic_init.argtypes = [vp, C.POINTER(Plane), i, i, vp]
        lib.hvd_gpu_variational_refine.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(SolverParams),
                                                   C.POINTER(C.c_int), C.POINTER(C.c_float), vp]
        lib.hvd_gpu_variational_refine_noalloc.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(SolverParams),
                                                           C.POINTER(C.c_int), C.POINTER(C.c_float), vp]
        lib.hvd_gpu_woven_size.argtypes = [C.POINTER(FrameParams), i, i, C.POINTER(i), C.POINTER(i)]
        lib.hvd_gpu_woven_picture.argtypes = [vp, C.POINTER(YcFrame), C.POINTER(FrameParams), i, i, vp, vp, vp, vp,
                                              vp]
        lib.hvd_gpu_variational_refine_temporal_batch.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(Neighbor),
                                                                  C.POINTER(C.c_int), C.POINTER(SolverParams),
                                                                  C.c_float, C.c_float, vp, vp]
        lib.hvd_gpu_variational_refine_joint.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(Neighbor),
                                                         C.POINTER(C.c_int), C.POINTER(Anchor),
                                                         C.POINTER(SolverParams), C.c_float, C.c_float, C.c_float,
                                                         vp, vp]
        lib.hvd_gpu_verify_motion.argtypes = [vp, pp, pp, i, i, i, pp, pp, i, pp, pp, pp, vp]
        lib.hvd_gpu_difference.argtypes = [vp, i, i, vp, vp, i, i, vp, vp]
        lib.hvd_gpu_estimate_noise_ire.argtypes = [vp, vp, i, i, C.POINTER(C.c_float), vp]
        lib.hvd_gpu_fft2d.argtypes = [vp, vp, i, i, i, vp, vp]
        lib.hvd_gpu_ambiguous_boxes.argtypes = [vp, vp, vp, i, i, C.c_float, C.POINTER(i), i, C.POINTER(i), vp, vp]
        lib.hvd_gpu_trajectory_fit.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_float), pp, pp, pp, i, i,
                                               vp, vp, vp]
        lib.hvd_gpu_trajectory_snap.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_float), pp, pp, pp, i, i, vp]
        lib.hvd_gpu_drizzle_frame.argtypes = [vp, i, i, pp, pp, C.POINTER(C.c_int), i, i, i, i, i, C.c_float, i, vp,
                                              vp, vp]
        lib.hvd_gpu_decode_window.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(C.c_int), i, C.POINTER(C.c_int),
                                              C.POINTER(SolverParams), C.POINTER(WindowParams),
                                              C.POINTER(WindowDiag), vp]
        lib.hvd_gpu_decode_fields_u16.argtypes = [vp, pp, i, C.POINTER(Geometry), C.c_float, C.c_float,
                                                  C.POINTER(SolverParams), C.c_float, pp, pp, pp, vp]
        lib.hvd_gpu_decode_fields_ire.argtypes = [vp, pp, i, C.POINTER(Geometry), C.POINTER(SolverParams),
                                                  C.c_float, pp, pp, pp, vp]
        lib.hvd_gpu_block_match_sad_u16.argtypes = [vp, pp, pp, i, i, i, i, i, i, i, pp, pp, pp, vp]
        lib.hvd_gpu_estimate_motion.argtypes = [vp, pp, pp, i, i, i, i, i, pp, pp, pp, vp]
        lib.hvd_gpu_decode_frame_window_s16.argtypes = [vp, pp, i, i, i, C.POINTER(FrameParams),
                                                        C.POINTER(SolverParams), C.POINTER(BridgeParams),
                                                        C.POINTER(YcFrame), C.POINTER(C.c_float), vp]
        lib.hvd_gpu_decode_frame_window_ex_s16.argtypes = [vp, pp, i, i, i, C.POINTER(FrameParams),
                                                           C.POINTER(SolverParams), C.POINTER(BridgeParams),
                                                           C.POINTER(WindowParams), C.POINTER(YcFrame),
                                                           C.POINTER(C.c_float), C.POINTER(WindowDiag), vp]
        lib.hvd_gpu_split_frame_fields.argtypes = [vp, vp, C.POINTER(FrameParams), C.c_float, vp, vp, vp]
        lib.hvd_gpu_field_parity_vote.argtypes = [vp, vp, vp, C.POINTER(FrameParams), C.POINTER(C.c_int), vp]
        lib.hvd_gpu_make_carrier.argtypes = [vp, vp, vp, i, C.POINTER(Geometry), vp, vp]
        lib.hvd_gpu_tridiag_smooth.argtypes = [vp, vp, vp, i, C.c_float, vp, vp]
        lib.hvd_gpu_box_blur.argtypes = [vp, vp, i, i, i, vp, vp]
        lib.hvd_gpu_chroma_plane.argtypes = [vp, i, vp, vp, vp, vp, C.c_size_t, vp, vp]
        lib.hvd_gpu_weave_frame_s16.argtypes = [vp, vp, C.POINTER(FrameParams), i, i, vp, vp, vp, C.c_float,
                                                C.POINTER(YcFrame), vp]
        lib.hvd_gpu_variational_refine_stages.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(SolverParams), pp, vp]
        lib.hvd_gpu_vectors_per_pixel.argtypes = [vp, vp, vp, i, i, i, i, i, vp, vp, vp]
        lib.hvd_gpu_warp_by_tiles.argtypes = [vp, vp, i, i, i, vp, vp, vp, vp]
        lib.hvd_gpu_warp_bilinear.argtypes = [vp, vp, i, i, vp, vp, C.c_float, i, i, vp, vp]
        lib.hvd_gpu_upsample_confidence.argtypes = [vp, vp, i, i, i, i, i, i, vp, vp]
        lib.hvd_gpu_complex_coherence.argtypes = [vp, vp, vp, i, i, i, vp, vp]
        lib.hvd_gpu_motion_compensate_prev.argtypes = [vp, vp, vp, i, i, vp, vp, vp, i, i, i, i, vp, vp, vp, vp]
        lib.hvd_gpu_variational_refine_temporal.argtypes = [vp, C.POINTER(Plane), C.POINTER(Neighbor), i,
                                                            C.POINTER(SolverParams), C.c_float, C.c_float, vp, vp]
        lib.hvd_gpu_refine_adjoint.argtypes = [vp, C.POINTER(Plane), i, C.POINTER(SolverParams), pp, pp, pp, i,
                                               C.c_double, pp, pp, C.POINTER(C.c_double), vp]
        lib.hvd_gpu_set_profiling.argtypes = [vp, i]
        lib.hvd_gpu_cg_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_double)]
        lib.hvd_gpu_set_solver_tuning.argtypes = [vp, C.POINTER(SolverTuning)]
        lib.hvd_gpu_set_graphs.argtypes = [vp, i]
        lib.hvd_gpu_autotune_solver.argtypes = [vp, i, i, i, i, i, C.POINTER(SolverTuning), C.POINTER(C.c_double)]
        _lib = lib
    return _lib


def _ptrs(tensors) -> C.Array:
    arr = (C.c_void_p * max(len(tensors), 1))()
    for k, t in enumerate(tensors):
        arr[k] = None if t is None else t.data_ptr()
    return arr


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Engine:
    """One GPU context (the analogue of an hvd::HvdEngine: one per thread)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.ctx = C.c_void_p()
        rc = self.lib.hvd_gpu_create(device, C.byref(self.ctx))
        if rc != 0:
            raise HvdGpuError(f"hvd_gpu_create({device}) failed with status {rc}")
        self.device = device

    def close(self):
        if self.ctx:
            self.lib.hvd_gpu_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.hvd_gpu_last_error(self.ctx).decode()
            raise HvdGpuError(f"{what}: status {rc}: {msg}")

    def synchronize(self):
        self._check(self.lib.hvd_gpu_synchronize(self.ctx, _stream()), "synchronize")

    def set_profiling(self, on: bool):
        self._check(self.lib.hvd_gpu_set_profiling(self.ctx, int(on)), "set_profiling")

    def cg_timing(self):
        """(CG-loop ms, CG iterations launched per solve group, pixel-iterations) since profiling was enabled."""
        ms, nl, pxit = C.c_double(), C.c_longlong(), C.c_double()
        self._check(self.lib.hvd_gpu_cg_timing(self.ctx, C.byref(ms), C.byref(nl), C.byref(pxit)), "cg_timing")
        return ms.value, nl.value, pxit.value

    # ---- solver launch shape ------------------------------------------
    def set_graphs(self, on: bool):
        """Replay each forward solve group's launches from one cached HIP graph (default on)."""
        self._check(self.lib.hvd_gpu_set_graphs(self.ctx, int(on)), "set_graphs")

    def set_solver_tuning(self, strip_rows: int = 0, group_pixels: int = 0):
        """Fix the solve's strip height and solve-group size (zeros: autotuned / built-in)."""
        t = SolverTuning(strip_rows, group_pixels)
        self._check(self.lib.hvd_gpu_set_solver_tuning(self.ctx, C.byref(t)), "set_solver_tuning")

    def autotune_solver(self, lines: int, width: int, is_pal: bool, nfields: int, iterations: int = 8):
        """Time every candidate launch shape on synthetic fields of this shape and record the fastest.

        Returns (strip_rows, group_pixels, ns per pixel-iteration)."""
        t, ns = SolverTuning(), C.c_double()
        self._check(self.lib.hvd_gpu_autotune_solver(self.ctx, lines, width, int(is_pal), nfields, iterations,
                                                     C.byref(t), C.byref(ns)), "autotune_solver")
        return t.strip_rows, t.group_pixels, ns.value

    # ---- level conversion ----------------------------------------------
    def sample_to_ire_u16(self, raw, blanking: float, white: float):
        import torch
        out = torch.empty(raw.shape, dtype=torch.float32, device=raw.device)
        self._check(self.lib.hvd_gpu_sample_to_ire_u16(self.ctx, C.c_void_p(raw.data_ptr()), raw.numel(),
                                                       blanking, white, C.c_void_p(out.data_ptr()), _stream()),
                    "sample_to_ire_u16")
        return out

    # ---- lock-in -----------------------------------------------------------
    def burst_lockin_phase(self, fields, geoms):
        """fields: list of (fh, fw) float32 IRE tensors.  Returns (theta list,
        vswitch list (None for NTSC))."""
        import torch
        n = len(fields)
        G = (Geometry * n)(*[Geometry.of(g) for g in geoms])
        th = [torch.empty(g.field_height, dtype=torch.float32, device=f.device) for f, g in zip(fields, geoms)]
        vs = [torch.empty(g.field_height, dtype=torch.int8, device=f.device) if g.standard in (1, 2) else None
              for f, g in zip(fields, geoms)]
        self._check(self.lib.hvd_gpu_burst_lockin_phase(self.ctx, _ptrs(fields), n, G, _ptrs(th), _ptrs(vs),
                                                        _stream()), "burst_lockin_phase")
        return th, vs

    def burst_amplitude_ire(self, fields, geoms):
        import torch
        n = len(fields)
        G = (Geometry * n)(*[Geometry.of(g) for g in geoms])
        out = torch.empty(n, dtype=torch.float32, device=fields[0].device)
        self._check(self.lib.hvd_gpu_burst_amplitude_ire(self.ctx, _ptrs(fields), n, G, C.c_void_p(out.data_ptr()),
                                                         _stream()), "burst_amplitude_ire")
        return out

    # ---- init / solve --------------------------------------------------------
    @staticmethod
    def plane(s, chi, y=None, theta=None, vswitch=None, carrier=None, a0=0, w0=0.5 * math.pi,
              sample_rate=4.0 * 315.0e6 / 88.0) -> Plane:
        """chi/carrier: (lines, width, 2) float32 (interleaved complex)."""
        lines, width = s.shape
        p = Plane(lines, width, s.data_ptr(), None if theta is None else theta.data_ptr(),
                  None if vswitch is None else vswitch.data_ptr(),
                  None if carrier is None else carrier.data_ptr(), a0, w0, sample_rate,
                  chi.data_ptr(), None if y is None else y.data_ptr())
        # the struct holds raw device pointers: keep the tensors alive with it
        p._keep = (s, chi, y, theta, vswitch, carrier)
        return p

    def holographic_init(self, planes, symmetry_variant=False):
        n = len(planes)
        P = (Plane * n)(*planes)
        self._check(self.lib.hvd_gpu_holographic_init(self.ctx, P, n, int(symmetry_variant), _stream()),
                    "holographic_init")

    def variational_refine(self, planes, params: SolverParams, want_stats=False, no_alloc=False):
        """no_alloc: through hvd_gpu_variational_refine_noalloc (the per-GPU
        fixed scratch instead of the context's workspace)."""
        n = len(planes)
        P = (Plane * n)(*planes)
        outer = max(params.irls_outer, 1)
        iters = (C.c_int * (n * outer))() if want_stats else None
        aniso = (C.c_float * n)() if want_stats else None
        fn = self.lib.hvd_gpu_variational_refine_noalloc if no_alloc else self.lib.hvd_gpu_variational_refine
        self._check(fn(self.ctx, P, n, C.byref(params), iters, aniso, _stream()),
                    "variational_refine_noalloc" if no_alloc else "variational_refine")
        if want_stats:
            return np.array(iters[:], dtype=np.int32).reshape(n, outer), np.array(aniso[:], dtype=np.float32)
        return None

    def variational_refine_temporal(self, plane, neighbours, params: SolverParams, temporal_strength,
                                    temporal_eps, want_stats=False):
        """VariationalRefine of one plane with temporal neighbours
        (variational.cpp:198-556): neighbours is a list of (composite,
        carrier (lines, width, 2), confidence) float32 device tensors on the
        plane's grid.  Returns the per-stage CG iterations when want_stats."""
        nn = len(neighbours)
        keep = [t for nb in neighbours for t in nb]
        N = (Neighbor * max(nn, 1))(*[Neighbor(a.data_ptr(), b.data_ptr(), c.data_ptr()) for a, b, c in neighbours])
        outer = max(params.irls_outer, 1)
        iters = (C.c_int * outer)() if want_stats else None
        self._check(self.lib.hvd_gpu_variational_refine_temporal(
            self.ctx, C.byref(plane), N, nn, C.byref(params), float(temporal_strength), float(temporal_eps), iters,
            _stream()), "variational_refine_temporal")
        del keep
        if want_stats:
            return np.array(iters[:], dtype=np.int32)
        return None

    def _batch_args(self, planes, neighbours):
        n = len(planes)
        P = (Plane * n)(*planes)
        flat = [nb for per in neighbours for nb in per]
        N = (Neighbor * max(len(flat), 1))(*[Neighbor(a.data_ptr(), b.data_ptr(), c.data_ptr()) for a, b, c in flat])
        nn = (C.c_int * n)(*[len(per) for per in neighbours])
        return P, N, nn, flat

    def variational_refine_temporal_batch(self, planes, neighbours, params: SolverParams, temporal_strength,
                                          temporal_eps, want_stats=False):
        """The batched solve with neighbours: neighbours[i] lists plane i's
        (composite, carrier, confidence) device tensors."""
        P, N, nn, keep = self._batch_args(planes, neighbours)
        outer = max(params.irls_outer, 1)
        iters = (C.c_int * (len(planes) * outer))() if want_stats else None
        self._check(self.lib.hvd_gpu_variational_refine_temporal_batch(
            self.ctx, P, len(planes), N, nn, C.byref(params), float(temporal_strength), float(temporal_eps), iters,
            _stream()), "variational_refine_temporal_batch")
        del keep
        if want_stats:
            return np.array(iters[:], dtype=np.int32).reshape(len(planes), outer)
        return None

    def variational_refine_joint(self, planes, neighbours, anchors, params: SolverParams, temporal_strength,
                                 temporal_eps, nr_anchor, want_stats=False):
        """VariationalRefineJoint of a batch: planes[i].y holds Y0 and
        receives Y; anchors[i] is (luma, chroma, weight) device tensors or
        None."""
        P, N, nn, keep = self._batch_args(planes, neighbours)
        A = (Anchor * len(planes))(*[Anchor(None, None, None) if a is None else
                                     Anchor(a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr()) for a in anchors])
        outer = max(params.irls_outer, 1)
        iters = (C.c_int * (len(planes) * outer))() if want_stats else None
        self._check(self.lib.hvd_gpu_variational_refine_joint(
            self.ctx, P, len(planes), N, nn, A, C.byref(params), float(temporal_strength), float(temporal_eps),
            float(nr_anchor), iters, _stream()), "variational_refine_joint")
        del keep
        if want_stats:
            return np.array(iters[:], dtype=np.int32).reshape(len(planes), outer)
        return None

    # ---- gradients.h, EstimateNoiseIre, Fft2d -----------------------------------
    DIFF_OPS = {"Dx": 0, "Dy": 1, "DxT": 2, "DyT": 3, "D1": 4, "D1T": 5, "D2": 6, "D2T": 7}

    def difference(self, op, a, weight=None):
        """gradients.h's op on a (h, w) float32 or (h, w, 2) complex device tensor."""
        import torch
        cplx = a.dim() == 3
        h, w = a.shape[0], a.shape[1]
        out = torch.empty_like(a)
        self._check(self.lib.hvd_gpu_difference(self.ctx, self.DIFF_OPS[op], int(cplx), a.data_ptr(),
                                                None if weight is None else weight.data_ptr(), h, w, out.data_ptr(),
                                                _stream()), "difference")
        return out

    def estimate_noise_ire(self, s):
        v = C.c_float(0.0)
        self._check(self.lib.hvd_gpu_estimate_noise_ire(self.ctx, s.data_ptr(), s.shape[0], s.shape[1], C.byref(v),
                                                        _stream()), "estimate_noise_ire")
        return v.value

    def fft2d(self, z, inverse=False):
        """(h, w, 2) float32 device tensor -> its numpy-convention 2-D DFT."""
        import torch
        out = torch.empty_like(z)
        self._check(self.lib.hvd_gpu_fft2d(self.ctx, z.data_ptr(), z.shape[0], z.shape[1], int(inverse),
                                           out.data_ptr(), _stream()), "fft2d")
        return out

    def trajectory_fit(self, offsets, shifts, dys, dxs, confs):
        import torch
        m = len(offsets)
        th, tw = dys[0].shape
        vy = torch.empty((th, tw), dtype=torch.float32, device=dys[0].device)
        vx = torch.empty_like(vy)
        self._check(self.lib.hvd_gpu_trajectory_fit(self.ctx, m, (C.c_int * m)(*offsets), (C.c_float * m)(*shifts),
                                                    _ptrs(dys), _ptrs(dxs), _ptrs(confs), th, tw, vy.data_ptr(),
                                                    vx.data_ptr(), _stream()), "trajectory_fit")
        return vy, vx

    def trajectory_snap(self, offsets, shifts, dys, dxs, confs):
        """In place on the dys, dxs device tensors."""
        m = len(offsets)
        th, tw = dys[0].shape
        self._check(self.lib.hvd_gpu_trajectory_snap(self.ctx, m, (C.c_int * m)(*offsets), (C.c_float * m)(*shifts),
                                                     _ptrs(dys), _ptrs(dxs), _ptrs(confs), th, tw, _stream()),
                    "trajectory_snap")

    def drizzle_frame(self, j0, ys, chis, parities, mc_tile=32, mc_search=16, nr_radius=2, nr_eps=0.0, scale=2):
        import torch
        n = len(ys)
        h, w = ys[0].shape
        oy = torch.empty((2 * h * scale, w), dtype=torch.float32, device=ys[0].device)
        oc = torch.empty((2 * h * scale, w, 2), dtype=torch.float32, device=ys[0].device)
        self._check(self.lib.hvd_gpu_drizzle_frame(self.ctx, j0, n, _ptrs(ys), _ptrs(chis),
                                                   (C.c_int * n)(*[int(v) for v in parities]), h, w, mc_tile,
                                                   mc_search, nr_radius, nr_eps, scale, oy.data_ptr(), oc.data_ptr(),
                                                   _stream()), "drizzle_frame")
        return oy, oc

    def verify_motion(self, refs, curs, pdys, pdxs, tile=32):
        import torch
        n = len(refs)
        h, w = refs[0].shape
        th, tw = -(-h // tile), -(-w // tile)
        outs = [[torch.zeros((th, tw), dtype=torch.float32, device=refs[0].device) for _ in range(n)]
                for _ in range(3)]
        self._check(self.lib.hvd_gpu_verify_motion(self.ctx, _ptrs(refs), _ptrs(curs), n, h, w, _ptrs(pdys),
                                                   _ptrs(pdxs), tile, _ptrs(outs[0]), _ptrs(outs[1]),
                                                   _ptrs(outs[2]), _stream()), "verify_motion")
        return outs

    def decode_window(self, planes, parities, is_pal, params: SolverParams, wparams: WindowParams, y_given=None):
        """hvd-core's DecodeFieldWindowWithInits on the device: planes with s,
        an explicit carrier, chi (chi0 in / chi out) and y (Y0 in where
        y_given[j] / Y out).  Returns the diagnostics dict."""
        n = len(planes)
        P = (Plane * n)(*planes)
        par = (C.c_int * n)(*[int(v) for v in parities])
        yg = None if y_given is None else (C.c_int * n)(*[int(bool(v)) for v in y_given])
        d = WindowDiag()
        self._check(self.lib.hvd_gpu_decode_window(self.ctx, P, n, par, int(is_pal), yg, C.byref(params),
                                                   C.byref(wparams), C.byref(d), _stream()), "decode_window")
        return {k: float(getattr(d, k)) for k, _ in WindowDiag._fields_}

    def ambiguous_boxes(self, y, chi, max_area=0.45, with_scores=False):
        """AmbiguousBoxes of one decoded field (y: (h, w), chi: (h, w, 2)
        float32 on the device): None for ambiguity everywhere (the whole
        window in 3-D), else the list of (fy0, fy1, x0, x1) boxes (empty: the
        window stays 2-D); with_scores also returns the tile scores."""
        h, w = y.shape
        cap = 256
        bx = (C.c_int * (4 * cap))()
        nb = C.c_int()
        th, tw = (h + 15) // 16, (w + 15) // 16
        sc = np.zeros((th, tw), np.float32) if with_scores else None
        self._check(self.lib.hvd_gpu_ambiguous_boxes(self.ctx, y.data_ptr(), chi.data_ptr(), h, w,
                                                     C.c_float(max_area), bx, cap, C.byref(nb),
                                                     sc.ctypes.data if with_scores else None, _stream()),
                    "ambiguous_boxes")
        boxes = None if nb.value < 0 else [tuple(bx[4 * k:4 * k + 4]) for k in range(min(nb.value, cap))]
        return (boxes, sc) if with_scores else boxes

    # ---- temporal neighbours' warps (temporal.cpp) ----------------------------
    def vectors_per_pixel(self, mdy, mdx, tile, h, w):
        import torch
        dy = torch.empty((h, w), dtype=torch.float32, device=mdy.device)
        dx = torch.empty_like(dy)
        self._check(self.lib.hvd_gpu_vectors_per_pixel(self.ctx, mdy.data_ptr(), mdx.data_ptr(), mdy.shape[0],
                                                       mdy.shape[1], tile, h, w, dy.data_ptr(), dx.data_ptr(),
                                                       _stream()), "vectors_per_pixel")
        return dy, dx

    def warp_by_tiles(self, a, dy, dx):
        """a: (h, w) real or (h, w, 2) interleaved complex float32."""
        import torch
        out = torch.empty_like(a)
        ch = 2 if a.dim() == 3 else 1
        self._check(self.lib.hvd_gpu_warp_by_tiles(self.ctx, a.data_ptr(), a.shape[0], a.shape[1], ch, dy.data_ptr(),
                                                   dx.data_ptr(), out.data_ptr(), _stream()), "warp_by_tiles")
        return out

    def warp_bilinear(self, a, dy, dx, row_offset):
        import torch
        out = torch.empty(dy.shape, dtype=torch.float32, device=a.device)
        self._check(self.lib.hvd_gpu_warp_bilinear(self.ctx, a.data_ptr(), a.shape[0], a.shape[1], dy.data_ptr(),
                                                   dx.data_ptr(), float(row_offset), dy.shape[0], dy.shape[1],
                                                   out.data_ptr(), _stream()), "warp_bilinear")
        return out

    def upsample_confidence(self, conf, tile, h, w, fast=False):
        import torch
        out = torch.empty((h, w), dtype=torch.float32, device=conf.device)
        self._check(self.lib.hvd_gpu_upsample_confidence(self.ctx, conf.data_ptr(), conf.shape[0], conf.shape[1],
                                                         tile, h, w, int(fast), out.data_ptr(), _stream()),
                    "upsample_confidence")
        return out

    def complex_coherence(self, z1, z2, r):
        import torch
        out = torch.empty(z1.shape[:2], dtype=torch.float32, device=z1.device)
        self._check(self.lib.hvd_gpu_complex_coherence(self.ctx, z1.data_ptr(), z2.data_ptr(), z1.shape[0],
                                                       z1.shape[1], r, out.data_ptr(), _stream()),
                    "complex_coherence")
        return out

    def motion_compensate_prev(self, comp, car, mdy, mdx, mconf, tile, fast=False):
        """A neighbour's (composite, carrier (h, w, 2), confidence) warped by a
        tile motion field (MotionCompensatePrev)."""
        import torch
        oc, ocar, ocf = torch.empty_like(comp), torch.empty_like(car), torch.empty_like(comp)
        self._check(self.lib.hvd_gpu_motion_compensate_prev(
            self.ctx, comp.data_ptr(), car.data_ptr(), comp.shape[0], comp.shape[1], mdy.data_ptr(), mdx.data_ptr(),
            mconf.data_ptr(), mdy.shape[0], mdy.shape[1], tile, int(fast), oc.data_ptr(), ocar.data_ptr(),
            ocf.data_ptr(), _stream()), "motion_compensate_prev")
        return oc, ocar, ocf

    def variational_refine_stages(self, planes, params: SolverParams):
        """The solve, returning chi at the start of every IRLS stage: a list of
        (irls_outer, lines, width, 2) float32 tensors (the adjoint's input)."""
        import torch
        n = len(planes)
        P = (Plane * n)(*planes)
        outer = max(params.irls_outer, 1)
        dev = torch.device("cuda", torch.cuda.current_device())
        stages = [torch.empty((outer, p.lines, p.width, 2), dtype=torch.float32, device=dev) for p in planes]
        self._check(self.lib.hvd_gpu_variational_refine_stages(self.ctx, P, n, C.byref(params), _ptrs(stages),
                                                               _stream()), "variational_refine_stages")
        return stages

    def refine_adjoint(self, planes, params: SolverParams, stages, gchi, gy=None, adj_iterations=200,
                       adj_tol=1e-6):
        """Gradients through the solve (planes[i].chi holds the final chi):
        returns (dL/dS list, dL/dchi0 list, (n, 2) float64 array of
        dL/dlambda_y, dL/dlambda_c)."""
        import torch
        n = len(planes)
        P = (Plane * n)(*planes)
        dev = gchi[0].device
        gs = [torch.empty((p.lines, p.width), dtype=torch.float32, device=dev) for p in planes]
        gchi0 = [torch.empty((p.lines, p.width, 2), dtype=torch.float32, device=dev) for p in planes]
        glam = (C.c_double * (2 * n))()
        self._check(self.lib.hvd_gpu_refine_adjoint(self.ctx, P, n, C.byref(params), _ptrs(stages), _ptrs(gchi),
                                                    _ptrs(gy) if gy is not None else None, int(adj_iterations),
                                                    float(adj_tol), _ptrs(gs), _ptrs(gchi0), glam, _stream()),
                    "refine_adjoint")
        return gs, gchi0, np.array(glam[:], dtype=np.float64).reshape(n, 2)

    # ---- field driver -------------------------------------------------------
    def decode_fields_u16(self, raws, geoms, params: SolverParams, blanking=20480.0, white=61440.0,
                          chroma_phase_deg=0.0, want_theta=False, outputs=None):
        """raws: list of (fh, fw) uint16 tensors (torch.uint16 or int16 views).
        Returns (ys, chis, thetas)."""
        import torch
        n = len(raws)
        G = (Geometry * n)(*[Geometry.of(g) for g in geoms])
        if outputs is None:
            ys = [torch.empty((g.lines, g.width), dtype=torch.float32, device=r.device) for r, g in zip(raws, geoms)]
            chis = [torch.empty((g.lines, g.width, 2), dtype=torch.float32, device=r.device)
                    for r, g in zip(raws, geoms)]
        else:
            ys, chis = outputs
        ths = ([torch.empty(g.field_height, dtype=torch.float32, device=r.device) for r, g in zip(raws, geoms)]
               if want_theta else None)
        self._check(self.lib.hvd_gpu_decode_fields_u16(self.ctx, _ptrs(raws), n, G, blanking, white,
                                                       C.byref(params), chroma_phase_deg, _ptrs(ys), _ptrs(chis),
                                                       _ptrs(ths) if ths else None, _stream()),
                    "decode_fields_u16")
        return ys, chis, ths

    def decode_fields_ire(self, fields, geoms, params: SolverParams, chroma_phase_deg=0.0, want_theta=False):
        import torch
        n = len(fields)
        G = (Geometry * n)(*[Geometry.of(g) for g in geoms])
        ys = [torch.empty((g.lines, g.width), dtype=torch.float32, device=f.device) for f, g in zip(fields, geoms)]
        chis = [torch.empty((g.lines, g.width, 2), dtype=torch.float32, device=f.device)
                for f, g in zip(fields, geoms)]
        ths = ([torch.empty(g.field_height, dtype=torch.float32, device=f.device) for f, g in zip(fields, geoms)]
               if want_theta else None)
        self._check(self.lib.hvd_gpu_decode_fields_ire(self.ctx, _ptrs(fields), n, G, C.byref(params),
                                                       chroma_phase_deg, _ptrs(ys), _ptrs(chis),
                                                       _ptrs(ths) if ths else None, _stream()),
                    "decode_fields_ire")
        return ys, chis, ths

    # ---- frame bridge ---------------------------------------------------------
    def decode_frame_window_s16(self, frames, core_begin, core_end, fp: FrameParams, params: SolverParams,
                                bridge: BridgeParams, want_gain=False, wparams: "WindowParams" = None,
                                diag: dict = None):
        """frames: list of flat int16 device tensors (field-sequential frames).
        Returns a list of (luma, chroma, u, v) device tensors per core frame and
        the applied chroma gain (None unless want_gain).  With wparams (or a
        diag dict to fill) the _ex entry point runs with those window knobs."""
        import torch
        dev = frames[0].device
        n_px = fp.frame_width * fp.frame_height
        outs, descs = [], []
        for _ in range(core_end - core_begin):
            o = (torch.empty(n_px, dtype=torch.int16, device=dev), torch.empty(n_px, dtype=torch.int16, device=dev),
                 torch.empty(n_px, dtype=torch.float64, device=dev), torch.empty(n_px, dtype=torch.float64, device=dev))
            outs.append(o)
            descs.append(YcFrame(*[t.data_ptr() for t in o]))
        arr = (YcFrame * max(len(descs), 1))(*descs)
        gain = C.c_float(0.0)
        if wparams is not None or diag is not None:
            d = WindowDiag()
            self._check(self.lib.hvd_gpu_decode_frame_window_ex_s16(
                self.ctx, _ptrs(frames), len(frames), core_begin, core_end, C.byref(fp), C.byref(params),
                C.byref(bridge), C.byref(wparams) if wparams is not None else None, arr,
                C.byref(gain) if want_gain else None, C.byref(d), _stream()), "decode_frame_window_ex_s16")
            if diag is not None:
                diag.update({k: float(getattr(d, k)) for k, _ in WindowDiag._fields_})
            return outs, (gain.value if want_gain else None)
        self._check(self.lib.hvd_gpu_decode_frame_window_s16(
            self.ctx, _ptrs(frames), len(frames), core_begin, core_end, C.byref(fp), C.byref(params),
            C.byref(bridge), arr, C.byref(gain) if want_gain else None, _stream()), "decode_frame_window_s16")
        return outs, (gain.value if want_gain else None)

    # ---- frame-bridge and component pieces ------------------------------------
    def split_frame_fields(self, frame, fp: FrameParams, zero_code: float):
        """Block 1 / block 2 of a field-sequential int16 frame (device tensor)
        as (field1_lines, frame_width) float32 IRE planes."""
        import torch
        f1 = [torch.empty((fp.field1_lines, fp.frame_width), dtype=torch.float32, device=frame.device)
              for _ in range(2)]
        self._check(self.lib.hvd_gpu_split_frame_fields(self.ctx, frame.data_ptr(), C.byref(fp), zero_code,
                                                        f1[0].data_ptr(), f1[1].data_ptr(), _stream()),
                    "split_frame_fields")
        return f1[0], f1[1]

    def field_parity_vote(self, block1, block2, fp: FrameParams) -> int:
        v = C.c_int(0)
        self._check(self.lib.hvd_gpu_field_parity_vote(self.ctx, block1.data_ptr(), block2.data_ptr(), C.byref(fp),
                                                       C.byref(v), _stream()), "field_parity_vote")
        return v.value

    def make_carrier(self, theta, vswitch, g):
        """Explicit (lines, active_width, 2) carrier plane from per-line theta."""
        import torch
        lines = theta.shape[0]
        width = g.active_video_end - g.active_video_start
        out = torch.empty((lines, width, 2), dtype=torch.float32, device=theta.device)
        geo = Geometry.of(g)
        self._check(self.lib.hvd_gpu_make_carrier(self.ctx, theta.data_ptr(),
                                                  vswitch.data_ptr() if vswitch is not None else None, lines,
                                                  C.byref(geo), out.data_ptr(), _stream()), "make_carrier")
        return out

    def tridiag_smooth(self, d, a, lam):
        import torch
        out = torch.empty_like(d)
        self._check(self.lib.hvd_gpu_tridiag_smooth(self.ctx, d.data_ptr(), a.data_ptr(), d.numel(), lam,
                                                    out.data_ptr(), _stream()), "tridiag_smooth")
        return out

    def box_blur(self, plane, r):
        import torch
        out = torch.empty_like(plane)
        h, w = plane.shape
        self._check(self.lib.hvd_gpu_box_blur(self.ctx, plane.data_ptr(), h, w, r, out.data_ptr(), _stream()),
                    "box_blur")
        return out

    def chroma_plane(self, mode, s=None, y=None, chi=None, carrier=None):
        import torch
        ref = s if mode == 0 else chi
        n = ref.shape[0] * ref.shape[1]
        out = torch.empty(ref.shape[:2], dtype=torch.float32, device=ref.device)
        ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        self._check(self.lib.hvd_gpu_chroma_plane(self.ctx, mode, ptr(s), ptr(y), ptr(chi), ptr(carrier), n,
                                                  out.data_ptr(), _stream()), "chroma_plane")
        return out

    def weave_frame_s16(self, src, fp: FrameParams, y, chroma, chi, gain):
        """Woven (rows, width) planes -> flat (luma, chroma, u, v) frame tensors."""
        import torch
        dev = src.device
        n_px = fp.frame_width * fp.frame_height
        o = (torch.empty(n_px, dtype=torch.int16, device=dev), torch.empty(n_px, dtype=torch.int16, device=dev),
             torch.empty(n_px, dtype=torch.float64, device=dev), torch.empty(n_px, dtype=torch.float64, device=dev))
        desc = YcFrame(*[t.data_ptr() for t in o])
        rows, width = chroma.shape
        self._check(self.lib.hvd_gpu_weave_frame_s16(self.ctx, src.data_ptr(), C.byref(fp), rows, width,
                                                     y.data_ptr() if y is not None else None, chroma.data_ptr(),
                                                     chi.data_ptr(), gain, C.byref(desc), _stream()),
                    "weave_frame_s16")
        return o

    # ---- motion ---------------------------------------------------------------
    def woven_picture(self, luma, u, v, fp: FrameParams, full_raster=False, field=-1):
        """ReorderToWoven + RGB24 of a field-sequential YcFrameS16 (device
        tensors luma int16, u / v float64, frame_height x frame_width) ->
        (y, u, v float64, rgb uint8 [h, w, 3]) device tensors."""
        import torch
        w, h = C.c_int(), C.c_int()
        self._check(self.lib.hvd_gpu_woven_size(C.byref(fp), int(full_raster), int(field), C.byref(w), C.byref(h)),
                    "woven_size")
        dev = luma.device
        yo = torch.zeros((h.value, w.value), dtype=torch.float64, device=dev)
        uo, vo = torch.zeros_like(yo), torch.zeros_like(yo)
        rgb = torch.zeros((h.value, w.value, 3), dtype=torch.uint8, device=dev)
        yc = YcFrame(luma.data_ptr(), None, u.data_ptr(), v.data_ptr())
        if h.value and w.value:
            self._check(self.lib.hvd_gpu_woven_picture(self.ctx, C.byref(yc), C.byref(fp), int(full_raster),
                                                       int(field), yo.data_ptr(), uo.data_ptr(), vo.data_ptr(),
                                                       rgb.data_ptr(), _stream()), "woven_picture")
        return yo, uo, vo, rgb

    def estimate_motion(self, refs, curs, tile=32, search=8):
        """hvd-core EstimateMotion for pairs of (h, w) float32 planes: returns
        (dy list, dx list, confidence list), each ceil(h/tile) x ceil(w/tile)."""
        import torch
        n = len(refs)
        h, w = refs[0].shape
        th, tw = -(-h // tile), -(-w // tile)
        dev = refs[0].device
        mk = lambda: [torch.empty((th, tw), dtype=torch.float32, device=dev) for _ in range(n)]  # noqa: E731
        dy, dx, cf = mk(), mk(), mk()
        self._check(self.lib.hvd_gpu_estimate_motion(self.ctx, _ptrs(refs), _ptrs(curs), n, h, w, tile, search,
                                                     _ptrs(dy), _ptrs(dx), _ptrs(cf), _stream()), "estimate_motion")
        return dy, dx, cf

    def block_match_sad_u16(self, refs, curs, bh=8, bw=16, sy=4, sx=16):
        import torch
        n = len(refs)
        h, w = refs[0].shape
        nby, nbx = -(-h // bh), -(-w // bw)
        dev = refs[0].device
        vy = [torch.empty((nby, nbx), dtype=torch.int16, device=dev) for _ in range(n)]
        vx = [torch.empty((nby, nbx), dtype=torch.int16, device=dev) for _ in range(n)]
        sad = [torch.empty((nby, nbx), dtype=torch.int32, device=dev) for _ in range(n)]
        self._check(self.lib.hvd_gpu_block_match_sad_u16(self.ctx, _ptrs(refs), _ptrs(curs), n, h, w, bh, bw, sy, sx,
                                                         _ptrs(vy), _ptrs(vx), _ptrs(sad), _stream()),
                    "block_match_sad_u16")
        return vy, vx, sad


def geometry_phase_per_sample(g) -> float:
    if g.subcarrier_hz <= 0:
        return 0.5 * math.pi
    return 2 * math.pi * g.subcarrier_hz / g.sample_rate
