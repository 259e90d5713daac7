"""ctypes front end for the CPU oracle (liboracle64/32.so) and, where it has
been built, hvd-core's own FFT-free engine (oracle/_ref/libhvdref.so).

TEST INFRASTRUCTURE ONLY: imported by tests, the smoke check and the
benchmark's CPU-baseline leg, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB64 = HERE / "lib" / "liboracle64.so"
LIB32 = HERE / "lib" / "liboracle32.so"
REFLIB = HERE / "_ref" / "libhvdref.so"
REFLIB_OMP = HERE / "_ref" / "libhvdref_omp.so"
REF_SRC = pathlib.Path("/root/reference/hvd-core/src")


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and the hvd-core shim when the reference checkout
    is present)."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "all"], check=True)
    if with_ref is None:
        with_ref = REF_SRC.is_dir()
    if with_ref:
        subprocess.run(["make", "-s", "-B", "-C", str(HERE), "ref"], check=True)
        # hvd-core's own test programs against the package's C++ API (needs
        # the package's libhvd_core.so, built first by __graft_entry__.build)
        if (HERE.parent / "hvd-cvbs-decoding_amd" / "lib" / "libhvd_core.so").exists():
            subprocess.run(["make", "-s", "-B", "-C", str(HERE), "reftests"], check=True)


class Geom(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "field_width", "field_height", "active_video_start", "active_video_end",
        "colour_burst_start", "colour_burst_end", "first_active_field_line",
        "last_active_field_line")] + [("sample_rate", C.c_double), ("subcarrier_hz", C.c_double),
                                      ("standard", C.c_int)]

    @classmethod
    def of(cls, g) -> "Geom":
        return cls(g.field_width, g.field_height, g.active_video_start, g.active_video_end,
                   g.colour_burst_start, g.colour_burst_end, g.first_active_field_line,
                   g.last_active_field_line, g.sample_rate, g.subcarrier_hz, g.standard)


class Cfg(C.Structure):
    _fields_ = [("lambda_c", C.c_double), ("chroma_aniso", C.c_double),
                ("charbonnier_eps", C.c_double), ("chroma_eps", C.c_double),
                ("structure_coupling", C.c_double), ("diag_prior", C.c_double),
                ("cg_iterations", C.c_int), ("irls_outer", C.c_int), ("cg_tol", C.c_double),
                ("fast", C.c_int), ("is_pal", C.c_int), ("lambda_y", C.c_double)]

    @classmethod
    def make(cls, **kw) -> "Cfg":
        d = dict(lambda_c=1.0, chroma_aniso=0.0, charbonnier_eps=0.5, chroma_eps=1.0,
                 structure_coupling=0.25, diag_prior=0.0, cg_iterations=2, irls_outer=4,
                 cg_tol=0.0, fast=1, is_pal=0, lambda_y=1.0)
        d.update(kw)
        return cls(**d)


class SeqCfg(C.Structure):
    """The HvdConfig fields of the window driver (sequence.cpp) beyond the
    solver's; defaults are hvd-core's (engine/hvd_config.h)."""
    _fields_ = [("base", Cfg)] + [(n, C.c_double) for n in (
        "temporal_strength", "temporal_eps", "nr_anchor", "nr_eps", "coherence_gate", "odd_gate_floor")] + \
        [(n, C.c_int) for n in ("passes", "nr_radius", "mc_tile", "mc_search", "bidirectional", "extended_temporal",
                                "trajectory_fit", "psi_init", "output_fidelity", "parallel_across_fields")]

    @classmethod
    def make(cls, **kw) -> "SeqCfg":
        names = {f[0] for f in Cfg._fields_}
        base = Cfg.make(**{k: v for k, v in kw.items() if k in names})
        d = dict(temporal_strength=0.0, temporal_eps=0.0, nr_anchor=1.0, nr_eps=0.0, coherence_gate=0.6,
                 odd_gate_floor=0.35, passes=2, nr_radius=2, mc_tile=32, mc_search=16, bidirectional=1,
                 extended_temporal=1, trajectory_fit=1, psi_init=0, output_fidelity=1, parallel_across_fields=1)
        d.update({k: v for k, v in kw.items() if k not in names})
        return cls(base=base, **d)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The restatement, FP64 (`fp32=False`) or FP32 with FP64 reductions."""

    def __init__(self, fp32: bool = False, path: str | os.PathLike | None = None):
        """`path` loads another build of the same source (the init-pin tests
        load builds with a planted mistake)."""
        if path is None:
            path = LIB32 if fp32 else LIB64
            if not path.exists():
                build(with_ref=False)
        self.lib = C.CDLL(str(path))
        self.dt = np.float32 if fp32 else np.float64
        self.pre = "or32_" if fp32 else "or64_"
        self.R = C.c_float if fp32 else C.c_double
        for name, res in (("phase_per_sample", C.c_double), ("line_advance", C.c_double),
                          ("burst_amplitude_ire", self.R), ("acc_gain", self.R),
                          ("resolve_chroma_aniso", self.R), ("variational_refine", self.R),
                          ("variational_refine_ex", self.R), ("variational_refine_nbr", self.R),
                          ("decode_field", self.R)):
            getattr(self.lib, self.pre + name).restype = res

    def f(self, name):
        return getattr(self.lib, self.pre + name)

    def arr(self, a, dtype=None):
        return np.ascontiguousarray(a, dtype=dtype or self.dt)

    # -- lock-in -------------------------------------------------------------
    def burst_lockin(self, field, x0, x1, w0):
        field = self.arr(field)
        h, w = field.shape
        z = np.zeros(2 * h, self.dt)
        amp = np.zeros(h, self.dt)
        self.f("burst_lockin")(_p(field), h, w, x0, x1, C.c_double(w0), _p(z), _p(amp))
        return z[0::2] + 1j * z[1::2], amp

    def tridiag_smooth(self, d, a, lam):
        d, a = self.arr(d), self.arr(a)
        x = np.zeros_like(d)
        self.f("tridiag_smooth")(_p(d), _p(a), len(d), self.R(lam), _p(x))
        return x

    def burst_lockin_phase(self, field, g):
        field = self.arr(field)
        th = np.zeros(g.field_height, self.dt)
        self.f("burst_lockin_phase")(_p(field), C.byref(Geom.of(g)), _p(th))
        return th

    def burst_lockin_phase_pal(self, field, g):
        field = self.arr(field)
        th = np.zeros(g.field_height, self.dt)
        vs = np.zeros(g.field_height, np.int8)
        self.f("burst_lockin_phase_pal")(_p(field), C.byref(Geom.of(g)), _p(th), _p(vs))
        return th, vs

    def burst_amplitude_ire(self, field, g):
        return float(self.f("burst_amplitude_ire")(_p(self.arr(field)), C.byref(Geom.of(g))))

    def acc_gain(self, measured, nominal):
        return float(self.f("acc_gain")(self.R(measured), self.R(nominal)))

    def make_carrier(self, theta_active, vswitch_active, g):
        th = self.arr(theta_active)
        lines = len(th)
        out = np.zeros((lines, g.width * 2), self.dt)
        vs = None if vswitch_active is None else np.ascontiguousarray(vswitch_active, np.int8)
        self.f("make_carrier")(_p(th), None if vs is None else _p(vs), lines, C.byref(Geom.of(g)), _p(out))
        return out[:, 0::2] + 1j * out[:, 1::2]

    # -- FFT / operators -----------------------------------------------------
    def fft2d(self, a, sign):
        a = np.asarray(a)
        h, w = a.shape
        buf = np.zeros((h, 2 * w), self.dt)
        buf[:, 0::2], buf[:, 1::2] = a.real, a.imag
        self.f("fft2d")(_p(buf), h, w, sign)
        return buf[:, 0::2] + 1j * buf[:, 1::2]

    def real_op(self, op, a):
        a = self.arr(a)
        h, w = a.shape
        d = np.zeros_like(a)
        self.f(("dx", "dy", "dxt", "dyt")[op])(_p(a), h, w, _p(d))
        return d

    def cplx_op(self, op, a):
        a = np.asarray(a)
        h, w = a.shape
        buf = self.cbuf(a)
        out = np.zeros_like(buf)
        self.f("cplx_op")(op, _p(buf), h, w, _p(out))
        return self.uncbuf(out)

    def cbuf(self, a):
        a = np.asarray(a)
        buf = np.zeros(a.shape[:-1] + (2 * a.shape[-1],), self.dt)
        buf[..., 0::2], buf[..., 1::2] = a.real, a.imag
        return buf

    @staticmethod
    def uncbuf(b):
        return b[..., 0::2] + 1j * b[..., 1::2]

    # -- init / solve ----------------------------------------------------------
    def holographic_init(self, s, carrier, sample_rate, symmetry_variant=False):
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        y0 = np.zeros_like(s)
        chi = np.zeros((h, 2 * w), self.dt)
        self.f("holographic_init")(_p(s), _p(cb), h, w, C.c_double(sample_rate), int(symmetry_variant),
                                   _p(y0), _p(chi))
        return y0, self.uncbuf(chi)

    def init_variants(self, s, carrier, sample_rate, symmetry_variant=False):
        """The init's crop variants before the blend: [nvar, h, w] complex."""
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        out = np.zeros((3, h, 2 * w), self.dt)
        nvar = self.f("init_variants")(_p(s), _p(cb), h, w, C.c_double(sample_rate), int(symmetry_variant),
                                       _p(out))
        return self.uncbuf(out[:nvar])

    def init_blend(self, s, carrier, chi_v):
        """The init's edge-energy blend of given variants -> (y0, chi0)."""
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        vb = self.cbuf(np.asarray(chi_v))
        y0 = np.zeros_like(s)
        chi = np.zeros((h, 2 * w), self.dt)
        self.f("init_blend")(_p(s), _p(cb), _p(vb), len(chi_v), h, w, _p(y0), _p(chi))
        return y0, self.uncbuf(chi)

    def pad_rows_mirror(self, a, pad):
        a = self.cbuf(a)
        h, w2 = a.shape
        out = np.zeros((h + 2 * pad, w2), self.dt)
        self.f("pad_rows_mirror")(_p(a), h, w2 // 2, pad, _p(out))
        return self.uncbuf(out)

    def box_blur_zero(self, a, r):
        a = self.arr(a).copy()
        h, w = a.shape
        self.f("box_blur_zero")(_p(a), h, w, r)
        return a

    def resolve_chroma_aniso(self, chi0, is_pal, forced=0.0):
        chi0 = np.asarray(chi0)
        h, w = chi0.shape
        return float(self.f("resolve_chroma_aniso")(_p(self.cbuf(chi0)), h, w, int(is_pal), self.R(forced)))

    def variational_refine(self, s, carrier, chi0, cfg: Cfg):
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        chi = self.cbuf(chi0)
        y = np.zeros_like(s)
        iters = np.zeros(max(cfg.irls_outer, 1), np.int32)
        aniso = self.f("variational_refine")(_p(s), _p(cb), h, w, C.byref(cfg), _p(chi), _p(y), _p(iters))
        return y, self.uncbuf(chi), float(aniso), iters

    def variational_refine_nbr(self, s, carrier, chi0, cfg: Cfg, neighbours, nu, eps_t):
        """The solve with temporal neighbours: a list of (composite S_w,
        carrier c_w, confidence) planes (hvd-core's NeighborTerm), at
        temporal strength nu and temporal_eps eps_t."""
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        chi = self.cbuf(chi0)
        y = np.zeros_like(s)
        nn = len(neighbours)
        sw = np.ascontiguousarray(np.stack([self.arr(n[0]) for n in neighbours]) if nn else np.zeros((1, h, w)),
                                  self.dt)
        cw = np.ascontiguousarray(np.stack([self.cbuf(n[1]) for n in neighbours]) if nn else np.zeros((1, h, 2 * w)),
                                  self.dt)
        cf = np.ascontiguousarray(np.stack([self.arr(n[2]) for n in neighbours]) if nn else np.zeros((1, h, w)),
                                  self.dt)
        iters = np.zeros(max(cfg.irls_outer, 1), np.int32)
        aniso = self.f("variational_refine_nbr")(_p(s), _p(cb), h, w, C.byref(cfg), nn, _p(sw), _p(cw), _p(cf),
                                                 self.R(nu), self.R(eps_t), _p(chi), _p(y), _p(iters))
        return y, self.uncbuf(chi), float(aniso), iters

    def variational_refine_stages(self, s, carrier, chi0, cfg: Cfg, hh_floor=1e-12, gg_floor=1e-10):
        """The solve, also returning chi at the start of every IRLS stage
        (irls_outer complex planes), as the adjoint consumes them; the floors
        are the absolute CG exits (hvd-core's by default)."""
        s = self.arr(s)
        h, w = s.shape
        cb = self.cbuf(carrier)
        chi = self.cbuf(chi0)
        y = np.zeros_like(s)
        k = max(cfg.irls_outer, 1)
        iters = np.zeros(k, np.int32)
        stages = np.zeros((k, h, 2 * w), self.dt)
        aniso = self.f("variational_refine_ex")(_p(s), _p(cb), h, w, C.byref(cfg), _p(chi), _p(y), _p(iters),
                                                _p(stages), C.c_double(hh_floor), C.c_double(gg_floor))
        return y, self.uncbuf(chi), self.uncbuf(stages), float(aniso), iters

    def refine_adjoint(self, s, carrier, cfg: Cfg, stages, chi_final, gchi, gy=None, adj_iters=2000,
                       adj_tol=1e-13):
        """Gradients of L through the solve from dL/dchi and dL/dY:
        returns (dL/dS, dL/dchi0, [dL/dlambda_y, dL/dlambda_c], adjoint CG
        iterations)."""
        s = self.arr(s)
        h, w = s.shape
        st = self.cbuf(stages)
        gs = np.zeros_like(s)
        gchi0 = np.zeros((h, 2 * w), self.dt)
        glam = np.zeros(2, np.float64)
        gyb = None if gy is None else self.arr(gy)
        it = self.f("refine_adjoint")(_p(s), _p(self.cbuf(carrier)), h, w, C.byref(cfg), _p(st),
                                      _p(self.cbuf(chi_final)), _p(self.cbuf(gchi)),
                                      None if gyb is None else _p(gyb), int(adj_iters), C.c_double(adj_tol),
                                      _p(gs), _p(gchi0), _p(glam))
        return gs, self.uncbuf(gchi0), glam, int(it)

    def prepare_field(self, field_ire, g, chroma_phase_deg=0.0):
        field = self.arr(field_ire)
        s = np.zeros((g.lines, g.width), self.dt)
        car = np.zeros((g.lines, 2 * g.width), self.dt)
        th = np.zeros(g.field_height, self.dt)
        vs = np.zeros(g.field_height, np.int8)
        self.f("prepare_field")(_p(field), C.byref(Geom.of(g)), self.R(chroma_phase_deg), _p(s), _p(car),
                                _p(th), _p(vs))
        return s, self.uncbuf(car), th, vs

    def decode_field(self, field_ire, g, cfg: Cfg, chroma_phase_deg=0.0):
        field = self.arr(field_ire)
        y = np.zeros((g.lines, g.width), self.dt)
        chi = np.zeros((g.lines, 2 * g.width), self.dt)
        th = np.zeros(g.field_height, self.dt)
        vs = np.zeros(g.field_height, np.int8)
        iters = np.zeros(max(cfg.irls_outer, 1), np.int32)
        aniso = self.f("decode_field")(_p(field), C.byref(Geom.of(g)), C.byref(cfg), self.R(chroma_phase_deg),
                                       _p(y), _p(chi), _p(th), _p(vs), _p(iters))
        return dict(y=y, chi=self.uncbuf(chi), theta=th, vswitch=vs, aniso=float(aniso), iters=iters)

    # -- motion ---------------------------------------------------------------
    def box_blur(self, a, r):
        a = self.arr(a).copy()
        h, w = a.shape
        self.f("box_blur_prefix")(_p(a), h, w, r)
        return a

    # -- temporal neighbours' warps (temporal.cpp) ---------------------------
    def vectors_per_pixel(self, mdy, mdx, tile, h, w):
        mdy, mdx = self.arr(mdy), self.arr(mdx)
        dy, dx = np.zeros((h, w), self.dt), np.zeros((h, w), self.dt)
        self.f("vectors_per_pixel")(_p(mdy), _p(mdx), mdy.shape[0], mdy.shape[1], tile, h, w, _p(dy), _p(dx))
        return dy, dx

    def warp_by_tiles(self, a, dy, dx):
        """a: (h, w) real or complex; dy, dx: per-pixel vectors."""
        cplx = np.iscomplexobj(a)
        h, w = a.shape
        src = self.cbuf(a) if cplx else self.arr(a)
        out = np.zeros_like(src)
        self.f("warp_by_tiles")(_p(src), h, w, 2 if cplx else 1, _p(self.arr(dy)), _p(self.arr(dx)), _p(out))
        return self.uncbuf(out) if cplx else out

    def warp_bilinear(self, a, dy, dx, row_offset):
        a = self.arr(a)
        out_h, out_w = dy.shape
        out = np.zeros((out_h, out_w), self.dt)
        self.f("warp_bilinear")(_p(a), a.shape[0], a.shape[1], _p(self.arr(dy)), _p(self.arr(dx)),
                                self.R(row_offset), out_h, out_w, _p(out))
        return out

    def upsample_confidence(self, conf, tile, h, w, fast=False):
        conf = self.arr(conf)
        out = np.zeros((h, w), self.dt)
        self.f("upsample_confidence")(_p(conf), conf.shape[0], conf.shape[1], tile, h, w, int(fast), _p(out))
        return out

    def complex_coherence(self, z1, z2, r):
        h, w = z1.shape
        out = np.zeros((h, w), self.dt)
        self.f("complex_coherence")(_p(self.cbuf(z1)), _p(self.cbuf(z2)), h, w, r, _p(out))
        return out

    def motion_compensate_prev(self, comp, car, mdy, mdx, mconf, tile, fast=False):
        comp = self.arr(comp)
        h, w = comp.shape
        mdy, mdx, mconf = self.arr(mdy), self.arr(mdx), self.arr(mconf)
        oc, ocar, ocf = np.zeros((h, w), self.dt), np.zeros((h, 2 * w), self.dt), np.zeros((h, w), self.dt)
        self.f("motion_compensate_prev")(_p(comp), _p(self.cbuf(car)), h, w, _p(mdy), _p(mdx), _p(mconf),
                                         mdy.shape[0], mdy.shape[1], tile, int(fast), _p(oc), _p(ocar), _p(ocf))
        return oc, self.uncbuf(ocar), ocf

    def estimate_motion(self, y_ref, y_cur, tile=32, search=8):
        y_ref, y_cur = self.arr(y_ref), self.arr(y_cur)
        h, w = y_ref.shape
        th, tw = -(-h // tile), -(-w // tile)
        dy, dx, cf = (np.zeros((th, tw), self.dt) for _ in range(3))
        self.f("estimate_motion")(_p(y_ref), _p(y_cur), h, w, tile, search, _p(dy), _p(dx), _p(cf))
        return dy, dx, cf

    # -- 3-D sequence pieces --------------------------------------------------
    def estimate_noise_ire(self, s):
        s = self.arr(s)
        self.f("estimate_noise_ire").restype = self.R
        return float(self.f("estimate_noise_ire")(_p(s), s.shape[0], s.shape[1]))

    def verify_motion(self, y_ref, y_cur, pdy, pdx, tile=32):
        y_ref, y_cur, pdy, pdx = self.arr(y_ref), self.arr(y_cur), self.arr(pdy), self.arr(pdx)
        h, w = y_ref.shape
        dy, dx, cf = (np.zeros_like(pdy) for _ in range(3))
        self.f("verify_motion")(_p(y_ref), _p(y_cur), h, w, _p(pdy), _p(pdx), tile, _p(dy), _p(dx), _p(cf))
        return dy, dx, cf

    def trajectory_snap(self, offsets, pshift, dy, dx, conf):
        """dy, dx, conf: [m, th, tw]; returns snapped copies of dy, dx."""
        dy, dx, conf = self.arr(dy).copy(), self.arr(dx).copy(), self.arr(conf)
        m = dy.shape[0]
        off = np.ascontiguousarray(offsets, np.int32)
        ps = self.arr(pshift)
        self.f("trajectory_snap")(m, _p(off), _p(ps), _p(dy), _p(dx), _p(conf), dy[0].size)
        return dy, dx

    def fit_trajectory(self, offsets, pshift, dy, dx, conf):
        dy, dx, conf = self.arr(dy), self.arr(dx), self.arr(conf)
        vy, vx = np.zeros_like(dy[0]), np.zeros_like(dy[0])
        self.f("fit_trajectory")(dy.shape[0], _p(np.ascontiguousarray(offsets, np.int32)), _p(self.arr(pshift)),
                                 _p(dy), _p(dx), _p(conf), dy[0].size, _p(vy), _p(vx))
        return vy, vx

    def joint_refine(self, s, carrier, y0, chi0, cfg: Cfg, neighbours=(), nu=0.0, eps_t=1.0, anchor=None,
                     nu_a=0.0):
        """VariationalRefineJoint -> (y, chi, aniso, stage iterations);
        neighbours: (S_w, c_w, conf) triples; anchor: (Y_a, chi_a, w_a)."""
        s = self.arr(s)
        h, w = s.shape
        y = self.arr(y0).copy()
        chi = self.cbuf(chi0)
        nn = len(neighbours)
        if nn:
            sw = np.ascontiguousarray(np.stack([self.arr(n[0]) for n in neighbours]))
            cw = np.ascontiguousarray(np.stack([self.cbuf(n[1]) for n in neighbours]))
            cf = np.ascontiguousarray(np.stack([self.arr(n[2]) for n in neighbours]))
        else:
            sw = cw = cf = np.zeros(1, self.dt)
        if anchor is not None:
            al, ac, aw = self.arr(anchor[0]), self.cbuf(anchor[1]), self.arr(anchor[2])
        iters = np.zeros(max(cfg.irls_outer, 1), np.int32)
        fn = self.f("joint_refine")
        fn.restype = self.R
        aniso = fn(_p(s), _p(self.cbuf(carrier)), h, w, C.byref(cfg), _p(y), _p(chi), nn, _p(sw), _p(cw), _p(cf),
                   self.R(nu), self.R(eps_t), _p(al) if anchor is not None else None,
                   _p(ac) if anchor is not None else None, _p(aw) if anchor is not None else None, self.R(nu_a),
                   _p(iters))
        return y, self.uncbuf(chi), float(aniso), iters

    def block_match_sad_u16(self, ref, cur, bh=8, bw=16, sy=4, sx=16):
        ref = np.ascontiguousarray(ref, np.uint16)
        cur = np.ascontiguousarray(cur, np.uint16)
        h, w = ref.shape
        nby, nbx = -(-h // bh), -(-w // bw)
        vy = np.zeros((nby, nbx), np.int16)
        vx = np.zeros((nby, nbx), np.int16)
        sad = np.zeros((nby, nbx), np.int32)
        self.f("block_match_sad_u16")(_p(ref), _p(cur), h, w, bh, bw, sy, sx, _p(vy), _p(vx), _p(sad))
        return vy, vx, sad


class Reference:
    """hvd-core's own engine code (FFT-free units), FP32 as it ships."""

    def __init__(self, openmp: bool = False):
        path = REFLIB_OMP if openmp else REFLIB
        if not path.exists():
            raise FileNotFoundError(str(path))
        # lazy binding: engine/sequence.cpp's DecodeFieldWindow names
        # HolographicInit, which needs FFTW and is not in the library; no
        # function exported here reaches it
        self.lib = C.CDLL(str(path), mode=os.RTLD_LAZY)
        self.lib.ref_burst_amplitude_ire.restype = C.c_float
        self.lib.ref_acc_gain.restype = C.c_float

    @staticmethod
    def available() -> bool:
        return REFLIB.exists()

    @staticmethod
    def f32(a):
        return np.ascontiguousarray(a, np.float32)

    def burst_lockin_phase(self, field, g):
        field = self.f32(field)
        th = np.zeros(g.field_height, np.float32)
        self.lib.ref_burst_lockin_phase(_p(field), C.byref(Geom.of(g)), _p(th))
        return th

    def burst_lockin_phase_pal(self, field, g):
        field = self.f32(field)
        th = np.zeros(g.field_height, np.float32)
        vs = np.zeros(g.field_height, np.int8)
        self.lib.ref_burst_lockin_phase_pal(_p(field), C.byref(Geom.of(g)), _p(th), _p(vs))
        return th, vs

    def burst_amplitude_ire(self, field, g):
        return float(self.lib.ref_burst_amplitude_ire(_p(self.f32(field)), C.byref(Geom.of(g))))

    def acc_gain(self, measured, nominal):
        return float(self.lib.ref_acc_gain(C.c_float(measured), C.c_float(nominal)))

    def tridiag_smooth(self, d, a, lam):
        d, a = self.f32(d), self.f32(a)
        x = np.zeros_like(d)
        self.lib.ref_tridiag_smooth(_p(d), _p(a), len(d), C.c_float(lam), _p(x))
        return x

    def make_carrier(self, theta_active, vswitch_active, g):
        th = self.f32(theta_active)
        out = np.zeros((len(th), 2 * g.width), np.float32)
        vs = None if vswitch_active is None else np.ascontiguousarray(vswitch_active, np.int8)
        self.lib.ref_make_carrier(_p(th), None if vs is None else _p(vs), len(th), C.byref(Geom.of(g)), _p(out))
        return out[:, 0::2] + 1j * out[:, 1::2]

    def real_op(self, op, a):
        a = self.f32(a)
        d = np.zeros_like(a)
        self.lib.ref_real_op(op, _p(a), a.shape[0], a.shape[1], _p(d))
        return d

    def cplx_op(self, op, a):
        a = np.asarray(a)
        buf = np.zeros(a.shape[:-1] + (2 * a.shape[-1],), np.float32)
        buf[..., 0::2], buf[..., 1::2] = a.real, a.imag
        out = np.zeros_like(buf)
        self.lib.ref_cplx_op(op, _p(buf), a.shape[0], a.shape[1], _p(out))
        return out[..., 0::2] + 1j * out[..., 1::2]

    def variational_refine(self, s, carrier, chi0, cfg: Cfg):
        s = self.f32(s)
        h, w = s.shape
        cb = np.zeros((h, 2 * w), np.float32)
        cb[:, 0::2], cb[:, 1::2] = carrier.real, carrier.imag
        chi = np.zeros((h, 2 * w), np.float32)
        chi[:, 0::2], chi[:, 1::2] = chi0.real, chi0.imag
        y = np.zeros_like(s)
        self.lib.ref_variational_refine(_p(s), _p(cb), h, w, C.byref(cfg), _p(chi), _p(y))
        return y, chi[:, 0::2] + 1j * chi[:, 1::2]

    def variational_refine_nbr(self, s, carrier, chi0, cfg: Cfg, neighbours, nu, eps_t):
        s = self.f32(s)
        h, w = s.shape

        def cbuf(z):
            b = np.zeros((h, 2 * w), np.float32)
            b[:, 0::2], b[:, 1::2] = z.real, z.imag
            return b
        cb, chi = cbuf(carrier), cbuf(chi0)
        nn = len(neighbours)
        sw = np.ascontiguousarray(np.stack([self.f32(n[0]) for n in neighbours]), np.float32)
        cw = np.ascontiguousarray(np.stack([cbuf(n[1]) for n in neighbours]), np.float32)
        cf = np.ascontiguousarray(np.stack([self.f32(n[2]) for n in neighbours]), np.float32)
        y = np.zeros_like(s)
        self.lib.ref_variational_refine_nbr(_p(s), _p(cb), h, w, C.byref(cfg), nn, _p(sw), _p(cw), _p(cf),
                                            C.c_float(nu), C.c_float(eps_t), _p(chi), _p(y))
        return y, chi[:, 0::2] + 1j * chi[:, 1::2]

    @staticmethod
    def cbuf32(a):
        a = np.asarray(a)
        b = np.zeros((a.shape[0], 2 * a.shape[1]), np.float32)
        b[:, 0::2], b[:, 1::2] = a.real, a.imag
        return b

    def vectors_per_pixel(self, mdy, mdx, tile, h, w):
        mdy, mdx = self.f32(mdy), self.f32(mdx)
        dy, dx = np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)
        self.lib.ref_vectors_per_pixel(_p(mdy), _p(mdx), mdy.shape[0], mdy.shape[1], tile, h, w, _p(dy), _p(dx))
        return dy, dx

    def warp_by_tiles(self, a, mdy, mdx, tile):
        """hvd-core's WarpByTiles[Complex] from the TILE motion field."""
        cplx = np.iscomplexobj(a)
        h, w = a.shape
        src = self.cbuf32(a) if cplx else self.f32(a)
        out = np.zeros_like(src)
        mdy, mdx = self.f32(mdy), self.f32(mdx)
        self.lib.ref_warp_by_tiles(_p(src), h, w, 2 if cplx else 1, _p(mdy), _p(mdx), mdy.shape[0], mdy.shape[1],
                                   tile, _p(out))
        return out[:, 0::2] + 1j * out[:, 1::2] if cplx else out

    def warp_bilinear(self, a, mdy, mdx, tile, row_offset, out_h, out_w):
        a, mdy, mdx = self.f32(a), self.f32(mdy), self.f32(mdx)
        out = np.zeros((out_h, out_w), np.float32)
        self.lib.ref_warp_bilinear(_p(a), a.shape[0], a.shape[1], _p(mdy), _p(mdx), mdy.shape[0], mdy.shape[1], tile,
                                   C.c_float(row_offset), out_h, out_w, _p(out))
        return out

    def upsample_confidence(self, conf, tile, h, w, fast=False):
        conf = self.f32(conf)
        out = np.zeros((h, w), np.float32)
        self.lib.ref_upsample_confidence(_p(conf), conf.shape[0], conf.shape[1], tile, h, w, int(fast), _p(out))
        return out

    def complex_coherence(self, z1, z2, r):
        h, w = z1.shape
        out = np.zeros((h, w), np.float32)
        self.lib.ref_complex_coherence(_p(self.cbuf32(z1)), _p(self.cbuf32(z2)), h, w, r, _p(out))
        return out

    def motion_compensate_prev(self, comp, car, mdy, mdx, mconf, tile, fast=False):
        comp, mdy, mdx, mconf = self.f32(comp), self.f32(mdy), self.f32(mdx), self.f32(mconf)
        h, w = comp.shape
        oc, ocar, ocf = np.zeros((h, w), np.float32), np.zeros((h, 2 * w), np.float32), np.zeros((h, w), np.float32)
        self.lib.ref_motion_compensate_prev(_p(comp), _p(self.cbuf32(car)), h, w, _p(mdy), _p(mdx), _p(mconf), mdy.shape[0],
                                            mdy.shape[1], tile, int(fast), _p(oc), _p(ocar), _p(ocf))
        return oc, ocar[:, 0::2] + 1j * ocar[:, 1::2], ocf

    def box_blur(self, a, r):
        a = self.f32(a).copy()
        self.lib.ref_box_blur(_p(a), a.shape[0], a.shape[1], r)
        return a

    def estimate_motion(self, y_ref, y_cur, tile=32, search=8):
        y_ref, y_cur = self.f32(y_ref), self.f32(y_cur)
        h, w = y_ref.shape
        th, tw = -(-h // tile), -(-w // tile)
        dy, dx, cf = (np.zeros((th, tw), np.float32) for _ in range(3))
        self.lib.ref_estimate_motion(_p(y_ref), _p(y_cur), h, w, tile, search, _p(dy), _p(dx), _p(cf))
        return dy, dx, cf


def _ref_extras():
    """Methods of Reference for the window driver's pieces."""

    def estimate_noise_ire(self, s):
        s = self.f32(s)
        self.lib.ref_estimate_noise_ire.restype = C.c_float
        return float(self.lib.ref_estimate_noise_ire(_p(s), s.shape[0], s.shape[1]))

    def verify_motion(self, y_ref, y_cur, pdy, pdx, tile=32):
        y_ref, y_cur, pdy, pdx = self.f32(y_ref), self.f32(y_cur), self.f32(pdy), self.f32(pdx)
        h, w = y_ref.shape
        dy, dx, cf = (np.zeros_like(pdy) for _ in range(3))
        self.lib.ref_verify_motion(_p(y_ref), _p(y_cur), h, w, _p(pdy), _p(pdx), pdy.shape[0], pdy.shape[1], tile,
                                   _p(dy), _p(dx), _p(cf))
        return dy, dx, cf

    def trajectory_snap(self, offsets, pshift, dy, dx, conf):
        dy, dx, conf = self.f32(dy).copy(), self.f32(dx).copy(), self.f32(conf)
        self.lib.ref_trajectory_snap(dy.shape[0], _p(np.ascontiguousarray(offsets, np.int32)), _p(self.f32(pshift)),
                                     _p(dy), _p(dx), _p(conf), dy.shape[1], dy.shape[2])
        return dy, dx

    def joint_refine(self, s, carrier, y0, chi0, cfg: Cfg, neighbours=(), nu=0.0, eps_t=1.0, anchor=None,
                     nu_a=0.0):
        s = self.f32(s)
        h, w = s.shape
        y = self.f32(y0).copy()
        chi = self.cbuf32(chi0)
        nn = len(neighbours)
        if nn:
            sw = np.ascontiguousarray(np.stack([self.f32(n[0]) for n in neighbours]))
            cw = np.ascontiguousarray(np.stack([self.cbuf32(n[1]) for n in neighbours]))
            cf = np.ascontiguousarray(np.stack([self.f32(n[2]) for n in neighbours]))
        else:
            sw = cw = cf = np.zeros(1, np.float32)
        if anchor is not None:
            al, ac, aw = self.f32(anchor[0]), self.cbuf32(anchor[1]), self.f32(anchor[2])
        self.lib.ref_joint_refine(_p(s), _p(self.cbuf32(carrier)), h, w, C.byref(cfg), nn, _p(sw), _p(cw), _p(cf),
                                  C.c_float(nu), C.c_float(eps_t), _p(al) if anchor is not None else None,
                                  _p(ac) if anchor is not None else None, _p(aw) if anchor is not None else None,
                                  C.c_float(nu_a), _p(y), _p(chi))
        return y, chi[:, 0::2] + 1j * chi[:, 1::2]

    def decode_window_with_inits(self, fields, parities, inits_chi, cfg: "SeqCfg", g=None, inits_y=None):
        """hvd-core's DecodeFieldWindowWithInits: fields = [(S, carrier)],
        -> (Y planes, chi planes, diagnostics dict)."""
        nf = len(fields)
        h, w = fields[0][0].shape
        s = np.ascontiguousarray(np.stack([self.f32(f[0]) for f in fields]))
        car = np.ascontiguousarray(np.stack([self.cbuf32(f[1]) for f in fields]))
        chi0 = np.ascontiguousarray(np.stack([self.cbuf32(c) for c in inits_chi]))
        y0 = None if inits_y is None else np.ascontiguousarray(np.stack([self.f32(y) for y in inits_y]))
        par = np.ascontiguousarray(parities, np.int32)
        oy = np.zeros((nf, h, w), np.float32)
        oc = np.zeros((nf, h, 2 * w), np.float32)
        diag = np.zeros(5, np.float32)
        geom = Geom() if g is None else Geom.of(g)
        n = self.lib.ref_decode_window_with_inits(nf, h, w, _p(s), _p(car), _p(par), _p(chi0),
                                                  None if y0 is None else _p(y0), C.byref(geom), C.byref(cfg),
                                                  _p(oy), _p(oc), _p(diag))
        assert n == nf
        keys = ("resolved_strength", "ambiguity_ire", "sigma_ire", "temporal_eps", "nr_eps")
        return oy, oc[..., 0::2] + 1j * oc[..., 1::2], dict(zip(keys, diag.tolist()))

    def drizzle_frame(self, j0, ys, chis, parities, mc_tile=32, mc_search=16, nr_radius=2, nr_eps=0.0, scale=2):
        """hvd-core's DrizzleFrame: (luma, chroma) on the (2 h scale) x w grid."""
        n = len(ys)
        h, w = ys[0].shape
        y = np.ascontiguousarray(np.stack([self.f32(a) for a in ys]))
        c = np.ascontiguousarray(np.stack([self.cbuf32(a) for a in chis]))
        par = np.ascontiguousarray(parities, np.int32)
        oy = np.zeros((2 * h * scale, w), np.float32)
        oc = np.zeros((2 * h * scale, 2 * w), np.float32)
        self.lib.ref_drizzle_frame(j0, n, h, w, _p(y), _p(c), _p(par), mc_tile, mc_search, nr_radius,
                                   C.c_float(nr_eps), scale, _p(oy), _p(oc))
        return oy, oc[..., 0::2] + 1j * oc[..., 1::2]

    for fn in (estimate_noise_ire, verify_motion, trajectory_snap, joint_refine, decode_window_with_inits,
               drizzle_frame):
        setattr(Reference, fn.__name__, fn)


_ref_extras()


def threads_hint() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "1"))
