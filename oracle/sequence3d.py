"""Oracle of hvd-core's window driver with its 3-D machinery:
DecodeFieldWindowWithInits (hvd-core/src/engine/sequence.cpp:371-1114) with
SynthReference (:52-150), PsiClosedForm (:160-226) and DrizzleFrame
(:285-369).

TEST INFRASTRUCTURE ONLY.  The per-plane operators come from the C oracle
(liboracle32: EstimateMotion, VerifyMotion, TrajectorySnap, the warps,
MotionCompensatePrev, ComplexCoherence, BoxBlur, EstimateNoiseIre and the two
solvers, each bit-exact with hvd-core); the driver's own per-pixel and
per-band arithmetic is restated here in numpy float32 in hvd-core's order of
operations (std::complex<float> products written out; FP64 where hvd-core
accumulates in double), so that on the FP32 build the whole window is
hvd-core's result bit for bit (checked against hvd-core compiled from source
in tests/test_oracle_sequence3d.py).
"""
from __future__ import annotations

import dataclasses

import numpy as np

import oracle as O

F = np.float32


def f32(x):
    return np.float32(x)


def median_np(v):
    """MedianOf / Median (sequence.cpp:28-38, motion.cpp:80-91): numpy median
    in float (the mean of the two middle elements for even counts)."""
    v = np.sort(np.asarray(v, F).ravel())
    n = len(v)
    if n == 0:
        return F(0)
    m = v[n // 2]
    if n % 2 == 0:
        m = F(0.5) * (m + v[n // 2 - 1])
    return F(m)


def cmul_re(a, b):
    """Re[a b] of complex64 arrays as std::complex<float> forms it."""
    return (a.real.astype(F) * b.real.astype(F) - a.imag.astype(F) * b.imag.astype(F)).astype(F)


def cabs(z):
    return np.hypot(z.real.astype(F), z.imag.astype(F)).astype(F)


@dataclasses.dataclass
class WindowCfg:
    """HvdConfig fields the driver reads (engine/hvd_config.h defaults)."""
    lambda_c: float = 1.0
    chroma_aniso: float = 0.0
    charbonnier_eps: float = 0.5
    chroma_eps: float = 1.0
    structure_coupling: float = 0.25
    diag_prior: float = 0.0
    cg_iterations: int = 2
    irls_outer: int = 4
    cg_tol: float = 0.0
    fast: bool = True
    temporal_strength: float = 0.0
    temporal_eps: float = 0.0
    nr_anchor: float = 1.0
    nr_eps: float = 0.0
    nr_radius: int = 2
    passes: int = 2
    mc_tile: int = 32
    mc_search: int = 16
    bidirectional: bool = True
    extended_temporal: bool = True
    trajectory_fit: bool = True
    coherence_gate: float = 0.6
    odd_gate_floor: float = 0.35
    psi_init: bool = False
    output_fidelity: bool = True
    parallel_across_fields: bool = True

    def solver(self, is_pal):
        return O.Cfg.make(lambda_c=self.lambda_c, chroma_aniso=self.chroma_aniso,
                          charbonnier_eps=self.charbonnier_eps, chroma_eps=self.chroma_eps,
                          structure_coupling=self.structure_coupling, diag_prior=self.diag_prior,
                          cg_iterations=self.cg_iterations, irls_outer=self.irls_outer, cg_tol=self.cg_tol,
                          fast=int(self.fast), is_pal=int(is_pal))

    def seq(self, is_pal=False):
        """The same configuration for hvd-core's own driver (oracle.SeqCfg)."""
        d = dataclasses.asdict(self)
        for k in ("fast", "bidirectional", "extended_temporal", "trajectory_fit", "psi_init", "output_fidelity",
                  "parallel_across_fields"):
            d[k] = int(d[k])
        return O.SeqCfg.make(is_pal=int(is_pal), **d)


@dataclasses.dataclass
class Motion:
    dy: np.ndarray
    dx: np.ndarray
    conf: np.ndarray


# ---------------------------------------------------------------- per-window pieces
def ambiguity(fields, stride, nr_eps):
    """The adaptive-strength scan (sequence.cpp:526-652): per same-parity pair
    (j, j + stride) the triangle-7 demod difference energy over decimated
    static samples, p95 of the 8-row bands, median over pairs."""
    kw = np.array([1, 2, 3, 4, 3, 2, 1], F)
    ambs = []
    for j in range(len(fields) - stride):
        Sj, cj = fields[j][0], fields[j][1]
        Sk, ck = fields[j + stride][0], fields[j + stride][1]
        hh, ww = Sj.shape
        ys = np.arange(2, hh - 2, 3)
        xs = np.arange(4, ww - 7, 4)
        if len(ys) == 0 or len(xs) == 0:
            continue
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        mj = np.zeros(Y.shape, F)
        mk = np.zeros(Y.shape, F)
        for k4 in range(4):
            mj = (mj + Sj[Y, X + k4]).astype(F)
            mk = (mk + Sk[Y, X + k4]).astype(F)
        keep = ~(np.abs(mj - mk).astype(F) * F(0.25) > F(nr_eps))

        def demod(S, car):
            dr = np.zeros(Y.shape, F)
            di = np.zeros(Y.shape, F)
            for k in range(7):
                a = (kw[k] * S[Y, X + k]).astype(F)
                c = car[Y, X + k]
                dr = (dr + a * c.real.astype(F)).astype(F)
                di = (di + a * (-c.imag.astype(F))).astype(F)
            return dr * F(1.0 / 16.0), di * F(1.0 / 16.0)

        djr, dji = demod(Sj, cj)
        dkr, dki = demod(Sk, ck)
        er, ei = (djr - dkr).astype(F), (dji - dki).astype(F)
        e = 0.25 * (er * er + ei * ei).astype(F).astype(np.float64)
        acc, n = 0.0, 0
        band_acc, band_n = [], []
        for yi in range(len(ys)):
            band = yi // 8
            if band >= len(band_acc):
                band_acc.append(0.0)
                band_n.append(0)
            for xi in range(len(xs)):
                if not keep[yi, xi]:
                    continue
                acc += e[yi, xi]
                n += 1
                band_acc[band] += e[yi, xi]
                band_n[band] += 1
        if n > 0:
            brms = [F(band_acc[b] / band_n[b]) for b in range(len(band_acc)) if band_n[b] >= 8]
            if brms:
                k95 = min(len(brms) - 1, int(0.95 * len(brms)))
                ambs.append(np.sqrt(np.sort(np.array(brms, F))[k95]).astype(F))
            else:
                ambs.append(np.sqrt(F(acc / n)).astype(F))
    return median_np(ambs)


def smooth_rows(src):
    """The envelope's horizontal 5-tap sliding RMS (sequence.cpp:855-888)."""
    hh, ww = src.shape
    out = np.zeros((hh, ww), F)
    acc = np.zeros(hh, F)
    for x in range(min(3, ww)):
        acc = (acc + src[:, x]).astype(F)
    for x in range(ww):
        lo, hi = x - 2, x + 2
        if x > 0:
            if hi < ww:
                acc = (acc + src[:, hi]).astype(F)
            if lo - 1 >= 0:
                acc = (acc - src[:, lo - 1]).astype(F)
        n_taps = min(hi, ww - 1) - max(lo, 0) + 1
        out[:, x] = np.sqrt((acc / F(n_taps)).astype(F))
    return out


def envelope(Y, chi):
    """vgrad and lapsm of the odd-offset envelope gate (sequence.cpp:
    790-940): one-sided max vertical difference and second difference of
    the decoded baseband, row-smoothed, noise-floored by 1.5 x the median."""
    hh, ww = Y.shape
    idx = np.arange(hh)
    ym, yp = np.maximum(idx - 1, 0), np.minimum(idx + 1, hh - 1)
    dyu = (Y - Y[ym]).astype(F)
    dyd = (Y[yp] - Y).astype(F)
    dcu = cabs(chi - chi[ym])
    dcd = cabs(chi[yp] - chi)
    mu = (dyu * dyu + dcu * dcu).astype(F)
    md = (dyd * dyd + dcd * dcd).astype(F)
    gmag = np.maximum(mu, md)
    ly = (Y - F(0.5) * (Y[ym] + Y[yp]).astype(F)).astype(F)
    half = (chi[ym] + chi[yp]).astype(np.complex64)
    half = (F(0.5) * half.real.astype(F) + 1j * (F(0.5) * half.imag.astype(F))).astype(np.complex64)
    lc = cabs((chi - half).astype(np.complex64))
    lapmag = (ly * ly + lc * lc).astype(F)
    vgrad = smooth_rows(gmag)
    med = median_np(vgrad.ravel()[::7])
    fl2 = F(2.25) * med * med
    vgrad = np.sqrt(np.maximum(F(0), (vgrad * vgrad - fl2).astype(F))).astype(F)
    t_raw = smooth_rows(lapmag)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(vgrad > F(1e-6), np.clip((t_raw / vgrad).astype(F), F(0), F(1)), F(0)).astype(F)
    lapsm = (F(1) - t * t).astype(F)
    return vgrad, lapsm


def synth_reference(orc, j, Ys, chis, fields, parities, cfg, eps_b, mcache):
    """SynthReference (sequence.cpp:52-150): the motion-compensated robust
    blend of the decoded neighbours within nr_radius, and its anchor weight
    from the re-encoded residual."""
    S, car = fields[j][0], fields[j][1]
    hh, ww = Ys[j].shape
    tile = cfg.mc_tile
    accY = Ys[j].astype(F).copy()
    accCr = chis[j].real.astype(F).copy()
    accCi = chis[j].imag.astype(F).copy()
    accW = np.ones((hh, ww), F)
    e2 = F(eps_b) * F(eps_b)
    for k in range(j - cfg.nr_radius, j + cfg.nr_radius + 1):
        if k == j or k < 0 or k >= len(Ys):
            continue
        if mcache is not None and (j, k) in mcache:
            mo = mcache[(j, k)]
        else:
            dy, dx, cf = orc.estimate_motion(Ys[k], Ys[j], tile, cfg.mc_search)
            mo = Motion(dy, dx, cf)
            if mcache is not None:
                mcache[(j, k)] = mo
        conf = orc.upsample_confidence(mo.conf, tile, hh, ww)
        row_off = (F(parities[j]) - F(parities[k])) / F(2)
        vdy, vdx = orc.vectors_per_pixel(mo.dy, mo.dx, tile, hh, ww)
        Yw = orc.warp_bilinear(Ys[k], vdy, vdx, row_off)
        wre = orc.warp_bilinear(chis[k].real.astype(F), vdy, vdx, row_off)
        wim = orc.warp_bilinear(chis[k].imag.astype(F), vdy, vdx, row_off)
        d = (Yw - Ys[j]).astype(F)
        cr = (wre - chis[j].real.astype(F)).astype(F)
        ci = (wim - chis[j].imag.astype(F)).astype(F)
        d2 = (d * d + (cr * cr + ci * ci).astype(F)).astype(F)
        w = ((conf * F(eps_b)).astype(F) * F(eps_b)).astype(F) / (d2 + e2).astype(F)
        w = w.astype(F)
        accY = (accY + w * Yw).astype(F)
        accCr = (accCr + w * wre).astype(F)
        accCi = (accCi + w * wim).astype(F)
        accW = (accW + w).astype(F)
    luma = (accY / accW).astype(F)
    chroma = ((accCr / accW).astype(F) + 1j * (accCi / accW).astype(F)).astype(np.complex64)
    s_hat = (luma + cmul_re(chroma, car)).astype(F)
    resid = np.abs((S - s_hat).astype(F)).astype(F)
    r = orc.box_blur(resid, 1)
    weight = (e2 / (r * r + e2).astype(F)).astype(F)
    return luma, chroma, weight


def psi_closed_form(S, car, nbrs, chi_fb):
    """PsiClosedForm (sequence.cpp:160-226): the per-pixel 3 x 3 weighted
    normal equations over (own, neighbour) observations, solved by Cramer,
    blended with the fallback by the total neighbour weight."""
    kEpsT = F(6.0)
    rc0, ic0 = car.real.astype(np.float64), -car.imag.astype(np.float64)
    S64 = S.astype(np.float64)
    a11 = np.ones_like(S64)
    a12, a13 = rc0.copy(), ic0.copy()
    a22, a23, a33 = rc0 * rc0, rc0 * ic0, ic0 * ic0
    b1, b2, b3 = S64.copy(), S64 * rc0, S64 * ic0
    wsum = np.zeros_like(S64)
    for sw, cw, cf in nbrs:
        dcr = (car.real.astype(F) - cw.real.astype(F)).astype(F)
        dci = (car.imag.astype(F) - cw.imag.astype(F)).astype(F)
        dc = (dcr + 1j * dci).astype(np.complex64)
        rt = ((sw - S).astype(F) + cmul_re(chi_fb.astype(np.complex64), dc)).astype(F)
        wk = ((cf * kEpsT).astype(F) * kEpsT).astype(F) / (rt * rt + kEpsT * kEpsT).astype(F)
        wsk = (wk.astype(F) * sw.astype(F)).astype(F).astype(np.float64)  # wk * Sk in float
        wk = wk.astype(F).astype(np.float64)
        rc, ic = cw.real.astype(np.float64), -cw.imag.astype(np.float64)
        a11 += wk
        a12 += wk * rc
        a13 += wk * ic
        a22 += wk * rc * rc
        a23 += wk * rc * ic
        a33 += wk * ic * ic
        b1 += wsk
        b2 += wsk * rc
        b3 += wsk * ic
        wsum += wk
    det = a11 * (a22 * a33 - a23 * a23) - a12 * (a12 * a33 - a23 * a13) + a13 * (a12 * a23 - a22 * a13)
    a11c = np.maximum(a11, 1e-9)
    ok = det > 1e-3 * a11c * a11c * a11c * 0.01
    d = np.where(ok, det, 1.0)
    q1 = (a11 * (b2 * a33 - b3 * a23) - b1 * (a12 * a33 - a23 * a13) + a13 * (a12 * b3 - b2 * a13)) / d
    q2 = (a11 * (a22 * b3 - a23 * b2) - a12 * (a12 * b3 - b2 * a13) + b1 * (a12 * a23 - a22 * a13)) / d
    t = np.where(ok, np.clip((wsum / 2.0).astype(F), F(0), F(1)), F(0)).astype(F)
    pr, pi = q1.astype(F), q2.astype(F)
    one_t = (F(1) - t).astype(F)
    outr = ((t * pr).astype(F) + (one_t * chi_fb.real.astype(F)).astype(F)).astype(F)
    outi = ((t * pi).astype(F) + (one_t * chi_fb.imag.astype(F)).astype(F)).astype(F)
    return (outr + 1j * outi).astype(np.complex64)


# ---------------------------------------------------------------- the driver
def decode_window_with_inits(orc, fields, parities, inits_chi, cfg: WindowCfg, is_pal=False, inits_y=None,
                             diag=None):
    """DecodeFieldWindowWithInits (sequence.cpp:371-1114).  fields: [(S,
    carrier)] float32 / complex64 planes; returns (Y list, chi list)."""
    nf = len(fields)
    fields = [(np.asarray(s, F), np.asarray(c, np.complex64)) for s, c in fields]
    c = dataclasses.replace(cfg)
    # self-calibration (:439-449)
    sigma = median_np([orc.estimate_noise_ire(s) for s, _ in fields])
    if c.temporal_eps <= 0:
        c.temporal_eps = float(np.clip(F(7) * sigma, F(4), F(20)))
    if c.nr_eps <= 0:
        c.nr_eps = float(np.clip(F(3) * sigma, F(3), F(12)))
    if c.extended_temporal:
        offs = ([-4, -3, -2, -1, 1, 2, 3, 4] if c.bidirectional else [-4, -3, -2, -1]) if is_pal else \
               ([-3, -2, -1, 1, 2, 3] if c.bidirectional else [-3, -2, -1])
    else:
        offs = [-2, -1, 1, 2] if c.bidirectional else [-2, -1]
    if is_pal:
        if F(c.coherence_gate) == F(0.6):
            c.coherence_gate = 1.0
        if c.nr_radius == 2:
            c.nr_radius = 4
    Ys, chis = [], []
    for j in range(nf):
        chis.append(np.asarray(inits_chi[j], np.complex64).copy())
        if inits_y is not None and inits_y[j] is not None:
            Ys.append(np.asarray(inits_y[j], F).copy())
        else:
            Ys.append((fields[j][0] - cmul_re(chis[j], fields[j][1])).astype(F))
    mcache_fast = {}
    amb_stride = 4 if is_pal else 2
    amb = F(0)
    if F(c.temporal_strength) == 0 and c.cg_iterations > 0 and nf > amb_stride:
        amb = ambiguity(fields, amb_stride, F(c.nr_eps))
        sigma_d = F(c.nr_eps) / F(3) if F(c.nr_eps) > 0 else F(1)
        floor = sigma_d * F(0.29)
        amb = np.sqrt(np.maximum(F(0), (amb * amb - floor * floor).astype(F))).astype(F)
        c.temporal_strength = float(np.clip((F(0.5) * amb - F(0.05)).astype(F), F(0), F(1.5)))
    if F(c.temporal_strength) < 0:
        c.temporal_strength = 0.0
    if diag is not None:
        diag.update(resolved_strength=max(0.0, float(F(c.temporal_strength))), ambiguity_ire=float(amb),
                    sigma_ire=float(sigma), temporal_eps=float(F(c.temporal_eps)), nr_eps=float(F(c.nr_eps)))
    has_temporal = F(c.temporal_strength) > 0 and c.cg_iterations > 0
    n_passes = max(1, c.passes) if has_temporal else 1
    scfg = c.solver(is_pal)
    nu, eps_t = F(c.temporal_strength), F(c.temporal_eps)
    tile = c.mc_tile

    for pss in range(n_passes):
        mcache = mcache_fast if c.fast else {}

        def solve_field(j):
            S, car = fields[j]
            pj = parities[j]
            pair = {}
            if has_temporal:
                full_os = [-1, 1, 2] if c.fast else offs
                for o in offs:
                    k = j + o
                    if k < 0 or k >= nf:
                        continue
                    if (j, k) in mcache:
                        pair[o] = mcache[(j, k)]
                    elif not c.trajectory_fit or o in full_os:
                        dy, dx, cf = orc.estimate_motion(Ys[k], Ys[j], tile, c.mc_search)
                        mcache[(j, k)] = pair[o] = Motion(dy, dx, cf)
                if c.fast and c.trajectory_fit and pair:
                    keys = sorted(pair)
                    sy = np.stack([(pair[o].dy - (F(parities[j + o]) - F(pj)) / F(2)).astype(F) / F(o)
                                   for o in keys]).astype(F)
                    sx = np.stack([(pair[o].dx / F(o)).astype(F) for o in keys])
                    th, tw = pair[keys[0]].dy.shape
                    vym0 = np.zeros((th, tw), F)
                    vxm0 = np.zeros((th, tw), F)
                    for i in range(th * tw):
                        yy, xx = divmod(i, tw)
                        vym0[yy, xx] = median_np(sy[:, yy, xx])
                        vxm0[yy, xx] = median_np(sx[:, yy, xx])
                    for o in offs:
                        k = j + o
                        if o in pair or k < 0 or k >= nf:
                            continue
                        if (j, k) in mcache:
                            pair[o] = mcache[(j, k)]
                            continue
                        ho = (F(parities[k]) - F(pj)) / F(2)
                        pdy = (F(o) * vym0 + ho).astype(F)
                        pdx = (F(o) * vxm0).astype(F)
                        dy, dx, cf = orc.verify_motion(Ys[k], Ys[j], pdy, pdx, tile)
                        mcache[(j, k)] = pair[o] = Motion(dy, dx, cf)
                if c.trajectory_fit and len(pair) >= 2:
                    keys = sorted(pair)
                    ps = np.array([(F(parities[j + o]) - F(pj)) / F(2) for o in keys], F)
                    dy, dx = orc.trajectory_snap(keys, ps, np.stack([pair[o].dy for o in keys]),
                                                 np.stack([pair[o].dx for o in keys]),
                                                 np.stack([pair[o].conf for o in keys]))
                    for n_, o in enumerate(keys):
                        pair[o] = Motion(dy[n_], dx[n_], pair[o].conf)
            nbrs = []
            vgrad = lapsm = None
            if has_temporal:
                if any(0 <= j + o < nf and (o & 1) for o in offs):
                    vgrad, lapsm = envelope(Ys[j], chis[j])
                hh, ww = Ys[j].shape
                for o in offs:
                    k = j + o
                    if k < 0 or k >= nf:
                        continue
                    mo = pair[o]
                    gate_on = F(c.coherence_gate) > 0 and not (c.fast and pss == 0)
                    comp, cw, conf = orc.motion_compensate_prev(fields[k][0], fields[k][1], mo.dy, mo.dx, mo.conf,
                                                                tile, c.fast)
                    comp = comp.astype(F)
                    cw = cw.astype(np.complex64)
                    conf = conf.astype(F)
                    if gate_on:
                        vdy, vdx = orc.vectors_per_pixel(mo.dy, mo.dx, tile, hh, ww)
                        wre = orc.warp_bilinear(chis[k].real.astype(F), vdy, vdx, 0.0)
                        wim = orc.warp_bilinear(chis[k].imag.astype(F), vdy, vdx, 0.0)
                        gamma = orc.complex_coherence(chis[j], (wre + 1j * wim).astype(np.complex64), 6).astype(F)
                        a = F(c.coherence_gate)
                        conf = (conf * ((F(1) - a) + (a * gamma).astype(F)).astype(F)).astype(F)
                    if (o & 1) and vgrad is not None:
                        e2 = F(c.temporal_eps) * F(c.temporal_eps)
                        fmax = np.clip(F(c.odd_gate_floor), F(0), F(1))
                        floor_i = (fmax * lapsm).astype(F)
                        g = (e2 / (e2 + (vgrad * vgrad).astype(F)).astype(F)).astype(F)
                        conf = (conf * np.maximum(floor_i, g)).astype(F)
                    nbrs.append((comp, cw, conf))
            if pss == 0 and c.psi_init and nbrs:
                chis[j] = psi_closed_form(S, car, nbrs, chis[j])
            if pss >= 1 and F(c.nr_anchor) > 0 and nf > 1:
                anchor = synth_reference(orc, j, Ys, chis, fields, parities, c, F(c.nr_eps),
                                         mcache_fast if c.fast else None)
                y, chi, _, _ = orc.joint_refine(S, car, Ys[j], chis[j], scfg, nbrs, nu, eps_t, anchor,
                                                F(c.nr_anchor))
            else:
                y, chi, _, _ = orc.variational_refine_nbr(S, car, chis[j], scfg, nbrs, nu, eps_t) \
                    if nbrs else orc.variational_refine(S, car, chis[j], scfg)
            Ys[j] = np.asarray(y, F)
            chis[j] = np.asarray(chi).astype(np.complex64)

        if c.fast or not has_temporal:
            max_off = max(abs(o) for o in offs)
            reach = max(max_off, c.nr_radius if F(c.nr_anchor) > 0 else 0)
            stride = reach + 1 if has_temporal else 1
            if c.parallel_across_fields:
                for color in range(stride):
                    for j in range(color, nf, stride):
                        solve_field(j)
            else:
                for j in range(nf):
                    solve_field(j)
        else:
            for j in range(nf):
                solve_field(j)

    out_y, out_chi = [], []
    for j in range(nf):
        out_chi.append(chis[j])
        if c.output_fidelity:
            out_y.append((fields[j][0] - cmul_re(chis[j], fields[j][1])).astype(F))
        else:
            out_y.append(Ys[j])
    return out_y, out_chi
