"""Losses of the fit (losses_jax.py:13-80) on torch tensors."""
from __future__ import annotations

import torch

from .fee import adc2charge  # noqa: F401  (re-export, losses_jax.py:380-383)


def rbf_kernel(x, y, sigma):
    """exp(-|x - y|^2 / (2 sigma^2)) (losses_jax.py:13-18)."""
    d = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return torch.exp(-d / (2 * sigma ** 2))


def mmd(x, y, px, py, sigma):
    """Probability-weighted maximum mean discrepancy (losses_jax.py:21-39)."""
    kxx = (rbf_kernel(x, x, sigma) * px[:, None] * px[None, :]).sum()
    kyy = (rbf_kernel(y, y, sigma) * py[:, None] * py[None, :]).sum()
    kxy = (rbf_kernel(x, y, sigma) * px[:, None] * py[None, :]).sum()
    sx, sy = px.sum(), py.sum()
    return kxx / sx ** 2 + kyy / sy ** 2 - 2 * kxy / (sx * sy)


def mse_adc(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_ticks, ref_hit_prob, ref_event,
            sigma=1, lambda_Q=1):
    """MMD between weighted hit clouds plus relative total-charge error (losses_jax.py:58-80)."""
    w_ref = ref_Q * ref_hit_prob
    w = Q * hit_prob
    ref_s = torch.stack((ref_x + ref_event * 1e5, ref_y, ref_z), dim=-1)
    s = torch.stack((x + event * 1e5, y, z), dim=-1)
    m = mmd(s, ref_s, w, w_ref, sigma)
    tot_ref, tot = w_ref.sum(), w.sum()
    charge_loss = ((tot - tot_ref) / (tot_ref + 1e-6)) ** 2
    aux = dict(charge_loss=charge_loss, mmd_loss_term=m, Q=Q, ref_Q=ref_Q, ref_hit_prob=ref_hit_prob, hit_prob=hit_prob)
    return m + lambda_Q * charge_loss, aux


def mse_loss(adcs, pIDs, adcs_ref, pIDs_ref):
    """Per-pixel squared difference of (P, H) hit arrays (losses_jax.py:41-56)."""
    allp = torch.unique(torch.cat([pIDs, pIDs_ref]))
    sig = torch.zeros((allp.numel(), adcs.shape[1]), dtype=adcs.dtype, device=adcs.device)
    sig = sig.index_add(0, torch.searchsorted(allp, pIDs), adcs)
    sig = sig.index_add(0, torch.searchsorted(allp, pIDs_ref), -adcs_ref)
    return (sig ** 2).sum(), dict()


def mse_time(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref):
    """Per-pixel squared difference of the hit ticks, ticks of empty ADC slots set to 0
    (losses_jax.py:82-87).  (mse_time_adc, :89-92, calls mse_adc with the wrong number of
    arguments in the reference and cannot run there; it is not provided.)"""
    t = torch.where(adcs > 0, ticks, torch.zeros_like(ticks))
    t_ref = torch.where(ref > 0, ticks_ref, torch.zeros_like(ticks_ref))
    return mse_loss(t, pixels, t_ref, pixels_ref)


def pixel_charge_loss(params, Q, hit_pixels, ref_Q, ref_hit_pixels):
    """Per-pixel charge loss: mse_loss on the summed hit charge of each pixel."""
    return mse_loss(Q[:, None], hit_pixels, ref_Q[:, None], ref_hit_pixels)


def _event_slices(e):
    """Sorted order and [begin, end) of each event in it (one host sync)."""
    order = torch.argsort(e, stable=True)
    ev, cnt = torch.unique_consecutive(e[order], return_counts=True)
    ends = torch.cumsum(cnt, 0)
    return order, dict(zip(ev.tolist(), zip((ends - cnt).tolist(), ends.tolist())))


def _kernel_rows(a, b, inv, c):
    """exp(-|a_i - b_j|^2 inv) @ c, kernel values in float32, sums in float64."""
    d = torch.cdist(a.float(), b.float()).square_()
    return (torch.exp_(d.mul_(-inv)).double() @ c)


def _pad4(a):
    """(N, 3) -> contiguous (N, 4) float32 (fourth column zero); (N, 4) -> contiguous float32."""
    if a.shape[1] == 4:
        return a.to(torch.float32).contiguous()
    out = torch.zeros((a.shape[0], 4), dtype=torch.float32, device=a.device)
    out[:, :3] = a
    return out


def _rbf_sums_hip(xs, lo, hi, ys, c, sigma):
    """out[i] = sum_{j=lo_i}^{hi_i-1} exp(-|xs_i - ys_j|^2 / 2 sigma^2) c_j on the GPU (lsim_rbf_sums)."""
    import ctypes as C

    from . import _lib

    lib = _lib.load()
    xs4, ys4 = _pad4(xs), _pad4(ys)
    c4 = c.to(torch.float32).contiguous()
    lo, hi = lo.to(torch.int32).contiguous(), hi.to(torch.int32).contiguous()
    out = torch.empty((xs.shape[0], 4), dtype=torch.float64, device=xs.device)
    st = C.c_void_p(torch.cuda.current_stream(xs.device).cuda_stream)
    _lib.check(lib.lsim_rbf_sums(C.c_void_p(xs4.data_ptr()), xs.shape[0], C.c_void_p(lo.data_ptr()),
                                 C.c_void_p(hi.data_ptr()), C.c_void_p(ys4.data_ptr()), C.c_void_p(c4.data_ptr()),
                                 ys.shape[0], C.c_float(float(sigma)), C.c_void_p(out.data_ptr()), st))
    return out


def _cell_keys(p4, e, lo, h, G):
    """Grid key ((event * Gx + cx) * Gy + cy) * Gz + cz of each point, cells clamped to the grid."""
    cell = torch.floor((p4[:, :3] - lo) / h)
    cell = torch.minimum(torch.clamp(cell, min=0.0), G - 1).to(torch.int64)
    g = G.to(torch.int64)
    return ((e.to(torch.int64) * g[0] + cell[:, 0]) * g[1] + cell[:, 1]) * g[2] + cell[:, 2]


def _block_sums(x, ex, y, ey, c, sigma, dims=3):
    """Per point of x: kernel-weighted sums of the 4 columns of c over the points of y in the same
    event, on the GPU (lsim_rbf_sums_grid, no host synchronisation).  Both clouds go on one grid of
    cells about a third of the kernel's cull radius wide (at most 64 per axis), y sorted by
    (event, cell) for the kernel's column spans and x by the same key so each wave's points are
    close together; dims 3 ignores the fourth coordinate."""
    import ctypes as C

    from . import _lib

    out = torch.zeros((x.shape[0], 4), dtype=torch.float64, device=x.device)
    if x.shape[0] == 0 or y.shape[0] == 0:
        return out
    lib = _lib.load()
    x4, y4 = _pad4(x), _pad4(y)
    pts = torch.cat([x4[:, :3], y4[:, :3]])
    lo = pts.amin(dim=0)
    ext = pts.amax(dim=0) - lo
    rc = float(lib.lsim_rbf_cull_radius(C.c_float(float(sigma))))
    G = torch.clamp(torch.ceil(ext / (rc / 3.0)), 1.0, 64.0)
    h = torch.where(ext > 0, ext / G, torch.ones_like(ext))
    grid = torch.cat([lo, h, G]).to(torch.float32).contiguous()
    kx, ky = _cell_keys(x4, ex, lo, h, G), _cell_keys(y4, ey, lo, h, G)
    ox, oy = torch.argsort(kx, stable=True), torch.argsort(ky, stable=True)
    xs, es = x4[ox].contiguous(), ex[ox].to(torch.int32).contiguous()
    ys, cs, ks = y4[oy].contiguous(), c[oy].to(torch.float32).contiguous(), ky[oy].contiguous()
    res = torch.empty((x.shape[0], 4), dtype=torch.float64, device=x.device)
    st = C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(lib.lsim_rbf_sums_grid(C.c_void_p(xs.data_ptr()), C.c_void_p(es.data_ptr()), xs.shape[0],
                                      C.c_void_p(ys.data_ptr()), C.c_void_p(cs.data_ptr()), C.c_void_p(ks.data_ptr()),
                                      ys.shape[0], C.c_void_p(grid.data_ptr()), int(dims), C.c_float(float(sigma)),
                                      C.c_void_p(res.data_ptr()), st))
    out[ox] = res
    return out


class _BlockMMDTerms(torch.autograd.Function):
    """Kernel sums of the event-blocked RBF MMD without keeping the kernel matrices.

    For point clouds x (N, 3) with weights px and y (M, 3) with weights py, where points of
    different events never interact (mse_adc offsets x by event * 1e5 cm, so those kernel values
    are exactly 0), returns (A, C) = (sum K(x, x) px px, sum K(x, y) px py) summed within events.
    Rows are processed in chunks per event; the backward reuses the per-point sums (K px,
    K (px x), K_xy py, K_xy (py y)) kept from the forward.  Gradients flow to x and px (the
    simulated hits); the target cloud is a constant.  Kernel values are evaluated in float32
    (relative error ~1e-7 per term), sums in float64.  On the GPU the sums come from the HIP
    kernel lsim_rbf_sums (one pass, no host synchronisation); CPU tensors use chunked torch code.
    """

    @staticmethod
    def forward(ctx, x, px, y, py, ex, ey, sigma, chunk):
        dt = torch.float64
        x, px, y, py = x.to(dt), px.to(dt), y.to(dt), py.to(dt)
        inv = 1.0 / (2.0 * sigma * sigma)
        N = x.shape[0]
        if x.is_cuda:
            kx = _block_sums(x, ex, x, ex, torch.cat([px[:, None], px[:, None] * x], dim=1), sigma)
            ky = _block_sums(x, ex, y, ey, torch.cat([py[:, None], py[:, None] * y], dim=1), sigma)
            A = (px * kx[:, 0]).sum()
            C = (px * ky[:, 0]).sum()
            ctx.save_for_backward(x, px, kx, ky)
            ctx.inv = inv
            return A, C
        kx = torch.zeros((N, 4), dtype=dt, device=x.device)  # [K_xx px, K_xx (px x)]
        ky = torch.zeros((N, 4), dtype=dt, device=x.device)  # [K_xy py, K_xy (py y)]
        ox, sx_ = _event_slices(ex)
        oy, sy_ = _event_slices(ey)
        for e, (b0, b1) in sx_.items():
            ix = ox[b0:b1]
            xe, pe = x[ix], px[ix]
            cx = torch.cat([pe[:, None], pe[:, None] * xe], dim=1)
            yb = sy_.get(e)
            if yb is not None:
                iy = oy[yb[0]:yb[1]]
                ye, qe = y[iy], py[iy]
                cy = torch.cat([qe[:, None], qe[:, None] * ye], dim=1)
            for r0 in range(0, len(ix), chunk):
                xs, rows = xe[r0:r0 + chunk], ix[r0:r0 + chunk]
                kx[rows] = _kernel_rows(xs, xe, inv, cx)
                if yb is not None:
                    ky[rows] = _kernel_rows(xs, ye, inv, cy)
        A = (px * kx[:, 0]).sum()
        C = (px * ky[:, 0]).sum()
        ctx.save_for_backward(x, px, kx, ky)
        ctx.inv = inv
        return A, C

    @staticmethod
    def backward(ctx, gA, gC):
        x, px, kx, ky = ctx.saved_tensors
        two_inv = 2.0 * ctx.inv  # 1 / sigma^2
        # d A / d px_i = 2 (K px)_i ; d C / d px_i = (K_xy py)_i
        g_px = 2.0 * gA * kx[:, 0] + gC * ky[:, 0]
        # d A / d x_i = -(2 px_i / s^2) [x_i (K px)_i - (K (px x))_i]; d C / d x_i likewise with K_xy
        dA = -(2.0 * two_inv) * px[:, None] * (x * kx[:, :1] - kx[:, 1:])
        dC = -two_inv * px[:, None] * (x * ky[:, :1] - ky[:, 1:])
        g_x = gA * dA + gC * dC
        return g_x, g_px, None, None, None, None, None, None


def _target_term(y, py, ey, sigma, chunk):
    """sum K(y, y) py py within events (no gradient)."""
    inv = 1.0 / (2.0 * sigma * sigma)
    y, py = y.to(torch.float64), py.to(torch.float64)
    if y.is_cuda:
        c = torch.zeros((y.shape[0], 4), dtype=torch.float64, device=y.device)
        c[:, 0] = py
        return (py * _block_sums(y, ey, y, ey, c, sigma)[:, 0]).sum()
    oy, sy_ = _event_slices(ey)
    B = torch.zeros((), dtype=torch.float64, device=y.device)
    for _, (b0, b1) in sy_.items():
        iy = oy[b0:b1]
        ye, qe = y[iy], py[iy]
        for r0 in range(0, len(iy), chunk):
            B = B + (qe[r0:r0 + chunk] * _kernel_rows(ye[r0:r0 + chunk], ye, inv, qe[:, None])[:, 0]).sum()
    return B


_TARGET_CACHE: dict = {}


def mmd_blockwise(x, y, px, py, sigma, ex, ey, chunk=8192):
    """``mmd`` (losses_jax.py:21-39) for clouds whose points interact only within their event
    (ex, ey): the same value, memory O(chunk x points per event).  The target-only kernel sum
    depends on the target cloud alone, which a fit keeps fixed; it is cached and reused while
    the target's positions, weights and events are equal in value to the cached ones (the
    callers rebuild those tensors on every call, so identity cannot serve as the key)."""
    hit = False
    if _TARGET_CACHE:
        cy, cpy, cey, cs, cB = _TARGET_CACHE["entry"]
        hit = (cs == float(sigma) and cy.shape == y.shape and cy.device == y.device and torch.equal(cy, y)
               and torch.equal(cpy, py) and torch.equal(cey, ey))
    if hit:
        B = cB
    else:
        B = _target_term(y, py, ey, float(sigma), int(chunk))
        _TARGET_CACHE["entry"] = (y.detach().clone(), py.detach().clone(), ey.detach().clone(), float(sigma), B)
    A, C = _BlockMMDTerms.apply(x, px, y, py, ex, ey, float(sigma), int(chunk))
    sx, sy = px.to(torch.float64).sum(), py.to(torch.float64).sum()
    return A / sx ** 2 + B / sy ** 2 - 2 * C / (sx * sy)


def mse_adc_blockwise(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_ticks, ref_hit_prob,
                      ref_event, sigma=1, lambda_Q=1):
    """``mse_adc`` (losses_jax.py:58-80) with the event-blocked MMD: the same loss (the reference
    separates events by offsetting x by event * 1e5 cm, which zeroes every cross-event kernel).
    Positions keep the reference's float32 rounding of x + event * 1e5 (event-local, see
    ``event_positions``)."""
    w_ref = (ref_Q * ref_hit_prob).to(torch.float64)
    w = (Q * hit_prob).to(torch.float64)
    s = event_positions(x, y, z, event, 1e5).to(torch.float64)
    ref_s = event_positions(ref_x, ref_y, ref_z, ref_event, 1e5).to(torch.float64)
    m = mmd_blockwise(s, ref_s, w, w_ref, sigma, event.to(torch.int64), ref_event.to(torch.int64))
    tot_ref, tot = w_ref.sum(), w.sum()
    charge_loss = ((tot - tot_ref) / (tot_ref + 1e-6)) ** 2
    aux = dict(charge_loss=charge_loss, mmd_loss_term=m)
    return m + lambda_Q * charge_loss, aux


# --- chamfer_3d and nll_loss (losses_jax.py:195-249, 329-361) ---------------------------------

def event_positions(x, y, z, event, offset):
    """Event-local positions (fl(x + event * offset) - event * offset, y, z) in float32.

    The reference stacks x + event * offset in float32, which for event >= 1 and offset 1e9
    rounds x to a multiple of 64 cm; the rounded value is kept here, and the offset (exactly
    representable, and the subtraction exact) is taken off again.  Differences between points of
    one event, the only ones that matter, are therefore the reference's bit for bit, while the
    gradients avoid the cancellation of 1e9-sized coordinates.  d/dx is 1, as in the reference.
    """
    off = event.to(torch.float32) * offset
    xo = (x.to(torch.float32) + off) - off
    return torch.stack((xo, y.to(torch.float32), z.to(torch.float32)), dim=-1)


def _sorted_blocks(qe, ce):
    """Stable event order of queries and candidates, and each sorted query's [lo, hi) of its event
    among the sorted candidates."""
    oq = torch.argsort(qe, stable=True)
    oc = torch.argsort(ce, stable=True)
    qs, cs = qe[oq].contiguous(), ce[oc].contiguous()
    lo = torch.searchsorted(cs, qs, right=False).to(torch.int32)
    hi = torch.searchsorted(cs, qs, right=True).to(torch.int32)
    return oq, oc, qs, lo, hi


def _nearest(pq, eq, pc, ec, origin):
    """For each query point, the index (into pc) of its nearest candidate of the same event (first on
    ties) and the squared distance, on the GPU (lsim_nearest).  With ``origin`` the queries of
    event 0 also see the zero-weight padding point at the origin (index -1 when it wins)."""
    import ctypes as C

    from . import _lib

    lib = _lib.load()
    oq, oc, qs, lo, hi = _sorted_blocks(eq.to(torch.int64), ec.to(torch.int64))
    q4, c4 = _pad4(pq[oq]), _pad4(pc[oc])
    n0 = int((qs == 0).sum().item()) if origin else 0
    idx = torch.empty(pq.shape[0], dtype=torch.int32, device=pq.device)
    d = torch.empty(pq.shape[0], dtype=torch.float32, device=pq.device)
    st = C.c_void_p(torch.cuda.current_stream(pq.device).cuda_stream)
    _lib.check(lib.lsim_nearest(C.c_void_p(q4.data_ptr()), q4.shape[0], C.c_void_p(lo.data_ptr()),
                                C.c_void_p(hi.data_ptr()), C.c_void_p(c4.data_ptr()), c4.shape[0], n0,
                                C.c_void_p(idx.data_ptr()), C.c_void_p(d.data_ptr()), st))
    i_sorted = idx.to(torch.int64)
    i_orig = torch.where(i_sorted >= 0, oc[i_sorted.clamp(min=0)], torch.full_like(i_sorted, -1))
    out_i = torch.empty_like(i_orig)
    out_d = torch.empty_like(d)
    out_i[oq] = i_orig
    out_d[oq] = d
    return out_i, out_d


def _take_or_zero(v, i, m):
    """v[i] where m, else 0 (v may be empty when m is all false)."""
    if v.shape[0] == 0:
        return torch.zeros(i.shape[0], dtype=v.dtype, device=v.device)
    return torch.where(m, v[i.clamp(min=0)], 0.0)


class _ChamferFn(torch.autograd.Function):
    """chamfer_distance_3d_og (losses_jax.py:195-229) between the simulated hits a and the target
    hits b, with points interacting only within their event (the reference's 1e9 cm event offset
    puts every cross-event distance above its 1e4 cut).  Gradients flow to the simulated positions
    and weights through the nearest pairs, with the argmins fixed, as reverse-mode AD does."""

    @staticmethod
    def forward(ctx, pa, wa, pb, wb, ea, eb, pad_a, pad_b):
        if not pa.is_cuda:
            raise RuntimeError("chamfer_3d runs on the GPU (lsim_nearest)")
        ia, da = _nearest(pa, ea, pb, eb, pad_b)
        ib, db = _nearest(pb, eb, pa, ea, pad_a)
        wa64, wb64 = wa.to(torch.float64), wb.to(torch.float64)
        ma = (ia >= 0) & (da < 1e4)
        mb = (ib >= 0) & (db < 1e4)
        wb_at = _take_or_zero(wb64, ia, ma)
        wa_at = _take_or_zero(wa64, ib, mb)
        la = torch.where(ma, da.to(torch.float64), 0.0) * wa64 * wb_at
        lb = torch.where(mb, db.to(torch.float64), 0.0) * wb64 * wa_at
        ctx.save_for_backward(pa, pb, wa64, wb64, ia, ib, da, db, ma, mb, wb_at)
        ctx.wa_dtype = wa.dtype
        return la.sum() + lb.sum()

    @staticmethod
    def backward(ctx, g):
        pa, pb, wa, wb, ia, ib, da, db, ma, mb, wb_at = ctx.saved_tensors
        pa64, pb64 = pa.to(torch.float64), pb.to(torch.float64)
        # a -> b: d(d_a wa wb*) = 2 (pa - pb*) wa wb* dpa + d_a wb* dwa
        diff_a = pa64 - pb64[ia.clamp(min=0)] if pb64.shape[0] else torch.zeros_like(pa64)
        g_pa = torch.where(ma[:, None], 2.0 * diff_a * (wa * wb_at)[:, None], 0.0)
        g_wa = torch.where(ma, da.to(torch.float64) * wb_at, 0.0)
        # b -> a: the nearest a* of each b gets 2 (pa* - pb) wb wa* and d_b wb
        sel = torch.nonzero(mb).squeeze(1)
        if sel.numel():
            a_star = ib[sel]
            diff_b = pa64[a_star] - pb64[sel]
            g_pa = g_pa.index_add(0, a_star, 2.0 * diff_b * (wb[sel] * wa[a_star])[:, None])
            g_wa = g_wa.index_add(0, a_star, db[sel].to(torch.float64) * wb[sel])
        return (g * g_pa).to(pa.dtype), (g * g_wa).to(ctx.wa_dtype), None, None, None, None, None, None


def chamfer_3d(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_ticks, ref_hit_prob,
               ref_event, adc_norm=10., match_z=False):
    """Weighted chamfer distance between simulated and target hits (losses_jax.py:231-249).

    Positions are (x + event * 1e9, y, z) in float32 (kept event-local, see ``event_positions``)
    and weights Q / adc_norm, as in the reference;
    the padding that ``pad_size`` adds there (zero-weight points at the origin) is honoured: an
    event-0 hit whose nearest point is such a padding point contributes nothing.  The pairwise
    minimum runs on the GPU per event (``lsim_nearest``).  Returns (loss, {})."""
    from .sim import pad_size

    n_ref, n_cur = int(ref_x.shape[0]), int(x.shape[0])
    new_ref, new_cur = pad_size((n_ref, n_cur), "chamfer_3d")
    pa = event_positions(x, y, z, event, 1e9)
    pb = event_positions(ref_x, ref_y, ref_z, ref_event, 1e9)
    wa = Q.to(torch.float32) / adc_norm
    wb = ref_Q.to(torch.float32) / adc_norm
    loss = _ChamferFn.apply(pa, wa, pb, wb, event.to(torch.int64), ref_event.to(torch.int64), new_cur > n_cur,
                            new_ref > n_ref)
    return loss, dict()


class _NLLFn(torch.autograd.Function):
    """-sum_j log(lambda_j + 1e-10) [ref_p_j > 0.5] with lambda_j = sum_i p_i K(c_i, r_j), the 4-D
    Gaussian-kernel intensity of nll_loss (losses_jax.py:329-361), summed within events (the 1e9 cm
    event offset makes every cross-event kernel value exactly 0 in float32).  Kernel values in
    float32 and sums in float64 on the GPU (lsim_rbf_sums); gradients to the simulated points and
    probabilities."""

    @staticmethod
    def forward(ctx, c4, p, r4, rp, ec, er, sigma):
        if not c4.is_cuda:
            raise RuntimeError("nll_loss runs on the GPU (lsim_rbf_sums)")
        coef = torch.zeros((c4.shape[0], 4), dtype=torch.float64, device=c4.device)
        coef[:, 0] = p.to(torch.float64)
        lam = _block_sums(r4, er, c4, ec, coef, sigma, dims=4)[:, 0]
        m = (rp > 0.5).to(torch.float64)
        term1 = (torch.log(lam + 1e-10) * m).sum()
        ctx.save_for_backward(c4, p, r4, ec, er, lam, m)
        ctx.sigma = sigma
        return -term1

    @staticmethod
    def backward(ctx, g):
        c4, p, r4, ec, er, lam, m = ctx.saved_tensors
        s2 = ctx.sigma * ctx.sigma
        gj = m / (lam + 1e-10)
        r64 = r4.to(torch.float64)
        coef = torch.stack((gj, gj * r64[:, 0], gj * r64[:, 1], gj * r64[:, 2]), dim=1)
        S = _block_sums(c4, ec, r4, er, coef, ctx.sigma, dims=4)
        coef2 = torch.zeros_like(coef)
        coef2[:, 0] = gj * r64[:, 3]
        S3 = _block_sums(c4, ec, r4, er, coef2, ctx.sigma, dims=4)[:, 0]
        c64, p64 = c4.to(torch.float64), p.to(torch.float64)
        # d(-term1)/dp_i = -(K g)_i ; d(-term1)/dc_i = (p_i / s^2) (c_i (K g)_i - (K g r)_i)
        g_p = -S[:, 0]
        Sr = torch.cat((S[:, 1:], S3[:, None]), dim=1)
        g_c = (p64 / s2)[:, None] * (c64 * S[:, :1] - Sr)
        return (g * g_c).to(c4.dtype), (g * g_p).to(p.dtype), None, None, None, None, None


def nll_loss(params, Q, x, y, z, ticks, hit_prob, event, ref_Q, ref_x, ref_y, ref_z, ref_ticks, ref_hit_prob,
             ref_event, adc_norm=10., sigma=1., match_z=False):
    """Poisson-process negative log-likelihood of the target hits under the simulated hits
    (losses_jax.py:329-361): sum(hit_prob) - sum_j log(lambda_j + 1e-10) over target hits with
    hit_prob > 0.5, lambda_j = sum_i hit_prob_i exp(-|c_i - r_j|^2 / 2 sigma^2) over the 4-D points
    (x + event 1e9, y, z, Q / adc_norm).  Returns (loss, {})."""
    from .sim import pad_size

    pad_size((int(ref_x.shape[0]), int(x.shape[0])), "nll_loss")  # the reference's padding history
    c4 = torch.cat((event_positions(x, y, z, event, 1e9), (Q.to(torch.float32) / adc_norm)[:, None]), dim=1)
    r4 = torch.cat((event_positions(ref_x, ref_y, ref_z, ref_event, 1e9),
                    (ref_Q.to(torch.float32) / adc_norm)[:, None]), dim=1)
    p = hit_prob.to(torch.float32)
    t1 = _NLLFn.apply(c4, p, r4, ref_hit_prob, event.to(torch.int64), ref_event.to(torch.int64), float(sigma))
    return p.to(torch.float64).sum() + t1, dict()


# --- soft-DTW losses (losses_jax.py:309-327, softdtw_jax.py:205-275) -------------------------------

def _sdtw_hip(x, y, gamma, want_grad):
    import ctypes as C

    from . import _lib

    lib = _lib.load()
    x = x.detach().to(torch.float32).contiguous()
    y = y.detach().to(torch.float32).contiguous()
    if not x.is_cuda:
        raise _lib.LsimError("soft-DTW runs on the GPU (lsim_sdtw)")
    n, m = x.numel(), y.numel()
    ws = C.c_size_t()
    _lib.check(lib.lsim_sdtw_query(n, m, C.byref(ws)))
    buf = torch.empty(ws.value, dtype=torch.uint8, device=x.device)
    val = torch.empty(1, dtype=torch.float32, device=x.device)
    g = torch.empty(n, dtype=torch.float64, device=x.device) if want_grad else None
    st = C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(lib.lsim_sdtw(C.c_void_p(x.data_ptr()), n, C.c_void_p(y.data_ptr()), m, C.c_float(float(gamma)),
                             C.c_void_p(buf.data_ptr()), buf.numel(), C.c_void_p(val.data_ptr()),
                             C.c_void_p(g.data_ptr()) if g is not None else None, st))
    return val[0], g


class _SoftDTWFn(torch.autograd.Function):
    """Soft-DTW value of two 1-D sequences; gradients to either (the second by symmetry:
    sdtw(x, y) = sdtw(y, x))."""

    @staticmethod
    def forward(ctx, x, y, gamma):
        v, gx = _sdtw_hip(x, y, gamma, ctx.needs_input_grad[0])
        gy = _sdtw_hip(y, x, gamma, True)[1] if ctx.needs_input_grad[1] else None
        ctx.save_for_backward(*(t for t in (gx, gy) if t is not None))
        ctx.has = (gx is not None, gy is not None)
        ctx.dtypes = (x.dtype, y.dtype)
        return v.to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        saved = list(ctx.saved_tensors)
        gx = (g * saved.pop(0)).to(ctx.dtypes[0]) if ctx.has[0] else None
        gy = (g * saved.pop(0)).to(ctx.dtypes[1]) if ctx.has[1] else None
        return gx, gy, None


class SoftDTW:
    """Soft dynamic time warping cost (softdtw_jax.py:205-275) with the squared-Euclidean ground
    cost, on the GPU (``lsim_sdtw``); ``pairwise(x, y)`` takes two 1-D sequences."""

    def __init__(self, gamma=1.0, ground_cost=None, debiased=False, **_):
        if ground_cost is not None:
            raise NotImplementedError("only the squared-Euclidean ground cost is implemented")
        self.gamma = float(gamma)
        self.debiased = debiased

    def pairwise(self, x, y):
        x, y = x.reshape(-1), y.reshape(-1)
        c = _SoftDTWFn.apply(x, y, self.gamma)
        if self.debiased:
            return c - 0.5 * (_SoftDTWFn.apply(x, x.detach(), self.gamma) * 2 - _SoftDTWFn.apply(x.detach(), x.detach(), self.gamma)
                              + _SoftDTWFn.apply(y, y.detach(), self.gamma) * 2 - _SoftDTWFn.apply(y.detach(), y.detach(), self.gamma))
        return c


def sdtw_loss(adcs, ref, dstw):
    """Soft-DTW between the positive entries of the flattened arrays, in order (losses_jax.py:309-316)."""
    a, r = adcs.reshape(-1), ref.reshape(-1)
    return dstw.pairwise(a[a > 0], r[r > 0]), dict()


def sdtw_adc(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref, dstw):
    """losses_jax.py:318-319."""
    return sdtw_loss(adcs, ref, dstw)


def sdtw_time(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref, dstw):
    """losses_jax.py:321-322."""
    return sdtw_loss(ticks, ticks_ref, dstw)


def sdtw_time_adc(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref, dstw, alpha=0.5):
    """losses_jax.py:324-327."""
    la, _ = sdtw_adc(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref, dstw)
    lt, _ = sdtw_time(params, adcs, pixels, ticks, ref, pixels_ref, ticks_ref, dstw)
    return alpha * la + (1 - alpha) * lt, dict()


def loss_functions(adc_norm=10.0, match_z=False):
    """The named fit losses of the reference's driver (optimize/fit_params.py:302-312) that run
    here, as (function, keyword arguments) for ``GenericLossStrategy(fn, **kw)``."""
    return {
        "mse_adc": (mse_adc_blockwise, {}),
        "chamfer_3d": (chamfer_3d, {"adc_norm": adc_norm, "match_z": match_z}),
        "nll": (nll_loss, {"adc_norm": adc_norm, "sigma": 1.0}),
        # the soft-DTW losses take a SoftDTW object as ``dstw`` (fit_params.py:334-335)
        "sdtw_adc": (sdtw_adc, {"dstw": SoftDTW(gamma=1.0)}),
        "sdtw_time": (sdtw_time, {"dstw": SoftDTW(gamma=1.0)}),
        "sdtw_time_adc": (sdtw_time_adc, {"dstw": SoftDTW(gamma=1.0), "alpha": 0.5}),
    }
