"""Front-end helpers on torch tensors (fee_jax.py:55-71; losses_jax.py:380-383)."""
import torch


def digitize(params, integral_list):
    """ADC counts from collected charge (fee_jax.py:55-71): clamped, not rounded."""
    x = torch.clamp(integral_list * params.GAIN + params.V_PEDESTAL - params.V_CM, min=0)
    return torch.clamp(x * params.ADC_COUNTS / (params.V_REF - params.V_CM), max=params.ADC_COUNTS)


def adc2charge(dw, params):
    """ADC counts to charge in ke- (losses_jax.py:380-383)."""
    return (dw / params.ADC_COUNTS * (params.V_REF - params.V_CM) + params.V_CM - params.V_PEDESTAL) / params.GAIN * 1e-3


# --- probabilistic readout (fee_jax.py:334-495, sim_jax.py:771-813) ---------------------------

def _pf_settings(params):
    V, S = int(params.fee_paths_scaling), int(params.MAX_ADC_VALUES)
    interval = round((3 * params.CLOCK_CYCLE + params.ADC_HOLD_DELAY * params.CLOCK_CYCLE) / params.t_sampling)
    return (V, S, interval, float(params.DISCRIMINATION_THRESHOLD), float(params.RESET_NOISE_CHARGE),
            float(params.t_sampling))


def _pf_forward(w, settings, stop_threshold):
    import ctypes as C

    from . import _lib

    lib = _lib.load()
    P, T = w.shape
    V, S, interval, theta, sigma, ts = settings
    ws = C.c_size_t()
    _lib.check(lib.lsim_prob_fee_query(P, T, V, S, C.byref(ws)))
    buf = torch.empty(max(ws.value, 256), dtype=torch.uint8, device=w.device)
    lp = torch.empty((P, S, T - 1), dtype=torch.float32, device=w.device)
    q = torch.empty((P, S, T - 1), dtype=torch.float32, device=w.device)
    st = C.c_void_p(torch.cuda.current_stream(w.device).cuda_stream)
    _lib.check(lib.lsim_prob_fee(C.c_void_p(w.data_ptr()), P, T, V, S, interval, C.c_float(theta), C.c_float(sigma),
                                 C.c_float(ts), C.c_float(float(stop_threshold)), C.c_void_p(buf.data_ptr()),
                                 buf.numel(), C.c_void_p(lp.data_ptr()), C.c_void_p(q.data_ptr()), st))
    return lp, q, buf


class ProbFeeFunction(torch.autograd.Function):
    """(log_prob, charge) of the probabilistic readout with its hand-written adjoint
    (``lsim_prob_fee_backward``): the path selections are held fixed, as the reference's
    stop_gradient around top_k holds them (fee_jax.py:377)."""

    @staticmethod
    def forward(ctx, w, settings, stop_threshold):
        lp, q, buf = _pf_forward(w, settings, stop_threshold)
        ctx.settings, ctx.buf, ctx.shape = settings, buf, tuple(w.shape)
        ctx.save_for_backward(lp)
        return lp, q

    @staticmethod
    def backward(ctx, g_lp, g_q):
        import ctypes as C

        from . import _lib

        (lp,) = ctx.saved_tensors
        P, T = ctx.shape
        V, S, interval, theta, sigma, ts = ctx.settings
        g_lp = torch.zeros_like(lp) if g_lp is None else g_lp.to(torch.float32).contiguous()
        g_q = torch.zeros_like(lp) if g_q is None else g_q.to(torch.float32).contiguous()
        d_w = torch.empty((P, T), dtype=torch.float32, device=lp.device)
        st = C.c_void_p(torch.cuda.current_stream(lp.device).cuda_stream)
        _lib.check(_lib.load().lsim_prob_fee_backward(P, T, V, S, interval, C.c_float(theta), C.c_float(sigma),
                                                      C.c_float(ts), C.c_void_p(ctx.buf.data_ptr()), ctx.buf.numel(),
                                                      C.c_void_p(lp.data_ptr()), C.c_void_p(g_lp.data_ptr()),
                                                      C.c_void_p(g_q.data_ptr()), C.c_void_p(d_w.data_ptr()), st))
        ctx.buf = None  # single backward
        return d_w, None, None


def get_adc_values_average_noise_vmap(params, wfs, stop_threshold=1e-9):
    """Beam search of the self-trigger under reset noise (fee_jax.py:386-461) on the GPU
    (``lsim_prob_fee``): returns (log_prob, charge), each (P, MAX_ADC_VALUES, T - 1) float32, the
    log probability that hit s starts at tick t and its expected charge in electrons.  If ``wfs``
    requires gradients both outputs carry autograd history through ``lsim_prob_fee_backward``."""
    from . import _lib

    w = torch.as_tensor(wfs).to(torch.float32).contiguous()
    if not w.is_cuda:
        raise _lib.LsimError("the probabilistic readout runs on the GPU (lsim_prob_fee)")
    settings = _pf_settings(params)
    if w.requires_grad and torch.is_grad_enabled():
        return ProbFeeFunction.apply(w, settings, stop_threshold)
    lp, q, _ = _pf_forward(w, settings, stop_threshold)
    return lp, q


def get_average_hit_values(ticks_prob, adcs_distrib):
    """Per (pixel, hit): expected tick, expected ADC and total weight over ticks (fee_jax.py:464-495)."""
    lam = ticks_prob.sum(dim=2)
    t = torch.arange(ticks_prob.shape[2], device=ticks_prob.device, dtype=ticks_prob.dtype)
    den = torch.clamp(lam, min=1e-10)
    return (t * ticks_prob).sum(dim=2) / den, (adcs_distrib * ticks_prob).sum(dim=2) / den, lam


def simulate_probabilistic(params, wfs, unique_pixels):
    """sim_jax.py:771-813: (adcs_distrib, pixel_x, pixel_y, ticks_prob, event) with ticks_prob the
    log probabilities of get_adc_values_average_noise_vmap and adcs_distrib its digitised
    expected charges."""
    from .detsim import get_pixel_coordinates, id2pixel

    ticks_prob, charge = get_adc_values_average_noise_vmap(params, wfs)
    adcs = digitize(params, charge)
    px, py, plane, event = id2pixel(params, torch.as_tensor(unique_pixels))
    xy = get_pixel_coordinates(params, px, py, plane)
    return adcs, xy[:, 0], xy[:, 1], ticks_prob, event
