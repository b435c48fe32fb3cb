"""Front-end helpers on torch tensors (fee_jax.py:55-71; losses_jax.py:380-383)."""
import torch


def digitize(params, integral_list):
    """ADC counts from collected charge (fee_jax.py:55-71): clamped, not rounded."""
    x = torch.clamp(integral_list * params.GAIN + params.V_PEDESTAL - params.V_CM, min=0)
    return torch.clamp(x * params.ADC_COUNTS / (params.V_REF - params.V_CM), max=params.ADC_COUNTS)


def adc2charge(dw, params):
    """ADC counts to charge in ke- (losses_jax.py:380-383)."""
    return (dw / params.ADC_COUNTS * (params.V_REF - params.V_CM) + params.V_CM - params.V_PEDESTAL) / params.GAIN * 1e-3
