"""torch.autograd.Function over the C ABI (forward = synthesis, backward = adjoint).

The forward/backward pair is cmbl_forward_device / cmbl_backward_device on
torch's current HIP stream; torch only provides device memory and the stream.
"""
import torch

from . import Map


def _stream():
    return torch.cuda.current_stream().cuda_stream


class Synthesis(torch.autograd.Function):
    """map = Y B a(x) for a packed a_lm tensor x (float64, on the GPU)."""

    @staticmethod
    def forward(ctx, x, mp):
        if x.dtype != torch.float64 or not x.is_cuda or x.numel() != mp.nparams:
            raise ValueError("x must be a float64 CUDA tensor with (lmax+1)^2 values")
        x = x.contiguous()
        out = torch.empty(mp.npix, dtype=torch.float64, device=x.device)
        mp.forward_device(x.data_ptr(), out.data_ptr(), _stream())
        ctx.mp = mp
        return out

    @staticmethod
    def backward(ctx, grad_map):
        mp = ctx.mp
        v = grad_map.contiguous().to(torch.float64)
        g = torch.empty(mp.nparams, dtype=torch.float64, device=v.device)
        mp.backward_device(v.data_ptr(), g.data_ptr(), _stream())
        return g, None


def synthesize(x, mp: Map):
    return Synthesis.apply(x, mp)
