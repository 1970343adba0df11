"""torch.autograd bridge for the exact adjoint (cg_tb_adjoint).

WinProbability.apply(p4, tb, initial, winstates) returns the initial boards'
mean win probability P(p4) (as a 0-d float64 tensor) computed by the HIP
solve of a generated device-resident tablebase; its backward multiplies the
incoming gradient by dP/dp4 from the upward occupation-measure sweep.  The
tensor is plumbing only: all arithmetic runs in libcablegen.so.
"""
import torch


class WinProbability(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p4, tb, initial, winstates):
        p = float(p4.detach().cpu().item())
        tb.solve(winstates, p4=p, keep_argmax=True)
        P, dP, _ = tb.adjoint(initial)
        ctx.save_for_backward(torch.tensor(dP, dtype=torch.float64))
        return torch.tensor(P, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        (dP,) = ctx.saved_tensors
        return grad * dP, None, None, None


def win_probability(p4, tb, initial, winstates):
    return WinProbability.apply(p4, tb, initial, winstates)
