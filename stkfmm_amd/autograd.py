"""torch.autograd.Function over the HIP direct-sum P2P and its hand-written adjoint kernels.

u = P2PFunction.apply(uker, precision, src_coord, src_value, trg_coord) returns the [n_trg, kt]
direct sum (scale factor applied, coincident pairs masked).  backward() calls
stkfmm_p2p_vjp_device, which returns dL/d(src_value), dL/d(src_coord) and dL/d(trg_coord) for the
incoming cotangent.  Supported micro-kernels: LAPLACE_P, LAPLACE_PGRAD, STOKES_VEL (Stokeslet),
STOKES_REGVEL (gradient with respect to eps included), RPY_U, RPY_ULAPU, STK_ULAPU, STOKES_PVEL,
STOKES_DOUBLEPVEL, STOKES_PVEL_FUSED.  The reference has no derivatives; this is new API.
"""
import torch

from . import PREC, UKER, uker_device, uker_dims, vjp_device

_ADJOINT_UKERS = (UKER.LAPLACE_P, UKER.LAPLACE_PGRAD, UKER.STOKES_VEL, UKER.STOKES_REGVEL, UKER.RPY_U,
                  UKER.RPY_ULAPU, UKER.STK_ULAPU, UKER.STOKES_PVEL, UKER.STOKES_DOUBLEPVEL, UKER.STOKES_PVEL_FUSED)


class P2PFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, uker, precision, src_coord, src_value, trg_coord):
        uker = UKER(int(uker))
        if uker not in _ADJOINT_UKERS:
            raise ValueError(f"no adjoint kernel for {uker.name}")
        _, kt = uker_dims(uker)
        sc, sv, tc = src_coord.contiguous(), src_value.contiguous(), trg_coord.contiguous()
        out = torch.zeros((tc.shape[0], kt), dtype=torch.float64, device=tc.device)
        uker_device(uker, sc, sv, tc, out, precision=precision)
        ctx.save_for_backward(sc, sv, tc)
        ctx.uker = uker
        ctx.precision = precision
        return out

    @staticmethod
    def backward(ctx, g_out):
        sc, sv, tc = ctx.saved_tensors
        need_sc, need_sv, need_tc = ctx.needs_input_grad[2], ctx.needs_input_grad[3], ctx.needs_input_grad[4]
        gsv, gsc, gtc = vjp_device(ctx.uker, sc, sv, tc, g_out.contiguous(), precision=ctx.precision,
                                   need=(need_sv, need_sc, need_tc))
        return None, None, gsc, gsv, gtc


def p2p(uker, src_coord, src_value, trg_coord, precision=PREC.FP64):
    """Differentiable direct-sum P2P (see P2PFunction)."""
    return P2PFunction.apply(int(uker), int(precision), src_coord, src_value, trg_coord)
