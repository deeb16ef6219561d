"""torch.autograd.Function over the library's forward record and reverse sweep.

    psi_T = EulerEvolve.apply(psi0, eps, k, M, nsteps, ckpt_every, base_desc_kwargs)

Forward runs `nsteps` steps of the selected integrator (euler by default;
euler_mobility and pseudospectral through desc_kwargs) on the GPU
(ch_sys_record, with checkpoints); backward calls ch_sys_adjoint_device with
the incoming dL/dpsi_T and returns dL/dpsi0 and dL/d(eps, k, M).  The tensors live on the GPU in the system's
precision (float64 -> FP64 fields, float32 -> FP32 fields).  This wrapper is
for checking gradients (torch.autograd.gradcheck, composition with torch
losses); the step and its adjoint are the HIP kernels.
"""
import torch

from .system import Context, System, make_desc

_ctx = None


def _context():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


class EulerEvolve(torch.autograd.Function):
    @staticmethod
    def forward(fctx, psi0, eps, k, M, nsteps, ckpt_every, desc_kwargs):
        if not psi0.is_cuda or psi0.dtype not in (torch.float64, torch.float32):
            raise ValueError("psi0 must be a float32/float64 GPU tensor")
        prec = "fp64" if psi0.dtype == torch.float64 else "fp32"
        kw = dict(desc_kwargs)
        desc = make_desc(precision=prec, epsilon=float(eps), k=float(k), M=float(M), **kw)
        ctx = _context()
        sysm = System(ctx, desc)
        x = psi0.detach().contiguous()
        torch.cuda.synchronize()
        sysm.upload_device(x.data_ptr())
        sysm.record(int(nsteps), int(ckpt_every))
        out = torch.empty_like(x)
        sysm.download_device(out.data_ptr())
        fctx.sysm = sysm
        fctx.param_dtype = [t.dtype if torch.is_tensor(t) else torch.float64 for t in (eps, k, M)]
        fctx.param_device = [t.device if torch.is_tensor(t) else psi0.device for t in (eps, k, M)]
        return out

    @staticmethod
    def backward(fctx, grad_out):
        g = grad_out.detach().contiguous()
        gin = torch.empty_like(g)
        torch.cuda.synchronize()
        # the recorded system stays alive with the graph: backward may run more
        # than once (retain_graph, gradcheck's row-by-row vector-Jacobian products)
        grads = fctx.sysm.adjoint_device(g.data_ptr(), gin.data_ptr())
        pg = [torch.tensor(grads[i], dtype=fctx.param_dtype[i], device=fctx.param_device[i]) for i in range(3)]
        return gin, pg[0], pg[1], pg[2], None, None, None


def euler_evolve(psi0, eps, k, M, nsteps, ckpt_every=100, **desc_kwargs):
    """Differentiable nsteps of an integrator (euler unless desc_kwargs names another)."""
    desc_kwargs.setdefault("integrator", "euler")
    return EulerEvolve.apply(psi0, eps, k, M, nsteps, ckpt_every, tuple(sorted(desc_kwargs.items())))
