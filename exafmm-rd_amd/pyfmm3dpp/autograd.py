"""torch.autograd binding of the periodic FMM over the C-ABI (optional; needs torch with ROCm).

`potential(plan, x, q)` returns phi (n,) for positions x (n, 3) and charges q (n,), float64 CUDA
tensors on the plan's device.  Its backward is the hand-written adjoint (fmm3dpp_evaluate_adjoint_device):
one more FMM evaluation with the upstream gradient as charges, combined by a HIP kernel into
dL/dq and dL/dx.  `energy(plan, x, q)` is E = 1/2 sum q phi, whose gradients are dE/dq = phi and
dE/dx = q grad(phi).  The plan must use FMM3DPP_DIPOLE_OFF or FMM3DPP_DIPOLE_PHYSICAL.
"""
import torch


def _soa(x):
    return [x[:, k].contiguous() for k in range(3)]


class _Potential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, plan, x, q):
        if x.dtype != torch.float64 or q.dtype != torch.float64 or not x.is_cuda or not q.is_cuda:
            raise TypeError("positions and charges must be float64 tensors on the GPU")
        n = x.shape[0]
        xs, ys, zs = _soa(x.detach())
        qc = q.detach().contiguous()
        phi = torch.empty(n, dtype=torch.float64, device=x.device)
        grad = torch.empty((n, 3), dtype=torch.float64, device=x.device)
        torch.cuda.synchronize(x.device)
        plan.set_particles_device(n, xs.data_ptr(), ys.data_ptr(), zs.data_ptr())
        plan.evaluate_device(qc.data_ptr(), phi.data_ptr(), grad.data_ptr())
        torch.cuda.synchronize(x.device)   # the plan runs on its own stream
        ctx.plan = plan
        ctx.save_for_backward(x.detach(), qc, grad)
        return phi

    @staticmethod
    def backward(ctx, gphi):
        x, q, grad = ctx.saved_tensors
        plan = ctx.plan
        n = x.shape[0]
        w = gphi.detach().to(torch.float64).contiguous()
        xs, ys, zs = _soa(x)
        dq = torch.empty(n, dtype=torch.float64, device=x.device)
        dx = torch.empty((n, 3), dtype=torch.float64, device=x.device)
        torch.cuda.synchronize(x.device)
        # the plan's tree may have been rebuilt by another forward since: rebuild for these positions
        plan.set_particles_device(n, xs.data_ptr(), ys.data_ptr(), zs.data_ptr())
        plan.evaluate_adjoint_device(q.data_ptr(), w.data_ptr(), grad.data_ptr(), dq.data_ptr(), dx.data_ptr())
        torch.cuda.synchronize(x.device)
        return None, dx, dq


def potential(plan, x, q):
    return _Potential.apply(plan, x, q)


def energy(plan, x, q):
    return 0.5 * torch.sum(q * potential(plan, x, q))
