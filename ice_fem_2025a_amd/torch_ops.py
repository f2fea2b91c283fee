"""torch.autograd.Function over the engine's assembly residual.

    r = assembly_residual(phys, Gc_e, A_e, n)     # forward: ife_physics_assemble_all
    (lam * r).sum().backward()                     # backward: ife_physics_adjoint

r is the concatenation of every step's residual f (step order, each in the step's global
dof order) as a float64 CUDA tensor; Gc_e and A_e are per-element float64 tensors of the
Fracture/PhaseField model, n the Glen exponent (0-d tensor).  The backward pass is the
hand-written adjoint kernel: it recomputes the quadrature-point state and contracts the
cotangent with d(r)/d(Gc_e, A_e, n), so nothing of the forward pass is stored.  All device
work runs on torch's current stream.  torch is plumbing here (device memory, streams,
autograd bookkeeping); every number comes from libife.so."""
import ctypes as C

import torch

from . import hip


class _TorchStream:
    """The hipStream_t of torch's current stream, in the form the Physics methods take."""

    def __init__(self):
        self.handle = C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _set_params(phys, model, Gc_e, A_e, n):
    phys.set_element_params(model, Gc_e.detach().double().cpu().numpy(), A_e.detach().double().cpu().numpy())
    phys.set_param(model, "n", float(n))


class AssemblyResidual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phys, model, Gc_e, A_e, n):
        _set_params(phys, model, Gc_e, A_e, n)
        st = _TorchStream()
        phys.assemble_all(stream=st)
        rows = [phys.n_rows(s) for s in range(phys.nsteps)]
        out = torch.empty(sum(rows), dtype=torch.float64, device="cuda")
        off = 0
        for s, nr in enumerate(rows):
            _, fp = phys.Kf_ptrs(s)
            if nr:
                hip.copy_d2d(out.data_ptr() + 8 * off, fp, 8 * nr, st.handle)
            off += nr
        ctx.phys, ctx.model, ctx.rows = phys, model, rows
        ctx.save_for_backward(Gc_e.detach(), A_e.detach(), n.detach())
        return out

    @staticmethod
    def backward(ctx, grad):
        phys, model, rows = ctx.phys, ctx.model, ctx.rows
        Gc_e, A_e, n = ctx.saved_tensors
        _set_params(phys, model, Gc_e, A_e, n)  # the point the forward pass evaluated
        lam = grad.detach().to(dtype=torch.float64, device="cuda").contiguous()
        ptrs, off = [], 0
        for nr in rows:
            ptrs.append(lam.data_ptr() + 8 * off if nr else None)
            off += nr
        ne = Gc_e.numel()
        g = torch.empty((3, ne), dtype=torch.float64, device="cuda")
        phys.adjoint_device(ptrs, g[0].data_ptr(), g[1].data_ptr(), g[2].data_ptr(), stream=_TorchStream())
        return None, None, g[0].clone(), g[1].clone(), g[2].sum().reshape(n.shape)


def assembly_residual(phys, Gc_e, A_e, n, model=0):
    return AssemblyResidual.apply(phys, model, Gc_e, A_e, n)
