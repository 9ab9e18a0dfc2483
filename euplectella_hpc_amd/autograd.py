"""PyTorch binding of the grey-lattice drag objective and its discrete adjoint.

``GreyDrag.apply(s, lattice, t0, t1, every, norm)`` returns the time-averaged drag coefficient
J(s) = sum_{t=t0..t1} F_x(a^t) / ((t1 - t0 + 1) norm) of a grey-mode :class:`Lattice` whose
initial populations are already set; ``J.backward()`` delivers dJ/ds computed by the library's
hand-written adjoint kernels (checkpoint every ``every`` steps, recompute per segment).  The solid
fraction and the gradient stay on the GPU: the tensors' device pointers go straight through the
C-ABI (eup_set_grey_device, eup_grey_objective_device).
"""
import torch


class GreyDrag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, lattice, t0, t1, every, norm):
        if not s.is_cuda:
            raise ValueError("the solid fraction must be a CUDA (HIP) tensor")
        dt = {torch.float64: "f64", torch.float32: "f32"}.get(s.dtype)
        if dt is None:
            raise ValueError("the solid fraction must be float32 or float64")
        sc = s.detach().contiguous()
        torch.cuda.current_stream(sc.device).synchronize()
        lattice.set_grey_device(sc.data_ptr(), dt)
        if ctx.needs_input_grad[0]:
            g = torch.empty_like(sc)
            J = lattice.grey_objective_device(t0, t1, every, norm, g.data_ptr(), dt)
            ctx.save_for_backward(g)
        else:
            J = lattice.grey_objective_device(t0, t1, every, norm, None, dt)
        return torch.tensor(J, dtype=s.dtype, device=s.device)

    @staticmethod
    def backward(ctx, gJ):
        (g,) = ctx.saved_tensors
        return gJ * g, None, None, None, None, None


def grey_drag(s, lattice, t0, t1, every, norm):
    return GreyDrag.apply(s, lattice, t0, t1, every, norm)
