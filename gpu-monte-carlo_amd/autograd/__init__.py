"""Differentiable XY and Kawamura lattice energies (torch.autograd over HIP kernels).

``lattice_energy(state, J, twist, model)`` returns the per-replica Hamiltonian

    H_r = -J sum_bonds w cos(theta_p - theta_q - Delta_r [x bonds only])

with the reference's bond weights (w = 1 for XY, models/xy/measure.py:71-96;
w = (1 + tau_p tau_q)(d_p . d_q) for Kawamura, models/kawamura/measure.py:71-99)
and a uniform boundary twist Delta_r on the x bonds (x = axis 1, the bond
direction of the reference's twist_cos_x / twist_sin_x channels).  Gradients
flow to the state (all five Kawamura channels), to J and to the twist; second
derivatives flow through (J, twist) for both models and through the state for
XY (Hessian-vector product kernel).  The closed forms at Delta = 0 are
dH/dDelta = -J twist_sin_x, d2H/dDelta2 = J twist_cos_x, dH/dJ = H / J.

All per-site work is done by the HIP kernels behind ``mc_*_twist_terms``,
``mc_*_twist_grad`` and ``mc_xy_twist_hvp``; torch only holds the tensors
(on the current HIP stream) and forms per-replica sums of kernel outputs.
``helicity_modulus`` turns snapshots into the helicity-modulus estimator of
models/xy/analysis.py:50-53 through these derivatives.
"""

import ctypes

import torch

from mcgpu import _lib

if len(_lib.runtime_paths()) > 1:
    raise ImportError(
        "two HIP runtimes are loaded (" + ", ".join(_lib.runtime_paths()) + "): torch was imported after "
        "libmcgpu bound to ROCm's runtime; import torch (or autograd) before the first GPU call")

__all__ = ["lattice_energy", "twist_derivatives", "helicity_modulus"]

_KERNELS = {
    "xy": ("mc_xy_twist_terms", "mc_xy_twist_grad", "mc_xy_twist_hvp", 3),
    "kawamura": ("mc_kawamura_twist_terms", "mc_kawamura_twist_grad", None, 4),
}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _code(t):
    if t.dtype == torch.float32:
        return _lib.MC_F32
    if t.dtype == torch.float64:
        return _lib.MC_F64
    raise TypeError(f"lattice energies take float32 or float64 states, got {t.dtype}")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _shape(state, model):
    R, L = state.shape[0], state.shape[1]
    return R, L


def _terms(model, state, twist):
    """(3, R) float64: Cx, Cy, Sx."""
    R, L = _shape(state, model)
    out = torch.empty((3, R), dtype=torch.float64, device=state.device)
    _lib.call(_KERNELS[model][0], _stream(), _code(state), _ptr(state), R, L, _ptr(twist), _ptr(out))
    return out


def _coef(x, R, like):
    if x is None:
        return None
    x = torch.as_tensor(x, dtype=torch.float64, device=like.device)
    return x.expand(R).contiguous()


def _grad(model, state, twist, alpha=None, beta=None, gamma=None):
    """State gradient of alpha Cx + beta Cy + gamma Sx (per-replica coefficients)."""
    R, L = _shape(state, model)
    a, b, g = _coef(alpha, R, state), _coef(beta, R, state), _coef(gamma, R, state)
    out = torch.empty_like(state)
    _lib.call(_KERNELS[model][1], _stream(), _code(state), _ptr(state), R, L, _ptr(twist), _ptr(a),
              _ptr(b), _ptr(g), _ptr(out))
    return out


def _hvp(model, state, twist, u, alpha=None, beta=None, gamma=None):
    if _KERNELS[model][2] is None:
        raise NotImplementedError(f"second derivatives with respect to the {model} state are not implemented")
    R, L = _shape(state, model)
    a, b, g = _coef(alpha, R, state), _coef(beta, R, state), _coef(gamma, R, state)
    u = u.contiguous().to(state.dtype)
    out = torch.empty_like(state)
    _lib.call(_KERNELS[model][2], _stream(), _code(state), _ptr(state), R, L, _ptr(twist), _ptr(a), _ptr(b),
              _ptr(g), _ptr(u), _ptr(out))
    return out


def _site_sum(x, model):
    """Per-replica sum over every site dimension."""
    return x.to(torch.float64).reshape(x.shape[0], -1).sum(dim=1)


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, J, twist, model):
        terms = _terms(model, state, twist)
        ctx.model = model
        ctx.save_for_backward(state, J, twist, terms)
        return -J * (terms[0] + terms[1])

    @staticmethod
    def backward(ctx, g):
        state, J, twist, terms = ctx.saved_tensors
        need = ctx.needs_input_grad[:3]
        gs, gj, gt = _EnergyBackward.apply(g.contiguous(), state, J, twist, terms, ctx.model, need)
        return (gs if need[0] else None, gj if need[1] else None, gt if need[2] else None, None)


class _EnergyBackward(torch.autograd.Function):
    """(g, state, J, twist) -> (dH/dstate * g, sum g dH/dJ, g dH/dtwist)."""

    @staticmethod
    def forward(ctx, g, state, J, twist, terms, model, need):
        cx, cy, sx = terms[0], terms[1], terms[2]
        ctx.model = model
        ctx.save_for_backward(g, state, J, twist, terms)
        ctx.set_materialize_grads(False)  # outputs nobody differentiates arrive as None
        gs =_grad(model, state, twist, alpha=-J * g, beta=-J * g) if need[0] else state.new_zeros(())
        gj = -(g * (cx + cy)).sum()
        gt = g * (-J * sx)
        return gs, gj, gt

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, u_state, u_J, u_twist):
        g, state, J, twist, terms = ctx.saved_tensors
        model = ctx.model
        cx, cy, sx = terms[0], terms[1], terms[2]
        need_g, need_state, need_J, need_twist = ctx.needs_input_grad[:4]
        R = g.shape[0]
        uJ = torch.zeros((), dtype=torch.float64, device=g.device) if u_J is None else u_J.to(torch.float64)
        uT = torch.zeros(R, dtype=torch.float64, device=g.device) if u_twist is None else u_twist.to(torch.float64)
        use_state = u_state is not None and u_state.dim() == state.dim()

        grad_g = uJ * (-(cx + cy)) + uT * (-J * sx)
        grad_J = (uT * (-g * sx)).sum()
        grad_twist = uJ * (-g * sx) + uT * (J * g * cx)
        if use_state and (need_g or need_J or need_twist):
            d_cxy = _site_sum(u_state * _grad(model, state, twist, alpha=1.0, beta=1.0), model)
            grad_g = grad_g + (-J) * d_cxy
            grad_J = grad_J + (-(g * d_cxy)).sum()
            if need_twist:
                d_sx = _site_sum(u_state * _grad(model, state, twist, gamma=1.0), model)
                grad_twist = grad_twist + (-J) * g * d_sx
        grad_state = None
        if need_state:
            grad_state = _grad(model, state, twist, alpha=-g * uJ, beta=-g * uJ, gamma=-J * g * uT)
            if use_state:
                grad_state = grad_state + _hvp(model, state, twist, u_state, alpha=-J * g, beta=-J * g)
        return (grad_g if need_g else None, grad_state, grad_J if need_J else None,
                grad_twist if need_twist else None, None, None, None)


def _check_state(state, model):
    if model not in _KERNELS:
        raise ValueError(f"model must be one of {sorted(_KERNELS)}, got {model!r}")
    if not isinstance(state, torch.Tensor) or not state.is_cuda:
        raise ValueError("state must be a torch tensor on the GPU")
    want = 3 if model == "xy" else 4
    if state.dim() != want or state.shape[1] != state.shape[2] or (model == "kawamura" and state.shape[3] != 5):
        raise ValueError(f"invalid {model} state shape {tuple(state.shape)}")
    if state.shape[1] < 2 or state.shape[0] < 1:
        raise ValueError("need at least one replica and L >= 2")
    _code(state)


def lattice_energy(state, J=1.0, twist=0.0, model="xy"):
    """Per-replica H (float64, shape (R,)), differentiable in state, J and twist.

    state: (R, L, L) angles (XY) or (R, L, L, 5) (theta, tau, dx, dy, dz)
    (Kawamura), float32 or float64, on the GPU.  J: float or 0-d tensor.
    twist: float, 0-d tensor or (R,) tensor (radians, on the x bonds).
    """
    _check_state(state, model)
    dev = state.device
    R = state.shape[0]
    J = torch.as_tensor(J, dtype=torch.float64, device=dev)
    if J.dim() != 0:
        raise ValueError("J must be a scalar")
    twist = torch.as_tensor(twist, dtype=torch.float64, device=dev)
    if twist.dim() == 0:
        twist = twist.expand(R)
    if twist.shape != (R,):
        raise ValueError(f"twist must be a scalar or have shape ({R},)")
    return _Energy.apply(state.contiguous(), J, twist.contiguous(), model)


def twist_derivatives(state, J=1.0, model="xy"):
    """dH/dDelta, d2H/dDelta2 and dH/dJ per replica at Delta = 0, through autograd."""
    R = state.shape[0]
    twist = torch.zeros(R, dtype=torch.float64, device=state.device, requires_grad=True)
    Jt = torch.tensor(float(J), dtype=torch.float64, device=state.device, requires_grad=True)
    with torch.enable_grad():
        H = lattice_energy(state.detach(), Jt, twist, model)
        d1, dJ = torch.autograd.grad(H.sum(), (twist, Jt), create_graph=True)
        (d2,) = torch.autograd.grad(d1.sum(), twist)
    per_J = _terms(model, state.detach().contiguous(), None)
    return {"energy": H.detach(), "dH_dtwist": d1.detach(), "d2H_dtwist2": d2.detach(),
            "dH_dJ": -(per_J[0] + per_J[1]), "dH_dJ_total": dJ.detach()}


def helicity_modulus(snapshots, temperatures, J=1.0, model="xy"):
    """Helicity modulus per temperature from a sequence of snapshots.

    snapshots: iterable of (R, L, L[, 5]) states in temperature order.
    Upsilon = [<d2H/dDelta2> - (<(dH/dDelta)^2> - <dH/dDelta>^2) / T] / L^2,
    the estimator of models/xy/analysis.py:50-53 (J cx - J^2/T Var(qx)) / L^2.
    """
    first, second, second_sq, n, L = None, None, None, 0, None
    for state in snapshots:
        d = twist_derivatives(state, J, model)
        a, b = d["dH_dtwist"], d["d2H_dtwist2"]
        first = a if first is None else first + a
        second_sq = a * a if second_sq is None else second_sq + a * a
        second = b if second is None else second + b
        n += 1
        L = state.shape[1]
    if n == 0:
        raise ValueError("need at least one snapshot")
    T = torch.as_tensor(temperatures, dtype=torch.float64, device=first.device)
    mean1, mean2, mean_sq = first / n, second / n, second_sq / n
    return (mean2 - (mean_sq - mean1 * mean1) / T) / (L * L)
