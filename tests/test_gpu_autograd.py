"""GPU checks of the differentiable XY / Kawamura energies (HIP backward kernels
behind torch.autograd).

Pinned by what the mathematics fixes, not by a restated formula:
* central finite differences (h = 1e-4, float64) of H along random state
  directions, in the twist and in J, and second differences in the twist;
* torch.autograd.gradcheck / gradgradcheck (numerical Jacobians, float64);
* closed forms against the drop-in measurement kernels: at Delta = 0,
  H = energy, dH/dDelta = -J twist_sin_x, d2H/dDelta2 = J twist_cos_x;
* the helicity modulus from autograd equals the reference estimator
  (models/xy/analysis.py:50-53) on the same snapshots;
* float32 gradients agree with float64 ones at L = 256, R = 64.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda", 0)


def _xy(R, L, dtype, seed):
    g = np.random.default_rng(seed)
    return torch.tensor(g.uniform(-np.pi, np.pi, (R, L, L)), dtype=dtype, device=_dev())


def _kawamura(R, L, dtype, seed):
    g = np.random.default_rng(seed)
    d = g.normal(size=(R, L, L, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    s = np.concatenate([g.uniform(-np.pi, np.pi, (R, L, L, 1)), g.choice([-1.0, 1.0], (R, L, L, 1)), d], -1)
    return torch.tensor(s, dtype=dtype, device=_dev())


def _state(model, R, L, dtype, seed):
    return _xy(R, L, dtype, seed) if model == "xy" else _kawamura(R, L, dtype, seed)


def test_xy_and_kawamura_closed_forms_against_measurement_kernels(gpu):
    from autograd import twist_derivatives
    from mcgpu import to_device
    from models.kawamura.model import KawamuraModel
    from models.xy.model import XYModel
    J = 1.3
    for model, cls in (("xy", XYModel), ("kawamura", KawamuraModel)):
        state = _state(model, 3, 16, torch.float64, 5)
        m = cls(J=J, dtype=np.float64).measure(to_device(state.cpu().numpy()))
        d = twist_derivatives(state, J, model)
        np.testing.assert_allclose(d["energy"].cpu().numpy(), m["energy"], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(d["dH_dtwist"].cpu().numpy(), -J * m["twist_sin_x"], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(d["d2H_dtwist2"].cpu().numpy(), J * m["twist_cos_x"], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(d["dH_dJ"].cpu().numpy(), m["energy"] / J, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("model,R,L", [("xy", 3, 16), ("kawamura", 2, 12)])
def test_central_differences_fp64(gpu, model, R, L):
    from autograd import lattice_energy
    h = 1e-4
    J = 0.9
    state = _state(model, R, L, torch.float64, 21).requires_grad_(True)
    twist = torch.tensor([0.0, 0.3, -0.7][:R], dtype=torch.float64, device=_dev(), requires_grad=True)
    Jt = torch.tensor(J, dtype=torch.float64, device=_dev(), requires_grad=True)
    H = lattice_energy(state, Jt, twist, model)
    gs, gj, gt = torch.autograd.grad(H.sum(), (state, Jt, twist), create_graph=True)
    (g2,) = torch.autograd.grad(gt.sum(), twist)
    with torch.no_grad():
        v = torch.randn(state.shape, dtype=torch.float64, device=_dev(), generator=torch.Generator(_dev()).manual_seed(3))
        fd = (lattice_energy(state + h * v, J, twist, model) - lattice_energy(state - h * v, J, twist, model)) / (2 * h)
        an = (gs * v).reshape(R, -1).sum(1)
        assert torch.allclose(fd, an, rtol=1e-7, atol=1e-7), (fd - an).abs().max()
        H0 = lattice_energy(state, J, twist, model)
        Hp = lattice_energy(state, J, twist + h, model)
        Hm = lattice_energy(state, J, twist - h, model)
        assert torch.allclose((Hp - Hm) / (2 * h), gt, rtol=1e-7, atol=1e-7)
        assert torch.allclose((Hp - 2 * H0 + Hm) / h**2, g2, rtol=1e-5, atol=1e-4)
        assert torch.allclose(gj, (H0 / J).sum(), rtol=1e-12)


def test_gradcheck_and_gradgradcheck(gpu):
    from autograd import lattice_energy
    th = _xy(2, 4, torch.float64, 8).requires_grad_(True)
    J = torch.tensor(1.1, dtype=torch.float64, device=_dev(), requires_grad=True)
    tw = torch.tensor([0.2, -0.4], dtype=torch.float64, device=_dev(), requires_grad=True)
    f = lambda s, j, t: lattice_energy(s, j, t, "xy")  # noqa: E731
    assert torch.autograd.gradcheck(f, (th, J, tw))
    assert torch.autograd.gradgradcheck(f, (th, J, tw))
    ks = _kawamura(2, 4, torch.float64, 9).requires_grad_(True)
    fk = lambda s, j, t: lattice_energy(s, j, t, "kawamura")  # noqa: E731
    assert torch.autograd.gradcheck(fk, (ks, J, tw))
    fixed = ks.detach()
    assert torch.autograd.gradgradcheck(lambda j, t: lattice_energy(fixed, j, t, "kawamura"), (J, tw))
    # second derivatives through the Kawamura state are not implemented: loud, not silent
    H = lattice_energy(ks, J, tw, "kawamura")
    (g,) = torch.autograd.grad(H.sum(), ks, create_graph=True)
    with pytest.raises(NotImplementedError):
        torch.autograd.grad((g * g).sum(), ks)


@pytest.mark.parametrize("model", ["xy", "kawamura"])
def test_l256_r64_fp32_and_fp64(gpu, model):
    """The differentiable-energy workload: L = 256, 64 replicas, both precisions."""
    from autograd import lattice_energy
    R, L, h = 64, 256, 1e-4
    s64 = _state(model, R, L, torch.float64, 44).requires_grad_(True)
    H64 = lattice_energy(s64, 1.0, 0.0, model)
    (g64,) = torch.autograd.grad(H64.sum(), s64)
    with torch.no_grad():
        v = torch.randn(s64.shape, dtype=torch.float64, device=_dev(), generator=torch.Generator(_dev()).manual_seed(1))
        fd = (lattice_energy(s64 + h * v, 1.0, 0.0, model) - lattice_energy(s64 - h * v, 1.0, 0.0, model)) / (2 * h)
        an = (g64 * v).reshape(R, -1).sum(1)
        assert torch.allclose(fd, an, rtol=1e-6, atol=1e-5), (fd - an).abs().max()
    s32 = s64.detach().to(torch.float32).requires_grad_(True)
    H32 = lattice_energy(s32, 1.0, 0.0, model)
    (g32,) = torch.autograd.grad(H32.sum(), s32)
    assert g32.dtype == torch.float32
    assert torch.allclose(H32, H64.detach(), rtol=1e-5, atol=1e-2)
    assert (g32.double() - g64).abs().max() < 1e-4


def test_helicity_modulus_equals_reference_estimator(gpu):
    from autograd import helicity_modulus
    from core.runtime import Runtime
    from models.xy.model import XYModel
    temps = np.linspace(0.6, 1.4, 6)
    J = 1.0
    rt = Runtime(XYModel(J=J, dtype=np.float64, start="random"), 16, temps, seed=58)
    for _ in range(300):
        rt.sweep()
    snaps, cx, qx = [], [], []
    for _ in range(60):
        for _ in range(5):
            rt.sweep()
        snaps.append(torch.from_numpy(rt.snapshot()).to(_dev()))
        m = rt.measure()
        cx.append(m["twist_cos_x"])
        qx.append(m["twist_sin_x"])
    cx, qx = np.array(cx), np.array(qx)
    ref = (J * cx.mean(0) - J**2 / temps * (np.mean(qx**2, 0) - qx.mean(0) ** 2)) / 16**2
    got = helicity_modulus(snaps, temps, J, "xy").cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)
    assert got[0] > got[-1]  # stiff at low T, soft above the BKT range


def test_autograd_on_a_side_stream_alongside_drop_in_measurements(gpu):
    """Reductions on torch's side stream and on the drop-in path's stream use
    separate partial buffers: results equal the sequential ones."""
    from autograd import lattice_energy
    from core.runtime import Runtime
    from models.xy.model import XYModel
    rt = Runtime(XYModel(start="random"), 64, np.linspace(0.8, 1.0, 8), seed=11)
    alone = rt.measure()
    state = _xy(16, 128, torch.float64, 5)
    J = torch.tensor(1.0, dtype=torch.float64, device=_dev())
    twist = torch.zeros(16, dtype=torch.float64, device=_dev())
    expected = lattice_energy(state, J, twist, "xy").cpu()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        got = [lattice_energy(state, J, twist, "xy") for _ in range(20)]
    measured = [rt.measure() for _ in range(20)]  # queued while the side stream works
    side.synchronize()
    for g in got:
        torch.testing.assert_close(g.cpu(), expected, rtol=0, atol=0)
    for m in measured:
        for k in alone:
            np.testing.assert_array_equal(m[k], alone[k])


def test_float32_terms_accurate_for_angles_of_many_turns(gpu):
    """The float32 kernels reduce whole turns before their bounded sine and
    cosine: angles spread over +-1000 rad give the float64 energy and
    gradient to float32 accuracy (the inputs' own rounding included)."""
    from autograd import lattice_energy
    g = np.random.default_rng(17)
    host = g.uniform(-1000, 1000, (8, 64, 64)).astype(np.float32)
    s32 = torch.tensor(host, device=_dev()).requires_grad_(True)
    s64 = torch.tensor(host.astype(np.float64), device=_dev()).requires_grad_(True)
    H32 = lattice_energy(s32, 1.0, 0.3, "xy")
    H64 = lattice_energy(s64, 1.0, 0.3, "xy")
    (g32,) = torch.autograd.grad(H32.sum(), s32)
    (g64,) = torch.autograd.grad(H64.sum(), s64)
    assert torch.allclose(H32, H64, rtol=0, atol=2e-2), (H32 - H64).abs().max()
    assert (g32.double() - g64).abs().max() < 3e-4


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_an_unaligned_state_view_takes_the_scalar_kernels(gpu, dtype):
    """The band kernels read and write rows in 16-byte vectors; a state view
    that starts one element past a 16-byte boundary takes the whole-row
    kernels' scalar accesses instead, with the same energy and gradient as the
    aligned copy (to each kernel pair's rounding)."""
    from autograd import lattice_energy
    R, L = 3, 128
    base = _xy(R, L, dtype, 5)
    flat = torch.empty(R * L * L + 1, dtype=dtype, device=_dev())
    flat[1:] = base.reshape(-1)
    view = flat[1:].view(R, L, L)
    assert view.data_ptr() % 16 != 0 and base.data_ptr() % 16 == 0
    out = []
    for s in (base, view):
        x = s.detach().requires_grad_(True)
        H = lattice_energy(x, 0.8, torch.tensor([0.0, 0.3, -1.1], dtype=torch.float64, device=_dev()), "xy")
        (g,) = torch.autograd.grad(H.sum(), x)
        out.append((H.detach().cpu().numpy(), g.double().cpu().numpy()))
    tol_h, tol_g = (1e-3 * L, 2e-5) if dtype == torch.float32 else (1e-10, 1e-12)
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=0, atol=tol_h)
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=0, atol=tol_g)


def test_float64_sincos_of_the_kernels_at_large_angles(gpu):
    """The float64 kernels' own sine and cosine (a three-part reduction by pi/2
    and fdlibm's polynomials) at angles spread over +-1e6 rad: the energy and
    gradient equal a numpy float64 evaluation of the same bond identities from
    each site's np.cos and np.sin to 1e-12, so only the sine and cosine are
    compared.  An angle of 2^51 rad (past the reduction's range) gives NaN."""
    from autograd import lattice_energy
    g = np.random.default_rng(23)
    R, L = 4, 64
    host = g.uniform(-1e6, 1e6, (R, L, L))
    tw = np.array([0.0, 0.3, -1.1, 2.0])
    th = torch.tensor(host, device=_dev()).requires_grad_(True)
    H = lattice_energy(th, 1.0, torch.tensor(tw, device=_dev()), "xy")
    (gr,) = torch.autograd.grad(H.sum(), th)
    c, s = np.cos(host), np.sin(host)
    cd, sd = np.cos(tw)[:, None, None], np.sin(tw)[:, None, None]

    def bond(c1, s1, c2, s2):  # cos and sin of theta_1 - theta_2
        return c1 * c2 + s1 * s2, s1 * c2 - c1 * s2

    cxa, sxa = bond(c, s, np.roll(c, -1, 1), np.roll(s, -1, 1))  # (i, j) - (i+1, j)
    cx, sx = cxa * cd + sxa * sd, sxa * cd - cxa * sd             # twisted by -Delta
    cy, sy = bond(c, s, np.roll(c, -1, 2), np.roll(s, -1, 2))     # (i, j) - (i, j+1)
    H_ref = -(cx.sum((1, 2)) + cy.sum((1, 2)))
    g_ref = -(-sx + np.roll(sx, 1, 1) - sy + np.roll(sy, 1, 2))
    np.testing.assert_allclose(H.detach().cpu().numpy(), H_ref, rtol=0, atol=1e-12 * L * L)
    np.testing.assert_allclose(gr.cpu().numpy(), g_ref, rtol=0, atol=1e-12)
    host[1, 5, 7] = 2.0 ** 51
    H2 = lattice_energy(torch.tensor(host, device=_dev()), 1.0, 0.0, "xy").cpu().numpy()
    assert np.isnan(H2[1]) and np.isfinite(H2[[0, 2, 3]]).all()


@pytest.mark.parametrize("L", [64, 128, 256, 1024])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_whole_row_xy_kernels_against_closed_forms(gpu, L, dtype):
    """The XY kernels of whole rows against the bond sums and their gradient
    evaluated in numpy float64, with a twist per replica (zero for replica 0,
    which takes the gradient's zero-twist loop): L = 64, 128, 256 run the band
    kernels (one wavefront per band of rows, 1, 2 or 4 columns per lane, the
    periodic row closed by DPP rotates), L = 1024 the whole-row workgroups (16
    wavefronts per row, the bonds across wavefront edges from LDS)."""
    from autograd import lattice_energy
    R = 3
    th64 = _xy(R, L, torch.float64, 91 + L)
    tw = torch.tensor([0.0, 0.3, -1.1], dtype=torch.float64, device=_dev())
    J = 0.8
    th = th64.to(dtype).requires_grad_(True)
    H = lattice_energy(th, J, tw, "xy")
    (g,) = torch.autograd.grad(H.sum(), th)
    t = th.detach().double().cpu().numpy()
    d = tw.cpu().numpy()[:, None, None]
    ax = t - np.roll(t, -1, axis=1) - d  # (i, j) - (i+1, j), twisted
    ay = t - np.roll(t, -1, axis=2)      # (i, j) - (i, j+1)
    H_ref = -J * (np.cos(ax).sum((1, 2)) + np.cos(ay).sum((1, 2)))
    g_ref = -J * (-np.sin(ax) + np.roll(np.sin(ax), 1, axis=1) - np.sin(ay) + np.roll(np.sin(ay), 1, axis=2))
    if dtype == torch.float64:
        np.testing.assert_allclose(H.detach().cpu().numpy(), H_ref, rtol=1e-12)
        np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=0, atol=1e-12)
    else:
        np.testing.assert_allclose(H.detach().cpu().numpy(), H_ref, rtol=0, atol=1e-3 * L)
        np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=0, atol=2e-5)
