"""GPU tests of the tangent-linear (forward-mode differentiated) form of the
solid path, through the C ABI (tj_tangent_*):

* values: every value of a tangent run is the FP64 form's, bit-exact against
  the oracle (probe trace and final state);
* derivatives: against central finite differences of the oracle,
  h = 1e-6 |theta| (the reference gradient; the reference itself has no
  differentiated form), the closed-form derivative of the initial energy,
  and the assembled loss gradient along a mixed direction;
* direction groups (chunk 1, 2, 3, 6) give bit-identical results;
* the torch.autograd wrapper returns the tangent gradient of the loss;
* refusals of the levels and directions outside the form.

The finite-difference bars: the probe trace is piecewise smooth (MUSCL
limiter, yield test and 5-wave branch switches), so a difference quotient
over +-h can cross a kink, and its own error is O(h^2) plus rounding over h.
Measured on the 32^2 plate impact over 50 steps: 4e-8 to 1.5e-6 of the
trace's derivative scale for the six directions; the bar is 1e-4.  (With
steps of 1e-3 theta the kinks crossed leave first-order remainders, which is
why the check is a central difference at 1e-6 and not a Taylor series.)
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, STEPS = 32, 50
FD_BAR = 1e-4


def _problem(theta=None, n=N, steps=STEPS, dim=2):
    from tenjin_amd import configs

    if dim == 2:
        return configs.plate_impact_differentiated(n=n, theta=theta, steps=steps)
    th = configs.plate_params() if theta is None else theta
    cu = dict(zip(configs.PLATE_PARAMS, th[:3]))
    al = dict(zip(configs.PLATE_PARAMS, th[3:6]))
    return configs.plate_impact(dim=3, n=n, box=n, seed=4, speed=500.0, cu=cu, al=al)


def _oracle_trace(pb, steps, probe):
    from oracle.oracle import Oracle

    o = Oracle(pb)
    comp, i, j, k = probe
    t, out = 0.0, []
    for q in range(steps):
        dt = o.driver_dt(q, t)
        o.step(dt)
        t += dt
        out.append(o.get()[comp, k, j, i])
    return np.array(out), o.get()


def _fd(p, steps, probe, dim=2, n=N, rel=1e-6):
    from tenjin_amd import configs

    th = configs.plate_params()
    h = rel * abs(th[p])
    tp, tm = list(th), list(th)
    tp[p] += h
    tm[p] -= h
    up, sp = _oracle_trace(_problem(tp, n, steps, dim), steps, probe)
    um, sm = _oracle_trace(_problem(tm, n, steps, dim), steps, probe)
    return (up - um) / (2 * h), (sp - sm) / (2 * h)


@pytest.fixture(scope="module")
def base():
    from tenjin_amd import configs, tangent

    pb = _problem()
    probe = configs.plate_probe(pb)
    print(f"\n[case] tangent plate impact 2D n={N} steps={STEPS} chunk=1", flush=True)
    tr = tangent.TangentRun(pb, tangent.ALL_DIRECTIONS, chunk=1)
    trace, done = tr.run(STEPS, probe)
    st = tr.status()
    S = tr.state()
    tr.close()
    return pb, probe, trace, done, st, S


def test_values_bit_exact_against_oracle(base):
    from tests.parity_util import exact_fraction

    pb, probe, trace, done, st, S = base
    assert done == STEPS and st == 0
    ov, So = _oracle_trace(pb, STEPS, probe)
    assert np.array_equal(trace[:, 0], ov)
    assert exact_fraction(pb, So, S) == 1.0


@pytest.mark.parametrize("p", range(6))
def test_probe_derivative_against_oracle_fd(base, p):
    from tenjin_amd import configs

    pb, probe, trace, *_ = base
    name = configs.PLATE_PARAMS[p % 3]
    fd, _ = _fd(p, STEPS, probe)
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(trace[:, 1 + p] - fd)) / scale
    print(f"\n[case] direction {p} ({name} of material {p // 3}): max|tangent - fd| / max|fd| = {err:.3e}, "
          f"scale {scale:.3e}", flush=True)
    assert scale > 0.0 and err <= FD_BAR


@pytest.mark.parametrize("chunk", [2, 3, 6])
def test_direction_groups_are_bit_identical(base, chunk):
    from tenjin_amd import tangent

    pb, probe, trace, *_ = base
    print(f"\n[case] tangent chunk={chunk}", flush=True)
    tr = tangent.TangentRun(pb, tangent.ALL_DIRECTIONS, chunk=chunk)
    t2, done = tr.run(STEPS, probe)
    tr.close()
    assert done == STEPS and np.array_equal(t2, trace)


def test_loss_gradient_along_a_mixed_direction():
    """g . v of the assembled loss gradient (2 sum_n r_n du_n/dtheta) against
    the central difference of the loss along v, a seeded mix of all six
    parameters scaled by theta, from FP64 runs of the GPU form."""
    from tenjin_amd import configs, tangent

    steps = 40
    th = np.array(configs.plate_params())
    pb = _problem(steps=steps)
    probe = configs.plate_probe(pb)
    target = tangent.forward_trace(_problem(configs.plate_target_theta(), steps=steps), steps, probe)
    print(f"\n[case] loss gradient along a mixed direction, steps={steps}", flush=True)
    loss, grad = tangent.loss_and_grad(th, target, n=N, steps=steps, chunk=2)
    v = np.random.RandomState(11).uniform(-1.0, 1.0, 6) * th
    h = 1e-6
    lp = tangent.probe_loss(tangent.forward_trace(_problem(th + h * v, steps=steps), steps, probe), target)
    lm = tangent.probe_loss(tangent.forward_trace(_problem(th - h * v, steps=steps), steps, probe), target)
    fd = (lp - lm) / (2 * h)
    gv = float(grad @ v)
    print(f"[case] g.v {gv:.9e} fd {fd:.9e} loss {loss:.6e}", flush=True)
    assert abs(gv - fd) <= 1e-4 * abs(fd)


def test_final_state_derivative_against_oracle_fd():
    """d(final state)/d(s of aluminium), every state component."""
    from tenjin_amd import configs, tangent

    steps = 30
    pb = _problem(steps=steps)
    probe = configs.plate_probe(pb)
    print(f"\n[case] final-state derivative, s of aluminium, steps={steps}", flush=True)
    tr = tangent.TangentRun(pb, [(1, "s")])
    tr.run(steps, probe)
    dS = tr.state_derivative(0)
    S = tr.state()
    tr.close()
    _, fdS = _fd(4, steps, probe)
    h = 1e-6 * configs.plate_params()[4]
    L = pb.layout
    worst = 0.0
    for c in L.state_comps():
        scale = np.max(np.abs(fdS[c]))
        # rounding of the two runs' values over the difference step
        noise = 8.0 * np.finfo(float).eps * np.max(np.abs(S[c])) / h
        if scale <= noise:
            assert np.max(np.abs(dS[c])) <= 2.0 * noise, L.names[c]
            continue
        err = np.max(np.abs(dS[c] - fdS[c])) / scale
        worst = max(worst, err)
        assert err <= FD_BAR + noise / scale, (L.names[c], err)
    print(f"[case] worst component {worst:.3e}", flush=True)


@pytest.mark.parametrize("m", [0, 1])
def test_initial_energy_derivative_closed_form(m):
    """After setInitialConditions (no step): dE/ds against the closed form of
    the Us-up EOS, and dE/dGamma0 = 0 (the pressure shift cancels it)."""
    from tenjin_amd import configs, tangent
    from tests.test_tangent import _closed_form_dE_ds

    pb = _problem(n=24)
    print(f"\n[case] initial dE/ds, dE/dGamma0 of material {m}", flush=True)
    tr = tangent.TangentRun(pb, [(m, "s"), (m, "Gamma0")], chunk=2)
    tr.run(0, configs.plate_probe(pb))
    cE = pb.layout.comp(configs.TOTAL_E)
    dE_ds = tr.state_derivative(0)[cE]
    dE_dg = tr.state_derivative(1)[cE]
    E = tr.state()[cE]
    tr.close()
    ref = _closed_form_dE_ds(pb, m)
    assert np.max(np.abs(dE_ds - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.max(np.abs(dE_dg)) <= 1e-12 * np.max(np.abs(E))


def _oracle_run_to(pb, final_time, probe):
    """The driver loop of main.cpp:121-191 to a final time (dt clipped to it)."""
    from oracle.oracle import Oracle

    o = Oracle(pb)
    comp, i, j, k = probe
    t, n, out = 0.0, 0, []
    while t < final_time:
        dt = o.driver_dt(n, t, final_time)
        o.step(dt)
        t += dt
        n += 1
        out.append(o.get()[comp, k, j, i])
    return np.array(out), o.get(), n


def test_final_time_clip_derivative():
    """A run to a fixed final time: the last dt is clipped to it, so its
    derivative is minus the accumulated derivative of t, and the final state's
    derivative is the one at that fixed time (central differences of the
    oracle's runs to the same final time)."""
    from tenjin_amd import configs, tangent

    from oracle.oracle import Oracle

    pb = _problem(steps=40)
    probe = configs.plate_probe(pb)
    # a final time halfway through step 26 of the unclipped run
    o = Oracle(pb)
    t = 0.0
    for q in range(25):
        dt = o.driver_dt(q, t)
        o.step(dt)
        t += dt
    T = t + 0.5 * o.driver_dt(25, t)
    print(f"\n[case] tangent run to final time {T:.6e}", flush=True)
    tr = tangent.TangentRun(pb, [(1, "s"), (0, "Gamma0")], chunk=2)
    trace, done = tr.run(40, probe, final_time=T)
    dS = [tr.state_derivative(0), tr.state_derivative(1)]
    S = tr.state()
    tr.close()
    ov, So, n_o = _oracle_run_to(pb, T, probe)
    assert done == n_o == 26
    assert np.array_equal(trace[:done, 0], ov)
    from tests.parity_util import exact_fraction

    assert exact_fraction(pb, So, S) == 1.0
    th = configs.plate_params()
    L = pb.layout
    for col, p in ((0, 4), (1, 0)):
        h = 1e-6 * abs(th[p])
        st = []
        for sgn in (1.0, -1.0):
            tp = list(th)
            tp[p] += sgn * h
            _, Sx, nx_ = _oracle_run_to(_problem(tp, steps=40), T, probe)
            assert nx_ == n_o
            st.append(Sx)
        fd = (st[0] - st[1]) / (2 * h)
        for c in (L.comp(configs.TOTAL_E), L.comp(configs.RHOU, 0, 0), L.comp(configs.P)):
            scale = np.max(np.abs(fd[c]))
            err = np.max(np.abs(dS[col][c] - fd[c])) / scale
            print(f"[case] direction {p} comp {L.names[c]}: {err:.3e}", flush=True)
            assert err <= FD_BAR


def test_reflective_boundaries():
    """Reflective x-low and y-high faces (mirrored states, flipped normal
    velocity): values bit-exact, derivatives against the oracle."""
    from tenjin_amd import configs, tangent

    def refl(theta=None):
        pb = _problem(theta, steps=40)
        pb.bc_lo = (1, 0, 0)
        pb.bc_hi = (0, 1, 0)
        return pb

    steps = 40
    pb = refl()
    comp = pb.layout.comp(configs.VELOCITY, 0, 0)
    probe = (comp, 3, N - 2, 0)  # next to both reflective faces
    print(f"\n[case] tangent with reflective x-low / y-high, steps={steps}", flush=True)
    tr = tangent.TangentRun(pb, [(0, "s"), (1, "Gamma0")], chunk=2)
    trace, done = tr.run(steps, probe)
    tr.close()
    ov, _ = _oracle_trace(pb, steps, probe)
    assert done == steps and np.array_equal(trace[:, 0], ov)
    th = configs.plate_params()
    for col, p in ((1, 1), (2, 3)):
        h = 1e-6 * abs(th[p])
        us = []
        for sgn in (1.0, -1.0):
            tp = list(th)
            tp[p] += sgn * h
            us.append(_oracle_trace(refl(tp), steps, probe)[0])
        fd = (us[0] - us[1]) / (2 * h)
        err = np.max(np.abs(trace[:, col] - fd)) / np.max(np.abs(fd))
        print(f"[case] reflective direction {p}: {err:.3e}", flush=True)
        assert err <= FD_BAR


def test_three_dimensional_probe_derivative():
    from tenjin_amd import configs, tangent

    n, steps = 12, 16
    pb = _problem(n=n, steps=steps, dim=3)
    nz, ny, nx = pb.shape
    probe = (pb.layout.comp(configs.VELOCITY, 0, 0), nx // 2 + 2, ny // 2, nz // 2)
    print(f"\n[case] tangent plate impact 3D n={n} steps={steps}", flush=True)
    tr = tangent.TangentRun(pb, [(1, "s"), (0, "Gamma0")], chunk=2)
    trace, done = tr.run(steps, probe)
    tr.close()
    assert done == steps
    ov, _ = _oracle_trace(pb, steps, probe)
    assert np.array_equal(trace[:, 0], ov)
    for col, p in ((1, 4), (2, 0)):
        fd, _ = _fd(p, steps, probe, dim=3, n=n)
        err = np.max(np.abs(trace[:, col] - fd)) / np.max(np.abs(fd))
        print(f"[case] 3D direction {p}: {err:.3e}", flush=True)
        assert err <= FD_BAR


def test_full_size_run_has_finite_derivatives():
    """The differentiated plate impact at full size (512^2, 200 steps): ahead of the wave the state's
    differences decay to the subnormals, where the MUSCL ratio's derivative
    must stay finite (no TJ_ST_TANGENT_NAN)."""
    from tenjin_amd import configs, tangent

    pb = configs.plate_impact_differentiated(n=512, steps=200)
    print("\n[case] tangent run 512^2, 200 steps, two directions", flush=True)
    tr = tangent.TangentRun(pb, [(0, "Y"), (1, "s")], chunk=2)
    trace, done = tr.run(200, configs.plate_probe(pb))
    st = tr.status()
    dS = tr.state_derivative(1)
    tr.close()
    assert done == 200 and st == 0
    assert np.all(np.isfinite(trace)) and np.all(np.isfinite(dS[pb.layout.state_comps()]))


def test_autograd_wrapper():
    import torch

    from tenjin_amd import configs, tangent

    steps = 20
    pb = _problem(steps=steps)
    probe = configs.plate_probe(pb)
    target = tangent.forward_trace(_problem(configs.plate_target_theta(), steps=steps), steps, probe)
    theta = torch.tensor(configs.plate_params(), dtype=torch.float64, requires_grad=True)
    loss = tangent.probe_loss_torch(theta, target, n=N, steps=steps, chunk=3)
    loss.backward()
    ref_loss, ref_grad = tangent.loss_and_grad(configs.plate_params(), target, n=N, steps=steps, chunk=1)
    u = tangent.forward_trace(pb, steps, probe)
    assert float(loss) == ref_loss == tangent.probe_loss(u, target)
    assert np.array_equal(theta.grad.numpy(), ref_grad)
    assert np.all(np.isfinite(ref_grad)) and np.any(ref_grad != 0.0)


def test_libtorch_autograd_function():
    """The C++ torch::autograd::Function gives the loss and gradient of the
    Python path (same tangent run; the sums differ in order only)."""
    import torch

    from tenjin_amd import configs, tangent

    steps = 20
    pb = _problem(steps=steps)
    probe = configs.plate_probe(pb)
    target = tangent.forward_trace(_problem(configs.plate_target_theta(), steps=steps), steps, probe)
    theta = torch.tensor(configs.plate_params(), dtype=torch.float64, requires_grad=True)
    print("\n[case] libtorch ProbeLoss", flush=True)
    loss = tangent.probe_loss_libtorch(theta, target, n=N, steps=steps, chunk=2)
    (3.0 * loss).backward()
    ref_loss, ref_grad = tangent.loss_and_grad(configs.plate_params(), target, n=N, steps=steps, chunk=1)
    assert abs(float(loss) - ref_loss) <= 1e-12 * ref_loss
    assert np.allclose(theta.grad.numpy(), 3.0 * ref_grad, rtol=1e-10, atol=1e-12 * np.max(np.abs(3.0 * ref_grad)))


def test_autograd_theta_on_gpu_any_shape():
    """Both autograd Functions accept theta on the GPU and of shape [2, 3]
    (per material: Gamma0, s, Y): theta.grad has theta's device and shape
    and equals loss_and_grad's gradient."""
    import torch

    from tenjin_amd import configs, tangent

    steps = 8
    pb = _problem(steps=steps)
    probe = configs.plate_probe(pb)
    target = tangent.forward_trace(_problem(configs.plate_target_theta(), steps=steps), steps, probe)
    ref_loss, ref_grad = tangent.loss_and_grad(configs.plate_params(), target, n=N, steps=steps, chunk=1)
    for name, fn in (("python", lambda th: tangent.probe_loss_torch(th, target, n=N, steps=steps, chunk=3)),
                     ("libtorch", lambda th: tangent.probe_loss_libtorch(th, target, n=N, steps=steps, chunk=3))):
        print(f"\n[case] theta on cuda, shape [2, 3], {name}", flush=True)
        theta = torch.tensor(configs.plate_params(), dtype=torch.float64, device="cuda").view(2, 3)
        theta.requires_grad_(True)
        loss = fn(theta)
        assert loss.device == theta.device
        (2.0 * loss).backward()
        assert theta.grad.shape == (2, 3) and theta.grad.device == theta.device
        got = theta.grad.detach().cpu().numpy().reshape(-1)
        assert abs(float(loss.detach().cpu()) - ref_loss) <= 1e-12 * ref_loss
        assert np.allclose(got, 2.0 * ref_grad, rtol=1e-10, atol=1e-12 * np.max(np.abs(ref_grad)))


def _fluid_fd(make, m, field, rel, steps, probe):
    """Central difference of the oracle's probe trace in material m's
    `field` ("gamma" = adiabaticIndex, i.e. Gamma0 + 1, or "pref")."""
    pb0 = make()
    v0 = getattr(pb0.materials[m], field)
    h = rel * abs(v0)
    us = []
    for sgn in (1.0, -1.0):
        pb = make()
        setattr(pb.materials[m], field, v0 + sgn * h)
        us.append(_oracle_trace(pb, steps, probe)[0])
    return (us[0] - us[1]) / (2 * h)


def _tube_thinc():
    from tenjin_amd import configs

    pb = configs.shock_tube_1d(200, 0.7, 64)  # 64-cell boxes: box-edge history planes
    pb.THINC = 1
    return pb


@pytest.mark.parametrize("case", ["tube1d", "inclusion2d", "tube1d_thinc", "inclusion2d_thinc"])
def test_fluid_derivatives_against_oracle_fd(case):
    """The fluid path (MUSCL, or MUSCL + THINC/BVD with its history planes,
    and HLLC): values bit-exact, derivatives with respect to the Gruneisen
    coefficients and the water's pref against the oracle."""
    from tenjin_amd import configs, tangent

    if case.startswith("tube1d"):
        make = _tube_thinc if case.endswith("thinc") else (lambda: configs.shock_tube_1d(200))
        steps = 60
        pb = make()
        probe = (pb.layout.comp(configs.P), 150, 0, 0)
        dirs = [(0, "Gamma0"), (0, "pref"), (1, "Gamma0")]
    else:
        th = 1 if case.endswith("thinc") else 0
        make = lambda: configs.shock_inclusion_2d(nx=64, ny=32, box=16, thinc=th)  # noqa: E731
        steps = 40
        pb = make()
        probe = (pb.layout.comp(configs.VELOCITY, 0, 0), 26, 13, 0)
        dirs = [(0, "Gamma0"), (1, "Gamma0")]
    print(f"\n[case] fluid tangent {case}, steps={steps}", flush=True)
    tr = tangent.TangentRun(pb, dirs, chunk=len(dirs))
    trace, done = tr.run(steps, probe)
    st = tr.status()
    S = tr.state()
    tr.close()
    ov, So = _oracle_trace(pb, steps, probe)
    from tests.parity_util import exact_fraction

    assert done == steps and st == 0
    assert np.array_equal(trace[:, 0], ov) and exact_fraction(pb, So, S) == 1.0
    for col, (m, par) in enumerate(dirs, start=1):
        fd = _fluid_fd(make, m, "gamma" if par == "Gamma0" else "pref", 1e-6, steps, probe)
        err = np.max(np.abs(trace[:, col] - fd)) / np.max(np.abs(fd))
        print(f"[case] {case} material {m} {par}: {err:.3e} (scale {np.max(np.abs(fd)):.3e})", flush=True)
        assert err <= FD_BAR


def test_thinc_direction_groups_are_bit_identical():
    from tenjin_amd import configs, tangent

    pb = configs.shock_inclusion_2d(nx=64, ny=32, box=16, thinc=1)
    probe = (pb.layout.comp(configs.VELOCITY, 0, 0), 26, 13, 0)
    dirs = [(0, "Gamma0"), (1, "Gamma0"), (0, "pref")]
    print("\n[case] THINC tangent chunk 1 vs 3", flush=True)
    out = []
    for chunk in (1, 3):
        tr = tangent.TangentRun(pb, dirs, chunk=chunk)
        out.append(tr.run(30, probe)[0])
        tr.close()
    assert np.array_equal(out[0], out[1])


def test_refusals():
    from tenjin_amd import TenjinError, configs, tangent

    with pytest.raises(TenjinError, match="mixture"):
        tangent.TangentRun(configs.reactive_tube(1, 64), [(1, "Gamma0")])
    with pytest.raises(TenjinError, match="Us-up"):
        tangent.TangentRun(configs.shock_tube_1d(64), [(0, "s")])
    with pytest.raises(TenjinError, match="pref"):
        tangent.TangentRun(_problem(n=16), [(0, "pref")])
    with pytest.raises(TenjinError, match="geometric source"):
        tangent.TangentRun(configs.rod_impact_axisymmetric(16, 16), [(0, "s")])
    elastic = configs.plate_impact(dim=2, n=16, box=16, plastic=False)
    with pytest.raises(TenjinError, match="plastic"):
        tangent.TangentRun(elastic, [(0, "Y")])
    with pytest.raises(TenjinError, match="chunk"):
        tangent.TangentRun(_problem(n=16), [(0, "s")], chunk=7)
    tr = tangent.TangentRun(_problem(n=16), [(0, "s")])
    with pytest.raises(TenjinError, match="probe"):
        tr.run(2, (10, 16, 0, 0))
    tr.vP0 = tr.g.vS  # the initial primitives must not alias the state
    with pytest.raises(TenjinError, match="distinct"):
        tr.run(2, (10, 3, 3, 0))
    tr.close()
