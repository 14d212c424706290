"""The adjoint kernel (st_adjoint) on the GPU: gradients of an ensemble spin
loss with respect to the 15 field parameters (gradient coefficients, E, B0)
against central finite differences of the forward path on the GPU and of
the CPU oracle at the same seed, and the torch.autograd binding."""
import numpy as np
import pytest

import spintrack as st
from spintrack import adjoint, configs
from conftest import to_oracle_opt

pytestmark = pytest.mark.gpu

# scale of each parameter for finite-difference steps where its base value is 0
SCALE = np.array([1e-8] * 9 + [7.5e6] * 3 + [3e-6] * 3)


def workload(n=128, tf=1.0, ioutInt=0.1, diffuse=0.0, seed=4, fixed_h=0.0):
    o = configs.neutron_cell(n=n, tf=tf, ioutInt=ioutInt, seed=seed)
    theta = 1e-8 * (2.0 * st.rng.uniforms(seed, 0, 9) - 1.0)
    st.set_fields(o, yi=(0.0, 1.0, 0.0), diffuse=diffuse, Gx=tuple(theta[0:3]), Gy=tuple(theta[3:6]),
                  Gz=tuple(theta[6:9]))
    if fixed_h:
        # Fixed steps make the forward map smooth in theta: with the adaptive
        # controller, accept/reject decisions move with theta and leave a
        # ~1e-13 noise floor in the loss, above the 1e-6 resolution of the
        # weakly coupled parameters' differences.  hmin = h keeps the steps
        # at h: the reference carries a segment's truncated last step into
        # the next segment (src/integrator.cpp:137-140, 244-245), and the hmin clamp
        # restores it.
        st.set_fields(o, fixedStepSize=True, h=fixed_h, hmin=fixed_h)
    return o


def gpu_loss(opt, theta, variant=st.VARIANT_FAST):
    r = adjoint.ForwardRun(adjoint.with_params(opt, theta), variant=variant, chunk_records=16)
    try:
        m = r.run()
        return adjoint.phase_loss(m, r.times, opt.t0)[0]
    finally:
        r.close()


def oracle_loss(O, opt, theta):
    o = adjoint.with_params(opt, theta)
    E = O.Ensemble(to_oracle_opt(o), o.numParticles)
    prev, means = o.t0, []
    times = st.output_times(o)
    for t in times:
        E.advance_to(t, prev_t=prev)
        prev = t
        s = O.xyz(E.snapshot()["S"])
        means.append(s.mean(axis=0))
    return adjoint.phase_loss(np.array(means), np.array(times), o.t0)[0]


def fd_steps(theta):
    return np.where(theta != 0.0, 1e-3 * np.abs(theta), 1e-3 * SCALE)


def central_fd(f, theta, scale=1.0):
    g = np.zeros(15)
    d = scale * fd_steps(theta)
    for p in range(15):
        e = np.zeros(15)
        e[p] = d[p]
        g[p] = (f(theta + e) - f(theta - e)) / (2 * d[p])
        print(f"[fd] step x{scale:g} {st._lib.GRAD_PARAMS[p]} {g[p]:.10e}", flush=True)
    return g


def resolution(L, theta, fd):
    """Relative resolution of richardson_fd's extrapolated difference.  With
    each loss evaluation good to e = 2 ulp, a central difference at step d
    is good to e/d, so (4*g(d/2) - g(d))/3 is good to (4*2e/d + e/d)/3 =
    3e/d, relative to |dL/dtheta| (the weakest coupling here moves the loss,
    some tens of rad, by ~2e-9 rad over 2d)."""
    return 3.0 * 2.0 * np.spacing(abs(L)) / (fd_steps(theta) * np.abs(fd))


def richardson_fd(f, theta):
    """Central differences at steps d and d/2, extrapolated: removes the
    O(d^2) truncation term, which dominates for the B0 components transverse
    to the precession axis (they tilt it: second- and third-order effects
    scaled by the large sensitivity to the field magnitude)."""
    g1 = central_fd(f, theta)
    g2 = central_fd(f, theta, 0.5)
    return (4.0 * g2 - g1) / 3.0, g1


def adjoint_gradient(opt, variant=st.VARIANT_FAST, chunk_records=16, adjoint_window=0):
    r = adjoint.ForwardRun(opt, variant=variant, chunk_records=chunk_records, adjoint_window=adjoint_window)
    try:
        m = r.run()
        L, seeds = adjoint.phase_loss(m, r.times, opt.t0)
        return L, r.gradient(seeds)
    finally:
        r.close()


def report(name, g, fd):
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)
    for p in range(15):
        print(f"[{name}] {st._lib.GRAD_PARAMS[p]:5s} adj {g[p]: .10e} fd {fd[p]: .10e} rel {rel[p]:.2e}", flush=True)
    return rel


def test_adjoint_matches_gpu_finite_differences():
    # fixed steps of 2e-5 s (omega*h = 0.011): all 15 parameters to 1e-6
    opt = workload(fixed_h=2e-5)
    theta = adjoint.params_of(opt)
    L, g = adjoint_gradient(opt)
    fd, fd_raw = richardson_fd(lambda th: gpu_loss(opt, th), theta)
    report("gpu-fd raw", g, fd_raw)
    rel = report("gpu-fd", g, fd)
    assert np.all(rel < np.maximum(1e-6, resolution(L, theta, fd))), rel


def test_adjoint_matches_oracle_finite_differences(oracle):
    # the CPU reference path (the oracle, exact arithmetic) differentiated by
    # finite differences at the same seed, fixed steps of 5e-5 s
    opt = workload(n=64, fixed_h=5e-5)
    theta = adjoint.params_of(opt)
    L, g = adjoint_gradient(opt)
    fd, _ = richardson_fd(lambda th: oracle_loss(oracle, opt, th), theta)
    rel = report("cpu-fd", g, fd)
    assert np.all(rel < np.maximum(1e-6, resolution(L, theta, fd))), rel


def test_adjoint_adaptive_steps_within_difference_noise():
    # adaptive DOP853 at rtol = 1e-12: differences carry a noise floor of
    # ~1e-12 in the loss (accept/reject decisions move with theta), so each
    # parameter must agree to 1e-6 or to that floor over its loss change
    opt = workload()
    theta = adjoint.params_of(opt)
    _, g = adjoint_gradient(opt)
    fd, _ = richardson_fd(lambda th: gpu_loss(opt, th), theta)
    rel = report("gpu-fd adaptive", g, fd)
    floor = 1e-11 / (2 * fd_steps(theta) * np.abs(fd))
    assert np.all(rel < np.maximum(1e-6, floor)), (rel, floor)


def test_adjoint_checkpoint_interval_and_diffuse_walls():
    # the gradient does not depend on where the checkpoints fall; diffuse
    # walls (random re-emission, replayed bit-identically) are handled
    opt = workload(n=96, diffuse=0.01)
    _, g16 = adjoint_gradient(opt, chunk_records=16)
    _, g256 = adjoint_gradient(opt, chunk_records=256)
    assert np.allclose(g16, g256, rtol=1e-9, atol=0.0), (g16, g256)


def test_adjoint_window_gives_the_same_gradient():
    # replaying one checkpoint chunk per adjoint launch, three (the last
    # window partial), or all of them walks the same records backwards in
    # the same order: the same gradient bit for bit
    opt = workload(n=96, diffuse=0.01)
    _, g1 = adjoint_gradient(opt, chunk_records=16, adjoint_window=1)
    _, g3 = adjoint_gradient(opt, chunk_records=16, adjoint_window=3)
    _, g0 = adjoint_gradient(opt, chunk_records=16)
    assert g1.tobytes() == g3.tobytes() == g0.tobytes(), (g1, g3, g0)


def test_torch_autograd_binding():
    from spintrack import autograd as sa
    import torch

    opt = workload(n=64)
    theta0 = adjoint.params_of(opt)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    means = sa.ensemble_spins(theta, opt, chunk_records=16)
    times = st.output_times(opt)
    loss = sa.phase_loss(means, times, opt.t0)
    loss.backward()
    L, g = adjoint_gradient(opt)
    assert abs(loss.item() - L) <= 1e-12 * max(1.0, abs(L))
    assert np.allclose(theta.grad.numpy(), g, rtol=1e-10, atol=0.0)
