"""Timesteps driven by a bct.dat inflow: the resistance cylinder stepped on
the GPU with its inlet velocities read from a bct.dat file (bct.read_bct,
bct.BCt, BCint at each step's time) against the same cylinder with the
same pulsatile profile given as a function of time.  The file samples the
profile at the step times, so BCint returns the samples (the interpolation
weight is zero up to the rounding of the times): BC within 1e-12 relative,
the states after 3 steps within 1e-10 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bct_inflow_steps_equal_functional_inflow(tmp_path):
    import torch
    import crimsonflowsolver_amd as cfa
    from crimsonflowsolver_amd import bct
    from crimsonflowsolver_amd import mesh as M
    from crimsonflowsolver_amd.timestep import ResistanceCylinder, ns_solver
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    m = M.cylinder_mesh(5, 5, 8)
    p = cfa.form_params()
    wave = lambda t: 10.0 + 4.0 * np.sin(2.0 * np.pi * t / (40 * p.Delt))
    times = np.arange(41) * p.Delt
    inl = m.inlet & ~m.wall
    rr = m.x[:, 0] ** 2 + m.x[:, 1] ** 2
    prof = 2.0 * np.clip(1.0 - rr / m.radius ** 2, 0.0, None)
    path = tmp_path / "bct.dat"
    with open(path, "w") as f:
        nodes = np.nonzero(inl)[0]
        f.write(f"{nodes.size} {times.size}\n")
        for n in nodes:
            f.write(f"{m.x[n, 0]:.12f} {m.x[n, 1]:.12f} {m.x[n, 2]:.12f} {times.size}\n")
            for t in times:
                f.write(f"0.0 0.0 {wave(t) * prof[n]:.17e} {t:.17e}\n")
    ibc, _ = M.dirichlet_codes(m)
    b = bct.BCt(bct.read_bct(path), m.x, ibc)
    assert b.nodes.size == np.count_nonzero(inl)
    mk = lambda **kw: ResistanceCylinder(m, params=p, ls=ns_solver(relTol=1e-10, maxItr=10), **kw)
    cf, cb = mk(inflow=wave), mk(bct=b)
    for t in times[1:6]:
        want, got = cf.bc(t), cb.bc(t)
        assert torch.allclose(got, want, rtol=0, atol=1e-12 * float(want.abs().max())), t
    y, ac = cf.state()
    yf, af, yb, ab = y, ac, y.clone(), ac.clone()
    for n in range(3):
        yf, af, _ = cf.step(yf, af, (n + 1) * p.Delt, iters=3)
        yb, ab, _ = cb.step(yb, ab, (n + 1) * p.Delt, iters=3)
    err = float((yf - yb).abs().max() / yf.abs().max())
    print("state after 3 steps: relative difference", err, flush=True)
    assert err <= 1e-10
