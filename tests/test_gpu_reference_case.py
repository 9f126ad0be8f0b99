"""The reference's basic case end to end on the GPU (phasta_case.PhastaCase:
the formation with the deformable wall, the boundary loop, ElmpvsQ with the
RCR outlet, memLS's NS solve, the generalized-alpha updates, all through the
C ABI) against the reference's own output of the same run
(tests/golden/basic_outputs.npz; the CPU replay is test_reference_case.py):
the RCR outlet's flow after step 5 within the reference's own test
tolerance (testMain.cxx:19, 1e-7), every step's outlet flow and RCR pressure
within 1e-9 relative, and the state after 5 steps (restart.5.1) within
1e-9 (velocity, displacement), 1e-10 (pressure), 1e-8 (acceleration).
Then, from the reference's restart.5.1 and its PHistRCR row 5, 5 more steps
against its restart test (testMain.cxx:25-42: PHistRCR row 10 =
11625.41088 within 1e-4).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_basic_case_five_steps_against_the_reference():
    import crimsonflowsolver_amd as cfa
    from crimsonflowsolver_amd.phasta_case import PhastaCase, wall_from_solver_inp
    if cfa.device_count() < 1:
        pytest.fail("no GPU visible: the HIP path must run here")
    z = np.load(os.path.join(HERE, "golden", "basic_case.npz"))
    b = np.load(os.path.join(HERE, "golden", "basic_bct.npz"))
    o = np.load(os.path.join(HERE, "golden", "basic_outputs.npz"))
    recs = [(b["x"][k], np.concatenate([b["u"][k], b["t"][:, None]], axis=1)) for k in range(b["x"].shape[0])]
    rcr = {3: (o["rcr"][0], o["rcr"][1], o["rcr"][2], o["pdist"])}
    c = PhastaCase(z["x"], z["ien"], z["ienb"], z["iBCB"], z["BCB"], z["iBC"], z["y"], z["ac"], z["u"], z["uref"],
                   z["nodetag"], z["ndsurf"], recs, rcr, rho=0.00106, mu=0.004, Delt=0.001, rhoinf=0.0,
                   wall=wall_from_solver_inp(), dirichlet_srfs=(2,))
    assert c.rcr[3].P == o["presshist"][0, 1] and c.rcr[3].Q == 0.0
    for n in range(5):
        r = c.step()
        q, pr = r[3]["Q"], r[3]["P"]
        eq = abs(q - o["qhist"][n, 1]) / abs(o["qhist"][n, 1])
        ep = abs(pr - o["phist"][n, 1]) / abs(o["phist"][n, 1])
        print(f"step {n + 1}: Q {q!r} (reference {o['qhist'][n, 1]!r}, {eq:.1e}), P {pr!r} ({ep:.1e}), "
              f"iterations {[(round(a, 4), int(b_), int(c_), int(d)) for a, b_, c_, d in r['iterations']]}",
              flush=True)
        assert eq <= 1e-9 and ep <= 1e-9, (n + 1, eq, ep)
        if n == 3:  # the values testMain.cxx:102-159 hold for RCR-emulating coronary and netlist outlets
            assert abs(r[3]["p_mean"] - 10645.5858581080) <= 1e-7 and abs(q - 478.116982120136) <= 1e-7
    assert abs(q - 714.921079082528) <= 1e-7  # testMain.cxx:19
    y, ac, u = c.state()
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    errs = {"u": rel(y[:, :3], o["y5"][:, :3]), "p": rel(y[:, 3], o["y5"][:, 3]), "ac": rel(ac[:, :3], o["ac5"][:, :3]),
            "disp": rel(u, o["u5"])}
    print("restart.5.1", errs)
    assert errs["u"] <= 1e-9 and errs["p"] <= 1e-10 and errs["ac"] <= 1e-8 and errs["disp"] <= 1e-9, errs
    # the restart test: restart.5.1, currentTimestepIndex 5, PHistRCR row 5
    c.restart(o["y5"], o["ac5"], o["u5"], 5, {3: o["phist"][4, 1]})
    ps = [c.step()[3]["P"] for _ in range(5)]
    print("PHistRCR rows 6-10:", ps)
    assert abs(ps[-1] - 0.1162541088e5) <= 10e-5  # testMain.cxx:38
