"""SPDHGPoisson on the GPU vs the FP64 oracle restatement of spdhg_poisson.py (same subset order from the same
seed): iterate, objective, dual state; warm start; the same result for a NumPy and a torch x0.
Tolerance: FP32 iterates within 1e-4 (relative L2) of the FP64 iterates after 8 epochs, objectives within 1e-5."""
import numpy as np
import pytest

from oracle.model import OracleSPDHG
from pairs import build, cfg, rel_l2

pytestmark = pytest.mark.gpu


def _problem(K=12, seed=7, gain=None):
    op, orc, _ = build(cfg(nx=16, ny=16, my=17, n_omega=24, angle_range=(0, 180), cor=0.0), K=K, sigma_deg=4.0,
                       pf_mode="sparse")
    rng = np.random.default_rng(seed)
    xt = (rng.random(op.coeff_shape) * (rng.random(op.coeff_shape) < 0.2) * op.support_mask()[None]
          ).astype(np.float32)
    ax = orc.direct(xt)
    b = rng.poisson(ax / ax.mean() * 50).astype(np.float32)
    return op, orc, b


@pytest.mark.parametrize("subsets,weights", [(4, None), (3, "eta")])
def test_spdhg_matches_oracle(subsets, weights):
    from diffractom_amd import SPDHGPoisson
    print(f"case subsets={subsets} weights={weights}", flush=True)
    op, orc, b = _problem()
    w = None
    if weights == "eta":
        w = np.ones((op.N_eta, op.N_peaks), np.float32)
        w[5:9] = 0
    sp = SPDHGPoisson(op, subsets=subsets, seed=1)
    x = sp.run(np.zeros(op.coeff_shape, np.float32), b, 8, weights=w, diagnostics_interval=2)
    ref = OracleSPDHG(orc, subsets=subsets, seed=1)
    x_ref = ref.run(np.zeros(op.coeff_shape), b, 8, weights=w, diagnostics_interval=2)
    assert abs(sp.rho_used / ref.rho_used - 1) < 1e-6
    assert rel_l2(x, x_ref) < 1e-4
    f = [s["f"] for s in sp.iter_stats if s["f"] is not None]
    f_ref = [s["f"] for s in ref.iter_stats if s["f"] is not None]
    np.testing.assert_allclose(f, f_ref, rtol=1e-5)
    assert rel_l2(sp.state["y"], ref.state["y"]) < 1e-4
    assert rel_l2(sp.state["z"], ref.state["z"]) < 1e-4


def test_spdhg_resume_and_torch_x0():
    """tests/test_subsets.py:148-171: two runs continued from solver.state equal one run; a torch x0 gives the
    NumPy x0's iterates."""
    import torch
    from diffractom_amd import SPDHGPoisson
    op, _, b = _problem()
    one = SPDHGPoisson(op, subsets=3, seed=1)
    x_one = one.run(np.zeros(op.coeff_shape, np.float32), b, 6, diagnostics_interval=2)
    xt = torch.zeros(op.coeff_shape, dtype=torch.float32, device="cuda")
    tor = SPDHGPoisson(op, subsets=3, seed=1)
    out = tor.run(xt, b, 6, diagnostics_interval=2)
    assert out is xt
    assert np.array_equal(xt.cpu().numpy(), x_one)
    two = SPDHGPoisson(op, subsets=3, seed=1)
    x = two.run(np.zeros(op.coeff_shape, np.float32), b, 4, diagnostics_interval=4)
    x = two.run(x, b, 2, diagnostics_interval=2, state=two.state)
    assert two.state["epoch"] == 6
    np.testing.assert_allclose(x, x_one, rtol=1e-5, atol=1e-6 * np.abs(x_one).max())
    st = {"y": two.state["y"], "epoch": 6}
    three = SPDHGPoisson(op, subsets=3, seed=2)
    three.run(x.copy(), b, 1, state=st)
    assert three.state["epoch"] == 7
