"""Writes golden/c0_oracle.npz: the FP64 oracle (oracle/model.py) on the full C0 workload (tests/c0.py).

  fwd_*   the forward projection A x_true (180 x 103 x 900): seeded entry sample, norms, per-angle sums
  adj_*   the adjoint A^T b of the Poisson counts (2000 x 99 x 99): the same figures (per-orientation sums)
  spdhg_f the SPDHG objective after 5, 10, 15 and 20 epochs (10 subsets, seed 0, x0 = 0, 200 subset steps)
  x_*     the SPDHG iterate after the 200 steps: the same figures
  rho, T and sigma step-size norms

The counts are not stored: the tests draw them again from the oracle's float64 forward (tests/c0.counts).
Takes about 10 minutes on 8 cores: python tests/golden/make_c0_fixture.py
"""
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import c0  # noqa: E402


def main():
    from oracle.model import OracleSPDHG
    t0 = time.time()
    cfg, U, sigma, xt, p, _ = c0.problem()
    orc = c0.oracle_operator(cfg, U, sigma)
    print(f"oracle operator {time.time() - t0:.0f} s", flush=True)
    ax = orc.direct(xt)
    b = c0.counts(ax, p)
    adj = orc.adjoint(b)
    print(f"forward / adjoint {time.time() - t0:.0f} s", flush=True)
    sol = OracleSPDHG(orc, subsets=p["subsets"], seed=0)
    x = sol.run(np.zeros(orc.coeff_shape), b, c0.EPOCHS, diagnostics_interval=c0.DIAG)
    print(f"SPDHG {c0.EPOCHS} epochs {time.time() - t0:.0f} s", flush=True)
    out = {"b_sum": float(np.asarray(b, np.float64).sum()), "b_l2": float(np.linalg.norm(b)),
           "spdhg_epochs": np.array([s["epoch"] for s in sol.iter_stats if s["f"] is not None]),
           "spdhg_f": np.array([s["f"] for s in sol.iter_stats if s["f"] is not None]),
           "rho": sol.rho_used, "T_l2": np.linalg.norm(sol.T), "sig_l2": np.linalg.norm(sol.sig)}
    for name, a, seed in (("fwd", ax, 1), ("adj", adj, 2), ("x", x, 3)):
        for k, v in c0.summary(a, seed).items():
            if not (name == "fwd" and k == "sum0"):         # (103 x 900 sums: left out for size)
                out[f"{name}_{k}"] = v
    np.savez_compressed(c0.FIXTURE, **out)
    print(f"wrote {c0.FIXTURE} ({c0.FIXTURE.stat().st_size} bytes): objectives {out['spdhg_f']}", flush=True)


if __name__ == "__main__":
    main()
