"""The FP64 oracle reproduces the objective the reference printed for its example (examples/
example_diffractom_pipeline.ipynb cells 11-13: tailored grid, 543 orientations, sigma 0.5 deg, SPDHG with 8 subsets,
seed 0, eta bins 40-49 and 130-139 masked): f = 1.14836889e6 after 10 epochs. The reference ran float32 on a GPU;
the oracle is float64, so the bar is 2e-5. This pins the whole chain: reflections, geometry, quadrature (2 x 3 Gauss
nodes per bin), pole figure, projector, step sizes, subset order, dual prox and deviance."""
import numpy as np
import pytest

from notebook import load, tailored_grid
from oracle.model import OracleOperator, OracleSPDHG


def test_notebook_objective_after_10_epochs():
    p = load()
    op = OracleOperator(p["cfg"], p["lattice"], "cubic", p["hkl"], p["wavelength"], tailored_grid(),
                        np.deg2rad(0.5))
    assert op.rule == (2, 1, 3, 1) and op.K == 543
    sp = OracleSPDHG(op, subsets=8, seed=0)
    sp.run(np.zeros(op.coeff_shape), p["b"], 10, weights=p["weights"], diagnostics_interval=10)
    f = sp.iter_stats[-1]["f"]
    golden = p["golden"]["tailored"]["f"]["10"]
    assert abs(f / golden - 1) < 2e-5, (f, golden)
