"""Products with the coefficients in host memory (HostStreamer): forward and adjoint of host arrays equal the
operator's device products (direct_cl / adjoint_cl: the same kernels in the same order) bit for bit, with several
orientation batches (two staging slots reused), repeated and interleaved calls, for the FP32 CSR and the BF16
block storage."""
import numpy as np
import pytest

from pairs import build, cfg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pf_mode,precision", [("sparse", "fp32"), ("sparse", "bf16_residual")])
def test_host_streamed_products_equal_device_products(pf_mode, precision):
    import torch
    from diffractom_amd.streaming import HostStreamer
    print(f"case host streaming {pf_mode} {precision}", flush=True)
    c = cfg(nx=20, ny=18, my=23, n_omega=24, angle_range=(0, 180), cor=0.3, n_eta=24)
    op, _, _ = build(c, K=300, pf_mode=pf_mode, pf_precision=precision, max_gb=2e-4, oracle=False)
    assert len(op.batches) >= 3
    rng = np.random.default_rng(3)
    hs = HostStreamer(op)
    for rep in range(2):
        x = (rng.random(op.coeff_shape) * op.support_mask()[None]).astype(np.float32)
        y = rng.random(op.data_shape).astype(np.float32)
        d_ref = torch.empty(op.data_shape, device="cuda")
        op.direct_cl(torch.from_numpy(x).cuda(), d_ref)
        a_ref = torch.empty(op.coeff_shape, device="cuda")
        op.adjoint_cl(torch.from_numpy(y).cuda(), a_ref)
        d = torch.empty(op.data_shape, device="cuda")
        hs.forward(x, d)
        a = np.empty(op.coeff_shape, np.float32)
        hs.adjoint(torch.from_numpy(y).cuda(), a)
        torch.cuda.synchronize()
        assert torch.equal(d, d_ref), rep
        assert np.array_equal(a, a_ref.cpu().numpy()), rep
    with pytest.raises(ValueError):
        hs.forward(np.zeros((3, 3), np.float32), d)
    hs.release()
