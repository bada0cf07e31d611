"""The two launch shapes of the scalar LDS-resident Jacobi (MSQ_JACOBI_SCALAR)
for padded orders 32..80:
"wide" (one thread per off-diagonal block, used while every workgroup of the
launch is resident) and "narrow" (several blocks per thread, used above
256 CUs x 20 waves of workgroups).  A batch above each order's threshold
runs narrow, the same matrix alone runs wide: with fixed sweeps every
rotation and update is the same arithmetic on the same values, so the
results agree bit for bit."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, jacobi_method, say
from msq import capi
from msq import generators as G

pytestmark = pytest.mark.gpu

# (order, matrices): above 2560 / 1024 / 640 / 393 matrices for padded
# orders 32 / 48 / 64 / 80 on a 256-CU device.
CASES = [(30, 2600), (45, 1100), (60, 700), (75, 400)]


def run(runtime, mats, opt):
    orders, flat = capi.pack_batch(mats)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    R = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    with jacobi_method(runtime, capi.JACOBI_SCALAR):
        runtime.jacobi_batched(orders, dev(flat), V, D, R, opt)
    return capi.unpack_batch(orders, host(V)), host(D), voffs


@pytest.mark.parametrize("order,count", CASES)
def test_narrow_batch_equals_wide_solo(runtime, order, count):
    rng = np.random.default_rng(order)
    base = [G.single_matrix(14, order, 10.0 ** -rng.uniform(1, 4), index=i)[0]
            for i in range(16)]
    mats = [base[i % len(base)] for i in range(count)]
    opt = capi.jacobi_options(fixed_sweeps=4)
    say(f"narrow shape: {count} matrices of order {order}")
    Vb, db, voffs = run(runtime, mats, opt)
    for k in range(len(base)):
        Va, da, _ = run(runtime, [base[k]], opt)
        for i in (k, k + len(base) * (count // len(base) - 1)):
            assert np.array_equal(Va[0], Vb[i]), (order, k, i)
            assert np.array_equal(da[:order], db[voffs[i]:voffs[i] + order]), (order, k, i)
