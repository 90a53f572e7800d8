"""The HIP cost-model kernels (dd_mlp_*) against the numpy reference
(oracle/mlp.py): forward pass, loss, gradient and Adam step."""
import ctypes

import numpy as np
import pytest

from dedispersion_amd import load
from dedispersion_amd._lib import check
from oracle import mlp

pytestmark = pytest.mark.gpu

SHAPE = (12, 64, 3)  # the tuner's network: 12 features, 3 x 64 ReLU units


def f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Net:
    def __init__(self, shape=SHAPE):
        self.lib = load()
        self.shape = shape
        h = ctypes.c_void_p()
        check(self.lib.dd_mlp_create(ctypes.byref(h), *shape))
        self.h = h
        n = ctypes.c_uint32()
        check(self.lib.dd_mlp_param_count(self.h, ctypes.byref(n)))
        self.count = n.value

    def set(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self.lib.dd_mlp_set_params(self.h, f32p(p)))

    def get(self):
        p = np.zeros(self.count, dtype=np.float32)
        check(self.lib.dd_mlp_get_params(self.h, f32p(p)))
        return p

    def close(self):
        self.lib.dd_mlp_destroy(self.h)


@pytest.fixture
def net(torch_cuda):
    n = Net()
    yield n
    n.close()


def random_case(rng, n_rows, count):
    p = (rng.standard_normal(count) * 0.2).astype(np.float32)
    x = rng.standard_normal((n_rows, SHAPE[0])).astype(np.float32)
    t = rng.standard_normal(n_rows).astype(np.float32)
    return p, x, t


def test_parameter_count_and_round_trip(net):
    assert net.count == mlp.param_count(*SHAPE) == 9217
    p = np.arange(net.count, dtype=np.float32) * 1e-3
    net.set(p)
    np.testing.assert_array_equal(net.get(), p)


@pytest.mark.parametrize("n_rows", [1, 37, 3000])
def test_forward_matches_oracle(torch_cuda, net, n_rows):
    torch = torch_cuda
    rng = np.random.default_rng(n_rows)
    p, x, _ = random_case(rng, n_rows, net.count)
    net.set(p)
    dx = torch.from_numpy(x).cuda()
    dy = torch.empty(n_rows, dtype=torch.float32, device="cuda")
    check(net.lib.dd_mlp_forward(net.h, dx.data_ptr(), n_rows, dy.data_ptr(), None))
    want, _ = mlp.forward(p, x, *SHAPE)
    got = dy.cpu().numpy()
    # float32 sums of <= 64 terms: relative error to the output scale
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("n_rows", [5, 256, 2880])
def test_gradient_matches_oracle(torch_cuda, net, n_rows):
    torch = torch_cuda
    rng = np.random.default_rng(100 + n_rows)
    p, x, t = random_case(rng, n_rows, net.count)
    net.set(p)
    dx, dt = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    dloss = torch.empty(1, dtype=torch.float32, device="cuda")
    dgrad = torch.empty(net.count, dtype=torch.float32, device="cuda")
    check(net.lib.dd_mlp_loss_grad(net.h, dx.data_ptr(), dt.data_ptr(), n_rows, dloss.data_ptr(), dgrad.data_ptr(),
                                   None))
    loss, grad = mlp.loss_grad(p, x, t, *SHAPE)
    assert abs(dloss.item() - loss) <= 1e-5 * max(1.0, loss)
    g = dgrad.cpu().numpy()
    # per layer, relative to the largest gradient of that layer
    for w, a, b, bo in mlp.layout(*SHAPE):
        for lo, hi in ((w, w + a * b), (bo, bo + b)):
            scale = max(np.max(np.abs(grad[lo:hi])), 1e-12)
            assert np.max(np.abs(g[lo:hi] - grad[lo:hi])) <= 2e-4 * scale, (lo, hi)


def test_adam_step_matches_oracle(torch_cuda, net):
    torch = torch_cuda
    rng = np.random.default_rng(7)
    p, _, _ = random_case(rng, 1, net.count)
    net.set(p)
    g = (rng.standard_normal(net.count) * 0.1).astype(np.float32)
    dg = torch.from_numpy(g).cuda()
    m = np.zeros(net.count)
    v = np.zeros(net.count)
    want = p.astype(np.float64)
    for step in (1, 2, 3):
        check(net.lib.dd_mlp_adam_step(net.h, dg.data_ptr(), ctypes.c_float(1e-2), None))
        want, m, v = mlp.adam_step(want, g.astype(np.float64), m, v, step, 1e-2)
    torch.cuda.synchronize()
    np.testing.assert_allclose(net.get(), want, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_function(torch_cuda, net):
    """Full-batch training on the device drives the loss of a smooth target
    (y = sin(x0) + 0.5 x1^2 - x2 x3) down by two orders of magnitude."""
    torch = torch_cuda
    rng = np.random.default_rng(11)
    x = rng.uniform(-1.5, 1.5, (2000, SHAPE[0])).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - x[:, 2] * x[:, 3]).astype(np.float32)
    check(net.lib.dd_mlp_init(net.h, 5))
    dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    losses = np.zeros(1500, dtype=np.float32)
    check(net.lib.dd_mlp_train(net.h, dx.data_ptr(), dy.data_ptr(), len(y), len(losses), ctypes.c_float(3e-3),
                               f32p(losses), None))
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0] / 100, (losses[0], losses[-1])
