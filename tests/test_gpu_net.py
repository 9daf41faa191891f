"""The learned preconditioner's network on the GPU (BF16 MFMA kernels)
against the CPU restatement oracle/pynet.py.

Two references: the oracle in the GPU's number model (rnd=bf16: BF16
activations, activation gradients and inner-layer weights, FP64 sums) and
the exact FP64 network.  With ReLU the GPU differs from the number model
only by FP32 summation order (and the odd BF16 value that rounds the other
way because of it): 2e-3 relative L2 for the output and 1e-2 per layer for
the weight gradients.  Without activations (linear, tied) every value of
both signs is re-rounded at each layer, so an FP32-ulp difference before a
rounding spreads: there the GPU must stay inside the number model's own
envelope, no farther from exact FP64 than 1.5x the model is, and a sparse
input (one impulse, few roundings near ties) must match the model to 1e-3.
The weight gradients of a random network under a random output gradient
cancel heavily, so the BF16 model itself is 3e-2..2e-1 away from FP64
there.  Scaling the input by 2 is exact in BF16 and FP32, so the
homogeneity check is bitwise."""
import numpy as np
import pytest

from hydro_amd.hydro import Net
from oracle import pynet

pytestmark = pytest.mark.gpu

NX, NY = 48, 32


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def weights():
    rng = np.random.default_rng(5)
    w = pynet.init_weights(seed=4)
    w[-9 * 16:] = rng.normal(0.0, np.sqrt(2.0 / (9 * 16)), 9 * 16)
    return w.astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def net(gpu, weights):
    n = Net(NX, NY)
    assert n.num_weights == pynet.num_weights()
    n.set_weights(weights)
    yield n
    n.close()


def test_weights_round_trip(net, weights):
    np.testing.assert_array_equal(net.get_weights(), weights.astype(np.float32))


MODES = {"relu": {}, "linear": {"linear": True}, "tied": {"tied": True}}


@pytest.fixture(scope="module")
def nets(gpu, weights):
    out = {}
    for mode, kw in MODES.items():
        n = Net(NX, NY)
        n.configure(**kw)
        n.set_weights(weights)
        out[mode] = n
    yield out
    for n in out.values():
        n.close()


@pytest.mark.parametrize("mode", list(MODES))
def test_forward_matches_oracle(nets, weights, mode):
    rng = np.random.default_rng(6)
    x = rng.normal(size=(NY, NX))
    print(f"case: net forward {NX}x{NY} {mode}", flush=True)
    y = nets[mode].forward(x)
    model = pynet.forward(weights, x, rnd=pynet.bf16, **MODES[mode])
    exact = pynet.forward(weights, x, **MODES[mode])
    e_model, e_exact = rel(y, model), rel(y, exact)
    envelope = rel(model, exact)
    print(f"forward rel L2 vs bf16 model {e_model:.3e}, vs fp64 {e_exact:.3e} "
          f"(model vs fp64 {envelope:.3e})", flush=True)
    assert e_exact < 2e-2
    if mode == "relu":
        assert e_model < 2e-3
    else:
        assert e_model < 1e-2 and e_exact < 1.5 * envelope
    imp = np.zeros((NY, NX))
    imp[NY // 3, NX // 2] = 1.0
    e_imp = rel(nets[mode].forward(imp), pynet.forward(weights, imp, rnd=pynet.bf16,
                                                       **MODES[mode]))
    print(f"impulse response rel L2 vs bf16 model {e_imp:.3e}", flush=True)
    assert e_imp < 1e-3


def test_forward_homogeneous_bitwise(net):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(NY, NX)).astype(np.float32)
    y1 = net.forward(x)
    y2 = net.forward(2.0 * x)
    np.testing.assert_array_equal(y2, 2.0 * y1)
    assert not np.any(net.forward(np.zeros_like(x)))


@pytest.mark.parametrize("mode", list(MODES))
def test_gradient_matches_oracle(nets, weights, mode):
    rng = np.random.default_rng(8)
    x = rng.normal(size=(NY, NX))
    gy = rng.normal(size=(NY, NX))
    print(f"case: net gradient {NX}x{NY} {mode}", flush=True)
    gw = nets[mode].gradient(x, gy).astype(np.float64)
    model, _ = pynet.backward(weights, x, gy, rnd=pynet.bf16, **MODES[mode])
    exact, _ = pynet.backward(weights, x, gy, **MODES[mode])
    off = 0
    for name, ci, co, _ in pynet.layer_shapes():
        k = 9 * ci * co
        sl = slice(off, off + k)
        if mode == "tied" and name[0] == "D":
            assert not np.any(gw[sl]), name  # tied decoder slots carry no gradient
            off += k
            continue
        e_model, e_exact = rel(gw[sl], model[sl]), rel(gw[sl], exact[sl])
        envelope = rel(model[sl], exact[sl])
        print(f"  {name} {ci}->{co}: rel L2 vs bf16 model {e_model:.3e}, vs fp64 {e_exact:.3e} "
              f"(model vs fp64 {envelope:.3e})", flush=True)
        assert e_exact < 0.3, name
        if mode == "relu":
            assert e_model < 1e-2, name
        else:
            assert e_exact < 1.5 * envelope + 1e-3, name
        off += k


def test_tied_network_symmetric(nets):
    """(u, N v) = (N u, v) for the tied network, to BF16 rounding."""
    rng = np.random.default_rng(12)
    u, v = rng.normal(size=(NY, NX)), rng.normal(size=(NY, NX))
    n = nets["tied"]
    a = float(np.sum(u * n.forward(v)))
    b = float(np.sum(n.forward(u) * v))
    assert abs(a - b) < 1e-2 * (abs(a) + abs(b))
    assert float(np.sum(u * n.forward(u))) > 0


def test_first_adam_step(gpu, weights):
    """Loss before the step matches the oracle; the first Adam step moves
    every weight with a clear gradient by -lr*sign(g)."""
    rng = np.random.default_rng(9)
    x = rng.normal(size=(NY, NX))
    t = rng.normal(size=(NY, NX))
    n = Net(NX, NY)
    n.set_weights(weights)
    lr = 1e-3
    loss = n.train_step(x, t, lr)
    y = pynet.forward(weights, x, rnd=pynet.bf16)
    ref_loss = float(np.mean((y - t) ** 2))
    assert abs(loss - ref_loss) < 1e-3 * ref_loss
    g, _ = pynet.backward(weights, x, 2.0 * (y - t) / x.size, rnd=pynet.bf16)
    w1 = n.get_weights().astype(np.float64)
    n.close()
    clear = np.abs(g) > 0.05 * np.abs(g).max()
    step = w1[clear] - weights[clear]
    np.testing.assert_allclose(step, -lr * np.sign(g[clear]), rtol=2e-3, atol=1e-7)


def test_training_reduces_loss(gpu):
    """A realisable target (another network's output) is fitted by Adam."""
    rng = np.random.default_rng(10)
    w_teacher = pynet.init_weights(seed=11)
    w_teacher[-9 * 16:] = rng.normal(0.0, 0.1, 9 * 16)
    xs = [rng.normal(size=(NY, NX)) for _ in range(4)]
    ts = [pynet.forward(w_teacher, x) for x in xs]
    n = Net(NX, NY)
    n.set_weights(pynet.init_weights(seed=12))
    losses = [n.train_step(xs[k % 4], ts[k % 4], 2e-3) for k in range(120)]
    n.close()
    first, last = np.mean(losses[:8]), np.mean(losses[-8:])
    print(f"training loss {first:.4e} -> {last:.4e}", flush=True)
    assert last < 0.5 * first
