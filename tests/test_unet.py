"""The learned warm start's U-Net: the FP64 oracle pinned by finite
differences and invariants (CPU), and the GPU network (BF16 on the matrix
cores, FP32 accumulation) against the oracle (gpu)."""
import ctypes

import numpy as np
import pytest

from oracle import unet_ref


def small_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in unet_ref.LAYERS:
        out.append(rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci)) * scale)
    return unet_ref.join(out, rng.standard_normal(unet_ref.HEAD) * 0.25)


def problem(n, seed):
    rng = np.random.default_rng(seed)
    field = rng.standard_normal(n ** 3)
    target = rng.standard_normal(n ** 3) * 0.1
    return field, target


def test_reference_gradient_matches_finite_differences():
    n = 8
    params = small_params(1)
    field, target = problem(n, 2)
    x0 = unet_ref.make_input(field, n)
    loss, grad = unet_ref.loss_and_grad(params, x0, target)
    rng = np.random.default_rng(3)
    # a few parameters of every layer and of the head
    offsets = np.cumsum([0] + [co * 27 * ci for ci, co in unet_ref.LAYERS])
    picks = []
    for l in range(len(unet_ref.LAYERS)):
        picks += list(rng.integers(offsets[l], offsets[l + 1], 3))
    picks += list(offsets[-1] + rng.integers(0, unet_ref.HEAD, 3))
    h = 1e-6
    checked = 0
    for i in picks:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        num = (unet_ref.loss_and_grad(p1, x0, target)[0] - unet_ref.loss_and_grad(p2, x0, target)[0]) / (2 * h)
        if abs(num) < 1e-12 and abs(grad[i]) < 1e-12:
            continue  # a dead ReLU path: both zero
        assert grad[i] == pytest.approx(num, rel=1e-4, abs=1e-9), (i, grad[i], num)
        checked += 1
    assert checked >= 20


def test_reference_is_positively_homogeneous():
    n = 8
    params = small_params(4)
    field, _ = problem(n, 5)
    x0 = unet_ref.make_input(field, n)
    out1, _ = unet_ref.forward(params, x0)
    x2 = x0.copy()
    x2[..., 0] *= 3.0
    x2[..., 1] *= 3.0
    out2, _ = unet_ref.forward(params, x2)
    assert np.allclose(out2, 3.0 * out1, rtol=1e-12, atol=1e-12)


def test_reference_adam_first_step_is_a_signed_step():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    q, m, v = unet_ref.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # t = 1: mhat = g, vhat = g^2, so the step is lr * g / (|g| + eps)
    assert np.allclose(q, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


# ------------------------------------------------------------------- GPU


def gpu_lib():
    from vapor_amd.abi import load_library
    return load_library()


class GpuUnet:
    def __init__(self, seed=0):
        self.lib = gpu_lib()
        h = ctypes.c_void_p()
        assert self.lib.vgx_unet_create(seed, ctypes.byref(h)) == 0, self.lib.vgx_last_error()
        self.h = h
        self.P = self.lib.vgx_unet_param_count()

    def close(self):
        self.lib.vgx_unet_destroy(self.h)

    def _f(self, a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self):
        out = np.empty(self.P, np.float32)
        assert self.lib.vgx_unet_get_params(self.h, self._f(out)) == 0
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert self.lib.vgx_unet_set_params(self.h, self._f(p)) == 0

    def grads(self):
        out = np.empty(self.P, np.float32)
        assert self.lib.vgx_unet_get_grads(self.h, self._f(out)) == 0
        return out

    def set_grads(self, g):
        g = np.ascontiguousarray(g, np.float32)
        assert self.lib.vgx_unet_set_grads(self.h, self._f(g)) == 0

    def predict(self, n, field, in_scale=1.0, out_scale=1.0):
        f = np.ascontiguousarray(field, np.float32)
        out = np.empty(n ** 3, np.float32)
        assert self.lib.vgx_unet_predict(self.h, n, self._f(f), in_scale, out_scale, self._f(out)) == 0, \
            self.lib.vgx_last_error()
        return out

    def loss_grad(self, n, field, target, accumulate=False):
        f = np.ascontiguousarray(field, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = ctypes.c_double()
        assert self.lib.vgx_unet_loss_grad(self.h, n, self._f(f), 1.0, self._f(t), int(accumulate),
                                           ctypes.byref(loss)) == 0, self.lib.vgx_last_error()
        return loss.value

    def adam(self, lr, grad_scale=1.0):
        assert self.lib.vgx_unet_adam_step(self.h, lr, grad_scale) == 0


def bf16_round(a):
    """Round to BF16 (nearest even), as the GPU stores weights and activations."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def rel_l2(got, want):
    return float(np.linalg.norm(np.asarray(got, np.float64) - want) / max(np.linalg.norm(want), 1e-30))


@pytest.mark.gpu
def test_gpu_unet_forward_matches_reference():
    n = 16
    net = GpuUnet(7)
    params = net.params()
    field, _ = problem(n, 8)
    got = net.predict(n, field, 1.0, 1.0)
    # the reference with the GPU's BF16 weights (the head stays FP32)
    p = params.astype(np.float64).copy()
    p[:-unet_ref.HEAD] = bf16_round(params[:-unet_ref.HEAD])
    want, _ = unet_ref.forward(p, unet_ref.make_input(bf16_round(field), n))
    err = rel_l2(got, want.reshape(-1))
    print("forward rel L2", err)
    assert err < 3e-2
    # the output scale is applied exactly (power-of-two scales round exactly)
    got2 = net.predict(n, field, 1.0, 0.5)
    assert np.array_equal(got2, np.float32(0.5) * got)
    net.close()


@pytest.mark.gpu
def test_gpu_unet_gradients_match_reference():
    n = 16
    net = GpuUnet(9)
    params = net.params()
    field, target = problem(n, 10)
    loss = net.loss_grad(n, field, target)
    got = net.grads()
    p = params.astype(np.float64).copy()
    p[:-unet_ref.HEAD] = bf16_round(params[:-unet_ref.HEAD])
    want_loss, want = unet_ref.loss_and_grad(p, unet_ref.make_input(bf16_round(field), n), target)
    assert loss == pytest.approx(want_loss, rel=3e-2)
    offsets = np.cumsum([0] + [co * 27 * ci for ci, co in unet_ref.LAYERS] + [unet_ref.HEAD])
    for l in range(len(offsets) - 1):
        g, w = got[offsets[l]:offsets[l + 1]], want[offsets[l]:offsets[l + 1]]
        cos = float(np.dot(g, w) / max(np.linalg.norm(g) * np.linalg.norm(w), 1e-30))
        print("layer", l, "rel L2", rel_l2(g, w), "cos", cos)
        assert rel_l2(g, w) < 0.1 and cos > 0.99, l
    # accumulate adds a second identical gradient
    net.loss_grad(n, field, target, accumulate=True)
    assert np.allclose(net.grads(), 2 * got, rtol=1e-5, atol=1e-12)
    net.close()


@pytest.mark.gpu
def test_gpu_adam_matches_reference():
    net = GpuUnet(11)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(12)
    m, v = np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 4):
        g = rng.standard_normal(p.size) * 1e-2
        net.set_grads(g)
        net.adam(1e-3, 0.5)
        p, m, v = unet_ref.adam(p, 0.5 * g, m, v, t, lr=1e-3)
    got = net.params()
    assert np.allclose(got, p, rtol=1e-5, atol=1e-7)
    net.close()


@pytest.mark.gpu
def test_gpu_training_reduces_the_loss():
    n = 16
    net = GpuUnet(13)
    field, _ = problem(n, 14)
    # a smooth target: the field blurred (what an inverse Laplacian looks like)
    f = field.reshape(n, n, n)
    target = f.copy()
    for _ in range(4):
        target = (np.roll(target, 1, 0) + np.roll(target, -1, 0) + np.roll(target, 1, 1) + np.roll(target, -1, 1) +
                  np.roll(target, 1, 2) + np.roll(target, -1, 2) + 2 * target) / 8
    target = target.reshape(-1)
    first = net.loss_grad(n, field, target)
    net.adam(3e-3)
    for _ in range(40):
        net.loss_grad(n, field, target)
        net.adam(3e-3)
    last = net.loss_grad(n, field, target)
    print("loss", first, "->", last)
    assert last < 0.5 * first
    net.close()


@pytest.mark.gpu
def test_learned_warm_start_records_trains_and_converges():
    """The solver records its own (residual, correction) pairs, the network
    trains on them, and with the learned start every projection still
    reaches the tolerance; the fields stay those of the plain warm start to
    the precision the two converged solves allow."""
    from vapor_amd import mac3d
    n = 32

    def params():
        p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
        p.update(emitter_radius=0.12, sor_omega=1.8, solve_mode=1, tolerance=1e-5, max_sweeps=4000,
                 check_interval=16, warm_start=1)
        return p

    data = mac3d.Dataset(n, 64)
    rec = mac3d.GpuMacSimulation3D(n, 1.0, params())
    rec.record(data)
    for _ in range(6):
        rec.advance(1 / 60)
    rec.record(None)
    assert len(data) == 12  # two warm projections per reflection step
    net = mac3d.UNet(3)
    first = net.train(data, 1, 4, 1e-3, seed=1)
    last = net.train(data, 4, 4, 1e-3, seed=2)
    print("training loss", first, "->", last)
    assert np.isfinite(last) and last < first

    plain = mac3d.GpuMacSimulation3D(n, 1.0, params())
    learned = mac3d.GpuMacSimulation3D(n, 1.0, params())
    learned.setUnet(net)
    decisions = []
    for _ in range(4):
        plain.advance(1 / 60)
        learned.advance(1 / 60)
        for k in (1, 0):
            r = learned.lastProjection(k)
            assert r.converged and r.final_residual_l2 <= 1e-5 * 1.01
            decisions.append(r.learned_start)
    print("learned start decisions", decisions)
    assert all(d in (1, 2) for d in decisions)
    for got, want in zip(learned.fields(), plain.fields()):
        scale = max(1.0, float(np.max(np.abs(want))))
        assert float(np.max(np.abs(got.astype(np.float64) - want))) <= 1e-3 * scale
    learned.setUnet(None)
    learned.advance(1 / 60)
    assert learned.lastProjection().learned_start == 0
