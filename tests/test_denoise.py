"""Dose denoiser (include/dxmc_denoise.h).

CPU: the C-ABI's argument checks, the feature maps, and the PyTorch
restatement of the network's backward pass against autograd (which pins the
reference the GPU tests use).  GPU: the matrix-core register layout, each
convolution kernel (forward, input gradient, weight gradient) against
torch.nn.functional.conv3d and torch.nn.grad on BF16-rounded inputs, the
whole network's prediction, loss and every parameter gradient against the
restatement, and that training lowers the loss.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from opendxmc_amd import _native as N
from opendxmc_amd import denoise as dn

import denoise_reference as ref


def _bf16(a: np.ndarray) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _probe(mode, x, w, extra, n, dims):
    dims_a = np.array(dims, dtype=np.uint64)
    out = np.zeros(ref.LAYER_P if mode == 2 else x.size, dtype=np.float32)
    extra_p = _pf(extra) if extra is not None else None
    rc = N.lib().dxg_conv3d_probe(mode, _pf(x), _pf(w), extra_p, n, dims_a.ctypes.data_as(C.POINTER(N.u64)), _pf(out))
    N.check(rc)
    return out


# ------------------------------------------------------------------ CPU

def test_parameter_count():
    # 18 layers x (32 x 27 x 32 weights + 32 biases), padded channels included
    assert dn.Denoiser.num_params() == 18 * (32 * 27 * 32 + 32)
    # the unpadded network: 4->32, 16 x 32->32, 32->1
    live = (4 * 27 * 32 + 32) + 16 * (32 * 27 * 32 + 32) + (32 * 27 + 1)
    assert live == 447_233


def test_probe_rejects_bad_shapes():
    x = np.zeros(2 * 4 * 4 * 48 * 32, dtype=np.float32)
    w = np.zeros(ref.LAYER_W, dtype=np.float32)
    b = np.zeros(32, dtype=np.float32)
    with pytest.raises(N.NativeError) as e:
        _probe(0, x, w, b, 1, (4, 4, 48))
    assert e.value.code == N.DXG_ERR_ARGUMENT and "multiple of 64" in str(e.value)
    x = np.zeros(1 * 2 * 6 * 64 * 32, dtype=np.float32)
    with pytest.raises(N.NativeError) as e:
        _probe(2, x, w, x, 1, (2, 6, 64))
    assert e.value.code == N.DXG_ERR_ARGUMENT and "multiple of 4" in str(e.value)
    with pytest.raises(N.NativeError) as e:
        N.check(N.lib().dxg_conv3d_probe(4, _pf(x), _pf(w), _pf(b), 1,
                                         np.array([2, 6, 64], dtype=np.uint64).ctypes.data_as(C.POINTER(N.u64)),
                                         _pf(x)))
    assert e.value.code == N.DXG_ERR_ARGUMENT


def test_create_without_device_fails_loudly():
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(N.NativeError) as e:
        dn.Denoiser(seed=1)
    assert e.value.code == N.DXG_ERR_DEVICE


def test_feature_maps():
    shape = (2, 3, 4)
    dose = np.full(shape, 2.0)
    dose[0, 0, 0] = 0.0
    var = (0.1 * dose) ** 2
    material = np.ones(shape, dtype=np.uint8)
    material[1, 2, 3] = 0
    density = np.full(shape, 1.05)
    emb = np.array([0.9, 1.1], dtype=np.float32)
    f, dref = dn.features(dn.DoseVolume(dose, var, density, material, emb))
    body = material != 0
    assert dref == pytest.approx(dose[body].mean())
    assert f.shape == shape + (4,) and f.dtype == np.float32
    assert f[1, 1, 1, 0] == pytest.approx(2.0 / dref, rel=1e-6)
    assert f[0, 0, 0, 0] == 0.0
    assert f[1, 1, 1, 3] == pytest.approx(0.1, rel=1e-6)
    assert f[0, 0, 0, 3] == 1.0                       # no events: full uncertainty
    assert f[1, 2, 3, 2] == pytest.approx(0.9) and f[0, 1, 1, 2] == pytest.approx(1.1)
    back = dn.dose_from_relative(dn.relative_dose(dose, dref), dref)
    np.testing.assert_allclose(back, dose, rtol=1e-6, atol=1e-9)


def test_material_embedding_water_is_one():
    emb = dn.material_embedding([{1: 0.111894, 8: 0.888106}, {7: 0.8, 8: 0.2}, {20: 1.0}])
    assert emb[0] == pytest.approx(1.0, rel=1e-12)
    assert 0.8 < emb[1] < 1.0          # air: slightly lower mass attenuation than water at 60 keV
    assert emb[2] > 3.0                # calcium: photoelectric dominated


def test_pad_and_error_metric():
    a = np.ones((2, 3, 70, 4))
    p = dn.pad_x(a)
    assert p.shape == (2, 3, 128, 4) and p[:, :, 70:].sum() == 0
    r = np.array([1.0, 2.5, 10.0, 0.1])
    d = np.array([1.1, 2.5, 9.0, 5.0])
    # voxels above 20 % of max: 2.5 and 10.0 (1.0 and 0.1 are below 2.0)
    assert dn.mean_relative_error(d, r) == pytest.approx((0.0 + 0.1) / 2)


def test_reference_backward_matches_autograd():
    """The hand-written backward of the restatement (unrounded, float64) equals
    autograd of its own forward, so a wrong gate, a transposed or unflipped
    weight or a missing residual term in it would fail here."""
    rng = np.random.default_rng(5)
    params = ref.random_params(3, scale=0.7)
    f = rng.normal(0.0, 1.0, (1, 2, 4, 4, 4))
    target = rng.normal(0.0, 1.0, (1, 2, 4, 4))
    mask = (rng.random((1, 2, 4, 4)) < 0.8).astype(np.float64)
    loss, _, grads = ref.loss_and_grads(params, f, target, mask, rounding=False, dtype=torch.float64)

    layers = [(w.to(torch.float64).requires_grad_(), b.to(torch.float64).requires_grad_()) for w, b in ref.split(params)]
    x0 = ref.pad_features(f, torch.float64)
    a = torch.relu(ref.conv(x0, *layers[0]))
    for k in range(ref.BLOCKS):
        h = torch.relu(ref.conv(a, *layers[2 * k + 1]))
        a = torch.relu(ref.conv(h, *layers[2 * k + 2]) + a)
    pred = torch.from_numpy(f[..., 0]) + ref.conv(a, *layers[-1])[..., 0]
    m = torch.from_numpy(mask)
    l2 = (m * (pred - torch.from_numpy(target)) ** 2).sum() / m.sum()
    l2.backward()
    auto = np.concatenate([ref.flat_grad(w.grad, b.grad) for w, b in layers])
    assert loss == pytest.approx(float(l2), rel=1e-12)
    np.testing.assert_allclose(grads, auto, rtol=1e-9, atol=1e-12 * np.abs(auto).max())


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
def test_mfma_register_layout():
    rng = np.random.default_rng(1)
    a = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    b = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    d = np.zeros((16, 16), dtype=np.float32)
    N.check(N.lib().dxg_mfma_probe(_pf(a), _pf(b), _pf(d)))
    np.testing.assert_array_equal(d, a @ b)


CONV_CASES = [(2, (5, 8, 64)), (1, (3, 4, 128)), (1, (1, 4, 64))]


def _conv_inputs(seed, n, dims):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n,) + dims + (32,)).astype(np.float32)
    w = (rng.normal(0.0, 1.0, (32, 27, 32)) / np.sqrt(27 * 32)).astype(np.float32)
    b = rng.normal(0.0, 0.1, 32).astype(np.float32)
    return x, w, b


def _torch_w(w):
    return torch.from_numpy(_bf16(w).reshape(32, 3, 3, 3, 32).transpose(0, 4, 1, 2, 3).copy())


@pytest.mark.gpu
@pytest.mark.parametrize("n,dims", CONV_CASES)
def test_conv_forward_matches_torch(n, dims):
    x, w, b = _conv_inputs(11, n, dims)
    out = _probe(0, x, w.reshape(-1), b, n, dims).reshape(x.shape)
    want = ref.conv(torch.from_numpy(_bf16(x)), _torch_w(w), torch.from_numpy(b)).numpy()
    # FP32 accumulation of 864 BF16 products in another order: ~1e-6 relative
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-5 * np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("n,dims", CONV_CASES + [(1, (3, 5, 192))])
def test_first_layer_kernel_matches_torch(n, dims):
    """The first layer's kernel (four input channels packed over the taps):
    relu(conv + bias) rounded to BF16, against torch on the same BF16 inputs.
    The extra case has an odd row count (a half-filled row pair) and an x tile
    with neighbours on both sides."""
    x, w, b = _conv_inputs(15, n, dims)
    x[..., 4:] = 0.0
    out = _probe(3, x, w.reshape(-1), b, n, dims).reshape(x.shape)
    want = torch.relu(ref.conv(torch.from_numpy(_bf16(x)), _torch_w(w), torch.from_numpy(b))).numpy()
    # FP32 sums of 108 products in another order, then BF16 rounding: within one
    # BF16 step (2^-8 relative) of the rounded reference
    want_bf = _bf16(want)
    tol = np.abs(want_bf) * 2.0**-7 + 2e-5 * np.abs(want).max()
    bad = np.abs(out - want_bf) > tol
    assert not bad.any(), f"{int(bad.sum())} of {out.size} outputs off"
    # and the channels feed it: an input in channel 3 alone changes the output
    x2 = x.copy()
    x2[..., 3] += 1.0
    assert np.abs(_probe(3, x2, w.reshape(-1), b, n, dims) - out.ravel()).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n,dims", CONV_CASES)
def test_conv_input_grad_matches_torch(n, dims):
    g, w, _ = _conv_inputs(12, n, dims)
    out = _probe(1, g, w.reshape(-1), None, n, dims).reshape(g.shape)
    want = ref.conv_input_grad(torch.from_numpy(_bf16(g)), _torch_w(w)).numpy()
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-5 * np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("n,dims", CONV_CASES)
def test_conv_weight_grad_matches_torch(n, dims):
    x, w, _ = _conv_inputs(13, n, dims)
    g = np.random.default_rng(14).normal(0.0, 1.0, x.shape).astype(np.float32)
    out = _probe(2, x, w.reshape(-1), g, n, dims)
    xt, gt = torch.from_numpy(_bf16(x)), torch.from_numpy(_bf16(g))
    want = ref.flat_grad(ref.conv_weight_grad(xt, gt), gt.sum(dim=(0, 1, 2, 3)))
    # sums over up to 6144 voxels in FP32
    np.testing.assert_allclose(out, want, rtol=0, atol=5e-5 * np.abs(want).max())


def _net_case(seed, batch, dims):
    rng = np.random.default_rng(seed)
    f = np.empty((batch,) + dims + (4,), dtype=np.float32)
    f[..., 0] = rng.normal(0.0, 1.0, (batch,) + dims)
    f[..., 1] = rng.uniform(0.0, 2.0, (batch,) + dims)
    f[..., 2] = rng.uniform(0.8, 1.6, (batch,) + dims)
    f[..., 3] = rng.uniform(0.0, 1.0, (batch,) + dims)
    target = (f[..., 0] * 0.5 + rng.normal(0.0, 0.1, (batch,) + dims)).astype(np.float32)
    mask = (rng.random((batch,) + dims) < 0.7).astype(np.float32)
    return f, target, mask


def _rounding_noise(a_gpu, a_bf16, a_exact):
    """GPU deviation from the BF16-rounded restatement, and the rounded
    restatement's own deviation from the exact (FP64, unrounded) network.
    Rounding flips compound over 18 layers, so two BF16 evaluations that sum
    in different orders differ by about as much as either differs from the
    exact result; a wrong tap, gate or bias would differ by far more."""
    return float(np.mean(np.abs(a_gpu - a_bf16))), float(np.mean(np.abs(a_bf16 - a_exact)))


@pytest.mark.gpu
def test_network_prediction_matches_reference():
    params = ref.random_params(21, scale=0.8)
    net = dn.Denoiser(seed=1)
    net.set_params(params)
    f, _, _ = _net_case(22, 1, (4, 8, 64))
    pred = net.infer(f[0]).astype(np.float64)
    want = ref.forward(params, f)[0].numpy()[0].astype(np.float64)
    exact = ref.forward(params, f.astype(np.float64), rounding=False, dtype=torch.float64)[0].numpy()[0]
    dev, noise = _rounding_noise(pred, want, exact)
    scale = np.abs(exact).max()
    print(f"prediction: GPU vs BF16 restatement {dev:.2e}, restatement vs exact {noise:.2e}, max {scale:.2f}")
    assert dev <= 2.0 * noise + 1e-6 * scale
    assert np.abs(pred - exact).max() <= 0.03 * scale


@pytest.mark.gpu
def test_network_loss_and_gradients_match_reference():
    params = ref.random_params(31, scale=0.8)
    net = dn.Denoiser(seed=1)
    net.set_params(params)
    f, target, mask = _net_case(32, 2, (4, 8, 64))
    loss = net.train_step(f, target, mask, lr=1e-6)
    grads = net.grads().astype(np.float64)
    rloss, _, rgrads = ref.loss_and_grads(params, f, target, mask)
    eloss, _, egrads = ref.loss_and_grads(params, f.astype(np.float64), target.astype(np.float64),
                                          mask.astype(np.float64), rounding=False, dtype=torch.float64)
    print(f"loss: GPU {loss:.6f}, BF16 restatement {rloss:.6f}, exact {eloss:.6f}")
    assert abs(loss - rloss) <= 2.0 * abs(rloss - eloss) + 1e-4 * eloss
    for l in range(ref.LAYERS):
        sl = slice(l * ref.LAYER_P, (l + 1) * ref.LAYER_P)
        g, r, e = grads[sl], rgrads[sl], egrads[sl]
        scale = np.abs(e).max()
        assert scale > 0, l
        dev, noise = _rounding_noise(g, r, e)
        print(f"layer {l:2d}: GPU vs BF16 restatement {dev / scale:.2e}, restatement vs exact {noise / scale:.2e}")
        assert dev <= 2.0 * noise + 1e-5 * scale, l
        assert np.abs(g - e).max() <= 0.05 * scale, l
        # the padded channels stay exactly zero
        if l == 0:
            assert np.all(g[:ref.LAYER_W].reshape(32, 27, 32)[:, :, 4:] == 0)
        if l == ref.LAYERS - 1:
            assert np.all(g[:ref.LAYER_W].reshape(32, 27, 32)[1:] == 0) and np.all(g[ref.LAYER_W + 1:] == 0)


@pytest.mark.gpu
def test_training_lowers_the_loss():
    """Noisy copies of a smooth dose field: the untrained network is the
    identity (loss = noise variance); training must remove part of the noise."""
    rng = np.random.default_rng(41)
    dims = (8, 16, 64)
    z, y, x = np.meshgrid(*[np.linspace(-1, 1, n) for n in dims], indexing="ij")
    net = dn.Denoiser(seed=3)
    losses = []
    for step in range(60):
        f = np.zeros((2,) + dims + (4,), dtype=np.float32)
        t = np.zeros((2,) + dims, dtype=np.float32)
        for b in range(2):
            c = rng.uniform(-0.5, 0.5, 3)
            field = np.exp(-2.0 * ((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2))
            t[b] = field
            f[b, ..., 0] = field + rng.normal(0.0, 0.5, dims)
            f[b, ..., 1] = 1.0
            f[b, ..., 2] = 1.0
            f[b, ..., 3] = 0.5
        losses.append(net.train_step(f, t, np.ones_like(t), lr=1e-3))
    assert losses[0] == pytest.approx(0.25, rel=0.1)
    assert np.mean(losses[-10:]) < 0.5 * losses[0], losses
    assert net.last_ms() > 0


@pytest.mark.gpu
def test_world_features_and_identity_apply():
    """Device features from the world's tally equal the host features from
    DoseScore; an untrained network (zero correction layer) returns the dose."""
    import opendxmc_amd as dx
    from opendxmc_amd import phantoms

    ph = phantoms.ct_body((96, 64, 24), (0.4, 0.4, 0.4), seed=3)     # x padded 96 -> 128
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    beam = dx.CTSpiralBeam((0, 0, -4), (0, 0, 4))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(10)
    beam.setTubeVoltage(120)
    beam.setNumberOfParticlesPerExposure(4000)
    dx.Transport(seed=5)(world, beam, None, True)

    net = dn.Denoiser(seed=1)
    fdev = net.world_features(grid)
    assert fdev.shape == (24, 64, 128, 4)
    assert np.all(fdev[:, :, 96:] == 0)
    vol = dn.DoseVolume.from_grid(grid, ph.dims, ph.density, ph.material, dn.material_embedding(ph.materials))
    fhost, dref = dn.features(vol)
    fd = fdev[:, :, :96]
    # device density is the packed FP32 word (2^-15 relative truncation)
    np.testing.assert_allclose(fd[..., 1], fhost[..., 1], rtol=4e-5)
    np.testing.assert_allclose(fd[..., 2], fhost[..., 2], rtol=1e-6)
    np.testing.assert_allclose(fd[..., 0], fhost[..., 0], rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(fd[..., 3], fhost[..., 3], rtol=0, atol=1e-5)

    dose = net.apply(grid)
    d, _, _ = grid.doseScored()
    np.testing.assert_allclose(dose, d, rtol=1e-4, atol=1e-7 * dref)
    assert net.last_ms() > 0


@pytest.mark.gpu
def test_apply_on_body_tiles_equals_full_volume(monkeypatch):
    """dxg_denoiser_apply computes only the 64-voxel tiles within the
    network's reach of the body; body-voxel doses equal the full computation."""
    import opendxmc_amd as dx
    from opendxmc_amd import phantoms

    ph = phantoms.ct_body((448, 96, 40), (0.4, 0.4, 0.4), seed=4)   # wide field: most x tiles are air
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(ph.dims, ph.density, ph.material, [dx.Material.byWeight(m) for m in ph.materials])
    grid.setSpacing(ph.spacing)
    world.build()
    beam = dx.CTSpiralBeam((0, 0, -6), (0, 0, 6))
    beam.setCollimation(4.0)
    beam.setStepAngleDeg(10)
    beam.setNumberOfParticlesPerExposure(3000)
    dx.Transport(seed=9)(world, beam, None, True)

    net = dn.Denoiser(seed=1)
    net.set_params(ref.random_params(51, scale=0.6))
    monkeypatch.setenv("DXG_DENOISE_SKIP", "0")
    full = net.apply(grid)
    t_full = net.last_ms()
    monkeypatch.setenv("DXG_DENOISE_SKIP", "1")
    part = net.apply(grid)
    part2 = net.apply(grid)          # cached tile list and zeroed buffers
    t_part = net.last_ms()
    body = ph.material != 0
    # the network outputs agree exactly; D_ref is a sum of FP64 atomics, so the
    # doses agree to its rounding (~1e-16)
    np.testing.assert_allclose(part[body], full[body], rtol=1e-12, atol=0)
    np.testing.assert_allclose(part2[body], full[body], rtol=1e-12, atol=0)
    print(f"full volume {t_full:.2f} ms, body tiles {t_part:.2f} ms")


def _built_world(dims, spacing, material, density, mats=({7: 0.76, 8: 0.24}, {1: 0.111894, 8: 0.888106})):
    import opendxmc_amd as dx
    world = dx.World()
    grid = world.addItem(dx.AAVoxelGrid(correction=1))
    grid.setData(dims, density.ravel(), material.ravel(), [dx.Material.byWeight(m) for m in mats])
    grid.setSpacing(spacing)
    world.build()
    return world, grid


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["phantom", "ragged", "empty", "full", "thin", "specks"])
def test_device_tile_list_equals_host_scan(case):
    """The row-pair tile lists built on the device (tile flags, dilation by one
    x tile and 19 - layer rows / planes, ordered compaction) equal the host
    scan's, entry for entry: a seeded body, x not a multiple of 64 with an odd
    row count, no body at all, body everywhere, fewer planes than the reach,
    and single body voxels near the volume's faces and corners."""
    from opendxmc_amd import phantoms
    rng = np.random.default_rng(7)
    mats = ({7: 0.76, 8: 0.24}, {1: 0.111894, 8: 0.888106})
    if case == "phantom":
        ph = phantoms.ct_body((448, 96, 40), (0.4, 0.4, 0.4), seed=4)
        dims, mat, mats = ph.dims, ph.material.reshape(ph.dims[::-1]).astype(np.uint8), ph.materials
    else:
        dims = {"ragged": (100, 37, 30), "empty": (128, 16, 8), "full": (64, 9, 5), "thin": (200, 50, 3),
                "specks": (300, 81, 45)}[case]
        nx, ny, nz = dims
        mat = np.zeros((nz, ny, nx), dtype=np.uint8)
        if case == "ragged":
            mat[rng.random(mat.shape) < 0.002] = 1
        elif case == "full":
            mat[:] = 1
        elif case == "thin":
            mat[1, 25, 130] = 1
        elif case == "specks":
            for z, y, x in [(0, 0, 0), (44, 80, 299), (20, 40, 64), (0, 80, 63), (44, 0, 192), (22, 1, 255)]:
                mat[z, y, x] = 1
    dens = np.where(mat != 0, 1.0, 0.0012)
    world, grid = _built_world(dims, (0.3, 0.3, 0.3), mat, dens, mats)
    net = dn.Denoiser(seed=1)
    sizes = []
    for layer in (0, 1, 9, 16, 17):   # the lists shrink with the layers' remaining reach (19 - layer voxels)
        dev = net.debug_tiles(grid, layer=layer)
        host = net.debug_tiles(grid, host=True, layer=layer)
        np.testing.assert_array_equal(dev, host)
        sizes.append(host.size)
        if case == "empty":
            assert host.size == 0
        if case == "full":
            nxt, nyp = 1, (dims[1] + 1) // 2
            assert host.size == dims[2] * nyp * nxt
    print(f"{case}: {sizes} row pairs for layers 0, 1, 9, 16, 17")
    assert sizes == sorted(sizes, reverse=True)
