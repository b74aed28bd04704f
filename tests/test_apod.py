"""Learned apodization: the MLP checker pinned on the CPU, and the HIP
training/inference kernels against it on the GPU.

The network has no reference counterpart. The checker (oracle/apod.py) is
pinned here by closed forms and by central finite differences of its own
float64 loss; the GPU gradients and frames are then compared with the
checker evaluated with the kernels' bf16 rounding points (tight bars) and
without them (loose bars: bf16 arithmetic itself).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import bf_config, make_case, oracle_inputs
from conftest import rel_l2

from oracle import apod as A

F = A.F


def _rand_params(rng, scale=0.08, bias=0.05, bf16_weights=True):
    Ws = [rng.normal(0, scale, (F, F)) for _ in range(A.LAYERS)]
    if bf16_weights:
        Ws = [A.bf16(w) for w in Ws]
    bs = [rng.normal(0, bias, F).astype(np.float32).astype(np.float64) for _ in range(A.LAYERS)]
    return A.pack(Ws, bs).astype(np.float32)


def _rand_batch(rng, n, n_active_min=4):
    x = A.bf16(rng.normal(0, 1, (n, F)))
    g = rng.normal(1, 0.5, (n, F)).astype(np.float32)
    mask = np.zeros((n, 4), np.uint32)
    for b in range(n):
        lo = rng.integers(0, 100)
        hi = min(128, lo + rng.integers(n_active_min, 60))
        for ch in range(lo, hi):
            mask[b, ch // 32] |= np.uint32(1 << (ch % 32))
    act = A.active_features(mask)
    x = x * act
    na = act[:, ::2].sum(1)
    sw = np.where(rng.random(n) < 0.9, 1.0 / na, 0.0).astype(np.float32)
    return x, g, mask, sw


# ---------------------------------------------------------------------------
# CPU: the checker itself
# ---------------------------------------------------------------------------

def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 65504.0,
                  1.0 + 2 ** -8 + 2 ** -20, np.inf], np.float32)
    r = A.bf16(v)
    # ties to even at 1 + 2^-8 (down) and 1 + 3 * 2^-8 (up); above the tie rounds up
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7]
    assert r[4] == -2.5 and r[5] == 65536.0 and r[6] == 1.0 + 2 ** -7 and np.isinf(r[7])
    assert np.array_equal(A.from_bits(A.bf16_bits(v)), r)


def test_gelu_closed_forms():
    z = np.array([-30.0, -1.0, 0.0, 1.0, 30.0])
    g = A.gelu(z)
    assert g[2] == 0.0 and g[0] == pytest.approx(0.0, abs=1e-12) and g[4] == pytest.approx(30.0)
    # tanh form at 1: 0.5 (1 + tanh(sqrt(2/pi) 1.044715))
    assert g[3] == pytest.approx(0.5 * (1 + np.tanh(np.sqrt(2 / np.pi) * 1.044715)), rel=1e-15)
    assert g[3] - g[1] == pytest.approx(1.0, rel=1e-12)  # gelu(z) - gelu(-z) = z
    h = 1e-6
    zz = np.linspace(-4, 4, 41)
    fd = (A.gelu(zz + h) - A.gelu(zz - h)) / (2 * h)
    assert np.max(np.abs(fd - A.gelu_grad(zz))) < 1e-8
    assert A.gelu_grad(np.array([0.0]))[0] == 0.5


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _rand_params(rng, scale=0.1, bias=0.1, bf16_weights=False).astype(np.float64)
    x, g, mask, sw = _rand_batch(rng, 6)
    loss, grad = A.loss_and_grad(p, x, g, mask, sw, emulate=False)
    assert loss > 0
    h = 1e-5
    picks = []
    for l in range(A.LAYERS):
        for _ in range(6):
            picks.append(l * F * F + int(rng.integers(0, F * F)))
        for _ in range(4):
            picks.append(A.LAYERS * F * F + l * F + int(rng.integers(0, F)))
    for i in picks:
        q = p.copy()
        q[i] += h
        lp, _ = A.loss_and_grad(q, x, g, mask, sw, emulate=False)
        q[i] -= 2 * h
        lm, _ = A.loss_and_grad(q, x, g, mask, sw, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_zero_network_closed_form():
    """W = 0, b = 0: the weights are all 1 (DAS), the loss is the mean of
    sw |1 - g|^2 over active channels and only the last layer's bias (and
    weights through H_3 = gelu(0) = 0: none) gets a gradient."""
    rng = np.random.default_rng(4)
    x, g, mask, sw = _rand_batch(rng, 10)
    p = np.zeros(A.N_PARAMS)
    loss, grad = A.loss_and_grad(p, x, g, mask, sw, emulate=False)
    act = A.active_features(mask)
    one = np.tile([1.0, 0.0], F // 2)
    e = (one - g) * act
    assert loss == pytest.approx(np.sum(sw.astype(np.float64)[:, None] * e * e) / 10, rel=1e-12)
    gb3 = grad[A.LAYERS * F * F + 3 * F:]
    assert np.allclose(gb3, (2.0 / 10 * sw.astype(np.float64)[:, None] * e).sum(0), rtol=1e-12, atol=1e-15)
    assert np.all(grad[:A.LAYERS * F * F] == 0)
    assert np.all(grad[A.LAYERS * F * F:A.LAYERS * F * F + 3 * F] == 0)


def test_features_are_phase_and_scale_free():
    rng = np.random.default_rng(5)
    y = rng.normal(size=(7, 128)) + 1j * rng.normal(size=(7, 128))
    _, _, mask, _ = _rand_batch(rng, 7)
    x = A.features(y, mask)
    x2 = A.features(y * (3.7 * np.exp(1j * 0.9)), mask)
    assert np.allclose(x, x2, atol=1e-12)
    act = A.active_features(mask)
    na = act[:, ::2].sum(1)
    assert np.allclose((x ** 2).sum(1), na)  # unit mean power over active channels
    xc = x[:, 0::2] + 1j * x[:, 1::2]
    assert np.allclose(xc.sum(1).imag, 0, atol=1e-12) and np.all(xc.sum(1).real >= 0)
    assert np.all(x[act == 0] == 0)


# ---------------------------------------------------------------------------
# GPU: the HIP kernels against the checker
# ---------------------------------------------------------------------------

def _runner(usx, case, combiner="MV", subarray=16):
    pc = usx.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bf_config(usx, combiner, subarray=subarray)
    return usx.DeviceRunner(case.grid, pc, case.raw), pc


@pytest.fixture(scope="module")
def apod_case(usx_mod, O):
    # 128 channels (the network's width), a small grid
    return make_case(usx_mod, O, n_el=128, n_samples=1024, n_tx=3, n_scat=1500,
                     grid=(24, 0.008, 0.030, 64), box=(-0.02, 0.02, 0.005, 0.032))


def _split(grads):
    nw = A.LAYERS * F * F
    return ([grads[l * F * F:(l + 1) * F * F] for l in range(A.LAYERS)] +
            [grads[nw + l * F:nw + (l + 1) * F] for l in range(A.LAYERS)])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 300])
def test_gradient_matches_oracle(usx_mod, apod_case, n):
    rng = np.random.default_rng(10 + n)
    p = _rand_params(rng)
    x, g, mask, sw = _rand_batch(rng, n)
    run, _ = _runner(usx_mod, apod_case)
    run.set_apodization(p)
    loss, grads = run.gradient(A.bf16_bits(x).reshape(n, F), g, mask, sw)
    loss_e, ref_e = A.loss_and_grad(p.astype(np.float64), x, g, mask, sw, emulate=True)
    loss_x, ref_x = A.loss_and_grad(p.astype(np.float64), x, g, mask, sw, emulate=False)
    assert loss == pytest.approx(loss_e, rel=1e-5)
    for i, (a, b, c) in enumerate(zip(_split(grads), _split(ref_e), _split(ref_x))):
        assert rel_l2(a, b) < 2e-3, (i, rel_l2(a, b))   # bf16 rounding points emulated
        assert rel_l2(a, c) < 3e-2, (i, rel_l2(a, c))   # against exact float64


@pytest.mark.gpu
def test_mv_weights_match_oracle_and_reproduce_the_frame(usx_mod, port, O, apod_case):
    run, pc = _runner(usx_mod, apod_case)
    frame, w, st = run.mv_weights(apod_case.raw)
    s, txs = oracle_inputs(O, apod_case)
    ref, ex = port.process_frame(s, txs, np.asarray(apod_case.raw.rf), apod_case.gspec,
                                 O.Demod(decimation=4), O.Bf.of(pc.beamformer))
    assert rel_l2(frame, ref) < 1e-4
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(apod_case.raw, d)
    s2 = O.Setup.of(apod_case.probe, apod_case.medium, iq.acq)
    ref2, ex2 = port.beamform(s2, txs, np.asarray(iq.iq), apod_case.gspec, O.Bf.of(pc.beamformer),
                              fdem=iq.demodulation_frequency, weights=True, ych=True, info=True)
    wr = ex2["weights"].reshape(w.shape)
    assert rel_l2(w, wr) < 1e-4
    assert np.array_equal(w != 0, wr != 0)
    # out == sum_ch g_ch Y_ch on the oracle's channel vectors
    y = ex2["ych"].reshape(w.shape)
    assert rel_l2((w * y).sum(-1), ref2) < 1e-4
    fb = ex2["info"][:, 2].reshape(st.shape)
    assert np.array_equal(st == 0, (fb == 0) & (ex2["info"][:, 0].reshape(st.shape) >= 4))


@pytest.fixture(scope="module")
def apod_ragged(usx_mod, O):
    # 23 x 61 = 1403 pixels: the inference kernel's last 128-pixel tile is
    # short (123 pixels), so its loads past the frame must read zeros
    return make_case(usx_mod, O, n_el=128, n_samples=1024, n_tx=3, n_scat=1500,
                     grid=(23, 0.008, 0.030, 61), box=(-0.02, 0.02, 0.005, 0.032))


@pytest.mark.gpu
@pytest.mark.parametrize("case_name", ["apod_case", "apod_ragged"])
def test_learned_frame_matches_oracle(usx_mod, port, O, case_name, request):
    apod_case = request.getfixturevalue(case_name)
    run_mv, pc = _runner(usx_mod, apod_case)
    d = usx_mod.DemodConfig()
    d.decimation = 4
    iq = usx_mod.demodulate(apod_case.raw, d)
    s2 = O.Setup.of(apod_case.probe, apod_case.medium, iq.acq)
    _, txs = oracle_inputs(O, apod_case)
    _, ex = port.beamform(s2, txs, np.asarray(iq.iq), apod_case.gspec, O.Bf.of(pc.beamformer),
                          fdem=iq.demodulation_frequency, weights=True, ych=True)
    y = ex["ych"]
    on = ex["weights"] != 0
    mask = np.zeros((y.shape[0], 4), np.uint32)
    for q in range(4):
        mask[:, q] = (on[:, 32 * q:32 * q + 32].astype(np.uint64) << np.arange(32, dtype=np.uint64)).sum(1).astype(np.uint32)
    rng = np.random.default_rng(7)
    pcl = usx_mod.PipelineConfig()
    pcl.demod.decimation = 4
    pcl.beamformer = bf_config(usx_mod, "Learned", subarray=16)
    pipe = usx_mod.ImagingPipeline(apod_case.grid, pcl)
    # zero network: DAS with the rectangular aperture
    pipe.set_apodization(np.zeros(A.N_PARAMS, np.float32))
    got0 = np.asarray(pipe.process_frame(apod_case.raw).data).ravel()
    assert rel_l2(got0, (y * on).sum(1)) < 1e-5
    p = _rand_params(rng, scale=0.06, bias=0.05)
    pipe.set_apodization(p)
    got = np.asarray(pipe.process_frame(apod_case.raw).data).ravel()
    ref_e = A.apply_weights(p.astype(np.float64), y, mask, emulate=True)
    ref_x = A.apply_weights(p.astype(np.float64), y, mask, emulate=False)
    assert rel_l2(got, ref_e) < 2e-3, rel_l2(got, ref_e)
    assert rel_l2(got, ref_x) < 3e-2, rel_l2(got, ref_x)


@pytest.mark.gpu
def test_learned_frame_pixel_limit(usx_mod, apod_case):
    """The inference kernel reaches the channel vectors (1 KiB a pixel) through
    32-bit buffer offsets: a frame of more than 2^22 - 128 pixels is refused
    before anything runs."""
    pcl = usx_mod.PipelineConfig()
    pcl.demod.decimation = 4
    pcl.beamformer = bf_config(usx_mod, "Learned", subarray=16)
    big = usx_mod.ScanGrid.cartesian(-0.01, 0.01, 2048, 0.005, 0.03, 2048)
    pipe = usx_mod.ImagingPipeline(big, pcl)
    pipe.set_apodization(np.zeros(A.N_PARAMS, np.float32))
    with pytest.raises(ValueError, match="pixels per frame"):
        pipe.process_frame(apod_case.raw)


@pytest.mark.gpu
def test_learned_combiner_falls_back_to_mv(usx_mod, apod_case):
    pc = usx_mod.PipelineConfig()
    pc.demod.decimation = 4
    pc.beamformer = bf_config(usx_mod, "Learned", subarray=16)
    pipe = usx_mod.ImagingPipeline(apod_case.grid, pc)
    assert not pipe.has_apodization
    got = np.asarray(pipe.process_frame(apod_case.raw).data)
    pc.beamformer = bf_config(usx_mod, "MV", subarray=16)
    mv = np.asarray(usx_mod.ImagingPipeline(apod_case.grid, pc).process_frame(apod_case.raw).data)
    assert np.array_equal(got, mv)


@pytest.mark.gpu
def test_training_on_mv_weights_lowers_the_loss(usx_mod, apod_case):
    run, _ = _runner(usx_mod, apod_case)
    rng = np.random.default_rng(11)
    npix = apod_case.grid.n_axis0 * apod_case.grid.n_axis1
    run.dataset_reserve(npix)
    frame = run.dataset_append(apod_case.raw, np.arange(npix, dtype=np.int32))
    assert run.dataset_size == npix and np.isfinite(frame).all()
    p0 = _rand_params(rng, scale=0.05, bias=0.0)
    nw = A.LAYERS * F * F
    p0[3 * F * F:nw] = 0.0  # last layer starts at zero: the network starts as DAS
    run.set_apodization(p0)
    losses = run.train(150, batch=512, lr=1e-3, seed=1)
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < 0.7 * losses[:10].mean(), (losses[:10].mean(), losses[-10:].mean())
    p1 = run.get_apodization()
    assert not np.array_equal(p0, p1)
