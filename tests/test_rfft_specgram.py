"""r2c / c2r through a half-length complex FFT (examples/demo_specialized_fft.py:
156-331, its tests :600-640) and the spectrogram transformation chain
(examples/demo_specgram.py:196-267), against numpy FP64."""

import numpy
import pytest

import oracle
from reikna_amd import Array, ArrayMetadata
from reikna_amd.fft.rfft import IRFFT, RFFT
from reikna_amd.fft.spectrogram import Spectrogram
from reikna_amd.runtime import VirtualDevice

RSHAPES = [(1024,), (8,), (4, 3000), (2, 3, 256), (3, 1994), (2, 2**16)]


def specgram_reference(x, fft_size, noverlap, pad_to):
    """|FFT| of von-Hann-windowed overlapping frames, non-negative
    frequencies, (frequencies, times): matplotlib's specgram(mode='magnitude')
    times the window sum, in FP64."""
    step = fft_size - noverlap
    frames = (x.size - noverlap) // step
    size = fft_size if pad_to is None else pad_to
    n = numpy.arange(fft_size)
    win = 0.5 * (1 - numpy.cos(2 * numpy.pi * n / (fft_size - 1)))
    rows = numpy.zeros((frames, size))
    for i in range(frames):
        rows[i, :fft_size] = x[i * step: i * step + fft_size].astype(numpy.float64) * win
    spec = numpy.abs(numpy.fft.fft(rows, axis=1))[:, : size // 2 + 1]
    return spec.T


@pytest.mark.parametrize("params", [dict(fft_size=256, noverlap=128, pad_to=None),
                                    dict(fft_size=1024, noverlap=900, pad_to=2048),
                                    dict(fft_size=100, noverlap=0, pad_to=128)])
def test_specgram_reference_against_matplotlib(params):
    """The FP64 restatement above is matplotlib's spectrogram (the reference
    demo's own comparison, demo_specgram.py:59,306): mlab.specgram in
    magnitude mode divides |FFT| by the window sum; detrending off, the
    symmetric Hann window, one-sided spectrum.  A planted mistake (the
    periodic Hann window) is caught."""
    mlab = pytest.importorskip("matplotlib.mlab")
    rng = numpy.random.default_rng(12)
    x = rng.standard_normal(20000)
    spec, _, _ = mlab.specgram(x, NFFT=params["fft_size"], Fs=2.0, noverlap=params["noverlap"],
                               pad_to=params["pad_to"], mode="magnitude")
    win_sum = numpy.hanning(params["fft_size"]).sum()
    ref = specgram_reference(x, **params)
    assert ref.shape == spec.shape
    assert numpy.allclose(ref, spec * win_sum, rtol=1e-10, atol=1e-10 * numpy.abs(ref).max())
    # planted mistake: the periodic window 0.5 - 0.5 cos(2 pi n / N) differs
    n = numpy.arange(params["fft_size"])
    periodic = 0.5 - 0.5 * numpy.cos(2 * numpy.pi * n / params["fft_size"])
    assert not numpy.allclose(periodic, numpy.hanning(params["fft_size"]))


def test_hann_window_is_numpy_hanning():
    from reikna_amd.fft.spectrogram import hann_window

    for n in (1, 2, 7, 256, 1000):
        assert numpy.allclose(hann_window(n), numpy.hanning(n), rtol=0, atol=1e-15)


def test_errors():
    with pytest.raises(ValueError):
        RFFT(ArrayMetadata((15,), numpy.float32))
    with pytest.raises(ValueError):
        RFFT(ArrayMetadata((16,), numpy.complex64))
    with pytest.raises(ValueError):
        IRFFT(ArrayMetadata((9,), numpy.float32))


@pytest.mark.parametrize("shape", RSHAPES[:4])
def test_compile(shape):
    vdev = VirtualDevice()
    RFFT(ArrayMetadata(shape, numpy.float32)).compile(vdev)
    IRFFT(ArrayMetadata((*shape[:-1], shape[-1] // 2 + 1), numpy.complex64)).compile(vdev)


def test_spectrogram_compiles():
    x = ArrayMetadata((8000,), numpy.float32)
    sp = Spectrogram(x, fft_size=256, noverlap=200, pad_to=512)
    assert sp.parameter.output.shape == (257, (8000 - 200) // 56)
    sp.compile(VirtualDevice())


@pytest.mark.gpu
@pytest.mark.parametrize("shape", RSHAPES, ids=[str(s) for s in RSHAPES])
def test_rfft_gpu(queue, shape):
    a = numpy.random.default_rng(shape[-1]).normal(size=shape).astype(numpy.float32)
    a_dev = Array.from_host(queue, a)
    rf = RFFT(a_dev)
    out = Array.empty_like(queue.device, rf.parameter.output)
    rf.compile(queue.device)(queue, out, a_dev)
    ref = oracle.fftn(a.astype(numpy.complex64), (len(shape) - 1,))[..., : shape[-1] // 2 + 1]
    err = oracle.rel_l2(out.get(queue), ref)
    assert err <= oracle.tolerance((shape[-1],)), err
    # the reference demo's own check (demo_specialized_fft.py:611)
    assert numpy.allclose(numpy.fft.rfft(a), out.get(queue), atol=1e-4 * numpy.sqrt(shape[-1]), rtol=1e-4)


@pytest.mark.gpu
def test_rfft_dont_store_last(queue):
    a = numpy.random.default_rng(5).normal(size=(4, 512)).astype(numpy.float32)
    a_dev = Array.from_host(queue, a)
    rf = RFFT(a_dev, dont_store_last=True)
    assert rf.parameter.output.shape == (4, 256)
    out = Array.empty_like(queue.device, rf.parameter.output)
    rf.compile(queue.device)(queue, out, a_dev)
    ref = oracle.fftn(a.astype(numpy.complex64), (1,))[:, :256]
    assert oracle.rel_l2(out.get(queue), ref) <= oracle.tolerance((512,))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", RSHAPES, ids=[str(s) for s in RSHAPES])
def test_irfft_gpu(queue, shape):
    rng = numpy.random.default_rng(shape[-1] + 1)
    spec_shape = (*shape[:-1], shape[-1] // 2 + 1)
    fa = (rng.normal(size=spec_shape) + 1j * rng.normal(size=spec_shape)).astype(numpy.complex64)
    fa_dev = Array.from_host(queue, fa)
    irf = IRFFT(fa_dev)
    out = Array.empty_like(queue.device, irf.parameter.output)
    irf.compile(queue.device)(queue, out, fa_dev)
    ref = numpy.fft.irfft(fa.astype(numpy.complex128), axis=-1)
    err = numpy.linalg.norm(out.get(queue) - ref) / numpy.linalg.norm(ref)
    assert err <= oracle.tolerance((shape[-1],)), err


@pytest.mark.gpu
@pytest.mark.parametrize("params", [dict(fft_size=1024, noverlap=900, pad_to=2048), dict(fft_size=256, noverlap=128, pad_to=None)])
def test_spectrogram_gpu(queue, params):
    dt = 0.0005
    t = numpy.arange(0.0, 20.0, dt)
    rng = numpy.random.default_rng(3)
    x = (numpy.sin(2 * numpy.pi * 100 * t) + 2 * numpy.sin(2 * numpy.pi * 400 * t) * ((t > 10) & (t < 12))
         + 0.01 * rng.standard_normal(t.size)).astype(numpy.float32)
    x_dev = Array.from_host(queue, x)
    sp = Spectrogram(x_dev, **params)
    out = Array.empty_like(queue.device, sp.parameter.output)
    sp.compile(queue.device)(queue, out, x_dev)
    ref = specgram_reference(x, **params)
    res = out.get(queue)
    assert res.shape == ref.shape
    # the reference demo's check (demo_specgram.py:251)
    assert numpy.allclose(res, ref, atol=1e-4 * params["fft_size"], rtol=1e-4)
    assert numpy.linalg.norm(res - ref) / numpy.linalg.norm(ref) <= oracle.tolerance((res.shape[0] * 2,))


@pytest.mark.gpu
def test_spectrogram_custom_window_gpu(queue):
    """A window other than Hann (a Blackman window as an array, and a
    function of the length) reaches the fused input transformation."""
    rng = numpy.random.default_rng(4)
    x = rng.standard_normal(30000).astype(numpy.float32)
    x_dev = Array.from_host(queue, x)
    for window in (numpy.blackman(512), numpy.bartlett):
        sp = Spectrogram(x_dev, fft_size=512, noverlap=384, pad_to=1024, window=window)
        out = Array.empty_like(queue.device, sp.parameter.output)
        sp.compile(queue.device)(queue, out, x_dev)
        w = window if isinstance(window, numpy.ndarray) else window(512)
        step = 512 - 384
        frames = (x.size - 384) // step
        rows = numpy.zeros((frames, 1024))
        for i in range(frames):
            rows[i, :512] = x[i * step: i * step + 512].astype(numpy.float64) * w
        ref = numpy.abs(numpy.fft.fft(rows, axis=1))[:, :513].T
        res = out.get(queue)
        assert res.shape == ref.shape
        assert numpy.linalg.norm(res - ref) / numpy.linalg.norm(ref) <= oracle.tolerance((1024,))
