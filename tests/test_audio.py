"""IR post-processing of `simulate` (simulator.cpp:1029-1082): resampling,
normalisation and the WAV writer.  The reference's resampler (kfr 6.3.0) is
not available, so its output is unpinned; these checks pin the
replacement's own properties: passband gain, stop-band attenuation, exact
DC, zero phase (no added delay), and the reference's sample-count and
input-rate rules."""
import numpy as np
import pytest

from virs_amd import audio, grid as G

FS = 305533.0  # the reference simulation rate, 1/dt at h = 343/88200 m, whole Hz


def _tone(f, fs, n, phase=0.3):
    return np.sin(2 * np.pi * f * np.arange(n) / fs + phase)


@pytest.mark.parametrize("f", [100.0, 5000.0, 15000.0, 19500.0])
def test_passband_tones_come_through_unchanged(f):
    n = 40000
    y = audio.resample(_tone(f, FS, n).astype(np.float32), FS, 48000.0)
    ref = _tone(f, 48000.0, len(y))
    mid = slice(300, len(y) - 300)
    assert np.max(np.abs(y[mid] - ref[mid])) < 1e-4


@pytest.mark.parametrize("f", [26000.0, 60000.0, 140000.0])
def test_stop_band_is_attenuated_100_db(f):
    n = 40000
    y = audio.resample(_tone(f, FS, n).astype(np.float32), FS, 48000.0)
    mid = slice(300, len(y) - 300)
    assert np.max(np.abs(y[mid])) < 1e-5


def test_constant_is_reproduced_exactly_away_from_the_ends():
    y = audio.resample(np.ones(20000, np.float32), FS, 48000.0)
    mid = slice(300, len(y) - 300)
    assert np.max(np.abs(y[mid] - 1.0)) < 1e-6


def test_zero_phase_impulse_response():
    """An impulse at input time 6K/288000 s lands centred on output sample K
    with a symmetric response: the resampler adds no delay."""
    fs_in, K = 288000.0, 700
    x = np.zeros(12000, np.float32)
    x[6 * K] = 1.0
    y = audio.resample(x, fs_in, 48000.0)
    assert int(np.argmax(y)) == K
    m = np.arange(1, 60)
    np.testing.assert_allclose(y[K - m], y[K + m], rtol=0, atol=1e-7)
    # its area is the input's: band-limited impulse of unit area at the new rate
    assert abs(float(np.sum(y)) * fs_in / 48000.0 - 1.0) < 1e-3


def test_upsampling_keeps_a_tone():
    y = audio.resample(_tone(1000.0, 48000.0, 4800).astype(np.float32), 48000.0, 96000.0)
    ref = _tone(1000.0, 96000.0, len(y))
    mid = slice(400, len(y) - 400)
    assert np.max(np.abs(y[mid] - ref[mid])) < 1e-4


def test_reference_sample_count_rate_and_normalisation():
    """num_samples = (size_t)(48000 * steps * dt) with dt = 0.5*h/(343*sqrt(3))
    in double, and the simulation rate truncated to whole Hz
    (simulator.cpp:930-933,1031-1063); the output peaks at exactly 1."""
    h = np.float32(G.reference_cell_size())
    steps = 20000
    dt = 0.5 * float(h) / (343.0 * np.sqrt(3.0))
    assert int(1.0 / dt) == int(FS)
    rng = np.random.default_rng(0)
    ir = (rng.standard_normal(steps) * np.exp(-np.arange(steps) / 3000.0)).astype(np.float32)
    y = audio.ir_audio(ir, h)
    assert len(y) == int(48000 * (steps * dt))
    assert float(np.max(np.abs(y))) == 1.0
    ref = audio.resample(ir, int(1.0 / dt), 48000.0, len(y))
    np.testing.assert_allclose(y, ref / np.max(np.abs(ref)), rtol=0, atol=1e-7)


def test_wav_round_trip(tmp_path):
    from scipy.io import wavfile
    x = np.linspace(-1, 1, 4801, dtype=np.float32)
    path = tmp_path / "output_audio.wav"
    audio.write_wav(path, x, 48000)
    rate, data = wavfile.read(path)
    assert rate == 48000 and data.dtype == np.float32
    np.testing.assert_array_equal(data, x)
    raw = path.read_bytes()
    assert raw[:4] == b"RIFF" and raw[8:12] == b"WAVE"
    assert int.from_bytes(raw[20:22], "little") == 3  # IEEE float
