"""Seeded synthetic seismic workloads for tests and benchmarks.

Band-passed Gaussian noise channels with DC offsets, Hann-tapered
band-limited template bursts, moveouts and injections.  Everything is
reproducible from one integer seed: Gaussian samples come from numpy's
PCG64 generator (stable across machines for a given numpy release) and
the Butterworth band-pass is designed here by the bilinear transform and
applied as a causal cascade of biquads.

"4-pole band-pass" means four poles in total: a second-order low-pass
prototype mapped to a band-pass (two biquads).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.signal import sosfilt

SEED_BASE = 0x4D464C00


def butter_bandpass_sos(f_lo: float, f_hi: float, fs: float) -> np.ndarray:
    """Second-order Butterworth low-pass prototype -> band-pass (4 poles),
    bilinear transform with pre-warping.  Returns a (2, 6) SOS array."""
    w1 = 2 * fs * math.tan(math.pi * f_lo / fs)
    w2 = 2 * fs * math.tan(math.pi * f_hi / fs)
    bw = w2 - w1
    w0 = math.sqrt(w1 * w2)
    # prototype poles of a 2nd-order Butterworth: exp(i*pi*(2k+n+1)/(2n)), n=2
    proto = [complex(math.cos(math.pi * (2 * k + 3) / 4), math.sin(math.pi * (2 * k + 3) / 4))
             for k in range(2)]
    # low-pass -> band-pass: each prototype pole p gives roots of
    # s^2 - p*bw*s + w0^2 = 0
    poles = []
    for p in proto:
        a = p * bw / 2
        d = np.sqrt(a * a - w0 * w0 + 0j)
        poles += [a + d, a - d]
    # bilinear transform z = (2fs + s)/(2fs - s)
    fs2 = 2 * fs
    zp = [(fs2 + s) / (fs2 - s) for s in poles]
    # band-pass zeros: two at s=0 (z=1) and two at s=inf (z=-1)
    # pair poles into conjugate biquads
    zp = sorted(zp, key=lambda z: (round(abs(z.imag), 12), z.real))
    upper = [z for z in zp if z.imag > 0]
    if len(upper) != 2:
        raise ValueError("unexpected pole layout")
    sos = np.zeros((2, 6))
    for i, z in enumerate(upper):
        sos[i, 3:] = [1.0, -2 * z.real, abs(z) ** 2]
    sos[0, :3] = [1.0, 2.0, 1.0]     # zeros at z = -1
    sos[1, :3] = [1.0, -2.0, 1.0]    # zeros at z = +1
    # unit gain at the geometric centre frequency
    fc = math.sqrt(f_lo * f_hi)
    zc = np.exp(2j * math.pi * fc / fs)
    g = 1.0
    for s in sos:
        g *= abs((s[0] * zc * zc + s[1] * zc + s[2]) / (s[3] * zc * zc + s[4] * zc + s[5]))
    sos[0, :3] /= g
    return sos


def hann(n: int) -> np.ndarray:
    if n == 1:
        return np.ones(1)
    k = np.arange(n)
    return 0.5 * (1 - np.cos(2 * math.pi * k / (n - 1)))


def bandpassed_noise(rng, n: int, fs: float, f_lo: float, f_hi: float, dtype=np.float64):
    sos = butter_bandpass_sos(f_lo, f_hi, fs)
    return sosfilt(sos, rng.standard_normal(n)).astype(dtype, copy=False)


def template_burst(rng, n: int, fs: float, f_lo: float, f_hi: float) -> np.ndarray:
    """Hann-tapered band-limited noise burst with unit energy (sum t^2 = 1)."""
    pad = 4 * n
    t = bandpassed_noise(rng, n + pad, fs, f_lo, f_hi)[pad:] * hann(n)
    return t / np.linalg.norm(t)


@dataclass
class MultiChannelWorkload:
    """Templates x channels detection workload (the multi-channel configs)."""
    fs: float
    signals: list                     # C arrays (float32)
    templates: list                   # [event][channel] FP64 arrays
    moveouts: np.ndarray              # [event, channel] integer samples
    origins: list = field(default_factory=list)   # [event] injection origin samples
    noise_rms: np.ndarray = None

    @property
    def n_events(self):
        return len(self.templates)

    @property
    def n_channels(self):
        return len(self.signals)


def multichannel_day(seed: int = SEED_BASE + 2, n_events: int = 16, n_channels: int = 24,
                     n_samples: int = 8_640_000, tlen: int = 512, fs: float = 100.0,
                     n_inject: int = 50, max_moveout: int = 200, dc: float = 100.0,
                     amp=(0.3, 1.5)) -> MultiChannelWorkload:
    """16 templates x 24 channels, one day at 100 Hz, FP32 signals.
    Channels: band-passed (1-10 Hz) N(0,1) noise + DC offset U(-dc, dc).
    Templates: per (event, channel) unit-energy Hann-tapered bursts.
    Moveouts: integers U[0, max_moveout].  Each event is injected n_inject
    times; per-channel amplitude U(amp) times the channel noise RMS times
    sqrt(tlen) (so the amplitude is a per-sample RMS ratio)."""
    rng = np.random.default_rng(seed)
    noise = []
    for c in range(n_channels):
        crng = np.random.default_rng([seed, 1, c])
        noise.append(bandpassed_noise(crng, n_samples, fs, 1.0, 10.0))
    rms = np.array([float(np.sqrt(np.mean(x[:100000] ** 2))) for x in noise])
    templates = []
    for e in range(n_events):
        templates.append([template_burst(np.random.default_rng([seed, 2, e, c]), tlen, fs, 1.0, 10.0)
                          for c in range(n_channels)])
    moveouts = rng.integers(0, max_moveout + 1, size=(n_events, n_channels))
    origins = []
    span = n_samples - tlen - max_moveout - 1
    for e in range(n_events):
        orig = np.sort(rng.integers(0, span, size=n_inject))
        origins.append(orig)
        amps = rng.uniform(amp[0], amp[1], size=(n_inject, n_channels))
        for c in range(n_channels):
            t = templates[e][c] * math.sqrt(tlen) * rms[c]
            for k, o in enumerate(orig):
                i0 = int(o) + int(moveouts[e, c])
                noise[c][i0:i0 + tlen] += amps[k, c] * t
    offsets = rng.uniform(-dc, dc, size=n_channels)
    signals = [(noise[c] + offsets[c]).astype(np.float32) for c in range(n_channels)]
    return MultiChannelWorkload(fs=fs, signals=signals, templates=templates, moveouts=moveouts,
                                origins=origins, noise_rms=rms)


def single_channel_hour(seed: int = SEED_BASE + 1, n: int = 360_000, tlen: int = 400,
                        fs: float = 100.0, n_inject: int = 20):
    """One hour at 100 Hz, FP64 band-passed N(0,1) noise, one 400-sample
    unit-energy template injected 20 times with amplitude U(0.5, 2) and SNR
    U(3, 10) (SNR = injected peak amplitude / noise RMS)."""
    rng = np.random.default_rng(seed)
    x = bandpassed_noise(np.random.default_rng([seed, 1]), n, fs, 1.0, 10.0)
    t = template_burst(np.random.default_rng([seed, 2]), tlen, fs, 1.0, 10.0)
    rms = float(np.sqrt(np.mean(x ** 2)))
    pos = np.sort(rng.integers(0, n - tlen, size=n_inject))
    amp = rng.uniform(0.5, 2.0, size=n_inject)
    snr = rng.uniform(3.0, 10.0, size=n_inject)
    for p, a, s in zip(pos, amp, snr):
        x[p:p + tlen] += (s * rms / np.abs(t).max()) * (a / 1.25) * t
    return x, t, pos


@dataclass
class ServingWorkload:
    """Served stream of detection requests: a template library and a set of
    channel streams from which requests cut their windows."""
    fs: float
    library: list          # FP64 unit-energy templates
    streams: list          # C float32 channel streams
    min_window: int
    max_window: int
    max_channels: int
    max_templates: int

    def sample_request(self, rng):
        """One request: window length log-uniform in [min_window, max_window],
        1..max_channels distinct channels, 1..max_templates distinct library
        templates; returns (channels, start, length, template ids)."""
        lo, hi = math.log(self.min_window), math.log(self.max_window)
        n = int(round(math.exp(rng.uniform(lo, hi))))
        n = max(n, max(len(t) for t in self.library) + 1)
        n = min(n, self.streams[0].size)
        nc = int(rng.integers(1, self.max_channels + 1))
        chans = np.sort(rng.choice(len(self.streams), nc, replace=False))
        start = int(rng.integers(0, self.streams[0].size - n + 1))
        nt = int(rng.integers(1, self.max_templates + 1))
        ids = np.sort(rng.choice(len(self.library), nt, replace=False))
        return chans.tolist(), start, n, ids.tolist()


def serving_stream(seed: int = SEED_BASE + 5, n_templates: int = 1000, n_channels: int = 48,
                   stream_len: int = 360_000, fs: float = 100.0, lt=(200, 2000),
                   min_window: int = 6_000, max_window: int = 360_000, max_channels: int = 48,
                   max_templates: int = 100, n_inject: int = 40) -> ServingWorkload:
    """Library: n_templates unit-energy Hann-tapered 1-10 Hz bursts with
    lengths U[lt].  Streams: n_channels one-hour float32 band-passed noise
    channels (DC offset U(-100, 100)), each with n_inject library templates
    injected at amplitude U(2, 8) x noise RMS x sqrt(Lt).  Requests cut
    windows of 60 s - 1 h (log-uniform) on 1-48 channels and match 1-100
    library templates."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lt[0], lt[1] + 1, size=n_templates)
    library = [template_burst(np.random.default_rng([seed, 2, i]), int(L), fs, 1.0, 10.0)
               for i, L in enumerate(lens)]
    streams = []
    for c in range(n_channels):
        x = bandpassed_noise(np.random.default_rng([seed, 1, c]), stream_len, fs, 1.0, 10.0)
        rms = float(np.sqrt(np.mean(x ** 2)))
        for _ in range(n_inject):
            t = library[int(rng.integers(0, n_templates))]
            p = int(rng.integers(0, stream_len - len(t)))
            x[p:p + len(t)] += rng.uniform(2, 8) * rms * math.sqrt(len(t)) * t
        streams.append((x + rng.uniform(-100, 100)).astype(np.float32))
    return ServingWorkload(fs=fs, library=library, streams=streams, min_window=min_window,
                           max_window=max_window, max_channels=max_channels,
                           max_templates=max_templates)


@dataclass
class UnevenBatch:
    """Templates of uneven length over channels of uneven duration."""
    fs: float
    signals: list                 # C float32 channels of different lengths
    templates: list               # [template] -> {channel: FP64 waveform}
    lengths: np.ndarray           # template lengths
    moveouts: list                # [template] -> {channel: shift}
    origins: list                 # [template] -> injection origins

    def rows(self):
        """(template, channel) pairs in template order."""
        return [(t, c) for t in range(len(self.templates)) for c in sorted(self.templates[t])]


def uneven_batch(seed: int = SEED_BASE + 4, n_templates: int = 200, n_channels: int = 50,
                 lt=(200, 3000), hours=(1.0, 24.0), fs: float = 100.0, max_moveout: int = 500,
                 max_inject: int = 30, snr=(2.0, 8.0)) -> UnevenBatch:
    """200 templates with lengths U[lt], each on 1..n_channels channels
    (count uniform, channels a random subset); n_channels float32 channels
    of band-passed (1-10 Hz) noise lasting U[hours] at fs; moveouts
    U[0, max_moveout]; each template injected 0..max_inject times inside the
    span all its channels cover, at per-sample amplitude U[snr] x noise RMS."""
    rng = np.random.default_rng(seed)
    ns = (rng.uniform(hours[0], hours[1], n_channels) * 3600 * fs).astype(np.int64)
    signals = []
    for c in range(n_channels):
        x = bandpassed_noise(np.random.default_rng([seed, 1, c]), int(ns[c]), fs, 1.0, 10.0,
                             dtype=np.float32)
        signals.append(x)
    rms = np.array([float(np.sqrt(np.mean(s[:200000].astype(np.float64) ** 2))) for s in signals])
    lengths = rng.integers(lt[0], lt[1] + 1, n_templates)
    templates, moveouts, origins = [], [], []
    for t in range(n_templates):
        k = int(rng.integers(1, n_channels + 1))
        chans = np.sort(rng.choice(n_channels, k, replace=False))
        L = int(lengths[t])
        templates.append({int(c): template_burst(np.random.default_rng([seed, 2, t, int(c)]), L, fs, 1.0, 10.0)
                          for c in chans})
        moveouts.append({int(c): int(rng.integers(0, max_moveout + 1)) for c in chans})
        span = int(min(ns[c] for c in chans)) - L - max_moveout - 1
        n_inj = int(rng.integers(0, max_inject + 1)) if span > 0 else 0
        orig = np.sort(rng.integers(0, span, n_inj)) if n_inj else np.zeros(0, dtype=np.int64)
        origins.append(orig)
        for o in orig:
            for c in chans:
                w = templates[t][int(c)] / np.abs(templates[t][int(c)]).max()
                i0 = int(o) + moveouts[t][int(c)]
                signals[c][i0:i0 + L] += (rng.uniform(snr[0], snr[1]) * rms[c] * w).astype(np.float32)
    for c in range(n_channels):
        signals[c] += np.float32(rng.uniform(-100, 100))
    return UnevenBatch(fs=fs, signals=signals, templates=templates, lengths=lengths,
                       moveouts=moveouts, origins=origins)


@dataclass
class PrecisionStress:
    """Short templates against channels with a large DC offset and gaps."""
    fs: float
    signals: list          # C float32 channels
    templates: list        # FP64 unit-energy templates
    gaps: list             # per channel: [(start, length)]
    injections: list       # per channel: [(template, position)]


def precision_stress(seed: int = SEED_BASE + 3, n_channels: int = 4, n: int = 2_000_000,
                     n_templates: int = 256, lt=(128, 256), fs: float = 100.0, dc: float = 1e4,
                     n_gaps: int = 20, gap=(500, 5000), n_inject: int = 30) -> PrecisionStress:
    """n_channels float32 channels of 2-20 Hz band-passed N(0,1) noise plus
    a DC offset of 1e4, each with n_gaps gap segments of 500-5000 samples
    holding a constant (DC + U(-50, 50)) plus 1e-6 N(0,1); n_templates
    unit-energy Hann-tapered 2-20 Hz bursts of U[lt] samples; each channel
    carries n_inject injected templates at 5x the noise RMS per sample."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lt[0], lt[1] + 1, n_templates)
    templates = [template_burst(np.random.default_rng([seed, 2, i]), int(L), fs, 2.0, 20.0)
                 for i, L in enumerate(lens)]
    signals, gaps, injections = [], [], []
    for c in range(n_channels):
        x = bandpassed_noise(np.random.default_rng([seed, 1, c]), n, fs, 2.0, 20.0)
        rms = float(np.sqrt(np.mean(x[:200000] ** 2)))
        inj = []
        for _ in range(n_inject):
            t = int(rng.integers(0, n_templates))
            p = int(rng.integers(0, n - len(templates[t])))
            w = templates[t] / np.abs(templates[t]).max()
            x[p:p + len(w)] += 5.0 * rms * w
            inj.append((t, p))
        x += dc
        g = []
        for _ in range(n_gaps):
            L = int(rng.integers(gap[0], gap[1] + 1))
            s0 = int(rng.integers(0, n - L))
            x[s0:s0 + L] = dc + rng.uniform(-50, 50) + 1e-6 * rng.standard_normal(L)
            g.append((s0, L))
        signals.append(x.astype(np.float32))
        gaps.append(g)
        injections.append(inj)
    return PrecisionStress(fs=fs, signals=signals, templates=templates, gaps=gaps, injections=injections)
