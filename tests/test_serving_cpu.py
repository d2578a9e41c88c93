"""CPU checks of the serving loop's host logic: the seeded Poisson stream
and the footprint used for HBM admission (pmx_engine_footprint needs no
device)."""
import numpy as np
import pytest

from patchmatch_cuda_amd import capi, configs, serving


def test_stream_is_seeded_poisson():
    a = serving.make_stream(4.0, 600.0, seed=11)
    b = serving.make_stream(4.0, 600.0, seed=11)
    assert [(r.arrival, r.size, r.n_src) for r in a] == [(r.arrival, r.size, r.n_src) for r in b]
    n = len(a)
    # 2400 expected; Poisson sd ~ 49
    assert abs(n - 2400) < 250
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 0.25) < 0.02
    # exponential gaps: sd == mean
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1
    assert all(4 <= r.n_src <= 20 for r in a)
    assert all(0 <= r.size < len(serving.SIZES) for r in a)
    assert np.all(np.diff([r.arrival for r in a]) > 0)
    counts = np.bincount([r.size for r in a], minlength=len(serving.SIZES))
    assert counts.min() > 0.15 * n


def test_footprint_closed_form():
    _, rp = configs.C2_REALISTIC
    W, H, V = 2048, 1536, 10
    f32 = capi.engine_footprint(rp, W, H, [(W, H)] * V)
    f16 = capi.engine_footprint(rp, W, H, [(W, H)] * V, capi.STORE_F16)
    fp = capi.engine_footprint(rp, W, H, [(W, H)] * V, capi.STORE_F16, proposal=True)
    pitch = (W + 4 + 63) // 64 * 64
    layer = pitch * (H + 4)
    # FP16 storage holds 2 bytes per padded texel per view in place of the
    # 8-byte FP32 texel pairs
    g = lambda b: (b + 255) // 256 * 256  # noqa: E731
    assert f32 - f16 == g(8 * layer * V) - g(2 * layer * V)
    # proposal planes: depth and spread, row- and column-major; the
    # proposal's network input (64 FP16 channels at quarter resolution) and
    # its two full-resolution output planes
    ld_r, ld_c = (W + 63) // 64 * 64, (H + 63) // 64 * 64
    W0, H0 = (W + 3) // 4, (H + 3) // 4
    assert fp - f16 == 2 * g(4 * ld_r * H) + 2 * g(4 * ld_c * W) + g(2 * W0 * H0 * 64) + g(8 * W * H)
    # dominant terms: two FP32 source blocks and the texel pairs, the
    # messages and the output costs, two cost layouts
    dominant = 2 * 4 * layer * V + 8 * layer * V + 8 * W * H * V + 2 * 4 * W * H * V
    assert dominant < f32 < dominant * 1.5
    # more views, more bytes
    assert capi.engine_footprint(rp, W, H, [(W, H)] * (V + 1)) > f32
    with pytest.raises(capi.PmxError):
        capi.engine_footprint(rp, 0, H, [(W, H)] * V)
