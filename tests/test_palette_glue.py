"""The palette parameterisation of the finetune model (gsplat/gsmodel.py:33-49
hsl_to_rgb, :203-212 get_palette_rgb), restated in rgp.gsmodel, pinned by
closed forms and by Python's own colorsys.hls_to_rgb (the textbook HSL to RGB
conversion), in float64 on the CPU."""
import colorsys
import math

import numpy as np
import pytest
import torch

from rgp.gsmodel import get_palette_rgb, hsl_to_rgb

D = torch.float64


def test_six_sector_hues_are_the_primaries_and_secondaries():
    h = torch.tensor([k * math.pi / 3 for k in range(6)], dtype=D)
    rgb = hsl_to_rgb(h, torch.ones(6, dtype=D), torch.full((6,), 0.5, dtype=D))
    expect = torch.tensor([[1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1], [1, 0, 1]], dtype=D)
    assert torch.allclose(rgb, expect, atol=1e-12, rtol=0)


def test_zero_saturation_is_grey_and_lightness_ends_are_black_and_white():
    rng = np.random.default_rng(0)
    h = torch.tensor(rng.uniform(0, 2 * math.pi, 50), dtype=D)
    l = torch.tensor(rng.uniform(0, 1, 50), dtype=D)
    grey = hsl_to_rgb(h, torch.zeros(50, dtype=D), l)
    assert torch.allclose(grey, l[:, None].expand(50, 3), atol=1e-12, rtol=0)
    s = torch.tensor(rng.uniform(0, 1, 50), dtype=D)
    assert torch.allclose(hsl_to_rgb(h, s, torch.zeros(50, dtype=D)), torch.zeros(50, 3, dtype=D), atol=1e-12)
    assert torch.allclose(hsl_to_rgb(h, s, torch.ones(50, dtype=D)), torch.ones(50, 3, dtype=D), atol=1e-12)


def test_hsl_to_rgb_matches_colorsys():
    rng = np.random.default_rng(1)
    n = 4000
    h = rng.uniform(0, 2 * math.pi, n)
    h[:12] = [k * math.pi / 3 for k in range(6)] + [2 * math.pi - 1e-9, 1e-9, math.pi, 5.0, 6.0, 0.1]
    s, l = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    got = hsl_to_rgb(torch.tensor(h, dtype=D), torch.tensor(s, dtype=D), torch.tensor(l, dtype=D)).numpy()
    ref = np.array([colorsys.hls_to_rgb(hh / (2 * math.pi), ll, ss) for hh, ss, ll in zip(h, s, l)])
    assert np.abs(got - ref).max() < 1e-12


def test_palette_equal_gaps_give_equally_spaced_hues():
    """Equal gap parameters split the hue circle evenly from hue_base; at full
    saturation and half lightness three colours from hue 0 are red, green,
    blue, and six from hue 0 are the six sector colours."""
    big = torch.full((3,), 40.0, dtype=D)  # sigmoid(40) = 1 in float64
    rgb = get_palette_rgb(torch.tensor([0.0], dtype=D), torch.zeros(3, dtype=D), big, torch.zeros(3, dtype=D))
    assert torch.allclose(rgb, torch.eye(3, dtype=D), atol=1e-12, rtol=0)
    rgb6 = get_palette_rgb(torch.tensor([4 * math.pi], dtype=D), torch.full((6,), -1.3, dtype=D),
                           torch.full((6,), 40.0, dtype=D), torch.zeros(6, dtype=D))
    h = torch.tensor([k * math.pi / 3 for k in range(6)], dtype=D)
    assert torch.allclose(rgb6, hsl_to_rgb(h, torch.ones(6, dtype=D), torch.full((6,), 0.5, dtype=D)), atol=1e-12)


@pytest.mark.parametrize("seed", [2, 3])
def test_palette_matches_independent_formula(seed):
    """hues = (base mod 2pi + cumsum of [0, gaps[:-1]]) mod 2pi with gaps =
    2pi softplus(g) / sum softplus(g), s = sigmoid, l = sigmoid, then the
    textbook conversion: numpy and colorsys only."""
    rng = np.random.default_rng(seed)
    k = 5
    base, gap, sat, lit = rng.normal(0, 3, 1), rng.normal(0, 1, k), rng.normal(0, 1, k), rng.normal(0, 1, k)
    sp = np.logaddexp(0.0, gap)
    gaps = 2 * math.pi * sp / sp.sum()
    hues = (np.mod(base[0], 2 * math.pi) + np.concatenate([[0.0], np.cumsum(gaps[:-1])])) % (2 * math.pi)
    sig = lambda x: 1 / (1 + np.exp(-x))
    ref = np.array([colorsys.hls_to_rgb(hh / (2 * math.pi), ll, ss) for hh, ss, ll in zip(hues, sig(sat), sig(lit))])
    got = get_palette_rgb(*(torch.tensor(a, dtype=D) for a in (base, gap, sat, lit))).numpy()
    assert np.abs(got - ref).max() < 1e-12
