"""Shipped weight files: bit-exact against the reference text, and the invariants the
reference's weights satisfy (which any transcription slip would break)."""
import json
import os
import tempfile

import numpy as np
import pytest

from conftest import PKG
from oracle.common import load_weights

WDIR = os.path.join(PKG, "weights")


def test_reextract_is_bit_exact(reference_dir):
    import importlib.util
    spec = importlib.util.spec_from_file_location("extract", os.path.join(WDIR, "extract.py"))
    ext = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ext)
    with tempfile.TemporaryDirectory() as d:
        assert ext.main(["--ref", reference_dir, "--out", d]) == 0
        man = json.load(open(os.path.join(WDIR, "manifest.json")))["weights"]
        for name in man:
            a = open(os.path.join(d, name + ".bin"), "rb").read()
            b = open(os.path.join(WDIR, name + ".bin"), "rb").read()
            assert a == b, name


def test_manifest_hashes():
    import hashlib
    man = json.load(open(os.path.join(WDIR, "manifest.json")))["weights"]
    assert len(man) == 22
    for name, m in man.items():
        arr = load_weights(name).astype("<f4")
        assert list(arr.shape) == m["shape"], name
        assert hashlib.sha256(arr.tobytes()).hexdigest() == m["data_sha256"], name


@pytest.mark.parametrize("nns", [16, 32, 64, 128, 256])
@pytest.mark.parametrize("wh", [4, 6])
def test_nnedi3_neurons_sum_to_zero(nns, wh):
    t = load_weights("nnedi3_nns%d_win8x%d" % (nns, wh))
    K = 8 * wh
    assert t.shape == (nns, 2 * K + 2)
    # every neuron's weights (both halves) sum to ~0: the network is invariant to the
    # window mean, which the FP16 contraction exploits by centring the window
    assert np.abs(t[:, :K].sum(1)).max() < 2e-6
    assert np.abs(t[:, K:2 * K].sum(1)).max() < 2e-6


@pytest.mark.parametrize("r", [2, 3, 4])
def test_lite_subpixel_weights_sum_to_one(r):
    lut = load_weights("ravu_lite_lut%d" % r)      # (288, ntex, 4)
    n = 2 * r - 1
    c = (n * n - 1) // 2
    # sum over taps of w_i[k] + w_i.wzyx[k] (mirrored) + centre = total weight of sub-pixel k
    tot = lut[:, :c, :].sum(1) + lut[:, :c, ::-1].sum(1) + lut[:, c, :]
    assert np.abs(tot - 1).max() < 2e-3


@pytest.mark.parametrize("r,base", [(2, 5), (3, 14)])
def test_zoom_phase_zero_is_unit_on_base_tap(r, base):
    lut = load_weights("ravu_zoom_lut%d" % r)      # (2592, 9*groups, 4)
    groups = lut.shape[1] // 9
    for b in range(288):
        w = np.array([lut[9 * b, 9 * g, c] for g in range(groups) for c in range(4)])
        e = np.zeros_like(w)
        e[base] = 1.0
        assert np.array_equal(w, e), b
