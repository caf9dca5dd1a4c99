"""Encoding/network layout: libnlod's C++ JSON front end vs the oracle's
restatement of tiny-cuda-nn's composite/padding rules, plus hand-derived
layout tables for the path's configurations (no GPU needed)."""
import json

import numpy as np
import pytest

from neural_lod_amd import configs
from oracle import model as M


def _lib_layout(cfg, precision=0):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], precision)
    return m, m.describe()


def _compare(cfg):
    m, d = _lib_layout(cfg)
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    assert d["enc_width"] == L.enc_width
    assert d["n_params"] == L.n_params == m.n_params
    assert d["n_net_params"] == L.n_net_params
    assert len(d["segments"]) == len(L.segs)
    for i, (a, b) in enumerate(zip(d["segments"], L.segs)):
        for k in ("type", "in_offset", "out_offset", "n_out", "n_pad", "param_offset"):
            if k == "param_offset" and a["type"] != 0:
                continue
            assert a[k] == getattr(b, k), f"segment {i} field {k}: {a[k]} != {getattr(b, k)}"
        if a["type"] == 0:
            offs, _, res = L.grid_levels(i)
            assert a["levels"] == offs.tolist()
            assert a["res"] == res.tolist()
    return d, L


@pytest.mark.parametrize("name", sorted(configs.ALL))
def test_layout_matches_oracle(name):
    _compare(configs.ALL[name])


def test_visibility_layout_table():
    d, L = _compare(configs.VISIBILITY)
    # hash 16 levels x 2 = 32 rows, SH degree 4 = 16 rows, 48 total (multiple of 16: no pad)
    assert d["enc_width"] == 48
    segs = d["segments"]
    assert [(s["out_offset"], s["n_out"], s["n_pad"]) for s in segs] == [(0, 32, 0), (32, 16, 0)]
    assert d["n_net_params"] == 64 * 48 + 3 * 64 * 64 + 16 * 64
    # levels 0..4 are exact multiples 16 * 1.5^l; the rest are hashed (2^19 entries)
    res = segs[0]["res"]
    assert res[:5] == [16, 24, 36, 54, 81]
    assert res[5] == 122
    offs = np.array(segs[0]["levels"])
    sizes = np.diff(offs)
    dense = [16 ** 3, 24 ** 3, 36 ** 3, 54 ** 3]
    assert sizes[:4].tolist() == [(v + 7) // 8 * 8 for v in dense]
    assert (sizes[4:] == 2 ** 19).all()


def test_throughput_layout_table():
    d, _ = _compare(configs.THROUGHPUT)
    # 32 + 16 + 16 + 1 = 65 -> 80; the 15 pad rows (1.0) belong to the last (identity) encoding
    assert d["enc_width"] == 80
    segs = d["segments"]
    assert [(s["type"], s["in_offset"], s["out_offset"], s["n_out"], s["n_pad"]) for s in segs] == [
        (0, 0, 0, 32, 0), (1, 3, 32, 16, 0), (1, 6, 48, 16, 0), (2, 9, 64, 1, 15)]


def test_reference_visibility_repeated_composite():
    d, _ = _compare(configs.REFERENCE_VISIBILITY)
    # one shared 7x2 grid applied to p_in and p_out: 28 -> 32, 2 zero rows per repetition
    assert d["enc_width"] == 32
    segs = d["segments"]
    assert [(s["in_offset"], s["out_offset"], s["n_out"], s["n_pad"]) for s in segs] == [(0, 0, 14, 2), (3, 16, 14, 2)]
    assert segs[0]["param_offset"] == segs[1]["param_offset"]
    # 8^3, 16^3, 32^3, 64^3 dense; 3 hashed levels of 2^19
    assert segs[0]["n_params"] == 2 * (512 + 4096 + 32768 + 262144 + 3 * 2 ** 19)


def test_reference_throughput_config_fixture(golden_dir):
    cfg = json.loads((golden_dir / "reference_throughput_config.json").read_text())
    d, L = _compare({"n_input_dims": 9, "n_output_dims": 3, "encoding": cfg["encoding"], "network": cfg["network"]})
    # 14 + 64 + 64 = 142 -> 144: the last SH gets 2 pad rows of 1.0 placed before its values
    assert d["enc_width"] == 144
    assert [(s["out_offset"], s["n_out"], s["n_pad"]) for s in d["segments"]] == [(0, 14, 0), (14, 64, 0), (78, 64, 2)]
    assert d["n_net_params"] == 128 * 144 + 3 * 128 * 128 + 16 * 128


@pytest.mark.parametrize("grid_type", ["Dense", "Tiled"])
def test_dense_and_tiled_grids(grid_type):
    cfg = {"n_input_dims": 3, "n_output_dims": 1,
           "encoding": {"otype": "Grid", "type": grid_type, "n_levels": 6, "n_features_per_level": 4,
                        "log2_hashmap_size": 14, "base_resolution": 4, "per_level_scale": 1.7},
           "network": {"otype": "FullyFusedMLP", "n_neurons": 32, "n_hidden_layers": 2}}
    d, _ = _compare(cfg)
    assert d["enc_width"] == 32  # 24 rows -> 32 (8 zero rows)


@pytest.mark.parametrize("bad, msg", [
    ({"otype": "HashGrid", "n_features_per_level": 3}, "n_features_per_level"),
    ({"otype": "SphericalHarmonics", "degree": 9}, "degree 8"),
    ({"otype": "OneBlob"}, "not supported"),
    ({"otype": "Composite"}, "nested"),
])
def test_config_errors(bad, msg):
    from neural_lod_amd._lib import NlodError
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    with pytest.raises(NlodError, match=msg):
        NetworkWithInputEncoding(3, 1, bad, {"otype": "FullyFusedMLP", "n_neurons": 64})


def test_network_errors():
    from neural_lod_amd._lib import NlodError
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    enc = {"otype": "Identity"}
    with pytest.raises(NlodError, match="16, 32, 64, and 128"):
        NetworkWithInputEncoding(3, 1, enc, {"otype": "FullyFusedMLP", "n_neurons": 48})
    with pytest.raises(NlodError, match="at least 1 hidden layer"):
        NetworkWithInputEncoding(3, 1, enc, {"otype": "FullyFusedMLP", "n_neurons": 64, "n_hidden_layers": 0})
    with pytest.raises(NlodError, match="malformed JSON"):
        NetworkWithInputEncoding(3, 1, "{not json", {"otype": "FullyFusedMLP"})


def test_init_params_tcnn_rules():
    """Deterministic in the seed; Xavier bounds per matrix; grid U[-1e-4, 1e-4]."""
    cfg = configs.VISIBILITY
    m, d = _lib_layout(cfg)
    p = m.initial_params(1337)
    assert np.array_equal(p, m.initial_params(1337))
    assert not np.array_equal(p, m.initial_params(7))
    W, K = 64, 48
    shapes = [(W, K), (W, W), (W, W), (W, W), (16, W)]
    off = 0
    for rows, cols in shapes:
        block = p[off: off + rows * cols]
        bound = np.sqrt(6.0 / (rows + cols))
        assert np.abs(block).max() <= bound
        assert np.abs(block).max() > 0.95 * bound
        assert abs(block.mean()) < 0.1 * bound
        off += rows * cols
    grid = p[off:]
    assert grid.size == m.n_params - m.n_network_params
    assert np.abs(grid).max() <= 1e-4 and np.abs(grid).max() > 0.99e-4
    assert abs(grid.mean()) < 1e-6
    # pcg32 floats are k/2^23 before scaling: distinct values, uniform spread
    assert np.unique(grid[:100000]).size > 99000
