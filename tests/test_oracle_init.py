"""tcnn parameter initialisation: the oracle's pcg32 / seed_seq / Xavier /
generate_random_uniform restatement pinned by the published PCG32 known-answer
vector, libstdc++'s std::seed_seq, brute-force simulation of the uniform kernel's
thread layout and exactness of its one-rounding transform; then libnlod's
nlod_model_init_params (host code, no GPU) against it bit for bit."""
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from neural_lod_amd import configs
from oracle import model as M


def test_pcg32_known_answer():
    """pcg32_srandom_r(42, 54), first six outputs of the PCG reference demo
    (pcg-c-basic, pcg32-demo.c)."""
    r = M.Pcg32(42, 54)
    assert [r.next_uint() for _ in range(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B,
                                                 0xCBED606E]


@pytest.mark.parametrize("delta", [0, 1, 2, 3, 127, 128, 1000, 4097])
def test_pcg32_advance_and_vector_equal_stepping(delta):
    a, b, c = M.Pcg32(1337, 1), M.Pcg32(1337, 1), M.Pcg32(1337, 1)
    stepped = [a.next_uint() for _ in range(delta)]
    b.advance(delta)
    assert b.state == a.state
    assert [int(v) for v in c.next_uints(delta)] == stepped
    assert c.state == a.state


def test_pcg32_float_is_bit_construction():
    """next_float is the float with bits (u >> 9) | 0x3f800000, minus 1."""
    r, s = M.Pcg32(7), M.Pcg32(7)
    for _ in range(2000):
        u = r.next_uint()
        ref = np.array([(u >> 9) | 0x3F800000], np.uint32).view(np.float32)[0] - np.float32(1.0)
        assert s.next_float() == ref


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_seed_seq_matches_libstdcxx(tmp_path):
    src = tmp_path / "ss.cpp"
    src.write_text(
        "#include <random>\n#include <cstdio>\n#include <vector>\n#include <cstdlib>\n"
        "int main(int c, char** v) { for (int i = 1; i < c; ++i) { std::seed_seq q{(unsigned)strtoul(v[i], 0, 10)};"
        " for (int n : {1, 2, 5, 8, 40, 70}) { std::vector<unsigned> o(n); q.generate(o.begin(), o.end());"
        " for (unsigned x : o) printf(\"%u \", x); } printf(\"\\n\"); } }\n")
    exe = tmp_path / "ss"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    seeds = [0, 1, 42, 1337, 0xFFFFFFFF]
    lines = subprocess.run([str(exe)] + [str(s) for s in seeds], check=True, capture_output=True,
                           text=True).stdout.splitlines()
    for s, line in zip(seeds, lines):
        ours = [w for n in (1, 2, 5, 8, 40, 70) for w in M.seed_seq_generate([s], n)]
        assert [int(w) for w in line.split()] == ours


def _uniform_kernel_brute(rng, n, lo, hi):
    """Thread-by-thread simulation of generate_random_kernel<float, pcg32, 4>
    launched with n_blocks_linear(ceil(n/4)) blocks of 128 threads."""
    n_thr = -(-(-(-n // 4)) // 128) * 128
    out = [None] * n
    span = Fraction(float(np.float32(np.float32(hi) - np.float32(lo))))
    for i in range(n_thr):
        r = M.Pcg32.__new__(M.Pcg32)
        r.state, r.inc = rng.state, rng.inc
        for _ in range(4 * i):
            r.next_uint()
        for j in range(4):
            idx = i + n_thr * j
            if idx >= n:
                break
            v = Fraction(int(r.next_uint() >> 9), 1 << 23)
            out[idx] = np.float32(float(v * span + Fraction(float(np.float32(lo)))))
    return np.array(out, np.float32)


@pytest.mark.parametrize("n", [1, 5, 511, 512, 513, 1000, 2051])
def test_uniform_table_thread_layout(n):
    rng, ref_rng = M.Pcg32(99), M.Pcg32(99)
    got = M.uniform_table(rng, n, -1e-4, 1e-4)
    assert np.array_equal(got, _uniform_kernel_brute(ref_rng, n, -1e-4, 1e-4))
    ref_rng.advance(n)
    assert rng.state == ref_rng.state


def test_uniform_transform_single_rounding():
    """v * span + lo is exact in float64 for every 23-bit v: the float32 result
    is the once-rounded (FMA) value."""
    span = np.float64(np.float32(np.float32(1e-4) - np.float32(-1e-4)))
    lo = np.float64(np.float32(-1e-4))
    for k in [0, 1, 2, 3, 12345, (1 << 22) - 1, 1 << 22, (1 << 23) - 2, (1 << 23) - 1]:
        v = np.float64(k) * 2.0 ** -23
        assert Fraction(float(v * span + lo)) == Fraction(float(v)) * Fraction(float(span)) + Fraction(float(lo))


@pytest.mark.parametrize("name", ["visibility", "reference_throughput"])
def test_library_init_matches_oracle(name):
    from neural_lod_amd.tcnn import NetworkWithInputEncoding
    cfg = configs.ALL[name]
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    m = NetworkWithInputEncoding(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"], 0)
    for seed in (1337, 42):
        ref = M.tcnn_initial_params(L, seed)
        got = m.initial_params(seed)
        bad = np.flatnonzero(got.view(np.uint32) != ref.view(np.uint32))
        assert bad.size == 0, f"{name} seed {seed}: {bad.size} params differ, first at {bad[:4]}"


def test_visibility_grid_needs_thread_rounding():
    """The visibility grid's ceil(n/4) is not a multiple of 128, so its table
    tells the rounded thread count from an unrounded one."""
    cfg = configs.VISIBILITY
    L = M.Layout(cfg["n_input_dims"], cfg["n_output_dims"], cfg["encoding"], cfg["network"])
    i = next(k for k, s in enumerate(L.segs) if s.type == M.SEG_GRID)
    offs, _, _ = L.grid_levels(i)
    n = int(offs[-1]) * L.segs[i].n_features
    assert ((n + 3) // 4) % 128 != 0
