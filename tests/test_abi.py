"""The C ABI: libnlod.so loads without a GPU and exports every function that
include/nlod.h declares; the ctypes table covers exactly those functions.
No compute call is made here."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared() -> set[str]:
    text = (ROOT / "include" / "nlod.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # comments
    return set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nlod_\w+)\s*\(", text, flags=re.M))


def test_header_declares_the_surface():
    names = _declared()
    for must in ("nlod_voxelize", "nlod_model_inference", "nlod_model_backward", "nlod_trainer_step",
                 "nlod_wavefront_run_batches", "nlod_labeler_visibility", "nlod_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from neural_lod_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libnlod.so is not built")
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, f"declared in nlod.h but not exported: {missing}"


def test_ctypes_table_matches_header():
    from neural_lod_amd import _lib
    assert set(_lib.SIGNATURES) == _declared()


def test_version_and_error_without_gpu():
    from neural_lod_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libnlod.so is not built")
    L = _lib.lib()
    assert L.nlod_version()
    assert L.nlod_last_error() is not None
