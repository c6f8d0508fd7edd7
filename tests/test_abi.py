"""The C ABI library loads without a GPU and exports every function that
include/kz.h declares; the Python binding covers the same set."""

import re
from pathlib import Path

import openkhuzdulgraph_amd as kz
from openkhuzdulgraph_amd import _ffi

HEADER = Path(__file__).resolve().parent.parent / "include" / "kz.h"


def declared_functions(path=HEADER):
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+[\s\*]+\**(kz_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = declared_functions()
    assert "kz_count_cliques" in names and "kz_dgraph_load" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_ffi.exported_symbols()) == declared_functions()


def test_no_device_errors_cleanly():
    import torch

    if torch.cuda.is_available():
        return
    try:
        kz.init(0)
    except kz.KzError as e:
        assert e.code in (_ffi.KZ_ERR_DEVICE, _ffi.KZ_ERR_INVALID)
    else:  # pragma: no cover - only on a GPU host
        kz.shutdown()


def test_calls_before_init_fail_with_state_error():
    g = kz.gen_erdos_renyi(50, 100, 1)
    try:
        g.to_device()
    except kz.KzError as e:
        assert e.code in (_ffi.KZ_ERR_STATE, _ffi.KZ_ERR_DEVICE)
