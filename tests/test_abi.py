"""CPU-side checks of the C-ABI library: it loads (no GPU needed to load) and
exports every function that the headers in include/ declare."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "for", "while")))


@pytest.mark.parametrize("header", ["gmpt_amd.h", "libgmpt.h"])
def test_library_exports_declared_symbols(gmpt, header):
    lib = gmpt.gmpt.lib()
    names = _declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_lists_match_headers(gmpt):
    assert sorted(gmpt.gmpt.GMPT_SYMBOLS) == _declared("gmpt_amd.h")
    assert sorted(gmpt.gmpt.LIBGMPT_SYMBOLS) == _declared("libgmpt.h")


def test_missing_library_fails_loudly(gmpt, tmp_path, monkeypatch):
    monkeypatch.setattr(gmpt.gmpt, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(gmpt.gmpt, "_lib", None)
    with pytest.raises(gmpt.GmptError):
        gmpt.gmpt.lib()


def test_batch_layout(workloads):
    b = workloads.batch_from_bytes([b"doe", b"dog"], [b"reindeer", b"puppy"])
    assert b.keys_indexs.tolist() == [0, 6, 7, 13]
    assert b.values_indexs.tolist() == [0, 7, 8, 12]
    assert b.key(1) == bytes([6, 4, 6, 15, 6, 7, 16])
