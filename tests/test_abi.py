"""CPU checks of the C-ABI library: it loads, exports every entry point include/grafeo_hip.h
declares, carries code objects built from the current kernel sources, and reports the
reference's error categories. No device work (these run without a GPU)."""

from __future__ import annotations

import os
import re
import subprocess

import pytest

import code_objects
from grafeo_cuda import _native

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "grafeo_hip.h")


def declared_functions() -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return set(re.findall(r"\b(gc_[a-z0-9_]+)\s*\(", text))


def exported_symbols() -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_entry_point():
    declared = declared_functions()
    assert declared, "no declarations parsed"
    missing = declared - exported_symbols()
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert set(_native.EXPORTED) == declared


def test_library_exports_nothing_but_the_abi():
    extra = {s for s in exported_symbols() if not s.startswith("gc_")}
    assert not extra, f"unexpected exports: {sorted(extra)[:10]}"


def test_embedded_code_objects_are_built_from_current_sources():
    """Counterpart of embedded_ptx_is_the_committed_ptx (crates/grafeo-cuda-algos/src/lib.rs:28-63)."""
    embedded = {c["name"]: c for c in _native.code_objects()}
    sources = dict(code_objects.kernel_sources())
    assert set(embedded) == set(sources)
    for name, path in sources.items():
        assert embedded[name]["source_sha256"] == code_objects.source_digest(path), name
        assert embedded[name]["size"] > 0
        assert len(embedded[name]["object_sha256"]) == 64


def test_code_object_manifest_matches_sources():
    """Counterpart of `cargo xtask check-ptx` (xtask/src/main.rs:47-67)."""
    assert code_objects.verify(code_objects.hipcc_version()) == []


def test_manifest_parse_round_trip_and_drift_detection(tmp_path, monkeypatch):
    text = code_objects.render("1.2.3", [("a", "00" * 32), ("b", "11" * 32)])
    parsed = code_objects.parse(text)
    assert parsed["hipcc"] == "1.2.3" and parsed["target"] == "gfx950"
    assert parsed["kernels"]["b"]["source_sha256"] == "11" * 32
    with pytest.raises(ValueError):
        code_objects.parse("hipcc = 1.2.3")
    manifest = tmp_path / "code_objects.toml"
    manifest.write_text(text)
    monkeypatch.setattr(code_objects, "MANIFEST", manifest)
    problems = code_objects.verify("1.2.3")
    assert any("source changed" in p or "not in the manifest" in p or "gone" in p for p in problems)


def test_hash_is_line_ending_insensitive():
    assert code_objects.sha256_hex(b"a\r\nb\r\n") == code_objects.sha256_hex(b"a\nb\n")
    assert code_objects.sha256_hex(b"") == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"


@pytest.mark.skipif(_native.device_count() > 0, reason="a GPU is present")
def test_no_device_is_reported_as_no_device():
    import grafeo_cuda

    with pytest.raises(RuntimeError, match="no GPU device found") as info:
        grafeo_cuda.GpuDevice(0)
    assert info.value.kind == "NoDevice"


def test_argument_type_errors():
    import grafeo_cuda

    with pytest.raises(OverflowError):
        _native.as_u32([-1], "offsets")
    with pytest.raises(OverflowError):
        _native.as_u32([1 << 32], "offsets")
    with pytest.raises(TypeError):
        _native.as_u32([1.5], "offsets")
    with pytest.raises(TypeError):
        grafeo_cuda.GpuGraph.from_csr([0], [], device=None)
