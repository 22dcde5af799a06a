"""Code-object manifest: which HIP sources the gfx950 code objects come from.

Replaces the reference's PTX bookkeeping (crates/grafeo-cuda-algos/build_support/ptx_common.rs,
crates/grafeo-cuda-algos/build.rs:19-78 and `cargo xtask check-ptx` / `regenerate-ptx`,
xtask/src/main.rs:47-140). The reference commits PTX and a manifest of SHA-256 hashes;
here the code objects are built from source every time (they are never committed), so the
committed manifest `code_objects.toml` pins what a build must start from:

  - the hipcc version, the target (gfx950) and the compile flags,
  - per kernel family (one `csrc/kernels/<name>.hip` file): the SHA-256 of its source
    together with the shared device headers it includes.

`python code_objects.py check` fails when a kernel source, the flags or the toolchain drifted
from the manifest without `regenerate` being run (the counterpart of `check-ptx`);
`python code_objects.py regenerate` rewrites the manifest. Hashes are taken after CRLF -> LF
normalization, as the reference does, so checkouts on any platform hash the same. The built
library carries the same source hashes plus the SHA-256 of each embedded code object
(gc_code_object_info), so a test can tell whether a library was built from these sources.
"""

from __future__ import annotations

import hashlib
import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
KERNEL_DIR = HERE / "csrc" / "kernels"
MANIFEST = HERE / "code_objects.toml"
TARGET = "gfx950"
HIPCC = "/opt/rocm/bin/hipcc"
# Packed FP32 instructions (v_pk_add/mul/fma_f32) are disabled for every kernel: some of their
# forms have given wrong results on this hardware without any error.
HIPCC_FLAGS = [
    "--genco",
    f"--offload-arch={TARGET}",
    "-O3",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def normalize_newlines(data: bytes) -> bytes:
    return data.replace(b"\r\n", b"\n")


def sha256_hex(data: bytes) -> str:
    return hashlib.sha256(normalize_newlines(data)).hexdigest()


def kernel_sources() -> list[tuple[str, Path]]:
    return sorted((p.stem, p) for p in KERNEL_DIR.glob("*.hip"))


def shared_headers() -> list[Path]:
    return sorted(KERNEL_DIR.glob("*.hpp"))


def source_digest(source: Path) -> str:
    """SHA-256 over the kernel source followed by every shared device header."""
    h = hashlib.sha256()
    h.update(normalize_newlines(source.read_bytes()))
    for header in shared_headers():
        h.update(b"\0" + header.name.encode() + b"\0")
        h.update(normalize_newlines(header.read_bytes()))
    return h.hexdigest()


def hipcc_version() -> str:
    out = subprocess.run([HIPCC, "--version"], capture_output=True, text=True, check=True).stdout
    m = re.search(r"HIP version:\s*(\S+)", out)
    return m.group(1) if m else "unknown"


def render(version: str, entries: list[tuple[str, str]]) -> str:
    lines = [
        "# Generated by `python code_objects.py regenerate`. Do not edit by hand.",
        "# Hashes are SHA-256 of the kernel source plus the shared device headers,",
        "# with CRLF line endings normalized to LF.",
        f'hipcc = "{version}"',
        f'target = "{TARGET}"',
        f'flags = "{" ".join(HIPCC_FLAGS)}"',
        "",
    ]
    for name, digest in entries:
        lines += [f"[kernels.{name}]", f'source_sha256 = "{digest}"', ""]
    return "\n".join(lines)


def parse(text: str) -> dict:
    """Parses the manifest format written by `render` (a TOML subset)."""
    top: dict = {}
    kernels: dict[str, dict[str, str]] = {}
    section = None
    for raw in text.splitlines():
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        m = re.fullmatch(r"\[kernels\.([A-Za-z0-9_]+)\]", line)
        if m:
            section = kernels.setdefault(m.group(1), {})
            continue
        m = re.fullmatch(r'([A-Za-z0-9_]+)\s*=\s*"([^"]*)"', line)
        if not m:
            raise ValueError(f"malformed manifest line: {raw!r}")
        (section if section is not None else top)[m.group(1)] = m.group(2)
    top["kernels"] = kernels
    return top


def verify(version: str | None = None) -> list[str]:
    """Returns the list of problems (empty when the manifest matches)."""
    if not MANIFEST.exists():
        return [f"{MANIFEST.name} is missing"]
    manifest = parse(MANIFEST.read_text())
    problems = []
    if manifest.get("target") != TARGET:
        problems.append(f"manifest target {manifest.get('target')!r} is not {TARGET}")
    if manifest.get("flags") != " ".join(HIPCC_FLAGS):
        problems.append("compile flags differ from the manifest")
    if version is not None and manifest.get("hipcc") != version:
        problems.append(f"hipcc {version} differs from the manifest's {manifest.get('hipcc')}")
    recorded = manifest["kernels"]
    current = {name: source_digest(path) for name, path in kernel_sources()}
    for name in sorted(set(recorded) | set(current)):
        if name not in current:
            problems.append(f"{name}: in the manifest but kernels/{name}.hip is gone")
        elif name not in recorded:
            problems.append(f"{name}: kernels/{name}.hip is not in the manifest")
        elif recorded[name].get("source_sha256") != current[name]:
            problems.append(f"{name}: source changed since the manifest was generated")
    return problems


def regenerate() -> None:
    entries = [(name, source_digest(path)) for name, path in kernel_sources()]
    MANIFEST.write_text(render(hipcc_version(), entries))


def main(argv: list[str]) -> int:
    if len(argv) != 2 or argv[1] not in ("check", "regenerate"):
        print("usage: python code_objects.py <check | regenerate>", file=sys.stderr)
        return 2
    if argv[1] == "regenerate":
        regenerate()
        print(f"wrote {MANIFEST.name}")
    problems = verify(hipcc_version())
    if problems:
        print("the code-object manifest does not match the kernel sources:", file=sys.stderr)
        for p in problems:
            print(f"  - {p}", file=sys.stderr)
        print("run `python code_objects.py regenerate` and commit code_objects.toml", file=sys.stderr)
        return 1
    print(f"code-object manifest matches {len(kernel_sources())} kernel sources ({TARGET})")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
