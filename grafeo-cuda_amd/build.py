"""Builds libgrafeo_hip.so: gfx950 code objects embedded in a C-ABI host library.

Steps
  1. hipcc --genco each csrc/kernels/<family>.hip into _build/<family>.hsaco (flags from
     code_objects.py, packed FP32 disabled).
  2. Generate _build/code_objects.S (.incbin of every code object) and
     _build/code_objects.cpp (the table the module cache loads from, with the source and
     object SHA-256 of each).
  3. g++ the host runtime and algorithm drivers (C++17, HIP runtime API only).
  4. Link grafeo_cuda/libgrafeo_hip.so against libamdhip64.

Usage: python build.py [--force] [--jobs N]. Without --force, unchanged objects are reused.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import code_objects as co  # noqa: E402

REPO = HERE.parent
BUILD = HERE / "_build"
OUT = HERE / "grafeo_cuda" / "libgrafeo_hip.so"
ROCM = Path("/opt/rocm")
HOST_SOURCES = sorted((HERE / "csrc" / "runtime").glob("*.cpp")) + sorted(
    (HERE / "csrc" / "algos").glob("*.cpp")
)
HOST_HEADERS = sorted((HERE / "csrc" / "runtime").glob("*.hpp")) + [REPO / "include" / "grafeo_hip.h"]
# Host sources built by hipcc: the static device scratch of the allocation-free entry point
# (a __device__ array whose address the HIP runtime resolves), with its gfx950 code object.
HIP_HOST_SOURCES = sorted((HERE / "csrc" / "runtime").glob("*.hip"))
CXX = "g++"
CXXFLAGS = [
    "-std=c++17",
    "-O2",
    "-fPIC",
    "-Wall",
    "-Wextra",
    "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}",
    f"-I{REPO / 'include'}",
    f"-I{HERE / 'csrc' / 'runtime'}",
]


def run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}{proc.stderr}")


def stale(target: Path, deps: list[Path], force: bool) -> bool:
    if force or not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build_code_object(name: str, src: Path, force: bool) -> Path:
    out = BUILD / f"{name}.hsaco"
    if stale(out, [src, *co.shared_headers(), Path(co.__file__)], force):
        run([co.HIPCC, *co.HIPCC_FLAGS, f"-I{co.KERNEL_DIR}", str(src), "-o", str(out)])
    return out


def write_blob(entries: list[tuple[str, Path, str]]) -> tuple[Path, Path]:
    asm = ["  .section .rodata"]
    for name, obj, _ in entries:
        asm += [
            "  .balign 4096",
            f"  .global gc_co_{name}_begin",
            f"gc_co_{name}_begin:",
            f'  .incbin "{obj}"',
            f"  .global gc_co_{name}_end",
            f"gc_co_{name}_end:",
        ]
    asm.append('  .section .note.GNU-stack,"",@progbits')
    cpp = ['#include "code_objects.hpp"', ""]
    for name, _, _ in entries:
        cpp.append(f'extern "C" const unsigned char gc_co_{name}_begin[], gc_co_{name}_end[];')
    cpp += ["", 'extern "C" const gc::CodeObject gc_embedded_code_objects[] = {']
    for name, obj, src_digest in entries:
        obj_digest = hashlib.sha256(obj.read_bytes()).hexdigest()
        cpp.append(
            f'  {{"{name}", gc_co_{name}_begin, gc_co_{name}_end, "{src_digest}", "{obj_digest}"}},'
        )
    cpp += ["};", f'extern "C" const uint32_t gc_embedded_code_object_count = {len(entries)};', ""]
    s_path, c_path = BUILD / "code_objects.S", BUILD / "code_objects.cpp"
    s_text, c_text = "\n".join(asm) + "\n", "\n".join(cpp)
    for path, text in ((s_path, s_text), (c_path, c_text)):
        if not path.exists() or path.read_text() != text:
            path.write_text(text)
    return s_path, c_path


def compile_host(src: Path, force: bool) -> Path:
    rel = src.relative_to(HERE).with_suffix(".o")
    out = BUILD / str(rel).replace(os.sep, "_")
    if stale(out, [src, *HOST_HEADERS, *sorted((HERE / "csrc" / "runtime").glob("*.hpp"))], force):
        run([CXX, *CXXFLAGS, "-c", str(src), "-o", str(out)])
    return out


def compile_hip_host(src: Path, force: bool) -> Path:
    rel = src.relative_to(HERE).with_suffix(".o")
    out = BUILD / str(rel).replace(os.sep, "_")
    if stale(out, [src, *HOST_HEADERS, Path(co.__file__)], force):
        flags = [f for f in co.HIPCC_FLAGS if f != "--genco"]
        run([co.HIPCC, *flags, "-fPIC", "-D__HIP_PLATFORM_AMD__", f"-I{HERE / 'csrc' / 'runtime'}", "-c",
             str(src), "-o", str(out)])
    return out


def build(force: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 4)
    sources = co.kernel_sources()
    if code_objects_manifest_problems := co.verify():
        print("note: code_objects.toml does not match the kernel sources "
              f"({len(code_objects_manifest_problems)} differences); run code_objects.py regenerate",
              file=sys.stderr)
    with cf.ThreadPoolExecutor(jobs) as pool:
        co_futs = {name: pool.submit(build_code_object, name, src, force) for name, src in sources}
        host_futs = [pool.submit(compile_host, src, force) for src in HOST_SOURCES]
        host_futs += [pool.submit(compile_hip_host, src, force) for src in HIP_HOST_SOURCES]
        objects = {name: f.result() for name, f in co_futs.items()}
        entries = [(name, objects[name], co.source_digest(src)) for name, src in sources]
        s_path, c_path = write_blob(entries)
        blob_s = pool.submit(
            lambda: (run(["gcc", "-c", str(s_path), "-o", str(BUILD / "code_objects_blob.o")]),
                     BUILD / "code_objects_blob.o")[1]
        )
        blob_c = pool.submit(
            lambda: (run([CXX, *CXXFLAGS, "-c", str(c_path), "-o", str(BUILD / "code_objects_table.o")]),
                     BUILD / "code_objects_table.o")[1]
        )
        host_objs = [f.result() for f in host_futs] + [blob_s.result(), blob_c.result()]
    tmp = OUT.with_suffix(".so.tmp")
    run(
        [
            CXX,
            "-shared",
            "-o",
            str(tmp),
            *map(str, host_objs),
            f"-L{ROCM / 'lib'}",
            "-lamdhip64",
            f"-Wl,-rpath,{ROCM / 'lib'}",
            f"-Wl,--version-script={HERE / 'csrc' / 'exports.map'}",
            "-pthread",
        ]
    )
    os.replace(tmp, OUT)
    return OUT


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--force", action="store_true", help="rebuild everything")
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args()
    out = build(force=args.force, jobs=args.jobs)
    print(f"built {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
