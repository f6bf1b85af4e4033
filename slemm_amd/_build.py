"""Build of the gfx950 engine library ``slemm_amd/libslemm_gpu.so`` with hipcc.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled to an object in ``slemm_amd/_build``
(in parallel) and linked into one shared library exposing the C ABI of
``include/slemm_gpu.h``.  The library is built in-tree so that it travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libslemm_gpu.so"
ARCH = os.environ.get("SLEMM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


# per-file device flags: the MFMA-ceiling probe keeps its accumulators in VGPRs, as the block-product
# kernels do by themselves (with AGPR accumulators the register allocator rotated the probe's
# independent chains through accvgpr copies inside its loop)
EXTRA = {"probe.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _generate() -> None:
    """Generated sources into OBJ: the MT19937 jump polynomials (csrc/gen_mt_jump.py, a few
    seconds), remade when the generator is newer than its output."""
    gen, inc = CSRC / "gen_mt_jump.py", OBJ / "mt_jump.inc"
    if inc.exists() and inc.stat().st_mtime >= gen.stat().st_mtime:
        return
    p = subprocess.run([sys.executable, str(gen)], capture_output=True, text=True, cwd=str(CSRC))
    if p.returncode != 0:
        raise RuntimeError("gen_mt_jump.py failed:\n" + p.stderr)
    tmp = inc.with_suffix(".tmp")
    tmp.write_text(p.stdout)
    os.replace(tmp, inc)


def _compile(src: Path) -> tuple[Path, str]:
    obj = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *EXTRA.get(src.name, []), "-c", str(src), "-o", str(obj)]
    else:
        # host-only translation units: C++ with the HIP runtime headers, no device pass; the
        # generated sources in OBJ on the include path
        cmd = [HIPCC, *[f for f in COMMON if not f.startswith("--offload-arch")], "-x", "c++",
               "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{OBJ}", "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless
    log = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{log}")
    return obj, log


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile all sources and link ``libslemm_gpu.so``; returns the library path."""
    OBJ.mkdir(exist_ok=True)
    _generate()
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, log in ex.map(_compile, sources):
            objs.append(obj)
            if verbose and log.strip():
                print(log, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    # rocBLAS / rocSOLVER: the FP64 window systems of the LMM quadratic forms (gstat.hip)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           "-L/opt/rocm/lib", "-lrocblas", "-lrocsolver", "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
