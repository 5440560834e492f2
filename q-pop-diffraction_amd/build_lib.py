"""Build libqpd.so for gfx950 in-tree (qpop_diffraction/lib/libqpd.so).

    python q-pop-diffraction_amd/build_lib.py [--verbose] [-j JOBS]

hipcc compiles the HIP sources for gfx950 only, with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).  The host code and
generic kernels (qpd.hip) and the parts of the compile-time-length kernel registry
(qpd_fast_sizes.hip, once per part) are compiled in parallel, then linked.
"""

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
FAST_PARTS = 8  # kFastParts in csrc/qpd_launch.hpp
OUT = HERE / "qpop_diffraction" / "lib" / "libqpd.so"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def units(objdir):
    """(command, object) of every translation unit."""
    inc = [f"-I{ROOT / 'include'}"]
    out = [([hipcc()] + FLAGS + inc + ["-c", str(CSRC / "qpd.hip"), "-o", str(objdir / "qpd.o")], objdir / "qpd.o")]
    for part in range(FAST_PARTS):
        obj = objdir / f"qpd_fast_{part}.o"
        out.append(([hipcc()] + FLAGS + inc + [f"-DQPD_FAST_PART={part}", "-c", str(CSRC / "qpd_fast_sizes.hip"),
                                               "-o", str(obj)], obj))
    return out


def build(verbose=False, jobs=None):
    OUT.parent.mkdir(parents=True, exist_ok=True)
    jobs = jobs or min(len(units(Path("."))), os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory(prefix="qpd_build_") as tmpd:
        objdir = Path(tmpd)
        tus = units(objdir)

        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(max_workers=jobs) as pool:
            results = list(pool.map(run, [c for c, _ in tus]))
        for (cmd, _), res in zip(tus, results):
            if res.returncode != 0:
                sys.stderr.write(res.stderr)
                raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd[-4:])}")
        tmp = OUT.with_suffix(".so.tmp")
        link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for _, o in tus]
        res = run(link)
        if res.returncode != 0:
            sys.stderr.write(res.stderr)
            raise RuntimeError(f"link failed ({res.returncode})")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    print(build(verbose="--verbose" in sys.argv, jobs=j))
