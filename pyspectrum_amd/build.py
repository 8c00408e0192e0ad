"""Build libpsx.so (gfx950) in-tree: ``python -m pyspectrum_amd.build``.

Each ``csrc/*.hip`` is compiled by hipcc into an object under ``_build/``
and the objects are linked into ``pyspectrum_amd/libpsx.so``.  Packed FP32
instructions are disabled for all device code.  The tile-shape definitions
chosen by the autotuner (``pyspectrum_amd/tuning.json``, written by
``python -m pyspectrum_amd.tune``) are applied when present.
"""
import concurrent.futures as cf
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libpsx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PSX_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns about the device-only feature flag; keep real diagnostics
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, err))
    return err


TUNING = os.path.join(HERE, "tuning.json")


def tuned_defines():
    """-D flags of the autotuner's choice (empty without tuning.json)."""
    if not os.path.isfile(TUNING):
        return []
    with open(TUNING) as f:
        d = json.load(f).get("defines", {})
    return ["-D%s=%d" % (k, int(v)) for k, v in sorted(d.items())]


def build(verbose=False, jobs=8, defines=None, lib=None, objdir=None):
    """Compile and link the library; ``defines`` (list of -D flags) default to
    the tuned ones, ``lib``/``objdir`` to the in-tree paths."""
    defines = tuned_defines() if defines is None else list(defines)
    lib = lib or LIB
    objdir = objdir or BUILD
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(objdir, s[:-4] + ".o") for s in srcs]

    def one(src, obj):
        return _run([HIPCC] + CFLAGS + defines + ["-c", os.path.join(CSRC, src), "-o", obj])

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(one, s, o) for s, o in zip(srcs, objs)]
        for s, f in zip(srcs, futs):
            msg = f.result()
            if verbose and msg.strip():
                print(s, msg, file=sys.stderr)
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
