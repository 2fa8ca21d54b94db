"""Builds libfmm_hip.so for gfx950, in tree.

Steps: (1) the operator generator writes the HIP device operator headers
into ``csrc/_gen``; (2) hipcc compiles every ``csrc/*.hip`` translation
unit for gfx950 in parallel; (3) the objects are linked into
``fmmgen_amd/_build/libfmm_hip.so``.

Usage: python -m fmmgen_amd.build [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "_gen")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libfmm_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Compiled operator variants: generate_code() arguments of the device
# header and the name of the variant table in the library.
VARIANTS = {
    "so0_pf": dict(order=13, source_order=0, potential=True, field=True),
    "so0_p": dict(order=13, source_order=0, potential=True, field=False),
    "so0_f": dict(order=13, source_order=0, potential=False, field=True),
    "so1_f": dict(order=9, source_order=1, potential=False, field=True),
    "so2_pf": dict(order=9, source_order=2, potential=True, field=True),
}

HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I", CSRC,
]


def _write_if_changed(path, text):
    if os.path.exists(path):
        with open(path) as f:
            if f.read() == text:
                return
    with open(path, "w") as f:
        f.write(text)


def generate():
    """Device operator headers plus one base and one per-order unit per variant."""
    sys.path.insert(0, os.path.dirname(HERE))
    from fmmgen_amd.codegen.writer import device_source
    os.makedirs(GEN, exist_ok=True)
    units = []
    for stem, kw in VARIANTS.items():
        text = device_source(kw["order"], kw["potential"], kw["field"], kw["source_order"],
                             harmonic_derivs=True)
        _write_if_changed(os.path.join(GEN, f"dev_{stem}.h"), text)
        head = (f'#include "_gen/dev_{stem}.h"\n'
                f"#define FMM_VARIANT_NS v_{stem}\n")
        base = os.path.join(GEN, f"variant_{stem}.hip")
        _write_if_changed(base, head + f"#define FMM_VARIANT_NAME variant_{stem}\n"
                          '#include "variant_base.h"\n')
        units.append(base)
        start = kw["source_order"] + (1 if kw["field"] else 0)
        for p in range(start, kw["order"]):
            u = os.path.join(GEN, f"variant_{stem}_o{p}.hip")
            _write_if_changed(u, head + '#include "variant_order.h"\n' + f"FMM_INSTANTIATE_ORDER({p})\n")
            units.append(u)
    return units


def _up_to_date(src, obj):
    """True when obj and its dependency file are newer than every input."""
    dep = obj[:-2] + ".d"
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    with open(dep) as f:
        text = f.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    t = os.path.getmtime(obj)
    return all(os.path.exists(x) and os.path.getmtime(x) <= t for x in files + [src])


def _compile(src, obj, incremental=False):
    if incremental and _up_to_date(src, obj):
        return obj
    cmd = [HIPCC, *HIPFLAGS, "-MMD", "-MF", obj[:-2] + ".d", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr[-6000:]}")
    return obj


def build(jobs=None, verbose=False, incremental=False):
    """Compile every unit (incremental=True: only those whose inputs changed)."""
    units = generate()
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")] + units

    def weight(path):   # longest compiles first: high-order units
        stem = os.path.basename(path)[:-4]
        tail = stem.rsplit("_o", 1)
        return -int(tail[1]) if len(tail) == 2 and tail[1].isdigit() else 0
    srcs.sort(key=weight)
    objs = []
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for s in srcs:
            obj = os.path.join(OUT, os.path.basename(s)[:-4] + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, s, obj, incremental))
        for f in futs:
            f.result()
            if verbose:
                print("compiled", f.result(), flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    print(build(a.j, verbose=True, incremental=a.incremental))
