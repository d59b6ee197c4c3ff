"""Build recipe of libcfhip.so (HIP for gfx950) — the product library.

    python channelflow.jl_amd/build.py

Compiles every source under csrc/ with hipcc for gfx950 into
channelflow.jl_amd/lib/libcfhip.so.  The library exports the C ABI declared
in include/channelflow_hip.h and has no torch dependency.
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcfhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# heaviest translation units first (the wall-normal layouts, then the x/z passes)
SOURCES = [*[f"kernels_y_p{p}.hip" for p in (6, 7, 4, 3, 2, 1, 5)], "kernels_xz.hip", "kernels_closure.hip", "kernels_y.hip",
           "kernels_rng.hip", "scratch.hip", "setup.cpp", "api.cpp"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 code generation (v_pk_*_f32) anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Werror=return-type", "-Werror=uninitialized",
]


RESOURCES = os.path.join(LIBDIR, "kernel_resources.txt")
OBJDIR = os.path.join(LIBDIR, "obj")


def _resource_table(remarks):
    """Per-kernel VGPR/AGPR/scratch/occupancy from the compiler's
    kernel-resource-usage remarks."""
    rows = []
    for blk in remarks.split("Function Name: ")[1:]:
        name = blk.split()[0]

        def get(key):
            m = re.search(key + r": (\d+)", blk)
            return int(m.group(1)) if m else -1

        rows.append((name, get("VGPRs"), get("AGPRs"), get(r"ScratchSize \[bytes/lane\]"),
                     get(r"Occupancy \[waves/SIMD\]")))
    return rows


def _compile(src, verbose):
    """One translation unit -> object file; returns (returncode, stderr)."""
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-Rpass-analysis=kernel-resource-usage", "-I", CSRC, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stderr, obj


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", str(os.cpu_count() or 6)))))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda s: _compile(s, verbose), srcs))
    stderr = "\n".join(r[1] for r in results)
    diag = [l for l in stderr.splitlines() if ("warning:" in l or "error:" in l) and "packed-fp32-ops" not in l]
    if diag:
        print("\n".join(diag), file=sys.stderr, flush=True)
    for (rc, _, _), src in zip(results, srcs):
        if rc != 0:
            raise subprocess.CalledProcessError(rc, src)
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[r[2] for r in results], "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    remarks = "\n".join(l.split("remark:", 1)[1] for l in stderr.splitlines() if "remark:" in l)
    rows = sorted(set(_resource_table(remarks)))
    with open(RESOURCES, "w") as f:
        f.write("kernel vgpr agpr scratch_bytes_per_lane occupancy_waves_per_simd\n")
        for row in rows:
            f.write(" ".join(str(x) for x in row) + "\n")
    flagged = [row[0] for row in rows if row[3] > 0]
    spills = _scratch_accesses(flagged, srcs, results, verbose) if flagged else []
    if spills:  # scratch on the device path is a register spill or an addressable local
        os.remove(tmp)
        raise RuntimeError("kernels using scratch memory: " + ", ".join(spills))
    os.replace(tmp, LIB)
    return LIB


def _scratch_accesses(kernels, srcs, results, verbose):
    """Of the kernels whose frame has a private segment, those whose code
    touches it.  The compiler can reserve a frame slot (for instance a
    register-scavenging slot of SGPR spills) that no instruction uses; such a
    kernel runs without scratch traffic and is accepted.  Decided from the
    device assembly of the sources that define the kernels (compiled in
    parallel)."""
    from concurrent.futures import ThreadPoolExecutor

    todo = []
    for src, (_, err, _) in zip(srcs, results):
        mine = [k for k in kernels if ("Function Name: " + k) in err]
        if mine:
            todo.append((src, mine))

    def check(item):
        src, mine = item
        asm = os.path.join(OBJDIR, os.path.basename(src) + ".s")
        cmd = [HIPCC, *FLAGS, "--cuda-device-only", "-S", "-I", CSRC, src, "-o", asm]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, capture_output=True)
        text = open(asm).read()
        os.remove(asm)
        used, notes = [], []
        for k in mine:
            start = text.find("\n" + k + ":")
            end = text.find(".Lfunc_end", start)
            body = text[start:end] if start >= 0 else ""
            if start < 0 or re.search(r"\n\s+(scratch_|buffer_|s_swappc)", body):
                used.append(k)
            else:
                notes.append(f"{k}: reserved private frame, no scratch access in its code")
        return used, notes

    used = []
    with ThreadPoolExecutor(max(1, len(todo))) as pool:
        for u, notes in pool.map(check, todo):
            used += u
            if verbose and notes:
                print("\n".join(notes), flush=True)
    return used


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
