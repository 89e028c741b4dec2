"""Builds libgevpm.so (the HIP library behind include/gev_pm.h) for gfx950.

Each .hip source is compiled to an object with hipcc in parallel, then the
objects are linked into gevolution-1.2_amd/lib/libgevpm.so, in-tree so the
library travels with the source tree.  Run as a script or through
__graft_entry__.build().
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libgevpm.so")
# per-kernel register / scratch / LDS / occupancy report of the last build
RESOURCES = os.path.join(LIB_DIR, "kernel_resources.tsv")

ARCH = os.environ.get("GEV_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add contraction, so every kernel
# evaluates the reference's expressions with the same roundings as a plain
# CPU build (bit-identical results where inputs agree).
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIPFLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Rpass-analysis=kernel-resource-usage",
]

_FIELDS = (("VGPRs", "vgprs"), ("AGPRs", "agprs"), ("TotalSGPRs", "sgprs"),
           ("ScratchSize [bytes/lane]", "scratch"), ("Occupancy [waves/SIMD]", "occupancy"),
           ("LDS Size [bytes/block]", "lds"))


def _resources(stderr):
    """Kernel rows from the compiler's kernel-resource-usage remarks."""
    rows, cur = [], None
    for line in stderr.splitlines():
        if "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split(" [-Rpass-analysis")[0].strip()
        if body.startswith("Function Name:"):
            cur = {"kernel": body.split(":", 1)[1].strip()}
            rows.append(cur)
        elif cur is not None:
            for label, key in _FIELDS:
                if body.startswith(label + ":"):
                    cur[key] = body.split(":", 1)[1].strip()
    return rows


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc()] + HIPFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    # keep warnings other than the host pass not knowing the packed-fp32
    # feature and the resource remarks (collected separately)
    lines = [l for l in r.stderr.splitlines()
             if "packed-fp32-ops" not in l and "remark:" not in l and "kernel-resource-usage" not in l
             and not l.lstrip().startswith(("|", "^")) and not re_location(l) and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")
    return obj, _resources(r.stderr)


def re_location(line):
    """Source echo lines the compiler prints under each remark ("  123 | code")."""
    t = line.strip()
    head = t.split("|", 1)[0].strip()
    return "|" in t and (head.isdigit() or head == "")


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        done = list(ex.map(_compile, srcs))
    objs = [o for o, _ in done]
    rows = [r for _, rs in done for r in rs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    with open(RESOURCES, "w") as f:
        f.write("kernel\t" + "\t".join(k for _, k in _FIELDS) + "\n")
        for r in sorted(rows, key=lambda r: r["kernel"]):
            f.write(r["kernel"] + "\t" + "\t".join(r.get(k, "") for _, k in _FIELDS) + "\n")
    spilled = [r["kernel"] for r in rows if r.get("scratch", "0") not in ("0", "")]
    if spilled:
        sys.stderr.write("warning: kernels using scratch memory (register spills):\n  " + "\n  ".join(spilled) + "\n")
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
