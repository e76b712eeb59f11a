"""In-tree build of libfunlbm_hip.so for gfx950 with hipcc.

Every HIP source under csrc/ is compiled to an object (in parallel) and linked
into funlbm_amd/_lib/libfunlbm_hip.so, which the ctypes loader opens.
"""

import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libfunlbm_hip.so")

ARCH = "gfx950"
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # the kernels follow the reference's float operation order; no FMA contraction
    "-ffp-contract=off",
    # packed fp32 instructions disabled
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build funlbm_amd")
    return path


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip")
    )


# Kernels whose register allocation decides the step's speed: the build
# fails if the compiler spills them to private memory (the allocation is
# sensitive to unrelated code; a spill costs several times the step time).
HOT_KERNELS = re.compile(r"k_fused_step|k_batch_fused")
RESOURCE_REMARKS = ["-Rpass-analysis=kernel-resource-usage"]


def resource_report(stderr: str):
    """{kernel: {field: value}} from the kernel-resource-usage remarks."""
    out, name = {}, None
    for line in stderr.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark:\s+([^:]+): (\S+) \[-Rpass-analysis", line)
        if m and name:
            out[name][m.group(1).strip()] = m.group(2)
    return out


def check_hot_kernels(report, src):
    bad = []
    for name, r in report.items():
        if not HOT_KERNELS.search(name):
            continue
        # (scalar spills go to vector-register lanes, not to memory)
        scratch = int(r.get("ScratchSize [bytes/lane]", "0"))
        spills = int(r.get("VGPRs Spill", "0"))
        if scratch or spills:
            bad.append(f"{name}: scratch {scratch} B/lane, VGPR spills {spills}")
    if bad:
        raise RuntimeError(f"hot kernels of {src} spill to private memory:\n" + "\n".join(bad))


def _compile(src, obj):
    cmd = [hipcc(), *HIPCC_FLAGS, *RESOURCE_REMARKS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{res.stderr}")
    check_hot_kernels(resource_report(res.stderr), src)
    return obj


def build_native(verbose: bool = False, jobs: int = 8) -> str:
    """Compiles every csrc/*.hip for gfx950 and links the shared library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    with tempfile.TemporaryDirectory(prefix="funlbm_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]
        with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
            list(ex.map(_compile, srcs, objs))
        out_tmp = LIB_PATH + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out_tmp]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{res.stderr}")
        os.replace(out_tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}", file=sys.stderr)
    return LIB_PATH


if __name__ == "__main__":
    build_native(verbose=True)
