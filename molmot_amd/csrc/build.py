"""Build libmolmot_hip.so (gfx950) from the .hip sources in this folder.

    python -m molmot_amd.csrc.build        # or __graft_entry__.build()

Each source is compiled to an object by hipcc in parallel, then linked into
molmot_amd/libmolmot_hip.so.  Flags:
  -ffp-contract=off   the detuning stage must keep the reference's rounding;
                      fused multiply-adds are written explicitly instead
  -target-feature -packed-fp32-ops
                      no packed FP32 (v_pk_*_f32) instructions are generated
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libmolmot_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("MOLMOT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# device headers that the run-time compiled (hipRTC) kernels of
# rate_spec.hip include; their text is embedded in the library
RTC_HEADERS = ["mm_plan.h", "mm_queue_dev.h", "philox.h", "rate_core.h", "mm_traj.h", "rate_adjoint.h",
               "mm_tangent.h", "mm_vjp.h", "mm_vsq.h", "mm_axis3d.h"]


def sources():
    return sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".hip"))


def _embed_headers():
    """Write _build/mm_rtc_headers.hip: the RTC_HEADERS as string literals."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    out = os.path.join(OBJ_DIR, "mm_rtc_headers.hip")
    lines = ["// generated by molmot_amd/csrc/build.py: device headers for hipRTC", ""]
    srcs = []
    for h in RTC_HEADERS:
        with open(os.path.join(HERE, h)) as f:
            text = f.read()
        if ")MMSRC\"" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        srcs.append('R"MMSRC(' + text + ')MMSRC"')
    lines.append("extern const char* const mm_rtc_header_names[] = {" + ", ".join(f'"{h}"' for h in RTC_HEADERS) + "};")
    lines.append("extern const char* const mm_rtc_header_srcs[] = {" + ",\n".join(srcs) + "};")
    lines.append(f"extern const int mm_rtc_n_headers = {len(RTC_HEADERS)};")
    text = "\n".join(lines) + "\n"
    old = None
    if os.path.exists(out):
        with open(out) as f:
            old = f.read()
    if old != text:
        with open(out, "w") as f:
            f.write(text)
    return out


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the MI355X kernels cannot be built")
    return exe


def _compile(src):
    os.makedirs(OBJ_DIR, exist_ok=True)
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    return obj


def build(verbose: bool = False) -> str:
    srcs = sources() + [_embed_headers()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
