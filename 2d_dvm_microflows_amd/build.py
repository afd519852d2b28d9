"""Builds libdvmhip.so for gfx950 in-tree (hipcc, no JIT cache)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 parity: no contraction of a*b+c into FMA, as in the reference build
    "-ffp-contract=off",
    # hardware FP atomics (ds_add_f64 / global_atomic_add_f64)
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_hip(verbose=False):
    src = os.path.join(HERE, "csrc", "dvm_capi.hip")
    out_dir = os.path.join(HERE, "lib")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libdvmhip.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build_hip(verbose=True)
