"""Build step: compiles libigr_hip.so for gfx950 in-tree (hipcc, via the
Makefile in csrc/).  Always recompiles what changed; `force=True` rebuilds
everything."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(force=False, jobs=8):
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    env = dict(os.environ)
    env.setdefault("HIPCC", "/opt/rocm/bin/hipcc")
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True, env=env)
    return os.path.join(_HERE, "lib", "libigr_hip.so")


if __name__ == "__main__":
    print(build())
