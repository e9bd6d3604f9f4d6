"""Build the C++ test programs of tests/cpp/ against include/ and
libsygraph_hip into tests/cpp/_bin/ (git-ignored; the binaries travel with
the tree to the GPU machine).  Called by __graft_entry__.build().

    python tests/cpp/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O2", "-std=c++20", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-w"]


def build() -> list[str]:
    """Compile every tests/cpp/*.cpp; returns the names that failed."""
    os.makedirs(OUT, exist_ok=True)
    failed = []
    for src in sorted(f for f in os.listdir(HERE) if f.endswith(".cpp")):
        name = src[:-4]
        cmd = [HIPCC, *FLAGS, "-I", os.path.join(REPO, "include"), os.path.join(HERE, src),
               "-L", os.path.join(REPO, "sygraph_amd"), "-lsygraph_hip", "-Wl,-rpath,$ORIGIN/../../../sygraph_amd",
               "-o", os.path.join(OUT, name)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            print(r.stdout, r.stderr, file=sys.stderr)
            failed.append(name)
    return failed


if __name__ == "__main__":
    bad = build()
    print("failed:" if bad else "ok", *bad)
    sys.exit(1 if bad else 0)
