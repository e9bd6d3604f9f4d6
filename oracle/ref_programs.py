"""Compile the reference's own example programs and C++ tests, unchanged,
against this repository's headers (include/) and libsygraph_hip.

Test infrastructure only.  The reference is header-only SYCL and cannot be
built as such here (no SYCL compiler in the image); what can be built are its
user-side programs -- examples/*/*.cpp and tests/*/*.cpp -- because they only
use the operator/algorithm API that include/ re-implements for HIP.  Running
them on an MI355X shows the drop-in boundary holds for the reference's own
callers.  Sources are read in place from the reference tree; binaries go to
oracle/_ref/ only (git-ignored).

Not built: tests/graph/graph.cpp (sycl::buffer / accessor / host_accessor,
which the queue facade does not provide).

tests/frontier/bitmap_frontier.cpp is built with -DNDEBUG, as a Release build
of the reference compiles it: its asserts contradict each other (line 13
asserts getNumElems() == 0 and line 15 getNumElems() == 128 on the same
frontier), so with asserts enabled it cannot pass on any implementation.  It
still runs its device-side insert kernel and the host remove.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_ref")
DEFAULT_REF = "/root/reference"

EXAMPLES = ["bfs", "sssp", "cc", "tc", "bc"]
TESTS = [
    "frontier/bitmap_frontier", "frontier/mlb_frontier", "formats/csr", "formats/coo2csr",
    "formats/coo2csr_unweighted", "formats/properties", "graph/graph_teardown", "operators/advance",
    "operators/advance_graph", "operators/advance_pull", "operators/filter_compute", "operators/intersection",
    "algorithms/bfs", "algorithms/sssp", "algorithms/cc", "algorithms/tc", "algorithms/bc",
]

EXTRA = {"frontier/bitmap_frontier": ["-DNDEBUG"]}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O2", "-std=c++20", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-w"]


def _link():
    lib_dir = os.path.join(REPO, "sygraph_amd")
    # $ORIGIN-relative rpath so the binaries find the library wherever the tree is copied
    return ["-L", lib_dir, "-lsygraph_hip", "-Wl,-rpath,$ORIGIN/../../../sygraph_amd"]


def _jobs(ref: str):
    jobs = []
    for ex in EXAMPLES:
        for prof in (False, True):
            name = ex + ("_profiling" if prof else "")
            cmd = [HIPCC, *FLAGS, "-DGRAPH_LOCATION=1", *(["-DENABLE_PROFILING"] if prof else []),
                   "-I", os.path.join(REPO, "include"), os.path.join(ref, "examples", ex, ex + ".cpp"), *_link(),
                   "-o", os.path.join(OUT, "examples", name)]
            jobs.append((name, cmd))
    # the graph in managed memory with the prefetch of memoryAlloc<shared>
    # (examples/CMakeLists.txt GRAPH_LOCATION / ENABLE_PREFETCH options)
    cmd = [HIPCC, *FLAGS, "-DGRAPH_LOCATION=2", "-DENABLE_PREFETCH", "-I", os.path.join(REPO, "include"),
           os.path.join(ref, "examples", "bfs", "bfs.cpp"), *_link(), "-o", os.path.join(OUT, "examples", "bfs_shared")]
    jobs.append(("bfs_shared", cmd))
    for t in TESTS:
        name = t.replace("/", "_")
        # asserts stay enabled (no -DNDEBUG): the tests check with assert()
        cmd = [HIPCC, *FLAGS, *EXTRA.get(t, []), "-I", os.path.join(REPO, "include"), "-I", os.path.join(ref, "tests", "include"),
               os.path.join(ref, "tests", t + ".cpp"), *_link(), "-o", os.path.join(OUT, "tests", name)]
        jobs.append((name, cmd))
    return jobs


def build(ref: str = DEFAULT_REF, workers: int = 8) -> list[str]:
    """Compile every program; returns the names that failed (empty on success)."""
    if not os.path.isdir(ref):
        return []
    os.makedirs(os.path.join(OUT, "examples"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "tests"), exist_ok=True)
    failed = []

    def run(job):
        name, cmd = job
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            return name, r.stderr[-4000:]
        return name, None

    with ThreadPoolExecutor(max_workers=workers) as ex:
        for name, err in ex.map(run, _jobs(ref)):
            if err is not None:
                failed.append(name)
                sys.stderr.write(f"[ref_programs] {name} failed:\n{err}\n")
    return failed


if __name__ == "__main__":
    bad = build(sys.argv[1] if len(sys.argv) > 1 else DEFAULT_REF)
    print("failed:", bad)
    sys.exit(1 if bad else 0)
