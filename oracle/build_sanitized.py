"""Build recipe of the sanitized host-only checks (test infrastructure only).

Compiles, with AddressSanitizer + UndefinedBehaviorSanitizer (kind "asan")
or ThreadSanitizer (kind "tsan"), host code only (no .hip file, no device):
  * this directory's oracle (elas_oracle.cpp, elas_oracle_post.cpp) and the
    reference's own src/triangle.cpp under namespace elas_oracle_tri, read
    where it lies;
  * the library's host Delaunay step: csrc/elas_triangle.cpp (worker pool,
    split merge, block cache) and the ELAS_TRI_INSTANCES namespaced copies of
    csrc/elas_tri_unit.cpp, each including the reference's triangle.cpp;
  * sanitize_main.cpp, which drives both (see its header);
into oracle/_build/san_<kind>/elas_sanitized.  Rebuilt only when a source is
newer than the binary.  Needs the reference checkout (Triangle's source).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "cuda-stereo-matching_amd", "csrc")
REF_SRC = "/root/reference/src"
TRI_INSTANCES = 16
FLAGS = {
    "asan": ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
    "tsan": ["-fsanitize=thread"],
}


def _sources():
    return [os.path.join(HERE, f) for f in ("elas_oracle.cpp", "elas_oracle_post.cpp", "sanitize_main.cpp")] + [
        os.path.join(CSRC, f) for f in ("elas_triangle.cpp", "elas_triangle.h", "elas_tri_unit.cpp")] + [
        os.path.join(ROOT, "include", "elas_hip.h"), os.path.join(REF_SRC, "triangle.cpp"), os.path.abspath(__file__)]


def binary(kind):
    return os.path.join(HERE, "_build", "san_" + kind, "elas_sanitized")


def build(kind="asan", jobs=None):
    """Compile and link; returns the binary's path."""
    if not os.path.isfile(os.path.join(REF_SRC, "triangle.cpp")):
        raise RuntimeError("the sanitized build needs the reference checkout (%s)" % REF_SRC)
    out = binary(kind)
    if os.path.isfile(out) and os.path.getmtime(out) > max(os.path.getmtime(s) for s in _sources()):
        return out
    bdir = os.path.dirname(out)
    os.makedirs(bdir, exist_ok=True)
    base = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fPIC", "-pthread",
            "-w"] + FLAGS[kind]
    cmds, objs = [], []

    def add(src, extra, name):
        o = os.path.join(bdir, name + ".o")
        cmds.append(base + extra + ["-c", src, "-o", o])
        objs.append(o)

    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", REF_SRC]
    add(os.path.join(HERE, "elas_oracle.cpp"), inc, "elas_oracle")
    add(os.path.join(HERE, "elas_oracle_post.cpp"), inc, "elas_oracle_post")
    add(os.path.join(REF_SRC, "triangle.cpp"), ["-DTRILIB=elas_oracle_tri"], "triangle_oracle")
    add(os.path.join(CSRC, "elas_triangle.cpp"), inc, "elas_triangle")
    for k in range(TRI_INSTANCES):
        add(os.path.join(CSRC, "elas_tri_unit.cpp"), inc + ["-DTRILIB=elas_tri%d" % k], "tri_unit_%d" % k)
    add(os.path.join(HERE, "sanitize_main.cpp"), inc, "sanitize_main")
    jobs = jobs or min(16, os.cpu_count() or 4)

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        return cmd, r.returncode, r.stdout + r.stderr

    with cf.ThreadPoolExecutor(jobs) as ex:
        for cmd, rc, text in ex.map(run, cmds):
            if rc != 0:
                raise RuntimeError("sanitized build failed: %s\n%s" % (" ".join(cmd), text))
    cmd, rc, text = run(base + objs + ["-o", out + ".tmp"])
    if rc != 0:
        raise RuntimeError("sanitized link failed: %s\n%s" % (" ".join(cmd), text))
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    for k in sys.argv[1:] or ["asan"]:
        print(build(k))
