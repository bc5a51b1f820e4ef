"""Build recipe of the CPU oracle (test infrastructure only).

Compiles
  * the reference's own Triangle (``src/triangle.cpp`` of the reference
    repository, unmodified, read where it lies) into ``oracle/_ref/`` under the
    namespace ``elas_oracle_tri``;
  * this directory's C++ restatement of the reference kernels;
into ``oracle/_build/libelas_oracle.so``.  Without the reference checkout the
already-built library is kept only while it is newer than this directory's
sources (the GPU box only loads it); otherwise the build fails.
FP contraction is off everywhere: Triangle's exact predicates and the float
plane evaluation both depend on it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_SRC = "/root/reference/src"
REF_DIR = os.path.join(HERE, "_ref")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libelas_oracle.so")

CXXFLAGS = ["-O3", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c++17"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


def build(verbose=False):
    """Compile the oracle; returns the library path."""
    if not os.path.isfile(os.path.join(REF_SRC, "triangle.cpp")):
        if not os.path.isfile(LIB):
            raise RuntimeError("reference checkout missing and no prebuilt oracle at " + LIB)
        srcs = [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith((".cpp", ".h", ".py"))]
        srcs.append(os.path.join(ROOT, "include", "elas_hip.h"))
        if os.path.getmtime(LIB) < max(os.path.getmtime(s) for s in srcs if os.path.isfile(s)):
            raise RuntimeError("oracle sources changed after %s was built, and it cannot be rebuilt "
                               "without the reference checkout (%s holds Triangle)" % (LIB, REF_SRC))
        return LIB
    os.makedirs(REF_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    tri_o = os.path.join(REF_DIR, "triangle_oracle.o")
    cmds = [["g++", "-O2", "-ffp-contract=off", "-fPIC", "-w", "-DTRILIB=elas_oracle_tri", "-c",
             os.path.join(REF_SRC, "triangle.cpp"), "-o", tri_o]]
    objs = [tri_o]
    for src in ("elas_oracle.cpp", "elas_oracle_post.cpp"):
        obj = os.path.join(OUT_DIR, src.replace(".cpp", ".o"))
        cmds.append(["g++"] + CXXFLAGS + ["-I", os.path.join(ROOT, "include"), "-I", REF_SRC, "-c",
                                          os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for c in cmds]
    for c, p in zip(cmds, procs):
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("oracle build failed: %s\n%s%s" % (" ".join(c), out, err))
        if verbose and err:
            print(err, file=sys.stderr)
    _run(["g++", "-shared", "-o", LIB + ".tmp"] + objs)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
