"""Build the ccfj native parts in-tree (no pip install, no JIT cache).

    python cc-fjpy_amd/build.py

Produces, next to the Python package so they travel with the source tree:
  ccfj/libccfj.so          HIP kernels + C ABI (include/ccfj.h), gfx950
  ccfj/_ccfj.*.so          Cython shim linked against libccfj.so
  ccfj/libccfj_emul.so     host build of the kernel arithmetic (CPU tests)
  ccfj/libccfj_calib.so    instruction-issue calibration kernels (bench.py --valu-calib)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(HERE, "ccfj")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("CCFJ_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are kept out of generated code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = ["fj_kernels.hip", "fj_kernels_linseg.hip", "ccfj_capi.hip", "fj_plain64.hip",
               "fj_batched.hip", "ccfj_batch.hip", "fj_mwfj.hip", "ccfj_mwfj.hip",
               "fj_hilbert.hip", "fj_cnn.hip", "ccfj_cnn.hip", "fj_synth.hip", "fj_cc.hip",
               "ccfj_cc.hip", "ccfj_stage.hip"]
# per-source options on top of HIP_FLAGS: the dense linear-segment kernels
# and the persistent batched kernels are compiled without machine-level
# loop-invariant code motion, which otherwise keeps the small-argument
# polynomial constants live in vector registers through the station loop
# and spills them (fj_kernels_linseg.hip, fj_batched.hip)
NO_MACHINE_LICM = ["-mllvm", "-disable-machine-licm"]
HIP_SOURCE_FLAGS = {"fj_kernels_linseg.hip": NO_MACHINE_LICM, "fj_batched.hip": NO_MACHINE_LICM}
BUILD_DIR = os.path.join(HERE, "_build")


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_hip(jobs=8):
    """Each source to an object (in parallel, with its own options), then
    one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(PKG, "libccfj.so")
    os.makedirs(BUILD_DIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, *HIP_SOURCE_FLAGS.get(src, []), "-I", INCLUDE, "-c",
              "-o", obj, os.path.join(CSRC, src)])
        return obj

    with ThreadPoolExecutor(max(1, jobs)) as pool:
        objs = list(pool.map(compile_one, HIP_SOURCES))
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, *objs,
          "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"])
    return out


def build_calib():
    """Instruction-issue calibration kernels (bench.py --valu-calib); a
    measurement aid, kept out of libccfj.so."""
    out = os.path.join(PKG, "libccfj_calib.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, os.path.join(CSRC, "valu_calib.hip")])
    return out


def build_cython():
    import numpy
    pyx = os.path.join(PKG, "_ccfj.pyx")
    cpp = os.path.join(PKG, "_ccfj.cpp")
    _run([sys.executable, "-m", "cython", "-3", "--cplus", "-o", cpp, pyx])
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "_ccfj" + ext)
    pyinc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp",
          "-DNPY_NO_DEPRECATED_API=NPY_1_7_API_VERSION",
          "-I", pyinc, "-I", numpy.get_include(), "-I", INCLUDE,
          "-o", out, cpp, "-L", PKG, "-lccfj", "-Wl,-rpath,$ORIGIN"])
    os.remove(cpp)
    return out


def build_emul():
    out = os.path.join(PKG, "libccfj_emul.so")
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-I", INCLUDE,
          "-o", out, os.path.join(CSRC, "fj_host_emul.cpp")])
    return out


def build_all():
    build_hip()
    build_cython()
    build_emul()
    build_calib()


if __name__ == "__main__":
    build_all()
