"""Build recipes for the native parts, all in-tree.

    libppf.so      HIP kernels for gfx950 + the C ABI of include/ppf.h (hipcc)
    libppfcal.so   copy kernel of known traffic for measurement calibration
    libppfio.so    HDF5 file contract of include/ppf_io.h (gcc + libhdf5)
    ppf            the drop-in stepping program (gcc, links both)
    oracle/liboracle.so  the CPU oracle (test infrastructure, gcc)

Every call compiles: nothing is skipped because an output already exists.
"""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(REPO, "include")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the bitwise FP64 kernels must round every product and
# sum on their own, as the reference's x86-64 build does; the FP32 and fast
# FP64 kernels (PhysC in csrc/ppf_physics.h) write their fused multiply-adds
# explicitly with fma().  Packed FP32 instructions are switched off.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=REPO):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n"
                           f"{r.stdout[-4000:]}\n{r.stderr[-4000:]}")
    return r


def build_libppf():
    """The three translation units are compiled in parallel (each holds its
    own device code; no device symbol crosses them), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    out = os.path.join(PKG, "libppf.so")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = ("ppf.hip", "ppf_ens.hip", "unet.hip")
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.splitext(f)[0] + ".o") for f in srcs]
    with ThreadPoolExecutor(len(srcs)) as ex:
        list(ex.map(lambda so: _run([HIPCC, *flags, f"-I{INC}", f"-I{CSRC}", "-c",
                                     os.path.join(CSRC, so[0]), "-o", so[1]]), zip(srcs, objs)))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_libppfcal():
    out = os.path.join(PKG, "libppfcal.so")
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "calib.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def have_hdf5():
    return os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")) and \
        os.path.exists(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))


def build_libppfio():
    out = os.path.join(PKG, "libppfio.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra", "-Werror",
          f"-I{INC}", f"-I{HDF5_DIR}/include", os.path.join(CSRC, "ppf_h5.c"),
          f"-L{HDF5_DIR}/lib", f"-Wl,-rpath,{HDF5_DIR}/lib", "-lhdf5", "-o", out])
    return out


def build_ppf_exe():
    out = os.path.join(REPO, "ppf")
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", f"-I{INC}",
          os.path.join(CSRC, "ppf_main.c"), f"-L{PKG}", f"-Wl,-rpath,{PKG}", "-lppf", "-lppfio",
          "-lm", "-o", out])
    return out


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(REPO, "oracle"), "liboracle.so"])
    return os.path.join(REPO, "oracle", "liboracle.so")


def build_all(verbose=True):
    built = [build_oracle(), build_libppf(), build_libppfcal()]
    if have_hdf5():
        built.append(build_libppfio())
        if os.path.exists(os.path.join(CSRC, "ppf_main.c")):
            built.append(build_ppf_exe())
    elif verbose:
        print(f"HDF5 not found under {HDF5_DIR}: libppfio.so and ppf not built")
    if verbose:
        for b in built:
            print("built", os.path.relpath(b, REPO))
    return built


if __name__ == "__main__":
    build_all()
