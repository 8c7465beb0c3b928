"""Build the HIP library libimhd.so in-tree for gfx950 (hipcc, no CMake), and, where a
Fortran compiler is present, the iso_c_binding interface and the example Fortran drivers
that reach it through the solver shim modules (isothermal_mhd_amd/fortran/)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libimhd.so")
OBJ = os.path.join(HERE, "_build")
# (source, extra flags): the FP32 kernels may contract to FMA; everything else may not.
SOURCES = [("vl_kernels_f64.hip", ["-ffp-contract=off"]),
           ("vl_kernels_f32.hip", ["-ffp-contract=fast"]),
           ("tvd_kernels_f64.hip", ["-ffp-contract=off"]),
           ("tvd_kernels_f32.hip", ["-ffp-contract=fast"]),
           ("aux_kernels.hip", ["-ffp-contract=off"]),
           ("forcing_kernels.hip", ["-ffp-contract=off"]),
           ("spectra_kernels.hip", ["-ffp-contract=off"]),
           ("ic_kernels.hip", ["-ffp-contract=off"]),
           ("imhd_api.cpp", ["-ffp-contract=off"])]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# FP64 sources: -ffp-contract=off (no FMA contraction), so the FP64 path reproduces the
# reference's operation order exactly.  Packed-FP32 instructions are disabled everywhere.
FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build(verbose=False, jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    procs, objs = [], []
    for s, extra in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, s + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + extra + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for s, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        txt = "\n".join(l for l in txt.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {s}\n{txt}\n")
    if failed:
        raise SystemExit("libimhd build failed")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT] + objs + ["-L/opt/rocm/lib", "-lhipfft",
                                                                            "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise SystemExit("link failed:\n" + r.stdout + r.stderr)
    return OUT


FORTRAN = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
# example configurations: (name, constants module)
F_EXAMPLES = [("vl", "example_constants_vl.f90"), ("tvd", "example_constants_tvd.f90"),
              ("vl_drive", "example_constants_vl_drive.f90")]


def fortran_compiler():
    return FLANG if os.path.exists(FLANG) else None


def build_fortran(verbose=False):
    """imhd_api.f90 once, then per example configuration its constants/grid modules, the
    shim modules and the example driver -> fortran/_build/imhd_example_<cfg> (rpath to
    libimhd.so).  Returns the executables built, or [] without a Fortran compiler."""
    fc = fortran_compiler()
    if fc is None:
        return []
    exes = []
    for name, consts in F_EXAMPLES:
        bdir = os.path.join(FBUILD, name)
        os.makedirs(bdir, exist_ok=True)
        srcs = [os.path.join(FORTRAN, "example", consts), os.path.join(FORTRAN, "example", "example_grid_data.f90"),
                os.path.join(FORTRAN, "imhd_api.f90"), os.path.join(FORTRAN, "imhd_shim.f90"),
                os.path.join(FORTRAN, "imhd_shim_turb.f90"),
                os.path.join(FORTRAN, "example", "imhd_example_driver.f90")]
        objs = []
        for src in srcs:
            obj = os.path.join(bdir, os.path.basename(src) + ".o")
            cmd = [fc, "-O2", "-c", "-J", bdir, "-I", bdir, src, "-o", obj]
            if verbose:
                print(" ".join(cmd))
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
            if r.returncode != 0:
                raise SystemExit(f"Fortran build failed ({src}):\n" + r.stdout + r.stderr)
            objs.append(obj)
        exe = os.path.join(FBUILD, f"imhd_example_{name}")
        cmd = [fc, "-o", exe] + objs + ["-L" + HERE, "-limhd", "-Wl,-rpath,$ORIGIN/../.."]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
        if r.returncode != 0:
            raise SystemExit("Fortran link failed:\n" + r.stdout + r.stderr)
        exes.append(exe)
    return exes


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
