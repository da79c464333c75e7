"""Builds the Fortran bridge and driver with flang, linked to libfp4.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
DRIVER = os.path.join(HERE, "fp4_driver")


def build(verbose=False):
    objdir = os.path.join(PKG, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O2", "-module-dir", objdir]
    mod_o = os.path.join(objdir, "fp4_processes.o")
    drv_o = os.path.join(objdir, "fp4_driver.o")
    cmds = [
        [FLANG, *common, "-c", os.path.join(HERE, "fp4_processes.f90"), "-o", mod_o],
        [FLANG, *common, "-c", os.path.join(HERE, "fp4_driver.f90"), "-o", drv_o],
        [FLANG, mod_o, drv_o, "-o", DRIVER + ".tmp", "-L" + PKG, "-lfp4", "-Wl,-rpath,$ORIGIN/..",
         "-Wl,-rpath,/opt/rocm/lib"],
    ]
    for c in cmds:
        if verbose:
            print(" ".join(c))
        subprocess.run(c, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
