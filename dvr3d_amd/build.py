"""In-tree build of the native parts (run on the CPU; hipcc cross-compiles).

  dvr3d_amd/libdvr3d_amd.so      HIP kernels + C ABI (hipcc, gfx950)
  dvr3d_amd/libdvr3d_host.so     Fortran host set-up with bind(C) entry points
  dvr3d_amd/host/dvr3drjz_amd    drop-in executable (Fortran host + POTV)
  dvr3d_amd/host/rotlev3_amd     drop-in ROTLEV3 (Fortran host)
  dvr3d_amd/host/dipole3_amd     drop-in DIPOLE3 (Fortran host + example DIPD)
  dvr3d_amd/host/xpect3_amd      drop-in XPECT3 (Fortran host + example PROP)

Every call recompiles.  The oracle (reference executable) is built separately
by oracle/Makefile.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
HIP_SOURCES = ["k_grid.hip", "k_jacobi.hip", "k_syrk.hip", "k_syev.hip", "k_syev_wg.hip", "k_gemm.hip", "k_ldl.hip", "k_dav.hip", "k_mlp.hip", "dvr_davidson.hip", "dvr_mlp.hip", "dvr_refine.hip",
               "k_syev_lds.hip", "k_twostage.hip", "dvr_nfmain.hip", "dvr_rotlev3.hip", "dvr_dipole3.hip", "dvr_xpect3.hip", "dvr_noalloc.hip", "dvr_capi.hip"]
ARCH = os.environ.get("DVR_OFFLOAD_ARCH", "gfx950")
# MFMA accumulators in ordinary VGPRs: without this the compiler keeps them
# in AGPRs and copies them to VGPRs and back around every k slice of the GEMM
# loops (with a wait for the MFMA pipeline to drain each time)
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form=1"]


def _run(cmd, cwd=None, verbose=False):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run(["hipcc", *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose=verbose)
        objs.append(obj)
    out = os.path.join(HERE, "libdvr3d_amd.so")
    tmp = out + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp], verbose=verbose)
    os.replace(tmp, out)
    return out


def build_host(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    mod = os.path.join(BUILD, "mod")
    os.makedirs(mod, exist_ok=True)
    fc = ["amdflang", "-O2", "-fPIC", "-module-dir", mod]
    setup_o = os.path.join(BUILD, "dvrh_setup.o")
    api_o = os.path.join(BUILD, "dvr3d_amd_api.o")
    potv_o = os.path.join(BUILD, "potv_example.o")
    noise_o = os.path.join(BUILD, "potv_noise.o")
    main_o = os.path.join(BUILD, "dvr3drjz_amd.o")
    _run(fc + ["-c", os.path.join(HOST, "dvrh_setup.f90"), "-o", setup_o], verbose=verbose)
    _run(fc + ["-c", os.path.join(HOST, "dvr3d_amd_api.f90"), "-o", api_o], verbose=verbose)
    _run(fc + ["-c", os.path.join(HOST, "potv_example.f90"), "-o", potv_o], verbose=verbose)
    _run(["gcc", "-O2", "-fPIC", "-c", os.path.join(HOST, "potv_noise.c"), "-o", noise_o],
         verbose=verbose)
    _run(fc + ["-c", os.path.join(HOST, "dvr3drjz_amd.f90"), "-o", main_o], verbose=verbose)
    grid_o = os.path.join(BUILD, "dvrh_potgrid.o")
    _run(fc + ["-c", os.path.join(HOST, "dvrh_potgrid.f90"), "-o", grid_o], verbose=verbose)
    lib = os.path.join(HERE, "libdvr3d_host.so")
    # the host library carries the example POTV, so that callers of
    # dvrh_potential_grid see the surface the executable is linked with
    _run(["amdflang", "-shared", setup_o, grid_o, potv_o, noise_o, "-lm", "-o", lib + ".tmp"],
         verbose=verbose)
    os.replace(lib + ".tmp", lib)
    exe = os.path.join(HOST, "dvr3drjz_amd")
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    _run(["amdflang", main_o, setup_o, api_o, potv_o, noise_o, "-lm", "-L" + HERE, "-ldvr3d_amd",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + rocm_lib, "-o", exe], verbose=verbose)
    rot_o = os.path.join(BUILD, "rotlev3_amd.o")
    _run(fc + ["-c", os.path.join(HOST, "rotlev3_amd.f90"), "-o", rot_o], verbose=verbose)
    rot = os.path.join(HOST, "rotlev3_amd")
    _run(["amdflang", rot_o, api_o, "-L" + HERE, "-ldvr3d_amd", "-Wl,-rpath,$ORIGIN/..",
          "-Wl,-rpath," + rocm_lib, "-o", rot], verbose=verbose)
    dip_o = os.path.join(BUILD, "dipole3_amd.o")
    dipd_o = os.path.join(BUILD, "dipd_example.o")
    _run(fc + ["-c", os.path.join(HOST, "dipole3_amd.f90"), "-o", dip_o], verbose=verbose)
    _run(fc + ["-c", os.path.join(HOST, "dipd_example.f90"), "-o", dipd_o], verbose=verbose)
    dip = os.path.join(HOST, "dipole3_amd")
    _run(["amdflang", dip_o, api_o, dipd_o, "-L" + HERE, "-ldvr3d_amd", "-Wl,-rpath,$ORIGIN/..",
          "-Wl,-rpath," + rocm_lib, "-o", dip], verbose=verbose)
    xp_o = os.path.join(BUILD, "xpect3_amd.o")
    prop_o = os.path.join(BUILD, "prop_example.o")
    _run(fc + ["-c", os.path.join(HOST, "xpect3_amd.f90"), "-o", xp_o], verbose=verbose)
    _run(fc + ["-c", os.path.join(HOST, "prop_example.f90"), "-o", prop_o], verbose=verbose)
    xp = os.path.join(HOST, "xpect3_amd")
    _run(["amdflang", xp_o, api_o, prop_o, "-L" + HERE, "-ldvr3d_amd", "-Wl,-rpath,$ORIGIN/..",
          "-Wl,-rpath," + rocm_lib, "-o", xp], verbose=verbose)
    return lib, exe


def build(verbose=False):
    build_hip(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
