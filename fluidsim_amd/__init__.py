"""fluidsim_amd -- MI355X-native FluidSim step pipeline.

The product is native code: HIP kernels for gfx950 behind the C ABI of
include/fluidsim_amd.h (lib/libfluidsim_amd.so) and the C++ host layer of
include/fluidsim_host.h (lib/libfluidsim_host.so, bin/fluidsim).  This
package builds them (`fluidsim_amd.build`), binds them with ctypes
(`fluidsim_amd._native`, `fluidsim_amd.sim`) and generates scenes in the
reference's file formats (`fluidsim_amd.scenes`).
"""
from ._native import FLUID, SOLID, VIRT, FsaError, FsaParams  # noqa: F401

__all__ = ["FLUID", "SOLID", "VIRT", "FsaError", "FsaParams"]
