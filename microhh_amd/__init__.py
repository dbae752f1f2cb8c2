"""MI355X (gfx950) dynamical core of MicroHH: 2i5/2nd-order advection,
Smagorinsky/constant diffusion, the 2nd-order FFT + tridiagonal pressure
projection and the RK3 time loop, behind the C ABI in include/.

The compute path is libmicrohh_amd.so (HIP kernels + C++ host layer); the
Python modules here only bind it (abi, model) and build it (build)."""
__all__ = ["abi", "model", "build"]
