"""pyfwi_amd: MI355X (gfx950) implementation of PyFWI's elastic propagator path.

Drop-in modules mirroring PyFWI v0.1.10:
  pyfwi_amd.wave_propagation  WavePropagator (forward modelling, gradient)
  pyfwi_amd.torchfwi          Fwi autograd function
  pyfwi_amd.grad_switcher     gradient parameterisations
  pyfwi_amd.acquisition       geometry, Source/Ricker, seismic_section
  pyfwi_amd.fwi_tools         Fdm, expand_model, PML profiles, CostFunction
  pyfwi_amd.processing        prepare_residual
The compute path is libpyfwi_hip.so (HIP kernels + C ABI, include/pyfwi_hip.h).
"""
__version__ = "0.1.0"
