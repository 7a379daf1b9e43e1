"""vcbct: MI355X-native virtual cone-beam CT chain (material-based polychromatic forward projector + FDK).

The compute runs in hand-written gfx950 HIP kernels behind the C ABI of include/vcbct.h (libvcbct.so);
this package mirrors the reference's operator interface on top of it:

  projector  initializeGPU / reprojectionOnGPU / terminateGPU, plus log-projection and adjoint
  fdk        prepareConstMap / filterSinogram / reconstructImage
  io         the reference's txt, flexmap, slice-stack and spectrum/xraydb formats
  autograd   torch.autograd.Function for the differentiable projector
  batch      many phantoms in one persistent launch
  phantoms   seeded synthetic inputs in the reference formats
"""
from ._lib import VcbctError, lib  # noqa: F401

__all__ = ["VcbctError", "lib"]
