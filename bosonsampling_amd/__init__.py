"""MI355X-native beam-splitter site update for lossy boson-sampling MPOs.

Hot path (C ABI: include/bsmpo.h, library: libbsmpo.so, gfx950 HIP):
  Fock BS elements -> charge-blocked sector contraction on MFMA -> FP64
  sector SVD (Gram + Jacobi) -> global top-chi truncation -> scatter.

Python modules:
  engine        device-resident MPO, update driver, observables, init
  cuda_kernels  drop-in for the reference's src/cuda_kernels.py API
  worker        drop-in for the reference's Worker (ExperimentInit/Simulate)
  mesh          Haar/Clements meshes and the reference mesh parameters
  build         in-tree hipcc build of libbsmpo.so
"""
__version__ = "0.1.0"
