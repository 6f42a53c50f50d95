"""seisproc_amd -- MI355X-native Kirchhoff prestack time migration (PSTM).

A drop-in for SeisProc's ``processors/migration`` path: the same migration
job configuration, velocity/anisotropy model objects and
``BaseProcessor.process(SeismicData)`` surface, with the migration itself in
hand-written HIP kernels for gfx950 (``libpstm_hip.so``) reached through a thin
C ABI (``include/pstm_abi.h``) via ctypes.  PyTorch-ROCm tensors carry the
device buffers and streams.

Layout
  abi_types.py          ctypes mirror of the C ABI structs/enums
  build.py              hipcc build of libpstm_hip.so (gfx950)
  csrc/                 HIP kernels and the extern "C" entry points
  models/               SeismicData, VelocityModel, AnisotropyModel,
                        OutputGrid/MigrationConfig, MigrationJobConfig
  processors/           BaseProcessor, registry, the single HIP kernel backend
  processors/migration/ engine, processor, migrators, autograd, LSM
  synthetic.py          seeded synthetic surveys (Ricker events at exact DSR times)
"""

__version__ = "0.1.0"
