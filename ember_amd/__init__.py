"""ember_amd: MI355X-native batched kernels for Ember's split-step hot path.

- chemistry production stage (SourceSystemQSS / SourceSystemCVODE semantics),
  one wavefront per grid point, persistent work queue;
- diffusion stage (DiffusionSystem + TridiagonalIntegrator), partitioned
  tridiagonal solves over Ember's point-major state.

The C ABI is include/ember_amd.h; ``ember_amd.capi`` is its Python binding.
"""
__version__ = "0.1.0"
