"""MI355X batched-sweep engine (HIP kernels behind a C ABI)."""
