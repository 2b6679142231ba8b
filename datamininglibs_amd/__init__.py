"""MI355X-native tree-ensemble build + predict (random forests and extra
trees) behind a C ABI (include/dte.h).  `capi` binds the library, `synth`
generates the benchmark workloads."""
