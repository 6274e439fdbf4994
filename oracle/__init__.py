"""The oracle -- TEST INFRASTRUCTURE ONLY.

CPU restatements of the reference renderers, used by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py as the checker. The product (the us3d package and libus3d.so) never imports, links or runs
anything here.

  ug_oracle.cpp / ug.py  UltraGauss mapper + CUDA-semantics shader, FP32 (canonical, bit-exact keys) and FP64
  fallback.py            the reference PyTorch fallback renderer (pipeline_god_file.py:556-603), FP32/FP64
  ur_oracle.cpp / ur.py  UltraG-Ray mapper + B-mode shader
"""
