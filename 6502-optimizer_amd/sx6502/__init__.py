"""sx6502 -- host side of the MI355X search core for the 6502 superoptimizer.

Modules:
  isa     instruction model, disassembler and assembly parser (pkg/inst, CLI parsers)
  core    ctypes binding of lib6502core.so (include/sx6502.h)
  search  pkg/search-shaped API on the device
  result  rule records, Go JSON and CUDA-tool JSONL writers
"""
from . import core, isa, result, search  # noqa: F401

__all__ = ["core", "isa", "result", "search"]
