"""Execution strategies and context (butterfly/utils/context.mojo:12-292).

The reference's strategies 0..7 are its CPU kernels; this build adds
ExecutionStrategy.GPU = 8, selectable exactly like the others
(`ctx.execution_strategy = ExecutionStrategy.GPU`, or `execution_strategy=8`
in a Butterfly config file, utils/context.mojo:199-203).  The GPU knobs are
read through the same key=value config pattern (utils/config.mojo:16-55).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, replace
from typing import Dict, Optional


class ExecutionStrategy:
    SCALAR = 0
    SCALAR_PARALLEL = 1
    SIMD = 2
    SIMD_PARALLEL = 3
    GRID = 4
    GRID_PARALLEL = 5
    GRID_FUSED = 6
    GRID_PARALLEL_FUSED = 7
    GPU = 8

    _names = ["SCALAR", "SCALAR_PARALLEL", "SIMD", "SIMD_PARALLEL", "GRID", "GRID_PARALLEL",
              "GRID_FUSED", "GRID_PARALLEL_FUSED", "GPU"]

    @staticmethod
    def from_int(value: int) -> int:
        if not is_valid_execution_strategy(value):
            raise ValueError(f"Unknown execution strategy: {value}")
        return value


def is_valid_execution_strategy(value: int) -> bool:
    """utils/context.mojo:43-44, extended by one value for the GPU."""
    return ExecutionStrategy.SCALAR <= value <= ExecutionStrategy.GPU


def execution_strategy_name(value: int) -> str:
    return ExecutionStrategy._names[value] if is_valid_execution_strategy(value) else "UNKNOWN"


GPU_MODES = {"fused": 0, "unfused": 1, "block": 2, "auto": 3}


@dataclass
class ExecContext:
    """Runtime knobs (utils/context.mojo:67-116) plus the GPU strategy's own."""

    threads: int = 0
    quantum_simd_parallel_chunks: int = 0
    parallel_flat: bool = False
    simd_use_specialized_h: bool = True
    simd_use_specialized_p: bool = True
    simd_use_specialized_cp: bool = True
    simd_use_specialized_x: bool = True
    simd_use_specialized_cx: bool = True
    simd_use_specialized_ry: bool = True
    simd_use_specialized_cry: bool = True
    grid_use_parallel: bool = True
    grid_col_bits_min: int = 3
    grid_col_bits_slack: int = 3
    validate_circuit: bool = True
    execution_strategy: int = ExecutionStrategy.GPU
    # GPU strategy knobs
    gpu_mode: str = "fused"      # fused | unfused | block | auto
    gpu_tile_bits: int = 0       # 0 = tuned default
    gpu_low_bits: int = 0
    gpu_fuse_k: int = 0          # block width for gpu_mode=block (0 = 5)
    gpu_max_gates_per_pass: int = 0
    gpu_renorm_every: int = 0
    measure_seed: int = -1       # apply_measure's configured seed (state.mojo:509); -1 = clock

    def copy(self) -> "ExecContext":
        return replace(self)

    @staticmethod
    def from_config(cfg: Dict[str, str]) -> "ExecContext":
        """Build from key=value pairs (utils/context.mojo:118-292 pattern)."""
        ctx = ExecContext()
        ints = ["threads", "quantum_simd_parallel_chunks", "grid_col_bits_min", "grid_col_bits_slack",
                "gpu_tile_bits", "gpu_low_bits", "gpu_fuse_k", "gpu_max_gates_per_pass", "gpu_renorm_every",
                "measure_seed"]
        bools = ["parallel_flat", "simd_use_specialized_h", "simd_use_specialized_p", "simd_use_specialized_cp",
                 "simd_use_specialized_x", "simd_use_specialized_cx", "simd_use_specialized_ry",
                 "simd_use_specialized_cry", "grid_use_parallel", "validate_circuit"]
        for k in ints:
            if k in cfg:
                setattr(ctx, k, int(cfg[k]))
        for k in bools:
            if k in cfg:
                setattr(ctx, k, cfg[k].strip().lower() in ("1", "true", "yes", "on"))
        if "execution_strategy" in cfg:
            ctx.execution_strategy = ExecutionStrategy.from_int(int(cfg["execution_strategy"]))
        if "gpu_mode" in cfg:
            if cfg["gpu_mode"] not in GPU_MODES:
                raise ValueError(f"unknown gpu_mode {cfg['gpu_mode']}")
            ctx.gpu_mode = cfg["gpu_mode"]
        return ctx


def load_config(path: Optional[str] = None) -> Dict[str, str]:
    """Parse a Butterfly key=value config (utils/config.mojo:16-55); the
    default path comes from BUTTERFLY_CONFIG_PATH, else butterfly.conf."""
    path = path or os.environ.get("BUTTERFLY_CONFIG_PATH", "butterfly.conf")
    out: Dict[str, str] = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line or "=" not in line:
                continue
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out
