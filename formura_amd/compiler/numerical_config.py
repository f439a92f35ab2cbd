"""Numerical configuration (`.yaml`) of a Formura program.

Mirrors src/Formura/NumericalConfig.hs:17-53: keys `intra_node_shape`,
`mpi_grid_shape`, `temporal_blocking_interval` (default 1), `monitor_interval`
(default 1), `initial_walls`, `wall_inverted`, `option_strings`; `nbu<axis>N`
option strings give the nanoblock unroll per axis (default 1).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import yaml


@dataclass
class NumericalConfig:
    intra_node_shape: List[int] = field(default_factory=list)
    mpi_grid_shape: List[int] = field(default_factory=list)
    temporal_blocking_interval: int = 1
    monitor_interval: int = 1
    initial_walls: Dict[str, List[int]] = field(default_factory=dict)
    wall_inverted: Optional[bool] = None
    option_strings: List[str] = field(default_factory=list)

    @staticmethod
    def from_yaml_text(text: str) -> "NumericalConfig":
        d = yaml.safe_load(text) or {}
        nc = NumericalConfig()
        for k in ("intra_node_shape", "mpi_grid_shape", "temporal_blocking_interval",
                  "monitor_interval", "initial_walls", "wall_inverted", "option_strings"):
            if k in d:
                setattr(nc, k, d[k])
        nc.intra_node_shape = [int(v) for v in nc.intra_node_shape]
        nc.mpi_grid_shape = [int(v) for v in nc.mpi_grid_shape] or [1] * len(nc.intra_node_shape)
        nc.temporal_blocking_interval = int(nc.temporal_blocking_interval)
        nc.monitor_interval = int(nc.monitor_interval)
        nc.option_strings = [str(s) for s in nc.option_strings]
        return nc

    @staticmethod
    def from_file(path: str) -> "NumericalConfig":
        with open(path, encoding="utf-8") as f:
            return NumericalConfig.from_yaml_text(f.read())

    def global_extents(self) -> List[int]:
        """`NX` etc.: intra_node_shape * mpi_grid_shape (Parser.hs:488)."""
        return [i * g for i, g in zip(self.intra_node_shape, self.mpi_grid_shape)]

    def nbu_size(self, axis: str) -> int:
        """Nanoblock unroll `nbu<axis>N` (NumericalConfig.hs:48-53)."""
        kwd = "nbu" + axis
        for opt in self.option_strings:
            if opt.startswith(kwd):
                return int(opt[len(kwd):])
        return 1

    def monitor_interval2(self) -> int:
        """Steps per `Formura_Forward`: the monitor interval rounded up to a
        multiple of 2T (MPICxx/Translate.hs:1108-1113)."""
        two_t = 2 * self.temporal_blocking_interval
        m = self.monitor_interval
        while m % two_t != 0:
            m += 1
        return m
