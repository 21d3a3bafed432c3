"""Workload presets: generator settings + placer parameters per configuration.

mesh_c1      : 576-cell 24x24 mesh, 32x32 bins, rho_t 1.0, gamma 0.5 bin widths (fixed)
stdcell_200k : 200k cells, 512x512 bins, rho_t 0.9
stdcell_2m   : 2M cells, 1024x1024 bins, rho_t 0.9 (the measured workload); cells stored
               in random order, as a LEF/DEF netlist may list its components
batch_64     : 64 instances, 5k-150k cells, 128/256/512 bins
learned_1m   : 1M cells, 1024x1024 bins (learned-step form)
"""
from __future__ import annotations

from . import api

CONFIGS = {
    "mesh_c1": dict(gen=dict(kind="mesh"), bins=32, target_density=1.0, gamma_bins=0.5,
                    gamma_schedule=0),
    "stdcell_200k": dict(gen=dict(kind="stdcell", n_movable=200_000), bins=512,
                         target_density=0.9, gamma_bins=4.0, gamma_schedule=1),
    "stdcell_2m": dict(gen=dict(kind="stdcell", n_movable=2_000_000, shuffle_storage=1), bins=1024,
                       target_density=0.9, gamma_bins=4.0, gamma_schedule=1),
    "learned_1m": dict(gen=dict(kind="stdcell", n_movable=1_000_000), bins=1024,
                       target_density=0.9, gamma_bins=4.0, gamma_schedule=1),
}


def bins_for_cells(n: int) -> int:
    """Bin grid matched to instance size (batched form): 128/256/512.

    The 200k-cell workload runs at 512x512 bins, 0.76 cells per bin; the
    thresholds are the geometric midpoints of that density between grid sizes
    (0.76 * 128 * 256 ~ 25k and 0.76 * 256 * 512 ~ 100k cells), so every
    instance sees 0.3-1.5 cells per bin.
    """
    if n < 25_000:
        return 128
    if n < 100_000:
        return 256
    return 512


def placer_params(cfg: dict, **kw) -> api.EpParams:
    b = cfg["bins"]
    p = api.default_params(bins_x=b, bins_y=b, target_density=cfg["target_density"],
                           gamma_bins=cfg["gamma_bins"], gamma_schedule=cfg["gamma_schedule"])
    for k, v in kw.items():
        setattr(p, k, v)
    return p
