"""entropical_amd -- MI355X-native Epanechnikov KDE hot path of lettis/entropical.

The compute path is the HIP library ``libentropical_kde.so`` (C ABI in
``include/entropical_kde.h``); this package is its Python mirror.
"""
from ._lib import EntropicalError  # noqa: F401
from .kde import (  # noqa: F401
    Context,
    DeviceArray,
    batch_autotune,
    batch_entropy,
    autoent,
    hestimate,
    kldiv,
    negs,
    read_xtc,
    mi_autotune,
    te_autotune,
    last_driver_ms,
    mi_matrix,
    te_matrix,
    batch_densities_dev,
    bin_column,
    cell_sort,
    combined_densities,
    combined_densities_dev,
    default_context,
    device_count,
    epa_convolution,
    prob_dens_coord_prep,
)

__all__ = [
    "EntropicalError",
    "Context",
    "combined_densities",
    "combined_densities_dev",
    "epa_convolution",
    "prob_dens_coord_prep",
    "bin_column",
    "cell_sort",
    "batch_densities_dev",
    "default_context",
    "device_count",
    "DeviceArray",
    "mi_matrix",
    "te_matrix",
    "batch_entropy",
    "last_driver_ms",
    "autoent",
    "hestimate",
    "kldiv",
    "negs",
    "read_xtc",
    "mi_autotune",
    "te_autotune",
    "batch_autotune",
]
