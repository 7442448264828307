"""MI355X-native photon-history engine with the C ABI of XMI-MSIM's
xmi_main_msim path (see include/xmimsim_amd.h).

    from xmimsim_amd import api, workloads
    prep = api.Prepared(workloads.single_layer_steel(10000))
    prep.compute_solid_angle()
    channels, var_red, brute, stats = prep.run(seed=1)
"""
from . import api, workloads  # noqa: F401

__all__ = ["api", "workloads"]
