"""MI355X-native force-and-step engine for Langevin dynamics of grafted
polyelectrolyte brushes (the hot path of wangshaoyun/MD_Brushes)."""
from .abi import (Batch, MD_FP32, MD_FP64, MD_NOISE_ETA, MD_NOISE_NONE, MD_NOISE_PHILOX, Engine, MDError,
                  MdParams, autotune_geometry, build_topology, derive, initialize_position,
                  initialize_velocity, load_library, replica_bytes)
from .configs import config, reference_defaults, replica

__all__ = ["Batch", "Engine", "MDError", "MdParams", "MD_FP32", "MD_FP64", "MD_NOISE_ETA", "MD_NOISE_NONE",
           "MD_NOISE_PHILOX", "autotune_geometry", "build_topology", "config", "derive", "initialize_position",
           "initialize_velocity", "load_library", "reference_defaults", "replica", "replica_bytes"]
