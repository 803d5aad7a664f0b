"""livox-warp_amd: the scan-to-map odometry hot path of livox-warp, rebuilt for MI355X (gfx950).

Import with importlib (the directory name carries a hyphen):

    lw = importlib.import_module("livox-warp_amd")
    odo = lw.Odometry(lw.PoseTable())

Submodules: odom (the drop-in Odometry), batch (many sequences stepped together), engine (one context driven directly), pose_table (the pose sink),
imap (the integration map the frames feed, with free-space carving), worker (the odometry on its own thread
and stream), lvxr (LVXR recordings decoded on the device), mapgrid (the prior map's distance-field grid),
localize (the coarse stages of global localisation over it), loc (batched global localisation), prior_map
(bulk voxelisation of a prior map), prior (the learned pose-increment prior), synth (seeded synthetic scenes, scans and walks), _abi (ctypes bindings of
liblwamd.so).
"""

from . import _abi, batch, engine, imap, loc, localize, lvxr, mapgrid, odom, pose_table, prior, prior_map, synth, worker  # noqa: F401
from .batch import BatchOdometry  # noqa: F401
from .engine import Engine  # noqa: F401
from .imap import IntegrationMap  # noqa: F401
from .localize import CoarseSearch, GlobalLocalizer  # noqa: F401
from .mapgrid import PriorGrid  # noqa: F401
from .odom import Odometry  # noqa: F401
from .pose_table import PoseTable  # noqa: F401
from .prior_map import VoxTable, voxel_downsample  # noqa: F401
from .worker import OdomWorker  # noqa: F401

__all__ = ["Odometry", "BatchOdometry", "batch", "Engine", "PoseTable", "IntegrationMap", "OdomWorker", "odom", "engine", "imap", "loc",
           "pose_table", "prior", "synth", "worker", "lvxr", "mapgrid", "localize", "PriorGrid", "CoarseSearch",
           "GlobalLocalizer", "prior_map", "VoxTable", "voxel_downsample"]
