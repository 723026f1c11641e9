"""QueryNode: per-frame statistics of a box of the grid, for parameter studies
(src/simulation/scene/nodes/QueryNode.cpp, include/simulation/scene/nodes/QueryNode.h).

The box statistics are reduced on the device by efs_query_box; this module keeps the node's
host logic: the box snapping of SimSpaceNode::update (SimSpaceNode.cpp:17-25), the three modes
of QueryNode::update (QueryNode.cpp:86-126) and the CSV export (QueryNode.cpp:133-180).

The reference reads and writes global properties from update(): the frame counter
(SIMULATION_FRAMES), the physical gravity (GRAVITY_FORCE_PHYSICAL), PAUSED and
SIMULATION_RESET.  Here they live in a RunState the caller owns and passes in.
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from . import _abi
from ._abi import check

SINGLE_SIMULATION, CHANGE_GRAVITY, CHANGE_Y_POS = 0, 1, 2


@dataclass
class Metrics:
    """QueryNode.h:10-17."""
    max_position: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    min_position: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    avg_position: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    velocity: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    pressure: float = 0.0
    gravity: float = 0.0

    def csv(self) -> str:
        """operator<< of QueryNode.h:19-28 (ostream's default float format, %g)."""
        vals = list(self.max_position) + list(self.min_position) + list(self.avg_position) + list(self.velocity)
        vals += [self.pressure, self.gravity]
        return ",".join(_g(v) for v in vals)


def _g(v) -> str:
    return "%g" % float(np.float32(v))


@dataclass
class RunState:
    """The global properties QueryNode::update reads and sets."""
    frames: int = 0          # SIMULATION_FRAMES
    gravity: float = 9.81    # GRAVITY_FORCE_PHYSICAL
    paused: bool = False     # PAUSED
    reset: bool = False      # SIMULATION_RESET (set by the node; the caller clears it)


def _round_half_away(v):
    return math.copysign(math.floor(abs(v) + 0.5), v)


class QueryNode:
    """A query box in grid units (position, scale) with the reference's defaults
    (QueryNode.h:49-61)."""

    def __init__(self, position=(0.0, 0.0, 0.0), scale=(1.0, 1.0, 1.0), enabled=True, mode=SINGLE_SIMULATION, avg_frames=5,
                 delay_frames=20, change_value=0.05, iterations=200, only_fluid=True, name="QueryNode"):
        self.name = name
        self.position = np.asarray(position, np.float32).copy()
        self.scale = np.asarray(scale, np.float32).copy()
        self.enabled = bool(enabled)
        self.mode = int(mode)
        self.avg_frames = int(avg_frames)
        self.delay_frames = int(delay_frames)
        self.change_value = float(change_value)
        self.iterations = int(iterations)
        self.only_fluid = bool(only_fluid)
        self.history: List[Metrics] = []
        self._avg_history: List[Metrics] = []

    @classmethod
    def from_json(cls, node: dict):
        """QueryNode(const nlohmann::json) (QueryNode.cpp:21-29) with the position / scale the
        SimSpaceNode base reads."""
        return cls(position=node.get("position", [0, 0, 0])[:3], scale=node.get("scale", [1, 1, 1])[:3],
                   enabled=node.get("enabled", True), mode=node.get("mode", SINGLE_SIMULATION),
                   avg_frames=node.get("avgFrames", 5), delay_frames=node.get("delayFrames", 20),
                   change_value=node.get("changeValue", 0.05), iterations=node.get("iterations", 200),
                   only_fluid=node.get("onlyFluid", True), name=node.get("name", "QueryNode"))

    def snap(self, dims):
        """SimSpaceNode::update: position = clamp(round(position), 0, dim), scale =
        clamp(round(scale), 1, dim - position) (round half away from zero, as glm::round)."""
        dim = np.asarray(dims, np.float32)
        pos = np.array([_round_half_away(float(v)) for v in self.position], np.float32)
        self.position = np.clip(pos, 0.0, dim).astype(np.float32)
        scl = np.array([_round_half_away(float(v)) for v in self.scale], np.float32)
        self.scale = np.minimum(np.maximum(scl, 1.0), dim - self.position).astype(np.float32)

    def box(self):
        """(lo, hi) of the loops of QueryNode.cpp:58-60 before clipping to [0, n - 1)."""
        lo = [int(v) for v in self.position]
        hi = [int(np.float32(p + s)) for p, s in zip(self.position, self.scale)]
        return lo, hi

    def measure(self, sim, gravity=0.0) -> Metrics:
        """One frame's statistics of the box, reduced on the device (efs_query_box)."""
        lo, hi = self.box()
        r = _abi.QueryResult()
        check(sim.lib.efs_query_box(sim.h, (ctypes.c_int32 * 3)(*lo), (ctypes.c_int32 * 3)(*hi), int(self.only_fluid),
                                    ctypes.byref(r)))
        return Metrics(np.array(r.max_pos, np.float32), np.array(r.min_pos, np.float32), np.array(r.avg_pos, np.float32),
                       np.array(r.vel, np.float32), float(r.pressure), float(np.float32(gravity)))

    def update(self, sim, state: RunState, dims=None):
        """QueryNode::update (QueryNode.cpp:31-126) for one frame of `sim`."""
        self.snap(dims if dims is not None else sim.dims)
        if not self.enabled or state.paused:
            return
        m = self.measure(sim, state.gravity)
        self.record(m, state)

    def record(self, m: Metrics, state: RunState):
        """The mode logic of QueryNode.cpp:86-126 for one frame's metrics."""
        if self.mode == SINGLE_SIMULATION:
            self.history.append(m)
            return
        if state.frames < self.delay_frames:
            return
        if state.frames >= self.delay_frames + self.avg_frames:
            n = np.float32(len(self._avg_history))
            with np.errstate(divide="ignore", invalid="ignore"):
                acc = Metrics()
                for h in self._avg_history:
                    acc.max_position = (acc.max_position + h.max_position).astype(np.float32)
                    acc.min_position = (acc.min_position + h.min_position).astype(np.float32)
                    acc.avg_position = (acc.avg_position + h.avg_position).astype(np.float32)
                    acc.velocity = (acc.velocity + h.velocity).astype(np.float32)
                    acc.pressure = float(np.float32(acc.pressure) + np.float32(h.pressure))
                    acc.gravity = float(np.float32(acc.gravity) + np.float32(h.gravity))
                acc.max_position = acc.max_position / n
                acc.min_position = acc.min_position / n
                acc.avg_position = acc.avg_position / n
                acc.velocity = acc.velocity / n
                acc.pressure = float(np.float32(acc.pressure) / n)
                acc.gravity = float(np.float32(acc.gravity) / n)
            self.history.append(acc)
            self._avg_history.clear()
            state.reset = True
            if self.mode == CHANGE_GRAVITY:
                state.gravity = float(np.float32(state.gravity) + np.float32(self.change_value))
            elif self.mode == CHANGE_Y_POS:
                self.position = self.position + np.array([0.0, self.change_value, 0.0], np.float32)
            self.iterations -= 1
        else:
            self._avg_history.append(m)
        if self.iterations == 0:
            state.paused = True

    def reset(self):
        """QueryNode::reset."""
        self.history.clear()

    def export_csv(self, path, dt, cell_size, dims, properties: Dict[str, object] = None):
        """QueryNode::exportData (QueryNode.cpp:133-180): a header, then one line per recorded
        frame; the first line also carries dt, cell size, the grid size and the given
        properties (name -> int / float / bool, written as std::to_string does)."""
        props = properties or {}
        with open(path, "w") as f:
            f.write("frame,max_pos_x,max_pos_y,max_pos_z,min_pos_x,min_pos_y,min_pos_z,avg_pos_x,avg_pos_y,avg_pos_z,"
                    "vel_x,vel_y,vel_z,pressure,gravity,dt,cell_size,grid_width,grid_height,grid_depth")
            for k in props:
                f.write("," + k)
            f.write("\n")
            for i, m in enumerate(self.history):
                f.write(f"{i},{m.csv()}")
                if i == 0:
                    f.write(f",{_g(dt)},{_g(cell_size)},{int(dims[0])},{int(dims[1])},{int(dims[2])}")
                    for v in props.values():
                        f.write("," + _to_string(v))
                f.write("\n")


def _to_string(v) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return "%f" % float(np.float32(v))
    return "nan"
