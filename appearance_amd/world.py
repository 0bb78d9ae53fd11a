"""VisibleWorldAction wire formats (crates/appearance-world/src/visible_world_action.rs).

Each action is a type tag plus a POD payload; `to_bytes()` gives exactly the
bytes VisibleWorldActionType::as_bytes produces, and `from_ty_and_bytes`
mirrors the reference's decoder (visible_world_action.rs:87-96).
"""
from __future__ import annotations

import struct
import uuid as _uuid
from dataclasses import dataclass, field

import numpy as np

CAMERA_UPDATE, SPAWN_MODEL, TRANSFORM_MODEL, DESTROY_MODEL, CLEAR = 0, 1, 2, 3, 4
DATA_SIZE = {CAMERA_UPDATE: 80, SPAWN_MODEL: 336, TRANSFORM_MODEL: 80, DESTROY_MODEL: 16, CLEAR: 4}
MUST_SYNC = {CAMERA_UPDATE: False, SPAWN_MODEL: True, TRANSFORM_MODEL: False, DESTROY_MODEL: True,
             CLEAR: True}


def _mat_bytes(m) -> bytes:
    m = np.asarray(m, dtype=np.float32)
    assert m.shape == (4, 4)
    return m.T.astype("<f4").tobytes()  # glam Mat4 is column-major


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """World transform of a camera at `eye` looking at `target` (it looks down -Z)."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    up2 = np.cross(back, right)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, up2, back, eye
    return m.astype(np.float32)


@dataclass
class CameraUpdate:
    transform: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    fov: float = 60.0
    near: float = 0.01
    far: float = 1000.0
    ty: int = CAMERA_UPDATE

    def to_bytes(self) -> bytes:
        return _mat_bytes(self.transform) + struct.pack("<fffI", self.fov, self.near, self.far, 0)


@dataclass
class SpawnModel:
    asset_path: str
    transform: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))
    entity_uuid: _uuid.UUID = field(default_factory=_uuid.uuid4)
    ty: int = SPAWN_MODEL

    def to_bytes(self) -> bytes:
        p = self.asset_path.encode()
        assert len(p) < 256
        return _mat_bytes(self.transform) + self.entity_uuid.bytes + p + b"\0" * (256 - len(p))


@dataclass
class TransformModel:
    entity_uuid: _uuid.UUID
    transform: np.ndarray
    ty: int = TRANSFORM_MODEL

    def to_bytes(self) -> bytes:
        return _mat_bytes(self.transform) + self.entity_uuid.bytes


@dataclass
class DestroyModel:
    entity_uuid: _uuid.UUID
    ty: int = DESTROY_MODEL

    def to_bytes(self) -> bytes:
        return self.entity_uuid.bytes


@dataclass
class Clear:
    value: int = 0
    ty: int = CLEAR

    def to_bytes(self) -> bytes:
        return struct.pack("<I", self.value)


def from_ty_and_bytes(ty: int, data: bytes):
    if ty not in DATA_SIZE or len(data) != DATA_SIZE[ty]:
        raise ValueError(f"bad world action ty={ty} len={len(data)}")
    if ty == CAMERA_UPDATE:
        m = np.frombuffer(data[:64], "<f4").reshape(4, 4).T.copy()
        fov, near, far, _ = struct.unpack("<fffI", data[64:80])
        return CameraUpdate(m, fov, near, far)
    if ty == SPAWN_MODEL:
        m = np.frombuffer(data[:64], "<f4").reshape(4, 4).T.copy()
        path = data[80:336].split(b"\0", 1)[0].decode()
        return SpawnModel(path, m, _uuid.UUID(bytes=data[64:80]))
    if ty == TRANSFORM_MODEL:
        m = np.frombuffer(data[:64], "<f4").reshape(4, 4).T.copy()
        return TransformModel(_uuid.UUID(bytes=data[64:80]), m)
    if ty == DESTROY_MODEL:
        return DestroyModel(_uuid.UUID(bytes=data))
    return Clear(struct.unpack("<I", data)[0])


@dataclass
class StartRenderData:
    """The render-node job (crates/appearance-render-loop/src/host.rs:85-93),
    20 bytes: width, height, row_start, row_end, frame_idx (u32 each)."""
    width: int
    height: int
    row_start: int
    row_end: int
    frame_idx: int = 0

    def to_bytes(self) -> bytes:
        return struct.pack("<5I", self.width, self.height, self.row_start, self.row_end, self.frame_idx)

    @classmethod
    def from_bytes(cls, b: bytes) -> "StartRenderData":
        if len(b) != 20:
            raise ValueError(f"StartRenderData carries 20 bytes, got {len(b)}")
        return cls(*struct.unpack("<5I", b))
