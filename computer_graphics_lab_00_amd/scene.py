"""Scene-file input: the fluids of a reference scene file, ready for the solver.

Reads the JSON scene format of src/scene/scene_io.cpp: `fluids[]` (:562-585, defaults of
scene::Fluid, src/scene/scene.h:186-216), `objects[]` (:697-717: parent, position,
rotation [x, y, z, w], scale, visible, and the `fluid` / `rigidBody` / `forceField`
components), `rigidBodies[]` (:548-560) and `forceFields[]` (:607-616). For each fluid it
builds what gfx::FluidSimulator hands to physics::deriveFluidParams
(src/gfx/fluid.cpp:507-509): the settings, the emitter world matrix (the world matrix of
the object carrying the fluid, Scene::worldMatrix, src/scene/scene.cpp:413-419, with
Transform::matrix, :74-78), the colliders of visible shape rigid bodies
(physics::collectShapeColliders, src/physics/fluid_sph.cpp:102-126, with
scene::colliderPose, scene.cpp:254-269) and the visible force fields
(physics::collectForceFields, src/physics/force_field.h:72-92).

Matrix arithmetic is FP32 in glm's evaluation order (mat4 products summed left to right,
glm::mat4_cast's quaternion formula). Poses of rigid bodies are taken from the file's
initial transforms: moving bodies need the host to refresh the colliders per step, as the
reference's renderer does each frame.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .capi import SphCollider, SphForceField, SphParams, SphSettings
from .fluid import cell_count_for, derive_fluid_params

F32 = np.float32
_SHAPES = {"sphere": 0, "box": 1, "plane": 2, "cylinder": 3, "capsule": 4, "mesh": 5}
_FIELDS = {"wind": 0, "vortex": 1, "point": 2}


def _vec3(value, default):
    if isinstance(value, list) and len(value) == 3:
        return [F32(v) for v in value]
    return [F32(v) for v in default]


def _quat(value):
    """(w, x, y, z) from the file's [x, y, z, w] (scene_io.cpp:49-54)."""
    if isinstance(value, list) and len(value) == 4:
        return (F32(value[3]), F32(value[0]), F32(value[1]), F32(value[2]))
    return (F32(1), F32(0), F32(0), F32(0))


def mat4_cast(q) -> np.ndarray:
    """glm::mat4_cast (column-major 4 x 4, m[col][row])."""
    w, x, y, z = q
    one, two = F32(1), F32(2)
    qxx, qyy, qzz = x * x, y * y, z * z
    qxz, qxy, qyz = x * z, x * y, y * z
    qwx, qwy, qwz = w * x, w * y, w * z
    m = np.eye(4, dtype=F32)
    m[0, 0] = one - two * (qyy + qzz)
    m[0, 1] = two * (qxy + qwz)
    m[0, 2] = two * (qxz - qwy)
    m[1, 0] = two * (qxy - qwz)
    m[1, 1] = one - two * (qxx + qzz)
    m[1, 2] = two * (qyz + qwx)
    m[2, 0] = two * (qxz + qwy)
    m[2, 1] = two * (qyz - qwx)
    m[2, 2] = one - two * (qxx + qyy)
    return m


def mat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """glm mat4 * mat4: column i = ((a0 b_i0 + a1 b_i1) + a2 b_i2) + a3 b_i3."""
    out = np.empty((4, 4), F32)
    for i in range(4):
        col = a[0] * b[i, 0]
        for k in range(1, 4):
            col = col + a[k] * b[i, k]
        out[i] = col
    return out


def transform_matrix(position, rotation, scale) -> np.ndarray:
    """scene::Transform::matrix: translate(I, p) * mat4_cast(q), then glm::scale."""
    t = np.eye(4, dtype=F32)
    t[3, :3] = position
    m = mat_mul(t, mat4_cast(rotation))
    for k in range(3):
        m[k] = m[k] * scale[k]
    return m


@dataclass
class SceneFluid:
    """One fluid of a scene file, with what deriveFluidParams needs."""
    index: int
    settings: SphSettings
    world: list[float]
    colliders: list = field(default_factory=list)
    fields: list = field(default_factory=list)
    display: str = "particles"
    surface_resolution: int = 48
    surface_iso: float = 0.5

    @property
    def count(self) -> int:
        return int(self.settings.particleCount)

    def params(self, cell_count: int | None = None) -> SphParams:
        c = cell_count if cell_count is not None else cell_count_for(self.count)
        return derive_fluid_params(self.settings, self.world, self.count, c, self.colliders, self.fields)


def _collider(shape: int, body: dict, world: np.ndarray) -> SphCollider:
    """collectShapeColliders' entry for a rotation + positive-scale world matrix
    (colliderPose decomposes the matrix into position, rotation and scale)."""
    position = world[3, :3].copy()
    scale = np.array([math.sqrt(float(np.dot(world[k, :3].astype(np.float64), world[k, :3]))) for k in range(3)],
                     F32)
    rot = np.eye(4, dtype=F32)
    for k in range(3):
        rot[k, :3] = world[k, :3] / scale[k]
    radius = F32(body.get("radius", 0.5))
    half = _vec3(body.get("halfExtents"), [0.5, 0.5, 0.5])
    c = SphCollider()
    c.shape = shape
    if shape in (3, 4):   # cylinder / capsule: the round section takes the X and Z scales
        c.radius = float(radius * max(scale[0], scale[2]))
    else:
        c.radius = float(radius * max(scale))
    c.halfExtents[:] = [float(half[k] * scale[k]) for k in range(3)]
    rigid = rot.copy()
    rigid[3, :3] = position
    inv = np.eye(4, dtype=F32)
    inv[:3, :3] = rot[:3, :3].T
    for k in range(3):   # -R^T p
        inv[3, k] = -(((rot[k, 0] * position[0]) + rot[k, 1] * position[1]) + rot[k, 2] * position[2])
    c.world[:] = [float(v) for v in rigid.reshape(-1)]
    c.inverseWorld[:] = [float(v) for v in inv.reshape(-1)]
    return c


def load_scene(path: str | Path) -> list[SceneFluid]:
    doc = json.loads(Path(path).read_text())
    fluids = doc.get("fluids", [])
    bodies = doc.get("rigidBodies", [])
    forcefields = doc.get("forceFields", [])
    objects = doc.get("objects", [])

    def handle(v, size):
        return v if isinstance(v, int) and 0 <= v < size else -1

    local = [transform_matrix(_vec3(o.get("position"), [0, 0, 0]), _quat(o.get("rotation")),
                              _vec3(o.get("scale"), [1, 1, 1])) for o in objects]
    parents = [handle(o.get("parent", -1), len(objects)) for o in objects]

    def world(i):
        m = local[i]
        p = parents[i]
        seen = 0
        while p >= 0 and seen < len(objects):
            m = mat_mul(local[p], m)
            p = parents[p]
            seen += 1
        return m

    def visible(i):
        seen = 0
        while i >= 0 and seen <= len(objects):
            if not objects[i].get("visible", True):
                return False
            i = parents[i]
            seen += 1
        return True

    colliders, fields = [], []
    for i, o in enumerate(objects):
        b = handle(o.get("rigidBody", -1), len(bodies))
        if b >= 0 and visible(i) and len(colliders) < 8:
            shape = _SHAPES.get(bodies[b].get("shape", "sphere"), 0)
            if shape != 5:   # meshes are not fluid colliders
                colliders.append(_collider(shape, bodies[b], world(i)))
        f = handle(o.get("forceField", -1), len(forcefields))
        if f >= 0 and visible(i) and len(fields) < 8:
            spec = forcefields[f]
            w = world(i)
            ff = SphForceField()
            ff.type = _FIELDS.get(spec.get("type", "wind"), 0)
            ff.position[:] = [float(v) for v in w[3, :3]]
            axis = w[1, :3] if ff.type == 1 else -w[2, :3]
            length = F32(math.sqrt(float((axis[0] * axis[0] + axis[1] * axis[1]) + axis[2] * axis[2])))
            ff.axis[:] = [float(v / length) for v in axis] if length > F32(1e-6) else [0.0, 1.0, 0.0]
            ff.strength = float(spec.get("strength", 10.0))
            ff.radius = max(float(spec.get("radius", 0.0)), 0.0)
            ff.falloff = max(float(spec.get("falloff", 1.0)), 0.0)
            fields.append(ff)

    out = []
    for k, spec in enumerate(fluids):
        s = SphSettings(
            emitterHalfExtents=_vec3(spec.get("emitterHalfExtents"), [0.5, 0.5, 0.5]),
            particleCount=int(spec.get("particleCount", 8192)),
            particleRadius=float(spec.get("particleRadius", 0.025)),
            smoothingRadius=float(spec.get("smoothingRadius", 0.1)),
            restDensity=float(spec.get("restDensity", 1000.0)),
            stiffness=float(spec.get("stiffness", 50.0)),
            viscosity=float(spec.get("viscosity", 0.5)),
            containerMin=_vec3(spec.get("containerMin"), [-1, 0, -1]),
            containerMax=_vec3(spec.get("containerMax"), [1, 2, 1]),
            gravity=_vec3(spec.get("gravity"), [0, -9.81, 0]))
        owner = next((i for i, o in enumerate(objects) if handle(o.get("fluid", -1), len(fluids)) == k), -1)
        w = world(owner) if owner >= 0 else np.eye(4, dtype=F32)
        res = int(spec.get("surfaceResolution", 48))
        out.append(SceneFluid(k, s, [float(v) for v in w.reshape(-1)], list(colliders), list(fields),
                              spec.get("display", "particles"), min(max(res, 8), 128),
                              float(spec.get("surfaceIso", 0.5))))
    return out
