"""Oracle of the learned surrogate — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg may import this module; the
product path never does. The surrogate has no counterpart in the reference (it sits on
top of FluidSolver::step, src/physics/fluid_sph.cpp:151-303), so everything here is a
restatement of this project's own definitions, pinned as follows:

  bf16()            round-to-nearest-even FP32 -> BF16; pinned against torch's
                    independent conversion (tests/test_surrogate_oracle.py).
  knn()             brute-force 8 nearest in-range neighbours by (r2, index), FP32
                    arithmetic in the order the density pass uses; pinned by invariants
                    (sorted distances, in range, no self) and by the kd-tree of scipy.
  walk()            the density pass's neighbour walk (27 buckets in dz, dy, dx order,
                    members by index, 32 per bucket, duplicate visits kept —
                    src/physics/fluid_sph.cpp:194-223) with the surrogate's extra sums:
                    the poly6 density gradient and the in-range count. Pinned by its
                    density equalling the FP32 SPH oracle's bit for bit and its gradient
                    equalling central differences of the FP64 density estimate.
  features()        the feature row from the walk and knn outputs.
  forward()/gradients(), exact mode (FP64, no rounding): pinned against torch.autograd in
                    float64 and central finite differences.
  forward()/gradients(), mirror mode: the same, with BF16 rounding where the GPU rounds
                    (weights, hidden activations, output and hidden gradients); pinned
                    against torch.autograd with BF16-rounding autograd functions placed
                    at those points, and to the exact mode within the BF16 error bound.
  adam():           Kingma & Ba (2015) Algorithm 1 with bias corrections; pinned by its
                    closed form for a constant gradient.
  targets():        mean SPH acceleration (force / density, gravity excluded) over the
                    four substeps of a macro-step, from the FP32 SPH oracle's forces.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
# (inputs, outputs, weight offset, bias offset) of each layer; W[out][in] row-major.
LAYERS = ((64, 128, 0, 8192), (128, 128, 8320, 24704), (128, 128, 24832, 41216), (128, 32, 41344, 45440))
PARAMS = 45472
OUTPUTS = 3
FEATURES = 64
MACRO = 4
ACCEL_SCALE = 98.1
DENSITY_SCALE = 100.0
VELOCITY_SCALE = 1.0
LENGTH_SCALE = 0.01


def bf16(x) -> np.ndarray:
    """FP32 values rounded to BF16 (round to nearest, ties to even), returned as FP32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=F32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    nan &= (u & 0x007FFFFF) != 0
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(F32).reshape(a.shape)


def bf16_bits(x) -> np.ndarray:
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(F32)


def frozen_mask() -> np.ndarray:
    m = np.zeros(PARAMS, dtype=bool)
    nin, nout, w, b = LAYERS[3]
    m[w + OUTPUTS * nin: w + nout * nin] = True
    m[b + OUTPUTS: b + nout] = True
    return m


def split(params):
    """[(W[out][in], b[out]) per layer] as FP64 arrays."""
    p = np.asarray(params, dtype=np.float64)
    return [(p[w: w + nout * nin].reshape(nout, nin), p[b: b + nout]) for nin, nout, w, b in LAYERS]


def _round(a, mirror):
    return bf16(a.astype(F32)).astype(np.float64) if mirror else a


def forward(params, x, mirror: bool = False):
    """y (rows x 3) and the activations [x, h1, h2, h3] for feature rows x (rows x 64).

    mirror: weights rounded to BF16 and hidden activations rounded to BF16 after ReLU,
    as the GPU path stores them; products are then exact (FP64) — the GPU accumulates in
    FP32, which stays within the tolerance the tests state."""
    a = np.asarray(x, dtype=np.float64)
    acts = [a]
    layers = split(params)
    for i, (w, b) in enumerate(layers):
        wq = _round(w, mirror)
        z = a @ wq.T + b
        if i < 3:
            a = _round(np.maximum(z, 0.0), mirror)
            acts.append(a)
        else:
            return z[:, :OUTPUTS], acts
    raise AssertionError


def gradients(params, x, t, mirror: bool = False):
    """(gradient of the mean squared error over rows x 3 outputs, loss)."""
    y, acts = forward(params, x, mirror)
    rows = y.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    layers = split(params)
    grad = np.zeros(PARAMS, dtype=np.float64)
    dz = np.zeros((rows, LAYERS[3][1]))
    if mirror:   # the GPU forms scale * (y - t) in FP32 and stores it as BF16
        yf = y.astype(F32)
        d32 = yf - np.asarray(t, dtype=F32)
        dz[:, :OUTPUTS] = bf16(F32(2.0 / (rows * OUTPUTS)) * d32)
    else:
        dz[:, :OUTPUTS] = 2.0 / (rows * OUTPUTS) * d
    for i in range(3, -1, -1):
        nin, nout, wo, bo = LAYERS[i]
        a = acts[i]
        grad[wo: wo + nout * nin] = (dz.T @ a).reshape(-1)
        grad[bo: bo + nout] = dz.sum(axis=0)
        if i == 0:
            break
        wq = _round(layers[i][0], mirror)
        da = (dz @ wq) * (acts[i] > 0.0)
        dz = _round(da, mirror)
    grad[frozen_mask()] = 0.0
    return grad, loss


def adam(w, m1, m2, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns the new (w, m1, m2)."""
    m1 = b1 * m1 + (1.0 - b1) * g
    m2 = b2 * m2 + (1.0 - b2) * g * g
    mhat = m1 / (1.0 - b1 ** step)
    vhat = m2 / (1.0 - b2 ** step)
    return w - lr * mhat / (np.sqrt(vhat) + eps), m1, m2


def knn(pos3, h):
    """Slots of the 8 nearest neighbours within h of each particle, by (r2, index), with
    r2 = (dx dx + dy dy) + dz dz in FP32, dx = p_self - p_other; -1 pads. Brute force,
    so it equals the grid walk's choice when no bucket is over its 32-particle cap."""
    p = np.asarray(pos3, dtype=F32)
    n = p.shape[0]
    h2 = F32(h) * F32(h)
    out = np.full((n, 8), -1, dtype=np.int64)
    for i in range(n):
        d = p[i][None, :] - p
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        ok = np.nonzero(r2 < h2)[0]
        ok = ok[ok != i]
        order = np.lexsort((ok, r2[ok]))[:8]
        out[i, : len(order)] = ok[order]
    return out


def _hash(cells, cell_count):
    c = cells.astype(np.int64) & 0xFFFFFFFF
    h = ((c[:, 0] * 73856093) & 0xFFFFFFFF) ^ ((c[:, 1] * 19349663) & 0xFFFFFFFF) ^ \
        ((c[:, 2] * 83492791) & 0xFFFFFFFF)
    return h & (cell_count - 1)


def walk(pos4, h, cell_count, mass, c_poly6, cap=32):
    """(density, grad rho (n x 3), in-range count) of the density pass, summed in FP32 in
    the pass's visiting order. c_poly6 must be the pass's FP32 constant."""
    p = np.asarray(pos4, dtype=F32)[:, :3]
    n = p.shape[0]
    hf = F32(h)
    cells = np.floor(p / hf).astype(np.int64)
    buckets = _hash(cells, cell_count)
    order = np.lexsort((np.arange(n), buckets))
    members = {}
    for b in np.unique(buckets):
        members[int(b)] = []
    for i in order:
        lst = members[int(buckets[i])]
        if len(lst) < cap:
            lst.append(int(i))
    offsets = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    cands = []
    for i in range(n):
        lst = []
        for off in offsets:
            b = int(_hash((cells[i] + np.array(off))[None, :], cell_count)[0])
            lst.extend(members.get(b, ()))
        cands.append(lst)
    width = max(len(c) for c in cands)
    cand = np.full((n, width), -1, dtype=np.int64)
    for i, c in enumerate(cands):
        cand[i, : len(c)] = c
    h2 = hf * hf
    m = F32(mass)
    c6 = F32(c_poly6)
    rho = np.zeros(n, F32)
    g = np.zeros((n, 3), F32)
    listed = np.zeros(n, np.int64)
    for c in range(width):
        u = cand[:, c]
        valid = u >= 0
        q = p[np.where(valid, u, 0)]
        d3 = p - q
        r2 = (d3[:, 0] * d3[:, 0] + d3[:, 1] * d3[:, 1]) + d3[:, 2] * d3[:, 2]
        inr = valid & (r2 < h2)
        d = h2 - r2
        rho = np.where(inr, rho + m * (((c6 * d) * d) * d), rho)
        d2 = d * d
        g = np.where(inr[:, None], g + d2[:, None] * d3, g)
        listed += inr
    rho = np.where(rho < F32(1e-6), F32(1e-6), rho)
    gscale = (F32(-6.0) * c6) * m
    return rho, g * gscale, listed


def features(pos4, vel4, density, grad, listed, h, rest_density, nbr=None, v_scale=VELOCITY_SCALE):
    """Feature rows (n x 64, FP32 values before BF16 rounding) of every particle."""
    pos = np.asarray(pos4, dtype=F32)[:, :3]
    vel = np.asarray(vel4, dtype=F32)[:, :3]
    rho = np.asarray(density, dtype=F32)
    if nbr is None:
        nbr = knn(pos, h)
    n = pos.shape[0]
    inv_h = F32(1.0) / F32(h)
    inv_v = F32(1.0) / F32(v_scale)
    inv_rho = F32(1.0) / F32(rest_density)
    f = np.zeros((n, FEATURES), dtype=F32)
    for q in range(8):
        u = nbr[:, q]
        has = u >= 0
        uu = np.where(has, u, 0)
        f[:, 7 * q: 7 * q + 3] = np.where(has[:, None], (pos[uu] - pos) * inv_h, 0)
        f[:, 7 * q + 3: 7 * q + 6] = np.where(has[:, None], (vel[uu] - vel) * inv_v, 0)
        f[:, 7 * q + 6] = np.where(has, (rho[uu] * inv_rho - F32(1.0)) * F32(DENSITY_SCALE), 0)
    f[:, 56] = (rho * inv_rho - F32(1.0)) * F32(DENSITY_SCALE)
    gs = (F32(h) * inv_rho) * F32(DENSITY_SCALE)
    f[:, 57:60] = np.asarray(grad, dtype=F32) * gs
    f[:, 60] = F32(h) / F32(LENGTH_SCALE)
    f[:, 61] = np.asarray(listed).astype(F32) * F32(1.0 / 32.0)
    return f


def targets(forces_and_density):
    """Mean over the substeps of force / density, in units of 98.1 m/s^2; the input is a
    list of (force (n x 3), density (n,)) per substep, indexed by particle."""
    acc = None
    for f, rho in forces_and_density:
        rho = np.maximum(np.asarray(rho, dtype=F32), F32(1e-6))
        a = np.asarray(f, dtype=F32)[:, :3] / rho[:, None]
        acc = a if acc is None else acc + a
    return acc * (F32(1.0) / (F32(len(forces_and_density)) * F32(ACCEL_SCALE)))
