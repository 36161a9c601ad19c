"""Oracle of the learned surrogate — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg may import this module; the
product path never does. The surrogate has no counterpart in the reference (it sits on
top of FluidSolver::step, src/physics/fluid_sph.cpp:151-303), so everything here is a
restatement of this project's own definitions, pinned as follows:

  bf16()            round-to-nearest-even FP32 -> BF16; pinned against torch's
                    independent conversion (tests/test_surrogate_oracle.py).
  walk()            the density pass's neighbour walk (27 buckets in dz, dy, dx order,
                    members by index, 32 per bucket, duplicate visits kept —
                    src/physics/fluid_sph.cpp:194-223) with the surrogate's neighbourhood
                    sums accumulated in that order. Pinned by its density equalling the FP32
                    SPH oracle's bit for bit, by its sums equalling FP64 brute-force sums over
                    the particles within h (where no bucket is capped or folded), and by
                    closed forms (a symmetric neighbourhood has zero odd moments).
  features()        the feature row from the walk's sums.
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


def _hash(cells, cell_count):
    c = cells.astype(np.int64) & 0xFFFFFFFF
    h = ((c[:, 0] * 73856093) & 0xFFFFFFFF) ^ ((c[:, 1] * 19349663) & 0xFFFFFFFF) ^ \
        ((c[:, 2] * 83492791) & 0xFFFFFFFF)
    return h & (cell_count - 1)


def walk(pos4, vel4, h, cell_count, mass, c_poly6, cap=32):
    """(density, sums (n x 20), in-range count) of the density pass: the density and the
    neighbourhood sums of the surrogate, summed in FP32 in the pass's visiting order.
    c_poly6 must be the pass's FP32 constant. sums columns: 0-2 sum s^2 e, 3 count,
    4-9 sum s^3 (xx, yy, zz, xy, xz, yz), 10 sum s^2 (w . e), 11 sum s, 12-14 sum s w,
    15 0, 16-18 sum s^2 (w x e), 19 0; s = (h2 - r2) / h2 as d * (1 / h2), e = (x_i -
    x_j) * (1 / h), w = v_j - v_i."""
    p = np.asarray(pos4, dtype=F32)[:, :3]
    v = np.asarray(vel4, dtype=F32)[:, :3]
    n = p.shape[0]
    hf = F32(h)
    cells = np.floor(p / hf).astype(np.int64)
    buckets = _hash(cells, cell_count)
    order = np.lexsort((np.arange(n), buckets))
    members = {}
    for i in order:
        lst = members.setdefault(int(buckets[i]), [])
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
    inv_h2 = F32(1) / h2
    inv_h = F32(1) / hf
    m = F32(mass)
    c6 = F32(c_poly6)
    rho = np.zeros(n, F32)
    ax = np.zeros((n, 20), F32)
    one = F32(1)
    for c in range(width):
        u = cand[:, c]
        valid = u >= 0
        uu = np.where(valid, u, 0)
        q = p[uu]
        d3 = p - q
        r2 = (d3[:, 0] * d3[:, 0] + d3[:, 1] * d3[:, 1]) + d3[:, 2] * d3[:, 2]
        inr = valid & (r2 < h2)
        d = h2 - r2
        rho = np.where(inr, rho + m * (((c6 * d) * d) * d), rho)
        sw = d * inv_h2
        e = d3 * inv_h
        s2 = sw * sw
        s3 = s2 * sw
        w = v[uu] - v
        terms = np.zeros((n, 20), F32)
        terms[:, 0:3] = s2[:, None] * e
        terms[:, 3] = one
        terms[:, 4] = s3 * (e[:, 0] * e[:, 0])
        terms[:, 5] = s3 * (e[:, 1] * e[:, 1])
        terms[:, 6] = s3 * (e[:, 2] * e[:, 2])
        terms[:, 7] = s3 * (e[:, 0] * e[:, 1])
        terms[:, 8] = s3 * (e[:, 0] * e[:, 2])
        terms[:, 9] = s3 * (e[:, 1] * e[:, 2])
        terms[:, 10] = s2 * ((w[:, 0] * e[:, 0] + w[:, 1] * e[:, 1]) + w[:, 2] * e[:, 2])
        terms[:, 11] = sw
        terms[:, 12:15] = sw[:, None] * w
        terms[:, 16] = s2 * (w[:, 1] * e[:, 2] - w[:, 2] * e[:, 1])
        terms[:, 17] = s2 * (w[:, 2] * e[:, 0] - w[:, 0] * e[:, 2])
        terms[:, 18] = s2 * (w[:, 0] * e[:, 1] - w[:, 1] * e[:, 0])
        live = np.array([k not in (15, 19) for k in range(20)])
        ax[:, live] = np.where(inr[:, None], ax[:, live] + terms[:, live], ax[:, live])
    rho = np.where(rho < F32(1e-6), F32(1e-6), rho)
    return rho, ax


def features(density, sums, h, rest_density, v_scale=VELOCITY_SCALE):
    """Feature rows (n x 64, FP32 values before BF16 rounding) from walk()'s outputs."""
    rho = np.asarray(density, dtype=F32)
    a = np.asarray(sums, dtype=F32)
    n = rho.shape[0]
    inv_v = F32(1.0) / F32(v_scale)
    f = np.zeros((n, FEATURES), dtype=F32)
    f[:, 0] = (rho / F32(rest_density) - F32(1.0)) * F32(DENSITY_SCALE)
    f[:, 1:4] = a[:, 0:3]
    f[:, 4:10] = a[:, 4:10]
    f[:, 10:13] = a[:, 12:15] * inv_v
    f[:, 13] = a[:, 10] * inv_v
    f[:, 14:17] = a[:, 16:19] * inv_v
    f[:, 17] = a[:, 3] * F32(1.0 / 32.0)
    f[:, 18] = a[:, 11] * F32(1.0 / 8.0)
    f[:, 19] = F32(h) / F32(LENGTH_SCALE)
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


def knn_features(pos4, vel4, density, h, cell_count, rest_density, cap=32, v_scale=VELOCITY_SCALE, k=8):
    """The 8-nearest-neighbour rows (n x 64, FP32 values before BF16 rounding) in particle
    order: among the in-range candidates of the density walk (27 buckets in the reference
    order, the first `cap` members of each bucket by index, r2 < h2 in the pass's FP32
    order), the distinct particles j != i with the k smallest (r2, j); for each, (x_i - x_j)
    * (1 / h) and (v_j - v_i) * (1 / v_scale); then 100 (rho / rho0 - 1) and h / 1 cm."""
    p = np.asarray(pos4, dtype=F32)[:, :3]
    v = np.asarray(vel4, dtype=F32)[:, :3]
    rho = np.asarray(density, dtype=F32)
    n = p.shape[0]
    hf = F32(h)
    cells = np.floor(p / hf).astype(np.int64)
    buckets = _hash(cells, cell_count)
    order = np.lexsort((np.arange(n), buckets))
    members = {}
    for i in order:
        lst = members.setdefault(int(buckets[i]), [])
        if len(lst) < cap:
            lst.append(int(i))
    offsets = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    h2 = hf * hf
    inv_h = F32(1) / hf
    inv_v = F32(1) / F32(v_scale)
    f = np.zeros((n, FEATURES), F32)
    for i in range(n):
        seen = {}
        for off in offsets:
            b = int(_hash((cells[i] + np.array(off))[None, :], cell_count)[0])
            for j in members.get(b, ()):
                if j == i or j in seen:
                    continue
                d3 = p[i] - p[j]
                r2 = (d3[0] * d3[0] + d3[1] * d3[1]) + d3[2] * d3[2]
                if r2 < h2:
                    seen[j] = r2
        near = sorted(seen.items(), key=lambda kv: (kv[1], kv[0]))[:k]
        for e, (j, _) in enumerate(near):
            f[i, 6 * e: 6 * e + 3] = (p[i] - p[j]) * inv_h
            f[i, 6 * e + 3: 6 * e + 6] = (v[j] - v[i]) * inv_v
    f[:, 6 * k] = (rho / F32(rest_density) - F32(1.0)) * F32(DENSITY_SCALE)
    f[:, 6 * k + 1] = hf / F32(LENGTH_SCALE)
    return f
